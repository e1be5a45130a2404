/*
 * transcript.c -- ChaCha8 stream, Poseidon2 permutation and the Fiat-Shamir transcript.
 * TEST INFRASTRUCTURE ONLY (checker for nightstream_amd; see neo_oracle.h).
 *
 * Restates:
 *   neo-transcript/src/poseidon2.rs:20-335  (overwrite sponge, rate 4, labels absorbed one
 *                                            byte per element, challenge/digest squeezes)
 *   neo-ccs/src/crypto/poseidon2_goldilocks.rs:40-43 (permutation from ChaCha8Rng::from_seed(SEED))
 *   neo-params/src/poseidon2_goldilocks.rs:60  (SEED)
 *   neo-reductions/src/engines/utils.rs:214-321 (CCS digest, width-16 Poseidon2, seed_from_u64)
 * Third-party algorithms restated from their published descriptions (not present in the
 * reference checkout; pinned versions in Cargo.lock):
 *   rand_chacha 0.9.0 ChaCha8Rng: ChaCha block function (RFC 8439 quarter round, 8 rounds),
 *     64-bit block counter + 64-bit stream id, 4-block output buffer, BlockRng next_u32 /
 *     next_u64 / fill_bytes consumption rules; rand_core 0.9.3 seed_from_u64 (PCG32 fill).
 *   p3-poseidon2 / p3-goldilocks 0.4.1: x^7 S-box, 8 external + 22 internal rounds, external
 *     layer = M4 = circ(2,3,1,1) per 4-lane chunk followed by adding the per-lane column sums,
 *     internal layer = diag(mu) + all-ones, constants drawn by new_from_rng_128 as
 *     [initial external][terminal external][internal] with rejection sampling x < q.
 *
 * Parity status:
 *   ChaCha core: pinned (tests run the 20-round block against `openssl enc -chacha20`,
 *     the same code path with rounds=8 is used for ChaCha8).
 *   ChaCha8Rng buffer/word layout and seed_from_u64: parity unpinned.
 *   Poseidon2: structure as above; the internal-layer diagonal mu is a STAND-IN (p3's fixed
 *     MATRIX_DIAG constants are not in this checkout and are not filled from memory): mu is
 *     drawn from the same ChaCha8 stream after the round constants.  parity unpinned.
 *   Transcript framing (labels, lengths, rate-4 overwrite, squeeze rules): restated
 *     line-by-line; parity with the reference's bytes is unpinned because the permutation is.
 */
#include <string.h>
#include "neo_oracle.h"

/* ---------------- ChaCha ---------------- */
#define ROTL32(x, n) (((x) << (n)) | ((x) >> (32 - (n))))
#define QR(a, b, c, d)              \
  a += b; d ^= a; d = ROTL32(d, 16); \
  c += d; b ^= c; b = ROTL32(b, 12); \
  a += b; d ^= a; d = ROTL32(d, 8);  \
  c += d; b ^= c; b = ROTL32(b, 7);

static void chacha_core(const uint32_t in[16], int rounds, uint32_t out[16]) {
  uint32_t x[16];
  memcpy(x, in, sizeof(x));
  for (int i = 0; i < rounds; i += 2) {
    QR(x[0], x[4], x[8], x[12]);
    QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]);
    QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]);
    QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]);
    QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

static const uint32_t SIGMA[4] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};

void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds, uint32_t out[16]) {
  uint32_t in[16];
  for (int i = 0; i < 4; i++) in[i] = SIGMA[i];
  for (int i = 0; i < 8; i++) in[4 + i] = key[i];
  in[12] = (uint32_t)counter;
  in[13] = (uint32_t)(counter >> 32);
  in[14] = (uint32_t)stream;
  in[15] = (uint32_t)(stream >> 32);
  chacha_core(in, rounds, out);
}

void or_chacha_ietf_block(const uint32_t key[8], uint32_t counter, const uint32_t nonce[3], int rounds,
                          uint32_t out[16]) {
  uint32_t in[16];
  for (int i = 0; i < 4; i++) in[i] = SIGMA[i];
  for (int i = 0; i < 8; i++) in[4 + i] = key[i];
  in[12] = counter;
  in[13] = nonce[0];
  in[14] = nonce[1];
  in[15] = nonce[2];
  chacha_core(in, rounds, out);
}

void or_chacha_init(or_chacha* r, const uint8_t seed[32], int rounds) {
  for (int i = 0; i < 8; i++)
    r->key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) | ((uint32_t)seed[4 * i + 2] << 16) |
                ((uint32_t)seed[4 * i + 3] << 24);
  r->counter = 0;
  r->stream = 0;
  r->idx = 64;
  r->rounds = rounds;
}

/* rand_core 0.9 SeedableRng::seed_from_u64: PCG32 output fills the seed 4 bytes at a time. */
void or_chacha_seed_from_u64(or_chacha* r, uint64_t state, int rounds) {
  uint8_t seed[32];
  for (int c = 0; c < 8; c++) {
    state = state * 6364136223846793005ULL + 11634580027462260723ULL;
    uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    uint32_t x = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    seed[4 * c] = (uint8_t)x;
    seed[4 * c + 1] = (uint8_t)(x >> 8);
    seed[4 * c + 2] = (uint8_t)(x >> 16);
    seed[4 * c + 3] = (uint8_t)(x >> 24);
  }
  or_chacha_init(r, seed, rounds);
}

static void chacha_refill(or_chacha* r) {
  for (int b = 0; b < 4; b++) or_chacha_block(r->key, r->counter + (uint64_t)b, r->stream, r->rounds, r->buf + 16 * b);
  r->counter += 4;
}

uint32_t or_chacha_next_u32(or_chacha* r) {
  if (r->idx >= 64) {
    chacha_refill(r);
    r->idx = 0;
  }
  return r->buf[r->idx++];
}

uint64_t or_chacha_next_u64(or_chacha* r) {
  if (r->idx < 63) {
    uint64_t v = (uint64_t)r->buf[r->idx] | ((uint64_t)r->buf[r->idx + 1] << 32);
    r->idx += 2;
    return v;
  } else if (r->idx >= 64) {
    chacha_refill(r);
    r->idx = 2;
    return (uint64_t)r->buf[0] | ((uint64_t)r->buf[1] << 32);
  } else {
    uint64_t x = r->buf[63];
    chacha_refill(r);
    r->idx = 1;
    return ((uint64_t)r->buf[0] << 32) | x;
  }
}

void or_chacha_fill_bytes(or_chacha* r, uint8_t* dst, size_t len) {
  size_t done = 0;
  while (done < len) {
    if (r->idx >= 64) {
      chacha_refill(r);
      r->idx = 0;
    }
    /* consume whole words; a partial trailing word is consumed entirely */
    while (done < len && r->idx < 64) {
      uint32_t w = r->buf[r->idx++];
      for (int b = 0; b < 4 && done < len; b++) dst[done++] = (uint8_t)(w >> (8 * b));
    }
  }
}

/* ---------------- Poseidon2 ---------------- */
static fq sample_fq(or_chacha* rng) {
  for (;;) {
    uint64_t x = or_chacha_next_u64(rng);
    if (x < NEO_Q) return x;
  }
}

void or_p2_init_from_seed(or_p2* p, int width, or_chacha* rng) {
  memset(p, 0, sizeof(*p));
  p->width = width;
  p->rounds_f = 8;
  p->rounds_p = 22;
  for (int r = 0; r < 4; r++)
    for (int i = 0; i < width; i++) p->ext_init[r][i] = sample_fq(rng);
  for (int r = 0; r < 4; r++)
    for (int i = 0; i < width; i++) p->ext_term[r][i] = sample_fq(rng);
  for (int r = 0; r < p->rounds_p; r++) p->internal_rc[r] = sample_fq(rng);
  /* STAND-IN internal diagonal (see file header). */
  for (int i = 0; i < width; i++) p->diag[i] = sample_fq(rng);
}

static inline fq sbox7(fq x) {
  fq x2 = or_fq_mul(x, x);
  fq x3 = or_fq_mul(x2, x);
  fq x4 = or_fq_mul(x2, x2);
  return or_fq_mul(x3, x4);
}

static void mat4(fq* x) {
  /* circ(2,3,1,1) */
  fq t01 = or_fq_add(x[0], x[1]);
  fq t23 = or_fq_add(x[2], x[3]);
  fq t0123 = or_fq_add(t01, t23);
  fq t01123 = or_fq_add(t0123, x[1]);
  fq t01233 = or_fq_add(t0123, x[3]);
  fq y3 = or_fq_add(t01233, or_fq_add(x[0], x[0]));
  fq y1 = or_fq_add(t01123, or_fq_add(x[2], x[2]));
  fq y0 = or_fq_add(t01123, t01);
  fq y2 = or_fq_add(t01233, t23);
  x[0] = y0;
  x[1] = y1;
  x[2] = y2;
  x[3] = y3;
}

static void ext_linear(fq* st, int w) {
  for (int c = 0; c < w; c += 4) mat4(st + c);
  fq sums[4] = {0, 0, 0, 0};
  for (int c = 0; c < w; c += 4)
    for (int k = 0; k < 4; k++) sums[k] = or_fq_add(sums[k], st[c + k]);
  for (int i = 0; i < w; i++) st[i] = or_fq_add(st[i], sums[i % 4]);
}

void or_p2_permute(const or_p2* p, fq* st) {
  int w = p->width;
  ext_linear(st, w);
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < w; i++) st[i] = sbox7(or_fq_add(st[i], p->ext_init[r][i]));
    ext_linear(st, w);
  }
  for (int r = 0; r < p->rounds_p; r++) {
    st[0] = sbox7(or_fq_add(st[0], p->internal_rc[r]));
    fq sum = 0;
    for (int i = 0; i < w; i++) sum = or_fq_add(sum, st[i]);
    for (int i = 0; i < w; i++) st[i] = or_fq_add(or_fq_mul(st[i], p->diag[i]), sum);
  }
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < w; i++) st[i] = sbox7(or_fq_add(st[i], p->ext_term[r][i]));
    ext_linear(st, w);
  }
}

static or_p2 g_perm8, g_perm16;
static int g_perm8_ready = 0, g_perm16_ready = 0;

const or_p2* or_p2_transcript_perm(void) {
  if (!g_perm8_ready) {
    static const uint8_t SEED[32] = "neo_poseidon2_goldilocks_seed___";
    or_chacha rng;
    or_chacha_init(&rng, SEED, 8);
    or_p2_init_from_seed(&g_perm8, 8, &rng);
    g_perm8_ready = 1;
  }
  return &g_perm8;
}

const or_p2* or_p2_digest_perm(void) {
  if (!g_perm16_ready) {
    or_chacha rng;
    or_chacha_seed_from_u64(&rng, 0x434353445F4D4154ULL, 8);
    or_p2_init_from_seed(&g_perm16, 16, &rng);
    g_perm16_ready = 1;
  }
  return &g_perm16;
}

void or_p2_permute8(fq* state) { or_p2_permute(or_p2_transcript_perm(), state); }

/* ---------------- transcript (poseidon2.rs) ---------------- */
static uint64_t g_tr_perms = 0;
uint64_t or_tr_permutation_count(void) { return g_tr_perms; }

static void tr_permute(or_tr* t) {
  g_tr_perms++;
  or_p2_permute(or_p2_transcript_perm(), t->st);
  t->absorbed = 0;
}

static void tr_absorb_elem(or_tr* t, fq x) { /* poseidon2.rs:21-27 */
  if (t->absorbed >= 4) tr_permute(t);
  t->st[t->absorbed++] = x;
}

static void tr_absorb_slice(or_tr* t, const fq* in, size_t len) { /* poseidon2.rs:30-61 */
  size_t src = 0;
  while (t->absorbed < 4 && src < len) t->st[t->absorbed++] = in[src++];
  if (t->absorbed == 4) tr_permute(t);
  while (len - src >= 4) {
    t->st[0] = in[src];
    t->st[1] = in[src + 1];
    t->st[2] = in[src + 2];
    t->st[3] = in[src + 3];
    tr_permute(t);
    src += 4;
  }
  while (src < len) t->st[t->absorbed++] = in[src++];
}

static void tr_absorb_label(or_tr* t, const uint8_t* label, size_t n) {
  for (size_t i = 0; i < n; i++) tr_absorb_elem(t, label[i]);
}

void or_tr_new(or_tr* t, const uint8_t* app_label, size_t app_len) { /* poseidon2.rs:77-86 */
  memset(t, 0, sizeof(*t));
  static const char* APP_DOMAIN = "neo/transcript/v1|poseidon2-goldilocks-w8-r4";
  tr_absorb_label(t, (const uint8_t*)APP_DOMAIN, strlen(APP_DOMAIN));
  tr_absorb_elem(t, app_len);
  tr_absorb_label(t, app_label, app_len);
}

void or_tr_append_message(or_tr* t, const char* label, const uint8_t* msg, size_t len) { /* 88-100 */
  tr_absorb_label(t, (const uint8_t*)label, strlen(label));
  tr_absorb_elem(t, len);
  for (size_t i = 0; i < len; i++) tr_absorb_elem(t, msg[i]);
}

void or_tr_append_fields(or_tr* t, const char* label, const fq* fs, size_t n) { /* 102-113 */
  tr_absorb_label(t, (const uint8_t*)label, strlen(label));
  tr_absorb_elem(t, n);
  tr_absorb_slice(t, fs, n);
}

void or_tr_append_u64s(or_tr* t, const char* label, const uint64_t* us, size_t n) { /* 245-256 */
  tr_absorb_label(t, (const uint8_t*)label, strlen(label));
  tr_absorb_elem(t, n);
  for (size_t i = 0; i < n; i++) tr_absorb_elem(t, us[i] % NEO_Q);
}

fq or_tr_challenge_field(or_tr* t, const char* label) { /* 140-153 */
  or_tr_append_message(t, "chal/label", (const uint8_t*)label, strlen(label));
  tr_absorb_elem(t, 1);
  tr_permute(t);
  return t->st[0];
}

void or_tr_challenge_fields(or_tr* t, const char* label, size_t n, fq* out) { /* 305-320 */
  or_tr_append_message(t, "chal/label", (const uint8_t*)label, strlen(label));
  size_t got = 0;
  while (got < n) {
    tr_absorb_elem(t, 1);
    tr_permute(t);
    for (size_t i = 0; i < 4 && got < n; i++) out[got++] = t->st[i];
  }
}

void or_tr_digest32(or_tr* t, uint8_t out[32]) { /* 161-178 */
  tr_absorb_elem(t, 1);
  tr_permute(t);
  for (int i = 0; i < 4; i++)
    for (int b = 0; b < 8; b++) out[8 * i + b] = (uint8_t)(t->st[i] >> (8 * b));
}
