// transcript.cpp -- ChaCha8, Poseidon2 and the Fiat-Shamir transcript (host).
// See transcript.h for the construction and its parity status.
#include "transcript.h"

#include <algorithm>
#include <chrono>

#include <immintrin.h>
#include <string.h>

namespace ns {

static inline uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
#define QR(a, b, c, d)                     \
  a += b; d ^= a; d = rotl(d, 16);         \
  c += d; b ^= c; b = rotl(b, 12);         \
  a += b; d ^= a; d = rotl(d, 8);          \
  c += d; b ^= c; b = rotl(b, 7);

void chacha8_block(const uint32_t key[8], uint64_t counter, uint32_t out[16]) {
  uint32_t in[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574, key[0], key[1], key[2], key[3],
                     key[4],     key[5],     key[6],     key[7],     (uint32_t)counter, (uint32_t)(counter >> 32),
                     0,          0};
  uint32_t x[16];
  memcpy(x, in, sizeof(x));
  for (int r = 0; r < 8; r += 2) {
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

ChaCha8::ChaCha8(const uint8_t seed[32]) {
  for (int i = 0; i < 8; i++)
    key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) | ((uint32_t)seed[4 * i + 2] << 16) |
             ((uint32_t)seed[4 * i + 3] << 24);
}

ChaCha8 ChaCha8::from_u64(uint64_t state) {
  uint8_t seed[32];
  for (int c = 0; c < 8; c++) {
    state = state * 6364136223846793005ULL + 11634580027462260723ULL;
    uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    uint32_t x = (xs >> rot) | (xs << ((32 - rot) & 31));
    for (int b = 0; b < 4; b++) seed[4 * c + b] = (uint8_t)(x >> (8 * b));
  }
  return ChaCha8(seed);
}

void ChaCha8::refill() {
  for (int b = 0; b < 4; b++) chacha8_block(key, counter + b, buf + 16 * b);
  counter += 4;
}

uint32_t ChaCha8::next_u32() {
  if (idx >= 64) {
    refill();
    idx = 0;
  }
  return buf[idx++];
}

uint64_t ChaCha8::next_u64() {
  if (idx < 63) {
    uint64_t v = (uint64_t)buf[idx] | ((uint64_t)buf[idx + 1] << 32);
    idx += 2;
    return v;
  }
  if (idx >= 64) {
    refill();
    idx = 2;
    return (uint64_t)buf[0] | ((uint64_t)buf[1] << 32);
  }
  uint64_t lo = buf[63];
  refill();
  idx = 1;
  return ((uint64_t)buf[0] << 32) | lo;
}

void ChaCha8::fill_bytes(uint8_t* dst, size_t len) {
  size_t done = 0;
  while (done < len) {
    if (idx >= 64) {
      refill();
      idx = 0;
    }
    while (done < len && idx < 64) {
      uint32_t w = buf[idx++];
      for (int b = 0; b < 4 && done < len; b++) dst[done++] = (uint8_t)(w >> (8 * b));
    }
  }
}

static uint64_t sample_fq(ChaCha8& rng) {
  for (;;) {
    uint64_t x = rng.next_u64();
    if (x < Q) return x;
  }
}

__attribute__((target("avx512f"))) static void permute8_avx512(const Poseidon2& P, uint64_t* st);

void Poseidon2::init(int w, ChaCha8& rng) {
  width = w;
  memset(ext_init, 0, sizeof(ext_init));
  memset(ext_term, 0, sizeof(ext_term));
  memset(diag, 0, sizeof(diag));
  for (int r = 0; r < 4; r++)
    for (int i = 0; i < w; i++) ext_init[r][i] = sample_fq(rng);
  for (int r = 0; r < 4; r++)
    for (int i = 0; i < w; i++) ext_term[r][i] = sample_fq(rng);
  for (int r = 0; r < 22; r++) internal_rc[r] = sample_fq(rng);
  for (int i = 0; i < w; i++) diag[i] = sample_fq(rng);  // stand-in diagonal, see transcript.h
  diag0p1 = diag[0] + 1;                                    // diag[0] < q, so no wrap
  if (w == 8) {
    for (int r = 0; r < 4; r++)
      for (int i = 0; i < 8; i++) {
        vec_rc[r][i] = ext_init[r][i];
        vec_rc[4 + r][i] = ext_term[r][i];
      }

    use_avx512 = false;
    if (__builtin_cpu_supports("avx512f")) {
      // enable the vector path only if it reproduces the scalar permutation on a set of states
      use_avx512 = true;
      uint64_t a[8], b[8], seed = 0x9E3779B97F4A7C15ull;
      for (int t = 0; t < 64 && use_avx512; t++) {
        for (int i = 0; i < 8; i++) {
          seed = seed * 6364136223846793005ull + 1442695040888963407ull;
          a[i] = b[i] = (t & 1) ? seed % Q : Q - 1 - (seed & 15);
        }
        permute_scalar(a);
        permute8_avx512(*this, b);
        if (memcmp(a, b, sizeof a) != 0) use_avx512 = false;
      }
      // and only where it is clearly the faster one: on Zen 5 hosts the unrolled scalar permutation
      // (throughput-bound, ~470 ns) beats the vector one (latency-bound, ~500 ns), and a start-up
      // timing within 10% is noise
      if (use_avx512) {
        auto best_ns = [&](bool vec) {
          double best = 1e30;
          uint64_t t[8] = {1, 2, 3, 4, 5, 6, 7, 8};
          for (int trial = 0; trial < 3; trial++) {
            const auto t0 = std::chrono::steady_clock::now();
            for (int k = 0; k < 2000; k++) {
              if (vec) permute8_avx512(*this, t);
              else permute_scalar(t);
            }
            best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
          }
          return best;
        };
        best_ns(true);
        const double tv = best_ns(true), ts = best_ns(false);
        use_avx512 = tv < 0.9 * ts;
      }
    }
  }
}

typedef unsigned __int128 u128;
// Partial reduction of any 128-bit value to a congruent u64 (possibly in [q, 2^64)): the borrow of
// lo - hi_hi has probability ~2^-32 for product outputs and stays a predictable branch; the carry
// of the second addition is folded branch-free.  Permutation outputs are canonicalized once.
static inline uint64_t red(u128 x) {
  uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  uint64_t hi_hi = hi >> 32, hi_lo = hi & EPS;
  uint64_t t0;
  if (__builtin_expect(__builtin_sub_overflow(lo, hi_hi, &t0), 0)) t0 -= EPS;
  uint64_t t1 = (hi_lo << 32) - hi_lo;
  uint64_t t2;
  bool c = __builtin_add_overflow(t0, t1, &t2);
  return t2 + ((uint64_t)0 - (uint64_t)c & EPS);
}

// A congruent u64 of a 128-bit sum whose high word is below 2^31 (sums of a few u64 values):
// lo + hi (2^32 - 1), one carry correction (the wrapped value is below 2^63, so adding 2^32 - 1
// cannot carry again).
static inline uint64_t red_small(u128 x) {
  const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  uint64_t s;
  const bool c = __builtin_add_overflow(lo, (hi << 32) - hi, &s);
  return s + ((uint64_t)0 - (uint64_t)c & EPS);
}

// a + rc for a < 2^64 and rc < q: a + rc < 2^65 - 2^32, so one correction gives a congruent u64
static inline uint64_t add_rc(uint64_t a, uint64_t rc) {
  uint64_t s;
  const bool c = __builtin_add_overflow(a, rc, &s);
  return s + ((uint64_t)0 - (uint64_t)c & EPS);
}

static inline uint64_t sbox7(uint64_t x) {
  uint64_t x2 = red((u128)x * x);
  uint64_t x3 = red((u128)x2 * x);
  uint64_t x4 = red((u128)x2 * x2);
  return red((u128)x3 * x4);
}

// External linear layer: M4 = circ(2,3,1,1) on every 4-lane chunk, then out[i] += sum of lane
// (i mod 4) over all chunks.  Accumulated in 128 bits and reduced once per output.
template <int W>
static inline void ext_layer(uint64_t* st) {
  u128 y[W];
  for (int c = 0; c < W; c += 4) {
    u128 x0 = st[c], x1 = st[c + 1], x2 = st[c + 2], x3 = st[c + 3];
    u128 t01 = x0 + x1, t23 = x2 + x3, t0123 = t01 + t23;
    y[c + 0] = t0123 + t01 + x1;       // 2 x0 + 3 x1 + x2 + x3
    y[c + 1] = t0123 + x1 + 2 * x2;    // x0 + 2 x1 + 3 x2 + x3
    y[c + 2] = t0123 + t23 + x3;       // x0 + x1 + 2 x2 + 3 x3
    y[c + 3] = t0123 + x3 + 2 * x0;    // 3 x0 + x1 + x2 + 2 x3
  }
  u128 sums[4];
  for (int k = 0; k < 4; k++) {
    sums[k] = 0;
    for (int c = 0; c < W; c += 4) sums[k] += y[c + k];
  }
  for (int i = 0; i < W; i++) st[i] = red_small(y[i] + sums[i & 3]);  // below 35 * 2^64
}

template <int W>
static inline void permute_w(const Poseidon2& P, uint64_t* stp) {
  // the state lives in locals: stores through stp could alias P's constants and force reloads
  uint64_t st[W];
#pragma unroll
  for (int i = 0; i < W; i++) st[i] = stp[i];
  ext_layer<W>(st);
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < W; i++) st[i] = sbox7(add_rc(st[i], P.ext_init[r][i]));
    ext_layer<W>(st);
  }
  // Internal rounds: s0' = x7 d0 + sum = x7 (d0 + 1) + rest with rest = sum_{i>0} s_i, so lane 0
  // does not wait for sum; x7 (d0 + 1) = (x^2 * x (d0 + 1)) * x^4 keeps it three multiplies deep.
  uint64_t x = add_rc(st[0], P.internal_rc[0]);
  for (int r = 0; r < 22; r++) {
    u128 rest = 0;
#pragma unroll
    for (int i = 1; i < W; i++) rest += st[i];
    const uint64_t xc = red((u128)x * P.diag0p1);
    const uint64_t x2 = red((u128)x * x);
    const uint64_t x3 = red((u128)x2 * x);
    const uint64_t x3c = red((u128)x2 * xc);
    const uint64_t x4 = red((u128)x2 * x2);
    const uint64_t x7 = red((u128)x3 * x4);
    // a product of two u64 plus two more u64 stays below 2^128, so lane 0 takes one reduction
    // after its last multiply (the rest is reduced beside the chain)
    u128 nx = (u128)x3c * x4 + red_small(rest);
    if (r < 21) nx += P.internal_rc[r + 1];
    const uint64_t sum = red_small(rest + x7);  // below 16 * 2^64
#pragma unroll
    for (int i = 1; i < W; i++) st[i] = red((u128)st[i] * P.diag[i] + sum);
    x = red(nx);
  }
  st[0] = x;
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < W; i++) st[i] = sbox7(add_rc(st[i], P.ext_term[r][i]));
    ext_layer<W>(st);
  }
#pragma unroll
  for (int i = 0; i < W; i++) stp[i] = st[i] >= Q ? st[i] - Q : st[i];
}

void Poseidon2::permute_scalar(uint64_t* st) const {
  if (width == 8) permute_w<8>(*this, st);
  else permute_w<16>(*this, st);
}

// ---- width 8 with AVX-512F: the 8-lane state is one zmm register ----
// External rounds run on the vector (Goldilocks products from four 32x32 multiplies, the
// circulant layer on 32-bit halves so every sum is exact, one reduction per lane).  In the
// internal rounds the s-box chain of lane 0 stays scalar while lanes 1..7 are updated in the
// vector, off the chain: rest' = sum_{i>0} mu_i s_i + 7 * sum.  Same outputs as permute_w<8>.
#define NS_AVX512 __attribute__((target("avx512f")))
NS_AVX512 static inline __m512i v_red(__m512i lo, __m512i hi) {
  const __m512i m32 = _mm512_set1_epi64(0xFFFFFFFFll);
  __m512i hhi = _mm512_srli_epi64(hi, 32), hlo = _mm512_and_si512(hi, m32);
  __m512i t0 = _mm512_sub_epi64(lo, hhi);
  t0 = _mm512_mask_sub_epi64(t0, _mm512_cmplt_epu64_mask(lo, hhi), t0, m32);
  __m512i t1 = _mm512_sub_epi64(_mm512_slli_epi64(hlo, 32), hlo);
  __m512i t2 = _mm512_add_epi64(t0, t1);
  return _mm512_mask_add_epi64(t2, _mm512_cmplt_epu64_mask(t2, t1), t2, m32);
}
NS_AVX512 static inline __m512i v_mul(__m512i a, __m512i b) {
  const __m512i m32 = _mm512_set1_epi64(0xFFFFFFFFll);
  __m512i ah = _mm512_srli_epi64(a, 32), bh = _mm512_srli_epi64(b, 32);
  __m512i ll = _mm512_mul_epu32(a, b), lh = _mm512_mul_epu32(a, bh), hl = _mm512_mul_epu32(ah, b),
          hh = _mm512_mul_epu32(ah, bh);
  __m512i t = _mm512_add_epi64(lh, _mm512_srli_epi64(ll, 32));
  __m512i u = _mm512_add_epi64(hl, _mm512_and_si512(t, m32));
  __m512i lo = _mm512_or_si512(_mm512_slli_epi64(u, 32), _mm512_and_si512(ll, m32));
  __m512i hi = _mm512_add_epi64(_mm512_add_epi64(hh, _mm512_srli_epi64(t, 32)), _mm512_srli_epi64(u, 32));
  return v_red(lo, hi);
}
// a + b for a < 2^64 and b < q: a wrap past 2^64 adds 2^64 mod q = 2^32 - 1 and cannot wrap again
NS_AVX512 static inline __m512i v_addq(__m512i a, __m512i b) {
  __m512i s = _mm512_add_epi64(a, b);
  return _mm512_mask_add_epi64(s, _mm512_cmplt_epu64_mask(s, b), s, _mm512_set1_epi64(0xFFFFFFFFll));
}
NS_AVX512 static inline u128 v_hsum(__m512i x) {
  const __m512i m32 = _mm512_set1_epi64(0xFFFFFFFFll);
  uint64_t L = (uint64_t)_mm512_reduce_add_epi64(_mm512_and_si512(x, m32));
  uint64_t H = (uint64_t)_mm512_reduce_add_epi64(_mm512_srli_epi64(x, 32));
  return ((u128)H << 32) + L;
}
NS_AVX512 static inline __m512i v_sbox(__m512i x) {
  __m512i x2 = v_mul(x, x);
  return v_mul(v_mul(x2, x), v_mul(x2, x2));
}
// circ(2,3,1,1) per 4-lane chunk, then out_i = y_i + (y_i + y_{i xor 4}), plus rc; exact on halves
NS_AVX512 static inline __m512i v_ext(__m512i x, __m512i rc) {
  const __m512i m32 = _mm512_set1_epi64(0xFFFFFFFFll);
  const __m512i p1 = _mm512_set_epi64(4, 7, 6, 5, 0, 3, 2, 1);
  const __m512i p2 = _mm512_set_epi64(5, 4, 7, 6, 1, 0, 3, 2);
  const __m512i p3 = _mm512_set_epi64(6, 5, 4, 7, 2, 1, 0, 3);
  const __m512i sw = _mm512_set_epi64(3, 2, 1, 0, 7, 6, 5, 4);
  __m512i L = _mm512_and_si512(x, m32), H = _mm512_srli_epi64(x, 32);
  __m512i l1 = _mm512_permutexvar_epi64(p1, L), h1 = _mm512_permutexvar_epi64(p1, H);
  __m512i SL = _mm512_add_epi64(_mm512_add_epi64(L, l1),
                                _mm512_add_epi64(_mm512_permutexvar_epi64(p2, L), _mm512_permutexvar_epi64(p3, L)));
  __m512i SH = _mm512_add_epi64(_mm512_add_epi64(H, h1),
                                _mm512_add_epi64(_mm512_permutexvar_epi64(p2, H), _mm512_permutexvar_epi64(p3, H)));
  __m512i yL = _mm512_add_epi64(_mm512_add_epi64(SL, L), _mm512_add_epi64(l1, l1));
  __m512i yH = _mm512_add_epi64(_mm512_add_epi64(SH, H), _mm512_add_epi64(h1, h1));
  __m512i oL = _mm512_add_epi64(_mm512_add_epi64(yL, yL), _mm512_permutexvar_epi64(sw, yL));
  __m512i oH = _mm512_add_epi64(_mm512_add_epi64(yH, yH), _mm512_permutexvar_epi64(sw, yH));
  oL = _mm512_add_epi64(oL, _mm512_and_si512(rc, m32));
  oH = _mm512_add_epi64(oH, _mm512_srli_epi64(rc, 32));
  // oH * 2^32 + oL with oH, oL < 2^38: low 64 bits plus a small top part times 2^64 = 2^32 - 1
  __m512i lo = _mm512_add_epi64(_mm512_slli_epi64(oH, 32), oL);
  __mmask8 c = _mm512_cmplt_epu64_mask(lo, oL);
  __m512i top = _mm512_add_epi64(_mm512_srli_epi64(oH, 32), _mm512_maskz_set1_epi64(c, 1));
  __m512i add = _mm512_sub_epi64(_mm512_slli_epi64(top, 32), top);
  __m512i r = _mm512_add_epi64(lo, add);
  return _mm512_mask_add_epi64(r, _mm512_cmplt_epu64_mask(r, add), r, m32);
}
NS_AVX512 static void permute8_avx512(const Poseidon2& P, uint64_t* st) {
  __m512i x = _mm512_loadu_si512(st);
  x = v_ext(x, _mm512_load_si512(P.vec_rc[0]));
  for (int r = 0; r < 4; r++) x = v_ext(v_sbox(x), r < 3 ? _mm512_load_si512(P.vec_rc[r + 1]) : _mm512_setzero_si512());
  const __m512i mu = _mm512_mask_blend_epi64(1, _mm512_loadu_si512(P.diag), _mm512_setzero_si512());
  uint64_t st0 = (uint64_t)_mm_cvtsi128_si64(_mm512_castsi512_si128(x));
  __m512i v = _mm512_mask_blend_epi64(1, x, _mm512_setzero_si512());
  uint64_t s0 = red((u128)st0 + P.internal_rc[0]);
  uint64_t restr = red(v_hsum(v));
  for (int r = 0; r < 22; r++) {
    const __m512i mv = v_mul(mu, v);
    const u128 W = v_hsum(mv);
    // s0 -> s0^7 (diag0 + 1) in three dependent multiplies: s0 (diag0 + 1) is formed beside s0^2
    // and carried through s0^3 (diag0 + 1) = s0^2 * it; s0^7 itself comes off the chain
    const uint64_t xc = red((u128)s0 * P.diag0p1);
    const uint64_t x2 = red((u128)s0 * s0);
    const uint64_t x3 = red((u128)x2 * s0);
    const uint64_t x3c = red((u128)x2 * xc);
    const uint64_t x4 = red((u128)x2 * x2);
    const uint64_t x7 = red((u128)x3 * x4);
    // x3c * x4 + restr + rc < (2^64 - 1)^2 + 2 (2^64 - 1) = 2^128 - 1: one reduction on the chain
    u128 acc = (u128)x3c * x4 + restr;
    const uint64_t sum = red((u128)restr + x7);
    if (r < 21) acc += P.internal_rc[r + 1];
    s0 = red(acc);
    restr = red(W + (u128)sum * 7);
    v = _mm512_mask_blend_epi64(1, v_addq(mv, _mm512_set1_epi64((long long)sum)), _mm512_setzero_si512());
  }
  x = _mm512_mask_blend_epi64(1, v, _mm512_set1_epi64((long long)s0));
  for (int r = 0; r < 4; r++)
    x = v_ext(v_sbox(v_addq(x, _mm512_load_si512(P.vec_rc[4 + r]))), _mm512_setzero_si512());
  const __m512i q = _mm512_set1_epi64((long long)Q);
  x = _mm512_mask_sub_epi64(x, _mm512_cmpge_epu64_mask(x, q), x, q);
  _mm512_storeu_si512(st, x);
}

void Poseidon2::permute(uint64_t* st) const {
  if (width == 8 && use_avx512) permute8_avx512(*this, st);
  else permute_scalar(st);
}

const Poseidon2& transcript_perm() {
  static Poseidon2 p = [] {
    static const uint8_t SEED[33] = "neo_poseidon2_goldilocks_seed___";
    ChaCha8 rng(SEED);
    Poseidon2 q;
    q.init(8, rng);
    return q;
  }();
  return p;
}

const Poseidon2& digest_perm() {
  static Poseidon2 p = [] {
    ChaCha8 rng = ChaCha8::from_u64(0x434353445F4D4154ULL);
    Poseidon2 q;
    q.init(16, rng);
    return q;
  }();
  return p;
}

// ---------------- transcript ----------------
Transcript::Transcript(const char* app_label) {
  memset(st_, 0, sizeof(st_));
  perm_ = &transcript_perm();
  append_message("neo/transcript/v1|poseidon2-goldilocks-w8-r4", (const uint8_t*)app_label, strlen(app_label));
}

void Transcript::permute() {
  perm_->permute(st_);
  absorbed_ = 0;
  nperm_++;
}

void Transcript::absorb(uint64_t x) {
  if (absorbed_ >= 4) permute();
  st_[absorbed_++] = x;
}

void Transcript::absorb_label(const char* label) {
  for (const char* p = label; *p; p++) absorb((uint8_t)*p);
}

void Transcript::absorb_slice(const uint64_t* in, size_t len) {
  size_t src = 0;
  while (absorbed_ < 4 && src < len) st_[absorbed_++] = in[src++];
  if (absorbed_ == 4) permute();
  while (len - src >= 4) {
    st_[0] = in[src];
    st_[1] = in[src + 1];
    st_[2] = in[src + 2];
    st_[3] = in[src + 3];
    permute();
    src += 4;
  }
  while (src < len) st_[absorbed_++] = in[src++];
}

void Transcript::append_message(const char* label, const uint8_t* msg, size_t len) {
  absorb_label(label);
  absorb(len);
  for (size_t i = 0; i < len; i++) absorb(msg[i]);
}

void Transcript::append_msg_u64(const char* label, uint64_t v) {
  uint8_t b[8];
  for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
  append_message(label, b, 8);
}

void Transcript::append_fields(const char* label, const uint64_t* fs, size_t n) {
  absorb_label(label);
  absorb(n);
  absorb_slice(fs, n);
}

void Transcript::append_u64s(const char* label, const uint64_t* us, size_t n) {
  absorb_label(label);
  absorb(n);
  for (size_t i = 0; i < n; i++) absorb(us[i] % Q);
}

uint64_t Transcript::challenge_field(const char* label) {
  append_message("chal/label", (const uint8_t*)label, strlen(label));
  absorb(1);
  permute();
  return st_[0];
}

void Transcript::challenge_fields(const char* label, size_t n, uint64_t* out) {
  append_message("chal/label", (const uint8_t*)label, strlen(label));
  size_t got = 0;
  while (got < n) {
    absorb(1);
    permute();
    for (size_t i = 0; i < 4 && got < n; i++) out[got++] = st_[i];
  }
}

void Transcript::digest32(uint8_t out[32]) {
  absorb(1);
  permute();
  for (int i = 0; i < 4; i++)
    for (int b = 0; b < 8; b++) out[8 * i + b] = (uint8_t)(st_[i] >> (8 * b));
}

}  // namespace ns
