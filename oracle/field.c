/*
 * field.c -- Goldilocks F_q, K = F_q[u]/(u^2-7), and R_q = F_q[X]/(X^54+X^27+1).
 * TEST INFRASTRUCTURE ONLY (checker for nightstream_amd; see neo_oracle.h).
 *
 * Restates: neo-math/src/field.rs:10-89 (Fq = p3 Goldilocks, K = BinomialExtensionField<Fq,2>
 * with W = 7), neo-math/src/ring.rs:54-66 (schoolbook mul), ring.rs:187-203
 * (reduce_mod_phi_81), neo-ajtai/src/commit.rs:46-54 (rot_step_phi_81),
 * neo-ajtai/src/util.rs:20-29 (to_balanced_i64).
 *
 * Pinned by: tests/test_oracle_field.py (Python big-int arithmetic for Fq/K products and
 * inverses, 7 a non-residue (neo-math/tests/field_ext.rs:19), ring product against a
 * big-int polynomial product reduced by Phi_81 by long division, rot_step against
 * multiplication by X).
 */
#include "neo_oracle.h"

static inline fq reduce128(unsigned __int128 x) {
  uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  uint64_t hi_hi = hi >> 32, hi_lo = hi & 0xFFFFFFFFULL;
  /* x = lo + 2^64 hi_lo + 2^96 hi_hi ; 2^64 = 2^32-1, 2^96 = -1 (mod q).  Carries and borrows
     become masks (random data would otherwise mispredict). */
  uint64_t t0 = lo - hi_hi;
  t0 -= (0 - (uint64_t)(lo < hi_hi)) & 0xFFFFFFFFULL;
  uint64_t t1 = (hi_lo << 32) - hi_lo;
  uint64_t t2 = t0 + t1;
  t2 += (0 - (uint64_t)(t2 < t1)) & 0xFFFFFFFFULL;
  return t2 >= NEO_Q ? t2 - NEO_Q : t2;
}

fq or_fq_add(fq a, fq b) {
  uint64_t s = a + b;
  s += (0 - (uint64_t)(s < a)) & 0xFFFFFFFFULL;
  return s >= NEO_Q ? s - NEO_Q : s;
}
fq or_fq_sub(fq a, fq b) {
  uint64_t d = a - b;
  return d - ((0 - (uint64_t)(a < b)) & 0xFFFFFFFFULL);
}
fq or_fq_neg(fq a) { return a ? NEO_Q - a : 0; }
fq or_fq_mul(fq a, fq b) { return reduce128((unsigned __int128)a * b); }
fq or_fq_pow(fq a, uint64_t e) {
  fq r = 1;
  while (e) {
    if (e & 1) r = or_fq_mul(r, a);
    a = or_fq_mul(a, a);
    e >>= 1;
  }
  return r;
}
fq or_fq_inv(fq a) { return or_fq_pow(a, NEO_Q - 2); }
fq or_fq_from_i64(int64_t v) {
  if (v >= 0) return (fq)v % NEO_Q;
  uint64_t m = (uint64_t)(-(v + 1)) + 1; /* |v| without overflow */
  return or_fq_neg(m % NEO_Q);
}
int64_t or_fq_to_balanced(fq a) {
  const uint64_t half = (NEO_Q - 1) / 2;
  if (a <= half) return (int64_t)a;
  return -(int64_t)(NEO_Q - a);
}

kx or_k_add(kx a, kx b) { kx r = {or_fq_add(a.c0, b.c0), or_fq_add(a.c1, b.c1)}; return r; }
kx or_k_sub(kx a, kx b) { kx r = {or_fq_sub(a.c0, b.c0), or_fq_sub(a.c1, b.c1)}; return r; }
kx or_k_neg(kx a) { kx r = {or_fq_neg(a.c0), or_fq_neg(a.c1)}; return r; }
kx or_k_mul(kx a, kx b) {
  /* (a0 + a1 u)(b0 + b1 u) = (a0 b0 + 7 a1 b1) + (a0 b1 + a1 b0) u */
  fq a0b0 = or_fq_mul(a.c0, b.c0), a1b1 = or_fq_mul(a.c1, b.c1);
  kx r;
  r.c0 = or_fq_add(a0b0, or_fq_mul(7, a1b1));
  r.c1 = or_fq_add(or_fq_mul(a.c0, b.c1), or_fq_mul(a.c1, b.c0));
  return r;
}
kx or_k_scale(kx a, fq s) { kx r = {or_fq_mul(a.c0, s), or_fq_mul(a.c1, s)}; return r; }
kx or_k_inv(kx a) {
  /* 1/(a0 + a1 u) = (a0 - a1 u) / (a0^2 - 7 a1^2) */
  fq n = or_fq_sub(or_fq_mul(a.c0, a.c0), or_fq_mul(7, or_fq_mul(a.c1, a.c1)));
  fq ni = or_fq_inv(n);
  kx r = {or_fq_mul(a.c0, ni), or_fq_mul(or_fq_neg(a.c1), ni)};
  return r;
}
kx or_k_from(fq a) { kx r = {a, 0}; return r; }
int or_k_eq(kx a, kx b) { return a.c0 == b.c0 && a.c1 == b.c1; }

void or_fq_mul_batch(const fq* a, const fq* b, fq* out, size_t n) {
  for (size_t i = 0; i < n; i++) out[i] = or_fq_mul(a[i], b[i]);
}
void or_k_mul_batch(const fq* a, const fq* b, fq* out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    kx x = {a[2 * i], a[2 * i + 1]}, y = {b[2 * i], b[2 * i + 1]};
    kx r = or_k_mul(x, y);
    out[2 * i] = r.c0;
    out[2 * i + 1] = r.c1;
  }
}
void or_k_inv_batch(const fq* a, fq* out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    kx x = {a[2 * i], a[2 * i + 1]};
    kx r = or_k_inv(x);
    out[2 * i] = r.c0;
    out[2 * i + 1] = r.c1;
  }
}

/* ring.rs:54-66 schoolbook product, then ring.rs:187-203 single downward reduction pass */
void or_rq_mul(const fq* a, const fq* b, fq* out) {
  fq tmp[2 * NEO_D - 1];
  for (int i = 0; i < 2 * NEO_D - 1; i++) tmp[i] = 0;
  for (int i = 0; i < NEO_D; i++)
    for (int j = 0; j < NEO_D; j++) tmp[i + j] = or_fq_add(tmp[i + j], or_fq_mul(a[i], b[j]));
  for (int i = 2 * NEO_D - 2; i >= NEO_D; i--) {
    fq t = tmp[i];
    tmp[i] = 0;
    tmp[i - NEO_D] = or_fq_sub(tmp[i - NEO_D], t);
    int i27 = i - 27;
    if (i27 < NEO_D) {
      tmp[i27] = or_fq_sub(tmp[i27], t);
    } else {
      tmp[i27 - NEO_D] = or_fq_add(tmp[i27 - NEO_D], t);
      if (i27 - 27 < NEO_D) tmp[i27 - 27] = or_fq_add(tmp[i27 - 27], t);
    }
  }
  for (int i = 0; i < NEO_D; i++) out[i] = tmp[i];
}

/* commit.rs:46-54: next = X * cur mod Phi_81 */
void or_rot_step(const fq* cur, fq* next) {
  fq last = cur[NEO_D - 1];
  next[0] = 0;
  for (int k = 1; k < NEO_D; k++) next[k] = cur[k - 1];
  next[0] = or_fq_sub(next[0], last);
  next[27] = or_fq_sub(next[27], last);
}
