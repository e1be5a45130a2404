/*
 * params.c -- NeoParams Goldilocks preset, auto-lambda for R1CS, extension policy v1.
 * TEST INFRASTRUCTURE ONLY (checker for nightstream_amd; see neo_oracle.h).
 *
 * Restates neo-params/src/lib.rs:81-144 (new + guard), 151-167 (goldilocks_127),
 * 186-222 (goldilocks_auto_r1cs_ccs_with, min_lambda 96, margin 2), 226-304
 * (bitlen, s_min, extension_check with exact u128 comparisons).
 * Pinned by: tests/test_oracle_params.py (independent Python big-int computation of the
 * same inequalities, and lambda = 119 for n in [2^10, 2^18] as derived by hand from lib.rs).
 */
#include "neo_oracle.h"

typedef unsigned __int128 u128;

static uint32_t bitlen128(u128 x) {
  uint32_t n = 0;
  while (x) {
    n++;
    x >>= 1;
  }
  return n;
}

static uint32_t ceil_log2_128(u128 x) { return x <= 1 ? 0 : bitlen128(x - 1); }

static uint32_t s_min(const or_params* p, uint32_t ell, uint32_t d_sc) {
  u128 ld = (u128)ell * d_sc;
  if (p->lambda < 128) {
    u128 pow2 = (u128)1 << p->lambda;
    /* checked_mul: overflow => skip the s=1 test */
    if (ld == 0 || pow2 <= (~(u128)0) / ld) {
      u128 rhs = pow2 * ld;
      if ((u128)p->q >= rhs) return 1;
    }
  }
  u128 q2 = (u128)p->q * p->q;
  if (p->lambda >= 128) return 3;
  u128 pow2 = (u128)1 << p->lambda;
  if (ld != 0 && pow2 > (~(u128)0) / ld) return 3;
  u128 rhs = pow2 * ld;
  return q2 >= rhs ? 2 : 3;
}

int or_extension_check(const or_params* p, uint32_t ell, uint32_t d_sc, uint32_t* smin_out, int32_t* slack_out) {
  uint32_t sm = s_min(p, ell, d_sc);
  if (sm > 2) return -1;
  u128 q2 = (u128)p->q * p->q;
  u128 ld = (u128)ell * d_sc;
  if (p->lambda >= 128) return -1;
  u128 pow2 = (u128)1 << p->lambda;
  if (ld != 0 && pow2 > (~(u128)0) / ld) return -1;
  u128 rhs = pow2 * ld;
  int32_t slack;
  if (q2 < rhs) {
    slack = -1;
  } else {
    slack = (int32_t)bitlen128(q2) - (int32_t)bitlen128(rhs);
    /* rhs.checked_shl(slack): shl never "overflows" in Rust's checked_shl sense unless
       slack >= 128; the comparison uses the wrapped value exactly like u128::checked_shl. */
    if (slack >= 0 && slack < 128) {
      u128 shifted = rhs << (uint32_t)slack;
      if (q2 < shifted) slack -= 1;
    }
  }
  if (smin_out) *smin_out = sm;
  if (slack_out) *slack_out = slack;
  return 0;
}

int or_params_auto_r1cs(size_t n_rows, or_params* p) {
  p->q = NEO_Q;
  p->eta = 81;
  p->d = 54;
  p->kappa = 16;
  p->m = 1ULL << 24;
  p->b = 2;
  p->k_rho = 12;
  p->B = 4096;
  p->T = 216;
  p->s = 2;
  p->lambda = 127;
  size_t padded = 2;
  if (n_rows != 0) {
    padded = 1;
    while (padded < n_rows) padded <<= 1;
    if (padded < 2) padded = 2;
  }
  uint32_t ell = ceil_log2_128((u128)p->d * padded);
  uint32_t d_sc = 1 + 4; /* 1 + max(u=2, 2b=4, 2) */
  uint32_t lam = p->lambda < 96 ? 96 : p->lambda;
  while (lam >= 96) {
    p->lambda = lam;
    int32_t slack;
    uint32_t sm;
    if (or_extension_check(p, ell, d_sc, &sm, &slack) == 0 && slack >= 2) return 0;
    lam -= 1;
  }
  return -1;
}
