/*
 * ajtai.c -- Ajtai public parameters, balanced b=2 decomposition, commitment, S-action.
 * TEST INFRASTRUCTURE ONLY (checker for nightstream_amd; see neo_oracle.h).
 *
 * Restates:
 *   neo-ajtai/src/commit.rs:21-30   sample_uniform_fq (rejection x < q on next_u64)
 *   neo-ajtai/src/commit.rs:133-156 setup (sequential stream: rows i, columns j, 54 coeffs)
 *   neo-ajtai/src/decomp.rs:118-158 decomp_b_row_major, b = 2, Balanced
 *   neo-ajtai/src/commit.rs:679-744 commit_precomp_ct_row_major:
 *       c_i = sum_j sum_t Z[t][j] * col_t(rot(a_ij)),  col_{t+1} = rot_step(col_t)
 *   neo-ajtai/src/commit.rs:458-493 s_mul / s_lincomb
 * Pinned by: tests/test_oracle_ajtai.py (Python big-int ring products sum_j a_ij * zhat_j
 * reduced mod Phi_81 on small m; decomposition round trip and the digit rule on
 * |z| <= 2^54-1; the linearity commit(Z1)+commit(Z2) = commit(Z1+Z2)).
 * setup's byte stream is parity unpinned (depends on the ChaCha8Rng layout).
 */
#include <stdlib.h>
#include <string.h>
#include "neo_oracle.h"

void or_decomp_b2_row_major(const fq* z, size_t m, int8_t* Z) {
  int64_t* a = (int64_t*)malloc(m * sizeof(int64_t));
  for (size_t c = 0; c < m; c++) a[c] = or_fq_to_balanced(z[c]);
  for (int row = 0; row < NEO_D; row++) {
    for (size_t c = 0; c < m; c++) {
      int64_t a0 = a[c];
      int64_t digit = (a0 & 1) == 0 ? 0 : (a0 >= 0 ? 1 : -1);
      Z[(size_t)row * m + c] = (int8_t)digit;
      a[c] = (a0 - digit) >> 1;
    }
  }
  free(a);
}

void or_ajtai_setup(const uint8_t seed[32], size_t kappa, size_t m, fq* pp) {
  or_chacha rng;
  or_chacha_init(&rng, seed, 8);
  size_t total = kappa * m * NEO_D;
  for (size_t i = 0; i < total; i++) {
    for (;;) {
      uint64_t x = or_chacha_next_u64(&rng);
      if (x < NEO_Q) {
        pp[i] = x;
        break;
      }
    }
  }
}

void or_ajtai_commit(const fq* pp, size_t kappa, size_t m, const int8_t* Z, fq* out) {
  for (size_t i = 0; i < kappa; i++) {
    fq acc_total[NEO_D];
    memset(acc_total, 0, sizeof(acc_total));
#pragma omp parallel
    {
      fq acc[NEO_D];
      fq col[NEO_D], nxt[NEO_D];
      memset(acc, 0, sizeof(acc));
#pragma omp for schedule(static)
      for (long j = 0; j < (long)m; j++) {
        memcpy(col, pp + ((size_t)i * m + (size_t)j) * NEO_D, sizeof(col));
        for (int t = 0; t < NEO_D; t++) {
          int8_t dg = Z[(size_t)t * m + (size_t)j];
          if (dg == 1) {
            for (int r = 0; r < NEO_D; r++) acc[r] = or_fq_add(acc[r], col[r]);
          } else if (dg == -1) {
            for (int r = 0; r < NEO_D; r++) acc[r] = or_fq_sub(acc[r], col[r]);
          }
          or_rot_step(col, nxt);
          memcpy(col, nxt, sizeof(col));
        }
      }
#pragma omp critical
      {
        for (int r = 0; r < NEO_D; r++) acc_total[r] = or_fq_add(acc_total[r], acc[r]);
      }
    }
    memcpy(out + i * NEO_D, acc_total, sizeof(acc_total));
  }
}

void or_s_mul(const fq* rho_coeffs, const fq* c, size_t kappa, fq* out) {
  for (size_t col = 0; col < kappa; col++) or_rq_mul(rho_coeffs, c + col * NEO_D, out + col * NEO_D);
}
