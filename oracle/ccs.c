/*
 * ccs.c -- CCS structure from R1CS, CSC storage, CCS matrix digest.
 * TEST INFRASTRUCTURE ONLY (checker for nightstream_amd; see neo_oracle.h).
 *
 * Restates:
 *   neo-ccs/src/r1cs.rs:13-43          identity-first [I, A, B, C], f = X1*X2 - X3 shifted by one var
 *   neo-ccs/src/relations.rs:27-64     Identity hint -> CcsMatrix::Identity, otherwise CSC
 *   neo-reductions/src/engines/utils.rs:214-321  digest_ccs_matrices (width-16 Poseidon2,
 *                                      15-element overwrite rate, row-major nonzero order)
 * Pinned by: tests/test_oracle_ccs.py (CSC round trip against the CSR input; row-major emit
 * order).  The digest value is parity unpinned (Poseidon2 constants, see transcript.c).
 */
#include <stdlib.h>
#include <string.h>
#include "neo_oracle.h"

static void csr_to_csc(size_t n, size_t m, const uint64_t* rp, const uint32_t* ci, const fq* v, or_mat* out) {
  size_t nnz = (size_t)rp[n];
  out->n = n;
  out->m = m;
  out->identity = 0;
  out->nnz = nnz;
  out->col_ptr = (size_t*)calloc(m + 1, sizeof(size_t));
  out->row_idx = (uint32_t*)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
  out->vals = (fq*)malloc((nnz ? nnz : 1) * sizeof(fq));
  for (size_t k = 0; k < nnz; k++) out->col_ptr[ci[k] + 1]++;
  for (size_t c = 0; c < m; c++) out->col_ptr[c + 1] += out->col_ptr[c];
  size_t* pos = (size_t*)malloc((m + 1) * sizeof(size_t));
  memcpy(pos, out->col_ptr, (m + 1) * sizeof(size_t));
  /* rows ascending inside each column (dense column-major scan order) */
  for (size_t r = 0; r < n; r++)
    for (uint64_t k = rp[r]; k < rp[r + 1]; k++) {
      size_t c = ci[k];
      out->row_idx[pos[c]] = (uint32_t)r;
      out->vals[pos[c]] = v[k];
      pos[c]++;
    }
  free(pos);
}

or_ccs* or_ccs_from_r1cs(size_t n, size_t m, const uint64_t* a_rp, const uint32_t* a_ci, const fq* a_v,
                         const uint64_t* b_rp, const uint32_t* b_ci, const fq* b_v, const uint64_t* c_rp,
                         const uint32_t* c_ci, const fq* c_v) {
  or_ccs* s = (or_ccs*)calloc(1, sizeof(or_ccs));
  s->n = n;
  s->m = m;
  s->t = 4;
  s->mats[0].identity = 1;
  s->mats[0].n = n;
  s->mats[0].m = m;
  csr_to_csc(n, m, a_rp, a_ci, a_v, &s->mats[1]);
  csr_to_csc(n, m, b_rp, b_ci, b_v, &s->mats[2]);
  csr_to_csc(n, m, c_rp, c_ci, c_v, &s->mats[3]);
  s->nterms = 2;
  s->term_coeff[0] = 1;
  s->term_exps[0][0] = 0;
  s->term_exps[0][1] = 1;
  s->term_exps[0][2] = 1;
  s->term_exps[0][3] = 0;
  s->term_coeff[1] = NEO_Q - 1;
  s->term_exps[1][0] = 0;
  s->term_exps[1][1] = 0;
  s->term_exps[1][2] = 0;
  s->term_exps[1][3] = 1;
  return s;
}

void or_ccs_free(or_ccs* s) {
  if (!s) return;
  for (size_t j = 0; j < s->t; j++) {
    free(s->mats[j].col_ptr);
    free(s->mats[j].row_idx);
    free(s->mats[j].vals);
  }
  free(s);
}

typedef struct {
  fq st[16];
  int absorbed;
  const or_p2* p;
} dg_state;

static void dg_emit(dg_state* d, uint64_t row, uint64_t col, uint64_t val) {
  if (d->absorbed + 3 > 15) {
    or_p2_permute(d->p, d->st);
    d->absorbed = 0;
  }
  d->st[d->absorbed] = row % NEO_Q;
  d->st[d->absorbed + 1] = col % NEO_Q;
  d->st[d->absorbed + 2] = val % NEO_Q;
  d->absorbed += 3;
}

void or_ccs_digest(const or_ccs* s, fq out[4]) {
  dg_state d;
  memset(&d, 0, sizeof(d));
  d.p = or_p2_digest_perm();
  const char* dom = "neo/ccs/matrices/v1";
  for (size_t i = 0; i < strlen(dom); i++) {
    if (d.absorbed >= 15) {
      or_p2_permute(d.p, d.st);
      d.absorbed = 0;
    }
    d.st[d.absorbed++] = (uint8_t)dom[i];
  }
  if (d.absorbed + 3 >= 16) {
    or_p2_permute(d.p, d.st);
    d.absorbed = 0;
  }
  d.st[d.absorbed] = s->n;
  d.st[d.absorbed + 1] = s->m;
  d.st[d.absorbed + 2] = s->t;
  or_p2_permute(d.p, d.st);
  for (size_t j = 0; j < s->t; j++) {
    d.absorbed = 0;
    d.st[0] = j;
    d.absorbed = 1;
    const or_mat* M = &s->mats[j];
    if (M->identity) {
      for (size_t r = 0; r < s->n; r++) dg_emit(&d, r, r, 1);
    } else {
      /* row-major enumeration of the CSC nonzeros; columns ascending within a row */
      size_t nrows = M->n, nnz = M->nnz;
      size_t* off = (size_t*)calloc(nrows + 1, sizeof(size_t));
      for (size_t k = 0; k < nnz; k++) off[M->row_idx[k] + 1]++;
      for (size_t r = 0; r < nrows; r++) off[r + 1] += off[r];
      size_t* wp = (size_t*)malloc((nrows + 1) * sizeof(size_t));
      memcpy(wp, off, (nrows + 1) * sizeof(size_t));
      uint64_t* ec = (uint64_t*)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
      uint64_t* ev = (uint64_t*)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
      for (size_t c = 0; c < M->m; c++)
        for (size_t k = M->col_ptr[c]; k < M->col_ptr[c + 1]; k++) {
          size_t r = M->row_idx[k];
          ec[wp[r]] = c;
          ev[wp[r]] = M->vals[k];
          wp[r]++;
        }
      for (size_t r = 0; r < nrows; r++)
        for (size_t e = off[r]; e < off[r + 1]; e++) dg_emit(&d, r, ec[e], ev[e]);
      free(off);
      free(wp);
      free(ec);
      free(ev);
    }
    or_p2_permute(d.p, d.st);
  }
  for (int i = 0; i < 4; i++) out[i] = d.st[i];
}
