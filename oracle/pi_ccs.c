/*
 * pi_ccs.c -- Pi_CCS round oracle: row-phase streaming tables, round polynomials, in-place
 * fold, precompute_for_r, and the sum-check interpolation helpers.
 * TEST INFRASTRUCTURE ONLY (checker for nightstream_amd; see neo_oracle.h).
 *
 * Restates neo-reductions/src/engines/optimized_engine/oracle.rs:
 *   82-414     RowStreamState::build (chi tables, z1 recomposition, f-var SpMV tables,
 *              NC digit tables, w_gamma_nc, Eval table, gamma^k)
 *   417-503    fold_inplace (T[i] <- T[2i] + (T[2i+1]-T[2i]) r for every table)
 *   1047-1178  evals_row_phase_impl, b = 2 generic-K path: per row pair, the univariate
 *              eq_beta(X)*(f(M z1)(X) + sum_{i,rho} w[i,rho] N(y(X))) + gamma^k eq_rin(X) Eval(X),
 *              N(y) = y^3 - y, coefficients low->high, evaluated at xs
 *   1517-1532  chi_tail_weights
 *   1659-1783  precompute_for_r (v_j = M_j^T chi_r, y_eval[i][j] = Z_i v_j)
 * and neo-reductions/src/sumcheck.rs:48-58 (poly_eval_k), 79-121 (interpolate_from_evals).
 *
 * Pinned by: tests/test_oracle_pi_ccs.py -- the round polynomial of every round equals a
 * brute-force sum of Q over the Boolean hypercube computed from the paper formula
 * (paper_exact.rs, re-derived in Python on tiny n); p(0)+p(1) equals the running sum; the
 * folded tables equal direct multilinear evaluation; y_eval equals a dense Z * M^T chi_r.
 */
#include <stdlib.h>
#include <string.h>
#include "neo_oracle.h"

static const kx K0 = {0, 0};
static const kx K1 = {1, 0};

void or_chi_table(const kx* bits, size_t nbits, kx* w) {
  size_t len = (size_t)1 << nbits;
  for (size_t i = 0; i < len; i++) w[i] = K0;
  w[0] = K1;
  for (size_t i = 0; i < nbits; i++) {
    size_t step = (size_t)1 << i;
    kx om = or_k_sub(K1, bits[i]);
    for (size_t mask = 0; mask < step; mask++) {
      kx v = w[mask];
      w[mask] = or_k_mul(v, om);
      w[mask + step] = or_k_mul(v, bits[i]);
    }
  }
}

static kx eq_mask(size_t mask, const kx* pts, size_t n) {
  kx p = K1;
  for (size_t b = 0; b < n; b++) p = or_k_mul(p, ((mask >> b) & 1) ? pts[b] : or_k_sub(K1, pts[b]));
  return p;
}

struct or_rowstate {
  size_t cur_len, n_pad, k_total, m, nf;
  kx* eqb;
  kx* eqr;       /* may be NULL */
  kx* fv[8];     /* f-var tables */
  int fv_j[8];   /* CCS matrix index of each f-var table */
  kx* nc;        /* [w][row][rho], row < cur_len */
  kx* ev;        /* may be NULL */
  kx* wg;        /* k_total * 54 */
  kx gamma_k;
  /* compiled f terms */
  size_t nterms;
  kx tcoeff[16];
  int tnv[16];
  int tvar[16][8];
  int texp[16][8];
  int deg_max;
};

static void spmv_csc(const or_mat* M, const kx* x, size_t n_eff, kx* y /* len n_pad, zeroed */) {
  for (size_t c = 0; c < M->m; c++) {
    kx xc = x[c];
    if (or_k_eq(xc, K0)) continue;
    for (size_t k = M->col_ptr[c]; k < M->col_ptr[c + 1]; k++) {
      size_t r = M->row_idx[k];
      if (r < n_eff) y[r] = or_k_add(y[r], or_k_scale(xc, M->vals[k]));
    }
  }
}

or_rowstate* or_rowstate_build(const or_ccs* s, const int8_t* const* Zs, size_t k_total, const kx* alpha,
                               const kx* beta_a, const kx* beta_r, kx gamma, size_t ell_d, size_t ell_n,
                               const kx* r_inputs) {
  or_rowstate* st = (or_rowstate*)calloc(1, sizeof(or_rowstate));
  size_t n_pad = (size_t)1 << ell_n, n_eff = s->n, m = s->m, D = NEO_D;
  st->cur_len = n_pad;
  st->n_pad = n_pad;
  st->k_total = k_total;
  st->m = m;
  st->eqb = (kx*)malloc(n_pad * sizeof(kx));
  or_chi_table(beta_r, ell_n, st->eqb);
  if (r_inputs) {
    st->eqr = (kx*)malloc(n_pad * sizeof(kx));
    or_chi_table(r_inputs, ell_n, st->eqr);
  }
  /* compile f (oracle.rs:127-186) */
  int used[8] = {0};
  for (size_t tt = 0; tt < s->nterms; tt++)
    for (size_t j = 0; j < s->t; j++)
      if (s->term_exps[tt][j]) used[j] = 1;
  int pos_by_j[8];
  st->nf = 0;
  for (size_t j = 0; j < s->t; j++) {
    pos_by_j[j] = -1;
    if (used[j]) {
      pos_by_j[j] = (int)st->nf;
      st->fv_j[st->nf++] = (int)j;
    }
  }
  int fdeg = 0;
  st->nterms = s->nterms;
  for (size_t tt = 0; tt < s->nterms; tt++) {
    st->tcoeff[tt] = or_k_from(s->term_coeff[tt]);
    st->tnv[tt] = 0;
    int d = 0;
    for (size_t j = 0; j < s->t; j++)
      if (s->term_exps[tt][j]) {
        st->tvar[tt][st->tnv[tt]] = pos_by_j[j];
        st->texp[tt][st->tnv[tt]] = (int)s->term_exps[tt][j];
        st->tnv[tt]++;
        d += (int)s->term_exps[tt][j];
      }
    if (d > fdeg) fdeg = d;
  }
  st->deg_max = fdeg + 1 > 4 ? fdeg + 1 : 4;
  /* w_gamma_nc (oracle.rs:199-209) */
  st->wg = (kx*)malloc(k_total * D * sizeof(kx));
  {
    kx g = gamma;
    for (size_t i = 0; i < k_total; i++) {
      for (size_t rho = 0; rho < D; rho++) st->wg[i * D + rho] = or_k_mul(eq_mask(rho, beta_a, ell_d), g);
      g = or_k_mul(g, gamma);
    }
  }
  /* z1 = sum_rho 2^rho Z1[rho, :] */
  kx* z1 = (kx*)calloc(m, sizeof(kx));
  {
    fq pw = 1;
    for (size_t rho = 0; rho < D; rho++) {
      for (size_t c = 0; c < m; c++) {
        int8_t dg = Zs[0][rho * m + c];
        if (dg == 1) z1[c].c0 = or_fq_add(z1[c].c0, pw);
        else if (dg == -1) z1[c].c0 = or_fq_sub(z1[c].c0, pw);
      }
      pw = or_fq_add(pw, pw);
    }
  }
  for (size_t f = 0; f < st->nf; f++) {
    st->fv[f] = (kx*)calloc(n_pad, sizeof(kx));
    const or_mat* M = &s->mats[st->fv_j[f]];
    if (M->identity) {
      size_t cap = n_eff < m ? n_eff : m;
      memcpy(st->fv[f], z1, cap * sizeof(kx));
    } else {
      spmv_csc(M, z1, n_eff, st->fv[f]);
    }
  }
  /* NC tables */
  st->nc = (kx*)calloc(k_total * n_pad * D, sizeof(kx));
  {
    size_t cap = n_eff < n_pad ? n_eff : n_pad;
    for (size_t w = 0; w < k_total; w++)
      for (size_t row = 0; row < cap; row++)
        for (size_t rho = 0; rho < D; rho++)
          st->nc[(w * n_pad + row) * D + rho] = or_k_from(or_fq_from_i64(Zs[w][rho * m + row]));
  }
  /* gamma^k and Eval table (oracle.rs:304-390) */
  st->gamma_k = K1;
  for (size_t i = 0; i < k_total; i++) st->gamma_k = or_k_mul(st->gamma_k, gamma);
  if (k_total >= 2 && r_inputs) {
    kx w_alpha[NEO_D];
    for (size_t rho = 0; rho < D; rho++) w_alpha[rho] = eq_mask(rho, alpha, ell_d);
    st->ev = (kx*)calloc(n_pad, sizeof(kx));
    kx* s_alpha = (kx*)malloc(m * sizeof(kx));
    kx* tmp = (kx*)malloc(n_pad * sizeof(kx));
    kx gpi = K1;
    for (size_t i = 1; i < k_total; i++) {
      gpi = or_k_mul(gpi, gamma);
      for (size_t c = 0; c < m; c++) s_alpha[c] = K0;
      for (size_t rho = 0; rho < D; rho++)
        for (size_t c = 0; c < m; c++) {
          int8_t dg = Zs[i][rho * m + c];
          if (dg == 1) s_alpha[c] = or_k_add(s_alpha[c], w_alpha[rho]);
          else if (dg == -1) s_alpha[c] = or_k_sub(s_alpha[c], w_alpha[rho]);
        }
      kx gkj = K1;
      for (size_t j = 0; j < s->t; j++) {
        kx coeff = or_k_mul(gpi, gkj);
        const or_mat* M = &s->mats[j];
        if (M->identity) {
          size_t cap = n_eff < m ? n_eff : m;
          for (size_t r = 0; r < cap; r++) st->ev[r] = or_k_add(st->ev[r], or_k_mul(coeff, s_alpha[r]));
        } else {
          for (size_t r = 0; r < n_pad; r++) tmp[r] = K0;
          spmv_csc(M, s_alpha, n_eff, tmp);
          for (size_t r = 0; r < n_pad; r++) st->ev[r] = or_k_add(st->ev[r], or_k_mul(coeff, tmp[r]));
        }
        gkj = or_k_mul(gkj, st->gamma_k);
      }
    }
    free(s_alpha);
    free(tmp);
  }
  free(z1);
  return st;
}

void or_rowstate_free(or_rowstate* st) {
  if (!st) return;
  free(st->eqb);
  free(st->eqr);
  for (size_t f = 0; f < st->nf; f++) free(st->fv[f]);
  free(st->nc);
  free(st->ev);
  free(st->wg);
  free(st);
}

size_t or_rowstate_len(const or_rowstate* st) { return st->cur_len; }

static void poly_mul_affine(kx* poly, int n, kx a, kx b) {
  kx prev = K0;
  for (int i = 0; i < n; i++) {
    kx old = poly[i];
    poly[i] = or_k_add(or_k_mul(a, old), or_k_mul(b, prev));
    prev = old;
  }
}

size_t or_rowstate_round_coeffs(const or_rowstate* st, kx* coeffs) {
  int dm = st->deg_max;
  size_t D = NEO_D, tail = st->cur_len / 2;
  const kx three = {3, 0};
  for (int d = 0; d <= dm; d++) coeffs[d] = K0;
  kx inner[16], term[16];
  for (size_t t = 0; t < tail; t++) {
    kx e0 = st->eqb[2 * t], e1 = or_k_sub(st->eqb[2 * t + 1], e0);
    for (int d = 0; d <= dm; d++) inner[d] = K0;
    for (size_t tt = 0; tt < st->nterms; tt++) {
      for (int d = 0; d <= dm; d++) term[d] = K0;
      term[0] = st->tcoeff[tt];
      for (int v = 0; v < st->tnv[tt]; v++) {
        const kx* tbl = st->fv[st->tvar[tt][v]];
        kx a = tbl[2 * t], b = or_k_sub(tbl[2 * t + 1], a);
        for (int e = 0; e < st->texp[tt][v]; e++) poly_mul_affine(term, dm + 1, a, b);
      }
      for (int d = 0; d <= dm; d++) inner[d] = or_k_add(inner[d], term[d]);
    }
    for (size_t w = 0; w < st->k_total; w++) {
      const kx* lo = st->nc + (w * st->n_pad + 2 * t) * D;
      const kx* hi = lo + D;
      for (size_t rho = 0; rho < D; rho++) {
        kx wg = st->wg[w * D + rho];
        kx a = lo[rho], b = or_k_sub(hi[rho], a);
        kx a2 = or_k_mul(a, a), a3 = or_k_mul(a2, a), b2 = or_k_mul(b, b), b3 = or_k_mul(b2, b);
        kx t0 = or_k_sub(a3, a);
        kx t1 = or_k_sub(or_k_mul(or_k_mul(a2, b), three), b);
        kx t2 = or_k_mul(or_k_mul(a, b2), three);
        inner[0] = or_k_add(inner[0], or_k_mul(wg, t0));
        inner[1] = or_k_add(inner[1], or_k_mul(wg, t1));
        inner[2] = or_k_add(inner[2], or_k_mul(wg, t2));
        inner[3] = or_k_add(inner[3], or_k_mul(wg, b3));
      }
    }
    coeffs[0] = or_k_add(coeffs[0], or_k_mul(e0, inner[0]));
    for (int d = 1; d <= dm; d++)
      coeffs[d] = or_k_add(coeffs[d], or_k_add(or_k_mul(e0, inner[d]), or_k_mul(e1, inner[d - 1])));
    if (st->eqr && st->ev) {
      kx r0 = st->eqr[2 * t], r1 = or_k_sub(st->eqr[2 * t + 1], r0);
      kx v0 = st->ev[2 * t], v1 = or_k_sub(st->ev[2 * t + 1], v0);
      kx g = st->gamma_k;
      coeffs[0] = or_k_add(coeffs[0], or_k_mul(g, or_k_mul(r0, v0)));
      coeffs[1] = or_k_add(coeffs[1], or_k_mul(g, or_k_add(or_k_mul(r0, v1), or_k_mul(r1, v0))));
      coeffs[2] = or_k_add(coeffs[2], or_k_mul(g, or_k_mul(r1, v1)));
    }
  }
  return (size_t)dm + 1;
}

kx or_poly_eval(const kx* c, size_t n, kx x) {
  if (n == 0) return K0;
  kx r = c[n - 1];
  for (size_t i = n - 1; i-- > 0;) r = or_k_add(or_k_mul(r, x), c[i]);
  return r;
}

void or_rowstate_evals(const or_rowstate* st, const kx* xs, size_t nx, kx* out) {
  kx coeffs[16];
  size_t n = or_rowstate_round_coeffs(st, coeffs);
  for (size_t i = 0; i < nx; i++) out[i] = or_poly_eval(coeffs, n, xs[i]);
}

static void fold_tbl(kx* t, size_t len, kx r) {
  for (size_t i = 0; i < len / 2; i++) t[i] = or_k_add(t[2 * i], or_k_mul(or_k_sub(t[2 * i + 1], t[2 * i]), r));
}

void or_rowstate_fold(or_rowstate* st, kx r) {
  size_t L = st->cur_len, D = NEO_D;
  fold_tbl(st->eqb, L, r);
  if (st->eqr) fold_tbl(st->eqr, L, r);
  for (size_t f = 0; f < st->nf; f++) fold_tbl(st->fv[f], L, r);
  for (size_t w = 0; w < st->k_total; w++) {
    kx* base = st->nc + w * st->n_pad * D;
    for (size_t i = 0; i < L / 2; i++)
      for (size_t rho = 0; rho < D; rho++) {
        kx lo = base[(2 * i) * D + rho], hi = base[(2 * i + 1) * D + rho];
        base[i * D + rho] = or_k_add(lo, or_k_mul(or_k_sub(hi, lo), r));
      }
  }
  if (st->ev) fold_tbl(st->ev, L, r);
  st->cur_len = L / 2;
}

size_t or_rowstate_export(const or_rowstate* st, int which, size_t idx, kx* out, size_t cap) {
  size_t L = st->cur_len, D = NEO_D;
  const kx* src = NULL;
  if (which == 0) src = st->eqb;
  else if (which == 1) src = st->eqr;
  else if (which == 2) src = st->ev;
  else if (which == 3 && idx < st->nf) src = st->fv[idx];
  else if (which == 4 && idx < st->k_total) {
    /* NC witness idx as [rho][row] */
    size_t n = L * D < cap ? L * D : cap;
    for (size_t rho = 0; rho < D; rho++)
      for (size_t row = 0; row < L; row++)
        if (rho * L + row < n) out[rho * L + row] = st->nc[(idx * st->n_pad + row) * D + rho];
    return n;
  }
  if (!src) return 0;
  size_t n = L < cap ? L : cap;
  memcpy(out, src, n * sizeof(kx));
  return n;
}

/* ---------------- precompute_for_r ---------------- */
void or_vjs(const or_ccs* s, const kx* r, size_t ell_n, kx* vjs) {
  size_t n_sz = (size_t)1 << ell_n, m = s->m;
  size_t n_eff = s->n < n_sz ? s->n : n_sz;
  kx* chi = (kx*)malloc(n_sz * sizeof(kx));
  or_chi_table(r, ell_n, chi);
  for (size_t j = 0; j < s->t; j++) {
    kx* v = vjs + j * m;
    for (size_t c = 0; c < m; c++) v[c] = K0;
    const or_mat* M = &s->mats[j];
    if (M->identity) {
      size_t cap = m < n_eff ? m : n_eff;
      memcpy(v, chi, cap * sizeof(kx));
    } else {
      for (size_t c = 0; c < M->m; c++) {
        kx acc = K0;
        for (size_t k = M->col_ptr[c]; k < M->col_ptr[c + 1]; k++) {
          size_t row = M->row_idx[k];
          if (row < n_eff) acc = or_k_add(acc, or_k_scale(chi[row], M->vals[k]));
        }
        v[c] = acc;
      }
    }
  }
  free(chi);
}

void or_precompute_y(const or_ccs* s, const int8_t* const* Zs, size_t k_total, const kx* r, size_t ell_n,
                     kx* y_eval) {
  or_precompute_y_par(s, Zs, k_total, r, ell_n, y_eval, 1);
}

/* parallel = 0 runs on one thread, as the reference's precompute_for_r does (oracle.rs:1674-1677);
 * the DEC children's y products are rayon-parallel there (common.rs:1547) and pass 1. */
void or_precompute_y_par(const or_ccs* s, const int8_t* const* Zs, size_t k_total, const kx* r, size_t ell_n,
                         kx* y_eval, int parallel) {
  size_t m = s->m, t = s->t, D = NEO_D;
  kx* vjs = (kx*)malloc(t * m * sizeof(kx));
  or_vjs(s, r, ell_n, vjs);
#pragma omp parallel for collapse(2) schedule(static) if (parallel)
  for (long i = 0; i < (long)k_total; i++)
    for (long j = 0; j < (long)t; j++) {
      const kx* v = vjs + (size_t)j * m;
      for (size_t rho = 0; rho < D; rho++) {
        kx acc = K0;
        const int8_t* zr = Zs[i] + rho * m;
        for (size_t c = 0; c < m; c++) {
          if (zr[c] == 1) acc = or_k_add(acc, v[c]);
          else if (zr[c] == -1) acc = or_k_sub(acc, v[c]);
        }
        y_eval[((size_t)i * t + (size_t)j) * D + rho] = acc;
      }
    }
  free(vjs);
}

/* ---------------- sum-check helpers ---------------- */
void or_interpolate(const kx* xs, const kx* ys, size_t n, kx* coeffs) {
  for (size_t i = 0; i < n; i++) coeffs[i] = K0;
  kx numer[32], next[32];
  for (size_t i = 0; i < n; i++) {
    for (size_t d = 0; d < n; d++) numer[d] = K0;
    numer[0] = K1;
    size_t cur = 0;
    for (size_t j = 0; j < n; j++) {
      if (i == j) continue;
      for (size_t d = 0; d < n; d++) next[d] = K0;
      for (size_t d = 0; d <= cur; d++) {
        next[d + 1] = or_k_add(next[d + 1], numer[d]);
        next[d] = or_k_add(next[d], or_k_mul(or_k_neg(xs[j]), numer[d]));
      }
      memcpy(numer, next, n * sizeof(kx));
      cur++;
    }
    kx denom = K1;
    for (size_t j = 0; j < n; j++)
      if (i != j) denom = or_k_mul(denom, or_k_sub(xs[i], xs[j]));
    kx scale = or_k_mul(ys[i], or_k_inv(denom));
    for (size_t d = 0; d <= cur; d++) coeffs[d] = or_k_add(coeffs[d], or_k_mul(scale, numer[d]));
  }
}
