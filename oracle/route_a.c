/*
 * route_a.c -- CPU restatement of the Route A memory sidecar of a shard fold step (TEST
 * INFRASTRUCTURE ONLY: the checker the MI355X library is compared against; the product never
 * loads it).
 *
 * Follows, function by function:
 *   neo-fold/src/memory_sidecar/memory.rs
 *     32-136     absorb_step_memory (public LUT / MEM parameters, table and init digests)
 *     232-428    prove_twist_addr_pre_time (read/write address lanes, batched per instance)
 *     430-621    prove_shout_addr_pre_time (address lookup lanes, one batch over active lanes)
 *     905-1025   build_route_a_memory_oracles (value, adapter, bitness; read/write check, bitness)
 *     1284-1670  finalize_route_a_memory_prover (val-eval batch, CPU ME claims at r_val)
 *   neo-fold/src/memory_sidecar/route_a_time.rs:28-152  claim order of the batched time sum-check
 *   neo-fold/src/memory_sidecar/transcript.rs:4-64      digest_fields, claim-sum binding
 *   neo-fold/src/memory_sidecar/utils.rs:43-72          bitness weights
 *   neo-fold/src/memory_sidecar/sumcheck_ds.rs:11-63    domain-separated sum-check wrappers
 *   neo-reductions/src/sumcheck.rs:350-456              run_batched_sumcheck_prover
 *   neo-fold/src/memory_sidecar/cpu_bus.rs:90-181, 969-1016  bus layout, bus time columns
 *   neo-memory/src/mem_init.rs:133-165                 eval_init_at_r_addr
 *   neo-memory/src/ts_common.rs:22-37                  sample_ext_point
 *
 * Pinned by: every Twist/Shout oracle is pinned against brute-force hypercube sums
 * (tests/test_oracle_twist.py); the composition here is pinned by the sum-check invariants the
 * prover checks on every claim in every round (a claimed sum that does not equal the oracle's
 * hypercube sum fails round 0), by the read/write claims of the time lanes equalling the final
 * values of the address lanes, and by tests/test_oracle_route_a.py (a corrupted trace is
 * rejected; the obligations' recompositions hold).  Transcript bytes: parity unpinned (the
 * transcript's Poseidon2 constants).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "neo_oracle.h"
#include "route_a.h"

static const kx K0 = {0, 0};
static const kx K1 = {1, 0};

void ra_msg_u64(or_tr* t, const char* label, uint64_t v) {
  uint8_t b[8];
  for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
  or_tr_append_message(t, label, b, 8);
}
static void msg_str(or_tr* t, const char* label, const char* s) {
  or_tr_append_message(t, label, (const uint8_t*)s, strlen(s));
}
static void tr_k(or_tr* t, const char* label, kx x) {
  fq f[2] = {x.c0, x.c1};
  or_tr_append_fields(t, label, f, 2);
}

/* ---------------- bus layout (cpu_bus.rs:73-168, bus_layout.rs:98-191) ---------------- */
int ra_infer_layout(const or_step_mem* mem, size_t nsteps, size_t n, size_t m, size_t m_in, ra_layout* L,
                    char* err, size_t errlen) {
  memset(L, 0, sizeof(*L));
  L->m = m;
  L->m_in = m_in;
  size_t chunk = 0;
  for (size_t s = 0; s < nsteps; s++) {
    for (size_t i = 0; i < mem[s].n_lut; i++)
      if (mem[s].luts[i].steps > chunk) chunk = mem[s].luts[i].steps;
    for (size_t i = 0; i < mem[s].n_mem; i++)
      if (mem[s].mems[i].steps > chunk) chunk = mem[s].mems[i].steps;
  }
  if (chunk == 0) chunk = 1;
  L->chunk = chunk;
  if (nsteps == 0) return 0;
  L->n_shout = mem[0].n_lut;
  L->n_twist = mem[0].n_mem;
  if (L->n_shout > RA_MAX_INST || L->n_twist > RA_MAX_INST) {
    snprintf(err, errlen, "too many memory instances");
    return -1;
  }
  for (size_t s = 1; s < nsteps; s++) {
    int bad = mem[s].n_lut != L->n_shout || mem[s].n_mem != L->n_twist;
    for (size_t i = 0; !bad && i < L->n_shout; i++)
      bad = mem[s].luts[i].d * mem[s].luts[i].ell != mem[0].luts[i].d * mem[0].luts[i].ell;
    for (size_t i = 0; !bad && i < L->n_twist; i++)
      bad = mem[s].mems[i].d * mem[s].mems[i].ell != mem[0].mems[i].d * mem[0].mems[i].ell;
    if (bad) {
      snprintf(err, errlen, "shared CPU bus layout mismatch across steps (step 0 vs step %zu)", s);
      return -1;
    }
  }
  size_t col = 0;
  for (size_t i = 0; i < L->n_shout; i++) {
    size_t e = mem[0].luts[i].d * mem[0].luts[i].ell;
    L->sh[i].addr0 = col;
    L->sh[i].ell = e;
    L->sh[i].has_lookup = col + e;
    L->sh[i].val = col + e + 1;
    col += e + 2;
  }
  for (size_t i = 0; i < L->n_twist; i++) {
    size_t e = mem[0].mems[i].d * mem[0].mems[i].ell;
    L->tw[i].ra0 = col;
    L->tw[i].wa0 = col + e;
    L->tw[i].ell = e;
    L->tw[i].has_read = col + 2 * e;
    L->tw[i].has_write = col + 2 * e + 1;
    L->tw[i].wv = col + 2 * e + 2;
    L->tw[i].rv = col + 2 * e + 3;
    L->tw[i].inc = col + 2 * e + 4;
    col += 2 * e + 5;
  }
  L->bus_cols = col;
  if (col * chunk > m) {
    snprintf(err, errlen, "BusLayout: bus region too large: bus_cols(%zu) * chunk_size(%zu) > m(%zu)", col, chunk, m);
    return -1;
  }
  L->bus_base = m - col * chunk;
  if (L->bus_base < m_in) {
    snprintf(err, errlen, "BusLayout: bus_base overlaps public inputs");
    return -1;
  }
  if (col > 0 && m_in + chunk > n) {
    snprintf(err, errlen, "bus time rows out of range: m_in + chunk_size > n");
    return -1;
  }
  return 0;
}

/* ---------------- sparse bus columns ---------------- */
void ra_col_free(ra_col* c) {
  free(c->idx);
  free(c->vals);
  memset(c, 0, sizeof(*c));
}
or_sparse_in ra_in(const ra_col* c) {
  or_sparse_in s = {c->nnz, c->idx, (const uint64_t*)c->vals};
  return s;
}
static int bus_col(const fq* z, const ra_layout* L, size_t col, size_t steps_len, size_t pow2, ra_col* out) {
  out->idx = (uint64_t*)malloc((L->chunk + 1) * sizeof(uint64_t));
  out->vals = (kx*)malloc((L->chunk + 1) * sizeof(kx));
  long n = or_bus_time_col(z, L->m, L->m_in, L->bus_base, L->chunk, L->bus_cols, col, steps_len, pow2, out->idx,
                           out->vals);
  if (n < 0) return -1;
  out->nnz = (size_t)n;
  return 0;
}

int ra_decode_twist(const fq* z, const ra_layout* L, size_t i, size_t steps, size_t ell_n, ra_twist_cols* T) {
  const size_t pow2 = (size_t)1 << ell_n;
  const ra_twist_lay* tl = &L->tw[i];
  memset(T, 0, sizeof(*T));
  T->ell = tl->ell;
  for (size_t b = 0; b < tl->ell; b++) {
    if (bus_col(z, L, tl->ra0 + b, steps, pow2, &T->ra[b])) return -1;
    if (bus_col(z, L, tl->wa0 + b, steps, pow2, &T->wa[b])) return -1;
  }
  if (bus_col(z, L, tl->has_read, steps, pow2, &T->has_read) || bus_col(z, L, tl->has_write, steps, pow2, &T->has_write) ||
      bus_col(z, L, tl->wv, steps, pow2, &T->wv) || bus_col(z, L, tl->rv, steps, pow2, &T->rv) ||
      bus_col(z, L, tl->inc, steps, pow2, &T->inc))
    return -1;
  return 0;
}
void ra_twist_cols_free(ra_twist_cols* T) {
  for (size_t b = 0; b < T->ell; b++) {
    ra_col_free(&T->ra[b]);
    ra_col_free(&T->wa[b]);
  }
  ra_col_free(&T->has_read);
  ra_col_free(&T->has_write);
  ra_col_free(&T->wv);
  ra_col_free(&T->rv);
  ra_col_free(&T->inc);
}
static int decode_shout(const fq* z, const ra_layout* L, size_t i, size_t steps, size_t ell_n, ra_shout_cols* S) {
  const size_t pow2 = (size_t)1 << ell_n;
  memset(S, 0, sizeof(*S));
  S->ell = L->sh[i].ell;
  for (size_t b = 0; b < S->ell; b++)
    if (bus_col(z, L, L->sh[i].addr0 + b, steps, pow2, &S->addr[b])) return -1;
  if (bus_col(z, L, L->sh[i].has_lookup, steps, pow2, &S->has_lookup) || bus_col(z, L, L->sh[i].val, steps, pow2, &S->val))
    return -1;
  return 0;
}
static void shout_cols_free(ra_shout_cols* S) {
  for (size_t b = 0; b < S->ell; b++) ra_col_free(&S->addr[b]);
  ra_col_free(&S->has_lookup);
  ra_col_free(&S->val);
}

/* ---------------- transcript pieces ---------------- */
/* digest_fields (memory_sidecar/transcript.rs:4-10) */
static void digest_fields(const char* label, const fq* fs, size_t n, uint8_t out[32]) {
  or_tr h;
  or_tr_new(&h, (const uint8_t*)"memory/public_digest", 20);
  msg_str(&h, "digest/label", label);
  ra_msg_u64(&h, "digest/len", n);
  or_tr_append_fields(&h, "digest/fields", fs, n);
  or_tr_digest32(&h, out);
}

/* absorb_step_memory_impl (memory.rs:82-124), no table_spec */
void ra_absorb_step_memory(or_tr* tr, const or_step_mem* sm) {
  size_t nl = sm ? sm->n_lut : 0, nm = sm ? sm->n_mem : 0;
  or_tr_append_message(tr, "step/absorb_memory_start", NULL, 0);
  ra_msg_u64(tr, "step/lut_count", nl);
  for (size_t i = 0; i < nl; i++) {
    const or_lut_inst* in = &sm->luts[i];
    ra_msg_u64(tr, "step/lut_idx", i);
    ra_msg_u64(tr, "shout/k", in->k);
    ra_msg_u64(tr, "shout/d", in->d);
    ra_msg_u64(tr, "shout/n_side", in->n_side);
    ra_msg_u64(tr, "shout/steps", in->steps);
    ra_msg_u64(tr, "shout/ell", in->ell);
    uint8_t dg[32];
    digest_fields("shout/table", in->table, in->k, dg);
    or_tr_append_message(tr, "shout/table_digest", dg, 32);
  }
  ra_msg_u64(tr, "step/mem_count", nm);
  for (size_t i = 0; i < nm; i++) {
    const or_mem_inst* in = &sm->mems[i];
    ra_msg_u64(tr, "step/mem_idx", i);
    ra_msg_u64(tr, "twist/k", in->k);
    ra_msg_u64(tr, "twist/d", in->d);
    ra_msg_u64(tr, "twist/n_side", in->n_side);
    ra_msg_u64(tr, "twist/steps", in->steps);
    ra_msg_u64(tr, "twist/ell", in->ell);
    uint8_t dg[32];
    if (!in->sparse) {
      digest_fields("twist/init/zero", NULL, 0, dg);
    } else {
      fq* fs = (fq*)malloc((2 * in->ninit + 1) * sizeof(fq));
      for (size_t q = 0; q < in->ninit; q++) {
        fs[2 * q] = in->init_addrs[q] % NEO_Q;
        fs[2 * q + 1] = in->init_vals[q];
      }
      digest_fields("twist/init/sparse", fs, 2 * in->ninit, dg);
      free(fs);
    }
    or_tr_append_message(tr, "twist/init_digest", dg, 32);
  }
  or_tr_append_message(tr, "step/absorb_memory_done", NULL, 0);
}

/* bind_batched_claim_sums (transcript.rs:12-25) */
static void bind_claim_sums(or_tr* tr, const char* prefix, const kx* sums, const char* const* labels, size_t n) {
  ra_msg_u64(tr, prefix, n);
  for (size_t i = 0; i < n; i++) {
    msg_str(tr, "addr_batch/label", labels[i]);
    ra_msg_u64(tr, "addr_batch/idx", i);
    tr_k(tr, "addr_batch/claimed_sum", sums[i]);
  }
}

/* sample_ext_point (ts_common.rs:22-37) */
static void sample_ext_point(or_tr* tr, const char* label, const char* l0, const char* l1, size_t len, kx* out) {
  for (size_t i = 0; i < len; i++) {
    ra_msg_u64(tr, label, i);
    out[i].c0 = or_tr_challenge_field(tr, l0);
    out[i].c1 = or_tr_challenge_field(tr, l1);
  }
}

/* bitness_weights (utils.rs:43-72) */
static void bitness_weights(const kx* r_cycle, size_t ell_n, size_t n, uint64_t sep, kx* out) {
  kx base = or_k_from(sep % NEO_Q);
  if (ell_n) base = or_k_add(r_cycle[0], base);
  const kx neg1 = {NEO_Q - 1, 0};
  if (or_k_eq(base, K0) || or_k_eq(base, K1) || or_k_eq(base, neg1)) base = or_k_add(base, K1);
  if (or_k_eq(base, K0)) base = K1;
  kx cur = K1;
  for (size_t i = 0; i < n; i++) {
    out[i] = cur;
    cur = or_k_mul(cur, base);
  }
}

/* eval_init_at_r_addr (mem_init.rs:133-165) */
static kx eval_init(const or_mem_inst* in, const kx* r_addr, size_t ell) {
  if (!in->sparse) return K0;
  kx acc = K0;
  for (size_t q = 0; q < in->ninit; q++) {
    kx chi = K1;
    for (size_t b = 0; b < ell; b++) chi = or_k_mul(chi, ((in->init_addrs[q] >> b) & 1) ? r_addr[b] : or_k_sub(K1, r_addr[b]));
    acc = or_k_add(acc, or_k_scale(chi, in->init_vals[q]));
  }
  return acc;
}

/* ---------------- batched sum-check (sumcheck.rs:350-456 with sumcheck_ds.rs:51-63) ---------------- */
void ra_claim_evals(ra_claim* c, const kx* xs, size_t n, kx* ys) {
  if (c->kind == RA_ROWSTATE) or_rowstate_evals((or_rowstate*)c->o, xs, n, ys);
  else or_ts_evals((or_ts*)c->o, xs, n, ys);
}
static void claim_fold(ra_claim* c, kx r) {
  if (c->kind == RA_ROWSTATE) or_rowstate_fold((or_rowstate*)c->o, r);
  else or_ts_fold((or_ts*)c->o, r);
}

int ra_batched_sumcheck(or_tr* tr, const char* domain, size_t inst_idx, ra_claim* cl, size_t n, size_t rounds,
                        kx* chals, char* err, size_t errlen) {
  msg_str(tr, "sc/domain", domain);
  ra_msg_u64(tr, "sc/inst_idx", inst_idx);
  for (size_t i = 0; i < n; i++) {
    cl[i].running = cl[i].claimed_sum;
    cl[i].nrounds = rounds;
    cl[i].polys = (kx*)calloc(rounds * (cl[i].deg + 1) + 1, sizeof(kx));
  }
  kx xs[64], ys[64];
  for (size_t rd = 0; rd < rounds; rd++) {
    ra_msg_u64(tr, "batched/round_idx", rd);
    for (size_t i = 0; i < n; i++) {
      const size_t np = cl[i].deg + 1;
      for (size_t x = 0; x < np; x++) xs[x] = or_k_from(x);
      ra_claim_evals(&cl[i], xs, np, ys);
      if (!or_k_eq(or_k_add(ys[0], ys[1]), cl[i].running)) {
        snprintf(err, errlen, "%s: claim %zu (%s) round %zu invariant failed", domain, i, cl[i].label, rd);
        return -1;
      }
      or_interpolate(xs, ys, np, cl[i].polys + rd * np);
    }
    for (size_t i = 0; i < n; i++) {
      const size_t np = cl[i].deg + 1;
      msg_str(tr, "batched/claim_label", cl[i].label);
      ra_msg_u64(tr, "batched/claim_idx", i);
      for (size_t x = 0; x < np; x++) tr_k(tr, "batched/round/coeff", cl[i].polys[rd * np + x]);
    }
    kx r;
    r.c0 = or_tr_challenge_field(tr, "batched/challenge/0");
    r.c1 = or_tr_challenge_field(tr, "batched/challenge/1");
    chals[rd] = r;
    for (size_t i = 0; i < n; i++) {
      const size_t np = cl[i].deg + 1;
      cl[i].running = or_poly_eval(cl[i].polys + rd * np, np, r);
      claim_fold(&cl[i], r);
    }
  }
  msg_str(tr, "sc/domain_end", domain);
  ra_msg_u64(tr, "sc/inst_idx_end", inst_idx);
  return 0;
}

void ra_claim_free(ra_claim* c) {
  free(c->polys);
  c->polys = NULL;
}

/* ---------------- addr-pre sum-checks ---------------- */
int ra_shout_addr_pre(or_tr* tr, const or_step_mem* sm, const ra_layout* L, const fq* zdec, size_t ell_n,
                      const kx* r_cycle, size_t step_idx, ra_shout_pre* out, char* err, size_t errlen) {
  memset(out, 0, sizeof(*out));
  const size_t nl = sm->n_lut;
  if (nl == 0) return 0;
  if (nl > 64) {
    snprintf(err, errlen, "Shout(Route A): skip mask supports up to 64 Shout instances");
    return -1;
  }
  out->n_lut = nl;
  out->cols = (ra_shout_cols*)calloc(nl, sizeof(ra_shout_cols));
  out->claimed = (kx*)calloc(nl, sizeof(kx));
  ra_claim* cl = (ra_claim*)calloc(nl, sizeof(ra_claim));
  size_t ell_addr = 0, nact = 0;
  for (size_t i = 0; i < nl; i++) {
    const or_lut_inst* in = &sm->luts[i];
    const size_t e = in->d * in->ell;
    if (i == 0) ell_addr = e;
    else if (e != ell_addr) {
      snprintf(err, errlen, "Shout(Route A): batched addr-pre requires uniform ell_addr");
      return -1;
    }
    if (in->steps > ((size_t)1 << ell_n)) {
      snprintf(err, errlen, "Shout(Route A): steps exceeds 2^ell_cycle");
      return -1;
    }
    if (decode_shout(zdec, L, i, in->steps, ell_n, &out->cols[i])) {
      snprintf(err, errlen, "bus column decode failed (lut %zu)", i);
      return -1;
    }
    const ra_shout_cols* S = &out->cols[i];
    int any = 0;
    for (size_t q = 0; q < S->has_lookup.nnz; q++)
      if (!or_k_eq(S->has_lookup.vals[q], K0)) any = 1;
    if (any) {
      or_sparse_in bits[64];
      for (size_t b = 0; b < e; b++) bits[b] = ra_in(&S->addr[b]);
      kx* tab = (kx*)malloc(in->k * sizeof(kx));
      for (size_t q = 0; q < in->k; q++) tab[q] = or_k_from(in->table[q]);
      kx claim;
      or_ts* o = or_ts_addr_lookup(bits, e, ra_in(&S->has_lookup), tab, in->k, r_cycle, ell_n, &claim);
      free(tab);
      out->claimed[i] = claim;
      out->active_mask |= (uint64_t)1 << i;
      cl[nact].kind = RA_TS;
      cl[nact].o = o;
      cl[nact].claimed_sum = claim;
      cl[nact].deg = or_ts_degree_bound(o);
      cl[nact].label = "shout/addr_pre";
      nact++;
    }
  }
  out->ell_addr = ell_addr;
  ra_msg_u64(tr, "shout/addr_pre_time/step_idx", step_idx);
  const char** labels = (const char**)malloc(nl * sizeof(char*));
  for (size_t i = 0; i < nl; i++) labels[i] = "shout/addr_pre";
  bind_claim_sums(tr, "shout/addr_pre_time/claimed_sums", out->claimed, labels, nl);
  free(labels);
  out->r_addr = (kx*)calloc(ell_addr + 1, sizeof(kx));
  int rc = 0;
  if (nact == 0) {
    ra_msg_u64(tr, "shout/addr_pre_time/no_sumcheck", step_idx);
    sample_ext_point(tr, "shout/addr_pre_time/no_sumcheck/r_addr", "shout/addr_pre_time/no_sumcheck/r_addr/0",
                     "shout/addr_pre_time/no_sumcheck/r_addr/1", ell_addr, out->r_addr);
  } else {
    for (size_t i = 0; i < nact; i++)
      if (or_ts_num_rounds((or_ts*)cl[i].o) != ell_addr) {
        snprintf(err, errlen, "shout addr-pre: oracle rounds != ell_addr");
        rc = -1;
      }
    if (!rc) rc = ra_batched_sumcheck(tr, "shout/addr_pre_time", step_idx, cl, nact, ell_addr, out->r_addr, err, errlen);
  }
  out->n_active = nact;
  out->rounds = (kx**)calloc(nact + 1, sizeof(kx*));
  out->degs = (size_t*)calloc(nact + 1, sizeof(size_t));
  for (size_t i = 0; i < nact; i++) {
    out->rounds[i] = cl[i].polys;
    out->degs[i] = cl[i].deg;
    cl[i].polys = NULL;
    or_ts_free((or_ts*)cl[i].o);
  }
  free(cl);
  return rc;
}

void ra_shout_pre_free(ra_shout_pre* s) {
  for (size_t i = 0; i < s->n_lut; i++) shout_cols_free(&s->cols[i]);
  free(s->cols);
  free(s->claimed);
  free(s->r_addr);
  for (size_t i = 0; i < s->n_active; i++) free(s->rounds[i]);
  free(s->rounds);
  free(s->degs);
  memset(s, 0, sizeof(*s));
}

int ra_twist_addr_pre(or_tr* tr, const or_step_mem* sm, const ra_layout* L, const fq* zdec, size_t ell_n,
                      const kx* r_cycle, ra_twist_pre* out, char* err, size_t errlen) {
  const size_t nm = sm->n_mem;
  for (size_t i = 0; i < nm; i++) {
    const or_mem_inst* in = &sm->mems[i];
    ra_twist_pre* P = &out[i];
    memset(P, 0, sizeof(*P));
    const size_t e = in->d * in->ell;
    if (in->steps > ((size_t)1 << ell_n)) {
      snprintf(err, errlen, "Twist(Route A): steps exceeds 2^ell_cycle");
      return -1;
    }
    if (ra_decode_twist(zdec, L, i, in->steps, ell_n, &P->cols)) {
      snprintf(err, errlen, "bus column decode failed (mem %zu)", i);
      return -1;
    }
    ra_twist_cols* T = &P->cols;
    uint64_t* ia = (uint64_t*)malloc((in->ninit + 1) * sizeof(uint64_t));
    kx* iv = (kx*)malloc((in->ninit + 1) * sizeof(kx));
    size_t ni = in->sparse ? in->ninit : 0;
    for (size_t q = 0; q < ni; q++) {
      if (in->init_addrs[q] >= in->k) {
        snprintf(err, errlen, "Twist: init address out of range");
        return -1;
      }
      ia[q] = in->init_addrs[q];
      iv[q] = or_k_from(in->init_vals[q]);
    }
    or_sparse_in ra[64], wa[64];
    for (size_t b = 0; b < e; b++) {
      ra[b] = ra_in(&T->ra[b]);
      wa[b] = ra_in(&T->wa[b]);
    }
    ra_claim cl[2];
    memset(cl, 0, sizeof(cl));
    cl[0].kind = cl[1].kind = RA_TS;
    cl[0].o = or_ts_read_addr(ia, iv, ni, r_cycle, ell_n, ra_in(&T->has_read), ra_in(&T->rv), ra, ra_in(&T->has_write),
                              wa, ra_in(&T->inc), e);
    cl[1].o = or_ts_write_addr(ia, iv, ni, r_cycle, ell_n, ra_in(&T->has_write), ra_in(&T->wv), wa, ra_in(&T->inc), e);
    free(ia);
    free(iv);
    cl[0].label = "twist/read_addr_pre";
    cl[1].label = "twist/write_addr_pre";
    for (int c = 0; c < 2; c++) {
      cl[c].deg = or_ts_degree_bound((or_ts*)cl[c].o);
      cl[c].claimed_sum = K0;
    }
    const kx zero2[2] = {K0, K0};
    const char* labels[2] = {"twist/read_addr_pre", "twist/write_addr_pre"};
    ra_msg_u64(tr, "twist/addr_pre_time/claim_idx", i);
    bind_claim_sums(tr, "twist/addr_pre_time/claimed_sums", zero2, labels, 2);
    P->ell_addr = e;
    P->r_addr = (kx*)calloc(e + 1, sizeof(kx));
    int rc = ra_batched_sumcheck(tr, "twist/addr_pre_time", i, cl, 2, e, P->r_addr, err, errlen);
    for (int c = 0; c < 2; c++) {
      P->rounds[c] = cl[c].polys;
      P->degs[c] = cl[c].deg;
      or_ts_free((or_ts*)cl[c].o);
    }
    if (rc) return rc;
    P->read_claim = cl[0].running;
    P->write_claim = cl[1].running;
  }
  return 0;
}

void ra_twist_pre_free(ra_twist_pre* P) {
  ra_twist_cols_free(&P->cols);
  free(P->r_addr);
  free(P->rounds[0]);
  free(P->rounds[1]);
  memset(P, 0, sizeof(*P));
}

/* ---------------- Route A time oracles (memory.rs:905-1025) ---------------- */
int ra_build_time_claims(const or_step_mem* sm, size_t ell_n, const kx* r_cycle, const ra_shout_pre* sp,
                         const ra_twist_pre* tp, ra_claim* cl, size_t* ncl, char* err, size_t errlen) {
  size_t n = 0;
  for (size_t i = 0; i < sm->n_lut; i++) {
    const ra_shout_cols* S = &sp->cols[i];
    const size_t e = S->ell;
    or_sparse_in bits[66];
    for (size_t b = 0; b < e; b++) bits[b] = ra_in(&S->addr[b]);
    kx vclaim, aclaim;
    ra_claim* c = &cl[n];
    memset(c, 0, 3 * sizeof(ra_claim));
    c[0].o = or_ts_shout_value(r_cycle, ell_n, ra_in(&S->has_lookup), ra_in(&S->val), &vclaim);
    c[0].claimed_sum = vclaim;
    c[0].label = "shout/value";
    c[0].dynamic = 1;
    c[1].o = or_ts_index_adapter(r_cycle, ell_n, ra_in(&S->has_lookup), bits, e, sp->r_addr, &aclaim);
    c[1].claimed_sum = aclaim;
    c[1].label = "shout/adapter";
    c[1].dynamic = 1;
    bits[e] = ra_in(&S->has_lookup);
    kx w[66];
    bitness_weights(r_cycle, ell_n, e + 1, 0x53484F5554ull + i, w);
    c[2].o = or_ts_bitness(r_cycle, ell_n, bits, w, e + 1);
    c[2].claimed_sum = K0;
    c[2].label = "shout/bitness";
    c[2].dynamic = 0;
    n += 3;
  }
  for (size_t i = 0; i < sm->n_mem; i++) {
    const ra_twist_pre* P = &tp[i];
    const ra_twist_cols* T = &P->cols;
    const size_t e = P->ell_addr;
    const kx init_at = eval_init(&sm->mems[i], P->r_addr, e);
    or_sparse_in ra[64], wa[64], bits[132];
    for (size_t b = 0; b < e; b++) {
      ra[b] = ra_in(&T->ra[b]);
      wa[b] = ra_in(&T->wa[b]);
      bits[b] = ra[b];
      bits[e + b] = wa[b];
    }
    bits[2 * e] = ra_in(&T->has_read);
    bits[2 * e + 1] = ra_in(&T->has_write);
    ra_claim* c = &cl[n];
    memset(c, 0, 3 * sizeof(ra_claim));
    c[0].o = or_ts_twist_read(r_cycle, ell_n, ra_in(&T->has_read), ra_in(&T->rv), ra, ra_in(&T->has_write),
                              ra_in(&T->inc), wa, e, P->r_addr, init_at);
    c[0].claimed_sum = P->read_claim;
    c[0].label = "twist/read_check";
    c[0].dynamic = 1;
    c[1].o = or_ts_twist_write(r_cycle, ell_n, ra_in(&T->has_write), ra_in(&T->wv), ra_in(&T->inc), wa, e, P->r_addr,
                               init_at);
    c[1].claimed_sum = P->write_claim;
    c[1].label = "twist/write_check";
    c[1].dynamic = 1;
    kx w[132];
    bitness_weights(r_cycle, ell_n, 2 * e + 2, 0x5457495354ull + i, w);
    c[2].o = or_ts_bitness(r_cycle, ell_n, bits, w, 2 * e + 2);
    c[2].claimed_sum = K0;
    c[2].label = "twist/bitness";
    c[2].dynamic = 0;
    n += 3;
  }
  for (size_t i = 0; i < n; i++) {
    cl[i].kind = RA_TS;
    cl[i].deg = or_ts_degree_bound((or_ts*)cl[i].o);
    if (or_ts_num_rounds((or_ts*)cl[i].o) != ell_n) {
      snprintf(err, errlen, "Route A time oracle %s has %zu rounds, expected ell_n = %zu", cl[i].label,
               or_ts_num_rounds((or_ts*)cl[i].o), ell_n);
      return -1;
    }
  }
  *ncl = n;
  return 0;
}

/* bind_batched_dynamic_claims (transcript.rs:36-64) */
void ra_bind_dynamic_claims(or_tr* tr, const ra_claim* cl, size_t n) {
  ra_msg_u64(tr, "batched/dynamic_bind/len", n);
  for (size_t i = 0; i < n; i++) {
    msg_str(tr, "batched/dynamic_bind/claim_label", cl[i].label);
    ra_msg_u64(tr, "batched/dynamic_bind/claim_idx", i);
    ra_msg_u64(tr, "batched/dynamic_bind/degree_bound", cl[i].deg);
    uint8_t d = cl[i].dynamic ? 1 : 0;
    or_tr_append_message(tr, "batched/dynamic_bind/is_dynamic", &d, 1);
    if (cl[i].dynamic) tr_k(tr, "batched/dynamic_bind/claimed_sum", cl[i].claimed_sum);
  }
}

/* ---------------- val-eval batch (memory.rs:1382-1535) ---------------- */
int ra_val_eval(or_tr* tr, const or_step_mem* sm, const ra_twist_pre* tp, const ra_twist_cols* prev, size_t ell_n,
                const kx* r_time, size_t step_idx, ra_val_pre* out, char* err, size_t errlen) {
  memset(out, 0, sizeof(*out));
  const size_t nm = sm->n_mem;
  if (nm == 0) return 0;
  const int has_prev = prev != NULL;
  const size_t per = has_prev ? 3 : 2, ncl = per * nm;
  ra_claim* cl = (ra_claim*)calloc(ncl, sizeof(ra_claim));
  uint8_t* tags = (uint8_t*)malloc(ncl);
  for (size_t i = 0; i < nm; i++) {
    const ra_twist_pre* P = &tp[i];
    const size_t e = P->ell_addr;
    or_sparse_in wa[64];
    for (size_t b = 0; b < e; b++) wa[b] = ra_in(&P->cols.wa[b]);
    ra_claim* c = &cl[per * i];
    kx s;
    c[0].o = or_ts_val_eval(ell_n, wa, e, ra_in(&P->cols.has_write), ra_in(&P->cols.inc), P->r_addr, r_time, &s);
    c[0].claimed_sum = s;
    c[0].label = "twist/val_eval_lt";
    tags[per * i] = 0;
    c[1].o = or_ts_total_inc(ell_n, wa, e, ra_in(&P->cols.has_write), ra_in(&P->cols.inc), P->r_addr, &s);
    c[1].claimed_sum = s;
    c[1].label = "twist/val_eval_total";
    tags[per * i + 1] = 1;
    if (has_prev) {
      or_sparse_in pwa[64];
      if (prev[i].ell != e) {
        snprintf(err, errlen, "Twist rollover requires stable geometry");
        return -1;
      }
      for (size_t b = 0; b < e; b++) pwa[b] = ra_in(&prev[i].wa[b]);
      c[2].o = or_ts_total_inc(ell_n, pwa, e, ra_in(&prev[i].has_write), ra_in(&prev[i].inc), P->r_addr, &s);
      c[2].claimed_sum = s;
      c[2].label = "twist/rollover_prev_total";
      tags[per * i + 2] = 2;
    }
  }
  for (size_t i = 0; i < ncl; i++) {
    cl[i].kind = RA_TS;
    cl[i].deg = or_ts_degree_bound((or_ts*)cl[i].o);
  }
  ra_msg_u64(tr, "twist/val_eval/batch_start", nm);
  ra_msg_u64(tr, "twist/val_eval/step_idx", step_idx);
  ra_msg_u64(tr, "twist/val_eval/claimed_sums_len", ncl);
  for (size_t i = 0; i < ncl; i++) {
    ra_msg_u64(tr, "twist/val_eval/claim_idx", i);
    or_tr_append_message(tr, "twist/val_eval/claim_kind", &tags[i], 1);
    tr_k(tr, "twist/val_eval/claimed_sum", cl[i].claimed_sum);
  }
  free(tags);
  out->r_val = (kx*)calloc(ell_n + 1, sizeof(kx));
  int rc = ra_batched_sumcheck(tr, "twist/val_eval_batch", step_idx, cl, ncl, ell_n, out->r_val, err, errlen);
  out->n_mem = nm;
  out->per = per;
  out->claims = cl;
  for (size_t i = 0; i < ncl; i++) {
    or_ts_free((or_ts*)cl[i].o);
    cl[i].o = NULL;
  }
  if (rc) return rc;
  or_tr_append_message(tr, "twist/val_eval/batch_done", NULL, 0);
  return 0;
}

void ra_val_pre_free(ra_val_pre* v) {
  if (v->claims)
    for (size_t i = 0; i < v->per * v->n_mem; i++) ra_claim_free(&v->claims[i]);
  free(v->claims);
  free(v->r_val);
  memset(v, 0, sizeof(*v));
}

/* append_bus_openings_to_me_instance (cpu_bus.rs:195-320) for an int32 matrix (Z_mix) */
int ra_bus_openings_i32(const int32_t* Z, const ra_layout* L, const kx* r, size_t ell_n, kx* y, kx* ys) {
  if (L->bus_cols == 0) return 0;
  kx* w = (kx*)malloc((L->chunk + 1) * sizeof(kx));
  for (size_t j = 0; j < L->chunk; j++) {
    size_t idx = L->m_in + j;
    if (idx >= ((size_t)1 << ell_n)) {
      free(w);
      return -1;
    }
    kx acc = K1;
    for (size_t b = 0; b < ell_n; b++) acc = or_k_mul(acc, ((idx >> b) & 1) ? r[b] : or_k_sub(K1, r[b]));
    w[j] = acc;
  }
  for (size_t col = 0; col < L->bus_cols; col++) {
    kx* row = y + col * 64;
    for (size_t rho = 0; rho < 64; rho++) row[rho] = K0;
    for (size_t rho = 0; rho < NEO_D; rho++) {
      kx acc = K0;
      for (size_t j = 0; j < L->chunk; j++) {
        int32_t d = Z[rho * L->m + L->bus_base + col * L->chunk + j];
        if (d) acc = or_k_add(acc, or_k_scale(w[j], or_fq_from_i64(d)));
      }
      row[rho] = acc;
    }
    kx s = K0;
    fq pw = 1;
    for (size_t rho = 0; rho < NEO_D; rho++) {
      s = or_k_add(s, or_k_scale(row[rho], pw));
      pw = or_fq_add(pw, pw);
    }
    ys[col] = s;
  }
  free(w);
  return 0;
}

/* ---------------- output binding: the output sum-check (output_check.rs) ---------------- */
static kx ob_challenge(or_tr* tr) {
  /* sample_k_challenge (output_check.rs:20-25) */
  const fq re = or_tr_challenge_field(tr, "output_check/chal/re");
  const fq im = or_tr_challenge_field(tr, "output_check/chal/im");
  kx k = {re, im};
  return k;
}

int ra_output_sumcheck(or_tr* tr, const or_ob_config* ob, kx* rounds, kx* r_prime, char* err, size_t errlen) {
  const size_t nb = ob->num_bits;
  if (nb > 30) {
    snprintf(err, errlen, "output binding: num_bits above 30");
    return -1;
  }
  const size_t k = (size_t)1 << nb;
  for (size_t i = 0; i < ob->n_claims; i++)
    if (ob->addrs[i] >= k || (i && ob->addrs[i] <= ob->addrs[i - 1])) {
      snprintf(err, errlen, "output binding: claimed addresses must be increasing and below 2^num_bits");
      return -1;
    }
  /* ProgramIO::absorb_into_transcript (output_check.rs:162-171): claims in address order */
  ra_msg_u64(tr, "output_check/num_claims", ob->n_claims);
  for (size_t i = 0; i < ob->n_claims; i++) {
    ra_msg_u64(tr, "output_check/addr", ob->addrs[i]);
    or_tr_append_fields(tr, "output_check/value", &ob->values[i], 1);
  }
  kx* r_addr = (kx*)malloc((nb + 1) * sizeof(kx));
  for (size_t i = 0; i < nb; i++) {
    ra_msg_u64(tr, "output_check/r_addr/idx", i);
    r_addr[i] = ob_challenge(tr);
  }
  /* tables (output_check.rs:319-357): chi(r_addr), the claim mask, the final state, the claims */
  kx* eq = (kx*)malloc(k * sizeof(kx));
  kx* mask = (kx*)calloc(k, sizeof(kx));
  kx* vf = (kx*)malloc(k * sizeof(kx));
  kx* vio = (kx*)calloc(k, sizeof(kx));
  or_chi_table(r_addr, nb, eq);
  for (size_t a = 0; a < k; a++) vf[a] = or_k_from(ob->final_state[a]);
  for (size_t i = 0; i < ob->n_claims; i++) {
    mask[ob->addrs[i]] = or_k_from(1);
    vio[ob->addrs[i]] = or_k_from(ob->values[i]);
  }
  /* run_sumcheck_prover (output_check.rs:600-620): evaluations at 0..3 (LSB-first pairs),
   * interpolated, each coefficient absorbed, then the challenge */
  kx xs[4], ys[4];
  for (int x = 0; x < 4; x++) xs[x] = or_k_from((fq)x);
  size_t n = k;
  for (size_t rd = 0; rd < nb; rd++) {
    const size_t half = n / 2;
    for (int x = 0; x < 4; x++) {
      kx acc = or_k_from(0);
      for (size_t i = 0; i < half; i++) {
        const size_t a0 = 2 * i, a1 = a0 + 1;
        const kx X = xs[x];
        const kx e = or_k_add(eq[a0], or_k_mul(or_k_sub(eq[a1], eq[a0]), X));
        const kx mk = or_k_add(mask[a0], or_k_mul(or_k_sub(mask[a1], mask[a0]), X));
        const kx f = or_k_add(vf[a0], or_k_mul(or_k_sub(vf[a1], vf[a0]), X));
        const kx v = or_k_add(vio[a0], or_k_mul(or_k_sub(vio[a1], vio[a0]), X));
        acc = or_k_add(acc, or_k_mul(or_k_mul(e, mk), or_k_sub(f, v)));
      }
      ys[x] = acc;
    }
    kx* co = &rounds[rd * 4];
    or_interpolate(xs, ys, 4, co);
    for (int c = 0; c < 4; c++) tr_k(tr, "output_check/round_coeff", co[c]);
    const kx r = ob_challenge(tr);
    r_prime[rd] = r;
    for (size_t i = 0; i < half; i++) {
      const size_t a0 = 2 * i, a1 = a0 + 1;
      eq[i] = or_k_add(eq[a0], or_k_mul(or_k_sub(eq[a1], eq[a0]), r));
      mask[i] = or_k_add(mask[a0], or_k_mul(or_k_sub(mask[a1], mask[a0]), r));
      vf[i] = or_k_add(vf[a0], or_k_mul(or_k_sub(vf[a1], vf[a0]), r));
      vio[i] = or_k_add(vio[a0], or_k_mul(or_k_sub(vio[a1], vio[a0]), r));
    }
    n = half;
  }
  free(r_addr);
  free(eq);
  free(mask);
  free(vf);
  free(vio);
  (void)err;
  (void)errlen;
  return 0;
}
