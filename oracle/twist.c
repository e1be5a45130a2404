/*
 * twist.c -- Twist/Shout sum-check oracles (time lanes and address lanes), CPU restatement.
 * TEST INFRASTRUCTURE ONLY (checker for nightstream_amd; see neo_oracle.h).
 *
 * Restates neo-memory/src/twist_oracle.rs:
 *   56-183     ProductRoundOracle (dense product of multilinear factors)
 *   185-198    chi_cycle_children, gather_pairs_from_sparse
 *   200-306    ShoutValueOracleSparse
 *   313-361    fill_time_point, lt_eval_at_bool_index, val_pre_from_inc_terms,
 *              build_inc_terms_at_r_addr
 *   384-507    IndexAdapterOracleSparseTime
 *   515-619    LazyWeightedBitnessOracleSparseTime
 *   622-764    TwistReadCheckOracleSparseTime
 *   767-909    TwistWriteCheckOracleSparseTime
 *   911-1052   TwistValEvalOracleSparseTime
 *   1054-1196  TwistTotalIncOracleSparseTime
 *   1222-1465  update_prefix_weights_in_place, addr_from_sparse_bits_at_time,
 *              merge_sparse_time_indices, TwistReadCheckAddrOracleSparseTime
 *   1468-1660  TwistWriteCheckAddrOracleSparseTime
 *   1662-1736  AddressLookupOracle
 * neo-memory/src/sparse_time.rs:10-127 (SparseIdxVec, fold_round_in_place),
 * neo-memory/src/mle.rs:7-24 (lt_eval), 70-95 (eq_single, chi_at_index),
 * neo-memory/src/bit_ops.rs:5-9 (eq_bit_affine).
 * The address-lane memory map (a HashMap in the reference) is a dense array indexed by
 * addr >> bit_idx; absent keys read as zero in both.
 *
 * Pinned by: tests/test_oracle_twist.py -- every round polynomial of every oracle equals a
 * brute-force sum over the Boolean hypercube of the dense multilinear extensions of the same
 * columns (the identity the reference checks in ts_oracle_identity_tests.rs /
 * track_a_sparse_equivalence.rs), the claimed sums equal the dense sums, and the
 * address-lane oracles equal a direct replay of memory over time.
 */
#include <stdlib.h>
#include <string.h>
#include "neo_oracle.h"

static const kx K0 = {0, 0};
static const kx K1 = {1, 0};

static int kz(kx a) { return a.c0 == 0 && a.c1 == 0; }
static kx kmul(kx a, kx b) { return or_k_mul(a, b); }
static kx kadd(kx a, kx b) { return or_k_add(a, b); }
static kx ksub(kx a, kx b) { return or_k_sub(a, b); }
static kx interp(kx f0, kx f1, kx x) { return kadd(f0, kmul(ksub(f1, f0), x)); }
/* mle.rs:71-73 */
static kx eq_single(kx a, kx b) { return kadd(kmul(ksub(K1, a), ksub(K1, b)), kmul(a, b)); }
/* bit_ops.rs:5-9: eq(bit, u) = bit*(2u-1) + (1-u) */
static kx eq_bit_affine(kx bit, kx u) { return kadd(kmul(bit, ksub(kadd(u, u), K1)), ksub(K1, u)); }
/* mle.rs:85-95 */
static kx chi_at_index(const kx* r, size_t ell, uint64_t idx) {
  kx acc = K1;
  for (size_t b = 0; b < ell; b++) acc = kmul(acc, ((idx >> b) & 1) ? r[b] : ksub(K1, r[b]));
  return acc;
}

/* ---------------- SparseIdxVec (sparse_time.rs:10-127) ---------------- */
typedef struct {
  size_t len, n;
  uint64_t* idx;
  kx* v;
} spv;

static int cmp_entry(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}

static void spv_init(spv* s, size_t len, or_sparse_in in) {
  s->len = len;
  s->n = 0;
  s->idx = (uint64_t*)malloc((in.nnz + 1) * sizeof(uint64_t));
  s->v = (kx*)malloc((in.nnz + 1) * sizeof(kx));
  /* from_entries: retain nonzero, sort by index */
  uint64_t* tmp = (uint64_t*)malloc((in.nnz + 1) * 3 * sizeof(uint64_t));
  size_t m = 0;
  for (size_t i = 0; i < in.nnz; i++) {
    kx v = {in.vals[2 * i], in.vals[2 * i + 1]};
    if (kz(v)) continue;
    tmp[3 * m] = in.idx[i];
    tmp[3 * m + 1] = v.c0;
    tmp[3 * m + 2] = v.c1;
    m++;
  }
  qsort(tmp, m, 3 * sizeof(uint64_t), cmp_entry);
  for (size_t i = 0; i < m; i++) {
    s->idx[i] = tmp[3 * i];
    s->v[i].c0 = tmp[3 * i + 1];
    s->v[i].c1 = tmp[3 * i + 2];
  }
  s->n = m;
  free(tmp);
}

static void spv_free(spv* s) {
  free(s->idx);
  free(s->v);
}

static kx spv_get(const spv* s, uint64_t idx) {
  size_t lo = 0, hi = s->n;
  while (lo < hi) {
    size_t mid = (lo + hi) / 2;
    if (s->idx[mid] < idx) lo = mid + 1;
    else hi = mid;
  }
  return (lo < s->n && s->idx[lo] == idx) ? s->v[lo] : K0;
}

static void spv_fold(spv* s, kx r) {
  kx omr = ksub(K1, r);
  size_t out = 0, i = 0;
  while (i < s->n) {
    uint64_t idx = s->idx[i];
    uint64_t parent = idx >> 1;
    kx acc = K0;
    if ((idx & 1) == 0) {
      acc = kadd(acc, kmul(s->v[i], omr));
      i++;
      if (i < s->n && s->idx[i] == idx + 1) {
        acc = kadd(acc, kmul(s->v[i], r));
        i++;
      }
    } else {
      acc = kadd(acc, kmul(s->v[i], r));
      i++;
    }
    if (!kz(acc)) {
      if (out > 0 && s->idx[out - 1] == parent) {
        s->v[out - 1] = kadd(s->v[out - 1], acc);
      } else {
        s->idx[out] = parent;
        s->v[out] = acc;
        out++;
      }
    }
  }
  s->n = out;
  s->len >>= 1;
}

static kx spv_single(const spv* s) { return s->n ? s->v[0] : K0; }

/* ---------------- shared helpers (twist_oracle.rs:185-361) ---------------- */
static void chi_cycle_children(const kx* r_cycle, size_t ell, size_t bit_idx, kx prefix_eq, uint64_t pair, kx* c0,
                               kx* c1) {
  kx suffix = K1;
  size_t shift = 0;
  for (size_t b = bit_idx + 1; b < ell; b++) {
    int bit = (pair >> shift) & 1;
    suffix = kmul(suffix, bit ? r_cycle[b] : ksub(K1, r_cycle[b]));
    shift++;
  }
  kx r = r_cycle[bit_idx];
  *c0 = kmul(kmul(prefix_eq, ksub(K1, r)), suffix);
  *c1 = kmul(kmul(prefix_eq, r), suffix);
}

/* distinct parents of a sorted entry list */
static size_t gather_pairs(const spv* s, uint64_t* out) {
  size_t n = 0;
  for (size_t i = 0; i < s->n; i++) {
    uint64_t p = s->idx[i] >> 1;
    if (n == 0 || out[n - 1] != p) out[n++] = p;
  }
  return n;
}

static void fill_time_point(kx* out, size_t ell, const kx* prefix, size_t bit_idx, kx bit_value, uint64_t pair) {
  for (size_t i = 0; i < bit_idx; i++) out[i] = prefix[i];
  out[bit_idx] = bit_value;
  size_t shift = 0;
  for (size_t b = bit_idx + 1; b < ell; b++) {
    out[b] = ((pair >> shift) & 1) ? K1 : K0;
    shift++;
  }
}

static kx lt_eval_at_bool_index(uint64_t idx, const kx* point, size_t ell) {
  kx suffix = K1, acc = K0;
  for (size_t ii = ell; ii-- > 0;) {
    if (((idx >> ii) & 1) == 0) {
      acc = kadd(acc, kmul(point[ii], suffix));
      suffix = kmul(suffix, ksub(K1, point[ii]));
    } else {
      suffix = kmul(suffix, point[ii]);
    }
  }
  return acc;
}

/* mle.rs:7-24 */
static kx lt_eval(const kx* jp, const kx* j, size_t ell) {
  kx suffix[65];
  suffix[ell] = K1;
  for (size_t i = ell; i-- > 0;) suffix[i] = kmul(suffix[i + 1], eq_single(jp[i], j[i]));
  kx acc = K0;
  for (size_t i = 0; i < ell; i++) acc = kadd(acc, kmul(kmul(ksub(K1, jp[i]), j[i]), suffix[i + 1]));
  return acc;
}

typedef struct {
  uint64_t t;
  kx v;
} term;

static kx val_pre_from_terms(kx init, const term* terms, size_t nt, const kx* tpoint, size_t ell) {
  kx acc = init;
  for (size_t i = 0; i < nt; i++) acc = kadd(acc, kmul(terms[i].v, lt_eval_at_bool_index(terms[i].t, tpoint, ell)));
  return acc;
}

/* ---------------- oracle object ---------------- */
enum {
  TS_PRODUCT = 0,
  TS_SHOUT_VALUE,
  TS_INDEX_ADAPTER,
  TS_BITNESS,
  TS_TWIST_READ,
  TS_TWIST_WRITE,
  TS_VAL_EVAL,
  TS_TOTAL_INC,
  TS_READ_ADDR,
  TS_WRITE_ADDR,
  TS_ADDR_LOOKUP
};

struct or_ts {
  int kind;
  size_t ell_n, bit_idx, degree_bound, ell_addr;
  kx* r_cycle;  /* or r_time for VAL_EVAL */
  kx* r_addr;
  kx prefix_eq;
  kx* chals;
  size_t nchals;
  /* time-lane columns */
  spv gate, a, b;  /* gate = has_lookup/has_read/has_write; a, b per kind */
  spv* bits;       /* addr bit columns (ell_addr) or bitness columns */
  size_t nbits;
  kx* weights;
  kx init_at_r_addr;
  term* terms;
  size_t nterms;
  kx* tc0;
  kx* tc1;
  /* product oracle */
  kx** factors;
  size_t nf, flen;
  /* address lanes */
  size_t nev;
  kx *eq_cycle, *ev_hr, *ev_rv, *ev_hw, *ev_inc, *ra_w, *wa_w;
  uint64_t *ra_addr, *wa_addr;
  size_t ninit;
  uint64_t* init_addr;
  kx *init_val, *init_w;
  kx* mem;
};

static or_ts* ts_new(int kind, const kx* r_cycle, size_t ell_n) {
  or_ts* o = (or_ts*)calloc(1, sizeof(or_ts));
  o->kind = kind;
  o->ell_n = ell_n;
  o->prefix_eq = K1;
  o->r_cycle = (kx*)malloc((ell_n + 1) * sizeof(kx));
  if (r_cycle) memcpy(o->r_cycle, r_cycle, ell_n * sizeof(kx));
  o->chals = (kx*)malloc((ell_n + 32) * sizeof(kx));
  o->tc0 = (kx*)malloc((ell_n + 1) * sizeof(kx));
  o->tc1 = (kx*)malloc((ell_n + 1) * sizeof(kx));
  return o;
}

static void set_bits(or_ts* o, const or_sparse_in* cols, size_t n, size_t len) {
  o->nbits = n;
  o->bits = (spv*)calloc(n ? n : 1, sizeof(spv));
  for (size_t b = 0; b < n; b++) spv_init(&o->bits[b], len, cols[b]);
}

static void set_raddr(or_ts* o, const kx* r_addr, size_t ell_addr) {
  o->ell_addr = ell_addr;
  o->r_addr = (kx*)malloc((ell_addr + 1) * sizeof(kx));
  if (ell_addr) memcpy(o->r_addr, r_addr, ell_addr * sizeof(kx));
}

static kx eq_addr_at(const or_ts* o, uint64_t t) {
  kx acc = K1;
  for (size_t b = 0; b < o->nbits; b++) acc = kmul(acc, eq_bit_affine(spv_get(&o->bits[b], t), o->r_addr[b]));
  return acc;
}

static void build_terms(or_ts* o, const spv* wa_bits, size_t ell_addr, const spv* has_write, const spv* inc) {
  o->terms = (term*)malloc((has_write->n + 1) * sizeof(term));
  o->nterms = 0;
  for (size_t i = 0; i < has_write->n; i++) {
    uint64_t t = has_write->idx[i];
    kx hw = has_write->v[i];
    kx inc_t = spv_get(inc, t);
    if (kz(hw) || kz(inc_t)) continue;
    kx eq = K1;
    for (size_t b = 0; b < ell_addr; b++) eq = kmul(eq, eq_bit_affine(spv_get(&wa_bits[b], t), o->r_addr[b]));
    kx v = kmul(kmul(hw, inc_t), eq);
    if (!kz(v)) {
      o->terms[o->nterms].t = t;
      o->terms[o->nterms].v = v;
      o->nterms++;
    }
  }
}

/* ---- constructors ---- */
or_ts* or_ts_product(const kx* const* factors, size_t nf, size_t len, size_t degree_bound) {
  size_t ell = 0;
  while (((size_t)1 << ell) < len) ell++;
  or_ts* o = ts_new(TS_PRODUCT, NULL, ell);
  o->nf = nf;
  o->flen = len;
  o->factors = (kx**)malloc((nf ? nf : 1) * sizeof(kx*));
  for (size_t f = 0; f < nf; f++) {
    o->factors[f] = (kx*)malloc(len * sizeof(kx));
    memcpy(o->factors[f], factors[f], len * sizeof(kx));
  }
  o->degree_bound = degree_bound;
  return o;
}

or_ts* or_ts_shout_value(const kx* r_cycle, size_t ell_n, or_sparse_in has_lookup, or_sparse_in val, kx* claim) {
  or_ts* o = ts_new(TS_SHOUT_VALUE, r_cycle, ell_n);
  size_t len = (size_t)1 << ell_n;
  spv_init(&o->gate, len, has_lookup);
  spv_init(&o->a, len, val);
  kx c = K0;
  for (size_t i = 0; i < o->gate.n; i++) {
    uint64_t t = o->gate.idx[i];
    kx v = spv_get(&o->a, t);
    if (kz(v)) continue;
    c = kadd(c, kmul(kmul(chi_at_index(r_cycle, ell_n, t), o->gate.v[i]), v));
  }
  if (claim) *claim = c;
  o->degree_bound = 3;
  return o;
}

or_ts* or_ts_index_adapter(const kx* r_cycle, size_t ell_n, or_sparse_in has_lookup, const or_sparse_in* addr_bits,
                           size_t ell_addr, const kx* r_addr, kx* claim) {
  or_ts* o = ts_new(TS_INDEX_ADAPTER, r_cycle, ell_n);
  size_t len = (size_t)1 << ell_n;
  spv_init(&o->gate, len, has_lookup);
  set_bits(o, addr_bits, ell_addr, len);
  set_raddr(o, r_addr, ell_addr);
  kx c = K0;
  for (size_t i = 0; i < o->gate.n; i++) {
    uint64_t t = o->gate.idx[i];
    c = kadd(c, kmul(kmul(chi_at_index(r_cycle, ell_n, t), o->gate.v[i]), eq_addr_at(o, t)));
  }
  if (claim) *claim = c;
  o->degree_bound = 2 + ell_addr;
  return o;
}

or_ts* or_ts_bitness(const kx* r_cycle, size_t ell_n, const or_sparse_in* cols, const kx* weights, size_t ncols) {
  or_ts* o = ts_new(TS_BITNESS, r_cycle, ell_n);
  size_t len = (size_t)1 << ell_n;
  set_bits(o, cols, ncols, len);
  o->weights = (kx*)malloc((ncols + 1) * sizeof(kx));
  if (ncols) memcpy(o->weights, weights, ncols * sizeof(kx));
  o->degree_bound = 3;
  return o;
}

or_ts* or_ts_twist_read(const kx* r_cycle, size_t ell_n, or_sparse_in has_read, or_sparse_in rv,
                        const or_sparse_in* ra_bits, or_sparse_in has_write, or_sparse_in inc,
                        const or_sparse_in* wa_bits, size_t ell_addr, const kx* r_addr, kx init_at_r_addr) {
  or_ts* o = ts_new(TS_TWIST_READ, r_cycle, ell_n);
  size_t len = (size_t)1 << ell_n;
  spv_init(&o->gate, len, has_read);
  spv_init(&o->a, len, rv);
  set_bits(o, ra_bits, ell_addr, len);
  set_raddr(o, r_addr, ell_addr);
  spv hw, ic;
  spv_init(&hw, len, has_write);
  spv_init(&ic, len, inc);
  spv* wb = (spv*)calloc(ell_addr ? ell_addr : 1, sizeof(spv));
  for (size_t b = 0; b < ell_addr; b++) spv_init(&wb[b], len, wa_bits[b]);
  build_terms(o, wb, ell_addr, &hw, &ic);
  for (size_t b = 0; b < ell_addr; b++) spv_free(&wb[b]);
  free(wb);
  spv_free(&hw);
  spv_free(&ic);
  o->init_at_r_addr = init_at_r_addr;
  o->degree_bound = 3 + ell_addr;
  return o;
}

or_ts* or_ts_twist_write(const kx* r_cycle, size_t ell_n, or_sparse_in has_write, or_sparse_in wv, or_sparse_in inc,
                         const or_sparse_in* wa_bits, size_t ell_addr, const kx* r_addr, kx init_at_r_addr) {
  or_ts* o = ts_new(TS_TWIST_WRITE, r_cycle, ell_n);
  size_t len = (size_t)1 << ell_n;
  spv_init(&o->gate, len, has_write);
  spv_init(&o->a, len, wv);
  spv_init(&o->b, len, inc);
  set_bits(o, wa_bits, ell_addr, len);
  set_raddr(o, r_addr, ell_addr);
  build_terms(o, o->bits, ell_addr, &o->gate, &o->b);
  o->init_at_r_addr = init_at_r_addr;
  o->degree_bound = 3 + ell_addr;
  return o;
}

or_ts* or_ts_val_eval(size_t ell_n, const or_sparse_in* wa_bits, size_t ell_addr, or_sparse_in has_write,
                      or_sparse_in inc, const kx* r_addr, const kx* r_time, kx* claim) {
  or_ts* o = ts_new(TS_VAL_EVAL, r_time, ell_n);
  size_t len = (size_t)1 << ell_n;
  spv_init(&o->gate, len, has_write);
  spv_init(&o->b, len, inc);
  set_bits(o, wa_bits, ell_addr, len);
  set_raddr(o, r_addr, ell_addr);
  kx c = K0;
  for (size_t i = 0; i < o->gate.n; i++) {
    uint64_t t = o->gate.idx[i];
    kx g = o->gate.v[i], it = spv_get(&o->b, t);
    if (kz(g) || kz(it)) continue;
    kx lt = lt_eval_at_bool_index(t, r_time, ell_n);
    c = kadd(c, kmul(kmul(kmul(g, it), lt), eq_addr_at(o, t)));
  }
  if (claim) *claim = c;
  o->degree_bound = 3 + ell_addr;
  return o;
}

or_ts* or_ts_total_inc(size_t ell_n, const or_sparse_in* wa_bits, size_t ell_addr, or_sparse_in has_write,
                       or_sparse_in inc, const kx* r_addr, kx* claim) {
  or_ts* o = ts_new(TS_TOTAL_INC, NULL, ell_n);
  size_t len = (size_t)1 << ell_n;
  spv_init(&o->gate, len, has_write);
  spv_init(&o->b, len, inc);
  set_bits(o, wa_bits, ell_addr, len);
  set_raddr(o, r_addr, ell_addr);
  kx c = K0;
  for (size_t i = 0; i < o->gate.n; i++) {
    uint64_t t = o->gate.idx[i];
    kx g = o->gate.v[i], it = spv_get(&o->b, t);
    if (kz(g) || kz(it)) continue;
    c = kadd(c, kmul(kmul(g, it), eq_addr_at(o, t)));
  }
  if (claim) *claim = c;
  o->degree_bound = 2 + ell_addr;
  return o;
}

static uint64_t addr_from_bits(const spv* cols, size_t n, uint64_t t) {
  uint64_t out = 0;
  for (size_t b = 0; b < n; b++)
    if (or_k_eq(spv_get(&cols[b], t), K1)) out |= (uint64_t)1 << b;
  return out;
}

static void set_init(or_ts* o, const uint64_t* addrs, const kx* vals, size_t ninit) {
  o->ninit = ninit;
  o->init_addr = (uint64_t*)malloc((ninit + 1) * sizeof(uint64_t));
  o->init_val = (kx*)malloc((ninit + 1) * sizeof(kx));
  o->init_w = (kx*)malloc((ninit + 1) * sizeof(kx));
  for (size_t i = 0; i < ninit; i++) {
    o->init_addr[i] = addrs[i];
    o->init_val[i] = vals[i];
    o->init_w[i] = K1;
  }
  o->mem = (kx*)malloc(((size_t)1 << o->ell_addr) * sizeof(kx) + sizeof(kx));
}

or_ts* or_ts_read_addr(const uint64_t* init_addrs, const kx* init_vals, size_t ninit, const kx* r_cycle, size_t ell_n,
                       or_sparse_in has_read, or_sparse_in rv, const or_sparse_in* ra_bits, or_sparse_in has_write,
                       const or_sparse_in* wa_bits, or_sparse_in inc, size_t ell_addr) {
  or_ts* o = ts_new(TS_READ_ADDR, r_cycle, ell_n);
  size_t len = (size_t)1 << ell_n;
  o->ell_addr = ell_addr;
  spv hr, r_v, hw, ic;
  spv_init(&hr, len, has_read);
  spv_init(&r_v, len, rv);
  spv_init(&hw, len, has_write);
  spv_init(&ic, len, inc);
  spv* rb = (spv*)calloc(ell_addr ? ell_addr : 1, sizeof(spv));
  spv* wb = (spv*)calloc(ell_addr ? ell_addr : 1, sizeof(spv));
  for (size_t b = 0; b < ell_addr; b++) {
    spv_init(&rb[b], len, ra_bits[b]);
    spv_init(&wb[b], len, wa_bits[b]);
  }
  /* merge_sparse_time_indices(has_read, has_write) */
  uint64_t* times = (uint64_t*)malloc((hr.n + hw.n + 1) * sizeof(uint64_t));
  size_t nt = 0, i = 0, j = 0;
  while (i < hr.n || j < hw.n) {
    uint64_t next;
    if (i < hr.n && j < hw.n) next = hr.idx[i] <= hw.idx[j] ? hr.idx[i++] : hw.idx[j++];
    else if (i < hr.n) next = hr.idx[i++];
    else next = hw.idx[j++];
    if (nt == 0 || times[nt - 1] != next) times[nt++] = next;
  }
  o->nev = nt;
  o->eq_cycle = (kx*)malloc((nt + 1) * sizeof(kx));
  o->ev_hr = (kx*)malloc((nt + 1) * sizeof(kx));
  o->ev_rv = (kx*)malloc((nt + 1) * sizeof(kx));
  o->ev_hw = (kx*)malloc((nt + 1) * sizeof(kx));
  o->ev_inc = (kx*)malloc((nt + 1) * sizeof(kx));
  o->ra_addr = (uint64_t*)malloc((nt + 1) * sizeof(uint64_t));
  o->wa_addr = (uint64_t*)malloc((nt + 1) * sizeof(uint64_t));
  o->ra_w = (kx*)malloc((nt + 1) * sizeof(kx));
  o->wa_w = (kx*)malloc((nt + 1) * sizeof(kx));
  for (size_t e = 0; e < nt; e++) {
    uint64_t t = times[e];
    o->eq_cycle[e] = chi_at_index(r_cycle, ell_n, t);
    o->ev_hr[e] = spv_get(&hr, t);
    o->ev_hw[e] = spv_get(&hw, t);
    o->ev_rv[e] = spv_get(&r_v, t);
    o->ev_inc[e] = spv_get(&ic, t);
    o->ra_addr[e] = kz(o->ev_hr[e]) ? 0 : addr_from_bits(rb, ell_addr, t);
    o->wa_addr[e] = kz(o->ev_hw[e]) ? 0 : addr_from_bits(wb, ell_addr, t);
    o->ra_w[e] = K1;
    o->wa_w[e] = K1;
  }
  free(times);
  for (size_t b = 0; b < ell_addr; b++) {
    spv_free(&rb[b]);
    spv_free(&wb[b]);
  }
  free(rb);
  free(wb);
  spv_free(&hr);
  spv_free(&r_v);
  spv_free(&hw);
  spv_free(&ic);
  set_init(o, init_addrs, init_vals, ninit);
  o->degree_bound = 2;
  return o;
}

or_ts* or_ts_write_addr(const uint64_t* init_addrs, const kx* init_vals, size_t ninit, const kx* r_cycle,
                        size_t ell_n, or_sparse_in has_write, or_sparse_in wv, const or_sparse_in* wa_bits,
                        or_sparse_in inc, size_t ell_addr) {
  or_ts* o = ts_new(TS_WRITE_ADDR, r_cycle, ell_n);
  size_t len = (size_t)1 << ell_n;
  o->ell_addr = ell_addr;
  spv hw, w_v, ic;
  spv_init(&hw, len, has_write);
  spv_init(&w_v, len, wv);
  spv_init(&ic, len, inc);
  spv* wb = (spv*)calloc(ell_addr ? ell_addr : 1, sizeof(spv));
  for (size_t b = 0; b < ell_addr; b++) spv_init(&wb[b], len, wa_bits[b]);
  size_t nt = hw.n;
  o->nev = nt;
  o->eq_cycle = (kx*)malloc((nt + 1) * sizeof(kx));
  o->ev_hw = (kx*)malloc((nt + 1) * sizeof(kx));
  o->ev_rv = (kx*)malloc((nt + 1) * sizeof(kx)); /* wv */
  o->ev_inc = (kx*)malloc((nt + 1) * sizeof(kx));
  o->wa_addr = (uint64_t*)malloc((nt + 1) * sizeof(uint64_t));
  o->wa_w = (kx*)malloc((nt + 1) * sizeof(kx));
  for (size_t e = 0; e < nt; e++) {
    uint64_t t = hw.idx[e];
    o->eq_cycle[e] = chi_at_index(r_cycle, ell_n, t);
    o->ev_hw[e] = hw.v[e];
    o->ev_rv[e] = spv_get(&w_v, t);
    o->ev_inc[e] = spv_get(&ic, t);
    o->wa_addr[e] = addr_from_bits(wb, ell_addr, t);
    o->wa_w[e] = K1;
  }
  for (size_t b = 0; b < ell_addr; b++) spv_free(&wb[b]);
  free(wb);
  spv_free(&hw);
  spv_free(&w_v);
  spv_free(&ic);
  set_init(o, init_addrs, init_vals, ninit);
  o->degree_bound = 2;
  return o;
}

or_ts* or_ts_addr_lookup(const or_sparse_in* addr_bits, size_t ell_addr, or_sparse_in has_lookup, const kx* table,
                         size_t table_len, const kx* r_cycle, size_t ell_n, kx* claim) {
  size_t len = (size_t)1 << ell_n, pow2_addr = (size_t)1 << ell_addr;
  spv hl;
  spv_init(&hl, len, has_lookup);
  spv* ab = (spv*)calloc(ell_addr ? ell_addr : 1, sizeof(spv));
  for (size_t b = 0; b < ell_addr; b++) spv_init(&ab[b], len, addr_bits[b]);
  kx* weight = (kx*)calloc(pow2_addr, sizeof(kx));
  for (size_t i = 0; i < hl.n; i++) {
    kx g = hl.v[i];
    if (kz(g)) continue;
    uint64_t t = hl.idx[i];
    kx w = kmul(chi_at_index(r_cycle, ell_n, t), g);
    if (kz(w)) continue;
    uint64_t a = addr_from_bits(ab, ell_addr, t);
    if (a < pow2_addr) weight[a] = kadd(weight[a], w);
  }
  kx c = K0;
  size_t lim = pow2_addr < table_len ? pow2_addr : table_len;
  for (size_t a = 0; a < lim; a++) c = kadd(c, kmul(table[a], weight[a]));
  if (claim) *claim = c;
  kx* tk = (kx*)calloc(pow2_addr, sizeof(kx));
  for (size_t a = 0; a < lim; a++) tk[a] = table[a];
  const kx* fs[2] = {tk, weight};
  or_ts* o = or_ts_product(fs, 2, pow2_addr, 2);
  o->kind = TS_ADDR_LOOKUP;
  free(tk);
  free(weight);
  for (size_t b = 0; b < ell_addr; b++) spv_free(&ab[b]);
  free(ab);
  spv_free(&hl);
  return o;
}

/* ---- RoundOracle ---- */
size_t or_ts_num_rounds(const or_ts* o) {
  switch (o->kind) {
    case TS_PRODUCT:
    case TS_ADDR_LOOKUP: {
      size_t l = 0;
      while (((size_t)1 << l) < o->flen) l++;
      return l;
    }
    case TS_TOTAL_INC: {
      size_t l = 0;
      while (((size_t)1 << l) < o->gate.len) l++;
      return l;
    }
    case TS_READ_ADDR:
    case TS_WRITE_ADDR:
      return o->ell_addr > o->bit_idx ? o->ell_addr - o->bit_idx : 0;
    default:
      return o->ell_n > o->bit_idx ? o->ell_n - o->bit_idx : 0;
  }
}

size_t or_ts_degree_bound(const or_ts* o) { return o->degree_bound; }

static void product_evals(or_ts* o, const kx* xs, size_t nx, kx* ys) {
  size_t rounds = or_ts_num_rounds(o);
  if (rounds == 0) {
    kx v = K1;
    for (size_t f = 0; f < o->nf; f++) v = kmul(v, o->factors[f][0]);
    for (size_t i = 0; i < nx; i++) ys[i] = v;
    return;
  }
  size_t half = (size_t)1 << (rounds - 1);
  for (size_t i = 0; i < nx; i++) {
    kx acc = K0;
    for (size_t p = 0; p < half; p++) {
      kx prod = K1;
      for (size_t f = 0; f < o->nf; f++) prod = kmul(prod, interp(o->factors[f][2 * p], o->factors[f][2 * p + 1], xs[i]));
      acc = kadd(acc, prod);
    }
    ys[i] = acc;
  }
}

static void addr_evals(or_ts* o, const kx* xs, size_t nx, kx* ys) {
  int is_read = o->kind == TS_READ_ADDR;
  for (size_t j = 0; j < nx; j++) ys[j] = K0;
  if (or_ts_num_rounds(o) == 0) {
    kx mem = K0;
    for (size_t i = 0; i < o->ninit; i++) mem = kadd(mem, kmul(o->init_val[i], o->init_w[i]));
    kx sum = K0;
    for (size_t e = 0; e < o->nev; e++) {
      if (is_read) {
        if (!kz(o->ev_hr[e]))
          sum = kadd(sum, kmul(kmul(kmul(o->eq_cycle[e], o->ev_hr[e]), o->ra_w[e]), ksub(mem, o->ev_rv[e])));
      } else {
        kx delta = ksub(ksub(o->ev_rv[e], mem), o->ev_inc[e]);
        sum = kadd(sum, kmul(kmul(kmul(o->eq_cycle[e], o->ev_hw[e]), o->wa_w[e]), delta));
      }
      if (!kz(o->ev_hw[e])) mem = kadd(mem, kmul(kmul(o->ev_inc[e], o->ev_hw[e]), o->wa_w[e]));
    }
    for (size_t j = 0; j < nx; j++) ys[j] = sum;
    return;
  }
  size_t bit = o->bit_idx;
  size_t nb = ((size_t)1 << o->ell_addr) >> bit;
  for (size_t x = 0; x < nb; x++) o->mem[x] = K0;
  for (size_t i = 0; i < o->ninit; i++) {
    kx c = kmul(o->init_val[i], o->init_w[i]);
    if (!kz(c)) o->mem[o->init_addr[i] >> bit] = kadd(o->mem[o->init_addr[i] >> bit], c);
  }
  for (size_t e = 0; e < o->nev; e++) {
    kx eq_t = o->eq_cycle[e];
    kx gate = is_read ? o->ev_hr[e] : o->ev_hw[e];
    if (!kz(gate)) {
      uint64_t a = is_read ? o->ra_addr[e] : o->wa_addr[e];
      uint64_t base = a >> (bit + 1);
      kx v0 = o->mem[base * 2], v1 = o->mem[base * 2 + 1];
      kx dv = ksub(v1, v0);
      kx pre = is_read ? o->ra_w[e] : o->wa_w[e];
      int bt = (a >> bit) & 1;
      for (size_t j = 0; j < nx; j++) {
        kx x = xs[j];
        kx vx = kadd(v0, kmul(dv, x));
        kx af = bt ? x : ksub(K1, x);
        kx diff = is_read ? ksub(vx, o->ev_rv[e]) : ksub(ksub(o->ev_rv[e], vx), o->ev_inc[e]);
        ys[j] = kadd(ys[j], kmul(kmul(kmul(kmul(eq_t, gate), pre), af), diff));
      }
    }
    if (!kz(o->ev_hw[e])) {
      uint64_t idx = o->wa_addr[e] >> bit;
      kx delta = kmul(kmul(o->ev_inc[e], o->ev_hw[e]), o->wa_w[e]);
      if (!kz(delta)) o->mem[idx] = kadd(o->mem[idx], delta);
    }
  }
}

void or_ts_evals(or_ts* o, const kx* xs, size_t nx, kx* ys) {
  if (o->kind == TS_PRODUCT || o->kind == TS_ADDR_LOOKUP) {
    product_evals(o, xs, nx, ys);
    return;
  }
  if (o->kind == TS_READ_ADDR || o->kind == TS_WRITE_ADDR) {
    addr_evals(o, xs, nx, ys);
    return;
  }
  for (size_t i = 0; i < nx; i++) ys[i] = K0;
  const size_t ell = o->ell_n, bi = o->bit_idx;
  if (o->kind == TS_BITNESS) {
    if (o->nbits == 0) return;
    if (o->bits[0].len == 1) {
      kx acc = K0;
      for (size_t c = 0; c < o->nbits; c++) {
        kx b = spv_single(&o->bits[c]);
        acc = kadd(acc, kmul(kmul(o->weights[c], b), ksub(b, K1)));
      }
      kx v = kmul(o->prefix_eq, acc);
      for (size_t i = 0; i < nx; i++) ys[i] = v;
      return;
    }
    size_t tot = 0;
    for (size_t c = 0; c < o->nbits; c++) tot += o->bits[c].n;
    uint64_t* pairs = (uint64_t*)malloc((tot + 1) * sizeof(uint64_t));
    size_t np = 0;
    for (size_t c = 0; c < o->nbits; c++)
      for (size_t e = 0; e < o->bits[c].n; e++) pairs[np++] = o->bits[c].idx[e] >> 1;
    qsort(pairs, np, sizeof(uint64_t), cmp_entry);
    size_t nd = 0;
    for (size_t e = 0; e < np; e++)
      if (nd == 0 || pairs[nd - 1] != pairs[e]) pairs[nd++] = pairs[e];
    for (size_t q = 0; q < nd; q++) {
      uint64_t pair = pairs[q], c0i = 2 * pair, c1i = c0i + 1;
      kx chi0, chi1;
      chi_cycle_children(o->r_cycle, ell, bi, o->prefix_eq, pair, &chi0, &chi1);
      for (size_t i = 0; i < nx; i++) {
        kx chi_x = interp(chi0, chi1, xs[i]);
        if (kz(chi_x)) continue;
        kx bs = K0;
        for (size_t c = 0; c < o->nbits; c++) {
          kx b0 = spv_get(&o->bits[c], c0i), b1 = spv_get(&o->bits[c], c1i);
          if (kz(b0) && kz(b1)) continue;
          kx bx = interp(b0, b1, xs[i]);
          bs = kadd(bs, kmul(kmul(o->weights[c], bx), ksub(bx, K1)));
        }
        ys[i] = kadd(ys[i], kmul(chi_x, bs));
      }
    }
    free(pairs);
    return;
  }
  if (o->gate.len == 1) {
    kx eq = K1;
    for (size_t b = 0; b < o->nbits; b++) eq = kmul(eq, eq_bit_affine(spv_single(&o->bits[b]), o->r_addr[b]));
    kx v = K0;
    switch (o->kind) {
      case TS_SHOUT_VALUE:
        v = kmul(kmul(o->prefix_eq, spv_single(&o->gate)), spv_single(&o->a));
        break;
      case TS_INDEX_ADAPTER:
        v = kmul(kmul(o->prefix_eq, spv_single(&o->gate)), eq);
        break;
      case TS_TWIST_READ: {
        kx vp = val_pre_from_terms(o->init_at_r_addr, o->terms, o->nterms, o->chals, o->nchals);
        v = kmul(kmul(kmul(o->prefix_eq, spv_single(&o->gate)), ksub(vp, spv_single(&o->a))), eq);
        break;
      }
      case TS_TWIST_WRITE: {
        kx vp = val_pre_from_terms(o->init_at_r_addr, o->terms, o->nterms, o->chals, o->nchals);
        kx delta = ksub(ksub(spv_single(&o->a), vp), spv_single(&o->b));
        v = kmul(kmul(kmul(o->prefix_eq, spv_single(&o->gate)), delta), eq);
        break;
      }
      case TS_VAL_EVAL: {
        kx lt = lt_eval(o->chals, o->r_cycle, ell);
        v = kmul(kmul(kmul(spv_single(&o->gate), spv_single(&o->b)), eq), lt);
        break;
      }
      case TS_TOTAL_INC:
        v = kmul(kmul(spv_single(&o->gate), spv_single(&o->b)), eq);
        break;
    }
    for (size_t i = 0; i < nx; i++) ys[i] = v;
    return;
  }
  uint64_t* pairs = (uint64_t*)malloc((o->gate.n + 1) * sizeof(uint64_t));
  size_t np = gather_pairs(&o->gate, pairs);
  kx* e0s = (kx*)malloc((o->nbits + 1) * sizeof(kx));
  kx* des = (kx*)malloc((o->nbits + 1) * sizeof(kx));
  for (size_t q = 0; q < np; q++) {
    uint64_t pair = pairs[q], c0i = 2 * pair, c1i = c0i + 1;
    kx g0 = spv_get(&o->gate, c0i), g1 = spv_get(&o->gate, c1i);
    if (kz(g0) && kz(g1)) continue;
    kx f0 = K0, f1 = K0; /* kind-specific extra linear factor (val / diff / delta / inc) */
    kx h0 = K1, h1 = K1; /* second extra factor (lt for VAL_EVAL) */
    int has_chi = o->kind != TS_VAL_EVAL && o->kind != TS_TOTAL_INC;
    kx chi0 = K1, chi1 = K1;
    if (has_chi) chi_cycle_children(o->r_cycle, ell, bi, o->prefix_eq, pair, &chi0, &chi1);
    switch (o->kind) {
      case TS_SHOUT_VALUE:
        f0 = spv_get(&o->a, c0i);
        f1 = spv_get(&o->a, c1i);
        break;
      case TS_INDEX_ADAPTER:
        f0 = K1;
        f1 = K1;
        break;
      case TS_TWIST_READ:
      case TS_TWIST_WRITE: {
        fill_time_point(o->tc0, ell, o->chals, bi, K0, pair);
        fill_time_point(o->tc1, ell, o->chals, bi, K1, pair);
        kx vp0 = val_pre_from_terms(o->init_at_r_addr, o->terms, o->nterms, o->tc0, ell);
        kx vp1 = val_pre_from_terms(o->init_at_r_addr, o->terms, o->nterms, o->tc1, ell);
        if (o->kind == TS_TWIST_READ) {
          f0 = ksub(vp0, spv_get(&o->a, c0i));
          f1 = ksub(vp1, spv_get(&o->a, c1i));
        } else {
          f0 = ksub(ksub(spv_get(&o->a, c0i), vp0), spv_get(&o->b, c0i));
          f1 = ksub(ksub(spv_get(&o->a, c1i), vp1), spv_get(&o->b, c1i));
        }
        break;
      }
      case TS_VAL_EVAL:
        f0 = spv_get(&o->b, c0i);
        f1 = spv_get(&o->b, c1i);
        fill_time_point(o->tc0, ell, o->chals, bi, K0, pair);
        fill_time_point(o->tc1, ell, o->chals, bi, K1, pair);
        h0 = lt_eval(o->tc0, o->r_cycle, ell);
        h1 = lt_eval(o->tc1, o->r_cycle, ell);
        break;
      case TS_TOTAL_INC:
        f0 = spv_get(&o->b, c0i);
        f1 = spv_get(&o->b, c1i);
        break;
    }
    for (size_t b = 0; b < o->nbits; b++) {
      kx e0 = eq_bit_affine(spv_get(&o->bits[b], c0i), o->r_addr[b]);
      kx e1 = eq_bit_affine(spv_get(&o->bits[b], c1i), o->r_addr[b]);
      e0s[b] = e0;
      des[b] = ksub(e1, e0);
    }
    for (size_t i = 0; i < nx; i++) {
      kx x = xs[i];
      kx prod = kmul(interp(chi0, chi1, x), interp(g0, g1, x));
      if (o->kind != TS_INDEX_ADAPTER) prod = kmul(prod, interp(f0, f1, x));
      if (o->kind == TS_VAL_EVAL) prod = kmul(prod, interp(h0, h1, x));
      for (size_t b = 0; b < o->nbits; b++) prod = kmul(prod, kadd(e0s[b], kmul(des[b], x)));
      ys[i] = kadd(ys[i], prod);
    }
  }
  free(e0s);
  free(des);
  free(pairs);
}

static void upd_prefix(kx* w, const uint64_t* addrs, size_t n, size_t bit, kx r) {
  kx r0 = ksub(K1, r);
  for (size_t i = 0; i < n; i++) w[i] = kmul(w[i], ((addrs[i] >> bit) & 1) ? r : r0);
}

void or_ts_fold(or_ts* o, kx r) {
  switch (o->kind) {
    case TS_PRODUCT:
    case TS_ADDR_LOOKUP: {
      size_t rounds = or_ts_num_rounds(o);
      if (rounds == 0) return;
      size_t half = (size_t)1 << (rounds - 1);
      for (size_t f = 0; f < o->nf; f++)
        for (size_t i = 0; i < half; i++) o->factors[f][i] = interp(o->factors[f][2 * i], o->factors[f][2 * i + 1], r);
      o->flen = half;
      o->chals[o->nchals++] = r;
      return;
    }
    case TS_READ_ADDR:
    case TS_WRITE_ADDR:
      if (or_ts_num_rounds(o) == 0) return;
      upd_prefix(o->init_w, o->init_addr, o->ninit, o->bit_idx, r);
      if (o->kind == TS_READ_ADDR) upd_prefix(o->ra_w, o->ra_addr, o->nev, o->bit_idx, r);
      upd_prefix(o->wa_w, o->wa_addr, o->nev, o->bit_idx, r);
      o->bit_idx++;
      return;
    case TS_TOTAL_INC:
      if (o->gate.len == 1) return;
      spv_fold(&o->gate, r);
      spv_fold(&o->b, r);
      for (size_t b = 0; b < o->nbits; b++) spv_fold(&o->bits[b], r);
      return;
    default:
      break;
  }
  if (or_ts_num_rounds(o) == 0) return;
  if (o->kind != TS_VAL_EVAL) o->prefix_eq = kmul(o->prefix_eq, eq_single(r, o->r_cycle[o->bit_idx]));
  if (o->kind != TS_BITNESS) spv_fold(&o->gate, r);
  if (o->kind == TS_SHOUT_VALUE || o->kind == TS_TWIST_READ || o->kind == TS_TWIST_WRITE) spv_fold(&o->a, r);
  if (o->kind == TS_TWIST_WRITE || o->kind == TS_VAL_EVAL) spv_fold(&o->b, r);
  for (size_t b = 0; b < o->nbits; b++) spv_fold(&o->bits[b], r);
  o->chals[o->nchals++] = r;
  o->bit_idx++;
}

void or_ts_free(or_ts* o) {
  if (!o) return;
  if (o->gate.idx) spv_free(&o->gate);
  if (o->a.idx) spv_free(&o->a);
  if (o->b.idx) spv_free(&o->b);
  for (size_t b = 0; b < o->nbits; b++) spv_free(&o->bits[b]);
  free(o->bits);
  free(o->weights);
  free(o->terms);
  free(o->r_cycle);
  free(o->r_addr);
  free(o->chals);
  free(o->tc0);
  free(o->tc1);
  for (size_t f = 0; f < o->nf; f++) free(o->factors[f]);
  free(o->factors);
  free(o->eq_cycle);
  free(o->ev_hr);
  free(o->ev_rv);
  free(o->ev_hw);
  free(o->ev_inc);
  free(o->ra_w);
  free(o->wa_w);
  free(o->ra_addr);
  free(o->wa_addr);
  free(o->init_addr);
  free(o->init_val);
  free(o->init_w);
  free(o->mem);
  free(o);
}
