// twist.hip -- Twist/Shout round oracles on the device (neo-memory/src/twist_oracle.rs).
//
// Time lanes (ShoutValue, IndexAdapter, LazyWeightedBitness, Twist read/write checks, ValEval,
// TotalInc; twist_oracle.rs:200-1196): one union support of the sparse columns, a pair kernel
// that builds each sibling pair's round polynomial as a product of linear factors and reduces
// the coefficients, and a fold that writes parents through an exclusive-scan compaction
// (sparse_time.rs:69-116).  Twist's Val_pre(t) at a pair's children is
// init + sum_{u < block} inc_u + sum_{u in block} inc_u * LT_low(u, t): terms below the pair's
// time block contribute exactly 1, terms above exactly 0 (twist_oracle.rs:328-342), so each
// round touches every increment term once instead of once per pair.
// Address lanes (twist_oracle.rs:1259-1660): per round, a stable radix sort of the write events
// by address bucket and a K-valued inclusive scan replace the sequential hash-map replay; each
// read queries its two sibling buckets by binary search.
// Product / address-lookup (twist_oracle.rs:56-183, 1662-1736): dense factors, pair kernel and
// in-place fold; the lookup's weight table is built by sorting lookups by address.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <map>
#include <mutex>

#include "tspool.h"
#include "twist.h"

namespace ns {

constexpr int TB = 256;
constexpr int TMAXBLK = 256;

__device__ __forceinline__ K2 kinterp(K2 f0, K2 f1, K2 x) { return kadd(f0, kmul(ksub(f1, f0), x)); }
__device__ __forceinline__ K2 eq1(K2 a, K2 b) {
  return kadd(kmul(ksub(kone(), a), ksub(kone(), b)), kmul(a, b));
}
// bit_ops.rs:5-9
__device__ __forceinline__ K2 eq_bit_affine(K2 bit, K2 u) {
  return kadd(kmul(bit, ksub(kadd(u, u), kone())), ksub(kone(), u));
}

template <int MAXC>
__device__ __forceinline__ void poly_mul_lin(K2 (&p)[MAXC], int& deg, K2 a, K2 b) {
  // p <- p * (a + b x)
#pragma unroll
  for (int d = MAXC - 1; d >= 1; d--)
    if (d <= deg + 1) p[d] = kadd(kmul(p[d], a), kmul(p[d - 1], b));
  p[0] = kmul(p[0], a);
  deg++;
}

__device__ __forceinline__ K2 wave_sum(K2 v) {
  for (int off = 32; off > 0; off >>= 1) {
    K2 o;
    o.c0 = __shfl_down(v.c0, off, 64);
    o.c1 = __shfl_down(v.c1, off, 64);
    v = kadd(v, o);
  }
  return v;
}

// Block-reduce `ncoef` coefficients into partials[blk * ncoef + d].
template <int MAXC>
__device__ void block_store_at(K2 (&acc)[MAXC], int ncoef, K2* partials, uint32_t blk) {
  __shared__ K2 red[TB / 64][MAXC];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < MAXC; d++) {
    if (d < ncoef) {
      K2 v = wave_sum(acc[d]);
      if (lane == 0) red[wave][d] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < ncoef) {
    K2 v = kzero();
    for (int w = 0; w < TB / 64; w++) v = kadd(v, red[w][threadIdx.x]);
    partials[blk * ncoef + threadIdx.x] = v;
  }
}
template <int MAXC>
__device__ void block_store(K2 (&acc)[MAXC], int ncoef, K2* partials) {
  block_store_at<MAXC>(acc, ncoef, partials, blockIdx.x);
}

// ---------------- one-workgroup paths for supports of at most ONEBLK_N entries ----------------
constexpr int ONEBLK_T = 1024, ONEBLK_PER = 8, ONEBLK_N = ONEBLK_T * ONEBLK_PER;

__device__ __forceinline__ uint32_t shfl_up_u32(uint32_t v, int off) { return __shfl_up(v, off, 64); }
__device__ __forceinline__ K2 shfl_up_k2(K2 v, int off) {
  return K2{__shfl_up(v.c0, off, 64), __shfl_up(v.c1, off, 64)};
}
__device__ __forceinline__ uint32_t add_u32(uint32_t a, uint32_t b) { return a + b; }
__device__ __forceinline__ K2 add_k2(K2 a, K2 b) { return kadd(a, b); }

// exclusive prefix of v over a ONEBLK_T-thread block, in thread order; `wsum` holds
// ONEBLK_T / 64 values of scratch LDS.  Ends with a barrier, so wsum can be reused.
template <class T, T (*SHFL)(T, int), T (*ADD)(T, T)>
__device__ __forceinline__ T block_exclusive(T v, T zero, T* wsum, T& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const T y = SHFL(x, off);
    if (lane >= off) x = ADD(x, y);
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  T before = zero, tot = zero;
  for (int j = 0; j < ONEBLK_T / 64; j++) {
    if (j < w) before = ADD(before, wsum[j]);
    tot = ADD(tot, wsum[j]);
  }
  total = tot;
  T ex = SHFL(x, 1);
  if (lane == 0) ex = zero;
  __syncthreads();
  return ADD(before, ex);
}

__global__ void k_ts_sum(const K2* __restrict__ partials, int nblk, int ncoef, K2* __restrict__ out) {
  const int d = blockIdx.x;
  K2 v = kzero();
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) v = kadd(v, partials[b * ncoef + d]);
  __shared__ K2 red[4];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[d] = kadd(kadd(red[0], red[1]), kadd(red[2], red[3]));
}

// ---------------- time lanes ----------------
struct TsTimeArgs {
  int kind;
  uint32_t n, stride;
  const uint32_t* idx;
  const K2* cols;  // [col][stride]: 0 gate, col_a, col_b, then nbits bit columns at bit0
  int col_a, col_b, bit0, nbits;
  const K2* consts;  // r_cycle[ell] | chals[ell] | r_addr[nbits] | weights[nbits]
  int ell, bit_idx;
  K2 prefix_eq;
  K2 init_at;
  const uint32_t* term_t;
  const K2* term_v;
  const K2* term_p;
  const uint32_t* term_rank;  // [len + 1]: number of terms before each time, or null (search)
  const K2* lt_s0;  // inclusive prefix sums over the terms of term_v * lt_low(term_t, bit_idx, c)
  const K2* lt_s1;
  uint32_t nterms;
  int ncoef;
  const uint32_t* n_dev;  // when set, the support size is read here (n is then an upper bound)
  int cgroups;            // bitness: threads per pair, each summing every cgroups-th bit column
};

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint64_t key) {
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// LT(u, t) restricted to positions b..0 (t_b = c Boolean, t_i = chals[i] below):
// lt_eval_at_bool_index (twist_oracle.rs:328-342) once every higher bit matched.
__device__ K2 lt_low(uint32_t u, int b, int c, const K2* chals) {
  K2 suffix = kone(), acc = kzero();
  if (((u >> b) & 1) == 0) {
    if (c) acc = suffix;
    suffix = c ? kzero() : suffix;
  } else {
    suffix = c ? suffix : kzero();
  }
  for (int i = b - 1; i >= 0; i--) {
    K2 t = chals[i];
    if (((u >> i) & 1) == 0) {
      acc = kadd(acc, kmul(t, suffix));
      suffix = kmul(suffix, ksub(kone(), t));
    } else {
      suffix = kmul(suffix, t);
    }
  }
  return acc;
}

// val_pre at both child time points (twist_oracle.rs:313-361): every term before the pair's time
// block contributes its full value (exclusive prefix term_p), the terms inside the block their
// LT_low-weighted value, summed as a difference of the per-term prefix sums lt_s0 / lt_s1.  The
// block's term range comes from the rank table when there is one, else from two binary searches.
__device__ __forceinline__ void val_pre_pair(const TsTimeArgs& A, uint64_t pair, K2& vp0, K2& vp1) {
  const int b = A.bit_idx;
  const uint64_t lo_key = pair << (b + 1), hi_key = (pair + 1) << (b + 1);
  uint32_t lo, hi;
  if (A.term_rank) {
    lo = A.term_rank[lo_key];
    hi = A.term_rank[hi_key];
  } else {
    lo = lower_bound_u32(A.term_t, 0, A.nterms, lo_key);
    hi = lower_bound_u32(A.term_t, lo, A.nterms, hi_key);
  }
  const K2 v = kadd(A.init_at, A.term_p[lo]);
  vp0 = vp1 = v;
  if (hi > lo) {
    if (lo > 0) {
      vp0 = kadd(v, ksub(A.lt_s0[hi - 1], A.lt_s0[lo - 1]));
      vp1 = kadd(v, ksub(A.lt_s1[hi - 1], A.lt_s1[lo - 1]));
    } else {
      vp0 = kadd(v, A.lt_s0[hi - 1]);
      vp1 = kadd(v, A.lt_s1[hi - 1]);
    }
  }
}

// per-term LT_low contributions of one round, for both children (one thread per term)
__device__ __forceinline__ void lt_contrib_at(const uint32_t* __restrict__ term_t, const K2* __restrict__ term_v,
                                              uint32_t nterms, int b, const K2* __restrict__ chals,
                                              K2* __restrict__ out0, K2* __restrict__ out1, uint32_t k) {
  if (k >= nterms) return;
  const K2 v = term_v[k];
  out0[k] = kmul(v, lt_low(term_t[k], b, 0, chals));
  out1[k] = kmul(v, lt_low(term_t[k], b, 1, chals));
}
__global__ void k_lt_contrib(const uint32_t* __restrict__ term_t, const K2* __restrict__ term_v, uint32_t nterms,
                             int b, const K2* __restrict__ chals, K2* __restrict__ out0, K2* __restrict__ out1) {
  lt_contrib_at(term_t, term_v, nterms, b, chals, out0, out1, blockIdx.x * blockDim.x + threadIdx.x);
}

// inclusive prefix sums of two K arrays of n <= ONEBLK_N entries in one launch (thread t owns
// entries [t*PER, t*PER+PER))
__device__ __forceinline__ void scan2_1blk_body(const K2* __restrict__ in0, const K2* __restrict__ in1, uint32_t n,
                                                K2* __restrict__ out0, K2* __restrict__ out1) {
  __shared__ K2 wsum[ONEBLK_T / 64];
  const uint32_t k0 = threadIdx.x * ONEBLK_PER;
  for (int c = 0; c < 2; c++) {
    const K2* in = c ? in1 : in0;
    K2 loc[ONEBLK_PER];
    K2 run = kzero();
#pragma unroll
    for (int j = 0; j < ONEBLK_PER; j++) {
      if (k0 + j < n) run = kadd(run, in[k0 + j]);
      loc[j] = run;
    }
    K2 total;
    const K2 base = block_exclusive<K2, shfl_up_k2, add_k2>(run, kzero(), wsum, total);
    K2* out = c ? out1 : out0;
#pragma unroll
    for (int j = 0; j < ONEBLK_PER; j++)
      if (k0 + j < n) out[k0 + j] = kadd(base, loc[j]);
  }
}
__global__ __launch_bounds__(ONEBLK_T) void k_scan2_1blk(const K2* __restrict__ in0, const K2* __restrict__ in1,
                                                         uint32_t n, K2* __restrict__ out0, K2* __restrict__ out1) {
  scan2_1blk_body(in0, in1, n, out0, out1);
}

// lt_eval (mle.rs:7-24) at the child time point (chals | c | bits of pair) against r_time
__device__ K2 lt_child(const TsTimeArgs& A, uint64_t pair, int c, const K2* r_time, const K2* chals) {
  K2 suffix = kone(), acc = kzero();
  for (int i = A.ell - 1; i >= 0; i--) {
    K2 jp;
    if (i < A.bit_idx) jp = chals[i];
    else if (i == A.bit_idx) jp = c ? kone() : kzero();
    else jp = ((pair >> (i - A.bit_idx - 1)) & 1) ? kone() : kzero();
    K2 j = r_time[i];
    acc = kadd(acc, kmul(kmul(ksub(kone(), jp), j), suffix));
    suffix = kmul(suffix, eq1(jp, j));
  }
  return acc;
}

// pair of union-support entry i and its children's positions (0xFFFFFFFF where absent); false
// for the odd sibling of a present even child (the even one owns the pair)
__device__ __forceinline__ bool ts_pair_at(const TsTimeArgs& A, uint32_t i, uint64_t& pair, uint32_t& i0,
                                           uint32_t& i1) {
  const uint32_t id = A.idx[i];
  pair = id >> 1;
  if (i > 0 && (A.idx[i - 1] >> 1) == pair) return false;
  i0 = i1 = 0xFFFFFFFFu;
  if ((id & 1) == 0) {
    i0 = i;
    if (i + 1 < A.n && A.idx[i + 1] == id + 1) i1 = i + 1;
  } else {
    i1 = i;
  }
  return true;
}

__device__ __forceinline__ K2 ts_col(const TsTimeArgs& A, int c, uint32_t which) {
  return which != 0xFFFFFFFFu ? A.cols[(size_t)c * A.stride + which] : kzero();
}

// chi_cycle_children (twist_oracle.rs:185-198)
__device__ __forceinline__ void ts_chi(const TsTimeArgs& A, uint64_t pair, K2& chi0, K2& chi1) {
  const K2* r_cycle = A.consts;
  K2 suffix = kone();
  int shift = 0;
  for (int b = A.bit_idx + 1; b < A.ell; b++, shift++)
    suffix = kmul(suffix, ((pair >> shift) & 1) ? r_cycle[b] : ksub(kone(), r_cycle[b]));
  const K2 r = r_cycle[A.bit_idx];
  chi0 = kmul(kmul(A.prefix_eq, ksub(kone(), r)), suffix);
  chi1 = kmul(kmul(A.prefix_eq, r), suffix);
}

// The summand of entry i's pair, for every kind but bitness, as a product of factors linear in
// the round variable X: fac(value at X = 0, slope) once per factor.  false when another entry
// owns the pair or the gate is zero on both children (the summand is zero).
template <class F>
__device__ __forceinline__ bool ts_pair_factors(const TsTimeArgs& A, uint32_t i, F&& fac) {
  const K2* r_cycle = A.consts;
  const K2* chals = A.consts + A.ell;
  const K2* r_addr = A.consts + 2 * A.ell;
  uint64_t pair;
  uint32_t i0, i1;
  if (!ts_pair_at(A, i, pair, i0, i1)) return false;
  const K2 g0 = ts_col(A, 0, i0), g1 = ts_col(A, 0, i1);
  if (kiszero(g0) && kiszero(g1)) return false;
  fac(g0, ksub(g1, g0));
  if (A.kind != TSK_VAL_EVAL && A.kind != TSK_TOTAL_INC) {
    K2 chi0, chi1;
    ts_chi(A, pair, chi0, chi1);
    fac(chi0, ksub(chi1, chi0));
  }
  K2 f0 = kzero(), f1 = kzero();
  bool has_f = true;
  switch (A.kind) {
    case TSK_SHOUT_VALUE:
      f0 = ts_col(A, A.col_a, i0);
      f1 = ts_col(A, A.col_a, i1);
      break;
    case TSK_INDEX_ADAPTER:
      has_f = false;
      break;
    case TSK_TWIST_READ: {
      K2 vp0, vp1;
      val_pre_pair(A, pair, vp0, vp1);
      f0 = ksub(vp0, ts_col(A, A.col_a, i0));
      f1 = ksub(vp1, ts_col(A, A.col_a, i1));
      break;
    }
    case TSK_TWIST_WRITE: {
      K2 vp0, vp1;
      val_pre_pair(A, pair, vp0, vp1);
      f0 = ksub(ksub(ts_col(A, A.col_a, i0), vp0), ts_col(A, A.col_b, i0));
      f1 = ksub(ksub(ts_col(A, A.col_a, i1), vp1), ts_col(A, A.col_b, i1));
      break;
    }
    case TSK_VAL_EVAL: {
      f0 = ts_col(A, A.col_b, i0);
      f1 = ts_col(A, A.col_b, i1);
      const K2 h0 = lt_child(A, pair, 0, r_cycle, chals), h1 = lt_child(A, pair, 1, r_cycle, chals);
      fac(h0, ksub(h1, h0));
      break;
    }
    case TSK_TOTAL_INC:
      f0 = ts_col(A, A.col_b, i0);
      f1 = ts_col(A, A.col_b, i1);
      break;
    default:
      break;
  }
  if (has_f) fac(f0, ksub(f1, f0));
  for (int b = 0; b < A.nbits; b++) {
    const K2 e0 = eq_bit_affine(ts_col(A, A.bit0 + b, i0), r_addr[b]);
    const K2 e1 = eq_bit_affine(ts_col(A, A.bit0 + b, i1), r_addr[b]);
    fac(e0, ksub(e1, e0));
  }
  return true;
}

// running product of linear factors as a coefficient vector (degree < MAXC)
template <int MAXC>
struct LinProd {
  K2 p[MAXC];
  int deg;
  __device__ __forceinline__ void operator()(K2 a, K2 b) { poly_mul_lin<MAXC>(p, deg, a, b); }
};

// One time-lane round as coefficients (degree < MAXC): every thread multiplies out its pairs'
// products; bitness, whose summand is not a product of linear factors, runs only here.  The
// work of block blk of a grid of nblk blocks.
template <int MAXC>
__device__ __forceinline__ void ts_time_round_body(TsTimeArgs& A, uint32_t blk, uint32_t nblk, K2* partials) {
  if (A.n_dev) A.n = *A.n_dev;
  const K2* weights = A.consts + 2 * A.ell + A.nbits;
  K2 acc[MAXC];
#pragma unroll
  for (int d = 0; d < MAXC; d++) acc[d] = kzero();
  const uint32_t cg = (uint32_t)A.cgroups;
  for (uint32_t t = blk * blockDim.x + threadIdx.x; t < A.n * cg; t += nblk * blockDim.x) {
    const uint32_t i = t / cg;
    LinProd<MAXC> lp;
    K2(&p)[MAXC] = lp.p;
#pragma unroll
    for (int d = 0; d < MAXC; d++) p[d] = kzero();
    p[0] = kone();
    lp.deg = 0;
    if (A.kind == TSK_BITNESS) {
      // chi(x) * sum_c w_c b_c(x) (b_c(x) - 1), b_c(x) = b0 + db x
      uint64_t pair;
      uint32_t i0, i1;
      if (!ts_pair_at(A, i, pair, i0, i1)) continue;
      K2 chi0, chi1;
      ts_chi(A, pair, chi0, chi1);
      K2 q0 = kzero(), q1 = kzero(), q2 = kzero();
      // this thread's share of the columns (a pair's cgroups threads sum disjoint column sets;
      // the summand is linear in the column sum)
      for (int c = (int)(t % cg); c < A.nbits; c += (int)cg) {
        const K2 b0 = ts_col(A, A.bit0 + c, i0), b1 = ts_col(A, A.bit0 + c, i1);
        if (kiszero(b0) && kiszero(b1)) continue;
        const K2 db = ksub(b1, b0), w = weights[c];
        const K2 b0m1 = ksub(b0, kone());
        q0 = kadd(q0, kmul(w, kmul(b0, b0m1)));
        q1 = kadd(q1, kmul(w, kadd(kmul(b0, db), kmul(db, b0m1))));
        q2 = kadd(q2, kmul(w, kmul(db, db)));
      }
      const K2 dc = ksub(chi1, chi0);
      p[0] = kmul(chi0, q0);
      p[1] = kadd(kmul(chi0, q1), kmul(dc, q0));
      p[2] = kadd(kmul(chi0, q2), kmul(dc, q1));
      p[3] = kmul(dc, q2);
    } else if (!ts_pair_factors(A, i, lp)) {
      continue;
    }
#pragma unroll
    for (int d = 0; d < MAXC; d++) acc[d] = kadd(acc[d], p[d]);
  }
  block_store_at<MAXC>(acc, A.ncoef, partials, blk);
}
template <int MAXC>
__global__ __launch_bounds__(TB) void k_ts_time_round(TsTimeArgs A, K2* __restrict__ partials) {
  ts_time_round_body<MAXC>(A, blockIdx.x, gridDim.x, partials);
}

// One time-lane round by evaluation at X = blockIdx.y (0 .. ncoef-1): each one-wave block sums
// its pairs' products of linear factors at X (one multiply-add per factor, no coefficient
// arrays); partials[blockIdx.x * ncoef + X] receives the block's sum and the host interpolates
// the ncoef sums to the round's coefficients.
__device__ __forceinline__ void ts_time_points_body(TsTimeArgs& A, uint32_t blk, uint32_t nblk, uint32_t x,
                                                    K2* partials) {
  if (A.n_dev) A.n = *A.n_dev;
  const K2 X{(uint64_t)x, 0};
  K2 acc = kzero();
  for (uint32_t i = blk * 64 + threadIdx.x; i < A.n; i += nblk * 64) {
    K2 prod = kone();
    if (ts_pair_factors(A, i, [&](K2 a, K2 b) { prod = kmul(prod, kadd(a, kmul(b, X))); })) acc = kadd(acc, prod);
  }
  acc = wave_sum(acc);
  if (threadIdx.x == 0) partials[blk * A.ncoef + x] = acc;
}
__global__ __launch_bounds__(64) void k_ts_time_points(TsTimeArgs A, K2* __restrict__ partials) {
  ts_time_points_body(A, blockIdx.x, gridDim.x, blockIdx.y, partials);
}

// ---------------- ragged groups of time lanes ----------------
// The time lanes of one batched sum-check share a launch per kernel: up to TS_GROUP oracles,
// arguments by value, block b belonging to member o with blk0[o] <= b < blk0[o + 1] and running
// as block b - blk0[o] of a grid of blk0[o + 1] - blk0[o] blocks, as in the member's own launch.
struct TsRoundGroup {
  TsTimeArgs a[TS_GROUP];
  K2* partials[TS_GROUP];
  uint32_t blk0[TS_GROUP + 1];
  int count;
};
static_assert(sizeof(TsRoundGroup) <= 4096, "kernel arguments of a round group");
__device__ __forceinline__ int group_member(const uint32_t* blk0, int count, uint32_t b) {
  int o = 0;
  while (o + 1 < count && blk0[o + 1] <= b) o++;
  return o;
}
template <int MAXC>
__global__ __launch_bounds__(TB) void k_ts_time_round_group(TsRoundGroup G) {
  const int o = group_member(G.blk0, G.count, blockIdx.x);
  TsTimeArgs A = G.a[o];
  ts_time_round_body<MAXC>(A, blockIdx.x - G.blk0[o], G.blk0[o + 1] - G.blk0[o], G.partials[o]);
}
// grid.y = the largest member's coefficient count; rows past a member's own count have no work
__global__ __launch_bounds__(64) void k_ts_time_points_group(TsRoundGroup G) {
  const int o = group_member(G.blk0, G.count, blockIdx.x);
  TsTimeArgs A = G.a[o];
  if ((int)blockIdx.y >= A.ncoef) return;
  ts_time_points_body(A, blockIdx.x - G.blk0[o], G.blk0[o + 1] - G.blk0[o], blockIdx.y, G.partials[o]);
}
struct TsSumGroup {
  const K2* partials[TS_GROUP];
  K2* out[TS_GROUP];
  int nblk[TS_GROUP], ncoef[TS_GROUP];
  int count;
};
// k_ts_sum of every member: block (d, o) sums coefficient d of member o into its mailbox
__global__ void k_ts_sum_group(TsSumGroup S) {
  const int o = blockIdx.y, d = blockIdx.x;
  if (o >= S.count || d >= S.ncoef[o]) return;
  const int nblk = S.nblk[o], ncoef = S.ncoef[o];
  const K2* partials = S.partials[o];
  K2 v = kzero();
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) v = kadd(v, partials[b * ncoef + d]);
  __shared__ K2 red[4];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) S.out[o][d] = kadd(kadd(red[0], red[1]), kadd(red[2], red[3]));
}
// The Twist val_pre terms' LT contributions of a round and their prefix sums (at most ONEBLK_N
// terms per member): ragged blocks of 256 for the contributions, one block per member for the
// two scans.
struct TsLtArgs {
  const uint32_t* term_t;
  const K2* term_v;
  uint32_t nterms;
  int bit;
  const K2* chals;
  K2 *c0, *c1, *s0, *s1;
};
struct TsLtGroup {
  TsLtArgs m[TS_GROUP];
  uint32_t blk0[TS_GROUP + 1];
  int count;
};
static_assert(sizeof(TsLtGroup) <= 4096, "kernel arguments of an LT group");
__global__ void k_lt_contrib_group(TsLtGroup G) {
  const int o = group_member(G.blk0, G.count, blockIdx.x);
  const TsLtArgs& L = G.m[o];
  lt_contrib_at(L.term_t, L.term_v, L.nterms, L.bit, L.chals, L.c0, L.c1,
                (blockIdx.x - G.blk0[o]) * blockDim.x + threadIdx.x);
}
__global__ __launch_bounds__(ONEBLK_T) void k_scan2_1blk_group(TsLtGroup G) {
  const TsLtArgs& L = G.m[blockIdx.x];
  scan2_1blk_body(L.c0, L.c1, L.nterms, L.s0, L.s1);
}

// one K value into device memory (a round challenge into an oracle's constants block)
__global__ void k_put_k2(K2* __restrict__ dst, K2 v) {
  if (threadIdx.x == 0) *dst = v;
}

__global__ void k_ts_heads(const uint32_t* __restrict__ idx, uint32_t n, uint32_t* __restrict__ flags) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = (i == 0 || (idx[i - 1] >> 1) != (idx[i] >> 1)) ? 1u : 0u;
}

// parent p of the pair headed by entry i: its time index and every column folded at r
__device__ __forceinline__ void ts_fold_entry(const uint32_t* __restrict__ idx, const K2* __restrict__ cols,
                                              uint32_t n, uint32_t stride, int ncol, uint32_t i, uint32_t p, K2 r,
                                              uint32_t* __restrict__ nidx, K2* __restrict__ ncols) {
  const uint32_t id = idx[i];
  uint32_t i0 = 0xFFFFFFFFu, i1 = 0xFFFFFFFFu;
  if ((id & 1) == 0) {
    i0 = i;
    if (i + 1 < n && idx[i + 1] == id + 1) i1 = i + 1;
  } else {
    i1 = i;
  }
  nidx[p] = id >> 1;
  for (int c = 0; c < ncol; c++) {
    K2 v0 = i0 != 0xFFFFFFFFu ? cols[(size_t)c * stride + i0] : kzero();
    K2 v1 = i1 != 0xFFFFFFFFu ? cols[(size_t)c * stride + i1] : kzero();
    ncols[(size_t)c * stride + p] = kinterp(v0, v1, r);
  }
}

__device__ __forceinline__ void ts_fold_body(const uint32_t* __restrict__ idx, const K2* __restrict__ cols, uint32_t n,
                                             uint32_t stride, int ncol, const uint32_t* __restrict__ flags,
                                             const uint32_t* __restrict__ pos, K2 r, uint32_t* __restrict__ nidx,
                                             K2* __restrict__ ncols, uint32_t* __restrict__ nout,
                                             uint32_t* __restrict__ ndev, uint32_t i) {
  if (i >= n) return;
  if (i == n - 1) nout[0] = ndev[0] = pos[i] + flags[i];
  if (!flags[i]) return;
  ts_fold_entry(idx, cols, n, stride, ncol, i, pos[i], r, nidx, ncols);
}
__global__ void k_ts_fold(const uint32_t* __restrict__ idx, const K2* __restrict__ cols, uint32_t n, uint32_t stride,
                          int ncol, const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, K2 r,
                          uint32_t* __restrict__ nidx, K2* __restrict__ ncols, uint32_t* __restrict__ nout,
                          uint32_t* __restrict__ ndev) {
  ts_fold_body(idx, cols, n, stride, ncol, flags, pos, r, nidx, ncols, nout, ndev, blockIdx.x * blockDim.x + threadIdx.x);
}

// pair heads of a time-lane fold and their exclusive scan in one launch (n <= ONEBLK_N; thread t
// owns entries [t*PER, t*PER+PER)): flags[i], pos[i] and the folded support size pos[n]; thread
// 0 also stores the round challenge into the constants block when chal_dst is set
__device__ __forceinline__ void ts_heads_1blk_body(const uint32_t* __restrict__ idx, uint32_t n,
                                                   uint32_t* __restrict__ flags, uint32_t* __restrict__ pos,
                                                   K2* __restrict__ chal_dst, K2 r) {
  __shared__ uint32_t wsum[ONEBLK_T / 64];
  const uint32_t i0 = threadIdx.x * ONEBLK_PER;
  uint32_t f[ONEBLK_PER], cnt = 0;
#pragma unroll
  for (int k = 0; k < ONEBLK_PER; k++) {
    const uint32_t i = i0 + k;
    f[k] = (i < n && (i == 0 || (idx[i - 1] >> 1) != (idx[i] >> 1))) ? 1u : 0u;
    cnt += f[k];
  }
  uint32_t total;
  uint32_t p = block_exclusive<uint32_t, shfl_up_u32, add_u32>(cnt, 0u, wsum, total);
#pragma unroll
  for (int k = 0; k < ONEBLK_PER; k++)
    if (i0 + k < n) {
      flags[i0 + k] = f[k];
      pos[i0 + k] = p;
      p += f[k];
    }
  if (threadIdx.x == 0) {
    pos[n] = total;
    if (chal_dst) *chal_dst = r;
  }
}
__global__ __launch_bounds__(ONEBLK_T) void k_ts_heads_1blk(const uint32_t* __restrict__ idx, uint32_t n,
                                                            uint32_t* __restrict__ flags, uint32_t* __restrict__ pos,
                                                            K2* __restrict__ chal_dst, K2 r) {
  ts_heads_1blk_body(idx, n, flags, pos, chal_dst, r);
}

// A time-lane fold of at most ONEBLK_N entries (the one-workgroup heads scan, then the
// compaction), for the members of a group: one heads block per member, then the compactions of
// all members as ragged blocks of 256 threads.
struct TsFoldArgs {
  const uint32_t* idx;
  const K2* cols;
  uint32_t n, stride;
  int ncol;
  uint32_t* flags;
  uint32_t* pos;
  uint32_t* nidx;
  K2* ncols;
  uint32_t* nout;  // pinned mailbox word (the folded support size, for the host)
  uint32_t* ndev;  // device slot (the same, for a round enqueued behind the fold)
  K2* chal_dst;    // the challenge's slot in the constants block, or null
};
constexpr int TS_FOLD_GROUP = 32;  // folds per group launch (smaller arguments than a round's)
struct TsFoldGroup {
  TsFoldArgs f[TS_FOLD_GROUP];
  uint32_t blk0[TS_FOLD_GROUP + 1];
  int count;
  K2 r;
};
static_assert(sizeof(TsFoldGroup) <= 4096, "kernel arguments of a fold group");
__global__ __launch_bounds__(ONEBLK_T) void k_ts_heads_1blk_group(TsFoldGroup G) {
  const TsFoldArgs& F = G.f[blockIdx.x];
  ts_heads_1blk_body(F.idx, F.n, F.flags, F.pos, F.chal_dst, G.r);
}
__global__ __launch_bounds__(256) void k_ts_fold_group(TsFoldGroup G) {
  const int o = group_member(G.blk0, G.count, blockIdx.x);
  const TsFoldArgs& F = G.f[o];
  ts_fold_body(F.idx, F.cols, F.n, F.stride, F.ncol, F.flags, F.pos, G.r, F.nidx, F.ncols, F.nout, F.ndev,
               (blockIdx.x - G.blk0[o]) * 256 + threadIdx.x);
}

// sum_k term_v[k] * lt_eval_at_bool_index(term_t[k], chals) (the fully bound Val_pre)
__global__ __launch_bounds__(TB) void k_ts_valpre_full(const uint32_t* __restrict__ tt, const K2* __restrict__ tv,
                                                       uint32_t nterms, const K2* __restrict__ chals, int ell,
                                                       K2* __restrict__ partials) {
  K2 acc[1] = {kzero()};
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nterms; k += gridDim.x * blockDim.x) {
    uint32_t u = tt[k];
    K2 suffix = kone(), lt = kzero();
    for (int i = ell - 1; i >= 0; i--) {
      K2 t = chals[i];
      if (((u >> i) & 1) == 0) {
        lt = kadd(lt, kmul(t, suffix));
        suffix = kmul(suffix, ksub(kone(), t));
      } else {
        suffix = kmul(suffix, t);
      }
    }
    acc[0] = kadd(acc[0], kmul(tv[k], lt));
  }
  block_store<1>(acc, 1, partials);
}

// chi_at_index(r, t) * scale[i] for a list of times (scale may be null)
__global__ void k_ts_chi_at(const uint32_t* __restrict__ ts, uint32_t n, const K2* __restrict__ r, int ell,
                            const K2* __restrict__ scale, K2* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t t = ts[i];
  K2 acc = kone();
  for (int b = 0; b < ell; b++) acc = kmul(acc, ((t >> b) & 1) ? r[b] : ksub(kone(), r[b]));
  out[i] = scale ? kmul(acc, scale[i]) : acc;
}

// ---------------- product / lookup ----------------
template <int MAXC>
__global__ __launch_bounds__(TB) void k_prod_round(const K2* __restrict__ fac, int nf, size_t stride, size_t half,
                                                   K2* __restrict__ partials) {
  K2 acc[MAXC];
#pragma unroll
  for (int d = 0; d < MAXC; d++) acc[d] = kzero();
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < half; i += (size_t)gridDim.x * blockDim.x) {
    K2 p[MAXC];
#pragma unroll
    for (int d = 0; d < MAXC; d++) p[d] = kzero();
    p[0] = kone();
    int deg = 0;
    for (int f = 0; f < nf; f++) {
      K2 a = fac[f * stride + 2 * i], b = fac[f * stride + 2 * i + 1];
      poly_mul_lin<MAXC>(p, deg, a, ksub(b, a));
    }
#pragma unroll
    for (int d = 0; d < MAXC; d++) acc[d] = kadd(acc[d], p[d]);
  }
  block_store<MAXC>(acc, nf + 1, partials);
}

__global__ void k_prod_fold(const K2* __restrict__ src, K2* __restrict__ dst, int nf, size_t stride, size_t half, K2 r) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  int f = blockIdx.y;
  if (i >= half || f >= nf) return;
  dst[f * stride + i] = kinterp(src[f * stride + 2 * i], src[f * stride + 2 * i + 1], r);
}

// W[a] = sum of weights of the lookups at address a (lookups sorted by address)
__global__ void k_lookup_weights(const uint32_t* __restrict__ saddr, const uint32_t* __restrict__ sidx, uint32_t n,
                                 const K2* __restrict__ w, uint32_t pow2_addr, K2* __restrict__ W) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= pow2_addr) return;
  uint32_t lo = lower_bound_u32(saddr, 0, n, a), hi = lower_bound_u32(saddr, lo, n, (uint64_t)a + 1);
  K2 acc = kzero();
  for (uint32_t j = lo; j < hi; j++) acc = kadd(acc, w[sidx[j]]);
  W[a] = acc;
}

// ---------------- address lanes ----------------
struct KAddOp {
  __device__ __forceinline__ K2 operator()(const K2& a, const K2& b) const { return kadd(a, b); }
};

__global__ void k_addr_keys(const K2* __restrict__ hw, const K2* __restrict__ inc, const K2* __restrict__ ww,
                            const uint32_t* __restrict__ wa, uint32_t nev, int bit, uint32_t sentinel,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, K2* __restrict__ dval) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  bool w = !kiszero(hw[e]);
  keys[e] = w ? (bit >= 0 ? (wa[e] >> bit) : 0u) : sentinel;
  vals[e] = e;
  dval[e] = w ? kmul(kmul(inc[e], hw[e]), ww[e]) : kzero();
}

__global__ void k_addr_gather(const uint32_t* __restrict__ svals, const K2* __restrict__ dval, uint32_t nev,
                              K2* __restrict__ out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nev) return;
  out[j] = dval[svals[j]];
}

// k_addr_keys + the stable sort by bucket + k_addr_gather + the K scan in one workgroup, for at
// most ONEBLK_N events and ell_addr <= 18: the sort key is (bucket << 13) | event, unique, so the
// order is (bucket, event) whatever the sort's stability; padding keys are all ones (above every
// real key, whose bucket is at most 2^ell_addr).  Thread t ends with sorted ranks t*8 .. t*8+7.
constexpr int AP_EBITS = 13;  // event index bits (ONEBLK_N = 8192)
__global__ __launch_bounds__(ONEBLK_T) void k_addr_prep_1blk(const K2* __restrict__ hw, const K2* __restrict__ inc,
                                                             const K2* __restrict__ ww, const uint32_t* __restrict__ wa,
                                                             uint32_t nev, int bit, uint32_t sentinel, int kbits,
                                                             uint32_t* __restrict__ skeys, uint32_t* __restrict__ svals,
                                                             K2* __restrict__ scan) {
  using Sort = hipcub::BlockRadixSort<uint32_t, ONEBLK_T, ONEBLK_PER>;
  __shared__ typename Sort::TempStorage tmp;
  __shared__ K2 wsum[ONEBLK_T / 64];
  const int nbits = AP_EBITS + kbits;
  const uint32_t pad = nbits >= 32 ? 0xFFFFFFFFu : ((1u << nbits) - 1), emask = (1u << AP_EBITS) - 1;
  uint32_t key[ONEBLK_PER];
#pragma unroll
  for (int j = 0; j < ONEBLK_PER; j++) {
    const uint32_t e = threadIdx.x * ONEBLK_PER + j;
    if (e < nev) {
      const uint32_t b = !kiszero(hw[e]) ? (bit >= 0 ? (wa[e] >> bit) : 0u) : sentinel;
      key[j] = (b << AP_EBITS) | e;
    } else {
      key[j] = pad;
    }
  }
  Sort(tmp).Sort(key, 0, nbits);
  K2 loc[ONEBLK_PER];
  K2 run = kzero();
#pragma unroll
  for (int j = 0; j < ONEBLK_PER; j++) {
    if (threadIdx.x * ONEBLK_PER + j < nev) {
      const uint32_t e = key[j] & emask;
      if (!kiszero(hw[e])) run = kadd(run, kmul(kmul(inc[e], hw[e]), ww[e]));
    }
    loc[j] = run;
  }
  K2 total;
  const K2 base = block_exclusive<K2, shfl_up_k2, add_k2>(run, kzero(), wsum, total);
#pragma unroll
  for (int j = 0; j < ONEBLK_PER; j++) {
    const uint32_t i = threadIdx.x * ONEBLK_PER + j;
    if (i < nev) {
      skeys[i] = key[j] >> AP_EBITS;
      svals[i] = key[j] & emask;
      scan[i] = kadd(base, loc[j]);
    }
  }
}

// mem value of bucket k just before event e: it[k] + sum of bucket-k write deltas of earlier events
__device__ K2 bucket_before(const uint32_t* skeys, const uint32_t* svals, const K2* scan, uint32_t nev, uint32_t k,
                            uint32_t e, const K2* it) {
  uint32_t lo = lower_bound_u32(skeys, 0, nev, k), hi = lower_bound_u32(skeys, lo, nev, (uint64_t)k + 1);
  uint32_t j = lower_bound_u32(svals, lo, hi, e);  // first event >= e in the bucket
  K2 s = kzero();
  if (j > lo) s = ksub(scan[j - 1], lo > 0 ? scan[lo - 1] : kzero());
  return kadd(it[k], s);
}

struct AddrArgs {
  uint32_t nev;
  int is_read, bit, final_;
  const K2 *eq, *hr, *rv, *hw, *inc, *rw, *ww;  // rv holds wv for the write check
  const uint32_t *ra, *wa;
  const uint32_t *skeys, *svals;
  const K2* scan;
  const K2* it;
};

__global__ __launch_bounds__(TB) void k_addr_round(AddrArgs A, K2* __restrict__ partials) {
  K2 acc[3] = {kzero(), kzero(), kzero()};
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < A.nev; e += gridDim.x * blockDim.x) {
    K2 gate = A.is_read ? A.hr[e] : A.hw[e];
    if (kiszero(gate)) continue;
    uint32_t addr = A.is_read ? A.ra[e] : A.wa[e];
    K2 c = kmul(kmul(A.eq[e], gate), A.is_read ? A.rw[e] : A.ww[e]);
    if (A.final_) {
      K2 mem = bucket_before(A.skeys, A.svals, A.scan, A.nev, 0, e, A.it);
      K2 diff = A.is_read ? ksub(mem, A.rv[e]) : ksub(ksub(A.rv[e], mem), A.inc[e]);
      acc[0] = kadd(acc[0], kmul(c, diff));
      continue;
    }
    uint32_t k0 = (addr >> (A.bit + 1)) << 1;
    K2 v0 = bucket_before(A.skeys, A.svals, A.scan, A.nev, k0, e, A.it);
    K2 v1 = bucket_before(A.skeys, A.svals, A.scan, A.nev, k0 + 1, e, A.it);
    K2 dv = ksub(v1, v0);
    // read: (v0 - rv) + dv x ; write: (wv - inc - v0) - dv x
    K2 base = A.is_read ? ksub(v0, A.rv[e]) : ksub(ksub(A.rv[e], A.inc[e]), v0);
    K2 slope = A.is_read ? dv : kneg(dv);
    if ((addr >> A.bit) & 1) {  // x * (base + slope x)
      acc[1] = kadd(acc[1], kmul(c, base));
      acc[2] = kadd(acc[2], kmul(c, slope));
    } else {  // (1 - x) * (base + slope x)
      acc[0] = kadd(acc[0], kmul(c, base));
      acc[1] = kadd(acc[1], kmul(c, ksub(slope, base)));
      acc[2] = ksub(acc[2], kmul(c, slope));
    }
  }
  block_store<3>(acc, 3, partials);
}

__global__ void k_addr_fold_w(K2* __restrict__ w, const uint32_t* __restrict__ addr, uint32_t n, int bit, K2 r) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  w[e] = kmul(w[e], ((addr[e] >> bit) & 1) ? r : ksub(kone(), r));
}

// ---------------- shared CPU bus (memory_sidecar/cpu_bus.rs) ----------------
// z[c] = sum_rho 2^rho Z[rho][c]; |z| < 2^54 so the signed sum is exact in int64
__device__ __forceinline__ uint64_t decode_col(const int8_t* __restrict__ Z, size_t m, size_t c) {
  int64_t acc = 0;
  for (int rho = D - 1; rho >= 0; rho--) acc = 2 * acc + Z[(size_t)rho * m + c];
  return from_i64(acc);
}

__global__ void k_bus_decode(const int8_t* __restrict__ Z, size_t m, uint64_t* __restrict__ z) {
  size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c < m) z[c] = decode_col(Z, m, c);
}

// One block per bus column: decode its chunk cells and compact the nonzero ones (time-ordered).
__global__ __launch_bounds__(256) void k_bus_time_cols(const int8_t* __restrict__ Z, size_t m, uint32_t m_in,
                                                       size_t bus_base, uint32_t chunk, uint32_t steps_len,
                                                       uint32_t* __restrict__ idx_out, K2* __restrict__ vals_out,
                                                       uint32_t* __restrict__ nnz_out) {
  const uint32_t col = blockIdx.x;
  __shared__ uint32_t wave_cnt[4];
  __shared__ uint32_t base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t j0 = 0; j0 < steps_len; j0 += 256) {
    uint32_t j = j0 + threadIdx.x;
    uint64_t v = 0;
    if (j < steps_len) v = decode_col(Z, m, bus_base + (size_t)col * chunk + j);
    bool nz = j < steps_len && v != 0;
    unsigned long long bal = __ballot(nz);
    uint32_t before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wave_cnt[wave] = __popcll(bal);
    __syncthreads();
    uint32_t off = base;
    for (int w = 0; w < wave; w++) off += wave_cnt[w];
    if (nz) {
      uint32_t p = off + before;
      idx_out[(size_t)col * chunk + p] = m_in + j;
      vals_out[(size_t)col * chunk + p] = K2{v, 0};
    }
    __syncthreads();
    if (threadIdx.x == 0) base += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) nnz_out[col] = base;
}

// y[col][rho] = sum_j w[j] Z[rho][bus_cell(col, j)]: one block per (col, rho)
__global__ __launch_bounds__(256) void k_bus_open(const int8_t* __restrict__ Z, size_t m, size_t bus_base,
                                                  uint32_t chunk, const K2* __restrict__ w, K2* __restrict__ y) {
  const uint32_t col = blockIdx.x, rho = blockIdx.y;
  const int8_t* row = Z + (size_t)rho * m + bus_base + (size_t)col * chunk;
  K2 acc = kzero();
  for (uint32_t j = threadIdx.x; j < chunk; j += blockDim.x) {
    int8_t d = row[j];
    if (d > 0) acc = kadd(acc, w[j]);
    else if (d < 0) acc = ksub(acc, w[j]);
  }
  __shared__ K2 red[4];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) y[(size_t)col * 64 + rho] = kadd(kadd(red[0], red[1]), kadd(red[2], red[3]));
}

// y_scalar[col] = sum_rho 2^rho y[col][rho]; rows rho >= 54 of the 64-padded y are zeroed
__global__ void k_bus_scalars(K2* __restrict__ y, uint32_t bus_cols, K2* __restrict__ ys) {
  uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= bus_cols) return;
  K2 s = kzero();
  uint64_t pw = 1;
  for (int rho = 0; rho < D; rho++) {
    s = kadd(s, kscale(y[(size_t)col * 64 + rho], pw));
    pw = fadd(pw, pw);
  }
  for (int rho = D; rho < 64; rho++) y[(size_t)col * 64 + rho] = kzero();
  ys[col] = s;
}

// ======================= host =======================
static unsigned nblocks(size_t n, size_t cap = TMAXBLK) {
  size_t b = (n + TB - 1) / TB;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

// ---------------- block / stream / mailbox cache (one device per process) ----------------
namespace {
struct HipPoolBackend {
  using Stream = hipStream_t;
  static bool dev_alloc(void** p, size_t n) { return hipMalloc(p, n) == hipSuccess; }
  static void alloc_failed(size_t n) {
    (void)hipGetLastError();
    throw Error(1, "out of device memory for a " + std::to_string(n) + "-byte oracle block");
  }
  static void dev_free(void* p) { (void)hipFree(p); }
  static void host_alloc(void** p, size_t n) { NS_CHECK(hipHostMalloc(p, n, hipHostMallocDefault)); }
  static void host_free(void* p) { (void)hipHostFree(p); }
  static hipStream_t stream_create() {
    hipStream_t s = nullptr;
    NS_CHECK(hipStreamCreate(&s));
    return s;
  }
  static void clear_error() { (void)hipGetLastError(); }
};
struct TsPool {
  using Pool = BlockPool<HipPoolBackend, K2>;
  // never destroyed: no HIP calls during process exit
  static Pool& get() {
    static Pool* pool = new Pool(TS_MAX_COEF + 1, (size_t)4 << 30, (size_t)1 << 30);
    return *pool;
  }
  static size_t size_class(size_t n) { return Pool::size_class(n); }
};
}  // namespace

TsBuf::~TsBuf() {
  if (p) TsPool::get().give(p, bytes);
}

void TsBuf::ensure(size_t n) {
  if (n <= bytes && p) return;
  if (p) {
    // the old block may still be read by work in flight on the owner's stream
    NS_CHECK(hipDeviceSynchronize());
    TsPool::get().give(p, bytes);
    p = nullptr;
    bytes = 0;
  }
  const size_t cap = TsPool::size_class(n);
  p = TsPool::get().take(cap);
  bytes = cap;
}

// value of a sparse column at times queried in nondecreasing order (a forward walk instead of a
// binary search per query)
struct ColCursor {
  const SparseCol* c;
  size_t k = 0;
  explicit ColCursor(const SparseCol& col) : c(&col) {}
  K2 at(uint32_t t) {
    const auto& ix = c->idx;
    while (k < ix.size() && ix[k] < t) k++;
    return (k < ix.size() && ix[k] == t) ? c->val[k] : kzero();
  }
};

// Host-to-device copies of the constructors go through pinned blocks (a DMA copy, no runtime
// staging through pageable memory); the blocks stay with the oracle until its stream has
// synchronized, then return to the cache.
void* TsOracleGpu::stage(size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (stage_.empty() || stage_off_ + bytes > stage_.back().second) {
    const size_t cap = TsPool::size_class(std::max(bytes, (size_t)1 << 20));
    stage_.emplace_back(TsPool::get().pinned_take(cap), cap);
    stage_off_ = 0;
  }
  void* p = (char*)stage_.back().first + stage_off_;
  stage_off_ += bytes;
  return p;
}

void TsOracleGpu::stage_release() {
  for (auto& b : stage_) TsPool::get().pinned_give(b.first, b.second);
  stage_.clear();
  stage_off_ = 0;
}

void TsOracleGpu::h2d_staged(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  void* h = stage(bytes);
  memcpy(h, src, bytes);
  NS_CHECK(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, st_));
}

template <class T>
void TsOracleGpu::upload(TsBuf& d, const std::vector<T>& h, size_t min_elems) {
  size_t n = std::max(h.size(), min_elems);
  d.ensure(n * sizeof(T));
  h2d_staged(d.p, h.data(), h.size() * sizeof(T));
}

TsOracleGpu::TsOracleGpu() {
  own_st_ = st_ = TsPool::get().stream();
  coef_h_ = TsPool::get().mailbox();
}

// The stream of a batch: every member's work goes there from the batch's first round on, and
// the stream returns to the cache (drained) with its last member.
struct TsSharedStream {
  hipStream_t s = nullptr;
  TsSharedStream() : s(TsPool::get().stream()) {}
  ~TsSharedStream() {
    (void)hipStreamSynchronize(s);
    TsPool::get().put_stream(s);
  }
};

// the stream the oracle's work went to drains before its own stream, the mailbox and (as
// members, after the body) the device blocks go back to the cache
TsOracleGpu::~TsOracleGpu() {
  if (st_) (void)hipStreamSynchronize(st_);
  if (own_st_) TsPool::get().put_stream(own_st_);
  stage_release();
  if (coef_h_) TsPool::get().put_mailbox(coef_h_);
}

void ts_join(TsOracleGpu* const* os, size_t n) {
  if (n == 0) return;
  bool same = os[0]->group_ != nullptr;
  for (size_t i = 1; i < n && same; i++) same = os[i]->group_ == os[0]->group_;
  if (same) return;
  auto g = std::make_shared<TsSharedStream>();
  for (size_t i = 0; i < n; i++) {
    // whatever the oracle enqueued so far (its constructor's uploads, an earlier batch) is done
    NS_CHECK(hipStreamSynchronize(os[i]->st_));
    os[i]->group_ = g;
    os[i]->st_ = g->s;
  }
}

static SparseCol empty_col() { return SparseCol{}; }

void TsOracleGpu::init_product(const std::vector<std::vector<K2>>& factors, size_t degree_bound) {
  kind_ = TSK_PRODUCT;
  nf_ = factors.size();
  if (nf_ == 0 || nf_ + 1 > (size_t)TS_MAX_COEF) throw Error(1, "product oracle: 1..39 factors required");
  flen_ = factors[0].size();
  if (flen_ == 0 || (flen_ & (flen_ - 1))) throw Error(1, "product oracle: factor length must be a power of two");
  std::vector<K2> flat(nf_ * flen_);
  for (size_t f = 0; f < nf_; f++) {
    if (factors[f].size() != flen_) throw Error(1, "product oracle: all factors must have the same length");
    std::copy(factors[f].begin(), factors[f].end(), flat.begin() + f * flen_);
  }
  upload(fac_[0], flat);
  fac_[1].ensure(flat.size() * sizeof(K2));
  cur_ = 0;
  degree_bound_ = degree_bound;
  size_t l = 0;
  while (((size_t)1 << l) < flen_) l++;
  ell_n_ = l;
  len_ = flen_;
}

// union support of the columns, values aligned; columns: gate, [a], [b], bits...
void TsOracleGpu::init_time(TsKind kind, const std::vector<K2>& r_cycle, size_t ell_n, const SparseCol* gate,
                            const SparseCol* a, const SparseCol* b, const std::vector<SparseCol>& bits,
                            const std::vector<K2>& r_addr, const std::vector<K2>& weights, K2 init_at_r_addr,
                            const SparseCol* term_write, const SparseCol* term_inc,
                            const std::vector<SparseCol>* term_bits) {
  kind_ = kind;
  if (ell_n < 1 || ell_n > 30) throw Error(1, "time-lane oracle: 1 <= ell_n <= 30");
  ell_n_ = ell_n;
  len_ = (size_t)1 << ell_n;
  r_cycle_ = r_cycle;
  r_cycle_.resize(ell_n, kzero());
  r_addr_ = r_addr;
  nbits_ = bits.size();
  if (kind != TSK_BITNESS && r_addr_.size() != nbits_) throw Error(1, "time-lane oracle: r_addr length != bit columns");
  std::vector<const SparseCol*> cs;
  SparseCol none = empty_col();
  cs.push_back(gate ? gate : &none);
  col_a_ = col_b_ = -1;
  if (a) {
    col_a_ = (int)cs.size();
    cs.push_back(a);
  }
  if (b) {
    col_b_ = (int)cs.size();
    cs.push_back(b);
  }
  for (auto& c : bits) cs.push_back(&c);
  ncol_ = cs.size();
  // union support: a byte map over time when the domain is not much larger than the entries,
  // else sort + unique
  size_t total = 0;
  for (auto* c : cs) {
    if (c->idx.size() != c->val.size()) throw Error(1, "sparse column: idx/val length mismatch");
    for (size_t j = 0; j < c->idx.size(); j++) {
      if (c->idx[j] >= len_) throw Error(1, "sparse column index out of range");
      if (j && c->idx[j] <= c->idx[j - 1]) throw Error(1, "sparse column indices must be strictly increasing");
    }
    total += c->idx.size();
  }
  std::vector<uint32_t> u;
  u.reserve(total);
  if (len_ <= 16 * total + 4096) {
    std::vector<uint8_t> mark(len_, 0);
    for (auto* c : cs)
      for (auto t : c->idx) mark[t] = 1;
    for (size_t t = 0; t < len_; t++)
      if (mark[t]) u.push_back((uint32_t)t);
  } else {
    for (auto* c : cs) u.insert(u.end(), c->idx.begin(), c->idx.end());
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
  }
  n_ = u.size();
  cap_ = std::max<size_t>(n_, 1);
  std::vector<K2> vals(ncol_ * cap_, kzero());
  for (size_t c = 0; c < ncol_; c++) {
    const SparseCol* col = cs[c];
    size_t p = 0;  // every index of the column is in u, both ascending
    for (size_t j = 0; j < col->idx.size(); j++) {
      while (u[p] < col->idx[j]) p++;
      vals[c * cap_ + p] = col->val[j];
    }
  }
  upload(idx_[0], u);
  upload(cols_[0], vals);
  idx_[1].ensure(std::max<size_t>(n_, 1) * 4);
  cols_[1].ensure(vals.size() * sizeof(K2));
  flags_.ensure(std::max<size_t>(n_, 1) * 4 + 8);
  pos_.ensure(std::max<size_t>(n_, 1) * 4 + 8);
  size_t tmp = 0;
  NS_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)std::max<size_t>(n_, 1), st_));
  scan_tmp_.ensure(tmp + 16);
  scan_tmp_bytes_ = tmp;
  cur_ = 0;
  // constants: r_cycle | chals | r_addr | weights
  consts_.ensure((2 * ell_n + 2 * std::max(nbits_, weights.size()) + 4) * sizeof(K2));
  std::vector<K2> cst(r_cycle_);
  cst.resize(2 * ell_n, kzero());
  if (kind == TSK_BITNESS) {
    cst.resize(2 * ell_n + nbits_, kzero());
    cst.insert(cst.end(), weights.begin(), weights.end());
  } else {
    cst.insert(cst.end(), r_addr_.begin(), r_addr_.end());
  }
  h2d_staged(consts_.p, cst.data(), cst.size() * sizeof(K2));
  switch (kind) {
    case TSK_SHOUT_VALUE: degree_bound_ = 3; break;
    case TSK_INDEX_ADAPTER: degree_bound_ = 2 + nbits_; break;
    case TSK_BITNESS: degree_bound_ = 3; break;
    case TSK_TWIST_READ:
    case TSK_TWIST_WRITE:
    case TSK_VAL_EVAL: degree_bound_ = 3 + nbits_; break;
    case TSK_TOTAL_INC: degree_bound_ = 2 + nbits_; break;
    default: throw Error(1, "not a time-lane oracle kind");
  }
  if (degree_bound_ + 1 > (size_t)TS_MAX_COEF) throw Error(1, "time-lane oracle: too many address bits");
  partials_.ensure((size_t)TMAXBLK * TS_MAX_COEF * sizeof(K2));
  coef_dev_.ensure(TS_MAX_COEF * sizeof(K2));
  // Twist Val_pre increment terms at r_addr (twist_oracle.rs:344-371), sorted by time
  init_at_ = init_at_r_addr;
  nterms_ = 0;
  if (kind == TSK_TWIST_READ || kind == TSK_TWIST_WRITE) {
    const SparseCol& hw = *term_write;
    const SparseCol& ic = *term_inc;
    const std::vector<SparseCol>& wb = *term_bits;
    if (wb.size() != r_addr_.size()) throw Error(1, "twist: wa_bits length != r_addr length");
    ColCursor icc(ic);
    std::vector<ColCursor> wbc;
    for (auto& c : wb) wbc.emplace_back(c);
    term_t_h_.clear();
    term_v_h_.clear();
    for (size_t i = 0; i < hw.idx.size(); i++) {
      uint32_t t = hw.idx[i];
      K2 h = hw.val[i], it = icc.at(t);
      if (kiszero(h) || kiszero(it)) continue;
      K2 eq = kone();
      for (size_t bb = 0; bb < wb.size(); bb++) {
        K2 bit = wbc[bb].at(t), uu = r_addr_[bb];
        eq = kmul(eq, kadd(kmul(bit, ksub(kadd(uu, uu), kone())), ksub(kone(), uu)));
      }
      K2 v = kmul(kmul(h, it), eq);
      if (!kiszero(v)) {
        term_t_h_.push_back(t);
        term_v_h_.push_back(v);
      }
    }
    nterms_ = term_t_h_.size();
    std::vector<K2> pre(nterms_ + 1, kzero());
    for (size_t k = 0; k < nterms_; k++) pre[k + 1] = kadd(pre[k], term_v_h_[k]);
    upload(term_t_, term_t_h_);
    upload(term_v_, term_v_h_);
    upload(term_p_, pre);
    // rank table: terms before each time 0..len (4 (len + 1) bytes), for domains up to 2^20
    has_rank_ = ell_n <= 20;
    if (has_rank_) {
      std::vector<uint32_t> rank(len_ + 1);
      size_t k = 0;
      for (size_t t = 0; t <= len_; t++) {
        while (k < nterms_ && term_t_h_[k] < t) k++;
        rank[t] = (uint32_t)k;
      }
      upload(term_rank_, rank);
    }
  }
  NS_CHECK(hipStreamSynchronize(st_));
  stage_release();
}

size_t TsOracleGpu::num_rounds() const {
  switch (kind_) {
    case TSK_PRODUCT:
    case TSK_ADDR_LOOKUP: {
      size_t l = 0;
      while (((size_t)1 << l) < len_) l++;
      return l;
    }
    case TSK_TOTAL_INC: {
      size_t l = 0;
      while (((size_t)1 << l) < len_) l++;
      return l;
    }
    case TSK_READ_ADDR:
    case TSK_WRITE_ADDR:
      return ell_addr_ > bit_idx_ ? ell_addr_ - bit_idx_ : 0;
    default:
      return ell_n_ > bit_idx_ ? ell_n_ - bit_idx_ : 0;
  }
}

// enqueue one round of a time-lane oracle (the challenges so far are already in the constants
// block: launch_fold stores each one); false when the support is empty and the round is zero
// The round's arguments and grid (the Twist val_pre prefix sums are enqueued here); false when
// the support is empty (a zero round).  Sets round_points_: by evaluation (k_ts_time_points,
// blocks of 64) above degree 7, else as coefficients (k_ts_time_round<8>, blocks of TB).
bool TsOracleGpu::time_round_prepare(TsTimeArgs& A, unsigned& nb, TsLtArgs* lt, bool* lt_described) {
  if (lt_described) *lt_described = false;
  const int ncoef = (int)degree_bound_ + 1;
  if (n_ == 0) return false;
  partials_.ensure((size_t)TMAXBLK * TS_MAX_COEF * sizeof(K2));
  A.kind = kind_;
  A.n = (uint32_t)n_;
  A.stride = (uint32_t)cap_;
  A.idx = idx_[cur_].as<uint32_t>();
  A.cols = cols_[cur_].as<K2>();
  A.col_a = col_a_;
  A.col_b = col_b_;
  A.bit0 = (int)(ncol_ - nbits_);
  A.nbits = (int)nbits_;
  A.consts = consts_.as<K2>();
  A.ell = (int)ell_n_;
  A.bit_idx = (int)bit_idx_;
  A.prefix_eq = prefix_eq_;
  A.init_at = init_at_;
  A.term_t = term_t_.as<uint32_t>();
  A.term_v = term_v_.as<K2>();
  A.term_p = term_p_.as<K2>();
  A.term_rank = has_rank_ ? term_rank_.as<uint32_t>() : nullptr;
  A.nterms = (uint32_t)nterms_;
  A.ncoef = ncoef;
  A.n_dev = n_on_device_ ? nslot_.as<uint32_t>() : nullptr;
  A.lt_s0 = A.lt_s1 = nullptr;
  if ((kind_ == TSK_TWIST_READ || kind_ == TSK_TWIST_WRITE) && nterms_ > 0) {
    const uint32_t nt = (uint32_t)nterms_;
    for (int c = 0; c < 2; c++) {
      lt_c_[c].ensure(nterms_ * sizeof(K2));
      lt_s_[c].ensure(nterms_ * sizeof(K2));
    }
    if (lt && nt <= (uint32_t)ONEBLK_N) {
      // described for the group's LT launches, which run before its round kernels
      *lt = TsLtArgs{term_t_.as<uint32_t>(), term_v_.as<K2>(), nt, (int)bit_idx_, consts_.as<K2>() + ell_n_,
                     lt_c_[0].as<K2>(), lt_c_[1].as<K2>(), lt_s_[0].as<K2>(), lt_s_[1].as<K2>()};
      *lt_described = true;
    } else {
      hipLaunchKernelGGL(k_lt_contrib, dim3((nt + 255) / 256), dim3(256), 0, st_, term_t_.as<uint32_t>(),
                         term_v_.as<K2>(), nt, (int)bit_idx_, consts_.as<K2>() + ell_n_, lt_c_[0].as<K2>(),
                         lt_c_[1].as<K2>());
      if (nt <= (uint32_t)ONEBLK_N) {
        hipLaunchKernelGGL(k_scan2_1blk, dim3(1), dim3(ONEBLK_T), 0, st_, lt_c_[0].as<K2>(), lt_c_[1].as<K2>(), nt,
                           lt_s_[0].as<K2>(), lt_s_[1].as<K2>());
      } else {
        size_t tb = 0;
        NS_CHECK(hipcub::DeviceScan::InclusiveScan(nullptr, tb, lt_c_[0].as<K2>(), lt_s_[0].as<K2>(), KAddOp(),
                                                   (int)nt, st_));
        lt_tmp_.ensure(tb + 16);
        for (int c = 0; c < 2; c++)
          NS_CHECK(hipcub::DeviceScan::InclusiveScan(lt_tmp_.p, tb, lt_c_[c].as<K2>(), lt_s_[c].as<K2>(), KAddOp(),
                                                     (int)nt, st_));
      }
    }
    A.lt_s0 = lt_s_[0].as<K2>();
    A.lt_s1 = lt_s_[1].as<K2>();
  }
  A.cgroups = 1;
  if (kind_ == TSK_BITNESS && nbits_ > 4) A.cgroups = (int)std::min<size_t>(8, (nbits_ + 3) / 4);
  round_points_ = ncoef > 8;
  if (!round_points_) {
    nb = nblocks(n_ * (size_t)A.cgroups);
  } else {
    if (kind_ == TSK_BITNESS) throw Error(1, "bitness round: degree above 7");
    nb = (unsigned)std::min<size_t>((n_ + 63) / 64, TMAXBLK);
  }
  return true;
}

bool TsOracleGpu::time_round_launch() {
  TsTimeArgs A;
  unsigned nb = 0;
  if (!time_round_prepare(A, nb, nullptr, nullptr)) return false;
  if (!round_points_)
    hipLaunchKernelGGL(k_ts_time_round<8>, dim3(nb), dim3(TB), 0, st_, A, partials_.as<K2>());
  else
    hipLaunchKernelGGL(k_ts_time_points, dim3(nb, A.ncoef), dim3(64), 0, st_, A, partials_.as<K2>());
  // the sums go straight to the pinned mailbox (read after the stream synchronizes)
  hipLaunchKernelGGL(k_ts_sum, dim3(A.ncoef), dim3(256), 0, st_, partials_.as<K2>(), (int)nb, A.ncoef, coef_h_);
  return true;
}

// enqueue a time-lane fold; the folded support size arrives in the mailbox (finish_fold).
// chal_dst: where the challenge goes in the constants block (null: not stored).  With grp set, a
// fold of at most ONEBLK_N entries is described there for a group launch instead (true).
bool TsOracleGpu::time_fold_launch(K2 r, K2* chal_dst, TsFoldArgs* grp) {
  nslot_.ensure(sizeof(uint32_t));
  if (n_ > 0 && n_ <= (size_t)ONEBLK_N) {
    const uint32_t n = (uint32_t)n_;
    uint32_t* nout = (uint32_t*)(coef_h_ + TS_MAX_COEF);  // pinned mailbox, read by finish_fold
    if (grp) {
      *grp = TsFoldArgs{idx_[cur_].as<uint32_t>(), cols_[cur_].as<K2>(), n, (uint32_t)cap_, (int)ncol_,
                        flags_.as<uint32_t>(), pos_.as<uint32_t>(), idx_[cur_ ^ 1].as<uint32_t>(),
                        cols_[cur_ ^ 1].as<K2>(), nout, nslot_.as<uint32_t>(), chal_dst};
    } else {
      hipLaunchKernelGGL(k_ts_heads_1blk, dim3(1), dim3(ONEBLK_T), 0, st_, idx_[cur_].as<uint32_t>(), n,
                         flags_.as<uint32_t>(), pos_.as<uint32_t>(), chal_dst, r);
      hipLaunchKernelGGL(k_ts_fold, dim3((n + 255) / 256), dim3(256), 0, st_, idx_[cur_].as<uint32_t>(),
                         cols_[cur_].as<K2>(), n, (uint32_t)cap_, (int)ncol_, flags_.as<uint32_t>(),
                         pos_.as<uint32_t>(), r, idx_[cur_ ^ 1].as<uint32_t>(), cols_[cur_ ^ 1].as<K2>(), nout,
                         nslot_.as<uint32_t>());
    }
    fold_pending_ = true;
    cur_ ^= 1;
    len_ >>= 1;
    return grp != nullptr;
  }
  if (chal_dst) hipLaunchKernelGGL(k_put_k2, dim3(1), dim3(64), 0, st_, chal_dst, r);
  if (n_ > 0) {
    const uint32_t n = (uint32_t)n_;
    const uint32_t stride = (uint32_t)cap_;
    unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_ts_heads, dim3(g), dim3(256), 0, st_, idx_[cur_].as<uint32_t>(), n, flags_.as<uint32_t>());
    size_t tmp = 0;
    NS_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flags_.as<uint32_t>(), pos_.as<uint32_t>(), (int)n, st_));
    if (tmp > scan_tmp_.bytes) throw Error(1, "scan workspace too small");
    NS_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp_.p, tmp, flags_.as<uint32_t>(), pos_.as<uint32_t>(), (int)n, st_));
    uint32_t* nout = (uint32_t*)(coef_h_ + TS_MAX_COEF);  // pinned mailbox, read by finish_fold
    hipLaunchKernelGGL(k_ts_fold, dim3(g), dim3(256), 0, st_, idx_[cur_].as<uint32_t>(), cols_[cur_].as<K2>(), n,
                       stride, (int)ncol_, flags_.as<uint32_t>(), pos_.as<uint32_t>(), r, idx_[cur_ ^ 1].as<uint32_t>(),
                       cols_[cur_ ^ 1].as<K2>(), nout, nslot_.as<uint32_t>());
    fold_pending_ = true;
    cur_ ^= 1;
  }
  len_ >>= 1;
  return false;
}

// value of a fully folded time-lane oracle (twist_oracle.rs singleton branches)
K2 TsOracleGpu::time_final() {
  std::vector<K2> v(ncol_, kzero());
  if (n_ > 0) {
    const uint32_t stride = (uint32_t)cap_;
    for (size_t c = 0; c < ncol_; c++)
      NS_CHECK(hipMemcpyAsync(&v[c], cols_[cur_].as<K2>() + c * stride, sizeof(K2), hipMemcpyDeviceToHost, st_));
    NS_CHECK(hipStreamSynchronize(st_));
  }
  const size_t bit0 = ncol_ - nbits_;
  if (kind_ == TSK_BITNESS) {
    // weights live in the constants block after r_cycle | chals | (nbits slots)
    std::vector<K2> w(nbits_);
    if (nbits_)
      NS_CHECK(hipMemcpy(w.data(), consts_.as<K2>() + 2 * ell_n_ + nbits_, nbits_ * sizeof(K2), hipMemcpyDeviceToHost));
    K2 acc = kzero();
    for (size_t c = 0; c < nbits_; c++) {
      K2 b = v[bit0 + c];
      acc = kadd(acc, kmul(kmul(w[c], b), ksub(b, kone())));
    }
    return kmul(prefix_eq_, acc);
  }
  K2 eq = kone();
  for (size_t b = 0; b < nbits_; b++) {
    K2 bit = v[bit0 + b], u = r_addr_[b];
    eq = kmul(eq, kadd(kmul(bit, ksub(kadd(u, u), kone())), ksub(kone(), u)));
  }
  K2 g = v[0];
  auto colv = [&](int c) { return c >= 0 ? v[c] : kzero(); };
  switch (kind_) {
    case TSK_SHOUT_VALUE: return kmul(kmul(prefix_eq_, g), colv(col_a_));
    case TSK_INDEX_ADAPTER: return kmul(kmul(prefix_eq_, g), eq);
    case TSK_TWIST_READ:
    case TSK_TWIST_WRITE: {
      K2 vp = init_at_;
      if (nterms_) {
        unsigned nb = nblocks(nterms_);
        hipLaunchKernelGGL(k_ts_valpre_full, dim3(nb), dim3(TB), 0, st_, term_t_.as<uint32_t>(), term_v_.as<K2>(),
                           (uint32_t)nterms_, consts_.as<K2>() + ell_n_, (int)ell_n_, partials_.as<K2>());
        hipLaunchKernelGGL(k_ts_sum, dim3(1), dim3(256), 0, st_, partials_.as<K2>(), (int)nb, 1, coef_dev_.as<K2>());
        K2 s;
        NS_CHECK(hipMemcpyAsync(&s, coef_dev_.p, sizeof(K2), hipMemcpyDeviceToHost, st_));
        NS_CHECK(hipStreamSynchronize(st_));
        vp = kadd(vp, s);
      }
      if (kind_ == TSK_TWIST_READ) return kmul(kmul(kmul(prefix_eq_, g), ksub(vp, colv(col_a_))), eq);
      return kmul(kmul(kmul(prefix_eq_, g), ksub(ksub(colv(col_a_), vp), colv(col_b_))), eq);
    }
    case TSK_VAL_EVAL: {
      // lt_eval(challenges, r_time), mle.rs:7-24
      K2 suffix = kone(), lt = kzero();
      for (size_t i = ell_n_; i-- > 0;) {
        K2 jp = chals_[i], j = r_cycle_[i];
        lt = kadd(lt, kmul(kmul(ksub(kone(), jp), j), suffix));
        suffix = kmul(suffix, kadd(kmul(ksub(kone(), jp), ksub(kone(), j)), kmul(jp, j)));
      }
      return kmul(kmul(kmul(g, colv(col_b_)), eq), lt);
    }
    case TSK_TOTAL_INC: return kmul(kmul(g, colv(col_b_)), eq);
    default: return kzero();
  }
}

bool TsOracleGpu::product_round_launch() {
  const int ncoef = (int)nf_ + 1;
  size_t half = len_ / 2;
  unsigned nb = nblocks(half);
  partials_.ensure((size_t)TMAXBLK * TS_MAX_COEF * sizeof(K2));
  coef_dev_.ensure(TS_MAX_COEF * sizeof(K2));
  if (ncoef <= 8)
    hipLaunchKernelGGL(k_prod_round<8>, dim3(nb), dim3(TB), 0, st_, fac_[cur_].as<K2>(), (int)nf_, flen_, half,
                       partials_.as<K2>());
  else
    hipLaunchKernelGGL(k_prod_round<TS_MAX_COEF>, dim3(nb), dim3(TB), 0, st_, fac_[cur_].as<K2>(), (int)nf_, flen_, half,
                       partials_.as<K2>());
  // the sums go straight to the pinned mailbox (read after the stream synchronizes)
  hipLaunchKernelGGL(k_ts_sum, dim3(ncoef), dim3(256), 0, st_, partials_.as<K2>(), (int)nb, ncoef, coef_h_);
  return true;
}

void TsOracleGpu::init_lookup(const std::vector<SparseCol>& addr_bits, const SparseCol& has_lookup,
                              const std::vector<K2>& table, const std::vector<K2>& r_cycle, size_t ell_n) {
  const size_t ell_addr = addr_bits.size();
  if (ell_addr > 24) throw Error(1, "address lookup: ell_addr <= 24");
  const size_t pow2_addr = (size_t)1 << ell_addr;
  const size_t len = (size_t)1 << ell_n;
  // lookups: nonzero gates; addresses from the bit columns (a bit counts when it equals ONE)
  std::vector<uint32_t> ts, addr;
  std::vector<K2> gates;
  std::vector<ColCursor> bc;
  for (auto& c : addr_bits) bc.emplace_back(c);
  for (size_t i = 0; i < has_lookup.idx.size(); i++) {
    if (has_lookup.idx[i] >= len) throw Error(1, "sparse column index out of range");
    if (i && has_lookup.idx[i] <= has_lookup.idx[i - 1])
      throw Error(1, "sparse column indices must be strictly increasing");
    if (kiszero(has_lookup.val[i])) continue;
    uint32_t t = has_lookup.idx[i], a = 0;
    for (size_t b = 0; b < ell_addr; b++)
      if (keq(bc[b].at(t), kone())) a |= 1u << b;
    ts.push_back(t);
    addr.push_back(a);
    gates.push_back(has_lookup.val[i]);
  }
  const uint32_t n = (uint32_t)ts.size();
  std::vector<K2> tab(pow2_addr, kzero());
  for (size_t a = 0; a < std::min(pow2_addr, table.size()); a++) tab[a] = table[a];
  nf_ = 2;
  flen_ = pow2_addr;
  fac_[0].ensure(2 * pow2_addr * sizeof(K2));
  fac_[1].ensure(2 * pow2_addr * sizeof(K2));
  h2d_staged(fac_[0].p, tab.data(), pow2_addr * sizeof(K2));
  K2* W = fac_[0].as<K2>() + pow2_addr;
  if (n == 0) {
    NS_CHECK(hipMemsetAsync(W, 0, pow2_addr * sizeof(K2), st_));
  } else {
    TsBuf &dts = scratch_[0], &dg = scratch_[1], &dw = scratch_[2], &da = scratch_[3], &dv = scratch_[4],
          &sa = scratch_[5], &sv = scratch_[6], &rc = scratch_[7], &tmp = scratch_[8];
    upload(dts, ts);
    upload(dg, gates);
    upload(da, addr);
    std::vector<K2> rcy(r_cycle);
    rcy.resize(std::max<size_t>(ell_n, 1), kzero());
    upload(rc, rcy);
    dw.ensure(n * sizeof(K2));
    hipLaunchKernelGGL(k_ts_chi_at, dim3((n + 255) / 256), dim3(256), 0, st_, dts.as<uint32_t>(), n, rc.as<K2>(),
                       (int)ell_n, dg.as<K2>(), dw.as<K2>());
    std::vector<uint32_t> iota(n);
    for (uint32_t i = 0; i < n; i++) iota[i] = i;
    upload(dv, iota);
    sa.ensure(n * 4);
    sv.ensure(n * 4);
    size_t tb = 0;
    NS_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, da.as<uint32_t>(), sa.as<uint32_t>(), dv.as<uint32_t>(),
                                                sv.as<uint32_t>(), (int)n, 0, (int)ell_addr + 1, st_));
    tmp.ensure(tb + 16);
    NS_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, da.as<uint32_t>(), sa.as<uint32_t>(), dv.as<uint32_t>(),
                                                sv.as<uint32_t>(), (int)n, 0, (int)ell_addr + 1, st_));
    hipLaunchKernelGGL(k_lookup_weights, dim3((unsigned)((pow2_addr + 255) / 256)), dim3(256), 0, st_,
                       sa.as<uint32_t>(), sv.as<uint32_t>(), n, dw.as<K2>(), (uint32_t)pow2_addr, W);
    NS_CHECK(hipStreamSynchronize(st_));
  }
  NS_CHECK(hipStreamSynchronize(st_));
  stage_release();
  kind_ = TSK_ADDR_LOOKUP;
  degree_bound_ = 2;
  cur_ = 0;
  len_ = pow2_addr;
  ell_n_ = ell_addr;
}

void TsOracleGpu::init_addr(TsKind kind, const std::vector<uint32_t>& init_addrs, const std::vector<K2>& init_vals,
                            const std::vector<K2>& r_cycle, size_t ell_n, const SparseCol* has_read,
                            const SparseCol* rv, const std::vector<SparseCol>* ra_bits, const SparseCol& has_write,
                            const SparseCol& wv_or_inc, const std::vector<SparseCol>& wa_bits, const SparseCol& inc,
                            size_t ell_addr) {
  kind_ = kind;
  if (ell_addr > 24) throw Error(1, "address lane: ell_addr <= 24");
  ell_addr_ = ell_addr;
  ell_n_ = ell_n;
  const size_t len = (size_t)1 << ell_n, pow2_addr = (size_t)1 << ell_addr;
  const bool rd = kind == TSK_READ_ADDR;
  // every column is read at the event times in ascending order
  auto check_col = [len](const SparseCol& c) {
    for (size_t i = 0; i < c.idx.size(); i++) {
      if (c.idx[i] >= len) throw Error(1, "sparse column index out of range");
      if (i && c.idx[i] <= c.idx[i - 1]) throw Error(1, "sparse column indices must be strictly increasing");
    }
  };
  check_col(has_write);
  check_col(inc);
  check_col(wv_or_inc);
  for (auto& c : wa_bits) check_col(c);
  if (rd) {
    check_col(*has_read);
    check_col(*rv);
    for (auto& c : *ra_bits) check_col(c);
  }
  auto cursors = [](const std::vector<SparseCol>& bits) {
    std::vector<ColCursor> v;
    for (auto& c : bits) v.emplace_back(c);
    return v;
  };
  auto addr_at = [](std::vector<ColCursor>& bits, uint32_t t) {
    uint32_t a = 0;
    for (size_t b = 0; b < bits.size(); b++)
      if (keq(bits[b].at(t), kone())) a |= 1u << b;
    return a;
  };
  auto nz_times = [](const SparseCol& c) {
    std::vector<uint32_t> v;
    for (size_t i = 0; i < c.idx.size(); i++)
      if (!kiszero(c.val[i])) v.push_back(c.idx[i]);
    return v;
  };
  std::vector<uint32_t> times = nz_times(has_write);
  if (rd) {
    std::vector<uint32_t> r = nz_times(*has_read), m;
    std::merge(r.begin(), r.end(), times.begin(), times.end(), std::back_inserter(m));
    m.erase(std::unique(m.begin(), m.end()), m.end());
    times.swap(m);
  }
  nev_ = times.size();
  std::vector<K2> hr(nev_), rvv(nev_), hw(nev_), ic(nev_), one(nev_, kone());
  std::vector<uint32_t> ra(nev_, 0), wa(nev_, 0);
  ColCursor hwc(has_write), icc(inc), wvc(wv_or_inc);
  std::vector<ColCursor> wac = cursors(wa_bits), rac;
  std::unique_ptr<ColCursor> hrc, rvc;
  if (rd) {
    rac = cursors(*ra_bits);
    hrc = std::make_unique<ColCursor>(*has_read);
    rvc = std::make_unique<ColCursor>(*rv);
  }
  for (size_t e = 0; e < nev_; e++) {
    uint32_t t = times[e];
    hw[e] = hwc.at(t);
    ic[e] = icc.at(t);
    if (rd) {
      hr[e] = hrc->at(t);
      rvv[e] = rvc->at(t);
      ra[e] = kiszero(hr[e]) ? 0 : addr_at(rac, t);
      wa[e] = kiszero(hw[e]) ? 0 : addr_at(wac, t);
    } else {
      hr[e] = kzero();
      rvv[e] = wvc.at(t);  // wv
      wa[e] = addr_at(wac, t);
    }
  }
  const size_t ne = std::max<size_t>(nev_, 1);
  upload(ev_hr_, hr, 1);
  upload(ev_rv_, rvv, 1);
  upload(ev_hw_, hw, 1);
  upload(ev_inc_, ic, 1);
  upload(ev_ra_, ra, 1);
  upload(ev_wa_, wa, 1);
  upload(ev_rw_, one, 1);
  upload(ev_ww_, one, 1);
  ev_eq_.ensure(ne * sizeof(K2));
  TsBuf &dts = scratch_[0], &rc = scratch_[1];
  upload(dts, times, 1);
  std::vector<K2> rcy(r_cycle);
  rcy.resize(std::max<size_t>(ell_n, 1), kzero());
  upload(rc, rcy);
  if (nev_)
    hipLaunchKernelGGL(k_ts_chi_at, dim3((unsigned)((nev_ + 255) / 256)), dim3(256), 0, st_, dts.as<uint32_t>(),
                       (uint32_t)nev_, rc.as<K2>(), (int)ell_n, (const K2*)nullptr, ev_eq_.as<K2>());
  // dense init table over addresses
  std::vector<K2> it(pow2_addr, kzero());
  for (size_t i = 0; i < init_addrs.size(); i++) {
    if (init_addrs[i] >= pow2_addr) throw Error(1, "init address out of range");
    it[init_addrs[i]] = kadd(it[init_addrs[i]], init_vals[i]);
  }
  upload(it_[0], it);
  it_[1].ensure(pow2_addr * sizeof(K2));
  ninit_tab_ = pow2_addr;
  for (int i = 0; i < 2; i++) {
    keys_[i].ensure(ne * 4);
    vals_[i].ensure(ne * 4);
  }
  dsort_.ensure(ne * sizeof(K2) * 2);
  dscan_.ensure(ne * sizeof(K2));
  size_t tb = 0, tb2 = 0;
  NS_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys_[0].as<uint32_t>(), keys_[1].as<uint32_t>(),
                                              vals_[0].as<uint32_t>(), vals_[1].as<uint32_t>(), (int)ne, 0,
                                              (int)ell_addr + 1, st_));
  NS_CHECK(hipcub::DeviceScan::InclusiveScan(nullptr, tb2, dsort_.as<K2>(), dscan_.as<K2>(), KAddOp(), (int)ne, st_));
  sort_tmp_bytes_ = std::max(tb, tb2);
  sort_tmp_.ensure(sort_tmp_bytes_ + 16);
  partials_.ensure((size_t)TMAXBLK * TS_MAX_COEF * sizeof(K2));
  coef_dev_.ensure(TS_MAX_COEF * sizeof(K2));
  cur_ = 0;
  bit_idx_ = 0;
  degree_bound_ = 2;
  NS_CHECK(hipStreamSynchronize(st_));
  stage_release();
}

bool TsOracleGpu::addr_round_launch(bool final) {
  if (nev_ == 0) return false;
  const uint32_t n = (uint32_t)nev_;
  const int bit = final ? -1 : (int)bit_idx_;
  const uint32_t sentinel = final ? 1u : (uint32_t)((ninit_tab_ >> bit_idx_));
  unsigned g = (n + 255) / 256;
  if (n <= (uint32_t)ONEBLK_N && ell_addr_ + 1 + AP_EBITS <= 32) {
    hipLaunchKernelGGL(k_addr_prep_1blk, dim3(1), dim3(ONEBLK_T), 0, st_, ev_hw_.as<K2>(), ev_inc_.as<K2>(),
                       ev_ww_.as<K2>(), ev_wa_.as<uint32_t>(), n, bit, sentinel, (int)ell_addr_ + 1,
                       keys_[1].as<uint32_t>(), vals_[1].as<uint32_t>(), dscan_.as<K2>());
  } else {
    K2* dval = dsort_.as<K2>() + n;
    hipLaunchKernelGGL(k_addr_keys, dim3(g), dim3(256), 0, st_, ev_hw_.as<K2>(), ev_inc_.as<K2>(), ev_ww_.as<K2>(),
                       ev_wa_.as<uint32_t>(), n, bit, sentinel, keys_[0].as<uint32_t>(), vals_[0].as<uint32_t>(), dval);
    size_t tb = sort_tmp_bytes_;
    NS_CHECK(hipcub::DeviceRadixSort::SortPairs(sort_tmp_.p, tb, keys_[0].as<uint32_t>(), keys_[1].as<uint32_t>(),
                                                vals_[0].as<uint32_t>(), vals_[1].as<uint32_t>(), (int)n, 0,
                                                (int)ell_addr_ + 1, st_));
    hipLaunchKernelGGL(k_addr_gather, dim3(g), dim3(256), 0, st_, vals_[1].as<uint32_t>(), dval, n, dsort_.as<K2>());
    tb = sort_tmp_bytes_;
    NS_CHECK(hipcub::DeviceScan::InclusiveScan(sort_tmp_.p, tb, dsort_.as<K2>(), dscan_.as<K2>(), KAddOp(), (int)n, st_));
  }
  AddrArgs A;
  A.nev = n;
  A.is_read = kind_ == TSK_READ_ADDR;
  A.bit = (int)bit_idx_;
  A.final_ = final ? 1 : 0;
  A.eq = ev_eq_.as<K2>();
  A.hr = ev_hr_.as<K2>();
  A.rv = ev_rv_.as<K2>();
  A.hw = ev_hw_.as<K2>();
  A.inc = ev_inc_.as<K2>();
  A.rw = ev_rw_.as<K2>();
  A.ww = ev_ww_.as<K2>();
  A.ra = ev_ra_.as<uint32_t>();
  A.wa = ev_wa_.as<uint32_t>();
  A.skeys = keys_[1].as<uint32_t>();
  A.svals = vals_[1].as<uint32_t>();
  A.scan = dscan_.as<K2>();
  A.it = it_[cur_].as<K2>();
  unsigned nb = nblocks(n);
  hipLaunchKernelGGL(k_addr_round, dim3(nb), dim3(TB), 0, st_, A, partials_.as<K2>());
  hipLaunchKernelGGL(k_ts_sum, dim3(3), dim3(256), 0, st_, partials_.as<K2>(), (int)nb, 3, coef_h_);
  return true;
}

void TsOracleGpu::addr_round(std::vector<K2>& c, bool final) {
  finish_fold();
  round_points_ = false;
  round_ncoef_ = 3;
  round_dev_ = addr_round_launch(final);
  finish_round(c);
}

bool TsOracleGpu::time_lane() const {
  return kind_ != TSK_PRODUCT && kind_ != TSK_ADDR_LOOKUP && kind_ != TSK_READ_ADDR && kind_ != TSK_WRITE_ADDR;
}

// What every round does before its kernels.  After a batched fold the round is enqueued without
// waiting for it: its kernels read the folded support size on the device (the host's n_ is an
// upper bound until finish_round).  A round the claim already computed is reused.
void TsOracleGpu::pre_round() {
  if (fold_pending_ && defer_fold_) {
    fold_pending_ = false;
    n_on_device_ = true;
  }
  defer_fold_ = false;
  finish_fold();
  round_points_ = false;
  if (claim_round_valid_) round_from_claim_ = true;
}

void TsOracleGpu::launch_round() {
  pre_round();
  if (round_from_claim_) return;
  switch (kind_) {
    case TSK_PRODUCT:
    case TSK_ADDR_LOOKUP:
      round_ncoef_ = (int)nf_ + 1;
      round_dev_ = product_round_launch();
      return;
    case TSK_READ_ADDR:
    case TSK_WRITE_ADDR:
      round_ncoef_ = 3;
      round_dev_ = addr_round_launch(false);
      return;
    default:
      round_ncoef_ = (int)degree_bound_ + 1;
      round_dev_ = time_round_launch();
  }
}

// Coefficients of the polynomial of degree < n through (x, y[x]), x = 0 .. n-1: Newton forward
// differences over the falling factorials x (x-1) ... (x-k+1), expanded to monomials.  Exact in
// the field, so it returns the coefficient vector itself for any polynomial of degree < n.
static void coeffs_from_points(const K2* y, int n, K2* c) {
  static const std::vector<uint64_t> inv_fact = [] {
    std::vector<uint64_t> v(TS_MAX_COEF + 1);
    uint64_t f = 1;
    for (int k = 0; k <= TS_MAX_COEF; k++) {
      if (k) f = fmul(f, (uint64_t)k);
      v[k] = finv(f);
    }
    return v;
  }();
  std::vector<K2> d(y, y + n), B(n, kzero());
  for (int k = 1; k < n; k++)
    for (int j = n - 1; j >= k; j--) d[j] = ksub(d[j], d[j - 1]);
  B[0] = kone();
  for (int j = 0; j < n; j++) c[j] = kzero();
  for (int k = 0; k < n; k++) {
    if (k > 0) {
      const uint64_t m = (uint64_t)(k - 1);  // B <- B * (x - m)
      for (int j = k; j >= 1; j--) B[j] = ksub(B[j - 1], K2{fmul(B[j].c0, m), fmul(B[j].c1, m)});
      B[0] = ksub(kzero(), K2{fmul(B[0].c0, m), fmul(B[0].c1, m)});
    }
    const K2 w{fmul(d[k].c0, inv_fact[k]), fmul(d[k].c1, inv_fact[k])};
    for (int j = 0; j <= k; j++) c[j] = kadd(c[j], kmul(w, B[j]));
  }
}

void TsOracleGpu::finish_round(std::vector<K2>& c) {
  if (round_from_claim_) {
    c = claim_round_;
    round_from_claim_ = false;
    return;
  }
  c.assign(round_ncoef_, kzero());
  if (round_dev_ || n_on_device_) {
    NS_CHECK(hipStreamSynchronize(st_));
    if (n_on_device_) {  // the fold before this round is complete
      n_ = *(const uint32_t*)(coef_h_ + TS_MAX_COEF);
      n_on_device_ = false;
    }
  }
  if (round_dev_) {
    if (round_points_)
      coeffs_from_points(coef_h_, round_ncoef_, c.data());
    else
      std::copy(coef_h_, coef_h_ + round_ncoef_, c.begin());
  }
  round_dev_ = false;
  round_points_ = false;
}

void TsOracleGpu::round_coeffs(std::vector<K2>& c) {
  launch_round();
  finish_round(c);
}

void TsOracleGpu::finish_fold() {
  if (!fold_pending_) return;
  NS_CHECK(hipStreamSynchronize(st_));
  n_ = *(const uint32_t*)(coef_h_ + TS_MAX_COEF);
  fold_pending_ = false;
}

void TsOracleGpu::evals(const K2* xs, size_t n, K2* out) {
  finish_fold();
  if (num_rounds() == 0) {
    K2 v = kzero();
    if (kind_ == TSK_PRODUCT || kind_ == TSK_ADDR_LOOKUP) {
      std::vector<K2> f(nf_);
      for (size_t i = 0; i < nf_; i++)
        NS_CHECK(hipMemcpyAsync(&f[i], fac_[cur_].as<K2>() + i * flen_, sizeof(K2), hipMemcpyDeviceToHost, st_));
      NS_CHECK(hipStreamSynchronize(st_));
      v = kone();
      for (auto& x : f) v = kmul(v, x);
    } else if (kind_ == TSK_READ_ADDR || kind_ == TSK_WRITE_ADDR) {
      std::vector<K2> c;
      addr_round(c, true);
      v = c[0];
    } else {
      v = time_final();
    }
    for (size_t i = 0; i < n; i++) out[i] = v;
    return;
  }
  std::vector<K2> c;
  round_coeffs(c);
  for (size_t i = 0; i < n; i++) {
    K2 acc = kzero();
    for (size_t d = c.size(); d-- > 0;) acc = kadd(kmul(acc, xs[i]), c[d]);
    out[i] = acc;
  }
}

K2 TsOracleGpu::claim() {
  std::vector<K2> c;
  round_coeffs(c);
  claim_round_ = c;
  claim_round_valid_ = true;
  // p(0) + p(1) = 2 c0 + c1 + ... + c_d
  K2 s = c.empty() ? kzero() : c[0];
  for (auto& x : c) s = kadd(s, x);
  return s;
}

void TsOracleGpu::fold(K2 r) {
  launch_fold(r);
  finish_fold();
}

// every fold is stream-ordered before the next round's kernels; only a time-lane fold leaves
// host state (the support size) to collect
void TsOracleGpu::launch_fold(K2 r, bool defer) { (void)fold_begin(r, defer, nullptr); }

// launch_fold; with grp set, a small time-lane fold is described there for a group launch (true)
bool TsOracleGpu::fold_begin(K2 r, bool defer, TsFoldArgs* grp) {
  finish_fold();
  defer_fold_ = defer;
  claim_round_valid_ = false;
  if (num_rounds() == 0) return false;
  switch (kind_) {
    case TSK_PRODUCT:
    case TSK_ADDR_LOOKUP: {
      size_t half = len_ / 2;
      hipLaunchKernelGGL(k_prod_fold, dim3((unsigned)((half + 255) / 256), (unsigned)nf_), dim3(256), 0, st_,
                         fac_[cur_].as<K2>(), fac_[cur_ ^ 1].as<K2>(), (int)nf_, flen_, half, r);
      cur_ ^= 1;
      len_ = half;
      chals_.push_back(r);
      return false;
    }
    case TSK_READ_ADDR:
    case TSK_WRITE_ADDR: {
      const uint32_t n = (uint32_t)nev_;
      if (n) {
        unsigned g = (n + 255) / 256;
        if (kind_ == TSK_READ_ADDR)
          hipLaunchKernelGGL(k_addr_fold_w, dim3(g), dim3(256), 0, st_, ev_rw_.as<K2>(), ev_ra_.as<uint32_t>(), n,
                             (int)bit_idx_, r);
        hipLaunchKernelGGL(k_addr_fold_w, dim3(g), dim3(256), 0, st_, ev_ww_.as<K2>(), ev_wa_.as<uint32_t>(), n,
                           (int)bit_idx_, r);
      }
      size_t half = (ninit_tab_ >> bit_idx_) / 2;
      hipLaunchKernelGGL(k_prod_fold, dim3((unsigned)((half + 255) / 256), 1), dim3(256), 0, st_, it_[cur_].as<K2>(),
                         it_[cur_ ^ 1].as<K2>(), 1, (size_t)0, half, r);
      cur_ ^= 1;
      bit_idx_++;
      return false;
    }
    case TSK_TOTAL_INC:
      return time_fold_launch(r, nullptr, grp);
    default:
      break;
  }
  if (kind_ != TSK_VAL_EVAL) {
    K2 rc = r_cycle_[bit_idx_];
    prefix_eq_ = kmul(prefix_eq_, kadd(kmul(ksub(kone(), r), ksub(kone(), rc)), kmul(r, rc)));
  }
  // constants block: r_cycle (ell_n) | challenges (ell_n slots); at most ell_n folds
  const bool described = time_fold_launch(r, consts_.as<K2>() + ell_n_ + chals_.size(), grp);
  chals_.push_back(r);
  bit_idx_++;
  return described;
}

static void batch_check(TsOracleGpu* const* os, size_t n) {
  if (n == 0) throw Error(1, "batched round: no oracles");
  for (size_t i = 0; i < n; i++)
    if (!os[i]) throw Error(1, "batched round: null oracle");
  const size_t rounds = os[0]->num_rounds();
  for (size_t i = 0; i < n; i++)
    if (os[i]->num_rounds() != rounds) throw Error(1, "batched round: every oracle must have the same number of rounds");
  if (rounds == 0) throw Error(1, "batched round: no rounds left");
}

// One group launch of up to TS_GROUP prepared time-lane rounds of one family, then their sums.
static void launch_round_group(hipStream_t st, const std::vector<TsTimeArgs>& args, const std::vector<unsigned>& nbs,
                               const std::vector<K2*>& partials, const std::vector<K2*>& outs, size_t first,
                               size_t count, bool points) {
  TsRoundGroup G{};
  TsSumGroup S{};
  G.count = S.count = (int)count;
  uint32_t b = 0;
  int maxc = 1;
  for (size_t j = 0; j < count; j++) {
    const TsTimeArgs& A = args[first + j];
    G.a[j] = A;
    G.partials[j] = partials[first + j];
    G.blk0[j] = b;
    b += nbs[first + j];
    S.partials[j] = partials[first + j];
    S.out[j] = outs[first + j];
    S.nblk[j] = (int)nbs[first + j];
    S.ncoef[j] = A.ncoef;
    maxc = std::max(maxc, A.ncoef);
  }
  G.blk0[count] = b;
  (void)hipGetLastError();  // launch errors below are this group's
  if (points)
    hipLaunchKernelGGL(k_ts_time_points_group, dim3(b, (unsigned)maxc), dim3(64), 0, st, G);
  else
    hipLaunchKernelGGL(k_ts_time_round_group<8>, dim3(b), dim3(TB), 0, st, G);
  NS_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_ts_sum_group, dim3((unsigned)maxc, (unsigned)count), dim3(256), 0, st, S);
  NS_CHECK(hipGetLastError());
}

void ts_batch_launch_round(TsOracleGpu* const* os, size_t n) {
  if (n == 0) throw Error(1, "batched round: no oracles");
  for (size_t i = 0; i < n; i++)
    if (!os[i] || os[i]->num_rounds() == 0) throw Error(1, "batched round: null oracle or no rounds left");
  // the time lanes share a stream and their launches; the other oracles keep their own streams
  std::vector<TsOracleGpu*> tl;
  for (size_t i = 0; i < n; i++)
    if (os[i]->time_lane()) tl.push_back(os[i]);
  ts_join(tl.data(), tl.size());
  const hipStream_t st = tl.empty() ? nullptr : tl[0]->st_;
  // the time lanes' rounds by family (coefficients, points); every other oracle launches its own
  std::vector<TsTimeArgs> args[2];
  std::vector<unsigned> nbs[2];
  std::vector<K2*> partials[2], outs[2];
  std::vector<TsLtArgs> lts;
  for (size_t i = 0; i < n; i++) {
    TsOracleGpu* o = os[i];
    if (!o->time_lane()) {
      o->launch_round();
      continue;
    }
    o->pre_round();
    if (o->round_from_claim_) continue;
    o->round_ncoef_ = (int)o->degree_bound_ + 1;
    TsTimeArgs A;
    unsigned nb = 0;
    TsLtArgs lt;
    bool lt_used = false;
    o->round_dev_ = o->time_round_prepare(A, nb, &lt, &lt_used);
    if (lt_used) lts.push_back(lt);
    if (!o->round_dev_) continue;
    const int f = o->round_points_ ? 1 : 0;
    args[f].push_back(A);
    nbs[f].push_back(nb);
    partials[f].push_back(o->partials_.as<K2>());
    outs[f].push_back(o->coef_h_);
  }
  // the Twist val_pre prefix sums first (the round kernels read them)
  for (size_t j = 0; j < lts.size(); j += TS_GROUP) {
    const size_t count = std::min<size_t>(TS_GROUP, lts.size() - j);
    TsLtGroup G{};
    G.count = (int)count;
    uint32_t b = 0;
    for (size_t q = 0; q < count; q++) {
      G.m[q] = lts[j + q];
      G.blk0[q] = b;
      b += (lts[j + q].nterms + 255) / 256;
    }
    G.blk0[count] = b;
    (void)hipGetLastError();  // launch errors below are this group's
    hipLaunchKernelGGL(k_lt_contrib_group, dim3(b), dim3(256), 0, st, G);
    NS_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_scan2_1blk_group, dim3((unsigned)count), dim3(ONEBLK_T), 0, st, G);
    NS_CHECK(hipGetLastError());
  }
  for (int f = 0; f < 2; f++)
    for (size_t j = 0; j < args[f].size(); j += TS_GROUP)
      launch_round_group(st, args[f], nbs[f], partials[f], outs[f], j,
                         std::min<size_t>(TS_GROUP, args[f].size() - j), f == 1);
}

void ts_batch_round(TsOracleGpu* const* os, size_t n, std::vector<std::vector<K2>>& coeffs) {
  batch_check(os, n);
  ts_batch_launch_round(os, n);
  coeffs.resize(n);
  for (size_t i = 0; i < n; i++) os[i]->finish_round(coeffs[i]);
}

void ts_batch_fold(TsOracleGpu* const* os, size_t n, K2 r) {
  for (size_t i = 0; i < n; i++)
    if (!os[i]) throw Error(1, "batched fold: null oracle");
  if (n == 0) return;
  std::vector<TsOracleGpu*> tl;
  for (size_t i = 0; i < n; i++)
    if (os[i]->time_lane()) tl.push_back(os[i]);
  ts_join(tl.data(), tl.size());
  const hipStream_t st = tl.empty() ? nullptr : tl[0]->st_;
  // the folds' support sizes are collected by the next batched round (or by any other use); the
  // small time-lane folds of the batch share one heads launch and one compaction launch per group
  std::vector<TsFoldArgs> fa;
  for (size_t i = 0; i < n; i++) {
    TsFoldArgs F;
    if (os[i]->fold_begin(r, true, &F)) fa.push_back(F);
  }
  for (size_t j = 0; j < fa.size(); j += TS_FOLD_GROUP) {
    const size_t count = std::min<size_t>(TS_FOLD_GROUP, fa.size() - j);
    TsFoldGroup G{};
    G.count = (int)count;
    G.r = r;
    uint32_t b = 0;
    for (size_t q = 0; q < count; q++) {
      G.f[q] = fa[j + q];
      G.blk0[q] = b;
      b += (fa[j + q].n + 255) / 256;
    }
    G.blk0[count] = b;
    (void)hipGetLastError();  // launch errors below are this group's
    hipLaunchKernelGGL(k_ts_heads_1blk_group, dim3((unsigned)count), dim3(ONEBLK_T), 0, st, G);
    NS_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ts_fold_group, dim3(b), dim3(256), 0, st, G);
    NS_CHECK(hipGetLastError());
  }
}

// ---------------- shared CPU bus (host) ----------------
static void bus_check(size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols, size_t pow2) {
  if (chunk == 0) throw Error(1, "bus: chunk_size must be >= 1");
  if (chunk > 0xFFFFFFFFull || m_in + chunk > pow2) throw Error(1, "bus: time index out of range for ell_n");
  if (bus_base + bus_cols * chunk > m) throw Error(1, "bus: bus region exceeds the witness");
}

void bus_decode(const int8_t* Z_host, size_t m, uint64_t* z_out) {
  hipStream_t st;
  NS_CHECK(hipStreamCreate(&st));
  DevBuf dz(D * m), dout(m * 8);
  NS_CHECK(hipMemcpyAsync(dz.p, Z_host, D * m, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_bus_decode, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, dz.as<int8_t>(), m,
                     dout.as<uint64_t>());
  NS_CHECK(hipMemcpyAsync(z_out, dout.p, m * 8, hipMemcpyDeviceToHost, st));
  NS_CHECK(hipStreamSynchronize(st));
  NS_CHECK(hipStreamDestroy(st));
}

void bus_time_cols(const int8_t* Z_host, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                   size_t steps_len, size_t ell_n, uint64_t* idx_out, uint64_t* vals_out, size_t* nnz_out) {
  if (ell_n > 31) throw Error(1, "bus: ell_n <= 31");
  bus_check(m, m_in, bus_base, chunk, bus_cols, (size_t)1 << ell_n);
  if (steps_len > chunk) throw Error(1, "bus: steps_len > chunk_size");
  if (bus_cols == 0) return;
  hipStream_t st;
  NS_CHECK(hipStreamCreate(&st));
  DevBuf dz(D * m), di(bus_cols * chunk * 4), dv(bus_cols * chunk * sizeof(K2)), dn(bus_cols * 4);
  NS_CHECK(hipMemcpyAsync(dz.p, Z_host, D * m, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_bus_time_cols, dim3((unsigned)bus_cols), dim3(256), 0, st, dz.as<int8_t>(), m, (uint32_t)m_in,
                     bus_base, (uint32_t)chunk, (uint32_t)steps_len, di.as<uint32_t>(), dv.as<K2>(), dn.as<uint32_t>());
  std::vector<uint32_t> hi(bus_cols * chunk), hn(bus_cols);
  NS_CHECK(hipMemcpyAsync(hi.data(), di.p, hi.size() * 4, hipMemcpyDeviceToHost, st));
  NS_CHECK(hipMemcpyAsync(vals_out, dv.p, bus_cols * chunk * sizeof(K2), hipMemcpyDeviceToHost, st));
  NS_CHECK(hipMemcpyAsync(hn.data(), dn.p, bus_cols * 4, hipMemcpyDeviceToHost, st));
  NS_CHECK(hipStreamSynchronize(st));
  NS_CHECK(hipStreamDestroy(st));
  for (size_t c = 0; c < bus_cols; c++) {
    nnz_out[c] = hn[c];
    for (size_t j = 0; j < hn[c]; j++) idx_out[c * chunk + j] = hi[c * chunk + j];
  }
}

void bus_openings(const int8_t* Z_host, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                  const K2* r, size_t ell_n, K2* y_out, K2* ys_out) {
  if (ell_n == 0 || ell_n > 31) throw Error(1, "bus openings require 1 <= ell_n <= 31");
  bus_check(m, m_in, bus_base, chunk, bus_cols, (size_t)1 << ell_n);
  if (bus_cols == 0) return;
  hipStream_t st;
  NS_CHECK(hipStreamCreate(&st));
  DevBuf dz(D * m), dt(chunk * 4), dr(ell_n * sizeof(K2)), dw(chunk * sizeof(K2)), dy(bus_cols * 64 * sizeof(K2)),
      dys(bus_cols * sizeof(K2));
  std::vector<uint32_t> ts(chunk);
  for (size_t j = 0; j < chunk; j++) ts[j] = (uint32_t)(m_in + j);
  NS_CHECK(hipMemcpyAsync(dz.p, Z_host, D * m, hipMemcpyHostToDevice, st));
  NS_CHECK(hipMemcpyAsync(dt.p, ts.data(), chunk * 4, hipMemcpyHostToDevice, st));
  NS_CHECK(hipMemcpyAsync(dr.p, r, ell_n * sizeof(K2), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_ts_chi_at, dim3((unsigned)((chunk + 255) / 256)), dim3(256), 0, st, dt.as<uint32_t>(),
                     (uint32_t)chunk, dr.as<K2>(), (int)ell_n, (const K2*)nullptr, dw.as<K2>());
  hipLaunchKernelGGL(k_bus_open, dim3((unsigned)bus_cols, D), dim3(256), 0, st, dz.as<int8_t>(), m, bus_base,
                     (uint32_t)chunk, dw.as<K2>(), dy.as<K2>());
  hipLaunchKernelGGL(k_bus_scalars, dim3((unsigned)((bus_cols + 63) / 64)), dim3(64), 0, st, dy.as<K2>(),
                     (uint32_t)bus_cols, dys.as<K2>());
  NS_CHECK(hipMemcpyAsync(y_out, dy.p, bus_cols * 64 * sizeof(K2), hipMemcpyDeviceToHost, st));
  NS_CHECK(hipMemcpyAsync(ys_out, dys.p, bus_cols * sizeof(K2), hipMemcpyDeviceToHost, st));
  NS_CHECK(hipStreamSynchronize(st));
  NS_CHECK(hipStreamDestroy(st));
}

}  // namespace ns
