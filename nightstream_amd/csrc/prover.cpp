// prover.cpp -- MI355X shard prover: host fold loop over device-resident state.
//
// Restates, in transcript order, neo-fold/src/shard.rs:1725-2283 (fold_shard_prove_impl) for
// CCS-only steps, with the Pi_CCS round oracle, precompute_for_r, the Ajtai commitment,
// Pi_RLC's Z_mix and Pi_DEC's digit split and child products on the GPU.  Host-side pieces:
// transcript (neo-transcript), sum-check drivers and Lagrange interpolation
// (neo-reductions/src/sumcheck.rs:79-209, 350-456), Ajtai-bit rounds
// (optimized_engine/oracle.rs:1812-1951), ME outputs (oracle.rs:1955-2072), Pi_RLC on X, y and
// commitments (optimized_engine/common.rs:1221-1379, session.rs:301-324) and the DEC checks
// (optimized_engine/common.rs:1575-1628).
#include <array>

#include "prover.h"
#include "route_a.h"

#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>

namespace ns {

// ---------------- device memory ----------------
void DevBuf::alloc(size_t n) {
  release();
  if (n == 0) n = 256;
  NS_CHECK(hipMalloc(&p, n));
  bytes = n;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

template <class T>
static void h2d(void* dst, const T* src, size_t n, hipStream_t s) {
  NS_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
// Device-to-host copies land in a pinned staging buffer (a direct DMA, no runtime staging
// through pageable memory) and are then copied to the caller's memory.  A prover context owns
// its stage (grown on first use, so a warmed context allocates no pinned memory while proving).
void* PinnedStage::get(size_t n) {
  if (n > bytes) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    size_t want = n < ((size_t)1 << 20) ? ((size_t)1 << 20) : n;
    NS_CHECK(hipHostMalloc(&p, want, hipHostMallocDefault));
    bytes = want;
  }
  return p;
}
PinnedStage::~PinnedStage() {
  if (p) (void)hipHostFree(p);
}

template <class T>
static void d2h(PinnedStage& stage, T* dst, const void* src, size_t n, hipStream_t s) {
  const size_t bytes = n * sizeof(T);
  if (bytes == 0) return;
  void* h = stage.get(bytes);
  NS_CHECK(hipMemcpyAsync(h, src, bytes, hipMemcpyDeviceToHost, s));
  NS_CHECK(hipStreamSynchronize(s));
  memcpy(dst, h, bytes);
}

// ---------------- kernel timers ----------------
namespace {
struct KTimers {
  struct Rec {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      NS_CHECK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  void collect(double* ms, uint64_t* cnt) {
    for (int c = 0; c < KT_N; c++) {
      ms[c] = 0;
      cnt[c] = 0;
    }
    for (auto& r : recs) {
      float t = 0;
      NS_CHECK(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.cls] += t;
      cnt[r.cls]++;
    }
  }
  ~KTimers() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
thread_local KTimers* g_kt = nullptr;
}  // namespace

bool ktimer_events(int cls, hipEvent_t* start, hipEvent_t* stop) {
  if (!g_kt) return false;
  *start = g_kt->get();
  *stop = g_kt->get();
  g_kt->recs.push_back({cls, *start, *stop});
  return true;
}

// ---------------- params (neo-params/src/lib.rs) ----------------
typedef unsigned __int128 u128;
static uint32_t bitlen128(u128 x) {
  uint32_t n = 0;
  while (x) {
    n++;
    x >>= 1;
  }
  return n;
}

bool Params::extension_check(uint32_t ell, uint32_t d_sc, uint32_t* s_min_out, int32_t* slack_out) const {
  u128 ld = (u128)ell * d_sc;
  uint32_t smin = 3;
  if (lambda < 128) {
    u128 pow2 = (u128)1 << lambda;
    bool ovf = ld != 0 && pow2 > (~(u128)0) / ld;
    if (!ovf && (u128)q >= pow2 * ld) smin = 1;
    else if (!ovf && (u128)q * q >= pow2 * ld) smin = 2;
  }
  if (smin > 2) return false;
  u128 q2 = (u128)q * q, rhs = ((u128)1 << lambda) * ld;
  int32_t slack;
  if (q2 < rhs) {
    slack = -1;
  } else {
    slack = (int32_t)bitlen128(q2) - (int32_t)bitlen128(rhs);
    if (slack >= 0 && slack < 128 && q2 < (rhs << (uint32_t)slack)) slack -= 1;
  }
  if (s_min_out) *s_min_out = smin;
  if (slack_out) *slack_out = slack;
  return true;
}

Params Params::goldilocks_auto_r1cs_ccs(size_t n_rows) {
  Params p;
  size_t padded = 2;
  if (n_rows) {
    padded = 1;
    while (padded < n_rows) padded <<= 1;
    if (padded < 2) padded = 2;
  }
  u128 prod = (u128)p.d * padded;
  uint32_t ell = prod <= 1 ? 0 : bitlen128(prod - 1);
  uint32_t d_sc = 5;
  for (uint32_t lam = p.lambda; lam >= 96; lam--) {
    p.lambda = lam;
    int32_t slack;
    uint32_t sm;
    if (p.extension_check(ell, d_sc, &sm, &slack) && slack >= 2) return p;
  }
  throw Error(2, "goldilocks_auto_r1cs_ccs: extension degree s >= 3 required");
}

// ---------------- CCS ----------------
size_t Ccs::f_degree() const {
  size_t d = 0;
  for (auto& e : term_exps) {
    size_t s = 0;
    for (auto x : e) s += x;
    d = std::max(d, s);
  }
  return d;
}

static void ccs_digest(const Ccs& s, uint64_t out[4]) {
  // neo-reductions/src/engines/utils.rs:214-321
  const Poseidon2& P = digest_perm();
  uint64_t st[16] = {0};
  int absorbed = 0;
  const char* dom = "neo/ccs/matrices/v1";
  for (const char* c = dom; *c; c++) {
    if (absorbed >= 15) {
      P.permute(st);
      absorbed = 0;
    }
    st[absorbed++] = (uint8_t)*c;
  }
  if (absorbed + 3 >= 16) {
    P.permute(st);
    absorbed = 0;
  }
  st[absorbed] = s.n;
  st[absorbed + 1] = s.m;
  st[absorbed + 2] = s.t;
  P.permute(st);
  auto emit = [&](uint64_t r, uint64_t c, uint64_t v) {
    if (absorbed + 3 > 15) {
      P.permute(st);
      absorbed = 0;
    }
    st[absorbed] = r % Q;
    st[absorbed + 1] = c % Q;
    st[absorbed + 2] = v % Q;
    absorbed += 3;
  };
  for (size_t j = 0; j < s.t; j++) {
    absorbed = 0;
    st[0] = j;
    absorbed = 1;
    if (j == 0) {
      for (size_t r = 0; r < s.n; r++) emit(r, r, 1);
    } else {
      // row-major nonzero order with ascending columns inside a row
      const HostCsr& M = s.host[j];
      std::vector<std::pair<uint32_t, uint64_t>> row;
      for (size_t r = 0; r < s.n; r++) {
        row.clear();
        for (uint64_t k = M.ptr[r]; k < M.ptr[r + 1]; k++)
          if (M.val[k] != 0) row.emplace_back(M.col[k], M.val[k]);
        std::sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (auto& e : row) emit(r, e.first, e.second);
      }
    }
    P.permute(st);
  }
  for (int i = 0; i < 4; i++) out[i] = st[i];
}

std::unique_ptr<Ccs> Ccs::from_r1cs(size_t n, size_t m, const HostCsr& a, const HostCsr& b, const HostCsr& c) {
  // neo-ccs/src/r1cs.rs:13-43 (identity-first, f = X1 X2 - X3 with a leading dummy variable)
  if (n != m) throw Error(1, "Pi_CCS requires a square CCS (n == m) with M_0 = I");
  if (n < 2 || (n & (n - 1)) != 0) throw Error(1, "n must be a power of two >= 2 on this backend");
  auto s = std::make_unique<Ccs>();
  s->n = n;
  s->m = m;
  s->t = 4;
  s->host.resize(4);
  s->host[1] = a;
  s->host[2] = b;
  s->host[3] = c;
  s->term_coeff = {1, Q - 1};
  s->term_exps = {{0, 1, 1, 0}, {0, 0, 0, 1}};
  // compile f over used variables
  std::vector<int> pos(s->t, -1);
  for (auto& e : s->term_exps)
    for (size_t j = 0; j < s->t; j++)
      if (e[j] && pos[j] < 0) pos[j] = 0;
  int nf = 0;
  for (size_t j = 0; j < s->t; j++)
    if (pos[j] == 0) {
      pos[j] = nf++;
      s->fvar_mat.push_back((int)j);
    }
  FPoly& f = s->fpoly;
  f.nterms = (int)s->term_coeff.size();
  for (int tt = 0; tt < f.nterms; tt++) {
    f.coeff[tt] = s->term_coeff[tt];
    f.nv[tt] = 0;
    for (size_t j = 0; j < s->t; j++)
      if (s->term_exps[tt][j]) {
        f.var[tt][f.nv[tt]] = pos[j];
        f.exp[tt][f.nv[tt]] = (int)s->term_exps[tt][j];
        f.nv[tt]++;
      }
  }
  f.deg_max = (int)std::max<size_t>(4, s->f_degree() + 1);
  if (f.deg_max != 4) throw Error(1, "f degree > 3 is not supported by the round kernels");
  for (size_t j = 1; j < 4; j++) {
    const HostCsr& H = s->host[j];
    if (H.ptr.size() != n + 1) throw Error(1, "CSR row pointer has wrong length");
    for (size_t k = 0; k < H.ptr[n]; k++)
      if (H.col[k] >= m) throw Error(1, "CSR column index out of range");
  }
  ccs_digest(*s, s->digest);
  return s;
}

// The matrices in HBM (CSR and CSC, heavy-column lists), uploaded on first use by a prover or an
// oracle: a verifier needs only the host side, so constructing a CCS needs no device.
void Ccs::ensure_device() {
  std::lock_guard<std::mutex> g(dev_mu);
  if (on_device) return;
  dev.assign(4, DevSparse{});
  dev[0].identity = true;
  dev[0].n = (uint32_t)n;
  dev[0].m = (uint32_t)m;
  for (size_t j = 1; j < 4; j++) {
    const HostCsr& H = host[j];
    size_t nnz = H.ptr[n];
    std::vector<uint32_t> rp(n + 1), cp(m + 1, 0), rows(nnz);
    std::vector<uint64_t> cv(nnz);
    for (size_t r = 0; r <= n; r++) rp[r] = (uint32_t)H.ptr[r];
    for (size_t k = 0; k < nnz; k++) cp[H.col[k] + 1]++;
    for (size_t cc = 0; cc < m; cc++) cp[cc + 1] += cp[cc];
    std::vector<uint32_t> pos2(cp.begin(), cp.end() - 1);
    for (size_t r = 0; r < n; r++)
      for (uint64_t k = H.ptr[r]; k < H.ptr[r + 1]; k++) {
        uint32_t cc = H.col[k];
        rows[pos2[cc]] = (uint32_t)r;
        cv[pos2[cc]] = H.val[k];
        pos2[cc]++;
      }
    DevSparse& D_ = dev[j];
    D_.n = (uint32_t)n;
    D_.m = (uint32_t)m;
    D_.nnz = (uint32_t)nnz;
    auto up = [&](const void* src, size_t bytes) {
      bufs.emplace_back(bytes);
      NS_CHECK(hipMemcpy(bufs.back().p, src, bytes, hipMemcpyHostToDevice));
      return bufs.back().p;
    };
    D_.csr_ptr = (uint32_t*)up(rp.data(), rp.size() * 4);
    D_.csr_col = (uint32_t*)up(H.col.data(), std::max<size_t>(nnz, 1) * 4);
    D_.csr_val = (uint64_t*)up(H.val.data(), std::max<size_t>(nnz, 1) * 8);
    D_.csc_ptr = (uint32_t*)up(cp.data(), cp.size() * 4);
    D_.csc_row = (uint32_t*)up(rows.data(), std::max<size_t>(nnz, 1) * 4);
    D_.csc_val = (uint64_t*)up(cv.data(), std::max<size_t>(nnz, 1) * 8);
    std::vector<uint32_t> heavy;
    for (size_t cc = 0; cc < m; cc++)
      if (cp[cc + 1] - cp[cc] > SPMV_HEAVY) heavy.push_back((uint32_t)cc);
    D_.n_heavy = (uint32_t)heavy.size();
    if (!heavy.empty()) D_.heavy_cols = (uint32_t*)up(heavy.data(), heavy.size() * 4);
  }
  on_device = true;
}

// ---------------- PP ----------------
__global__ void k_count_rejects(const uint64_t* x, size_t n, unsigned long long* cnt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && x[i] >= Q) atomicAdd(cnt, 1ull);
}

static std::unique_ptr<Pp> finish_pp(std::unique_ptr<Pp> pp, bool keep, hipStream_t s) {
  pp->limbs.alloc((size_t)KAPPA * 8 * D * pp->m);
  launch_pp_limbs(pp->coeffs.as<uint64_t>(), KAPPA, pp->m, pp->limbs.as<int8_t>(), s);
  NS_CHECK(hipStreamSynchronize(s));
  if (!keep) pp->coeffs.release();
  return pp;
}

std::unique_ptr<Pp> Pp::setup_seeded(const uint8_t seed[32], size_t m, bool keep_coeffs, hipStream_t s) {
  // neo-ajtai/src/commit.rs:133-156: one ChaCha8 stream, rows i, columns j, 54 coefficients,
  // each a next_u64 draw accepted when < q.  Draw k without rejections is keystream words
  // (2k, 2k+1); the rare rejected draws are compacted on the host.
  if (m == 0 || m % 64 != 0) throw Error(1, "Ajtai m must be a nonzero multiple of 64");
  auto pp = std::make_unique<Pp>();
  pp->m = m;
  const size_t N = (size_t)KAPPA * m * D, extra = 1024;
  const size_t draws = N + extra, nblocks = (2 * draws + 15) / 16;
  DevBuf words(nblocks * 64);
  ChaCha8 rng(seed);
  launch_chacha8_words(rng.key, 0, nblocks, words.as<uint32_t>(), s);
  DevBuf cnt(8);
  NS_CHECK(hipMemsetAsync(cnt.p, 0, 8, s));
  hipLaunchKernelGGL(k_count_rejects, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, words.as<uint64_t>(), N,
                     cnt.as<unsigned long long>());
  unsigned long long rej = 0;
  PinnedStage stage;
  d2h(stage, &rej, cnt.p, 1, s);
  if (rej == 0) {
    pp->coeffs = std::move(words);
  } else {
    std::vector<uint64_t> x(draws), out;
    out.reserve(N);
    d2h(stage, x.data(), words.p, draws, s);
    size_t used = 0;
    while (out.size() < N) {
      if (used == x.size()) {
        // continue the stream on the host
        ChaCha8 tail(seed);
        for (size_t i = 0; i < used; i++) (void)tail.next_u64();
        while (out.size() < N) {
          uint64_t v = tail.next_u64();
          if (v < Q) out.push_back(v);
        }
        break;
      }
      uint64_t v = x[used++];
      if (v < Q) out.push_back(v);
    }
    pp->coeffs.alloc(N * 8);
    h2d(pp->coeffs.p, out.data(), N, s);
  }
  return finish_pp(std::move(pp), keep_coeffs, s);
}

std::unique_ptr<Pp> Pp::from_coeffs(const uint64_t* host, size_t m, bool keep_coeffs, hipStream_t s) {
  if (m == 0 || m % 64 != 0) throw Error(1, "Ajtai m must be a nonzero multiple of 64");
  auto pp = std::make_unique<Pp>();
  pp->m = m;
  pp->coeffs.alloc((size_t)KAPPA * m * D * 8);
  h2d(pp->coeffs.p, host, (size_t)KAPPA * m * D, s);
  return finish_pp(std::move(pp), keep_coeffs, s);
}

// ---------------- host helpers ----------------
static std::vector<K2> chi_host(const K2* bits, size_t nbits) {
  size_t len = (size_t)1 << nbits;
  std::vector<K2> w(len, kzero());
  w[0] = kone();
  for (size_t i = 0; i < nbits; i++) {
    size_t step = (size_t)1 << i;
    K2 om = ksub(kone(), bits[i]);
    for (size_t mask = 0; mask < step; mask++) {
      K2 v = w[mask];
      w[mask] = kmul(v, om);
      w[mask + step] = kmul(v, bits[i]);
    }
  }
  return w;
}

static K2 eq_mask(size_t mask, const K2* pts, size_t n) {
  K2 p = kone();
  for (size_t b = 0; b < n; b++) p = kmul(p, ((mask >> b) & 1) ? pts[b] : ksub(kone(), pts[b]));
  return p;
}

static K2 eq_lin(K2 a, K2 b) { return kadd(kmul(ksub(kone(), a), ksub(kone(), b)), kmul(a, b)); }

static K2 poly_eval(const K2* c, size_t n, K2 x) {
  K2 r = c[n - 1];
  for (size_t i = n - 1; i-- > 0;) r = kadd(kmul(r, x), c[i]);
  return r;
}

// ---------------- Ajtai-bit rounds ----------------
void AjtaiPhase::init(const Ccs& s, std::vector<K2> y, size_t k, const K2* alpha_, const K2* beta_a_,
                      const K2* beta_r, K2 gamma_, const std::vector<K2>& r_time, const K2* r_inputs) {
  kt = k;
  t = s.t;
  y_eval = std::move(y);
  alpha.assign(alpha_, alpha_ + ell_d);
  beta_a.assign(beta_a_, beta_a_ + ell_d);
  gamma = gamma_;
  chals.clear();
  // F' = f(m_vals), m_vals[j] = sum_rho 2^rho y_eval[0][j][rho] (= <z1, v_j>)
  std::vector<K2> mv(t, kzero());
  K2 pw = kone();
  for (int rho = 0; rho < D; rho++) {
    for (size_t j = 0; j < t; j++) mv[j] = kadd(mv[j], kmul(y_eval[(0 * t + j) * D + rho], pw));
    pw = kadd(pw, pw);
  }
  f_prime = kzero();
  for (size_t tt = 0; tt < s.term_coeff.size(); tt++) {
    K2 term = K2{s.term_coeff[tt], 0};
    for (size_t j = 0; j < t; j++)
      for (uint32_t e = 0; e < s.term_exps[tt][j]; e++) term = kmul(term, mv[j]);
    f_prime = kadd(f_prime, term);
  }
  eq_beta_r = kone();
  for (size_t i = 0; i < r_time.size(); i++) eq_beta_r = kmul(eq_beta_r, eq_lin(r_time[i], beta_r[i]));
  has_inputs = r_inputs != nullptr;
  eq_r_in = kzero();
  if (has_inputs) {
    eq_r_in = kone();
    for (size_t i = 0; i < r_time.size(); i++) eq_r_in = kmul(eq_r_in, eq_lin(r_time[i], r_inputs[i]));
  }
}

// out = sum_i a_i(X) v_i(X) mod Phi_81 (X^54 = -X^27 - 1) for a_i with 54 coefficients in
// [-2, 2] (a[i*54 + u]) and v_i in F_q^54.  The 32-bit halves of v are accumulated in int64
// (|sum| < 2^33 * 16 * 54 < 2^43, no carries, vectorizable adds) and reduced once per
// coefficient of the product.
static void mix_small_coeffs(const int8_t* a, const uint64_t* const* v, size_t kt, uint64_t* out) {
  int64_t lo[2 * D - 1] = {0}, hi[2 * D - 1] = {0};
  for (size_t i = 0; i < kt; i++) {
    uint32_t vl[D], vh[D];
    for (int x = 0; x < D; x++) {
      vl[x] = (uint32_t)v[i][x];
      vh[x] = (uint32_t)(v[i][x] >> 32);
    }
    for (int u = 0; u < D; u++) {
      const int au = a[i * D + u];
      int64_t* L = lo + u;
      int64_t* H = hi + u;
      switch (au) {
        case 1:
          for (int x = 0; x < D; x++) L[x] += vl[x], H[x] += vh[x];
          break;
        case -1:
          for (int x = 0; x < D; x++) L[x] -= vl[x], H[x] -= vh[x];
          break;
        case 2:
          for (int x = 0; x < D; x++) L[x] += 2 * (int64_t)vl[x], H[x] += 2 * (int64_t)vh[x];
          break;
        case -2:
          for (int x = 0; x < D; x++) L[x] -= 2 * (int64_t)vl[x], H[x] -= 2 * (int64_t)vh[x];
          break;
        default:
          break;
      }
    }
  }
  uint64_t P[2 * D - 1];
  for (int q = 0; q < 2 * D - 1; q++) P[q] = fadd(from_i64(lo[q]), fmul(from_i64(hi[q]), 1ull << 32));
  for (int q = 0; q < D; q++) {
    uint64_t c = P[q];
    if (q <= 25) c = fadd(c, P[q + 81]);
    if (q >= 27) c = fsub(c, P[q + 27]);
    if (q <= 26) c = fsub(c, P[q + 54]);
    out[q] = c;
  }
}

void AjtaiPhase::fold(K2 r) {
  // fold digit bit j = chals.size() at the entries that later rounds read (multiples of 2^(j+1))
  const size_t j = chals.size(), stride = (size_t)1 << j, step = stride << 1;
  for (size_t q = 0; q < kt * t; q++) {
    K2* y = &y_eval[q * D];
    for (size_t i0 = 0; i0 < (size_t)D; i0 += step) {
      const K2 lo = y[i0], hi = i0 + stride < (size_t)D ? y[i0 + stride] : kzero();
      y[i0] = kadd(lo, kmul(ksub(hi, lo), r));
    }
  }
  chals.push_back(r);
}

// Round j of the Ajtai-bit sum-check (optimized_engine/oracle.rs:1812-1951) as a polynomial in
// X: with a = y[id], b = y[id + 2^j] - a for id = tt 2^(j+1) (y folded by the earlier
// challenges), N(a + bX) = (a^3 - a) + (3a^2 b - b) X + 3ab^2 X^2 + b^3 X^3, and the eq factors
// and the Eval term are linear in X.  The values at xs are evaluations of that polynomial.
std::vector<K2> AjtaiPhase::evals(const std::vector<K2>& xs) const {
  const std::vector<K2> o = coeffs();
  std::vector<K2> ys(xs.size());
  for (size_t xi = 0; xi < xs.size(); xi++) ys[xi] = poly_eval(o.data(), o.size(), xs[xi]);
  return ys;
}

std::vector<K2> AjtaiPhase::coeffs() const {
  const size_t j = chals.size(), free_a = ell_d - j - 1, tail = (size_t)1 << free_a, hs = (size_t)1 << (j + 1);
  const size_t half = (size_t)1 << j;
  std::vector<K2> wbt = chi_host(beta_a.data() + j + 1, free_a), wat = chi_host(alpha.data() + j + 1, free_a);
  K2 eqb_pref = kone(), eqa_pref = kone();
  for (size_t i = 0; i < j; i++) {
    eqb_pref = kmul(eqb_pref, eq_lin(chals[i], beta_a[i]));
    eqa_pref = kmul(eqa_pref, eq_lin(chals[i], alpha[i]));
  }
  std::vector<K2> gpi(kt);
  gpi[0] = kone();
  for (size_t i = 1; i < kt; i++) gpi[i] = kmul(gpi[i - 1], gamma);
  K2 gk = kone();
  for (size_t i = 0; i < kt; i++) gk = kmul(gk, gamma);
  std::vector<K2> gkj(t);
  gkj[0] = kone();
  for (size_t q = 1; q < t; q++) gkj[q] = kmul(gkj[q - 1], gk);
  auto ab = [&](size_t q, size_t tt, K2& a, K2& b) {
    const size_t id = tt * hs;
    const K2* y = &y_eval[q * D];
    a = y[id];
    b = ksub(id + half < (size_t)D ? y[id + half] : kzero(), a);
  };
  // NC(X) = sum_i gamma^(i+1) sum_tt wbt[tt] N(a + bX), cubic
  K2 nc[4] = {kzero(), kzero(), kzero(), kzero()};
  K2 g = gamma;
  for (size_t i = 0; i < kt; i++) {
    K2 c[4] = {kzero(), kzero(), kzero(), kzero()};
    for (size_t tt = 0; tt < tail && tt * hs < (size_t)D; tt++) {
      K2 a, b;
      ab(i * t + 0, tt, a, b);
      const K2 a2 = ksqr(a), b2 = ksqr(b), w = wbt[tt];
      const K2 wa = kmul(w, a), wb = kmul(w, b);
      c[0] = kadd(c[0], ksub(kmul(wa, a2), wa));
      const K2 t3 = kmul(wb, a2);
      c[1] = kadd(c[1], ksub(kadd(kadd(t3, t3), t3), wb));
      const K2 t5 = kmul(wa, b2);
      c[2] = kadd(c[2], kadd(kadd(t5, t5), t5));
      c[3] = kadd(c[3], kmul(wb, b2));
    }
    for (int d = 0; d < 4; d++) nc[d] = kadd(nc[d], kmul(g, c[d]));
    g = kmul(g, gamma);
  }
  // eq_beta(X) = eq_beta_r eqb_pref ((1 - beta) + (2 beta - 1) X), likewise eq_ar with alpha
  const K2 bj = beta_a[j], aj_ = alpha[j];
  const K2 cb = kmul(eq_beta_r, eqb_pref), ca = has_inputs ? kmul(eq_r_in, eqa_pref) : kzero();
  const K2 eb0 = kmul(cb, ksub(kone(), bj)), eb1 = kmul(cb, ksub(kadd(bj, bj), kone()));
  const K2 ea0 = kmul(ca, ksub(kone(), aj_)), ea1 = kmul(ca, ksub(kadd(aj_, aj_), kone()));
  K2 o[5] = {kzero(), kzero(), kzero(), kzero(), kzero()};
  K2 fn[4] = {kadd(f_prime, nc[0]), nc[1], nc[2], nc[3]};
  for (int d = 0; d < 4; d++) {
    o[d] = kadd(o[d], kmul(eb0, fn[d]));
    o[d + 1] = kadd(o[d + 1], kmul(eb1, fn[d]));
  }
  if (kt >= 2 && has_inputs) {
    // inner(X) = sum_jm sum_{i >= 1} gamma^i gk^jm sum_tt wat[tt] (a + bX), linear
    K2 in0 = kzero(), in1 = kzero();
    for (size_t jm = 0; jm < t; jm++)
      for (size_t i = 1; i < kt; i++) {
        K2 s0 = kzero(), s1 = kzero();
        for (size_t tt = 0; tt < tail && tt * hs < (size_t)D; tt++) {
          K2 a, b;
          ab(i * t + jm, tt, a, b);
          s0 = kadd(s0, kmul(wat[tt], a));
          s1 = kadd(s1, kmul(wat[tt], b));
        }
        const K2 wgt = kmul(gpi[i], gkj[jm]);
        in0 = kadd(in0, kmul(wgt, s0));
        in1 = kadd(in1, kmul(wgt, s1));
      }
    in0 = kmul(gk, in0);
    in1 = kmul(gk, in1);
    o[0] = kadd(o[0], kmul(ea0, in0));
    o[1] = kadd(o[1], kadd(kmul(ea0, in1), kmul(ea1, in0)));
    o[2] = kadd(o[2], kmul(ea1, in1));
  }
  return std::vector<K2>(o, o + 5);
}

// ---------------- CCS oracle on the device ----------------
CcsOracle::CcsOracle(const Ccs& s, const int8_t* const* Zs, size_t k, const K2* alpha, const K2* beta_a,
                     const K2* beta_r, K2 gamma, size_t ell_d, size_t ell_n, const K2* r_inputs, hipStream_t st,
                     OracleWs* ws)
    : s_(s), st_(st), k_(k), ell_n_(ell_n) {
  (void)ell_d;  // the Ajtai-bit rounds run on the host
  const_cast<Ccs&>(s).ensure_device();
  const size_t n_pad = (size_t)1 << ell_n, m = s.m, n = s.n;
  if (n_pad != n) throw Error(1, "row domain must equal n (power of two)");
  if (!ws) {
    own_ = std::make_unique<OracleWs>();
    ws = own_.get();
  }
  w_ = ws;
  len_ = n_pad;
  zs_.assign(Zs, Zs + k);
  w_->Zptrs.ensure(k * sizeof(void*));
  h2d(w_->Zptrs.p, (const void* const*)Zs, k, st);
  nf_ = (int)s.fvar_mat.size();
  has_eval_ = k >= 2 && r_inputs != nullptr;
  nsc_ = 1 + nf_ + (has_eval_ ? 2 : 0);
  // scalar tables: [eqb][f vars][eqr][ev]
  w_->sc[0].ensure((size_t)nsc_ * n_pad * sizeof(K2));
  w_->sc[1].ensure((size_t)nsc_ * (n_pad / 2) * sizeof(K2));
  K2* T = w_->sc[0].as<K2>();
  w_->bits.ensure(128 * sizeof(K2));
  K2* bits = w_->bits.as<K2>();
  h2d(bits, beta_r, ell_n, st);
  launch_chi(bits, (int)ell_n, T, st);
  w_->z1.ensure(m * sizeof(K2));
  launch_z1(Zs[0], m, w_->z1.as<K2>(), st);
  for (int f = 0; f < nf_; f++) {
    const DevSparse& M = s.dev[s.fvar_mat[f]];
    launch_spmv(M, w_->z1.as<K2>(), T + (size_t)(1 + f) * n_pad, kone(), false, st);
  }
  gamma_k_ = kone();
  for (size_t i = 0; i < k; i++) gamma_k_ = kmul(gamma_k_, gamma);
  if (has_eval_) {
    h2d(bits + 64, r_inputs, ell_n, st);
    launch_chi(bits + 64, (int)ell_n, T + (size_t)(1 + nf_) * n_pad, st);
    // S_alpha = sum_{i>=1} gamma^i sum_rho chi_alpha[rho] Z_i[rho,:]
    std::vector<K2> W((k - 1) * D);
    K2 g = kone();
    for (size_t i = 1; i < k; i++) {
      g = kmul(g, gamma);
      for (int rho = 0; rho < D; rho++) W[(i - 1) * D + rho] = kmul(g, eq_mask(rho, alpha, ell_d));
    }
    w_->Wd.ensure(W.size() * sizeof(K2));
    w_->Sa.ensure(m * sizeof(K2));
    h2d(w_->Wd.p, W.data(), W.size(), st);
    w_->cs_ws.ensure(colsum_ws_bytes(k - 1, m));
    launch_weighted_digit_colsum(w_->Zptrs.as<const int8_t*>() + 1, k - 1, w_->Wd.as<K2>(), m, w_->Sa.as<K2>(),
                                 w_->cs_ws.p, st);
    K2* ev = T + (size_t)(2 + nf_) * n_pad;
    K2 gkj = kone();
    for (size_t j = 0; j < s.t; j++) {
      launch_spmv(s.dev[j], w_->Sa.as<K2>(), ev, gkj, j > 0, st);
      gkj = kmul(gkj, gamma_k_);
    }
  }
  // NC weights wg[w*54 + rho] = chi_{beta_a}(rho) gamma^{w+1}
  std::vector<K2> wg(k * D);
  {
    K2 g = gamma;
    std::vector<K2> wb(D);
    for (int rho = 0; rho < D; rho++) wb[rho] = eq_mask(rho, beta_a, ell_d);
    for (size_t w = 0; w < k; w++) {
      for (int rho = 0; rho < D; rho++) wg[w * D + rho] = kmul(wb[rho], g);
      g = kmul(g, gamma);
    }
  }
  w_->wg.ensure(wg.size() * sizeof(K2));
  h2d(w_->wg.p, wg.data(), wg.size(), st);
  const std::vector<int8_t> wl = weight_limbs(wg.data(), k);
  w_->wl.ensure(wl.size());
  h2d(w_->wl.p, wl.data(), wl.size(), st);
  // chi tables of the variables after each fused round j >= 3 (2^(ell_n - j - 1) entries)
  beta_r_.assign(beta_r, beta_r + ell_n);
  if (has_eval_) rin_.assign(r_inputs, r_inputs + ell_n);
  g_off_.assign(ell_n + 1, 0);
  g_total_ = 0;
  for (size_t j = 3; j < ell_n; j++) {
    g_off_[j] = g_total_;
    g_total_ += (size_t)1 << (ell_n - j - 1);
  }
  if (g_total_) {
    w_->gt.ensure(2 * g_total_ * sizeof(K2));
    launch_chi_suffix(bits, (int)ell_n, 3, g_total_, has_eval_ ? 2 : 1, w_->gt.as<K2>(), st);
  }
  digit2_ = n_pad >= 8 && digit_round2_fits(n_pad / 4, (int)k) && m % 8 == 0 && n_pad <= m;
  w_->nc[0].ensure(k * D * (n_pad / 2) * sizeof(K2));
  // nc[1] also holds round 2's 120 x n/8 digit-monomial sums (larger than the tables when k = 1)
  w_->nc[1].ensure(std::max(k * D * std::max<size_t>(n_pad / 4, 1), 15 * n_pad) * sizeof(K2));
  w_->coeffs.ensure(8 * sizeof(K2));
  w_->ws.ensure(round_ws_bytes());
  NS_CHECK(hipStreamSynchronize(st));  // host vectors above go out of scope
}

// Waits for a round's ready word (k_sum_partials_sig) by polling the coherent mailbox.  After
// 50 ms (a very long round, or a failed launch) the stream wait takes over and reports errors.
static void wait_round(const RoundSignal& sg, hipStream_t st) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  for (uint32_t spin = 1;; spin++) {
    if (__atomic_load_n(sg.flag, __ATOMIC_ACQUIRE) == sg.seq) return;
    __builtin_ia32_pause();
    if ((spin & 1023) == 0 && clk::now() - t0 > std::chrono::milliseconds(50)) break;
  }
  NS_CHECK(hipStreamSynchronize(st));
  if (__atomic_load_n(sg.flag, __ATOMIC_ACQUIRE) != sg.seq) throw Error(1, "round coefficients not published");
}

// ---------------- shared fused rounds (FusedHub) ----------------
namespace {
std::atomic<bool> g_shared_rounds{false};
std::atomic<uint64_t> g_hub_launches{0}, g_hub_jobs{0};

class FusedHub {
 public:
  static FusedHub& current() {
    static std::mutex mu;
    static std::vector<std::unique_ptr<FusedHub>> hubs;
    int dev = 0;
    NS_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    if ((size_t)dev >= hubs.size()) hubs.resize(dev + 1);
    if (!hubs[dev]) hubs[dev].reset(new FusedHub());
    return *hubs[dev];
  }
  hipStream_t stream() const { return st_; }
  // queue a round; returns once it has been issued (by this thread or the combiner at work)
  void submit(const FusedJob& J, hipEvent_t wait_first) {
    std::unique_lock<std::mutex> lk(mu_);
    pending_.push_back({J, wait_first});
    if (combining_) return;
    combining_ = true;
    while (!pending_.empty()) {
      std::vector<Queued> batch;
      batch.swap(pending_);
      lk.unlock();
      try {
        issue(batch);
      } catch (...) {
        lk.lock();
        combining_ = false;
        throw;
      }
      lk.lock();
    }
    combining_ = false;
  }

 private:
  struct Queued {
    FusedJob job;
    hipEvent_t wait_first;
  };
  static constexpr int SLOTS = 8;
  std::mutex mu_;
  std::vector<Queued> pending_;
  bool combining_ = false;
  hipStream_t st_ = nullptr;
  FusedJob* host_[SLOTS] = {};
  FusedJob* dev_[SLOTS] = {};
  hipEvent_t ev_[SLOTS] = {};
  bool live_[SLOTS] = {};
  int next_ = 0;

  FusedHub() {
    NS_CHECK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    for (int i = 0; i < SLOTS; i++) {
      NS_CHECK(hipHostMalloc((void**)&host_[i], FUSED_MAX_JOBS * sizeof(FusedJob), hipHostMallocDefault));
      NS_CHECK(hipMalloc((void**)&dev_[i], FUSED_MAX_JOBS * sizeof(FusedJob)));
      NS_CHECK(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
    }
  }
  // only the combiner calls this (combining_ is held), so slots need no lock
  void issue(std::vector<Queued>& batch) {
    for (auto& q : batch)
      if (q.wait_first) NS_CHECK(hipStreamWaitEvent(st_, q.wait_first, 0));
    for (size_t at = 0; at < batch.size(); at += FUSED_MAX_JOBS) {
      const size_t nj = std::min(batch.size() - at, (size_t)FUSED_MAX_JOBS);
      g_hub_launches++;
      g_hub_jobs += nj;
      if (nj == 1) {
        launch_fused_job(batch[at].job, st_);
        continue;
      }
      const int s = next_;
      next_ = (next_ + 1) % SLOTS;
      if (live_[s]) NS_CHECK(hipEventSynchronize(ev_[s]));  // the slot's last launch has read its jobs
      size_t blocks = 0;
      for (size_t i = 0; i < nj; i++) {
        host_[s][i] = batch[at + i].job;
        blocks += (size_t)batch[at + i].job.nblk;
      }
      NS_CHECK(hipMemcpyAsync(dev_[s], host_[s], nj * sizeof(FusedJob), hipMemcpyHostToDevice, st_));
      launch_fused_jobs(dev_[s], (int)nj, blocks, st_);
      NS_CHECK(hipEventRecord(ev_[s], st_));
      live_[s] = true;
    }
  }
};
}  // namespace

void set_shared_rounds(bool on) { g_shared_rounds = on; }
bool shared_rounds() { return g_shared_rounds; }
FusedHubStats fused_hub_stats() {
  FusedHubStats s;
  s.launches = g_hub_launches;
  s.jobs = g_hub_jobs;
  return s;
}

void CcsOracle::join_hub() {
  if (!hub_used_) return;
  hub_used_ = false;
  if (!w_->hub_out) NS_CHECK(hipEventCreateWithFlags(&w_->hub_out, hipEventDisableTiming));
  NS_CHECK(hipEventRecord(w_->hub_out, FusedHub::current().stream()));
  NS_CHECK(hipStreamWaitEvent(st_, w_->hub_out, 0));
}

RoundSignal CcsOracle::next_signal() {
  RoundSignal sg;
  sg.flag = (uint64_t*)((char*)w_->mbox.get<K2>(8) + 192);
  sg.seq = ++w_->round_seq;
  return sg;
}

std::vector<K2> CcsOracle::round_coeffs() {
  if (pending_) {
    FusedArgs F;
    F.len = len_;
    F.sc_src = w_->sc[cur_sc_].as<K2>();
    F.sc_dst = w_->sc[cur_sc_ ^ 1].as<K2>();
    F.nsc = nsc_;
    F.nf = nf_;
    F.has_eval = has_eval_;
    F.nc_src = w_->nc[cur_nc_].as<K2>();
    F.nc_dst = w_->nc[cur_nc_ ^ 1].as<K2>();
    F.k = (int)k_;
    F.gamma_k = gamma_k_;
    F.f = s_.fpoly;
    F.r = pending_r_;
    F.r7 = fmul7(pending_r_.c1);
    const size_t j = round_;
    if (j < 3 || j >= ell_n_) throw Error(1, "fused round outside the chi tables");
    F.gb = w_->gt.as<K2>() + g_off_[j];
    F.gr = has_eval_ ? w_->gt.as<K2>() + g_total_ + g_off_[j] : nullptr;
    K2* mb = w_->mbox.get<K2>(8);
    const RoundSignal sg = next_signal();
    hipStream_t wst = st_;
    if (shared_rounds()) {
      // the first shared round waits for this stream's work; later ones follow on the hub stream
      hipEvent_t first = nullptr;
      if (!hub_used_) {
        if (!w_->hub_in) NS_CHECK(hipEventCreateWithFlags(&w_->hub_in, hipEventDisableTiming));
        NS_CHECK(hipEventRecord(w_->hub_in, st_));
        first = w_->hub_in;
      }
      FusedHub& hub = FusedHub::current();
      hub_used_ = true;
      hub.submit(fused_job(F, w_->wg.as<K2>(), mb, w_->ws.p, sg), first);
      wst = hub.stream();
    } else {
      launch_round_fused(F, w_->wg.as<K2>(), mb, w_->ws.p, st_, &sg);
    }
    cur_sc_ ^= 1;
    cur_nc_ ^= 1;
    pending_ = false;
    wait_round(sg, wst);
    // eq_beta(t; X) = C eq(beta_j, X) g(t): multiply P by C ((1 - beta_j) + (2 beta_j - 1) X)
    // (a degree-5 term would need f of degree 4, as in the unfused rounds it is dropped), and Q
    // by gamma^k C' ((1 - r_in_j) + (2 r_in_j - 1) X)
    const K2* P = mb;
    const K2 bj = beta_r_[j];
    const K2 b0 = kmul(cb_, ksub(kone(), bj)), b1 = kmul(cb_, ksub(kadd(bj, bj), kone()));
    std::vector<K2> o(5, kzero());
    for (int d = 0; d < 5; d++) {
      o[d] = kmul(b0, P[d]);
      if (d) o[d] = kadd(o[d], kmul(b1, P[d - 1]));
    }
    if (has_eval_) {
      const K2 rj = rin_[j], cg = kmul(gamma_k_, cr_);
      const K2 a0 = kmul(cg, ksub(kone(), rj)), a1 = kmul(cg, ksub(kadd(rj, rj), kone()));
      o[0] = kadd(o[0], kmul(a0, mb[5]));
      o[1] = kadd(o[1], kadd(kmul(a0, mb[6]), kmul(a1, mb[5])));
      o[2] = kadd(o[2], kmul(a1, mb[6]));
    }
    return o;
  }
  RowTables T;
  const K2* sc = w_->sc[cur_sc_].as<K2>();
  T.len = len_;
  T.eqb = sc;
  T.fv = sc + len_;
  T.nf = nf_;
  T.eqr = has_eval_ ? sc + (size_t)(1 + nf_) * len_ : nullptr;
  T.ev = has_eval_ ? sc + (size_t)(2 + nf_) * len_ : nullptr;
  T.nc = round_ <= 1 || (round_ == 2 && digit2_) ? nullptr : w_->nc[cur_nc_].as<K2>();
  T.digits = w_->Zptrs.as<const int8_t*>();
  T.m = s_.m;
  T.k = (int)k_;
  T.gamma_k = gamma_k_;
  // rounds 0-2 from the digits: the coefficient polynomials of the digit monomials, for the int8
  // MFMA path (and round 1's vector path).  With y(X) = sum_i L_i(X) d_i (digits in {-1, 0, 1}, so
  // d^2 = |d| and d^3 = d), N(y) = y^3 - y = sum_i (L_i^3 - L_i) d_i + 3 sum_{i != j} L_i^2 L_j |d_i| d_j
  // + 6 sum_{i < j < l} L_i L_j L_l d_i d_j d_l.
  typedef std::array<K2, 4> P4;
  auto lin = [](K2 c0, K2 c1) { return P4{c0, c1, kzero(), kzero()}; };
  auto mul = [](const P4& a, const P4& b) {
    P4 o{kzero(), kzero(), kzero(), kzero()};
    for (int i = 0; i < 4; i++)
      for (int j = 0; i + j < 4; j++) o[i + j] = kadd(o[i + j], kmul(a[i], b[j]));
    return o;
  };
  auto scale = [](const P4& a, uint64_t c) { return P4{kscale(a[0], c), kscale(a[1], c), kscale(a[2], c), kscale(a[3], c)}; };
  std::vector<K2> pc;
  auto push = [&](const P4& p) { pc.insert(pc.end(), p.begin(), p.end()); };
  auto monomial_polys = [&](const std::vector<P4>& L) {
    const int nd = (int)L.size();
    for (int i = 0; i < nd; i++) {
      P4 c = mul(mul(L[i], L[i]), L[i]);
      for (int d = 0; d < 4; d++) c[d] = ksub(c[d], L[i][d]);
      push(c);
    }
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++)
        if (i != j) push(scale(mul(mul(L[i], L[i]), L[j]), 3));
    for (int i = 0; i < nd; i++)
      for (int j = i + 1; j < nd; j++)
        for (int l = j + 1; l < nd; l++) push(scale(mul(mul(L[i], L[j]), L[l]), 6));
  };
  if (round_ == 0) {
    // lo, hi, |lo| hi, lo |hi| of a digit pair: N(lo + (hi - lo) X) = lo A + hi B + 3 |lo| hi C + 3 lo |hi| D
    auto ki = [](int64_t v) { return K2{from_i64(v), 0}; };
    push(P4{kzero(), ki(-2), ki(3), ki(-1)});
    push(P4{kzero(), ki(-1), kzero(), ki(1)});
    push(P4{kzero(), ki(3), ki(-6), ki(3)});
    push(P4{kzero(), kzero(), ki(3), ki(-3)});
  } else if (round_ == 1) {
    // a = (1-r0) d0 + r0 d1, a' = (1-r0) d2 + r0 d3, y = (1-X) a + X a'
    const K2 al = ksub(kone(), r0_), be = r0_;
    monomial_polys({lin(al, kneg(al)), lin(be, kneg(be)), lin(kzero(), al), lin(kzero(), be)});
    T.digit_round1 = true;
  } else if (round_ == 2 && digit2_) {
    // the pair's entries fold digits 8t..8t+3 and 8t+4..8t+7 by chi_(r0, r1)
    const K2 a0 = ksub(kone(), r0_), a1 = ksub(kone(), r1_);
    const K2 w[4] = {kmul(a0, a1), kmul(r0_, a1), kmul(a0, r1_), kmul(r0_, r1_)};
    std::vector<P4> L;
    for (int i = 0; i < 4; i++) L.push_back(lin(w[i], kneg(w[i])));
    for (int i = 0; i < 4; i++) L.push_back(lin(kzero(), w[i]));
    monomial_polys(L);
    T.digit_round2 = true;
  }
  if (!pc.empty()) {
    w_->pc.ensure(pc.size() * sizeof(K2));
    h2d(w_->pc.p, pc.data(), pc.size(), st_);
    T.pc = w_->pc.as<K2>();
    // the monomial sums go where no live table is: nc[0] before the tables exist, else the other one
    DevBuf& ub = round_ <= 1 ? w_->nc[0] : w_->nc[cur_nc_ ^ 1];
    T.wl = w_->wl.as<const int8_t>();
    T.ubuf = ub.as<K2>();
    T.ubuf_bytes = ub.bytes;
  }
  K2* mb = w_->mbox.get<K2>(8);
  const RoundSignal sg = next_signal();
  launch_round_eval(T, s_.fpoly, w_->wg.as<K2>(), round_ == 0, mb, w_->ws.p, st_, &sg);
  wait_round(sg, st_);
  return std::vector<K2>(mb, mb + 5);
}

void CcsOracle::flush() {
  join_hub();
  if (!pending_) return;
  const size_t prev = len_ * 2;
  launch_fold_k(w_->sc[cur_sc_].as<K2>(), (size_t)nsc_, prev, w_->sc[cur_sc_ ^ 1].as<K2>(), pending_r_, st_);
  launch_fold_k(w_->nc[cur_nc_].as<K2>(), k_ * D, prev, w_->nc[cur_nc_ ^ 1].as<K2>(), pending_r_, st_);
  cur_sc_ ^= 1;
  cur_nc_ ^= 1;
  pending_ = false;
}

void CcsOracle::fold(K2 r) {
  flush();
  if (round_ < ell_n_) {
    cb_ = kmul(cb_, eq_lin(r, beta_r_[round_]));
    if (has_eval_) cr_ = kmul(cr_, eq_lin(r, rin_[round_]));
  }
  const size_t half = len_ / 2;
  if (round_ >= 2 && nsc_ <= FUSED_MAX_SC && half >= 1 && !(round_ == 2 && digit2_)) {
    pending_ = true;
    pending_r_ = r;
    len_ = half;
    round_++;
    return;
  }
  int nsc_dst = cur_sc_ ^ 1;
  launch_fold_k(w_->sc[cur_sc_].as<K2>(), (size_t)nsc_, len_, w_->sc[nsc_dst].as<K2>(), r, st_);
  if (round_ == 0) {
    // the NC tables stay as digits through round 1 (evaluated from them with r0)
    r0_ = r;
    if (half < 2) {  // a one-round table: materialize now
      launch_fold_digits(w_->Zptrs.as<const int8_t*>(), (int)k_, s_.m, len_, w_->nc[0].as<K2>(), r, st_);
      cur_nc_ = 0;
    }
  } else if (round_ == 1 && digit2_) {
    r1_ = r;  // the NC tables stay as digits through round 2
  } else if (round_ == 2 && digit2_) {
    launch_fold_digits3(w_->Zptrs.as<const int8_t*>(), (int)k_, s_.m, len_ * 4, w_->nc[0].as<K2>(), r0_, r1_, r, st_);
    cur_nc_ = 0;
  } else if (round_ == 1 && len_ * 2 >= 4) {
    r1_ = r;
    launch_fold_digits2(w_->Zptrs.as<const int8_t*>(), (int)k_, s_.m, len_ * 2, w_->nc[0].as<K2>(), r0_, r, st_);
    cur_nc_ = 0;
  } else {
    int dst = cur_nc_ ^ 1;
    launch_fold_k(w_->nc[cur_nc_].as<K2>(), k_ * D, len_, w_->nc[dst].as<K2>(), r, st_);
    cur_nc_ = dst;
  }
  cur_sc_ = nsc_dst;
  len_ = half;
  round_++;
}

std::vector<K2> CcsOracle::y_eval(const std::vector<K2>& r, DevBuf& ws, DevBuf& vjs) {
  flush();
  const size_t m = s_.m, t = s_.t, n = s_.n;
  w_->bits.ensure(128 * sizeof(K2));
  w_->chi.ensure(((size_t)1 << ell_n_) * sizeof(K2));
  h2d(w_->bits.p, r.data(), r.size(), st_);
  launch_chi(w_->bits.as<K2>(), (int)ell_n_, w_->chi.as<K2>(), st_);
  vjs.ensure(t * m * sizeof(K2));
  size_t n_eff = std::min(n, (size_t)1 << ell_n_);
  for (size_t j = 0; j < t; j++) launch_spmv_t(s_.dev[j], w_->chi.as<K2>(), n_eff, vjs.as<K2>() + j * m, st_);
  ws.ensure(ydigits_ws_bytes(k_, t, m));
  w_->ydev.ensure(k_ * t * D * sizeof(K2));
  launch_ydigits(w_->Zptrs.as<const int8_t*>(), k_, vjs.as<K2>(), t, m, w_->ydev.as<K2>(), ws.p, st_);
  std::vector<K2> out(k_ * t * D);
  d2h(w_->stage, out.data(), w_->ydev.p, out.size(), st_);
  return out;
}

// ---------------- canonical encoding ----------------
struct Enc {
  std::vector<uint64_t> v;
  void u(uint64_t x) { v.push_back(x); }
  void k(K2 x) {
    v.push_back(x.c0);
    v.push_back(x.c1);
  }
  void digest(const uint8_t d[32]) {
    for (int i = 0; i < 4; i++) {
      uint64_t w = 0;
      for (int j = 0; j < 8; j++) w |= (uint64_t)d[8 * i + j] << (8 * j);
      v.push_back(w);
    }
  }
  void me(const MeHost& e) {
    for (int i = 0; i < KAPPA * D; i++) u(e.c[i]);
    u(e.m_in);
    for (auto x : e.X) u(x);
    u(e.r.size());
    for (auto& x : e.r) k(x);
    u(e.t);
    for (size_t j = 0; j < e.t; j++) {
      u(DPAD);
      for (int i = 0; i < DPAD; i++) k(e.y[j * DPAD + i]);
    }
    u(e.t);
    for (auto& x : e.y_scalars) k(x);
    digest(e.fold_digest);
    u(0);
    u(0);
    u(0);
  }
};

// ---------------- prover ----------------
Prover::Prover(const Ccs& s, const Pp& pp) : s_(&s), pp_(&pp) {
  const_cast<Ccs&>(s).ensure_device();
  if (pp.m != s.m) throw Error(1, "PP width does not match CCS m");
  params_ = Params::goldilocks_auto_r1cs_ccs(s.n);
  NS_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

void Prover::rebind(const Ccs& s, const Pp& pp) {
  const_cast<Ccs&>(s).ensure_device();
  if (pp.m != s.m) throw Error(1, "PP width does not match CCS m");
  NS_CHECK(hipStreamSynchronize(stream_));
  s_ = &s;
  pp_ = &pp;
  params_ = Params::goldilocks_auto_r1cs_ccs(s.n);
}

Prover::~Prover() { (void)hipStreamDestroy(stream_); }

void Prover::commit_launch(const int8_t* const* Zs_dev, size_t nw) {
  ws_ajtai_.ensure(ajtai_workspace_bytes(s_->m, nw));
  ptrs_.ensure(nw * sizeof(void*));
  cout_.ensure(nw * KAPPA * D * 8);
  h2d(ptrs_.p, (const void* const*)Zs_dev, nw, stream_);
  ajtai_commit(pp_->limbs.as<int8_t>(), KAPPA, s_->m, ptrs_.as<const int8_t*>(), nw, cout_.as<uint64_t>(),
               ws_ajtai_.p, stream_);
}

void Prover::commit_fetch(size_t nw, uint64_t* host_out) {
  d2h(ows_.stage, host_out, cout_.p, nw * KAPPA * D, stream_);
}

void Prover::commit_device(const int8_t* const* Zs_dev, size_t nw, uint64_t* host_out) {
  commit_launch(Zs_dev, nw);
  commit_fetch(nw, host_out);
}

static size_t log2_pow2(size_t n) {
  size_t l = 0;
  while (((size_t)1 << l) < n) l++;
  return l < 1 ? 1 : l;
}

// ---------------- transcript steps shared by the prover and the verifier ----------------
// bind_header_and_instances_with_digest + bind_me_inputs (neo-reductions/src/engines/utils.rs:83-181)
void bind_step_header(Transcript& tr, const Ccs& s, const Params& params, size_t ell, size_t d_sc, int32_t slack,
                      const uint64_t* x, size_t m_in, const uint64_t* c_mcs, const std::vector<MeHost>& acc) {
  tr.append_message("phase/pi-ccs", (const uint8_t*)nullptr, 0);
  tr.append_message("neo/ccs/header/v1", (const uint8_t*)nullptr, 0);
  uint64_t hdr[6] = {64, 2, params.lambda, ell, d_sc, (uint64_t)(slack < 0 ? -slack : slack)};
  tr.append_u64s("ccs/header", hdr, 6);
  uint8_t sign = slack >= 0 ? 1 : 0;
  tr.append_message("ccs/slack_sign", &sign, 1);
  tr.append_message("neo/ccs/instances", (const uint8_t*)nullptr, 0);
  uint64_t dd[3] = {s.n, s.m, s.t};
  tr.append_u64s("dims", dd, 3);
  for (int i = 0; i < 4; i++) tr.append_fields("mat_digest", &s.digest[i], 1);
  tr.append_message("neo/ccs/poly", (const uint8_t*)nullptr, 0);
  uint64_t ar = s.t, tl = s.term_coeff.size();
  tr.append_u64s("arity", &ar, 1);
  tr.append_u64s("terms_len", &tl, 1);
  std::vector<size_t> order(tl);
  for (size_t i = 0; i < tl; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return s.term_exps[a] < s.term_exps[b]; });
  for (size_t i : order) {
    tr.append_fields("coeff", &s.term_coeff[i], 1);
    std::vector<uint64_t> ex(s.term_exps[i].begin(), s.term_exps[i].end());
    tr.append_u64s("exps", ex.data(), ex.size());
  }
  tr.append_fields("x", x, m_in);
  uint64_t mi = m_in;
  tr.append_u64s("m_in", &mi, 1);
  tr.append_fields("c_data", c_mcs, KAPPA * D);
  tr.append_message("neo/ccs/me_inputs/v2", (const uint8_t*)nullptr, 0);
  uint64_t cnt = acc.size();
  tr.append_u64s("me_count", &cnt, 1);
  std::vector<uint64_t> buf;
  for (auto& e : acc) {
    tr.append_fields("c_data_in", e.c, KAPPA * D);
    mi = e.m_in;
    tr.append_u64s("m_in_in", &mi, 1);
    buf.clear();
    for (auto& v : e.r) {
      buf.push_back(v.c0);
      buf.push_back(v.c1);
    }
    tr.append_fields("r_in", buf.data(), buf.size());
    buf.clear();
    for (auto& v : e.y) {
      buf.push_back(v.c0);
      buf.push_back(v.c1);
    }
    tr.append_fields("y_elem", buf.data(), buf.size());
  }
}

// sample_challenges (utils.rs:184-212): alpha (ell_d), beta (ell), gamma, each K = two challenges
StepChallenges sample_challenges(Transcript& tr, size_t ell_d, size_t ell) {
  StepChallenges ch;
  tr.append_message("neo/ccs/chals/v1", (const uint8_t*)nullptr, 0);
  auto draw = [&]() {
    uint64_t c2[2];
    tr.challenge_fields("chal/k", 2, c2);
    return K2{c2[0], c2[1]};
  };
  for (size_t i = 0; i < ell_d; i++) ch.alpha.push_back(draw());
  for (size_t i = 0; i < ell; i++) ch.beta.push_back(draw());
  ch.gamma = draw();
  return ch;
}

// claimed initial sum T from the ME inputs (optimized_engine/common.rs:839-947), core rows only
K2 claimed_initial_sum(const Ccs& s, const StepChallenges& ch, const std::vector<MeHost>& acc, size_t ell_d) {
  const size_t nacc = acc.size(), kt = nacc + 1;
  if (kt < 2) return kzero();
  std::vector<K2> chi_a = chi_host(ch.alpha.data(), ell_d);
  K2 gk = kone();
  for (size_t i = 0; i < kt; i++) gk = kmul(gk, ch.gamma);
  K2 inner = kzero();
  for (size_t j = 0; j < s.t; j++) {
    K2 gkj = kone();
    for (size_t q = 0; q < j; q++) gkj = kmul(gkj, gk);
    K2 gi = kone();
    for (size_t a = 0; a < nacc; a++) {
      gi = kmul(gi, ch.gamma);  // gamma^{i-1}, i = a + 2
      K2 ye = kzero();
      for (int rho = 0; rho < DPAD; rho++) ye = kadd(ye, kmul(acc[a].y[j * DPAD + rho], chi_a[rho]));
      inner = kadd(inner, kmul(kmul(gi, gkj), ye));
    }
  }
  return kmul(gk, inner);
}

// bind_rlc_inputs (neo-fold/src/shard.rs:1125-1168)
void bind_rlc_inputs(Transcript& tr, const char* lane, size_t idx, const std::vector<MeHost>& ins) {
  tr.append_message("fold/rlc_inputs/v1", lane);
  uint64_t v = idx;
  tr.append_u64s("step_idx", &v, 1);
  v = ins.size();
  tr.append_u64s("me_count", &v, 1);
  for (auto& e : ins) {
    tr.append_fields("c_data", e.c, KAPPA * D);
    v = e.m_in;
    tr.append_u64s("m_in", &v, 1);
    tr.append_message("me_fold_digest", e.fold_digest, 32);
    for (auto& x : e.r) tr.append_k("r_limb", x);
    tr.append_fields("X", e.X.data(), e.X.size());
    for (auto& x : e.y) tr.append_k("y_elem", x);
    for (auto& x : e.y_scalars) tr.append_k("y_scalar", x);
    v = 0;
    tr.append_u64s("c_step_coords_len", &v, 1);
    tr.append_fields("c_step_coords", nullptr, 0);
    tr.append_u64s("u_offset", &v, 1);
    tr.append_u64s("u_len", &v, 1);
  }
}

// sample_rot_rhos_n (neo-reductions/src/common.rs:410-573): coefficient vectors a_i in [-2, 2]
// drawn from transcript digests, rho_i = rot(a_i) (the identity when there is one input)
void sample_rot_rhos(Transcript& tr, size_t kt, std::vector<int8_t>& rho_a, std::vector<uint64_t>& rhos) {
  rho_a.assign(kt * D, 0);
  rhos.assign(kt * D * D, 0);
  for (size_t i = 0; i < kt; i++) {
    tr.append_msg_u64("rlc/rot/index", i);
    // draw_alphabet_vector (neo-reductions/src/common.rs:410-439)
    static const int8_t A5[5] = {-2, -1, 0, 1, 2};
    const uint32_t bucket = (65536u / 5u) * 5u;
    size_t got = 0;
    uint64_t ctr = i;
    while (got < (size_t)D) {
      tr.append_msg_u64("rlc/rot/chunk", ctr);
      uint8_t dig[32];
      tr.digest32(dig);
      for (int w = 0; w < 16 && got < (size_t)D; w++) {
        uint32_t x = (uint32_t)dig[2 * w] | ((uint32_t)dig[2 * w + 1] << 8);
        if (x < bucket) rho_a[i * D + got++] = A5[x % 5];
      }
      ctr++;
    }
    uint64_t* R = &rhos[i * D * D];
    if (kt == 1) {
      for (int q = 0; q < D; q++) R[q * D + q] = 1;
    } else {
      // rot_from_coeffs: column j = X^j * a mod Phi_81
      uint64_t col[D], nxt[D];
      for (int r = 0; r < D; r++) col[r] = from_i64(rho_a[i * D + r]);
      for (int j = 0; j < D; j++) {
        for (int r = 0; r < D; r++) R[r * D + j] = col[r];
        uint64_t last = col[D - 1];
        nxt[0] = fneg(last);
        for (int r = 1; r < D; r++) nxt[r] = col[r - 1];
        nxt[27] = fsub(nxt[27], last);
        memcpy(col, nxt, sizeof(col));
      }
    }
  }
}

// ---------------- shared-bus openings on the device ----------------
void Prover::bus_weights(const std::vector<K2>& r, const BusLayout& L) {
  rbits_.ensure(std::max<size_t>(r.size(), 1) * sizeof(K2));
  busw_.ensure(L.chunk * sizeof(K2));
  h2d(rbits_.p, r.data(), r.size(), stream_);
  launch_bus_weights((uint32_t)L.m_in, (uint32_t)L.chunk, rbits_.as<K2>(), (int)r.size(), busw_.as<K2>(), stream_);
}

// append_bus_openings_to_me_instance (cpu_bus.rs:195-320) for every (plane, ME) pair at the
// point whose weights bus_weights() left in busw_
void Prover::open_bus_planes(const std::vector<const int8_t*>& planes, const BusLayout& L,
                             const std::vector<MeHost*>& outs, size_t t_core) {
  const size_t bc = L.bus_cols;
  if (!bc) return;
  for (size_t p0 = 0; p0 < planes.size(); p0 += 16) {
    const size_t np = std::min<size_t>(16, planes.size() - p0);
    PlanePtrs pl{};
    for (size_t q = 0; q < np; q++) pl.p[q] = planes[p0 + q];
    busy_.ensure(np * bc * 64 * sizeof(K2));
    busys_.ensure(np * bc * sizeof(K2));
    launch_bus_open_planes(pl, (int)np, s_->m, L.bus_base, (uint32_t)L.chunk, (uint32_t)bc, busw_.as<K2>(),
                           busy_.as<K2>(), busys_.as<K2>(), stream_);
    std::vector<K2> y(np * bc * 64), ys(np * bc);
    d2h(ows_.stage, y.data(), busy_.p, y.size(), stream_);
    d2h(ows_.stage, ys.data(), busys_.p, ys.size(), stream_);
    for (size_t q = 0; q < np; q++) {
      MeHost& e = *outs[p0 + q];
      std::copy(y.begin() + q * bc * 64, y.begin() + (q + 1) * bc * 64, e.y.begin() + t_core * DPAD);
      std::copy(ys.begin() + q * bc, ys.begin() + (q + 1) * bc, e.y_scalars.begin() + t_core);
    }
  }
}

void Prover::open_bus_i32(const int32_t* zmix, const BusLayout& L, MeHost& out, size_t t_core) {
  const size_t bc = L.bus_cols;
  if (!bc) return;
  busy_.ensure(bc * 64 * sizeof(K2));
  busys_.ensure(bc * sizeof(K2));
  launch_bus_open_i32(zmix, s_->m, L.bus_base, (uint32_t)L.chunk, (uint32_t)bc, busw_.as<K2>(), busy_.as<K2>(),
                      busys_.as<K2>(), stream_);
  d2h(ows_.stage, out.y.data() + t_core * DPAD, busy_.p, bc * 64, stream_);
  d2h(ows_.stage, out.y_scalars.data() + t_core, busys_.p, bc, stream_);
}

// v_j = M_j^T chi_r for every matrix (the vectors precompute_for_r contracts the digits with)
void Prover::vjs_at(const std::vector<K2>& r, DevBuf& vjs) {
  const Ccs& s = *s_;
  const size_t ell = r.size();
  rbits_.ensure(std::max<size_t>(ell, 1) * sizeof(K2));
  rchi_.ensure(((size_t)1 << ell) * sizeof(K2));
  h2d(rbits_.p, r.data(), ell, stream_);
  launch_chi(rbits_.as<K2>(), (int)ell, rchi_.as<K2>(), stream_);
  vjs.ensure(s.t * s.m * sizeof(K2));
  const size_t n_eff = std::min(s.n, (size_t)1 << ell);
  for (size_t j = 0; j < s.t; j++) launch_spmv_t(s.dev[j], rchi_.as<K2>(), n_eff, vjs.as<K2>() + j * s.m, stream_);
}

// y[w][j][rho] = sum_c Z_w[rho][c] v_j[c] for device digit planes (int8 MFMA contraction)
void Prover::y_launch(const std::vector<const int8_t*>& planes, const DevBuf& vjs) {
  const Ccs& s = *s_;
  const size_t nw = planes.size();
  yptr_.ensure(nw * sizeof(void*));
  h2d(yptr_.p, (const void* const*)planes.data(), nw, stream_);
  ws_yd_.ensure(ydigits_ws_bytes(nw, s.t, s.m));
  ydev_.ensure(nw * s.t * D * sizeof(K2));
  launch_ydigits(yptr_.as<const int8_t*>(), nw, vjs.as<K2>(), s.t, s.m, ydev_.as<K2>(), ws_yd_.p, stream_);
}

std::vector<K2> Prover::y_fetch(size_t nw) {
  std::vector<K2> out(nw * s_->t * D);
  d2h(ows_.stage, out.data(), ydev_.p, out.size(), stream_);
  return out;
}

std::vector<K2> Prover::y_products(const std::vector<const int8_t*>& planes, const DevBuf& vjs) {
  y_launch(planes, vjs);
  return y_fetch(planes.size());
}

void Prover::head_cols(const int8_t* const* planes, size_t np, size_t m_in, std::vector<std::vector<uint64_t>*> Xs) {
  for (size_t p = 0; p < np; p++) Xs[p]->assign((size_t)D * m_in, 0);
  if (!m_in || !np) return;
  for (size_t p0 = 0; p0 < np; p0 += 16) {
    const size_t cnt = std::min<size_t>(16, np - p0);
    PlanePtrs pl{};
    for (size_t p = 0; p < cnt; p++) pl.p[p] = planes[p0 + p];
    head_.ensure(cnt * D * m_in);
    launch_head_cols(pl, (int)cnt, s_->m, (int)m_in, head_.as<int8_t>(), stream_);
    std::vector<int8_t> tmp(cnt * D * m_in);
    d2h(ows_.stage, tmp.data(), head_.p, tmp.size(), stream_);
    for (size_t p = 0; p < cnt; p++)
      for (size_t i = 0; i < (size_t)D * m_in; i++) (*Xs[p0 + p])[i] = from_i64(tmp[p * D * m_in + i]);
  }
}

// bind_rlc_inputs (shard.rs:1125-1168) + sample_rot_rhos_n + Pi_RLC + Pi_DEC of one lane
// (prove_rlc_dec_lane, shard.rs:1170-1331).  ins' witnesses are Zs (device planes); the children's
// digit planes go to `bank` (16 planes); vjs: M_j^T chi_r at the lane's point r = ins[0].r.
void Prover::rlc_dec_lane(Transcript& tr, const char* lane, size_t idx, const std::vector<MeHost>& ins,
                          const std::vector<const int8_t*>& Zs, DevBuf* bank, const DevBuf& vjs, const BusLayout& L,
                          LaneOut& out) {
  const Ccs& s = *s_;
  const size_t m = s.m, t = s.t, k_dec = params_.k_rho, kt = ins.size(), m_in = ins[0].m_in;
  const size_t t_tot = t + L.bus_cols;
  hipStream_t st = stream_;
  bind_rlc_inputs(tr, lane, idx, ins);
  if ((uint64_t)kt * 216u * (params_.b - 1) >= ((uint64_t)1 << params_.k_rho))
    throw Error(5, "RLC norm bound violated");
  std::vector<int8_t> rho_a;
  sample_rot_rhos(tr, kt, rho_a, out.rhos);
  std::vector<int8_t> rho_i8(kt * 64 * 64, 0);  // [i][row][col] padded to 64 x 64
  for (size_t i = 0; i < kt; i++)
    for (int q = 0; q < D * D; q++) {
      const int64_t v = to_balanced(out.rhos[i * D * D + q]);
      if (v < -127 || v > 127) throw Error(5, "rotation entry out of int8 range");
      rho_i8[(i * 64 + q / D) * 64 + q % D] = (int8_t)v;
    }
  std::vector<K2> bpow(D);
  bpow[0] = kone();
  for (int i = 1; i < D; i++) bpow[i] = kadd(bpow[i - 1], bpow[i - 1]);
  MeHost& parent = out.parent;
  parent.m_in = m_in;
  parent.t = kt == 1 ? ins[0].t : t_tot;
  parent.r = ins[0].r;
  memcpy(parent.fold_digest, ins[0].fold_digest, 32);
  parent.X.assign((size_t)D * m_in, 0);
  parent.y.assign(parent.t * DPAD, kzero());
  parent.y_scalars.assign(parent.t, kzero());
  // Z_mix on the device (the identity rotation when kt = 1: Z_mix = Z_0)
  zmix_.ensure((size_t)D * m * 4);
  rho_dev_.ensure(rho_i8.size());
  zptr_.ensure(kt * sizeof(void*));
  h2d(rho_dev_.p, rho_i8.data(), rho_i8.size(), st);
  h2d(zptr_.p, (const void* const*)Zs.data(), kt, st);
  launch_rlc_zmix(rho_dev_.as<int8_t>(), zptr_.as<const int8_t*>(), kt, m, zmix_.as<int32_t>(), st);
  const int32_t* zsrc = zmix_.as<int32_t>();
  if (kt == 1) {
    memcpy(parent.c, ins[0].c, sizeof(parent.c));
    parent.X = ins[0].X;
    parent.y = ins[0].y;
    for (size_t j = 0; j < parent.t; j++) {
      K2 sc = kzero();
      for (int rho = 0; rho < D; rho++) sc = kadd(sc, kmul(parent.y[j * DPAD + rho], bpow[rho]));
      parent.y_scalars[j] = sc;
    }
  } else {
    for (size_t i = 0; i < kt; i++) {
      const uint64_t* R = &out.rhos[i * D * D];
      for (int rr = 0; rr < D; rr++)
        for (int kk = 0; kk < D; kk++) {
          uint64_t co = R[rr * D + kk];
          if (!co) continue;
          for (size_t c = 0; c < m_in; c++)
            parent.X[rr * m_in + c] = fadd(parent.X[rr * m_in + c], fmul(co, ins[i].X[kk * m_in + c]));
        }
    }
    // y_j = sum_i rot(a_i) y_ij and c = sum_i rot(a_i) c_i: rot(a) v is a(X) v(X) mod Phi_81
    // with the small coefficients a_i in [-2, 2] (mix_small_coeffs); core rows only
    std::vector<const uint64_t*> vp(kt);
    std::vector<uint64_t> vbuf(kt * D * 2), outc(D);
    for (size_t j = 0; j < t; j++) {
      for (int comp = 0; comp < 2; comp++) {
        for (size_t i = 0; i < kt; i++) {
          const K2* y = &ins[i].y[j * DPAD];
          for (int v = 0; v < D; v++) vbuf[i * D + v] = comp ? y[v].c1 : y[v].c0;
          vp[i] = &vbuf[i * D];
        }
        mix_small_coeffs(rho_a.data(), vp.data(), kt, outc.data());
        for (int q = 0; q < D; q++) (comp ? parent.y[j * DPAD + q].c1 : parent.y[j * DPAD + q].c0) = outc[q];
      }
    }
    for (int col = 0; col < KAPPA; col++) {
      for (size_t i = 0; i < kt; i++) vp[i] = &ins[i].c[col * D];
      mix_small_coeffs(rho_a.data(), vp.data(), kt, &parent.c[col * D]);
    }
    for (size_t j = 0; j < t; j++) {
      K2 sc = kzero();
      for (int rho = 0; rho < D; rho++) sc = kadd(sc, kmul(parent.y[j * DPAD + rho], bpow[rho]));
      parent.y_scalars[j] = sc;
    }
    if (L.bus_cols) {
      bus_weights(parent.r, L);
      open_bus_i32(zsrc, L, parent, t);
    }
  }

  // ---- Pi_DEC ----
  if (k_dec > 16) throw Error(1, "k_rho > 16 is not supported");
  std::vector<const int8_t*> pptr(k_dec);
  DecPlanes dec_pl{};
  for (size_t i = 0; i < k_dec; i++) {
    bank[i].ensure((size_t)D * m);
    pptr[i] = bank[i].as<int8_t>();
    dec_pl.p[i] = bank[i].as<int8_t>();
  }
  flags_.ensure(4);
  launch_dec_split(zsrc, (size_t)D * m, (int)k_dec, dec_pl, flags_.as<int>(), st);
  // the children's commitments and y products queue behind the split; the host waits once (the
  // range check reads the split's flags after them: an out-of-range split fails the step either way)
  commit_launch(pptr.data(), k_dec);
  y_launch(pptr, vjs);
  uint32_t dmask = 0;
  d2h(ows_.stage, &dmask, flags_.p, 1, st);
  if (dec_out_of_range(dmask, (int)k_dec)) throw Error(6, "step " + std::to_string(idx) + ": DEC split out of range");
  std::vector<uint64_t> child_c(k_dec * KAPPA * D, 0);
  commit_fetch(k_dec, child_c.data());
  std::vector<K2> y_child = y_fetch(k_dec);
  out.children.assign(k_dec, MeHost{});
  {
    std::vector<std::vector<uint64_t>*> xs;
    for (auto& c : out.children) xs.push_back(&c.X);
    head_cols(pptr.data(), k_dec, m_in, xs);
  }
  for (size_t i = 0; i < k_dec; i++) {
    MeHost& e = out.children[i];
    if (dec_plane_nonzero(dmask, (int)i)) memcpy(e.c, &child_c[i * KAPPA * D], sizeof(e.c));
    else memset(e.c, 0, sizeof(e.c));
    e.m_in = m_in;
    e.t = t_tot;
    e.r = parent.r;
    e.y.assign(t_tot * DPAD, kzero());
    e.y_scalars.assign(t_tot, kzero());
    for (size_t j = 0; j < t; j++) {
      K2 sc = kzero();
      for (int rho = 0; rho < D; rho++) {
        K2 v = y_child[(i * t + j) * D + rho];
        e.y[j * DPAD + rho] = v;
        sc = kadd(sc, kmul(v, bpow[rho]));
      }
      e.y_scalars[j] = sc;
    }
    memcpy(e.fold_digest, parent.fold_digest, 32);
  }
  if (L.bus_cols) {
    bus_weights(parent.r, L);
    std::vector<MeHost*> outs;
    for (auto& c : out.children) outs.push_back(&c);
    open_bus_planes(pptr, L, outs, t);
  }
  {
    bool ok_y = true, ok_x = true, ok_c = true;
    for (size_t j = 0; j < t && ok_y; j++)
      for (int q = 0; q < DPAD; q++) {
        K2 lhs = kzero();
        uint64_t pw = 1;
        for (size_t i = 0; i < k_dec; i++) {
          lhs = kadd(lhs, kscale(out.children[i].y[j * DPAD + q], pw));
          pw = fadd(pw, pw);
        }
        if (!keq(lhs, parent.y[j * DPAD + q])) {
          ok_y = false;
          break;
        }
      }
    for (size_t q = 0; q < (size_t)D * m_in; q++) {
      uint64_t lhs = 0, pw = 1;
      for (size_t i = 0; i < k_dec; i++) {
        lhs = fadd(lhs, fmul(out.children[i].X[q], pw));
        pw = fadd(pw, pw);
      }
      if (lhs != parent.X[q]) ok_x = false;
    }
    for (int q = 0; q < KAPPA * D; q++) {
      uint64_t lhs = out.children[0].c[q], pw = 2;
      for (size_t i = 1; i < k_dec; i++) {
        lhs = fadd(lhs, fmul(out.children[i].c[q], pw));
        pw = fadd(pw, pw);
      }
      if (lhs != parent.c[q]) ok_c = false;
    }
    if (!(ok_y && ok_x && ok_c))
      throw Error(7, std::string(strcmp(lane, "val") ? "DEC" : "DEC(val)") + " public check failed at step " +
                         std::to_string(idx) + " (y=" + std::to_string(ok_y) + ", X=" + std::to_string(ok_x) +
                         ", c=" + std::to_string(ok_c) + ")");
  }
  out.planes = pptr;
}

// ---- lane-level Pi_RLC / Pi_DEC (rlc_with_commit, dec_children_with_commit; api.rs:131-324) ----
// parent = sum_i rot(a_i) * ins[i] (every y row, so bus openings mix with the core rows), commitment
// mixed as mix_rhos_commits (session.rs:311-319); Z_mix = sum_i rot(a_i) Z_i into zmix_dev.
// One input is the reference's fast path: the parent is the input (y_scalars recomputed).
void Prover::rlc_api(const std::vector<MeHost>& ins, const int8_t* rho_coeffs, const int8_t* const* Z_dev,
                     MeHost& parent, int32_t* zmix_dev) {
  const size_t kt = ins.size(), m = s_->m;
  if (kt == 0) throw Error(1, "RLC needs at least one input");
  const size_t t = ins[0].t, m_in = ins[0].m_in;
  for (auto& e : ins)
    if (e.t != t || e.m_in != m_in || e.r.size() != ins[0].r.size())
      throw Error(1, "RLC inputs must share t, m_in and the point r");
  std::vector<int8_t> rho_i8(kt * 64 * 64, 0);
  std::vector<int8_t> a(kt * D, 0);
  for (size_t i = 0; i < kt; i++) {
    for (int q = 0; q < D; q++) {
      a[i * D + q] = kt == 1 ? (q == 0 ? 1 : 0) : rho_coeffs[i * D + q];
      if (a[i * D + q] < -2 || a[i * D + q] > 2) throw Error(1, "rotation coefficients must lie in [-2, 2]");
    }
    // rot(a_i): column j = X^j * a_i mod Phi_81 (small integers: |entry| <= 4)
    int64_t col[D], nxt[D];
    for (int r = 0; r < D; r++) col[r] = a[i * D + r];
    for (int j = 0; j < D; j++) {
      for (int r = 0; r < D; r++) rho_i8[(i * 64 + r) * 64 + j] = (int8_t)col[r];
      const int64_t last = col[D - 1];
      nxt[0] = -last;
      for (int r = 1; r < D; r++) nxt[r] = col[r - 1];
      nxt[27] -= last;
      memcpy(col, nxt, sizeof(col));
    }
  }
  std::vector<K2> bpow(D);
  bpow[0] = kone();
  for (int i = 1; i < D; i++) bpow[i] = kadd(bpow[i - 1], bpow[i - 1]);
  parent = MeHost{};
  parent.m_in = m_in;
  parent.t = t;
  parent.r = ins[0].r;
  memcpy(parent.fold_digest, ins[0].fold_digest, 32);
  parent.X.assign((size_t)D * m_in, 0);
  parent.y.assign(t * DPAD, kzero());
  parent.y_scalars.assign(t, kzero());
  if (kt == 1) {
    memcpy(parent.c, ins[0].c, sizeof(parent.c));
    parent.X = ins[0].X;
    parent.y = ins[0].y;
  } else {
    std::vector<const uint64_t*> vp(kt);
    std::vector<uint64_t> vbuf(kt * D), outc(D);
    // X: rot(a_i) applied to every column of X_i (a column is a ring element)
    for (size_t c = 0; c < m_in; c++) {
      for (size_t i = 0; i < kt; i++) {
        for (int q = 0; q < D; q++) vbuf[i * D + q] = ins[i].X[q * m_in + c];
        vp[i] = &vbuf[i * D];
      }
      mix_small_coeffs(a.data(), vp.data(), kt, outc.data());
      for (int q = 0; q < D; q++) parent.X[q * m_in + c] = outc[q];
    }
    for (size_t j = 0; j < t; j++)
      for (int comp = 0; comp < 2; comp++) {
        for (size_t i = 0; i < kt; i++) {
          const K2* y = &ins[i].y[j * DPAD];
          for (int v = 0; v < D; v++) vbuf[i * D + v] = comp ? y[v].c1 : y[v].c0;
          vp[i] = &vbuf[i * D];
        }
        mix_small_coeffs(a.data(), vp.data(), kt, outc.data());
        for (int q = 0; q < D; q++) (comp ? parent.y[j * DPAD + q].c1 : parent.y[j * DPAD + q].c0) = outc[q];
      }
    for (int col = 0; col < KAPPA; col++) {
      for (size_t i = 0; i < kt; i++) vp[i] = &ins[i].c[col * D];
      mix_small_coeffs(a.data(), vp.data(), kt, &parent.c[col * D]);
    }
  }
  for (size_t j = 0; j < t; j++) {
    K2 sc = kzero();
    for (int rho = 0; rho < D; rho++) sc = kadd(sc, kmul(parent.y[j * DPAD + rho], bpow[rho]));
    parent.y_scalars[j] = sc;
  }
  rho_dev_.ensure(rho_i8.size());
  zptr_.ensure(kt * sizeof(void*));
  h2d(rho_dev_.p, rho_i8.data(), rho_i8.size(), stream_);
  h2d(zptr_.p, (const void* const*)Z_dev, kt, stream_);
  launch_rlc_zmix(rho_dev_.as<int8_t>(), zptr_.as<const int8_t*>(), kt, m, zmix_dev, stream_);
  NS_CHECK(hipStreamSynchronize(stream_));
}

// children of a parent: balanced base-2 split of Z_mix into planes_dev, child commitments
// (zero for an all-zero plane), child X and y (core rows from v_j = M_j^T chi_r; the rows past
// the CCS's t are bus openings of each plane when bus_chunk > 0), and the public checks.
void Prover::dec_api(const MeHost& parent, const int32_t* zmix_dev, size_t k_dec, int8_t* const* planes_dev,
                     size_t bus_chunk, std::vector<MeHost>& children, bool ok[3]) {
  const Ccs& s = *s_;
  const size_t m = s.m, t = s.t, m_in = parent.m_in;
  if (k_dec == 0 || k_dec > 16) throw Error(1, "k_dec must be 1..16");
  if (parent.t < t) throw Error(1, "parent has fewer y rows than the CCS has matrices");
  BusLayout L;
  L.m = m;
  L.m_in = m_in;
  L.bus_cols = parent.t - t;
  if (L.bus_cols) {
    if (!bus_chunk || L.bus_cols * bus_chunk > m) throw Error(1, "bus rows need a chunk size that fits the witness");
    L.chunk = bus_chunk;
    L.bus_base = m - L.bus_cols * bus_chunk;
  }
  DecPlanes dec_pl{};
  std::vector<const int8_t*> pptr(k_dec);
  for (size_t i = 0; i < k_dec; i++) {
    dec_pl.p[i] = planes_dev[i];
    pptr[i] = planes_dev[i];
  }
  flags_.ensure(4);
  launch_dec_split(zmix_dev, (size_t)D * m, (int)k_dec, dec_pl, flags_.as<int>(), stream_);
  uint32_t dmask = 0;
  d2h(ows_.stage, &dmask, flags_.p, 1, stream_);
  if (dec_out_of_range(dmask, (int)k_dec)) throw Error(6, "DEC split out of range");
  std::vector<uint64_t> child_c(k_dec * KAPPA * D, 0);
  commit_device(pptr.data(), k_dec, child_c.data());
  vjs_at(parent.r, vjs_val_);
  std::vector<K2> yc = y_products(pptr, vjs_val_);
  std::vector<K2> bpow(D);
  bpow[0] = kone();
  for (int i = 1; i < D; i++) bpow[i] = kadd(bpow[i - 1], bpow[i - 1]);
  children.assign(k_dec, MeHost{});
  {
    std::vector<std::vector<uint64_t>*> xs;
    for (auto& c : children) xs.push_back(&c.X);
    head_cols(pptr.data(), k_dec, m_in, xs);
  }
  for (size_t i = 0; i < k_dec; i++) {
    MeHost& e = children[i];
    if (dec_plane_nonzero(dmask, (int)i)) memcpy(e.c, &child_c[i * KAPPA * D], sizeof(e.c));
    e.m_in = m_in;
    e.t = parent.t;
    e.r = parent.r;
    e.y.assign(e.t * DPAD, kzero());
    e.y_scalars.assign(e.t, kzero());
    for (size_t j = 0; j < t; j++) {
      K2 sc = kzero();
      for (int rho = 0; rho < D; rho++) {
        const K2 v = yc[(i * t + j) * D + rho];
        e.y[j * DPAD + rho] = v;
        sc = kadd(sc, kmul(v, bpow[rho]));
      }
      e.y_scalars[j] = sc;
    }
    memcpy(e.fold_digest, parent.fold_digest, 32);
  }
  if (L.bus_cols) {
    bus_weights(parent.r, L);
    std::vector<MeHost*> outs;
    for (auto& c : children) outs.push_back(&c);
    open_bus_planes(pptr, L, outs, t);
  }
  ok[0] = ok[1] = ok[2] = true;
  for (size_t q = 0; q < t * DPAD && ok[0]; q++) {
    K2 lhs = kzero();
    uint64_t pw = 1;
    for (size_t i = 0; i < k_dec; i++) {
      lhs = kadd(lhs, kscale(children[i].y[q], pw));
      pw = fadd(pw, pw);
    }
    ok[0] = keq(lhs, parent.y[q]);
  }
  for (size_t q = 0; q < (size_t)D * m_in; q++) {
    uint64_t lhs = 0, pw = 1;
    for (size_t i = 0; i < k_dec; i++) {
      lhs = fadd(lhs, fmul(children[i].X[q], pw));
      pw = fadd(pw, pw);
    }
    if (lhs != parent.X[q]) ok[1] = false;
  }
  for (int q = 0; q < KAPPA * D; q++) {
    uint64_t lhs = 0, pw = 1;
    for (size_t i = 0; i < k_dec; i++) {
      lhs = fadd(lhs, fmul(children[i].c[q], pw));
      pw = fadd(pw, pw);
    }
    if (lhs != parent.c[q]) ok[2] = false;
  }
}

static void enc_rounds(Enc& pb, const std::vector<K2>& polys, size_t rounds, size_t np) {
  pb.u(rounds);
  for (size_t rd = 0; rd < rounds; rd++) {
    pb.u(np);
    for (size_t i = 0; i < np; i++) pb.k(polys[rd * np + i]);
  }
}

static void enc_lane(Enc& pb, const LaneOut& o, size_t k_dec) {
  pb.u(o.rhos.size() / (D * D));
  for (auto x : o.rhos) pb.u(x);
  pb.me(o.parent);
  pb.u(k_dec);
  for (auto& e : o.children) pb.me(e);
}

std::vector<uint64_t> Prover::prove_shard(const uint64_t* const* z_steps, size_t nsteps, size_t m_in,
                                          bool z_on_device, const StepMemH* mem, const ObCfgH* ob) {
  using clk = std::chrono::steady_clock;
  std::unique_ptr<KTimers> kt;
  if (timing) {
    kt = std::make_unique<KTimers>();
    g_kt = kt.get();
  }
  struct Reset {
    ~Reset() { g_kt = nullptr; }
  } reset_guard;
  const Ccs& s = *s_;
  const size_t m = s.m, t = s.t, k_dec = params_.k_rho;
  const size_t ell_d = 6, ell_n = log2_pow2(s.n), ell = ell_d + ell_n;
  const size_t d_sc = std::max<size_t>(s.f_degree() + 1, std::max<size_t>(2, 2 * params_.b + 2));
  uint32_t smin;
  int32_t slack;
  if (!params_.extension_check((uint32_t)ell, (uint32_t)d_sc, &smin, &slack) || slack < 0)
    throw Error(2, "extension policy failed");
  if (m_in > m) throw Error(1, "m_in exceeds m");
  if (ob) {
    if (nsteps == 0 || !mem) throw Error(1, "output binding requires >= 1 step with memory");
    ob_validate(*ob, mem[nsteps - 1]);
  }
  hipStream_t st = stream_;
  auto t0 = clk::now();
  for (int i = 0; i < NPHASE; i++) phase_s[i] = 0, phase_perm[i] = 0;
  auto tmark = clk::now();
  int cur_phase = -1;
  const Transcript* trp = nullptr;
  uint64_t pmark = 0;
  auto phase_mark = [&](int next) {
    auto now = clk::now();
    const uint64_t pn = trp ? trp->permutations() : 0;
    if (cur_phase >= 0) {
      phase_s[cur_phase] += std::chrono::duration<double>(now - tmark).count();
      phase_perm[cur_phase] += pn - pmark;
    }
    tmark = now;
    pmark = pn;
    cur_phase = next;
  };
  phase_mark(10);  // step intake

  const StepMemH no_mem{};
  BusLayout L;
  if (mem) {
    L = infer_bus_layout(mem, nsteps, s.n, m, m_in);
  } else {
    L.m = m;
    L.m_in = m_in;
  }
  const size_t t_tot = t + L.bus_cols;

  // ---- step intake: digits and MCS commitments for all steps ----
  if (zmcs_.size() < nsteps) zmcs_.resize(nsteps);
  std::vector<DevBuf>& Zmcs = zmcs_;
  std::vector<std::vector<uint64_t>> x_host(nsteps);
  {
    zbuf_.ensure(m * 8);
    DevBuf& zbuf = zbuf_;
    for (size_t i = 0; i < nsteps; i++) {
      Zmcs[i].ensure((size_t)D * m);
      const uint64_t* zsrc = z_steps[i];
      if (!z_on_device) {
        h2d(zbuf.p, z_steps[i], m, st);
        zsrc = zbuf.as<uint64_t>();
        x_host[i].assign(z_steps[i], z_steps[i] + m_in);
      } else {
        x_host[i].resize(m_in);
        if (m_in) d2h(ows_.stage, x_host[i].data(), z_steps[i], m_in, st);
      }
      launch_decomp_b2(zsrc, m, Zmcs[i].as<int8_t>(), st);
      if (!z_on_device) NS_CHECK(hipStreamSynchronize(st));
    }
    NS_CHECK(hipStreamSynchronize(st));
  }
  if (mem) check_bus_guardrails(s, L, x_host);
  std::vector<uint64_t> cmcs(nsteps * KAPPA * D);
  {
    std::vector<const int8_t*> ptrs(nsteps);
    for (size_t i = 0; i < nsteps; i++) ptrs[i] = Zmcs[i].as<int8_t>();
    commit_device(ptrs.data(), nsteps, cmcs.data());
  }

  Enc pb;
  pb.u(0x314652504f52504eULL);
  pb.u(nsteps);
  Transcript tr("neo.fold/session");
  trp = &tr;
  pmark = tr.permutations();
  DigestCache digests;

  std::vector<MeHost> acc, val_obl;
  std::vector<const int8_t*> acc_wit;  // digit planes of the accumulator (planes_[cur_planes])
  int cur_planes = 0;
  DevBuf& vjs = vjs_;
  DevBuf& ws_yd = ws_yd_;
  const size_t xs_n = d_sc + 1;
  std::vector<K2> xs(xs_n);
  for (size_t i = 0; i < xs_n; i++) xs[i] = K2{i, 0};
  std::vector<K2> bpow(D);
  bpow[0] = kone();
  for (int i = 1; i < D; i++) bpow[i] = kadd(bpow[i - 1], bpow[i - 1]);
  std::vector<TwistCols> prev_tw;
  bool have_prev = false;
  std::vector<uint64_t> zbus;
  std::vector<std::vector<K2>> ob_rounds;  // OutputSumcheckProof.round_polys
  std::vector<K2> ob_r_prime;

  for (size_t idx = 0; idx < nsteps; idx++) {
    const StepMemH& sm = mem ? mem[idx] : no_mem;
    const size_t nacc = acc.size(), kt = nacc + 1;
    absorb_step_memory(tr, sm, digests);
    // output binding on the last step, before r_time is sampled (shard.rs:1803-1862)
    const bool include_ob = ob && idx + 1 == nsteps;
    if (include_ob) {
      tr.append_msg_u64("shard/output_binding_start", idx);
      const uint64_t mi = ob->mem_idx, nb = ob->num_bits;
      tr.append_u64s("output_binding/mem_idx", &mi, 1);
      tr.append_u64s("output_binding/num_bits", &nb, 1);
      ob_prove(tr, *ob, ob_rounds, ob_r_prime);
    }
    phase_mark(0);
    bind_step_header(tr, s, params_, ell, d_sc, slack, x_host[idx].data(), m_in, &cmcs[idx * KAPPA * D], acc);
    const StepChallenges ch = sample_challenges(tr, ell_d, ell);
    const std::vector<K2>& alpha = ch.alpha;
    const K2* beta_a = ch.beta.data();
    const K2* beta_r = ch.beta.data() + ell_d;
    const K2 gamma = ch.gamma;
    const K2 T = claimed_initial_sum(s, ch, acc, ell_d);
    tr.append_k("sumcheck/initial_sum", T);
    std::vector<K2> r_cycle(ell_n);
    for (size_t i = 0; i < ell_n; i++) {
      tr.append_msg_u64("route_a/r_cycle", i);
      r_cycle[i] = tr.challenge_k("route_a/cycle/0", "route_a/cycle/1");
    }
    phase_mark(1);
    // device oracle
    std::vector<const int8_t*> Zs(kt);
    Zs[0] = Zmcs[idx].as<int8_t>();
    for (size_t a = 0; a < nacc; a++) Zs[1 + a] = acc_wit[a];
    CcsOracle oracle(s, Zs.data(), kt, alpha.data(), beta_a, beta_r, gamma, ell_d, ell_n,
                     nacc ? acc[0].r.data() : nullptr, st, &ows_);
    // Route A memory: the bus region decoded from the step's digit planes, then the address-lane
    // sum-checks (Shout first) and the time-lane oracles (memory.rs:232-1025)
    ShoutPre sp;
    std::vector<TwistPre> tp;
    std::vector<std::unique_ptr<TsOracleGpu>> mem_owned;
    std::vector<RaClaim> claims(1);
    claims[0].ccs = &oracle;
    claims[0].deg = d_sc;
    claims[0].label = "ccs/time";
    claims[0].dyn = true;
    claims[0].sum = T;
    if (sm.n_lut || sm.n_mem) {
      const size_t blen = L.bus_cols * L.chunk;
      zbus.resize(blen);
      busz_.ensure(blen * 8);
      launch_bus_decode_range(Zmcs[idx].as<int8_t>(), m, L.bus_base, blen, busz_.as<uint64_t>(), st);
      d2h(ows_.stage, zbus.data(), busz_.p, blen, st);
      shout_addr_pre(tr, sm, L, zbus.data(), ell_n, r_cycle, idx, sp);
      twist_addr_pre(tr, sm, L, zbus.data(), ell_n, r_cycle, tp);
      build_time_claims(sm, ell_n, r_cycle, sp, tp, mem_owned, claims);
    }
    if (include_ob) {
      // TwistTotalIncOracleSparseTime at r', the last batched time claim (shard.rs:1990-2004)
      const TwistCols& T = tp[ob->mem_idx].cols;
      mem_owned.push_back(std::make_unique<TsOracleGpu>());
      TsOracleGpu* o = mem_owned.back().get();
      o->init_time(TSK_TOTAL_INC, {}, ell_n, &T.has_write, nullptr, &T.inc, T.wa, ob_r_prime, {}, kzero(), nullptr,
                   nullptr, nullptr);
      RaClaim c;
      c.ts = o;
      c.label = "output_binding/inc_total";
      c.dyn = true;
      c.sum = o->claim();
      c.deg = o->degree_bound();
      claims.push_back(c);
    }
    phase_mark(2);
    // the batched time sum-check: the CCS claim and the memory time lanes (route_a_time.rs:28-152)
    bind_dynamic_claims(tr, claims);
    std::vector<K2> r_time;
    run_batched(tr, "shard/batched_time", idx, claims, ell_n, r_time);
    K2 running = claims[0].running;
    const std::vector<K2>& time_coeffs = claims[0].polys;
    for (auto& c : claims) c.ts = nullptr;
    mem_owned.clear();

    phase_mark(3);
    // precompute_for_r on the device
    std::vector<K2> y_eval = oracle.y_eval(r_time, ws_yd, vjs);
    phase_mark(4);
    // Ajtai-bit rounds (host)
    AjtaiPhase aj;
    aj.init(s, y_eval, kt, alpha.data(), beta_a, beta_r, gamma, r_time, nacc ? acc[0].r.data() : nullptr);
    tr.append_message("sc/domain", "ccs/ajtai");
    tr.append_msg_u64("sc/inst_idx", idx);
    tr.append_k("sc/claimed_sum", running);
    std::vector<K2> aj_coeffs(ell_d * xs_n);
    for (size_t rd = 0; rd < ell_d; rd++) {
      const std::vector<K2> c = aj.coeffs();
      K2* co = &aj_coeffs[rd * xs_n];
      for (size_t i = 0; i < xs_n; i++) co[i] = i < c.size() ? c[i] : kzero();
      if (!keq(kadd(poly_eval(co, xs_n, xs[0]), poly_eval(co, xs_n, xs[1])), running))
        throw Error(4, "step " + std::to_string(idx) + ": ajtai round " + std::to_string(rd) + " invariant failed");
      for (size_t i = 0; i < xs_n; i++) tr.append_k("sumcheck/round/coeff", co[i]);
      K2 r = tr.challenge_k("sumcheck/challenge/0", "sumcheck/challenge/1");
      running = poly_eval(co, xs_n, r);
      aj.fold(r);
    }
    const std::vector<K2>& aj_chals = aj.chals;
    tr.append_message("sc/domain_end", "ccs/ajtai");
    tr.append_msg_u64("sc/inst_idx_end", idx);
    uint8_t fold_digest[32];
    tr.digest32(fold_digest);

    phase_mark(5);
    // ME outputs, with the shared-bus openings appended (shard.rs:2082-2120)
    std::vector<MeHost> ccs_out(kt);
    for (size_t i = 0; i < kt; i++) {
      MeHost& e = ccs_out[i];
      e.m_in = m_in;
      e.t = t_tot;
      if (i == 0) {
        memcpy(e.c, &cmcs[idx * KAPPA * D], sizeof(e.c));
        const int8_t* pl0 = Zmcs[idx].as<int8_t>();
        head_cols(&pl0, 1, m_in, {&e.X});
      } else {
        memcpy(e.c, acc[i - 1].c, sizeof(e.c));
        e.X = acc[i - 1].X;
      }
      e.r = r_time;
      e.y.assign(t_tot * DPAD, kzero());
      e.y_scalars.assign(t_tot, kzero());
      for (size_t j = 0; j < t; j++) {
        K2 sc = kzero();
        for (int rho = 0; rho < D; rho++) {
          K2 v = y_eval[(i * t + j) * D + rho];
          e.y[j * DPAD + rho] = v;
          sc = kadd(sc, kmul(v, bpow[rho]));
        }
        e.y_scalars[j] = sc;
      }
      memcpy(e.fold_digest, fold_digest, 32);
    }
    if (L.bus_cols) {
      bus_weights(r_time, L);
      std::vector<MeHost*> outs;
      for (auto& e : ccs_out) outs.push_back(&e);
      open_bus_planes(Zs, L, outs, t);
    }
    // finalize_route_a_memory_prover: val-eval batch, CPU ME claims at r_val (memory.rs:1284-1670)
    ValPre vp;
    std::vector<MeHost> vclaims;
    std::vector<const int8_t*> vZ;
    if (sm.n_mem) {
      val_eval_batch(tr, sm, tp, (idx > 0 && have_prev) ? &prev_tw : nullptr, ell_n, r_time, idx, vp);
      Transcript fork = tr;
      fork.append_message("fork", "cpu_bus/me_digest_val");
      uint8_t vdig[32];
      fork.digest32(vdig);
      const size_t nvc = idx > 0 ? 2 : 1;
      vjs_at(vp.r_val, vjs_val_);
      for (size_t q = 0; q < nvc; q++) vZ.push_back(Zmcs[idx - q].as<int8_t>());
      std::vector<K2> yv = y_products(vZ, vjs_val_);
      vclaims.resize(nvc);
      for (size_t q = 0; q < nvc; q++) {
        MeHost& e = vclaims[q];
        const size_t stp = idx - q;
        memcpy(e.c, &cmcs[stp * KAPPA * D], sizeof(e.c));
        e.m_in = m_in;
        e.t = t_tot;
        head_cols(&vZ[q], 1, m_in, {&e.X});
        e.r = vp.r_val;
        e.y.assign(t_tot * DPAD, kzero());
        e.y_scalars.assign(t_tot, kzero());
        for (size_t j = 0; j < t; j++) {
          K2 sc = kzero();
          for (int rho = 0; rho < D; rho++) {
            K2 v = yv[(q * t + j) * D + rho];
            e.y[j * DPAD + rho] = v;
            sc = kadd(sc, kmul(v, bpow[rho]));
          }
          e.y_scalars[j] = sc;
        }
        memcpy(e.fold_digest, vdig, 32);
      }
      if (L.bus_cols) {
        bus_weights(vp.r_val, L);
        std::vector<MeHost*> outs;
        for (auto& e : vclaims) outs.push_back(&e);
        open_bus_planes(vZ, L, outs, t);
      }
    }

    phase_mark(6);
    // ---- main lane, then the val lane (shard.rs:2171-2254) ----
    const int nxt_planes = cur_planes ^ 1;
    LaneOut mainl, vall;
    rlc_dec_lane(tr, "main", idx, ccs_out, Zs, planes_[nxt_planes], vjs, L, mainl);
    phase_mark(8);
    if (!vclaims.empty()) {
      tr.append_msg_u64("fold/val_lane_start", idx);
      rlc_dec_lane(tr, "val", idx, vclaims, vZ, valplanes_, vjs_val_, L, vall);
    }
    phase_mark(9);
    tr.append_msg_u64("fold/step_done", idx);

    // ---- encode StepProof ----
    pb.u(kt);
    for (auto& e : ccs_out) pb.me(e);
    pb.u(ell_n + ell_d);
    for (size_t rd = 0; rd < ell_n; rd++) {
      pb.u(xs_n);
      for (size_t i = 0; i < xs_n; i++) pb.k(time_coeffs[rd * xs_n + i]);
    }
    for (size_t rd = 0; rd < ell_d; rd++) {
      pb.u(xs_n);
      for (size_t i = 0; i < xs_n; i++) pb.k(aj_coeffs[rd * xs_n + i]);
    }
    pb.k(T);
    pb.u(ell_n + ell_d);
    for (auto& x : r_time) pb.k(x);
    for (auto& x : aj_chals) pb.k(x);
    pb.u(ell_d);
    for (auto& x : alpha) pb.k(x);
    pb.u(ell_d);
    for (size_t i = 0; i < ell_d; i++) pb.k(beta_a[i]);
    pb.u(ell_n);
    for (size_t i = 0; i < ell_n; i++) pb.k(beta_r[i]);
    pb.k(gamma);
    pb.k(running);
    pb.digest(fold_digest);
    enc_lane(pb, mainl, k_dec);
    // mem: MemSidecarProof { cpu_me_claims_val, shout_addr_pre, proofs }
    pb.u(vclaims.size());
    for (auto& e : vclaims) pb.me(e);
    pb.u(sm.n_lut);
    for (auto& x : sp.claimed) pb.k(x);
    pb.u(sp.active_mask);
    pb.u(sp.rounds.size());
    for (size_t i = 0; i < sp.rounds.size(); i++) enc_rounds(pb, sp.rounds[i], sp.ell_addr, sp.degs[i] + 1);
    pb.u(sp.r_addr.size());
    for (auto& x : sp.r_addr) pb.k(x);
    pb.u(sm.n_lut + sm.n_mem);
    for (size_t i = 0; i < sm.n_lut; i++) {
      pb.u(1);  // MemOrLutProof::Shout(ShoutProof::default())
      pb.u(0);
      pb.u(0);
      pb.u(0);
    }
    for (size_t i = 0; i < sm.n_mem; i++) {
      const TwistPre& P = tp[i];
      pb.u(0);  // MemOrLutProof::Twist
      pb.u(2);
      pb.k(kzero());
      pb.k(kzero());
      pb.u(2);
      for (int c = 0; c < 2; c++) enc_rounds(pb, P.rounds[c], P.ell_addr, P.degs[c] + 1);
      pb.u(P.r_addr.size());
      for (auto& x : P.r_addr) pb.k(x);
      const RaClaim* vc = &vp.claims[vp.per * i];
      pb.u(1);
      pb.k(vc[0].sum);
      enc_rounds(pb, vc[0].polys, ell_n, vc[0].deg + 1);
      pb.k(vc[1].sum);
      enc_rounds(pb, vc[1].polys, ell_n, vc[1].deg + 1);
      if (vp.per == 3) {
        pb.u(1);
        pb.k(vc[2].sum);
        pb.u(1);
        enc_rounds(pb, vc[2].polys, ell_n, vc[2].deg + 1);
      } else {
        pb.u(0);
        pb.u(0);
      }
    }
    // batched_time: claimed_sums, degree_bounds, labels, round_polys
    pb.u(claims.size());
    for (auto& c : claims) pb.k(c.sum);
    pb.u(claims.size());
    for (auto& c : claims) pb.u(c.deg);
    pb.u(claims.size());
    for (auto& c : claims) {
      const size_t ln = strlen(c.label);
      pb.u(ln);
      for (size_t q = 0; q < ln; q++) pb.u((uint8_t)c.label[q]);
    }
    pb.u(claims.size());
    for (auto& c : claims) enc_rounds(pb, c.polys, ell_n, c.deg + 1);
    if (!vclaims.empty()) {
      pb.u(1);
      enc_lane(pb, vall, k_dec);
      for (auto& e : vall.children) val_obl.push_back(e);
    } else {
      pb.u(0);
    }

    acc = std::move(mainl.children);
    acc_wit = mainl.planes;
    cur_planes = nxt_planes;
    prev_tw.clear();
    for (auto& P : tp) prev_tw.push_back(std::move(P.cols));
    have_prev = sm.n_mem > 0;
    phase_mark(11);
  }
  phase_mark(-1);
  // ShardProof.output_proof (shard_proof_types.rs:165-171)
  if (ob) {
    pb.u(1);
    pb.u(ob_rounds.size());
    for (auto& c : ob_rounds) {
      pb.u(c.size());
      for (auto& x : c) pb.k(x);
    }
  } else {
    pb.u(0);
  }
  pb.u(acc.size());
  for (auto& e : acc) pb.me(e);
  pb.u(val_obl.size());
  for (auto& e : val_obl) pb.me(e);
  NS_CHECK(hipStreamSynchronize(st));
  last_seconds = std::chrono::duration<double>(clk::now() - t0).count();
  last_permutations = tr.permutations();
  if (kt) kt->collect(kernel_ms, kernel_count);
  return pb.v;
}

}  // namespace ns
