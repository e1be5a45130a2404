// capi.cpp -- extern "C" entry points declared in include/nightstream.h.
#include <string.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/nightstream.h"
#include "prover.h"
#include "twist.h"
#include "route_a.h"

using namespace ns;

struct ns_ccs {
  std::unique_ptr<Ccs> s;
};
struct ns_pp {
  std::unique_ptr<Pp> pp;
  // commit-call buffers kept across calls (hipMalloc / hipFree of the GEMM's split planes cost
  // more than the GEMM itself at small batches); one call at a time per handle
  std::mutex mu;
  DevBuf ptrs, outd, ws;
};
struct ns_prover {
  std::unique_ptr<Prover> p;
};
struct ns_oracle {
  const Ccs* s;
  hipStream_t st;
  std::vector<DevBuf> zbufs;
  std::unique_ptr<CcsOracle> o;
  size_t ell_d, ell_n, d_sc, k;
  std::vector<K2> alpha, beta_a, beta_r, r_inputs;
  K2 gamma;
  std::vector<K2> row_chals;
  bool ajtai_ready = false;
  AjtaiPhase aj;
  DevBuf ws, vjs;
};

static thread_local std::string g_err;

template <class F>
static int guard(F&& f) {
  try {
    f();
    return NS_OK;
  } catch (const Error& e) {
    g_err = e.what();
    switch (e.code) {
      case 2: return NS_EPARAMS;
      case 4: return NS_ESUMCHECK;
      case 6: return NS_ERANGE;
      case 5:
      case 7: return NS_EPROTOCOL;
      default: return NS_EINVAL;
    }
  } catch (const HipError& e) {
    g_err = std::string("HIP error ") + hipGetErrorString(e.err) + " at " + e.what + " (" + e.file + ":" +
            std::to_string(e.line) + ")";
    return e.err == hipErrorNoDevice || e.err == hipErrorInvalidDevice ? NS_ENODEV : NS_EHIP;
  } catch (const std::exception& e) {
    g_err = e.what();
    return NS_EINVAL;
  }
}

static void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw HipError(hipErrorNoDevice, "hipGetDeviceCount", __FILE__, __LINE__);
}

static hipStream_t default_stream() {
  static hipStream_t s = [] {
    hipStream_t x;
    NS_CHECK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    return x;
  }();
  return s;
}

extern "C" {

const char* ns_last_error(void) { return g_err.c_str(); }

int ns_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void ns_free(void* p) { free(p); }

int ns_params_goldilocks_auto_r1cs_ccs(size_t n_rows, uint64_t out[11]) {
  return guard([&] {
    Params p = Params::goldilocks_auto_r1cs_ccs(n_rows);
    uint64_t v[11] = {p.q, p.eta, p.d, p.kappa, p.m, p.b, p.k_rho, p.B, p.T, p.s, p.lambda};
    memcpy(out, v, sizeof(v));
  });
}

int ns_ccs_from_r1cs(size_t n, size_t m, const uint64_t* a_ptr, const uint32_t* a_col, const uint64_t* a_val,
                     const uint64_t* b_ptr, const uint32_t* b_col, const uint64_t* b_val, const uint64_t* c_ptr,
                     const uint32_t* c_col, const uint64_t* c_val, ns_ccs** out) {
  return guard([&] {
    auto mk = [&](const uint64_t* p, const uint32_t* c, const uint64_t* v) {
      HostCsr h;
      h.ptr.assign(p, p + n + 1);
      size_t nnz = p[n];
      h.col.assign(c, c + nnz);
      h.val.assign(v, v + nnz);
      for (auto& x : h.val)
        if (x >= Q) throw Error(1, "matrix value not canonical");
      return h;
    };
    auto s = Ccs::from_r1cs(n, m, mk(a_ptr, a_col, a_val), mk(b_ptr, b_col, b_val), mk(c_ptr, c_col, c_val));
    *out = new ns_ccs{std::move(s)};
  });
}

int ns_ccs_digest(const ns_ccs* s, uint64_t out[4]) {
  return guard([&] { memcpy(out, s->s->digest, 32); });
}

void ns_ccs_free(ns_ccs* s) { delete s; }

int ns_pp_setup_seeded(const uint8_t seed[32], size_t m, int keep_coeffs, ns_pp** out) {
  return guard([&] {
    require_device();
    auto pp = Pp::setup_seeded(seed, m, keep_coeffs != 0, default_stream());
    *out = new ns_pp{std::move(pp)};
  });
}

int ns_pp_from_coeffs(const uint64_t* coeffs, size_t m, int keep_coeffs, ns_pp** out) {
  return guard([&] {
    require_device();
    auto pp = Pp::from_coeffs(coeffs, m, keep_coeffs != 0, default_stream());
    *out = new ns_pp{std::move(pp)};
  });
}

int ns_pp_coeffs(const ns_pp* pp, uint64_t* out) {
  return guard([&] {
    if (!pp->pp->coeffs.p) throw Error(1, "PP coefficients were not kept");
    NS_CHECK(hipMemcpy(out, pp->pp->coeffs.p, (size_t)KAPPA * pp->pp->m * D * 8, hipMemcpyDeviceToHost));
  });
}

void ns_pp_free(ns_pp* pp) { delete pp; }

int ns_decomp_b(const uint64_t* z, size_t m, uint32_t b, int8_t* Z_out) {
  return guard([&] {
    require_device();
    if (b != 2) throw Error(1, "only b = 2 is implemented");
    hipStream_t s = default_stream();
    DevBuf zd(m * 8), Zd((size_t)D * m);
    NS_CHECK(hipMemcpyAsync(zd.p, z, m * 8, hipMemcpyHostToDevice, s));
    launch_decomp_b2(zd.as<uint64_t>(), m, Zd.as<int8_t>(), s);
    NS_CHECK(hipMemcpyAsync(Z_out, Zd.p, (size_t)D * m, hipMemcpyDeviceToHost, s));
    NS_CHECK(hipStreamSynchronize(s));
  });
}

// the batched commitment of n_wit device digit matrices with the handle's cached buffers
static void commit_on(ns_pp* pp, const int8_t* const* Z_dev, size_t n_wit, uint64_t* out, hipStream_t s) {
  const size_t m = pp->pp->m;
  std::lock_guard<std::mutex> g(pp->mu);
  pp->ptrs.ensure(n_wit * sizeof(void*));
  pp->outd.ensure(n_wit * KAPPA * D * 8);
  pp->ws.ensure(ajtai_workspace_bytes(m, n_wit));
  NS_CHECK(hipMemcpyAsync(pp->ptrs.p, Z_dev, n_wit * sizeof(void*), hipMemcpyHostToDevice, s));
  ajtai_commit(pp->pp->limbs.as<int8_t>(), KAPPA, m, pp->ptrs.as<const int8_t*>(), n_wit, pp->outd.as<uint64_t>(),
               pp->ws.p, s);
  NS_CHECK(hipMemcpyAsync(out, pp->outd.p, n_wit * KAPPA * D * 8, hipMemcpyDeviceToHost, s));
  NS_CHECK(hipStreamSynchronize(s));
}

static std::vector<DevBuf> upload_digits(const int8_t* const* Z, size_t n, size_t m, hipStream_t s) {
  std::vector<DevBuf> out;
  for (size_t i = 0; i < n; i++) {
    out.emplace_back((size_t)D * m);
    NS_CHECK(hipMemcpyAsync(out.back().p, Z[i], (size_t)D * m, hipMemcpyHostToDevice, s));
  }
  return out;
}

int ns_ajtai_commit_batch(const ns_pp* pp, const int8_t* const* Z, size_t n_wit, uint64_t* out) {
  return guard([&] {
    require_device();
    if (!pp) throw Error(1, "null PP");
    if (n_wit && (!Z || !out)) throw Error(1, "null witness array or output");
    const size_t m = pp->pp->m;
    // The GEMM's int32 K-split partial sums are exact for digits in {-1, 0, 1} (at most
    // 128 * m <= 2^27 per split for m <= 2^20); wider digits are refused, not wrapped.
    for (size_t w = 0; w < n_wit; w++) {
      if (!Z[w]) throw Error(1, "null witness");
      const uint8_t* z = (const uint8_t*)Z[w];
      uint8_t bad = 0;
      for (size_t i = 0; i < (size_t)D * m; i++) bad |= (uint8_t)((uint8_t)(z[i] + 1) > 2);
      if (bad) throw Error(1, "witness " + std::to_string(w) + " has a digit outside {-1, 0, 1}");
    }
    hipStream_t s = default_stream();
    auto bufs = upload_digits(Z, n_wit, m, s);
    std::vector<const int8_t*> ptrs;
    for (auto& b : bufs) ptrs.push_back(b.as<int8_t>());
    commit_on(const_cast<ns_pp*>(pp), ptrs.data(), n_wit, out, s);
  });
}

int ns_ajtai_commit_batch_device(const ns_pp* pp, const int8_t* const* Z_dev, size_t n_wit, uint64_t* out) {
  return guard([&] {
    require_device();
    if (!pp) throw Error(1, "null PP");
    if (n_wit && (!Z_dev || !out)) throw Error(1, "null witness array or output");
    for (size_t i = 0; i < n_wit; i++)
      if (!Z_dev[i]) throw Error(1, "null device witness");
    commit_on(const_cast<ns_pp*>(pp), Z_dev, n_wit, out, default_stream());
  });
}

int ns_ajtai_force_variant(int variant) {
  return guard([&] { ajtai_force_variant(variant); });
}

int ns_set_shared_rounds(int on) {
  return guard([&] { set_shared_rounds(on != 0); });
}

int ns_shared_round_stats(uint64_t* launches, uint64_t* rounds) {
  return guard([&] {
    if (!launches || !rounds) throw Error(1, "null output");
    const FusedHubStats st = fused_hub_stats();
    *launches = st.launches;
    *rounds = st.jobs;
  });
}

static std::vector<K2> kvec(const uint64_t* p, size_t n) {
  std::vector<K2> v(n);
  for (size_t i = 0; i < n; i++) v[i] = K2{p[2 * i], p[2 * i + 1]};
  return v;
}

int ns_ccs_oracle_new(const ns_ccs* s, const int8_t* const* Z, size_t k, const uint64_t* alpha,
                      const uint64_t* beta_a, const uint64_t* beta_r, const uint64_t gamma[2], size_t ell_d,
                      size_t ell_n, const uint64_t* r_inputs, ns_oracle** out) {
  return guard([&] {
    require_device();
    if (ell_d != 6) throw Error(1, "ell_d must be 6 (D = 54)");
    if (k == 0) throw Error(1, "need at least one witness");
    auto o = std::make_unique<ns_oracle>();
    o->s = s->s.get();
    o->st = default_stream();
    o->zbufs = upload_digits(Z, k, s->s->m, o->st);
    std::vector<const int8_t*> ptrs;
    for (auto& b : o->zbufs) ptrs.push_back(b.as<int8_t>());
    o->alpha = kvec(alpha, ell_d);
    o->beta_a = kvec(beta_a, ell_d);
    o->beta_r = kvec(beta_r, ell_n);
    if (r_inputs) o->r_inputs = kvec(r_inputs, ell_n);
    o->gamma = K2{gamma[0], gamma[1]};
    o->ell_d = ell_d;
    o->ell_n = ell_n;
    o->k = k;
    o->d_sc = std::max<size_t>(s->s->f_degree() + 1, 6);
    o->o = std::make_unique<CcsOracle>(*s->s, ptrs.data(), k, o->alpha.data(), o->beta_a.data(), o->beta_r.data(),
                                       o->gamma, ell_d, ell_n, r_inputs ? o->r_inputs.data() : nullptr, o->st);
    *out = o.release();
  });
}

int ns_oracle_evals(ns_oracle* o, const uint64_t* xs, size_t n, uint64_t* out) {
  return guard([&] {
    std::vector<K2> x = kvec(xs, n), y(n);
    if (o->row_chals.size() < o->ell_n) {
      std::vector<K2> c = o->o->round_coeffs();
      for (size_t i = 0; i < n; i++) {
        K2 r = c.back();
        for (size_t d = c.size() - 1; d-- > 0;) r = kadd(kmul(r, x[i]), c[d]);
        y[i] = r;
      }
    } else {
      if (!o->ajtai_ready) {
        auto ye = o->o->y_eval(o->row_chals, o->ws, o->vjs);
        o->aj.init(*o->s, std::move(ye), o->k, o->alpha.data(), o->beta_a.data(), o->beta_r.data(), o->gamma,
                   o->row_chals, o->r_inputs.empty() ? nullptr : o->r_inputs.data());
        o->ajtai_ready = true;
      }
      if (o->aj.chals.size() >= o->ell_d) throw Error(1, "evals after the last round");
      y = o->aj.evals(x);
    }
    for (size_t i = 0; i < n; i++) {
      out[2 * i] = y[i].c0;
      out[2 * i + 1] = y[i].c1;
    }
  });
}

int ns_oracle_fold(ns_oracle* o, const uint64_t r[2]) {
  return guard([&] {
    K2 x{r[0], r[1]};
    if (o->row_chals.size() < o->ell_n) {
      o->row_chals.push_back(x);
      o->o->fold(x);
    } else {
      if (!o->ajtai_ready) {
        auto ye = o->o->y_eval(o->row_chals, o->ws, o->vjs);
        o->aj.init(*o->s, std::move(ye), o->k, o->alpha.data(), o->beta_a.data(), o->beta_r.data(), o->gamma,
                   o->row_chals, o->r_inputs.empty() ? nullptr : o->r_inputs.data());
        o->ajtai_ready = true;
      }
      o->aj.fold(x);
    }
  });
}

size_t ns_oracle_num_rounds(const ns_oracle* o) { return o->ell_n + o->ell_d; }
size_t ns_oracle_degree_bound(const ns_oracle* o) { return o->d_sc; }

int ns_oracle_y_eval(ns_oracle* o, const uint64_t* r, uint64_t* out) {
  return guard([&] {
    std::vector<K2> rr = kvec(r, o->ell_n);
    auto y = o->o->y_eval(rr, o->ws, o->vjs);
    memcpy(out, y.data(), y.size() * sizeof(K2));
  });
}

void ns_oracle_free(ns_oracle* o) { delete o; }

int ns_rlc_zmix(const int8_t* rho_coeffs, const int8_t* const* Z, size_t k, size_t m, int32_t* zmix_out) {
  return guard([&] {
    require_device();
    hipStream_t s = default_stream();
    // rot(a_i) as int8 [k][64][64] (row r, column j; zero beyond 54)
    std::vector<int8_t> rho(k * 64 * 64, 0);
    for (size_t i = 0; i < k; i++) {
      int64_t col[D], nxt[D];
      for (int r = 0; r < D; r++) {
        col[r] = rho_coeffs[i * D + r];
        if (col[r] < -2 || col[r] > 2) throw Error(1, "rho coefficient outside the alphabet [-2, 2]");
      }
      for (int j = 0; j < D; j++) {
        for (int r = 0; r < D; r++) rho[(i * 64 + r) * 64 + j] = (int8_t)col[r];
        int64_t last = col[D - 1];
        nxt[0] = -last;
        for (int r = 1; r < D; r++) nxt[r] = col[r - 1];
        nxt[27] -= last;
        memcpy(col, nxt, sizeof(col));
      }
    }
    auto bufs = upload_digits(Z, k, m, s);
    std::vector<const int8_t*> ptrs;
    for (auto& b : bufs) ptrs.push_back(b.as<int8_t>());
    DevBuf pd(k * sizeof(void*)), rd(rho.size()), zd((size_t)D * m * 4);
    NS_CHECK(hipMemcpyAsync(pd.p, ptrs.data(), k * sizeof(void*), hipMemcpyHostToDevice, s));
    NS_CHECK(hipMemcpyAsync(rd.p, rho.data(), rho.size(), hipMemcpyHostToDevice, s));
    launch_rlc_zmix(rd.as<int8_t>(), pd.as<const int8_t*>(), k, m, zd.as<int32_t>(), s);
    NS_CHECK(hipMemcpyAsync(zmix_out, zd.p, (size_t)D * m * 4, hipMemcpyDeviceToHost, s));
    NS_CHECK(hipStreamSynchronize(s));
  });
}

int ns_dec_split(const int32_t* zmix, size_t m, size_t k_dec, int8_t* const* outs, int* nonzero) {
  return guard([&] {
    require_device();
    if (k_dec == 0 || k_dec > (size_t)DEC_MAX_PLANES) throw Error(1, "k_dec out of range");
    hipStream_t s = default_stream();
    const size_t len = (size_t)D * m;
    DevBuf zd(len * 4), flags(4);
    std::vector<DevBuf> planes;
    DecPlanes pl{};
    for (size_t i = 0; i < k_dec; i++) {
      planes.emplace_back(len);
      pl.p[i] = planes.back().as<int8_t>();
    }
    NS_CHECK(hipMemcpyAsync(zd.p, zmix, len * 4, hipMemcpyHostToDevice, s));
    launch_dec_split(zd.as<int32_t>(), len, (int)k_dec, pl, flags.as<int>(), s);
    uint32_t mask = 0;
    NS_CHECK(hipMemcpyAsync(&mask, flags.p, 4, hipMemcpyDeviceToHost, s));
    NS_CHECK(hipStreamSynchronize(s));
    if (dec_out_of_range(mask, (int)k_dec)) throw Error(6, "DEC split: value out of range for k_rho digits");
    for (size_t i = 0; i < k_dec; i++) {
      NS_CHECK(hipMemcpy(outs[i], planes[i].p, len, hipMemcpyDeviceToHost));
      nonzero[i] = dec_plane_nonzero(mask, (int)i) ? 1 : 0;
    }
  });
}

int ns_dec_split_device_noalloc(const int32_t* zmix_dev, size_t m, size_t k_dec, int8_t* const* outs_dev,
                                int* nonzero) {
  return guard([&] {
    require_device();
    if (k_dec == 0 || k_dec > (size_t)DEC_MAX_PLANES) throw Error(1, "k_dec out of range (1..30)");
    if (!nonzero || (m && !zmix_dev) || !outs_dev) throw Error(1, "null argument");
    for (size_t i = 0; i < k_dec; i++)
      if (!outs_dev[i]) throw Error(1, "null output plane");
    hipStream_t s = default_stream();
    int* flags = dec_noalloc_flags();  // this device's copy of the static scratch (no allocation)
    DecPlanes pl{};
    for (size_t i = 0; i < k_dec; i++) pl.p[i] = outs_dev[i];
    launch_dec_split(zmix_dev, (size_t)D * m, (int)k_dec, pl, flags, s);
    uint32_t mask = 0;
    NS_CHECK(hipMemcpyAsync(&mask, flags, sizeof(mask), hipMemcpyDeviceToHost, s));
    NS_CHECK(hipStreamSynchronize(s));
    if (dec_out_of_range(mask, (int)k_dec)) throw Error(6, "DEC split: value out of range for k_rho digits");
    for (size_t i = 0; i < k_dec; i++) nonzero[i] = dec_plane_nonzero(mask, (int)i) ? 1 : 0;
  });
}

int ns_prover_new(const ns_ccs* s, const ns_pp* pp, ns_prover** out) {
  return guard([&] {
    require_device();
    auto p = std::make_unique<Prover>(*s->s, *pp->pp);
    *out = new ns_prover{std::move(p)};
  });
}

int ns_prover_rebind(ns_prover* p, const ns_ccs* s, const ns_pp* pp) {
  return guard([&] {
    if (!p || !s || !pp) throw Error(1, "null handle");
    p->p->rebind(*s->s, *pp->pp);
  });
}

int ns_prove_shard(ns_prover* p, const uint64_t* const* z_steps, size_t n_steps, size_t m_in, uint8_t** proof,
                   size_t* proof_len, double* seconds) {
  return guard([&] {
    if (n_steps == 0) throw Error(1, "need at least one step");
    std::vector<uint64_t> enc = p->p->prove_shard(z_steps, n_steps, m_in);
    uint8_t* buf = (uint8_t*)malloc(enc.size() * 8);
    if (!buf) throw Error(1, "out of host memory");
    memcpy(buf, enc.data(), enc.size() * 8);
    *proof = buf;
    *proof_len = enc.size() * 8;
    if (seconds) *seconds = p->p->last_seconds;
  });
}

int ns_prove_shard_device(ns_prover* p, const uint64_t* const* z_steps_dev, size_t n_steps, size_t m_in,
                          uint8_t** proof, size_t* proof_len, double* seconds) {
  return guard([&] {
    if (n_steps == 0) throw Error(1, "need at least one step");
    std::vector<uint64_t> enc = p->p->prove_shard(z_steps_dev, n_steps, m_in, true);
    uint8_t* buf = (uint8_t*)malloc(enc.size() * 8);
    if (!buf) throw Error(1, "out of host memory");
    memcpy(buf, enc.data(), enc.size() * 8);
    *proof = buf;
    *proof_len = enc.size() * 8;
    if (seconds) *seconds = p->p->last_seconds;
  });
}

// the per-step instances as the prover's host structs (the arrays stay the caller's)
struct StepMemConv {
  std::vector<std::vector<LutInstH>> luts;
  std::vector<std::vector<MemInstH>> mems;
  std::vector<StepMemH> steps;
  StepMemConv(const ns_step_mem* mem, size_t n_steps, bool check) : luts(n_steps), mems(n_steps), steps(n_steps) {
    for (size_t s = 0; mem && s < n_steps; s++) {
      for (size_t i = 0; i < mem[s].n_lut; i++) {
        const ns_lut_inst& a = mem[s].luts[i];
        if (check && !a.table && a.k) throw Error(1, "LUT instance without a table");
        if (check && (a.d * a.ell > 32 || a.k != ((size_t)1 << (a.d * a.ell))))
          throw Error(1, "LUT table size must be 2^(d*ell)");
        luts[s].push_back(LutInstH{a.k, a.d, a.n_side, a.steps, a.ell, a.table});
      }
      for (size_t i = 0; i < mem[s].n_mem; i++) {
        const ns_mem_inst& a = mem[s].mems[i];
        if (check && a.d * a.ell > 32) throw Error(1, "Twist address width above 32 bits");
        for (size_t q = 1; check && a.init_sparse && q < a.ninit; q++)
          if (a.init_addrs[q] <= a.init_addrs[q - 1])
            throw Error(1, "MemInit::Sparse must be strictly increasing by address");
        mems[s].push_back(MemInstH{a.k, a.d, a.n_side, a.steps, a.ell, a.init_sparse != 0,
                                   a.init_sparse ? a.ninit : 0, a.init_addrs, a.init_vals});
      }
      steps[s] = StepMemH{luts[s].size(), luts[s].data(), mems[s].size(), mems[s].data()};
    }
  }
};

static ObCfgH ob_in(const ns_output_binding& a) {
  ObCfgH o;
  o.mem_idx = a.mem_idx;
  o.num_bits = a.num_bits;
  o.n_claims = a.n_claims;
  o.addrs = a.addrs;
  o.values = a.values;
  o.final_state = a.final_state;
  return o;
}

static void prove_shard_impl(ns_prover* p, const uint64_t* const* z_steps, int z_on_device, size_t n_steps,
                             size_t m_in, const ns_step_mem* mem, const ns_output_binding* ob, uint8_t** proof,
                             size_t* proof_len, double* seconds) {
  if (!p || !z_steps || !proof || !proof_len) throw Error(1, "null argument");
  if (n_steps == 0) throw Error(1, "need at least one step");
  StepMemConv cv(mem, n_steps, true);
  ObCfgH oc;
  if (ob) oc = ob_in(*ob);
  std::vector<uint64_t> enc = p->p->prove_shard(z_steps, n_steps, m_in, z_on_device != 0,
                                                mem ? cv.steps.data() : nullptr, ob ? &oc : nullptr);
  uint8_t* buf = (uint8_t*)malloc(enc.size() * 8);
  if (!buf) throw Error(1, "out of host memory");
  memcpy(buf, enc.data(), enc.size() * 8);
  *proof = buf;
  *proof_len = enc.size() * 8;
  if (seconds) *seconds = p->p->last_seconds;
}

int ns_prove_shard_mem(ns_prover* p, const uint64_t* const* z_steps, int z_on_device, size_t n_steps, size_t m_in,
                       const ns_step_mem* mem, uint8_t** proof, size_t* proof_len, double* seconds) {
  return guard([&] { prove_shard_impl(p, z_steps, z_on_device, n_steps, m_in, mem, nullptr, proof, proof_len, seconds); });
}

int ns_prove_shard_ob(ns_prover* p, const uint64_t* const* z_steps, int z_on_device, size_t n_steps, size_t m_in,
                      const ns_step_mem* mem, const ns_output_binding* ob, uint8_t** proof, size_t* proof_len,
                      double* seconds) {
  return guard([&] {
    if (!ob) throw Error(1, "null output binding");
    prove_shard_impl(p, z_steps, z_on_device, n_steps, m_in, mem, ob, proof, proof_len, seconds);
  });
}

int ns_shard_verify(const ns_ccs* s, const uint64_t* const* x_steps, size_t n_steps, size_t m_in,
                    const ns_step_mem* mem, const uint64_t* mcs_commitments, const uint8_t* proof, size_t proof_len) {
  return ns_shard_verify_ob(s, x_steps, n_steps, m_in, mem, nullptr, mcs_commitments, proof, proof_len);
}

int ns_shard_verify_ob(const ns_ccs* s, const uint64_t* const* x_steps, size_t n_steps, size_t m_in,
                       const ns_step_mem* mem, const ns_output_binding* ob, const uint64_t* mcs_commitments,
                       const uint8_t* proof, size_t proof_len) {
  return guard([&] {
    if (!s || !proof || (n_steps && !x_steps)) throw Error(1, "null argument");
    StepMemConv cv(mem, n_steps, false);
    ObCfgH oc;
    if (ob) oc = ob_in(*ob);
    verify_shard(*s->s, x_steps, n_steps, m_in, mem ? cv.steps.data() : nullptr, mcs_commitments, proof, proof_len,
                 ob ? &oc : nullptr);
  });
}

static MeHost me_in(const ns_me& a) {
  MeHost e;
  if ((a.m_in && !a.X) || (a.ell_n && !a.r) || (a.t && (!a.y || !a.y_scalars))) throw Error(1, "ME arrays missing");
  memcpy(e.c, a.c, sizeof(e.c));
  e.m_in = a.m_in;
  e.X.assign(a.X, a.X + (size_t)D * a.m_in);
  for (size_t i = 0; i < a.ell_n; i++) e.r.push_back(K2{a.r[2 * i], a.r[2 * i + 1]});
  e.t = a.t;
  for (size_t i = 0; i < a.t * DPAD; i++) e.y.push_back(K2{a.y[2 * i], a.y[2 * i + 1]});
  for (size_t i = 0; i < a.t; i++) e.y_scalars.push_back(K2{a.y_scalars[2 * i], a.y_scalars[2 * i + 1]});
  memcpy(e.fold_digest, a.fold_digest, 32);
  return e;
}

static void me_out(const MeHost& e, ns_me& a) {
  if (a.m_in != e.m_in || a.ell_n != e.r.size() || a.t != e.t) throw Error(1, "output ME arrays have other shapes");
  if ((a.m_in && !a.X) || (a.ell_n && !a.r) || (a.t && (!a.y || !a.y_scalars))) throw Error(1, "ME arrays missing");
  memcpy(a.c, e.c, sizeof(a.c));
  std::copy(e.X.begin(), e.X.end(), a.X);
  for (size_t i = 0; i < e.r.size(); i++) a.r[2 * i] = e.r[i].c0, a.r[2 * i + 1] = e.r[i].c1;
  for (size_t i = 0; i < e.y.size(); i++) a.y[2 * i] = e.y[i].c0, a.y[2 * i + 1] = e.y[i].c1;
  for (size_t i = 0; i < e.t; i++) a.y_scalars[2 * i] = e.y_scalars[i].c0, a.y_scalars[2 * i + 1] = e.y_scalars[i].c1;
  memcpy(a.fold_digest, e.fold_digest, 32);
}

int ns_rlc(ns_prover* p, const ns_me* ins, size_t k, const int8_t* rho_coeffs, const int8_t* const* Z_dev,
           ns_me* parent, int32_t* zmix_dev) {
  return guard([&] {
    require_device();
    if (!p || !ins || !k || !Z_dev || !parent || !zmix_dev || (k > 1 && !rho_coeffs)) throw Error(1, "null argument");
    if (k > 64) throw Error(1, "at most 64 RLC inputs");
    std::vector<MeHost> v;
    for (size_t i = 0; i < k; i++) v.push_back(me_in(ins[i]));
    MeHost par;
    p->p->rlc_api(v, rho_coeffs, Z_dev, par, zmix_dev);
    me_out(par, *parent);
  });
}

int ns_dec(ns_prover* p, const ns_me* parent, const int32_t* zmix_dev, size_t k_dec, int8_t* const* planes_dev,
           size_t bus_chunk, ns_me* children, int* ok) {
  return guard([&] {
    require_device();
    if (!p || !parent || !zmix_dev || !planes_dev || !children || !ok) throw Error(1, "null argument");
    std::vector<MeHost> ch;
    bool okb[3];
    p->p->dec_api(me_in(*parent), zmix_dev, k_dec, planes_dev, bus_chunk, ch, okb);
    for (size_t i = 0; i < k_dec; i++) me_out(ch[i], children[i]);
    for (int i = 0; i < 3; i++) ok[i] = okb[i] ? 1 : 0;
  });
}

int ns_prover_set_timing(ns_prover* p, int on) {
  return guard([&] { p->p->timing = on != 0; });
}

int ns_prover_kernel_stats(const ns_prover* p, double* ms, uint64_t* counts, size_t n) {
  return guard([&] {
    for (size_t i = 0; i < n && i < (size_t)KT_N; i++) {
      ms[i] = p->p->kernel_ms[i];
      counts[i] = p->p->kernel_count[i];
    }
  });
}

int ns_prover_phase_stats(const ns_prover* p, double* seconds, size_t n) {
  return guard([&] {
    for (size_t i = 0; i < n && i < (size_t)Prover::NPHASE; i++) seconds[i] = p->p->phase_s[i];
  });
}

int ns_prover_phase_perms(const ns_prover* p, uint64_t* perms, size_t n) {
  return guard([&] {
    for (size_t i = 0; i < n && i < (size_t)Prover::NPHASE; i++) perms[i] = p->p->phase_perm[i];
  });
}

int ns_poseidon2_permute(uint64_t* state, size_t width, int force_scalar) {
  return guard([&] {
    if (width != 8 && width != 16) throw Error(1, "Poseidon2 width must be 8 or 16");
    for (size_t i = 0; i < width; i++)
      if (state[i] >= Q) throw Error(1, "state element is not a canonical field element");
    const Poseidon2& P = width == 8 ? transcript_perm() : digest_perm();
    if (force_scalar) P.permute_scalar(state);
    else P.permute(state);
  });
}

void ns_prover_free(ns_prover* p) { delete p; }

}  // extern "C"

// ---------------- Twist / Shout oracles ----------------
struct ns_ts {
  std::unique_ptr<TsOracleGpu> o;
};

static SparseCol to_col(const ns_sparse_col& c) {
  SparseCol s;
  s.idx.resize(c.nnz);
  s.val.resize(c.nnz);
  for (size_t i = 0; i < c.nnz; i++) {
    if (c.idx[i] > 0xFFFFFFFFull) throw Error(1, "sparse column index exceeds 32 bits");
    s.idx[i] = (uint32_t)c.idx[i];
    s.val[i] = K2{c.vals[2 * i], c.vals[2 * i + 1]};
    if (s.val[i].c0 >= Q || s.val[i].c1 >= Q) throw Error(1, "sparse column value is not canonical");
  }
  return s;
}

static std::vector<SparseCol> to_cols(const ns_sparse_col* c, size_t n) {
  std::vector<SparseCol> v;
  for (size_t i = 0; i < n; i++) v.push_back(to_col(c[i]));
  return v;
}

template <class F>
static int ts_make(ns_ts** out, uint64_t* claim, F&& f) {
  return guard([&] {
    if (!out) throw Error(1, "null output handle");
    auto h = std::make_unique<ns_ts>();
    h->o = std::make_unique<TsOracleGpu>();
    f(*h->o);
    if (claim) {
      K2 c = h->o->claim();
      claim[0] = c.c0;
      claim[1] = c.c1;
    }
    *out = h.release();
  });
}

extern "C" int ns_ts_product_new(const uint64_t* const* factors, size_t nf, size_t len, size_t degree_bound,
                                 ns_ts** out) {
  return ts_make(out, nullptr, [&](TsOracleGpu& o) {
    std::vector<std::vector<K2>> fs;
    for (size_t f = 0; f < nf; f++) fs.push_back(kvec(factors[f], len));
    o.init_product(fs, degree_bound);
  });
}

extern "C" int ns_ts_shout_value_new(const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_lookup,
                                     ns_sparse_col val, uint64_t claim[2], ns_ts** out) {
  return ts_make(out, claim, [&](TsOracleGpu& o) {
    SparseCol g = to_col(has_lookup), a = to_col(val);
    o.init_time(TSK_SHOUT_VALUE, kvec(r_cycle, ell_n), ell_n, &g, &a, nullptr, {}, {}, {}, kzero(), nullptr, nullptr,
                nullptr);
  });
}

extern "C" int ns_ts_index_adapter_new(const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_lookup,
                                       const ns_sparse_col* addr_bits, size_t ell_addr, const uint64_t* r_addr,
                                       uint64_t claim[2], ns_ts** out) {
  return ts_make(out, claim, [&](TsOracleGpu& o) {
    SparseCol g = to_col(has_lookup);
    o.init_time(TSK_INDEX_ADAPTER, kvec(r_cycle, ell_n), ell_n, &g, nullptr, nullptr, to_cols(addr_bits, ell_addr),
                kvec(r_addr, ell_addr), {}, kzero(), nullptr, nullptr, nullptr);
  });
}

extern "C" int ns_ts_bitness_new(const uint64_t* r_cycle, size_t ell_n, const ns_sparse_col* cols,
                                 const uint64_t* weights, size_t ncols, ns_ts** out) {
  return ts_make(out, nullptr, [&](TsOracleGpu& o) {
    o.init_time(TSK_BITNESS, kvec(r_cycle, ell_n), ell_n, nullptr, nullptr, nullptr, to_cols(cols, ncols), {},
                kvec(weights, ncols), kzero(), nullptr, nullptr, nullptr);
  });
}

extern "C" int ns_ts_twist_read_new(const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_read, ns_sparse_col rv,
                                    const ns_sparse_col* ra_bits, ns_sparse_col has_write, ns_sparse_col inc,
                                    const ns_sparse_col* wa_bits, size_t ell_addr, const uint64_t* r_addr,
                                    const uint64_t init_at_r_addr[2], ns_ts** out) {
  return ts_make(out, nullptr, [&](TsOracleGpu& o) {
    SparseCol g = to_col(has_read), a = to_col(rv), hw = to_col(has_write), ic = to_col(inc);
    std::vector<SparseCol> wb = to_cols(wa_bits, ell_addr);
    o.init_time(TSK_TWIST_READ, kvec(r_cycle, ell_n), ell_n, &g, &a, nullptr, to_cols(ra_bits, ell_addr),
                kvec(r_addr, ell_addr), {}, K2{init_at_r_addr[0], init_at_r_addr[1]}, &hw, &ic, &wb);
  });
}

extern "C" int ns_ts_twist_write_new(const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_write,
                                     ns_sparse_col wv, ns_sparse_col inc, const ns_sparse_col* wa_bits,
                                     size_t ell_addr, const uint64_t* r_addr, const uint64_t init_at_r_addr[2],
                                     ns_ts** out) {
  return ts_make(out, nullptr, [&](TsOracleGpu& o) {
    SparseCol g = to_col(has_write), a = to_col(wv), b = to_col(inc);
    std::vector<SparseCol> wb = to_cols(wa_bits, ell_addr);
    o.init_time(TSK_TWIST_WRITE, kvec(r_cycle, ell_n), ell_n, &g, &a, &b, wb, kvec(r_addr, ell_addr), {},
                K2{init_at_r_addr[0], init_at_r_addr[1]}, &g, &b, &wb);
  });
}

extern "C" int ns_ts_val_eval_new(size_t ell_n, const ns_sparse_col* wa_bits, size_t ell_addr,
                                  ns_sparse_col has_write, ns_sparse_col inc, const uint64_t* r_addr,
                                  const uint64_t* r_time, uint64_t claim[2], ns_ts** out) {
  return ts_make(out, claim, [&](TsOracleGpu& o) {
    SparseCol g = to_col(has_write), b = to_col(inc);
    o.init_time(TSK_VAL_EVAL, kvec(r_time, ell_n), ell_n, &g, nullptr, &b, to_cols(wa_bits, ell_addr),
                kvec(r_addr, ell_addr), {}, kzero(), nullptr, nullptr, nullptr);
  });
}

extern "C" int ns_ts_total_inc_new(size_t ell_n, const ns_sparse_col* wa_bits, size_t ell_addr,
                                   ns_sparse_col has_write, ns_sparse_col inc, const uint64_t* r_addr,
                                   uint64_t claim[2], ns_ts** out) {
  return ts_make(out, claim, [&](TsOracleGpu& o) {
    SparseCol g = to_col(has_write), b = to_col(inc);
    o.init_time(TSK_TOTAL_INC, {}, ell_n, &g, nullptr, &b, to_cols(wa_bits, ell_addr), kvec(r_addr, ell_addr), {},
                kzero(), nullptr, nullptr, nullptr);
  });
}

static void init_lists(const uint64_t* init_addrs, const uint64_t* init_vals, size_t ninit, std::vector<uint32_t>& a,
                       std::vector<K2>& v) {
  for (size_t i = 0; i < ninit; i++) {
    if (init_addrs[i] > 0xFFFFFFFFull) throw Error(1, "init address exceeds 32 bits");
    a.push_back((uint32_t)init_addrs[i]);
    v.push_back(K2{init_vals[2 * i], init_vals[2 * i + 1]});
  }
}

extern "C" int ns_ts_read_addr_new(const uint64_t* init_addrs, const uint64_t* init_vals, size_t ninit,
                                   const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_read, ns_sparse_col rv,
                                   const ns_sparse_col* ra_bits, ns_sparse_col has_write,
                                   const ns_sparse_col* wa_bits, ns_sparse_col inc, size_t ell_addr, ns_ts** out) {
  return ts_make(out, nullptr, [&](TsOracleGpu& o) {
    std::vector<uint32_t> ia;
    std::vector<K2> iv;
    init_lists(init_addrs, init_vals, ninit, ia, iv);
    SparseCol hr = to_col(has_read), r = to_col(rv), hw = to_col(has_write), ic = to_col(inc);
    std::vector<SparseCol> rb = to_cols(ra_bits, ell_addr), wb = to_cols(wa_bits, ell_addr);
    o.init_addr(TSK_READ_ADDR, ia, iv, kvec(r_cycle, ell_n), ell_n, &hr, &r, &rb, hw, ic, wb, ic, ell_addr);
  });
}

extern "C" int ns_ts_write_addr_new(const uint64_t* init_addrs, const uint64_t* init_vals, size_t ninit,
                                    const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_write, ns_sparse_col wv,
                                    const ns_sparse_col* wa_bits, ns_sparse_col inc, size_t ell_addr, ns_ts** out) {
  return ts_make(out, nullptr, [&](TsOracleGpu& o) {
    std::vector<uint32_t> ia;
    std::vector<K2> iv;
    init_lists(init_addrs, init_vals, ninit, ia, iv);
    SparseCol hw = to_col(has_write), w = to_col(wv), ic = to_col(inc);
    std::vector<SparseCol> wb = to_cols(wa_bits, ell_addr);
    o.init_addr(TSK_WRITE_ADDR, ia, iv, kvec(r_cycle, ell_n), ell_n, nullptr, nullptr, nullptr, hw, w, wb, ic,
                ell_addr);
  });
}

extern "C" int ns_ts_addr_lookup_new(const ns_sparse_col* addr_bits, size_t ell_addr, ns_sparse_col has_lookup,
                                     const uint64_t* table, size_t table_len, const uint64_t* r_cycle, size_t ell_n,
                                     uint64_t claim[2], ns_ts** out) {
  return ts_make(out, claim, [&](TsOracleGpu& o) {
    o.init_lookup(to_cols(addr_bits, ell_addr), to_col(has_lookup), kvec(table, table_len), kvec(r_cycle, ell_n),
                  ell_n);
  });
}

extern "C" int ns_ts_evals(ns_ts* o, const uint64_t* xs, size_t n, uint64_t* out) {
  return guard([&] {
    if (!o) throw Error(1, "null oracle");
    std::vector<K2> x = kvec(xs, n), y(n);
    o->o->evals(x.data(), n, y.data());
    for (size_t i = 0; i < n; i++) {
      out[2 * i] = y[i].c0;
      out[2 * i + 1] = y[i].c1;
    }
  });
}

extern "C" int ns_ts_fold(ns_ts* o, const uint64_t r[2]) {
  return guard([&] {
    if (!o) throw Error(1, "null oracle");
    o->o->fold(K2{r[0], r[1]});
  });
}

static std::vector<TsOracleGpu*> ts_handles(ns_ts* const* os, size_t n) {
  if (n && !os) throw Error(1, "null oracle array");
  std::vector<TsOracleGpu*> v(n);
  for (size_t i = 0; i < n; i++) {
    if (!os[i]) throw Error(1, "null oracle");
    v[i] = os[i]->o.get();
  }
  return v;
}

extern "C" int ns_ts_batch_round(ns_ts* const* os, size_t n, uint64_t* out) {
  return guard([&] {
    std::vector<TsOracleGpu*> v = ts_handles(os, n);
    std::vector<std::vector<K2>> cs;
    ts_batch_round(v.data(), n, cs);
    size_t w = 0;
    for (size_t i = 0; i < n; i++) {
      const size_t deg = v[i]->degree_bound();
      // values at x = 0..deg of the coefficient vector (Horner), deg + 1 per oracle
      for (size_t x = 0; x <= deg; x++) {
        const K2 xk{(uint64_t)x, 0};
        K2 acc = kzero();
        for (size_t d = cs[i].size(); d-- > 0;) acc = kadd(kmul(acc, xk), cs[i][d]);
        out[w++] = acc.c0;
        out[w++] = acc.c1;
      }
    }
  });
}

extern "C" int ns_ts_batch_fold(ns_ts* const* os, size_t n, const uint64_t r[2]) {
  return guard([&] {
    std::vector<TsOracleGpu*> v = ts_handles(os, n);
    ts_batch_fold(v.data(), n, K2{r[0], r[1]});
  });
}

extern "C" size_t ns_ts_num_rounds(const ns_ts* o) { return o ? o->o->num_rounds() : 0; }
extern "C" size_t ns_ts_degree_bound(const ns_ts* o) { return o ? o->o->degree_bound() : 0; }
extern "C" void ns_ts_free(ns_ts* o) { delete o; }

// ---------------- shared CPU bus ----------------
extern "C" int ns_bus_decode(const int8_t* Z, size_t m, uint64_t* z_out) {
  return guard([&] {
    require_device();
    bus_decode(Z, m, z_out);
  });
}

extern "C" int ns_bus_time_cols(const int8_t* Z, size_t m, size_t m_in, size_t bus_base, size_t chunk,
                                size_t bus_cols, size_t steps_len, size_t ell_n, uint64_t* idx_out, uint64_t* vals_out,
                                size_t* nnz_out) {
  return guard([&] {
    require_device();
    bus_time_cols(Z, m, m_in, bus_base, chunk, bus_cols, steps_len, ell_n, idx_out, vals_out, nnz_out);
  });
}

extern "C" int ns_bus_openings(const int8_t* Z, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                               const uint64_t* r, size_t ell_n, uint64_t* y_out, uint64_t* y_scalars_out) {
  return guard([&] {
    require_device();
    std::vector<K2> rr = kvec(r, ell_n), y(bus_cols * 64), ys(bus_cols);
    bus_openings(Z, m, m_in, bus_base, chunk, bus_cols, rr.data(), ell_n, y.data(), ys.data());
    for (size_t i = 0; i < y.size(); i++) {
      y_out[2 * i] = y[i].c0;
      y_out[2 * i + 1] = y[i].c1;
    }
    for (size_t i = 0; i < ys.size(); i++) {
      y_scalars_out[2 * i] = ys[i].c0;
      y_scalars_out[2 * i + 1] = ys[i].c1;
    }
  });
}

// ---------------- device memory ----------------
extern "C" int ns_device_alloc(size_t bytes, void** out) {
  return guard([&] {
    require_device();
    if (!out) throw Error(1, "null output pointer");
    *out = nullptr;
    NS_CHECK(hipMalloc(out, bytes ? bytes : 1));
  });
}

extern "C" int ns_device_free(void* p) {
  return guard([&] {
    if (p) NS_CHECK(hipFree(p));
  });
}

extern "C" int ns_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes) {
  return guard([&] { NS_CHECK(hipMemcpy(dst_dev, src_host, bytes, hipMemcpyHostToDevice)); });
}

extern "C" int ns_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes) {
  return guard([&] { NS_CHECK(hipMemcpy(dst_host, src_dev, bytes, hipMemcpyDeviceToHost)); });
}

extern "C" int ns_device_synchronize(void) {
  return guard([&] { NS_CHECK(hipDeviceSynchronize()); });
}
