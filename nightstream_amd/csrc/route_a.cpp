// route_a.cpp -- host side of the Route A memory sidecar (see route_a.h).
#include "route_a.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_set>

namespace ns {

// ---------------- layout and guardrails ----------------
BusLayout infer_bus_layout(const StepMemH* mem, size_t nsteps, size_t n, size_t m, size_t m_in) {
  BusLayout L;
  L.m = m;
  L.m_in = m_in;
  size_t chunk = 0;
  for (size_t s = 0; s < nsteps; s++) {
    for (size_t i = 0; i < mem[s].n_lut; i++) chunk = std::max(chunk, mem[s].luts[i].steps);
    for (size_t i = 0; i < mem[s].n_mem; i++) chunk = std::max(chunk, mem[s].mems[i].steps);
  }
  L.chunk = chunk ? chunk : 1;
  if (nsteps == 0) return L;
  for (size_t s = 1; s < nsteps; s++) {
    bool bad = mem[s].n_lut != mem[0].n_lut || mem[s].n_mem != mem[0].n_mem;
    for (size_t i = 0; !bad && i < mem[0].n_lut; i++)
      bad = mem[s].luts[i].d * mem[s].luts[i].ell != mem[0].luts[i].d * mem[0].luts[i].ell;
    for (size_t i = 0; !bad && i < mem[0].n_mem; i++)
      bad = mem[s].mems[i].d * mem[s].mems[i].ell != mem[0].mems[i].d * mem[0].mems[i].ell;
    if (bad) throw Error(1, "shared CPU bus layout mismatch across steps (step 0 vs step " + std::to_string(s) + ")");
  }
  size_t col = 0;
  for (size_t i = 0; i < mem[0].n_lut; i++) {
    const size_t e = mem[0].luts[i].d * mem[0].luts[i].ell;
    L.sh.push_back({col, e, col + e, col + e + 1});
    col += e + 2;
  }
  for (size_t i = 0; i < mem[0].n_mem; i++) {
    const size_t e = mem[0].mems[i].d * mem[0].mems[i].ell;
    L.tw.push_back({col, col + e, e, col + 2 * e, col + 2 * e + 1, col + 2 * e + 2, col + 2 * e + 3, col + 2 * e + 4});
    col += 2 * e + 5;
  }
  L.bus_cols = col;
  if (col * L.chunk > m)
    throw Error(1, "BusLayout: bus region too large: bus_cols(" + std::to_string(col) + ") * chunk_size(" +
                       std::to_string(L.chunk) + ") > m(" + std::to_string(m) + ")");
  L.bus_base = m - col * L.chunk;
  if (L.bus_base < m_in) throw Error(1, "BusLayout: bus_base overlaps public inputs");
  if (col > 0 && m_in + L.chunk > n) throw Error(1, "bus time rows out of range: m_in + chunk_size > n");
  return L;
}

static size_t cell(const BusLayout& L, size_t col, size_t j) { return L.bus_base + col * L.chunk + j; }

void check_bus_guardrails(const Ccs& s, const BusLayout& L, const std::vector<std::vector<uint64_t>>& xs) {
  if (L.bus_cols == 0) return;
  if (s.host.size() < 4) throw Error(1, "shared_cpu_bus=true but CPU CCS has fewer than 3 matrices");
  // public constant-one columns (cpu_bus.rs:694-721)
  std::vector<size_t> ones;
  for (size_t c = 0; c < L.m_in; c++) {
    bool all = !xs.empty();
    for (auto& x : xs) all = all && c < x.size() && x[c] == 1;
    if (all) ones.push_back(c);
  }
  if (ones.empty()) throw Error(1, "shared_cpu_bus=true but no public constant-one column found");
  const HostCsr &A = s.host[1], &B = s.host[2], &C = s.host[3];
  const size_t n = s.n, m = s.m;
  // canonical padding rows: A = (one - flag), B = field, C = 0 (cpu_bus.rs:723-939)
  std::unordered_set<uint64_t> present;
  std::vector<uint8_t> padding(n, 0);
  for (size_t r = 0; r < n; r++) {
    if (C.ptr[r + 1] != C.ptr[r]) continue;
    if (B.ptr[r + 1] - B.ptr[r] != 1) continue;
    if (A.ptr[r + 1] - A.ptr[r] != 2) continue;
    const size_t a0 = A.ptr[r];
    const uint32_t c1 = A.col[a0], c2 = A.col[a0 + 1];
    const uint64_t v1 = A.val[a0], v2 = A.val[a0 + 1];
    if (v1 != fneg(v2)) continue;
    const bool one1 = std::find(ones.begin(), ones.end(), (size_t)c1) != ones.end();
    const bool one2 = std::find(ones.begin(), ones.end(), (size_t)c2) != ones.end();
    size_t flag;
    if (one1) flag = c2;
    else if (one2) flag = c1;
    else continue;
    present.insert((uint64_t)flag * m + B.col[B.ptr[r]]);
    padding[r] = 1;
  }
  auto need_pad = [&](size_t flag_col, size_t field_col, const std::string& what) {
    for (size_t j = 0; j < L.chunk; j++)
      if (!present.count((uint64_t)cell(L, flag_col, j) * m + cell(L, field_col, j)))
        throw Error(1, "shared_cpu_bus=true but CPU CCS is missing the padding constraint " + what + " at j=" +
                           std::to_string(j));
  };
  for (size_t i = 0; i < L.sh.size(); i++) {
    need_pad(L.sh[i].has_lookup, L.sh[i].val, "(1-has_lookup)*val");
    for (size_t b = 0; b < L.sh[i].ell; b++) need_pad(L.sh[i].has_lookup, L.sh[i].addr0 + b, "(1-has_lookup)*addr_bits");
  }
  for (auto& t : L.tw) {
    need_pad(t.has_read, t.rv, "(1-has_read)*rv");
    for (size_t b = 0; b < t.ell; b++) need_pad(t.has_read, t.ra0 + b, "(1-has_read)*ra_bits");
    need_pad(t.has_write, t.wv, "(1-has_write)*wv");
    need_pad(t.has_write, t.inc, "(1-has_write)*inc_at_write_addr");
    for (size_t b = 0; b < t.ell; b++) need_pad(t.has_write, t.wa0 + b, "(1-has_write)*wa_bits");
  }
  // every bus cell referenced by an active matrix, and every one but inc outside padding rows
  // (A, B, C are the matrices f uses: f = M1 M2 - M3)
  std::vector<uint8_t> any(m, 0), outside(m, 0);
  for (const HostCsr* M : {&A, &B, &C})
    for (size_t r = 0; r < n; r++)
      for (size_t q = M->ptr[r]; q < M->ptr[r + 1]; q++) {
        any[M->col[q]] = 1;
        if (!padding[r]) outside[M->col[q]] = 1;
      }
  std::vector<uint8_t> is_inc(L.bus_cols, 0);
  for (auto& t : L.tw) is_inc[t.inc] = 1;
  for (size_t col = 0; col < L.bus_cols; col++)
    for (size_t j = 0; j < L.chunk; j++) {
      const size_t z = cell(L, col, j);
      if (!any[z])
        throw Error(1, "shared_cpu_bus=true but CPU CCS does not reference bus column " + std::to_string(col) +
                           " in any active constraint matrix");
      if (!is_inc[col] && !outside[z])
        throw Error(1, "shared_cpu_bus=true but CPU CCS never references bus column " + std::to_string(col) +
                           " outside the canonical padding rows");
    }
}

// ---------------- transcript pieces ----------------
static void msg_str(Transcript& tr, const char* label, const char* s) { tr.append_message(label, s); }

// digest_fields (memory_sidecar/transcript.rs:4-10)
static std::vector<uint8_t> digest_fields(const char* label, const uint64_t* fs, size_t n) {
  Transcript h("memory/public_digest");
  msg_str(h, "digest/label", label);
  h.append_msg_u64("digest/len", n);
  h.append_fields("digest/fields", fs, n);
  std::vector<uint8_t> out(32);
  h.digest32(out.data());
  return out;
}

void absorb_step_memory(Transcript& tr, const StepMemH& sm, DigestCache& cache) {
  tr.append_message("step/absorb_memory_start", (const uint8_t*)nullptr, 0);
  tr.append_msg_u64("step/lut_count", sm.n_lut);
  for (size_t i = 0; i < sm.n_lut; i++) {
    const LutInstH& in = sm.luts[i];
    tr.append_msg_u64("step/lut_idx", i);
    tr.append_msg_u64("shout/k", in.k);
    tr.append_msg_u64("shout/d", in.d);
    tr.append_msg_u64("shout/n_side", in.n_side);
    tr.append_msg_u64("shout/steps", in.steps);
    tr.append_msg_u64("shout/ell", in.ell);
    auto key = std::make_pair(in.table, in.k);
    auto it = cache.tables.find(key);
    if (it == cache.tables.end()) it = cache.tables.emplace(key, digest_fields("shout/table", in.table, in.k)).first;
    tr.append_message("shout/table_digest", it->second.data(), 32);
  }
  tr.append_msg_u64("step/mem_count", sm.n_mem);
  for (size_t i = 0; i < sm.n_mem; i++) {
    const MemInstH& in = sm.mems[i];
    tr.append_msg_u64("step/mem_idx", i);
    tr.append_msg_u64("twist/k", in.k);
    tr.append_msg_u64("twist/d", in.d);
    tr.append_msg_u64("twist/n_side", in.n_side);
    tr.append_msg_u64("twist/steps", in.steps);
    tr.append_msg_u64("twist/ell", in.ell);
    std::vector<uint8_t> dg;
    if (!in.sparse) {
      dg = digest_fields("twist/init/zero", nullptr, 0);
    } else {
      std::vector<uint64_t> fs(2 * in.ninit);
      for (size_t q = 0; q < in.ninit; q++) {
        fs[2 * q] = in.init_addrs[q] % Q;
        fs[2 * q + 1] = in.init_vals[q];
      }
      dg = digest_fields("twist/init/sparse", fs.data(), fs.size());
    }
    tr.append_message("twist/init_digest", dg.data(), 32);
  }
  tr.append_message("step/absorb_memory_done", (const uint8_t*)nullptr, 0);
}

// bind_batched_claim_sums (transcript.rs:12-25)
static void bind_claim_sums(Transcript& tr, const char* prefix, const std::vector<K2>& sums,
                            const std::vector<const char*>& labels) {
  tr.append_msg_u64(prefix, sums.size());
  for (size_t i = 0; i < sums.size(); i++) {
    msg_str(tr, "addr_batch/label", labels[i]);
    tr.append_msg_u64("addr_batch/idx", i);
    tr.append_k("addr_batch/claimed_sum", sums[i]);
  }
}

void bind_dynamic_claims(Transcript& tr, const std::vector<RaClaim>& cl) {
  tr.append_msg_u64("batched/dynamic_bind/len", cl.size());
  for (size_t i = 0; i < cl.size(); i++) {
    msg_str(tr, "batched/dynamic_bind/claim_label", cl[i].label);
    tr.append_msg_u64("batched/dynamic_bind/claim_idx", i);
    tr.append_msg_u64("batched/dynamic_bind/degree_bound", cl[i].deg);
    const uint8_t d = cl[i].dyn ? 1 : 0;
    tr.append_message("batched/dynamic_bind/is_dynamic", &d, 1);
    if (cl[i].dyn) tr.append_k("batched/dynamic_bind/claimed_sum", cl[i].sum);
  }
}

// bitness_weights (memory_sidecar/utils.rs:43-72)
static std::vector<K2> bitness_weights(const std::vector<K2>& r_cycle, size_t n, uint64_t sep) {
  K2 base{sep % Q, 0};
  if (!r_cycle.empty()) base = kadd(r_cycle[0], base);
  const K2 neg1{Q - 1, 0};
  if (keq(base, kzero()) || keq(base, kone()) || keq(base, neg1)) base = kadd(base, kone());
  if (keq(base, kzero())) base = kone();
  std::vector<K2> w(n);
  K2 cur = kone();
  for (size_t i = 0; i < n; i++) {
    w[i] = cur;
    cur = kmul(cur, base);
  }
  return w;
}

// eval_init_at_r_addr (neo-memory/src/mem_init.rs:133-165)
K2 eval_init(const MemInstH& in, const std::vector<K2>& r_addr) {
  if (!in.sparse) return kzero();
  K2 acc = kzero();
  for (size_t q = 0; q < in.ninit; q++) {
    K2 chi = kone();
    for (size_t b = 0; b < r_addr.size(); b++)
      chi = kmul(chi, ((in.init_addrs[q] >> b) & 1) ? r_addr[b] : ksub(kone(), r_addr[b]));
    acc = kadd(acc, kscale(chi, in.init_vals[q]));
  }
  return acc;
}

// ---------------- bus time columns (cpu_bus.rs:976-1016) ----------------
static SparseCol bus_col(const BusLayout& L, const uint64_t* zbus, size_t col, size_t steps, size_t ell_n) {
  if (col >= L.bus_cols) throw Error(1, "bus col_id out of range");
  if (steps > L.chunk) throw Error(1, "steps_len > bus.chunk_size");
  if (L.m_in + L.chunk > ((size_t)1 << ell_n)) throw Error(1, "bus time index out of range for ell_n");
  SparseCol c;
  const uint64_t* z = zbus + col * L.chunk;
  for (size_t j = 0; j < steps; j++)
    if (z[j]) {
      c.idx.push_back((uint32_t)(L.m_in + j));
      c.val.push_back(K2{z[j], 0});
    }
  return c;
}

// ---------------- batched sum-check ----------------
static K2 poly_at(const std::vector<K2>& c, size_t off, size_t n, K2 x) {
  K2 acc = kzero();
  for (size_t i = n; i-- > 0;) acc = kadd(kmul(acc, x), c[off + i]);
  return acc;
}

void run_batched(Transcript& tr, const char* domain, size_t inst, std::vector<RaClaim>& cl, size_t rounds,
                 std::vector<K2>& chals) {
  msg_str(tr, "sc/domain", domain);
  tr.append_msg_u64("sc/inst_idx", inst);
  chals.assign(rounds, kzero());
  for (auto& c : cl) {
    c.running = c.sum;
    c.polys.assign(rounds * (c.deg + 1), kzero());
    const size_t want = c.ts ? c.ts->num_rounds() : c.ccs->ell_n() - c.ccs->round();
    if (want < rounds)
      throw Error(7, std::string(domain) + ": claim " + c.label + " has fewer rounds than the batch");
  }
  std::vector<K2> co;
  std::vector<TsOracleGpu*> tso;  // the device oracles step as one batch (shared launches)
  for (auto& c : cl)
    if (c.ts) tso.push_back(c.ts);
  for (size_t rd = 0; rd < rounds; rd++) {
    tr.append_msg_u64("batched/round_idx", rd);
    if (!tso.empty()) ts_batch_launch_round(tso.data(), tso.size());
    for (size_t i = 0; i < cl.size(); i++) {
      RaClaim& c = cl[i];
      if (c.ts) c.ts->finish_round(co);
      else co = c.ccs->round_coeffs();
      const size_t np = c.deg + 1;
      for (size_t q = np; q < co.size(); q++)
        if (!keq(co[q], kzero())) throw Error(7, std::string(c.label) + ": round polynomial exceeds its degree bound");
      for (size_t q = 0; q < np; q++) c.polys[rd * np + q] = q < co.size() ? co[q] : kzero();
      const K2 s01 = kadd(poly_at(c.polys, rd * np, np, kzero()), poly_at(c.polys, rd * np, np, kone()));
      if (!keq(s01, c.running))
        throw Error(4, std::string(domain) + ": claim " + std::to_string(i) + " (" + c.label + ") round " +
                           std::to_string(rd) + " invariant failed");
    }
    for (size_t i = 0; i < cl.size(); i++) {
      const RaClaim& c = cl[i];
      const size_t np = c.deg + 1;
      msg_str(tr, "batched/claim_label", c.label);
      tr.append_msg_u64("batched/claim_idx", i);
      for (size_t q = 0; q < np; q++) tr.append_k("batched/round/coeff", c.polys[rd * np + q]);
    }
    const K2 r = tr.challenge_k("batched/challenge/0", "batched/challenge/1");
    chals[rd] = r;
    for (auto& c : cl) {
      c.running = poly_at(c.polys, rd * (c.deg + 1), c.deg + 1, r);
      if (!c.ts) c.ccs->fold(r);
    }
    if (!tso.empty()) ts_batch_fold(tso.data(), tso.size(), r);
  }
  for (auto& c : cl)
    if (c.ts) c.ts->finish_fold();
  msg_str(tr, "sc/domain_end", domain);
  tr.append_msg_u64("sc/inst_idx_end", inst);
}

// ---------------- addr-pre (memory.rs:232-621) ----------------
void shout_addr_pre(Transcript& tr, const StepMemH& sm, const BusLayout& L, const uint64_t* zbus, size_t ell_n,
                    const std::vector<K2>& r_cycle, size_t step_idx, ShoutPre& out) {
  out = ShoutPre{};
  const size_t nl = sm.n_lut;
  if (nl == 0) return;
  if (nl > 64) throw Error(1, "Shout(Route A): skip mask supports up to 64 Shout instances");
  out.claimed.assign(nl, kzero());
  out.cols.resize(nl);
  std::vector<std::unique_ptr<TsOracleGpu>> owned;
  std::vector<RaClaim> cl;
  for (size_t i = 0; i < nl; i++) {
    const LutInstH& in = sm.luts[i];
    const size_t e = in.d * in.ell;
    if (i == 0) out.ell_addr = e;
    else if (e != out.ell_addr) throw Error(1, "Shout(Route A): batched addr-pre requires uniform ell_addr");
    if (in.steps > ((size_t)1 << ell_n)) throw Error(1, "Shout(Route A): steps exceeds 2^ell_cycle");
    ShoutCols& S = out.cols[i];
    for (size_t b = 0; b < e; b++) S.addr.push_back(bus_col(L, zbus, L.sh[i].addr0 + b, in.steps, ell_n));
    S.has_lookup = bus_col(L, zbus, L.sh[i].has_lookup, in.steps, ell_n);
    S.val = bus_col(L, zbus, L.sh[i].val, in.steps, ell_n);
    if (!S.has_lookup.idx.empty()) {
      std::vector<K2> tab(in.k);
      for (size_t q = 0; q < in.k; q++) tab[q] = K2{in.table[q], 0};
      owned.push_back(std::make_unique<TsOracleGpu>());
      TsOracleGpu* o = owned.back().get();
      o->init_lookup(S.addr, S.has_lookup, tab, r_cycle, ell_n);
      RaClaim c;
      c.ts = o;
      c.deg = o->degree_bound();
      c.label = "shout/addr_pre";
      c.sum = o->claim();
      out.claimed[i] = c.sum;
      out.active_mask |= (uint64_t)1 << i;
      cl.push_back(c);
    }
  }
  tr.append_msg_u64("shout/addr_pre_time/step_idx", step_idx);
  bind_claim_sums(tr, "shout/addr_pre_time/claimed_sums", out.claimed, std::vector<const char*>(nl, "shout/addr_pre"));
  if (cl.empty()) {
    tr.append_msg_u64("shout/addr_pre_time/no_sumcheck", step_idx);
    out.r_addr.resize(out.ell_addr);
    for (size_t i = 0; i < out.ell_addr; i++) {
      tr.append_msg_u64("shout/addr_pre_time/no_sumcheck/r_addr", i);
      out.r_addr[i] =
          tr.challenge_k("shout/addr_pre_time/no_sumcheck/r_addr/0", "shout/addr_pre_time/no_sumcheck/r_addr/1");
    }
    return;
  }
  for (auto& c : cl)
    if (c.ts->num_rounds() != out.ell_addr) throw Error(7, "shout addr-pre: oracle rounds != ell_addr");
  run_batched(tr, "shout/addr_pre_time", step_idx, cl, out.ell_addr, out.r_addr);
  for (auto& c : cl) {
    out.rounds.push_back(std::move(c.polys));
    out.degs.push_back(c.deg);
  }
}

void twist_addr_pre(Transcript& tr, const StepMemH& sm, const BusLayout& L, const uint64_t* zbus, size_t ell_n,
                    const std::vector<K2>& r_cycle, std::vector<TwistPre>& out) {
  out.clear();
  out.resize(sm.n_mem);
  for (size_t i = 0; i < sm.n_mem; i++) {
    const MemInstH& in = sm.mems[i];
    TwistPre& P = out[i];
    const size_t e = in.d * in.ell;
    if (in.steps > ((size_t)1 << ell_n)) throw Error(1, "Twist(Route A): steps exceeds 2^ell_cycle");
    TwistCols& T = P.cols;
    const BusLayout::Tw& tl = L.tw[i];
    T.ell = e;
    for (size_t b = 0; b < e; b++) {
      T.ra.push_back(bus_col(L, zbus, tl.ra0 + b, in.steps, ell_n));
      T.wa.push_back(bus_col(L, zbus, tl.wa0 + b, in.steps, ell_n));
    }
    T.has_read = bus_col(L, zbus, tl.has_read, in.steps, ell_n);
    T.has_write = bus_col(L, zbus, tl.has_write, in.steps, ell_n);
    T.wv = bus_col(L, zbus, tl.wv, in.steps, ell_n);
    T.rv = bus_col(L, zbus, tl.rv, in.steps, ell_n);
    T.inc = bus_col(L, zbus, tl.inc, in.steps, ell_n);
    std::vector<uint32_t> ia;
    std::vector<K2> iv;
    if (in.sparse)
      for (size_t q = 0; q < in.ninit; q++) {
        if (in.init_addrs[q] >= in.k) throw Error(1, "Twist: init address out of range");
        ia.push_back((uint32_t)in.init_addrs[q]);
        iv.push_back(K2{in.init_vals[q], 0});
      }
    TsOracleGpu ro, wo;
    ro.init_addr(TSK_READ_ADDR, ia, iv, r_cycle, ell_n, &T.has_read, &T.rv, &T.ra, T.has_write, T.inc, T.wa, T.inc, e);
    wo.init_addr(TSK_WRITE_ADDR, ia, iv, r_cycle, ell_n, nullptr, nullptr, nullptr, T.has_write, T.wv, T.wa, T.inc, e);
    std::vector<RaClaim> cl(2);
    cl[0].ts = &ro;
    cl[0].label = "twist/read_addr_pre";
    cl[1].ts = &wo;
    cl[1].label = "twist/write_addr_pre";
    for (auto& c : cl) c.deg = c.ts->degree_bound();
    tr.append_msg_u64("twist/addr_pre_time/claim_idx", i);
    bind_claim_sums(tr, "twist/addr_pre_time/claimed_sums", {kzero(), kzero()},
                    {"twist/read_addr_pre", "twist/write_addr_pre"});
    P.ell_addr = e;
    run_batched(tr, "twist/addr_pre_time", i, cl, e, P.r_addr);
    for (int c = 0; c < 2; c++) {
      P.rounds[c] = std::move(cl[c].polys);
      P.degs[c] = cl[c].deg;
    }
    P.read_claim = cl[0].running;
    P.write_claim = cl[1].running;
  }
}

// ---------------- time lanes (memory.rs:905-1025) ----------------
void build_time_claims(const StepMemH& sm, size_t ell_n, const std::vector<K2>& r_cycle, const ShoutPre& sp,
                       const std::vector<TwistPre>& tp, std::vector<std::unique_ptr<TsOracleGpu>>& owned,
                       std::vector<RaClaim>& cl) {
  auto add = [&](const char* label, bool dyn) -> TsOracleGpu* {
    owned.push_back(std::make_unique<TsOracleGpu>());
    RaClaim c;
    c.ts = owned.back().get();
    c.label = label;
    c.dyn = dyn;
    cl.push_back(c);
    return c.ts;
  };
  for (size_t i = 0; i < sm.n_lut; i++) {
    const ShoutCols& S = sp.cols[i];
    const size_t e = S.addr.size();
    TsOracleGpu* v = add("shout/value", true);
    v->init_time(TSK_SHOUT_VALUE, r_cycle, ell_n, &S.has_lookup, &S.val, nullptr, {}, {}, {}, kzero(), nullptr, nullptr,
                 nullptr);
    cl.back().sum = v->claim();
    TsOracleGpu* a = add("shout/adapter", true);
    a->init_time(TSK_INDEX_ADAPTER, r_cycle, ell_n, &S.has_lookup, nullptr, nullptr, S.addr, sp.r_addr, {}, kzero(),
                 nullptr, nullptr, nullptr);
    cl.back().sum = a->claim();
    std::vector<SparseCol> bits = S.addr;
    bits.push_back(S.has_lookup);
    TsOracleGpu* b = add("shout/bitness", false);
    b->init_time(TSK_BITNESS, r_cycle, ell_n, nullptr, nullptr, nullptr, bits, {},
                 bitness_weights(r_cycle, e + 1, 0x53484F5554ull + i), kzero(), nullptr, nullptr, nullptr);
  }
  for (size_t i = 0; i < sm.n_mem; i++) {
    const TwistPre& P = tp[i];
    const TwistCols& T = P.cols;
    const K2 init_at = eval_init(sm.mems[i], P.r_addr);
    TsOracleGpu* r = add("twist/read_check", true);
    r->init_time(TSK_TWIST_READ, r_cycle, ell_n, &T.has_read, &T.rv, nullptr, T.ra, P.r_addr, {}, init_at,
                 &T.has_write, &T.inc, &T.wa);
    cl.back().sum = P.read_claim;
    TsOracleGpu* w = add("twist/write_check", true);
    w->init_time(TSK_TWIST_WRITE, r_cycle, ell_n, &T.has_write, &T.wv, &T.inc, T.wa, P.r_addr, {}, init_at,
                 &T.has_write, &T.inc, &T.wa);
    cl.back().sum = P.write_claim;
    std::vector<SparseCol> bits = T.ra;
    bits.insert(bits.end(), T.wa.begin(), T.wa.end());
    bits.push_back(T.has_read);
    bits.push_back(T.has_write);
    TsOracleGpu* b = add("twist/bitness", false);
    b->init_time(TSK_BITNESS, r_cycle, ell_n, nullptr, nullptr, nullptr, bits, {},
                 bitness_weights(r_cycle, bits.size(), 0x5457495354ull + i), kzero(), nullptr, nullptr, nullptr);
  }
  for (auto& c : cl) {
    if (c.ccs) continue;
    c.deg = c.ts->degree_bound();
    if (c.ts->num_rounds() != ell_n)
      throw Error(7, std::string("Route A time oracle ") + c.label + " does not have ell_n rounds");
  }
}

// ---------------- val-eval (memory.rs:1382-1535) ----------------
void val_eval_batch(Transcript& tr, const StepMemH& sm, const std::vector<TwistPre>& tp,
                    const std::vector<TwistCols>* prev, size_t ell_n, const std::vector<K2>& r_time, size_t step_idx,
                    ValPre& out) {
  out = ValPre{};
  const size_t nm = sm.n_mem;
  if (nm == 0) return;
  out.per = prev ? 3 : 2;
  std::vector<std::unique_ptr<TsOracleGpu>> owned;
  std::vector<uint8_t> tags;
  for (size_t i = 0; i < nm; i++) {
    const TwistPre& P = tp[i];
    auto add = [&](const char* label, uint8_t tag) -> TsOracleGpu* {
      owned.push_back(std::make_unique<TsOracleGpu>());
      RaClaim c;
      c.ts = owned.back().get();
      c.label = label;
      out.claims.push_back(c);
      tags.push_back(tag);
      return c.ts;
    };
    TsOracleGpu* lt = add("twist/val_eval_lt", 0);
    lt->init_time(TSK_VAL_EVAL, r_time, ell_n, &P.cols.has_write, nullptr, &P.cols.inc, P.cols.wa, P.r_addr, {},
                  kzero(), nullptr, nullptr, nullptr);
    out.claims.back().sum = lt->claim();
    TsOracleGpu* tot = add("twist/val_eval_total", 1);
    tot->init_time(TSK_TOTAL_INC, {}, ell_n, &P.cols.has_write, nullptr, &P.cols.inc, P.cols.wa, P.r_addr, {}, kzero(),
                   nullptr, nullptr, nullptr);
    out.claims.back().sum = tot->claim();
    if (prev) {
      const TwistCols& pc = (*prev)[i];
      if (pc.ell != P.ell_addr) throw Error(1, "Twist rollover requires stable geometry");
      TsOracleGpu* pt = add("twist/rollover_prev_total", 2);
      pt->init_time(TSK_TOTAL_INC, {}, ell_n, &pc.has_write, nullptr, &pc.inc, pc.wa, P.r_addr, {}, kzero(), nullptr,
                    nullptr, nullptr);
      out.claims.back().sum = pt->claim();
    }
  }
  for (auto& c : out.claims) c.deg = c.ts->degree_bound();
  tr.append_msg_u64("twist/val_eval/batch_start", nm);
  tr.append_msg_u64("twist/val_eval/step_idx", step_idx);
  tr.append_msg_u64("twist/val_eval/claimed_sums_len", out.claims.size());
  for (size_t i = 0; i < out.claims.size(); i++) {
    tr.append_msg_u64("twist/val_eval/claim_idx", i);
    tr.append_message("twist/val_eval/claim_kind", &tags[i], 1);
    tr.append_k("twist/val_eval/claimed_sum", out.claims[i].sum);
  }
  run_batched(tr, "twist/val_eval_batch", step_idx, out.claims, ell_n, out.r_val);
  for (auto& c : out.claims) c.ts = nullptr;
  tr.append_message("twist/val_eval/batch_done", (const uint8_t*)nullptr, 0);
}

// ---------------- output binding (output_check.rs, shard.rs:1803-1862) ----------------
void ob_validate(const ObCfgH& ob, const StepMemH& sm) {
  if (ob.mem_idx >= sm.n_mem) throw Error(1, "output binding mem_idx out of range");
  if (ob.num_bits > 30) throw Error(1, "output binding: num_bits above 30");
  const MemInstH& mi = sm.mems[ob.mem_idx];
  if (mi.k != ((size_t)1 << ob.num_bits)) throw Error(1, "output binding: 2^num_bits differs from the Twist k");
  if (mi.d * mi.ell != ob.num_bits) throw Error(1, "output binding: num_bits differs from the Twist ell_addr");
  if (ob.n_claims && (!ob.addrs || !ob.values)) throw Error(1, "output binding: claims missing");
  for (size_t i = 0; i < ob.n_claims; i++) {
    if (ob.addrs[i] >> ob.num_bits) throw Error(1, "output binding: claimed address out of range");
    if (i && ob.addrs[i] <= ob.addrs[i - 1]) throw Error(1, "output binding: claims must be increasing by address");
    if (ob.values[i] >= Q) throw Error(1, "output binding: non-canonical claimed value");
  }
}

// chi(r, a) = prod_i (bit i of a ? r_i : 1 - r_i) for every a < 2^|r| (output_check.rs:179-185)
static std::vector<K2> chi_table(const std::vector<K2>& r) {
  std::vector<K2> w((size_t)1 << r.size(), kzero());
  w[0] = kone();
  for (size_t i = 0; i < r.size(); i++) {
    const size_t step = (size_t)1 << i;
    const K2 om = ksub(kone(), r[i]);
    for (size_t j = 0; j < step; j++) {
      const K2 v = w[j];
      w[j] = kmul(v, om);
      w[j + step] = kmul(v, r[i]);
    }
  }
  return w;
}

static K2 ob_challenge(Transcript& tr) { return tr.challenge_k("output_check/chal/re", "output_check/chal/im"); }

std::vector<K2> ob_sample_r_addr(Transcript& tr, const ObCfgH& ob) {
  tr.append_msg_u64("output_check/num_claims", ob.n_claims);
  for (size_t i = 0; i < ob.n_claims; i++) {
    tr.append_msg_u64("output_check/addr", ob.addrs[i]);
    tr.append_fields("output_check/value", &ob.values[i], 1);
  }
  std::vector<K2> r(ob.num_bits);
  for (size_t i = 0; i < ob.num_bits; i++) {
    tr.append_msg_u64("output_check/r_addr/idx", i);
    r[i] = ob_challenge(tr);
  }
  return r;
}

// The round polynomial of eq * mask * (val_final - val_io) equals that of the product of the
// three multilinear factors eq, mask and d = val_final - val_io (d's extension is the difference
// of the two extensions), so the four tables of the reference become three device factors.
void ob_prove(Transcript& tr, const ObCfgH& ob, std::vector<std::vector<K2>>& rounds, std::vector<K2>& r_prime) {
  if (!ob.final_state) throw Error(1, "output binding: final memory state missing");
  const std::vector<K2> r_addr = ob_sample_r_addr(tr, ob);
  const size_t k = (size_t)1 << ob.num_bits;
  std::vector<std::vector<K2>> fac(3);
  fac[0] = chi_table(r_addr);
  fac[1].assign(k, kzero());
  fac[2].resize(k);
  for (size_t a = 0; a < k; a++) {
    if (ob.final_state[a] >= Q) throw Error(1, "output binding: non-canonical final memory value");
    fac[2][a] = K2{ob.final_state[a], 0};
  }
  for (size_t i = 0; i < ob.n_claims; i++) {
    fac[1][ob.addrs[i]] = kone();
    fac[2][ob.addrs[i]] = ksub(fac[2][ob.addrs[i]], K2{ob.values[i], 0});
  }
  rounds.assign(ob.num_bits, {});
  r_prime.assign(ob.num_bits, kzero());
  if (ob.num_bits == 0) return;
  TsOracleGpu o;
  o.init_product(fac, 3);
  for (size_t rd = 0; rd < ob.num_bits; rd++) {
    std::vector<K2>& c = rounds[rd];
    o.round_coeffs(c);
    c.resize(4, kzero());
    for (auto& x : c) tr.append_k("output_check/round_coeff", x);
    r_prime[rd] = ob_challenge(tr);
    if (rd + 1 < ob.num_bits) o.fold(r_prime[rd]);
  }
}

}  // namespace ns
