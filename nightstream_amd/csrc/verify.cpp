// verify.cpp -- shard verifier over the canonical proof encoding (fold_shard_verify_impl,
// neo-fold/src/shard.rs:2482-2994), host only.
//
// It replays the Fiat-Shamir transcript with the same labels and absorb order as the prover and
// re-derives every challenge -- alpha, beta, gamma, r_cycle, the address-lane, batched-time, Ajtai
// and val-eval round challenges, the RLC rotations of both lanes and every digest -- and refuses a
// proof whose recorded challenge, claim or round polynomial disagrees.  On top of the replay it
// checks the public algebra:
//   - every claim of every batched sum-check: degree bound, p(0) + p(1) = running sum
//     (neo-reductions/src/sumcheck.rs:459-559);
//   - Pi_CCS: the claimed initial sum from the carried ME inputs, and the terminal identity
//     (optimized_engine/common.rs:688-829) from the ME outputs at (r', alpha');
//   - Route A links: the Twist address lanes start from zero and their final values are the
//     read/write time claims; bitness claims are zero; the val-lane claims sit at r_val with the
//     transcript-forked digest (memory.rs:1284-1670);
//   - Pi_RLC (api.rs:333-453): parent X, y (core and bus rows) and commitment recomputed from the
//     inputs and the re-derived rotations; Pi_DEC (api.rs:458-587): y, X and c recompose from the
//     children in base 2, children share r and the digest;
//   - the obligations: main = the last step's children, val = every step's val-lane children;
//   - output binding, when supplied (shard.rs:2566-2614, 2863-2925): the output sum-check from
//     claim zero, the inc_total claim's terminal value against the Twist time openings at r_time,
//     and eq(r_addr, r') * io_mask(r') * (val_init(r') + inc_total - val_io(r')) against the
//     output sum-check's final value.
// Not checked: the terminal identities of the memory time lanes (their oracles' final values
// against the bus openings, memory.rs:1671-2380) and the Ajtai openings of commitments (the
// verifier holds no witness).
#include <cstring>
#include <string>

#include "prover.h"
#include "route_a.h"

namespace ns {
namespace {

struct Reject : std::runtime_error {
  explicit Reject(const std::string& m) : std::runtime_error(m) {}
};
void need(bool c, const std::string& m) {
  if (!c) throw Reject(m);
}

struct Reader {
  const uint64_t* w;
  size_t n, i = 0;
  uint64_t u() {
    need(i < n, "proof truncated");
    return w[i++];
  }
  K2 k() {
    K2 x;
    x.c0 = u();
    x.c1 = u();
    need(x.c0 < Q && x.c1 < Q, "non-canonical field element");
    return x;
  }
  std::vector<K2> ks() {
    size_t c = u();
    need(c <= n, "bad length");
    std::vector<K2> v(c);
    for (auto& x : v) x = k();
    return v;
  }
  std::vector<std::vector<K2>> rounds() {
    size_t c = u();
    need(c <= n, "bad length");
    std::vector<std::vector<K2>> v(c);
    for (auto& r : v) {
      size_t l = u();
      need(l <= 64, "round polynomial too long");
      r.resize(l);
      for (auto& x : r) x = k();
    }
    return v;
  }
  void digest(uint8_t d[32]) {
    for (int q = 0; q < 4; q++) {
      uint64_t x = u();
      for (int b = 0; b < 8; b++) d[8 * q + b] = (uint8_t)(x >> (8 * b));
    }
  }
  MeHost me() {
    MeHost e;
    for (int q = 0; q < KAPPA * D; q++) e.c[q] = u();
    e.m_in = u();
    need(e.m_in <= 4096, "m_in too large");
    e.X.resize((size_t)D * e.m_in);
    for (auto& x : e.X) x = u();
    e.r = ks();
    e.t = u();
    need(e.t <= 4096, "too many y rows");
    e.y.resize(e.t * DPAD);
    for (size_t j = 0; j < e.t; j++) {
      need(u() == DPAD, "y row length");
      for (int q = 0; q < DPAD; q++) e.y[j * DPAD + q] = k();
    }
    need(u() == e.t, "y_scalars length");
    e.y_scalars.resize(e.t);
    for (auto& x : e.y_scalars) x = k();
    digest(e.fold_digest);
    need(u() == 0 && u() == 0 && u() == 0, "c_step_coords / u_offset / u_len");
    return e;
  }
};

K2 peval(const std::vector<K2>& c, K2 x) {
  K2 acc = kzero();
  for (size_t i = c.size(); i-- > 0;) acc = kadd(kmul(acc, x), c[i]);
  return acc;
}
K2 eq_lin(K2 a, K2 b) { return kadd(kmul(ksub(kone(), a), ksub(kone(), b)), kmul(a, b)); }
K2 eq_pts(const K2* p, const K2* q, size_t n) {
  K2 acc = kone();
  for (size_t i = 0; i < n; i++) acc = kmul(acc, eq_lin(p[i], q[i]));
  return acc;
}
std::vector<K2> chi_tab(const K2* bits, size_t nb) {
  std::vector<K2> w{kone()};
  for (size_t b = 0; b < nb; b++) {
    std::vector<K2> nw(w.size() * 2);
    for (size_t i = 0; i < w.size(); i++) {
      nw[i] = kmul(w[i], ksub(kone(), bits[b]));
      nw[i + w.size()] = kmul(w[i], bits[b]);
    }
    w.swap(nw);
  }
  return w;
}

bool me_eq(const MeHost& a, const MeHost& b) {
  return !memcmp(a.c, b.c, sizeof(a.c)) && a.m_in == b.m_in && a.X == b.X && a.t == b.t &&
         a.r.size() == b.r.size() && std::equal(a.r.begin(), a.r.end(), b.r.begin(), keq) &&
         std::equal(a.y.begin(), a.y.end(), b.y.begin(), keq) &&
         std::equal(a.y_scalars.begin(), a.y_scalars.end(), b.y_scalars.begin(), keq) &&
         !memcmp(a.fold_digest, b.fold_digest, 32);
}

// one claim of a batched sum-check replayed from its recorded round polynomials
struct VClaim {
  const char* label;
  size_t deg;
  K2 running;
  const std::vector<std::vector<K2>>* rounds;
};

// the verifier's side of run_batched_sumcheck_prover_ds: absorb every claim's round polynomial,
// draw the shared challenge, advance every running sum; returns the challenges
std::vector<K2> replay_batched(Transcript& tr, const char* domain, size_t inst, std::vector<VClaim>& cl,
                               size_t rounds) {
  tr.append_message("sc/domain", domain);
  tr.append_msg_u64("sc/inst_idx", inst);
  std::vector<K2> chals(rounds);
  for (auto& c : cl) need(c.rounds->size() == rounds, std::string(domain) + ": " + c.label + " round count");
  for (size_t rd = 0; rd < rounds; rd++) {
    tr.append_msg_u64("batched/round_idx", rd);
    for (size_t i = 0; i < cl.size(); i++) {
      const std::vector<K2>& co = (*cl[i].rounds)[rd];
      need(co.size() == cl[i].deg + 1, std::string(domain) + ": " + cl[i].label + " degree bound");
      need(keq(kadd(peval(co, kzero()), peval(co, kone())), cl[i].running),
           std::string(domain) + ": claim " + std::to_string(i) + " (" + cl[i].label + ") round " +
               std::to_string(rd) + ": p(0) + p(1) != running sum");
      tr.append_message("batched/claim_label", cl[i].label);
      tr.append_msg_u64("batched/claim_idx", i);
      for (auto& x : co) tr.append_k("batched/round/coeff", x);
    }
    chals[rd] = tr.challenge_k("batched/challenge/0", "batched/challenge/1");
    for (auto& c : cl) c.running = peval((*c.rounds)[rd], chals[rd]);
  }
  tr.append_message("sc/domain_end", domain);
  tr.append_msg_u64("sc/inst_idx_end", inst);
  return chals;
}

bool kvec_eq(const std::vector<K2>& a, const std::vector<K2>& b) {
  return a.size() == b.size() && std::equal(a.begin(), a.end(), b.begin(), keq);
}

// RLC public recomputation (api.rs:333-453) and DEC public checks (api.rs:458-587)
void check_rlc_dec(const std::vector<MeHost>& ins, const std::vector<uint64_t>& rhos, const std::vector<int8_t>& rho_a,
                   const MeHost& parent, const std::vector<MeHost>& ch, size_t k_dec, const std::string& where) {
  const size_t kt = ins.size(), m_in = ins[0].m_in, t = parent.t;
  need(kvec_eq(parent.r, ins[0].r), where + ": parent r");
  if (kt == 1) {
    need(parent.t == ins[0].t && kvec_eq(parent.y, ins[0].y) && parent.X == ins[0].X &&
             !memcmp(parent.c, ins[0].c, sizeof(parent.c)),
         where + ": k = 1 parent differs from its input");
  } else {
    for (auto& e : ins) need(e.t == t && e.m_in == m_in, where + ": input shapes");
    std::vector<K2> y(t * DPAD, kzero());
    std::vector<uint64_t> X((size_t)D * m_in, 0), c(KAPPA * D, 0);
    for (size_t i = 0; i < kt; i++) {
      const uint64_t* R = &rhos[i * D * D];
      for (size_t j = 0; j < t; j++)
        for (int rr = 0; rr < D; rr++) {
          K2 acc = kzero();
          for (int kk = 0; kk < D; kk++)
            if (R[rr * D + kk]) acc = kadd(acc, kscale(ins[i].y[j * DPAD + kk], R[rr * D + kk]));
          y[j * DPAD + rr] = kadd(y[j * DPAD + rr], acc);
        }
      for (int rr = 0; rr < D; rr++)
        for (int kk = 0; kk < D; kk++)
          if (R[rr * D + kk])
            for (size_t q = 0; q < m_in; q++)
              X[rr * m_in + q] = fadd(X[rr * m_in + q], fmul(R[rr * D + kk], ins[i].X[kk * m_in + q]));
      // c += rot(a_i) c_i column by column: a(X) * c(X) mod Phi_81
      for (int col = 0; col < KAPPA; col++) {
        uint64_t tmp[2 * D] = {0};
        for (int a = 0; a < D; a++) {
          const int64_t av = rho_a[i * D + a];
          if (!av) continue;
          for (int b = 0; b < D; b++) tmp[a + b] = fadd(tmp[a + b], fmul(from_i64(av), ins[i].c[col * D + b]));
        }
        for (int q = 2 * D - 2; q >= D; q--) {
          const uint64_t v = tmp[q];
          tmp[q] = 0;
          tmp[q - D] = fsub(tmp[q - D], v);
          tmp[q - 27] = fsub(tmp[q - 27], v);
        }
        for (int q = 0; q < D; q++) c[col * D + q] = fadd(c[col * D + q], tmp[q]);
      }
    }
    need(kvec_eq(y, parent.y), where + ": RLC y mismatch");
    need(X == parent.X, where + ": RLC X mismatch");
    need(!memcmp(c.data(), parent.c, sizeof(parent.c)), where + ": RLC commitment mismatch");
  }
  // y_scalars of parent and children
  auto scal_ok = [&](const MeHost& e) {
    for (size_t j = 0; j < e.t; j++) {
      K2 sc = kzero();
      uint64_t pw = 1;
      for (int rho = 0; rho < D; rho++) {
        sc = kadd(sc, kscale(e.y[j * DPAD + rho], pw));
        pw = fadd(pw, pw);
      }
      if (!keq(sc, e.y_scalars[j])) return false;
    }
    return true;
  };
  need(scal_ok(parent), where + ": parent y_scalars");
  need(ch.size() == k_dec, where + ": child count");
  for (auto& e : ch) {
    need(e.t == t && e.m_in == m_in && kvec_eq(e.r, parent.r), where + ": child shape / r");
    need(!memcmp(e.fold_digest, parent.fold_digest, 32), where + ": child digest");
    need(scal_ok(e), where + ": child y_scalars");
  }
  for (size_t q = 0; q < t * DPAD; q++) {
    K2 lhs = kzero();
    uint64_t pw = 1;
    for (size_t i = 0; i < k_dec; i++) {
      lhs = kadd(lhs, kscale(ch[i].y[q], pw));
      pw = fadd(pw, pw);
    }
    need(keq(lhs, parent.y[q]), where + ": DEC y recomposition");
  }
  for (size_t q = 0; q < (size_t)D * m_in; q++) {
    uint64_t lhs = 0, pw = 1;
    for (size_t i = 0; i < k_dec; i++) {
      lhs = fadd(lhs, fmul(ch[i].X[q], pw));
      pw = fadd(pw, pw);
    }
    need(lhs == parent.X[q], where + ": DEC X recomposition");
  }
  for (int q = 0; q < KAPPA * D; q++) {
    uint64_t lhs = 0, pw = 1;
    for (size_t i = 0; i < k_dec; i++) {
      lhs = fadd(lhs, fmul(ch[i].c[q], pw));
      pw = fadd(pw, pw);
    }
    need(lhs == parent.c[q], where + ": DEC commitment recomposition");
  }
}

// rhs_terminal_identity_paper_exact (optimized_engine/common.rs:688-829) from the ME outputs
K2 terminal_rhs(const Ccs& s, const std::vector<MeHost>& out, const K2* r_prime, size_t ell_n, const K2* a_prime,
                const StepChallenges& ch, size_t ell_d, const std::vector<K2>* r_inputs) {
  const size_t t = s.t, k = out.size();
  const K2* beta_a = ch.beta.data();
  const K2* beta_r = ch.beta.data() + ell_d;
  const K2 eq_beta = kmul(eq_pts(a_prime, beta_a, ell_d), eq_pts(r_prime, beta_r, ell_n));
  const K2 eq_ar = r_inputs ? kmul(eq_pts(a_prime, ch.alpha.data(), ell_d), eq_pts(r_prime, r_inputs->data(), ell_n))
                            : kzero();
  std::vector<K2> mv(t);
  for (size_t j = 0; j < t; j++) {
    K2 acc = kzero();
    uint64_t pw = 1;
    for (int rho = 0; rho < D; rho++) {
      acc = kadd(acc, kscale(out[0].y[j * DPAD + rho], pw));
      pw = fadd(pw, pw);
    }
    mv[j] = acc;
  }
  K2 Fp = kzero();
  for (size_t tt = 0; tt < s.term_coeff.size(); tt++) {
    K2 term{s.term_coeff[tt], 0};
    for (size_t j = 0; j < t; j++)
      for (uint32_t e = 0; e < s.term_exps[tt][j]; e++) term = kmul(term, mv[j]);
    Fp = kadd(Fp, term);
  }
  const std::vector<K2> ca = chi_tab(a_prime, ell_d);
  K2 nc = kzero(), g = ch.gamma;
  for (auto& e : out) {
    K2 ye = kzero();
    for (int rho = 0; rho < DPAD; rho++) ye = kadd(ye, kmul(e.y[rho], ca[rho]));
    nc = kadd(nc, kmul(g, kmul(ye, ksub(kmul(ye, ye), kone()))));
    g = kmul(g, ch.gamma);
  }
  K2 gk = kone();
  for (size_t i = 0; i < k; i++) gk = kmul(gk, ch.gamma);
  K2 ev = kzero();
  if (r_inputs && k >= 2) {
    for (size_t j = 0; j < t; j++)
      for (size_t i = 1; i < k; i++) {
        K2 ye = kzero();
        for (int rho = 0; rho < DPAD; rho++) ye = kadd(ye, kmul(out[i].y[j * DPAD + rho], ca[rho]));
        K2 w = kone();
        for (size_t q = 0; q < i; q++) w = kmul(w, ch.gamma);
        for (size_t q = 0; q < j; q++) w = kmul(w, gk);
        ev = kadd(ev, kmul(w, ye));
      }
  }
  return kadd(kmul(eq_beta, kadd(Fp, nc)), kmul(eq_ar, kmul(gk, ev)));
}

}  // namespace

// Returns normally when the proof is accepted; throws Error(7, reason) otherwise.
void verify_shard(const Ccs& s, const uint64_t* const* x_steps, size_t nsteps, size_t m_in, const StepMemH* mem,
                  const uint64_t* mcs_c, const uint8_t* proof, size_t proof_len, const ObCfgH* ob) {
  if (ob) {
    if (nsteps == 0 || !mem) throw Error(1, "output binding requires >= 1 step with memory");
    ob_validate(*ob, mem[nsteps - 1]);
  }
  try {
    need(proof_len % 8 == 0, "proof length is not a multiple of 8");
    Reader R{(const uint64_t*)proof, proof_len / 8};
    need(R.u() == 0x314652504f52504eULL, "not a canonical shard proof");
    need(R.u() == nsteps, "step count");
    const Params params = Params::goldilocks_auto_r1cs_ccs(s.n);
    const size_t t = s.t, k_dec = params.k_rho;
    size_t ell_n = 0;
    while (((size_t)1 << ell_n) < s.n) ell_n++;
    const size_t ell_d = 6, ell = ell_d + ell_n;
    const size_t d_sc = std::max<size_t>(s.f_degree() + 1, std::max<size_t>(2, 2 * params.b + 2));
    uint32_t smin;
    int32_t slack;
    need(params.extension_check((uint32_t)ell, (uint32_t)d_sc, &smin, &slack) && slack >= 0, "extension policy");
    const StepMemH no_mem{};
    BusLayout L;
    if (mem) {
      L = infer_bus_layout(mem, nsteps, s.n, s.m, m_in);
    } else {
      L.m = s.m;
      L.m_in = m_in;
    }
    const size_t t_tot = t + L.bus_cols;
    Transcript tr("neo.fold/session");
    DigestCache digests;
    std::vector<MeHost> acc, val_all;
    for (size_t idx = 0; idx < nsteps; idx++) {
      const std::string at = "step " + std::to_string(idx);
      const StepMemH& sm = mem ? mem[idx] : no_mem;
      const size_t kt = acc.size() + 1;
      // ---- parse the step ----
      need(R.u() == kt, at + ": ccs_out count");
      std::vector<MeHost> ccs_out(kt);
      for (auto& e : ccs_out) e = R.me();
      std::vector<std::vector<K2>> sc_rounds = R.rounds();
      const K2 sc_init = R.k();
      const std::vector<K2> sc_chals = R.ks();
      const std::vector<K2> p_alpha = R.ks(), p_beta_a = R.ks(), p_beta_r = R.ks();
      const K2 p_gamma = R.k(), sc_final = R.k();
      uint8_t hd[32];
      R.digest(hd);
      auto read_lane = [&](std::vector<uint64_t>& rhos, MeHost& par, std::vector<MeHost>& ch) {
        const size_t k = R.u();
        need(k <= 64, "rho count");
        rhos.resize(k * D * D);
        for (auto& x : rhos) x = R.u();
        par = R.me();
        const size_t nc = R.u();
        need(nc == k_dec, "child count");
        ch.resize(nc);
        for (auto& e : ch) e = R.me();
        return k;
      };
      std::vector<uint64_t> rhos;
      MeHost parent;
      std::vector<MeHost> children;
      const size_t n_rho = read_lane(rhos, parent, children);
      // mem
      const size_t nvc = R.u();
      need(nvc <= 2, at + ": CPU ME claims at r_val");
      std::vector<MeHost> vclaims(nvc);
      for (auto& e : vclaims) e = R.me();
      const std::vector<K2> sh_sums = R.ks();
      const uint64_t sh_mask = R.u();
      const size_t n_act = R.u();
      need(n_act <= 64, "shout lanes");
      std::vector<std::vector<std::vector<K2>>> sh_polys(n_act);
      for (auto& p : sh_polys) p = R.rounds();
      const std::vector<K2> sh_r = R.ks();
      const size_t nproofs = R.u();
      need(nproofs == sm.n_lut + sm.n_mem, at + ": memory proof count");
      struct TwV {
        std::vector<K2> sums, r_addr;
        std::vector<std::vector<std::vector<K2>>> polys;
        bool has_ve = false, has_prev = false;
        K2 lt_sum, tot_sum, prev_sum;
        std::vector<std::vector<K2>> lt_r, tot_r, prev_r;
      };
      std::vector<TwV> tws;
      for (size_t q = 0; q < nproofs; q++) {
        const uint64_t tag = R.u();
        TwV tv;
        tv.sums = R.ks();
        const size_t np = R.u();
        need(np <= 2, "address lanes");
        for (size_t c = 0; c < np; c++) tv.polys.push_back(R.rounds());
        tv.r_addr = R.ks();
        if (q < sm.n_lut) {
          need(tag == 1 && tv.sums.empty() && tv.polys.empty() && tv.r_addr.empty(), at + ": Shout proof entry");
          continue;
        }
        need(tag == 0, at + ": Twist proof entry");
        tv.has_ve = R.u() != 0;
        if (tv.has_ve) {
          tv.lt_sum = R.k();
          tv.lt_r = R.rounds();
          tv.tot_sum = R.k();
          tv.tot_r = R.rounds();
          const bool ps = R.u() != 0;
          if (ps) tv.prev_sum = R.k();
          const bool pr = R.u() != 0;
          if (pr) tv.prev_r = R.rounds();
          need(ps == pr, at + ": rollover claim without rounds");
          tv.has_prev = ps;
        }
        tws.push_back(std::move(tv));
      }
      const std::vector<K2> bt_sums = R.ks();
      const size_t nd = R.u();
      need(nd == bt_sums.size(), at + ": degree bound count");
      std::vector<size_t> bt_degs(nd);
      for (auto& d : bt_degs) d = R.u();
      const size_t nl = R.u();
      need(nl == bt_sums.size(), at + ": label count");
      std::vector<std::string> bt_labels(nl);
      for (auto& lb : bt_labels) {
        const size_t len = R.u();
        need(len <= 64, "label length");
        for (size_t q = 0; q < len; q++) lb.push_back((char)R.u());
      }
      const size_t npl = R.u();
      need(npl == bt_sums.size(), at + ": batched round polynomial count");
      std::vector<std::vector<std::vector<K2>>> bt_polys(npl);
      for (auto& p : bt_polys) p = R.rounds();
      const bool has_val = R.u() != 0;
      std::vector<uint64_t> vrhos;
      MeHost vparent;
      std::vector<MeHost> vchildren;
      if (has_val) read_lane(vrhos, vparent, vchildren);
      // ShardProof.output_proof follows the last step (shard_proof_types.rs:165-171)
      const bool include_ob = ob && idx + 1 == nsteps;
      std::vector<std::vector<K2>> ob_polys;
      if (idx + 1 == nsteps) {
        const bool has_ob = R.u() != 0;
        need(has_ob == (ob != nullptr), has_ob ? "proof has output binding, but no binding was supplied"
                                               : "output binding supplied, but the proof has none");
        if (has_ob) ob_polys = R.rounds();
      }

      // ---- replay ----
      absorb_step_memory(tr, sm, digests);
      // output binding (shard.rs:2566-2614): the output sum-check from claim zero, its state kept
      // for the final check below
      std::vector<K2> ob_r_prime, ob_r_addr;
      K2 ob_final = kzero();
      if (include_ob) {
        tr.append_msg_u64("shard/output_binding_start", idx);
        const uint64_t mi = ob->mem_idx, nb = ob->num_bits;
        tr.append_u64s("output_binding/mem_idx", &mi, 1);
        tr.append_u64s("output_binding/num_bits", &nb, 1);
        ob_r_addr = ob_sample_r_addr(tr, *ob);
        need(ob_polys.size() == ob->num_bits, at + ": output sum-check round count");
        K2 run = kzero();
        for (size_t rd = 0; rd < ob_polys.size(); rd++) {
          const std::vector<K2>& co = ob_polys[rd];
          need(co.size() == 4, at + ": output sum-check round degree");
          need(keq(kadd(peval(co, kzero()), peval(co, kone())), run), at + ": output sum-check round p(0) + p(1)");
          for (auto& x : co) tr.append_k("output_check/round_coeff", x);
          const K2 r = tr.challenge_k("output_check/chal/re", "output_check/chal/im");
          ob_r_prime.push_back(r);
          run = peval(co, r);
        }
        ob_final = run;
      }
      if (mcs_c) need(!memcmp(ccs_out[0].c, mcs_c + idx * KAPPA * D, sizeof(ccs_out[0].c)), at + ": MCS commitment");
      bind_step_header(tr, s, params, ell, d_sc, slack, x_steps[idx], m_in, ccs_out[0].c, acc);
      const StepChallenges ch = sample_challenges(tr, ell_d, ell);
      need(kvec_eq(p_alpha, ch.alpha), at + ": alpha does not follow from the transcript");
      need(kvec_eq(p_beta_a, std::vector<K2>(ch.beta.begin(), ch.beta.begin() + ell_d)) &&
               kvec_eq(p_beta_r, std::vector<K2>(ch.beta.begin() + ell_d, ch.beta.end())),
           at + ": beta does not follow from the transcript");
      need(keq(p_gamma, ch.gamma), at + ": gamma does not follow from the transcript");
      const K2 T = claimed_initial_sum(s, ch, acc, ell_d);
      need(keq(T, sc_init), at + ": claimed initial sum");
      tr.append_k("sumcheck/initial_sum", T);
      std::vector<K2> r_cycle(ell_n);
      for (size_t i = 0; i < ell_n; i++) {
        tr.append_msg_u64("route_a/r_cycle", i);
        r_cycle[i] = tr.challenge_k("route_a/cycle/0", "route_a/cycle/1");
      }
      // Shout address lanes (memory.rs:430-621 / verify_shout_addr_pre_time 623-816)
      if (sm.n_lut) {
        need(sh_sums.size() == sm.n_lut, at + ": Shout claimed sums");
        const size_t ell_addr = sm.luts[0].d * sm.luts[0].ell;
        tr.append_msg_u64("shout/addr_pre_time/step_idx", idx);
        tr.append_msg_u64("shout/addr_pre_time/claimed_sums", sm.n_lut);
        for (size_t i = 0; i < sm.n_lut; i++) {
          tr.append_message("addr_batch/label", "shout/addr_pre");
          tr.append_msg_u64("addr_batch/idx", i);
          tr.append_k("addr_batch/claimed_sum", sh_sums[i]);
        }
        need(sm.n_lut >= 64 || (sh_mask >> sm.n_lut) == 0, at + ": Shout active mask out of range");
        if (sh_mask == 0) {
          need(n_act == 0, at + ": inactive Shout lanes with rounds");
          tr.append_msg_u64("shout/addr_pre_time/no_sumcheck", idx);
          std::vector<K2> r(ell_addr);
          for (size_t i = 0; i < ell_addr; i++) {
            tr.append_msg_u64("shout/addr_pre_time/no_sumcheck/r_addr", i);
            r[i] = tr.challenge_k("shout/addr_pre_time/no_sumcheck/r_addr/0", "shout/addr_pre_time/no_sumcheck/r_addr/1");
          }
          need(kvec_eq(r, sh_r), at + ": Shout r_addr");
        } else {
          std::vector<VClaim> cl;
          size_t a = 0;
          for (size_t i = 0; i < sm.n_lut; i++)
            if ((sh_mask >> i) & 1) {
              need(a < n_act, at + ": Shout active lanes");
              cl.push_back(VClaim{"shout/addr_pre", 2, sh_sums[i], &sh_polys[a++]});
            } else {
              need(kiszero(sh_sums[i]), at + ": inactive Shout lane with a nonzero claim");
            }
          need(a == n_act, at + ": Shout active lanes");
          const std::vector<K2> r = replay_batched(tr, "shout/addr_pre_time", idx, cl, ell_addr);
          need(kvec_eq(r, sh_r), at + ": Shout r_addr does not follow from the transcript");
        }
      } else {
        need(sh_sums.empty() && sh_mask == 0 && n_act == 0 && sh_r.empty(), at + ": Shout pre-time must be empty");
      }
      // Twist address lanes (memory.rs:232-428 / verify_twist_addr_pre_time 818-903)
      std::vector<K2> read_claims, write_claims;
      for (size_t i = 0; i < sm.n_mem; i++) {
        TwV& tv = tws[i];
        const size_t e = sm.mems[i].d * sm.mems[i].ell;
        need(tv.sums.size() == 2 && kiszero(tv.sums[0]) && kiszero(tv.sums[1]), at + ": Twist address claims");
        need(tv.polys.size() == 2, at + ": Twist address lanes");
        tr.append_msg_u64("twist/addr_pre_time/claim_idx", i);
        tr.append_msg_u64("twist/addr_pre_time/claimed_sums", 2);
        const char* lbl[2] = {"twist/read_addr_pre", "twist/write_addr_pre"};
        for (int c = 0; c < 2; c++) {
          tr.append_message("addr_batch/label", lbl[c]);
          tr.append_msg_u64("addr_batch/idx", c);
          tr.append_k("addr_batch/claimed_sum", kzero());
        }
        std::vector<VClaim> cl{{lbl[0], 2, kzero(), &tv.polys[0]}, {lbl[1], 2, kzero(), &tv.polys[1]}};
        const std::vector<K2> r = replay_batched(tr, "twist/addr_pre_time", i, cl, e);
        need(kvec_eq(r, tv.r_addr), at + ": Twist r_addr does not follow from the transcript");
        read_claims.push_back(cl[0].running);
        write_claims.push_back(cl[1].running);
      }
      // the batched time sum-check (route_a_time.rs:28-152 / 159-272)
      std::vector<std::string> want_lb{"ccs/time"};
      std::vector<size_t> want_deg{d_sc};
      for (size_t i = 0; i < sm.n_lut; i++) {
        const size_t e = sm.luts[i].d * sm.luts[i].ell;
        want_lb.insert(want_lb.end(), {"shout/value", "shout/adapter", "shout/bitness"});
        want_deg.insert(want_deg.end(), {3, e + 2, 3});
      }
      for (size_t i = 0; i < sm.n_mem; i++) {
        const size_t e = sm.mems[i].d * sm.mems[i].ell;
        want_lb.insert(want_lb.end(), {"twist/read_check", "twist/write_check", "twist/bitness"});
        want_deg.insert(want_deg.end(), {e + 3, e + 3, 3});
      }
      if (include_ob) {
        want_lb.push_back("output_binding/inc_total");
        want_deg.push_back(2 + ob->num_bits);
      }
      need(bt_labels == want_lb, at + ": batched time labels drift");
      need(bt_degs == want_deg, at + ": batched time degree bounds drift");
      need(keq(bt_sums[0], T), at + ": CCS time claim");
      for (size_t i = 0; i < sm.n_lut; i++) need(kiszero(bt_sums[1 + 3 * i + 2]), at + ": Shout bitness claim");
      for (size_t i = 0; i < sm.n_mem; i++) {
        const size_t b = 1 + 3 * sm.n_lut + 3 * i;
        need(keq(bt_sums[b], read_claims[i]) && keq(bt_sums[b + 1], write_claims[i]),
             at + ": Twist time claims are not the address lanes' final values");
        need(kiszero(bt_sums[b + 2]), at + ": Twist bitness claim");
      }
      std::vector<VClaim> tcl;
      std::vector<RaClaim> bind(nl);
      for (size_t i = 0; i < nl; i++) {
        tcl.push_back(VClaim{want_lb[i].c_str(), bt_degs[i], bt_sums[i], &bt_polys[i]});
        bind[i].label = want_lb[i].c_str();
        bind[i].deg = bt_degs[i];
        bind[i].sum = bt_sums[i];
        bind[i].dyn = want_lb[i].find("bitness") == std::string::npos;
      }
      bind_dynamic_claims(tr, bind);
      const std::vector<K2> r_time = replay_batched(tr, "shard/batched_time", idx, tcl, ell_n);
      need(sc_rounds.size() == ell_n + ell_d && sc_chals.size() == ell_n + ell_d, at + ": Pi_CCS round count");
      for (size_t rd = 0; rd < ell_n; rd++)
        need(kvec_eq(sc_rounds[rd], bt_polys[0][rd]), at + ": Pi_CCS rounds differ from the batched CCS claim");
      need(kvec_eq(std::vector<K2>(sc_chals.begin(), sc_chals.begin() + ell_n), r_time),
           at + ": r_time does not follow from the transcript");
      // Ajtai rounds (run_sumcheck_prover_ds "ccs/ajtai")
      K2 running = tcl[0].running;
      tr.append_message("sc/domain", "ccs/ajtai");
      tr.append_msg_u64("sc/inst_idx", idx);
      tr.append_k("sc/claimed_sum", running);
      std::vector<K2> a_prime(ell_d);
      for (size_t rd = 0; rd < ell_d; rd++) {
        const std::vector<K2>& co = sc_rounds[ell_n + rd];
        need(co.size() == d_sc + 1, at + ": Ajtai round degree bound");
        need(keq(kadd(peval(co, kzero()), peval(co, kone())), running), at + ": Ajtai round p(0) + p(1)");
        for (auto& x : co) tr.append_k("sumcheck/round/coeff", x);
        a_prime[rd] = tr.challenge_k("sumcheck/challenge/0", "sumcheck/challenge/1");
        need(keq(a_prime[rd], sc_chals[ell_n + rd]), at + ": Ajtai challenge does not follow from the transcript");
        running = peval(co, a_prime[rd]);
      }
      tr.append_message("sc/domain_end", "ccs/ajtai");
      tr.append_msg_u64("sc/inst_idx_end", idx);
      need(keq(running, sc_final), at + ": sumcheck_final");
      uint8_t fd[32];
      tr.digest32(fd);
      need(!memcmp(fd, hd, 32), at + ": header digest does not follow from the transcript");
      // Pi_CCS outputs and the terminal identity
      for (size_t i = 0; i < kt; i++) {
        const MeHost& e = ccs_out[i];
        need(e.t == t_tot && e.m_in == m_in && kvec_eq(e.r, r_time), at + ": ccs_out shape / r");
        need(!memcmp(e.fold_digest, fd, 32), at + ": ccs_out digest");
        if (i >= 1)
          need(!memcmp(e.c, acc[i - 1].c, sizeof(e.c)) && e.X == acc[i - 1].X, at + ": carried c / X");
      }
      if (include_ob) {
        // inc_total's terminal value from the Twist time openings at r_time, then
        // val_final(r') = val_init(r') + inc_total and the output sum-check's final value
        // (shard.rs:2863-2925, output_binding.rs:63-81)
        const BusLayout::Tw& tw = L.tw[ob->mem_idx];
        const std::vector<K2>& ys = ccs_out[0].y_scalars;
        need(ys.size() == t_tot && tw.ell == ob->num_bits, at + ": Twist time openings");
        K2 inc_term = kmul(ys[t + tw.has_write], ys[t + tw.inc]);
        for (size_t b = 0; b < tw.ell; b++) inc_term = kmul(inc_term, eq_lin(ys[t + tw.wa0 + b], ob_r_prime[b]));
        need(keq(tcl.back().running, inc_term), at + ": inc_total terminal mismatch");
        const size_t nb = ob->num_bits;
        K2 mask = kzero(), vio = kzero();
        for (size_t q = 0; q < ob->n_claims; q++) {
          K2 chi = kone();
          for (size_t b = 0; b < nb; b++)
            chi = kmul(chi, ((ob->addrs[q] >> b) & 1) ? ob_r_prime[b] : ksub(kone(), ob_r_prime[b]));
          mask = kadd(mask, chi);
          vio = kadd(vio, kscale(chi, ob->values[q]));
        }
        const K2 eq_eval = eq_pts(ob_r_addr.data(), ob_r_prime.data(), nb);
        const K2 val_final = kadd(eval_init(sm.mems[ob->mem_idx], ob_r_prime), bt_sums.back());
        need(keq(kmul(kmul(eq_eval, mask), ksub(val_final, vio)), ob_final), at + ": output binding final check failed");
      }
      const K2 rhs = terminal_rhs(s, ccs_out, r_time.data(), ell_n, a_prime.data(), ch, ell_d,
                                  acc.empty() ? nullptr : &acc[0].r);
      need(keq(rhs, running), at + ": terminal identity failed");
      // finalize_route_a_memory (val-eval batch, ME claims at r_val)
      std::vector<K2> r_val;
      if (sm.n_mem) {
        const bool has_prev = idx > 0;
        std::vector<VClaim> vcl;
        std::vector<uint8_t> tags;
        std::vector<K2> sums;
        for (size_t i = 0; i < sm.n_mem; i++) {
          TwV& tv = tws[i];
          const size_t e = sm.mems[i].d * sm.mems[i].ell;
          need(tv.has_ve && tv.has_prev == has_prev, at + ": val-eval proof shape");
          vcl.push_back(VClaim{"twist/val_eval_lt", e + 3, tv.lt_sum, &tv.lt_r});
          vcl.push_back(VClaim{"twist/val_eval_total", e + 2, tv.tot_sum, &tv.tot_r});
          tags.insert(tags.end(), {0, 1});
          if (has_prev) {
            vcl.push_back(VClaim{"twist/rollover_prev_total", e + 2, tv.prev_sum, &tv.prev_r});
            tags.push_back(2);
          }
        }
        tr.append_msg_u64("twist/val_eval/batch_start", sm.n_mem);
        tr.append_msg_u64("twist/val_eval/step_idx", idx);
        tr.append_msg_u64("twist/val_eval/claimed_sums_len", vcl.size());
        for (size_t i = 0; i < vcl.size(); i++) {
          tr.append_msg_u64("twist/val_eval/claim_idx", i);
          tr.append_message("twist/val_eval/claim_kind", &tags[i], 1);
          tr.append_k("twist/val_eval/claimed_sum", vcl[i].running);
        }
        r_val = replay_batched(tr, "twist/val_eval_batch", idx, vcl, ell_n);
        tr.append_message("twist/val_eval/batch_done", (const uint8_t*)nullptr, 0);
        Transcript fork = tr;
        fork.append_message("fork", "cpu_bus/me_digest_val");
        uint8_t vdig[32];
        fork.digest32(vdig);
        need(nvc == (has_prev ? 2u : 1u), at + ": CPU ME claims at r_val");
        for (size_t q = 0; q < nvc; q++) {
          const MeHost& e = vclaims[q];
          need(e.t == t_tot && kvec_eq(e.r, r_val) && !memcmp(e.fold_digest, vdig, 32), at + ": val claim point / digest");
          if (mcs_c) need(!memcmp(e.c, mcs_c + (idx - q) * KAPPA * D, sizeof(e.c)), at + ": val claim commitment");
        }
      } else {
        need(nvc == 0 && !has_val, at + ": val lane without Twist instances");
      }
      // main lane, then the val lane
      {
        bind_rlc_inputs(tr, "main", idx, ccs_out);
        std::vector<int8_t> ra;
        std::vector<uint64_t> rr;
        sample_rot_rhos(tr, kt, ra, rr);
        need(n_rho == kt && rr == rhos, at + ": RLC rotations do not follow from the transcript");
        check_rlc_dec(ccs_out, rhos, ra, parent, children, k_dec, at + " (main)");
      }
      if (nvc) {
        need(has_val, at + ": val lane missing");
        tr.append_msg_u64("fold/val_lane_start", idx);
        bind_rlc_inputs(tr, "val", idx, vclaims);
        std::vector<int8_t> ra;
        std::vector<uint64_t> rr;
        sample_rot_rhos(tr, nvc, ra, rr);
        need(rr == vrhos, at + ": val-lane rotations do not follow from the transcript");
        check_rlc_dec(vclaims, vrhos, ra, vparent, vchildren, k_dec, at + " (val)");
        for (auto& e : vchildren) val_all.push_back(e);
      }
      tr.append_msg_u64("fold/step_done", idx);
      acc = std::move(children);
    }
    if (nsteps == 0) need(R.u() == 0, "output binding without steps");
    need(R.u() == acc.size(), "main obligations count");
    for (auto& e : acc) need(me_eq(R.me(), e), "main obligations differ from the last children");
    need(R.u() == val_all.size(), "val obligations count");
    for (auto& e : val_all) need(me_eq(R.me(), e), "val obligations differ from the val-lane children");
    need(R.i == R.n, "trailing words after the proof");
  } catch (const Reject& e) {
    throw Error(7, std::string("proof rejected: ") + e.what());
  }
}

}  // namespace ns
