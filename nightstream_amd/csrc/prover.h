// prover.h -- host orchestration of the MI355X Neo folding prover: device-resident CCS,
// Ajtai PP, witnesses and Pi_CCS tables, with the Fiat-Shamir transcript on the host.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.h"
#include "transcript.h"

namespace ns {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Owned device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n);
  void ensure(size_t n) {
    if (n > bytes) alloc(n);
  }
  void release();
  template <class T>
  T* as() const { return (T*)p; }
};

struct Params {
  uint64_t q = Q;
  uint32_t eta = 81, d = 54, kappa = 16;
  uint64_t m = 1ull << 24;
  uint32_t b = 2, k_rho = 12;
  uint64_t B = 4096;
  uint32_t T = 216, s = 2, lambda = 127;
  static Params goldilocks_auto_r1cs_ccs(size_t n_rows);  // neo-params/src/lib.rs:186-222
  bool extension_check(uint32_t ell, uint32_t d_sc, uint32_t* s_min, int32_t* slack) const;
};

struct HostCsr {
  std::vector<uint64_t> ptr;
  std::vector<uint32_t> col;
  std::vector<uint64_t> val;
};

// CCS structure: identity-first [I, M_1 .. M_{t-1}] with the sparse polynomial f.
struct Ccs {
  size_t n = 0, m = 0, t = 0;
  std::vector<HostCsr> host;     // host[j] for j >= 1 (index 0 = identity, empty)
  std::vector<DevSparse> dev;    // dev[j]
  std::vector<DevBuf> bufs;
  // f terms
  std::vector<uint64_t> term_coeff;
  std::vector<std::vector<uint32_t>> term_exps;
  FPoly fpoly;                   // compiled over used variables
  std::vector<int> fvar_mat;     // CCS matrix index of each f-var table
  uint64_t digest[4];
  std::mutex dev_mu;
  bool on_device = false;
  // host structure and digest only; the matrices reach HBM on the first ensure_device()
  static std::unique_ptr<Ccs> from_r1cs(size_t n, size_t m, const HostCsr& a, const HostCsr& b, const HostCsr& c);
  void ensure_device();
  size_t f_degree() const;
};

// Ajtai public parameters resident in HBM as int8 limb planes (see ajtai.hip).
struct Pp {
  size_t kappa = KAPPA, m = 0;
  DevBuf limbs;   // kappa*8*54 rows x m
  DevBuf coeffs;  // optional u64 [kappa][m][54]
  static std::unique_ptr<Pp> setup_seeded(const uint8_t seed[32], size_t m, bool keep_coeffs, hipStream_t s);
  static std::unique_ptr<Pp> from_coeffs(const uint64_t* host_coeffs, size_t m, bool keep_coeffs, hipStream_t s);
};

struct MeHost {
  uint64_t c[KAPPA * D];
  size_t m_in = 0;
  std::vector<uint64_t> X;   // D x m_in
  std::vector<K2> r;
  size_t t = 0;
  std::vector<K2> y;         // t x 64
  std::vector<K2> y_scalars; // t
  uint8_t fold_digest[32];
};

// Device buffers of a Pi_CCS oracle, reused across fold steps (grow-only).
// Pinned host memory the device writes directly (round coefficients): no copy, and the host
// reads it after the stream synchronizes.
struct HostMailbox {
  void* p = nullptr;
  HostMailbox() = default;
  HostMailbox(const HostMailbox&) = delete;
  HostMailbox& operator=(const HostMailbox&) = delete;
  ~HostMailbox() {
    if (p) (void)hipHostFree(p);
  }
  template <class T>
  T* get(size_t n) {
    // coherent: kernels publish round coefficients and a ready word here while they run
    if (!p) NS_CHECK(hipHostMalloc(&p, std::max<size_t>(n * sizeof(T), 256), hipHostMallocCoherent));
    return (T*)p;
  }
};

// Pinned staging for device-to-host copies, grown on demand (at least 1 MiB).
struct PinnedStage {
  void* p = nullptr;
  size_t bytes = 0;
  PinnedStage() = default;
  PinnedStage(const PinnedStage&) = delete;
  PinnedStage& operator=(const PinnedStage&) = delete;
  ~PinnedStage();
  void* get(size_t n);
};

struct OracleWs {
  DevBuf Zptrs, sc[2], nc[2], wg, wl, coeffs, ws, z1, bits, Wd, Sa, chi, ydev, cs_ws, pc, gt;
  HostMailbox mbox;   // 8 K2: the round polynomial's coefficients; the ready word at byte 192
  uint64_t round_seq = 0;  // last ready-word value handed to a round launch
  PinnedStage stage;  // the context's device-to-host staging
  // the stream order between this context's stream and the shared-round stream (see FusedHub)
  hipEvent_t hub_in = nullptr, hub_out = nullptr;
  OracleWs() = default;
  OracleWs(const OracleWs&) = delete;
  OracleWs& operator=(const OracleWs&) = delete;
  ~OracleWs() {
    if (hub_in) (void)hipEventDestroy(hub_in);
    if (hub_out) (void)hipEventDestroy(hub_out);
  }
};

// Shared launches for the fused rounds of concurrent sessions on one GPU (flat combining).  A
// session queues its round as a FusedJob; the first session to find no combiner at work becomes
// the combiner and issues every queued round -- a lone job as its own launch pair, several as one
// ragged fused launch and one ragged partial-sum launch over a device job array -- on the hub's
// stream, until the queue is empty.  Rounds that arrive while a launch is being issued ride on
// the next one: an uncontended session pays nothing, contended sessions pay one launch pair per
// group of rounds instead of one per round.  Every session then waits on its own ready word.
// Enabled process-wide by set_shared_rounds (prover pools with several contexts turn it on).
void set_shared_rounds(bool on);
bool shared_rounds();
struct FusedHubStats {
  uint64_t launches = 0, jobs = 0;
};
FusedHubStats fused_hub_stats();

struct StepMemH;
struct ObCfgH;
struct BusLayout;

// Fiat-Shamir steps shared by the prover and the verifier (same labels and absorb order)
struct StepChallenges {
  std::vector<K2> alpha, beta;  // beta = beta_a (ell_d) | beta_r (ell_n)
  K2 gamma{0, 0};
};
struct Ccs;
void bind_step_header(Transcript& tr, const Ccs& s, const Params& params, size_t ell, size_t d_sc, int32_t slack,
                      const uint64_t* x, size_t m_in, const uint64_t* c_mcs, const std::vector<MeHost>& acc);
StepChallenges sample_challenges(Transcript& tr, size_t ell_d, size_t ell);
K2 claimed_initial_sum(const Ccs& s, const StepChallenges& ch, const std::vector<MeHost>& acc, size_t ell_d);
void bind_rlc_inputs(Transcript& tr, const char* lane, size_t idx, const std::vector<MeHost>& ins);
void sample_rot_rhos(Transcript& tr, size_t kt, std::vector<int8_t>& rho_a, std::vector<uint64_t>& rhos);

// One RLC -> DEC lane's output: the rotations, the parent, the k_rho children and their planes.
struct LaneOut {
  std::vector<uint64_t> rhos;  // k x D x D
  MeHost parent;
  std::vector<MeHost> children;
  std::vector<const int8_t*> planes;
};

struct StepTiming {
  double transcript_s = 0, gpu_wait_s = 0;
};

class Prover {
 public:
  Prover(const Ccs& s, const Pp& pp);
  ~Prover();
  // Fold `nsteps` steps given their full z vectors (host, length m each).  Returns the
  // canonical proof encoding (u64 little-endian words; same layout as the CPU restatement).
  // z_on_device: z_steps are device pointers (inputs already resident in HBM).
  // mem (optional, one entry per step): the step's Twist/Shout instances over the shared CPU bus
  // (Route A memory sidecar, neo-fold/src/memory_sidecar/memory.rs).
  // ob (optional): output binding on the last step (fold_shard_prove_with_output_binding,
  // neo-fold/src/shard.rs:2352-2385); the proof then carries the output sum-check.
  std::vector<uint64_t> prove_shard(const uint64_t* const* z_steps, size_t nsteps, size_t m_in,
                                    bool z_on_device = false, const StepMemH* mem = nullptr,
                                    const ObCfgH* ob = nullptr);
  double last_seconds = 0;
  // Per-kernel-class timing of the last prove (HIP events on the prover stream).
  bool timing = false;
  // Host wall time per phase of the last prove: 0 bind header/ME inputs/challenges, 1 oracle
  // build, 2 row rounds, 3 y products, 4 Ajtai rounds, 5 ME outputs, 6 RLC binding + rhos,
  // 7 Z_mix, 8 DEC, 9 encode, 10 step intake (decomp + MCS commits), 11 accumulator swap.
  static constexpr int NPHASE = 12;
  double phase_s[NPHASE] = {0};
  uint64_t phase_perm[NPHASE] = {0};
  double kernel_ms[KT_N] = {0};
  uint64_t kernel_count[KT_N] = {0};
  uint64_t last_permutations = 0;
  // commit helpers exposed to the C ABI
  void commit_device(const int8_t* const* Zs_host_ptrs_dev, size_t nw, uint64_t* host_out);
  // the same in two halves: the launches (queued on the prover's stream), then the copy out
  void commit_launch(const int8_t* const* Zs_host_ptrs_dev, size_t nw);
  void commit_fetch(size_t nw, uint64_t* host_out);
  hipStream_t stream() const { return stream_; }
  const Ccs& ccs() const { return *s_; }
  const Pp& pp() const { return *pp_; }
  // lane-level Pi_RLC / Pi_DEC on this context's stream and buffers (the C ABI's ns_rlc / ns_dec)
  void rlc_api(const std::vector<MeHost>& ins, const int8_t* rho_coeffs, const int8_t* const* Z_dev, MeHost& parent,
               int32_t* zmix_dev);
  void dec_api(const MeHost& parent, const int32_t* zmix_dev, size_t k_dec, int8_t* const* planes_dev,
               size_t bus_chunk, std::vector<MeHost>& children, bool ok[3]);
  // Reuse this prover's device buffers (grown on demand) for another CCS / PP pair.
  void rebind(const Ccs& s, const Pp& pp);
  DevBuf& ws_ajtai() { return ws_ajtai_; }

 private:
  const Ccs* s_;
  const Pp* pp_;
  Params params_;
  hipStream_t stream_;
  DevBuf ws_ajtai_, ws_round_, ws_yd_;
  // buffers reused across fold steps
  OracleWs ows_;
  DevBuf planes_[2][16], zmix_, rho_dev_, zptr_, pptr_dev_, flags_, ydev_, ptrs_, cout_, zbuf_, vjs_, head_;
  std::vector<DevBuf> zmcs_;
  // Route A: val-lane child planes, M_j^T chi at r_val, bus weights / openings / decoded bus region
  DevBuf valplanes_[16], vjs_val_, rbits_, rchi_, busw_, busy_, busys_, busz_, yptr_;
  void bus_weights(const std::vector<K2>& r, const BusLayout& L);
  void open_bus_planes(const std::vector<const int8_t*>& planes, const BusLayout& L, const std::vector<MeHost*>& outs,
                       size_t t_core);
  void open_bus_i32(const int32_t* zmix, const BusLayout& L, MeHost& out, size_t t_core);
  void vjs_at(const std::vector<K2>& r, DevBuf& vjs);
  std::vector<K2> y_products(const std::vector<const int8_t*>& planes, const DevBuf& vjs);
  void y_launch(const std::vector<const int8_t*>& planes, const DevBuf& vjs);
  std::vector<K2> y_fetch(size_t nw);
  void head_cols(const int8_t* const* planes, size_t np, size_t m_in, std::vector<std::vector<uint64_t>*> Xs);
  void rlc_dec_lane(Transcript& tr, const char* lane, size_t idx, const std::vector<MeHost>& ins,
                    const std::vector<const int8_t*>& Zs, DevBuf* bank, const DevBuf& vjs, const BusLayout& L,
                    LaneOut& out);
};

// Ajtai-bit rounds of Pi_CCS on the host (optimized_engine/oracle.rs:1812-1951), from the
// precompute_for_r products y_eval[w][j][rho] at the folded row point.
// y_eval is folded by each round's challenge as it arrives (only the entries at multiples of
// 2^round stay meaningful), and a round's values come from the polynomial in X built once per
// round, so each round costs O(kt * t * 54 / 2^round) K products.
struct AjtaiPhase {
  size_t kt = 0, t = 0, ell_d = 6;
  std::vector<K2> y_eval, alpha, beta_a, chals;
  K2 gamma, f_prime, eq_beta_r, eq_r_in;
  bool has_inputs = false;
  void init(const Ccs& s, std::vector<K2> y, size_t k, const K2* alpha, const K2* beta_a, const K2* beta_r, K2 gamma,
            const std::vector<K2>& r_time, const K2* r_inputs);
  std::vector<K2> evals(const std::vector<K2>& xs) const;
  std::vector<K2> coeffs() const;  // the round polynomial, degree <= 4, low to high
  void fold(K2 r);
};

// Device-resident Pi_CCS round oracle (RoundOracle, neo-reductions/src/sumcheck.rs:17-34).
class CcsOracle {
 public:
  CcsOracle(const Ccs& s, const int8_t* const* Zs_dev_host, size_t k, const K2* alpha, const K2* beta_a,
            const K2* beta_r, K2 gamma, size_t ell_d, size_t ell_n, const K2* r_inputs, hipStream_t st,
            OracleWs* ws = nullptr);
  // coefficients of the current row round polynomial (deg_max + 1)
  std::vector<K2> round_coeffs();
  void fold(K2 r);
  size_t round() const { return round_; }
  size_t ell_n() const { return ell_n_; }
  // precompute_for_r at the folded row point: y_eval[w][j][rho]
  std::vector<K2> y_eval(const std::vector<K2>& r, DevBuf& ws, DevBuf& vjs_out);

 private:
  const Ccs& s_;
  hipStream_t st_;
  size_t k_, ell_n_, len_, round_ = 0;
  int nf_, nsc_;
  bool has_eval_;
  K2 gamma_k_;
  std::unique_ptr<OracleWs> own_;
  OracleWs* w_;           // Zptrs: k digit pointers; sc: scalar tables [nsc][len]; nc: NC [k*54][len]
  int cur_sc_ = 0, cur_nc_ = 0;
  std::vector<const int8_t*> zs_;
  K2 r0_ = K2{0, 0};  // round-0 challenge: round 1 is evaluated from the digits with it
  K2 r1_ = K2{0, 0};  // round-1 challenge: with r0, round 2 is evaluated from the digits
  // Round 2 from the digits (int8 MFMA): the NC tables are first materialized after round 2,
  // folded by r0, r1, r2 at once, and round 3 is evaluated unfused on them
  bool digit2_ = false;
  // From round 2 on, fold(r) only records r: the next round_coeffs folds and evaluates in one
  // launch (launch_round_fused), and anything else that reads the tables flushes it first.
  bool pending_ = false;
  K2 pending_r_ = K2{0, 0};
  void flush();
  RoundSignal next_signal();  // the mailbox's ready word and the value the next round stores
  // shared rounds: the first fused round of the oracle on the hub stream waits for this stream's
  // work (hub_pending_in_); before this stream reads what hub rounds wrote, it waits for them
  bool hub_used_ = false;
  void join_hub();
  // eq factors of the fused rounds: beta_r and r_in per row variable, the products of
  // eq(beta_i, r_i) and eq(r_in_i, r_i) over the folded variables, and the offsets of the chi
  // tables of the later variables (gt: every fused round's beta table, then its r_in table)
  std::vector<K2> beta_r_, rin_;
  K2 cb_ = K2{1, 0}, cr_ = K2{1, 0};
  std::vector<size_t> g_off_;
  size_t g_total_ = 0;
};

}  // namespace ns
