// route_a.h -- the Route A memory sidecar of a shard fold step on the host side of the MI355X
// prover: shared-bus layout and guardrails, bus time columns, the Twist/Shout address-lane
// sum-checks, the memory time lanes of the batched time sum-check and the Twist val-eval batch
// (neo-fold/src/memory_sidecar/{memory,route_a_time,cpu_bus,transcript,utils}.rs).  Every
// oracle is a device-resident TsOracleGpu; only round polynomials and challenges cross PCIe.
#pragma once
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "twist.h"

namespace ns {

// Memory instances of one step (neo-memory/src/witness.rs:23-96), tables materialized.
struct LutInstH {
  size_t k, d, n_side, steps, ell;
  const uint64_t* table;  // k F_q values
};
struct MemInstH {
  size_t k, d, n_side, steps, ell;
  bool sparse;  // MemInit::Sparse (else MemInit::Zero)
  size_t ninit;
  const uint64_t* init_addrs;
  const uint64_t* init_vals;
};
struct StepMemH {
  size_t n_lut = 0;
  const LutInstH* luts = nullptr;
  size_t n_mem = 0;
  const MemInstH* mems = nullptr;
};

// OutputBindingConfig (neo-fold/src/output_binding.rs:23-35) with the ProgramIO claims
// (neo-memory/src/output_check.rs:60-171) and the prover's final memory state of instance mem_idx.
struct ObCfgH {
  size_t mem_idx = 0, num_bits = 0, n_claims = 0;
  const uint64_t* addrs = nullptr;        // strictly increasing, < 2^num_bits
  const uint64_t* values = nullptr;       // F_q
  const uint64_t* final_state = nullptr;  // 2^num_bits F_q values (prover only)
};
// eval_init_at_r_addr (neo-memory/src/mem_init.rs:133-165): the initial memory at r_addr
K2 eval_init(const MemInstH& in, const std::vector<K2>& r_addr);
// Checks of the binding against the last step's Twist instance and the claims (shard.rs:1820-1846,
// output_check.rs:146-160); throws Error(1, ...) on a violation.
void ob_validate(const ObCfgH& ob, const StepMemH& sm);
// The ProgramIO absorb and r_addr sampling (output_check.rs:162-171, 260-277).
std::vector<K2> ob_sample_r_addr(Transcript& tr, const ObCfgH& ob);
// generate_output_sumcheck_proof_and_challenges (output_check.rs:623-634) on the device: the
// summand eq(r_addr, .) * io_mask * (val_final - val_io) as a three-factor product oracle.
void ob_prove(Transcript& tr, const ObCfgH& ob, std::vector<std::vector<K2>>& rounds, std::vector<K2>& r_prime);

// BusLayout (neo-memory/src/cpu/bus_layout.rs:29-56)
struct BusLayout {
  struct Sh {
    size_t addr0, ell, has_lookup, val;
  };
  struct Tw {
    size_t ra0, wa0, ell, has_read, has_write, wv, rv, inc;
  };
  size_t m = 0, m_in = 0, chunk = 1, bus_cols = 0, bus_base = 0;
  std::vector<Sh> sh;
  std::vector<Tw> tw;
};

// infer_bus_layout_for_steps (cpu_bus.rs:90-168)
BusLayout infer_bus_layout(const StepMemH* mem, size_t nsteps, size_t n, size_t m, size_t m_in);
// ensure_ccs_has_shared_bus_padding_for_steps + ensure_ccs_binds_shared_bus_for_steps
// (cpu_bus.rs:171-181, 435-967); xs: every step's public inputs.  Throws on a violation.
void check_bus_guardrails(const Ccs& s, const BusLayout& L, const std::vector<std::vector<uint64_t>>& xs);

// absorb_step_memory_witness (memory.rs:82-136); table digests cached by table address within
// one prove (a table digest absorbs every entry: 2^16 entries are ~16k permutations)
struct DigestCache {
  std::map<std::pair<const uint64_t*, size_t>, std::vector<uint8_t>> tables;
};
void absorb_step_memory(Transcript& tr, const StepMemH& sm, DigestCache& cache);

struct TwistCols {
  size_t ell = 0;
  std::vector<SparseCol> ra, wa;
  SparseCol has_read, has_write, wv, rv, inc;
};
struct ShoutCols {
  std::vector<SparseCol> addr;
  SparseCol has_lookup, val;
};

// one claim of a batched sum-check (BatchedClaim, neo-reductions/src/sumcheck.rs:315-324)
struct RaClaim {
  TsOracleGpu* ts = nullptr;
  CcsOracle* ccs = nullptr;
  size_t deg = 0;
  const char* label = "";
  bool dyn = false;
  K2 sum{0, 0}, running{0, 0};
  std::vector<K2> polys;  // rounds x (deg + 1) coefficients, low to high
};
// run_batched_sumcheck_prover_ds (memory_sidecar/sumcheck_ds.rs:51-63 over sumcheck.rs:350-456):
// every device oracle's round is enqueued before the host waits for any of them
void run_batched(Transcript& tr, const char* domain, size_t inst, std::vector<RaClaim>& cl, size_t rounds,
                 std::vector<K2>& chals);
// bind_batched_dynamic_claims (memory_sidecar/transcript.rs:36-64)
void bind_dynamic_claims(Transcript& tr, const std::vector<RaClaim>& cl);

struct ShoutPre {
  size_t ell_addr = 0;
  uint64_t active_mask = 0;
  std::vector<K2> claimed, r_addr;
  std::vector<std::vector<K2>> rounds;  // per active lane: ell_addr x (deg + 1)
  std::vector<size_t> degs;
  std::vector<ShoutCols> cols;
};
struct TwistPre {
  size_t ell_addr = 0;
  TwistCols cols;
  std::vector<K2> r_addr;
  std::vector<K2> rounds[2];
  size_t degs[2] = {0, 0};
  K2 read_claim{0, 0}, write_claim{0, 0};
};
// zbus: the decoded bus region z[bus_base .. m) of the step
void shout_addr_pre(Transcript& tr, const StepMemH& sm, const BusLayout& L, const uint64_t* zbus, size_t ell_n,
                    const std::vector<K2>& r_cycle, size_t step_idx, ShoutPre& out);
void twist_addr_pre(Transcript& tr, const StepMemH& sm, const BusLayout& L, const uint64_t* zbus, size_t ell_n,
                    const std::vector<K2>& r_cycle, std::vector<TwistPre>& out);
// build_route_a_memory_oracles (memory.rs:905-1025) in the batched time claim order
void build_time_claims(const StepMemH& sm, size_t ell_n, const std::vector<K2>& r_cycle, const ShoutPre& sp,
                       const std::vector<TwistPre>& tp, std::vector<std::unique_ptr<TsOracleGpu>>& owned,
                       std::vector<RaClaim>& cl);

struct ValPre {
  size_t per = 0;
  std::vector<RaClaim> claims;
  std::vector<K2> r_val;
};
// the val-eval batch of finalize_route_a_memory_prover (memory.rs:1382-1535); prev: the previous
// step's decoded Twist columns (rollover) or null
void val_eval_batch(Transcript& tr, const StepMemH& sm, const std::vector<TwistPre>& tp,
                    const std::vector<TwistCols>* prev, size_t ell_n, const std::vector<K2>& r_time, size_t step_idx,
                    ValPre& out);

// fold_shard_verify over the canonical encoding (verify.cpp): returns when the proof is accepted,
// throws Error(7, reason) otherwise.  mcs_c: expected MCS commitments (n_steps x kappa*D) or null.
// ob: the output binding the proof must carry (fold_shard_verify_with_output_binding), or null.
void verify_shard(const Ccs& s, const uint64_t* const* x_steps, size_t nsteps, size_t m_in, const StepMemH* mem,
                  const uint64_t* mcs_c, const uint8_t* proof, size_t proof_len, const ObCfgH* ob = nullptr);

}  // namespace ns
