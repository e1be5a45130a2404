/*
 * nightstream.h -- C ABI of the MI355X Neo folding backend (libnightstream.so).
 *
 * The reference (LFDT-Nightstream, Rust) has no FFI layer: its hot-path boundaries are
 * in-process traits.  Each entry point below replaces one of them with plain pointers and
 * sizes so a Rust `extern "C"` crate (or ctypes) can bind it.  Conventions:
 *   - F_q elements are uint64_t canonical values < q = 2^64 - 2^32 + 1.
 *   - K = F_q[u]/(u^2 - 7) elements are two uint64_t (c0, c1).
 *   - Digit matrices are int8 row-major D x m (Z[rho*m + c]), D = 54, entries in {-1,0,1}.
 *   - Commitments are kappa*D uint64_t, data[col*D + row] (neo-ajtai/src/types.rs:17-22).
 *   - Every function returns 0 on success or a negative NS_E* code; ns_last_error() returns
 *     the message of the calling thread's last failure (the reference panics or returns
 *     Err(PiCcsError) in these cases, neo-reductions/src/error.rs:6-21).
 *   - All work runs on the calling process's current HIP device; nothing falls back to the
 *     CPU: without a usable GPU every compute entry point returns NS_ENODEV.
 */
#ifndef NIGHTSTREAM_H
#define NIGHTSTREAM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NS_OK 0
#define NS_EINVAL -1      /* invalid input (PiCcsError::InvalidInput) */
#define NS_EPARAMS -2     /* parameter / extension policy failure (ParamsError) */
#define NS_ESUMCHECK -4   /* sum-check invariant failed (PiCcsError::SumcheckError) */
#define NS_ERANGE -6      /* DEC split out of range (PiCcsError::ProtocolError) */
#define NS_EPROTOCOL -7   /* other protocol check failed (PiCcsError::ProtocolError) */
#define NS_ENODEV -10     /* no usable HIP device */
#define NS_EHIP -11       /* HIP runtime error */

typedef struct ns_ccs ns_ccs;
typedef struct ns_pp ns_pp;
typedef struct ns_prover ns_prover;
typedef struct ns_oracle ns_oracle;

/* ---- runtime ---- */
const char* ns_last_error(void);
int ns_device_count(void);
/* Free a buffer returned by this library. */
void ns_free(void* p);
/* Device memory on this library's HIP runtime, for callers that keep witnesses resident in HBM
 * (ns_prove_shard_device).  Runtime plumbing with no reference counterpart. */
int ns_device_alloc(size_t bytes, void** out);
int ns_device_free(void* p);
int ns_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int ns_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int ns_device_synchronize(void);

/* ---- parameters: NeoParams::goldilocks_auto_r1cs_ccs (neo-params/src/lib.rs:186-222) ----
 * out[0..10] = q, eta, d, kappa, m, b, k_rho, B, T, s, lambda */
int ns_params_goldilocks_auto_r1cs_ccs(size_t n_rows, uint64_t out[11]);

/* ---- CCS structure: r1cs_to_ccs (neo-ccs/src/r1cs.rs:13-43), identity-first, n == m ----
 * (host only: the matrices reach HBM when a prover or an oracle first uses the structure)
 * CSR per matrix: row_ptr[n+1], col[nnz], val[nnz].  The digest is digest_ccs_matrices
 * (neo-reductions/src/engines/utils.rs:214-321). */
int ns_ccs_from_r1cs(size_t n, size_t m, const uint64_t* a_ptr, const uint32_t* a_col, const uint64_t* a_val,
                     const uint64_t* b_ptr, const uint32_t* b_col, const uint64_t* b_val, const uint64_t* c_ptr,
                     const uint32_t* c_col, const uint64_t* c_val, ns_ccs** out);
int ns_ccs_digest(const ns_ccs* s, uint64_t out[4]);
void ns_ccs_free(ns_ccs* s);

/* ---- Ajtai public parameters: neo_ajtai::setup (neo-ajtai/src/commit.rs:133-156) ----
 * Samples M in R_q^{kappa x m} from ChaCha8Rng::from_seed(seed) directly in HBM. */
int ns_pp_setup_seeded(const uint8_t seed[32], size_t m, int keep_coeffs, ns_pp** out);
/* PP from host coefficients pp[(i*m + j)*54 + u] (kappa = 16). */
int ns_pp_from_coeffs(const uint64_t* coeffs, size_t m, int keep_coeffs, ns_pp** out);
/* Copy coefficients of PP rows back (requires keep_coeffs): out[(i*m + j)*54 + u]. */
int ns_pp_coeffs(const ns_pp* pp, uint64_t* out);
void ns_pp_free(ns_pp* pp);

/* ---- decomposition: decomp_b_row_major, b = 2 balanced (neo-ajtai/src/decomp.rs:118-158) ---- */
int ns_decomp_b(const uint64_t* z, size_t m, uint32_t b, int8_t* Z_out);

/* ---- SModuleHomomorphism::commit (neo-ccs/src/traits.rs:8-14; AjtaiSModule::commit,
 *      neo-ajtai/src/s_module.rs:335-364 -> commit_precomp_ct_row_major, commit.rs:679-744).
 * Commits n_wit host digit matrices in one batched launch; out[w*kappa*D + col*D + row].
 * Digits must lie in {-1, 0, 1} (balanced b = 2): the int8 GEMM accumulates per-K-split int32
 * partial sums that are exact only in that range; other digits return NS_EINVAL. */
int ns_ajtai_commit_batch(const ns_pp* pp, const int8_t* const* Z, size_t n_wit, uint64_t* out);
/* The same commitments for digit matrices already resident in HBM: Z_dev[w] is a device pointer
 * to the D x m int8 matrix of witness w (from ns_device_alloc); only the commitments cross PCIe.
 * Same {-1, 0, 1} digit contract; device digits are not scanned (the caller's decomposition,
 * ns_decomp_b or the prover's own, produces only those). */
int ns_ajtai_commit_batch_device(const ns_pp* pp, const int8_t* const* Z_dev, size_t n_wit, uint64_t* out);
/* Tile shape of the commitment GEMM: 0 (default) times the three shapes (16 x 16 x 64 MFMAs in
 * 192 x 352 blocks; 32 x 32 x 32 MFMAs in 256 x 256 blocks on int8 digits; the same blocks on
 * digits packed to 2-bit codes, which relies on the {-1, 0, 1} digit contract) and every K-split
 * count on the first commit of each (m, n_wit) and keeps the fastest; 1, 2 or 3 forces that
 * shape.  Clears the cached choices.  The commitments do not depend on it.  No counterpart in the reference: a tuning and
 * testing hook for callers that want a fixed kernel (reproducible timing, per-shape tests). */
int ns_ajtai_force_variant(int variant);

/* ---- RoundOracle for Pi_CCS (neo-reductions/src/sumcheck.rs:17-34; OptimizedOracle,
 *      optimized_engine/oracle.rs:1597-1639 new_with_sparse, 2075-2104 trait impl).
 * Z: k host digit matrices (MCS first, then ME witnesses), uploaded once and kept in HBM.
 * alpha[ell_d], beta_a[ell_d], beta_r[ell_n], r_inputs[ell_n] (nullable) are K elements. */
int ns_ccs_oracle_new(const ns_ccs* s, const int8_t* const* Z, size_t k, const uint64_t* alpha,
                      const uint64_t* beta_a, const uint64_t* beta_r, const uint64_t gamma[2], size_t ell_d,
                      size_t ell_n, const uint64_t* r_inputs, ns_oracle** out);
/* evals_at for the row rounds: out[2*i..] = round polynomial at xs[i] (n points). */
int ns_oracle_evals(ns_oracle* o, const uint64_t* xs, size_t n, uint64_t* out);
int ns_oracle_fold(ns_oracle* o, const uint64_t r[2]);
size_t ns_oracle_num_rounds(const ns_oracle* o);
size_t ns_oracle_degree_bound(const ns_oracle* o);
/* precompute_for_r (oracle.rs:1659-1783) at row point r[ell_n]: out[(w*t + j)*54 + rho] (K). */
int ns_oracle_y_eval(ns_oracle* o, const uint64_t* r, uint64_t* out);
void ns_oracle_free(ns_oracle* o);

/* ---- Pi_RLC Z_mix (optimized_engine/common.rs:1329-1363): zmix = sum_i rot(a_i) Z_i,
 *      a_i = rho_coeffs[i*54 .. ] in [-2, 2] (neo-reductions/src/common.rs:370-403). */
int ns_rlc_zmix(const int8_t* rho_coeffs, const int8_t* const* Z, size_t k, size_t m, int32_t* zmix_out);
/* ---- Pi_DEC split (neo-reductions/src/common.rs:75-268): balanced base-2 digits of zmix
 *      into k_dec planes; nonzero[i] = plane i has a nonzero digit.  NS_ERANGE if |v| >= 2^k_dec. */
int ns_dec_split(const int32_t* zmix, size_t m, size_t k_dec, int8_t* const* outs, int* nonzero);
/* ---- no reference counterpart: the Pi_DEC split for callers that cannot allocate device memory
 *      (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * zmix_dev: device int32 [54][m]; outs_dev[i] (host array of k_dec device pointers): device int8
 * [54][m] planes; nonzero: host int[k_dec].  Same digits, flags and NS_ERANGE as ns_dec_split.
 * Allocates nothing: its counters live in a static per-GPU scratch array of 64 ints (so
 * 1 <= k_dec <= 30), and every call on one GPU shares that array -- calls that overlap on one
 * GPU race. */
int ns_dec_split_device_noalloc(const int32_t* zmix_dev, size_t m, size_t k_dec, int8_t* const* outs_dev,
                                int* nonzero);

/* ---- lane-level Pi_RLC / Pi_DEC returning ME instances (neo-reductions/src/api.rs:131-324:
 *      rlc_with_commit, dec_children_with_commit), on a prover context's stream and buffers.
 * An MeInstance (neo-ccs/src/relations.rs) is passed with caller-owned arrays: X [54][m_in] F_q,
 * r [ell_n] K, y [t][64] K (core matrices first, then one opening per shared-bus column), y_scalars
 * [t] K.  Outputs use the same shapes as the inputs (the caller sizes their arrays). */
typedef struct {
  uint64_t c[16 * 54];
  size_t m_in;
  uint64_t* X;
  size_t ell_n;
  uint64_t* r;
  size_t t;
  uint64_t* y;
  uint64_t* y_scalars;
  uint8_t fold_digest[32];
} ns_me;
/* parent = sum_i rot(a_i) ins[i] with a_i = rho_coeffs[i*54 ..] in [-2, 2] (sample_rot_rhos_n's
 * vectors; one input: the identity fast path), and Z_mix = sum_i rot(a_i) Z_i into zmix_dev
 * (device int32 [54][m]); Z_dev[i]: device digit planes of the inputs. */
int ns_rlc(ns_prover* p, const ns_me* ins, size_t k, const int8_t* rho_coeffs, const int8_t* const* Z_dev,
           ns_me* parent, int32_t* zmix_dev);
/* The k_dec children of parent: balanced base-2 planes of zmix_dev into planes_dev[i] (device
 * int8 [54][m]), child commitments, X and y (rows past the CCS's t are bus openings of a shared
 * bus of bus_chunk cycles per column; 0 when there is none).  ok[0..2] = the public checks on y,
 * X and c (dec_children_with_commit's ok_y, ok_X, ok_c).  NS_ERANGE when Z_mix does not fit. */
int ns_dec(ns_prover* p, const ns_me* parent, const int32_t* zmix_dev, size_t k_dec, int8_t* const* planes_dev,
           size_t bus_chunk, ns_me* children, int* ok);

/* ---- shard prover: fold_shard_prove (neo-fold/src/shard.rs:2285-2316) driven as
 *      FoldingSession::{add_step_io, fold_and_prove} (neo-fold/src/session.rs:1018-1311).
 * z_steps[s] is step s's full z = [x | w] (length m); x = z[0..m_in].  On success *proof
 * holds the canonical encoding (little-endian u64 words, nightstream_amd/proof.py); free it
 * with ns_free. */
int ns_prover_new(const ns_ccs* s, const ns_pp* pp, ns_prover** out);
/* Point an existing prover at another CCS / PP pair, keeping its device buffers (grown on
 * demand): a serving worker proves job after job without reallocating.  No reference
 * counterpart (the reference builds a FoldingSession per shard, session.rs:668-1953). */
int ns_prover_rebind(ns_prover* p, const ns_ccs* s, const ns_pp* pp);
int ns_prove_shard(ns_prover* p, const uint64_t* const* z_steps, size_t n_steps, size_t m_in, uint8_t** proof,
                   size_t* proof_len, double* seconds);
/* Same, with z_steps_dev[s] device pointers (inputs already resident in HBM). */
int ns_prove_shard_device(ns_prover* p, const uint64_t* const* z_steps_dev, size_t n_steps, size_t m_in,
                          uint8_t** proof, size_t* proof_len, double* seconds);
/* ---- the same with Twist/Shout memory over the shared CPU bus (Route A): fold_shard_prove with
 *      StepWitnessBundle lut_instances / mem_instances (neo-memory/src/witness.rs:23-205,
 *      neo-fold/src/shard.rs:1725-2283, memory_sidecar/memory.rs).  mem[s] describes step s's
 *      instances (metadata-only: their columns are the bus tail of z, layout
 *      neo-memory/src/cpu/bus_layout.rs:98-191, which the CCS must pad and bind as
 *      memory_sidecar/cpu_bus.rs:171-181 requires; violations return NS_EINVAL).  The proof adds
 *      each step's MemSidecarProof, the batched time claims of the memory lanes and the val-lane
 *      RLC/DEC, and ends with both obligation lanes (ShardObligations{main, val}).
 *      z_on_device: z_steps are device pointers.  mem == NULL proves CCS-only steps. */
typedef struct {
  size_t k, d, n_side, steps, ell;
  const uint64_t* table;  /* k = 2^(d*ell) F_q entries (LutInstance with no table_spec) */
} ns_lut_inst;
typedef struct {
  size_t k, d, n_side, steps, ell;
  int init_sparse;        /* 0: MemInit::Zero; 1: MemInit::Sparse(init_addrs[i], init_vals[i]) */
  size_t ninit;
  const uint64_t* init_addrs; /* strictly increasing, < k */
  const uint64_t* init_vals;  /* F_q */
} ns_mem_inst;
typedef struct {
  size_t n_lut;
  const ns_lut_inst* luts;
  size_t n_mem;
  const ns_mem_inst* mems;
} ns_step_mem;
int ns_prove_shard_mem(ns_prover* p, const uint64_t* const* z_steps, int z_on_device, size_t n_steps, size_t m_in,
                       const ns_step_mem* mem, uint8_t** proof, size_t* proof_len, double* seconds);
/* ---- fold_shard_verify (neo-fold/src/shard.rs:2482-2994): replays the Fiat-Shamir transcript
 *      of a canonical proof and re-derives every challenge (alpha, beta, gamma, r_cycle, every
 *      round challenge, the RLC rotations, the digests), then checks the sum-check chains, the
 *      Pi_CCS terminal identity, the Route A claim links, Pi_RLC / Pi_DEC of both lanes and both
 *      obligation lanes.  x_steps[s]: step s's public input (m_in values); mem as for
 *      ns_prove_shard_mem (NULL for CCS-only shards); mcs_commitments: n_steps x kappa*D expected
 *      MCS commitments, or NULL to take them from the proof.  Host only (no device needed).
 *      Returns 0 when the proof is accepted, NS_EPROTOCOL with the reason otherwise. */
int ns_shard_verify(const ns_ccs* s, const uint64_t* const* x_steps, size_t n_steps, size_t m_in,
                    const ns_step_mem* mem, const uint64_t* mcs_commitments, const uint8_t* proof, size_t proof_len);
/* ---- output binding: fold_shard_prove_with_output_binding / fold_shard_verify_with_output_binding
 *      (neo-fold/src/shard.rs:2352-2385, 3032-3108; OutputBindingConfig output_binding.rs:23-35,
 *      ProgramIO neo-memory/src/output_check.rs:60-171).  On the last step, right after its memory
 *      is absorbed, the prover runs the output sum-check over eq(r_addr, .) * io_mask *
 *      (val_final - val_io) (num_bits rounds of degree 3) and adds the batched time claim
 *      "output_binding/inc_total" linking the final state to the Twist increments of instance
 *      mem_idx; the proof carries the round polynomials after the steps.  Requirements (else
 *      NS_EINVAL): n_steps >= 1 with memory, mem_idx < n_mem of the last step, that instance's
 *      k == 2^num_bits and d*ell == num_bits, num_bits <= 30, claim addresses strictly increasing
 *      and < 2^num_bits.  final_state (prover only): 2^num_bits F_q values, the memory after the
 *      last step; the verifier ignores it. */
typedef struct {
  size_t mem_idx;
  size_t num_bits;
  size_t n_claims;
  const uint64_t* addrs;       /* claimed addresses, strictly increasing */
  const uint64_t* values;      /* claimed values (F_q) */
  const uint64_t* final_state; /* 2^num_bits F_q values (prover) or NULL (verifier) */
} ns_output_binding;
int ns_prove_shard_ob(ns_prover* p, const uint64_t* const* z_steps, int z_on_device, size_t n_steps, size_t m_in,
                      const ns_step_mem* mem, const ns_output_binding* ob, uint8_t** proof, size_t* proof_len,
                      double* seconds);
int ns_shard_verify_ob(const ns_ccs* s, const uint64_t* const* x_steps, size_t n_steps, size_t m_in,
                       const ns_step_mem* mem, const ns_output_binding* ob, const uint64_t* mcs_commitments,
                       const uint8_t* proof, size_t proof_len);
/* Shared launches for the fused Pi_CCS rounds (rounds >= 3) of concurrent provers on one GPU
 * (process-wide; off by default): a prover queues its round, and whichever prover finds no launch
 * being issued issues every queued round -- one alone as its own launch, several as one ragged
 * launch with a job array -- on a shared stream.  Proofs do not depend on it.  The counters
 * report launches issued and rounds carried since the process started.  No counterpart in the
 * reference (its sessions do not share a device): for servers and batch provers running many
 * sessions at once. */
int ns_set_shared_rounds(int on);
int ns_shared_round_stats(uint64_t* launches, uint64_t* rounds);
/* Per-kernel timing of the next proves (HIP events on the prover's stream); classes:
 * 0 round-0 eval, 1 round-2 eval, 2 table fold, 3 Ajtai GEMM, 4 digit x K GEMM, 5 RLC Z_mix,
 * 6 DEC split, 7 round-1 eval, 8 fused fold + round (rounds >= 3).  ms[i] is the summed
 * duration and counts[i] the number of launches. */
int ns_prover_set_timing(ns_prover* p, int on);
int ns_prover_kernel_stats(const ns_prover* p, double* ms, uint64_t* counts, size_t n);
/* Host wall-clock seconds per phase of the last prove (12 phases: bind inputs, oracle build,
 * row rounds, y products, Ajtai rounds, ME outputs, RLC binding, Z_mix, DEC, encode, intake,
 * accumulator swap). */
int ns_prover_phase_stats(const ns_prover* p, double* seconds, size_t n);
/* Poseidon2 permutations of the host transcript per phase of the last prove (same 12 phases). */
int ns_prover_phase_perms(const ns_prover* p, uint64_t* perms, size_t n);

/* ---- host Poseidon2 (neo-ccs/src/crypto/poseidon2_goldilocks.rs:40-53 permutation()) ----
 * Permutes state[width] in place: width 8 is the transcript permutation
 * (neo-transcript/src/poseidon2.rs), width 16 the CCS-digest one.  force_scalar = 1 runs the
 * portable path even where the AVX-512 one is selected (both give the same output).  Host only. */
int ns_poseidon2_permute(uint64_t* state, size_t width, int force_scalar);
void ns_prover_free(ns_prover* p);

/* ---- Twist/Shout round oracles (neo-memory/src/twist_oracle.rs).  Each handle implements the
 *      RoundOracle contract (neo-reductions/src/sumcheck.rs:17-34) through ns_ts_evals /
 *      ns_ts_fold / ns_ts_num_rounds / ns_ts_degree_bound.  Sparse time columns are SparseIdxVec
 *      (neo-memory/src/sparse_time.rs:10-24): nnz strictly increasing indices in [0, 2^ell_n)
 *      and their K values (vals[2i], vals[2i+1]).  claim, where given, receives the sum the
 *      reference constructor returns. */
typedef struct {
  size_t nnz;
  const uint64_t* idx;
  const uint64_t* vals;
} ns_sparse_col;
typedef struct ns_ts ns_ts;
/* ProductRoundOracle::new (twist_oracle.rs:56-76): nf dense factors of length len (power of 2). */
int ns_ts_product_new(const uint64_t* const* factors, size_t nf, size_t len, size_t degree_bound, ns_ts** out);
/* ShoutValueOracleSparse::new (twist_oracle.rs:213-245). */
int ns_ts_shout_value_new(const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_lookup, ns_sparse_col val,
                          uint64_t claim[2], ns_ts** out);
/* IndexAdapterOracleSparseTime::new_with_gate (twist_oracle.rs:396-442). */
int ns_ts_index_adapter_new(const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_lookup,
                            const ns_sparse_col* addr_bits, size_t ell_addr, const uint64_t* r_addr, uint64_t claim[2],
                            ns_ts** out);
/* LazyWeightedBitnessOracleSparseTime::new_with_cycle (twist_oracle.rs:525-541). */
int ns_ts_bitness_new(const uint64_t* r_cycle, size_t ell_n, const ns_sparse_col* cols, const uint64_t* weights,
                      size_t ncols, ns_ts** out);
/* TwistReadCheckOracleSparseTime::new (twist_oracle.rs:641-684). */
int ns_ts_twist_read_new(const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_read, ns_sparse_col rv,
                         const ns_sparse_col* ra_bits, ns_sparse_col has_write, ns_sparse_col inc,
                         const ns_sparse_col* wa_bits, size_t ell_addr, const uint64_t* r_addr,
                         const uint64_t init_at_r_addr[2], ns_ts** out);
/* TwistWriteCheckOracleSparseTime::new (twist_oracle.rs:787-822). */
int ns_ts_twist_write_new(const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_write, ns_sparse_col wv,
                          ns_sparse_col inc, const ns_sparse_col* wa_bits, size_t ell_addr, const uint64_t* r_addr,
                          const uint64_t init_at_r_addr[2], ns_ts** out);
/* TwistValEvalOracleSparseTime::new (twist_oracle.rs:927-972); r_time has ell_n entries. */
int ns_ts_val_eval_new(size_t ell_n, const ns_sparse_col* wa_bits, size_t ell_addr, ns_sparse_col has_write,
                       ns_sparse_col inc, const uint64_t* r_addr, const uint64_t* r_time, uint64_t claim[2],
                       ns_ts** out);
/* TwistTotalIncOracleSparseTime::new (twist_oracle.rs:1063-1104). */
int ns_ts_total_inc_new(size_t ell_n, const ns_sparse_col* wa_bits, size_t ell_addr, ns_sparse_col has_write,
                        ns_sparse_col inc, const uint64_t* r_addr, uint64_t claim[2], ns_ts** out);
/* TwistReadCheckAddrOracleSparseTime::new (twist_oracle.rs:1282-1378); init: (addr, K) pairs. */
int ns_ts_read_addr_new(const uint64_t* init_addrs, const uint64_t* init_vals, size_t ninit, const uint64_t* r_cycle,
                        size_t ell_n, ns_sparse_col has_read, ns_sparse_col rv, const ns_sparse_col* ra_bits,
                        ns_sparse_col has_write, const ns_sparse_col* wa_bits, ns_sparse_col inc, size_t ell_addr,
                        ns_ts** out);
/* TwistWriteCheckAddrOracleSparseTime::new (twist_oracle.rs:1489-1547). */
int ns_ts_write_addr_new(const uint64_t* init_addrs, const uint64_t* init_vals, size_t ninit,
                         const uint64_t* r_cycle, size_t ell_n, ns_sparse_col has_write, ns_sparse_col wv,
                         const ns_sparse_col* wa_bits, ns_sparse_col inc, size_t ell_addr, ns_ts** out);
/* AddressLookupOracle::new (twist_oracle.rs:1684-1736): table of table_len K entries. */
int ns_ts_addr_lookup_new(const ns_sparse_col* addr_bits, size_t ell_addr, ns_sparse_col has_lookup,
                          const uint64_t* table, size_t table_len, const uint64_t* r_cycle, size_t ell_n,
                          uint64_t claim[2], ns_ts** out);
int ns_ts_evals(ns_ts* o, const uint64_t* xs, size_t n, uint64_t* out);
int ns_ts_fold(ns_ts* o, const uint64_t r[2]);
/* Threads: the ns_ts_*_new constructors may run concurrently on several host threads (each
 * oracle gets its own stream; the library's device-block cache is locked); one oracle must not be
 * used from two threads at once, as with the reference's &mut self. */
/* One round of the batched sum-check with shared challenges
 * (run_batched_sumcheck_prover, neo-reductions/src/sumcheck.rs:350-456, the evals_at loop at
 * :390-407): every oracle's round is enqueued on its own stream before the host waits for any.
 * For each oracle in order, its round polynomial's values at x = 0, 1, ..., degree_bound are
 * written to out (degree_bound + 1 K values, 2 words each, oracles back to back).  Every oracle
 * must have the same number of rounds left, at least one (the reference panics otherwise). */
int ns_ts_batch_round(ns_ts* const* os, size_t n, uint64_t* out);
/* Fold every oracle with the shared challenge r (sumcheck.rs:436-442). */
int ns_ts_batch_fold(ns_ts* const* os, size_t n, const uint64_t r[2]);
size_t ns_ts_num_rounds(const ns_ts* o);
size_t ns_ts_degree_bound(const ns_ts* o);
void ns_ts_free(ns_ts* o);

/* ---- shared CPU bus of the memory sidecar.  Z: host digit matrix [54][m] of the CPU witness;
 *      bus_cell(col, j) = bus_base + col*chunk + j, time_index(j) = m_in + j
 *      (neo-memory/src/cpu/bus_layout.rs:63-80). */
/* decode_vector (neo-memory/src/ajtai.rs:31-55): z[c] = sum_rho 2^rho Z[rho][c]. */
int ns_bus_decode(const int8_t* Z, size_t m, uint64_t* z_out);
/* decode_cpu_z_to_k + build_time_sparse_from_bus_col for every column (cpu_bus.rs:969-1016):
 * column c's nonzero entries go to idx_out[c*chunk ..] / vals_out[2*(c*chunk) ..], count nnz_out[c]. */
int ns_bus_time_cols(const int8_t* Z, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                     size_t steps_len, size_t ell_n, uint64_t* idx_out, uint64_t* vals_out, size_t* nnz_out);
/* append_bus_openings_to_me_instance (cpu_bus.rs:195-320): y_out[(c*64 + rho)*2 ..] and
 * y_scalars_out[c*2 ..] for every bus column at the ME point r[ell_n]. */
int ns_bus_openings(const int8_t* Z, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                    const uint64_t* r, size_t ell_n, uint64_t* y_out, uint64_t* y_scalars_out);

#ifdef __cplusplus
}
#endif
#endif
