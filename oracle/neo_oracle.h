/*
 * neo_oracle.h -- CPU restatement of the Nightstream Neo folding hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X backend
 * in nightstream_amd/: tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it; the product path never does.  Every function names the
 * reference file:line it restates (paths relative to the reference checkout).
 *
 * Parity status (what pins each part) is recorded in the header comment of the
 * .c file that implements it.  Parts with no external pin say "parity unpinned".
 *
 * Conventions shared with the product library:
 *   fq   : canonical Goldilocks element, 0 <= x < q, q = 2^64 - 2^32 + 1
 *   kx   : K = F_q[u]/(u^2 - 7), stored (c0, c1) = c0 + c1*u
 *   digit matrices Z: int8, row-major D x m (Z[rho*m + c]), values in {-1,0,1}
 *   Commitment: kappa*D fq, data[col*D + row] (neo-ajtai/src/types.rs:17-22)
 */
#ifndef NEO_ORACLE_H
#define NEO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEO_D 54
#define NEO_Q 0xFFFFFFFF00000001ULL

typedef uint64_t fq;
typedef struct { fq c0, c1; } kx;

/* ---------------- field (neo-math/src/field.rs) ---------------- */
fq  or_fq_add(fq a, fq b);
fq  or_fq_sub(fq a, fq b);
fq  or_fq_neg(fq a);
fq  or_fq_mul(fq a, fq b);
fq  or_fq_pow(fq a, uint64_t e);
fq  or_fq_inv(fq a);
fq  or_fq_from_i64(int64_t v);
int64_t or_fq_to_balanced(fq a);           /* neo-ajtai/src/util.rs:20-29 */
kx  or_k_add(kx a, kx b);
kx  or_k_sub(kx a, kx b);
kx  or_k_neg(kx a);
kx  or_k_mul(kx a, kx b);
kx  or_k_scale(kx a, fq s);
kx  or_k_inv(kx a);
kx  or_k_from(fq a);
int or_k_eq(kx a, kx b);

/* exported wrappers for tests (pass by pointer) */
void or_fq_mul_batch(const fq* a, const fq* b, fq* out, size_t n);
void or_k_mul_batch(const fq* a, const fq* b, fq* out, size_t n); /* pairs */
void or_k_inv_batch(const fq* a, fq* out, size_t n);

/* ---------------- ring R_q = F_q[X]/(X^54+X^27+1) (neo-math/src/ring.rs) ---------------- */
void or_rq_mul(const fq* a, const fq* b, fq* out);           /* ring.rs:54-66, 187-203 */
void or_rot_step(const fq* cur, fq* next);                   /* neo-ajtai/src/commit.rs:46-54 */

/* ---------------- ChaCha (rand_chacha 0.9 ChaCha8Rng layout) ---------------- */
typedef struct {
  uint32_t key[8];
  uint64_t counter;  /* next block counter */
  uint64_t stream;
  uint32_t buf[64];  /* 4 blocks */
  int idx;           /* next word in buf; 64 => empty */
  int rounds;
} or_chacha;
void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds, uint32_t out[16]);
void or_chacha_ietf_block(const uint32_t key[8], uint32_t counter, const uint32_t nonce[3], int rounds, uint32_t out[16]);
void or_chacha_init(or_chacha* r, const uint8_t seed[32], int rounds);
void or_chacha_seed_from_u64(or_chacha* r, uint64_t state, int rounds);
uint32_t or_chacha_next_u32(or_chacha* r);
uint64_t or_chacha_next_u64(or_chacha* r);
void or_chacha_fill_bytes(or_chacha* r, uint8_t* dst, size_t len);

/* ---------------- Poseidon2 over Goldilocks (p3-poseidon2 0.4.1 structure) ---------------- */
typedef struct {
  int width, rounds_f, rounds_p;
  fq ext_init[4][16];
  fq ext_term[4][16];
  fq internal_rc[32];
  fq diag[16];
} or_p2;
void or_p2_init_from_seed(or_p2* p, int width, or_chacha* rng);
void or_p2_permute(const or_p2* p, fq* state);
const or_p2* or_p2_transcript_perm(void);  /* width 8, ChaCha8(SEED) */
const or_p2* or_p2_digest_perm(void);      /* width 16, ChaCha8::seed_from_u64(CCS_DIGEST_SEED) */
void or_p2_permute8(fq* state);            /* test hook */

/* ---------------- transcript (neo-transcript/src/poseidon2.rs) ---------------- */
typedef struct { fq st[8]; int absorbed; } or_tr;
void or_tr_new(or_tr* t, const uint8_t* app_label, size_t app_len);
void or_tr_append_message(or_tr* t, const char* label, const uint8_t* msg, size_t len);
void or_tr_append_fields(or_tr* t, const char* label, const fq* fs, size_t n);
void or_tr_append_u64s(or_tr* t, const char* label, const uint64_t* us, size_t n);
fq   or_tr_challenge_field(or_tr* t, const char* label);
void or_tr_challenge_fields(or_tr* t, const char* label, size_t n, fq* out);
void or_tr_digest32(or_tr* t, uint8_t out[32]);
uint64_t or_tr_permutation_count(void);  /* process-wide count of transcript permutations */

/* ---------------- params (neo-params/src/lib.rs) ---------------- */
typedef struct {
  uint64_t q; uint32_t eta, d, kappa; uint64_t m; uint32_t b, k_rho; uint64_t B; uint32_t T, s, lambda;
} or_params;
int or_params_auto_r1cs(size_t n_rows, or_params* out);
int or_extension_check(const or_params* p, uint32_t ell, uint32_t d_sc, uint32_t* s_min, int32_t* slack_bits);

/* ---------------- Ajtai (neo-ajtai) ---------------- */
void or_decomp_b2_row_major(const fq* z, size_t m, int8_t* Z);               /* decomp.rs:118-158 */
void or_ajtai_setup(const uint8_t seed[32], size_t kappa, size_t m, fq* pp); /* commit.rs:133-156; pp[(i*m+j)*54+u] */
/* Z: balanced base-2 digits in {-1, 0, 1} (row-major [54][m]); other values are ignored. */
void or_ajtai_commit(const fq* pp, size_t kappa, size_t m, const int8_t* Z, fq* out); /* commit.rs:679-744 */
void or_s_mul(const fq* rho_coeffs, const fq* c, size_t kappa, fq* out);     /* commit.rs:458-471 */

/* ---------------- CCS (neo-ccs) ---------------- */
typedef struct {
  size_t n, m;
  int identity;           /* matrix is I_n */
  /* CSC: col_ptr[m+1], row_idx[nnz], vals[nnz] */
  size_t* col_ptr; uint32_t* row_idx; fq* vals; size_t nnz;
} or_mat;
typedef struct {
  size_t n, m, t;
  or_mat mats[8];
  /* f terms: coeff + exps[t] */
  size_t nterms;
  fq term_coeff[16];
  uint32_t term_exps[16][8];
} or_ccs;
/* R1CS (CSR arrays for A,B,C, n x m with n==m) -> identity-first CCS (r1cs.rs:13-43) */
or_ccs* or_ccs_from_r1cs(size_t n, size_t m,
                         const uint64_t* a_rp, const uint32_t* a_ci, const fq* a_v,
                         const uint64_t* b_rp, const uint32_t* b_ci, const fq* b_v,
                         const uint64_t* c_rp, const uint32_t* c_ci, const fq* c_v);
void or_ccs_free(or_ccs* s);
void or_ccs_digest(const or_ccs* s, fq out[4]);                              /* engines/utils.rs:214-321 */

/* ---------------- Pi_CCS oracle pieces (optimized_engine/oracle.rs) ---------------- */
void or_chi_table(const kx* bits, size_t nbits, kx* out);                    /* oracle.rs:1517-1532 */
/* Row-phase streaming state; opaque handle. */
typedef struct or_rowstate or_rowstate;
or_rowstate* or_rowstate_build(const or_ccs* s, const int8_t* const* Zs, size_t k_total,
                               const kx* alpha, const kx* beta_a, const kx* beta_r, kx gamma,
                               size_t ell_d, size_t ell_n, const kx* r_inputs /* may be NULL */);
void or_rowstate_free(or_rowstate* st);
/* Round coefficients (low->high, deg_max+1 entries); returns count. oracle.rs:1047-1178 */
size_t or_rowstate_round_coeffs(const or_rowstate* st, kx* coeffs_out);
void or_rowstate_evals(const or_rowstate* st, const kx* xs, size_t nx, kx* out);
void or_rowstate_fold(or_rowstate* st, kx r);                                /* oracle.rs:417-503 */
size_t or_rowstate_len(const or_rowstate* st);
/* Export the folded tables (for kernel parity): which=0 eq_beta,1 eq_rin,2.. f vars, then nc, eval */
size_t or_rowstate_export(const or_rowstate* st, int which, size_t idx, kx* out, size_t cap);

/* precompute_for_r (oracle.rs:1659-1783): y_eval[i][j][rho] for all k_total witnesses */
void or_precompute_y_par(const or_ccs* s, const int8_t* const* Zs, size_t k_total, const kx* r, size_t ell_n,
                         kx* y_eval, int parallel);
/* 1: the loops the reference runs serially (Z_mix, precompute_for_r) run on one thread */
void or_set_serial_where_serial(int on);
void or_precompute_y(const or_ccs* s, const int8_t* const* Zs, size_t k_total, const kx* r, size_t ell_n,
                     kx* y_eval /* k_total*t*54 */);
void or_vjs(const or_ccs* s, const kx* r, size_t ell_n, kx* vjs /* t*m */);

/* ---------------- sum-check helpers (neo-reductions/src/sumcheck.rs) ---------------- */
void or_interpolate(const kx* xs, const kx* ys, size_t n, kx* coeffs);     /* sumcheck.rs:79-121 */
kx   or_poly_eval(const kx* coeffs, size_t n, kx x);                        /* sumcheck.rs:48-58 */

/* ---------------- RLC / DEC (optimized_engine/common.rs, neo-reductions/src/common.rs) ---------------- */
void or_rot_from_coeffs(const int8_t* a, fq* rho /* 54x54 row-major */);   /* common.rs:370-403 */
void or_rlc_zmix(const fq* const* rhos, const int8_t* const* Zs, size_t k, size_t m, int32_t* zmix);
int  or_split_b(const int32_t* zmix, size_t len, size_t k, int8_t* const* outs, int* nonzero);

/* ---------------- shard prover (neo-fold/src/shard.rs:1725-2283) ---------------- */
/* Memory instances of one step (neo-memory/src/witness.rs:23-96): LUT tables are materialized
 * (no table_spec); a MEM init is MemInit::Zero (sparse = 0) or MemInit::Sparse pairs. */
typedef struct { size_t k, d, n_side, steps, ell; const fq* table; } or_lut_inst;
typedef struct {
  size_t k, d, n_side, steps, ell;
  int sparse;
  size_t ninit;
  const uint64_t* init_addrs;
  const fq* init_vals;
} or_mem_inst;
typedef struct { size_t n_lut; const or_lut_inst* luts; size_t n_mem; const or_mem_inst* mems; } or_step_mem;
/* output binding of a shard's final memory (neo-fold/src/output_binding.rs:23-95,
 * neo-memory/src/output_check.rs): claims[i] = (addrs[i], values[i]) over the Twist instance
 * mem_idx of the last step, whose final state has 2^num_bits entries */
typedef struct {
  size_t mem_idx, num_bits, n_claims;
  const uint64_t* addrs; /* strictly increasing, < 2^num_bits */
  const fq* values;
  const fq* final_state;
} or_ob_config;
typedef struct or_prover or_prover;
or_prover* or_prover_new(const or_ccs* s, const uint8_t pp_seed[32], int n_threads);
void or_prover_free(or_prover* p);
/* Fold `nsteps` steps.  z_steps[s] is the full z vector (length m) of step s; m_in public inputs.
 * Writes the canonical proof serialization into *out (malloc'd) and its length. */
int or_prove_shard(or_prover* p, const fq* const* z_steps, size_t nsteps, size_t m_in,
                   uint8_t** out, size_t* out_len, double* seconds_per_step);
/* The same with each step's Twist/Shout instances over the shared CPU bus (Route A); mem may
 * be NULL (no instances in any step). */
int or_prove_shard_mem(or_prover* p, const fq* const* z_steps, const or_step_mem* mem, size_t nsteps, size_t m_in,
                       uint8_t** out, size_t* out_len, double* seconds_per_step);
/* The same with output binding on the last step (fold_shard_prove_with_output_binding,
 * neo-fold/src/shard.rs:2352-2385); ob may be NULL. */
int or_prove_shard_mem_ob(or_prover* p, const fq* const* z_steps, const or_step_mem* mem, size_t nsteps,
                          size_t m_in, const or_ob_config* ob, uint8_t** out, size_t* out_len,
                          double* seconds_per_step);
void or_free(void* ptr);
/* Wall time of fold step i of the last or_prove_shard call (the MCS commits of all steps are
 * done before step 0 and counted in neither). */
double or_last_step_seconds(size_t i);
int or_num_threads(void);
const char* or_last_error(void);


/* ---------------- Twist / Shout oracles (neo-memory/src/twist_oracle.rs) ----------------
 * Sparse time columns: nnz entries (idx[i], vals[2i..2i+1] as K), over a length-2^ell_n domain.
 * All oracles share evals/fold/num_rounds/degree_bound (RoundOracle, sumcheck.rs:17-34). */
typedef struct {
  size_t nnz;
  const uint64_t* idx;
  const uint64_t* vals;
} or_sparse_in;
typedef struct or_ts or_ts;
or_ts* or_ts_product(const kx* const* factors, size_t nf, size_t len, size_t degree_bound);
or_ts* or_ts_shout_value(const kx* r_cycle, size_t ell_n, or_sparse_in has_lookup, or_sparse_in val, kx* claim);
or_ts* or_ts_index_adapter(const kx* r_cycle, size_t ell_n, or_sparse_in has_lookup, const or_sparse_in* addr_bits,
                           size_t ell_addr, const kx* r_addr, kx* claim);
or_ts* or_ts_bitness(const kx* r_cycle, size_t ell_n, const or_sparse_in* cols, const kx* weights, size_t ncols);
or_ts* or_ts_twist_read(const kx* r_cycle, size_t ell_n, or_sparse_in has_read, or_sparse_in rv,
                        const or_sparse_in* ra_bits, or_sparse_in has_write, or_sparse_in inc,
                        const or_sparse_in* wa_bits, size_t ell_addr, const kx* r_addr, kx init_at_r_addr);
or_ts* or_ts_twist_write(const kx* r_cycle, size_t ell_n, or_sparse_in has_write, or_sparse_in wv, or_sparse_in inc,
                         const or_sparse_in* wa_bits, size_t ell_addr, const kx* r_addr, kx init_at_r_addr);
or_ts* or_ts_val_eval(size_t ell_n, const or_sparse_in* wa_bits, size_t ell_addr, or_sparse_in has_write,
                      or_sparse_in inc, const kx* r_addr, const kx* r_time, kx* claim);
or_ts* or_ts_total_inc(size_t ell_n, const or_sparse_in* wa_bits, size_t ell_addr, or_sparse_in has_write,
                       or_sparse_in inc, const kx* r_addr, kx* claim);
or_ts* or_ts_read_addr(const uint64_t* init_addrs, const kx* init_vals, size_t ninit, const kx* r_cycle, size_t ell_n,
                       or_sparse_in has_read, or_sparse_in rv, const or_sparse_in* ra_bits, or_sparse_in has_write,
                       const or_sparse_in* wa_bits, or_sparse_in inc, size_t ell_addr);
or_ts* or_ts_write_addr(const uint64_t* init_addrs, const kx* init_vals, size_t ninit, const kx* r_cycle,
                        size_t ell_n, or_sparse_in has_write, or_sparse_in wv, const or_sparse_in* wa_bits,
                        or_sparse_in inc, size_t ell_addr);
or_ts* or_ts_addr_lookup(const or_sparse_in* addr_bits, size_t ell_addr, or_sparse_in has_lookup, const kx* table,
                         size_t table_len, const kx* r_cycle, size_t ell_n, kx* claim);
void or_ts_evals(or_ts* o, const kx* xs, size_t nx, kx* ys);
void or_ts_fold(or_ts* o, kx r);
size_t or_ts_num_rounds(const or_ts* o);
size_t or_ts_degree_bound(const or_ts* o);
void or_ts_free(or_ts* o);

/* ---------------- shared CPU bus (neo-memory/src/cpu/bus_layout.rs, memory_sidecar/cpu_bus.rs) ---- */
void or_bus_decode(const int8_t* Z, size_t m, fq* z);
long or_bus_time_col(const fq* z, size_t zlen, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                     size_t col_id, size_t steps_len, size_t pow2_cycle, uint64_t* idx_out, kx* vals_out);
int or_bus_openings(const int8_t* Z, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                    const kx* r, size_t ell_n, kx* y, kx* y_scalars);

#ifdef __cplusplus
}
#endif
#endif
