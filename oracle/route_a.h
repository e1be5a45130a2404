/* route_a.h -- types of the CPU restatement of the Route A memory sidecar (route_a.c).
 * TEST INFRASTRUCTURE ONLY (the checker of the MI355X library). */
#ifndef ROUTE_A_H
#define ROUTE_A_H
#include "neo_oracle.h"

#define RA_MAX_INST 64
#define RA_MAX_ELL 64

typedef struct { size_t addr0, ell, has_lookup, val; } ra_shout_lay;
typedef struct { size_t ra0, wa0, ell, has_read, has_write, wv, rv, inc; } ra_twist_lay;
/* BusLayout (neo-memory/src/cpu/bus_layout.rs:29-56) */
typedef struct {
  size_t m, m_in, chunk, bus_cols, bus_base, n_shout, n_twist;
  ra_shout_lay sh[RA_MAX_INST];
  ra_twist_lay tw[RA_MAX_INST];
} ra_layout;

/* a decoded sparse time column (SparseIdxVec, neo-memory/src/sparse_time.rs:10-24) */
typedef struct { size_t nnz; uint64_t* idx; kx* vals; } ra_col;
typedef struct { size_t ell; ra_col addr[RA_MAX_ELL], has_lookup, val; } ra_shout_cols;
typedef struct { size_t ell; ra_col ra[RA_MAX_ELL], wa[RA_MAX_ELL], has_read, has_write, wv, rv, inc; } ra_twist_cols;

enum { RA_ROWSTATE = 0, RA_TS = 1 };
/* one claim of a batched sum-check (BatchedClaim, sumcheck.rs:315-324) */
typedef struct {
  int kind;
  void* o;
  size_t deg, nrounds;
  const char* label;
  int dynamic;
  kx claimed_sum, running;
  kx* polys; /* nrounds x (deg + 1) coefficients, low to high */
} ra_claim;

typedef struct {
  size_t n_lut, ell_addr, n_active;
  uint64_t active_mask;
  ra_shout_cols* cols;
  kx* claimed;   /* n_lut */
  kx* r_addr;    /* ell_addr */
  kx** rounds;   /* per active claim: ell_addr x (deg + 1) */
  size_t* degs;
} ra_shout_pre;

typedef struct {
  ra_twist_cols cols;
  size_t ell_addr;
  kx* r_addr;
  kx* rounds[2];
  size_t degs[2];
  kx read_claim, write_claim;
} ra_twist_pre;

typedef struct {
  size_t n_mem, per;
  ra_claim* claims; /* per * n_mem, oracles already freed, polys kept */
  kx* r_val;
} ra_val_pre;

void ra_msg_u64(or_tr* t, const char* label, uint64_t v);
int ra_infer_layout(const or_step_mem* mem, size_t nsteps, size_t n, size_t m, size_t m_in, ra_layout* L, char* err,
                    size_t errlen);
or_sparse_in ra_in(const ra_col* c);
void ra_col_free(ra_col* c);
int ra_decode_twist(const fq* z, const ra_layout* L, size_t i, size_t steps, size_t ell_n, ra_twist_cols* T);
void ra_twist_cols_free(ra_twist_cols* T);
void ra_absorb_step_memory(or_tr* tr, const or_step_mem* sm);
void ra_claim_evals(ra_claim* c, const kx* xs, size_t n, kx* ys);
int ra_batched_sumcheck(or_tr* tr, const char* domain, size_t inst_idx, ra_claim* cl, size_t n, size_t rounds,
                        kx* chals, char* err, size_t errlen);
void ra_claim_free(ra_claim* c);
int ra_shout_addr_pre(or_tr* tr, const or_step_mem* sm, const ra_layout* L, const fq* zdec, size_t ell_n,
                      const kx* r_cycle, size_t step_idx, ra_shout_pre* out, char* err, size_t errlen);
void ra_shout_pre_free(ra_shout_pre* s);
int ra_twist_addr_pre(or_tr* tr, const or_step_mem* sm, const ra_layout* L, const fq* zdec, size_t ell_n,
                      const kx* r_cycle, ra_twist_pre* out, char* err, size_t errlen);
void ra_twist_pre_free(ra_twist_pre* P);
int ra_build_time_claims(const or_step_mem* sm, size_t ell_n, const kx* r_cycle, const ra_shout_pre* sp,
                         const ra_twist_pre* tp, ra_claim* cl, size_t* ncl, char* err, size_t errlen);
void ra_bind_dynamic_claims(or_tr* tr, const ra_claim* cl, size_t n);
int ra_val_eval(or_tr* tr, const or_step_mem* sm, const ra_twist_pre* tp, const ra_twist_cols* prev, size_t ell_n,
                const kx* r_time, size_t step_idx, ra_val_pre* out, char* err, size_t errlen);
void ra_val_pre_free(ra_val_pre* v);
/* output sum-check of output binding (output_check.rs:252-640): absorbs the claims, samples
 * r_addr, proves sum_k eq(r_addr, k) mask(k) (final(k) - claimed(k)) = 0 over num_bits rounds;
 * rounds: num_bits x 4 coefficients, r_prime: the round challenges */
int ra_output_sumcheck(or_tr* tr, const or_ob_config* ob, kx* rounds, kx* r_prime, char* err, size_t errlen);
int ra_bus_openings_i32(const int32_t* Z, const ra_layout* L, const kx* r, size_t ell_n, kx* y, kx* ys);
#endif
