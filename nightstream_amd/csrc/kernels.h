// kernels.h -- launchers for the gfx950 kernels of the Neo folding hot path.
//
// All launchers are asynchronous on the given stream.  Device buffers are plain pointers;
// digit matrices are int8 row-major D x m (Z[rho*m + c]); K elements are ns::K2 (16 bytes).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "field.h"

namespace ns {

constexpr int D = 54;      // ring degree, neo-math/src/ring.rs:14
constexpr int DPAD = 64;   // 2^ell_d
constexpr int KAPPA = 16;  // Ajtai rank (Goldilocks preset)
constexpr int MAX_F_VARS = 8;
constexpr int MAX_F_TERMS = 8;
constexpr int MAX_DM = 7;  // max coefficient index of a row-phase round polynomial

#define NS_CHECK(x)                                                                              \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) throw ns::HipError(e_, #x, __FILE__, __LINE__);                        \
  } while (0)

struct HipError {
  hipError_t err;
  const char* what;
  const char* file;
  int line;
  HipError(hipError_t e, const char* w, const char* f, int l) : err(e), what(w), file(f), line(l) {}
};

// Per-kernel timing: when a prover enabled timing on the calling thread, ktimer_events hands out
// a start/stop event pair that hipExtLaunchKernelGGL records at the dispatch's own start and end
// (so host submission gaps are not counted); otherwise it returns false and the launch is plain.
enum KernelClass {
  KT_ROUND0 = 0,
  KT_ROUND,
  KT_FOLD,
  KT_AJTAI_GEMM,
  KT_YD_GEMM,
  KT_ZMIX,
  KT_SPLIT,
  KT_ROUND1,
  KT_FUSED,
  KT_N
};
bool ktimer_events(int cls, hipEvent_t* start, hipEvent_t* stop);
#define NS_LAUNCH_TIMED(cls, kernel, grid, block, shmem, stream, ...)                                   \
  do {                                                                                                \
    hipEvent_t ns_ev_a_ = nullptr, ns_ev_b_ = nullptr;                                                \
    if (ktimer_events((cls), &ns_ev_a_, &ns_ev_b_))                                                   \
      hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, ns_ev_a_, ns_ev_b_, 0, __VA_ARGS__); \
    else                                                                                              \
      hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                            \
  } while (0)

// Sparse matrix on the device: CSR (rows) for M x and CSC (columns) for M^T x.
struct DevSparse {
  uint32_t n = 0, m = 0, nnz = 0;
  bool identity = false;
  uint32_t* csr_ptr = nullptr;
  uint32_t* csr_col = nullptr;
  uint64_t* csr_val = nullptr;
  uint32_t* csc_ptr = nullptr;
  uint32_t* csc_row = nullptr;
  uint64_t* csc_val = nullptr;
  // columns with more than SPMV_HEAVY entries (a constant-one column of padding rows holds
  // hundreds of thousands): M^T chi reduces each of them with a whole workgroup
  uint32_t n_heavy = 0;
  uint32_t* heavy_cols = nullptr;
};
constexpr uint32_t SPMV_HEAVY = 1024;

// Compiled CCS polynomial f for the round kernels: sum_t coeff_t * prod_v x_{var}^{exp}.
struct FPoly {
  int nterms = 0;
  uint64_t coeff[MAX_F_TERMS];
  int nv[MAX_F_TERMS];
  int var[MAX_F_TERMS][4];
  int exp[MAX_F_TERMS][4];
  int deg_max = 4;  // max(4, deg f + 1)
};

// ---- decomposition (neo-ajtai/src/decomp.rs:118-158, b = 2 balanced) ----
void launch_decomp_b2(const uint64_t* z, size_t m, int8_t* Z, hipStream_t s);

// ---- Ajtai ----
// PP coefficients [kappa][m][54] (u64, device) -> int8 limb planes [(i*8+l)*54+u][m] (value - 128).
void launch_pp_limbs(const uint64_t* pp, size_t kappa, size_t m, int8_t* limbs, hipStream_t s);
// ChaCha8 keystream for setup(): words [n_words] from (key, counter0); see ajtai.hip.
void launch_chacha8_words(const uint32_t key[8], uint64_t first_block, size_t nblocks, uint32_t* out, hipStream_t s);
size_t ajtai_workspace_bytes(size_t m, size_t nw);
// Commit nw digit matrices (device pointers in Zs_dev, a device array of nw pointers).
// 0: timed choice of tile shape and K splits; 1 / 2: force the 16x16x64 / 32x32x32 tiles
void ajtai_force_variant(int v);
void ajtai_commit(const int8_t* limbs, size_t kappa, size_t m, const int8_t* const* Zs_dev, size_t nw,
                  uint64_t* out /* nw x kappa x 54 */, void* workspace, hipStream_t s);

// ---- chi tables, matrix-vector products ----
void launch_chi(const K2* bits, int nbits, K2* out, hipStream_t s);  // len 2^nbits
// chi of bits[j+1 .. ell) for every j in [j0, ell), tables concatenated (2^(ell - j - 1) entries
// each, total entries); ntab = 2 also builds the tables of bits[64 + ..] at out + total
void launch_chi_suffix(const K2* bits, int ell, int j0, size_t total, int ntab, K2* out, hipStream_t s);
void launch_z1(const int8_t* Z1, size_t m, K2* z1, hipStream_t s);
// y[r] = coeff * sum_k val[k] x[col[k]] (+ y[r] if accumulate), r < n
void launch_spmv(const DevSparse& M, const K2* x, K2* y, K2 coeff, bool accumulate, hipStream_t s);
// v[c] = sum_{k in col c, row < n_eff} val[k] chi[row]
void launch_spmv_t(const DevSparse& M, const K2* chi, size_t n_eff, K2* v, hipStream_t s);
// S[c] = sum_{w, rho} W[w*54+rho] * Z_w[rho][c]
void launch_weighted_digit_colsum(const int8_t* const* Zs_dev, size_t nw, const K2* W, size_t m, K2* S, void* ws,
                                  hipStream_t s);
size_t colsum_ws_bytes(size_t nw, size_t m);
// y[w][j][rho] = sum_c Z_w[rho][c] v_j[c]  (w < nw, j < t); partial workspace sized by ydigits_ws_bytes
size_t ydigits_ws_bytes(size_t nw, size_t t, size_t m);
void launch_ydigits(const int8_t* const* Zs_dev, size_t nw, const K2* v, size_t t, size_t m, K2* y, void* ws,
                    hipStream_t s);

// ---- Pi_CCS row phase ----
struct RowTables {
  size_t len;         // current table length L (power of two)
  const K2* eqb;      // [L]
  const K2* eqr;      // [L] or null
  const K2* ev;       // [L] or null
  const K2* fv;       // [nf][L]
  int nf;
  const K2* nc;       // [k*54][L] (rounds >= 1)
  const int8_t* const* digits;  // round 0: k device pointers to D x m digit matrices
  size_t m;           // digit row stride (round 0)
  int k;
  K2 gamma_k;
  bool digit_round1 = false;  // round 1 evaluated from the digits (tables folded once, by r0)
  bool digit_round2 = false;  // round 2 evaluated from the digits (int8 MFMA path only)
  // coefficient polynomials (X^0..X^3) of the digit monomials: [4][4] in round 0 (MFMA path),
  // [20][4] in round 1, [120][4] in round 2
  const K2* pc = nullptr;
  // rounds 0-2 on int8 MFMA: the NC weights as balanced base-256 limbs, [16][64 k] int8
  // (weight_limbs), and device scratch for the weighted digit-monomial sums (4, 20 or 120 x len / 2
  // K values in rounds 0, 1, 2); without them rounds 0 and 1 run on the vector ALUs and round 2
  // on the K tables
  const int8_t* wl = nullptr;
  K2* ubuf = nullptr;
  size_t ubuf_bytes = 0;
};
// Balanced base-256 limbs of the NC weights wg [k*54] for the digit-monomial MFMA: limb row
// 8 c + l (component c, byte l) of [16][64 k]; weight w*54 + rho sits in column 64 w + rho (the
// 10 columns after each instance's 54 are zero).  Each component is taken as the representative
// v of its class mod q with |v| <= 128 (2^64 - 1) / 255, so that v = sum_l 256^l s_l exactly
// with every s_l in [-128, 127].
std::vector<int8_t> weight_limbs(const K2* wg, size_t k);
size_t round_ws_bytes();
// Optional ready word for a round's coefficients in coherent pinned host memory: the partial-sum
// kernel stores the coefficients, then seq into *flag (system-scope release).
struct RoundSignal {
  uint64_t* flag = nullptr;
  uint64_t seq = 0;
};
// Writes the deg_max+1 coefficients (low->high) of the round polynomial to out (device).
void launch_round_eval(const RowTables& T, const FPoly& f, const K2* wg /* [k*54] device */, bool round0,
                       K2* out, void* ws, hipStream_t s, const RoundSignal* sig = nullptr);
// Rounds >= 3: fold the NC tables ([k*54][2 len] -> [k*54][len]) and the f-variable and Eval
// scalar tables by r (the eq tables are not read: their factors come from gb / gr and the host)
// and write P0..P4, Q0, Q1 to out: P(X) = sum_t gb[t] (f_t + NC_t)(X), Q(X) = sum_t gr[t] Eval_t(X)
// over the pairs t of the folded tables.  One fused fold + evaluate launch and the partial sum.
// nsc <= 8.
struct FusedArgs {
  size_t len;  // table length after the fold
  const K2* sc_src;
  K2* sc_dst;
  int nsc, nf;
  bool has_eval;
  const K2* nc_src;
  K2* nc_dst;
  int k;
  K2 gamma_k;
  FPoly f;
  K2 r;
  uint64_t r7;            // 7 r.c1
  const K2* gb = nullptr;  // [len / 2]: chi of beta over the variables after this round's
  const K2* gr = nullptr;  // [len / 2]: the same for r_in (has_eval)
  // set from f by fused_job: its degree (at most 4) and, for f = c0 x_a x_b + c1 x_c (the R1CS
  // shape), the variables a, b, c (r1cs[0] < 0 otherwise) and c0, c1
  int fdeg = 4;
  int r1cs[3] = {-1, -1, -1};
  uint64_t r1cs_c[2] = {0, 0};
  int split = 0;  // set by fused_job for short rounds: NC rows and the R1CS f item split by evaluation point
};
constexpr int FUSED_MAX_SC = 8;
void launch_round_fused(const FusedArgs& A, const K2* wg, K2* out, void* ws, hipStream_t s,
                        const RoundSignal* sig = nullptr);
// One session's fused round as a self-contained job (its arguments, weights, partial-sum scratch,
// coefficient mailbox and ready word), so that the rounds of concurrent sessions can share a
// launch: launch_fused_jobs runs up to FUSED_MAX_JOBS jobs (the array in device memory) in one
// fused launch and one partial-sum launch; every job's coefficients and ready word land as in
// its own launch.
struct FusedJob {
  FusedArgs A;
  const K2* wg;
  K2* partials;
  K2* out;
  uint64_t* flag;
  uint64_t seq;
  int G, nblk;
};
constexpr int FUSED_MAX_JOBS = 32;
FusedJob fused_job(const FusedArgs& A, const K2* wg, K2* out, void* ws, const RoundSignal& sig);
void launch_fused_job(const FusedJob& J, hipStream_t s);
void launch_fused_jobs(const FusedJob* jobs_dev, int njobs, size_t total_blocks, hipStream_t s);
// Fold every table with challenge r: dst[i] = src[2i] + r (src[2i+1] - src[2i]).
void launch_fold_k(const K2* src, size_t rows, size_t len, K2* dst, K2 r, hipStream_t s);
void launch_fold_digits2(const int8_t* const* digits_dev, int k, size_t m, size_t len, K2* dst, K2 r0, K2 r1,
                         hipStream_t s);
void launch_fold_digits(const int8_t* const* digits_dev, int k, size_t m, size_t len, K2* dst, K2 r, hipStream_t s);
// Tables after rounds 0, 1 and 2 straight from the digits (round 2 evaluated from them too):
// T3[i] = sum_b chi_(r0, r1, r2)(b) D[8 i + b], [k*54][len / 8] for the digit rows of length len.
void launch_fold_digits3(const int8_t* const* digits_dev, int k, size_t m, size_t len, K2* dst, K2 r0, K2 r1, K2 r2,
                         hipStream_t s);
// Whether round 2 of a table of `len` entries (len / 2 pairs) runs from the digits on int8 MFMA
// (launch_round_eval with digit_round2): its block size and the instance limit
bool digit_round2_fits(size_t len, int k);

// ---- Pi_RLC / Pi_DEC ----
// zmix[r][c] = sum_i sum_kk rho_i[r][kk] Z_i[kk][c]; rho given as int8 [k][54][54]
// rho_dev: the k rotation matrices as int8 [k][64][64] (row r, column kk; zero beyond 54)
void launch_rlc_zmix(const int8_t* rho_dev, const int8_t* const* Zs_dev, size_t k, size_t m, int32_t* zmix,
                     hipStream_t s);
// out[e] = sum over K splits of part[sp * n + e] (int32)
void launch_sum_splits(const int32_t* part, size_t splits, size_t n, int32_t* out, hipStream_t s);
// Gather the first m_in columns of up to 16 digit planes [54][m]: out[(p*54 + rho)*m_in + c].
struct PlanePtrs {
  const int8_t* p[16];
};
void launch_head_cols(const PlanePtrs& planes, int np, size_t m, int m_in, int8_t* out, hipStream_t s);

// ---- shared CPU bus (bus.hip) ----
// w[j] = chi_r(m_in + j) for j < chunk (r: ell K values on the device)
void launch_bus_weights(uint32_t m_in, uint32_t chunk, const K2* r_dev, int ell, K2* w, hipStream_t s);
// out[i] = decoded z[base + i] from the digit planes Z [54][m]
void launch_bus_decode_range(const int8_t* Z, size_t m, size_t base, size_t len, uint64_t* out, hipStream_t s);
// bus openings of np digit planes: y[(p*bus_cols + col)*64 + rho], ys[p*bus_cols + col]
void launch_bus_open_planes(const PlanePtrs& planes, int np, size_t m, size_t bus_base, uint32_t chunk,
                            uint32_t bus_cols, const K2* w, K2* y, K2* ys, hipStream_t s);
// the same for one int32 matrix [54][m] (Z_mix)
void launch_bus_open_i32(const int32_t* Z, size_t m, size_t bus_base, uint32_t chunk, uint32_t bus_cols, const K2* w,
                         K2* y, K2* ys, hipStream_t s);
// balanced base-2 digits of zmix into k_dec planes (by value: no pointer table in device memory);
// flags[0] is cleared by the launch, then holds the OR of every |entry|: plane i is nonzero iff
// bit i is set, and a bit at or above k_dec means an entry out of range (dec_out_of_range)
constexpr int DEC_MAX_PLANES = 30;
struct DecPlanes {
  int8_t* p[DEC_MAX_PLANES];
};
void launch_dec_split(const int32_t* zmix, size_t len, int k_dec, const DecPlanes& outs, int* flags, hipStream_t s);
inline bool dec_out_of_range(uint32_t mask, int k_dec) { return (mask >> k_dec) != 0; }
inline bool dec_plane_nonzero(uint32_t mask, int i) { return ((mask >> i) & 1u) != 0; }
// The current device's address of the DEC split's static scratch (counters), for the
// no-allocation entry point.
int* dec_noalloc_flags();

}  // namespace ns
