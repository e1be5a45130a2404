// ajtai.hip -- balanced decomposition, Ajtai public-parameter generation and the Ajtai
// commitment as an exact int8 MFMA contraction on gfx950.
//
// Commitment (neo-ajtai/src/commit.rs:679-744): for every Ajtai row i < kappa,
//   c_i = cf( sum_j a_ij * zhat_j  mod Phi_81 ),  zhat_j = sum_t Z[t][j] X^t,  Z[t][j] in {-1,0,1}.
// Every a_ij coefficient (u64 < q) is split into 8 byte limbs stored with a -128 offset so they
// fit signed int8.  The polynomial product sum is then
//   P_i[s] = sum_l 256^l sum_{u+t=s} ( G'_l[u][t] + 128 S[t] ),
//   G'_l[u][t] = sum_j (limb_l(a_ij[u]) - 128) Z[t][j],   S[t] = sum_j Z[t][j],
// a GEMM with M = kappa*8*54 limb rows, N = 54 * (#witnesses) digit rows and K = m, computed
// with v_mfma_i32_16x16x64_i8 and int32 accumulation (|G'| <= 128 m < 2^31 for m <= 2^23).
// The epilogue sums anti-diagonals in int64, combines limbs in F_q and reduces mod
// Phi_81 = X^54 + X^27 + 1 using X^81 = 1 and X^54 = -X^27 - 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "kernels.h"

namespace ns {

// ---------------- decomposition ----------------
__global__ void k_decomp_b2(const uint64_t* __restrict__ z, size_t m, int8_t* __restrict__ Z) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  int64_t a = to_balanced(z[c]);
  for (int row = 0; row < D; row++) {
    int64_t dg = (a & 1) == 0 ? 0 : (a >= 0 ? 1 : -1);
    Z[(size_t)row * m + c] = (int8_t)dg;
    a = (a - dg) >> 1;
  }
}

void launch_decomp_b2(const uint64_t* z, size_t m, int8_t* Z, hipStream_t s) {
  if (m == 0) return;
  hipLaunchKernelGGL(k_decomp_b2, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, z, m, Z);
}

// ---------------- PP limbs ----------------
// Row (i*8 + l)*54 + u of the 6912-row limb matrix holds byte_l(pp[(i*m + j)*54 + u]) - 128 at
// column j.  Stored K-step-major: limbs[((j / 64) * 6912 + row) * 64 + j % 64], so the 64-column
// slice of every row for one K step of the commit GEMM is one contiguous 442 KB region.
__global__ void k_pp_limbs(const uint64_t* __restrict__ pp, size_t kappa, size_t m, int8_t* __restrict__ limbs) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t i = blockIdx.y;
  if (j >= m) return;
  const uint64_t* a = pp + (i * m + j) * D;
  int8_t* col = limbs + (j >> 6) * (size_t)(KAPPA * 8 * D) * 64 + (j & 63);
  for (int u = 0; u < D; u++) {
    uint64_t v = a[u];
    for (int l = 0; l < 8; l++) {
      int b = (int)((v >> (8 * l)) & 0xFF) - 128;
      col[((i * 8 + l) * D + u) * 64] = (int8_t)b;
    }
  }
}

void launch_pp_limbs(const uint64_t* pp, size_t kappa, size_t m, int8_t* limbs, hipStream_t s) {
  hipLaunchKernelGGL(k_pp_limbs, dim3((unsigned)((m + 255) / 256), (unsigned)kappa), dim3(256), 0, s, pp, kappa, m,
                     limbs);
}

// ---------------- ChaCha8 keystream (rand_chacha block layout) ----------------
__device__ __forceinline__ uint32_t rotl32(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
#define NS_QR(a, b, c, d) \
  a += b; d ^= a; d = rotl32(d, 16); c += d; b ^= c; b = rotl32(b, 12); \
  a += b; d ^= a; d = rotl32(d, 8);  c += d; b ^= c; b = rotl32(b, 7);

struct ChaKey {
  uint32_t k[8];
};

__global__ void k_chacha8_words(ChaKey key, uint64_t first_block, size_t nblocks, uint32_t* __restrict__ out) {
  size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  uint64_t ctr = first_block + b;
  uint32_t in[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574, key.k[0], key.k[1], key.k[2], key.k[3],
                     key.k[4],   key.k[5],   key.k[6],   key.k[7],   (uint32_t)ctr, (uint32_t)(ctr >> 32), 0, 0};
  uint32_t x[16];
#pragma unroll
  for (int i = 0; i < 16; i++) x[i] = in[i];
#pragma unroll
  for (int r = 0; r < 8; r += 2) {
    NS_QR(x[0], x[4], x[8], x[12]);
    NS_QR(x[1], x[5], x[9], x[13]);
    NS_QR(x[2], x[6], x[10], x[14]);
    NS_QR(x[3], x[7], x[11], x[15]);
    NS_QR(x[0], x[5], x[10], x[15]);
    NS_QR(x[1], x[6], x[11], x[12]);
    NS_QR(x[2], x[7], x[8], x[13]);
    NS_QR(x[3], x[4], x[9], x[14]);
  }
#pragma unroll
  for (int i = 0; i < 16; i++) out[b * 16 + i] = x[i] + in[i];
}

void launch_chacha8_words(const uint32_t key[8], uint64_t first_block, size_t nblocks, uint32_t* out, hipStream_t s) {
  ChaKey k;
  for (int i = 0; i < 8; i++) k.k[i] = key[i];
  hipLaunchKernelGGL(k_chacha8_words, dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, s, k, first_block,
                     nblocks, out);
}

// ---------------- Ajtai commit ----------------
typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int AJ_MT = 3;   // 16-row M subtiles per wave
constexpr int AJ_NS = 11;  // 16-column N subtiles per wave
// Block tile 192 limb rows x 352 digit rows, 8 waves as 4 (M) x 2 (N), each wave 48 x 176: per
// 64-wide K step a block stages 34 KB into LDS for 264 MFMAs (a 96 x 704 tile, which reads each
// PP byte once instead of twice, stages 51 KB and measured slower).  The blocks resident on one
// XCD (see xcd_remap) work on one K range at a time, so the digit tile of a K step is served from
// that XCD's L2.  A and B slices go through double-buffered LDS (pitch 80 B: the 16 rows of a
// fragment read hit all 64 banks once); global loads run two K steps ahead through two register
// sets.
constexpr int AJ_BM = 192, AJ_BN = 352, AJ_BK = 64, AJ_LDP = 80;  // LDS row pitch (bytes)
constexpr int AJ_THREADS = 512;
constexpr int AJ_NBLK = AJ_BN;                   // digit rows per N group
constexpr int AJ_MROWS = KAPPA * 8 * D;          // 6912 = 36 x 192
constexpr int AJ_MBLK = AJ_MROWS / AJ_BM;        // 36
constexpr int AJ_ACH = AJ_BM * 4, AJ_BCH = AJ_BN * 4;  // 16-byte chunks per K step: 768, 1408
constexpr int AJ_ASL = (AJ_ACH + AJ_THREADS - 1) / AJ_THREADS;  // A chunks per thread
constexpr int AJ_BSL = (AJ_BCH + AJ_THREADS - 1) / AJ_THREADS;  // B chunks per thread

static size_t aj_groups(size_t nw) { return (nw * D + AJ_NBLK - 1) / AJ_NBLK; }
// K splits: one block per CU at a time (87 KB of LDS), so pick the split count that wastes the
// least of the last round of 256 blocks: minimise ceil(blocks / 256) / splits (fewest splits on
// ties, since every split adds a partial-sum plane), with at least 16 K steps per split.
static size_t aj_splits(size_t m, size_t nw) {
  const size_t nsteps = m / AJ_BK, per = (size_t)AJ_MBLK * aj_groups(nw);
  size_t best = 1;
  double best_t = 1e30;
  for (size_t sp = 1; sp <= 16 && sp * 16 <= nsteps; sp++) {
    const double t = (double)((per * sp + 255) / 256) / (double)sp;
    if (t < best_t - 1e-12) {
      best_t = t;
      best = sp;
    }
  }
  return best;
}

// Row-major digit rows a K step apart sit m bytes apart: at m >= 2^18 one K step of a 64-witness
// batch touches over a thousand 2 MB pages, so the digits are first copied to K-step-major order.
static bool aj_kmajor(size_t m) { return m >= ((size_t)1 << 18); }

// Largest K-split count the tuner tries: at most 16, with at least 16 K steps per split.
static size_t aj_max_splits(size_t m) { return std::max<size_t>(1, std::min<size_t>(16, (m / AJ_BK) / 16)); }

static size_t a3_groups_host(size_t nw) { return (nw * D + 255) / 256; }
static size_t aj_npad_max(size_t nw) { return std::max(aj_groups(nw) * AJ_NBLK, a3_groups_host(nw) * 256); }

// Workspace: partial-sum planes for up to aj_max_splits(m) splits | S (digit row sums) | the
// split-sum plane | K-step-major int8 digits (m >= 2^18) | packed 2-bit digits; every region
// 256-byte aligned and sized for either tile shape, so any tuned choice fits.
struct AjWs {
  size_t part, S, red, bt, bp, total;
};
static size_t aj_up256(size_t x) { return (x + 255) & ~(size_t)255; }
static AjWs aj_ws_layout(size_t m, size_t nw) {
  const size_t npad = aj_npad_max(nw), sp = aj_max_splits(m), plane = (size_t)AJ_MROWS * npad * 4;
  AjWs w;
  w.part = 0;
  w.S = aj_up256(sp * plane);
  w.red = aj_up256(w.S + nw * D * 4);
  w.bt = aj_up256(w.red + (sp > 1 ? plane : 0));
  w.bp = aj_up256(w.bt + (aj_kmajor(m) ? nw * D * m : 0));
  w.total = w.bp + nw * D * (m / 4) + 256;
  return w;
}
size_t ajtai_workspace_bytes(size_t m, size_t nw) { return aj_ws_layout(m, nw).total; }

// Workgroups go to the 8 XCDs round-robin (hardware block b runs on XCD b % 8).  The remap gives
// each XCD a contiguous run of logical tiles, ordered (K split, M tile, N group), so the blocks
// an XCD runs together share a K range.
__device__ inline unsigned xcd_remap(unsigned b, unsigned nb) {
  const unsigned per = (nb + 7) / 8, tall = (nb % 8) ? nb % 8 : 8;
  const unsigned x = b % 8, l = b / 8;
  return x < tall ? x * per + l : tall * per + (x - tall) * (per - 1) + l;
}

// 1-D grid of 36 x groups x S blocks; split z covers K steps [z nsteps / S, (z + 1) nsteps / S).
// Digits come row-major (Zs[w] + t m, a K step 64 bytes further) or, when Bt is set, K-step-major
// (Bt[(step * nw * 54 + w * 54 + t) * 64], see k_digits_kmajor).
__global__ __launch_bounds__(AJ_THREADS) void k_ajtai_gemm(const int8_t* __restrict__ limbs, size_t nsteps,
                                                          const int8_t* const* __restrict__ Zs,
                                                          const int8_t* __restrict__ Bt, size_t nw, size_t m,
                                                          size_t npad, int groups, int S,
                                                          int32_t* __restrict__ part) {
  typedef const __attribute__((address_space(1))) v4i gv4i;
  __shared__ __align__(16) int8_t As[2][AJ_BM * AJ_LDP];
  __shared__ __align__(16) int8_t Bs[2][AJ_BN * AJ_LDP];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int WMW = AJ_BM / 48;  // waves along M
  const int wm = wave % WMW, wn = wave / WMW;
  const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned per_split = (unsigned)AJ_MBLK * (unsigned)groups;
  const unsigned split = lid / per_split, rem = lid % per_split;
  const size_t m0 = (size_t)(rem / groups) * AJ_BM;
  const size_t n0 = (size_t)(rem % groups) * AJ_BN;
  const size_t s_begin = (size_t)split * nsteps / S, s_end = ((size_t)split + 1) * nsteps / S;
  const int kq = (lane >> 4) * 16, r16 = lane & 15;
  // staging: A chunk tid (valid below 384), B chunks tid + 512 c (valid below 2816 and for
  // digit rows of real witnesses); invalid slots load a valid address so every thread issues
  // the same loads, and store nothing (A) or zeros (B)
  // Chunk e of a K step is row crow(e), 16-byte column (e & 3): four lanes read one row's 64
  // contiguous bytes, and a 16-lane group writes rows {g, g+4, g+8, g+12} of its 16-row unit,
  // whose 80-byte LDS rows start on 16 distinct 4-bank groups (conflict-free stores).
  auto crow = [](int e) { return (e >> 6) * 16 + ((e >> 2) & 3) * 4 + ((e >> 4) & 3); };
  constexpr size_t ASTEP = (size_t)AJ_MROWS * 64;  // bytes per K step of the limb matrix
  const int8_t* asrc[AJ_ASL];
  int adst[AJ_ASL];
#pragma unroll
  for (int c = 0; c < AJ_ASL; c++) {
    const int chunk = tid + c * AJ_THREADS, ok = chunk < AJ_ACH;
    asrc[c] = limbs + (ok ? (m0 + crow(chunk)) * 64 + (chunk & 3) * 16 : 0);
    adst[c] = ok ? crow(chunk) * AJ_LDP + (chunk & 3) * 16 : -1;
  }
  const int8_t* bsrc[AJ_BSL];
  int bdst[AJ_BSL];
  bool bok[AJ_BSL];
#pragma unroll
  for (int c = 0; c < AJ_BSL; c++) {
    const int chunk = tid + c * AJ_THREADS, inb = chunk < AJ_BCH;
    const int row = inb ? crow(chunk) : 0, col = (chunk & 3) * 16;
    const size_t nc = n0 + row, w = nc / D, t = nc % D;
    bok[c] = inb && w < nw;
    if (Bt)
      bsrc[c] = Bt + (bok[c] ? nc * 64 : 0) + col;
    else
      bsrc[c] = bok[c] ? Zs[w] + t * m + col : Zs[0] + col;
    bdst[c] = inb ? row * AJ_LDP + col : -1;
  }
  const size_t bstep = Bt ? nw * D * 64 : AJ_BK;  // bytes from one K step of a digit row to the next
  v4i ra0[AJ_ASL], rb0[AJ_BSL], ra1[AJ_ASL], rb1[AJ_BSL];
#define AJ_GLOAD(RA, RB, st)                                                                \
  {                                                                                         \
    _Pragma("unroll") for (int c = 0; c < AJ_ASL; c++) RA[c] = *(gv4i*)(asrc[c] + (size_t)(st) * ASTEP); \
    _Pragma("unroll") for (int c = 0; c < AJ_BSL; c++) RB[c] = *(gv4i*)(bsrc[c] + (size_t)(st) * bstep); \
  }
#define AJ_LSTORE(RA, RB, buf)                                                              \
  {                                                                                         \
    _Pragma("unroll") for (int c = 0; c < AJ_ASL; c++) if (adst[c] >= 0) *(v4i*)(&As[buf][adst[c]]) = RA[c]; \
    _Pragma("unroll") for (int c = 0; c < AJ_BSL; c++)                                      \
      if (bdst[c] >= 0) *(v4i*)(&Bs[buf][bdst[c]]) = bok[c] ? RB[c] : v4i{0, 0, 0, 0};      \
  }

  v4i acc[AJ_MT][AJ_NS];
#pragma unroll
  for (int mt = 0; mt < AJ_MT; mt++)
#pragma unroll
    for (int ns_ = 0; ns_ < AJ_NS; ns_++) acc[mt][ns_] = v4i{0, 0, 0, 0};
  auto compute = [&](int buf) {
    const int8_t* A = As[buf] + (wm * 48 + r16) * AJ_LDP + kq;
    const int8_t* B = Bs[buf] + (wn * 176 + r16) * AJ_LDP + kq;
    v4i a[AJ_MT];
#pragma unroll
    for (int mt = 0; mt < AJ_MT; mt++) a[mt] = *(const v4i*)(A + mt * 16 * AJ_LDP);
#pragma unroll
    for (int ns_ = 0; ns_ < AJ_NS; ns_++) {
      const v4i b = *(const v4i*)(B + ns_ * 16 * AJ_LDP);
#pragma unroll
      for (int mt = 0; mt < AJ_MT; mt++)
        acc[mt][ns_] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[mt], b, acc[mt][ns_], 0, 0, 0);
    }
  };

  // Steps past s_end - 1 re-read step s_end - 1 (loaded, never stored).  The loop body has no
  // exit between a prefetch and its store, and the scheduling barriers keep each prefetch issued
  // before the MFMAs of the step it overlaps.
  const size_t last = s_end - 1;
  AJ_GLOAD(ra0, rb0, s_begin);
  AJ_GLOAD(ra1, rb1, min(s_begin + 1, last));
  AJ_LSTORE(ra0, rb0, 0);
  __syncthreads();
  size_t st = s_begin;
  for (; st + 2 < s_end; st += 2) {
    // step st from buffer 0 (set 1 holds st + 1): prefetch st + 2 into set 0
    AJ_GLOAD(ra0, rb0, st + 2);
    __builtin_amdgcn_sched_barrier(0);
    compute(0);
    AJ_LSTORE(ra1, rb1, 1);
    __syncthreads();
    // step st + 1 from buffer 1 (set 0 holds st + 2): prefetch st + 3 into set 1
    AJ_GLOAD(ra1, rb1, min(st + 3, last));
    __builtin_amdgcn_sched_barrier(0);
    compute(1);
    AJ_LSTORE(ra0, rb0, 0);
    __syncthreads();
  }
  // one or two steps left: st in buffer 0, st + 1 (if any) in set 1
  compute(0);
  if (st + 1 < s_end) {
    AJ_LSTORE(ra1, rb1, 1);
    __syncthreads();
    compute(1);
  }
#undef AJ_GLOAD
#undef AJ_LSTORE

  // D[row][col]: col = lane % 16, row = 4 * (lane / 16) + v
  int32_t* base = part + ((size_t)split * AJ_MROWS + m0 + wm * 48) * npad + n0 + wn * 176;
#pragma unroll
  for (int mt = 0; mt < AJ_MT; mt++)
#pragma unroll
    for (int ns_ = 0; ns_ < AJ_NS; ns_++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        size_t row = mt * 16 + (lane >> 4) * 4 + v;
        base[row * npad + ns_ * 16 + r16] = acc[mt][ns_][v];
      }
}

// ---- 32 x 32 x 32 variant: 4 waves of 128 x 128 (4 x 4 MFMA tiles each), direct-to-LDS loads ----
// Per 64-wide K step a block stages A (256 limb rows) and B (256 digit rows), 64 bytes per row,
// with global_load_lds_dwordx4: each load instruction fills 1 KB of LDS, lane j the 16 bytes at
// +16 j, i.e. row j / 4, slot j % 4.  Row r's 16-byte chunk c goes to slot c ^ swz(r), so the 16
// rows one fragment read touches per phase land on distinct banks (for 32- and 64-bank phases).
// Three stages rotate, two steps of loads in flight: at step s every wave waits for its own loads
// of step s, a bare barrier publishes them (and retires step s - 1's reads), step s + 2 is
// issued, then step s's 16 fragments are read and its 32 MFMAs issued (the second K half's
// reads interleaved with the first half's MFMAs).  Per step and
// wave: 16 ds_read_b128 fragments for 32 MFMAs of 32 x 32 x 32 (1 M MACs per 16 KB of LDS reads,
// against 0.54 M per 14 KB for the 48 x 176 tile of 16 x 16 x 64).  B rows past the last digit row
// are never loaded: their slots are zeroed once.  What bounds it (m = 2^20, 64 witnesses): the
// loads with their barriers alone take 14.3 ms and the fragment reads with the MFMAs alone 15.4 ms,
// together 23.3 ms (2.1 POP/s; the MFMA pipes alone sustain 4.35 POP/s) -- the step's 32 KB of
// staging (about 14 TB/s across the GPU, 62% L2 hits) and its LDS traffic overlap the MFMAs only
// in part.
typedef int v16i __attribute__((ext_vector_type(16)));
constexpr int A3_BM = 256, A3_BN = 256, A3_THREADS = 256, A3_STAGES = 3;
constexpr int A3_ROWS = A3_BM + A3_BN, A3_STAGE_BYTES = A3_ROWS * 64;
constexpr int A3_MBLK = AJ_MROWS / A3_BM;  // 27
constexpr int A3_LOADS = A3_ROWS * 4 / 64 / (A3_THREADS / 64);  // load instructions per wave per step: 8
static size_t a3_groups(size_t nw) { return (nw * D + A3_BN - 1) / A3_BN; }
__device__ __forceinline__ int a3_swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

__global__ __launch_bounds__(A3_THREADS) void k_ajtai_gemm32(const int8_t* __restrict__ limbs, size_t nsteps,
                                                            const int8_t* const* __restrict__ Zs,
                                                            const int8_t* __restrict__ Bt, size_t nw, size_t m,
                                                            size_t npad, int groups, int S,
                                                            int32_t* __restrict__ part) {
  __shared__ __align__(16) int8_t lds[A3_STAGES * A3_STAGE_BYTES];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned per_split = (unsigned)A3_MBLK * (unsigned)groups;
  const unsigned split = lid / per_split, rem = lid % per_split;
  const size_t m0 = (size_t)(rem / groups) * A3_BM;
  const size_t n0 = (size_t)(rem % groups) * A3_BN;
  const size_t s_begin = (size_t)split * nsteps / S, s_end = ((size_t)split + 1) * nsteps / S;
  const size_t nrows_b = nw * D;
  // zero the slots of B rows past the last digit row (never loaded) in every stage
  for (int i = tid; i < A3_STAGES * A3_BN * 4; i += A3_THREADS) {
    const int st = i / (A3_BN * 4), rb = (i >> 2) % A3_BN, p = i & 3;
    if (n0 + rb >= nrows_b) *(v4i*)(lds + st * A3_STAGE_BYTES + (A3_BM + rb) * 64 + p * 16) = v4i{0, 0, 0, 0};
  }
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // published by the first step's barrier
  // this lane's source for each of the wave's load instructions (instruction q = wave * 8 + i
  // covers stage rows 16 q .. 16 q + 15)
  const int8_t* src[A3_LOADS];
  bool on[A3_LOADS];
  const size_t astep = (size_t)AJ_MROWS * 64, bstep = Bt ? nrows_b * 64 : 64;
#pragma unroll
  for (int i = 0; i < A3_LOADS; i++) {
    const int row = (wave * A3_LOADS + i) * 16 + (lane >> 2), p = lane & 3;
    if (row < A3_BM) {
      const int c = p ^ a3_swz(row);
      src[i] = limbs + (m0 + row) * 64 + c * 16;
      on[i] = true;
    } else {
      const int rb = row - A3_BM, c = p ^ a3_swz(rb);
      const size_t n = n0 + rb;
      on[i] = n < nrows_b;
      const size_t nn = on[i] ? n : 0;
      src[i] = Bt ? Bt + nn * 64 + c * 16 : Zs[nn / D] + (nn % D) * m + c * 16;
    }
  }
  auto load_stage = [&](size_t st, int buf) {
    int8_t* base = lds + buf * A3_STAGE_BYTES;
#pragma unroll
    for (int i = 0; i < A3_LOADS; i++)
      if (on[i])
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src[i] + st * ((wave * A3_LOADS + i) * 16 < A3_BM ? astep : bstep)),
                                         (__attribute__((address_space(3))) void*)(base + (wave * A3_LOADS + i) * 1024),
                                         16, 0, 0);
  };
  v16i acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int v = 0; v < 16; v++) acc[mt][nt][v] = 0;
  const int r32 = lane & 31, khalf = lane >> 5;
  auto compute = [&](int buf) {
    const int8_t* A = lds + buf * A3_STAGE_BYTES;
    const int8_t* B = A + A3_BM * 64;
    v4i a[2][4], b[2][4];
#pragma unroll
    for (int kh = 0; kh < 2; kh++) {
      const int c = kh * 2 + khalf;
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const int row = wm * 128 + mt * 32 + r32;
        a[kh][mt] = *(const v4i*)(A + row * 64 + ((c ^ a3_swz(row)) << 4));
      }
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const int row = wn * 128 + nt * 32 + r32;
        b[kh][nt] = *(const v4i*)(B + row * 64 + ((c ^ a3_swz(row)) << 4));
      }
    }
#pragma unroll
    for (int kh = 0; kh < 2; kh++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
          acc[mt][nt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[kh][mt], b[kh][nt], acc[mt][nt], 0, 0, 0);
    // the first half's 8 reads, then its 16 MFMAs with the second half's 8 reads interleaved (a
    // read, two MFMAs), then the second half's 16 MFMAs: the second half's LDS latency hides
    // behind the first half's matrix work
#pragma unroll
    for (int i = 0; i < 8; i++) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);
  };
  // load instructions this wave issues per step (one whose 16 rows are all padding is skipped)
  int nload = 0;
#pragma unroll
  for (int i = 0; i < A3_LOADS; i++) nload += __ballot(on[i]) != 0;
  nload = __builtin_amdgcn_readfirstlane(nload);
  for (int k = 0; k < A3_STAGES - 1; k++)
    if (s_begin + k < s_end) load_stage(s_begin + k, k);
  int buf = 0;
  for (size_t st = s_begin; st < s_end; st++) {
    // step st's loads are done once at most the later steps' loads are outstanding
    const size_t later = std::min<size_t>(A3_STAGES - 2, s_end - 1 - st);
    switch ((int)later * nload) {
#define A3_W(n) \
  case n: asm volatile("s_waitcnt vmcnt(" #n ")" ::: "memory"); break;
      A3_W(16) A3_W(15) A3_W(14) A3_W(13) A3_W(12) A3_W(11) A3_W(10) A3_W(9) A3_W(8) A3_W(7) A3_W(6) A3_W(5)
      A3_W(4) A3_W(3) A3_W(2) A3_W(1)
#undef A3_W
      default: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    }
    // a bare barrier: __syncthreads() would also wait for the later steps' loads
    asm volatile("s_barrier" ::: "memory");
    if (st + A3_STAGES - 1 < s_end) load_stage(st + A3_STAGES - 1, (buf + A3_STAGES - 1) % A3_STAGES);
    compute(buf);
    buf = (buf + 1) % A3_STAGES;
  }
  // D[i][j]: j = lane % 32, i = 8 (v / 4) + 4 (lane / 32) + v % 4
  int32_t* base = part + ((size_t)split * AJ_MROWS + m0 + wm * 128) * npad + n0 + wn * 128;
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const size_t row = mt * 32 + (v >> 2) * 8 + khalf * 4 + (v & 3);
        base[row * npad + nt * 32 + r32] = acc[mt][nt][v];
      }
}

// Digits to K-step-major order for large m: Bt[(step * R + r) * 64 + b] = Z_w[t][step * 64 + b]
// with r = w * 54 + t and R = nw * 54 rows.  Block (row tile of 64, run of KM_STEPS steps): each
// lane moves 16 bytes per step; per step the block writes 64 consecutive rows (4 KB contiguous)
// and reads 64 bytes of each row, whose 128-byte line the next step uses.
constexpr int KM_STEPS = 32;
__global__ __launch_bounds__(256) void k_digits_kmajor(const int8_t* const* __restrict__ Zs, size_t nw, size_t m,
                                                       int8_t* __restrict__ Bt) {
  const size_t R = nw * D, nsteps = m / 64;
  const size_t r = (size_t)blockIdx.x * 64 + (threadIdx.x >> 2);
  const int q = (threadIdx.x & 3) * 16;
  if (r >= R) return;
  const int8_t* src = Zs[r / D] + (r % D) * m + q;
  const size_t s0 = (size_t)blockIdx.y * KM_STEPS, s1 = min(s0 + KM_STEPS, nsteps);
  for (size_t st = s0; st < s1; st++)
    *(v4i*)(Bt + (st * R + r) * 64 + q) = *(const v4i*)(src + st * 64);
}

__device__ __forceinline__ int sum4_i8a(int x) {
  return (x << 24 >> 24) + (x << 16 >> 24) + (x << 8 >> 24) + (x >> 24);
}

// S[w*54 + t] = sum_j Z_w[t][j]; 16-byte loads (m is a multiple of 64)
__global__ __launch_bounds__(256) void k_digit_rowsum(const int8_t* const* __restrict__ Zs, size_t m,
                                                      int32_t* __restrict__ S) {
  size_t w = blockIdx.x / D, t = blockIdx.x % D;
  const int4* row = (const int4*)(Zs[w] + t * m);
  int32_t acc = 0;
  for (size_t j = threadIdx.x; j < m / 16; j += blockDim.x) {
    int4 v = row[j];
    acc += sum4_i8a(v.x) + sum4_i8a(v.y) + sum4_i8a(v.z) + sum4_i8a(v.w);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  __shared__ int32_t red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) S[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// ---- 32 x 32 x 32 variant with packed digits ----
// Digits carry two bits of information, so the B operand travels as 2-bit codes (d & 3: 0, 1, 3)
// and is expanded to int8 fragments in registers: per 64-wide K step and digit row, 16 bytes =
// four 32-bit words, word g holding digits 16 g .. 16 g + 15 with digit 16 g + 4 i + j at bit
// 8 j + 2 i, so (word >> 2 i) & 0x03030303 holds the codes of digits 16 g + 4 i .. + 3 in its
// bytes and one byte permute maps codes to int8 (0 -> 0, 1 -> 1, 3 -> -1).  Words are stored in
// the order g = 0, 2, 1, 3: lane half h of an MFMA B fragment needs words h (K half 0) and 2 + h
// (K half 1), one 8-byte read.  Per K step a block stages 16 KB of limbs and 4 KB of packed
// digits (against 32 KB for the int8 layout), and a wave reads 8 KB of limb fragments and 2 KB
// of digit words from LDS (against 16 KB).
constexpr int A4_STAGES = 4, A4_ABYTES = A3_BM * 64, A4_BBYTES = A3_BN * 16;
constexpr int A4_STAGE_BYTES = A4_ABYTES + A4_BBYTES;  // 20 KB
constexpr int A4_ALOADS = A4_ABYTES / 1024 / 4;         // limb load instructions per wave per step: 4
constexpr int A4_VM_PER_STEP = A4_ALOADS + 1;           // + one 1 KB load of 64 digit records

__global__ __launch_bounds__(256) void k_digits_pack(const int8_t* const* __restrict__ Zs, size_t nw, size_t m,
                                                     uint32_t* __restrict__ Bp, int32_t* __restrict__ S) {
  const size_t R = nw * D, nsteps = m / 64;
  const size_t r = (size_t)blockIdx.x * 64 + (threadIdx.x >> 2);
  const int q = threadIdx.x & 3;  // 16-digit quarter of the step = word g
  if (r >= R) return;             // a row's four lanes leave together
  const int8_t* src = Zs[r / D] + (r % D) * m + q * 16;
  const int pos = (q & 1) * 2 + (q >> 1);
  const size_t s0 = (size_t)blockIdx.y * KM_STEPS, s1 = min(s0 + KM_STEPS, nsteps);
  int sum = 0;
  for (size_t st = s0; st < s1; st++) {
    const v4i v = *(const v4i*)(src + st * 64);  // v[i] byte j = digit 4 i + j of the quarter
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      w |= ((uint32_t)v[i] & 0x03030303u) << (2 * i);
      sum += sum4_i8a(v[i]);
    }
    Bp[(st * R + r) * 4 + pos] = w;
  }
  sum += __shfl_xor(sum, 1, 64);
  sum += __shfl_xor(sum, 2, 64);
  if (q == 0) atomicAdd(&S[r], sum);
}

__global__ void k_zero_i32(int32_t* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// 16 digits (one packed word) to an int8 MFMA fragment
__device__ __forceinline__ v4i a4_unpack(uint32_t w) {
  constexpr uint32_t TAB = 0xFF000100u;  // selector 0 -> 0x00, 1 -> 0x01, 3 -> 0xFF
  v4i f;
  f[0] = (int)__builtin_amdgcn_perm(0u, TAB, w & 0x03030303u);
  f[1] = (int)__builtin_amdgcn_perm(0u, TAB, (w >> 2) & 0x03030303u);
  f[2] = (int)__builtin_amdgcn_perm(0u, TAB, (w >> 4) & 0x03030303u);
  f[3] = (int)__builtin_amdgcn_perm(0u, TAB, (w >> 6) & 0x03030303u);
  return f;
}

// Same grid, tiles and split layout as k_ajtai_gemm32 (256 x 256 blocks of four 128 x 128 waves).
// NST stages rotate (four are as fast as seven: the loads are not latency-bound).  Every step
// issues its 5 loads per wave unconditionally (steps past the end re-load the last step into a
// stage nobody reads again), so the vmcnt waits are exact.  The LDS reads are software-pipelined
// across steps: while the 16 MFMAs of step st's first K half run, the fragments of its second
// half are read; then each wave waits for its own loads of st + 1 (later steps' loads stay in
// flight), a bare barrier publishes every wave's slices, the loads of st + NST - 1 go into the
// stage read at st - 1, and the first-half fragments and digit records of st + 1 are read behind
// the 16 MFMAs of st's second half.  Digit rows past the last real row read whatever was loaded
// there and are masked to zero codes.  Measured at m = 2^18, 64 witnesses, the MFMAs alone run
// at 84% of peak, and each of the LDS reads, the unpacking and the loads adds its own time unless
// the schedule interleaves it with the MFMAs (one wave per SIMD): the interleaved schedule below
// took the kernel from 2.42 to 2.71 POP/s there.
template <int NST>
__global__ __launch_bounds__(256) void k_ajtai_gemm32p(const int8_t* __restrict__ limbs, size_t nsteps,
                                                      const uint32_t* __restrict__ Bp, size_t nrows_b, size_t npad,
                                                      int groups, int S, int32_t* __restrict__ part) {
  __shared__ __align__(16) int8_t lds[NST * A4_STAGE_BYTES];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const unsigned lid = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned per_split = (unsigned)A3_MBLK * (unsigned)groups;
  const unsigned split = lid / per_split, rem = lid % per_split;
  const size_t m0 = (size_t)(rem / groups) * A3_BM;
  const size_t n0 = (size_t)(rem % groups) * A3_BN;
  const size_t s_begin = (size_t)split * nsteps / S, s_end = ((size_t)split + 1) * nsteps / S;
  const size_t astep = (size_t)AJ_MROWS * 64, bstep = nrows_b * 16;  // bytes per K step
  const int8_t* asrc[A4_ALOADS];
#pragma unroll
  for (int i = 0; i < A4_ALOADS; i++) {
    const int row = (wave * A4_ALOADS + i) * 16 + (lane >> 2), c = (lane & 3) ^ a3_swz(row);
    asrc[i] = limbs + (m0 + row) * 64 + c * 16;
  }
  // digit records n0 + 64 wave + lane (row 0 for rows past the end: loaded, masked when read)
  const size_t nb = n0 + (size_t)wave * 64 + lane;
  const int8_t* bsrc = (const int8_t*)Bp + (nb < nrows_b ? nb : 0) * 16;
  uint32_t bmask[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) bmask[nt] = n0 + wn * 128 + nt * 32 + (lane & 31) < nrows_b ? 0xFFFFFFFFu : 0u;
  auto load_step = [&](size_t st, int buf) {
    int8_t* base = lds + buf * A4_STAGE_BYTES;
#pragma unroll
    for (int i = 0; i < A4_ALOADS; i++)
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(asrc[i] + st * astep),
                                       (__attribute__((address_space(3))) void*)(base + (wave * A4_ALOADS + i) * 1024),
                                       16, 0, 0);
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(bsrc + st * bstep),
                                     (__attribute__((address_space(3))) void*)(base + A4_ABYTES + wave * 1024), 16, 0, 0);
  };
  v16i acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int v = 0; v < 16; v++) acc[mt][nt][v] = 0;
  const int r32 = lane & 31, khalf = lane >> 5;
  // limb fragments of K half kh from stage buf
  auto read_a = [&](int buf, int kh, v4i (&a)[4]) {
    const int8_t* A = lds + buf * A4_STAGE_BYTES;
    const int c = kh * 2 + khalf;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const int row = wm * 128 + mt * 32 + r32;
      a[mt] = *(const v4i*)(A + row * 64 + ((c ^ a3_swz(row)) << 4));
    }
  };
  // this lane's digit records from stage buf (16 bytes per 32-row tile; words picked at use)
  auto read_b = [&](int buf, v4i (&rec)[4]) {
    const int8_t* B = lds + buf * A4_STAGE_BYTES + A4_ABYTES;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) rec[nt] = *(const v4i*)(B + (wn * 128 + nt * 32 + r32) * 16);
  };
  // digit fragments of K half kh from the records, and the 16 MFMAs of a K half (N tile outer,
  // so the first four need only the first fragment)
  auto unpack = [&](const v4i (&rec)[4], int kh, v4i (&b)[4]) {
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const uint32_t w = (uint32_t)(khalf ? rec[nt][2 + kh] : rec[nt][kh]) & bmask[nt];
      b[nt] = a4_unpack(w);
    }
  };
  auto mfmas = [&](const v4i (&a)[4], const v4i (&b)[4]) {
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
        acc[mt][nt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
  };
  static_assert(A4_VM_PER_STEP == 5 && NST >= 4 && 5 * (NST - 2) <= 63, "vmcnt immediates");
  const size_t last = s_end - 1;
#pragma unroll
  for (int k = 0; k < NST - 1; k++) load_step(min(s_begin + k, last), k);
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(5 * (NST - 2)) : "memory");
  asm volatile("s_barrier" ::: "memory");
  v4i a0[4], a1[4], rc[4], rn[4];
  read_a(0, 0, a0);
  read_b(0, rc);
  int buf = 0;
  for (size_t st = s_begin; st < s_end; st++) {
    // first K half of st: the first digit fragment, then each MFMA with a slice of the unpacking
    // of both halves, the second half's 4 limb fragments read beside the first MFMAs (a wave of
    // a SIMD to itself does not overlap its VALU work with its MFMAs on its own: the schedule
    // has to interleave them)
    v4i b0[4], b1[4];
    read_a(buf, 1, a1);
    unpack(rc, 0, b0);
    mfmas(a0, b0);
    unpack(rc, 1, b1);
    __builtin_amdgcn_sched_group_barrier(0x002, 13, 0);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (i < 4) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    const int nbuf = buf + 1 == NST ? 0 : buf + 1;
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(5 * (NST - 3)) : "memory");
    asm volatile("s_barrier" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    // second K half of st: the loads of st + NST - 1 and the reads of st + 1's first half and
    // digit records beside its MFMAs
    load_step(min(st + NST - 1, last), buf == 0 ? NST - 1 : buf - 1);
    read_a(nbuf, 0, a0);
    read_b(nbuf, rn);
    mfmas(a1, b1);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      if (i < 8) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      else if (i < 13) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) rc[nt] = rn[nt];
    buf = nbuf;
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the re-loads of the last steps
  int32_t* base = part + ((size_t)split * AJ_MROWS + m0 + wm * 128) * npad + n0 + wn * 128;
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const size_t row = mt * 32 + (v >> 2) * 8 + khalf * 4 + (v & 3);
        base[row * npad + nt * 32 + r32] = acc[mt][nt][v];
      }
}

// block = (w, i), 8 waves: wave l sums limb l's anti-diagonals G_l[s] (int64) for s < 107;
// then 107 threads combine the limbs in F_q and 54 threads reduce mod Phi_81.
__global__ __launch_bounds__(512) void k_ajtai_epilogue(const int32_t* __restrict__ part, size_t splits,
                                                        size_t npad, const int32_t* __restrict__ S,
                                                        uint64_t* __restrict__ out) {
  const size_t w = blockIdx.x / KAPPA, i = blockIdx.x % KAPPA;
  __shared__ int64_t G[8][2 * D - 1];
  __shared__ int64_t corr[2 * D - 1];
  __shared__ uint64_t P[2 * D - 1];
  const int lane = threadIdx.x & 63, l = threadIdx.x >> 6;
  for (int s = lane; s < 2 * D - 1; s += 64) {
    int u_lo = s - (D - 1) > 0 ? s - (D - 1) : 0;
    int u_hi = s < D - 1 ? s : D - 1;
    int64_t g = 0;
    for (size_t sp = 0; sp < splits; sp++) {
      const int32_t* base = part + (sp * AJ_MROWS + (i * 8 + l) * D) * npad + w * D + s;
      for (int u = u_lo; u <= u_hi; u++) g += base[(size_t)u * npad - u];
    }
    G[l][s] = g;
    if (l == 0) {
      int64_t c = 0;
      for (int u = u_lo; u <= u_hi; u++) c += S[w * D + (s - u)];
      corr[s] = c;
    }
  }
  __syncthreads();
  const int s = threadIdx.x;
  if (s < 2 * D - 1) {
    uint64_t acc = 0;
    for (int ll = 0; ll < 8; ll++) acc = fadd(acc, fmul(from_i64(G[ll][s] + 128 * corr[s]), 1ull << (8 * ll)));
    P[s] = acc;
  }
  __syncthreads();
  if (s < D) {
    uint64_t c = P[s];
    if (s <= 25) c = fadd(c, P[s + 81]);
    if (s >= 27) c = fsub(c, P[s + 27]);
    if (s <= 26) c = fsub(c, P[s + 54]);
    out[(w * KAPPA + i) * D + s] = c;
  }
}

// Tile shape and K-split autotuner: the first commit of each (m, nw) in the process times both
// tile shapes (16 x 16 x 64 MFMAs in 192 x 352 blocks, 32 x 32 x 32 MFMAs in 256 x 256 blocks) at
// every split count from 1 to aj_max_splits(m) on its own inputs -- the GEMM and the reduction of
// its split planes, which grows with the count (twice each, HIP events, the faster kept) -- and
// caches the fastest; later commits use the cached choice.  Every choice gives the same
// commitments (exact int32 partial sums, at most 2^20 * 128 in magnitude), so it only moves
// time.  A single lock serializes the first measurement of each shape.
struct AjChoice {
  int variant;  // 1: 16 x 16 x 64 tiles, 2: 32 x 32 x 32 tiles, 3: 32 x 32 x 32 tiles on packed digits
  size_t splits;
};
static std::mutex g_aj_mu;
static std::map<std::pair<size_t, size_t>, AjChoice> g_aj_chosen;
static int g_aj_force = 0;

void ajtai_force_variant(int v) {
  std::lock_guard<std::mutex> g(g_aj_mu);
  if (v < 0 || v > 3) throw std::invalid_argument("Ajtai tile variant must be 0 (timed), 1, 2 or 3");
  g_aj_force = v;
  g_aj_chosen.clear();
}

// the digit operand of each tile shape: int8 rows (Zs, or Bt when K-step-major) or packed words
struct AjDigits {
  const int8_t* const* Zs;
  const int8_t* Bt;
  const uint32_t* Bp;
};

static void aj_launch(int variant, const int8_t* limbs, size_t nsteps, const AjDigits& dg, size_t nw, size_t m,
                      size_t sp, int32_t* part, hipStream_t s, bool timed) {
  if (variant == 3) {
    const size_t groups = a3_groups(nw), npad = groups * A3_BN;
    const dim3 grid((unsigned)(A3_MBLK * groups * sp));
    if (timed)
      NS_LAUNCH_TIMED(KT_AJTAI_GEMM, k_ajtai_gemm32p<A4_STAGES>, grid, dim3(256), 0, s, limbs, nsteps, dg.Bp, nw * D, npad,
                      (int)groups, (int)sp, part);
    else
      hipLaunchKernelGGL(k_ajtai_gemm32p<A4_STAGES>, grid, dim3(256), 0, s, limbs, nsteps, dg.Bp, nw * D, npad, (int)groups,
                         (int)sp, part);
  } else if (variant == 2) {
    const size_t groups = a3_groups(nw), npad = groups * A3_BN;
    const dim3 grid((unsigned)(A3_MBLK * groups * sp));
    if (timed)
      NS_LAUNCH_TIMED(KT_AJTAI_GEMM, k_ajtai_gemm32, grid, dim3(A3_THREADS), 0, s, limbs, nsteps, dg.Zs, dg.Bt, nw, m,
                      npad, (int)groups, (int)sp, part);
    else
      hipLaunchKernelGGL(k_ajtai_gemm32, grid, dim3(A3_THREADS), 0, s, limbs, nsteps, dg.Zs, dg.Bt, nw, m, npad,
                         (int)groups, (int)sp, part);
  } else {
    const size_t groups = aj_groups(nw), npad = groups * AJ_NBLK;
    const dim3 grid((unsigned)(AJ_MBLK * groups * sp));
    if (timed)
      NS_LAUNCH_TIMED(KT_AJTAI_GEMM, k_ajtai_gemm, grid, dim3(AJ_THREADS), 0, s, limbs, nsteps, dg.Zs, dg.Bt, nw, m,
                      npad, (int)groups, (int)sp, part);
    else
      hipLaunchKernelGGL(k_ajtai_gemm, grid, dim3(AJ_THREADS), 0, s, limbs, nsteps, dg.Zs, dg.Bt, nw, m, npad,
                         (int)groups, (int)sp, part);
  }
}
static size_t aj_npad(int variant, size_t nw) {
  return variant >= 2 ? a3_groups(nw) * A3_BN : aj_groups(nw) * AJ_NBLK;
}

// Digit preparation of a tile shape: S (row sums) and the layout its GEMM reads.  The packed
// layout's kernel accumulates S itself (zeroed first); the int8 shapes take k_digit_rowsum and,
// for m >= 2^18, the K-step-major copy.
static void aj_prepare(int variant, const int8_t* const* Zs_dev, size_t nw, size_t m, const AjDigits& dg, int32_t* S,
                       hipStream_t s) {
  const size_t nsteps = m / AJ_BK;
  const dim3 tiles((unsigned)((nw * D + 63) / 64), (unsigned)((nsteps + KM_STEPS - 1) / KM_STEPS));
  if (variant >= 3) {
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((nw * D + 255) / 256)), dim3(256), 0, s, S, nw * D);
    hipLaunchKernelGGL(k_digits_pack, tiles, dim3(256), 0, s, Zs_dev, nw, m, const_cast<uint32_t*>(dg.Bp), S);
  } else {
    hipLaunchKernelGGL(k_digit_rowsum, dim3((unsigned)(nw * D)), dim3(256), 0, s, Zs_dev, m, S);
    if (dg.Bt) hipLaunchKernelGGL(k_digits_kmajor, tiles, dim3(256), 0, s, Zs_dev, nw, m, const_cast<int8_t*>(dg.Bt));
  }
}

// Tile shape and K-split autotuner: the first commit of each (m, nw) in the process prepares the
// digits in every layout, then times the three tile shapes (16 x 16 x 64 MFMAs in 192 x 352 blocks,
// 32 x 32 x 32 MFMAs in 256 x 256 blocks from int8 or from packed digits) at every split count
// from 1 to aj_max_splits(m) on its own inputs -- the GEMM and the reduction of its split planes,
// which grows with the count (twice each, HIP events, the faster kept) -- and caches the fastest;
// later commits prepare and use only the cached choice.  Every choice gives the same commitments
// (exact int32 partial sums, at most 2^20 * 128 in magnitude), so it only moves time.  A single
// lock serializes the first measurement of each shape.
static AjChoice aj_tuned(const int8_t* limbs, size_t m, const int8_t* const* Zs_dev, const AjDigits& dg, size_t nw,
                         int32_t* part, int32_t* S, int32_t* red, hipStream_t s, bool* prepared) {
  std::lock_guard<std::mutex> g(g_aj_mu);
  const auto key = std::make_pair(m, nw);
  auto it = g_aj_chosen.find(key);
  *prepared = false;
  if (it != g_aj_chosen.end()) return it->second;
  *prepared = true;
  const size_t nsteps = m / AJ_BK;
  hipEvent_t e0, e1;
  NS_CHECK(hipEventCreate(&e0));
  NS_CHECK(hipEventCreate(&e1));
  for (int variant = 3; variant >= 1; variant -= 2) aj_prepare(variant, Zs_dev, nw, m, dg, S, s);
  AjChoice best{g_aj_force ? g_aj_force : 1, aj_splits(m, nw)};
  float best_ms = 1e30f;
  for (int variant = 1; variant <= 3; variant++) {
    if (g_aj_force && variant != g_aj_force) continue;
    for (size_t sp = 1; sp <= aj_max_splits(m); sp++) {
      float ms = 1e30f;
      for (int rep = 0; rep < 2; rep++) {
        NS_CHECK(hipEventRecord(e0, s));
        aj_launch(variant, limbs, nsteps, dg, nw, m, sp, part, s, false);
        if (sp > 1) launch_sum_splits(part, sp, (size_t)AJ_MROWS * aj_npad(variant, nw), red, s);
        NS_CHECK(hipEventRecord(e1, s));
        NS_CHECK(hipEventSynchronize(e1));
        float t = 0;
        NS_CHECK(hipEventElapsedTime(&t, e0, e1));
        ms = std::min(ms, t);
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = AjChoice{variant, sp};
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g_aj_chosen[key] = best;
  return best;
}

void ajtai_commit(const int8_t* limbs, size_t kappa, size_t m, const int8_t* const* Zs_dev, size_t nw,
                  uint64_t* out, void* workspace, hipStream_t s) {
  if (nw == 0) return;
  const size_t nsteps = m / AJ_BK;
  const AjWs L = aj_ws_layout(m, nw);
  char* ws = (char*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
  if (ws != (char*)workspace) throw std::invalid_argument("Ajtai workspace must be 256-byte aligned");
  int32_t* part = (int32_t*)(ws + L.part);
  int32_t* S = (int32_t*)(ws + L.S);
  int32_t* red = (int32_t*)(ws + L.red);
  AjDigits dg{Zs_dev, aj_kmajor(m) ? (const int8_t*)(ws + L.bt) : nullptr, (const uint32_t*)(ws + L.bp)};
  bool prepared = false;  // a first commit of this shape prepares every layout while tuning
  const AjChoice ch = aj_tuned(limbs, m, Zs_dev, dg, nw, part, S, red, s, &prepared);
  if (!prepared) aj_prepare(ch.variant, Zs_dev, nw, m, dg, S, s);
  const size_t splits = ch.splits, npad = aj_npad(ch.variant, nw);
  aj_launch(ch.variant, limbs, nsteps, dg, nw, m, splits, part, s, true);
  const int32_t* G = part;
  if (splits > 1) {
    // reduce the K splits with coalesced loads before the anti-diagonal epilogue
    launch_sum_splits(part, splits, (size_t)AJ_MROWS * npad, red, s);
    G = red;
  }
  hipLaunchKernelGGL(k_ajtai_epilogue, dim3((unsigned)(nw * KAPPA)), dim3(512), 0, s, G, (size_t)1, npad, S, out);
  (void)kappa;
}

}  // namespace ns
