// ccs.hip -- Pi_CCS row-phase kernels: chi tables, z1 recomposition, sparse products, the
// Eval column sums, the per-round univariate polynomial and the in-place table folds, plus the
// digit x K products y[w][j][rho] = sum_c Z_w[rho][c] v_j[c] (precompute_for_r and DEC children).
//
// Round polynomial (optimized_engine/oracle.rs:1047-1178, b = 2): for every row pair
// (2t, 2t+1) and X the current row bit,
//   eq_beta(X) * ( f(M z1)(X) + sum_{w,rho} wg[w,rho] * N(y_{w,rho}(X)) ) + gamma^k eq_rin(X) Eval(X),
//   N(y) = y^3 - y, y(X) = lo + (hi - lo) X,
// summed over all pairs; the kernel returns the deg_max + 1 coefficients (low -> high).
// Round 0 reads the int8 digit matrices directly: with lo, hi in {-1,0,1},
//   N(y(X)) = lo A(X) + hi B(X) + 3 lo^2 hi C(X) + 3 lo hi^2 D(X),
//   A = -2X + 3X^2 - X^3, B = X^3 - X, C = X - 2X^2 + X^3, D = X^2 - X^3,
// so the NC part of a pair needs only signed selections of wg (no products).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <utility>
#include <stdexcept>

#include "kernels.h"

namespace ns {

typedef int v4i __attribute__((ext_vector_type(4)));

// ---------------- chi ----------------
__global__ void k_chi(const K2* __restrict__ bits, int nbits, K2* __restrict__ out) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ((size_t)1 << nbits)) return;
  K2 w = kone();
  for (int b = 0; b < nbits; b++) {
    K2 r = bits[b];
    w = kmul(w, ((idx >> b) & 1) ? r : ksub(kone(), r));
  }
  out[idx] = w;
}

void launch_chi(const K2* bits, int nbits, K2* out, hipStream_t s) {
  size_t len = (size_t)1 << nbits;
  hipLaunchKernelGGL(k_chi, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, bits, nbits, out);
}

// chi tables of the suffixes bits[j+1 .. ell) for j = j0 .. ell - 1, concatenated (table j holds
// 2^(ell - j - 1) entries); blockIdx.y = 0 takes bits[0 ..], 1 takes bits[64 ..] into out + total
__global__ void k_chi_suffix(const K2* __restrict__ bits, int ell, int j0, size_t total, K2* __restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const K2* b = bits + 64 * blockIdx.y;
  size_t off = 0;
  int j = j0;
  while (j < ell - 1 && idx >= off + ((size_t)1 << (ell - j - 1))) {
    off += (size_t)1 << (ell - j - 1);
    j++;
  }
  const size_t e = idx - off;
  K2 w = kone();
  for (int i = 0; i < ell - j - 1; i++) {
    const K2 r = b[j + 1 + i];
    w = kmul(w, ((e >> i) & 1) ? r : ksub(kone(), r));
  }
  out[blockIdx.y * total + idx] = w;
}

void launch_chi_suffix(const K2* bits, int ell, int j0, size_t total, int ntab, K2* out, hipStream_t s) {
  if (total == 0) return;
  hipLaunchKernelGGL(k_chi_suffix, dim3((unsigned)((total + 255) / 256), (unsigned)ntab), dim3(256), 0, s, bits, ell, j0,
                     total, out);
}

// ---------------- z1 ----------------
__global__ void k_z1(const int8_t* __restrict__ Z, size_t m, K2* __restrict__ z1) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  int64_t acc = 0;
  for (int rho = D - 1; rho >= 0; rho--) acc = 2 * acc + Z[(size_t)rho * m + c];
  z1[c] = K2{from_i64(acc), 0};
}

void launch_z1(const int8_t* Z1, size_t m, K2* z1, hipStream_t s) {
  hipLaunchKernelGGL(k_z1, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, Z1, m, z1);
}

// ---------------- sparse products ----------------
__global__ void k_spmv(const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ col,
                       const uint64_t* __restrict__ val, uint32_t n, const K2* __restrict__ x, K2* __restrict__ y,
                       K2 coeff, int accumulate) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  K2 acc = kzero();
  for (uint32_t k = ptr[r]; k < ptr[r + 1]; k++) acc = kadd(acc, kscale(x[col[k]], val[k]));
  acc = kmul(acc, coeff);
  y[r] = accumulate ? kadd(y[r], acc) : acc;
}

__global__ void k_copy_scaled(const K2* __restrict__ x, size_t n, K2* __restrict__ y, K2 coeff, int accumulate) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  K2 v = kmul(x[r], coeff);
  y[r] = accumulate ? kadd(y[r], v) : v;
}

void launch_spmv(const DevSparse& M, const K2* x, K2* y, K2 coeff, bool accumulate, hipStream_t s) {
  dim3 g((M.n + 255) / 256);
  if (M.identity) {
    size_t cap = M.n < M.m ? M.n : M.m;
    hipLaunchKernelGGL(k_copy_scaled, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, x, cap, y, coeff,
                       accumulate ? 1 : 0);
    return;
  }
  hipLaunchKernelGGL(k_spmv, g, dim3(256), 0, s, M.csr_ptr, M.csr_col, M.csr_val, M.n, x, y, coeff,
                     accumulate ? 1 : 0);
}

// v = M^T chi over the CSC arrays: one thread per column; columns with more than SPMV_HEAVY
// entries are left to k_spmv_t_heavy (a thread walking 500k entries would take ~0.2 s)
__global__ void k_spmv_t(const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ row,
                         const uint64_t* __restrict__ val, uint32_t m, const K2* __restrict__ chi, uint32_t n_eff,
                         K2* __restrict__ v) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  const uint32_t lo = ptr[c], hi = ptr[c + 1];
  if (hi - lo > SPMV_HEAVY) return;
  K2 acc = kzero();
  for (uint32_t k = lo; k < hi; k++) {
    uint32_t r = row[k];
    if (r < n_eff) acc = kadd(acc, kscale(chi[r], val[k]));
  }
  v[c] = acc;
}

// one 1024-thread workgroup per heavy column: strided partial sums, then a tree in LDS
__global__ __launch_bounds__(1024) void k_spmv_t_heavy(const uint32_t* __restrict__ ptr,
                                                       const uint32_t* __restrict__ row,
                                                       const uint64_t* __restrict__ val,
                                                       const uint32_t* __restrict__ cols, const K2* __restrict__ chi,
                                                       uint32_t n_eff, K2* __restrict__ v) {
  __shared__ K2 red[1024];
  const uint32_t c = cols[blockIdx.x];
  const uint32_t lo = ptr[c], hi = ptr[c + 1];
  K2 acc = kzero();
  for (uint32_t k = lo + threadIdx.x; k < hi; k += blockDim.x) {
    const uint32_t r = row[k];
    if (r < n_eff) acc = kadd(acc, kscale(chi[r], val[k]));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = kadd(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) v[c] = red[0];
}

__global__ void k_copy_prefix(const K2* __restrict__ x, size_t cap, size_t m, K2* __restrict__ v) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  v[c] = c < cap ? x[c] : kzero();
}

void launch_spmv_t(const DevSparse& M, const K2* chi, size_t n_eff, K2* v, hipStream_t s) {
  dim3 g((M.m + 255) / 256);
  if (M.identity) {
    size_t cap = M.m < n_eff ? M.m : n_eff;
    hipLaunchKernelGGL(k_copy_prefix, g, dim3(256), 0, s, chi, cap, (size_t)M.m, v);
    return;
  }
  hipLaunchKernelGGL(k_spmv_t, g, dim3(256), 0, s, M.csc_ptr, M.csc_row, M.csc_val, M.m, chi, (uint32_t)n_eff, v);
  if (M.n_heavy)
    hipLaunchKernelGGL(k_spmv_t_heavy, dim3(M.n_heavy), dim3(1024), 0, s, M.csc_ptr, M.csc_row, M.csc_val,
                       M.heavy_cols, chi, (uint32_t)n_eff, v);
}

// ---------------- Eval column sums ----------------
// ---------------- contractions over digit rows on int8 MFMA ----------------
// out[r][c] = sum_i sum_{kk<54} A_i[r][kk] Z_i[kk][c] (int32) for r < arows <= MT*16.  A_i is int8
// [MT*16][64] at A + i*MT*16*64, zero beyond arows rows and 54 columns.  A block (4 waves) owns
// 64 columns; wave w takes the instances i = w, w+4, ...  Lane L loads digit rows
// 16(L/16)..+15 of the columns c0 + 4(L%16) .. +3 as dwords (16 lanes read 64 contiguous bytes
// of a row) and transposes them in registers into four v_mfma_i32_16x16x64_i8 B fragments:
// byte j of fragment n is row 16(L/16)+j of column c0 + 4(L%16) + n, so column j of MFMA tile
// n is global column c0 + 4j + n and each lane ends up owning four adjacent output columns
// (one 16-byte store per row).  The waves' partial sums meet in LDS.  Requires m % 64 == 0.
typedef int v4i_t __attribute__((ext_vector_type(4)));
constexpr int RC_NB = 64;
constexpr int RC_AP = 64;  // padded row length of A_i
__device__ __forceinline__ uint32_t byte_col(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3, int n) {
  const int s = 8 * n;
  return ((v0 >> s) & 0xFFu) | (((v1 >> s) & 0xFFu) << 8) | (((v2 >> s) & 0xFFu) << 16) | (((v3 >> s) & 0xFFu) << 24);
}
template <int MT>
__global__ __launch_bounds__(256) void k_rowcontract(const int8_t* __restrict__ A, int arows,
                                                     const int8_t* const* __restrict__ Zs, int k, size_t m,
                                                     int32_t* __restrict__ out) {
  __shared__ v4i_t red[3][MT * 4][64];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const size_t c0 = (size_t)blockIdx.x * RC_NB;
  const int kq = lane >> 4, cg = lane & 15;
  v4i_t acc[MT][4];
#pragma unroll
  for (int mt = 0; mt < MT; mt++)
#pragma unroll
    for (int n = 0; n < 4; n++) acc[mt][n] = v4i_t{0, 0, 0, 0};
  for (int i = wave; i < k; i += 4) {
    typedef const __attribute__((address_space(1))) uint32_t gu32;
    const int8_t* Z = Zs[i] + c0 + 4 * cg;
    uint32_t v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int rho = 16 * kq + j;
      const uint32_t x = *(gu32*)(Z + (size_t)min(rho, D - 1) * m);  // rows past 53 read row 53
      v[j] = rho < D ? x : 0u;
    }
    v4i_t b[4];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
      for (int q = 0; q < 4; q++) b[n][q] = (int)byte_col(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3], n);
    const int8_t* Ai = A + (size_t)i * MT * 16 * RC_AP + cg * RC_AP + 16 * kq;
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
      const v4i_t a = *(const v4i_t*)(Ai + mt * 16 * RC_AP);
#pragma unroll
      for (int n = 0; n < 4; n++) acc[mt][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[n], acc[mt][n], 0, 0, 0);
    }
  }
  if (wave > 0)
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int n = 0; n < 4; n++) red[wave - 1][mt * 4 + n][lane] = acc[mt][n];
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int w = 0; w < 3; w++)
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int n = 0; n < 4; n++) acc[mt][n] += red[w][mt * 4 + n][lane];
#pragma unroll
  for (int mt = 0; mt < MT; mt++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = mt * 16 + 4 * kq + v;
      if (row < arows)
        *(v4i_t*)(out + (size_t)row * m + c0 + 4 * cg) = v4i_t{acc[mt][0][v], acc[mt][1][v], acc[mt][2][v], acc[mt][3][v]};
    }
}

// A_i rows for the weighted column sum: bytes of W.c0 and W.c1 (each minus 128), then a row of
// ones; int8 [32][64] per instance, zero beyond row 17 and column 54
__global__ void k_w_limbs(const K2* __restrict__ W, int nw, int8_t* __restrict__ A) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nw * 32 * RC_AP) return;
  int i = idx / (32 * RC_AP), r = (idx / RC_AP) % 32, rho = idx % RC_AP;
  int8_t v = 0;
  if (r < 17 && rho < D) {
    K2 w = W[i * D + rho];
    if (r < 8) v = (int8_t)((int)((w.c0 >> (8 * r)) & 0xFF) - 128);
    else if (r < 16) v = (int8_t)((int)((w.c1 >> (8 * (r - 8))) & 0xFF) - 128);
    else v = 1;
  }
  A[idx] = v;
}

// S[c] = (sum_l 256^l (G_l + 128 G_16), same for c1): undo the -128 byte offset with the digit sum
__global__ void k_colsum_epilogue(const int32_t* __restrict__ G, size_t m, K2* __restrict__ S) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  const int64_t corr = 128 * (int64_t)G[16 * m + c];
  uint64_t comp[2];
#pragma unroll
  for (int cp = 0; cp < 2; cp++) {
    uint64_t acc = 0;
#pragma unroll
    for (int l = 0; l < 8; l++) acc = fadd(acc, fmul(from_i64(G[(size_t)(cp * 8 + l) * m + c] + corr), 1ull << (8 * l)));
    comp[cp] = acc;
  }
  S[c] = K2{comp[0], comp[1]};
}

size_t colsum_ws_bytes(size_t nw, size_t m) { return nw * 32 * RC_AP + 256 + 17 * m * 4 + 256; }

__global__ __launch_bounds__(256) void k_weighted_colsum(const int8_t* const* __restrict__ Zs, int nw,
                                                         const K2* __restrict__ W, size_t m, K2* __restrict__ S) {
  __shared__ K2 w_s[16 * D];
  for (int i = threadIdx.x; i < nw * D; i += blockDim.x) w_s[i] = W[i];
  __syncthreads();
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  K2 pos = kzero(), neg = kzero();
  for (int w = 0; w < nw; w++) {
    const int8_t* z = Zs[w] + c;
    int8_t d[D];
#pragma unroll
    for (int rho = 0; rho < D; rho++) d[rho] = z[(size_t)rho * m];
#pragma unroll
    for (int rho = 0; rho < D; rho++) {
      K2 wv = w_s[w * D + rho];
      K2 zero = kzero();
      pos = kadd(pos, d[rho] > 0 ? wv : zero);
      neg = kadd(neg, d[rho] < 0 ? wv : zero);
    }
  }
  S[c] = ksub(pos, neg);
}

void launch_weighted_digit_colsum(const int8_t* const* Zs_dev, size_t nw, const K2* W, size_t m, K2* S, void* ws,
                                  hipStream_t s) {
  if (nw == 0) {
    NS_CHECK(hipMemsetAsync(S, 0, m * sizeof(K2), s));
    return;
  }
  if (m % RC_NB != 0 || !ws) {
    hipLaunchKernelGGL(k_weighted_colsum, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, Zs_dev, (int)nw, W, m,
                       S);
    return;
  }
  int8_t* A = (int8_t*)ws;
  int32_t* G = (int32_t*)(((uintptr_t)(A + nw * 32 * RC_AP) + 255) & ~(uintptr_t)255);
  int na = (int)(nw * 32 * RC_AP);
  hipLaunchKernelGGL(k_w_limbs, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, W, (int)nw, A);
  hipLaunchKernelGGL(k_rowcontract<2>, dim3((unsigned)(m / RC_NB)), dim3(256), 0, s, A, 17, Zs_dev, (int)nw, m, G);
  hipLaunchKernelGGL(k_colsum_epilogue, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, G, m, S);
}

// ---------------- y = Z v via int8 MFMA on byte limbs of v ----------------
// Vl[(j*2 + comp)*8 + l][c] = byte_l(v_j[c].comp) - 128 (64 rows per 4 matrices).
// Balanced base-256 limbs of v's K coordinates: component x in [0, q) is taken as the representative
// v of its class with |v| <= 128 (2^64 - 1) / 255 (x itself up to 0x7F..7F, else x - q) and written
// as 8 digits in [-128, 127] (digit-by-digit with carries), so the digit x limb products need no
// offset correction.  Limb row (j 2 + comp) 8 + l of [16 t][m].
__device__ __forceinline__ void balanced_limbs(uint64_t x, int8_t (&d)[8]) {
  if (x <= 0x7F7F7F7F7F7F7F7Full) {
    uint64_t u = x;  // non-negative: u = sum d_l 256^l
#pragma unroll
    for (int l = 0; l < 8; l++) {
      int r = (int)(u & 255);
      if (r >= 128) r -= 256;
      d[l] = (int8_t)r;
      u = (u - (uint64_t)(int64_t)r) >> 8;
    }
  } else {
    uint64_t w = Q - x;  // v = -w: the digits of -w, with w's magnitude carried
#pragma unroll
    for (int l = 0; l < 8; l++) {
      const int r = (int)(w & 255);
      if (r <= 128) {
        d[l] = (int8_t)(-r);
        w = (w - (uint64_t)r) >> 8;
      } else {
        d[l] = (int8_t)(256 - r);
        w = (w + (uint64_t)(256 - r)) >> 8;
      }
    }
  }
}
__global__ void k_v_limbs(const K2* __restrict__ v, size_t t, size_t m, int8_t* __restrict__ Vl) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t j = blockIdx.y;
  if (c >= m) return;
  K2 x = v[j * m + c];
  uint64_t comp[2] = {x.c0, x.c1};
#pragma unroll
  for (int cp = 0; cp < 2; cp++) {
    int8_t d[8];
    balanced_limbs(comp[cp], d);
#pragma unroll
    for (int l = 0; l < 8; l++) Vl[((j * 2 + cp) * 8 + l) * m + c] = d[l];
  }
}



constexpr int YD_SPLIT_MAX = 64;

// grid: x = 64-row tile of digit rows (w, rho), y = K split.  N = 16 * t limb rows (t <= 8).
template <int NS>
__global__ __launch_bounds__(256) void k_ydigits_gemm(const int8_t* const* __restrict__ Zs, size_t nrows,
                                                      const int8_t* __restrict__ Vl, size_t m, size_t kc,
                                                      int32_t* __restrict__ part, size_t mpad) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t r0 = (size_t)blockIdx.x * 64 + wave * 16;
  const size_t k_begin = (size_t)blockIdx.y * kc;
  const int kq = (lane >> 4) * 16, r16 = lane & 15;
  const size_t myrow = r0 + r16;
  const bool aval = myrow < nrows;
  const int8_t* arow = aval ? Zs[myrow / D] + (myrow % D) * m + k_begin + kq : nullptr;
  const int8_t* brow[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) brow[n] = Vl + (size_t)(n * 16 + r16) * m + k_begin + kq;
  v4i acc[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) acc[n] = v4i{0, 0, 0, 0};
  // four K steps' loads in flight at once, then their MFMAs (one step's loads alone leave the wave
  // waiting a memory latency per step)
  constexpr int UNR = 4;
  size_t k = 0;
  for (; k + UNR * 64 <= kc; k += UNR * 64) {
    v4i a[UNR], b[UNR][NS];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      a[u] = aval ? *(const v4i*)(arow + k + 64 * u) : v4i{0, 0, 0, 0};
#pragma unroll
      for (int n = 0; n < NS; n++) b[u][n] = *(const v4i*)(brow[n] + k + 64 * u);
    }
#pragma unroll
    for (int u = 0; u < UNR; u++)
#pragma unroll
      for (int n = 0; n < NS; n++) acc[n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[u], b[u][n], acc[n], 0, 0, 0);
  }
  for (; k < kc; k += 64) {
    v4i a = aval ? *(const v4i*)(arow + k) : v4i{0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < NS; n++) {
      v4i b = *(const v4i*)(brow[n] + k);
      acc[n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc[n], 0, 0, 0);
    }
  }
  int32_t* base = part + ((size_t)blockIdx.y * mpad + r0) * (NS * 16);
#pragma unroll
  for (int n = 0; n < NS; n++)
#pragma unroll
    for (int v = 0; v < 4; v++) base[((lane >> 4) * 4 + v) * (NS * 16) + n * 16 + r16] = acc[n][v];
}

// The same product with both operands staged through LDS: per 256-byte K chunk a block loads its
// 64 digit rows and the 16 NS limb rows of v with 16-byte loads, 16 threads per row (256
// contiguous bytes of one row per 16 lanes, where the direct kernel's lanes read 64 bytes from
// 16 rows apart), into a double-buffered tile (rows padded to 272 bytes: the 16-byte fragment
// reads of 16 lanes hit distinct banks); the next chunk's loads are in flight during this
// chunk's MFMAs.  kc must be a multiple of 256.
constexpr int YD_PITCH = 272;
template <int NS>
__global__ __launch_bounds__(256) void k_ydigits_gemm_lds(const int8_t* const* __restrict__ Zs, size_t nrows,
                                                         const int8_t* __restrict__ Vl, size_t m, size_t kc,
                                                         int32_t* __restrict__ part, size_t mpad) {
  __shared__ __align__(16) int8_t at[2][64 * YD_PITCH];
  __shared__ __align__(16) int8_t bt[2][NS * 16 * YD_PITCH];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t r0 = (size_t)blockIdx.x * 64;
  const size_t k_begin = (size_t)blockIdx.y * kc;
  const int kq = (lane >> 4) * 16, r16 = lane & 15;
  // this thread's load slots: A rows (4 of them) and B rows (NS of them), 16 bytes each
  const int seg = threadIdx.x & 15, rsub = threadIdx.x >> 4;  // rsub 0..15
  const int8_t* asrc[4];
  bool aok[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const size_t row = r0 + (size_t)(q * 16 + rsub);
    aok[q] = row < nrows;
    asrc[q] = aok[q] ? Zs[row / D] + (row % D) * m + k_begin + seg * 16 : nullptr;
  }
  v4i ra[4], rb[NS];
  auto gload = [&](size_t k) {
#pragma unroll
    for (int q = 0; q < 4; q++) ra[q] = aok[q] ? *(const v4i*)(asrc[q] + k) : v4i{0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < NS; n++) rb[n] = *(const v4i*)(Vl + (size_t)(n * 16 + rsub) * m + k_begin + k + seg * 16);
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) *(v4i*)(at[buf] + (q * 16 + rsub) * YD_PITCH + seg * 16) = ra[q];
#pragma unroll
    for (int n = 0; n < NS; n++) *(v4i*)(bt[buf] + (n * 16 + rsub) * YD_PITCH + seg * 16) = rb[n];
  };
  v4i acc[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) acc[n] = v4i{0, 0, 0, 0};
  const size_t nch = kc / 256;
  gload(0);
  lstore(0);
  __syncthreads();
  for (size_t c = 0; c < nch; c++) {
    const int buf = (int)(c & 1);
    if (c + 1 < nch) gload((c + 1) * 256);
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const v4i a = *(const v4i*)(at[buf] + (wave * 16 + r16) * YD_PITCH + st * 64 + kq);
#pragma unroll
      for (int n = 0; n < NS; n++) {
        const v4i b = *(const v4i*)(bt[buf] + (n * 16 + r16) * YD_PITCH + st * 64 + kq);
        acc[n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc[n], 0, 0, 0);
      }
    }
    if (c + 1 < nch) lstore(buf ^ 1);  // the other buffer's readers finished before the last barrier
    __syncthreads();
  }
  int32_t* base = part + ((size_t)blockIdx.y * mpad + r0 + wave * 16) * (NS * 16);
#pragma unroll
  for (int n = 0; n < NS; n++)
#pragma unroll
    for (int v = 0; v < 4; v++) base[((lane >> 4) * 4 + v) * (NS * 16) + n * 16 + r16] = acc[n][v];
}

// thread per (row, j): y[row][j] = sum_comp,l 256^l G (balanced limbs: no offset correction)
__global__ void k_ydigits_epilogue(const int32_t* __restrict__ part, size_t splits, size_t mpad, size_t nrows,
                                   size_t t, K2* __restrict__ y) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nrows * t) return;
  size_t row = idx / t, j = idx % t;
  const size_t ncols = 16 * t;
  uint64_t comp[2];
  for (int cp = 0; cp < 2; cp++) {
    uint64_t acc = 0;
    for (int l = 0; l < 8; l++) {
      int64_t g = 0;
      size_t col = (j * 2 + cp) * 8 + l;
      for (size_t sp = 0; sp < splits; sp++) g += part[(sp * mpad + row) * ncols + col];
      acc = fadd(acc, fmul(from_i64(g), 1ull << (8 * l)));
    }
    comp[cp] = acc;
  }
  // output layout y[w][j][rho]
  size_t w = row / D, rho = row % D;
  y[(w * t + j) * D + rho] = K2{comp[0], comp[1]};
}

// out[e] = sum_sp part[sp * n + e]: the K-split partials reduced with coalesced loads
__global__ void k_sum_splits_i32(const int32_t* __restrict__ part, size_t splits, size_t n, int32_t* __restrict__ out) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int32_t acc = 0;
  for (size_t sp = 0; sp < splits; sp++) acc += part[sp * n + e];
  out[e] = acc;
}

void launch_sum_splits(const int32_t* part, size_t splits, size_t n, int32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_splits_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, splits, n, out);
}

static size_t yd_splits(size_t m) {
  size_t s = m / 2048;
  if (s < 1) s = 1;
  if (s > YD_SPLIT_MAX) s = YD_SPLIT_MAX;
  while (s > 1 && (m / s) % 64 != 0) s--;
  return s;
}

size_t ydigits_ws_bytes(size_t nw, size_t t, size_t m) {
  size_t nrows = nw * D, mpad = (nrows + 63) / 64 * 64;
  return (yd_splits(m) + 1) * mpad * 16 * t * 4 + 16 * t * m + nrows * 4 + 2048;
}

void launch_ydigits(const int8_t* const* Zs_dev, size_t nw, const K2* v, size_t t, size_t m, K2* y, void* ws,
                    hipStream_t s) {
  const size_t nrows = nw * D, mpad = (nrows + 63) / 64 * 64, splits = yd_splits(m), kc = m / splits;
  int32_t* part = (int32_t*)ws;
  int8_t* Vl = (int8_t*)(part + splits * mpad * 16 * t);
  int32_t* rowsum = (int32_t*)(((uintptr_t)(Vl + 16 * t * m) + 255) & ~(uintptr_t)255);
  int32_t* red = (int32_t*)(((uintptr_t)(rowsum + nrows) + 255) & ~(uintptr_t)255);
  hipLaunchKernelGGL(k_v_limbs, dim3((unsigned)((m + 255) / 256), (unsigned)t), dim3(256), 0, s, v, t, m, Vl);
  dim3 g((unsigned)(mpad / 64), (unsigned)splits);
#define NS_YD(T)                                                                                              \
  do {                                                                                                        \
    if (kc % 256 == 0 && (T) <= 4)                                                                             \
      NS_LAUNCH_TIMED(KT_YD_GEMM, k_ydigits_gemm_lds<T>, g, dim3(256), 0, s, Zs_dev, nrows, Vl, m, kc, part, mpad); \
    else                                                                                                      \
      NS_LAUNCH_TIMED(KT_YD_GEMM, k_ydigits_gemm<T>, g, dim3(256), 0, s, Zs_dev, nrows, Vl, m, kc, part, mpad);     \
  } while (0)
  switch (t) {
    case 1: NS_YD(1); break;
    case 2: NS_YD(2); break;
    case 3: NS_YD(3); break;
    case 4: NS_YD(4); break;
    case 5: NS_YD(5); break;
    case 6: NS_YD(6); break;
    case 7: NS_YD(7); break;
    default: NS_YD(8); break;
  }
#undef NS_YD
  size_t nt = nrows * t;
  launch_sum_splits(part, splits, mpad * 16 * t, red, s);
  hipLaunchKernelGGL(k_ydigits_epilogue, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, red, (size_t)1, mpad,
                     nrows, t, y);
}

// ---------------- round polynomial ----------------
constexpr int RB = 256;        // threads per block
constexpr int RMAXBLK = 2048;  // partial slots
constexpr int FUSED_NCOEF = 7;  // P0..P4, Q0, Q1 of a fused round

struct RoundArgs {
  size_t len;
  const K2* eqb;
  const K2* eqr;
  const K2* ev;
  const K2* fv;
  int nf;
  const K2* nc;
  const int8_t* const* digits;
  size_t m;
  int k;
  K2 gamma_k;
  FPoly f;
  const K2* pc;  // round 1: [20][4] coefficient polynomials of the digit monomials
  const K2* U;   // modes 3, 4: [monomial][pair] weighted digit-monomial sums (k_digit_monomials)
  int r1cs[3];   // f = c0 x_a x_b + c1 x_c: a, b, c (r1cs[0] < 0 for other shapes), c0, c1
  uint64_t r1cs_c[2];
  int nb_u, nslice;  // mode 6: E-sum blocks (monomials x pair slices) ahead of the item blocks
};

// Wave sum of K values, valid in lane 63 (DPP_LANE): each component is summed lazily as a 64-bit
// value plus a carry count with DPP row shifts and row broadcasts (register-to-register moves
// across lanes, no LDS round trips), then reduced once.  Lanes a DPP step reads from outside the
// row (or that its row mask leaves out) add zero.
constexpr int DPP_LANE = 63;
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, true);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void dpp_lazy_add(uint64_t& v, uint32_t& c) {
  const uint32_t lo = dpp_u32<CTRL, ROW_MASK>((uint32_t)v), hi = dpp_u32<CTRL, ROW_MASK>((uint32_t)(v >> 32));
  const uint32_t cc = dpp_u32<CTRL, ROW_MASK>(c);
  const uint64_t o = ((uint64_t)hi << 32) | lo;
  const uint64_t t = v + o;
  c += cc + (uint32_t)(t < o);
  v = t;
}
__device__ __forceinline__ uint64_t wave_sum_fq(uint64_t v) {
  uint32_t c = 0;
  dpp_lazy_add<0x111, 0xf>(v, c);  // row_shr:1
  dpp_lazy_add<0x112, 0xf>(v, c);  // row_shr:2
  dpp_lazy_add<0x114, 0xf>(v, c);  // row_shr:4
  dpp_lazy_add<0x118, 0xf>(v, c);  // row_shr:8: lane 15 of each row holds the row's sum
  dpp_lazy_add<0x142, 0xa>(v, c);  // row_bcast:15 into rows 1 and 3
  dpp_lazy_add<0x143, 0xc>(v, c);  // row_bcast:31 into rows 2 and 3: lane 63 holds the total
  return reduce128_dev(v, c);
}
__device__ __forceinline__ K2 wave_sum(K2 v) { return K2{wave_sum_fq(v.c0), wave_sum_fq(v.c1)}; }

// c0 (a + a' X)(b + b' X) + c1 (e + e' X) into in[0..4] (the f polynomial of the R1CS shape on one
// pair's values)
__device__ __forceinline__ void r1cs_poly(K2 al, K2 ad, K2 bl, K2 bd, K2 el, K2 ed, uint64_t c0, uint64_t c1,
                                          K2 (&in)[5]) {
  const uint64_t al7 = fmul7(al.c1), ad7 = fmul7(ad.c1);
  K2 p0 = kmul_w7(al, al7, bl), p1 = kadd(kmul_w7(al, al7, bd), kmul_w7(ad, ad7, bl)), p2 = kmul_w7(ad, ad7, bd);
  if (c0 != 1) {
    p0 = kscale(p0, c0);
    p1 = kscale(p1, c0);
    p2 = kscale(p2, c0);
  }
  if (c1 == Q - 1) {
    in[0] = ksub(p0, el);
    in[1] = ksub(p1, ed);
  } else {
    in[0] = kadd(p0, kscale(el, c1));
    in[1] = kadd(p1, kscale(ed, c1));
  }
  in[2] = p2;
  in[3] = in[4] = kzero();
}

// ---- rounds 0 and 1 on int8 MFMA ----
// Both rounds need, for every digit monomial j and pair t, the weighted sum over the k*54 digit
// rows U_j[t] = sum_row wg[row] mono_j(row, t), mono_j in {-1, 0, 1}: round 0 has the four
// monomials lo, hi, |lo| hi, lo |hi| of a digit pair, round 1 the twenty of a digit quad (see
// k_round_eval).  That is a GEMM of the weights by the monomials: A = the weights' balanced
// base-256 limbs (16 limb rows, 8 per F_q component, in [-128, 127]; instance w's 54 rows at
// k = 64 w + rho, the 10 rows after them zero), B = the monomials (k = row, n = pair), on
// v_mfma_i32_16x16x64_i8 with int32 accumulation (|sum| < 2^18 per limb).  A wave owns 16 pairs
// (round 1) or 32 (round 0, two n tiles) and all rows; lane (n = lane % 16, k group = lane / 16)
// loads one digit word per row for rows 16 (lane / 16) .. + 15 of each instance -- 64 contiguous
// bytes per row across the 16 lanes -- transposes four rows' words into digit-major bytes with
// byte permutes and forms the monomials on packed bytes: for d in {0x00, 0x01, 0xFF},
// |d_i| d_j = d_j & mask(d_i) and d_i d_j d_l = (d_i ^ d_j ^ d_l) & masks (the xor of three
// nonzero signs is their product).  The epilogue turns each pair's 16 limb sums into the K value
// U_j[t] (limbs 0..3 + 2^32 limbs 4..7 per component, modulo q) and stores it.
__device__ __forceinline__ uint32_t mono_mask(uint32_t x) {
  const uint32_t e = x & 0x01010101u;  // |d| per byte
  return (e << 8) - e;                 // 0xFF where d != 0
}
__device__ __forceinline__ uint64_t from_i64_dev(int64_t v) { return v >= 0 ? (uint64_t)v : Q - (uint64_t)(-v); }

template <int MODE>  // 0: round 0 (digit pairs, 32 pairs per wave), 1: round 1 (digit quads, 16 pairs per wave)
__global__ __launch_bounds__(256) void k_digit_monomials(const int8_t* const* __restrict__ digits, int k, size_t m,
                                                        const int8_t* __restrict__ wl, size_t pairs,
                                                        K2* __restrict__ U) {
  constexpr int NT = MODE == 0 ? 2 : 1;   // n tiles per wave
  constexpr int J = MODE == 0 ? 4 : 20;   // monomials
  constexpr int NACC = J * NT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  const size_t tb = ((size_t)blockIdx.x * 4 + wave) * (16 * NT);
  if (tb >= pairs) return;  // whole waves only: pairs is a multiple of 16 NT
  // digit byte offset of this lane's word: 4 digits of pair tb + r16 (round 1), or 2 digits of
  // each of the pairs tb + 2 r16, tb + 2 r16 + 1 (round 0)
  const size_t off = MODE == 0 ? 2 * (tb + 2 * (size_t)r16) : 4 * (tb + (size_t)r16);
  const int lim = min(16, D - 16 * kg);  // rows of this lane's group that exist (kg = 3: 6)
  v4i acc[NACC];
#pragma unroll
  for (int a = 0; a < NACC; a++) acc[a] = v4i{0, 0, 0, 0};
  const size_t ldk = (size_t)64 * k;  // limb row length
  uint32_t cur[16], nxt[16];
  auto load = [&](int w, uint32_t (&x)[16]) {
    const int8_t* base = digits[w] + (size_t)(16 * kg) * m + off;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = *(const uint32_t*)(base + (size_t)(i < lim ? i : lim - 1) * m);
  };
  load(0, cur);
  for (int w = 0; w < k; w++) {
    if (w + 1 < k) load(w + 1, nxt);
    const v4i a = *(const v4i*)(wl + (size_t)r16 * ldk + (size_t)w * 64 + 16 * kg);
    // digit-major bytes of rows 4q .. 4q+3: T[q][i] byte b = digit i of row 4q + b
    uint32_t T[4][4], M[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t x0 = cur[4 * q], x1 = cur[4 * q + 1], x2 = cur[4 * q + 2], x3 = cur[4 * q + 3];
      const uint32_t l01 = __builtin_amdgcn_perm(x1, x0, 0x05010400u), h01 = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
      const uint32_t l23 = __builtin_amdgcn_perm(x3, x2, 0x05010400u), h23 = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
      T[q][0] = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
      T[q][1] = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
      T[q][2] = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
      T[q][3] = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
#pragma unroll
      for (int i = 0; i < 4; i++) M[q][i] = mono_mask(T[q][i]);
    }
#define NS_MONO(slot, expr)                                                   \
  {                                                                           \
    v4i b;                                                                    \
    _Pragma("unroll") for (int q = 0; q < 4; q++) {                           \
      const uint32_t* t_ = T[q];                                              \
      const uint32_t* m_ = M[q];                                              \
      (void)t_;                                                               \
      (void)m_;                                                               \
      b[q] = (int)(expr);                                                     \
    }                                                                         \
    acc[slot] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc[slot], 0, 0, 0); \
  }
    if (MODE == 0) {
      // n tile 0: pairs tb + 2 r16 (digits 0, 1 of the word), tile 1: tb + 2 r16 + 1 (digits 2, 3)
      NS_MONO(0, t_[0])
      NS_MONO(1, t_[1])
      NS_MONO(2, t_[1] & m_[0])
      NS_MONO(3, t_[0] & m_[1])
      NS_MONO(4, t_[2])
      NS_MONO(5, t_[3])
      NS_MONO(6, t_[3] & m_[2])
      NS_MONO(7, t_[2] & m_[3])
    } else {
      NS_MONO(0, t_[0])
      NS_MONO(1, t_[1])
      NS_MONO(2, t_[2])
      NS_MONO(3, t_[3])
      NS_MONO(4, t_[1] & m_[0])
      NS_MONO(5, t_[2] & m_[0])
      NS_MONO(6, t_[3] & m_[0])
      NS_MONO(7, t_[0] & m_[1])
      NS_MONO(8, t_[2] & m_[1])
      NS_MONO(9, t_[3] & m_[1])
      NS_MONO(10, t_[0] & m_[2])
      NS_MONO(11, t_[1] & m_[2])
      NS_MONO(12, t_[3] & m_[2])
      NS_MONO(13, t_[0] & m_[3])
      NS_MONO(14, t_[1] & m_[3])
      NS_MONO(15, t_[2] & m_[3])
      NS_MONO(16, (t_[0] ^ t_[1] ^ t_[2]) & m_[0] & m_[1] & m_[2])
      NS_MONO(17, (t_[0] ^ t_[1] ^ t_[3]) & m_[0] & m_[1] & m_[3])
      NS_MONO(18, (t_[0] ^ t_[2] ^ t_[3]) & m_[0] & m_[2] & m_[3])
      NS_MONO(19, (t_[1] ^ t_[2] ^ t_[3]) & m_[1] & m_[2] & m_[3])
    }
#undef NS_MONO
    if (w + 1 < k) {
#pragma unroll
      for (int i = 0; i < 16; i++) cur[i] = nxt[i];
    }
  }
  // lane holds limbs 4 kg .. 4 kg + 3 of its pair column: component kg / 2, bytes 4 (kg & 1) ..
#pragma unroll
  for (int a = 0; a < NACC; a++) {
    const int64_t p = (int64_t)acc[a][0] + ((int64_t)acc[a][1] << 8) + ((int64_t)acc[a][2] << 16) +
                      ((int64_t)acc[a][3] << 24);
    const int plo = __shfl_xor((int)(uint32_t)p, 16, 64), phi = __shfl_xor((int)(uint32_t)((uint64_t)p >> 32), 16, 64);
    const int64_t hi = (int64_t)(((uint64_t)(uint32_t)phi << 32) | (uint32_t)plo);
    if ((kg & 1) == 0) {
      const uint64_t v = fadd(from_i64_dev(p), fmul(from_i64_dev(hi), (uint64_t)1 << 32));
      const int j = a % J, tile = a / J;
      const size_t t = MODE == 0 ? tb + 2 * (size_t)r16 + tile : tb + (size_t)r16;
      uint64_t* dst = (uint64_t*)(U + (size_t)j * pairs + t);
      dst[kg >> 1] = v;
    }
  }
}

// Round 2 from the digits: each pair's two table entries fold digits 8t .. 8t+3 and 8t+4 .. 8t+7,
// so N(y) expands into 120 monomials of 8 digits: d_i (8), |d_i| d_j (i != j, 56, i-major) and
// d_i d_j d_l (i < j < l, 56, lexicographic).  A block owns 16 pairs; its six waves take 20
// monomials each.  Per instance the block loads the 64 x 16 8-byte digit words once (thread u:
// pair u % 16, rows 4 (u / 16) .. + 3), transposes them to digit-major bytes with their masks
// into LDS (double-buffered, 20-dword rows so that the 16-byte reads of a 16-lane group hit
// distinct banks), and every wave builds its monomials' B fragments from there.
struct Mono8 {
  int n, a, b, c;
};
__host__ __device__ constexpr Mono8 mono8_of(int j) {
  if (j < 8) return Mono8{1, j, 0, 0};
  if (j < 64) {
    const int p = j - 8, i = p / 7, r = p % 7;
    return Mono8{2, i, r < i ? r : r + 1, 0};
  }
  int idx = j - 64;
  for (int a = 0; a < 8; a++)
    for (int b = a + 1; b < 8; b++)
      for (int c = b + 1; c < 8; c++) {
        if (idx == 0) return Mono8{3, a, b, c};
        idx--;
      }
  return Mono8{0, 0, 0, 0};
}
template <int J8>
__device__ __forceinline__ uint32_t mono8(const uint32_t* T, const uint32_t* M) {
  constexpr Mono8 mo = mono8_of(J8);
  if constexpr (mo.n == 1)
    return T[mo.a];
  else if constexpr (mo.n == 2)
    return T[mo.b] & M[mo.a];
  else
    return (T[mo.a] ^ T[mo.b] ^ T[mo.c]) & M[mo.a] & M[mo.b] & M[mo.c];
}
constexpr int M8_GROUP = 20, M8_NGROUPS = 6, M8_LDS_ROW = 20;
// monomials 20 G .. 20 G + 19 of the four 4-row words (T[q][0..7], masks T[q][8..15]), each
// straight into its MFMA
template <int G, int... I>
__device__ __forceinline__ void mono8_mfma(const uint32_t (&T)[4][16], v4i a, v4i (&acc)[M8_GROUP],
                                           std::integer_sequence<int, I...>) {
  ((acc[I] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
        a,
        v4i{(int)mono8<M8_GROUP * G + I>(T[0], T[0] + 8), (int)mono8<M8_GROUP * G + I>(T[1], T[1] + 8),
            (int)mono8<M8_GROUP * G + I>(T[2], T[2] + 8), (int)mono8<M8_GROUP * G + I>(T[3], T[3] + 8)},
        acc[I], 0, 0, 0)),
   ...);
}

__global__ __launch_bounds__(64 * M8_NGROUPS) void k_digit_monomials8(const int8_t* const* __restrict__ digits, int k,
                                                                     size_t m, const int8_t* __restrict__ wl,
                                                                     size_t pairs, K2* __restrict__ U) {
  __shared__ __align__(16) uint32_t tm[2][16][16][M8_LDS_ROW];  // [buffer][row quad][pair][T 0..7, M 0..7]
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  const size_t tb = (size_t)blockIdx.x * 16;
  // transpose unit of thread u < 256: pair u % 16, row quad u / 16 (rows past 53 re-read row 53:
  // their weight limbs are zero)
  const int u = threadIdx.x, un = u & 15, uq = u >> 4;
  const bool unit = u < 256;
  const size_t uoff = 8 * (tb + (size_t)un);
  uint2 x[4];
  auto load = [&](int w) {
    const int8_t* base = digits[w] + uoff;
#pragma unroll
    for (int i = 0; i < 4; i++) x[i] = *(const uint2*)(base + (size_t)min(4 * uq + i, D - 1) * m);
  };
  if (unit) load(0);
  v4i acc[M8_GROUP];
#pragma unroll
  for (int a = 0; a < M8_GROUP; a++) acc[a] = v4i{0, 0, 0, 0};
  const size_t ldk = (size_t)64 * k;
  for (int w = 0; w < k; w++) {
    const int buf = w & 1;
    if (unit) {
      uint32_t T[16];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint32_t x0 = h ? x[0].y : x[0].x, x1 = h ? x[1].y : x[1].x;
        const uint32_t x2 = h ? x[2].y : x[2].x, x3 = h ? x[3].y : x[3].x;
        const uint32_t l01 = __builtin_amdgcn_perm(x1, x0, 0x05010400u), h01 = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
        const uint32_t l23 = __builtin_amdgcn_perm(x3, x2, 0x05010400u), h23 = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
        T[4 * h + 0] = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
        T[4 * h + 1] = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
        T[4 * h + 2] = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
        T[4 * h + 3] = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
      }
#pragma unroll
      for (int i = 0; i < 8; i++) T[8 + i] = mono_mask(T[i]);
      uint4* dst = (uint4*)tm[buf][uq][un];
#pragma unroll
      for (int v = 0; v < 4; v++) dst[v] = uint4{T[4 * v], T[4 * v + 1], T[4 * v + 2], T[4 * v + 3]};
      if (w + 1 < k) load(w + 1);  // in flight during this instance's MFMAs
    }
    __syncthreads();  // buffer `buf` written; the other buffer's readers (instance w - 1) are done
    const v4i a = *(const v4i*)(wl + (size_t)r16 * ldk + (size_t)w * 64 + 16 * kg);
    uint32_t T[4][16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4* src = (const uint4*)tm[buf][4 * kg + q][r16];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint4 y = src[v];
        T[q][4 * v] = y.x;
        T[q][4 * v + 1] = y.y;
        T[q][4 * v + 2] = y.z;
        T[q][4 * v + 3] = y.w;
      }
    }
    constexpr auto seq = std::make_integer_sequence<int, M8_GROUP>{};
    switch (grp) {
      case 0: mono8_mfma<0>(T, a, acc, seq); break;
      case 1: mono8_mfma<1>(T, a, acc, seq); break;
      case 2: mono8_mfma<2>(T, a, acc, seq); break;
      case 3: mono8_mfma<3>(T, a, acc, seq); break;
      case 4: mono8_mfma<4>(T, a, acc, seq); break;
      default: mono8_mfma<5>(T, a, acc, seq); break;
    }
  }
#pragma unroll
  for (int jj = 0; jj < M8_GROUP; jj++) {
    const int64_t p = (int64_t)acc[jj][0] + ((int64_t)acc[jj][1] << 8) + ((int64_t)acc[jj][2] << 16) +
                      ((int64_t)acc[jj][3] << 24);
    const int plo = __shfl_xor((int)(uint32_t)p, 16, 64), phi = __shfl_xor((int)(uint32_t)((uint64_t)p >> 32), 16, 64);
    const int64_t hi = (int64_t)(((uint64_t)(uint32_t)phi << 32) | (uint32_t)plo);
    if ((kg & 1) == 0) {
      const uint64_t v = fadd(from_i64_dev(p), fmul(from_i64_dev(hi), (uint64_t)1 << 32));
      uint64_t* dst = (uint64_t*)(U + (size_t)(grp * M8_GROUP + jj) * pairs + tb + r16);
      dst[kg >> 1] = v;
    }
  }
}


// coefficient vector of width 5 (deg_max = 4).  Work item = (row pair t, row group g): the NC
// sum over the k*54 digit rows of a pair is split into G groups so that short tables still fill
// the GPU; every group multiplies its partial NC polynomial by the pair's eq_beta(X) (the
// round polynomial is linear in the NC sum), and group 0 adds the f and Eval terms.
// MODE 0: round 0 over the int8 digits; MODE 1: round 1 over the int8 digits (tables folded once
// by r0, see below); MODE 2: K tables; MODE 6: no rows (G = 0): the NC part from the weighted
// monomial sums U of k_digit_monomials(8) in the first A.nb_u blocks, the f and Eval items after.
template <int MODE>
__global__ __launch_bounds__(RB) void k_round_eval(RoundArgs A, const K2* __restrict__ wg, K2* __restrict__ partials,
                                                   int G) {
  // LDS: wg [k*54], round-1 monomial polynomials [80], reduction scratch [20], then -wg [k*54]
  // (rounds 0 and 1) or 7 wg.c1 [k*54] (K tables), then 7 pc.c1 [80]
  extern __shared__ K2 sh[];
  const int nrow = A.k * D;
  K2* pcs = sh + nrow;
  K2* nsh = sh + nrow + 100;
  uint64_t* w7s = (uint64_t*)(sh + nrow + 100);
  uint64_t* pcs7 = (uint64_t*)(sh + 2 * nrow + 100);
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
    const K2 w = wg[i];
    sh[i] = w;
    if (MODE == 2)
      w7s[i] = fmul7(w.c1);
    else
      nsh[i] = kneg(w);
  }
  __shared__ const int8_t* dptr[32];  // digit planes of the k instances (rounds 0 and 1)
  if ((MODE == 0 || MODE == 1) && threadIdx.x < A.k && threadIdx.x < 32) dptr[threadIdx.x] = A.digits[threadIdx.x];
  if (MODE == 1)
    for (int i = threadIdx.x; i < 80; i += blockDim.x) {
      const K2 p = A.pc[i];
      pcs[i] = p;
      pcs7[i] = fmul7(p.c1);
    }
  __syncthreads();
  const size_t pairs = A.len / 2;
  const int R = G > 0 ? (nrow + G - 1) / G : 0;  // G = 0 (mode 6): the f and Eval items only
  // groups 0..G-1 take R digit rows each; group G takes the f and Eval terms of the pair
  K2 c[5];
#pragma unroll
  for (int d = 0; d < 5; d++) c[d] = kzero();
  K2 ev0 = kzero(), ev1 = kzero(), ev2 = kzero();
  // round 1 splits each (pair, row group) item into two halves of the 20 digit monomials
  const int H = MODE == 1 ? 2 : 1;
  // items per pair: G * H row items, then the f item and (with r_in) the Eval item
  const int nfe = A.eqr ? 2 : 1;
  const size_t total = pairs * ((size_t)G * H + nfe);
  const int lp = __builtin_ctzll(pairs);  // pairs is a power of two (tables are 2^ell_n long)
  // mode 6: blocks below A.nb_u take the NC part from U (E-sums of monomial j over a slice of
  // pairs, times pc_j), the rest take the f and Eval items
  unsigned bid = blockIdx.x, nbk = gridDim.x;
  if (MODE == 6 && blockIdx.x < (unsigned)A.nb_u) {
    const int j = blockIdx.x / A.nslice, sl = blockIdx.x % A.nslice;
    const size_t per = pairs / A.nslice, t0 = (size_t)sl * per;
    KS2 a0 = ks2_zero(), a1 = ks2_zero();
    for (size_t t = t0 + threadIdx.x; t < t0 + per; t += 2 * blockDim.x) {
      // two pairs' loads in flight at once (per is a multiple of 2 blockDim or below blockDim)
      const size_t t2 = t + blockDim.x;
      const bool two = t2 < t0 + per;
      const K2 e0 = A.eqb[2 * t], e0h = A.eqb[2 * t + 1], u = A.U[(size_t)j * pairs + t];
      K2 f0 = kzero(), f0h = kzero(), v = kzero();
      if (two) {
        f0 = A.eqb[2 * t2];
        f0h = A.eqb[2 * t2 + 1];
        v = A.U[(size_t)j * pairs + t2];
      }
      ks2_mac(a0, e0, u);
      ks2_mac(a1, ksub(e0h, e0), u);
      ks2_mac(a0, f0, v);
      ks2_mac(a1, ksub(f0h, f0), v);
    }
    // block totals of E0, E1, then coefficient d = pc_j[d] E0 + pc_j[d - 1] E1 by thread d
    K2* red = sh + nrow + 80;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const K2 s0 = wave_sum(ks2_reduce(a0)), s1 = wave_sum(ks2_reduce(a1));
    if (lane == DPP_LANE) {
      red[wave] = s0;
      red[4 + wave] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
      const int d = threadIdx.x;
      const K2 E0 = kadd(kadd(red[0], red[1]), kadd(red[2], red[3]));
      const K2 E1 = kadd(kadd(red[4], red[5]), kadd(red[6], red[7]));
      K2 v = d < 4 ? kmul(A.pc[j * 4 + d], E0) : kzero();
      if (d >= 1) v = kadd(v, kmul(A.pc[j * 4 + d - 1], E1));
      partials[(size_t)blockIdx.x * 5 + d] = v;
    }
    return;  // block-uniform
  }
  bool items = true;
  if (MODE == 6) {
    items = blockIdx.x >= (unsigned)A.nb_u;  // the E-sum blocks take no items
    bid = items ? blockIdx.x - A.nb_u : 0;
    nbk -= A.nb_u;
  }
  for (size_t idx = items ? (size_t)bid * blockDim.x + threadIdx.x : total; idx < total;
       idx += (size_t)nbk * blockDim.x) {
    const size_t t = idx & (pairs - 1);
    const int gh = (int)(idx >> lp);
    const bool fe = gh >= G * H;  // f item (gh = G H) or Eval item (gh = G H + 1)
    if (fe && gh > G * H) {
      // gamma^k eq_rin(X) Eval(X), summed apart and scaled once
      const K2 r0 = A.eqr[2 * t], r1 = ksub(A.eqr[2 * t + 1], r0);
      const K2 v0 = A.ev[2 * t], v1 = ksub(A.ev[2 * t + 1], v0);
      ev0 = kadd(ev0, kmul(r0, v0));
      ev1 = kadd(ev1, kadd(kmul(r0, v1), kmul(r1, v0)));
      ev2 = kadd(ev2, kmul(r1, v1));
      continue;
    }
    const int g = fe ? G : gh / H, h = fe ? 0 : gh % H;
    const int r_lo = min(nrow, g * R), r_hi = min(nrow, r_lo + R);
    const K2 e0 = A.eqb[2 * t];
    const K2 e1 = ksub(A.eqb[2 * t + 1], e0);
    K2 in[5];
#pragma unroll
    for (int d = 0; d < 5; d++) in[d] = kzero();
    if (fe && A.r1cs[0] >= 0) {
      // f(M z1)(X) of the R1CS shape, straight-line
      const K2* ta = A.fv + (size_t)A.r1cs[0] * A.len + 2 * t;
      const K2* tb = A.fv + (size_t)A.r1cs[1] * A.len + 2 * t;
      const K2* tc = A.fv + (size_t)A.r1cs[2] * A.len + 2 * t;
      const K2 al = ta[0], ah = ta[1], bl = tb[0], bh = tb[1], el = tc[0], eh = tc[1];
      r1cs_poly(al, ksub(ah, al), bl, ksub(bh, bl), el, ksub(eh, el), A.r1cs_c[0], A.r1cs_c[1], in);
    } else if (fe) {
      // f(M z1)(X)
      for (int tt = 0; tt < A.f.nterms; tt++) {
        K2 p[5];
        p[0] = K2{A.f.coeff[tt], 0};
#pragma unroll
        for (int d = 1; d < 5; d++) p[d] = kzero();
        for (int v = 0; v < A.f.nv[tt]; v++) {
          const K2* tbl = A.fv + (size_t)A.f.var[tt][v] * A.len;
          K2 a = tbl[2 * t], b = ksub(tbl[2 * t + 1], a);
          for (int e = 0; e < A.f.exp[tt][v]; e++) {
#pragma unroll
            for (int d = 4; d >= 1; d--) p[d] = kadd(kmul(a, p[d]), kmul(b, p[d - 1]));
            p[0] = kmul(a, p[0]);
          }
        }
#pragma unroll
        for (int d = 0; d < 5; d++) in[d] = kadd(in[d], p[d]);
      }
    } else if (MODE == 0) {
      // signed selections of wg, summed lazily (+w or q - w) and reduced once per work item
      KS1 zlo = ks1_zero(), zhi = ks1_zero(), z3 = ks1_zero(), z4 = ks1_zero();
      for (int r0 = r_lo; r0 < r_hi; r0 += 8) {
        // eight rows' digit pairs in flight at once (rows past r_hi re-read row r_hi - 1, then count as 0)
        uint32_t qd[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int rr = min(r0 + u, r_hi - 1), w = rr / D, rho = rr - w * D;
          const uint32_t x = *(const uint16_t*)(dptr[w] + (size_t)rho * A.m + 2 * t);
          qd[u] = r0 + u < r_hi ? x : 0u;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
        const int r = min(r0 + u, r_hi - 1);
        if (qd[u] == 0) continue;
        const int lo = (int)(int8_t)(qd[u] & 0xFF), hi = (int)(int8_t)(qd[u] >> 8);
        const K2 wv = sh[r], nw = nsh[r];
        const K2 vlo = lo > 0 ? wv : (lo < 0 ? nw : kzero());
        const K2 vhi = hi > 0 ? wv : (hi < 0 ? nw : kzero());
        const bool both = lo != 0 && hi != 0;
        ks1_add(zlo, vlo);
        ks1_add(zhi, vhi);
        ks1_add(z3, both ? vhi : kzero());  // lo^2 hi = hi
        ks1_add(z4, both ? vlo : kzero());  // lo hi^2 = lo
        }
      }
      const K2 slo = ks1_reduce(zlo), shi = ks1_reduce(zhi), s3 = ks1_reduce(z3), s4 = ks1_reduce(z4);
      // c1 = -2 slo - shi + 3 s3 ; c2 = 3 slo - 6 s3 + 3 s4 ; c3 = -slo + shi + 3 s3 - 3 s4
      K2 s3x3 = kadd(kadd(s3, s3), s3);
      K2 s4x3 = kadd(kadd(s4, s4), s4);
      K2 slo2 = kadd(slo, slo);
      in[1] = kadd(in[1], ksub(ksub(s3x3, slo2), shi));
      in[2] = kadd(in[2], ksub(kadd(kadd(slo2, slo), s4x3), kadd(s3x3, s3x3)));
      in[3] = kadd(in[3], ksub(kadd(shi, s3x3), kadd(slo, s4x3)));
    } else if (MODE == 1) {
      // Round 1 straight from the digits: with alpha = 1 - r0, beta = r0 the pair's table values
      // are a = alpha d0 + beta d1, a' = alpha d2 + beta d3 (d_i = digits 4t..4t+3), and
      // N(a + (a' - a) X) = sum over the 20 digit monomials {d_i, |d_i| d_j (i != j),
      // d_i d_j d_l (i < j < l)} of a fixed cubic in X (expanding (sum_i L_i(X) d_i)^3 with
      // d^3 = d, d^2 = |d|).  Each row adds +-w to the monomials it hits; the cubics are applied
      // once per work item.
#define NS_UPD(acc, mv)                                         \
  {                                                             \
    const int mm = (mv);                                        \
    ks1_add(acc, mm > 0 ? wv : (mm < 0 ? nw : kzero()));        \
  }
#define NS_APPLY(j, acc)                                                         \
  {                                                                              \
    const K2 av = ks1_reduce(acc);                                               \
    in[0] = kadd(in[0], kmul_w7(pcs[(j)*4 + 0], pcs7[(j)*4 + 0], av));           \
    in[1] = kadd(in[1], kmul_w7(pcs[(j)*4 + 1], pcs7[(j)*4 + 1], av));           \
    in[2] = kadd(in[2], kmul_w7(pcs[(j)*4 + 2], pcs7[(j)*4 + 2], av));           \
    in[3] = kadd(in[3], kmul_w7(pcs[(j)*4 + 3], pcs7[(j)*4 + 3], av));           \
  }
      if (h == 0) {
        KS1 a0 = ks1_zero(), a1 = ks1_zero(), a2 = ks1_zero(), a3 = ks1_zero(), a4 = ks1_zero(), a5 = ks1_zero(),
            a6 = ks1_zero(), a7 = ks1_zero(), a8 = ks1_zero(), a9 = ks1_zero();
        for (int r0 = r_lo; r0 < r_hi; r0 += 8) {
          uint32_t qd[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int rr = min(r0 + u, r_hi - 1), w = rr / D, rho = rr - w * D;
            const uint32_t x = *(const uint32_t*)(dptr[w] + (size_t)rho * A.m + 4 * t);
            qd[u] = r0 + u < r_hi ? x : 0u;
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
          const int r = min(r0 + u, r_hi - 1);
          const uint32_t q = qd[u];
          if (q == 0) continue;  // most digit words of the DEC children are zero
          const int d0 = (int)(int8_t)(q & 0xFF), d1 = (int)(int8_t)((q >> 8) & 0xFF),
                    d2 = (int)(int8_t)((q >> 16) & 0xFF), d3 = (int)(int8_t)(q >> 24);
          const int e0 = d0 & 1, e1 = d1 & 1, e2 = d2 & 1, e3 = d3 & 1;
          (void)e0; (void)e1; (void)e2; (void)e3;
          const K2 wv = sh[r], nw = nsh[r];
          NS_UPD(a0, d0)
          NS_UPD(a1, d1)
          NS_UPD(a2, d2)
          NS_UPD(a3, d3)
          NS_UPD(a4, e0 * d1)
          NS_UPD(a5, e0 * d2)
          NS_UPD(a6, e0 * d3)
          NS_UPD(a7, e1 * d0)
          NS_UPD(a8, e1 * d2)
          NS_UPD(a9, e1 * d3)
          }
        }
        NS_APPLY(0, a0)
        NS_APPLY(1, a1)
        NS_APPLY(2, a2)
        NS_APPLY(3, a3)
        NS_APPLY(4, a4)
        NS_APPLY(5, a5)
        NS_APPLY(6, a6)
        NS_APPLY(7, a7)
        NS_APPLY(8, a8)
        NS_APPLY(9, a9)
      }
      else {
        KS1 a10 = ks1_zero(), a11 = ks1_zero(), a12 = ks1_zero(), a13 = ks1_zero(), a14 = ks1_zero(),
            a15 = ks1_zero(), a16 = ks1_zero(), a17 = ks1_zero(), a18 = ks1_zero(), a19 = ks1_zero();
        for (int r0 = r_lo; r0 < r_hi; r0 += 8) {
          uint32_t qd[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int rr = min(r0 + u, r_hi - 1), w = rr / D, rho = rr - w * D;
            const uint32_t x = *(const uint32_t*)(dptr[w] + (size_t)rho * A.m + 4 * t);
            qd[u] = r0 + u < r_hi ? x : 0u;
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
          const int r = min(r0 + u, r_hi - 1);
          const uint32_t q = qd[u];
          if (q == 0) continue;  // most digit words of the DEC children are zero
          const int d0 = (int)(int8_t)(q & 0xFF), d1 = (int)(int8_t)((q >> 8) & 0xFF),
                    d2 = (int)(int8_t)((q >> 16) & 0xFF), d3 = (int)(int8_t)(q >> 24);
          const int e0 = d0 & 1, e1 = d1 & 1, e2 = d2 & 1, e3 = d3 & 1;
          (void)e0; (void)e1; (void)e2; (void)e3;
          const K2 wv = sh[r], nw = nsh[r];
          NS_UPD(a10, e2 * d0)
          NS_UPD(a11, e2 * d1)
          NS_UPD(a12, e2 * d3)
          NS_UPD(a13, e3 * d0)
          NS_UPD(a14, e3 * d1)
          NS_UPD(a15, e3 * d2)
          NS_UPD(a16, d0 * d1 * d2)
          NS_UPD(a17, d0 * d1 * d3)
          NS_UPD(a18, d0 * d2 * d3)
          NS_UPD(a19, d1 * d2 * d3)
          }
        }
        NS_APPLY(10, a10)
        NS_APPLY(11, a11)
        NS_APPLY(12, a12)
        NS_APPLY(13, a13)
        NS_APPLY(14, a14)
        NS_APPLY(15, a15)
        NS_APPLY(16, a16)
        NS_APPLY(17, a17)
        NS_APPLY(18, a18)
        NS_APPLY(19, a19)
      }
#undef NS_UPD
#undef NS_APPLY
    } else if (MODE == 2) {
      // w N(a + b X) = w (a^3 - a) + w (3 a^2 b - b) X + 3 w a b^2 X^2 + w b^3 X^3: wa, wb, a^2, b^2
      // reduced per row, the four products and two sums accumulated lazily over the rows
      KS2 z1 = ks2_zero(), z3 = ks2_zero(), z5 = ks2_zero(), z6 = ks2_zero();
      KS1 z2 = ks1_zero(), z4 = ks1_zero();
      const K2* rowp = A.nc + (size_t)r_lo * A.len + 2 * t;
      K2 n0 = r_lo < r_hi ? rowp[0] : kzero(), n1 = r_lo < r_hi ? rowp[1] : kzero();
      for (int r = r_lo; r < r_hi; r++) {
        const K2 a = n0;
        const K2 b = ksub(n1, a);
        if (r + 1 < r_hi) {  // next row's pair in flight during this row's arithmetic
          rowp += A.len;
          n0 = rowp[0];
          n1 = rowp[1];
        }
        const K2 wv = sh[r];
        const uint64_t w7 = w7s[r];
        const K2 wa = kmul_w7(wv, w7, a), wb = kmul_w7(wv, w7, b);
        const K2 a2 = ksqr_dev(a), b2 = ksqr_dev(b);
        const uint64_t was = ksum(wa), wbs = ksum(wb), a2s = ksum(a2), b2s = ksum(b2);
        ks2k_mac(z1, wa, was, a2, a2s);
        ks1_add(z2, wa);
        ks2k_mac(z3, wb, wbs, a2, a2s);
        ks1_add(z4, wb);
        ks2k_mac(z5, wa, was, b2, b2s);
        ks2k_mac(z6, wb, wbs, b2, b2s);
      }
      const K2 s1 = ks2k_reduce(z1), s2 = ks1_reduce(z2), s3 = ks2k_reduce(z3), s4 = ks1_reduce(z4),
               s5 = ks2k_reduce(z5), s6 = ks2k_reduce(z6);
      in[0] = kadd(in[0], ksub(s1, s2));
      in[1] = kadd(in[1], ksub(kadd(kadd(s3, s3), s3), s4));
      in[2] = kadd(in[2], kadd(kadd(s5, s5), s5));
      in[3] = kadd(in[3], s6);
    }
    // times eq_beta(X) = e0 + e1 X (in[4] is nonzero only for the f group)
    const uint64_t e07 = fmul7(e0.c1), e17 = fmul7(e1.c1);
    c[0] = kadd(c[0], kmul_w7(e0, e07, in[0]));
#pragma unroll
    for (int d = 1; d < 4; d++) c[d] = kadd(c[d], kadd(kmul_w7(e0, e07, in[d]), kmul_w7(e1, e17, in[d - 1])));
    c[4] = kadd(c[4], kmul_w7(e1, e17, in[3]));
    if (fe) c[4] = kadd(c[4], kmul_w7(e0, e07, in[4]));
  }
  if (A.eqr) {
    c[0] = kadd(c[0], kmul(A.gamma_k, ev0));
    c[1] = kadd(c[1], kmul(A.gamma_k, ev1));
    c[2] = kadd(c[2], kmul(A.gamma_k, ev2));
  }
  // block reduction
  K2* red = sh + nrow + 80;  // [4 waves][5]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
#pragma unroll
  for (int d = 0; d < 5; d++) {
    K2 v = wave_sum(c[d]);
    if (lane == DPP_LANE) red[wave * 5 + d] = v;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    K2 v = kzero();
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) v = kadd(v, red[w * 5 + threadIdx.x]);
    partials[blockIdx.x * 5 + threadIdx.x] = v;
  }
}

// one block per coefficient: strided partial sums, wave reductions, then the 4 wave results
__global__ __launch_bounds__(256) void k_sum_partials(const K2* __restrict__ partials, int nblk, int ncoef,
                                                      K2* __restrict__ out) {
  const int d = blockIdx.x;
  K2 v = kzero();
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) v = kadd(v, partials[b * ncoef + d]);
  __shared__ K2 red[4];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == DPP_LANE) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[d] = kadd(kadd(red[0], red[1]), kadd(red[2], red[3]));
}

// The same sum in one block (wave d sums coefficient d), published to pinned host memory with a
// ready word: lane 0 of wave 0 stores the coefficients, then `seq` into *flag with a system-scope
// release, so a host that reads the flag with acquire sees the coefficients.  The host polls the
// word instead of waiting on the stream (the stream wait's wake-up cost ~30 us per round).
__global__ __launch_bounds__(512) void k_sum_partials_sig(const K2* __restrict__ partials, int nblk, int ncoef,
                                                          K2* __restrict__ out, uint64_t* flag, uint64_t seq) {
  __shared__ K2 res[8];
  const int d = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (d < ncoef) {
    K2 v = kzero();
    for (int b = lane; b < nblk; b += 64) v = kadd(v, partials[b * ncoef + d]);
    v = wave_sum(v);
    if (lane == DPP_LANE) res[d] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < ncoef; i++) out[i] = res[i];
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Block j: the sum of job j's partial coefficients, published with its ready word.
__global__ __launch_bounds__(512) void k_sum_partials_multi(const FusedJob* __restrict__ jobs) {
  __shared__ K2 res[8];
  const FusedJob& J = jobs[blockIdx.x];
  const int d = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (d < FUSED_NCOEF) {
    K2 v = kzero();
    for (int b = lane; b < J.nblk; b += 64) v = kadd(v, J.partials[b * FUSED_NCOEF + d]);
    v = wave_sum(v);
    if (lane == DPP_LANE) res[d] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < FUSED_NCOEF; i++) J.out[i] = res[i];
    __hip_atomic_store(J.flag, J.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

static void launch_sum_partials(const K2* partials, size_t nblk, int ncoef, K2* out, const RoundSignal* sig,
                                hipStream_t s) {
  if (sig && sig->flag)
    hipLaunchKernelGGL(k_sum_partials_sig, dim3(1), dim3(512), 0, s, partials, (int)nblk, ncoef, out, sig->flag,
                       sig->seq);
  else
    hipLaunchKernelGGL(k_sum_partials, dim3(ncoef), dim3(256), 0, s, partials, (int)nblk, ncoef, out);
}

size_t round_ws_bytes() { return (size_t)RMAXBLK * 7 * sizeof(K2); }

std::vector<int8_t> weight_limbs(const K2* wg, size_t k) {
  const size_t ld = 64 * k;
  std::vector<int8_t> out(16 * ld, 0);
  for (size_t w = 0; w < k; w++)
    for (int rho = 0; rho < D; rho++) {
      const K2 x = wg[w * D + rho];
      const uint64_t comp[2] = {x.c0, x.c1};
      for (int c = 0; c < 2; c++) {
        __int128 v = comp[c] <= 0x7F7F7F7F7F7F7F7Full ? (__int128)comp[c] : (__int128)comp[c] - (__int128)Q;
        for (int l = 0; l < 8; l++) {
          int d = (int)(v & 255);
          if (d >= 128) d -= 256;
          out[(size_t)(c * 8 + l) * ld + w * 64 + rho] = (int8_t)d;
          v = (v - d) >> 8;
        }
        if (v != 0) throw std::logic_error("weight_limbs: value outside the balanced range");
      }
    }
  return out;
}

// The degree of f (capped at 4, as the rounds drop higher terms) and its R1CS shape, if it has one:
// f = c0 x_a x_b + c1 x_c with a, b, c < nf (r1cs[0] = -1 otherwise).
static void classify_fpoly(const FPoly& f, int nf, int& fdeg, int (&r1cs)[3], uint64_t (&c)[2]) {
  int deg = 0;
  for (int tt = 0; tt < f.nterms; tt++) {
    int d = 0;
    for (int v = 0; v < f.nv[tt]; v++) d += f.exp[tt][v];
    deg = std::max(deg, std::min(d, 4));
  }
  fdeg = deg;
  r1cs[0] = r1cs[1] = r1cs[2] = -1;
  c[0] = c[1] = 0;
  if (f.nterms != 2) return;
  for (int p = 0; p < 2; p++) {
    const int l = 1 - p;  // p: the product term, l: the linear term
    if (f.nv[p] != 2 || f.exp[p][0] != 1 || f.exp[p][1] != 1 || f.nv[l] != 1 || f.exp[l][0] != 1) continue;
    const int v[3] = {f.var[p][0], f.var[p][1], f.var[l][0]};
    bool ok = true;
    for (int i = 0; i < 3; i++) ok = ok && v[i] >= 0 && v[i] < nf;
    if (!ok) return;
    for (int i = 0; i < 3; i++) r1cs[i] = v[i];
    c[0] = f.coeff[p];
    c[1] = f.coeff[l];
    return;
  }
}

void launch_round_eval(const RowTables& T, const FPoly& f, const K2* wg, bool round0, K2* out, void* ws,
                       hipStream_t s, const RoundSignal* sig) {
  if (T.len < 2 || (T.len & (T.len - 1))) throw std::invalid_argument("round: table length must be a power of two >= 2");
  RoundArgs A;
  A.len = T.len;
  A.eqb = T.eqb;
  A.eqr = T.eqr;
  A.ev = T.ev;
  A.fv = T.fv;
  A.nf = T.nf;
  A.nc = T.nc;
  A.digits = T.digits;
  A.m = T.m;
  A.k = T.k;
  A.gamma_k = T.gamma_k;
  A.f = f;
  int fdeg = 0;
  classify_fpoly(f, T.nf, fdeg, A.r1cs, A.r1cs_c);
  size_t pairs = T.len / 2;
  K2* partials = (K2*)ws;
  A.pc = T.pc;
  A.U = nullptr;
  A.nb_u = 0;
  A.nslice = 1;
  // rounds 0, 1 and 2 on int8 MFMA when the tables allow whole 128-pair blocks (round 2: 64) and U
  // fits the buffer given for it: k_digit_monomials(8) -> U, then mode 6 (the NC part from U, the
  // f and Eval items), one partial sum
  const int dmode = round0 ? 0 : (T.digit_round1 ? 1 : (T.digit_round2 ? 2 : -1));
  const size_t nmono = dmode == 0 ? 4 : (dmode == 1 ? 20 : 120);
  const size_t min_pairs = dmode == 2 ? 64 : 128;
  const size_t digits_per_pair = dmode == 0 ? 2 : (dmode == 1 ? 4 : 8);  // bytes read per row and pair
  if (dmode >= 0 && T.wl && T.ubuf && T.pc && pairs >= min_pairs && pairs % min_pairs == 0 && T.k <= 32 &&
      T.ubuf_bytes >= nmono * pairs * sizeof(K2) && digits_per_pair * pairs <= T.m) {
    const int kt = dmode == 0 ? KT_ROUND0 : (dmode == 1 ? KT_ROUND1 : KT_ROUND);
    if (dmode == 2) {
      NS_LAUNCH_TIMED(kt, k_digit_monomials8, dim3((unsigned)(pairs / 16)), dim3(64 * M8_NGROUPS), 0, s, T.digits,
                      T.k, T.m, T.wl, pairs, T.ubuf);
    } else if (dmode == 1) {
      NS_LAUNCH_TIMED(kt, k_digit_monomials<1>, dim3((unsigned)(pairs / 64)), dim3(256), 0, s, T.digits, T.k, T.m,
                      T.wl, pairs, T.ubuf);
    } else {
      NS_LAUNCH_TIMED(kt, k_digit_monomials<0>, dim3((unsigned)(pairs / 128)), dim3(256), 0, s, T.digits, T.k, T.m,
                      T.wl, pairs, T.ubuf);
    }
    size_t nb_fe = (pairs * (T.eqr ? 2 : 1) + RB - 1) / RB;
    if (nb_fe > RMAXBLK / 4) nb_fe = RMAXBLK / 4;
    // E-sum slices of 512 pairs (two per thread), as many as the partial slots allow (powers of two)
    size_t nslice = std::max<size_t>(1, pairs / 512);
    while (nslice > 1 && nmono * nslice + nb_fe > RMAXBLK) nslice /= 2;
    const size_t nb_u = nmono * nslice;
    A.U = T.ubuf;
    A.nb_u = (int)nb_u;
    A.nslice = (int)nslice;
    const size_t shmem = ((size_t)T.k * D * 2 + 80 + 4 * 5 + 40) * sizeof(K2);
    NS_LAUNCH_TIMED(kt, k_round_eval<6>, dim3((unsigned)(nb_u + nb_fe)), dim3(RB), shmem, s, A, wg, partials, 0);
    launch_sum_partials(partials, nb_u + nb_fe, 5, out, sig, s);
    return;
  }
  if (dmode == 2) throw std::logic_error("round 2 from the digits: tables outside the MFMA path's sizes");
  // Split the k*54 NC rows of a pair across G work items: long tables get long items (each item
  // ends with a few reductions and, in round 1, 40 coefficient products), short tables get as
  // many items as it takes to fill the GPU, down to one row per item.
  const size_t nrow = (size_t)T.k * D;
  const size_t target = round0 ? (size_t)RB * 1024 : (size_t)RB * 512;
  size_t G = 1;
  while (G < nrow && pairs * G < target) G *= 2;
  if (G > nrow) G = nrow;
  G = (nrow + (nrow + G - 1) / G - 1) / ((nrow + G - 1) / G);  // no empty groups
  const size_t H = T.digit_round1 ? 2 : 1;
  size_t nblk = (pairs * (G * H + (T.eqr ? 2 : 1)) + RB - 1) / RB;
  if (nblk > RMAXBLK) nblk = RMAXBLK;
  if (nblk < 1) nblk = 1;
  size_t shmem = ((size_t)T.k * D * 2 + 80 + 4 * 5 + 40) * sizeof(K2);
  if (round0)
    NS_LAUNCH_TIMED(KT_ROUND0, k_round_eval<0>, dim3((unsigned)nblk), dim3(RB), shmem, s, A, wg, partials, (int)G);
  else if (T.digit_round1)
    NS_LAUNCH_TIMED(KT_ROUND1, k_round_eval<1>, dim3((unsigned)nblk), dim3(RB), shmem, s, A, wg, partials, (int)G);
  else
    NS_LAUNCH_TIMED(KT_ROUND, k_round_eval<2>, dim3((unsigned)nblk), dim3(RB), shmem, s, A, wg, partials, (int)G);
  launch_sum_partials(partials, nblk, 5, out, sig, s);
}

// ---------------- fused fold + round on the K tables (rounds >= 3) ----------------
// The eq tables are products: after folding the variables before j, eq_beta[2t + x] =
// C_j eq(beta_j, x) g_j(t) with g_j = chi of beta_{j+1..} over the pairs t and C_j the product of
// eq(beta_i, r_i) over the folded variables (likewise eq_{r_in}).  So
//   sum_t eq_beta(t; X) (f_t + NC_t)(X) = C_j eq(beta_j, X) sum_t g_j(t) (f_t + NC_t)(X)
// and the device computes P(X) = sum_t g_j(t) (f_t + NC_t)(X) (degree <= 4) and
// Q(X) = sum_t g'_j(t) Eval_t(X) (degree 1) from the g tables; the host multiplies by the linear
// eq factors (C_j, gamma^k included).  The eq tables are neither read nor folded here.
//
// Work item (pair t, row group g < G) folds its rows' entries 4t..4t+3 of the previous tables by
// the previous challenge r into entries 2t, 2t+1 of the new tables (written for the next round)
// and evaluates its part of P on them; item g = G evaluates f(Mz) and the Eval term from the
// scalar tables (folded in registers), and items g = G + 1 + i fold the i-th f-variable table
// (or the Eval table) for the next round, so no thread waits on more than one round of loads per
// table.  Each block leaves its 7 partial coefficients (P0..P4, Q0, Q1) for k_sum_partials (the
// kernel boundary orders them; a device-scope fence per block would write back L2).
__device__ __forceinline__ K2 fold2(K2 lo, K2 hi, K2 r, uint64_t r7) { return kadd(lo, kmul_w7(r, r7, ksub(hi, lo))); }


// scalar table q of pair t folded at 2t, 2t+1 for the next round
__device__ __forceinline__ void fused_fold_scalar(const FusedArgs& A, size_t t, int q) {
  const size_t L2 = 2 * A.len;
  const K2* src = A.sc_src + (size_t)q * L2 + 4 * t;
  const K2 lo = fold2(src[0], src[1], A.r, A.r7), hi = fold2(src[2], src[3], A.r, A.r7);
  K2* dst = A.sc_dst + (size_t)q * A.len + 2 * t;
  dst[0] = lo;
  dst[1] = hi;
}

// f(Mz)(X) of pair t into in[0..4], on the f-variable tables folded in registers: every load is
// issued at once, and a variable is picked by a register select (a runtime-indexed array would
// live in scratch memory)
__device__ __forceinline__ void fused_f_item(const FusedArgs& A, size_t t, K2 (&in)[5]) {
  const size_t L2 = 2 * A.len;
  const K2 r = A.r;
  const uint64_t r7 = A.r7;
  constexpr int NFV = FUSED_MAX_SC - 1;  // nsc = 1 + nf (+ 2 Eval tables) <= FUSED_MAX_SC
  K2 flo[NFV], fhi[NFV];
#pragma unroll
  for (int q = 0; q < NFV; q++) {
    flo[q] = fhi[q] = kzero();
    if (q < A.nf) {
      const K2* src = A.sc_src + (size_t)(1 + q) * L2 + 4 * t;
      flo[q] = fold2(src[0], src[1], r, r7);
      fhi[q] = fold2(src[2], src[3], r, r7);
    }
  }
#pragma unroll
  for (int d = 0; d < 5; d++) in[d] = kzero();
  for (int tt = 0; tt < A.f.nterms; tt++) {
    // p = coeff * prod (lo + (hi - lo) X)^exp over the term's variables, degree by degree: the
    // first factor scales the F_q coefficient, later ones touch only coefficients up to deg + 1
    // (deg is uniform across the grid, so the guards are scalar branches)
    const uint64_t cf = A.f.coeff[tt];
    K2 p[5];
    p[0] = K2{cf, 0};
#pragma unroll
    for (int d = 1; d < 5; d++) p[d] = kzero();
    int deg = 0;
    bool scalar = true;
    for (int v = 0; v < A.f.nv[tt]; v++) {
      const int q = A.f.var[tt][v];
      K2 lo = flo[0], hi = fhi[0];
#pragma unroll
      for (int qq = 1; qq < NFV; qq++)
        if (q == qq) {
          lo = flo[qq];
          hi = fhi[qq];
        }
      const K2 a = lo, b = ksub(hi, lo);
      for (int e = 0; e < A.f.exp[tt][v]; e++) {
        if (scalar) {
          p[0] = kscale(a, cf);
          p[1] = kscale(b, cf);
          scalar = false;
        } else {
#pragma unroll
          for (int d = 4; d >= 1; d--)
            if (d <= deg + 1) p[d] = d <= deg ? kadd(kmul(a, p[d]), kmul(b, p[d - 1])) : kmul(b, p[d - 1]);
          p[0] = kmul(a, p[0]);
        }
        deg = deg < 4 ? deg + 1 : 4;  // as in the full product, terms past X^4 are dropped
      }
    }
#pragma unroll
    for (int d = 0; d < 5; d++)
      if (d <= deg) in[d] = kadd(in[d], p[d]);
  }
}

// f = c0 x_a x_b + c1 x_c (the R1CS shape) of pair t, straight-line: the twelve table entries are
// loaded at once, then P = c0 (a + a' X)(b + b' X) + c1 (c + c' X) on the folded values
__device__ __forceinline__ void fused_f_r1cs(const FusedArgs& A, size_t t, K2 (&in)[5]) {
  const size_t L2 = 2 * A.len;
  const K2* pa = A.sc_src + (size_t)(1 + A.r1cs[0]) * L2 + 4 * t;
  const K2* pb = A.sc_src + (size_t)(1 + A.r1cs[1]) * L2 + 4 * t;
  const K2* pc = A.sc_src + (size_t)(1 + A.r1cs[2]) * L2 + 4 * t;
  const K2 a0 = pa[0], a1 = pa[1], a2 = pa[2], a3 = pa[3];
  const K2 b0 = pb[0], b1 = pb[1], b2 = pb[2], b3 = pb[3];
  const K2 e0 = pc[0], e1 = pc[1], e2 = pc[2], e3 = pc[3];
  const K2 r = A.r;
  const uint64_t r7 = A.r7;
  const K2 al = fold2(a0, a1, r, r7), ad = ksub(fold2(a2, a3, r, r7), al);
  const K2 bl = fold2(b0, b1, r, r7), bd = ksub(fold2(b2, b3, r, r7), bl);
  const K2 el = fold2(e0, e1, r, r7), ed = ksub(fold2(e2, e3, r, r7), el);
  r1cs_poly(al, ad, bl, bd, el, ed, A.r1cs_c[0], A.r1cs_c[1], in);
}

// NC rows r_lo .. r_hi - 1 of pair t: folded into the next tables, and the four coefficients of
// sum_row w_row N(y_row(X)) written to in[0..3] (lazy accumulators, one reduction per term).
// SCALE: the weights are taken times gt (7 gt.c1 = g7), for items of one or a few rows where
// scaling each weight is cheaper than scaling the four sums.
template <bool SCALE>
__device__ __forceinline__ void fused_nc_rows(const FusedArgs& A, const K2* __restrict__ wg, size_t t, int r_lo,
                                              int r_hi, K2 (&in)[4]) {
  const size_t L2 = 2 * A.len;
  const K2 r = A.r;
  const uint64_t r7 = A.r7;
  KS2 z1 = ks2_zero(), z3 = ks2_zero(), z5 = ks2_zero(), z6 = ks2_zero();
  KS1 z2 = ks1_zero(), z4 = ks1_zero();
  for (int row = r_lo; row < r_hi; row++) {
    const K2* src = A.nc_src + (size_t)row * L2 + 4 * t;
    const K2 u0 = src[0], u1 = src[1], u2 = src[2], u3 = src[3];
    const K2 a = fold2(u0, u1, r, r7), hv = fold2(u2, u3, r, r7);
    K2* dst = A.nc_dst + (size_t)row * A.len + 2 * t;
    dst[0] = a;
    dst[1] = hv;
    const K2 b = ksub(hv, a);
    K2 wv = wg[row];
    if (SCALE) {  // loaded beside the row so that its latency overlaps the row's
      const K2 gt = A.gb[t];
      wv = kmul_w7(gt, fmul7(gt.c1), wv);
    }
    const uint64_t w7 = fmul7(wv.c1);
    const K2 wa = kmul_w7(wv, w7, a), wb = kmul_w7(wv, w7, b);
    const K2 a2 = ksqr_dev(a), b2 = ksqr_dev(b);
    const uint64_t was = ksum(wa), wbs = ksum(wb), a2s = ksum(a2), b2s = ksum(b2);
    ks2k_mac(z1, wa, was, a2, a2s);
    ks1_add(z2, wa);
    ks2k_mac(z3, wb, wbs, a2, a2s);
    ks1_add(z4, wb);
    ks2k_mac(z5, wa, was, b2, b2s);
    ks2k_mac(z6, wb, wbs, b2, b2s);
  }
  const K2 s1 = ks2k_reduce(z1), s2 = ks1_reduce(z2), s3 = ks2k_reduce(z3), s4 = ks1_reduce(z4),
           s5 = ks2k_reduce(z5), s6 = ks2k_reduce(z6);
  in[0] = ksub(s1, s2);
  in[1] = ksub(kadd(kadd(s3, s3), s3), s4);
  in[2] = kadd(kadd(s5, s5), s5);
  in[3] = s6;
}

// x / 2 in F_q per component (x odd: (x + q) / 2)
__device__ __forceinline__ K2 khalf(K2 x) {
  return K2{(x.c0 >> 1) + ((x.c0 & 1) ? (Q + 1) / 2 : 0), (x.c1 >> 1) + ((x.c1 & 1) ? (Q + 1) / 2 : 0)};
}

// Short rounds (few pairs): one NC row of pair t split four ways by evaluation point, so that a
// lane's chain is one fold, one square and two products.  Lane s evaluates w g N(y(X)) at X = 0
// (s = 0), 1 (s = 1), -1 (s = 3) and takes the leading coefficient w g b^3 (s = 2), all from the
// folded line y = a + b X (b, y(-1) = 2a - y(1) fold linearly from the previous entries); with
// V0, V1, V-1, Vinf those values the cubic's coefficients are P0 = V0, P3 = Vinf,
// P1 = (V1 - V-1)/2 - Vinf, P2 = (V1 + V-1)/2 - V0, and each lane adds its share.  Lanes 0 and 1
// write the folded entries for the next round.
__device__ __forceinline__ void fused_nc_split(const FusedArgs& A, const K2* __restrict__ wg, size_t t, int row, int s,
                                               K2 (&c)[FUSED_NCOEF]) {
  const K2* src = A.nc_src + (size_t)row * (2 * A.len) + 4 * t;
  const K2 u0 = src[0], u1 = src[1], u2 = src[2], u3 = src[3];
  const K2 gt = A.gb[t], wv = wg[row];
  const K2 lo = s == 0 ? u0 : (s == 1 ? u2 : (s == 2 ? ksub(u2, u0) : ksub(kadd(u0, u0), u2)));
  const K2 hi = s == 0 ? u1 : (s == 1 ? u3 : (s == 2 ? ksub(u3, u1) : ksub(kadd(u1, u1), u3)));
  const K2 w = kmul_w7(gt, fmul7(gt.c1), wv);
  const K2 y = fold2(lo, hi, A.r, A.r7);
  if (s < 2) A.nc_dst[(size_t)row * A.len + 2 * t + s] = y;
  const K2 y2 = ksqr_dev(y);
  const K2 wy = kmul_w7(w, fmul7(w.c1), y);
  const K2 v = kmul(wy, s == 2 ? y2 : ksub(y2, kone()));
  const K2 hv = khalf(v), nv = kneg(v), nh = kneg(hv);
  c[0] = kadd(c[0], s == 0 ? v : kzero());
  c[1] = kadd(c[1], s == 1 ? hv : (s == 2 ? nv : (s == 3 ? nh : kzero())));
  c[2] = kadd(c[2], s == 0 ? nv : (s == 2 ? kzero() : hv));
  c[3] = kadd(c[3], s == 2 ? v : kzero());
}

// The R1CS f item of a short round split three ways: lane s evaluates g P at X = 0 (s = 0), 1
// (s = 1) and takes the leading coefficient g c0 a' b' (s = 2); P1 = P(1) - P0 - P2.
__device__ __forceinline__ void fused_f_split(const FusedArgs& A, size_t t, int s, K2 (&c)[FUSED_NCOEF]) {
  const size_t L2 = 2 * A.len;
  const K2* pa = A.sc_src + (size_t)(1 + A.r1cs[0]) * L2 + 4 * t;
  const K2* pb = A.sc_src + (size_t)(1 + A.r1cs[1]) * L2 + 4 * t;
  const K2* pc = A.sc_src + (size_t)(1 + A.r1cs[2]) * L2 + 4 * t;
  const K2 a0 = pa[0], a1 = pa[1], a2 = pa[2], a3 = pa[3];
  const K2 b0 = pb[0], b1 = pb[1], b2 = pb[2], b3 = pb[3];
  const K2 e0 = pc[0], e1 = pc[1], e2 = pc[2], e3 = pc[3];
  const K2 gt = A.gb[t];
  const K2 alo = s == 0 ? a0 : (s == 1 ? a2 : ksub(a2, a0)), ahi = s == 0 ? a1 : (s == 1 ? a3 : ksub(a3, a1));
  const K2 blo = s == 0 ? b0 : (s == 1 ? b2 : ksub(b2, b0)), bhi = s == 0 ? b1 : (s == 1 ? b3 : ksub(b3, b1));
  const K2 elo = s == 0 ? e0 : e2, ehi = s == 0 ? e1 : e3;
  const K2 ya = fold2(alo, ahi, A.r, A.r7), yb = fold2(blo, bhi, A.r, A.r7), ye = fold2(elo, ehi, A.r, A.r7);
  K2 p = kmul_w7(ya, fmul7(ya.c1), yb);
  const uint64_t c0 = A.r1cs_c[0], c1 = A.r1cs_c[1];
  if (c0 != 1) p = kscale(p, c0);
  if (s < 2) p = c1 == Q - 1 ? ksub(p, ye) : kadd(p, kscale(ye, c1));
  const K2 v = kmul_w7(gt, fmul7(gt.c1), p);
  c[0] = kadd(c[0], s == 0 ? v : kzero());
  c[1] = kadd(c[1], s == 1 ? v : kneg(v));
  c[2] = kadd(c[2], s == 2 ? v : kzero());
}

// One block's share of a fused round: block bid of nb blocks over the work items of A.  Shared by
// the single-session launch and the ragged launch over several sessions' rounds.
__device__ __forceinline__ void fused_block(const FusedArgs& A, const K2* __restrict__ wg, K2* __restrict__ partials,
                                           int G, unsigned bid, unsigned nb, K2* red) {
  const int nrow = A.k * D;
  const size_t pairs = A.len / 2;
  const int R = (nrow + G - 1) / G;
  const int nfold = A.nf + (A.has_eval ? 1 : 0);
  K2 c[FUSED_NCOEF];
#pragma unroll
  for (int d = 0; d < FUSED_NCOEF; d++) c[d] = kzero();
  // NC items first, then (from a wave boundary, so that no wave runs both kinds) the f item and
  // the Eval item of every pair and the scalar fold items.  Short rounds (A.split) take one row
  // per NC item split four ways and the R1CS f item split three ways.
  const int ne = A.has_eval ? 1 : 0;
  const int spl = A.split ? 4 : 1, fs = A.split && A.r1cs[0] >= 0 ? 3 : 1;
  const size_t nnc = pairs * (size_t)G * spl, base_s = (nnc + 63) & ~(size_t)63;
  int has_s = 0;
  const size_t total = base_s + pairs * (fs + ne + (size_t)nfold);
  const int fdeg = A.fdeg;  // degree of f (uniform): coefficients above it are zero
  const int lp = __builtin_ctzll(pairs);  // pairs is a power of two (tables are 2^ell_n long)
  for (size_t idx = (size_t)bid * blockDim.x + threadIdx.x; idx < total; idx += (size_t)nb * blockDim.x) {
    if (idx >= nnc && idx < base_s) continue;
    const bool nc_item = idx < nnc;
    has_s |= !nc_item;
    const size_t li = nc_item ? idx : idx - base_s;
    const size_t t = li & (pairs - 1);
    const size_t hi_ = li >> lp;
    if (nc_item && A.split) {  // one row, sub-item s
      fused_nc_split(A, wg, t, (int)(hi_ >> 2), (int)(hi_ & 3), c);
      continue;
    }
    if (!nc_item && fs == 3 && hi_ < 3) {
      fused_f_split(A, t, (int)hi_, c);
      continue;
    }
    // g = G: the f item, G + 1: the Eval item (with r_in), then the fold items
    const int g = nc_item ? (int)hi_ : G + (int)hi_ - (fs - 1);
    if (g > G + ne) {  // fold items: f-variable tables 1..nf, then the Eval table 2 + nf
      const int i = g - G - 1 - ne;
      fused_fold_scalar(A, t, i < A.nf ? 1 + i : 2 + A.nf);
      continue;
    }
    if (g == G) {
      const K2 gt = A.gb[t];
      K2 in[5];
      if (A.r1cs[0] >= 0)
        fused_f_r1cs(A, t, in);
      else
        fused_f_item(A, t, in);
      const uint64_t g7 = fmul7(gt.c1);
#pragma unroll
      for (int d = 0; d < 5; d++)
        if (d <= fdeg) c[d] = kadd(c[d], kmul_w7(gt, g7, in[d]));
      continue;
    }
    if (g == G + 1) {  // Eval item: Q += g'(t) (vl + (vh - vl) X) on the folded Eval table
      const size_t L2 = 2 * A.len;
      const K2* sv = A.sc_src + (size_t)(2 + A.nf) * L2 + 4 * t;
      const K2 gr = A.gr[t];
      const K2 vl = fold2(sv[0], sv[1], A.r, A.r7), vh = fold2(sv[2], sv[3], A.r, A.r7);
      const uint64_t gr7 = fmul7(gr.c1);
      c[5] = kadd(c[5], kmul_w7(gr, gr7, vl));
      c[6] = kadd(c[6], kmul_w7(gr, gr7, ksub(vh, vl)));
      continue;
    }
    const int r_lo = min(nrow, g * R), r_hi = min(nrow, r_lo + R);
    K2 in[4];
    if (R <= 3) {
      fused_nc_rows<true>(A, wg, t, r_lo, r_hi, in);
#pragma unroll
      for (int d = 0; d < 4; d++) c[d] = kadd(c[d], in[d]);
    } else {
      const K2 gt = A.gb[t];
      fused_nc_rows<false>(A, wg, t, r_lo, r_hi, in);
      const uint64_t g7 = fmul7(gt.c1);
#pragma unroll
      for (int d = 0; d < 4; d++) c[d] = kadd(c[d], kmul_w7(gt, g7, in[d]));
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // blocks of NC items only (most of them) leave P4, Q0 and Q1 at zero
  const bool any_s = __syncthreads_or(has_s);
#pragma unroll
  for (int d = 0; d < FUSED_NCOEF; d++) {
    if (d >= 4 && !any_s) {
      if (lane == DPP_LANE) red[wave * FUSED_NCOEF + d] = kzero();
      continue;
    }
    K2 v = wave_sum(c[d]);
    if (lane == DPP_LANE) red[wave * FUSED_NCOEF + d] = v;
  }
  __syncthreads();
  if (threadIdx.x < FUSED_NCOEF) {
    K2 v = kzero();
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) v = kadd(v, red[w * FUSED_NCOEF + threadIdx.x]);
    partials[bid * FUSED_NCOEF + threadIdx.x] = v;
  }
}

__global__ __launch_bounds__(RB) void k_round_fused(FusedArgs A, const K2* __restrict__ wg, K2* __restrict__ partials,
                                                     int G) {
  // row weights come from global memory (11 KB, cache-resident): a short round's block needs a
  // few of them, so staging all k*54 in LDS would cost more than it saves
  __shared__ K2 red[4 * FUSED_NCOEF];
  fused_block(A, wg, partials, G, blockIdx.x, gridDim.x, red);
}

// Several sessions' fused rounds in one launch: job j owns blocks [off_j, off_j + nblk_j) in the
// order of the array (at most FUSED_MAX_JOBS jobs, found by a scan of the block counts).
__global__ __launch_bounds__(RB) void k_round_fused_multi(const FusedJob* __restrict__ jobs, int njobs) {
  __shared__ K2 red[4 * FUSED_NCOEF];
  unsigned b = blockIdx.x;
  int j = 0;
  while (j + 1 < njobs && b >= (unsigned)jobs[j].nblk) b -= (unsigned)jobs[j++].nblk;
  const FusedJob& J = jobs[j];
  fused_block(J.A, J.wg, J.partials, J.G, b, (unsigned)J.nblk, red);
}

static void classify_f(FusedArgs& A) { classify_fpoly(A.f, A.nf, A.fdeg, A.r1cs, A.r1cs_c); }

FusedJob fused_job(const FusedArgs& A, const K2* wg, K2* out, void* ws, const RoundSignal& sig) {
  const size_t pairs = A.len / 2, nrow = (size_t)A.k * D;
  if (pairs == 0 || (pairs & (pairs - 1))) throw std::invalid_argument("fused round: table length must be a power of two >= 2");
  if (A.nf + 1 > FUSED_MAX_SC - (A.has_eval ? 2 : 0)) throw std::invalid_argument("fused round: too many scalar tables");
  if (!A.gb || (A.has_eval && !A.gr)) throw std::invalid_argument("fused round: g tables missing");
  FusedJob J;
  J.A = A;
  classify_f(J.A);
  // rounds of at most 32 pairs leave the GPU mostly idle and wait on one work item's chain of
  // dependent products: they take one row per NC item, split four ways (and the R1CS f item
  // three ways)
  J.A.split = pairs <= 32 ? 1 : 0;
  const size_t target = (size_t)RB * 512;
  size_t G = 1;
  while (G < nrow && pairs * G < target) G *= 2;
  if (G > nrow || J.A.split) G = nrow;
  G = (nrow + (nrow + G - 1) / G - 1) / ((nrow + G - 1) / G);  // no empty groups
  const size_t nfold = (size_t)A.nf + (A.has_eval ? 1 : 0);
  const size_t spl = J.A.split ? 4 : 1, fs = J.A.split && J.A.r1cs[0] >= 0 ? 3 : 1;
  const size_t base_s = (pairs * G * spl + 63) & ~(size_t)63;
  size_t nblk = (base_s + pairs * (fs + (A.has_eval ? 1 : 0) + nfold) + RB - 1) / RB;
  if (nblk > RMAXBLK) nblk = RMAXBLK;
  if (nblk < 1) nblk = 1;
  J.wg = wg;
  J.partials = (K2*)ws;
  J.out = out;
  J.flag = sig.flag;
  J.seq = sig.seq;
  J.G = (int)G;
  J.nblk = (int)nblk;
  return J;
}

void launch_fused_job(const FusedJob& J, hipStream_t s) {
  NS_LAUNCH_TIMED(KT_FUSED, k_round_fused, dim3((unsigned)J.nblk), dim3(RB), 0, s, J.A, J.wg, J.partials, J.G);
  RoundSignal sig;
  sig.flag = J.flag;
  sig.seq = J.seq;
  launch_sum_partials(J.partials, (size_t)J.nblk, FUSED_NCOEF, J.out, J.flag ? &sig : nullptr, s);
}

void launch_fused_jobs(const FusedJob* jobs_dev, int njobs, size_t total_blocks, hipStream_t s) {
  if (njobs < 1 || njobs > FUSED_MAX_JOBS) throw std::invalid_argument("fused jobs: 1..FUSED_MAX_JOBS per launch");
  hipLaunchKernelGGL(k_round_fused_multi, dim3((unsigned)total_blocks), dim3(RB), 0, s, jobs_dev, njobs);
  hipLaunchKernelGGL(k_sum_partials_multi, dim3((unsigned)njobs), dim3(512), 0, s, jobs_dev);
}

void launch_round_fused(const FusedArgs& A, const K2* wg, K2* out, void* ws, hipStream_t s,
                        const RoundSignal* sig) {
  RoundSignal none;
  launch_fused_job(fused_job(A, wg, out, ws, sig ? *sig : none), s);
}

// ---------------- folds ----------------
__global__ void k_fold_k(const K2* __restrict__ src, size_t len, K2* __restrict__ dst, K2 r) {
  size_t half = len / 2;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t row = blockIdx.y;
  if (i >= half) return;
  const K2* s = src + row * len;
  K2 lo = s[2 * i], hi = s[2 * i + 1];
  dst[row * half + i] = kadd(lo, kmul(ksub(hi, lo), r));
}

void launch_fold_k(const K2* src, size_t rows, size_t len, K2* dst, K2 r, hipStream_t s) {
  size_t half = len / 2;
  if (rows == 0 || half == 0) return;
  NS_LAUNCH_TIMED(KT_FOLD, k_fold_k, dim3((unsigned)((half + 255) / 256), (unsigned)rows), dim3(256), 0, s, src, len,
                  dst, r);
}

__global__ void k_fold_digits(const int8_t* const* __restrict__ digits, size_t m, size_t len, K2* __restrict__ dst,
                              K2 r) {
  size_t half = len / 2;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t row = blockIdx.y;  // w*54 + rho
  if (i >= half) return;
  const int8_t* z = digits[row / D] + (row % D) * m;
  // rows beyond the witness width (m < len) are zero (RowStreamState::build pads to 2^ell_n)
  int lo = (2 * i < m) ? z[2 * i] : 0;
  int hi = (2 * i + 1 < m) ? z[2 * i + 1] : 0;
  K2 v = K2{from_i64(lo), 0};
  v = kadd(v, kmul_small(r, hi - lo));
  dst[row * half + i] = v;
}

// Tables after rounds 0 and 1 straight from the digits: T2[i] = sum_b chi_{(r0, r1)}(b) D[4i + b].
__global__ void k_fold_digits2(const int8_t* const* __restrict__ digits, size_t m, size_t len, K2* __restrict__ dst,
                               K2 w00, K2 w10, K2 w01, K2 w11) {
  size_t quarter = len / 4;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t row = blockIdx.y;
  if (i >= quarter) return;
  const int8_t* z = digits[row / D] + (row % D) * m;
  int d[4];
  if ((m & 3) == 0 && 4 * i + 3 < m) {  // rows start 4-byte aligned: one load for the four digits
    const uint32_t w4 = *reinterpret_cast<const uint32_t*>(z + 4 * i);
#pragma unroll
    for (int b = 0; b < 4; b++) d[b] = (int)(int8_t)(w4 >> (8 * b));
  } else {
#pragma unroll
    for (int b = 0; b < 4; b++) d[b] = (4 * i + b < m) ? z[4 * i + b] : 0;
  }
  const K2 w[4] = {w00, w10, w01, w11};
  K2 v = kzero();
#pragma unroll
  for (int b = 0; b < 4; b++) v = kadd(v, d[b] > 0 ? w[b] : (d[b] < 0 ? kneg(w[b]) : kzero()));
  dst[row * quarter + i] = v;
}

void launch_fold_digits2(const int8_t* const* digits_dev, int k, size_t m, size_t len, K2* dst, K2 r0, K2 r1,
                         hipStream_t s) {
  size_t quarter = len / 4;
  const K2 a0 = ksub(kone(), r0), a1 = ksub(kone(), r1);
  NS_LAUNCH_TIMED(KT_FOLD, k_fold_digits2, dim3((unsigned)((quarter + 255) / 256), (unsigned)(k * D)), dim3(256), 0,
                  s, digits_dev, m, len, dst, kmul(a0, a1), kmul(r0, a1), kmul(a0, r1), kmul(r0, r1));
}

// Tables after rounds 0-2 from the digits: 8 digits per entry, weights chi_(r0, r1, r2)(b) for
// b = bit0 r0 + bit1 r1 + bit2 r2 order (digit 8i + b), signed selections summed lazily.
struct Chi8 {
  K2 w[8];
};
__global__ void k_fold_digits3(const int8_t* const* __restrict__ digits, size_t m, size_t eighth, K2* __restrict__ dst,
                               Chi8 wts) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t row = blockIdx.y;
  if (i >= eighth) return;
  const uint2 w8 = *reinterpret_cast<const uint2*>(digits[row / D] + (row % D) * m + 8 * i);
  KS1 acc = ks1_zero();
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const int d = (int)(int8_t)((b < 4 ? w8.x : w8.y) >> (8 * (b & 3)));
    const K2 w = wts.w[b];
    ks1_add(acc, d > 0 ? w : (d < 0 ? kneg(w) : kzero()));
  }
  dst[row * eighth + i] = ks1_reduce(acc);
}

bool digit_round2_fits(size_t len, int k) {
  const size_t pairs = len / 2;
  return pairs >= 64 && pairs % 64 == 0 && k <= 32;
}

void launch_fold_digits3(const int8_t* const* digits_dev, int k, size_t m, size_t len, K2* dst, K2 r0, K2 r1, K2 r2,
                         hipStream_t s) {
  if (len % 8 != 0 || m % 8 != 0 || len > m) throw std::invalid_argument("fold_digits3: len and m multiples of 8, len <= m");
  const size_t eighth = len / 8;
  const K2 a0 = ksub(kone(), r0), a1 = ksub(kone(), r1), a2 = ksub(kone(), r2);
  Chi8 w;
  for (int b = 0; b < 8; b++) w.w[b] = kmul(kmul(b & 1 ? r0 : a0, b & 2 ? r1 : a1), b & 4 ? r2 : a2);
  NS_LAUNCH_TIMED(KT_FOLD, k_fold_digits3, dim3((unsigned)((eighth + 255) / 256), (unsigned)(k * D)), dim3(256), 0, s,
                  digits_dev, m, eighth, dst, w);
}

void launch_fold_digits(const int8_t* const* digits_dev, int k, size_t m, size_t len, K2* dst, K2 r, hipStream_t s) {
  size_t half = len / 2;
  hipLaunchKernelGGL(k_fold_digits, dim3((unsigned)((half + 255) / 256), (unsigned)(k * D)), dim3(256), 0, s,
                     digits_dev, m, len, dst, r);
}

// ---------------- Pi_RLC / Pi_DEC ----------------
__global__ __launch_bounds__(256) void k_rlc_zmix(const int8_t* __restrict__ rho, const int8_t* const* __restrict__ Zs,
                                                  int k, size_t m, int32_t* __restrict__ zmix) {
  extern __shared__ int8_t rho_s[];  // [i][kk][r] from the padded [i][64][64] rows
  for (int idx = threadIdx.x; idx < k * D * D; idx += blockDim.x) {
    int i = idx / (D * D), rem = idx % (D * D), kk = rem / D, r = rem % D;
    rho_s[idx] = rho[((size_t)i * RC_AP + r) * RC_AP + kk];
  }
  __syncthreads();
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  int32_t acc[D];
#pragma unroll
  for (int r = 0; r < D; r++) acc[r] = 0;
  for (int i = 0; i < k; i++) {
    const int8_t* z = Zs[i] + c;
    int8_t dz[D];
#pragma unroll
    for (int kk = 0; kk < D; kk++) dz[kk] = z[(size_t)kk * m];
    for (int kk = 0; kk < D; kk++) {
      int d = dz[kk];
      if (d == 0) continue;
      const int8_t* col = rho_s + (i * D + kk) * D;
#pragma unroll
      for (int r = 0; r < D; r++) acc[r] += d * (int)col[r];
    }
  }
#pragma unroll
  for (int r = 0; r < D; r++) zmix[(size_t)r * m + c] = acc[r];
}

void launch_rlc_zmix(const int8_t* rho_dev, const int8_t* const* Zs_dev, size_t k, size_t m, int32_t* zmix,
                     hipStream_t s) {
  if (m % RC_NB == 0) {
    NS_LAUNCH_TIMED(KT_ZMIX, k_rowcontract<4>, dim3((unsigned)(m / RC_NB)), dim3(256), 0, s, rho_dev, (int)D, Zs_dev,
                    (int)k, m, zmix);
    return;
  }
  size_t shmem = k * D * D;
  NS_LAUNCH_TIMED(KT_ZMIX, k_rlc_zmix, dim3((unsigned)((m + 255) / 256)), dim3(256), shmem, s, rho_dev, Zs_dev, (int)k,
                  m, zmix);
}

// Balanced base-2 digits (neo-reductions/src/common.rs:75-268): repeatedly r = v % 2 (the sign
// of v when v is odd), v = (v - r) / 2.  Both steps keep the sign of v and halve |v|, so digit i
// is sign(v) * bit i of |v|, and v is in range exactly when |v| < 2^k_dec.  Each thread takes two
// runs of 16 consecutive entries (four 16-byte loads each) and writes one 16-byte word per plane
// and run (two runs per thread: 65% of HBM peak at 2^16 x 54 entries, against 61% with one); plane i is
// non-zero exactly when bit i of the OR of all |v| is set, so one OR word per block carries every
// flag.  vec = 0 (an unaligned pointer from a caller, or len not a multiple of 16) takes byte
// stores instead.
constexpr int DEC_ITER = 2;
__global__ __launch_bounds__(256) void k_dec_split(const int32_t* __restrict__ zmix, size_t len, int k_dec,
                                                   DecPlanes outs, int* __restrict__ flags, int vec) {
  uint32_t orv = 0;
  // DEC_ITER runs of 16 entries per thread, a block's runs one block-width apart
  for (int it = 0; it < DEC_ITER; it++) {
    const size_t base = (((size_t)blockIdx.x * DEC_ITER + it) * blockDim.x + threadIdx.x) * 16;
    if (base >= len) break;
    const bool full = vec && base + 16 <= len;
    uint32_t a[16], negw[4] = {0, 0, 0, 0};
    if (full) {
      const int4* src = reinterpret_cast<const int4*>(zmix + base);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int4 x = src[q];
        const int32_t v4[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int32_t v = v4[j];
          a[4 * q + j] = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
          negw[q] |= (v < 0 ? 0xFFu : 0u) << (8 * j);
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int32_t v = base + e < len ? zmix[base + e] : 0;
        a[e] = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
        negw[e >> 2] |= (v < 0 ? 0xFFu : 0u) << (8 * (e & 3));
      }
    }
#pragma unroll
    for (int e = 0; e < 16; e++) orv |= a[e];
    for (int i = 0; i < k_dec; i++) {
      uint32_t w[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        // one byte per entry: 0x01 for a set bit, widened to 0xFF where the entry is negative
        const uint32_t bits = ((a[4 * q] >> i) & 1u) | (((a[4 * q + 1] >> i) & 1u) << 8) |
                              (((a[4 * q + 2] >> i) & 1u) << 16) | (((a[4 * q + 3] >> i) & 1u) << 24);
        w[q] = bits | ((bits * 0xFFu) & negw[q]);
      }
      int8_t* dst = outs.p[i] + base;
      if (full) {
        *reinterpret_cast<uint4*>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
      } else {
        for (int e = 0; e < 16 && base + e < len; e++) dst[e] = (int8_t)(w[e >> 2] >> (8 * (e & 3)));
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) orv |= __shfl_down(orv, off, 64);
  __shared__ uint32_t red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = orv;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t o = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) o |= red[w];
    // one word for every plane: most blocks find their bits already set and skip the contended
    // atomic after a single load (one load per plane made each block's tail a dozen L2 round trips)
    uint32_t* mask = reinterpret_cast<uint32_t*>(flags);
    if (o & ~__atomic_load_n(mask, __ATOMIC_RELAXED)) atomicOr(mask, o);
  }
}

__global__ void k_fill_i32(int* __restrict__ p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_dec_split(const int32_t* zmix, size_t len, int k_dec, const DecPlanes& outs, int* flags, hipStream_t s) {
  if (k_dec < 1 || k_dec > DEC_MAX_PLANES) throw std::invalid_argument("DEC split: 1..30 planes");
  // flags start at zero: a fill kernel, not a memset node (callers may capture this sequence)
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, flags, 1, 0);
  if (len == 0) return;
  bool aligned = (((uintptr_t)zmix) & 15) == 0 && (len & 15) == 0;
  for (int i = 0; i < k_dec; i++) aligned = aligned && (((uintptr_t)outs.p[i]) & 15) == 0;
  const size_t threads = (len + 15) / 16;
  const unsigned blocks = (unsigned)((threads + 256 * DEC_ITER - 1) / (256 * DEC_ITER));
  NS_LAUNCH_TIMED(KT_SPLIT, k_dec_split, dim3(blocks), dim3(256), 0, s, zmix, len, k_dec, outs, flags, aligned ? 1 : 0);
}

// Static scratch of the no-allocation DEC entry point: the split's OR-of-entries word (one int used
// of 64).  Each GPU holds its own copy for the life of the process; every call on one GPU shares
// it, so calls that overlap on one GPU race.
static __device__ __align__(256) int scratch[64];

// This device's address of `scratch` (the host-side name is only a stand-in, never a device
// address), looked up once per device with that device current and kept; later calls reuse it.
int* dec_noalloc_flags() {
  static std::mutex mu;
  static int* per_dev[256] = {nullptr};
  int dev = 0;
  NS_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 256) throw std::invalid_argument("device ordinal out of range");
  std::lock_guard<std::mutex> g(mu);
  if (!per_dev[dev]) {
    void* ptr = nullptr;
    NS_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    per_dev[dev] = (int*)ptr;
  }
  return per_dev[dev];
}
static_assert(DEC_MAX_PLANES < 31, "the DEC split's OR word keeps a range bit above every plane");

// ---------------- head columns (ME public inputs X) ----------------
__global__ void k_head_cols(PlanePtrs planes, int np, size_t m, int m_in, int8_t* __restrict__ out) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= np * D * m_in) return;
  int p = idx / (D * m_in), r = (idx / m_in) % D, c = idx % m_in;
  out[idx] = planes.p[p][(size_t)r * m + c];
}

void launch_head_cols(const PlanePtrs& planes, int np, size_t m, int m_in, int8_t* out, hipStream_t s) {
  int n = np * D * m_in;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_head_cols, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, planes, np, m, m_in, out);
}

}  // namespace ns
