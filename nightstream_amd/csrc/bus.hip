// bus.hip -- shared-CPU-bus kernels of the shard prover: the bus region of a step's witness
// decoded from its resident digit planes, and the bus openings appended to ME instances
// (append_bus_openings_to_me_instance, neo-fold/src/memory_sidecar/cpu_bus.rs:195-320) for
// several digit planes, or for an int32 matrix (Z_mix), in one launch.
//
// Bus cell (col, j) is z[bus_base + col * chunk + j] and lives at time index m_in + j; the
// opening of column col at the point r is y[col][rho] = sum_j chi_r(m_in + j) Z[rho][cell(col, j)]
// and y_scalar[col] = sum_rho 2^rho y[col][rho].
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ns {

// w[j] = chi_r(m_in + j): the multilinear basis at the time index, bits little-endian
__global__ void k_bus_weights(uint32_t m_in, uint32_t chunk, const K2* __restrict__ r, int ell, K2* __restrict__ w) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= chunk) return;
  const uint32_t t = m_in + j;
  K2 acc = kone();
  for (int b = 0; b < ell; b++) acc = kmul(acc, ((t >> b) & 1) ? r[b] : ksub(kone(), r[b]));
  w[j] = acc;
}

// balanced digits of column c recomposed: sum_rho 2^rho Z[rho][c] (decode_vector)
__global__ void k_bus_decode_range(const int8_t* __restrict__ Z, size_t m, size_t base, size_t len,
                                   uint64_t* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  const size_t c = base + i;
  uint64_t acc = 0, pw = 1;
  for (int rho = 0; rho < D; rho++) {
    const int8_t d = Z[(size_t)rho * m + c];
    if (d > 0) acc = fadd(acc, pw);
    else if (d < 0) acc = fsub(acc, pw);
    pw = fadd(pw, pw);
  }
  out[i] = acc;
}

__device__ __forceinline__ K2 bus_wave_sum(K2 v) {
  for (int off = 32; off > 0; off >>= 1) {
    K2 o;
    o.c0 = __shfl_down(v.c0, off, 64);
    o.c1 = __shfl_down(v.c1, off, 64);
    v = kadd(v, o);
  }
  return v;
}

__device__ __forceinline__ K2 block_sum256(K2 acc, K2* red) {
  acc = bus_wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  return kadd(kadd(red[0], red[1]), kadd(red[2], red[3]));
}

// block (col, rho, plane): y[(plane * bus_cols + col) * 64 + rho]
__global__ __launch_bounds__(256) void k_bus_open_planes(PlanePtrs planes, size_t m, size_t bus_base, uint32_t chunk,
                                                         uint32_t bus_cols, const K2* __restrict__ w,
                                                         K2* __restrict__ y) {
  const uint32_t col = blockIdx.x, rho = blockIdx.y, p = blockIdx.z;
  const int8_t* row = planes.p[p] + (size_t)rho * m + bus_base + (size_t)col * chunk;
  K2 acc = kzero();
  for (uint32_t j = threadIdx.x; j < chunk; j += blockDim.x) {
    const int8_t d = row[j];
    if (d > 0) acc = kadd(acc, w[j]);
    else if (d < 0) acc = ksub(acc, w[j]);
  }
  __shared__ K2 red[4];
  acc = block_sum256(acc, red);
  if (threadIdx.x == 0) y[((size_t)p * bus_cols + col) * 64 + rho] = acc;
}

// the same for one int32 matrix (entries scale the weights as field elements)
__global__ __launch_bounds__(256) void k_bus_open_i32(const int32_t* __restrict__ Z, size_t m, size_t bus_base,
                                                      uint32_t chunk, const K2* __restrict__ w, K2* __restrict__ y) {
  const uint32_t col = blockIdx.x, rho = blockIdx.y;
  const int32_t* row = Z + (size_t)rho * m + bus_base + (size_t)col * chunk;
  K2 acc = kzero();
  for (uint32_t j = threadIdx.x; j < chunk; j += blockDim.x) {
    const int32_t d = row[j];
    if (d != 0) {
      const uint64_t s = d > 0 ? (uint64_t)d : Q - (uint64_t)(-(int64_t)d);
      acc = kadd(acc, kscale(w[j], s));
    }
  }
  __shared__ K2 red[4];
  acc = block_sum256(acc, red);
  if (threadIdx.x == 0) y[(size_t)col * 64 + rho] = acc;
}

// y_scalar of every opened column; the padded rows 54..63 are zeroed
__global__ void k_bus_open_scalars(K2* __restrict__ y, uint32_t ncols, K2* __restrict__ ys) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  K2 s = kzero();
  uint64_t pw = 1;
  for (int rho = 0; rho < D; rho++) {
    s = kadd(s, kscale(y[(size_t)c * 64 + rho], pw));
    pw = fadd(pw, pw);
  }
  for (int rho = D; rho < 64; rho++) y[(size_t)c * 64 + rho] = kzero();
  ys[c] = s;
}

void launch_bus_weights(uint32_t m_in, uint32_t chunk, const K2* r_dev, int ell, K2* w, hipStream_t s) {
  hipLaunchKernelGGL(k_bus_weights, dim3((chunk + 255) / 256), dim3(256), 0, s, m_in, chunk, r_dev, ell, w);
}

void launch_bus_decode_range(const int8_t* Z, size_t m, size_t base, size_t len, uint64_t* out, hipStream_t s) {
  if (!len) return;
  hipLaunchKernelGGL(k_bus_decode_range, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, Z, m, base, len, out);
}

void launch_bus_open_planes(const PlanePtrs& planes, int np, size_t m, size_t bus_base, uint32_t chunk,
                            uint32_t bus_cols, const K2* w, K2* y, K2* ys, hipStream_t s) {
  if (!np || !bus_cols) return;
  hipLaunchKernelGGL(k_bus_open_planes, dim3(bus_cols, D, (unsigned)np), dim3(256), 0, s, planes, m, bus_base, chunk,
                     bus_cols, w, y);
  const uint32_t nc = (uint32_t)np * bus_cols;
  hipLaunchKernelGGL(k_bus_open_scalars, dim3((nc + 63) / 64), dim3(64), 0, s, y, nc, ys);
}

void launch_bus_open_i32(const int32_t* Z, size_t m, size_t bus_base, uint32_t chunk, uint32_t bus_cols, const K2* w,
                         K2* y, K2* ys, hipStream_t s) {
  if (!bus_cols) return;
  hipLaunchKernelGGL(k_bus_open_i32, dim3(bus_cols, D), dim3(256), 0, s, Z, m, bus_base, chunk, w, y);
  hipLaunchKernelGGL(k_bus_open_scalars, dim3((bus_cols + 63) / 64), dim3(64), 0, s, y, bus_cols, ys);
}

}  // namespace ns
