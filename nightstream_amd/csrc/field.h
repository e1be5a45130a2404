// field.h -- Goldilocks F_q and K = F_q[u]/(u^2 - 7) for host and gfx950 device code.
//
// F_q: q = 2^64 - 2^32 + 1, canonical representatives in [0, q).  Reduction of a 128-bit
// product uses 2^64 = 2^32 - 1 and 2^96 = -1 (mod q).  K follows neo-math/src/field.rs:10-16
// (BinomialExtensionField<Goldilocks, 2>, u^2 = W = 7).  Everything is exact, so any
// evaluation order gives the same canonical result.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NS_HD __host__ __device__ __forceinline__

namespace ns {

constexpr uint64_t Q = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod q

NS_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Reduce lo + 2^64 hi (any 128-bit value) to [0, q).  Branch-free: carries and borrows are
// turned into masks so random data does not cost mispredicted branches (host) or divergence
// (device).
NS_HD uint64_t reduce128(uint64_t lo, uint64_t hi) {
  uint64_t hi_hi = hi >> 32, hi_lo = hi & EPS;
  uint64_t t0 = lo - hi_hi;
  t0 -= (uint64_t)0 - (uint64_t)(lo < hi_hi) & EPS;
  uint64_t t1 = (hi_lo << 32) - hi_lo;
  uint64_t t2 = t0 + t1;
  t2 += (uint64_t)0 - (uint64_t)(t2 < t1) & EPS;
  return t2 >= Q ? t2 - Q : t2;
}

NS_HD uint64_t fadd(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  s += (uint64_t)0 - (uint64_t)(s < a) & EPS;
  return s >= Q ? s - Q : s;
}
NS_HD uint64_t fsub(uint64_t a, uint64_t b) {
  uint64_t d = a - b;
  d -= (uint64_t)0 - (uint64_t)(a < b) & EPS;  // d + q == d - (2^64 - q) mod 2^64
  return d;
}
NS_HD uint64_t fneg(uint64_t a) { return a ? Q - a : 0; }
#if defined(__HIP_DEVICE_COMPILE__)
// 64 x 64 -> 128 from four 32 x 32 + 64 multiply-adds (v_mad_u64_u32); every partial sum fits in
// 64 bits, so no carry chains are needed.
__device__ __forceinline__ void mul_wide(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  const uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32), bl = (uint32_t)b, bh = (uint32_t)(b >> 32);
  const uint64_t p0 = (uint64_t)al * bl;
  const uint64_t p1 = (uint64_t)al * bh + (p0 >> 32);
  const uint64_t p2 = (uint64_t)ah * bl + (uint32_t)p1;
  hi = (uint64_t)ah * bh + (p1 >> 32) + (p2 >> 32);
  lo = (p2 << 32) | (uint32_t)p0;
}
// reduce128 with hi_lo * (2^32 - 1) formed from 32-bit pieces instead of a multiply
__device__ __forceinline__ uint64_t reduce128_dev(uint64_t lo, uint64_t hi) {
  const uint32_t hl = (uint32_t)hi, hh = (uint32_t)(hi >> 32);
  uint64_t t0 = lo - hh;
  t0 -= (uint64_t)0 - (uint64_t)(lo < hh) & EPS;
  const uint64_t t1 = ((uint64_t)(hl - (uint32_t)(hl != 0)) << 32) | (uint32_t)(0u - hl);
  uint64_t t2 = t0 + t1;
  t2 += (uint64_t)0 - (uint64_t)(t2 < t1) & EPS;
  return t2 >= Q ? t2 - Q : t2;
}
NS_HD uint64_t fmul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul_wide(a, b, lo, hi);
  return reduce128_dev(lo, hi);
}
// 7 a = 8 a - a as a 128-bit value
NS_HD uint64_t fmul7(uint64_t a) {
  const uint64_t a8 = a << 3;
  const uint64_t lo = a8 - a;
  const uint64_t hi = (a >> 61) - (uint64_t)(a8 < a);
  return reduce128_dev(lo, hi);
}
#else
NS_HD uint64_t fmul(uint64_t a, uint64_t b) { return reduce128(a * b, mulhi64(a, b)); }
NS_HD uint64_t fmul7(uint64_t a) { return reduce128(a * 7ull, mulhi64(a, 7ull)); }
#endif
NS_HD uint64_t fdouble(uint64_t a) { return fadd(a, a); }
// Signed small integer into F_q.
NS_HD uint64_t from_i64(int64_t v) { return v >= 0 ? (uint64_t)v : Q - (uint64_t)(-v); }
NS_HD int64_t to_balanced(uint64_t a) { return a <= (Q - 1) / 2 ? (int64_t)a : -(int64_t)(Q - a); }

inline uint64_t fpow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = fmul(r, a);
    a = fmul(a, a);
    e >>= 1;
  }
  return r;
}
inline uint64_t finv(uint64_t a) { return fpow(a, Q - 2); }

struct K2 {
  uint64_t c0, c1;
};

NS_HD K2 kmake(uint64_t a, uint64_t b) { return K2{a, b}; }
NS_HD K2 kzero() { return K2{0, 0}; }
NS_HD K2 kone() { return K2{1, 0}; }
NS_HD K2 kadd(K2 a, K2 b) { return K2{fadd(a.c0, b.c0), fadd(a.c1, b.c1)}; }
NS_HD K2 ksub(K2 a, K2 b) { return K2{fsub(a.c0, b.c0), fsub(a.c1, b.c1)}; }
NS_HD K2 kneg(K2 a) { return K2{fneg(a.c0), fneg(a.c1)}; }
NS_HD K2 kscale(K2 a, uint64_t s) { return K2{fmul(a.c0, s), fmul(a.c1, s)}; }
NS_HD bool keq(K2 a, K2 b) { return a.c0 == b.c0 && a.c1 == b.c1; }
NS_HD bool kiszero(K2 a) { return (a.c0 | a.c1) == 0; }
// (a0 + a1 u)(b0 + b1 u) = (a0 b0 + 7 a1 b1) + ((a0 + a1)(b0 + b1) - a0 b0 - a1 b1) u
NS_HD K2 kmul(K2 a, K2 b) {
  uint64_t p0 = fmul(a.c0, b.c0);
  uint64_t p1 = fmul(a.c1, b.c1);
  uint64_t p2 = fmul(fadd(a.c0, a.c1), fadd(b.c0, b.c1));
  return K2{fadd(p0, fmul7(p1)), fsub(fsub(p2, p0), p1)};
}
NS_HD K2 ksqr(K2 a) {
  // (a0^2 + 7 a1^2) + 2 a0 a1 u
  uint64_t p0 = fmul(a.c0, a.c0);
  uint64_t p1 = fmul(a.c1, a.c1);
  uint64_t x = fmul(a.c0, a.c1);
  return K2{fadd(p0, fmul7(p1)), fadd(x, x)};
}
// a * s for a small signed integer s in [-8, 8]
NS_HD K2 kmul_small(K2 a, int s) {
  K2 r = kzero();
  K2 b = s < 0 ? kneg(a) : a;
  int m = s < 0 ? -s : s;
  for (int i = 0; i < m; i++) r = kadd(r, b);
  return r;
}
inline K2 kinv(K2 a) {
  uint64_t n = fsub(fmul(a.c0, a.c0), fmul7(fmul(a.c1, a.c1)));
  uint64_t ni = finv(n);
  return K2{fmul(a.c0, ni), fmul(fneg(a.c1), ni)};
}

}  // namespace ns
