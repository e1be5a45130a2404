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
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t lo, hi;
  mul_wide(a, b, lo, hi);
  return reduce128_dev(lo, hi);
#else
  return reduce128(a * b, mulhi64(a, b));
#endif
}
// 7 a = 8 a - a as a 128-bit value
NS_HD uint64_t fmul7(uint64_t a) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t a8 = a << 3;
  const uint64_t lo = a8 - a;
  const uint64_t hi = (a >> 61) - (uint64_t)(a8 < a);
  return reduce128_dev(lo, hi);
#else
  return reduce128(a * 7ull, mulhi64(a, 7ull));
#endif
}
// ---- lazy sums (device): values are added without reduction and reduced once at the end ----
// S1: a sum of field elements, v + 2^64 c (c counts the carries out of v).
struct S1 {
  uint64_t v;
  uint32_t c;
};
__device__ __forceinline__ S1 s1_zero() { return S1{0, 0}; }
__device__ __forceinline__ void s1_add(S1& s, uint64_t x) {
  const uint64_t t = s.v + x;
  s.c += (uint32_t)(t < x);
  s.v = t;
}
__device__ __forceinline__ uint64_t s1_reduce(const S1& s) { return reduce128_dev(s.v, s.c); }
// S2: a sum of 64 x 64-bit products, lo + 2^64 (hi + cl) + 2^128 ch.
struct S2 {
  uint64_t lo, hi;
  uint32_t cl, ch;
};
__device__ __forceinline__ S2 s2_zero() { return S2{0, 0, 0, 0}; }
__device__ __forceinline__ void s2_mac(S2& s, uint64_t x, uint64_t y) {
  uint64_t pl, ph;
  mul_wide(x, y, pl, ph);
  const uint64_t t = s.lo + pl;
  s.cl += (uint32_t)(t < pl);
  s.lo = t;
  const uint64_t u = s.hi + ph;
  s.ch += (uint32_t)(u < ph);
  s.hi = u;
}
// 2^128 = (2^32 - 1)^2 = -2^32 (mod q); ch < 2^32 keeps ch * 2^32 below q.
__device__ __forceinline__ uint64_t s2_reduce(const S2& s) {
  const uint64_t h = s.hi + s.cl;
  const uint32_t ch = s.ch + (uint32_t)(h < s.hi);
  return fsub(reduce128_dev(s.lo, h), (uint64_t)ch << 32);
}
// x0 y0 + x1 y1 reduced once (two 128-bit products, at most one carry into 2^128)
__device__ __forceinline__ uint64_t fdot2(uint64_t x0, uint64_t y0, uint64_t x1, uint64_t y1) {
  uint64_t l0, h0, l1, h1;
  mul_wide(x0, y0, l0, h0);
  mul_wide(x1, y1, l1, h1);
  const uint64_t lo = l0 + l1;
  const uint64_t hs = h0 + h1;  // h0, h1 <= 2^64 - 2
  const uint64_t hi = hs + (uint64_t)(lo < l1);
  const bool top = (hs < h1) | (hi < hs);
  const uint64_t r = reduce128_dev(lo, hi);
  return top ? fsub(r, (uint64_t)1 << 32) : r;
}
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
// (w0 + w1 u)(a0 + a1 u) with w7 = 7 w1 given: (w0 a0 + w7 a1) + (w0 a1 + w1 a0) u, each
// component a single reduction of two wide products (for a multiplicand reused many times).
__device__ __forceinline__ K2 kmul_w7(K2 w, uint64_t w7, K2 a) {
  return K2{fdot2(w.c0, a.c0, w7, a.c1), fdot2(w.c0, a.c1, w.c1, a.c0)};
}
// a^2 = (a0^2 + 7 a1^2) + 2 a0 a1 u
__device__ __forceinline__ K2 ksqr_dev(K2 a) {
  const uint64_t x = fmul(a.c0, a.c1);
  return K2{fdot2(a.c0, a.c0, fmul7(a.c1), a.c1), fadd(x, x)};
}
// lazy K sums: KS1 of K elements, KS2 of K products (x0 y0, x1 y1 and x0 y1 + x1 y0 kept apart,
// the 7 applied once at the end)
struct KS1 {
  S1 c0, c1;
};
struct KS2 {
  S2 p00, p11, px;
};
__device__ __forceinline__ KS1 ks1_zero() { return KS1{s1_zero(), s1_zero()}; }
__device__ __forceinline__ KS2 ks2_zero() { return KS2{s2_zero(), s2_zero(), s2_zero()}; }
__device__ __forceinline__ void ks1_add(KS1& s, K2 x) {
  s1_add(s.c0, x.c0);
  s1_add(s.c1, x.c1);
}
__device__ __forceinline__ K2 ks1_reduce(const KS1& s) { return K2{s1_reduce(s.c0), s1_reduce(s.c1)}; }
__device__ __forceinline__ void ks2_mac(KS2& s, K2 x, K2 y) {
  s2_mac(s.p00, x.c0, y.c0);
  s2_mac(s.p11, x.c1, y.c1);
  s2_mac(s.px, x.c0, y.c1);
  s2_mac(s.px, x.c1, y.c0);
}
__device__ __forceinline__ K2 ks2_reduce(const KS2& s) {
  return K2{fadd(s2_reduce(s.p00), fmul7(s2_reduce(s.p11))), s2_reduce(s.px)};
}
// Karatsuba form of the same lazy sum: px holds sum (x0 + x1)(y0 + y1) and the cross term is
// px - p00 - p11, so a product costs three wide products; the operand sums (ksum, a congruent u64)
// are formed once per operand and shared by every product it enters.
__device__ __forceinline__ uint64_t ksum(K2 x) {
  uint64_t s;
  const bool c = __builtin_add_overflow(x.c0, x.c1, &s);
  return s + ((uint64_t)0 - (uint64_t)c & EPS);  // x0 + x1 < 2^65 - 2^33: one fold of 2^64 = EPS
}
__device__ __forceinline__ void ks2k_mac(KS2& s, K2 x, uint64_t xs, K2 y, uint64_t ys) {
  s2_mac(s.p00, x.c0, y.c0);
  s2_mac(s.p11, x.c1, y.c1);
  s2_mac(s.px, xs, ys);
}
__device__ __forceinline__ K2 ks2k_reduce(const KS2& s) {
  const uint64_t r00 = s2_reduce(s.p00), r11 = s2_reduce(s.p11);
  return K2{fadd(r00, fmul7(r11)), fsub(fsub(s2_reduce(s.px), r00), r11)};
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
