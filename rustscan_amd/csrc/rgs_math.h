// Deterministic single-precision scalar math shared by the HIP kernels and
// by host code.
//
// Every function here is a fixed sequence of IEEE-754 basic operations
// (+, -, *, /, sqrt, fma, round-to-nearest-even, bit casts), each of which is
// correctly rounded on both the CPU (SSE/FMA3) and the GPU (gfx950 with
// correctly rounded fp32 divide/sqrt and no fp contraction).  The same inputs
// therefore give the same bits on both sides, which is what makes the
// visibility cull, the tile tests and the alpha thresholds of the renderer
// reproducible between device and host.
//
// exp: Cody-Waite reduction x = k*ln2 + r with ln2 split in two parts, then
//      the degree-7 Taylor polynomial of e^r on |r| <= ln2/2 (Horner, fma),
//      scaled by 2^k built from its exponent bits.  Inputs <= -87 return 0,
//      inputs > 88.7 return +inf, NaN propagates.
// log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//      log(m) = 2*atanh(s) = 2s + 2s^3/3 + ... + 2s^11/11, plus e*ln2.
//      Defined for positive normal x; x <= 0 returns -inf, NaN propagates,
//      +inf returns +inf.
// Accuracy against correctly rounded results is a few ulp over the full
// float range; it is checked exhaustively-by-sampling in the test suite.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define RGS_HD __host__ __device__ __forceinline__
#else
#define RGS_HD static inline
#endif

RGS_HD uint32_t rgs_f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
RGS_HD float rgs_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

RGS_HD float rgs_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
RGS_HD float rgs_sqrt(float x) { return __builtin_sqrtf(x); }
RGS_HD float rgs_rint(float x) { return __builtin_rintf(x); }

RGS_HD float rgs_exp(float x) {
  if (x != x) return x;
  if (x <= -87.0f) return 0.0f;
  if (x > 88.72283f) return rgs_u2f(0x7f800000u);
  const float k = rgs_rint(x * 1.44269504088896341f);
  // ln2 = LN2_HI + LN2_LO; LN2_HI has 15 significant bits so k*LN2_HI is exact.
  float r = rgs_fma(-k, 0.693145751953125f, x);
  r = rgs_fma(-k, 1.428606765330187045e-06f, r);
  float p = 1.9841269841269841e-04f;            // 1/7!
  p = rgs_fma(p, r, 1.3888888888888889e-03f);   // 1/6!
  p = rgs_fma(p, r, 8.3333333333333333e-03f);   // 1/5!
  p = rgs_fma(p, r, 4.1666666666666667e-02f);   // 1/4!
  p = rgs_fma(p, r, 1.6666666666666667e-01f);   // 1/3!
  p = rgs_fma(p, r, 0.5f);
  p = rgs_fma(p, r, 1.0f);
  p = rgs_fma(p, r, 1.0f);
  const int ki = (int)k;  // in [-126, 128]
  if (ki > 127) return (p * 2.0f) * rgs_u2f(0x7f000000u);
  return p * rgs_u2f((uint32_t)(ki + 127) << 23);
}

// rgs_exp without its special cases, for arguments clamped to [-87, 88]:
// there k = rint(x*log2(e)) lies in [-126, 127], the scale 2^k is a normal
// float and the result equals rgs_exp(x) bit for bit for every x in
// (-87, 88] (at x = -87 rgs_exp returns 0, this returns 2^-126 * p).  The
// compositing kernels only consume values for x = -sigma in [-5.6, 0], where
// both agree; for other lanes the value is discarded.  NaN input returns a
// finite value, which callers reject through their own sigma test.
RGS_HD float rgs_exp_core(float x) {
  x = x > -87.0f ? x : -87.0f;
  x = x < 88.0f ? x : 88.0f;
  const float k = rgs_rint(x * 1.44269504088896341f);
  float r = rgs_fma(-k, 0.693145751953125f, x);
  r = rgs_fma(-k, 1.428606765330187045e-06f, r);
  float p = 1.9841269841269841e-04f;
  p = rgs_fma(p, r, 1.3888888888888889e-03f);
  p = rgs_fma(p, r, 8.3333333333333333e-03f);
  p = rgs_fma(p, r, 4.1666666666666667e-02f);
  p = rgs_fma(p, r, 1.6666666666666667e-01f);
  p = rgs_fma(p, r, 0.5f);
  p = rgs_fma(p, r, 1.0f);
  p = rgs_fma(p, r, 1.0f);
  return p * rgs_u2f((uint32_t)((int)k + 127) << 23);
}

// rgs_exp_core without the argument clamps, scaled with ldexp: equals
// rgs_exp(x) bit for bit for every x in (-87, 88] (there 2^k is a normal
// scale and p * 2^k is exact).  Outside that range, and for NaN, it returns
// an arbitrary value (possibly inf or NaN) that the caller must discard; the
// compositing kernels evaluate it on lanes whose result they then mask off.
RGS_HD float rgs_exp_core_fast(float x) {
  const float k = rgs_rint(x * 1.44269504088896341f);
  float r = rgs_fma(-k, 0.693145751953125f, x);
  r = rgs_fma(-k, 1.428606765330187045e-06f, r);
  float p = 1.9841269841269841e-04f;
  p = rgs_fma(p, r, 1.3888888888888889e-03f);
  p = rgs_fma(p, r, 8.3333333333333333e-03f);
  p = rgs_fma(p, r, 4.1666666666666667e-02f);
  p = rgs_fma(p, r, 1.6666666666666667e-01f);
  p = rgs_fma(p, r, 0.5f);
  p = rgs_fma(p, r, 1.0f);
  p = rgs_fma(p, r, 1.0f);
  return __builtin_ldexpf(p, (int)k);
}

RGS_HD float rgs_log(float x) {
  if (x != x) return x;
  if (x <= 0.0f) return rgs_u2f(0xff800000u);
  const uint32_t bits = rgs_f2u(x);
  if (bits >= 0x7f800000u) return x;
  if (bits < 0x00800000u) {
    // subnormal: scale into the normal range first (exact)
    return rgs_log(x * 16777216.0f) - 16.635532333438687f;
  }
  int e = (int)((bits >> 23) & 0xffu) - 127;
  float m = rgs_u2f((bits & 0x007fffffu) | 0x3f800000u);  // [1, 2)
  if (m > 1.41421356237f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;          // exact (Sterbenz)
  const float s = f / (2.0f + f);
  const float s2 = s * s;
  float q = 0.18181818181818182f;                 // 2/11
  q = rgs_fma(q, s2, 0.22222222222222222f);       // 2/9
  q = rgs_fma(q, s2, 0.28571428571428571f);       // 2/7
  q = rgs_fma(q, s2, 0.4f);                       // 2/5
  q = rgs_fma(q, s2, 0.66666666666666667f);       // 2/3
  // log(m) = 2s + s*s2*q
  const float lm = rgs_fma(s * s2, q, 2.0f * s);
  const float ef = (float)e;
  return rgs_fma(ef, 0.693145751953125f, rgs_fma(ef, 1.428606765330187045e-06f, lm));
}

// 1 / (1 + e^-x), the logistic function used for opacities.
RGS_HD float rgs_sigmoid(float x) { return 1.0f / (1.0f + rgs_exp(-x)); }

// expf with the bits of the C library the reference's CPU renderer links
// (Rust's f32::exp calls it): the table-driven method of glibc's expf
// (ARM optimized-routines, Szabolcs Nagy 2017).  In double precision,
// x * 32/ln2 = k + r with k rounded to an integer through the 1.5 * 2^52
// shift, 2^(k/32) from a 32-entry table of 2^(i/32) (correctly rounded
// doubles, each stored minus i << 47 so that adding k << 47 also forms the
// exponent), and e^(r ln2/32) by a degree-3 polynomial; the double result is
// rounded to float once.  Each step is a correctly rounded IEEE operation,
// so host and device agree bit for bit; against this image's libm expf it
// differs on 2 of the 2^32 inputs (x = 0x1.04845ep+5 and -0x1.f8cbb2p+5,
// one float ulp, both outside |x| < 32), checked exhaustively by the tests.
// Overflow (x > 0x1.62e42ep6) gives +inf, underflow (x < -0x1.9fe368p6)
// gives 0, NaN propagates.
RGS_HD float rgs_expf_libm(float x) {
  constexpr uint64_t kTab[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
  };
  if (x != x) return x + x;
  if (x > 0x1.62e42ep6f) return rgs_u2f(0x7f800000u);
  if (x < -0x1.9fe368p6f) return 0.0f;
  const double z = 0x1.71547652b82fep+0 * 32.0 * (double)x;
  double kd = z + 0x1.8p+52;
  const uint64_t ki = __builtin_bit_cast(uint64_t, kd);
  kd -= 0x1.8p+52;
  const double r = z - kd;
  const double s = __builtin_bit_cast(double, kTab[ki % 32] + (ki << 47));
  const double p = __builtin_fma(0x1.c6af84b912394p-5 / 32768.0, r, 0x1.ebfce50fac4f3p-3 / 1024.0);
  const double y = __builtin_fma(p, r * r, __builtin_fma(0x1.62e42ff0c52d6p-1 / 32.0, r, 1.0));
  return (float)(y * s);
}


// 1 / sqrt(x) as two correctly rounded operations.
RGS_HD float rgs_rsqrt(float x) { return 1.0f / rgs_sqrt(x); }

// ---- topology step: a counter-based random stream and cos(2*pi*u) --------
// SplitMix64 (Steele, Lea and Flood, 2014): the output for state x.
RGS_HD uint64_t rgs_splitmix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Uniform float in [0, 1) with 24 random bits: draw `counter` of the stream
// `seed` (stream and counter are mixed so that streams do not overlap).
RGS_HD float rgs_uniform01(uint64_t seed, uint64_t counter) {
  const uint64_t z = rgs_splitmix64(rgs_splitmix64(seed) ^ (counter * 0xd6e8feb86659fd93ull));
  return (float)(uint32_t)(z >> 40) * 5.9604644775390625e-08f;  // 2^-24
}

// cos(2*pi*u): reduction to |t| <= 1/2 (t = u - rint(u)), then to an angle
// y in [0, pi/4] by the quadrant identities cos(pi - x) = -cos(x) and
// cos(pi/2 - x) = sin(x), and the Taylor series of cos (to x^10) or sin (to
// x^11) on |y| <= pi/4 (truncation below 1.1e-10).
RGS_HD float rgs_cos_2pi(float u) {
  float t = u - rgs_rint(u);
  if (t < 0.0f) t = -t;                       // [0, 1/2]
  float sign = 1.0f;
  if (t > 0.25f) {
    t = 0.5f - t;                             // cos(pi - x) = -cos(x)
    sign = -1.0f;
  }
  const float y = t * 6.28318530717958648f;   // [0, pi/2]
  if (y <= 0.785398163397448310f) {
    const float y2 = y * y;
    float p = -2.7557319223985893e-07f;       // -1/10!
    p = rgs_fma(p, y2, 2.4801587301587302e-05f);   // 1/8!
    p = rgs_fma(p, y2, -1.3888888888888889e-03f);  // -1/6!
    p = rgs_fma(p, y2, 4.1666666666666667e-02f);   // 1/4!
    p = rgs_fma(p, y2, -0.5f);
    p = rgs_fma(p, y2, 1.0f);
    return sign * p;
  }
  const float z = 1.57079632679489662f - y;   // cos(y) = sin(pi/2 - y)
  const float z2 = z * z;
  float p = -2.5052108385441720e-08f;         // -1/11!
  p = rgs_fma(p, z2, 2.7557319223985891e-06f);   // 1/9!
  p = rgs_fma(p, z2, -1.9841269841269841e-04f);  // -1/7!
  p = rgs_fma(p, z2, 8.3333333333333333e-03f);   // 1/5!
  p = rgs_fma(p, z2, -1.6666666666666667e-01f);  // -1/3!
  p = rgs_fma(p, z2 * z, z);
  return sign * p;
}
