// Deterministic single-precision elementary functions shared by the gfx950
// transport kernels and their CPU replay.
//
// Vendor libm / device math libraries differ between host and device, so the
// mixed-precision engine uses only these routines, built from IEEE-754 +, -,
// *, /, sqrt and fused multiply-add, all correctly rounded on both sides.
// Every translation unit that includes this header is compiled with
// -ffp-contract=off, so the compiler forms no FMA of its own; the polynomial
// steps below ask for one explicitly (fma_f: v_fma_f32 on the device, the
// correctly rounded fmaf of the C library on the host).
// Methods:
//   log   : exponent/mantissa split with m in [sqrt(1/2), sqrt(2)), then
//           log(m) = 2 atanh(s), s = (m-1)/(m+1), atanh series to s^9;
//           ln2 split in two parts (Cody & Waite, "Software Manual for the
//           Elementary Functions", 1980).
//   exp   : Cody-Waite reduction y = k ln2 + r, |r| <= ln2/2, Taylor series
//           of e^r to r^8, scaling by 2^k through the exponent field.
//   expm1 : Taylor series to y^9 for |y| < ln2/2, exp(y) - 1 otherwise.
//   sincos of 2*pi*u : exact quadrant reduction of 4u, Taylor series of sin
//           and cos on [-pi/4, pi/4].
// Accuracy is a few ulp, far below Monte Carlo noise; what matters is that
// host and device produce the same bits.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>
#include "hd.h"

namespace storm {
namespace fpm {

STORM_HD uint32_t f2u(float x) {
  uint32_t u;
  memcpy(&u, &x, sizeof(u));
  return u;
}
STORM_HD float u2f(uint32_t u) {
  float x;
  memcpy(&x, &u, sizeof(x));
  return x;
}
STORM_HD uint64_t d2u(double x) {
  uint64_t u;
  memcpy(&u, &x, sizeof(u));
  return u;
}
STORM_HD double u2d(uint64_t u) {
  double x;
  memcpy(&x, &u, sizeof(x));
  return x;
}

// a * b + c with one rounding
STORM_HD float fma_f(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fmaf(a, b, c);
#else
  return ::fmaf(a, b, c);
#endif
}

constexpr float kLn2Hi = 0.693115234375f;          // 12 significant bits
constexpr float kLn2Lo = 3.194618329871446e-05f;   // ln2 - kLn2Hi
constexpr float kLog2e = 1.4426950216293335f;
constexpr float kHalfPi = 1.5707963705062866f;
constexpr float kSqrt2 = 1.4142135381698608f;

// Natural log for finite x > 0 (normal or subnormal).  Returns -inf for 0.
STORM_HD float logf_d(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -u2f(0x7f800000u) : u2f(0x7fc00000u);
  uint32_t b = f2u(x);
  int e = 0;
  if (b < 0x00800000u) {  // subnormal: scale by 2^24 first
    x = x * 16777216.0f;
    b = f2u(x);
    e = -24;
  }
  e += static_cast<int>((b >> 23) & 0xffu) - 127;
  float m = u2f((b & 0x007fffffu) | 0x3f800000u);  // [1, 2)
  if (m > kSqrt2) {
    m = m * 0.5f;
    e += 1;
  }
  float f = m - 1.0f;
  float s = f / (2.0f + f);
  float z = s * s;
  float p = 1.0f / 9.0f;
  p = fma_f(z, p, 1.0f / 7.0f);
  p = fma_f(z, p, 1.0f / 5.0f);
  p = fma_f(z, p, 1.0f / 3.0f);
  p = fma_f(z, p, 1.0f);
  float lm = 2.0f * s * p;
  float ef = static_cast<float>(e);
  return fma_f(ef, kLn2Hi, fma_f(ef, kLn2Lo, lm));
}

// e^y.  Flushes to 0 below -87 and saturates to +inf above 88.
STORM_HD float expf_d(float y) {
  if (y < -87.0f) return 0.0f;
  if (y > 88.0f) return u2f(0x7f800000u);
  float kf = y * kLog2e;
  int k = static_cast<int>(kf + (kf >= 0.0f ? 0.5f : -0.5f));
  float fk = static_cast<float>(k);
  float r = fma_f(-fk, kLn2Lo, fma_f(-fk, kLn2Hi, y));
  float p = 1.0f / 40320.0f;
  p = fma_f(r, p, 1.0f / 5040.0f);
  p = fma_f(r, p, 1.0f / 720.0f);
  p = fma_f(r, p, 1.0f / 120.0f);
  p = fma_f(r, p, 1.0f / 24.0f);
  p = fma_f(r, p, 1.0f / 6.0f);
  p = fma_f(r, p, 0.5f);
  p = fma_f(r, p, 1.0f);
  p = fma_f(r, p, 1.0f);
  // scale by 2^k in two halves so that k down to -126-ish stays representable
  int k1 = k / 2;
  int k2 = k - k1;
  float s1 = u2f(static_cast<uint32_t>(k1 + 127) << 23);
  float s2 = u2f(static_cast<uint32_t>(k2 + 127) << 23);
  return (p * s1) * s2;
}

// e^y - 1.
STORM_HD float expm1f_d(float y) {
  float ay = y < 0.0f ? -y : y;
  if (ay < 0.34657359f) {
    float p = 1.0f / 362880.0f;
    p = fma_f(y, p, 1.0f / 40320.0f);
    p = fma_f(y, p, 1.0f / 5040.0f);
    p = fma_f(y, p, 1.0f / 720.0f);
    p = fma_f(y, p, 1.0f / 120.0f);
    p = fma_f(y, p, 1.0f / 24.0f);
    p = fma_f(y, p, 1.0f / 6.0f);
    p = fma_f(y, p, 0.5f);
    return fma_f(y * y, p, y);
  }
  return expf_d(y) - 1.0f;
}

// sin and cos of 2*pi*u for u in [0, 1].
STORM_HD void sincos2pi(float u, float &s_out, float &c_out) {
  float v = 4.0f * u;  // quadrant coordinate, exact
  int q = static_cast<int>(v + 0.5f);
  float t = v - static_cast<float>(q);  // [-0.5, 0.5], exact
  float a = t * kHalfPi;
  float a2 = a * a;
  float sp = 1.0f / 362880.0f;
  sp = fma_f(a2, sp, -1.0f / 5040.0f);
  sp = fma_f(a2, sp, 1.0f / 120.0f);
  sp = fma_f(a2, sp, -1.0f / 6.0f);
  float s = fma_f(a * a2, sp, a);
  float cp = -1.0f / 3628800.0f;
  cp = fma_f(a2, cp, 1.0f / 40320.0f);
  cp = fma_f(a2, cp, -1.0f / 720.0f);
  cp = fma_f(a2, cp, 1.0f / 24.0f);
  cp = fma_f(a2, cp, -0.5f);
  float c = fma_f(a2, cp, 1.0f);
  switch (q & 3) {
    case 0: s_out = s; c_out = c; break;
    case 1: s_out = c; c_out = -s; break;
    case 2: s_out = -s; c_out = -c; break;
    default: s_out = -c; c_out = s; break;
  }
}

// Exact power of two as a double, 2^k; k is clamped to the normal range
// [-1022, 1023].
STORM_HD double pow2d(int k) {
  k = k < -1022 ? -1022 : (k > 1023 ? 1023 : k);
  return u2d(static_cast<uint64_t>(k + 1023) << 52);
}

}  // namespace fpm
}  // namespace storm
