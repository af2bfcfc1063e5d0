// Arithmetic policy of the two engines, host/device.
//   float  : the deterministic own routines of fpmath.h (mixed engine; the CPU
//            replay reproduces them bit for bit)
//   double : the platform libm (plain FP64 engine)
#pragma once

#include <stdint.h>
#include <math.h>
#include "fpmath.h"
#include "hd.h"
#include "philox.h"

namespace storm {

template <typename Real>
struct MathPolicy;

template <>
struct MathPolicy<float> {
  STORM_HD static float log(float x) { return fpm::logf_d(x); }
  STORM_HD static float exp(float x) { return fpm::expf_d(x); }
  STORM_HD static float expm1(float x) { return fpm::expm1f_d(x); }
  STORM_HD static void sincos2pi(float u, float &s, float &c) { fpm::sincos2pi(u, s, c); }
  STORM_HD static float sqrt(float x) { return ::sqrtf(x); }
  STORM_HD static float uniform(uint32_t x) { return u01_f32(x); }
  STORM_HD static float inf() { return fpm::u2f(0x7f800000u); }
};

template <>
struct MathPolicy<double> {
  STORM_HD static double log(double x) { return ::log(x); }
  STORM_HD static double exp(double x) { return ::exp(x); }
  STORM_HD static double expm1(double x) { return ::expm1(x); }
  STORM_HD static void sincos2pi(double u, double &s, double &c) {
    double a = 6.283185307179586476925286766559 * u;
    s = ::sin(a);
    c = ::cos(a);
  }
  STORM_HD static double sqrt(double x) { return ::sqrt(x); }
  STORM_HD static double uniform(uint32_t x) { return (static_cast<double>(x) + 0.5) * 2.3283064365386963e-10; }
  STORM_HD static double inf() { return fpm::u2d(0x7ff0000000000000ull); }
};

// Isotropic unit vector from two uniforms: mu = 2 u0 - 1, phi = 2 pi u1
// (isotropicVelocity, radiation/RadiationOpacityModel.hpp:139-150).
template <typename Real>
STORM_HD void isotropic_dir(Real u0, Real u1, Real o[3]) {
  Real mu = Real(2) * u0 - Real(1);
  Real st2 = Real(1) - mu * mu;
  Real st = MathPolicy<Real>::sqrt(st2 > Real(0) ? st2 : Real(0));
  Real sp, cp;
  MathPolicy<Real>::sincos2pi(u1, sp, cp);
  o[0] = st * cp;
  o[1] = st * sp;
  o[2] = mu;
}

}  // namespace storm
