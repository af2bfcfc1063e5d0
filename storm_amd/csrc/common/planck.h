// Normalised Planck integral, for the group structure of multigroup IMC:
//   planck_integral(a, b) = 15 / pi^4 * int_a^b x^3 / (e^x - 1) dx,
// the fraction of a black body's energy between photon energies a kT and
// b kT (the quantity the reference takes from the external planck_integral
// library, radiation/RadiationIMC.hpp:1022, boundary/SideTemperature.hpp:72).
//
// Written here from two textbook expansions of F(x) = int_0^x t^3/(e^t-1) dt:
//   x < 3/2: t/(e^t - 1) = sum_k B_k t^k / k!  (Bernoulli numbers), so
//            F(x) = sum_k B_k x^(k+3) / (k! (k+3)); the series converges for
//            x < 2 pi and the terms up to k = 22 leave ~1e-15 at x = 3/2;
//   x >= 3/2: 1/(e^t - 1) = sum_n e^(-n t), so int_x^inf t^3 e^(-n t) dt gives
//            int_x^inf = sum_n e^(-n x) (x^3/n + 3x^2/n^2 + 6x/n^3 + 6/n^4),
//            28 terms for x = 3/2 and fewer above.
// Above x = 3/2 the difference of two upper tails is used, so that narrow
// high-energy groups keep their relative precision.
#pragma once

#include <math.h>
#include "hd.h"

namespace storm {
namespace planck {

constexpr double kNorm = 15.0 / (3.14159265358979323846 * 3.14159265358979323846 * 3.14159265358979323846 *
                                 3.14159265358979323846);

// int_0^x t^3/(e^t - 1) dt for 0 <= x < 3/2 (series in Bernoulli numbers)
STORM_HD double lower_series(double x) {
  // B_k / k! for k = 0, 1, 2, 4, ..., 22 (odd k > 1 vanish)
  const double c[13] = {1.0,
                        -0.5,
                        1.0 / 12.0,
                        -1.0 / 720.0,
                        1.0 / 30240.0,
                        -1.0 / 1209600.0,
                        1.0 / 47900160.0,
                        -691.0 / 1307674368000.0,
                        1.0 / 74724249600.0,
                        -3617.0 / 10670622842880000.0,
                        43867.0 / 5109094217170944000.0,
                        -174611.0 / 802857662698291200000.0,
                        77683.0 / 14101100039391805440000.0};
  const double x2 = x * x, x3 = x2 * x;
  // k = 0 and k = 1 terms
  double s = c[0] * x3 / 3.0 + c[1] * x3 * x / 4.0;
  double xp = x3 * x2;  // x^(k+3) for k = 2
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int q = 2; q < 13; ++q) {
    const int k = 2 * (q - 1);
    s += c[q] * xp / static_cast<double>(k + 3);
    xp *= x2;
  }
  return s;
}

// int_x^inf t^3/(e^t - 1) dt for x >= 3/2 (exponential series)
STORM_HD double upper_series(double x) {
  const double x2 = x * x, x3 = x2 * x;
  double s = 0.0;
  for (int n = 1; n <= 28; ++n) {
    const double dn = static_cast<double>(n);
    const double e = exp(-dn * x);
    if (e == 0.0) break;
    s += e * (x3 / dn + 3.0 * x2 / (dn * dn) + 6.0 * x / (dn * dn * dn) + 6.0 / (dn * dn * dn * dn));
    if (e < 1e-18) break;
  }
  return s;
}

// fraction of the black-body energy below x = E / kT
STORM_HD double cdf(double x) {
  if (!(x > 0.0)) return 0.0;
  if (x < 1.5) return kNorm * lower_series(x);
  return 1.0 - kNorm * upper_series(x);
}

// fraction of the black-body energy above x = E / kT
STORM_HD double tail(double x) {
  if (!(x > 0.0)) return 1.0;
  if (x < 1.5) return 1.0 - kNorm * lower_series(x);
  return kNorm * upper_series(x);
}

// fraction between a kT and b kT
STORM_HD double integral(double a, double b) {
  if (!(b > a)) return 0.0;
  if (a >= 1.5) return tail(a) - tail(b);
  return cdf(b) - cdf(a);
}

}  // namespace planck
}  // namespace storm
