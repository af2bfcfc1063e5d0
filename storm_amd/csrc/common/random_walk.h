// Random-walk tables and samplers (Fleck & Canfield diffusion in a sphere),
// host/device.
//
// Tables (radiation/RandomWalk.hpp:20-79): 1024 dimensionless times tau
// log-spaced on [1e-8, 64]; the survival probability S(tau) of a walker
// started at the centre of a unit sphere with D = 1,
//     S(tau) = 2 sum_{n>=1} (-1)^(n+1) exp(-n^2 pi^2 tau),
// and for each tau the inverse of the radial CDF of the walker's position
// given survival, tabulated on 256 uniform levels xi in [0, 1]:
//     G(x, tau) = 2 / S(tau) sum_n exp(-n^2 pi^2 tau) (sin(n pi x) / (n pi) - x cos(n pi x))
// (free-space form erf(a) - 2 a exp(-a^2) / sqrt(pi), a = x / (2 sqrt(tau)),
// below tau = 0.01), inverted by linear interpolation on a 512-point x grid.
// The series are truncated after 500 terms or once n^2 pi^2 tau > 700, as
// the reference does.
//
// Samplers (RandomWalk.hpp:81-137): radius by bilinear interpolation in
// (log tau, xi), as the reference does.  The leak time inverts the same
// S(tau) through a table of the inverse itself (build_leak_table), where the
// reference searches its S table: the same distribution (the table is within
// 2e-5 in tau of the exact inverse, the search of the S table within 1.3e-3),
// with one log and two neighbouring loads in place of a chain
// of ten dependent loads per random-walk step.
#pragma once

#include <stdint.h>
#include <math.h>
#include "hd.h"
#include "mathpolicy.h"

namespace storm {
namespace rw {

constexpr int kTableSize = 1024;
constexpr int kRadiusSize = 256;
constexpr double kTauMin = 1e-8;
constexpr double kTauMax = 64.0;
// Inverse leak-time table: T(y), y = -ln S(T), on kLeakSize points uniform
// in ln y over [kLeakLogYMin, kLeakLogYMax] (y from 1.4e-11 to 4.48).
constexpr int kLeakSize = 2049;
constexpr double kLeakLogYMin = -25.0;
constexpr double kLeakLogYMax = 1.5;

// ln kTauMin and ln kTauMax
constexpr double kLogTauMin = -18.420680743952367;
constexpr double kLogTauMax = 4.1588830833596715;

template <typename Real>
struct Tables {
  const Real *tau;       // [kTableSize]
  const Real *leak;      // [kLeakSize], increasing (build_leak_table)
  const Real *radius;    // [kTableSize * kRadiusSize]
};

// The three tables in one buffer of kAllSize numbers: leak, tau, radius.
constexpr int kTauOffset = kLeakSize;
constexpr int kRadiusOffset = kLeakSize + kTableSize;
constexpr int kAllSize = kRadiusOffset + kTableSize * kRadiusSize;
template <typename Real>
STORM_HD Tables<Real> tables_at(const Real *all) {
  return Tables<Real>{all + kTauOffset, all, all + kRadiusOffset};
}

// Leak time for a uniform xi: the T with S(T) = 1 - xi.  With
// y = -ln(1 - xi), T is interpolated linearly in ln y between the two
// neighbouring table points.  Past the table the first mode alone,
// S = 2 exp(-pi^2 T), gives T = (y + ln 2) / pi^2; the next mode is below
// exp(-3 pi^2 T) < 2e-7 of it there.  Below the table (xi < 1.4e-11) T is
// the first entry; xi = 1 gives kTauMax, as the reference's table end does.
template <typename Real>
STORM_HD Real sample_leak_time(const Tables<Real> &t, Real xi) {
  using M = MathPolicy<Real>;
  const Real s = Real(1) - xi;
  if (!(s > Real(0))) return Real(kTauMax);
  // small xi: the series of -ln(1 - xi) (relative error below xi^5 / 6),
  // since 1 - xi keeps too few of xi's bits in float
  const Real y = xi < Real(0.015625)
                     ? xi * (Real(1) + xi * (Real(0.5) + xi * (Real(1) / Real(3) + xi * (Real(0.25) + xi * Real(0.2)))))
                     : -M::log(s);
  const Real ly = y > Real(0) ? M::log(y) : Real(kLeakLogYMin);
  if (ly >= Real(kLeakLogYMax)) {
    const Real T = (y + Real(0.69314718055994530942)) * Real(0.10132118364233777144);  // 1 / pi^2
    return T < Real(kTauMax) ? T : Real(kTauMax);
  }
  Real pos = (ly - Real(kLeakLogYMin)) * Real((kLeakSize - 1) / (kLeakLogYMax - kLeakLogYMin));
  if (pos < Real(0)) pos = Real(0);
  int i = static_cast<int>(pos);
  if (i > kLeakSize - 2) i = kLeakSize - 2;
  const Real f = pos - static_cast<Real>(i);
  const Real t0 = t.leak[i], t1 = t.leak[i + 1];
  return t0 + f * (t1 - t0);
}

template <typename Real>
STORM_HD Real sample_radius(const Tables<Real> &t, Real tau, Real xi) {
  using M = MathPolicy<Real>;
  const Real tmin = t.tau[0], tmax = t.tau[kTableSize - 1];
  const Real tc = tau < tmin ? tmin : (tau > tmax ? tmax : tau);
  const Real pos = (M::log(tc) - Real(kLogTauMin)) / (Real(kLogTauMax) - Real(kLogTauMin)) * Real(kTableSize - 1);
  int ti = static_cast<int>(pos);
  if (ti > kTableSize - 2) ti = kTableSize - 2;
  if (ti < 0) ti = 0;
  const Real tf = pos - static_cast<Real>(ti);
  const Real xc = xi < Real(0) ? Real(0) : (xi > Real(1) ? Real(1) : xi);
  const Real xp = xc * Real(kRadiusSize - 1);
  int xj = static_cast<int>(xp);
  if (xj > kRadiusSize - 2) xj = kRadiusSize - 2;
  const Real xf = xp - static_cast<Real>(xj);
  const Real *r0 = t.radius + ti * kRadiusSize;
  const Real *r1 = r0 + kRadiusSize;
  const Real a = r0[xj] + xf * (r0[xj + 1] - r0[xj]);
  const Real b = r1[xj] + xf * (r1[xj + 1] - r1[xj]);
  const Real r = a + tf * (b - a);
  return r < Real(0) ? Real(0) : (r > Real(1) ? Real(1) : r);
}

// ---------------------------------------------------------------- host build
// (plain host functions: built once on the host and uploaded)
inline double survival_series(double tau) {
  if (tau <= 0.0 || tau < 1e-4) return 1.0;
  const double pi = 3.14159265358979323846;
  double sum = 0.0;
  for (int n = 1; n <= 500; ++n) {
    double arg = -static_cast<double>(n * n) * pi * pi * tau;
    if (arg < -700.0) break;
    double term = exp(arg);
    sum += (n % 2 == 1) ? term : -term;
  }
  double s = 2.0 * sum;
  return s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
}

inline double radial_cdf(double x, double tau) {
  const double pi = 3.14159265358979323846;
  if (tau <= 0.0) return x > 0.0 ? 1.0 : 0.0;
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  double v;
  if (tau < 0.01) {
    double a = x / (2.0 * sqrt(tau));
    v = erf(a) - 2.0 * a * exp(-a * a) / sqrt(pi);
  } else {
    double S = survival_series(tau);
    if (S < 1e-30) return 1.0;
    double sum = 0.0;
    for (int n = 1; n <= 500; ++n) {
      double arg = -static_cast<double>(n * n) * pi * pi * tau;
      if (arg < -700.0) break;
      double npi = static_cast<double>(n) * pi, npx = npi * x;
      sum += exp(arg) * (sin(npx) / npi - x * cos(npx));
    }
    v = 2.0 * sum / S;
  }
  return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

// Fills tau[1024], survival[1024], radius[1024*256] in double precision.
inline void build_tables(double *tau, double *survival, double *radius) {
  const double lmin = log(kTauMin), lmax = log(kTauMax);
  for (int i = 0; i < kTableSize; ++i) {
    double f = static_cast<double>(i) / static_cast<double>(kTableSize - 1);
    tau[i] = exp(lmin + f * (lmax - lmin));
    survival[i] = survival_series(tau[i]);
  }
  constexpr int X = 512;
  double g[X];
  for (int i = 0; i < kTableSize; ++i) {
    for (int k = 0; k < X; ++k) g[k] = radial_cdf(static_cast<double>(k) / (X - 1), tau[i]);
    for (int j = 0; j < kRadiusSize; ++j) {
      double xi = static_cast<double>(j) / (kRadiusSize - 1);
      double &out = radius[i * kRadiusSize + j];
      if (xi <= g[0]) {
        out = 0.0;
      } else if (xi >= g[X - 1]) {
        out = 1.0;
      } else {
        int idx = 0;  // first k with g[k] >= xi
        int lo = 0, hi = X - 1;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (g[mid] < xi) lo = mid + 1;
          else hi = mid;
        }
        idx = lo < 1 ? 1 : lo;
        double g0 = g[idx - 1], g1 = g[idx];
        double x0 = static_cast<double>(idx - 1) / (X - 1), x1 = static_cast<double>(idx) / (X - 1);
        double fr = (g1 > g0) ? (xi - g0) / (g1 - g0) : 0.5;
        out = x0 + fr * (x1 - x0);
      }
    }
  }
}

// T(y_i) for the leak sampler: the root of survival_series(T) = exp(-y_i)
// at y_i = exp(kLeakLogYMin + i h), by bisection in ln T (T grows with y, so
// each root brackets the next from below; S(1e-3) is 1 to double precision
// and S(kTauMax) ~ 1e-274, so [1e-3, kTauMax] holds every root).
inline void build_leak_table(double *leak) {
  double lo0 = log(1e-3);
  const double hi0 = log(kTauMax);
  for (int i = 0; i < kLeakSize; ++i) {
    const double ly = kLeakLogYMin + (kLeakLogYMax - kLeakLogYMin) * static_cast<double>(i) / (kLeakSize - 1);
    const double target = exp(-exp(ly));
    double lo = lo0, hi = hi0;
    for (int it = 0; it < 64 && hi - lo > 1e-15; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (survival_series(exp(mid)) > target) lo = mid;
      else hi = mid;
    }
    leak[i] = exp(0.5 * (lo + hi));
    lo0 = lo;
  }
}

// Fills `all` [kAllSize] with the three tables (tables_at's layout).
inline void build_all_tables(double *all) {
  double survival[kTableSize];
  build_leak_table(all);
  build_tables(all + kTauOffset, survival, all + kRadiusOffset);
}

}  // namespace rw
}  // namespace storm
