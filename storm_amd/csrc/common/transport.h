// Per-packet IMC transport on a Cartesian mesh in cell-local coordinates.
//
// One function, compiled for gfx950 (the transport kernels) and for the host
// (the bit-exact replay used by the tests), advances one packet through up to
// `max_segments` segments.  A segment is one call of the reference's
// RadiationIMC::step (radiation/RadiationIMC.hpp:1700-1881): the distance to
// the nearest cell face, to the next effective-scatter event and to census is
// compared, the packet moves, its weight is attenuated by the continuous
// absorption factor, the absorbed energy and the time-integrated radiation
// energy are tallied in the cell, and the event is applied (face crossing with
// the manager's cell move or boundary call, MonteCarloManagerSerial.hpp:318-357;
// isotropic redirection; census; roulette below 1e-3 of the step's starting
// weight, RadiationIMC.hpp:1812-1820).
//
// Layout decisions (MI355X): the position is the fraction of the cell width
// on each axis (u in [0,1]^3), so a face crossing sets one coordinate to 0 or
// 1 exactly instead of nudging by 1e-8 as the FP64 reference does; the
// direction is a unit vector; path lengths are in cm (time * c).  The Real
// type is float for the mixed engine (deterministic own math, fixed-point
// tallies through the Sink) and double for the plain FP64 engine.
#pragma once

#include <stdint.h>
#include <math.h>
#include "constants.h"
#include "fixed.h"
#include "fpmath.h"
#include "hd.h"
#include "philox.h"

namespace storm {

// Packed cell coordinates: i (12 bits) | j (10 bits) | k (10 bits).
STORM_HD uint32_t pack_ijk(uint32_t i, uint32_t j, uint32_t k) { return (i << 20) | (j << 10) | k; }
STORM_HD uint32_t ijk_i(uint32_t p) { return p >> 20; }
STORM_HD uint32_t ijk_j(uint32_t p) { return (p >> 10) & 0x3ffu; }
STORM_HD uint32_t ijk_k(uint32_t p) { return p & 0x3ffu; }

// Mesh as seen by the transport: per-axis widths (cm) and the fate of a
// packet at each box face (0 reflect, 1 escape), order x-, x+, y-, y+, z-, z+.
template <typename Real>
struct MeshView {
  int32_t nx, ny, nz;
  int32_t face_kind[6];
  const Real *width[3];  // width[a][idx] for a = x, y, z
};

template <typename Real>
STORM_HD uint32_t flat_cell(const MeshView<Real> &m, uint32_t ijk) {
  return (ijk_i(ijk) * static_cast<uint32_t>(m.ny) + ijk_j(ijk)) * static_cast<uint32_t>(m.nz) + ijk_k(ijk);
}

// Per-cell opacities used by a segment (cm^-1):
//   sig_abs = f * sigma_a            continuous absorption (weight attenuation)
//   sig_ev  = sigma_s + (1-f)*sigma_a  effective-scatter event opacity
template <typename Real>
struct CellXS {
  Real sig_abs, sig_ev;
};

template <typename Real>
struct Packet {
  Real u[3];      // position as a fraction of the cell width per axis
  Real s;         // path length left in the step, cm (= timeLeft * c)
  Real o[3];      // unit direction
  Real wmin;      // roulette threshold, 1e-3 of the weight at step start
  double w;       // weight, erg
  uint32_t ijk;   // packed cell coordinates
  uint32_t id;    // packet id within the step (keys the random stream)
  uint32_t ctr;   // Philox blocks consumed by transport so far
  uint32_t flags;
};

// Scalars shared by every packet of a step.
struct StepConsts {
  uint64_t seed;
  uint32_t step;
  int32_t eq;      // energy quantum exponent, Q = 2^eq erg
  int32_t et;      // path-integral quantum exponent, 2^et erg*cm
};

enum TransportOutcome : int { kAlive = 0, kCensus = 1, kGone = 2, kEscaped = 3 };

// ---------------------------------------------------------------- math policy
template <typename Real>
struct MathPolicy;

template <>
struct MathPolicy<float> {
  STORM_HD static float log(float x) { return fpm::logf_d(x); }
  STORM_HD static float expm1(float x) { return fpm::expm1f_d(x); }
  STORM_HD static void sincos2pi(float u, float &s, float &c) { fpm::sincos2pi(u, s, c); }
  STORM_HD static float sqrt(float x) { return ::sqrtf(x); }
  STORM_HD static float uniform(uint32_t x) { return u01_f32(x); }
  STORM_HD static float inf() { return fpm::u2f(0x7f800000u); }
};

template <>
struct MathPolicy<double> {
  STORM_HD static double log(double x) { return ::log(x); }
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

// ---------------------------------------------------------------- transport
// Sink interface (duck-typed):
//   visit(cell, dep_limbs_or_double, path_integral, segments)   cell visit ends
//   escape(cell, weight)                                         packet escapes
// For Real=float the energies are passed as fixed-point limbs (exact); for
// Real=double as plain doubles.
template <typename Real, typename Sink>
STORM_HD int transport_packet(Packet<Real> &p, const MeshView<Real> &m, const CellXS<Real> *xs,
                              const StepConsts &k, int max_segments, Sink &sink) {
  using M = MathPolicy<Real>;
  constexpr bool fixed_point = sizeof(Real) == 4;
  const PacketStream rs = PacketStream::make(k.seed, p.id, k.step, kTagTransport);

  uint32_t cell = flat_cell(m, p.ijk);
  double w_enter = p.w;      // weight when the current cell visit began
  double path_int = 0.0;     // sum of w * (1 - e^{-x}) / sig_abs over the visit, erg*cm
  uint32_t visit_segs = 0;
  int outcome = kAlive;

  auto flush = [&](double w_exit) {
    if constexpr (fixed_point) {
      fx::Limbs d = fx::sub(fx::limbs_of(w_enter, k.eq), fx::limbs_of(w_exit, k.eq));
      fx::Limbs e = fx::limbs_of(fx::quantize_nearest(path_int, k.et), k.et);
      sink.visit(cell, d, e, visit_segs);
    } else {
      sink.visit_fp(cell, w_enter - w_exit, path_int, visit_segs);
    }
    visit_segs = 0;
    path_int = 0.0;
  };

  for (int seg = 0; seg < max_segments; ++seg) {
    const uint32_t i = ijk_i(p.ijk), j = ijk_j(p.ijk), kk = ijk_k(p.ijk);
    const Real width[3] = {m.width[0][i], m.width[1][j], m.width[2][kk]};
    const CellXS<Real> x = xs[cell];
    ++visit_segs;

    // Random numbers of this segment: one Philox block.
    Philox4 r = rs.block(p.ctr++);

    // Distance to the nearest face along the direction.
    Real dface = M::inf(), dsecond = M::inf();
    int fax = -1, sax = -1;
    for (int a = 0; a < 3; ++a) {
      Real oa = p.o[a];
      if (oa != Real(0)) {
        Real frac = oa > Real(0) ? (Real(1) - p.u[a]) : p.u[a];
        Real d = frac * width[a] / (oa > Real(0) ? oa : -oa);
        if (d < dface) {
          dsecond = dface;
          sax = fax;
          dface = d;
          fax = a;
        } else if (d < dsecond) {
          dsecond = d;
          sax = a;
        }
      }
    }
    if constexpr (fixed_point) {
      // Two faces within 1e-6 of each other along the path (an edge or corner
      // crossing): decide which one is hit first in FP64.
      if (sax >= 0 && dsecond - dface <= Real(1e-6) * dsecond) {
        double df[2];
        int ax2[2] = {fax, sax};
        for (int q = 0; q < 2; ++q) {
          int a = ax2[q];
          double oa = static_cast<double>(p.o[a]);
          double fr = oa > 0.0 ? (1.0 - static_cast<double>(p.u[a])) : static_cast<double>(p.u[a]);
          df[q] = fr * static_cast<double>(width[a]) / (oa > 0.0 ? oa : -oa);
        }
        if (df[1] < df[0]) {
          fax = sax;
          dface = static_cast<Real>(df[1]);
        }
      }
    }
    // Distance to the next effective-scatter event (-ln xi / sigma_ev).
    Real xi = M::uniform(r.v[0]);
    Real dcol = x.sig_ev > Real(0) ? (-M::log(xi)) / x.sig_ev : M::inf();
    // Event: face < scatter < census on ties, as the reference's min_element.
    int ev = 0;
    Real d = dface;
    if (dcol < d) {
      ev = 1;
      d = dcol;
    }
    if (p.s < d) {
      ev = 2;
      d = p.s;
    }
    p.s -= d;
    if (ev == 2 || p.s < Real(0)) p.s = Real(0);

    // Continuous absorption over the segment.
    Real xabs = x.sig_abs * d;
    Real frac = -M::expm1(-xabs);
    double w0 = p.w;
    if (xabs >= Real(1e-12)) {
      path_int += w0 * static_cast<double>(frac / x.sig_abs);
    } else {
      path_int += w0 * static_cast<double>(d);
    }
    p.w = w0 - w0 * static_cast<double>(frac);

    // Move (the face axis is set exactly below).
    for (int a = 0; a < 3; ++a) {
      if (ev == 0 && a == fax) continue;
      Real nu = p.u[a] + p.o[a] * d / width[a];
      p.u[a] = nu < Real(0) ? Real(0) : (nu > Real(1) ? Real(1) : nu);
    }

    // Roulette (RadiationIMC.hpp:1812-1820): the remaining weight goes to the material.
    if (p.w < static_cast<double>(p.wmin)) {
      flush(0.0);
      p.w = 0.0;
      outcome = kGone;
      break;
    }

    if (ev == 0) {
      const bool up = p.o[fax] > Real(0);
      uint32_t idx = fax == 0 ? i : (fax == 1 ? j : kk);
      uint32_t n = static_cast<uint32_t>(fax == 0 ? m.nx : (fax == 1 ? m.ny : m.nz));
      bool outside = up ? (idx + 1u >= n) : (idx == 0u);
      if (!outside) {
        // interior face: the packet enters the neighbour on the opposite face
        double wq = fixed_point ? fx::quantize_down(p.w, k.eq) : p.w;
        flush(wq);
        p.w = wq;
        w_enter = wq;
        uint32_t ni = i, nj = j, nk = kk;
        uint32_t nidx = up ? idx + 1u : idx - 1u;
        if (fax == 0) ni = nidx; else if (fax == 1) nj = nidx; else nk = nidx;
        p.ijk = pack_ijk(ni, nj, nk);
        p.u[fax] = up ? Real(0) : Real(1);
        cell = flat_cell(m, p.ijk);
      } else {
        int face = 2 * fax + (up ? 1 : 0);
        p.u[fax] = up ? Real(1) : Real(0);
        if (m.face_kind[face] == 1) {
          // escape (REMOVE); the weight leaves the problem
          double wq = fixed_point ? fx::quantize_down(p.w, k.eq) : p.w;
          flush(wq);
          sink.escape(cell, wq);
          p.w = 0.0;
          outcome = kEscaped;
          break;
        }
        // specular reflection on the box face (BoundaryCondition.hpp:87-108)
        p.o[fax] = -p.o[fax];
      }
    } else if (ev == 1) {
      // effective scatter: isotropic redirection (RadiationIMC.hpp:1827-1874)
      isotropic_dir<Real>(M::uniform(r.v[2]), M::uniform(r.v[3]), p.o);
    } else {
      // census: the weight is kept as a multiple of the quantum in both
      // engines so that population control can work in exact integers
      double wq = fx::quantize_down(p.w, k.eq);
      flush(wq);
      p.w = wq;
      outcome = kCensus;
      break;
    }
  }
  if (outcome == kAlive) {
    // segment budget used up: close the visit so the next pass starts clean
    double wq = fixed_point ? fx::quantize_down(p.w, k.eq) : p.w;
    flush(wq);
    p.w = wq;
  }
  return outcome;
}

}  // namespace storm
