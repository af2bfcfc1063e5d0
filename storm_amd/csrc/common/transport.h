// Per-packet IMC transport on a Cartesian mesh in cell-local coordinates.
//
// One function, compiled for gfx950 (the transport kernels) and for the host
// (the bit-exact replay used by the tests), advances one packet through up to
// `max_segments` segments.  A segment is one call of the reference's
// RadiationIMC::step (radiation/RadiationIMC.hpp:1700-1881), counted like
// cellsStepsCounters (MonteCarloManagerSerial.hpp:302-303):
//   * in a random-walk cell, a random-walk step if the packet is deep enough
//     (tryRandomWalkStep, :1065-1276);
//   * else in a DDMC cell, a discrete-diffusion step (tryDDMCStep, :1415-1612);
//   * else analytic tracking: the distance to the nearest cell face, to the
//     next effective-scatter event and to census is compared, the packet
//     moves, its weight is attenuated by the continuous absorption factor, the
//     absorbed energy and the time-integrated radiation energy are tallied in
//     the cell, and the event is applied (face crossing with the manager's
//     cell move or boundary call, MonteCarloManagerSerial.hpp:318-357;
//     isotropic redirection; census).  Roulette below 1e-3 of the step's
//     starting weight (1e-4 in a random-walk step) deposits the rest.
//
// Layout decisions (MI355X): the position is the fraction of the cell width
// on each axis (u in [0,1]^3), so a face crossing sets one coordinate to 0 or
// 1 exactly instead of nudging by 1e-8 as the FP64 reference does; the
// direction is a unit vector; times are carried as path lengths in cm
// (time * c).  The Real type is float for the mixed engine (deterministic own
// math, fixed-point tallies) and double for the plain FP64 engine.
#pragma once

#include <stdint.h>
#include <math.h>
#include "constants.h"
#include "fixed.h"
#include "fpmath.h"
#include "hd.h"
#include "mathpolicy.h"
#include "philox.h"

namespace storm {

// Packed cell coordinates: i (12 bits) | j (10 bits) | k (10 bits).
STORM_HD uint32_t pack_ijk(uint32_t i, uint32_t j, uint32_t k) { return (i << 20) | (j << 10) | k; }
STORM_HD uint32_t ijk_i(uint32_t p) { return p >> 20; }
STORM_HD uint32_t ijk_j(uint32_t p) { return (p >> 10) & 0x3ffu; }
STORM_HD uint32_t ijk_k(uint32_t p) { return p & 0x3ffu; }

// Multigroup spectrum as seen by the transport (groups < 2: gray).  Photon
// energies are in keV.  A packet of energy E in cell c sees
//   sigma_a(c, E) = sigma_a(c, 1 keV) * max(E, e_lo)^(-nu_pow)
// (the cell table then holds f sigma_a(c, 1 keV) and (1-f) sigma_a(c, 1 keV))
// and re-emits after an effective scatter at an energy drawn from the cell's
// emission CDF: cdf[c*(groups+1) + g] over the group bounds, linear within a
// group (DensmoreOpacity::GetThermalEnergy, examples/densmore2012/DensmoreOpacity.hpp:66-77).
template <typename Real>
struct SpectrumView {
  int32_t groups;
  Real nu_pow;
  Real e_lo, e_hi;     // bounds[0], bounds[groups]
  const Real *bounds;  // groups + 1
  const Real *cdf;     // per cell, groups + 1, cdf[0] = 0 and cdf[groups] = 1
  const Real *sig_s;   // per cell scattering opacity (gray)
  int32_t eg;          // 1: tally the path integral per cell and group (withEgTimeAvg)
};

// RadiationOpacityModel::findGroup (radiation/RadiationOpacityModel.hpp:76-86):
// the first group whose upper bound lies above E, else the last.
template <typename Real>
STORM_HD int32_t group_of(Real e, const Real *bounds, int32_t groups) {
  int32_t lo = 0, hi = groups - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (e < bounds[mid + 1]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// The draw for the unbiased rounding of a per-group path integral: a hash of
// the packet, its stream position, the cell and the group.
STORM_HD uint32_t group_tally_draw(uint32_t id, uint32_t ctr, uint32_t cell, int32_t g) {
  uint32_t h = id * 0x85EBCA6Bu ^ (ctr + 0x2F1B9A3Du) * 0xC2B2AE35u ^ cell * 0x27D4EB2Du ^
               static_cast<uint32_t>(g) * 0x165667B1u;
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  return h >> 8;
}

// Energy (keV) at the quantile `u` of a normalised group CDF, linear within
// the group (LinearInterpolation, utils/LinearInterpolation.hpp:17-42, of
// the CDF onto the bounds).
template <typename Real>
STORM_HD Real spectrum_quantile(const Real *cdf, const Real *bounds, int32_t groups, Real u) {
  // smallest i in [1, groups] with cdf[i] > u
  int32_t lo = 0, hi = groups;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (cdf[mid] > u) hi = mid;
    else lo = mid;
  }
  const Real c0 = cdf[hi - 1], c1 = cdf[hi];
  const Real b0 = bounds[hi - 1], b1 = bounds[hi];
  const Real t = c1 > c0 ? (u - c0) / (c1 - c0) : Real(0.5);
  const Real e = b0 + t * (b1 - b0);
  return e < b0 ? b0 : (e > b1 ? b1 : e);
}

// Mesh as seen by the transport: per-axis widths (cm), the fate of a packet
// at each box face (order x-, x+, y-, y+, z-, z+) and the spectrum.
template <typename Real>
struct MeshView {
  int32_t nx, ny, nz;
  uint32_t escape_mask;  // bit f set: box face f removes the packet (else it reflects)
  // the cell widths and their reciprocals, interleaved (width_pairs), axis
  // x then y then z: index i along axis a at width_at(m, a, i)
  const Real *width;
  SpectrumView<Real> spec;
};

// (width, 1 / width) of index i along axis a
template <typename Real>
STORM_HD const Real *width_at(const MeshView<Real> &m, int a, uint32_t i) {
  const uint32_t off = a == 0 ? 0u : (a == 1 ? static_cast<uint32_t>(m.nx) : static_cast<uint32_t>(m.nx + m.ny));
  return m.width + 2u * (off + i);
}

// Host: widths w[0..n) and their reciprocals, interleaved into out[0..2n).
// The reciprocal is the one IEEE division a segment would otherwise do on
// entering the cell, so the tracking is unchanged bit for bit.
template <typename Real>
inline void width_pairs(const Real *w, int n, Real *out) {
  for (int i = 0; i < n; ++i) {
    out[2 * i] = w[i];
    out[2 * i + 1] = Real(1) / w[i];
  }
}

template <typename Real>
STORM_HD uint32_t flat_cell(const MeshView<Real> &m, uint32_t ijk) {
  return (ijk_i(ijk) * static_cast<uint32_t>(m.ny) + ijk_j(ijk)) * static_cast<uint32_t>(m.nz) + ijk_k(ijk);
}

// Per-cell opacities used by a segment (cm^-1):
//   sig_abs = f * sigma_a              continuous absorption (weight attenuation)
//   sig_ev  = sigma_s + (1-f)*sigma_a  effective-scatter event opacity
//   inv_ev  = 1 / sig_ev (0 where sig_ev is 0), for the event distance
template <typename Real>
struct CellXS {
  Real sig_abs, sig_ev, inv_ev;
};

template <typename Real>
STORM_HD CellXS<Real> make_xs(Real sig_abs, Real sig_ev) {
  CellXS<Real> x;
  x.sig_abs = sig_abs;
  x.sig_ev = sig_ev;
  x.inv_ev = sig_ev > Real(0) ? Real(1) / sig_ev : Real(0);
  return x;
}

// Per-cell data of the diffusion accelerations:
//   flags   bit 0 random-walk cell, bit 1 DDMC cell
//   sig_t   sigma_a + sigma_s (cm^-1); multigroup: the Planck mean of
//           sigma_a(E) + sigma_s over the diffusive groups (the walk's
//           particle threshold opacity)
//   rate    DDMC leak rate per face / c (cm^-1), order x-, x+, y-, y+, z-, z+
//   total   sum of rate
//   gcut    multigroup: the diffusive groups are 0 .. gcut-1
//   sig_d   opacity of the walk's diffusion coefficient, D = c / (3 sig_d)
//           (gray: sig_t; multigroup: the Planck-weighted harmonic mean of
//           sigma_a(E) + sigma_s over the diffusive groups)
//   sig_abs multigroup: f times the Planck mean of sigma_a(E) over the
//           diffusive groups, the continuous absorption during a walk
//   up_rate multigroup: (1 - f) sigma_a,mean (1 - gamma) (cm^-1), the rate of
//           re-emission above the diffusive groups, which ends a walk
//   e_cut   multigroup: lower bound (keV) of group gcut; packets at or above
//           it do not walk (+inf gray)
//   ddmc_sig_t  DDMC: sigma_P + sigma_s, the Planck-mean total opacity (the
//           particle threshold's opacity; equal to sig_t in a gray problem)
//   ddmc_abs    DDMC in a multigroup problem: f sigma_P, the continuous
//           absorption during a DDMC step (a gray step uses the cell's own);
//           with ddmc_multigroup_pgrw f sigma_A of the diffusive groups
//   ddmc_e_cut  DDMC with ddmc_multigroup_pgrw: lower bound (keV) of the
//           first non-diffusive group; packets at or above it take no DDMC
//           step (+inf otherwise)
// (precomputeRandomWalkData, radiation/RadiationIMC.hpp:983-1063;
// precomputeDDMCData without ddmcUseMultigroupPGRW, :1282-1351)
template <typename Real>
struct CellAccel {
  Real sig_t;
  Real total;
  Real rate[6];
  uint32_t flags;
  uint32_t gcut;
  Real sig_d;
  Real sig_abs;
  Real up_rate;
  Real e_cut;
  Real ddmc_sig_t;
  Real ddmc_abs;
  Real ddmc_e_cut;
};

// Random-walk tables (see random_walk.h) and the particle thresholds.
template <typename Real>
struct AccelView {
  const CellAccel<Real> *cell;  // null: no random walk and no DDMC in this problem
  const Real *rw;           // [rw::kAllSize]: the tables (rw::tables_at)
  Real rw_particle_tau;    // rwMinParticleOpticalDepth
  Real ddmc_particle_tau;  // ddmcMinParticleOpticalDepth
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
  Real nu;        // photon energy, keV (multigroup; 0 when gray)
};

// Scalars shared by every packet of a step.
struct StepConsts {
  uint64_t seed;
  uint32_t step;
  int32_t eq;      // energy quantum exponent, Q = 2^eq erg
  int32_t et;      // path-integral quantum exponent, 2^et erg*cm
};

enum TransportOutcome : int { kAlive = 0, kCensus = 1, kGone = 2, kEscaped = 3 };

}  // namespace storm

#include "random_walk.h"

namespace storm {

// ---------------------------------------------------------------- transport
// Component access by a runtime axis without indexing a local array (a
// runtime index into a register array would put it in scratch memory).
template <typename T>
STORM_HD T sel3(int a, T x, T y, T z) {
  return a == 0 ? x : (a == 1 ? y : z);
}
template <typename T>
STORM_HD void put3(int a, T &x, T &y, T &z, T v) {
  x = a == 0 ? v : x;
  y = a == 1 ? v : y;
  z = a == 2 ? v : z;
}

// Cell-local coordinates are fractions of the cell width.  A float moved by
// less than half a unit in its last place keeps its value, and near u = 1
// that unit is 2^-24 of the width: in an optically thick cell a collision
// flight is far shorter.  A packet lying on a face (after a reflection, say)
// would then stay pinned there, at distance 0 from the face, never eligible
// for the random walk, and with its time budget no longer decreasing either
// once a flight is below the budget's own unit: a history without end (the
// FP64 reference pins the same way once a flight is below the unit of its
// absolute coordinates).  So a packet on a face that moves inward by less
// than a unit is taken one unit inside (the smallest normal number above 0
// on the lower face), far below the cell size but enough for the random walk
// to take over in such a cell; elsewhere a sub-unit move of `delta` is taken
// as one whole unit with probability |delta| / unit (from 8 random bits
// `r8`), so that the expected move is exact.
STORM_HD float move_coord(float u, float n, float delta, uint32_t r8) {
  if (n != u || delta == 0.0f) return n;
  if (delta > 0.0f ? !(u < 1.0f) : !(u > 0.0f)) return n;
  if (u == 0.0f) return 1.17549435e-38f;
  if (u == 1.0f) return 0.99999994f;
  uint32_t b;
  __builtin_memcpy(&b, &u, sizeof(b));
  b = delta > 0.0f ? b + 1u : b - 1u;
  float next;
  __builtin_memcpy(&next, &b, sizeof(next));
  const float unit = delta > 0.0f ? next - u : u - next;
  const float p = (delta > 0.0f ? delta : -delta) / unit;
  return (static_cast<float>(r8) + 0.5f) * 0.00390625f < p ? next : u;
}

// The next representable number above a positive x.
STORM_HD float next_above(float x) {
  uint32_t b;
  __builtin_memcpy(&b, &x, sizeof(b));
  ++b;
  __builtin_memcpy(&x, &b, sizeof(x));
  return x;
}
STORM_HD double next_above(double x) {
  uint64_t b;
  __builtin_memcpy(&b, &x, sizeof(b));
  ++b;
  __builtin_memcpy(&x, &b, sizeof(x));
  return x;
}

// Sink interface (duck-typed):
//   visit(cell, dep_limbs, path_limbs, segments)  (Real = float) or
//   visit_fp(cell, dep, path, segments)           (Real = double): a cell visit ends
//   roulette(cell, weight)                          Real = float: weight rouletted into the material
//   escape_fx(cell, weight) / escape_fp(...)        the packet leaves the problem (mixed / FP64)
//   rw_step() / ddmc_step()                         acceleration counters
// The mixed engine tallies the path integral only in cells without
// absorption: elsewhere it equals the absorbed energy divided by f sigma_a
// (both are sums of w (1 - e^{-x}) over the same segments), which the
// postStep evaluates from the exact deposit tally.
//
// Tracker<Real, F> holds one packet's history between segments, so that a
// kernel can advance a packet one segment at a time and give an idle lane a
// new packet (transport_packet below runs a whole pass of one packet).
// F selects the code compiled in: kFeatAccel (random walk and DDMC) and
// kFeatMultigroup; a problem without them runs a leaner kernel with the same
// results.
enum : uint32_t { kFeatAccel = 1u, kFeatMultigroup = 2u, kFeatAll = 3u };

template <typename Real, uint32_t F = kFeatAll>
struct Tracker {
  using M = MathPolicy<Real>;
  static constexpr bool fixed_point = sizeof(Real) == 4;

  PacketStream rs;
  Real u0, u1, u2, o0, o1, o2, s, wmin, nu;
  Real io0, io1, io2;  // 1/|o| per axis (+inf along an axis the packet does not move on)
  // the current cell's data, loaded when the packet enters it (and, in
  // multigroup, when its energy changes); reciprocals so that a segment
  // multiplies instead of dividing
  Real cell_sig_abs, cell_sig_ev, cell_sig_s, inv_sig_ev, cw0, cw1, cw2, icw0, icw1, icw2;
  double w, w_enter, path_int;  // w_enter: weight when the current cell visit began
  uint32_t ijk, ctr, cell, visit_segs, id;
  int outcome;
  bool derive_pint;
  // per-group tally (multigroup with withEgTimeAvg): the packet's group and
  // the part of the visit's path integral already handed to it
  bool eg_track;
  int32_t grp;
  double eg_mark;

  STORM_HD void begin(const Packet<Real> &p, const MeshView<Real> &m, const CellXS<Real> *xs, const StepConsts &k) {
    rs = PacketStream::make(k.seed, p.id, k.step, kTagTransport);
    u0 = p.u[0];
    u1 = p.u[1];
    u2 = p.u[2];
    set_dir(p.o[0], p.o[1], p.o[2]);
    s = p.s;
    w = p.w;
    ijk = p.ijk;
    ctr = p.ctr;
    wmin = p.wmin;
    id = p.id;
    nu = p.nu;
    eg_track = (F & kFeatMultigroup) != 0u && m.spec.groups >= 2 && m.spec.eg != 0;
    grp = 0;
    eg_mark = 0.0;
    load_cell(m, xs);
    w_enter = w;
    path_int = 0.0;  // sum of w * (1 - e^{-x}) / sig_abs over the visit, erg*cm
    visit_segs = 0;
    outcome = kAlive;
  }

  // Writes the state back; a packet still alive closes its visit so that the
  // next pass starts clean.
  template <typename Sink>
  STORM_HD void end(Packet<Real> &p, const StepConsts &k, Sink &sink) {
    if (outcome == kAlive) {
      double wq = fixed_point ? fx::carry_round(w, w_enter, k.eq, id, ctr, cell) : w;
      flush(wq, k, sink);
      w = wq;
    }
    p.u[0] = u0;
    p.u[1] = u1;
    p.u[2] = u2;
    p.o[0] = o0;
    p.o[1] = o1;
    p.o[2] = o2;
    p.s = s;
    p.w = w;
    p.wmin = wmin;
    p.ijk = ijk;
    p.ctr = ctr;
    p.id = id;
    p.nu = nu;
  }

  // The mixed engine tallies the path integral only where postStep cannot
  // derive it from the deposit (gray cells without absorption).
  STORM_HD bool derives_path_integral(const MeshView<Real> &m) const {
    if constexpr ((F & kFeatMultigroup) != 0u) {
      if (m.spec.groups >= 2) return false;
    }
    return fixed_point && cell_sig_abs > Real(0);
  }

  // Hands the path integral since the last call to the packet's group
  // (RadiationIMC.hpp:1166-1173, 1485-1492, 1802-1809): at the end of each
  // visit and before the photon energy changes.
  template <typename Sink>
  STORM_HD void group_split(const StepConsts &k, Sink &sink) {
    if constexpr ((F & kFeatMultigroup) != 0u) {
      if (eg_track) {
        const double v = path_int - eg_mark;
        if (v > 0.0) sink.group_path(cell, grp, v, k, group_tally_draw(id, ctr, cell, grp));
        eg_mark = path_int;
      }
    }
  }

  template <typename Sink>
  STORM_HD void flush(double w_exit, const StepConsts &k, Sink &sink) {
    group_split(k, sink);
    eg_mark = 0.0;
    if constexpr (fixed_point) {
      fx::Limbs d = fx::sub(fx::limbs_of(w_enter, k.eq), fx::limbs_of(w_exit, k.eq));
      fx::Limbs e{0u, 0u};
      if (!derive_pint) {
        // the rounding draw: a hash of the packet, its stream position and
        // the cell (no random-stream draw, so the history is unchanged)
        uint32_t h = id * 0x9E3779B1u ^ (ctr + 0x7F4A7C15u) * 0x85EBCA77u ^ cell * 0xC2B2AE3Du;
        h ^= h >> 15;
        h *= 0x2C1B3C6Du;
        h ^= h >> 12;
        e = fx::limbs_of(fx::quantize_unbiased(path_int, k.et, h >> 8), k.et);
      }
      sink.visit(cell, d, e, visit_segs);
    } else {
      sink.visit_fp(cell, w_enter - w_exit, path_int, visit_segs);
    }
    visit_segs = 0;
    path_int = 0.0;
  }

  // close the visit and move into the neighbour across face `fax`/`up`
  template <typename Sink>
  STORM_HD void enter_neighbour(int fax, bool up, const MeshView<Real> &m, const CellXS<Real> *xs,
                                const StepConsts &k, Sink &sink) {
    double wq = fixed_point ? fx::carry_round(w, w_enter, k.eq, id, ctr, cell) : w;
    flush(wq, k, sink);
    w = wq;
    w_enter = wq;
    uint32_t ni = ijk_i(ijk), nj = ijk_j(ijk), nk = ijk_k(ijk);
    if (fax == 0) ni = up ? ni + 1u : ni - 1u;
    else if (fax == 1) nj = up ? nj + 1u : nj - 1u;
    else nk = up ? nk + 1u : nk - 1u;
    ijk = pack_ijk(ni, nj, nk);
    put3<Real>(fax, u0, u1, u2, up ? Real(0) : Real(1));
    load_cell(m, xs);
  }

  // The same move with the neighbour's data already loaded (the analytic
  // step loads it before the segment's arithmetic, to overlap the latency):
  // `nijk`/`ncell` its index, `nw` its width along `fax`, `nx`/`nsig_s` its
  // opacities.
  template <typename Sink>
  STORM_HD void enter_loaded_neighbour(int fax, bool up, uint32_t nijk, uint32_t ncell, Real nw, Real niw,
                                       const CellXS<Real> &nx, Real nsig_s, const MeshView<Real> &m,
                                       const StepConsts &k, Sink &sink) {
    double wq = fixed_point ? fx::carry_round(w, w_enter, k.eq, id, ctr, cell) : w;
    flush(wq, k, sink);
    w = wq;
    w_enter = wq;
    ijk = nijk;
    cell = ncell;
    put3<Real>(fax, u0, u1, u2, up ? Real(0) : Real(1));
    put3<Real>(fax, cw0, cw1, cw2, nw);
    put3<Real>(fax, icw0, icw1, icw2, niw);
    set_opacity(m, nx, nsig_s);
  }

  STORM_HD void load_cell(const MeshView<Real> &m, const CellXS<Real> *xs) {
    cell = flat_cell(m, ijk);
    const Real *p0 = width_at(m, 0, ijk_i(ijk));
    const Real *p1 = width_at(m, 1, ijk_j(ijk));
    const Real *p2 = width_at(m, 2, ijk_k(ijk));
    cw0 = p0[0];
    icw0 = p0[1];
    cw1 = p1[0];
    icw1 = p1[1];
    cw2 = p2[0];
    icw2 = p2[1];
    load_opacity(m, xs);
  }

  STORM_HD void set_dir(Real a, Real b, Real c) {
    o0 = a;
    o1 = b;
    o2 = c;
    io0 = a != Real(0) ? Real(1) / (a > Real(0) ? a : -a) : M::inf();
    io1 = b != Real(0) ? Real(1) / (b > Real(0) ? b : -b) : M::inf();
    io2 = c != Real(0) ? Real(1) / (c > Real(0) ? c : -c) : M::inf();
  }

  // opacities of the current cell at the packet's energy
  STORM_HD void load_opacity(const MeshView<Real> &m, const CellXS<Real> *xs) {
    const bool mg = (F & kFeatMultigroup) != 0u && m.spec.groups >= 2;
    set_opacity(m, xs[cell], mg ? m.spec.sig_s[cell] : Real(0));
  }
  // from the cell's gray opacities `x` and scattering opacity `sig_s`
  // (multigroup only)
  STORM_HD void set_opacity(const MeshView<Real> &m, const CellXS<Real> &x, Real sig_s) {
    cell_sig_abs = x.sig_abs;
    cell_sig_ev = x.sig_ev;
    cell_sig_s = Real(0);
    inv_sig_ev = x.inv_ev;
    if ((F & kFeatMultigroup) != 0u && m.spec.groups >= 2) {
      // clampFrequencyToBounds (RadiationIMC.hpp:868-872) and E^-nu_pow
      Real e = nu < m.spec.e_lo ? m.spec.e_lo : (nu > m.spec.e_hi ? m.spec.e_hi : nu);
      const Real fac = M::exp(-m.spec.nu_pow * M::log(e));
      cell_sig_s = sig_s;
      cell_sig_abs = x.sig_abs * fac;
      cell_sig_ev = cell_sig_s + x.sig_ev * fac;
      if (eg_track) grp = group_of<Real>(e, m.spec.bounds, m.spec.groups);
      inv_sig_ev = cell_sig_ev > Real(0) ? Real(1) / cell_sig_ev : Real(0);
    }
    derive_pint = derives_path_integral(m);
  }

  // continuous absorption over a path d (cm); `tally_free` adds w * d to the
  // path integral when nothing is absorbed (the analytic step and DDMC do,
  // the random-walk step does not, RadiationIMC.hpp:1163-1165)
  STORM_HD void absorb(Real sig_abs, Real d, bool tally_free) {
    Real xabs = sig_abs * d;
    Real frac = -M::expm1(-xabs);
    double w0 = w;
    w = w0 - w0 * static_cast<double>(frac);
    // the mixed engine derives the path integral of absorbing gray cells
    // from their deposit (see flush)
    if (derive_pint) return;
    if (xabs >= Real(1e-12)) {
      path_int += w0 * static_cast<double>(frac / sig_abs);
    } else if (tally_free) {
      path_int += w0 * static_cast<double>(d);
    } else if (sig_abs > Real(0)) {
      path_int += w0 * static_cast<double>(frac / sig_abs);
    }
  }

  // roulette: the remaining weight goes to the material (RadiationIMC.hpp:1812-1820)
  template <typename Sink>
  STORM_HD void roulette(const StepConsts &k, Sink &sink) {
    double wq = fixed_point ? fx::quantize_down(w, k.eq) : w;
    flush(fixed_point ? wq : 0.0, k, sink);
    if constexpr (fixed_point) sink.roulette(cell, wq);
    w = 0.0;
    outcome = kGone;
  }

  template <typename Sink>
  STORM_HD void census(const StepConsts &k, Sink &sink) {
    // the weight is kept as a multiple of the quantum in both engines so
    // that population control can work in exact integers
    double wq = fx::carry_round(w, w_enter, k.eq, id, ctr, cell);
    flush(wq, k, sink);
    w = wq;
    outcome = kCensus;
  }

  // One segment of the history (a random-walk step, a DDMC step or an
  // analytic step to the next face, scatter or census event); returns the
  // outcome, kAlive while the packet goes on.
  template <typename Sink>
  STORM_HD int segment(const MeshView<Real> &m, const CellXS<Real> *xs, const AccelView<Real> &acc,
                       const StepConsts &k, Sink &sink) {
    const uint32_t i = ijk_i(ijk), j = ijk_j(ijk), kk = ijk_k(ijk);
    const Real w0w = cw0, w1w = cw1, w2w = cw2;
    const Real sig_abs = cell_sig_abs;
    ++visit_segs;

    // Random numbers of a segment: one Philox block, block ctr (two for a
    // random-walk step, three for a multigroup re-emission), drawn in each
    // branch below after the loads it does not depend on are issued.
    if ((F & kFeatAccel) != 0u && acc.cell) {
      const CellAccel<Real> ca = acc.cell[cell];
      if (ca.flags) {
        // distance to the nearest face (computeMinDistanceToFaces, :956-969)
        Real Ro;
        {
          Real a0 = u0 * w0w, b0 = (Real(1) - u0) * w0w;
          Real a1 = u1 * w1w, b1 = (Real(1) - u1) * w1w;
          Real a2 = u2 * w2w, b2 = (Real(1) - u2) * w2w;
          Real m0 = a0 < b0 ? a0 : b0, m1 = a1 < b1 ? a1 : b1, m2 = a2 < b2 ? a2 : b2;
          Ro = m0 < m1 ? m0 : m1;
          Ro = m2 < Ro ? m2 : Ro;
        }
        const Real tauR = Ro * ca.sig_t;
        // multigroup: the partial-gray walk, for packets in the diffusive groups only (:1094-1114)
        const bool pgrw = (F & kFeatMultigroup) != 0u && m.spec.groups >= 2;
        const bool walks = !pgrw || nu < ca.e_cut;
        if ((ca.flags & 1u) && walks && Ro > Real(0) && ca.sig_t > Real(0) && tauR >= acc.rw_particle_tau) {
          // ---------------- random-walk step (tryRandomWalkStep)
          const Philox4 r = rs.block(ctr++);
          const Philox4 r2 = rs.block(ctr++);
          const rw::Tables<Real> tb = rw::tables_at(acc.rw);
          const Real tauLeak = rw::sample_leak_time<Real>(tb, M::uniform(r.v[0]));
          // tLeak = tau Ro^2 / D with D = c / (3 sig_d): as a path, tau Ro^2 3 sig_d
          const Real sig_d = pgrw ? ca.sig_d : ca.sig_t;
          const Real sLeak = tauLeak * Ro * Ro * (Real(3) * sig_d);
          // multigroup: re-emission above the diffusive groups (:1126-1131)
          Real sUp = M::inf();
          if (pgrw && ca.up_rate > Real(0)) sUp = (-M::log(M::uniform(r2.v[2]))) / ca.up_rate;
          // leak, census, upscatter; ties in that order (:1133-1152)
          int rev = 0;
          Real d = sLeak;
          if (!(sLeak <= s && sLeak <= sUp)) {
            rev = s <= sUp ? 1 : 2;
            d = rev == 1 ? s : sUp;
          }
          const bool leak = rev == 0;
          absorb(pgrw ? ca.sig_abs : sig_abs, d, false);
          s = rev == 1 ? Real(0) : s - d;
          if (s < Real(0)) s = Real(0);
          sink.rw_step();
          if (w < static_cast<double>(wmin * Real(0.1))) {
            roulette(k, sink);
            return outcome;
          }
          // new position on the sphere (leak) or inside it (census)
          Real dir[3];
          isotropic_dir<Real>(M::uniform(r.v[1]), M::uniform(r.v[2]), dir);
          Real disp = Ro;
          if (!leak) {
            const Real tauPos = d / (Real(3) * sig_d * Ro * Ro);
            disp = Ro * rw::sample_radius<Real>(tb, tauPos, M::uniform(r.v[3]));
          }
          if (disp > Ro) disp = Ro;
          const Real c0 = u0, c1 = u1, c2 = u2;
          // nudge toward the cell centre (RadiationIMC.hpp:1214-1215)
          Real n0 = (c0 + disp * dir[0] * icw0) * Real(1 - 1e-6) + Real(1e-6) * Real(0.5);
          Real n1 = (c1 + disp * dir[1] * icw1) * Real(1 - 1e-6) + Real(1e-6) * Real(0.5);
          Real n2 = (c2 + disp * dir[2] * icw2) * Real(1 - 1e-6) + Real(1e-6) * Real(0.5);
          for (int guard = 0; guard < 2000; ++guard) {
            bool inside = n0 >= Real(0) && n0 <= Real(1) && n1 >= Real(0) && n1 <= Real(1) && n2 >= Real(0) &&
                          n2 <= Real(1);
            if (inside) break;
            disp = disp * Real(0.99);
            n0 = c0 + disp * dir[0] * icw0;
            n1 = c1 + disp * dir[1] * icw1;
            n2 = c2 + disp * dir[2] * icw2;
          }
          u0 = n0 < Real(0) ? Real(0) : (n0 > Real(1) ? Real(1) : n0);
          u1 = n1 < Real(0) ? Real(0) : (n1 > Real(1) ? Real(1) : n1);
          u2 = n2 < Real(0) ? Real(0) : (n2 > Real(1) ? Real(1) : n2);
          Real nd[3];
          isotropic_dir<Real>(M::uniform(r2.v[0]), M::uniform(r2.v[1]), nd);
          set_dir(nd[0], nd[1], nd[2]);
          if ((F & kFeatMultigroup) != 0u && rev == 2) {
            // a new energy from the cell's emission spectrum above the
            // diffusive groups (:1237-1252)
            const int32_t G = m.spec.groups;
            const Real *cdfc = m.spec.cdf + static_cast<size_t>(cell) * static_cast<size_t>(G + 1);
            const Real lo = cdfc[ca.gcut];
            group_split(k, sink);
            if (lo < Real(1)) {
              nu = spectrum_quantile<Real>(cdfc, m.spec.bounds, G, lo + M::uniform(r2.v[3]) * (Real(1) - lo));
            } else {
              nu = next_above(m.spec.bounds[ca.gcut]);
            }
            nu = nu < m.spec.e_lo ? m.spec.e_lo : (nu > m.spec.e_hi ? m.spec.e_hi : nu);
            load_opacity(m, xs);
          }
          if (rev == 1) census(k, sink);
          return outcome;
        }
        if ((ca.flags & 2u) && ca.total > Real(0) && Ro * ca.ddmc_sig_t >= acc.ddmc_particle_tau &&
            ((F & kFeatMultigroup) == 0u || nu < ca.ddmc_e_cut)) {
          // ---------------- DDMC step (tryDDMCStep; in a multigroup problem
          // with the Planck-mean data and the photon energy kept, as the
          // reference does without ddmcUseMultigroupPGRW)
          const Philox4 r = rs.block(ctr++);
          const Real sEvent = (-M::log(M::uniform(r.v[0]))) / ca.total;
          const bool cen = s <= sEvent;
          const Real d = cen ? s : sEvent;
          absorb((F & kFeatMultigroup) != 0u ? ca.ddmc_abs : sig_abs, d, true);
          s = cen ? Real(0) : s - d;
          if (s < Real(0)) s = Real(0);
          sink.ddmc_step();
          if (w < static_cast<double>(wmin)) {
            roulette(k, sink);
            return outcome;
          }
          if (cen) {
            census(k, sink);
            return outcome;
          }
          // leak through a face chosen in proportion to its rate
          Real pick = M::uniform(r.v[1]) * ca.total;
          int face = -1, lastf = -1;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
          for (int f = 0; f < 6; ++f) {
            if (ca.rate[f] > Real(0)) {
              lastf = f;
              pick -= ca.rate[f];
              if (face < 0 && pick <= Real(0)) face = f;
            }
          }
          if (face < 0) face = lastf;
          const int fax = face >> 1;
          const bool up = (face & 1) != 0;
          // direction: mu = sqrt(U) about the outward normal, basis as in :1565-1586
          const Real mu = M::sqrt(M::uniform(r.v[2]));
          const Real st2 = Real(1) - mu * mu;
          const Real st = M::sqrt(st2 > Real(0) ? st2 : Real(0));
          Real sp, cp;
          M::sincos2pi(M::uniform(r.v[3]), sp, cp);
          const Real sg = up ? Real(1) : Real(-1);
          if (fax == 0) {          // e1 = y, e2 = n x y = sg z
            set_dir(sg * mu, st * cp, sg * st * sp);
          } else if (fax == 1) {   // e1 = x, e2 = n x x = -sg z
            set_dir(st * cp, sg * mu, -sg * st * sp);
          } else {                 // e1 = x, e2 = n x x = sg y
            set_dir(st * cp, sg * st * sp, sg * mu);
          }
          // placed at the face centroid (:1588), in the neighbour
          u0 = Real(0.5);
          u1 = Real(0.5);
          u2 = Real(0.5);
          enter_neighbour(fax, up, m, xs, k, sink);
          return outcome;
        }
      }
    }

    // ---------------- analytic step
    // Distance to the nearest face along the direction (+inf along an axis
    // the packet does not move on); ties go to the lower axis.
    auto face_dist = [](Real ua, Real oa, Real wa, Real ioa) {
      Real frac = oa > Real(0) ? (Real(1) - ua) : ua;
      Real d = frac * wa * ioa;
      return oa != Real(0) ? d : M::inf();
    };
    const Real d0 = face_dist(u0, o0, w0w, io0), d1 = face_dist(u1, o1, w1w, io1),
               d2 = face_dist(u2, o2, w2w, io2);
    int fax = 0, sax = 1;
    Real dface = d0, dsecond = d1;
    if (d1 < d0) {
      fax = 1;
      sax = 0;
      dface = d1;
      dsecond = d0;
    }
    if (d2 < dface) {
      sax = fax;
      dsecond = dface;
      fax = 2;
      dface = d2;
    } else if (d2 < dsecond) {
      sax = 2;
      dsecond = d2;
    }
    if constexpr (fixed_point) {
      // Two faces within 1e-6 of each other along the path (an edge or corner
      // crossing): decide which one is hit first in FP64.
      if (dsecond < M::inf() && dsecond - dface <= Real(1e-6) * dsecond) {
        auto dist64 = [&](int a) {
          double oa = static_cast<double>(sel3(a, o0, o1, o2));
          double ua = static_cast<double>(sel3(a, u0, u1, u2));
          double fr = oa > 0.0 ? (1.0 - ua) : ua;
          return fr * static_cast<double>(sel3(a, w0w, w1w, w2w)) / (oa > 0.0 ? oa : -oa);
        };
        double dA = dist64(fax), dB = dist64(sax);
        if (dB < dA) {
          fax = sax;
          dface = static_cast<Real>(dB);
        }
      }
    }
    // The neighbour across the nearest face (this cell again when that face
    // is on the box boundary): its data is loaded now and used only if the
    // segment ends on that face, so that the load latency overlaps the
    // random numbers and the attenuation below.
    const Real of = sel3(fax, o0, o1, o2);
    const bool up = of > Real(0);
    const uint32_t idx = sel3<uint32_t>(fax, i, j, kk);
    const uint32_t nax = static_cast<uint32_t>(sel3<int32_t>(fax, m.nx, m.ny, m.nz));
    const bool outside = up ? (idx + 1u >= nax) : (idx == 0u);
    const uint32_t nidx = outside ? idx : (up ? idx + 1u : idx - 1u);
    const uint32_t snz = static_cast<uint32_t>(m.nz);
    const uint32_t cstride = sel3<uint32_t>(fax, static_cast<uint32_t>(m.ny) * snz, snz, 1u);
    const uint32_t istride = sel3<uint32_t>(fax, 1u << 20, 1u << 10, 1u);
    const uint32_t ncell = outside ? cell : (up ? cell + cstride : cell - cstride);
    const uint32_t nijk = outside ? ijk : (up ? ijk + istride : ijk - istride);
    const CellXS<Real> nxs = xs[ncell];
    const Real *nwp = width_at(m, fax, nidx);
    const Real nw = nwp[0], niw = nwp[1];
    Real nsig_s = Real(0);
    if ((F & kFeatMultigroup) != 0u && m.spec.groups >= 2) nsig_s = m.spec.sig_s[ncell];

    const Philox4 r = rs.block(ctr++);
    // Distance to the next effective-scatter event (-ln xi / sigma_ev).
    Real xi = M::uniform(r.v[0]);
    Real dcol = cell_sig_ev > Real(0) ? (-M::log(xi)) * inv_sig_ev : M::inf();
    // Event: face < scatter < census on ties, as the reference's min_element.
    int ev = 0;
    Real d = dface;
    if (dcol < d) {
      ev = 1;
      d = dcol;
    }
    if (s < d) {
      ev = 2;
      d = s;
    }
    s -= d;
    if (ev == 2 || s < Real(0)) s = Real(0);

    absorb(sig_abs, d, true);

    // Move (the face axis is set exactly below).
    {
      const Real m0 = o0 * d * icw0, m1 = o1 * d * icw1, m2 = o2 * d * icw2;
      Real n0 = u0 + m0, n1 = u1 + m1, n2 = u2 + m2;
      n0 = n0 < Real(0) ? Real(0) : (n0 > Real(1) ? Real(1) : n0);
      n1 = n1 < Real(0) ? Real(0) : (n1 > Real(1) ? Real(1) : n1);
      n2 = n2 < Real(0) ? Real(0) : (n2 > Real(1) ? Real(1) : n2);
      if constexpr (fixed_point) {
        // the low 8 bits of these words are not used by the uniforms above
        n0 = move_coord(u0, n0, m0, r.v[1] & 0xffu);
        n1 = move_coord(u1, n1, m1, r.v[2] & 0xffu);
        n2 = move_coord(u2, n2, m2, r.v[3] & 0xffu);
      }
      if (!(ev == 0 && fax == 0)) u0 = n0;
      if (!(ev == 0 && fax == 1)) u1 = n1;
      if (!(ev == 0 && fax == 2)) u2 = n2;
    }

    // Roulette (RadiationIMC.hpp:1812-1820): the remaining weight goes to the material.
    if (w < static_cast<double>(wmin)) {
      roulette(k, sink);
      return outcome;
    }

    if (ev == 0) {
      if (!outside) {
        // interior face: the packet enters the neighbour on the opposite face
        enter_loaded_neighbour(fax, up, nijk, ncell, nw, niw, nxs, nsig_s, m, k, sink);
      } else {
        const int face = 2 * fax + (up ? 1 : 0);
        put3<Real>(fax, u0, u1, u2, up ? Real(1) : Real(0));
        if ((m.escape_mask >> face) & 1u) {
          // escape (REMOVE); the weight leaves the problem
          double wq = fixed_point ? fx::carry_round(w, w_enter, k.eq, id, ctr, cell) : w;
          flush(wq, k, sink);
          if constexpr (fixed_point) sink.escape_fx(cell, wq);
          else sink.escape_fp(cell, wq);
          w = 0.0;
          outcome = kEscaped;
        } else {
          // specular reflection on the box face (BoundaryCondition.hpp:87-108)
          put3<Real>(fax, o0, o1, o2, -of);
        }
      }
    } else if (ev == 1) {
      // scatter: isotropic redirection (RadiationIMC.hpp:1827-1874)
      Real nd[3];
      isotropic_dir<Real>(M::uniform(r.v[2]), M::uniform(r.v[3]), nd);
      set_dir(nd[0], nd[1], nd[2]);
      if ((F & kFeatMultigroup) != 0u && m.spec.groups >= 2) {
        // an effective (absorption + re-emission) scatter takes a new energy
        // from the cell's emission spectrum (:1843-1852); an elastic one
        // (probability sigma_s / event opacity) keeps it
        const bool elastic = M::uniform(r.v[1]) * cell_sig_ev < cell_sig_s;
        if (!elastic) {
          group_split(k, sink);
          const Philox4 r2 = rs.block(ctr++);
          const int32_t G = m.spec.groups;
          nu = spectrum_quantile<Real>(m.spec.cdf + static_cast<size_t>(cell) * static_cast<size_t>(G + 1),
                                       m.spec.bounds, G, M::uniform(r2.v[0]));
          load_opacity(m, xs);
        }
      }
    } else {
      census(k, sink);
    }
    return outcome;
  }
};

// The tracker of a mesh type (the polyhedral mesh's is in unstructured.h).
template <typename Real, uint32_t F, typename MeshT>
struct TrackerOf {
  using type = Tracker<Real, F>;
};

// One transport pass of a packet: at most `max_segments` segments, then the
// packet is written back (alive, in census, absorbed or escaped).
template <typename Real, uint32_t F = kFeatAll, typename Sink>
STORM_HD int transport_packet(Packet<Real> &p, const MeshView<Real> &m, const CellXS<Real> *xs,
                              const AccelView<Real> &acc, const StepConsts &k, int max_segments, Sink &sink) {
  Tracker<Real, F> t;
  t.begin(p, m, xs, k);
  for (int seg = 0; seg < max_segments && t.outcome == kAlive; ++seg) t.segment(m, xs, acc, k, sink);
  t.end(p, k, sink);
  return t.outcome;
}

}  // namespace storm
