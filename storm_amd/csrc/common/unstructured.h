// Per-packet IMC transport on a mesh of convex polyhedral cells (a Voronoi
// tessellation, or any mesh flattened to per-cell face tables), shared by
// the gfx950 kernels and the CPU replay.
//
// The physics of a segment is the Cartesian tracker's (transport.h):
// RadiationIMC::step (radiation/RadiationIMC.hpp:1700-1881), the random walk
// (tryRandomWalkStep, :1065-1276) and DDMC (tryDDMCStep, :1415-1612).  What
// changes is the geometry, which follows the reference's general-face code:
//   * distance to the nearest face: the smallest non-negative
//     (d_f - n_f . x) / (n_f . o) over the faces moving outward
//     (Particle::distanceToNearestFace, particle/Particle.hpp:188-279);
//   * the random walk's sphere radius: the smallest plane distance
//     d_f - n_f . x (computeMinDistanceToFaces, RadiationIMC.hpp:956-969);
//     after a walk the position is nudged towards the mesh point by 1e-6
//     and the displacement shrunk by 0.99 while outside a face (:1211-1228);
//   * a DDMC leak through face f lands on the face centroid with
//     mu = sqrt(U) about the outward normal in the basis of :1561-1586;
//   * a box face removes or reflects the packet (escape mask per side).
//
// Layout (MI355X): a packet's position is kept relative to its cell's mesh
// point (the Voronoi generator), so FP32 carries it to ~1e-7 of the cell
// size; crossing a face adds the face's frame shift (mesh point of the cell
// minus that of the neighbour).  Each face record is 32 bytes in FP32
// (normal, plane offset in the cell's frame, neighbour, frame shift).
#pragma once

#include <stdint.h>
#include <math.h>
#include "hd.h"
#include "transport.h"

namespace storm {

template <typename Real>
struct UFace {
  Real n[3];     // outward unit normal
  Real d;        // plane offset in the cell's frame: n . x = d on the face
  int32_t nbr;   // neighbour cell, or -1 - box side (x-, x+, y-, y+, z-, z+)
  Real sh[3];    // mesh point of this cell minus that of the neighbour (0 on the box)
};

template <typename Real>
struct UMeshView {
  uint32_t ncells;
  uint32_t escape_mask;            // bit s: box side s removes the packet (else it reflects)
  const uint32_t *face_start;      // ncells + 1
  const UFace<Real> *face;         // per face
  const Real *face_centroid;       // 3 per face, in the cell's frame (DDMC leaks)
  const Real *ddmc_rate;           // per face, leak rate / c (cm^-1); null without DDMC
  SpectrumView<Real> spec;
};

// A tetrahedron of a cell's fan decomposition: the mesh point (the frame's
// origin) and a triangle of a face fan, a, b, c in the cell's frame; `cum`
// the cell's cumulative volume fraction up to and including it.
template <typename Real>
struct UTet {
  Real cum;
  Real a[3], b[3], c[3];
};

// A triangle of a source face's fan, vertices in the cell's frame; `cum` the
// face's cumulative area fraction.
template <typename Real>
struct UTri {
  Real cum;
  Real a[3], b[3], c[3];
};

// A boundary source face: its cell, its triangles and its packets
// [first, first + count) of the step's boundary packets.
struct USource {
  uint32_t cell;
  uint32_t tri_start, tri_count;
  uint32_t first, count;
  int32_t xhi;  // 1: an x+ face (packets enter along -x)
};

// Stochastic rounding of a sub-unit move of an absolute coordinate: when
// x + delta rounds back to x, take one unit in the direction of delta with
// probability |delta| / unit (8 random bits), so the expected move is exact
// and a packet resting on a face of a very opaque cell leaves it.
STORM_HD float ulp_step(float x, bool upward) {
  if (x == 0.0f) return upward ? 1.17549435e-38f : -1.17549435e-38f;  // the smallest normal number
  uint32_t b;
  __builtin_memcpy(&b, &x, sizeof(b));
  const bool positive = (b >> 31) == 0u;
  b = positive == upward ? b + 1u : b - 1u;
  float y;
  __builtin_memcpy(&y, &b, sizeof(y));
  return y;
}
STORM_HD float umove_coord(float x, float nx, float delta, uint32_t r8) {
  if (nx != x || delta == 0.0f) return nx;
  const float up = ulp_step(x, delta > 0.0f);
  const float unit = delta > 0.0f ? up - x : x - up;
  const float p = (delta > 0.0f ? delta : -delta) / unit;
  return (static_cast<float>(r8) + 0.5f) * 0.00390625f < p ? up : x;
}

template <typename Real, uint32_t F = kFeatAll>
struct UTracker {
  using M = MathPolicy<Real>;
  static constexpr bool fixed_point = sizeof(Real) == 4;

  PacketStream rs;
  Real x0, x1, x2, o0, o1, o2, s, wmin, nu;
  Real cell_sig_abs, cell_sig_ev, cell_sig_s, inv_sig_ev;
  double w, w_enter, path_int;
  uint32_t cell, ctr, visit_segs, id, fbeg, fend;
  int outcome;
  bool derive_pint;
  // per-group tally (multigroup with withEgTimeAvg), as in Tracker
  bool eg_track;
  int32_t grp;
  double eg_mark;

  STORM_HD void begin(const Packet<Real> &p, const UMeshView<Real> &m, const CellXS<Real> *xs, const StepConsts &k) {
    rs = PacketStream::make(k.seed, p.id, k.step, kTagTransport);
    x0 = p.u[0];
    x1 = p.u[1];
    x2 = p.u[2];
    o0 = p.o[0];
    o1 = p.o[1];
    o2 = p.o[2];
    s = p.s;
    w = p.w;
    cell = p.ijk;
    ctr = p.ctr;
    wmin = p.wmin;
    id = p.id;
    nu = p.nu;
    eg_track = (F & kFeatMultigroup) != 0u && m.spec.groups >= 2 && m.spec.eg != 0;
    grp = 0;
    eg_mark = 0.0;
    load_cell(m, xs);
    w_enter = w;
    path_int = 0.0;
    visit_segs = 0;
    outcome = kAlive;
  }

  template <typename Sink>
  STORM_HD void end(Packet<Real> &p, const StepConsts &k, Sink &sink) {
    if (outcome == kAlive) {
      double wq = fixed_point ? fx::carry_round(w, w_enter, k.eq, id, ctr, cell) : w;
      flush(wq, k, sink);
      w = wq;
    }
    p.u[0] = x0;
    p.u[1] = x1;
    p.u[2] = x2;
    p.o[0] = o0;
    p.o[1] = o1;
    p.o[2] = o2;
    p.s = s;
    p.w = w;
    p.wmin = wmin;
    p.ijk = cell;
    p.ctr = ctr;
    p.id = id;
    p.nu = nu;
  }

  STORM_HD bool derives_path_integral(const UMeshView<Real> &m) const {
    if constexpr ((F & kFeatMultigroup) != 0u) {
      if (m.spec.groups >= 2) return false;
    }
    return fixed_point && cell_sig_abs > Real(0);
  }

  // Hands the path integral since the last call to the packet's group: at
  // the end of each visit and before the photon energy changes.
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

  STORM_HD void load_cell(const UMeshView<Real> &m, const CellXS<Real> *xs) {
    fbeg = m.face_start[cell];
    fend = m.face_start[cell + 1u];
    load_opacity(m, xs);
  }
  STORM_HD void load_opacity(const UMeshView<Real> &m, const CellXS<Real> *xs) {
    const bool mg = (F & kFeatMultigroup) != 0u && m.spec.groups >= 2;
    set_opacity(m, xs[cell], mg ? m.spec.sig_s[cell] : Real(0));
  }
  STORM_HD void set_opacity(const UMeshView<Real> &m, const CellXS<Real> &xc, Real sig_s) {
    cell_sig_abs = xc.sig_abs;
    cell_sig_ev = xc.sig_ev;
    cell_sig_s = Real(0);
    inv_sig_ev = xc.inv_ev;
    if ((F & kFeatMultigroup) != 0u && m.spec.groups >= 2) {
      Real e = nu < m.spec.e_lo ? m.spec.e_lo : (nu > m.spec.e_hi ? m.spec.e_hi : nu);
      const Real fac = M::exp(-m.spec.nu_pow * M::log(e));
      cell_sig_s = sig_s;
      cell_sig_abs = xc.sig_abs * fac;
      cell_sig_ev = cell_sig_s + xc.sig_ev * fac;
      if (eg_track) grp = group_of<Real>(e, m.spec.bounds, m.spec.groups);
      inv_sig_ev = cell_sig_ev > Real(0) ? Real(1) / cell_sig_ev : Real(0);
    } else if constexpr ((F & kFeatAccel) != 0u) {
      // the same value, divided here: loading it made the register
      // allocation of the gray random-walk/DDMC kernel spill more (18%
      // slower on the batch of slabs, measured)
      inv_sig_ev = cell_sig_ev > Real(0) ? Real(1) / cell_sig_ev : Real(0);
    }
    derive_pint = derives_path_integral(m);
  }

  STORM_HD void absorb(Real sig_abs, Real d, bool tally_free) {
    Real xabs = sig_abs * d;
    Real frac = -M::expm1(-xabs);
    double w0 = w;
    w = w0 - w0 * static_cast<double>(frac);
    if (derive_pint) return;
    if (xabs >= Real(1e-12)) {
      path_int += w0 * static_cast<double>(frac / sig_abs);
    } else if (tally_free) {
      path_int += w0 * static_cast<double>(d);
    } else if (sig_abs > Real(0)) {
      path_int += w0 * static_cast<double>(frac / sig_abs);
    }
  }

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
    double wq = fx::carry_round(w, w_enter, k.eq, id, ctr, cell);
    flush(wq, k, sink);
    w = wq;
    outcome = kCensus;
  }

  // Close the visit and continue in the cell across face `f`.
  template <typename Sink>
  STORM_HD void cross(const UFace<Real> &fc, const UMeshView<Real> &m, const CellXS<Real> *xs, const StepConsts &k,
                      Sink &sink) {
    double wq = fixed_point ? fx::carry_round(w, w_enter, k.eq, id, ctr, cell) : w;
    flush(wq, k, sink);
    w = wq;
    w_enter = wq;
    x0 += fc.sh[0];
    x1 += fc.sh[1];
    x2 += fc.sh[2];
    cell = static_cast<uint32_t>(fc.nbr);
    load_cell(m, xs);
  }

  template <typename Sink>
  STORM_HD int segment(const UMeshView<Real> &m, const CellXS<Real> *xs, const AccelView<Real> &acc,
                       const StepConsts &k, Sink &sink) {
    const Real sig_abs = cell_sig_abs;
    ++visit_segs;
    if ((F & kFeatAccel) != 0u && acc.cell) {
      const CellAccel<Real> ca = acc.cell[cell];
      if (ca.flags) {
        // distance to the nearest face plane (computeMinDistanceToFaces)
        Real Ro = M::inf();
        for (uint32_t f = fbeg; f < fend; ++f) {
          const UFace<Real> fc = m.face[f];
          const Real dist = fc.d - (fc.n[0] * x0 + fc.n[1] * x1 + fc.n[2] * x2);
          Ro = dist < Ro ? dist : Ro;
        }
        if (!(Ro > Real(0))) Ro = Real(0);
        const Real tauR = Ro * ca.sig_t;
        const bool pgrw = (F & kFeatMultigroup) != 0u && m.spec.groups >= 2;
        const bool walks = !pgrw || nu < ca.e_cut;
        if ((ca.flags & 1u) && walks && Ro > Real(0) && ca.sig_t > Real(0) && tauR >= acc.rw_particle_tau) {
          // ---------------- random-walk step (tryRandomWalkStep)
          const Philox4 r = rs.block(ctr++);
          const Philox4 r2 = rs.block(ctr++);
          const rw::Tables<Real> tb = rw::tables_at(acc.rw);
          const Real tauLeak = rw::sample_leak_time<Real>(tb, M::uniform(r.v[0]));
          const Real sig_d = pgrw ? ca.sig_d : ca.sig_t;
          const Real sLeak = tauLeak * Ro * Ro * (Real(3) * sig_d);
          Real sUp = M::inf();
          if (pgrw && ca.up_rate > Real(0)) sUp = (-M::log(M::uniform(r2.v[2]))) / ca.up_rate;
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
          Real dir[3];
          isotropic_dir<Real>(M::uniform(r.v[1]), M::uniform(r.v[2]), dir);
          Real disp = Ro;
          if (!leak) {
            const Real tauPos = d / (Real(3) * sig_d * Ro * Ro);
            disp = Ro * rw::sample_radius<Real>(tb, tauPos, M::uniform(r.v[3]));
          }
          if (disp > Ro) disp = Ro;
          const Real c0 = x0, c1 = x1, c2 = x2;
          // nudge towards the mesh point, the frame's origin (:1214-1215)
          Real n0 = (c0 + disp * dir[0]) * Real(1 - 1e-6);
          Real n1 = (c1 + disp * dir[1]) * Real(1 - 1e-6);
          Real n2 = (c2 + disp * dir[2]) * Real(1 - 1e-6);
          for (int guard = 0; guard < 2000; ++guard) {
            bool inside = true;
            for (uint32_t f = fbeg; f < fend && inside; ++f) {
              const UFace<Real> fc = m.face[f];
              inside = fc.n[0] * n0 + fc.n[1] * n1 + fc.n[2] * n2 <= fc.d;
            }
            if (inside) break;
            disp = disp * Real(0.99);
            n0 = c0 + disp * dir[0];
            n1 = c1 + disp * dir[1];
            n2 = c2 + disp * dir[2];
          }
          x0 = n0;
          x1 = n1;
          x2 = n2;
          Real nd[3];
          isotropic_dir<Real>(M::uniform(r2.v[0]), M::uniform(r2.v[1]), nd);
          o0 = nd[0];
          o1 = nd[1];
          o2 = nd[2];
          if ((F & kFeatMultigroup) != 0u && rev == 2) {
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
            ((F & kFeatMultigroup) == 0u || nu < ca.ddmc_e_cut) && m.ddmc_rate) {
          // ---------------- DDMC step (tryDDMCStep)
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
          uint32_t face = 0xFFFFFFFFu, lastf = 0xFFFFFFFFu;
          for (uint32_t f = fbeg; f < fend; ++f) {
            const Real rt = m.ddmc_rate[f];
            if (rt > Real(0)) {
              lastf = f;
              pick -= rt;
              if (face == 0xFFFFFFFFu && pick <= Real(0)) face = f;
            }
          }
          if (face == 0xFFFFFFFFu) face = lastf;
          const UFace<Real> fc = m.face[face];
          // mu = sqrt(U) about the outward normal, basis as in :1565-1586
          const Real mu = M::sqrt(M::uniform(r.v[2]));
          const Real st2 = Real(1) - mu * mu;
          const Real st = M::sqrt(st2 > Real(0) ? st2 : Real(0));
          Real sp, cp;
          M::sincos2pi(M::uniform(r.v[3]), sp, cp);
          const Real nx = fc.n[0], ny = fc.n[1], nz = fc.n[2];
          const bool hx = (nx < Real(0) ? -nx : nx) < Real(0.9);
          // helper (1,0,0) or (0,1,0), e1 = helper - (helper . n) n, e2 = n x e1
          Real e10 = (hx ? Real(1) : Real(0)) - (hx ? nx : ny) * nx;
          Real e11 = (hx ? Real(0) : Real(1)) - (hx ? nx : ny) * ny;
          Real e12 = -(hx ? nx : ny) * nz;
          const Real e1n = M::sqrt(e10 * e10 + e11 * e11 + e12 * e12);
          if (e1n > Real(0)) {
            e10 /= e1n;
            e11 /= e1n;
            e12 /= e1n;
          }
          const Real e20 = ny * e12 - nz * e11, e21 = nz * e10 - nx * e12, e22 = nx * e11 - ny * e10;
          Real d0 = mu * nx + st * cp * e10 + st * sp * e20;
          Real d1 = mu * ny + st * cp * e11 + st * sp * e21;
          Real d2 = mu * nz + st * cp * e12 + st * sp * e22;
          const Real dn = M::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
          if (dn > Real(0)) {
            d0 /= dn;
            d1 /= dn;
            d2 /= dn;
          }
          o0 = d0;
          o1 = d1;
          o2 = d2;
          // on the face centroid (:1588), then into the neighbour
          x0 = m.face_centroid[3 * face];
          x1 = m.face_centroid[3 * face + 1];
          x2 = m.face_centroid[3 * face + 2];
          cross(fc, m, xs, k, sink);
          return outcome;
        }
      }
    }

    // ---------------- analytic step
    // nearest face along the direction, and the next one for ties
    Real dface = M::inf(), dsecond = M::inf();
    uint32_t fa = fbeg, fb = fbeg;
    for (uint32_t f = fbeg; f < fend; ++f) {
      const UFace<Real> fc = m.face[f];
      const Real den = fc.n[0] * o0 + fc.n[1] * o1 + fc.n[2] * o2;
      if (!(den > Real(0))) continue;
      Real t = (fc.d - (fc.n[0] * x0 + fc.n[1] * x1 + fc.n[2] * x2)) / den;
      t = t > Real(0) ? t : Real(0);
      if (t < dface) {
        dsecond = dface;
        fb = fa;
        dface = t;
        fa = f;
      } else if (t < dsecond) {
        dsecond = t;
        fb = f;
      }
    }
    if constexpr (fixed_point) {
      // two faces within 1e-6 along the path (an edge or a vertex): decide in FP64
      if (dsecond < M::inf() && dsecond - dface <= Real(1e-6) * dsecond) {
        auto dist64 = [&](uint32_t f) {
          const UFace<Real> fc = m.face[f];
          const double den = static_cast<double>(fc.n[0]) * o0 + static_cast<double>(fc.n[1]) * o1 +
                             static_cast<double>(fc.n[2]) * o2;
          const double num = static_cast<double>(fc.d) - (static_cast<double>(fc.n[0]) * x0 +
                                                          static_cast<double>(fc.n[1]) * x1 +
                                                          static_cast<double>(fc.n[2]) * x2);
          const double t = num / den;
          return t > 0.0 ? t : 0.0;
        };
        const double dA = dist64(fa), dB = dist64(fb);
        if (dB < dA) {
          fa = fb;
          dface = static_cast<Real>(dB);
        }
      }
    }
    const UFace<Real> fcx = m.face[fa];
    const Philox4 r = rs.block(ctr++);
    Real xi = M::uniform(r.v[0]);
    Real dcol = cell_sig_ev > Real(0) ? (-M::log(xi)) * inv_sig_ev : M::inf();
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
    {
      const Real m0 = o0 * d, m1 = o1 * d, m2 = o2 * d;
      Real n0 = x0 + m0, n1 = x1 + m1, n2 = x2 + m2;
      if constexpr (fixed_point) {
        n0 = umove_coord(x0, n0, m0, r.v[1] & 0xffu);
        n1 = umove_coord(x1, n1, m1, r.v[2] & 0xffu);
        n2 = umove_coord(x2, n2, m2, r.v[3] & 0xffu);
      }
      x0 = n0;
      x1 = n1;
      x2 = n2;
    }
    if (w < static_cast<double>(wmin)) {
      roulette(k, sink);
      return outcome;
    }
    if (ev == 0) {
      if (fcx.nbr >= 0) {
        cross(fcx, m, xs, k, sink);
      } else {
        const uint32_t side = static_cast<uint32_t>(-1 - fcx.nbr);
        if ((m.escape_mask >> side) & 1u) {
          double wq = fixed_point ? fx::carry_round(w, w_enter, k.eq, id, ctr, cell) : w;
          flush(wq, k, sink);
          if constexpr (fixed_point) sink.escape_fx(cell, wq);
          else sink.escape_fp(cell, wq);
          w = 0.0;
          outcome = kEscaped;
        } else {
          // specular reflection on the box face (BoundaryCondition.hpp:87-108)
          const Real on = o0 * fcx.n[0] + o1 * fcx.n[1] + o2 * fcx.n[2];
          o0 -= Real(2) * on * fcx.n[0];
          o1 -= Real(2) * on * fcx.n[1];
          o2 -= Real(2) * on * fcx.n[2];
        }
      }
    } else if (ev == 1) {
      Real nd[3];
      isotropic_dir<Real>(M::uniform(r.v[2]), M::uniform(r.v[3]), nd);
      o0 = nd[0];
      o1 = nd[1];
      o2 = nd[2];
      if ((F & kFeatMultigroup) != 0u && m.spec.groups >= 2) {
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

template <typename Real, uint32_t F>
struct TrackerOf<Real, F, UMeshView<Real>> {
  using type = UTracker<Real, F>;
};

template <typename Real, uint32_t F = kFeatAll, typename Sink>
STORM_HD int transport_packet(Packet<Real> &p, const UMeshView<Real> &m, const CellXS<Real> *xs,
                              const AccelView<Real> &acc, const StepConsts &k, int max_segments, Sink &sink) {
  UTracker<Real, F> t;
  t.begin(p, m, xs, k);
  for (int seg = 0; seg < max_segments && t.outcome == kAlive; ++seg) t.segment(m, xs, acc, k, sink);
  t.end(p, k, sink);
  return t.outcome;
}

}  // namespace storm
