// Packet birth for the two sources of a gray IMC step, shared by the gfx950
// sourcing kernels and the CPU replay.
//
//   emission : RadiationIMC::generateParticles / generateSingleParticle
//              (radiation/RadiationIMC.hpp:2405-2553): a uniform point in the
//              cell (RandomInCellPositionSampler is uniform in a box,
//              utils/RandomInCell.hpp:13-66), isotropic direction
//              (RadiationOpacityModel.hpp:139-150), timeLeft = dt * U, weight
//              E_i / n_i, initialWeight = weight.
//   boundary : SideTemperature::generateNewBoundaryParticles
//              (boundary/SideTemperature.hpp:114-163): a uniform point on the
//              x- box face of the cell, mu = sqrt(U) about +x, phi = 2 pi U,
//              timeLeft = dt * U, weight sigma_SB T^4 A dt / Npercell.
// Multigroup: an emitted packet takes its energy from the cell's emission
// CDF (GetThermalEnergy, RadiationIMC.hpp:2496-2500), a boundary packet from
// the Planck fractions of the source temperature with the energy uniform in
// the group (examples/densmore2012/DensmoreBoundary.hpp:113-125).
// Each packet draws from its own stream keyed by (seed, id) with the birth tag.
#pragma once

#include <stdint.h>
#include "hd.h"
#include "philox.h"
#include "transport.h"

namespace storm {

template <typename Real>
STORM_HD void birth_emission(Packet<Real> &p, uint32_t id, uint32_t ijk, double w, Real s0, uint64_t seed,
                             uint32_t step, const SpectrumView<Real> &sp, uint32_t cell) {
  using M = MathPolicy<Real>;
  PacketStream rs = PacketStream::make(seed, id, step, kTagBirth);
  Philox4 r0 = rs.block(0), r1 = rs.block(1);
  p.u[0] = M::uniform(r0.v[0]);
  p.u[1] = M::uniform(r0.v[1]);
  p.u[2] = M::uniform(r0.v[2]);
  p.s = s0 * M::uniform(r0.v[3]);
  isotropic_dir<Real>(M::uniform(r1.v[0]), M::uniform(r1.v[1]), p.o);
  p.w = w;
  p.wmin = static_cast<Real>(w * 1e-3);
  p.ijk = ijk;
  p.id = id;
  p.ctr = 0u;
  p.flags = 0u;
  p.nu = Real(0);
  if (sp.groups >= 2) {
    const int32_t G = sp.groups;
    p.nu = spectrum_quantile<Real>(sp.cdf + static_cast<size_t>(cell) * static_cast<size_t>(G + 1), sp.bounds, G,
                                   M::uniform(r1.v[2]));
  }
}

// `i` is the cell column of the face: 0 for the x- face (entering along
// +x), nx - 1 for the x+ face of TwoSidesTemperature (entering along -x,
// boundary/TwoSidesTemperature.hpp:176-180).  `bcdf`: normalised Planck
// fractions of the source temperature over the groups (groups + 1
// entries), or null when gray.
template <typename Real>
STORM_HD void birth_boundary(Packet<Real> &p, uint32_t id, bool xhi, uint32_t i, uint32_t j, uint32_t k, double w,
                             Real s0, uint64_t seed, uint32_t step, const SpectrumView<Real> &spec,
                             const Real *bcdf) {
  using M = MathPolicy<Real>;
  PacketStream rs = PacketStream::make(seed, id, step, kTagBirth);
  Philox4 r0 = rs.block(0), r1 = rs.block(1);
  p.u[0] = xhi ? Real(1) : Real(0);
  p.u[1] = M::uniform(r0.v[0]);
  p.u[2] = M::uniform(r0.v[1]);
  Real mu = M::sqrt(M::uniform(r0.v[2]));
  Real st2 = Real(1) - mu * mu;
  Real st = M::sqrt(st2 > Real(0) ? st2 : Real(0));
  Real sp, cp;
  M::sincos2pi(M::uniform(r0.v[3]), sp, cp);
  p.o[0] = xhi ? -mu : mu;
  p.o[1] = st * cp;
  p.o[2] = st * sp;
  p.s = s0 * M::uniform(r1.v[0]);
  p.w = w;
  p.wmin = static_cast<Real>(w * 1e-3);
  p.ijk = pack_ijk(i, j, k);
  p.id = id;
  p.ctr = 0u;
  p.flags = 0u;
  p.nu = Real(0);
  if (spec.groups >= 2 && bcdf) {
    const int32_t G = spec.groups;
    const Real r = M::uniform(r1.v[1]);
    int32_t g = 0;
    while (g + 1 < G && bcdf[g + 1] < r) ++g;
    p.nu = spec.bounds[g] + M::uniform(r1.v[2]) * (spec.bounds[g + 1] - spec.bounds[g]);
  }
}

}  // namespace storm
