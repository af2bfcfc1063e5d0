// gfx950 kernels of the unstructured-mesh (polyhedral cell) path: packet
// birth by tetrahedra and boundary triangles, and the DDMC leak rates per
// face.  The transport itself is k_transport over UMeshView (UTracker,
// common/unstructured.h); census, comb, preStep and postStep are the
// Cartesian path's per-cell kernels run with flat cell indices.
#pragma once

#include "common/unstructured.h"
#include "kernels.h"

namespace storm {
namespace gpu {

template <typename Real>
__device__ __forceinline__ uint32_t pick_cum(const Real *cum_base, uint32_t stride_words, uint32_t n, Real u) {
  // first entry with cum > u (cum of the last entry is 1)
  uint32_t lo = 0, hi = n - 1u;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cum_base[static_cast<size_t>(mid) * stride_words] > u) hi = mid;
    else lo = mid + 1u;
  }
  return lo;
}

// Emission in a polyhedral cell: RandomInCellPositionSampler
// (utils/RandomInCell.hpp:10-67): a tetrahedron of the fan about the mesh
// point chosen by volume, then a uniform point in it from three sorted
// uniforms; isotropic direction; timeLeft = dt * U; as birth_emission
// otherwise (same random blocks, the position's draws in r0.v[0..2] and
// r1.v[3]).
template <typename Real>
__device__ void birth_emission_u(Packet<Real> &p, uint32_t id, uint32_t cell, double w, Real s0, uint64_t seed,
                                 uint32_t step, const SpectrumView<Real> &sp, const UTet<Real> *tets,
                                 uint32_t tet_first, uint32_t tet_count) {
  using M = MathPolicy<Real>;
  PacketStream rs = PacketStream::make(seed, id, step, kTagBirth);
  Philox4 r0 = rs.block(0), r1 = rs.block(1);
  const UTet<Real> *T = tets + tet_first;
  const uint32_t t = pick_cum<Real>(&T[0].cum, sizeof(UTet<Real>) / sizeof(Real), tet_count, M::uniform(r0.v[0]));
  Real x1 = M::uniform(r0.v[1]), x2 = M::uniform(r0.v[2]), x3 = M::uniform(r1.v[3]);
  // sort x1 <= x2 <= x3
  if (x1 > x2) { const Real q = x1; x1 = x2; x2 = q; }
  if (x2 > x3) { const Real q = x2; x2 = x3; x3 = q; }
  if (x1 > x2) { const Real q = x1; x1 = x2; x2 = q; }
  const Real la = x2 - x1, lb = x3 - x2, lc = Real(1) - x3;  // the mesh point takes x1
  const UTet<Real> &tt = T[t];
  for (int a = 0; a < 3; ++a) p.u[a] = la * tt.a[a] + lb * tt.b[a] + lc * tt.c[a];
  p.s = s0 * M::uniform(r0.v[3]);
  isotropic_dir<Real>(M::uniform(r1.v[0]), M::uniform(r1.v[1]), p.o);
  p.w = w;
  p.wmin = static_cast<Real>(w * 1e-3);
  p.ijk = cell;
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

// Emission packets in cell order (flat cells: the Z-order code is the cell).
template <typename Real>
__global__ __launch_bounds__(256) void k_emit_u(uint32_t n_emit, uint32_t M, const uint64_t *__restrict__ offsets,
                                                const double *__restrict__ wemit, PacketSoA<Real> out, uint32_t base,
                                                Real s0, uint64_t seed, uint32_t step, uint32_t *__restrict__ begin,
                                                SpectrumView<Real> sp, const UTet<Real> *__restrict__ tets,
                                                const uint32_t *__restrict__ tet_start) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < n_emit;
  uint32_t c = 0xFFFFFFFFu;
  if (valid) {
    uint32_t lo = 0, hi = M;  // offsets[M] = n_emit
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (offsets[mid] <= q) lo = mid;
      else hi = mid;
    }
    c = lo;
    Packet<Real> p;
    birth_emission_u<Real>(p, base + q, c, wemit[c], s0, seed, step, sp, tets, tet_start[c],
                           tet_start[c + 1] - tet_start[c]);
    store_packet(out, base + q, p);
  }
  uint32_t one = valid ? 1u : 0u;
  if (wave_run_sum(c, one) && valid) atomicAdd(&begin[c], one);
}

// Planck surface source on the x- (and x+) box faces of a polyhedral mesh
// (SideTemperature::generateNewBoundaryParticles, boundary/SideTemperature.hpp:114-163,
// with RandomPointOnFace, utils/RandomOnFace.hpp:14-66: a triangle of the
// face's fan chosen by area, a uniform point in it), mu = sqrt(U) about the
// inward normal, phi = 2 pi U, timeLeft = dt * U.  wface: weight of a packet
// of each source face.
template <typename Real>
__global__ __launch_bounds__(256) void k_boundary_u(uint32_t n_bc, uint32_t nsrc, const USource *__restrict__ src,
                                                    const UTri<Real> *__restrict__ tris,
                                                    const double *__restrict__ wface, PacketSoA<Real> out,
                                                    uint32_t base, Real s0, uint64_t seed, uint32_t step,
                                                    uint32_t *__restrict__ begin, SpectrumView<Real> sp,
                                                    const Real *__restrict__ bcdf, uint32_t bcdf_xhi) {
  using M = MathPolicy<Real>;
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < n_bc;
  uint32_t cell = 0xFFFFFFFFu;
  if (valid) {
    // source face of packet q: last face with first <= q
    uint32_t lo = 0, hi = nsrc;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (src[mid].first <= q) lo = mid;
      else hi = mid;
    }
    const USource S = src[lo];
    cell = S.cell;
    const uint32_t id = base + q;
    PacketStream rs = PacketStream::make(seed, id, step, kTagBirth);
    Philox4 r0 = rs.block(0), r1 = rs.block(1);
    Packet<Real> p;
    const UTri<Real> *T = tris + S.tri_start;
    const uint32_t t = pick_cum<Real>(&T[0].cum, sizeof(UTri<Real>) / sizeof(Real), S.tri_count, M::uniform(r1.v[3]));
    const Real su = M::sqrt(M::uniform(r0.v[0])), v = M::uniform(r0.v[1]);
    const Real la = Real(1) - su, lb = su * (Real(1) - v), lc = su * v;
    const UTri<Real> &tt = T[t];
    for (int a = 0; a < 3; ++a) p.u[a] = la * tt.a[a] + lb * tt.b[a] + lc * tt.c[a];
    const Real mu = M::sqrt(M::uniform(r0.v[2]));
    const Real st2 = Real(1) - mu * mu;
    const Real st = M::sqrt(st2 > Real(0) ? st2 : Real(0));
    Real sps, cps;
    M::sincos2pi(M::uniform(r0.v[3]), sps, cps);
    p.o[0] = S.xhi ? -mu : mu;
    p.o[1] = st * cps;
    p.o[2] = st * sps;
    p.s = s0 * M::uniform(r1.v[0]);
    const double w = wface[lo];
    p.w = w;
    p.wmin = static_cast<Real>(w * 1e-3);
    p.ijk = cell;
    p.id = id;
    p.ctr = 0u;
    p.flags = 0u;
    p.nu = Real(0);
    if (sp.groups >= 2 && bcdf) {
      const int32_t G = sp.groups;
      const Real *bc = bcdf + (S.xhi ? static_cast<size_t>(bcdf_xhi) * static_cast<size_t>(G + 1) : 0u);
      const Real r = M::uniform(r1.v[1]);
      int32_t g = 0;
      while (g + 1 < G && bc[g + 1] < r) ++g;
      p.nu = sp.bounds[g] + M::uniform(r1.v[2]) * (sp.bounds[g + 1] - sp.bounds[g]);
    }
    store_packet(out, id, p);
  }
  uint32_t one = valid ? 1u : 0u;
  if (wave_run_sum(cell, one) && valid) atomicAdd(&begin[cell], one);
}

// Leak rates of the DDMC cells of a polyhedral mesh (precomputeDDMCData,
// radiation/RadiationIMC.hpp:1359-1404): per face towards a cell,
// D_f A / (V d) / c with D_f the harmonic mean of the two cells' D (the
// cell's own when the neighbour's is 0) and d the distance from the mesh
// point to the face plane; faces on the box are skipped.  face_geo holds
// (area, distance) per face.
template <typename Real>
__global__ __launch_bounds__(256) void k_accel_rates_u(uint32_t N, const uint32_t *__restrict__ face_start,
                                                       const UFace<Real> *__restrict__ face,
                                                       const double2 *__restrict__ face_geo,
                                                       const double *__restrict__ V, const double *__restrict__ Dtil,
                                                       CellAccel<Real> *__restrict__ out, Real *__restrict__ rate) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  CellAccel<Real> a = out[c];
  const uint32_t f0 = face_start[c], f1 = face_start[c + 1];
  if (!(a.flags & 2u)) {
    for (uint32_t f = f0; f < f1; ++f) rate[f] = Real(0);
    return;
  }
  const double D1 = Dtil[c];
  double total = 0.0;
  for (uint32_t f = f0; f < f1; ++f) {
    rate[f] = Real(0);
    const int32_t nb = face[f].nbr;
    if (nb < 0) continue;
    const double D2 = Dtil[nb];
    double Df = D1;
    if (D2 > 0.0) Df = (D1 > 0.0 && D2 > 0.0) ? 2.0 * D1 * D2 / (D1 + D2) : fmax(D1, D2);
    const double2 g = face_geo[f];
    const double r = Df * g.x / (V[c] * g.y);
    if (r > 0.0 && isfinite(r)) {
      rate[f] = static_cast<Real>(r);
      total += r;
    }
  }
  a.total = static_cast<Real>(total);
  if (!(total > 0.0)) {
    a.flags &= ~2u;
    for (uint32_t f = f0; f < f1; ++f) rate[f] = Real(0);
  }
  out[c] = a;
}

}  // namespace gpu
}  // namespace storm
