// gfx950 kernels of one IMC step.  Every kernel is templated on the tracking
// type Real (float: mixed engine, double: plain FP64 engine).
//
//   k_prestep      RadiationIMC::preStep per-cell part (radiation/RadiationIMC.hpp:1632-1662)
//                  and the emission energies of generateParticles (:2046-2060)
//   k_alloc        photon allocation n_i (:2076-2087) and packet weights E_i/n_i (:2211)
//   k_import       census re-entry (MonteCarloManagerSerial.hpp:393-405)
//   k_emit         emission packets (generateSingleParticle :2516-2553, timeLeft :2408)
//   k_boundary     Planck surface source (boundary/SideTemperature.hpp:114-163)
//   k_transport    the segment loop (MonteCarloManagerSerial.hpp:266-383 + RadiationIMC::step)
//   k_comb_*       CombPopulationControl::activate (population/CombPopulationControl.hpp:79-253)
//   k_post         RadiationIMC::postStep (:1887-2029)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common/constants.h"
#include "common/fixed.h"
#include "common/philox.h"
#include "common/planck.h"
#include "common/sourcing.h"
#include "common/transport.h"
#include "device_util.h"

namespace storm {
namespace gpu {

template <typename Real>
struct V4;
template <>
struct V4<float> {
  using T = float4;
};
template <>
struct V4<double> {
  using T = double4;
};

// Packet store (structure of arrays, one HBM stream per field group).
template <typename Real>
struct PacketSoA {
  typename V4<Real>::T *pos;  // ux, uy, uz, s
  typename V4<Real>::T *dir;  // ox, oy, oz, wmin
  double *w;
  uint4 *meta;                // ijk, id, ctr, flags
  Real *nu;                   // photon energy, keV (multigroup only, else null)
};

// Census store: what survives to the end of the step.
template <typename Real>
struct CensusSoA {
  typename V4<Real>::T *pos;  // ux, uy, uz, photon energy (keV, multigroup)
  typename V4<Real>::T *dir;  // ox, oy, oz, -
  double *w;
  uint2 *meta;                // ijk, id
};

template <typename Real>
__device__ __forceinline__ Packet<Real> load_packet(const PacketSoA<Real> &s, uint32_t i) {
  Packet<Real> p;
  auto a = s.pos[i];
  auto b = s.dir[i];
  p.u[0] = a.x;
  p.u[1] = a.y;
  p.u[2] = a.z;
  p.s = a.w;
  p.o[0] = b.x;
  p.o[1] = b.y;
  p.o[2] = b.z;
  p.wmin = b.w;
  p.w = s.w[i];
  uint4 m = s.meta[i];
  p.ijk = m.x;
  p.id = m.y;
  p.ctr = m.z;
  p.flags = m.w;
  p.nu = s.nu ? s.nu[i] : Real(0);
  return p;
}

template <typename Real>
__device__ __forceinline__ void store_packet(const PacketSoA<Real> &s, uint32_t i, const Packet<Real> &p) {
  typename V4<Real>::T a, b;
  a.x = p.u[0];
  a.y = p.u[1];
  a.z = p.u[2];
  a.w = p.s;
  b.x = p.o[0];
  b.y = p.o[1];
  b.z = p.o[2];
  b.w = p.wmin;
  s.pos[i] = a;
  s.dir[i] = b;
  s.w[i] = p.w;
  s.meta[i] = make_uint4(p.ijk, p.id, p.ctr, p.flags);
  if (s.nu) s.nu[i] = p.nu;
}

template <typename Real>
__device__ __forceinline__ void store_census(const CensusSoA<Real> &c, uint32_t i, const Packet<Real> &p) {
  typename V4<Real>::T a, b;
  a.x = p.u[0];
  a.y = p.u[1];
  a.z = p.u[2];
  a.w = p.nu;
  b.x = p.o[0];
  b.y = p.o[1];
  b.z = p.o[2];
  b.w = Real(0);
  c.pos[i] = a;
  c.dir[i] = b;
  c.w[i] = p.w;
  c.meta[i] = make_uint2(p.ijk, p.id);
}

template <typename Real>
__device__ __forceinline__ void copy_census(const CensusSoA<Real> &from, uint32_t i, const CensusSoA<Real> &to,
                                            uint32_t j) {
  to.pos[j] = from.pos[i];
  to.dir[j] = from.dir[i];
  to.w[j] = from.w[i];
  to.meta[j] = from.meta[i];
}

// Per-cell integer tallies: pairs (lo sum, hi sum) of fixed-point limbs.
struct Tallies {
  unsigned long long *dep;   // [2N] material deposit by continuous absorption
  unsigned long long *pint;  // [2N] path integral of the weight, erg*cm (cells without absorption)
  unsigned long long *esc;   // [2N] escaped through REMOVE faces
  unsigned long long *cin;   // [2N] census weight entering transport
  unsigned long long *cpre;  // [2N] census weight before population control (k_comb_keys)
  unsigned long long *cout;  // [2N] census weight after population control
  unsigned long long *rem;   // [2N] comb remainders (go to the material)
  unsigned long long *rdep;  // [2N] rouletted weight (goes to the material)
  unsigned long long *cdrop; // [2N] stratified comb: weight of the dropped bins (leaves the problem)
  double *dep_fp;            // [N]  FP64 engine: deposit
  double *pint_fp;           // [N]  FP64 engine: path integral
  double *esc_fp;            // [N]  FP64 engine: escaped
  uint32_t *segs;            // [N]  tracked segments (cellsStepsCounters)
  uint32_t *begin;           // [N]  packets per cell at transport start
  // Replicas 1..kTallyReplicas-1 of the per-visit tallies (dep, pint, segs;
  // dep_fp, pint_fp in the FP64 engine), one per XCD, each laid out as its
  // replica-0 array above; k_fold_replicas adds them into replica 0 after
  // the transport.  A wavefront adds into the replica of the XCD it runs on,
  // so that the atomics of different XCDs never share a cache line.
  unsigned long long *dep_rep, *pint_rep;  // [(R-1) * 2N]
  double *dep_fp_rep, *pint_fp_rep;        // [(R-1) * N]
  uint32_t *segs_rep;                      // [(R-1) * N]
  uint32_t ncell;
  // per-group path integral (withEgTimeAvg), [cell * eg_groups + g]: quanta
  // of 2^et in the mixed engine, FP64 in the FP64 engine (null: off)
  unsigned long long *eg;
  double *eg_fp;
  uint32_t eg_groups;
};

// Tally replicas: one per XCD of an MI355X (8 XCDs of 32 CUs).
constexpr uint32_t kTallyReplicas = 8;

// Index of the XCD this wavefront runs on: hardware register XCC_ID
// (s_getreg id 20), bits [3:0].
__device__ __forceinline__ uint32_t xcd_index() {
  return static_cast<uint32_t>(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20)) & 0xfu;
}

__device__ __forceinline__ void tally_limbs(unsigned long long *t, uint32_t c, fx::Limbs v) {
  if (v.lo) atomicAdd(&t[2 * c], static_cast<unsigned long long>(v.lo));
  if (v.hi) atomicAdd(&t[2 * c + 1], static_cast<unsigned long long>(v.hi));
}
__device__ __forceinline__ void tally_limbs_n(unsigned long long *t, uint32_t c, fx::Limbs v, uint64_t n) {
  if (v.lo) atomicAdd(&t[2 * c], static_cast<unsigned long long>(v.lo * n));
  if (v.hi) atomicAdd(&t[2 * c + 1], static_cast<unsigned long long>(v.hi * n));
}

// Device material model (storm_material_desc).
struct MatDev {
  const double *sa0, *ss0, *rho;
  double Tref, Tfloor, apow, arho, spow, srho;
  double eos_f, eos_beta, eos_mu, cvTfloor;

  __device__ double density(uint32_t c) const { return rho ? rho[c] : 1.0; }
  __device__ double tfac(double T, double p) const {
    if (p == 0.0) return 1.0;
    return pow(fmax(T, Tfloor) / Tref, -p);
  }
  __device__ double rfac(uint32_t c, double p) const {
    if (p == 0.0 || !rho) return 1.0;
    return pow(rho[c], p);
  }
  __device__ double sigma_a(uint32_t c, double T) const { return sa0[c] * tfac(T, apow) * rfac(c, arho); }
  __device__ double sigma_s(uint32_t c, double T) const {
    return ss0 ? ss0[c] * tfac(T, spow) * rfac(c, srho) : 0.0;
  }
  __device__ double dT2cv(double rhoc, double T) const {
    double cvVol = eos_f * eos_beta * pow(fmax(T, cvTfloor), eos_beta - 1.0) * pow(rhoc, 1.0 - eos_mu);
    return cvVol / rhoc;
  }
  __device__ double de2T(double rhoc, double specific) const {
    double eVol = specific * rhoc;
    double ratio = fmax(eVol, 0.0) / (eos_f * pow(rhoc, 1.0 - eos_mu));
    return pow(fmax(ratio, 0.0), 1.0 / eos_beta);
  }
};

// Error flags raised by kernels (checked by the host after the step).
enum : uint32_t { kErrFleck = 1u, kErrNegativeEnergy = 2u, kErrTallyOverflow = 4u, kErrCapacity = 8u };

// Cell layout for the per-cell kernels: a Cartesian nx x ny x nz mesh whose
// packets carry packed (i, j, k), or (flat = 1) a polyhedral mesh of nx cells
// whose packets carry the cell index itself (ny = nz = 1).
struct MeshDims {
  uint32_t nx, ny, nz;
  uint32_t flat;
};
// flat cell index of a packet's cell code
__device__ __forceinline__ uint32_t cell_of_code(uint32_t code, MeshDims d) {
  return d.flat ? code : (ijk_i(code) * d.ny + ijk_j(code)) * d.nz + ijk_k(code);
}

__device__ __forceinline__ uint32_t ijk_of_flat(uint32_t c, MeshDims d) {
  uint32_t k = c % d.nz;
  uint32_t t = c / d.nz;
  uint32_t j = t % d.ny;
  uint32_t i = t / d.ny;
  return pack_ijk(i, j, k);
}

// Z-order (Morton) code of a cell: the bits of k, j and i interleaved from
// the lowest up, an axis dropping out when its bits are used up (bits[a] =
// ceil(log2(n_a))).  The combed census is written in this order, so the
// packets one transport block takes are a compact 3-D region of the mesh
// rather than a line of cells, and touch fewer distinct cells.
struct Morton {
  uint32_t bits[3];  // i, j, k
  uint32_t total;
  uint32_t flat;     // cell codes are flat cell indices (polyhedral meshes, numbered in Z order already)
};
__host__ __device__ inline uint32_t morton_encode(uint32_t i, uint32_t j, uint32_t k, Morton m) {
  uint32_t code = 0, pos = 0;
  for (uint32_t b = 0; pos < m.total; ++b) {
    if (b < m.bits[2]) code |= ((k >> b) & 1u) << pos++;
    if (b < m.bits[1]) code |= ((j >> b) & 1u) << pos++;
    if (b < m.bits[0]) code |= ((i >> b) & 1u) << pos++;
  }
  return code;
}
__host__ __device__ inline void morton_ijk(uint32_t code, Morton m, uint32_t &i, uint32_t &j, uint32_t &k) {
  uint32_t pos = 0;
  i = j = k = 0;
  for (uint32_t b = 0; pos < m.total; ++b) {
    if (b < m.bits[2]) k |= ((code >> pos++) & 1u) << b;
    if (b < m.bits[1]) j |= ((code >> pos++) & 1u) << b;
    if (b < m.bits[0]) i |= ((code >> pos++) & 1u) << b;
  }
}
__host__ __device__ inline uint32_t morton_cell(uint32_t code, Morton m, MeshDims d) {
  uint32_t i, j, k;
  morton_ijk(code, m, i, j, k);
  return (i * d.ny + j) * d.nz + k;
}

// Photon counts in Z order: out[code] for every code below 2^total (0 where
// the code is outside the mesh), so that emission is laid out like the
// combed census.
__global__ __launch_bounds__(256) void k_morton_counts(uint32_t M, Morton mo, MeshDims d,
                                                       const uint32_t *__restrict__ nphot,
                                                       uint32_t *__restrict__ out) {
  const uint32_t code = blockIdx.x * blockDim.x + threadIdx.x;
  if (code > M) return;
  if (code == M) {  // a zero past the last code: the exclusive scan of M + 1 counts ends in the total
    out[M] = 0u;
    return;
  }
  uint32_t i, j, k;
  morton_ijk(code, mo, i, j, k);
  out[code] = (i < d.nx && j < d.ny && k < d.nz) ? nphot[(i * d.ny + j) * d.nz + k] : 0u;
}


// ------------------------------------------------------------------ preStep
template <typename Real>
__global__ __launch_bounds__(256) void k_prestep(uint32_t N, const double *__restrict__ T, MatDev mat,
                                                 const double *__restrict__ host_sa,
                                                 const double *__restrict__ host_ss,
                                                 const double *__restrict__ host_cv,
                                                 const double *__restrict__ V, double dt,
                                                 CellXS<Real> *__restrict__ xs, double *__restrict__ fleck,
                                                 double *__restrict__ sig_a_out, double *__restrict__ sig_s_out,
                                                 double *__restrict__ Eemit, uint32_t *__restrict__ err) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double Tc = T[c];
  const double sa = host_sa ? host_sa[c] : mat.sigma_a(c, Tc);
  const double ss = host_ss ? host_ss[c] : mat.sigma_s(c, Tc);
  const double cv = host_cv ? host_cv[c] : mat.dT2cv(mat.density(c), Tc);
  const double T3 = Tc * Tc * Tc;
  const double f = 1.0 / (1.0 + (4.0 * units::arad * T3 * sa * units::clight * dt) / cv);
  if (!(f >= 0.0 && f <= 1.0)) atomicOr(err, kErrFleck);
  fleck[c] = f;
  sig_a_out[c] = sa;
  sig_s_out[c] = ss;
  Eemit[c] = f * V[c] * units::arad * (T3 * Tc) * sa * dt * units::clight;
  xs[c] = make_xs<Real>(static_cast<Real>(f * sa), static_cast<Real>(ss + (1.0 - f) * sa));
}

// Multigroup preStep per cell (storm_group_desc): the Planck opacity
// sum_g sigma(Ec_g) B_g / sum_g B_g (DensmoreOpacity::CalcPlanckOpacity,
// examples/densmore2012/DensmoreOpacity.hpp:38-55) feeds the Fleck factor and
// the emission energy; the normalised cumulative of sigma(Ec_g) B_g is the
// emission spectrum (:66-77, :128-148); the cell table holds f sigma_a and
// (1-f) sigma_a at 1 keV, scaled per packet by E^-nu_pow.
template <typename Real>
__global__ __launch_bounds__(256) void k_prestep_mg(
    uint32_t N, const double *__restrict__ T, MatDev mat, const double *__restrict__ host_sa,
    const double *__restrict__ host_ss, const double *__restrict__ host_cv, const double *__restrict__ V, double dt,
    int32_t G, const double *__restrict__ bounds, double nu_pow, double nu_ref, CellXS<Real> *__restrict__ xs,
    Real *__restrict__ cdf, Real *__restrict__ sig_s_cell, double *__restrict__ fleck,
    double *__restrict__ sig_a_out, double *__restrict__ sig_s_out, double *__restrict__ Eemit,
    uint32_t *__restrict__ err, const double *__restrict__ chord_c, double rw_cell_tau,
    CellAccel<Real> *__restrict__ rw_out, double ddmc_pgrw_tau, double *__restrict__ Dtil) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double Tc = T[c];
  const double sa_gray = host_sa ? host_sa[c] : mat.sigma_a(c, Tc);
  const double ss = host_ss ? host_ss[c] : mat.sigma_s(c, Tc);
  const double cv = host_cv ? host_cv[c] : mat.dT2cv(mat.density(c), Tc);
  const double kT = units::k_boltz * fmax(Tc, mat.Tfloor);
  Real *my = cdf + static_cast<size_t>(c) * static_cast<size_t>(G + 1);
  double acc = 0.0, sumB = 0.0;
  // partial-gray random walk (precomputeRandomWalkData, RadiationIMC.hpp:1008-1060):
  // the groups from the lowest up while each is optically thick over the
  // cell's mean chord; Planck-weighted sums over them
  const double chord = rw_out ? chord_c[c] : 0.0;
  const double kTrw = units::k_boltz * Tc;
  double allAB = 0.0, dB = 0.0, dBA = 0.0, dBT = 0.0, dBoT = 0.0;
  uint32_t cut = 0;
  bool thin = false;
  // partial-gray DDMC (ddmcUseMultigroupPGRW, RadiationIMC.hpp:1301-1342):
  // the diffusive groups from the lowest up with equal weights 1/G
  double qB = 0.0, qBA = 0.0, qBT = 0.0, qBoT = 0.0;
  uint32_t qcut = 0;
  bool qthin = !(ddmc_pgrw_tau >= 0.0) || !(Tc > 0.0);
  my[0] = Real(0);
  for (int32_t g = 0; g < G; ++g) {
    const double a = bounds[g] / kT, b = bounds[g + 1] / kT;
    const double Bg = planck::integral(a, b);
    const double Ec = 0.5 * (bounds[g] + bounds[g + 1]);
    const double sg = sa_gray * pow(fmax(Ec, bounds[0]) / nu_ref, -nu_pow);
    const double wg = (sg > 0.0 && isfinite(sg) && isfinite(Bg)) ? sg * Bg : 0.0;
    sumB += Bg;
    acc += wg;
    my[g + 1] = static_cast<Real>(acc);
    if (rw_out) {
      const double ar = bounds[g] / kTrw, br = bounds[g + 1] / kTrw;
      const double Br = (ar > 0.0 && br > ar) ? planck::integral(ar, br) : 1.0 / static_cast<double>(G);
      const double st = sg + ss;
      allAB += sg * Br;
      if (!thin && st * chord >= rw_cell_tau) {
        cut = static_cast<uint32_t>(g) + 1u;
        dB += Br;
        dBA += Br * sg;
        dBT += Br * st;
        if (st > 0.0) dBoT += Br / st;
      } else {
        thin = true;
      }
      if (!qthin && st * chord >= ddmc_pgrw_tau) {
        const double Bq = 1.0 / static_cast<double>(G);
        qcut = static_cast<uint32_t>(g) + 1u;
        qB += Bq;
        qBA += Bq * sg;
        qBT += Bq * st;
        if (st > 0.0) qBoT += Bq / st;
      } else {
        qthin = true;
      }
    }
  }
  const bool valid = acc > 0.0 && isfinite(acc);
  for (int32_t g = 1; g <= G; ++g) {
    my[g] = valid ? static_cast<Real>(static_cast<double>(my[g]) / acc)
                  : static_cast<Real>((bounds[g] - bounds[0]) / (bounds[G] - bounds[0]));
  }
  my[G] = Real(1);
  const double sP = sumB > 0.0 ? acc / sumB : 1e-20;
  const double T3 = Tc * Tc * Tc;
  const double f = 1.0 / (1.0 + (4.0 * units::arad * T3 * sP * units::clight * dt) / cv);
  if (!(f >= 0.0 && f <= 1.0)) atomicOr(err, kErrFleck);
  fleck[c] = f;
  sig_a_out[c] = sP;
  sig_s_out[c] = ss;
  Eemit[c] = f * V[c] * units::arad * (T3 * Tc) * sP * dt * units::clight;
  const double s1 = sa_gray * pow(units::kev / nu_ref, -nu_pow);  // sigma_a at 1 keV
  xs[c] = make_xs<Real>(static_cast<Real>(f * s1), static_cast<Real>((1.0 - f) * s1));
  sig_s_cell[c] = static_cast<Real>(ss);
  if (rw_out) {
    CellAccel<Real> r;
    r.total = Real(0);
    for (int q = 0; q < 6; ++q) r.rate[q] = Real(0);
    r.flags = 0u;
    r.gcut = 0u;
    r.sig_t = r.sig_d = r.sig_abs = r.up_rate = Real(0);
    r.e_cut = static_cast<Real>(INFINITY);
    r.ddmc_sig_t = r.ddmc_abs = Real(0);
    r.ddmc_e_cut = static_cast<Real>(INFINITY);
    if (qcut > 0u && qB > 0.0) {
      const double qT = qBT / qB, qD = qBoT / (3.0 * qB);
      r.ddmc_sig_t = static_cast<Real>(qT);
      r.ddmc_abs = static_cast<Real>(f * (qBA / qB));
      r.ddmc_e_cut = static_cast<Real>(bounds[qcut] / units::kev);
      Dtil[c] = qD;
      if (qT > 0.0 && qD > 0.0) r.flags |= 4u;
    } else if (ddmc_pgrw_tau >= 0.0) {
      Dtil[c] = 0.0;
    }
    if (cut > 0u && dB > 0.0) {
      const double sA = dBA / dB, sT = dBT / dB;
      const double gamma = allAB > 0.0 ? dBA / allAB : 1.0;
      r.flags |= 1u;
      r.gcut = cut;
      r.sig_t = static_cast<Real>(sT);
      r.sig_d = static_cast<Real>(dBoT > 0.0 ? dB / dBoT : 0.0);
      r.sig_abs = static_cast<Real>(f * sA);
      r.up_rate = static_cast<Real>((gamma < 1.0 && sA > 0.0 && f > 0.0) ? (1.0 - f) * sA * (1.0 - gamma) : 0.0);
      r.e_cut = static_cast<Real>(bounds[cut] / units::kev);
    }
    rw_out[c] = r;
  }
}

// Energy quantum exponents of a step: every energy the step can move (the
// census energy of the last step, this step's emission total and the boundary
// source total) is below 2^62 quanta; q[0] for energies, q[1] for path
// integrals (energy x c dt).  The host reads them back.
__global__ void k_quantum(const double *__restrict__ Etot, double census_prev, double source_total, double c,
                          double dt, int32_t *__restrict__ q) {
  if (threadIdx.x != 0u || blockIdx.x != 0u) return;
  const double Ebound = census_prev + *Etot + source_total;
  q[0] = Ebound > 0.0 ? fx::ceil_log2(Ebound) - fx::kHeadroomBits : -1000;
  q[1] = Ebound > 0.0 ? fx::ceil_log2(Ebound * c * dt) - fx::kHeadroomBits : -1000;
}

// Batches of independent instances (inst.cell non-null).  Each instance counts energy in its own unit, 2^-shift erg (shift >= 0, set
// every step by k_inst_quantum): packet weights and integer tallies of its
// cells are in that unit, so the step's one quantum 2^eq is 2^(eq - shift)
// erg for it, the quantum a stand-alone run of the instance would use.  The
// transport is linear in the weights and never mixes instances, so it runs
// unchanged on the scaled weights.
struct InstanceView {
  const uint32_t *cell;      // instance of each cell (null: one problem)
  const uint32_t *ncells;    // per instance
  const uint64_t *nphot;     // newPhotonsPerCell per instance
  const uint64_t *nmin;      // comb Nmin per instance
  const double *factor;      // comb totalParticlesFactor per instance
  const double *etot;        // emission total per instance, erg (k_inst_sum)
  const double *wtot;        // census weight per instance, in its unit (k_inst_census)
  const int32_t *shift;      // unit exponent of this step, per instance
  const int32_t *shift_prev; // unit exponent of the last step (the census' unit)
};

// Quantum and unit exponents of a step of a batch: instance k's energy bound
// E_k (last census + emission + boundary source, erg) gives its own exponent
// e_k = ceil(log2 E_k) - 62, the one k_quantum gives a stand-alone problem;
// the step's quantum is eq = max_k e_k and instance k's unit 2^-(eq - e_k)
// erg, so every instance has E_k below 2^62 quanta.  The path-integral
// quantum is eq + ceil(log2(c dt)).  The last step's shifts move to
// shift_prev first.  One block.
__global__ __launch_bounds__(256) void k_inst_quantum(uint32_t K, const double *__restrict__ census,
                                                      const double *__restrict__ etot,
                                                      const double *__restrict__ src, double cdt,
                                                      int32_t *__restrict__ shift, int32_t *__restrict__ shift_prev,
                                                      int32_t *__restrict__ q) {
  __shared__ int32_t sh[256];
  int32_t m = -1000;
  for (uint32_t k = threadIdx.x; k < K; k += 256u) {
    shift_prev[k] = shift[k];
    const double E = census[k] + etot[k] + src[k];
    if (E > 0.0) {
      const int32_t e = fx::ceil_log2(E) - fx::kHeadroomBits;
      m = e > m ? e : m;
    }
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (uint32_t w = 128u; w > 0u; w >>= 1) {
    if (threadIdx.x < w && sh[threadIdx.x + w] > sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + w];
    __syncthreads();
  }
  const int32_t eq = sh[0];
  for (uint32_t k = threadIdx.x; k < K; k += 256u) {
    const double E = census[k] + etot[k] + src[k];
    shift[k] = E > 0.0 ? eq - (fx::ceil_log2(E) - fx::kHeadroomBits) : 0;
  }
  if (threadIdx.x == 0) {
    q[0] = eq;
    q[1] = eq == -1000 ? -1000 : eq + fx::ceil_log2(cdt);
  }
}

// photons per cell and packet weight, rounded to the step quantum without bias.
// Batches: each instance allocates over its own cells with its own
// newPhotonsPerCell and emission total, weights in its unit.
__global__ __launch_bounds__(256) void k_alloc(uint32_t N, const double *__restrict__ Eemit,
                                               const double *__restrict__ Etot_p, uint64_t n_min_all,
                                               const int32_t *__restrict__ q, uint32_t *__restrict__ nphot,
                                               double *__restrict__ wemit,
                                               unsigned long long *__restrict__ emit_limbs, InstanceView inst,
                                               uint32_t step) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double Etot = *Etot_p;
  uint64_t Ncells = N, n_min = n_min_all;
  if (inst.cell) {
    const uint32_t k = inst.cell[c];
    Etot = inst.etot[k];
    Ncells = inst.ncells[k];
    n_min = inst.nphot[k];
  }
  const int32_t eq = q[0];
  const double E = Eemit[c];
  const uint64_t total = Ncells * n_min * 10ull;
  uint64_t share = Etot > 0.0 ? static_cast<uint64_t>(E / Etot * static_cast<double>(total)) : n_min;
  uint64_t n = 0;
  if (E > 0.0) {
    n = share < 20ull * n_min ? share : 20ull * n_min;
    if (n < n_min) n = n_min;
  }
  double w = 0.0;
  if (n) {
    const double per = E / static_cast<double>(n);
    w = fx::emission_round(inst.cell ? ldexp(per, inst.shift[inst.cell[c]]) : per, eq, c, step);
  }
  if (w == 0.0) n = 0;  // below one quantum: nothing to emit
  nphot[c] = static_cast<uint32_t>(n);
  wemit[c] = w;
  fx::Limbs L = n ? fx::limbs_of(w, eq) : fx::Limbs{0, 0};
  emit_limbs[2 * c] = L.lo * n;
  emit_limbs[2 * c + 1] = L.hi * n;
}

// ------------------------------------------------------------------ sourcing
template <typename Real>
__device__ __forceinline__ Real uni(uint32_t x) {
  return MathPolicy<Real>::uniform(x);
}


// Census re-entry: timeLeft = dt, initialWeight = |w| (MonteCarloManagerSerial.hpp:393-405).
// Weights are re-quantised if the quantum grew; the remainder (in the old
// quantum) goes to the material.
// Batches: the census is in its instances' units of the last step; a weight
// moves to this step's unit (x 2^(shift - shift_prev), exact) after the
// re-quantisation, which runs in the old unit (new quantum 2^(eq_new - delta)).
template <typename Real>
__global__ __launch_bounds__(256) void k_import(uint32_t n, CensusSoA<Real> cen, PacketSoA<Real> out,
                                                Real s0, int32_t eq_old, int32_t eq_new, MeshDims d,
                                                Tallies t, unsigned long long *__restrict__ requant,
                                                InstanceView inst) {
  // no early return: the per-cell sums below need every lane of the wavefront
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  uint32_t c = 0xFFFFFFFFu;
  fx::Limbs L{0u, 0u};
  if (valid) {
    Packet<Real> p;
    auto a = cen.pos[i];
    auto b = cen.dir[i];
    uint2 m = cen.meta[i];
    p.u[0] = a.x;
    p.u[1] = a.y;
    p.u[2] = a.z;
    p.s = s0;
    p.o[0] = b.x;
    p.o[1] = b.y;
    p.o[2] = b.z;
    double w = cen.w[i];
    c = cell_of_code(m.x, d);
    int32_t delta = 0;
    if (inst.cell) {
      const uint32_t k = inst.cell[c];
      delta = inst.shift[k] - inst.shift_prev[k];
    }
    if (eq_new - delta > eq_old) {
      double wq = fx::quantize_down(w, eq_new - delta);
      if (wq != w) tally_limbs(requant, c, fx::limbs_of(w - wq, eq_old));  // exact in the old quantum
      w = wq;
    }
    if (delta) w = ldexp(w, delta);
    p.w = w;
    p.wmin = static_cast<Real>(w * 1e-3);
    p.ijk = m.x;
    p.id = i;
    p.ctr = 0u;
    p.flags = 0u;
    p.nu = a.w;
    store_packet(out, i, p);
    L = fx::limbs_of(w, eq_new);
  }
  // the census comes sorted by cell after population control: one atomic per run
  unsigned long long lo = L.lo, hi = L.hi;
  uint32_t one = valid ? 1u : 0u;
  wave_run_sum(c, lo);
  wave_run_sum(c, hi);
  if (wave_run_sum(c, one) && valid) {
    if (lo) atomicAdd(&t.cin[2 * c], lo);
    if (hi) atomicAdd(&t.cin[2 * c + 1], hi);
    atomicAdd(&t.begin[c], one);
  }
}

// Host-made new packets (BoundaryCondition::generateNewBoundaryParticles,
// radiation/RadiationIMC.hpp:1683-1692): one record each, position in the
// cell's frame (Cartesian: fractions of the cell; polyhedral: relative to its
// mesh point), unit direction, weight already on the step's quantum, path
// length left (c * timeLeft), photon energy (keV).
struct HostPacket {
  double u[3];
  double o[3];
  double w;
  double s;
  float nu;
  uint32_t cell;  // flat cell index
  uint32_t code;  // the packet's cell code (packed ijk, or the flat index)
  uint32_t pad;
};

template <typename Real>
__global__ __launch_bounds__(256) void k_host_packets(uint32_t n, const HostPacket *__restrict__ in,
                                                      PacketSoA<Real> out, uint32_t base,
                                                      uint32_t *__restrict__ begin) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < n;
  uint32_t cell = 0xFFFFFFFFu;
  if (valid) {
    const HostPacket h = in[q];
    Packet<Real> p;
    for (int a = 0; a < 3; ++a) {
      p.u[a] = static_cast<Real>(h.u[a]);
      p.o[a] = static_cast<Real>(h.o[a]);
    }
    p.s = static_cast<Real>(h.s);
    p.w = h.w;
    p.wmin = static_cast<Real>(h.w * 1e-3);
    p.ijk = h.code;
    p.id = base + q;
    p.ctr = 0u;
    p.flags = 0u;
    p.nu = static_cast<Real>(h.nu);
    store_packet(out, base + q, p);
    cell = h.cell;
  }
  uint32_t one = valid ? 1u : 0u;
  if (wave_run_sum(cell, one) && valid) atomicAdd(&begin[cell], one);
}

// Emission packets: uniform position in the cell (RandomInCellPositionSampler
// is uniform in the box), isotropic direction, timeLeft = dt * U.
template <typename Real>
// Emission packets in Z order of their cells: `offsets` is the exclusive scan
// of the photon counts over the M = 2^total codes (k_morton_counts).
__global__ __launch_bounds__(256) void k_emit(uint32_t n_emit, uint32_t M, const uint64_t *__restrict__ offsets,
                                              const double *__restrict__ wemit, PacketSoA<Real> out,
                                              uint32_t base, Real s0, MeshDims d, Morton mo, uint64_t seed,
                                              uint32_t step, uint32_t *__restrict__ begin, SpectrumView<Real> sp) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < n_emit;
  uint32_t c = 0xFFFFFFFFu;
  if (valid) {
    // code owning packet q: last code with offsets[code] <= q
    uint32_t lo = 0, hi = M;  // offsets[M] = n_emit
    while (hi - lo > 1u) {
      uint32_t mid = (lo + hi) >> 1;
      if (offsets[mid] <= q) lo = mid;
      else hi = mid;
    }
    c = morton_cell(lo, mo, d);
    Packet<Real> p;
    birth_emission<Real>(p, base + q, ijk_of_flat(c, d), wemit[c], s0, seed, step, sp, c);
    store_packet(out, base + q, p);
  }
  uint32_t one = valid ? 1u : 0u;
  if (wave_run_sum(c, one) && valid) atomicAdd(&begin[c], one);
}

// Planck surface source on the x- box face: per face Npc packets of weight
// sigma_SB T^4 A dt / Npc, uniform on the face, mu = sqrt(U) about +x.
template <typename Real>
// Faces f = s * ny * nz + j * nz + k for the source sides s (0 .. nsides-1);
// xhi_from: the first side index that is the x+ face (nsides if none);
// bcdf holds groups + 1 entries per side.
__global__ __launch_bounds__(256) void k_boundary(uint32_t n_bc, uint32_t npc, const double *__restrict__ wface,
                                                  PacketSoA<Real> out, uint32_t base, Real s0, MeshDims d,
                                                  uint64_t seed, uint32_t step, uint32_t *__restrict__ begin,
                                                  SpectrumView<Real> sp, const Real *__restrict__ bcdf,
                                                  uint32_t xhi_from) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < n_bc;
  uint32_t cell = 0xFFFFFFFFu;
  if (valid) {
    const uint32_t f = q / npc;
    const uint32_t per_side = d.ny * d.nz;
    const uint32_t side = f / per_side, jk = f % per_side;
    const uint32_t j = jk / d.nz, k = jk % d.nz;
    const bool xhi = side >= xhi_from;
    const uint32_t i = xhi ? d.nx - 1u : 0u;
    cell = (i * d.ny + j) * d.nz + k;
    Packet<Real> p;
    birth_boundary<Real>(p, base + q, xhi, i, j, k, wface[f], s0, seed, step, sp,
                         bcdf ? bcdf + static_cast<size_t>(side) * static_cast<size_t>(sp.groups + 1) : nullptr);
    store_packet(out, base + q, p);
  }
  uint32_t one = valid ? 1u : 0u;
  if (wave_run_sum(cell, one) && valid) atomicAdd(&begin[cell], one);
}

// ------------------------------------------------------------------ transport
// Packets enter the transport ordered by cell (the census comes sorted from
// population control, emission is generated cell by cell).  Both transport
// kernels keep that order: a block takes a contiguous range of the input, so
// its visits fall in few cells, which its LDS tally cache exploits.

// Block-local cache of the per-visit tallies in LDS.  Global atomics on
// device memory are performed at the memory controller (every one a DRAM
// read-modify-write), so a visit's three atomics are the transport's most
// expensive operation.  The packets of one block start in neighbouring cells
// (the input is ordered by cell) and in optically thick media stay close for
// many segments, so a block's visits fall in few cells: they are summed in a
// small open-addressed table in LDS (cell -> sums) and written to device
// memory once per cell when the block's last wavefront ends.  A visit whose
// cell finds no slot within kCacheProbes probes goes straight to device
// memory.  Integer sums are exact in any order; the FP64 engine's sums
// (doubles in v[0], v[1]) change only their order.
// The table lives in dynamic LDS: `slots` (a power of two) slots of `nf`
// 64-bit sums, a tag and a segment count each.  nf = 1 in the mixed gray
// engine (the deposit, in quanta; the path integral, tallied only in cells
// without absorption, goes to device memory), 2 in the mixed multigroup
// engine (deposit and path integral), 2 in the FP64 engine (deposit and
// path integral as doubles).  The host sizes it to what a block's share of
// the CU's 160 KB of LDS holds (cache_slots).
constexpr uint32_t kCacheProbes = 2;
inline __host__ __device__ uint32_t cache_bytes(uint32_t slots, uint32_t nf) {
  return slots * (8u * nf + 8u) + 16u;
}
// Largest power-of-two slot count within 4 per packet of the block's range
// and the LDS share of one block when `blocks_per_cu` blocks are resident.
inline uint32_t cache_slots(uint32_t packets, uint32_t nf, uint32_t blocks_per_cu) {
  const uint32_t budget = (160u * 1024u) / (blocks_per_cu ? blocks_per_cu : 1u) - 64u;
  uint32_t s = 1;
  while (2 * s <= 4 * packets && cache_bytes(2 * s, nf) <= budget) s *= 2;
  return s;
}
struct TallyCache {
  uint32_t bits, mask, nf;
  unsigned long long *v;  // v[f * slots + s], f < nf
  uint32_t *tag;          // cell + 1; 0: free
  uint32_t *segs;
  uint32_t *waves_done;
  uint32_t *next_chunk;   // k_transport's chunks handed out so far
  uint32_t *staged;       // k_transport_refill: census packets staged by the block
  __device__ __forceinline__ unsigned long long &at(uint32_t f, uint32_t s) const { return v[f * (mask + 1u) + s]; }
};

// Lays the table out in `lds` (cache_bytes(slots, nf) bytes) and clears it;
// the caller synchronises the block before first use.
__device__ __forceinline__ TallyCache cache_init(unsigned char *lds, uint32_t slots, uint32_t nf) {
  TallyCache L;
  L.bits = 31u - static_cast<uint32_t>(__clz(static_cast<int>(slots)));
  L.mask = slots - 1u;
  L.nf = nf;
  L.v = reinterpret_cast<unsigned long long *>(lds);
  L.tag = reinterpret_cast<uint32_t *>(lds + 8 * nf * slots);
  L.segs = L.tag + slots;
  L.waves_done = L.segs + slots;
  L.next_chunk = L.waves_done + 1;
  L.staged = L.waves_done + 2;
  for (uint32_t s = threadIdx.x; s < slots; s += blockDim.x) {
    L.tag[s] = 0u;
    L.segs[s] = 0u;
    for (uint32_t f = 0; f < nf; ++f) L.at(f, s) = 0ull;
  }
  if (threadIdx.x == 0) {
    *L.waves_done = 0u;
    *L.next_chunk = 0u;
    *L.staged = 0u;
  }
  return L;
}

// Adds into the table (workgroup scope: the table is the block's own).
template <typename T>
__device__ __forceinline__ void lds_add(T *p, T v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Slot of cell c (claimed if free), or -1 when the probed slots hold other cells.
__device__ __forceinline__ int cache_slot(uint32_t *tag, uint32_t bits, uint32_t mask, uint32_t c) {
  const uint32_t key = c + 1u;
  const uint32_t h = (c * 0x9E3779B1u) >> (32u - bits);
  int found = -1;
  for (uint32_t probe = 0; probe < kCacheProbes && found < 0; ++probe) {
    const uint32_t s = (h + probe) & mask;
    uint32_t cur = __hip_atomic_load(&tag[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == 0u) cur = atomicCAS(&tag[s], 0u, key);
    if (cur == 0u || cur == key) found = static_cast<int>(s);
  }
  return found;
}

struct DeviceSink {
  Tallies t;
  int32_t eq;
  uint32_t rw = 0, ddmc = 0;
  // this wavefront's replica of the per-visit tallies
  unsigned long long *dep, *pint;
  double *dep_fp, *pint_fp;
  uint32_t *segs;
  TallyCache cache{};  // the block's LDS cache (cache.v null: none)
  __device__ DeviceSink(const Tallies &tt, int32_t e) : t(tt), eq(e) {
    // the integer replicas exist in the mixed engine, the FP64 ones in the
    // FP64 engine (the other pair of pointers is null)
    const uint32_t r = xcd_index() % kTallyReplicas;
    const size_t n = tt.ncell;
    dep = (r && tt.dep_rep) ? tt.dep_rep + (r - 1) * 2 * n : tt.dep;
    pint = (r && tt.pint_rep) ? tt.pint_rep + (r - 1) * 2 * n : tt.pint;
    segs = r ? tt.segs_rep + (r - 1) * n : tt.segs;
    dep_fp = (r && tt.dep_fp_rep) ? tt.dep_fp_rep + (r - 1) * n : tt.dep_fp;
    pint_fp = (r && tt.pint_fp_rep) ? tt.pint_fp_rep + (r - 1) * n : tt.pint_fp;
  }
  // Whole values (below 2^62 quanta, as is every sum of them) into the low
  // word of the cell's pair: one atomic per tally.
  __device__ void visit(uint32_t c, fx::Limbs d, fx::Limbs e, uint32_t n) {
    const unsigned long long dv = fx::value_of(d), ev = fx::value_of(e);
    const int s = cache.v ? cache_slot(cache.tag, cache.bits, cache.mask, c) : -1;
    if (s >= 0) {
      if (dv) lds_add(&cache.at(0, s), dv);
      if (cache.nf >= 2) {
        if (ev) lds_add(&cache.at(1, s), ev);
      } else if (ev) {
        atomicAdd(&pint[2 * c], ev);
      }
      lds_add(&cache.segs[s], n);
      return;
    }
    if (dv) atomicAdd(&dep[2 * c], dv);
    if (ev) atomicAdd(&pint[2 * c], ev);
    atomicAdd(&segs[c], n);
  }
  __device__ void roulette(uint32_t c, double w) { tally_limbs(t.rdep, c, fx::limbs_of(w, eq)); }
  __device__ void visit_fp(uint32_t c, double d, double e, uint32_t n) {
    const int s = cache.v ? cache_slot(cache.tag, cache.bits, cache.mask, c) : -1;
    if (s >= 0) {
      if (d != 0.0) lds_add(reinterpret_cast<double *>(&cache.at(0, s)), d);
      if (e != 0.0) lds_add(reinterpret_cast<double *>(&cache.at(1, s)), e);
      lds_add(&cache.segs[s], n);
      return;
    }
    if (d != 0.0) atomicAdd(&dep_fp[c], d);
    if (e != 0.0) atomicAdd(&pint_fp[c], e);
    atomicAdd(&segs[c], n);
  }
  // Called by every wavefront of the block once its packets are done: the
  // last one writes the cache to device memory (and returns true).
  __device__ bool finish_block() {
    if (!cache.v) return false;
    uint32_t done = 0u;
    if (lane_rank(__builtin_amdgcn_ballot_w64(true)) == 0u)
      done = __hip_atomic_fetch_add(cache.waves_done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    done = __builtin_amdgcn_readfirstlane(done);
    if (done + 1u != (blockDim.x + 63u) / 64u) return false;
    for (uint32_t s = __lane_id(); s <= cache.mask; s += 64u) {
      const uint32_t key = __hip_atomic_load(&cache.tag[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (!key) continue;
      const uint32_t c = key - 1u;
      if (dep_fp) {
        const double d = __longlong_as_double(static_cast<long long>(cache.at(0, s)));
        const double e = __longlong_as_double(static_cast<long long>(cache.at(1, s)));
        if (d != 0.0) atomicAdd(&dep_fp[c], d);
        if (e != 0.0) atomicAdd(&pint_fp[c], e);
      } else {
        if (cache.at(0, s)) atomicAdd(&dep[2 * c], cache.at(0, s));
        if (cache.nf >= 2 && cache.at(1, s)) atomicAdd(&pint[2 * c], cache.at(1, s));
      }
      if (cache.segs[s]) atomicAdd(&segs[c], cache.segs[s]);
    }
    return true;
  }
  // escaped weight: in quanta (mixed engine) or as a double (FP64 engine)
  __device__ void escape_fx(uint32_t c, double w) { tally_limbs(t.esc, c, fx::limbs_of(w, eq)); }
  __device__ void escape_fp(uint32_t c, double w) { atomicAdd(&t.esc_fp[c], w); }
  __device__ void rw_step() { ++rw; }
  __device__ void ddmc_step() { ++ddmc; }
  // path integral v (erg cm) of cell c in group g: rounded without bias to
  // the path-integral quantum in the mixed engine (u24: the draw)
  __device__ void group_path(uint32_t c, int32_t g, double v, const StepConsts &k, uint32_t u24) {
    const size_t i = static_cast<size_t>(c) * t.eg_groups + static_cast<uint32_t>(g);
    if (t.eg_fp) {
      atomicAdd(&t.eg_fp[i], v);
    } else if (t.eg) {
      const double q = fx::quantize_unbiased(v, k.et, u24);
      if (q > 0.0) atomicAdd(&t.eg[i], static_cast<unsigned long long>(fx::value_of(fx::limbs_of(q, k.et))));
    }
  }
};

// Sum over the active lanes of a wavefront, added once to *dst.
__device__ __forceinline__ void wave_add_u32(uint32_t v, unsigned long long *dst) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (v && lane_rank(__builtin_amdgcn_ballot_w64(true)) == 0u) atomicAdd(dst, static_cast<unsigned long long>(v));
}

// One launch over n_in packets.  Block b owns the packets
// [b * chunks * 64, (b + 1) * chunks * 64) and hands them to its wavefronts
// 64 at a time (one packet per lane, tracked until it ends or has taken
// `max_segments` segments); a wavefront whose packets ended takes the next
// chunk.  The packets come ordered by cell, so the block's range is a compact
// region and its LDS tally table serves all of it; each cell it touched is
// written to device memory once, when the block's last wavefront ends.
// stats: [0] escapes, [1] random-walk steps, [2] DDMC steps, [4] longest history
#define STORM_TRANSPORT_ARGS                                                                                  \
  uint32_t n_in, PacketSoA<Real> in, PacketSoA<Real> out, uint32_t *__restrict__ n_out, CensusSoA<Real> cen,  \
      uint32_t *__restrict__ n_cen, MeshT mesh, const CellXS<Real> *__restrict__ xs,                          \
      AccelView<Real> acc, StepConsts k, int32_t max_segments, Tallies t,                                     \
      unsigned long long *__restrict__ stats, uint32_t cache_slots, uint32_t cache_fields, uint32_t chunks
template <typename Real, uint32_t F, typename MeshT>
__device__ __forceinline__ void transport_chunks(STORM_TRANSPORT_ARGS) {
  extern __shared__ __align__(16) unsigned char lds[];
  const TallyCache cache = cache_init(lds, cache_slots, cache_fields);
  __syncthreads();
  DeviceSink sink(t, k.eq);
  sink.cache = cache;
  const uint32_t lane = __lane_id();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * chunks * 64u;
  uint32_t escaped = 0, longest = 0;
  for (;;) {
    uint32_t ch = 0;
    if (lane == 0u) ch = __hip_atomic_fetch_add(cache.next_chunk, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    ch = __builtin_amdgcn_readfirstlane(ch);
    if (ch >= chunks) break;
    const uint64_t i = base + static_cast<uint64_t>(ch) * 64u + lane;
    const bool valid = i < n_in;
    int outcome = kGone;
    Packet<Real> p;
    if (valid) {
      p = load_packet(in, static_cast<uint32_t>(i));
      outcome = transport_packet<Real, F>(p, mesh, xs, acc, k, max_segments, sink);
    }
    const bool alive = valid && outcome == kAlive;
    const bool census = valid && outcome == kCensus;
    uint32_t slot = wave_append(alive, n_out);
    if (alive) store_packet(out, slot, p);
    uint32_t cslot = wave_append(census, n_cen);
    if (census) store_census(cen, cslot, p);
    escaped += (valid && outcome == kEscaped) ? 1u : 0u;
    if (valid && p.ctr > longest) longest = p.ctr;
    if (base + static_cast<uint64_t>(ch + 1u) * 64u >= n_in) break;  // the last chunk of the input
  }
  sink.finish_block();
  wave_add_u32(escaped, &stats[0]);
  wave_max_u32(longest, &stats[4]);
  if (acc.cell) {
    wave_add_u32(sink.rw, &stats[1]);
    wave_add_u32(sink.ddmc, &stats[2]);
  }
}

// Waves per SIMD the history kernel is compiled for.  FP32: 6 (80 VGPRs, a
// few dozen bytes of spill), measured 17% faster than the compiler's own
// choice of 4 (98-118 VGPRs) on the 64^3 cube with the random walk and 5%
// faster than 5 on the gray cube.  FP64: the compiler's choice (2-3 waves);
// at 6 it spills 200-500 bytes per lane.
template <typename Real>
struct TransportWaves {
  static constexpr int kMin = 1;
};
template <>
struct TransportWaves<float> {
  static constexpr int kMin = 6;
};

template <typename Real, uint32_t F, typename MeshT = MeshView<Real>>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(TransportWaves<Real>::kMin, 8))) void
k_transport(STORM_TRANSPORT_ARGS) {
  transport_chunks<Real, F, MeshT>(n_in, in, out, n_out, cen, n_cen, mesh, xs, acc, k, max_segments, t, stats,
                                   cache_slots, cache_fields, chunks);
}

// Block-local lane refill (the refill schedule).  The block owns the same
// range of the cell-ordered input as k_transport and sums its visits in the
// same LDS table, but each wavefront advances its 64 packets one segment at
// a time: once at least kRefillIdle of its lanes are idle (their packet
// ended or took `max_segments` segments), they take the next packets of the
// block's range, 64 at a time from the block's LDS counter.  A lane then idles only until
// enough others have finished, not until the longest history of its chunk
// ends; a refill of several lanes costs one memory round trip for all of
// them.  Per packet the result is that of k_transport with the same
// `max_segments` (a history does not depend on the lane that runs it).
// Census packets are staged in the block's own part of `stage` (its input
// range: a packet ends in census at most once) through an LDS counter, and
// the block's last wavefront reserves their census slots with one atomic and
// copies them there.  Appending them one or two at a time as lanes finish
// would put an atomic per census packet on the one census counter that every
// wavefront of the launch shares.
// Idle lanes that take new packets together: 2, 4 and 16 measured within 1%
// of each other (cube, step slab, batch of slabs).
constexpr uint32_t kRefillIdle = 4;
template <typename Real, uint32_t F, typename MeshT>
__device__ __forceinline__ void transport_refill(STORM_TRANSPORT_ARGS, CensusSoA<Real> stage) {
  extern __shared__ __align__(16) unsigned char lds[];
  const TallyCache cache = cache_init(lds, cache_slots, cache_fields);
  __syncthreads();
  DeviceSink sink(t, k.eq);
  sink.cache = cache;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * chunks * 64u;
  const uint64_t lim = base + static_cast<uint64_t>(chunks) * 64u < n_in ? base + static_cast<uint64_t>(chunks) * 64u
                                                                          : static_cast<uint64_t>(n_in);
  const uint32_t span = static_cast<uint32_t>(lim - base);  // packets of this block
  typename TrackerOf<Real, F, MeshT>::type tr;
  bool busy = false;
  int32_t nseg = 0;
  // wave-uniform: the chunk being handed out, [pool_next, pool_end) of the
  // block's range, and whether the block's range is exhausted
  uint32_t pool_next = 0, pool_end = 0;
  bool drained = false;
  uint32_t escaped = 0, longest = 0;
  for (;;) {
    if (!drained) {
      const uint64_t idle = __builtin_amdgcn_ballot_w64(!busy);
      const uint32_t nidle = static_cast<uint32_t>(__popcll(idle));
      if (nidle >= kRefillIdle) {
        const uint32_t r = lane_rank(idle);
        uint32_t given = 0;
        while (given < nidle) {
          if (pool_next >= pool_end) {
            uint32_t ch = 0;
            if (lane_rank(__builtin_amdgcn_ballot_w64(true)) == 0u)
              ch = __hip_atomic_fetch_add(cache.next_chunk, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            ch = __builtin_amdgcn_readfirstlane(ch);
            if (ch >= chunks || ch * 64u >= span) {
              drained = true;
              break;
            }
            pool_next = ch * 64u;
            pool_end = pool_next + 64u < span ? pool_next + 64u : span;
          }
          const uint32_t avail = pool_end - pool_next;
          const uint32_t take = nidle - given < avail ? nidle - given : avail;
          if (!busy && r >= given && r < given + take) {
            const Packet<Real> p = load_packet(in, static_cast<uint32_t>(base + pool_next + (r - given)));
            tr.begin(p, mesh, xs, k);
            busy = true;
            nseg = 0;
          }
          pool_next += take;
          given += take;
        }
      }
    }
    // every lane idle: the block's range was exhausted on this or an earlier round
    if (__builtin_amdgcn_ballot_w64(busy) == 0ull) break;
    int outcome = kAlive;
    if (busy) {
      outcome = tr.segment(mesh, xs, acc, k, sink);
      ++nseg;
    }
    const bool fin = busy && (outcome != kAlive || nseg >= max_segments);
    Packet<Real> p;
    if (fin) {
      tr.end(p, k, sink);
      busy = false;
    }
    const bool census = fin && outcome == kCensus;
    const bool survivor = fin && outcome == kAlive;
    escaped += (fin && outcome == kEscaped) ? 1u : 0u;
    if (fin && p.ctr > longest) longest = p.ctr;
    const uint32_t oslot = wave_append(survivor, n_out);
    if (survivor) store_packet(out, oslot, p);
    const uint64_t want = __builtin_amdgcn_ballot_w64(census);
    if (want) {
      uint32_t first = 0u;
      if (lane_rank(__builtin_amdgcn_ballot_w64(true)) == 0u)
        first = __hip_atomic_fetch_add(cache.staged, static_cast<uint32_t>(__popcll(want)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
      first = __builtin_amdgcn_readfirstlane(first);
      if (census) store_census(stage, static_cast<uint32_t>(base) + first + lane_rank(want), p);
    }
  }
  // the staged census becomes visible to the block's last wavefront
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (sink.finish_block()) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t nst = __hip_atomic_load(cache.staged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    uint32_t g = 0u;
    if (nst && __lane_id() == 0u) g = atomicAdd(n_cen, nst);
    g = __builtin_amdgcn_readfirstlane(g);
    for (uint32_t q = __lane_id(); q < nst; q += 64u) copy_census(stage, static_cast<uint32_t>(base) + q, cen, g + q);
  }
  wave_add_u32(escaped, &stats[0]);
  wave_max_u32(longest, &stats[4]);
  if (acc.cell) {
    wave_add_u32(sink.rw, &stats[1]);
    wave_add_u32(sink.ddmc, &stats[2]);
  }
}

// Waves per SIMD the refill kernel is compiled for.  FP32: 5 (96 VGPRs),
// measured 17% faster than 6 on the cube (whose spills it reloads in its
// per-segment loop: 77 B of HBM traffic per segment at 6) and 9% faster than
// 4; on the step slab 4 and 5 are 8% and 5% faster than 6.  FP64: the
// compiler's choice.
template <typename Real>
struct RefillWaves {
  static constexpr int kMin = 1;
};
template <>
struct RefillWaves<float> {
  static constexpr int kMin = 5;
};

template <typename Real, uint32_t F, typename MeshT = MeshView<Real>>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(RefillWaves<Real>::kMin, 8))) void
k_transport_refill(STORM_TRANSPORT_ARGS, CensusSoA<Real> stage) {
  transport_refill<Real, F, MeshT>(n_in, in, out, n_out, cen, n_cen, mesh, xs, acc, k, max_segments, t, stats,
                                   cache_slots, cache_fields, chunks, stage);
}
#undef STORM_TRANSPORT_ARGS

// ------------------------------------------------------------------ accelerations
// Per-cell random-walk / DDMC data (precomputeRandomWalkData, RadiationIMC.hpp:982-1063;
// precomputeDDMCData, :1282-1413), gray.  Pass 1: sigma_t, D/c = 1/(3 sigma_t), flags.
template <typename Real>
// In a multigroup problem (`walk_set`) the random-walk fields come from
// k_prestep_mg and only the DDMC ones are set here, from the Planck mean
// sigma_P = sig_a (the reference's DDMC without ddmcUseMultigroupPGRW).
__global__ __launch_bounds__(256) void k_accel_cells(uint32_t N, const double *__restrict__ chord_c,
                                                     const double *__restrict__ sig_a,
                                                     const double *__restrict__ sig_s,
                                                     const double *__restrict__ fleck, int32_t rw_on,
                                                     double rw_cell_tau, int32_t ddmc_on, double ddmc_cell_tau,
                                                     int32_t walk_set, int32_t ddmc_pgrw,
                                                     double *__restrict__ Dtil, CellAccel<Real> *__restrict__ out) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double chord = chord_c[c];  // mean chord 4 V / S (RadiationIMC.hpp:1297-1298)
  const double st = sig_a[c] + sig_s[c];
  CellAccel<Real> a;
  uint32_t pre = 0u;
  if (walk_set) {
    a = out[c];
    pre = a.flags;
    a.flags &= 1u;
  } else {
    a.sig_t = static_cast<Real>(st);
    a.flags = 0u;
    a.gcut = 0u;
    a.sig_d = a.sig_t;
    a.sig_abs = Real(0);
    a.up_rate = Real(0);
    a.e_cut = static_cast<Real>(INFINITY);
    if (rw_on && st * chord >= rw_cell_tau) a.flags |= 1u;
  }
  a.total = Real(0);
  for (int f = 0; f < 6; ++f) a.rate[f] = Real(0);
  if (walk_set && ddmc_pgrw) {
    // partial-gray DDMC data and D / c from k_prestep_mg (bit 2: eligible)
    if (ddmc_on && (pre & 4u)) a.flags |= 2u;
    out[c] = a;
    return;
  }
  a.ddmc_sig_t = static_cast<Real>(st);
  a.ddmc_abs = static_cast<Real>(fleck[c] * sig_a[c]);
  a.ddmc_e_cut = static_cast<Real>(INFINITY);
  const double Dt = st > 0.0 ? 1.0 / (3.0 * st) : 0.0;
  Dtil[c] = Dt;
  if (ddmc_on && st * chord >= ddmc_cell_tau && Dt > 0.0) a.flags |= 2u;
  out[c] = a;
}

// Pass 2: leak rate per face, D_f A / (V d) / c with the harmonic mean of
// the two cells' D; faces on the box are skipped (RadiationIMC.hpp:1368-1371).
template <typename Real>
__global__ __launch_bounds__(256) void k_accel_rates(uint32_t N, MeshDims d, MeshView<Real> mv,
                                                     const double *__restrict__ Dtil,
                                                     CellAccel<Real> *__restrict__ out) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  CellAccel<Real> a = out[c];
  if (!(a.flags & 2u)) return;
  const uint32_t ijk = ijk_of_flat(c, d);
  const uint32_t idx[3] = {ijk_i(ijk), ijk_j(ijk), ijk_k(ijk)};
  const uint32_t n[3] = {d.nx, d.ny, d.nz};
  const uint32_t stride[3] = {d.ny * d.nz, d.nz, 1u};
  const double D1 = Dtil[c];
  double total = 0.0;
  for (int f = 0; f < 6; ++f) {
    const int ax = f >> 1;
    const bool up = f & 1;
    a.rate[f] = Real(0);
    if (up ? (idx[ax] + 1u >= n[ax]) : (idx[ax] == 0u)) continue;
    const uint32_t nb = up ? c + stride[ax] : c - stride[ax];
    const double D2 = Dtil[nb];
    double Df = D1;
    if (D2 > 0.0) Df = (D1 > 0.0 && D2 > 0.0) ? 2.0 * D1 * D2 / (D1 + D2) : fmax(D1, D2);
    const double w = width_at(mv, ax, idx[ax])[0];
    const double rate = Df / (w * (0.5 * w));  // A / V = 1 / w, face distance w / 2
    if (rate > 0.0 && isfinite(rate)) {
      a.rate[f] = static_cast<Real>(rate);
      total += rate;
    }
  }
  a.total = static_cast<Real>(total);
  if (!(total > 0.0)) {
    a.flags &= ~2u;
    for (int f = 0; f < 6; ++f) a.rate[f] = Real(0);
  }
  out[c] = a;
}

// ------------------------------------------------------------------ comb
// Bijective 32-bit mix (odd multipliers and xor-shifts are invertible on
// 2^32): gives every census packet a distinct pseudo-random rank within its
// cell, replacing the reference's sort-by-id + std::shuffle
// (CombPopulationControl.hpp:196-202) by a deterministic permutation.
__device__ __forceinline__ uint32_t perm32(uint32_t x, uint32_t salt) {
  x ^= salt;
  x *= 0x9E3779B1u;
  x ^= x >> 16;
  x *= 0x85EBCA77u;
  x ^= x >> 13;
  x *= 0xC2B2AE3Du;
  x ^= x >> 16;
  return x;
}

// Sort keys of the census: the cell's Morton code in the high bits and the
// top `pbits` bits of the packet's rank within its cell below it (32-bit
// keys: four radix passes).  The value carries the full 32-bit rank above
// the packet's index, for k_comb_ties.  Weight-0 packets get code
// 2^total and sort last.
template <typename Real>
// Stratified comb (gbits > 0): the packet's photon group (group_of over
// `bounds`, `groups` of them) sits between the cell code and the rank bits,
// so each cell's run splits into runs of its groups.
__global__ __launch_bounds__(256) void k_comb_keys(uint32_t n, CensusSoA<Real> cen, Morton mo, uint32_t salt,
                                                   uint32_t pbits, uint32_t *__restrict__ keys,
                                                   uint32_t *__restrict__ vals, uint32_t gbits,
                                                   const Real *__restrict__ bounds, int32_t groups) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint2 m = cen.meta[i];
  const uint32_t rank = perm32(m.y, salt);
  vals[i] = i;
  // weight-0 packets are dropped (CombPopulationControl.hpp:136-137)
  const bool drop = cen.w[i] == 0.0;
  const uint32_t code = drop ? (1u << mo.total)
                             : (mo.flat ? m.x : morton_encode(ijk_i(m.x), ijk_j(m.x), ijk_k(m.x), mo));
  uint32_t key = code;
  if (gbits) {
    const Real e = cen.pos[i].w;
    key = (key << gbits) | (drop ? 0u : static_cast<uint32_t>(group_of<Real>(e, bounds, groups)));
  }
  keys[i] = pbits ? (key << pbits) | (rank >> (32u - pbits)) : key;
}

// Orders each run of equal sort keys by the full rank of its packets
// (perm32 of the packet id; ranks are distinct), so that the order inside a
// cell is the order of the full 32-bit ranks whatever order the census was
// written in.  The values are census indices; a packet's rank is recomputed
// from its id where the sort compares it.  Runs are short (a few packets
// share the top rank bits of a cell); long ones, from a cell holding far more
// packets than 2^pbits, are heap-sorted.
__global__ __launch_bounds__(256) void k_comb_ties(uint32_t n, const uint32_t *__restrict__ keys,
                                                   uint32_t *__restrict__ vals, uint32_t dropped_key,
                                                   const uint2 *__restrict__ meta, uint32_t salt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1u >= n) return;
  const uint32_t k = keys[i];
  if (k >= dropped_key || keys[i + 1u] != k || (i > 0u && keys[i - 1u] == k)) return;
  uint32_t j = i + 2u;
  while (j < n && keys[j] == k) ++j;
  uint32_t *v = vals + i;
  auto rank = [&](uint32_t idx) { return perm32(meta[idx].y, salt); };
  const uint32_t L = j - i;
  if (L <= 32u) {
    for (uint32_t a = 1; a < L; ++a) {
      const uint32_t x = v[a];
      const uint32_t rx = rank(x);
      uint32_t b = a;
      while (b > 0u && rank(v[b - 1u]) > rx) {
        v[b] = v[b - 1u];
        --b;
      }
      v[b] = x;
    }
    return;
  }
  // heap sort by rank (ascending)
  auto sift = [&](uint32_t root, uint32_t len) {
    const uint32_t x = v[root];
    const uint32_t rx = rank(x);
    for (;;) {
      uint32_t ch = 2u * root + 1u;
      if (ch >= len) break;
      uint32_t rc = rank(v[ch]);
      if (ch + 1u < len) {
        const uint32_t r1 = rank(v[ch + 1u]);
        if (r1 > rc) {
          ++ch;
          rc = r1;
        }
      }
      if (rc <= rx) break;
      v[root] = v[ch];
      root = ch;
    }
    v[root] = x;
  };
  for (uint32_t r = L / 2u; r-- > 0u;) sift(r, L);
  for (uint32_t e = L - 1u; e > 0u; --e) {
    const uint32_t t = v[0];
    v[0] = v[e];
    v[e] = t;
    sift(0u, e);
  }
}

// Over the sorted census: each weight in sorted order (`ws`, read by the
// later comb kernels in place of a gather through the sort's values) and in
// quanta (`wq`, for the prefix sums), and each cell's run [start, end) of
// sorted indices (end stays 0 for a cell with no packet).
// The comb's unit of a sorted key: the cell, or with gbits > 0 (the
// stratified comb) the (cell, group) bin, numbered cell * groups + group.
struct CombUnits {
  uint32_t pbits, gbits, groups;
  __device__ uint32_t unit(uint32_t key, const Morton &mo, const MeshDims &d, uint32_t &cell) const {
    const uint32_t code = key >> (pbits + gbits);
    cell = morton_cell(code, mo, d);
    if (!gbits) return cell;
    return cell * groups + ((key >> pbits) & ((1u << gbits) - 1u));
  }
  __device__ bool dropped(uint32_t key, const Morton &mo) const { return (key >> (pbits + gbits)) >> mo.total; }
};

template <typename Real>
__global__ __launch_bounds__(256) void k_comb_sorted(uint32_t n, const uint32_t *__restrict__ keys,
                                                     const uint32_t *__restrict__ vals, CombUnits cu,
                                                     CensusSoA<Real> cen, int32_t eq, Morton mo, MeshDims d,
                                                     double *__restrict__ ws, uint64_t *__restrict__ wq,
                                                     uint32_t *__restrict__ start, uint32_t *__restrict__ end) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t ukey = keys[i] >> cu.pbits;
  const double w = cen.w[vals[i]];
  ws[i] = w;
  wq[i] = fx::value_of(fx::limbs_of(w, eq));
  if (cu.dropped(keys[i], mo)) return;
  uint32_t c;
  const uint32_t u = cu.unit(keys[i], mo, d, c);
  if (i == 0u || (keys[i - 1u] >> cu.pbits) != ukey) start[u] = i;
  if (i + 1u == n || (keys[i + 1u] >> cu.pbits) != ukey) end[u] = i + 1u;
}

struct CombCell {
  double w_out;     // COMB: tooth weight; SPLIT: ideal weight W_c / N_c
  double r;         // comb offset in (0,1)
  double Wd;        // W_c as double
  uint32_t target;  // N_c
  uint32_t mode;    // 0 empty, 1 split/pass-through, 2 comb, 3 dropped (stratified comb)
};

// Per-instance sum of a per-cell value over contiguous cell ranges
// [first[k], first[k + 1]) (one block per instance, fixed tree: the result
// does not depend on scheduling).
__global__ __launch_bounds__(256) void k_inst_sum(const uint32_t *__restrict__ first, const double *__restrict__ v,
                                                  double *__restrict__ out) {
  __shared__ double sh[256];
  const uint32_t k = blockIdx.x;
  double s = 0.0;
  for (uint32_t c = first[k] + threadIdx.x; c < first[k + 1]; c += 256u) s += v[c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t w = 128u; w > 0u; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = sh[0];
}

// Inclusive prefix sum of the sorted census weights in quanta (below 2^62,
// as is every sum of weights on the step's quantum), widened for the
// 128-bit arithmetic of the comb.
__device__ __forceinline__ unsigned __int128 cum_value(const uint64_t *cum, uint32_t i) {
  return static_cast<unsigned __int128>(cum[i]);
}

// Largest integer <= v with at most 53 significant bits (exactly representable
// as a double).
__device__ __forceinline__ unsigned __int128 floor53(unsigned __int128 v) {
  const uint64_t hi = static_cast<uint64_t>(v >> 64), lo = static_cast<uint64_t>(v);
  const int bits = hi ? 128 - __clzll(hi) : (lo ? 64 - __clzll(lo) : 0);
  if (bits <= 53) return v;
  const int drop = bits - 53;
  return (v >> drop) << drop;
}

// Exact double of an integer with at most 53 significant bits.
__device__ __forceinline__ double u128_exact_double(unsigned __int128 v) {
  return static_cast<double>(static_cast<uint64_t>(v >> 64)) * 18446744073709551616.0 +
         static_cast<double>(static_cast<uint64_t>(v));
}

// Remainder limbs of W - q * n (requires q * n <= W).
__device__ __forceinline__ fx::Limbs remainder_limbs(unsigned __int128 W, unsigned __int128 q, uint64_t n) {
  unsigned __int128 rv = W - q * static_cast<unsigned __int128>(n);
  fx::Limbs rl;
  rl.lo = static_cast<uint64_t>(rv) & fx::kLoMask;
  rl.hi = static_cast<uint64_t>(rv >> fx::kLoBits);
  return rl;
}

__device__ __forceinline__ double u128_to_double(unsigned __int128 v) {
  return static_cast<double>(static_cast<uint64_t>(v >> 64)) * 18446744073709551616.0 +
         static_cast<double>(static_cast<uint64_t>(v));
}

// One unit of the comb (a cell, or a bin of the stratified comb) holding
// `count` packets of total weight W (quanta) combed to `target` packets:
// kept and split (count <= target) or combed with teeth of floor(W/target)
// quanta, the remainder into the material of `cell`, the offset drawn from
// the stream of (seed, unit, step).
__device__ void comb_unit_setup(CombCell &cc, unsigned __int128 W, uint32_t count, uint64_t target, int32_t eq,
                                uint64_t seed, uint32_t unit, uint32_t step, unsigned long long *__restrict__ rem,
                                uint32_t cell) {
  cc.Wd = u128_to_double(W) * fpm::pow2d(eq);
  cc.target = static_cast<uint32_t>(target);
  if (count <= target) {
    cc.mode = 1u;
    cc.w_out = cc.Wd / static_cast<double>(target);  // weight_ideal
  } else {
    cc.mode = 2u;
    // tooth weight: floor(W / N) quanta, cut to 53 bits so it is a double;
    // the remainder W - N * tooth goes to the material
    const unsigned __int128 q = floor53(W / static_cast<unsigned __int128>(target));
    cc.w_out = u128_exact_double(q) * fpm::pow2d(eq);
    tally_limbs(rem, cell, remainder_limbs(W, q, target));
    PacketStream rs = PacketStream::make(seed, unit, step, kTagComb);
    Philox4 r = rs.block(0);
    cc.r = u01_f64(r.v[0], r.v[1]);
  }
}

// Census weight of each instance (erg): the sorted census is in cell order
// and an instance's cells are contiguous, so its packets are the run from
// the first packet of its first occupied cell to the last of its last one;
// the weight is the difference of the exact prefix sums at the two ends.
__global__ __launch_bounds__(256) void k_inst_census(const uint32_t *__restrict__ first,
                                                     const uint32_t *__restrict__ start,
                                                     const uint32_t *__restrict__ end,
                                                     const uint64_t *__restrict__ cum, int32_t eq,
                                                     double *__restrict__ wtot) {
  __shared__ uint32_t lo_s[256], hi_s[256];
  const uint32_t k = blockIdx.x;
  uint32_t lo = 0xFFFFFFFFu, hi = 0u;
  for (uint32_t c = first[k] + threadIdx.x; c < first[k + 1]; c += 256u) {
    const uint32_t e = end[c];
    if (e) {
      lo = start[c] < lo ? start[c] : lo;
      hi = e > hi ? e : hi;
    }
  }
  lo_s[threadIdx.x] = lo;
  hi_s[threadIdx.x] = hi;
  __syncthreads();
  for (uint32_t w = 128u; w > 0u; w >>= 1) {
    if (threadIdx.x < w) {
      lo_s[threadIdx.x] = lo_s[threadIdx.x + w] < lo_s[threadIdx.x] ? lo_s[threadIdx.x + w] : lo_s[threadIdx.x];
      hi_s[threadIdx.x] = hi_s[threadIdx.x + w] > hi_s[threadIdx.x] ? hi_s[threadIdx.x + w] : hi_s[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t a = lo_s[0], b = hi_s[0];
    double W = 0.0;
    if (b > 0u) {
      const unsigned __int128 v = cum_value(cum, b - 1u) - (a ? cum_value(cum, a - 1u) : static_cast<unsigned __int128>(0));
      W = u128_to_double(v) * fpm::pow2d(eq);
    }
    wtot[k] = W;
  }
}

// Per-cell comb targets (CombPopulationControl.hpp:157-206).  The cell's
// census weight W_c (CombPopulationControl.hpp:139-141) is the difference of
// the prefix sums over its run of sorted packets; it is also stored, as limbs,
// for the energy ledger.
__global__ __launch_bounds__(256) void k_comb_cells(uint32_t N, const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ end,
                                                    const uint64_t *__restrict__ cum, uint32_t n, uint64_t nmin_all,
                                                    double factor_all, int32_t eq, int32_t comb_on, uint64_t seed,
                                                    uint32_t step, CombCell *__restrict__ cells,
                                                    unsigned long long *__restrict__ W,
                                                    unsigned long long *__restrict__ rem, InstanceView inst) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  uint64_t nmin = nmin_all, Ncells = N;
  double factor = factor_all;
  if (inst.cell) {
    const uint32_t k = inst.cell[c];
    nmin = inst.nmin[k];
    factor = inst.factor[k];
    Ncells = inst.ncells[k];
  }
  CombCell cc;
  cc.w_out = 0.0;
  cc.r = 0.5;
  cc.target = 0;
  cc.mode = 0;
  const uint32_t e = end[c];
  const uint32_t s0 = e ? start[c] : 0u;
  const uint32_t count = e - s0;
  // total census weight (exact integer), as a double in erg
  const double Wtot = inst.cell ? inst.wtot[inst.cell[c]] : u128_to_double(cum_value(cum, n - 1u)) * fpm::pow2d(eq);
  const unsigned __int128 Wc =
      count ? cum_value(cum, e - 1u) - (s0 ? cum_value(cum, s0 - 1u) : static_cast<unsigned __int128>(0))
            : static_cast<unsigned __int128>(0);
  W[2 * c] = static_cast<unsigned long long>(static_cast<uint64_t>(Wc) & fx::kLoMask);
  W[2 * c + 1] = static_cast<unsigned long long>(static_cast<uint64_t>(Wc >> fx::kLoBits));
  const double Wd = u128_to_double(Wc);
  cc.Wd = Wd * fpm::pow2d(eq);
  if (count == 0u || !comb_on) {
    cc.mode = count ? 1u : 0u;
    cc.target = count;
    cc.w_out = 1e308;  // never split
    cells[c] = cc;
    return;
  }
  const uint64_t Ntotal = static_cast<uint64_t>(static_cast<double>(Ncells * nmin) * factor);
  uint64_t share = static_cast<uint64_t>(static_cast<double>(Ntotal) * cc.Wd / Wtot);
  uint64_t NinCell = share > nmin ? share : nmin;
  if (NinCell > 20ull * nmin) NinCell = 20ull * nmin;
  comb_unit_setup(cc, Wc, count, NinCell, eq, seed, c, step, rem, c);
  cells[c] = cc;
}

// StratifiedCombPopulationControl::activate, per cell
// (population/CombPopulationControl.hpp:415-466): the cell's target as in
// k_comb_cells, shared among its occupied (cell, group) bins -- minPerGroup
// each when they all fit, else one each for the `target` heaviest (ties:
// lower group first) and none for the rest -- then the remainder in
// proportion to the bin weights, the largest fractional parts rounded up
// (ties: the first in that order); each bin is then split or combed as a
// cell of the plain comb (:346-404).  A bin given no packets is dropped: its
// weight leaves the problem (tallied in `drop`), as the reference discards
// those packets (:348-349, :409-413).  `bins` has `groups` entries per cell.
constexpr uint32_t kMaxStratGroups = 64;

__global__ __launch_bounds__(256) void k_comb_strat(uint32_t N, uint32_t groups, const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ end,
                                                    const uint64_t *__restrict__ cum, uint32_t n, uint64_t nmin,
                                                    double factor, uint64_t min_per, int32_t eq, uint64_t seed,
                                                    uint32_t step, CombCell *__restrict__ bins,
                                                    unsigned long long *__restrict__ W,
                                                    unsigned long long *__restrict__ rem,
                                                    unsigned long long *__restrict__ drop) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const uint32_t G = groups;
  auto bin_weight = [&](uint32_t g) {
    const uint32_t u = c * G + g, e = end[u];
    if (!e) return static_cast<unsigned __int128>(0);
    const uint32_t s0 = start[u];
    return cum_value(cum, e - 1u) - (s0 ? cum_value(cum, s0 - 1u) : static_cast<unsigned __int128>(0));
  };
  double wb[kMaxStratGroups];
  uint32_t tg[kMaxStratGroups];
  double frac[kMaxStratGroups];
  uint8_t order[kMaxStratGroups];
  unsigned __int128 Wc = 0;
  uint32_t nact = 0;
  for (uint32_t g = 0; g < G; ++g) {
    const unsigned __int128 b = bin_weight(g);
    wb[g] = end[c * G + g] ? u128_to_double(b) : -1.0;  // -1: no packets
    tg[g] = 0u;
    frac[g] = 0.0;
    Wc += b;
    if (end[c * G + g]) ++nact;
  }
  W[2 * c] = static_cast<unsigned long long>(static_cast<uint64_t>(Wc) & fx::kLoMask);
  W[2 * c + 1] = static_cast<unsigned long long>(static_cast<uint64_t>(Wc >> fx::kLoBits));
  uint64_t target = 0;
  if (nact) {
    const double Wcd = u128_to_double(Wc);
    const double Wtot = u128_to_double(cum_value(cum, n - 1u));
    const uint64_t Ntotal = static_cast<uint64_t>(static_cast<double>(static_cast<uint64_t>(N) * nmin) * factor);
    const uint64_t share = static_cast<uint64_t>(static_cast<double>(Ntotal) * Wcd / Wtot);
    target = share > nmin ? share : nmin;
    if (target > 20ull * nmin) target = 20ull * nmin;
    uint64_t given = 0;
    const bool all_fit = static_cast<uint64_t>(nact) * min_per <= target;
    // the order the reference walks its active groups in: by group, or by
    // weight (heaviest first) when they do not all fit
    uint32_t k = 0;
    for (uint32_t g = 0; g < G; ++g) {
      if (wb[g] < 0.0) continue;
      uint32_t pos = k++;
      if (!all_fit) {
        pos = 0;
        for (uint32_t h = 0; h < G; ++h)
          if (wb[h] >= 0.0 && (wb[h] > wb[g] || (wb[h] == wb[g] && h < g))) ++pos;
      }
      order[pos] = static_cast<uint8_t>(g);
    }
    for (uint32_t j = 0; j < nact; ++j) {
      const uint32_t g = order[j];
      const uint64_t t = all_fit ? min_per : (j < target ? 1u : 0u);
      tg[g] = static_cast<uint32_t>(t);
      given += t;
    }
    const uint64_t rest = target > given ? target - given : 0;
    uint64_t extra = 0;
    for (uint32_t j = 0; j < nact; ++j) {
      const uint32_t g = order[j];
      if (!tg[g]) continue;
      const double x = static_cast<double>(rest) * wb[g] / Wcd;
      const double e = floor(x);
      tg[g] += static_cast<uint32_t>(e);
      extra += static_cast<uint64_t>(e);
      frac[g] = x - e;
    }
    while (extra < rest) {
      int best = -1;
      double bf = -1.0;
      for (uint32_t j = 0; j < nact; ++j) {
        const uint32_t g = order[j];
        if (tg[g] && frac[g] > bf) {
          best = static_cast<int>(g);
          bf = frac[g];
        }
      }
      if (best < 0) break;
      ++tg[best];
      frac[best] = 0.0;
      ++extra;
    }
  }
  for (uint32_t g = 0; g < G; ++g) {
    const uint32_t u = c * G + g;
    CombCell cc;
    cc.w_out = 0.0;
    cc.r = 0.5;
    cc.target = 0;
    cc.mode = 0u;
    cc.Wd = 0.0;
    const uint32_t e = end[u];
    if (e) {
      const uint32_t count = e - start[u];
      const unsigned __int128 b = bin_weight(g);
      if (!tg[g]) {
        cc.mode = 3u;
        cc.Wd = u128_to_double(b) * fpm::pow2d(eq);
        tally_limbs(drop, c, fx::Limbs{static_cast<uint64_t>(b) & fx::kLoMask, static_cast<uint64_t>(b >> fx::kLoBits)});
      } else {
        comb_unit_setup(cc, b, count, tg[g], eq, seed, u, step, rem, c);
      }
    }
    bins[u] = cc;
  }
}

__device__ __forceinline__ uint64_t comb_teeth(double C, double Wc, double target, double r) {
  // #{k >= 0 : (k + r) * W/N < C} = ceil(C*N/W - r), clamped to [0, N]
  double x = ceil(C * target / Wc - r);
  if (!(x > 0.0)) return 0;
  if (x > target) return static_cast<uint64_t>(target);
  return static_cast<uint64_t>(x);
}

// Output copies per sorted census packet (0 for the dropped weight-0 packets
// at the end).  `ws` holds the weights in sorted order, `cum` the inclusive
// prefix sums of their quanta; [start, end) is each cell's run.
__global__ __launch_bounds__(256) void k_comb_count(uint32_t n, const uint32_t *__restrict__ keys, CombUnits cu,
                                                    const double *__restrict__ ws,
                                                    const CombCell *__restrict__ cells,
                                                    const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ end,
                                                    const uint64_t *__restrict__ cum, int32_t eq, Morton mo,
                                                    MeshDims d, uint32_t *__restrict__ ncopy) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (cu.dropped(keys[i], mo)) {
    ncopy[i] = 0u;
    return;
  }
  uint32_t c;
  const uint32_t u = cu.unit(keys[i], mo, d, c);
  const CombCell cc = cells[u];
  const double w = ws[i];
  uint32_t k = 1;
  if (cc.mode == 3u) {
    k = 0;  // a bin the stratified comb drops
  } else if (cc.mode == 1u) {
    if (w > 2.0 * cc.w_out) k = static_cast<uint32_t>(ceil(w / cc.w_out));
  } else if (cc.mode == 2u) {
    const uint32_t s = start[u];
    const bool last = (i + 1u == end[u]);
    const unsigned __int128 Cbase = s > 0u ? cum_value(cum, s - 1u) : static_cast<unsigned __int128>(0);
    fx::Limbs wl = fx::limbs_of(w, eq);
    unsigned __int128 wv = (static_cast<unsigned __int128>(wl.hi) << fx::kLoBits) + wl.lo;
    unsigned __int128 Cincl = cum_value(cum, i) - Cbase;
    unsigned __int128 Cexcl = Cincl - wv;
    const double Wc = cc.Wd * fpm::pow2d(-eq);
    const double tgt = static_cast<double>(cc.target);
    uint64_t k1 = last ? cc.target : comb_teeth(u128_to_double(Cincl), Wc, tgt, cc.r);
    uint64_t k0 = comb_teeth(u128_to_double(Cexcl), Wc, tgt, cc.r);
    k = static_cast<uint32_t>(k1 > k0 ? k1 - k0 : 0);
  }
  ncopy[i] = k;
}

// Writes the combed census.  The census weight out per cell is summed over
// each run of equal cells inside the wavefront (the keys are sorted, so a
// cell's lanes are side by side) and added once per run, instead of once per
// packet into the same two words.
template <typename Real>
__global__ __launch_bounds__(256) void k_comb_write(uint32_t n, const uint32_t *__restrict__ keys,
                                                    const uint32_t *__restrict__ vals, CombUnits cu,
                                                    CensusSoA<Real> cen, const double *__restrict__ ws,
                                                    const CombCell *__restrict__ cells,
                                                    const uint32_t *__restrict__ ncopy,
                                                    const uint64_t *__restrict__ offs, CensusSoA<Real> out,
                                                    int32_t eq, Morton mo, MeshDims d, Tallies t) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t c = 0xFFFFFFFFu;  // lanes past the end and dropped packets: no cell
  unsigned long long lo = 0, hi = 0;
  if (i < n) {
    uint32_t u = 0;
    if (!cu.dropped(keys[i], mo)) u = cu.unit(keys[i], mo, d, c);
    const uint32_t k = ncopy[i];
    if (k) {
      const CombCell cc = cells[u];
      const uint32_t src = vals[i];
      const double w = ws[i];
      double wout = w;
      if (cc.mode == 2u) {
        wout = cc.w_out;
      } else if (k > 1u) {
        // split copies: floor(w / k) quanta each, remainder to the material
        fx::Limbs wl = fx::limbs_of(w, eq);
        const unsigned __int128 wv = (static_cast<unsigned __int128>(wl.hi) << fx::kLoBits) + wl.lo;
        const unsigned __int128 q = floor53(wv / static_cast<unsigned __int128>(k));
        wout = u128_exact_double(q) * fpm::pow2d(eq);
        tally_limbs(t.rem, c, remainder_limbs(wv, q, k));
      }
      const auto a = cen.pos[src];
      const auto b = cen.dir[src];
      const uint2 m = cen.meta[src];
      const uint64_t o = offs[i];
      for (uint32_t q = 0; q < k; ++q) {
        out.pos[o + q] = a;
        out.dir[o + q] = b;
        out.w[o + q] = wout;
        out.meta[o + q] = make_uint2(m.x, static_cast<uint32_t>(o + q));
      }
      const fx::Limbs l = fx::limbs_of(wout, eq);
      lo = static_cast<unsigned long long>(l.lo) * k;
      hi = static_cast<unsigned long long>(l.hi) * k;
    }
  }
  wave_run_sum(c, lo);
  if (wave_run_sum(c, hi) && c != 0xFFFFFFFFu) {
    if (lo) atomicAdd(&t.cout[2 * c], lo);
    if (hi) atomicAdd(&t.cout[2 * c + 1], hi);
  }
}

// ------------------------------------------------------------------ postStep
template <typename Real>
__global__ __launch_bounds__(256) void k_post(uint32_t N, Tallies t, int32_t eq, int32_t et, int32_t eq_old,
                                              const unsigned long long *__restrict__ emit_limbs,
                                              const unsigned long long *__restrict__ requant, double dt,
                                              const double *__restrict__ V, MatDev mat, int32_t device_eos,
                                              double *__restrict__ T, double *__restrict__ e_mat,
                                              double *__restrict__ e_sync, double *__restrict__ erad_avg,
                                              double *__restrict__ census_erad, double *__restrict__ clamp,
                                              uint32_t *__restrict__ err, int32_t fp64,
                                              const CellXS<Real> *__restrict__ xs, int32_t gray,
                                              InstanceView inst) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  if (*err & kErrTallyOverflow) return;  // the step's tallies are not trusted: commit nothing
  int32_t sh = 0;  // batches: the cell's instance unit is 2^-sh erg (2^-sh_prev for the census' requant)
  if (inst.cell) {
    const uint32_t k = inst.cell[c];
    sh = inst.shift[k];
    eq_old -= inst.shift_prev[k];
    eq -= sh;
    et -= sh;
  }
  const double emit = fx::to_double(emit_limbs[2 * c], emit_limbs[2 * c + 1], eq);
  double dep, pint;
  if (fp64) {
    dep = ldexp(t.dep_fp[c], -sh);
    pint = ldexp(t.pint_fp[c], -sh);
  } else {
    dep = fx::to_double(t.dep[2 * c], t.dep[2 * c + 1], eq);
    // Each segment absorbs w (1 - e^{-x}) and adds w (1 - e^{-x}) / sig_abs to
    // the path integral, so where sig_abs > 0 the integral is the absorbed
    // energy over sig_abs (the transport skips that tally there).
    const double sabs = static_cast<double>(xs[c].sig_abs);
    pint = (gray && sabs > 0.0) ? dep / sabs : fx::to_double(t.pint[2 * c], t.pint[2 * c + 1], et);
    dep += fx::to_double(t.rdep[2 * c], t.rdep[2 * c + 1], eq);
  }
  dep += fx::to_double(t.rem[2 * c], t.rem[2 * c + 1], eq);
  dep += fx::to_double(requant[2 * c], requant[2 * c + 1], eq_old);
  const double Vc = V[c];
  erad_avg[c] = pint / (units::clight * dt * Vc);  // RadiationIMC.hpp:1895 (time integral = path / c)
  census_erad[c] = fx::to_double(t.cout[2 * c], t.cout[2 * c + 1], eq);
  double E = (e_mat[c] - emit) + dep;
  double cl = 0.0;
  if (E < 0.0) {  // throwIfNegativeInternalEnergy (RadiationIMC.hpp:771-817)
    const double Tc = T[c];
    double thermal = units::arad * (Tc * Tc * Tc * Tc) * Vc;
    if (thermal < 1e-30) thermal = 1e-30;
    if (fabs(E) / thermal < 0.1) {
      double ne = fmax(0.0, e_sync[c]);
      cl = ne - E;
      E = ne;
    } else {
      atomicOr(err, kErrNegativeEnergy);
    }
  }
  clamp[c] = cl;
  e_mat[c] = E;
  e_sync[c] = E;
  if (device_eos) {
    const double rho = mat.density(c);
    T[c] = mat.de2T(rho, E / (rho * Vc));  // synchronizeMaterialCell (:830-866)
  }
}

// out[0] += sum of v[0..N) (integer, order independent)
__global__ __launch_bounds__(256) void k_total_u32(uint32_t N, const uint32_t *__restrict__ v,
                                                   unsigned long long *__restrict__ out) {
  unsigned long long s = 0;
  for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < N; c += gridDim.x * 256u) s += v[c];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ unsigned long long sh[4];
  if ((threadIdx.x & 63u) == 0u) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    s = sh[0] + sh[1] + sh[2] + sh[3];
    if (s) atomicAdd(out, s);
  }
}

// Up to 8 device scalars of 4 or 8 bytes, gathered for one copy to the host.
struct ScalarRefs {
  const void *p[8];
  uint32_t bytes[8];
};
__global__ void k_gather_scalars(ScalarRefs r, uint32_t n, unsigned long long *__restrict__ out) {
  const uint32_t t = threadIdx.x;
  if (t >= n) return;
  out[t] = r.bytes[t] == 8u ? *static_cast<const unsigned long long *>(r.p[t])
                            : static_cast<unsigned long long>(*static_cast<const uint32_t *>(r.p[t]));
}

__global__ void k_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, uint64_t n, double v) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Zeroes up to kZeroRanges ranges of 32-bit words in one launch (blockIdx.y
// picks the range, the x blocks stride over it): the step's tallies and
// counters are cleared by one kernel instead of one per buffer.
constexpr int kZeroRanges = 8;
struct ZeroRanges {
  uint32_t *p[kZeroRanges];
  uint64_t n[kZeroRanges];  // words
};
__global__ __launch_bounds__(256) void k_zero_ranges(ZeroRanges r) {
  uint32_t *__restrict__ p = r.p[blockIdx.y];
  const uint64_t n = r.n[blockIdx.y];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256u;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256u + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// Adds tally replicas 1..R-1 into replica 0, cell by cell in replica order
// (integer sums are exact; the FP64 engine's sums are in a fixed order).
// Also the check of the two-limb tallies' bound: a cell's low limbs take up
// to 2^28 additions of values below 2^36 within 64 bits, so a cell with 2^28
// or more segments in the step sets kErrTallyOverflow here, before the comb
// and postStep, and postStep then commits nothing.  The segment count
// saturates at 2^32 - 1.
__global__ __launch_bounds__(256) void k_fold_replicas(Tallies t, uint32_t reps, uint32_t *__restrict__ err) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= t.ncell) return;
  const size_t n = t.ncell;
  uint64_t sg = t.segs[c];
  for (uint32_t r = 0; r + 1 < reps; ++r) sg += t.segs_rep[r * n + c];
  t.segs[c] = sg > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(sg);
  if (sg >= (1ull << 28)) atomicOr(err, kErrTallyOverflow);
  if (t.dep_rep) {
    unsigned long long d0 = t.dep[2 * c], d1 = t.dep[2 * c + 1];
    unsigned long long p0 = t.pint[2 * c], p1 = t.pint[2 * c + 1];
    for (uint32_t r = 0; r + 1 < reps; ++r) {
      d0 += t.dep_rep[r * 2 * n + 2 * c];
      d1 += t.dep_rep[r * 2 * n + 2 * c + 1];
      p0 += t.pint_rep[r * 2 * n + 2 * c];
      p1 += t.pint_rep[r * 2 * n + 2 * c + 1];
    }
    t.dep[2 * c] = d0;
    t.dep[2 * c + 1] = d1;
    t.pint[2 * c] = p0;
    t.pint[2 * c + 1] = p1;
  }
  if (t.dep_fp_rep) {
    double a = t.dep_fp[c], b = t.pint_fp[c];
    for (uint32_t r = 0; r + 1 < reps; ++r) {
      a += t.dep_fp_rep[r * n + c];
      b += t.pint_fp_rep[r * n + c];
    }
    t.dep_fp[c] = a;
    t.pint_fp[c] = b;
  }
}

// Plain copy of n 16-byte words, one per thread: a kernel whose HBM traffic
// is known exactly (16 n bytes read, 16 n written), to scale byte counters.
__global__ __launch_bounds__(256) void k_copy_calibration(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                          uint64_t n) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace gpu
}  // namespace storm
