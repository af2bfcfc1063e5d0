// CPU replay of the mixed-precision engine's transport and sourcing.
//
// TEST INFRASTRUCTURE ONLY.  It compiles the engine's shared host/device
// headers (storm_amd/csrc/common/*.h) with g++ -ffp-contract=off and replays
// packets from their Philox streams, so the tests can check that
//   (1) the gfx950 kernels reproduce the integer tallies bit-exactly
//       (replay_transport_f32 on the packets the GPU step started from), and
//   (2) the engine's first-step physics agrees statistically with the FP64
//       oracle before any GPU is involved (cpu_first_step).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../include/storm_problem.h"
#include "../storm_amd/csrc/common/fixed.h"
#include "../storm_amd/csrc/common/planck.h"
#include "../storm_amd/csrc/common/random_walk.h"
#include "../storm_amd/csrc/common/sourcing.h"
#include "../storm_amd/csrc/common/transport.h"

using namespace storm;

namespace {
thread_local std::string g_err;

struct HostSink {
  uint64_t *dep, *rdep, *pint, *esc;
  double *dep_fp, *pint_fp, *esc_fp;
  uint32_t *segs;
  int32_t eq;
  uint64_t rw = 0, ddmc = 0;
  // whole values into the low word, as the engine's transport adds them
  void visit(uint32_t c, fx::Limbs d, fx::Limbs e, uint32_t s) {
    dep[2 * c] += fx::value_of(d);
    pint[2 * c] += fx::value_of(e);
    segs[c] += s;
  }
  void roulette(uint32_t c, double w) {
    fx::Limbs l = fx::limbs_of(w, eq);
    rdep[2 * c] += l.lo;
    rdep[2 * c + 1] += l.hi;
  }
  void visit_fp(uint32_t c, double d, double e, uint32_t s) {
    dep_fp[c] += d;
    pint_fp[c] += e;
    segs[c] += s;
  }
  void escape_fx(uint32_t c, double w) {
    fx::Limbs l = fx::limbs_of(w, eq);
    esc[2 * c] += l.lo;
    esc[2 * c + 1] += l.hi;
  }
  void escape_fp(uint32_t c, double w) { esc_fp[c] += w; }
  void rw_step() { ++rw; }
  void ddmc_step() { ++ddmc; }
  // the replay does not check the per-group tally (the problems it replays
  // leave it off)
  void group_path(uint32_t, int32_t, double, const StepConsts &, uint32_t) {}
};

// Random-walk tables in the engine's precision (same host build as the engine).
template <typename Real>
struct HostTables {
  std::vector<Real> all;  // rw::tables_at's layout
  HostTables() {
    std::vector<double> a(rw::kAllSize);
    rw::build_all_tables(a.data());
    all.assign(a.begin(), a.end());
  }
};

// The engine's device layouts from the replay's plain inputs: widths
// (nx + ny + nz) into interleaved (w, 1/w) pairs per axis, and per-cell
// (sig_abs, sig_ev) pairs into CellXS with the derived reciprocal.
struct HostMesh {
  std::vector<float> wp;
  std::vector<CellXS<float>> xs;
  HostMesh(int32_t nx, int32_t ny, int32_t nz, const float *widths, const float *xs2) {
    wp.resize(2 * static_cast<size_t>(nx + ny + nz));
    width_pairs<float>(widths, nx + ny + nz, wp.data());
    const size_t N = static_cast<size_t>(nx) * ny * nz;
    xs.resize(N);
    for (size_t c = 0; c < N; ++c) xs[c] = make_xs<float>(xs2[2 * c], xs2[2 * c + 1]);
  }
  void view(MeshView<float> &mv) const { mv.width = wp.data(); }
};

template <typename Real>
AccelView<Real> make_accel_view(const CellAccel<Real> *cells, const HostTables<Real> *tb, int rw_on, int ddmc_on,
                                double rw_ptau, double ddmc_ptau) {
  AccelView<Real> a;
  a.cell = cells;
  a.rw = tb ? tb->all.data() : nullptr;
  a.rw_particle_tau = rw_on ? static_cast<Real>(rw_ptau) : MathPolicy<Real>::inf();
  a.ddmc_particle_tau = ddmc_on ? static_cast<Real>(ddmc_ptau) : MathPolicy<Real>::inf();
  return a;
}

// Host restatement of the engine's per-cell acceleration data.
// The engine's DDMC cell data (k_accel_cells, k_accel_rates): total opacity
// st (gray: sigma_a + sigma_s; multigroup: the Planck mean sigma_P +
// sigma_s) and continuous absorption fabs per cell.
// With `pg` (ddmc_multigroup_pgrw) the per-cell D / c, eligibility and
// energy cutoff come from the partial-gray sums instead.
struct PartialGrayDDMC {
  std::vector<double> Dt, ecut;
  std::vector<char> elig;
};
template <typename Real>
void add_ddmc(const storm_problem *p, std::vector<CellAccel<Real>> &out, const std::vector<double> &st,
              const std::vector<double> &fabs, const PartialGrayDDMC *pg = nullptr) {
  const int nx = p->mesh.nx, ny = p->mesh.ny, nz = p->mesh.nz;
  const uint32_t N = static_cast<uint32_t>(nx) * ny * nz;
  std::vector<double> Dt(N);
  auto w = [&](int ax, int idx) {
    const double *e = ax == 0 ? p->mesh.x_edges : (ax == 1 ? p->mesh.y_edges : p->mesh.z_edges);
    return static_cast<double>(static_cast<Real>(e[idx + 1] - e[idx]));
  };
  for (uint32_t c = 0; c < N; ++c) {
    int k = static_cast<int>(c % nz), j = static_cast<int>((c / nz) % ny),
        i = static_cast<int>(c / (static_cast<uint32_t>(ny) * nz));
    double dx = w(0, i), dy = w(1, j), dz = w(2, k);
    double V = dx * dy * dz, S = 2.0 * (dy * dz + dx * dz + dx * dy), chord = 4.0 * V / S;
    CellAccel<Real> &a = out[c];
    a.total = Real(0);
    for (int f = 0; f < 6; ++f) a.rate[f] = Real(0);
    a.ddmc_sig_t = static_cast<Real>(st[c]);
    a.ddmc_abs = static_cast<Real>(fabs[c]);
    if (pg) {
      a.ddmc_e_cut = static_cast<Real>(pg->ecut[c]);
      Dt[c] = pg->Dt[c];
      if (p->imc.with_ddmc && pg->elig[c]) a.flags |= 2u;
      continue;
    }
    a.ddmc_e_cut = std::numeric_limits<Real>::infinity();
    Dt[c] = st[c] > 0.0 ? 1.0 / (3.0 * st[c]) : 0.0;
    if (p->imc.with_ddmc && st[c] * chord >= p->imc.ddmc_min_cell_optical_depth && Dt[c] > 0.0) a.flags |= 2u;
  }
  const uint32_t n3[3] = {static_cast<uint32_t>(nx), static_cast<uint32_t>(ny), static_cast<uint32_t>(nz)};
  const uint32_t stride[3] = {static_cast<uint32_t>(ny * nz), static_cast<uint32_t>(nz), 1u};
  for (uint32_t c = 0; c < N; ++c) {
    CellAccel<Real> &a = out[c];
    if (!(a.flags & 2u)) continue;
    uint32_t idx[3] = {c / stride[0], (c / nz) % ny, c % nz};
    double total = 0.0;
    for (int f = 0; f < 6; ++f) {
      int ax = f >> 1;
      bool up = f & 1;
      if (up ? (idx[ax] + 1u >= n3[ax]) : (idx[ax] == 0u)) continue;
      uint32_t nb = up ? c + stride[ax] : c - stride[ax];
      double D1 = Dt[c], D2 = Dt[nb], Df = D1;
      if (D2 > 0.0) Df = (D1 > 0.0 && D2 > 0.0) ? 2.0 * D1 * D2 / (D1 + D2) : std::fmax(D1, D2);
      double ww = w(ax, static_cast<int>(idx[ax]));
      double rate = Df / (ww * (0.5 * ww));
      if (rate > 0.0 && std::isfinite(rate)) {
        a.rate[f] = static_cast<Real>(rate);
        total += rate;
      }
    }
    a.total = static_cast<Real>(total);
    if (!(total > 0.0)) {
      a.flags &= ~2u;
      for (int f = 0; f < 6; ++f) a.rate[f] = Real(0);
    }
  }
}

template <typename Real>
std::vector<CellAccel<Real>> build_accel(const storm_problem *p, const std::vector<double> &sa,
                                         const std::vector<double> &ss, const std::vector<double> &f) {
  const int nx = p->mesh.nx, ny = p->mesh.ny, nz = p->mesh.nz;
  const uint32_t N = static_cast<uint32_t>(nx) * ny * nz;
  std::vector<CellAccel<Real>> out(N);
  std::vector<double> st(N), fabs(N);
  auto w = [&](int ax, int idx) {
    const double *e = ax == 0 ? p->mesh.x_edges : (ax == 1 ? p->mesh.y_edges : p->mesh.z_edges);
    return static_cast<double>(static_cast<Real>(e[idx + 1] - e[idx]));
  };
  for (uint32_t c = 0; c < N; ++c) {
    int k = static_cast<int>(c % nz), j = static_cast<int>((c / nz) % ny),
        i = static_cast<int>(c / (static_cast<uint32_t>(ny) * nz));
    double dx = w(0, i), dy = w(1, j), dz = w(2, k);
    double V = dx * dy * dz, S = 2.0 * (dy * dz + dx * dz + dx * dy), chord = 4.0 * V / S;
    st[c] = sa[c] + ss[c];
    fabs[c] = f[c] * sa[c];
    CellAccel<Real> &a = out[c];
    a.sig_t = static_cast<Real>(st[c]);
    a.flags = 0u;
    a.gcut = 0u;
    a.sig_d = a.sig_t;
    a.sig_abs = Real(0);
    a.up_rate = Real(0);
    a.e_cut = std::numeric_limits<Real>::infinity();
    if (p->imc.with_random_walk && st[c] * chord >= p->imc.rw_min_cell_optical_depth) a.flags |= 1u;
  }
  add_ddmc<Real>(p, out, st, fabs);
  return out;
}

// The engine's partial-gray random-walk cell data (k_prestep_mg): T, gray
// sigma_a and sigma_s, Fleck factors per cell.
template <typename Real>
std::vector<CellAccel<Real>> build_accel_mg(const storm_problem *p, const std::vector<double> &sa,
                                            const std::vector<double> &ss, const std::vector<double> &f,
                                            const std::vector<double> &sP) {
  const int nx = p->mesh.nx, ny = p->mesh.ny, nz = p->mesh.nz;
  const uint32_t N = static_cast<uint32_t>(nx) * ny * nz;
  const int32_t G = p->groups.num_groups;
  const double *gb = p->groups.bounds;
  std::vector<CellAccel<Real>> out(N);
  const bool pgrw = p->imc.with_ddmc && p->imc.ddmc_multigroup_pgrw;
  PartialGrayDDMC pg;
  std::vector<double> qst(N, 0.0), qabs(N, 0.0);
  if (pgrw) {
    pg.Dt.assign(N, 0.0);
    pg.ecut.assign(N, 0.0);
    pg.elig.assign(N, 0);
  }
  auto w = [&](int ax, int idx) {
    const double *e = ax == 0 ? p->mesh.x_edges : (ax == 1 ? p->mesh.y_edges : p->mesh.z_edges);
    return static_cast<double>(static_cast<Real>(e[idx + 1] - e[idx]));
  };
  for (uint32_t c = 0; c < N; ++c) {
    int k = static_cast<int>(c % nz), j = static_cast<int>((c / nz) % ny),
        i = static_cast<int>(c / (static_cast<uint32_t>(ny) * nz));
    const double dx = w(0, i), dy = w(1, j), dz = w(2, k);
    const double chord = 4.0 * dx * dy * dz / (2.0 * (dy * dz + dx * dz + dx * dy));
    const double kTrw = units::k_boltz * p->T0[c];
    double allAB = 0.0, dB = 0.0, dBA = 0.0, dBT = 0.0, dBoT = 0.0;
    uint32_t cut = 0;
    bool thin = false;
    double qB = 0.0, qBA = 0.0, qBT = 0.0, qBoT = 0.0;
    uint32_t qcut = 0;
    bool qthin = !pgrw || !(p->T0[c] > 0.0);
    for (int32_t g = 0; g < G; ++g) {
      const double Ec = 0.5 * (gb[g] + gb[g + 1]);
      const double sg = sa[c] * std::pow(std::fmax(Ec, gb[0]) / p->groups.nu_ref, -p->groups.nu_pow);
      const double ar = gb[g] / kTrw, br = gb[g + 1] / kTrw;
      const double Br = (ar > 0.0 && br > ar) ? planck::integral(ar, br) : 1.0 / static_cast<double>(G);
      const double st = sg + ss[c];
      allAB += sg * Br;
      if (!thin && p->imc.with_random_walk && st * chord >= p->imc.rw_min_cell_optical_depth) {
        cut = static_cast<uint32_t>(g) + 1u;
        dB += Br;
        dBA += Br * sg;
        dBT += Br * st;
        if (st > 0.0) dBoT += Br / st;
      } else {
        thin = true;
      }
      if (!qthin && st * chord >= p->imc.ddmc_min_cell_optical_depth) {
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
    if (pgrw) {
      pg.ecut[c] = std::numeric_limits<double>::infinity();
      if (qcut > 0u && qB > 0.0) {
        qst[c] = qBT / qB;
        qabs[c] = f[c] * (qBA / qB);
        pg.Dt[c] = qBoT / (3.0 * qB);
        pg.ecut[c] = static_cast<double>(static_cast<Real>(gb[qcut] / units::kev));
        pg.elig[c] = (qst[c] > 0.0 && pg.Dt[c] > 0.0) ? 1 : 0;
      }
    }
    CellAccel<Real> &r = out[c];
    r.total = Real(0);
    for (int q = 0; q < 6; ++q) r.rate[q] = Real(0);
    r.flags = 0u;
    r.gcut = 0u;
    r.sig_t = r.sig_d = r.sig_abs = r.up_rate = Real(0);
    r.e_cut = std::numeric_limits<Real>::infinity();
    r.ddmc_sig_t = r.ddmc_abs = Real(0);
    if (cut > 0u && dB > 0.0) {
      const double sA = dBA / dB, sT = dBT / dB;
      const double gamma = allAB > 0.0 ? dBA / allAB : 1.0;
      r.flags = 1u;
      r.gcut = cut;
      r.sig_t = static_cast<Real>(sT);
      r.sig_d = static_cast<Real>(dBoT > 0.0 ? dB / dBoT : 0.0);
      r.sig_abs = static_cast<Real>(f[c] * sA);
      r.up_rate = static_cast<Real>((gamma < 1.0 && sA > 0.0 && f[c] > 0.0) ? (1.0 - f[c]) * sA * (1.0 - gamma) : 0.0);
      r.e_cut = static_cast<Real>(gb[cut] / units::kev);
    }
  }
  if (p->imc.with_ddmc && pgrw) {
    add_ddmc<Real>(p, out, qst, qabs, &pg);  // partial-gray DDMC (k_prestep_mg)
  } else if (p->imc.with_ddmc) {
    // DDMC on the Planck means (k_accel_cells after k_prestep_mg)
    std::vector<double> st(N), fabs(N);
    for (uint32_t c = 0; c < N; ++c) {
      st[c] = sP[c] + ss[c];
      fabs[c] = f[c] * sP[c];
    }
    add_ddmc<Real>(p, out, st, fabs);
  }
  return out;
}
}  // namespace

extern "C" {

const char *replay_last_error(void) { return g_err.c_str(); }

// Bytes of one storm::CellAccel<float> record (the layout the replay entry
// points and storm_gpu_debug_accel use).
uint64_t replay_cell_accel_bytes(void) { return sizeof(CellAccel<float>); }

// The engine's unbiased rounding of path integrals (fx::quantize_unbiased).
double replay_quantize_unbiased(double v, int32_t e, uint32_t u24) { return fx::quantize_unbiased(v, e, u24); }

// Philox4x32-10 block (for known-answer tests).
void replay_philox(const uint32_t *ctr4, const uint32_t *key2, uint32_t *out4) {
  Philox4 r = philox4x32_10(ctr4[0], ctr4[1], ctr4[2], ctr4[3], key2[0], key2[1]);
  for (int i = 0; i < 4; ++i) out4[i] = r.v[i];
}

// The engine's FP32 elementary functions on a batch (for accuracy tests):
// which 0 log, 1 exp, 2 expm1, 3 sin(2 pi x), 4 cos(2 pi x).
void replay_fpmath(int32_t which, int64_t n, const float *x, float *y) {
  for (int64_t i = 0; i < n; ++i) {
    float s, c;
    switch (which) {
      case 0: y[i] = fpm::logf_d(x[i]); break;
      case 1: y[i] = fpm::expf_d(x[i]); break;
      case 2: y[i] = fpm::expm1f_d(x[i]); break;
      case 3: fpm::sincos2pi(x[i], s, c); y[i] = s; break;
      default: fpm::sincos2pi(x[i], s, c); y[i] = c; break;
    }
  }
}

// The engine's random-walk tables (double) and its samplers in `precision`
// (0 float, 1 double) on a batch.
void replay_rw_tables(double *tau, double *survival, double *radius) { rw::build_tables(tau, survival, radius); }

// The engine's inverse leak-time table T(y), rw::kLeakSize doubles.
void replay_rw_leak_table(double *leak) { rw::build_leak_table(leak); }

// The engine's normalised Planck integral over [a_i, b_i] (in units of kT).
void replay_planck(int64_t n, const double *a, const double *b, double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = planck::integral(a[i], b[i]);
}

void replay_rw_sample(int32_t precision, int64_t n, const double *xi, const double *tau, double *leak,
                      double *radius) {
  auto go = [&](auto tag) {
    using Real = decltype(tag);
    HostTables<Real> tb;
    const rw::Tables<Real> t = rw::tables_at(tb.all.data());
    for (int64_t i = 0; i < n; ++i) {
      leak[i] = static_cast<double>(rw::sample_leak_time<Real>(t, static_cast<Real>(xi[i])));
      radius[i] = static_cast<double>(rw::sample_radius<Real>(t, static_cast<Real>(tau[i]), static_cast<Real>(xi[i])));
    }
  };
  if (precision == 0) go(float{});
  else go(double{});
}

// Exact replay of `n` packets (storm::Packet<float>) through the transport.
// widths: nx + ny + nz floats; xs: 2 floats per cell; consts: seed, step, eq, et;
// accel: N CellAccel<float> records or NULL; acc_params: rw_on, ddmc_on,
// rw particle tau, ddmc particle tau.  Multigroup (groups >= 2): nu_pow,
// group bounds in keV (groups + 1), the per-cell emission CDFs
// (N * (groups + 1)) and scattering opacities (N).  Tallies are accumulated
// into the given arrays (dep/rdep/pint/esc: 2 per cell, segs: 1 per cell);
// counts[0..3] outcomes, counts[4] random-walk steps, counts[5] DDMC steps.
int replay_transport_f32(int64_t n, const void *packets, int32_t nx, int32_t ny, int32_t nz, const float *widths,
                         const int32_t *face_kind, const float *xs, const void *accel, const double *acc_params,
                         const int64_t *consts, int32_t K, int32_t groups, double nu_pow, const float *bounds_kev,
                         const float *cdf, const float *sig_s, uint64_t *dep, uint64_t *rdep, uint64_t *pint,
                         uint64_t *esc, uint32_t *segs, int64_t *counts) {
  try {
    MeshView<float> mv;
    mv.nx = nx;
    mv.ny = ny;
    mv.nz = nz;
    mv.escape_mask = 0u;
    for (int f = 0; f < 6; ++f)
      if (face_kind[f] == STORM_FACE_ESCAPE) mv.escape_mask |= 1u << f;
    const HostMesh hm(nx, ny, nz, widths, xs);
    hm.view(mv);
    mv.spec = SpectrumView<float>{};
    mv.spec.groups = groups;
    if (groups >= 2) {
      mv.spec.nu_pow = static_cast<float>(nu_pow);
      mv.spec.e_lo = bounds_kev[0];
      mv.spec.e_hi = bounds_kev[groups];
      mv.spec.bounds = bounds_kev;
      mv.spec.cdf = cdf;
      mv.spec.sig_s = sig_s;
    }
    StepConsts kc;
    kc.seed = static_cast<uint64_t>(consts[0]);
    kc.step = static_cast<uint32_t>(consts[1]);
    kc.eq = static_cast<int32_t>(consts[2]);
    kc.et = static_cast<int32_t>(consts[3]);
    const int rw_on = acc_params ? static_cast<int>(acc_params[0]) : 0;
    const int ddmc_on = acc_params ? static_cast<int>(acc_params[1]) : 0;
    std::unique_ptr<HostTables<float>> tb(rw_on ? new HostTables<float>() : nullptr);
    AccelView<float> av = make_accel_view<float>(static_cast<const CellAccel<float> *>(accel), tb.get(), rw_on,
                                                 ddmc_on, acc_params ? acc_params[2] : 0.0,
                                                 acc_params ? acc_params[3] : 0.0);
    std::vector<Packet<float>> pk(static_cast<size_t>(n));
    std::memcpy(pk.data(), packets, static_cast<size_t>(n) * sizeof(Packet<float>));
    HostSink sink{dep, rdep, pint, esc, nullptr, nullptr, nullptr, segs, kc.eq};
    for (auto &p : pk) {
      int o = kAlive;
      while (o == kAlive)
        o = transport_packet<float>(p, mv, hm.xs.data(), av, kc, K, sink);
      counts[o] += 1;
    }
    counts[4] += static_cast<int64_t>(sink.rw);
    counts[5] += static_cast<int64_t>(sink.ddmc);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// Segment-by-segment trace of one packet through the engine's transport
// (mixed engine): after each segment, 12 doubles: u0 u1 u2 o0 o1 o2 s w cell
// ijk ctr outcome.  At most max_seg segments; returns the number written, or
// -1 on error.
int64_t replay_trace_f32(const void *packet, int32_t nx, int32_t ny, int32_t nz, const float *widths,
                         const int32_t *face_kind, const float *xs, const void *accel, const double *acc_params,
                         const int64_t *consts, int64_t max_seg, double *trace) {
  try {
    MeshView<float> mv;
    mv.nx = nx;
    mv.ny = ny;
    mv.nz = nz;
    mv.escape_mask = 0u;
    for (int f = 0; f < 6; ++f)
      if (face_kind[f] == STORM_FACE_ESCAPE) mv.escape_mask |= 1u << f;
    const HostMesh hm(nx, ny, nz, widths, xs);
    hm.view(mv);
    mv.spec = SpectrumView<float>{};
    StepConsts kc;
    kc.seed = static_cast<uint64_t>(consts[0]);
    kc.step = static_cast<uint32_t>(consts[1]);
    kc.eq = static_cast<int32_t>(consts[2]);
    kc.et = static_cast<int32_t>(consts[3]);
    const int rw_on = acc_params ? static_cast<int>(acc_params[0]) : 0;
    const int ddmc_on = acc_params ? static_cast<int>(acc_params[1]) : 0;
    std::unique_ptr<HostTables<float>> tb(rw_on ? new HostTables<float>() : nullptr);
    AccelView<float> av = make_accel_view<float>(static_cast<const CellAccel<float> *>(accel), tb.get(), rw_on,
                                                 ddmc_on, acc_params ? acc_params[2] : 0.0,
                                                 acc_params ? acc_params[3] : 0.0);
    Packet<float> p;
    std::memcpy(&p, packet, sizeof(p));
    const size_t N = static_cast<size_t>(nx) * ny * nz;
    std::vector<uint64_t> d(2 * N), r(2 * N), q(2 * N), e(2 * N);
    std::vector<uint32_t> sg(N);
    HostSink sink{d.data(), r.data(), q.data(), e.data(), nullptr, nullptr, nullptr, sg.data(), kc.eq};
    Tracker<float> t;
    t.begin(p, mv, hm.xs.data(), kc);
    int64_t k = 0;
    while (k < max_seg && t.outcome == kAlive) {
      t.segment(mv, hm.xs.data(), av, kc, sink);
      double *o = trace + 12 * k;
      o[0] = t.u0; o[1] = t.u1; o[2] = t.u2;
      o[3] = t.o0; o[4] = t.o1; o[5] = t.o2;
      o[6] = t.s; o[7] = t.w; o[8] = t.cell; o[9] = t.ijk; o[10] = t.ctr; o[11] = t.outcome;
      ++k;
    }
    return k;
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return -1;
  }
}

// First IMC step of the engine on the CPU (no census): preStep, photon
// allocation, emission + boundary births and transport, in the engine's
// arithmetic (precision 0: float + fixed point, 1: double).  Outputs per cell:
// deposited energy (erg), Erad time average (erg/cm^3), segments; totals:
// escaped, census energy, packets, emitted, source, census packets,
// random-walk steps, DDMC steps.
int cpu_first_step(const storm_problem *p, double dt, int32_t precision, uint32_t step, double *dep_out,
                   double *erad_out, double *segs_out, double *totals_out) {
  try {
    const int nx = p->mesh.nx, ny = p->mesh.ny, nz = p->mesh.nz;
    const uint32_t N = static_cast<uint32_t>(nx) * ny * nz;
    std::vector<double> V(N), E(N), f(N), sa(N), ss(N);
    auto widthd = [&](const double *e, int i) { return e[i + 1] - e[i]; };
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k)
          V[(static_cast<size_t>(i) * ny + j) * nz + k] =
              widthd(p->mesh.x_edges, i) * widthd(p->mesh.y_edges, j) * widthd(p->mesh.z_edges, k);
    const storm_material_desc &m = p->material;
    const int32_t G = p->groups.num_groups >= 2 ? p->groups.num_groups : 0;
    const double *gb = p->groups.bounds;
    std::vector<double> cdfd(G ? static_cast<size_t>(N) * (G + 1) : 0), s1(N, 0.0), sPc(N, 0.0);
    double Etot = 0.0;
    for (uint32_t c = 0; c < N; ++c) {
      double T = p->T0[c];
      double rho = m.density ? m.density[c] : 1.0;
      auto tfac = [&](double pw) {
        return pw == 0.0 ? 1.0 : std::pow(std::fmax(T, m.opacity_T_floor) / m.opacity_T_ref, -pw);
      };
      auto rfac = [&](double pw) { return (pw == 0.0 || !m.density) ? 1.0 : std::pow(rho, pw); };
      sa[c] = m.sigma_a0[c] * tfac(m.a_temp_pow) * rfac(m.a_rho_pow);
      ss[c] = m.sigma_s0 ? m.sigma_s0[c] * tfac(m.s_temp_pow) * rfac(m.s_rho_pow) : 0.0;
      double cv = m.eos_f * m.eos_beta * std::pow(std::fmax(T, m.eos_cv_T_floor), m.eos_beta - 1.0) *
                  std::pow(rho, 1.0 - m.eos_mu) / rho;
      double T3 = T * T * T;
      double sP = sa[c];
      if (G) {
        // the engine's multigroup preStep (k_prestep_mg)
        const double kT = units::k_boltz * std::fmax(T, m.opacity_T_floor);
        double acc = 0.0, sumB = 0.0;
        double *my = &cdfd[static_cast<size_t>(c) * (G + 1)];
        my[0] = 0.0;
        for (int32_t g = 0; g < G; ++g) {
          const double Bg = planck::integral(gb[g] / kT, gb[g + 1] / kT);
          const double Ec = 0.5 * (gb[g] + gb[g + 1]);
          const double sg = sa[c] * std::pow(std::fmax(Ec, gb[0]) / p->groups.nu_ref, -p->groups.nu_pow);
          const double wg = (sg > 0.0 && std::isfinite(sg) && std::isfinite(Bg)) ? sg * Bg : 0.0;
          sumB += Bg;
          acc += wg;
          my[g + 1] = acc;
        }
        const bool valid = acc > 0.0 && std::isfinite(acc);
        for (int32_t g = 1; g <= G; ++g) my[g] = valid ? my[g] / acc : (gb[g] - gb[0]) / (gb[G] - gb[0]);
        my[G] = 1.0;
        sP = sumB > 0.0 ? acc / sumB : 1e-20;
        sPc[c] = sP;
        s1[c] = sa[c] * std::pow(units::kev / p->groups.nu_ref, -p->groups.nu_pow);
      }
      f[c] = 1.0 / (1.0 + (4.0 * units::arad * T3 * sP * units::clight * dt) / cv);
      E[c] = f[c] * V[c] * units::arad * (T3 * T) * sP * dt * units::clight;
      if (!G) s1[c] = sa[c];
      Etot += E[c];
    }
    // source sides in the engine's order: x- then x+
    std::vector<double> sideT;
    std::vector<int> sideHi;
    if (p->boundary.source_face == STORM_XLO) {
      sideT.push_back(p->boundary.source_temperature);
      sideHi.push_back(0);
    }
    if (p->boundary.source_xhi) {
      sideT.push_back(p->boundary.source_temperature_xhi);
      sideHi.push_back(1);
    }
    const uint32_t npc = sideT.empty() ? 0u : static_cast<uint32_t>(p->boundary.source_packets_per_face);
    const size_t per_side = static_cast<size_t>(ny) * nz;
    double src = 0.0;
    std::vector<double> wface(npc ? per_side * sideT.size() : 0);
    for (size_t fi = 0; fi < wface.size(); ++fi) {
      const size_t s = fi / per_side, jk = fi % per_side;
      const double T4 = sideT[s] * sideT[s] * sideT[s] * sideT[s];
      int j = static_cast<int>(jk / nz), k = static_cast<int>(jk % nz);
      wface[fi] = units::sigma_sb * T4 * widthd(p->mesh.y_edges, j) * widthd(p->mesh.z_edges, k) * dt / npc;
      src += wface[fi] * npc;
    }
    const double Ebound = Etot + src;
    const int32_t eq = fx::ceil_log2(Ebound) - fx::kHeadroomBits;
    const int32_t et = fx::ceil_log2(Ebound * units::clight * dt) - fx::kHeadroomBits;
    for (auto &w : wface) w = fx::quantize_down(w, eq);
    const uint64_t nmin = p->imc.new_photons_per_cell;
    const uint64_t total = static_cast<uint64_t>(N) * nmin * 10ull;
    std::vector<uint32_t> nph(N);
    std::vector<double> wem(N);
    for (uint32_t c = 0; c < N; ++c) {
      uint64_t share = Etot > 0.0 ? static_cast<uint64_t>(E[c] / Etot * static_cast<double>(total)) : nmin;
      uint64_t n = 0;
      if (E[c] > 0.0) {
        n = share < 20ull * nmin ? share : 20ull * nmin;
        if (n < nmin) n = nmin;
      }
      double w = n ? fx::emission_round(E[c] / static_cast<double>(n), eq, c, step) : 0.0;
      if (w == 0.0) n = 0;
      nph[c] = static_cast<uint32_t>(n);
      wem[c] = w;
    }
    std::vector<uint64_t> dep(2 * N, 0), rdep(2 * N, 0), pint(2 * N, 0), esc(2 * N, 0);
    std::vector<float> sabs32(N);
    for (uint32_t c = 0; c < N; ++c) sabs32[c] = G ? 0.0f : static_cast<float>(f[c] * sa[c]);
    std::vector<double> depf(N, 0), pintf(N, 0), escf(N, 0);
    std::vector<uint32_t> segs(N, 0);
    int64_t outc[4] = {0, 0, 0, 0};
    uint64_t nrw = 0, nddmc = 0;
    double census = 0.0;
    StepConsts kc{p->seed, step, eq, et};
    const bool accel_on = p->imc.with_random_walk || p->imc.with_ddmc;
    auto go = [&](auto tag) {
      using Real = decltype(tag);
      std::vector<Real> w0(nx), w1(ny), w2(nz);
      for (int i = 0; i < nx; ++i) w0[i] = static_cast<Real>(widthd(p->mesh.x_edges, i));
      for (int i = 0; i < ny; ++i) w1[i] = static_cast<Real>(widthd(p->mesh.y_edges, i));
      for (int i = 0; i < nz; ++i) w2[i] = static_cast<Real>(widthd(p->mesh.z_edges, i));
      MeshView<Real> mv;
      mv.nx = nx;
      mv.ny = ny;
      mv.nz = nz;
      mv.escape_mask = 0u;
      for (int q = 0; q < 6; ++q)
        if (p->boundary.face_kind[q] == STORM_FACE_ESCAPE) mv.escape_mask |= 1u << q;
      std::vector<Real> wp(2 * static_cast<size_t>(nx + ny + nz));
      width_pairs<Real>(w0.data(), nx, wp.data());
      width_pairs<Real>(w1.data(), ny, wp.data() + 2 * nx);
      width_pairs<Real>(w2.data(), nz, wp.data() + 2 * (nx + ny));
      mv.width = wp.data();
      std::vector<CellXS<Real>> xs(N);
      for (uint32_t c = 0; c < N; ++c)
        xs[c] = make_xs<Real>(static_cast<Real>(f[c] * s1[c]),
                              G ? static_cast<Real>((1.0 - f[c]) * s1[c]) : static_cast<Real>(ss[c] + (1.0 - f[c]) * sa[c]));
      std::vector<Real> cdfr(cdfd.begin(), cdfd.end()), ssr(ss.begin(), ss.end()), bk(G ? G + 1 : 0), bcdf;
      for (int32_t g = 0; G && g <= G; ++g) bk[g] = static_cast<Real>(gb[g] / units::kev);
      mv.spec = SpectrumView<Real>{};
      mv.spec.groups = G;
      if (G) {
        mv.spec.nu_pow = static_cast<Real>(p->groups.nu_pow);
        mv.spec.e_lo = bk[0];
        mv.spec.e_hi = bk[G];
        mv.spec.bounds = bk.data();
        mv.spec.cdf = cdfr.data();
        mv.spec.sig_s = ssr.data();
        for (double Ts : sideT) {
          const double kT = units::k_boltz * Ts;
          std::vector<double> acc(G + 1, 0.0);
          for (int32_t g = 0; g < G; ++g) acc[g + 1] = acc[g] + planck::integral(gb[g] / kT, gb[g + 1] / kT);
          for (int32_t g = 0; g <= G; ++g) bcdf.push_back(acc[G] > 0.0 ? static_cast<Real>(acc[g] / acc[G]) : Real(0));
        }
      }
      std::vector<CellAccel<Real>> acc_cells;
      std::unique_ptr<HostTables<Real>> tb;
      if (accel_on) {
        acc_cells = G ? build_accel_mg<Real>(p, sa, ss, f, sPc) : build_accel<Real>(p, sa, ss, f);
        if (p->imc.with_random_walk) tb.reset(new HostTables<Real>());
      }
      AccelView<Real> av = make_accel_view<Real>(accel_on ? acc_cells.data() : nullptr, tb.get(),
                                                 p->imc.with_random_walk, p->imc.with_ddmc,
                                                 p->imc.rw_min_particle_optical_depth,
                                                 p->imc.ddmc_min_particle_optical_depth);
      const Real s0 = static_cast<Real>(units::clight * dt);
      HostSink sink{dep.data(), rdep.data(), pint.data(), esc.data(), nullptr, nullptr, nullptr, segs.data(), eq};
      if (sizeof(Real) == 8) {
        sink.dep_fp = depf.data();
        sink.pint_fp = pintf.data();
        sink.esc_fp = escf.data();
      }
      uint32_t id = 0;
      std::vector<Packet<Real>> pk;
      for (uint32_t c = 0; c < N; ++c) {
        uint32_t k = c % nz, j = (c / nz) % ny, i = c / (static_cast<uint32_t>(ny) * nz);
        for (uint32_t q = 0; q < nph[c]; ++q) {
          Packet<Real> pp;
          birth_emission<Real>(pp, id++, pack_ijk(i, j, k), wem[c], s0, p->seed, step, mv.spec, c);
          pk.push_back(pp);
        }
      }
      for (size_t fi = 0; fi < wface.size(); ++fi) {
        for (uint32_t q = 0; q < npc; ++q) {
          Packet<Real> pp;
          const size_t s = fi / per_side, jk = fi % per_side;
          const bool hi = sideHi[s] != 0;
          birth_boundary<Real>(pp, id++, hi, hi ? static_cast<uint32_t>(nx - 1) : 0u, static_cast<uint32_t>(jk / nz),
                               static_cast<uint32_t>(jk % nz), wface[fi], s0, p->seed, step, mv.spec,
                               G ? bcdf.data() + s * static_cast<size_t>(G + 1) : nullptr);
          pk.push_back(pp);
        }
      }
      for (auto &pp : pk) {
        int o = kAlive;
        while (o == kAlive) o = transport_packet<Real>(pp, mv, xs.data(), av, kc, 1 << 20, sink);
        outc[o] += 1;
        if (o == kCensus) census += pp.w;
      }
      nrw = sink.rw;
      nddmc = sink.ddmc;
    };
    if (precision == 0)
      go(float{});
    else
      go(double{});
    double escaped = 0.0;
    for (uint32_t c = 0; c < N; ++c) {
      double d = depf[c], e = pintf[c];
      if (precision == 0) {
        // path integral = absorbed / sig_abs where the cell absorbs (the engine's postStep)
        d = fx::to_double(dep[2 * c], dep[2 * c + 1], eq);
        e = sabs32[c] > 0.0f ? d / static_cast<double>(sabs32[c]) : fx::to_double(pint[2 * c], pint[2 * c + 1], et);
        d += fx::to_double(rdep[2 * c], rdep[2 * c + 1], eq);
      }
      escaped += precision == 0 ? fx::to_double(esc[2 * c], esc[2 * c + 1], eq) : escf[c];
      dep_out[c] = d;
      erad_out[c] = e / (units::clight * dt * V[c]);
      segs_out[c] = segs[c];
    }
    double emitted = 0.0;
    for (uint32_t c = 0; c < N; ++c) emitted += wem[c] * nph[c];
    totals_out[0] = escaped;
    totals_out[1] = census;
    totals_out[2] = static_cast<double>(outc[0] + outc[1] + outc[2] + outc[3]);
    totals_out[3] = emitted;
    totals_out[4] = src;
    totals_out[5] = static_cast<double>(outc[kCensus]);
    totals_out[6] = static_cast<double>(nrw);
    totals_out[7] = static_cast<double>(nddmc);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

}  // extern "C"
