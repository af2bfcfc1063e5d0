// C entry points of the FP64 CPU oracle, for ctypes (tests, smoke, bench's
// cpu_baseline).  TEST INFRASTRUCTURE ONLY -- never linked into the product.
#include <chrono>
#include <cstring>
#include <memory>
#include <string>

#include "oracle.hpp"

using namespace storm_oracle;

namespace {
thread_local std::string g_err;

CartesianGrid makeGrid(const storm_mesh_desc &m) {
  std::vector<double> xe(m.x_edges, m.x_edges + m.nx + 1);
  std::vector<double> ye(m.y_edges, m.y_edges + m.ny + 1);
  std::vector<double> ze(m.z_edges, m.z_edges + m.nz + 1);
  return CartesianGrid(std::move(xe), std::move(ye), std::move(ze));
}

Material makeMaterial(const storm_material_desc &d, size_t N, const storm_group_desc *g = nullptr) {
  Material m;
  if (g && g->num_groups >= 2) {
    m.bounds.assign(g->bounds, g->bounds + g->num_groups + 1);
    m.nu_pow = g->nu_pow;
    m.nu_ref = g->nu_ref;
  }
  m.sa0.assign(d.sigma_a0, d.sigma_a0 + N);
  if (d.sigma_s0) m.ss0.assign(d.sigma_s0, d.sigma_s0 + N);
  if (d.density) m.rho.assign(d.density, d.density + N);
  m.Tref = d.opacity_T_ref;
  m.Tfloor = d.opacity_T_floor;
  m.apow = d.a_temp_pow;
  m.arho = d.a_rho_pow;
  m.spow = d.s_temp_pow;
  m.srho = d.s_rho_pow;
  m.eos_f = d.eos_f;
  m.eos_beta = d.eos_beta;
  m.eos_mu = d.eos_mu;
  m.cvTfloor = d.eos_cv_T_floor;
  return m;
}
}  // namespace

namespace {
// Per-step ledger layout (doubles): emitted, source, escaped, clamp,
// censusIn, censusOut, censusComb, materialBefore, materialAfter, segments,
// startParticles, endParticles, seconds, combDropped
constexpr size_t kLedgerStride = 14;

template <typename Grid>
int run_schedule_on(const Grid &grid, const storm_problem *prob, int nsteps, const double *dts, const double *source_T,
                    double *T_out, double *e_out, double *erad_out, double *census_out, double *segments_out,
                    double *ledger_out, double *eg_out) {
  try {
    size_t N = grid.GetPointNo();
    Material mat = makeMaterial(prob->material, N, &prob->groups);
    std::vector<Cell> cells(N);
    std::vector<Extensives> ext(N);
    for (size_t i = 0; i < N; ++i) {
      double rho = mat.density(i), V = grid.GetVolume(i);
      cells[i].temperature = prob->T0[i];
      cells[i].internalEnergy = mat.eFromT(prob->T0[i], rho) * V;
      ext[i].mass = rho * V;
      ext[i].internal_energy = cells[i].internalEnergy;
    }
    storm_boundary_desc bd = prob->boundary;
    FaceTableBoundary<Grid> bc(grid, bd, prob->seed);
    if (mat.multigroup()) bc.setGroupBounds(mat.bounds);
    IMCParams ip;
    ip.newPhotonsPerCell = prob->imc.new_photons_per_cell;
    ip.withRandomWalk = prob->imc.with_random_walk != 0;
    ip.withDDMC = prob->imc.with_ddmc != 0;
    ip.rwMinCellOpticalDepth = prob->imc.rw_min_cell_optical_depth;
    ip.rwMinParticleOpticalDepth = prob->imc.rw_min_particle_optical_depth;
    ip.ddmcMinCellOpticalDepth = prob->imc.ddmc_min_cell_optical_depth;
    ip.ddmcMinParticleOpticalDepth = prob->imc.ddmc_min_particle_optical_depth;
    ip.ddmcUseMultigroupPGRW = prob->imc.ddmc_multigroup_pgrw != 0;
    ip.withEgTimeAvg = prob->groups.with_eg_time_avg != 0;
    RadiationIMC<Grid> imc(grid, &bc, cells, ext, mat, ip, prob->seed);
    std::unique_ptr<PopulationControl> comb;
    StratifiedComb<Grid> *strat = nullptr;
    if (prob->imc.comb == 2) {
      auto sc = std::make_unique<StratifiedComb<Grid>>(grid, mat.bounds, prob->imc.comb_nmin, prob->imc.comb_factor,
                                                        prob->imc.comb_min_per_group, prob->seed);
      strat = sc.get();
      comb = std::move(sc);
    } else if (prob->imc.comb) {
      comb = std::make_unique<Comb<Grid>>(grid, prob->imc.comb_nmin, prob->imc.comb_factor, prob->seed);
    }
    Manager<Grid> mgr(grid, imc, comb.get(), bc);
    std::vector<Particle> census;
    for (int s = 0; s < nsteps; ++s) {
      auto t0 = std::chrono::steady_clock::now();
      double censusIn = 0, matBefore = 0;
      for (const Particle &p : census) censusIn += p.weight;
      for (size_t i = 0; i < N; ++i) matBefore += ext[i].internal_energy;
      imc.ledger = EnergyLedger{};
      bc.escapedEnergy = 0;
      bc.sourceEnergy = 0;
      if (source_T) bc.setSourceTemperature(source_T[s]);
      census = mgr.step(std::move(census), dts[s]);
      auto t1 = std::chrono::steady_clock::now();
      double censusOut = 0, censusComb = 0, matAfter = 0, segs = 0;
      for (const Particle &p : mgr.lastPreCombCensus()) censusOut += p.weight;
      for (const Particle &p : census) censusComb += p.weight;
      for (size_t i = 0; i < N; ++i) matAfter += ext[i].internal_energy;
      for (size_t i = 0; i < N; ++i) segs += static_cast<double>(mgr.cellsStepsCounters[i]);
      if (ledger_out) {
        double *L = ledger_out + static_cast<size_t>(s) * kLedgerStride;
        L[0] = imc.ledger.emitted;
        L[1] = bc.sourceEnergy;
        L[2] = bc.escapedEnergy;
        L[3] = imc.ledger.clamp;
        L[4] = censusIn;
        L[5] = censusOut;
        L[6] = censusComb;
        L[7] = matBefore;
        L[8] = matAfter;
        L[9] = segs;
        L[10] = static_cast<double>(mgr.startParticleCount);
        L[11] = static_cast<double>(mgr.endParticleCount);
        L[12] = std::chrono::duration<double>(t1 - t0).count();
        L[13] = strat ? strat->dropped : 0.0;
      }
    }
    for (size_t i = 0; i < N; ++i) {
      if (T_out) T_out[i] = cells[i].temperature;
      if (e_out) e_out[i] = ext[i].internal_energy;
      if (erad_out) erad_out[i] = imc.eradTimeAvg()[i];
      if (census_out) census_out[i] = ext[i].Erad;
      if (segments_out) segments_out[i] = static_cast<double>(mgr.cellsStepsCounters[i]);
    }
    if (eg_out) {
      const std::vector<double> &eg = imc.egTimeAvg();
      std::copy(eg.begin(), eg.end(), eg_out);
    }
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

}  // namespace

extern "C" {

// Per-step ledger layout (doubles): emitted, source, escaped, clamp,
// censusIn, censusOut, censusComb, materialBefore, materialAfter, segments,
// startParticles, endParticles, seconds, combDropped (the stratified comb's
// dropped bins)
#define ORACLE_LEDGER_STRIDE 14

const char *oracle_last_error(void) { return g_err.c_str(); }

// Run `nsteps` IMC steps of `prob` with time step `dt`.  `source_T` (nullable)
// overrides the boundary temperature per step.  Outputs are per cell after the
// last step (any may be NULL): T, internal energy, Erad time average,
// census Erad (extensive), segment counts of the last step (as double), and a
// per-step ledger (nsteps * ORACLE_LEDGER_STRIDE).
int oracle_run_schedule(const storm_problem *prob, int nsteps, const double *dts, const double *source_T,
                        double *T_out, double *e_out, double *erad_out, double *census_out,
                        double *segments_out, double *ledger_out);

// oracle_run_schedule with the per-group time-averaged radiation energy of
// the last step as well (eg_out: ncells * num_groups, [cell * G + g]; filled
// when the problem asks for it, storm_group_desc.with_eg_time_avg).
int oracle_run_groups(const storm_problem *prob, int nsteps, const double *dts, const double *source_T,
                      double *T_out, double *e_out, double *erad_out, double *census_out, double *segments_out,
                      double *ledger_out, double *eg_out);

int oracle_run(const storm_problem *prob, int nsteps, double dt, const double *source_T,
               double *T_out, double *e_out, double *erad_out, double *census_out,
               double *segments_out, double *ledger_out) {
  std::vector<double> dts(static_cast<size_t>(nsteps > 0 ? nsteps : 0), dt);
  return oracle_run_schedule(prob, nsteps, dts.data(), source_T, T_out, e_out, erad_out, census_out, segments_out,
                             ledger_out);
}

// oracle_run with a time step per step (dts[s]), as the reference's example
// drivers change dt between steps (examples/marshak_wave/MarshakCommon.hpp:343-357).
int oracle_run_schedule(const storm_problem *prob, int nsteps, const double *dts, const double *source_T,
                        double *T_out, double *e_out, double *erad_out, double *census_out,
                        double *segments_out, double *ledger_out) {
  return oracle_run_groups(prob, nsteps, dts, source_T, T_out, e_out, erad_out, census_out, segments_out, ledger_out,
                           nullptr);
}

int oracle_run_groups(const storm_problem *prob, int nsteps, const double *dts, const double *source_T,
                      double *T_out, double *e_out, double *erad_out, double *census_out, double *segments_out,
                      double *ledger_out, double *eg_out) {
  try {
    if (prob->unstructured) {
      UnstructuredGrid grid(*prob->unstructured);
      return run_schedule_on(grid, prob, nsteps, dts, source_T, T_out, e_out, erad_out, census_out, segments_out,
                             ledger_out, eg_out);
    }
    CartesianGrid grid = makeGrid(prob->mesh);
    return run_schedule_on(grid, prob, nsteps, dts, source_T, T_out, e_out, erad_out, census_out, segments_out,
                           ledger_out, eg_out);
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// Per-cell preStep quantities for one step from T0 (no transport): Planck
// opacity, Fleck factor, emission energy E_i, photons n_i.
int oracle_prestep(const storm_problem *prob, double dt, double *sigma_out, double *fleck_out,
                   double *emit_out, double *nphot_out) {
  try {
    CartesianGrid grid = makeGrid(prob->mesh);
    size_t N = grid.GetPointNo();
    Material mat = makeMaterial(prob->material, N, &prob->groups);
    std::vector<Cell> cells(N);
    std::vector<Extensives> ext(N);
    for (size_t i = 0; i < N; ++i) {
      double rho = mat.density(i), V = grid.GetVolume(i);
      cells[i].temperature = prob->T0[i];
      cells[i].internalEnergy = mat.eFromT(prob->T0[i], rho) * V;
      ext[i].mass = rho * V;
      ext[i].internal_energy = cells[i].internalEnergy;
    }
    IMCParams ip;
    ip.newPhotonsPerCell = prob->imc.new_photons_per_cell;
    RadiationIMC<CartesianGrid> imc(grid, nullptr, cells, ext, mat, ip, prob->seed);
    imc.preStep(dt);
    for (size_t i = 0; i < N; ++i) {
      if (sigma_out) sigma_out[i] = imc.planckOpacities()[i];
      if (fleck_out) fleck_out[i] = imc.factorFleck()[i];
      if (emit_out) emit_out[i] = imc.emissionEnergy()[i];
      if (nphot_out) nphot_out[i] = static_cast<double>(imc.photonsPerCell()[i]);
    }
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// Photon allocation of generateParticles for a given emission-energy vector.
int oracle_alloc(int64_t n, const double *E, uint64_t new_photons, int64_t *out) {
  try {
    std::vector<double> e(E, E + n);
    double total = 0.0;
    for (double v : e) total += v;
    std::vector<size_t> c = photonAllocation(e, total, static_cast<size_t>(new_photons));
    for (int64_t i = 0; i < n; ++i) out[i] = static_cast<int64_t>(c[static_cast<size_t>(i)]);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

namespace {
struct OracleState {
  CartesianGrid grid;
  Material mat;
  std::vector<Cell> cells;
  std::vector<Extensives> ext;
  std::unique_ptr<RadiationIMC<CartesianGrid>> imc;
  OracleState(const storm_problem *prob, double dt)
      : grid(makeGrid(prob->mesh)), mat(makeMaterial(prob->material, grid.GetPointNo(), &prob->groups)) {
    const size_t N = grid.GetPointNo();
    cells.resize(N);
    ext.resize(N);
    for (size_t i = 0; i < N; ++i) {
      double rho = mat.density(i), V = grid.GetVolume(i);
      cells[i].temperature = prob->T0[i];
      cells[i].internalEnergy = mat.eFromT(prob->T0[i], rho) * V;
      ext[i].mass = rho * V;
      ext[i].internal_energy = cells[i].internalEnergy;
    }
    IMCParams ip;
    ip.newPhotonsPerCell = prob->imc.new_photons_per_cell;
    ip.withRandomWalk = false;
    ip.withDDMC = true;
    ip.ddmcMinCellOpticalDepth = prob->imc.ddmc_min_cell_optical_depth;
    ip.ddmcMinParticleOpticalDepth = prob->imc.ddmc_min_particle_optical_depth;
    ip.ddmcUseMultigroupPGRW = prob->imc.ddmc_multigroup_pgrw != 0;
    imc = std::make_unique<RadiationIMC<CartesianGrid>>(grid, nullptr, cells, ext, mat, ip, prob->seed);
    imc->preStep(dt);
  }
};
}  // namespace

// DDMC cell data of preStep (precomputeDDMCData) from T0, DDMC on and the
// random walk off: per cell eligible (0/1), sigma_A, sigma_T, D (cm^2/s),
// total leak rate (1/s), the number of face leaks and up to 6 of them
// (neighbour cell, rate in 1/s).
int oracle_ddmc_data(const storm_problem *prob, double dt, double *eligible, double *sigma_a, double *sigma_t,
                     double *D, double *total, int64_t *nleaks, int64_t *leak_next, double *leak_rate) {
  try {
    OracleState st(prob, dt);
    const auto &dd = st.imc->ddmcData();
    for (size_t i = 0; i < dd.size(); ++i) {
      eligible[i] = dd[i].eligible ? 1.0 : 0.0;
      sigma_a[i] = dd[i].sigmaA;
      sigma_t[i] = dd[i].sigmaT;
      D[i] = dd[i].diffusionCoefficient;
      total[i] = dd[i].totalLeakRate;
      nleaks[i] = static_cast<int64_t>(dd[i].faceLeaks.size());
      for (size_t q = 0; q < 6; ++q) {
        const bool has = q < dd[i].faceLeaks.size();
        leak_next[6 * i + q] = has ? static_cast<int64_t>(dd[i].faceLeaks[q].nextCellIndex) : -1;
        leak_rate[6 * i + q] = has ? dd[i].faceLeaks[q].rate : 0.0;
      }
    }
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// n independent DDMC attempts (tryDDMCStep) on a packet of weight 1 at
// (x, y, z) in `cell` with `time_left` s: per attempt the outcome (-1: no
// DDMC step, else the particle status), the next cell, the new direction
// (3 per attempt), the time used and the weight after.
int oracle_ddmc_samples(const storm_problem *prob, double dt, int64_t cell, double x, double y, double z,
                        double time_left, int64_t n, int64_t *status, int64_t *next, double *dir, double *t_used,
                        double *w_after) {
  try {
    OracleState st(prob, dt);
    for (int64_t k = 0; k < n; ++k) {
      Particle p;
      p.cellIndex = static_cast<size_t>(cell);
      p.location = Vec3(x, y, z);
      p.velocity = Vec3(units::clight, 0.0, 0.0);
      p.timeLeft = time_left;
      p.weight = 1.0;
      p.initialWeight = 1.0;
      p.frequency = 0.0;
      StepResult fn;
      const bool took = st.imc->ddmcStep(p, fn);
      status[k] = took ? static_cast<int64_t>(fn.change) : -1;
      next[k] = took && fn.change == CELL_MOVE ? static_cast<int64_t>(fn.nextCellIndex) : -1;
      dir[3 * k] = p.velocity.x / units::clight;
      dir[3 * k + 1] = p.velocity.y / units::clight;
      dir[3 * k + 2] = p.velocity.z / units::clight;
      t_used[k] = time_left - p.timeLeft;
      w_after[k] = p.weight;
    }
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// Random-walk tables: tau[1024], survival[1024], radius[1024*256].
int oracle_rw_tables(double *tau, double *survival, double *radius) {
  try {
    RandomWalk rw;
    std::memcpy(tau, rw.tauTable().data(), rw.tauTable().size() * sizeof(double));
    std::memcpy(survival, rw.survivalTable().data(), rw.survivalTable().size() * sizeof(double));
    std::memcpy(radius, rw.radiusTable().data(), rw.radiusTable().size() * sizeof(double));
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// RandomWalk::sampleLeakTime / sampleRadius for a batch of inputs.
int oracle_rw_sample(int n, const double *xi, const double *tau, double *leak_out, double *radius_out) {
  try {
    static RandomWalk rw;
    for (int i = 0; i < n; ++i) {
      if (leak_out) leak_out[i] = rw.sampleLeakTime(xi[i]);
      if (radius_out) radius_out[i] = rw.sampleRadius(tau[i], xi[i]);
    }
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// Comb on a synthetic census in one cell-indexed grid: inputs are per-packet
// cell and weight; outputs per-cell output count and output weight sum.
int oracle_comb(const storm_problem *prob, int64_t n, const int64_t *cell, const double *weight,
                double *count_out, double *wsum_out) {
  try {
    CartesianGrid grid = makeGrid(prob->mesh);
    size_t N = grid.GetPointNo();
    std::vector<Particle> ps(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
      ps[i].id = static_cast<size_t>(i);
      ps[i].cellIndex = static_cast<size_t>(cell[i]);
      ps[i].location = grid.GetCellCM(ps[i].cellIndex);
      ps[i].weight = weight[i];
    }
    Comb<CartesianGrid> comb(grid, prob->imc.comb_nmin, prob->imc.comb_factor, prob->seed);
    std::vector<Particle> out = comb.activate(ps);
    for (size_t i = 0; i < N; ++i) {
      count_out[i] = 0;
      wsum_out[i] = 0;
    }
    for (const Particle &p : out) {
      count_out[p.cellIndex] += 1;
      wsum_out[p.cellIndex] += p.weight;
    }
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

// StratifiedCombPopulationControl::activate on a hand-built census of the
// problem's Cartesian mesh: packet i in cell[i] with photon energy freq[i]
// (erg) and weight[i]; per cell and group (cell * G + g, G the problem's
// groups, 1 when gray) the count and weight of the packets out, and the
// weight of the dropped bins.
int oracle_strat_comb(const storm_problem *prob, int64_t n, const int64_t *cell, const double *freq,
                      const double *weight, double *count_out, double *wsum_out, double *dropped_out) {
  try {
    CartesianGrid grid = makeGrid(prob->mesh);
    const size_t N = grid.GetPointNo();
    std::vector<double> bounds;
    if (prob->groups.num_groups >= 2) bounds.assign(prob->groups.bounds, prob->groups.bounds + prob->groups.num_groups + 1);
    const size_t G = bounds.size() > 1 ? bounds.size() - 1 : 1;
    std::vector<Particle> ps(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
      ps[i].id = static_cast<size_t>(i);
      ps[i].cellIndex = static_cast<size_t>(cell[i]);
      ps[i].location = grid.GetCellCM(ps[i].cellIndex);
      ps[i].weight = weight[i];
      ps[i].frequency = freq[i];
    }
    StratifiedComb<CartesianGrid> comb(grid, bounds, prob->imc.comb_nmin, prob->imc.comb_factor,
                                       prob->imc.comb_min_per_group, prob->seed);
    std::vector<Particle> out = comb.activate(ps);
    for (size_t i = 0; i < N * G; ++i) {
      count_out[i] = 0;
      wsum_out[i] = 0;
    }
    for (const Particle &p : out) {
      const size_t b = p.cellIndex * G + comb.classify(p);
      count_out[b] += 1;
      wsum_out[b] += p.weight;
    }
    *dropped_out = comb.dropped;
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

}  // extern "C"
