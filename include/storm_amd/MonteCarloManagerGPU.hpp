// Drop-in manager for one MI355X: the public interface of the reference's
// MonteCarloManagerSerial (manager/MonteCarloManagerSerial.hpp:25-117) over
// the gfx950 engine of include/storm_gpu.h.
//
//   MonteCarloManagerSerial(grid, physics, populationControl, boundaryCondition)   :38-40
//   std::vector<Particle> step(std::vector<Particle> &&census, dt)                 :44, :385-469
//   GetCellsStepsCounters / GetStartParticleCount / GetInitialParticleCount /
//   GetPreStepParticleCount / GetEndParticleCount / GetBeginningParticleCount /
//   GetPureComputeTime / GetHandlerMemoryBytes / getTracker / resetTracker       :46-83
//
// What the manager takes from the plugins, once per step on the host:
//   * the mesh: a Cartesian Grid (GetPointNo, GetCellCM, GetVolume,
//     GetBoxCoordinates) with cells ordered (i*ny + j)*nz + k;
//   * the physics: RadiationIMC (include/storm_amd/plugins.hpp), whose opacity
//     model and EOS are tabulated per cell before the step and which receives
//     the new internal energies, temperatures and Erad after it;
//   * the boundary: SideTemperature, VacuumBoundary or RigidBoundary describe
//     their faces themselves; any other BoundaryCondition is probed once
//     through apply() with a packet leaving each box face (REMOVE = escape,
//     REFLECT = specular), and the packets its generateNewBoundaryParticles
//     makes (a MarshakBoundary, say) are made on the host every step and
//     joined to the step's sources on the GPU, as RadiationIMC::preStep joins
//     them (RadiationIMC.hpp:1683-1692);
//   * population control: CombPopulationControl (Nmin, factor) and
//     StratifiedCombPopulationControl (classified by photon group) run on the
//     GPU, NoPopulationControl passes the census through; other controls are
//     rejected;
//   * multigroup physics (NumGroups >= 2): the groups, the opacity's energy
//     law (checked by RadiationIMC) and, with withEgTimeAvg, the per-group
//     time-averaged radiation energy back after each step.
// The packets themselves never leave HBM during a step.  By default step()
// returns the census as std::vector<Particle> like the reference and takes it
// back on the next call; setResidentCensus(true) keeps it in HBM instead
// (step() then returns an empty vector and an empty argument continues from
// the resident census).
#pragma once

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <type_traits>
#include <string>
#include <utility>
#include <vector>

#include "../storm_gpu.h"
#include "Particle.hpp"
#include "StormError.hpp"
#include "plugins.hpp"

namespace STORM {

enum class GpuPrecision { Mixed = STORM_GPU_MIXED, FP64 = STORM_GPU_FP64 };

namespace detail {
template <typename G, typename = void>
struct has_axis_edges : std::false_type {};
template <typename G>
struct has_axis_edges<G, std::void_t<decltype(std::declval<const G &>().xEdges()),
                                     decltype(std::declval<const G &>().yEdges()),
                                     decltype(std::declval<const G &>().zEdges())>> : std::true_type {};
}  // namespace detail

template <typename T, typename Grid>
class MonteCarloManagerGPU {
 public:
  using MCParticle = Particle<T, Grid>;

  struct MonteCarloStepFinalData {
    std::vector<MCParticle> remaining;
    size_t leavingCount = 0;
  };

  // Packet routes are not recorded on the GPU; the tracker stays empty.
  class Tracker {
   public:
    void Reset(void) {}
    std::vector<MCParticle> GetTrackParticleRoute(size_t) const { return {}; }
  };

  MonteCarloManagerGPU(const Grid &grid, const std::shared_ptr<MonteCarloPhysics<T, Grid>> &physics,
                       const std::shared_ptr<PopulationControl<T, Grid>> &populationControl,
                       const std::shared_ptr<BoundaryCondition<T, Grid>> &boundaryCondition,
                       GpuPrecision precision = GpuPrecision::Mixed)
      : grid_(grid), physics_(physics), population_(populationControl), boundary_(boundaryCondition),
        precision_(precision) {
    if (!physics_) throw StormError("MonteCarloManagerGPU: physics is null");
    if (!population_) throw StormError("MonteCarloManagerGPU: population control is null");
    if (!boundary_) throw StormError("MonteCarloManagerGPU: boundary condition is null");
    imc_ = dynamic_cast<GpuRadiationPhysics *>(physics_.get());
    if (!imc_) throw StormError("MonteCarloManagerGPU runs RadiationIMC physics (plugins.hpp)");
    Ncells_ = grid_.GetPointNo();
    describeMesh();
    describeBoundary();
    describePopulation();
  }

  MonteCarloManagerGPU(const MonteCarloManagerGPU &) = delete;
  MonteCarloManagerGPU &operator=(const MonteCarloManagerGPU &) = delete;
  virtual ~MonteCarloManagerGPU() {
    if (eng_) storm_gpu_destroy(eng_);
  }

  // One IMC step: census in, census out (MonteCarloManagerSerial.hpp:385-469).
  std::vector<MCParticle> step(std::vector<MCParticle> &&particleList, dt_t fullDt) {
    if (!(fullDt > 0.0)) throw StormError("MonteCarloManagerGPU::step: fullDt must be positive");
    physics_->updateGridData();
    ensureEngine();
    // host plugins, tabulated once per cell (RadiationIMC.hpp:1645-1648, :1753)
    imc_->gpuTabulate(T_.data(), sa_.data(), ss_.data(), cv_.data(), e_.data(), mass_.data());
    check(storm_gpu_set_cell_state(eng_, T_.data(), sa_.data(), ss_.data(), cv_.data(), e_.data()));
    if (sourceT_) check(storm_gpu_set_source_temperature(eng_, sourceT_()));
    if (!resident_ || !particleList.empty()) importCensus(particleList);
    particleList.clear();
    if (hostSource_) queueBoundaryParticles(boundary_->generateNewBoundaryParticles(fullDt));
    check(storm_gpu_step(eng_, fullDt));
    check(storm_gpu_get_counters(eng_, &counters_));
    std::vector<double> erad(Ncells_), cen(Ncells_), fleck(Ncells_), planck(Ncells_);
    check(storm_gpu_get_cells(eng_, nullptr, e_.data(), erad.data(), cen.data(), fleck.data(), planck.data()));
    imc_->gpuAbsorb(e_.data(), erad.data(), cen.data(), fleck.data(), planck.data());
    if (egOn_) {
      std::vector<double> eg(Ncells_ * (bounds_.size() - 1));
      check(storm_gpu_get_group_erad(eng_, eg.data()));
      imc_->gpuGroupErad(eg.data());
    }
    imc_->gpuCounters(counters_.random_walk_steps, counters_.ddmc_steps);
    if (auto *vac = dynamic_cast<VacuumBoundary<T, Grid> *>(boundary_.get())) {
      storm_gpu_ledger L;
      check(storm_gpu_get_ledger(eng_, &L));
      const double Q = std::ldexp(1.0, L.eq);
      vac->addEscapedEnergy((static_cast<double>(L.escaped.hi) * 18446744073709551616.0 +
                             static_cast<double>(L.escaped.lo)) * Q);
    }
    steps_.assign(Ncells_, 0);
    begin_.assign(Ncells_, 0);
    std::vector<uint64_t> tmp(Ncells_);
    check(storm_gpu_get_segments(eng_, tmp.data()));
    std::copy(tmp.begin(), tmp.end(), steps_.begin());
    check(storm_gpu_get_beginning_counts(eng_, tmp.data()));
    std::copy(tmp.begin(), tmp.end(), begin_.begin());
    if (resident_) return {};
    return exportCensus();
  }

  // --------------------------------------------------------------- reference getters
  const Tracker &getTracker(void) { return tracker_; }
  void resetTracker(void) { tracker_.Reset(); }
  const std::vector<size_t> &GetCellsStepsCounters(void) const { return steps_; }
  std::vector<size_t> &GetCellsStepsCounters(void) { return steps_; }
  size_t GetStartParticleCount(void) const { return counters_.start_particles; }
  size_t GetInitialParticleCount(void) const { return counters_.initial_particles; }
  size_t GetPreStepParticleCount(void) const { return counters_.prestep_particles; }
  size_t GetEndParticleCount(void) const { return counters_.end_particles; }
  // GPU time of the transport kernels of the last step, s
  double GetPureComputeTime(void) const { return counters_.seconds_transport; }
  const std::vector<size_t> &GetBeginningParticleCount(void) const { return begin_; }
  std::vector<size_t> &GetBeginningParticleCount(void) { return begin_; }
  size_t GetHandlerMemoryBytes(void) const { return 0; }

  // --------------------------------------------------------------- GPU extras
  void setResidentCensus(bool on) { resident_ = on; }
  size_t residentCensusSize() {
    ensureEngine();
    uint64_t n = 0;
    check(storm_gpu_census_size(eng_, &n));
    return n;
  }
  std::vector<MCParticle> exportCensus() {
    ensureEngine();
    uint64_t n = 0, w = 0;
    check(storm_gpu_census_size(eng_, &n));
    std::vector<storm_packet_record> recs(n);
    check(storm_gpu_export_census(eng_, recs.data(), n, &w));
    std::vector<MCParticle> out;
    out.reserve(w);
    for (uint64_t i = 0; i < w; ++i) {
      const storm_packet_record &r = recs[i];
      MCParticle p(r.id, T(r.x, r.y, r.z), T(r.vx, r.vy, r.vz), 0.0);
      p.cellIndex = r.cell;
      p.weight = r.weight;
      p.initialWeight = r.weight;
      p.frequency = r.frequency;
      out.push_back(p);
    }
    return out;
  }
  const storm_gpu_counters &counters() const { return counters_; }
  storm_gpu_ledger ledger() {
    storm_gpu_ledger L;
    check(storm_gpu_get_ledger(eng_, &L));
    return L;
  }
  void setTuning(int block, int segmentsPerPass, int schedule) {
    ensureEngine();
    check(storm_gpu_set_tuning(eng_, block, segmentsPerPass, schedule));
  }
  // Boundary temperature read before every step (MarshakBoundary-style drives).
  template <typename F>
  void setSourceTemperatureProvider(F &&f) {
    sourceT_ = std::forward<F>(f);
  }
  const std::vector<double> &xEdges() const { return xe_; }
  const std::vector<double> &yEdges() const { return ye_; }
  const std::vector<double> &zEdges() const { return ze_; }
  int faceKind(int face) const { return bnd_.face_kind[face]; }

 private:
  static void check(int rc) {
    if (rc != 0) throw StormError(std::string("gfx950 engine: ") + storm_gpu_last_error());
  }

  // Axis edges: taken as they are from grids that expose them, else
  // recovered from the cell centres (on a Cartesian grid each axis' centres
  // are the midpoints of consecutive edges, starting at the box corner).
  void describeMesh() {
    if constexpr (detail::has_axis_edges<Grid>::value) {
      auto copy = [](const auto &src, std::vector<double> &dst) { dst.assign(src.begin(), src.end()); };
      copy(grid_.xEdges(), xe_);
      copy(grid_.yEdges(), ye_);
      copy(grid_.zEdges(), ze_);
      checkCellOrder();
      return;
    }
    const auto box = grid_.GetBoxCoordinates();
    auto axis = [&](auto get, double lo, double hi, std::vector<double> &edges) {
      std::vector<double> c;
      c.reserve(Ncells_);
      for (size_t i = 0; i < Ncells_; ++i) c.push_back(get(grid_.GetCellCM(i)));
      std::sort(c.begin(), c.end());
      const double tol = 1e-9 * (hi - lo);
      std::vector<double> u;
      for (double v : c)
        if (u.empty() || v - u.back() > tol) u.push_back(v);
      edges.assign(1, lo);
      for (double m : u) edges.push_back(2.0 * m - edges.back());
      edges.back() = hi;
      for (size_t q = 1; q < edges.size(); ++q)
        if (!(edges[q] > edges[q - 1])) throw StormError("MonteCarloManagerGPU: the grid is not Cartesian");
    };
    axis([](const T &p) { return static_cast<double>(p.x); }, box.first.x, box.second.x, xe_);
    axis([](const T &p) { return static_cast<double>(p.y); }, box.first.y, box.second.y, ye_);
    axis([](const T &p) { return static_cast<double>(p.z); }, box.first.z, box.second.z, ze_);
    checkCellOrder();
  }

  void checkCellOrder() {
    const size_t nx = xe_.size() - 1, ny = ye_.size() - 1, nz = ze_.size() - 1;
    if (nx * ny * nz != Ncells_) throw StormError("MonteCarloManagerGPU: the grid is not a full Cartesian box");
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j)
        for (size_t k = 0; k < nz; ++k) {
          const size_t c = (i * ny + j) * nz + k;
          const T cm = grid_.GetCellCM(c);
          const double V = (xe_[i + 1] - xe_[i]) * (ye_[j + 1] - ye_[j]) * (ze_[k + 1] - ze_[k]);
          const double dx = std::abs(cm.x - 0.5 * (xe_[i] + xe_[i + 1]));
          const double dy = std::abs(cm.y - 0.5 * (ye_[j] + ye_[j + 1]));
          const double dz = std::abs(cm.z - 0.5 * (ze_[k] + ze_[k + 1]));
          if (dx + dy + dz > 1e-9 * (xe_.back() - xe_.front()) || std::abs(grid_.GetVolume(c) - V) > 1e-9 * V)
            throw StormError("MonteCarloManagerGPU: cells must be ordered (i*ny + j)*nz + k");
        }
  }

  void describeBoundary() {
    GpuBoundaryInfo info;
    if (!boundary_->describeForGpu(info)) info = probeBoundary();
    for (int f = 0; f < 6; ++f) bnd_.face_kind[f] = info.face_kind[f];
    bnd_.source_face = info.source ? STORM_XLO : -1;
    bnd_.freeze_source_T4 = info.freeze_T4 ? 1 : 0;
    bnd_.source_temperature = info.temperature;
    bnd_.source_packets_per_face = info.packets_per_face;
    bnd_.source_xhi = info.source_xhi ? 1 : 0;
    bnd_.source_temperature_xhi = info.temperature_xhi;
    if (auto *st = dynamic_cast<SideTemperature<T, Grid> *>(boundary_.get()))
      sourceT_ = [st]() { return st->GetTemperature(); };
  }

  // A packet at the centre of each box face, leaving the box.
  GpuBoundaryInfo probeBoundary() {
    GpuBoundaryInfo info;
    const auto box = grid_.GetBoxCoordinates();
    const double lo[3] = {box.first.x, box.first.y, box.first.z}, hi[3] = {box.second.x, box.second.y, box.second.z};
    for (int f = 0; f < 6; ++f) {
      const int ax = f / 2;
      const bool up = (f & 1) != 0;
      double x[3], v[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < 3; ++a) x[a] = 0.5 * (lo[a] + hi[a]);
      x[ax] = up ? hi[ax] : lo[ax];
      v[ax] = up ? 1.0 : -1.0;
      MCParticle p(0, T(x[0], x[1], x[2]), T(v[0], v[1], v[2]), 0.0);
      p.weight = 0.0;
      p.initialWeight = 0.0;
      p.cellIndex = grid_.GetContainingCell(T(x[0], x[1], x[2]));
      const ParticleStatus s = boundary_->apply(p);
      if (s == ParticleStatus::REMOVE) info.face_kind[f] = STORM_FACE_ESCAPE;
      else if (s == ParticleStatus::REFLECT) info.face_kind[f] = STORM_FACE_REFLECT;
      else throw StormError("MonteCarloManagerGPU: boundary apply() must return REMOVE or REFLECT");
    }
    // a boundary that makes its own packets: made on the host every step
    hostSource_ = !boundary_->generateNewBoundaryParticles(1e-30).empty();
    return info;
  }

  void describePopulation() {
    imc_params_ = storm_imc_params{};
    imc_->gpuParameters(imc_params_);
    groups_ = imc_->gpuGroups(bounds_, nuPow_, nuRef_, egOn_);
    if (auto *sc = dynamic_cast<StratifiedCombPopulationControl<T, Grid> *>(population_.get())) {
      // the GPU classifies by photon group: the classifier must agree on a
      // packet at the centre of every group
      const size_t G = groups_ ? bounds_.size() - 1 : 1;
      if (sc->getGroupCount() != G)
        throw StormError("MonteCarloManagerGPU: the stratified comb's group count must be the problem's groups");
      for (size_t g = 0; groups_ && g < G; ++g) {
        MCParticle p;
        p.frequency = 0.5 * (bounds_[g] + bounds_[g + 1]);
        p.cellIndex = 0;
        if (sc->classify(p) != g)
          throw StormError("MonteCarloManagerGPU: the stratified comb runs on the GPU classified by photon group");
      }
      imc_params_.comb = 2;
      imc_params_.comb_nmin = sc->getNmin();
      imc_params_.comb_factor = sc->getTotalParticlesFactor();
      imc_params_.comb_min_per_group = sc->getMinParticlesPerGroup();
    } else if (auto *comb = dynamic_cast<CombPopulationControl<T, Grid> *>(population_.get())) {
      imc_params_.comb = 1;
      imc_params_.comb_nmin = comb->getNmin();
      imc_params_.comb_factor = comb->getTotalParticlesFactor();
    } else if (dynamic_cast<NoPopulationControl<T, Grid> *>(population_.get())) {
      imc_params_.comb = 0;
    } else {
      throw StormError("MonteCarloManagerGPU: population control must be Comb, StratifiedComb or NoPopulationControl");
    }
  }

  void ensureEngine() {
    if (eng_) return;
    T_.resize(Ncells_);
    sa_.resize(Ncells_);
    ss_.resize(Ncells_);
    cv_.resize(Ncells_);
    e_.resize(Ncells_);
    mass_.resize(Ncells_);
    imc_->gpuTabulate(T_.data(), sa_.data(), ss_.data(), cv_.data(), e_.data(), mass_.data());
    storm_problem prob{};
    prob.mesh.nx = static_cast<int32_t>(xe_.size() - 1);
    prob.mesh.ny = static_cast<int32_t>(ye_.size() - 1);
    prob.mesh.nz = static_cast<int32_t>(ze_.size() - 1);
    prob.mesh.x_edges = xe_.data();
    prob.mesh.y_edges = ye_.data();
    prob.mesh.z_edges = ze_.data();
    // material values come from the host plugins every step
    prob.material.sigma_a0 = sa_.data();
    prob.material.opacity_T_ref = 1.0;
    prob.material.eos_f = 1.0;
    prob.material.eos_beta = 1.0;
    prob.boundary = bnd_;
    prob.imc = imc_params_;
    if (groups_) {
      prob.groups.num_groups = static_cast<int32_t>(bounds_.size() - 1);
      prob.groups.with_eg_time_avg = egOn_ ? 1 : 0;
      prob.groups.bounds = bounds_.data();
      prob.groups.nu_pow = nuPow_;
      prob.groups.nu_ref = nuRef_;
      // photon temperatures of the Planck fractions floored at 1 K (the
      // reference's multigroup models take max(T, 1 K), e.g.
      // examples/densmore2012/DensmoreOpacity.hpp:40)
      prob.material.opacity_T_floor = 1.0;
    }
    prob.T0 = T_.data();
    prob.seed = imc_->gpuSeed();
    check(storm_gpu_create(&prob, static_cast<int32_t>(precision_), STORM_GPU_MATERIAL_HOST, &eng_));
  }

  void queueBoundaryParticles(const std::vector<MCParticle> &particles) {
    std::vector<storm_packet_record> recs(particles.size());
    std::vector<double> tl(particles.size());
    for (size_t i = 0; i < particles.size(); ++i) {
      const MCParticle &p = particles[i];
      storm_packet_record &r = recs[i];
      r.x = p.location.x;
      r.y = p.location.y;
      r.z = p.location.z;
      r.vx = p.velocity.x;
      r.vy = p.velocity.y;
      r.vz = p.velocity.z;
      r.weight = p.weight;
      r.frequency = p.frequency;
      r.cell = p.cellIndex < Ncells_ ? p.cellIndex : grid_.GetContainingCell(p.location);
      r.id = p.id;
      tl[i] = p.timeLeft;
    }
    check(storm_gpu_queue_source_packets(eng_, recs.data(), tl.data(), recs.size()));
  }

  void importCensus(const std::vector<MCParticle> &particles) {
    std::vector<storm_packet_record> recs(particles.size());
    for (size_t i = 0; i < particles.size(); ++i) {
      const MCParticle &p = particles[i];
      storm_packet_record &r = recs[i];
      r.x = p.location.x;
      r.y = p.location.y;
      r.z = p.location.z;
      r.vx = p.velocity.x;
      r.vy = p.velocity.y;
      r.vz = p.velocity.z;
      r.weight = p.weight;
      r.frequency = p.frequency;
      r.cell = p.cellIndex < Ncells_ ? p.cellIndex : grid_.GetContainingCell(p.location);
      r.id = p.id;
    }
    check(storm_gpu_import_census(eng_, recs.data(), recs.size()));
  }

  const Grid &grid_;
  std::shared_ptr<MonteCarloPhysics<T, Grid>> physics_;
  std::shared_ptr<PopulationControl<T, Grid>> population_;
  std::shared_ptr<BoundaryCondition<T, Grid>> boundary_;
  GpuRadiationPhysics *imc_ = nullptr;
  GpuPrecision precision_;
  size_t Ncells_ = 0;
  std::vector<double> xe_, ye_, ze_;
  storm_boundary_desc bnd_{};
  storm_imc_params imc_params_{};
  storm_gpu_engine *eng_ = nullptr;
  bool resident_ = false;
  bool hostSource_ = false;  // the boundary makes its packets on the host
  bool groups_ = false, egOn_ = false;  // multigroup; with the per-group tally
  std::vector<double> bounds_;
  double nuPow_ = 0.0, nuRef_ = 1.0;
  std::function<double()> sourceT_;
  std::vector<double> T_, sa_, ss_, cv_, e_, mass_;
  std::vector<size_t> steps_, begin_;
  storm_gpu_counters counters_{};
  Tracker tracker_;
};

// The reference's name for the manager of this path.
template <typename T, typename Grid>
using MonteCarloManagerSerial = MonteCarloManagerGPU<T, Grid>;

}  // namespace STORM
