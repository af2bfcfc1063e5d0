// The reference's plugin surface for the IMC path, for use with
// MonteCarloManagerGPU: same class names, template parameters, constructors
// and public methods as
//   physics/MonteCarloPhysics.hpp:16-53            MonteCarloPhysics
//   boundary/BoundaryCondition.hpp:30-51           BoundaryCondition
//   boundary/SideTemperature.hpp:20-167            SideTemperature
//   boundary/TwoSidesTemperature.hpp:20-58         TwoSidesTemperature
//   boundary/Vacuum.hpp:10-45                      VacuumBoundary
//   boundary/RigidBoundary.hpp:9-24                RigidBoundary
//   population/PopulationControl.hpp:8-22          PopulationControl
//   population/CombPopulationControl.hpp:20-32     CombPopulationControl
//   population/NoPopulationControl.hpp:8-28        NoPopulationControl
//   radiation/RadiationIMCParameters.hpp:12-56     RadiationIMCParameters
//   radiation/RadiationCell.hpp:6-19               RadiationCell, SimpleExtensives
//   radiation/RadiationOpacityModel.hpp:18-151     RadiationOpacityModel
//   radiation/RadiationIMC.hpp:332-467             RadiationIMC
// The physics itself runs in the gfx950 engine (include/storm_gpu.h); these
// classes carry the configuration, the user's opacity / EOS objects (called
// once per cell per step on the host, never per segment) and the results.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <random>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../storm_problem.h"
#include "Particle.hpp"
#include "StormError.hpp"

namespace STORM {

// ------------------------------------------------------------------ boundary
enum class DDMCBoundaryFaceBehavior { Unsupported, ReflectingRigid };

// What the GPU needs to know about a boundary: the fate of a packet on each
// box face (x-, x+, y-, y+, z-, z+) and an optional Planck source on x-.
struct GpuBoundaryInfo {
  int face_kind[6] = {STORM_FACE_REFLECT, STORM_FACE_REFLECT, STORM_FACE_REFLECT,
                      STORM_FACE_REFLECT, STORM_FACE_REFLECT, STORM_FACE_REFLECT};
  bool source = false;        // Planck source on x-
  double temperature = 0.0;
  size_t packets_per_face = 0;
  bool freeze_T4 = false;
  bool source_xhi = false;    // Planck source on x+ as well (same packets per face)
  double temperature_xhi = 0.0;
};

template <typename T, typename Grid>
class BoundaryCondition {
 public:
  explicit BoundaryCondition(const Grid &grid) : grid(grid) {}
  virtual ~BoundaryCondition() = default;
  virtual ParticleStatus apply(Particle<T, Grid> &particle) = 0;
  virtual std::vector<Particle<T, Grid>> generateNewBoundaryParticles(double fullDt) = 0;
  virtual DDMCBoundaryFaceBehavior getDDMCBoundaryFaceBehavior(size_t, size_t, size_t) const {
    return DDMCBoundaryFaceBehavior::Unsupported;
  }
  // Face table for the GPU; boundaries that do not describe themselves are
  // probed through apply() by the manager.
  virtual bool describeForGpu(GpuBoundaryInfo &) const { return false; }

 protected:
  const Grid &grid;
};

// Planck surface source on the x- face; packets reaching x- leave, the other
// faces reflect (boundary/SideTemperature.hpp:84-163).  As in the reference,
// T^4 is taken at the first step (`static const double T4`, :117) unless
// freezeT4 is false (the MarshakBoundary behaviour,
// examples/marshak_wave/MarshakBoundary.hpp:150-156).
template <typename T, typename Grid>
class SideTemperature : public BoundaryCondition<T, Grid> {
 public:
  SideTemperature(const Grid &grid, double temperature, size_t Npercell,
                  const std::vector<double> &energyBoundaries = {}, bool freezeT4 = true)
      : BoundaryCondition<T, Grid>(grid), temperature(temperature), Npercell(Npercell), freezeT4(freezeT4) {
    // multigroup: the GPU draws the photon energy from the source's Planck
    // fractions over the problem's groups, linear within a group, as the
    // reference does with these boundaries (SideTemperature.hpp:148-152)
    (void)energyBoundaries;
  }
  ParticleStatus apply(Particle<T, Grid> &) override {
    throw StormError("SideTemperature::apply is evaluated on the GPU (MonteCarloManagerGPU)");
  }
  std::vector<Particle<T, Grid>> generateNewBoundaryParticles(double) override {
    throw StormError("SideTemperature sources are generated on the GPU (MonteCarloManagerGPU)");
  }
  DDMCBoundaryFaceBehavior getDDMCBoundaryFaceBehavior(size_t, size_t, size_t) const override {
    return DDMCBoundaryFaceBehavior::ReflectingRigid;
  }
  bool describeForGpu(GpuBoundaryInfo &info) const override {
    info = GpuBoundaryInfo{};
    info.face_kind[STORM_XLO] = STORM_FACE_ESCAPE;
    info.source = Npercell > 0;
    info.temperature = temperature;
    info.packets_per_face = Npercell;
    info.freeze_T4 = freezeT4;
    return true;
  }
  void SetTemperature(double temp) { temperature = temp; }
  double GetTemperature() const { return temperature; }

 private:
  double temperature;
  size_t Npercell;
  bool freezeT4;
};

// Planck sources on both x faces, which remove packets; the other faces
// reflect.  T^4 of each side is taken at the first step
// (`static const double T4_L/T4_R`, boundary/TwoSidesTemperature.hpp:128-129).
template <typename T, typename Grid>
class TwoSidesTemperature : public BoundaryCondition<T, Grid> {
 public:
  TwoSidesTemperature(const Grid &grid, double temperatureLeft, double temperatureRight, size_t Npercell,
                      const std::vector<double> &energyBoundaries = {})
      : BoundaryCondition<T, Grid>(grid), temperatureLeft(temperatureLeft), temperatureRight(temperatureRight),
        Npercell(Npercell) {
    if (!energyBoundaries.empty())
      throw StormError("TwoSidesTemperature: multigroup sources go through storm_group_desc, not the C++ manager");
  }
  ParticleStatus apply(Particle<T, Grid> &) override {
    throw StormError("TwoSidesTemperature::apply is evaluated on the GPU (MonteCarloManagerGPU)");
  }
  std::vector<Particle<T, Grid>> generateNewBoundaryParticles(double) override {
    throw StormError("TwoSidesTemperature sources are generated on the GPU (MonteCarloManagerGPU)");
  }
  bool describeForGpu(GpuBoundaryInfo &info) const override {
    info = GpuBoundaryInfo{};
    info.face_kind[STORM_XLO] = STORM_FACE_ESCAPE;
    info.face_kind[STORM_XHI] = STORM_FACE_ESCAPE;
    info.source = Npercell > 0;
    info.source_xhi = Npercell > 0;
    info.temperature = temperatureLeft;
    info.temperature_xhi = temperatureRight;
    info.packets_per_face = Npercell;
    info.freeze_T4 = true;
    return true;
  }

 private:
  double temperatureLeft, temperatureRight;
  size_t Npercell;
};

// Every face removes the packet and the escaped energy is tallied (boundary/Vacuum.hpp:18-22).
template <typename T, typename Grid>
class VacuumBoundary : public BoundaryCondition<T, Grid> {
 public:
  using MCParticle = Particle<T, Grid>;
  explicit VacuumBoundary(const Grid &grid) : BoundaryCondition<T, Grid>(grid) {}
  ParticleStatus apply(MCParticle &particle) override {
    escapedEnergy_ += particle.weight;
    return ParticleStatus::REMOVE;
  }
  std::vector<MCParticle> generateNewBoundaryParticles(double) override { return {}; }
  bool describeForGpu(GpuBoundaryInfo &info) const override {
    info = GpuBoundaryInfo{};
    for (int f = 0; f < 6; ++f) info.face_kind[f] = STORM_FACE_ESCAPE;
    return true;
  }
  double getEscapedEnergy() const { return escapedEnergy_; }
  void resetEscapedEnergy() { escapedEnergy_ = 0.0; }
  void addEscapedEnergy(double e) { escapedEnergy_ += e; }

 private:
  double escapedEnergy_ = 0.0;
};

// Every face reflects (boundary/RigidBoundary.hpp:42-58).
template <typename T, typename Grid>
class RigidBoundary : public BoundaryCondition<T, Grid> {
 public:
  explicit RigidBoundary(const Grid &grid) : BoundaryCondition<T, Grid>(grid) {}
  ParticleStatus apply(Particle<T, Grid> &) override {
    throw StormError("RigidBoundary::apply is evaluated on the GPU (MonteCarloManagerGPU)");
  }
  std::vector<Particle<T, Grid>> generateNewBoundaryParticles(double) override { return {}; }
  DDMCBoundaryFaceBehavior getDDMCBoundaryFaceBehavior(size_t, size_t, size_t) const override {
    return DDMCBoundaryFaceBehavior::ReflectingRigid;
  }
  bool describeForGpu(GpuBoundaryInfo &info) const override {
    info = GpuBoundaryInfo{};
    return true;
  }
};

// ------------------------------------------------------------------ physics
template <typename T, typename Grid>
class MonteCarloPhysics {
 public:
  using MCParticle = Particle<T, Grid>;
  MonteCarloPhysics(const Grid &grid, const std::shared_ptr<BoundaryCondition<T, Grid>> &boundary)
      : grid(grid), boundary(boundary) {}
  virtual ~MonteCarloPhysics() = default;
  virtual void updateGridData(void) {}
  virtual std::vector<MCParticle> preStep(double fullDt) = 0;
  virtual StepResult<T, Grid> step(MCParticle &particle, std::vector<MCParticle> &particlesToAdd) = 0;
  virtual void postStep(const std::vector<MCParticle> &particles, double fullDt) = 0;
  virtual size_t getRandomWalkStepCount() const { return 0; }
  virtual size_t getDDMCStepCount() const { return 0; }
  virtual size_t getDDMCLeakCount() const { return 0; }
  virtual size_t getDDMCCensusCount() const { return 0; }
  virtual size_t getDDMCUpscatterCount() const { return 0; }
  virtual size_t getDDMCFallbackCount() const { return 0; }
  virtual std::string getAccelerationDebugInfo(size_t, double) const { return std::string(); }

 protected:
  const Grid &grid;
  std::shared_ptr<BoundaryCondition<T, Grid>> boundary;
};

// ------------------------------------------------------------------ population control
template <typename T, typename Grid>
class PopulationControl {
 public:
  using MCParticle = Particle<T, Grid>;
  explicit PopulationControl(const Grid &grid) : grid(grid) {}
  virtual ~PopulationControl() = default;
  virtual std::vector<MCParticle> activate(const std::vector<MCParticle> &particles) = 0;

 protected:
  const Grid &grid;
};

// Comb population control: run on the GPU by the manager
// (population/CombPopulationControl.hpp:79-253).
template <typename T, typename Grid>
class CombPopulationControl : public PopulationControl<T, Grid> {
 public:
  CombPopulationControl(const Grid &grid, size_t Nmin = 20, double totalParticlesFactor = 2.0)
      : PopulationControl<T, Grid>(grid), Nmin(Nmin), totalParticlesFactor(totalParticlesFactor) {}
  std::vector<Particle<T, Grid>> activate(const std::vector<Particle<T, Grid>> &) override {
    throw StormError("CombPopulationControl runs on the GPU (MonteCarloManagerGPU)");
  }
  size_t getNmin() const { return Nmin; }
  double getTotalParticlesFactor() const { return totalParticlesFactor; }

 private:
  size_t Nmin;
  double totalParticlesFactor;
};

template <typename T, typename Grid>
class NoPopulationControl : public PopulationControl<T, Grid> {
 public:
  using MCParticle = Particle<T, Grid>;
  explicit NoPopulationControl(const Grid &grid) : PopulationControl<T, Grid>(grid) {}
  std::vector<MCParticle> activate(const std::vector<MCParticle> &particles) override { return particles; }
};

// StratifiedCombPopulationControl (population/CombPopulationControl.hpp:35-72,
// 256-511): the comb by cell and class.  The GPU classifies a packet by its
// photon group; the manager checks the classifier against that on packets
// at the groups' centres and rejects any other.
template <typename T, typename Grid>
class StratifiedCombPopulationControl : public PopulationControl<T, Grid> {
 public:
  using MCParticle = Particle<T, Grid>;
  using Classifier = std::function<size_t(const MCParticle &)>;
  StratifiedCombPopulationControl(const Grid &grid, size_t groupCount, Classifier classifier, size_t Nmin = 20,
                                  double totalParticlesFactor = 2.0, size_t minParticlesPerGroup = 2)
      : PopulationControl<T, Grid>(grid), groupCount(std::max<size_t>(1, groupCount)),
        classifier(std::move(classifier)), Nmin(Nmin), totalParticlesFactor(totalParticlesFactor),
        minParticlesPerGroup(minParticlesPerGroup) {}
  std::vector<MCParticle> activate(const std::vector<MCParticle> &) override {
    throw StormError("StratifiedCombPopulationControl runs on the GPU (MonteCarloManagerGPU)");
  }
  size_t getGroupCount() const { return groupCount; }
  size_t getNmin() const { return Nmin; }
  double getTotalParticlesFactor() const { return totalParticlesFactor; }
  size_t getMinParticlesPerGroup() const { return minParticlesPerGroup; }
  size_t classify(const MCParticle &p) const { return classifier ? classifier(p) : 0; }

 private:
  size_t groupCount;
  Classifier classifier;
  size_t Nmin;
  double totalParticlesFactor;
  size_t minParticlesPerGroup;
};

// ------------------------------------------------------------------ radiation
template <std::size_t NumGroups>
struct RadiationIMCParameters {
  std::size_t newPhotonsPerCell = 0;
  bool withHydro = false;
  bool diffusionPressureGradient = false;
  bool MMC = false;
  bool withMultigroupOpacity = false;
  bool withRandomWalk = false;
  double rwMinCellOpticalDepth = 25.0;
  double rwMinParticleOpticalDepth = 5.0;
  bool withDDMC = false;
  double ddmcMinCellOpticalDepth = 15.0;
  double ddmcMinParticleOpticalDepth = 5.0;
  bool ddmcUseMultigroupPGRW = false;
  bool noHydroFeedback = false;
  bool withEgTimeAvg = false;
  bool withCompton = false;
  bool comptonUseInduced = true;
  bool comptonAllowNZeroFallback = true;
  bool comptonDebugParityCheck = false;
  bool comptonCheckSignedTallies = false;
  bool comptonDiagnostics = false;
  bool comptonAngleDependent = true;
  double comptonSignedTallyTolerance = 1e-10;
  std::size_t comptonMatrixSamples = 200000;
  std::array<double, NumGroups + 1> energyBoundaries{};
  bool energyBoundariesProvided = false;
  struct PostProcessParameters {
    bool enabled = false;
    double sourceDt = 0.0;
    double transportTime = 0.0;
    bool useCellVelocities = true;
    struct PolarizationParameters {
      bool enabled = false;
      int manualScatteringsAfterAcceleration = 0;
      double depolarizationScatterings = 0.0;
      std::string acceleratedClosure;
    } polarization;
  } postProcess;
};

struct RadiationCell {
  double temperature = 0;     // K
  double cv = 0;              // unused by RadiationIMC
  double internalEnergy = 0;  // erg (previous step's value, the negative-energy floor)
  double Erad = 0;            // erg/g
};

struct SimpleExtensives {
  double internal_energy = 0;  // erg
  double mass = 0;             // g
  double Erad = 0;             // erg
};

template <typename PointT, typename GridT, typename CellT, std::size_t NumGroups>
class RadiationOpacityModel {
 public:
  using Point = PointT;
  using Grid = GridT;
  using Cell = CellT;
  using GroupArray = std::array<double, NumGroups>;
  using GroupBoundaries = std::array<double, NumGroups + 1>;
  virtual ~RadiationOpacityModel() = default;
  virtual double CalcPlanckOpacity(const CellT &cell) = 0;
  virtual double CalcAbsorptionOpacity(const CellT &cell, double) { return CalcPlanckOpacity(cell); }
  virtual double CalcScatteringOpacity(const CellT &) { return 0.0; }
  virtual double CalcScatteringOpacity(const CellT &cell, double) { return CalcScatteringOpacity(cell); }
  // Emission and scattering directions.  The GPU samples both isotropically;
  // RadiationIMC rejects at construction a model whose overrides behave
  // otherwise (radiation/RadiationOpacityModel.hpp:50-68).
  virtual PointT getRandomVelocity(const CellT &, std::mt19937_64 &rng, std::uniform_real_distribution<double> &dist) {
    return isotropicVelocity(rng, dist);
  }
  virtual PointT getNewScatterVelocity(const CellT &cell, const PointT &, double, std::mt19937_64 &rng,
                                       std::uniform_real_distribution<double> &dist) {
    return getRandomVelocity(cell, rng, dist);
  }
  virtual bool ComptonIncludedInTransport() const { return false; }
  virtual void reseed(std::uint64_t) {}
  virtual std::size_t findGroup(double frequency, const GroupBoundaries &b) const {
    for (std::size_t g = 0; g < NumGroups; ++g)
      if (frequency < b[g + 1]) return g;
    return NumGroups - 1;
  }
  // Thermal photon energies: uniform over the whole range by default
  // (:88-98), a group's share proportional to its width (:100-115), the
  // cumulative opacity g+1 over G (:117-126), the group centres (:128-136).
  virtual double GetThermalEnergy(const CellT &, double random, const GroupBoundaries &b) const {
    return b[0] + random * (b[NumGroups] - b[0]);
  }
  virtual double SampleThermalEnergyInGroup(const CellT &, std::size_t group, double random,
                                            const GroupBoundaries &b) const {
    return b[group] + random * (b[group + 1] - b[group]);
  }
  virtual GroupArray GetThermalGroupPdf(const CellT &, const GroupBoundaries &b) const {
    GroupArray pdf{};
    const double span = b[NumGroups] - b[0];
    if (!(span > 0.0)) return pdf;
    for (std::size_t g = 0; g < NumGroups; ++g) pdf[g] = (b[g + 1] - b[g]) / span;
    return pdf;
  }
  virtual GroupArray GetCumulativeOpacity(const CellT &, const GroupBoundaries &) const {
    GroupArray c{};
    for (std::size_t g = 0; g < NumGroups; ++g) c[g] = static_cast<double>(g + 1) / static_cast<double>(NumGroups);
    return c;
  }
  virtual GroupArray getEnergyCenters(const GroupBoundaries &b) const {
    GroupArray c{};
    for (std::size_t g = 0; g < NumGroups; ++g) c[g] = 0.5 * (b[g] + b[g + 1]);
    return c;
  }

 protected:
  // phi = 2 pi U, mu = 2 U - 1, speed c (:139-150)
  static PointT isotropicVelocity(std::mt19937_64 &rng, std::uniform_real_distribution<double> &dist) {
    constexpr double kTwoPi = 6.283185307179586476925286766559;
    constexpr double kC = 2.99792458e10;
    const double phi = kTwoPi * dist(rng);
    const double mu = 2.0 * dist(rng) - 1.0;
    const double st = std::sqrt(std::max(0.0, 1.0 - mu * mu));
    return PointT(st * std::cos(phi) * kC, st * std::sin(phi) * kC, mu * kC);
  }
};

// Where a new packet is placed in its cell (utils/RandomInCell.hpp:10-67 and
// radiation/RadiationIMC.hpp:317-329).  The GPU samples uniformly in the
// cell itself, which is what RandomInCellPositionSampler does; these types
// select that behaviour (RadiationIMC accepts no other sampler).
template <typename PointT, typename GridT>
struct RandomInCellPositionSampler {};
template <typename PointT, typename GridT>
struct CellCenterPositionSampler {};

namespace detail {
template <typename T, typename = void>
struct has_tracers : std::false_type {};
template <typename T>
struct has_tracers<T, std::void_t<decltype(std::declval<const T &>().tracers)>> : std::true_type {};
template <typename T, typename = void>
struct has_static_tracer_names : std::false_type {};
template <typename T>
struct has_static_tracer_names<T, std::void_t<decltype(T::tracerNames)>> : std::true_type {};
template <typename T, typename = void>
struct has_member_tracer_names : std::false_type {};
template <typename T>
struct has_member_tracer_names<T, std::void_t<decltype(std::declval<const T &>().tracerNames)>> : std::true_type {};
inline const std::vector<double> &no_tracers() {
  static const std::vector<double> v;
  return v;
}
inline const std::vector<std::string> &no_tracer_names() {
  static const std::vector<std::string> v;
  return v;
}
}  // namespace detail

// The cell accessors RadiationIMC uses for the EOS calls
// (radiation/RadiationIMCTraits.hpp:58-120): a cell's tracers and their
// names if the cell type has them, empty vectors otherwise.
template <typename PointT, typename CellT, typename ExtensivesT, std::size_t NumGroups>
struct DirectRadiationIMCTraits {
  using Point = PointT;
  using Cell = CellT;
  using Extensives = ExtensivesT;
  const std::vector<double> &tracers(const CellT &cell) const {
    if constexpr (detail::has_tracers<CellT>::value) return cell.tracers;
    else return detail::no_tracers();
  }
  const std::vector<std::string> &tracerNames(const CellT &cell) const {
    if constexpr (detail::has_static_tracer_names<CellT>::value) return CellT::tracerNames;
    else if constexpr (detail::has_member_tracer_names<CellT>::value) return cell.tracerNames;
    else return detail::no_tracer_names();
  }
};

namespace detail {
template <typename C, typename = void>
struct has_density : std::false_type {};
template <typename C>
struct has_density<C, std::void_t<decltype(std::declval<const C &>().density)>> : std::true_type {};
template <typename C, typename = void>
struct has_internal_energy_density : std::false_type {};
template <typename C>
struct has_internal_energy_density<C, std::void_t<decltype(std::declval<C &>().internalEnergy)>> : std::true_type {};
template <typename C, typename = void>
struct has_erad : std::false_type {};
template <typename C>
struct has_erad<C, std::void_t<decltype(std::declval<C &>().Erad)>> : std::true_type {};
}  // namespace detail

// What the manager needs from the physics, independent of its template
// arguments: per-cell host tabulation before a step and the results after.
class GpuRadiationPhysics {
 public:
  virtual ~GpuRadiationPhysics() = default;
  virtual void gpuParameters(storm_imc_params &p) const = 0;
  virtual std::uint64_t gpuSeed() const = 0;
  // T, Planck opacity, scattering opacity, dT2cv, internal energy, mass per cell
  virtual void gpuTabulate(double *T, double *sa, double *ss, double *cv, double *e, double *mass) = 0;
  // new internal energy, Erad time average, census Erad (extensive), Fleck, Planck opacity
  virtual void gpuAbsorb(const double *e, const double *eradAvg, const double *censusErad, const double *fleck,
                         const double *planck) = 0;
  virtual void gpuCounters(std::size_t rw, std::size_t ddmc) = 0;
  // multigroup problems: the groups and the opacity's energy law (false: gray)
  virtual bool gpuGroups(std::vector<double> &, double &, double &, bool &) const { return false; }
  // the per-group time-averaged radiation energy of the step, [cell * G + g]
  virtual void gpuGroupErad(const double *) {}
};

template <typename PointT, typename GridT, typename CellT, typename ExtensivesT, typename EOST,
          std::size_t NumGroups,
          typename TraitsT = DirectRadiationIMCTraits<PointT, CellT, ExtensivesT, NumGroups>,
          typename PositionSamplerT = RandomInCellPositionSampler<PointT, GridT>>
class RadiationIMC : public MonteCarloPhysics<PointT, GridT>, public GpuRadiationPhysics {
 public:
  static_assert(std::is_same<PositionSamplerT, RandomInCellPositionSampler<PointT, GridT>>::value,
                "the GPU places new packets uniformly in their cell (RandomInCellPositionSampler)");
  using Base = MonteCarloPhysics<PointT, GridT>;
  using MCParticle = Particle<PointT, GridT>;
  using Parameters = RadiationIMCParameters<NumGroups>;
  using OpacityModel = RadiationOpacityModel<PointT, GridT, CellT, NumGroups>;
  using GroupArray = std::array<double, NumGroups>;
  using GroupBoundaries = std::array<double, NumGroups + 1>;
  using Traits = TraitsT;
  using Sampler = PositionSamplerT;

  // radiation/RadiationIMC.hpp:410-419, arguments in the same order
  RadiationIMC(const GridT &grid, const std::shared_ptr<BoundaryCondition<PointT, GridT>> &boundary,
               std::vector<CellT> &cells, std::vector<ExtensivesT> &extensives, std::shared_ptr<EOST> eos,
               std::shared_ptr<OpacityModel> opacity, Parameters parameters, TraitsT traits = TraitsT(),
               PositionSamplerT positionSampler = PositionSamplerT(), std::uint64_t seed = 42)
      : Base(grid, boundary), cells_(cells), extensives_(extensives), eos_(std::move(eos)), opacity_(std::move(opacity)),
        parameters_(parameters), traits_(std::move(traits)), seed_(seed) {
    (void)positionSampler;
    // constructor checks of radiation/RadiationIMC.hpp:585-602, 689-694
    if (parameters_.newPhotonsPerCell == 0) throw StormError("RadiationIMC requires newPhotonsPerCell > 0");
    if (!eos_) throw StormError("RadiationIMC requires a non-null EOS");
    if (!opacity_) throw StormError("RadiationIMC requires a non-null opacity model");
    if (cells_.size() < grid.GetPointNo()) throw StormError("RadiationIMC cells vector is smaller than the grid cell count");
    if (extensives_.size() < grid.GetPointNo())
      throw StormError("RadiationIMC extensives vector is smaller than the grid cell count");
    if (parameters_.withCompton) throw StormError("RadiationIMC option withCompton is not implemented");
    if (parameters_.postProcess.enabled) throw StormError("RadiationIMC option postProcess.enabled is not implemented");
    if (parameters_.withHydro || parameters_.MMC) throw StormError("RadiationIMC: hydro coupling is not supported on the GPU");
    if (NumGroups >= 2 && !parameters_.withMultigroupOpacity)
      throw StormError("RadiationIMC: NumGroups >= 2 needs withMultigroupOpacity");
    if (NumGroups < 2 && parameters_.withMultigroupOpacity)
      throw StormError("RadiationIMC: withMultigroupOpacity needs NumGroups >= 2");
    rejectDirectionOverrides();
    opacity_->reseed(seed + 1ULL);
    Erad_time_avg_.assign(grid.GetPointNo(), 0.0);
    if constexpr (NumGroups >= 2) probeGroupModel();
  }

  std::vector<MCParticle> preStep(double) override { throw StormError(kGpuOnly); }
  StepResult<PointT, GridT> step(MCParticle &, std::vector<MCParticle> &) override { throw StormError(kGpuOnly); }
  void postStep(const std::vector<MCParticle> &, double) override { throw StormError(kGpuOnly); }

  // generateInitialParticles (radiation/RadiationIMC.hpp:2559-2612), gray:
  // particlesPerCell packets per cell carrying the cell's radiation energy
  // (extensives' Erad, else cells' Erad times mass), uniform in the cell box,
  // isotropic, for the first call of the manager's step.
  std::vector<MCParticle> generateInitialParticles(std::size_t particlesPerCell) {
    std::vector<MCParticle> out;
    if (particlesPerCell == 0) return out;
    const std::size_t N = this->grid.GetPointNo();
    std::mt19937_64 rng(seed_ ^ 0x9E3779B97F4A7C15ull);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (std::size_t i = 0; i < N; ++i) {
      double E = 0.0;
      if constexpr (detail::has_erad<ExtensivesT>::value) E = extensives_[i].Erad;
      else if constexpr (detail::has_erad<CellT>::value) E = cells_[i].Erad * extensives_[i].mass;
      if (!(E > 0.0)) continue;
      // the cell's box: per axis, the extreme face centroids
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (auto f : this->grid.GetCellFaces(i)) {
        const PointT fc = this->grid.FaceCM(f);
        const double v[3] = {fc.x, fc.y, fc.z};
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::min(lo[a], v[a]);
          hi[a] = std::max(hi[a], v[a]);
        }
      }
      for (std::size_t j = 0; j < particlesPerCell; ++j) {
        MCParticle p;
        p.cellIndex = i;
        p.frequency = 0.0;
        // multigroup: an energy from the cell's thermal spectrum (:2604-2605)
        if constexpr (NumGroups >= 2) p.frequency = opacity_->GetThermalEnergy(cells_[i], U(rng), parameters_.energyBoundaries);
        // uniform in the cell (RandomInCellPositionSampler on a box cell)
        p.location = PointT(lo[0] + U(rng) * (hi[0] - lo[0]), lo[1] + U(rng) * (hi[1] - lo[1]),
                            lo[2] + U(rng) * (hi[2] - lo[2]));
        // isotropic (RadiationOpacityModel.hpp:139-150)
        const double phi = 6.283185307179586 * U(rng), mu = 2.0 * U(rng) - 1.0;
        const double st = std::sqrt(std::max(0.0, 1.0 - mu * mu));
        constexpr double c = 2.99792458e10;  // cm/s
        p.velocity = PointT(st * std::cos(phi) * c, st * std::sin(phi) * c, mu * c);
        p.weight = E / static_cast<double>(particlesPerCell);
        p.initialWeight = p.weight;
        out.push_back(p);
      }
    }
    return out;
  }

  const std::vector<double> &getFactorFleck() const { return factorFleck_; }
  const std::vector<double> &getPlanckOpacities() const { return planck_; }
  const std::vector<double> &getEradTimeAvg() const { return Erad_time_avg_; }
  std::vector<double> &getEradTimeAvg() { return Erad_time_avg_; }
  // RadiationIMC.hpp:433-434 (withEgTimeAvg in a multigroup problem)
  const std::vector<GroupArray> &getEgTimeAvg() const { return Eg_time_avg_; }
  std::vector<GroupArray> &getEgTimeAvg() { return Eg_time_avg_; }
  const Parameters &getParameters() const { return parameters_; }
  std::size_t getRandomWalkStepCount() const override { return rw_; }
  std::size_t getDDMCStepCount() const override { return ddmc_; }
  void setNewPhotonsPerCell(std::size_t n) { parameters_.newPhotonsPerCell = n; }

  // ---------------------------------------------------------------- GpuRadiationPhysics
  void gpuParameters(storm_imc_params &p) const override {
    p.new_photons_per_cell = parameters_.newPhotonsPerCell;
    p.with_random_walk = parameters_.withRandomWalk ? 1 : 0;
    p.with_ddmc = parameters_.withDDMC ? 1 : 0;
    p.rw_min_cell_optical_depth = parameters_.rwMinCellOpticalDepth;
    p.rw_min_particle_optical_depth = parameters_.rwMinParticleOpticalDepth;
    p.ddmc_min_cell_optical_depth = parameters_.ddmcMinCellOpticalDepth;
    p.ddmc_min_particle_optical_depth = parameters_.ddmcMinParticleOpticalDepth;
    p.ddmc_multigroup_pgrw = parameters_.ddmcUseMultigroupPGRW ? 1 : 0;
  }
  std::uint64_t gpuSeed() const override { return seed_; }

  void gpuTabulate(double *T, double *sa, double *ss, double *cv, double *e, double *mass) override {
    const std::size_t N = this->grid.GetPointNo();
    for (std::size_t i = 0; i < N; ++i) {
      const CellT &cell = cells_[i];
      T[i] = cell.temperature;
      // gray: the Planck opacity; multigroup: sigma_a(cell, T) at nu_ref, the
      // factor of the energy law the GPU applies per packet
      if constexpr (NumGroups >= 2) sa[i] = opacity_->CalcAbsorptionOpacity(cell, nu_ref_);
      else sa[i] = opacity_->CalcPlanckOpacity(cell);
      ss[i] = opacity_->CalcScatteringOpacity(cell);
      cv[i] = eos_->dT2cv(density(i), cell.temperature, traits_.tracers(cell), traits_.tracerNames(cell));
      e[i] = extensives_[i].internal_energy;
      mass[i] = extensives_[i].mass;
      if constexpr (detail::has_internal_energy_density<CellT>::value) {
        // the negative-energy floor of the step is the last synchronised value
        // (RadiationIMC.hpp:794-797); the engine keeps e as that floor
      }
    }
  }

  void gpuAbsorb(const double *e, const double *eradAvg, const double *censusErad, const double *fleck,
                 const double *planck) override {
    const std::size_t N = this->grid.GetPointNo();
    Erad_time_avg_.assign(eradAvg, eradAvg + N);
    factorFleck_.assign(fleck, fleck + N);
    planck_.assign(planck, planck + N);
    for (std::size_t i = 0; i < N; ++i) {
      // synchronizeMaterialCell (radiation/RadiationIMC.hpp:830-866)
      extensives_[i].internal_energy = e[i];
      CellT &cell = cells_[i];
      const double specific = e[i] / extensives_[i].mass;
      if constexpr (detail::has_internal_energy_density<CellT>::value) cell.internalEnergy = e[i];
      cell.temperature = eos_->de2T(density(i), specific, traits_.tracers(cell), traits_.tracerNames(cell));
      // census radiation energy (RadiationIMC.hpp:1989-2028)
      if constexpr (detail::has_erad<ExtensivesT>::value) extensives_[i].Erad = censusErad[i];
      if constexpr (detail::has_erad<CellT>::value)
        cell.Erad = extensives_[i].mass > 0.0 ? censusErad[i] / extensives_[i].mass : 0.0;
    }
  }
  void gpuCounters(std::size_t rw, std::size_t ddmc) override {
    rw_ = rw;
    ddmc_ = ddmc;
  }
  bool gpuGroups(std::vector<double> &bounds, double &nu_pow, double &nu_ref, bool &eg) const override {
    if constexpr (NumGroups < 2) {
      return false;
    } else {
      bounds.assign(parameters_.energyBoundaries.begin(), parameters_.energyBoundaries.end());
      nu_pow = nu_pow_;
      nu_ref = nu_ref_;
      eg = parameters_.withEgTimeAvg;
      return true;
    }
  }
  void gpuGroupErad(const double *eg) override {
    const std::size_t N = this->grid.GetPointNo();
    Eg_time_avg_.assign(N, GroupArray{});
    for (std::size_t i = 0; i < N; ++i)
      for (std::size_t g = 0; g < NumGroups; ++g) Eg_time_avg_[i][g] = eg[i * NumGroups + g];
  }

 private:
  static constexpr const char *kGpuOnly = "RadiationIMC runs on the GPU: drive it with MonteCarloManagerGPU";

  // The GPU draws emission and scattering directions isotropically, as the
  // opacity model's defaults do (radiation/RadiationOpacityModel.hpp:50-68,
  // 139-150).  A model that overrides getRandomVelocity or
  // getNewScatterVelocity with another law, or that puts Compton scattering
  // in the transport, cannot run here: each override is called beside the
  // base class's own body (a qualified, non-virtual call) from equal random
  // states on a few cells and directions, and any difference in the result
  // or in the random numbers drawn is an error, not a silent change of
  // physics.
  void rejectDirectionOverrides() {
    if (opacity_->ComptonIncludedInTransport())
      throw StormError("RadiationIMC: opacity models with Compton scattering in the transport are not supported on the GPU");
    const std::size_t n = std::min<std::size_t>(this->grid.GetPointNo(), 4);
    auto same = [](const PointT &a, const PointT &b) { return a.x == b.x && a.y == b.y && a.z == b.z; };
    for (std::size_t i = 0; i < n; ++i) {
      for (std::uint64_t trial = 0; trial < 4; ++trial) {
        std::mt19937_64 ra(0x5EEDull + 977 * i + trial), rb(0x5EEDull + 977 * i + trial);
        std::uniform_real_distribution<double> da(0.0, 1.0), db(0.0, 1.0);
        const PointT got = opacity_->getRandomVelocity(cells_[i], ra, da);
        const PointT want = opacity_->OpacityModel::getRandomVelocity(cells_[i], rb, db);
        if (!same(got, want) || ra() != rb())
          throw StormError("RadiationIMC: the opacity model's getRandomVelocity is not isotropic; the GPU engine "
                           "samples emission directions isotropically");
        const PointT old(0.6 * 2.99792458e10, 0.0, 0.8 * 2.99792458e10);
        const double nu = 1e-9 * static_cast<double>(trial + 1);
        const PointT gs = opacity_->getNewScatterVelocity(cells_[i], old, nu, ra, da);
        const PointT ws = opacity_->OpacityModel::getNewScatterVelocity(cells_[i], old, nu, rb, db);
        if (!same(gs, ws) || ra() != rb())
          throw StormError("RadiationIMC: the opacity model's getNewScatterVelocity is not the isotropic default; "
                           "the GPU engine scatters isotropically");
      }
    }
  }

  // The GPU's multigroup opacity is sigma_a(cell, T) (max(E, b_0) / nu_ref)^-p
  // (storm_group_desc, storm_problem.h), scattering gray.  The model is
  // checked to be of that form on its first cells -- three energies across
  // the groups give p, the middle one must follow it to 1e-9, an energy
  // below b_0 must give b_0's opacity, scattering must not depend on the
  // energy -- and sigma_a(cell, T) is tabulated at nu_ref every step.
  void probeGroupModel() {
    const auto &b = parameters_.energyBoundaries;
    for (std::size_t g = 0; g < NumGroups; ++g)
      if (!(b[g + 1] > b[g]) || !(b[0] > 0.0))
        throw StormError("RadiationIMC: energyBoundaries must be positive and increasing");
    const double E1 = 0.5 * (b[0] + b[1]), E3 = 0.5 * (b[NumGroups - 1] + b[NumGroups]), E2 = std::sqrt(E1 * E3);
    nu_ref_ = E2;
    const std::size_t n = std::min<std::size_t>(this->grid.GetPointNo(), 8);
    for (std::size_t i = 0; i < n; ++i) {
      const CellT &cell = cells_[i];
      const double s1 = opacity_->CalcAbsorptionOpacity(cell, E1), s2 = opacity_->CalcAbsorptionOpacity(cell, E2),
                   s3 = opacity_->CalcAbsorptionOpacity(cell, E3);
      if (!(s1 > 0.0 && s2 > 0.0 && s3 > 0.0) || !std::isfinite(s1) || !std::isfinite(s3))
        throw StormError("RadiationIMC: the multigroup opacity must be positive and finite");
      const double p = -std::log(s3 / s1) / std::log(E3 / E1);
      if (i == 0) nu_pow_ = p;
      const bool law = std::abs(p - nu_pow_) <= 1e-9 * std::max(1.0, std::abs(p)) &&
                       std::abs(s2 - s1 * std::pow(E2 / E1, -p)) <= 1e-9 * s2;
      const double below = opacity_->CalcAbsorptionOpacity(cell, 0.5 * b[0]), at = opacity_->CalcAbsorptionOpacity(cell, b[0]);
      const double sc1 = opacity_->CalcScatteringOpacity(cell, E1), sc3 = opacity_->CalcScatteringOpacity(cell, E3);
      if (!law || std::abs(below - at) > 1e-9 * at || std::abs(sc1 - sc3) > 1e-12 * std::max(std::abs(sc1), 1e-300))
        throw StormError("RadiationIMC: the GPU engine's multigroup opacity is sigma_a(cell, T) * E^-p for E above the "
                         "lowest group bound (constant below it) with gray scattering; this opacity model is not");
    }
  }

  double density(std::size_t i) const {
    if constexpr (detail::has_density<CellT>::value) return cells_[i].density;
    else return extensives_[i].mass / this->grid.GetVolume(i);
  }

  std::vector<CellT> &cells_;
  std::vector<ExtensivesT> &extensives_;
  std::shared_ptr<EOST> eos_;
  std::shared_ptr<OpacityModel> opacity_;
  Parameters parameters_;
  TraitsT traits_;
  std::uint64_t seed_;
  std::vector<double> Erad_time_avg_, factorFleck_, planck_;
  std::vector<GroupArray> Eg_time_avg_;
  double nu_pow_ = 0.0, nu_ref_ = 1.0;
  std::size_t rw_ = 0, ddmc_ = 0;
};

}  // namespace STORM
