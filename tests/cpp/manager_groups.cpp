// Multigroup user code written against the reference's manager interface,
// driving the gfx950 engine through include/storm_amd/MonteCarloManagerGPU.hpp:
// the Densmore step slab of examples/densmore2012 in 30 groups (1e-4 to 1e2
// keV, log-spaced) on nx cells over [0, 3] cm, sigma(E) = sigma0 /
// (sqrt(kT) E^3) with sigma0 = 10 keV^3.5/cm below x = 2 cm and 1000 above,
// cv = 1e15 erg/(cm^3 keV), T0 = 0.3 keV, a 1 keV SideTemperature source on
// x- (T^4 taken every step), withEgTimeAvg.
//
// usage: manager_groups <precision 0|1> <steps> <seed0> <nseeds> <nx>
//                       <boundary packets> <photons per cell> <comb Nmin>
//                       <comb factor> <stratified 0|1>
// Prints one JSON object: per seed the cell temperatures (K), Erad and
// census Erad, and the largest relative gap between the per-group tally
// summed over the groups and Erad.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "storm_amd/CartesianGrid.hpp"
#include "storm_amd/MonteCarloManagerGPU.hpp"

namespace {

struct Vec3 {
  using coord_type = double;
  double x = 0, y = 0, z = 0;
  Vec3() = default;
  explicit Vec3(double v) : x(v), y(v), z(v) {}
  Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
  friend std::ostream &operator<<(std::ostream &os, const Vec3 &v) {
    return os << "(" << v.x << ", " << v.y << ", " << v.z << ")";
  }
};

using Grid = STORM::CartesianGrid<Vec3>;
constexpr std::size_t G = 30;
constexpr double kKev = 1.602176634e-9, kBoltz = 1.380649e-16, kKevK = kKev / kBoltz;
constexpr double kCvVol = 1e15 / kKevK;  // erg / (cm^3 K)

struct ConstantCvEOS {
  double dT2cv(double density, double, const std::vector<double> &, const std::vector<std::string> &) const {
    return kCvVol / density;
  }
  double de2T(double density, double specific, const std::vector<double> &, const std::vector<std::string> &) const {
    return specific * density / kCvVol;
  }
};

// sigma_a(cell, E) = sigma0 / (sqrt(k max(T, 1 K)) max(E, E_0)^3): the
// step's two materials by cell index, gray zero scattering.
class StepOpacity : public STORM::RadiationOpacityModel<Vec3, Grid, STORM::RadiationCell, G> {
 public:
  StepOpacity(const std::vector<STORM::RadiationCell> &cells, std::vector<double> sigma0, double e0)
      : cells_(cells), sigma0_(std::move(sigma0)), e0_(e0) {}
  double CalcAbsorptionOpacity(const STORM::RadiationCell &cell, double E) override {
    const double kT = kBoltz * std::max(cell.temperature, 1.0);
    const double e = std::max(E, e0_);
    return sigma0_[static_cast<std::size_t>(&cell - cells_.data())] / (std::sqrt(kT) * e * e * e);
  }
  // the gray mean at the Planck peak; the engine forms its own Planck means
  double CalcPlanckOpacity(const STORM::RadiationCell &cell) override {
    return CalcAbsorptionOpacity(cell, 2.821 * kBoltz * std::max(cell.temperature, 1.0));
  }

 private:
  const std::vector<STORM::RadiationCell> &cells_;
  std::vector<double> sigma0_;
  double e0_;
};

using IMC = STORM::RadiationIMC<Vec3, Grid, STORM::RadiationCell, STORM::SimpleExtensives, ConstantCvEOS, G>;

void print_array(const char *name, const std::vector<double> &a) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < a.size(); ++i) std::printf(i ? ", %.17g" : "%.17g", a[i]);
  std::printf("], ");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 11) {
    std::fprintf(stderr, "usage: %s precision steps seed0 nseeds nx npc nphot nmin factor stratified\n", argv[0]);
    return 2;
  }
  const int precision = std::atoi(argv[1]);
  const int steps = std::atoi(argv[2]);
  const long seed0 = std::atol(argv[3]);
  const int nseeds = std::atoi(argv[4]);
  const size_t nx = static_cast<size_t>(std::atol(argv[5]));
  const size_t npc = static_cast<size_t>(std::atol(argv[6]));
  const size_t nphot = static_cast<size_t>(std::atol(argv[7]));
  const size_t nmin = static_cast<size_t>(std::atol(argv[8]));
  const double factor = std::atof(argv[9]);
  const bool stratified = std::atoi(argv[10]) != 0;
  const double L = 3.0, h = L / static_cast<double>(nx), dt = 5e-12;
  std::array<double, G + 1> bounds{};
  for (size_t g = 0; g <= G; ++g) bounds[g] = kKev * 1e-4 * std::pow(1e6, static_cast<double>(g) / G);
  bounds[G] = kKev * 1e2;
  try {
    std::printf("{\"runs\": [");
    for (int s = 0; s < nseeds; ++s) {
      std::vector<double> xe(nx + 1);
      for (size_t i = 0; i <= nx; ++i) xe[i] = static_cast<double>(i) * h;
      Grid grid(xe, 0.0, h, 1, 0.0, h, 1);
      const size_t N = grid.GetPointNo();
      std::vector<STORM::RadiationCell> cells(N);
      std::vector<STORM::SimpleExtensives> ext(N);
      std::vector<double> sigma0(N);
      for (size_t c = 0; c < N; ++c) {
        cells[c].temperature = 0.3 * kKevK;
        ext[c].mass = grid.GetVolume(c);
        ext[c].internal_energy = kCvVol * cells[c].temperature * grid.GetVolume(c);
        sigma0[c] = (grid.GetCellCM(c).x < 2.0 ? 10.0 : 1000.0) * std::pow(kKev, 3.5);
      }
      STORM::RadiationIMCParameters<G> params;
      params.newPhotonsPerCell = nphot;
      params.withMultigroupOpacity = true;
      params.withEgTimeAvg = true;
      params.energyBoundaries = bounds;
      params.energyBoundariesProvided = true;
      auto boundary = std::make_shared<STORM::SideTemperature<Vec3, Grid>>(
          grid, kKevK, npc, std::vector<double>(bounds.begin(), bounds.end()), false);
      auto physics = std::make_shared<IMC>(grid, boundary, cells, ext, std::make_shared<ConstantCvEOS>(),
                                           std::make_shared<StepOpacity>(cells, sigma0, bounds[0]), params,
                                           IMC::Traits(), IMC::Sampler(), static_cast<std::uint64_t>(seed0 + s));
      std::shared_ptr<STORM::PopulationControl<Vec3, Grid>> pc;
      if (stratified) {
        auto classify = [bounds](const STORM::Particle<Vec3, Grid> &p) {
          for (size_t g = 0; g < G; ++g)
            if (p.frequency < bounds[g + 1]) return g;
          return G - 1;
        };
        pc = std::make_shared<STORM::StratifiedCombPopulationControl<Vec3, Grid>>(grid, G, classify, nmin, factor, 2);
      } else {
        pc = std::make_shared<STORM::CombPopulationControl<Vec3, Grid>>(grid, nmin, factor);
      }
      STORM::MonteCarloManagerSerial<Vec3, Grid> manager(
          grid, physics, pc, boundary, precision == 1 ? STORM::GpuPrecision::FP64 : STORM::GpuPrecision::Mixed);
      std::vector<STORM::Particle<Vec3, Grid>> census;
      for (int k = 0; k < steps; ++k) census = manager.step(std::move(census), dt);
      std::vector<double> T(N), cen(N, 0.0);
      for (size_t c = 0; c < N; ++c) T[c] = cells[c].temperature;
      for (const auto &p : census) cen[p.cellIndex] += p.weight;
      double gap = 0.0, emax = 0.0;
      const auto &erad = physics->getEradTimeAvg();
      for (size_t c = 0; c < N; ++c) emax = std::max(emax, erad[c]);
      for (size_t c = 0; c < N; ++c) {
        double sum = 0.0;
        for (double v : physics->getEgTimeAvg()[c]) sum += v;
        gap = std::max(gap, std::abs(sum - erad[c]) / emax);
      }
      std::printf(s ? ", {" : "{");
      print_array("T", T);
      print_array("erad", erad);
      print_array("census", cen);
      std::printf("\"eg_gap\": %.3g, \"census_packets\": %zu}", gap, census.size());
    }
    std::printf("]}\n");
  } catch (const STORM::StormError &e) {
    std::fprintf(stderr, "StormError: %s\n", e.what());
    return 1;
  }
  return 0;
}
