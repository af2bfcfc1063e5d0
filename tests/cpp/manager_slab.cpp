// User code written against the reference's manager interface, driving the
// gfx950 engine through include/storm_amd/MonteCarloManagerGPU.hpp: a gray
// slab nx x 1 x 1 over [0, 1] cm, sigma_a = 1 / cm, constant cv, a 1 keV
// SideTemperature on x-, comb population control (the set-up of
// examples/serial_cartesian/main.cpp with the slab of the parity tests).
//
// usage: manager_slab <precision 0|1> <resident 0|1> <steps> <seed0> <nseeds>
//                     <nx> <boundary packets> <photons per cell> <comb Nmin>
//                     [<x+ temperature, keV>: TwoSidesTemperature instead |
//                      "user": a boundary of this program that makes its own
//                      packets, the way the reference's Marshak example does]
// Prints one JSON object: per seed the cell temperatures (K), time-averaged
// Erad (erg/cm^3) and census Erad (erg), plus the manager's counters.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
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

constexpr double kKevK = 1.602176634e-9 / 1.380649e-16;
constexpr double kCvVol = 1e15 / kKevK;  // erg / (cm^3 K)

// e = cv T per volume, density 1 g/cm^3
struct ConstantCvEOS {
  double dT2cv(double density, double, const std::vector<double> &, const std::vector<std::string> &) const {
    return kCvVol / density;
  }
  double de2T(double density, double specific, const std::vector<double> &, const std::vector<std::string> &) const {
    return specific * density / kCvVol;
  }
};

struct UnitOpacity : STORM::RadiationOpacityModel<Vec3, Grid, STORM::RadiationCell, 1> {
  double CalcPlanckOpacity(const STORM::RadiationCell &) override { return 1.0; }
};

using IMC = STORM::RadiationIMC<Vec3, Grid, STORM::RadiationCell, STORM::SimpleExtensives, ConstantCvEOS, 1>;

// A user boundary that makes its own packets, as the reference's Marshak
// example does (examples/marshak_wave/MarshakBoundary.hpp:87-130): on every
// cell face of the x- side of the box, Npercell packets of energy
// sigma_SB T^4 A dt / Npercell each, at a uniform point of the face, mu =
// sqrt(U) into the box, azimuth 2 pi U, timeLeft = dt U.  Packets reaching
// x- leave; the other faces reflect.
class FaceSourceBoundary : public STORM::BoundaryCondition<Vec3, Grid> {
 public:
  FaceSourceBoundary(const Grid &grid, double T, size_t Npercell, std::uint64_t seed)
      : STORM::BoundaryCondition<Vec3, Grid>(grid), T_(T), n_(Npercell), re_(seed) {}

  STORM::ParticleStatus apply(STORM::Particle<Vec3, Grid> &p) override {
    const auto box = this->grid.GetBoxCoordinates();
    if (p.location.x <= box.first.x) return STORM::ParticleStatus::REMOVE;
    return STORM::ParticleStatus::REFLECT;
  }

  std::vector<STORM::Particle<Vec3, Grid>> generateNewBoundaryParticles(double fullDt) override {
    constexpr double kSigmaSB = 5.670374419e-5, kC = 2.99792458e10, kPi = 3.14159265358979323846;
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<STORM::Particle<Vec3, Grid>> out;
    const double T4 = T_ * T_ * T_ * T_;
    const size_t N = this->grid.GetPointNo();
    for (size_t i = 0; i < N; ++i) {
      const Vec3 c = this->grid.GetMeshPoint(i);
      for (size_t f : this->grid.GetCellFaces(i)) {
        const auto nb = this->grid.GetFaceNeighbors(f);
        const size_t other = nb.first == i ? nb.second : nb.first;
        if (!this->grid.IsPointOutsideBox(other)) continue;
        const Vec3 g = this->grid.GetMeshPoint(other);
        const double dx = g.x - c.x, dy = g.y - c.y, dz = g.z - c.z;
        if (dx / std::sqrt(dx * dx + dy * dy + dz * dz) > -0.99) continue;  // not the x- side
        const auto corners = this->grid.GetPointsInFace(f);
        const auto &v = this->grid.GetFacePoints();
        double ylo = v[corners[0]].y, yhi = ylo, zlo = v[corners[0]].z, zhi = zlo;
        for (size_t q : corners) {
          ylo = std::min(ylo, v[q].y);
          yhi = std::max(yhi, v[q].y);
          zlo = std::min(zlo, v[q].z);
          zhi = std::max(zhi, v[q].z);
        }
        const double E = kSigmaSB * T4 * this->grid.GetArea(f) * fullDt / static_cast<double>(n_);
        for (size_t k = 0; k < n_; ++k) {
          STORM::Particle<Vec3, Grid> p;
          p.location = Vec3(v[corners[0]].x, ylo + U(re_) * (yhi - ylo), zlo + U(re_) * (zhi - zlo));
          const double mu = std::sqrt(U(re_)), st = std::sqrt(1.0 - mu * mu), th = 2.0 * kPi * U(re_);
          p.velocity = Vec3(kC * mu, kC * st * std::cos(th), kC * st * std::sin(th));
          p.frequency = 0.0;
          p.weight = E;
          p.initialWeight = E;
          p.timeLeft = fullDt * U(re_);
          p.cellIndex = i;
          out.push_back(p);
        }
      }
    }
    return out;
  }

 private:
  double T_;
  size_t n_;
  std::mt19937_64 re_;
};

void print_array(const char *name, const std::vector<double> &a, bool last = false) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < a.size(); ++i) std::printf(i ? ", %.17g" : "%.17g", a[i]);
  std::printf("]%s", last ? "" : ", ");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 10 && argc != 11) {
    std::fprintf(stderr, "usage: %s precision resident steps seed0 nseeds nx npc nphot nmin [T_right_keV]\n", argv[0]);
    return 2;
  }
  const bool user_boundary = argc == 11 && std::strcmp(argv[10], "user") == 0;
  const bool two_sides = argc == 11 && !user_boundary;
  const double T_right = two_sides ? std::atof(argv[10]) * kKevK : 0.0;
  const int precision = std::atoi(argv[1]);
  const bool resident = std::atoi(argv[2]) != 0;
  const int steps = std::atoi(argv[3]);
  const long seed0 = std::atol(argv[4]);
  const int nseeds = std::atoi(argv[5]);
  const size_t nx = static_cast<size_t>(std::atol(argv[6]));
  const size_t npc = static_cast<size_t>(std::atol(argv[7]));
  const size_t nphot = static_cast<size_t>(std::atol(argv[8]));
  const size_t nmin = static_cast<size_t>(std::atol(argv[9]));
  const double dt = 1e-11, h = 1.0 / static_cast<double>(nx);
  try {
    std::printf("{\"runs\": [");
    for (int s = 0; s < nseeds; ++s) {
      std::vector<double> xe(nx + 1);
      for (size_t i = 0; i <= nx; ++i) xe[i] = static_cast<double>(i) * h;
      Grid grid(xe, 0.0, h, 1, 0.0, h, 1);
      const size_t N = grid.GetPointNo();
      std::vector<STORM::RadiationCell> cells(N);
      std::vector<STORM::SimpleExtensives> ext(N);
      for (size_t c = 0; c < N; ++c) {
        cells[c].temperature = 1e-3 * kKevK;
        ext[c].mass = grid.GetVolume(c);
        ext[c].internal_energy = kCvVol * cells[c].temperature * grid.GetVolume(c);
      }
      STORM::RadiationIMCParameters<1> params;
      params.newPhotonsPerCell = nphot;
      std::shared_ptr<STORM::BoundaryCondition<Vec3, Grid>> boundary;
      if (user_boundary)
        boundary = std::make_shared<FaceSourceBoundary>(grid, kKevK, npc, static_cast<std::uint64_t>(seed0 + s) * 7919u);
      else if (two_sides)
        boundary = std::make_shared<STORM::TwoSidesTemperature<Vec3, Grid>>(grid, kKevK, T_right, npc);
      else
        boundary = std::make_shared<STORM::SideTemperature<Vec3, Grid>>(grid, kKevK, npc);
      auto physics = std::make_shared<IMC>(grid, boundary, cells, ext, std::make_shared<ConstantCvEOS>(),
                                           std::make_shared<UnitOpacity>(), params, IMC::Traits(),
                                           IMC::Sampler(), static_cast<std::uint64_t>(seed0 + s));
      auto comb = std::make_shared<STORM::CombPopulationControl<Vec3, Grid>>(grid, nmin, 2.0);
      STORM::MonteCarloManagerSerial<Vec3, Grid> manager(
          grid, physics, comb, boundary,
          precision == 1 ? STORM::GpuPrecision::FP64 : STORM::GpuPrecision::Mixed);
      manager.setResidentCensus(resident);
      std::vector<STORM::Particle<Vec3, Grid>> census;
      size_t segments = 0;
      for (int k = 0; k < steps; ++k) {
        census = manager.step(std::move(census), dt);
        for (size_t v : manager.GetCellsStepsCounters()) segments += v;
      }
      std::vector<double> T(N), cen(N, 0.0);
      for (size_t c = 0; c < N; ++c) T[c] = cells[c].temperature;
      if (resident) census = manager.exportCensus();
      for (const auto &p : census) cen[p.cellIndex] += p.weight;
      std::printf(s ? ", {" : "{");
      print_array("T", T);
      print_array("erad", physics->getEradTimeAvg());
      print_array("census", cen);
      std::printf("\"census_packets\": %zu, \"end_particles\": %zu, \"segments\": %zu}", census.size(),
                  manager.GetEndParticleCount(), segments);
    }
    std::printf("]}\n");
  } catch (const STORM::StormError &e) {
    std::fprintf(stderr, "StormError: %s\n", e.what());
    return 1;
  }
  return 0;
}
