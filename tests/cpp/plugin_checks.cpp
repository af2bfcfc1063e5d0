// Plugin-surface checks of the C++ drop-in that need no GPU: RadiationIMC
// takes the reference's constructor arguments (radiation/RadiationIMC.hpp:410-419:
// parameters, traits, position sampler, seed), calls the EOS with the
// traits' tracers, and rejects at construction an opacity model whose
// direction overrides the GPU cannot honour
// (radiation/RadiationOpacityModel.hpp:50-68); a multigroup RadiationIMC
// accepts an opacity model of the engine's family sigma(T) E^-p (p found
// from the model) and rejects others.
// Prints one line per case, "<case> ok" or "<case> FAIL: ...", and exits 0
// only if every case behaves as expected.
#include <algorithm>
#include <cmath>
#include <cstdio>
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
};
using Grid = STORM::CartesianGrid<Vec3>;
using Cell = STORM::RadiationCell;

// An EOS that records the tracer vectors it was given.
struct TracerEOS {
  mutable std::size_t calls = 0;
  double dT2cv(double, double, const std::vector<double> &tr, const std::vector<std::string> &names) const {
    ++calls;
    return tr.empty() && names.empty() ? 1.0 : -1.0;
  }
  double de2T(double, double e, const std::vector<double> &, const std::vector<std::string> &) const { return e; }
};

struct Isotropic : STORM::RadiationOpacityModel<Vec3, Grid, Cell, 1> {
  double CalcPlanckOpacity(const Cell &) override { return 1.0; }
};
// Overrides with the default law (same draws, same result): accepted.
struct IsotropicOverride : Isotropic {
  Vec3 getNewScatterVelocity(const Cell &c, const Vec3 &, double, std::mt19937_64 &rng,
                             std::uniform_real_distribution<double> &dist) override {
    return getRandomVelocity(c, rng, dist);
  }
};
// Forward-peaked scattering, written without `override`: must be rejected.
struct ForwardPeaked : Isotropic {
  Vec3 getNewScatterVelocity(const Cell &, const Vec3 &old, double, std::mt19937_64 &,
                             std::uniform_real_distribution<double> &) {
    return old;
  }
};
// An emission law that is not isotropic: must be rejected.
struct Beamed : Isotropic {
  Vec3 getRandomVelocity(const Cell &, std::mt19937_64 &rng, std::uniform_real_distribution<double> &dist) override {
    (void)dist(rng);
    (void)dist(rng);
    return Vec3(2.99792458e10, 0.0, 0.0);
  }
};
struct WithCompton : Isotropic {
  bool ComptonIncludedInTransport() const override { return true; }
};

using IMC = STORM::RadiationIMC<Vec3, Grid, Cell, STORM::SimpleExtensives, TracerEOS, 1>;

template <typename Opacity>
std::string construct() {
  std::vector<double> xe = {0.0, 0.5, 1.0};
  Grid grid(xe, 0.0, 0.5, 1, 0.0, 0.5, 1);
  std::vector<Cell> cells(2);
  std::vector<STORM::SimpleExtensives> ext(2);
  for (auto &e : ext) e.mass = 1.0;
  STORM::RadiationIMCParameters<1> p;
  p.newPhotonsPerCell = 1;
  auto bc = std::make_shared<STORM::VacuumBoundary<Vec3, Grid>>(grid);
  auto eos = std::make_shared<TracerEOS>();
  try {
    IMC imc(grid, bc, cells, ext, eos, std::make_shared<Opacity>(), p,
            STORM::DirectRadiationIMCTraits<Vec3, Cell, STORM::SimpleExtensives, 1>(),
            STORM::RandomInCellPositionSampler<Vec3, Grid>(), 7);
    double T[2], sa[2], ss[2], cv[2], e[2], m[2];
    imc.gpuTabulate(T, sa, ss, cv, e, m);
    if (imc.gpuSeed() != 7) return "seed not taken from the tenth argument";
    if (cv[0] != 1.0 || eos->calls != 2) return "EOS not called with the traits' (empty) tracers";
    return "";
  } catch (const STORM::StormError &err) {
    return std::string("StormError: ") + err.what();
  }
}

// Multigroup: 4 groups over 1e-3..10 keV (erg), two cells.
constexpr std::size_t kG = 4;
constexpr double kKev = 1.602176634e-9;
struct PowerLaw : STORM::RadiationOpacityModel<Vec3, Grid, Cell, kG> {
  double p = 3.0;
  double CalcPlanckOpacity(const Cell &) override { return 1.0; }
  double CalcAbsorptionOpacity(const Cell &c, double E) override {
    const double e = std::max(E, 1e-3 * kKev) / kKev;
    return (1.0 + c.temperature) * std::pow(e, -p);
  }
};
struct PowerLaw2 : PowerLaw {
  PowerLaw2() { p = 2.0; }
};
// E^-3 below 1 keV, E^-2 above: not one power law
struct BrokenLaw : PowerLaw {
  double CalcAbsorptionOpacity(const Cell &, double E) override {
    const double e = std::max(E, 1e-3 * kKev) / kKev;
    return e < 1.0 ? std::pow(e, -3.0) : std::pow(e, -2.0);
  }
};
// energy-dependent scattering
struct ColouredScatter : PowerLaw {
  double CalcScatteringOpacity(const Cell &, double E) override { return E / kKev; }
};
using IMCG = STORM::RadiationIMC<Vec3, Grid, Cell, STORM::SimpleExtensives, TracerEOS, kG>;

template <typename Opacity>
std::string construct_groups(double want_p) {
  std::vector<double> xe = {0.0, 0.5, 1.0};
  Grid grid(xe, 0.0, 0.5, 1, 0.0, 0.5, 1);
  std::vector<Cell> cells(2);
  cells[0].temperature = 1.0;
  cells[1].temperature = 3.0;
  std::vector<STORM::SimpleExtensives> ext(2);
  for (auto &e : ext) e.mass = 1.0;
  STORM::RadiationIMCParameters<kG> p;
  p.newPhotonsPerCell = 1;
  p.withMultigroupOpacity = true;
  for (std::size_t g = 0; g <= kG; ++g) p.energyBoundaries[g] = kKev * 1e-3 * std::pow(1e4, static_cast<double>(g) / kG);
  auto bc = std::make_shared<STORM::VacuumBoundary<Vec3, Grid>>(grid);
  try {
    IMCG imc(grid, bc, cells, ext, std::make_shared<TracerEOS>(), std::make_shared<Opacity>(), p);
    std::vector<double> b;
    double pow_ = 0.0, ref = 0.0;
    bool eg = true;
    if (!imc.gpuGroups(b, pow_, ref, eg) || b.size() != kG + 1 || eg) return "groups not handed to the manager";
    if (std::abs(pow_ - want_p) > 1e-9) return "wrong exponent " + std::to_string(pow_);
    double T[2], sa[2], ss[2], cv[2], e[2], m[2];
    imc.gpuTabulate(T, sa, ss, cv, e, m);
    // sigma(T) at nu_ref: (1 + T) (nu_ref / keV)^-p
    if (std::abs(sa[1] - 4.0 * std::pow(ref / kKev, -want_p)) > 1e-12 * sa[1]) return "sigma(T) not tabulated at nu_ref";
    return "";
  } catch (const STORM::StormError &err) {
    return std::string("StormError: ") + err.what();
  }
}

}  // namespace

int main() {
  int bad = 0;
  auto expect = [&](const char *name, const std::string &got, bool want_error) {
    const bool is_error = got.rfind("StormError", 0) == 0;
    const bool ok = want_error ? is_error : got.empty();
    std::printf("%s %s%s\n", name, ok ? "ok" : "FAIL: ", ok ? "" : (got.empty() ? "accepted" : got.c_str()));
    bad += ok ? 0 : 1;
  };
  expect("isotropic", construct<Isotropic>(), false);
  expect("isotropic_override", construct<IsotropicOverride>(), false);
  expect("forward_peaked_scatter", construct<ForwardPeaked>(), true);
  expect("beamed_emission", construct<Beamed>(), true);
  expect("compton_in_transport", construct<WithCompton>(), true);
  expect("groups_power_law_3", construct_groups<PowerLaw>(3.0), false);
  expect("groups_power_law_2", construct_groups<PowerLaw2>(2.0), false);
  expect("groups_broken_law", construct_groups<BrokenLaw>(0.0), true);
  expect("groups_coloured_scatter", construct_groups<ColouredScatter>(3.0), true);
  return bad ? 1 : 0;
}
