// Physical constants in CGS (temperature in K), CODATA 2018.
//
// The reference takes these from the external `units` library
// (radiation/RadiationIMC.hpp:25); that library is not vendored, so the values
// are restated here from CODATA 2018.  kev_kelvin = keV / k_B reproduces the
// 1e-3 keV tail (1.160451812155e4 K) of examples/marshak_wave_1/reference.txt.
#pragma once

namespace storm {
namespace units {

constexpr double clight = 2.99792458e10;            // cm / s (exact)
constexpr double k_boltz = 1.380649e-16;            // erg / K (exact)
constexpr double kev = 1.602176634e-9;              // erg (exact)
constexpr double kev_kelvin = kev / k_boltz;        // K per keV = 1.160451812e7
constexpr double sigma_sb = 5.670374419e-5;         // erg cm^-2 s^-1 K^-4
constexpr double arad = 4.0 * sigma_sb / clight;    // erg cm^-3 K^-4
constexpr double inv_clight2 = 1.0 / (clight * clight);

}  // namespace units
}  // namespace storm
