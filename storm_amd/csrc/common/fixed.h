// Exact fixed-point bookkeeping for packet weights and cell tallies.
//
// A step has an energy quantum Q = 2^eq erg chosen so that every energy the
// step can move is below 2^62 Q.  Packet weights are FP64 values that are
// kept integer multiples of Q at every point where they are tallied, so the
// conversion weight -> integer is exact, and any sum of the step's energy
// flows (a cell's deposits, the census, the ledger's totals) fits one u64.
// (Q is 2^-62 of the step's energy: a weight of 1e-7 of it keeps 39 bits,
// far finer than the FP32 tracking.)  A value is written as two limbs,
// value = hi * 2^36 + lo with lo < 2^36; tallies are per-cell pairs of u64
// accumulators, value = hi_sum * 2^36 + lo_sum.  The transport's per-visit
// tallies add whole values (value_of) into the low word, one atomic each; the
// other tallies add limb by limb (no carries), exact as long as a cell
// receives fewer than 2^28 of those additions per step.  Integer sums are
// exact and independent of the order of the additions.
#pragma once

#include <stdint.h>
#include <math.h>
#include "fpmath.h"
#include "hd.h"

namespace storm {
namespace fx {

constexpr int kLoBits = 36;
constexpr uint64_t kLoMask = (1ull << kLoBits) - 1ull;
constexpr int kHeadroomBits = 62;  // energies bounded by 2^62 quanta

struct Limbs {
  uint64_t lo, hi;
};

// Round a non-negative double down to a multiple of 2^e (exact).
STORM_HD double quantize_down(double w, int e) {
  return floor(w * fpm::pow2d(-e)) * fpm::pow2d(e);
}
// Round to the nearest multiple of 2^e (exact; ties to even).
STORM_HD double quantize_nearest(double w, int e) {
  return rint(w * fpm::pow2d(-e)) * fpm::pow2d(e);
}

// v >= 0 rounded to a multiple of 2^e without bias: up with probability equal
// to the fraction left over, taking u24 (24 uniform bits) as the draw.  A
// path integral of a very opaque cell can be far below one quantum; rounding
// it to nearest would drop it every time.
STORM_HD double quantize_unbiased(double v, int e, uint32_t u24) {
  const double y = v * fpm::pow2d(-e);
  const double fl = floor(y);
  const double frac = y - fl;  // exact
  const double u = static_cast<double>(u24 & 0xFFFFFFu) * 0x1p-24;
  return (u < frac ? fl + 1.0 : fl) * fpm::pow2d(e);
}

// The weight a packet carries on after a cell visit (into the next cell, to
// census or out of the problem), as a multiple of 2^e: w rounded without
// bias, the draw a hash of the packet id, its random-stream position and the
// cell (no stream draw, so the history is unchanged), and never above
// w_enter, the weight the visit began with.  Rounding down every time would
// take up to one quantum from a packet at each face it crosses, which ends
// packets of a few quanta long before their physics does.
STORM_HD double carry_round(double w, double w_enter, int e, uint32_t id, uint32_t ctr, uint32_t cell) {
  uint32_t h = id * 0x2545F491u ^ (ctr + 0x3C6EF372u) * 0x9E3779B9u ^ cell * 0x27D4EB2Fu;
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  const double wq = quantize_unbiased(w, e, h >> 8);
  return wq > w_enter ? quantize_down(w, e) : wq;
}

// Weight of each of a cell's emission packets, E / n rounded to a multiple of
// 2^e without bias (the draw a hash of the cell and the step), so that the
// energy a cell emits is E in expectation however few quanta a packet holds.
STORM_HD double emission_round(double per, int e, uint32_t cell, uint32_t step) {
  uint32_t h = cell * 0x9E3779B1u ^ (step + 0x165667B1u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return quantize_unbiased(per, e, h >> 8);
}

// Limbs of w / 2^e for a w that is an integer multiple of 2^e below 2^62.
STORM_HD Limbs limbs_of(double w, int e) {
  double y = w * fpm::pow2d(-e);            // exact integer
  double h = floor(y * fpm::pow2d(-kLoBits));
  double l = y - h * fpm::pow2d(kLoBits);   // exact
  Limbs r;
  r.hi = static_cast<uint64_t>(h);
  r.lo = static_cast<uint64_t>(l);
  return r;
}

// The integer hi * 2^36 + lo (below 2^62 for the step's energies).
STORM_HD uint64_t value_of(Limbs l) { return (l.hi << kLoBits) + l.lo; }

// a - b for a >= b, both normalised (lo < 2^36); result normalised.
STORM_HD Limbs sub(Limbs a, Limbs b) {
  Limbs r;
  if (a.lo >= b.lo) {
    r.lo = a.lo - b.lo;
    r.hi = a.hi - b.hi;
  } else {
    r.lo = a.lo + (1ull << kLoBits) - b.lo;
    r.hi = a.hi - b.hi - 1ull;
  }
  return r;
}

// Value of an accumulator pair (hi, lo sums) as a double times 2^e.
STORM_HD double to_double(uint64_t lo_sum, uint64_t hi_sum, int e) {
  // hi_sum * 2^36 + lo_sum, rounded once per term; the same expression is
  // evaluated on the host and on the device.
  double h = static_cast<double>(hi_sum) * fpm::pow2d(kLoBits);
  double l = static_cast<double>(lo_sum);
  return (h + l) * fpm::pow2d(e);
}

// Exponent of the smallest power of two >= x (x > 0, normal).
STORM_HD int ceil_log2(double x) {
  uint64_t b = fpm::d2u(x);
  int e = static_cast<int>((b >> 52) & 0x7ffull) - 1023;
  if ((b & ((1ull << 52) - 1ull)) != 0ull) e += 1;
  return e;
}

}  // namespace fx
}  // namespace storm
