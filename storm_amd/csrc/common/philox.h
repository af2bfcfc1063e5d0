// Counter-based Philox4x32-10 generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), written from the
// published round function.  The same header is compiled for gfx950 and for
// the host, so a packet's random stream -- keyed by (seed, packet id) and
// counted by (step, stream tag, draw index) -- is reproduced bit-exactly by a
// CPU replay, independent of thread scheduling.
//
// Uniforms:
//   u01_f32 : 24 random bits -> float in the open interval (0, 1)
//   u01_f64 : 53 random bits (two words) -> double in the open interval (0, 1)
#pragma once

#include <stdint.h>
#include "hd.h"

namespace storm {

struct Philox4 {
  uint32_t v[4];
};

namespace philox_detail {
constexpr uint32_t kM0 = 0xD2511F53u;
constexpr uint32_t kM1 = 0xCD9E8D57u;
constexpr uint32_t kW0 = 0x9E3779B9u;  // golden ratio
constexpr uint32_t kW1 = 0xBB67AE85u;  // sqrt(3) - 1

STORM_HD void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  uint64_t p = static_cast<uint64_t>(a) * static_cast<uint64_t>(b);
  hi = static_cast<uint32_t>(p >> 32);
  lo = static_cast<uint32_t>(p);
}
}  // namespace philox_detail

// One Philox4x32 block with 10 rounds.
STORM_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                               uint32_t k0, uint32_t k1) {
  using namespace philox_detail;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(kM0, c0, hi0, lo0);
    mulhilo(kM1, c2, hi1, lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += kW0;
    k1 += kW1;
  }
  Philox4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// (0,1) float from the top 24 bits of a word: (k + 0.5) * 2^-24.
STORM_HD float u01_f32(uint32_t x) {
  return (static_cast<float>(x >> 8) + 0.5f) * 5.9604644775390625e-08f;
}

// (0,1) double from 53 bits of two words: (k + 0.5) * 2^-53.
STORM_HD double u01_f64(uint32_t a, uint32_t b) {
  uint64_t k = (static_cast<uint64_t>(a) << 21) ^ static_cast<uint64_t>(b >> 11);
  k &= (1ull << 53) - 1ull;
  return (static_cast<double>(k) + 0.5) * 1.1102230246251565e-16;
}

// Stream tags (third counter word) separating the uses of one packet key.
enum : uint32_t {
  kTagBirth = 1u,       // emission / boundary sourcing draws
  kTagTransport = 2u,   // per-segment transport draws
  kTagComb = 3u,        // comb offsets (keyed by cell)
  kTagCellField = 4u,   // synthetic per-cell random fields (problem setup)
  kTagRandomWalk = 5u,  // random-walk / DDMC draws
};

// A packet's random stream: key = (seed, packet id); counter =
// (draw index, step, tag, instance).
struct PacketStream {
  uint32_t k0, k1;
  uint32_t step, tag, inst;

  STORM_HD static PacketStream make(uint64_t seed, uint32_t packet_id, uint32_t step,
                                    uint32_t tag, uint32_t inst = 0u) {
    PacketStream s;
    s.k0 = packet_id;
    s.k1 = static_cast<uint32_t>(seed) ^ (static_cast<uint32_t>(seed >> 32) * 0x9E3779B9u);
    s.step = step;
    s.tag = tag;
    s.inst = inst;
    return s;
  }

  STORM_HD Philox4 block(uint32_t n) const {
    return philox4x32_10(n, step, tag, inst, k0, k1);
  }
};

}  // namespace storm
