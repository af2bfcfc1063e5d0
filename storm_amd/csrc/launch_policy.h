// The transport's launch policy, host only: the launch shape (block size,
// schedule, chunks per wavefront of each schedule), the on-line tuner and the
// schedule watch.  A packet's history does not depend on any of them, so the
// policy only ever changes how fast a step runs.  Kept apart from the engine
// so that its decisions can be checked on the CPU with made-up rates.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "storm_gpu.h"

namespace storm {
namespace gpu {

// The pass schedule's 64-packet chunks per wavefront of a block by default
// (a block owns chunks * block packets of the cell-ordered input and one LDS
// tally table for all of them); 1, 2, 4, 8 or 16 by
// storm_gpu_set_chunks_per_wave or the on-line tuner.
constexpr uint32_t kChunksPerWave = 2;

struct LaunchPolicy {
  int32_t block = 256;
  int32_t schedule = STORM_GPU_SCHEDULE_PASS;
  // 64-packet chunks per wavefront of a block's range, per schedule (pass,
  // refill): refill gains from a larger range (its lanes idle only once the
  // range is drained), pass does not (its wavefront waits for its longest
  // history whatever the range)
  int32_t chunks_of[2] = {static_cast<int32_t>(kChunksPerWave), 8};
  int32_t chunks_per_wave() const { return chunks_of[schedule == STORM_GPU_SCHEDULE_REFILL ? 1 : 0]; }
  int32_t &chunks_ref() { return chunks_of[schedule == STORM_GPU_SCHEDULE_REFILL ? 1 : 0]; }

  // On-line tuning of the transport's launch (storm_gpu_set_tuning with
  // block 0).  A packet's history does not depend on the block size, the
  // chunks a block hands out or the schedule, so trying them changes no
  // result.  Three phases, one parameter each with the others at their best
  // so far: the schedule (pass, refill, each with its own chunks), the block
  // size (128, 256, 512), the chunks per wavefront of the chosen schedule (1,
  // 2, 4, 8, 16).  A phase runs its candidates forwards then backwards, one
  // step each (an order that cancels a steady trend of the workload), and
  // keeps the one with the highest mean rate of transport segments per
  // second; 20 steps in all.
  struct LaunchShape {
    int32_t block, chunks, schedule;
  };
  int32_t tune_phase = -1;  // -1: off
  std::vector<LaunchShape> tune_seq;
  std::vector<double> tune_sum;
  size_t tune_pos = 0, tune_ncand = 0;
  void tune_start() {
    tune_phase = 0;
    tune_plan();
  }
  void tune_plan() {
    std::vector<LaunchShape> c;
    if (tune_phase == 0) {
      c.push_back({block, chunks_of[0], STORM_GPU_SCHEDULE_PASS});
      c.push_back({block, chunks_of[1], STORM_GPU_SCHEDULE_REFILL});
    } else if (tune_phase == 1) {
      for (int32_t b : {128, 256, 512}) c.push_back({b, chunks_per_wave(), schedule});
    } else {
      for (int32_t k : {1, 2, 4, 8, 16}) c.push_back({block, k, schedule});
    }
    tune_ncand = c.size();
    tune_seq = c;
    for (size_t k = c.size(); k-- > 0;) tune_seq.push_back(c[k]);
    tune_sum.assign(tune_ncand, 0.0);
    tune_pos = 0;
  }
  void apply(const LaunchShape &t) {
    block = t.block;
    schedule = t.schedule;
    chunks_ref() = t.chunks;
  }

  // The schedule watch.  The workload of a problem changes as it runs (a
  // cold start that the random walk dominates, a heated region that then
  // spreads), and so may the faster schedule (with a refill kernel compiled
  // for 6 waves per SIMD, refill was 24% faster on the 64^3 cube over steps
  // 3-12 and pass 12% faster over steps 13-50; at 5 waves refill leads
  // throughout).  Every watch_every steps one step runs in the other
  // schedule; its rate is compared with the mean of the steps either side of
  // it (which cancels a steady trend) and the engine switches when the other
  // is more than 3% faster.  A probe that loses by more than 10% doubles the
  // interval (16 steps at first, at most 128), a switch resets it: a problem
  // runs hundreds to thousands of steps and its workload drifts over tens,
  // while each probe of a clearly slower schedule costs its margin.  On
  // unless the launch was fixed (storm_gpu_set_tuning with a block size);
  // results do not depend on it.
  static constexpr uint32_t kWatchFirst = 16, kWatchMax = 128;
  bool watch = true;
  int32_t watch_state = 0;  // 0: normal, 1: this step probes, 2: the step after the probe
  uint32_t watch_n = 0, watch_every = kWatchFirst;
  double watch_prev = 0.0, watch_alt = 0.0;
  void watch_reset() {
    watch_state = 0;
    watch_n = 0;
    watch_every = kWatchFirst;
    watch_prev = 0.0;
  }
  static int32_t other(int32_t s) {
    return s == STORM_GPU_SCHEDULE_REFILL ? STORM_GPU_SCHEDULE_PASS : STORM_GPU_SCHEDULE_REFILL;
  }

  void tune_before_step() {
    if (tune_phase >= 0) {
      apply(tune_seq[tune_pos]);
      return;
    }
    if (watch && watch_state == 0 && watch_n >= watch_every && watch_prev > 0.0) {
      watch_state = 1;
      schedule = other(schedule);
    }
  }
  void tune_after_step(double segments, double seconds) {
    const double rate = seconds > 0.0 ? segments / seconds : 0.0;
    if (tune_phase >= 0) {
      const size_t k = tune_pos < tune_ncand ? tune_pos : 2 * tune_ncand - 1 - tune_pos;
      tune_sum[k] += rate;
      if (++tune_pos < tune_seq.size()) return;
      size_t best = 0;
      for (size_t q = 1; q < tune_ncand; ++q)
        if (tune_sum[q] > tune_sum[best]) best = q;
      apply(tune_seq[best]);
      if (++tune_phase > 2) {
        tune_phase = -1;
        watch_reset();
        return;
      }
      tune_plan();
      return;
    }
    if (!watch) return;
    if (watch_state == 0) {
      watch_prev = rate;
      ++watch_n;
    } else if (watch_state == 1) {
      watch_alt = rate;
      schedule = other(schedule);
      watch_state = 2;
    } else {
      const double around = 0.5 * (watch_prev + rate);
      if (watch_alt > 1.03 * around) {
        schedule = other(schedule);
        watch_prev = watch_alt;
        watch_every = kWatchFirst;
      } else {
        watch_prev = rate;
        if (watch_alt < 0.9 * around) watch_every = std::min(2 * watch_every, kWatchMax);
      }
      watch_state = 0;
      watch_n = 0;
    }
  }
};

}  // namespace gpu
}  // namespace storm
