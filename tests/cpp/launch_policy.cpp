// The transport's launch policy (storm_amd/csrc/launch_policy.h) on made-up
// transport rates, without a GPU: the tuner's candidate sequence and choice,
// the schedule watch's probes, switches, back-off and its cancellation of a
// steady trend, and a fixed launch.  Prints one "ok" line per check; exits 1
// on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "launch_policy.h"

using storm::gpu::LaunchPolicy;

namespace {

void check(bool c, const char *what) {
  if (!c) {
    std::printf("FAIL %s\n", what);
    std::exit(1);
  }
}

struct Shape {
  int32_t block, chunks, schedule;
};

// one step: the policy picks the launch, the model gives its rate
template <typename Model>
Shape run_step(LaunchPolicy &p, int step, Model rate) {
  p.tune_before_step();
  const Shape s{p.block, p.chunks_per_wave(), p.schedule};
  p.tune_after_step(rate(s, step), 1.0);
  return s;
}

double block_f(int32_t b) { return b == 128 ? 0.9 : (b == 256 ? 1.0 : 1.1); }
double chunk_f(int32_t k) { return k == 1 ? 0.8 : (k == 2 ? 0.9 : (k == 4 ? 1.0 : (k == 8 ? 1.2 : 1.1))); }

}  // namespace

int main() {
  // tuner: refill 30% faster, 512 threads and 8 chunks best, on a workload
  // that grows 2% per step
  {
    LaunchPolicy p;
    p.schedule = STORM_GPU_SCHEDULE_REFILL;
    p.tune_start();
    auto rate = [](const Shape &s, int step) {
      return (1.0 + 0.02 * step) * (s.schedule == STORM_GPU_SCHEDULE_REFILL ? 1.3 : 1.0) * block_f(s.block) *
             chunk_f(s.chunks);
    };
    std::vector<Shape> seq;
    for (int k = 0; k < 20; ++k) seq.push_back(run_step(p, k, rate));
    const int sch[4] = {0, 1, 1, 0};
    for (int k = 0; k < 4; ++k) check(seq[k].schedule == sch[k], "tuner: schedules pass, refill, refill, pass");
    const int blk[6] = {128, 256, 512, 512, 256, 128};
    for (int k = 0; k < 6; ++k) check(seq[4 + k].block == blk[k], "tuner: blocks 128..512..128");
    const int chk[10] = {1, 2, 4, 8, 16, 16, 8, 4, 2, 1};
    for (int k = 0; k < 10; ++k) check(seq[10 + k].chunks == chk[k], "tuner: chunks 1..16..1");
    check(p.tune_phase == -1, "tuner: done after 20 steps");
    check(p.schedule == STORM_GPU_SCHEDULE_REFILL && p.block == 512 && p.chunks_per_wave() == 8,
          "tuner: keeps refill, 512 threads, 8 chunks");
    check(p.chunks_of[0] == static_cast<int32_t>(storm::gpu::kChunksPerWave), "tuner: pass keeps its own chunks");
    std::printf("tuner sequence and choice ok\n");
  }
  // watch: pass 25% slower throughout; probes after 16 steps, then the
  // interval doubles to 128
  {
    LaunchPolicy p;
    p.schedule = STORM_GPU_SCHEDULE_REFILL;
    auto rate = [](const Shape &s, int) { return s.schedule == STORM_GPU_SCHEDULE_REFILL ? 1.0 : 0.75; };
    std::vector<int> probes;
    for (int k = 0; k < 420; ++k) {
      const Shape s = run_step(p, k, rate);
      if (s.schedule == STORM_GPU_SCHEDULE_PASS) probes.push_back(k);
    }
    check(p.schedule == STORM_GPU_SCHEDULE_REFILL, "watch: keeps the faster schedule");
    // probe k+1 follows the step after probe k by its interval
    const int want[5] = {16, 16 + 2 + 32, 16 + 2 + 32 + 2 + 64, 16 + 2 + 32 + 2 + 64 + 2 + 128,
                         16 + 2 + 32 + 2 + 64 + 2 + 128 + 2 + 128};
    check(probes.size() == 5, "watch: five probes in 420 steps");
    for (int k = 0; k < 5; ++k) check(probes[k] == want[k], "watch: probes back off 16, 32, 64, 128, 128");
    std::printf("watch back-off ok\n");
  }
  // watch: pass becomes 10% faster from step 40: switch at the first probe
  // after it, and the interval is 16 again
  {
    LaunchPolicy p;
    p.schedule = STORM_GPU_SCHEDULE_REFILL;
    auto rate = [](const Shape &s, int step) {
      const double pass = step < 40 ? 0.85 : 1.10;
      return s.schedule == STORM_GPU_SCHEDULE_REFILL ? 1.0 : pass;
    };
    int switched = -1;
    std::vector<Shape> seq;
    for (int k = 0; k < 120; ++k) {
      seq.push_back(run_step(p, k, rate));
      if (switched < 0 && p.schedule == STORM_GPU_SCHEDULE_PASS && p.watch_state == 0) switched = k;
    }
    // probes at 16 (lost by 15%: next in 32) and 50 (won): switch after step 51
    check(seq[16].schedule == STORM_GPU_SCHEDULE_PASS && seq[17].schedule == STORM_GPU_SCHEDULE_REFILL,
          "switch: first probe at step 16");
    check(seq[50].schedule == STORM_GPU_SCHEDULE_PASS && seq[51].schedule == STORM_GPU_SCHEDULE_REFILL,
          "switch: second probe 32 steps later");
    check(switched == 51, "switch: after the step that follows the winning probe");
    check(seq[52].schedule == STORM_GPU_SCHEDULE_PASS, "switch: runs the faster schedule from then on");
    check(seq[52 + 16].schedule == STORM_GPU_SCHEDULE_REFILL && seq[52 + 17].schedule == STORM_GPU_SCHEDULE_PASS,
          "switch: next probe 16 steps later, of the other schedule");
    check(p.schedule == STORM_GPU_SCHEDULE_PASS, "switch: stays on pass");
    std::printf("watch switch ok\n");
  }
  // watch: equal schedules on a workload growing 3% per step: the probes
  // are compared with the mean of their neighbours, so no switch
  {
    LaunchPolicy p;
    p.schedule = STORM_GPU_SCHEDULE_REFILL;
    auto rate = [](const Shape &, int step) { return std::pow(1.03, step); };
    for (int k = 0; k < 200; ++k) {
      run_step(p, k, rate);
      check(!(p.schedule == STORM_GPU_SCHEDULE_PASS && p.watch_state == 0), "trend: no switch on a growing workload");
    }
    std::printf("watch trend ok\n");
  }
  // fixed launch: the watch is off
  {
    LaunchPolicy p;
    p.schedule = STORM_GPU_SCHEDULE_PASS;
    p.watch = false;
    auto rate = [](const Shape &s, int) { return s.schedule == STORM_GPU_SCHEDULE_REFILL ? 2.0 : 1.0; };
    for (int k = 0; k < 100; ++k) check(run_step(p, k, rate).schedule == STORM_GPU_SCHEDULE_PASS, "fixed: no probes");
    std::printf("fixed launch ok\n");
  }
  return 0;
}
