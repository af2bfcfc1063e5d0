// Wavefront-level helpers for gfx950 (64 lanes): aggregated queue appends
// and deterministic reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace storm {
namespace gpu {

// Rank of this lane among the lanes whose bit is set in `mask` below it.
__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Append: every active lane calls this; lanes with pred get consecutive slots
// from `counter` with one atomic per wavefront.  Returns the slot or ~0u.
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t *counter) {
  const uint64_t exec = __builtin_amdgcn_ballot_w64(true);
  const uint64_t want = __builtin_amdgcn_ballot_w64(pred);
  const uint32_t total = static_cast<uint32_t>(__popcll(want));
  uint32_t base = 0u;
  if (lane_rank(exec) == 0u && total != 0u) base = atomicAdd(counter, total);
  base = __builtin_amdgcn_readfirstlane(base);
  return pred ? base + lane_rank(want) : ~0u;
}

// Segmented wavefront sum: adds `v` over each run of consecutive lanes with
// equal `key` (inclusive scan restarted at each run head, 6 shuffle steps).
// All 64 lanes must be active.  Returns true on the last lane of a run,
// whose `v` then holds the run total: one atomic per run instead of one per
// lane when the keys come sorted (packets of one cell side by side).
template <typename T>
__device__ __forceinline__ bool wave_run_sum(uint32_t key, T &v) {
  const uint32_t lane = __lane_id();
  const uint32_t prev = __shfl_up(key, 1u, 64);
  const uint32_t next = __shfl_down(key, 1u, 64);
  const uint64_t heads = __builtin_amdgcn_ballot_w64(lane == 0u || prev != key);
  const uint64_t upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
  const uint32_t start = 63u - static_cast<uint32_t>(__clzll(static_cast<long long>(heads & upto)));
#pragma unroll
  for (uint32_t off = 1; off < 64u; off <<= 1) {
    const T o = __shfl_up(v, off, 64);
    if (lane >= start + off) v += o;
  }
  return lane == 63u || next != key;
}

// Maximum over the wavefront's lanes, one atomic per wavefront.
__device__ __forceinline__ void wave_max_u32(uint32_t v, unsigned long long *dst) {
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  if (v && lane_rank(__builtin_amdgcn_ballot_w64(true)) == 0u) atomicMax(dst, static_cast<unsigned long long>(v));
}

// Integer sums are order independent: per-block LDS reduction, one atomic.
template <int BLOCK>
__device__ __forceinline__ void block_atomic_add_u64(uint64_t v, unsigned long long *dst) {
  __shared__ uint64_t sh[BLOCK];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && sh[0] != 0ull) atomicAdd(dst, static_cast<unsigned long long>(sh[0]));
  __syncthreads();
}

// Deterministic double sum: fixed grid, fixed per-thread stride order and a
// fixed LDS tree; a second single-block launch adds the block partials.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sum_partials(const double *__restrict__ in, uint64_t n,
                                                        double *__restrict__ partials) {
  __shared__ double sh[BLOCK];
  double acc = 0.0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += stride) acc += in[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

}  // namespace gpu
}  // namespace storm
