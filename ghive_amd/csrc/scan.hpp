// Block- and wave-level prefix sums for wave64 workgroups.
#pragma once

#include "device.hpp"

namespace ghive {

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t o = __shfl_up(v, d, kWave);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// Exclusive scan across a workgroup of blockDim.x threads (a multiple of 64, at most 1024).
// lds must hold blockDim.x/64 uint32. Returns this thread's exclusive prefix; total = sum.
__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t *lds, uint32_t &total) {
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
  uint32_t inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) lds[wave] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t w = 0; w < nw; ++w) {
    uint32_t s = lds[w];
    if (w < wave) before += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return before + inc - v;
}

}  // namespace ghive
