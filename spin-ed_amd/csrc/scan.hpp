// scan.hpp -- exclusive scan of 32-bit counts into 64-bit offsets in one
// workgroup of 1024 threads (each thread sums a contiguous slice, then a
// Hillis-Steele scan over the 1024 slice sums).  Used for the basis
// builder's block counts and the gather cache's per-wave term counts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spined {
namespace {

// Exclusive scan of n block counts into offsets (single block of 1024).
__global__ void __launch_bounds__(1024)
    scan_counts_kernel(const uint32_t *__restrict__ counts, uint64_t n,
                       uint64_t *__restrict__ offsets, uint64_t *__restrict__ total) {
  __shared__ uint64_t part[1024];
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t b = threadIdx.x * per;
  const uint64_t e = b + per < n ? b + per : n;
  uint64_t s = 0;
  for (uint64_t i = b; i < e; ++i) s += counts[i];
  part[threadIdx.x] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan over 1024 partial sums
  for (int off = 1; off < 1024; off <<= 1) {
    const uint64_t v = threadIdx.x >= static_cast<unsigned>(off) ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
  for (uint64_t i = b; i < e; ++i) {
    offsets[i] = run;
    run += counts[i];
  }
  if (threadIdx.x == 1023) *total = part[1023];
}

} // namespace
} // namespace spined
