// scan.hpp -- exclusive scan of 32-bit counts into 64-bit offsets.  Up to
// kScanTile counts: one workgroup of 1024 threads (each thread sums a
// contiguous slice, then a Hillis-Steele scan over the 1024 slice sums).
// Longer inputs: tile sums, the one-workgroup scan over them, then each tile
// scanned from its offset.  Used for the basis builder's block counts and
// the gather cache's per-wave term counts (13.5 M of them at D = 8.6e8).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

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

// 256 threads x 16 consecutive counts per tile
constexpr int kScanThreads = 256, kScanPer = 16;
constexpr uint64_t kScanTile = kScanThreads * kScanPer;

__global__ void __launch_bounds__(kScanThreads)
    scan_tile_sums_kernel(const uint32_t *__restrict__ counts, uint64_t n,
                          uint32_t *__restrict__ tile_sums) {
  __shared__ uint64_t red[kScanThreads];
  const uint64_t b = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  uint64_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q)
    if (b + q < n) s += counts[b + q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = kScanThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off)) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  // at most kScanTile counts of at most 2^32 - 1 each would overflow 32 bits;
  // the callers' counts (states per 256-word block, terms per wave) are far
  // below 2^20, so a tile sum stays below 2^32
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = static_cast<uint32_t>(red[0]);
}

__global__ void __launch_bounds__(kScanThreads)
    scan_tile_apply_kernel(const uint32_t *__restrict__ counts, uint64_t n,
                           const uint64_t *__restrict__ tile_offsets,
                           uint64_t *__restrict__ offsets) {
  __shared__ uint64_t part[kScanThreads];
  const uint64_t b = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  uint32_t c[kScanPer];
  uint64_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    c[q] = b + q < n ? counts[b + q] : 0u;
    s += c[q];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint64_t v = threadIdx.x >= static_cast<unsigned>(off) ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t run = tile_offsets[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    if (b + q < n) offsets[b + q] = run;
    run += c[q];
  }
}

// offsets[i] = counts[0] + ... + counts[i-1]; *total = the sum of all n
// (device pointers, on stream s).  Every count must be below 2^20.
void scan_counts(const uint32_t *counts, uint64_t n, uint64_t *offsets, uint64_t *total,
                 hipStream_t s) {
  if (n <= kScanTile) {
    scan_counts_kernel<<<1, 1024, 0, s>>>(counts, n, offsets, total);
    SPINED_CHECK_LAUNCH();
    return;
  }
  const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
  DeviceBuffer<uint32_t> tsum(tiles);
  DeviceBuffer<uint64_t> toff(tiles);
  scan_tile_sums_kernel<<<static_cast<unsigned>(tiles), kScanThreads, 0, s>>>(counts, n,
                                                                              tsum.ptr);
  SPINED_CHECK_LAUNCH();
  scan_counts_kernel<<<1, 1024, 0, s>>>(tsum.ptr, tiles, toff.ptr, total);
  SPINED_CHECK_LAUNCH();
  scan_tile_apply_kernel<<<static_cast<unsigned>(tiles), kScanThreads, 0, s>>>(counts, n,
                                                                               toff.ptr, offsets);
  SPINED_CHECK_LAUNCH();
  // the tile temporaries are freed on return: let the kernels finish first
  SPINED_HIP(hipStreamSynchronize(s));
}

} // namespace
} // namespace spined
