// Batched path-segment back-propagation on gfx950: the training samples of
// many recorded paths (PathSegmentDataStorage::prepareSamples,
// data/PathSegmentDataStorage.h:160-315, applied to each path), written as
// PGLSampleData records, surface and volume apart, in the order the
// per-path calls would append them to one sample storage.
//
// One thread per path (a path is short and its samples depend on a running
// backward chain): pass 1 counts each path's surface and volume samples,
// an exclusive scan of the packed counts gives every path's output offsets,
// pass 2 recomputes and writes.  Both passes run the same routine
// (path_samples.h), so the counts and the records agree.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <hipcub/device/device_scan.hpp>

#include "kernels.h"
#include "path_samples.h"

namespace pglgpu {

namespace {

constexpr int kPathThreads = 256;

__global__ void __launch_bounds__(kPathThreads) k_path_count(const PathVertex* __restrict__ segs,
                                                             const uint32_t* __restrict__ offsets, uint32_t numPaths,
                                                             uint32_t numSegments, PathSampleOptions o,
                                                             unsigned long long* __restrict__ packed,
                                                             unsigned long long* __restrict__ totals) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= numPaths) return;
  const uint32_t b = offsets[p], e = offsets[p + 1];
  uint32_t surf = 0, vol = 0;
  if (b > e || e > numSegments) {
    totals[2] = 1ull;  // malformed range: nothing is read, the host reports the error
  } else if (e > b) {
    path_samples(segs + b, (int)(e - b), o, [&](const PathSample& s) {
      if (s.flags & 2u) vol++; else surf++;
    });
  }
  packed[p] = (unsigned long long)surf | ((unsigned long long)vol << 32);
}

__global__ void k_path_totals(const unsigned long long* __restrict__ packed,
                              const unsigned long long* __restrict__ scanned, uint32_t numPaths,
                              unsigned long long* __restrict__ totals) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const unsigned long long t = numPaths ? scanned[numPaths - 1] + packed[numPaths - 1] : 0ull;
    totals[0] = t & 0xffffffffull;
    totals[1] = t >> 32;
  }
}

__global__ void __launch_bounds__(kPathThreads) k_path_emit(const PathVertex* __restrict__ segs,
                                                            const uint32_t* __restrict__ offsets, uint32_t numPaths,
                                                            uint32_t numSegments, PathSampleOptions o,
                                                            const unsigned long long* __restrict__ scanned,
                                                            PathSample* __restrict__ surfOut,
                                                            PathSample* __restrict__ volOut) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= numPaths) return;
  const uint32_t b = offsets[p], e = offsets[p + 1];
  if (e <= b || e > numSegments) return;
  uint32_t si = (uint32_t)(scanned[p] & 0xffffffffull);
  uint32_t vi = (uint32_t)(scanned[p] >> 32);
  path_samples(segs + b, (int)(e - b), o, [&](const PathSample& s) {
    if (s.flags & 2u) volOut[vi++] = s; else surfOut[si++] = s;
  });
}

}  // namespace

size_t path_scan_temp_bytes(uint32_t numPaths) {
  size_t bytes = 0;
  const unsigned long long* in = nullptr;
  unsigned long long* out = nullptr;
  PGL_KERNEL_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)numPaths));
  return bytes;
}

// totals: [0] surface samples, [1] volume samples, [2] 1 if a path range was malformed
void launch_path_count(const PathVertex* segs, const uint32_t* offsets, uint32_t numPaths, uint32_t numSegments,
                       const PathSampleOptions& o, unsigned long long* packed, unsigned long long* scanned,
                       void* scanTemp, size_t scanBytes, unsigned long long* totals, hipStream_t s) {
  PGL_KERNEL_CHECK(hipMemsetAsync(totals, 0, 3 * sizeof(unsigned long long), s));
  if (numPaths == 0) return;
  const uint32_t blocks = (numPaths + kPathThreads - 1) / kPathThreads;
  hipLaunchKernelGGL(k_path_count, dim3(blocks), dim3(kPathThreads), 0, s, segs, offsets, numPaths, numSegments, o,
                     packed, totals);
  PGL_KERNEL_CHECK(hipcub::DeviceScan::ExclusiveSum(scanTemp, scanBytes, packed, scanned, (int)numPaths, s));
  hipLaunchKernelGGL(k_path_totals, dim3(1), dim3(64), 0, s, packed, scanned, numPaths, totals);
}

void launch_path_emit(const PathVertex* segs, const uint32_t* offsets, uint32_t numPaths, uint32_t numSegments,
                      const PathSampleOptions& o, const unsigned long long* scanned, PathSample* surfOut,
                      PathSample* volOut, hipStream_t s) {
  if (numPaths == 0) return;
  const uint32_t blocks = (numPaths + kPathThreads - 1) / kPathThreads;
  hipLaunchKernelGGL(k_path_emit, dim3(blocks), dim3(kPathThreads), 0, s, segs, offsets, numPaths, numSegments, o,
                     scanned, surfOut, volOut);
}

}  // namespace pglgpu
