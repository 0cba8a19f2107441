// Batched path-segment back-propagation on gfx950: the training samples of
// many recorded paths (PathSegmentDataStorage::prepareSamples,
// data/PathSegmentDataStorage.h:160-315, applied to each path), written as
// PGLSampleData records, surface and volume apart, in the order the
// per-path calls would append them to one sample storage.
//
// One thread per vertex: the lanes of a wave cover consecutive vertices, so
// they read one contiguous stretch of the vertex array (lanes of one path
// read the same later vertices) instead of one path per lane with loads
// scattered over many paths.  A vertex's thread redoes the path's backward
// distance walk from the last vertex down to its own (the exact sequence of
// float operations of the serial walk) and then its own contribution
// (path_samples.h), and stores its record and a packed surface | volume << 32
// flag.  An inclusive scan of the flags gives every record its place: the
// records of earlier paths, plus those of the same path at later vertices
// (the walk emits vertex n-2 first).
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <hipcub/device/device_scan.hpp>

#include "kernels.h"
#include "path_samples.h"

namespace pglgpu {

namespace {

constexpr int kPathThreads = 256;
constexpr uint32_t kNoPath = 0xffffffffu;

// the path holding vertex v: the last p with offsets[p] <= v
__device__ __forceinline__ uint32_t path_of(const uint32_t* __restrict__ offsets, uint32_t numPaths, uint32_t v) {
  uint32_t lo = 0, hi = numPaths;
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (offsets[mid] <= v) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(kPathThreads) k_path_check(const uint32_t* __restrict__ offsets, uint32_t numPaths,
                                                             uint32_t numSegments,
                                                             unsigned long long* __restrict__ totals) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= numPaths) return;
  if (offsets[p] > offsets[p + 1] || offsets[p + 1] > numSegments) totals[2] = 1ull;
}

__global__ void __launch_bounds__(kPathThreads) k_path_vertex(const PathVertex* __restrict__ segs, uint32_t numSegments,
                                                              const uint32_t* __restrict__ offsets, uint32_t numPaths,
                                                              PathSampleOptions o, PathSample* __restrict__ records,
                                                              unsigned long long* __restrict__ flags,
                                                              uint32_t* __restrict__ pathOf) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < numSegments; v += stride) {
    unsigned long long flag = 0ull;
    uint32_t p = kNoPath;
    const uint32_t q = path_of(offsets, numPaths, v);
    const uint32_t b = offsets[q], e = offsets[q + 1];
    if (b <= v && v < e && e <= numSegments) {
      p = q;
      const int n = (int)(e - b), i = (int)(v - b);
      PathSample s;
      if (i <= n - 2 && path_vertex_sample(segs + b, n, i, o, s)) {
        records[v] = s;
        flag = (s.flags & 2u) ? (1ull << 32) : 1ull;
      }
    }
    flags[v] = flag;
    pathOf[v] = p;
  }
}

__global__ void k_path_totals(const unsigned long long* __restrict__ incl, uint32_t numSegments,
                              unsigned long long* __restrict__ totals) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const unsigned long long t = numSegments ? incl[numSegments - 1] : 0ull;
    totals[0] = t & 0xffffffffull;
    totals[1] = t >> 32;
  }
}

__global__ void __launch_bounds__(kPathThreads) k_path_scatter(const PathSample* __restrict__ records,
                                                               const unsigned long long* __restrict__ flags,
                                                               const uint32_t* __restrict__ pathOf,
                                                               const unsigned long long* __restrict__ incl,
                                                               const uint32_t* __restrict__ offsets,
                                                               uint32_t numSegments, PathSample* __restrict__ surfOut,
                                                               PathSample* __restrict__ volOut) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < numSegments; v += stride) {
    const unsigned long long f = flags[v];
    if (!f) continue;
    const uint32_t p = pathOf[v];
    const uint32_t b = offsets[p], e = offsets[p + 1];
    const unsigned long long before = b > 0 ? incl[b - 1] : 0ull;  // records of earlier paths
    const unsigned long long last = incl[e - 1], mine = incl[v];
    const int hi = (f >> 32) ? 32 : 0;
    // records of earlier paths + those of this path at later vertices
    const uint32_t pos = (uint32_t)(before >> hi) + ((uint32_t)(last >> hi) - (uint32_t)(mine >> hi));
    if (hi) volOut[pos] = records[v]; else surfOut[pos] = records[v];
  }
}

uint32_t path_grid(uint32_t n) {
  const uint32_t blocks = (n + kPathThreads - 1) / kPathThreads;
  return blocks < 65536u ? (blocks ? blocks : 1u) : 65536u;
}

}  // namespace

size_t path_scan_temp_bytes(uint32_t numSegments) {
  size_t bytes = 0;
  const unsigned long long* in = nullptr;
  unsigned long long* out = nullptr;
  PGL_KERNEL_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, (int)numSegments));
  return bytes;
}

void launch_path_records(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
                         const PathSampleOptions& o, PathSample* records, unsigned long long* flags, uint32_t* pathOf,
                         unsigned long long* incl, void* scanTemp, size_t scanBytes, unsigned long long* totals,
                         hipStream_t s) {
  PGL_KERNEL_CHECK(hipMemsetAsync(totals, 0, 3 * sizeof(unsigned long long), s));
  if (numPaths == 0 || numSegments == 0) return;
  hipLaunchKernelGGL(k_path_check, dim3((numPaths + kPathThreads - 1) / kPathThreads), dim3(kPathThreads), 0, s,
                     offsets, numPaths, numSegments, totals);
  hipLaunchKernelGGL(k_path_vertex, dim3(path_grid(numSegments)), dim3(kPathThreads), 0, s, segs, numSegments, offsets,
                     numPaths, o, records, flags, pathOf);
  PGL_KERNEL_CHECK(hipcub::DeviceScan::InclusiveSum(scanTemp, scanBytes, flags, incl, (int)numSegments, s));
  hipLaunchKernelGGL(k_path_totals, dim3(1), dim3(64), 0, s, incl, numSegments, totals);
}

void launch_path_scatter(const PathSample* records, const unsigned long long* flags, const uint32_t* pathOf,
                         const unsigned long long* incl, const uint32_t* offsets, uint32_t numSegments,
                         PathSample* surfOut, PathSample* volOut, hipStream_t s) {
  if (numSegments == 0) return;
  hipLaunchKernelGGL(k_path_scatter, dim3(path_grid(numSegments)), dim3(kPathThreads), 0, s, records, flags, pathOf,
                     incl, offsets, numSegments, surfOut, volOut);
}

}  // namespace pglgpu
