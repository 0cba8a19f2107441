// Batched path-segment back-propagation on gfx950: the training samples of
// many recorded paths (PathSegmentDataStorage::prepareSamples,
// data/PathSegmentDataStorage.h:160-315, applied to each path), written as
// PGLSampleData records, surface and volume apart, in the order the
// per-path calls would append them to one sample storage.
//
// One thread per vertex, 256 consecutive vertices per workgroup (a tile).
// The tile's vertices, and the rest of the path its last vertex belongs to
// (up to 64 more), are first copied to LDS with coalesced loads, together
// with the offsets of the tile's paths; each vertex's thread then finds its
// path in LDS, redoes the path's backward distance walk from the last vertex
// down to its own (the exact sequence of float operations of the serial
// walk) and its own contribution (path_samples.h) reading LDS, and stores its
// record and a packed surface | volume << 32 flag.  A path reaching beyond
// the staged vertices is read from global memory.  The first path of every
// tile is found beforehand, one thread per tile.  An inclusive scan of the
// flags gives every record its place: the records of earlier paths, plus
// those of the same path at later vertices (the walk emits vertex n-2 first).
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <hipcub/device/device_scan.hpp>

#include <mutex>
#include <stdexcept>
#include <string>

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

__global__ void k_path_clear(unsigned long long* __restrict__ totals) {
  if (threadIdx.x < 3) totals[threadIdx.x] = 0ull;
}

__global__ void __launch_bounds__(kPathThreads) k_path_check(const uint32_t* __restrict__ offsets, uint32_t numPaths,
                                                             uint32_t numSegments,
                                                             unsigned long long* __restrict__ totals) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= numPaths) return;
  if (offsets[p] > offsets[p + 1] || offsets[p + 1] > numSegments) totals[2] = 1ull;
}

constexpr uint32_t kStage = kPathThreads + 64;  // staged vertices of a tile
constexpr uint32_t kTileOffsets = 512;           // path offsets of a tile kept in LDS
constexpr uint32_t kStepBuckets = 64;            // walk lengths >= 63 share the last bucket

// tileFirst[t] = path of vertex t * kPathThreads, tileFirst[numTiles] = path
// of the last vertex (numSegments >= 1)
__global__ void __launch_bounds__(kPathThreads) k_path_tiles(const uint32_t* __restrict__ offsets, uint32_t numPaths,
                                                             uint32_t numSegments, uint32_t numTiles,
                                                             uint32_t* __restrict__ tileFirst) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > numTiles) return;
  tileFirst[t] = path_of(offsets, numPaths, t < numTiles ? t * kPathThreads : numSegments - 1);
}

// the last p in [0, count) with offs[p] <= v (0 if none)
__device__ __forceinline__ uint32_t tile_path(const uint32_t* offs, uint32_t count, uint32_t v) {
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (offs[mid] <= v) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(kPathThreads) k_path_vertex(const PathVertex* __restrict__ segs, uint32_t numSegments,
                                                              const uint32_t* __restrict__ offsets, uint32_t numPaths,
                                                              const uint32_t* __restrict__ tileFirst,
                                                              PathSampleOptions o, PathSample* __restrict__ records,
                                                              unsigned long long* __restrict__ flags,
                                                              uint32_t* __restrict__ pathOf) {
  __shared__ PathVertex stage[kStage];
  __shared__ uint32_t offs[kTileOffsets];
  __shared__ uint32_t stageEnd;
  __shared__ uint32_t hist[kStepBuckets];
  __shared__ uint32_t vb[kPathThreads], ve[kPathThreads], vq[kPathThreads];
  __shared__ uint16_t order[kPathThreads];
  const uint32_t tid = threadIdx.x, t0 = blockIdx.x * kPathThreads;
  const uint32_t t1 = min(t0 + kPathThreads, numSegments);
  const uint32_t v = t0 + tid;
  // the tile's paths p0 .. p1 (monotone offsets); their offsets p0 .. p1 + 1
  const uint32_t p0 = tileFirst[blockIdx.x];
  const uint32_t p1 = max(p0, min(tileFirst[blockIdx.x + 1], numPaths - 1));
  const uint32_t np = p1 - p0 + 2;
  const bool inLds = np <= kTileOffsets;
  if (inLds)
    for (uint32_t i = tid; i < np; i += kPathThreads) offs[i] = offsets[p0 + i];
  if (tid == 0) stageEnd = t1;
  __syncthreads();
  uint32_t q = kNoPath, b = 0, e = 0;
  if (v < t1) {
    q = inLds ? p0 + tile_path(offs, np - 1, v) : path_of(offsets, numPaths, v);
    b = inLds ? offs[q - p0] : offsets[q];
    e = inLds ? offs[q - p0 + 1] : offsets[q + 1];
    // the tile's last vertex extends the staged range to the end of its path
    if (v + 1 == t1 && e > t1 && e <= numSegments) stageEnd = min(e, t0 + kStage);
  }
  __syncthreads();
  const uint32_t nStage = stageEnd - t0;
  {
    // 8-byte words (the vertex alignment); every thread issues all its loads
    // before its first store, so a tile pays one memory latency
    constexpr uint32_t kWords = (uint32_t)(sizeof(PathVertex) / 8);
    constexpr uint32_t kIt = (kStage * kWords + kPathThreads - 1) / kPathThreads;
    const uint2* src = reinterpret_cast<const uint2*>(segs + t0);
    uint2* dst = reinterpret_cast<uint2*>(stage);
    const uint32_t words = nStage * kWords;
    uint2 tmp[kIt];
#pragma unroll
    for (uint32_t k = 0; k < kIt; k++) {
      const uint32_t i = tid + k * kPathThreads;
      if (i < words) tmp[k] = src[i];
    }
#pragma unroll
    for (uint32_t k = 0; k < kIt; k++) {
      const uint32_t i = tid + k * kPathThreads;
      if (i < words) dst[i] = tmp[k];
    }
  }
  // a vertex's walk takes e - v - 1 steps: the tile's vertices are handed to
  // the threads in decreasing order of that length (counting sort in LDS), so
  // that the lanes of a wave take about the same number of steps
  const bool valid = v < t1 && b <= v && v < e && e <= numSegments;
  const uint32_t steps = valid && v + 1 < e ? e - v - 1 : 0u;
  const uint32_t key = kStepBuckets - 1 - min(steps, kStepBuckets - 1);
  if (tid < kStepBuckets) hist[tid] = 0u;
  vb[tid] = b;
  ve[tid] = e;
  vq[tid] = valid ? q : kNoPath;
  __syncthreads();
  const uint32_t rank = v < t1 ? atomicAdd(&hist[key], 1u) : 0u;
  __syncthreads();
  if (tid < kStepBuckets) {  // exclusive prefix of the bucket counts (one wave)
    const uint32_t c = hist[tid];
    uint32_t x = c;
#pragma unroll
    for (int d = 1; d < (int)kStepBuckets; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, (int)kStepBuckets);
      if ((int)tid >= d) x += y;
    }
    hist[tid] = x - c;
  }
  __syncthreads();
  if (v < t1) order[hist[key] + rank] = (uint16_t)tid;
  __syncthreads();
  if (v >= t1) return;
  // this thread's vertex (the order within a bucket varies; each record does not)
  const uint32_t w = order[tid], u = t0 + w;
  const uint32_t ub = vb[w], ue = ve[w], uq = vq[w];
  unsigned long long flag = 0ull;
  if (uq != kNoPath) {
    PathSample s;
    bool ok = false;
    if (u + 1 < ue) {  // vertices n-1 (the last of a path) yield no sample
      // the walk reads vertices u .. e-1 only: vertex u is index 0 of the suffix
      if (ue <= t0 + nStage) ok = path_vertex_sample(stage + w, (int)(ue - u), 0, o, s);
      else ok = path_vertex_sample(segs + u, (int)(ue - u), 0, o, s);
    }
    if (ok) {
      records[u] = s;
      flag = (s.flags & 2u) ? (1ull << 32) : 1ull;
    }
  }
  (void)ub;
  flags[u] = flag;
  pathOf[u] = uq;
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

// ---- records routed by their EInsideVolume flag (SampleDataStorage.h:72-90):
// flag = surface | volume << 32 per record, inclusive scan, stable scatter
__global__ void __launch_bounds__(kPathThreads) k_route_flags(const PathSample* __restrict__ recs, uint32_t n,
                                                              unsigned long long* __restrict__ flags) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    flags[i] = (recs[i].flags & 2u) ? (1ull << 32) : 1ull;
}

__global__ void __launch_bounds__(kPathThreads) k_route_scatter(const PathSample* __restrict__ recs,
                                                                const unsigned long long* __restrict__ incl,
                                                                uint32_t n, PathSample* __restrict__ surfOut,
                                                                PathSample* __restrict__ volOut) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const PathSample r = recs[i];
    const unsigned long long c = incl[i];
    if (r.flags & 2u) volOut[(uint32_t)(c >> 32) - 1u] = r;
    else surfOut[(uint32_t)(c & 0xffffffffull) - 1u] = r;
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

size_t path_tile_words(uint32_t numSegments) { return (numSegments + kPathThreads - 1) / kPathThreads + 1; }

void launch_path_records(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
                         const PathSampleOptions& o, PathSample* records, unsigned long long* flags, uint32_t* pathOf,
                         uint32_t* tileFirst, unsigned long long* incl, void* scanTemp, size_t scanBytes,
                         unsigned long long* totals, hipStream_t s) {
  hipLaunchKernelGGL(k_path_clear, dim3(1), dim3(64), 0, s, totals);
  if (numPaths == 0 || numSegments == 0) return;
  hipLaunchKernelGGL(k_path_check, dim3((numPaths + kPathThreads - 1) / kPathThreads), dim3(kPathThreads), 0, s,
                     offsets, numPaths, numSegments, totals);
  const uint32_t numTiles = (numSegments + kPathThreads - 1) / kPathThreads;
  hipLaunchKernelGGL(k_path_tiles, dim3(numTiles / kPathThreads + 1), dim3(kPathThreads), 0, s, offsets, numPaths,
                     numSegments, numTiles, tileFirst);
  hipLaunchKernelGGL(k_path_vertex, dim3(numTiles), dim3(kPathThreads), 0, s, segs, numSegments, offsets, numPaths,
                     tileFirst, o, records, flags, pathOf);
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

void launch_route_count(const PathSample* recs, uint32_t n, unsigned long long* flags, unsigned long long* incl,
                        void* scanTemp, size_t scanBytes, unsigned long long* totals, hipStream_t s) {
  hipLaunchKernelGGL(k_path_clear, dim3(1), dim3(64), 0, s, totals);
  if (n == 0) return;
  hipLaunchKernelGGL(k_route_flags, dim3(path_grid(n)), dim3(kPathThreads), 0, s, recs, n, flags);
  PGL_KERNEL_CHECK(hipcub::DeviceScan::InclusiveSum(scanTemp, scanBytes, flags, incl, (int)n, s));
  hipLaunchKernelGGL(k_path_totals, dim3(1), dim3(64), 0, s, incl, n, totals);
}

void launch_route_scatter(const PathSample* recs, const unsigned long long* incl, uint32_t n, PathSample* surfOut,
                          PathSample* volOut, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_route_scatter, dim3(path_grid(n)), dim3(kPathThreads), 0, s, recs, incl, n, surfOut, volOut);
}

// ---- the variant that allocates nothing
//
// Its scratch (records, flags, their scan, path of each vertex, tile table,
// scan temporary storage, totals) is carved from one file-scope device array
// of a fixed size; each GPU holds its own copy for the life of the process.
// Host code reaches it only through the address hipGetSymbolAddress returns
// on that device.

namespace {

constexpr size_t na_align(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kNaRecords = 0;
constexpr size_t kNaFlags = kNaRecords + na_align(size_t(kNoAllocMaxVertices) * sizeof(PathSample));
constexpr size_t kNaIncl = kNaFlags + na_align(size_t(kNoAllocMaxVertices) * 8);
constexpr size_t kNaPathOf = kNaIncl + na_align(size_t(kNoAllocMaxVertices) * 8);
constexpr size_t kNaTiles = kNaPathOf + na_align(size_t(kNoAllocMaxVertices) * 4);
constexpr size_t kNaTotals = kNaTiles + na_align(size_t(kNoAllocMaxVertices / kPathThreads + 1) * 4);
constexpr size_t kNaScan = kNaTotals + 256;
constexpr size_t kNaScanBytes = size_t(1) << 18;  // the scan's temporary storage (checked at set-up)
constexpr size_t kNaBytes = kNaScan + kNaScanBytes;
static_assert(kNaBytes == kNoAllocScratchBytes, "scratch layout and its declared size differ");
constexpr int kNaMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// this device's copy of scratch, looked up once per device (with it current)
static unsigned char* noalloc_scratch() {
  static std::mutex lock;
  static unsigned char* ptr[kNaMaxDevices] = {};
  int dev = 0;
  PGL_KERNEL_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) throw std::runtime_error("openpgl_amd: device index out of range");
  std::lock_guard<std::mutex> g(lock);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw std::runtime_error(std::string("openpgl_amd: hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e));
    if (path_scan_temp_bytes(kNoAllocMaxVertices) > kNaScanBytes)
      throw std::runtime_error("openpgl_amd: the path scan needs more temporary storage than the static scratch holds");
    ptr[dev] = static_cast<unsigned char*>(p);
  }
  return ptr[dev];
}

void path_prepare_noalloc(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
                          const PathSampleOptions& o, PathSample* surfOut, size_t surfCap, PathSample* volOut,
                          size_t volCap, size_t& numSurf, size_t& numVol, hipStream_t s) {
  numSurf = numVol = 0;
  if (numSegments > kNoAllocMaxVertices)
    throw std::runtime_error("openpgl_amd: the allocation-free path batch takes at most " +
                             std::to_string(kNoAllocMaxVertices) + " vertices per call (got " +
                             std::to_string(numSegments) + ")");
  unsigned char* base = noalloc_scratch();
  const uint32_t n = numPaths ? numSegments : 0;
  PathSample* records = reinterpret_cast<PathSample*>(base + kNaRecords);
  unsigned long long* flags = reinterpret_cast<unsigned long long*>(base + kNaFlags);
  unsigned long long* incl = reinterpret_cast<unsigned long long*>(base + kNaIncl);
  uint32_t* pathOf = reinterpret_cast<uint32_t*>(base + kNaPathOf);
  uint32_t* tiles = reinterpret_cast<uint32_t*>(base + kNaTiles);
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(base + kNaTotals);
  const size_t scanBytes = path_scan_temp_bytes(n > 0 ? n : 1u);
  if (scanBytes > kNaScanBytes) throw std::runtime_error("openpgl_amd: path scan storage exceeds the static scratch");
  launch_path_records(segs, n, offsets, numPaths, o, records, flags, pathOf, tiles, incl, base + kNaScan, scanBytes,
                      totals, s);
  unsigned long long t[3] = {0ull, 0ull, 0ull};
  PGL_KERNEL_CHECK(hipMemcpyAsync(t, totals, sizeof(t), hipMemcpyDeviceToHost, s));
  PGL_KERNEL_CHECK(hipStreamSynchronize(s));
  if (t[2]) throw std::runtime_error("openpgl_amd: path offsets must be non-decreasing and within the segments");
  numSurf = (size_t)t[0];
  numVol = (size_t)t[1];
  if (numSurf > surfCap || numVol > volCap)
    throw std::runtime_error("openpgl_amd: path samples exceed the output capacity (" + std::to_string(numSurf) +
                             " surface, " + std::to_string(numVol) + " volume)");
  if (numSurf == 0 && numVol == 0) return;
  launch_path_scatter(records, flags, pathOf, incl, offsets, n, surfOut, volOut, s);
  PGL_KERNEL_CHECK(hipStreamSynchronize(s));
}

}  // namespace pglgpu
