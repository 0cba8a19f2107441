// Stable counting sort of (region key, sample index) pairs for keys in
// [0, numKeys), numKeys <= kSortMaxKeys: the region-contiguous order of the
// update's samples.  Three kernels:
//   1. per block, the histogram of the keys of its contiguous range of tiles;
//   2. per key, the exclusive scan of the block counts starting at the key's
//      first output position (the running sum of the key counts, given);
//   3. per block, tile by tile in order: a stable LDS radix sort of the
//      tile's (key, index) pairs, the rank of each element inside its run of
//      equal keys (block max-scan of the run heads) and the scatter to
//      running position + rank.
// The output is identical to a stable radix sort of the pairs.
#include <hip/hip_runtime.h>

#include <hipcub/block/block_radix_sort.hpp>
#include <hipcub/block/block_reduce.hpp>
#include <hipcub/block/block_scan.hpp>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <algorithm>

#include "kernels.h"

namespace pglgpu {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 4;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;

// key of element i: keys[i], or the region of the leaf node nodeOf[i]
// (keyNodes[nodeOf[i]].idx()) when keys is null
struct SortKeys {
  const uint32_t* __restrict__ keys;
  const uint32_t* __restrict__ nodeOf;
  const KDNode* __restrict__ keyNodes;
  __device__ __forceinline__ uint32_t operator()(uint64_t i) const {
    return keys ? keys[i] : keyNodes[nodeOf[i]].idx();
  }
};

__global__ void __launch_bounds__(kSortThreads) k_sort_hist(SortKeys keys, uint32_t n,
                                                            uint32_t tilesPerBlock, uint32_t numKeys,
                                                            uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[kSortMaxKeys];
  for (uint32_t k = threadIdx.x; k < numKeys; k += kSortThreads) hist[k] = 0u;
  __syncthreads();
  const uint64_t begin = (uint64_t)blockIdx.x * tilesPerBlock * kSortTile;
  const uint64_t end = min<uint64_t>((uint64_t)n, begin + (uint64_t)tilesPerBlock * kSortTile);
  for (uint64_t i = begin + threadIdx.x; i < end; i += kSortThreads) {
    const uint32_t k = keys(i);
    if (k < numKeys) atomicAdd(&hist[k], 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < numKeys; k += kSortThreads) counts[(size_t)blockIdx.x * numKeys + k] = hist[k];
}

// one workgroup per key: offsets[b][key] = keyBegin[key] + sum_{b' < b} counts[b'][key]
__global__ void __launch_bounds__(kSortThreads) k_sort_scan(const uint32_t* __restrict__ counts, uint32_t blocks,
                                                            uint32_t numKeys, const uint32_t* __restrict__ keyBegin,
                                                            uint32_t* __restrict__ offsets) {
  typedef hipcub::BlockScan<uint32_t, kSortThreads> Scan;
  __shared__ typename Scan::TempStorage tmp;
  __shared__ uint32_t carry;
  const uint32_t key = blockIdx.x;
  if (threadIdx.x == 0) carry = keyBegin[key];
  __syncthreads();
  for (uint32_t b0 = 0; b0 < blocks; b0 += kSortThreads) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t c = b < blocks ? counts[(size_t)b * numKeys + key] : 0u;
    uint32_t ex, total;
    Scan(tmp).ExclusiveSum(c, ex, total);
    if (b < blocks) offsets[(size_t)b * numKeys + key] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += total;
    __syncthreads();
  }
}

struct MaxOp {
  __device__ __forceinline__ int operator()(int a, int b) const { return a > b ? a : b; }
};

// element (key k, index i) at sorted position dst: the (key, index) pair
struct PairOut {
  uint32_t* __restrict__ keysOut;
  uint32_t* __restrict__ valsOut;
  __device__ __forceinline__ void operator()(uint32_t dst, uint32_t k, uint32_t i) const {
    keysOut[dst] = k;
    valsOut[dst] = i;
  }
};

template <class Out>
__global__ void __launch_bounds__(kSortThreads) k_sort_scatter(SortKeys keys, uint32_t n,
                                                               uint32_t tilesPerBlock, uint32_t numKeys, int bits,
                                                               const uint32_t* __restrict__ offsets, Out out) {
  typedef hipcub::BlockRadixSort<uint32_t, kSortThreads, kSortItems, uint32_t> Sort;
  typedef hipcub::BlockScan<int, kSortThreads> Scan;
  __shared__ union {
    typename Sort::TempStorage sort;
    typename Scan::TempStorage scan;
  } tmp;
  __shared__ uint32_t running[kSortMaxKeys];
  __shared__ uint32_t firstKey[kSortThreads], firstVal[kSortThreads], lastKey[kSortThreads];
  const uint32_t pad = 0xffffffffu;
  for (uint32_t k = threadIdx.x; k < numKeys; k += kSortThreads)
    running[k] = offsets[(size_t)blockIdx.x * numKeys + k];
  __syncthreads();
  const uint64_t first = (uint64_t)blockIdx.x * tilesPerBlock * kSortTile;
  const uint32_t t = threadIdx.x;
  for (uint32_t tile = 0; tile < tilesPerBlock; tile++) {
    const uint64_t base = first + (uint64_t)tile * kSortTile;
    if (base >= n) break;  // uniform over the block
    uint32_t k[kSortItems], v[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
      const uint64_t i = base + t * kSortItems + j;
      const bool ok = i < n;
      const uint32_t key = ok ? keys(i) : 0u;
      const bool real = ok && key < numKeys;  // keys out of range are dropped (never written out of bounds)
      k[j] = real ? key : ((1u << bits) - 1u);  // padding sorts after every real element of its key
      v[j] = real ? (uint32_t)i : pad;
    }
    Sort(tmp.sort).Sort(k, v, 0, bits);  // blocked arrangement: thread t holds positions 4t..4t+3
    __syncthreads();
    firstKey[t] = k[0];
    firstVal[t] = v[0];
    lastKey[t] = k[kSortItems - 1];
    __syncthreads();
    // run heads (position 0 is always one) and their max-scan: the start of each element's run
    int head[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
      const int pos = (int)(t * kSortItems + j);
      const bool isHead = j > 0 ? k[j] != k[j - 1] : (t > 0 && k[0] != lastKey[t - 1]);
      head[j] = isHead ? pos : 0;
    }
    int runStart[kSortItems];
    Scan(tmp.scan).InclusiveScan(head, runStart, MaxOp());
    // scatter to the key's running position + rank in its run
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
      if (v[j] == pad) continue;
      const int pos = (int)(t * kSortItems + j);
      const uint32_t dst = running[k[j]] + (uint32_t)(pos - runStart[j]);
      if (dst < n) out(dst, k[j], v[j]);
    }
    __syncthreads();
    // the last real element of each run advances its key's position
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
      if (v[j] == pad) continue;
      const int pos = (int)(t * kSortItems + j);
      uint32_t nk, nv;
      if (j + 1 < kSortItems) {
        nk = k[j + 1];
        nv = v[j + 1];
      } else if (t + 1 < kSortThreads) {
        nk = firstKey[t + 1];
        nv = firstVal[t + 1];
      } else {
        nk = pad;
        nv = pad;
      }
      if (nv == pad || nk != k[j]) running[k[j]] += (uint32_t)(pos - runStart[j] + 1);
    }
    __syncthreads();
  }
}

}  // namespace

size_t counting_sort_temp_bytes(uint32_t n, uint32_t numKeys) {
  const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
  const uint32_t blocks = std::min<uint32_t>(std::max<uint32_t>(tiles, 1), 1024);
  return 2 * (size_t)blocks * numKeys * sizeof(uint32_t);
}

template <class Out>
static void counting_sort(void* temp, SortKeys keys, const uint32_t* keyBegin, uint32_t numKeys, uint32_t n, Out out,
                          hipStream_t s) {
  if (n == 0 || numKeys == 0) return;
  const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
  const uint32_t blocks = std::min<uint32_t>(std::max<uint32_t>(tiles, 1), 1024);
  const uint32_t tilesPerBlock = (tiles + blocks - 1) / blocks;
  int bits = 1;
  while ((1u << bits) < numKeys && bits < 31) bits++;
  uint32_t* counts = (uint32_t*)temp;
  uint32_t* offsets = counts + (size_t)blocks * numKeys;
  hipLaunchKernelGGL(k_sort_hist, dim3(blocks), dim3(kSortThreads), 0, s, keys, n, tilesPerBlock, numKeys, counts);
  hipLaunchKernelGGL(k_sort_scan, dim3(numKeys), dim3(kSortThreads), 0, s, counts, blocks, numKeys, keyBegin, offsets);
  hipLaunchKernelGGL(k_sort_scatter<Out>, dim3(blocks), dim3(kSortThreads), 0, s, keys, n, tilesPerBlock, numKeys,
                     bits, offsets, out);
}

void counting_sort_pairs(void* temp, const uint32_t* keys, const uint32_t* keyBegin, uint32_t numKeys, uint32_t n,
                         uint32_t* keysOut, uint32_t* valsOut, hipStream_t s) {
  counting_sort(temp, SortKeys{keys, nullptr, nullptr}, keyBegin, numKeys, n, PairOut{keysOut, valsOut}, s);
}

void counting_sort_leaves(void* temp, const uint32_t* nodeOf, const KDNode* keyNodes, const uint32_t* keyBegin,
                          uint32_t numKeys, uint32_t n, uint32_t* keysOut, uint32_t* valsOut, hipStream_t s) {
  counting_sort(temp, SortKeys{nullptr, nodeOf, keyNodes}, keyBegin, numKeys, n, PairOut{keysOut, valsOut}, s);
}

// ------------------------------------------------------------ deterministic order
// Field::fitRegions / updateRegions sort every region's samples with
// std::sort(SampleDataLess) in deterministic mode (field/Field.h:409-412,
// 469-472).  SampleDataLess (data/SampleData.h:82-93) is lexicographic over
// weight, pdf, distance, position x/y/z, direction x/y/z (flags ignored;
// -0 and +0 compare equal).  On the device: a radix sort of
// (region << 32 | key(weight)) pairs, then, only while equal keys remain,
// refinement passes that sort (group << 32 | key(next field)) where group
// numbers the runs of equal keys so far.  The result is the order of the
// comparator; records equal in all nine fields differ at most in their flags,
// which training does not read, so their relative order does not matter.
namespace {

// order-preserving map of a float to uint32 (-0 mapped to +0)
__device__ __forceinline__ uint32_t order_key(float x) {
  uint32_t b = __float_as_uint(x);
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// float offset inside the 40-byte record of the comparator's i-th field
__device__ __forceinline__ int less_field(int i) {
  // weight, pdf, distance, pos.x, pos.y, pos.z, dir.x, dir.y, dir.z
  return i < 3 ? 6 + i : i - 3;
}

__global__ void __launch_bounds__(256) k_det_first(SortKeys keys, const float* __restrict__ recs, uint32_t n,
                                                   uint64_t* __restrict__ keyOut, uint32_t* __restrict__ valOut) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t r = keys(i);
    keyOut[i] = (r << 32) | order_key(recs[(size_t)i * 10 + less_field(0)]);
    valOut[i] = i;
  }
}

// head[j] = 1 where a run of equal keys starts; counter += positions that
// continue a run; regionOut (first pass only) = key >> 32
__global__ void __launch_bounds__(256) k_det_heads(const uint64_t* __restrict__ key, uint32_t n,
                                                   uint32_t* __restrict__ head, uint32_t* __restrict__ counter,
                                                   uint32_t* __restrict__ regionOut) {
  typedef hipcub::BlockReduce<uint32_t, 256> Reduce;
  __shared__ typename Reduce::TempStorage tmp;
  uint32_t ties = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const uint64_t k = key[j];
    const bool h = j == 0 || key[j - 1] != k;
    head[j] = h ? 1u : 0u;
    ties += h ? 0u : 1u;
    if (regionOut) regionOut[j] = (uint32_t)(k >> 32);
  }
  const uint32_t t = Reduce(tmp).Sum(ties);
  if (threadIdx.x == 0 && t) atomicAdd(counter, t);
}

// next pass: (group of equal keys so far << 32) | key(field), same values
__global__ void __launch_bounds__(256) k_det_refine(const float* __restrict__ recs,
                                                    const uint32_t* __restrict__ valIn,
                                                    const uint32_t* __restrict__ groupIncl, int field, uint32_t n,
                                                    uint64_t* __restrict__ keyOut, uint32_t* __restrict__ valOut) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const uint32_t v = valIn[j];
    const uint64_t g = (uint64_t)(groupIncl[j] - 1u);
    keyOut[j] = (g << 32) | order_key(recs[(size_t)v * 10 + less_field(field)]);
    valOut[j] = v;
  }
}

struct DetLayout {
  size_t key[2], val, head, incl, sortTmp, scanTmp, sortBytes, scanBytes, total;
};

DetLayout det_layout(uint32_t n) {
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  DetLayout L;
  size_t sortBytes = 0, scanBytes = 0;
  hipcub::DoubleBuffer<uint64_t> kb(nullptr, nullptr);
  hipcub::DoubleBuffer<uint32_t> vb(nullptr, nullptr);
  PGL_KERNEL_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, kb, vb, (int)n, 0, 64));
  PGL_KERNEL_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, scanBytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                    (int)n));
  size_t off = 0;
  L.key[0] = off; off += up(size_t(n) * 8);
  L.key[1] = off; off += up(size_t(n) * 8);
  L.val = off; off += up(size_t(n) * 4);
  L.head = off; off += up(size_t(n) * 4);
  L.incl = off; off += up(size_t(n) * 4);
  L.sortTmp = off; off += up(sortBytes);
  L.scanTmp = off; off += up(scanBytes);
  L.sortBytes = sortBytes;
  L.scanBytes = scanBytes;
  L.total = off;
  return L;
}

int bits_for(uint64_t count) {
  int b = 0;
  while (b < 32 && (1ull << b) < count) b++;
  return b;
}

// the sort of both entry points; returns the number of refinement passes
int det_sort(void* temp, SortKeys keys, const void* records, uint32_t numKeys, uint32_t n, uint32_t* keysOut,
             uint32_t* valsOut, uint32_t* counterDev, uint32_t* counterHost, hipStream_t s) {
  if (n == 0) return 0;
  const DetLayout L = det_layout(n);
  char* t = (char*)temp;
  uint64_t* kA = (uint64_t*)(t + L.key[0]);
  uint64_t* kB = (uint64_t*)(t + L.key[1]);
  uint32_t* vB = (uint32_t*)(t + L.val);
  uint32_t* head = (uint32_t*)(t + L.head);
  uint32_t* incl = (uint32_t*)(t + L.incl);
  const float* recs = (const float*)records;
  const uint32_t g = std::min<uint32_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_det_first, dim3(g), dim3(256), 0, s, keys, recs, n, kA, valsOut);
  hipcub::DoubleBuffer<uint64_t> K(kA, kB);
  hipcub::DoubleBuffer<uint32_t> V(valsOut, vB);
  size_t sortBytes = L.sortBytes;
  PGL_KERNEL_CHECK(hipcub::DeviceRadixSort::SortPairs(t + L.sortTmp, sortBytes, K, V, (int)n, 0,
                                                      32 + bits_for(numKeys), s));
  int passes = 0;
  for (int field = 0;; field++) {
    PGL_KERNEL_CHECK(hipMemsetAsync(counterDev, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_det_heads, dim3(g), dim3(256), 0, s, K.Current(), n, head, counterDev,
                       field == 0 ? keysOut : nullptr);
    PGL_KERNEL_CHECK(hipMemcpyAsync(counterHost, counterDev, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PGL_KERNEL_CHECK(hipStreamSynchronize(s));
    const uint32_t ties = *counterHost;
    if (ties == 0 || field == 8) break;
    size_t scanBytes = L.scanBytes;
    PGL_KERNEL_CHECK(hipcub::DeviceScan::InclusiveSum(t + L.scanTmp, scanBytes, head, incl, (int)n, s));
    hipLaunchKernelGGL(k_det_refine, dim3(g), dim3(256), 0, s, recs, V.Current(), incl, field + 1, n,
                       K.Alternate(), V.Alternate());
    K.selector ^= 1;
    V.selector ^= 1;
    PGL_KERNEL_CHECK(hipcub::DeviceRadixSort::SortPairs(t + L.sortTmp, sortBytes, K, V, (int)n, 0,
                                                        32 + bits_for(uint64_t(n) - ties), s));
    passes++;
  }
  if (V.Current() != valsOut)
    PGL_KERNEL_CHECK(hipMemcpyAsync(valsOut, V.Current(), size_t(n) * 4, hipMemcpyDeviceToDevice, s));
  return passes;
}

}  // namespace

size_t det_sort_temp_bytes(uint32_t n) { return det_layout(std::max<uint32_t>(n, 1)).total; }

int det_sort_leaves(void* temp, const void* records, const uint32_t* nodeOf, const KDNode* keyNodes,
                    uint32_t numKeys, uint32_t n, uint32_t* keysOut, uint32_t* valsOut, uint32_t* counterDev,
                    uint32_t* counterHost, hipStream_t s) {
  return det_sort(temp, SortKeys{nullptr, nodeOf, keyNodes}, records, numKeys, n, keysOut, valsOut, counterDev,
                  counterHost, s);
}

int det_sort_keys(void* temp, const void* records, const uint32_t* keys, uint32_t numKeys, uint32_t n,
                  uint32_t* keysOut, uint32_t* valsOut, uint32_t* counterDev, uint32_t* counterHost, hipStream_t s) {
  return det_sort(temp, SortKeys{keys, nullptr, nullptr}, records, numKeys, n, keysOut, valsOut, counterDev,
                  counterHost, s);
}

}  // namespace pglgpu
