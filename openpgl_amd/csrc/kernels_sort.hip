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
#include <hipcub/block/block_scan.hpp>

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

}  // namespace pglgpu
