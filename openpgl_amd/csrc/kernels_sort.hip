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

// ------------------------------------------------------------ deterministic order
// Field::fitRegions / updateRegions sort every region's samples with
// std::sort(SampleDataLess) in deterministic mode (field/Field.h:409-412,
// 469-472).  SampleDataLess (data/SampleData.h:82-93) is lexicographic over
// weight, pdf, distance, position x/y/z, direction x/y/z (flags ignored;
// -0 and +0 compare equal).  Records equal in all nine fields differ at most
// in their flags, which training does not read, so any order of them gives
// the same result; here they keep their index order (the total order used
// below is the nine fields, then the record index).
//
// Each region's index range is sorted on the device without host
// round trips:
//   * ranges of up to 16384 samples in one workgroup: a stable LDS radix sort
//     by the weight's order key; only if two weights of the range are equal,
//     stable LDS sorts by every field from the last to the first (least
//     significant first), which from index order gives the total order;
//   * longer ranges as sorted blocks of 16384, then merge passes in which
//     every element finds its place in the partner block by binary search
//     under the total order.
namespace {

// order-preserving map of a float to uint32 (-0 mapped to +0)
__device__ __forceinline__ uint32_t order_key(float x) {
  uint32_t b = __float_as_uint(x);
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// float offset inside the 40-byte record of the comparator's i-th field:
// weight, pdf, distance, pos.x, pos.y, pos.z, dir.x, dir.y, dir.z
__device__ __forceinline__ int less_field(int i) { return i < 3 ? 6 + i : i - 3; }

__device__ __forceinline__ uint32_t field_key(const float* __restrict__ recs, uint32_t idx, int field) {
  return idx == 0xffffffffu ? 0xffffffffu : order_key(recs[(size_t)idx * 10 + less_field(field)]);
}

// SampleDataLess, then the record index
__device__ __forceinline__ bool record_less(const float* __restrict__ recs, uint32_t a, uint32_t b) {
  for (int f = 0; f < 9; f++) {
    const uint32_t ka = field_key(recs, a, f), kb = field_key(recs, b, f);
    if (ka != kb) return ka < kb;
  }
  return a < b;
}

struct DetSegment {
  uint32_t begin, count;
};

constexpr int kDetLocalRun = 32;  // longer runs of equal weights: full sort by every field

template <int THREADS, int ITEMS>
__global__ void __launch_bounds__(THREADS) k_det_block_sort(const float* __restrict__ recs,
                                                          const DetSegment* __restrict__ segs,
                                                          uint32_t* __restrict__ vals) {
  typedef hipcub::BlockRadixSort<uint32_t, THREADS, ITEMS, uint32_t> Sort;
  constexpr int CAP = THREADS * ITEMS;
  __shared__ union {
    typename Sort::TempStorage sort;
    uint32_t idx[CAP];
  } sh;
  __shared__ uint32_t sKey[CAP];
  __shared__ int longRun;
  __shared__ uint32_t keyAnd, keyOr;
  const DetSegment sg = segs[blockIdx.x];
  const uint32_t t = threadIdx.x;
  uint32_t key[ITEMS], idx[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t pos = t * ITEMS + j;
    idx[j] = pos < sg.count ? vals[sg.begin + pos] : 0xffffffffu;
    key[j] = field_key(recs, idx[j], 0);
  }
  if (t == 0) {
    longRun = 0;
    keyAnd = 0xffffffffu;
    keyOr = 0u;
  }
  __syncthreads();
  // only the low bits up to the highest bit in which the range's weight keys
  // differ are sorted (the padding's all-ones keys still sort last: the sort
  // is stable and the padding comes last)
  uint32_t ka = 0xffffffffu, ko = 0u;
#pragma unroll
  for (int j = 0; j < ITEMS; j++)
    if (t * ITEMS + j < sg.count) {
      ka &= key[j];
      ko |= key[j];
    }
  for (int off = 32; off > 0; off >>= 1) {
    ka &= __shfl_xor(ka, off);
    ko |= __shfl_xor(ko, off);
  }
  if ((t & 63) == 0) {
    atomicAnd(&keyAnd, ka);
    atomicOr(&keyOr, ko);
  }
  __syncthreads();
  const uint32_t differ = keyAnd ^ keyOr;
  const int endBit = differ ? 32 - __clz(differ) : 1;
  Sort(sh.sort).Sort(key, idx, 0, endBit);  // blocked arrangement in and out; stable
  __syncthreads();
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    sKey[t * ITEMS + j] = key[j];
    sh.idx[t * ITEMS + j] = idx[j];
  }
  __syncthreads();
  // runs of equal weights (rare for continuous weights): short runs put in
  // order by the thread owning their first element, with the full comparator
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t p = t * ITEMS + j;
    if (p + 1 >= sg.count || sKey[p + 1] != sKey[p] || (p > 0 && sKey[p - 1] == sKey[p])) continue;
    uint32_t L = 2;
    while (p + L < sg.count && sKey[p + L] == sKey[p] && L <= kDetLocalRun) L++;
    if (L > kDetLocalRun) {
      longRun = 1;
      continue;
    }
    for (uint32_t a = 1; a < L; a++) {  // insertion sort of sh.idx[p, p + L)
      const uint32_t v = sh.idx[p + a];
      uint32_t b = a;
      while (b > 0 && record_less(recs, v, sh.idx[p + b - 1])) {
        sh.idx[p + b] = sh.idx[p + b - 1];
        b--;
      }
      sh.idx[p + b] = v;
    }
  }
  __syncthreads();
  if (longRun) {  // uniform over the block: stable sorts by every field, least significant first
    for (int f = 8; f >= 0; f--) {
#pragma unroll
      for (int j = 0; j < ITEMS; j++) key[j] = field_key(recs, idx[j], f);
      __syncthreads();
      Sort(sh.sort).Sort(key, idx);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) sh.idx[t * ITEMS + j] = idx[j];
    __syncthreads();
  }
  for (uint32_t p = t; p < sg.count; p += THREADS) vals[sg.begin + p] = sh.idx[p];
}

// one merge pass over the segments of the table (blockIdx.y): runs
// [k*2w, k*2w + w) and [k*2w + w, (k+1)*2w) of each segment are merged (a
// segment of at most w elements is copied); every element binary-searches
// its rank in the partner run under the total order
__global__ void __launch_bounds__(256) k_det_merge(const float* __restrict__ recs, const DetSegment* __restrict__ segs,
                                                   uint32_t w, const uint32_t* __restrict__ inBase,
                                                   uint32_t* __restrict__ outBase) {
  const DetSegment sg = segs[blockIdx.y];
  const uint32_t n = sg.count;
  const uint32_t* __restrict__ in = inBase + sg.begin;
  uint32_t* __restrict__ out = outBase + sg.begin;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t pair = i / (2 * w), start = pair * 2 * w;
    const uint32_t aEnd = min(start + w, n), bEnd = min(start + 2 * w, n);
    const uint32_t v = in[i];
    uint32_t lo, hi, rankBase;
    if (i < aEnd) {  // element of run A: the B elements before it
      lo = aEnd;
      hi = bEnd;
      rankBase = i - start;
    } else {  // element of run B: the A elements before it
      lo = start;
      hi = aEnd;
      rankBase = i - aEnd;
    }
    const uint32_t first = lo;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (record_less(recs, in[mid], v)) lo = mid + 1;
      else hi = mid;
    }
    out[start + rankBase + (lo - first)] = v;
  }
}

__global__ void __launch_bounds__(256) k_det_copy(const DetSegment* __restrict__ segs, const uint32_t* __restrict__ in,
                                                  uint32_t* __restrict__ out) {
  const DetSegment sg = segs[blockIdx.y];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sg.count; i += stride)
    out[sg.begin + i] = in[sg.begin + i];
}

}  // namespace

constexpr uint32_t kDetSmall = 256 * 4, kDetMid = 256 * 16, kDetBig = 512 * 16, kDetLarge = 512 * 32;
constexpr int kDetTables = 5;

static size_t det_table_entries(uint32_t n, uint32_t numSegments) { return (size_t)numSegments + n / kDetLarge + 3; }

size_t det_sort_stage_bytes(uint32_t n, uint32_t numSegments) {
  return det_table_entries(n, numSegments) * kDetTables * sizeof(DetSegment);
}

size_t det_sort_temp_bytes(uint32_t n, uint32_t numSegments) {
  // the segment tables and the ping-pong index buffer of the merge passes
  return ((det_sort_stage_bytes(n, numSegments) + 255) & ~size_t(255)) + size_t(n) * 4 + 256;
}

void det_sort_segments(void* temp, void* pinnedHost, size_t pinnedBytes, const void* records,
                       const uint32_t* begin, const uint32_t* count, uint32_t numSegments, uint32_t n,
                       uint32_t* vals, hipStream_t s) {
  if (n == 0 || numSegments == 0) return;
  const float* recs = (const float*)records;
  // tables: 0 ranges <= 1024, 1 <= 4096, 2 <= 8192, 3 blocks of <= 16384
  // (ranges up to 16384 and the blocks of longer ranges), 4 ranges longer
  // than 16384
  const size_t cap = det_table_entries(n, numSegments);
  if (pinnedBytes < det_sort_stage_bytes(n, numSegments))
    throw std::runtime_error("det_sort_segments: staging too small");
  DetSegment* host = (DetSegment*)pinnedHost;
  DetSegment* cls[kDetTables];
  for (int c = 0; c < kDetTables; c++) cls[c] = host + c * cap;
  uint32_t num[kDetTables] = {0, 0, 0, 0, 0};
  uint32_t longest = 0;
  for (uint32_t r = 0; r < numSegments; r++) {
    const uint32_t c = count[r];
    if (c < 2) continue;
    if (c <= kDetSmall) {
      cls[0][num[0]++] = DetSegment{begin[r], c};
    } else if (c <= kDetMid) {
      cls[1][num[1]++] = DetSegment{begin[r], c};
    } else if (c <= kDetBig) {
      cls[2][num[2]++] = DetSegment{begin[r], c};
    } else {
      for (uint32_t o = 0; o < c; o += kDetLarge) cls[3][num[3]++] = DetSegment{begin[r] + o, std::min(kDetLarge, c - o)};
      if (c > kDetLarge) {
        cls[4][num[4]++] = DetSegment{begin[r], c};
        longest = std::max(longest, c);
      }
    }
  }
  DetSegment* dev = (DetSegment*)temp;
  PGL_KERNEL_CHECK(hipMemcpyAsync(dev, host, det_sort_stage_bytes(n, numSegments), hipMemcpyHostToDevice, s));
  if (num[0]) hipLaunchKernelGGL((k_det_block_sort<256, 4>), dim3(num[0]), dim3(256), 0, s, recs, dev, vals);
  if (num[1]) hipLaunchKernelGGL((k_det_block_sort<256, 16>), dim3(num[1]), dim3(256), 0, s, recs, dev + cap, vals);
  if (num[2]) hipLaunchKernelGGL((k_det_block_sort<512, 16>), dim3(num[2]), dim3(512), 0, s, recs, dev + 2 * cap, vals);
  if (num[3]) hipLaunchKernelGGL((k_det_block_sort<512, 32>), dim3(num[3]), dim3(512), 0, s, recs, dev + 3 * cap, vals);
  if (num[4] == 0) return;
  // merge passes over all ranges longer than one block, ping-pong through the temp buffer
  uint32_t* tmpIdx =
      (uint32_t*)((char*)temp + ((det_sort_stage_bytes(n, numSegments) + 255) & ~size_t(255)));
  const dim3 grid(std::min<uint32_t>((longest + 255) / 256, 256), num[4]);
  uint32_t *a = vals, *b = tmpIdx;
  for (uint32_t w = kDetLarge; w < longest; w *= 2) {
    hipLaunchKernelGGL(k_det_merge, grid, dim3(256), 0, s, recs, dev + 4 * cap, w, a, b);
    std::swap(a, b);
  }
  if (a != vals) hipLaunchKernelGGL(k_det_copy, grid, dim3(256), 0, s, dev + 4 * cap, a, vals);
}

}  // namespace pglgpu
