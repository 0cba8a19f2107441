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
#include "sample_gather.h"

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
// and, with carryIn, carryIn[i]
struct PairOut {
  static constexpr bool kRecords = false;
  uint32_t* __restrict__ keysOut;
  uint32_t* __restrict__ valsOut;
  const uint32_t* __restrict__ carryIn;
  uint32_t* __restrict__ carryOut;
  const uint2* recsIn = nullptr;
  __device__ __forceinline__ void operator()(uint32_t dst, uint32_t k, uint32_t i, const uint2*) const {
    keysOut[dst] = k;
    valsOut[dst] = i;
    if (carryIn) carryOut[dst] = carryIn[i];
  }
};

// the same with the record itself moved to its sorted position: recsOut[dst]
// = recsIn[i] (the 40-byte record as five 8-byte words, read from the tile's
// copy in LDS), posOut[dst] = dst (optional: the identity order the
// deterministic sort starts from)
struct RecordOut {
  static constexpr bool kRecords = true;
  uint32_t* __restrict__ keysOut;
  uint32_t* __restrict__ posOut;
  const uint32_t* __restrict__ carryIn;
  uint32_t* __restrict__ carryOut;
  const uint2* __restrict__ recsIn;
  uint2* __restrict__ recsOut;
  __device__ __forceinline__ void operator()(uint32_t dst, uint32_t k, uint32_t i, const uint2* rec) const {
    keysOut[dst] = k;
    if (posOut) posOut[dst] = dst;
    if (carryIn) carryOut[dst] = carryIn[i];
    uint2 w[5];
#pragma unroll
    for (int q = 0; q < 5; q++) w[q] = rec[q];
    uint2* d = recsOut + (size_t)dst * 5;
#pragma unroll
    for (int q = 0; q < 5; q++) d[q] = w[q];
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
  __shared__ uint32_t sDst[kSortTile], sKey[kSortTile], sVal[kSortTile];
  // the tile's records (RecordOut), loaded with coalesced reads
  __shared__ uint2 recStage[Out::kRecords ? kSortTile * 5 : 1];
  const uint32_t pad = 0xffffffffu;
  for (uint32_t k = threadIdx.x; k < numKeys; k += kSortThreads)
    running[k] = offsets[(size_t)blockIdx.x * numKeys + k];
  __syncthreads();
  const uint64_t first = (uint64_t)blockIdx.x * tilesPerBlock * kSortTile;
  const uint32_t t = threadIdx.x;
  for (uint32_t tile = 0; tile < tilesPerBlock; tile++) {
    const uint64_t base = first + (uint64_t)tile * kSortTile;
    if (base >= n) break;  // uniform over the block
    if constexpr (Out::kRecords) {
      const uint32_t words = (uint32_t)min<uint64_t>((uint64_t)kSortTile, (uint64_t)n - base) * 5u;
      const uint2* src = out.recsIn + base * 5;
      constexpr int kIt = (int)(kSortTile * 5 / kSortThreads);
      uint2 tmpw[kIt];
#pragma unroll
      for (int q = 0; q < kIt; q++) {
        const uint32_t w = t + (uint32_t)q * kSortThreads;
        if (w < words) tmpw[q] = src[w];
      }
#pragma unroll
      for (int q = 0; q < kIt; q++) {
        const uint32_t w = t + (uint32_t)q * kSortThreads;
        if (w < words) recStage[w] = tmpw[q];
      }
      // visible to the scatter below through the barriers of the sort
    }
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
    // destination of every sorted position: the key's running position + rank
    // in its run (pad for none), written out below in striped order so that
    // consecutive lanes write consecutive destinations of a run
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
      const int pos = (int)(t * kSortItems + j);
      const uint32_t dst = running[k[j]] + (uint32_t)(pos - runStart[j]);
      sDst[pos] = (v[j] == pad || dst >= n) ? pad : dst;
      sKey[pos] = k[j];
      sVal[pos] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
      const int pos = (int)(t + (uint32_t)j * kSortThreads);
      const uint32_t dst = sDst[pos];
      if (dst != pad) {
        const uint32_t vv = sVal[pos];
        out(dst, sKey[pos], vv, recStage + (Out::kRecords ? (size_t)(vv - base) * 5 : 0));
      }
    }
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
  counting_sort(temp, SortKeys{keys, nullptr, nullptr}, keyBegin, numKeys, n,
                PairOut{keysOut, valsOut, nullptr, nullptr}, s);
}

void counting_sort_leaves(void* temp, const uint32_t* nodeOf, const KDNode* keyNodes, const uint32_t* keyBegin,
                          uint32_t numKeys, uint32_t n, uint32_t* keysOut, uint32_t* valsOut, hipStream_t s,
                          const uint32_t* carryIn, uint32_t* carryOut) {
  counting_sort(temp, SortKeys{nullptr, nodeOf, keyNodes}, keyBegin, numKeys, n,
                PairOut{keysOut, valsOut, carryIn, carryOut}, s);
}

void counting_sort_leaf_records(void* temp, const uint32_t* nodeOf, const KDNode* keyNodes, const uint32_t* keyBegin,
                                uint32_t numKeys, uint32_t n, const void* recsIn, void* recsOut, uint32_t* keysOut,
                                uint32_t* posOut, const uint32_t* carryIn, uint32_t* carryOut, hipStream_t s) {
  counting_sort(temp, SortKeys{nullptr, nodeOf, keyNodes}, keyBegin, numKeys, n,
                RecordOut{keysOut, posOut, carryIn, carryOut, (const uint2*)recsIn, (uint2*)recsOut}, s);
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
// Each region's index range is sorted on the device without host round
// trips, by kernels small enough to share the CUs with the region kernel of
// the previous update (the spatial phase of an update overlaps it):
//   * tiles of up to 2048 samples in one workgroup: a
//     stable LDS radix sort by the weight's order key over the bits in which
//     the tile's keys differ; runs of equal weights (rare for continuous
//     weights) are put in order with the full comparator by one thread each,
//     and a run longer than 32 makes the tile sort by every field, least
//     significant first (from index order that gives the total order);
//   * ranges longer than one tile: merge passes (merge path: a workgroup
//     finds where its output range starts and ends in the two runs, stages
//     those inputs in LDS and merges them there), on the weight keys and the
//     full comparator where the weights are equal.
namespace {

// float offset inside the 40-byte record of the comparator's i-th field:
// weight, pdf, distance, pos.x, pos.y, pos.z, dir.x, dir.y, dir.z
__device__ __forceinline__ int less_field(int i) { return i < 3 ? 6 + i : i - 3; }

__device__ __forceinline__ uint32_t field_key(const float* __restrict__ recs, uint32_t idx, int field) {
  return idx == 0xffffffffu ? 0xffffffffu : float_order_key(recs[(size_t)idx * 10 + less_field(field)]);
}

// SampleDataLess after the weight (fields 1..8), then the record index
__device__ __forceinline__ bool record_less_after_weight(const float* __restrict__ recs, uint32_t a, uint32_t b) {
  for (int f = 1; f < 9; f++) {
    const uint32_t ka = field_key(recs, a, f), kb = field_key(recs, b, f);
    if (ka != kb) return ka < kb;
  }
  return a < b;
}

// the same, out of line: equal weights are rare, and sixteen inlined copies
// in the merge's unrolled searches make its code outgrow the instruction cache
__device__ __noinline__ bool record_less_tie(const float* __restrict__ recs, uint32_t a, uint32_t b) {
  return record_less_after_weight(recs, a, b);
}

struct DetSegment {
  uint32_t begin, count;
};

constexpr int kDetLocalRun = 32;  // longer runs of equal weights: the tile sorts by every field

template <int THREADS, int ITEMS>
__global__ void __launch_bounds__(THREADS) k_det_tile_sort(const float* __restrict__ recs,
                                                         const DetSegment* __restrict__ segs,
                                                         uint32_t* __restrict__ vals, const uint32_t* __restrict__ keysIn,
                                                         uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsAlt,
                                                         uint32_t* __restrict__ keysAlt) {
  typedef hipcub::BlockRadixSort<uint32_t, THREADS, ITEMS, uint32_t> Sort;
  constexpr int CAP = THREADS * ITEMS;
  __shared__ union {
    typename Sort::TempStorage sort;
    uint32_t idx[CAP];
  } sh;
  __shared__ uint32_t sKey[CAP];
  __shared__ int longRun;
  __shared__ uint32_t keyAnd, keyOr;
  const DetSegment sg0 = segs[blockIdx.x];
  // bit 31 of the count: the sorted tile goes to valsAlt / keysAlt (the first
  // merge pass reads it there)
  const DetSegment sg{sg0.begin, sg0.count & 0x7fffffffu};
  const bool alt = (sg0.count >> 31) != 0;
  const uint32_t t = threadIdx.x;
  uint32_t key[ITEMS], idx[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t pos = t * ITEMS + j;
    idx[j] = pos < sg.count ? vals[sg.begin + pos] : 0xffffffffu;
  }
  if (keysIn) {
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const uint32_t pos = t * ITEMS + j;
      key[j] = pos < sg.count ? keysIn[sg.begin + pos] : 0xffffffffu;
    }
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; j++) key[j] = field_key(recs, idx[j], 0);
  }
  if (t == 0) {
    longRun = 0;
    keyAnd = 0xffffffffu;
    keyOr = 0u;
  }
  __syncthreads();
  // only the low bits up to the highest bit in which the tile's weight keys
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
  // runs of equal weights: short runs put in order by the thread owning their first element
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
      while (b > 0 && record_less_after_weight(recs, v, sh.idx[p + b - 1])) {
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
  uint32_t* __restrict__ valsDst = alt ? valsAlt : vals;
  uint32_t* __restrict__ keysDst = alt ? keysAlt : keysOut;
  for (uint32_t p = t; p < sg.count; p += THREADS) {
    valsDst[sg.begin + p] = sh.idx[p];
    keysDst[sg.begin + p] = sKey[p];
  }
}

// One merge pass over the segments of the table (blockIdx.y): sorted runs
// [k*2w, k*2w + w) and [k*2w + w, (k+1)*2w) of each segment become one run (a
// segment of at most w elements is copied).  Merge path: workgroup x writes
// outputs [x*kMergeTile, (x+1)*kMergeTile) of its segment.  The number of
// elements of the first run among the first d outputs of a pair is the first
// a at which "A[a] before B[d-1-a]" turns false along the diagonal; one wave
// finds it for the workgroup's first and last output with 64 probes per round
// (a range shrinks 64-fold per round of loads).  The workgroup then stages
// its inputs in LDS, every thread finds its own split in LDS and merges
// kMergeItems outputs sequentially, and the outputs leave through LDS in
// order.  Order: weight keys, the full comparator where they are equal (no
// two elements are equal in that total order).
constexpr int kMergeItems = 8, kMergeTile = 256 * kMergeItems;

__device__ __forceinline__ bool elem_less(const float* __restrict__ recs, uint32_t ka, uint32_t ia, uint32_t kb,
                                          uint32_t ib) {
  if (ka != kb) return ka < kb;
  return record_less_tie(recs, ia, ib);
}

// first a in [lo, hi) with !(A[a] before B[d-1-a]) (hi if none); uniform over the wave
__device__ uint32_t wave_merge_split(const float* __restrict__ recs, const uint32_t* __restrict__ kA,
                                     const uint32_t* __restrict__ iA, const uint32_t* __restrict__ kB,
                                     const uint32_t* __restrict__ iB, uint32_t d, uint32_t lo, uint32_t hi) {
  const uint32_t lane = threadIdx.x & 63;
  while (lo < hi) {
    const uint32_t step = (hi - lo + 63) / 64;
    const uint32_t a = lo + lane * step;
    bool before = false;
    if (a < hi) {
      const uint32_t b = d - 1 - a;
      before = elem_less(recs, kA[a], iA[a], kB[b], iB[b]);
    }
    const uint32_t c = (uint32_t)__popcll(__ballot(before));  // probes 0 .. c-1 are before (monotone)
    if (c == 0) {
      hi = lo;
    } else {
      const uint32_t nlo = lo + (c - 1) * step + 1;
      hi = min(hi, lo + c * step);
      lo = nlo;
    }
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_det_merge(const float* __restrict__ recs, const DetSegment* __restrict__ segs,
                                                   uint32_t w, const uint32_t* __restrict__ idxIn,
                                                   const uint32_t* __restrict__ keyIn, uint32_t* __restrict__ idxOut,
                                                   uint32_t* __restrict__ keyOut) {
  __shared__ uint32_t sK[kMergeTile], sI[kMergeTile];
  __shared__ uint32_t sSplit[2];
  const DetSegment sg = segs[blockIdx.y];
  const uint32_t n = sg.count;
  const uint32_t o0 = blockIdx.x * kMergeTile;
  if (o0 >= n) return;  // uniform over the workgroup
  const uint32_t* __restrict__ inI = idxIn + sg.begin;
  const uint32_t* __restrict__ inK = keyIn + sg.begin;
  const uint32_t ps = o0 / (2 * w) * (2 * w);  // the pair (kMergeTile divides 2w)
  const uint32_t aBeg = ps, aEnd = min(ps + w, n), bEnd = min(ps + 2 * w, n);
  const uint32_t lenA = aEnd - aBeg, lenB = bEnd - aEnd;
  const uint32_t d0 = o0 - ps, d1 = min(o0 + kMergeTile, bEnd) - ps;
  const uint32_t wave = threadIdx.x >> 6;
  if (wave < 2) {
    const uint32_t d = wave == 0 ? d0 : d1;
    const uint32_t lo = d > lenB ? d - lenB : 0u, hi = min(d, lenA);
    const uint32_t a = wave_merge_split(recs, inK + aBeg, inI + aBeg, inK + aEnd, inI + aEnd, d, lo, hi);
    if ((threadIdx.x & 63) == 0) sSplit[wave] = a;
  }
  __syncthreads();
  const uint32_t a0 = sSplit[0], a1 = sSplit[1];
  const uint32_t na = a1 - a0, nb = (d1 - a1) - (d0 - a0);
  // stage A[a0, a1) then B[d0 - a0, d1 - a1): all loads before the LDS writes
  uint32_t tk[kMergeItems], ti[kMergeItems];
#pragma unroll
  for (int u = 0; u < kMergeItems; u++) {
    const uint32_t j = threadIdx.x + u * 256;
    tk[u] = ti[u] = 0;
    if (j < na + nb) {
      const uint32_t src = j < na ? aBeg + a0 + j : aEnd + (d0 - a0) + (j - na);
      tk[u] = inK[src];
      ti[u] = inI[src];
    }
  }
#pragma unroll
  for (int u = 0; u < kMergeItems; u++) {
    const uint32_t j = threadIdx.x + u * 256;
    if (j < na + nb) {
      sK[j] = tk[u];
      sI[j] = ti[u];
    }
  }
  __syncthreads();
  // this thread's outputs [dl, dl + kMergeItems) of the workgroup's range
  const uint32_t dl = threadIdx.x * kMergeItems, dn = na + nb;
  uint32_t outK[kMergeItems], outI[kMergeItems];
  if (dl < dn) {
    uint32_t lo = dl > nb ? dl - nb : 0u, hi = min(dl, na);
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1, b = na + dl - 1 - mid;
      if (elem_less(recs, sK[mid], sI[mid], sK[b], sI[b])) lo = mid + 1;
      else hi = mid;
    }
    uint32_t ia = lo, ib = na + (dl - lo);
    uint32_t hk = ia < na ? sK[ia] : 0u, hi2 = ia < na ? sI[ia] : 0u;
    uint32_t gk = ib < dn ? sK[ib] : 0u, gi = ib < dn ? sI[ib] : 0u;
#pragma unroll
    for (int u = 0; u < kMergeItems; u++) {
      const bool takeA = ia < na && (ib >= dn || elem_less(recs, hk, hi2, gk, gi));
      outK[u] = takeA ? hk : gk;
      outI[u] = takeA ? hi2 : gi;
      if (takeA) {
        ia++;
        if (ia < na) { hk = sK[ia]; hi2 = sI[ia]; }
      } else {
        ib++;
        if (ib < dn) { gk = sK[ib]; gi = sI[ib]; }
      }
    }
  }
  __syncthreads();
  if (dl < dn) {
#pragma unroll
    for (int u = 0; u < kMergeItems; u++) {
      sK[dl + u] = outK[u];
      sI[dl + u] = outI[u];
    }
  }
  __syncthreads();
  // outputs in order, coalesced
#pragma unroll
  for (int u = 0; u < kMergeItems; u++) {
    const uint32_t j = threadIdx.x + u * 256;
    if (j < dn) {
      idxOut[sg.begin + o0 + j] = sI[j];
      keyOut[sg.begin + o0 + j] = sK[j];
    }
  }
}


}  // namespace

constexpr uint32_t kDetTile = 256 * 8;
constexpr int kDetTables = 2;

static size_t det_table_entries(uint32_t n, uint32_t numSegments) { return (size_t)numSegments + n / kDetTile + 3; }
static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t det_sort_stage_bytes(uint32_t n, uint32_t numSegments) {
  return det_table_entries(n, numSegments) * kDetTables * sizeof(DetSegment);
}

size_t det_sort_temp_bytes(uint32_t n, uint32_t numSegments) {
  // the segment tables, the tiles' weight keys and the merge passes' second index and key buffers
  return align256(det_sort_stage_bytes(n, numSegments)) + 3 * align256(size_t(n) * 4);
}

void det_sort_segments(void* temp, void* pinnedHost, size_t pinnedBytes, const void* records,
                       const uint32_t* begin, const uint32_t* count, uint32_t numSegments, uint32_t n,
                       uint32_t* vals, uint32_t* weightKeys, hipStream_t s) {
  if (n == 0 || numSegments == 0) return;
  const float* recs = (const float*)records;
  // tables: 0 tiles of at most kDetTile (every range up to that and the tiles
  // of longer ranges), 1 ranges longer than one tile
  const size_t cap = det_table_entries(n, numSegments);
  if (pinnedBytes < det_sort_stage_bytes(n, numSegments))
    throw std::runtime_error("det_sort_segments: staging too small");
  uint32_t longest = 0;
  for (uint32_t r = 0; r < numSegments; r++) longest = std::max(longest, count[r]);
  // merge passes: the sorted tiles of long ranges start in the buffer that
  // makes the last pass end in vals
  uint32_t passes = 0;
  for (uint32_t w = kDetTile; w < longest; w *= 2) passes++;
  const uint32_t altFlag = (passes & 1) ? 0x80000000u : 0u;
  DetSegment* host = (DetSegment*)pinnedHost;
  DetSegment* cls[kDetTables] = {host, host + cap};
  uint32_t num[kDetTables] = {0, 0};
  for (uint32_t r = 0; r < numSegments; r++) {
    const uint32_t c = count[r];
    if (c < 2) continue;
    if (c <= kDetTile) {
      cls[0][num[0]++] = DetSegment{begin[r], c};
    } else {
      for (uint32_t o = 0; o < c; o += kDetTile)
        cls[0][num[0]++] = DetSegment{begin[r] + o, std::min(kDetTile, c - o) | altFlag};
      cls[1][num[1]++] = DetSegment{begin[r], c};
    }
  }
  char* t = (char*)temp;
  DetSegment* dev = (DetSegment*)t;
  t += align256(det_sort_stage_bytes(n, numSegments));
  uint32_t* keysA = weightKeys ? weightKeys : (uint32_t*)t;
  t += align256(size_t(n) * 4);
  uint32_t* idxB = (uint32_t*)t;
  t += align256(size_t(n) * 4);
  uint32_t* keysB = (uint32_t*)t;
  // only the filled parts of the two tables
  const size_t used = num[1] ? cap + num[1] : num[0];
  PGL_KERNEL_CHECK(hipMemcpyAsync(dev, host, used * sizeof(DetSegment), hipMemcpyHostToDevice, s));
  const uint32_t* keysIn = weightKeys;  // in place: a tile reads its keys before writing them
  if (num[0])
    hipLaunchKernelGGL((k_det_tile_sort<256, 8>), dim3(num[0]), dim3(256), 0, s, recs, dev, vals, keysIn, keysA, idxB,
                       keysB);
  if (num[1] == 0) return;
  // merge passes over all ranges longer than one tile, ping-pong between the buffers
  static_assert((2 * kDetTile) % kMergeTile == 0, "a merge workgroup's outputs lie in one pair of runs");
  const dim3 grid((longest + kMergeTile - 1) / kMergeTile, num[1]);
  uint32_t *ia = vals, *ka = keysA, *ib = idxB, *kb = keysB;
  if (altFlag) {
    std::swap(ia, ib);
    std::swap(ka, kb);
  }
  for (uint32_t w = kDetTile; w < longest; w *= 2) {
    hipLaunchKernelGGL(k_det_merge, grid, dim3(256), 0, s, recs, dev + cap, w, ia, ka, ib, kb);
    std::swap(ia, ib);
    std::swap(ka, kb);
  }
}

}  // namespace pglgpu
