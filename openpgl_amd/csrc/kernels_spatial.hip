// Spatial-side kernels of Field::Update on gfx950: AoS->SoA ingest, scene
// bounds, the level-synchronous kd push-down with integer statistics, leaf
// assignment, region ranges and the gather + parallax reprojection into
// region-contiguous SoA streams.
//
// Reference behaviour:
//   IntegerSampleStatistics::addSample  data/SampleStatistics.h:261-277
//   partition predicate p[dim] < pivot  spatial/kdtree/KDTreeBuilder.h:193
//   estimateSceneBounds (skips the last sample)  field/Field.h:326-344
//   reprojectSample                      directional/vmm/ParallaxAwareVonMisesFisherWeightedEMFactory.h:1151-1181
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "common.h"
#include "kernels.h"
#include "sample_gather.h"

namespace pglgpu {


// One coalesced pass over the 40-byte records: positions to SoA for the kd
// rounds and, in deterministic mode (wkey not null), the order key of each
// weight for the within-region sort.
__global__ void __launch_bounds__(256) k_ingest(const SampleAoS* __restrict__ in, uint32_t n,
                                                float* __restrict__ px, float* __restrict__ py,
                                                float* __restrict__ pz, uint32_t* __restrict__ wkey) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const SampleAoS s = in[i];
    px[i] = s.px;
    py[i] = s.py;
    pz[i] = s.pz;
    if (wkey) wkey[i] = float_order_key(s.w);
  }
}

// min/max of positions over [0, n) (the caller passes n-1 to reproduce the
// reference's skipped last sample); one partial per block.
__global__ void __launch_bounds__(256) k_bounds(const float* __restrict__ px, const float* __restrict__ py,
                                                const float* __restrict__ pz, uint32_t n,
                                                float* __restrict__ partial /* [blocks][6] */) {
  __shared__ float red[6][256];
  float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
  float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float p[3] = {px[i], py[i], pz[i]};
    for (int j = 0; j < 3; j++) {
      lo[j] = fminf(lo[j], p[j]);
      hi[j] = fmaxf(hi[j], p[j]);
    }
  }
  for (int j = 0; j < 3; j++) {
    red[j][threadIdx.x] = lo[j];
    red[3 + j][threadIdx.x] = hi[j];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      for (int j = 0; j < 3; j++) {
        red[j][threadIdx.x] = fminf(red[j][threadIdx.x], red[j][threadIdx.x + s]);
        red[3 + j][threadIdx.x] = fmaxf(red[3 + j][threadIdx.x], red[3 + j][threadIdx.x + s]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Integer statistics of one point w.r.t. a node's (centre, 1/half-extent):
// t = (p - c) * inv; mean term trunc(t * 4096), variance term trunc((t * t) * 4096).
// Explicit round-to-nearest ops keep the sequence free of contraction.
__device__ __forceinline__ void int_terms(float p, float c, float inv, long long& m, long long& v) {
  const float t = __fmul_rn(__fsub_rn(p, c), inv);
  const float tv = __fmul_rn(__fmul_rn(t, t), 4096.f);
  const float ts = __fmul_rn(t, 4096.f);
  m = (long long)ts;  // C++ conversion truncates toward zero, as the reference's Vec3<int64> ctor
  v = (long long)tv;
}

static const int kSlotTile = 448;  // LDS accumulator slots per pass (44.8 KiB)
static const uint32_t kLdsNodes = 2048;  // trees up to this many nodes are walked from LDS (16 KiB)
static const uint32_t kLdsSlots = 4096;  // node -> frontier slot tables up to this size are read from LDS (16 KiB)
static const int kStatsItems = 4;        // samples per thread and iteration of the statistics passes
// 256 threads with at most 104 VGPRs: one wave per SIMD fits beside two
// workgroups of the region kernel (the spatial phase overlaps it)
static const uint32_t kStatsThreads = 256;

// Block-private statistics table in LDS, one array per field, so that slot s
// of a field and slot s + 1 sit in neighbouring banks.  (With one 128-byte
// record per slot, field f of every slot maps to the same two banks and the
// lanes' atomics on random slots serialise.)
struct SlotTable {
  uint32_t n[kSlotTile];
  long long sum[3][kSlotTile];
  long long sq[3][kSlotTile];
  long long lo[3][kSlotTile];
  long long hi[3][kSlotTile];
};

__device__ __forceinline__ void slot_init(SlotTable& tb, int c) {
  tb.n[c] = 0;
  for (int j = 0; j < 3; j++) {
    tb.sum[j][c] = 0;
    tb.sq[j][c] = 0;
    tb.lo[j][c] = 2147483647ll;   //  std::numeric_limits<int>::max()
    tb.hi[j][c] = -2147483647ll;  // -std::numeric_limits<int>::max()
  }
}

__device__ __forceinline__ void acc_init(IntStatsAcc* a) {
  a->n = 0;
  for (int j = 0; j < 3; j++) {
    a->sum[j] = 0;
    a->sq[j] = 0;
    a->lo[j] = 2147483647ll;
    a->hi[j] = -2147483647ll;
  }
}

// Accumulation into the block's LDS table.  When every active lane of the
// wave targets the same slot (samples arrive grouped by region after the
// first frames) the wave reduces with xor-shuffles and one lane issues the
// LDS atomics; otherwise each lane issues its own LDS atomics.  Every lane of
// the wave must call this; inactive lanes pass slot < 0.
__device__ void wave_accumulate(SlotTable& tb, int slot, const long long m[3], const long long v[3]) {
  const int lane = threadIdx.x & 63;
  const unsigned long long active = __ballot(slot >= 0);
  if (!active) return;
  const int leader = __ffsll((long long)active) - 1;
  const int t = __shfl(slot, leader);
  if (__ballot(slot >= 0 && slot != t) != 0) {
    if (slot >= 0) {
      atomicAdd(&tb.n[slot], 1u);
      for (int j = 0; j < 3; j++) {
        atomicAdd((unsigned long long*)&tb.sum[j][slot], (unsigned long long)m[j]);
        atomicAdd((unsigned long long*)&tb.sq[j][slot], (unsigned long long)v[j]);
        atomicMin(&tb.lo[j][slot], m[j]);
        atomicMax(&tb.hi[j][slot], m[j]);
      }
    }
    return;
  }
  const bool mine = slot >= 0;
  long long sm[3], sq[3], lo[3], hi[3];
  for (int j = 0; j < 3; j++) {
    sm[j] = mine ? m[j] : 0ll;
    sq[j] = mine ? v[j] : 0ll;
    lo[j] = mine ? m[j] : 0x7fffffffffffffffll;
    hi[j] = mine ? m[j] : (long long)0x8000000000000000ull;
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int j = 0; j < 3; j++) {
      sm[j] += __shfl_xor(sm[j], off);
      sq[j] += __shfl_xor(sq[j], off);
      const long long ol = __shfl_xor(lo[j], off), oh = __shfl_xor(hi[j], off);
      lo[j] = ol < lo[j] ? ol : lo[j];
      hi[j] = oh > hi[j] ? oh : hi[j];
    }
  }
  if (lane == leader) {
    atomicAdd(&tb.n[t], (uint32_t)__popcll(active));
    for (int j = 0; j < 3; j++) {
      atomicAdd((unsigned long long*)&tb.sum[j][t], (unsigned long long)sm[j]);
      atomicAdd((unsigned long long*)&tb.sq[j][t], (unsigned long long)sq[j]);
      atomicMin(&tb.lo[j][t], lo[j]);
      atomicMax(&tb.hi[j][t], hi[j]);
    }
  }
}

// Leaf parameters for the push-down to existing leaves: integer statistics of
// a leaf are taken w.r.t. its parent's bounds (the root leaf: the tree bounds).
struct LeafParams {
  float center[3];
  float invHalf[3];
};

// Statistics pass over all samples with a block-private accumulator table for
// slots [slotBase, slotBase + kSlotTile); each block writes its table to
// partial[block][slot] and k_reduce_partials combines the blocks (no global
// atomics: device-scope atomics serialise across the XCDs).
//   kDescend: every sample walks the existing inner nodes to its leaf
//             (KDTreeBuilder.h:292-349 for the levels that already exist);
//             statistics of the leaf w.r.t. its parent's bounds.  A tree of
//             at most kLdsNodes nodes is staged in LDS first (each level of
//             the walk is a dependent load).
//   else:     one level of the push-down below the frontier nodes: left iff
//             p[dim] < pivot; child statistics w.r.t. the frontier node's bounds.
template <bool kDescend>
__global__ void __launch_bounds__(kStatsThreads) k_spatial_stats(const float* __restrict__ px, const float* __restrict__ py,
                                                       const float* __restrict__ pz, uint32_t n,
                                                       const KDNode* __restrict__ nodes,
                                                       const int32_t* __restrict__ slotOfNode,
                                                       const LeafParams* __restrict__ leafParams,
                                                       const FrontierNode* __restrict__ frontier,
                                                       uint32_t* __restrict__ nodeOf, int writeNodes, int slotBase,
                                                       int numSlots, IntStatsAcc* __restrict__ partial,
                                                       uint32_t numNodes) {
  __shared__ SlotTable acc;
  // descend: the tree, every leaf's index replaced by its slot, and the
  // tile's leaf parameters; else: the slot of every node and the tile's
  // frontier nodes.  Each per-sample lookup is then an LDS access instead of
  // a dependent global load.
  __shared__ KDNode tree[kDescend ? kLdsNodes : 1];
  __shared__ LeafParams sPar[kDescend ? kSlotTile : 1];
  __shared__ int32_t sSlot[kDescend ? 1 : kLdsSlots];
  __shared__ FrontierNode sFr[kDescend ? 1 : kSlotTile / 2];
  const bool inLds = numNodes <= (kDescend ? kLdsNodes : kLdsSlots);
  const int f0 = slotBase / 2, nf = (numSlots + 1) / 2;  // frontier nodes of the tile (slotBase is even)
  // Staging: every load of a table is issued before the first LDS write (a
  // loop that waits for each load in turn pays the memory latency per
  // iteration).  Stores to memory come last in every iteration below for the
  // same reason: the loads that follow a store wait for it.
  if (kDescend) {
    if (inLds) {
      constexpr int kPer = kLdsNodes / kStatsThreads;
      KDNode tn[kPer];
      int32_t ts[kPer];
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const uint32_t c = threadIdx.x + u * kStatsThreads;
        if (c < numNodes) {
          tn[u] = nodes[c];
          ts[u] = slotOfNode[c];
        }
      }
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const uint32_t c = threadIdx.x + u * kStatsThreads;
        if (c < numNodes) {
          KDNode k = tn[u];
          if (k.isLeaf()) k.bits = (3u << 30) | (uint32_t)ts[u];
          tree[c] = k;
        }
      }
    }
    // leaf parameters as words (a struct copy through a register array goes through scratch)
    constexpr int kW = sizeof(LeafParams) / 4;
    constexpr int kPerP = (kSlotTile * kW + kStatsThreads - 1) / kStatsThreads;
    const float* src = (const float*)(leafParams + slotBase);
    float* dst = (float*)sPar;
    float tp[kPerP];
#pragma unroll
    for (int u = 0; u < kPerP; u++) {
      const int c = (int)threadIdx.x + u * (int)kStatsThreads;
      if (c < numSlots * kW) tp[u] = src[c];
    }
#pragma unroll
    for (int u = 0; u < kPerP; u++) {
      const int c = (int)threadIdx.x + u * (int)kStatsThreads;
      if (c < numSlots * kW) dst[c] = tp[u];
    }
  } else {
    if (inLds) {
      constexpr int kPer = kLdsSlots / kStatsThreads;
      int32_t ts[kPer];
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const uint32_t c = threadIdx.x + u * kStatsThreads;
        if (c < numNodes) ts[u] = slotOfNode[c];
      }
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const uint32_t c = threadIdx.x + u * kStatsThreads;
        if (c < numNodes) sSlot[c] = ts[u];
      }
    }
    constexpr int kW = sizeof(FrontierNode) / 4;
    constexpr int kPerF = (kSlotTile / 2 * kW + kStatsThreads - 1) / kStatsThreads;
    const uint32_t* src = (const uint32_t*)(frontier + f0);
    uint32_t* dst = (uint32_t*)sFr;
    uint32_t tf[kPerF];
#pragma unroll
    for (int u = 0; u < kPerF; u++) {
      const int c = (int)threadIdx.x + u * (int)kStatsThreads;
      if (c < nf * kW) tf[u] = src[c];
    }
#pragma unroll
    for (int u = 0; u < kPerF; u++) {
      const int c = (int)threadIdx.x + u * (int)kStatsThreads;
      if (c < nf * kW) dst[c] = tf[u];
    }
  }
  for (int c = threadIdx.x; c < numSlots; c += blockDim.x) slot_init(acc, c);
  __syncthreads();
  // kStatsItems samples per thread and iteration, their loads issued together
  const uint32_t step = blockDim.x * kStatsItems, stride = gridDim.x * step;
  for (uint32_t base = blockIdx.x * step; base < n; base += stride) {
    float p[kStatsItems][3];
    uint32_t nd[kStatsItems];
#pragma unroll
    for (int q = 0; q < kStatsItems; q++) {
      const uint32_t i = base + q * blockDim.x + threadIdx.x;
      nd[q] = 0;
      p[q][0] = p[q][1] = p[q][2] = 0.f;
      if (i < n) {
        p[q][0] = px[i];
        p[q][1] = py[i];
        p[q][2] = pz[i];
        if (!kDescend) nd[q] = nodeOf[i];
      }
    }
    // lookups of all samples first (the walks advance in lockstep, so their
    // dependent loads overlap), then the accumulation, then the stores
    int slot[kStatsItems];
    uint32_t node[kStatsItems];
    float cen[kStatsItems][3], inv[kStatsItems][3];  // centre and 1/half extent of the slot's bounds
    if (kDescend) {
      KDNode k[kStatsItems];
      if (inLds) {
#pragma unroll
        for (int q = 0; q < kStatsItems; q++) {
          node[q] = 0;
          k[q] = tree[0];
        }
        for (;;) {
          bool more = false;
#pragma unroll
          for (int q = 0; q < kStatsItems; q++) {
            if (k[q].isLeaf()) continue;
            more = true;
            const uint32_t d = k[q].dim();
            const float pd = d == 0 ? p[q][0] : (d == 1 ? p[q][1] : p[q][2]);
            node[q] = k[q].idx() + ((pd < k[q].pivot) ? 0u : 1u);
            k[q] = tree[node[q]];
          }
          if (!more) break;
        }
#pragma unroll
        for (int q = 0; q < kStatsItems; q++) slot[q] = (int)k[q].idx();
      } else {
#pragma unroll
        for (int q = 0; q < kStatsItems; q++) {
          node[q] = 0;
          k[q] = nodes[0];
        }
        for (;;) {
          bool more = false;
#pragma unroll
          for (int q = 0; q < kStatsItems; q++) {
            if (k[q].isLeaf()) continue;
            more = true;
            const uint32_t d = k[q].dim();
            const float pd = d == 0 ? p[q][0] : (d == 1 ? p[q][1] : p[q][2]);
            node[q] = k[q].idx() + ((pd < k[q].pivot) ? 0u : 1u);
            k[q] = nodes[node[q]];
          }
          if (!more) break;
        }
#pragma unroll
        for (int q = 0; q < kStatsItems; q++) slot[q] = slotOfNode[node[q]];
      }
#pragma unroll
      for (int q = 0; q < kStatsItems; q++) {
        const uint32_t i = base + q * blockDim.x + threadIdx.x;
        if (i >= n) slot[q] = -1;
        for (int j = 0; j < 3; j++) cen[q][j] = inv[q][j] = 0.f;
        if (slot[q] >= slotBase && slot[q] < slotBase + numSlots) {
          const LeafParams& lp = sPar[slot[q] - slotBase];
          for (int j = 0; j < 3; j++) { cen[q][j] = lp.center[j]; inv[q][j] = lp.invHalf[j]; }
        }
      }
    } else {
      int32_t f[kStatsItems];
      if (inLds) {
#pragma unroll
        for (int q = 0; q < kStatsItems; q++) f[q] = sSlot[nd[q]];
      } else {
#pragma unroll
        for (int q = 0; q < kStatsItems; q++) f[q] = slotOfNode[nd[q]];
      }
#pragma unroll
      for (int q = 0; q < kStatsItems; q++) {
        const uint32_t i = base + q * blockDim.x + threadIdx.x;
        slot[q] = -1;
        node[q] = 0;
        for (int j = 0; j < 3; j++) cen[q][j] = inv[q][j] = 0.f;
        if (i < n && f[q] >= 0) {
          const int fl = f[q] - f0;
          const FrontierNode fn = (fl >= 0 && fl < nf) ? sFr[fl] : frontier[f[q]];
          const float pd = fn.dim == 0 ? p[q][0] : (fn.dim == 1 ? p[q][1] : p[q][2]);
          const uint32_t side = (pd < fn.pivot) ? 0u : 1u;
          node[q] = fn.left + side;
          slot[q] = 2 * f[q] + (int)side;
          for (int j = 0; j < 3; j++) { cen[q][j] = fn.center[j]; inv[q][j] = fn.invHalf[j]; }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kStatsItems; q++) {
      int local = -1;
      long long m[3] = {0, 0, 0}, v[3] = {0, 0, 0};
      if (slot[q] >= slotBase && slot[q] < slotBase + numSlots) {
        local = slot[q] - slotBase;
        for (int j = 0; j < 3; j++) int_terms(p[q][j], cen[q][j], inv[q][j], m[j], v[j]);
      }
      wave_accumulate(acc, local, m, v);
    }
    if (writeNodes) {
#pragma unroll
      for (int q = 0; q < kStatsItems; q++) {
        const uint32_t i = base + q * blockDim.x + threadIdx.x;
        // descend: every sample; a round: the samples below the frontier
        if (i < n && (kDescend || slot[q] >= 0)) nodeOf[i] = node[q];
      }
    }
  }
  __syncthreads();
  IntStatsAcc* out = partial + (size_t)blockIdx.x * numSlots;
  for (int c = threadIdx.x; c < numSlots; c += blockDim.x) {
    IntStatsAcc a;
    a.n = acc.n[c];
    for (int j = 0; j < 3; j++) {
      a.sum[j] = acc.sum[j][c];
      a.sq[j] = acc.sq[j][c];
      a.lo[j] = acc.lo[j][c];
      a.hi[j] = acc.hi[j][c];
      a.pad[j] = 0;
    }
    out[c] = a;
  }
}

// out[slot] = combination of the per-block tables: one workgroup per slot,
// threads stride over the blocks, then a tree reduction in LDS (integer sums
// and min/max: exact and independent of the order).
__global__ void __launch_bounds__(256) k_reduce_partials(const IntStatsAcc* __restrict__ partial, uint32_t blocks,
                                                         int numSlots, IntStatsAcc* __restrict__ out) {
  __shared__ IntStatsAcc red[256];
  const int c = blockIdx.x;
  IntStatsAcc r;
  acc_init(&r);
  // two rows per step, both loaded before either is combined
  for (uint32_t b = threadIdx.x; b < blocks; b += 2 * blockDim.x) {
    const uint32_t b2 = b + blockDim.x;
    IntStatsAcc p[2];
    p[0] = partial[(size_t)b * numSlots + c];
    if (b2 < blocks) p[1] = partial[(size_t)b2 * numSlots + c];
    else acc_init(&p[1]);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      r.n += p[h].n;
      for (int j = 0; j < 3; j++) {
        r.sum[j] += p[h].sum[j];
        r.sq[j] += p[h].sq[j];
        r.lo[j] = p[h].lo[j] < r.lo[j] ? p[h].lo[j] : r.lo[j];
        r.hi[j] = p[h].hi[j] > r.hi[j] ? p[h].hi[j] : r.hi[j];
      }
    }
  }
  red[threadIdx.x] = r;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      IntStatsAcc& a = red[threadIdx.x];
      const IntStatsAcc& b = red[threadIdx.x + h];
      a.n += b.n;
      for (int j = 0; j < 3; j++) {
        a.sum[j] += b.sum[j];
        a.sq[j] += b.sq[j];
        a.lo[j] = b.lo[j] < a.lo[j] ? b.lo[j] : a.lo[j];
        a.hi[j] = b.hi[j] > a.hi[j] ? b.hi[j] : a.hi[j];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// sample -> region key (leaf data index) and identity index for the sort
__global__ void __launch_bounds__(256) k_leaf_keys(const uint32_t* __restrict__ nodeOf, const KDNode* __restrict__ nodes,
                                                   uint32_t n, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    keys[i] = nodes[nodeOf[i]].idx();
    vals[i] = i;
  }
}

__global__ void __launch_bounds__(256) k_region_ranges(const uint32_t* __restrict__ sortedKeys, uint32_t n,
                                                       uint32_t* __restrict__ begin, uint32_t* __restrict__ end) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const uint32_t k = sortedKeys[j];
    if (j == 0 || sortedKeys[j - 1] != k) begin[k] = j;
    if (j + 1 == n || sortedKeys[j + 1] != k) end[k] = j + 1;
  }
}

// Gather into region-contiguous SoA, each sample reprojected to its region's
// pivot (gather_sample).  Workgroup b runs on XCD b % 8 (round-robin
// dispatch); it takes range (b % 8) * (G / 8) + b / 8 of G contiguous ranges
// of `chunk` positions, so each XCD walks one contiguous eighth of the output
// and the records of a region (a permutation of one contiguous range when
// the records are in region order) are read into one XCD's L2 only.
__global__ void __launch_bounds__(256) k_gather_reproject(const SampleAoS* __restrict__ in,
                                                          const uint32_t* __restrict__ sortedKeys,
                                                          const uint32_t* __restrict__ sortedIdx, uint32_t n,
                                                          uint32_t chunk,
                                                          const float4* __restrict__ regionPivot /* xyz, minDist */,
                                                          float* __restrict__ odx, float* __restrict__ ody,
                                                          float* __restrict__ odz, float* __restrict__ ow,
                                                          float* __restrict__ opdf, float* __restrict__ odist,
                                                          int reproject) {
  const uint32_t G = gridDim.x;  // a multiple of 8
  const uint32_t range = (blockIdx.x & 7u) * (G >> 3) + (blockIdx.x >> 3);
  const uint64_t begin = (uint64_t)range * chunk;
  const uint64_t end = min<uint64_t>((uint64_t)n, begin + chunk);
  for (uint64_t jj = begin + threadIdx.x; jj < end; jj += 256) {
    const uint32_t j = (uint32_t)jj;
    gather_sample(in[sortedIdx ? sortedIdx[j] : j], regionPivot[sortedKeys[j]], reproject, j, odx, ody, odz, ow,
                  opdf, odist);
  }
}

// Brute-force 8 nearest region pivots per region (exact, ties by index), the
// precomputed neighbourhood of KNearestRegionsSearchTree::buildRegionNeighbours
// (spatial/KNN.h:212-266).
// One wave per region: every lane keeps the 8 nearest of candidates lane,
// lane + 64, ... in (d^2, index) order; 8 wave-wide arg-min rounds over the
// lanes' heads merge them.  Same total order as a serial scan, so the lists
// are identical.
__global__ void __launch_bounds__(256) k_knn8(const float4* __restrict__ pivots, uint32_t numRegions,
                                              uint32_t* __restrict__ nbIds, float4* __restrict__ nbPts) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= numRegions) return;  // uniform over the wave
  const float4 q = pivots[r];
  float bd[8];
  uint32_t bi[8];
  for (int i = 0; i < 8; i++) { bd[i] = 3.402823466e38f; bi[i] = 0xffffffffu; }
  for (uint32_t i = lane; i < numRegions; i += 64) {
    const float4 p = pivots[i];
    const float dx = __fsub_rn(q.x, p.x), dy = __fsub_rn(q.y, p.y), dz = __fsub_rn(q.z, p.z);
    const float d = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
    if (d < bd[7] || (d == bd[7] && i < bi[7])) {
      int k = 7;
      while (k > 0 && (d < bd[k - 1] || (d == bd[k - 1] && i < bi[k - 1]))) {
        bd[k] = bd[k - 1];
        bi[k] = bi[k - 1];
        k--;
      }
      bd[k] = d;
      bi[k] = i;
    }
  }
  int head = 0;
  for (int k = 0; k < 8; k++) {
    float d = 3.402823466e38f;
    uint32_t id = 0xffffffffu;
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (j == head) { d = bd[j]; id = bi[j]; }
    float wd = d;
    uint32_t wi = id;
    for (int off = 32; off > 0; off >>= 1) {
      const float od = __shfl_xor(wd, off);
      const uint32_t oi = __shfl_xor(wi, off);
      if (od < wd || (od == wd && oi < wi)) { wd = od; wi = oi; }
    }
    if (wi != 0xffffffffu && wi == id && head < 8) head++;
    if (lane == 0) {
      nbIds[r * 8 + k] = wi;
      nbPts[r * 8 + k] = wi == 0xffffffffu ? make_float4(0.f, 0.f, 0.f, 0.f) : pivots[wi];
    }
  }
}

// ------------------------------------------------------------------ launchers
static inline uint32_t grid_for(uint32_t n, uint32_t block, uint32_t maxBlocks = 4096) {
  uint32_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return g > maxBlocks ? maxBlocks : g;
}

void launch_ingest(const void* aos, uint32_t n, float* px, float* py, float* pz, uint32_t* wkey, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_ingest, dim3(grid_for(n, 256)), dim3(256), 0, s, (const SampleAoS*)aos, n, px, py, pz, wkey);
}

uint32_t launch_bounds(const float* px, const float* py, const float* pz, uint32_t n, float* partial, hipStream_t s) {
  const uint32_t g = grid_for(n, 256, 1024);
  hipLaunchKernelGGL(k_bounds, dim3(g), dim3(256), 0, s, px, py, pz, n, partial);
  return g;
}

// every block's table is one row of the partial tables that k_reduce_partials combines
static uint32_t stats_blocks(uint32_t n) { return grid_for(n, kStatsThreads * kStatsItems, 512); }

size_t spatial_partial_bytes(uint32_t n) { return (size_t)stats_blocks(n) * kSlotTile * sizeof(IntStatsAcc); }

static void run_stats_passes(bool descend, const float* px, const float* py, const float* pz, uint32_t n,
                             const KDNode* nodes, uint32_t numNodes, const int32_t* slotOfNode,
                             const void* leafParams, const FrontierNode* frontier, uint32_t* nodeOf,
                             uint32_t numSlots, IntStatsAcc* partial, IntStatsAcc* out, hipStream_t s) {
  const uint32_t blocks = stats_blocks(n);
  for (uint32_t base = 0; base < numSlots || base == 0; base += kSlotTile) {
    const int tile = (int)std::min<uint32_t>(kSlotTile, numSlots - base);
    // the round kernel reads nodeOf to find its frontier node: write it in the last pass only
    const int writeNodes = base + kSlotTile >= numSlots ? 1 : 0;
    if (descend)
      hipLaunchKernelGGL(k_spatial_stats<true>, dim3(blocks), dim3(kStatsThreads), 0, s, px, py, pz, n, nodes, slotOfNode,
                         (const LeafParams*)leafParams, frontier, nodeOf, writeNodes, (int)base, tile, partial,
                         numNodes);
    else
      hipLaunchKernelGGL(k_spatial_stats<false>, dim3(blocks), dim3(kStatsThreads), 0, s, px, py, pz, n, nodes, slotOfNode,
                         (const LeafParams*)leafParams, frontier, nodeOf, writeNodes, (int)base, tile, partial,
                         numNodes);
    if (tile > 0)
      hipLaunchKernelGGL(k_reduce_partials, dim3(tile), dim3(256), 0, s, partial, blocks, tile, out + base);
    if (numSlots == 0) break;
  }
}

void launch_descend_stats(const float* px, const float* py, const float* pz, uint32_t n, const KDNode* nodes,
                          uint32_t numNodes, const int32_t* leafSlot, const void* params, uint32_t numLeaves,
                          uint32_t* nodeOf, IntStatsAcc* partial, IntStatsAcc* acc, hipStream_t s) {
  run_stats_passes(true, px, py, pz, n, nodes, numNodes, leafSlot, params, nullptr, nodeOf, numLeaves, partial, acc,
                   s);
}

void launch_kd_round(const float* px, const float* py, const float* pz, uint32_t n, uint32_t* nodeOf,
                     const int32_t* frontierSlot, uint32_t numNodes, const FrontierNode* frontier,
                     uint32_t numFrontier, IntStatsAcc* partial, IntStatsAcc* childAcc, hipStream_t s) {
  run_stats_passes(false, px, py, pz, n, nullptr, numNodes, frontierSlot, nullptr, frontier, nodeOf, 2 * numFrontier,
                   partial, childAcc, s);
}

void launch_leaf_keys(const uint32_t* nodeOf, const KDNode* nodes, uint32_t n, uint32_t* keys, uint32_t* vals,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_leaf_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, nodeOf, nodes, n, keys, vals);
}

size_t sort_pairs_temp_bytes(uint32_t n) {
  size_t bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
  return bytes;
}

void sort_pairs(void* temp, size_t tempBytes, const uint32_t* keysIn, uint32_t* keysOut, const uint32_t* valsIn,
                uint32_t* valsOut, uint32_t n, int endBit, hipStream_t s) {
  hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, keysIn, keysOut, valsIn, valsOut, (int)n, 0, endBit, s);
}

void launch_region_ranges(const uint32_t* sortedKeys, uint32_t n, uint32_t* begin, uint32_t* end, hipStream_t s) {
  hipLaunchKernelGGL(k_region_ranges, dim3(grid_for(n, 256)), dim3(256), 0, s, sortedKeys, n, begin, end);
}

// Keys of region-contiguous records: key[i] = region r with begin[r] <= i < begin[r+1], value[i] = i.
__global__ void __launch_bounds__(256) k_region_keys(const uint32_t* __restrict__ begin, uint32_t numRegions,
                                                     uint32_t n, uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = numRegions;  // last r with begin[r] <= i
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (begin[mid] <= i) lo = mid; else hi = mid;
  }
  keys[i] = lo;
  vals[i] = i;
}

void launch_region_keys(const uint32_t* begin, uint32_t numRegions, uint32_t n, uint32_t* keys, uint32_t* vals,
                        hipStream_t s) {
  if (n == 0 || numRegions == 0) return;
  hipLaunchKernelGGL(k_region_keys, dim3((n + 255) / 256), dim3(256), 0, s, begin, numRegions, n, keys, vals);
}

void launch_gather_reproject(const void* aos, const uint32_t* sortedKeys, const uint32_t* sortedIdx, uint32_t n,
                             const float4* regionPivot, float* dx, float* dy, float* dz, float* w, float* pdf,
                             float* dist, bool reproject, hipStream_t s) {
  if (n == 0) return;
  // G ranges of `chunk` positions (a multiple of 256), G a multiple of 8
  const uint32_t tiles = (n + 255) / 256;
  const uint32_t G0 = std::min<uint32_t>(tiles, 4096);
  const uint32_t tilesPerRange = (tiles + G0 - 1) / G0;
  const uint32_t chunk = tilesPerRange * 256;
  const uint32_t ranges = (n + chunk - 1) / chunk;
  const uint32_t G = (ranges + 7) / 8 * 8;
  hipLaunchKernelGGL(k_gather_reproject, dim3(G), dim3(256), 0, s, (const SampleAoS*)aos, sortedKeys, sortedIdx, n,
                     chunk, regionPivot, dx, dy, dz, w, pdf, dist, reproject ? 1 : 0);
}

void launch_knn8(const float4* pivots, uint32_t numRegions, uint32_t* nbIds, float4* nbPts, hipStream_t s) {
  if (numRegions == 0) return;
  hipLaunchKernelGGL(k_knn8, dim3((numRegions + 3) / 4), dim3(256), 0, s, pivots, numRegions, nbIds, nbPts);
}

}  // namespace pglgpu
