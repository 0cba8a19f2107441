// Directional training of Field::Update on gfx950: the whole adaptive
// split-and-merge update of a region — weighted EM (E-step over the region's
// sample stream, MAP M-step), chi-square split statistics, recursive/partial
// splitting with partial re-fit, pairwise chi-square merging and the
// harmonic-mean parallax distances — runs on the device with no host round
// trip.  All regions of an update run in one persistent launch.
//
// Teams: a region is worked on by a team of T workgroups (T = 1 for small
// regions).  Its samples are cut into fixed chunks of kChunk samples; every
// sample pass computes one partial per chunk (team rank r takes chunks r,
// r+T, ...), the partials are folded in chunk order with Kahan compensation
// (the log-likelihood in FP64), and every rank of the team then runs the same
// per-region logic on identical LDS state.  Because the chunking and the fold
// order do not depend on T, the result is bit-identical for every team size
// and schedule.  Ranks exchange partials through global memory (double
// buffered) and meet at a counter barrier with a bounded wait.  Work items
// (region, rank) are taken from a queue in order, with every rank of a team
// queued consecutively, by a grid that is co-resident on the device.
//
// Work layout inside a chunk: a quad of 4 lanes evaluates one sample; lane q
// of the quad owns mixture components k = 4j + q (j < 8), so a wave evaluates
// 16 samples x 32 components per step with the component parameters held in
// registers.  The per-sample pdf is a 2-step quad reduction; per-lane sums over
// the chunk are reduced over quads (xor-shuffles) and waves (fixed order).
// Per-component M-step / split / merge logic runs in wave 0 with one lane per
// component, serial decisions in lane 0.  The mean-cosine -> kappa inversion
// and the log-likelihood convergence test are evaluated in FP64; everything
// else is FP32.
//
// Reference behaviour followed (directional/vmm/, Open PGL 0.6.0):
//   AdaptiveSplitandMergeFactory.h          fit 303-347, update 352-462
//   ParallaxAwareVonMisesFisherWeightedEMFactory.h
//     SufficientStatistics 93-593, InitUniformVMM 632-700, updateMixture 751-795,
//     partialUpdateMixture 798-844, weightedExpectationStep 847-892, MAP 895-999,
//     partial MAP 1015-1148, component distances 1199-1326
//   ParallaxAwareVonMisesFisherMixture.h    softAssignment 697-733, split/merge 165-289,
//     performRelativeParallaxShift 809-831, normalisation 852-876, kappa maps 910-928
//   VMMChiSquareComponentSplitter.h         PerformRecursiveSplitting 293-356,
//     UpdateSplitStatistics 439-523, SplitComponent 527-665, mergeComponentStats 994-1100
//   VMMChiSquareComponentMerger.h           PerformMerging 67-89, CalculateMergeCost 93-177,
//     ThresholdedMergeNext 218-279, _IntegratedDivision/_Product 338-396
//   field/Field.h                           updateRegions 441-511 (decay, parallax shift)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "common.h"
#include "kernels.h"

#ifndef PGL_REGION_OCC
#define PGL_REGION_OCC 3
#endif
#if PGL_REGION_OCC == 3
#define PGL_REGION_KERNEL k_region_update
// The split statistics pass uses all 168 VGPRs; called as a function it
// saves the 33 callee-saved registers among them to scratch per lane on every
// call, inlined into the update path it spills nothing.  (The EM loop, which
// saves 47, runs slower inlined and stays a call.)  The fit path (first update
// of a region) keeps its calls.
#define PGL_REGION_CALL __forceinline__
#else
#define PGL_REGION_KERNEL k_region_update_wide
#define PGL_REGION_CALL
#endif

namespace pglgpu {

#define RT 256
#define RWAVES 4
#define K_ESTEP_COPY 32
// value layout of a pass result: field f of component k at f * 32 + k, scalars from kScal
static const int kScal = 192;
static const float kTwoPi = 6.283185307179586f;

struct RegionSamples {
  const float* __restrict__ dx;
  const float* __restrict__ dy;
  const float* __restrict__ dz;
  const float* __restrict__ w;
  const float* __restrict__ pdf;
  const float* __restrict__ dist;
};

// Team context of a workgroup (kept in LDS; thread 0 advances the epoch).
struct Team {
  uint32_t rank, T, nc;
  float* part;          // 2 x nc x kPartStride floats of this region
  uint32_t* counter;    // barrier counter of this region
  uint32_t epoch;       // barriers passed
  int* err;
  unsigned long long timeout;  // longest barrier wait, wall-clock ticks
};

struct Shared {
  RegionState S;
  SuffStats cur;
  SuffStats tmp;
  SplitStatsState lsplit;
  float red[RWAVES][32][8];
  alignas(16) float frame[32][8];  // split statistics: tangent frame of each component (x axis, y axis, 2 unused)
  float cost[32 * 32];
  float selfp[32][6];      // merging: product of each component with itself (direction, kappa, norm, scale)
  float term[64][3];       // merging: the three cost terms of up to 64 pairs
  float pre[10];           // merging: weight, direction and mean cosine of the chosen pair before the merge
  float scal[RWAVES][8];
  float stage[5][kChunk];  // sample pass: the current chunk's dx, dy, dz, w and pdf or distance
  uint32_t stageKey;       // first sample (absolute index) of the chunk held in stage[0..3], ~0: none
  uint32_t stageCnt;       // its samples
  const float* stageExtra; // the array whose values stage[4] holds for that chunk (null: none)
  double ll;
  double prevLL, invPrevLL;
  float un[4];
  float nAssigned;
  uint32_t mask;
  int32_t converged;
  int32_t numSplits;
  int32_t found, ci, cj;
  int32_t order[32];
  int32_t W;  // SIMD width of the emulated reference device
  float fin[kPartStride];  // folded pass results (layout: see the pass functions)
  double finLL;            // folded log-likelihood
  int32_t abort;           // a team barrier timed out: finish without waiting
  uint32_t item;           // queue item of this workgroup
  FactoryConfig cfg;       // launch constants, read from LDS by the region logic
  RegionSamples rs;
  Team tm;                 // this workgroup's place in its region's team
  uint32_t jobRegion, jobIndex;
};


// ----------------------------------------------------------------- helpers
__device__ __forceinline__ float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
  return fmaf(ax, bx, fmaf(ay, by, az * bz));
}

// full-wave sum with the result of lane 0 broadcast (identical on all lanes)
__device__ __forceinline__ float wave_sum(float v) {
  v += __shfl_xor(v, 32);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 8);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 1);
  return __shfl(v, 0);
}

__device__ __forceinline__ float mc_to_kappa(float r) {  // (3r - r^3) / (1 - r^2) in FP64
  const double rd = r;
  const double r2 = rd * rd;
  return (float)((rd * 3.0 - rd * r2) / (1.0 - r2));
}

__device__ __forceinline__ float kappa_to_mc(float kappa) {  // coth(k) - 1/k
  if (!(kappa > 0.f)) return 0.f;
  return 1.f / tanhf(kappa) - 1.f / kappa;
}

// embree::frame(N): x axis from the larger of (0,Nz,-Ny) / (-Nz,0,Nx), y = N x x
__device__ __forceinline__ void frame_of(float nx, float ny, float nz, float& xx, float& xy, float& xz, float& yx,
                                         float& yy, float& yz) {
  const float a0x = 0.f, a0y = nz, a0z = -ny;
  const float a1x = -nz, a1y = 0.f, a1z = nx;
  const bool first = dot3f(a0x, a0y, a0z, a0x, a0y, a0z) > dot3f(a1x, a1y, a1z, a1x, a1y, a1z);
  float dx = first ? a0x : a1x, dy = first ? a0y : a1y, dz = first ? a0z : a1z;
  float l = 1.f / sqrtf(dot3f(dx, dy, dz, dx, dy, dz));
  xx = dx * l;
  xy = dy * l;
  xz = dz * l;
  const float cx = ny * xz - nz * xy, cy = nz * xx - nx * xz, cz = nx * xy - ny * xx;
  l = 1.f / sqrtf(dot3f(cx, cy, cz, cx, cy, cz));
  yx = cx * l;
  yy = cy * l;
  yz = cz * l;
}

__device__ void calc_normalization_lane(VMMParams& V, int k) {  // PAVMM:852-863
  const float e = expf(-2.f * V.kappa[k]);
  V.em2k[k] = e;
  V.norm[k] = V.kappa[k] > 0.f ? V.kappa[k] / (kTwoPi * (1.f - e)) : kOneOverFourPi;
}

__device__ void clear_comp(VMMParams& V, int k) {  // PAVMM:276-289
  V.w[k] = 0.f; V.kappa[k] = 0.f; V.em2k[k] = 1.f; V.mc[k] = 0.f; V.norm[k] = kOneOverFourPi;
  V.mux[k] = 0.f; V.muy[k] = 0.f; V.muz[k] = 1.f; V.dist[k] = 0.f;
}

__device__ __forceinline__ void suff_zero_slot(SuffStats& s, int k) {
  s.swdx[k] = 0.f; s.swdy[k] = 0.f; s.swdz[k] = 0.f; s.sw[k] = 0.f; s.sdw[k] = 0.f;
}
__device__ __forceinline__ void split_zero_slot(SplitStatsState& s, int k) {
  s.chi[k] = 0.f; s.covx[k] = 0.f; s.covy[k] = 0.f; s.covz[k] = 0.f; s.sn[k] = 0.f; s.ssw[k] = 0.f; s.ssa[k] = 0.f;
}

__device__ __forceinline__ int vec_lim(int W, int K) { int l = ((K + W - 1) / W) * W; return l > 32 ? 32 : l; }

// ------------------------------------------------------------ sample passes
// Loads this lane's component parameters (k = 4j + q) into registers.
struct CompRegs {
  float x[8], y[8], z[8], kap2[8], wn[8];  // kap2 = kappa log2(e)
};

__device__ __forceinline__ void load_comps(const VMMParams& V, int q, CompRegs& c) {
  const int K = V.K;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int k = 4 * j + q;
    const bool a = k < K;
    c.x[j] = a ? V.mux[k] : 0.f;
    c.y[j] = a ? V.muy[k] : 0.f;
    c.z[j] = a ? V.muz[k] : 1.f;
    c.kap2[j] = a ? V.kappa[k] * 1.44269504088896341f : 0.f;
    c.wn[j] = a ? V.w[k] * V.norm[k] : 0.f;
  }
}

// sum over the 4 lanes of a quad (DPP quad permutes: lane ^ 1, then lane ^ 2);
// every lane of the quad gets the same bits, as IEEE addition commutes
__device__ __forceinline__ float quad_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
  return v;
}

// Calls f with std::integral_constant<int, JT> for the smallest even JT >= J
// (J = component groups in use), so that the sample loops are unrolled over
// a fixed number of groups with no branch between components; the groups in
// [J, JT) hold zero-weight components.
template <class F>
__device__ __forceinline__ void with_groups(int J, F&& f) {
  if (J > 6) f(std::integral_constant<int, 8>());
  else if (J > 4) f(std::integral_constant<int, 6>());
  else if (J > 2) f(std::integral_constant<int, 4>());
  else f(std::integral_constant<int, 2>());
}

// soft_eval for the first JT component groups of this lane (e[j] = 0 for
// j >= JT), with no branch between the components
template <int JT>
__device__ __forceinline__ float soft_eval_t(const CompRegs& c, float dx, float dy, float dz, float e[8]) {
  float part = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (j < JT) {
      const float ct = dot3f(dx, dy, dz, c.x[j], c.y[j], c.z[j]);
      e[j] = c.wn[j] * __builtin_amdgcn_exp2f(fminf(fmaf(c.kap2[j], ct, -c.kap2[j]), 0.f));
      part += e[j];
    } else {
      e[j] = 0.f;
    }
  }
  return quad_sum(part);
}

// e_k = w_k norm_k exp(kappa_k min(mu_k.d - 1, 0)), evaluated as
// w_k norm_k 2^min(kap2_k mu_k.d - kap2_k, 0); returns the quad-wide pdf
__device__ __forceinline__ float soft_eval(const CompRegs& c, int J, float dx, float dy, float dz, float e[8]) {
  float part = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (j < J) {
      const float ct = dot3f(dx, dy, dz, c.x[j], c.y[j], c.z[j]);
      e[j] = c.wn[j] * __builtin_amdgcn_exp2f(fminf(fmaf(c.kap2[j], ct, -c.kap2[j]), 0.f));
      part += e[j];
    } else {
      e[j] = 0.f;
    }
  }
  return quad_sum(part);
}

// Cross-lane sums without LDS traffic: DPP row rotations inside each row of
// 16 lanes, then the gfx950 row swaps (permlane16_swap exchanges the odd rows
// of its first operand with the even rows of its second, permlane32_swap the
// upper half of the first with the lower half of the second; with both
// operands equal, the two results added give lane i the sum with lane i ^ 16
// or i ^ 32).
#define PGL_ROW_ROR(v, n) __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + (n), 0xF, 0xF, false))
__device__ __forceinline__ float cross_row_sum(float v) {
  const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// sum over the 16 lanes of a wave with the same lane % 4 (the quads' lane q);
// lanes 0..3 hold the sums of q = 0..3
__device__ __forceinline__ float quad_group_sum(float v) {
  v += PGL_ROW_ROR(v, 4);
  v += PGL_ROW_ROR(v, 8);
  return cross_row_sum(v);
}

// sum over the 64 lanes of a wave, valid in lane 0
__device__ __forceinline__ float wave_sum0(float v) {
  v += PGL_ROW_ROR(v, 1);
  v += PGL_ROW_ROR(v, 2);
  v += PGL_ROW_ROR(v, 4);
  v += PGL_ROW_ROR(v, 8);
  return cross_row_sum(v);
}

// Kahan step of the cross-chunk fold
__device__ __forceinline__ void kfold(float& sum, float& comp, float x) {
  const float y = x - comp;
  const float t = sum + y;
  comp = (t - sum) - y;
  sum = t;
}

// value v of the current chunk from the per-wave partials (fixed order)
__device__ __forceinline__ float chunk_value(const Shared& sh, int v, int F, int NS) {
  if (v < 32 * F) {
    const int f = v >> 5, k = v & 31;
    return (sh.red[0][k][f] + sh.red[1][k][f]) + (sh.red[2][k][f] + sh.red[3][k][f]);
  }
  if (v >= kScal && v < kScal + NS) {
    const int i = v - kScal;
    return (sh.scal[0][i] + sh.scal[1][i]) + (sh.scal[2][i] + sh.scal[3][i]);
  }
  return 0.f;
}

// Team barrier.  Partials are written with device-coherent (agent-scope
// relaxed atomic, write-through) stores and read with device-coherent loads,
// so no cache writeback or invalidation is needed: each wave waits for its
// stores to complete (s_waitcnt), the workgroup arrives on the counter and
// waits until all T ranks have arrived.  The wait is bounded: after
// tm.timeout ticks of the constant-rate wall clock, or as soon as any
// workgroup has raised the error flag, the workgroup raises the flag and stops
// waiting at this and every later barrier (it still arrives, so no teammate
// waits for it).  The host turns the flag into an error.
__device__ void team_barrier(Shared& sh, Team& tm) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    tm.epoch++;  // read by the other threads after the closing barrier
    const unsigned long long tb = wall_clock64();
    atomicAdd(tm.counter, 1u);
    if (!sh.abort) {
      const uint32_t target = tm.epoch * tm.T;
      const unsigned long long t0 = tb;
      while (__hip_atomic_load(tm.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(2);
        if (__hip_atomic_load(tm.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            wall_clock64() - t0 > tm.timeout) {
          atomicExch(tm.err, 1);
          sh.abort = 1;
          break;
        }
      }
    }
    sh.S.prof[4] += (float)(wall_clock64() - tb);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  __syncthreads();
}

// End of a chunk: value v = threadIdx.x of chunk c goes into the running fold
// (T = 1, chunks arrive in order) or to the team's partial buffer.
__device__ __forceinline__ void chunk_out(const Shared& sh, const Team& tm, uint32_t c, int F, int NS, int llIdx,
                                          float& fs, float& fc, double& fd) {
  const int v = threadIdx.x;
  const float x = chunk_value(sh, v, F, NS);
  if (tm.T == 1) {
    if (v == llIdx) fd += (double)x;
    else kfold(fs, fc, x);
  } else {
    __hip_atomic_store(tm.part + ((size_t)(tm.epoch & 1u) * tm.nc + c) * kPartStride + v, x, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}

// End of a pass: fold the chunk partials of all ranks in chunk order, publish in sh.fin.
__device__ void pass_finish(Shared& sh, Team& tm, int llIdx, float fs, float fc, double fd) {
  const int v = threadIdx.x;
  if (tm.T > 1) {
    const uint32_t buf = tm.epoch & 1u;
    team_barrier(sh, tm);
    const unsigned long long tF = wall_clock64();
    const float* p = tm.part + (size_t)buf * tm.nc * kPartStride + v;
    // eight loads in flight at a time, folded in chunk order
    for (uint32_t c0 = 0; c0 < tm.nc; c0 += 8) {
      float x[8];
#pragma unroll
      for (int u = 0; u < 8; u++)
        x[u] = c0 + u < tm.nc ? __hip_atomic_load(p + (size_t)(c0 + u) * kPartStride, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : 0.f;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (c0 + u < tm.nc) {
          if (v == llIdx) fd += (double)x[u];
          else kfold(fs, fc, x[u]);
        }
      }
    }
    if (v == 0) sh.S.prof[10] += (float)(wall_clock64() - tF);
  }
  sh.fin[v] = fs;
  if (v == llIdx) sh.finLL = fd;
  __syncthreads();
}

// Copies samples [first, first + cnt) of the region's stream into the
// staging arrays (dx, dy, dz, w, and pdf or distance when extra is set):
// every thread issues all of its coalesced loads at once, so a chunk costs
// one memory latency instead of one per sample step.  A chunk already
// staged (every pass reads the same samples, and a rank takes the same
// chunks in every pass) is not loaded again; the extra array is loaded when
// the stage holds another one.  Callers synchronise before reading the
// stage and then record what it holds with stage_mark.
__device__ __forceinline__ void stage_chunk(Shared& sh, const RegionSamples& rs, uint32_t first, uint32_t cnt,
                                            const float* __restrict__ extra) {
  const int tid = threadIdx.x;
  const bool have = sh.stageKey == first && sh.stageCnt == cnt;
  const bool loadMain = !have;
  const bool loadExtra = extra != nullptr && !(have && sh.stageExtra == extra);
  if (!loadMain && !loadExtra) return;
  float v[kChunk / RT][5];
#pragma unroll
  for (int m = 0; m < (int)(kChunk / RT); m++) {
    const uint32_t o = tid + m * RT;
    if (o < cnt) {
      const uint32_t i = first + o;
      if (loadMain) {
        v[m][0] = rs.dx[i];
        v[m][1] = rs.dy[i];
        v[m][2] = rs.dz[i];
        v[m][3] = rs.w[i];
      }
      if (loadExtra) v[m][4] = extra[i];
    }
  }
#pragma unroll
  for (int m = 0; m < (int)(kChunk / RT); m++) {
    const uint32_t o = tid + m * RT;
    if (o < cnt) {
      if (loadMain) {
#pragma unroll
        for (int a = 0; a < 4; a++) sh.stage[a][o] = v[m][a];
      }
      if (loadExtra) sh.stage[4][o] = v[m][4];
    }
  }
}

// What the stage holds after stage_chunk (thread 0, after the barrier that
// follows stage_chunk; every thread read the previous values before it).
__device__ __forceinline__ void stage_mark(Shared& sh, uint32_t first, uint32_t cnt, const float* extra) {
  if (threadIdx.x != 0) return;
  const bool have = sh.stageKey == first && sh.stageCnt == cnt;
  if (!have) sh.stageExtra = extra;
  else if (extra) sh.stageExtra = extra;
  sh.stageKey = first;
  sh.stageCnt = cnt;
}

// This rank's chunks of a pass: rank, rank + T, ...; with a team of more
// than one rank they are taken in reverse order when the last of them is
// still staged from the previous pass (partials are folded in chunk order
// from the team buffer, so the order does not change any result; a lone
// rank folds as it goes and keeps the order).
struct RankChunks {
  uint32_t count;
  bool reverse;
  __device__ __forceinline__ uint32_t chunk(const Team& tm, uint32_t i) const {
    return tm.rank + (reverse ? count - 1 - i : i) * tm.T;
  }
};
__device__ __forceinline__ RankChunks rank_chunks(const Shared& sh, const Team& tm, uint32_t base) {
  RankChunks r;
  r.count = tm.rank < tm.nc ? (tm.nc - tm.rank + tm.T - 1) / tm.T : 0u;
  r.reverse = tm.T > 1 && r.count > 1 && sh.stageKey == base + (tm.rank + (r.count - 1) * tm.T) * kChunk;
  return r;
}

// Per-lane E-step sums of a chunk.
struct EStepAcc {
  float* bx;
  float* by;
  float* bz;
  float* bw;
  float& ll;
  float& u0;
  float& u1;
  float& u2;
  float& u3;
};

// One sample of the E-step, without branches: a sample the mixture does not
// cover (pdf <= 1e-16) adds zero to the component sums and its weight to the
// unassigned sums (lane q = 0 of the quad, like the log-likelihood).  e and
// pdf come from soft_eval_t.
template <int JT>
__device__ __forceinline__ void estep_accumulate(int q, float dx, float dy, float dz, float w, const float e[8],
                                                 float pdf, EStepAcc& a) {
  const bool ok = pdf > 1e-16f;
  const float wu = (!ok && q == 0) ? w : 0.f;
  a.u0 += wu;
  a.u1 = fmaf(dx, wu, a.u1);
  a.u2 = fmaf(dy, wu, a.u2);
  a.u3 = fmaf(dz, wu, a.u3);
  const float lp = fmaf(w, logf(pdf), a.ll);
  a.ll = (ok && q == 0) ? lp : a.ll;
  const float r = ok ? w * __builtin_amdgcn_rcpf(pdf) : 0.f;
  const float rx = r * dx, ry = r * dy, rz = r * dz;
#pragma unroll
  for (int j = 0; j < JT; j++) {
    a.bx[j] = fmaf(e[j], rx, a.bx[j]);
    a.by[j] = fmaf(e[j], ry, a.by[j]);
    a.bz[j] = fmaf(e[j], rz, a.bz[j]);
    a.bw[j] = fmaf(e[j], r, a.bw[j]);
  }
}

// The staged chunk's samples s = quad, quad + 64, ... in that order.  Two
// samples per step: both are evaluated before either is accumulated, so the
// two exp/dot chains overlap; the sums still take the samples one after the
// other, in sample order.
template <int JT>
__device__ __forceinline__ void estep_chunk(const Shared& sh, const CompRegs& c, int q, int quad, uint32_t cnt,
                                            EStepAcc& a) {
  uint32_t s = quad;
  for (; s + RT / 4 < cnt; s += RT / 2) {
    const uint32_t t = s + RT / 4;
    const float ax = sh.stage[0][s], ay = sh.stage[1][s], az = sh.stage[2][s], aw = sh.stage[3][s];
    const float bx = sh.stage[0][t], by = sh.stage[1][t], bz = sh.stage[2][t], bw = sh.stage[3][t];
    float ea[8], eb[8];
    const float pa = soft_eval_t<JT>(c, ax, ay, az, ea);
    const float pb = soft_eval_t<JT>(c, bx, by, bz, eb);
    estep_accumulate<JT>(q, ax, ay, az, aw, ea, pa, a);
    estep_accumulate<JT>(q, bx, by, bz, bw, eb, pb, a);
  }
  if (s < cnt) {
    const float ax = sh.stage[0][s], ay = sh.stage[1][s], az = sh.stage[2][s], aw = sh.stage[3][s];
    float ea[8];
    const float pa = soft_eval_t<JT>(c, ax, ay, az, ea);
    estep_accumulate<JT>(q, ax, ay, az, aw, ea, pa, a);
  }
}

// E-step (WEM:847-892).  Result: fin[f*32+k] = sum over samples of
// gamma_k w (dx, dy, dz, 1) for f = 0..3, finLL = sum w log pdf,
// fin[kScal+0..3] = unassigned (w, w dx, w dy, w dz).
__device__ void pass_estep(Shared& sh, Team& tm, const RegionSamples& rs, uint32_t base, uint32_t n) {
  const unsigned long long tPass = wall_clock64();
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, q = tid & 3, quad = tid >> 2;
  const VMMParams& V = sh.S.vmm;
  const int K = V.K;
  const int J = (K + 3) >> 2;
  CompRegs c;
  load_comps(V, q, c);
  float fs = 0.f, fc = 0.f;
  double fd = 0.0;
  const RankChunks rc = rank_chunks(sh, tm, base);
  for (uint32_t ci = 0; ci < rc.count; ci++) {
    const uint32_t ch = rc.chunk(tm, ci);
    const uint32_t cb = ch * kChunk, ce = min(n, cb + kChunk);
    stage_chunk(sh, rs, base + cb, ce - cb, nullptr);
    __syncthreads();
    stage_mark(sh, base + cb, ce - cb, nullptr);
    float bx[8], by[8], bz[8], bw[8];
#pragma unroll
    for (int j = 0; j < 8; j++) bx[j] = by[j] = bz[j] = bw[j] = 0.f;
    float ll = 0.f, u0 = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f;
    EStepAcc a{bx, by, bz, bw, ll, u0, u1, u2, u3};
    with_groups(J, [&](auto jt) { estep_chunk<decltype(jt)::value>(sh, c, q, quad, ce - cb, a); });
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float vx = quad_group_sum(bx[j]);
      const float vy = quad_group_sum(by[j]);
      const float vz = quad_group_sum(bz[j]);
      const float vw = quad_group_sum(bw[j]);
      if (lane < 4) {
        const int k = 4 * j + lane;
        sh.red[wave][k][0] = vx;
        sh.red[wave][k][1] = vy;
        sh.red[wave][k][2] = vz;
        sh.red[wave][k][3] = vw;
      }
    }
    // lanes q != 0 hold zeros for the per-sample scalars
    const float vll = wave_sum0(ll);
    const float v0 = wave_sum0(u0), v1 = wave_sum0(u1), v2 = wave_sum0(u2), v3 = wave_sum0(u3);
    if (lane == 0) {
      sh.scal[wave][0] = vll;
      sh.scal[wave][1] = v0;
      sh.scal[wave][2] = v1;
      sh.scal[wave][3] = v2;
      sh.scal[wave][4] = v3;
    }
    __syncthreads();
    chunk_out(sh, tm, ch, 4, 5, kScal, fs, fc, fd);
    __syncthreads();
  }
  pass_finish(sh, tm, kScal, fs, fc, fd);
  if (tid == 0) {
    sh.S.estepEvals += (float)n * (float)K;
    sh.S.prof[1] += (float)(wall_clock64() - tPass);
    sh.S.prof[8] += 1.f;
  }
}

// Distances pass (WEM:1199-1326).  Result: fin[k] = sum of w g_k (g_k pdf / w_k) / dist,
// fin[32+k] = sum of w g_k (g_k pdf / w_k).
__device__ void pass_distances(Shared& sh, Team& tm, const RegionSamples& rs, uint32_t base, uint32_t n) {
  const unsigned long long tPass = wall_clock64();
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, q = tid & 3, quad = tid >> 2;
  const VMMParams& V = sh.S.vmm;
  const int K = V.K;
  const int J = (K + 3) >> 2;
  CompRegs c;
  load_comps(V, q, c);
  float invw[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int k = 4 * j + q;
    invw[j] = k < K ? 1.f / V.w[k] : 0.f;
  }
  float fs = 0.f, fc = 0.f;
  double fd = 0.0;
  const RankChunks rc = rank_chunks(sh, tm, base);
  for (uint32_t ci = 0; ci < rc.count; ci++) {
    const uint32_t ch = rc.chunk(tm, ci);
    const uint32_t cb = ch * kChunk, ce = min(n, cb + kChunk);
    stage_chunk(sh, rs, base + cb, ce - cb, rs.dist);
    __syncthreads();
    stage_mark(sh, base + cb, ce - cb, rs.dist);
    float ad[8], aw[8];
#pragma unroll
    for (int j = 0; j < 8; j++) ad[j] = aw[j] = 0.f;
    const uint32_t cnt = ce - cb;
    with_groups(J, [&](auto jt) {
      constexpr int JT = decltype(jt)::value;
      for (uint32_t s = quad; s < cnt; s += RT / 4) {
        const float dx = sh.stage[0][s], dy = sh.stage[1][s], dz = sh.stage[2][s], w = sh.stage[3][s];
        float e[8];
        const float pdf = soft_eval_t<JT>(c, dx, dy, dz, e);
        const bool ok = pdf > 1e-16f;  // other samples add zero
        // w g_k (g_k pdf / w_k) = (w / pdf) e_k^2 / w_k; 1 / w_k is applied to the chunk sums
        const float wip = ok ? w * __builtin_amdgcn_rcpf(pdf) : 0.f;
        const float wipd = ok ? wip * __builtin_amdgcn_rcpf(sh.stage[4][s]) : 0.f;
#pragma unroll
        for (int j = 0; j < JT; j++) {
          const float e2 = e[j] * e[j];
          ad[j] = fmaf(e2, wipd, ad[j]);
          aw[j] = fmaf(e2, wip, aw[j]);
        }
      }
    });
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float vd = quad_group_sum(ad[j]) * invw[j];
      const float vw = quad_group_sum(aw[j]) * invw[j];
      if (lane < 4) {
        sh.red[wave][4 * j + lane][0] = vd;
        sh.red[wave][4 * j + lane][1] = vw;
      }
    }
    __syncthreads();
    chunk_out(sh, tm, ch, 2, 0, -1, fs, fc, fd);
    __syncthreads();
  }
  pass_finish(sh, tm, -1, fs, fc, fd);
  if (tid == 0) { sh.S.auxEvals += (float)n * (float)K; sh.S.prof[3] += (float)(wall_clock64() - tPass); }
}

// Split-statistics pass (Splitter:439-523): chi^2 terms, assigned mass and the
// zero-mean tangent-plane covariance in each component's frame.  Result:
// fin[f*32+k] for f = chi, assigned, weight, cov xx, cov yy, cov xy;
// fin[kScal] = number of samples with pdf > 0.
// Per-lane split statistics sums of a chunk.
struct SplitAcc {
  float* chi;
  float* sa;
  float* sw;
  float* cx;
  float* cy;
  float* cz;
  float& cnt;
};

// Split statistics of one sample against component group j (frame f0, f1):
// vmf = e[j], ip = 1 / pdf, cf = chi-square factor, wip = w / pdf.
__device__ __forceinline__ void split_add(SplitAcc& a, int j, float4 f0, float2 f1, float dx, float dy, float dz,
                                          float vmf, float ip, float cf, float wip) {
  a.chi[j] += vmf > 0.f ? vmf * cf : 0.f;
  a.sa[j] = fmaf(vmf, ip, a.sa[j]);
  const float lx = dot3f(f0.x, f0.y, f0.z, dx, dy, dz);
  const float ly = dot3f(f0.w, f1.x, f1.y, dx, dy, dz);
  const float aw = vmf * wip;
  a.sw[j] += aw;
  a.cx[j] = fmaf(aw, lx * lx, a.cx[j]);
  a.cy[j] = fmaf(aw, ly * ly, a.cy[j]);
  a.cz[j] = fmaf(aw, lx * ly, a.cz[j]);
}

// Per-sample factors of the split statistics.
struct SplitFactors {
  float ip, cf, wip;
};
__device__ __forceinline__ SplitFactors split_factors(float pdf, float w, float spdf, float inv_mc) {
  const float ip = __builtin_amdgcn_rcpf(pdf);
  const float value = w * spdf;
  const float av = value * inv_mc * ip;  // value / (mc pdf)
  // (value^2 vmf / (mc pdf)^2 - 2 vmf value / (mc pdf) + vmf) / spdf = vmf cf
  const float cf = (av * av - 2.f * av + 1.f) * __builtin_amdgcn_rcpf(spdf);
  const float wip = w * ip;
  return {ip, cf, wip};
}

// The staged chunk's samples s = quad, quad + 64, ... in that order, for the
// first JT component groups with no branch between the components (groups in
// [J, JT) hold zero-weight components, whose sums are not read).  The frame
// of component k is one 16 B and one 8 B read of sh.frame[k]; the next
// group's frame is read while the current group's sums are formed.
template <int JT>
__device__ __forceinline__ void split_chunk(const Shared& sh, const CompRegs& c, int q, int quad, uint32_t cnt,
                                            float inv_mc, SplitAcc& a) {
  for (uint32_t s = quad; s < cnt; s += RT / 4) {
    const float dx = sh.stage[0][s], dy = sh.stage[1][s], dz = sh.stage[2][s];
    float e[8];
    const float pdf = soft_eval_t<JT>(c, dx, dy, dz, e);
    if (pdf > 1e-16f) {
      const SplitFactors F = split_factors(pdf, sh.stage[3][s], sh.stage[4][s], inv_mc);
      a.cnt += 1.f;
      float4 f0 = *reinterpret_cast<const float4*>(&sh.frame[q][0]);
      float2 f1 = *reinterpret_cast<const float2*>(&sh.frame[q][4]);
#pragma unroll
      for (int j = 0; j < JT; j++) {
        float4 n0 = f0;
        float2 n1 = f1;
        if (j + 1 < JT) {
          n0 = *reinterpret_cast<const float4*>(&sh.frame[4 * (j + 1) + q][0]);
          n1 = *reinterpret_cast<const float2*>(&sh.frame[4 * (j + 1) + q][4]);
        }
        split_add(a, j, f0, f1, dx, dy, dz, e[j], F.ip, F.cf, F.wip);
        f0 = n0;
        f1 = n1;
      }
    }
  }
}

__device__ void pass_split(Shared& sh, Team& tm, const RegionSamples& rs, uint32_t base, uint32_t n, float mcEst) {
  const unsigned long long tPass = wall_clock64();
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, q = tid & 3, quad = tid >> 2;
  const VMMParams& V = sh.S.vmm;
  const int K = V.K;
  const int J = (K + 3) >> 2;
  CompRegs c;
  load_comps(V, q, c);
  const float inv_mc = 1.f / mcEst;
  float fs = 0.f, fc = 0.f;
  double fd = 0.0;
  const RankChunks rc = rank_chunks(sh, tm, base);
  for (uint32_t ci = 0; ci < rc.count; ci++) {
    const uint32_t ch = rc.chunk(tm, ci);
    const uint32_t cb = ch * kChunk, ce = min(n, cb + kChunk);
    stage_chunk(sh, rs, base + cb, ce - cb, rs.pdf);
    __syncthreads();
    stage_mark(sh, base + cb, ce - cb, rs.pdf);
    float achi[8], asa[8], asw[8], acx[8], acy[8], acz[8];
#pragma unroll
    for (int j = 0; j < 8; j++) achi[j] = asa[j] = asw[j] = acx[j] = acy[j] = acz[j] = 0.f;
    float cnt = 0.f;
    SplitAcc a{achi, asa, asw, acx, acy, acz, cnt};
    with_groups(J, [&](auto jt) { split_chunk<decltype(jt)::value>(sh, c, q, quad, ce - cb, inv_mc, a); });
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float v0 = quad_group_sum(achi[j]);
      const float v1 = quad_group_sum(asa[j]);
      const float v2 = quad_group_sum(asw[j]);
      const float v3 = quad_group_sum(acx[j]);
      const float v4 = quad_group_sum(acy[j]);
      const float v5 = quad_group_sum(acz[j]);
      if (lane < 4) {
        const int k = 4 * j + lane;
        sh.red[wave][k][0] = v0;
        sh.red[wave][k][1] = v1;
        sh.red[wave][k][2] = v2;
        sh.red[wave][k][3] = v3;
        sh.red[wave][k][4] = v4;
        sh.red[wave][k][5] = v5;
      }
    }
    // every lane of a quad counted the sample: divide the wave sum by 4
    const float vc = wave_sum0(cnt) * 0.25f;
    if (lane == 0) sh.scal[wave][0] = vc;
    __syncthreads();
    chunk_out(sh, tm, ch, 6, 1, -1, fs, fc, fd);
    __syncthreads();
  }
  pass_finish(sh, tm, -1, fs, fc, fd);
  if (tid == 0) { sh.S.splitEvals += (float)n * (float)K; sh.S.prof[2] += (float)(wall_clock64() - tPass); }
}

__device__ __forceinline__ float fin(const Shared& sh, int k, int f) { return sh.fin[f * 32 + k]; }

// ------------------------------------------------------ region-level steps
// UpdateSplitStatistics into SS (S.split or the local recursive-split copy).
__device__ PGL_REGION_CALL void update_split_stats(Shared& sh, Team& tm, SplitStatsState& SS,
                                                   const RegionSamples& rs, uint32_t base, uint32_t n, float mcEst) {
  const int tid = threadIdx.x;
  const VMMParams& V = sh.S.vmm;
  if (tid < 32) {
    const int k = tid;
    float xx, xy, xz, yx, yy, yz;
    frame_of(V.mux[k], V.muy[k], V.muz[k], xx, xy, xz, yx, yy, yz);
    // components k >= K have zero weight: a zero frame keeps their unread sums finite
    const bool a = k < V.K;
    float* f = sh.frame[k];
    f[0] = a ? xx : 0.f; f[1] = a ? xy : 0.f; f[2] = a ? xz : 0.f;
    f[3] = a ? yx : 0.f; f[4] = a ? yy : 0.f; f[5] = a ? yz : 0.f;
  }
  __syncthreads();
  pass_split(sh, tm, rs, base, n, mcEst);
  if (tid < 32) {
    const int k = tid;
    const float nA = sh.fin[kScal];
    const int lim = vec_lim(sh.W, SS.K);
    if (k < SS.K) {
      const float nOld = SS.sn[k];
      const float nNew = nOld + nA;
      SS.ssa[k] += fin(sh, k, 1);
      SS.ssw[k] += fin(sh, k, 2);
      SS.covx[k] += fin(sh, k, 3);
      SS.covy[k] += fin(sh, k, 4);
      SS.covz[k] += fin(sh, k, 5);
      // running mean chi += (x - chi)/count over the assigned samples, in closed form
      if (nA > 0.f) SS.chi[k] = (SS.chi[k] * nOld + fin(sh, k, 0)) / nNew;
      SS.sn[k] = nNew;
    } else if (k < lim) {
      SS.sn[k] += nA;
    }
  }
  __syncthreads();
}

__device__ void distances(Shared& sh, Team& tm, SuffStats& ST, const RegionSamples& rs, uint32_t base, uint32_t n,
                          bool init) {
  pass_distances(sh, tm, rs, base, n);
  const int tid = threadIdx.x;
  VMMParams& V = sh.S.vmm;
  if (tid < 32) {
    const int k = tid;
    const int lim = vec_lim(sh.W, V.K);
    if (k < V.K) {
      const float bd = fin(sh, k, 0), bw = fin(sh, k, 1);
      if (init) {
        ST.sdw[k] = bw;
        V.dist[k] = ST.sdw[k] / bd;
      } else {
        float sumInv = bd;
        sumInv += V.dist[k] > 0.f ? (ST.sdw[k] / V.dist[k]) : 0.f;
        ST.sdw[k] += bw;
        V.dist[k] = ST.sdw[k] / sumInv;
      }
    } else if (k < lim) {
      V.dist[k] = 0.f;
      ST.sdw[k] = 0.f;
    }
  }
  __syncthreads();
}

// EM loop (updateMixture / partialUpdateMixture). prev = region statistics or
// the partial-refit temporary; masked = partial fit over sh.mask.
__device__ int em(Shared& sh, Team& tm, SuffStats& prev, const RegionSamples& rs, uint32_t base,
                                  uint32_t n, const FactoryConfig& cfg, bool masked) {
  const int tid = threadIdx.x;
  VMMParams& V = sh.S.vmm;
  SuffStats& cur = sh.cur;
  if (tid < 32) suff_zero_slot(cur, tid);
  if (tid == 0) {
    cur.K = 32; cur.sumWeights = 0.f; cur.numSamples = 0.f; cur.overall = 0.f; cur.normalized = 0;
    sh.converged = 0; sh.prevLL = 0.0; sh.invPrevLL = 1.0;
  }
  __syncthreads();
  int it = 0;
  while (!sh.converged && it < cfg.maxEMIterations) {
    if (tid < 32) suff_zero_slot(cur, tid);
    if (tid == 0) { cur.K = V.K; cur.numSamples = (float)n; cur.sumWeights = 0.f; cur.normalized = 0; }
    __syncthreads();
    pass_estep(sh, tm, rs, base, n);
    const unsigned long long tM = wall_clock64();
    if (tid < K_ESTEP_COPY) {
      const int k = tid;
      if (k < V.K) {
        cur.swdx[k] = fin(sh, k, 0); cur.swdy[k] = fin(sh, k, 1); cur.swdz[k] = fin(sh, k, 2); cur.sw[k] = fin(sh, k, 3);
      }
    }
    if (tid == 0) sh.ll = sh.finLL;
    __syncthreads();
    // unassigned mass becomes a new component (WEM:732-748)
    const float un0 = sh.fin[kScal + 1];
    if (tid == 0 && un0 > 0.f && cur.K < kMaxComponents) {
      const int k = cur.K;
      cur.K++;
      cur.sw[k] = un0; cur.swdx[k] = sh.fin[kScal + 2]; cur.swdy[k] = sh.fin[kScal + 3]; cur.swdz[k] = sh.fin[kScal + 4];
      prev.K++;
      prev.sw[k] = 0.f; prev.swdx[k] = 0.f; prev.swdy[k] = 0.f; prev.swdz[k] = 0.f;
      V.K++;
      if (masked) sh.mask |= 1u << (V.K - 1);
    }
    __syncthreads();
    if (tid < 64) {
      const int k = tid;
      const int K = V.K;
      const bool act = k < K;
      // normalize(numSamples) (WEM:553-572)
      const float tot = wave_sum(act ? cur.sw[k] : 0.f);
      const float ns = cur.numSamples;
      const float nrm = ns / tot;
      if (act) { cur.swdx[k] *= nrm; cur.swdy[k] *= nrm; cur.swdz[k] *= nrm; cur.sw[k] *= nrm; }
      // MAP weights (WEM:895-927 / 1030-1078)
      const float wp = cfg.weightPrior;
      const float nsum = cur.numSamples + prev.numSamples;
      float wt = 0.f;
      if (act) wt = (wp + (cur.sw[k] + prev.sw[k])) / ((wp * (float)K) + nsum);
      const bool m = masked ? (act && ((sh.mask >> k) & 1u)) : act;
      if (masked) {
        const float sumP = wave_sum(m ? wt : 0.f);
        const float sumW = wave_sum((act && !m) ? V.w[k] : 0.f);
        float inv = 1.f / sumP;
        inv *= 1.f - sumW;
        if (k < 32) V.w[k] = m ? wt * inv : (act ? V.w[k] : 0.f);
      } else if (k < 32) {
        V.w[k] = wt;
      }
      // MAP mean direction and concentration (WEM:930-999 / 1081-1148)
      if (k < 32) {
        const float cn = cur.numSamples, pn = prev.numSamples;
        const float nss = cn + pn;
        const float overall = cur.numSamples + prev.overall;
        const float ce = cn / nss;
        const float pe = 1.f - ce;
        if (act) {
          const float partialN = V.w[k] * overall;
          const float cmx = cur.sw[k] > 0.f ? cur.swdx[k] / cur.sw[k] : 0.f;
          const float cmy = cur.sw[k] > 0.f ? cur.swdy[k] / cur.sw[k] : 0.f;
          const float cmz = cur.sw[k] > 0.f ? cur.swdz[k] / cur.sw[k] : 0.f;
          const float pmx = prev.sw[k] > 0.f ? prev.swdx[k] / prev.sw[k] : 0.f;
          const float pmy = prev.sw[k] > 0.f ? prev.swdy[k] / prev.sw[k] : 0.f;
          const float pmz = prev.sw[k] > 0.f ? prev.swdz[k] / prev.sw[k] : 0.f;
          const float mdx = cmx * ce + pmx * pe, mdy = cmy * ce + pmy * pe, mdz = cmz * ce + pmz * pe;
          float mcos = sqrtf(dot3f(mdx, mdy, mdz, mdx, mdy, mdz));
          if (m) {
            if (mcos > 0.f) { V.mux[k] = mdx / mcos; V.muy[k] = mdy / mcos; V.muz[k] = mdz / mcos; }
            mcos = (cfg.meanCosinePrior * cfg.meanCosinePriorStrength + mcos * partialN) /
                   (cfg.meanCosinePriorStrength + partialN);
            mcos = fminf(cfg.maxMeanCosine, mcos);
            V.mc[k] = mcos;
            V.kappa[k] = mc_to_kappa(mcos);
          }
        } else {
          V.mux[k] = 0.f; V.muy[k] = 0.f; V.muz[k] = 1.f; V.mc[k] = 0.f; V.kappa[k] = 0.f;
        }
        calc_normalization_lane(V, k);
      }
      if (tid == 0) { cur.numSamples = ns; cur.sumWeights = tot; cur.normalized = 1; }
    }
    __syncthreads();
    it++;
    if (tid == 0 && it > 1) {
      const double rel = fabs(sh.ll - sh.prevLL) * sh.invPrevLL;
      if (rel < (double)cfg.convergenceThreshold) sh.converged = 1;
      sh.prevLL = sh.ll;
      sh.invPrevLL = 1.0 / fabs(sh.ll);
    }
    if (tid == 0) sh.S.prof[9] += (float)(wall_clock64() - tM);
    __syncthreads();
  }
  // previousStats += currentStats (WEM:575-593)
  if (tid < 32) {
    const int k = tid;
    prev.swdx[k] += cur.swdx[k]; prev.swdy[k] += cur.swdy[k]; prev.swdz[k] += cur.swdz[k];
    prev.sw[k] += cur.sw[k]; prev.sdw[k] += cur.sdw[k];
  }
  if (tid == 0) {
    prev.sumWeights += cur.sumWeights;
    prev.numSamples += cur.numSamples;
    prev.overall += cur.numSamples;
    sh.S.lastLogLikelihood = (float)sh.prevLL;
  }
  __syncthreads();
  return it;
}

// ---- splitting (serial, lane 0)
__device__ __forceinline__ void exp_map(float vx, float vy, float& ox, float& oy, float& oz) {  // Splitter:175-190
  const float len = sqrtf(vx * vx + vy * vy);
  const float sinc = sinf(len) / len;
  ox = len > 0.f ? vx * sinc : 0.f;
  oy = len > 0.f ? vy * sinc : 0.f;
  oz = cosf(len);
}
__device__ __forceinline__ void log_map(float x, float y, float z, float& ox, float& oy) {  // Splitter:161-172
  const float zc = fminf(1.f, fmaxf(-1.f, z));
  const float alpha = acosf(zc);
  const float invSinc = alpha / sinf(alpha);
  ox = alpha > 0.f ? x * invSinc : 0.f;
  oy = alpha > 0.f ? y * invSinc : 0.f;
}

__device__ bool split_component(VMMParams& V, SplitStatsState& SS, SuffStats& ST, int idx, const FactoryConfig& cfg) {
  const float numAssigned = SS.ssa[idx];
  const float invSw = 1.f / SS.ssw[idx];
  const float cx = SS.covx[idx] * invSw, cy = SS.covy[idx] * invSw, cz = SS.covz[idx] * invSw;
  const float D = sqrtf((cx - cy) * (cx - cy) + (cz * cz * 4.f)) * 0.5f;
  float ev0 = (float)((double)(cx + cy) * 0.5);
  ev0 += D;
  float e0x = -cz, e0y = cx - ev0;
  const float l0 = sqrtf(e0x * e0x + e0y * e0y);
  e0x /= l0;
  e0y /= l0;
  const float weight = V.w[idx], meanCosine = V.mc[idx], kappa = V.kappa[idx];
  if ((double)kappa >= (double)kMaxKappaSplit * 0.9) return false;
  const float mx = V.mux[idx], my = V.muy[idx], mz = V.muz[idx];
  const float nw = weight * 0.5f;
  float m0x = mx, m0y = my, m0z = mz, m1x = mx, m1y = my, m1z = mz;
  float nc0 = meanCosine, nc1 = meanCosine * meanCosine;
  if (D > 1e-8f) {
    float xx, xy, xz, yx, yy, yz;
    frame_of(mx, my, mz, xx, xy, xz, yx, yy, yz);
    const float ax = e0x * ev0 * 0.5f, ay = e0y * ev0 * 0.5f;
    float lx, ly, lz;
    exp_map(ax, ay, lx, ly, lz);
    m0x = xx * lx + yx * ly + mx * lz;
    m0y = xy * lx + yy * ly + my * lz;
    m0z = xz * lx + yz * ly + mz * lz;
    nc0 = meanCosine / fabsf(dot3f(mx, my, mz, m0x, m0y, m0z));
    nc0 = fminf(nc0, cfg.maxMeanCosineSplit);
    nc1 = nc0;
    exp_map(-ax, -ay, lx, ly, lz);
    m1x = xx * lx + yx * ly + mx * lz;
    m1y = xy * lx + yy * ly + my * lz;
    m1z = xz * lx + yz * ly + mz * lz;
  } else if (numAssigned < 2.f) {
    return false;
  }
  const int K = V.K;
  // PAVMM::splitComponent (PAVMM:165-190)
  V.dist[K] = V.dist[idx];
  V.w[idx] = nw; V.mc[idx] = nc0; V.kappa[idx] = mc_to_kappa(nc0);
  V.mux[idx] = m0x; V.muy[idx] = m0y; V.muz[idx] = m0z;
  V.w[K] = nw; V.mc[K] = nc1; V.kappa[K] = mc_to_kappa(nc1);
  V.mux[K] = m1x; V.muy[K] = m1y; V.muz[K] = m1z;
  V.K = K + 1;
  // PAVMM:165-190 renormalises every component here; a component's
  // normalisation depends only on its own kappa and nothing reads it before
  // the split loop ends, so the caller renormalises all of them once after
  // the loop (the same values)
  // SufficientStatistics::splitComponentsStats (WEM:520-550)
  float s = ST.sw[idx];
  s /= 2.f;
  ST.sw[idx] = s;
  ST.swdx[idx] = m0x * nc0 * s; ST.swdy[idx] = m0y * nc0 * s; ST.swdz[idx] = m0z * nc0 * s;
  ST.sw[K] = s;
  ST.swdx[K] = m1x * nc1 * s; ST.swdy[K] = m1y * nc1 * s; ST.swdz[K] = m1z * nc1 * s;
  const float t = ST.sdw[idx] * 0.5f;
  ST.sdw[idx] = t;
  ST.sdw[K] = t;
  ST.K += 1;
  split_zero_slot(SS, idx);
  split_zero_slot(SS, K);
  SS.K = K + 1;
  return true;
}

// candidates in descending chi^2 (stable for equal values)
__device__ void sort_candidates(const SplitStatsState& SS, int* order) {
  const int K = SS.K;
  for (int i = 0; i < K; i++) order[i] = i;
  for (int i = 1; i < K; i++) {
    const int v = order[i];
    const float c = SS.chi[v];
    int j = i - 1;
    while (j >= 0 && SS.chi[order[j]] < c) { order[j + 1] = order[j]; j--; }
    order[j + 1] = v;
  }
}

// The same order on wave 0: lane k's place is the number of candidates with
// a larger value plus the earlier ones with an equal value (thread 0 runs the
// insertion sort instead when a value is NaN).  Callers synchronise after.
__device__ void rank_candidates(Shared& sh, const SplitStatsState& SS) {
  const int tid = threadIdx.x;
  if (tid >= 64) return;
  const int K = SS.K;
  const bool mine = tid < K;
  const float c = mine ? SS.chi[tid] : 0.f;
  const bool anyNaN = __ballot(mine && c != c) != 0;
  if (!anyNaN && mine) {
    int r = 0;
    for (int j = 0; j < K; j++) {
      const float cj = SS.chi[j];
      r += (cj > c || (cj == c && j < tid)) ? 1 : 0;
    }
    sh.order[r] = tid;
  }
  if (anyNaN && tid == 0) sort_candidates(SS, sh.order);
}

// maskedReplace (WEM:448-468)
__device__ void masked_replace(Shared& sh, SuffStats& dst, const SuffStats& src) {
  const int tid = threadIdx.x;
  if (tid < 64) {
    const int k = tid;
    float v = 0.f;
    if (k < 32) {
      if ((sh.mask >> k) & 1u) {
        dst.swdx[k] = src.swdx[k]; dst.swdy[k] = src.swdy[k]; dst.swdz[k] = src.swdz[k];
        dst.sw[k] = src.sw[k]; dst.sdw[k] = src.sdw[k];
      }
      v = dst.sw[k];
    }
    const float sum = wave_sum(v);
    if (tid == 0) { if (dst.normalized) dst.numSamples = sum; else dst.sumWeights = sum; }
  }
  __syncthreads();
}

__device__ void copy_suff(Shared& sh, SuffStats& dst, const SuffStats& src) {
  const int tid = threadIdx.x;
  if (tid < 32) {
    dst.swdx[tid] = src.swdx[tid]; dst.swdy[tid] = src.swdy[tid]; dst.swdz[tid] = src.swdz[tid];
    dst.sw[tid] = src.sw[tid]; dst.sdw[tid] = src.sdw[tid];
  }
  if (tid == 0) {
    dst.sumWeights = src.sumWeights; dst.numSamples = src.numSamples; dst.overall = src.overall;
    dst.K = src.K; dst.normalized = src.normalized;
  }
  __syncthreads();
}

// SufficientStatistics::clear(K) (WEM:372-389)
__device__ void clear_suff(Shared& sh, SuffStats& s, int K) {
  const int tid = threadIdx.x;
  if (tid < vec_lim(sh.W, K)) suff_zero_slot(s, tid);
  if (tid == 0) { s.K = K; s.sumWeights = 0.f; s.numSamples = 0.f; s.normalized = 0; }
  __syncthreads();
}
__device__ void clear_split(Shared& sh, SplitStatsState& s, int K) {
  const int tid = threadIdx.x;
  if (tid < vec_lim(sh.W, K)) split_zero_slot(s, tid);
  if (tid == 0) s.K = K;
  __syncthreads();
}

// PerformRecursiveSplitting (Splitter:293-356)
__device__ int recursive_splitting(Shared& sh, Team& tm, const RegionSamples& rs, uint32_t base, uint32_t n,
                                   float mcEst, const FactoryConfig& cfg, int& pits) {
  const int tid = threadIdx.x;
  VMMParams& V = sh.S.vmm;
  int total = 0;
  copy_suff(sh, sh.tmp, sh.S.suff);
  if (tid == 0) sh.numSplits = -1;
  __syncthreads();
  while (V.K < kMaxComponents && sh.numSplits != 0) {
    clear_split(sh, sh.lsplit, 32);
    clear_split(sh, sh.lsplit, V.K);
    update_split_stats(sh, tm, sh.lsplit, rs, base, n, mcEst);
    const unsigned long long tS = wall_clock64();
    rank_candidates(sh, sh.lsplit);
    __syncthreads();
    if (tid == 0) {
      sh.mask = 0u;
      int ns = 0;
      const int numComp = V.K;
      for (int c = 0; c < numComp; c++) {
        const int idx = sh.order[c];
        if (sh.lsplit.chi[idx] > cfg.splittingThreshold && V.K < kMaxComponents) {
          const bool ok = split_component(V, sh.lsplit, sh.tmp, idx, cfg);
          sh.mask |= 1u << idx;
          sh.mask |= 1u << (V.K - 1);
          if (ok) ns++;
        }
      }
      sh.numSplits = ns;
    }
    __syncthreads();
    if (sh.numSplits > 0) {
      if (tid < 32) calc_normalization_lane(V, tid);
      __syncthreads();
    }
    if (tid == 0) sh.S.prof[6] += (float)(wall_clock64() - tS);
    if (sh.numSplits > 0) {
      clear_suff(sh, sh.tmp, V.K);
      pits += em(sh, tm, sh.tmp, rs, base, n, cfg, true);
      if (tid == 0) sh.S.suff.K = V.K;
      __syncthreads();
      masked_replace(sh, sh.S.suff, sh.tmp);
    }
    total += sh.numSplits;
    __syncthreads();
  }
  return total;
}

// ---- merging
__device__ float product_scale(float m0x, float m0y, float m0z, float k0, float n0, float m1x, float m1y, float m1z,
                               float k1, float n1, float& pmx, float& pmy, float& pmz, float& pk, float& pn) {
  pmx = k0 * m0x + k1 * m1x;
  pmy = k0 * m0y + k1 * m1y;
  pmz = k0 * m0z + k1 * m1z;
  // both outcomes computed and selected: no branch, so that independent
  // chains of the merge step can be interleaved
  const float r = sqrtf(dot3f(pmx, pmy, pmz, pmx, pmy, pmz));
  const bool big = r > 1e-3f;
  const float e = expf(-2.f * r);
  pn = big ? r / (kTwoPi * (1.f - e)) : 1.f / (4.f * 3.14159265358979f);
  pmx = big ? pmx / r : m0x;
  pmy = big ? pmy / r : m0y;
  pmz = big ? pmz / r : m0z;
  pk = big ? r : 0.f;
  float scale = (n0 * n1) / pn;
  const float c0 = dot3f(m0x, m0y, m0z, pmx, pmy, pmz), c1 = dot3f(m1x, m1y, m1z, pmx, pmy, pmz);
  scale *= expf(k0 * (c0 - 1.f) + k1 * (c1 - 1.f));
  return scale;
}

__device__ float integrated_division(float m0x, float m0y, float m0z, float k0, float n0, float m1x, float m1y,
                                     float m1z, float k1, float n1, float em2k1) {
  float pmx = k0 * m0x + k1 * m1x, pmy = k0 * m0y + k1 * m1y, pmz = k0 * m0z + k1 * m1z;
  const float r = sqrtf(dot3f(pmx, pmy, pmz, pmx, pmy, pmz));
  const bool big = r > 1e-3f;  // both outcomes computed and selected (no branch)
  const float e = expf(-2.f * r);
  const float pn = big ? r / (kTwoPi * (1.f - e)) : 1.f / (4.f * 3.14159265358979f);
  pmx = big ? pmx / r : m0x;
  pmy = big ? pmy / r : m0y;
  pmz = big ? pmz / r : m0z;
  const float pk = big ? r : 0.f;
  float scale = (n0 * n1) / pn;
  const float c0 = dot3f(m0x, m0y, m0z, pmx, pmy, pmz), c1 = dot3f(m1x, m1y, m1z, pmx, pmy, pmz);
  scale *= (4.f * 3.14159265358979f * 3.14159265358979f * (1.f - em2k1)) / (k1 * k1);
  scale *= expf((k0 * (c0 - 1.f) + k1 * (c1 - 1.f)) + (2.f * k1));
  return scale;
}

// CalculateMergeCost (Merger:93-177), split into its parts: the
// moment-matched merged lobe of the pair, the product of each component with
// itself (cached per component while merging) and the three integrated
// divisions, each a term of the chi-square estimate.
struct MergedLobe {
  float w, dx, dy, dz, kappa, nrm, em2k, mc;
};

__device__ __forceinline__ MergedLobe merged_lobe_of(float w0, float c0, float m0x, float m0y, float m0z, float w1,
                                                     float c1, float m1x, float m1y, float m1z) {
  MergedLobe m;
  m.w = w0 + w1;
  const float dx = (w0 * c0 * m0x + w1 * c1 * m1x) / m.w;
  const float dy = (w0 * c0 * m0y + w1 * c1 * m1y) / m.w;
  const float dz = (w0 * c0 * m0z + w1 * c1 * m1z) / m.w;
  const float mcos = dx * dx + dy * dy + dz * dz;
  // both outcomes computed and selected (no branch)
  const bool pos = mcos > 0.f;
  const float r = sqrtf(mcos);
  float kap = mc_to_kappa(r);
  kap = kap < 1e-3f ? 0.f : kap;
  const float e = expf(-2.f * kap);
  m.mc = pos ? r : mcos;
  m.kappa = pos ? kap : 0.f;
  m.em2k = pos ? e : 1.f;
  m.nrm = pos ? kap / (kTwoPi * (1.f - e)) : kOneOverFourPi;
  m.dx = pos ? dx / r : m0x;
  m.dy = pos ? dy / r : m0y;
  m.dz = pos ? dz / r : m0z;
  return m;
}

__device__ __forceinline__ MergedLobe merged_lobe(const VMMParams& V, int i0, int i1) {
  return merged_lobe_of(V.w[i0], V.mc[i0], V.mux[i0], V.muy[i0], V.muz[i0], V.w[i1], V.mc[i1], V.mux[i1], V.muy[i1],
                        V.muz[i1]);
}

// product of component k with itself: {direction x, y, z, kappa, norm, scale}
__device__ __forceinline__ void self_product(const VMMParams& V, int k, float* out) {
  const float mx = V.mux[k], my = V.muy[k], mz = V.muz[k], kk = V.kappa[k], nn = V.norm[k];
  out[5] = product_scale(mx, my, mz, kk, nn, mx, my, mz, kk, nn, out[0], out[1], out[2], out[3], out[4]);
}

// term t of the cost of pair (i0, i1): 0 and 1 the self terms of i0 and i1,
// 2 the cross term (both orders counted)
__device__ __forceinline__ float merge_term(const VMMParams& V, const float (*selfp)[6], int i0, int i1, int t,
                                            const MergedLobe& m) {
  if (t < 2) {
    const int k = t == 0 ? i0 : i1;
    const float wk = V.w[k];
    const float* sp = selfp[k];
    return ((wk * wk) / m.w) *
           (sp[5] * integrated_division(sp[0], sp[1], sp[2], sp[3], sp[4], -m.dx, -m.dy, -m.dz, m.kappa, m.nrm, m.em2k));
  }
  float cx, cy, cz, ck, cn;
  const float s01 = product_scale(V.mux[i0], V.muy[i0], V.muz[i0], V.kappa[i0], V.norm[i0], V.mux[i1], V.muy[i1],
                                  V.muz[i1], V.kappa[i1], V.norm[i1], cx, cy, cz, ck, cn);
  return 2.f * ((V.w[i0] * V.w[i1]) / m.w) *
         (s01 * integrated_division(cx, cy, cz, ck, cn, -m.dx, -m.dy, -m.dz, m.kappa, m.nrm, m.em2k));
}

// chi = ((t0 + t1) + t2) - w, each term rounded on its own
__device__ __forceinline__ float merge_chi(float t0, float t1, float t2, float w) {
#pragma clang fp contract(off)
  float chi = 0.f;
  chi += t0;
  chi += t1;
  chi += t2;
  chi -= w;
  return chi;
}

// ComponentSplitStatistics::mergeComponentStats, zero-mean variant
// (Splitter:994-1100): the statistics of slot i after merging slot j into it.
struct SplitSlot {
  float covx, covy, covz, ssw, sn, ssa, chi;
};
__device__ __forceinline__ SplitSlot split_merged_slot(const SplitStatsState& SS, int i, int j, float wI, float mIx,
                                                       float mIy, float mIz, float wJ, float mJx, float mJy,
                                                       float mJz, float wK, float mKx, float mKy, float mKz) {
  float xx, xy, xz, yx, yy, yz;
  frame_of(mKx, mKy, mKz, xx, xy, xz, yx, yy, yz);
  float aix, aiy, ajx, ajy;
  log_map(dot3f(xx, xy, xz, mIx, mIy, mIz), dot3f(yx, yy, yz, mIx, mIy, mIz), dot3f(mKx, mKy, mKz, mIx, mIy, mIz),
          aix, aiy);
  log_map(dot3f(xx, xy, xz, mJx, mJy, mJz), dot3f(yx, yy, yz, mJx, mJy, mJz), dot3f(mKx, mKy, mKz, mJx, mJy, mJz),
          ajx, ajy);
  const float invWK = wK > 0.f ? 1.f / wK : 1.f;
  const float swI = SS.ssw[i], swJ = SS.ssw[j], swK = swI + swJ;
  const float rI = 1.f / swI, rJ = 1.f / swJ;  // used where the weight is positive (no branch)
  const bool pI = swI > 0.f, pJ = swJ > 0.f;
  const float cI[3] = {pI ? SS.covx[i] * rI : 0.f, pI ? SS.covy[i] * rI : 0.f, pI ? SS.covz[i] * rI : 0.f};
  const float cJ[3] = {pJ ? SS.covx[j] * rJ : 0.f, pJ ? SS.covy[j] * rJ : 0.f, pJ ? SS.covz[j] * rJ : 0.f};
  const float mII[3] = {aix * aix, aiy * aiy, aix * aiy};
  const float mJJ[3] = {ajx * ajx, ajy * ajy, ajx * ajy};
  float cK[3];
  for (int d = 0; d < 3; d++) cK[d] = (wI * cI[d] + wI * mII[d] + wJ * cJ[d] + wJ * mJJ[d]) * invWK;
  SplitSlot o;
  o.ssa = SS.ssa[i] + SS.ssa[j];
  o.sn = invWK * (wI * SS.sn[i] + wJ * SS.sn[j]);
  o.chi = SS.chi[i] + SS.chi[j];
  o.covx = cK[0] * swK; o.covy = cK[1] * swK; o.covz = cK[2] * swK;
  o.ssw = swK;
  return o;
}

// lexicographic (cost, pair index) minimum
__device__ __forceinline__ void argmin_take(float& best, int& bestP, float ob, int op) {
  if (ob < best || (ob == best && op < bestP)) { best = ob; bestP = op; }
}
#define PGL_ROW_ROR_I(v, n) __builtin_amdgcn_mov_dpp((v), 0x120 + (n), 0xF, 0xF, false)

// wave-wide lexicographic minimum of (best, bestP): DPP row rotations, then
// the row swaps; every lane ends with the same pair index
__device__ __forceinline__ int wave_argmin(float best, int bestP) {
  argmin_take(best, bestP, PGL_ROW_ROR(best, 1), PGL_ROW_ROR_I(bestP, 1));
  argmin_take(best, bestP, PGL_ROW_ROR(best, 2), PGL_ROW_ROR_I(bestP, 2));
  argmin_take(best, bestP, PGL_ROW_ROR(best, 4), PGL_ROW_ROR_I(bestP, 4));
  argmin_take(best, bestP, PGL_ROW_ROR(best, 8), PGL_ROW_ROR_I(bestP, 8));
  {
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(best), __float_as_uint(best), false, false);
    const auto b = __builtin_amdgcn_permlane16_swap((unsigned)bestP, (unsigned)bestP, false, false);
    best = __uint_as_float(a[0]);
    bestP = (int)b[0];
    argmin_take(best, bestP, __uint_as_float(a[1]), (int)b[1]);
  }
  {
    const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(best), __float_as_uint(best), false, false);
    const auto b = __builtin_amdgcn_permlane32_swap((unsigned)bestP, (unsigned)bestP, false, false);
    best = __uint_as_float(a[0]);
    bestP = (int)b[0];
    argmin_take(best, bestP, __uint_as_float(a[1]), (int)b[1]);
  }
  return __builtin_amdgcn_readfirstlane(bestP);
}

// One merge of pair (ci, cj) into slot ci, the last slot of each state
// moving into slot cj, on two waves at once (lanes 0..31 own component
// slots; every lane computes the merged values alike; each slot lane reads
// all old values before it writes, and LDS operations of one wave complete in
// order).  Wave 0: the mixture (PAVMM::mergeComponents PAVMM:191-257), the
// components' self products and the sufficient statistics
// (SufficientStatistics::mergeComponents WEM:485-518).  Wave 1: the split
// statistics (mergeComponentStats), from the pair's values stashed in sh.pre
// before the mixture changes.
__device__ void merge_step_mixture(Shared& sh, SuffStats& ST, int ci, int cj) {
  const int lane = threadIdx.x & 63;
  VMMParams& V = sh.S.vmm;
  const int LV = V.K - 1, LT = ST.K - 1;
  const MergedLobe m = merged_lobe(V, ci, cj);
  float msp[6];  // product of the merged lobe with itself
  msp[5] = product_scale(m.dx, m.dy, m.dz, m.kappa, m.nrm, m.dx, m.dy, m.dz, m.kappa, m.nrm, msp[0], msp[1], msp[2],
                         msp[3], msp[4]);
  const float w0 = V.w[ci], w1 = V.w[cj];
  const float d0 = V.dist[ci], d1 = V.dist[cj];
  float nd = w0 * d0 + w1 * d1;
  nd /= (w0 + w1);
  // merged sufficient statistics of slot ci
  const float tdx = ST.swdx[ci] + ST.swdx[cj], tdy = ST.swdy[ci] + ST.swdy[cj], tdz = ST.swdz[ci] + ST.swdz[cj];
  const float tsw = ST.sw[ci] + ST.sw[cj], tsd = ST.sdw[ci] + ST.sdw[cj];
  if (lane < 32) {
    const int k = lane;
    // the last slots as they stand once slot ci holds the merge
    const float lw = V.w[LV], lk = V.kappa[LV], lmc = V.mc[LV], ln = V.norm[LV], le = V.em2k[LV];
    const float lx = V.mux[LV], ly = V.muy[LV], lz = V.muz[LV], ld = V.dist[LV];
    const bool tI = LT == ci;
    const float udx = tI ? tdx : ST.swdx[LT], udy = tI ? tdy : ST.swdy[LT], udz = tI ? tdz : ST.swdz[LT];
    const float usw = tI ? tsw : ST.sw[LT], usd = tI ? tsd : ST.sdw[LT];
    float lsp[6];
    for (int a = 0; a < 6; a++) lsp[a] = sh.selfp[LV][a];
    if (k == LV) {
      clear_comp(V, k);
    } else if (k == cj) {
      V.w[k] = lw; V.kappa[k] = lk; V.mc[k] = lmc; V.norm[k] = ln; V.em2k[k] = le;
      V.mux[k] = lx; V.muy[k] = ly; V.muz[k] = lz; V.dist[k] = ld;
      for (int a = 0; a < 6; a++) sh.selfp[k][a] = lsp[a];
    } else if (k == ci) {
      V.w[k] = m.w; V.kappa[k] = m.kappa; V.mc[k] = m.mc; V.norm[k] = m.nrm; V.em2k[k] = m.em2k;
      V.mux[k] = m.dx; V.muy[k] = m.dy; V.muz[k] = m.dz; V.dist[k] = nd;
      for (int a = 0; a < 6; a++) sh.selfp[k][a] = msp[a];
    }
    if (k == LT) {
      suff_zero_slot(ST, k);
    } else if (k == cj) {
      ST.swdx[k] = udx; ST.swdy[k] = udy; ST.swdz[k] = udz; ST.sw[k] = usw; ST.sdw[k] = usd;
    } else if (k == ci) {
      ST.swdx[k] = tdx; ST.swdy[k] = tdy; ST.swdz[k] = tdz; ST.sw[k] = tsw; ST.sdw[k] = tsd;
    }
  }
  if (lane == 0) {
    V.K = LV;
    ST.K = LT;
  }
}

__device__ void merge_step_split(Shared& sh, SplitStatsState& SS, int ci, int cj) {
  const int lane = threadIdx.x & 63;
  const int LS = SS.K - 1;
  const float* pre = sh.pre;  // w, mu x, y, z of ci; the same of cj; mean cosines of ci, cj
  const MergedLobe m = merged_lobe_of(pre[0], pre[8], pre[1], pre[2], pre[3], pre[4], pre[9], pre[5], pre[6], pre[7]);
  const SplitSlot ns = split_merged_slot(SS, ci, cj, pre[0], pre[1], pre[2], pre[3], pre[4], pre[5], pre[6], pre[7],
                                         m.w, m.dx, m.dy, m.dz);
  if (lane < 32) {
    const int k = lane;
    const bool sI = LS == ci;
    const float schi = sI ? ns.chi : SS.chi[LS], scx = sI ? ns.covx : SS.covx[LS], scy = sI ? ns.covy : SS.covy[LS];
    const float scz = sI ? ns.covz : SS.covz[LS], ssn = sI ? ns.sn : SS.sn[LS], ssw = sI ? ns.ssw : SS.ssw[LS];
    const float ssa = sI ? ns.ssa : SS.ssa[LS];
    if (k == LS) {
      split_zero_slot(SS, k);
    } else if (k == cj) {
      SS.chi[k] = schi; SS.covx[k] = scx; SS.covy[k] = scy; SS.covz[k] = scz; SS.sn[k] = ssn; SS.ssw[k] = ssw;
      SS.ssa[k] = ssa;
    } else if (k == ci) {
      SS.chi[k] = ns.chi; SS.covx[k] = ns.covx; SS.covy[k] = ns.covy; SS.covz[k] = ns.covz; SS.sn[k] = ns.sn;
      SS.ssw[k] = ns.ssw; SS.ssa[k] = ns.ssa;
    }
  }
  if (lane == 0) SS.K = LS;
}

// PerformMerging with ThresholdedMergeNext (Merger:45-89, 218-279)
//
// Pair costs live in sh.cost[i * 32 + j] (i < j).  After a merge only the
// pairs of the merged slot ci and of slot cj (which receives the last
// component) change: every other pair has unchanged inputs, so the table
// equals a full recomputation bit for bit.  The merge loop is sequential by
// definition.  A step: wave 0 takes the arg-min over the pairs whose cost did
// not change (least cost below the threshold with both components holding
// split samples, ties to the first pair in (i, j) order -- the serial scan of
// ThresholdedMergeNext) while waves 1..3 compute the three cost terms of the
// changed pairs, one pair per lane; wave 0 then writes the changed costs,
// takes them into the arg-min and stashes the chosen pair's values; the
// merge runs on waves 0 and 1 at once.  Each component's product with itself
// is cached and moves with the component.
__device__ int perform_merging(Shared& sh, SuffStats& ST, SplitStatsState& SS, const FactoryConfig& cfg) {
  const unsigned long long tMerge = wall_clock64();
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  VMMParams& V = sh.S.vmm;
  int total = 0;
  if (V.K > sh.W) {
    for (int k = tid; k < V.K; k += RT) self_product(V, k, sh.selfp[k]);
    __syncthreads();
    // full table
    for (int p = tid; p < 32 * 32; p += RT) {
      const int i = p >> 5, j = p & 31;
      if (i < j && j < V.K) {
        const MergedLobe m = merged_lobe(V, i, j);
        const float t0 = merge_term(V, sh.selfp, i, j, 0, m);
        const float t1 = merge_term(V, sh.selfp, i, j, 1, m);
        const float t2 = merge_term(V, sh.selfp, i, j, 2, m);
        sh.cost[p] = merge_chi(t0, t1, t2, m.w);
      }
    }
    __syncthreads();
    // ci, cj: the pair merged in the previous step (-1 before the first);
    // pairs with either index changed, every other pair kept its cost
    int ci = -1, cj = -1;
    for (;;) {
      const int K = V.K;
      // the changed pair of this lane (lanes 0..31: (ci, q), lanes 32..63:
      // (cj, q - 32) unless that pair is listed already)
      int pi = -1, pj = -1;
      if (ci >= 0) {
        const int a0 = lane < 32 ? ci : cj, b0 = lane & 31;
        if (a0 < K && b0 < K && b0 != a0 && !(lane >= 32 && b0 == ci)) { pi = min(a0, b0); pj = max(a0, b0); }
      }
      float best = 3.402823466e38f;
      int bestP = 0x7fffffff;
      uint32_t hasSplit = 0;
      if (wave == 0) {
        // arg-min over the unchanged pairs while waves 1..3 compute the
        // changed pairs' terms; lane l scans pairs l, l + 64, ... in
        // increasing order (rows 2r and 2r + 1), reading four at a time
        hasSplit = (uint32_t)__ballot(lane < 32 && SS.sn[lane & 31] > 0.f);
        const float thr = cfg.mergingThreshold;
        for (int r0 = 0; r0 < 16 && 2 * r0 < K; r0 += 4) {
          float cs[4];
#pragma unroll
          for (int u = 0; u < 4; u++) cs[u] = sh.cost[lane + 64 * (r0 + u)];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int p = lane + 64 * (r0 + u);
            const int i = p >> 5, j = p & 31;
            const bool changed = i == ci || i == cj || j == ci || j == cj;
            const float cst = cs[u];
            if (i < j && j < K && !changed && ((hasSplit >> i) & (hasSplit >> j) & 1u) && cst < thr && cst < best) {
              best = cst;
              bestP = p;
            }
          }
        }
      } else if (pi >= 0) {
        sh.term[lane][wave - 1] = merge_term(V, sh.selfp, pi, pj, wave - 1, merged_lobe(V, pi, pj));
      }
      __syncthreads();
      if (wave == 0) {
        // the changed pairs' costs into the table and into the arg-min
        if (pi >= 0) {
          const float cst = merge_chi(sh.term[lane][0], sh.term[lane][1], sh.term[lane][2], V.w[pi] + V.w[pj]);
          sh.cost[pi * 32 + pj] = cst;
          if (((hasSplit >> pi) & (hasSplit >> pj) & 1u) && cst < cfg.mergingThreshold)
            argmin_take(best, bestP, cst, pi * 32 + pj);
        }
        const int pm = wave_argmin(best, bestP);
        if (lane == 0) {
          sh.found = pm != 0x7fffffff;
          if (sh.found) {
            const int mi = pm >> 5, mj = pm & 31;
            sh.ci = mi;
            sh.cj = mj;
            sh.pre[0] = V.w[mi]; sh.pre[1] = V.mux[mi]; sh.pre[2] = V.muy[mi]; sh.pre[3] = V.muz[mi];
            sh.pre[4] = V.w[mj]; sh.pre[5] = V.mux[mj]; sh.pre[6] = V.muy[mj]; sh.pre[7] = V.muz[mj];
            sh.pre[8] = V.mc[mi]; sh.pre[9] = V.mc[mj];
          }
        }
      }
      __syncthreads();
      if (!sh.found) break;
      ci = sh.ci;
      cj = sh.cj;
      if (wave == 0) merge_step_mixture(sh, ST, ci, cj);
      else if (wave == 1) merge_step_split(sh, SS, ci, cj);
      __syncthreads();
      total++;
      if (V.K <= sh.W) break;
    }
  }
  if (tid == 0) sh.S.prof[5] += (float)(wall_clock64() - tMerge);
  return total;
}

// ---- validity (PAVMM:421-513, WEM:285-343)
__device__ bool region_valid(Shared& sh, bool checkSuff) {
  const int tid = threadIdx.x;
  int ok = 1;
  if (tid < 32) {
    const VMMParams& V = sh.S.vmm;
    const int k = tid;
    const float vals[9] = {V.w[k], V.kappa[k], V.mux[k], V.muy[k], V.muz[k], V.norm[k], V.em2k[k], V.mc[k], V.dist[k]};
    for (int i = 0; i < 9; i++) if (!isfinite(vals[i])) ok = 0;
    if (k < V.K) {
      if (V.w[k] < 0.f || V.w[k] > 1.f + 1e-6f) ok = 0;
      if (V.kappa[k] < 0.f || V.mc[k] < 0.f || V.mc[k] > 1.f || V.norm[k] < 0.f || V.dist[k] < 0.f) ok = 0;
      if (V.mux[k] < -1.f || V.mux[k] > 1.f || V.muy[k] < -1.f || V.muy[k] > 1.f || V.muz[k] < -1.f || V.muz[k] > 1.f) ok = 0;
    } else {
      if (V.w[k] != 0.f || V.kappa[k] != 0.f || V.mux[k] != 0.f || V.muy[k] != 0.f || V.muz[k] != 1.f || V.mc[k] != 0.f ||
          fabsf(V.norm[k] - kOneOverFourPi) >= 1e-6f || V.em2k[k] != 1.f || V.dist[k] != 0.f)
        ok = 0;
    }
    if (checkSuff) {
      const SuffStats& s = sh.S.suff;
      if (!isfinite(s.swdx[k]) || !isfinite(s.swdy[k]) || !isfinite(s.swdz[k]) || !isfinite(s.sw[k]) || !isfinite(s.sdw[k])) ok = 0;
      if (k < s.K) { if (s.sw[k] < 0.f || s.sdw[k] < 0.f) ok = 0; }
      else if (s.swdx[k] != 0.f || s.swdy[k] != 0.f || s.swdz[k] != 0.f || s.sw[k] != 0.f || s.sdw[k] != 0.f) ok = 0;
      if (k == 0 && (!isfinite(s.numSamples) || s.numSamples < 0.f || !isfinite(s.sumWeights) || s.sumWeights < 0.f)) ok = 0;
    }
  }
  return __syncthreads_and(ok) != 0;
}

// ---- ASM::fit (ASM:303-347)
__device__ __noinline__ void region_fit(Shared& sh, Team& tm, const RegionSamples& rs, uint32_t base, uint32_t n,
                                        const FactoryConfig& cfg, const float* __restrict__ uniformDirs) {
  const int tid = threadIdx.x;
  VMMParams& V = sh.S.vmm;
  const int K0 = sh.S.initK > 0 ? min(sh.S.initK, kMaxComponents) : cfg.initK;
  clear_suff(sh, sh.S.suff, K0);
  clear_split(sh, sh.S.split, K0);
  // InitUniformVMM (WEM:632-666): spherical Fibonacci means for K0 components
  if (tid < 32) {
    const int k = tid;
    const float* d = uniformDirs + ((K0 - 1) * 32 + k) * 3;
    V.mux[k] = d[0]; V.muy[k] = d[1]; V.muz[k] = d[2];
    if (k < K0) {
      V.kappa[k] = cfg.initKappa;
      V.w[k] = 1.f / (float)K0;
    } else {
      V.kappa[k] = 0.f; V.w[k] = 0.f; V.norm[k] = kOneOverFourPi; V.em2k[k] = 1.f; V.mc[k] = 0.f;
    }
    calc_normalization_lane(V, k);
    const float t = 1.f - 2.f / (expf(2.f * V.kappa[k]) - 1.f);
    const float m = 1.f / t - 1.f / V.kappa[k];
    V.mc[k] = V.kappa[k] > 0.f ? m : 0.f;
  }
  if (tid == 0) {
    V.K = K0;
    sh.S.afterSplitLo = sh.S.afterSplitHi = sh.S.afterMergeLo = sh.S.afterMergeHi = 0;
  }
  __syncthreads();
  clear_suff(sh, sh.S.suff, K0);
  if (tid == 0) sh.S.suff.normalized = 1;
  __syncthreads();
  const int its = em(sh, tm, sh.S.suff, rs, base, n, cfg, false);
  distances(sh, tm, sh.S.suff, rs, base, n, true);
  int splits = 0, merges = 0, pits = 0;
  if (cfg.useSplitAndMerge) {
    const float mcEst = sh.S.suff.sumWeights / sh.S.suff.numSamples;
    splits = recursive_splitting(sh, tm, rs, base, n, mcEst, cfg, pits);
    clear_split(sh, sh.S.split, V.K);
    update_split_stats(sh, tm, sh.S.split, rs, base, n, mcEst);
    merges = perform_merging(sh, sh.S.suff, sh.S.split, cfg);
  }
  distances(sh, tm, sh.S.suff, rs, base, n, true);
  if (tid == 0) {
    sh.S.afterSplitLo = sh.S.afterSplitHi = sh.S.afterMergeLo = sh.S.afterMergeHi = 0;
    sh.S.emIterations = its; sh.S.partialEmIterations = pits; sh.S.numSplits = splits; sh.S.numMerges = merges;
  }
  __syncthreads();
}

__device__ __forceinline__ unsigned long long get64(uint32_t lo, uint32_t hi) {
  return ((unsigned long long)hi << 32) | lo;
}

// ---- ASM::update (ASM:352-462)
__device__ void region_update_body(Shared& sh, Team& tm, const RegionSamples& rs, uint32_t base, uint32_t n,
                                   const FactoryConfig& cfg) {
  const int tid = threadIdx.x;
  VMMParams& V = sh.S.vmm;
  RegionState& S = sh.S;
  const int prevK = V.K;
  const int its = em(sh, tm, S.suff, rs, base, n, cfg, false);
  if (tid == 0 && prevK < V.K) S.split.K = V.K;
  __syncthreads();
  int splits = 0, merges = 0, pits = 0;
  if (cfg.useSplitAndMerge) {
    const float mcEst = S.suff.sumWeights / S.suff.numSamples;
    unsigned long long afterSplit = get64(S.afterSplitLo, S.afterSplitHi) + n;
    unsigned long long afterMerge = get64(S.afterMergeLo, S.afterMergeHi) + n;
    update_split_stats(sh, tm, S.split, rs, base, n, mcEst);
    if (afterSplit >= (unsigned long long)(cfg.minSamplesForSplitting < 0 ? 0 : cfg.minSamplesForSplitting)) {
      const unsigned long long tS = wall_clock64();
      const int kBefore = V.K;
      rank_candidates(sh, S.split);
      __syncthreads();
      if (tid == 0) {
        sh.mask = 0u;
        int ns = 0;
        const int numCand = S.split.K;
        for (int c = 0; c < numCand; c++) {
          const int idx = sh.order[c];
          if (S.split.chi[idx] > cfg.splittingThreshold && V.K < kMaxComponents) {
            split_component(V, S.split, S.suff, idx, cfg);
            sh.mask |= 1u << idx;
            sh.mask |= 1u << (V.K - 1);
            ns++;
          }
        }
        sh.numSplits = ns;
      }
      __syncthreads();
      // the renormalisation of every component that each split does
      // (PAVMM:165-190), once after the loop: nothing reads it before
      if (V.K > kBefore) {
        if (tid < 32) calc_normalization_lane(V, tid);
        __syncthreads();
      }
      if (tid == 0) S.prof[6] += (float)(wall_clock64() - tS);
      splits = sh.numSplits;
      if (splits > 0 && cfg.partialReFit && n >= (uint32_t)(cfg.minSamplesForPartialRefitting < 0 ? 0 : cfg.minSamplesForPartialRefitting)) {
        copy_suff(sh, sh.tmp, S.suff);
        clear_suff(sh, sh.tmp, V.K);
        pits = em(sh, tm, sh.tmp, rs, base, n, cfg, true);
        if (tid == 0) S.suff.K = V.K;
        __syncthreads();
        masked_replace(sh, S.suff, sh.tmp);
        if (tid == 0) S.split.K = V.K;
        __syncthreads();
      }
      afterSplit = 0;
    }
    if (afterMerge >= (unsigned long long)(cfg.minSamplesForMerging < 0 ? 0 : cfg.minSamplesForMerging)) {
      merges = perform_merging(sh, S.suff, S.split, cfg);
      afterMerge = 0;
    }
    if (tid == 0) {
      S.afterSplitLo = (uint32_t)afterSplit; S.afterSplitHi = (uint32_t)(afterSplit >> 32);
      S.afterMergeLo = (uint32_t)afterMerge; S.afterMergeHi = (uint32_t)(afterMerge >> 32);
    }
    __syncthreads();
  }
  distances(sh, tm, S.suff, rs, base, n, false);
  if (tid == 0) { S.emIterations = its; S.partialEmIterations = pits; S.numSplits = splits; S.numMerges = merges; }
  __syncthreads();
}

// One region on one team rank: load the state into LDS, run fit or update,
// rank 0 writes the state back.
__device__ void region_job(Shared& sh, Team& tm, RegionState* __restrict__ states, uint32_t r, const RegionSamples& rs,
                           const FactoryConfig& cfg, const float* __restrict__ uniformDirs) {
  const unsigned long long tJob = wall_clock64();
  const int tid = threadIdx.x;
  {
    const uint32_t* src = (const uint32_t*)(states + r);
    uint32_t* dst = (uint32_t*)&sh.S;
    for (int i = tid; i < (int)(sizeof(RegionState) / 4); i += RT) dst[i] = src[i];
  }
  __syncthreads();
  RegionState& S = sh.S;
  const int mode = S.mode;
  const uint32_t n = S.sampleCount, base = S.sampleBegin;
  if (tid == 0) {
    S.emIterations = 0; S.partialEmIterations = 0; S.numSplits = 0; S.numMerges = 0;
    S.estepEvals = 0.f; S.auxEvals = 0.f; S.splitEvals = 0.f;
    for (int i = 0; i < 12; i++) S.prof[i] = 0.f;
    S.prof[7] = (float)tm.T;
    S.prof[11] = (float)S.sampleCount;
    sh.W = cfg.vectorWidth;
  }
  __syncthreads();
  if (mode == 2 && S.splitFlag) {  // Field.h:457-461: decay the training statistics of split regions
    const float a = cfg.decayOnSpatialSplit;
    if (tid < 32) {
      const int k = tid;
      S.suff.swdx[k] *= a; S.suff.swdy[k] *= a; S.suff.swdz[k] *= a; S.suff.sw[k] *= a; S.suff.sdw[k] *= a;
      if (k < vec_lim(sh.W, S.split.K)) {
        S.split.covx[k] *= a; S.split.covy[k] *= a; S.split.covz[k] *= a;
        S.split.sn[k] *= a; S.split.ssw[k] *= a; S.split.ssa[k] *= a;
      }
    }
    if (tid == 0) { S.suff.numSamples *= a; S.suff.sumWeights *= a; S.splitFlag = 0; }
    __syncthreads();
  }
  if (n > 0) {
    if (mode == 2) {
      // parallax shift of the statistics and of the mixture to the new pivot
      // (WEM:265-283, PAVMM:809-831); the plain VMM keeps its fit-time pivot
      // (Field.h:476)
      if (!cfg.noParallax) {
        const float shx = S.vmm.pivot[0] - S.ssMean[0], shy = S.vmm.pivot[1] - S.ssMean[1],
                    shz = S.vmm.pivot[2] - S.ssMean[2];
        if (tid < 32) {
          const int k = tid;
          VMMParams& V = S.vmm;
          if (k < V.K) {
            const float dk = V.dist[k];
            if (dk > 0.f) {
              float sx = S.suff.swdx[k], sy = S.suff.swdy[k], sz = S.suff.swdz[k];
              const float mcos = sqrtf(dot3f(sx, sy, sz, sx, sy, sz));
              if (mcos > 0.f) {
                sx = (sx / mcos) * dk + shx;
                sy = (sy / mcos) * dk + shy;
                sz = (sz / mcos) * dk + shz;
                const float l = sqrtf(dot3f(sx, sy, sz, sx, sy, sz));
                S.suff.swdx[k] = (sx / l) * mcos;
                S.suff.swdy[k] = (sy / l) * mcos;
                S.suff.swdz[k] = (sz / l) * mcos;
              }
            }
            const float px = V.mux[k] * dk + shx, py = V.muy[k] * dk + shy, pz = V.muz[k] * dk + shz;
            const float l = sqrtf(dot3f(px, py, pz, px, py, pz));
            if (dk > 0.f && isfinite(dk)) { V.mux[k] = px / l; V.muy[k] = py / l; V.muz[k] = pz / l; }
            V.dist[k] = dk > 0.f ? l : 0.f;
          }
        }
        __syncthreads();
        if (tid == 0) { S.vmm.pivot[0] -= shx; S.vmm.pivot[1] -= shy; S.vmm.pivot[2] -= shz; }
        __syncthreads();
      }
      region_update_body(sh, tm, rs, base, n, cfg);
      const bool ok = region_valid(sh, true);
      if (tid == 0) S.valid = ok ? 1 : 0;
    } else if (mode == 1) {
      region_fit(sh, tm, rs, base, n, cfg, uniformDirs);
      if (tid == 0) { S.vmm.pivot[0] = S.ssMean[0]; S.vmm.pivot[1] = S.ssMean[1]; S.vmm.pivot[2] = S.ssMean[2]; }
      const bool ok = region_valid(sh, false);
      if (tid == 0) { S.valid = ok ? 1 : 0; S.splitFlag = 0; S.fitted = 1; }
    }
  } else if (mode == 1) {
    if (tid == 0) { S.valid = 0; S.splitFlag = 0; }
  }
  if (tid == 0) sh.S.prof[0] = (float)(wall_clock64() - tJob);
  __syncthreads();
  if (tm.rank == 0) {
    uint32_t* dst = (uint32_t*)(states + r);
    const uint32_t* src = (const uint32_t*)&sh.S;
    for (int i = tid; i < (int)(sizeof(RegionState) / 4); i += RT) dst[i] = src[i];
  }
}


// Persistent region kernel: each workgroup takes queue items (job, rank) in
// order until the queue is empty; the ranks of a job are queued
// consecutively, so a team is always completed by the next free workgroups
// and every wait ends (given T <= grid, which the host guarantees, and a
// grid that fits the device).  This file is compiled twice (each build with
// its own device functions): PGL_REGION_OCC = 3 fits three workgroups on a CU
// (168 VGPRs) and serves the default launch of two per CU, which leaves the
// room of the third to other streams' kernels (the next update's spatial
// phase); PGL_REGION_OCC = 2 (248 VGPRs, the faster code, k_region_update_wide)
// serves launches of at most one workgroup per CU, which leave half of every
// CU to other work (serving next to training).
__global__ void __launch_bounds__(RT) __attribute__((amdgpu_waves_per_eu(PGL_REGION_OCC, PGL_REGION_OCC)))
PGL_REGION_KERNEL(RegionState* __restrict__ states,
                                                      const RegionJob* __restrict__ jobs,
                                                      const uint2* __restrict__ items, uint32_t numItems,
                                                      uint32_t* __restrict__ queue, uint32_t* __restrict__ counters,
                                                      float* __restrict__ partial, int* __restrict__ err,
                                                      unsigned long long timeout,
                                                      unsigned long long* __restrict__ stats,
                                                      float* __restrict__ costHint, RegionSamples rs,
                                                      FactoryConfig cfg, const float* __restrict__ uniformDirs) {
  __shared__ Shared sh;
  const int tid = threadIdx.x;
  if (tid == 0) {
    sh.abort = 0;
    sh.cfg = cfg;
    sh.rs = rs;
    sh.stageKey = 0xffffffffu;
    sh.stageCnt = 0;
    sh.stageExtra = nullptr;
  }
  for (;;) {
    if (tid == 0) {
      const uint32_t it = atomicAdd(queue, 1u);
      sh.item = it;
      if (it < numItems) {
        const uint2 item = items[it];
        const RegionJob job = jobs[item.x];
        Team& tm = sh.tm;
        tm.rank = item.y;
        tm.T = job.team;
        tm.nc = job.chunks;
        tm.part = partial + job.partialOffset;
        tm.counter = counters + item.x;
        tm.epoch = 0;
        tm.err = err;
        tm.timeout = timeout;
        sh.jobRegion = job.region;
        sh.jobIndex = item.x;
      }
    }
    __syncthreads();
    if (sh.item >= numItems) break;
    Team& tm = sh.tm;
    region_job(sh, tm, states, sh.jobRegion, sh.rs, sh.cfg, uniformDirs);
    if (tid == 0 && tm.rank == 0) {
      const RegionState& S = sh.S;
      atomicAdd(&stats[0], (unsigned long long)S.estepEvals);
      atomicAdd(&stats[1], (unsigned long long)S.auxEvals);
      atomicAdd(&stats[2], (unsigned long long)S.splitEvals);
      atomicAdd(&stats[3], (unsigned long long)(S.emIterations + S.partialEmIterations));
      if (costHint) {
        const float n = (float)S.sampleCount;
        costHint[sh.jobRegion] = n > 0.f ? (S.estepEvals + S.auxEvals + S.splitEvals) / n : 0.f;
      }
    }
    __syncthreads();
  }
}

#if PGL_REGION_OCC == 3
// the PGL_REGION_OCC = 2 build of the kernel (its own translation unit)
const void* region_kernel_wide();

// clone the directional state of the ancestor into regions created by spatial splits
__global__ void __launch_bounds__(256) k_copy_states(RegionState* __restrict__ states, const uint32_t* __restrict__ pairs,
                                                     uint32_t numPairs) {
  const uint32_t p = blockIdx.x;
  if (p >= numPairs) return;
  const uint32_t dst = pairs[2 * p], src = pairs[2 * p + 1];
  const uint4* s = (const uint4*)(states + src);
  uint4* d = (uint4*)(states + dst);
  for (uint32_t i = threadIdx.x; i < sizeof(RegionState) / 16; i += blockDim.x) d[i] = s[i];
}

__global__ void __launch_bounds__(256) k_set_inputs(RegionState* __restrict__ states, const RegionInput* __restrict__ in,
                                                    uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const RegionInput v = in[i];
  RegionState& S = states[v.region];
  for (int j = 0; j < 3; j++) { S.ssMean[j] = v.ssMean[j]; S.ssVar[j] = v.ssVar[j]; }
  S.ssN = v.ssN;
  S.sampleBegin = v.begin;
  S.sampleCount = v.count;
  S.mode = v.mode;
  S.splitFlag = v.splitFlag;
  S.initK = v.initK;
  S.emIterations = 0; S.partialEmIterations = 0; S.numSplits = 0; S.numMerges = 0;
  S.estepEvals = 0.f; S.auxEvals = 0.f; S.splitEvals = 0.f;
}

// Reads the launch's tables straight from pinned host memory (no copy
// operation before it): all blocks copy the job and item tables (copy16
// 16-byte words) to device memory for the region kernel; block 0 zeroes the
// launch's counters; block b < numInputs prepares the region of inputs[b].  A
// new region's ancestor may receive its own inputs while the copy runs: only
// input fields can differ, and the new region overwrites all of them.
__global__ void __launch_bounds__(256) k_prepare_regions(RegionState* __restrict__ states,
                                                         const uint4* __restrict__ hostTables, uint4* __restrict__ tables,
                                                         uint32_t copy16, const RegionInput* __restrict__ in,
                                                         uint32_t numInputs, uint32_t* __restrict__ zero,
                                                         uint32_t numZero, int* __restrict__ err,
                                                         unsigned long long* __restrict__ stats) {
  const uint32_t b = blockIdx.x;
  for (uint32_t i = b * blockDim.x + threadIdx.x; i < copy16; i += gridDim.x * blockDim.x) tables[i] = hostTables[i];
  if (b == 0) {
    for (uint32_t i = threadIdx.x; i < numZero; i += blockDim.x) zero[i] = 0u;
    if (threadIdx.x == 0) *err = 0;
    if (threadIdx.x < 4) stats[threadIdx.x] = 0ull;
  }
  if (in == nullptr || b >= numInputs) return;
  const RegionInput v = in[b];
  if (v.copyFrom >= 0) {  // uniform over the block
    const uint4* src = (const uint4*)(states + v.copyFrom);
    uint4* dst = (uint4*)(states + v.region);
    for (uint32_t i = threadIdx.x; i < sizeof(RegionState) / 16; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    RegionState& S = states[v.region];
    for (int j = 0; j < 3; j++) { S.ssMean[j] = v.ssMean[j]; S.ssVar[j] = v.ssVar[j]; }
    S.ssN = v.ssN;
    S.sampleBegin = v.begin;
    S.sampleCount = v.count;
    S.mode = v.mode;
    S.splitFlag = v.splitFlag;
    S.initK = v.initK;
    S.emIterations = 0; S.partialEmIterations = 0; S.numSplits = 0; S.numMerges = 0;
    S.estepEvals = 0.f; S.auxEvals = 0.f; S.splitEvals = 0.f;
  }
}

void launch_prepare_regions(RegionState* states, const void* hostTables, void* tables, size_t tableBytes,
                            const RegionInput* inputs, uint32_t numInputs, uint32_t* zero, uint32_t numZero, int* err,
                            unsigned long long* stats, hipStream_t s) {
  const uint32_t copy16 = (uint32_t)((tableBytes + 15) / 16);
  uint32_t blocks = inputs && numInputs ? numInputs : 1u;
  blocks = std::max<uint32_t>(blocks, std::min<uint32_t>((copy16 + 255) / 256, 1024u));
  hipLaunchKernelGGL(k_prepare_regions, dim3(blocks), dim3(256), 0, s, states, (const uint4*)hostTables,
                     (uint4*)tables, copy16, inputs, numInputs, zero, numZero, err, stats);
}

void launch_copy_states(RegionState* states, const uint32_t* pairs, uint32_t numPairs, hipStream_t s) {
  if (numPairs == 0) return;
  hipLaunchKernelGGL(k_copy_states, dim3(numPairs), dim3(256), 0, s, states, pairs, numPairs);
}

void launch_set_inputs(RegionState* states, const RegionInput* inputs, uint32_t count, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_set_inputs, dim3((count + 255) / 256), dim3(256), 0, s, states, inputs, count);
}

__global__ void __launch_bounds__(256) k_zero_u32(uint32_t* __restrict__ p, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

int region_kernel_blocks_per_cu() {
  int b = 0;
  PGL_KERNEL_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_region_update, RT, 0));
  return b;
}

void launch_region_update(const RegionLaunch& L, const FactoryConfig& cfg, hipStream_t s) {
  if (L.numItems == 0) return;
  const uint32_t nz = L.numJobs + 1;
  if (L.zeroQueue) hipLaunchKernelGGL(k_zero_u32, dim3((nz + 255) / 256), dim3(256), 0, s, L.queueAndCounters, nz);
  RegionSamples rs{L.dx, L.dy, L.dz, L.w, L.pdf, L.dist};
  RegionState* states = L.states;
  const RegionJob* jobs = L.jobs;
  const uint2* items = L.items;
  uint32_t numItems = L.numItems;
  uint32_t* queue = L.queueAndCounters;
  uint32_t* counters = L.queueAndCounters + 1;
  float* partial = L.partial;
  int* err = L.err;
  unsigned long long timeout = L.timeoutTicks;
  unsigned long long* stats = L.stats;
  float* costHint = L.costHint;
  const float* uniformDirs = L.uniformDirs;
  FactoryConfig c = cfg;
  // at most one workgroup per CU: the variant compiled for two per CU
  int cus = 0, dev = 0;
  PGL_KERNEL_CHECK(hipGetDevice(&dev));
  PGL_KERNEL_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const bool wide = L.grid <= (uint32_t)cus;
  void* args[] = {&states, &jobs,    &items, &numItems, &queue, &counters,   &partial,
                  &err,    &timeout, &stats, &costHint, &rs,    (void*)&c, &uniformDirs};
  const void* fn = wide ? region_kernel_wide() : (const void*)k_region_update;
  if (L.cooperative) PGL_KERNEL_CHECK(hipLaunchCooperativeKernel(fn, dim3(L.grid), dim3(RT), args, 0, s));
  else PGL_KERNEL_CHECK(hipLaunchKernel(fn, dim3(L.grid), dim3(RT), args, 0, s));
}

#else
const void* region_kernel_wide() { return (const void*)k_region_update_wide; }
#endif

}  // namespace pglgpu
