// Host orchestration of Field::Update on the MI355X.
//
// The spatial decisions (split test, split dimension and position, region
// statistics merge/split/decay) are taken on the host in exact IEEE float
// arithmetic from integer statistics that the GPU accumulates; everything
// proportional to the number of samples runs in HIP kernels.
//
// Reference behaviour:
//   KDTreePartitionBuilder::updateTree/updateTreeNode  spatial/kdtree/KDTreeBuilder.h:76-118, 233-350
//   SampleStatistics / IntegerSampleStatistics          data/SampleStatistics.h:12-342
//   Field::buildField/updateField/fitRegions/updateRegions  field/Field.h:147-220, 389-511
//   Field::estimateSceneBounds                           field/Field.h:326-344
//   Field::getRegion                                     field/Field.h:109-145
#pragma clang fp contract(off)
#include "field.h"

#include <immintrin.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>

namespace pglgpu {

static const float kFltMax = 3.402823466e38f;

// ---------------------------------------------------------------- statistics
void HostSampleStats::merge(const HostSampleStats& b) {  // SampleStatistics.h:101-118
  float mA[3], mB[3], vA[3], vB[3];
  for (int i = 0; i < 3; i++) { mA[i] = mean[i]; mB[i] = b.mean[i]; vA[i] = var[i]; vB[i] = b.var[i]; }
  const float nA = n, nB = b.n;
  for (int i = 0; i < 3; i++) mean[i] = mA[i] * nA + mB[i] * nB;
  n += nB;
  for (int i = 0; i < 3; i++) mean[i] /= n;
  for (int i = 0; i < 3; i++) {
    const float a = (vA[i] + (nA * mA[i]) * mA[i]) + vB[i];
    const float s = a + (nB * mB[i]) * mB[i];
    var[i] = s - (n * mean[i]) * mean[i];
  }
  for (int i = 0; i < 3; i++) { lo[i] = std::min(lo[i], b.lo[i]); hi[i] = std::max(hi[i], b.hi[i]); }
}

void HostSampleStats::split(int dim, float pos, float decay, bool splitLower) {  // SampleStatistics.h:68-99
  if (n > 0.f) {
    const float variance = var[dim] / n;
    const float sd = std::sqrt(variance);
    const float nv = variance - variance / 4.0f;
    var[dim] = nv * n;
    if (splitLower) {
      lo[dim] = std::max(pos, lo[dim]);
      mean[dim] = std::min(hi[dim], mean[dim] + sd / 2.0f);
    } else {
      hi[dim] = std::min(pos, hi[dim]);
      mean[dim] = std::max(lo[dim], mean[dim] - sd / 2.0f);
    }
    n *= decay;
    for (int i = 0; i < 3; i++) var[i] *= decay;
  }
}

void HostSampleStats::variance(float v[3]) const {
  for (int i = 0; i < 3; i++) v[i] = var[i] / n;
}

// embree::rcp (common/math/emath.h:61-88) as each reference device type
// compiles it (openpgl/CMakeLists.txt:125-131): CPU_4 (SSE4.1) rcpss and a
// Newton step without FMA, CPU_8 (AVX2) rcpss and a fused Newton step,
// CPU_16 (AVX-512VL, AVX2) rcp14 and a fused Newton step.
__attribute__((target("sse4.1"))) static float rcp_sse4(float x) {
  const __m128 a = _mm_set_ss(x);
  const __m128 r = _mm_rcp_ss(a);
  return _mm_cvtss_f32(_mm_mul_ss(r, _mm_sub_ss(_mm_set_ss(2.0f), _mm_mul_ss(r, a))));
}
__attribute__((target("avx2,fma"))) static float rcp_avx2(float x) {
  const __m128 a = _mm_set_ss(x);
  const __m128 r = _mm_rcp_ss(a);
  return _mm_cvtss_f32(_mm_mul_ss(r, _mm_fnmadd_ss(r, a, _mm_set_ss(2.0f))));
}
__attribute__((target("avx512f,avx512vl,avx2,fma"))) static float rcp_avx512(float x) {
  const __m128 a = _mm_set_ss(x);
  const __m128 r = _mm_rcp14_ss(_mm_set_ss(0.0f), a);
  return _mm_cvtss_f32(_mm_mul_ss(r, _mm_fnmadd_ss(r, a, _mm_set_ss(2.0f))));
}

bool host_supports_width(int vectorWidth) {
  if (vectorWidth == 16) return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512vl");
  if (vectorWidth == 8) return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
  return __builtin_cpu_supports("sse4.1");
}

float host_rcp(float x, int vectorWidth) {
  if (vectorWidth == 4) return rcp_sse4(x);
  if (vectorWidth == 16) return rcp_avx512(x);
  return rcp_avx2(x);
}

// IntegerSampleStatistics::getSampleStatistics (SampleStatistics.h:297-340)
HostSampleStats int_stats_to_float(const IntStatsAcc& a, const float center[3], const float half[3]) {
  HostSampleStats s;
  if (a.n > 0) {
    const uint32_t n = (uint32_t)a.n;
    const float invN = 1.f / float(n);
    for (int i = 0; i < 3; i++) {
      float m = (float(a.sum[i]) / 4096.f) * invN;
      float v = (float(a.sq[i]) / 4096.f) * invN;
      v = v - m * m;
      v = std::fabs(v);
      m = m * half[i];
      m = m + center[i];
      v = v * (half[i] * half[i]);
      s.mean[i] = m;
      s.var[i] = v * float(n);
      float l = float(a.lo[i]) - 0.0f;
      l = (l / 4096.f) * half[i] + center[i];
      float u = float(a.hi[i]) + 0.0f;
      u = (u / 4096.f) * half[i] + center[i];
      s.lo[i] = l;
      s.hi[i] = u;
    }
    s.n = float(n);
  }
  return s;
}

static void bounds_params(const float lo[3], const float hi[3], float center[3], float half[3], float inv[3],
                          int vectorWidth) {
  for (int i = 0; i < 3; i++) {
    const float ext = hi[i] - lo[i];
    half[i] = ext * 0.5f;
    inv[i] = host_rcp(half[i], vectorWidth);
    center[i] = lo[i] + half[i];
  }
}

// ------------------------------------------------------------------ DevBuf
// Growth is geometric (2x, at least 64 KiB): hipFree synchronises the device,
// so buffers whose size follows the region count must not reallocate often.
void DevBuf::ensure(size_t b, bool keep) {
  if (b <= bytes) return;
  size_t nb = std::max(std::max(b, bytes * 2), size_t(65536));
  void* np = nullptr;
  PGL_HIP_CHECK(hipMalloc(&np, nb));
  if (keep && p && bytes) PGL_HIP_CHECK(hipMemcpy(np, p, bytes, hipMemcpyDeviceToDevice));
  if (p) (void)hipFree(p);
  p = np;
  bytes = nb;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// ------------------------------------------------------------ KindField
float reprojection_min_distance(const float var[3]) {  // WEM:1184-1196, |.| with fused dot
  const float l = std::sqrt(std::fmaf(var[0], var[0], std::fmaf(var[1], var[1], var[2] * var[2])));
  return 3.f * 3.f * std::sqrt(l);
}

void uniform_directions(std::vector<float>& out) {  // WEM:668-700 spherical Fibonacci
  out.assign(32 * 32 * 3, 0.f);
  const float gr = 1.618033988749895f;
  for (int l = 0; l < 32; l++)
    for (int n = 0; n < 32; n++) {
      float* d = &out[(l * 32 + n) * 3];
      if (n < l + 1) {
        const float phi = 2.0f * float(M_PI) * ((float)n / gr);
        const float z = 1.0f - ((2.0f * n + 1.0f) / float(l + 1));
        const float theta = std::acos(z);
        d[0] = std::sin(theta) * std::cos(phi);
        d[1] = std::sin(theta) * std::sin(phi);
        d[2] = std::cos(theta);
      } else {
        d[0] = 0.f; d[1] = 0.f; d[2] = 1.f;
      }
    }
}

KindField::KindField(const KdSettings& kd, const FactoryConfig& fc, bool deterministic, int vectorWidth)
    : kd_(kd), fc_(fc), deterministic_(deterministic), vectorWidth_(vectorWidth) {
  if (!host_supports_width(vectorWidth))
    throw std::runtime_error("openpgl_amd: the host CPU lacks the instruction set of a SIMD-width-" +
                             std::to_string(vectorWidth) +
                             " device (its kd-tree reciprocals follow that device's arithmetic)");
  for (int i = 0; i < 3; i++) { sceneLo_[i] = sceneHi_[i] = treeLo_[i] = treeHi_[i] = 0.f; }
  kdDown_.coherent = true;
  for (int p = 0; p < 2; p++) {
    for (auto& e : evSp_[p]) PGL_HIP_CHECK(hipEventCreate(&e));
    PGL_HIP_CHECK(hipEventCreateWithFlags(&evDir_[p], hipEventDisableTiming));
    for (auto& e : evKnn_[p]) PGL_HIP_CHECK(hipEventCreate(&e));
  }
  PGL_HIP_CHECK(hipEventCreateWithFlags(&evKd_, hipEventDisableTiming));
  PGL_HIP_CHECK(hipStreamCreate(&sp_));
  std::vector<float> ud;
  uniform_directions(ud);
  dUniformDirs_.ensure(ud.size() * 4);
  PGL_HIP_CHECK(hipMemcpy(dUniformDirs_.p, ud.data(), ud.size() * 4, hipMemcpyHostToDevice));
}

KindField::~KindField() {
  drain();
  (void)hipStreamSynchronize(sp_);
  for (int p = 0; p < 2; p++) {
    for (auto& e : evSp_[p]) (void)hipEventDestroy(e);
    (void)hipEventDestroy(evDir_[p]);
    for (auto& e : evKnn_[p]) (void)hipEventDestroy(e);
  }
  (void)hipEventDestroy(evKd_);
  (void)hipStreamDestroy(sp_);
}

void KindField::setSceneBounds(const float lo[3], const float hi[3]) {
  for (int i = 0; i < 3; i++) { sceneLo_[i] = lo[i]; sceneHi_[i] = hi[i]; }
  boundsSet_ = true;
}

void KindField::sceneBounds(float lo[3], float hi[3]) const {
  for (int i = 0; i < 3; i++) { lo[i] = sceneLo_[i]; hi[i] = sceneHi_[i]; }
}

void KindField::reset() {  // Field::resetField (Field.h:222-232)
  drain();
  failed_ = false;
  failMsg_.clear();
  iteration_ = 0;
  boundsSet_ = false;
  initialized_ = false;
  nodes_.clear();
  regions_.clear();
  hostStates_.clear();
  hintRegions_ = 0;
  mirrorDirty_ = true;
}

// Field::estimateSceneBounds: AABB over samples[0, n-2] (the container's end()
// points at the last element), upper starting at FLT_MIN, scaled 3x about the centre.
void KindField::estimateSceneBounds(const float* px, const float* py, const float* pz, uint32_t n, hipStream_t s) {
  for (int i = 0; i < 3; i++) { sceneLo_[i] = kFltMax; sceneHi_[i] = std::numeric_limits<float>::min(); }
  boundsSet_ = false;
  if (n == 0) return;
  const uint32_t m = n - 1;
  if (m > 0) {
    dPartial_.ensure(1024 * 6 * sizeof(float));
    const uint32_t blocks = launch_bounds(px, py, pz, m, dPartial_.as<float>(), s);
    std::vector<float> part(blocks * 6);
    PGL_HIP_CHECK(hipMemcpyAsync(part.data(), dPartial_.p, blocks * 6 * sizeof(float), hipMemcpyDeviceToHost, s));
    PGL_HIP_CHECK(hipStreamSynchronize(s));
    for (uint32_t b = 0; b < blocks; b++)
      for (int i = 0; i < 3; i++) {
        sceneLo_[i] = std::min(sceneLo_[i], part[b * 6 + i]);
        sceneHi_[i] = std::max(sceneHi_[i], part[b * 6 + 3 + i]);
      }
  }
  for (int i = 0; i < 3; i++) {
    const float c = (sceneLo_[i] + sceneHi_[i]) * 0.5f;
    const float lo = c + 3.0f * (sceneLo_[i] - c);
    const float hi = c + 3.0f * (sceneHi_[i] - c);
    sceneLo_[i] = lo;
    sceneHi_[i] = hi;
  }
  boundsSet_ = true;
}

struct FrontierItem {
  uint32_t node, depth;
  float lo[3], hi[3];
};
struct FrontierItemList {
  std::vector<FrontierItem> items;
};

static uint8_t max_dim(const float v[3]) {  // KDTreeBuilder.h:223-226
  const int a = v[1] > v[0];
  return v[a] > v[2] ? uint8_t(a) : uint8_t(2);
}

static KDNode make_leaf(uint32_t data) { return KDNode{0.f, (3u << 30) | data}; }
static KDNode make_inner(uint32_t dim, float pivot, uint32_t left) { return KDNode{pivot, (dim << 30) | left}; }

// leaf visited with `count` new samples whose statistics are `st` (KDTreeBuilder.h:248-291)
void KindField::visitLeaf(uint32_t node, uint32_t depth, const float lo[3], const float hi[3], uint32_t count,
                          const HostSampleStats& st, FrontierItemList& next) {
  const uint32_t data = nodes_[node].idx();
  RegionHost& reg = regions_[data];
  if (depth < kd_.maxDepth && reg.ss.n + float(count) > float(kd_.maxSamples)) {
    HostSampleStats merged = reg.ss;
    merged.merge(st);
    float var[3];
    merged.variance(var);
    const uint8_t dim = max_dim(var);
    const float pos = merged.mean[dim];
    RegionHost right = reg;
    reg.ss.split(dim, pos, 0.25f, false);
    right.ss.split(dim, pos, 0.25f, true);
    reg.splitFlag = true;
    right.splitFlag = true;
    reg.hi[dim] = pos;
    right.lo[dim] = pos;
    right.copyFrom = reg.copyFrom >= 0 ? reg.copyFrom : (int32_t)data;
    right.count = 0;
    regions_.push_back(right);
    const uint32_t rightIdx = (uint32_t)regions_.size() - 1;
    if (splitRegion_.size() < nodes_.size() + 2) splitRegion_.resize(nodes_.size() + 2, -1);
    splitRegion_[node] = (int32_t)rightIdx;
    const uint32_t left = (uint32_t)nodes_.size();
    nodes_.push_back(make_leaf(data));
    nodes_.push_back(make_leaf(rightIdx));
    nodes_[node] = make_inner(dim, pos, left);
    FrontierItem it;
    it.node = node;
    it.depth = depth;
    for (int i = 0; i < 3; i++) { it.lo[i] = lo[i]; it.hi[i] = hi[i]; }
    next.items.push_back(it);
  } else {
    reg.ss.merge(st);
    reg.count = count;
  }
}

// Waits for the work queued on s so far by polling an event recorded behind
// it: the kd rounds' waits are on the update's critical path, and a blocking
// stream wait may hand the thread to the OS scheduler and wake it late.
static void spin_wait(hipEvent_t ev, hipStream_t s) {
  PGL_HIP_CHECK(hipEventRecord(ev, s));
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) PGL_HIP_CHECK(e);
  }
}

void KindField::kdUpdate(uint32_t n, hipStream_t s) {
  FrontierItemList frontier, next;
  uint32_t rounds = 1;
  // Leaf table of the current tree: every sample first descends through the
  // existing inner nodes to its leaf (those levels take no decisions), with
  // the leaf's statistics taken w.r.t. its parent's bounds.
  struct LeafInfo {
    uint32_t node, depth;
    float lo[3], hi[3];
    float center[3], half[3], invHalf[3];
  };
  std::vector<LeafInfo> leaves;
  std::vector<int32_t> leafSlot(nodes_.size(), -1);
  {
    struct Item { uint32_t node, depth; float lo[3], hi[3], plo[3], phi[3]; };
    std::vector<Item> stack;
    Item root;
    root.node = 0;
    root.depth = 1;
    for (int i = 0; i < 3; i++) { root.lo[i] = root.plo[i] = treeLo_[i]; root.hi[i] = root.phi[i] = treeHi_[i]; }
    stack.push_back(root);
    while (!stack.empty()) {
      const Item it = stack.back();
      stack.pop_back();
      const KDNode nd = nodes_[it.node];
      if (nd.isLeaf()) {
        LeafInfo li;
        li.node = it.node;
        li.depth = it.depth;
        for (int i = 0; i < 3; i++) { li.lo[i] = it.lo[i]; li.hi[i] = it.hi[i]; }
        bounds_params(it.plo, it.phi, li.center, li.half, li.invHalf, vectorWidth_);
        leafSlot[it.node] = (int32_t)leaves.size();
        leaves.push_back(li);
        continue;
      }
      const uint32_t d = nd.dim(), l = nd.idx();
      Item a, b;
      a.node = l; b.node = l + 1;
      a.depth = b.depth = it.depth + 1;
      for (int i = 0; i < 3; i++) {
        a.lo[i] = b.lo[i] = a.plo[i] = b.plo[i] = it.lo[i];
        a.hi[i] = b.hi[i] = a.phi[i] = b.phi[i] = it.hi[i];
      }
      a.hi[d] = nd.pivot;
      b.lo[d] = nd.pivot;
      stack.push_back(b);
      stack.push_back(a);
    }
  }
  const uint32_t L = (uint32_t)leaves.size();
  std::vector<float> params(6 * L);
  for (uint32_t i = 0; i < L; i++)
    for (int j = 0; j < 3; j++) { params[6 * i + j] = leaves[i].center[j]; params[6 * i + 3 + j] = leaves[i].invHalf[j]; }
  // the reductions write the statistics straight into coherent pinned memory
  kdDown_.ensure(std::max<size_t>(L, 2) * sizeof(IntStatsAcc));
  // one upload: nodes | leaf slots | leaf parameters (16-byte aligned parts)
  const size_t nb = nodes_.size() * sizeof(KDNode), sb = leafSlot.size() * sizeof(int32_t), pb = params.size() * 4;
  const size_t offSlot = (nb + 15) & ~size_t(15), offPar = (offSlot + sb + 15) & ~size_t(15);
  const size_t upB = offPar + std::max<size_t>(pb, sizeof(FrontierNode));
  kdUp_.ensure(upB);
  dKdUp_.ensure(upB);
  {
    char* up = kdUp_.as<char>();
    memcpy(up, nodes_.data(), nb);
    memcpy(up + offSlot, leafSlot.data(), sb);
    memcpy(up + offPar, params.data(), pb);
    PGL_HIP_CHECK(hipMemcpyAsync(dKdUp_.p, up, offPar + pb, hipMemcpyHostToDevice, s));
  }
  char* const kd = dKdUp_.as<char>();
  dStatPartial_.ensure(spatial_partial_bytes(n));
  launch_descend_stats(dPx_.as<float>(), dPy_.as<float>(), dPz_.as<float>(), n, (const KDNode*)kd,
                       (uint32_t)nodes_.size(), (const int32_t*)(kd + offSlot), kd + offPar, L, dNodeOf_.as<uint32_t>(),
                       dStatPartial_.as<IntStatsAcc>(), (IntStatsAcc*)kdDown_.dev, s);
  {
    const IntStatsAcc* acc = kdDown_.as<IntStatsAcc>();
    spin_wait(evKd_, s);
    for (uint32_t i = 0; i < L; i++) {
      if (acc[i].n == 0) continue;
      const LeafInfo& li = leaves[i];
      const HostSampleStats st = int_stats_to_float(acc[i], li.center, li.half);
      visitLeaf(li.node, li.depth, li.lo, li.hi, (uint32_t)acc[i].n, st, frontier);
    }
  }
  std::vector<FrontierNode> fn;
  std::vector<int32_t> slot;
  std::vector<float> halves;
  while (!frontier.items.empty()) {
    rounds++;
    const uint32_t F = (uint32_t)frontier.items.size();
    fn.resize(F);
    halves.resize(3 * F);
    slot.assign(nodes_.size(), -1);
    for (uint32_t f = 0; f < F; f++) {
      const FrontierItem& it = frontier.items[f];
      const KDNode nd = nodes_[it.node];
      fn[f].node = it.node;
      fn[f].left = nd.idx();
      fn[f].dim = nd.dim();
      fn[f].pivot = nd.pivot;
      bounds_params(it.lo, it.hi, fn[f].center, &halves[3 * f], fn[f].invHalf, vectorWidth_);
      slot[it.node] = (int32_t)f;
    }
    kdDown_.ensure(2 * F * sizeof(IntStatsAcc));
    // one upload: frontier nodes | slots
    const size_t fb = F * sizeof(FrontierNode), rsb = slot.size() * sizeof(int32_t);
    const size_t offS = (fb + 15) & ~size_t(15);
    kdUp_.ensure(offS + rsb);
    dKdUp_.ensure(offS + rsb);
    {
      char* up = kdUp_.as<char>();
      memcpy(up, fn.data(), fb);
      memcpy(up + offS, slot.data(), rsb);
      PGL_HIP_CHECK(hipMemcpyAsync(dKdUp_.p, up, offS + rsb, hipMemcpyHostToDevice, s));
    }
    char* const kr = dKdUp_.as<char>();
    launch_kd_round(dPx_.as<float>(), dPy_.as<float>(), dPz_.as<float>(), n, dNodeOf_.as<uint32_t>(),
                    (const int32_t*)(kr + offS), (uint32_t)slot.size(), (const FrontierNode*)kr, F,
                    dStatPartial_.as<IntStatsAcc>(), (IntStatsAcc*)kdDown_.dev, s);
    const IntStatsAcc* acc = kdDown_.as<IntStatsAcc>();
    spin_wait(evKd_, s);
    next.items.clear();
    for (uint32_t f = 0; f < F; f++) {
      const FrontierItem& it = frontier.items[f];
      const FrontierNode& p = fn[f];
      for (uint32_t side = 0; side < 2; side++) {
        const IntStatsAcc& a = acc[2 * f + side];
        if (a.n == 0) continue;  // empty range: updateTreeNode returns (KDTreeBuilder.h:235-238)
        const uint32_t child = p.left + side;
        float lo[3], hi[3];
        for (int i = 0; i < 3; i++) { lo[i] = it.lo[i]; hi[i] = it.hi[i]; }
        if (side == 0) hi[p.dim] = p.pivot; else lo[p.dim] = p.pivot;
        if (nodes_[child].isLeaf()) {
          const HostSampleStats st = int_stats_to_float(a, p.center, &halves[3 * f]);
          visitLeaf(child, it.depth + 1, lo, hi, (uint32_t)a.n, st, next);
        } else {
          FrontierItem c;
          c.node = child;
          c.depth = it.depth + 1;
          for (int i = 0; i < 3; i++) { c.lo[i] = lo[i]; c.hi[i] = hi[i]; }
          next.items.push_back(c);
        }
      }
    }
    std::swap(frontier, next);
  }
  lastKdRounds_ = rounds;
}

void KindField::ensureRegionCapacity(uint32_t r, hipStream_t s) {
  if (r <= stateCapacity_) return;
  uint32_t cap = std::max<uint32_t>(r, stateCapacity_ * 2);
  // reallocation synchronises the stream: start with room for 4096 regions
  // (16 MiB) so that a growing field does not pay it every few updates
  cap = std::max<uint32_t>(cap, 4096);
  DevBuf nb;
  nb.ensure(size_t(cap) * sizeof(RegionState));
  PGL_HIP_CHECK(hipMemsetAsync(nb.p, 0, size_t(cap) * sizeof(RegionState), s));
  if (stateCapacity_ > 0)
    PGL_HIP_CHECK(hipMemcpyAsync(nb.p, dStates_.p, size_t(stateCapacity_) * sizeof(RegionState),
                                 hipMemcpyDeviceToDevice, s));
  PGL_HIP_CHECK(hipStreamSynchronize(s));
  std::swap(dStates_.p, nb.p);
  std::swap(dStates_.bytes, nb.bytes);
  stateCapacity_ = cap;
}

// The serial reference recursion (KDTreeBuilder.h:346-349 with one thread)
// creates child node pairs and right-hand regions in depth-first pre-order of
// the split events.  The level-synchronous push-down creates them level by
// level; renumber the ones created in this update into the reference order so
// node and region indices (and therefore GetId) match the reference.
void KindField::renumberDepthFirst(uint32_t nodesBefore, uint32_t regionsBefore) {
  const uint32_t N = (uint32_t)nodes_.size(), R = (uint32_t)regions_.size();
  if (N == nodesBefore) { keyNodes_ = nodes_; return; }
  std::vector<uint32_t> nodeMap(N), regionMap(R);
  for (uint32_t i = 0; i < N; i++) nodeMap[i] = i;
  for (uint32_t i = 0; i < R; i++) regionMap[i] = i;
  uint32_t nextNode = nodesBefore, nextRegion = regionsBefore;
  splitRegion_.resize(N, -1);
  std::vector<uint32_t> stack(1, 0u);
  while (!stack.empty()) {
    const uint32_t nd = stack.back();
    stack.pop_back();
    const KDNode& x = nodes_[nd];
    if (x.isLeaf()) continue;
    const uint32_t l = x.idx();
    if (l >= nodesBefore) {
      nodeMap[l] = nextNode;
      nodeMap[l + 1] = nextNode + 1;
      nextNode += 2;
    }
    if (splitRegion_[nd] >= 0) regionMap[(uint32_t)splitRegion_[nd]] = nextRegion++;
    stack.push_back(l + 1);
    stack.push_back(l);
  }
  // leaf-key view: push-down node numbering, final region ids
  keyNodes_ = nodes_;
  for (auto& k : keyNodes_)
    if (k.isLeaf()) k.bits = (3u << 30) | regionMap[k.idx()];
  std::vector<KDNode> nn(N);
  for (uint32_t i = 0; i < N; i++) {
    KDNode k = nodes_[i];
    k.bits = k.isLeaf() ? ((3u << 30) | regionMap[k.idx()]) : ((k.bits & 0xC0000000u) | nodeMap[k.idx()]);
    nn[nodeMap[i]] = k;
  }
  nodes_.swap(nn);
  std::vector<RegionHost> nr(R);
  for (uint32_t i = 0; i < R; i++) nr[regionMap[i]] = regions_[i];
  regions_.swap(nr);
}

// Field::updateField (Field.h:187-220) for records already in device memory.
//
// Pipelined over consecutive updates: the spatial phase (ingest, kd update
// with its host decisions, leaf keys, sort, gather + reprojection) runs on
// the field's spatial stream and writes one of two region-ordered sample
// buffers; the directional phase (state copies and inputs, the region kernel)
// is queued on the field's stream `s` behind an event of the gather, and
// behind the previous update's directional phase, which may still be running
// then.  After queueing it the call finishes the previous update (waits for
// it, collects its statistics) and returns; every read of device state
// finishes the pending update.
// Buffers whose size follows the number of regions and nodes are sized for
// at least kReserveRegions regions (and twice as many nodes) at once, and for
// twice the current count beyond that.  A buffer that grows is freed, and
// freeing device or pinned memory waits for the whole device: in the middle
// of an update that would stall the overlap of this update's spatial phase
// with the previous update's region kernel.
static const uint32_t kReserveRegions = 4096;

void KindField::reserve(uint32_t n, uint32_t regions) {
  const uint32_t Rc = std::max<uint32_t>(kReserveRegions, 2 * regions);
  const size_t nodes = 2 * size_t(Rc);
  const size_t acc = 2 * size_t(Rc) * sizeof(IntStatsAcc);
  kdDown_.ensure(acc);
  const size_t up = nodes * sizeof(KDNode) + nodes * sizeof(int32_t) + 6 * size_t(Rc) * 4 + 4 * 16 +
                    2 * size_t(Rc) * sizeof(FrontierNode);
  kdUp_.ensure(up);
  dKdUp_.ensure(up);
  const size_t tab = nodes * sizeof(KDNode) + size_t(Rc) * 4 + 2 * size_t(Rc) * sizeof(float4) + 64;
  for (int p = 0; p < 2; p++) {
    pin_[p][0].ensure(tab);
    dTables_[p].ensure(tab);
  }
  const uint32_t Rs = std::min<uint32_t>(Rc, kSortMaxKeys);
  dCountSortTmp_.ensure(counting_sort_temp_bytes(n, Rs));
  if (deterministic_) {
    dDetTmp_.ensure(det_sort_temp_bytes(n, Rc));
    for (int p = 0; p < 2; p++) detPin_[p].ensure(det_sort_stage_bytes(n, Rc));
  }
  dNbIds_.ensure(size_t(Rc) * 8 * 4);
  dNbPts_.ensure(size_t(Rc) * 8 * sizeof(float4));
  dNodes_.ensure(nodes * sizeof(KDNode));
}

void KindField::update(const void* deviceSamples, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  if (failed_) finish();  // throws: a failed field takes no update until reset()
  const auto t0 = std::chrono::steady_clock::now();
  reserve(n, (uint32_t)regions_.size());
  const int P = parity_;
  UpdateInfo inf;
  inf.samples = n;
  if (n > kMaxUpdateSamples) throw std::runtime_error("openpgl_amd: at most 2^30 samples per update");
  const size_t n4 = size_t(n) * 4;
  dPx_.ensure(n4); dPy_.ensure(n4); dPz_.ensure(n4); dNodeOf_.ensure(n4);
  dKeys_.ensure(n4); dVals_.ensure(n4); dKeys2_.ensure(n4); dVals2_.ensure(n4);
  for (int j = 0; j < 6; j++) dSoa_[P][j].ensure(n4);
  const size_t sortBytes = sort_pairs_temp_bytes(n);
  dSortTmp_.ensure(sortBytes);
  hipStream_t sp = sp_;

  // ---------------------------------------------------------------- spatial phase
  PGL_HIP_CHECK(hipEventRecord(evSp_[P][0], sp));
  if (deterministic_) {
    dWkey_.ensure(n4);
    dWkeySorted_.ensure(n4);
  }
  launch_ingest(deviceSamples, n, dPx_.as<float>(), dPy_.as<float>(), dPz_.as<float>(),
                deterministic_ ? dWkey_.as<uint32_t>() : nullptr, sp);
  const bool first = !initialized_;
  if (first) {
    if (!boundsSet_) estimateSceneBounds(dPx_.as<float>(), dPy_.as<float>(), dPz_.as<float>(), n, sp);
    for (int i = 0; i < 3; i++) { treeLo_[i] = sceneLo_[i]; treeHi_[i] = sceneHi_[i]; }
    nodes_.assign(1, make_leaf(0));
    regions_.assign(1, RegionHost());
    for (int i = 0; i < 3; i++) { regions_[0].lo[i] = sceneLo_[i]; regions_[0].hi[i] = sceneHi_[i]; }
  }
  for (auto& r : regions_) { r.count = 0; r.copyFrom = -1; }
  const uint32_t regionsBefore = (uint32_t)regions_.size();
  const uint32_t nodesBefore = first ? 1u : (uint32_t)nodes_.size();
  splitRegion_.assign(nodes_.size(), -1);
  kdUpdate(n, sp);
  inf.kdRounds = lastKdRounds_;
  renumberDepthFirst(nodesBefore, regionsBefore);
  const uint32_t R = (uint32_t)regions_.size();
  PGL_HIP_CHECK(hipEventRecord(evSp_[P][1], sp));

  // host tables of this update: every region's inputs and pivots, the sort's
  // region starts and the KNN pivots, uploaded with the leaf-key nodes at once
  hInputs_.assign(R, RegionInput());
  hPivots_.assign(R, make_float4(0.f, 0.f, 0.f, 0.f));
  hWork_.clear();
  hBegin_.resize(R);
  hCount_.resize(R);
  uint32_t begin = 0;
  for (uint32_t r = 0; r < R; r++) {
    const RegionHost& rg = regions_[r];
    RegionInput& in = hInputs_[r];
    in.region = r;
    for (int i = 0; i < 3; i++) { in.ssMean[i] = rg.ss.mean[i]; in.ssVar[i] = rg.ss.var[i]; }
    in.ssN = rg.ss.n;
    in.begin = begin;
    in.count = rg.count;
    in.mode = first ? 1 : 2;
    in.splitFlag = rg.splitFlag ? 1 : 0;
    in.initK = 0;
    in.copyFrom = r >= regionsBefore ? rg.copyFrom : -1;  // new regions start from their ancestor's state
    hBegin_[r] = begin;
    hCount_[r] = rg.count;
    begin += rg.count;
    float var[3];
    rg.ss.variance(var);
    hPivots_[r] = make_float4(rg.ss.mean[0], rg.ss.mean[1], rg.ss.mean[2], reprojection_min_distance(var));
    // teams sized by the region's samples (the number of EM iterations a region
    // needs is not predictable from its last update)
    if (first || rg.count > 0 || rg.splitFlag) hWork_.push_back(RegionWork{r, rg.count, (float)rg.count});
  }
  if (begin != n) throw std::runtime_error("kd update lost samples");
  if (kd_.knnLookup) {
    hKnnPiv_.resize(R);
    for (uint32_t r = 0; r < R; r++)
      hKnnPiv_[r] = make_float4(regions_[r].ss.mean[0], regions_[r].ss.mean[1], regions_[r].ss.mean[2], 0.f);
  }
  const size_t keyB = keyNodes_.size() * sizeof(KDNode), pivB = size_t(R) * sizeof(float4);
  const size_t offBegin = (keyB + 15) & ~size_t(15), offPiv = (offBegin + size_t(R) * 4 + 15) & ~size_t(15);
  const size_t offKnn = offPiv + pivB, tabB = offKnn + (kd_.knnLookup ? pivB : 0);
  pin_[P][0].ensure(tabB);
  dTables_[P].ensure(tabB);
  {
    char* h = pin_[P][0].as<char>();
    memcpy(h, keyNodes_.data(), keyB);
    memcpy(h + offBegin, hBegin_.data(), size_t(R) * 4);
    memcpy(h + offPiv, hPivots_.data(), pivB);
    if (kd_.knnLookup) memcpy(h + offKnn, hKnnPiv_.data(), pivB);
    PGL_HIP_CHECK(hipMemcpyAsync(dTables_[P].p, h, tabB, hipMemcpyHostToDevice, sp));
  }
  char* const tab = dTables_[P].as<char>();
  knnPivots_ = (const float4*)(tab + offKnn);

  // region-contiguous order: stable sort of (region, sample index)
  const void* gatherRecs = deviceSamples;  // records the gather (and the deterministic sort) read
  const uint32_t* gatherIdx = dVals2_.as<uint32_t>();
  if (R <= kSortMaxKeys) {
    // stable counting sort keyed by the region of each sample's leaf (every
    // region's first position is known from the kd counts), moving the
    // records themselves into region order: the passes after it read each
    // region from one contiguous range instead of gathering 40-byte records
    // from all over the input
    dCountSortTmp_.ensure(counting_sort_temp_bytes(n, R));
    dRecs_.ensure(size_t(n) * kSampleRecordBytes);
    counting_sort_leaf_records(dCountSortTmp_.p, dNodeOf_.as<uint32_t>(), (const KDNode*)tab,
                               (const uint32_t*)(tab + offBegin), R, n, deviceSamples, dRecs_.p,
                               dKeys2_.as<uint32_t>(), deterministic_ ? dVals2_.as<uint32_t>() : nullptr,
                               deterministic_ ? dWkey_.as<uint32_t>() : nullptr,
                               deterministic_ ? dWkeySorted_.as<uint32_t>() : nullptr, sp);
    gatherRecs = dRecs_.p;
    gatherIdx = deterministic_ ? dVals2_.as<uint32_t>() : nullptr;
  } else {
    launch_leaf_keys(dNodeOf_.as<uint32_t>(), (const KDNode*)tab, n, dKeys_.as<uint32_t>(), dVals_.as<uint32_t>(),
                     sp);
    int endBit = 1;
    while ((1u << endBit) < R && endBit < 32) endBit++;
    sort_pairs(dSortTmp_.p, sortBytes, dKeys_.as<uint32_t>(), dKeys2_.as<uint32_t>(), dVals_.as<uint32_t>(),
               dVals2_.as<uint32_t>(), n, endBit, sp);
  }
  // deterministic mode: every region's samples in SampleDataLess order
  // (Field.h:409-412, 469-472)
  if (deterministic_) {
    dDetTmp_.ensure(det_sort_temp_bytes(n, R));
    detPin_[P].ensure(det_sort_stage_bytes(n, R));
    // the weight keys travelled with the counting sort (above 4096 regions
    // the tiles read the weights from the records)
    det_sort_segments(dDetTmp_.p, detPin_[P].p, detPin_[P].bytes, gatherRecs, hBegin_.data(), hCount_.data(), R,
                      n, dVals2_.as<uint32_t>(), R <= kSortMaxKeys ? dWkeySorted_.as<uint32_t>() : nullptr, sp);
  }
  PGL_HIP_CHECK(hipEventRecord(evSp_[P][2], sp));

  // gather into the region-ordered SoA, reprojected to the region pivots
  float* sv[6];
  for (int j = 0; j < 6; j++) sv[j] = dSoa_[P][j].as<float>();
  float* const* soa = sv;
  launch_gather_reproject(gatherRecs, dKeys2_.as<uint32_t>(), gatherIdx, n,
                          (const float4*)(tab + offPiv), soa[0], soa[1], soa[2], soa[3], soa[4], soa[5],
                          fc_.noParallax == 0, sp);
  PGL_HIP_CHECK(hipEventRecord(evSp_[P][3], sp));
  lastGatherParity_ = P;
  // the directional launch's host work (jobs, teams, pinned tables) while the
  // previous update's region kernel still runs
  RegionRunner& rr = runner();
  rr.stage(hWork_, RegionSamplesView{soa[0], soa[1], soa[2], soa[3], soa[4], soa[5]}, fc_,
           dUniformDirs_.as<float>(), nullptr, hInputs_.data(), R);

  // ---------------------------------------------------------- directional phase
  // queued on s behind the previous update's directional phase (which may
  // still run: the stream starts this launch when that one ends, without a
  // round trip through the host); waits for this update's gather
  if (first) {
    ensureRegionCapacity(1, s);
    PGL_HIP_CHECK(hipMemsetAsync(dStates_.p, 0, sizeof(RegionState), s));
  }
  ensureRegionCapacity(R, s);
  PGL_HIP_CHECK(hipStreamWaitEvent(s, evSp_[P][3], 0));
  // the new regions' state copies, the inputs and the launch (one upload, one preparing kernel)
  rr.submit(dStates_.as<RegionState>(), s);
  const int slot = rr.lastSlot();
  inf.regionLaunches = rr.lastLaunches;
  inf.regionJobs = rr.lastJobs;
  inf.regionItems = rr.lastItems;
  inf.maxTeam = rr.lastMaxTeam;
  inf.regionGrid = rr.lastGrid;
  PGL_HIP_CHECK(hipEventRecord(evDir_[P], s));

  // KNN neighbourhoods of the region pivots (sample-statistics means, uploaded
  // with the spatial tables) and the tree in its final numbering for the query
  // kernels: they need only this update's spatial results, so they run on the
  // spatial stream beside the region kernels, queued after the directional
  // launch; every reader of either first finishes the pending update, which
  // waits for them
  PGL_HIP_CHECK(hipEventRecord(evKnn_[P][0], sp));
  if (kd_.knnLookup) {
    dNbIds_.ensure(size_t(R) * 8 * 4);
    dNbPts_.ensure(size_t(R) * 8 * sizeof(float4));
    launch_knn8(knnPivots_, R, dNbIds_.as<uint32_t>(), dNbPts_.as<float4>(), sp);
  }
  dNodes_.ensure(nodes_.size() * sizeof(KDNode));
  PGL_HIP_CHECK(hipMemcpyAsync(dNodes_.p, pinned(pin_[P][5], nodes_.data(), nodes_.size()),
                               nodes_.size() * sizeof(KDNode), hipMemcpyHostToDevice, sp));
  PGL_HIP_CHECK(hipEventRecord(evKnn_[P][1], sp));


  // ------------------------------------------------ finish the previous update
  try {
    finish();
  } catch (...) {
    // the previous update failed: the field is invalid until reset(); this
    // update's queued work still runs to its end before anything is released
    (void)hipEventSynchronize(evDir_[P]);
    (void)hipEventSynchronize(evKnn_[P][1]);
    throw;
  }

  for (auto& r : regions_) { r.splitFlag = false; r.copyFrom = -1; }
  initialized_ = true;
  iteration_++;
  mirrorDirty_ = true;
  pendingInfo_ = inf;
  pendingCopyMs_ = copyMs_;
  copyMs_ = 0.f;
  pendingRegions_ = R;
  pendingT0_ = t0;
  pendingParity_ = P;
  pendingSlot_ = slot;
  pending_ = true;
  parity_ = 1 - P;
}

void KindField::finish() const {
  if (failed_)
    throw std::runtime_error("openpgl_amd: " + failMsg_ + "; the field is invalid until pglFieldReset");
  finishPending();
}

void KindField::drain() const noexcept {
  try {
    finishPending();
  } catch (...) {
  }
}

// Completes the queued directional phase: waits for it, checks the region
// kernel's error flag and records the update's figures.  A failed update
// leaves region states that no later call may use: the field is marked
// failed until reset().
void KindField::finishPending() const {
  if (!pending_) return;
  const int P = pendingParity_;
  PGL_HIP_CHECK(hipEventSynchronize(evDir_[P]));
  PGL_HIP_CHECK(hipEventSynchronize(evKnn_[P][1]));
  pending_ = false;
  RegionRunner& rr = *runner_;
  hintRegions_ = pendingRegions_;
  UpdateInfo& i = pendingInfo_;
  const RegionLaunchStats& ls = rr.stats(pendingSlot_);
  i.estepEvals = (double)ls.estepEvals;
  i.auxEvals = (double)ls.distEvals;
  i.splitEvals = (double)ls.splitEvals;
  i.emIterations = (double)ls.emIterations;
  float ms = 0.f;
  PGL_HIP_CHECK(hipEventElapsedTime(&ms, evSp_[P][0], evSp_[P][1])); i.spatialMs = ms;
  PGL_HIP_CHECK(hipEventElapsedTime(&ms, evSp_[P][1], evSp_[P][2])); i.sortMs = ms;
  PGL_HIP_CHECK(hipEventElapsedTime(&ms, evSp_[P][2], evSp_[P][3])); i.gatherMs = ms;

  PGL_HIP_CHECK(hipEventElapsedTime(&ms, evKnn_[P][0], evKnn_[P][1])); i.knnMs = ms;
  i.regionKernelMs = rr.kernelMs(pendingSlot_);
  i.directionalMs = i.regionKernelMs;
  i.regions = pendingRegions_;
  i.wallMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - pendingT0_).count();
  info_ = i;
  lastTimes_[0] = (float)i.wallMs;
  lastTimes_[1] = pendingCopyMs_;  // 0 for records already in device memory
  lastTimes_[2] = (float)(i.spatialMs + i.sortMs + i.gatherMs);
  lastTimes_[3] = (float)(i.directionalMs + i.knnMs);
  history_.push_back(i);
  if (history_.size() > 256) history_.erase(history_.begin());
  try {
    rr.check(pendingSlot_);
  } catch (const std::exception& e) {
    failed_ = true;
    failMsg_ = std::string("the last update failed (") + e.what() + ")";
    throw;
  }
}

void KindField::updateFromHost(const void* hostSamples, uint32_t n) {
  if (n == 0) return;
  finish();  // a failed field takes no update (and the staging buffer is free)
  const auto t0 = std::chrono::steady_clock::now();
  dAos_.ensure(size_t(n) * kSampleRecordBytes);
  // on the spatial stream, after the previous update's gather read the staging buffer
  PGL_HIP_CHECK(hipMemcpyAsync(dAos_.p, hostSamples, size_t(n) * kSampleRecordBytes, hipMemcpyHostToDevice, sp_));
  PGL_HIP_CHECK(hipStreamSynchronize(sp_));
  copyMs_ = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

KindField::Statistics KindField::statistics(hipStream_t s) const {
  syncMirror(s);
  Statistics st;
  st.numRegions = regions_.size();
  st.regionsReserved = std::max<size_t>(stateCapacity_, st.numRegions);
  st.sizePerRegion = sizeof(RegionState);
  st.timeUpdate = lastTimes_[0];
  st.timeCopySamples = lastTimes_[1];
  st.timeSpatial = lastTimes_[2];
  st.timeDirectional = lastTimes_[3];
  // KDTree::calculateMaxDepth (KDTree.h:651-662): a leaf counts 1
  if (!nodes_.empty()) {
    std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 1u}};
    while (!stack.empty()) {
      const auto [n, d] = stack.back();
      stack.pop_back();
      if (nodes_[n].isLeaf()) {
        st.kdMaxDepth = std::max<size_t>(st.kdMaxDepth, d);
      } else {
        stack.push_back({nodes_[n].idx(), d + 1});
        stack.push_back({nodes_[n].idx() + 1, d + 1});
      }
    }
  }
  st.kdNodes = nodes_.size();
  st.kdNodesReserved = std::max(nodes_.capacity(), dNodes_.bytes / sizeof(KDNode));
  st.kdSizePerNode = sizeof(KDNode);
  st.sizePerDistribution = sizeof(VMMParams);
  // Field::getStatistics: float accumulation in region order, then the mean
  // and the square root of the mean square (0 / 0 for an empty field)
  float mn = 1e10f, mx = 0.f, avg = 0.f, sec = 0.f;
  for (size_t r = 0; r < st.numRegions; r++) {
    const int k = (int)hostStates_[r].vmm.K;
    mn = std::min(mn, (float)k);
    mx = std::max(mx, (float)k);
    avg += k;
    sec += k * k;
  }
  avg /= float(st.numRegions);
  sec /= float(st.numRegions);
  st.minK = mn;
  st.maxK = mx;
  st.averageK = avg;
  st.secondMomentK = std::sqrt(sec);
  return st;
}

void PinnedBuf::ensure(size_t b) {
  if (b <= bytes) return;
  const size_t nb = std::max(std::max(b, bytes * 2), size_t(65536));
  void* np = nullptr;
  PGL_HIP_CHECK(hipHostMalloc(&np, nb, coherent ? hipHostMallocCoherent : hipHostMallocDefault));
  if (p) {
    std::memcpy(np, p, bytes);
    (void)hipHostFree(p);
  }
  p = np;
  bytes = nb;
  dev = nullptr;
  if (coherent) PGL_HIP_CHECK(hipHostGetDevicePointer(&dev, p, 0));
}

PinnedBuf::~PinnedBuf() {
  if (p) (void)hipHostFree(p);
}

RegionRunner& KindField::runner() {
  if (!runner_) runner_.reset(new RegionRunner());
  return *runner_;
}

void KindField::syncMirror(hipStream_t s) const {
  finish();
  std::lock_guard<std::mutex> lock(mirrorMutex_);
  if (!mirrorDirty_) return;
  const uint32_t R = (uint32_t)regions_.size();
  hostStates_.resize(R);
  if (R) PGL_HIP_CHECK(hipMemcpyAsync(hostStates_.data(), dStates_.p, size_t(R) * sizeof(RegionState), hipMemcpyDeviceToHost, s));
  if (kd_.knnLookup && R) {
    hostNbIds_.resize(size_t(R) * 8);
    hostNbPts_.resize(size_t(R) * 8);
    PGL_HIP_CHECK(hipMemcpyAsync(hostNbIds_.data(), dNbIds_.p, size_t(R) * 8 * 4, hipMemcpyDeviceToHost, s));
    PGL_HIP_CHECK(hipMemcpyAsync(hostNbPts_.data(), dNbPts_.p, size_t(R) * 8 * sizeof(float4), hipMemcpyDeviceToHost, s));
  }
  PGL_HIP_CHECK(hipStreamSynchronize(s));
  mirrorDirty_ = false;
}

int KindField::lookupRegion(const float p[3], float* u, uint32_t& id) const {
  if (iteration_ == 0 || !initialized_) return -1;
  for (int j = 0; j < 3; j++)
    if (!(p[j] >= treeLo_[j] && p[j] <= treeHi_[j])) return -1;
  uint32_t i = 0;
  while (!nodes_[i].isLeaf()) {
    const uint32_t d = nodes_[i].dim();
    i = nodes_[i].idx() + (p[d] >= nodes_[i].pivot ? 1u : 0u);
  }
  const uint32_t leaf = nodes_[i].idx();
  id = leaf;
  if (kd_.knnLookup && *u >= 0.f) {
    uint32_t ids[8];
    float px[8], py[8], pz[8];
    for (int k = 0; k < 8; k++) {
      ids[k] = hostNbIds_[leaf * 8 + k];
      px[k] = hostNbPts_[leaf * 8 + k].x;
      py[k] = hostNbPts_[leaf * 8 + k].y;
      pz[k] = hostNbPts_[leaf * 8 + k].z;
    }
    const uint32_t nbSize = std::min<uint32_t>(8, (uint32_t)regions_.size());
    return (int)knn_pick(ids, px, py, pz, nbSize, p[0], p[1], p[2], u, vectorWidth_ != 4);
  }
  return (int)leaf;
}

QueryField KindField::queryView() const {
  finish();
  QueryField q;
  q.nodes = dNodes_.as<KDNode>();
  q.regions = dStates_.as<RegionState>();
  q.nbIds = dNbIds_.as<uint32_t>();
  q.nbPts = dNbPts_.as<float4>();
  q.nbSize = std::min<uint32_t>(8, (uint32_t)regions_.size());
  q.knn = kd_.knnLookup ? 1 : 0;
  q.trained = (initialized_ && iteration_ > 0) ? 1 : 0;
  q.parallax = fc_.noParallax ? 0 : 1;
  q.knnFused = vectorWidth_ != 4 ? 1 : 0;
  for (int i = 0; i < 3; i++) { q.lo[i] = treeLo_[i]; q.hi[i] = treeHi_[i]; }
  return q;
}

void KindField::canonicalOrder(std::vector<KDNode>& out, std::vector<uint32_t>& order) const {
  out.clear();
  order.clear();
  if (nodes_.empty()) return;
  // iterative depth-first traversal with the reference's numbering (KDTree.h:256-283)
  out.push_back(nodes_[0]);
  struct Item { uint32_t src, dst; };
  std::vector<Item> stack;
  stack.push_back({0, 0});
  while (!stack.empty()) {
    const Item it = stack.back();
    stack.pop_back();
    const KDNode n = nodes_[it.src];
    if (!n.isLeaf()) {
      const uint32_t l = n.idx();
      const uint32_t ni = (uint32_t)out.size();
      out[it.dst].bits = (out[it.dst].bits & 0xC0000000u) | ni;
      out.push_back(nodes_[l]);
      out.push_back(nodes_[l + 1]);
      stack.push_back({l + 1, ni + 1});
      stack.push_back({l, ni});
    } else {
      out[it.dst] = make_leaf((uint32_t)order.size());
      order.push_back(n.idx());
    }
  }
}

void export_state(const RegionState& S, PGLRegionRecordHost* o) {
  memset(o, 0, sizeof(*o));
  o->K = S.vmm.K;
  for (int k = 0; k < 32; k++) {
    o->weights[k] = S.vmm.w[k]; o->kappas[k] = S.vmm.kappa[k];
    o->meanDirections[k][0] = S.vmm.mux[k]; o->meanDirections[k][1] = S.vmm.muy[k]; o->meanDirections[k][2] = S.vmm.muz[k];
    o->normalizations[k] = S.vmm.norm[k]; o->eMinus2Kappa[k] = S.vmm.em2k[k]; o->meanCosines[k] = S.vmm.mc[k];
    o->distances[k] = S.vmm.dist[k];
    o->sumOfWeightedDirections[k][0] = S.suff.swdx[k]; o->sumOfWeightedDirections[k][1] = S.suff.swdy[k];
    o->sumOfWeightedDirections[k][2] = S.suff.swdz[k];
    o->sumOfWeightedStats[k] = S.suff.sw[k]; o->sumOfDistanceWeights[k] = S.suff.sdw[k];
    o->chiSquareMCEstimates[k] = S.split.chi[k];
    o->splitCovariances[k][0] = S.split.covx[k]; o->splitCovariances[k][1] = S.split.covy[k];
    o->splitCovariances[k][2] = S.split.covz[k];
    o->splitNumSamples[k] = S.split.sn[k]; o->splitSumWeights[k] = S.split.ssw[k]; o->splitSumAssignedSamples[k] = S.split.ssa[k];
  }
  for (int j = 0; j < 3; j++) o->pivot[j] = S.vmm.pivot[j];
  o->sumWeights = S.suff.sumWeights; o->numSamples = S.suff.numSamples; o->overallNumSamples = S.suff.overall;
  o->suffK = S.suff.K; o->normalized = S.suff.normalized; o->splitK = S.split.K;
  o->numSamplesAfterLastSplit = ((uint64_t)S.afterSplitHi << 32) | S.afterSplitLo;
  o->numSamplesAfterLastMerge = ((uint64_t)S.afterMergeHi << 32) | S.afterMergeLo;
  o->valid = S.valid;
  o->numSamplesThisUpdate = S.sampleCount;
  o->emIterations = S.emIterations; o->partialEmIterations = S.partialEmIterations;
  o->numSplits = S.numSplits; o->numMerges = S.numMerges;
}

void KindField::exportRegion(uint32_t r, PGLRegionRecordHost* o) const {
  const RegionState& S = hostStates_[r];
  const RegionHost& h = regions_[r];
  export_state(S, o);
  for (int j = 0; j < 3; j++) {
    o->ssMean[j] = h.ss.mean[j]; o->ssVariance[j] = h.ss.var[j]; o->ssLower[j] = h.ss.lo[j]; o->ssUpper[j] = h.ss.hi[j];
    o->regionLower[j] = h.lo[j]; o->regionUpper[j] = h.hi[j];
  }
  o->ssNumSamples = h.ss.n;
  o->splitFlag = h.splitFlag;
}

void KindField::knnHost(std::vector<uint32_t>& ids) const { ids = hostNbIds_; }

bool KindField::validate() const {  // Field::isValid (Field.h:307-317)
  for (uint32_t r = 0; r < (uint32_t)regions_.size(); r++) {
    if (!hostStates_[r].valid) return false;
  }
  return true;
}

GuidingField::GuidingField(const KdSettings& kd, const FactoryConfig& fc, bool deterministic, int vectorWidth)
    : surface(kd, fc, deterministic, vectorWidth), volume(kd, fc, deterministic, vectorWidth) {
  PGL_HIP_CHECK(hipStreamCreate(&stream));
  dHG.ensure(128 * sizeof(HGRep));
  PGL_HIP_CHECK(hipMemcpy(dHG.p, host_hg_table(), 128 * sizeof(HGRep), hipMemcpyHostToDevice));
}

static const float kHGTable[128 * 18] = {
#include "hg_table.inc"
};

const HGRep* host_hg_table() {
  static_assert(sizeof(HGRep) == 18 * sizeof(float), "HGRep row is 18 floats");
  return reinterpret_cast<const HGRep*>(kHGTable);
}

}  // namespace pglgpu
