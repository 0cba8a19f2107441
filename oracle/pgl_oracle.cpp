// ORACLE — test infrastructure only (see pgl_oracle.h).
//
// Reference files followed (paths relative to the Open PGL 0.6.0 tree):
//   data/SampleStatistics.h          SampleStatistics 12-193, IntegerSampleStatistics 195-342
//   spatial/kdtree/KDTreeBuilder.h   updateTree 76-118, getSplitDimensionAndPosition 218-230,
//                                    updateTreeNode 233-350
//   spatial/kdtree/KDTree.h          KDNode 21-170, getDataIdxAtPos 398-426, rearrange 256-283
//   spatial/KNN.h                    draw 22-28, RegionNeighbours<8> 116-149, neighbours 212-266
//   field/Field.h                    getRegion 109-145, buildField 147-185, updateField 187-220,
//                                    estimateSceneBounds 326-344, fitRegions 389-439,
//                                    updateRegions 441-511
//   directional/vmm/ParallaxAwareVonMisesFisherMixture.h   product 593-671, pdf 676-694,
//                                    softAssignment 697-733, sample 737-806, shift 809-831,
//                                    split/merge/swap/clear 165-289, normalisation 852-876,
//                                    KappaToMeanCosine/MeanCosineToKappa 910-928
//   directional/vmm/ParallaxAwareVonMisesFisherWeightedEMFactory.h
//                                    SufficientStatistics 93-593, InitUniformVMM 632-700,
//                                    updateMixture 751-795, partialUpdateMixture 798-844,
//                                    E-step 847-892, MAP 895-999, partial MAP 1015-1148,
//                                    reproject/prepare 1151-1196, distances 1199-1326
//   directional/vmm/AdaptiveSplitandMergeFactory.h   fit 303-347, update 352-462
//   directional/vmm/VMMChiSquareComponentSplitter.h  PerformRecursiveSplitting 293-356,
//                                    UpdateSplitStatistics 439-523, SplitComponent 527-665,
//                                    mergeComponentStats 994-1100, getSplitCandidates 1103-1117
//   directional/vmm/VMMChiSquareComponentMerger.h    PerformMerging 67-89, cost 93-177,
//                                    ThresholdedMergeNext 218-279, products 338-396
//   directional/vmm/VMM{Surface,Volume}SamplingDistribution.h, VMMPhaseFunctions.h:44-57
//
// How each part is pinned (tests/):
//   * IntegerSampleStatistics / rcp: golden vectors written by the reference's own
//     SampleStatistics.h compiled in oracle/_ref (tests/test_oracle.py), plus closed forms;
//     the rcp of each device width against the reference built for that ISA
//     (oracle/_ref/libpglref_w4.so, libpglref_w16.so; tests/test_oracle_spatial.py).
//   * kd build: unbuildable reference (needs oneTBB headers); pinned by invariants
//     (partition/lookup agreement, leaf capacity, split = mean of argmax variance of
//     the merged statistics) and, over several frames, by an independent
//     re-derivation of each Update from the previous tree (parent bounds, strict
//     threshold, depth limit, history merge) that must match bit for bit and that
//     each mutated rule breaks (tests/test_oracle_spatial.py).
//   * KNN: brute-force 8-NN recomputation; the pick (draw, fused or unfused squared
//     distances, slot ids in the low mantissa bits, rescaled random number) on
//     hand-made neighbourhoods with known distances (tests/test_oracle_spatial.py).
//   * directional fit/update, products, sample, pdf: the reference's own directional
//     headers compiled in oracle/_ref (golden KATs in tests/golden), K=1 closed-form
//     MLE, FP64 twin agreement, integral of pdf = 1.
//   * plain VMM field type (orc_config.parallax = 0): the reference's
//     ParallaxAwareVonMisesFisherMixture<8, 32, false> path compiled in oracle/_ref
//     (tests/golden/region_kat_np.npz and a live comparison).
#include "pgl_oracle.h"

#include <immintrin.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <atomic>
#include <thread>
#include <vector>

namespace orc {

static const int MAXK = 32;
static const float OPENPGL_MAX_KAPPA_F = 32000.0f;   // WEM:13
static const double ONE_OVER_FOUR_PI_D = 0.07957747154594767;

// ---------------------------------------------------------------- host rcp
// embree::rcp (emath.h:61-88) under each device's flags (CMakeLists.txt:125-131):
// CPU_8 -mavx2 -mfma: rcpss + one fused Newton step; CPU_4 -msse4.1: rcpss +
// a Newton step without FMA; CPU_16 -mavx512vl (with AVX2): rcp14 + fused step.
__attribute__((target("avx2,fma"))) static float rcp_avx2(float x) {
  const __m128 a = _mm_set_ss(x);
  const __m128 r = _mm_rcp_ss(a);
  return _mm_cvtss_f32(_mm_mul_ss(r, _mm_fnmadd_ss(r, a, _mm_set_ss(2.0f))));
}
__attribute__((target("sse4.1"))) static float rcp_sse4(float x) {
  const __m128 a = _mm_set_ss(x);
  const __m128 r = _mm_rcp_ss(a);
  return _mm_cvtss_f32(_mm_mul_ss(r, _mm_sub_ss(_mm_set_ss(2.0f), _mm_mul_ss(r, a))));
}
__attribute__((target("avx512f,avx512vl,avx2,fma"))) static float rcp_avx512(float x) {
  const __m128 a = _mm_set_ss(x);
  const __m128 r = _mm_rcp14_ss(_mm_set_ss(0.0f), a);
  return _mm_cvtss_f32(_mm_mul_ss(r, _mm_fnmadd_ss(r, a, _mm_set_ss(2.0f))));
}
static float rcp_width(float x, int W) { return W == 4 ? rcp_sse4(x) : (W == 16 ? rcp_avx512(x) : rcp_avx2(x)); }

// ------------------------------------------------------- sample statistics
struct SampleStats {
  float mean[3] = {0, 0, 0};
  float var[3] = {0, 0, 0};
  float n = 0.f;
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};

  void clear() { *this = SampleStats(); }
  void getVariance(float v[3]) const {
    for (int i = 0; i < 3; i++) v[i] = var[i] / n;
  }
  // SampleStatistics.h:68-99
  void split(int dim, float pos, float decay, bool splitLower) {
    if (n > 0.f) {
      const float variance = var[dim] / n;
      const float sd = std::sqrt(variance);
      const float newVariance = variance - variance / 4.0f;
      var[dim] = newVariance * n;
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
  // SampleStatistics.h:101-118 (evaluated without contraction)
  void merge(const SampleStats& b) {
    float mA[3], mB[3], vA[3], vB[3];
    for (int i = 0; i < 3; i++) { mA[i] = mean[i]; mB[i] = b.mean[i]; vA[i] = var[i]; vB[i] = b.var[i]; }
    const float nA = n, nB = b.n;
    for (int i = 0; i < 3; i++) {
      float t0 = mA[i] * nA;
      float t1 = mB[i] * nB;
      mean[i] = t0 + t1;
    }
    n += nB;
    for (int i = 0; i < 3; i++) mean[i] /= n;
    for (int i = 0; i < 3; i++) {
      float a = nA * mA[i];
      float a2 = a * mA[i];
      float b2 = nB * mB[i];
      float b3 = b2 * mB[i];
      float s = vA[i] + a2;
      s = s + vB[i];
      s = s + b3;
      float c = n * mean[i];
      float c2 = c * mean[i];
      var[i] = s - c2;
    }
    for (int i = 0; i < 3; i++) { lo[i] = std::min(lo[i], b.lo[i]); hi[i] = std::max(hi[i], b.hi[i]); }
  }
};

struct IntStats {
  int64_t mean[3] = {0, 0, 0};
  int64_t var[3] = {0, 0, 0};
  uint32_t n = 0;
  int64_t lo[3], hi[3];
  float center[3], half[3], invHalf[3];
  explicit IntStats(const float lower[3], const float upper[3], int W = 8) {
    for (int i = 0; i < 3; i++) {
      lo[i] = std::numeric_limits<int>::max();
      hi[i] = -std::numeric_limits<int>::max();
      const float ext = upper[i] - lower[i];
      half[i] = ext * 0.5f;
      invHalf[i] = rcp_width(half[i], W);
      center[i] = lower[i] + half[i];
    }
  }
  // SampleStatistics.h:261-277
  void add(const float p[3]) {
    n++;
    for (int i = 0; i < 3; i++) {
      float t = (p[i] - center[i]);
      t = t * invHalf[i];
      float tv = t * t;
      tv = tv * 4096.f;
      float ts = t * 4096.f;
      const int64_t is = (int64_t)ts;
      mean[i] += is;
      var[i] += (int64_t)tv;
      lo[i] = std::min(lo[i], is);
      hi[i] = std::max(hi[i], is);
    }
  }
  void merge(const IntStats& b) {
    for (int i = 0; i < 3; i++) {
      mean[i] += b.mean[i]; var[i] += b.var[i];
      lo[i] = std::min(lo[i], b.lo[i]); hi[i] = std::max(hi[i], b.hi[i]);
    }
    n += b.n;
  }
  // SampleStatistics.h:297-340
  SampleStats toSampleStats() const {
    SampleStats s;
    if (n > 0) {
      const float invN = 1.f / float(n);
      for (int i = 0; i < 3; i++) {
        float m = (float(mean[i]) / 4096.f);
        m = m * invN;
        float v = (float(var[i]) / 4096.f);
        v = v * invN;
        float mm = m * m;
        v = v - mm;
        v = std::fabs(v);
        m = m * half[i];
        m = m + center[i];
        float hh = half[i] * half[i];
        v = v * hh;
        s.mean[i] = m;
        s.var[i] = v * float(n);
        float l = float(lo[i]) - 0.0f;
        l = l / 4096.f;
        l = l * half[i];
        l = l + center[i];
        float u = float(hi[i]) + 0.0f;
        u = u / 4096.f;
        u = u * half[i];
        u = u + center[i];
        s.lo[i] = l;
        s.hi[i] = u;
      }
      s.n = float(n);
    }
    return s;
  }
};

// --------------------------------------------------------------- kd tree
struct Node {
  float pivot = 0.f;
  uint32_t bits = 0;  // dim<<30 | idx
  bool isLeaf() const { return (bits >> 30) == 3; }
  uint32_t dim() const { return bits >> 30; }
  uint32_t idx() const { return (bits << 2) >> 2; }
  void setLeaf(uint32_t data) { bits = (3u << 30) | data; }
  void setInner(uint32_t d, float p, uint32_t left) { pivot = p; bits = (d << 30) | left; }
};

// ------------------------------------------------------------- directional
template <class R> struct VMM {
  int K = 32;
  R w[MAXK], kappa[MAXK], mu[MAXK][3], norm[MAXK], em2k[MAXK], mc[MAXK], dist[MAXK];
  R pivot[3] = {0, 0, 0};
  VMM() { for (int k = 0; k < MAXK; k++) clearComp(k); }
  void clearComp(int k) {
    w[k] = 0; kappa[k] = 0; em2k[k] = 1; mc[k] = 0; norm[k] = R(ONE_OVER_FOUR_PI_D);
    mu[k][0] = 0; mu[k][1] = 0; mu[k][2] = 1; dist[k] = 0;
  }
};

template <class R> struct Suff {
  R swd[MAXK][3], sw[MAXK], sdw[MAXK];
  R sumWeights = 0, numSamples = 0, overall = 0;
  int K = MAXK;
  bool normalized = false;
  Suff() { for (int k = 0; k < MAXK; k++) { swd[k][0] = swd[k][1] = swd[k][2] = 0; sw[k] = 0; sdw[k] = 0; } }
  // clear(_numComponents) clears whole SIMD vectors (WEM:372-389); vectors are 8 wide
  void clear(int nk, int W) {
    K = nk;
    const int lim = std::min(MAXK, ((nk + W - 1) / W) * W);
    for (int k = 0; k < lim; k++) { swd[k][0] = swd[k][1] = swd[k][2] = 0; sw[k] = 0; sdw[k] = 0; }
    sumWeights = 0; numSamples = 0; normalized = false;
  }
};

template <class R> struct SplitStats {
  R chi[MAXK], mean[MAXK][2], cov[MAXK][3], n[MAXK], sw[MAXK], sa[MAXK];
  int K = 0;
  SplitStats() { for (int k = 0; k < MAXK; k++) zero(k); }
  void zero(int k) { chi[k] = 0; mean[k][0] = mean[k][1] = 0; cov[k][0] = cov[k][1] = cov[k][2] = 0; n[k] = 0; sw[k] = 0; sa[k] = 0; }
  void clear(int nk, int W) {
    K = nk;
    const int lim = std::min(MAXK, ((nk + W - 1) / W) * W);
    for (int k = 0; k < lim; k++) zero(k);
  }
};

template <class R> struct Training {
  Suff<R> suff;
  SplitStats<R> split;
  size_t afterSplit = 0, afterMerge = 0;
};

struct Cfg {
  orc_config c;
};

template <class R> static R dot3(const R a[3], const R b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <class R> static R KappaToMeanCosine(R kappa) {
  if (!(kappa > R(0))) return R(0);
  return R(1) / std::tanh(kappa) - R(1) / kappa;
}
template <class R> static R MeanCosineToKappa(R r) {
  const R r2 = r * r;
  return (r * R(3) - r * r2) / (R(1) - r2);
}

template <class R> static void calcNormalization(VMM<R>& v) {
  for (int k = 0; k < MAXK; k++) {
    v.em2k[k] = std::exp(R(-2) * v.kappa[k]);
    const R nrm = v.kappa[k] / (R(2.0 * M_PI) * (R(1) - v.em2k[k]));
    v.norm[k] = v.kappa[k] > R(0) ? nrm : R(ONE_OVER_FOUR_PI_D);
  }
}
template <class R> static void calcMeanCosines(VMM<R>& v) {
  for (int k = 0; k < MAXK; k++) {
    const R t = R(1) - R(2) / (std::exp(R(2) * v.kappa[k]) - R(1));
    const R m = R(1) / t - R(1) / v.kappa[k];
    v.mc[k] = v.kappa[k] > R(0) ? m : R(0);
  }
}

// coordinate frame around N (linearspace3.h:117-124): columns dx, dy, N
template <class R> static void frame(const R N[3], R dx[3], R dy[3]) {
  const R d0[3] = {R(0), N[2], -N[1]};
  const R d1[3] = {-N[2], R(0), N[0]};
  const R* d = dot3(d0, d0) > dot3(d1, d1) ? d0 : d1;
  R l = R(1) / std::sqrt(dot3(d, d));
  for (int i = 0; i < 3; i++) dx[i] = d[i] * l;
  R c[3] = {N[1] * dx[2] - N[2] * dx[1], N[2] * dx[0] - N[0] * dx[2], N[0] * dx[1] - N[1] * dx[0]};
  l = R(1) / std::sqrt(dot3(c, c));
  for (int i = 0; i < 3; i++) dy[i] = c[i] * l;
}

// softAssignment (PAVMM:697-733): returns false when pdf <= 1e-16
template <class R> static bool softAssign(const VMM<R>& v, const R d[3], R* a, R& pdf) {
  pdf = 0;
  for (int k = 0; k < v.K; k++) {
    const R c = dot3(d, v.mu[k]);
    const R cm1 = std::min(c - R(1), R(0));
    a[k] = v.w[k] * (v.norm[k] * std::exp(v.kappa[k] * cm1));
    pdf += a[k];
  }
  if (pdf <= R(1e-16)) return false;
  const R inv = R(1) / pdf;
  for (int k = 0; k < v.K; k++) a[k] *= inv;
  return true;
}

template <class R> static R vmmPdf(const VMM<R>& v, const R d[3]) {
  R pdf = 0;
  for (int k = 0; k < v.K; k++) {
    const R c = dot3(d, v.mu[k]);
    const R cm1 = std::min(c - R(1), R(0));
    pdf += v.w[k] * (v.norm[k] * std::exp(v.kappa[k] * cm1));
  }
  return pdf;
}

template <class R> struct PreparedSample { R d[3]; R w, pdf, dist; };

// ----------------------------------------------------------- the factory
template <class R> struct Factory {
  const orc_config c;  // by value: callers may free their config after construction
  R uniformDirs[MAXK][MAXK][3];
  R maxMeanCosine;
  int W;
  int stopPhase = 0;  // debugging aid: stop fit() after a phase (0 = run to the end)
  explicit Factory(const orc_config& cfg) : c(cfg) {
    W = cfg.vectorWidth;
    maxMeanCosine = KappaToMeanCosine<R>(R(cfg.maxKappa));
    if (sizeof(R) == 4) maxMeanCosine = KappaToMeanCosine<float>(cfg.maxKappa);
    // WEM:668-700 spherical Fibonacci directions
    const R gr = R(1.618033988749895);
    for (int l = 0; l < MAXK; l++)
      for (int n = 0; n < MAXK; n++) {
        if (n < l + 1) {
          const R phi = R(2.0 * M_PI) * (R(n) / gr);
          const R z = R(1) - ((R(2) * R(n) + R(1)) / R(l + 1));
          const R theta = std::acos(z);
          uniformDirs[l][n][0] = std::sin(theta) * std::cos(phi);
          uniformDirs[l][n][1] = std::sin(theta) * std::sin(phi);
          uniformDirs[l][n][2] = std::cos(theta);
        } else {
          uniformDirs[l][n][0] = 0; uniformDirs[l][n][1] = 0; uniformDirs[l][n][2] = 1;
        }
      }
  }

  void initUniform(VMM<R>& v, int K, R kappa) const {
    v.K = K;
    const R weight = R(1) / R(K);
    for (int k = 0; k < MAXK; k++) {
      for (int i = 0; i < 3; i++) v.mu[k][i] = uniformDirs[K - 1][k][i];
      if (k < K) { v.kappa[k] = kappa; v.w[k] = weight; }
      else { v.kappa[k] = 0; v.w[k] = 0; v.norm[k] = R(ONE_OVER_FOUR_PI_D); v.em2k[k] = 1; v.mc[k] = 0; }
    }
    calcNormalization(v);
    calcMeanCosines(v);
  }

  struct Unassigned { R sw = 0; R swd[3] = {0, 0, 0}; };

  // weightedExpectationStep (WEM:847-892)
  R estep(const VMM<R>& v, Suff<R>& st, Unassigned& un, const PreparedSample<R>* s, size_t n) const {
    un = Unassigned();
    st.clear(v.K, W);
    st.K = v.K;
    st.numSamples = R(n);
    R ll = 0;
    R a[MAXK];
    for (size_t i = 0; i < n; i++) {
      R pdf;
      if (!softAssign(v, s[i].d, a, pdf)) {
        un.sw += s[i].w;
        for (int j = 0; j < 3; j++) un.swd[j] += s[i].d[j] * s[i].w;
        continue;
      }
      ll += s[i].w * std::log(pdf);
      for (int k = 0; k < v.K; k++) {
        for (int j = 0; j < 3; j++) st.swd[k][j] += (s[i].d[j] * a[k]) * s[i].w;
        st.sw[k] += a[k] * s[i].w;
      }
    }
    return ll;
  }

  static void normalize(Suff<R>& st, R ns) {
    st.numSamples = ns;
    R sum = 0;
    for (int k = 0; k < st.K; k++) sum += st.sw[k];
    st.sumWeights = sum;
    const R nrm = ns / sum;
    for (int k = 0; k < st.K; k++) {
      for (int j = 0; j < 3; j++) st.swd[k][j] *= nrm;
      st.sw[k] *= nrm;
    }
    st.normalized = true;
  }

  static void addStats(Suff<R>& a, const Suff<R>& b) {  // operator+= (WEM:575-593)
    a.sumWeights += b.sumWeights;
    a.numSamples += b.numSamples;
    a.overall += b.numSamples;
    for (int k = 0; k < a.K; k++) {
      for (int j = 0; j < 3; j++) a.swd[k][j] += b.swd[k][j];
      a.sw[k] += b.sw[k];
      a.sdw[k] += b.sdw[k];
    }
  }

  void mapWeights(VMM<R>& v, const Suff<R>& cur, const Suff<R>& prev) const {
    const R wp = R(c.weightPrior);
    const R ns = cur.numSamples + prev.numSamples;
    for (int k = 0; k < v.K; k++) {
      const R wt = cur.sw[k] + prev.sw[k];
      v.w[k] = (wp + wt) / ((wp * R(v.K)) + ns);
    }
    for (int k = v.K; k < MAXK; k++) v.w[k] = 0;
  }

  void meanDir(const Suff<R>& cur, const Suff<R>& prev, int k, R ce, R pe, R md[3]) const {
    R cm[3], pm[3];
    for (int j = 0; j < 3; j++) {
      cm[j] = cur.sw[k] > R(0) ? cur.swd[k][j] / cur.sw[k] : R(0);
      pm[j] = prev.sw[k] > R(0) ? prev.swd[k][j] / prev.sw[k] : R(0);
      md[j] = cm[j] * ce + pm[j] * pe;
    }
  }

  R mcToKappa(R r) const {
    // FP64 inversion of the mean cosine (the GPU path inverts in FP64 too);
    // the oracle keeps the reference's FP32 formula in float mode.
    return MeanCosineToKappa<R>(r);
  }

  void mapMeanKappa(VMM<R>& v, const Suff<R>& cur, const Suff<R>& prev, const bool* mask) const {
    const R cn = cur.numSamples, pn = prev.numSamples;
    const R ns = cn + pn;
    const R overall = cur.numSamples + prev.overall;
    const R ce = cn / ns;
    const R pe = R(1) - ce;
    const R prior = R(c.meanCosinePrior), ps = R(c.meanCosinePriorStrength);
    for (int k = 0; k < v.K; k++) {
      const R partial = v.w[k] * overall;
      R md[3];
      meanDir(cur, prev, k, ce, pe, md);
      R mcos = std::sqrt(dot3(md, md));
      const bool m = mask ? mask[k] : true;
      if (m && mcos > R(0)) for (int j = 0; j < 3; j++) v.mu[k][j] = md[j] / mcos;
      mcos = (prior * ps + mcos * partial) / (ps + partial);
      mcos = std::min(maxMeanCosine, mcos);
      if (m) { v.mc[k] = mcos; v.kappa[k] = mcToKappa(mcos); }
    }
    for (int k = v.K; k < MAXK; k++) {
      v.mu[k][0] = 0; v.mu[k][1] = 0; v.mu[k][2] = 1; v.mc[k] = 0; v.kappa[k] = 0;
      if (!mask) { v.norm[k] = R(ONE_OVER_FOUR_PI_D); v.em2k[k] = 1; }
    }
    calcNormalization(v);
  }

  void partialMapWeights(VMM<R>& v, const bool* mask, const Suff<R>& cur, const Suff<R>& prev) const {
    const R wp = R(c.weightPrior);
    const R ns = cur.numSamples + prev.numSamples;
    R sumW = 0, sumP = 0;
    for (int k = 0; k < v.K; k++) {
      R wt = cur.sw[k] + prev.sw[k];
      wt = (wp + wt) / ((wp * R(v.K)) + ns);
      if (mask[k]) { sumP += wt; v.w[k] = wt; } else sumW += v.w[k];
    }
    R inv = R(1) / sumP;
    inv *= R(1) - sumW;
    for (int k = 0; k < v.K; k++) if (mask[k]) v.w[k] = v.w[k] * inv;
    for (int k = v.K; k < MAXK; k++) v.w[k] = 0;
  }

  void handleUnassigned(const Unassigned& un, VMM<R>& v, Suff<R>& cur, Suff<R>& prev) const {
    const int k = cur.K;
    cur.K++;
    cur.sw[k] = un.sw;
    for (int j = 0; j < 3; j++) cur.swd[k][j] = un.swd[j];
    prev.K++;
    prev.sw[k] = 0;
    for (int j = 0; j < 3; j++) prev.swd[k][j] = 0;
    v.K++;
  }

  // updateMixture / partialUpdateMixture (WEM:751-844)
  int em(VMM<R>& v, Suff<R>& prev, const PreparedSample<R>* s, size_t n, bool* mask) const {
    Suff<R> cur;
    cur.clear(MAXK, W);
    int it = 0;
    bool converged = false;
    R prevLL = 0, invPrevLL = 1;
    Unassigned un;
    while (!converged && it < int(c.maxEMIterations)) {
      const R ll = estep(v, cur, un, s, n);
      if (un.sw > R(0) && cur.K < MAXK) {
        handleUnassigned(un, v, cur, prev);
        if (mask) mask[v.K - 1] = true;
      }
      normalize(cur, cur.numSamples);
      if (mask) {
        partialMapWeights(v, mask, cur, prev);
        mapMeanKappa(v, cur, prev, mask);
      } else {
        mapWeights(v, cur, prev);
        mapMeanKappa(v, cur, prev, nullptr);
      }
      it++;
      if (it > 1) {
        const R rel = std::fabs(ll - prevLL) * invPrevLL;
        if (getenv("ORC_DEBUG_EM")) fprintf(stderr, "em it %d K %d ll %.9g rel %.9g\n", it, v.K, double(ll), double(rel));
        if (rel < R(c.convergenceThreshold)) converged = true;
        prevLL = ll;
        invPrevLL = R(1) / std::fabs(ll);
      }
    }
    addStats(prev, cur);
    return it;
  }

  // init/updateComponentDistances (WEM:1199-1326), harmonic mean
  void distances(VMM<R>& v, Suff<R>& st, const PreparedSample<R>* s, size_t n, bool init) const {
    R bd[MAXK], bw[MAXK], a[MAXK];
    for (int k = 0; k < MAXK; k++) bd[k] = bw[k] = 0;
    for (size_t i = 0; i < n; i++) {
      const R sd = R(1) / s[i].dist;
      R pdf;
      if (softAssign(v, s[i].d, a, pdf)) {
        for (int k = 0; k < v.K; k++) {
          const R wt = s[i].w * a[k] * ((a[k] * pdf) / v.w[k]);
          bd[k] += wt * sd;
          bw[k] += wt;
        }
      }
    }
    const int lim = std::min(MAXK, ((v.K + W - 1) / W) * W);
    for (int k = 0; k < lim; k++) {
      if (init) {
        st.sdw[k] = bw[k];
        v.dist[k] = st.sdw[k] / bd[k];
      } else {
        R sumInv = bd[k];
        sumInv += v.dist[k] > R(0) ? (st.sdw[k] / v.dist[k]) : R(0);
        st.sdw[k] += bw[k];
        v.dist[k] = st.sdw[k] / sumInv;
      }
    }
    for (int k = v.K; k < lim; k++) { v.dist[k] = 0; st.sdw[k] = 0; }
  }

  // UpdateSplitStatistics (Splitter:439-523), zero-mean covariance, running chi^2 mean
  void updateSplitStats(const VMM<R>& v, SplitStats<R>& ss, R mc, const PreparedSample<R>* s, size_t n) const {
    R a[MAXK];
    R fx[MAXK][3], fy[MAXK][3];
    for (int k = 0; k < ss.K; k++) frame(v.mu[k], fx[k], fy[k]);
    for (size_t i = 0; i < n; i++) {
      R pdf;
      if (!softAssign(v, s[i].d, a, pdf)) continue;
      const R value = s[i].w * s[i].pdf;
      for (int k = 0; k < ss.K; k++) {
        const R vmf = a[k] * pdf;
        R part = vmf * value;
        part /= (mc * pdf);
        R chi = value * value * vmf;
        chi /= mc * mc * pdf * pdf;
        chi -= R(2) * part;
        chi += vmf;
        chi /= s[i].pdf;
        if (!(a[k] > R(0))) chi = 0;
        ss.sa[k] += a[k];
        ss.n[k] += R(1);
        ss.chi[k] += (chi - ss.chi[k]) / ss.n[k];
        // local frame: frame(mu).inverse() * d = (dx.d, dy.d, N.d) for an orthonormal frame
        const R lx = dot3(fx[k], s[i].d), ly = dot3(fy[k], s[i].d);
        const R aw = a[k] * s[i].w;
        ss.sw[k] += aw;
        ss.cov[k][0] += aw * (lx * lx);
        ss.cov[k][1] += aw * (ly * ly);
        ss.cov[k][2] += aw * (lx * ly);
      }
      // unused lanes of the last SIMD vector still count the sample (Splitter:484-487)
      const int lim = std::min(MAXK, ((ss.K + W - 1) / W) * W);
      for (int k = ss.K; k < lim; k++) ss.n[k] += R(1);
    }
  }

  static void expMap(const R v2[2], R out[3]) {  // Map2DTo3D, log-map variant (Splitter:175-190)
    const R len = std::sqrt(v2[0] * v2[0] + v2[1] * v2[1]);
    const R sinc = std::sin(len) / len;
    out[0] = len > R(0) ? v2[0] * sinc : R(0);
    out[1] = len > R(0) ? v2[1] * sinc : R(0);
    out[2] = std::cos(len);
  }
  static void logMap(const R v3[3], R out[2]) {  // Map3DTo2D (Splitter:161-172)
    const R z = std::max(R(-1), std::min(R(1), v3[2]));
    const R alpha = std::acos(z);
    const R invSinc = alpha / std::sin(alpha);
    out[0] = alpha > R(0) ? v3[0] * invSinc : R(0);
    out[1] = alpha > R(0) ? v3[1] * invSinc : R(0);
  }

  static void vmmSplitComponent(VMM<R>& v, int i0, int i1, R w0, R w1, const R m0[3], const R m1[3], R c0, R c1) {
    v.dist[i1] = v.dist[i0];
    v.w[i0] = w0; v.mc[i0] = c0; v.kappa[i0] = MeanCosineToKappa<R>(c0);
    for (int j = 0; j < 3; j++) v.mu[i0][j] = m0[j];
    v.w[i1] = w1; v.mc[i1] = c1; v.kappa[i1] = MeanCosineToKappa<R>(c1);
    for (int j = 0; j < 3; j++) v.mu[i1][j] = m1[j];
    if (i1 == v.K) v.K++;
    calcNormalization(v);
  }

  static void suffSplit(Suff<R>& st, int i, int j, const R m0[3], const R m1[3], R c0, R c1) {
    R s = st.sw[i];
    s /= R(2);
    st.sw[i] = s;
    for (int d = 0; d < 3; d++) st.swd[i][d] = m0[d] * c0 * s;
    st.sw[j] = s;
    for (int d = 0; d < 3; d++) st.swd[j][d] = m1[d] * c1 * s;
    const R t = st.sdw[i] * R(0.5);
    st.sdw[i] = t;
    st.sdw[j] = t;
    st.K += 1;
  }

  // SplitComponent (Splitter:527-665)
  bool splitComponent(VMM<R>& v, SplitStats<R>& ss, Suff<R>& st, int idx) const {
    const R numAssigned = ss.sa[idx];
    const R invSw = R(1) / ss.sw[idx];
    const R cx = ss.cov[idx][0] * invSw, cy = ss.cov[idx][1] * invSw, cz = ss.cov[idx][2] * invSw;
    const R D = std::sqrt((cx - cy) * (cx - cy) + (cz * cz * R(4))) * R(0.5);
    R ev0 = (cx + cy) * R(0.5);
    ev0 += D;
    R e0[2] = {-cz, cx - ev0};
    const R l0 = std::sqrt(e0[0] * e0[0] + e0[1] * e0[1]);
    e0[0] /= l0; e0[1] /= l0;
    const R weight = v.w[idx], meanCosine = v.mc[idx], kappa = v.kappa[idx];
    if (kappa >= R(OPENPGL_MAX_KAPPA_F * 0.9)) return false;
    const R md[3] = {v.mu[idx][0], v.mu[idx][1], v.mu[idx][2]};
    const R nw0 = weight * R(0.5), nw1 = nw0;
    R md0[3] = {md[0], md[1], md[2]}, md1[3] = {md[0], md[1], md[2]};
    R nc0 = meanCosine, nc1 = meanCosine * meanCosine;
    if (D > R(1e-8)) {
      R fx[3], fy[3];
      frame(md, fx, fy);
      const R m2a[2] = {e0[0] * ev0 * R(0.5), e0[1] * ev0 * R(0.5)};
      R l3[3];
      expMap(m2a, l3);
      for (int j = 0; j < 3; j++) md0[j] = fx[j] * l3[0] + fy[j] * l3[1] + md[j] * l3[2];
      nc0 = meanCosine / std::fabs(dot3(md, md0));
      nc0 = std::min(nc0, KappaToMeanCosine<R>(R(OPENPGL_MAX_KAPPA_F)));
      nc1 = nc0;
      const R m2b[2] = {-(e0[0] * ev0 * R(0.5)), -(e0[1] * ev0 * R(0.5))};
      expMap(m2b, l3);
      for (int j = 0; j < 3; j++) md1[j] = fx[j] * l3[0] + fy[j] * l3[1] + md[j] * l3[2];
    } else {
      if (numAssigned < R(2)) return false;
    }
    const int K = v.K;
    vmmSplitComponent(v, idx, K, nw0, nw1, md0, md1, nc0, nc1);
    suffSplit(st, idx, K, md0, md1, nc0, nc1);
    ss.zero(idx);
    ss.zero(K);
    ss.K = K + 1;
    return true;
  }

  static void maskedReplace(Suff<R>& dst, const bool* mask, const Suff<R>& src) {
    R sum = 0;
    for (int k = 0; k < MAXK; k++) {
      if (mask[k]) {
        for (int j = 0; j < 3; j++) dst.swd[k][j] = src.swd[k][j];
        dst.sw[k] = src.sw[k];
        dst.sdw[k] = src.sdw[k];
      }
      sum += dst.sw[k];
    }
    if (dst.normalized) dst.numSamples = sum; else dst.sumWeights = sum;
  }

  struct Cand { int idx; R chi; };
  std::vector<Cand> candidates(const SplitStats<R>& ss) const {
    std::vector<Cand> c;
    for (int k = 0; k < ss.K; k++) c.push_back({k, ss.chi[k]});
    std::stable_sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) { return a.chi > b.chi; });
    return c;
  }

  // PerformRecursiveSplitting (Splitter:293-356)
  int recursiveSplitting(VMM<R>& v, Suff<R>& st, R mc, const PreparedSample<R>* s, size_t n, int& emIts) const {
    int total = 0;
    int numSplits = -1;
    Suff<R> temp = st;
    SplitStats<R> ss;
    while (v.K < MAXK && numSplits != 0) {
      numSplits = 0;
      ss.clear(MAXK, W);
      ss.clear(v.K, W);
      updateSplitStats(v, ss, mc, s, n);
      auto cands = candidates(ss);
      bool mask[MAXK] = {false};
      const int numComp = v.K;
      for (int k = 0; k < numComp; k++) {
        if (cands[k].chi > R(c.splittingThreshold) && v.K < MAXK) {
          const bool ok = splitComponent(v, ss, temp, cands[k].idx);
          mask[cands[k].idx] = true;
          mask[v.K - 1] = true;
          if (ok) numSplits++;
        }
      }
      if (numSplits > 0) {
        temp.clear(v.K, W);
        emIts += em(v, temp, s, n, mask);
        st.K = v.K;
        maskedReplace(st, mask, temp);
      }
      total += numSplits;
    }
    return total;
  }

  // ---- merging (Merger)
  static R productScale(const R m0[3], R k0, R n0, const R m1[3], R k1, R n1, R pm[3], R& pk, R& pn) {
    for (int j = 0; j < 3; j++) pm[j] = k0 * m0[j] + k1 * m1[j];
    pk = std::sqrt(dot3(pm, pm));
    pn = R(1.0 / (4.0 * M_PI));
    if (pk > R(1e-3)) {
      const R e = std::exp(R(-2) * pk);
      pn = pk / (R(2.0 * M_PI) * (R(1) - e));
      for (int j = 0; j < 3; j++) pm[j] /= pk;
    } else {
      pk = 0;
      for (int j = 0; j < 3; j++) pm[j] = m0[j];
    }
    R scale = (n0 * n1) / pn;
    const R c0 = dot3(m0, pm), c1 = dot3(m1, pm);
    scale *= std::exp(k0 * (c0 - R(1)) + k1 * (c1 - R(1)));
    return scale;
  }
  static R integratedDivision(const R m0[3], R k0, R n0, const R m1[3], R k1, R n1, R em2k1) {
    R pm[3];
    for (int j = 0; j < 3; j++) pm[j] = k0 * m0[j] + k1 * m1[j];
    R pk = std::sqrt(dot3(pm, pm));
    R pn = R(1.0 / (4.0 * M_PI));
    if (pk > R(1e-3)) {
      const R e = std::exp(R(-2) * pk);
      pn = pk / (R(2.0 * M_PI) * (R(1) - e));
      for (int j = 0; j < 3; j++) pm[j] /= pk;
    } else {
      pk = 0;
      for (int j = 0; j < 3; j++) pm[j] = m0[j];
    }
    R scale = (n0 * n1) / pn;
    const R c0 = dot3(m0, pm), c1 = dot3(m1, pm);
    scale *= (R(4.0 * M_PI * M_PI) * (R(1) - em2k1)) / (k1 * k1);
    scale *= std::exp((k0 * (c0 - R(1)) + k1 * (c1 - R(1))) + (R(2) * k1));
    return scale;
  }
  R mergeCost(const VMM<R>& v, int i0, int i1) const {
    R w0 = v.w[i0], w1 = v.w[i1];
    const R k0 = v.kappa[i0], k1 = v.kappa[i1];
    const R *m0 = v.mu[i0], *m1 = v.mu[i1];
    const R c0 = v.mc[i0], c1 = v.mc[i1], n0 = v.norm[i0], n1 = v.norm[i1];
    if (i0 == i1) { w0 /= R(2); w1 /= R(2); }
    const R w = w0 + w1;
    R kappa = 0, nrm = R(ONE_OVER_FOUR_PI_D), em2k = 1;
    R md[3];
    for (int j = 0; j < 3; j++) md[j] = (w0 * c0 * m0[j] + w1 * c1 * m1[j]) / w;
    R mcos = dot3(md, md);
    if (mcos > R(0)) {
      mcos = std::sqrt(mcos);
      kappa = MeanCosineToKappa<R>(mcos);
      kappa = kappa < R(1e-3) ? R(0) : kappa;
      em2k = std::exp(R(-2) * kappa);
      nrm = kappa / (R(2.0 * M_PI) * (R(1) - em2k));
      for (int j = 0; j < 3; j++) md[j] /= mcos;
    } else {
      for (int j = 0; j < 3; j++) md[j] = m0[j];
    }
    const R negMd[3] = {-md[0], -md[1], -md[2]};
    R m00[3], k00, n00, m11[3], k11, n11, m01[3], k01, n01;
    const R s00 = productScale(m0, k0, n0, m0, k0, n0, m00, k00, n00);
    const R s11 = productScale(m1, k1, n1, m1, k1, n1, m11, k11, n11);
    const R s01 = productScale(m0, k0, n0, m1, k1, n1, m01, k01, n01);
    R chi = 0;
    chi += ((w0 * w0) / w) * (s00 * integratedDivision(m00, k00, n00, negMd, kappa, nrm, em2k));
    chi += ((w1 * w1) / w) * (s11 * integratedDivision(m11, k11, n11, negMd, kappa, nrm, em2k));
    chi += R(2) * ((w0 * w1) / w) * (s01 * integratedDivision(m01, k01, n01, negMd, kappa, nrm, em2k));
    chi -= w;
    return chi;
  }

  static void vmmSwap(VMM<R>& v, int a, int b) {
    if (a == b) return;
    std::swap(v.w[a], v.w[b]); std::swap(v.kappa[a], v.kappa[b]); std::swap(v.em2k[a], v.em2k[b]);
    std::swap(v.mc[a], v.mc[b]); std::swap(v.norm[a], v.norm[b]);
    for (int j = 0; j < 3; j++) std::swap(v.mu[a][j], v.mu[b][j]);
    std::swap(v.dist[a], v.dist[b]);
  }
  // PAVMM::mergeComponents (PAVMM:259-343)
  static void vmmMerge(VMM<R>& v, int i0, int i1) {
    if (i0 == i1) return;
    const R w0 = v.w[i0], w1 = v.w[i1], c0 = v.mc[i0], c1 = v.mc[i1];
    R kappa = 0, nrm = R(ONE_OVER_FOUR_PI_D), e = 1;
    const R w = w0 + w1;
    R md[3];
    for (int j = 0; j < 3; j++) md[j] = w0 * c0 * v.mu[i0][j] + w1 * c1 * v.mu[i1][j];
    for (int j = 0; j < 3; j++) md[j] /= w;
    R mcos = md[0] * md[0] + md[1] * md[1] + md[2] * md[2];
    if (mcos > R(0)) {
      mcos = std::sqrt(mcos);
      kappa = MeanCosineToKappa<R>(mcos);
      kappa = kappa < R(1e-3) ? R(0) : kappa;
      e = std::exp(R(-2) * kappa);
      nrm = kappa / (R(2.0 * M_PI) * (R(1) - e));
      for (int j = 0; j < 3; j++) md[j] /= mcos;
    } else {
      for (int j = 0; j < 3; j++) md[j] = v.mu[i0][j];
    }
    v.w[i0] = w; v.kappa[i0] = kappa; v.mc[i0] = mcos; v.norm[i0] = nrm; v.em2k[i0] = e;
    for (int j = 0; j < 3; j++) v.mu[i0][j] = md[j];
    R nd = w0 * v.dist[i0] + w1 * v.dist[i1];
    nd /= (w0 + w1);
    v.dist[i0] = nd;
    vmmSwap(v, i1, v.K - 1);
    v.clearComp(v.K - 1);
    v.K -= 1;
  }

  // ComponentSplitStatistics::mergeComponentStats (Splitter:994-1100), zero-mean variant
  static void splitStatsMerge(SplitStats<R>& ss, int i, int j, R wI, const R mI[3], R wJ, const R mJ[3], R wK, const R mK[3]) {
    const int L = ss.K - 1;
    R fx[3], fy[3];
    frame(mK, fx, fy);
    const R li[3] = {dot3(fx, mI), dot3(fy, mI), dot3(mK, mI)};
    const R lj[3] = {dot3(fx, mJ), dot3(fy, mJ), dot3(mK, mJ)};
    R ai[2], aj[2];
    logMap(li, ai);
    logMap(lj, aj);
    const R invWK = (wK > R(0)) ? R(1) / wK : R(1);
    const R swI = ss.sw[i], swJ = ss.sw[j], swK = swI + swJ;
    R cI[3] = {0, 0, 0}, cJ[3] = {0, 0, 0};
    if (swI > R(0)) { const R r = R(1) / swI; for (int d = 0; d < 3; d++) cI[d] = ss.cov[i][d] * r; }
    if (swJ > R(0)) { const R r = R(1) / swJ; for (int d = 0; d < 3; d++) cJ[d] = ss.cov[j][d] * r; }
    const R mII[3] = {ai[0] * ai[0], ai[1] * ai[1], ai[0] * ai[1]};
    const R mJJ[3] = {aj[0] * aj[0], aj[1] * aj[1], aj[0] * aj[1]};
    R cK[3];
    for (int d = 0; d < 3; d++) cK[d] = (wI * cI[d] + wI * mII[d] + wJ * cJ[d] + wJ * mJJ[d]) * invWK;
    const R saK = ss.sa[i] + ss.sa[j];
    const R nK = invWK * (wI * ss.n[i] + wJ * ss.n[j]);
    const R chiK = ss.chi[i] + ss.chi[j];
    ss.mean[i][0] = 0; ss.mean[i][1] = 0;
    for (int d = 0; d < 3; d++) ss.cov[i][d] = cK[d] * swK;
    ss.sw[i] = swK; ss.n[i] = nK; ss.sa[i] = saK; ss.chi[i] = chiK;
    ss.mean[j][0] = ss.mean[L][0]; ss.mean[j][1] = ss.mean[L][1];
    for (int d = 0; d < 3; d++) ss.cov[j][d] = ss.cov[L][d];
    ss.sw[j] = ss.sw[L]; ss.n[j] = ss.n[L]; ss.sa[j] = ss.sa[L]; ss.chi[j] = ss.chi[L];
    ss.zero(L);
    ss.K--;
  }
  static void suffMerge(Suff<R>& st, int i, int j) {
    const int L = st.K - 1;
    for (int d = 0; d < 3; d++) st.swd[i][d] += st.swd[j][d];
    st.sw[i] += st.sw[j];
    st.sdw[i] += st.sdw[j];
    for (int d = 0; d < 3; d++) st.swd[j][d] = st.swd[L][d];
    st.sw[j] = st.sw[L];
    st.sdw[j] = st.sdw[L];
    for (int d = 0; d < 3; d++) st.swd[L][d] = 0;
    st.sw[L] = 0;
    st.sdw[L] = 0;
    st.K--;
  }

  int performMerging(VMM<R>& v, Suff<R>& st, SplitStats<R>& ss) const {
    int total = 0;
    while (v.K > W) {
      int ci = 0, cj = 0;
      R minCost = std::numeric_limits<R>::max();
      bool found = false;
      for (int i = 0; i < v.K - 1; i++)
        for (int j = i + 1; j < v.K; j++) {
          const R cost = mergeCost(v, i, j);
          if (cost < R(c.mergingThreshold) && cost < minCost && ss.n[i] > R(0) && ss.n[j] > R(0)) {
            ci = i; cj = j; minCost = cost; found = true;
          }
        }
      if (!found) break;
      if (getenv("ORC_DEBUG_MERGE")) fprintf(stderr, "merge %d %d cost %g K %d\n", ci, cj, double(minCost), v.K);
      R mI[3] = {v.mu[ci][0], v.mu[ci][1], v.mu[ci][2]};
      const R wI = v.w[ci];
      R mJ[3] = {v.mu[cj][0], v.mu[cj][1], v.mu[cj][2]};
      const R wJ = v.w[cj];
      vmmMerge(v, ci, cj);
      R mK[3] = {v.mu[ci][0], v.mu[ci][1], v.mu[ci][2]};
      const R wK = v.w[ci];
      splitStatsMerge(ss, ci, cj, wI, mI, wJ, mJ, wK, mK);
      suffMerge(st, ci, cj);
      total++;
    }
    return total;
  }

  // ASM::fit (ASM:303-347)
  void fit(VMM<R>& v, Training<R>& t, const PreparedSample<R>* s, size_t n, orc_region_record* info) const {
    const int K0 = c.initK;
    t.suff.clear(K0, W);
    t.split.clear(K0, W);
    t.afterSplit = 0; t.afterMerge = 0;
    initUniform(v, K0, R(c.initKappa));
    t.suff.clear(K0, W);
    t.suff.normalized = true;
    int its = em(v, t.suff, s, n, nullptr);
    distances(v, t.suff, s, n, true);
    if (stopPhase == 1) return;
    int splits = 0, merges = 0, pits = 0;
    if (c.useSplitAndMerge) {
      const R mc = t.suff.sumWeights / t.suff.numSamples;
      splits = recursiveSplitting(v, t.suff, mc, s, n, pits);
      if (stopPhase == 2) return;
      t.split.clear(v.K, W);
      updateSplitStats(v, t.split, mc, s, n);
      if (stopPhase == 3) return;
      merges = performMerging(v, t.suff, t.split);
      if (stopPhase == 4) return;
    }
    t.afterSplit = 0; t.afterMerge = 0;
    distances(v, t.suff, s, n, true);
    if (info) { info->emIterations = its; info->partialEmIterations = pits; info->numSplits = splits; info->numMerges = merges; }
  }

  // ASM::update (ASM:352-462)
  void update(VMM<R>& v, Training<R>& t, const PreparedSample<R>* s, size_t n, orc_region_record* info) const {
    const int prevK = v.K;
    int its = em(v, t.suff, s, n, nullptr);
    if (prevK < v.K) t.split.K = v.K;
    int splits = 0, merges = 0, pits = 0;
    if (c.useSplitAndMerge) {
      const R mc = t.suff.sumWeights / t.suff.numSamples;
      t.afterSplit += n;
      t.afterMerge += n;
      updateSplitStats(v, t.split, mc, s, n);
      if (t.afterSplit >= size_t(std::max(0, c.minSamplesForSplitting))) {
        bool mask[MAXK] = {false};
        auto cands = candidates(t.split);
        for (size_t k = 0; k < cands.size(); k++) {
          if (cands[k].chi > R(c.splittingThreshold) && v.K < MAXK) {
            splitComponent(v, t.split, t.suff, cands[k].idx);
            mask[cands[k].idx] = true;
            mask[v.K - 1] = true;
            splits++;
          }
        }
        if (splits > 0 && c.partialReFit && n >= size_t(std::max(0, c.minSamplesForPartialRefitting))) {
          Suff<R> temp = t.suff;
          temp.clear(v.K, W);
          pits = em(v, temp, s, n, mask);
          t.suff.K = v.K;
          maskedReplace(t.suff, mask, temp);
          t.split.K = v.K;
        }
        t.afterSplit = 0;
      }
      if (t.afterMerge >= size_t(std::max(0, c.minSamplesForMerging))) {
        merges = performMerging(v, t.suff, t.split);
        t.afterMerge = 0;
      }
    }
    distances(v, t.suff, s, n, false);
    if (info) { info->emIterations = its; info->partialEmIterations = pits; info->numSplits = splits; info->numMerges = merges; }
  }
};

// SufficientStatistics::applyParallaxShift (WEM:265-283)
template <class R> static void suffShift(Suff<R>& st, const VMM<R>& v, const R shift[3]) {
  for (int k = 0; k < v.K; k++) {
    if (!(v.dist[k] > R(0))) continue;
    R d[3] = {st.swd[k][0], st.swd[k][1], st.swd[k][2]};
    const R mcos = std::sqrt(dot3(d, d));
    if (!(mcos > R(0))) continue;
    for (int j = 0; j < 3; j++) d[j] = (d[j] / mcos) * v.dist[k] + shift[j];
    const R l = std::sqrt(dot3(d, d));
    for (int j = 0; j < 3; j++) st.swd[k][j] = (d[j] / l) * mcos;
  }
}
// performRelativeParallaxShift (PAVMM:809-831)
template <class R> static void vmmShift(VMM<R>& v, const R shift[3]) {
  for (int k = 0; k < v.K; k++) {
    R m[3];
    for (int j = 0; j < 3; j++) m[j] = v.mu[k][j] * v.dist[k] + shift[j];
    const R l = std::sqrt(dot3(m, m));
    if (v.dist[k] > R(0) && std::isfinite(v.dist[k]))
      for (int j = 0; j < 3; j++) v.mu[k][j] = m[j] / l;
    v.dist[k] = v.dist[k] > R(0) ? l : R(0);
  }
  for (int j = 0; j < 3; j++) v.pivot[j] -= shift[j];
}

template <class R> static bool suffValid(const Suff<R>& s) {
  for (int k = 0; k < MAXK; k++) {
    if (!std::isfinite(s.swd[k][0]) || !std::isfinite(s.swd[k][1]) || !std::isfinite(s.swd[k][2])) return false;
    if (!std::isfinite(s.sw[k]) || !std::isfinite(s.sdw[k])) return false;
    if (k < s.K) { if (s.sw[k] < R(0) || s.sdw[k] < R(0)) return false; }
    else if (s.swd[k][0] != R(0) || s.swd[k][1] != R(0) || s.swd[k][2] != R(0) || s.sw[k] != R(0) || s.sdw[k] != R(0)) return false;
  }
  return std::isfinite(s.numSamples) && s.numSamples >= R(0) && std::isfinite(s.sumWeights) && s.sumWeights >= R(0);
}
template <class R> static bool vmmValid(const VMM<R>& v) {
  for (int k = 0; k < MAXK; k++) {
    const R vals[9] = {v.w[k], v.kappa[k], v.mu[k][0], v.mu[k][1], v.mu[k][2], v.norm[k], v.em2k[k], v.mc[k], v.dist[k]};
    for (R x : vals) if (!std::isfinite(x)) return false;
    if (k < v.K) {
      if (v.w[k] < R(0) || v.w[k] > R(1) + R(1e-6)) return false;
      if (v.kappa[k] < R(0) || v.mc[k] < R(0) || v.mc[k] > R(1) || v.norm[k] < R(0) || v.dist[k] < R(0)) return false;
      for (int j = 0; j < 3; j++) if (v.mu[k][j] < R(-1) || v.mu[k][j] > R(1)) return false;
    } else {
      if (v.w[k] != R(0) || v.kappa[k] != R(0) || v.mu[k][0] != R(0) || v.mu[k][1] != R(0) || v.mu[k][2] != R(1) ||
          v.mc[k] != R(0) || std::fabs(v.norm[k] - R(ONE_OVER_FOUR_PI_D)) >= R(1e-6) || v.em2k[k] != R(1) || v.dist[k] != R(0))
        return false;
    }
  }
  return true;
}

// --------------------------------------------------------------- regions
// ------------------------------------------------------------------- KNN pick
static uint32_t knnDraw(float* u, uint32_t size) {  // KNN.h:22-28
  size = std::min<uint32_t>(4, size);
  const uint32_t sel = uint32_t(*u * float(size));
  *u = (*u - float(sel) / float(size)) * float(size);
  return std::min(sel, size - 1);
}

// RegionNeighbours<8>::sampleApproximateClosestRegionIdx (KNN.h:134-148);
// RegionNeighbours<4> (KNN.h:86-113) merges two sorted halves to the same
// pick.  d^2 = embree::dot(d, d) = madd(x, x, madd(y, y, z * z)) (vec3.h:204),
// madd fused under AVX2 / AVX-512 and a * b + c under SSE4 (vfloat4_sse2.h:439-444).
// ids: 8 slots (~0u = empty), pts: x[8], y[8], z[8].
static uint32_t knnPickSlots(const uint32_t* ids, const float* pts, uint32_t size, const float p[3], float* u,
                             bool fused) {
  const uint32_t sel = knnDraw(u, size);
  float dist[8];
  for (int i = 0; i < 8; i++) {
    if (ids[i] == ~0u) { dist[i] = std::numeric_limits<float>::infinity(); continue; }
    const float dx = pts[0 + i] - p[0];
    const float dy = pts[8 + i] - p[1];
    const float dz = pts[16 + i] - p[2];
    float d;
    if (fused) {
      d = std::fma(dx, dx, std::fma(dy, dy, dz * dz));
    } else {
      const float zz = dz * dz;
      const float yy = dy * dy;
      const float inner = yy + zz;
      const float xx = dx * dx;
      d = xx + inner;
    }
    uint32_t bits;
    memcpy(&bits, &d, 4);
    bits = (bits & ~7u) | uint32_t(i);
    memcpy(&d, &bits, 4);
    dist[i] = d;
  }
  std::sort(dist, dist + 8);
  uint32_t bits;
  memcpy(&bits, &dist[sel], 4);
  return ids[bits & 7u];
}

template <class R> struct Region {
  VMM<R> vmm;
  Training<R> ts;
  SampleStats ss;
  float lower[3], upper[3];
  bool splitFlag = false;
  bool valid = true;
  size_t begin = 0, end = 0;
  orc_region_record info;
  Region() { memset(&info, 0, sizeof(info)); }
};

static bool sampleLess(const orc_sample& a, const orc_sample& b) {  // SampleData.h:82-93
  if (a.weight != b.weight) return a.weight < b.weight;
  if (a.pdf != b.pdf) return a.pdf < b.pdf;
  if (a.distance != b.distance) return a.distance < b.distance;
  if (a.px != b.px) return a.px < b.px;
  if (a.py != b.py) return a.py < b.py;
  if (a.pz != b.pz) return a.pz < b.pz;
  if (a.dx != b.dx) return a.dx < b.dx;
  if (a.dy != b.dy) return a.dy < b.dy;
  return a.dz < b.dz;
}

// reprojectSample / prepareSamples (WEM:1151-1196)
template <class R>
static void prepare(const orc_sample* in, size_t n, const SampleStats& ss, std::vector<PreparedSample<R>>& out,
                    bool parallax) {
  out.resize(n);
  float var[3];
  ss.getVariance(var);
  float minDist = std::sqrt(var[0] * var[0] + var[1] * var[1] + var[2] * var[2]);
  minDist = 3.f * 3.f * std::sqrt(minDist);
  const R piv[3] = {R(ss.mean[0]), R(ss.mean[1]), R(ss.mean[2])};
  for (size_t i = 0; i < n; i++) {
    const orc_sample& s = in[i];
    PreparedSample<R>& o = out[i];
    o.w = s.weight; o.pdf = s.pdf;
    o.d[0] = s.dx; o.d[1] = s.dy; o.d[2] = s.dz;
    o.dist = s.distance;
    if (!parallax) continue;  // the plain VMM trains on the samples as recorded
    if (std::isinf(s.distance)) continue;
    if (!(s.distance > 0.f)) continue;
    const R dd = std::max(R(minDist), R(s.distance));
    R org[3] = {R(s.px) + R(s.dx) * dd, R(s.py) + R(s.dy) * dd, R(s.pz) + R(s.dz) * dd};
    R nd[3] = {org[0] - piv[0], org[1] - piv[1], org[2] - piv[2]};
    const R len = std::sqrt(dot3(nd, nd));
    for (int j = 0; j < 3; j++) o.d[j] = nd[j] / len;
    o.dist = len;
  }
}

// --------------------------------------------------------------- field
struct HGRep { int K; float g; float w[4], mc[4], kappa[4], norm[4]; };
static HGRep g_hg[128];
static bool g_hgSet = false;

template <class R> struct Field {
  orc_config cfg;
  Factory<R> factory;
  std::vector<Node> nodes;
  std::vector<Region<R>> regions;
  float sceneLo[3], sceneHi[3];
  bool boundsSet = false;
  bool initialized = false;
  uint32_t iteration = 0;
  float treeLo[3], treeHi[3];
  std::vector<orc_sample> samples;
  // KNN: per region 8 neighbours (storage ids) + pivots
  std::vector<uint32_t> nbIds;
  std::vector<float> nbPts;
  std::vector<uint32_t> nbSize;

  explicit Field(const orc_config& c) : cfg(c), factory(c) {}

  void estimateSceneBounds() {  // Field.h:326-344 incl. the N-1 quirk and FLT_MIN upper
    for (int i = 0; i < 3; i++) { sceneLo[i] = FLT_MAX; sceneHi[i] = std::numeric_limits<float>::min(); }
    boundsSet = false;
    if (!samples.empty()) {
      for (size_t i = 0; i + 1 < samples.size(); i++) {
        const float p[3] = {samples[i].px, samples[i].py, samples[i].pz};
        for (int j = 0; j < 3; j++) { sceneLo[j] = std::min(sceneLo[j], p[j]); sceneHi[j] = std::max(sceneHi[j], p[j]); }
      }
      for (int j = 0; j < 3; j++) {
        const float c = (sceneLo[j] + sceneHi[j]) * 0.5f;
        float a = sceneLo[j] - c;
        a = 3.0f * a;
        sceneLo[j] = c + a;
        float b = sceneHi[j] - c;
        b = 3.0f * b;
        sceneHi[j] = c + b;
      }
      boundsSet = true;
    }
  }

  static uint8_t maxDim(const float v[3]) {  // KDTreeBuilder.h:223-226
    const int a = v[1] > v[0];
    return v[a] > v[2] ? uint8_t(a) : uint8_t(2);
  }

  SampleStats intStats(size_t b, size_t e, const float lo[3], const float hi[3]) const {
    IntStats is(lo, hi, int(cfg.vectorWidth));
    for (size_t i = b; i < e; i++) { const float p[3] = {samples[i].px, samples[i].py, samples[i].pz}; is.add(p); }
    return is.toSampleStats();
  }

  // updateTreeNode (KDTreeBuilder.h:233-350), serial: left subtree before right
  void updateNode(uint32_t nodeIdx, size_t depth, const float blo[3], const float bhi[3], size_t b, size_t e,
                  const SampleStats& stats) {
    if (e <= b) return;
    uint32_t dim;
    float pos;
    uint32_t left;
    if (nodes[nodeIdx].isLeaf()) {
      const uint32_t dataIdx = nodes[nodeIdx].idx();
      Region<R>& reg = regions[dataIdx];
      if (depth < cfg.maxDepth && reg.ss.n + float(e - b) > float(cfg.maxSamples)) {
        SampleStats merged = reg.ss;
        merged.merge(stats);
        float var[3];
        merged.getVariance(var);
        dim = maxDim(var);
        pos = merged.mean[dim];
        Region<R> right = reg;
        regions[dataIdx].ss.split(dim, pos, 0.25f, false);
        right.ss.split(dim, pos, 0.25f, true);
        regions[dataIdx].splitFlag = true;
        right.splitFlag = true;
        regions[dataIdx].upper[dim] = pos;
        right.lower[dim] = pos;
        regions.push_back(right);
        const uint32_t rightIdx = uint32_t(regions.size() - 1);
        left = uint32_t(nodes.size());
        nodes.push_back(Node());
        nodes.push_back(Node());
        nodes[nodeIdx].setInner(dim, pos, left);
        nodes[left].setLeaf(dataIdx);
        nodes[left + 1].setLeaf(rightIdx);
      } else {
        reg.ss.merge(stats);
        reg.begin = b;
        reg.end = e;
        return;
      }
    } else {
      dim = nodes[nodeIdx].dim();
      pos = nodes[nodeIdx].pivot;
      left = nodes[nodeIdx].idx();
    }
    float lLo[3], lHi[3], rLo[3], rHi[3];
    for (int i = 0; i < 3; i++) { lLo[i] = rLo[i] = blo[i]; lHi[i] = rHi[i] = bhi[i]; }
    lHi[dim] = pos;
    rLo[dim] = pos;
    auto mid = std::stable_partition(samples.begin() + b, samples.begin() + e, [&](const orc_sample& s) {
      const float p[3] = {s.px, s.py, s.pz};
      return p[dim] < pos;
    });
    const size_t m = size_t(mid - samples.begin());
    SampleStats sl, sr;
    if (nodes[left].isLeaf() || nodes[left + 1].isLeaf()) {
      sl = intStats(b, m, blo, bhi);
      sr = intStats(m, e, blo, bhi);
    }
    updateNode(left, depth + 1, lLo, lHi, b, m, sl);
    updateNode(left + 1, depth + 1, rLo, rHi, m, e, sr);
  }

  void updateTree() {  // KDTreeBuilder.h:76-118
    SampleStats st;
    if (nodes[0].isLeaf()) st = intStats(0, samples.size(), treeLo, treeHi);
    updateNode(0, 1, treeLo, treeHi, 0, samples.size(), st);
  }

  void buildKNN() {  // KNN.h:188-266 (exact 8-NN by brute force; ties by index)
    const size_t R_ = regions.size();
    nbIds.assign(R_ * 8, ~0u);
    nbPts.assign(R_ * 24, 0.f);
    nbSize.assign(R_, 0);
    std::vector<std::pair<float, uint32_t>> d(R_);
    for (size_t r = 0; r < R_; r++) {
      const float* q = regions[r].ss.mean;
      for (size_t i = 0; i < R_; i++) {
        const float* p = regions[i].ss.mean;
        float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
        float a = dx * dx;
        float b2 = dy * dy;
        float c2 = dz * dz;
        float s = a + b2;
        s = s + c2;
        d[i] = {s, uint32_t(i)};
      }
      const size_t k = std::min<size_t>(8, R_);
      std::partial_sort(d.begin(), d.begin() + k, d.end(), [](const std::pair<float, uint32_t>& a, const std::pair<float, uint32_t>& b) {
        return a.first < b.first || (a.first == b.first && a.second < b.second);
      });
      nbSize[r] = uint32_t(k);
      for (size_t i = 0; i < k; i++) {
        nbIds[r * 8 + i] = d[i].second;
        for (int j = 0; j < 3; j++) nbPts[r * 24 + j * 8 + i] = regions[d[i].second].ss.mean[j];
      }
    }
  }

  void fitOrUpdateRegion(size_t r, bool fit) {
    Region<R>& reg = regions[r];
    if (!fit && reg.splitFlag) {  // Field.h:457-461
      const R a = R(cfg.decayOnSpatialSplit);
      Suff<R>& s = reg.ts.suff;
      for (int k = 0; k < MAXK; k++) { for (int j = 0; j < 3; j++) s.swd[k][j] *= a; s.sw[k] *= a; s.sdw[k] *= a; }
      s.numSamples *= a; s.sumWeights *= a;
      // ComponentSplitStatistics::decay (Splitter:1146-1162): whole SIMD blocks of the device width
      SplitStats<R>& sp = reg.ts.split;
      const int W = int(cfg.vectorWidth);
      const int lim = std::min(MAXK, ((sp.K + W - 1) / W) * W);
      for (int k = 0; k < lim; k++) { sp.cov[k][0] *= a; sp.cov[k][1] *= a; sp.cov[k][2] *= a; sp.n[k] *= a; sp.sw[k] *= a; sp.sa[k] *= a; }
      reg.splitFlag = false;
    }
    const size_t n = reg.end - reg.begin;
    reg.info.numSamplesThisUpdate = uint32_t(n);
    if (n > 0) {
      if (cfg.deterministic) std::sort(samples.begin() + reg.begin, samples.begin() + reg.end, sampleLess);
      const float sm[3] = {reg.ss.mean[0], reg.ss.mean[1], reg.ss.mean[2]};
      std::vector<PreparedSample<R>> ps;
      if (!fit && cfg.parallax) {
        const R shift[3] = {reg.vmm.pivot[0] - R(sm[0]), reg.vmm.pivot[1] - R(sm[1]), reg.vmm.pivot[2] - R(sm[2])};
        suffShift(reg.ts.suff, reg.vmm, shift);
        vmmShift(reg.vmm, shift);
      }
      prepare<R>(samples.data() + reg.begin, n, reg.ss, ps, cfg.parallax != 0);
      if (fit) {
        factory.fit(reg.vmm, reg.ts, ps.data(), n, &reg.info);
        for (int j = 0; j < 3; j++) reg.vmm.pivot[j] = R(sm[j]);
        reg.valid = vmmValid(reg.vmm);
        reg.splitFlag = false;
      } else {
        factory.update(reg.vmm, reg.ts, ps.data(), n, &reg.info);
        reg.valid = vmmValid(reg.vmm) && suffValid(reg.ts.suff);
      }
    } else if (fit) {
      reg.valid = false;
      reg.splitFlag = false;
    }
    reg.begin = reg.end = 0;
  }

  void update(const orc_sample* in, size_t n, int threads) {
    if (n == 0) { iteration++; return; }
    samples.assign(in, in + n);
    const bool first = !initialized;
    if (first) {
      if (!boundsSet) estimateSceneBounds();
      for (int i = 0; i < 3; i++) { treeLo[i] = sceneLo[i]; treeHi[i] = sceneHi[i]; }
      nodes.clear();
      nodes.push_back(Node());
      nodes[0].setLeaf(0);
      regions.clear();
      regions.resize(1);
      for (int i = 0; i < 3; i++) { regions[0].lower[i] = sceneLo[i]; regions[0].upper[i] = sceneHi[i]; }
    }
    for (auto& r : regions) { r.begin = r.end = 0; r.info.numSamplesThisUpdate = 0; }
    updateTree();
    if (cfg.knnLookup) buildKNN();
    const size_t NR = regions.size();
    threads = std::max(1, threads);
    std::vector<std::thread> pool;
    std::atomic_size_t next(0);
    auto work = [&]() {
      for (;;) {
        const size_t r = next.fetch_add(1);
        if (r >= NR) break;
        fitOrUpdateRegion(r, first);
      }
    };
    for (int t = 0; t < threads; t++) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    initialized = true;
    iteration++;
  }

  // canonical DFS order (KDTree.h:256-283)
  void canonical(std::vector<Node>& out, std::vector<uint32_t>& dataOrder) const {
    out.clear();
    dataOrder.clear();
    if (nodes.empty()) return;
    out.push_back(nodes[0]);
    rearrange(nodes[0], 0, out, dataOrder);
  }
  void rearrange(const Node& n, uint32_t idx, std::vector<Node>& out, std::vector<uint32_t>& data) const {
    if (!n.isLeaf()) {
      const uint32_t l = n.idx();
      const Node nl = nodes[l], nr = nodes[l + 1];
      const uint32_t ni = uint32_t(out.size());
      out[idx].bits = (out[idx].bits & 0xC0000000u) | ni;
      out.push_back(nl);
      out.push_back(nr);
      rearrange(nl, ni, out, data);
      rearrange(nr, ni + 1, out, data);
    } else {
      const uint32_t nd = uint32_t(data.size());
      data.push_back(n.idx());
      out[idx].setLeaf(nd);
    }
  }

  uint32_t lookup(const float p[3]) const {  // KDTree.h:398-426 (node walk; same result as treelets)
    uint32_t i = 0;
    while (!nodes[i].isLeaf()) {
      const uint32_t d = nodes[i].dim();
      const float pv = nodes[i].pivot;
      i = nodes[i].idx() + (p[d] >= pv ? 1u : 0u);
    }
    return nodes[i].idx();
  }


  uint32_t knnPick(uint32_t leaf, const float p[3], float* u) const {
    return knnPickSlots(&nbIds[leaf * 8], &nbPts[leaf * 24], nbSize[leaf], p, u, cfg.vectorWidth != 4);
  }

  const Region<R>* getRegion(const float p[3], float* u, uint32_t& id) const {  // Field.h:109-145
    if (iteration == 0 || !initialized) return nullptr;
    for (int j = 0; j < 3; j++) if (!(p[j] >= treeLo[j] && p[j] <= treeHi[j])) return nullptr;
    const uint32_t leaf = lookup(p);
    id = leaf;
    if (cfg.knnLookup && *u >= 0.f) return &regions[knnPick(leaf, p, u)];
    return &regions[leaf];
  }
};

// ------------------------------------------------------- query distribution
template <class R> struct QueryDist {
  VMM<R> d[4];
  R weights[4];
  int num = 0;
  VMM<R> li;
  R integral = 1;
};

// PAVMM::product (PAVMM:609-671)
template <class R> static R vmmProduct(VMM<R>& v, R weight, const R md[3], R kappa, R normalization) {
  R integral = 0;
  for (int k = 0; k < v.K; k++) {
    R nm[3];
    for (int j = 0; j < 3; j++) nm[j] = v.kappa[k] * v.mu[k][j] + kappa * md[j];
    R nk = std::sqrt(dot3(nm, nm));
    const bool ok = nk > R(1e-3);
    if (!ok) nk = 0;
    for (int j = 0; j < 3; j++) nm[j] = ok ? nm[j] / nk : v.mu[k][j];
    const R ne = std::exp(R(-2) * nk);
    R nn = nk / (R(2.0 * M_PI) * (R(1) - ne));
    if (!ok) nn = R(ONE_OVER_FOUR_PI_D);
    R scale = (v.norm[k] * normalization) / nn;
    const R c0 = dot3(v.mu[k], nm), c1 = dot3(md, nm);
    const R ev = std::exp(v.kappa[k] * (c0 - R(1)) + kappa * (c1 - R(1)));
    scale *= ev;
    scale *= v.w[k] * weight;
    v.w[k] = scale;
    v.kappa[k] = nk;
    for (int j = 0; j < 3; j++) v.mu[k][j] = nm[j];
    v.norm[k] = nn;
    v.em2k[k] = ne;
    integral += v.w[k];
  }
  for (int k = 0; k < v.K; k++) v.w[k] /= integral;
  return integral;
}

// PAVMM::sample (PAVMM:737-806)
template <class R> static void vmmSample(const VMM<R>& v, R u0, R u1, R out[3]) {
  int sel = 0;
  R sum = 0, cdf = 0;
  for (;;) {
    cdf = v.w[sel];
    if (sum + cdf >= u1 || sel + 1 >= v.K) break;
    sum += cdf;
    sel++;
  }
  const R s1 = std::min(R(1) - R(FLT_EPSILON), (u1 - sum) / cdf);
  R local[3];
  const R kappa = v.kappa[sel];
  if (kappa == R(0)) {
    const R z = R(1) - R(2) * s1;
    const R r = std::sqrt(std::max(R(0), R(1) - z * z));
    const R phi = R(2.0 * M_PI) * u0;
    local[0] = r * std::cos(phi); local[1] = r * std::sin(phi); local[2] = z;
  } else {
    R ct = R(1) + std::log(R(1) + ((v.em2k[sel] - R(1)) * u0)) / kappa;
    ct = std::min(R(1), std::max(ct, R(-1)));
    const R st = std::sqrt(R(1) - ct * ct);
    const R phi = R(2.0 * M_PI) * s1;
    local[0] = st * std::cos(phi); local[1] = st * std::sin(phi); local[2] = ct;
  }
  R fx[3], fy[3];
  frame(v.mu[sel], fx, fy);
  for (int j = 0; j < 3; j++) out[j] = fx[j] * local[0] + fy[j] * local[1] + v.mu[sel][j] * local[2];
}

template <class R> static bool initQuery(const Field<R>& f, const float pos[3], float* u, QueryDist<R>& q, uint32_t& id) {
  const Region<R>* reg = f.getRegion(pos, u, id);
  if (!reg || !reg->valid) return false;
  q.li = reg->vmm;
  if (f.cfg.parallax) {
    const R shift[3] = {q.li.pivot[0] - R(pos[0]), q.li.pivot[1] - R(pos[1]), q.li.pivot[2] - R(pos[2])};
    vmmShift(q.li, shift);
  }
  q.d[0] = q.li;
  q.weights[0] = 1;
  q.num = 1;
  q.integral = 1;
  return true;
}

template <class R> static void applyCosine(QueryDist<R>& q, const float n[3]) {  // VMMSurfaceSamplingDistribution.h:52-60
  const float ck = 2.18853f;
  const float cn = 2.18853f / (2.0f * float(M_PI) * (1.0f - std::exp(-2.0f * 2.18853f)));
  const R md[3] = {R(n[0]), R(n[1]), R(n[2])};
  q.integral = vmmProduct(q.d[0], R(1), md, R(ck), R(cn));
}

template <class R> static void applyHG(QueryDist<R>& q, const float dir[3], float g) {  // VMMVolumeSamplingDistribution.h:54-80
  const float ag = std::fabs(g);
  const float step = (0.99f - 0.0f) / float(128);
  int idx = int(std::floor((ag - 0.0f) / step));
  idx = std::min(idx, 127);
  const HGRep& rep = g_hg[idx];
  R sum = 0;
  for (int i = 0; i < rep.K; i++) {
    q.d[i] = q.li;
    const float sgn = (g * rep.mc[i] > 0.f) ? 1.f : -1.f;
    const R od[3] = {R(dir[0] * sgn), R(dir[1] * sgn), R(dir[2] * sgn)};
    q.weights[i] = vmmProduct(q.d[i], R(rep.w[i]), od, R(rep.kappa[i]), R(rep.norm[i]));
    sum += q.weights[i];
  }
  for (int i = 0; i < rep.K; i++) q.weights[i] /= sum;
  q.num = rep.K;
  q.integral = sum;
}

template <class R> static R queryPdf(const QueryDist<R>& q, const R d[3]) {
  R pdf = 0;
  for (int i = 0; i < q.num; i++) pdf += q.weights[i] * vmmPdf(q.d[i], d);
  return pdf;
}
template <class R> static void querySample(const QueryDist<R>& q, R u0, R u1, R out[3]) {
  R w = 0;
  int i = 0;
  for (; i < q.num - 1; ++i) {
    const R nw = w + q.weights[i];
    if (nw > u0) break;
    w = nw;
  }
  vmmSample(q.d[i], (u0 - w) / q.weights[i], u1, out);
}

// ----------------------------------------------------------------- export
template <class R> static void exportRegion(const Region<R>& r, orc_region_record* o) {
  const uint32_t a = r.info.emIterations, b = r.info.partialEmIterations, c = r.info.numSplits, d = r.info.numMerges;
  const uint32_t nthis = r.info.numSamplesThisUpdate;
  memset(o, 0, sizeof(*o));
  o->K = r.vmm.K;
  for (int k = 0; k < MAXK; k++) {
    o->weights[k] = float(r.vmm.w[k]); o->kappas[k] = float(r.vmm.kappa[k]);
    for (int j = 0; j < 3; j++) o->meanDirections[k][j] = float(r.vmm.mu[k][j]);
    o->normalizations[k] = float(r.vmm.norm[k]); o->eMinus2Kappa[k] = float(r.vmm.em2k[k]);
    o->meanCosines[k] = float(r.vmm.mc[k]); o->distances[k] = float(r.vmm.dist[k]);
    for (int j = 0; j < 3; j++) o->sumOfWeightedDirections[k][j] = float(r.ts.suff.swd[k][j]);
    o->sumOfWeightedStats[k] = float(r.ts.suff.sw[k]); o->sumOfDistanceWeights[k] = float(r.ts.suff.sdw[k]);
    o->chiSquareMCEstimates[k] = float(r.ts.split.chi[k]);
    o->splitMeans[k][0] = float(r.ts.split.mean[k][0]); o->splitMeans[k][1] = float(r.ts.split.mean[k][1]);
    for (int j = 0; j < 3; j++) o->splitCovariances[k][j] = float(r.ts.split.cov[k][j]);
    o->splitNumSamples[k] = float(r.ts.split.n[k]); o->splitSumWeights[k] = float(r.ts.split.sw[k]);
    o->splitSumAssignedSamples[k] = float(r.ts.split.sa[k]);
  }
  for (int j = 0; j < 3; j++) o->pivot[j] = float(r.vmm.pivot[j]);
  o->sumWeights = float(r.ts.suff.sumWeights); o->numSamples = float(r.ts.suff.numSamples);
  o->overallNumSamples = float(r.ts.suff.overall); o->suffK = r.ts.suff.K; o->normalized = r.ts.suff.normalized;
  o->splitK = r.ts.split.K;
  o->numSamplesAfterLastSplit = r.ts.afterSplit; o->numSamplesAfterLastMerge = r.ts.afterMerge;
  for (int j = 0; j < 3; j++) {
    o->ssMean[j] = r.ss.mean[j]; o->ssVariance[j] = r.ss.var[j]; o->ssLower[j] = r.ss.lo[j]; o->ssUpper[j] = r.ss.hi[j];
    o->regionLower[j] = r.lower[j]; o->regionUpper[j] = r.upper[j];
  }
  o->ssNumSamples = r.ss.n;
  o->splitFlag = r.splitFlag; o->valid = r.valid;
  o->numSamplesThisUpdate = nthis;
  o->emIterations = a; o->partialEmIterations = b; o->numSplits = c; o->numMerges = d;
}

}  // namespace orc

// ===================================================================== C API
using namespace orc;

struct orc_field_s {
  int precision;
  Field<float>* f32 = nullptr;
  Field<double>* f64 = nullptr;
};

struct orc_region_s {
  int precision;
  orc_config cfg;
  Factory<float>* fac32 = nullptr;
  Factory<double>* fac64 = nullptr;
  Region<float> r32;
  Region<double> r64;
  bool first = true;
};

#define DISPATCH(F, EXPR) \
  do { if ((F)->precision == 0) { auto& fld = *(F)->f32; EXPR; } else { auto& fld = *(F)->f64; EXPR; } } while (0)

extern "C" {

void orc_config_defaults(orc_config* c, uint32_t maxSamplesPerLeaf) {
  memset(c, 0, sizeof(*c));
  c->knnLookup = 1; c->maxSamples = maxSamplesPerLeaf; c->maxDepth = 32;
  c->initK = 16; c->initKappa = 0.5f; c->maxK = 32; c->maxEMIterations = 100;
  c->maxKappa = 320000.f; c->convergenceThreshold = 0.005f; c->weightPrior = 0.01f;
  c->meanCosinePriorStrength = 0.2f; c->meanCosinePrior = 0.f;
  c->useSplitAndMerge = 1; c->splittingThreshold = 0.5f; c->mergingThreshold = 0.025f;
  c->partialReFit = 1;
  c->minSamplesForSplitting = int32_t(maxSamplesPerLeaf / 8);
  c->minSamplesForPartialRefitting = int32_t(maxSamplesPerLeaf / 8);
  c->minSamplesForMerging = int32_t(maxSamplesPerLeaf / 4);
  c->decayOnSpatialSplit = 0.25f; c->deterministic = 1; c->vectorWidth = 8;
  c->parallax = 1;
}

orc_field orc_field_new(const orc_config* cfg, int precision) {
  orc_field f = new orc_field_s;
  f->precision = precision;
  if (precision == 0) f->f32 = new Field<float>(*cfg); else f->f64 = new Field<double>(*cfg);
  return f;
}
void orc_field_release(orc_field f) { delete f->f32; delete f->f64; delete f; }
void orc_field_set_scene_bounds(orc_field f, const float* lo, const float* hi) {
  DISPATCH(f, { for (int i = 0; i < 3; i++) { fld.sceneLo[i] = lo[i]; fld.sceneHi[i] = hi[i]; } fld.boundsSet = true; });
}
void orc_field_get_scene_bounds(orc_field f, float* lo, float* hi) {
  DISPATCH(f, { for (int i = 0; i < 3; i++) { lo[i] = fld.sceneLo[i]; hi[i] = fld.sceneHi[i]; } });
}
void orc_field_update(orc_field f, const orc_sample* s, size_t n, int threads) { DISPATCH(f, fld.update(s, n, threads)); }
uint32_t orc_field_iteration(orc_field f) { uint32_t r = 0; DISPATCH(f, r = fld.iteration); return r; }
uint32_t orc_field_num_nodes(orc_field f) { uint32_t r = 0; DISPATCH(f, r = uint32_t(fld.nodes.size())); return r; }
uint32_t orc_field_num_regions(orc_field f) { uint32_t r = 0; DISPATCH(f, r = uint32_t(fld.regions.size())); return r; }
void orc_field_export_nodes(orc_field f, float* piv, uint32_t* dims, uint32_t* cod) {
  DISPATCH(f, {
    std::vector<Node> out; std::vector<uint32_t> data;
    fld.canonical(out, data);
    for (size_t i = 0; i < out.size(); i++) { piv[i] = out[i].pivot; dims[i] = out[i].dim(); cod[i] = out[i].idx(); }
  });
}
void orc_field_region_canonical_map(orc_field f, uint32_t* s2c) {
  DISPATCH(f, {
    std::vector<Node> out; std::vector<uint32_t> data;
    fld.canonical(out, data);
    for (size_t i = 0; i < fld.regions.size(); i++) s2c[i] = ~0u;
    for (size_t c = 0; c < data.size(); c++) s2c[data[c]] = uint32_t(c);
  });
}
void orc_field_export_region(orc_field f, uint32_t ci, orc_region_record* o) {
  DISPATCH(f, {
    std::vector<Node> out; std::vector<uint32_t> data;
    fld.canonical(out, data);
    exportRegion(fld.regions[data[ci]], o);
  });
}
void orc_field_sample_leaf_ids(orc_field f, const float* pos, size_t n, uint32_t* ids) {
  DISPATCH(f, {
    std::vector<Node> out; std::vector<uint32_t> data;
    fld.canonical(out, data);
    std::vector<uint32_t> s2c(fld.regions.size(), ~0u);
    for (size_t c = 0; c < data.size(); c++) s2c[data[c]] = uint32_t(c);
    for (size_t i = 0; i < n; i++) ids[i] = s2c[fld.lookup(pos + 3 * i)];
  });
}
void orc_field_knn(orc_field f, uint32_t* nb) {
  DISPATCH(f, {
    std::vector<Node> out; std::vector<uint32_t> data;
    fld.canonical(out, data);
    std::vector<uint32_t> s2c(fld.regions.size(), ~0u);
    for (size_t c = 0; c < data.size(); c++) s2c[data[c]] = uint32_t(c);
    for (size_t c = 0; c < data.size(); c++)
      for (int i = 0; i < 8; i++) {
        const uint32_t id = fld.nbIds.empty() ? ~0u : fld.nbIds[data[c] * 8 + i];
        nb[c * 8 + i] = id == ~0u ? ~0u : s2c[id];
      }
  });
}

}  // extern "C"

template <class R>
static int queryImpl(const Field<R>& fld, int volume, const float* pos, const float* nrm, float g, const float* u2, float u1,
                     float* dirOut, float* pdfOut, uint32_t* idOut, float* integral, const float* pdfDir) {
  QueryDist<R> q;
  uint32_t id = ~0u;
  float u = u1;
  if (!initQuery(fld, pos, &u, q, id)) return 0;
  if (volume) applyHG(q, nrm, g); else applyCosine(q, nrm);
  if (pdfDir) {
    const R d[3] = {R(pdfDir[0]), R(pdfDir[1]), R(pdfDir[2])};
    *pdfOut = float(queryPdf(q, d));
    return 1;
  }
  R d[3];
  querySample(q, R(u2[0]), R(u2[1]), d);
  for (int j = 0; j < 3; j++) dirOut[j] = float(d[j]);
  *pdfOut = float(queryPdf(q, d));
  if (idOut) *idOut = id;
  if (integral) *integral = float(q.integral);
  return 1;
}

// The scalar SamplingDistribution calls of a renderer (SurfaceVolumeField.h:50-85,
// VMM{Surface,Volume}SamplingDistribution.h): Init (u1 updated in place by the
// KNN draw), `products` cosine / HG products, SamplePDF(u2), then PDF and
// IncomingRadiancePDF at evalDir.  out: dir[3], pdf(dir), pdf(evalDir),
// pdfLi(evalDir), product integral.
template <class R>
static int scalarSequence(const Field<R>& fld, int volume, const float* pos, float* u1, const float* nrm, float g,
                          int products, const float* u2, const float* ev, float* out, uint32_t* idOut) {
  QueryDist<R> q;
  uint32_t id = ~0u;
  if (!initQuery(fld, pos, u1, q, id)) return 0;
  for (int p = 0; p < products; p++) {
    if (volume) applyHG(q, nrm, g); else applyCosine(q, nrm);
  }
  R d[3];
  querySample(q, R(u2[0]), R(u2[1]), d);
  float df[3];
  for (int j = 0; j < 3; j++) { df[j] = float(d[j]); out[j] = df[j]; }
  const R dd[3] = {R(df[0]), R(df[1]), R(df[2])};
  out[3] = float(queryPdf(q, dd));
  const R e[3] = {R(ev[0]), R(ev[1]), R(ev[2])};
  out[4] = float(queryPdf(q, e));
  out[5] = float(vmmPdf(q.li, e));
  out[6] = float(q.integral);
  if (idOut) *idOut = id;
  return 1;
}

template <class R>
static int pdfAtDirs(const Field<R>& fld, int volume, const float* pos, const float* nrm, float g, size_t n,
                     const float* dirs, float* pdfOut) {
  QueryDist<R> q;
  uint32_t id = ~0u;
  float u = -1.f;
  if (!initQuery(fld, pos, &u, q, id)) return 0;
  if (volume) applyHG(q, nrm, g); else applyCosine(q, nrm);
  for (size_t i = 0; i < n; i++) {
    const R d[3] = {R(dirs[3 * i]), R(dirs[3 * i + 1]), R(dirs[3 * i + 2])};
    pdfOut[i] = float(queryPdf(q, d));
  }
  return 1;
}

extern "C" {

int orc_field_scalar_sequence(orc_field f, int volume, const float* pos, float* u1, const float* nrm, float g,
                              int products, const float* u2, const float* evalDir, float* out7, uint32_t* id) {
  int r = 0;
  DISPATCH(f, r = scalarSequence(fld, volume, pos, u1, nrm, g, products, u2, evalDir, out7, id));
  return r;
}

// pdf of one query's distribution (Init at pos without KNN, one product) at n directions
int orc_field_query_pdf_dirs(orc_field f, int volume, const float* pos, const float* nrm, float g, size_t n,
                             const float* dirs, float* pdfOut) {
  int r = 0;
  DISPATCH(f, r = pdfAtDirs(fld, volume, pos, nrm, g, n, dirs, pdfOut));
  return r;
}

int orc_field_query(orc_field f, int volume, const float* pos, const float* nrm, float g, const float* u2, float u1,
                    float* dirOut, float* pdfOut, uint32_t* idOut, float* integral) {
  int r = 0;
  DISPATCH(f, r = queryImpl(fld, volume, pos, nrm, g, u2, u1, dirOut, pdfOut, idOut, integral, nullptr));
  return r;
}
int orc_field_query_pdf(orc_field f, int volume, const float* pos, const float* nrm, float g, float u1, const float* dir, float* pdfOut) {
  int r = 0;
  DISPATCH(f, r = queryImpl(fld, volume, pos, nrm, g, nullptr, u1, nullptr, pdfOut, nullptr, nullptr, dir));
  return r;
}
void orc_field_query_batch(orc_field f, int volume, size_t n, const float* pos, const float* nrm, const float* g,
                           const float* u2, const float* u1, float* dirOut, float* pdfOut, uint32_t* idOut,
                           float* integralOut, int threads) {
  threads = std::max(1, threads);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&, t]() {
      const size_t b = n * t / threads, e = n * (t + 1) / threads;
      for (size_t i = b; i < e; i++) {
        uint32_t id = ~0u;
        float integ = 0.f;
        int ok = 0;
        DISPATCH(f, ok = queryImpl(fld, volume, pos + 3 * i, nrm + 3 * i, g ? g[i] : 0.f, u2 + 2 * i, u1 ? u1[i] : -1.f,
                                   dirOut + 3 * i, pdfOut + i, &id, &integ, nullptr));
        if (!ok) { pdfOut[i] = 0.f; id = ~0u; dirOut[3 * i] = dirOut[3 * i + 1] = dirOut[3 * i + 2] = 0.f; }
        if (idOut) idOut[i] = id;
        if (integralOut) integralOut[i] = integ;
      }
    });
  for (auto& t : pool) t.join();
}

orc_region orc_region_new(const orc_config* cfg, int precision) {
  orc_region r = new orc_region_s;
  r->precision = precision;
  r->cfg = *cfg;
  if (precision == 0) r->fac32 = new Factory<float>(r->cfg); else r->fac64 = new Factory<double>(r->cfg);
  return r;
}
void orc_region_release(orc_region r) { delete r->fac32; delete r->fac64; delete r; }

}  // extern "C"

template <class R>
static void regionStep(Region<R>& reg, const Factory<R>& fac, bool& first, const orc_sample* s, size_t n, const float* m,
                       const float* v, float ns, int splitFlag) {
  for (int j = 0; j < 3; j++) { reg.ss.mean[j] = m[j]; reg.ss.var[j] = v[j]; }
  reg.ss.n = ns;
  std::vector<orc_sample> sorted(s, s + n);
  std::sort(sorted.begin(), sorted.end(), sampleLess);
  std::vector<PreparedSample<R>> ps;
  reg.info.numSamplesThisUpdate = uint32_t(n);
  if (first) {
    prepare<R>(sorted.data(), n, reg.ss, ps, fac.c.parallax != 0);
    fac.fit(reg.vmm, reg.ts, ps.data(), n, &reg.info);
    for (int j = 0; j < 3; j++) reg.vmm.pivot[j] = R(m[j]);
    reg.valid = vmmValid(reg.vmm);
    first = false;
  } else {
    if (splitFlag) {
      const R a = R(0.25);
      Suff<R>& st = reg.ts.suff;
      for (int k = 0; k < MAXK; k++) { for (int j = 0; j < 3; j++) st.swd[k][j] *= a; st.sw[k] *= a; st.sdw[k] *= a; }
      st.numSamples *= a; st.sumWeights *= a;
      // ComponentSplitStatistics::decay (Splitter:1146-1162): whole SIMD blocks of the device width
      SplitStats<R>& sp = reg.ts.split;
      const int W = int(fac.c.vectorWidth);
      const int lim = std::min(MAXK, ((sp.K + W - 1) / W) * W);
      for (int k = 0; k < lim; k++) { sp.cov[k][0] *= a; sp.cov[k][1] *= a; sp.cov[k][2] *= a; sp.n[k] *= a; sp.sw[k] *= a; sp.sa[k] *= a; }
    }
    if (fac.c.parallax) {
      const R shift[3] = {reg.vmm.pivot[0] - R(m[0]), reg.vmm.pivot[1] - R(m[1]), reg.vmm.pivot[2] - R(m[2])};
      suffShift(reg.ts.suff, reg.vmm, shift);
      vmmShift(reg.vmm, shift);
    }
    prepare<R>(sorted.data(), n, reg.ss, ps, fac.c.parallax != 0);
    fac.update(reg.vmm, reg.ts, ps.data(), n, &reg.info);
    reg.valid = vmmValid(reg.vmm) && suffValid(reg.ts.suff);
  }
}

extern "C" {

void orc_region_step(orc_region r, const orc_sample* s, size_t n, const float* m, const float* v, float ns, int splitFlag) {
  if (r->precision == 0) regionStep(r->r32, *r->fac32, r->first, s, n, m, v, ns, splitFlag);
  else regionStep(r->r64, *r->fac64, r->first, s, n, m, v, ns, splitFlag);
}
void orc_region_fit_phase(orc_region r, const orc_sample* s, size_t n, const float* m, const float* v, float ns, int phase) {
  if (r->precision == 0) r->fac32->stopPhase = phase; else r->fac64->stopPhase = phase;
  orc_region_step(r, s, n, m, v, ns, 0);
  if (r->precision == 0) r->fac32->stopPhase = 0; else r->fac64->stopPhase = 0;
}

void orc_region_export(orc_region r, orc_region_record* o) {
  if (r->precision == 0) exportRegion(r->r32, o); else exportRegion(r->r64, o);
}
float orc_region_pdf(orc_region r, const float* d) {
  if (r->precision == 0) { const float dd[3] = {d[0], d[1], d[2]}; return vmmPdf(r->r32.vmm, dd); }
  const double dd[3] = {d[0], d[1], d[2]};
  return float(vmmPdf(r->r64.vmm, dd));
}

}  // extern "C"

template <class R>
static float regionQuery(const Region<R>& reg, bool parallax, int volume, const float* pos, const float* nrm, float g,
                         const float* u2, const float* pdfDir, float* pdfAtDir, float* dirOut, float* pdfOut) {
  QueryDist<R> q;
  q.li = reg.vmm;
  if (parallax) {
    const R shift[3] = {q.li.pivot[0] - R(pos[0]), q.li.pivot[1] - R(pos[1]), q.li.pivot[2] - R(pos[2])};
    vmmShift(q.li, shift);
  }
  q.d[0] = q.li; q.weights[0] = 1; q.num = 1; q.integral = 1;
  if (volume) applyHG(q, nrm, g); else applyCosine(q, nrm);
  const R pd[3] = {R(pdfDir[0]), R(pdfDir[1]), R(pdfDir[2])};
  *pdfAtDir = float(queryPdf(q, pd));
  R d[3];
  querySample(q, R(u2[0]), R(u2[1]), d);
  for (int j = 0; j < 3; j++) dirOut[j] = float(d[j]);
  *pdfOut = float(queryPdf(q, d));
  return float(q.integral);
}

extern "C" {

float orc_region_query(orc_region r, int volume, const float* pos, const float* nrm, float g, const float* u2,
                       const float* pdfDir, float* pdfAtDir, float* dirOut, float* pdfOut) {
  const bool par = r->cfg.parallax != 0;
  if (r->precision == 0) return regionQuery(r->r32, par, volume, pos, nrm, g, u2, pdfDir, pdfAtDir, dirOut, pdfOut);
  return regionQuery(r->r64, par, volume, pos, nrm, g, u2, pdfDir, pdfAtDir, dirOut, pdfOut);
}

void orc_stats_merge_split(const float* a, const float* b, int dim, float pos, float* m, float* l, float* r) {
  auto load = [](const float* x) {
    SampleStats s;
    for (int j = 0; j < 3; j++) { s.mean[j] = x[j]; s.var[j] = x[3 + j]; s.lo[j] = x[7 + j]; s.hi[j] = x[10 + j]; }
    s.n = x[6];
    return s;
  };
  auto store = [](const SampleStats& s, float* x) {
    for (int j = 0; j < 3; j++) { x[j] = s.mean[j]; x[3 + j] = s.var[j]; x[7 + j] = s.lo[j]; x[10 + j] = s.hi[j]; }
    x[6] = s.n;
  };
  SampleStats sa = load(a), sb = load(b);
  SampleStats sm = sa;
  sm.merge(sb);
  store(sm, m);
  SampleStats sl = sa, sr = sa;
  sl.split(dim, pos, 0.25f, false);
  sr.split(dim, pos, 0.25f, true);
  store(sl, l);
  store(sr, r);
}

void orc_integer_stats(const float* pos, size_t n, const float* lo, const float* hi, float* out, int64_t* raw) {
  IntStats is(lo, hi);
  for (size_t i = 0; i < n; i++) is.add(pos + 3 * i);
  SampleStats s = is.toSampleStats();
  for (int j = 0; j < 3; j++) { out[j] = s.mean[j]; out[3 + j] = s.var[j]; out[7 + j] = s.lo[j]; out[10 + j] = s.hi[j]; }
  out[6] = s.n;
  if (raw) {
    for (int j = 0; j < 3; j++) { raw[j] = is.mean[j]; raw[3 + j] = is.var[j]; raw[7 + j] = is.lo[j]; raw[10 + j] = is.hi[j]; }
    raw[6] = is.n;
  }
}
float orc_rcp(float x) { return rcp_avx2(x); }
float orc_rcp_width(float x, int W) { return rcp_width(x, W); }
void orc_integer_stats_width(const float* pos, size_t n, const float* lo, const float* hi, int W, int64_t* raw) {
  IntStats is(lo, hi, W);
  for (size_t i = 0; i < n; i++) is.add(pos + 3 * i);
  for (int j = 0; j < 3; j++) { raw[j] = is.mean[j]; raw[3 + j] = is.var[j]; raw[7 + j] = is.lo[j]; raw[10 + j] = is.hi[j]; }
  raw[6] = is.n;
}
uint32_t orc_knn_pick(const uint32_t* ids8, const float* pts24, uint32_t size, const float* p, float* u, int fused) {
  return knnPickSlots(ids8, pts24, size, p, u, fused != 0);
}

void orc_set_hg_table(const float* t) {
  for (int i = 0; i < 128; i++) {
    const float* e = t + 18 * i;
    g_hg[i].K = int(e[0]); g_hg[i].g = e[1];
    for (int k = 0; k < 4; k++) { g_hg[i].w[k] = e[2 + k]; g_hg[i].mc[k] = e[6 + k]; g_hg[i].kappa[k] = e[10 + k]; g_hg[i].norm[k] = e[14 + k]; }
  }
  g_hgSet = true;
}
int orc_hg_rep(float g, float* w, float* mc, float* kappa, float* norm) {
  const float ag = std::fabs(g);
  const float step = (0.99f - 0.0f) / float(128);
  int idx = int(std::floor((ag - 0.0f) / step));
  idx = std::min(idx, 127);
  for (int k = 0; k < 4; k++) { w[k] = g_hg[idx].w[k]; mc[k] = g_hg[idx].mc[k]; kappa[k] = g_hg[idx].kappa[k]; norm[k] = g_hg[idx].norm[k]; }
  return g_hg[idx].K;
}

}  // extern "C"

// ---------------------------------------------------------------- path segments
// PathSegmentDataStorage::prepareSamples / calculatePixelEstimate
// (data/PathSegmentDataStorage.h:160-393) for one path.  Vector dot products
// are embree's madd chain (fused under the CPU_8 build), everything else
// separate IEEE operations in the reference's order.
namespace orc {

static inline float vdot(const float a[3], const float b[3]) { return std::fma(a[0], b[0], std::fma(a[1], b[1], a[2] * b[2])); }

}  // namespace orc

extern "C" {

size_t orc_path_prepare_samples(const orc_path_segment* segs, size_t n, float maxDistance, int useNEEMiWeights,
                                int guideDirectLight, int rrAffectsDirectContribution, orc_sample* out, size_t cap) {
  using orc::vdot;
  size_t written = 0;
  float lastDistance = 0.0f;
  for (long i = long(n) - 2; i >= 0; --i) {
    const orc_path_segment& cur = segs[i];
    float d[3];
    for (int k = 0; k < 3; k++) d[k] = segs[i + 1].position[k] - cur.position[k];
    const float currentDistance = std::sqrt(vdot(d, d));
    const float distance = std::fmin(currentDistance + lastDistance, 2.0f * maxDistance);
    if (!cur.isDelta && cur.roughness >= 0.3f) {
      lastDistance = 0.0f;
    } else {
      lastDistance = distance;
      if (cur.eta != 1.0f) {
        const float cosThetaI = vdot(cur.normal, cur.directionOut);
        const float cosThetaO = vdot(cur.normal, cur.directionIn);
        lastDistance *= std::fabs(cosThetaI / (cosThetaO * cur.eta));
      }
    }
    if (cur.isDelta || !(cur.roughness > 0.01f)) continue;
    const float pdf = std::max(0.1f, cur.pdfDirectionIn);
    float throughput[3] = {1.f, 1.f, 1.f}, contribution[3] = {0.f, 0.f, 0.f};
    float previousRR = 1.0f;
    for (size_t j = size_t(i) + 1; j < n; ++j) {
      const orc_path_segment& nxt = segs[j];
      float clamped[3];
      for (int k = 0; k < 3; k++) {
        throughput[k] = throughput[k] * segs[j - 1].transmittanceWeight[k];
        clamped[k] = throughput[k] < 10.0f ? throughput[k] : 10.0f;  // embree::min
        const float add = clamped[k] * nxt.scatteredContribution[k];
        contribution[k] = contribution[k] + add;
      }
      float direct[3];
      for (int k = 0; k < 3; k++) direct[k] = rrAffectsDirectContribution ? nxt.directContribution[k]
                                                                          : nxt.directContribution[k] * previousRR;
      const bool first = j == size_t(i) + 1;
      if (first && !useNEEMiWeights) {
        if (guideDirectLight)
          for (int k = 0; k < 3; k++) contribution[k] = contribution[k] + clamped[k] * direct[k];
      } else if (!first || guideDirectLight) {
        for (int k = 0; k < 3; k++) {
          const float cw = clamped[k] * nxt.miWeight;
          contribution[k] = contribution[k] + cw * direct[k];
        }
      }
      for (int k = 0; k < 3; k++) throughput[k] = throughput[k] * nxt.scatteringWeight[k];
      if (nxt.russianRouletteProbability > 0.f) {
        for (int k = 0; k < 3; k++) throughput[k] = throughput[k] / nxt.russianRouletteProbability;
      } else {
        for (int k = 0; k < 3; k++) throughput[k] = 0.f;
      }
      previousRR = nxt.russianRouletteProbability;
    }
    const bool positive = contribution[0] > 0.0f || contribution[1] > 0.0f || contribution[2] > 0.0f;
    if (!positive || !(distance > 0)) continue;
    if (written < cap) {
      orc_sample& s = out[written];
      s.px = cur.position[0]; s.py = cur.position[1]; s.pz = cur.position[2];
      s.dx = cur.directionIn[0]; s.dy = cur.directionIn[1]; s.dz = cur.directionIn[2];
      const float spectrum = (contribution[0] + contribution[1] + contribution[2]) / 3.0f;
      s.weight = spectrum / pdf;
      s.pdf = pdf;
      s.distance = distance;
      s.flags = cur.volumeScatter ? 2u : 0u;
    }
    written++;
  }
  return written;
}

size_t orc_path_prepare_batch(const orc_path_segment* segs, const uint32_t* offsets, size_t numPaths,
                              float maxDistance, int useNEEMiWeights, int guideDirectLight,
                              int rrAffectsDirectContribution, orc_sample* out, size_t cap) {
  size_t total = 0;
  for (size_t p = 0; p < numPaths; p++) {
    const size_t b = offsets[p], e = offsets[p + 1];
    const size_t room = total < cap ? cap - total : 0;
    total += orc_path_prepare_samples(segs + b, e - b, maxDistance, useNEEMiWeights, guideDirectLight,
                                      rrAffectsDirectContribution, out + (total < cap ? total : cap), room);
  }
  return total;
}

void orc_path_pixel_estimate(const orc_path_segment* segs, size_t n, int rrAffectsDirectContribution, float* rgb) {
  rgb[0] = rgb[1] = rgb[2] = 0.f;
  if (n == 0) return;
  float throughput[3] = {1.f, 1.f, 1.f}, contribution[3] = {0.f, 0.f, 0.f};
  float previousRR = 1.0f;
  for (size_t j = 1; j < n; ++j) {
    const orc_path_segment& nxt = segs[j];
    for (int k = 0; k < 3; k++) {
      throughput[k] = throughput[k] * segs[j - 1].transmittanceWeight[k];
      contribution[k] = contribution[k] + throughput[k] * nxt.scatteredContribution[k];
    }
    for (int k = 0; k < 3; k++) {
      const float direct = rrAffectsDirectContribution ? nxt.directContribution[k] : nxt.directContribution[k] * previousRR;
      const float tw = throughput[k] * nxt.miWeight;
      contribution[k] = contribution[k] + tw * direct;
    }
    for (int k = 0; k < 3; k++) throughput[k] = throughput[k] * nxt.scatteringWeight[k];
    if (nxt.russianRouletteProbability > 0.f) {
      for (int k = 0; k < 3; k++) throughput[k] = throughput[k] / nxt.russianRouletteProbability;
    } else {
      for (int k = 0; k < 3; k++) throughput[k] = 0.f;
    }
    previousRR = nxt.russianRouletteProbability;
  }
  for (int k = 0; k < 3; k++) {
    const float own = segs[0].directContribution[k] + segs[0].scatteredContribution[k];
    rgb[k] = n == 1 ? own : own + segs[0].scatteringWeight[k] * contribution[k];
  }
}

}  // extern "C"
