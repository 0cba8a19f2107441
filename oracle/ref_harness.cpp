// ORACLE — test infrastructure only.
//
// Harness that compiles the reference's own directional code (Open PGL 0.6.0
// headers under the reference tree, used where they lie) into
// oracle/_ref/libpglref.so, so tests can produce golden vectors and pin the
// restatement in pgl_oracle.cpp.  The reference's kd-tree builder, KNN and
// Field need oneTBB headers, which this image does not have: those parts are
// not built here.
//
// Mirrors, for a single region, the per-region body of Field::fitRegions
// (field/Field.h:389-439) and Field::updateRegions (field/Field.h:441-511),
// and exposes the reference's own Region serialisation (spatial/Region.h:63-80,
// data/Range.h:29-38) for SIMD widths 4 and 8 to pin the field-file layout.
#include <sstream>
#include <vector>
#include <algorithm>
#include <cstring>
#include <cstdlib>
#include <new>

#include "data/SampleStatistics.h"
#include "directional/vmm/ParallaxAwareVonMisesFisherMixture.h"
#include "directional/vmm/ParallaxAwareVonMisesFisherWeightedEMFactory.h"
#include "directional/vmm/AdaptiveSplitandMergeFactory.h"
#include "directional/vmm/VMMSurfaceSamplingDistribution.h"
#include "directional/vmm/VMMVolumeSamplingDistribution.h"
#include "spatial/Region.h"
#include "data/Range.h"

#include "pgl_oracle.h"

using namespace openpgl;

// P = parallax compensation: the reference's PARALLAX_AWARE_VMM (true) and
// VMM (false) field types (device/Device.h:93-184)
template <bool P>
struct RefRegionT {
  typedef ParallaxAwareVonMisesFisherMixture<8, 32, P> VMM;
  typedef AdaptiveSplitAndMergeFactory<VMM> ASM;
  VMM vmm;
  typename ASM::Statistics stats;
  SampleStatistics ss;
  typename ASM::Configuration cfg;
  bool first = true;
  bool valid = false;
  typename ASM::FittingStatistics fit;
  RefRegionT() : vmm(), stats(), ss() {}
};
typedef RefRegionT<true> RefRegion;
typedef RefRegion::VMM VMM8;
typedef RefRegion::ASM ASM8;

template <class Cfg>
static void toCfg(const orc_config* c, Cfg& a) {
  a.weightedEMCfg.initK = c->initK;
  a.weightedEMCfg.initKappa = c->initKappa;
  a.weightedEMCfg.maxK = c->maxK;
  a.weightedEMCfg.maxEMIterrations = c->maxEMIterations;
  a.weightedEMCfg.maxKappa = c->maxKappa;
  a.weightedEMCfg.maxMeanCosine = KappaToMeanCosine<float>(c->maxKappa);
  a.weightedEMCfg.convergenceThreshold = c->convergenceThreshold;
  a.weightedEMCfg.weightPrior = c->weightPrior;
  a.weightedEMCfg.meanCosinePriorStrength = c->meanCosinePriorStrength;
  a.weightedEMCfg.meanCosinePrior = c->meanCosinePrior;
  a.splittingThreshold = c->splittingThreshold;
  a.mergingThreshold = c->mergingThreshold;
  a.partialReFit = c->partialReFit;
  a.maxSplitItr = 1;
  a.useSplitAndMerge = c->useSplitAndMerge;
  a.minSamplesForSplitting = c->minSamplesForSplitting;
  a.minSamplesForPartialRefitting = c->minSamplesForPartialRefitting;
  a.minSamplesForMerging = c->minSamplesForMerging;
}


template <bool P>
static void* region_new(const orc_config* c) {
  static bool hgInit = false;
  if (!hgInit) { VMMSingleLobeHenyeyGreensteinOracle::init(); hgInit = true; }
  void* mem = nullptr;
  if (posix_memalign(&mem, 64, sizeof(RefRegionT<P>)) != 0) return nullptr;
  RefRegionT<P>* r = new (mem) RefRegionT<P>();  // vfloat<8> members need 32-byte alignment
  toCfg(c, r->cfg);
  return r;
}
template <bool P>
static void region_release(void* p) {
  ((RefRegionT<P>*)p)->~RefRegionT<P>();
  free(p);
}

// the per-region body of Field::fitRegions / Field::updateRegions
// (field/Field.h:405-438, 455-511)
template <bool P>
static void region_step(void* p, const orc_sample* s, size_t n, const float* m, const float* v, float ns, int splitFlag) {
  typedef RefRegionT<P> RR;
  RR* r = (RR*)p;
  r->ss.mean = Point3(m[0], m[1], m[2]);
  r->ss.sampleVariance = Vector3(v[0], v[1], v[2]);
  r->ss.numSamples = ns;
  std::vector<SampleData> data(n);
  memcpy(data.data(), s, n * sizeof(SampleData));
  std::sort(data.begin(), data.end(), SampleDataLess);
  typename RR::ASM factory;
  r->fit = typename RR::ASM::FittingStatistics();
  const Point3 sampleMean = r->ss.mean;
  if (r->first) {
    factory.prepareSamples(data.data(), n, r->ss, r->cfg);
    factory.fit(r->vmm, r->stats, data.data(), n, r->cfg, r->fit);
    r->vmm._pivotPosition = sampleMean;
    r->valid = r->vmm.isValid();
    r->first = false;
  } else {
    if (splitFlag) r->stats.decay(0.25f);
    if (RR::VMM::ParallaxCompensation == 1) {
      r->stats.sufficientStatistics.applyParallaxShift(r->vmm, r->vmm._pivotPosition - sampleMean);
      r->vmm.performRelativeParallaxShift(r->vmm._pivotPosition - sampleMean);
    }
    factory.prepareSamples(data.data(), n, r->ss, r->cfg);
    factory.update(r->vmm, r->stats, data.data(), n, r->cfg, r->fit);
    r->valid = r->vmm.isValid() && r->stats.isValid();
  }
}

template <bool P>
static void region_export(void* p, orc_region_record* o) {
  RefRegionT<P>* r = (RefRegionT<P>*)p;
  memset(o, 0, sizeof(*o));
  const typename RefRegionT<P>::VMM& v = r->vmm;
  o->K = v._numComponents;
  const auto& su = r->stats.sufficientStatistics;
  const auto& sp = r->stats.splittingStatistics;
  for (int k = 0; k < 32; k++) {
    const int q = k / 8, l = k % 8;
    o->weights[k] = v._weights[q][l];
    o->kappas[k] = v._kappas[q][l];
    o->meanDirections[k][0] = v._meanDirections[q].x[l];
    o->meanDirections[k][1] = v._meanDirections[q].y[l];
    o->meanDirections[k][2] = v._meanDirections[q].z[l];
    o->normalizations[k] = v._normalizations[q][l];
    o->eMinus2Kappa[k] = v._eMinus2Kappa[q][l];
    o->meanCosines[k] = v._meanCosines[q][l];
    o->distances[k] = v._distances[q][l];
    o->sumOfWeightedDirections[k][0] = su.sumOfWeightedDirections[q].x[l];
    o->sumOfWeightedDirections[k][1] = su.sumOfWeightedDirections[q].y[l];
    o->sumOfWeightedDirections[k][2] = su.sumOfWeightedDirections[q].z[l];
    o->sumOfWeightedStats[k] = su.sumOfWeightedStats[q][l];
    o->sumOfDistanceWeights[k] = su.sumOfDistanceWeightes[q][l];
    o->chiSquareMCEstimates[k] = sp.chiSquareMCEstimates[q][l];
    o->splitMeans[k][0] = sp.splitMeans[q].x[l];
    o->splitMeans[k][1] = sp.splitMeans[q].y[l];
    o->splitCovariances[k][0] = sp.splitWeightedSampleCovariances[q].x[l];
    o->splitCovariances[k][1] = sp.splitWeightedSampleCovariances[q].y[l];
    o->splitCovariances[k][2] = sp.splitWeightedSampleCovariances[q].z[l];
    o->splitNumSamples[k] = sp.numSamples[q][l];
    o->splitSumWeights[k] = sp.sumWeights[q][l];
    o->splitSumAssignedSamples[k] = sp.sumAssignedSamples[q][l];
  }
  o->pivot[0] = v._pivotPosition.x; o->pivot[1] = v._pivotPosition.y; o->pivot[2] = v._pivotPosition.z;
  o->sumWeights = su.sumWeights; o->numSamples = su.numSamples; o->overallNumSamples = su.overallNumSamples;
  o->suffK = su.numComponents; o->normalized = su.normalized;
  o->splitK = sp.numComponents;
  o->numSamplesAfterLastSplit = r->stats.numSamplesAfterLastSplit;
  o->numSamplesAfterLastMerge = r->stats.numSamplesAfterLastMerge;
  for (int j = 0; j < 3; j++) {
    o->ssMean[j] = r->ss.mean[j]; o->ssVariance[j] = r->ss.sampleVariance[j];
  }
  o->ssNumSamples = r->ss.numSamples;
  o->valid = r->valid;
  o->numSplits = r->fit.numSplits;
  o->numMerges = r->fit.numMerges;
  o->emIterations = r->fit.numUpdateWEMIterations;
  o->partialEmIterations = r->fit.numPartialUpdateWEMIterations;
}

template <bool P>
static float region_pdf(void* p, const float* d) {
  RefRegionT<P>* r = (RefRegionT<P>*)p;
  return r->vmm.pdf(Vector3(d[0], d[1], d[2]));
}

// init at pos (parallax shift) + cosine (surface) or HG (volume) product, then
// pdf(dir) and sample(u) -> dirOut/pdfOut.  Returns the product integral.
template <bool P>
static float region_query(void* p, int volume, const float* pos, const float* nrm, float g, const float* u2, const float* pdfDir,
                       float* pdfAtDir, float* dirOut, float* pdfOut) {
  RefRegionT<P>* r = (RefRegionT<P>*)p;
  float integral = 1.f;
  if (volume) {
    VMMVolumeSamplingDistribution<typename RefRegionT<P>::VMM, P> d;
    d.init(&r->vmm, Point3(pos[0], pos[1], pos[2]));
    d.applySingleLobeHenyeyGreensteinProduct(Vector3(nrm[0], nrm[1], nrm[2]), g);
    integral = d.m_productIntegral;
    *pdfAtDir = d.pdf(Vector3(pdfDir[0], pdfDir[1], pdfDir[2]));
    Vector3 o = d.sample(Point2(u2[0], u2[1]));
    dirOut[0] = o.x; dirOut[1] = o.y; dirOut[2] = o.z;
    *pdfOut = d.pdf(o);
  } else {
    VMMSurfaceSamplingDistribution<typename RefRegionT<P>::VMM, P> d;
    d.init(&r->vmm, Point3(pos[0], pos[1], pos[2]));
    d.applyCosineProduct(Vector3(nrm[0], nrm[1], nrm[2]));
    integral = d.m_productIntegral;
    *pdfAtDir = d.pdf(Vector3(pdfDir[0], pdfDir[1], pdfDir[2]));
    Vector3 o = d.sample(Point2(u2[0], u2[1]));
    dirOut[0] = o.x; dirOut[1] = o.y; dirOut[2] = o.z;
    *pdfOut = d.pdf(o);
  }
  return integral;
}

// ---------------------------------------------------------------- field-file records
// A storage entry of Field::m_regionStorageContainer (Region + Range) of a
// device of SIMD width W, trained by the reference's own factory.
template <int W>
struct FileRegion {
  typedef ParallaxAwareVonMisesFisherMixture<W, 32, true> VMM;
  typedef AdaptiveSplitAndMergeFactory<VMM> ASM;
  Region<VMM, typename ASM::Statistics> region;
  Range range;
  typename ASM::Configuration cfg;
  bool first = true;
};

template <int W>
static void* file_region_new(const orc_config* c) {
  typedef FileRegion<W> FR;
  void* mem = nullptr;
  if (posix_memalign(&mem, 64, sizeof(FR)) != 0) return nullptr;
  FR* r = new (mem) FR();
  auto& a = r->cfg;
  a.weightedEMCfg.initK = c->initK;
  a.weightedEMCfg.initKappa = c->initKappa;
  a.weightedEMCfg.maxK = c->maxK;
  a.weightedEMCfg.maxEMIterrations = c->maxEMIterations;
  a.weightedEMCfg.maxKappa = c->maxKappa;
  a.weightedEMCfg.maxMeanCosine = KappaToMeanCosine<float>(c->maxKappa);
  a.weightedEMCfg.convergenceThreshold = c->convergenceThreshold;
  a.weightedEMCfg.weightPrior = c->weightPrior;
  a.weightedEMCfg.meanCosinePriorStrength = c->meanCosinePriorStrength;
  a.weightedEMCfg.meanCosinePrior = c->meanCosinePrior;
  a.splittingThreshold = c->splittingThreshold;
  a.mergingThreshold = c->mergingThreshold;
  a.partialReFit = c->partialReFit;
  a.maxSplitItr = 1;
  a.useSplitAndMerge = c->useSplitAndMerge;
  a.minSamplesForSplitting = c->minSamplesForSplitting;
  a.minSamplesForPartialRefitting = c->minSamplesForPartialRefitting;
  a.minSamplesForMerging = c->minSamplesForMerging;
  return r;
}

// fit (first call) or update of the region with the given sample statistics
// (13 floats: mean, sampleVariance, numSamples, lower, upper) and region bounds
template <int W>
static void file_region_step(void* p, const orc_sample* s, size_t n, const float* ss13, const float* bounds6,
                             int splitFlag) {
  typedef FileRegion<W> FR;
  FR* r = (FR*)p;
  auto& R = r->region;
  R.sampleStatistics.mean = Point3(ss13[0], ss13[1], ss13[2]);
  R.sampleStatistics.sampleVariance = Vector3(ss13[3], ss13[4], ss13[5]);
  R.sampleStatistics.numSamples = ss13[6];
  R.sampleStatistics.sampleBounds.lower = Vector3(ss13[7], ss13[8], ss13[9]);
  R.sampleStatistics.sampleBounds.upper = Vector3(ss13[10], ss13[11], ss13[12]);
  R.regionBounds.lower = Vector3(bounds6[0], bounds6[1], bounds6[2]);
  R.regionBounds.upper = Vector3(bounds6[3], bounds6[4], bounds6[5]);
  std::vector<SampleData> data(n);
  memcpy(data.data(), s, n * sizeof(SampleData));
  std::sort(data.begin(), data.end(), SampleDataLess);
  typename FR::ASM factory;
  typename FR::ASM::FittingStatistics fit;
  const Point3 sampleMean = R.sampleStatistics.mean;
  if (r->first) {
    factory.prepareSamples(data.data(), n, R.sampleStatistics, r->cfg);
    factory.fit(R.distribution, R.trainingStatistics, data.data(), n, r->cfg, fit);
    R.distribution._pivotPosition = sampleMean;
    R.valid = R.distribution.isValid();
    R.splitFlag = false;
    r->first = false;
  } else {
    if (splitFlag) R.trainingStatistics.decay(0.25f);
    R.trainingStatistics.sufficientStatistics.applyParallaxShift(R.distribution, R.distribution._pivotPosition - sampleMean);
    R.distribution.performRelativeParallaxShift(R.distribution._pivotPosition - sampleMean);
    factory.prepareSamples(data.data(), n, R.sampleStatistics, r->cfg);
    factory.update(R.distribution, R.trainingStatistics, data.data(), n, r->cfg, fit);
    R.valid = R.distribution.isValid() && R.trainingStatistics.isValid();
    R.splitFlag = false;
  }
  r->range.reset();
}

template <int W>
static size_t file_region_serialize(void* p, char* buf, size_t cap) {
  FileRegion<W>* r = (FileRegion<W>*)p;
  std::ostringstream os;
  r->region.serialize(os);
  r->range.serialize(os);
  const std::string b = os.str();
  if (buf && cap >= b.size()) memcpy(buf, b.data(), b.size());
  return b.size();
}

template <int W>
static void file_region_export(void* p, orc_region_record* o) {
  FileRegion<W>* r = (FileRegion<W>*)p;
  const auto& R = r->region;
  memset(o, 0, sizeof(*o));
  const auto& v = R.distribution;
  const auto& su = R.trainingStatistics.sufficientStatistics;
  const auto& sp = R.trainingStatistics.splittingStatistics;
  o->K = v._numComponents;
  for (int k = 0; k < 32; k++) {
    const int q = k / W, l = k % W;
    o->weights[k] = v._weights[q][l];
    o->kappas[k] = v._kappas[q][l];
    o->meanDirections[k][0] = v._meanDirections[q].x[l];
    o->meanDirections[k][1] = v._meanDirections[q].y[l];
    o->meanDirections[k][2] = v._meanDirections[q].z[l];
    o->normalizations[k] = v._normalizations[q][l];
    o->eMinus2Kappa[k] = v._eMinus2Kappa[q][l];
    o->meanCosines[k] = v._meanCosines[q][l];
    o->distances[k] = v._distances[q][l];
    o->sumOfWeightedDirections[k][0] = su.sumOfWeightedDirections[q].x[l];
    o->sumOfWeightedDirections[k][1] = su.sumOfWeightedDirections[q].y[l];
    o->sumOfWeightedDirections[k][2] = su.sumOfWeightedDirections[q].z[l];
    o->sumOfWeightedStats[k] = su.sumOfWeightedStats[q][l];
    o->sumOfDistanceWeights[k] = su.sumOfDistanceWeightes[q][l];
    o->chiSquareMCEstimates[k] = sp.chiSquareMCEstimates[q][l];
    o->splitMeans[k][0] = sp.splitMeans[q].x[l];
    o->splitMeans[k][1] = sp.splitMeans[q].y[l];
    o->splitCovariances[k][0] = sp.splitWeightedSampleCovariances[q].x[l];
    o->splitCovariances[k][1] = sp.splitWeightedSampleCovariances[q].y[l];
    o->splitCovariances[k][2] = sp.splitWeightedSampleCovariances[q].z[l];
    o->splitNumSamples[k] = sp.numSamples[q][l];
    o->splitSumWeights[k] = sp.sumWeights[q][l];
    o->splitSumAssignedSamples[k] = sp.sumAssignedSamples[q][l];
  }
  o->pivot[0] = v._pivotPosition.x; o->pivot[1] = v._pivotPosition.y; o->pivot[2] = v._pivotPosition.z;
  o->sumWeights = su.sumWeights; o->numSamples = su.numSamples; o->overallNumSamples = su.overallNumSamples;
  o->suffK = su.numComponents; o->normalized = su.normalized;
  o->splitK = sp.numComponents;
  o->numSamplesAfterLastSplit = R.trainingStatistics.numSamplesAfterLastSplit;
  o->numSamplesAfterLastMerge = R.trainingStatistics.numSamplesAfterLastMerge;
  for (int j = 0; j < 3; j++) {
    o->ssMean[j] = R.sampleStatistics.mean[j];
    o->ssVariance[j] = R.sampleStatistics.sampleVariance[j];
    o->ssLower[j] = R.sampleStatistics.sampleBounds.lower[j];
    o->ssUpper[j] = R.sampleStatistics.sampleBounds.upper[j];
    o->regionLower[j] = R.regionBounds.lower[j];
    o->regionUpper[j] = R.regionBounds.upper[j];
  }
  o->ssNumSamples = R.sampleStatistics.numSamples;
  o->splitFlag = R.splitFlag;
  o->valid = R.valid;
}

template <int W>
static float file_region_pdf(void* p, const float* d) {
  return ((FileRegion<W>*)p)->region.distribution.pdf(Vector3(d[0], d[1], d[2]));
}

template <int W>
static void file_region_release(void* p) {
  ((FileRegion<W>*)p)->~FileRegion<W>();
  free(p);
}

// a storage entry read back with the reference's Region/Range deserialisation
template <int W>
static void* file_region_deserialize(const char* buf, size_t n) {
  typedef FileRegion<W> FR;
  void* mem = nullptr;
  if (posix_memalign(&mem, 64, sizeof(FR)) != 0) return nullptr;
  FR* r = new (mem) FR();
  std::istringstream is(std::string(buf, n));
  r->region.deserialize(is);
  r->range.deserialize(is);
  r->first = false;
  if (!is) {
    file_region_release<W>(r);
    return nullptr;
  }
  return r;
}

extern "C" {

void* ref_region_new(const orc_config* c) { return region_new<true>(c); }
void ref_region_release(void* p) { region_release<true>(p); }
void ref_region_step(void* p, const orc_sample* s, size_t n, const float* m, const float* v, float ns, int splitFlag) {
  region_step<true>(p, s, n, m, v, ns, splitFlag);
}
void* ref_region_np_new(const orc_config* c) { return region_new<false>(c); }
void ref_region_np_release(void* p) { region_release<false>(p); }
void ref_region_np_step(void* p, const orc_sample* s, size_t n, const float* m, const float* v, float ns, int splitFlag) {
  region_step<false>(p, s, n, m, v, ns, splitFlag);
}

// ASM::fit stopped after a phase: 1 EM + distances, 2 recursive splitting,
// 3 split statistics, 4 merging (AdaptiveSplitandMergeFactory.h:303-347)
void ref_region_fit_phase(void* p, const orc_sample* s, size_t n, const float* m, const float* v, float ns, int phase) {
  RefRegion* r = (RefRegion*)p;
  r->ss.mean = Point3(m[0], m[1], m[2]);
  r->ss.sampleVariance = Vector3(v[0], v[1], v[2]);
  r->ss.numSamples = ns;
  std::vector<SampleData> data(n);
  memcpy(data.data(), s, n * sizeof(SampleData));
  std::sort(data.begin(), data.end(), SampleDataLess);
  ASM8 factory;
  factory.prepareSamples(data.data(), n, r->ss, r->cfg);
  typedef ASM8::WeightedEMFactory WEM;
  const auto& cfg = r->cfg;
  r->stats.clear(cfg.weightedEMCfg.initK);
  WEM wem;
  typename WEM::FittingStatistics fs;
  wem.fitMixture(r->vmm, r->stats.sufficientStatistics, data.data(), n, cfg.weightedEMCfg, fs);
  wem.initComponentDistances(r->vmm, r->stats.sufficientStatistics, data.data(), n);
  if (phase >= 2) {
    float mc = r->stats.sufficientStatistics.getSumWeights() / r->stats.sufficientStatistics.getNumSamples();
    ASM8::Splitter splitter;
    splitter.PerformRecursiveSplitting(r->vmm, r->stats.sufficientStatistics, cfg.splittingThreshold, mc, data.data(), n, cfg.weightedEMCfg);
    if (phase >= 3) splitter.CalculateSplitStatistics(r->vmm, r->stats.splittingStatistics, mc, data.data(), n);
    if (phase >= 4) {
      ASM8::Merger merger;
      merger.PerformMerging(r->vmm, cfg.mergingThreshold, r->stats.sufficientStatistics, r->stats.splittingStatistics);
    }
  }
  r->first = false;
}

// one ThresholdedMergeNext step (VMMChiSquareComponentMerger.h:218-279); returns the merge cost or -1
float ref_region_merge_step(void* p) {
  RefRegion* r = (RefRegion*)p;
  ASM8::Merger merger;
  float cost = -1.f;
  const bool found = merger.ThresholdedMergeNext(r->vmm, r->cfg.mergingThreshold, cost, r->stats.sufficientStatistics,
                                                 r->stats.splittingStatistics);
  return found ? cost : 1e30f;
}
float ref_region_merge_cost(void* p, int i, int j) {
  RefRegion* r = (RefRegion*)p;
  ASM8::Merger merger;
  return merger.CalculateMergeCost(r->vmm, i, j);
}

void ref_region_export(void* p, orc_region_record* o) { region_export<true>(p, o); }
float ref_region_pdf(void* p, const float* d) { return region_pdf<true>(p, d); }
float ref_region_query(void* p, int volume, const float* pos, const float* nrm, float g, const float* u2, const float* pdfDir,
                       float* pdfAtDir, float* dirOut, float* pdfOut) {
  return region_query<true>(p, volume, pos, nrm, g, u2, pdfDir, pdfAtDir, dirOut, pdfOut);
}
void ref_region_np_export(void* p, orc_region_record* o) { region_export<false>(p, o); }
float ref_region_np_pdf(void* p, const float* d) { return region_pdf<false>(p, d); }
float ref_region_np_query(void* p, int volume, const float* pos, const float* nrm, float g, const float* u2,
                          const float* pdfDir, float* pdfAtDir, float* dirOut, float* pdfOut) {
  return region_query<false>(p, volume, pos, nrm, g, u2, pdfDir, pdfAtDir, dirOut, pdfOut);
}

void ref_integer_stats(const float* pos, size_t n, const float* lo, const float* hi, float* out, int64_t* raw) {
  BBox b(Vector3(lo[0], lo[1], lo[2]), Vector3(hi[0], hi[1], hi[2]));
  IntegerSampleStatistics is(b);
  for (size_t i = 0; i < n; i++) is.addSample(Point3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
  SampleStatistics s = is.getSampleStatistics();
  for (int j = 0; j < 3; j++) {
    out[j] = s.mean[j]; out[3 + j] = s.sampleVariance[j];
    out[7 + j] = s.sampleBounds.lower[j]; out[10 + j] = s.sampleBounds.upper[j];
    raw[j] = is.mean[j]; raw[3 + j] = is.variance[j];
    raw[7 + j] = is.sampleBounds.lower[j]; raw[10 + j] = is.sampleBounds.upper[j];
  }
  out[6] = s.numSamples;
  raw[6] = is.numSamples;
}

// SampleStatistics::merge / split exercised on given inputs (KDTreeBuilder.h:254-263)
void ref_stats_merge_split(const float* a13, const float* b13, int dim, float pos, float* merged13, float* left13, float* right13) {
  auto load = [](const float* x) {
    SampleStatistics s;
    s.mean = Point3(x[0], x[1], x[2]); s.sampleVariance = Vector3(x[3], x[4], x[5]); s.numSamples = x[6];
    s.sampleBounds.lower = Vector3(x[7], x[8], x[9]); s.sampleBounds.upper = Vector3(x[10], x[11], x[12]);
    return s;
  };
  auto store = [](const SampleStatistics& s, float* x) {
    for (int j = 0; j < 3; j++) { x[j] = s.mean[j]; x[3 + j] = s.sampleVariance[j]; x[7 + j] = s.sampleBounds.lower[j]; x[10 + j] = s.sampleBounds.upper[j]; }
    x[6] = s.numSamples;
  };
  SampleStatistics a = load(a13), b = load(b13);
  SampleStatistics m = a;
  m.merge(b);
  store(m, merged13);
  SampleStatistics l = a, r = a;
  l.split(dim, pos, 0.25f, false);
  r.split(dim, pos, 0.25f, true);
  store(l, left13);
  store(r, right13);
}

float ref_rcp(float x) { return embree::rcp(x); }

// sizeof(PGLPathSegmentData) and the offsets of its members in declaration order
// (include/openpgl/data.h:50-113).  PathSegmentDataStorage itself is not built
// here: it includes SampleDataStorage.h, which needs oneTBB.
void ref_path_segment_layout(size_t* out) {
  out[0] = sizeof(PGLPathSegmentData);
  out[1] = offsetof(PGLPathSegmentData, position);
  out[2] = offsetof(PGLPathSegmentData, directionIn);
  out[3] = offsetof(PGLPathSegmentData, directionOut);
  out[4] = offsetof(PGLPathSegmentData, normal);
  out[5] = offsetof(PGLPathSegmentData, volumeScatter);
  out[6] = offsetof(PGLPathSegmentData, pdfDirectionIn);
  out[7] = offsetof(PGLPathSegmentData, isDelta);
  out[8] = offsetof(PGLPathSegmentData, scatteringWeight);
  out[9] = offsetof(PGLPathSegmentData, transmittanceWeight);
  out[10] = offsetof(PGLPathSegmentData, directContribution);
  out[11] = offsetof(PGLPathSegmentData, miWeight);
  out[12] = offsetof(PGLPathSegmentData, scatteredContribution);
  out[13] = offsetof(PGLPathSegmentData, russianRouletteProbability);
  out[14] = offsetof(PGLPathSegmentData, eta);
  out[15] = offsetof(PGLPathSegmentData, roughness);
  out[16] = offsetof(PGLPathSegmentData, regionPtr);
}

// field-file storage entries for SIMD width W (4 or 8; 16 needs AVX-512 builds)
void* ref_file_region_new(const orc_config* c, int W) {
  return W == 4 ? file_region_new<4>(c) : (W == 8 ? file_region_new<8>(c) : nullptr);
}
void ref_file_region_step(void* p, int W, const orc_sample* s, size_t n, const float* ss13, const float* bounds6,
                          int splitFlag) {
  if (W == 4) file_region_step<4>(p, s, n, ss13, bounds6, splitFlag);
  else file_region_step<8>(p, s, n, ss13, bounds6, splitFlag);
}
size_t ref_file_region_serialize(void* p, int W, char* buf, size_t cap) {
  return W == 4 ? file_region_serialize<4>(p, buf, cap) : file_region_serialize<8>(p, buf, cap);
}
void* ref_file_region_deserialize(const char* buf, size_t n, int W) {
  return W == 4 ? file_region_deserialize<4>(buf, n) : (W == 8 ? file_region_deserialize<8>(buf, n) : nullptr);
}
void ref_file_region_export(void* p, int W, orc_region_record* o) {
  if (W == 4) file_region_export<4>(p, o);
  else file_region_export<8>(p, o);
}
float ref_file_region_pdf(void* p, int W, const float* d) {
  return W == 4 ? file_region_pdf<4>(p, d) : file_region_pdf<8>(p, d);
}
void ref_file_region_release(void* p, int W) {
  if (W == 4) file_region_release<4>(p);
  else file_region_release<8>(p);
}
// AdaptiveSplitAndMergeFactory::Configuration::serialize of the given settings
size_t ref_file_config_serialize(const orc_config* c, char* buf, size_t cap) {
  FileRegion<8>* r = (FileRegion<8>*)file_region_new<8>(c);
  std::ostringstream os;
  r->cfg.serialize(os);
  file_region_release<8>(r);
  const std::string b = os.str();
  if (buf && cap >= b.size()) memcpy(buf, b.data(), b.size());
  return b.size();
}

void ref_hg_table(float* t) {
  VMMSingleLobeHenyeyGreensteinOracle::init();
  for (int i = 0; i < 128; i++) {
    const VMMPhaseFunctionRepresentation& rep = VMMSingleLobeHenyeyGreensteinOracle::representations[i];
    float* e = t + 18 * i;
    e[0] = float(rep.K); e[1] = rep.g;
    for (int k = 0; k < 4; k++) { e[2 + k] = rep.weights[k]; e[6 + k] = rep.meanCosines[k]; e[10 + k] = rep.kappas[k]; e[14 + k] = rep.normalizations[k]; }
  }
}

}  // extern "C"
