// ORACLE — test infrastructure only.
//
// CPU restatement of Open PGL 0.6.0's guiding-field training path
// (Field::Update: kd-tree build/update + parallax-aware vMF weighted EM with
// adaptive split-and-merge) and of its query path (region lookup, KNN pick,
// parallax shift, cosine / Henyey-Greenstein products, sample, pdf).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this code, and only as the checker (or the reported CPU baseline).  The
// product library (openpgl_amd/) never includes, links or calls it.
//
// Numerics:
//   * kd-tree, integer sample statistics, scene bounds and KNN are restated
//     with the reference's exact IEEE float semantics (no contraction; the
//     host's rcpss+Newton reciprocal as in embree::rcp under AVX2), so tree,
//     leaf ids and region sample statistics are bit-exact.
//   * the directional part is templated on the real type: R=float follows
//     the reference's FP32 algorithm sequentially (sample order = the
//     deterministic SampleDataLess sort), R=double is the FP64 twin used to
//     state the error of both the reference-order FP32 and the GPU results.
//     Transcendentals use libm (the reference uses ispc fastapprox
//     polynomials for exp/log/acos/sin; those differ by ~1e-7 relative).
//
// Parity pins for each function are listed in the header of pgl_oracle.cpp.
#pragma once
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float px, py, pz;
  float dx, dy, dz;
  float weight, pdf, distance;
  uint32_t flags;
} orc_sample;  // same 40-byte layout as PGLSampleData (include/openpgl/data.h:21-49)

typedef struct {
  // spatial (PGLKDTreeArguments, config.h:33-39)
  uint32_t knnLookup, maxSamples, maxDepth;
  // directional (PGLVMMFactoryArguments, config.h:41-81)
  uint32_t initK; float initKappa; uint32_t maxK; uint32_t maxEMIterations;
  float maxKappa; float convergenceThreshold; float weightPrior;
  float meanCosinePriorStrength; float meanCosinePrior;
  uint32_t useSplitAndMerge; float splittingThreshold; float mergingThreshold;
  uint32_t partialReFit; int32_t minSamplesForSplitting;
  int32_t minSamplesForPartialRefitting; int32_t minSamplesForMerging;
  // field (Device.h:100-104)
  float decayOnSpatialSplit; uint32_t deterministic;
  // merge loop stops at K <= vectorWidth (VMMChiSquareComponentMerger.h:72)
  uint32_t vectorWidth;
  // 1: parallax-aware mixture (PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM);
  // 0: plain VMM (PGL_DIRECTIONAL_DISTRIBUTION_VMM, Device.h:139-184): no
  // reprojection (WEM:1186), no shift before updates (Field.h:476) or queries
  // (VMMSurfaceSamplingDistribution.h:41, VMMVolumeSamplingDistribution.h:43)
  uint32_t parallax;
} orc_config;

// one recorded path vertex, same 136-byte layout as PGLPathSegmentData
// (include/openpgl/data.h:50-113)
typedef struct {
  float position[3], directionIn[3], directionOut[3], normal[3];
  uint8_t volumeScatter; uint8_t pad0[3];
  float pdfDirectionIn;
  uint8_t isDelta; uint8_t pad1[3];
  float scatteringWeight[3], transmittanceWeight[3], directContribution[3];
  float miWeight;
  float scatteredContribution[3];
  float russianRouletteProbability, eta, roughness;
  uint32_t pad2;
  uint64_t regionPtr;
} orc_path_segment;

// flat per-region export record used by tests (canonical DFS order)
typedef struct {
  // VMM
  uint32_t K;
  float weights[32], kappas[32], meanDirections[32][3], normalizations[32],
      eMinus2Kappa[32], meanCosines[32], distances[32];
  float pivot[3];
  // sufficient statistics
  float sumOfWeightedDirections[32][3], sumOfWeightedStats[32], sumOfDistanceWeights[32];
  float sumWeights, numSamples, overallNumSamples; uint32_t suffK; uint32_t normalized;
  // split statistics
  float chiSquareMCEstimates[32], splitMeans[32][2], splitCovariances[32][3],
      splitNumSamples[32], splitSumWeights[32], splitSumAssignedSamples[32];
  uint32_t splitK;
  uint64_t numSamplesAfterLastSplit, numSamplesAfterLastMerge;
  // sample statistics (float, bit-exact vs reference)
  float ssMean[3], ssVariance[3], ssNumSamples, ssLower[3], ssUpper[3];
  float regionLower[3], regionUpper[3];
  uint32_t splitFlag, valid;
  uint32_t numSamplesThisUpdate;  // samples assigned in the last update
  uint32_t emIterations, partialEmIterations, numSplits, numMerges;
} orc_region_record;

typedef struct orc_field_s* orc_field;

void orc_config_defaults(orc_config* cfg, uint32_t maxSamplesPerLeaf);

// precision: 0 = FP32 (reference order), 1 = FP64 twin
orc_field orc_field_new(const orc_config* cfg, int precision);
void orc_field_release(orc_field f);
void orc_field_set_scene_bounds(orc_field f, const float* lower3, const float* upper3);
void orc_field_get_scene_bounds(orc_field f, float* lower3, float* upper3);
// updates the surface or volume field with n samples (routing by flag is the
// caller's job, as in SampleDataStorage); numThreads parallelises the
// per-region directional fits
void orc_field_update(orc_field f, const orc_sample* samples, size_t n, int numThreads);
uint32_t orc_field_iteration(orc_field f);
// canonical (DFS) export: nodes as (pivot, dim(0-2) or 3 = leaf, child-or-region)
uint32_t orc_field_num_nodes(orc_field f);
uint32_t orc_field_num_regions(orc_field f);
void orc_field_export_nodes(orc_field f, float* pivots, uint32_t* dims, uint32_t* childOrData);
void orc_field_export_region(orc_field f, uint32_t canonicalIdx, orc_region_record* out);
// storage index -> canonical index map (size num_regions)
void orc_field_region_canonical_map(orc_field f, uint32_t* storageToCanonical);
void orc_field_sample_leaf_ids(orc_field f, const float* pos, size_t n, uint32_t* canonicalIds);
// KNN neighbour lists (canonical region ids, 8 per region, ~0u = empty)
void orc_field_knn(orc_field f, uint32_t* neighbours);

// query (surface if volume == 0): init + product + sample + pdf.  g is used
// only for volume queries.  u1 < 0 disables KNN.  Returns 1 if valid.
int orc_field_query(orc_field f, int volume, const float* pos, const float* normalOrDir, float g,
                    const float* u2, float u1, float* dirOut, float* pdfOut, uint32_t* idOut,
                    float* productIntegral);
// batched form of orc_field_query (outputs pdf = 0, id = ~0 for invalid)
void orc_field_query_batch(orc_field f, int volume, size_t n, const float* pos, const float* nrm,
                           const float* g, const float* u2, const float* u1, float* dirOut,
                           float* pdfOut, uint32_t* idOut, float* integralOut, int numThreads);
// pdf of a given direction for an initialised query (used for PDF compare)
int orc_field_query_pdf(orc_field f, int volume, const float* pos, const float* normalOrDir, float g,
                        float u1, const float* dir, float* pdfOut);
// the scalar call sequence of a renderer: Init (u1 updated in place by the
// KNN draw), `products` cosine (surface) or HG (volume) products,
// SamplePDF(u2), PDF and IncomingRadiancePDF at evalDir.  out7: sampled
// dir[3], its pdf, pdf(evalDir), pdfLi(evalDir), product integral.
int orc_field_scalar_sequence(orc_field f, int volume, const float* pos, float* u1, const float* normalOrDir,
                              float g, int products, const float* u2, const float* evalDir, float* out7,
                              uint32_t* id);
// pdf of one query (Init without KNN + one product) at n directions; 0 if Init fails
int orc_field_query_pdf_dirs(orc_field f, int volume, const float* pos, const float* normalOrDir, float g,
                             size_t n, const float* dirs, float* pdfOut);

// single-region directional KATs (no spatial structure): fit on the first
// call, update afterwards, exactly as Field::fitRegions/updateRegions do for
// one region whose sample statistics are given.
typedef struct orc_region_s* orc_region;
orc_region orc_region_new(const orc_config* cfg, int precision);
void orc_region_release(orc_region r);
void orc_region_step(orc_region r, const orc_sample* samples, size_t n, const float* ssMean,
                     const float* ssVariance, float ssNumSamples, int splitFlag);
void orc_region_export(orc_region r, orc_region_record* out);
float orc_region_pdf(orc_region r, const float* dir);

// init at pos + cosine (volume=0) or HG (volume=1) product on a single region,
// pdf at pdfDir, one sample (dirOut, pdfOut); returns the product integral
float orc_region_query(orc_region r, int volume, const float* pos, const float* nrm, float g, const float* u2,
                       const float* pdfDir, float* pdfAtDir, float* dirOut, float* pdfOut);
// SampleStatistics::merge(a, b) and split of a at (dim, pos) into lower/upper halves (13 floats each)
void orc_stats_merge_split(const float* a13, const float* b13, int dim, float pos, float* merged13, float* left13,
                           float* right13);

// integer sample statistics of a point set with given bounds, converted to
// SampleStatistics (KDTreeBuilder semantics).  out: mean3 var3 n lower3 upper3
void orc_integer_stats(const float* pos, size_t n, const float* lower3, const float* upper3,
                       float* out13, int64_t* rawOut13);
float orc_rcp(float x);
// embree::rcp of the device of SIMD width W (4: SSE4, 8: AVX2, 16: AVX-512VL)
float orc_rcp_width(float x, int W);
// integer statistics (raw int64 sums, counts, min/max) with the reciprocal of width W
void orc_integer_stats_width(const float* pos, size_t n, const float* lower3, const float* upper3, int W,
                             int64_t* rawOut13);
// KNN pick on one neighbourhood: ids8 (~0u = empty slot), pts24 = x[8] y[8] z[8],
// size = neighbours found, p = query position, u updated in place; fused:
// squared distances with FMA (CPU_8 / CPU_16) or without (CPU_4)
uint32_t orc_knn_pick(const uint32_t* ids8, const float* pts24, uint32_t size, const float* p, float* u, int fused);

// HG phase-function representation lookup (fills K and 4x{w, meanCosine, kappa, norm})
int orc_hg_rep(float g, float* w, float* mc, float* kappa, float* norm);
void orc_set_hg_table(const float* table);  // 128 x (K, g, 4w, 4mc, 4kappa, 4norm) = 18 floats

// PathSegmentDataStorage::prepareSamples of one path (n vertices): writes at
// most cap samples, returns how many the path yields
size_t orc_path_prepare_samples(const orc_path_segment* segs, size_t n, float maxDistance, int useNEEMiWeights,
                                int guideDirectLight, int rrAffectsDirectContribution, orc_sample* out, size_t cap);
// prepareSamples of numPaths paths [offsets[p], offsets[p+1]) in order, one
// thread; returns the total (at most cap written to out)
size_t orc_path_prepare_batch(const orc_path_segment* segs, const uint32_t* offsets, size_t numPaths,
                              float maxDistance, int useNEEMiWeights, int guideDirectLight,
                              int rrAffectsDirectContribution, orc_sample* out, size_t cap);
// PathSegmentDataStorage::calculatePixelEstimate of one path
void orc_path_pixel_estimate(const orc_path_segment* segs, size_t n, int rrAffectsDirectContribution, float* rgb);

#ifdef __cplusplus
}
#endif
