/*
 * Open PGL-compatible C ABI of the MI355X guiding-field library
 * (libopenpgl_amd.so).  Every function keeps the name, argument meaning and
 * error behaviour of the Open PGL 0.6.0 C API it replaces; each declaration
 * cites the reference header line (paths relative to the reference tree,
 * openpgl/include/openpgl/).  Functions marked EXTENSION have no counterpart
 * in the reference.
 *
 * As in the reference, a few signatures take C++ references; from C they are
 * the same pointer at the ABI level.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
#define PGL_REF(T) T&
extern "C" {
#else
#define PGL_REF(T) T*
#endif

#define PGL_VMM_MAX_COMPONENTS 32      /* common.h:26 */
#define PGL_VMM_MAX_KAPPA 320000.f     /* common.h:27 */
#define PGL_TREE_MAX_SAMPLE_PER_LEAF 32000 /* config.h:25 */

/* common.h:29-47 */
typedef struct { float x, y, z; } pgl_vec3f;
typedef pgl_vec3f pgl_point3f;
typedef struct { float x, y; } pgl_vec2f;
typedef pgl_vec2f pgl_point2f;
typedef struct { pgl_vec3f lower, upper; } pgl_box3f;

/* data.h:21-49: 40-byte training record; flags bit 1 (EInsideVolume) routes to the volume field */
typedef struct PGLSampleData {
  pgl_point3f position;
  pgl_vec3f direction;
  float weight;
  float pdf;
  float distance;
  uint32_t flags;
} PGLSampleData;
#define PGL_SAMPLE_FLAG_SPLATTED 1u
#define PGL_SAMPLE_FLAG_INSIDE_VOLUME 2u

/* data.h:50-113: one recorded path vertex (136 bytes).  A default-constructed
 * vertex (pglPathSegmentStorageNextSegment) has directionIn (0,1,0),
 * directionOut (1,0,0), normal (0,0,1), pdfDirectionIn, miWeight,
 * russianRouletteProbability, eta, roughness 1, unit weights and zero
 * contributions. */
typedef struct PGLPathSegmentData {
  pgl_point3f position;
  pgl_vec3f directionIn;
  pgl_vec3f directionOut;
  pgl_vec3f normal;
  bool volumeScatter;
  float pdfDirectionIn;
  bool isDelta;
  pgl_vec3f scatteringWeight;
  pgl_vec3f transmittanceWeight;
  pgl_vec3f directContribution;
  float miWeight;
  pgl_vec3f scatteredContribution;
  float russianRouletteProbability;
  float eta;
  float roughness;
  const void* regionPtr;
} PGLPathSegmentData;

/* data.h:106-109 */
typedef struct PGLString { char* m_str; size_t m_size; } PGLString;

/* types.h */
enum PGL_SPATIAL_STRUCTURE_TYPE { PGL_SPATIAL_STRUCTURE_KDTREE = 0 };
enum PGL_DIRECTIONAL_DISTRIBUTION_TYPE {
  PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM = 0,
  PGL_DIRECTIONAL_DISTRIBUTION_QUADTREE,
  PGL_DIRECTIONAL_DISTRIBUTION_VMM
};

/* config.h:25-31.  The CPU_* types select the reference device whose SIMD
 * width (4/8/16) defines the merge floor; GPU_MI355X behaves as CPU_8.
 * EXTENSION: PGL_DEVICE_TYPE_GPU_MI355X is appended without renumbering. */
enum PGL_DEVICE_TYPE {
  PGL_DEVICE_TYPE_CPU_4,
  PGL_DEVICE_TYPE_CPU_8,
  PGL_DEVICE_TYPE_CPU_16,
  PGL_DEVICE_TYPE_NONE,
  PGL_DEVICE_TYPE_GPU_MI355X
};

/* config.h:33-39 */
typedef struct PGLKDTreeArguments {
  bool knnLookup;
  size_t minSamples;
  size_t maxSamples;
  size_t maxDepth;
} PGLKDTreeArguments;

/* config.h:41-81 */
typedef struct PGLVMMFactoryArguments {
  size_t initK;
  float initKappa;
  size_t maxK;
  size_t maxEMIterrations;
  float maxKappa;
  float convergenceThreshold;
  float weightPrior;
  float meanCosinePriorStrength;
  float meanCosinePrior;
  bool useSplitAndMerge;
  float splittingThreshold;
  float mergingThreshold;
  bool partialReFit;
  int maxSplitItr;
  int minSamplesForSplitting;
  int minSamplesForPartialRefitting;
  int minSamplesForMerging;
} PGLVMMFactoryArguments;

/* config.h:99-118 */
typedef struct PGLDebugArguments { bool fitRegions; } PGLDebugArguments;
typedef struct PGLFieldArguments {
  enum PGL_SPATIAL_STRUCTURE_TYPE spatialStructureType;
  void* spatialSturctureArguments;
  enum PGL_DIRECTIONAL_DISTRIBUTION_TYPE directionalDistributionType;
  void* directionalDistributionArguments;
  bool deterministic;
  PGLDebugArguments debugArguments;
} PGLFieldArguments;

typedef struct PGLDeviceImpl* PGLDevice;
typedef struct PGLFieldImpl* PGLField;
typedef struct PGLSampleStorageImpl* PGLSampleStorage;
typedef struct PGLSurfaceSamplingDistributionImpl* PGLSurfaceSamplingDistribution;
typedef struct PGLVolumeSamplingDistributionImpl* PGLVolumeSamplingDistribution;
typedef struct PGLRegionImpl* PGLRegion;
typedef struct PGLFieldStatisticsImpl* PGLFieldStatistics;
typedef struct PGLPathSegmentStorageImpl* PGLPathSegmentStorage;

/* config.h:120, api.cpp:755-787: heap-allocates the two argument structs;
 * pglDeviceNewField takes ownership and deletes them (Device.h:109,136) */
void pglFieldArgumentsSetDefaults(PGL_REF(PGLFieldArguments) fieldArguments,
                                  enum PGL_SPATIAL_STRUCTURE_TYPE spatialType,
                                  enum PGL_DIRECTIONAL_DISTRIBUTION_TYPE directionalType, bool deterministic,
                                  size_t maxSamplesPerLeaf);
void pglReleaseString(PGLString str);                                            /* data.h:115 */

/* device.h:22-31 */
PGLDevice pglNewDevice(enum PGL_DEVICE_TYPE deviceType, size_t numThreads);
PGLField pglDeviceNewField(PGLDevice device, PGLFieldArguments args);
PGLField pglDeviceNewFieldFromFile(PGLDevice device, const char* fieldFileName);
void pglReleaseDevice(PGLDevice device);

/* field.h:36-70 */
void pglReleaseField(PGLField field);
bool pglFieldStoreToFile(PGLField field, const char* fieldFileName);
size_t pglFieldGetIteration(PGLField field);
void pglFieldSetSceneBounds(PGLField field, pgl_box3f bounds);
pgl_box3f pglFieldGetSceneBounds(PGLField field);
void pglFieldUpdate(PGLField field, PGLSampleStorage sampleStorage);
void pglFieldUpdateSurface(PGLField field, PGLSampleStorage sampleStorage);
void pglFieldUpdateVolume(PGLField field, PGLSampleStorage sampleStorage);
void pglFieldReset(PGLField field);
PGLSurfaceSamplingDistribution pglFieldNewSurfaceSamplingDistribution(PGLField field);
bool pglFieldInitSurfaceSamplingDistribution(PGLField field, PGLSurfaceSamplingDistribution surfaceSamplingDistribution,
                                             pgl_point3f position, float* sample1D);
PGLVolumeSamplingDistribution pglFieldNewVolumeSamplingDistribution(PGLField field);
bool pglFieldInitVolumeSamplingDistribution(PGLField field, PGLVolumeSamplingDistribution volumeSamplingDistribution,
                                            pgl_point3f position, float* sample1D);
bool pglFieldValidate(PGLField field);
bool pglFieldCompare(PGLField fieldA, PGLField fieldB);
PGLFieldStatistics pglFieldGetSurfaceStatistics(PGLField field);
PGLFieldStatistics pglFieldGetVolumeStatistics(PGLField field);

/* fieldstatistics.h:21-27 */
void pglReleaseFieldStatistics(PGLFieldStatistics fieldStatistics);
PGLString pglFieldStatisticsToString(PGLFieldStatistics fieldStatistics);
PGLString pglFieldStatisticsHeaderCSVString(PGLFieldStatistics fieldStatistics);
PGLString pglFieldStatisticsToCSVString(PGLFieldStatistics fieldStatistics);
/* EXTENSION: the FieldStatistics figures as numbers (field/FieldStatistics.h:16-30):
 * [0] numCacheRegions [1] numCacheRegionsReserved [2] sizePerCacheRegions
 * [3] timeLastUpdate [4] timeLastUpdateCopySamples [5] ...SpatialStructureUpdate
 * [6] ...DirectionalDistriubtionUpdate (ms) [7] kd maxDepth [8] numberOfNodes
 * [9] numberOfReservedNodes [10] sizePerNode [11] sizePerDistribution
 * [12] min [13] max [14] average [15] second moment (root mean square) of the
 * component counts; count <= 16 values are written. */
void pglFieldStatisticsGetValues(PGLFieldStatistics fieldStatistics, double* out, size_t count);

/* samplestorage.h:35-65 */
PGLSampleStorage pglNewSampleStorage(void);
PGLSampleStorage pglNewSampleStorageFromFile(const char* sampleStorageFileName);
void pglReleaseSampleStorage(PGLSampleStorage sampleStorage);
bool pglSampleStorageStoreToFile(PGLSampleStorage sampleStorage, const char* sampleStorageFileName);
void pglSampleStorageAddSample(PGLSampleStorage sampleStorage, PGL_REF(PGLSampleData) sample);
void pglSampleStorageAddSamples(PGLSampleStorage sampleStorage, const PGLSampleData* samples, size_t numSamples);
void pglSampleStorageReserve(PGLSampleStorage sampleStorage, const size_t sizeSurface, const size_t sizeVolume);
void pglSampleStorageClear(PGLSampleStorage sampleStorage);
void pglSampleStorageClearSurface(PGLSampleStorage sampleStorage);
void pglSampleStorageClearVolume(PGLSampleStorage sampleStorage);
size_t pglSampleStorageGetSizeSurface(PGLSampleStorage sampleStorage);
size_t pglSampleStorageGetSizeVolume(PGLSampleStorage sampleStorage);
PGLSampleData pglSampleStorageGetSampleSurface(PGLSampleStorage sampleStorage, const int idx);
PGLSampleData pglSampleStorageGetSampleVolume(PGLSampleStorage sampleStorage, const int idx);
bool pglSampleStorageValidate(PGLSampleStorage sampleStorage);
bool pglSampleStorageCompare(PGLSampleStorage sampleStorageA, PGLSampleStorage sampleStorageB);

/* pathsegmentstorage.h:35-68 (the C++ default arguments of the reference are
 * useNEEMiWeights = false, guideDirectLight = false,
 * rrAffectsDirectContribution = true).  Per-path storage on the host, as in
 * the reference; vertices and samples beyond the reserved capacity are
 * dropped. */
PGLPathSegmentStorage pglNewPathSegmentStorage(void);
void pglReleasePathSegmentStorage(PGLPathSegmentStorage pathSegmentStorage);
void pglPathSegmentStorageReserve(PGLPathSegmentStorage pathSegmentStorage, size_t size);
void pglPathSegmentStorageClear(PGLPathSegmentStorage pathSegmentStorage);
void pglPathSegmentSetMaxDistance(PGLPathSegmentStorage pathSegmentStorage, float maxDistance);
float pglPathSegmentGetMaxDistance(PGLPathSegmentStorage pathSegmentStorage);
int pglPathSegmentGetNumSegments(PGLPathSegmentStorage pathSegmentStorage);
int pglPathSegmentGetNumSamples(PGLPathSegmentStorage pathSegmentStorage);
size_t pglPathSegmentStoragePrepareSamples(PGLPathSegmentStorage pathSegmentStorage, const bool useNEEMiWeights,
                                           const bool guideDirectLight, const bool rrAffectsDirectContribution);
pgl_vec3f pglPathSegmentStorageCalculatePixelEstimate(PGLPathSegmentStorage pathSegmentStorage,
                                                      const bool rrAffectsDirectContribution);
const PGLSampleData* pglPathSegmentStorageGetSamples(PGLPathSegmentStorage pathSegmentStorage,
                                                     PGL_REF(size_t) nSamples);
void pglPathSegmentStorageAddSample(PGLPathSegmentStorage pathSegmentStorage, PGLSampleData sample);
PGLPathSegmentData* pglPathSegmentStorageNextSegment(PGLPathSegmentStorage pathSegmentStorage);
void pglPathSegmentStorageAddSegment(PGLPathSegmentStorage pathSegmentStorage, PGLPathSegmentData segment);
bool pglPathSegmentStorageValidateSamples(PGLPathSegmentStorage pathSegmentStorage);
bool pglPathSegmentStorageValidateSegments(PGLPathSegmentStorage pathSegmentStorage);
void pglPathSegmentStoragePropagateSamples(PGLPathSegmentStorage pathSegmentStorage, PGLSampleStorage sampleStorage,
                                           bool guideDirectLight, bool useNEEMiWeights,
                                           bool rrAffectsDirectContribution);

/* surfacesamplingdistribution.h:24-46 */
void pglReleaseSurfaceSamplingDistribution(PGLSurfaceSamplingDistribution surfaceSamplingDistribution);
void pglSurfaceSamplingDistributionApplyCosineProduct(PGLSurfaceSamplingDistribution surfaceSamplingDistribution,
                                                      pgl_vec3f normal);
bool pglSurfaceSamplingDistributionSupportsApplyCosineProduct(PGLSurfaceSamplingDistribution surfaceSamplingDistribution);
pgl_vec3f pglSurfaceSamplingDistributionSample(PGLSurfaceSamplingDistribution surfaceSamplingDistribution,
                                               pgl_point2f sample);
float pglSurfaceSamplingDistributionPDF(PGLSurfaceSamplingDistribution surfaceSamplingDistribution, pgl_vec3f direction);
float pglSurfaceSamplingDistributionSamplePDF(PGLSurfaceSamplingDistribution surfaceSamplingDistribution,
                                              pgl_point2f sample, PGL_REF(pgl_vec3f) direction);
float pglSurfaceSamplingDistributionIncomingRadiancePDF(PGLSurfaceSamplingDistribution surfaceSamplingDistribution,
                                                        pgl_vec3f direction);
uint32_t pglSurfaceSamplingDistributionGetId(PGLSurfaceSamplingDistribution surfaceSamplingDistribution);
bool pglSurfaceSamplingDistributionValidate(PGLSurfaceSamplingDistribution surfaceSamplingDistribution);
void pglSurfaceSamplingDistributionClear(PGLSurfaceSamplingDistribution surfaceSamplingDistribution);
PGLRegion pglSurfaceSamplingGetRegion(PGLSurfaceSamplingDistribution surfaceSamplingDistribution);

/* volumesamplingdistribution.h:24-46 */
void pglReleaseVolumeSamplingDistribution(PGLVolumeSamplingDistribution volumeSamplingDistribution);
pgl_vec3f pglVolumeSamplingDistributionSample(PGLVolumeSamplingDistribution volumeSamplingDistribution,
                                              pgl_point2f sample);
float pglVolumeSamplingDistributionPDF(PGLVolumeSamplingDistribution volumeSamplingDistribution, pgl_vec3f direction);
float pglVolumeSamplingDistributionSamplePDF(PGLVolumeSamplingDistribution volumeSamplingDistribution,
                                             pgl_point2f sample, PGL_REF(pgl_vec3f) direction);
float pglVolumeSamplingDistributionIncomingRadiancePDF(PGLVolumeSamplingDistribution volumeSamplingDistribution,
                                                       pgl_vec3f direction);
uint32_t pglVolumeSamplingDistributionGetId(PGLVolumeSamplingDistribution volumeSamplingDistribution);
bool pglVolumeSamplingDistributionValidate(PGLVolumeSamplingDistribution volumeSamplingDistribution);
void pglVolumeSamplingDistributionClear(PGLVolumeSamplingDistribution volumeSamplingDistribution);
void pglVolumeSamplingDistributionApplySingleLobeHenyeyGreensteinProduct(
    PGLVolumeSamplingDistribution volumeSamplingDistribution, pgl_vec3f dir, const float meanCosine);
bool pglVolumeSamplingDistributionSupportsApplySingleLobeHenyeyGreensteinProduct(
    PGLVolumeSamplingDistribution volumeSamplingDistribution);
PGLRegion pglVolumeSamplingGetRegion(PGLVolumeSamplingDistribution volumeSamplingDistribution);

/* ------------------------------------------------------------------------
 * EXTENSIONS (no counterpart in the reference)
 * ---------------------------------------------------------------------- */

/* EXTENSION: Field::Update of one field kind from records already resident in
 * device memory (same semantics as pglFieldUpdateSurface/Volume; no flag
 * routing).  Used by GPU-side producers and by the benchmark.  Updates are
 * queued: the call returns once the training work is queued on the device;
 * the next update, any query or read of the field, and pglReleaseField wait
 * for it.  The device records must stay unchanged until then. */
void pglFieldUpdateSurfaceDevice(PGLField field, const PGLSampleData* deviceSamples, size_t numSamples);
void pglFieldUpdateVolumeDevice(PGLField field, const PGLSampleData* deviceSamples, size_t numSamples);
/* EXTENSION: pglFieldUpdate (api/field.h, Field::updateField) with mixed
 * surface and volume records in device memory: they are split on the
 * device by PGL_SAMPLE_FLAG_INSIDE_VOLUME, each part in input order, as a
 * sample storage routes them (SampleDataStorage.h:72-90), then the surface
 * and the volume field are updated (an empty part leaves its field as it
 * is).  Returns once the records have been split (they may be reused then);
 * the training work is queued as for pglFieldUpdateSurfaceDevice. */
void pglFieldUpdateDevice(PGLField field, const PGLSampleData* deviceSamples, size_t numSamples);

/* EXTENSION: batched Init + product + Sample + PDF for n queries on the GPU.
 * All pointers are device pointers: pos/normalOrDir (3n floats), g (n, HG mean
 * cosine, volume queries only; may be NULL), u2 (2n), u1 (n; < 0 disables the
 * KNN pick; may be NULL = all disabled), isVolume (n bytes; NULL = all
 * surface).  Outputs dirOut (3n), pdfOut (n; 0 for a failed Init), idOut (n,
 * kd leaf id or 0xffffffff; may be NULL), integralOut (n, product integral; may
 * be NULL).  stream = hipStream_t (NULL = default stream); asynchronous. */
void pglFieldSampleBatchDevice(PGLField field, size_t n, const float* pos, const float* normalOrDir, const float* g,
                               const float* u2, const float* u1, const uint8_t* isVolume, float* dirOut,
                               float* pdfOut, uint32_t* idOut, float* integralOut, void* stream);
/* EXTENSION: host-pointer variant of pglFieldSampleBatchDevice (synchronous). */
void pglFieldSampleBatch(PGLField field, size_t n, const float* pos, const float* normalOrDir, const float* g,
                         const float* u2, const float* u1, const uint8_t* isVolume, float* dirOut, float* pdfOut,
                         uint32_t* idOut, float* integralOut);

/* EXTENSION: flat region record used to compare fields against a checker
 * (canonical depth-first region order of KDTree::rearrangeNodesForCompare). */
typedef struct PGLRegionRecord {
  uint32_t K;
  float weights[32], kappas[32], meanDirections[32][3], normalizations[32], eMinus2Kappa[32], meanCosines[32],
      distances[32];
  float pivot[3];
  float sumOfWeightedDirections[32][3], sumOfWeightedStats[32], sumOfDistanceWeights[32];
  float sumWeights, numSamples, overallNumSamples;
  uint32_t suffK, normalized;
  float chiSquareMCEstimates[32], splitMeans[32][2], splitCovariances[32][3], splitNumSamples[32],
      splitSumWeights[32], splitSumAssignedSamples[32];
  uint32_t splitK;
  uint64_t numSamplesAfterLastSplit, numSamplesAfterLastMerge;
  float ssMean[3], ssVariance[3], ssNumSamples, ssLower[3], ssUpper[3];
  float regionLower[3], regionUpper[3];
  uint32_t splitFlag, valid;
  uint32_t numSamplesThisUpdate;
  uint32_t emIterations, partialEmIterations, numSplits, numMerges;
} PGLRegionRecord;

/* EXTENSION: introspection (kind: 0 = surface field, 1 = volume field) */
uint32_t pglFieldGetNumNodes(PGLField field, int kind);
uint32_t pglFieldGetNumRegions(PGLField field, int kind);
void pglFieldExportNodes(PGLField field, int kind, float* pivots, uint32_t* dims, uint32_t* childOrRegion);
void pglFieldExportRegion(PGLField field, int kind, uint32_t canonicalIdx, PGLRegionRecord* out);
void pglFieldSampleLeafIds(PGLField field, int kind, const float* positions, size_t n, uint32_t* canonicalIds);
void pglFieldGetKNN(PGLField field, int kind, uint32_t* neighbours /* 8 per region, canonical ids */);
/* EXTENSION: figures of the last update of one kind (times in milliseconds):
 * [0] total wall, [1] spatial (GPU), [2] directional region kernel (GPU),
 * [3] KNN (GPU), [4] gather+reproject (GPU), [5] sort (GPU), [6] kd rounds,
 * [7] region kernel launches, [8] E-step sample x component evaluations,
 * [9] distance-pass evaluations, [10] EM iterations (sum), [11] samples,
 * [12] split-statistics evaluations, [13] regions, [14] region jobs,
 * [15] queue items (sum of team sizes), [16] largest team, [17] grid,
 * [18] region kernel GPU time (events around its launch only).
 * Writes min(count, 19) values. */
void pglFieldGetLastUpdateInfo(PGLField field, int kind, double* out, size_t count);
/* EXTENSION (testing): the next update of the field kind (0 surface, 1 volume)
 * is reported as failed once its kernels have run, as a region kernel whose
 * team barrier timed out is.  A failed update is reported by the call that
 * completes it; from then on every update, query, export and file write of
 * the field fails (prints the error, returns false / 0 / nothing) until
 * pglFieldReset. */
void pglFieldDebugFailNextUpdate(PGLField field, int kind);
/* EXTENSION: the same figures for the last min(maxEntries, 256) updates of
 * one kind, oldest first, 19 doubles each ([13] is 0); returns the count.
 * Updates are queued (pglFieldUpdate* returns once the directional phase is
 * queued, the next update or any read of the field waits for it), so this
 * is how a caller gets per-update figures without waiting after each one. */
size_t pglFieldGetUpdateHistory(PGLField field, int kind, double* out, size_t maxEntries);

/* EXTENSION: scheduling of the directional update.  None of the fields
 * changes the result (bit-identical for every schedule).
 *   perRegionLaunches  1: one launch of one workgroup per region (the plain
 *                      per-region GPU path); 2: one launch per region with a
 *                      team of one workgroup per 1024-sample chunk (up to
 *                      maxTeam and the device), each region alone on the
 *                      device; 0: one persistent launch
 *   grid               workgroups of the persistent launch (0: two per CU,
 *                      which leaves room for other streams' kernels; at
 *                      most the workgroups that fit at once, three per CU)
 *   balance            a team gets one workgroup per
 *                      total samples / (grid * balance) of its region
 *   minChunksPerRank   at least this many 1024-sample chunks per workgroup
 *   maxTeam            largest team
 *   cooperative        1: cooperative launch (the runtime guarantees that the
 *                      whole grid is resident and keeps other kernels off
 *                      the device meanwhile); 0 (default): ordinary launch
 *                      (the grid fits the device, so every workgroup becomes
 *                      resident once other streams' kernels end; waits stay
 *                      bounded; other streams' kernels share the CUs that
 *                      finished workgroups free) */
typedef struct {
  int perRegionLaunches;
  uint32_t grid;
  float balance;
  uint32_t minChunksPerRank;
  uint32_t maxTeam;
  int cooperative;
} PGLRegionSchedule;
void pglFieldSetRegionSchedule(PGLField field, const PGLRegionSchedule* schedule);
/* EXTENSION: per-region profile of the last update, 12 floats per region in
 * storage order: microseconds of [0] the whole region, [1] E-step passes,
 * [2] split-statistics passes, [3] distance passes, [4] team barrier waits,
 * [5] merging, [6] splitting logic, [9] M-steps and convergence tests,
 * [10] cross-chunk folds; [7] team size, [8] E-step passes, [11] samples. */
void pglFieldGetRegionProfile(PGLField field, int kind, float* out);
void pglFieldGetRegionSchedule(PGLField field, PGLRegionSchedule* schedule);

/* ------------------------------------------------------------------------
 * EXTENSION: batched directional training of pre-partitioned regions.
 * The Field::fitRegions / updateRegions step (field/Field.h:389-511) for a
 * caller-given set of regions, without a spatial structure, in one persistent
 * GPU launch.  Arguments as for pglDeviceNewField (the sub-structs are
 * deleted; only the directional arguments are used).
 *   pglRegionBatchStep: deviceRecords = region-contiguous PGLSampleData in
 *   device memory (region r starts at the sum of counts[0..r-1]);
 *   sampleStats7 = per region the float sample statistics of the positions
 *   (mean[3], variance * n[3], n), host memory; initK (optional) = initial
 *   components per region (0 = the arguments' initK); splitFlags
 *   (optional). The first step fits, later steps update.
 *   pglRegionBatchAutotune re-runs the last step's launch from its input
 *   state for each candidate schedule (NULL = the built-in set from
 *   pglRegionScheduleCandidates, where grid == 1 means one workgroup per CU
 *   and grid == 2 all that fit at once),
 *   keeps the fastest of reps runs per candidate, selects the fastest
 *   candidate and returns its index (-1 on error); msOut receives the kernel
 *   time of every candidate.  Schedules never change results.
 * ------------------------------------------------------------------------ */
typedef struct PGLRegionBatchImpl* PGLRegionBatch;
PGLRegionBatch pglNewRegionBatch(PGLDevice device, PGLFieldArguments args);
void pglReleaseRegionBatch(PGLRegionBatch batch);
void pglRegionBatchStep(PGLRegionBatch batch, uint32_t numRegions, const PGLSampleData* deviceRecords,
                        const uint32_t* counts, const float* sampleStats7, const uint32_t* initK,
                        const int* splitFlags);
uint32_t pglRegionBatchGetNumRegions(PGLRegionBatch batch);
void pglRegionBatchExportRegion(PGLRegionBatch batch, uint32_t region, PGLRegionRecord* out);
void pglRegionBatchSetSchedule(PGLRegionBatch batch, const PGLRegionSchedule* schedule);
void pglRegionBatchGetSchedule(PGLRegionBatch batch, PGLRegionSchedule* schedule);
size_t pglRegionScheduleCandidates(PGLRegionSchedule* out, size_t maxCount);
int pglRegionBatchAutotune(PGLRegionBatch batch, const PGLRegionSchedule* candidates, size_t numCandidates, int reps,
                           float* msOut);
/* same layout as pglFieldGetLastUpdateInfo ([1], [3]-[6] are 0) */
void pglRegionBatchGetLastInfo(PGLRegionBatch batch, double* out, size_t count);

/* ------------------------------------------------------------------------
 * EXTENSION: continuous-batching query server.
 * A request is n queries (the batched Init + cosine/HG product + sample +
 * pdf of pglFieldSampleBatch) with host pointers that stay valid until the
 * request is done; pinned host memory (hipHostMalloc) gives asynchronous
 * copies.  The device-memory budget is split into one arena per in-flight
 * slot (68 bytes per query), allocated at creation.  Requests are queued as
 * pieces of at most 32768 queries; whenever a slot is free, the pending
 * pieces are admitted in arrival order while its arena has room and run as
 * one batch, so batches on different slots overlap on the device and large
 * requests pipeline across slots.  Each query's result depends only on its own inputs,
 * so a request's results are the same alone or in any mix.  The server
 * serves a snapshot of the field taken at creation and by pglServerRefresh
 * (call it from the thread that updates the field, after the update), so the
 * field may be updated while requests are served.
 *   pglServerSubmit returns the request id (0 if rejected: larger than the
 *   budget); pglServerWait returns 1 when done, 0 on
 *   timeout (timeoutMs < 0 waits forever); pglServerGetTiming: milliseconds
 *   since server creation of submit, admission (batch issue) and completion
 *   (results in host memory); pglServerGetStats: [0] requests done,
 *   [1] queries done, [2] batches, [3] peak in-flight bytes, [4] peak pending
 *   requests, [5] largest batch.
 * ------------------------------------------------------------------------ */
typedef struct {
  size_t memoryBudgetBytes; /* 0: 512 MiB */
  uint32_t slots;           /* 0: 4 */
  uint32_t maxBatchQueries; /* 0: 2^20 */
} PGLServerConfig;
typedef struct PGLServerImpl* PGLServer;
PGLServer pglNewServer(PGLField field, const PGLServerConfig* config);
void pglReleaseServer(PGLServer server);
uint64_t pglServerSubmit(PGLServer server, size_t n, const float* pos, const float* normalOrDir, const float* g,
                         const float* u2, const float* u1, const uint8_t* isVolume, float* dirOut, float* pdfOut,
                         uint32_t* idOut, float* integralOut);
int pglServerPoll(PGLServer server, uint64_t request);
int pglServerWait(PGLServer server, uint64_t request, double timeoutMs);
void pglServerDrain(PGLServer server);
int pglServerGetTiming(PGLServer server, uint64_t request, double* out3);
void pglServerForget(PGLServer server, uint64_t request);
void pglServerRefresh(PGLServer server);
void pglServerGetStats(PGLServer server, double* out, size_t count);

/* ---------------------------------------------------------------------------
 * EXTENSION: path segments of many paths on the GPU.  Path p is the vertex
 * range [pathOffsets[p], pathOffsets[p+1]) of the numSegments vertices
 * `segments` (device memory, the PGLPathSegmentData layout; ranges must be
 * non-decreasing and within numSegments, else the call fails); its training samples are those of
 * pglPathSegmentStoragePrepareSamples on that path with the given max
 * distance (pglPathSegmentSetMaxDistance, 1e6 by default) and flags.
 *
 * pglPathSegmentBatchPrepareSamplesDevice writes them to device arrays,
 * surface and volume samples apart, in the order per-path
 * pglPathSegmentStoragePropagateSamples calls in path order would append
 * them to one sample storage; returns false (with a message) when a
 * capacity is too small, writing nothing.  Blocks until written.
 *
 * pglFieldUpdateFromPathSegmentsDevice trains the field with those samples
 * without a host round trip: pglFieldUpdate semantics (surface and volume
 * field, one iteration).  The path arrays may be reused once it returns.
 * ------------------------------------------------------------------------ */
typedef struct PGLPathSegmentBatchImpl* PGLPathSegmentBatch;
PGLPathSegmentBatch pglNewPathSegmentBatch(void);
void pglReleasePathSegmentBatch(PGLPathSegmentBatch batch);
bool pglPathSegmentBatchPrepareSamplesDevice(PGLPathSegmentBatch batch, const PGLPathSegmentData* segments,
                                             size_t numSegments, const uint32_t* pathOffsets, uint32_t numPaths,
                                             float maxDistance,
                                             bool useNEEMiWeights, bool guideDirectLight,
                                             bool rrAffectsDirectContribution, PGLSampleData* surfaceOut,
                                             size_t surfaceCapacity, PGLSampleData* volumeOut, size_t volumeCapacity,
                                             size_t* numSurfaceOut, size_t* numVolumeOut);
void pglFieldUpdateFromPathSegmentsDevice(PGLField field, const PGLPathSegmentData* segments, size_t numSegments,
                                          const uint32_t* pathOffsets, uint32_t numPaths, float maxDistance,
                                          bool useNEEMiWeights, bool guideDirectLight,
                                          bool rrAffectsDirectContribution);

/* EXTENSION, no counterpart in the reference: the samples of
 * pglPathSegmentBatchPrepareSamplesDevice (same arguments, results and
 * errors) without allocating anything, for callers that cannot allocate
 * device or pinned memory (render or serving loops held to a fixed memory
 * footprint, hosts that forbid allocation inside the frame).  Its scratch is one static
 * device array of 15.25 MiB per GPU, held until the process ends; at most
 * 262144 vertices (numSegments) per call, larger calls fail with a message
 * and write nothing.  Work is queued on `stream` (hipStream_t, NULL = the
 * default stream) and the call blocks until the records are written.  All
 * calls on one GPU share that array: calls that overlap on one GPU (from
 * several threads) race and must be serialised by the caller. */
bool pglPathSegmentsPrepareSamplesDeviceNoAlloc(const PGLPathSegmentData* segments, size_t numSegments,
                                                const uint32_t* pathOffsets, uint32_t numPaths, float maxDistance,
                                                bool useNEEMiWeights, bool guideDirectLight,
                                                bool rrAffectsDirectContribution, PGLSampleData* surfaceOut,
                                                size_t surfaceCapacity, PGLSampleData* volumeOut,
                                                size_t volumeCapacity, size_t* numSurfaceOut, size_t* numVolumeOut,
                                                void* stream);

#ifdef __cplusplus
}  /* extern "C" */
#endif
