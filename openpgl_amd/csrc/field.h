// Host side of the MI355X guiding field: configuration, the host copy of the
// kd tree and of the per-region float sample statistics (the bit-exact part of
// the spatial decisions), device buffers and the update pipeline.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <iosfwd>
#include <algorithm>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "query_math.h"
#include "region_sched.h"

namespace pglgpu {

#define PGL_HIP_CHECK(expr)                                                                          \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + \
                               ":" + std::to_string(__LINE__));                                      \
  } while (0)

// SampleStatistics (data/SampleStatistics.h:12-193), float, evaluated without contraction.
struct HostSampleStats {
  float mean[3] = {0.f, 0.f, 0.f};
  float var[3] = {0.f, 0.f, 0.f};
  float n = 0.f;
  float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
  float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
  void merge(const HostSampleStats& b);
  void split(int dim, float pos, float decay, bool splitLower);
  void variance(float v[3]) const;
};

// embree::rcp (emath.h:61-88) of the reference device of this SIMD width:
// 4 = SSE4 (rcpss, Newton step without FMA), 8 = AVX2 (rcpss, fused Newton
// step), 16 = AVX-512VL (rcp14, fused Newton step)
float host_rcp(float x, int vectorWidth);
bool host_supports_width(int vectorWidth);
HostSampleStats int_stats_to_float(const IntStatsAcc& a, const float center[3], const float half[3]);

struct KdSettings {
  uint32_t maxSamples = 32000;
  uint32_t maxDepth = 32;
  uint32_t minSamples = 100;
  bool knnLookup = true;
};

// Settings a field file carries that the training kernels do not read
// (WeightedEMFactory::Configuration maxK/maxKappa, WEM:33-42;
// AdaptiveSplitAndMergeFactory::Configuration maxSplitItr,
// AdaptiveSplitandMergeFactory.h:46; Field::m_fitRegions, Field.h:650).
struct FieldFileSettings {
  uint64_t maxK = 32;
  float maxKappa = 320000.f;
  int32_t maxSplitItr = 1;
  bool fitRegions = true;
};

// Growable device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b, bool keep = false);
  void release();
  ~DevBuf() { release(); }
  template <class T> T* as() const { return (T*)p; }
};

// Growable pinned host allocation (hipHostMalloc).  A coherent one is not
// cached by the GPU: kernels write their results into it directly (through
// `dev`, its device address), visible to the host once the kernel completes.
struct PinnedBuf {
  void* p = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
  bool coherent = false;
  void ensure(size_t b);
  ~PinnedBuf();
  template <class T> T* as() const { return (T*)p; }
};

struct UpdateInfo {
  double wallMs = 0, spatialMs = 0, directionalMs = 0, knnMs = 0, gatherMs = 0, sortMs = 0;
  double kdRounds = 0, regionLaunches = 0, estepEvals = 0, auxEvals = 0, emIterations = 0, samples = 0;
  double regionJobs = 0, regionItems = 0, maxTeam = 0, regionGrid = 0, splitEvals = 0, regionKernelMs = 0;
  double regions = 0;  // regions after the update
};

// One guiding field (surface or volume): Field<> of field/Field.h.
class KindField {
 public:
  KindField(const KdSettings& kd, const FactoryConfig& fc, bool deterministic, int vectorWidth);
  ~KindField();

  void setSceneBounds(const float lo[3], const float hi[3]);
  // queues an update (see field.cpp); deviceSamples must stay valid until the next call
  void update(const void* deviceSamples, uint32_t n, hipStream_t s);
  // update from host records (copied to the field's own staging buffer first)
  void updateHost(const void* hostSamples, uint32_t n, hipStream_t s) {
    updateFromHost(hostSamples, n);
    update(dAos_.p, n, s);
  }
  // waits for a queued update and collects its figures; throws if that
  // update failed, and from then on at every use until reset()
  void finish() const;
  // completes a queued update without throwing (release paths)
  void drain() const noexcept;
  bool failed() const { return failed_; }
  void reset();

  // FieldStatistics (field/FieldStatistics.h:16-30, Field.h:604-638):
  // region storage, the last update's times, KDTreeStatistics and
  // DirectionalDistributionStatistics
  struct Statistics {
    size_t numRegions = 0, regionsReserved = 0, sizePerRegion = 0;
    float timeUpdate = 0.f, timeCopySamples = 0.f, timeSpatial = 0.f, timeDirectional = 0.f;
    size_t kdMaxDepth = 0, kdNodes = 0, kdNodesReserved = 0, kdSizePerNode = 0;
    size_t sizePerDistribution = 0;
    float minK = 0.f, maxK = 0.f, averageK = 0.f, secondMomentK = 0.f;
  };
  Statistics statistics(hipStream_t s) const;

  // host mirror of the device state for scalar queries and exports
  void syncMirror(hipStream_t s) const;
  const RegionState& mirrorRegion(uint32_t r) const { return hostStates_[r]; }
  uint32_t numRegions() const { return (uint32_t)regions_.size(); }
  const std::vector<KDNode>& nodes() const { return nodes_; }
  bool initialized() const { return initialized_; }
  uint32_t iteration() const { return iteration_; }
  bool sceneBoundsSet() const { return boundsSet_; }
  void sceneBounds(float lo[3], float hi[3]) const;
  // Field::getRegion (field/Field.h:109-145) on the host mirror; -1 = none
  int lookupRegion(const float p[3], float* u, uint32_t& id) const;
  QueryField queryView() const;
  void canonicalOrder(std::vector<KDNode>& nodesOut, std::vector<uint32_t>& regionOrder) const;
  void exportRegion(uint32_t storageIdx, struct PGLRegionRecordHost* out) const;
  void knnHost(std::vector<uint32_t>& ids) const;
  const UpdateInfo& lastInfo() const { finish(); return info_; }
  const std::vector<UpdateInfo>& history() const { finish(); return history_; }
  RegionSchedule& schedule() { return runner().schedule; }
  RegionRunner& runner();
  bool validate() const;
  int vectorWidth() const { return vectorWidth_; }
  // false for the plain VMM field type (PGL_DIRECTIONAL_DISTRIBUTION_VMM)
  bool parallaxAware() const { return fc_.noParallax == 0; }
  // completes when the last queued update has read its input records (null before the first)
  hipEvent_t lastGatherEvent() const { return lastGatherParity_ < 0 ? nullptr : evSp_[lastGatherParity_][3]; }

  // Field::serialize / Field::deserialize (field/Field.h:240-305) in the
  // layout of a reference device of SIMD width W (4, 8 or 16); readFile
  // replaces the whole state and uploads it to the device.
  void writeFile(std::ostream& os, int W, bool isSurface, hipStream_t s) const;
  void readFile(std::istream& is, int W, bool parallaxAware, hipStream_t s);
  FieldFileSettings fileSettings;

  struct RegionHost {
    HostSampleStats ss;
    float lo[3], hi[3];
    bool splitFlag = false;
    uint32_t count = 0;       // samples of this update
    int32_t copyFrom = -1;    // device state to clone (new regions of this update)
  };

 private:
  void estimateSceneBounds(const float* px, const float* py, const float* pz, uint32_t n, hipStream_t s);
  void kdUpdate(uint32_t n, hipStream_t s);
  void visitLeaf(uint32_t node, uint32_t depth, const float lo[3], const float hi[3], uint32_t count,
                 const HostSampleStats& st, struct FrontierItemList& next);
  void ensureRegionCapacity(uint32_t r, hipStream_t s);
  void renumberDepthFirst(uint32_t nodesBefore, uint32_t regionsBefore);
  void updateFromHost(const void* hostSamples, uint32_t n);
  void reserve(uint32_t n, uint32_t regions);

  KdSettings kd_;
  FactoryConfig fc_;
  bool deterministic_;
  int vectorWidth_;

  std::vector<KDNode> nodes_;
  std::vector<RegionHost> regions_;
  std::vector<int32_t> splitRegion_;  // per node: region created when it split in this update (-1 = none)
  std::vector<KDNode> keyNodes_;      // node array of the push-down numbering with final region ids
  float sceneLo_[3], sceneHi_[3];
  bool boundsSet_ = false;
  float treeLo_[3], treeHi_[3];
  bool initialized_ = false;
  uint32_t iteration_ = 0;

  // device buffers
  DevBuf dPx_, dPy_, dPz_, dNodeOf_, dKeys_, dVals_, dKeys2_, dVals2_, dSortTmp_;
  DevBuf dSoa_[2][6];     // region-ordered dx, dy, dz, w, pdf, dist of two consecutive updates
  DevBuf dAos_;           // staging of host records
  DevBuf dRecs_;          // the update's records in region order (counting-sort path)
  DevBuf dTables_[2];          // per parity: leaf-key nodes | region starts | pivots | KNN pivots
  const float4* knnPivots_ = nullptr;  // KNN pivots of the pending update (in dTables_)
  hipStream_t sp_ = nullptr;  // spatial phase
  int parity_ = 0;
  int lastGatherParity_ = -1;
  hipEvent_t evSp_[2][4];
  hipEvent_t evDir_[2];      // end of the directional phase (region kernel and totals copy) on s
  hipEvent_t evKd_ = nullptr;  // the kd rounds' host waits
  hipEvent_t evKnn_[2][2];  // around the KNN build and the tree upload (spatial stream)
  // pinned staging of the per-update uploads, by frame parity: [0] inputs,
  // [1] pivots, [2] copy pairs, [3] KNN pivots, [4] key nodes, [5] nodes,
  // [6] region begins
  PinnedBuf pin_[2][7];
  DevBuf dCountSortTmp_;
  DevBuf dDetTmp_;           // deterministic-mode sort
  DevBuf dWkey_, dWkeySorted_;  // its weight keys in input and in region order
  PinnedBuf detPin_[2];      // its segment tables, by frame parity
  std::vector<uint32_t> hCount_;
  std::vector<uint32_t> hBegin_;
  PinnedBuf kdUp_, kdDown_;  // kd rounds (synchronous)
  template <class T> const T* pinned(PinnedBuf& b, const T* src, size_t count) {
    b.ensure(std::max<size_t>(count, 1) * sizeof(T));
    if (count) memcpy(b.p, src, count * sizeof(T));
    return b.as<T>();
  }
  std::vector<RegionInput> hInputs_;
  std::vector<float4> hPivots_, hKnnPiv_;
  std::vector<RegionWork> hWork_;
  uint32_t lastKdRounds_ = 0;
  mutable bool pending_ = false;
  mutable UpdateInfo pendingInfo_;
  mutable uint32_t pendingRegions_ = 0;
  mutable int pendingParity_ = 0;
  mutable int pendingSlot_ = 0;  // the region runner's result slot of the pending update
  std::chrono::steady_clock::time_point pendingT0_;
  mutable std::vector<UpdateInfo> history_;
  DevBuf dNodes_, dKdUp_, dPartial_, dStatPartial_;  // dKdUp_: the kd rounds' uploads
  std::unique_ptr<RegionRunner> runner_;
  mutable uint32_t hintRegions_ = 0;
  DevBuf dStates_, dNbIds_, dNbPts_, dUniformDirs_;
  uint32_t stateCapacity_ = 0;

  mutable std::vector<RegionState> hostStates_;
  mutable std::vector<uint32_t> hostNbIds_;
  mutable std::vector<float4> hostNbPts_;
  mutable bool mirrorDirty_ = true;
  mutable std::mutex mirrorMutex_;
  mutable UpdateInfo info_;
  // Field::m_timeLastUpdate, ...CopySamples, ...SpatialStructureUpdate,
  // ...DirectionalDistriubtionUpdate (ms): the last update's, or a loaded file's
  mutable float lastTimes_[4] = {0.f, 0.f, 0.f, 0.f};
  float copyMs_ = 0.f;                 // host-to-device copy of the next update's records
  mutable float pendingCopyMs_ = 0.f;  // of the queued update
  mutable bool failed_ = false;
  mutable std::string failMsg_;
  void finishPending() const;
};

// Region record layout of the EXTENSION export (mirrors PGLRegionRecord).
struct PGLRegionRecordHost {
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
};

// SurfaceVolumeField (field/SurfaceVolumeField.h): the surface and the volume field.
class GuidingField {
 public:
  GuidingField(const KdSettings& kd, const FactoryConfig& fc, bool deterministic, int vectorWidth);
  KindField surface, volume;
  uint32_t iteration = 0;
  hipStream_t stream = nullptr;
  DevBuf dHG;
  DevBuf dQuery;  // staging for the host-pointer batch API
};

const HGRep* host_hg_table();
// spherical-Fibonacci initial means (InitUniformVMM, WEM:632-666): 32 x 32 x 3 floats
void uniform_directions(std::vector<float>& out);
// field files in the layout of a reference device of SIMD width W
void store_field_file(const GuidingField& gf, const char* path, int W);
void load_field_file(GuidingField& gf, const char* path, int W);
// directional part of a region record
void export_state(const RegionState& S, PGLRegionRecordHost* out);
// minimum reprojection distance of prepareSamples (WEM:1184-1196): 9 sqrt(|variance|) with a fused dot
float reprojection_min_distance(const float var[3]);

}  // namespace pglgpu
