// Batched directional training of many pre-partitioned regions (no spatial
// structure): the Field::fitRegions / updateRegions step (field/Field.h:389-511)
// for a caller-given set of regions, in one persistent launch.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "field.h"
#include "region_sched.h"

namespace pglgpu {

class RegionBatch {
 public:
  // deterministic: every region's samples are first sorted by SampleDataLess
  // (Field.h:409-412, 469-472)
  RegionBatch(const FactoryConfig& fc, bool deterministic);
  ~RegionBatch();

  // One step for numRegions regions: the first call fits (ASM::fit), later
  // calls update (ASM::update).  records: device PGLSampleData, region r's
  // records at [begin_r, begin_r + counts[r]) with begin the running sum of
  // counts.  stats7: per region the float sample statistics of the positions
  // (mean[3], variance * n[3], n) as the kd build keeps them.  initK: per
  // region components of the initial mixture (nullptr or 0: the factory's
  // initK).  splitFlags (optional): decay the training statistics first.
  void step(uint32_t numRegions, const void* records, const uint32_t* counts, const float* stats7,
            const uint32_t* initK, const int* splitFlags, hipStream_t s);

  // Re-runs the directional launch of the last step from its input state for
  // each candidate schedule (reps times, keeping the fastest run of each),
  // selects the fastest and returns its index.  The result state equals the
  // step's (the schedule does not change results).  msOut[i]: kernel time.
  int autotune(const std::vector<RegionSchedule>& candidates, int reps, std::vector<float>& msOut, hipStream_t s);

  uint32_t numRegions() const { return numRegions_; }
  void exportRegion(uint32_t r, PGLRegionRecordHost* out, hipStream_t s);
  RegionRunner& runner() { return runner_; }
  const UpdateInfo& lastInfo() const { return info_; }

 private:
  FactoryConfig fc_;
  bool deterministic_;
  RegionRunner runner_;
  uint32_t numRegions_ = 0;
  uint32_t steps_ = 0;
  uint32_t lastSamples_ = 0;
  std::vector<RegionWork> lastWork_;
  DevBuf dStates_, dSnapshot_, dInputs_, dPivots_, dBegin_, dKeys_, dVals_;
  DevBuf dDx_, dDy_, dDz_, dW_, dPdf_, dDist_, dUniformDirs_;
  DevBuf dDetTmp_;
  PinnedBuf detPin_;
  std::vector<RegionState> hostStates_;
  UpdateInfo info_;
  hipEvent_t ev_[2];
};

}  // namespace pglgpu
