// Scheduling of the directional update: turns a list of (region, samples)
// into team jobs for the persistent region kernel and launches it.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace pglgpu {

struct DevBuf;

// Schedule parameters.  None of them changes the numerical result: the
// chunking and the fold order of the region kernel do not depend on them.
struct RegionSchedule {
  int perRegionLaunches = 0;  // 1: one launch of one workgroup per region (plain per-region path)
  uint32_t grid = 0;          // workgroups of the persistent launch (0: all that are co-resident)
  float balance = 1.f;        // team target = total samples / (grid * balance), at least minChunksPerRank chunks
  uint32_t minChunksPerRank = 1;
  uint32_t maxTeam = 64;
};

struct RegionSamplesView {
  const float *dx, *dy, *dz, *w, *pdf, *dist;
};

class RegionRunner {
 public:
  RegionRunner();
  ~RegionRunner();
  RegionRunner(const RegionRunner&) = delete;
  RegionRunner& operator=(const RegionRunner&) = delete;

  // work: (samples, region) with the region's samples at state.sampleBegin
  // of the sample arrays.  Jobs are ordered by samples, largest first.
  void launch(RegionState* states, const std::vector<std::pair<uint32_t, uint32_t>>& work,
              const RegionSamplesView& rs, const FactoryConfig& cfg, const float* uniformDirs, hipStream_t s);
  // after the stream has completed: throws if a team barrier timed out
  void check() const;

  RegionSchedule schedule;
  uint32_t residentGrid() const { return resident_; }
  // of the last launch
  uint32_t lastLaunches = 0, lastJobs = 0, lastItems = 0, lastMaxTeam = 0, lastGrid = 0;

 private:
  uint32_t resident_ = 0;
  bool cooperative_ = false;
  unsigned long long timeoutTicks_ = 0;
  void* dJobs_ = nullptr;
  size_t jobsBytes_ = 0;
  void* dItems_ = nullptr;
  size_t itemsBytes_ = 0;
  void* dQueue_ = nullptr;
  size_t queueBytes_ = 0;
  void* dPartial_ = nullptr;
  size_t partialBytes_ = 0;
  int* dErr_ = nullptr;
  int* hErr_ = nullptr;
  std::vector<RegionJob> jobs_;
  std::vector<uint2> items_;
};

}  // namespace pglgpu
