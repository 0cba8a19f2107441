// Scheduling of the directional update: turns a list of (region, samples)
// into team jobs for the persistent region kernel and launches it.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"
#include "kernels.h"

namespace pglgpu {

struct DevBuf;

// Schedule parameters.  None of them changes the numerical result: the
// chunking and the fold order of the region kernel do not depend on them.
struct RegionSchedule {
  int perRegionLaunches = 0;  // 1: one launch of one workgroup per region (plain per-region path)
  uint32_t grid = 0;          // workgroups of the persistent launch (0: two per CU; at most all that fit at once)
  float balance = 1.f;        // team target = total samples / (grid * balance), at least minChunksPerRank chunks
  uint32_t minChunksPerRank = 1;
  uint32_t maxTeam = 64;
  // 1: hipLaunchCooperativeKernel (the runtime makes the whole grid resident
  // before it starts and keeps other streams' kernels off the device
  // meanwhile); 0: ordinary launch.  The grid never exceeds what is
  // resident at once and every other kernel ends, so every workgroup is
  // eventually resident and every team wait ends; the waits stay bounded in
  // any case.  The ordinary launch lets the next update's spatial kernels use
  // the CUs that finished workgroups free.
  int cooperative = 0;
};

struct RegionSamplesView {
  const float *dx, *dy, *dz, *w, *pdf, *dist;
};

// One region of a launch: its samples start at state.sampleBegin; cost is the
// predicted work (any unit, e.g. samples x evaluations per sample).
struct RegionWork {
  uint32_t region;
  uint32_t samples;
  float cost;
};

// Totals of a launch (read after the stream has completed).
struct RegionLaunchStats {
  unsigned long long estepEvals, distEvals, splitEvals, emIterations;
};

class RegionRunner {
 public:
  RegionRunner();
  ~RegionRunner();
  RegionRunner(const RegionRunner&) = delete;
  RegionRunner& operator=(const RegionRunner&) = delete;

  // Jobs are ordered by cost, largest first; team sizes follow the cost.
  // costHint (optional, device, per region) receives the evaluations per
  // sample of this launch.  inputs (optional, host, numInputs entries) are
  // applied to the states before the launch (launch_prepare_regions), also
  // when there is no work.  The previous launch must have completed.
  void launch(RegionState* states, std::vector<RegionWork> work, const RegionSamplesView& rs,
              const FactoryConfig& cfg, const float* uniformDirs, float* costHint, hipStream_t s,
              const RegionInput* inputs = nullptr, uint32_t numInputs = 0);
  // launch() in two steps: stage() does the host work (jobs, teams, the
  // pinned tables; it waits only for the previous launch's preparing kernel,
  // which reads them), so it may run while the previous launch still runs;
  // submit() queues the preparing kernel and the launch on s, behind the
  // previous launch if that is still queued or running.  Each launch keeps
  // its figures (check, stats, kernelMs) in one of two result slots:
  // lastSlot() after submit() names it, and they stay valid until the
  // second submit() after it.
  void stage(std::vector<RegionWork> work, const RegionSamplesView& rs, const FactoryConfig& cfg,
             const float* uniformDirs, float* costHint, const RegionInput* inputs = nullptr,
             uint32_t numInputs = 0);
  void submit(RegionState* states, hipStream_t s);
  int lastSlot() const { return cur_; }
  // after the launch has completed: throws if a team barrier timed out
  void check(int slot) const;
  void check() const { check(cur_); }
  // testing: the next check() reports a failed launch (the kernels run normally)
  mutable bool failNext = false;
  const RegionLaunchStats& stats(int slot) const { return hTot_[slot].stats; }
  const RegionLaunchStats& stats() const { return stats(cur_); }
  // GPU time of the region kernel(s) of a launch (HIP events around the
  // kernel launches only), valid after the launch has completed
  float kernelMs(int slot) const;
  float kernelMs() const { return kernelMs(cur_); }

  RegionSchedule schedule;
  uint32_t residentGrid() const { return resident_; }
  // of the last launch
  uint32_t lastLaunches = 0, lastJobs = 0, lastItems = 0, lastMaxTeam = 0, lastGrid = 0;

 private:
  uint32_t resident_ = 0;     // workgroups that fit on the device at once
  uint32_t defaultGrid_ = 0;  // two per CU (one more fits per CU, for other streams' kernels)
  bool cooperative_ = false;
  unsigned long long timeoutTicks_ = 0;
  void* dStage_ = nullptr;  // jobs | items | inputs of the current launch
  size_t stageBytes_ = 0;
  void* dQueue_ = nullptr;
  size_t queueBytes_ = 0;
  void* dPartial_ = nullptr;
  size_t partialBytes_ = 0;
  // the error flag and the totals of a launch, adjacent (one copy back into
  // the launch's result slot)
  struct LaunchTotals {
    int err;
    int pad;
    RegionLaunchStats stats;
  };
  LaunchTotals* dTot_ = nullptr;  // zeroed by each launch's preparing kernel
  LaunchTotals* hTot_ = nullptr;  // [2] result slots (pinned)
  int* dErr_ = nullptr;
  RegionLaunchStats* dStats_ = nullptr;
  hipEvent_t evBegin_[2] = {nullptr, nullptr}, evEnd_[2] = {nullptr, nullptr};
  bool timed_[2] = {false, false};
  int cur_ = 1;  // slot of the last submit()
  hipEvent_t evPrep_ = nullptr;  // after the last preparing kernel (which reads the pinned tables)
  bool prepQueued_ = false;
  bool submitted_ = false;  // evEnd_[cur_] has been recorded
  // the staged launch: what submit() needs besides the tables
  struct Staged {
    bool valid = false, empty = false;
    RegionSamplesView rs{};
    FactoryConfig cfg{};
    const float* uniformDirs = nullptr;
    float* costHint = nullptr;
    uint32_t numInputs = 0, J = 0, G = 0;
    size_t offItems = 0, offIn = 0, inB = 0;
    bool perRegion = false, perRegionTeams = false;
  } staged_;
  uint32_t stagedMaxTeam_ = 0;
  void* hStageMapped_ = nullptr;  // hStage_ whose device address is hStageDev_
  void* hStageDev_ = nullptr;
  std::vector<RegionJob> jobs_;
  std::vector<uint2> items_;
  // pinned staging of the upload (reused once the previous launch has completed)
  void* hStage_ = nullptr;
  size_t hStageBytes_ = 0;
};

}  // namespace pglgpu
