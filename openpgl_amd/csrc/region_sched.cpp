// Team planning and launch of the persistent region kernel.
#include "region_sched.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace pglgpu {

#define PGL_SCHED_CHECK(expr)                                                                             \
  do {                                                                                                    \
    const hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                                                   " in " #expr);                                         \
  } while (0)

static void grow_pinned(void*& p, size_t& have, size_t need) {
  if (need <= have) return;
  if (p) PGL_SCHED_CHECK(hipHostFree(p));
  p = nullptr;
  need = std::max<size_t>(std::max(need, have * 2), 65536);
  PGL_SCHED_CHECK(hipHostMalloc(&p, need, hipHostMallocDefault));
  have = need;
}

static void grow(void*& p, size_t& have, size_t need) {
  if (need <= have) return;
  if (p) PGL_SCHED_CHECK(hipFree(p));
  p = nullptr;
  need = std::max<size_t>(std::max(need, have * 2), 65536);
  PGL_SCHED_CHECK(hipMalloc(&p, need));
  have = need;
}

RegionRunner::RegionRunner() {
  int dev = 0, cus = 0, coop = 0;
  PGL_SCHED_CHECK(hipGetDevice(&dev));
  PGL_SCHED_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PGL_SCHED_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  int rateKHz = 0;
  PGL_SCHED_CHECK(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, dev));
  if (rateKHz <= 0) rateKHz = 100000;
  timeoutTicks_ = (unsigned long long)rateKHz * 1000ull * 2ull;  // 2 s
  const int perCU = region_kernel_blocks_per_cu();
  if (perCU <= 0 || cus <= 0) throw std::runtime_error("region kernel cannot be resident on this device");
  resident_ = (uint32_t)(perCU * cus);
  defaultGrid_ = (uint32_t)(std::min(perCU, 2) * cus);
  cooperative_ = coop != 0;
  PGL_SCHED_CHECK(hipMalloc((void**)&dTot_, sizeof(LaunchTotals)));
  PGL_SCHED_CHECK(hipHostMalloc((void**)&hTot_, 2 * sizeof(LaunchTotals), hipHostMallocDefault));
  dErr_ = &dTot_->err;
  dStats_ = &dTot_->stats;
  for (int k = 0; k < 2; k++) {
    hTot_[k].err = 0;
    hTot_[k].stats = RegionLaunchStats{0, 0, 0, 0};
    PGL_SCHED_CHECK(hipEventCreate(&evBegin_[k]));
    PGL_SCHED_CHECK(hipEventCreate(&evEnd_[k]));
  }
  PGL_SCHED_CHECK(hipEventCreateWithFlags(&evPrep_, hipEventDisableTiming));
}

RegionRunner::~RegionRunner() {
  (void)hipFree(dStage_);
  (void)hipHostFree(hStage_);
  (void)hipFree(dQueue_);
  (void)hipFree(dPartial_);
  (void)hipFree(dTot_);
  (void)hipHostFree(hTot_);
  for (int k = 0; k < 2; k++) {
    if (evBegin_[k]) (void)hipEventDestroy(evBegin_[k]);
    if (evEnd_[k]) (void)hipEventDestroy(evEnd_[k]);
  }
  if (evPrep_) (void)hipEventDestroy(evPrep_);
}

void RegionRunner::launch(RegionState* states, std::vector<RegionWork> work, const RegionSamplesView& rs,
                          const FactoryConfig& cfg, const float* uniformDirs, float* costHint, hipStream_t s,
                          const RegionInput* inputs, uint32_t numInputs) {
  stage(std::move(work), rs, cfg, uniformDirs, costHint, inputs, numInputs);
  submit(states, s);
}

void RegionRunner::stage(std::vector<RegionWork> work, const RegionSamplesView& rs, const FactoryConfig& cfg,
                         const float* uniformDirs, float* costHint, const RegionInput* inputs, uint32_t numInputs) {
  // the pinned staging is read by the previous launch's preparing kernel
  if (prepQueued_) PGL_SCHED_CHECK(hipEventSynchronize(evPrep_));
  prepQueued_ = false;
  Staged& st = staged_;
  st = Staged();
  st.valid = true;
  st.rs = rs;
  st.cfg = cfg;
  st.uniformDirs = uniformDirs;
  st.costHint = costHint;
  st.numInputs = inputs ? numInputs : 0;
  stagedMaxTeam_ = 0;
  if (work.empty()) {
    st.empty = true;
    if (inputs && numInputs) {  // the states still take their inputs
      if (submitted_ && size_t(numInputs) * sizeof(RegionInput) > hStageBytes_)
        PGL_SCHED_CHECK(hipEventSynchronize(evEnd_[cur_]));
      grow_pinned(hStage_, hStageBytes_, size_t(numInputs) * sizeof(RegionInput));
      memcpy(hStage_, inputs, size_t(numInputs) * sizeof(RegionInput));
    }
    return;
  }
  std::stable_sort(work.begin(), work.end(), [](const RegionWork& a, const RegionWork& b) { return a.cost > b.cost; });
  const uint32_t J = (uint32_t)work.size();
  // perRegionLaunches: 1 = one launch of one workgroup per region; 2 = one
  // launch per region with a team of one workgroup per chunk (up to maxTeam
  // and the resident grid), i.e. each region alone on the device
  const bool perRegion = schedule.perRegionLaunches != 0;
  const bool perRegionTeams = schedule.perRegionLaunches == 2;
  const uint32_t G = perRegion ? 1u : (schedule.grid ? std::min(schedule.grid, resident_) : defaultGrid_);
  double total = 0.0;
  for (const auto& w : work) total += std::max(w.cost, 0.f);
  const double bal = schedule.balance > 0.f ? schedule.balance : 1.f;
  const double target = std::max(total / ((double)G * bal), 1e-30);
  const uint32_t minChunks = std::max<uint32_t>(1, schedule.minChunksPerRank);
  const uint32_t teamCap = std::max<uint32_t>(1, std::min(schedule.maxTeam, perRegionTeams ? resident_ : G));

  jobs_.resize(J);
  items_.clear();
  unsigned long long partialFloats = 0;
  for (uint32_t j = 0; j < J; j++) {
    const uint32_t n = work[j].samples;
    RegionJob& jb = jobs_[j];
    jb.region = work[j].region;
    jb.chunks = std::max<uint32_t>(1, (n + kChunk - 1) / kChunk);
    uint32_t T = 1;
    if (!perRegion || perRegionTeams) {
      const double want = perRegionTeams ? (double)jb.chunks : std::ceil(std::max(work[j].cost, 0.f) / target);
      T = (uint32_t)std::min<double>(std::max(want, 1.0), (double)std::max<uint32_t>(1, jb.chunks / minChunks));
      T = std::max<uint32_t>(1, std::min(T, teamCap));
      // fewest ranks with the same largest number of chunks per rank
      const uint32_t per = (jb.chunks + T - 1) / T;
      T = (jb.chunks + per - 1) / per;
    }
    jb.team = T;
    jb.pad = 0;
    jb.partialOffset = T > 1 ? partialFloats : 0ull;
    if (T > 1) partialFloats += 2ull * jb.chunks * kPartStride;
    stagedMaxTeam_ = std::max(stagedMaxTeam_, T);
    if (!perRegion || perRegionTeams)
      for (uint32_t r = 0; r < T; r++) items_.push_back(make_uint2(j, r));
  }
  if (perRegion && !perRegionTeams) items_.assign(1, make_uint2(0, 0));

  // jobs | items | inputs (16-byte aligned parts) in pinned memory, read by the preparing kernel
  const size_t jobsB = J * sizeof(RegionJob), itemsB = items_.size() * sizeof(uint2);
  const size_t inB = inputs ? size_t(numInputs) * sizeof(RegionInput) : 0;
  const size_t offItems = (jobsB + 15) & ~size_t(15), offIn = (offItems + itemsB + 15) & ~size_t(15);
  const size_t stageB = offIn + inB;
  // a buffer that grows is freed: first wait until the last launch (and with
  // it every earlier one on its stream), which may still be queued or
  // running and reading it, has completed
  const size_t partB = std::max<unsigned long long>(partialFloats, 1) * 4;
  if (submitted_ && (offIn > stageBytes_ || size_t(J + 1) * 4 > queueBytes_ || partB > partialBytes_ ||
                     stageB > hStageBytes_))
    PGL_SCHED_CHECK(hipEventSynchronize(evEnd_[cur_]));
  grow(dStage_, stageBytes_, offIn);
  grow(dQueue_, queueBytes_, (size_t)(J + 1) * 4);
  grow(dPartial_, partialBytes_, partB);
  grow_pinned(hStage_, hStageBytes_, stageB);
  char* hs = (char*)hStage_;
  memcpy(hs, jobs_.data(), jobsB);
  memcpy(hs + offItems, items_.data(), itemsB);
  if (inB) memcpy(hs + offIn, inputs, inB);
  st.J = J;
  st.G = G;
  st.offItems = offItems;
  st.offIn = offIn;
  st.inB = inB;
  st.perRegion = perRegion;
  st.perRegionTeams = perRegionTeams;
}

void RegionRunner::submit(RegionState* states, hipStream_t s) {
  if (!staged_.valid) throw std::runtime_error("region runner: nothing staged");
  Staged& st = staged_;
  st.valid = false;
  lastLaunches = lastJobs = lastItems = lastGrid = 0;
  lastMaxTeam = stagedMaxTeam_;
  // the result slot of this launch (the other one holds the previous launch's)
  const int k = 1 - cur_;
  cur_ = k;
  hTot_[k].err = 0;
  hTot_[k].stats = RegionLaunchStats{0, 0, 0, 0};
  timed_[k] = false;
  if (hStage_ != hStageMapped_) {
    PGL_SCHED_CHECK(hipHostGetDevicePointer(&hStageDev_, hStage_, 0));
    hStageMapped_ = hStage_;
  }
  char* hdev = (char*)hStageDev_;
  if (st.empty) {
    if (st.numInputs) {
      launch_prepare_regions(states, nullptr, nullptr, 0, (const RegionInput*)hdev, st.numInputs, nullptr, 0, dErr_,
                             (unsigned long long*)dStats_, s);
      PGL_SCHED_CHECK(hipEventRecord(evPrep_, s));
      prepQueued_ = true;
    }
    // marks the end of everything queued so far (the buffers' last use)
    PGL_SCHED_CHECK(hipEventRecord(evEnd_[k], s));
    submitted_ = true;
    return;
  }
  const uint32_t J = st.J;
  char* ds = (char*)dStage_;
  launch_prepare_regions(states, hdev, ds, st.offIn, st.inB ? (const RegionInput*)(hdev + st.offIn) : nullptr,
                         st.inB ? st.numInputs : 0, (uint32_t*)dQueue_, J + 1, dErr_, (unsigned long long*)dStats_, s);
  PGL_SCHED_CHECK(hipEventRecord(evPrep_, s));
  prepQueued_ = true;

  const FactoryConfig& cfg = st.cfg;
  RegionLaunch L;
  L.states = states;
  L.jobs = (const RegionJob*)ds;
  L.numJobs = J;
  L.items = (const uint2*)(ds + st.offItems);
  L.numItems = (uint32_t)items_.size();
  L.grid = std::min<uint32_t>(st.G, L.numItems);
  L.queueAndCounters = (uint32_t*)dQueue_;
  L.partial = (float*)dPartial_;
  L.err = dErr_;
  L.timeoutTicks = timeoutTicks_;
  L.stats = (unsigned long long*)dStats_;
  L.costHint = st.costHint;
  L.dx = st.rs.dx; L.dy = st.rs.dy; L.dz = st.rs.dz; L.w = st.rs.w; L.pdf = st.rs.pdf; L.dist = st.rs.dist;
  L.uniformDirs = st.uniformDirs;
  L.cooperative = cooperative_ && schedule.cooperative != 0;
  L.zeroQueue = st.perRegion;
  PGL_SCHED_CHECK(hipEventRecord(evBegin_[k], s));
  if (st.perRegionTeams) {
    // one launch per job: its consecutive items, the queue and counters zeroed before each
    L.cooperative = false;
    L.zeroQueue = true;
    const uint2* allItems = L.items;
    uint32_t first = 0;
    for (uint32_t j = 0; j < J; j++) {
      const uint32_t T = jobs_[j].team;
      L.items = allItems + first;
      L.numItems = T;
      L.grid = T;
      launch_region_update(L, cfg, s);
      first += T;
      lastGrid = std::max(lastGrid, T);
    }
    lastLaunches = J;
  } else if (st.perRegion) {
    L.cooperative = false;
    L.numJobs = 1;
    L.numItems = 1;
    L.grid = 1;
    for (uint32_t j = 0; j < J; j++) {
      L.jobs = (const RegionJob*)ds + j;
      launch_region_update(L, cfg, s);
    }
    lastLaunches = J;
    lastGrid = 1;
  } else {
    launch_region_update(L, cfg, s);
    lastLaunches = 1;
    lastGrid = L.grid;
  }
  PGL_SCHED_CHECK(hipEventRecord(evEnd_[k], s));
  submitted_ = true;
  timed_[k] = true;
  lastJobs = J;
  lastItems = (uint32_t)items_.size();
  // the error flag and the totals are adjacent: one copy back
  PGL_SCHED_CHECK(hipMemcpyAsync(&hTot_[k], dTot_, sizeof(LaunchTotals), hipMemcpyDeviceToHost, s));
}

float RegionRunner::kernelMs(int slot) const {
  if (!timed_[slot]) return 0.f;
  float ms = 0.f;
  PGL_SCHED_CHECK(hipEventElapsedTime(&ms, evBegin_[slot], evEnd_[slot]));
  return ms;
}

void RegionRunner::check(int slot) const {
  if (failNext) {
    failNext = false;
    throw std::runtime_error("directional update: failure injected for testing");
  }
  if (hTot_[slot].err != 0)
    throw std::runtime_error("directional update: a region team barrier timed out (workgroups not co-resident)");
}

}  // namespace pglgpu
