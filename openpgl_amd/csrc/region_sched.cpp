// Team planning and launch of the persistent region kernel.
#include "region_sched.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace pglgpu {

#define PGL_SCHED_CHECK(expr)                                                                             \
  do {                                                                                                    \
    const hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                                                   " in " #expr);                                         \
  } while (0)

static void grow(void*& p, size_t& have, size_t need) {
  if (need <= have) return;
  if (p) PGL_SCHED_CHECK(hipFree(p));
  p = nullptr;
  need = std::max<size_t>(need, 256);
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
  cooperative_ = coop != 0;
  PGL_SCHED_CHECK(hipMalloc((void**)&dErr_, sizeof(int)));
  PGL_SCHED_CHECK(hipHostMalloc((void**)&hErr_, sizeof(int), hipHostMallocDefault));
  *hErr_ = 0;
}

RegionRunner::~RegionRunner() {
  (void)hipFree(dJobs_);
  (void)hipFree(dItems_);
  (void)hipFree(dQueue_);
  (void)hipFree(dPartial_);
  (void)hipFree(dErr_);
  (void)hipHostFree(hErr_);
}

void RegionRunner::launch(RegionState* states, const std::vector<std::pair<uint32_t, uint32_t>>& work,
                          const RegionSamplesView& rs, const FactoryConfig& cfg, const float* uniformDirs,
                          hipStream_t s) {
  lastLaunches = lastJobs = lastItems = lastMaxTeam = lastGrid = 0;
  *hErr_ = 0;
  if (work.empty()) return;
  const uint32_t J = (uint32_t)work.size();
  const bool perRegion = schedule.perRegionLaunches != 0;
  const uint32_t G = perRegion ? 1u : (schedule.grid ? std::min(schedule.grid, resident_) : resident_);
  uint64_t total = 0;
  for (const auto& w : work) total += w.first;
  const double bal = schedule.balance > 0.f ? schedule.balance : 1.f;
  const uint64_t minTarget = (uint64_t)kChunk * std::max<uint32_t>(1, schedule.minChunksPerRank);
  const uint64_t target = std::max<uint64_t>(minTarget, (uint64_t)((double)total / ((double)G * bal)) + 1);
  const uint32_t teamCap = std::max<uint32_t>(1, std::min(schedule.maxTeam, G));

  jobs_.resize(J);
  items_.clear();
  unsigned long long partialFloats = 0;
  for (uint32_t j = 0; j < J; j++) {
    const uint32_t n = work[j].first;
    RegionJob& jb = jobs_[j];
    jb.region = work[j].second;
    jb.chunks = std::max<uint32_t>(1, (n + kChunk - 1) / kChunk);
    uint32_t T = 1;
    if (!perRegion) {
      T = (uint32_t)std::min<uint64_t>((n + target - 1) / target, jb.chunks);
      T = std::max<uint32_t>(1, std::min(T, teamCap));
      // fewest ranks with the same largest number of chunks per rank
      const uint32_t per = (jb.chunks + T - 1) / T;
      T = (jb.chunks + per - 1) / per;
    }
    jb.team = T;
    jb.pad = 0;
    jb.partialOffset = T > 1 ? partialFloats : 0ull;
    if (T > 1) partialFloats += 2ull * jb.chunks * kPartStride;
    lastMaxTeam = std::max(lastMaxTeam, T);
    if (!perRegion)
      for (uint32_t r = 0; r < T; r++) items_.push_back(make_uint2(j, r));
  }
  if (perRegion) items_.assign(1, make_uint2(0, 0));

  grow(dJobs_, jobsBytes_, J * sizeof(RegionJob));
  grow(dItems_, itemsBytes_, items_.size() * sizeof(uint2));
  grow(dQueue_, queueBytes_, (size_t)(J + 1) * 4);
  grow(dPartial_, partialBytes_, std::max<unsigned long long>(partialFloats, 1) * 4);
  PGL_SCHED_CHECK(hipMemcpyAsync(dJobs_, jobs_.data(), J * sizeof(RegionJob), hipMemcpyHostToDevice, s));
  PGL_SCHED_CHECK(hipMemcpyAsync(dItems_, items_.data(), items_.size() * sizeof(uint2), hipMemcpyHostToDevice, s));
  PGL_SCHED_CHECK(hipMemsetAsync(dErr_, 0, sizeof(int), s));

  const RegionJob* dj = (const RegionJob*)dJobs_;
  const uint2* di = (const uint2*)dItems_;
  if (perRegion) {
    for (uint32_t j = 0; j < J; j++)
      launch_region_update(states, dj + j, 1, di, 1, 1, (uint32_t*)dQueue_, (float*)dPartial_, dErr_, timeoutTicks_, rs.dx, rs.dy,
                           rs.dz, rs.w, rs.pdf, rs.dist, cfg, uniformDirs, false, s);
    lastLaunches = J;
    lastGrid = 1;
  } else {
    const uint32_t grid = std::min<uint32_t>(G, (uint32_t)items_.size());
    launch_region_update(states, dj, J, di, (uint32_t)items_.size(), grid, (uint32_t*)dQueue_, (float*)dPartial_,
                         dErr_, timeoutTicks_, rs.dx, rs.dy, rs.dz, rs.w, rs.pdf, rs.dist, cfg, uniformDirs, cooperative_, s);
    lastLaunches = 1;
    lastGrid = grid;
  }
  lastJobs = J;
  lastItems = (uint32_t)items_.size();
  PGL_SCHED_CHECK(hipMemcpyAsync(hErr_, dErr_, sizeof(int), hipMemcpyDeviceToHost, s));
}

void RegionRunner::check() const {
  if (*hErr_ != 0)
    throw std::runtime_error("directional update: a region team barrier timed out (workgroups not co-resident)");
}

}  // namespace pglgpu
