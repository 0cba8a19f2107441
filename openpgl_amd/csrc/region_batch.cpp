// Batched directional training of pre-partitioned regions.
#include "region_batch.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>

namespace pglgpu {

RegionBatch::RegionBatch(const FactoryConfig& fc, bool deterministic) : fc_(fc), deterministic_(deterministic) {
  std::vector<float> ud;
  uniform_directions(ud);
  dUniformDirs_.ensure(ud.size() * 4);
  PGL_HIP_CHECK(hipMemcpy(dUniformDirs_.p, ud.data(), ud.size() * 4, hipMemcpyHostToDevice));
  for (auto& e : ev_) PGL_HIP_CHECK(hipEventCreate(&e));
}

RegionBatch::~RegionBatch() {
  for (auto& e : ev_) (void)hipEventDestroy(e);
}

void RegionBatch::step(uint32_t R, const void* records, const uint32_t* counts, const float* stats7,
                       const uint32_t* initK, const int* splitFlags, hipStream_t s) {
  if (R == 0) return;
  if (steps_ > 0 && R != numRegions_)
    throw std::runtime_error("region batch: the number of regions changed between steps");
  if (!records || !counts || !stats7) throw std::runtime_error("region batch: null input");
  const auto t0 = std::chrono::steady_clock::now();
  const bool first = steps_ == 0;
  if (first) {
    numRegions_ = R;
    dStates_.ensure(size_t(R) * sizeof(RegionState));
    dSnapshot_.ensure(size_t(R) * sizeof(RegionState));
    PGL_HIP_CHECK(hipMemsetAsync(dStates_.p, 0, size_t(R) * sizeof(RegionState), s));
  }
  std::vector<RegionInput> inputs(R);
  std::vector<float4> pivots(R);
  std::vector<uint32_t> begin(R);
  std::vector<RegionWork> work;
  work.reserve(R);
  uint64_t total = 0;
  for (uint32_t r = 0; r < R; r++) {
    begin[r] = (uint32_t)total;
    RegionInput& in = inputs[r];
    in.region = r;
    HostSampleStats ss;
    for (int j = 0; j < 3; j++) { ss.mean[j] = stats7[r * 7 + j]; ss.var[j] = stats7[r * 7 + 3 + j]; }
    ss.n = stats7[r * 7 + 6];
    for (int j = 0; j < 3; j++) { in.ssMean[j] = ss.mean[j]; in.ssVar[j] = ss.var[j]; }
    in.ssN = ss.n;
    in.begin = (uint32_t)total;
    in.count = counts[r];
    in.mode = first ? 1 : 2;
    in.splitFlag = splitFlags ? (splitFlags[r] ? 1 : 0) : 0;
    in.initK = initK ? (int32_t)initK[r] : 0;
    if (in.initK < 0 || in.initK > kMaxComponents) throw std::runtime_error("region batch: initK must be in [0, 32]");
    float var[3];
    ss.variance(var);
    pivots[r] = make_float4(ss.mean[0], ss.mean[1], ss.mean[2], reprojection_min_distance(var));
    total += counts[r];
    if (first || counts[r] > 0 || in.splitFlag) work.push_back(RegionWork{r, counts[r], (float)counts[r]});
  }
  if (total > kMaxUpdateSamples) throw std::runtime_error("region batch: more than 2^30 samples");
  const uint32_t n = (uint32_t)total;
  dInputs_.ensure(size_t(R) * sizeof(RegionInput));
  dPivots_.ensure(size_t(R) * sizeof(float4));
  dBegin_.ensure(size_t(R) * 4);
  for (DevBuf* b : {&dKeys_, &dVals_, &dDx_, &dDy_, &dDz_, &dW_, &dPdf_, &dDist_}) b->ensure(std::max<size_t>(n, 1) * 4);
  PGL_HIP_CHECK(hipMemcpyAsync(dInputs_.p, inputs.data(), R * sizeof(RegionInput), hipMemcpyHostToDevice, s));
  PGL_HIP_CHECK(hipMemcpyAsync(dPivots_.p, pivots.data(), R * sizeof(float4), hipMemcpyHostToDevice, s));
  PGL_HIP_CHECK(hipMemcpyAsync(dBegin_.p, begin.data(), R * 4, hipMemcpyHostToDevice, s));
  launch_set_inputs(dStates_.as<RegionState>(), dInputs_.as<RegionInput>(), R, s);
  launch_region_keys(dBegin_.as<uint32_t>(), R, n, dKeys_.as<uint32_t>(), dVals_.as<uint32_t>(), s);
  const uint32_t* gatherKeys = dKeys_.as<uint32_t>();
  const uint32_t* gatherIdx = dVals_.as<uint32_t>();
  if (deterministic_ && n > 0) {  // every region's samples in SampleDataLess order (Field.h:409-412, 469-472)
    dDetTmp_.ensure(det_sort_temp_bytes(n, R));
    detPin_.ensure(det_sort_stage_bytes(n, R));
    det_sort_segments(dDetTmp_.p, detPin_.p, detPin_.bytes, records, begin.data(), counts, R, n, dVals_.as<uint32_t>(),
                      nullptr, s);
  }
  launch_gather_reproject(records, gatherKeys, gatherIdx, n, dPivots_.as<float4>(),
                          dDx_.as<float>(), dDy_.as<float>(), dDz_.as<float>(), dW_.as<float>(), dPdf_.as<float>(),
                          dDist_.as<float>(), fc_.noParallax == 0, s);
  PGL_HIP_CHECK(hipMemcpyAsync(dSnapshot_.p, dStates_.p, size_t(R) * sizeof(RegionState), hipMemcpyDeviceToDevice, s));
  PGL_HIP_CHECK(hipEventRecord(ev_[0], s));
  runner_.launch(dStates_.as<RegionState>(), work,
                 RegionSamplesView{dDx_.as<float>(), dDy_.as<float>(), dDz_.as<float>(), dW_.as<float>(),
                                   dPdf_.as<float>(), dDist_.as<float>()},
                 fc_, dUniformDirs_.as<float>(), nullptr, s);
  PGL_HIP_CHECK(hipEventRecord(ev_[1], s));
  PGL_HIP_CHECK(hipStreamSynchronize(s));
  runner_.check();
  info_ = UpdateInfo();
  float ms = 0.f;
  PGL_HIP_CHECK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
  info_.directionalMs = ms;
  info_.regionKernelMs = runner_.kernelMs();
  info_.samples = n;
  info_.regionLaunches = runner_.lastLaunches;
  info_.regionJobs = runner_.lastJobs;
  info_.regionItems = runner_.lastItems;
  info_.maxTeam = runner_.lastMaxTeam;
  info_.regionGrid = runner_.lastGrid;
  info_.estepEvals = (double)runner_.stats().estepEvals;
  info_.auxEvals = (double)runner_.stats().distEvals;
  info_.splitEvals = (double)runner_.stats().splitEvals;
  info_.emIterations = (double)runner_.stats().emIterations;
  info_.wallMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  lastWork_ = work;
  lastSamples_ = n;
  hostStates_.clear();
  steps_++;
}

int RegionBatch::autotune(const std::vector<RegionSchedule>& cands, int reps, std::vector<float>& msOut, hipStream_t s) {
  if (steps_ == 0) throw std::runtime_error("region batch: autotune needs a completed step");
  if (cands.empty()) throw std::runtime_error("region batch: no candidate schedules");
  const RegionSchedule keep = runner_.schedule;
  const size_t bytes = size_t(numRegions_) * sizeof(RegionState);
  msOut.assign(cands.size(), 0.f);
  int best = -1;
  for (size_t c = 0; c < cands.size(); c++) {
    runner_.schedule = cands[c];
    float bestMs = 0.f;
    for (int rep = 0; rep < std::max(1, reps); rep++) {
      PGL_HIP_CHECK(hipMemcpyAsync(dStates_.p, dSnapshot_.p, bytes, hipMemcpyDeviceToDevice, s));
      runner_.launch(dStates_.as<RegionState>(), lastWork_,
                     RegionSamplesView{dDx_.as<float>(), dDy_.as<float>(), dDz_.as<float>(), dW_.as<float>(),
                                       dPdf_.as<float>(), dDist_.as<float>()},
                     fc_, dUniformDirs_.as<float>(), nullptr, s);
      PGL_HIP_CHECK(hipStreamSynchronize(s));
      runner_.check();
      const float ms = runner_.kernelMs();
      if (rep == 0 || ms < bestMs) bestMs = ms;
    }
    msOut[c] = bestMs;
    if (best < 0 || bestMs < msOut[best]) best = (int)c;
  }
  runner_.schedule = best >= 0 ? cands[best] : keep;
  hostStates_.clear();
  return best;
}

void RegionBatch::exportRegion(uint32_t r, PGLRegionRecordHost* out, hipStream_t s) {
  if (r >= numRegions_) throw std::runtime_error("region batch: region index out of range");
  if (hostStates_.size() != numRegions_) {
    hostStates_.resize(numRegions_);
    PGL_HIP_CHECK(hipMemcpyAsync(hostStates_.data(), dStates_.p, size_t(numRegions_) * sizeof(RegionState),
                                 hipMemcpyDeviceToHost, s));
    PGL_HIP_CHECK(hipStreamSynchronize(s));
  }
  export_state(hostStates_[r], out);
  const RegionState& S = hostStates_[r];
  for (int j = 0; j < 3; j++) { out->ssMean[j] = S.ssMean[j]; out->ssVariance[j] = S.ssVar[j]; }
  out->ssNumSamples = S.ssN;
}

}  // namespace pglgpu
