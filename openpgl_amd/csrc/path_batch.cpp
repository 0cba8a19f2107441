// Host side of the batched path-segment back-propagation (kernels_path.hip).
#include "path_batch.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace pglgpu {

static_assert(sizeof(PathSample) == kSampleRecordBytes, "PathSample is the 40-byte PGLSampleData record");

PathSampleBatch::PathSampleBatch() {
  PGL_HIP_CHECK(hipStreamCreate(&stream_));
}

PathSampleBatch::~PathSampleBatch() {
  (void)hipStreamSynchronize(stream_);
  (void)hipStreamDestroy(stream_);
}

void PathSampleBatch::count(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
                            const PathSampleOptions& o, size_t& numSurf, size_t& numVol) {
  numSurf = numVol = 0;
  offsets_ = offsets;
  numSegments_ = numPaths ? numSegments : 0;
  totals_.ensure(24);
  hTotals_.ensure(24);
  const size_t V = std::max<size_t>(numSegments_, 1);
  records_.ensure(V * sizeof(PathSample));
  flags_.ensure(V * 8);
  incl_.ensure(V * 8);
  pathOf_.ensure(V * 4);
  tiles_.ensure(path_tile_words((uint32_t)V) * 4);
  const size_t scanBytes = path_scan_temp_bytes(std::max<uint32_t>(numSegments_, 1));
  scanTemp_.ensure(std::max<size_t>(scanBytes, 16));
  launch_path_records(segs, numSegments_, offsets, numPaths, o, records_.as<PathSample>(),
                      flags_.as<unsigned long long>(), pathOf_.as<uint32_t>(), tiles_.as<uint32_t>(),
                      incl_.as<unsigned long long>(),
                      scanTemp_.p, scanBytes, totals_.as<unsigned long long>(), stream_);
  PGL_HIP_CHECK(hipMemcpyAsync(hTotals_.p, totals_.p, 24, hipMemcpyDeviceToHost, stream_));
  PGL_HIP_CHECK(hipStreamSynchronize(stream_));
  const unsigned long long* t = hTotals_.as<unsigned long long>();
  if (t[2]) throw std::runtime_error("openpgl_amd: path offsets must be non-decreasing and within the segments");
  numSurf = (size_t)t[0];
  numVol = (size_t)t[1];
}

void PathSampleBatch::scatter(PathSample* surfOut, PathSample* volOut) {
  launch_path_scatter(records_.as<PathSample>(), flags_.as<unsigned long long>(), pathOf_.as<uint32_t>(),
                      incl_.as<unsigned long long>(), offsets_, numSegments_, surfOut, volOut, stream_);
  PGL_HIP_CHECK(hipStreamSynchronize(stream_));
}

void PathSampleBatch::prepare(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets,
                              uint32_t numPaths, const PathSampleOptions& o, PathSample* surfOut, size_t surfCap,
                              PathSample* volOut, size_t volCap, size_t& numSurf, size_t& numVol) {
  count(segs, numSegments, offsets, numPaths, o, numSurf, numVol);
  if (numSurf > surfCap || numVol > volCap)
    throw std::runtime_error("openpgl_amd: path samples exceed the output capacity (" + std::to_string(numSurf) +
                             " surface, " + std::to_string(numVol) + " volume)");
  if (numSurf == 0 && numVol == 0) return;
  scatter(surfOut, volOut);
}

void PathSampleBatch::prepareOwned(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets,
                                   uint32_t numPaths, const PathSampleOptions& o, const PathSample*& surf,
                                   size_t& numSurf, const PathSample*& vol, size_t& numVol) {
  const int b = next_;
  count(segs, numSegments, offsets, numPaths, o, numSurf, numVol);
  next_ = 1 - next_;
  acquireOwned(b, numSurf, numVol);
  surf = out_[b][0].as<PathSample>();
  vol = out_[b][1].as<PathSample>();
  if (numSurf || numVol) scatter(out_[b][0].as<PathSample>(), out_[b][1].as<PathSample>());
  lastBuffer_ = b;
}

void PathSampleBatch::acquireOwned(int b, size_t numSurf, size_t numVol) {
  // the buffer's previous consumers (the gathers of an earlier update) must be done with it
  for (int k = 0; k < 2; k++)
    if (pendingRelease_[b][k]) {
      PGL_HIP_CHECK(hipStreamWaitEvent(stream_, released_[b][k], 0));
      pendingRelease_[b][k] = false;
    }
  const size_t ns = std::max<size_t>(numSurf, 1), nv = std::max<size_t>(numVol, 1);
  if (out_[b][0].bytes < ns * sizeof(PathSample) || out_[b][1].bytes < nv * sizeof(PathSample))
    PGL_HIP_CHECK(hipStreamSynchronize(stream_));  // a reallocation frees memory the wait above protects
  out_[b][0].ensure(ns * sizeof(PathSample));
  out_[b][1].ensure(nv * sizeof(PathSample));
}

void PathSampleBatch::routeOwned(const PathSample* recs, uint32_t n, const PathSample*& surf, size_t& numSurf,
                                 const PathSample*& vol, size_t& numVol) {
  const int b = next_;
  next_ = 1 - next_;
  totals_.ensure(24);
  hTotals_.ensure(24);
  const size_t V = std::max<size_t>(n, 1);
  flags_.ensure(V * 8);
  incl_.ensure(V * 8);
  const size_t scanBytes = path_scan_temp_bytes((uint32_t)V);
  scanTemp_.ensure(std::max<size_t>(scanBytes, 16));
  launch_route_count(recs, n, flags_.as<unsigned long long>(), incl_.as<unsigned long long>(), scanTemp_.p,
                     scanBytes, totals_.as<unsigned long long>(), stream_);
  PGL_HIP_CHECK(hipMemcpyAsync(hTotals_.p, totals_.p, 24, hipMemcpyDeviceToHost, stream_));
  PGL_HIP_CHECK(hipStreamSynchronize(stream_));
  const unsigned long long* t = hTotals_.as<unsigned long long>();
  numSurf = (size_t)t[0];
  numVol = (size_t)t[1];
  acquireOwned(b, numSurf, numVol);
  surf = out_[b][0].as<PathSample>();
  vol = out_[b][1].as<PathSample>();
  if (n) {
    launch_route_scatter(recs, incl_.as<unsigned long long>(), n, out_[b][0].as<PathSample>(),
                         out_[b][1].as<PathSample>(), stream_);
    PGL_HIP_CHECK(hipStreamSynchronize(stream_));
  }
  lastBuffer_ = b;
}

void PathSampleBatch::setReleased(hipEvent_t surfaceDone, hipEvent_t volumeDone) {
  const int b = lastBuffer_;
  released_[b][0] = surfaceDone;
  released_[b][1] = volumeDone;
  pendingRelease_[b][0] = surfaceDone != nullptr;
  pendingRelease_[b][1] = volumeDone != nullptr;
}

}  // namespace pglgpu
