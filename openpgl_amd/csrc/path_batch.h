// Path segments -> training samples for many paths at once on the GPU
// (the batched form of PathSegmentDataStorage::prepareSamples +
// propagateSamples, data/PathSegmentDataStorage.h:160-315, 518-530).
#pragma once
#include <hip/hip_runtime.h>

#include "field.h"
#include "path_samples.h"

namespace pglgpu {

class PathSampleBatch {
 public:
  PathSampleBatch();
  ~PathSampleBatch();
  PathSampleBatch(const PathSampleBatch&) = delete;
  PathSampleBatch& operator=(const PathSampleBatch&) = delete;

  // Counts the samples of paths [offsets[p], offsets[p+1]) (device arrays;
  // a range outside [0, numSegments] is an error),
  // then writes them to surfOut / volOut (device, capacities in records).
  // Throws if a capacity is too small (nothing is written then).  Blocks
  // until the records are written.
  void prepare(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
               const PathSampleOptions& o, PathSample* surfOut, size_t surfCap, PathSample* volOut, size_t volCap,
               size_t& numSurf, size_t& numVol);

  // Samples of the paths written into the batch's own double-buffered output
  // (for a field update that reads them after this call returns).  Before a
  // buffer is reused, the stream waits for `released` events recorded by
  // the previous consumer of that buffer.
  void prepareOwned(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
                    const PathSampleOptions& o, const PathSample*& surf, size_t& numSurf, const PathSample*& vol,
                    size_t& numVol);
  // Mixed surface / volume records (device) split by their EInsideVolume
  // flag into the batch's double-buffered output, each part in input order,
  // as a sample storage routes them (SampleDataStorage.h:72-90).
  void routeOwned(const PathSample* recs, uint32_t n, const PathSample*& surf, size_t& numSurf,
                  const PathSample*& vol, size_t& numVol);
  // events after which the buffer handed out by the last prepareOwned / routeOwned may be overwritten
  void setReleased(hipEvent_t surfaceDone, hipEvent_t volumeDone);

  hipStream_t stream() const { return stream_; }

 private:
  void count(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
             const PathSampleOptions& o, size_t& numSurf, size_t& numVol);
  void scatter(PathSample* surfOut, PathSample* volOut);
  // output buffer b for ns / nv records, once its previous consumers are done
  void acquireOwned(int b, size_t numSurf, size_t numVol);

  hipStream_t stream_ = nullptr;
  DevBuf records_, flags_, pathOf_, tiles_, incl_, scanTemp_, totals_;
  const uint32_t* offsets_ = nullptr;  // of the last count()
  uint32_t numSegments_ = 0;
  PinnedBuf hTotals_;
  DevBuf out_[2][2];  // [buffer][surface, volume]
  hipEvent_t released_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  bool pendingRelease_[2][2] = {{false, false}, {false, false}};
  int next_ = 0;
  int lastBuffer_ = 0;
};

}  // namespace pglgpu
