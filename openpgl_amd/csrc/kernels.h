// Launchers of the gfx950 kernels (implemented in the *.hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "common.h"

#define PGL_KERNEL_CHECK(x)                                                                       \
  do {                                                                                            \
    const hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace pglgpu {

struct HGRep;

// Device view of one trained field used by the query kernel.
struct QueryField {
  const KDNode* nodes;
  const RegionState* regions;
  const uint32_t* nbIds;
  const float4* nbPts;
  uint32_t nbSize;
  int32_t knn;
  int32_t trained;
  float lo[3], hi[3];
  int32_t parallax;  // 0: plain VMM field, no shift to the query position
  int32_t knnFused;  // KNN squared distances with FMA (CPU_8/CPU_16 devices) or without (CPU_4)
};

// spatial
// positions to SoA; wkey (optional): order key of every record's weight (deterministic mode)
void launch_ingest(const void* aos, uint32_t n, float* px, float* py, float* pz, uint32_t* wkey, hipStream_t s);
uint32_t launch_bounds(const float* px, const float* py, const float* pz, uint32_t n, float* partial, hipStream_t s);
// leafParams: per leaf slot {center[3], invHalf[3]} of the leaf's parent bounds.
// partial: scratch of spatial_partial_bytes(n) bytes for the per-block tables.
size_t spatial_partial_bytes(uint32_t n);
void launch_descend_stats(const float* px, const float* py, const float* pz, uint32_t n, const KDNode* nodes,
                          uint32_t numNodes, const int32_t* leafSlot, const void* leafParams, uint32_t numLeaves,
                          uint32_t* nodeOf, IntStatsAcc* partial, IntStatsAcc* acc, hipStream_t s);
void launch_kd_round(const float* px, const float* py, const float* pz, uint32_t n, uint32_t* nodeOf,
                     const int32_t* frontierSlot, uint32_t numNodes, const FrontierNode* frontier,
                     uint32_t numFrontier, IntStatsAcc* partial, IntStatsAcc* childAcc, hipStream_t s);
void launch_leaf_keys(const uint32_t* nodeOf, const KDNode* nodes, uint32_t n, uint32_t* keys, uint32_t* vals,
                      hipStream_t s);
size_t sort_pairs_temp_bytes(uint32_t n);
void sort_pairs(void* temp, size_t tempBytes, const uint32_t* keysIn, uint32_t* keysOut, const uint32_t* valsIn,
                uint32_t* valsOut, uint32_t n, int endBit, hipStream_t s);
void launch_region_ranges(const uint32_t* sortedKeys, uint32_t n, uint32_t* begin, uint32_t* end, hipStream_t s);
// Stable sort of (key, index) pairs for keys < numKeys <= kSortMaxKeys
// (kernels_sort.hip); keyBegin[k] = first output position of key k (the
// running sum of the key counts); keys >= numKeys are dropped.
static const uint32_t kSortMaxKeys = 4096;
size_t counting_sort_temp_bytes(uint32_t n, uint32_t numKeys);
void counting_sort_pairs(void* temp, const uint32_t* keys, const uint32_t* keyBegin, uint32_t numKeys, uint32_t n,
                         uint32_t* keysOut, uint32_t* valsOut, hipStream_t s);
// the same with the key of sample i computed as keyNodes[nodeOf[i]].idx()
// (the region of its leaf; launch_leaf_keys and the sort in one); with
// carryIn (optional), carryOut[j] = carryIn[valsOut[j]] is written as well
void counting_sort_leaves(void* temp, const uint32_t* nodeOf, const KDNode* keyNodes, const uint32_t* keyBegin,
                          uint32_t numKeys, uint32_t n, uint32_t* keysOut, uint32_t* valsOut, hipStream_t s,
                          const uint32_t* carryIn = nullptr, uint32_t* carryOut = nullptr);
// the same moving the 40-byte records into region order: recsOut[j] =
// recsIn[i] for the sample i sorted to position j, keysOut[j] its region,
// posOut[j] = j (optional), carryOut[j] = carryIn[i] (optional).  Later
// passes then read each region's records from one contiguous range.
void counting_sort_leaf_records(void* temp, const uint32_t* nodeOf, const KDNode* keyNodes, const uint32_t* keyBegin,
                                uint32_t numKeys, uint32_t n, const void* recsIn, void* recsOut, uint32_t* keysOut,
                                uint32_t* posOut, const uint32_t* carryIn, uint32_t* carryOut, hipStream_t s);
// Deterministic-mode order (kernels_sort.hip): every region's range of a
// region-contiguous index array put in SampleDataLess order
// (data/SampleData.h:82-93), as Field::fitRegions/updateRegions sort them
// (field/Field.h:409-412, 469-472).  records: the PGLSampleData records the
// indices refer to; region r's indices are vals[begin[r], begin[r] +
// count[r]) (host arrays), sorted in place.  temp: det_sort_temp_bytes(n,
// numSegments) bytes of device memory; pinnedHost: det_sort_stage_bytes(n,
// numSegments) bytes of pinned host memory for the segment tables (not
// reused before the stream passes this point).  No host synchronisation.
size_t det_sort_temp_bytes(uint32_t n, uint32_t numSegments);
size_t det_sort_stage_bytes(uint32_t n, uint32_t numSegments);
// weightKeys (optional): float_order_key of the weight of the record vals[j]
// at position j (in the same layout as vals; overwritten); null: read from
// the records.
void det_sort_segments(void* temp, void* pinnedHost, size_t pinnedBytes, const void* records,
                       const uint32_t* begin, const uint32_t* count, uint32_t numSegments, uint32_t n,
                       uint32_t* vals, uint32_t* weightKeys, hipStream_t s);
// keys/values of region-contiguous records (begin: numRegions ascending offsets)
void launch_region_keys(const uint32_t* begin, uint32_t numRegions, uint32_t n, uint32_t* keys, uint32_t* vals,
                        hipStream_t s);
// SoA output j = record sortedIdx[j] (sortedIdx null: record j), reprojected
// to the pivot of region sortedKeys[j]
void launch_gather_reproject(const void* aos, const uint32_t* sortedKeys, const uint32_t* sortedIdx, uint32_t n,
                             const float4* regionPivot, float* dx, float* dy, float* dz, float* w, float* pdf,
                             float* dist, bool reproject, hipStream_t s);
void launch_knn8(const float4* pivots, uint32_t numRegions, uint32_t* nbIds, float4* nbPts, hipStream_t s);

// directional
// Directional update of the regions listed in jobs (persistent kernel, see
// kernels_region.hip).  items: (job, rank) pairs, ranks of a job consecutive;
// queueAndCounters: numJobs + 1 words (cleared by the launcher); partial: per
// job 2 x chunks x kPartStride floats at job.partialOffset; err: set to 1 when
// a team barrier waited longer than timeoutTicks of the wall clock
// (wall_clock64); stats: E-step, distance and split evaluations and EM
// iterations are added to stats[0..3]; costHint (optional, per region):
// evaluations per sample of this update.  Every job's team must be <= grid.
struct RegionLaunch {
  RegionState* states;
  const RegionJob* jobs;
  uint32_t numJobs;
  const uint2* items;
  uint32_t numItems;
  uint32_t grid;
  uint32_t* queueAndCounters;
  float* partial;
  int* err;
  unsigned long long timeoutTicks;
  unsigned long long* stats;
  float* costHint;
  const float *dx, *dy, *dz, *w, *pdf, *dist;
  const float* uniformDirs;
  bool cooperative;
  bool zeroQueue;  // zero the queue and team counters before the launch (else done by launch_prepare_regions)
};
int region_kernel_blocks_per_cu();
void launch_region_update(const RegionLaunch& launch, const FactoryConfig& cfg, hipStream_t s);

void launch_copy_states(RegionState* states, const uint32_t* pairs /* dst, src */, uint32_t numPairs, hipStream_t s);
void launch_set_inputs(RegionState* states, const RegionInput* inputs, uint32_t count, hipStream_t s);
// One kernel before a persistent region launch, reading its tables from
// pinned host memory: copies tableBytes (a multiple of 16) from hostTables to
// tables, zeroes numZero words at zero (queue and team counters), the error
// flag and the four launch totals; with inputs (host, pinned), every region's
// state is first copied from its ancestor where copyFrom >= 0 and then given
// its inputs (launch_copy_states followed by launch_set_inputs).
void launch_prepare_regions(RegionState* states, const void* hostTables, void* tables, size_t tableBytes,
                            const RegionInput* inputs, uint32_t numInputs, uint32_t* zero, uint32_t numZero, int* err,
                            unsigned long long* stats, hipStream_t s);

// query
// One piece of a query batch: queries [begin, begin + n) of the batch read
// their inputs and write their results through these pointers (device or
// device-accessible pinned host memory); g, u1, vol, id, integral may be null.
struct QueryPiece {
  const float *pos, *nrm, *g, *u2, *u1;
  const uint8_t* vol;
  float *dir, *pdf, *integral;
  uint32_t* id;
  uint32_t begin, n;
};
void launch_query_pieces(const QueryField& surf, const QueryField& vol, const HGRep* hg, const QueryPiece* pieces,
                         uint32_t numPieces, uint32_t n, hipStream_t s);
void launch_query(const QueryField& surf, const QueryField& vol, const HGRep* hg, uint32_t n, const float* pos,
                  const float* nrm, const float* g, const float* u2, const float* u1, const uint8_t* isVolume,
                  float* dirOut, float* pdfOut, uint32_t* idOut, float* integralOut, hipStream_t s);

// path segments -> training samples (kernels_path.hip).  Per vertex: its
// record, a flag packed as surface | volume << 32 and its path; totals:
// [0] surface records, [1] volume records, [2] 1 if a path range was malformed
struct PathSampleOptions;
struct PathSample;
size_t path_scan_temp_bytes(uint32_t numSegments);
size_t path_tile_words(uint32_t numSegments);  // entries of tileFirst
void launch_path_records(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
                         const PathSampleOptions& o, PathSample* records, unsigned long long* flags, uint32_t* pathOf,
                         uint32_t* tileFirst, unsigned long long* incl, void* scanTemp, size_t scanBytes,
                         unsigned long long* totals, hipStream_t s);
void launch_path_scatter(const PathSample* records, const unsigned long long* flags, const uint32_t* pathOf,
                         const unsigned long long* incl, const uint32_t* offsets, uint32_t numSegments,
                         PathSample* surfOut, PathSample* volOut, hipStream_t s);

// Records split by their EInsideVolume flag, each part in input order
// (SampleDataStorage.h:72-90): flags + scan (totals[0] surface, [1] volume),
// then the scatter.
void launch_route_count(const PathSample* recs, uint32_t n, unsigned long long* flags, unsigned long long* incl,
                        void* scanTemp, size_t scanBytes, unsigned long long* totals, hipStream_t s);
void launch_route_scatter(const PathSample* recs, const unsigned long long* incl, uint32_t n, PathSample* surfOut,
                          PathSample* volOut, hipStream_t s);

// The same two passes with every scratch buffer in one static device array
// per GPU (15.25 MiB, held until the process ends): at most
// kNoAllocMaxVertices vertices per call, no device or pinned allocation.
// Calls that overlap on one GPU share that array and race.  Blocks until the
// records are written; throws on a size above the maximum, malformed ranges
// or a capacity that is too small (nothing written then).
constexpr uint32_t kNoAllocMaxVertices = 1u << 18;
constexpr size_t kNoAllocScratchBytes = 15995392;
void path_prepare_noalloc(const PathVertex* segs, uint32_t numSegments, const uint32_t* offsets, uint32_t numPaths,
                          const PathSampleOptions& o, PathSample* surfOut, size_t surfCap, PathSample* volOut,
                          size_t volCap, size_t& numSurf, size_t& numVol, hipStream_t s);

}  // namespace pglgpu
