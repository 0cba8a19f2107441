// Shared host/device layouts of the MI355X guiding field.
//
// Everything the kernels read or write in HBM is declared here as plain
// structs of 32-bit words so that host code (openpgl_amd/csrc/*.cpp) and HIP
// kernels (*.hip) agree on one layout.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PGL_HD __host__ __device__
#else
#define PGL_HD
#endif

namespace pglgpu {

static const int kMaxComponents = 32;   // PGL_VMM_MAX_COMPONENTS (include/openpgl/common.h:26)
// Directional passes: samples of a region are reduced in fixed chunks of
// kChunk samples; one chunk partial holds kPartStride floats.
static const uint32_t kChunk = 1024;
static const int kPartStride = 256;
static const size_t kSampleRecordBytes = 40;  // PGLSampleData (include/openpgl/data.h:21-49)
// samples of one update: sample indices, sort keys and the region kernel's
// sample offsets are 32-bit; 2^30 keeps every grid-stride index below 2^31
static const uint32_t kMaxUpdateSamples = 1u << 30;
static const int kMergeWidth = 8;       // merge loop floor = SIMD width of the CPU_8 device
static const float kOneOverFourPi = 0.07957747154594767f;
static const float kMaxKappaSplit = 32000.0f;  // OPENPGL_MAX_KAPPA used by the splitter

// kd node, same 8-byte encoding as KDNode (spatial/kdtree/KDTree.h:21-31):
// bits 31..30 = split dim (3 = leaf), bits 29..0 = left child / region index.
struct KDNode {
  float pivot;
  uint32_t bits;
  PGL_HD bool isLeaf() const { return (bits >> 30) == 3u; }
  PGL_HD uint32_t dim() const { return bits >> 30; }
  PGL_HD uint32_t idx() const { return bits & 0x3fffffffu; }
};

// Training configuration of the directional factory (PGLVMMFactoryArguments,
// include/openpgl/config.h:41-81) as consumed by the region kernel.
struct FactoryConfig {
  int32_t initK;
  float initKappa;
  int32_t maxEMIterations;
  float maxMeanCosine;           // KappaToMeanCosine(maxKappa)
  float convergenceThreshold;
  float weightPrior;
  float meanCosinePriorStrength;
  float meanCosinePrior;
  int32_t useSplitAndMerge;
  float splittingThreshold;
  float mergingThreshold;
  int32_t partialReFit;
  int32_t minSamplesForSplitting;
  int32_t minSamplesForPartialRefitting;
  int32_t minSamplesForMerging;
  float decayOnSpatialSplit;
  float maxMeanCosineSplit;      // KappaToMeanCosine(32000)
  int32_t vectorWidth;           // SIMD width of the emulated reference device (merge floor, clear granularity)
  int32_t noParallax;            // 1: plain VMM field (PGL_DIRECTIONAL_DISTRIBUTION_VMM): no shift before updates
  float pad[1];
};

// ParallaxAwareVonMisesFisherMixture (PAVMM:22-87) as SoA over 32 slots.
struct VMMParams {
  float w[32], kappa[32], mux[32], muy[32], muz[32], norm[32], em2k[32], mc[32], dist[32];
  float pivot[3];
  int32_t K;
};

// SufficientStatistics (WEM:93-168).
struct SuffStats {
  float swdx[32], swdy[32], swdz[32], sw[32], sdw[32];
  float sumWeights, numSamples, overall;
  int32_t K;
  int32_t normalized;
};

// ComponentSplitStatistics (Splitter:57-104); the zero-mean variant keeps no means.
struct SplitStatsState {
  float chi[32], covx[32], covy[32], covz[32], sn[32], ssw[32], ssa[32];
  int32_t K;
};

// One path vertex as the renderer records it: PGLPathSegmentData
// (include/openpgl/data.h:50-113), 136 bytes with the same member offsets.
struct PathVertex {
  float px, py, pz;        // position
  float dix, diy, diz;     // directionIn (towards the next vertex)
  float dox, doy, doz;     // directionOut (towards the previous vertex)
  float nx, ny, nz;        // normal
  uint8_t volumeScatter;
  uint8_t pad0[3];
  float pdfDirectionIn;
  uint8_t isDelta;
  uint8_t pad1[3];
  float scatteringWeight[3];
  float transmittanceWeight[3];
  float directContribution[3];
  float miWeight;
  float scatteredContribution[3];
  float russianRouletteProbability;
  float eta;
  float roughness;
  uint32_t pad2;
  uint64_t regionPtr;
};
static_assert(sizeof(PathVertex) == 136, "PathVertex is PGLPathSegmentData (136 B)");

// Per-region directional state kept resident in HBM (one 4 KiB record per region).
struct RegionState {
  VMMParams vmm;
  SuffStats suff;
  SplitStatsState split;
  uint32_t afterSplitLo, afterSplitHi, afterMergeLo, afterMergeHi;  // size_t counters
  // --- per-update inputs written by the host before the region kernel
  float ssMean[3];
  float ssVar[3];       // sampleVariance (= variance * n)
  float ssN;
  uint32_t sampleBegin, sampleCount;
  int32_t mode;         // 0 = skip, 1 = fit (first frame), 2 = update
  int32_t splitFlag;    // region was split spatially since the last update
  // --- outputs
  int32_t valid;
  int32_t emIterations, partialEmIterations, numSplits, numMerges;
  int32_t fitted;
  float lastLogLikelihood;
  float estepEvals;     // sample x component evaluations in E-steps of the last update
  float auxEvals;       // sample x component evaluations in distance passes
  float splitEvals;     // sample x component evaluations in split-statistics passes
  // profile of the last update (rank 0 of the region's team), wall-clock
  // ticks: [0] total, [1] E-step passes, [2] split passes, [3] distance
  // passes, [4] team barrier waits, [5] merging, [6] splitting logic,
  // [9] M-step and convergence test, [10] cross-chunk folds;
  // counts: [7] team size, [8] E-step passes, [11] samples
  float prof[12];
  int32_t initK;        // components of the initial mixture of a fit (0: the factory's initK)
  uint32_t pad[304];
};
static_assert(sizeof(RegionState) == 4096, "RegionState must be 4 KiB");

// Integer sample statistics accumulator of one kd node (data/SampleStatistics.h:195-342)
struct IntStatsAcc {
  unsigned long long n;
  long long sum[3];
  long long sq[3];
  long long lo[3];
  long long hi[3];
  long long pad[3];
};
static_assert(sizeof(IntStatsAcc) == 128, "IntStatsAcc is 128 B");

// Per-update inputs of one region written into its RegionState.
struct RegionInput {
  uint32_t region;
  float ssMean[3], ssVar[3], ssN;
  uint32_t begin, count;
  int32_t mode, splitFlag;
  int32_t initK;          // 0: the factory's initK
  int32_t copyFrom = -1;  // >= 0: a new region whose state starts as a copy of this ancestor's
};

// One region of a directional launch: team = workgroups working on it,
// chunks = ceil(samples / kChunk), partialOffset = float offset of its
// 2 x chunks x kPartStride partial buffer.
struct RegionJob {
  uint32_t region, team, chunks, pad;
  unsigned long long partialOffset;
};

// Node being pushed down one level in a kd round.
struct FrontierNode {
  uint32_t node;       // inner node id
  uint32_t left;       // left child id (right = left + 1)
  uint32_t dim;
  float pivot;
  float center[3];     // bounds centre of this node (integer statistics of both children)
  float invHalf[3];    // embree::rcp(half extent), computed on the host
};

}  // namespace pglgpu
