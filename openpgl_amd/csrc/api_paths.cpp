// C ABI of the path-segment storage (openpgl/api/api.cpp:444-562 in the
// reference) and its batched GPU form.
//
// The per-path storage is the renderer-side object of the reference: a
// fixed-capacity array of path vertices per path (one per render thread),
// turned into training samples by PrepareSamples and appended to a sample
// storage by PropagateSamples.  It runs on the host, as in the reference.
// The batched EXTENSION functions take the vertices of many paths in device
// memory and produce the samples (or a field update) on the GPU with the
// same per-path routine (path_samples.h).
#pragma clang fp contract(off)
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "api_impl.h"
#include "path_batch.h"
#include "path_samples.h"

using namespace pglgpu;

static_assert(sizeof(PGLPathSegmentData) == sizeof(PathVertex), "PGLPathSegmentData layout");
static_assert(offsetof(PGLPathSegmentData, pdfDirectionIn) == offsetof(PathVertex, pdfDirectionIn), "layout");
static_assert(offsetof(PGLPathSegmentData, isDelta) == offsetof(PathVertex, isDelta), "layout");
static_assert(offsetof(PGLPathSegmentData, scatteringWeight) == offsetof(PathVertex, scatteringWeight), "layout");
static_assert(offsetof(PGLPathSegmentData, miWeight) == offsetof(PathVertex, miWeight), "layout");
static_assert(offsetof(PGLPathSegmentData, roughness) == offsetof(PathVertex, roughness), "layout");
static_assert(offsetof(PGLPathSegmentData, regionPtr) == offsetof(PathVertex, regionPtr), "layout");
static_assert(sizeof(PGLSampleData) == sizeof(PathSample), "PGLSampleData layout");

// PathSegmentDataStorage (data/PathSegmentDataStorage.h:15-530): capacity
// set by Reserve, vertices appended by NextSegment / AddSegment, samples by
// PrepareSamples / AddSample; additions beyond the capacity are dropped.
struct PGLPathSegmentStorageImpl {
  std::vector<PGLPathSegmentData> segments;
  std::vector<PGLSampleData> samples;
  size_t capacity = 0;
  size_t numSegments = 0;
  size_t numSamples = 0;
  float maxDistance = 1e6f;
};

// default-constructed PGLPathSegmentData (include/openpgl/data.h:55-112)
static PGLPathSegmentData default_segment() {
  PGLPathSegmentData d;
  memset(&d, 0, sizeof(d));
  d.directionIn.y = 1.f;
  d.directionOut.x = 1.f;
  d.normal.z = 1.f;
  d.pdfDirectionIn = 1.f;
  d.scatteringWeight = {1.f, 1.f, 1.f};
  d.transmittanceWeight = {1.f, 1.f, 1.f};
  d.miWeight = 1.f;
  d.russianRouletteProbability = 1.f;
  d.eta = 1.f;
  d.roughness = 1.f;
  return d;
}

static size_t prepare_host(PGLPathSegmentStorageImpl* s, bool useNEEMiWeights, bool guideDirectLight,
                           bool rrAffectsDirectContribution) {
  const PathSampleOptions o{s->maxDistance, useNEEMiWeights, guideDirectLight, rrAffectsDirectContribution};
  path_samples((const PathVertex*)s->segments.data(), (int)s->numSegments, o, [&](const PathSample& x) {
    if (s->numSamples < s->capacity) memcpy(&s->samples[s->numSamples++], &x, sizeof(x));
  });
  return s->numSamples;
}

// finite(x) (embree::isvalid)
static bool fin(float x) { return std::isfinite(x); }

// isValid(PathSegmentData) (data/PathSegmentData.h:19-103)
static bool segment_valid(const PGLPathSegmentData& p) {
  auto unit = [](const pgl_vec3f& v) {
    return fin(v.x) && v.x <= 1.f && v.x >= -1.f && fin(v.y) && v.y <= 1.f && v.y >= -1.f && fin(v.z) &&
           v.z <= 1.f && v.z >= -1.f;
  };
  auto nonneg = [](const pgl_vec3f& v) {
    return fin(v.x) && v.x >= 0.f && fin(v.y) && v.y >= 0.f && fin(v.z) && v.z >= 0.f;
  };
  bool ok = fin(p.position.x) && fin(p.position.y) && fin(p.position.z);
  ok = ok && unit(p.directionIn) && unit(p.directionOut) && unit(p.normal);
  ok = ok && fin(p.pdfDirectionIn) && p.pdfDirectionIn >= 0.f;
  ok = ok && nonneg(p.scatteringWeight) && nonneg(p.transmittanceWeight) && nonneg(p.directContribution) &&
       nonneg(p.scatteredContribution);
  ok = ok && fin(p.miWeight) && p.miWeight >= 0.f && p.miWeight <= 1.f;
  // the reference tests miWeight >= 0 here, not the probability (PathSegmentData.h:92)
  ok = ok && fin(p.russianRouletteProbability) && p.miWeight >= 0.f && p.russianRouletteProbability <= 1.f;
  ok = ok && fin(p.eta) && p.eta >= 0.f;
  ok = ok && fin(p.roughness) && p.roughness >= 0.f && p.roughness <= 1.f;
  return ok;
}

// isValid(SampleData) (data/SampleData.h:25-46)
static bool sample_valid(const PGLSampleData& d) {
  return fin(d.position.x) && fin(d.position.y) && fin(d.position.z) && fin(d.direction.x) && fin(d.direction.y) &&
         fin(d.direction.z) && fin(d.weight) && d.weight >= 0.f && fin(d.pdf) && d.pdf > 0.f && fin(d.distance) &&
         d.distance > 0.f;
}

extern "C" {

PGLPathSegmentStorage pglNewPathSegmentStorage(void) PGL_CATCH_BEGIN { return new PGLPathSegmentStorageImpl(); }
PGL_CATCH_END(nullptr)

void pglReleasePathSegmentStorage(PGLPathSegmentStorage storage) PGL_CATCH_BEGIN { delete storage; }
PGL_CATCH_END_VOID

void pglPathSegmentStorageReserve(PGLPathSegmentStorage s, size_t size) {
  if (s->capacity == size) return;
  s->segments.assign(size, default_segment());
  s->samples.assign(size, PGLSampleData());
  s->capacity = size;
  s->numSegments = 0;
  s->numSamples = 0;
}

void pglPathSegmentStorageClear(PGLPathSegmentStorage s) {
  s->numSegments = 0;
  s->numSamples = 0;
}

void pglPathSegmentSetMaxDistance(PGLPathSegmentStorage s, float maxDistance) { s->maxDistance = maxDistance; }
float pglPathSegmentGetMaxDistance(PGLPathSegmentStorage s) { return s->maxDistance; }
int pglPathSegmentGetNumSegments(PGLPathSegmentStorage s) { return (int)s->numSegments; }
int pglPathSegmentGetNumSamples(PGLPathSegmentStorage s) { return (int)s->numSamples; }

size_t pglPathSegmentStoragePrepareSamples(PGLPathSegmentStorage s, const bool useNEEMiWeights,
                                           const bool guideDirectLight, const bool rrAffectsDirectContribution) {
  return prepare_host(s, useNEEMiWeights, guideDirectLight, rrAffectsDirectContribution);
}

pgl_vec3f pglPathSegmentStorageCalculatePixelEstimate(PGLPathSegmentStorage s, const bool rrAffectsDirectContribution) {
  float c[3];
  path_pixel_estimate((const PathVertex*)s->segments.data(), (int)s->numSegments, rrAffectsDirectContribution, c);
  pgl_vec3f r;
  r.x = c[0];
  r.y = c[1];
  r.z = c[2];
  return r;
}

const PGLSampleData* pglPathSegmentStorageGetSamples(PGLPathSegmentStorage s, size_t& nSamples) {
  nSamples = s->numSamples;
  return s->samples.data();
}

void pglPathSegmentStorageAddSample(PGLPathSegmentStorage s, PGLSampleData sample) {
  if (s->numSamples < s->capacity) s->samples[s->numSamples++] = sample;
}

PGLPathSegmentData* pglPathSegmentStorageNextSegment(PGLPathSegmentStorage s) {
  if (s->numSegments >= s->capacity) return nullptr;
  PGLPathSegmentData* d = &s->segments[s->numSegments++];
  *d = default_segment();
  return d;
}

void pglPathSegmentStorageAddSegment(PGLPathSegmentStorage s, PGLPathSegmentData segment) {
  if (s->numSegments < s->capacity) s->segments[s->numSegments++] = segment;
}

bool pglPathSegmentStorageValidateSamples(PGLPathSegmentStorage s) {
  bool ok = true;
  for (size_t i = 0; i < s->numSamples; i++) ok = ok && sample_valid(s->samples[i]);
  return ok;
}

bool pglPathSegmentStorageValidateSegments(PGLPathSegmentStorage s) {
  bool ok = true;
  for (size_t i = 0; i < s->numSegments; i++) ok = ok && segment_valid(s->segments[i]);
  return ok;
}

// PathSegmentDataStorage::propagateSamples (data/PathSegmentDataStorage.h:518-530)
void pglPathSegmentStoragePropagateSamples(PGLPathSegmentStorage s, PGLSampleStorage storage, bool guideDirectLight,
                                           bool useNEEMiWeights, bool rrAffectsDirectContribution) {
  prepare_host(s, useNEEMiWeights, guideDirectLight, rrAffectsDirectContribution);
  {
    std::lock_guard<std::mutex> lock(storage->mutex);
    for (size_t i = 0; i < s->numSamples; i++) {
      const PGLSampleData& d = s->samples[i];
      if (d.flags & PGL_SAMPLE_FLAG_INSIDE_VOLUME) storage->volume.push_back(d);
      else storage->surface.push_back(d);
    }
  }
  s->numSegments = 0;
  s->numSamples = 0;
}

// ---- EXTENSION: batched path segments on the GPU

PGLPathSegmentBatch pglNewPathSegmentBatch(void) PGL_CATCH_BEGIN { return (PGLPathSegmentBatch) new PathSampleBatch(); }
PGL_CATCH_END(nullptr)

void pglReleasePathSegmentBatch(PGLPathSegmentBatch batch) PGL_CATCH_BEGIN { delete (PathSampleBatch*)batch; }
PGL_CATCH_END_VOID

bool pglPathSegmentBatchPrepareSamplesDevice(PGLPathSegmentBatch batch, const PGLPathSegmentData* segments,
                                             size_t numSegments, const uint32_t* pathOffsets, uint32_t numPaths,
                                             float maxDistance,
                                             bool useNEEMiWeights, bool guideDirectLight,
                                             bool rrAffectsDirectContribution, PGLSampleData* surfaceOut,
                                             size_t surfaceCapacity, PGLSampleData* volumeOut, size_t volumeCapacity,
                                             size_t* numSurfaceOut, size_t* numVolumeOut) PGL_CATCH_BEGIN {
  THROW_IF_NULL(batch, "handle");
  if (numPaths > 0) {
    THROW_IF_NULL(segments, "segments");
    THROW_IF_NULL(pathOffsets, "path offsets");
  }
  if (numSegments > kMaxUpdateSamples) throw std::runtime_error("openpgl_amd: at most 2^30 path vertices per batch");
  const PathSampleOptions o{maxDistance, useNEEMiWeights, guideDirectLight, rrAffectsDirectContribution};
  size_t ns = 0, nv = 0;
  ((PathSampleBatch*)batch)
      ->prepare((const PathVertex*)segments, (uint32_t)numSegments, pathOffsets, numPaths, o, (PathSample*)surfaceOut,
                surfaceOut ? surfaceCapacity : 0, (PathSample*)volumeOut, volumeOut ? volumeCapacity : 0, ns, nv);
  if (numSurfaceOut) *numSurfaceOut = ns;
  if (numVolumeOut) *numVolumeOut = nv;
  return true;
}
PGL_CATCH_END(false)

// pglFieldUpdate (SurfaceVolumeField::updateField) with the records in
// device memory: split by EInsideVolume on the device as a sample storage
// would route them (SampleDataStorage.h:72-90), then one iteration of the
// surface and the volume field
void pglFieldUpdateDevice(PGLField field, const PGLSampleData* deviceSamples, size_t numSamples) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  if (numSamples > 0) THROW_IF_NULL(deviceSamples, "samples");
  if (numSamples > kMaxUpdateSamples) throw std::runtime_error("openpgl_amd: at most 2^30 samples per update");
  if (!field->paths) field->paths = new PathSampleBatch();
  const PathSample *surf = nullptr, *vol = nullptr;
  size_t ns = 0, nv = 0;
  field->paths->routeOwned((const PathSample*)deviceSamples, (uint32_t)numSamples, surf, ns, vol, nv);
  GuidingField& gf = *field->gf;
  hipEvent_t evS = nullptr, evV = nullptr;
  if (ns) {
    gf.surface.update(surf, (uint32_t)ns, gf.stream);
    evS = gf.surface.lastGatherEvent();
  }
  if (nv) {
    gf.volume.update(vol, (uint32_t)nv, gf.stream);
    evV = gf.volume.lastGatherEvent();
  }
  field->paths->setReleased(evS, evV);
  gf.iteration++;
}
PGL_CATCH_END_VOID

bool pglPathSegmentsPrepareSamplesDeviceNoAlloc(const PGLPathSegmentData* segments, size_t numSegments,
                                                const uint32_t* pathOffsets, uint32_t numPaths, float maxDistance,
                                                bool useNEEMiWeights, bool guideDirectLight,
                                                bool rrAffectsDirectContribution, PGLSampleData* surfaceOut,
                                                size_t surfaceCapacity, PGLSampleData* volumeOut,
                                                size_t volumeCapacity, size_t* numSurfaceOut, size_t* numVolumeOut,
                                                void* stream) PGL_CATCH_BEGIN {
  if (numSurfaceOut) *numSurfaceOut = 0;
  if (numVolumeOut) *numVolumeOut = 0;
  if (numPaths > 0) {
    THROW_IF_NULL(segments, "segments");
    THROW_IF_NULL(pathOffsets, "path offsets");
  }
  if (numSegments > kNoAllocMaxVertices)
    throw std::runtime_error("openpgl_amd: the allocation-free path batch takes at most " +
                             std::to_string(kNoAllocMaxVertices) + " vertices per call (got " +
                             std::to_string(numSegments) + ")");
  const PathSampleOptions o{maxDistance, useNEEMiWeights, guideDirectLight, rrAffectsDirectContribution};
  size_t ns = 0, nv = 0;
  path_prepare_noalloc((const PathVertex*)segments, (uint32_t)numSegments, pathOffsets, numPaths, o,
                       (PathSample*)surfaceOut, surfaceOut ? surfaceCapacity : 0, (PathSample*)volumeOut,
                       volumeOut ? volumeCapacity : 0, ns, nv, (hipStream_t)stream);
  if (numSurfaceOut) *numSurfaceOut = ns;
  if (numVolumeOut) *numVolumeOut = nv;
  return true;
}
PGL_CATCH_END(false)

// the samples of the paths train the field: SurfaceVolumeField::updateField
// with the surface and volume samples of all paths (one iteration)
void pglFieldUpdateFromPathSegmentsDevice(PGLField field, const PGLPathSegmentData* segments, size_t numSegments,
                                          const uint32_t* pathOffsets, uint32_t numPaths, float maxDistance,
                                          bool useNEEMiWeights, bool guideDirectLight,
                                          bool rrAffectsDirectContribution) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  if (numPaths > 0) {
    THROW_IF_NULL(segments, "segments");
    THROW_IF_NULL(pathOffsets, "path offsets");
  }
  if (numSegments > kMaxUpdateSamples) throw std::runtime_error("openpgl_amd: at most 2^30 path vertices per update");
  if (!field->paths) field->paths = new PathSampleBatch();
  const PathSampleOptions o{maxDistance, useNEEMiWeights, guideDirectLight, rrAffectsDirectContribution};
  const PathSample *surf = nullptr, *vol = nullptr;
  size_t ns = 0, nv = 0;
  field->paths->prepareOwned((const PathVertex*)segments, (uint32_t)numSegments, pathOffsets, numPaths, o, surf, ns,
                             vol, nv);
  if (ns > kMaxUpdateSamples || nv > kMaxUpdateSamples)
    throw std::runtime_error("openpgl_amd: at most 2^30 samples per update");
  GuidingField& gf = *field->gf;
  hipEvent_t evS = nullptr, evV = nullptr;
  if (ns) {
    gf.surface.update(surf, (uint32_t)ns, gf.stream);
    evS = gf.surface.lastGatherEvent();
  }
  if (nv) {
    gf.volume.update(vol, (uint32_t)nv, gf.stream);
    evV = gf.volume.lastGatherEvent();
  }
  field->paths->setReleased(evS, evV);
  gf.iteration++;
}
PGL_CATCH_END_VOID

}  // extern "C"
