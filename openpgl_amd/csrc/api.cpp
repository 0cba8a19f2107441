// C ABI of libopenpgl_amd.so: the Open PGL 0.6.0 pgl* entry points
// (openpgl/api/api.cpp in the reference) backed by the MI355X field.
//
// Error behaviour follows the reference: creation, update, reset, release and
// file functions catch exceptions, print them to stdout and return
// nullptr/false/0 (api.cpp:33-72).  Field queries, exports and statistics
// do the same here, since they may have to report a device error or an
// update that failed (the field then refuses every use until
// pglFieldReset); the scalar distribution calls cannot fail.
// Scalar Init/Sample/PDF calls are served from a host mirror of the device
// field (refreshed lazily after each update); batched queries run on the GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/openpgl_amd/openpgl.h"
#include "api_impl.h"
#include "field.h"
#include "region_batch.h"
#include "server.h"

using namespace pglgpu;

struct HostDist {
  int kind = 0;
  const KindField* field = nullptr;
  const RegionState* region = nullptr;
  uint32_t id = 0xffffffffu;
  int num = 0;
  int K = 0;
  float mw[4] = {0, 0, 0, 0};
  float w[4][32], kappa[4][32], mx[4][32], my[4][32], mz[4][32], norm[4][32], em2k[4][32];
  float lw[32], lkappa[32], lmx[32], lmy[32], lmz[32], lnorm[32], lem2k[32];
  float integral = 1.f;
};
struct PGLSurfaceSamplingDistributionImpl : HostDist {};
struct PGLVolumeSamplingDistributionImpl : HostDist {};

struct PGLFieldStatisticsImpl {
  KindField::Statistics st;
};

static PGLString make_string(const std::string& s) {
  PGLString r;
  r.m_size = s.size();
  r.m_str = new char[s.size() + 1];
  memcpy(r.m_str, s.c_str(), s.size() + 1);
  return r;
}

// ---- host distribution (VMMSurfaceSamplingDistribution / VMMVolumeSamplingDistribution)
static bool dist_init(HostDist* d, const KindField& f, hipStream_t s, const float pos[3], float* sample1D) {
  f.syncMirror(s);
  uint32_t id = 0xffffffffu;
  float u = sample1D ? *sample1D : -1.f;
  const int r = f.lookupRegion(pos, &u, id);
  if (sample1D) *sample1D = u;
  if (r < 0) return false;
  const RegionState& S = f.mirrorRegion((uint32_t)r);
  if (!S.valid) return false;
  d->field = &f;
  d->region = &S;
  d->id = id;
  const VMMParams& V = S.vmm;
  d->K = V.K;
  const float sx = V.pivot[0] - pos[0], sy = V.pivot[1] - pos[1], sz = V.pivot[2] - pos[2];
  for (int k = 0; k < V.K; k++) {
    float dist;
    shift_comp(V, k, sx, sy, sz, d->lmx[k], d->lmy[k], d->lmz[k], dist, f.parallaxAware());
    d->lw[k] = V.w[k];
    d->lkappa[k] = V.kappa[k];
    d->lnorm[k] = V.norm[k];
    d->lem2k[k] = V.em2k[k];
    d->w[0][k] = d->lw[k]; d->kappa[0][k] = d->lkappa[k]; d->mx[0][k] = d->lmx[k]; d->my[0][k] = d->lmy[k];
    d->mz[0][k] = d->lmz[k]; d->norm[0][k] = d->lnorm[k]; d->em2k[0][k] = d->lem2k[k];
  }
  d->num = 1;
  d->mw[0] = 1.f;
  d->integral = 1.f;
  return true;
}

// PAVMM::product (ParallaxAwareVonMisesFisherMixture.h:609-671) of mixture m
// with one lobe: of the incident-radiance mixture copied into slot m
// (fromLi, the volume distribution's HG lobes, VMMVolumeSamplingDistribution.h:64)
// or of slot m itself in place (the surface distribution's cosine product,
// VMMSurfaceSamplingDistribution.h:52-60, so repeated products compound).
static float dist_product(HostDist* d, int m, bool fromLi, float weight, float nx, float ny, float nz, float kn,
                          float nn) {
  float integral = 0.f;
  for (int k = 0; k < d->K; k++) {
    const ProdComp p =
        fromLi ? product_comp(d->lw[k], d->lkappa[k], d->lmx[k], d->lmy[k], d->lmz[k], d->lnorm[k], weight, nx, ny, nz,
                              kn, nn)
               : product_comp(d->w[m][k], d->kappa[m][k], d->mx[m][k], d->my[m][k], d->mz[m][k], d->norm[m][k],
                              weight, nx, ny, nz, kn, nn);
    d->w[m][k] = p.scale; d->kappa[m][k] = p.kappa; d->mx[m][k] = p.mx; d->my[m][k] = p.my; d->mz[m][k] = p.mz;
    d->norm[m][k] = p.norm; d->em2k[m][k] = p.em2k;
    integral += p.scale;
  }
  for (int k = 0; k < d->K; k++) d->w[m][k] /= integral;
  return integral;
}

static void dist_sample(const HostDist* d, float ux, float uy, float out[3]) {
  float acc = 0.f;
  int m = 0;
  for (; m < d->num - 1; ++m) {
    const float nw = acc + d->mw[m];
    if (nw > ux) break;
    acc = nw;
  }
  const float u0 = (ux - acc) / d->mw[m];
  float sum = 0.f, cdf = 0.f;
  int k = 0;
  for (;; k++) {
    cdf = d->w[m][k];
    if (sum + cdf >= uy || k + 1 >= d->K) break;
    sum += cdf;
  }
  const float us = std::min(1.f - 1.19209290e-7f, (uy - sum) / cdf);
  sample_comp(d->kappa[m][k], d->em2k[m][k], d->mx[m][k], d->my[m][k], d->mz[m][k], u0, us, out);
}

static float dist_pdf(const HostDist* d, const float dir[3]) {
  float pdf = 0.f;
  for (int m = 0; m < d->num; m++) {
    float p = 0.f;
    for (int k = 0; k < d->K; k++)
      p += eval_comp(d->w[m][k], d->norm[m][k], d->kappa[m][k], d->mx[m][k], d->my[m][k], d->mz[m][k], dir[0], dir[1], dir[2]);
    pdf += d->mw[m] * p;
  }
  return pdf;
}

static float dist_pdf_li(const HostDist* d, const float dir[3]) {
  float p = 0.f;
  for (int k = 0; k < d->K; k++)
    p += eval_comp(d->lw[k], d->lnorm[k], d->lkappa[k], d->lmx[k], d->lmy[k], d->lmz[k], dir[0], dir[1], dir[2]);
  return p;
}

void* storage_alloc(size_t bytes) {
  // one header ahead of the samples records how the block was allocated
  constexpr size_t kHdr = 64;
  static const bool pinned = [] {
    int n = 0;
    const bool ok = hipGetDeviceCount(&n) == hipSuccess && n > 0;
    (void)hipGetLastError();
    return ok;
  }();
  char* raw = nullptr;
  uint32_t kind = 0;
  if (pinned && hipHostMalloc((void**)&raw, bytes + kHdr, hipHostMallocDefault) == hipSuccess) {
    kind = 1;
  } else {
    (void)hipGetLastError();
    raw = static_cast<char*>(std::malloc(bytes + kHdr));
    if (!raw) throw std::bad_alloc();
  }
  *reinterpret_cast<uint32_t*>(raw) = kind;
  return raw + kHdr;
}

void storage_free(void* p) noexcept {
  if (!p) return;
  char* raw = static_cast<char*>(p) - 64;
  if (*reinterpret_cast<uint32_t*>(raw)) (void)hipHostFree(raw);
  else std::free(raw);
}

static void field_update_host(PGLFieldImpl* F, KindField& kf, const SampleVec& samples) {
  if (samples.empty()) return;
  if (samples.size() > kMaxUpdateSamples) throw std::runtime_error("openpgl_amd: at most 2^30 samples per update");
  kf.updateHost(samples.data(), (uint32_t)samples.size(), F->gf->stream);
}

extern "C" {

void pglFieldArgumentsSetDefaults(PGLFieldArguments& a, enum PGL_SPATIAL_STRUCTURE_TYPE spatialType,
                                  enum PGL_DIRECTIONAL_DISTRIBUTION_TYPE directionalType, bool deterministic,
                                  size_t maxSamplesPerLeaf) {
  (void)spatialType;
  a.spatialStructureType = PGL_SPATIAL_STRUCTURE_KDTREE;
  PGLKDTreeArguments* kd = new PGLKDTreeArguments();
  kd->knnLookup = true; kd->minSamples = 100; kd->maxSamples = maxSamplesPerLeaf; kd->maxDepth = 32;
  a.spatialSturctureArguments = kd;
  a.deterministic = deterministic;
  a.debugArguments.fitRegions = true;
  a.directionalDistributionType = directionalType == PGL_DIRECTIONAL_DISTRIBUTION_QUADTREE
                                      ? PGL_DIRECTIONAL_DISTRIBUTION_QUADTREE
                                      : (directionalType == PGL_DIRECTIONAL_DISTRIBUTION_VMM ? PGL_DIRECTIONAL_DISTRIBUTION_VMM
                                                                                              : PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM);
  PGLVMMFactoryArguments* v = new PGLVMMFactoryArguments();
  v->initK = PGL_VMM_MAX_COMPONENTS / 2; v->initKappa = 0.5f; v->maxK = PGL_VMM_MAX_COMPONENTS;
  v->maxEMIterrations = 100; v->maxKappa = PGL_VMM_MAX_KAPPA; v->convergenceThreshold = 0.005f;
  v->weightPrior = 0.01f; v->meanCosinePriorStrength = 0.2f; v->meanCosinePrior = 0.0f;
  v->useSplitAndMerge = true; v->splittingThreshold = 0.5f; v->mergingThreshold = 0.025f; v->partialReFit = true;
  v->maxSplitItr = 1;
  v->minSamplesForSplitting = (int)(maxSamplesPerLeaf / 8);
  v->minSamplesForPartialRefitting = (int)(maxSamplesPerLeaf / 8);
  v->minSamplesForMerging = (int)(maxSamplesPerLeaf / 4);
  a.directionalDistributionArguments = v;
}

void pglReleaseString(PGLString str) { delete[] str.m_str; }

PGLDevice pglNewDevice(enum PGL_DEVICE_TYPE deviceType, size_t numThreads) PGL_CATCH_BEGIN {
  (void)numThreads;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw std::runtime_error("openpgl_amd: no HIP device available (this library runs on the GPU only)");
  PGL_HIP_CHECK(hipSetDevice(0));
  PGLDeviceImpl* d = new PGLDeviceImpl();
  d->type = deviceType;
  d->vectorWidth = deviceType == PGL_DEVICE_TYPE_CPU_4 ? 4 : (deviceType == PGL_DEVICE_TYPE_CPU_16 ? 16 : 8);
  return d;
}
PGL_CATCH_END(nullptr)

void pglReleaseDevice(PGLDevice device) PGL_CATCH_BEGIN {
  THROW_IF_NULL(device, "handle");
  delete device;
}
PGL_CATCH_END_VOID

static FactoryConfig factory_config(const PGLVMMFactoryArguments* v, int vectorWidth) {
  FactoryConfig fc{};
  fc.initK = (int)v->initK;
  fc.initKappa = v->initKappa;
  fc.maxEMIterations = (int)v->maxEMIterrations;
  fc.maxMeanCosine = 1.f / std::tanh(v->maxKappa) - 1.f / v->maxKappa;  // KappaToMeanCosine<float>
  fc.convergenceThreshold = v->convergenceThreshold;
  fc.weightPrior = v->weightPrior;
  fc.meanCosinePriorStrength = v->meanCosinePriorStrength;
  fc.meanCosinePrior = v->meanCosinePrior;
  fc.useSplitAndMerge = v->useSplitAndMerge;
  fc.splittingThreshold = v->splittingThreshold;
  fc.mergingThreshold = v->mergingThreshold;
  fc.partialReFit = v->partialReFit;
  fc.minSamplesForSplitting = v->minSamplesForSplitting;
  fc.minSamplesForPartialRefitting = v->minSamplesForPartialRefitting;
  fc.minSamplesForMerging = v->minSamplesForMerging;
  fc.decayOnSpatialSplit = 0.25f;
  fc.maxMeanCosineSplit = 1.f / std::tanh(32000.0f) - 1.f / 32000.0f;
  fc.vectorWidth = vectorWidth;
  if (fc.initK < 1 || fc.initK > 32) throw std::runtime_error("openpgl_amd: initK must be in [1, 32]");
  return fc;
}

// Device::newField for the parallax-aware VMM (device/Device.h:90-138)
static PGLFieldImpl* new_field(PGLDevice device, PGLFieldArguments args) {
  if (args.directionalDistributionType != PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM &&
      args.directionalDistributionType != PGL_DIRECTIONAL_DISTRIBUTION_VMM)
    throw std::runtime_error("openpgl_amd: only the VMM field types are implemented (no quadtree)");
  PGLKDTreeArguments* kd = (PGLKDTreeArguments*)args.spatialSturctureArguments;
  PGLVMMFactoryArguments* v = (PGLVMMFactoryArguments*)args.directionalDistributionArguments;
  THROW_IF_NULL(kd, "spatial structure arguments");
  THROW_IF_NULL(v, "directional distribution arguments");
  KdSettings ks;
  ks.knnLookup = kd->knnLookup;
  ks.minSamples = (uint32_t)kd->minSamples;
  ks.maxSamples = (uint32_t)kd->maxSamples;
  ks.maxDepth = (uint32_t)kd->maxDepth;
  FactoryConfig fc = factory_config(v, device->vectorWidth);
  // the plain VMM (Device.h:139-184): ParallaxAwareVonMisesFisherMixture<W, 32, false>
  fc.noParallax = args.directionalDistributionType == PGL_DIRECTIONAL_DISTRIBUTION_VMM ? 1 : 0;
  FieldFileSettings fs;
  fs.maxK = v->maxK;
  fs.maxKappa = v->maxKappa;
  fs.maxSplitItr = v->maxSplitItr;
  fs.fitRegions = args.debugArguments.fitRegions;
  // the reference deletes the argument structs it was handed (Device.h:109, 136)
  delete kd;
  delete v;
  PGLFieldImpl* F = new PGLFieldImpl();
  F->gf = new GuidingField(ks, fc, args.deterministic, device->vectorWidth);
  F->gf->surface.fileSettings = fs;
  F->gf->volume.fileSettings = fs;
  F->deviceType = device->type;
  return F;
}

static void release_field(PGLFieldImpl* field) {
  delete field->paths;
  field->gf->surface.drain();
  field->gf->volume.drain();
  if (field->gf->stream) (void)hipStreamDestroy(field->gf->stream);
  delete field->gf;
  delete field;
}

PGLField pglDeviceNewField(PGLDevice device, PGLFieldArguments args) PGL_CATCH_BEGIN {
  THROW_IF_NULL(device, "handle");
  return new_field(device, args);
}
PGL_CATCH_END(nullptr)

// Device::newFieldFromFile (device/Device.h:217-271): the file is read in the
// layout of the device's SIMD width
PGLField pglDeviceNewFieldFromFile(PGLDevice device, const char* fieldFileName) PGL_CATCH_BEGIN {
  THROW_IF_NULL(device, "handle");
  THROW_IF_NULL(fieldFileName, "string");
  PGLFieldArguments args;
  pglFieldArgumentsSetDefaults(args, PGL_SPATIAL_STRUCTURE_KDTREE, PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM,
                               false, 32000);
  PGLFieldImpl* F = new_field(device, args);
  try {
    load_field_file(*F->gf, fieldFileName, device->vectorWidth);
  } catch (...) {
    release_field(F);
    throw;
  }
  return F;
}
PGL_CATCH_END(nullptr)

void pglReleaseField(PGLField field) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  release_field(field);
}
PGL_CATCH_END_VOID

// SurfaceVolumeField::storeToFile (field/SurfaceVolumeField.h:197-213) in the
// layout of the field's device's SIMD width
bool pglFieldStoreToFile(PGLField field, const char* fieldFileName) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  THROW_IF_NULL(fieldFileName, "string");
  store_field_file(*field->gf, fieldFileName, field->gf->surface.vectorWidth());
  return true;
}
PGL_CATCH_END(false)

size_t pglFieldGetIteration(PGLField field) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  return field->gf->iteration;
}
PGL_CATCH_END(0)

void pglFieldSetSceneBounds(PGLField field, pgl_box3f bounds) PGL_CATCH_BEGIN {
  // SurfaceVolumeField::setSceneBounds: enlarge_by(1.01f) returns a new box and is discarded
  const float lo[3] = {bounds.lower.x, bounds.lower.y, bounds.lower.z};
  const float hi[3] = {bounds.upper.x, bounds.upper.y, bounds.upper.z};
  field->gf->surface.setSceneBounds(lo, hi);
  field->gf->volume.setSceneBounds(lo, hi);
}
PGL_CATCH_END_VOID

pgl_box3f pglFieldGetSceneBounds(PGLField field) PGL_CATCH_BEGIN {
  float a[3], b[3], c[3], d[3];
  field->gf->surface.sceneBounds(a, b);
  field->gf->volume.sceneBounds(c, d);
  pgl_box3f r;
  r.lower.x = std::min(a[0], c[0]); r.lower.y = std::min(a[1], c[1]); r.lower.z = std::min(a[2], c[2]);
  r.upper.x = std::max(b[0], d[0]); r.upper.y = std::max(b[1], d[1]); r.upper.z = std::max(b[2], d[2]);
  return r;
}
PGL_CATCH_END(pgl_box3f{})

void pglFieldUpdate(PGLField field, PGLSampleStorage storage) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  THROW_IF_NULL(storage, "handle");
  field_update_host(field, field->gf->surface, storage->surface);
  field_update_host(field, field->gf->volume, storage->volume);
  field->gf->iteration++;
}
PGL_CATCH_END_VOID

void pglFieldUpdateSurface(PGLField field, PGLSampleStorage storage) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  THROW_IF_NULL(storage, "handle");
  field_update_host(field, field->gf->surface, storage->surface);
  field->gf->iteration++;
}
PGL_CATCH_END_VOID

void pglFieldUpdateVolume(PGLField field, PGLSampleStorage storage) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  THROW_IF_NULL(storage, "handle");
  field_update_host(field, field->gf->volume, storage->volume);
  field->gf->iteration++;
}
PGL_CATCH_END_VOID

void pglFieldUpdateSurfaceDevice(PGLField field, const PGLSampleData* samples, size_t n) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  if (n > kMaxUpdateSamples) throw std::runtime_error("openpgl_amd: at most 2^30 samples per update");
  field->gf->surface.update(samples, (uint32_t)n, field->gf->stream);
  field->gf->iteration++;
}
PGL_CATCH_END_VOID

void pglFieldUpdateVolumeDevice(PGLField field, const PGLSampleData* samples, size_t n) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  if (n > kMaxUpdateSamples) throw std::runtime_error("openpgl_amd: at most 2^30 samples per update");
  field->gf->volume.update(samples, (uint32_t)n, field->gf->stream);
  field->gf->iteration++;
}
PGL_CATCH_END_VOID

void pglFieldReset(PGLField field) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  field->gf->iteration = 0;
  field->gf->surface.reset();
  field->gf->volume.reset();
}
PGL_CATCH_END_VOID

PGLSurfaceSamplingDistribution pglFieldNewSurfaceSamplingDistribution(PGLField field) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  PGLSurfaceSamplingDistribution d = new PGLSurfaceSamplingDistributionImpl();
  d->kind = 0;
  return d;
}
PGL_CATCH_END(nullptr)

bool pglFieldInitSurfaceSamplingDistribution(PGLField field, PGLSurfaceSamplingDistribution d, pgl_point3f position,
                                             float* sample1D) PGL_CATCH_BEGIN {
  const float p[3] = {position.x, position.y, position.z};
  return dist_init(d, field->gf->surface, field->gf->stream, p, sample1D);
}
PGL_CATCH_END(false)

PGLVolumeSamplingDistribution pglFieldNewVolumeSamplingDistribution(PGLField field) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  PGLVolumeSamplingDistribution d = new PGLVolumeSamplingDistributionImpl();
  d->kind = 1;
  return d;
}
PGL_CATCH_END(nullptr)

bool pglFieldInitVolumeSamplingDistribution(PGLField field, PGLVolumeSamplingDistribution d, pgl_point3f position,
                                            float* sample1D) PGL_CATCH_BEGIN {
  const float p[3] = {position.x, position.y, position.z};
  return dist_init(d, field->gf->volume, field->gf->stream, p, sample1D);
}
PGL_CATCH_END(false)

bool pglFieldValidate(PGLField field) PGL_CATCH_BEGIN {
  bool ok = true;
  if (field->gf->surface.initialized()) { field->gf->surface.syncMirror(field->gf->stream); ok = ok && field->gf->surface.validate(); }
  if (field->gf->volume.initialized()) { field->gf->volume.syncMirror(field->gf->stream); ok = ok && field->gf->volume.validate(); }
  return ok;
}
PGL_CATCH_END(false)

static bool kind_equal(const KindField& a, const KindField& b, hipStream_t sa, hipStream_t sb) {
  if (a.initialized() != b.initialized() || a.iteration() != b.iteration()) return false;
  if (!a.initialized()) return true;
  a.syncMirror(sa);
  b.syncMirror(sb);
  std::vector<KDNode> na, nb;
  std::vector<uint32_t> oa, ob;
  a.canonicalOrder(na, oa);
  b.canonicalOrder(nb, ob);
  if (na.size() != nb.size()) return false;
  for (size_t i = 0; i < na.size(); i++)
    if (na[i].pivot != nb[i].pivot || na[i].bits != nb[i].bits) return false;
  for (size_t c = 0; c < oa.size(); c++) {
    PGLRegionRecordHost ra, rb;
    a.exportRegion(oa[c], &ra);
    b.exportRegion(ob[c], &rb);
    // Region::operator== (spatial/Region.h:106-129) ignores the valid flag;
    // the figures of the last update are not part of a region's state
    for (PGLRegionRecordHost* r : {&ra, &rb}) {
      r->valid = 0;
      r->numSamplesThisUpdate = r->emIterations = r->partialEmIterations = r->numSplits = r->numMerges = 0;
    }
    if (memcmp(&ra, &rb, sizeof(ra)) != 0) return false;
  }
  return true;
}

bool pglFieldCompare(PGLField fieldA, PGLField fieldB) PGL_CATCH_BEGIN {  // Field::operator== after the canonical reorder
  return kind_equal(fieldA->gf->surface, fieldB->gf->surface, fieldA->gf->stream, fieldB->gf->stream) &&
         kind_equal(fieldA->gf->volume, fieldB->gf->volume, fieldA->gf->stream, fieldB->gf->stream) &&
         fieldA->gf->iteration == fieldB->gf->iteration;
}
PGL_CATCH_END(false)

static PGLFieldStatistics make_stats(const KindField& k, hipStream_t s) {
  PGLFieldStatistics r = new PGLFieldStatisticsImpl();
  try {
    r->st = k.statistics(s);
  } catch (...) {
    delete r;
    throw;
  }
  return r;
}
PGLFieldStatistics pglFieldGetSurfaceStatistics(PGLField field) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  return make_stats(field->gf->surface, field->gf->stream);
}
PGL_CATCH_END(nullptr)
PGLFieldStatistics pglFieldGetVolumeStatistics(PGLField field) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  return make_stats(field->gf->volume, field->gf->stream);
}
PGL_CATCH_END(nullptr)
void pglReleaseFieldStatistics(PGLFieldStatistics s) { delete s; }

// FieldStatistics::toString / headerCSVString / toCSVString with
// KDTreeStatistics and DirectionalDistributionStatistics
// (field/FieldStatistics.h:34-97, KDTreeStatistics.h, DirectionalDistributionStatistics.h)
PGLString pglFieldStatisticsToString(PGLFieldStatistics s) PGL_CATCH_BEGIN {
  THROW_IF_NULL(s, "handle");
  const KindField::Statistics& t = s->st;
  std::ostringstream o;
  o << "FieldStatistics:\n";
  o << "\tnumCacheRegions = " << t.numRegions << "\n";
  o << "\tnumCacheRegionsReserved = " << t.regionsReserved << "\n";
  o << "\tsizePerCacheRegions = " << t.sizePerRegion << " bs\n";
  o << "\tsizeAllCacheRegionsUsed = " << float(t.sizePerRegion * t.numRegions) / 1024 / 1024 << " Mbs\n";
  o << "\tsizeAllCacheRegionsReserved = " << float(t.sizePerRegion * t.regionsReserved) / 1024 / 1024 << " Mbs\n";
  o << "\ttimeUpdate = " << t.timeUpdate << " ms\n";
  o << "\ttimeCopySamples = " << t.timeCopySamples << " ms\n";
  o << "\ttimeSpatialStructureUpdate = " << t.timeSpatial << " ms\n";
  o << "\ttimeDirectionalDistriubtionUpdate= " << t.timeDirectional << " ms\n";
  o << "KDTreeStatistics: \n";
  o << "\tnumberOfNodes            = " << t.kdNodes << "\n";
  o << "\tnumberOfReservedNodes    = " << t.kdNodesReserved << "\n";
  o << "\tmaxDepth                 = " << t.kdMaxDepth << "\n";
  o << "\tsizePerNode              = " << t.kdSizePerNode << " bs\n";
  o << "\tsizeAllNodesUsed         = " << float(t.kdNodes * t.kdSizePerNode) / 1024 << " kbs\n";
  o << "\tsizeAllNodesReserved     = " << float(t.kdNodesReserved * t.kdSizePerNode) / 1024 << " kbs\n";
  o << "DirectionalDistributionStatistics: \n";
  o << "\tsizePerDistribution              = " << t.sizePerDistribution << " bs\n";
  o << "\tminNumberOfComponents            = " << t.minK << "\n";
  o << "\tmaxNumberOfComponents            = " << t.maxK << "\n";
  o << "\taverageNumberOfComponents        = " << t.averageK << "\n";
  o << "\tsecondMomentNumberOfComponents   = " << t.secondMomentK << "\n";
  return make_string(o.str());
}
PGL_CATCH_END(PGLString{})
PGLString pglFieldStatisticsHeaderCSVString(PGLFieldStatistics s) PGL_CATCH_BEGIN {
  THROW_IF_NULL(s, "handle");
  return make_string(
      "FieldStatistics: , numCacheRegions , numCacheRegionsReserved , sizePerCacheRegions(bs) , "
      "sizeAllCacheRegionsUsed(Mbs) , sizeAllCacheRegionsReserved(Mbs) , timeUpdate(ms) , timeCopySamples(ms) , "
      "timeSpatialStructureUpdate(ms) , timeDirectionalDistriubtionUpdate(ms) , KDTreeStatistics: , numberOfNodes , "
      "numberOfReservedNodes , maxDepth , sizePerNode , sizeAllNodesUsed , sizeAllNodesReserved , "
      "DirectionalDistributionStatistics: , sizePerDistribution , minNumberOfComponents , maxNumberOfComponents , "
      "averageNumberOfComponents , secondMomentNumberOfComponents , ");
}
PGL_CATCH_END(PGLString{})
PGLString pglFieldStatisticsToCSVString(PGLFieldStatistics s) PGL_CATCH_BEGIN {
  THROW_IF_NULL(s, "handle");
  const KindField::Statistics& t = s->st;
  std::ostringstream o;
  o << "  , " << t.numRegions << " , " << t.regionsReserved << " , " << t.sizePerRegion << " , "
    << float(t.sizePerRegion * t.numRegions) / 1024 / 1024 << " , "
    << float(t.sizePerRegion * t.regionsReserved) / 1024 / 1024 << " , " << t.timeUpdate << " , "
    << t.timeCopySamples << " , " << t.timeSpatial << " , " << t.timeDirectional << " , ";
  o << "  , " << t.kdNodes << " , " << t.kdNodesReserved << " , " << t.kdMaxDepth << " , " << t.kdSizePerNode
    << " , " << t.kdNodes * t.kdSizePerNode << " , " << t.kdNodesReserved * t.kdSizePerNode << " , ";
  o << "  , " << t.sizePerDistribution << " , " << t.minK << " , " << t.maxK << " , " << t.averageK << " , "
    << t.secondMomentK << " , ";
  return make_string(o.str());
}
PGL_CATCH_END(PGLString{})

// EXTENSION (testing): the figures of FieldStatistics as numbers
void pglFieldStatisticsGetValues(PGLFieldStatistics s, double* out, size_t count) PGL_CATCH_BEGIN {
  THROW_IF_NULL(s, "handle");
  const KindField::Statistics& t = s->st;
  const double v[16] = {(double)t.numRegions, (double)t.regionsReserved, (double)t.sizePerRegion,
                        t.timeUpdate, t.timeCopySamples, t.timeSpatial, t.timeDirectional,
                        (double)t.kdMaxDepth, (double)t.kdNodes, (double)t.kdNodesReserved,
                        (double)t.kdSizePerNode, (double)t.sizePerDistribution, t.minK, t.maxK, t.averageK,
                        t.secondMomentK};
  for (size_t j = 0; j < count && j < 16; j++) out[j] = v[j];
}
PGL_CATCH_END_VOID

// ---- sample storage
PGLSampleStorage pglNewSampleStorage(void) PGL_CATCH_BEGIN { return new PGLSampleStorageImpl(); }
PGL_CATCH_END(nullptr)

static const char kStorageHeader[] = "OPENPGL_0.6.0_SAMPLE_STORAGE";

PGLSampleStorage pglNewSampleStorageFromFile(const char* name) PGL_CATCH_BEGIN {  // SampleDataStorage.h:30-70, 196-216
  THROW_IF_NULL(name, "string");
  std::ifstream is(name, std::ios::binary);
  if (!is.is_open()) throw std::runtime_error("error: couldn't open file");
  char buf[256];
  is.read(buf, sizeof(kStorageHeader));
  if (!is) throw std::runtime_error("error: invalid file header");
  PGLSampleStorageImpl* s = new PGLSampleStorageImpl();
  for (int kind = 0; kind < 2; kind++) {
    size_t n = 0;
    is.read((char*)&n, sizeof(size_t));
    auto& v = kind == 0 ? s->surface : s->volume;
    v.resize(n);
    if (n) is.read((char*)v.data(), n * sizeof(PGLSampleData));
  }
  return s;
}
PGL_CATCH_END(nullptr)

void pglReleaseSampleStorage(PGLSampleStorage s) PGL_CATCH_BEGIN {
  THROW_IF_NULL(s, "handle");
  delete s;
}
PGL_CATCH_END_VOID

bool pglSampleStorageStoreToFile(PGLSampleStorage s, const char* name) PGL_CATCH_BEGIN {  // SampleDataStorage.h:176-194
  THROW_IF_NULL(s, "handle");
  THROW_IF_NULL(name, "string");
  std::ofstream os(name, std::ios::binary);
  if (!os.is_open()) throw std::runtime_error("error: couldn't open file!");
  os.write(kStorageHeader, sizeof(kStorageHeader));
  for (int kind = 0; kind < 2; kind++) {
    const auto& v = kind == 0 ? s->surface : s->volume;
    const size_t n = v.size();
    os.write((const char*)&n, sizeof(size_t));
    if (n) os.write((const char*)v.data(), n * sizeof(PGLSampleData));
  }
  return true;
}
PGL_CATCH_END(false)

void pglSampleStorageAddSample(PGLSampleStorage s, PGLSampleData& sample) {  // SampleDataStorage.h:72-80
  std::lock_guard<std::mutex> l(s->mutex);
  if (sample.flags & PGL_SAMPLE_FLAG_INSIDE_VOLUME) s->volume.push_back(sample);
  else s->surface.push_back(sample);
}

void pglSampleStorageAddSamples(PGLSampleStorage s, const PGLSampleData* samples, size_t n) {  // :82-90
  std::lock_guard<std::mutex> l(s->mutex);
  size_t nv = 0;
  for (size_t i = 0; i < n; i++) nv += (samples[i].flags & PGL_SAMPLE_FLAG_INSIDE_VOLUME) ? 1 : 0;
  storage_grow(s->surface, n - nv);
  storage_grow(s->volume, nv);
  for (size_t i = 0; i < n; i++) {
    if (samples[i].flags & PGL_SAMPLE_FLAG_INSIDE_VOLUME) s->volume.push_back(samples[i]);
    else s->surface.push_back(samples[i]);
  }
}

void pglSampleStorageReserve(PGLSampleStorage s, const size_t a, const size_t b) {
  std::lock_guard<std::mutex> l(s->mutex);
  s->surface.reserve(a);
  s->volume.reserve(b);
}
void pglSampleStorageClear(PGLSampleStorage s) { std::lock_guard<std::mutex> l(s->mutex); s->surface.clear(); s->volume.clear(); }
void pglSampleStorageClearSurface(PGLSampleStorage s) { std::lock_guard<std::mutex> l(s->mutex); s->surface.clear(); }
void pglSampleStorageClearVolume(PGLSampleStorage s) { std::lock_guard<std::mutex> l(s->mutex); s->volume.clear(); }
size_t pglSampleStorageGetSizeSurface(PGLSampleStorage s) { return s->surface.size(); }
size_t pglSampleStorageGetSizeVolume(PGLSampleStorage s) { return s->volume.size(); }
PGLSampleData pglSampleStorageGetSampleSurface(PGLSampleStorage s, const int idx) { return s->surface[idx]; }
PGLSampleData pglSampleStorageGetSampleVolume(PGLSampleStorage s, const int idx) { return s->volume[idx]; }

static bool sample_valid(const PGLSampleData& d) {  // SampleData.h:25-46
  const float v[9] = {d.position.x, d.position.y, d.position.z, d.direction.x, d.direction.y, d.direction.z, d.weight, d.pdf, d.distance};
  for (float x : v) if (!std::isfinite(x)) return false;
  return d.weight >= 0.f && d.pdf > 0.f && d.distance > 0.f;
}
bool pglSampleStorageValidate(PGLSampleStorage s) {
  for (const auto& d : s->surface) if (!sample_valid(d)) return false;
  for (const auto& d : s->volume) if (!sample_valid(d)) return false;
  return true;
}
bool pglSampleStorageCompare(PGLSampleStorage a, PGLSampleStorage b) {
  if (a->surface.size() != b->surface.size() || a->volume.size() != b->volume.size()) return false;
  return memcmp(a->surface.data(), b->surface.data(), a->surface.size() * sizeof(PGLSampleData)) == 0 &&
         memcmp(a->volume.data(), b->volume.data(), a->volume.size() * sizeof(PGLSampleData)) == 0;
}

// ---- surface distribution
void pglReleaseSurfaceSamplingDistribution(PGLSurfaceSamplingDistribution d) { delete d; }
void pglSurfaceSamplingDistributionApplyCosineProduct(PGLSurfaceSamplingDistribution d, pgl_vec3f n) {
  if (d->num > 0) d->integral = dist_product(d, 0, false, 1.f, n.x, n.y, n.z, kCosineKappa, cosine_normalization());
}
bool pglSurfaceSamplingDistributionSupportsApplyCosineProduct(PGLSurfaceSamplingDistribution) { return true; }
pgl_vec3f pglSurfaceSamplingDistributionSample(PGLSurfaceSamplingDistribution d, pgl_point2f u) {
  float o[3];
  dist_sample(d, u.x, u.y, o);
  return pgl_vec3f{o[0], o[1], o[2]};
}
float pglSurfaceSamplingDistributionPDF(PGLSurfaceSamplingDistribution d, pgl_vec3f dir) {
  const float v[3] = {dir.x, dir.y, dir.z};
  return dist_pdf(d, v);
}
float pglSurfaceSamplingDistributionSamplePDF(PGLSurfaceSamplingDistribution d, pgl_point2f u, pgl_vec3f& dir) {
  float o[3];
  dist_sample(d, u.x, u.y, o);
  dir = pgl_vec3f{o[0], o[1], o[2]};
  return dist_pdf(d, o);
}
float pglSurfaceSamplingDistributionIncomingRadiancePDF(PGLSurfaceSamplingDistribution d, pgl_vec3f dir) {
  const float v[3] = {dir.x, dir.y, dir.z};
  return dist_pdf_li(d, v);
}
uint32_t pglSurfaceSamplingDistributionGetId(PGLSurfaceSamplingDistribution d) { return d->id; }
bool pglSurfaceSamplingDistributionValidate(PGLSurfaceSamplingDistribution d) { return d->num > 0; }
void pglSurfaceSamplingDistributionClear(PGLSurfaceSamplingDistribution d) { d->num = 0; }
PGLRegion pglSurfaceSamplingGetRegion(PGLSurfaceSamplingDistribution d) { return (PGLRegion)d->region; }

// ---- volume distribution
void pglReleaseVolumeSamplingDistribution(PGLVolumeSamplingDistribution d) { delete d; }
pgl_vec3f pglVolumeSamplingDistributionSample(PGLVolumeSamplingDistribution d, pgl_point2f u) {
  float o[3];
  dist_sample(d, u.x, u.y, o);
  return pgl_vec3f{o[0], o[1], o[2]};
}
float pglVolumeSamplingDistributionPDF(PGLVolumeSamplingDistribution d, pgl_vec3f dir) {
  const float v[3] = {dir.x, dir.y, dir.z};
  return dist_pdf(d, v);
}
float pglVolumeSamplingDistributionSamplePDF(PGLVolumeSamplingDistribution d, pgl_point2f u, pgl_vec3f& dir) {
  float o[3];
  dist_sample(d, u.x, u.y, o);
  dir = pgl_vec3f{o[0], o[1], o[2]};
  return dist_pdf(d, o);
}
float pglVolumeSamplingDistributionIncomingRadiancePDF(PGLVolumeSamplingDistribution d, pgl_vec3f dir) {
  const float v[3] = {dir.x, dir.y, dir.z};
  return dist_pdf_li(d, v);
}
uint32_t pglVolumeSamplingDistributionGetId(PGLVolumeSamplingDistribution d) { return d->id; }
bool pglVolumeSamplingDistributionValidate(PGLVolumeSamplingDistribution d) { return d->num > 0; }
void pglVolumeSamplingDistributionClear(PGLVolumeSamplingDistribution d) { d->num = 0; }
void pglVolumeSamplingDistributionApplySingleLobeHenyeyGreensteinProduct(PGLVolumeSamplingDistribution d, pgl_vec3f dir,
                                                                         const float meanCosine) {
  // VMMVolumeSamplingDistribution.h:54-80
  const HGRep& rep = host_hg_table()[hg_index(meanCosine)];
  float sum = 0.f;
  const int KP = (int)rep.K;
  for (int i = 0; i < KP; i++) {
    const float sg = meanCosine * rep.mc[i] > 0.f ? 1.f : -1.f;
    d->mw[i] = dist_product(d, i, true, rep.w[i], dir.x * sg, dir.y * sg, dir.z * sg, rep.kappa[i], rep.norm[i]);
    sum += d->mw[i];
  }
  for (int i = 0; i < KP; i++) d->mw[i] /= sum;
  d->num = KP;
  d->integral = sum;
}
bool pglVolumeSamplingDistributionSupportsApplySingleLobeHenyeyGreensteinProduct(PGLVolumeSamplingDistribution) {
  return true;
}
PGLRegion pglVolumeSamplingGetRegion(PGLVolumeSamplingDistribution d) { return (PGLRegion)d->region; }

// ---- batched queries
void pglFieldSampleBatchDevice(PGLField field, size_t n, const float* pos, const float* nrm, const float* g,
                               const float* u2, const float* u1, const uint8_t* isVolume, float* dirOut, float* pdfOut,
                               uint32_t* idOut, float* integralOut, void* stream) PGL_CATCH_BEGIN {
  const QueryField sf = field->gf->surface.queryView();
  const QueryField vf = field->gf->volume.queryView();
  launch_query(sf, vf, field->gf->dHG.as<HGRep>(), (uint32_t)n, pos, nrm, g, u2, u1, isVolume, dirOut, pdfOut, idOut,
               integralOut, (hipStream_t)stream);
}
PGL_CATCH_END_VOID

void pglFieldSampleBatch(PGLField field, size_t n, const float* pos, const float* nrm, const float* g, const float* u2,
                         const float* u1, const uint8_t* isVolume, float* dirOut, float* pdfOut, uint32_t* idOut,
                         float* integralOut) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  if (n == 0) return;
  // staging layout: pos 3n, nrm 3n, g n, u2 2n, u1 n, dir 3n, pdf n, id n, integral n (floats), vol n bytes
  const size_t nf = 16 * n;
  DevBuf& b = field->gf->dQuery;
  b.ensure(nf * 4 + n + 64);
  float* base = b.as<float>();
  float *dPos = base, *dNrm = base + 3 * n, *dG = base + 6 * n, *dU2 = base + 7 * n, *dU1 = base + 9 * n;
  float *dDir = base + 10 * n, *dPdf = base + 13 * n, *dInt = base + 15 * n;
  uint32_t* dId = (uint32_t*)(base + 14 * n);
  uint8_t* dVol = (uint8_t*)(base + nf);
  hipStream_t s = field->gf->stream;
  PGL_HIP_CHECK(hipMemcpyAsync(dPos, pos, 12 * n, hipMemcpyHostToDevice, s));
  PGL_HIP_CHECK(hipMemcpyAsync(dNrm, nrm, 12 * n, hipMemcpyHostToDevice, s));
  if (g) PGL_HIP_CHECK(hipMemcpyAsync(dG, g, 4 * n, hipMemcpyHostToDevice, s));
  PGL_HIP_CHECK(hipMemcpyAsync(dU2, u2, 8 * n, hipMemcpyHostToDevice, s));
  if (u1) PGL_HIP_CHECK(hipMemcpyAsync(dU1, u1, 4 * n, hipMemcpyHostToDevice, s));
  if (isVolume) PGL_HIP_CHECK(hipMemcpyAsync(dVol, isVolume, n, hipMemcpyHostToDevice, s));
  pglFieldSampleBatchDevice(field, n, dPos, dNrm, g ? dG : nullptr, dU2, u1 ? dU1 : nullptr, isVolume ? dVol : nullptr,
                            dDir, dPdf, dId, dInt, s);
  PGL_HIP_CHECK(hipGetLastError());
  PGL_HIP_CHECK(hipMemcpyAsync(dirOut, dDir, 12 * n, hipMemcpyDeviceToHost, s));
  PGL_HIP_CHECK(hipMemcpyAsync(pdfOut, dPdf, 4 * n, hipMemcpyDeviceToHost, s));
  if (idOut) PGL_HIP_CHECK(hipMemcpyAsync(idOut, dId, 4 * n, hipMemcpyDeviceToHost, s));
  if (integralOut) PGL_HIP_CHECK(hipMemcpyAsync(integralOut, dInt, 4 * n, hipMemcpyDeviceToHost, s));
  PGL_HIP_CHECK(hipStreamSynchronize(s));
}
PGL_CATCH_END_VOID

// ---- introspection
static KindField& kind_of(PGLField f, int kind) { return kind == 0 ? f->gf->surface : f->gf->volume; }

// both complete a queued update first (and so report a failed one)
uint32_t pglFieldGetNumNodes(PGLField field, int kind) PGL_CATCH_BEGIN {
  const KindField& k = kind_of(field, kind);
  k.finish();
  return (uint32_t)k.nodes().size();
}
PGL_CATCH_END(0)
uint32_t pglFieldGetNumRegions(PGLField field, int kind) PGL_CATCH_BEGIN {
  const KindField& k = kind_of(field, kind);
  k.finish();
  return k.numRegions();
}
PGL_CATCH_END(0)

void pglFieldExportNodes(PGLField field, int kind, float* pivots, uint32_t* dims, uint32_t* childOrRegion) PGL_CATCH_BEGIN {
  std::vector<KDNode> out;
  std::vector<uint32_t> order;
  kind_of(field, kind).canonicalOrder(out, order);
  for (size_t i = 0; i < out.size(); i++) { pivots[i] = out[i].pivot; dims[i] = out[i].dim(); childOrRegion[i] = out[i].idx(); }
}
PGL_CATCH_END_VOID

void pglFieldExportRegion(PGLField field, int kind, uint32_t canonicalIdx, PGLRegionRecord* out) PGL_CATCH_BEGIN {
  KindField& k = kind_of(field, kind);
  k.syncMirror(field->gf->stream);
  std::vector<KDNode> nodes;
  std::vector<uint32_t> order;
  k.canonicalOrder(nodes, order);
  static_assert(sizeof(PGLRegionRecord) == sizeof(PGLRegionRecordHost), "record layouts agree");
  k.exportRegion(order[canonicalIdx], (PGLRegionRecordHost*)out);
}
PGL_CATCH_END_VOID

void pglFieldSampleLeafIds(PGLField field, int kind, const float* positions, size_t n, uint32_t* ids) PGL_CATCH_BEGIN {
  KindField& k = kind_of(field, kind);
  std::vector<KDNode> nodes;
  std::vector<uint32_t> order;
  k.canonicalOrder(nodes, order);
  for (size_t i = 0; i < n; i++) {
    uint32_t j = 0;
    const float* p = positions + 3 * i;
    while (!nodes[j].isLeaf()) j = nodes[j].idx() + (p[nodes[j].dim()] >= nodes[j].pivot ? 1u : 0u);
    ids[i] = nodes[j].idx();
  }
}
PGL_CATCH_END_VOID

void pglFieldGetKNN(PGLField field, int kind, uint32_t* nb) PGL_CATCH_BEGIN {
  KindField& k = kind_of(field, kind);
  k.syncMirror(field->gf->stream);
  std::vector<KDNode> nodes;
  std::vector<uint32_t> order;
  k.canonicalOrder(nodes, order);
  std::vector<uint32_t> s2c(k.numRegions(), 0xffffffffu);
  for (size_t c = 0; c < order.size(); c++) s2c[order[c]] = (uint32_t)c;
  std::vector<uint32_t> ids;
  k.knnHost(ids);
  for (size_t c = 0; c < order.size(); c++)
    for (int i = 0; i < 8; i++) {
      const uint32_t id = ids.empty() ? 0xffffffffu : ids[order[c] * 8 + i];
      nb[c * 8 + i] = id == 0xffffffffu ? 0xffffffffu : s2c[id];
    }
}
PGL_CATCH_END_VOID

static void info_values(const UpdateInfo& i, double* o) {
  o[0] = i.wallMs; o[1] = i.spatialMs; o[2] = i.directionalMs; o[3] = i.knnMs; o[4] = i.gatherMs; o[5] = i.sortMs;
  o[6] = i.kdRounds; o[7] = i.regionLaunches;
  o[8] = i.estepEvals; o[9] = i.auxEvals; o[10] = i.emIterations; o[11] = i.samples; o[12] = i.splitEvals;
  o[13] = i.regions;
  o[14] = i.regionJobs; o[15] = i.regionItems; o[16] = i.maxTeam; o[17] = i.regionGrid;
  o[18] = i.regionKernelMs;
}

size_t pglFieldGetUpdateHistory(PGLField field, int kind, double* out, size_t maxEntries) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  KindField& k = kind_of(field, kind);
  const std::vector<UpdateInfo>& h = k.history();
  const size_t n = std::min(h.size(), maxEntries);
  for (size_t e = 0; e < n; e++) info_values(h[h.size() - n + e], out + 19 * e);
  return n;
}
PGL_CATCH_END(0)

void pglFieldGetLastUpdateInfo(PGLField field, int kind, double* out, size_t count) PGL_CATCH_BEGIN {
  KindField& k = kind_of(field, kind);
  const UpdateInfo& i = k.lastInfo();
  double o[19];
  info_values(i, o);
  for (size_t j = 0; j < count && j < 19; j++) out[j] = o[j];
}
PGL_CATCH_END_VOID

void pglFieldDebugFailNextUpdate(PGLField field, int kind) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  kind_of(field, kind).runner().failNext = true;
}
PGL_CATCH_END_VOID

void pglFieldGetRegionProfile(PGLField field, int kind, float* out) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "field");
  KindField& k = kind_of(field, kind);
  k.syncMirror(field->gf->stream);
  int dev = 0, rateKHz = 0;
  PGL_HIP_CHECK(hipGetDevice(&dev));
  PGL_HIP_CHECK(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, dev));
  const float toUs = rateKHz > 0 ? 1000.f / (float)rateKHz : 0.01f;
  for (uint32_t r = 0; r < k.numRegions(); r++) {
    const RegionState& S = k.mirrorRegion(r);
    for (int i = 0; i < 12; i++) {
      const bool count = i == 7 || i == 8 || i == 11;
      out[r * 12 + i] = count ? S.prof[i] : S.prof[i] * toUs;
    }
  }
}
PGL_CATCH_END_VOID

void pglFieldSetRegionSchedule(PGLField field, const PGLRegionSchedule* sc) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "field");
  THROW_IF_NULL(sc, "schedule");
  for (KindField* k : {&field->gf->surface, &field->gf->volume}) {
    RegionSchedule& r = k->schedule();
    r.perRegionLaunches = sc->perRegionLaunches;
    r.grid = sc->grid;
    r.balance = sc->balance;
    r.minChunksPerRank = sc->minChunksPerRank;
    r.maxTeam = sc->maxTeam;
    r.cooperative = sc->cooperative;
  }
}
PGL_CATCH_END_VOID

void pglFieldGetRegionSchedule(PGLField field, PGLRegionSchedule* sc) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "field");
  THROW_IF_NULL(sc, "schedule");
  const RegionSchedule& r = field->gf->surface.schedule();
  sc->perRegionLaunches = r.perRegionLaunches;
  sc->grid = r.grid;
  sc->balance = r.balance;
  sc->minChunksPerRank = r.minChunksPerRank;
  sc->maxTeam = r.maxTeam;
  sc->cooperative = r.cooperative;
}
PGL_CATCH_END_VOID


/* ------------------------------------------------------------------ batched regions (EXTENSION) */

struct PGLRegionBatchImpl {
  RegionBatch* b = nullptr;
  hipStream_t stream = nullptr;
};

static void to_sched(const PGLRegionSchedule* sc, RegionSchedule& r) {
  r.perRegionLaunches = sc->perRegionLaunches;
  r.grid = sc->grid;
  r.balance = sc->balance;
  r.minChunksPerRank = sc->minChunksPerRank;
  r.maxTeam = sc->maxTeam;
  r.cooperative = sc->cooperative;
}
static void from_sched(const RegionSchedule& r, PGLRegionSchedule* sc) {
  sc->perRegionLaunches = r.perRegionLaunches;
  sc->grid = r.grid;
  sc->balance = r.balance;
  sc->minChunksPerRank = r.minChunksPerRank;
  sc->maxTeam = r.maxTeam;
  sc->cooperative = r.cooperative;
}

PGLRegionBatch pglNewRegionBatch(PGLDevice device, PGLFieldArguments args) PGL_CATCH_BEGIN {
  THROW_IF_NULL(device, "handle");
  if (args.directionalDistributionType != PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM &&
      args.directionalDistributionType != PGL_DIRECTIONAL_DISTRIBUTION_VMM)
    throw std::runtime_error("openpgl_amd: only the VMM field types are implemented (no quadtree)");
  PGLKDTreeArguments* kd = (PGLKDTreeArguments*)args.spatialSturctureArguments;
  PGLVMMFactoryArguments* v = (PGLVMMFactoryArguments*)args.directionalDistributionArguments;
  THROW_IF_NULL(v, "directional distribution arguments");
  FactoryConfig fc = factory_config(v, device->vectorWidth);
  fc.noParallax = args.directionalDistributionType == PGL_DIRECTIONAL_DISTRIBUTION_VMM ? 1 : 0;
  delete kd;
  delete v;
  PGLRegionBatchImpl* B = new PGLRegionBatchImpl();
  PGL_HIP_CHECK(hipStreamCreate(&B->stream));
  B->b = new RegionBatch(fc, args.deterministic);
  return B;
}
PGL_CATCH_END(nullptr)

void pglReleaseRegionBatch(PGLRegionBatch batch) PGL_CATCH_BEGIN {
  THROW_IF_NULL(batch, "handle");
  delete batch->b;
  (void)hipStreamDestroy(batch->stream);
  delete batch;
}
PGL_CATCH_END_VOID

void pglRegionBatchStep(PGLRegionBatch batch, uint32_t numRegions, const PGLSampleData* deviceRecords,
                        const uint32_t* counts, const float* sampleStats7, const uint32_t* initK,
                        const int* splitFlags) PGL_CATCH_BEGIN {
  THROW_IF_NULL(batch, "handle");
  batch->b->step(numRegions, deviceRecords, counts, sampleStats7, initK, splitFlags, batch->stream);
}
PGL_CATCH_END_VOID

uint32_t pglRegionBatchGetNumRegions(PGLRegionBatch batch) PGL_CATCH_BEGIN { return batch->b->numRegions(); }
PGL_CATCH_END(0)

void pglRegionBatchExportRegion(PGLRegionBatch batch, uint32_t region, PGLRegionRecord* out) PGL_CATCH_BEGIN {
  THROW_IF_NULL(batch, "handle");
  THROW_IF_NULL(out, "record");
  static_assert(sizeof(PGLRegionRecord) == sizeof(PGLRegionRecordHost), "record layouts differ");
  batch->b->exportRegion(region, (PGLRegionRecordHost*)out, batch->stream);
}
PGL_CATCH_END_VOID

void pglRegionBatchSetSchedule(PGLRegionBatch batch, const PGLRegionSchedule* sc) PGL_CATCH_BEGIN {
  THROW_IF_NULL(batch, "handle");
  THROW_IF_NULL(sc, "schedule");
  to_sched(sc, batch->b->runner().schedule);
}
PGL_CATCH_END_VOID

void pglRegionBatchGetSchedule(PGLRegionBatch batch, PGLRegionSchedule* sc) PGL_CATCH_BEGIN {
  THROW_IF_NULL(batch, "handle");
  THROW_IF_NULL(sc, "schedule");
  from_sched(batch->b->runner().schedule, sc);
}
PGL_CATCH_END_VOID

size_t pglRegionScheduleCandidates(PGLRegionSchedule* out, size_t maxCount) {
  // search space of the autotuner: launch kind (ordinary or cooperative),
  // grid (the default two per CU, one per CU, or all that fit at once), team
  // balance, chunks per rank, largest team
  std::vector<PGLRegionSchedule> c;
  for (int coop = 0; coop < 2; coop++)
    for (uint32_t gridSel : {0u, 1u, 2u})
      for (float bal : {0.5f, 1.f, 2.f, 4.f})
        for (uint32_t mc : {1u, 2u, 4u})
          for (uint32_t mt : {16u, 64u}) c.push_back(PGLRegionSchedule{0, gridSel, bal, mc, mt, coop});
  if (out)
    for (size_t i = 0; i < c.size() && i < maxCount; i++) out[i] = c[i];
  return c.size();
}

int pglRegionBatchAutotune(PGLRegionBatch batch, const PGLRegionSchedule* candidates, size_t numCandidates, int reps,
                           float* msOut) PGL_CATCH_BEGIN {
  THROW_IF_NULL(batch, "handle");
  std::vector<PGLRegionSchedule> cs;
  if (candidates) {
    cs.assign(candidates, candidates + numCandidates);
  } else {
    cs.resize(pglRegionScheduleCandidates(nullptr, 0));
    pglRegionScheduleCandidates(cs.data(), cs.size());
  }
  // grid == 1 in a candidate means one workgroup per CU, grid == 2 all that fit at once
  const uint32_t resident = batch->b->runner().residentGrid();
  int cus = 0, dev = 0;
  PGL_HIP_CHECK(hipGetDevice(&dev));
  PGL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  std::vector<RegionSchedule> rs(cs.size());
  for (size_t i = 0; i < cs.size(); i++) {
    to_sched(&cs[i], rs[i]);
    if (rs[i].grid == 1) rs[i].grid = std::min<uint32_t>(resident, (uint32_t)cus);
    else if (rs[i].grid == 2) rs[i].grid = resident;
  }
  std::vector<float> ms;
  const int best = batch->b->autotune(rs, reps, ms, batch->stream);
  if (msOut)
    for (size_t i = 0; i < ms.size(); i++) msOut[i] = ms[i];
  return best;
}
PGL_CATCH_END(-1)

void pglRegionBatchGetLastInfo(PGLRegionBatch batch, double* out, size_t count) PGL_CATCH_BEGIN {
  THROW_IF_NULL(batch, "handle");
  const UpdateInfo& i = batch->b->lastInfo();
  double o[19] = {i.wallMs, 0.0, i.directionalMs, 0.0, 0.0, 0.0, 0.0, i.regionLaunches, i.estepEvals, i.auxEvals,
                  i.emIterations, i.samples, i.splitEvals, (double)batch->b->numRegions(), i.regionJobs,
                  i.regionItems, i.maxTeam, i.regionGrid, i.regionKernelMs};
  for (size_t j = 0; j < count && j < 19; j++) out[j] = o[j];
}
PGL_CATCH_END_VOID


/* ------------------------------------------------------------------ query server (EXTENSION) */

struct PGLServerImpl {
  QueryServer* s = nullptr;
};

PGLServer pglNewServer(PGLField field, const PGLServerConfig* cfg) PGL_CATCH_BEGIN {
  THROW_IF_NULL(field, "handle");
  ServerConfig c;
  if (cfg) {
    if (cfg->memoryBudgetBytes) c.memoryBudget = cfg->memoryBudgetBytes;
    if (cfg->slots) c.slots = cfg->slots;
    if (cfg->maxBatchQueries) c.maxBatchQueries = cfg->maxBatchQueries;
  }
  PGLServerImpl* S = new PGLServerImpl();
  S->s = new QueryServer(field->gf, c);
  return S;
}
PGL_CATCH_END(nullptr)

void pglReleaseServer(PGLServer server) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  delete server->s;
  delete server;
}
PGL_CATCH_END_VOID

uint64_t pglServerSubmit(PGLServer server, size_t n, const float* pos, const float* normalOrDir, const float* g,
                         const float* u2, const float* u1, const uint8_t* isVolume, float* dirOut, float* pdfOut,
                         uint32_t* idOut, float* integralOut) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  QueryRequest r;
  r.n = n; r.pos = pos; r.nrm = normalOrDir; r.g = g; r.u2 = u2; r.u1 = u1; r.isVolume = isVolume;
  r.dir = dirOut; r.pdf = pdfOut; r.id = idOut; r.integral = integralOut;
  return server->s->submit(r);
}
PGL_CATCH_END(0)

int pglServerPoll(PGLServer server, uint64_t request) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  return server->s->poll(request) ? 1 : 0;
}
PGL_CATCH_END(-1)

int pglServerWait(PGLServer server, uint64_t request, double timeoutMs) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  return server->s->wait(request, timeoutMs) ? 1 : 0;
}
PGL_CATCH_END(-1)

void pglServerDrain(PGLServer server) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  server->s->drain();
}
PGL_CATCH_END_VOID

int pglServerGetTiming(PGLServer server, uint64_t request, double* out3) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  return server->s->timing(request, out3[0], out3[1], out3[2]) ? 1 : 0;
}
PGL_CATCH_END(-1)

void pglServerForget(PGLServer server, uint64_t request) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  server->s->forget(request);
}
PGL_CATCH_END_VOID

void pglServerRefresh(PGLServer server) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  server->s->refresh();
}
PGL_CATCH_END_VOID

void pglServerGetStats(PGLServer server, double* out, size_t count) PGL_CATCH_BEGIN {
  THROW_IF_NULL(server, "handle");
  server->s->stats(out, count);
}
PGL_CATCH_END_VOID

}  // extern "C"
