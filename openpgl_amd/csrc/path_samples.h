// Training samples of one recorded path: the back-propagation of
// PathSegmentDataStorage::prepareSamples (data/PathSegmentDataStorage.h:160-315),
// shared by the host path-segment storage and the batched GPU kernels so that
// both produce the same records.
//
// Float semantics follow the reference's CPU_8 build: no contraction of
// separate operations, the vector dot product fused as
// fma(a.x, b.x, fma(a.y, b.y, a.z * b.z)) (embree madd under AVX2), correctly
// rounded division and square root.  Every file that includes this header
// compiles it with contraction off.
#pragma once
#include <math.h>

#include "common.h"

namespace pglgpu {

struct PathSampleOptions {
  float maxDistance;  // PathSegmentDataStorage::m_max_distance (1e6 by default)
  bool useNEEMiWeights;
  bool guideDirectLight;
  bool rrAffectsDirectContribution;
};

// 40-byte training record (PGLSampleData, include/openpgl/data.h:21-49)
struct PathSample {
  float px, py, pz, dx, dy, dz, weight, pdf, distance;
  uint32_t flags;
};

PGL_HD inline float path_fdot(float ax, float ay, float az, float bx, float by, float bz) {
  return fmaf(ax, bx, fmaf(ay, by, az * bz));
}

// One step of the backward walk at vertex c (next vertex nx): the sample
// distance of c and the length carried on to the previous vertex.
PGL_HD inline float path_step_distance(const PathVertex& c, const PathVertex& nx, float maxDistance,
                                       float& lastDistance) {
  // distance to the next vertex plus the length carried through specular chains
  const float ex = nx.px - c.px, ey = nx.py - c.py, ez = nx.pz - c.pz;
  const float currentDistance = sqrtf(path_fdot(ex, ey, ez, ex, ey, ez));
  const float distance = fminf(currentDistance + lastDistance, 2.0f * maxDistance);
  if (!c.isDelta && c.roughness >= 0.3f) {
    lastDistance = 0.0f;
  } else {
    lastDistance = distance;
    if (c.eta != 1.0f) {
      const float cosThetaI = path_fdot(c.nx, c.ny, c.nz, c.dox, c.doy, c.doz);
      const float cosThetaO = path_fdot(c.nx, c.ny, c.nz, c.dix, c.diy, c.diz);
      lastDistance *= fabsf(cosThetaI / (cosThetaO * c.eta));
    }
  }
  return distance;
}

// The sample of vertex i (sample distance given), or false when vertex i
// yields none: delta or near-specular vertex, no radiance brought back, or a
// non-positive distance.
PGL_HD inline bool path_vertex_record(const PathVertex* seg, int n, int i, float distance, const PathSampleOptions& o,
                                      PathSample& s) {
  const float minPdf = 0.1f;
  const float maxThroughput = 10.0f;
  const PathVertex& c = seg[i];
  if (c.isDelta || !(c.roughness > 0.01f)) return false;
  const float pdf = minPdf < c.pdfDirectionIn ? c.pdfDirectionIn : minPdf;
  // incident radiance carried back from the rest of the path
  float t[3] = {1.0f, 1.0f, 1.0f};
  float contrib[3] = {0.0f, 0.0f, 0.0f};
  float previousRR = 1.0f;
  for (int j = i + 1; j < n; ++j) {
    const PathVertex& prev = seg[j - 1];
    const PathVertex& v = seg[j];
    float ct[3];
    for (int k = 0; k < 3; k++) {
      t[k] = t[k] * prev.transmittanceWeight[k];
      ct[k] = t[k] < maxThroughput ? t[k] : maxThroughput;
      contrib[k] += ct[k] * v.scatteredContribution[k];
    }
    float direct[3] = {v.directContribution[0], v.directContribution[1], v.directContribution[2]};
    if (!o.rrAffectsDirectContribution)
      for (int k = 0; k < 3; k++) direct[k] *= previousRR;
    if (j == i + 1 && !o.useNEEMiWeights) {
      if (o.guideDirectLight)
        for (int k = 0; k < 3; k++) contrib[k] += ct[k] * direct[k];
    } else if (j > i + 1 || o.guideDirectLight) {
      for (int k = 0; k < 3; k++) contrib[k] += (ct[k] * v.miWeight) * direct[k];
    }
    for (int k = 0; k < 3; k++) t[k] = t[k] * v.scatteringWeight[k];
    const float rr = v.russianRouletteProbability;
    if (rr > 0.f) {
      for (int k = 0; k < 3; k++) t[k] /= rr;
    } else {
      t[0] = t[1] = t[2] = 0.0f;
    }
    previousRR = rr;
  }
  if (!(contrib[0] > 0.0f || contrib[1] > 0.0f || contrib[2] > 0.0f)) return false;
  if (!(distance > 0.0f)) return false;
  s.px = c.px;
  s.py = c.py;
  s.pz = c.pz;
  s.dx = c.dix;
  s.dy = c.diy;
  s.dz = c.diz;
  s.weight = ((contrib[0] + contrib[1]) + contrib[2]) / 3.0f / pdf;
  s.pdf = pdf;
  s.distance = distance;
  s.flags = c.volumeScatter ? 2u : 0u;  // EInsideVolume
  return true;
}

// Calls emit(const PathSample&) for every training sample of the path
// seg[0 .. n-1], in the reference's order (vertex n-2 down to 0); returns the
// number of samples.
template <class Emit>
PGL_HD inline uint32_t path_samples(const PathVertex* seg, int n, const PathSampleOptions& o, Emit&& emit) {
  uint32_t count = 0;
  float lastDistance = 0.0f;
  for (int i = n - 2; i >= 0; --i) {
    const float distance = path_step_distance(seg[i], seg[i + 1], o.maxDistance, lastDistance);
    PathSample s;
    if (path_vertex_record(seg, n, i, distance, o, s)) {
      emit(s);
      count++;
    }
  }
  return count;
}

// The sample of vertex i alone (0 <= i <= n-2): the same walk from vertex n-2
// down to i, so the record is the one path_samples gives for vertex i.
PGL_HD inline bool path_vertex_sample(const PathVertex* seg, int n, int i, const PathSampleOptions& o, PathSample& s) {
  float lastDistance = 0.0f;
  float distance = 0.0f;
  for (int k = n - 2; k >= i; --k) distance = path_step_distance(seg[k], seg[k + 1], o.maxDistance, lastDistance);
  return path_vertex_record(seg, n, i, distance, o, s);
}

// PathSegmentDataStorage::calculatePixelEstimate (data/PathSegmentDataStorage.h:317-393)
PGL_HD inline void path_pixel_estimate(const PathVertex* seg, int n, bool rrAffectsDirectContribution, float out[3]) {
  out[0] = out[1] = out[2] = 0.0f;
  if (n == 0) return;
  float t[3] = {1.0f, 1.0f, 1.0f};
  float contrib[3] = {0.0f, 0.0f, 0.0f};
  float previousRR = 1.0f;
  for (int j = 1; j < n; ++j) {
    const PathVertex& prev = seg[j - 1];
    const PathVertex& v = seg[j];
    for (int k = 0; k < 3; k++) {
      t[k] = t[k] * prev.transmittanceWeight[k];
      contrib[k] += t[k] * v.scatteredContribution[k];
    }
    float direct[3] = {v.directContribution[0], v.directContribution[1], v.directContribution[2]};
    if (!rrAffectsDirectContribution)
      for (int k = 0; k < 3; k++) direct[k] *= previousRR;
    for (int k = 0; k < 3; k++) contrib[k] += (t[k] * v.miWeight) * direct[k];
    for (int k = 0; k < 3; k++) t[k] = t[k] * v.scatteringWeight[k];
    const float rr = v.russianRouletteProbability;
    if (rr > 0.f) {
      for (int k = 0; k < 3; k++) t[k] /= rr;
    } else {
      t[0] = t[1] = t[2] = 0.0f;
    }
    previousRR = rr;
  }
  const PathVertex& f = seg[0];
  if (n == 1) {
    for (int k = 0; k < 3; k++) out[k] = f.directContribution[k] + f.scatteredContribution[k];
  } else {
    for (int k = 0; k < 3; k++)
      out[k] = (f.directContribution[k] + f.scatteredContribution[k]) + f.scatteringWeight[k] * contrib[k];
  }
}

}  // namespace pglgpu
