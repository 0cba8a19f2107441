// Per-component query math shared by the host scalar API and the batched
// GPU query kernel: parallax shift, vMF product with a cosine or phase-
// function lobe, component sampling and evaluation.
//
// Reference behaviour (directional/vmm/, Open PGL 0.6.0):
//   performRelativeParallaxShift  ParallaxAwareVonMisesFisherMixture.h:809-831
//   product                       ParallaxAwareVonMisesFisherMixture.h:609-671
//   pdf                           ParallaxAwareVonMisesFisherMixture.h:676-694
//   sample                        ParallaxAwareVonMisesFisherMixture.h:737-806
//   cosine lobe kappa 2.18853     VMMSurfaceSamplingDistribution.h:52-60
//   HG representation lookup      VMMPhaseFunctions.h:44-57
#pragma once
#include <math.h>
#include <stdint.h>

#include "common.h"

namespace pglgpu {

static const float kCosineKappa = 2.18853f;

struct HGRep {
  float K, g, w[4], mc[4], kappa[4], norm[4];
};

PGL_HD inline float qdot(float ax, float ay, float az, float bx, float by, float bz) {
  return fmaf(ax, bx, fmaf(ay, by, az * bz));
}

PGL_HD inline float cosine_normalization() {
  return 2.18853f / (2.0f * 3.14159265358979f * (1.0f - expf(-2.0f * 2.18853f)));
}

// component k of a region's mixture after the parallax shift to a query
// position (shift = pivot - position); unshifted for the plain VMM field
// (VMMSurfaceSamplingDistribution.h:41, VMMVolumeSamplingDistribution.h:43)
PGL_HD inline void shift_comp(const VMMParams& V, int k, float sx, float sy, float sz, float& mx, float& my,
                              float& mz, float& dist, bool parallax = true) {
  const float d = V.dist[k];
  mx = V.mux[k];
  my = V.muy[k];
  mz = V.muz[k];
  if (!parallax) {
    dist = d;
    return;
  }
  const float px = mx * d + sx, py = my * d + sy, pz = mz * d + sz;
  const float l = sqrtf(qdot(px, py, pz, px, py, pz));
  if (d > 0.f && isfinite(d)) {
    mx = px / l;
    my = py / l;
    mz = pz / l;
  }
  dist = d > 0.f ? l : 0.f;
}

struct ProdComp {
  float scale;  // unnormalised product weight
  float kappa, mx, my, mz, norm, em2k;
};

// product of one component with the lobe (weight, n, kn, nn)
PGL_HD inline ProdComp product_comp(float w, float kappa, float mx, float my, float mz, float norm, float weight,
                                    float nx, float ny, float nz, float kn, float nn) {
  ProdComp p;
  float ax = kappa * mx + kn * nx, ay = kappa * my + kn * ny, az = kappa * mz + kn * nz;
  float nk = sqrtf(qdot(ax, ay, az, ax, ay, az));
  const bool ok = nk > 1e-3f;
  nk = ok ? nk : 0.f;
  ax = ok ? ax / nk : mx;
  ay = ok ? ay / nk : my;
  az = ok ? az / nk : mz;
  const float e = expf(-2.f * nk);
  float nrm = nk / (2.f * 3.14159265358979f * (1.f - e));
  nrm = ok ? nrm : kOneOverFourPi;
  float scale = (norm * nn) / nrm;
  const float c0 = qdot(mx, my, mz, ax, ay, az);
  const float c1 = qdot(nx, ny, nz, ax, ay, az);
  scale *= expf(kappa * (c0 - 1.f) + kn * (c1 - 1.f));
  scale *= w * weight;
  p.scale = scale;
  p.kappa = nk;
  p.mx = ax;
  p.my = ay;
  p.mz = az;
  p.norm = nrm;
  p.em2k = e;
  return p;
}

PGL_HD inline float eval_comp(float w, float norm, float kappa, float mx, float my, float mz, float dx, float dy,
                              float dz) {
  const float c = qdot(dx, dy, dz, mx, my, mz);
  return w * (norm * expf(kappa * fminf(c - 1.f, 0.f)));
}

// embree::frame(N) columns dx, dy
PGL_HD inline void qframe(float nx, float ny, float nz, float* fx, float* fy) {
  const float a0[3] = {0.f, nz, -ny};
  const float a1[3] = {-nz, 0.f, nx};
  const bool first = qdot(a0[0], a0[1], a0[2], a0[0], a0[1], a0[2]) > qdot(a1[0], a1[1], a1[2], a1[0], a1[1], a1[2]);
  const float* d = first ? a0 : a1;
  float l = 1.f / sqrtf(qdot(d[0], d[1], d[2], d[0], d[1], d[2]));
  fx[0] = d[0] * l;
  fx[1] = d[1] * l;
  fx[2] = d[2] * l;
  const float c[3] = {ny * fx[2] - nz * fx[1], nz * fx[0] - nx * fx[2], nx * fx[1] - ny * fx[0]};
  l = 1.f / sqrtf(qdot(c[0], c[1], c[2], c[0], c[1], c[2]));
  fy[0] = c[0] * l;
  fy[1] = c[1] * l;
  fy[2] = c[2] * l;
}

// sample one vMF component with (u0, u1) (u1 already rescaled by the CDF walk)
PGL_HD inline void sample_comp(float kappa, float em2k, float mx, float my, float mz, float u0, float u1, float* out) {
  float lx, ly, lz;
  if (kappa == 0.f) {
    const float z = 1.f - 2.f * u1;
    const float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    const float phi = 2.f * 3.14159265358979f * u0;
    lx = r * cosf(phi);
    ly = r * sinf(phi);
    lz = z;
  } else {
    float ct = 1.f + logf(1.f + ((em2k - 1.f) * u0)) / kappa;
    ct = fminf(1.f, fmaxf(ct, -1.f));
    const float st = sqrtf(1.f - ct * ct);
    const float phi = 2.f * 3.14159265358979f * u1;
    lx = st * cosf(phi);
    ly = st * sinf(phi);
    lz = ct;
  }
  float fx[3], fy[3];
  qframe(mx, my, mz, fx, fy);
  out[0] = fx[0] * lx + fy[0] * ly + mx * lz;
  out[1] = fx[1] * lx + fy[1] * ly + my * lz;
  out[2] = fx[2] * lx + fy[2] * ly + mz * lz;
}

PGL_HD inline int hg_index(float g) {
  const float ag = fabsf(g);
  const float step = (0.99f - 0.0f) / 128.f;
  int idx = (int)floorf((ag - 0.0f) / step);
  return idx < 127 ? idx : 127;
}

PGL_HD inline float knn_d2_unfused(float dx, float dy, float dz) {
#pragma clang fp contract(off)
  return dx * dx + (dy * dy + dz * dz);
}

// KNN draw(): selects one of min(4, size) neighbours and rescales u in place (spatial/KNN.h:22-28)
PGL_HD inline uint32_t knn_draw(float* u, uint32_t size) {
  size = size < 4u ? size : 4u;
  const uint32_t sel = (uint32_t)(*u * (float)size);
  *u = (*u - (float)sel / (float)size) * (float)size;
  return sel < size - 1 ? sel : size - 1;
}

// embree::dot(d, d) = madd(x, x, madd(y, y, z * z)) (math/vec3.h:204) with
// madd fused (AVX2 / AVX-512 builds) or a * b + c rounded twice (SSE4 build,
// simd/vfloat4_sse2.h:439-444)
PGL_HD inline float knn_d2(float dx, float dy, float dz, bool fused) {
  if (fused) return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  return knn_d2_unfused(dx, dy, dz);
}

// RegionNeighbours<8>::sampleApproximateClosestRegionIdx (spatial/KNN.h:134-148;
// RegionNeighbours<4>, :86-113, merges two sorted halves to the same pick):
// squared distances with the low 3 mantissa bits replaced by the slot index,
// invalid slots at +inf, sorted ascending, pick the selected one.
PGL_HD inline uint32_t knn_pick(const uint32_t* ids, const float* px, const float* py, const float* pz, uint32_t size,
                                float x, float y, float z, float* u, bool fusedD2) {
  const uint32_t sel = knn_draw(u, size);
  float d[8];
  for (int i = 0; i < 8; i++) {
    if (ids[i] == 0xffffffffu) {
      d[i] = INFINITY;
      continue;
    }
    const float dx = px[i] - x, dy = py[i] - y, dz = pz[i] - z;
    float dd = knn_d2(dx, dy, dz, fusedD2);
    union { float f; uint32_t u; } cv;
    cv.f = dd;
    cv.u = (cv.u & ~7u) | (uint32_t)i;
    d[i] = cv.f;
  }
  for (int i = 1; i < 8; i++) {  // insertion sort of 8 keys
    const float v = d[i];
    int j = i - 1;
    while (j >= 0 && d[j] > v) {
      d[j + 1] = d[j];
      j--;
    }
    d[j + 1] = v;
  }
  union { float f; uint32_t u; } cv;
  cv.f = d[sel];
  return ids[cv.u & 7u];
}

}  // namespace pglgpu
