// One training sample moved into the region-ordered SoA arrays of the
// directional update, reprojected to its region's pivot (shared by the
// gather kernel and the sort that gathers while it scatters).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pglgpu {

// PGLSampleData (include/openpgl/data.h:21-49)
struct SampleAoS {
  float px, py, pz, dx, dy, dz, w, pdf, dist;
  uint32_t flags;
};
static_assert(sizeof(SampleAoS) == 40, "PGLSampleData is 40 B");

// Order-preserving map of a float to uint32 (-0 mapped to +0, which
// SampleDataLess treats as equal): a < b as floats iff key(a) < key(b).
__device__ __forceinline__ uint32_t float_order_key(float x) {
  uint32_t b = __float_as_uint(x);
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Reprojection to the region's pivot: origin = p + d * max(minDist, dist);
// d' = normalise(origin - pivot).  distance == inf keeps the direction
// (position moves to the pivot); distance <= 0 keeps the sample unchanged.
// Without reprojection (plain VMM field, WEM:1186) the sample is written as
// recorded.
__device__ __forceinline__ void gather_sample(const SampleAoS& s, float4 pv /* xyz, minDist */, int reproject,
                                              uint32_t j, float* __restrict__ odx, float* __restrict__ ody,
                                              float* __restrict__ odz, float* __restrict__ ow,
                                              float* __restrict__ opdf, float* __restrict__ odist) {
  float dx = s.dx, dy = s.dy, dz = s.dz, dist = s.dist;
  if (reproject && !isinf(s.dist) && s.dist > 0.f) {
    const float dd = fmaxf(pv.w, s.dist);
    const float ox = s.px + s.dx * dd - pv.x;
    const float oy = s.py + s.dy * dd - pv.y;
    const float oz = s.pz + s.dz * dd - pv.z;
    const float len = sqrtf(ox * ox + oy * oy + oz * oz);
    dx = ox / len;
    dy = oy / len;
    dz = oz / len;
    dist = len;
  }
  odx[j] = dx;
  ody[j] = dy;
  odz[j] = dz;
  ow[j] = s.w;
  opdf[j] = s.pdf;
  odist[j] = dist;
}

}  // namespace pglgpu
