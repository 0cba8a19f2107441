// Batched guiding queries on gfx950: for each query, region lookup (kd walk
// plus the stochastic KNN pick), parallax shift of the region mixture to the
// query position, product with the cosine lobe (surface) or the 3-lobe
// Henyey-Greenstein representation (volume), one sample and its pdf.
//
// A quad of four lanes per query, each lane holding the products of a quarter
// of the components in registers between the integral, the CDF walk and the
// pdf (surface); the region data (4 KiB per region) stays L2/MALL resident
// while queries stream.
//
// Reference behaviour: Field::getRegion field/Field.h:109-145,
// SurfaceVolumeField::init*SamplingDistribution field/SurfaceVolumeField.h:50-85,
// VMMSurfaceSamplingDistribution.h:37-97, VMMVolumeSamplingDistribution.h:37-123.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"
#include "query_math.h"

namespace pglgpu {

__device__ __forceinline__ int lookup_region(const QueryField& f, float x, float y, float z, float* u, uint32_t& id) {
  if (!f.trained) return -1;
  if (!(x >= f.lo[0] && x <= f.hi[0] && y >= f.lo[1] && y <= f.hi[1] && z >= f.lo[2] && z <= f.hi[2])) return -1;
  uint32_t i = 0;
  KDNode nd = f.nodes[0];
  while (!nd.isLeaf()) {
    const uint32_t d = nd.dim();
    const float p = d == 0 ? x : (d == 1 ? y : z);
    i = nd.idx() + (p >= nd.pivot ? 1u : 0u);
    nd = f.nodes[i];
  }
  const uint32_t leaf = nd.idx();
  id = leaf;
  if (f.knn && *u >= 0.f) {
    uint32_t ids[8];
    float px[8], py[8], pz[8];
    for (int s = 0; s < 8; s++) {
      ids[s] = f.nbIds[leaf * 8 + s];
      const float4 p = f.nbPts[leaf * 8 + s];
      px[s] = p.x;
      py[s] = p.y;
      pz[s] = p.z;
    }
    return (int)knn_pick(ids, px, py, pz, f.nbSize, x, y, z, u, f.knnFused != 0);
  }
  return (int)leaf;
}

// The four lanes of a quad (lanes 4i .. 4i + 3 of a wave) work on one query:
// lane q owns the components k = 4j + q (j < 8), so a query's component
// loops take a quarter of the serial steps and the component parameters of
// a region are read by four lanes at once.  Every lane of a quad runs the
// same control flow (the query decides it), so the quad sums below see four
// active lanes.

// sum over the quad; the same bits in its four lanes (IEEE addition commutes)
__device__ __forceinline__ float qsum4(float v) {
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));  // lanes q ^ 1
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));  // lanes q ^ 2
  return v;
}

// value of quad lane 0, 1, 2 or 3 (DPP quad broadcast)
template <int L>
__device__ __forceinline__ float qlane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), L * 0x55, 0xF, 0xF, false));
}

// This lane's components (k = 4j + q) after the parallax shift to the query
// position: their directions, formed once per query and shared by every
// product below.
struct LaneShift {
  float mx[8], my[8], mz[8];
};
__device__ __forceinline__ void lane_shift(const VMMParams& V, int q, bool par, float sx, float sy, float sz,
                                           LaneShift& h) {
  const int K = V.K;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int k = 4 * j + q;
    h.mx[j] = 0.f; h.my[j] = 0.f; h.mz[j] = 1.f;
    if (k < K) {
      float d;
      shift_comp(V, k, sx, sy, sz, h.mx[j], h.my[j], h.mz[j], d, par);
    }
  }
}

// Products of this lane's shifted components with the lobe (weight, n, kn,
// nn): scale 0 for k >= K.
struct LaneComps {
  float s[8], kap[8], mx[8], my[8], mz[8], nrm[8], e2k[8];
};
__device__ __forceinline__ void lane_products(const VMMParams& V, int q, const LaneShift& h, float weight, float nx,
                                              float ny, float nz, float kn, float nn, LaneComps& c) {
  const int K = V.K;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int k = 4 * j + q;
    c.s[j] = 0.f; c.kap[j] = 0.f; c.mx[j] = 0.f; c.my[j] = 0.f; c.mz[j] = 1.f; c.nrm[j] = 0.f; c.e2k[j] = 1.f;
    if (k < K) {
      const ProdComp p =
          product_comp(V.w[k], V.kappa[k], h.mx[j], h.my[j], h.mz[j], V.norm[k], weight, nx, ny, nz, kn, nn);
      c.s[j] = p.scale; c.kap[j] = p.kappa; c.mx[j] = p.mx; c.my[j] = p.my; c.mz[j] = p.mz; c.nrm[j] = p.norm;
      c.e2k[j] = p.em2k;
    }
  }
}

// integral of the product of the shifted mixture with the lobe (quad sum)
__device__ __forceinline__ float quad_integral(const LaneComps& c) {
  float part = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) part += c.s[j];
  return qsum4(part);
}

// pdf at direction d of the product held in c (quad sum)
__device__ __forceinline__ float quad_pdf(const VMMParams& V, int q, const LaneComps& c, float integral, float dx,
                                          float dy, float dz) {
  const float inv = 1.f / integral;
  float part = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++)
    if (4 * j + q < V.K) part += eval_comp(c.s[j] * inv, c.nrm[j], c.kap[j], c.mx[j], c.my[j], c.mz[j], dx, dy, dz);
  return qsum4(part);
}

// pdf at d of the product with a lobe, the products formed on the fly (quad sum)
__device__ __forceinline__ float quad_pdf_lobe(const VMMParams& V, int q, const LaneShift& h, float weight, float nx,
                                               float ny, float nz, float kn, float nn, float integral, float dx,
                                               float dy, float dz) {
  const float inv = 1.f / integral;
  float part = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int k = 4 * j + q;
    if (k < V.K) {
      const ProdComp p =
          product_comp(V.w[k], V.kappa[k], h.mx[j], h.my[j], h.mz[j], V.norm[k], weight, nx, ny, nz, kn, nn);
      part += eval_comp(p.scale * inv, p.norm, p.kappa, p.mx, p.my, p.mz, dx, dy, dz);
    }
  }
  return qsum4(part);
}

// One direction from the product held in c (PAVMM::sample): the component is
// the first k, in component order, whose cumulative share of the integral
// reaches u1 (the last one if rounding leaves the total short), then u0 and
// the rescaled u1 sample it.  The walk goes over groups of four components
// (one per lane) and then inside the chosen group.
__device__ __forceinline__ void quad_sample(const VMMParams& V, int q, int lane, const LaneComps& c, float integral,
                                            float u0, float u1, float* out) {
  const int K = V.K;
  const int J = (K + 3) >> 2;
  float acc = 0.f, accs = 0.f;  // share of the groups before group js
  int js = -1;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (j < J && js < 0) {
      const float g = qsum4(c.s[j] / integral);
      if (acc + g >= u1 || j == J - 1) {  // the last group is walked in any case
        js = j;
        accs = acc;
      } else {
        acc += g;
      }
    }
  }
  // this lane's share in group js and the values of its component
  float x = 0.f, kap = 0.f, e2k = 1.f, mx = 0.f, my = 0.f, mz = 1.f;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (j == js) {
      x = c.s[j] / integral; kap = c.kap[j]; e2k = c.e2k[j]; mx = c.mx[j]; my = c.my[j]; mz = c.mz[j];
    }
  }
  const float x0 = qlane<0>(x), x1 = qlane<1>(x), x2 = qlane<2>(x), x3 = qlane<3>(x);
  const int last = 4 * js + 3 < K ? 3 : ((K - 1) & 3);  // last component of the group
  // the walk inside the group: components 4 js + l, l = 0 .. last
  float sum = accs;
  int qs = last;
  if (last > 0 && sum + x0 >= u1) qs = 0;
  else {
    if (last > 0) sum += x0;
    if (last > 1 && sum + x1 >= u1) qs = 1;
    else {
      if (last > 1) sum += x1;
      if (last > 2 && sum + x2 >= u1) qs = 2;
      else if (last > 2) sum += x2;
    }
  }
  const float cdf = qs == 0 ? x0 : (qs == 1 ? x1 : (qs == 2 ? x2 : x3));
  const int src = (lane & ~3) | qs;
  kap = __shfl(kap, src);
  e2k = __shfl(e2k, src);
  mx = __shfl(mx, src);
  my = __shfl(my, src);
  mz = __shfl(mz, src);
  const float us = fminf(1.f - 1.19209290e-7f, (u1 - sum) / cdf);
  sample_comp(kap, e2k, mx, my, mz, u0, us, out);
}

// One query on the four lanes of a quad: lookup (+ KNN pick), parallax shift,
// cosine (surface) or 3-lobe HG (volume) product, one sample and its pdf.
// Invalid queries (outside the field, untrained field, invalid region) give
// dir 0, pdf 0, id ~0.  Every lane returns the same results.
__device__ __forceinline__ void query_quad(const QueryField& surf, const QueryField& vol, const HGRep* __restrict__ hg,
                                           bool volume, float x, float y, float z, float nx, float ny, float nz,
                                           float g, float ux, float uy, float u, float out[3], float& pdf,
                                           uint32_t& id, float& integral) {
  const QueryField& f = volume ? vol : surf;
  const int lane = threadIdx.x & 63, q = lane & 3;
  id = 0xffffffffu;
  pdf = 0.f;
  integral = 0.f;
  out[0] = out[1] = out[2] = 0.f;
  const int r = lookup_region(f, x, y, z, &u, id);
  if (r < 0 || !f.regions[r].valid) {
    id = 0xffffffffu;
    return;
  }
  const VMMParams& V = f.regions[r].vmm;
  const bool par = f.parallax != 0;
  const float sx = V.pivot[0] - x, sy = V.pivot[1] - y, sz = V.pivot[2] - z;
  LaneShift h;
  lane_shift(V, q, par, sx, sy, sz, h);
  LaneComps c;
  if (!volume) {
    const float kn = kCosineKappa, nn = cosine_normalization();
    lane_products(V, q, h, 1.f, nx, ny, nz, kn, nn, c);
    integral = quad_integral(c);
    quad_sample(V, q, lane, c, integral, ux, uy, out);
    pdf = quad_pdf(V, q, c, integral, out[0], out[1], out[2]);
  } else {
    const HGRep& rep = hg[hg_index(g)];
    const int KP = (int)rep.K;
    float integ[4], sgn[4];
    float sum = 0.f;
    for (int m = 0; m < KP; m++) {
      sgn[m] = (g * rep.mc[m] > 0.f) ? 1.f : -1.f;
      lane_products(V, q, h, rep.w[m], nx * sgn[m], ny * sgn[m], nz * sgn[m], rep.kappa[m], rep.norm[m], c);
      integ[m] = quad_integral(c);
      sum += integ[m];
    }
    float mw[4];
    for (int m = 0; m < KP; m++) mw[m] = integ[m] / sum;
    float acc = 0.f;
    int m = 0;
    for (; m < KP - 1; ++m) {
      const float nw = acc + mw[m];
      if (nw > ux) break;
      acc = nw;
    }
    lane_products(V, q, h, rep.w[m], nx * sgn[m], ny * sgn[m], nz * sgn[m], rep.kappa[m], rep.norm[m], c);
    quad_sample(V, q, lane, c, integ[m], (ux - acc) / mw[m], uy, out);
    for (int j = 0; j < KP; j++)
      pdf += mw[j] * (j == m ? quad_pdf(V, q, c, integ[j], out[0], out[1], out[2])
                             : quad_pdf_lobe(V, q, h, rep.w[j], nx * sgn[j], ny * sgn[j], nz * sgn[j], rep.kappa[j],
                                             rep.norm[j], integ[j], out[0], out[1], out[2]));
    integral = sum;
  }
}

__global__ void __launch_bounds__(256) k_query(QueryField surf, QueryField vol, const HGRep* __restrict__ hg,
                                               uint32_t n, const float* __restrict__ pos, const float* __restrict__ nrm,
                                               const float* __restrict__ gArr, const float* __restrict__ u2,
                                               const float* __restrict__ u1, const uint8_t* __restrict__ isVolume,
                                               float* __restrict__ dirOut, float* __restrict__ pdfOut,
                                               uint32_t* __restrict__ idOut, float* __restrict__ integralOut) {
  const uint32_t i = (uint32_t)(((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2);  // this quad's query
  if (i >= n) return;
  const bool volume = isVolume ? isVolume[i] != 0 : false;
  float out[3], pdf, integral;
  uint32_t id;
  query_quad(surf, vol, hg, volume, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], nrm[3 * i], nrm[3 * i + 1],
             nrm[3 * i + 2], gArr ? gArr[i] : 0.f, u2[2 * i], u2[2 * i + 1], u1 ? u1[i] : -1.f, out, pdf, id, integral);
  if ((threadIdx.x & 3) != 0) return;
  dirOut[3 * i] = out[0];
  dirOut[3 * i + 1] = out[1];
  dirOut[3 * i + 2] = out[2];
  pdfOut[i] = pdf;
  if (idOut) idOut[i] = id;
  if (integralOut) integralOut[i] = integral;
}

// Batch of pieces: the quad of query i of the batch finds its piece (pieces
// ordered by begin) and works on query i - begin of it through the piece's
// pointers.
__global__ void __launch_bounds__(256) k_query_pieces(QueryField surf, QueryField vol, const HGRep* __restrict__ hg,
                                                      const QueryPiece* __restrict__ pieces, uint32_t numPieces,
                                                      uint32_t n) {
  const uint32_t i = (uint32_t)(((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2);
  if (i >= n) return;
  uint32_t lo = 0, hi = numPieces;  // last piece with begin <= i
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pieces[mid].begin <= i) lo = mid; else hi = mid;
  }
  const QueryPiece& P = pieces[lo];
  const uint32_t j = i - P.begin;
  if (j >= P.n) return;
  const bool volume = P.vol ? P.vol[j] != 0 : false;
  float out[3], pdf, integral;
  uint32_t id;
  query_quad(surf, vol, hg, volume, P.pos[3 * j], P.pos[3 * j + 1], P.pos[3 * j + 2], P.nrm[3 * j], P.nrm[3 * j + 1],
             P.nrm[3 * j + 2], P.g ? P.g[j] : 0.f, P.u2[2 * j], P.u2[2 * j + 1], P.u1 ? P.u1[j] : -1.f, out, pdf, id,
             integral);
  if ((threadIdx.x & 3) != 0) return;
  P.dir[3 * j] = out[0];
  P.dir[3 * j + 1] = out[1];
  P.dir[3 * j + 2] = out[2];
  P.pdf[j] = pdf;
  if (P.id) P.id[j] = id;
  if (P.integral) P.integral[j] = integral;
}

void launch_query_pieces(const QueryField& surf, const QueryField& vol, const HGRep* hg, const QueryPiece* pieces,
                         uint32_t numPieces, uint32_t n, hipStream_t s) {
  if (n == 0 || numPieces == 0) return;
  hipLaunchKernelGGL(k_query_pieces, dim3((n + 63) / 64), dim3(256), 0, s, surf, vol, hg, pieces, numPieces, n);
}

void launch_query(const QueryField& surf, const QueryField& vol, const HGRep* hg, uint32_t n, const float* pos,
                  const float* nrm, const float* g, const float* u2, const float* u1, const uint8_t* isVolume,
                  float* dirOut, float* pdfOut, uint32_t* idOut, float* integralOut, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_query, dim3((n + 63) / 64), dim3(256), 0, s, surf, vol, hg, n, pos, nrm, g, u2, u1, isVolume,
                     dirOut, pdfOut, idOut, integralOut);
}

}  // namespace pglgpu
