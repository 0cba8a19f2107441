// Batched guiding queries on gfx950: for each query, region lookup (kd walk
// plus the stochastic KNN pick), parallax shift of the region mixture to the
// query position, product with the cosine lobe (surface) or the 3-lobe
// Henyey-Greenstein representation (volume), one sample and its pdf.
//
// One thread per query; mixture components are recomputed per pass (integral,
// CDF walk, pdf) instead of being stored, so nothing spills and the region
// data (4 KiB per region) stays L2/MALL resident while queries stream.
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

// integral of the product of the shifted mixture with lobe (weight, n, kn, nn)
__device__ float mixture_integral(const VMMParams& V, bool par, float sx, float sy, float sz, float weight, float nx,
                                  float ny, float nz, float kn, float nn) {
  float integral = 0.f;
  for (int k = 0; k < V.K; k++) {
    float mx, my, mz, d;
    shift_comp(V, k, sx, sy, sz, mx, my, mz, d, par);
    const ProdComp p = product_comp(V.w[k], V.kappa[k], mx, my, mz, V.norm[k], weight, nx, ny, nz, kn, nn);
    integral += p.scale;
  }
  return integral;
}

__device__ void mixture_sample(const VMMParams& V, bool par, float sx, float sy, float sz, float weight, float nx,
                               float ny, float nz, float kn, float nn, float integral, float u0, float u1, float* out) {
  float sum = 0.f;
  ProdComp sel;
  float cdf = 0.f;
  for (int k = 0; k < V.K; k++) {
    float mx, my, mz, d;
    shift_comp(V, k, sx, sy, sz, mx, my, mz, d, par);
    sel = product_comp(V.w[k], V.kappa[k], mx, my, mz, V.norm[k], weight, nx, ny, nz, kn, nn);
    cdf = sel.scale / integral;
    if (sum + cdf >= u1 || k + 1 >= V.K) break;
    sum += cdf;
  }
  const float us = fminf(1.f - 1.19209290e-7f, (u1 - sum) / cdf);
  sample_comp(sel.kappa, sel.em2k, sel.mx, sel.my, sel.mz, u0, us, out);
}

__device__ float mixture_pdf(const VMMParams& V, bool par, float sx, float sy, float sz, float weight, float nx,
                             float ny, float nz, float kn, float nn, float integral, float dx, float dy, float dz) {
  float pdf = 0.f;
  const float inv = 1.f / integral;
  for (int k = 0; k < V.K; k++) {
    float mx, my, mz, d;
    shift_comp(V, k, sx, sy, sz, mx, my, mz, d, par);
    const ProdComp p = product_comp(V.w[k], V.kappa[k], mx, my, mz, V.norm[k], weight, nx, ny, nz, kn, nn);
    pdf += eval_comp(p.scale * inv, p.norm, p.kappa, p.mx, p.my, p.mz, dx, dy, dz);
  }
  return pdf;
}

// One query: lookup (+ KNN pick), parallax shift, cosine (surface) or 3-lobe
// HG (volume) product, one sample and its pdf.  Invalid queries (outside the
// field, untrained field, invalid region) give dir 0, pdf 0, id ~0.
__device__ __forceinline__ void query_one(const QueryField& surf, const QueryField& vol, const HGRep* __restrict__ hg,
                                          bool volume, float x, float y, float z, float nx, float ny, float nz,
                                          float g, float ux, float uy, float u, float out[3], float& pdf,
                                          uint32_t& id, float& integral) {
  const QueryField& f = volume ? vol : surf;
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
  if (!volume) {
    const float kn = kCosineKappa, nn = cosine_normalization();
    integral = mixture_integral(V, par, sx, sy, sz, 1.f, nx, ny, nz, kn, nn);
    mixture_sample(V, par, sx, sy, sz, 1.f, nx, ny, nz, kn, nn, integral, ux, uy, out);
    pdf = mixture_pdf(V, par, sx, sy, sz, 1.f, nx, ny, nz, kn, nn, integral, out[0], out[1], out[2]);
  } else {
    const HGRep& rep = hg[hg_index(g)];
    const int KP = (int)rep.K;
    float integ[4], sgn[4];
    float sum = 0.f;
    for (int m = 0; m < KP; m++) {
      sgn[m] = (g * rep.mc[m] > 0.f) ? 1.f : -1.f;
      integ[m] = mixture_integral(V, par, sx, sy, sz, rep.w[m], nx * sgn[m], ny * sgn[m], nz * sgn[m], rep.kappa[m], rep.norm[m]);
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
    mixture_sample(V, par, sx, sy, sz, rep.w[m], nx * sgn[m], ny * sgn[m], nz * sgn[m], rep.kappa[m], rep.norm[m], integ[m],
                   (ux - acc) / mw[m], uy, out);
    for (int j = 0; j < KP; j++)
      pdf += mw[j] * mixture_pdf(V, par, sx, sy, sz, rep.w[j], nx * sgn[j], ny * sgn[j], nz * sgn[j], rep.kappa[j],
                                 rep.norm[j], integ[j], out[0], out[1], out[2]);
    integral = sum;
  }
}

__global__ void __launch_bounds__(256) k_query(QueryField surf, QueryField vol, const HGRep* __restrict__ hg,
                                               uint32_t n, const float* __restrict__ pos, const float* __restrict__ nrm,
                                               const float* __restrict__ gArr, const float* __restrict__ u2,
                                               const float* __restrict__ u1, const uint8_t* __restrict__ isVolume,
                                               float* __restrict__ dirOut, float* __restrict__ pdfOut,
                                               uint32_t* __restrict__ idOut, float* __restrict__ integralOut) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool volume = isVolume ? isVolume[i] != 0 : false;
  float out[3], pdf, integral;
  uint32_t id;
  query_one(surf, vol, hg, volume, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], nrm[3 * i], nrm[3 * i + 1],
            nrm[3 * i + 2], gArr ? gArr[i] : 0.f, u2[2 * i], u2[2 * i + 1], u1 ? u1[i] : -1.f, out, pdf, id, integral);
  dirOut[3 * i] = out[0];
  dirOut[3 * i + 1] = out[1];
  dirOut[3 * i + 2] = out[2];
  pdfOut[i] = pdf;
  if (idOut) idOut[i] = id;
  if (integralOut) integralOut[i] = integral;
}

// Batch of pieces: thread i of the batch finds its piece (pieces ordered by
// begin) and works on query i - begin of it through the piece's pointers.
__global__ void __launch_bounds__(256) k_query_pieces(QueryField surf, QueryField vol, const HGRep* __restrict__ hg,
                                                      const QueryPiece* __restrict__ pieces, uint32_t numPieces,
                                                      uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
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
  query_one(surf, vol, hg, volume, P.pos[3 * j], P.pos[3 * j + 1], P.pos[3 * j + 2], P.nrm[3 * j], P.nrm[3 * j + 1],
            P.nrm[3 * j + 2], P.g ? P.g[j] : 0.f, P.u2[2 * j], P.u2[2 * j + 1], P.u1 ? P.u1[j] : -1.f, out, pdf, id,
            integral);
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
  hipLaunchKernelGGL(k_query_pieces, dim3((n + 255) / 256), dim3(256), 0, s, surf, vol, hg, pieces, numPieces, n);
}

void launch_query(const QueryField& surf, const QueryField& vol, const HGRep* hg, uint32_t n, const float* pos,
                  const float* nrm, const float* g, const float* u2, const float* u1, const uint8_t* isVolume,
                  float* dirOut, float* pdfOut, uint32_t* idOut, float* integralOut, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_query, dim3((n + 255) / 256), dim3(256), 0, s, surf, vol, hg, n, pos, nrm, g, u2, u1, isVolume,
                     dirOut, pdfOut, idOut, integralOut);
}

}  // namespace pglgpu
