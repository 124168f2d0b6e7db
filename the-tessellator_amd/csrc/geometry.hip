// geometry.hip -- staged cell geometry -> CSR (the compaction behind
// tess_diagram_compute_geometry).  HBM-bound copies, one wavefront per
// staged cell.
#include "launch.h"
#include "tess_internal.h"

namespace tess {

namespace {

__global__ __launch_bounds__(256) void k_geo_compact(const uint32_t* __restrict__ slot_cell, int nslots, int vcap,
                                                     int hcap, const double* __restrict__ sv,
                                                     const uint16_t* __restrict__ sr,
                                                     const uint16_t* __restrict__ sro,
                                                     const uint32_t* __restrict__ count, int n,
                                                     const uint64_t* __restrict__ voff,
                                                     const uint64_t* __restrict__ rbase,
                                                     const uint64_t* __restrict__ foff, double* __restrict__ out_v,
                                                     uint16_t* __restrict__ out_ring,
                                                     uint64_t* __restrict__ out_roff) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= nslots) return;
  const uint32_t t = slot_cell[w];
  if (!t) return;  // nothing staged here (the cell overflowed this pass)
  const uint32_t o = t - 1u;
  const uint32_t nv = count[o], nr = count[(size_t)n + o];
  const uint64_t f0 = foff[o], nf = foff[o + 1] - f0, vb = voff[o], rb = rbase[o];
  const double* v = sv + (size_t)w * 3 * vcap;
  for (uint32_t k = lane; k < 3 * nv; k += 64) out_v[3 * vb + k] = v[k];
  const uint16_t* r = sr + (size_t)w * hcap;
  for (uint32_t k = lane; k < nr; k += 64) out_ring[rb + k] = r[k];
  const uint16_t* ro = sro + (size_t)w * kFaceStage;
  for (uint32_t k = lane; k < nf; k += 64) out_roff[f0 + k] = rb + ro[k];
}

}  // namespace

hipError_t launch_geo_compact(const uint32_t* slot_cell, int nslots, int vcap, int hcap, const double* sv,
                              const uint16_t* sr, const uint16_t* sro, const uint32_t* count, int n,
                              const uint64_t* voff, const uint64_t* rbase, const uint64_t* foff, double* out_v,
                              uint16_t* out_ring, uint64_t* out_roff, hipStream_t s) {
  if (nslots <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_geo_compact, dim3((unsigned)((nslots + 3) / 4)), dim3(256), 0, s, slot_cell, nslots, vcap,
                     hcap, sv, sr, sro, count, n, voff, rbase, foff, out_v, out_ring, out_roff);
  return hipGetLastError();
}

}  // namespace tess
