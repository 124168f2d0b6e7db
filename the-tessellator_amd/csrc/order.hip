// order.hip -- the reference's internal point order on the device:
// Diagram::morton_index (interface.rs:151-183) of every point and the stable
// sort by it that sort_particles intends (interface.rs:88-148).
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "launch.h"
#include "tess_internal.h"

namespace tess {

namespace {

// `(v / side) as usize` per axis: truncation, negative and NaN -> 0,
// saturation above (only the low 10 bits are used)
__device__ __forceinline__ uint32_t morton_axis(double v, double lo, double side) {
  const double t = (v - lo) / side;
  if (!(t > 0.0)) return 0u;
  if (t >= 4294967295.0) return 0xffffffffu;
  return (uint32_t)t;
}

// The diagram's bounding box starts as the default box at the origin and is
// grown to contain every point (interface.rs:52-57, vector3.rs:253-287): the
// points' bounds, extended to the origin.
__global__ void k_morton(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                         int n, const GridDev* __restrict__ gp, uint32_t* __restrict__ codes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GridDev g = *gp;
  const double lx = g.xmin < 0.0 ? g.xmin : 0.0, ly = g.ymin < 0.0 ? g.ymin : 0.0, lz = g.zmin < 0.0 ? g.zmin : 0.0;
  const double hx = g.xmax > 0.0 ? g.xmax : 0.0, hy = g.ymax > 0.0 ? g.ymax : 0.0, hz = g.zmax > 0.0 ? g.zmax : 0.0;
  const uint32_t ix = morton_axis(x[i], lx, (hx - lx) * (1.0 / 1024.0));
  const uint32_t iy = morton_axis(y[i], ly, (hy - ly) * (1.0 / 1024.0));
  const uint32_t iz = morton_axis(z[i], lz, (hz - lz) * (1.0 / 1024.0));
  uint32_t m = 0;
#pragma unroll
  for (int level = 0; level < 10; ++level)
    m |= (((ix >> level) & 1u) << (3 * level)) | (((iy >> level) & 1u) << (3 * level + 1)) |
         (((iz >> level) & 1u) << (3 * level + 2));
  codes[i] = m;
}

__global__ void k_iota(uint32_t* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

__global__ void k_inverse(const uint32_t* __restrict__ perm, int n, uint32_t* __restrict__ inv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) inv[perm[k]] = (uint32_t)k;
}

// Grid cell of each site (AoS xyz), clamped to the grid: sites inside the
// container box but outside the points' bounding box take the nearest cell.
__global__ void k_site_keys(const double* __restrict__ xyz, int m, const GridDev* __restrict__ g,
                            uint32_t* __restrict__ key) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double top = (double)(g->cpd - 1);
  const double fx = fmin(fmax(floor((xyz[3 * i] - g->xmin) * g->ix), 0.0), top);
  const double fy = fmin(fmax(floor((xyz[3 * i + 1] - g->ymin) * g->iy), 0.0), top);
  const double fz = fmin(fmax(floor((xyz[3 * i + 2] - g->zmin) * g->iz), 0.0), top);
  const uint32_t c = (uint32_t)g->cpd;
  key[i] = ((uint32_t)fz * c + (uint32_t)fy) * c + (uint32_t)fx;
}

inline unsigned blocks_of(int n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_morton(const double* x, const double* y, const double* z, int n, const GridDev* g,
                         uint32_t* codes, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton, dim3(blocks_of(n)), dim3(256), 0, s, x, y, z, n, g, codes);
  return hipGetLastError();
}

hipError_t launch_index_maps(const uint32_t* codes, int n, uint32_t* iota, uint32_t* keys_out,
                             uint32_t* original_of_internal, uint32_t* internal_of_original, void* tmp,
                             size_t* tmp_bytes, hipStream_t s) {
  if (!tmp)  // size query: LSD radix sort over the 30 code bits (stable)
    return rocprim::radix_sort_pairs(nullptr, *tmp_bytes, codes, keys_out, iota, original_of_internal, (size_t)n,
                                     0, 30, s);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3(blocks_of(n)), dim3(256), 0, s, iota, n);
  hipError_t e = rocprim::radix_sort_pairs(tmp, *tmp_bytes, codes, keys_out, iota, original_of_internal, (size_t)n,
                                           0, 30, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_inverse, dim3(blocks_of(n)), dim3(256), 0, s, original_of_internal, n, internal_of_original);
  return hipGetLastError();
}

hipError_t launch_site_order(const double* xyz, int m, const GridDev* g, uint32_t* keys, uint32_t* keys_out,
                             uint32_t* iota, uint32_t* order, void* tmp, size_t* tmp_bytes, hipStream_t s) {
  if (!tmp)  // size query: LSD radix sort over the 32 key bits (stable)
    return rocprim::radix_sort_pairs(nullptr, *tmp_bytes, keys, keys_out, iota, order, (size_t)m, 0, 32, s);
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_site_keys, dim3(blocks_of(m)), dim3(256), 0, s, xyz, m, g, keys);
  hipLaunchKernelGGL(k_iota, dim3(blocks_of(m)), dim3(256), 0, s, iota, m);
  return rocprim::radix_sort_pairs(tmp, *tmp_bytes, keys, keys_out, iota, order, (size_t)m, 0, 32, s);
}

}  // namespace tess
