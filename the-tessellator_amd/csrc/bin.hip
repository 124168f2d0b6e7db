// bin.hip -- uniform-grid counting sort on the device (celery.rs:231-414).
//
//   k_soa        AoS xyz -> SoA x, y, z
//   k_bounds_*   bounding box of the points (celery.rs:81-125) and the grid's
//                cell sizes / inverse sizes (celery.rs:153-181)
//   k_bin_count  cell id per point (celery.rs:269-325) and a histogram whose
//                atomics are aggregated per run of equal ids in a wavefront
//   k_scan_*     exclusive prefix scan of the histogram = delimiters
//                (celery.rs:372-414)
//   k_scatter    points to their cell's range (atomic cursor), or, with the
//                ranks the histogram's atomics hand out, without atomics
//   k_sort_cells ascending point index inside each cell (stable order, the
//                canonical form of celery.rs:357-369), or ascending
//                (Morton code, index) for the internal order
//   k_gather     positions in grid order (coalesced candidate loads)
//   k_compact    per-cell staged faces -> CSR
// All passes are HBM-bound streams over the points or the cells.
#include "launch.h"
#include "tess_internal.h"

namespace tess {

namespace {

__device__ __forceinline__ int axis_index(double v, double vmin, double vmax, double inv, int cpd) {
  if (v >= vmax) return cpd - 1;
  double t = (v - vmin) * inv;
  if (!(t > 0.0)) return 0;
  if (t >= 2147483647.0) return cpd - 1;
  int i = (int)t;
  return i < cpd - 1 ? i : cpd - 1;
}

}  // namespace

__global__ void k_soa(const double* __restrict__ xyz, int n, double* __restrict__ x,
                      double* __restrict__ y, double* __restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = xyz[3 * (size_t)i];
  y[i] = xyz[3 * (size_t)i + 1];
  z[i] = xyz[3 * (size_t)i + 2];
}

// Per-block min/max of the three axes -> part[block][6].
__global__ __launch_bounds__(256) void k_bounds_partial(const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ z, int n,
                                                        double* __restrict__ part) {
  __shared__ double sm[6][256];
  const int t = threadIdx.x;
  double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
  for (int i = blockIdx.x * 256 + t; i < n; i += gridDim.x * 256) {
    const double v[3] = {x[i], y[i], z[i]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = v[a] < lo[a] ? v[a] : lo[a];
      hi[a] = v[a] > hi[a] ? v[a] : hi[a];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    sm[a][t] = lo[a];
    sm[3 + a][t] = hi[a];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        sm[a][t] = sm[a][t + s] < sm[a][t] ? sm[a][t + s] : sm[a][t];
        sm[3 + a][t] = sm[3 + a][t + s] > sm[3 + a][t] ? sm[3 + a][t + s] : sm[3 + a][t];
      }
    }
    __syncthreads();
  }
  if (t < 6) part[blockIdx.x * 6 + t] = sm[t][0];
}

// Reduce the partials (one wavefront); cell sizes as celery.rs:164-170:
// (max-min)/cpd and cpd/(max-min).
__global__ __launch_bounds__(64) void k_bounds_final(const double* __restrict__ part, int nparts, int cpd,
                                                     GridDev* __restrict__ g) {
  const int t = threadIdx.x;
  double lo[3] = {part[0], part[1], part[2]}, hi[3] = {part[3], part[4], part[5]};
  for (int b = t; b < nparts; b += 64)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double l = part[6 * b + a], h = part[6 * b + 3 + a];
      lo[a] = l < lo[a] ? l : lo[a];
      hi[a] = h > hi[a] ? h : hi[a];
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double l = __shfl_xor(lo[a], o, 64), h = __shfl_xor(hi[a], o, 64);
      lo[a] = l < lo[a] ? l : lo[a];
      hi[a] = h > hi[a] ? h : hi[a];
    }
  if (t != 0) return;
  const double c = (double)cpd;
  g->xmin = lo[0];
  g->xmax = hi[0];
  g->ymin = lo[1];
  g->ymax = hi[1];
  g->zmin = lo[2];
  g->zmax = hi[2];
  g->sx = (hi[0] - lo[0]) / c;
  g->sy = (hi[1] - lo[1]) / c;
  g->sz = (hi[2] - lo[2]) / c;
  g->ix = c / (hi[0] - lo[0]);
  g->iy = c / (hi[1] - lo[1]);
  g->iz = c / (hi[2] - lo[2]);
  g->cpd = cpd;
  g->pad = 0;
}

// Cell id of every point; histogram with one atomic per run of equal ids
// among consecutive lanes (spatially ordered inputs collapse to few atomics).
// rank (nullable): each point's slot inside its cell's range, from the value
// the run's atomic returns (a permutation of 0 .. count - 1 per cell, in no
// set order: the in-cell sort fixes the order afterwards).
__global__ __launch_bounds__(256) void k_bin_count(const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z, int n,
                                                   const GridDev* __restrict__ gp,
                                                   uint32_t* __restrict__ bins,
                                                   uint32_t* __restrict__ cnt,
                                                   uint32_t* __restrict__ rank) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const GridDev g = *gp;
  const int lane = __lane_id();
  uint32_t b = 0xffffffffu;
  if (i < n) {
    const int a = axis_index(x[i], g.xmin, g.xmax, g.ix, g.cpd);
    const int c = axis_index(y[i], g.ymin, g.ymax, g.iy, g.cpd);
    const int d = axis_index(z[i], g.zmin, g.zmax, g.iz, g.cpd);
    b = (uint32_t)((a * g.cpd + c) * g.cpd + d);
    bins[i] = b;
  }
  const uint32_t prev = __shfl_up(b, 1, 64);
  const bool head = (i < n) && (lane == 0 || prev != b);
  // both ballots are taken by the whole wavefront (a ballot inside the
  // branch below would only see the head lanes)
  const uint64_t heads = __ballot(head);
  const uint64_t act = __ballot(i < n);
  uint32_t first = 0;  // the run's first slot (head lanes)
  if (head) {
    const uint64_t after = lane == 63 ? 0ull : (heads >> (lane + 1));
    const int last = 63 - __clzll((long long)act);  // last active lane
    const int nexth = after ? lane + 1 + (__ffsll((unsigned long long)after) - 1) : last + 1;
    first = atomicAdd(&cnt[b], (uint32_t)(nexth - lane));
  }
  if (rank) {
    // the run's head: the highest head lane at or below this one (an active
    // lane always has one: the first active lane is a head)
    const uint64_t upto = lane == 63 ? heads : (heads & ((2ull << lane) - 1ull));
    const int h = upto ? 63 - __clzll((long long)upto) : lane;
    const uint32_t f = (uint32_t)__shfl((int)first, h, 64);
    if (i < n) rank[i] = f + (uint32_t)(lane - h);
  }
}

// Three-phase exclusive scan of m counts (u32).  Block = 1024 elements.
__global__ __launch_bounds__(256) void k_scan_block(const uint32_t* __restrict__ in, int m,
                                                    uint32_t* __restrict__ out,
                                                    uint32_t* __restrict__ sums) {
  __shared__ uint32_t wsum[4];
  const int t = threadIdx.x, base = blockIdx.x * 1024;
  uint32_t v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int idx = base + 4 * t + k;
    v[k] = idx < m ? in[idx] : 0u;
    s += v[k];
  }
  // exclusive scan of per-thread sums across the block
  const int lane = t & 63, wid = t >> 6;
  uint32_t incl = s;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  uint32_t woff = 0;
  for (int w = 0; w < wid; ++w) woff += wsum[w];
  uint32_t run = woff + incl - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int idx = base + 4 * t + k;
    if (idx < m) out[idx] = run;
    run += v[k];
  }
  if (t == 255) sums[blockIdx.x] = woff + incl;
}

// Exclusive scan of the block sums in place (one block, sequential chunks).
__global__ __launch_bounds__(256) void k_scan_sums(uint32_t* __restrict__ sums, int nb,
                                                   uint32_t* __restrict__ total) {
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t carry;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 256) {
    const int idx = base + t;
    const uint32_t v = idx < nb ? sums[idx] : 0u;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = carry;
    for (int w = 0; w < wid; ++w) woff += wsum[w];
    if (idx < nb) sums[idx] = woff + incl - v;
    __syncthreads();
    if (t == 255) carry = woff + incl;
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, int m, const uint32_t* __restrict__ sums,
                           const uint32_t* __restrict__ total, uint32_t* __restrict__ last) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] += sums[i >> 10];
  if (i == 0 && last) *last = *total;
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, int m, uint32_t* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) dst[i] = src[i];
}

__global__ void k_scatter(const uint32_t* __restrict__ bins, int n, uint32_t* __restrict__ cursor,
                          uint32_t* __restrict__ sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t pos = atomicAdd(&cursor[bins[i]], 1u);
  sorted[pos] = (uint32_t)i;
}

// The same with the slots k_bin_count handed out: no atomics.
__global__ void k_scatter_ranked(const uint32_t* __restrict__ bins, const uint32_t* __restrict__ rank, int n,
                                 const uint32_t* __restrict__ delim, uint32_t* __restrict__ sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sorted[delim[bins[i]] + rank[i]] = (uint32_t)i;
}

// Ascending point index inside each cell: insertion sort for small cells,
// one wavefront per cell (rank by comparison) for large ones.
// Order of two points of a cell: (key, index) when keys are given (the
// internal Morton order), else the index.
__device__ __forceinline__ bool before(const uint32_t* __restrict__ key, uint32_t a, uint32_t b) {
  if (key) {
    const uint32_t ka = key[a], kb = key[b];
    if (ka != kb) return ka < kb;
  }
  return a < b;
}

__global__ void k_sort_cells(const uint32_t* __restrict__ delim, int nbins,
                             uint32_t* __restrict__ sorted, uint32_t* __restrict__ big,
                             uint32_t* __restrict__ nbig, const uint32_t* __restrict__ key) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  const uint32_t s = delim[b], e = delim[b + 1];
  if (e - s <= 1) return;
  if (e - s > 32) {
    const uint32_t k = atomicAdd(nbig, 1u);
    big[k] = (uint32_t)b;
    return;
  }
  for (uint32_t i = s + 1; i < e; ++i) {
    const uint32_t v = sorted[i];
    uint32_t j = i;
    while (j > s && before(key, v, sorted[j - 1])) {
      sorted[j] = sorted[j - 1];
      --j;
    }
    sorted[j] = v;
  }
}

// Large cells: rank of each element = number of elements before it in the
// cell's order (ids are distinct).  A cell of at most kSortLds points is
// staged in LDS (ids and keys) and ranked against the staged copy, each
// element written straight to its rank; a larger one is ranked against
// global memory into a scratch copy, then copied back.
constexpr uint32_t kSortLds = 4096;
__device__ void rank_sort_cell(uint32_t* __restrict__ sorted, uint32_t* __restrict__ scratch, uint32_t s,
                               uint32_t e, const uint32_t* __restrict__ key, uint32_t* sid, uint32_t* skey) {
  const uint32_t m = e - s;
  if (m <= kSortLds) {
    const uint32_t m4 = (m + 3) >> 2;
    for (uint32_t i = threadIdx.x; i < 4 * m4; i += blockDim.x) {
      // past the cell: the largest id, which no id of the cell exceeds
      const uint32_t v = i < m ? sorted[s + i] : 0xffffffffu;
      sid[i] = v;
      if (key && i < m) skey[i] = key[v];
    }
    __syncthreads();
    if (key) {
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t v = sid[i], kv = skey[i];
        uint32_t r = 0;
        for (uint32_t j = 0; j < m; ++j) {
          const uint32_t kj = skey[j];
          r += (kj < kv || (kj == kv && sid[j] < v)) ? 1u : 0u;
        }
        sorted[s + r] = v;
      }
    } else {
      // two elements per thread and pass, four staged ids per LDS read
      const uint4* s4 = reinterpret_cast<const uint4*>(sid);
      for (uint32_t i = threadIdx.x; i < m; i += 2 * blockDim.x) {
        const uint32_t i1 = i + blockDim.x;
        const uint32_t v0 = sid[i], v1 = i1 < m ? sid[i1] : 0u;
        uint32_t r0 = 0, r1 = 0;
        for (uint32_t j = 0; j < m4; ++j) {
          const uint4 w = s4[j];
          r0 += (w.x < v0 ? 1u : 0u) + (w.y < v0 ? 1u : 0u) + (w.z < v0 ? 1u : 0u) + (w.w < v0 ? 1u : 0u);
          r1 += (w.x < v1 ? 1u : 0u) + (w.y < v1 ? 1u : 0u) + (w.z < v1 ? 1u : 0u) + (w.w < v1 ? 1u : 0u);
        }
        sorted[s + r0] = v0;
        if (i1 < m) sorted[s + r1] = v1;
      }
    }
    __syncthreads();
    return;
  }
  for (uint32_t i = s + threadIdx.x; i < e; i += blockDim.x) {
    const uint32_t v = sorted[i];
    uint32_t r = 0;
    for (uint32_t j = s; j < e; ++j) r += before(key, sorted[j], v);
    scratch[s + r] = v;
  }
  __syncthreads();
  for (uint32_t i = s + threadIdx.x; i < e; i += blockDim.x) sorted[i] = scratch[i];
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_sort_big(const uint32_t* __restrict__ delim,
                                                  const uint32_t* __restrict__ big,
                                                  const uint32_t* __restrict__ nbig,
                                                  uint32_t* __restrict__ sorted,
                                                  uint32_t* __restrict__ scratch,
                                                  const uint32_t* __restrict__ key) {
  __shared__ __align__(16) uint32_t sid[kSortLds];
  __shared__ uint32_t skey[kSortLds];
  const uint32_t nb = *nbig;
  for (uint32_t q = blockIdx.x; q < nb; q += gridDim.x) {
    const uint32_t b = big[q];
    rank_sort_cell(sorted, scratch, delim[b], delim[b + 1], key, sid, skey);
  }
}

__global__ void k_gather(const uint32_t* __restrict__ sorted, int n, const double* __restrict__ x,
                         const double* __restrict__ y, const double* __restrict__ z,
                         double* __restrict__ sx, double* __restrict__ sy, double* __restrict__ sz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = sorted[k];
  sx[k] = x[i];
  sy[k] = y[i];
  sz[k] = z[i];
}

// ---- many sets at once (a batch): the same passes over every set's
// points and grid cells in one launch each, a thread finding its set by a
// binary search over the sets' first entries (base[nsets] = total).
template <class T>
__device__ __forceinline__ int set_of(const T* __restrict__ base, int nsets, uint64_t g) {
  int lo = 0, hi = nsets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((uint64_t)base[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Bounding boxes of every set: block q takes part q - pblk[set] of its set
// (grid-stride over the set's points, as k_bounds_partial does for one set)
// -> part[q][6]; then one wavefront per set reduces its parts and writes its
// grid (as k_bounds_final).
__global__ __launch_bounds__(256) void k_seg_bounds_partial(SegBinning a) {
  __shared__ double sm[6][256];
  const int t = threadIdx.x;
  const uint32_t q = blockIdx.x;
  const int set = set_of(a.pblk, a.nsets, q);
  const uint32_t r = q - a.pblk[set], nparts = a.pblk[set + 1] - a.pblk[set];
  const uint64_t p0 = a.pbase[set];
  const uint32_t n = (uint32_t)(a.pbase[set + 1] - p0);
  const double* x = a.x + p0;
  const double* y = a.y + p0;
  const double* z = a.z + p0;
  double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
  for (uint32_t i = r * 256 + t; i < n; i += nparts * 256) {
    const double v[3] = {x[i], y[i], z[i]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = v[k] < lo[k] ? v[k] : lo[k];
      hi[k] = v[k] > hi[k] ? v[k] : hi[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    sm[k][t] = lo[k];
    sm[3 + k][t] = hi[k];
  }
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sm[k][t] = sm[k][t + h] < sm[k][t] ? sm[k][t + h] : sm[k][t];
        sm[3 + k][t] = sm[3 + k][t + h] > sm[3 + k][t] ? sm[3 + k][t + h] : sm[3 + k][t];
      }
    }
    __syncthreads();
  }
  if (t < 6) a.part[6 * (size_t)q + t] = sm[t][0];
}

__global__ __launch_bounds__(64) void k_seg_bounds_final(SegBinning a) {
  const int t = threadIdx.x, set = blockIdx.x;
  const double* part = a.part + 6 * (size_t)a.pblk[set];
  const int nparts = (int)(a.pblk[set + 1] - a.pblk[set]);
  double lo[3] = {part[0], part[1], part[2]}, hi[3] = {part[3], part[4], part[5]};
  for (int b = t; b < nparts; b += 64)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double l = part[6 * b + k], h = part[6 * b + 3 + k];
      lo[k] = l < lo[k] ? l : lo[k];
      hi[k] = h > hi[k] ? h : hi[k];
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double l = __shfl_xor(lo[k], o, 64), h = __shfl_xor(hi[k], o, 64);
      lo[k] = l < lo[k] ? l : lo[k];
      hi[k] = h > hi[k] ? h : hi[k];
    }
  if (t != 0) return;
  GridDev* g = a.grid + set;
  const double c = (double)a.cpd[set];
  g->xmin = lo[0];
  g->xmax = hi[0];
  g->ymin = lo[1];
  g->ymax = hi[1];
  g->zmin = lo[2];
  g->zmax = hi[2];
  g->sx = (hi[0] - lo[0]) / c;
  g->sy = (hi[1] - lo[1]) / c;
  g->sz = (hi[2] - lo[2]) / c;
  g->ix = c / (hi[0] - lo[0]);
  g->iy = c / (hi[1] - lo[1]);
  g->iz = c / (hi[2] - lo[2]);
  g->cpd = a.cpd[set];
  g->pad = 0;
}

// Cell ids and the histogram of every set (k_bin_count over the batch; a run
// of equal ids never crosses sets, whose cells are numbered apart).
__global__ __launch_bounds__(256) void k_seg_bin_count(SegBinning a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = __lane_id();
  uint32_t b = 0xffffffffu;  // cell id over the batch
  if (i < a.n) {
    const int set = set_of(a.pbase, a.nsets, i);
    const GridDev g = a.grid[set];
    const int u = axis_index(a.x[i], g.xmin, g.xmax, g.ix, g.cpd);
    const int v = axis_index(a.y[i], g.ymin, g.ymax, g.iy, g.cpd);
    const int w = axis_index(a.z[i], g.zmin, g.zmax, g.iz, g.cpd);
    const uint32_t local = (uint32_t)((u * g.cpd + v) * g.cpd + w);
    a.bins[i] = local;
    b = a.cbase[set] + local;
  }
  const uint32_t prev = __shfl_up(b, 1, 64);
  const bool head = (i < a.n) && (lane == 0 || prev != b);
  const uint64_t heads = __ballot(head);
  const uint64_t act = __ballot(i < a.n);
  uint32_t first = 0;
  if (head) {
    const uint64_t after = lane == 63 ? 0ull : (heads >> (lane + 1));
    const int last = 63 - __clzll((long long)act);
    const int nexth = after ? lane + 1 + (__ffsll((unsigned long long)after) - 1) : last + 1;
    first = atomicAdd(&a.cnt[b], (uint32_t)(nexth - lane));
  }
  const uint64_t upto = lane == 63 ? heads : (heads & ((2ull << lane) - 1ull));
  const int h = upto ? 63 - __clzll((long long)upto) : lane;
  const uint32_t f = (uint32_t)__shfl((int)first, h, 64);
  if (i < a.n) a.rank[i] = f + (uint32_t)(lane - h);
}

// Each set's delimiters from the scan over the whole batch's cells: the
// scan at a set's cell counts every point of the sets before it too.
__global__ void k_seg_delim(SegBinning a, uint32_t ndelim) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ndelim) return;
  const int set = set_of(a.bbase, a.nsets, j);
  const uint32_t b = j - a.bbase[set], nbs = a.cbase[set + 1] - a.cbase[set];
  const uint32_t p0 = (uint32_t)a.pbase[set];
  a.delim[j] = b < nbs ? a.gscan[a.cbase[set] + b] - p0 : (uint32_t)a.pbase[set + 1] - p0;
}

__global__ void k_seg_scatter(SegBinning a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int set = set_of(a.pbase, a.nsets, i);
  const uint32_t p0 = (uint32_t)a.pbase[set];
  a.sorted[p0 + a.delim[a.bbase[set] + a.bins[i]] + a.rank[i]] = i - p0;
}

// In-cell order of every set's cells (k_sort_cells over the batch); larger
// cells are listed by batch cell id for k_seg_sort_big.
__global__ void k_seg_sort_cells(SegBinning a) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.nb) return;
  const int set = set_of(a.cbase, a.nsets, j);
  const uint32_t* d = a.delim + a.bbase[set] + (j - a.cbase[set]);
  const uint32_t s = d[0], e = d[1];
  if (e - s <= 1) return;
  if (e - s > 32) {
    a.big[atomicAdd(&a.misc[1], 1u)] = j;
    return;
  }
  uint32_t* srt = a.sorted + a.pbase[set];
  for (uint32_t i = s + 1; i < e; ++i) {
    const uint32_t v = srt[i];
    uint32_t k = i;
    while (k > s && v < srt[k - 1]) {
      srt[k] = srt[k - 1];
      --k;
    }
    srt[k] = v;
  }
}

__global__ __launch_bounds__(256) void k_seg_sort_big(SegBinning a) {
  __shared__ __align__(16) uint32_t sid[kSortLds];
  __shared__ uint32_t skey[kSortLds];
  const uint32_t nbig = a.misc[1];
  for (uint32_t q = blockIdx.x; q < nbig; q += gridDim.x) {
    const uint32_t j = a.big[q];
    const int set = set_of(a.cbase, a.nsets, j);
    const uint32_t* d = a.delim + a.bbase[set] + (j - a.cbase[set]);
    rank_sort_cell(a.sorted + a.pbase[set], a.rank + a.pbase[set], d[0], d[1], nullptr, sid, skey);
  }
}

__global__ void k_seg_gather(SegBinning a) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const uint64_t o = a.pbase[set_of(a.pbase, a.nsets, k)] + a.sorted[k];
  a.sx[k] = a.x[o];
  a.sy[k] = a.y[o];
  a.sz[k] = a.z[o];
}

// Exclusive scan of face counts into u64 offsets: block scan in u64.
__global__ __launch_bounds__(256) void k_scan64_block(const uint32_t* __restrict__ in, int m,
                                                      uint64_t* __restrict__ out,
                                                      uint64_t* __restrict__ sums) {
  __shared__ uint64_t wsum[4];
  const int t = threadIdx.x, base = blockIdx.x * 1024, lane = t & 63, wid = t >> 6;
  uint64_t v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int idx = base + 4 * t + k;
    v[k] = idx < m ? (uint64_t)in[idx] : 0ull;
    s += v[k];
  }
  uint64_t incl = s;
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  uint64_t woff = 0;
  for (int w = 0; w < wid; ++w) woff += wsum[w];
  uint64_t run = woff + incl - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int idx = base + 4 * t + k;
    if (idx < m) out[idx] = run;
    run += v[k];
  }
  if (t == 255) sums[blockIdx.x] = woff + incl;
}

__global__ __launch_bounds__(256) void k_scan64_sums(uint64_t* __restrict__ sums, int nb) {
  __shared__ uint64_t wsum[4];
  __shared__ uint64_t carry;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 256) {
    const int idx = base + t;
    const uint64_t v = idx < nb ? sums[idx] : 0ull;
    uint64_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      uint64_t u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint64_t woff = carry;
    for (int w = 0; w < wid; ++w) woff += wsum[w];
    if (idx < nb) sums[idx] = woff + incl - v;
    __syncthreads();
    if (t == 255) carry = woff + incl;
    __syncthreads();
  }
}

__global__ void k_scan64_add(uint64_t* __restrict__ out, int m, const uint64_t* __restrict__ sums,
                             const uint32_t* __restrict__ in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] += sums[i >> 10];
  if (i == m - 1) out[m] = out[i] + in[i];
}

// Staged faces -> CSR: 16 lanes per cell.
__global__ __launch_bounds__(256) void k_compact(const uint32_t* __restrict__ fcount,
                                                 const uint64_t* __restrict__ foff, int n,
                                                 const int32_t* __restrict__ snbr,
                                                 const double* __restrict__ sarea,
                                                 int32_t* __restrict__ nbr, double* __restrict__ area,
                                                 const uint64_t* __restrict__ fpos) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int cell = g >> 4, sub = g & 15;
  if (cell >= n) return;
  const uint32_t fc = fcount[cell];
  const uint64_t o = foff[cell];
  const uint64_t src = fpos ? (fc ? fpos[cell] : 0) : (uint64_t)cell * kFaceStage;
  for (uint32_t f = sub; f < fc; f += 16) {
    nbr[o + f] = snbr[src + f];
    area[o + f] = sarea[src + f];
  }
}

}  // namespace tess

namespace tess {

namespace {
__global__ void k_invert(const uint32_t* __restrict__ sorted, int n, uint32_t* __restrict__ inv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) inv[sorted[k]] = (uint32_t)k;
}
inline int blocks(long long n, int t) { return (int)((n + t - 1) / t); }
}  // namespace

hipError_t launch_soa(const double* xyz, int n, double* x, double* y, double* z, hipStream_t s) {
  hipLaunchKernelGGL(k_soa, dim3(blocks(n, 256)), dim3(256), 0, s, xyz, n, x, y, z);
  return hipGetLastError();
}

hipError_t launch_bounds(const double* x, const double* y, const double* z, int n, int cpd,
                         double* part, int nparts, GridDev* g, hipStream_t s) {
  hipLaunchKernelGGL(k_bounds_partial, dim3(nparts), dim3(256), 0, s, x, y, z, n, part);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, s, part, nparts, cpd, g);
  return hipGetLastError();
}

hipError_t launch_counting_sort(const double* x, const double* y, const double* z, int n,
                                const GridDev* g, int nbins, const SortBufs& b, hipStream_t s,
                                const uint32_t* key) {
  // fill kernels rather than memset nodes: this sequence may be captured
  // into a graph
  hipError_t e;
  if ((e = launch_fill_u32(b.cnt, (size_t)nbins, 0u, s)) != hipSuccess) return e;
  if ((e = launch_fill_u32(b.misc, 4, 0u, s)) != hipSuccess) return e;
  // the ranks go to b.scratch, which the in-cell sort reuses once the
  // scatter has read them
  hipLaunchKernelGGL(k_bin_count, dim3(blocks(n, 256)), dim3(256), 0, s, x, y, z, n, g, b.bins, b.cnt, b.scratch);
  const int nsb = blocks(nbins, 1024);
  hipLaunchKernelGGL(k_scan_block, dim3(nsb), dim3(256), 0, s, b.cnt, nbins, b.delim, b.sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, b.sums, nsb, b.misc);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks(nbins, 256)), dim3(256), 0, s, b.delim, nbins, b.sums,
                     b.misc, b.delim + nbins);
  hipLaunchKernelGGL(k_scatter_ranked, dim3(blocks(n, 256)), dim3(256), 0, s, b.bins, b.scratch, n, b.delim,
                     b.sorted);
  hipLaunchKernelGGL(k_sort_cells, dim3(blocks(nbins, 256)), dim3(256), 0, s, b.delim, nbins, b.sorted,
                     b.big, b.misc + 1, key);
  hipLaunchKernelGGL(k_sort_big, dim3(1024), dim3(256), 0, s, b.delim, b.big, b.misc + 1, b.sorted,
                     b.scratch, key);
  return hipGetLastError();
}

hipError_t launch_seg_binning(const SegBinning& a, hipStream_t s) {
  if (a.nsets <= 0 || a.n == 0) return hipSuccess;
  // fill kernels rather than memset nodes: this sequence may be captured
  hipError_t e;
  if ((e = launch_fill_u32(a.cnt, a.nb, 0u, s)) != hipSuccess || (e = launch_fill_u32(a.misc, 2, 0u, s)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_seg_bounds_partial, dim3(a.nparts), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_seg_bounds_final, dim3(a.nsets), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_seg_bin_count, dim3(blocks(a.n, 256)), dim3(256), 0, s, a);
  if ((e = launch_scan_u32(a.cnt, (int)a.nb, a.gscan, a.sums, a.misc, s)) != hipSuccess) return e;
  const uint32_t ndelim = a.nb + (uint32_t)a.nsets;
  hipLaunchKernelGGL(k_seg_delim, dim3(blocks(ndelim, 256)), dim3(256), 0, s, a, ndelim);
  hipLaunchKernelGGL(k_seg_scatter, dim3(blocks(a.n, 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_seg_sort_cells, dim3(blocks(a.nb, 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_seg_sort_big, dim3(1024), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_seg_gather, dim3(blocks(a.n, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_grid_order(const double* x, const double* y, const double* z, int n, const GridDev* g,
                             int nbins, const SortBufs& b, hipStream_t s) {
  hipError_t e;
  if ((e = launch_fill_u32(b.cnt, (size_t)nbins, 0u, s)) != hipSuccess) return e;
  if ((e = launch_fill_u32(b.misc, 4, 0u, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_bin_count, dim3(blocks(n, 256)), dim3(256), 0, s, x, y, z, n, g, b.bins, b.cnt, nullptr);
  const int nsb = blocks(nbins, 1024);
  hipLaunchKernelGGL(k_scan_block, dim3(nsb), dim3(256), 0, s, b.cnt, nbins, b.delim, b.sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, b.sums, nsb, b.misc);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks(nbins, 256)), dim3(256), 0, s, b.delim, nbins, b.sums,
                     b.misc, b.delim + nbins);
  hipLaunchKernelGGL(k_copy_u32, dim3(blocks(nbins, 256)), dim3(256), 0, s, b.delim, nbins, b.cursor);
  hipLaunchKernelGGL(k_scatter, dim3(blocks(n, 256)), dim3(256), 0, s, b.bins, n, b.cursor, b.sorted);
  return hipGetLastError();
}

hipError_t launch_scan_u32(const uint32_t* in, int m, uint32_t* out, uint32_t* sums, uint32_t* total,
                           hipStream_t s) {
  if (m <= 0) return hipSuccess;
  const int nsb = blocks(m, 1024);
  hipLaunchKernelGGL(k_scan_block, dim3(nsb), dim3(256), 0, s, in, m, out, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, sums, nsb, total);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks(m, 256)), dim3(256), 0, s, out, m, sums, total, nullptr);
  return hipGetLastError();
}

hipError_t launch_gather(const uint32_t* sorted, int n, const double* x, const double* y,
                         const double* z, double* sx, double* sy, double* sz, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(blocks(n, 256)), dim3(256), 0, s, sorted, n, x, y, z, sx, sy, sz);
  return hipGetLastError();
}

hipError_t launch_face_offsets(const uint32_t* fcount, int n, uint64_t* foff, uint64_t* sums,
                               hipStream_t s) {
  const int nsb = blocks(n, 1024);
  hipLaunchKernelGGL(k_scan64_block, dim3(nsb), dim3(256), 0, s, fcount, n, foff, sums);
  hipLaunchKernelGGL(k_scan64_sums, dim3(1), dim3(256), 0, s, sums, nsb);
  hipLaunchKernelGGL(k_scan64_add, dim3(blocks(n, 256)), dim3(256), 0, s, foff, n, sums, fcount);
  return hipGetLastError();
}

hipError_t launch_compact(const uint32_t* fcount, const uint64_t* foff, int n, const int32_t* snbr,
                          const double* sarea, int32_t* nbr, double* area, hipStream_t s, const uint64_t* fpos) {
  hipLaunchKernelGGL(k_compact, dim3(blocks((long long)n * 16, 256)), dim3(256), 0, s, fcount, foff, n,
                     snbr, sarea, nbr, area, fpos);
  return hipGetLastError();
}

hipError_t launch_invert(const uint32_t* sorted, int n, uint32_t* inv, hipStream_t s) {
  hipLaunchKernelGGL(k_invert, dim3(blocks(n, 256)), dim3(256), 0, s, sorted, n, inv);
  return hipGetLastError();
}

}  // namespace tess
