// query.hip -- radius queries over a diagram's grid (celery.rs:700-855,
// :1023-1075; interface.rs:348-365).
//
// One wavefront per query, two passes over the same walk: the first counts
// each query's results, a u64 scan turns the counts into CSR offsets, the
// second writes the ids in the reference's order.  Box queries walk the index
// box of the query in (i, j, k) loop order, 64 grid cells per step; the
// neighbour cloud walks the shell table in canonical order, 64 entries per
// step, until the first key above the radius.  The points of a step's cells
// are then walked one per lane, so loads and stores of neighbouring lanes
// are neighbouring addresses and a dense cell does not hold up the wave.
// Large query sets are taken in grid order of their positions, so the
// wavefronts running together read the same grid cells.  The walk is bound
// by the reads of delimiters, sorted indices and (for the real radius)
// gathered coordinates: integer and byte work on HBM.
#include "launch.h"
#include "tess_internal.h"

namespace tess {

namespace {

__device__ __forceinline__ int q_axis_index(double v, double vmin, double vmax, double inv, int cpd) {
  if (v >= vmax) return cpd - 1;  // celery.rs:269-282
  double t = (v - vmin) * inv;
  if (!(t > 0.0)) return 0;       // saturating cast, NaN -> 0
  if (t >= 2147483647.0) return cpd - 1;
  int i = (int)t;
  return i < cpd - 1 ? i : cpd - 1;
}

// celery.rs:682-697: the first argument only when strictly smaller / larger.
__device__ __forceinline__ double min_float(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double max_float(double a, double b) { return a > b ? a : b; }

__device__ __forceinline__ int cell_gap(int a, int b) {
  const int d = (a > b ? a - b : b - a) - 1;
  return d > 0 ? d : 0;
}

// Within one wavefront LDS operations execute in issue order: a phase
// boundary only has to keep the compiler from moving LDS accesses across it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Running maximum over the lanes (inclusive), values >= -1: DPP row shifts
// inside the rows of 16, then row broadcasts (lanes without a source: -1).
__device__ __forceinline__ int wave_max_scan(int v) {
  int t;
  t = __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xf, 0xf, false); v = t > v ? t : v;  // row_shr:1
  t = __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xf, 0xf, false); v = t > v ? t : v;  // row_shr:2
  t = __builtin_amdgcn_update_dpp(-1, v, 0x114, 0xf, 0xf, false); v = t > v ? t : v;  // row_shr:4
  t = __builtin_amdgcn_update_dpp(-1, v, 0x118, 0xf, 0xf, false); v = t > v ? t : v;  // row_shr:8
  t = __builtin_amdgcn_update_dpp(-1, v, 0x142, 0xa, 0xf, false); v = t > v ? t : v;  // row_bcast:15
  t = __builtin_amdgcn_update_dpp(-1, v, 0x143, 0xc, 0xf, false); v = t > v ? t : v;  // row_bcast:31
  return v;
}

// Exclusive prefix of v over the wavefront; *total = the sum.  DPP row
// shifts inside each row of 16 lanes, then the row totals broadcast by
// row_bcast:15 / :31 (lanes without a source add 0).
__device__ __forceinline__ uint32_t wave_excl(uint32_t v, int lane, uint32_t* total) {
  (void)lane;
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 -> rows 1, 3
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 -> rows 2, 3
  *total = (uint32_t)__builtin_amdgcn_readlane(x, 63);
  return (uint32_t)x - v;
}

// The points of the step's 64 grid cells, in cell order and, inside a cell,
// in grid (sorted-index) order: lane l's cell holds len points from grid
// position s0 on (len 0: no cell).  The concatenation is walked 64 points at
// a time, one point per lane; a lane finds its cell by binary search over the
// cells' exclusive starts, read from their lanes.  MODE 0 keeps every point
// (only the count is needed in the count pass), MODE 1 the points within the
// radius (celery.rs:700-702, :843-851), MODE 2 the points of the target group
// (interface.rs:358-361); kept points are written in walk order (ballot rank).
template <int MODE, bool FILL>
__device__ __forceinline__ void emit_points(const QueryArgs& a, uint32_t s0, uint32_t len, int lane, double qx,
                                            double qy, double qz, double r2, uint64_t& cursor, uint64_t end,
                                            uint8_t* head) {
  uint32_t total;
  const uint32_t start = wave_excl(len, lane, &total);
  if (MODE == 0 && !FILL) {
    cursor += total;
    return;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  int carry = 0;  // the cell holding the last position of the previous window
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t p = base + (uint32_t)lane;
    // the cell of position p: the last non-empty cell starting at or before
    // p (each writes its lane at its start in this window, then a running
    // maximum; before the window's first start, the carried cell)
    head[lane] = 0xff;
    wave_sync();
    if (len > 0 && start >= base && start - base < 64u) head[start - base] = (uint8_t)lane;
    wave_sync();
    const int h = head[lane];
    int c = wave_max_scan(h == 0xff ? -1 : h);
    c = c > carry ? c : carry;
    carry = __builtin_amdgcn_readlane(c, 63);
    wave_sync();
    const uint32_t cs0 = __shfl(s0, c, 64), cstart = __shfl(start, c, 64);
    const bool valid = p < total;
    const uint32_t s = cs0 + (p - cstart);
    uint32_t id = 0;
    bool keep = false;
    if (valid) {
      id = a.sorted[s];
      if (MODE == 0) {
        keep = true;
      } else if (MODE == 1) {
        const double dx = qx - a.sx[s], dy = qy - a.sy[s], dz = qz - a.sz[s];
        keep = (dx * dx + dy * dy) + dz * dz <= r2;
      } else {
        keep = a.groups[id] == a.target_group;
      }
    }
    if (MODE == 0) {
      if (valid && cursor + p < end) a.ids[cursor + p] = id;
    } else {
      const uint64_t m = __ballot(keep);
      if (FILL && keep) {
        const uint64_t pos = cursor + (uint64_t)__popcll(m & below);
        if (pos < end) a.ids[pos] = id;
      }
      cursor += (uint64_t)__popcll(m);
    }
  }
  if (MODE == 0) cursor += total;
}

// find_cells_in_radius (KIND 0) and the two neighbour searches built on it
// (KIND 1 cell radius, KIND 2 real radius): celery.rs:753-855.
template <int KIND, bool FILL>
__global__ __launch_bounds__(256) void k_query_box(QueryArgs a) {
  __shared__ uint8_t heads[4][64];
  uint8_t* head = heads[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= a.nq) return;
  const int qi = a.order ? (int)a.order[w] : w;
  const GridDev& G = a.g;
  const int cpd = G.cpd;
  const double qx = a.q[3 * (size_t)qi], qy = a.q[3 * (size_t)qi + 1], qz = a.q[3 * (size_t)qi + 2];
  const double r = a.radius[(size_t)qi * a.rstride];
  const int i0 = q_axis_index(max_float(qx - r, G.xmin), G.xmin, G.xmax, G.ix, cpd);
  const int j0 = q_axis_index(max_float(qy - r, G.ymin), G.ymin, G.ymax, G.iy, cpd);
  const int k0 = q_axis_index(max_float(qz - r, G.zmin), G.zmin, G.zmax, G.iz, cpd);
  const int i1 = q_axis_index(min_float(qx + r, G.xmax), G.xmin, G.xmax, G.ix, cpd);
  const int j1 = q_axis_index(min_float(qy + r, G.ymax), G.ymin, G.ymax, G.iy, cpd);
  const int k1 = q_axis_index(min_float(qz + r, G.zmax), G.zmin, G.zmax, G.iz, cpd);
  const int hx = q_axis_index(qx, G.xmin, G.xmax, G.ix, cpd);
  const int hy = q_axis_index(qy, G.ymin, G.ymax, G.iy, cpd);
  const int hz = q_axis_index(qz, G.zmin, G.zmax, G.iz, cpd);
  const double r2 = r * r;
  const int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
  const uint64_t njk = (uint64_t)(nj > 0 ? nj : 0) * (uint64_t)(nk > 0 ? nk : 0);
  // a NaN radius admits no cell (ds <= NaN is false): skip the walk
  const uint64_t total = (ni > 0 && r == r) ? (uint64_t)ni * njk : 0;
  uint64_t cursor = FILL ? a.offsets[qi] : 0;
  const uint64_t end = FILL ? a.offsets[qi + 1] : 0;
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint64_t base = 0; base < total; base += 64) {
    const uint64_t idx = base + (uint64_t)lane;
    bool in = false;
    uint32_t b = 0;
    if (idx < total) {
      const uint64_t rem = idx % njk;
      const int i = i0 + (int)(idx / njk), j = j0 + (int)(rem / (uint64_t)nk), k = k0 + (int)(rem % (uint64_t)nk);
      // celery.rs:708-743: gap distance of the query's cell to (i, j, k)
      const double gx = (double)cell_gap(hx, i) * G.sx, gy = (double)cell_gap(hy, j) * G.sy,
                   gz = (double)cell_gap(hz, k) * G.sz;
      in = (gx * gx + gy * gy) + gz * gz <= r2;
      b = ((uint32_t)i * (uint32_t)cpd + (uint32_t)j) * (uint32_t)cpd + (uint32_t)k;  // celery.rs:317-325
    }
    if (KIND == 0) {
      const uint64_t m = __ballot(in);
      if (FILL && in) {
        const uint64_t pos = cursor + (uint64_t)__popcll(m & below);
        if (pos < end) a.ids[pos] = b;
      }
      cursor += (uint64_t)__popcll(m);
    } else {
      uint32_t s0 = 0, len = 0;
      if (in) {
        s0 = a.delim[b];
        len = a.delim[b + 1] - s0;
      }
      emit_points<KIND == 1 ? 0 : 1, FILL>(a, s0, len, lane, qx, qy, qz, r2, cursor, end, head);
    }
  }
  if (lane == 0 && a.count) a.count[qi] = (uint32_t)(FILL ? cursor - a.offsets[qi] : cursor);
}

// Cell::compute_neighbor_cloud: expand_all_in_radius from the query's grid
// cell (celery.rs:1023-1075) + the group filter (interface.rs:358-361).  The
// table covers every key <= the largest radius (or is the full order).
template <bool FILL>
__global__ __launch_bounds__(256) void k_query_cloud(QueryArgs a) {
  __shared__ uint8_t heads[4][64];
  uint8_t* head = heads[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= a.nq) return;
  const int qi = a.order ? (int)a.order[w] : w;
  const GridDev& G = a.g;
  const int cpd = G.cpd;
  const double qx = a.q[3 * (size_t)qi], qy = a.q[3 * (size_t)qi + 1], qz = a.q[3 * (size_t)qi + 2];
  const double r = a.radius[(size_t)qi * a.rstride];
  const int hx = q_axis_index(qx, G.xmin, G.xmax, G.ix, cpd);
  const int hy = q_axis_index(qy, G.ymin, G.ymax, G.iy, cpd);
  const int hz = q_axis_index(qz, G.zmin, G.zmax, G.iz, cpd);
  uint64_t cursor = FILL ? a.offsets[qi] : 0;
  const uint64_t end = FILL ? a.offsets[qi + 1] : 0;
  for (int base = 0; base < a.tlen; base += 64) {
    const int t = base + lane;
    const bool have = t < a.tlen;
    const double key = have ? a.key[t] : 0.0;
    const bool beyond = have && key > r;  // celery.rs:1036: the loop ends here
    const bool take = have && !beyond;    // keys ascend: the taken lanes are a prefix
    uint32_t s0 = 0, len = 0;
    if (take) {
      int ox, oy, oz;
      unpack_off(a.off[t], ox, oy, oz);
      const int bx = hx + ox, by = hy + oy, bz = hz + oz;
      if (bx >= 0 && bx < cpd && by >= 0 && by < cpd && bz >= 0 && bz < cpd) {
        const uint32_t b = ((uint32_t)bx * (uint32_t)cpd + (uint32_t)by) * (uint32_t)cpd + (uint32_t)bz;
        s0 = a.delim[b];
        len = a.delim[b + 1] - s0;
      }
    }
    if (a.groups)
      emit_points<2, FILL>(a, s0, len, lane, qx, qy, qz, 0.0, cursor, end, head);
    else
      emit_points<0, FILL>(a, s0, len, lane, qx, qy, qz, 0.0, cursor, end, head);
    if (__ballot(beyond)) break;
  }
  if (lane == 0 && a.count) a.count[qi] = (uint32_t)(FILL ? cursor - a.offsets[qi] : cursor);
}

template <bool FILL>
hipError_t launch_query_pass(const QueryArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)((a.nq + 3) / 4)), block(256);
  switch (a.kind) {
    case 0:
      hipLaunchKernelGGL((k_query_box<0, FILL>), grid, block, 0, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((k_query_box<1, FILL>), grid, block, 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((k_query_box<2, FILL>), grid, block, 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((k_query_cloud<FILL>), grid, block, 0, s, a);
      break;
  }
  return hipGetLastError();
}

}  // namespace

// Ids of query q from src[from[q] ..] to dst[to[q] .. to[q + 1]), one
// wavefront per query.
__global__ __launch_bounds__(256) void k_query_compact(int nq, const uint64_t* __restrict__ from,
                                                       const uint64_t* __restrict__ to,
                                                       const uint32_t* __restrict__ src,
                                                       uint32_t* __restrict__ dst) {
  const int lane = threadIdx.x & 63;
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= nq) return;
  const uint64_t f = from[q], t = to[q], m = to[q + 1] - t;
  for (uint64_t k = lane; k < m; k += 64) dst[t + k] = src[f + k];
}

hipError_t launch_query_count(const QueryArgs& a, hipStream_t s) { return launch_query_pass<false>(a, s); }
hipError_t launch_query_fill(const QueryArgs& a, hipStream_t s) { return launch_query_pass<true>(a, s); }
hipError_t launch_query_compact(int nq, const uint64_t* from, const uint64_t* to, const uint32_t* src, uint32_t* dst,
                                hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_query_compact, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, nq, from, to, src, dst);
  return hipGetLastError();
}

}  // namespace tess
