// query.hip -- radius queries over a diagram's grid (celery.rs:700-855,
// :1023-1075; interface.rs:348-365).
//
// One wavefront per query, two passes over the same walk: the first counts
// each query's results, a u64 scan turns the counts into CSR offsets, the
// second writes the ids in the reference's order.  Box queries walk the index
// box of the query in (i, j, k) loop order, 64 grid cells per step; the
// neighbour cloud walks the shell table in canonical order, 64 entries per
// step, until the first key above the radius.  Each lane owns one grid cell
// of the step and lists that cell's points in sorted order, so a wave-wide
// exclusive scan of the per-lane counts gives every lane its output position.
// The walk is bound by the reads of delimiters, sorted indices and (for the
// real radius) gathered coordinates: integer and byte work on HBM.
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

// Exclusive prefix of v over the wavefront; *total = the sum.
__device__ __forceinline__ uint32_t wave_excl(uint32_t v, int lane, uint32_t* total) {
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  *total = __shfl(incl, 63, 64);
  return incl - v;
}

// Points of grid cell b that the query keeps: all (cell radius, cloud
// without group), the ones within the radius (real radius), or the ones of
// the target group (cloud).  Written from ids[pos] on when ids is set.
template <int KIND>
__device__ __forceinline__ uint32_t cell_points(const QueryArgs& a, uint32_t b, double qx, double qy, double qz,
                                                double r2, uint32_t* ids, uint64_t pos, uint64_t end) {
  const uint32_t s0 = a.delim[b], s1 = a.delim[b + 1];
  if (KIND == 1 || (KIND == 3 && !a.groups)) {
    if (ids)
      for (uint32_t s = s0; s < s1; ++s)
        if (pos + (s - s0) < end) ids[pos + (s - s0)] = a.sorted[s];
    return s1 - s0;
  }
  uint32_t c = 0;
  for (uint32_t s = s0; s < s1; ++s) {
    bool keep;
    if (KIND == 2) {
      const double dx = qx - a.sx[s], dy = qy - a.sy[s], dz = qz - a.sz[s];  // celery.rs:700-702
      keep = (dx * dx + dy * dy) + dz * dz <= r2;
    } else {
      keep = a.groups[a.sorted[s]] == a.target_group;  // interface.rs:358-361
    }
    if (keep) {
      if (ids && pos + c < end) ids[pos + c] = a.sorted[s];
      ++c;
    }
  }
  return c;
}

// find_cells_in_radius (KIND 0) and the two neighbour searches built on it
// (KIND 1 cell radius, KIND 2 real radius): celery.rs:753-855.
template <int KIND, bool FILL>
__global__ __launch_bounds__(256) void k_query_box(QueryArgs a) {
  const int lane = threadIdx.x & 63;
  const int qi = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (qi >= a.nq) return;
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
  const uint64_t total = ni > 0 ? (uint64_t)ni * njk : 0;
  uint64_t cursor = FILL ? a.offsets[qi] : 0;
  const uint64_t end = FILL ? a.offsets[qi + 1] : 0;
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
    uint32_t c = 0;
    if (in) c = KIND == 0 ? 1u : cell_points<KIND>(a, b, qx, qy, qz, r2, nullptr, 0, 0);
    uint32_t tot;
    const uint32_t ex = wave_excl(c, lane, &tot);
    if (FILL && c) {
      const uint64_t pos = cursor + ex;
      if (KIND == 0) {
        if (pos < end) a.ids[pos] = b;
      } else {
        cell_points<KIND>(a, b, qx, qy, qz, r2, a.ids, pos, end);
      }
    }
    cursor += tot;
  }
  if (!FILL && lane == 0) a.count[qi] = (uint32_t)cursor;
}

// Cell::compute_neighbor_cloud: expand_all_in_radius from the query's grid
// cell (celery.rs:1023-1075) + the group filter (interface.rs:358-361).  The
// table covers every key <= the largest radius (or is the full order).
template <bool FILL>
__global__ __launch_bounds__(256) void k_query_cloud(QueryArgs a) {
  const int lane = threadIdx.x & 63;
  const int qi = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (qi >= a.nq) return;
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
    uint32_t c = 0, b = 0;
    if (take) {
      int ox, oy, oz;
      unpack_off(a.off[t], ox, oy, oz);
      const int bx = hx + ox, by = hy + oy, bz = hz + oz;
      if (bx >= 0 && bx < cpd && by >= 0 && by < cpd && bz >= 0 && bz < cpd) {
        b = ((uint32_t)bx * (uint32_t)cpd + (uint32_t)by) * (uint32_t)cpd + (uint32_t)bz;
        c = cell_points<3>(a, b, qx, qy, qz, 0.0, nullptr, 0, 0);
      }
    }
    uint32_t tot;
    const uint32_t ex = wave_excl(c, lane, &tot);
    if (FILL && c) cell_points<3>(a, b, qx, qy, qz, 0.0, a.ids, cursor + ex, end);
    cursor += tot;
    if (__ballot(beyond)) break;
  }
  if (!FILL && lane == 0) a.count[qi] = (uint32_t)cursor;
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

hipError_t launch_query_count(const QueryArgs& a, hipStream_t s) { return launch_query_pass<false>(a, s); }
hipError_t launch_query_fill(const QueryArgs& a, hipStream_t s) { return launch_query_pass<true>(a, s); }

}  // namespace tess
