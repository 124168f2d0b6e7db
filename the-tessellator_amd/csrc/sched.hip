// sched.hip -- work lists of the learned schedule.
//
// Every grid position gets a key: its slot -- a vertex capacity (from the
// predicted candidate count and two thresholds), or the long-search slot
// when the predicted final radius reaches past the default shell table --
// then a cost bucket in descending predicted cost.  A stable counting sort by key (tiles of 256
// positions, per-wavefront rank by ballot matching) turns the keys into one
// list per slot, heaviest cells first and grid order inside a bucket, so the
// dispatcher starts the long cells early and neighbouring cells still run
// together.  The lists only order work: every cell is computed exactly once.
#include "launch.h"
#include "tess_internal.h"

namespace tess {

namespace {

constexpr int kTile = 256;

// Order-preserving u32 image of a float (larger float, larger integer).
__device__ __forceinline__ uint32_t ordered(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ uint32_t sched_key(const float* y, int stride, float scale, float shift, float t1, float t2,
                                              const SchedRadius& r, uint32_t o, float* vr) {
  const float v = y[(size_t)o * stride] * scale + shift;  // o: point of the cell
  int slot = v > t2 ? 2 : (v > t1 ? 1 : 0);
  *vr = -INFINITY;
  if (r.y) {  // predicted log(R / s_min) past the threshold: the long-search slot
    const float w = r.y[(size_t)o * stride] * r.scale + r.shift;
    if (w > r.threshold) {
      slot = 3;
      *vr = w;
    }
  }
  // 4 buckets per unit of log(1 + C), heaviest first; NaN lands in bucket 0
  float u = (16.0f - v) * 4.0f;
  u = fminf(fmaxf(u, 0.0f), (float)(kSchedBuckets - 1));
  return (uint32_t)(slot * kSchedBuckets + (int)u);
}

// Lanes of this wavefront whose key equals the lane's own (keys < 256).
__device__ __forceinline__ uint64_t match_key(uint32_t key, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint64_t bb = __ballot(valid && ((key >> b) & 1u));
    m &= ((key >> b) & 1u) ? bb : ~bb;
  }
  return m;
}

__global__ __launch_bounds__(kTile) void k_sched_hist(const float* __restrict__ y, int stride, float scale,
                                                      float shift, float t1, float t2, SchedRadius r,
                                                      const uint32_t* __restrict__ sorted, int n,
                                                      uint8_t* __restrict__ keys, uint32_t* __restrict__ hist,
                                                      int ntiles, uint32_t* __restrict__ rmax) {
  __shared__ uint32_t cnt[kTile / 64][kSchedKeys];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < (kTile / 64) * kSchedKeys; i += kTile) (&cnt[0][0])[i] = 0;
  __syncthreads();
  const int k = blockIdx.x * kTile + t;
  const bool valid = k < n;
  uint32_t key = 0;
  float vr = -INFINITY;
  if (valid) {
    key = sched_key(y, stride, scale, shift, t1, t2, r, sorted ? sorted[k] : (uint32_t)k, &vr);
    keys[k] = (uint8_t)key;
  }
  // the largest predicted radius of the long-search slot
  float wmax = vr;
  for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
  if (lane == 0 && wmax > -INFINITY) atomicMax(rmax, ordered(wmax));
  const uint64_t m = match_key(key, valid);
  if (valid && lane == __ffsll((unsigned long long)m) - 1) cnt[w][key] = (uint32_t)__popcll(m);
  __syncthreads();
  for (int q = t; q < kSchedKeys; q += kTile) {
    uint32_t s = 0;
#pragma unroll
    for (int ww = 0; ww < kTile / 64; ++ww) s += cnt[ww][q];
    hist[(size_t)q * ntiles + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kTile) void k_sched_scatter(const uint8_t* __restrict__ keys, int n,
                                                         const uint32_t* __restrict__ start, int ntiles,
                                                         uint32_t* __restrict__ list) {
  __shared__ uint32_t cnt[kTile / 64][kSchedKeys];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < (kTile / 64) * kSchedKeys; i += kTile) (&cnt[0][0])[i] = 0;
  __syncthreads();
  const int k = blockIdx.x * kTile + t;
  const bool valid = k < n;
  const uint32_t key = valid ? keys[k] : 0u;
  const uint64_t m = match_key(key, valid);
  if (valid && lane == __ffsll((unsigned long long)m) - 1) cnt[w][key] = (uint32_t)__popcll(m);
  __syncthreads();
  if (!valid) return;
  uint32_t pos = start[(size_t)key * ntiles + blockIdx.x];
  for (int ww = 0; ww < w; ++ww) pos += cnt[ww][key];
  pos += (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  list[pos] = (uint32_t)k;
}

// bounds[s] = first list position of slot s; bounds[kSchedSlots] = n.
__global__ void k_sched_bounds(const uint32_t* __restrict__ start, int ntiles, int n, uint32_t* __restrict__ bounds) {
  const int s = threadIdx.x;
  if (s < kSchedSlots) bounds[s] = start[(size_t)(s * kSchedBuckets) * ntiles];
  if (s == kSchedSlots) bounds[kSchedSlots] = (uint32_t)n;
}

__global__ void k_mark_status(const uint32_t* __restrict__ list, int m, const uint32_t* __restrict__ sorted,
                              uint32_t* __restrict__ status, tess_cell_stats* __restrict__ stats, uint32_t flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t o = sorted ? sorted[list[i]] : list[i];
  status[o] |= flag;
  if (stats) stats[o].status |= flag;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t m, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) p[i] = v;
}

// items[g] = (set << 32) | (g - base[set]) for g in [0, total); base has
// nsets + 1 entries (base[nsets] = total).
__global__ void k_batch_items(const uint64_t* __restrict__ base, int nsets, uint64_t total,
                              uint64_t* __restrict__ items) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = nsets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  items[g] = ((uint64_t)lo << 32) | (g - base[lo]);
}

// cost[g] = y[(base[set] + sorted[base[set] + pos]) * stride] * scale + shift
// for item g = (set, pos); sorted holds set-local indices.
__global__ void k_batch_costs(const uint64_t* __restrict__ items, uint64_t total, const uint64_t* __restrict__ base,
                              const uint32_t* __restrict__ sorted, const float* __restrict__ y, int stride,
                              float scale, float shift, float* __restrict__ cost) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const uint64_t it = items[g];
  const uint64_t b = base[it >> 32];
  const uint64_t o = b + sorted[b + (uint32_t)it];
  cost[g] = y[o * stride] * scale + shift;
}

// bad[0] += 1 for every set whose delimiters do not end at its size.
__global__ void k_batch_check(const uint32_t* __restrict__ delim, const uint64_t* __restrict__ dend,
                              const uint64_t* __restrict__ base, int nsets, uint32_t* __restrict__ bad) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsets) return;
  if ((uint64_t)delim[dend[s]] != base[s + 1] - base[s]) atomicAdd(bad, 1u);
}

inline int blocks(long long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

hipError_t launch_batch_items(const uint64_t* base, int nsets, uint64_t total, uint64_t* items, hipStream_t s) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_items, dim3(blocks((long long)total, 256)), dim3(256), 0, s, base, nsets, total, items);
  return hipGetLastError();
}

hipError_t launch_batch_costs(const uint64_t* items, uint64_t total, const uint64_t* base, const uint32_t* sorted,
                              const float* y, int stride, float scale, float shift, float* cost, hipStream_t s) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_costs, dim3(blocks((long long)total, 256)), dim3(256), 0, s, items, total, base, sorted,
                     y, stride, scale, shift, cost);
  return hipGetLastError();
}

hipError_t launch_batch_check(const uint32_t* delim, const uint64_t* dend, const uint64_t* base, int nsets,
                              uint32_t* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_check, dim3(blocks(nsets, 256)), dim3(256), 0, s, delim, dend, base, nsets, bad);
  return hipGetLastError();
}

hipError_t launch_schedule(const float* y, int stride, float scale, float shift, float t1, float t2,
                           const uint32_t* sorted, int n, uint8_t* key, uint32_t* hist, uint32_t* sums,
                           uint32_t* list, uint32_t* bounds, hipStream_t s, const SchedRadius* radius) {
  if (n <= 0) return hipSuccess;
  const int ntiles = blocks(n, kTile);
  const int m = kSchedKeys * ntiles;
  SchedRadius r;
  if (radius) r = *radius;
  hipError_t e0 = launch_fill_u32(bounds + kSchedSlots + 1, 1, 0u, s);
  if (e0 != hipSuccess) return e0;
  hipLaunchKernelGGL(k_sched_hist, dim3(ntiles), dim3(kTile), 0, s, y, stride, scale, shift, t1, t2, r, sorted, n, key,
                     hist, ntiles, bounds + kSchedSlots + 1);
  // exclusive scan in place: hist becomes the start of each (key, tile)
  hipError_t e = launch_scan_u32(hist, m, hist, sums, sums + blocks(m, 1024), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sched_scatter, dim3(ntiles), dim3(kTile), 0, s, key, n, hist, ntiles, list);
  hipLaunchKernelGGL(k_sched_bounds, dim3(1), dim3(64), 0, s, hist, ntiles, n, bounds);
  return hipGetLastError();
}

hipError_t launch_mark_status(const uint32_t* list, int m, const uint32_t* sorted, uint32_t* status,
                              tess_cell_stats* stats, uint32_t flag, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mark_status, dim3(blocks(m, 256)), dim3(256), 0, s, list, m, sorted, status, stats, flag);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, size_t m, uint32_t v, hipStream_t s) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks((long long)m, 256)), dim3(256), 0, s, p, m, v);
  return hipGetLastError();
}

}  // namespace tess
