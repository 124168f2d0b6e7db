// launch.h -- host-side launchers of the device passes (one per kernel file).
#pragma once
#include "tess_internal.h"

namespace tess {

// bin.hip
hipError_t launch_soa(const double* xyz, int n, double* x, double* y, double* z, hipStream_t s);
hipError_t launch_bounds(const double* x, const double* y, const double* z, int n, int cpd,
                         double* part, int nparts, GridDev* g, hipStream_t s);
// counting sort: bins/cnt/delim/cursor/sorted/scratch sized by the caller
struct SortBufs {
  uint32_t* bins;     // [n]
  uint32_t* cnt;      // [nbins]
  uint32_t* delim;    // [nbins + 1]
  uint32_t* cursor;   // [nbins] (grid-order binning only)
  uint32_t* sorted;   // [n]
  uint32_t* scratch;  // [n]: slots in the cell ranges, then the in-cell sort's copy
  uint32_t* sums;     // [ceil(nbins / 1024) + 1]
  uint32_t* big;      // [nbins]
  uint32_t* misc;     // [4]: total, nbig
};
// key (nullable): a cell's points ascend by (key[i], i) instead of i
hipError_t launch_counting_sort(const double* x, const double* y, const double* z, int n,
                                const GridDev* g, int nbins, const SortBufs& b, hipStream_t s,
                                const uint32_t* key = nullptr);
// The counting sort of many point sets at once (a batch): set after set in
// the SoA arrays, their grid cells numbered apart (set k's cells from
// cbase[k]), their delimiters and sorted indices set-local, as
// launch_bounds + launch_counting_sort + launch_gather would leave them set
// by set (input order inside a cell).  pbase, cbase, bbase: nsets + 1
// entries (first point / cell / delimiter of each set, then the totals).
struct SegBinning {
  int nsets = 0;
  uint32_t n = 0, nb = 0;  // points, grid cells over all sets
  const uint64_t* pbase = nullptr;
  const uint32_t* cbase = nullptr;
  const uint32_t* bbase = nullptr;
  const int32_t* cpd = nullptr;  // cells per dimension of each set
  const uint32_t* pblk = nullptr;  // [nsets + 1] first bounds block of each set
  uint32_t nparts = 0;             // bounds blocks over all sets
  double* part = nullptr;          // [nparts][6] their partial boxes
  const double *x = nullptr, *y = nullptr, *z = nullptr;
  GridDev* grid = nullptr;  // [nsets]
  uint32_t* bins = nullptr;     // [n] set-local cell of each point
  uint32_t* rank = nullptr;     // [n] slot in its cell, then the large cells' scratch
  uint32_t* cnt = nullptr;      // [nb]
  uint32_t* gscan = nullptr;    // [nb] exclusive scan of cnt over the batch
  uint32_t* sums = nullptr;     // [ceil(nb / 1024) + 1]
  uint32_t* delim = nullptr;    // [nb + nsets]
  uint32_t* sorted = nullptr;   // [n]
  uint32_t* big = nullptr;      // [nb]
  uint32_t* misc = nullptr;     // [2]: scan total, large cells
  double *sx = nullptr, *sy = nullptr, *sz = nullptr;  // [n] positions in grid order
};
hipError_t launch_seg_binning(const SegBinning& a, hipStream_t s);
// The same binning without the in-cell ordering: b.sorted lists the n
// positions grid cell by grid cell, in no set order inside a cell (only
// bins, cnt, delim, cursor, sorted, sums and misc are used)
hipError_t launch_grid_order(const double* x, const double* y, const double* z, int n, const GridDev* g,
                             int nbins, const SortBufs& b, hipStream_t s);
// exclusive scan of m u32 (sums: ceil(m / 1024) + 1 entries); *total = sum
hipError_t launch_scan_u32(const uint32_t* in, int m, uint32_t* out, uint32_t* sums, uint32_t* total,
                           hipStream_t s);
hipError_t launch_gather(const uint32_t* sorted, int n, const double* x, const double* y,
                         const double* z, double* sx, double* sy, double* sz, hipStream_t s);
hipError_t launch_face_offsets(const uint32_t* fcount, int n, uint64_t* foff, uint64_t* sums,
                               hipStream_t s);
// cell o's faces are read at snbr[fpos[o]] (fpos null: at o * kFaceStage)
hipError_t launch_compact(const uint32_t* fcount, const uint64_t* foff, int n, const int32_t* snbr,
                          const double* sarea, int32_t* nbr, double* area, hipStream_t s,
                          const uint64_t* fpos = nullptr);
hipError_t launch_invert(const uint32_t* sorted, int n, uint32_t* inv, hipStream_t s);

// cell.hip: vpl = vertex slots per lane (1, 2 or 4)
hipError_t launch_cells(int vpl, const CellArgs& a, hipStream_t s);
// the same cells staging their geometry (CellArgs::geo_*), with the cell
// outputs too when with_cells, vpl 1 or 4; staging capacity per work item:
// 64 vpl vertices, kCellHalfEdges(vpl) ring entries, kFaceStage faces
hipError_t launch_cells_geometry(int vpl, const CellArgs& a, bool with_cells, hipStream_t s);
constexpr int kCellHalfEdges(int vpl) { return vpl == 1 ? 256 : 768; }
// geometry.hip: staged geometry of nslots work items -> the CSR
hipError_t launch_geo_compact(const uint32_t* slot_cell, int nslots, int vcap, int hcap, const double* sv,
                              const uint16_t* sr, const uint16_t* sro, const uint32_t* count, int n,
                              const uint64_t* voff, const uint64_t* rbase, const uint64_t* foff, double* out_v,
                              uint16_t* out_ring, uint64_t* out_roff, hipStream_t s);
// many sets: one wavefront per item of a list of host-known length n ...
hipError_t launch_cells_items(int vpl, const BatchArgs& b, uint32_t n, hipStream_t s);
// ... or persistent over a device-counted list at capacity 128 (vpl 2) or
// 256 (vpl 4) (grid = resident blocks: cells_queue_blocks_per_cu(vpl) x CUs)
hipError_t launch_cells_queue(int vpl, const BatchArgs& b, int grid, hipStream_t s);
int cells_queue_blocks_per_cu(int vpl);
// cells too large for the LDS kernels, one block per workspace (p.wsd ...
// hold `blocks` workspaces), blocks taking items in turn
hipError_t launch_cells_spill(const SpillArgs& p, int blocks, hipStream_t s);
hipError_t launch_diag_arith(const double* a, const double* b, int n, double* sq, double* rc,
                             double* dv, hipStream_t s);

// sched.hip: work lists for the learned schedule
constexpr int kSchedBuckets = 64;                      // cost buckets per slot
constexpr int kSchedSlots = 4;                         // capacity 64, 128, 256; long search
constexpr int kSchedKeys = kSchedSlots * kSchedBuckets;  // slot-major keys (fit a byte)
static_assert(kSchedKeys <= 256, "schedule keys are bytes");
// The long-search slot: a cell goes there when y[o * stride] * scale + shift
// (its predicted log(R / s_min)) exceeds threshold (y null: no such slot).
struct SchedRadius {
  const float* y = nullptr;
  float scale = 1.f, shift = 0.f, threshold = 0.f;
};
// Key of every grid position from a per-point cost v = y[o * stride] * scale
// + shift (log(1 + candidates)), o = sorted[k] (k when sorted is null):
// capacity slot by thresholds t1 < t2 (or the long-search slot, radius),
// then descending cost.  Writes the stable key order to list[n], the start
// of each slot (and n) to bounds[0..kSchedSlots] and the largest predicted
// log radius of the long-search slot (order-preserving u32 bits, 0 when the
// slot is empty) to bounds[kSchedSlots + 1].  hist: kSchedKeys * ceil(n /
// 256) entries; sums: ceil(hist / 1024) + 1.
hipError_t launch_schedule(const float* y, int stride, float scale, float shift, float t1, float t2,
                           const uint32_t* sorted, int n, uint8_t* key, uint32_t* hist, uint32_t* sums,
                           uint32_t* list, uint32_t* bounds, hipStream_t s, const SchedRadius* radius = nullptr);
// batches: item list, per-item costs from per-point predictions, binning check
hipError_t launch_batch_items(const uint64_t* base, int nsets, uint64_t total, uint64_t* items, hipStream_t s);
hipError_t launch_batch_costs(const uint64_t* items, uint64_t total, const uint64_t* base, const uint32_t* sorted,
                              const float* y, int stride, float scale, float shift, float* cost, hipStream_t s);
hipError_t launch_batch_check(const uint32_t* delim, const uint64_t* dend, const uint64_t* base, int nsets,
                              uint32_t* bad, hipStream_t s);
// status[o] |= flag (and stats[o].status when stats is not null), o = sorted[list[i]]
// (o = list[i] when sorted is null)
hipError_t launch_mark_status(const uint32_t* list, int m, const uint32_t* sorted, uint32_t* status,
                              tess_cell_stats* stats, uint32_t flag, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t m, uint32_t v, hipStream_t s);

// query.hip: radius queries, one wavefront per query (count, then fill at
// offsets from launch_face_offsets over the counts; a fill with count set
// also writes the ids it kept per query, for fills into upper-bound slots
// that launch_query_compact then packs)
hipError_t launch_query_count(const QueryArgs& a, hipStream_t s);
hipError_t launch_query_fill(const QueryArgs& a, hipStream_t s);
hipError_t launch_query_compact(int nq, const uint64_t* from, const uint64_t* to, const uint32_t* src, uint32_t* dst,
                                hipStream_t s);

// order.hip: Diagram::morton_index of every point (the box: the grid's
// bounds grown to contain the origin) and the internal order it defines
hipError_t launch_morton(const double* x, const double* y, const double* z, int n, const GridDev* g,
                         uint32_t* codes, hipStream_t s);
// original_of_internal = stable sort of 0..n-1 by code; internal_of_original
// its inverse.  tmp: device scratch of *tmp_bytes (query with tmp = null).
hipError_t launch_index_maps(const uint32_t* codes, int n, uint32_t* iota, uint32_t* keys_out,
                             uint32_t* original_of_internal, uint32_t* internal_of_original, void* tmp,
                             size_t* tmp_bytes, hipStream_t s);
// order = the sites (AoS xyz, m) stably sorted by their grid cell (clamped),
// the work order of a pass over outside sites: neighbouring waves then read
// the same grid cells.  tmp as launch_index_maps.
hipError_t launch_site_order(const double* xyz, int m, const GridDev* g, uint32_t* keys, uint32_t* keys_out,
                             uint32_t* iota, uint32_t* order, void* tmp, size_t* tmp_bytes, hipStream_t s);

// mlp.hip: learned search-radius model (bf16 MFMA, fp32 accumulate)
hipError_t launch_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s);
hipError_t launch_features(const uint32_t* bins, const uint32_t* delim, int cpd, int n, uint16_t* feat,
                           hipStream_t s);
hipError_t launch_mlp_forward(const uint16_t* X, int n, const uint16_t* W, const float* bias, float* out,
                              hipStream_t s);
// one gradient slot per 32 samples: grad[(nsamp + 31) / 32][kMlpParams]
hipError_t launch_mlp_train_step(const uint16_t* X, const float* T, const int* idx, int nsamp, float scale,
                                 const uint16_t* W, const float* bias, float* grad, float* loss, hipStream_t s);
hipError_t launch_targets(const tess_cell_stats* st, int n, double smin, float* T, uint16_t* vmax, hipStream_t s);
hipError_t launch_grad_sum(const float* grad, int nslots, float* out, hipStream_t s);
hipError_t launch_standardise(const float* T, size_t n, const float norm[4], float* Tn, hipStream_t s);
hipError_t launch_weights_from_master(const float* master, uint16_t* wbf, float* bias, hipStream_t s);
hipError_t launch_adam(const float* grad, int nslots, float* master, float* m1, float* m2, uint16_t* wbf,
                       float* bias, float lr, float b1c, float b2c, hipStream_t s);

}  // namespace tess
