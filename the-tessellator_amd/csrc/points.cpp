// points.cpp -- cells of sites that are not points of the diagram
// (Diagram::get_cell_at_particle, interface.rs:211-232, with
// Cell::compute_voronoi_cell, :257-313), many at once on the device.
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>

#include "launch.h"
#include "runtime.h"

// The results; the call's temporaries belong to the diagram (stmp).
struct tess_site_cells {
  tess_ctx* ctx = nullptr;
  size_t m = 0;
  uint64_t total = 0;
  uint32_t reruns = 0;
  double ms[2] = {0, 0};  // cell passes, face offsets + compaction
  bool have_stats = false;
  DevBuf<double> volume, area;
  DevBuf<uint32_t> fcount, status;
  DevBuf<int32_t> nbr;  // CSR
  DevBuf<uint64_t> foff;
  DevBuf<tess_cell_stats> stats;
};

namespace {

// Site sets from this size on are taken in grid order.
constexpr size_t kSiteOrderMin = 4096;

struct Events {
  hipEvent_t e[3] = {nullptr, nullptr, nullptr};
  Events() {
    for (auto& x : e) (void)hipEventCreate(&x);
  }
  ~Events() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

}  // namespace

extern "C" {

tess_status tess_diagram_cells_at_particles(tess_diagram* d, const double* xyz, size_t m, double search_radius,
                                            int64_t target_group, tess_site_cells** out) {
  if (!d || !out) return TESS_E_INVALID;
  *out = nullptr;
  tess_ctx* c = d->ctx;
  if (m == 0 || !xyz) return c->fail(TESS_E_INVALID, "no sites");
  if (m > (size_t)INT32_MAX / kFaceStage) return c->fail(TESS_E_INVALID, "too many sites");
  if (target_group >= 0 && !d->have_groups) return c->fail(TESS_E_INVALID, "diagram has no groups");
  // the container polyhedron is translated to the site (interface.rs:265-266):
  // only sites inside the container give a cell
  for (size_t i = 0; i < m; ++i)
    for (int k = 0; k < 3; ++k) {
      const double v = xyz[3 * i + k];
      if (!(v >= d->lo[k] && v <= d->hi[k]))
        return c->fail(TESS_E_INVALID, "site " + std::to_string(i) + " is outside the container box");
    }
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  tess_status st;
  if (!d->binned && (st = bin(d)) != TESS_OK) return st;
  const GridDev& G = d->hgrid;
  Table* t = c->table(d->cpd, G.sx, G.sy, G.sz, kTableL, &st);
  if (!t) return st;
  std::unique_ptr<tess_site_cells> p(new (std::nothrow) tess_site_cells);
  if (!p) return c->fail(TESS_E_NOMEM, "host allocation failed");
  p->ctx = c;
  p->m = m;
  p->have_stats = d->opts.collect_stats != 0;
  auto& T = d->stmp;  // serialised by the context lock
  if (p->volume.alloc(m) != hipSuccess || p->fcount.alloc(m) != hipSuccess || p->status.alloc(m) != hipSuccess ||
      p->foff.alloc(m + 1) != hipSuccess || (p->have_stats && p->stats.alloc(m) != hipSuccess) ||
      T.sites.alloc(3 * m) != hipSuccess || T.snbr.alloc(arena_capacity(m)) != hipSuccess ||
      T.sarea.alloc(arena_capacity(m)) != hipSuccess || T.fsums.alloc((m + 1023) / 1024 + 1) != hipSuccess ||
      T.summary.alloc(1) != hipSuccess || T.redo.alloc(m + 1) != hipSuccess || T.tredo.alloc(m + 1) != hipSuccess ||
      T.spill.alloc(m + 1) != hipSuccess || T.fpos.alloc(m) != hipSuccess || T.arena.alloc(kArenaWords) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  Events ev;
  if ((st = c->hip(hipMemcpyAsync(T.sites.p, xyz, 24 * m, hipMemcpyHostToDevice, s), "upload")) != TESS_OK ||
      (st = c->hip(hipMemsetAsync(T.summary.p, 0, sizeof(Summary), s), "memset")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(T.redo.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(T.tredo.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(T.spill.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32((uint32_t*)T.arena.p, 2 * kArenaWords, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(hipEventRecord(ev.e[0], s), "event")) != TESS_OK)
    return st;
  CellArgs a = base_args(d, t);
  a.ext = T.sites.p;
  a.work = nullptr;
  a.nwork = (int)m;
  a.xcd_chunk = 0;
  // larger sets: first pass in grid order of the sites, each XCD taking runs
  // of them, so that waves running together read the same grid cells (the
  // outputs stay per site); small sets are not worth the sort's launches
  DevBuf<uint32_t>&keys = T.keys, &keys_out = T.keys_out, &iota = T.iota, &order = T.order;
  DevBuf<uint8_t>& tmp = T.sort_tmp;
  if (m >= kSiteOrderMin) {
    size_t tmp_bytes = 0;
    if ((st = c->hip(launch_site_order(nullptr, (int)m, a.grid, nullptr, nullptr, nullptr, nullptr, nullptr,
                                       &tmp_bytes, s),
                     "sort")) != TESS_OK)
      return st;
    if (keys.alloc(m) != hipSuccess || keys_out.alloc(m) != hipSuccess || iota.alloc(m) != hipSuccess ||
        order.alloc(m) != hipSuccess || tmp.alloc(tmp_bytes) != hipSuccess)
      return c->fail(TESS_E_NOMEM, "device allocation failed");
    if ((st = c->hip(launch_site_order(T.sites.p, (int)m, a.grid, keys.p, keys_out.p, iota.p, order.p, tmp.p,
                                       &tmp_bytes, s),
                     "sort")) != TESS_OK)
      return st;
    a.work = order.p;
    a.xcd_chunk = plain_xcd_chunk(d);
  }
  a.target_group = target_group;
  a.radius_limit = search_radius;
  a.volume = p->volume.p;
  a.fcount = p->fcount.p;
  a.snbr = T.snbr.p;
  a.sarea = T.sarea.p;
  a.fpos = T.fpos.p;
  a.arena_next = T.arena.p;
  a.arena_cap = arena_capacity(m);
  a.status = p->status.p;
  a.stats = p->have_stats ? p->stats.p : nullptr;
  a.summary = T.summary.p;
  const int vpl = pick_vpl(d);
  // capacity overflows below 256 vertices: the rerun list; at 256: the
  // spill kernel's; shell-table exhaustions: their own list
  a.redo = (vpl < 4 ? T.redo.p : T.spill.p) + 1;
  a.nredo = vpl < 4 ? T.redo.p : T.spill.p;
  a.tredo = T.tredo.p + 1;
  a.ntredo = T.tredo.p;
  a.defer_table = 1;
  const CellArgs a0 = a;
  if ((st = c->hip(launch_cells(vpl, a, s), "cell kernel")) != TESS_OK) return st;
  uint32_t nredo = 0;
  if ((st = c->hip(hipMemcpyAsync(&nredo, T.redo.p, 4, hipMemcpyDeviceToHost, s), "redo")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "cell kernel")) != TESS_OK)
    return st;
  if (nredo && vpl < 4) {  // sites that outgrew the first pass: the largest variant
    if (T.redo_list.alloc(nredo) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
    a.work = T.redo_list.p;
    a.nwork = (int)nredo;
    a.xcd_chunk = 0;
    a.redo = T.spill.p + 1;
    a.nredo = T.spill.p;
    if ((st = c->hip(hipMemcpyAsync(T.redo_list.p, T.redo.p + 1, 4ull * nredo, hipMemcpyDeviceToDevice, s),
                     "redo")) != TESS_OK ||
        (st = c->hip(launch_cells(4, a, s), "cell kernel")) != TESS_OK)
      return st;
  }
  // sites past the shell table or capacity 256
  std::vector<uint32_t> tab, spl;
  if ((st = read_list(c, T.tredo.p, tab)) != TESS_OK || (st = read_list(c, T.spill.p, spl)) != TESS_OK) return st;
  FinishCounts fin;
  if ((!tab.empty() || !spl.empty()) &&
      (st = finish_cells(d, a0, std::move(tab), std::move(spl), &fin)) != TESS_OK)
    return st;
  if (nredo && vpl < 4 &&
      (st = c->hip(launch_mark_status(T.redo_list.p, (int)nredo, nullptr, p->status.p, a.stats, TESS_CELL_RERUN, s),
                   "mark")) != TESS_OK)
    return st;
  if ((st = c->hip(hipEventRecord(ev.e[1], s), "event")) != TESS_OK ||
      (st = c->hip(launch_face_offsets(p->fcount.p, (int)m, p->foff.p, T.fsums.p, s), "face offsets")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&p->total, p->foff.p + m, 8, hipMemcpyDeviceToHost, s), "total")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "face offsets")) != TESS_OK)
    return st;
  if (p->nbr.alloc(p->total) != hipSuccess || p->area.alloc(p->total) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  Summary sm2;
  float ms[2] = {0, 0};
  if ((st = c->hip(launch_compact(p->fcount.p, p->foff.p, (int)m, T.snbr.p, T.sarea.p, p->nbr.p, p->area.p, s,
                                  T.fpos.p),
                   "compact")) != TESS_OK ||
      (st = c->hip(hipEventRecord(ev.e[2], s), "event")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&sm2, T.summary.p, sizeof(Summary), hipMemcpyDeviceToHost, s), "summary")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "compact")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[0], ev.e[0], ev.e[1]), "event")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[1], ev.e[1], ev.e[2]), "event")) != TESS_OK)
    return st;
  p->ms[0] = ms[0];
  p->ms[1] = ms[1];
  p->reruns = vpl < 4 ? nredo : 0;
  if (fin.failed) return c->fail(TESS_E_CAPACITY, std::to_string(fin.failed) + " cells exceeded the spill workspace limit");
  if (sm2.degenerate)
    return c->fail(TESS_E_DEGENERATE, std::to_string(sm2.degenerate) + " cells failed topology checks");
  *out = p.release();
  return TESS_OK;
}

size_t tess_site_cells_count(const tess_site_cells* p) { return p ? p->m : 0; }

tess_status tess_site_cells_total_faces(const tess_site_cells* p, uint64_t* out) {
  if (!p || !out) return TESS_E_INVALID;
  *out = p->total;
  return TESS_OK;
}

tess_status tess_site_cells_copy(const tess_site_cells* p, double* volume, uint32_t* face_count,
                                 uint64_t* face_offset, int32_t* neighbor, double* face_area) {
  if (!p) return TESS_E_INVALID;
  (void)hipSetDevice(p->ctx->device);
  hipError_t e = hipSuccess;
  tess_ctx* c = p->ctx;
  if (volume) e = download(c, volume, p->volume.p, 8 * p->m);
  if (face_count && e == hipSuccess) e = download(c, face_count, p->fcount.p, 4 * p->m);
  if (face_offset && e == hipSuccess) e = download(c, face_offset, p->foff.p, 8 * (p->m + 1));
  if (neighbor && p->total && e == hipSuccess) e = download(c, neighbor, p->nbr.p, 4 * p->total);
  if (face_area && p->total && e == hipSuccess) e = download(c, face_area, p->area.p, 8 * p->total);
  return p->ctx->hip(e, "download");
}

tess_status tess_site_cells_copy_stats(const tess_site_cells* p, tess_cell_stats* stats) {
  if (!p || !stats) return TESS_E_INVALID;
  if (!p->have_stats) return p->ctx->fail(TESS_E_STATE, "the diagram was created without collect_stats");
  (void)hipSetDevice(p->ctx->device);
  return p->ctx->hip(hipMemcpy(stats, p->stats.p, sizeof(tess_cell_stats) * p->m, hipMemcpyDeviceToHost),
                     "download");
}

tess_status tess_site_cells_timings(const tess_site_cells* p, double ms[2], uint32_t* capacity_reruns) {
  if (!p) return TESS_E_INVALID;
  if (ms) {
    ms[0] = p->ms[0];
    ms[1] = p->ms[1];
  }
  if (capacity_reruns) *capacity_reruns = p->reruns;
  return TESS_OK;
}

void tess_site_cells_destroy(tess_site_cells* p) {
  if (!p) return;
  (void)hipSetDevice(p->ctx->device);
  delete p;
}

}  // extern "C"
