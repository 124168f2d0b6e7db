// geometry.cpp -- vertices and face rings of every cell of a diagram
// (Cell::compute_vertices / VoronoiFace::compute_vertices,
// interface.rs:367-370, :403-406; polyhedron.rs:885-919) as CSR on the device.
#include <string>

#include "launch.h"
#include "runtime.h"

tess_status geo_staging_alloc(tess_diagram* d, GeoStaging& g, size_t items, int vpl) {
  tess_ctx* c = d->ctx;
  g.vpl = vpl;
  if (g.slot.alloc(items) != hipSuccess || g.v.alloc(3ull * 64 * vpl * items) != hipSuccess ||
      g.ring.alloc((size_t)kCellHalfEdges(vpl) * items) != hipSuccess ||
      g.roff.alloc((size_t)kFaceStage * items) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  return c->hip(launch_fill_u32(g.slot.p, items, 0u, c->stream), "fill");
}

namespace {

struct GeoEvents {
  hipEvent_t e[3] = {nullptr, nullptr, nullptr};
  bool ok = true;
  GeoEvents() {
    for (auto& x : e)
      if (hipEventCreate(&x) != hipSuccess) ok = false;
  }
  ~GeoEvents() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

// Per-cell counts (d->geo_count, staged with the cells) to CSR offsets, then
// the staged vertices and rings of every pass compacted into the CSR arrays.
// ev[1] is recorded here; ev[0], when recorded, opens the staging time.
tess_status finish_geometry(tess_diagram* d, const GeoStage& g, GeoEvents& ev, double stage_ms) {
  const GeoStaging& first = g.first;
  const GeoStaging& rerun = g.rerun;
  const uint32_t nredo = g.nredo;
  tess_ctx* c = d->ctx;
  hipStream_t s = c->stream;
  const size_t n = d->n;
  tess_status st;
  if ((st = c->hip(hipEventRecord(ev.e[1], s), "event")) != TESS_OK ||
      (st = c->hip(launch_face_offsets(d->geo_count.p, (int)n, d->geo_voff.p, d->geo_sums.p, s), "scan")) !=
          TESS_OK ||
      (st = c->hip(launch_face_offsets(d->geo_count.p + n, (int)n, d->geo_rbase.p, d->geo_sums.p, s), "scan")) !=
          TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&d->geo_nv, d->geo_voff.p + n, 8, hipMemcpyDeviceToHost, s), "download")) !=
          TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&d->geo_nr, d->geo_rbase.p + n, 8, hipMemcpyDeviceToHost, s), "download")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "geometry")) != TESS_OK)
    return st;
  const uint64_t F = d->total_faces;
  if (d->geo_v.alloc(3 * d->geo_nv) != hipSuccess || d->geo_roff.alloc(F + 1) != hipSuccess ||
      d->geo_ring.alloc(d->geo_nr) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  auto compact = [&](const GeoStaging& g, size_t items) {
    return c->hip(launch_geo_compact(g.slot.p, (int)items, 64 * g.vpl, kCellHalfEdges(g.vpl), g.v.p, g.ring.p,
                                     g.roff.p, d->geo_count.p, (int)n, d->geo_voff.p, d->geo_rbase.p, d->foff.p,
                                     d->geo_v.p, d->geo_ring.p, d->geo_roff.p, s),
                  "compact");
  };
  float ms[2] = {(float)stage_ms, 0.f};
  for (size_t k = 0; k < g.extra.size(); ++k)
    if ((st = compact(*g.extra[k], g.extra_items[k])) != TESS_OK) return st;
  if ((st = compact(first, n)) != TESS_OK || (nredo && (st = compact(rerun, nredo)) != TESS_OK) ||
      (st = c->hip(hipMemcpyAsync(d->geo_roff.p + F, &d->geo_nr, 8, hipMemcpyHostToDevice, s), "upload")) !=
          TESS_OK ||
      (st = c->hip(hipEventRecord(ev.e[2], s), "event")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "geometry")) != TESS_OK ||
      (stage_ms < 0 && (st = c->hip(hipEventElapsedTime(&ms[0], ev.e[0], ev.e[1]), "event")) != TESS_OK) ||
      (st = c->hip(hipEventElapsedTime(&ms[1], ev.e[1], ev.e[2]), "event")) != TESS_OK)
    return st;
  d->geo_ms[0] = ms[0];
  d->geo_ms[1] = ms[1];
  d->geo_ready = true;
  return TESS_OK;
}

}  // namespace

extern "C" {

tess_status tess_diagram_compute_geometry(tess_diagram* d) {
  if (!d) return TESS_E_INVALID;
  tess_ctx* c = d->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  tess_status st;
  const size_t n = d->n;
  GeoEvents ev;
  if (!ev.ok) return c->fail(TESS_E_DEVICE, "event create failed");
  if (d->geo_count.alloc(2 * n) != hipSuccess || d->geo_voff.alloc(n + 1) != hipSuccess ||
      d->geo_rbase.alloc(n + 1) != hipSuccess || d->geo_sums.alloc((n + 1023) / 1024 + 1) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  GeoStage g;
  g.count = d->geo_count.p;
  if (!d->computed && pick_vpl(d) == 1) {
    // not computed yet: the compute's own cell passes stage the geometry
    // (capacity 64, the overflows again at 256), then only the offsets and
    // the compaction remain
    if ((st = geo_staging_alloc(d, g.first, n, 1)) != TESS_OK ||
        (st = c->hip(launch_fill_u32(d->geo_count.p, 2 * n, 0u, s), "fill")) != TESS_OK)
      return st;
    d->geo_stage = &g;
    st = compute_cells(d, nullptr);
    d->geo_stage = nullptr;
    if (st != TESS_OK) return st;
    return finish_geometry(d, g, ev, d->tim.cells_ms);
  }
  if (!d->computed && (st = compute_cells(d, nullptr)) != TESS_OK) return st;
  if (d->degenerate) return c->fail(TESS_E_DEGENERATE, "the diagram has failed cells");
  const GridDev& G = d->hgrid;
  Table* t = c->table(d->cpd, G.sx, G.sy, G.sz, kTableL, &st);
  if (!t) return st;
  // One pass recomputes every cell like compute does (capacity 64, the
  // overflows again at 256) and stages its geometry per work item; the
  // counts give the CSR offsets and a compaction copies the staged data.
  if ((st = geo_staging_alloc(d, g.first, n, 1)) != TESS_OK) return st;
  CellArgs a = base_args(d, t);
  a.work = nullptr;
  a.nwork = (int)n;
  a.xcd_chunk = plain_xcd_chunk(d);
  a.stats = nullptr;
  a.summary = nullptr;
  a.redo = d->redo.p + 1;
  a.nredo = d->redo.p;
  a.tredo = d->tredo.p + 1;
  a.ntredo = d->tredo.p;
  a.defer_table = 1;
  a.geo_count = d->geo_count.p;
  a.geo_slot = g.first.slot.p;
  a.geo_v = g.first.v.p;
  a.geo_ring = g.first.ring.p;
  a.geo_roff = g.first.roff.p;
  uint32_t nredo = 0;
  if ((st = c->hip(launch_fill_u32(d->geo_count.p, 2 * n, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(d->redo.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(d->tredo.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(d->spill.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(hipEventRecord(ev.e[0], s), "event")) != TESS_OK ||
      (st = c->hip(launch_cells_geometry(1, a, false, s), "geometry")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&nredo, d->redo.p, 4, hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "geometry")) != TESS_OK)
    return st;
  if (nredo) {
    if (d->redo_list.alloc(nredo) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
    if ((st = geo_staging_alloc(d, g.rerun, nredo, 4)) != TESS_OK) return st;
    CellArgs r = a;
    r.work = d->redo_list.p;
    r.nwork = (int)nredo;
    r.xcd_chunk = 0;
    r.redo = d->spill.p + 1;
    r.nredo = d->spill.p;
    r.geo_slot = g.rerun.slot.p;
    r.geo_v = g.rerun.v.p;
    r.geo_ring = g.rerun.ring.p;
    r.geo_roff = g.rerun.roff.p;
    if ((st = c->hip(hipMemcpyAsync(d->redo_list.p, d->redo.p + 1, 4ull * nredo, hipMemcpyDeviceToDevice, s),
                     "redo")) != TESS_OK ||
        (st = c->hip(launch_cells_geometry(4, r, false, s), "geometry")) != TESS_OK)
      return st;
  }
  g.nredo = nredo;
  // cells past the shell table (staged by longer-table passes) or above 256
  // vertices (refused)
  std::vector<uint32_t> tab, spl;
  if ((st = read_list(c, d->tredo.p, tab)) != TESS_OK || (st = read_list(c, d->spill.p, spl)) != TESS_OK) return st;
  if (!tab.empty() || !spl.empty()) {
    g.with_cells = false;
    if ((st = finish_cells(d, a, std::move(tab), std::move(spl), nullptr, &g)) != TESS_OK) return st;
  }
  return finish_geometry(d, g, ev, -1.0);
}

tess_status tess_diagram_geometry_sizes(const tess_diagram* d, uint64_t* nvertices, uint64_t* nring) {
  if (!d) return TESS_E_INVALID;
  if (!d->geo_ready) return d->ctx->fail(TESS_E_STATE, "geometry read before tess_diagram_compute_geometry");
  if (nvertices) *nvertices = d->geo_nv;
  if (nring) *nring = d->geo_nr;
  return TESS_OK;
}

tess_status tess_diagram_geometry_timings(const tess_diagram* d, double ms[2]) {
  if (!d || !ms) return TESS_E_INVALID;
  if (!d->geo_ready) return d->ctx->fail(TESS_E_STATE, "geometry read before tess_diagram_compute_geometry");
  ms[0] = d->geo_ms[0];
  ms[1] = d->geo_ms[1];
  return TESS_OK;
}

tess_status tess_diagram_copy_geometry(const tess_diagram* d, uint64_t* vertex_offset, double* vertices,
                                       uint64_t* ring_offset, uint16_t* ring) {
  if (!d) return TESS_E_INVALID;
  tess_ctx* c = d->ctx;
  if (!d->geo_ready) return c->fail(TESS_E_STATE, "geometry read before tess_diagram_compute_geometry");
  (void)hipSetDevice(c->device);
  hipError_t e = hipSuccess;
  if (vertex_offset) e = download(c, vertex_offset, d->geo_voff.p, 8 * (d->n + 1));
  if (vertices && d->geo_nv && e == hipSuccess) e = download(c, vertices, d->geo_v.p, 24 * d->geo_nv);
  if (ring_offset && e == hipSuccess) e = download(c, ring_offset, d->geo_roff.p, 8 * (d->total_faces + 1));
  if (ring && d->geo_nr && e == hipSuccess) e = download(c, ring, d->geo_ring.p, 2 * d->geo_nr);
  return c->hip(e, "download");
}

}  // extern "C"
