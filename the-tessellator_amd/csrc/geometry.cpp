// geometry.cpp -- vertices and face rings of every cell of a diagram
// (Cell::compute_vertices / VoronoiFace::compute_vertices,
// interface.rs:367-370, :403-406; polyhedron.rs:885-919) as CSR on the device.
#include <string>

#include "launch.h"
#include "runtime.h"

extern "C" {

tess_status tess_diagram_compute_geometry(tess_diagram* d) {
  if (!d) return TESS_E_INVALID;
  tess_ctx* c = d->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  tess_status st;
  if (!d->computed && (st = compute_cells(d, nullptr)) != TESS_OK) return st;
  if (d->degenerate) return c->fail(TESS_E_DEGENERATE, "the diagram has failed cells");
  const size_t n = d->n;
  const GridDev& G = d->hgrid;
  Table* t = c->table(d->cpd, G.sx, G.sy, G.sz, kTableL, &st);
  if (!t) return st;
  // One pass recomputes every cell like compute does (capacity 64, the
  // overflows again at 256) and stages its geometry per work item; the
  // counts give the CSR offsets and a compaction copies the staged data.
  const size_t v1 = 64, h1 = kCellHalfEdges(1), v4 = 256, h4 = kCellHalfEdges(4);
  DevBuf<uint32_t> slot1, slot4;
  DevBuf<double> sv1, sv4;
  DevBuf<uint16_t> sr1, sr4, sro1, sro4;
  if (d->geo_count.alloc(2 * n) != hipSuccess || d->geo_voff.alloc(n + 1) != hipSuccess ||
      d->geo_rbase.alloc(n + 1) != hipSuccess || d->geo_sums.alloc((n + 1023) / 1024 + 1) != hipSuccess ||
      slot1.alloc(n) != hipSuccess || sv1.alloc(3 * v1 * n) != hipSuccess || sr1.alloc(h1 * n) != hipSuccess ||
      sro1.alloc((size_t)kFaceStage * n) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  CellArgs a = base_args(d, t);
  a.work = nullptr;
  a.nwork = (int)n;
  a.xcd_chunk = plain_xcd_chunk(d);
  a.stats = nullptr;
  a.summary = nullptr;
  a.redo = d->redo.p + 1;
  a.nredo = d->redo.p;
  a.geo_count = d->geo_count.p;
  a.geo_slot = slot1.p;
  a.geo_v = sv1.p;
  a.geo_ring = sr1.p;
  a.geo_roff = sro1.p;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) return c->fail(TESS_E_DEVICE, "event create failed");
  struct Release {
    hipEvent_t* e;
    ~Release() {
      for (int k = 0; k < 3; ++k) (void)hipEventDestroy(e[k]);
    }
  } release{ev};
  uint32_t nredo = 0;
  if ((st = c->hip(launch_fill_u32(d->geo_count.p, 2 * n, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(slot1.p, n, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(d->redo.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(hipEventRecord(ev[0], s), "event")) != TESS_OK ||
      (st = c->hip(launch_cells_geometry(1, a, s), "geometry")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&nredo, d->redo.p, 4, hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "geometry")) != TESS_OK)
    return st;
  if (nredo) {
    if (d->redo_list.alloc(nredo) != hipSuccess || slot4.alloc(nredo) != hipSuccess ||
        sv4.alloc(3 * v4 * nredo) != hipSuccess || sr4.alloc(h4 * nredo) != hipSuccess ||
        sro4.alloc((size_t)kFaceStage * nredo) != hipSuccess)
      return c->fail(TESS_E_NOMEM, "device allocation failed");
    CellArgs r = a;
    r.work = d->redo_list.p;
    r.nwork = (int)nredo;
    r.xcd_chunk = 0;
    r.redo = nullptr;
    r.nredo = nullptr;
    r.geo_slot = slot4.p;
    r.geo_v = sv4.p;
    r.geo_ring = sr4.p;
    r.geo_roff = sro4.p;
    if ((st = c->hip(hipMemcpyAsync(d->redo_list.p, d->redo.p + 1, 4ull * nredo, hipMemcpyDeviceToDevice, s),
                     "redo")) != TESS_OK ||
        (st = c->hip(launch_fill_u32(slot4.p, nredo, 0u, s), "fill")) != TESS_OK ||
        (st = c->hip(launch_cells_geometry(4, r, s), "geometry")) != TESS_OK)
      return st;
  }
  if ((st = c->hip(hipEventRecord(ev[1], s), "event")) != TESS_OK ||
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
  float ms[2] = {0, 0};
  if ((st = c->hip(launch_geo_compact(slot1.p, (int)n, (int)v1, (int)h1, sv1.p, sr1.p, sro1.p, d->geo_count.p,
                                      (int)n, d->geo_voff.p, d->geo_rbase.p, d->foff.p, d->geo_v.p, d->geo_ring.p,
                                      d->geo_roff.p, s),
                   "compact")) != TESS_OK ||
      (nredo && (st = c->hip(launch_geo_compact(slot4.p, (int)nredo, (int)v4, (int)h4, sv4.p, sr4.p, sro4.p,
                                                d->geo_count.p, (int)n, d->geo_voff.p, d->geo_rbase.p, d->foff.p,
                                                d->geo_v.p, d->geo_ring.p, d->geo_roff.p, s),
                             "compact")) != TESS_OK) ||
      (st = c->hip(hipMemcpyAsync(d->geo_roff.p + F, &d->geo_nr, 8, hipMemcpyHostToDevice, s), "upload")) !=
          TESS_OK ||
      (st = c->hip(hipEventRecord(ev[2], s), "event")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "geometry")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[0], ev[0], ev[1]), "event")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[1], ev[1], ev[2]), "event")) != TESS_OK)
    return st;
  d->geo_ms[0] = ms[0];
  d->geo_ms[1] = ms[1];
  d->geo_ready = true;
  return TESS_OK;
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
  if (vertex_offset) e = hipMemcpy(vertex_offset, d->geo_voff.p, 8 * (d->n + 1), hipMemcpyDeviceToHost);
  if (vertices && d->geo_nv && e == hipSuccess)
    e = hipMemcpy(vertices, d->geo_v.p, 24 * d->geo_nv, hipMemcpyDeviceToHost);
  if (ring_offset && e == hipSuccess)
    e = hipMemcpy(ring_offset, d->geo_roff.p, 8 * (d->total_faces + 1), hipMemcpyDeviceToHost);
  if (ring && d->geo_nr && e == hipSuccess) e = hipMemcpy(ring, d->geo_ring.p, 2 * d->geo_nr, hipMemcpyDeviceToHost);
  return c->hip(e, "download");
}

}  // extern "C"
