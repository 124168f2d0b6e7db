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
  if (d->geo_count.alloc(2 * n) != hipSuccess || d->geo_voff.alloc(n + 1) != hipSuccess ||
      d->geo_rbase.alloc(n + 1) != hipSuccess || d->geo_sums.alloc((n + 1023) / 1024 + 1) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  // count pass: live vertices and ring entries per cell.  The cells are
  // recomputed like compute does it: capacity 64 first, the overflows again
  // at 256 (the geometry kernels write nothing else).
  CellArgs a = base_args(d, t);
  a.work = nullptr;
  a.nwork = (int)n;
  a.xcd_chunk = plain_xcd_chunk(d);
  a.stats = nullptr;
  a.summary = nullptr;
  a.redo = d->redo.p + 1;
  a.nredo = d->redo.p;
  a.geo_count = d->geo_count.p;
  CellArgs r = a;  // the rerun of the overflows
  r.work = d->redo_list.p;
  r.xcd_chunk = 0;
  r.redo = nullptr;
  r.nredo = nullptr;
  uint32_t nredo = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) return c->fail(TESS_E_DEVICE, "event create failed");
  struct Release {
    hipEvent_t* e;
    ~Release() {
      for (int k = 0; k < 4; ++k) (void)hipEventDestroy(e[k]);
    }
  } release{ev};
  if ((st = c->hip(launch_fill_u32(d->geo_count.p, 2 * n, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(d->redo.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(hipEventRecord(ev[0], s), "event")) != TESS_OK ||
      (st = c->hip(launch_cells_geometry(1, a, s), "geometry count")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&nredo, d->redo.p, 4, hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "geometry count")) != TESS_OK)
    return st;
  if (nredo) {
    if (d->redo_list.alloc(nredo) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
    r.work = d->redo_list.p;
    r.nwork = (int)nredo;
    if ((st = c->hip(hipMemcpyAsync(d->redo_list.p, d->redo.p + 1, 4ull * nredo, hipMemcpyDeviceToDevice, s),
                     "redo")) != TESS_OK ||
        (st = c->hip(launch_cells_geometry(4, r, s), "geometry count")) != TESS_OK)
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
      (st = c->hip(hipStreamSynchronize(s), "geometry count")) != TESS_OK)
    return st;
  const uint64_t F = d->total_faces;
  if (d->geo_v.alloc(3 * d->geo_nv) != hipSuccess || d->geo_roff.alloc(F + 1) != hipSuccess ||
      d->geo_ring.alloc(d->geo_nr) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  // fill pass at the offsets (the overflows write nothing at 64 and are
  // filled by the rerun)
  for (CellArgs* p : {&a, &r}) {
    p->geo_count = nullptr;
    p->geo_voff = d->geo_voff.p;
    p->geo_foff = d->foff.p;
    p->geo_rbase = d->geo_rbase.p;
    p->geo_v = d->geo_v.p;
    p->geo_roff = d->geo_roff.p;
    p->geo_ring = d->geo_ring.p;
  }
  a.redo = nullptr;
  a.nredo = nullptr;
  float ms[2] = {0, 0};
  if ((st = c->hip(hipEventRecord(ev[2], s), "event")) != TESS_OK ||
      (st = c->hip(launch_cells_geometry(1, a, s), "geometry fill")) != TESS_OK ||
      (nredo && (st = c->hip(launch_cells_geometry(4, r, s), "geometry fill")) != TESS_OK) ||
      (st = c->hip(hipEventRecord(ev[3], s), "event")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(d->geo_roff.p + F, &d->geo_nr, 8, hipMemcpyHostToDevice, s), "upload")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "geometry fill")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[0], ev[0], ev[1]), "event")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[1], ev[2], ev[3]), "event")) != TESS_OK)
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
