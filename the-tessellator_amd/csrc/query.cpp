// query.cpp -- radius queries over a diagram's grid on the device
// (Celery::find_cells_in_radius / find_neighbors_in_{cell,real}_radius,
// celery.rs:700-855; Cell::compute_neighbor_cloud, interface.rs:348-365).
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "launch.h"
#include "runtime.h"

struct tess_query {
  tess_ctx* ctx = nullptr;
  size_t nq = 0;
  uint64_t total = 0;
  double ms[3] = {0, 0, 0};  // count pass, fill pass, whole call
  DevBuf<uint64_t> offsets;
  DevBuf<uint32_t> ids;
};

namespace {

// Shell-table extent for neighbour clouds: the smallest L = kTableL * 2^k
// whose cut-off min_a (L s_a)^2 exceeds every finite radius, so that the
// entries with key <= radius are a prefix of the table; NaN and +inf radii
// (no stop, celery.rs:1036) and large radii take the full order (L = cpd - 1).
int cloud_table_extent(int64_t cpd, const GridDev& G, const double* r, size_t m) {
  const int full = (int)(cpd - 1);
  double rmax = -std::numeric_limits<double>::infinity();
  for (size_t i = 0; i < m; ++i) {
    if (!(r[i] < std::numeric_limits<double>::infinity())) return full;
    rmax = std::max(rmax, r[i]);
  }
  auto sq = [](double t) { return t * t; };
  auto kcut = [&](int L) { return std::min(sq(L * G.sx), std::min(sq(L * G.sy), sq(L * G.sz))); };
  int L = kTableL;
  while (L < full && !(kcut(L) > rmax)) L *= 2;
  return L < full ? L : full;
}

// Query sets at least this large are taken in grid order.
constexpr size_t kQueryOrderMin = 4096;

// Largest upper-bound slot buffer of a filtered query set (8 GiB of ids);
// larger sets count exactly first.
constexpr uint64_t kBoundMax = 1ull << 31;

struct Events {
  hipEvent_t e[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  Events() {
    for (auto& x : e) (void)hipEventCreate(&x);
  }
  ~Events() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

tess_status query_impl(tess_diagram* d, int kind, const double* xyz, size_t nq, const double* radius,
                       size_t nradius, int64_t target_group, bool device_inputs, tess_query** out) {
  if (!d || !out) return TESS_E_INVALID;
  *out = nullptr;
  tess_ctx* c = d->ctx;
  if (kind < TESS_QUERY_CELLS || kind > TESS_QUERY_CLOUD) return c->fail(TESS_E_INVALID, "unknown query kind");
  if (nq && (!xyz || !radius)) return c->fail(TESS_E_INVALID, "null query positions or radii");
  if (nq && nradius != 1 && nradius != nq) return c->fail(TESS_E_INVALID, "nradius must be 1 or the query count");
  if (nq > (size_t)std::numeric_limits<int>::max() - 4) return c->fail(TESS_E_INVALID, "too many queries");
  if (target_group >= 0 && kind != TESS_QUERY_CLOUD)
    return c->fail(TESS_E_INVALID, "target_group applies to TESS_QUERY_CLOUD only");
  if (target_group >= 0 && !d->have_groups) return c->fail(TESS_E_INVALID, "diagram has no groups");
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  tess_status st;
  if (!d->binned && (st = bin(d)) != TESS_OK) return st;
  std::unique_ptr<tess_query> q(new (std::nothrow) tess_query);
  if (!q) return c->fail(TESS_E_NOMEM, "host allocation failed");
  q->ctx = c;
  q->nq = nq;
  if (q->offsets.alloc(nq + 1) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
  if (nq == 0) {
    const uint64_t zero = 0;
    if ((st = c->hip(hipMemcpy(q->offsets.p, &zero, 8, hipMemcpyHostToDevice), "upload")) != TESS_OK) return st;
    *out = q.release();
    return TESS_OK;
  }
  const GridDev& G = d->hgrid;
  const Table* t = nullptr;
  if (kind == TESS_QUERY_CLOUD) {
    // the table extent needs the radii on the host
    std::vector<double> hr;
    const double* r = radius;
    if (device_inputs) {
      hr.resize(nradius);
      if ((st = c->hip(hipMemcpy(hr.data(), radius, 8 * nradius, hipMemcpyDeviceToHost), "download")) != TESS_OK)
        return st;
      r = hr.data();
    }
    const int L = cloud_table_extent(d->cpd, G, r, nradius);
    if (L > 511) return c->fail(TESS_E_INVALID, "radius spans more than 511 grid cells");
    if (!(t = c->table(d->cpd, G.sx, G.sy, G.sz, L, &st))) return st;
  }
  DevBuf<double> qd, rd;
  auto& T = d->qtmp;  // serialised by the context lock
  DevBuf<uint32_t>& cnt = T.cnt;
  DevBuf<uint64_t>& sums = T.sums;
  if (cnt.alloc(nq) != hipSuccess || sums.alloc((nq + 1023) / 1024 + 1) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  if (device_inputs) {
    qd.view(const_cast<double*>(xyz), 3 * nq);
    rd.view(const_cast<double*>(radius), nradius);
  } else {
    if (qd.alloc(3 * nq) != hipSuccess || rd.alloc(nradius) != hipSuccess)
      return c->fail(TESS_E_NOMEM, "device allocation failed");
    if ((st = c->hip(hipMemcpyAsync(qd.p, xyz, 24 * nq, hipMemcpyHostToDevice, s), "upload")) != TESS_OK ||
        (st = c->hip(hipMemcpyAsync(rd.p, radius, 8 * nradius, hipMemcpyHostToDevice, s), "upload")) != TESS_OK)
      return st;
  }
  // Large sets: take the queries in grid order of their positions (the
  // results do not depend on the order the wavefronts take them in).
  DevBuf<double>&ox = T.ox, &oy = T.oy, &oz = T.oz;
  DevBuf<uint32_t>&obins = T.obins, &ocnt = T.ocnt, &odelim = T.odelim, &ocursor = T.ocursor, &oorder = T.oorder,
                   &osums = T.osums, &omisc = T.omisc;
  const bool ordered = nq >= kQueryOrderMin;
  if (ordered) {
    const size_t nb = (size_t)d->nbins;
    if (ox.alloc(nq) != hipSuccess || oy.alloc(nq) != hipSuccess || oz.alloc(nq) != hipSuccess ||
        obins.alloc(nq) != hipSuccess || oorder.alloc(nq) != hipSuccess || ocnt.alloc(nb) != hipSuccess ||
        odelim.alloc(nb + 1) != hipSuccess || ocursor.alloc(nb) != hipSuccess ||
        osums.alloc((nb + 1023) / 1024 + 1) != hipSuccess || omisc.alloc(4) != hipSuccess)
      return c->fail(TESS_E_NOMEM, "device allocation failed");
    SortBufs b{obins.p, ocnt.p, odelim.p, ocursor.p, oorder.p, nullptr, osums.p, nullptr, omisc.p};
    if ((st = c->hip(launch_soa(qd.p, (int)nq, ox.p, oy.p, oz.p, s), "query order")) != TESS_OK ||
        (st = c->hip(launch_grid_order(ox.p, oy.p, oz.p, (int)nq, d->grid.p, d->nbins, b, s), "query order")) !=
            TESS_OK)
      return st;
  }
  QueryArgs a{};
  a.g = G;
  a.order = ordered ? oorder.p : nullptr;
  a.sorted = d->sorted.p;
  a.delim = d->delim.p;
  a.sx = d->sx.p;
  a.sy = d->sy.p;
  a.sz = d->sz.p;
  a.groups = target_group >= 0 ? d->groups.p : nullptr;
  a.target_group = target_group;
  a.q = qd.p;
  a.radius = rd.p;
  a.rstride = nradius == 1 ? 0 : 1;
  a.nq = (int)nq;
  a.kind = kind;
  a.key = t ? t->key.p : nullptr;
  a.off = t ? t->off.p : nullptr;
  a.tlen = t ? t->len : 0;
  a.count = cnt.p;
  Events ev;
  // Filtered kinds (real radius, grouped clouds): the unfiltered walk gives
  // upper bounds from the delimiters alone, one fill pass writes the kept
  // ids into the bound slots and counts them, and a compaction packs them:
  // the points are read once instead of once per pass.
  const bool filtered = kind == TESS_QUERY_REAL_RADIUS || (kind == TESS_QUERY_CLOUD && target_group >= 0);
  if (filtered) {
    QueryArgs b = a;
    b.kind = kind == TESS_QUERY_REAL_RADIUS ? TESS_QUERY_CELL_RADIUS : TESS_QUERY_CLOUD;
    b.groups = nullptr;
    DevBuf<uint64_t>& boff = T.boff;
    uint64_t bound = 0;
    if (boff.alloc(nq + 1) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
    if ((st = c->hip(hipEventRecord(ev.e[0], s), "event")) != TESS_OK ||
        (st = c->hip(launch_query_count(b, s), "query bound")) != TESS_OK ||
        (st = c->hip(hipEventRecord(ev.e[1], s), "event")) != TESS_OK ||
        (st = c->hip(launch_face_offsets(cnt.p, (int)nq, boff.p, sums.p, s), "query scan")) != TESS_OK ||
        (st = c->hip(hipMemcpyAsync(&bound, boff.p + nq, 8, hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
        (st = c->hip(hipStreamSynchronize(s), "query bound")) != TESS_OK)
      return st;
    DevBuf<uint32_t>& slots = T.slots;
    if (bound <= kBoundMax && slots.alloc(bound) == hipSuccess) {
      a.offsets = boff.p;
      a.ids = slots.p;
      a.count = cnt.p;  // kept per query
      if ((st = c->hip(hipEventRecord(ev.e[2], s), "event")) != TESS_OK ||
          (st = c->hip(launch_query_fill(a, s), "query fill")) != TESS_OK ||
          (st = c->hip(hipEventRecord(ev.e[3], s), "event")) != TESS_OK ||
          (st = c->hip(launch_face_offsets(cnt.p, (int)nq, q->offsets.p, sums.p, s), "query scan")) != TESS_OK ||
          (st = c->hip(hipMemcpyAsync(&q->total, q->offsets.p + nq, 8, hipMemcpyDeviceToHost, s), "download")) !=
              TESS_OK ||
          (st = c->hip(hipStreamSynchronize(s), "query fill")) != TESS_OK)
        return st;
      if (q->ids.alloc(q->total) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
      float ms[3] = {0, 0, 0};
      if ((st = c->hip(launch_query_compact((int)nq, boff.p, q->offsets.p, slots.p, q->ids.p, s), "compact")) !=
              TESS_OK ||
          (st = c->hip(hipEventRecord(ev.e[4], s), "event")) != TESS_OK ||
          (st = c->hip(hipStreamSynchronize(s), "query compact")) != TESS_OK ||
          (st = c->hip(hipEventElapsedTime(&ms[0], ev.e[0], ev.e[1]), "event")) != TESS_OK ||
          (st = c->hip(hipEventElapsedTime(&ms[1], ev.e[2], ev.e[3]), "event")) != TESS_OK ||
          (st = c->hip(hipEventElapsedTime(&ms[2], ev.e[0], ev.e[4]), "event")) != TESS_OK)
        return st;
      for (int k = 0; k < 3; ++k) q->ms[k] = ms[k];
      *out = q.release();
      return TESS_OK;
    }
    a.count = cnt.p;  // too many slots: count exactly, then fill
  }
  if ((st = c->hip(hipEventRecord(ev.e[0], s), "event")) != TESS_OK ||
      (st = c->hip(launch_query_count(a, s), "query count")) != TESS_OK ||
      (st = c->hip(hipEventRecord(ev.e[1], s), "event")) != TESS_OK ||
      (st = c->hip(launch_face_offsets(cnt.p, (int)nq, q->offsets.p, sums.p, s), "query scan")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&q->total, q->offsets.p + nq, 8, hipMemcpyDeviceToHost, s), "download")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "query count")) != TESS_OK)
    return st;
  if (q->ids.alloc(q->total) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
  a.offsets = q->offsets.p;
  a.ids = q->ids.p;
  a.count = nullptr;
  float ms[3] = {0, 0, 0};
  if ((st = c->hip(hipEventRecord(ev.e[2], s), "event")) != TESS_OK ||
      (st = c->hip(launch_query_fill(a, s), "query fill")) != TESS_OK ||
      (st = c->hip(hipEventRecord(ev.e[3], s), "event")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "query fill")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[0], ev.e[0], ev.e[1]), "event")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[1], ev.e[2], ev.e[3]), "event")) != TESS_OK ||
      (st = c->hip(hipEventElapsedTime(&ms[2], ev.e[0], ev.e[3]), "event")) != TESS_OK)
    return st;
  for (int k = 0; k < 3; ++k) q->ms[k] = ms[k];
  *out = q.release();
  return TESS_OK;
}

}  // namespace

extern "C" {

tess_status tess_diagram_query(tess_diagram* d, int kind, const double* xyz, size_t nq, const double* radius,
                               size_t nradius, int64_t target_group, tess_query** out) {
  return query_impl(d, kind, xyz, nq, radius, nradius, target_group, false, out);
}

tess_status tess_diagram_query_device(tess_diagram* d, int kind, const double* device_xyz, size_t nq,
                                      const double* device_radius, size_t nradius, int64_t target_group,
                                      tess_query** out) {
  return query_impl(d, kind, device_xyz, nq, device_radius, nradius, target_group, true, out);
}

size_t tess_query_count(const tess_query* q) { return q ? q->nq : 0; }
uint64_t tess_query_total(const tess_query* q) { return q ? q->total : 0; }
tess_status tess_query_timings(const tess_query* q, double ms[3]) {
  if (!q || !ms) return TESS_E_INVALID;
  for (int k = 0; k < 3; ++k) ms[k] = q->ms[k];
  return TESS_OK;
}

tess_status tess_query_copy(const tess_query* q, uint64_t* offsets, uint32_t* ids) {
  if (!q) return TESS_E_INVALID;
  (void)hipSetDevice(q->ctx->device);
  hipError_t e = hipSuccess;
  if (offsets) e = download(q->ctx, offsets, q->offsets.p, 8 * (q->nq + 1));
  if (ids && q->total && e == hipSuccess) e = download(q->ctx, ids, q->ids.p, 4 * q->total);
  return q->ctx->hip(e, "download");
}

tess_status tess_query_device(const tess_query* q, const uint64_t** offsets, const uint32_t** ids) {
  if (!q) return TESS_E_INVALID;
  if (offsets) *offsets = q->offsets.p;
  if (ids) *ids = q->total ? q->ids.p : nullptr;
  return TESS_OK;
}

void tess_query_destroy(tess_query* q) {
  if (!q) return;
  (void)hipSetDevice(q->ctx->device);
  delete q;
}

}  // extern "C"
