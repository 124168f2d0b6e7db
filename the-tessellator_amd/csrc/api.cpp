// api.cpp -- host runtime behind include/tess.h.
//
// Owns device memory and the HIP stream, builds the canonical shell table
// (celery.rs:418-679) on the host once per grid shape, and sequences the
// device passes: bounds -> counting sort -> gather -> cell kernel (with a
// larger-capacity rerun for cells that outgrow the default) -> face offsets
// -> CSR compaction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "runtime.h"

using namespace tess;

namespace tess {

// ---------------------------------------------------------------- shell table
// Generation order of celery.rs:437-673: the centre, then offsets with three,
// two (xy, xz, yz) and one (x, y, z) non-zero components, each block looping
// over |o|-1 in lexicographic order with the sign patterns listed there.
uint64_t generation_index(int64_t cpd, int ox, int oy, int oz) {
  const uint64_t M = (uint64_t)(cpd - 1);
  const uint64_t a = ox ? (uint64_t)std::abs(ox) - 1 : 0;
  const uint64_t b = oy ? (uint64_t)std::abs(oy) - 1 : 0;
  const uint64_t c = oz ? (uint64_t)std::abs(oz) - 1 : 0;
  const uint64_t two = 1 + 8 * M * M * M, one = two + 12 * M * M;
  if (!ox && !oy && !oz) return 0;
  if (ox && oy && oz) {
    static const int order[8] = {0, 1, 2, 4, 3, 5, 6, 7};
    const int bits = ((ox < 0) << 2) | ((oy < 0) << 1) | (oz < 0);
    return 1 + 8 * ((a * M + b) * M + c) + order[bits];
  }
  if (ox && oy) return two + 4 * (a * M + b) + (((ox < 0) << 1) | (oy < 0));
  if (ox && oz) return two + 4 * M * M + 4 * (a * M + c) + (((ox < 0) << 1) | (oz < 0));
  if (oy && oz) return two + 8 * M * M + 4 * (b * M + c) + (((oy < 0) << 1) | (oz < 0));
  if (ox) return one + 2 * a + (ox < 0);
  if (oy) return one + 2 * M + 2 * b + (oy < 0);
  return one + 4 * M + 2 * c + (oz < 0);
}

// celery.rs:419-427: squared gap distance, (sq(gx sx) + sq(gy sy)) + sq(gz sz).
double shell_key(int ox, int oy, int oz, double sx, double sy, double sz) {
  if (!ox && !oy && !oz) return -1.0;
  auto sq = [](double t) { return t * t; };
  const double gx = ox ? std::abs(ox) - 1 : 0, gy = oy ? std::abs(oy) - 1 : 0,
               gz = oz ? std::abs(oz) - 1 : 0;
  return sq(gx * sx) + sq(gy * sy) + sq(gz * sz);
}

// Entries with |o|_inf <= L sorted by (key, generation index).  For L < cpd-1
// only keys below kcut = min_a (L s_a)^2 are kept: every offset outside the
// cube has a key >= kcut, so the result is a prefix of the full order.
void build_table_host(int64_t cpd, double sx, double sy, double sz, int L, std::vector<double>& key,
                      std::vector<int32_t>& off, double* kcut) {
  if (L > cpd - 1) L = (int)(cpd - 1);
  const bool full = (L == cpd - 1);
  double kc = std::numeric_limits<double>::infinity();
  if (!full) {
    auto sq = [](double t) { return t * t; };
    kc = std::min(sq(L * sx), std::min(sq(L * sy), sq(L * sz)));
  }
  struct E {
    double k;
    uint64_t g;
    int32_t o;
  };
  std::vector<E> es;
  es.reserve((size_t)(2 * L + 1) * (2 * L + 1) * (2 * L + 1));
  for (int i = -L; i <= L; ++i)
    for (int j = -L; j <= L; ++j)
      for (int k = -L; k <= L; ++k) {
        const double kk = shell_key(i, j, k, sx, sy, sz);
        if (!full && !(kk < kc)) continue;
        es.push_back({kk, generation_index(cpd, i, j, k), pack_off(i, j, k)});
      }
  std::sort(es.begin(), es.end(), [](const E& a, const E& b) { return a.k != b.k ? a.k < b.k : a.g < b.g; });
  key.resize(es.size());
  off.resize(es.size());
  for (size_t t = 0; t < es.size(); ++t) {
    key[t] = es[t].k;
    off[t] = es[t].o;
  }
  *kcut = kc;
}

uint64_t bits(double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
}

int64_t icbrt_floor(double x) {
  if (!(x >= 1.0)) return 0;
  int64_t k = (int64_t)std::cbrt(x);
  while (k > 0 && (double)k * (double)k * (double)k > x) --k;
  while ((double)(k + 1) * (double)(k + 1) * (double)(k + 1) <= x) ++k;
  return k;
}

}  // namespace tess

namespace tess {

void host_copy(char* dst, const char* src, size_t len) {
  if (len == 0) return;  // (memcpy takes no null pointers, even for no bytes)
  const size_t kPiece = size_t(4) << 20;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t t = std::min<size_t>({size_t(8), size_t(hw), (len + kPiece - 1) / kPiece});
  if (t <= 1) {
    std::memcpy(dst, src, len);
    return;
  }
  const size_t per = (len + t - 1) / t;
  std::vector<std::thread> th;
  th.reserve(t - 1);
  size_t k = 1;
  for (; k < t && k * per < len; ++k) {
    const size_t o = k * per, l = std::min(per, len - o);
    try {
      th.emplace_back([=] { std::memcpy(dst + o, src + o, l); });
    } catch (...) {  // no thread: this piece and the rest on this one
      break;
    }
  }
  for (size_t o = k * per; o < len; o += per) std::memcpy(dst + o, src + o, std::min(per, len - o));
  std::memcpy(dst, src, std::min(per, len));
  for (auto& x : th) x.join();
}

hipError_t download(tess_ctx* c, void* dst, const void* src, size_t bytes) {
  constexpr size_t kMin = size_t(16) << 20;
  if (bytes < kMin) return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  // at least four chunks (so that transfers and host copies overlap), of
  // 8 MiB to one staging buffer each
  const size_t kChunk = std::min(Staging::kBytes, std::max(size_t(8) << 20, (bytes / 4 + 4095) & ~size_t(4095)));
  Staging& S = c->stage;
  std::lock_guard<std::mutex> g(S.mu);
  for (int k = 0; k < 2; ++k) {
    if (!S.buf[k] && hipHostMalloc(&S.buf[k], Staging::kBytes, hipHostMallocDefault) != hipSuccess) {
      S.buf[k] = nullptr;
      return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);  // no pinned memory: the plain copy
    }
    if (!S.ev[k] && hipEventCreateWithFlags(&S.ev[k], hipEventDisableTiming) != hipSuccess) {
      S.ev[k] = nullptr;
      return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    }
  }
  const size_t nch = (bytes + kChunk - 1) / kChunk;
  auto len_of = [&](size_t i) { return std::min(kChunk, bytes - i * kChunk); };
  // chunk i goes to buffer i & 1; chunk i + 1 is issued before chunk i is
  // copied out, into the buffer chunk i - 1 left (copied out already)
  auto issue = [&](size_t i) {
    hipError_t e = hipMemcpyAsync(S.buf[i & 1], static_cast<const char*>(src) + i * kChunk, len_of(i),
                                  hipMemcpyDeviceToHost, c->stream);
    return e == hipSuccess ? hipEventRecord(S.ev[i & 1], c->stream) : e;
  };
  hipError_t e = issue(0);
  for (size_t i = 0; i < nch && e == hipSuccess; ++i) {
    if (i + 1 < nch && (e = issue(i + 1)) != hipSuccess) break;
    if ((e = hipEventSynchronize(S.ev[i & 1])) != hipSuccess) break;
    host_copy(static_cast<char*>(dst) + i * kChunk, static_cast<const char*>(S.buf[i & 1]), len_of(i));
  }
  if (e != hipSuccess) (void)hipStreamSynchronize(c->stream);  // nothing left in flight into the staging
  return e;
}

tess_status alloc_all(tess_diagram* d) {
  tess_ctx* c = d->ctx;
  const size_t n = d->n;
  const int nparts = (int)std::min<size_t>(256, (n + 255) / 256);
  const size_t nb = (size_t)d->nbins;
  bool ok = d->x.alloc(n) == hipSuccess && d->y.alloc(n) == hipSuccess && d->z.alloc(n) == hipSuccess &&
            d->grid.alloc(1) == hipSuccess && d->part.alloc(6 * (size_t)nparts) == hipSuccess &&
            d->bins.alloc(n) == hipSuccess && d->cnt.alloc(nb) == hipSuccess &&
            d->delim.alloc(nb + 1) == hipSuccess && d->cursor.alloc(nb) == hipSuccess &&
            d->sorted.alloc(n) == hipSuccess && d->scratch.alloc(n) == hipSuccess &&
            d->sums.alloc((nb + 1023) / 1024 + 1) == hipSuccess && d->big.alloc(nb) == hipSuccess &&
            d->misc.alloc(4) == hipSuccess && d->sx.alloc(n) == hipSuccess && d->sy.alloc(n) == hipSuccess &&
            (d->opts.point_order != TESS_ORDER_MORTON || d->codes.alloc(n) == hipSuccess) &&
            d->sz.alloc(n) == hipSuccess && d->volume.alloc(n) == hipSuccess &&
            d->fcount.alloc(n) == hipSuccess && d->status.alloc(n) == hipSuccess &&
            d->redo.alloc(n + 1) == hipSuccess && d->snbr.alloc(arena_capacity(n)) == hipSuccess &&
            d->sarea.alloc(arena_capacity(n)) == hipSuccess && d->foff.alloc(n + 1) == hipSuccess &&
            d->fsums.alloc((n + 1023) / 1024 + 1) == hipSuccess && d->work1.alloc(1) == hipSuccess &&
            d->summary.alloc(1) == hipSuccess && d->tredo.alloc(n + 1) == hipSuccess &&
            d->spill.alloc(n + 1) == hipSuccess && d->fpos.alloc(n) == hipSuccess &&
            d->arena_ctr.alloc(kArenaWords) == hipSuccess;
  d->arena_next = d->arena_ctr.p;
  d->arena_cap = arena_capacity(n);
  if (ok && d->opts.collect_stats) ok = d->stats.alloc(n) == hipSuccess;
  if (!ok) return c->fail(TESS_E_NOMEM, "device allocation failed");
  return TESS_OK;
}

tess_status bin_enqueue(tess_diagram* d, hipStream_t s) {
  tess_ctx* c = d->ctx;
  const int n = (int)d->n;
  const int nparts = (int)std::min<size_t>(256, (d->n + 255) / 256);
  tess_status st;
  if ((st = c->hip(launch_bounds(d->x.p, d->y.p, d->z.p, n, (int)d->cpd, d->part.p, nparts, d->grid.p, s),
                   "bounds")) != TESS_OK)
    return st;
  SortBufs b{d->bins.p, d->cnt.p, d->delim.p, d->cursor.p, d->sorted.p, d->scratch.p, d->sums.p, d->big.p,
             d->misc.p};
  const uint32_t* key = nullptr;
  if (d->opts.point_order == TESS_ORDER_MORTON) {  // a cell's points in the internal (Morton) order
    if ((st = c->hip(launch_morton(d->x.p, d->y.p, d->z.p, n, d->grid.p, d->codes.p, s), "morton")) != TESS_OK)
      return st;
    key = d->codes.p;
  }
  if ((st = c->hip(launch_counting_sort(d->x.p, d->y.p, d->z.p, n, d->grid.p, d->nbins, b, s, key), "sort")) !=
      TESS_OK)
    return st;
  return c->hip(launch_gather(d->sorted.p, n, d->x.p, d->y.p, d->z.p, d->sx.p, d->sy.p, d->sz.p, s), "gather");
}

tess_status bin(tess_diagram* d) {
  tess_ctx* c = d->ctx;
  hipStream_t s = c->stream;
  tess_status st;
  if ((st = bin_enqueue(d, s)) != TESS_OK) return st;
  if ((st = c->hip(hipMemcpyAsync(&d->hgrid, d->grid.p, sizeof(GridDev), hipMemcpyDeviceToHost, s),
                   "grid copy")) != TESS_OK)
    return st;
  // every point must have landed in exactly one grid cell before any pass
  // dereferences the sorted indices
  uint32_t total = 0;
  if ((st = c->hip(hipMemcpyAsync(&total, d->delim.p + d->nbins, 4, hipMemcpyDeviceToHost, s), "grid copy")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "binning")) != TESS_OK)
    return st;
  if (total != (uint32_t)d->n) return c->fail(TESS_E_DEVICE, "counting sort lost points");
  d->binned = true;
  d->inv_ready = false;
  return TESS_OK;
}

// Vertex capacity of the first pass (64 by default: highest occupancy);
// cells that outgrow it are rerun with 256.
int pick_vpl(const tess_diagram* d) {
  switch (d->opts.vertex_capacity) {
    case 128:
      return 2;
    case 256:
      return 4;
    default:
      return 1;
  }
}

CellArgs base_args(tess_diagram* d, const Table* t) {
  CellArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = (int)d->n;
  a.spx = d->sx.p;
  a.spy = d->sy.p;
  a.spz = d->sz.p;
  a.sorted = d->sorted.p;
  a.delim = d->delim.p;
  a.grid = d->grid.p;
  a.tkey = t->key.p;
  a.toff = t->off.p;
  a.tlen = t->len;
  a.kcut = t->kcut;
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = d->lo[k];
    a.hi[k] = d->hi[k];
  }
  a.groups = d->have_groups ? d->groups.p : nullptr;
  a.target_group = -1;
  a.radius_limit = std::numeric_limits<double>::quiet_NaN();
  a.volume = d->volume.p;
  a.fcount = d->fcount.p;
  a.snbr = d->snbr.p;
  a.sarea = d->sarea.p;
  a.fpos = d->fpos.p;
  a.arena_next = d->arena_next;
  a.arena_cap = d->arena_cap;
  a.status = d->status.p;
  a.stats = d->opts.collect_stats ? d->stats.p : nullptr;
  a.summary = d->summary.p;
  return a;
}

tess_status create_common(tess_ctx* ctx, size_t n, const double lo[3], const double hi[3],
                          const tess_opts* opts, tess_diagram** out) {
  if (!ctx || !out || !lo || !hi) return TESS_E_INVALID;
  *out = nullptr;
  if (n == 0) return ctx->fail(TESS_E_INVALID, "empty point set");
  if (n > (size_t)INT32_MAX / kFaceStage) return ctx->fail(TESS_E_INVALID, "too many points");
  for (int k = 0; k < 3; ++k)
    if (!(hi[k] > lo[k])) return ctx->fail(TESS_E_INVALID, "container box is empty");
  tess_diagram* d = new (std::nothrow) tess_diagram;
  if (!d) return ctx->fail(TESS_E_NOMEM, "host allocation failed");
  d->ctx = ctx;
  if (opts)
    d->opts = *opts;
  else
    tess_opts_default(&d->opts);
  if (!(d->opts.cell_density > 0)) d->opts.cell_density = 1.25;
  if (d->opts.point_order != TESS_ORDER_INPUT && d->opts.point_order != TESS_ORDER_MORTON) {
    delete d;
    return ctx->fail(TESS_E_INVALID, "unknown point_order");
  }
  d->n = n;
  d->cpd = icbrt_floor((double)n / d->opts.cell_density) + 1;  // celery.rs:161-162
  if (d->cpd > 1000) {
    delete d;
    return ctx->fail(TESS_E_INVALID, "grid too fine (cell_density too small)");
  }
  d->nbins = (int)(d->cpd * d->cpd * d->cpd);
  for (int k = 0; k < 3; ++k) {
    d->lo[k] = lo[k];
    d->hi[k] = hi[k];
  }
  std::memset(&d->tim, 0, sizeof(d->tim));
  tess_status st = alloc_all(d);
  if (st != TESS_OK) {
    delete d;
    return st;
  }
  *out = d;
  return TESS_OK;
}

// Run length of the XCD mapping of a whole-set pass: the caller's choice, else
// 256 cells (measured on one MI355X against 64 ... 16384 and one eighth of
// the set: neighbouring cells share an L2, dense regions spread over all XCDs).
int plain_xcd_chunk(const tess_diagram* d) { return d->opts.xcd_chunk > 0 ? d->opts.xcd_chunk : 256; }

tess_status run_cells(tess_diagram* d, Table* t, int vpl, const uint32_t* work, int nwork, bool push_redo) {
  tess_ctx* c = d->ctx;
  if (nwork <= 0) return TESS_OK;
  CellArgs a = base_args(d, t);
  a.work = work;
  a.nwork = nwork;
  // whole set in grid order: each XCD takes runs of consecutive cells
  a.xcd_chunk = work ? 0 : plain_xcd_chunk(d);
  // capacity overflows: below 256 vertices to the rerun list, at 256 to the
  // spill list; shell-table exhaustions to their own list
  DevBuf<uint32_t>& cap = vpl < 4 ? d->redo : d->spill;
  a.redo = push_redo ? cap.p + 1 : nullptr;
  a.nredo = push_redo ? cap.p : nullptr;
  a.tredo = d->tredo.p + 1;
  a.ntredo = d->tredo.p;
  a.defer_table = 1;
  if (GeoStage* g = d->geo_stage) {  // the same pass, staging the geometry too
    a.geo_count = g->count;
    GeoStaging& b = work ? g->rerun : g->first;
    a.geo_slot = b.slot.p;
    a.geo_v = b.v.p;
    a.geo_ring = b.ring.p;
    a.geo_roff = b.roff.p;
    return c->hip(launch_cells_geometry(vpl, a, true, c->stream), "cell kernel");
  }
  return c->hip(launch_cells(vpl, a, c->stream), "cell kernel");
}


namespace {

struct Events {
  hipEvent_t e[5] = {};
  bool ok = true;
  Events() {
    for (auto& x : e)
      if (hipEventCreate(&x) != hipSuccess) ok = false;
  }
  ~Events() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
  float ms(int a, int b) const {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e[a], e[b]);
    return t;
  }
};

}  // namespace

tess_status compute_cells(tess_diagram* d, const Planner* plan) {
  tess_ctx* c = d->ctx;
  hipStream_t s = c->stream;
  tess_status st;
  Events ev;
  if (!ev.ok) return c->fail(TESS_E_DEVICE, "event create failed");
  d->geo_ready = false;
  (void)hipEventRecord(ev.e[0], s);
  if ((st = bin(d)) != TESS_OK) return st;
  (void)hipEventRecord(ev.e[1], s);
  auto t0 = std::chrono::steady_clock::now();
  const GridDev& G = d->hgrid;
  Table* t = c->table(d->cpd, G.sx, G.sy, G.sz, kTableL, &st);
  auto t1 = std::chrono::steady_clock::now();
  if (!t) return st;
  const int n = (int)d->n;
  // work lists: one pass over every cell at the default capacity, or the
  // planner's lists, largest capacity first
  Schedule sch;
  // geometry staging indexes the first pass by grid position: whole-set pass only
  if (plan && d->geo_stage) return c->fail(TESS_E_STATE, "geometry staging needs the whole-set pass");
  if (plan) {
    if ((st = (*plan)(d, sch)) != TESS_OK) return st;
  } else {
    sch.count[pick_vpl(d) >> 1] = n;  // vpl 1, 2, 4 -> slot 0, 1, 2
  }
  (void)hipEventRecord(ev.e[2], s);
  if ((st = c->hip(hipMemsetAsync(d->summary.p, 0, sizeof(Summary), s), "memset")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(d->redo.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(d->tredo.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32(d->spill.p, 1, 0u, s), "fill")) != TESS_OK ||
      (st = c->hip(launch_fill_u32((uint32_t*)d->arena_next, 2 * kArenaWords, 0u, s), "fill")) != TESS_OK)
    return st;
  // the planner's long-search list first (capacity 256, a table sized from
  // the largest predicted radius), then the capacity lists, largest first
  if (sch.count[3]) {
    Table* tl = c->table(d->cpd, G.sx, G.sy, G.sz, sch.long_l, &st);
    if (!tl) return st;
    if ((st = run_cells(d, tl, 4, sch.list[3], sch.count[3], true)) != TESS_OK) return st;
  }
  static const int kVpl[3] = {1, 2, 4};
  for (int slot = 2; slot >= 0; --slot)
    if (sch.count[slot] && (st = run_cells(d, t, kVpl[slot], sch.list[slot], sch.count[slot], true)) != TESS_OK)
      return st;
  // cells that outgrew their capacity: rerun with the largest variant; after
  // a 64-vertex whole-set pass, first at 128 vertices (nearly all of them
  // fit, and that variant keeps more wavefronts resident) and the rest at 256
  uint32_t nredo = 0;
  if ((st = c->hip(hipMemcpyAsync(&nredo, d->redo.p, 4, hipMemcpyDeviceToHost, s), "redo")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "cell kernel")) != TESS_OK)
    return st;
  if (nredo) {
    if (d->redo_list.alloc(nredo) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
    if ((st = c->hip(hipMemcpyAsync(d->redo_list.p, d->redo.p + 1, 4ull * nredo, hipMemcpyDeviceToDevice, s),
                     "redo")) != TESS_OK)
      return st;
    if (!plan && !d->geo_stage && pick_vpl(d) == 1) {
      uint32_t nredo2 = 0;
      if ((st = c->hip(launch_fill_u32(d->redo.p, 1, 0u, s), "fill")) != TESS_OK ||
          (st = run_cells(d, t, 2, d->redo_list.p, (int)nredo, true)) != TESS_OK ||
          (st = c->hip(hipMemcpyAsync(&nredo2, d->redo.p, 4, hipMemcpyDeviceToHost, s), "redo")) != TESS_OK ||
          (st = c->hip(hipStreamSynchronize(s), "cell kernel")) != TESS_OK)
        return st;
      if (nredo2) {
        if (d->redo_list2.alloc(nredo2) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
        if ((st = c->hip(hipMemcpyAsync(d->redo_list2.p, d->redo.p + 1, 4ull * nredo2, hipMemcpyDeviceToDevice, s),
                         "redo")) != TESS_OK ||
            (st = run_cells(d, t, 4, d->redo_list2.p, (int)nredo2, true)) != TESS_OK)
          return st;
      }
    } else {
      if (GeoStage* g = d->geo_stage) {  // staging of the rerun items
        g->nredo = nredo;
        if ((st = geo_staging_alloc(d, g->rerun, nredo, 4)) != TESS_OK) return st;
      }
      if ((st = run_cells(d, t, 4, d->redo_list.p, (int)nredo, true)) != TESS_OK) return st;
    }
  }
  // cells past the shell table or the largest LDS capacity
  std::vector<uint32_t> tab, spl;
  if ((st = read_list(c, d->tredo.p, tab)) != TESS_OK || (st = read_list(c, d->spill.p, spl)) != TESS_OK) return st;
  FinishCounts fin;
  const int ntab = (int)tab.size();
  if (!tab.empty() || !spl.empty()) {
    CellArgs a = base_args(d, t);
    if ((st = finish_cells(d, a, std::move(tab), std::move(spl), &fin, d->geo_stage)) != TESS_OK) return st;
  }
  // planned passes: a cell whose search outran its table was under-sized by
  // the prediction too (marked after its rerun wrote its status)
  if (plan && ntab &&
      (st = c->hip(launch_mark_status(d->tredo.p + 1, ntab, d->sorted.p, d->status.p,
                                      d->opts.collect_stats ? d->stats.p : nullptr, TESS_CELL_FALLBACK, s),
                   "mark")) != TESS_OK)
    return st;
  if (nredo && (st = c->hip(launch_mark_status(d->redo_list.p, (int)nredo, d->sorted.p, d->status.p,
                                               d->opts.collect_stats ? d->stats.p : nullptr,
                                               TESS_CELL_RERUN | (plan ? TESS_CELL_FALLBACK : 0u), s),
                            "mark")) != TESS_OK)
    return st;
  (void)hipEventRecord(ev.e[3], s);
  if ((st = c->hip(launch_face_offsets(d->fcount.p, n, d->foff.p, d->fsums.p, s), "face offsets")) != TESS_OK)
    return st;
  uint64_t total = 0;
  unsigned long long used = 0;
  if ((st = c->hip(hipMemcpyAsync(&total, d->foff.p + n, 8, hipMemcpyDeviceToHost, s), "total")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&used, d->arena_next + (size_t)kArenaStride * (kArenaParts + 1), 8, hipMemcpyDeviceToHost, s),
                   "arena")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "face offsets")) != TESS_OK)
    return st;
  if (used) return c->fail(TESS_E_CAPACITY, "the face arena is full");
  if (d->nbr.alloc(total) != hipSuccess || d->area.alloc(total) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  if ((st = c->hip(launch_compact(d->fcount.p, d->foff.p, n, d->snbr.p, d->sarea.p, d->nbr.p, d->area.p, s,
                                  d->fpos.p),
                   "compact")) != TESS_OK)
    return st;
  (void)hipEventRecord(ev.e[4], s);
  Summary sm2;
  if ((st = c->hip(hipMemcpyAsync(&sm2, d->summary.p, sizeof(Summary), hipMemcpyDeviceToHost, s), "summary")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "compute")) != TESS_OK)
    return st;
  d->tim.bin_ms = ev.ms(0, 1);
  d->tim.plan_ms = ev.ms(1, 2);
  d->tim.cells_ms = ev.ms(2, 3);
  d->tim.compact_ms = ev.ms(3, 4);
  d->tim.total_ms = ev.ms(0, 4);
  d->tim.table_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  d->tim.capacity_reruns = nredo;
  d->tim.fallback_cells = plan ? nredo + (uint64_t)ntab : 0;
  d->tim.planned_long = sch.count[3];
  d->tim.flops = sm2.flops;
  d->tim.incident_cuts = sm2.incident_cuts;
  d->table_reruns = d->tim.table_reruns = fin.table;
  d->spill_cells = d->tim.spill_cells = fin.spill;
  d->total_faces = total;
  d->computed = true;
  const uint64_t bad = sm2.degenerate;
  d->degenerate = bad;
  if (fin.failed)
    return c->fail(TESS_E_CAPACITY, std::to_string(fin.failed) + " cells exceeded the spill workspace limit");
  if (bad) return c->fail(TESS_E_DEGENERATE, std::to_string(bad) + " cells failed topology checks");
  return TESS_OK;
}

}  // namespace tess

extern "C" {

void tess_opts_default(tess_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->cell_density = 1.25;
  o->mode = TESS_MODE_PLAIN;
  o->vertex_capacity = 0;
  o->collect_stats = 0;
}

tess_status tess_ctx_create(int dev, tess_ctx** out) {
  if (!out) return TESS_E_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || dev < 0 || dev >= count) return TESS_E_DEVICE;
  if (hipSetDevice(dev) != hipSuccess) return TESS_E_DEVICE;
  tess_ctx* c = new (std::nothrow) tess_ctx;
  if (!c) return TESS_E_NOMEM;
  c->device = dev;
  (void)noalloc_setup(dev);  // the scratch address of this device, once
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return TESS_E_DEVICE;
  }
  *out = c;
  return TESS_OK;
}

void tess_ctx_destroy(tess_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->tables.clear();
  c->diag.reset();
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* tess_last_error(const tess_ctx* c) { return c ? c->err.c_str() : "null context"; }
void* tess_ctx_stream(tess_ctx* c) { return c ? (void*)c->stream : nullptr; }
tess_status tess_ctx_synchronize(tess_ctx* c) {
  if (!c) return TESS_E_INVALID;
  return c->hip(hipStreamSynchronize(c->stream), "synchronize");
}

tess_status tess_diagram_create(tess_ctx* ctx, const double* xyz, size_t n, const double lo[3],
                                const double hi[3], const int64_t* groups, const tess_opts* opts,
                                tess_diagram** out) {
  if (!ctx) return TESS_E_INVALID;
  if (!xyz) return ctx->fail(TESS_E_INVALID, "null points");
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  tess_status st = create_common(ctx, n, lo, hi, opts, out);
  if (st != TESS_OK) return st;
  tess_diagram* d = *out;
  if (d->xyz_tmp.alloc(3 * n) != hipSuccess) {
    delete d;
    *out = nullptr;
    return ctx->fail(TESS_E_NOMEM, "device allocation failed");
  }
  hipStream_t s = ctx->stream;
  st = ctx->hip(hipMemcpyAsync(d->xyz_tmp.p, xyz, 24 * n, hipMemcpyHostToDevice, s), "upload");
  if (st == TESS_OK) st = ctx->hip(launch_soa(d->xyz_tmp.p, (int)n, d->x.p, d->y.p, d->z.p, s), "soa");
  if (st == TESS_OK && groups) {
    d->have_groups = true;
    if (d->groups.alloc(n) != hipSuccess)
      st = ctx->fail(TESS_E_NOMEM, "device allocation failed");
    else
      st = ctx->hip(hipMemcpyAsync(d->groups.p, groups, 8 * n, hipMemcpyHostToDevice, s), "upload");
  }
  if (st == TESS_OK) st = ctx->hip(hipStreamSynchronize(s), "upload");
  d->xyz_tmp.reset();
  if (st != TESS_OK) {
    delete d;
    *out = nullptr;
  }
  return st;
}

tess_status tess_diagram_create_device(tess_ctx* ctx, const double* dxyz, size_t n, const double lo[3],
                                       const double hi[3], const int64_t* dgroups, const tess_opts* opts,
                                       tess_diagram** out) {
  if (!ctx) return TESS_E_INVALID;
  if (!dxyz) return ctx->fail(TESS_E_INVALID, "null points");
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  tess_status st = create_common(ctx, n, lo, hi, opts, out);
  if (st != TESS_OK) return st;
  tess_diagram* d = *out;
  hipStream_t s = ctx->stream;
  st = ctx->hip(launch_soa(dxyz, (int)n, d->x.p, d->y.p, d->z.p, s), "soa");
  if (st == TESS_OK && dgroups) {
    d->have_groups = true;
    if (d->groups.alloc(n) != hipSuccess)
      st = ctx->fail(TESS_E_NOMEM, "device allocation failed");
    else
      st = ctx->hip(hipMemcpyAsync(d->groups.p, dgroups, 8 * n, hipMemcpyDeviceToDevice, s), "copy");
  }
  if (st == TESS_OK) st = ctx->hip(hipStreamSynchronize(s), "soa");
  if (st != TESS_OK) {
    delete d;
    *out = nullptr;
  }
  return st;
}

void tess_diagram_destroy(tess_diagram* d) {
  if (!d) return;
  (void)hipSetDevice(d->ctx->device);
  (void)hipStreamSynchronize(d->ctx->stream);
  delete d;
}

size_t tess_diagram_size(const tess_diagram* d) { return d ? d->n : 0; }

tess_status tess_diagram_compute(tess_diagram* d) {
  if (!d) return TESS_E_INVALID;
  std::lock_guard<std::mutex> g(d->ctx->mu);
  (void)hipSetDevice(d->ctx->device);
  return compute_cells(d, nullptr);
}

tess_status tess_diagram_total_faces(const tess_diagram* d, uint64_t* out) {
  if (!d || !out) return TESS_E_INVALID;
  if (!d->computed) return TESS_E_STATE;
  *out = d->total_faces;
  return TESS_OK;
}

tess_status tess_diagram_copy_cells(const tess_diagram* d, double* volume, uint32_t* fc, uint64_t* fo,
                                    int32_t* nbr, double* area) {
  if (!d) return TESS_E_INVALID;
  if (!d->computed) return d->ctx->fail(TESS_E_STATE, "results read before compute");
  tess_ctx* c = d->ctx;
  (void)hipSetDevice(c->device);
  const size_t n = d->n;
  hipError_t e = hipSuccess;
  if (volume && e == hipSuccess) e = download(c, volume, d->volume.p, 8 * n);
  if (fc && e == hipSuccess) e = download(c, fc, d->fcount.p, 4 * n);
  if (fo && e == hipSuccess) e = download(c, fo, d->foff.p, 8 * (n + 1));
  if (nbr && e == hipSuccess && d->total_faces) e = download(c, nbr, d->nbr.p, 4 * d->total_faces);
  if (area && e == hipSuccess && d->total_faces) e = download(c, area, d->area.p, 8 * d->total_faces);
  return c->hip(e, "download");
}

tess_status tess_diagram_device_cells(const tess_diagram* d, tess_cells_device* out) {
  if (!d || !out) return TESS_E_INVALID;
  if (!d->computed) return TESS_E_STATE;
  out->volume = d->volume.p;
  out->face_count = d->fcount.p;
  out->face_offset = d->foff.p;
  out->neighbor = d->nbr.p;
  out->face_area = d->area.p;
  out->total_faces = d->total_faces;
  out->n = d->n;
  return TESS_OK;
}

tess_status tess_diagram_copy_stats(const tess_diagram* d, tess_cell_stats* stats) {
  if (!d || !stats) return TESS_E_INVALID;
  if (!d->computed || !d->opts.collect_stats)
    return d->ctx->fail(TESS_E_STATE, "stats need collect_stats = 1 and a computed diagram");
  (void)hipSetDevice(d->ctx->device);
  return d->ctx->hip(download(d->ctx, stats, d->stats.p, sizeof(tess_cell_stats) * d->n), "download");
}

tess_status tess_diagram_timings(const tess_diagram* d, tess_timings* out) {
  if (!d || !out) return TESS_E_INVALID;
  *out = d->tim;
  return TESS_OK;
}

tess_status tess_diagram_grid(const tess_diagram* d, double b[6], double sz[3], int64_t* cpd) {
  if (!d) return TESS_E_INVALID;
  if (!d->binned) return d->ctx->fail(TESS_E_STATE, "grid read before compute");
  const GridDev& G = d->hgrid;
  if (b) {
    b[0] = G.xmin;
    b[1] = G.xmax;
    b[2] = G.ymin;
    b[3] = G.ymax;
    b[4] = G.zmin;
    b[5] = G.zmax;
  }
  if (sz) {
    sz[0] = G.sx;
    sz[1] = G.sy;
    sz[2] = G.sz;
  }
  if (cpd) *cpd = d->cpd;
  return TESS_OK;
}

tess_status tess_diagram_copy_grid(const tess_diagram* d, uint32_t* cell_of_point, uint32_t* sorted,
                                   uint64_t* delim) {
  if (!d) return TESS_E_INVALID;
  if (!d->binned) return d->ctx->fail(TESS_E_STATE, "grid read before compute");
  (void)hipSetDevice(d->ctx->device);
  hipError_t e = hipSuccess;
  if (cell_of_point) e = hipMemcpy(cell_of_point, d->bins.p, 4 * d->n, hipMemcpyDeviceToHost);
  if (sorted && e == hipSuccess) e = hipMemcpy(sorted, d->sorted.p, 4 * d->n, hipMemcpyDeviceToHost);
  if (delim && e == hipSuccess) {
    std::vector<uint32_t> tmp((size_t)d->nbins + 1);
    e = hipMemcpy(tmp.data(), d->delim.p, 4 * tmp.size(), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < tmp.size(); ++i) delim[i] = tmp[i];
  }
  return d->ctx->hip(e, "download");
}

tess_status tess_diagram_copy_index_maps(tess_diagram* d, uint32_t* original_of_internal,
                                         uint32_t* internal_of_original) {
  if (!d) return TESS_E_INVALID;
  tess_ctx* c = d->ctx;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  tess_status st;
  if (!d->binned && (st = bin(d)) != TESS_OK) return st;
  const int n = (int)d->n;
  DevBuf<uint32_t> codes, iota, keys, ooi, ioo;
  DevBuf<uint8_t> tmp;
  size_t tmp_bytes = 0;
  if (codes.alloc(n) != hipSuccess || iota.alloc(n) != hipSuccess || keys.alloc(n) != hipSuccess ||
      ooi.alloc(n) != hipSuccess || ioo.alloc(n) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  if ((st = c->hip(launch_index_maps(codes.p, n, iota.p, keys.p, ooi.p, ioo.p, nullptr, &tmp_bytes, s), "sort")) !=
      TESS_OK)
    return st;
  if (tmp.alloc(tmp_bytes) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
  if ((st = c->hip(launch_morton(d->x.p, d->y.p, d->z.p, n, d->grid.p, codes.p, s), "morton")) != TESS_OK ||
      (st = c->hip(launch_index_maps(codes.p, n, iota.p, keys.p, ooi.p, ioo.p, tmp.p, &tmp_bytes, s), "sort")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "index maps")) != TESS_OK)
    return st;
  hipError_t e = hipSuccess;
  if (original_of_internal) e = hipMemcpy(original_of_internal, ooi.p, 4ull * n, hipMemcpyDeviceToHost);
  if (internal_of_original && e == hipSuccess)
    e = hipMemcpy(internal_of_original, ioo.p, 4ull * n, hipMemcpyDeviceToHost);
  return c->hip(e, "download");
}

}  // extern "C"

namespace {

// One cell with its geometry through the single-cell export of the cell
// kernel: the point `index` of the diagram, or the outside site `site`
// (host, 3 doubles; index ignored).  Capacity 128 first, then 256.
tess_status single_cell(tess_diagram* d, size_t index, const double* site, double search_radius,
                        int64_t target_group, tess_cell** out) {
  if (!d || !out) return TESS_E_INVALID;
  *out = nullptr;
  tess_ctx* c = d->ctx;
  if (!site && index >= d->n) return c->fail(TESS_E_INVALID, "index out of range");
  if (site)
    for (int k = 0; k < 3; ++k)
      if (!(site[k] >= d->lo[k] && site[k] <= d->hi[k]))
        return c->fail(TESS_E_INVALID, "the site is outside the container box");
  if (target_group >= 0 && !d->have_groups) return c->fail(TESS_E_INVALID, "diagram has no groups");
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  tess_status st;
  if (!d->binned) {
    if ((st = bin(d)) != TESS_OK) return st;
    if ((st = c->hip(hipStreamSynchronize(s), "binning")) != TESS_OK) return st;
  }
  if (!site && !d->inv_ready) {
    if (d->inv.alloc(d->n) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
    if ((st = c->hip(launch_invert(d->sorted.p, (int)d->n, d->inv.p, s), "invert")) != TESS_OK) return st;
    d->inv_ready = true;
  }
  if (d->exp_v.alloc(3 * 256) != hipSuccess || d->exp_meta.alloc(2 + 256) != hipSuccess ||
      d->exp_ring.alloc(kFaceStage * 64) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  if (site && d->one_site.alloc(3) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
  const GridDev& G = d->hgrid;
  // the cell's outputs go to buffers of its own (the diagram's results stay
  // as they are); the kernel indexes them by the point's original index, so
  // the point's cell gets bases shifted back by that index
  if (d->one_vol.alloc(1) != hipSuccess || d->one_u32.alloc(2) != hipSuccess ||
      d->one_nbr.alloc(kFaceStage) != hipSuccess || d->one_area.alloc(kFaceStage) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  const size_t shift = site ? 0 : index;
  auto shifted = [shift](auto* p, size_t per) {
    return reinterpret_cast<decltype(p)>(reinterpret_cast<uintptr_t>(p) - shift * per);
  };
  double* vol = shifted(d->one_vol.p, sizeof(double));
  uint32_t* fcount = shifted(d->one_u32.p, sizeof(uint32_t));
  uint32_t* status_p = shifted(d->one_u32.p + 1, sizeof(uint32_t));
  int32_t* snbr = shifted(d->one_nbr.p, kFaceStage * sizeof(int32_t));
  double* sarea = shifted(d->one_area.p, kFaceStage * sizeof(double));
  const uint32_t oi = (uint32_t)shift;
  if (site && (st = c->hip(hipMemcpyAsync(d->one_site.p, site, 24, hipMemcpyHostToDevice, s), "upload")) != TESS_OK)
    return st;
  std::unique_ptr<tess_cell> cell(new (std::nothrow) tess_cell);
  if (!cell) return c->fail(TESS_E_NOMEM, "host allocation failed");
  uint32_t status = 0;
  // capacity 128, then 256; a security ball past the shell table takes a
  // table twice as long, up to the full order
  int L = kTableL, vpl = 2;
  for (;;) {
    Table* t = c->table(d->cpd, G.sx, G.sy, G.sz, L, &st);
    if (!t) return st;
    CellArgs a = base_args(d, t);
    a.fpos = nullptr;
    a.arena_next = nullptr;
    a.summary = nullptr;
    a.work = site ? nullptr : d->work1.p;
    a.ext = site ? d->one_site.p : nullptr;
    a.nwork = 1;
    a.xcd_chunk = 0;
    a.target_group = target_group;
    a.radius_limit = search_radius;
    a.volume = vol;
    a.fcount = fcount;
    a.status = status_p;
    a.snbr = snbr;
    a.sarea = sarea;
    a.stats = nullptr;
    a.exp_v = d->exp_v.p;
    a.exp_ring = d->exp_ring.p;
    a.exp_meta = d->exp_meta.p;
    a.redo = nullptr;
    a.nredo = nullptr;
    if ((!site && (st = c->hip(hipMemcpyAsync(d->work1.p, d->inv.p + oi, 4, hipMemcpyDeviceToDevice, s), "copy")) !=
                      TESS_OK) ||
        (st = c->hip(launch_cells(vpl, a, s), "cell kernel")) != TESS_OK ||
        (st = c->hip(hipMemcpyAsync(&status, d->one_u32.p + 1, 4, hipMemcpyDeviceToHost, s), "copy")) != TESS_OK ||
        (st = c->hip(hipStreamSynchronize(s), "cell kernel")) != TESS_OK)
      return st;
    if ((status & TESS_CELL_CAPACITY) && vpl < 4)
      vpl = 4;
    else if ((status & TESS_CELL_TABLE) && L < d->cpd - 1)
      L = (int)std::min<int64_t>(2 * (int64_t)L, d->cpd - 1);
    else
      break;
  }
  if (status & (TESS_CELL_CAPACITY | TESS_CELL_DEGENERATE | TESS_CELL_TABLE))
    return c->fail(status & TESS_CELL_DEGENERATE ? TESS_E_DEGENERATE : TESS_E_CAPACITY,
                   (site ? std::string("the site's cell") : "cell " + std::to_string(index)) + " failed (status " +
                       std::to_string(status) + ")");
  uint32_t fc = 0;
  std::vector<int32_t> nb(kFaceStage), ring(kFaceStage * 64), meta(2 + 256);
  std::vector<double> ar(kFaceStage), vv(3 * 256);
  hipError_t e = hipMemcpy(&cell->volume, d->one_vol.p, 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&fc, d->one_u32.p, 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(nb.data(), d->one_nbr.p, 4 * kFaceStage, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ar.data(), d->one_area.p, 8 * kFaceStage, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ring.data(), d->exp_ring.p, 4 * ring.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(meta.data(), d->exp_meta.p, 4 * meta.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(vv.data(), d->exp_v.p, 8 * vv.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return c->hip(e, "download");
  cell->original = site ? -1 : (int64_t)index;  // Cell::original_index: None for a site
  const int vc = meta[0];
  for (uint32_t f = 0; f < fc; ++f) {
    cell->nbr.push_back(nb[f]);
    cell->area.push_back(ar[f]);
    std::vector<double> r;
    for (int k = 0; k < 64 && ring[f * 64 + k] >= 0; ++k) {
      const int v = ring[f * 64 + k];
      r.push_back(vv[3 * v]);
      r.push_back(vv[3 * v + 1]);
      r.push_back(vv[3 * v + 2]);
    }
    cell->ring.push_back(std::move(r));
  }
  for (int v = 0; v < vc && v < 256; ++v)
    if (meta[2 + v]) {
      cell->verts.push_back(vv[3 * v]);
      cell->verts.push_back(vv[3 * v + 1]);
      cell->verts.push_back(vv[3 * v + 2]);
    }
  *out = cell.release();
  return TESS_OK;
}

}  // namespace

extern "C" {

tess_status tess_cell_at_index(tess_diagram* d, size_t index, double search_radius, int64_t target_group,
                               tess_cell** out) {
  return single_cell(d, index, nullptr, search_radius, target_group, out);
}

tess_status tess_cell_at_particle(tess_diagram* d, const double xyz[3], double search_radius, int64_t target_group,
                                  tess_cell** out) {
  if (!xyz) return d ? d->ctx->fail(TESS_E_INVALID, "null site") : TESS_E_INVALID;
  return single_cell(d, 0, xyz, search_radius, target_group, out);
}

void tess_cell_destroy(tess_cell* c) { delete c; }
int64_t tess_cell_original_index(const tess_cell* c) { return c ? c->original : -1; }
double tess_cell_volume(const tess_cell* c) { return c ? c->volume : std::nan(""); }
size_t tess_cell_neighbors(const tess_cell* c, int64_t* out, size_t cap) {
  if (!c) return 0;
  for (size_t i = 0; i < c->nbr.size() && i < cap && out; ++i) out[i] = c->nbr[i];
  return c->nbr.size();
}
size_t tess_cell_face_count(const tess_cell* c) { return c ? c->nbr.size() : 0; }
tess_status tess_cell_face(const tess_cell* c, size_t f, double* area, int64_t* neighbor) {
  if (!c || f >= c->nbr.size()) return TESS_E_INVALID;
  if (area) *area = c->area[f];
  if (neighbor) *neighbor = c->nbr[f];
  return TESS_OK;
}
size_t tess_cell_face_vertices(const tess_cell* c, size_t f, double* xyz, size_t cap) {
  if (!c || f >= c->ring.size()) return 0;
  const size_t k = c->ring[f].size() / 3;
  for (size_t i = 0; i < 3 * k && i < 3 * cap && xyz; ++i) xyz[i] = c->ring[f][i];
  return k;
}
size_t tess_cell_vertices(const tess_cell* c, double* xyz, size_t cap) {
  if (!c) return 0;
  const size_t k = c->verts.size() / 3;
  for (size_t i = 0; i < 3 * k && i < 3 * cap && xyz; ++i) xyz[i] = c->verts[i];
  return k;
}

tess_status tess_diag_arith(tess_ctx* ctx, const double* a, const double* b, size_t n, double* sq, double* rc,
                            double* dv) {
  if (!ctx || !a || !b || !sq || !rc || !dv || n == 0) return TESS_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (ctx->diag.alloc(5 * n) != hipSuccess) return ctx->fail(TESS_E_NOMEM, "device allocation failed");
  double* p = ctx->diag.p;
  hipStream_t s = ctx->stream;
  tess_status st;
  if ((st = ctx->hip(hipMemcpyAsync(p, a, 8 * n, hipMemcpyHostToDevice, s), "upload")) != TESS_OK) return st;
  if ((st = ctx->hip(hipMemcpyAsync(p + n, b, 8 * n, hipMemcpyHostToDevice, s), "upload")) != TESS_OK) return st;
  if ((st = ctx->hip(launch_diag_arith(p, p + n, (int)n, p + 2 * n, p + 3 * n, p + 4 * n, s), "arith")) != TESS_OK)
    return st;
  if ((st = ctx->hip(hipMemcpyAsync(sq, p + 2 * n, 8 * n, hipMemcpyDeviceToHost, s), "download")) != TESS_OK)
    return st;
  if ((st = ctx->hip(hipMemcpyAsync(rc, p + 3 * n, 8 * n, hipMemcpyDeviceToHost, s), "download")) != TESS_OK)
    return st;
  if ((st = ctx->hip(hipMemcpyAsync(dv, p + 4 * n, 8 * n, hipMemcpyDeviceToHost, s), "download")) != TESS_OK)
    return st;
  return ctx->hip(hipStreamSynchronize(s), "arith");
}

}  // extern "C"
