// batch.cpp -- many independent sets tessellated together.
//
// Every set keeps its own container box, grid and shell table, so its cells
// are bitwise those of a tess_diagram of the same points.  All sets live in
// combined device arrays; each set is a tess_diagram "view" into its slices,
// so the single-set binning and argument code is reused as is.
//
//   graph 1 (captured once): binning of every set + a check that every point
//            was binned
//   host:    grid download, shell tables (built in parallel, cached by grid
//            shape), per-set kernel arguments
//   plan:    item list in set order, or the learned order (features +
//            inference + cost sort over all sets)
//   graph 2 (captured per schedule shape, then replayed): the first pass over
//            the item list (one wavefront per cell, dispatched in list order),
//            a persistent rerun pass at the largest capacity for the cells the
//            first pass could not hold
//   host:    cells past the shell table or the largest LDS capacity, if
//            any, finished set by set (finish_cells); the face-offset scan,
//            total faces, then CSR compaction of all sets at once
// With config.eager the same passes are launched directly instead.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <thread>
#include <vector>

#include "runtime.h"

using namespace tess;

namespace {

struct HostTable {
  std::vector<double> key;
  std::vector<int32_t> off;
  double kcut = 0;
};

using TableKey = std::tuple<int64_t, uint64_t, uint64_t, uint64_t>;

struct Graph {
  hipGraphExec_t exec = nullptr;
  ~Graph() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
};

}  // namespace

// Host words the device results are read back into (pinned).
struct Mail {
  uint32_t bad, nredo, ntab, nspill;
  uint64_t total, arena;
};

struct tess_batch {
  tess_ctx* ctx = nullptr;
  tess_opts opts;
  tess_batch_config cfg{64, 0, 1, 0, 0};  // capacity 64, interleaved, one rerun cell per grab, list order, graphs
  size_t nsets = 0, N = 0;
  std::vector<size_t> n, pbase;                        // points per set, first point
  std::vector<size_t> bbase, cbase, sbase, partbase;  // delim, bins, scan sums, partial bounds
  // combined device arrays
  DevBuf<double> x, y, z, sx, sy, sz, part, volume, sarea, area;
  DevBuf<GridDev> grid;
  DevBuf<uint32_t> bins, cnt, delim, cursor, sorted, scratch, sums, big, misc, fcount, status, ctrl;
  DevBuf<int32_t> snbr, nbr;
  DevBuf<uint64_t> foff, fsums, items, redo, redo2, base_dev, dend_dev;
  // face arena of all sets, shell-table exhaustions, cells for the spill kernel
  DevBuf<uint64_t> fpos, tredo, spill;
  DevBuf<unsigned long long> arena;
  Mail* mail = nullptr;       // pinned
  GridDev* hgrid_pin = nullptr;  // pinned, nsets
  uint64_t table_reruns = 0, spill_cells = 0;
  ~tess_batch() {
    if (mail) (void)hipHostFree(mail);
    if (hgrid_pin) (void)hipHostFree(hgrid_pin);
  }
  DevBuf<tess_cell_stats> stats;
  DevBuf<Summary> summary;
  DevBuf<CellArgs> args;
  DevBuf<double> tkey;
  DevBuf<int32_t> toff;
  // learned order
  DevBuf<uint16_t> feat;
  DevBuf<float> pred, cost;
  DevBuf<uint32_t> skey, shist, ssums, order;
  std::vector<std::unique_ptr<tess_diagram>> views;
  std::vector<GridDev> hgrid;
  std::map<TableKey, std::shared_ptr<HostTable>> tables;
  std::vector<TableKey> table_of;  // per set, as uploaded
  Graph gbin, gcells;
  std::vector<int64_t> gcells_key;  // config + slot counts the cells graph was captured for
  // every set's first cell and first delimiter over the batch (nsets + 1
  // each) and its cells per dimension, for the binning of all sets at once
  DevBuf<uint32_t> cbase_dev, bbase_dev, pblk_dev;  // + first bounds block of each set
  DevBuf<int32_t> cpd_dev;
  uint64_t total_faces = 0;
  tess_timings tim;
  bool computed = false;
};

namespace {

// ctrl words: [0] rerun count, [1] rerun counter, [2] binning check, [3..8]
// slot bounds of the learned order (and its radius word), [10] shell-table
// exhaustions, [11] cells for the spill kernel, [12] count and [13] counter
// of the second rerun (cells past 128 vertices)
constexpr int kCtrlRedo = 0, kCtrlQueue = 1, kCtrlBad = 2, kCtrlBounds = 3, kCtrlTab = 10, kCtrlSpill = 11,
              kCtrlRedo2 = 12, kCtrlQueue2 = 13, kCtrlWords = 14;

tess_status make_views(tess_batch* b) {
  tess_ctx* c = b->ctx;
  for (size_t s = 0; s < b->nsets; ++s) {
    auto d = std::make_unique<tess_diagram>();
    d->ctx = c;
    d->opts = b->opts;
    d->n = b->n[s];
    d->cpd = icbrt_floor((double)d->n / b->opts.cell_density) + 1;  // celery.rs:161-162
    d->nbins = (int)(d->cpd * d->cpd * d->cpd);
    const size_t p = b->pbase[s], nb = (size_t)d->nbins, nparts = std::min<size_t>(256, (d->n + 255) / 256);
    d->x.view(b->x.p + p, d->n);
    d->y.view(b->y.p + p, d->n);
    d->z.view(b->z.p + p, d->n);
    d->sx.view(b->sx.p + p, d->n);
    d->sy.view(b->sy.p + p, d->n);
    d->sz.view(b->sz.p + p, d->n);
    d->grid.view(b->grid.p + s, 1);
    d->part.view(b->part.p + b->partbase[s], 6 * nparts);
    d->bins.view(b->bins.p + p, d->n);
    d->sorted.view(b->sorted.p + p, d->n);
    d->scratch.view(b->scratch.p + p, d->n);
    d->cnt.view(b->cnt.p + b->cbase[s], nb);
    d->cursor.view(b->cursor.p + b->cbase[s], nb);
    d->big.view(b->big.p + b->cbase[s], nb);
    d->delim.view(b->delim.p + b->bbase[s], nb + 1);
    d->sums.view(b->sums.p + b->sbase[s], (nb + 1023) / 1024 + 1);
    d->misc.view(b->misc.p + 4 * s, 4);
    d->volume.view(b->volume.p + p, d->n);
    d->fcount.view(b->fcount.p + p, d->n);
    d->status.view(b->status.p + p, d->n);
    // every set writes into the batch's face arena
    d->snbr.view(b->snbr.p, arena_capacity(b->N));
    d->sarea.view(b->sarea.p, arena_capacity(b->N));
    d->fpos.view(b->fpos.p + p, d->n);
    d->arena_next = b->arena.p;
    d->arena_cap = arena_capacity(b->N);
    d->summary.view(b->summary.p + s, 1);
    if (b->opts.collect_stats) d->stats.view(b->stats.p + p, d->n);
    for (int k = 0; k < 3; ++k) {
      d->lo[k] = 0;
      d->hi[k] = 0;
    }
    b->views.push_back(std::move(d));
  }
  return TESS_OK;
}

tess_status capture_begin(tess_ctx* c) {
  return c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "graph capture");
}

tess_status capture_end(tess_ctx* c, Graph& g) {
  hipGraph_t graph = nullptr;
  tess_status st = c->hip(hipStreamEndCapture(c->stream, &graph), "graph capture");
  if (st != TESS_OK) return st;
  g.reset();
  st = c->hip(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0), "graph instantiate");
  (void)hipGraphDestroy(graph);
  return st;
}

// The binning of every set in one launch per pass (the sets' grids,
// histograms, delimiters and sorted indices land in their slices of the
// combined arrays, as a set-by-set binning would leave them), then the
// check.
tess_status enqueue_binning(tess_batch* b) {
  tess_ctx* c = b->ctx;
  hipStream_t s = c->stream;
  tess_status st;
  if ((st = c->hip(launch_fill_u32(b->ctrl.p + kCtrlBad, 1, 0u, s), "fill")) != TESS_OK) return st;
  SegBinning a;
  a.nsets = (int)b->nsets;
  a.n = (uint32_t)b->N;
  a.nb = (uint32_t)b->cnt.n;
  a.pbase = b->base_dev.p;
  a.cbase = b->cbase_dev.p;
  a.bbase = b->bbase_dev.p;
  a.cpd = b->cpd_dev.p;
  a.pblk = b->pblk_dev.p;
  a.nparts = (uint32_t)(b->part.n / 6);
  a.part = b->part.p;
  a.x = b->x.p;
  a.y = b->y.p;
  a.z = b->z.p;
  a.grid = b->grid.p;
  a.bins = b->bins.p;
  a.rank = b->scratch.p;
  a.cnt = b->cnt.p;
  a.gscan = b->cursor.p;
  a.sums = b->sums.p;
  a.delim = b->delim.p;
  a.sorted = b->sorted.p;
  a.big = b->big.p;
  a.misc = b->misc.p;
  a.sx = b->sx.p;
  a.sy = b->sy.p;
  a.sz = b->sz.p;
  if ((st = c->hip(launch_seg_binning(a, s), "binning")) != TESS_OK) return st;
  return c->hip(launch_batch_check(b->delim.p, b->dend_dev.p, b->base_dev.p, (int)b->nsets, b->ctrl.p + kCtrlBad, s),
                "binning check");
}

// Binning of every set, then the check; captured once.
tess_status build_bin_graph(tess_batch* b) {
  tess_ctx* c = b->ctx;
  tess_status st = capture_begin(c);
  if (st != TESS_OK) return st;
  tess_status err = enqueue_binning(b);
  st = capture_end(c, b->gbin);
  return err != TESS_OK ? err : st;
}

// Shell tables for every set's grid (host, in parallel, cached by shape),
// uploaded as one combined table, and the per-set kernel arguments.
tess_status prepare_tables(tess_batch* b, double* table_ms) {
  tess_ctx* c = b->ctx;
  auto t0 = std::chrono::steady_clock::now();
  std::vector<TableKey> keys(b->nsets);
  std::vector<TableKey> missing;
  for (size_t s = 0; s < b->nsets; ++s) {
    const GridDev& G = b->hgrid[s];
    keys[s] = std::make_tuple((int64_t)G.cpd, bits(G.sx), bits(G.sy), bits(G.sz));
    if (!b->tables.count(keys[s])) {
      b->tables[keys[s]] = nullptr;
      missing.push_back(keys[s]);
    }
  }
  if (!missing.empty()) {
    std::vector<std::shared_ptr<HostTable>> built(missing.size());
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < hw; ++t)
      pool.emplace_back([&, t]() {
        for (size_t i = t; i < missing.size(); i += hw) {
          auto h = std::make_shared<HostTable>();
          const TableKey& k = missing[i];
          double sxv, syv, szv;
          uint64_t u;
          u = std::get<1>(k);
          std::memcpy(&sxv, &u, 8);
          u = std::get<2>(k);
          std::memcpy(&syv, &u, 8);
          u = std::get<3>(k);
          std::memcpy(&szv, &u, 8);
          build_table_host(std::get<0>(k), sxv, syv, szv, kTableL, h->key, h->off, &h->kcut);
          built[i] = h;
        }
      });
    for (auto& th : pool) th.join();
    for (size_t i = 0; i < missing.size(); ++i) b->tables[missing[i]] = built[i];
  }
  if (keys != b->table_of) {
    // combined upload and per-set arguments
    size_t total = 0;
    std::vector<size_t> tb(b->nsets);
    for (size_t s = 0; s < b->nsets; ++s) {
      tb[s] = total;
      total += b->tables[keys[s]]->key.size();
    }
    std::vector<double> hk(total);
    std::vector<int32_t> ho(total);
    for (size_t s = 0; s < b->nsets; ++s) {
      const HostTable& h = *b->tables[keys[s]];
      std::copy(h.key.begin(), h.key.end(), hk.begin() + tb[s]);
      std::copy(h.off.begin(), h.off.end(), ho.begin() + tb[s]);
    }
    if (b->tkey.alloc(total) != hipSuccess || b->toff.alloc(total) != hipSuccess)
      return c->fail(TESS_E_NOMEM, "device allocation failed");
    std::vector<CellArgs> ha(b->nsets);
    for (size_t s = 0; s < b->nsets; ++s) {
      const HostTable& h = *b->tables[keys[s]];
      Table t;
      t.key.view(b->tkey.p + tb[s], h.key.size());
      t.off.view(b->toff.p + tb[s], h.off.size());
      t.len = (int)h.key.size();
      t.kcut = h.kcut;
      t.L = kTableL;
      ha[s] = base_args(b->views[s].get(), &t);
      ha[s].defer_table = 1;  // the passes list them (BatchArgs::tredo)
    }
    tess_status st;
    if ((st = c->hip(hipMemcpy(b->tkey.p, hk.data(), 8 * total, hipMemcpyHostToDevice), "upload")) != TESS_OK ||
        (st = c->hip(hipMemcpy(b->toff.p, ho.data(), 4 * total, hipMemcpyHostToDevice), "upload")) != TESS_OK ||
        (st = c->hip(hipMemcpy(b->args.p, ha.data(), sizeof(CellArgs) * b->nsets, hipMemcpyHostToDevice),
                     "upload")) != TESS_OK)
      return st;
    b->table_of = keys;
    b->gcells_key.clear();  // arguments moved: recapture the cells graph
  }
  *table_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return TESS_OK;
}

// Learned order: features of every set, inference over all points, per-item
// costs and the stable cost sort.  Returns the slot bounds.
tess_status learned_order(tess_batch* b, const tess_model* m, uint32_t bounds[4]) {
  tess_ctx* c = b->ctx;
  hipStream_t s = c->stream;
  const size_t N = b->N;
  const size_t ntiles = (N + 255) / 256, mk = (size_t)kSchedKeys * ntiles;
  if (b->feat.alloc(N * kMlpIn) != hipSuccess || b->pred.alloc(N * kMlpOut) != hipSuccess ||
      b->cost.alloc(N) != hipSuccess || b->skey.alloc((N + 3) / 4) != hipSuccess || b->shist.alloc(mk) != hipSuccess ||
      b->ssums.alloc((mk + 1023) / 1024 + 1) != hipSuccess || b->order.alloc(N) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  tess_status st;
  for (size_t k = 0; k < b->nsets; ++k) {
    tess_diagram* d = b->views[k].get();
    if ((st = c->hip(launch_features(d->bins.p, d->delim.p, (int)d->cpd, (int)d->n, b->feat.p + b->pbase[k] * kMlpIn,
                                     s),
                     "features")) != TESS_OK)
      return st;
  }
  if ((st = c->hip(launch_mlp_forward(b->feat.p, (int)N, m->wbf.p, m->bias.p, b->pred.p, s), "mlp forward")) !=
          TESS_OK ||
      (st = c->hip(launch_batch_costs(b->items.p, N, b->base_dev.p, b->sorted.p, b->pred.p + 1, kMlpOut, m->norm[3],
                                      m->norm[2], b->cost.p, s),
                   "costs")) != TESS_OK ||
      (st = c->hip(launch_schedule(b->cost.p, 1, 1.f, 0.f, m->calib[0], m->calib[1], nullptr, (int)N,
                                   (uint8_t*)b->skey.p, b->shist.p, b->ssums.p, b->order.p,
                                   b->ctrl.p + kCtrlBounds, s),
                   "schedule")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(bounds, b->ctrl.p + kCtrlBounds, 16, hipMemcpyDeviceToHost, s), "schedule")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "schedule")) != TESS_OK)
    return st;
  if (!(bounds[0] == 0 && bounds[0] <= bounds[1] && bounds[1] <= bounds[2] && bounds[2] <= bounds[3] &&
        bounds[3] == (uint32_t)N))
    return c->fail(TESS_E_DEVICE, "schedule lost cells");
  return TESS_OK;
}

int vpl_of(int capacity) { return capacity >= 256 ? 4 : (capacity >= 128 ? 2 : 1); }

// First pass(es) and the rerun pass.
tess_status enqueue_cells(tess_batch* b, const uint32_t* list_base, const uint32_t bounds[4], bool learned) {
  tess_ctx* c = b->ctx;
  hipStream_t s = c->stream;
  tess_status st;
  if (launch_fill_u32(b->ctrl.p, 2, 0u, s) != hipSuccess || launch_fill_u32(b->ctrl.p + kCtrlTab, 4, 0u, s) != hipSuccess ||
      launch_fill_u32((uint32_t*)b->arena.p, 2 * kArenaWords, 0u, s) != hipSuccess ||
      launch_fill_u32((uint32_t*)b->summary.p, b->nsets * (sizeof(Summary) / 4), 0u, s) != hipSuccess)
    return c->fail(TESS_E_DEVICE, "fill");
  BatchArgs a;
  std::memset(&a, 0, sizeof(a));
  a.sets = b->args.p;
  a.items = b->items.p;
  a.tredo = b->tredo.p;
  a.ntredo = b->ctrl.p + kCtrlTab;
  // capacity overflows of the first passes below 256 vertices go to the
  // rerun list, at 256 to the spill kernel's
  auto overflow_to = [&](BatchArgs& x, int vpl) {
    x.redo = vpl < 4 ? b->redo.p : b->spill.p;
    x.nredo = b->ctrl.p + (vpl < 4 ? kCtrlRedo : kCtrlSpill);
  };
  if (learned) {
    static const int kVpl[3] = {1, 2, 4};
    for (int slot = 2; slot >= 0; --slot) {
      const uint32_t cnt = bounds[slot + 1] - bounds[slot];
      if (!cnt) continue;
      a.order = list_base + bounds[slot];
      overflow_to(a, kVpl[slot]);
      a.xcd_chunk = 0;
      if ((st = c->hip(launch_cells_items(kVpl[slot], a, cnt, s), "cell pass")) != TESS_OK) return st;
    }
  } else {
    const int vpl = vpl_of(b->cfg.first_capacity);
    a.order = nullptr;
    overflow_to(a, vpl);
    a.xcd_chunk = b->cfg.xcd_chunk;
    if ((st = c->hip(launch_cells_items(vpl, a, (uint32_t)b->N, s), "cell pass")) != TESS_OK) return st;
  }
  // reruns, persistent over device-counted lists: at 128 vertices (the
  // cells that outgrew 64 nearly all fit, and this variant keeps more
  // wavefronts resident), then at 256 for the cells past 128
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu <= 0)
    ncu = 256;
  for (int stage = 0; stage < 2; ++stage) {
    const int vpl = stage == 0 ? 2 : 4;
    BatchArgs r;
    std::memset(&r, 0, sizeof(r));
    r.sets = b->args.p;
    r.items = stage == 0 ? b->redo.p : b->redo2.p;
    r.count = b->ctrl.p + (stage == 0 ? kCtrlRedo : kCtrlRedo2);
    r.counter = b->ctrl.p + (stage == 0 ? kCtrlQueue : kCtrlQueue2);
    r.chunk = std::max(1, b->cfg.redo_chunk);
    if (stage == 0) {  // past 128 vertices: to the second rerun
      r.redo = b->redo2.p;
      r.nredo = b->ctrl.p + kCtrlRedo2;
    } else {
      overflow_to(r, 4);
    }
    r.tredo = b->tredo.p;
    r.ntredo = b->ctrl.p + kCtrlTab;
    if ((st = c->hip(launch_cells_queue(vpl, r, ncu * cells_queue_blocks_per_cu(vpl), s), "rerun pass")) != TESS_OK)
      return st;
  }
  return TESS_OK;
}

// Cells of one set that the batched passes could not finish (shell table,
// capacity 256), finished like a single diagram's.
tess_status finish_set(tess_batch* b, size_t set, std::vector<uint32_t>& tab, std::vector<uint32_t>& spl,
                       uint64_t* failed) {
  tess_ctx* c = b->ctx;
  tess_diagram* d = b->views[set].get();
  tess_status st;
  Table* t = c->table(d->cpd, d->hgrid.sx, d->hgrid.sy, d->hgrid.sz, kTableL, &st);
  if (!t) return st;
  CellArgs a = base_args(d, t);
  FinishCounts fc;
  if ((st = finish_cells(d, a, std::move(tab), std::move(spl), &fc)) != TESS_OK) return st;
  b->table_reruns += fc.table;
  b->spill_cells += fc.spill;
  *failed += fc.failed;
  return TESS_OK;
}

tess_status batch_compute(tess_batch* b, const tess_model* m) {
  tess_ctx* c = b->ctx;
  hipStream_t s = c->stream;
  tess_status st;
  hipEvent_t ev[5];
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) return c->fail(TESS_E_DEVICE, "event create failed");
  struct Guard {
    hipEvent_t* e;
    ~Guard() {
      for (int i = 0; i < 5; ++i) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  const bool eager = b->cfg.eager != 0;
  Mail* mail = b->mail;
  if (eager) {
    // the binning passes of every set, launched directly
    if ((st = c->hip(hipEventRecord(ev[0], s), "event")) != TESS_OK || (st = enqueue_binning(b)) != TESS_OK)
      return st;
  } else {
    if (!b->gbin.exec && (st = build_bin_graph(b)) != TESS_OK) return st;
    (void)hipEventRecord(ev[0], s);
    if ((st = c->hip(hipGraphLaunch(b->gbin.exec, s), "binning graph")) != TESS_OK) return st;
  }
  (void)hipEventRecord(ev[1], s);
  // read-backs go to pinned host memory
  if ((st = c->hip(hipMemcpyAsync(b->hgrid_pin, b->grid.p, sizeof(GridDev) * b->nsets, hipMemcpyDeviceToHost, s),
                   "grid copy")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&mail->bad, b->ctrl.p + kCtrlBad, 4, hipMemcpyDeviceToHost, s), "grid copy")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "binning")) != TESS_OK)
    return st;
  if (mail->bad) return c->fail(TESS_E_DEVICE, std::to_string(mail->bad) + " sets lost points in the counting sort");
  for (size_t k = 0; k < b->nsets; ++k) {
    b->hgrid[k] = b->hgrid_pin[k];
    b->views[k]->hgrid = b->hgrid[k];
  }
  double table_ms = 0;
  if ((st = prepare_tables(b, &table_ms)) != TESS_OK) return st;
  const bool learned = b->cfg.learned != 0;
  uint32_t bounds[4] = {0, (uint32_t)b->N, (uint32_t)b->N, (uint32_t)b->N};
  if (learned && (st = learned_order(b, m, bounds)) != TESS_OK) return st;
  (void)hipEventRecord(ev[2], s);
  if (eager) {
    if ((st = enqueue_cells(b, b->order.p, bounds, learned)) != TESS_OK) return st;
  } else {
    std::vector<int64_t> key = {b->cfg.first_capacity, b->cfg.xcd_chunk, b->cfg.redo_chunk, b->cfg.learned,
                                bounds[1], bounds[2]};
    if (!b->gcells.exec || key != b->gcells_key) {
      if ((st = capture_begin(c)) != TESS_OK) return st;
      tess_status err = enqueue_cells(b, b->order.p, bounds, learned);
      st = capture_end(c, b->gcells);
      if (err != TESS_OK) return err;
      if (st != TESS_OK) return st;
      b->gcells_key = key;
    }
    if ((st = c->hip(hipGraphLaunch(b->gcells.exec, s), "cells graph")) != TESS_OK) return st;
  }
  std::vector<Summary> sm(b->nsets);
  if ((st = c->hip(hipMemcpyAsync(&mail->nredo, b->ctrl.p + kCtrlRedo, 4, hipMemcpyDeviceToHost, s), "redo")) !=
          TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&mail->ntab, b->ctrl.p + kCtrlTab, 8, hipMemcpyDeviceToHost, s), "lists")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "cells")) != TESS_OK)
    return st;
  const uint32_t nredo = mail->nredo, ntab = mail->ntab, nspill = mail->nspill;
  // cells past the shell table or capacity 256 (rare): finished set by set
  uint64_t failed = 0;
  b->table_reruns = b->spill_cells = 0;
  if (ntab || nspill) {
    std::vector<uint64_t> t64(ntab), s64(nspill);
    if ((ntab && (st = c->hip(hipMemcpy(t64.data(), b->tredo.p, 8ull * ntab, hipMemcpyDeviceToHost), "download")) !=
                     TESS_OK) ||
        (nspill &&
         (st = c->hip(hipMemcpy(s64.data(), b->spill.p, 8ull * nspill, hipMemcpyDeviceToHost), "download")) != TESS_OK))
      return st;
    std::map<size_t, std::pair<std::vector<uint32_t>, std::vector<uint32_t>>> per;
    for (uint64_t it : t64) per[it >> 32].first.push_back((uint32_t)it);
    for (uint64_t it : s64) per[it >> 32].second.push_back((uint32_t)it);
    for (auto& kv : per)
      if ((st = finish_set(b, kv.first, kv.second.first, kv.second.second, &failed)) != TESS_OK) return st;
  }
  if ((st = c->hip(launch_face_offsets(b->fcount.p, (int)b->N, b->foff.p, b->fsums.p, s), "face offsets")) !=
      TESS_OK)
    return st;
  (void)hipEventRecord(ev[3], s);
  if ((st = c->hip(hipMemcpyAsync(&mail->total, b->foff.p + b->N, 8, hipMemcpyDeviceToHost, s), "total")) !=
          TESS_OK ||
      (st = c->hip(hipMemcpyAsync(&mail->arena, b->arena.p + (size_t)kArenaStride * (kArenaParts + 1), 8, hipMemcpyDeviceToHost, s),
                   "arena")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(sm.data(), b->summary.p, sizeof(Summary) * b->nsets, hipMemcpyDeviceToHost, s),
                   "summary")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "cells")) != TESS_OK)
    return st;
  const uint64_t total = mail->total;
  if (mail->arena) return c->fail(TESS_E_CAPACITY, "the face arena is full");
  if (b->nbr.alloc(total) != hipSuccess || b->area.alloc(total) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  if ((st = c->hip(launch_compact(b->fcount.p, b->foff.p, (int)b->N, b->snbr.p, b->sarea.p, b->nbr.p, b->area.p, s,
                                  b->fpos.p),
                   "compact")) != TESS_OK)
    return st;
  (void)hipEventRecord(ev[4], s);
  if ((st = c->hip(hipStreamSynchronize(s), "compute")) != TESS_OK) return st;
  float ms = 0;
  std::memset(&b->tim, 0, sizeof(b->tim));
  (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
  b->tim.bin_ms = ms;
  (void)hipEventElapsedTime(&ms, ev[1], ev[2]);
  b->tim.plan_ms = ms;
  (void)hipEventElapsedTime(&ms, ev[2], ev[3]);
  b->tim.cells_ms = ms;
  (void)hipEventElapsedTime(&ms, ev[3], ev[4]);
  b->tim.compact_ms = ms;
  (void)hipEventElapsedTime(&ms, ev[0], ev[4]);
  b->tim.total_ms = ms;
  b->tim.table_ms = table_ms;
  b->tim.capacity_reruns = nredo;
  b->tim.fallback_cells = learned ? nredo : 0;
  b->tim.table_reruns = b->table_reruns;
  b->tim.spill_cells = b->spill_cells;
  uint64_t bad_cells = 0, flops = 0, inc = 0;
  for (const Summary& x : sm) {
    bad_cells += x.degenerate;
    flops += x.flops;
    inc += x.incident_cuts;
  }
  b->tim.incident_cuts = inc;
  b->tim.flops = flops;
  b->total_faces = total;
  b->computed = true;
  if (failed) return c->fail(TESS_E_CAPACITY, std::to_string(failed) + " cells exceeded the spill workspace limit");
  if (bad_cells) return c->fail(TESS_E_DEGENERATE, std::to_string(bad_cells) + " cells failed topology checks");
  return TESS_OK;
}

bool valid_config(const tess_batch_config& k) {
  return (k.first_capacity == 64 || k.first_capacity == 128 || k.first_capacity == 256) &&
         k.xcd_chunk >= 0 && k.xcd_chunk <= (1 << 24) && k.redo_chunk >= 1 && k.redo_chunk <= 64 &&
         (k.learned == 0 || k.learned == 1) && (k.eager == 0 || k.eager == 1);
}

std::string signature(const tess_batch* b, bool learned) {
  hipDeviceProp_t p;
  std::string arch = "unknown";
  int ncu = 0;
  if (hipGetDeviceProperties(&p, b->ctx->device) == hipSuccess) {
    arch = p.gcnArchName;
    ncu = p.multiProcessorCount;
  }
  for (char& ch : arch)
    if (ch == ' ') ch = '_';
  const int lg = b->N ? (int)std::lround(std::log2((double)b->N)) : 0;
  std::ostringstream os;
  os << arch << ":cu" << ncu << ":sets" << b->nsets << ":log2n" << lg << ":density" << b->opts.cell_density
     << ":learned" << (learned ? 1 : 0);
  return os.str();
}

}  // namespace

extern "C" {

tess_status tess_batch_create(tess_ctx* ctx, size_t nsets, const double* const* xyz, const size_t* n,
                              const double* lo, const double* hi, const tess_opts* opts, tess_batch** out) {
  if (!ctx || !out) return TESS_E_INVALID;
  *out = nullptr;
  if (nsets == 0 || !xyz || !n || !lo || !hi) return ctx->fail(TESS_E_INVALID, "empty batch or null arrays");
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  auto b = std::make_unique<tess_batch>();
  b->ctx = ctx;
  if (opts)
    b->opts = *opts;
  else
    tess_opts_default(&b->opts);
  if (!(b->opts.cell_density > 0)) b->opts.cell_density = 1.25;
  if (b->opts.point_order != TESS_ORDER_INPUT) return ctx->fail(TESS_E_INVALID, "batches use the input point order");
  b->nsets = nsets;
  b->n.assign(n, n + nsets);
  size_t N = 0, B = 0, C = 0, S = 0, P = 0;
  for (size_t s = 0; s < nsets; ++s) {
    if (n[s] == 0 || !xyz[s]) return ctx->fail(TESS_E_INVALID, "empty point set in batch");
    if (n[s] > (size_t)INT32_MAX / kFaceStage) return ctx->fail(TESS_E_INVALID, "too many points in a set");
    for (int k = 0; k < 3; ++k)
      if (!(hi[3 * s + k] > lo[3 * s + k])) return ctx->fail(TESS_E_INVALID, "container box is empty");
    const int64_t cpd = icbrt_floor((double)n[s] / b->opts.cell_density) + 1;
    if (cpd > 1000) return ctx->fail(TESS_E_INVALID, "grid too fine (cell_density too small)");
    const size_t nb = (size_t)(cpd * cpd * cpd);
    b->pbase.push_back(N);
    b->bbase.push_back(B);
    b->cbase.push_back(C);
    b->sbase.push_back(S);
    b->partbase.push_back(P);
    N += n[s];
    B += nb + 1;
    C += nb;
    S += (nb + 1023) / 1024 + 1;
    P += 6 * std::min<size_t>(256, (n[s] + 255) / 256);
  }
  if (N > (size_t)UINT32_MAX || B > (size_t)UINT32_MAX) return ctx->fail(TESS_E_INVALID, "batch too large");
  b->N = N;
  bool ok = b->x.alloc(N) == hipSuccess && b->y.alloc(N) == hipSuccess && b->z.alloc(N) == hipSuccess &&
            b->sx.alloc(N) == hipSuccess && b->sy.alloc(N) == hipSuccess && b->sz.alloc(N) == hipSuccess &&
            b->part.alloc(P) == hipSuccess && b->grid.alloc(nsets) == hipSuccess && b->bins.alloc(N) == hipSuccess &&
            b->sorted.alloc(N) == hipSuccess && b->scratch.alloc(N) == hipSuccess && b->cnt.alloc(C) == hipSuccess &&
            b->cursor.alloc(C) == hipSuccess && b->big.alloc(C) == hipSuccess && b->delim.alloc(B) == hipSuccess &&
            b->sums.alloc(S) == hipSuccess && b->misc.alloc(4 * nsets) == hipSuccess &&
            b->volume.alloc(N) == hipSuccess && b->fcount.alloc(N) == hipSuccess && b->status.alloc(N) == hipSuccess &&
            b->snbr.alloc(arena_capacity(N)) == hipSuccess && b->sarea.alloc(arena_capacity(N)) == hipSuccess &&
            b->foff.alloc(N + 1) == hipSuccess && b->fsums.alloc((N + 1023) / 1024 + 1) == hipSuccess &&
            b->items.alloc(N) == hipSuccess && b->redo.alloc(N) == hipSuccess && b->redo2.alloc(N) == hipSuccess &&
            b->base_dev.alloc(nsets + 1) == hipSuccess && b->dend_dev.alloc(nsets) == hipSuccess &&
            b->cbase_dev.alloc(nsets + 1) == hipSuccess && b->bbase_dev.alloc(nsets + 1) == hipSuccess &&
            b->cpd_dev.alloc(nsets) == hipSuccess && b->pblk_dev.alloc(nsets + 1) == hipSuccess &&
            b->summary.alloc(nsets) == hipSuccess && b->args.alloc(nsets) == hipSuccess &&
            b->ctrl.alloc(kCtrlWords) == hipSuccess && b->fpos.alloc(N) == hipSuccess &&
            b->tredo.alloc(N) == hipSuccess && b->spill.alloc(N) == hipSuccess && b->arena.alloc(kArenaWords) == hipSuccess &&
            hipHostMalloc((void**)&b->mail, sizeof(Mail), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc((void**)&b->hgrid_pin, sizeof(GridDev) * nsets, hipHostMallocDefault) == hipSuccess;
  if (ok && b->opts.collect_stats) ok = b->stats.alloc(N) == hipSuccess;
  if (!ok) return ctx->fail(TESS_E_NOMEM, "device allocation failed");
  // points: one AoS upload, one SoA pass
  tess_status st;
  {
    DevBuf<double> aos;
    if (aos.alloc(3 * N) != hipSuccess) return ctx->fail(TESS_E_NOMEM, "device allocation failed");
    for (size_t s = 0; s < nsets; ++s)
      if ((st = ctx->hip(hipMemcpy(aos.p + 3 * b->pbase[s], xyz[s], 24 * n[s], hipMemcpyHostToDevice), "upload")) !=
          TESS_OK)
        return st;
    if ((st = ctx->hip(launch_soa(aos.p, (int)N, b->x.p, b->y.p, b->z.p, ctx->stream), "soa")) != TESS_OK ||
        (st = ctx->hip(hipStreamSynchronize(ctx->stream), "soa")) != TESS_OK)
      return st;
  }
  // base: first point of each set (and N); dend: index of each set's last
  // delimiter, which must equal the set's size after binning
  std::vector<uint64_t> base(nsets + 1), dend(nsets);
  std::vector<uint32_t> cb(nsets + 1), bb(nsets + 1), pb(nsets + 1);
  std::vector<int32_t> cpds(nsets);
  for (size_t s = 0; s < nsets; ++s) {
    base[s] = b->pbase[s];
    dend[s] = (s + 1 < nsets ? b->bbase[s + 1] : B) - 1;
    cb[s] = (uint32_t)b->cbase[s];
    bb[s] = (uint32_t)b->bbase[s];
    pb[s] = (uint32_t)(b->partbase[s] / 6);
    cpds[s] = (int32_t)(icbrt_floor((double)n[s] / b->opts.cell_density) + 1);
  }
  base[nsets] = N;
  cb[nsets] = (uint32_t)C;
  bb[nsets] = (uint32_t)B;
  pb[nsets] = (uint32_t)(P / 6);
  if ((st = ctx->hip(hipMemcpy(b->cbase_dev.p, cb.data(), 4 * cb.size(), hipMemcpyHostToDevice), "upload")) !=
          TESS_OK ||
      (st = ctx->hip(hipMemcpy(b->bbase_dev.p, bb.data(), 4 * bb.size(), hipMemcpyHostToDevice), "upload")) !=
          TESS_OK ||
      (st = ctx->hip(hipMemcpy(b->cpd_dev.p, cpds.data(), 4 * cpds.size(), hipMemcpyHostToDevice), "upload")) !=
          TESS_OK ||
      (st = ctx->hip(hipMemcpy(b->pblk_dev.p, pb.data(), 4 * pb.size(), hipMemcpyHostToDevice), "upload")) != TESS_OK)
    return st;
  if ((st = ctx->hip(hipMemcpy(b->base_dev.p, base.data(), 8 * base.size(), hipMemcpyHostToDevice), "upload")) !=
          TESS_OK ||
      (st = ctx->hip(hipMemcpy(b->dend_dev.p, dend.data(), 8 * dend.size(), hipMemcpyHostToDevice), "upload")) !=
          TESS_OK ||
      (st = ctx->hip(launch_batch_items(b->base_dev.p, (int)nsets, N, b->items.p, ctx->stream), "items")) !=
          TESS_OK ||
      (st = ctx->hip(hipStreamSynchronize(ctx->stream), "items")) != TESS_OK)
    return st;
  if ((st = make_views(b.get())) != TESS_OK) return st;
  for (size_t s = 0; s < nsets; ++s)
    for (int k = 0; k < 3; ++k) {
      b->views[s]->lo[k] = lo[3 * s + k];
      b->views[s]->hi[k] = hi[3 * s + k];
    }
  b->hgrid.resize(nsets);
  std::memset(&b->tim, 0, sizeof(b->tim));
  *out = b.release();
  return TESS_OK;
}

void tess_batch_destroy(tess_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  delete b;
}

size_t tess_batch_sets(const tess_batch* b) { return b ? b->nsets : 0; }

tess_status tess_batch_set_config(tess_batch* b, const tess_batch_config* cfg) {
  if (!b || !cfg) return TESS_E_INVALID;
  if (!valid_config(*cfg)) return b->ctx->fail(TESS_E_INVALID, "bad batch configuration");
  b->cfg = *cfg;
  return TESS_OK;
}

tess_status tess_batch_get_config(const tess_batch* b, tess_batch_config* cfg) {
  if (!b || !cfg) return TESS_E_INVALID;
  *cfg = b->cfg;
  return TESS_OK;
}

tess_status tess_batch_compute(tess_batch* b, const tess_model* m) {
  if (!b) return TESS_E_INVALID;
  tess_ctx* c = b->ctx;
  if (b->cfg.learned && !(m && m->wbf.p)) return c->fail(TESS_E_INVALID, "learned order needs a model");
  if (m && m->ctx != c) return c->fail(TESS_E_INVALID, "model and batch belong to different contexts");
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  return batch_compute(b, m);
}

tess_status tess_batch_tune(tess_batch* b, const tess_model* m, const char* cache_path, tess_batch_config* best) {
  if (!b) return TESS_E_INVALID;
  tess_ctx* c = b->ctx;
  const bool can_learn = m && m->wbf.p;
  const std::string sig = signature(b, can_learn);
  // cache lookup: "signature first_capacity xcd_chunk redo_chunk learned ms"
  if (cache_path) {
    std::ifstream in(cache_path);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream is(line);
      std::string s;
      tess_batch_config k;
      k.eager = 0;
      double ms;
      if (is >> s >> k.first_capacity >> k.xcd_chunk >> k.redo_chunk >> k.learned >> ms && s == sig &&
          valid_config(k) && (!k.learned || can_learn)) {
        b->cfg = k;
        if (best) *best = k;
        return TESS_OK;
      }
    }
  }
  std::vector<tess_batch_config> cand;
  for (int cap : {64, 128})
    for (int xcd : {0, 256, 4096}) cand.push_back({cap, xcd, 1, 0, 0});
  if (can_learn)
    for (int xcd : {0, 256}) cand.push_back({64, xcd, 1, 1, 0});
  double best_ms = 1e300;
  tess_batch_config bk = b->cfg;
  for (const tess_batch_config& k : cand) {
    b->cfg = k;
    double t = 1e300;
    for (int rep = 0; rep < 2; ++rep) {  // the first run may capture graphs
      std::lock_guard<std::mutex> g(c->mu);
      (void)hipSetDevice(c->device);
      tess_status st = batch_compute(b, m);
      if (st != TESS_OK) return st;
      t = std::min(t, b->tim.total_ms);
    }
    if (t < best_ms) {
      best_ms = t;
      bk = k;
    }
  }
  b->cfg = bk;
  if (best) *best = bk;
  if (cache_path) {
    std::ofstream out(cache_path, std::ios::app);
    out << sig << ' ' << bk.first_capacity << ' ' << bk.xcd_chunk << ' ' << bk.redo_chunk << ' ' << bk.learned
        << ' ' << best_ms << '\n';
  }
  return TESS_OK;
}

tess_status tess_batch_total_faces(const tess_batch* b, size_t set, uint64_t* out) {
  if (!b || !out) return TESS_E_INVALID;
  if (set >= b->nsets) return b->ctx->fail(TESS_E_INVALID, "set index out of range");
  if (!b->computed) return b->ctx->fail(TESS_E_STATE, "results read before compute");
  uint64_t f[2];
  (void)hipSetDevice(b->ctx->device);
  tess_status st = b->ctx->hip(hipMemcpy(&f[0], b->foff.p + b->pbase[set], 8, hipMemcpyDeviceToHost), "download");
  if (st == TESS_OK)
    st = b->ctx->hip(hipMemcpy(&f[1], b->foff.p + b->pbase[set] + b->n[set], 8, hipMemcpyDeviceToHost), "download");
  if (st == TESS_OK) *out = f[1] - f[0];
  return st;
}

tess_status tess_batch_copy_cells(const tess_batch* b, size_t set, double* volume, uint32_t* fc, uint64_t* fo,
                                  int32_t* nbr, double* area) {
  if (!b) return TESS_E_INVALID;
  tess_ctx* c = b->ctx;
  if (set >= b->nsets) return c->fail(TESS_E_INVALID, "set index out of range");
  if (!b->computed) return c->fail(TESS_E_STATE, "results read before compute");
  (void)hipSetDevice(c->device);
  const size_t p = b->pbase[set], n = b->n[set];
  std::vector<uint64_t> off(n + 1);
  hipError_t e = hipMemcpy(off.data(), b->foff.p + p, 8 * (n + 1), hipMemcpyDeviceToHost);
  const uint64_t f0 = off[0], nf = off[n] - off[0];
  if (e == hipSuccess && volume) e = download(c, volume, b->volume.p + p, 8 * n);
  if (e == hipSuccess && fc) e = download(c, fc, b->fcount.p + p, 4 * n);
  if (e == hipSuccess && fo)
    for (size_t i = 0; i <= n; ++i) fo[i] = off[i] - f0;
  if (e == hipSuccess && nbr && nf) e = download(c, nbr, b->nbr.p + f0, 4 * nf);
  if (e == hipSuccess && area && nf) e = download(c, area, b->area.p + f0, 8 * nf);
  return c->hip(e, "download");
}

tess_status tess_batch_timings(const tess_batch* b, tess_timings* out) {
  if (!b || !out) return TESS_E_INVALID;
  *out = b->tim;
  return TESS_OK;
}

}  // extern "C"
