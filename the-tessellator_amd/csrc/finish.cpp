// finish.cpp -- the cells an LDS cell pass could not finish.
//
// The reference streams every offset of the grid, (2 cpd - 1)^3 of them
// (celery.rs:418-679, :971-1018), and its Pool grows without bound
// (pool.rs:85-110).  The LDS passes hold a prefix of the shell order
// (|o|_inf <= 12) and at most 256 vertices / 128 faces per cell; the cells
// past either limit are listed by the passes and finished here:
//   * a security ball that reaches past the table: the cell is recomputed
//     at capacity 256 with a table twice as long (up to the full order,
//     which has no end), as often as needed;
//   * a cell above 256 vertices or 128 faces: the spill kernel, whose face
//     rings live in a global-memory workspace that grows until they fit.
// Both recompute the whole cell in the canonical order, so the outputs are
// those of an unbounded stream and storage.
#include <algorithm>
#include <string>

#include "launch.h"
#include "runtime.h"

using namespace tess;

namespace tess {

namespace {

// First spill workspace: 4096 ring entries (two buffers) and 512 faces per
// block, about 240 KB; grown 4x per retry up to 2^22 entries.
constexpr int kSpillRmax0 = 4096, kSpillFmax0 = 512;
constexpr int kSpillRmaxMax = 1 << 22;
// Blocks of one spill launch (each with its own workspace) and the bytes
// their workspaces may take together.
constexpr int kSpillBlocks = 256;
constexpr size_t kSpillBudget = size_t(1) << 30;

tess_status upload_list(tess_ctx* c, DevBuf<uint32_t>& buf, const std::vector<uint32_t>& v) {
  if (buf.alloc(v.size()) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
  return c->hip(hipMemcpyAsync(buf.p, v.data(), 4 * v.size(), hipMemcpyHostToDevice, c->stream), "upload");
}

}  // namespace

tess_status read_list(tess_ctx* c, const uint32_t* dev, std::vector<uint32_t>& v) {
  uint32_t n = 0;
  tess_status st;
  if ((st = c->hip(hipMemcpyAsync(&n, dev, 4, hipMemcpyDeviceToHost, c->stream), "download")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(c->stream), "download")) != TESS_OK)
    return st;
  v.resize(n);
  if (!n) return TESS_OK;
  if ((st = c->hip(hipMemcpyAsync(v.data(), dev + 1, 4ull * n, hipMemcpyDeviceToHost, c->stream), "download")) !=
      TESS_OK)
    return st;
  return c->hip(hipStreamSynchronize(c->stream), "download");
}

tess_status finish_cells(tess_diagram* d, const CellArgs& a0, std::vector<uint32_t> tab, std::vector<uint32_t> spl,
                         FinishCounts* out, GeoStage* geo) {
  tess_ctx* c = d->ctx;
  hipStream_t s = c->stream;
  tess_status st;
  FinishCounts fc;
  const GridDev& G = d->hgrid;
  const int64_t full = d->cpd - 1;
  int L = kTableL;
  int rmax = kSpillRmax0, fmax = kSpillFmax0;
  // status of an item's cell: the original index (grid position -> sorted)
  // or the site itself
  const uint32_t* sorted = a0.ext ? nullptr : a0.sorted;
  while (!tab.empty() || !spl.empty()) {
    if (!tab.empty()) {
      if (L >= full) return c->fail(TESS_E_DEVICE, "a cell reached past the full shell table");
      L = (int)std::min<int64_t>(2 * (int64_t)L, full);
      Table* t = c->table(d->cpd, G.sx, G.sy, G.sz, L, &st);
      if (!t) return st;
      const size_t m = tab.size();
      if ((st = upload_list(c, c->fin_list, tab)) != TESS_OK) return st;
      if (c->fin_ctr.alloc(4) != hipSuccess || c->fin_tab.alloc(m + 1) != hipSuccess ||
          c->fin_cap.alloc(m + 1) != hipSuccess)
        return c->fail(TESS_E_NOMEM, "device allocation failed");
      CellArgs a = a0;
      a.tkey = t->key.p;
      a.toff = t->off.p;
      a.tlen = t->len;
      a.kcut = t->kcut;
      a.work = c->fin_list.p;
      a.nwork = (int)m;
      a.xcd_chunk = 0;
      a.redo = c->fin_cap.p + 1;
      a.nredo = c->fin_cap.p;
      a.tredo = c->fin_tab.p + 1;
      a.ntredo = c->fin_tab.p;
      a.defer_table = 1;
      if ((st = c->hip(launch_fill_u32(c->fin_tab.p, 1, 0u, s), "fill")) != TESS_OK ||
          (st = c->hip(launch_fill_u32(c->fin_cap.p, 1, 0u, s), "fill")) != TESS_OK)
        return st;
      if (geo) {
        // the geometry of these cells, staged per item of this pass
        geo->extra.push_back(std::make_unique<GeoStaging>());
        GeoStaging& g = *geo->extra.back();
        geo->extra_items.push_back((uint32_t)m);
        if ((st = geo_staging_alloc(d, g, m, 4)) != TESS_OK) return st;
        a.geo_count = geo->count;
        a.geo_slot = g.slot.p;
        a.geo_v = g.v.p;
        a.geo_ring = g.ring.p;
        a.geo_roff = g.roff.p;
        if ((st = c->hip(launch_cells_geometry(4, a, geo->with_cells, s), "cell kernel")) != TESS_OK) return st;
      } else if ((st = c->hip(launch_cells(4, a, s), "cell kernel")) != TESS_OK) {
        return st;
      }
      if ((st = c->hip(launch_mark_status(c->fin_list.p, (int)m, sorted, a0.status, a0.stats, TESS_CELL_LONG_TABLE, s),
                       "mark")) != TESS_OK)
        return st;
      fc.table += m;
      std::vector<uint32_t> more_cap;
      if ((st = read_list(c, c->fin_tab.p, tab)) != TESS_OK || (st = read_list(c, c->fin_cap.p, more_cap)) != TESS_OK)
        return st;
      spl.insert(spl.end(), more_cap.begin(), more_cap.end());
    }
    if (!spl.empty()) {
      if (geo) return c->fail(TESS_E_CAPACITY, "geometry export holds at most 256 vertices per cell");
      if (a0.fpos == nullptr) return c->fail(TESS_E_CAPACITY, "a cell above 256 vertices needs the face arena");
      Table* t = c->table(d->cpd, G.sx, G.sy, G.sz, L, &st);
      if (!t) return st;
      const size_t m = spl.size();
      CellArgs a = a0;
      a.tkey = t->key.p;
      a.toff = t->off.p;
      a.tlen = t->len;
      a.kcut = t->kcut;
      a.work = nullptr;
      a.redo = nullptr;
      a.tredo = nullptr;
      const size_t per = 8 * spill_wsd_doubles(rmax, fmax) + 4 * spill_wsu_words(fmax) + spill_wsb_bytes(rmax, fmax);
      const int blocks = (int)std::max<size_t>(1, std::min<size_t>({m, (size_t)kSpillBlocks, kSpillBudget / per}));
      if ((st = upload_list(c, c->fin_list, spl)) != TESS_OK) return st;
      if (c->spill_args.alloc(1) != hipSuccess || c->fin_ctr.alloc(4) != hipSuccess ||
          c->fin_tab64.alloc(m) != hipSuccess || c->fin_cap64.alloc(m) != hipSuccess ||
          c->spill_d.alloc(blocks * spill_wsd_doubles(rmax, fmax)) != hipSuccess ||
          c->spill_u.alloc(blocks * spill_wsu_words(fmax)) != hipSuccess ||
          c->spill_b.alloc(blocks * spill_wsb_bytes(rmax, fmax)) != hipSuccess)
        return c->fail(TESS_E_NOMEM, "device allocation failed");
      SpillArgs P;
      std::memset(&P, 0, sizeof(P));
      P.sets = c->spill_args.p;
      P.list32 = c->fin_list.p;
      P.count = (int)m;
      P.rmax = rmax;
      P.fmax = fmax;
      P.wsd = c->spill_d.p;
      P.wsu = c->spill_u.p;
      P.wsb = c->spill_b.p;
      P.redo = c->fin_cap64.p;
      P.nredo = c->fin_ctr.p;
      P.tredo = c->fin_tab64.p;
      P.ntredo = c->fin_ctr.p + 1;
      uint32_t ctr[2] = {0, 0};
      if ((st = c->hip(hipMemcpyAsync(c->spill_args.p, &a, sizeof(CellArgs), hipMemcpyHostToDevice, s), "upload")) !=
              TESS_OK ||
          (st = c->hip(launch_fill_u32(c->fin_ctr.p, 2, 0u, s), "fill")) != TESS_OK ||
          (st = c->hip(launch_cells_spill(P, blocks, s), "spill kernel")) != TESS_OK ||
          (st = c->hip(hipMemcpyAsync(ctr, c->fin_ctr.p, 8, hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
          (st = c->hip(hipStreamSynchronize(s), "spill kernel")) != TESS_OK)
        return st;
      fc.spill += m;
      std::vector<uint64_t> cap64(ctr[0]), tab64(ctr[1]);
      if ((ctr[0] && (st = c->hip(hipMemcpy(cap64.data(), c->fin_cap64.p, 8ull * ctr[0], hipMemcpyDeviceToHost),
                                  "download")) != TESS_OK) ||
          (ctr[1] && (st = c->hip(hipMemcpy(tab64.data(), c->fin_tab64.p, 8ull * ctr[1], hipMemcpyDeviceToHost),
                                  "download")) != TESS_OK))
        return st;
      spl.clear();
      for (uint64_t it : tab64) tab.push_back((uint32_t)it);
      if (!cap64.empty()) {
        if (rmax >= kSpillRmaxMax) {
          fc.failed += cap64.size();
        } else {
          rmax *= 4;
          fmax *= 4;
          for (uint64_t it : cap64) spl.push_back((uint32_t)it);
        }
      }
    }
  }
  fc.table_l = L;
  if (out) *out = fc;
  return TESS_OK;
}

}  // namespace tess
