// runtime.h -- host-side objects behind include/tess.h, shared by the
// translation units of the runtime (api.cpp, model.cpp, batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "launch.h"
#include "tess_internal.h"

namespace tess {

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  bool own = true;  // false: a view into memory owned elsewhere
  ~DevBuf() { reset(); }
  void reset() {
    if (p && own) (void)hipFree(p);
    p = nullptr;
    n = 0;
    own = true;
  }
  void view(T* q, size_t count) {
    reset();
    p = q;
    n = count;
    own = false;
  }
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    reset();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * count);
    if (e == hipSuccess) n = count;
    return e;
  }
};

struct Table {
  DevBuf<double> key;
  DevBuf<int32_t> off;
  int len = 0;
  double kcut = 0;
  int L = 0;
};

// Shell table of a grid shape: entries with |o|_inf <= L in canonical order
// (celery.rs:418-679), see build_table_host.
void build_table_host(int64_t cpd, double sx, double sy, double sz, int L, std::vector<double>& key,
                      std::vector<int32_t>& off, double* kcut);
constexpr int kTableL = 12;
// Position of an offset in the generation loops of celery.rs:437-673, and its
// key (sq(gx sx) + sq(gy sy)) + sq(gz sz) of the gaps (celery.rs:419-427).
uint64_t generation_index(int64_t cpd, int ox, int oy, int oz);
double shell_key(int ox, int oy, int oz, double sx, double sy, double sz);
uint64_t bits(double v);
int64_t icbrt_floor(double x);

// Pinned host staging for large downloads into caller memory: the results
// are copied into the two buffers in turn while host threads move the
// previous chunk into the caller's (pageable) memory.
struct Staging {
  static constexpr size_t kBytes = size_t(64) << 20;  // per buffer
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::mutex mu;
  Staging() = default;
  Staging(const Staging&) = delete;
  Staging& operator=(const Staging&) = delete;
  ~Staging() {
    for (int k = 0; k < 2; ++k) {
      if (buf[k]) (void)hipHostFree(buf[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
  }
};

}  // namespace tess

using namespace tess;

struct tess_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::mutex mu;
  // shell tables keyed by (cpd, sx, sy, sz, L)
  std::map<std::tuple<int64_t, uint64_t, uint64_t, uint64_t, int>, std::unique_ptr<Table>> tables;
  DevBuf<double> diag;
  // spill kernel: workspaces (grown on demand), the pass's arguments, lists
  DevBuf<double> spill_d;
  DevBuf<uint32_t> spill_u;
  DevBuf<uint8_t> spill_b;
  DevBuf<CellArgs> spill_args;
  DevBuf<uint32_t> fin_list, fin_ctr;
  DevBuf<uint32_t> fin_tab, fin_cap;
  DevBuf<uint64_t> fin_tab64, fin_cap64;
  Staging stage;  // large downloads (tess::download)

  tess_status fail(tess_status s, const std::string& m) {
    err = m;
    return s;
  }
  tess_status hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return TESS_OK;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return TESS_E_DEVICE;
  }
  Table* table(int64_t cpd, double sx, double sy, double sz, int L, tess_status* st) {
    auto key = std::make_tuple(cpd, bits(sx), bits(sy), bits(sz), L);
    auto it = tables.find(key);
    if (it != tables.end()) return it->second.get();
    std::vector<double> k;
    std::vector<int32_t> o;
    double kc;
    build_table_host(cpd, sx, sy, sz, L, k, o, &kc);
    auto t = std::make_unique<Table>();
    if (t->key.alloc(k.size()) != hipSuccess || t->off.alloc(o.size()) != hipSuccess) {
      *st = fail(TESS_E_NOMEM, "shell table allocation failed");
      return nullptr;
    }
    if (hipMemcpyAsync(t->key.p, k.data(), 8 * k.size(), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(t->off.p, o.data(), 4 * o.size(), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      *st = fail(TESS_E_DEVICE, "shell table upload failed");
      return nullptr;
    }
    t->len = (int)k.size();
    t->kcut = kc;
    t->L = L;
    Table* r = t.get();
    tables[key] = std::move(t);
    return r;
  }
};

struct tess_cell {
  int64_t original = -1;
  double volume = 0;
  std::vector<int64_t> nbr;
  std::vector<double> area;
  std::vector<std::vector<double>> ring;  // site-relative coordinates per face
  std::vector<double> verts;
};

// Geometry of one cell pass staged per work item: slot (output index + 1),
// vertices, ring entries and ring offsets at the pass's capacity.
struct GeoStaging {
  int vpl = 1;
  DevBuf<uint32_t> slot;
  DevBuf<double> v;
  DevBuf<uint16_t> ring, roff;
};

// Staging filled by the cell passes of a compute itself (the geometry of a
// diagram not yet computed): the first pass per grid position, the capacity
// rerun per rerun item.
struct GeoStage {
  uint32_t* count = nullptr;  // [2n]: live vertices, ring entries per cell
  GeoStaging first, rerun;
  uint32_t nredo = 0;
  // passes with longer shell tables (cells whose security ball reached past
  // the default table), each staging its own items
  std::vector<std::unique_ptr<GeoStaging>> extra;
  std::vector<uint32_t> extra_items;
  bool with_cells = true;  // the staging passes also write the cell outputs
};

// Allocates g for `items` work items at capacity 64 vpl and clears its slots.
tess_status geo_staging_alloc(tess_diagram* d, GeoStaging& g, size_t items, int vpl);

struct tess_diagram {
  tess_ctx* ctx = nullptr;
  tess_opts opts;
  size_t n = 0;
  int64_t cpd = 0;
  int nbins = 0;
  double lo[3], hi[3];
  bool have_groups = false;
  // inputs
  DevBuf<double> x, y, z, xyz_tmp;
  DevBuf<int64_t> groups;
  // grid
  DevBuf<GridDev> grid;
  DevBuf<double> part;
  DevBuf<uint32_t> bins, cnt, delim, cursor, sorted, scratch, sums, big, misc, inv;
  DevBuf<double> sx, sy, sz;
  DevBuf<uint32_t> codes;  // Morton codes (TESS_ORDER_MORTON)
  GridDev hgrid;
  bool binned = false;
  bool computed = false;
  bool inv_ready = false;
  // outputs
  DevBuf<double> volume, sarea, area;
  DevBuf<uint32_t> fcount, status, redo;
  DevBuf<int32_t> snbr, nbr;
  DevBuf<uint64_t> foff, fsums;
  DevBuf<tess_cell_stats> stats;
  DevBuf<Summary> summary;
  DevBuf<uint32_t> work1, redo_list, redo_list2;
  // cells that exhausted the shell table / outgrew the largest LDS variant:
  // count at [0], grid positions from [1]
  DevBuf<uint32_t> tredo, spill;
  // face arena: cell o's faces start at arena position fpos[o]; a batch's
  // sets share the batch's arena (arena_next / arena_cap point there)
  DevBuf<uint64_t> fpos;
  DevBuf<unsigned long long> arena_ctr;
  unsigned long long* arena_next = nullptr;
  uint64_t arena_cap = 0;
  uint64_t table_reruns = 0, spill_cells = 0;
  // learned schedule: features, predictions, work lists
  DevBuf<uint16_t> feat;
  DevBuf<float> pred;
  DevBuf<uint32_t> sched_key, sched_list, sched_hist, sched_sums;
  DevBuf<double> exp_v;
  DevBuf<int32_t> exp_ring, exp_meta;
  // outputs of one outside site (tess_cell_at_particle)
  DevBuf<double> one_site, one_vol, one_area;
  DevBuf<uint32_t> one_u32;  // face count, status
  DevBuf<int32_t> one_nbr;
  // temporaries of the radius queries, kept between calls (grown on demand,
  // freed with the diagram): the per-query counts and their scans, the
  // grid order of large query sets, the kept ids' upper-bound slots
  struct {
    DevBuf<uint32_t> cnt, obins, ocnt, odelim, ocursor, oorder, osums, omisc, slots;
    DevBuf<uint64_t> sums, boff;
    DevBuf<double> ox, oy, oz;
  } qtmp;
  // temporaries of tess_diagram_cells_at_particles, kept between calls
  // (grown on demand, freed with the diagram): the uploaded sites, the face
  // arena and its counters, the rerun / shell-table / spill lists, the grid
  // order of the sites and the scan sums
  struct {
    DevBuf<double> sites, sarea;
    DevBuf<int32_t> snbr;
    DevBuf<uint64_t> fpos, fsums;
    DevBuf<unsigned long long> arena;
    DevBuf<uint32_t> redo, redo_list, tredo, spill, keys, keys_out, iota, order;
    DevBuf<uint8_t> sort_tmp;
    DevBuf<Summary> summary;
  } stmp;
  // geometry of every cell (tess_diagram_compute_geometry)
  DevBuf<uint32_t> geo_count;
  DevBuf<uint64_t> geo_voff, geo_rbase, geo_roff, geo_sums;
  DevBuf<double> geo_v;
  DevBuf<uint16_t> geo_ring;
  uint64_t geo_nv = 0, geo_nr = 0;
  double geo_ms[2] = {0, 0};  // staging pass (+ rerun), scans + compaction
  GeoStage* geo_stage = nullptr;  // set while a compute stages the geometry
  bool geo_ready = false;
  uint64_t total_faces = 0;
  tess_timings tim;
  uint64_t degenerate = 0;
};

#include <limits>

struct tess_model {
  tess_ctx* ctx = nullptr;
  DevBuf<float> master, m1, m2, bias;
  DevBuf<uint16_t> wbf;
  float norm[4] = {0.f, 1.f, 0.f, 1.f};
  float calib[4] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(), 0.f, 0.f};
  int adam_steps = 0;
  // training samples
  DevBuf<uint16_t> X;
  DevBuf<float> T;
  DevBuf<uint16_t> V;
  size_t count = 0;
};

namespace tess {

// dst <- src (device), bytes of it, ordered after the context stream's work
// and finished on return.  Copies of 16 MiB or more go through the
// context's pinned staging, chunk by chunk, overlapping the transfer of one
// chunk with the host copy of the one before; smaller ones are one
// hipMemcpy.
hipError_t download(tess_ctx* c, void* dst, const void* src, size_t bytes);
// dst <- src, len bytes, on up to eight host threads in pieces of at least
// 4 MiB (the staged download's copy out of pinned memory)
void host_copy(char* dst, const char* src, size_t len);

tess_status alloc_all(tess_diagram* d);
// bounds -> counting sort -> gather, enqueued only (graph-capturable)
tess_status bin_enqueue(tess_diagram* d, hipStream_t s);
// bin_enqueue + grid download + the check that every point was binned
tess_status bin(tess_diagram* d);
int pick_vpl(const tess_diagram* d);
CellArgs base_args(tess_diagram* d, const Table* t);
tess_status create_common(tess_ctx* ctx, size_t n, const double lo[3], const double hi[3],
                          const tess_opts* opts, tess_diagram** out);
// Cell kernel over grid positions work[0..nwork) (all cells when work is
// null); capacity overflows are appended to the diagram's redo list when
// push_redo is set.
tess_status run_cells(tess_diagram* d, Table* t, int vpl, const uint32_t* work, int nwork, bool push_redo);
// Work lists of grid positions per vertex capacity (slot 0: 64, 1: 128,
// 2: 256) and the long-search list (slot 3: capacity 256 with a shell table
// of long_l rings); a null list with a count means every cell.
struct Schedule {
  const uint32_t* list[4] = {nullptr, nullptr, nullptr, nullptr};
  int count[4] = {0, 0, 0, 0};
  int long_l = 0;
};
// Fills a schedule after binning (learned mode).
using Planner = std::function<tess_status(tess_diagram*, Schedule&)>;
// Binning, shell table, cell passes (with the capacity rerun), face offsets
// and compaction; fills the diagram's timings.  plan == nullptr: plain.
tess_status compute_cells(tess_diagram* d, const Planner* plan);
// XCD run length of whole-set cell passes (the caller's, else 256)
int plain_xcd_chunk(const tess_diagram* d);

// What finish_cells did.
struct FinishCounts {
  uint64_t table = 0;   // cells recomputed with a longer shell table
  uint64_t spill = 0;   // cells computed by the spill kernel
  uint64_t failed = 0;  // cells that still failed (workspace limit)
  int table_l = kTableL;
};
// The cells an LDS pass could not finish.  table: work items whose security
// ball reached past the shell table -- recomputed at capacity 256 with a
// table twice as long until none is left (the full table has no end);
// spill: work items above 256 vertices or 128 faces -- the global-memory
// kernel, its workspace grown until they fit.  `a` is the pass's template
// (inputs, outputs, filters, site list); items are what its work list
// holds.  geo: the geometry staging of the pass (table reruns then stage
// their geometry too; spill cells are refused there).
tess_status finish_cells(tess_diagram* d, const CellArgs& a, std::vector<uint32_t> table,
                         std::vector<uint32_t> spill, FinishCounts* out, GeoStage* geo = nullptr);
// The no-allocation entry point's scratch address on `device` (noalloc.hip),
// taken once per device with that device current.
tess_status noalloc_setup(int device);
// Reads a device list (count at [0], items from [1]) into v.
tess_status read_list(tess_ctx* c, const uint32_t* dev, std::vector<uint32_t>& v);

}  // namespace tess
