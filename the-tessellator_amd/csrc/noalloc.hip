// noalloc.hip -- tess_compute_noalloc: a diagram's cells (the work of
// tess_diagram_create_device + tess_diagram_compute) without allocating
// device memory, for callers that cannot: code under HIP graph capture,
// serving loops, harnesses that forbid allocation.  Every temporary lives in
// `scratch`, a fixed device array of this library that each GPU holds from
// the first use of the library until the process ends.  Calls on one GPU
// share it, so calls that overlap on one GPU race; it keeps the last call's
// contents, so the counters and lists are cleared by fill kernels at the
// start of every call.
//
// The passes are the library's own: bounds, counting sort and gather into
// the grid, the capacity-64 cell pass, the capacity-256 rerun, the spill
// kernel for cells above 256 vertices, the face-offset scan and the CSR
// compaction into the caller's buffers.  The shell table is the full
// canonical order (celery.rs:418-679) for the grid shape, so no cell needs a
// longer one; it is built on the host into static memory and uploaded once
// per grid shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <limits>
#include <mutex>

#include "launch.h"
#include "runtime.h"

namespace tess {

namespace {

constexpr size_t kNaPoints = TESS_NOALLOC_MAX_POINTS;  // points per call
constexpr int kNaCpd = 32;                              // grid cells per dimension
constexpr size_t kNaBins = (size_t)kNaCpd * kNaCpd * kNaCpd;
constexpr size_t kNaTable = (size_t)(2 * kNaCpd - 1) * (2 * kNaCpd - 1) * (2 * kNaCpd - 1);
constexpr size_t kNaFaces = 48;  // face arena slots per point (cells average about 15.5 faces)
constexpr int kNaParts = 256;
constexpr int kNaSpillBlocks = 2, kNaRmax = 8192, kNaFmax = 1024;
constexpr int kNaDevices = 64;

// The scratch layout (one per GPU).
struct NaArena {
  double x[kNaPoints], y[kNaPoints], z[kNaPoints];
  double sx[kNaPoints], sy[kNaPoints], sz[kNaPoints];
  double part[6 * kNaParts];
  double tkey[kNaTable];
  double sarea[kNaPoints * kNaFaces];
  double spill_d[kNaSpillBlocks * (6 * kNaRmax + 6 * kNaFmax)];
  uint64_t fpos[kNaPoints];
  uint64_t fsums[kNaPoints / 1024 + 2];
  unsigned long long arena_ctr[kArenaWords];
  uint64_t spill_fail[kNaPoints];
  GridDev grid;
  Summary summary;
  CellArgs args;
  int32_t toff[kNaTable];
  int32_t snbr[kNaPoints * kNaFaces];
  uint32_t bins[kNaPoints], sorted[kNaPoints], sort_tmp[kNaPoints], status[kNaPoints];
  uint32_t redo[kNaPoints + 1], redo_list[kNaPoints], spill[kNaPoints + 1], ctr[4];
  uint32_t cnt[kNaBins], delim[kNaBins + 1], cursor[kNaBins], big[kNaBins];
  uint32_t sums[kNaBins / 1024 + 2], misc[4];
  uint32_t spill_u[kNaSpillBlocks * (2 * (kNaFmax + 1) + 7 * kNaFmax)];
  uint8_t spill_b[kNaSpillBlocks * (kNaRmax + kNaFmax)];
};

}  // namespace

static __device__ __align__(256) unsigned char scratch[sizeof(NaArena)];

namespace {

static_assert(sizeof(NaArena) == TESS_NOALLOC_SCRATCH_BYTES, "the scratch size the header states");

// Per device: the scratch array's device address (from hipGetSymbolAddress,
// never the array's host-side name) and the grid shape whose table it holds.
struct NaDevice {
  bool tried = false;
  hipError_t err = hipSuccess;
  NaArena* arena = nullptr;
  int64_t cpd = -1;
  uint64_t sxb = 0, syb = 0, szb = 0;
  int tlen = 0;
};
NaDevice g_dev[kNaDevices];
std::mutex g_mu;  // g_dev and the host table buffers

// Host table buffers (static: the call allocates nothing).
struct NaEnt {
  double k;
  uint64_t g;
  int32_t o;
};
NaEnt g_ent[kNaTable];
double g_key[kNaTable];
int32_t g_off[kNaTable];

template <class T>
T* field(NaArena* a, size_t offset) {
  return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(a) + offset);
}
#define NA(a, m) field<std::remove_extent_t<decltype(NaArena::m)>>(a, offsetof(NaArena, m))

}  // namespace

// Called by tess_ctx_create with the context's device current: the scratch
// address of that device, once per device.
tess_status noalloc_setup(int device) {
  if (device < 0 || device >= kNaDevices) return TESS_OK;  // the entry point refuses such devices
  std::lock_guard<std::mutex> lk(g_mu);
  NaDevice& nd = g_dev[device];
  if (nd.tried) return TESS_OK;
  nd.tried = true;
  void* p = nullptr;
  nd.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (nd.err == hipSuccess) nd.arena = static_cast<NaArena*>(p);
  return TESS_OK;
}

}  // namespace tess

using namespace tess;

extern "C" tess_status tess_compute_noalloc(tess_ctx* ctx, const double* dxyz, size_t n, const double lo[3],
                                            const double hi[3], double cell_density, double* volume,
                                            uint32_t* face_count, uint64_t* face_offset, int32_t* neighbor,
                                            double* face_area, uint64_t face_cap, uint64_t* total_faces) {
  if (!ctx) return TESS_E_INVALID;
  if (!dxyz || !lo || !hi || !volume || !face_count || !face_offset || !total_faces ||
      (face_cap && (!neighbor || !face_area)))
    return ctx->fail(TESS_E_INVALID, "null argument");
  *total_faces = 0;
  if (n == 0) return ctx->fail(TESS_E_INVALID, "empty point set");
  if (n > kNaPoints)
    return ctx->fail(TESS_E_INVALID, "tess_compute_noalloc takes at most " + std::to_string(kNaPoints) +
                                         " points (" + std::to_string(n) + " given)");
  for (int k = 0; k < 3; ++k)
    if (!(hi[k] > lo[k])) return ctx->fail(TESS_E_INVALID, "container box is empty");
  const double density = cell_density > 0 ? cell_density : 1.25;
  const int64_t cpd = icbrt_floor((double)n / density) + 1;  // celery.rs:161-162
  if (cpd > kNaCpd)
    return ctx->fail(TESS_E_INVALID, "tess_compute_noalloc takes grids of at most " + std::to_string(kNaCpd) +
                                         " cells per dimension (cell_density too small)");
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (ctx->device < 0 || ctx->device >= kNaDevices) return ctx->fail(TESS_E_DEVICE, "device index out of range");
  NaArena* A;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    const NaDevice& nd = g_dev[ctx->device];
    if (!nd.arena)
      return ctx->fail(TESS_E_DEVICE, std::string("scratch address unavailable: ") +
                                          (nd.tried ? hipGetErrorString(nd.err) : "context not set up"));
    A = nd.arena;
  }
  hipStream_t s = ctx->stream;
  const int ni = (int)n, nbins = (int)(cpd * cpd * cpd);
  const int nparts = (int)std::min<size_t>(kNaParts, (n + 255) / 256);
  double *x = NA(A, x), *y = NA(A, y), *z = NA(A, z);
  double *sx = NA(A, sx), *sy = NA(A, sy), *sz = NA(A, sz);
  GridDev* grid = field<GridDev>(A, offsetof(NaArena, grid));
  uint32_t* delim = NA(A, delim);
  tess_status st;
  // binning (celery.rs:81-414)
  SortBufs b{NA(A, bins), NA(A, cnt), delim, NA(A, cursor), NA(A, sorted), NA(A, sort_tmp), NA(A, sums),
             NA(A, big), NA(A, misc)};
  GridDev hg;
  uint32_t binned = 0;
  if ((st = ctx->hip(launch_soa(dxyz, ni, x, y, z, s), "soa")) != TESS_OK ||
      (st = ctx->hip(launch_bounds(x, y, z, ni, (int)cpd, NA(A, part), nparts, grid, s), "bounds")) != TESS_OK ||
      (st = ctx->hip(launch_counting_sort(x, y, z, ni, grid, nbins, b, s), "sort")) != TESS_OK ||
      (st = ctx->hip(launch_gather(b.sorted, ni, x, y, z, sx, sy, sz, s), "gather")) != TESS_OK ||
      (st = ctx->hip(hipMemcpyAsync(&hg, grid, sizeof(GridDev), hipMemcpyDeviceToHost, s), "grid copy")) != TESS_OK ||
      (st = ctx->hip(hipMemcpyAsync(&binned, delim + nbins, 4, hipMemcpyDeviceToHost, s), "grid copy")) != TESS_OK ||
      (st = ctx->hip(hipStreamSynchronize(s), "binning")) != TESS_OK)
    return st;
  if (binned != (uint32_t)n) return ctx->fail(TESS_E_DEVICE, "counting sort lost points");
  // the full shell table of this grid shape, once per shape
  int tlen = 0;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    NaDevice& nd = g_dev[ctx->device];
    if (nd.cpd != cpd || nd.sxb != bits(hg.sx) || nd.syb != bits(hg.sy) || nd.szb != bits(hg.sz)) {
      nd.cpd = -1;
      const int L = (int)(cpd - 1);
      size_t m = 0;
      for (int i = -L; i <= L; ++i)
        for (int j = -L; j <= L; ++j)
          for (int k = -L; k <= L; ++k)
            g_ent[m++] = {shell_key(i, j, k, hg.sx, hg.sy, hg.sz), generation_index(cpd, i, j, k), pack_off(i, j, k)};
      std::sort(g_ent, g_ent + m, [](const NaEnt& p, const NaEnt& q) { return p.k != q.k ? p.k < q.k : p.g < q.g; });
      for (size_t t = 0; t < m; ++t) {
        g_key[t] = g_ent[t].k;
        g_off[t] = g_ent[t].o;
      }
      if ((st = ctx->hip(hipMemcpyAsync(NA(A, tkey), g_key, 8 * m, hipMemcpyHostToDevice, s), "table upload")) !=
              TESS_OK ||
          (st = ctx->hip(hipMemcpyAsync(NA(A, toff), g_off, 4 * m, hipMemcpyHostToDevice, s), "table upload")) !=
              TESS_OK ||
          (st = ctx->hip(hipStreamSynchronize(s), "table upload")) != TESS_OK)
        return st;
      nd.cpd = cpd;
      nd.sxb = bits(hg.sx);
      nd.syb = bits(hg.sy);
      nd.szb = bits(hg.sz);
      nd.tlen = (int)m;
    }
    tlen = nd.tlen;
  }
  // counters and lists start at zero
  uint32_t* redo = NA(A, redo);
  uint32_t* spill = NA(A, spill);
  uint32_t* ctr = NA(A, ctr);
  Summary* summary = field<Summary>(A, offsetof(NaArena, summary));
  unsigned long long* arena_next = NA(A, arena_ctr);
  if ((st = ctx->hip(launch_fill_u32(redo, 1, 0u, s), "fill")) != TESS_OK ||
      (st = ctx->hip(launch_fill_u32(spill, 1, 0u, s), "fill")) != TESS_OK ||
      (st = ctx->hip(launch_fill_u32(ctr, 4, 0u, s), "fill")) != TESS_OK ||
      (st = ctx->hip(launch_fill_u32(reinterpret_cast<uint32_t*>(arena_next), 2 * kArenaWords, 0u, s), "fill")) !=
          TESS_OK ||
      (st = ctx->hip(launch_fill_u32(reinterpret_cast<uint32_t*>(summary), sizeof(Summary) / 4, 0u, s), "fill")) !=
          TESS_OK)
    return st;
  CellArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = ni;
  a.spx = sx;
  a.spy = sy;
  a.spz = sz;
  a.sorted = b.sorted;
  a.delim = delim;
  a.grid = grid;
  a.tkey = NA(A, tkey);
  a.toff = NA(A, toff);
  a.tlen = tlen;
  a.kcut = std::numeric_limits<double>::infinity();
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = lo[k];
    a.hi[k] = hi[k];
  }
  a.nwork = ni;
  a.xcd_chunk = 256;
  a.target_group = -1;
  a.radius_limit = std::numeric_limits<double>::quiet_NaN();
  a.volume = volume;
  a.fcount = face_count;
  a.snbr = NA(A, snbr);
  a.sarea = NA(A, sarea);
  a.fpos = NA(A, fpos);
  a.arena_next = arena_next;
  a.arena_cap = (uint64_t)kNaPoints * kNaFaces;  // the whole scratch arena
  a.status = NA(A, status);
  a.summary = summary;
  a.redo = redo + 1;
  a.nredo = redo;
  // capacity 64, then the overflows at 256, then the spill kernel
  uint32_t nredo = 0, nspill = 0;
  if ((st = ctx->hip(launch_cells(1, a, s), "cell kernel")) != TESS_OK ||
      (st = ctx->hip(hipMemcpyAsync(&nredo, redo, 4, hipMemcpyDeviceToHost, s), "redo")) != TESS_OK ||
      (st = ctx->hip(hipStreamSynchronize(s), "cell kernel")) != TESS_OK)
    return st;
  if (nredo) {
    CellArgs r = a;
    r.work = NA(A, redo_list);
    r.nwork = (int)nredo;
    r.xcd_chunk = 0;
    r.redo = spill + 1;
    r.nredo = spill;
    if ((st = ctx->hip(hipMemcpyAsync(NA(A, redo_list), redo + 1, 4ull * nredo, hipMemcpyDeviceToDevice, s),
                       "redo")) != TESS_OK ||
        (st = ctx->hip(launch_cells(4, r, s), "cell kernel")) != TESS_OK ||
        (st = ctx->hip(hipMemcpyAsync(&nspill, spill, 4, hipMemcpyDeviceToHost, s), "spill")) != TESS_OK ||
        (st = ctx->hip(hipStreamSynchronize(s), "cell kernel")) != TESS_OK)
      return st;
  }
  if (nspill) {
    CellArgs* dargs = field<CellArgs>(A, offsetof(NaArena, args));
    SpillArgs P;
    std::memset(&P, 0, sizeof(P));
    P.sets = dargs;
    P.list32 = spill + 1;
    P.count = (int)nspill;
    P.rmax = kNaRmax;
    P.fmax = kNaFmax;
    P.wsd = NA(A, spill_d);
    P.wsu = NA(A, spill_u);
    P.wsb = NA(A, spill_b);
    P.redo = NA(A, spill_fail);
    P.nredo = ctr;
    uint32_t failed = 0;
    if ((st = ctx->hip(hipMemcpyAsync(dargs, &a, sizeof(CellArgs), hipMemcpyHostToDevice, s), "upload")) != TESS_OK ||
        (st = ctx->hip(launch_cells_spill(P, kNaSpillBlocks, s), "spill kernel")) != TESS_OK ||
        (st = ctx->hip(hipMemcpyAsync(&failed, ctr, 4, hipMemcpyDeviceToHost, s), "spill")) != TESS_OK ||
        (st = ctx->hip(hipStreamSynchronize(s), "spill kernel")) != TESS_OK)
      return st;
    if (failed)
      return ctx->fail(TESS_E_CAPACITY, std::to_string(failed) + " cells exceed the no-allocation spill workspace (" +
                                            std::to_string(kNaFmax) + " faces, " + std::to_string(kNaRmax) +
                                            " ring entries)");
  }
  // face offsets, then the caller's CSR
  uint64_t total = 0;
  unsigned long long used = 0;
  Summary sm;
  if ((st = ctx->hip(launch_face_offsets(face_count, ni, face_offset, NA(A, fsums), s), "face offsets")) != TESS_OK ||
      (st = ctx->hip(hipMemcpyAsync(&total, face_offset + n, 8, hipMemcpyDeviceToHost, s), "total")) != TESS_OK ||
      (st = ctx->hip(hipMemcpyAsync(&used, arena_next + (size_t)kArenaStride * (kArenaParts + 1), 8, hipMemcpyDeviceToHost, s),
                     "arena")) != TESS_OK ||
      (st = ctx->hip(hipMemcpyAsync(&sm, summary, sizeof(Summary), hipMemcpyDeviceToHost, s), "summary")) !=
          TESS_OK ||
      (st = ctx->hip(hipStreamSynchronize(s), "face offsets")) != TESS_OK)
    return st;
  if (used)
    return ctx->fail(TESS_E_CAPACITY, "the face arena (" + std::to_string(kNaFaces) + " faces per point) is full");
  *total_faces = total;
  if (sm.degenerate)
    return ctx->fail(TESS_E_DEGENERATE, std::to_string(sm.degenerate) + " cells failed topology checks");
  if (total > face_cap)
    return ctx->fail(TESS_E_CAPACITY, "the cells have " + std::to_string(total) + " faces, face_cap is " +
                                          std::to_string(face_cap));
  if ((st = ctx->hip(launch_compact(face_count, face_offset, ni, NA(A, snbr), NA(A, sarea), neighbor, face_area, s,
                                    NA(A, fpos)),
                     "compact")) != TESS_OK)
    return st;
  return ctx->hip(hipStreamSynchronize(s), "compact");
}
