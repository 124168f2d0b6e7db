// cell.hip -- one Voronoi cell per wavefront (gfx950, wave64).
//
// Per cell (Cell::compute_voronoi_cell, interface.rs:257-313):
//   * the container box, translated by -site, is the starting polyhedron
//     (polyhedron.rs:268-392, :859-867), held in LDS as vertices + half-edge
//     pairs (flip(h) = h ^ 1) + faces;
//   * candidates arrive in the reference's shell order (celery.rs:418-679,
//     :971-1018): 64 table entries per batch, their grid cells' points
//     gathered 64 per chunk (lane = candidate), bisector planes built
//     lane-parallel (vector3.rs:223-239);
//   * candidates are then applied in order: lanes classify vertices
//     (vector3.rs:170-193, tolerance 1e-12), a ballot decides whether the
//     cell changes, and a lane-parallel clip rewires the half-edges: lanes
//     find the Outside runs of every face, one lane per run places the exit
//     point intersection(outside, inside) (vector3.rs:213-219) and the bridge
//     half-edge, and the new face (tagged with the point id) is linked
//     through an exit-point map;
//   * the security-radius stop: a shell entry whose key exceeds
//     (4 R^2)(1 + 1e-9), R = max |vertex|, and every candidate with |p|^2
//     above the same bound, cannot cut the cell, so skipping them leaves the
//     result bit-identical to streaming every point;
//   * outputs (polyhedron.rs:776-855, interface.rs:408-410): fan-triangulated
//     weighted normals anchored at each face's lexicographically smallest
//     vertex, volume summed over faces in creation order, areas, neighbours
//     sorted ascending -- all written coalesced per cell.
// All f64 arithmetic is compiled with -ffp-contract=off so every operation
// is one IEEE-rounded operation in the same order as the CPU restatement.
#include <math.h>

#include <type_traits>

#include "launch.h"
#include "tess_internal.h"

namespace tess {

namespace {

constexpr double kTol = 1e-12;  // polyhedron.rs:221-223
enum : uint8_t { C_OUT = 0, C_INC = 1, C_IN = 2 };

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// One wavefront per workgroup: the wave's LDS operations execute in issue
// order, so a phase boundary only has to stop the compiler from moving LDS
// accesses across it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ int popc(uint64_t m) { return __popcll(m); }
__device__ __forceinline__ int ctz64(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }
// Set bits of m below this lane (lane = the lane's index in its wavefront).
__device__ __forceinline__ int mbcnt(uint64_t m, int lane) {
  (void)lane;
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// An empty asm that hands v back as an opaque per-lane value: the compiler
// keeps the slot ranks below as vector values and does not fold their
// compare into the masks they came from.
__device__ __forceinline__ void opaque(int& v) { asm volatile("" : "+v"(v)); }
// Bit `lane` of a wave-uniform mask as this lane's predicate: the mask is
// used as the lane mask it is, with no per-lane shift.
__device__ __forceinline__ bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
// Lane l's v as a scalar.
__device__ __forceinline__ double rl_d(double v, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffu), l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// A value every lane holds (e.g. an LDS broadcast read) as a scalar.
__device__ __forceinline__ int uni_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni_d(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(u & 0xffffffffu));
  const int hi = __builtin_amdgcn_readfirstlane((int)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// Running maximum of unsigned v over the lanes (inclusive): DPP row shifts
// inside the rows of 16 (lanes without a source read 0, the identity), then
// the row maxima broadcast to the later rows.
__device__ __forceinline__ uint32_t wave_max_scan_u32(uint32_t v) {
  uint32_t t;
  t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true); v = t > v ? t : v;  // row_shr:1
  t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true); v = t > v ? t : v;  // row_shr:2
  t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true); v = t > v ? t : v;  // row_shr:4
  t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true); v = t > v ? t : v;  // row_shr:8
  t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); v = t > v ? t : v;  // row_bcast:15 -> rows 1, 3
  t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); v = t > v ? t : v;  // row_bcast:31 -> rows 2, 3
  return v;
}
// Maximum of unsigned v over the wavefront, as a scalar.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_max_scan_u32(v), 63);
}
// Maximum over the wavefront of non-negative doubles (no -0, no NaN), as a
// scalar.  Such doubles order as their bit patterns do as unsigned
// integers: the largest high word, then the largest low word among the
// lanes that hold it.
__device__ __forceinline__ double wave_max_uni(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const uint32_t hi = (uint32_t)(u >> 32), lo = (uint32_t)(u & 0xffffffffu);
  const uint32_t mh = wave_max_u32(hi);
  const uint32_t ml = wave_max_u32(hi == mh ? lo : 0u);
  return __longlong_as_double((long long)(((unsigned long long)mh << 32) | ml));
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// Inclusive prefix sum over the wavefront with DPP row shifts inside each
// row of 16 lanes, then the row totals broadcast by row_bcast:15 / :31
// (lanes without a source add 0).
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
  (void)lane;
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 -> rows 1, 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 -> rows 2, 3
  return v;
}

// Position of the k-th (0-based) set bit of w; w must have more than k bits.
__device__ __forceinline__ int select_bit(uint64_t w, int k) {
  int pos = 0;
#pragma unroll
  for (int sh = 32; sh > 0; sh >>= 1) {
    int c = __popcll(w & ((1ull << sh) - 1ull));
    if (k >= c) {
      k -= c;
      w >>= sh;
      pos += sh;
    }
  }
  return pos;
}
template <int W>
__device__ __forceinline__ int select_multi(const uint64_t (&m)[W], int k) {
#pragma unroll
  for (int r = 0; r < W; ++r) {
    int c = __popcll(m[r]);
    if (k < c) return 64 * r + select_bit(m[r], k);
    k -= c;
  }
  return -1;
}
template <int W>
__device__ __forceinline__ int popc_multi(const uint64_t (&m)[W]) {
  int c = 0;
#pragma unroll
  for (int r = 0; r < W; ++r) c += __popcll(m[r]);
  return c;
}
// Lowest k set bits of m (k <= popcount(m)).
template <int W>
__device__ __forceinline__ void lowest_bits(const uint64_t (&m)[W], int k, uint64_t (&out)[W]) {
#pragma unroll
  for (int r = 0; r < W; ++r) {
    int c = __popcll(m[r]);
    if (k >= c) {
      out[r] = m[r];
      k -= c;
    } else if (k == 0) {
      out[r] = 0;
    } else {
      int pos = select_bit(m[r], k);  // first bit not taken
      out[r] = m[r] & ((1ull << pos) - 1ull);
      k = 0;
    }
  }
}

__device__ __forceinline__ int axis_index(double v, double vmin, double vmax, double inv, int cpd) {
  if (v >= vmax) return cpd - 1;  // celery.rs:269-282
  double t = (v - vmin) * inv;
  if (!(t > 0.0)) return 0;       // Rust `as usize` saturates, NaN -> 0
  if (t >= 2147483647.0) return cpd - 1;
  int i = (int)t;
  return i < cpd - 1 ? i : cpd - 1;
}

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b & 7).
// With runs of K list entries dealt to the XCDs in turn, XCD x's j-th block
// takes entry j % K of run (j / K) * 8 + x: neighbouring cells share an L2
// while the heavy regions of the list still spread over all XCDs.
__host__ __device__ __forceinline__ uint32_t xcd_map(uint32_t b, uint32_t K) {
  const uint32_t x = b & 7u, j = b >> 3;
  return ((j / K) * 8u + x) * K + (j % K);
}
// Blocks needed to cover n entries (whole rounds of 8 runs).
__host__ __device__ __forceinline__ uint32_t xcd_grid(uint32_t n, uint32_t K) {
  const uint32_t round = 8u * K;
  return (n + round - 1) / round * round;
}

// Distance from a site coordinate to grid cell b's slab along one axis, a
// lower bound for every point binned there: the slab [vmin + b*size,
// vmin + (b+1)*size] is widened by a millionth of its width, far more than
// the rounding of the binning and of the distance itself.
__device__ __forceinline__ double box_gap(int b, double vmin, double size, double site) {
  const double lo = (vmin + (double)b * size) - site, hi = lo + size, m = size * 1e-6;
  if (lo - m > 0.0) return lo - m;
  if (hi + m < 0.0) return -(hi + m);
  return 0.0;
}

__device__ __forceinline__ double stop_thr(double r2) { return (4.0 * r2) * (1.0 + 1e-9); }

// nf consecutive face slots of the arena: in the region of this wave's XCD,
// else in the shared overflow region; ~0 (and the no-room count bumped)
// when both are full.  Wave-uniform result.
__device__ __forceinline__ uint64_t arena_reserve(unsigned long long* ctr, uint64_t cap, uint32_t nf, int lane) {
  unsigned long long p = ~0ull;
  if (lane == 0) {
    const uint64_t part = arena_part_size(cap);
    const uint32_t x = blockIdx.x % kArenaParts;
    unsigned long long q = atomicAdd(ctr + (size_t)kArenaStride * x, (unsigned long long)nf);
    if (q + nf <= part) {
      p = (unsigned long long)x * part + q;
    } else {
      q = atomicAdd(ctr + (size_t)kArenaStride * kArenaParts, (unsigned long long)nf);
      if (kArenaParts * part + q + nf <= cap)
        p = kArenaParts * part + q;
      else
        atomicAdd(ctr + (size_t)kArenaStride * (kArenaParts + 1), 1ull);
    }
  }
  const uint32_t plo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(p & 0xffffffffull));
  const uint32_t phi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(p >> 32));
  return ((uint64_t)phi << 32) | plo;
}

// A chunk with at least this many staged candidates is pre-tested lane-parallel.
constexpr int kPreMin = 16;
// A chunk looks at up to 64 * WIN stream positions (WIN per lane, WIN at
// most kWinMax) and stages the first 64 kept candidates among them.
constexpr int kWinMax = 4;

__device__ __forceinline__ bool lex_less(double ax, double ay, double az, double bx, double by, double bz) {
  if (ax != bx) return ax < bx;
  if (ay != by) return ay < by;
  return az < bz;
}

// Container cube in half-edge pairs.  Vertices FDL FDR FUR FUL BDL BDR BUR
// BUL, faces F(-y) R(+x) B(+y) L(-x) U(+z) D(-z) (polyhedron.rs:73-392,
// with DR on face D).
__constant__ uint8_t kCubeTgt[24] = {3, 2, 0, 3, 1, 0, 2, 1, 2, 6, 5, 1, 6, 5, 6, 7, 4, 5, 7, 4, 7, 3, 0, 4};
__constant__ uint16_t kCubeNext[24] = {2,  9,  4, 20, 6,  23, 0,  10, 7,  15, 12, 5,
                                       8, 16, 13, 21, 18, 11, 14, 22, 19, 1,  3,  17};
__constant__ uint8_t kCubeFace[24] = {0, 4, 0, 3, 0, 5, 0, 1, 1, 4, 1, 5, 1, 2, 2, 4, 2, 5, 2, 3, 3, 4, 3, 5};
__constant__ uint16_t kCubeStart[6] = {0, 8, 14, 20, 1, 5};
__constant__ int32_t kWallTag[6] = {-3, -2, -4, -1, -6, -5};

template <int VPL, int PPL, int FPL>
struct Lds {
  static constexpr int VC = 64 * VPL, PC = 64 * PPL, HC = 2 * PC, FC = 64 * FPL;
  static_assert((VC & (VC - 1)) == 0, "vertex slot numbers are masked with VC - 1");
  // a pair slot number: one byte while there are at most 256 pairs
  using PairSlot = typename std::conditional<(PC > 256), uint16_t, uint8_t>::type;
  // a half-edge number: one byte while there are at most 256 half-edges
  using HalfEdge = typename std::conditional<(HC > 256), uint16_t, uint8_t>::type;
  double vx[VC], vy[VC], vz[VC];
  // kept candidates of the current chunk (in stream order); before them,
  // the window's entry heads; the output phase reuses the space for
  // per-face volume terms
  union {
    struct {
      double p[64][3];  // candidate - site
      int32_t tag[64];
      int32_t ent[64];
    } cand;
    uint8_t head[64 * kWinMax];  // the entry whose candidates start at window position p (0xff: none)
    double fcontrib[FC > 64 ? FC : 64];
  } u;
  int32_t f_tag[FC];
  HalfEdge he_next[HC];
  uint16_t f_start[FC];
  uint16_t f_seq[FC];
  uint16_t fentry[FC];  // fast clip: the Outside-run entry half-edge of each face
  uint16_t fexit[FC];   // ... and its exit half-edge
  // per-cut scratch of the general clip and of the fast clip, and the
  // staging's entry table (never live at the same time)
  union {
    struct {
      uint32_t bstart[64];  // first candidate position of each entry
      uint32_t bpref[64];   // its first chunk position
    } st;
    struct {
      uint16_t exit_of[VC];
      uint16_t run_a[64];
      uint8_t newv_of[HC];
    } g;
    struct {
      uint16_t cpair[64];    // crossing pair k: 2 * pair + (half-edge 2p points Inside)
      uint8_t vslot[64];     // k-th free vertex slot
      PairSlot pslot[64];    // k-th free pair slot
      uint8_t rank[PC];      // rank of a crossing pair
    } f;
  } cs;
  uint8_t he_tgt[HC];
  uint8_t he_face[HC];
  uint8_t vcls[VC];
  uint8_t fmark[FC];
};

enum { CUT_OK = 0, CUT_CAPACITY = 1, CUT_DEGENERATE = 2, CUT_GENERAL = 3 };

template <int VPL, int PPL, int FPL>
struct CellState {
  uint64_t VA[VPL];  // live vertex slots
  uint64_t PA[PPL];  // live half-edge pairs
  uint64_t FA[FPL];  // live faces
  int seq;           // next face creation number
  double rx[VPL], ry[VPL], rz[VPL];  // this lane's vertex slots (register copy)
};

// n.v of vertex v: the sequential test's arithmetic on the same operands,
// so bit for bit its value.
template <int VPL, int PPL, int FPL>
__device__ __forceinline__ double ndot(const Lds<VPL, PPL, FPL>& S, int v, double nx, double ny, double nz) {
  return (nx * S.vx[v] + ny * S.vy[v]) + nz * S.vz[v];
}

// Rank of pair p among the set bits of m (bits below p).
template <int W>
__device__ __forceinline__ int rank_multi(const uint64_t (&m)[W], int p) {
  int c = 0;
#pragma unroll
  for (int r = 0; r < W; ++r) {
    if (r < (p >> 6))
      c += __popcll(m[r]);
    else if (r == (p >> 6))
      c += __popcll(m[r] & ((1ull << (p & 63)) - 1ull));
  }
  return c;
}

// Clip for the common configuration: no Incident vertex is adjacent to an
// Outside one and every face crossed by the plane has one Outside run.  Then
// every Inside-Outside edge pair is both the entry of a run (on the face
// where it goes Inside -> Outside) and the exit of another (on the other
// face), so the pairs index everything: pair k of the crossing set owns new
// vertex k (its intersection point) and bridge pair k (the bridge on its
// entry face and that bridge's flip on the new face).  Returns CUT_GENERAL
// without touching the polyhedron when the configuration is not the common
// one.
template <int VPL, int PPL, int FPL>
__device__ __forceinline__ int clip_fast(Lds<VPL, PPL, FPL>& S, CellState<VPL, PPL, FPL>& C,
                                         const uint64_t (&OM)[VPL], double pnx, double pny, double pnz, double off,
                                         int32_t tag, int lane, int* newv) {
  // 1. scan pairs; crossing pairs get a rank (prefix of the crossing ballot)
  //    and are listed in rank order with their targets and faces
  uint64_t CM[PPL], PD[PPL];
  bool slow = false;
  int before = 0;
#pragma unroll
  for (int r = 0; r < PPL; ++r) {
    if (r > 0 && C.PA[r] == 0) {  // pair slots fill from the bottom
      CM[r] = 0;
      PD[r] = 0;
      continue;
    }
    const int p = lane + 64 * r;
    const bool al = lane_bit(C.PA[r]);
    // read on every lane (a dead pair's stale target, kept in range, is
    // masked off below): no divergent branch around the reads
    constexpr int VC = Lds<VPL, PPL, FPL>::VC;
    const uint32_t c0 = S.vcls[S.he_tgt[2 * p] & (VC - 1)];
    const uint32_t c1 = S.vcls[S.he_tgt[2 * p + 1] & (VC - 1)];
    const bool o0 = c0 == C_OUT, o1 = c1 == C_OUT;
    const bool cross = al && ((o0 && c1 == C_IN) || (o1 && c0 == C_IN));
    slow = slow || (al && ((o0 && c1 == C_INC) || (o1 && c0 == C_INC)));
    CM[r] = ballot(cross);
    PD[r] = ballot(al && (o0 || o1) && !cross);
    if (cross) {
      const int k = before + mbcnt(CM[r], lane);
      if (k < 64) {
        S.cs.f.cpair[k] = (uint16_t)(2 * p + (o0 ? 0 : 1));
        S.cs.f.rank[p] = (uint8_t)k;
      }
    }
    before += popc(CM[r]);
  }
  if (ballot(slow)) return CUT_GENERAL;
  const int ncross = before;
  if (ncross < 3) return CUT_GENERAL;
  if (ncross > 64) return CUT_CAPACITY;
  // 2. the first ncross free vertex slots (slots holding no vertex, or an
  //    Outside vertex, which the cut removes: new vertices are written only
  //    after every lane has read the old ones) and free pair slots (pairs
  //    that die are free)
  uint64_t vfree[VPL], pfree[PPL];
#pragma unroll
  for (int r = 0; r < VPL; ++r) vfree[r] = ~C.VA[r] | OM[r];
#pragma unroll
  for (int r = 0; r < PPL; ++r) pfree[r] = ~(C.PA[r] & ~PD[r]);
  if (ncross > popc_multi(vfree) || ncross > popc_multi(pfree)) return CUT_CAPACITY;
  int nf = -1;
#pragma unroll
  for (int r = FPL - 1; r >= 0; --r)
    if (~C.FA[r]) nf = 64 * r + ctz64(~C.FA[r]);
  if (nf < 0) return CUT_CAPACITY;
  // the slots taken (the lowest ncross free ones) as masks, from the same ranks
  uint64_t vnew[VPL], pnew[PPL];
  {
    int bv = 0;
#pragma unroll
    for (int r = 0; r < VPL; ++r) {
      int k = bv + mbcnt(vfree[r], lane);
      opaque(k);
      const bool take = lane_bit(vfree[r]) && k < ncross;
      if (take) S.cs.f.vslot[k] = (uint8_t)(lane + 64 * r);
      vnew[r] = ballot(take);
      bv += popc(vfree[r]);
    }
    int bp = 0;
#pragma unroll
    for (int r = 0; r < PPL; ++r) {
      int k = bp + mbcnt(pfree[r], lane);
      opaque(k);
      const bool take = lane_bit(pfree[r]) && k < ncross;
      if (take) S.cs.f.pslot[k] = (typename Lds<VPL, PPL, FPL>::PairSlot)(lane + 64 * r);
      pnew[r] = ballot(take);
      bp += popc(pfree[r]);
    }
  }
  wave_sync();
  // 3. lane k < ncross handles crossing pair k: intersection(outside,
  //    inside) into new vertex slot k, the face maps
  const bool mine = lane < ncross;
  int fin = 0, fout = 0, hin = 0, sv = 0, sp = 0;
  double nx = 0.0, ny = 0.0, nz = 0.0;
  if (mine) {
    hin = S.cs.f.cpair[lane];  // the half-edge that enters the Outside end
    sv = S.cs.f.vslot[lane];
    sp = S.cs.f.pslot[lane];
    const int o = S.he_tgt[hin], i = S.he_tgt[hin ^ 1];
    fin = S.he_face[hin];
    fout = S.he_face[hin ^ 1];
    const double dO = ndot(S, o, pnx, pny, pnz), dI = ndot(S, i, pnx, pny, pnz);
    const double t = (off - dO) / (dI - dO);
    const double ox = S.vx[o], oy = S.vy[o], oz = S.vz[o];
    nx = ox + (S.vx[i] - ox) * t;
    ny = oy + (S.vy[i] - oy) * t;
    nz = oz + (S.vz[i] - oz) * t;
    S.fentry[fin] = (uint16_t)hin;
    S.fexit[fout] = (uint16_t)(hin ^ 1);
  }
  wave_sync();
  // 4. one run per face: every writer reads back its own half-edge; the same
  //    round trip fetches the run's exit (on its entry face) and the entry of
  //    the run on its exit face
  bool clash = false;
  int b = 0, a2 = 0;
  if (mine) {
    const int e_in = S.fentry[fin], x_out = S.fexit[fout];
    b = S.fexit[fin];
    a2 = S.fentry[fout];
    clash = e_in != hin || x_out != (hin ^ 1);
  }
  if (ballot(clash)) return CUT_GENERAL;
  // 5. commit: bridge k on the entry face, its flip on the new face.  The
  //    new vertex may take the slot of an Outside one, so it is written only
  //    now: every read of the old vertices is done and the general clip can
  //    no longer be asked to start from the uncut cell.
  if (mine) {
    S.vx[sv] = nx;
    S.vy[sv] = ny;
    S.vz[sv] = nz;
    const int Xe = S.cs.f.vslot[S.cs.f.rank[b >> 1]];
    const int nxt = 2 * S.cs.f.pslot[S.cs.f.rank[a2 >> 1]] + 1;
    const int br = 2 * sp;
    S.he_tgt[hin] = (uint8_t)sv;
    S.he_next[hin] = (uint16_t)br;
    S.he_tgt[br] = (uint8_t)Xe;
    S.he_face[br] = (uint8_t)fin;
    S.he_next[br] = (uint16_t)b;
    S.he_tgt[br + 1] = (uint8_t)sv;
    S.he_face[br + 1] = (uint8_t)nf;
    S.he_next[br + 1] = (uint16_t)nxt;
    S.f_start[fin] = (uint16_t)br;
  }
#pragma unroll
  for (int r = 0; r < VPL; ++r) C.VA[r] = (C.VA[r] & ~OM[r]) | vnew[r];
#pragma unroll
  for (int r = 0; r < PPL; ++r) C.PA[r] = (C.PA[r] & ~PD[r]) | pnew[r];
  if (lane == 0) {
    S.f_tag[nf] = tag;
    S.f_seq[nf] = (uint16_t)C.seq;
    S.f_start[nf] = (uint16_t)(2 * S.cs.f.pslot[0] + 1);
  }
  C.seq++;
  // faces whose every vertex was Outside (only possible when pairs died)
  int anyd = 0;
#pragma unroll
  for (int r = 0; r < PPL; ++r) anyd |= (PD[r] != 0);
  if (anyd) {
#pragma unroll
    for (int r = 0; r < FPL; ++r) S.fmark[lane + 64 * r] = 0;
    wave_sync();
#pragma unroll
    for (int r = 0; r < PPL; ++r) {
      if (r > 0 && C.PA[r] == 0) continue;
      const int q = lane + 64 * r;
      if (lane_bit(C.PA[r])) {
        S.fmark[S.he_face[2 * q]] = 1;
        S.fmark[S.he_face[2 * q + 1]] = 1;
      }
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < FPL; ++r) C.FA[r] = ballot(S.fmark[lane + 64 * r] != 0);
  } else {
#pragma unroll
    for (int r = 0; r < FPL; ++r)
      if ((nf >> 6) == r) C.FA[r] |= 1ull << (nf & 63);
  }
  wave_sync();
#pragma unroll
  for (int r = 0; r < VPL; ++r) {
    const int v = lane + 64 * r;
    if (lane_bit(vnew[r])) {
      C.rx[r] = S.vx[v];
      C.ry[r] = S.vy[v];
      C.rz[r] = S.vz[v];
    }
  }
  *newv = ncross;
  return CUT_OK;
}

// Apply one bisector that puts at least one vertex Outside.  OM: Outside
// vertices; S.vcls holds every live vertex's class (n.v is recomputed).
template <int VPL, int PPL, int FPL>
__device__ __forceinline__ int clip(Lds<VPL, PPL, FPL>& S, CellState<VPL, PPL, FPL>& C, const uint64_t (&OM)[VPL],
                    double pnx, double pny, double pnz, double off, int32_t tag, int lane, int* newv) {
  constexpr int HC = Lds<VPL, PPL, FPL>::HC;
#pragma unroll
  for (int r = 0; r < VPL; ++r) S.cs.g.exit_of[lane + 64 * r] = 0xffff;

  // 1. scan pairs: Outside-run entries (P -> O, P not Outside) and pairs to
  //    delete (an Outside end whose other end is not Inside).
  uint64_t PD[PPL];
  int nruns = 0;
#pragma unroll
  for (int r = 0; r < PPL; ++r) {
    const int p = lane + 64 * r;
    const bool al = lane_bit(C.PA[r]);
    bool ent = false, del = false;
    int eh = 0;
    if (al) {
      const int t0 = S.he_tgt[2 * p], t1 = S.he_tgt[2 * p + 1];
      const uint8_t c0 = S.vcls[t0], c1 = S.vcls[t1];
      if (c0 == C_OUT && c1 != C_OUT) {
        ent = true;
        eh = 2 * p;
      } else if (c1 == C_OUT && c0 != C_OUT) {
        ent = true;
        eh = 2 * p + 1;
      }
      del = (c0 == C_OUT && c1 != C_IN) || (c1 == C_OUT && c0 != C_IN);
    }
    PD[r] = ballot(del);
    const uint64_t em = ballot(ent);
    const int pos = nruns + mbcnt(em, lane);
    if (ent && pos < 64) S.cs.g.run_a[pos] = (uint16_t)eh;
    nruns += popc(em);
  }
  if (nruns > 64) return CUT_CAPACITY;
  if (nruns < 1) return CUT_DEGENERATE;
  wave_sync();

  // 2. one lane per run: walk to the exit, classify the run's ends.
  const bool isrun = lane < nruns;
  int a = 0, b = 0, P = 0, Q = 0, Ok = 0, nb = 0, pred = 0, fa = 0;
  uint8_t cP = C_IN, cQ = C_IN;
  bool one = false, two = false, bad = false;
  if (isrun) {
    a = S.cs.g.run_a[lane];
    P = S.he_tgt[a ^ 1];
    cP = S.vcls[P];
    int x = S.he_next[a], g = 0;
    while (S.vcls[S.he_tgt[x]] == C_OUT) {
      x = S.he_next[x];
      if (++g > HC) {
        bad = true;
        break;
      }
    }
    b = x;
    Q = S.he_tgt[b];
    cQ = S.vcls[Q];
    Ok = S.he_tgt[b ^ 1];
    nb = S.he_next[b];
    fa = S.he_face[a];
    if (cP != C_IN && cQ != C_IN) {
      if (nb == a)
        one = true;  // face reduced to the single vertex P == Q
      else if (S.he_tgt[nb] == P && S.he_next[nb] == a)
        two = true;  // face reduced to the edge Q -> P
    }
    if (!one && !two && cP != C_IN) {
      int u = (cQ == C_IN) ? b : nb;
      int g2 = 0;
      while (S.he_next[u] != a) {
        u = S.he_next[u];
        if (++g2 > HC) {
          bad = true;
          break;
        }
      }
      pred = u;
    }
  }
  if (ballot(bad)) return CUT_DEGENERATE;
  const bool needX = isrun && cQ == C_IN;
  const bool normal = isrun && !one && !two;

  // 3. allocate vertex slots (reusing Outside slots), pairs, the new face.
  uint64_t vkeep[VPL], vfree[VPL], pkeep[PPL], pfree[PPL];
#pragma unroll
  for (int r = 0; r < VPL; ++r) {
    vkeep[r] = C.VA[r] & ~OM[r];
    vfree[r] = ~vkeep[r];
  }
#pragma unroll
  for (int r = 0; r < PPL; ++r) {
    pkeep[r] = C.PA[r] & ~PD[r];
    pfree[r] = ~pkeep[r];
  }
  const uint64_t xm = ballot(needX), bm = ballot(normal);
  const int nX = popc(xm), nB = popc(bm);
  if (nX > popc_multi(vfree) || nB > popc_multi(pfree)) return CUT_CAPACITY;
  int nf = -1;
#pragma unroll
  for (int r = FPL - 1; r >= 0; --r)
    if (~C.FA[r]) nf = 64 * r + ctz64(~C.FA[r]);
  if (nf < 0) return CUT_CAPACITY;
  const int vslot = needX ? select_multi(vfree, mbcnt(xm, lane)) : 0;
  const int pslot = normal ? select_multi(pfree, mbcnt(bm, lane)) : 0;

  // 4. exit points: intersection(outside, inside), vector3.rs:213-219
  double X0 = 0, X1 = 0, X2 = 0;
  if (needX) {
    const double dO = ndot(S, Ok, pnx, pny, pnz), dQ = ndot(S, Q, pnx, pny, pnz);
    const double t = (off - dO) / (dQ - dO);
    const double ox = S.vx[Ok], oy = S.vy[Ok], oz = S.vz[Ok];
    X0 = ox + (S.vx[Q] - ox) * t;
    X1 = oy + (S.vy[Q] - oy) * t;
    X2 = oz + (S.vz[Q] - oz) * t;
  }
  wave_sync();
  if (needX) {
    S.vx[vslot] = X0;
    S.vy[vslot] = X1;
    S.vz[vslot] = X2;
    S.cs.g.newv_of[b] = (uint8_t)vslot;
  }
  wave_sync();
  const int Xid = needX ? vslot : Q;
  const int Eid = isrun ? ((cP == C_IN) ? (int)S.cs.g.newv_of[a ^ 1] : P) : 0;

  // 5. rewire: bridge on the run's face, its flip on the new face.
  int nfh = -1;
  if (normal) {
    const int br = 2 * pslot, brf = br + 1;
    S.he_tgt[br] = (uint8_t)Xid;
    S.he_face[br] = (uint8_t)fa;
    S.he_next[br] = (uint16_t)(needX ? b : nb);
    S.he_tgt[brf] = (uint8_t)Eid;
    S.he_face[brf] = (uint8_t)nf;
    if (cP == C_IN) {
      S.he_tgt[a] = (uint8_t)Eid;
      S.he_next[a] = (uint16_t)br;
    } else {
      S.he_next[pred] = (uint16_t)br;
    }
    S.f_start[fa] = (uint16_t)br;
    S.cs.g.exit_of[Xid] = (uint16_t)brf;
    nfh = brf;
  } else if (two) {
    S.he_face[nb] = (uint8_t)nf;
    S.cs.g.exit_of[Q] = (uint16_t)nb;
    nfh = nb;
  }
  wave_sync();
  bool lbad = false;
  if (nfh >= 0) {
    const int nx = S.cs.g.exit_of[Eid];
    if (nx == 0xffff)
      lbad = true;
    else
      S.he_next[nfh] = (uint16_t)nx;
  }
  if (ballot(lbad)) return CUT_DEGENERATE;
  const uint64_t hm = ballot(nfh >= 0);
  if (!hm) return CUT_DEGENERATE;
  const int nfstart = rl_i(nfh, ctz64(hm));

  // 6. live sets
  uint64_t vnew[VPL], pnew[PPL];
  lowest_bits(vfree, nX, vnew);
  lowest_bits(pfree, nB, pnew);
#pragma unroll
  for (int r = 0; r < VPL; ++r) C.VA[r] = vkeep[r] | vnew[r];
#pragma unroll
  for (int r = 0; r < PPL; ++r) C.PA[r] = pkeep[r] | pnew[r];
#pragma unroll
  for (int r = 0; r < FPL; ++r) S.fmark[lane + 64 * r] = 0;
  if (lane == 0) {
    S.f_tag[nf] = tag;
    S.f_seq[nf] = (uint16_t)C.seq;
    S.f_start[nf] = (uint16_t)nfstart;
  }
  C.seq++;
  wave_sync();
#pragma unroll
  for (int r = 0; r < PPL; ++r) {
    const int p = lane + 64 * r;
    if (lane_bit(C.PA[r])) {
      S.fmark[S.he_face[2 * p]] = 1;
      S.fmark[S.he_face[2 * p + 1]] = 1;
    }
  }
  wave_sync();
#pragma unroll
  for (int r = 0; r < FPL; ++r) C.FA[r] = ballot(S.fmark[lane + 64 * r] != 0);
  // refresh this lane's vertex registers
#pragma unroll
  for (int r = 0; r < VPL; ++r) {
    const int v = lane + 64 * r;
    if (lane_bit(C.VA[r])) {
      C.rx[r] = S.vx[v];
      C.ry[r] = S.vy[v];
      C.rz[r] = S.vz[v];
    }
  }
  *newv = nX;
  return CUT_OK;
}

template <int VPL, int PPL, int FPL>
__device__ double cell_r2(const CellState<VPL, PPL, FPL>& C, int lane) {
  double m = 0.0;
#pragma unroll
  for (int r = 0; r < VPL; ++r)
    if (lane_bit(C.VA[r])) {
      double d = (C.rx[r] * C.rx[r] + C.ry[r] * C.ry[r]) + C.rz[r] * C.rz[r];
      m = d > m ? d : m;
    }
  return wave_max_uni(m);
}

// cell_r2 and a vertex slot that attains it.
template <int VPL, int PPL, int FPL>
__device__ double cell_r2_slot(const CellState<VPL, PPL, FPL>& C, int lane, int* slot) {
  double d[VPL];
  double m = 0.0;
#pragma unroll
  for (int r = 0; r < VPL; ++r) {
    d[r] = (C.rx[r] * C.rx[r] + C.ry[r] * C.ry[r]) + C.rz[r] * C.rz[r];
    if (lane_bit(C.VA[r])) m = d[r] > m ? d[r] : m;
  }
  m = wave_max_uni(m);
  int s = -1;
#pragma unroll
  for (int r = VPL - 1; r >= 0; --r) {
    const uint64_t b = ballot(lane_bit(C.VA[r]) && d[r] == m);
    if (b) s = 64 * r + ctz64(b);
  }
  *slot = s;
  return m;
}

}  // namespace

// One cell: the site at grid position gpos of the set A.  Returns the
// cell's status bits (wave-uniform); the caller queues capacity reruns.
// GEO: 0 the cell outputs, 1 the staged geometry only, 2 both.
// WIDE: the operation count accumulates in one 64-bit register pair instead
// of 32-bit scalar parts (the same count; which form compiles to the faster
// kernel differs between the plain and the batched launch).  WIN: stream
// slices per chunk (2 for the plain launch, 4 for the batched one).
template <int VPL, int PPL, int FPL, int GEO = 0, bool WIDE = false, int WIN = 2>
__device__ __forceinline__ uint32_t cell_body(const CellArgs& A, const int gpos, Lds<VPL, PPL, FPL>& S,
                                              const int lane, const int slot = 0) {
  using L = Lds<VPL, PPL, FPL>;
  constexpr int FC = L::FC;
  static_assert(WIN >= 1 && WIN <= kWinMax, "the window's heads fit the staging space");
  CellState<VPL, PPL, FPL> C;
  // a site of the set excludes itself by index (interface.rs:299-305); an
  // outside site (interface.rs:211-232) excludes nothing
  const bool ext = A.ext != nullptr;
  const uint32_t self = (uint32_t)uni_i((int)(ext ? 0xffffffffu : A.sorted[gpos]));
  // the site and the grid are wave-uniform: scalar registers (the grid read
  // through the constant address space), not vector ones
  const double sx = uni_d(ext ? A.ext[3 * gpos] : A.spx[gpos]);
  const double sy = uni_d(ext ? A.ext[3 * gpos + 1] : A.spy[gpos]);
  const double sz = uni_d(ext ? A.ext[3 * gpos + 2] : A.spz[gpos]);
  typedef const __attribute__((address_space(4))) GridDev ConstGrid;
  const GridDev& G = *(const GridDev*)(ConstGrid*)A.grid;
  const int cpd = G.cpd;
  const int ix = uni_i(axis_index(sx, G.xmin, G.xmax, G.ix, cpd));
  const int iy = uni_i(axis_index(sy, G.ymin, G.ymax, G.iy, cpd));
  const int iz = uni_i(axis_index(sz, G.zmin, G.zmax, G.iz, cpd));

  // ---- container cube translated by -site (polyhedron.rs:226-392, :859-867)
  {
    const double x0 = A.lo[0] + (-sx), x1 = A.hi[0] + (-sx);
    const double y0 = A.lo[1] + (-sy), y1 = A.hi[1] + (-sy);
    const double z0 = A.lo[2] + (-sz), z1 = A.hi[2] + (-sz);
#pragma unroll
    for (int r = 0; r < VPL; ++r) {
      const int v = lane + 64 * r;
      C.rx[r] = 0.0;  // slots without a vertex hold zeros (tests on them are masked)
      C.ry[r] = 0.0;
      C.rz[r] = 0.0;
      if (v < 8) {
        const double vx = (v == 1 || v == 2 || v == 5 || v == 6) ? x1 : x0;
        const double vy = (v >= 4) ? y1 : y0;
        const double vz = (v == 2 || v == 3 || v == 6 || v == 7) ? z1 : z0;
        S.vx[v] = vx;
        S.vy[v] = vy;
        S.vz[v] = vz;
        C.rx[r] = vx;
        C.ry[r] = vy;
        C.rz[r] = vz;
      }
    }
    if (lane < 24) {
      S.he_tgt[lane] = kCubeTgt[lane];
      S.he_next[lane] = kCubeNext[lane];
      S.he_face[lane] = kCubeFace[lane];
    }
    if (lane < 6) {
      S.f_tag[lane] = kWallTag[lane];
      S.f_start[lane] = kCubeStart[lane];
      S.f_seq[lane] = (uint16_t)lane;
    }
#pragma unroll
    for (int r = 0; r < VPL; ++r) C.VA[r] = 0;
#pragma unroll
    for (int r = 0; r < PPL; ++r) C.PA[r] = 0;
#pragma unroll
    for (int r = 0; r < FPL; ++r) C.FA[r] = 0;
    C.VA[0] = 0xffull;
    C.PA[0] = 0xfffull;
    C.FA[0] = 0x3full;
    C.seq = 6;
  }
  wave_sync();
  int r2_slot = 0;
  double r2 = cell_r2_slot(C, lane, &r2_slot);

  uint32_t streamed = 0, classified = 0, ncuts = 0, ninc = 0, status = 0;
  // the operation count in 32-bit scalar parts, combined at the end:
  // 13 + 6V per classification (V the vertex count it was tested against),
  // 12 per new vertex + 5V per cut
  uint32_t cls_base = 0;  // classifications already in f_nv
  uint32_t f_nv = 0, f_cut = 0;
  uint64_t f_wide = 0;  // WIDE: 13 + 6V per classification + 12 new + 5V per cut
  int maxv = 8;
  int e_lastcut = -1;
  const bool skip_cells = A.target_group < 0;
  int e_stop_final = 0;
  bool done = false;
  const double rlim = A.radius_limit;

  for (int ebase = 0; !done; ebase += 64) {
    if (ebase >= A.tlen) {
      // table exhausted: every later entry has key >= kcut, so the stream
      // has ended when kcut is past the security threshold or the radius
      // limit; otherwise the cell needs a longer table
      if (!(A.kcut > stop_thr(r2)) && !(A.kcut > rlim)) status |= TESS_CELL_TABLE;
      e_stop_final = A.tlen;
      break;
    }
    const int e = ebase + lane;
    const bool have = e < A.tlen;
    double key = 0.0;
    uint32_t bst = 0;
    int cnt = 0;
    float ebd2 = 0.0f;  // lower bound of |p - site|^2 over the grid cell's points
    if (have) {
      key = A.tkey[e];
      int ox, oy, oz;
      unpack_off(A.toff[e], ox, oy, oz);
      const int bx = ix + ox, by = iy + oy, bz = iz + oz;
      if (bx >= 0 && bx < cpd && by >= 0 && by < cpd && bz >= 0 && bz < cpd) {
        const int bin = (bx * cpd + by) * cpd + bz;
        bst = A.delim[bin];
        cnt = (int)(A.delim[bin + 1] - bst);
        const double gx = box_gap(bx, G.xmin, G.sx, sx), gy = box_gap(by, G.ymin, G.sy, sy),
                     gz = box_gap(bz, G.zmin, G.sz, sz);
        ebd2 = __double2float_rd((gx * gx + gy * gy) + gz * gz);
      }
    }
    const uint64_t sm = ballot(have && (key > stop_thr(r2) || key > rlim));
    const uint64_t em = ballot(!have);
    const int nE = sm ? ctz64(sm) : 64;
    const int nEnd = em ? ctz64(em) : 64;
    const int lim = nE < nEnd ? nE : nEnd;
    int e_stop = ebase + lim;  // entries [ebase, e_stop) are reached
    bool stop_here = (sm != 0) && nE <= nEnd;
    // The chunks walk the candidates of the reached entries in stream order
    // from a cursor (entry ce, offset co).  Before each chunk, grid cells
    // that the security ball no longer reaches (lower bound above the stop
    // threshold) are left out: every candidate in them would be skipped at
    // its turn, so the result and the counters are unchanged.  With a group
    // filter the streamed count needs each candidate's group, so nothing is
    // left out.
    int ce = 0, co = 0;
    for (;;) {
      if (done) break;
      const double thr_c = stop_thr(r2);
      int c = 0;
      if (lane >= ce && ebase + lane < e_stop) {
        c = cnt - (lane == ce ? co : 0);
        if (skip_cells && (double)ebd2 > thr_c) c = 0;
      }
      const int incl = wave_incl_scan(c, lane);
      const int total = rl_i(incl, 63);
      if (total == 0) break;
      const int pre = incl - c;
      S.cs.st.bpref[lane] = (uint32_t)pre;
      S.cs.st.bstart[lane] = bst + (uint32_t)(lane == ce ? co : 0);
      // The window: stream positions [0, min(total, 64 WIN)), lane + 64 j
      // in slice j.  Entry of position k: the last non-empty entry starting
      // at or before k (heads by position, then a running maximum carried
      // across the slices).
      const int nsl = total < 64 * WIN ? (total + 63) >> 6 : WIN;
#pragma unroll
      for (int j = 0; j < WIN; ++j) S.u.head[lane + 64 * j] = 0xff;
      wave_sync();
      if (c > 0 && pre < 64 * WIN) S.u.head[pre] = (uint8_t)lane;
      wave_sync();
      int ekw[WIN];
      {
        int carry = -1;
#pragma unroll
        for (int j = 0; j < WIN; ++j) {
          ekw[j] = 0;
          if (j >= nsl) continue;
          const int hk = S.u.head[lane + 64 * j];
          int e = (int)wave_max_scan_u32(hk == 0xff ? 0u : (uint32_t)hk + 1u) - 1;
          e = e > carry ? e : carry;
          ekw[j] = e;
          carry = rl_i(e, 63);
        }
      }
      wave_sync();  // the heads share space with the staged candidates
      // Stage the kept candidates in stream order, up to 64: the chunk ends
      // after the 64th (positions past it are streamed again by the next
      // chunk).  The sequential pass reads them back as wave-uniform values.
      int consumed = total < 64 * WIN ? total : 64 * WIN;
      int nkeep = 0;
      uint64_t vmw[WIN];  // valid positions of each slice (group-filter streamed count)
#pragma unroll
      for (int j = 0; j < WIN; ++j) {
        vmw[j] = 0;
        if (64 * j >= consumed) continue;
        const int k = lane + 64 * j;
        const int ek = ekw[j];
        bool valid = false;
        uint32_t jidx = 0;
        double px = 0, py = 0, pz = 0, d2 = 0;
        if (k < consumed) {
          const uint32_t s = S.cs.st.bstart[ek] + (uint32_t)(k - (int)S.cs.st.bpref[ek]);
          jidx = A.sorted[s];
          valid = (jidx != self) && (A.target_group < 0 || A.groups[jidx] == A.target_group);
          px = A.spx[s] - sx;
          py = A.spy[s] - sy;
          pz = A.spz[s] - sz;
          d2 = (px * px + py * py) + pz * pz;
        }
        // every position of the window lies in an entry before the stop one
        const bool keep = valid && !(d2 > stop_thr(r2));
        const uint64_t km = ballot(keep);
        const int q = nkeep + mbcnt(km, lane);
        if (keep && q < 64) {
          S.u.cand.p[q][0] = px;
          S.u.cand.p[q][1] = py;
          S.u.cand.p[q][2] = pz;
          S.u.cand.tag[q] = (int32_t)jidx;
          S.u.cand.ent[q] = ebase + ek;
        }
        vmw[j] = ballot(valid);
        const int kc = popc(km);
        if (nkeep + kc > 64) {
          consumed = 64 * j + select_bit(km, 63 - nkeep) + 1;  // just past the 64th
          nkeep = 64;
        } else {
          nkeep += kc;
          if (nkeep == 64 && 64 * (j + 1) < consumed) consumed = 64 * (j + 1);
        }
      }
      wave_sync();
      // bisector planes, lane = staged candidate (halfway_from_origin_to,
      // vector3.rs:223-239)
      // (kept in this lane's registers: the sequential pass reads candidate
      // q's plane from lane q)
      double cnx0 = 0, cny0 = 0, cnz0 = 0, coff0 = 0;
      if (lane < nkeep) {
        const double px = S.u.cand.p[lane][0], py = S.u.cand.p[lane][1], pz = S.u.cand.p[lane][2];
        const double cd2 = (px * px + py * py) + pz * pz;  // the staging's d2, bit for bit
        const double mag = sqrt(cd2);
        const double inv = 1.0 / mag;
        cnx0 = px * inv;
        cny0 = py * inv;
        cnz0 = pz * inv;
        coff0 = (cnx0 * (px * 0.5) + cny0 * (py * 0.5)) + cnz0 * (pz * 0.5);
      }
      wave_sync();
      const uint64_t all_kept = nkeep >= 64 ? ~0ull : ((1ull << nkeep) - 1ull);
      // Pre-test, lane = candidate: which staged candidates cut the cell as it
      // is now.  One that does not can never cut (the cell only shrinks), so
      // only the others need the sequential test; the ones skipped are still
      // counted as classified at their turn, exactly as the sequential pass
      // would count them.  Same arithmetic as the sequential test.
      uint64_t mc = all_kept;
      if (nkeep >= kPreMin) {
        bool cut = false;
        if (lane < nkeep) {
#pragma unroll
          for (int r = 0; r < VPL; ++r) {
            uint64_t m = C.VA[r];
            while (m) {
              const int v = 64 * r + ctz64(m);
              m &= m - 1ull;
              const double d = (cnx0 * S.vx[v] + cny0 * S.vy[v]) + cnz0 * S.vz[v];
              cut = cut || (d - coff0 > kTol);
            }
          }
        }
        mc = ballot(cut);
      }
      uint64_t pv = all_kept;  // pending and still within the stop entry / radius
      int nv = popc_multi(C.VA);
      for (;;) {
        const uint64_t next = pv & mc;
        const uint64_t before = pv & (next - 1ull) & ~next;  // pending below the next cutter (all if none)
        const int nb = popc(before);
        classified += (uint32_t)nb;
        pv &= ~before;
        if (!next) break;
        const int q = ctz64(next);
        pv &= pv - 1ull;
        const double cnx = rl_d(cnx0, q), cny = rl_d(cny0, q);
        const double cnz = rl_d(cnz0, q), coff = rl_d(coff0, q);
        classified++;
        // Outside test only; the three-way classification is needed only
        // when the candidate cuts (same arithmetic both times)
        uint64_t OM[VPL];
        double dv[VPL];
        bool any = false;
#pragma unroll
        for (int r = 0; r < VPL; ++r) {
          if (r > 0 && C.VA[r] == 0) {  // vertex slots fill from the bottom
            dv[r] = 0.0;
            OM[r] = 0;
            continue;
          }
          dv[r] = (cnx * C.rx[r] + cny * C.ry[r]) + cnz * C.rz[r];
          OM[r] = ballot(dv[r] - coff > kTol) & C.VA[r];
          any = any || OM[r] != 0;
        }
        if (!any) continue;
#pragma unroll
        for (int r = 0; r < VPL; ++r)
          if (lane_bit(C.VA[r])) {
            const double sd = dv[r] - coff;
            S.vcls[lane + 64 * r] = sd > kTol ? C_OUT : (sd < -kTol ? C_IN : C_INC);
          }
        wave_sync();
        int newv = 0;
        const int32_t ctag = uni_i(S.u.cand.tag[q]);
        const int ec = uni_i(S.u.cand.ent[q]);
        uint64_t kept[VPL];  // vertices the cut keeps
        bool max_cut = false;
#pragma unroll
        for (int r = 0; r < VPL; ++r) {
          kept[r] = C.VA[r] & ~OM[r];
          if ((r2_slot >> 6) == r) max_cut = (OM[r] >> (r2_slot & 63)) & 1ull;
        }
        int rc = clip_fast(S, C, OM, cnx, cny, cnz, coff, ctag, lane, &newv);
        if (rc == CUT_GENERAL) rc = clip(S, C, OM, cnx, cny, cnz, coff, ctag, lane, &newv);
        if (rc != CUT_OK) {
          status |= (rc == CUT_CAPACITY) ? TESS_CELL_CAPACITY : TESS_CELL_DEGENERATE;
          done = true;
          break;
        }
        ncuts++;
        ninc += newv == 0;  // no Inside-Outside edge: only Incident vertices border the cut
        if (WIDE)
          f_wide += (uint64_t)(classified - cls_base) * (13ull + 6ull * (uint64_t)nv);
        else
          f_nv += (classified - cls_base) * (uint32_t)nv;
        cls_base = classified;
        nv = popc_multi(C.VA);
        if (nv > maxv) maxv = nv;
        if (WIDE)
          f_wide += 12ull * (uint64_t)newv + 5ull * (uint64_t)nv;
        else
          f_cut += 12u * (uint32_t)newv + 5u * (uint32_t)nv;
        // the largest |v|^2 changes only if its vertex was cut off or a new
        // vertex exceeds it (new vertices lie on old edges, but rounding
        // may put one an ulp beyond)
        bool r2_same = !max_cut;
        if (r2_same) {
          uint64_t big = 0;
#pragma unroll
          for (int r = 0; r < VPL; ++r) {
            const bool isnew = lane_bit(C.VA[r] & ~kept[r]);
            const double d = (C.rx[r] * C.rx[r] + C.ry[r] * C.ry[r]) + C.rz[r] * C.rz[r];
            big |= ballot(isnew && d > r2);
          }
          r2_same = big == 0;
        }
        e_lastcut = ec;
        // With an unchanged radius the stop entry and the pending set stay as
        // they are: the entries between the last cut and the stop entry all
        // have keys within the unchanged threshold.
        if (!r2_same) {
          r2 = cell_r2_slot(C, lane, &r2_slot);
          const double thr = stop_thr(r2);
          const uint64_t sm2 = ballot(lane < lim && ebase + lane > e_lastcut && (key > thr || key > rlim));
          if (sm2) {
            e_stop = ebase + ctz64(sm2);
            stop_here = true;
          }
          // the smaller cell reaches fewer of the pending candidates
          bool reach = false;
          if (lane < nkeep && S.u.cand.ent[lane] < e_stop) {
            const double px = S.u.cand.p[lane][0], py = S.u.cand.p[lane][1], pz = S.u.cand.p[lane][2];
            reach = !(((px * px + py * py) + pz * pz) > thr);
          }
          pv &= ballot(reach);
        }
      }
      if (WIDE)
        f_wide += (uint64_t)(classified - cls_base) * (13ull + 6ull * (uint64_t)nv);
      else
        f_nv += (classified - cls_base) * (uint32_t)nv;
      cls_base = classified;
      if (!skip_cells) {
        // the valid positions consumed that lie before the stop entry (a
        // prefix of the window: positions follow the entries)
        const int es = e_stop - ebase;
        const int pe = es >= 64 ? consumed : rl_i(pre, es);
        const int lim_k = pe < consumed ? pe : consumed;
#pragma unroll
        for (int j = 0; j < WIN; ++j) {
          const int lk = lim_k - 64 * j;
          if (lk <= 0) break;
          streamed += (uint32_t)popc(lk >= 64 ? vmw[j] : (vmw[j] & ((1ull << lk) - 1ull)));
        }
      }
      // advance the cursor past the positions this chunk consumed
      if (consumed >= total) {
        ce = 64;
        co = 0;
      } else {
        // the entry holding position `consumed`: the last non-empty one
        // starting at or before it
        const uint64_t hm = ballot(c > 0 && pre <= consumed);
        const int lo = 63 - __clzll((long long)hm);
        co = consumed - rl_i(pre, lo) + (lo == ce ? co : 0);
        ce = lo;
      }
      wave_sync();
    }
    if (skip_cells) {
      // streamed: every candidate of the reached entries but the site itself
      const int rc = (ebase + lane < e_stop && lane < lim) ? cnt : 0;
      const int sc = rl_i(wave_incl_scan(rc, lane), 63);
      streamed += (uint32_t)(sc - (ebase == 0 && !ext ? 1 : 0));
    }
    wave_sync();
    if (stop_here || done) {
      e_stop_final = e_stop;
      done = true;
    }
  }
  const uint64_t flops = WIDE ? 64ull + 8ull * streamed + f_wide
                              : 64ull + 8ull * streamed + 13ull * classified + 6ull * f_nv + f_cut;

  // ---- outputs
  const uint32_t oi = ext ? (uint32_t)gpos : self;
  int nfaces = popc_multi(C.FA);
  const bool ok = (status & (TESS_CELL_CAPACITY | TESS_CELL_DEGENERATE)) == 0;
  if (ok) {
    const int nv = popc_multi(C.VA), ne = popc_multi(C.PA);
    if (nv - ne + nfaces != 2) status |= TESS_CELL_DEGENERATE;
  }
  // per-cell face slots hold kFaceStage faces; the arena has no such limit
  if (nfaces > kFaceStage && (GEO != 0 || A.fpos == nullptr)) status |= TESS_CELL_CAPACITY;
  // a cell that reached the end of the table is recomputed with a longer
  // one when the caller collects them: it writes nothing now
  const bool table_rerun = (status & TESS_CELL_TABLE) && A.defer_table;
  // where the faces go: an arena range reserved by one atomic, or the
  // cell's own slots
  uint64_t fbase = (uint64_t)oi * kFaceStage;
  if (GEO != 1 && A.fpos != nullptr && !table_rerun && (status & (TESS_CELL_CAPACITY | TESS_CELL_DEGENERATE)) == 0) {
    fbase = arena_reserve(A.arena_next, A.arena_cap, (uint32_t)nfaces, lane);
    if (fbase == ~0ull)
      status |= TESS_CELL_ARENA;  // arena full (counted; the caller reports it)
    else if (lane == 0)
      A.fpos[oi] = fbase;
  }
  uint32_t oflops = 0;  // this lane's faces
  if ((status & (TESS_CELL_CAPACITY | TESS_CELL_DEGENERATE | TESS_CELL_ARENA)) == 0 && !table_rerun) {
    double area[FPL], contrib[FPL];
    int32_t tagv[FPL];
    int seqv[FPL];
    bool fal[FPL];
    int hanc[FPL], ringk[FPL];  // geometry export: anchor half-edge, ring length
#pragma unroll
    for (int r = 0; r < FPL; ++r) {
      const int f = lane + 64 * r;
      fal[r] = lane_bit(C.FA[r]);
      area[r] = 0;
      contrib[r] = 0;
      tagv[r] = 0;
      seqv[r] = 0;
      hanc[r] = 0;
      ringk[r] = 0;
      if (fal[r]) {
        const int st = S.f_start[f];
        int h = st, ha = st, k = 0;
        int va = S.he_tgt[st];
        double ax = S.vx[va], ay = S.vy[va], az = S.vz[va];
        do {
          const int v = S.he_tgt[h];
          const double x = S.vx[v], y = S.vy[v], z = S.vz[v];
          if (lex_less(x, y, z, ax, ay, az)) {
            ax = x;
            ay = y;
            az = z;
            ha = h;
          }
          h = S.he_next[h];
          ++k;
        } while (h != st && k <= 255);
        // fan from the anchor: sum cross(V[j-1]-A, V[j]-A), j = 2..k-1
        double wx = 0.0, wy = 0.0, wz = 0.0;
        h = S.he_next[ha];
        int v = S.he_tgt[h];
        double cx = S.vx[v] - ax, cy = S.vy[v] - ay, cz = S.vz[v] - az;
        for (int j = 2; j < k; ++j) {
          h = S.he_next[h];
          v = S.he_tgt[h];
          const double qx = cx, qy = cy, qz = cz;
          cx = S.vx[v] - ax;
          cy = S.vy[v] - ay;
          cz = S.vz[v] - az;
          wx = wx + (qy * cz - qz * cy);
          wy = wy + (qz * cx - qx * cz);
          wz = wz + (qx * cy - qy * cx);
        }
        contrib[r] = (ax * wx + ay * wy) + az * wz;
        area[r] = 0.5 * sqrt((wx * wx + wy * wy) + wz * wz);
        tagv[r] = S.f_tag[f];
        seqv[r] = S.f_seq[f];
        hanc[r] = ha;
        ringk[r] = k;
        oflops += 3u * (k - 1) + 12u * (k - 2) + 13u;
      }
    }
    // rank faces by creation number (volume order) and by tag (output order)
    int rs[FPL], rt[FPL];
#pragma unroll
    for (int r = 0; r < FPL; ++r) {
      rs[r] = 0;
      rt[r] = 0;
    }
#pragma unroll
    for (int q = 0; q < FPL; ++q) {
      uint64_t m = C.FA[q];
      while (m) {
        const int g = 64 * q + ctz64(m);
        m &= m - 1;
        const int gs = S.f_seq[g];
        const int32_t gt = S.f_tag[g];
#pragma unroll
        for (int r = 0; r < FPL; ++r) {
          rs[r] += (gs < seqv[r]);
          rt[r] += (gt < tagv[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < FPL; ++r)
      if (fal[r]) S.u.fcontrib[rs[r]] = contrib[r];
    wave_sync();
    // (the geometry passes write only the geometry)
    if (GEO != 1 && lane == 0) {
      double vol = 0.0;
      for (int f = 0; f < nfaces; ++f) vol = vol + S.u.fcontrib[f];
      A.volume[oi] = vol / 6.0;
      A.fcount[oi] = (uint32_t)nfaces;
    }
    if (GEO != 1) {
#pragma unroll
      for (int r = 0; r < FPL; ++r)
        if (fal[r]) {
          A.snbr[fbase + rt[r]] = tagv[r];
          A.sarea[fbase + rt[r]] = area[r];
        }
    }
    if (A.exp_v) {
      // geometry export for a single-cell query: slots and rings in tag order
#pragma unroll
      for (int r = 0; r < VPL; ++r) {
        const int v = lane + 64 * r;
        A.exp_v[3 * v] = S.vx[v];
        A.exp_v[3 * v + 1] = S.vy[v];
        A.exp_v[3 * v + 2] = S.vz[v];
        A.exp_meta[2 + v] = (int)lane_bit(C.VA[r]);
      }
#pragma unroll
      for (int r = 0; r < FPL; ++r)
        if (fal[r]) {
          const int st = S.f_start[lane + 64 * r];
          int h = st, k = 0;
          do {
            A.exp_ring[rt[r] * 64 + k] = S.he_tgt[h];
            h = S.he_next[h];
            ++k;
          } while (h != st && k < 63);
          A.exp_ring[rt[r] * 64 + k] = -1;
        }
      if (lane == 0) {
        A.exp_meta[0] = L::VC;
        A.exp_meta[1] = nfaces;
      }
    }
    if constexpr (GEO != 0) {
      // Geometry of whole sets (compute_vertices / compute_face_vertices,
      // polyhedron.rs:885-919), staged at this work item's slot: the live
      // vertices in slot order, and per face in output order the ring of
      // local vertex numbers from the anchor (the lexicographically smallest
      // vertex) along the face's orientation, with the rings' offsets.
      int below = 0;  // live slots in the lower registers
      double* sv = A.geo_v + (size_t)slot * (3 * L::VC);
#pragma unroll
      for (int r = 0; r < VPL; ++r) {
        if (lane_bit(C.VA[r])) {
          const int q = below + mbcnt(C.VA[r], lane);
          const int v = lane + 64 * r;
          sv[3 * q] = S.vx[v];
          sv[3 * q + 1] = S.vy[v];
          sv[3 * q + 2] = S.vz[v];
        }
        below += popc(C.VA[r]);
      }
      // ring offsets in output order: lengths by rank, then a wave scan
      wave_sync();
#pragma unroll
      for (int r = 0; r < FPL; ++r)
        if (fal[r]) S.cs.st.bpref[rt[r]] = (uint32_t)ringk[r];
      wave_sync();
      const int kk = lane < nfaces ? (int)S.cs.st.bpref[lane] : 0;
      const int incl = wave_incl_scan(kk, lane);
      const int ex = incl - kk;
      if (lane < nfaces) A.geo_roff[(size_t)slot * kFaceStage + lane] = (uint16_t)ex;
      S.cs.st.bstart[lane] = (uint32_t)ex;
      wave_sync();
      uint16_t* sr = A.geo_ring + (size_t)slot * L::HC;
#pragma unroll
      for (int r = 0; r < FPL; ++r)
        if (fal[r]) {
          const int pos = (int)S.cs.st.bstart[rt[r]];
          int h = hanc[r], j = 0;
          do {
            const int v = S.he_tgt[h];
            const int vr = v >> 6;
            const uint64_t lower = (1ull << (v & 63)) - 1ull;
            int rank = 0;  // live slots below v
#pragma unroll
            for (int q = 0; q < VPL; ++q)
              rank += q < vr ? popc(C.VA[q]) : (q == vr ? popc(C.VA[q] & lower) : 0);
            sr[pos + j] = (uint16_t)rank;
            h = S.he_next[h];
            ++j;
          } while (h != hanc[r] && j < ringk[r]);
        }
      const int nring = rl_i(incl, 63);
      if (lane == 0) {
        A.geo_count[oi] = (uint32_t)below;
        A.geo_count[(size_t)A.n + oi] = (uint32_t)nring;
        A.geo_slot[slot] = oi + 1u;
      }
    }
  } else {
    nfaces = 0;
  }
  const uint64_t oflops_all = (uint32_t)rl_i(wave_incl_scan((int)oflops, lane), 63);
  if (lane == 0) {
    if (GEO != 1) {
      if (status & (TESS_CELL_CAPACITY | TESS_CELL_DEGENERATE | TESS_CELL_ARENA)) {
        A.volume[oi] = 0.0;
        A.fcount[oi] = 0;
      }
      A.status[oi] = status;
    }
    if (A.summary) {
      if (!(status & (TESS_CELL_CAPACITY | TESS_CELL_DEGENERATE)) && !table_rerun) {
        atomicAdd(&A.summary->flops, (unsigned long long)(flops + oflops_all + 1));
        atomicAdd(&A.summary->candidates, (unsigned long long)streamed);
        if (ninc) atomicAdd(&A.summary->incident_cuts, (unsigned long long)ninc);
      }
      if (status & TESS_CELL_DEGENERATE) atomicAdd(&A.summary->degenerate, 1u);
      if (status & TESS_CELL_CAPACITY) atomicAdd(&A.summary->capacity, 1u);
      if (status & TESS_CELL_TABLE) atomicAdd(&A.summary->table, 1u);
    }
    if (GEO != 1 && A.stats) {
      tess_cell_stats st;
      st.streamed = streamed;
      st.classified = classified;
      st.cuts = ncuts;
      st.entries = (uint32_t)e_stop_final;
      st.flops = flops + oflops_all + 1;
      st.r2 = r2;
      st.max_vertices = (uint32_t)maxv;
      st.status = status;
      st.incident_cuts = ninc;
      st.reserved = 0;
      A.stats[oi] = st;
    }
  }
  (void)FC;
  return status;
}

// Wavefronts per SIMD the register allocation aims at: six for the
// capacity-64 kernels (5,760 B of LDS each, up to 28 per CU), five for the larger
// ones (their LDS caps them lower).
template <int VPL> constexpr int kOcc = VPL == 1 ? 6 : 5;
template <int VPL> constexpr int kOccItems = VPL == 1 ? 6 : 5;

// One cell per wavefront over a set's grid positions (all, or a work list).
template <int VPL, int PPL, int FPL, int GEO = 0>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kOcc<VPL>))) void k_cells(CellArgs A) {
  __shared__ Lds<VPL, PPL, FPL> S;
  const int lane = threadIdx.x;
  const int w = A.xcd_chunk ? (int)xcd_map(blockIdx.x, (uint32_t)A.xcd_chunk) : (int)blockIdx.x;
  if (w >= A.nwork) return;
  const int gpos = A.work ? (int)A.work[w] : w;
  const uint32_t status = cell_body<VPL, PPL, FPL, GEO, true>(A, gpos, S, lane, w);
  if (lane == 0 && (status & TESS_CELL_CAPACITY) && A.redo) {
    const uint32_t slot = atomicAdd(A.nredo, 1u);
    A.redo[slot] = (uint32_t)gpos;
  }
  if (lane == 0 && (status & TESS_CELL_TABLE) && A.tredo) {
    const uint32_t slot = atomicAdd(A.ntredo, 1u);
    A.tredo[slot] = (uint32_t)gpos;
  }
}

// Cells of many sets in one launch.  An item is (set << 32) | grid position;
// the set's arguments are read through the constant address space, as
// scalar loads the compiler may repeat instead of keeping live.
template <int VPL, int PPL, int FPL>
__device__ __forceinline__ void batch_item(const BatchArgs& B, uint64_t it, Lds<VPL, PPL, FPL>& S, int lane) {
  typedef const __attribute__((address_space(4))) CellArgs ConstArgs;
  ConstArgs* cs = (ConstArgs*)(B.sets) + (uint32_t)(it >> 32);
  const CellArgs& A = *(const CellArgs*)cs;
  const uint32_t status = cell_body<VPL, PPL, FPL, 0, false, 4>(A, (int)(uint32_t)it, S, lane);
  if (lane == 0 && (status & TESS_CELL_CAPACITY) && B.redo) {
    const uint32_t slot = atomicAdd(B.nredo, 1u);
    B.redo[slot] = it;
  }
  if (lane == 0 && (status & TESS_CELL_TABLE) && B.tredo) {
    const uint32_t slot = atomicAdd(B.ntredo, 1u);
    B.tredo[slot] = it;
  }
}

// One wavefront per item of a list whose length the host knows (the grid):
// the hardware dispatcher hands out the items in list order as wavefront
// slots free up, which is the work queue of the pass.
template <int VPL, int PPL, int FPL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kOccItems<VPL>))) void k_cells_items(BatchArgs B, uint32_t n) {
  __shared__ Lds<VPL, PPL, FPL> S;
  const uint32_t i = B.xcd_chunk ? xcd_map(blockIdx.x, (uint32_t)B.xcd_chunk) : blockIdx.x;
  if (i >= n) return;
  batch_item(B, B.items[B.order ? B.order[i] : i], S, threadIdx.x);
}

// Persistent form for a list whose length is only known on the device
// (*B.count): every wavefront takes `chunk` items at a time from a shared
// counter until the list is exhausted, so the loop ends for every wavefront.
template <int VPL, int PPL, int FPL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(VPL == 2 ? 4 : 3))) void k_cells_queue(BatchArgs B) {
  __shared__ Lds<VPL, PPL, FPL> S;
  const int lane = threadIdx.x;
  const uint32_t n = *B.count;
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(B.counter, (uint32_t)B.chunk);
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
    if (base >= n) break;
    const uint32_t end = base + (uint32_t)B.chunk < n ? base + (uint32_t)B.chunk : n;
    for (uint32_t i = base; i < end; ++i) {
      wave_sync();
      batch_item(B, B.items[B.order ? B.order[i] : i], S, lane);
    }
  }
}

// ---------------------------------------------------------------- spill
// Cells beyond the LDS kernels (more than 256 vertices or 128 faces, e.g. a
// site ringed by hundreds of points).  One wavefront per cell; the
// polyhedron is its list of face rings (site-relative vertex coordinates,
// counter-clockwise seen from outside) in creation order, in a global-memory
// workspace that the host grows on demand.  A cut classifies every ring
// entry (vector3.rs:170-193), clips each ring on its own -- an Outside run is
// replaced by its entry and exit points, a point on an Inside-Outside edge
// being intersection(outside, inside) (vector3.rs:213-219) and an Incident
// end kept as is -- and closes the new face from the runs' exit -> entry
// segments, tagged with the cutting point (polyhedron.rs:438-642 with the
// walk's defects fixed).  The arithmetic and the face order are those of the
// LDS kernels, so a cell both can build is bitwise the same.

__device__ __forceinline__ void ws_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

struct SpillWs {
  double* ring[2];    // [rmax][3]
  double* seg;        // [fmax][6]: exit point, entry point; the output phase
                      // keeps face volume terms and areas here
  uint32_t* foff[2];  // [fmax + 1]
  int32_t* ftag[2];   // [fmax]
  int32_t* snext;     // [fmax]
  uint32_t* fterm;    // [fmax][4]: kept length, segments, ring offset, segment offset
  uint8_t* loc;       // [rmax]
  uint8_t* used;      // [fmax]
};

__device__ __forceinline__ void spill_ws(const SpillArgs& P, int blk, SpillWs& W) {
  double* d = P.wsd + (size_t)blk * spill_wsd_doubles(P.rmax, P.fmax);
  uint32_t* u = P.wsu + (size_t)blk * spill_wsu_words(P.fmax);
  uint8_t* b = P.wsb + (size_t)blk * spill_wsb_bytes(P.rmax, P.fmax);
  W.ring[0] = d;
  W.ring[1] = d + 3ull * P.rmax;
  W.seg = d + 6ull * P.rmax;
  W.foff[0] = u;
  W.foff[1] = u + (P.fmax + 1);
  W.ftag[0] = (int32_t*)(u + 2 * (P.fmax + 1));
  W.ftag[1] = W.ftag[0] + P.fmax;
  W.snext = W.ftag[1] + P.fmax;
  W.fterm = (uint32_t*)(W.snext + P.fmax);
  W.loc = b;
  W.used = b + P.rmax;
}

// Clips the ring R[a, a + k) (classes L[a, a + k)).  Counts the kept ring
// entries (olen), the exit -> entry segments (nseg) and the crossing edges
// created on their exit side (nnew); writes the kept ring to out and the
// segments to segs when those are not null.
__device__ void spill_clip_ring(const double* R, const uint8_t* L, int a, int k, double nx, double ny, double nz,
                                double off, double* out, double* segs, int& olen, int& nseg, int& nnew) {
  olen = 0;
  nseg = 0;
  nnew = 0;
  int start = -1, nout = 0;
  for (int i = 0; i < k; ++i) {
    if (L[a + i] == C_OUT)
      ++nout;
    else if (start < 0)
      start = i;
  }
  if (nout == 0) {
    if (out)
      for (int i = 0; i < 3 * k; ++i) out[i] = R[3 * a + i];
    olen = k;
    return;
  }
  if (start < 0) return;  // the whole face is cut off
  auto emit = [&](double x, double y, double z) {
    if (out) {
      out[3 * olen] = x;
      out[3 * olen + 1] = y;
      out[3 * olen + 2] = z;
    }
    ++olen;
  };
  // intersection(outside o, inside q), vector3.rs:213-219
  auto cross_pt = [&](const double* o, const double* q, double& x, double& y, double& z) {
    const double dO = (nx * o[0] + ny * o[1]) + nz * o[2];
    const double dQ = (nx * q[0] + ny * q[1]) + nz * q[2];
    const double t = (off - dO) / (dQ - dO);
    x = o[0] + (q[0] - o[0]) * t;
    y = o[1] + (q[1] - o[1]) * t;
    z = o[2] + (q[2] - o[2]) * t;
  };
  double ex = 0.0, ey = 0.0, ez = 0.0;  // entry point of the current Outside run
  int i = start;
  for (int t = 0; t < k; ++t) {
    int j = i + 1;
    if (j == k) j = 0;
    const uint8_t ci = L[a + i], cj = L[a + j];
    const double* vi = R + 3 * (a + i);
    const double* vj = R + 3 * (a + j);
    if (ci != C_OUT) {
      emit(vi[0], vi[1], vi[2]);
      if (cj == C_OUT) {
        if (ci == C_IN) {
          cross_pt(vj, vi, ex, ey, ez);
          emit(ex, ey, ez);
        } else {
          ex = vi[0];
          ey = vi[1];
          ez = vi[2];
        }
      }
    } else if (cj != C_OUT) {
      double x = vj[0], y = vj[1], z = vj[2];
      if (cj == C_IN) {
        cross_pt(vi, vj, x, y, z);
        emit(x, y, z);
        ++nnew;
      }
      if (!(x == ex && y == ey && z == ez)) {
        if (segs) {
          double* g = segs + 6 * nseg;
          g[0] = x;
          g[1] = y;
          g[2] = z;
          g[3] = ex;
          g[4] = ey;
          g[5] = ez;
        }
        ++nseg;
      }
    }
    i = j;
  }
}

enum { SPILL_NOCUT = -1, SPILL_CAPACITY = -2, SPILL_DEGENERATE = -3 };

// One bisector against the soup in buffer b (nf faces, ne2 ring entries).
// Returns the new vertices (>= 0) when the cell changed -- the soup is then
// in buffer b ^ 1 with nf, ne2 updated -- or SPILL_*.
__device__ int spill_cut(SpillWs& W, int b, int& nf, int& ne2, int rmax, int fmax, double nx, double ny, double nz,
                         double off, int32_t tag, int lane) {
  const double* R = W.ring[b];
  bool anyo = false;
  for (int i = lane; i < ne2; i += 64) {
    const double sd = ((nx * R[3 * i] + ny * R[3 * i + 1]) + nz * R[3 * i + 2]) - off;
    const uint8_t c = sd > kTol ? C_OUT : (sd < -kTol ? C_IN : C_INC);
    W.loc[i] = c;
    anyo = anyo || c == C_OUT;
  }
  if (!ballot(anyo)) return SPILL_NOCUT;
  ws_sync();
  const uint32_t* FO = W.foff[b];
  // pass 1: per face the kept length (0 when fewer than 3 entries remain),
  // its segments and new vertices; exclusive scans give the offsets
  int tlen = 0, tseg = 0, tkept = 0, tnew = 0;
  for (int f0 = 0; f0 < nf; f0 += 64) {
    const int f = f0 + lane;
    int olen = 0, nseg = 0, nnew = 0;
    if (f < nf) {
      const int a = (int)FO[f], k = (int)(FO[f + 1] - FO[f]);
      spill_clip_ring(R, W.loc, a, k, nx, ny, nz, off, nullptr, nullptr, olen, nseg, nnew);
      if (olen < 3) olen = 0;
    }
    const int il = wave_incl_scan(olen, lane), is = wave_incl_scan(nseg, lane);
    const int ik = wave_incl_scan(olen > 0 ? 1 : 0, lane);
    if (f < nf) {
      W.fterm[4 * f] = (uint32_t)olen;
      W.fterm[4 * f + 1] = (uint32_t)(tkept + ik - (olen > 0 ? 1 : 0));
      W.fterm[4 * f + 2] = (uint32_t)(tlen + il - olen);
      W.fterm[4 * f + 3] = (uint32_t)(tseg + is - nseg);
    }
    tlen += rl_i(il, 63);
    tseg += rl_i(is, 63);
    tkept += rl_i(ik, 63);
    tnew += rl_i(wave_incl_scan(nnew, lane), 63);
  }
  const int m = tseg;
  if (tlen + m > rmax || tkept + 1 > fmax || m > fmax) return SPILL_CAPACITY;
  if (m < 3) return SPILL_DEGENERATE;
  ws_sync();
  // pass 2: kept rings and segments into place
  double* RN = W.ring[b ^ 1];
  uint32_t* FN = W.foff[b ^ 1];
  int32_t* TN = W.ftag[b ^ 1];
  for (int f = lane; f < nf; f += 64) {
    const int a = (int)FO[f], k = (int)(FO[f + 1] - FO[f]);
    const int olen = (int)W.fterm[4 * f], kidx = (int)W.fterm[4 * f + 1];
    const int roff = (int)W.fterm[4 * f + 2], soff = (int)W.fterm[4 * f + 3];
    int ol, ns, nn;
    spill_clip_ring(R, W.loc, a, k, nx, ny, nz, off, olen > 0 ? RN + 3 * roff : nullptr, W.seg + 6 * soff, ol, ns,
                    nn);
    if (olen > 0) {
      FN[kidx] = (uint32_t)roff;
      TN[kidx] = W.ftag[b][f];
    }
  }
  ws_sync();
  // the new face: segments chained head (exit point) to tail (entry point)
  for (int g = lane; g < m; g += 64) {
    const double e0 = W.seg[6 * g + 3], e1 = W.seg[6 * g + 4], e2 = W.seg[6 * g + 5];
    int nxt = -1;
    for (int h = 0; h < m; ++h)
      if (W.seg[6 * h] == e0 && W.seg[6 * h + 1] == e1 && W.seg[6 * h + 2] == e2) {
        nxt = h;
        break;
      }
    W.snext[g] = nxt;
    W.used[g] = 0;
  }
  ws_sync();
  int err = 0;
  if (lane == 0) {
    int cur = 0, pos = tlen;
    for (int c = 0; c < m; ++c) {
      W.used[cur] = 1;
      RN[3 * pos] = W.seg[6 * cur];
      RN[3 * pos + 1] = W.seg[6 * cur + 1];
      RN[3 * pos + 2] = W.seg[6 * cur + 2];
      ++pos;
      const int nxt = W.snext[cur];
      if (nxt < 0) {
        err = 1;
        break;
      }
      if (nxt == 0) {
        if (c != m - 1) err = 1;
        break;
      }
      if (W.used[nxt]) {
        err = 1;
        break;
      }
      cur = nxt;
    }
    FN[tkept] = (uint32_t)tlen;
    TN[tkept] = tag;
    FN[tkept + 1] = (uint32_t)(tlen + m);
  }
  err = uni_i(rl_i(err, 0));
  ws_sync();
  if (err) return SPILL_DEGENERATE;
  nf = tkept + 1;
  ne2 = tlen + m;
  return tnew;
}

__device__ double spill_r2(const double* R, int ne2, int lane) {
  double m = 0.0;
  for (int i = lane; i < ne2; i += 64) {
    const double d = (R[3 * i] * R[3 * i] + R[3 * i + 1] * R[3 * i + 1]) + R[3 * i + 2] * R[3 * i + 2];
    m = d > m ? d : m;
  }
  return wave_max_uni(m);
}

// The cell of work item gpos of set A (as cell_body) in the spill workspace.
__device__ void spill_cell(const CellArgs& A, const int gpos, const SpillArgs& P, SpillWs& W, uint64_t item,
                           int lane) {
  const bool ext = A.ext != nullptr;
  const uint32_t self = (uint32_t)uni_i((int)(ext ? 0xffffffffu : A.sorted[gpos]));
  // the site and the grid are wave-uniform: scalar registers (the grid read
  // through the constant address space), not vector ones
  const double sx = uni_d(ext ? A.ext[3 * gpos] : A.spx[gpos]);
  const double sy = uni_d(ext ? A.ext[3 * gpos + 1] : A.spy[gpos]);
  const double sz = uni_d(ext ? A.ext[3 * gpos + 2] : A.spz[gpos]);
  typedef const __attribute__((address_space(4))) GridDev ConstGrid;
  const GridDev& G = *(const GridDev*)(ConstGrid*)A.grid;
  const int cpd = G.cpd;
  const int ix = uni_i(axis_index(sx, G.xmin, G.xmax, G.ix, cpd));
  const int iy = uni_i(axis_index(sy, G.ymin, G.ymax, G.iy, cpd));
  const int iz = uni_i(axis_index(sz, G.zmin, G.zmax, G.iz, cpd));
  // container cube translated by -site: face f's ring from its start half-edge
  {
    const double x0 = A.lo[0] + (-sx), x1 = A.hi[0] + (-sx);
    const double y0 = A.lo[1] + (-sy), y1 = A.hi[1] + (-sy);
    const double z0 = A.lo[2] + (-sz), z1 = A.hi[2] + (-sz);
    if (lane < 6) {
      int h = kCubeStart[lane];
      for (int k = 0; k < 4; ++k) {
        const int v = kCubeTgt[h];
        double* o = W.ring[0] + 3 * (4 * lane + k);
        o[0] = (v == 1 || v == 2 || v == 5 || v == 6) ? x1 : x0;
        o[1] = (v >= 4) ? y1 : y0;
        o[2] = (v == 2 || v == 3 || v == 6 || v == 7) ? z1 : z0;
        h = kCubeNext[h];
      }
      W.foff[0][lane] = (uint32_t)(4 * lane);
      W.ftag[0][lane] = kWallTag[lane];
    }
    if (lane == 0) W.foff[0][6] = 24;
  }
  ws_sync();
  int b = 0, nf = 6, ne2 = 24;
  double r2 = spill_r2(W.ring[0], ne2, lane);
  uint32_t streamed = 0, classified = 0, ncuts = 0, ninc = 0, status = TESS_CELL_SPILL;
  uint64_t flops = 24 + 40;
  int maxv = 8;
  bool done = false;
  const double rlim = A.radius_limit;
  int e = 0;
  for (; e < A.tlen; ++e) {
    const double key = A.tkey[e];
    if (key > stop_thr(r2) || key > rlim) break;
    int ox, oy, oz;
    unpack_off(A.toff[e], ox, oy, oz);
    const int bx = ix + ox, by = iy + oy, bz = iz + oz;
    if (bx < 0 || bx >= cpd || by < 0 || by >= cpd || bz < 0 || bz >= cpd) continue;
    const int bin = (bx * cpd + by) * cpd + bz;
    const uint32_t bst = A.delim[bin];
    const int cnt = (int)(A.delim[bin + 1] - bst);
    for (int k0 = 0; k0 < cnt && !done; k0 += 64) {
      const int k = k0 + lane;
      bool valid = false;
      uint32_t jidx = 0;
      double px = 0, py = 0, pz = 0, d2 = 0;
      if (k < cnt) {
        const uint32_t s = bst + (uint32_t)k;
        jidx = A.sorted[s];
        valid = (jidx != self) && (A.target_group < 0 || A.groups[jidx] == A.target_group);
        px = A.spx[s] - sx;
        py = A.spy[s] - sy;
        pz = A.spz[s] - sz;
        d2 = (px * px + py * py) + pz * pz;
      }
      double nx = 0, ny = 0, nz = 0, off = 0;
      if (valid) {  // halfway_from_origin_to, vector3.rs:223-239
        const double mag = sqrt(d2);
        const double inv = 1.0 / mag;
        nx = px * inv;
        ny = py * inv;
        nz = pz * inv;
        off = (nx * (px * 0.5) + ny * (py * 0.5)) + nz * (pz * 0.5);
      }
      uint64_t pend = ballot(valid);
      streamed += (uint32_t)popc(pend);
      while (pend) {
        const int q = ctz64(pend);
        pend &= pend - 1ull;
        if (rl_d(d2, q) > stop_thr(r2)) continue;
        const int nv0 = ne2 / 2 - nf + 2;
        classified++;
        flops += 13 + 6ull * (uint64_t)nv0;
        const int rc = spill_cut(W, b, nf, ne2, P.rmax, P.fmax, rl_d(nx, q), rl_d(ny, q), rl_d(nz, q),
                                 rl_d(off, q), (int32_t)rl_i((int)jidx, q), lane);
        if (rc == SPILL_NOCUT) continue;
        if (rc < 0) {
          status |= rc == SPILL_CAPACITY ? TESS_CELL_CAPACITY : TESS_CELL_DEGENERATE;
          done = true;
          break;
        }
        b ^= 1;
        ncuts++;
        ninc += rc == 0;
        const int nv = ne2 / 2 - nf + 2;
        if (nv > maxv) maxv = nv;
        flops += 12ull * (uint64_t)rc + 5ull * (uint64_t)nv;
        r2 = spill_r2(W.ring[b], ne2, lane);
      }
    }
    if (done) break;
  }
  if (!done && e >= A.tlen && !(A.kcut > stop_thr(r2)) && !(A.kcut > rlim)) status |= TESS_CELL_TABLE;
  flops += 8ull * streamed;

  const uint32_t oi = ext ? (uint32_t)gpos : self;
  const bool fail = (status & (TESS_CELL_CAPACITY | TESS_CELL_DEGENERATE | TESS_CELL_TABLE)) != 0;
  uint64_t fbase = (uint64_t)oi * kFaceStage;
  uint64_t oflops = 0;
  bool wrote = false;
  if (!fail) {
    if (A.fpos != nullptr) {
      fbase = arena_reserve(A.arena_next, A.arena_cap, (uint32_t)nf, lane);
      if (fbase == ~0ull) status |= TESS_CELL_ARENA;
    } else if (nf > kFaceStage) {
      status |= TESS_CELL_CAPACITY;
    }
  }
  if (!fail && !(status & (TESS_CELL_CAPACITY | TESS_CELL_ARENA))) {
    // per face: the fan from the lexicographically smallest vertex
    // (polyhedron.rs:776-808), its volume term and area (interface.rs:408-410)
    const double* R = W.ring[b];
    const uint32_t* FO = W.foff[b];
    const int32_t* FT = W.ftag[b];
    double* contrib = W.seg;
    double* area = W.seg + P.fmax;
    for (int f = lane; f < nf; f += 64) {
      const int a = (int)FO[f], k = (int)(FO[f + 1] - FO[f]);
      int an = 0;
      for (int i = 1; i < k; ++i)
        if (lex_less(R[3 * (a + i)], R[3 * (a + i) + 1], R[3 * (a + i) + 2], R[3 * (a + an)], R[3 * (a + an) + 1],
                     R[3 * (a + an) + 2]))
          an = i;
      const double ax = R[3 * (a + an)], ay = R[3 * (a + an) + 1], az = R[3 * (a + an) + 2];
      int j1 = an + 1;
      if (j1 == k) j1 = 0;
      double cx = R[3 * (a + j1)] - ax, cy = R[3 * (a + j1) + 1] - ay, cz = R[3 * (a + j1) + 2] - az;
      double wx = 0.0, wy = 0.0, wz = 0.0;
      int jj = j1;
      for (int j = 2; j < k; ++j) {
        ++jj;
        if (jj == k) jj = 0;
        const double qx = cx, qy = cy, qz = cz;
        cx = R[3 * (a + jj)] - ax;
        cy = R[3 * (a + jj) + 1] - ay;
        cz = R[3 * (a + jj) + 2] - az;
        wx = wx + (qy * cz - qz * cy);
        wy = wy + (qz * cx - qx * cz);
        wz = wz + (qx * cy - qy * cx);
      }
      contrib[f] = (ax * wx + ay * wy) + az * wz;
      area[f] = 0.5 * sqrt((wx * wx + wy * wy) + wz * wz);
      oflops += 3ull * (k - 1) + 12ull * (k - 2) + 13;
    }
    ws_sync();
    // neighbours ascending: a face's rank is the number of smaller tags
    for (int f = lane; f < nf; f += 64) {
      const int32_t t = FT[f];
      int rank = 0;
      for (int g = 0; g < nf; ++g) rank += FT[g] < t;
      A.snbr[fbase + rank] = t;
      A.sarea[fbase + rank] = area[f];
    }
    if (lane == 0) {
      double vol = 0.0;
      for (int f = 0; f < nf; ++f) vol = vol + contrib[f];
      A.volume[oi] = vol / 6.0;
      A.fcount[oi] = (uint32_t)nf;
      if (A.fpos) A.fpos[oi] = fbase;
    }
    wrote = true;
  }
  oflops = wave_sum_u64(oflops);
  if (lane == 0) {
    if (!wrote) {
      A.volume[oi] = 0.0;
      A.fcount[oi] = 0;
    }
    A.status[oi] = status;
    if (A.summary) {
      if (wrote) {
        atomicAdd(&A.summary->flops, (unsigned long long)(flops + oflops + 1));
        atomicAdd(&A.summary->candidates, (unsigned long long)streamed);
        if (ninc) atomicAdd(&A.summary->incident_cuts, (unsigned long long)ninc);
      }
      if (status & TESS_CELL_DEGENERATE) atomicAdd(&A.summary->degenerate, 1u);
    }
    if (A.stats) {
      tess_cell_stats st;
      st.streamed = streamed;
      st.classified = classified;
      st.cuts = ncuts;
      st.entries = (uint32_t)e;
      st.flops = flops + oflops + 1;
      st.r2 = r2;
      st.max_vertices = (uint32_t)maxv;
      st.status = status;
      st.incident_cuts = ninc;
      st.reserved = 0;
      A.stats[oi] = st;
    }
    if ((status & TESS_CELL_CAPACITY) && P.redo) P.redo[atomicAdd(P.nredo, 1u)] = item;
    if ((status & TESS_CELL_TABLE) && P.tredo) P.tredo[atomicAdd(P.ntredo, 1u)] = item;
  }
}

// Every block owns one workspace and takes items blockIdx.x, + gridDim.x, ...
__global__ __launch_bounds__(64) void k_cells_spill(SpillArgs P) {
  const int lane = threadIdx.x;
  SpillWs W;
  spill_ws(P, blockIdx.x, W);
  typedef const __attribute__((address_space(4))) CellArgs ConstArgs;
  for (int i = blockIdx.x; i < P.count; i += gridDim.x) {
    const uint64_t it = P.list64 ? P.list64[i] : (uint64_t)P.list32[i];
    ConstArgs* cs = (ConstArgs*)(P.sets) + (uint32_t)(it >> 32);
    spill_cell(*(const CellArgs*)cs, (int)(uint32_t)it, P, W, it, lane);
    ws_sync();
  }
}

template __global__ void k_cells<1, 2, 1>(CellArgs);
template __global__ void k_cells<2, 3, 2>(CellArgs);
template __global__ void k_cells<4, 6, 2>(CellArgs);
template __global__ void k_cells_items<1, 2, 1>(BatchArgs, uint32_t);
template __global__ void k_cells_items<2, 3, 2>(BatchArgs, uint32_t);
template __global__ void k_cells_items<4, 6, 2>(BatchArgs, uint32_t);
template __global__ void k_cells_queue<2, 3, 2>(BatchArgs);
template __global__ void k_cells_queue<4, 6, 2>(BatchArgs);

// f64 sqrt / reciprocal / division as compiled for the cell kernel.
__global__ void k_diag_arith(const double* a, const double* b, int n, double* s, double* r, double* q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  s[i] = sqrt(a[i]);
  r[i] = 1.0 / a[i];
  q[i] = a[i] / b[i];
}

}  // namespace tess

namespace tess {

hipError_t launch_cells_items(int vpl, const BatchArgs& b, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t nb = b.xcd_chunk ? xcd_grid(n, (uint32_t)b.xcd_chunk) : n;
  switch (vpl) {
    case 1:
      hipLaunchKernelGGL((k_cells_items<1, 2, 1>), dim3(nb), dim3(64), 0, s, b, n);
      break;
    case 2:
      hipLaunchKernelGGL((k_cells_items<2, 3, 2>), dim3(nb), dim3(64), 0, s, b, n);
      break;
    case 4:
      hipLaunchKernelGGL((k_cells_items<4, 6, 2>), dim3(nb), dim3(64), 0, s, b, n);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cells_queue(int vpl, const BatchArgs& b, int grid, hipStream_t s) {
  if (grid <= 0 || b.chunk <= 0 || (vpl != 2 && vpl != 4)) return hipErrorInvalidValue;
  if (vpl == 2)
    hipLaunchKernelGGL((k_cells_queue<2, 3, 2>), dim3(grid), dim3(64), 0, s, b);
  else
    hipLaunchKernelGGL((k_cells_queue<4, 6, 2>), dim3(grid), dim3(64), 0, s, b);
  return hipGetLastError();
}

int cells_queue_blocks_per_cu(int vpl) {
  int nb = 0;
  hipError_t e = vpl == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cells_queue<2, 3, 2>, 64, 0)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cells_queue<4, 6, 2>, 64, 0);
  return e == hipSuccess && nb > 0 ? nb : 1;
}

hipError_t launch_cells(int vpl, const CellArgs& a, hipStream_t s) {
  if (a.nwork <= 0) return hipSuccess;
  int nb = a.nwork;
  if (a.xcd_chunk) nb = (int)xcd_grid((uint32_t)nb, (uint32_t)a.xcd_chunk);
  switch (vpl) {
    case 1:
      hipLaunchKernelGGL((k_cells<1, 2, 1>), dim3(nb), dim3(64), 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((k_cells<2, 3, 2>), dim3(nb), dim3(64), 0, s, a);
      break;
    case 4:
      hipLaunchKernelGGL((k_cells<4, 6, 2>), dim3(nb), dim3(64), 0, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static_assert(Lds<1, 2, 1>::HC == kCellHalfEdges(1) && Lds<4, 6, 2>::HC == kCellHalfEdges(4),
              "geometry staging capacity");

hipError_t launch_cells_geometry(int vpl, const CellArgs& a, bool with_cells, hipStream_t s) {
  if (a.nwork <= 0) return hipSuccess;
  const int nb = a.xcd_chunk ? (int)xcd_grid((uint32_t)a.nwork, (uint32_t)a.xcd_chunk) : a.nwork;
  if (vpl == 1 && !with_cells)
    hipLaunchKernelGGL((k_cells<1, 2, 1, 1>), dim3(nb), dim3(64), 0, s, a);
  else if (vpl == 4 && !with_cells)
    hipLaunchKernelGGL((k_cells<4, 6, 2, 1>), dim3(nb), dim3(64), 0, s, a);
  else if (vpl == 1)
    hipLaunchKernelGGL((k_cells<1, 2, 1, 2>), dim3(nb), dim3(64), 0, s, a);
  else if (vpl == 4)
    hipLaunchKernelGGL((k_cells<4, 6, 2, 2>), dim3(nb), dim3(64), 0, s, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_cells_spill(const SpillArgs& p, int blocks, hipStream_t s) {
  if (p.count <= 0) return hipSuccess;
  if (blocks <= 0 || p.rmax < 24 || p.fmax < 8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_cells_spill, dim3(blocks), dim3(64), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_diag_arith(const double* a, const double* b, int n, double* sq, double* rc,
                             double* dv, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_arith, dim3((n + 255) / 256), dim3(256), 0, s, a, b, n, sq, rc, dv);
  return hipGetLastError();
}

}  // namespace tess
