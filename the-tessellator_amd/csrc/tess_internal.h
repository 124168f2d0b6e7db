// tess_internal.h -- structures shared by the host runtime and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tess.h"

namespace tess {

// Uniform grid over the points' bounding box (celery.rs:81-188).
struct GridDev {
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double sx, sy, sz;  // cell sizes
  double ix, iy, iz;  // inverse cell sizes
  int32_t cpd;
  int32_t pad;
};

// Shell-table entry offset packed as three 10-bit signed fields.
__host__ __device__ inline int32_t pack_off(int ox, int oy, int oz) {
  return ((ox + 512) << 20) | ((oy + 512) << 10) | (oz + 512);
}
__host__ __device__ inline void unpack_off(int32_t p, int& ox, int& oy, int& oz) {
  ox = ((p >> 20) & 1023) - 512;
  oy = ((p >> 10) & 1023) - 512;
  oz = (p & 1023) - 512;
}

// Launch-wide counters accumulated by the cell kernel (one atomic per cell).
struct Summary {
  unsigned long long flops;       // algorithmic f64 ops of all cells
  unsigned long long candidates;  // candidates streamed
  unsigned int degenerate;        // cells with TESS_CELL_DEGENERATE
  unsigned int capacity;          // cells with TESS_CELL_CAPACITY
  unsigned int table;             // cells with TESS_CELL_TABLE
  unsigned int fallback;          // learned mode: cells that fell back
  unsigned long long incident_cuts;  // cuts with no Inside vertex next to the Outside region
};

// Faces staged per cell before CSR compaction when a pass writes per-cell
// slots (single-cell queries, geometry staging); whole-set passes write
// their faces into a shared arena instead (CellArgs::fpos).
constexpr int kFaceStage = 64;

// The face arena is split into one region per XCD (workgroups are dealt to
// the 8 XCDs in turn, so a wave's region is its block number mod 8), each
// 7/64 of the capacity, and a shared overflow region (the last 1/8): the
// ranges that neighbouring waves reserve are written through one L2 and
// leave it as whole lines.  Counters sit one 128-byte line apart: the 8
// regions', the overflow's, then the count of cells that found no room.
constexpr int kArenaParts = 8;
constexpr int kArenaStride = 16;  // u64 words between counters
constexpr int kArenaWords = (kArenaParts + 2) * kArenaStride;
__host__ __device__ inline uint64_t arena_part_size(uint64_t cap) { return cap / 64 * 7; }
// Faces an arena of n cells holds: 64 per cell, at least 2^16 (so that every
// region of a small set still takes a cell of hundreds of faces).
__host__ __device__ inline uint64_t arena_capacity(uint64_t n) {
  return n * kFaceStage > (1ull << 16) ? n * kFaceStage : (1ull << 16);
}

// Learned search-radius model 27 -> 64 -> 64 -> 2.  Parameter vector layout
// (FP32 master copy): W1 [32][64] (rows 27..31 are the zero feature pad),
// W2 [64][64], W3 [64][16] (columns 2..15 unused), b1 [64], b2 [64],
// b3 [16].  Weights row-major [input][output].
constexpr int kMlpIn = 32;
constexpr int kMlpFeatures = 27;
constexpr int kMlpH = 64;
constexpr int kMlpOut = 2;
constexpr int kMlpOutPad = 16;
constexpr int kMlpWeights = kMlpIn * kMlpH + kMlpH * kMlpH + kMlpH * kMlpOutPad;
constexpr int kMlpBias = kMlpH + kMlpH + kMlpOutPad;
constexpr int kMlpParams = kMlpWeights + kMlpBias;
static_assert(kMlpParams == TESS_MLP_PARAMS, "parameter count");

struct CellArgs {
  int n;                   // points in the set
  const double* spx;       // positions in grid order
  const double* spy;
  const double* spz;
  const uint32_t* sorted;  // grid order -> original index
  const uint32_t* delim;   // CSR delimiters, cpd^3 + 1
  const GridDev* grid;
  const double* tkey;      // shell table (canonical order)
  const int32_t* toff;
  int tlen;
  double kcut;             // every entry past the table has key >= kcut
  double lo[3], hi[3];     // container box
  const uint32_t* work;    // grid positions to compute (nullptr: 0..nwork-1)
  int nwork;
  const double* ext;       // nullable: sites not in the set, AoS; work items
                           // then index these sites and the outputs
                           // (Diagram::get_cell_at_particle, no self)
  int xcd_chunk;           // > 0: XCDs take runs of this many consecutive cells
  const int64_t* groups;   // nullable
  int64_t target_group;    // < 0: none
  double radius_limit;     // NaN: none (expand_all_in_radius cut-off otherwise)
  // outputs, indexed by original point index
  double* volume;
  uint32_t* fcount;
  int32_t* snbr;           // face arena (fpos set) or per-cell slots [n][kFaceStage]
  double* sarea;
  uint64_t* fpos;          // nullable: cell o's faces start at snbr[fpos[o]], a
                           // range reserved by one atomic on the arena's
                           // counters (no per-cell face limit); null: at
                           // o * kFaceStage
  unsigned long long* arena_next;  // kArenaWords counters (kArenaParts regions,
                                   // overflow, cells without room)
  uint64_t arena_cap;      // faces the arena holds
  uint32_t* status;
  tess_cell_stats* stats;  // nullable
  Summary* summary;        // nullable
  uint32_t* redo;          // grid positions of cells needing a larger variant
  uint32_t* nredo;
  uint32_t* tredo;         // nullable: cells whose security ball reaches past
  uint32_t* ntredo;        // the shell table (rerun with a longer one)
  int defer_table;         // 1: such cells write no outputs (a rerun will)
  // geometry of whole sets (the geometry kernel only), staged per work item
  // w: vertices geo_v[w][64 VPL][3], rings geo_ring[w][half-edge capacity],
  // ring offsets in the cell geo_roff[w][kFaceStage]; geo_count[o] = live
  // vertices and geo_count[n + o] = ring entries of cell o; geo_slot[w] =
  // o + 1 once staged (0: nothing staged, e.g. a capacity overflow)
  uint32_t* geo_count;
  uint32_t* geo_slot;
  double* geo_v;
  uint16_t* geo_ring;
  uint16_t* geo_roff;
  // single-cell geometry export (nullable): vertices and face rings
  double* exp_v;           // [256][3]
  int32_t* exp_ring;       // [256 faces][64] vertex ids, -1 terminated
  int32_t* exp_meta;       // nv, nf
};

// Persistent cell pass over many sets (one wavefront per cell at a time).
struct BatchArgs {
  const CellArgs* sets;   // per-set arguments (device array)
  const uint64_t* items;  // (set << 32) | grid position
  const uint32_t* order;  // nullable: item i is items[order[i]]
  const uint32_t* count;  // number of items (device)
  uint32_t* counter;      // next item to take; 0 before the launch
  int chunk;              // items taken per atomic
  uint64_t* redo;         // capacity overflows (nullable)
  uint32_t* nredo;
  uint64_t* tredo;        // shell-table exhaustions (nullable)
  uint32_t* ntredo;
  int xcd_chunk;          // > 0: XCDs take runs of this many consecutive items
};

// Cells too large for the LDS kernels (more than 256 vertices or 128
// faces): one wavefront per cell with the polyhedron as face rings in a
// global-memory workspace of rmax ring entries and fmax faces per block.
struct SpillArgs {
  const CellArgs* sets;    // device: per-set arguments (a diagram: one)
  const uint32_t* list32;  // items of set 0 (grid positions / site indices)
  const uint64_t* list64;  // or (set << 32) | position
  int count;
  int rmax, fmax;
  double* wsd;             // per block: two ring buffers [rmax][3], segments [fmax][6]
  uint32_t* wsu;           // per block: face offsets 2 x [fmax + 1], tags 2 x [fmax],
                           // segment links [fmax], face terms [fmax][4]
  uint8_t* wsb;            // per block: ring-entry classes [rmax], segment flags [fmax]
  uint64_t* redo;          // items that outgrew the workspace
  uint32_t* nredo;
  uint64_t* tredo;         // items that exhausted the shell table
  uint32_t* ntredo;
};
// Workspace of one spill block.
__host__ __device__ inline size_t spill_wsd_doubles(int rmax, int fmax) { return 6ull * rmax + 6ull * fmax; }
__host__ __device__ inline size_t spill_wsu_words(int fmax) { return 2ull * (fmax + 1) + 7ull * fmax; }
__host__ __device__ inline size_t spill_wsb_bytes(int rmax, int fmax) { return (size_t)rmax + (size_t)fmax; }

// Radius queries over one grid (query.hip).
struct QueryArgs {
  GridDev g;
  const uint32_t* sorted;  // grid-ordered point ids [n]
  const uint32_t* delim;   // [cpd^3 + 1]
  const double *sx, *sy, *sz;  // gathered coordinates (grid order)
  const int64_t* groups;   // by point id; null: no group filter
  int64_t target_group;
  const double* q;         // query positions, AoS [3 nq]
  const uint32_t* order;   // nullable: wavefront w takes query order[w]
  const double* radius;    // radius of query i: radius[i * rstride]
  int rstride;
  int nq;
  int kind;                // 0 cells, 1 cell radius, 2 real radius, 3 neighbour cloud
  const double* key;       // shell table (kind 3)
  const int32_t* off;
  int tlen;
  uint32_t* count;         // count pass: results per query
  const uint64_t* offsets; // fill pass: CSR offsets [nq + 1]
  uint32_t* ids;
};

}  // namespace tess
