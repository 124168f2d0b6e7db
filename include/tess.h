/* tess.h -- C ABI of the MI355X Voronoi engine.
 *
 * The surface mirrors the-tessellator's public API (src/interface.rs):
 *   Diagram      -> tess_diagram   (interface.rs:24-47, built from points;
 *                                   the reference's add_particle_with_group /
 *                                   initialize, interface.rs:52-84, are private)
 *   Cell         -> tess_cell      (interface.rs:236-390)
 *   VoronoiFace  -> face index of a tess_cell (interface.rs:393-417)
 * plus the all-cells entry point (tess_diagram_compute) that evaluates
 * Cell::compute_voronoi_cell / compute_volume / compute_neighbors /
 * compute_faces for every point in one device pass, which is the path a
 * caller looping over get_cell_at_index takes in the reference.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Host pointers unless the name says
 *    "device".  No exception or panic crosses the ABI: every fallible call
 *    returns a tess_status and tess_last_error() explains it.
 *  - Neighbour ids are original point indices; the six container walls are
 *    -1..-6 for -x, +x, -y, +y, -z, +z (the reference panics on walls,
 *    polyhedron.rs:871-881).  Neighbours of a cell are sorted ascending and
 *    face areas are aligned with them.  face_count includes wall faces.
 *  - A context owns one HIP stream and is not thread-safe; diagrams are
 *    immutable after creation and may be queried from distinct contexts.
 */
#ifndef TESS_H
#define TESS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  TESS_OK = 0,
  TESS_E_INVALID = 1,    /* bad argument (null pointer, n == 0, empty box) */
  TESS_E_NOMEM = 2,      /* host or device allocation failed */
  TESS_E_DEVICE = 3,     /* HIP runtime or kernel launch error */
  TESS_E_CAPACITY = 4,   /* a cell outgrew the largest kernel capacity */
  TESS_E_DEGENERATE = 5, /* a cell failed its topology checks */
  TESS_E_STATE = 6,      /* call order: e.g. results read before compute */
  TESS_E_NOTFOUND = 7    /* model file missing or malformed */
} tess_status;

typedef struct tess_ctx tess_ctx;
typedef struct tess_diagram tess_diagram;
typedef struct tess_cell tess_cell;

enum {
  TESS_MODE_PLAIN = 0,   /* shell search in the reference's order */
  TESS_MODE_LEARNED = 1  /* same output; MLP pre-sizing + cost-sorted queue */
};

typedef struct {
  /* Points per grid cell.  1.25 is the reference's constant
   * (celery.rs:185-188); cells_per_dimension = floor(cbrt(n / density)) + 1
   * (celery.rs:161-162). */
  double cell_density;
  int32_t mode;            /* TESS_MODE_* */
  int32_t vertex_capacity; /* 0: automatic; else 64, 128 or 256 */
  int32_t collect_stats;   /* 1: keep per-cell search counters */
  int32_t reserved;
} tess_opts;

/* Per-cell search counters (collect_stats = 1). */
typedef struct {
  uint32_t streamed;   /* candidates taken from the shell stream (self excluded) */
  uint32_t classified; /* candidates whose bisector was tested against the cell */
  uint32_t cuts;       /* candidates that changed the cell */
  uint32_t entries;    /* shell-table entries visited */
  uint64_t flops;      /* algorithmic f64 operations */
  double r2;           /* final max |vertex - site|^2 */
  uint32_t max_vertices;
  uint32_t status;     /* 0 ok; TESS_CELL_* bits */
} tess_cell_stats;

enum {
  TESS_CELL_CAPACITY = 1u,   /* cell needed a larger kernel variant */
  TESS_CELL_TABLE = 2u,      /* shell table exhausted before the stop */
  TESS_CELL_DEGENERATE = 4u, /* topology check failed */
  TESS_CELL_FALLBACK = 8u    /* learned mode: prediction under-sized the batch */
};

/* Device pointers to a computed diagram's results (valid until destroy). */
typedef struct {
  const double* volume;       /* [n] */
  const uint32_t* face_count; /* [n] */
  const uint64_t* face_offset;/* [n+1] */
  const int32_t* neighbor;    /* [total_faces] */
  const double* face_area;    /* [total_faces] */
  uint64_t total_faces;
  size_t n;
} tess_cells_device;

/* Phase timings of the last tess_diagram_compute (milliseconds, HIP events). */
typedef struct {
  double bin_ms;     /* bounds + counting sort + gather */
  double cells_ms;   /* cell kernel(s) */
  double compact_ms; /* face scan + CSR compaction */
  double total_ms;   /* first event to last event */
  double table_ms;   /* host shell-table lookup/build (wall clock) */
  uint64_t flops;    /* algorithmic f64 ops of the cells (collect_stats = 1) */
  uint64_t fallback_cells; /* learned mode: cells that fell back to the plain search */
  uint64_t capacity_reruns;/* cells rerun with a larger capacity */
} tess_timings;

void tess_opts_default(tess_opts* opts);

/* ---- context ---------------------------------------------------------- */
tess_status tess_ctx_create(int hip_device, tess_ctx** out);
void tess_ctx_destroy(tess_ctx* ctx);
const char* tess_last_error(const tess_ctx* ctx);
/* The context's hipStream_t (returned as void*). */
void* tess_ctx_stream(tess_ctx* ctx);
tess_status tess_ctx_synchronize(tess_ctx* ctx);

/* ---- diagram (interface.rs:24-84) --------------------------------------
 * xyz: n points, AoS (x0,y0,z0,x1,...), copied for the call.  lo/hi: the
 * container box; every cell is the box clipped by bisectors (the reference's
 * container polyhedron, interface.rs:66-79).  The spatial grid is built over
 * the points' bounding box (celery.rs:81-125).  groups may be NULL. */
tess_status tess_diagram_create(tess_ctx* ctx, const double* xyz, size_t n, const double lo[3],
                                const double hi[3], const int64_t* groups, const tess_opts* opts,
                                tess_diagram** out);
/* Same with device-resident AoS points (no host copy). */
tess_status tess_diagram_create_device(tess_ctx* ctx, const double* device_xyz, size_t n,
                                       const double lo[3], const double hi[3],
                                       const int64_t* device_groups, const tess_opts* opts,
                                       tess_diagram** out);
void tess_diagram_destroy(tess_diagram* d);
size_t tess_diagram_size(const tess_diagram* d);

/* All cells: binning (celery.rs:231-250) + Cell::compute_voronoi_cell /
 * compute_volume / compute_neighbors / compute_faces for every point
 * (interface.rs:257-384).  Results stay in device memory; enqueued on the
 * context stream and completed before return. */
tess_status tess_diagram_compute(tess_diagram* d);
tess_status tess_diagram_total_faces(const tess_diagram* d, uint64_t* out);
/* Copy results to host arrays: volume[n], face_count[n], face_offset[n+1],
 * neighbor[total_faces], face_area[total_faces]; any pointer may be NULL. */
tess_status tess_diagram_copy_cells(const tess_diagram* d, double* volume, uint32_t* face_count,
                                    uint64_t* face_offset, int32_t* neighbor, double* face_area);
tess_status tess_diagram_device_cells(const tess_diagram* d, tess_cells_device* out);
/* Per-cell counters (needs collect_stats = 1): stats[n]. */
tess_status tess_diagram_copy_stats(const tess_diagram* d, tess_cell_stats* stats);
tess_status tess_diagram_timings(const tess_diagram* d, tess_timings* out);
/* Grid of the diagram (celery.rs:153-181): bounds[6] = xmin,xmax,ymin,ymax,
 * zmin,zmax; sizes[3]; cells_per_dimension. */
tess_status tess_diagram_grid(const tess_diagram* d, double bounds[6], double sizes[3],
                              int64_t* cells_per_dimension);
/* Celery::sorted_indices / delimiters (celery.rs:193-227): sorted[n],
 * delimiters[cpd^3 + 1] (pass NULL to skip). */
tess_status tess_diagram_copy_grid(const tess_diagram* d, uint32_t* cell_of_point,
                                   uint32_t* sorted_indices, uint64_t* delimiters);

/* ---- one cell (interface.rs:186-208, :257-390) -------------------------
 * search_radius: NaN for None (expanding search with the security-radius
 * stop); otherwise the reference's expand_all_in_radius cut-off
 * (celery.rs:1023-1075, key compared against the radius as written).
 * target_group: < 0 for None. */
tess_status tess_cell_at_index(tess_diagram* d, size_t index, double search_radius,
                               int64_t target_group, tess_cell** out);
void tess_cell_destroy(tess_cell* c);
int64_t tess_cell_original_index(const tess_cell* c);             /* interface.rs:387-389 */
double tess_cell_volume(const tess_cell* c);                      /* interface.rs:337-339 */
/* Neighbour ids (interface.rs:342-344); returns the count, writes <= cap. */
size_t tess_cell_neighbors(const tess_cell* c, int64_t* out, size_t cap);
size_t tess_cell_face_count(const tess_cell* c);                  /* interface.rs:373-384 */
/* VoronoiFace::compute_area / compute_neighbor (interface.rs:408-416). */
tess_status tess_cell_face(const tess_cell* c, size_t face, double* area, int64_t* neighbor);
/* VoronoiFace::compute_vertices (interface.rs:403-405): site-relative
 * coordinates, xyz[3*k]; returns the vertex count, writes <= cap. */
size_t tess_cell_face_vertices(const tess_cell* c, size_t face, double* xyz, size_t cap);
/* Cell::compute_vertices (interface.rs:368-370), site-relative. */
size_t tess_cell_vertices(const tess_cell* c, double* xyz, size_t cap);

/* ---- diagnostics (no counterpart in the reference) --------------------- */
/* Device f64 sqrt, reciprocal and division of n inputs (for checking that
 * they are correctly rounded like the host's). */
tess_status tess_diag_arith(tess_ctx* ctx, const double* a, const double* b, size_t n,
                            double* sqrt_a, double* recip_a, double* a_div_b);

#ifdef __cplusplus
}
#endif
#endif /* TESS_H */
