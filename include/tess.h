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
  int32_t xcd_chunk;       /* whole-set passes: cells each XCD takes in a row
                              (0: automatic, 256) */
  int32_t point_order;     /* TESS_ORDER_*: the order of a grid cell's points
                              in the candidate stream */
  int32_t reserved;
} tess_opts;

enum {
  /* ascending input index (the default) */
  TESS_ORDER_INPUT = 0,
  /* the reference's internal order: Diagram::sort_particles by
   * morton_index (interface.rs:88-183; 10 levels on a 1024 grid over the
   * points' bounding box grown from the origin), ties by input index.  The
   * Celery is built over the sorted points (interface.rs:64), so a cell's
   * candidates come in this order.  Ids stay original indices. */
  TESS_ORDER_MORTON = 1
};

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
  uint32_t incident_cuts; /* cuts whose Outside region touched no Inside vertex
                             (bordered by Incident vertices only): the
                             reference's find_outgoing_edge, polyhedron.rs:418-432,
                             needs an Inside -> Outside edge and skips them */
  uint32_t reserved;
} tess_cell_stats;

enum {
  TESS_CELL_CAPACITY = 1u,   /* cell needed a larger kernel variant */
  TESS_CELL_TABLE = 2u,      /* shell table exhausted before the stop */
  TESS_CELL_DEGENERATE = 4u, /* topology check failed */
  TESS_CELL_FALLBACK = 8u,   /* learned mode: prediction under-sized the batch */
  TESS_CELL_RERUN = 16u,     /* recomputed at the largest capacity after overflowing */
  TESS_CELL_LONG_TABLE = 32u,/* recomputed with a longer shell table: its security
                                ball reached past the default 25^3-offset prefix
                                (the reference streams all (2 cpd - 1)^3 offsets,
                                celery.rs:418-679, :971-1018) */
  TESS_CELL_SPILL = 64u,     /* more than 256 vertices or 128 faces: computed by
                                the global-memory kernel (no size limit but the
                                workspace, grown on demand like the reference's
                                Pool, pool.rs:85-110) */
  TESS_CELL_ARENA = 128u     /* no room left in the face arena: the call fails
                                with TESS_E_CAPACITY */
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

/* Phase timings of the last compute (milliseconds, HIP events). */
typedef struct {
  double bin_ms;     /* bounds + counting sort + gather */
  double cells_ms;   /* cell kernel(s) */
  double compact_ms; /* face scan + CSR compaction */
  double total_ms;   /* first event to last event */
  double table_ms;   /* host shell-table lookup/build (wall clock) */
  uint64_t flops;    /* algorithmic f64 ops of the cells (collect_stats = 1) */
  uint64_t fallback_cells; /* learned mode: cells that fell back to the plain search */
  uint64_t capacity_reruns;/* cells rerun with a larger capacity */
  double plan_ms;          /* learned mode: features + inference + work lists */
  uint64_t table_reruns;   /* cells recomputed with a longer shell table */
  uint64_t spill_cells;    /* cells computed by the global-memory spill kernel */
  uint64_t incident_cuts;  /* cuts bordered by Incident vertices only (see
                              tess_cell_stats.incident_cuts) */
  uint64_t planned_long;   /* learned mode: cells routed to the long-search pass
                              by their predicted radius */
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
 * neighbor[total_faces], face_area[total_faces]; any pointer may be NULL.
 * Arrays of 16 MiB or more (here and in every other *_copy* call) come down
 * through two 64 MiB pinned host buffers the context allocates on first use
 * and keeps until tess_ctx_destroy, copied out by up to eight host threads;
 * calls on one context take the buffers in turn. */
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
/* Diagram::original_indices / internal_indices (interface.rs:35-39, set by
 * sort_particles, :88-148): the internal order is the Morton order of
 * TESS_ORDER_MORTON (whatever point_order the diagram uses); either output
 * may be NULL.  Computed on the device (Morton codes + radix sort). */
tess_status tess_diagram_copy_index_maps(tess_diagram* d, uint32_t* original_of_internal,
                                         uint32_t* internal_of_original);

/* ---- one cell (interface.rs:186-208, :257-390) -------------------------
 * search_radius: NaN for None (expanding search with the security-radius
 * stop); otherwise the reference's expand_all_in_radius cut-off
 * (celery.rs:1023-1075, key compared against the radius as written).
 * target_group: < 0 for None. */
tess_status tess_cell_at_index(tess_diagram* d, size_t index, double search_radius,
                               int64_t target_group, tess_cell** out);
/* Diagram::get_cell_at_particle (interface.rs:211-232): the cell of a site
 * that is not a point of the diagram (inside the container box); no point
 * is excluded as its own, and tess_cell_original_index returns -1 (None). */
tess_status tess_cell_at_particle(tess_diagram* d, const double xyz[3], double search_radius,
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

/* ---- geometry of every cell (interface.rs:367-370, :403-406) -----------
 * Cell::compute_vertices and VoronoiFace::compute_vertices
 * (polyhedron.rs:885-919) for all cells at once, as CSR.  On a diagram not
 * yet computed it runs the compute, whose cell passes stage the geometry as
 * well (one cell pass for both, at the default vertex capacity); after a
 * compute it runs one staging pass.  Vertices are site-relative, per cell in the kernel's
 * slot order (the reference's pool order is not reproducible);
 * vertex_offset[n + 1] indexes them.  Each face (in the order of the
 * neighbour lists, so face f of the whole diagram is neighbor[f]) has a ring
 * ring[ring_offset[f] .. ring_offset[f + 1]) of vertex numbers within its
 * cell (global vertex = vertex_offset[cell] + entry), starting at the face's
 * lexicographically smallest vertex and counter-clockwise seen from outside.
 * ring_offset has total_faces + 1 entries. */
tess_status tess_diagram_compute_geometry(tess_diagram* d);
tess_status tess_diagram_geometry_sizes(const tess_diagram* d, uint64_t* nvertices, uint64_t* nring);
/* Device time of the last geometry call, ms: [0] the cell pass that stages
 * the geometry (with its capacity rerun), [1] the offsets and the copy into
 * the CSR. */
tess_status tess_diagram_geometry_timings(const tess_diagram* d, double ms[2]);
tess_status tess_diagram_copy_geometry(const tess_diagram* d, uint64_t* vertex_offset, double* vertices,
                                       uint64_t* ring_offset, uint16_t* ring);

/* ---- cells of outside sites (interface.rs:211-232) ---------------------
 * Diagram::get_cell_at_particle + Cell::compute_voronoi_cell for m sites
 * that are not points of the diagram (xyz: host AoS): each cell is the
 * container box clipped by the bisectors between the site and the diagram's
 * points (all of them, or those of target_group >= 0; no point is excluded
 * as the cell's own).  search_radius: NaN = None, else the
 * expand_all_in_radius cut-off of tess_cell_at_index.  Sites must lie in the
 * container box (TESS_E_INVALID otherwise).  Results are laid out as
 * tess_diagram_copy_cells, indexed by site.  The call's device temporaries
 * (the sites, a face arena of 64 faces per site, work lists) are kept by the
 * diagram for its next calls and freed with it. */
typedef struct tess_site_cells tess_site_cells;
tess_status tess_diagram_cells_at_particles(tess_diagram* d, const double* xyz, size_t m, double search_radius,
                                            int64_t target_group, tess_site_cells** out);
size_t tess_site_cells_count(const tess_site_cells* p);
tess_status tess_site_cells_total_faces(const tess_site_cells* p, uint64_t* out);
tess_status tess_site_cells_copy(const tess_site_cells* p, double* volume, uint32_t* face_count,
                                 uint64_t* face_offset, int32_t* neighbor, double* face_area);
/* Per-site counters (the diagram must have been created with collect_stats). */
tess_status tess_site_cells_copy_stats(const tess_site_cells* p, tess_cell_stats* stats);
/* ms[0]: cell passes, ms[1]: face offsets + compaction; reruns at capacity 256. */
tess_status tess_site_cells_timings(const tess_site_cells* p, double ms[2], uint32_t* capacity_reruns);
void tess_site_cells_destroy(tess_site_cells* p);

/* ---- radius queries (celery.rs:700-855, interface.rs:348-365) ---------
 * Many queries of one kind over a diagram's grid (the grid is built first
 * if no compute has built it), with results as CSR: the ids of query i are
 * ids[offsets[i] .. offsets[i + 1]), in the reference's order.
 *  TESS_QUERY_CELLS        Celery::find_cells_in_radius (celery.rs:753-797):
 *                          grid cell ids (i cpd + j) cpd + k of the cells in
 *                          the clamped index box whose gap distance from the
 *                          query's cell is <= radius, in i, j, k loop order.
 *  TESS_QUERY_CELL_RADIUS  find_neighbors_in_cell_radius (:802-819): the
 *                          points of those cells, cell by cell, each cell's
 *                          points in grid (sorted-index) order.
 *  TESS_QUERY_REAL_RADIUS  find_neighbors_in_real_radius (:825-855): the same
 *                          walk, keeping points with squared distance
 *                          <= radius^2.
 *  TESS_QUERY_CLOUD        Cell::compute_neighbor_cloud (interface.rs:348-365):
 *                          ExpandingSearch::expand_all_in_radius from the
 *                          query position (celery.rs:1023-1075: bins in
 *                          search order while the squared gap distance is
 *                          <= radius, compared unsquared as there), then the
 *                          target-group filter when target_group >= 0.
 * Point ids are original indices.  xyz: nq host AoS positions (any position:
 * its grid cell is clamped as in celery.rs:269-282).  radius: nradius = 1
 * (one radius for all) or nq values.  target_group >= 0 needs
 * TESS_QUERY_CLOUD and a diagram with groups.  A NaN radius lists nothing
 * for the box kinds and every point for the cloud, as in the reference.
 * The call's device temporaries (per-query counts, the grid order of large
 * query sets, the filtered kinds' upper-bound slots: 4 B per candidate
 * point) are kept by the diagram for its next queries and freed with it. */
enum {
  TESS_QUERY_CELLS = 0,
  TESS_QUERY_CELL_RADIUS = 1,
  TESS_QUERY_REAL_RADIUS = 2,
  TESS_QUERY_CLOUD = 3
};
typedef struct tess_query tess_query;
tess_status tess_diagram_query(tess_diagram* d, int kind, const double* xyz, size_t nq, const double* radius,
                               size_t nradius, int64_t target_group, tess_query** out);
/* Same with device-resident positions and radii (read during the call). */
tess_status tess_diagram_query_device(tess_diagram* d, int kind, const double* device_xyz, size_t nq,
                                      const double* device_radius, size_t nradius, int64_t target_group,
                                      tess_query** out);
size_t tess_query_count(const tess_query* q);   /* number of queries */
uint64_t tess_query_total(const tess_query* q); /* ids over all queries */
/* offsets: nq + 1 entries; ids: tess_query_total entries (either may be NULL) */
tess_status tess_query_copy(const tess_query* q, uint64_t* offsets, uint32_t* ids);
/* Device pointers owned by q (valid until tess_query_destroy). */
tess_status tess_query_device(const tess_query* q, const uint64_t** offsets, const uint32_t** ids);
/* Device time, ms: [0] the count pass, [1] the fill pass, [2] the whole
 * call from the count pass to the last kernel (with the scans, the sizing
 * of the id buffer and, for the real radius and grouped clouds, the packing
 * of the ids the fill pass wrote into upper-bound slots: for those kinds the
 * count pass counts the unfiltered walk from the delimiters alone). */
tess_status tess_query_timings(const tess_query* q, double ms[3]);
void tess_query_destroy(tess_query* q);

/* ---- learned search-radius model (no counterpart in the reference) -----
 * A 27 -> 64 -> 64 -> 2 ReLU network.  Input: log(1 + count) of the 3x3x3
 * grid cells around a point's grid cell (bf16, padded to TESS_MLP_IN).
 * Outputs: log(R / s) and log(1 + C), where R is the cell's final security
 * radius (max vertex distance), s the smallest grid-cell edge and C the
 * number of candidates the plain search classifies; both are standardised
 * with the mean/std kept in the model.  Weights are bf16 on the MFMA units
 * with fp32 accumulation; an fp32 master copy is trained with Adam.
 * Predictions only choose each cell's vertex capacity and its place in the
 * work queue: results are identical to TESS_MODE_PLAIN for any prediction.
 *
 * Parameter vector (fp32, TESS_MLP_PARAMS): W1[32][64], W2[64][64],
 * W3[64][16] (row-major [in][out]; columns 2..15 of W3 unused), b1[64],
 * b2[64], b3[16].  norm[4] = mean0, std0, mean1, std1 of the targets.
 * calib[4] = routing thresholds on log(1 + C): cells above calib[0] start
 * at vertex capacity 128, above calib[1] at 256; calib[2..3] reserved.
 * Training sets them to the predicted value above which at least one cell
 * in five of the training samples outgrew 64 (resp. 128) vertices. */
#define TESS_MLP_IN 32
#define TESS_MLP_PARAMS 7312
typedef struct tess_model tess_model;

tess_status tess_model_create(tess_ctx* ctx, uint64_t seed, tess_model** out);
void tess_model_destroy(tess_model* m);
tess_status tess_model_get(const tess_model* m, float* params, float norm[4], float calib[4]);
tess_status tess_model_set(tess_model* m, const float* params, const float norm[4], const float calib[4]);
tess_status tess_model_save(const tess_model* m, const char* path);
tess_status tess_model_load(tess_model* m, const char* path);
/* Append every cell of a diagram computed with collect_stats = 1 as a
 * training sample (features on the device, targets from its counters). */
tess_status tess_model_add_samples(tess_model* m, tess_diagram* d);
size_t tess_model_sample_count(const tess_model* m);
/* Standardise the targets, run `epochs` passes of Adam over the samples in
 * shuffled mini-batches on the GPU, then set the routing thresholds from the
 * samples' vertex counts.  loss_out[epochs]: mean loss per epoch (nullable). */
tess_status tess_model_train(tess_model* m, int epochs, int batch, float lr, uint64_t seed, float* loss_out);
/* Predictions for every point of a diagram (after compute or binning):
 * out[n][2] = predicted R (length units) and C, in original order. */
tess_status tess_diagram_predict(tess_diagram* d, const tess_model* m, float* out);
/* All cells, learned schedule (same results as tess_diagram_compute). */
tess_status tess_diagram_compute_learned(tess_diagram* d, const tess_model* m);
/* All cells, schedule from caller-given predictions log(1 + C) per point
 * (original order; any values, e.g. adversarial) and thresholds calib[2]. */
tess_status tess_diagram_compute_scheduled(tess_diagram* d, const float* log1p_c, const float calib[2]);

/* Low-level access for checking the network against a CPU reference.
 * x: bf16 bit patterns [n][TESS_MLP_IN]; t: standardised targets [n][2]. */
tess_status tess_diagram_features(tess_diagram* d, uint16_t* x);
tess_status tess_mlp_forward(tess_model* m, const uint16_t* x, size_t n, float* y);
/* Gradient of 0.5 * sum((y - t)^2) / n w.r.t. the parameters (bf16 weights
 * as stored), summed over the batch in a fixed order; loss = that value. */
tess_status tess_mlp_gradient(tess_model* m, const uint16_t* x, const float* t, size_t n, float* grad,
                              float* loss);
/* One 16x16x32 bf16 MFMA: d[16][16] = a[16][32] * b[32][16] (row-major). */
tess_status tess_mfma_probe(tess_ctx* ctx, const uint16_t* a, const uint16_t* b, float* d);

/* ---- batches of independent sets (no counterpart in the reference) -----
 * Many point sets, each with its own container box and grid, tessellated
 * together.  Per set the results are bitwise those of a tess_diagram of the
 * same points.  The binning of every set is captured once into a HIP graph;
 * the cells of all sets run as one list in one launch (plus a persistent
 * rerun pass for cells that outgrow the first pass's vertex capacity),
 * captured into a second graph that is replayed while the schedule is
 * unchanged. */
typedef struct tess_batch tess_batch;
typedef struct {
  int32_t first_capacity; /* vertex capacity of the first pass: 64, 128, 256 */
  int32_t xcd_chunk;      /* 0: list entries dealt to the XCDs one by one;
                             K > 0: in runs of K consecutive entries */
  int32_t redo_chunk;     /* cells a wavefront takes per grab in the rerun pass */
  int32_t learned;        /* 1: heaviest predicted cells first (needs a model) */
  int32_t eager;          /* 1: launch the passes directly on the stream
                             instead of capturing and replaying HIP graphs
                             (same kernels, same arguments) */
} tess_batch_config;

/* xyz[s]: host AoS points of set s (n[s] of them); lo/hi: [nsets][3] boxes. */
tess_status tess_batch_create(tess_ctx* ctx, size_t nsets, const double* const* xyz, const size_t* n,
                              const double* lo, const double* hi, const tess_opts* opts, tess_batch** out);
void tess_batch_destroy(tess_batch* b);
size_t tess_batch_sets(const tess_batch* b);
tess_status tess_batch_set_config(tess_batch* b, const tess_batch_config* cfg);
tess_status tess_batch_get_config(const tess_batch* b, tess_batch_config* cfg);
/* All cells of all sets; m may be NULL unless config.learned = 1. */
tess_status tess_batch_compute(tess_batch* b, const tess_model* m);
/* Time the candidate configurations on this batch and keep the fastest:
 * first_capacity 64 or 128 x xcd_chunk 0, 256 or 4096, and with a model the
 * learned order at capacity 64 with xcd_chunk 0 or 256 (redo_chunk 1).
 * cache_path (nullable): a text file of earlier results keyed by device and
 * batch shape; a hit skips the timing, a miss appends a line. */
tess_status tess_batch_tune(tess_batch* b, const tess_model* m, const char* cache_path, tess_batch_config* best);
tess_status tess_batch_total_faces(const tess_batch* b, size_t set, uint64_t* out);
/* One set's results, laid out as tess_diagram_copy_cells. */
tess_status tess_batch_copy_cells(const tess_batch* b, size_t set, double* volume, uint32_t* face_count,
                                  uint64_t* face_offset, int32_t* neighbor, double* face_area);
tess_status tess_batch_timings(const tess_batch* b, tess_timings* out);

/* ---- without device allocation (no counterpart in the reference) ------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc.  The cells of n
 * device points (AoS xyz, n <= TESS_NOALLOC_MAX_POINTS, a grid of at most 32
 * cells per dimension) in the box [lo, hi], as tess_diagram_create_device +
 * tess_diagram_compute + tess_diagram_device_cells would give them, bitwise,
 * written into the caller's device buffers volume[n], face_count[n],
 * face_offset[n + 1], neighbor[face_cap], face_area[face_cap]; *total_faces
 * receives the face count (past face_cap: TESS_E_CAPACITY, neighbor and
 * face_area untouched).  Every temporary lives in one fixed device array of
 * the library, TESS_NOALLOC_SCRATCH_BYTES per GPU, which each GPU holds from
 * the library's first use until the process ends.  Calls on one GPU share
 * it: calls that overlap on one GPU race (one context and its stream at a
 * time per device).  Larger inputs are refused with TESS_E_INVALID. */
#define TESS_NOALLOC_MAX_POINTS 16384
#define TESS_NOALLOC_SCRATCH_BYTES 15460688
tess_status tess_compute_noalloc(tess_ctx* ctx, const double* xyz_device, size_t n, const double lo[3],
                                 const double hi[3], double cell_density, double* volume, uint32_t* face_count,
                                 uint64_t* face_offset, int32_t* neighbor, double* face_area, uint64_t face_cap,
                                 uint64_t* total_faces);

/* ---- diagnostics (no counterpart in the reference) --------------------- */
/* Device f64 sqrt, reciprocal and division of n inputs (for checking that
 * they are correctly rounded like the host's). */
tess_status tess_diag_arith(tess_ctx* ctx, const double* a, const double* b, size_t n,
                            double* sqrt_a, double* recip_a, double* a_div_b);

#ifdef __cplusplus
}
#endif
#endif /* TESS_H */
