/* oracle.h -- CPU restatement of the-tessellator's per-cell Voronoi path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark call it; the product path
 * (libtess) never links, loads or calls it.
 *
 * The reference (mcomstock/the-tessellator, Rust) cannot be built in this
 * image and its clip panics after one cut (polyhedron.rs:438-642), so this is
 * a restatement of its arithmetic contract with the defects of the clip fixed:
 *   grid        celery.rs:81-188, :269-414
 *   shell order celery.rs:418-679 (ties: stable by generation order)
 *   stream      celery.rs:971-1018 (expand_all_no_radius)
 *   plane math  vector3.rs:38-77, :170-239 ; tolerance polyhedron.rs:221-223
 *   clip        polyhedron.rs:396-642, twice: the half-edge walk with its
 *               dangling-edge/free-list/recursion defects fixed, and a
 *               per-face polygon clip with the same arithmetic that also
 *               cuts Outside regions bordered only by Incident vertices
 *   outputs     polyhedron.rs:776-855, interface.rs:408-410
 */
#ifndef TESS_ORACLE_H
#define TESS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- vector / plane known answers (vector3.rs) ---- */
double orc_dot(const double a[3], const double b[3]);
void orc_cross(const double a[3], const double b[3], double out[3]);
/* plane: out4 = {nx, ny, nz, offset} */
void orc_halfway_from_origin_to(const double p[3], double out4[4]);
void orc_build_from_non_unit_normal_and_point(const double n[3], const double pt[3], double out4[4]);
/* 0 = Outside, 1 = Incident, 2 = Inside (vector3.rs:170-185) */
int orc_location(double signed_distance, double tol);
int orc_vector_location(const double plane4[4], const double v[3], double tol);
void orc_intersection(const double plane4[4], const double a[3], const double b[3], double out[3]);

/* ---- grid (celery.rs) ---- */
typedef struct {
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double sx, sy, sz;      /* cell sizes */
  double ix, iy, iz;      /* inverse cell sizes */
  int64_t cpd;            /* cells per dimension */
} orc_grid;

int64_t orc_cells_per_dimension(size_t n, double density);
/* bounds + cell info of n SoA points. */
void orc_grid_info(const double* x, const double* y, const double* z, size_t n, double density,
                   orc_grid* g);
/* per-axis cell index (celery.rs:269-314) */
int64_t orc_axis_index(double v, double vmin, double vmax, double inv, int64_t cpd);
/* bin id of every point, stable order by bin, CSR delimiters (length cpd^3+1). */
void orc_grid_build(const double* x, const double* y, const double* z, size_t n, const orc_grid* g,
                    uint32_t* bins, uint32_t* sorted, uint64_t* delim);

/* canonical search order restricted to |o|_inf <= L (L = cpd-1 gives the full
 * table).  Returns the entry count; if key/off are non-NULL they are filled
 * (off: 3 int32 per entry).  When L < cpd-1 only entries with key < kcut are
 * kept, which is a prefix of the full canonical order; *kcut_out receives the
 * cut (+inf for the full table). */
size_t orc_search_order(int64_t cpd, double sx, double sy, double sz, int64_t L, double* key,
                        int32_t* off, uint64_t* gen, double* kcut_out);
/* generation index of an offset in the reference's loops (celery.rs:437-673) */
uint64_t orc_generation_index(int64_t cpd, int32_t ox, int32_t oy, int32_t oz);

/* Candidate stream of a site: expand_all_no_radius (max_radius NaN, all n
 * points) or expand_all_in_radius(max_radius). Returns the count. */
/* Diagram::morton_index (interface.rs:151-183) of every point: 10 levels on a
 * 1024 grid over the points' bounding box grown from the origin. */
void orc_morton_codes(const double* x, const double* y, const double* z, size_t n, uint32_t* codes);
/* sort_particles (interface.rs:88-148) as intended: internal order by Morton
 * code, ties by input index; either output may be NULL. */
void orc_index_maps(const double* x, const double* y, const double* z, size_t n, uint32_t* original_of_internal,
                    uint32_t* internal_of_original);
/* orc_grid_build; order 1: a cell's points in internal (Morton) order. */
void orc_grid_build_ordered(const double* x, const double* y, const double* z, size_t n, const orc_grid* g,
                            int order, uint32_t* bins, uint32_t* sorted, uint64_t* delim);

/* Radius queries (celery.rs:700-855) on a built grid.  check_cell_in_range
 * (celery.rs:708-743): is grid cell (i, j, k) within radius of the cell of
 * (x, y, z), counting touching cells as distance 0.  find_cells_in_radius
 * (celery.rs:753-797): cell ids in i, j, k loop order over the clamped box.
 * find_neighbors (celery.rs:802-855): the points of those cells in cell
 * order then sorted order; real != 0 keeps only points with squared
 * distance <= radius^2.  out may be NULL (count only). */
int orc_check_cell_in_range(const orc_grid* g, double x, double y, double z, double radius, int64_t i, int64_t j,
                            int64_t k);
size_t orc_find_cells_in_radius(const orc_grid* g, double x, double y, double z, double radius, uint32_t* out);
size_t orc_find_neighbors(const double* px, const double* py, const double* pz, const orc_grid* g,
                          const uint32_t* sorted, const uint64_t* delim, double x, double y, double z, double radius,
                          int real, uint32_t* out);
/* Many radius queries over one grid built from the points (celery.rs:231-250):
 * kind 0 find_cells_in_radius, 1 find_neighbors_in_cell_radius, 2
 * find_neighbors_in_real_radius, 3 Cell::compute_neighbor_cloud
 * (interface.rs:348-365: expand_all_in_radius from the query position, then
 * the target-group filter when target_group >= 0; groups may be NULL then).
 * q: nq AoS positions; radius: nq radii.  offsets (nq + 1) are always
 * written; out (offsets[nq] entries) only when not NULL. */
void orc_queries(int kind, const double* x, const double* y, const double* z, size_t n, double density,
                 const int64_t* groups, int64_t target_group, const double* q, size_t nq, const double* radius,
                 uint64_t* offsets, uint32_t* out);
/* orc_queries over a grid whose cells list their points in the given order
 * (0: input index, 1: internal Morton order, see orc_grid_build_ordered). */
void orc_queries_ordered(int kind, const double* x, const double* y, const double* z, size_t n, double density,
                         const int64_t* groups, int64_t target_group, const double* q, size_t nq,
                         const double* radius, uint64_t* offsets, uint32_t* out, int order);
size_t orc_expand_all(const double* x, const double* y, const double* z, size_t n, double density,
                      double sxp, double syp, double szp, double max_radius, uint32_t* out);

/* ---- single polyhedron (polyhedron.rs) ---- */
typedef struct orc_poly orc_poly;
/* container cube in the reference's half-edge form (clip = walk) ... */
orc_poly* orc_poly_new(double xmin, double ymin, double zmin, double xmax, double ymax, double zmax);
/* ... or as face rings (clip = per-face). */
orc_poly* orc_poly_new_faces(double xmin, double ymin, double zmin, double xmax, double ymax, double zmax);
void orc_poly_free(orc_poly*);
int orc_poly_cut(orc_poly*, int64_t tag, const double plane4[4]);   /* 1 = cut */
int orc_poly_counts(const orc_poly*, int* nv, int* ne, int* nf);    /* live counts */
double orc_poly_volume(const orc_poly*);
/* faces in canonical order (creation order): tag, area, weighted normal, degree. Returns count. */
int orc_poly_faces(const orc_poly*, int64_t* tag, double* area, double* wn3, int* degree, int cap);
/* live vertices (slot order). */
int orc_poly_vertices(const orc_poly*, double* xyz, int cap);
/* 1 if every half-edge is consistent (flip/next/face), Euler V-E+F == 2. */
int orc_poly_check(const orc_poly*);
/* find_outgoing_edge: returns 1 and the source/target coordinates if found. */
int orc_poly_outgoing(const orc_poly*, const double plane4[4], double src[3], double dst[3]);

/* ---- whole-cell computation (interface.rs:257-344) ---- */
/* stream: every point in shell order (the reference's expand_all_no_radius)
 * or shell order with the security-radius stop (bit-identical results). */
enum { ORC_MODE_ALL_POINTS = 0, ORC_MODE_SECURITY = 1 };
/* clip: per-face polygon clip (default) or the half-edge walk of
 * polyhedron.rs:438-642; ORC_CHECK adds a manifold check per cell. */
enum { ORC_CLIP_FACES = 0, ORC_CLIP_WALK = 1, ORC_CHECK = 2 };

/* per-cell instrumentation */
typedef struct {
  uint64_t streamed;    /* candidates taken from the stream (self excluded) */
  uint64_t classified;  /* candidates whose plane was built and vertices classified */
  uint64_t cuts;        /* candidates that changed the cell */
  uint64_t new_vertices;/* intersection vertices created */
  uint64_t entries;     /* shell-table entries visited (incl. out-of-grid) */
  uint64_t flops;       /* algorithmic f64 ops (see oracle.cpp: count_*) */
  double r2;            /* final max |v|^2 */
  double stop_key;      /* key of the first unvisited table entry (+inf if exhausted) */
  int32_t max_vertices; /* largest live vertex count reached */
  int32_t max_faces;
  int32_t max_edges;    /* largest live half-edge count reached */
  int32_t status;       /* 0 ok, nonzero = topology failure */
  int32_t incident_cuts;/* cuts that created no intersection vertex: the Outside
                           region was bordered by Incident vertices only, which
                           polyhedron.rs:418-432 (find_outgoing_edge) skips */
  int32_t reserved;
} orc_cell_stats;

typedef struct orc_result orc_result;
/* Computes the cells of `cells` (or all n when cells == NULL) of the point set
 * in the box [lo, hi].  density: points per grid cell (1.25 = reference).
 * nthreads <= 0: all cores. */
orc_result* orc_compute(const double* x, const double* y, const double* z, size_t n,
                        const double lo[3], const double hi[3], double density, int mode, int clip,
                        const uint32_t* cells, size_t ncells, int nthreads);
/* Same with the reference's query options: max_radius (NaN = None) stops the
 * stream at the first shell key above it (celery.rs:1023-1075) and
 * target_group >= 0 keeps only points of that group (interface.rs:280-297). */
/* Cells of sites that are not points of the set (Diagram::get_cell_at_particle,
 * interface.rs:211-232, + compute_voronoi_cell): nsites AoS positions; no
 * point is excluded as the cell's own. */
orc_result* orc_compute_at_points(const double* x, const double* y, const double* z, size_t n,
                                  const double lo[3], const double hi[3], double density, int mode, int clip,
                                  const double* sites, size_t nsites, int nthreads, double max_radius,
                                  const int64_t* groups, int64_t target_group);
/* orc_compute_at_points with the candidate order of a grid cell chosen by
 * order (as orc_compute_ordered). */
orc_result* orc_compute_at_points_ordered(const double* x, const double* y, const double* z, size_t n,
                                          const double lo[3], const double hi[3], double density, int mode,
                                          int clip, const double* sites, size_t nsites, int nthreads,
                                          double max_radius, const int64_t* groups, int64_t target_group,
                                          int order);
/* orc_compute (security stop, per-face clip) plus the geometry of every cell:
 * the vertex ring of each face (compute_face_vertices, polyhedron.rs:895-919),
 * faces in output order, each ring from its lexicographically smallest
 * vertex, counter-clockwise seen from outside, site-relative. */
orc_result* orc_compute_geometry(const double* x, const double* y, const double* z, size_t n, const double lo[3],
                                 const double hi[3], double density, const uint32_t* cells, size_t ncells,
                                 int nthreads);
/* orc_compute_geometry with the candidate order of a grid cell chosen by
 * order (as orc_compute_ordered). */
orc_result* orc_compute_geometry_ordered(const double* x, const double* y, const double* z, size_t n,
                                         const double lo[3], const double hi[3], double density,
                                         const uint32_t* cells, size_t ncells, int nthreads, int order);
uint64_t orc_result_ring_entries(const orc_result* R);
/* ring_offset: total faces + 1 entries (into the entries); xyz: 3 per entry */
void orc_result_copy_rings(const orc_result* R, uint64_t* ring_offset, double* xyz);
/* orc_compute_query with the candidate order of a grid cell chosen by order
 * (0: input index, 1: internal Morton order, see orc_grid_build_ordered). */
orc_result* orc_compute_ordered(const double* x, const double* y, const double* z, size_t n,
                                const double lo[3], const double hi[3], double density, int mode, int clip,
                                const uint32_t* cells, size_t ncells, int nthreads, double max_radius,
                                const int64_t* groups, int64_t target_group, int order);
orc_result* orc_compute_query(const double* x, const double* y, const double* z, size_t n,
                              const double lo[3], const double hi[3], double density, int mode, int clip,
                              const uint32_t* cells, size_t ncells, int nthreads, double max_radius,
                              const int64_t* groups, int64_t target_group);
size_t orc_result_cells(const orc_result*);
uint64_t orc_result_total_faces(const orc_result*);
/* CSR copy; face_offset has ncells+1 entries. */
void orc_result_copy(const orc_result*, double* volume, uint32_t* face_count, uint64_t* face_offset,
                     int32_t* neighbor, double* area);
void orc_result_stats(const orc_result*, orc_cell_stats* stats);
double orc_result_seconds(const orc_result*);
void orc_result_free(orc_result*);

#ifdef __cplusplus
}
#endif
#endif
