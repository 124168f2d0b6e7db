// oracle.cpp -- CPU restatement of the-tessellator's per-cell Voronoi path.
//
// TEST INFRASTRUCTURE ONLY: this is the checker that the GPU library is
// compared against (and the CPU baseline the benchmark times).  Nothing in
// the product library links or calls it.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math, so
// every f64 operation is one IEEE-rounded operation in source order).
//
// Reference file:line citations are to mcomstock/the-tessellator.
//
// Decisions where the reference is broken or unspecified (the reference
// cannot run: its clip leaves dangling edges after one cut):
//  * tie order of the shell table: stable by generation order
//    (celery.rs:676 uses sort_unstable);
//  * order of points inside a grid cell: ascending input index
//    (celery.rs:366 uses sort_unstable_by);
//  * cube half-edge DR belongs to face D (polyhedron.rs:383 says B);
//  * the clip walk of polyhedron.rs:438-642 is kept (outgoing edge rule,
//    Incident copy, intersection(outside, inside), bridge edges) but the new
//    face is closed properly and cut-off parts are deleted exactly once;
//  * wall faces carry ids -1..-6 (-x,+x,-y,+y,-z,+z) instead of panicking
//    (polyhedron.rs:877);
//  * face order for the volume sum = creation order (walls in cube order,
//    then cut faces), face anchor = lexicographically smallest vertex of the
//    face (polyhedron.rs:776-808 anchors at the start edge's target, which
//    depends on pool slot reuse).  Both only change rounding.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kTol = 1e-12;  // polyhedron.rs:221-223
constexpr int OUT = 0, INC = 1, IN = 2;

struct V3 {
  double x, y, z;
};

// vector3.rs:38-40 -- evaluated left to right.
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vector3.rs:44-50
inline V3 cross(const V3& a, const V3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
// vector3.rs:53-59
inline V3 scale(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }

struct Plane {
  V3 n;
  double off;
};

// vector3.rs:223-239: n = p * (1/|p|), offset = n . (p * 0.5)
inline Plane halfway(const V3& p) {
  double m = std::sqrt(dot(p, p));
  double inv = 1.0 / m;
  V3 n = scale(p, inv);
  return {n, dot(n, scale(p, 0.5))};
}
// vector3.rs:170-178
inline int location(double s, double tol) { return s > tol ? OUT : (s < -tol ? IN : INC); }
// vector3.rs:191-193, :208-210
inline double signed_distance(const Plane& pl, const V3& v) { return dot(pl.n, v) - pl.off; }
// vector3.rs:213-219: a + (b - a) * ((off - n.a) / (n.b - n.a))
inline V3 intersection(const Plane& pl, const V3& a, const V3& b) {
  double ao = dot(pl.n, a);
  double bo = dot(pl.n, b);
  return add(a, scale(sub(b, a), (pl.off - ao) / (bo - ao)));
}

// lexicographic order on coordinates: canonical face anchor
inline bool lex_less(const V3& a, const V3& b) {
  if (a.x != b.x) return a.x < b.x;
  if (a.y != b.y) return a.y < b.y;
  return a.z < b.z;
}

// ---------------------------------------------------------------- polyhedron
struct HE {
  int flip, next, target, face;
};
struct Face {
  int64_t tag;
  int start;
};

struct Poly {
  std::vector<V3> v;
  std::vector<uint8_t> valive;
  std::vector<HE> e;
  std::vector<uint8_t> ealive;
  std::vector<Face> f;
  std::vector<uint8_t> falive;
  int nv = 0, ne = 0, nf = 0;  // live counts
  // scratch
  std::vector<uint8_t> loc, vdead;

  int add_vertex(const V3& p) {
    v.push_back(p);
    valive.push_back(1);
    ++nv;
    return (int)v.size() - 1;
  }
  int add_edge(const HE& h) {
    e.push_back(h);
    ealive.push_back(1);
    ++ne;
    return (int)e.size() - 1;
  }
  int add_face(int64_t tag, int start) {
    f.push_back({tag, start});
    falive.push_back(1);
    ++nf;
    return (int)f.size() - 1;
  }
};

// Wall ids of the container faces F(-y) R(+x) B(+y) L(-x) U(+z) D(-z).
constexpr int64_t kWallTag[6] = {-3, -2, -4, -1, -6, -5};

// Container cube, polyhedron.rs:268-392 (vertex, face and half-edge numbering
// kept; DR's face is D, fixing polyhedron.rs:383).
void build_cube(Poly& P, double x0, double y0, double z0, double x1, double y1, double z1) {
  P.v.clear();
  P.valive.clear();
  P.e.clear();
  P.ealive.clear();
  P.f.clear();
  P.falive.clear();
  P.nv = P.ne = P.nf = 0;
  // FDL FDR FUR FUL BDL BDR BUR BUL
  P.add_vertex({x0, y0, z0});
  P.add_vertex({x1, y0, z0});
  P.add_vertex({x1, y0, z1});
  P.add_vertex({x0, y0, z1});
  P.add_vertex({x0, y1, z0});
  P.add_vertex({x1, y1, z0});
  P.add_vertex({x1, y1, z1});
  P.add_vertex({x0, y1, z1});
  // {flip, target, next} per half-edge; 4 per face in face order F R B L U D.
  static const int tbl[24][3] = {
      {16, 3, 1},  {15, 0, 2},  {20, 1, 3},  {5, 2, 0},    // F: FU FL FD FR
      {17, 2, 5},  {3, 1, 6},   {23, 5, 7},  {9, 6, 4},    // R: RU RF RD RB
      {18, 6, 9},  {7, 5, 10},  {22, 4, 11}, {13, 7, 8},   // B: BU BR BD BL
      {19, 7, 13}, {11, 4, 14}, {21, 0, 15}, {1, 3, 12},   // L: LU LB LD LF
      {0, 2, 17},  {4, 6, 18},  {8, 7, 19},  {12, 3, 16},  // U: UF UR UB UL
      {2, 0, 21},  {14, 4, 22}, {10, 5, 23}, {6, 1, 20},   // D: DF DL DB DR
  };
  for (int fi = 0; fi < 6; ++fi) P.add_face(kWallTag[fi], 4 * fi);
  for (int h = 0; h < 24; ++h) P.add_edge({tbl[h][0], tbl[h][2], tbl[h][1], h / 4});
}

void translate(Poly& P, const V3& shift) {
  for (size_t i = 0; i < P.v.size(); ++i)
    if (P.valive[i]) P.v[i] = add(P.v[i], shift);  // polyhedron.rs:859-867
}

double max_r2(const Poly& P) {
  double r2 = 0.0;
  for (size_t i = 0; i < P.v.size(); ++i)
    if (P.valive[i]) {
      double d = dot(P.v[i], P.v[i]);
      if (d > r2) r2 = d;
    }
  return r2;
}

// drop dead slots when they dominate (keeps slot scans short)
void compact(Poly& P) {
  if (P.e.size() < 64 || P.e.size() < 3 * (size_t)P.ne) return;
  std::vector<int> vm(P.v.size(), -1), em(P.e.size(), -1), fm(P.f.size(), -1);
  std::vector<V3> v2;
  std::vector<HE> e2;
  std::vector<Face> f2;
  for (size_t i = 0; i < P.v.size(); ++i)
    if (P.valive[i]) {
      vm[i] = (int)v2.size();
      v2.push_back(P.v[i]);
    }
  for (size_t i = 0; i < P.e.size(); ++i)
    if (P.ealive[i]) {
      em[i] = (int)e2.size();
      e2.push_back(P.e[i]);
    }
  for (size_t i = 0; i < P.f.size(); ++i)
    if (P.falive[i]) {
      fm[i] = (int)f2.size();
      f2.push_back(P.f[i]);
    }
  for (auto& h : e2) {
    h.flip = em[h.flip];
    h.next = em[h.next];
    h.target = vm[h.target];
    h.face = fm[h.face];
  }
  for (auto& fc : f2) fc.start = em[fc.start];
  P.v.swap(v2);
  P.e.swap(e2);
  P.f.swap(f2);
  P.valive.assign(P.v.size(), 1);
  P.ealive.assign(P.e.size(), 1);
  P.falive.assign(P.f.size(), 1);
}

struct CutStats {
  uint64_t new_vertices = 0;
  int status = 0;
};

// find_outgoing_edge, polyhedron.rs:396-435: needs an Outside vertex, then the
// first half-edge (slot order) whose target is Inside and whose source is
// Outside; returns its flip (Inside -> Outside).
int find_outgoing(const Poly& P) {
  bool any_out = false;
  for (size_t i = 0; i < P.v.size(); ++i)
    if (P.valive[i] && P.loc[i] == OUT) {
      any_out = true;
      break;
    }
  if (!any_out) return -1;
  for (size_t h = 0; h < P.e.size(); ++h) {
    if (!P.ealive[h]) continue;
    const HE& eh = P.e[h];
    if (P.loc[eh.target] == IN && P.loc[P.e[eh.flip].target] == OUT) return eh.flip;
  }
  return -1;
}

void classify(Poly& P, const Plane& pl) {
  P.loc.resize(P.v.size());
  for (size_t i = 0; i < P.v.size(); ++i)
    if (P.valive[i]) P.loc[i] = (uint8_t)location(signed_distance(pl, P.v[i]), kTol);
}

// cut_with_plane, polyhedron.rs:438-642.  Returns true when the cell changed.
bool cut(Poly& P, int64_t tag, const Plane& pl, CutStats& cs) {
  classify(P, pl);
  const int first = find_outgoing(P);
  if (first < 0) return false;

  const int newface = P.add_face(tag, -1);
  std::vector<int> nf_edges;
  std::vector<int> dv;  // vertices passed by the walk (destroyed)
  int outgoing = first;
  int prev_int = -1;
  const int limit = 4 * (int)P.e.size() + 64;
  int guard = 0;
  for (;;) {
    int pv = P.e[outgoing].target;
    dv.push_back(pv);
    int ce = P.e[outgoing].next;
    int cv = P.e[ce].target;
    int ploc = P.loc[pv];
    bool need = (ploc == OUT);
    int cloc = P.loc[cv];
    // Keep going until back inside (polyhedron.rs:521-537).
    while (cloc != IN) {
      need = true;
      dv.push_back(cv);
      pv = cv;
      ploc = cloc;
      ce = P.e[ce].next;
      cv = P.e[ce].target;
      cloc = P.loc[cv];
      if (++guard > limit) {
        cs.status = 1;
        return false;
      }
    }
    if (prev_int >= 0) P.e[outgoing].target = prev_int;  // :550 (first one fixed below)
    if (need) {
      // Incident: copy the vertex (:555-565); else intersect outside -> inside (:567-570).
      int X = (ploc == INC) ? P.add_vertex(P.v[pv]) : P.add_vertex(intersection(pl, P.v[pv], P.v[cv]));
      if (ploc != INC) cs.new_vertices++;
      P.loc.push_back(INC);
      int fidx = P.e[outgoing].face;
      P.f[fidx].start = outgoing;  // :576-578
      int br = P.add_edge({-1, ce, X, fidx});
      int nf = P.add_edge({br, -1, prev_int, newface});
      P.e[br].flip = nf;
      P.e[outgoing].next = br;
      nf_edges.push_back(nf);
      prev_int = X;
    }
    outgoing = P.e[ce].flip;
    if (outgoing == first) break;
    if (++guard > limit) {
      cs.status = 2;
      return false;
    }
  }
  // Close the walk: the first face's entry is the last exit (fixes :550/:625-636),
  // and the new face is a closed loop, traversed opposite to the walk (fixes :478-598).
  P.e[first].target = prev_int;
  const int m = (int)nf_edges.size();
  P.e[nf_edges[0]].target = prev_int;
  for (int i = 0; i < m; ++i) P.e[nf_edges[i]].next = nf_edges[(i - 1 + m) % m];
  P.f[newface].start = nf_edges[0];

  // Delete the cut-off part exactly once (fixes :492, :645-730): every vertex
  // passed by the walk or Outside, every half-edge touching one of them, and
  // every face left without half-edges.
  P.vdead.assign(P.v.size(), 0);
  for (int x : dv) P.vdead[x] = 1;
  for (size_t i = 0; i < P.v.size(); ++i)
    if (P.valive[i] && P.loc[i] == OUT) P.vdead[i] = 1;
  for (size_t i = 0; i < P.v.size(); ++i)
    if (P.valive[i] && P.vdead[i]) {
      P.valive[i] = 0;
      --P.nv;
    }
  std::vector<uint8_t> fused(P.f.size(), 0);
  for (size_t h = 0; h < P.e.size(); ++h) {
    if (!P.ealive[h]) continue;
    const HE& eh = P.e[h];
    if (P.vdead[eh.target] || P.vdead[P.e[eh.flip].target]) {
      P.ealive[h] = 0;
      --P.ne;
    }
  }
  for (size_t h = 0; h < P.e.size(); ++h)
    if (P.ealive[h]) fused[P.e[h].face] = 1;
  for (size_t fi = 0; fi < P.f.size(); ++fi)
    if (P.falive[fi] && !fused[fi]) {
      P.falive[fi] = 0;
      --P.nf;
    }
  compact(P);
  return true;
}

bool check_poly(const Poly& P) {
  int ne = 0, nv = 0, nf = 0;
  std::vector<int> vseen(P.v.size(), 0);
  for (size_t h = 0; h < P.e.size(); ++h) {
    if (!P.ealive[h]) continue;
    ++ne;
    const HE& eh = P.e[h];
    if (eh.flip < 0 || eh.next < 0 || eh.target < 0 || eh.face < 0) return false;
    if (!P.ealive[eh.flip] || P.e[eh.flip].flip != (int)h) return false;
    if (!P.ealive[eh.next] || P.e[eh.next].face != eh.face) return false;
    if (!P.valive[eh.target] || !P.falive[eh.face]) return false;
    if (eh.target == P.e[eh.flip].target) return false;
    vseen[eh.target] = 1;
  }
  for (size_t i = 0; i < P.v.size(); ++i)
    if (P.valive[i]) {
      ++nv;
      if (!vseen[i]) return false;
    }
  for (size_t fi = 0; fi < P.f.size(); ++fi) {
    if (!P.falive[fi]) continue;
    ++nf;
    int s = P.f[fi].start;
    if (s < 0 || !P.ealive[s] || P.e[s].face != (int)fi) return false;
    int h = s, k = 0;
    do {
      h = P.e[h].next;
      if (++k > (int)P.e.size()) return false;
    } while (h != s);
    if (k < 3) return false;
  }
  if (ne % 2) return false;
  return nv - ne / 2 + nf == 2;
}

struct FaceOut {
  int64_t tag;
  double area;
  V3 wn;
  int degree;
};

// A face as its tag and vertex ring (counter-clockwise seen from outside).
struct Ring {
  int64_t tag;
  std::vector<V3> v;
};

// weighted_normal / compute_volume / compute_area (polyhedron.rs:776-855,
// interface.rs:408-410) over faces in creation order, each anchored at the
// lexicographically smallest vertex of its ring.
double ring_outputs(const std::vector<Ring>& rings, std::vector<FaceOut>* faces, uint64_t* flops) {
  double vol = 0.0;
  uint64_t fl = 0;
  for (const Ring& r : rings) {
    const int k = (int)r.v.size();
    int a = 0;
    for (int i = 1; i < k; ++i)
      if (lex_less(r.v[i], r.v[a])) a = i;
    const V3 A = r.v[a];
    V3 wn = {0.0, 0.0, 0.0};
    V3 cur = sub(r.v[(a + 1) % k], A);
    for (int j = 2; j < k; ++j) {
      V3 prev = cur;
      cur = sub(r.v[(a + j) % k], A);
      wn = add(wn, cross(prev, cur));
    }
    vol = vol + dot(A, wn);
    double area = 0.5 * std::sqrt(dot(wn, wn));
    fl += 3ull * (k - 1) + 12ull * (k - 2) + 6 + 7;
    if (faces) faces->push_back({r.tag, area, wn, k});
  }
  fl += 1;
  if (flops) *flops += fl;
  return vol / 6.0;
}

void poly_rings(const Poly& P, std::vector<Ring>& rings) {
  rings.clear();
  for (size_t fi = 0; fi < P.f.size(); ++fi) {
    if (!P.falive[fi]) continue;
    Ring r;
    r.tag = P.f[fi].tag;
    int s = P.f[fi].start, h = s;
    do {
      r.v.push_back(P.v[P.e[h].target]);
      h = P.e[h].next;
    } while (h != s);
    rings.push_back(std::move(r));
  }
}

// ------------------------------------------------------- per-face clipping
// The clip restated face by face.  Every face ring is clipped on its own
// (vertices classified by vector3.rs:170-193 with tolerance 1e-12; Outside
// runs replaced by their entry and exit points, where a point on an
// Inside-Outside edge is intersection(outside, inside) (vector3.rs:213-219,
// called as in polyhedron.rs:567-570) and an Incident end vertex is kept as
// is (polyhedron.rs:555-565)); the new face (tagged with the cutting point,
// polyhedron.rs:478-481) is the loop of exit -> entry segments.  For every
// cut the half-edge walk of polyhedron.rs:438-642 completes, this gives the
// same cell bit for bit; it also cuts an Outside region whose border is made
// of Incident vertices only, which polyhedron.rs:418-432 skips.
struct Soup {
  std::vector<Ring> f;  // creation order
  std::vector<Ring> tmp;
  std::vector<uint8_t> loc;
};

void soup_cube(Soup& S, double x0, double y0, double z0, double x1, double y1, double z1) {
  Poly P;
  build_cube(P, x0, y0, z0, x1, y1, z1);
  poly_rings(P, S.f);
}

void soup_translate(Soup& S, const V3& shift) {
  for (Ring& r : S.f)
    for (V3& v : r.v) v = add(v, shift);
}

double soup_r2(const Soup& S) {
  double r2 = 0.0;
  for (const Ring& r : S.f)
    for (const V3& v : r.v) {
      double d = dot(v, v);
      if (d > r2) r2 = d;
    }
  return r2;
}

// live vertex count from Euler (V = E - F + 2)
int soup_nv(const Soup& S) {
  int e2 = 0;
  for (const Ring& r : S.f) e2 += (int)r.v.size();
  return e2 / 2 - (int)S.f.size() + 2;
}

inline bool same(const V3& a, const V3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

bool cut_soup(Soup& S, int64_t tag, const Plane& pl, CutStats& cs) {
  bool any_out = false;
  for (const Ring& r : S.f) {
    for (const V3& v : r.v)
      if (location(signed_distance(pl, v), kTol) == OUT) {
        any_out = true;
        break;
      }
    if (any_out) break;
  }
  if (!any_out) return false;
  struct Seg {
    V3 x, e;
  };
  std::vector<Seg> segs;
  S.tmp.clear();
  uint64_t newv = 0;
  for (const Ring& r : S.f) {
    const int k = (int)r.v.size();
    S.loc.resize(k);
    int nout = 0, start = -1;
    for (int i = 0; i < k; ++i) {
      S.loc[i] = (uint8_t)location(signed_distance(pl, r.v[i]), kTol);
      if (S.loc[i] == OUT)
        ++nout;
      else if (start < 0)
        start = i;
    }
    if (nout == 0) {
      S.tmp.push_back(r);
      continue;
    }
    if (start < 0) continue;  // face entirely cut off
    Ring o;
    o.tag = r.tag;
    V3 entry = {0, 0, 0};
    for (int t = 0; t < k; ++t) {
      const int i = (start + t) % k, j = (start + t + 1) % k;
      if (S.loc[i] != OUT) {
        o.v.push_back(r.v[i]);
        if (S.loc[j] == OUT) {
          if (S.loc[i] == IN) {
            entry = intersection(pl, r.v[j], r.v[i]);
            o.v.push_back(entry);
          } else {
            entry = r.v[i];
          }
        }
      } else if (S.loc[j] != OUT) {
        V3 x = r.v[j];
        if (S.loc[j] == IN) {
          x = intersection(pl, r.v[i], r.v[j]);
          o.v.push_back(x);
          ++newv;  // each crossing edge is created once, on its exit side
        }
        if (!same(x, entry)) segs.push_back({x, entry});
      }
    }
    if (o.v.size() >= 3) S.tmp.push_back(std::move(o));
  }
  // new face: exit -> entry segments chained head to tail
  Ring nf;
  nf.tag = tag;
  const int m = (int)segs.size();
  if (m < 3) {
    cs.status = 3;
    return false;
  }
  std::vector<uint8_t> used(m, 0);
  int cur = 0;
  for (int c = 0; c < m; ++c) {
    used[cur] = 1;
    nf.v.push_back(segs[cur].x);
    int nxt = -1;
    for (int g = 0; g < m; ++g)
      if (same(segs[g].x, segs[cur].e)) {
        nxt = g;
        break;
      }
    if (nxt < 0) {
      cs.status = 4;
      return false;
    }
    if (nxt == 0) {
      if (c != m - 1) {
        cs.status = 5;
        return false;
      }
      break;
    }
    if (used[nxt]) {
      cs.status = 6;
      return false;
    }
    cur = nxt;
  }
  S.tmp.push_back(std::move(nf));
  S.f.swap(S.tmp);
  cs.new_vertices += newv;
  return true;
}

// ------------------------------------------------------------------- grid
int64_t icbrt_floor(double x) {
  if (!(x >= 1.0)) return 0;
  int64_t k = (int64_t)std::cbrt(x);
  while (k > 0 && (double)k * (double)k * (double)k > x) --k;
  while ((double)(k + 1) * (double)(k + 1) * (double)(k + 1) <= x) ++k;
  return k;
}

// Rust `f64 as usize`: saturating, NaN -> 0 (float.rs:138-141).
inline int64_t trunc_sat(double v) {
  if (!(v > 0.0)) return 0;
  if (v >= 9.2e18) return INT64_MAX;
  return (int64_t)v;
}

struct Grid {
  orc_grid g;
  std::vector<uint32_t> sorted;
  std::vector<uint64_t> delim;
  // canonical table (possibly truncated) + lazily built full table
  std::vector<double> key;
  std::vector<int32_t> off;
  double kcut = 0.0;
  std::vector<double> fkey;
  std::vector<int32_t> foff;
  std::once_flag full_once;
};

}  // namespace

extern "C" {

double orc_dot(const double a[3], const double b[3]) {
  return dot({a[0], a[1], a[2]}, {b[0], b[1], b[2]});
}
void orc_cross(const double a[3], const double b[3], double out[3]) {
  V3 c = cross({a[0], a[1], a[2]}, {b[0], b[1], b[2]});
  out[0] = c.x;
  out[1] = c.y;
  out[2] = c.z;
}
void orc_halfway_from_origin_to(const double p[3], double out4[4]) {
  Plane pl = halfway({p[0], p[1], p[2]});
  out4[0] = pl.n.x;
  out4[1] = pl.n.y;
  out4[2] = pl.n.z;
  out4[3] = pl.off;
}
// vector3.rs:243-248 + :229-239
void orc_build_from_non_unit_normal_and_point(const double n[3], const double pt[3], double out4[4]) {
  V3 nn = {n[0], n[1], n[2]};
  V3 u = scale(nn, 1.0 / std::sqrt(dot(nn, nn)));
  out4[0] = u.x;
  out4[1] = u.y;
  out4[2] = u.z;
  out4[3] = dot(u, {pt[0], pt[1], pt[2]});
}
int orc_location(double sd, double tol) { return location(sd, tol); }
int orc_vector_location(const double p4[4], const double v[3], double tol) {
  Plane pl = {{p4[0], p4[1], p4[2]}, p4[3]};
  return location(signed_distance(pl, {v[0], v[1], v[2]}), tol);
}
void orc_intersection(const double p4[4], const double a[3], const double b[3], double out[3]) {
  Plane pl = {{p4[0], p4[1], p4[2]}, p4[3]};
  V3 r = intersection(pl, {a[0], a[1], a[2]}, {b[0], b[1], b[2]});
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

// celery.rs:153-170: cpd = floor(cbrt(n / density)) + 1 (exact integer root).
int64_t orc_cells_per_dimension(size_t n, double density) {
  return icbrt_floor((double)n / density) + 1;
}

// celery.rs:81-125 (bounds of the points) + :153-181 (cell info).
void orc_grid_info(const double* x, const double* y, const double* z, size_t n, double density,
                   orc_grid* g) {
  double xmin = x[0], xmax = x[0], ymin = y[0], ymax = y[0], zmin = z[0], zmax = z[0];
  for (size_t i = 0; i < n; ++i) {
    if (x[i] < xmin) xmin = x[i];
    if (x[i] > xmax) xmax = x[i];
    if (y[i] < ymin) ymin = y[i];
    if (y[i] > ymax) ymax = y[i];
    if (z[i] < zmin) zmin = z[i];
    if (z[i] > zmax) zmax = z[i];
  }
  const int64_t cpd = orc_cells_per_dimension(n, density);
  const double c = (double)cpd;
  g->xmin = xmin;
  g->xmax = xmax;
  g->ymin = ymin;
  g->ymax = ymax;
  g->zmin = zmin;
  g->zmax = zmax;
  g->sx = (xmax - xmin) / c;
  g->sy = (ymax - ymin) / c;
  g->sz = (zmax - zmin) / c;
  g->ix = c / (xmax - xmin);
  g->iy = c / (ymax - ymin);
  g->iz = c / (zmax - zmin);
  g->cpd = cpd;
}

// celery.rs:269-282
int64_t orc_axis_index(double v, double vmin, double vmax, double inv, int64_t cpd) {
  if (v >= vmax) return cpd - 1;
  int64_t i = trunc_sat((v - vmin) * inv);
  return i < cpd - 1 ? i : cpd - 1;
}

static inline uint32_t bin_of(const orc_grid* g, double x, double y, double z, int64_t* ix,
                              int64_t* iy, int64_t* iz) {
  int64_t a = orc_axis_index(x, g->xmin, g->xmax, g->ix, g->cpd);
  int64_t b = orc_axis_index(y, g->ymin, g->ymax, g->iy, g->cpd);
  int64_t c = orc_axis_index(z, g->zmin, g->zmax, g->iz, g->cpd);
  if (ix) {
    *ix = a;
    *iy = b;
    *iz = c;
  }
  return (uint32_t)((a * g->cpd + b) * g->cpd + c);  // celery.rs:317-325
}

// celery.rs:342-414 with a stable counting sort (ascending input index in a cell).
void orc_grid_build(const double* x, const double* y, const double* z, size_t n, const orc_grid* g,
                    uint32_t* bins, uint32_t* sorted, uint64_t* delim) {
  const uint64_t nb = (uint64_t)g->cpd * g->cpd * g->cpd;
  std::vector<uint64_t> cnt(nb + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    bins[i] = bin_of(g, x[i], y[i], z[i], nullptr, nullptr, nullptr);
    cnt[bins[i] + 1]++;
  }
  for (uint64_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
  for (uint64_t b = 0; b <= nb; ++b) delim[b] = cnt[b];
  for (size_t i = 0; i < n; ++i) sorted[cnt[bins[i]]++] = (uint32_t)i;
}

// Diagram::morton_index (interface.rs:151-183): 10 levels on a 1024 grid over
// the diagram's bounding box, which starts as the default box at the origin
// and is grown to contain every point (interface.rs:52-57,
// vector3.rs:253-287).  The index per axis is `(offset / side) as usize`
// (saturating, NaN -> 0); bit 3l holds x's bit l, 3l+1 y's, 3l+2 z's.
void orc_morton_codes(const double* x, const double* y, const double* z, size_t n, uint32_t* codes) {
  double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
  for (size_t i = 0; i < n; ++i) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int a = 0; a < 3; ++a) {
      if (v[a] < lo[a]) lo[a] = v[a];
      if (v[a] > hi[a]) hi[a] = v[a];
    }
  }
  double side[3];
  for (int a = 0; a < 3; ++a) side[a] = (hi[a] - lo[a]) * (1.0 / 1024.0);
  for (size_t i = 0; i < n; ++i) {
    const double v[3] = {x[i], y[i], z[i]};
    uint64_t idx[3];
    for (int a = 0; a < 3; ++a) idx[a] = (uint64_t)trunc_sat((v[a] - lo[a]) / side[a]);
    uint32_t m = 0;
    int bit = 0;
    for (int level = 0; level < 10; ++level)
      for (int a = 0; a < 3; ++a) m |= (uint32_t)((idx[a] >> level) & 1u) << bit++;
    codes[i] = m;
  }
}

// sort_particles (interface.rs:88-148) as intended: internal order = points by
// Morton code, ties by input index (the reference's sort_unstable_by leaves
// them unspecified, and its in-place permutation loop indexes with codes).
void orc_index_maps(const double* x, const double* y, const double* z, size_t n, uint32_t* original_of_internal,
                    uint32_t* internal_of_original) {
  std::vector<uint32_t> code(n), ord(n);
  orc_morton_codes(x, y, z, n, code.data());
  for (size_t i = 0; i < n; ++i) ord[i] = (uint32_t)i;
  std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
  for (size_t k = 0; k < n; ++k) {
    if (original_of_internal) original_of_internal[k] = ord[k];
    if (internal_of_original) internal_of_original[ord[k]] = (uint32_t)k;
  }
}

// orc_grid_build with the points of a cell in internal (Morton) order when
// order == 1: the Celery of the reference is built over the sorted points
// (interface.rs:64, celery.rs:357-369), so its sorted indices ascend by
// internal index.  Ids stay original indices.
void orc_grid_build_ordered(const double* x, const double* y, const double* z, size_t n, const orc_grid* g,
                            int order, uint32_t* bins, uint32_t* sorted, uint64_t* delim) {
  orc_grid_build(x, y, z, n, g, bins, sorted, delim);
  if (order != 1) return;
  std::vector<uint32_t> ooi(n);
  orc_index_maps(x, y, z, n, ooi.data(), nullptr);
  const uint64_t nb = (uint64_t)g->cpd * g->cpd * g->cpd;
  std::vector<uint64_t> cur(delim, delim + nb);
  for (size_t k = 0; k < n; ++k) {
    const uint32_t i = ooi[k];
    sorted[cur[bins[i]]++] = i;
  }
}

// celery.rs:708-743.  sq(i sx) + sq(j sy) + sq(k sz) of the per-axis cell
// gaps max(0, |a - b| - 1), against radius * radius.
int orc_check_cell_in_range(const orc_grid* g, double x, double y, double z, double radius, int64_t i, int64_t j,
                            int64_t k) {
  int64_t xi, yi, zi;
  bin_of(g, x, y, z, &xi, &yi, &zi);
  auto gap = [](int64_t a, int64_t b) {
    const int64_t d = (a > b ? a - b : b - a) - 1;
    return d > 0 ? d : 0;
  };
  auto sq = [](double t) { return t * t; };
  const double ds = sq((double)gap(xi, i) * g->sx) + sq((double)gap(yi, j) * g->sy) + sq((double)gap(zi, k) * g->sz);
  return ds <= radius * radius;
}

// celery.rs:682-697: the first argument only when strictly smaller / larger.
static inline double min_float(double a, double b) { return a < b ? a : b; }
static inline double max_float(double a, double b) { return a > b ? a : b; }

// celery.rs:753-797.
size_t orc_find_cells_in_radius(const orc_grid* g, double x, double y, double z, double radius, uint32_t* out) {
  int64_t i0, j0, k0, i1, j1, k1;
  bin_of(g, max_float(x - radius, g->xmin), max_float(y - radius, g->ymin), max_float(z - radius, g->zmin), &i0, &j0,
         &k0);
  bin_of(g, min_float(x + radius, g->xmax), min_float(y + radius, g->ymax), min_float(z + radius, g->zmax), &i1, &j1,
         &k1);
  size_t c = 0;
  for (int64_t i = i0; i <= i1; ++i)
    for (int64_t j = j0; j <= j1; ++j)
      for (int64_t k = k0; k <= k1; ++k)
        if (orc_check_cell_in_range(g, x, y, z, radius, i, j, k)) {
          if (out) out[c] = (uint32_t)((i * g->cpd + j) * g->cpd + k);  // celery.rs:317-325
          ++c;
        }
  return c;
}

// celery.rs:802-819 (real == 0) and :825-855 (real != 0).
size_t orc_find_neighbors(const double* px, const double* py, const double* pz, const orc_grid* g,
                          const uint32_t* sorted, const uint64_t* delim, double x, double y, double z, double radius,
                          int real, uint32_t* out) {
  std::vector<uint32_t> cells(orc_find_cells_in_radius(g, x, y, z, radius, nullptr));
  orc_find_cells_in_radius(g, x, y, z, radius, cells.data());
  size_t c = 0;
  const double r2 = radius * radius;
  for (uint32_t b : cells)
    for (uint64_t s = delim[b]; s < delim[b + 1]; ++s) {
      const uint32_t p = sorted[s];
      if (real) {
        const double dx = x - px[p], dy = y - py[p], dz = z - pz[p];  // celery.rs:700-702
        if (!((dx * dx + dy * dy) + dz * dz <= r2)) continue;
      }
      if (out) out[c] = p;
      ++c;
    }
  return c;
}

// Position of an offset in the generation loops of celery.rs:437-673.
uint64_t orc_generation_index(int64_t cpd, int32_t ox, int32_t oy, int32_t oz) {
  const uint64_t M = (uint64_t)(cpd - 1);
  const uint64_t a = ox ? (uint64_t)std::abs(ox) - 1 : 0;
  const uint64_t b = oy ? (uint64_t)std::abs(oy) - 1 : 0;
  const uint64_t c = oz ? (uint64_t)std::abs(oz) - 1 : 0;
  const uint64_t base3 = 1, base2 = 1 + 8 * M * M * M, base1 = base2 + 12 * M * M;
  if (!ox && !oy && !oz) return 0;
  if (ox && oy && oz) {
    // (+,+,+) (+,+,-) (+,-,+) (-,+,+) (+,-,-) (-,+,-) (-,-,+) (-,-,-)
    static const int sidx[8] = {0, 1, 2, 4, 3, 5, 6, 7};  // by bits (x<0)<<2|(y<0)<<1|(z<0)
    int bits = ((ox < 0) << 2) | ((oy < 0) << 1) | (oz < 0);
    return base3 + 8 * ((a * M + b) * M + c) + sidx[bits];
  }
  if (ox && oy) return base2 + 4 * (a * M + b) + ((ox < 0) << 1 | (oy < 0));
  if (ox && oz) return base2 + 4 * M * M + 4 * (a * M + c) + ((ox < 0) << 1 | (oz < 0));
  if (oy && oz) return base2 + 8 * M * M + 4 * (b * M + c) + ((oy < 0) << 1 | (oz < 0));
  if (ox) return base1 + 2 * a + (ox < 0);
  if (oy) return base1 + 2 * M + 2 * b + (oy < 0);
  return base1 + 4 * M + 2 * c + (oz < 0);
}

// celery.rs:418-432: key = (sq(gx*sx) + sq(gy*sy)) + sq(gz*sz), gap g = max(|o|-1, 0).
static inline double shell_key(int32_t ox, int32_t oy, int32_t oz, double sx, double sy, double sz) {
  if (!ox && !oy && !oz) return -1.0;
  auto sq = [](double t) { return t * t; };
  double gx = (double)(ox ? std::abs(ox) - 1 : 0);
  double gy = (double)(oy ? std::abs(oy) - 1 : 0);
  double gz = (double)(oz ? std::abs(oz) - 1 : 0);
  return sq(gx * sx) + sq(gy * sy) + sq(gz * sz);
}

size_t orc_search_order(int64_t cpd, double sx, double sy, double sz, int64_t L, double* key,
                        int32_t* off, uint64_t* gen, double* kcut_out) {
  if (L > cpd - 1) L = cpd - 1;
  if (L < 0) L = 0;
  const bool full = (L == cpd - 1);
  double kcut = std::numeric_limits<double>::infinity();
  if (!full) {
    auto sq = [](double t) { return t * t; };
    double Ld = (double)L;
    kcut = std::min(sq(Ld * sx), std::min(sq(Ld * sy), sq(Ld * sz)));
  }
  struct Ent {
    double k;
    uint64_t g;
    int32_t o[3];
  };
  std::vector<Ent> ents;
  const int32_t l = (int32_t)L;
  ents.reserve((size_t)(2 * L + 1) * (2 * L + 1) * (2 * L + 1));
  for (int32_t i = -l; i <= l; ++i)
    for (int32_t j = -l; j <= l; ++j)
      for (int32_t k = -l; k <= l; ++k) {
        double kk = shell_key(i, j, k, sx, sy, sz);
        if (!full && !(kk < kcut)) continue;
        ents.push_back({kk, orc_generation_index(cpd, i, j, k), {i, j, k}});
      }
  std::sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    if (a.k != b.k) return a.k < b.k;
    return a.g < b.g;
  });
  if (key || off || gen)
    for (size_t t = 0; t < ents.size(); ++t) {
      if (key) key[t] = ents[t].k;
      if (off) {
        off[3 * t] = ents[t].o[0];
        off[3 * t + 1] = ents[t].o[1];
        off[3 * t + 2] = ents[t].o[2];
      }
      if (gen) gen[t] = ents[t].g;
    }
  if (kcut_out) *kcut_out = kcut;
  return ents.size();
}

}  // extern "C"

namespace {

constexpr int64_t kTableL = 16;

void grid_setup(Grid& G, const double* x, const double* y, const double* z, size_t n, double density,
                bool want_full, int order = 0) {
  orc_grid_info(x, y, z, n, density, &G.g);
  const uint64_t nb = (uint64_t)G.g.cpd * G.g.cpd * G.g.cpd;
  std::vector<uint32_t> bins(n);
  G.sorted.resize(n);
  G.delim.resize(nb + 1);
  orc_grid_build_ordered(x, y, z, n, &G.g, order, bins.data(), G.sorted.data(), G.delim.data());
  const int64_t L = want_full ? G.g.cpd - 1 : std::min<int64_t>(G.g.cpd - 1, kTableL);
  size_t m = orc_search_order(G.g.cpd, G.g.sx, G.g.sy, G.g.sz, L, nullptr, nullptr, nullptr, nullptr);
  G.key.resize(m);
  G.off.resize(3 * m);
  orc_search_order(G.g.cpd, G.g.sx, G.g.sy, G.g.sz, L, G.key.data(), G.off.data(), nullptr, &G.kcut);
}

void grid_full(Grid& G) {
  std::call_once(G.full_once, [&] {
    const int64_t L = G.g.cpd - 1;
    size_t m = orc_search_order(G.g.cpd, G.g.sx, G.g.sy, G.g.sz, L, nullptr, nullptr, nullptr, nullptr);
    G.fkey.resize(m);
    G.foff.resize(3 * m);
    orc_search_order(G.g.cpd, G.g.sx, G.g.sy, G.g.sz, L, G.fkey.data(), G.foff.data(), nullptr,
                     nullptr);
  });
}

// Security-radius threshold shared with the device: a candidate (or table
// entry) whose squared distance (key) exceeds (4 R^2)(1 + 1e-9) cannot put any
// vertex outside its bisector plane.
inline double stop_threshold(double r2) { return (4.0 * r2) * (1.0 + 1e-9); }

// The two clip implementations behind one interface.
struct WalkCell {  // half-edge walk of polyhedron.rs:438-642 (fixed)
  Poly P;
  void init(const double lo[3], const double hi[3], const V3& site) {
    build_cube(P, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
    translate(P, {-site.x, -site.y, -site.z});
  }
  bool cut_with(int64_t tag, const Plane& pl, CutStats& cs) { return cut(P, tag, pl, cs); }
  double r2() const { return max_r2(P); }
  int nv() const { return P.nv; }
  int nf() const { return P.nf; }
  int ne() const { return P.ne; }
  void rings(std::vector<Ring>& r) const { poly_rings(P, r); }
  bool check() const { return check_poly(P); }
};

struct FaceCell {  // per-face polygon clip
  Soup S;
  void init(const double lo[3], const double hi[3], const V3& site) {
    soup_cube(S, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
    soup_translate(S, {-site.x, -site.y, -site.z});
  }
  bool cut_with(int64_t tag, const Plane& pl, CutStats& cs) { return cut_soup(S, tag, pl, cs); }
  double r2() const { return soup_r2(S); }
  int nv() const { return soup_nv(S); }
  int nf() const { return (int)S.f.size(); }
  int ne() const {
    int e2 = 0;
    for (const Ring& r : S.f) e2 += (int)r.v.size();
    return e2;
  }
  void rings(std::vector<Ring>& r) const { r = S.f; }
  bool check() const {
    // closed 2-manifold: every directed edge appears once and its reverse once
    std::vector<std::pair<const V3*, const V3*>> es;
    for (const Ring& r : S.f) {
      if (r.v.size() < 3) return false;
      for (size_t i = 0; i < r.v.size(); ++i) es.push_back({&r.v[i], &r.v[(i + 1) % r.v.size()]});
    }
    for (auto& a : es) {
      int fwd = 0, rev = 0;
      for (auto& b : es) {
        if (same(*a.first, *b.first) && same(*a.second, *b.second)) ++fwd;
        if (same(*a.first, *b.second) && same(*a.second, *b.first)) ++rev;
      }
      if (fwd != 1 || rev != 1) return false;
    }
    return true;
  }
};

// compute_voronoi_cell (interface.rs:257-313) for point `i`, or for the
// outside site *ext (Diagram::get_cell_at_particle, interface.rs:211-232:
// no point is excluded as the cell's own).
struct Filter {
  double max_radius = std::numeric_limits<double>::quiet_NaN();  // expand_all_in_radius cut-off
  const int64_t* groups = nullptr;                               // interface.rs:280-297
  int64_t target = -1;
};

template <class Cell>
void compute_cell(Grid& G, const double* x, const double* y, const double* z, size_t i,
                  const double lo[3], const double hi[3], int mode, Cell& C,
                  std::vector<FaceOut>& faces, double* volume, orc_cell_stats* st, bool check,
                  const Filter& flt, const V3* ext = nullptr, std::vector<Ring>* rings_out = nullptr) {
  const V3 site = ext ? *ext : V3{x[i], y[i], z[i]};
  C.init(lo, hi, site);
  uint64_t flops = 24 + 40;
  double r2 = C.r2();
  int64_t ix, iy, iz;
  bin_of(&G.g, site.x, site.y, site.z, &ix, &iy, &iz);
  const int64_t cpd = G.g.cpd;
  CutStats cs;
  orc_cell_stats s;
  std::memset(&s, 0, sizeof(s));
  s.max_vertices = C.nv();
  s.max_faces = C.nf();
  s.max_edges = C.ne();
  s.stop_key = std::numeric_limits<double>::infinity();

  const bool all = (mode == ORC_MODE_ALL_POINTS);
  bool truncated = !all && std::isfinite(G.kcut);
  const double* key = G.key.data();
  const int32_t* off = G.off.data();
  size_t m = G.key.size();
  if (all && std::isfinite(G.kcut)) {
    grid_full(G);
    key = G.fkey.data();
    off = G.foff.data();
    m = G.fkey.size();
  }
  for (size_t t = 0;; ++t) {
    if (t == m) {
      if (!truncated) break;
      // The truncated table is a prefix of the canonical order and every
      // later entry has key >= kcut.
      if (G.kcut > stop_threshold(r2)) {
        s.stop_key = G.kcut;
        break;
      }
      grid_full(G);
      key = G.fkey.data();
      off = G.foff.data();
      m = G.fkey.size();
      truncated = false;
      if (t >= m) break;
    }
    if ((!all && key[t] > stop_threshold(r2)) || key[t] > flt.max_radius) {
      s.stop_key = key[t];
      break;
    }
    s.entries++;
    int64_t bx = ix + off[3 * t], by = iy + off[3 * t + 1], bz = iz + off[3 * t + 2];
    if (bx < 0 || bx >= cpd || by < 0 || by >= cpd || bz < 0 || bz >= cpd) continue;
    uint64_t b = (uint64_t)((bx * cpd + by) * cpd + bz);
    for (uint64_t k = G.delim[b]; k < G.delim[b + 1]; ++k) {
      uint32_t j = G.sorted[k];
      if (!ext && j == i) continue;  // self (interface.rs:299-305, by internal index)
      if (flt.groups && flt.target >= 0 && flt.groups[j] != flt.target) continue;
      s.streamed++;
      V3 p = {x[j] - site.x, y[j] - site.y, z[j] - site.z};
      double d2 = dot(p, p);
      flops += 8;
      if (!all && d2 > stop_threshold(r2)) continue;
      // halfway_from_origin_to (vector3.rs:223-225) reusing |p|^2 = dot(p,p)
      double mag = std::sqrt(d2);
      double inv = 1.0 / mag;
      V3 n = scale(p, inv);
      Plane pl = {n, dot(n, scale(p, 0.5))};
      s.classified++;
      flops += 13 + 6ull * C.nv();
      uint64_t nv0 = cs.new_vertices;
      if (C.cut_with((int64_t)j, pl, cs)) {
        s.cuts++;
        int nvv = C.nv();
        flops += 12ull * (cs.new_vertices - nv0) + 5ull * nvv;
        if (cs.new_vertices == nv0) s.incident_cuts++;  // no Inside-Outside edge was crossed
        r2 = C.r2();
        if (nvv > s.max_vertices) s.max_vertices = nvv;
        if (C.nf() > s.max_faces) s.max_faces = C.nf();
        if (C.ne() > s.max_edges) s.max_edges = C.ne();
      }
      if (cs.status) break;
    }
    if (cs.status) break;
  }
  faces.clear();
  std::vector<Ring> rings;
  C.rings(rings);
  *volume = ring_outputs(rings, &faces, &flops);
  std::sort(faces.begin(), faces.end(), [](const FaceOut& a, const FaceOut& b) { return a.tag < b.tag; });
  if (rings_out) {
    // compute_face_vertices (polyhedron.rs:895-919) per face in output (tag)
    // order, each ring from its lexicographically smallest vertex
    *rings_out = rings;
    std::sort(rings_out->begin(), rings_out->end(), [](const Ring& a, const Ring& b) { return a.tag < b.tag; });
    for (Ring& r : *rings_out) {
      size_t a = 0;
      for (size_t i = 1; i < r.v.size(); ++i)
        if (lex_less(r.v[i], r.v[a])) a = i;
      std::rotate(r.v.begin(), r.v.begin() + (long)a, r.v.end());
    }
  }
  s.new_vertices = cs.new_vertices;
  s.flops = flops;
  s.r2 = r2;
  s.status = cs.status ? cs.status : ((!check || C.check()) ? 0 : 10);
  if (st) *st = s;
}

}  // namespace

struct orc_poly {
  int clip = ORC_CLIP_WALK;
  Poly P;     // half-edge form (walk)
  Soup S;     // ring form (per-face clip)
};

struct orc_result {
  size_t ncells = 0;
  std::vector<uint64_t> ring_off;  // geometry: per face, into ring_xyz (3 doubles per entry)
  std::vector<double> ring_xyz;
  std::vector<double> volume;
  std::vector<uint32_t> fcount;
  std::vector<uint64_t> foff;
  std::vector<int32_t> nbr;
  std::vector<double> area;
  std::vector<orc_cell_stats> stats;
  double seconds = 0.0;
};

extern "C" {

// expand_all_no_radius (celery.rs:971-1018) when max_radius is NaN, else
// expand_all_in_radius (celery.rs:1023-1075): stop before the first entry
// whose key (a squared distance) exceeds max_radius as given.
size_t orc_expand_all(const double* x, const double* y, const double* z, size_t n, double density,
                      double sxp, double syp, double szp, double max_radius, uint32_t* out) {
  Grid G;
  grid_setup(G, x, y, z, n, density, true);
  int64_t ix, iy, iz;
  bin_of(&G.g, sxp, syp, szp, &ix, &iy, &iz);
  const int64_t cpd = G.g.cpd;
  size_t c = 0;
  for (size_t t = 0; t < G.key.size(); ++t) {
    if (G.key[t] > max_radius) break;
    int64_t bx = ix + G.off[3 * t], by = iy + G.off[3 * t + 1], bz = iz + G.off[3 * t + 2];
    if (bx < 0 || bx >= cpd || by < 0 || by >= cpd || bz < 0 || bz >= cpd) continue;
    uint64_t b = (uint64_t)((bx * cpd + by) * cpd + bz);
    for (uint64_t k = G.delim[b]; k < G.delim[b + 1]; ++k) out[c++] = G.sorted[k];
  }
  return c;
}

void orc_queries(int kind, const double* x, const double* y, const double* z, size_t n, double density,
                 const int64_t* groups, int64_t target_group, const double* q, size_t nq, const double* radius,
                 uint64_t* offsets, uint32_t* out) {
  orc_queries_ordered(kind, x, y, z, n, density, groups, target_group, q, nq, radius, offsets, out, 0);
}

void orc_queries_ordered(int kind, const double* x, const double* y, const double* z, size_t n, double density,
                         const int64_t* groups, int64_t target_group, const double* q, size_t nq,
                         const double* radius, uint64_t* offsets, uint32_t* out, int order) {
  Grid G;
  grid_setup(G, x, y, z, n, density, kind == 3, order);
  const int64_t cpd = G.g.cpd;
  uint64_t c = 0;
  std::vector<uint32_t> tmp;
  for (size_t a = 0; a < nq; ++a) {
    offsets[a] = c;
    const double qx = q[3 * a], qy = q[3 * a + 1], qz = q[3 * a + 2], r = radius[a];
    if (kind == 0) {
      const size_t m = orc_find_cells_in_radius(&G.g, qx, qy, qz, r, nullptr);
      if (out) orc_find_cells_in_radius(&G.g, qx, qy, qz, r, out + c);
      c += m;
    } else if (kind == 1 || kind == 2) {
      const size_t m = orc_find_neighbors(x, y, z, &G.g, G.sorted.data(), G.delim.data(), qx, qy, qz, r, kind == 2,
                                          nullptr);
      if (out)
        orc_find_neighbors(x, y, z, &G.g, G.sorted.data(), G.delim.data(), qx, qy, qz, r, kind == 2, out + c);
      c += m;
    } else {
      // celery.rs:1023-1075 from the bin of the query position.
      int64_t ix, iy, iz;
      bin_of(&G.g, qx, qy, qz, &ix, &iy, &iz);
      for (size_t t = 0; t < G.key.size(); ++t) {
        if (G.key[t] > r) break;
        const int64_t bx = ix + G.off[3 * t], by = iy + G.off[3 * t + 1], bz = iz + G.off[3 * t + 2];
        if (bx < 0 || bx >= cpd || by < 0 || by >= cpd || bz < 0 || bz >= cpd) continue;
        const uint64_t b = (uint64_t)((bx * cpd + by) * cpd + bz);
        for (uint64_t k = G.delim[b]; k < G.delim[b + 1]; ++k) {
          const uint32_t p = G.sorted[k];
          if (target_group >= 0 && groups[p] != target_group) continue;  // interface.rs:358-361
          if (out) out[c] = p;
          ++c;
        }
      }
    }
  }
  offsets[nq] = c;
}

orc_poly* orc_poly_new(double xmin, double ymin, double zmin, double xmax, double ymax, double zmax) {
  orc_poly* p = new orc_poly;
  build_cube(p->P, xmin, ymin, zmin, xmax, ymax, zmax);
  return p;
}
orc_poly* orc_poly_new_faces(double xmin, double ymin, double zmin, double xmax, double ymax,
                             double zmax) {
  orc_poly* p = new orc_poly;
  p->clip = ORC_CLIP_FACES;
  soup_cube(p->S, xmin, ymin, zmin, xmax, ymax, zmax);
  return p;
}
void orc_poly_free(orc_poly* p) { delete p; }
int orc_poly_cut(orc_poly* p, int64_t tag, const double p4[4]) {
  Plane pl = {{p4[0], p4[1], p4[2]}, p4[3]};
  CutStats cs;
  bool r = p->clip == ORC_CLIP_WALK ? cut(p->P, tag, pl, cs) : cut_soup(p->S, tag, pl, cs);
  if (cs.status) return -cs.status;
  return r ? 1 : 0;
}
int orc_poly_counts(const orc_poly* p, int* nv, int* ne, int* nf) {
  if (p->clip == ORC_CLIP_WALK) {
    *nv = p->P.nv;
    *ne = p->P.ne;
    *nf = p->P.nf;
  } else {
    FaceCell F;
    F.S = p->S;
    *nv = F.nv();
    *ne = F.ne();
    *nf = F.nf();
  }
  return 0;
}
static void poly_any_rings(const orc_poly* p, std::vector<Ring>& r) {
  if (p->clip == ORC_CLIP_WALK)
    poly_rings(p->P, r);
  else
    r = p->S.f;
}
double orc_poly_volume(const orc_poly* p) {
  std::vector<Ring> r;
  poly_any_rings(p, r);
  return ring_outputs(r, nullptr, nullptr);
}
int orc_poly_faces(const orc_poly* p, int64_t* tag, double* area, double* wn3, int* degree, int cap) {
  std::vector<FaceOut> faces;
  std::vector<Ring> r;
  poly_any_rings(p, r);
  ring_outputs(r, &faces, nullptr);
  int k = 0;
  for (const auto& f : faces) {
    if (k >= cap) break;
    tag[k] = f.tag;
    area[k] = f.area;
    wn3[3 * k] = f.wn.x;
    wn3[3 * k + 1] = f.wn.y;
    wn3[3 * k + 2] = f.wn.z;
    degree[k] = f.degree;
    ++k;
  }
  return (int)faces.size();
}
int orc_poly_vertices(const orc_poly* p, double* xyz, int cap) {
  int k = 0;
  if (p->clip != ORC_CLIP_WALK) {
    std::vector<V3> u;
    for (const Ring& r : p->S.f)
      for (const V3& v : r.v) {
        bool dup = false;
        for (const V3& w : u) dup = dup || same(v, w);
        if (!dup) u.push_back(v);
      }
    for (const V3& v : u) {
      if (k < cap) {
        xyz[3 * k] = v.x;
        xyz[3 * k + 1] = v.y;
        xyz[3 * k + 2] = v.z;
      }
      ++k;
    }
    return k;
  }
  for (size_t i = 0; i < p->P.v.size(); ++i)
    if (p->P.valive[i]) {
      if (k < cap) {
        xyz[3 * k] = p->P.v[i].x;
        xyz[3 * k + 1] = p->P.v[i].y;
        xyz[3 * k + 2] = p->P.v[i].z;
      }
      ++k;
    }
  return k;
}
int orc_poly_check(const orc_poly* p) {
  if (p->clip == ORC_CLIP_WALK) return check_poly(p->P) ? 1 : 0;
  FaceCell F;
  F.S = p->S;
  return F.check() ? 1 : 0;
}
int orc_poly_outgoing(const orc_poly* p, const double p4[4], double src[3], double dst[3]) {
  Poly& P = const_cast<Poly&>(p->P);
  Plane pl = {{p4[0], p4[1], p4[2]}, p4[3]};
  classify(P, pl);
  int h = find_outgoing(P);
  if (h < 0) return 0;
  V3 d = P.v[P.e[h].target], s = P.v[P.e[P.e[h].flip].target];
  src[0] = s.x;
  src[1] = s.y;
  src[2] = s.z;
  dst[0] = d.x;
  dst[1] = d.y;
  dst[2] = d.z;
  return 1;
}

orc_result* orc_compute(const double* x, const double* y, const double* z, size_t n,
                        const double lo[3], const double hi[3], double density, int mode, int clip,
                        const uint32_t* cells, size_t ncells, int nthreads) {
  return orc_compute_query(x, y, z, n, lo, hi, density, mode, clip, cells, ncells, nthreads,
                           std::numeric_limits<double>::quiet_NaN(), nullptr, -1);
}

orc_result* orc_compute_query(const double* x, const double* y, const double* z, size_t n,
                              const double lo[3], const double hi[3], double density, int mode, int clip,
                              const uint32_t* cells, size_t ncells, int nthreads, double max_radius,
                              const int64_t* groups, int64_t target_group) {
  return orc_compute_ordered(x, y, z, n, lo, hi, density, mode, clip, cells, ncells, nthreads, max_radius, groups,
                             target_group, 0);
}

orc_result* orc_compute_ordered(const double* x, const double* y, const double* z, size_t n,
                                const double lo[3], const double hi[3], double density, int mode, int clip,
                                const uint32_t* cells, size_t ncells, int nthreads, double max_radius,
                                const int64_t* groups, int64_t target_group, int order) {
  Filter flt;
  flt.max_radius = max_radius;
  flt.groups = groups;
  flt.target = target_group;
  orc_result* R = new orc_result;
  if (n == 0) return R;
  auto t0 = std::chrono::steady_clock::now();
  Grid G;
  grid_setup(G, x, y, z, n, density, mode == ORC_MODE_ALL_POINTS, order);
  if (!cells) ncells = n;
  R->ncells = ncells;
  R->volume.resize(ncells);
  R->fcount.resize(ncells);
  R->stats.resize(ncells);
  std::vector<std::vector<FaceOut>> per(ncells);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    WalkCell W;
    FaceCell F;
    std::vector<FaceOut> faces;
    const bool check = (clip & ORC_CHECK) != 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
    for (size_t c = 0; c < ncells; ++c) {
      size_t i = cells ? cells[c] : c;
      if ((clip & 1) == ORC_CLIP_WALK)
        compute_cell(G, x, y, z, i, lo, hi, mode, W, faces, &R->volume[c], &R->stats[c], check, flt);
      else
        compute_cell(G, x, y, z, i, lo, hi, mode, F, faces, &R->volume[c], &R->stats[c], check, flt);
      per[c] = faces;
      R->fcount[c] = (uint32_t)faces.size();
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  R->seconds = std::chrono::duration<double>(t1 - t0).count();
  R->foff.resize(ncells + 1);
  R->foff[0] = 0;
  for (size_t c = 0; c < ncells; ++c) R->foff[c + 1] = R->foff[c] + R->fcount[c];
  R->nbr.resize(R->foff[ncells]);
  R->area.resize(R->foff[ncells]);
  for (size_t c = 0; c < ncells; ++c)
    for (size_t k = 0; k < per[c].size(); ++k) {
      R->nbr[R->foff[c] + k] = (int32_t)per[c][k].tag;
      R->area[R->foff[c] + k] = per[c][k].area;
    }
  return R;
}

orc_result* orc_compute_at_points(const double* x, const double* y, const double* z, size_t n,
                                  const double lo[3], const double hi[3], double density, int mode, int clip,
                                  const double* sites, size_t nsites, int nthreads, double max_radius,
                                  const int64_t* groups, int64_t target_group) {
  return orc_compute_at_points_ordered(x, y, z, n, lo, hi, density, mode, clip, sites, nsites, nthreads, max_radius,
                                       groups, target_group, 0);
}

orc_result* orc_compute_at_points_ordered(const double* x, const double* y, const double* z, size_t n,
                                          const double lo[3], const double hi[3], double density, int mode,
                                          int clip, const double* sites, size_t nsites, int nthreads,
                                          double max_radius, const int64_t* groups, int64_t target_group,
                                          int order) {
  Filter flt;
  flt.max_radius = max_radius;
  flt.groups = groups;
  flt.target = target_group;
  orc_result* R = new orc_result;
  if (n == 0) return R;
  auto t0 = std::chrono::steady_clock::now();
  Grid G;
  grid_setup(G, x, y, z, n, density, mode == ORC_MODE_ALL_POINTS, order);
  const size_t ncells = nsites;
  R->ncells = ncells;
  R->volume.resize(ncells);
  R->fcount.resize(ncells);
  R->stats.resize(ncells);
  std::vector<std::vector<FaceOut>> per(ncells);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    WalkCell W;
    FaceCell F;
    std::vector<FaceOut> faces;
    const bool check = (clip & ORC_CHECK) != 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
    for (size_t c = 0; c < ncells; ++c) {
      const V3 site = {sites[3 * c], sites[3 * c + 1], sites[3 * c + 2]};
      if ((clip & 1) == ORC_CLIP_WALK)
        compute_cell(G, x, y, z, 0, lo, hi, mode, W, faces, &R->volume[c], &R->stats[c], check, flt, &site);
      else
        compute_cell(G, x, y, z, 0, lo, hi, mode, F, faces, &R->volume[c], &R->stats[c], check, flt, &site);
      per[c] = faces;
      R->fcount[c] = (uint32_t)faces.size();
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  R->seconds = std::chrono::duration<double>(t1 - t0).count();
  R->foff.resize(ncells + 1);
  R->foff[0] = 0;
  for (size_t c = 0; c < ncells; ++c) R->foff[c + 1] = R->foff[c] + R->fcount[c];
  R->nbr.resize(R->foff[ncells]);
  R->area.resize(R->foff[ncells]);
  for (size_t c = 0; c < ncells; ++c)
    for (size_t k = 0; k < per[c].size(); ++k) {
      R->nbr[R->foff[c] + k] = (int32_t)per[c][k].tag;
      R->area[R->foff[c] + k] = per[c][k].area;
    }
  return R;
}

orc_result* orc_compute_geometry(const double* x, const double* y, const double* z, size_t n, const double lo[3],
                                 const double hi[3], double density, const uint32_t* cells, size_t ncells,
                                 int nthreads) {
  return orc_compute_geometry_ordered(x, y, z, n, lo, hi, density, cells, ncells, nthreads, 0);
}

orc_result* orc_compute_geometry_ordered(const double* x, const double* y, const double* z, size_t n,
                                         const double lo[3], const double hi[3], double density,
                                         const uint32_t* cells, size_t ncells, int nthreads, int order) {
  orc_result* R = orc_compute_ordered(x, y, z, n, lo, hi, density, ORC_MODE_SECURITY, ORC_CLIP_FACES, cells, ncells,
                                      nthreads, std::numeric_limits<double>::quiet_NaN(), nullptr, -1, order);
  if (n == 0) return R;
  Grid G;
  grid_setup(G, x, y, z, n, density, false, order);
  const size_t nc = R->ncells;
  std::vector<std::vector<Ring>> per(nc);
  Filter flt;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    FaceCell F;
    std::vector<FaceOut> faces;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
    for (size_t c = 0; c < nc; ++c) {
      double vol;
      orc_cell_stats st;
      compute_cell(G, x, y, z, cells ? cells[c] : c, lo, hi, ORC_MODE_SECURITY, F, faces, &vol, &st, false, flt,
                   nullptr, &per[c]);
    }
  }
  R->ring_off.assign(1, 0);
  for (size_t c = 0; c < nc; ++c)
    for (const Ring& r : per[c]) {
      for (const V3& v : r.v) {
        R->ring_xyz.push_back(v.x);
        R->ring_xyz.push_back(v.y);
        R->ring_xyz.push_back(v.z);
      }
      R->ring_off.push_back(R->ring_xyz.size() / 3);
    }
  return R;
}

uint64_t orc_result_ring_entries(const orc_result* R) { return R->ring_xyz.size() / 3; }

void orc_result_copy_rings(const orc_result* R, uint64_t* ring_offset, double* xyz) {
  if (ring_offset) std::copy(R->ring_off.begin(), R->ring_off.end(), ring_offset);
  if (xyz) std::copy(R->ring_xyz.begin(), R->ring_xyz.end(), xyz);
}
size_t orc_result_cells(const orc_result* R) { return R->ncells; }
uint64_t orc_result_total_faces(const orc_result* R) { return R->ncells ? R->foff[R->ncells] : 0; }
void orc_result_copy(const orc_result* R, double* volume, uint32_t* face_count, uint64_t* face_offset,
                     int32_t* neighbor, double* area) {
  if (volume) std::copy(R->volume.begin(), R->volume.end(), volume);
  if (face_count) std::copy(R->fcount.begin(), R->fcount.end(), face_count);
  if (face_offset) std::copy(R->foff.begin(), R->foff.end(), face_offset);
  if (neighbor) std::copy(R->nbr.begin(), R->nbr.end(), neighbor);
  if (area) std::copy(R->area.begin(), R->area.end(), area);
}
void orc_result_stats(const orc_result* R, orc_cell_stats* st) {
  std::copy(R->stats.begin(), R->stats.end(), st);
}
double orc_result_seconds(const orc_result* R) { return R->seconds; }
void orc_result_free(orc_result* R) { delete R; }

}  // extern "C"
