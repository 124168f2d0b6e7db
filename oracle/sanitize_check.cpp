// sanitize_check.cpp -- the oracle (test infrastructure) under AddressSanitizer
// and UndefinedBehaviorSanitizer: a few small diagrams of each kind the tests
// use (uniform, clustered corner blobs with a void, a site ringed by points
// on a sphere, a container larger than the points, a jittered lattice, the
// walk clip and the all-points mode), each checked for the volume sum.
// Built and run by `make -C oracle sanitize` (exit status 0 = clean).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "oracle.h"

namespace {

struct Set {
  const char* name;
  std::vector<double> x, y, z;
  double lo[3], hi[3];
};

double box_volume(const Set& s) { return (s.hi[0] - s.lo[0]) * (s.hi[1] - s.lo[1]) * (s.hi[2] - s.lo[2]); }

int check(const Set& s, int mode, int clip) {
  orc_result* r = orc_compute(s.x.data(), s.y.data(), s.z.data(), s.x.size(), s.lo, s.hi, 1.25, mode, clip, nullptr,
                              0, 2);
  const size_t n = orc_result_cells(r);
  const uint64_t f = orc_result_total_faces(r);
  std::vector<double> vol(n), area(f);
  std::vector<uint32_t> fc(n);
  std::vector<uint64_t> fo(n + 1);
  std::vector<int32_t> nb(f);
  orc_result_copy(r, vol.data(), fc.data(), fo.data(), nb.data(), area.data());
  std::vector<orc_cell_stats> st(n);
  orc_result_stats(r, st.data());
  orc_result_free(r);
  double sum = 0.0;
  int bad = 0;
  for (size_t i = 0; i < n; ++i) {
    sum += vol[i];
    bad += st[i].status != 0;
  }
  const double want = box_volume(s);
  const bool ok = !bad && std::fabs(sum - want) <= 1e-9 * want;
  std::printf("%-22s mode %d clip %d: %zu cells, %llu faces, volume %.12g (box %.12g), failed cells %d -> %s\n",
              s.name, mode, clip, n, (unsigned long long)f, sum, want, bad, ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

Set uniform(const char* name, int n, unsigned seed, const double lo[3], const double hi[3], const double plo[3],
            const double phi[3]) {
  Set s;
  s.name = name;
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (int i = 0; i < n; ++i) {
    s.x.push_back(plo[0] + (phi[0] - plo[0]) * u(g));
    s.y.push_back(plo[1] + (phi[1] - plo[1]) * u(g));
    s.z.push_back(plo[2] + (phi[2] - plo[2]) * u(g));
  }
  for (int k = 0; k < 3; ++k) {
    s.lo[k] = lo[k];
    s.hi[k] = hi[k];
  }
  return s;
}

}  // namespace

int main() {
  const double z0[3] = {0, 0, 0}, o1[3] = {1, 1, 1}, m1[3] = {-1, -1, -1}, p2[3] = {2, 2, 2};
  int fails = 0;
  fails += check(uniform("uniform 1500", 1500, 1, z0, o1, z0, o1), ORC_MODE_SECURITY, ORC_CLIP_FACES);
  fails += check(uniform("uniform 300 walk", 300, 2, z0, o1, z0, o1), ORC_MODE_SECURITY, ORC_CLIP_WALK);
  fails += check(uniform("uniform 300 all points", 300, 3, z0, o1, z0, o1), ORC_MODE_ALL_POINTS, ORC_CLIP_FACES);
  fails += check(uniform("box [-1,2]^3", 1200, 4, m1, p2, z0, o1), ORC_MODE_SECURITY, ORC_CLIP_FACES);
  {
    const double a0[3] = {0, 0, 0}, a1[3] = {0.1, 0.1, 0.1}, b0[3] = {0.9, 0.9, 0.9};
    Set a = uniform("void", 700, 5, z0, o1, a0, a1), b = uniform("void", 700, 6, z0, o1, b0, o1);
    a.x.insert(a.x.end(), b.x.begin(), b.x.end());
    a.y.insert(a.y.end(), b.y.begin(), b.y.end());
    a.z.insert(a.z.end(), b.z.begin(), b.z.end());
    fails += check(a, ORC_MODE_SECURITY, ORC_CLIP_FACES);
  }
  {
    Set s;
    s.name = "sphere ring 200";
    std::mt19937_64 g(7);
    std::normal_distribution<double> nd(0.0, 1.0);
    s.x.push_back(0.5);
    s.y.push_back(0.5);
    s.z.push_back(0.5);
    for (int i = 0; i < 200; ++i) {
      double v[3] = {nd(g), nd(g), nd(g)};
      const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      s.x.push_back(0.5 + 0.3 * v[0] / r);
      s.y.push_back(0.5 + 0.3 * v[1] / r);
      s.z.push_back(0.5 + 0.3 * v[2] / r);
    }
    for (int k = 0; k < 3; ++k) {
      s.lo[k] = 0;
      s.hi[k] = 1;
    }
    fails += check(s, ORC_MODE_SECURITY, ORC_CLIP_FACES);
  }
  {
    Set s;
    s.name = "fcc 4^3 jitter";
    std::mt19937_64 g(8);
    std::normal_distribution<double> nd(0.0, 1e-3);
    const double b[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k)
          for (int q = 0; q < 4; ++q) {
            s.x.push_back(i + b[q][0] + 0.25 + nd(g));
            s.y.push_back(j + b[q][1] + 0.25 + nd(g));
            s.z.push_back(k + b[q][2] + 0.25 + nd(g));
          }
    for (int k = 0; k < 3; ++k) {
      s.lo[k] = 0;
      s.hi[k] = 4;
    }
    fails += check(s, ORC_MODE_SECURITY, ORC_CLIP_FACES);
  }
  std::printf("%s\n", fails ? "sanitized oracle run: FAILURES" : "sanitized oracle run: clean");
  return fails ? 1 : 0;
}
