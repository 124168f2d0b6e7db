// datagen.cpp -- seeded synthetic point sets for the tessellation workloads.
//
// PRNG: xoshiro256** seeded through SplitMix64; uniform doubles are
// (u64 >> 11) * 2^-53.  Gaussians use Box-Muller (libm log/cos/sin, so sets
// with Gaussians are regenerated on the host that checks them, never stored).
//
// Workloads:
//   uniform   n points i.i.d. uniform in [lo, hi)
//   clustered 80% from 64 isotropic Gaussians (centres uniform, sigma 0.02,
//             resampled until inside the unit cube), 20% uniform background
//   batch     set i: n = round(exp(U[ln nmin, ln nmax])), box sides U[0.5, 2],
//             points uniform in [0, side)
//   fcc       ncell^3 FCC cells (cubic constant a), 4 sites per cell at
//             (basis + 1/4) * a, each jittered by N(0, sigma^2) per axis
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& w : s) w = splitmix(x);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  // Box-Muller, one value per call (the second is discarded for simplicity).
  double normal() {
    double u1 = 1.0 - uniform();  // (0, 1]
    double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

}  // namespace

extern "C" {

void tg_uniform(uint64_t seed, size_t n, const double lo[3], const double hi[3], double* xyz) {
  Rng r(seed);
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) xyz[3 * i + a] = lo[a] + r.uniform() * (hi[a] - lo[a]);
}

void tg_clustered(uint64_t seed, size_t n, int nclusters, double sigma, double frac, double* xyz) {
  Rng r(seed);
  double c[1024][3];
  if (nclusters > 1024) nclusters = 1024;
  for (int k = 0; k < nclusters; ++k)
    for (int a = 0; a < 3; ++a) c[k][a] = r.uniform();
  for (size_t i = 0; i < n; ++i) {
    if (r.uniform() < frac) {
      int k = (int)(r.uniform() * nclusters);
      if (k >= nclusters) k = nclusters - 1;
      for (int a = 0; a < 3; ++a) {
        double v;
        do {
          v = c[k][a] + sigma * r.normal();
        } while (!(v >= 0.0 && v < 1.0));
        xyz[3 * i + a] = v;
      }
    } else {
      for (int a = 0; a < 3; ++a) xyz[3 * i + a] = r.uniform();
    }
  }
}

// Size and box of batch set `seed` (no points).
void tg_batch_shape(uint64_t seed, double nmin, double nmax, uint64_t* n, double side[3]) {
  Rng r(seed);
  double ln = std::log(nmin) + r.uniform() * (std::log(nmax) - std::log(nmin));
  *n = (uint64_t)std::llround(std::exp(ln));
  for (int a = 0; a < 3; ++a) side[a] = 0.5 + 1.5 * r.uniform();
}

// Points of batch set `seed` (same stream continued after the shape draws).
void tg_batch_points(uint64_t seed, double nmin, double nmax, double* xyz) {
  Rng r(seed);
  double ln = std::log(nmin) + r.uniform() * (std::log(nmax) - std::log(nmin));
  uint64_t n = (uint64_t)std::llround(std::exp(ln));
  double side[3];
  for (int a = 0; a < 3; ++a) side[a] = 0.5 + 1.5 * r.uniform();
  for (uint64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) xyz[3 * i + a] = r.uniform() * side[a];
}

size_t tg_fcc_count(int ncell) { return (size_t)4 * ncell * ncell * ncell; }

void tg_fcc(uint64_t seed, int ncell, double a, double sigma, double* xyz) {
  Rng r(seed);
  static const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  size_t i = 0;
  for (int cx = 0; cx < ncell; ++cx)
    for (int cy = 0; cy < ncell; ++cy)
      for (int cz = 0; cz < ncell; ++cz)
        for (int b = 0; b < 4; ++b) {
          const int cc[3] = {cx, cy, cz};
          for (int k = 0; k < 3; ++k) {
            double v = (cc[k] + basis[b][k] + 0.25) * a;
            if (sigma > 0.0) v += sigma * r.normal();
            xyz[3 * i + k] = v;
          }
          ++i;
        }
}

}  // extern "C"
