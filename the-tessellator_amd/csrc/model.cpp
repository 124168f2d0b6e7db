// model.cpp -- the learned search-radius model: samples, GPU training,
// inference, and the learned schedule of the cell passes.
//
// The network's predictions never change a result.  They choose, per cell,
// the vertex capacity of its first pass (64, 128 or 256 vertices) and its
// position in the work queue (heaviest predicted cells first).  A cell whose
// predicted capacity is too small overflows and is recomputed by the plain
// largest-capacity pass; those cells are flagged TESS_CELL_FALLBACK and
// counted in tess_timings.fallback_cells.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <numeric>
#include <vector>

#include "runtime.h"

using namespace tess;

namespace {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Grow a device buffer to hold `need` elements, keeping the first `keep`.
template <class T>
hipError_t grow(DevBuf<T>& b, size_t need, size_t keep, hipStream_t s) {
  if (b.p && b.n >= need) return hipSuccess;
  size_t cap = std::max(need, b.n * 2);
  DevBuf<T> nb;
  hipError_t e = nb.alloc(cap);
  if (e != hipSuccess) return e;
  if (keep && b.p) {
    e = hipMemcpyAsync(nb.p, b.p, sizeof(T) * keep, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
  }
  std::swap(b.p, nb.p);
  std::swap(b.n, nb.n);
  return hipSuccess;
}

tess_status refresh_weights(tess_model* m) {
  return m->ctx->hip(launch_weights_from_master(m->master.p, m->wbf.p, m->bias.p, m->ctx->stream), "weights");
}

double smallest_edge(const tess_diagram* d) {
  return std::min(d->hgrid.sx, std::min(d->hgrid.sy, d->hgrid.sz));
}

// Features and raw network outputs (standardised) for every point of a
// binned diagram, into d->feat / d->pred.
tess_status infer(tess_diagram* d, const tess_model* m) {
  tess_ctx* c = d->ctx;
  const int n = (int)d->n;
  if (d->feat.alloc((size_t)n * kMlpIn) != hipSuccess || d->pred.alloc((size_t)n * kMlpOut) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  tess_status st;
  if ((st = c->hip(launch_features(d->bins.p, d->delim.p, (int)d->cpd, n, d->feat.p, c->stream), "features")) !=
      TESS_OK)
    return st;
  return c->hip(launch_mlp_forward(d->feat.p, n, m->wbf.p, m->bias.p, d->pred.p, c->stream), "mlp forward");
}

// Work lists from per-point costs y[o * stride] * scale + shift.
// Shell rings a table needs so that its cut key exceeds the stop threshold of
// a cell of radius r: kcut(L) = min_a (L s_a)^2 > 4 r^2 (1 + 1e-9), with a
// 5% margin on r; the default table when that is enough, the full order at
// most.
int rings_for_radius(const tess_diagram* d, double r) {
  const double smin = smallest_edge(d);
  int L = kTableL;
  const int64_t full = d->cpd - 1;
  while (L < full) {
    const double kc = (L * smin) * (L * smin);
    if (kc > 4.0 * (1.05 * r) * (1.05 * r) * (1.0 + 1e-9)) break;
    L = (int)std::min<int64_t>(2 * (int64_t)L, full);
  }
  return L;
}

tess_status schedule(tess_diagram* d, const float* y, int stride, float scale, float shift, float t1, float t2,
                     Schedule& sch, const SchedRadius* radius = nullptr) {
  tess_ctx* c = d->ctx;
  const int n = (int)d->n;
  const size_t ntiles = (d->n + 255) / 256;
  const size_t m = (size_t)kSchedKeys * ntiles;
  if (d->sched_key.alloc((d->n + 3) / 4) != hipSuccess || d->sched_hist.alloc(m) != hipSuccess ||
      d->sched_sums.alloc((m + 1023) / 1024 + 1) != hipSuccess ||
      d->sched_list.alloc(d->n + kSchedSlots + 2) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  uint32_t* bounds_dev = d->sched_list.p + d->n;
  tess_status st;
  if ((st = c->hip(launch_schedule(y, stride, scale, shift, t1, t2, d->sorted.p, n, (uint8_t*)d->sched_key.p,
                                   d->sched_hist.p, d->sched_sums.p, d->sched_list.p, bounds_dev, c->stream,
                                   radius),
                   "schedule")) != TESS_OK)
    return st;
  uint32_t b[kSchedSlots + 2];
  if ((st = c->hip(hipMemcpyAsync(b, bounds_dev, sizeof(b), hipMemcpyDeviceToHost, c->stream), "schedule")) !=
          TESS_OK ||
      (st = c->hip(hipStreamSynchronize(c->stream), "schedule")) != TESS_OK)
    return st;
  bool ok = b[0] == 0 && b[kSchedSlots] == (uint32_t)n;
  for (int slot = 0; slot < kSchedSlots; ++slot) ok = ok && b[slot] <= b[slot + 1];
  if (!ok) return c->fail(TESS_E_DEVICE, "schedule lost cells");
  for (int slot = 0; slot < kSchedSlots; ++slot) {
    sch.list[slot] = d->sched_list.p + b[slot];
    sch.count[slot] = (int)(b[slot + 1] - b[slot]);
  }
  if (sch.count[3]) {
    // the table of the long-search slot: long enough for its largest
    // predicted radius (order-preserving bits back to the float)
    const uint32_t u = b[kSchedSlots + 1];
    uint32_t bitsv = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float w;
    std::memcpy(&w, &bitsv, 4);
    const double r = std::exp((double)w) * smallest_edge(d);
    sch.long_l = rings_for_radius(d, std::isfinite(r) ? r : 1e300);
  }
  return TESS_OK;
}

bool model_ready(const tess_model* m) { return m && m->master.p && m->wbf.p && m->bias.p; }

}  // namespace

extern "C" {

tess_status tess_model_create(tess_ctx* ctx, uint64_t seed, tess_model** out) {
  if (!ctx || !out) return TESS_E_INVALID;
  *out = nullptr;
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  tess_model* m = new (std::nothrow) tess_model;
  if (!m) return ctx->fail(TESS_E_NOMEM, "host allocation failed");
  m->ctx = ctx;
  if (m->master.alloc(kMlpParams) != hipSuccess || m->m1.alloc(kMlpParams) != hipSuccess ||
      m->m2.alloc(kMlpParams) != hipSuccess || m->bias.alloc(kMlpBias) != hipSuccess ||
      m->wbf.alloc(kMlpWeights) != hipSuccess) {
    delete m;
    return ctx->fail(TESS_E_NOMEM, "device allocation failed");
  }
  // He-uniform weights (fan-in 27, 64, 64), zero biases and pad rows/columns
  std::vector<float> p(kMlpParams, 0.f);
  uint64_t s = seed;
  auto fill = [&](int off, int rows, int cols, int used_rows, int used_cols, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int r = 0; r < rows; ++r)
      for (int k = 0; k < cols; ++k) {
        const double u = uniform01(s);
        p[off + r * cols + k] = (r < used_rows && k < used_cols) ? (float)((2.0 * u - 1.0) * a) : 0.f;
      }
  };
  fill(0, kMlpIn, kMlpH, kMlpFeatures, kMlpH, kMlpFeatures);
  fill(kMlpIn * kMlpH, kMlpH, kMlpH, kMlpH, kMlpH, kMlpH);
  fill(kMlpIn * kMlpH + kMlpH * kMlpH, kMlpH, kMlpOutPad, kMlpH, kMlpOut, kMlpH);
  tess_status st = ctx->hip(hipMemcpy(m->master.p, p.data(), 4 * p.size(), hipMemcpyHostToDevice), "upload");
  if (st == TESS_OK) st = ctx->hip(hipMemsetAsync(m->m1.p, 0, 4 * kMlpParams, ctx->stream), "memset");
  if (st == TESS_OK) st = ctx->hip(hipMemsetAsync(m->m2.p, 0, 4 * kMlpParams, ctx->stream), "memset");
  if (st == TESS_OK) st = refresh_weights(m);
  if (st == TESS_OK) st = ctx->hip(hipStreamSynchronize(ctx->stream), "model init");
  if (st != TESS_OK) {
    delete m;
    return st;
  }
  *out = m;
  return TESS_OK;
}

void tess_model_destroy(tess_model* m) {
  if (!m) return;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  delete m;
}

tess_status tess_model_get(const tess_model* m, float* params, float norm[4], float calib[4]) {
  if (!m) return TESS_E_INVALID;
  (void)hipSetDevice(m->ctx->device);
  if (params) {
    tess_status st =
        m->ctx->hip(hipMemcpy(params, m->master.p, 4 * kMlpParams, hipMemcpyDeviceToHost), "download");
    if (st != TESS_OK) return st;
  }
  if (norm) std::memcpy(norm, m->norm, sizeof(m->norm));
  if (calib) std::memcpy(calib, m->calib, sizeof(m->calib));
  return TESS_OK;
}

tess_status tess_model_set(tess_model* m, const float* params, const float norm[4], const float calib[4]) {
  if (!m) return TESS_E_INVALID;
  tess_ctx* c = m->ctx;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  if (norm) {
    if (!(norm[1] > 0.f) || !(norm[3] > 0.f)) return c->fail(TESS_E_INVALID, "target std must be positive");
    std::memcpy(m->norm, norm, sizeof(m->norm));
  }
  if (calib) std::memcpy(m->calib, calib, sizeof(m->calib));
  if (params) {
    tess_status st = c->hip(hipMemcpy(m->master.p, params, 4 * kMlpParams, hipMemcpyHostToDevice), "upload");
    if (st == TESS_OK) st = refresh_weights(m);
    if (st == TESS_OK) st = c->hip(hipStreamSynchronize(c->stream), "model set");
    if (st != TESS_OK) return st;
  }
  return TESS_OK;
}

tess_status tess_model_save(const tess_model* m, const char* path) {
  if (!m || !path) return TESS_E_INVALID;
  std::vector<float> p(kMlpParams);
  tess_status st = tess_model_get(m, p.data(), nullptr, nullptr);
  if (st != TESS_OK) return st;
  FILE* f = std::fopen(path, "wb");
  if (!f) return m->ctx->fail(TESS_E_INVALID, std::string("cannot write ") + path);
  const char magic[8] = {'T', 'E', 'S', 'S', 'M', 'L', 'P', '1'};
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(p.data(), 4, p.size(), f) == p.size() &&
            std::fwrite(m->norm, 4, 4, f) == 4 && std::fwrite(m->calib, 4, 4, f) == 4;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? TESS_OK : m->ctx->fail(TESS_E_INVALID, std::string("write failed: ") + path);
}

tess_status tess_model_load(tess_model* m, const char* path) {
  if (!m || !path) return TESS_E_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return m->ctx->fail(TESS_E_NOTFOUND, std::string("no model file ") + path);
  char magic[8];
  std::vector<float> p(kMlpParams);
  float norm[4], calib[4];
  const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "TESSMLP1", 8) == 0 &&
                  std::fread(p.data(), 4, p.size(), f) == p.size() && std::fread(norm, 4, 4, f) == 4 &&
                  std::fread(calib, 4, 4, f) == 4;
  std::fclose(f);
  if (!ok) return m->ctx->fail(TESS_E_NOTFOUND, std::string("malformed model file ") + path);
  return tess_model_set(m, p.data(), norm, calib);
}

tess_status tess_model_add_samples(tess_model* m, tess_diagram* d) {
  if (!m || !d) return TESS_E_INVALID;
  tess_ctx* c = m->ctx;
  if (d->ctx != c) return c->fail(TESS_E_INVALID, "model and diagram belong to different contexts");
  if (!d->computed || !d->opts.collect_stats)
    return c->fail(TESS_E_STATE, "training samples need a diagram computed with collect_stats = 1");
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const size_t n = d->n, at = m->count;
  if (grow(m->X, (at + n) * kMlpIn, at * kMlpIn, s) != hipSuccess ||
      grow(m->T, (at + n) * kMlpOut, at * kMlpOut, s) != hipSuccess || grow(m->V, at + n, at, s) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  tess_status st;
  if ((st = c->hip(launch_features(d->bins.p, d->delim.p, (int)d->cpd, (int)n, m->X.p + at * kMlpIn, s),
                   "features")) != TESS_OK ||
      (st = c->hip(launch_targets(d->stats.p, (int)n, smallest_edge(d), m->T.p + at * kMlpOut, m->V.p + at, s),
                   "targets")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "samples")) != TESS_OK)
    return st;
  m->count = at + n;
  return TESS_OK;
}

size_t tess_model_sample_count(const tess_model* m) { return m ? m->count : 0; }

tess_status tess_model_train(tess_model* m, int epochs, int batch, float lr, uint64_t seed, float* loss_out) {
  if (!m || epochs < 0 || batch < 32 || !(lr > 0.f)) return m ? m->ctx->fail(TESS_E_INVALID, "bad training options")
                                                              : TESS_E_INVALID;
  tess_ctx* c = m->ctx;
  if (m->count < (size_t)batch) return c->fail(TESS_E_STATE, "fewer samples than one batch");
  if (m->count > (size_t)INT32_MAX) return c->fail(TESS_E_INVALID, "too many samples");
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const size_t S = m->count;
  tess_status st;
  // target statistics (host, double)
  std::vector<float> T(2 * S);
  if ((st = c->hip(hipMemcpy(T.data(), m->T.p, 4 * T.size(), hipMemcpyDeviceToHost), "download")) != TESS_OK)
    return st;
  double mean[2] = {0, 0}, var[2] = {0, 0};
  for (size_t i = 0; i < S; ++i)
    for (int k = 0; k < 2; ++k) mean[k] += T[2 * i + k];
  for (int k = 0; k < 2; ++k) mean[k] /= (double)S;
  for (size_t i = 0; i < S; ++i)
    for (int k = 0; k < 2; ++k) var[k] += (T[2 * i + k] - mean[k]) * (T[2 * i + k] - mean[k]);
  for (int k = 0; k < 2; ++k) {
    const double sd = std::sqrt(var[k] / (double)S);
    m->norm[2 * k] = (float)mean[k];
    m->norm[2 * k + 1] = (float)(sd > 1e-6 ? sd : 1.0);
  }
  DevBuf<float> Tn, grad, loss;
  DevBuf<int> idx;
  const int nbatches = (int)(S / (size_t)batch);  // the remainder of a shuffle is left out
  const int slots = (batch + 31) / 32;
  if (Tn.alloc(2 * S) != hipSuccess || grad.alloc((size_t)slots * kMlpParams) != hipSuccess ||
      loss.alloc((size_t)nbatches * slots) != hipSuccess || idx.alloc(S) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  if ((st = c->hip(launch_standardise(m->T.p, S, m->norm, Tn.p, s), "standardise")) != TESS_OK) return st;
  std::vector<int> perm(S);
  std::vector<float> lh((size_t)nbatches * slots);
  uint64_t rs = seed;
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(perm.begin(), perm.end(), 0);
    for (size_t i = S - 1; i > 0; --i) std::swap(perm[i], perm[splitmix64(rs) % (i + 1)]);
    if ((st = c->hip(hipMemcpyAsync(idx.p, perm.data(), 4 * S, hipMemcpyHostToDevice, s), "upload")) != TESS_OK)
      return st;
    for (int b = 0; b < nbatches; ++b) {
      ++m->adam_steps;
      const float b1c = 1.f - std::pow(0.9f, (float)m->adam_steps);
      const float b2c = 1.f - std::pow(0.999f, (float)m->adam_steps);
      if ((st = c->hip(launch_mlp_train_step(m->X.p, Tn.p, idx.p + (size_t)b * batch, batch, 1.f / batch, m->wbf.p,
                                             m->bias.p, grad.p, loss.p + (size_t)b * slots, s),
                       "train step")) != TESS_OK ||
          (st = c->hip(launch_adam(grad.p, slots, m->master.p, m->m1.p, m->m2.p, m->wbf.p, m->bias.p, lr, b1c, b2c,
                                   s),
                       "adam")) != TESS_OK)
        return st;
    }
    if ((st = c->hip(hipMemcpyAsync(lh.data(), loss.p, 4 * lh.size(), hipMemcpyDeviceToHost, s), "loss")) !=
            TESS_OK ||
        (st = c->hip(hipStreamSynchronize(s), "train")) != TESS_OK)
      return st;
    double tot = 0;
    for (float v : lh) tot += v;
    if (loss_out) loss_out[ep] = (float)(tot / ((double)nbatches * batch));
  }
  // routing thresholds: predicted log(1 + C) of the samples against the
  // largest vertex count they reached.  Predictions are binned (width 1/4);
  // threshold r is the upper edge of the highest bin (with at least 20
  // samples) in which fewer than one cell in five outgrew caps[r] vertices,
  // so the cells routed to the larger capacity are those whose bins mostly
  // need it.  No bin at or above one in five: the route stays closed.
  DevBuf<float> Y;
  if (Y.alloc(2 * S) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
  std::vector<float> y(2 * S);
  std::vector<uint16_t> V(S);
  if ((st = c->hip(launch_mlp_forward(m->X.p, (int)S, m->wbf.p, m->bias.p, Y.p, s), "mlp forward")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(y.data(), Y.p, 4 * y.size(), hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(V.data(), m->V.p, 2 * S, hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "calibrate")) != TESS_OK)
    return st;
  constexpr int kBins = 80;  // log(1 + C) in [0, 20)
  const int caps[2] = {64, 128};
  for (int r = 0; r < 2; ++r) {
    std::vector<uint64_t> cnt(kBins, 0), over(kBins, 0);
    bool any = false;
    for (size_t i = 0; i < S; ++i) {
      const float v = y[2 * i + 1] * m->norm[3] + m->norm[2];
      const int bin = std::min(kBins - 1, std::max(0, (int)std::floor(v * 4.0f)));
      cnt[bin]++;
      if (V[i] > caps[r]) {
        over[bin]++;
        any = true;
      }
    }
    float t = std::numeric_limits<float>::infinity();
    if (any) {
      int hi = -1;  // highest populated bin below one in five
      bool high_seen = false;
      for (int bin = 0; bin < kBins; ++bin) {
        if (cnt[bin] < 20) continue;
        if (5 * over[bin] < cnt[bin])
          hi = bin;
        else
          high_seen = true;
      }
      if (high_seen) t = (hi + 1) / 4.0f;
    }
    m->calib[r] = t;
  }
  if (m->calib[1] < m->calib[0]) m->calib[1] = m->calib[0];
  return TESS_OK;
}

tess_status tess_diagram_predict(tess_diagram* d, const tess_model* m, float* out) {
  if (!d || !out || !model_ready(m)) return TESS_E_INVALID;
  tess_ctx* c = d->ctx;
  if (m->ctx != c) return c->fail(TESS_E_INVALID, "model and diagram belong to different contexts");
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  tess_status st;
  if (!d->binned && (st = bin(d)) != TESS_OK) return st;
  if ((st = infer(d, m)) != TESS_OK) return st;
  std::vector<float> y(2 * d->n);
  if ((st = c->hip(hipMemcpyAsync(y.data(), d->pred.p, 4 * y.size(), hipMemcpyDeviceToHost, c->stream),
                   "download")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(c->stream), "predict")) != TESS_OK)
    return st;
  const double smin = smallest_edge(d);
  for (size_t i = 0; i < d->n; ++i) {
    out[2 * i] = (float)(std::exp((double)y[2 * i] * m->norm[1] + m->norm[0]) * smin);
    out[2 * i + 1] = (float)std::expm1((double)y[2 * i + 1] * m->norm[3] + m->norm[2]);
  }
  return TESS_OK;
}

tess_status tess_diagram_compute_learned(tess_diagram* d, const tess_model* m) {
  if (!d || !model_ready(m)) return TESS_E_INVALID;
  tess_ctx* c = d->ctx;
  if (m->ctx != c) return c->fail(TESS_E_INVALID, "model and diagram belong to different contexts");
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  Planner plan = [m](tess_diagram* dd, Schedule& sch) -> tess_status {
    tess_status st = infer(dd, m);
    if (st != TESS_OK) return st;
    // the predicted final radius R routes a cell to the long-search list when
    // its security ball would reach past the default table: 4 R^2 (1 + 1e-9)
    // >= kcut = (12 s_min)^2, i.e. log(R / s_min) >= log 6, less a margin of
    // 20% for the prediction's error; a cell it under-sizes still finds its
    // table by the exact rerun
    SchedRadius r;
    r.y = dd->pred.p;
    r.scale = m->norm[1];
    r.shift = m->norm[0];
    r.threshold = (float)(std::log(kTableL / 2.0) + std::log(0.8));
    return schedule(dd, dd->pred.p + 1, kMlpOut, m->norm[3], m->norm[2], m->calib[0], m->calib[1], sch, &r);
  };
  return compute_cells(d, &plan);
}

tess_status tess_diagram_compute_scheduled(tess_diagram* d, const float* log1p_c, const float calib[2]) {
  if (!d || !log1p_c || !calib) return TESS_E_INVALID;
  tess_ctx* c = d->ctx;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  if (d->pred.alloc(d->n * kMlpOut) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
  tess_status st = c->hip(hipMemcpy(d->pred.p, log1p_c, 4 * d->n, hipMemcpyHostToDevice), "upload");
  if (st != TESS_OK) return st;
  const float t1 = calib[0], t2 = calib[1];
  Planner plan = [t1, t2](tess_diagram* dd, Schedule& sch) -> tess_status {
    return schedule(dd, dd->pred.p, 1, 1.f, 0.f, t1, t2, sch);
  };
  return compute_cells(d, &plan);
}

tess_status tess_diagram_features(tess_diagram* d, uint16_t* x) {
  if (!d || !x) return TESS_E_INVALID;
  tess_ctx* c = d->ctx;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  tess_status st;
  if (!d->binned && (st = bin(d)) != TESS_OK) return st;
  if (d->feat.alloc(d->n * kMlpIn) != hipSuccess) return c->fail(TESS_E_NOMEM, "device allocation failed");
  if ((st = c->hip(launch_features(d->bins.p, d->delim.p, (int)d->cpd, (int)d->n, d->feat.p, c->stream),
                   "features")) != TESS_OK)
    return st;
  return c->hip(hipMemcpy(x, d->feat.p, 2 * d->n * kMlpIn, hipMemcpyDeviceToHost), "download");
}

tess_status tess_mlp_forward(tess_model* m, const uint16_t* x, size_t n, float* y) {
  if (!model_ready(m) || !x || !y || n > (size_t)INT32_MAX) return TESS_E_INVALID;
  tess_ctx* c = m->ctx;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  DevBuf<uint16_t> X;
  DevBuf<float> Y;
  if (X.alloc(n * kMlpIn) != hipSuccess || Y.alloc(n * kMlpOut) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  tess_status st;
  if ((st = c->hip(hipMemcpy(X.p, x, 2 * n * kMlpIn, hipMemcpyHostToDevice), "upload")) != TESS_OK ||
      (st = c->hip(launch_mlp_forward(X.p, (int)n, m->wbf.p, m->bias.p, Y.p, c->stream), "mlp forward")) != TESS_OK)
    return st;
  return c->hip(hipMemcpy(y, Y.p, 4 * n * kMlpOut, hipMemcpyDeviceToHost), "download");
}

tess_status tess_mlp_gradient(tess_model* m, const uint16_t* x, const float* t, size_t n, float* grad, float* loss) {
  if (!model_ready(m) || !x || !t || !grad || n == 0 || n > (size_t)INT32_MAX) return TESS_E_INVALID;
  tess_ctx* c = m->ctx;
  std::lock_guard<std::mutex> g(c->mu);
  (void)hipSetDevice(c->device);
  const int slots = (int)((n + 31) / 32);
  DevBuf<uint16_t> X;
  DevBuf<float> T, G, Gs, L;
  DevBuf<int> idx;
  if (X.alloc(n * kMlpIn) != hipSuccess || T.alloc(n * kMlpOut) != hipSuccess ||
      G.alloc((size_t)slots * kMlpParams) != hipSuccess || Gs.alloc(kMlpParams) != hipSuccess ||
      L.alloc(slots) != hipSuccess || idx.alloc(n) != hipSuccess)
    return c->fail(TESS_E_NOMEM, "device allocation failed");
  std::vector<int> id(n);
  std::iota(id.begin(), id.end(), 0);
  hipStream_t s = c->stream;
  tess_status st;
  if ((st = c->hip(hipMemcpy(X.p, x, 2 * n * kMlpIn, hipMemcpyHostToDevice), "upload")) != TESS_OK ||
      (st = c->hip(hipMemcpy(T.p, t, 4 * n * kMlpOut, hipMemcpyHostToDevice), "upload")) != TESS_OK ||
      (st = c->hip(hipMemcpy(idx.p, id.data(), 4 * n, hipMemcpyHostToDevice), "upload")) != TESS_OK ||
      (st = c->hip(launch_mlp_train_step(X.p, T.p, idx.p, (int)n, 1.f / (float)n, m->wbf.p, m->bias.p, G.p, L.p, s),
                   "train step")) != TESS_OK ||
      (st = c->hip(launch_grad_sum(G.p, slots, Gs.p, s), "grad sum")) != TESS_OK)
    return st;
  std::vector<float> lh(slots);
  if ((st = c->hip(hipMemcpyAsync(grad, Gs.p, 4 * kMlpParams, hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
      (st = c->hip(hipMemcpyAsync(lh.data(), L.p, 4 * slots, hipMemcpyDeviceToHost, s), "download")) != TESS_OK ||
      (st = c->hip(hipStreamSynchronize(s), "gradient")) != TESS_OK)
    return st;
  if (loss) {
    double tot = 0;
    for (float v : lh) tot += v;
    *loss = (float)(tot / (double)n);
  }
  return TESS_OK;
}

tess_status tess_mfma_probe(tess_ctx* ctx, const uint16_t* a, const uint16_t* b, float* d) {
  if (!ctx || !a || !b || !d) return TESS_E_INVALID;
  std::lock_guard<std::mutex> g(ctx->mu);
  (void)hipSetDevice(ctx->device);
  DevBuf<uint16_t> A, B;
  DevBuf<float> D;
  if (A.alloc(16 * 32) != hipSuccess || B.alloc(32 * 16) != hipSuccess || D.alloc(256) != hipSuccess)
    return ctx->fail(TESS_E_NOMEM, "device allocation failed");
  tess_status st;
  if ((st = ctx->hip(hipMemcpy(A.p, a, 2 * 512, hipMemcpyHostToDevice), "upload")) != TESS_OK ||
      (st = ctx->hip(hipMemcpy(B.p, b, 2 * 512, hipMemcpyHostToDevice), "upload")) != TESS_OK ||
      (st = ctx->hip(launch_mfma_probe(A.p, B.p, D.p, ctx->stream), "mfma probe")) != TESS_OK)
    return st;
  return ctx->hip(hipMemcpy(d, D.p, 4 * 256, hipMemcpyDeviceToHost), "download");
}

}  // extern "C"
