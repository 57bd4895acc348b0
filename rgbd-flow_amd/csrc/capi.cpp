// capi.cpp -- the C ABI of librgbdflow: context, HBM buffer pool and the
// device-resident orchestration of the coarse-to-fine solve.
//
// The host enqueues a whole frame pair (every pyramid level, every outer and
// inner iteration) on the context's stream without waiting: early exits are
// decided on the device (SolveCtl), and the host reads the control blocks
// back once, after the pair, to report iteration counts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/rgbdflow/rgbdflow.h"
#include "internal.hpp"

using namespace rgbdflow;

namespace {

struct Buffer {
  void* p = nullptr;
  size_t bytes = 0;
};

struct LevelShape {
  int W, H;
};

}  // namespace

struct rgbdflow_ctx {
  int device = 0;
  int cu_count = 256;
  size_t lds_per_block = 65536;
  hipStream_t stream = nullptr;
  unsigned long long budget = 0;
  unsigned long long in_use = 0;
  std::string err;
  std::map<std::string, Buffer> bufs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool time_relax = false;
  std::vector<hipEvent_t> relax_ev;
  int relax_ev_used = 0;
  // loaded pair
  bool pair_loaded = false;
  bool pair_ran = false;
  int pW = 0, pH = 0;
  double p_ratio = 0;
  int p_min_width = 0;
  bool p_sigma_given = false;
  bool p_hpb_given = false;
  std::vector<LevelShape> p_levels;
  std::map<std::string, TileSched> plans;  // tile schedules by level shape and request
  int p_final_slot = 0;  // which of the two flow buffers holds the finest result
};

namespace {

const char* kNoCtx = "null context";

int fail(rgbdflow_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                         \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(ctx, RGBDFLOW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Grow-only named device buffer, charged against the context's HBM budget.
int ensure(rgbdflow_ctx* c, const std::string& name, size_t bytes, void** out) {
  Buffer& b = c->bufs[name];
  if (b.bytes >= bytes && b.p) {
    *out = b.p;
    return RGBDFLOW_OK;
  }
  if (b.p) {
    hipFree(b.p);
    c->in_use -= b.bytes;
    b.p = nullptr;
    b.bytes = 0;
  }
  bytes = (bytes + 255) & ~size_t(255);
  if (c->budget && c->in_use + bytes > c->budget)
    return fail(c, RGBDFLOW_ERR_BUDGET,
                "HBM budget exceeded: need " + std::to_string(c->in_use + bytes) + " bytes, budget " +
                    std::to_string(c->budget));
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
  if (e != hipSuccess) return fail(c, RGBDFLOW_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  b.p = p;
  b.bytes = bytes;
  c->in_use += bytes;
  *out = p;
  return RGBDFLOW_OK;
}

template <class T>
int get(rgbdflow_ctx* c, const std::string& name, size_t count, T** out) {
  void* p = nullptr;
  int rc = ensure(c, name, count * sizeof(T), &p);
  *out = static_cast<T*>(p);
  return rc;
}

#define TRY(expr)                 \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != RGBDFLOW_OK) return rc_; \
  } while (0)

// GaussianPyramid.cpp:20-63 level sizes.
std::vector<LevelShape> pyramid_shape(int W, int H, double ratio, int min_width, int* n_smooth, double* r_eff) {
  if (ratio > 0.98 || ratio < 0.4) ratio = 0.75;
  double w = W * ratio;
  int n = 1;
  while (w >= min_width) {
    n++;
    w *= ratio;
  }
  const int ns = std::log(0.25) / std::log(ratio);
  std::vector<LevelShape> L(n);
  L[0] = {W, H};
  for (int i = 1; i < n; i++) {
    if (i <= ns) {
      const double r = std::pow(ratio, i);
      L[i] = {(int)((double)W * r), (int)((double)H * r)};
    } else {
      const double rate = (double)std::pow(ratio, i) * W / L[i - ns].W;
      L[i] = {(int)((double)L[i - ns].W * rate), (int)((double)L[i - ns].H * rate)};
    }
  }
  if (n_smooth) *n_smooth = ns;
  if (r_eff) *r_eff = ratio;
  return L;
}

// Image.h:1179-1199: normalised Gaussian taps exp(-i^2 / (2 sigma^2)).
std::vector<double> gaussian_taps(double sigma, int fsize) {
  std::vector<double> g(2 * fsize + 1);
  double sum = 0;
  const double s2 = sigma * sigma * 2;
  for (int i = -fsize; i <= fsize; i++) {
    g[i + fsize] = std::exp(-(double)(i * i) / s2);
    sum += g[i + fsize];
  }
  for (int i = 0; i < 2 * fsize + 1; i++) g[i] /= sum;
  return g;
}

// Builds levels 1..L-1 of a Gaussian pyramid on the device.  lvl[k] points at
// level k's planes ([C][N_k]); lvl[0] is the source image.
void build_pyramid(hipStream_t s, const std::vector<LevelShape>& L, int ns, double ratio, int C,
                   std::vector<double*>& lvl) {
  const double base = (1 / ratio - 1);
  const double nSigma = base * ns;
  for (size_t i = 1; i < L.size(); i++) {
    if ((int)i <= ns) {
      const double sigma = base * i;
      const int fs = (int)(sigma * 3);
      const std::vector<double> g = gaussian_taps(sigma, fs);
      launch_pyramid_level(s, lvl[0], L[0].W, L[0].H, C, g.data(), fs, std::pow(ratio, (int)i), lvl[i], L[i].W,
                           L[i].H);
    } else {
      const int fs = (int)(nSigma * 3);
      const std::vector<double> g = gaussian_taps(nSigma, fs);
      const LevelShape& src = L[i - ns];
      const double rate = (double)std::pow(ratio, (int)i) * L[0].W / src.W;
      launch_pyramid_level(s, lvl[i - ns], src.W, src.H, C, g.data(), fs, rate, lvl[i], L[i].W, L[i].H);
    }
  }
}

size_t pyramid_elems(const std::vector<LevelShape>& L, int from) {
  size_t n = 0;
  for (size_t i = from; i < L.size(); i++) n += (size_t)L[i].W * L[i].H;
  return n;
}

// sceneFlowMultiscale.cpp:16-25 getDownsizedCamParams (focal length only is used).
float downsized_focal(const rgbdflow_camera& c, int width) {
  const double f = (double)width / c.xRes;
  float fl = c.focalLength;
  fl *= f;
  return fl;
}

// Level work buffers, sized for N pixels.
struct Work {
  double *I1, *I2, *Iw, *Dw, *u[2], *Ix, *Iy, *It, *Dx, *Dy, *Dt, *phi, *partials;
  float *du, *du_alt, *dW, *sW, *mW, *rW, *cW, *q, *coef;
  SolveCtl* ctl;  // kMaxLevels entries
};

int get_work(rgbdflow_ctx* c, size_t N, Work& w) {
  TRY(get(c, "I1", kFeat * N, &w.I1));
  TRY(get(c, "I2", kFeat * N, &w.I2));
  TRY(get(c, "Iw", kFeat * N, &w.Iw));
  TRY(get(c, "Dw", N, &w.Dw));
  TRY(get(c, "u0", 3 * N, &w.u[0]));
  TRY(get(c, "u1", 3 * N, &w.u[1]));
  TRY(get(c, "Ix", kFeat * N, &w.Ix));
  TRY(get(c, "Iy", kFeat * N, &w.Iy));
  TRY(get(c, "It", kFeat * N, &w.It));
  TRY(get(c, "Dx", N, &w.Dx));
  TRY(get(c, "Dy", N, &w.Dy));
  TRY(get(c, "Dt", N, &w.Dt));
  TRY(get(c, "phi", N, &w.phi));
  TRY(get(c, "partials", (size_t)stat_blocks((int)N) + 16, &w.partials));
  TRY(get(c, "du", 3 * N, &w.du));
  TRY(get(c, "du_alt", 3 * N, &w.du_alt));
  TRY(get(c, "dW", N, &w.dW));
  TRY(get(c, "sW", N, &w.sW));
  TRY(get(c, "mW", N, &w.mW));
  TRY(get(c, "rW", 2 * N, &w.rW));
  TRY(get(c, "cW", N, &w.cW));
  TRY(get(c, "q", N, &w.q));
  TRY(get(c, "coef", (size_t)kCoef * N, &w.coef));
  TRY(get(c, "ctl", kMaxLevels, &w.ctl));
  return RGBDFLOW_OK;
}

LevelView make_view(const Work& w, int W, int H, const double* I1, const double* I2, const double* D1,
                    const double* D2, double* u, int level) {
  LevelView v;
  v.W = W;
  v.H = H;
  v.N = W * H;
  v.I1 = I1;
  v.I2 = I2;
  v.Iw = w.Iw;
  v.D1 = D1;
  v.D2 = D2;
  v.Dw = w.Dw;
  v.u = u;
  v.du = w.du;
  v.du_alt = w.du_alt;
  v.dW = w.dW;
  v.sW = w.sW;
  v.mW = w.mW;
  v.rW = w.rW;
  v.cW = w.cW;
  v.q = w.q;
  v.Ix = w.Ix;
  v.Iy = w.Iy;
  v.It = w.It;
  v.Dx = w.Dx;
  v.Dy = w.Dy;
  v.Dt = w.Dt;
  v.phi = w.phi;
  v.coef = w.coef;
  v.partials = w.partials;
  v.ctl = w.ctl + level;
  return v;
}

int validate_iters(rgbdflow_ctx* c, int nOuter, int nInner, int nSOR) {
  if (nOuter < 0 || nInner < 0 || nSOR < 0) return fail(c, RGBDFLOW_ERR_ARG, "negative iteration count");
  if (nOuter > kMaxOuter) return fail(c, RGBDFLOW_ERR_ARG, "nOuterFPIterations above " + std::to_string(kMaxOuter));
  if ((long long)nInner * (nSOR / kCheckEvery + 1) > kMaxChecks)
    return fail(c, RGBDFLOW_ERR_ARG, "nInner * nSOR / 25 above " + std::to_string(kMaxChecks));
  return RGBDFLOW_OK;
}

void relax_event(rgbdflow_ctx* c, bool begin) {
  if (!c->time_relax) return;
  if (c->relax_ev_used >= (int)c->relax_ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    c->relax_ev.push_back(e);
  }
  hipEventRecord(c->relax_ev[c->relax_ev_used++], c->stream);
  (void)begin;
}

struct RelaxPlan {
  int variant = RGBDFLOW_RELAX_PLAIN;
  TileSched ts{};
};

// Tile schedule of the tiled relaxation for a W x H level.  Explicit tile_*
// parameters win; otherwise a cost model picks the workgroup shape, the tile
// and the iterations per launch: a tile covering the whole level runs a whole
// check interval per launch; otherwise cost ~ launches x (launch overhead +
// waves x per-workgroup work), where the per-workgroup work grows with the
// region (tile + halo of 2 T) per thread.
TileSched plan_tiles(const rgbdflow_ctx* c, int W, int H, int req_T, int req_w, int req_h, int req_cfg) {
  TileSched best{};
  double best_cost = 1e300;
  const int n_cfg = tiled_config_count();
  for (int cfg = 0; cfg < n_cfg; cfg++) {
    if (req_cfg >= 0 && cfg != req_cfg) continue;
    if (tiled_config_lds(cfg) > c->lds_per_block) continue;
    int nt, p;
    tiled_config_shape(cfg, &nt, &p);
    const int cap = nt * p;
    const int lds_occ = (int)std::max<size_t>(1, (160 * 1024) / tiled_config_lds(cfg));
    const int occ = std::max(1, std::min(lds_occ, 2048 / nt));
    const long long pairs_img = (long long)((W + 1) / 2) * H;
    if (pairs_img <= cap && req_w == 0) {
      const double per_pass = std::ceil((double)pairs_img / nt) * 0.06 + 0.25;
      const double cost = 2 * (4.0 + std::ceil((double)pairs_img / nt) * 0.6 + 50 * per_pass);
      if (cost < best_cost) {
        best_cost = cost;
        best = TileSched{W + (W & 1), H, 1, 1, 1, cap, cfg, kCheckEvery};
      }
      continue;
    }
    const int Ts[] = {2, 3, 4, 5, 6, 8};
    for (int T : Ts) {
      if (req_T > 0 && T != req_T) continue;
      const int h = 2 * T;
      for (int TW = 8; TW <= W + 7; TW += 8) {
        if (req_w > 0 && TW != req_w + (req_w & 1)) continue;
        for (int TH = 4; TH <= H + 3; TH += 4) {
          if (req_h > 0 && TH != req_h) continue;
          const int rw = std::min(TW + 2 * h, W + (W & 1)), rh = std::min(TH + 2 * h, H);
          const long long rp = (long long)((rw + 1) / 2) * rh;
          if (rp > cap) continue;
          const int tiles = ((W + TW - 1) / TW) * ((H + TH - 1) / TH);
          const double waves = std::ceil((double)tiles / (c->cu_count * occ));
          const double per_thread = std::ceil((double)rp / nt);
          const double per_wg = per_thread * 0.6 + 2 * T * (per_thread * 0.06 + 0.25);
          const double launches = std::ceil((double)kCheckEvery / T) * 2;
          const double cost = launches * (4.0 + waves * per_wg * occ);
          if (cost < best_cost) {
            best_cost = cost;
            best = TileSched{TW, TH, (W + TW - 1) / TW, (H + TH - 1) / TH, 0, cap, cfg, T};
          }
        }
      }
    }
  }
  return best;
}

std::string plan_key(int W, int H, int T, int tw, int th, int cfg) {
  char key[96];
  std::snprintf(key, sizeof key, "%d,%d,%d,%d,%d,%d", W, H, T, tw, th, cfg);
  return key;
}

RelaxPlan make_relax_plan(rgbdflow_ctx* c, int W, int H, const rgbdflow_params* p) {
  RelaxPlan r;
  if (!p || p->relax_variant != RGBDFLOW_RELAX_TILED) return r;
  const std::string key = plan_key(W, H, p->tile_iters, p->tile_w, p->tile_h, p->tile_config);
  auto it = c->plans.find(key);
  if (it == c->plans.end())
    it = c->plans.emplace(key, plan_tiles(c, W, H, p->tile_iters, p->tile_w, p->tile_h, p->tile_config)).first;
  r.ts = it->second;
  if (r.ts.cap == 0) return r;  // nothing fits: plain schedule
  r.variant = RGBDFLOW_RELAX_TILED;
  return r;
}

// nSOR red-black iterations of one (outer, inner), with the convergence
// checks of .cu:905-922 at iterations 24, 49, ...; the tiled schedule splits
// the iterations into launches that end at every check iteration.
void enqueue_relax(rgbdflow_ctx* c, const LevelView& v, const SolveScalars& k, const RelaxPlan& plan, int slot_base) {
  hipStream_t s = c->stream;
  if (plan.variant == RGBDFLOW_RELAX_TILED) {
    int it = 0;
    float* cur = v.du;
    float* nxt = v.du_alt;
    while (it < k.n_sor) {
      const int boundary = (it / kCheckEvery + 1) * kCheckEvery;
      const int T = std::min(std::min(plan.ts.tmax, boundary - it), k.n_sor - it);
      const int last = it + T - 1;
      const int n_prev = k.inner_exit ? it / kCheckEvery : 0;
      const int slot = (last % kCheckEvery == kCheckEvery - 1) ? slot_base + last / kCheckEvery : -1;
      relax_event(c, true);
      launch_relax_tiled(s, v, plan.ts, cur, nxt, T, slot_base, n_prev, slot);
      relax_event(c, false);
      std::swap(cur, nxt);
      it += T;
    }
    if (cur != v.du) launch_copy_f32(s, cur, (size_t)3 * v.N, v.du);
    return;
  }
  for (int it = 0; it < k.n_sor; it++) {
    const int n_prev = k.inner_exit ? it / kCheckEvery : 0;
    const int slot = (it % kCheckEvery == kCheckEvery - 1) ? slot_base + it / kCheckEvery : -1;
    relax_event(c, true);
    launch_relax_plain(s, v, 0, 0, slot_base, n_prev, slot);
    relax_event(c, false);
    relax_event(c, true);
    launch_relax_plain(s, v, 1, 0, slot_base, n_prev, slot);
    relax_event(c, false);
  }
}

// sceneFlowSingleScale's loops (.cu:755-954) enqueued on the device.
void enqueue_level_solve(rgbdflow_ctx* c, const LevelView& v, const SolveScalars& k, const RelaxPlan& plan) {
  hipStream_t s = c->stream;
  const int checks = k.n_sor / kCheckEvery;
  launch_ctl_reset(s, v.ctl);
  for (int o = 0; o < k.n_outer; o++) {
    launch_derivs(s, v, k.n_inner * checks);
    for (int h = 0; h < k.n_inner; h++) {
      launch_phi(s, v, o);
      launch_system(s, v, k, o);
      enqueue_relax(c, v, k, plan, h * checks);
    }
    launch_update_warp(s, v, k, o);
  }
}

void ctl_to_stats(const SolveCtl& ctl, const SolveScalars& k, int& outer_exec, int& sweeps, double& last_stat) {
  outer_exec = ctl.outer_done ? ctl.outer_exit_at + 1 : k.n_outer;
  sweeps = ctl.sweeps_exec;
  last_stat = (k.allow_early && outer_exec > 0) ? ctl.stat[outer_exec - 1] : -1.0;
}

SolveScalars scalars_from(const rgbdflow_params* p, double cw, double alpha, double gamma, int nOuter, int nInner,
                          int nSOR, int allow_early) {
  SolveScalars k;
  k.alpha = alpha;
  k.gamma = gamma;
  k.color_weight = cw;
  k.n_outer = nOuter;
  k.n_inner = nInner;
  k.n_sor = nSOR;
  k.allow_early = allow_early;
  k.inner_exit = p ? p->inner_exit : 1;
  k.outer_exit = p ? p->outer_exit : 1;
  return k;
}

int upload(rgbdflow_ctx* c, void* dst, const void* src, size_t bytes) {
  HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return RGBDFLOW_OK;
}
int download(rgbdflow_ctx* c, void* dst, const void* src, size_t bytes) {
  HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return RGBDFLOW_OK;
}

// Interleaved host [N][C] f64 -> device planes [C][N].
int upload_planes(rgbdflow_ctx* c, const double* host, size_t N, int C, double* dst) {
  double* stage;
  TRY(get(c, "stage", N * C, &stage));
  TRY(upload(c, stage, host, N * C * sizeof(double)));
  launch_deinterleave(c->stream, stage, (int)N, C, dst);
  return RGBDFLOW_OK;
}
// Device planes [C][N] -> interleaved host [N][C] f64.
int download_planes(rgbdflow_ctx* c, const double* src, size_t N, int C, double* host) {
  double* stage;
  TRY(get(c, "stage_out", N * C, &stage));
  launch_interleave(c->stream, src, (int)N, C, stage);
  TRY(download(c, host, stage, N * C * sizeof(double)));
  return RGBDFLOW_OK;
}

int sync(rgbdflow_ctx* c) {
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipGetLastError());
  return RGBDFLOW_OK;
}

void collect_relax_timing(rgbdflow_ctx* c, rgbdflow_stats* st) {
  if (!c->time_relax || !st) return;
  double tot = 0;
  int n = 0;
  for (int i = 0; i + 1 < c->relax_ev_used; i += 2) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->relax_ev[i], c->relax_ev[i + 1]) == hipSuccess) {
      tot += ms;
      n++;
    }
  }
  st->relax_ms = tot;
  st->relax_launches = n;
}

// Level-solve inputs supplied by the host (sceneFlowSingleScale's arguments).
int stage_level_inputs(rgbdflow_ctx* c, int W, int H, float focal_level, const double* Im1, const double* dm1,
                       const double* Im2, const double* dm2, const double* warpIm2, const double* warpDm2,
                       const double* u, const double* v, const double* w, const double* dW64, const double* sW64,
                       const double* mW64, const double* r64, Work& wk, double** D1, double** D2) {
  const size_t N = (size_t)W * H;
  TRY(get_work(c, N, wk));
  TRY(get(c, "lvD1", N, D1));
  TRY(get(c, "lvD2", N, D2));
  TRY(upload_planes(c, Im1, N, kFeat, wk.I1));
  TRY(upload_planes(c, Im2, N, kFeat, wk.I2));
  TRY(upload_planes(c, warpIm2, N, kFeat, wk.Iw));
  TRY(upload(c, *D1, dm1, N * 8));
  TRY(upload(c, *D2, dm2, N * 8));
  TRY(upload(c, wk.Dw, warpDm2, N * 8));
  TRY(upload(c, wk.u[0], u, N * 8));
  TRY(upload(c, wk.u[0] + N, v, N * 8));
  TRY(upload(c, wk.u[0] + 2 * N, w, N * 8));
  double *d, *s, *m, *r;
  TRY(get(c, "lvW0", N, &d));
  TRY(get(c, "lvW1", N, &s));
  TRY(get(c, "lvW2", N, &m));
  TRY(get(c, "lvW3", 2 * N, &r));
  TRY(upload(c, d, dW64, N * 8));
  TRY(upload(c, s, sW64, N * 8));
  TRY(upload(c, m, mW64, N * 8));
  TRY(upload(c, r, r64, 2 * N * 8));
  launch_level_weights_f64(c->stream, d, s, m, r, *D1, W, H, focal_level, wk.dW, wk.sW, wk.mW, wk.rW, wk.cW);
  return RGBDFLOW_OK;
}

bool bad_dims(int W, int H) { return W < 2 || H < 2 || (long long)W * H > (1LL << 28); }

}  // namespace

// ============================================================================
extern "C" {

int rgbdflow_default_params(rgbdflow_params* p) {
  if (!p) return RGBDFLOW_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->variant = RGBDFLOW_VARIANT_FULL;
  p->regularization_type = RGBDFLOW_REG_HPB;
  p->color_data_weight = 1;
  p->depth_vs_color_data_weight = 1;
  p->alpha = 0;
  p->gamma = 0;
  p->pyramid_ratio = .75;
  p->min_width = 20;
  p->n_outer = 7;
  p->n_inner = 1;
  p->n_sor = 30;
  p->relax_variant = RGBDFLOW_RELAX_PLAIN;
  p->inner_exit = 1;
  p->outer_exit = 1;
  p->tile_config = -1;
  return RGBDFLOW_OK;
}

int rgbdflow_create(rgbdflow_ctx** out, int device, unsigned long long budget) {
  if (!out) return RGBDFLOW_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return RGBDFLOW_ERR_HIP;
  if (device < 0 || device >= n) return RGBDFLOW_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return RGBDFLOW_ERR_HIP;
  rgbdflow_ctx* c = new rgbdflow_ctx();
  c->device = device;
  c->budget = budget;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return RGBDFLOW_ERR_HIP;
  }
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0) c->cu_count = v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
    c->lds_per_block = (size_t)v;
  *out = c;
  return RGBDFLOW_OK;
}

int rgbdflow_destroy(rgbdflow_ctx* c) {
  if (!c) return RGBDFLOW_ERR_ARG;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  for (auto& kv : c->bufs)
    if (kv.second.p) hipFree(kv.second.p);
  for (hipEvent_t e : c->relax_ev) hipEventDestroy(e);
  hipEventDestroy(c->ev0);
  hipEventDestroy(c->ev1);
  hipStreamDestroy(c->stream);
  delete c;
  return RGBDFLOW_OK;
}

const char* rgbdflow_last_error(const rgbdflow_ctx* c) { return c ? c->err.c_str() : kNoCtx; }
unsigned long long rgbdflow_bytes_in_use(const rgbdflow_ctx* c) { return c ? c->in_use : 0; }

int rgbdflow_synchronize(rgbdflow_ctx* c) {
  if (!c) return RGBDFLOW_ERR_ARG;
  return sync(c);
}

int rgbdflow_set_relax_timing(rgbdflow_ctx* c, int on) {
  if (!c) return RGBDFLOW_ERR_ARG;
  c->time_relax = on != 0;
  return RGBDFLOW_OK;
}

int rgbdflow_pyramid_levels(int W, int H, double ratio, int min_width, int* widths, int* heights) {
  if (bad_dims(W, H)) return RGBDFLOW_ERR_ARG;
  const std::vector<LevelShape> L = pyramid_shape(W, H, ratio, min_width, nullptr, nullptr);
  if ((int)L.size() > kMaxLevels) return RGBDFLOW_ERR_UNSUPPORTED;
  for (size_t i = 0; i < L.size(); i++) {
    if (widths) widths[i] = L[i].W;
    if (heights) heights[i] = L[i].H;
  }
  return (int)L.size();
}

int rgbdflow_solve_level(rgbdflow_ctx* c, int W, int H, float focal_level, const double* Im1, const double* dm1,
                         const double* Im2, const double* dm2, const double* warpIm2, const double* warpDm2,
                         double* u, double* v, double* w, double cw, const double* dW, double alpha,
                         const double* sW, double gamma, const double* mW, const double* rW, int nOuter, int nInner,
                         int nSOR, int allowEarly, const rgbdflow_params* opts, rgbdflow_stats* st) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H)) return fail(c, RGBDFLOW_ERR_ARG, "bad image size");
  if (!Im1 || !dm1 || !Im2 || !dm2 || !warpIm2 || !warpDm2 || !u || !v || !w || !dW || !sW || !mW || !rW)
    return fail(c, RGBDFLOW_ERR_ARG, "null input");
  TRY(validate_iters(c, nOuter, nInner, nSOR));
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  Work wk;
  double *D1, *D2;
  TRY(stage_level_inputs(c, W, H, focal_level, Im1, dm1, Im2, dm2, warpIm2, warpDm2, u, v, w, dW, sW, mW, rW, wk,
                         &D1, &D2));
  const SolveScalars k = scalars_from(opts, cw, alpha, gamma, nOuter, nInner, nSOR, allowEarly);
  LevelView view = make_view(wk, W, H, wk.I1, wk.I2, D1, D2, wk.u[0], 0);
  c->relax_ev_used = 0;
  HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
  enqueue_level_solve(c, view, k, make_relax_plan(c, W, H, opts));
  HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
  TRY(download(c, u, wk.u[0], N * 8));
  TRY(download(c, v, wk.u[0] + N, N * 8));
  TRY(download(c, w, wk.u[0] + 2 * N, N * 8));
  SolveCtl ctl;
  TRY(download(c, &ctl, view.ctl, sizeof(SolveCtl)));
  TRY(sync(c));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->n_levels = 1;
    st->width[0] = W;
    st->height[0] = H;
    ctl_to_stats(ctl, k, st->outer_exec[0], st->sweeps_exec[0], st->last_stat[0]);
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    st->device_ms = ms;
    collect_relax_timing(c, st);
  }
  return RGBDFLOW_OK;
}

int rgbdflow_level_system(rgbdflow_ctx* c, int W, int H, float focal_level, const double* Im1, const double* dm1,
                          const double* Im2, const double* dm2, const double* warpIm2, const double* warpDm2,
                          const double* u, const double* v, const double* w, double cw, const double* dW,
                          double alpha, const double* sW, double gamma, const double* mW, const double* rW,
                          double* coef13) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H)) return fail(c, RGBDFLOW_ERR_ARG, "bad image size");
  if (!Im1 || !dm1 || !Im2 || !dm2 || !warpIm2 || !warpDm2 || !u || !v || !w || !dW || !sW || !mW || !rW || !coef13)
    return fail(c, RGBDFLOW_ERR_ARG, "null input");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  Work wk;
  double *D1, *D2;
  TRY(stage_level_inputs(c, W, H, focal_level, Im1, dm1, Im2, dm2, warpIm2, warpDm2, u, v, w, dW, sW, mW, rW, wk,
                         &D1, &D2));
  const SolveScalars k = scalars_from(nullptr, cw, alpha, gamma, 1, 1, 0, 0);
  LevelView view = make_view(wk, W, H, wk.I1, wk.I2, D1, D2, wk.u[0], 0);
  launch_ctl_reset(c->stream, view.ctl);
  launch_derivs(c->stream, view, 0);
  launch_phi(c->stream, view, 0);
  launch_system(c->stream, view, k, 0);
  std::vector<float> host((size_t)kCoef * N);
  TRY(download(c, host.data(), wk.coef, host.size() * sizeof(float)));
  TRY(sync(c));
  for (size_t i = 0; i < host.size(); i++) coef13[i] = host[i];
  return RGBDFLOW_OK;
}

int rgbdflow_relax(rgbdflow_ctx* c, const double* coef13, int W, int H, int n_iter, const rgbdflow_params* sched,
                   double* x) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !coef13 || !x || n_iter < 0) return fail(c, RGBDFLOW_ERR_ARG, "bad relax arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  Work wk;
  TRY(get_work(c, N, wk));
  std::vector<float> hc((size_t)kCoef * N), hx(3 * N);
  for (size_t i = 0; i < hc.size(); i++) hc[i] = (float)coef13[i];
  for (size_t i = 0; i < hx.size(); i++) hx[i] = (float)x[i];
  TRY(upload(c, wk.coef, hc.data(), hc.size() * 4));
  TRY(upload(c, wk.du, hx.data(), hx.size() * 4));
  LevelView view = make_view(wk, W, H, wk.I1, wk.I2, nullptr, nullptr, wk.u[0], 0);
  launch_ctl_reset(c->stream, view.ctl);
  const RelaxPlan plan = make_relax_plan(c, W, H, sched);
  SolveScalars k{};
  k.n_sor = n_iter;
  k.inner_exit = 0;
  // checks are recorded but never stop this call (inner_exit = 0)
  if (n_iter / kCheckEvery + 1 > kMaxChecks) return fail(c, RGBDFLOW_ERR_ARG, "n_iter too large");
  enqueue_relax(c, view, k, plan, 0);
  TRY(download(c, hx.data(), wk.du, hx.size() * 4));
  TRY(sync(c));
  for (size_t i = 0; i < hx.size(); i++) x[i] = hx[i];
  return RGBDFLOW_OK;
}

int rgbdflow_depth_stdevs_pointwise(rgbdflow_ctx* c, const float* depth, int W, int H, int max_dist, float* out) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !depth || !out || max_dist < 0) return fail(c, RGBDFLOW_ERR_ARG, "bad sigma arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  float *d, *s;
  TRY(get(c, "sig_in", N, &d));
  TRY(get(c, "sig_out", N, &s));
  TRY(upload(c, d, depth, N * 4));
  launch_depth_sigma(c->stream, d, W, H, max_dist, s);
  TRY(download(c, out, s, N * 4));
  return sync(c);
}

int rgbdflow_gaussian_pyramid(rgbdflow_ctx* c, const double* img, int W, int H, int C, double ratio, int min_width,
                              double* out) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !img || !out || C < 1 || C > 8) return fail(c, RGBDFLOW_ERR_ARG, "bad pyramid arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  int ns;
  double r;
  const std::vector<LevelShape> L = pyramid_shape(W, H, ratio, min_width, &ns, &r);
  const size_t N = (size_t)W * H;
  double *base, *rest;
  TRY(get(c, "gp_base", N * C, &base));
  TRY(get(c, "gp_rest", pyramid_elems(L, 1) * C + 1, &rest));
  TRY(upload_planes(c, img, N, C, base));
  std::vector<double*> lvl(L.size());
  lvl[0] = base;
  size_t off = 0;
  for (size_t i = 1; i < L.size(); i++) {
    lvl[i] = rest + off;
    off += (size_t)L[i].W * L[i].H * C;
  }
  build_pyramid(c->stream, L, ns, r, C, lvl);
  size_t hoff = 0;
  for (size_t i = 0; i < L.size(); i++) {
    const size_t n = (size_t)L[i].W * L[i].H;
    TRY(download_planes(c, lvl[i], n, C, out + hoff));
    TRY(sync(c));
    hoff += n * C;
  }
  return RGBDFLOW_OK;
}

int rgbdflow_depth_map_pyramid(rgbdflow_ctx* c, const double* depth, int W, int H, double ratio, int min_width,
                               double* out) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !depth || !out) return fail(c, RGBDFLOW_ERR_ARG, "bad depth pyramid arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  double r;
  const std::vector<LevelShape> L = pyramid_shape(W, H, ratio, min_width, nullptr, &r);
  std::vector<LevelShape> S(L.size());
  size_t tot = 0;
  for (size_t m = 0; m < L.size(); m++) {
    const double lr = std::pow(r, (int)m);
    S[m] = {(int)(W * lr), (int)(H * lr)};
    tot += (size_t)S[m].W * S[m].H;
  }
  double* d;
  TRY(get(c, "dmp", tot, &d));
  TRY(upload(c, d, depth, (size_t)W * H * 8));
  size_t off = (size_t)W * H, prev = 0;
  for (size_t m = 1; m < S.size(); m++) {
    launch_depth_map_level(c->stream, d + prev, S[m - 1].W, S[m - 1].H, d + off, S[m].W, S[m].H);
    prev = off;
    off += (size_t)S[m].W * S[m].H;
  }
  TRY(download(c, out, d, tot * 8));
  return sync(c);
}

int rgbdflow_im2feature(rgbdflow_ctx* c, const double* bgr, int W, int H, double* out5) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !bgr || !out5) return fail(c, RGBDFLOW_ERR_ARG, "bad feature arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  double *in, *f;
  TRY(get(c, "feat_in", 3 * N, &in));
  TRY(get(c, "feat_out", kFeat * N, &f));
  TRY(upload_planes(c, bgr, N, 3, in));
  launch_features(c->stream, in, W, H, f);
  TRY(download_planes(c, f, N, kFeat, out5));
  return sync(c);
}

int rgbdflow_warp_by_3d_flow(rgbdflow_ctx* c, const double* I1, const double* D1, const double* I2, const double* D2,
                             const double* u, const double* v, const double* w, int W, int H, int C, double* wI,
                             double* wD) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (C != kFeat) return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "warp_by_3d_flow: C must be 5 (feature images)");
  if (bad_dims(W, H) || !I1 || !D1 || !I2 || !D2 || !u || !v || !w || !wI || !wD)
    return fail(c, RGBDFLOW_ERR_ARG, "bad warp arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  Work wk;
  TRY(get_work(c, N, wk));
  double *d1, *d2;
  TRY(get(c, "lvD1", N, &d1));
  TRY(get(c, "lvD2", N, &d2));
  TRY(upload_planes(c, I1, N, kFeat, wk.I1));
  TRY(upload_planes(c, I2, N, kFeat, wk.I2));
  TRY(upload(c, d1, D1, N * 8));
  TRY(upload(c, d2, D2, N * 8));
  TRY(upload(c, wk.u[0], u, N * 8));
  TRY(upload(c, wk.u[0] + N, v, N * 8));
  TRY(upload(c, wk.u[0] + 2 * N, w, N * 8));
  LevelView view = make_view(wk, W, H, wk.I1, wk.I2, d1, d2, wk.u[0], 0);
  launch_warp_cpu_style(c->stream, view);
  TRY(download_planes(c, wk.Iw, N, kFeat, wI));
  TRY(download(c, wD, wk.Dw, N * 8));
  return sync(c);
}

int rgbdflow_upsample_flow(rgbdflow_ctx* c, const double* u, const double* v, const double* w, int Ws, int Hs, int Wd,
                           int Hd, double uv_scale, double* uo, double* vo, double* wo) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(Ws, Hs) || bad_dims(Wd, Hd) || !u || !v || !w || !uo || !vo || !wo)
    return fail(c, RGBDFLOW_ERR_ARG, "bad upsample arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t Ns = (size_t)Ws * Hs, Nd = (size_t)Wd * Hd;
  double *a, *b;
  TRY(get(c, "ups_in", 3 * Ns, &a));
  TRY(get(c, "ups_out", 3 * Nd, &b));
  TRY(upload(c, a, u, Ns * 8));
  TRY(upload(c, a + Ns, v, Ns * 8));
  TRY(upload(c, a + 2 * Ns, w, Ns * 8));
  launch_upsample(c->stream, a, Ws, Hs, b, Wd, Hd, uv_scale);
  TRY(download(c, uo, b, Nd * 8));
  TRY(download(c, vo, b + Nd, Nd * 8));
  TRY(download(c, wo, b + 2 * Nd, Nd * 8));
  return sync(c);
}

// ---------------------------------------------------------------- pair path
int rgbdflow_pair_load(rgbdflow_ctx* c, const rgbdflow_params* p, int W, int H, const double* im1, const double* dm1,
                       const double* im2, const double* dm2, const float* sigma1, const float* hpb) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!p || bad_dims(W, H) || !im1 || !dm1 || !im2 || !dm2) return fail(c, RGBDFLOW_ERR_ARG, "bad pair arguments");
  if (p->regularization_type == RGBDFLOW_REG_HPB && !hpb)
    return fail(c, RGBDFLOW_ERR_ARG, "HPB regularisation needs the boundary maps (hpb)");
  HIP_TRY(c, hipSetDevice(c->device));
  c->pair_loaded = false;
  const size_t N = (size_t)W * H;
  const std::vector<LevelShape> L = pyramid_shape(W, H, p->pyramid_ratio, p->min_width, nullptr, nullptr);
  if ((int)L.size() > kMaxLevels) return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "too many pyramid levels");
  double *i1, *i2, *d1, *d2;
  float* s1;
  TRY(get(c, "p_im1", 3 * N, &i1));
  TRY(get(c, "p_im2", 3 * N, &i2));
  TRY(get(c, "p_dm1", N, &d1));
  TRY(get(c, "p_dm2", N, &d2));
  TRY(get(c, "p_sig1", N, &s1));
  TRY(upload_planes(c, im1, N, 3, i1));
  TRY(upload_planes(c, im2, N, 3, i2));
  TRY(upload(c, d1, dm1, N * 8));
  TRY(upload(c, d2, dm2, N * 8));
  if (sigma1) TRY(upload(c, s1, sigma1, N * 4));
  if (hpb) {
    float* h;
    const size_t nh = pyramid_elems(L, 0) * 2;
    TRY(get(c, "p_hpb", nh, &h));
    TRY(upload(c, h, hpb, nh * 4));
  }
  TRY(sync(c));
  c->pW = W;
  c->pH = H;
  c->p_ratio = p->pyramid_ratio;
  c->p_min_width = p->min_width;
  c->p_sigma_given = sigma1 != nullptr;
  c->p_hpb_given = hpb != nullptr;
  c->p_levels = L;
  c->pair_loaded = true;
  c->pair_ran = false;
  return RGBDFLOW_OK;
}

int rgbdflow_pair_run(rgbdflow_ctx* c, const rgbdflow_params* p, rgbdflow_stats* st) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!p) return fail(c, RGBDFLOW_ERR_ARG, "null params");
  if (!c->pair_loaded) return fail(c, RGBDFLOW_ERR_STATE, "pair_run before pair_load");
  if (p->pyramid_ratio != c->p_ratio || p->min_width != c->p_min_width)
    return fail(c, RGBDFLOW_ERR_STATE, "pyramid shape differs from pair_load's params");
  if ((int)p->cam.xRes != c->pW || (int)p->cam.yRes != c->pH)
    return fail(c, RGBDFLOW_ERR_ARG, "params.cam resolution differs from the image size");
  if (p->variant != RGBDFLOW_VARIANT_FULL) return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "bad scene flow variant");
  if (p->regularization_type == RGBDFLOW_REG_HPB && !c->p_hpb_given)
    return fail(c, RGBDFLOW_ERR_ARG, "HPB regularisation needs the boundary maps (hpb)");
  if (p->regularization_type < 0 || p->regularization_type > 2) return fail(c, RGBDFLOW_ERR_ARG, "bad regularization");
  TRY(validate_iters(c, p->n_outer, p->n_inner, p->n_sor));
  HIP_TRY(c, hipSetDevice(c->device));
  const int W = c->pW, H = c->pH;
  const size_t N = (size_t)W * H;
  int ns;
  double ratio;
  const std::vector<LevelShape> L = pyramid_shape(W, H, p->pyramid_ratio, p->min_width, &ns, &ratio);
  const int nL = (int)L.size();
  const size_t rest = pyramid_elems(L, 1);
  double *i1, *i2, *d1, *d2, *pi1, *pi2, *pd1, *pd2, *pdel, *psm, *del0, *sm0;
  float *s1, *hpb = nullptr, *dm1f;
  TRY(get(c, "p_im1", 3 * N, &i1));
  TRY(get(c, "p_im2", 3 * N, &i2));
  TRY(get(c, "p_dm1", N, &d1));
  TRY(get(c, "p_dm2", N, &d2));
  TRY(get(c, "p_sig1", N, &s1));
  if (c->p_hpb_given) TRY(get(c, "p_hpb", pyramid_elems(L, 0) * 2, &hpb));
  TRY(get(c, "py_im1", 3 * rest + 1, &pi1));
  TRY(get(c, "py_im2", 3 * rest + 1, &pi2));
  TRY(get(c, "py_d1", rest + 1, &pd1));
  TRY(get(c, "py_d2", rest + 1, &pd2));
  TRY(get(c, "py_del", rest + 1, &pdel));
  TRY(get(c, "py_sm", rest + 1, &psm));
  TRY(get(c, "p_del0", N, &del0));
  TRY(get(c, "p_sm0", N, &sm0));
  TRY(get(c, "p_dm1f", N, &dm1f));
  Work wk;
  TRY(get_work(c, N, wk));
  hipStream_t s = c->stream;
  c->relax_ev_used = 0;
  HIP_TRY(c, hipEventRecord(c->ev0, s));

  // per-pair weights: depth stdevs (runFlowTwoFrames.cpp:89-91) -> delta
  // (OpticalFlow.cpp:71-81); s = f^2 (OpticalFlow.cpp:90-97)
  if (!c->p_sigma_given) {
    launch_to_f32(s, d1, N, dm1f);
    launch_depth_sigma(s, dm1f, W, H, 4, s1);
  }
  launch_depth_data_weights(s, s1, (int)N, p->depth_vs_color_data_weight, del0);
  const double f0 = p->cam.focalLength;
  launch_fill_f64(s, sm0, N, f0 * f0);

  // the six pyramids of sceneFlowMultiscale.cpp:40-85 (the raw-depth pyramid
  // feeds only the boundary probabilities, which the caller supplies)
  auto levels_of = [&](double* base, double* restp, int C) {
    std::vector<double*> lv(nL);
    lv[0] = base;
    size_t off = 0;
    for (int i = 1; i < nL; i++) {
      lv[i] = restp + off;
      off += (size_t)L[i].W * L[i].H * C;
    }
    return lv;
  };
  std::vector<double*> P1 = levels_of(i1, pi1, 3), P2 = levels_of(i2, pi2, 3), PD1 = levels_of(d1, pd1, 1),
                       PD2 = levels_of(d2, pd2, 1), PDel = levels_of(del0, pdel, 1), PS = levels_of(sm0, psm, 1);
  build_pyramid(s, L, ns, ratio, 3, P1);
  build_pyramid(s, L, ns, ratio, 3, P2);
  build_pyramid(s, L, ns, ratio, 1, PD1);
  build_pyramid(s, L, ns, ratio, 1, PD2);
  build_pyramid(s, L, ns, ratio, 1, PDel);
  build_pyramid(s, L, ns, ratio, 1, PS);

  std::vector<size_t> hpb_off(nL, 0);
  for (int k = 1; k < nL; k++) hpb_off[k] = hpb_off[k - 1] + (size_t)L[k - 1].W * L[k - 1].H * 2;

  std::vector<SolveScalars> ks(nL);
  int cur = 0;
  for (int k = nL - 1; k >= 0; k--) {
    const int Wk = L[k].W, Hk = L[k].H;
    const size_t Nk = (size_t)Wk * Hk;
    launch_features(s, P1[k], Wk, Hk, wk.I1);
    launch_features(s, P2[k], Wk, Hk, wk.I2);
    const float fk = downsized_focal(p->cam, Wk);
    LevelView view;
    if (k == nL - 1) {
      cur = 0;
      view = make_view(wk, Wk, Hk, wk.I1, wk.I2, PD1[k], PD2[k], wk.u[cur], k);
      launch_fill_f64(s, wk.u[cur], 3 * Nk, 0.0);
      HIP_TRY(c, hipMemcpyAsync(wk.Iw, wk.I2, kFeat * Nk * 8, hipMemcpyDeviceToDevice, s));
      HIP_TRY(c, hipMemcpyAsync(wk.Dw, PD2[k], Nk * 8, hipMemcpyDeviceToDevice, s));
    } else {
      const int nxt = 1 - cur;
      launch_upsample(s, wk.u[cur], L[k + 1].W, L[k + 1].H, wk.u[nxt], Wk, Hk, 1 / p->pyramid_ratio);
      cur = nxt;
      view = make_view(wk, Wk, Hk, wk.I1, wk.I2, PD1[k], PD2[k], wk.u[cur], k);
      launch_warp_cpu_style(s, view);
    }
    launch_level_weights(s, PDel[k], PS[k], hpb ? hpb + hpb_off[k] : nullptr, p->regularization_type, PD1[k], Wk, Hk,
                         (int)p->cam.xRes, p->cam.focalLength, fk, wk.dW, wk.sW, wk.mW, wk.rW, wk.cW);
    ks[k] = scalars_from(p, p->color_data_weight, p->alpha, p->gamma, p->n_outer, p->n_inner, p->n_sor, k < nL - 1);
    enqueue_level_solve(c, view, ks[k], make_relax_plan(c, Wk, Hk, p));
  }
  c->p_final_slot = cur;
  HIP_TRY(c, hipEventRecord(c->ev1, s));
  HIP_TRY(c, hipGetLastError());
  c->pair_ran = true;
  if (!st) return RGBDFLOW_OK;
  std::vector<SolveCtl> ctl(nL);
  TRY(download(c, ctl.data(), wk.ctl, nL * sizeof(SolveCtl)));
  TRY(sync(c));
  std::memset(st, 0, sizeof(*st));
  st->n_levels = nL;
  for (int k = 0; k < nL; k++) {
    st->width[k] = L[k].W;
    st->height[k] = L[k].H;
    ctl_to_stats(ctl[k], ks[k], st->outer_exec[k], st->sweeps_exec[k], st->last_stat[k]);
  }
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  st->device_ms = ms;
  collect_relax_timing(c, st);
  return RGBDFLOW_OK;
}

int rgbdflow_pair_read(rgbdflow_ctx* c, double* u, double* v, double* w) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!c->pair_ran) return fail(c, RGBDFLOW_ERR_STATE, "pair_read before pair_run");
  if (!u || !v || !w) return fail(c, RGBDFLOW_ERR_ARG, "null output");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)c->pW * c->pH;
  double* uu;
  TRY(get(c, c->p_final_slot ? "u1" : "u0", 3 * N, &uu));
  TRY(download(c, u, uu, N * 8));
  TRY(download(c, v, uu + N, N * 8));
  TRY(download(c, w, uu + 2 * N, N * 8));
  return sync(c);
}

int rgbdflow_multiscale(rgbdflow_ctx* c, int W, int H, const double* im1, const double* dm1, const double* im2,
                        const double* dm2, const float* sigma1, const float* hpb, const rgbdflow_params* p, double* u,
                        double* v, double* w, rgbdflow_stats* st) {
  TRY(rgbdflow_pair_load(c, p, W, H, im1, dm1, im2, dm2, sigma1, hpb));
  rgbdflow_stats local;
  TRY(rgbdflow_pair_run(c, p, st ? st : &local));
  return rgbdflow_pair_read(c, u, v, w);
}

int rgbdflow_autotune_relax(rgbdflow_ctx* c, int W, int H, rgbdflow_tile_choice* out) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H)) return fail(c, RGBDFLOW_ERR_ARG, "bad autotune size");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  Work wk;
  TRY(get_work(c, N, wk));
  LevelView view = make_view(wk, W, H, wk.I1, wk.I2, nullptr, nullptr, wk.u[0], 0);
  hipStream_t s = c->stream;
  launch_ctl_reset(s, view.ctl);
  launch_fill_test_system(s, wk.coef, wk.du, W, H);
  SolveScalars k{};
  k.n_sor = kCheckEvery;  // one check interval: the unit the schedule repeats
  k.inner_exit = 0;
  // candidate schedules
  std::vector<TileSched> cand;
  const long long pairs_img = (long long)((W + 1) / 2) * H;
  const int Ts[] = {1, 2, 3, 4, 5, 6, 8, 12};
  const int TWs[] = {8, 16, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320};
  for (int cfg = 0; cfg < tiled_config_count(); cfg++) {
    if (tiled_config_lds(cfg) > c->lds_per_block) continue;
    int nt, pp;
    tiled_config_shape(cfg, &nt, &pp);
    const int cap = nt * pp;
    if (pairs_img <= cap) {
      cand.push_back(TileSched{W + (W & 1), H, 1, 1, 1, cap, cfg, kCheckEvery});
      continue;
    }
    for (int T : Ts) {
      const int h = 2 * T;
      for (int TW : TWs) {
        if (TW > W + (W & 1)) continue;
        const int rw = std::min(TW + 2 * h, W + (W & 1));
        const int rows = cap / ((rw + 1) / 2);
        const int thmax = std::min(rows - 2 * h, H);
        if (thmax < 2) continue;
        for (int TH : {thmax, thmax / 2}) {
          if (TH < 2) continue;
          if (((rw + 1) / 2) * (long long)std::min(TH + 2 * h, H) > cap) continue;
          cand.push_back(TileSched{TW, TH, (W + TW - 1) / TW, (H + TH - 1) / TH, 0, cap, cfg, T});
        }
      }
    }
  }
  if (cand.empty()) return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "no tiled schedule fits this level");
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  double best_us = 1e300;
  TileSched best{};
  const bool timing = c->time_relax;
  c->time_relax = false;
  for (const TileSched& ts : cand) {
    RelaxPlan plan;
    plan.variant = RGBDFLOW_RELAX_TILED;
    plan.ts = ts;
    enqueue_relax(c, view, k, plan, 0);  // warm-up
    double t_best = 1e300;
    for (int rep = 0; rep < 3; rep++) {
      hipEventRecord(e0, s);
      enqueue_relax(c, view, k, plan, 0);
      hipEventRecord(e1, s);
      hipEventSynchronize(e1);
      float ms = 0;
      hipEventElapsedTime(&ms, e0, e1);
      t_best = std::min(t_best, (double)ms);
    }
    const double us = t_best * 1e3 / kCheckEvery;
    if (us < best_us) {
      best_us = us;
      best = ts;
    }
  }
  c->time_relax = timing;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  TRY(sync(c));
  c->plans[plan_key(W, H, 0, 0, 0, -1)] = best;
  if (out) {
    out->tile_w = best.TW;
    out->tile_h = best.TH;
    out->tile_iters = best.tmax;
    out->tile_config = best.config;
    tiled_config_shape(best.config, &out->threads, &out->pairs_per_thread);
    out->single = best.single;
    out->candidates = (int)cand.size();
    out->us_per_iter = best_us;
  }
  return RGBDFLOW_OK;
}

int rgbdflow_with_init(rgbdflow_ctx* c, int W, int H, const double* im1, const double* dm1, const double* im2,
                       const double* dm2, const float* sigma1, const float* hpb0, const double* est,
                       const rgbdflow_params* p, double* u, double* v, double* w, rgbdflow_stats* st) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!p || bad_dims(W, H) || !im1 || !dm1 || !im2 || !dm2 || !est || !u || !v || !w)
    return fail(c, RGBDFLOW_ERR_ARG, "bad with_init arguments");
  if (p->regularization_type == RGBDFLOW_REG_HPB && !hpb0)
    return fail(c, RGBDFLOW_ERR_ARG, "HPB regularisation needs the boundary map (hpb0)");
  if ((int)p->cam.xRes != W || (int)p->cam.yRes != H)
    return fail(c, RGBDFLOW_ERR_ARG, "params.cam resolution differs from the image size");
  TRY(validate_iters(c, p->n_outer, p->n_inner, p->n_sor));
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  Work wk;
  TRY(get_work(c, N, wk));
  double *i1, *i2, *d1, *d2, *del0, *sm0, *est_d;
  float *s1, *h = nullptr, *dm1f;
  TRY(get(c, "p_im1", 3 * N, &i1));
  TRY(get(c, "p_im2", 3 * N, &i2));
  TRY(get(c, "p_dm1", N, &d1));
  TRY(get(c, "p_dm2", N, &d2));
  TRY(get(c, "p_sig1", N, &s1));
  TRY(get(c, "p_del0", N, &del0));
  TRY(get(c, "p_sm0", N, &sm0));
  TRY(get(c, "p_dm1f", N, &dm1f));
  TRY(get(c, "wi_est", 3 * N, &est_d));
  c->pair_loaded = false;
  c->pair_ran = false;
  hipStream_t s = c->stream;
  TRY(upload_planes(c, im1, N, 3, i1));
  TRY(upload_planes(c, im2, N, 3, i2));
  TRY(upload(c, d1, dm1, N * 8));
  TRY(upload(c, d2, dm2, N * 8));
  TRY(upload_planes(c, est, N, 3, wk.u[0]));
  if (hpb0) {
    TRY(get(c, "p_hpb", N * 2, &h));
    TRY(upload(c, h, hpb0, N * 2 * 4));
  }
  if (sigma1) {
    TRY(upload(c, s1, sigma1, N * 4));
  } else {
    launch_to_f32(s, d1, N, dm1f);
    launch_depth_sigma(s, dm1f, W, H, 4, s1);
  }
  c->relax_ev_used = 0;
  HIP_TRY(c, hipEventRecord(c->ev0, s));
  launch_depth_data_weights(s, s1, (int)N, p->depth_vs_color_data_weight, del0);
  const double f0 = p->cam.focalLength;
  launch_fill_f64(s, sm0, N, f0 * f0);
  launch_features(s, i1, W, H, wk.I1);
  launch_features(s, i2, W, H, wk.I2);
  LevelView view = make_view(wk, W, H, wk.I1, wk.I2, d1, d2, wk.u[0], 0);
  launch_warp_cpu_style(s, view);
  launch_level_weights(s, del0, sm0, h, p->regularization_type, d1, W, H, (int)p->cam.xRes, p->cam.focalLength,
                       p->cam.focalLength, wk.dW, wk.sW, wk.mW, wk.rW, wk.cW);
  const SolveScalars k = scalars_from(p, p->color_data_weight, p->alpha, p->gamma, p->n_outer, p->n_inner, p->n_sor, 0);
  enqueue_level_solve(c, view, k, make_relax_plan(c, W, H, p));
  HIP_TRY(c, hipEventRecord(c->ev1, s));
  TRY(download(c, u, wk.u[0], N * 8));
  TRY(download(c, v, wk.u[0] + N, N * 8));
  TRY(download(c, w, wk.u[0] + 2 * N, N * 8));
  SolveCtl ctl;
  TRY(download(c, &ctl, view.ctl, sizeof(SolveCtl)));
  TRY(sync(c));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->n_levels = 1;
    st->width[0] = W;
    st->height[0] = H;
    ctl_to_stats(ctl, k, st->outer_exec[0], st->sweeps_exec[0], st->last_stat[0]);
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    st->device_ms = ms;
    collect_relax_timing(c, st);
  }
  return RGBDFLOW_OK;
}

}  // extern "C"
