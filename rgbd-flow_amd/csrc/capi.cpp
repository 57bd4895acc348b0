// capi.cpp -- the C ABI of librgbdflow: context, HBM buffer pool and the
// device-resident orchestration of batched coarse-to-fine solves.
//
// A batch of frame pairs (any sizes) is solved in lock step: pyramids are
// aligned on their finest level, and step s of the batch runs, for every pair
// that has it, the level S-1-s (S = the largest level count).  Each launch
// serves all pairs of a step through a work-list (internal.hpp), so a batch
// of B pairs costs the launches of one pair.  The host enqueues the whole
// batch without waiting: early exits are decided on the device (SolveCtl per
// pair and level), and the host reads the control blocks back once, after
// the batch, to report iteration counts.  A single pair is a batch of one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rgbdflow/rgbdflow.h"
#include "capi_internal.hpp"
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

// Work buffers of one solve slot, sized for its finest level; a level of N
// pixels uses the first N entries of every plane.
struct Work {
  double *I1, *I2, *Iw, *Dw, *u[2], *Ix, *Iy, *It, *Dx, *Dy, *Dt, *phi, *partials;
  float *dW, *sW, *mW, *rW, *cW, *q;
  float4 *du, *du_alt;  // [N] (du, dv, dw, unused)
  float4* coef;         // [N][kRec] system records
  SolveCtl* ctl;  // kMaxLevels entries, one per level
  // coarse-grid correction (allocated when asked for): the coarse levels of
  // the level being solved, packed level after level
  float4 *cg_coef, *cg_du, *cg_du_alt;
  SolveCtl* cg_ctl;  // kMaxCg entries, one per coarse level
};

// One loaded frame pair: inputs, pyramids (levels >= 1 packed finest first)
// and its work buffers.
struct PairSlot {
  int W = 0, H = 0, min_width = 0;
  bool cam_given = false, sigma_given = false, hpb_given = false;
  bool hpb_device = false;  // boundary maps computed on the device (HPB and none given)
  bool pyr1_given = false;  // frame 1's colour and depth pyramids were loaded (not built)
  bool init = false;        // solved as sceneFlowWithInit: finest level only, from the estimate in wk.u[1]
  rgbdflow_camera cam{};
  std::vector<LevelShape> L;
  std::vector<size_t> off;   // level k >= 1: element offset (per channel) in the pyramid arrays
  std::vector<size_t> hoff;  // level k: float offset of its boundary map
  double *im1 = nullptr, *im2 = nullptr, *dm1 = nullptr, *dm2 = nullptr, *del0 = nullptr, *sm0 = nullptr;
  double *pim1 = nullptr, *pim2 = nullptr, *pd1 = nullptr, *pd2 = nullptr, *pdel = nullptr, *psm = nullptr;
  float *sigma = nullptr, *hpb = nullptr;
  double* pdm = nullptr;  // depth-map pyramid of dm1, levels >= 1 (hpb_device)
  double* dmh = nullptr;  // its level 0 when it is not dm1 (the colour-filled depth), else null
  Work wk{};
};

// Device work-lists of one step of a batch.
struct StepPlan {
  int n = 0;
  int pix = 0, plain = 0, tiled = 0, otile = 0;  // total blocks of each kernel family
  size_t items_off = 0, pix_off = 0, plain_off = 0, tiled_off = 0, otile_off = 0;
  Views vpix{}, vplain{}, vtiled{}, votile{};
  bool tiled_ok = false;
  int cfg = 0, slots = 0, T = 0;  // tiled launch shape: workgroup shape, LDS slots, iterations
  int grid = 0;                    // tiled launch grid (tiled_prepare, when the step is packed)
  int storage = 0;                 // system record storage of the step's items
  int img = 0;                     // feature-image storage of the step's items
  int relax_mode = kRelaxFixed;    // RelaxMode of the step's items
  bool any_early = false;
  // coarse-grid correction: the coarse levels 1, 2, ... of the step's items
  // (level l holds the items that have it; their operators), the same
  // levels' items that relax level by level (the rest of each item's levels
  // fit in LDS and run in k_cg_bottom), and the k_cg_bottom work-list
  std::vector<StepPlan> cg, cg_nf;
  size_t cg_bottom_off = 0;
  int cg_bottom_n = 0;
  Views vbottom{};
};

struct BatchPlan {
  bool valid = false;
  unsigned long long gen = 0, plans_gen = 0;
  std::vector<unsigned char> params;
  std::vector<char> host;  // packed descriptors, uploaded to "desc"
  std::vector<StepPlan> steps;
  size_t prep_items = 0, prep_start = 0;
  int prep_n = 0, prep_blocks = 0;
  struct Wave {
    size_t items, start;
    int n, blocks;
    WorkList<PyrJob> wl;
  };
  std::vector<Wave> waves;
  WorkList<PairPrep> prep{};
  // boundary maps computed on the device, level by level: the depth-map
  // pyramid level (from the next larger one; level 0 is dm1) and the level's
  // computeHeuristicPb
  struct DmStep {
    const double* larger;  // the next larger level
    int LW, LH;
    double* dm;            // the level's depth-map pyramid image
    int W, H;
  };
  std::vector<DmStep> dm_steps;
  // then the levels' computeHeuristicPb, several levels per launch
  std::vector<HpbBatch> hpb_groups;
  // the schedule captured as a HIP graph (params.graph)
  int eager_runs = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void drop_graph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    eager_runs = 0;
  }
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
  int relax_n = 0;  // relaxation launches inside the timed event pairs
  // loaded batch
  bool loaded = false, ran = false;
  double load_ratio = 0;
  int load_min_width = 0;
  std::vector<PairSlot> pairs;
  unsigned long long gen = 0;  // bumped by every load that may move buffers
  std::string dev_sig;         // shapes of the loaded device batch (batch_load_device), else empty
  BatchPlan plan;
  std::map<std::string, TileSched> plans;  // tile schedules by level shape and request
  unsigned long long plans_gen = 0;
};

namespace {

const char* kNoCtx = "null context";

// Scratch of rgbdflow_fill_depth_holes_noalloc: the fill's per-pixel state
// [N], its two lists of unfilled pixels [2N] and, at kNoAllocLeft, the count
// of pixels left unfilled, for frames of at most kNoAllocMaxPixels pixels
// (25 MB).  Each device holds its own copy until the process ends.  Host
// code never uses the array's name as an address: it uses the per-device
// address that hipGetSymbolAddress returns at context creation.
constexpr size_t kNoAllocMaxPixels = (size_t)1920 * 1088;
constexpr size_t kNoAllocLeft = 3 * kNoAllocMaxPixels;
__device__ __align__(256) int scratch[kNoAllocLeft + 64];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// Called with `device` current (rgbdflow_create); looks the address up once per device.
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess) g_scratch_ptr[device] = static_cast<int*>(p);
}

int* scratch_of(int device, hipError_t* err) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  *err = g_scratch_err[device];
  return g_scratch_ptr[device];
}

int fail(rgbdflow_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                         \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(ctx, RGBDFLOW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TRY(expr)                       \
  do {                                  \
    int rc_ = (expr);                   \
    if (rc_ != RGBDFLOW_OK) return rc_; \
  } while (0)

// Grow-only named device buffer, charged against the context's HBM budget.
// A buffer is freed only after the stream has drained (work in flight may
// still read it).
int ensure(rgbdflow_ctx* c, const std::string& name, size_t bytes, void** out) {
  Buffer& b = c->bufs[name];
  if (b.bytes >= bytes && b.p) {
    *out = b.p;
    return RGBDFLOW_OK;
  }
  if (b.p) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
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

// The job that makes level i (>= 1) of a Gaussian pyramid (GaussianPyramid.cpp:47-61):
// levels up to n_smooth smooth the finest image with sigma (1/ratio - 1) i and
// resize it by ratio^i; deeper levels re-smooth level i - n_smooth.  lvl[k]
// points at level k's planes.  Returns false if the filter is too wide.
bool pyramid_job(const std::vector<LevelShape>& L, int ns, double ratio, int C, const std::vector<double*>& lvl,
                 int i, PyrJob* j) {
  const double base = (1 / ratio - 1);
  const double sigma = (i <= ns) ? base * i : base * ns;
  const int fs = (int)(sigma * 3);
  if (2 * fs + 1 > kMaxTaps) return false;
  const std::vector<double> g = gaussian_taps(sigma, fs);
  std::memset(j, 0, sizeof(*j));
  for (int t = 0; t < 2 * fs + 1; t++) j->taps[t] = g[t];
  j->fs = fs;
  j->C = C;
  j->dst = lvl[i];
  j->Wd = L[i].W;
  j->Hd = L[i].H;
  if (i <= ns) {
    j->src = lvl[0];
    j->Ws = L[0].W;
    j->Hs = L[0].H;
    j->ratio = std::pow(ratio, i);
  } else {
    const LevelShape& s = L[i - ns];
    j->src = lvl[i - ns];
    j->Ws = s.W;
    j->Hs = s.H;
    j->ratio = (double)std::pow(ratio, i) * L[0].W / s.W;
  }
  return true;
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

int get_work(rgbdflow_ctx* c, const std::string& pre, size_t N, Work& w) {
  TRY(get(c, pre + "I1", kFeat * N, &w.I1));
  TRY(get(c, pre + "I2", kFeat * N, &w.I2));
  TRY(get(c, pre + "Iw", kFeat * N, &w.Iw));
  TRY(get(c, pre + "Dw", N, &w.Dw));
  TRY(get(c, pre + "u0", 3 * N, &w.u[0]));
  TRY(get(c, pre + "u1", 3 * N, &w.u[1]));
  TRY(get(c, pre + "Ix", kFeat * N, &w.Ix));
  TRY(get(c, pre + "Iy", kFeat * N, &w.Iy));
  TRY(get(c, pre + "It", kFeat * N, &w.It));
  TRY(get(c, pre + "Dx", N, &w.Dx));
  TRY(get(c, pre + "Dy", N, &w.Dy));
  TRY(get(c, pre + "Dt", N, &w.Dt));
  TRY(get(c, pre + "phi", N, &w.phi));
  // per-block partial sums: 1 per block for the outer statistic, 3 for the residual
  TRY(get(c, pre + "partials", 3 * (size_t)pix_blocks((int)N) + 16, &w.partials));
  TRY(get(c, pre + "du", N, &w.du));
  TRY(get(c, pre + "du_alt", N, &w.du_alt));
  TRY(get(c, pre + "dW", N, &w.dW));
  TRY(get(c, pre + "sW", N, &w.sW));
  TRY(get(c, pre + "mW", N, &w.mW));
  TRY(get(c, pre + "rW", 2 * N, &w.rW));
  TRY(get(c, pre + "cW", N, &w.cW));
  TRY(get(c, pre + "q", N, &w.q));
  TRY(get(c, pre + "coef", (size_t)kRec * N, &w.coef));
  TRY(get(c, pre + "ctl", kMaxLevels, &w.ctl));
  w.cg_coef = w.cg_du = w.cg_du_alt = nullptr;
  w.cg_ctl = nullptr;
  return RGBDFLOW_OK;
}

// Sizes of the coarse levels of a W x H level (at most `levels`).
std::vector<LevelShape> cg_shapes(int W, int H, int levels) {
  std::vector<LevelShape> L;
  while ((int)L.size() < levels && cg_coarse_ok(W, H)) {
    W = (W + 1) / 2;
    H = (H + 1) / 2;
    L.push_back({W, H});
  }
  return L;
}

// The coarse-grid correction's buffers of a work slot whose finest level is
// W x H (every level of the slot uses a prefix of them).
int get_cg_work(rgbdflow_ctx* c, const std::string& pre, int W, int H, Work& w) {
  size_t n = 0;
  for (const LevelShape& l : cg_shapes(W, H, kMaxCg)) n += (size_t)l.W * l.H;
  n += 1;
  TRY(get(c, pre + "cg_coef", (size_t)kRec * n, &w.cg_coef));
  TRY(get(c, pre + "cg_du", n, &w.cg_du));
  TRY(get(c, pre + "cg_du_alt", n, &w.cg_du_alt));
  TRY(get(c, pre + "cg_ctl", kMaxCg, &w.cg_ctl));
  return RGBDFLOW_OK;
}

// The solve fields of a level view over work buffers (level-entry fields zero).
LevelView make_view(const Work& w, int W, int H, const double* D1, const double* D2, double* u, int level) {
  LevelView v;
  std::memset(&v, 0, sizeof(v));
  v.W = W;
  v.H = H;
  v.N = W * H;
  v.I1 = w.I1;
  v.I2 = w.I2;
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
  v.uv_scale = 1;
  v.storage = RGBDFLOW_STORE_F32;
  set_omega(v, .7f);
  return v;
}

int validate_iters(rgbdflow_ctx* c, int nOuter, int nInner, int nSOR) {
  if (nOuter < 0 || nInner < 0 || nSOR < 0) return fail(c, RGBDFLOW_ERR_ARG, "negative iteration count");
  if (nOuter > kMaxOuter) return fail(c, RGBDFLOW_ERR_ARG, "nOuterFPIterations above " + std::to_string(kMaxOuter));
  if ((long long)nInner * (nSOR / kCheckEvery + 1) > kMaxChecks)
    return fail(c, RGBDFLOW_ERR_ARG, "nInner * nSOR / 25 above " + std::to_string(kMaxChecks));
  return RGBDFLOW_OK;
}

void relax_event(rgbdflow_ctx* c) {
  if (!c->time_relax) return;
  if (c->relax_ev_used >= (int)c->relax_ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    c->relax_ev.push_back(e);
  }
  hipEventRecord(c->relax_ev[c->relax_ev_used++], c->stream);
}

// Whether every tile of width TW (and its halo of 2T at cut edges) of a
// W-wide level fits a streamed region row.
bool stream_width_ok(int TW, int W, int T) {
  if (TW >= W) return (W + 1) / 2 <= stream_pairs();
  const int h = 2 * T;
  for (int ox = 0; ox < W; ox += TW) {
    const int rx0 = std::max(0, ox - h) & ~1, rx1 = std::min(W, std::min(ox + TW, W) + h);
    if ((rx1 - rx0 + 1) / 2 > stream_pairs()) return false;
  }
  return true;
}

// Tile schedule of the tiled relaxation for a W x H level.  Explicit tile_*
// parameters win; otherwise a cost model picks the workgroup shape, the tile
// and the iterations per launch: a tile covering the whole level runs a whole
// check interval per launch; otherwise cost ~ launches x (launch overhead +
// waves x per-workgroup work), where the per-workgroup work grows with the
// region (tile + halo of 2 T) per thread.  cap == 0: nothing fits.
// A plan request for any non-streamed workgroup shape.
constexpr int kAnyTiledConfig = -2;

TileSched plan_tiles(const rgbdflow_ctx* c, int W, int H, int req_T, int req_w, int req_h, int req_cfg) {
  TileSched best{};
  double best_cost = 1e300;
  const int n_cfg = tiled_config_count();
  for (int cfg = 0; cfg < n_cfg; cfg++) {
    if (req_cfg >= 0 && cfg != req_cfg) continue;
    int nt, p;
    tiled_config_shape(cfg, &nt, &p);
    const int tr = tiled_config_stream(cfg);
    if (req_cfg == kAnyTiledConfig && tr > 0) continue;
    if (tr > 0) {
      // streamed: a tile's height is free; its width plus the halo must fit
      // a region row.  Cost ~ launches x (launch overhead + waves of tiles x
      // steps per tile x time per step).
      if (req_T > tr) continue;
      const int T = req_T > 0 ? req_T : tr;
      const int h = 2 * T;
      const double step_us = 0.5 * ((tr * 2 + 1 + 3) / 4 * 0.142 + 0.05);
      const int tw0 = req_w > 0 ? req_w + (req_w & 1) : 8, tw1 = req_w > 0 ? tw0 : W + 7;
      for (int TW = tw0; TW <= tw1; TW += 2) {
        if (req_w <= 0 && (TW & 7) && TW < W) continue;
        if (!stream_width_ok(TW, W, T)) continue;
        const int th0 = req_h > 0 ? req_h : 4, th1 = req_h > 0 ? req_h : H + 3;
        for (int TH = th0; TH <= th1; TH += 4) {
          const int rh = std::min(TH + 2 * h, H);
          const int tiles = ((W + TW - 1) / TW) * ((H + TH - 1) / TH);
          const double waves = std::ceil((double)tiles / c->cu_count);
          const double per_wg = (rh + 4 * T - 1) * step_us;
          const double launches = std::ceil((double)kCheckEvery / T) * 2;
          const double cost = launches * (4.0 + waves * per_wg);
          if (cost < best_cost) {
            best_cost = cost;
            best = TileSched{TW, TH, (W + TW - 1) / TW, (H + TH - 1) / TH, 0, stream_pairs(), cfg, T};
          }
        }
      }
      continue;
    }
    const int cap = nt * p;
    const long long pairs_img = (long long)((W + 1) / 2) * H;
    if (pairs_img <= cap && req_w == 0) {
      if (tiled_lds_bytes(tiled_slots(TileSched{W + (W & 1), H, 1, 1, 1, cap, cfg, kCheckEvery}, W, H, 0)) >
          c->lds_per_block)
        continue;
      const double per_pass = std::ceil((double)pairs_img / nt) * 0.06 + 0.25;
      const double cost = 2 * (4.0 + std::ceil((double)pairs_img / nt) * 0.6 + 50 * per_pass);
      if (cost < best_cost) {
        best_cost = cost;
        best = TileSched{W + (W & 1), H, 1, 1, 1, cap, cfg, kCheckEvery};
      }
      continue;
    }
    const int Ts[] = {1, 2, 3, 4, 5, 6, 8, 12};
    for (int T : Ts) {
      if (req_T > 0 && T != req_T) continue;
      if (req_T <= 0 && (T == 1 || T == 12)) continue;  // the cost model's range
      const int h = 2 * T;
      for (int TW = 8; TW <= W + 7; TW += 8) {
        if (req_w > 0 && TW != req_w + (req_w & 1)) continue;
        for (int TH = 2; TH <= H + 3; TH += 2) {
          if (req_h > 0 && TH != req_h) continue;
          if (req_h <= 0 && (TH & 3)) continue;
          const int rw = std::min(TW + 2 * h, W + (W & 1)), rh = std::min(TH + 2 * h, H);
          const long long rp = (long long)((rw + 1) / 2) * rh;
          if (rp > cap) continue;
          const TileSched cand{TW, TH, (W + TW - 1) / TW, (H + TH - 1) / TH, 0, cap, cfg, T};
          const size_t lds = tiled_lds_bytes(tiled_slots(cand, W, H, T));
          if (lds > c->lds_per_block) continue;
          const int occ = std::max(1, std::min((int)((160 * 1024) / lds), 2048 / nt));
          const int tiles = ((W + TW - 1) / TW) * ((H + TH - 1) / TH);
          const double waves = std::ceil((double)tiles / (c->cu_count * occ));
          const double per_thread = std::ceil((double)rp / nt);
          const double per_wg = per_thread * 0.6 + 2 * T * (per_thread * 0.06 + 0.25);
          const double launches = std::ceil((double)kCheckEvery / T) * 2;
          const double cost = launches * (4.0 + waves * per_wg * occ);
          if (cost < best_cost) {
            best_cost = cost;
            best = cand;
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

// The key of a schedule the autotuner timed with one record storage.
std::string storage_key(const std::string& key, int storage) { return key + "/s" + std::to_string(storage); }

// Cached plan_tiles.  For a request without tile sizes (cfg -1 or
// kAnyTiledConfig), a schedule tuned with the given record storage comes
// first; else the one last tuned or installed for the shape.
TileSched cached_plan(rgbdflow_ctx* c, int W, int H, int T, int tw, int th, int cfg, int storage = -1) {
  const std::string key = plan_key(W, H, T, tw, th, cfg);
  if (storage >= 0 && T == 0 && tw == 0 && th == 0 && (cfg == -1 || cfg == kAnyTiledConfig)) {
    auto st = c->plans.find(storage_key(key, storage));
    if (st != c->plans.end()) return st->second;
  }
  auto it = c->plans.find(key);
  if (it == c->plans.end()) it = c->plans.emplace(key, plan_tiles(c, W, H, T, tw, th, cfg)).first;
  return it->second;
}

// Relaxation schedule of one step.  The tiled kernel takes one workgroup
// shape and one iteration count per launch, so the largest item's schedule
// (the caller's tile_* request, else the tuned or modelled one) fixes them
// and the other items are planned under them.  Falls back to the plain
// schedule when an item has no tiled schedule.
void plan_step_relax(rgbdflow_ctx* c, const rgbdflow_params* p, std::vector<LevelView>& items, StepPlan& st) {
  st.tiled_ok = false;
  if (!p || p->relax_variant != RGBDFLOW_RELAX_TILED || items.empty()) return;
  size_t lead = 0;
  for (size_t i = 1; i < items.size(); i++)
    if (items[i].N > items[lead].N) lead = i;
  const int LW = items[lead].W, LH = items[lead].H;
  const int storage = items[lead].storage;
  TileSched lt = cached_plan(c, LW, LH, p->tile_iters, p->tile_w, p->tile_h, p->tile_config, storage);
  // items of several sizes: tiled schedules only (a streamed workgroup's
  // time is its tiles' rows plus a ring fill and drain, which smaller items'
  // shorter tiles do not amortise)
  bool uniform = true;
  for (const LevelView& v : items) uniform = uniform && v.W == LW && v.H == LH;
  if (!uniform && lt.cap != 0 && tiled_config_stream(lt.config) > 0)
    lt = cached_plan(c, LW, LH, p->tile_iters, p->tile_w, p->tile_h, kAnyTiledConfig, storage);
  if (lt.cap == 0) return;
  std::vector<TileSched> ts(items.size());
  for (size_t i = 0; i < items.size(); i++) {
    if (items[i].W == LW && items[i].H == LH) {
      ts[i] = lt;
      continue;
    }
    // a smaller item: a whole-level tile under the lead's workgroup shape,
    // else tiles of the lead's iteration count
    const TileSched t = cached_plan(c, items[i].W, items[i].H, lt.single ? 0 : lt.tmax, 0, 0, lt.config);
    if (t.cap == 0 || (!t.single && (lt.single || t.tmax != lt.tmax))) return;
    ts[i] = t;
  }
  st.slots = 0;
  for (size_t i = 0; i < items.size(); i++) {
    items[i].ts = ts[i];
    st.slots = std::max(st.slots, tiled_slots(ts[i], items[i].W, items[i].H, lt.tmax));
  }
  if (tiled_lds_bytes(st.slots) > c->lds_per_block) return;
  st.tiled_ok = true;
  st.cfg = lt.config;
  st.T = lt.tmax;
}

// The steps of one level's items: each item size that kStepSplitItems items
// or more have (and that has a tiled schedule) gets a step of its own, so it
// relaxes under its own schedule -- a step of several sizes plans its items
// under its largest item's workgroup shape and iteration count, and tiled
// schedules only (a streamed launch has one tile shape).  The sizes of one
// item each share one step, so a served batch at light load (a pair or so of
// each size) keeps one launch per kernel and its latency.  Items keep their
// order within a step.
constexpr int kStepSplitItems = 2;
std::vector<std::vector<size_t>> level_steps(rgbdflow_ctx* c, const rgbdflow_params* p,
                                             const std::vector<LevelView>& items) {
  std::vector<std::vector<size_t>> out;
  std::vector<std::pair<int, int>> shapes;  // distinct sizes, first-seen order
  for (const LevelView& v : items)
    if (std::find(shapes.begin(), shapes.end(), std::make_pair(v.W, v.H)) == shapes.end())
      shapes.emplace_back(v.W, v.H);
  std::vector<size_t> rest;
  if (p && p->relax_variant == RGBDFLOW_RELAX_TILED && shapes.size() > 1) {
    std::vector<std::pair<int, int>> own;
    for (const auto& sh : shapes) {
      const TileSched t =
          cached_plan(c, sh.first, sh.second, p->tile_iters, p->tile_w, p->tile_h, p->tile_config, p->storage);
      if (t.cap == 0) continue;
      long long count = 0;
      for (const LevelView& v : items) count += v.W == sh.first && v.H == sh.second;
      if (count >= kStepSplitItems) own.push_back(sh);
    }
    for (const auto& sh : own) {
      std::vector<size_t> g;
      for (size_t i = 0; i < items.size(); i++)
        if (items[i].W == sh.first && items[i].H == sh.second) g.push_back(i);
      out.push_back(g);
    }
    for (size_t i = 0; i < items.size(); i++)
      if (std::find(own.begin(), own.end(), std::make_pair(items[i].W, items[i].H)) == own.end()) rest.push_back(i);
  } else {
    for (size_t i = 0; i < items.size(); i++) rest.push_back(i);
  }
  if (!rest.empty()) out.push_back(rest);
  return out;
}

size_t pack(std::vector<char>& host, const void* p, size_t bytes) {
  const size_t off = (host.size() + 255) & ~size_t(255);
  host.resize(off + bytes);
  std::memcpy(host.data() + off, p, bytes);
  return off;
}

// Packs a step's items and the block offsets of each kernel family.
void pack_step(std::vector<char>& host, const std::vector<LevelView>& items, StepPlan& st) {
  st.n = (int)items.size();
  st.storage = items.empty() ? 0 : items[0].storage;
  st.img = items.empty() ? 0 : items[0].img;
  bool fixed = true, gs = !items.empty(), block = !items.empty();
  for (const LevelView& v : items) {
    fixed = fixed && v.omega == .7f && v.pixel_gs == 0;
    gs = gs && v.pixel_gs == 1;
    block = block && v.pixel_gs == 2;
  }
  st.relax_mode = fixed ? kRelaxFixed : gs ? kRelaxGS : block ? kRelaxBlock : kRelaxFree;
  std::vector<int> pix(st.n + 1), plain(st.n + 1), tiled(st.n + 1), otile(st.n + 1);
  pix[0] = plain[0] = tiled[0] = otile[0] = 0;
  for (int i = 0; i < st.n; i++) {
    otile[i + 1] = otile[i] + otile_blocks(items[i].W, items[i].H);
    pix[i + 1] = pix[i] + pix_blocks(items[i].N);
    plain[i + 1] = plain[i] + plain_blocks(items[i].W, items[i].H);
    tiled[i + 1] = tiled[i] + (st.tiled_ok ? tiled_blocks(items[i].ts) : 0);
  }
  st.pix = pix[st.n];
  st.plain = plain[st.n];
  st.tiled = tiled[st.n];
  st.otile = otile[st.n];
  st.grid = st.tiled_ok && st.tiled > 0 ? tiled_prepare(st.cfg, st.storage, st.relax_mode, st.slots, st.tiled) : 0;
  st.any_early = false;
  for (const LevelView& v : items) st.any_early = st.any_early || v.allow_early;
  st.items_off = pack(host, items.data(), items.size() * sizeof(LevelView));
  st.pix_off = pack(host, pix.data(), pix.size() * sizeof(int));
  st.plain_off = pack(host, plain.data(), plain.size() * sizeof(int));
  st.tiled_off = pack(host, tiled.data(), tiled.size() * sizeof(int));
  st.otile_off = pack(host, otile.data(), otile.size() * sizeof(int));
}

// The coarse levels of a step for the coarse-grid correction (p's
// coarse_levels; none when 0): each item's levels over its work slot's
// coarse buffers (works[i]: item i's), planned and packed level by level.
void pack_coarse(rgbdflow_ctx* c, const rgbdflow_params* p, std::vector<char>& host,
                 const std::vector<LevelView>& items, const std::vector<const Work*>& works, StepPlan& st) {
  st.cg.clear();
  st.cg_nf.clear();
  st.cg_bottom_n = 0;
  if (!p || p->coarse_levels <= 0) return;
  for (const Work* w : works)
    if (!w->cg_coef) return;  // no buffers: the caller allocates them before planning
  std::vector<std::vector<LevelView>> lv, nf;
  std::vector<LevelView> bottom;
  for (size_t i = 0; i < items.size(); i++) {
    const LevelView& fv = items[i];
    const Work& w = *works[i];
    LevelView::Finer fin{fv.W, fv.H, fv.storage, fv.coef, fv.du, fv.du_alt, fv.ctl, fv.ctl, 1, 0, 0};
    size_t off = 0;
    const std::vector<LevelShape> L = cg_shapes(fv.W, fv.H, p->coarse_levels);
    // the levels from `fused` down fit in one workgroup's LDS
    size_t fused = L.size(), tail = 0;
    while (fused > 0 && tail + (size_t)L[fused - 1].W * L[fused - 1].H <= (size_t)kCgFusePx) {
      fused--;
      tail += (size_t)L[fused].W * L[fused].H;
    }
    for (size_t l = 0; l < L.size(); l++) {
      LevelView v;
      std::memset(&v, 0, sizeof(v));
      v.W = L[l].W;
      v.H = L[l].H;
      v.N = v.W * v.H;
      v.coef = w.cg_coef + (size_t)kRec * off;
      v.du = w.cg_du + off;
      v.du_alt = w.cg_du_alt + off;
      v.ctl = w.cg_ctl + l;
      v.partials = fv.partials;
      v.storage = RGBDFLOW_STORE_F32;
      v.img = fv.img;
      v.omega = fv.omega;
      v.omega_c = fv.omega_c;
      v.pixel_gs = fv.pixel_gs;
      v.finer = fin;
      fin = LevelView::Finer{v.W, v.H, RGBDFLOW_STORE_F32, v.coef, v.du, v.du_alt, v.ctl, fv.ctl, 0, (int)l + 1, 0};
      if (lv.size() <= l) lv.resize(l + 1);
      lv[l].push_back(v);
      if (l < fused) {
        if (nf.size() <= l) nf.resize(l + 1);
        nf[l].push_back(v);
      } else if (l == fused) {
        LevelView b = v;
        b.finer.fused = (int)(L.size() - fused);
        bottom.push_back(b);
      }
      off += (size_t)v.N;
    }
  }
  // the coarse relaxation: the step's variant, launches of coarse_sweeps iterations
  rgbdflow_params q = *p;
  q.tile_iters = p->coarse_sweeps;
  q.tile_w = q.tile_h = 0;
  q.tile_config = -1;
  for (std::vector<LevelView>& li : lv) {
    st.cg.emplace_back();
    pack_step(host, li, st.cg.back());
  }
  for (std::vector<LevelView>& li : nf) {
    st.cg_nf.emplace_back();
    plan_step_relax(c, &q, li, st.cg_nf.back());
    pack_step(host, li, st.cg_nf.back());
  }
  st.cg_bottom_n = (int)bottom.size();
  if (!bottom.empty()) st.cg_bottom_off = pack(host, bottom.data(), bottom.size() * sizeof(LevelView));
}

void bind_step(char* dev, StepPlan& st) {
  const LevelView* items = reinterpret_cast<const LevelView*>(dev + st.items_off);
  st.vpix = Views{items, reinterpret_cast<const int*>(dev + st.pix_off), st.n};
  st.vplain = Views{items, reinterpret_cast<const int*>(dev + st.plain_off), st.n};
  st.vtiled = Views{items, reinterpret_cast<const int*>(dev + st.tiled_off), st.n};
  st.votile = Views{items, reinterpret_cast<const int*>(dev + st.otile_off), st.n};
  for (StepPlan& cs : st.cg) bind_step(dev, cs);
  for (StepPlan& cs : st.cg_nf) bind_step(dev, cs);
  st.vbottom = Views{reinterpret_cast<const LevelView*>(dev + st.cg_bottom_off), nullptr, st.cg_bottom_n};
}

int upload(rgbdflow_ctx* c, void* dst, const void* src, size_t bytes) {
  HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return RGBDFLOW_OK;
}
int download(rgbdflow_ctx* c, void* dst, const void* src, size_t bytes) {
  HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return RGBDFLOW_OK;
}

int sync(rgbdflow_ctx* c) {
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipGetLastError());
  return RGBDFLOW_OK;
}

// Uploads packed descriptors to the named device buffer once the stream has
// drained (work in flight may read the buffer).
int upload_desc(rgbdflow_ctx* c, const char* name, const std::vector<char>& host, char** dev) {
  TRY(sync(c));
  TRY(get(c, name, host.size(), dev));
  HIP_TRY(c, hipMemcpy(*dev, host.data(), host.size(), hipMemcpyHostToDevice));
  return RGBDFLOW_OK;
}

// A one-step plan over the given items (the single-level entry points and the
// autotuner).  keep_ts: the items' tile schedules and st's launch shape are set.
int one_step(rgbdflow_ctx* c, const char* name, std::vector<LevelView>& items, const rgbdflow_params* sched,
             StepPlan& st, bool keep_ts = false, const Work* wk = nullptr) {
  if (!keep_ts) plan_step_relax(c, sched, items, st);
  std::vector<char> host;
  pack_step(host, items, st);
  if (wk && items.size() == 1) pack_coarse(c, sched, host, items, std::vector<const Work*>{wk}, st);
  char* dev;
  TRY(upload_desc(c, name, host, &dev));
  bind_step(dev, st);
  return RGBDFLOW_OK;
}

// nSOR red-black iterations of one (outer, inner), with the convergence
// checks of .cu:905-922 at iterations 24, 49, ...; the tiled schedule splits
// the iterations into launches that end at every check iteration.  With
// RGBDFLOW_EXIT_RESIDUAL the check is the FP64 relative residual, evaluated
// after the launch that ends at the check iteration (the relaxation launches
// then record no change maxima).
// copy_back: leave the result in du (a tiled schedule's last launch may have
// written du_alt; SolveCtl::du_in_alt says which) -- not needed when the next
// reader is k_update_warp, which reads whichever holds it.
// n red-black iterations of a coarse level from the unknowns in du_alt (alt)
// or du; returns where they end.  No exits: the relaxed level's stop
// (SolveCtl::outer_done of the coarse level, set by k_cg_restrict) skips them.
int coarse_sweeps(rgbdflow_ctx* c, const StepPlan& cs, int n, int alt) {
  hipStream_t s = c->stream;
  if (cs.tiled_ok) {
    for (int it = 0; it < n;) {
      const int T = std::min(cs.T, n - it);
      launch_relax_tiled(s, cs.vtiled, cs.tiled, cs.grid, cs.cfg, cs.storage, cs.relax_mode, cs.slots, alt, T, 0, 0,
                         -1);
      alt ^= 1;
      it += T;
    }
    return alt;
  }
  for (int it = 0; it < n; it++) {
    launch_relax_plain(s, cs.vplain, cs.plain, 0, 0, 0, -1);
    launch_relax_plain(s, cs.vplain, cs.plain, 1, 0, 0, -1);
  }
  return 0;
}

// One V-cycle of the coarse-grid correction on the relaxed level's unknowns
// (in du_alt if alt): down, the coarse rhs from the finer level's residual
// and the coarse level's sweeps; the levels that fit in LDS in one
// k_cg_bottom launch; up, the sweeps again and the correction added to the
// finer level.  n_prev: checks before it (the stop test).
void enqueue_vcycle(rgbdflow_ctx* c, const StepPlan& st, const SolveScalars& k, int slot_base, int n_prev, int alt) {
  hipStream_t s = c->stream;
  const int L = (int)st.cg_nf.size();
  int where[kMaxCg + 1];
  where[0] = alt;
  int mask = alt;
  for (int l = 1; l <= L; l++) {
    const StepPlan& cs = st.cg_nf[l - 1];
    launch_cg_restrict(s, cs.vpix, cs.pix, where[l - 1], slot_base, n_prev);
    where[l] = coarse_sweeps(c, cs, k.cg_sweeps, 0);
    mask |= where[l] << l;
  }
  launch_cg_bottom(s, st.vbottom, mask, slot_base, n_prev, k.cg_sweeps);
  for (int l = L; l >= 1; l--) {
    const StepPlan& cs = st.cg_nf[l - 1];
    where[l] = coarse_sweeps(c, cs, k.cg_sweeps, where[l]);
    launch_cg_prolong(s, cs.vpix, cs.pix, where[l], where[l - 1]);
  }
}

void enqueue_relax(rgbdflow_ctx* c, const StepPlan& st, const SolveScalars& k, int slot_base, bool copy_back = true) {
  hipStream_t s = c->stream;
  const bool resid = k.inner_exit == RGBDFLOW_EXIT_RESIDUAL;
  auto check_of = [&](int last) { return (last % kCheckEvery == kCheckEvery - 1) ? last / kCheckEvery : -1; };
  // coarse-grid correction: a V-cycle after iteration cg_at of every kCheckEvery
  const bool cg = !st.cg.empty() && k.n_sor > k.cg_at;
  if (cg)
    for (const StepPlan& cs : st.cg) launch_cg_operator(s, cs.vpix, cs.pix);
  if (st.tiled_ok) {
    // timed: one event pair around the launches of this (outer, inner), which
    // run back to back (around each launch where residual checks or V-cycles
    // come between)
    const bool each = resid || cg;
    if (!each && k.n_sor > 0) relax_event(c);
    int it = 0, flip = 0;
    while (it < k.n_sor) {
      const int boundary = (it / kCheckEvery + 1) * kCheckEvery;
      int T = std::min(std::min(st.T, boundary - it), k.n_sor - it);
      const int cg_it = it - it % kCheckEvery + k.cg_at;
      if (cg && cg_it >= it) T = std::min(T, cg_it - it + 1);
      const int last = it + T - 1;
      const int n_prev = k.inner_exit ? it / kCheckEvery : 0;
      const int chk = check_of(last);
      const int slot = chk >= 0 ? slot_base + chk : -1;
      if (each) relax_event(c);
      launch_relax_tiled(s, st.vtiled, st.tiled, st.grid, st.cfg, st.storage, st.relax_mode, st.slots, flip, T,
                         slot_base, n_prev, resid ? -1 : slot);
      if (each) relax_event(c);
      c->relax_n += c->time_relax ? 1 : 0;
      // the launch read du (flip 0) or du_alt (flip 1) and wrote the other
      if (resid && chk >= 0) launch_residual(s, st.vpix, st.pix, flip ? 0 : 1, slot_base, chk, slot, k.residual_tol);
      flip ^= 1;
      it += T;
      if (cg && last == cg_it) enqueue_vcycle(c, st, k, slot_base, k.inner_exit ? it / kCheckEvery : 0, flip);
    }
    if (!each && k.n_sor > 0) relax_event(c);
    if (k.n_sor > 0 && copy_back) launch_copy_du(s, st.vpix, st.pix);
    return;
  }
  for (int it = 0; it < k.n_sor; it++) {
    const int n_prev = k.inner_exit ? it / kCheckEvery : 0;
    const int chk = check_of(it);
    const int slot = chk >= 0 ? slot_base + chk : -1;
    relax_event(c);
    launch_relax_plain(s, st.vplain, st.plain, 0, slot_base, n_prev, resid ? -1 : slot);
    relax_event(c);
    relax_event(c);
    launch_relax_plain(s, st.vplain, st.plain, 1, slot_base, n_prev, resid ? -1 : slot);
    relax_event(c);
    c->relax_n += c->time_relax ? 2 : 0;
    if (resid && chk >= 0) launch_residual(s, st.vpix, st.pix, 0, slot_base, chk, slot, k.residual_tol);
    if (cg && it % kCheckEvery == k.cg_at) enqueue_vcycle(c, st, k, slot_base, k.inner_exit ? (it + 1) / kCheckEvery : 0, 0);
  }
}

// sceneFlowSingleScale's loops (.cu:755-954) for every item of a step.
void enqueue_level_solve(rgbdflow_ctx* c, const StepPlan& st, const SolveScalars& k) {
  hipStream_t s = c->stream;
  const int checks = k.n_sor / kCheckEvery;
  launch_ctl_reset(s, st.vpix);
  for (int o = 0; o < k.n_outer; o++) {
    if (k.n_inner == 0) launch_derivs(s, st.vpix, st.pix, 0, st.img);  // du = 0 (.cu:776), then the re-warp
    for (int h = 0; h < k.n_inner; h++) {
      if (h == 0) {
        launch_outer_first(s, st.votile, st.otile, k.n_inner * checks, k, k.n_inner > 1 ? 1 : 0, st.img);
      } else {
        launch_phi(s, st.vpix, st.pix, st.img);
        launch_system(s, st.vpix, st.pix, k, st.img);
      }
      enqueue_relax(c, st, k, h * checks, h + 1 < k.n_inner);
    }
    launch_update_warp(s, st.vpix, st.pix, st.any_early ? 1 : 0, st.img);
    if (st.any_early) launch_stat_final(s, st.vpix, o, k.outer_exit);
  }
}

void ctl_to_stats(const SolveCtl& ctl, int n_outer, bool allow_early, int& outer_exec, int& sweeps,
                  double& last_stat, double& margin) {
  outer_exec = ctl.outer_done ? ctl.outer_exit_at + 1 : n_outer;
  sweeps = ctl.sweeps_exec;
  last_stat = (allow_early && outer_exec > 0) ? ctl.stat[outer_exec - 1] : -1.0;
  // the statistic's closest approach to the early-exit threshold (.cu:948-951)
  margin = -1.0;
  if (allow_early)
    for (int o = 0; o < outer_exec; o++) {
      const double m = std::fabs(ctl.stat[o] - 4e-4) / 4e-4;
      if (margin < 0 || m < margin) margin = m;
    }
}

void ctl_residual_stats(const SolveCtl& ctl, rgbdflow_stats* st, int k) {
  st->residual_checks[k] = ctl.resid_checks;
  st->residual[k] = ctl.resid_checks ? ctl.resid : -1.0;
  st->energy[k] = ctl.resid_checks ? ctl.energy : 0.0;
}

SolveScalars scalars_from(const rgbdflow_params* p, double cw, double alpha, double gamma, int nOuter, int nInner,
                          int nSOR) {
  SolveScalars k;
  k.alpha = alpha;
  k.gamma = gamma;
  k.color_weight = cw;
  k.n_outer = nOuter;
  k.n_inner = nInner;
  k.n_sor = nSOR;
  k.inner_exit = p ? p->inner_exit : 1;
  k.outer_exit = p ? p->outer_exit : 1;
  k.residual_tol = (p && p->residual_tol > 0) ? p->residual_tol : 1e-4;
  k.cg_sweeps = p ? p->coarse_sweeps : 2;
  k.cg_at = p ? p->coarse_at : 9;
  return k;
}

// Interleaved host [N][C] f64 -> device planes [C][N].
int upload_planes(rgbdflow_ctx* c, const double* host, size_t N, int C, double* dst) {
  double* stage;
  TRY(get(c, "stage", N * C, &stage));
  TRY(upload(c, stage, host, N * C * sizeof(double)));
  launch_deinterleave(c->stream, stage, (int)N, C, dst);
  return RGBDFLOW_OK;
}
// Device planes [C][N] -> interleaved host [N][C] f64 (waits).
int download_planes(rgbdflow_ctx* c, const double* src, size_t N, int C, double* host) {
  double* stage;
  TRY(get(c, "stage_out", N * C, &stage));
  launch_interleave(c->stream, src, (int)N, C, stage);
  TRY(download(c, host, stage, N * C * sizeof(double)));
  return sync(c);
}

void collect_relax_timing(rgbdflow_ctx* c, rgbdflow_stats* st) {
  if (!c->time_relax || !st) return;
  double tot = 0;
  for (int i = 0; i + 1 < c->relax_ev_used; i += 2) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->relax_ev[i], c->relax_ev[i + 1]) == hipSuccess) tot += ms;
  }
  st->relax_ms = tot;
  st->relax_launches = c->relax_n;
}

bool bad_dims(int W, int H) { return W < 2 || H < 2 || (long long)W * H > (1LL << 28); }

// Level-solve inputs supplied by the host (sceneFlowSingleScale's arguments)
// in the single-level work set "d.": features, warp, depths, flow and the f64
// weights (converted to the solver's f32 copies by the level-weights kernel).
int stage_level_inputs(rgbdflow_ctx* c, int W, int H, float focal_level, const double* Im1, const double* dm1,
                       const double* Im2, const double* dm2, const double* warpIm2, const double* warpDm2,
                       const double* u, const double* v, const double* w, const double* dW64, const double* sW64,
                       const double* mW64, const double* r64, Work& wk, LevelView& view) {
  const size_t N = (size_t)W * H;
  TRY(get_work(c, "d.", N, wk));
  double *D1, *D2, *d, *s, *m, *r;
  TRY(get(c, "d.D1", N, &D1));
  TRY(get(c, "d.D2", N, &D2));
  TRY(get(c, "d.W0", N, &d));
  TRY(get(c, "d.W1", N, &s));
  TRY(get(c, "d.W2", N, &m));
  TRY(get(c, "d.W3", 2 * N, &r));
  TRY(upload_planes(c, Im1, N, kFeat, wk.I1));
  TRY(upload_planes(c, Im2, N, kFeat, wk.I2));
  TRY(upload_planes(c, warpIm2, N, kFeat, wk.Iw));
  TRY(upload(c, D1, dm1, N * 8));
  TRY(upload(c, D2, dm2, N * 8));
  TRY(upload(c, wk.Dw, warpDm2, N * 8));
  TRY(upload(c, wk.u[0], u, N * 8));
  TRY(upload(c, wk.u[0] + N, v, N * 8));
  TRY(upload(c, wk.u[0] + 2 * N, w, N * 8));
  TRY(upload(c, d, dW64, N * 8));
  TRY(upload(c, s, sW64, N * 8));
  TRY(upload(c, m, mW64, N * 8));
  TRY(upload(c, r, r64, 2 * N * 8));
  view = make_view(wk, W, H, D1, D2, wk.u[0], 0);
  view.delk = d;
  view.smk = s;
  view.mag64 = m;
  view.reg64 = r;
  view.f_level = focal_level;
  return RGBDFLOW_OK;
}

// ------------------------------------------------------------------ batches
// Shape and buffers of pair slot idx (W x H, camera, pyramid min width).
int alloc_pair(rgbdflow_ctx* c, const rgbdflow_params* p, int idx, int W, int H, const rgbdflow_camera& cam,
               int min_width, bool hpb, bool init, PairSlot& ps) {
  if (bad_dims(W, H)) return fail(c, RGBDFLOW_ERR_ARG, "bad size of pair " + std::to_string(idx));
  ps.W = W;
  ps.H = H;
  ps.cam_given = cam.xRes != 0;
  ps.cam = cam;
  ps.min_width = min_width > 0 ? min_width : p->min_width;
  ps.hpb_given = hpb;
  ps.pyr1_given = false;
  ps.L = pyramid_shape(W, H, p->pyramid_ratio, ps.min_width, nullptr, nullptr);
  // sceneFlowWithInit.cpp:7-81 refines at full resolution only
  ps.init = init;
  if (init) ps.L.resize(1);
  const int nL = (int)ps.L.size();
  if (nL > kMaxLevels) return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "too many pyramid levels");
  ps.off.assign(nL, 0);
  for (int k = 2; k < nL; k++) ps.off[k] = ps.off[k - 1] + (size_t)ps.L[k - 1].W * ps.L[k - 1].H;
  ps.hoff.assign(nL, 0);
  for (int k = 1; k < nL; k++) ps.hoff[k] = ps.hoff[k - 1] + (size_t)ps.L[k - 1].W * ps.L[k - 1].H * 2;
  const size_t N = (size_t)W * H, rest = pyramid_elems(ps.L, 1) + 1;
  const std::string pre = "b" + std::to_string(idx) + ".";
  TRY(get(c, pre + "im1", 3 * N, &ps.im1));
  TRY(get(c, pre + "im2", 3 * N, &ps.im2));
  TRY(get(c, pre + "dm1", N, &ps.dm1));
  TRY(get(c, pre + "dm2", N, &ps.dm2));
  TRY(get(c, pre + "del0", N, &ps.del0));
  TRY(get(c, pre + "sm0", N, &ps.sm0));
  TRY(get(c, pre + "sig", N, &ps.sigma));
  TRY(get(c, pre + "pim1", 3 * rest, &ps.pim1));
  TRY(get(c, pre + "pim2", 3 * rest, &ps.pim2));
  TRY(get(c, pre + "pd1", rest, &ps.pd1));
  TRY(get(c, pre + "pd2", rest, &ps.pd2));
  TRY(get(c, pre + "pdel", rest, &ps.pdel));
  TRY(get(c, pre + "psm", rest, &ps.psm));
  ps.hpb = nullptr;
  ps.pdm = nullptr;
  ps.hpb_device = !hpb && p->regularization_type == RGBDFLOW_REG_HPB;
  if (hpb || ps.hpb_device) TRY(get(c, pre + "hpb", pyramid_elems(ps.L, 0) * 2, &ps.hpb));
  if (ps.hpb_device) {
    // DepthMapPyramid level sizes are (int)(W r^m) (depthMapPyramid.cpp:15-30); the
    // boundary maps pair them with the colour pyramid's levels, which must match
    double r = 0;
    pyramid_shape(W, H, p->pyramid_ratio, ps.min_width, nullptr, &r);
    for (size_t m = 0; m < ps.L.size(); m++) {
      const double lr = std::pow(r, (int)m);
      if ((int)(W * lr) != ps.L[m].W || (int)(H * lr) != ps.L[m].H)
        return fail(c, RGBDFLOW_ERR_UNSUPPORTED,
                    "boundary maps on the device need depth-map and colour pyramid levels of equal sizes");
    }
    TRY(get(c, pre + "pdm", pyramid_elems(ps.L, 1) + 1, &ps.pdm));
  }
  TRY(get_work(c, pre, N, ps.wk));
  if (p->coarse_levels > 0) TRY(get_cg_work(c, pre, W, H, ps.wk));
  return RGBDFLOW_OK;
}

int load_pair(rgbdflow_ctx* c, const rgbdflow_params* p, int idx, const rgbdflow_pair_input& in, PairSlot& ps) {
  if (!in.im1 || !in.dm1 || !in.im2 || !in.dm2)
    return fail(c, RGBDFLOW_ERR_ARG, "bad arguments of pair " + std::to_string(idx));
  TRY(alloc_pair(c, p, idx, in.W, in.H, in.cam, in.min_width, in.hpb != nullptr, in.est != nullptr, ps));
  ps.sigma_given = in.sigma1 != nullptr;
  const size_t N = (size_t)in.W * in.H;
  if (in.est) TRY(upload_planes(c, in.est, N, 3, ps.wk.u[1]));
  TRY(upload_planes(c, in.im1, N, 3, ps.im1));
  TRY(upload_planes(c, in.im2, N, 3, ps.im2));
  TRY(upload(c, ps.dm1, in.dm1, N * 8));
  TRY(upload(c, ps.dm2, in.dm2, N * 8));
  if (in.sigma1) TRY(upload(c, ps.sigma, in.sigma1, N * 4));
  if (in.hpb) TRY(upload(c, ps.hpb, in.hpb, pyramid_elems(ps.L, 0) * 2 * 4));
  ps.dmh = nullptr;
  if (in.hpb_dm1 && ps.hpb_device) {
    TRY(get(c, "b" + std::to_string(idx) + ".dmh", N, &ps.dmh));
    TRY(upload(c, ps.dmh, in.hpb_dm1, N * 8));
  }
  return sync(c);  // the caller may release its arrays after return
}

// Level k's planes of one of a pair's pyramids (level 0 = the base image).
double* level_of(const PairSlot& ps, double* base, double* rest, int C, int k) {
  if (k == 0) return base;
  return rest + ps.off[k] * C;
}

// Builds (or reuses) the descriptors of a batch run under params p.
int batch_plan(rgbdflow_ctx* c, const rgbdflow_params* p) {
  BatchPlan& bp = c->plan;
  const unsigned char* pb = reinterpret_cast<const unsigned char*>(p);
  if (bp.valid && bp.gen == c->gen && bp.plans_gen == c->plans_gen && bp.params.size() == sizeof(*p) &&
      std::memcmp(bp.params.data(), pb, sizeof(*p)) == 0)
    return RGBDFLOW_OK;
  bp.valid = false;
  TRY(sync(c));
  bp.drop_graph();
  bp.host.clear();
  bp.steps.clear();
  bp.waves.clear();
  bp.dm_steps.clear();
  bp.hpb_groups.clear();
  const int B = (int)c->pairs.size();
  // the coarse-grid correction's buffers (pairs loaded without it have none)
  if (p->coarse_levels > 0)
    for (int i = 0; i < B; i++) {
      PairSlot& ps = c->pairs[i];
      if (!ps.wk.cg_coef) TRY(get_cg_work(c, "b" + std::to_string(i) + ".", ps.W, ps.H, ps.wk));
    }
  // per-pair preparation
  {
    std::vector<PairPrep> items(B);
    std::vector<int> start(B + 1, 0);
    for (int i = 0; i < B; i++) {
      const PairSlot& ps = c->pairs[i];
      const rgbdflow_camera& cam = ps.cam_given ? ps.cam : p->cam;
      PairPrep& q = items[i];
      q.dm1 = ps.dm1;
      q.sigma = ps.sigma;
      q.del0 = ps.del0;
      q.sm0 = ps.sm0;
      q.W = ps.W;
      q.H = ps.H;
      q.compute_sigma = ps.sigma_given ? 0 : 1;
      q.dvc = p->depth_vs_color_data_weight;
      const double f0 = cam.focalLength;
      q.f2 = f0 * f0;
      start[i + 1] = start[i] + pair_prep_blocks(ps.W, ps.H);
    }
    bp.prep_items = pack(bp.host, items.data(), items.size() * sizeof(PairPrep));
    bp.prep_start = pack(bp.host, start.data(), start.size() * sizeof(int));
    bp.prep_n = B;
    bp.prep_blocks = start[B];
  }
  // the six pyramids of sceneFlowMultiscale.cpp:40-85 (the raw-depth pyramid
  // feeds only the boundary probabilities, which the caller supplies), in
  // waves: wave w makes levels w ns + 1 .. (w + 1) ns of every pyramid, which
  // read only level 0 or levels of earlier waves
  {
    int ns = 0;
    double ratio = 0;
    pyramid_shape(16, 16, p->pyramid_ratio, 1 << 30, &ns, &ratio);
    std::vector<std::vector<PyrJob>> waves;
    for (int i = 0; i < B; i++) {
      PairSlot& ps = c->pairs[i];
      const int nL = (int)ps.L.size();
      struct Img {
        double *base, *rest;
        int C;
      } imgs[6] = {{ps.im1, ps.pim1, 3}, {ps.im2, ps.pim2, 3},  {ps.dm1, ps.pd1, 1},
                   {ps.dm2, ps.pd2, 1},  {ps.del0, ps.pdel, 1}, {ps.sm0, ps.psm, 1}};
      for (const Img& im : imgs) {
        if (ps.pyr1_given && (im.base == ps.im1 || im.base == ps.dm1)) continue;  // loaded with frame 1
        std::vector<double*> lvl(nL);
        for (int k = 0; k < nL; k++) lvl[k] = level_of(ps, im.base, im.rest, im.C, k);
        for (int k = 1; k < nL; k++) {
          PyrJob j;
          if (!pyramid_job(ps.L, ns, ratio, im.C, lvl, k, &j))
            return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "pyramid filter wider than 31 taps");
          const int w = (k - 1) / ns;
          if ((int)waves.size() <= w) waves.resize(w + 1);
          waves[w].push_back(j);
        }
      }
    }
    for (const std::vector<PyrJob>& jobs : waves) {
      std::vector<int> start(jobs.size() + 1, 0);
      for (size_t j = 0; j < jobs.size(); j++) start[j + 1] = start[j] + pyramid_job_blocks(jobs[j]);
      BatchPlan::Wave wv;
      wv.items = pack(bp.host, jobs.data(), jobs.size() * sizeof(PyrJob));
      wv.start = pack(bp.host, start.data(), start.size() * sizeof(int));
      wv.n = (int)jobs.size();
      wv.blocks = start.back();
      wv.wl = WorkList<PyrJob>{};
      bp.waves.push_back(wv);
    }
  }
  // boundary maps of pairs loaded without them (sceneFlowMultiscale.cpp:116):
  // the depth-map pyramids, then the levels in groups of up to kHpbMaxItems
  // (and about kHpbGroupFloats of scratch), each level with its own scratch;
  // groups run one after the other and share the scratch
  {
    constexpr size_t kHpbGroupFloats = size_t(64) << 20;
    struct Lv {
      const PairSlot* ps;
      int k;
      size_t off;
    };
    std::vector<std::vector<Lv>> groups;
    size_t scratch = 0, used = 0;
    for (const PairSlot& ps : c->pairs) {
      if (!ps.hpb_device) continue;
      // level 0: dm1, or the depth given for the maps (the colour-filled raw depth)
      double* d0 = ps.dmh ? ps.dmh : ps.dm1;
      for (int k = 0; k < (int)ps.L.size(); k++) {
        if (k > 0)
          bp.dm_steps.push_back({level_of(ps, d0, ps.pdm, 1, k - 1), ps.L[k - 1].W, ps.L[k - 1].H,
                                 level_of(ps, d0, ps.pdm, 1, k), ps.L[k].W, ps.L[k].H});
        const size_t need = (hpb_scratch_floats(ps.L[k].W, ps.L[k].H) + 63) & ~size_t(63);
        if (groups.empty() || groups.back().size() == (size_t)kHpbMaxItems ||
            (used > 0 && used + need > kHpbGroupFloats)) {
          groups.emplace_back();
          used = 0;
        }
        groups.back().push_back({&ps, k, used});
        used += need;
        scratch = std::max(scratch, used);
      }
    }
    float* hs = nullptr;
    if (scratch) TRY(get(c, "hpb.scratch", scratch, &hs));
    for (const std::vector<Lv>& g : groups) {
      HpbBatch hb;
      hb.n = 0;
      for (const Lv& e : g) {
        const PairSlot& ps = *e.ps;
        hpb_batch_add(hb, hpb_level(ps.L[e.k].W, ps.L[e.k].H, level_of(ps, ps.im1, ps.pim1, 3, e.k),
                                    level_of(ps, ps.dmh ? ps.dmh : ps.dm1, ps.pdm, 1, e.k), ps.hpb + ps.hoff[e.k],
                                    hs + e.off));
      }
      bp.hpb_groups.push_back(hb);
    }
  }
  // the level steps (sceneFlowMultiscale.cpp:95-196), aligned on the finest level
  int S = 0;
  for (const PairSlot& ps : c->pairs) S = std::max(S, (int)ps.L.size());
  for (int s = 0; s < S; s++) {
    const int k = S - 1 - s;
    std::vector<LevelView> items;
    std::vector<const Work*> works;
    for (int i = 0; i < B; i++) {
      PairSlot& ps = c->pairs[i];
      const int nL = (int)ps.L.size();
      if (k >= nL) continue;
      works.push_back(&ps.wk);
      const rgbdflow_camera& cam = ps.cam_given ? ps.cam : p->cam;
      const int Wk = ps.L[k].W, Hk = ps.L[k].H;
      const int slot = (nL - 1 - k) & 1;  // flow buffers alternate level by level
      LevelView v = make_view(ps.wk, Wk, Hk, level_of(ps, ps.dm1, ps.pd1, 1, k), level_of(ps, ps.dm2, ps.pd2, 1, k),
                              ps.wk.u[slot], k);
      v.allow_early = k < nL - 1;
      v.P1k = level_of(ps, ps.im1, ps.pim1, 3, k);
      v.P2k = level_of(ps, ps.im2, ps.pim2, 3, k);
      v.delk = level_of(ps, ps.del0, ps.pdel, 1, k);
      v.smk = level_of(ps, ps.sm0, ps.psm, 1, k);
      v.hpbk = ps.hpb ? ps.hpb + ps.hoff[k] : nullptr;
      v.uv_scale = 1 / p->pyramid_ratio;
      if (k < nL - 1) {
        v.u_prev = ps.wk.u[1 - slot];
        v.Wp = ps.L[k + 1].W;
        v.Hp = ps.L[k + 1].H;
      } else if (ps.init) {
        // the estimate enters the level as a same-size flow (sceneFlowWithInit.cpp:13-21)
        v.u_prev = ps.wk.u[1];
        v.Wp = Wk;
        v.Hp = Hk;
        v.uv_scale = 1;
      }
      v.storage = p->storage;
      v.img = p->image_storage;
      set_omega(v, p->omega, p->pixel_update);
      v.reg_type = p->regularization_type;
      v.W_full = (int)cam.xRes;
      v.f_full = cam.focalLength;
      v.f_level = downsized_focal(cam, Wk);
      items.push_back(v);
    }
    for (const std::vector<size_t>& g : level_steps(c, p, items)) {
      std::vector<LevelView> gi;
      std::vector<const Work*> gw;
      for (size_t i : g) {
        gi.push_back(items[i]);
        gw.push_back(works[i]);
      }
      StepPlan st;
      plan_step_relax(c, p, gi, st);
      pack_step(bp.host, gi, st);
      pack_coarse(c, p, bp.host, gi, gw, st);
      bp.steps.push_back(std::move(st));
    }
  }
  char* dev;
  TRY(upload_desc(c, "desc", bp.host, &dev));
  for (StepPlan& st : bp.steps) bind_step(dev, st);
  for (BatchPlan::Wave& wv : bp.waves)
    wv.wl = WorkList<PyrJob>{reinterpret_cast<const PyrJob*>(dev + wv.items),
                             reinterpret_cast<const int*>(dev + wv.start), wv.n};
  bp.prep = WorkList<PairPrep>{reinterpret_cast<const PairPrep*>(dev + bp.prep_items),
                               reinterpret_cast<const int*>(dev + bp.prep_start), bp.prep_n};
  bp.params.assign(pb, pb + sizeof(*p));
  bp.gen = c->gen;
  bp.plans_gen = c->plans_gen;
  bp.valid = true;
  return RGBDFLOW_OK;
}

// The block update needs f32 records: the 3x3 blocks are negative definite
// as assembled, but their condition reaches ~1e6 where the depth term's
// weight is large, and rounded to 16 bits many are not definite any more.
int check_block_storage(rgbdflow_ctx* c, int pixel_update, int storage) {
  if (pixel_update == RGBDFLOW_PIXEL_BLOCK && storage != RGBDFLOW_STORE_F32)
    return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "pixel_update BLOCK needs f32 system records");
  return RGBDFLOW_OK;
}

int check_coarse(rgbdflow_ctx* c, const rgbdflow_params* p) {
  if (p->coarse_levels < 0 || p->coarse_levels > kMaxCg) return fail(c, RGBDFLOW_ERR_ARG, "bad coarse_levels");
  if (p->coarse_levels > 0 && (p->coarse_sweeps < 1 || p->coarse_sweeps > 64))
    return fail(c, RGBDFLOW_ERR_ARG, "bad coarse_sweeps");
  if (p->coarse_levels > 0 && (p->coarse_at < 0 || p->coarse_at >= kCheckEvery - 1))
    return fail(c, RGBDFLOW_ERR_ARG, "bad coarse_at");
  return RGBDFLOW_OK;
}

int check_run_params(rgbdflow_ctx* c, const rgbdflow_params* p) {
  if (!p) return fail(c, RGBDFLOW_ERR_ARG, "null params");
  if (p->variant != RGBDFLOW_VARIANT_FULL) return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "bad scene flow variant");
  if (p->regularization_type < 0 || p->regularization_type > 2) return fail(c, RGBDFLOW_ERR_ARG, "bad regularization");
  if (p->storage < RGBDFLOW_STORE_F32 || p->storage > RGBDFLOW_STORE_BF16)
    return fail(c, RGBDFLOW_ERR_ARG, "bad system storage");
  if (p->inner_exit < RGBDFLOW_EXIT_NONE || p->inner_exit > RGBDFLOW_EXIT_RESIDUAL)
    return fail(c, RGBDFLOW_ERR_ARG, "bad inner_exit");
  if (p->pixel_update < RGBDFLOW_PIXEL_JACOBI || p->pixel_update > RGBDFLOW_PIXEL_BLOCK)
    return fail(c, RGBDFLOW_ERR_ARG, "bad pixel_update");
  TRY(check_block_storage(c, p->pixel_update, p->storage));
  if (p->image_storage < RGBDFLOW_IMAGES_F64 || p->image_storage > RGBDFLOW_IMAGES_BF16)
    return fail(c, RGBDFLOW_ERR_ARG, "bad image_storage");
  TRY(check_coarse(c, p));
  return validate_iters(c, p->n_outer, p->n_inner, p->n_sor);
}

// A pair whose frames are already on the device (the serving path).
int load_pair_device(rgbdflow_ctx* c, const rgbdflow_params* p, int idx, const PairDev& in, PairSlot& ps) {
  if (!in.f1.im || !in.f1.dm || !in.f2.im || !in.f2.dm)
    return fail(c, RGBDFLOW_ERR_ARG, "bad device frames of pair " + std::to_string(idx));
  TRY(alloc_pair(c, p, idx, in.W, in.H, in.cam, in.min_width, in.f1.hpb != nullptr, in.init != nullptr, ps));
  ps.sigma_given = in.f1.sigma != nullptr;
  const size_t N = (size_t)in.W * in.H, rest = pyramid_elems(ps.L, 1);
  hipStream_t s = c->stream;
  if (in.f1.sigma) HIP_TRY(c, hipMemcpyAsync(ps.sigma, in.f1.sigma, N * 4, hipMemcpyDeviceToDevice, s));
  ps.dmh = nullptr;
  if (in.f1.hpb_dm && ps.hpb_device) {
    TRY(get(c, "b" + std::to_string(idx) + ".dmh", N, &ps.dmh));
    HIP_TRY(c, hipMemcpyAsync(ps.dmh, in.f1.hpb_dm, N * 8, hipMemcpyDeviceToDevice, s));
  }
  HIP_TRY(c, hipMemcpyAsync(ps.im1, in.f1.im, 3 * N * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(ps.dm1, in.f1.dm, N * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(ps.im2, in.f2.im, 3 * N * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(ps.dm2, in.f2.dm, N * 8, hipMemcpyDeviceToDevice, s));
  if (in.f1.hpb)
    HIP_TRY(c, hipMemcpyAsync(ps.hpb, in.f1.hpb, pyramid_elems(ps.L, 0) * 2 * 4, hipMemcpyDeviceToDevice, s));
  if (in.init) HIP_TRY(c, hipMemcpyAsync(ps.wk.u[1], in.init, 3 * N * 8, hipMemcpyDeviceToDevice, s));
  if (in.f1.pim && in.f1.pd && rest > 0) {
    HIP_TRY(c, hipMemcpyAsync(ps.pim1, in.f1.pim, 3 * rest * 8, hipMemcpyDeviceToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(ps.pd1, in.f1.pd, rest * 8, hipMemcpyDeviceToDevice, s));
    ps.pyr1_given = true;
  }
  return RGBDFLOW_OK;
}

// Every launch of a batch run, on the context's stream.
void enqueue_batch(rgbdflow_ctx* c, const BatchPlan& bp, const SolveScalars& k) {
  hipStream_t s = c->stream;
  // per-pair weights: depth stdevs (runFlowTwoFrames.cpp:89-91) -> delta
  // (OpticalFlow.cpp:71-81); s = f^2 (OpticalFlow.cpp:90-97); then the pyramids
  launch_pair_prep(s, bp.prep, bp.prep_blocks);
  for (const BatchPlan::Wave& wv : bp.waves) launch_pyramid_jobs(s, wv.wl, wv.blocks);
  // boundary maps made on the device, from the colour pyramid just built
  for (const BatchPlan::DmStep& d : bp.dm_steps) launch_depth_map_level(s, d.larger, d.LW, d.LH, d.dm, d.W, d.H);
  for (const HpbBatch& hb : bp.hpb_groups) launch_hpb_batch(s, hb);
  // coarse to fine
  for (const StepPlan& sp : bp.steps) {
    launch_features(s, sp.vpix, sp.pix, sp.img);
    launch_level_entry(s, sp.vpix, sp.pix, sp.img);
    launch_level_weights(s, sp.vpix, sp.pix);
    enqueue_level_solve(c, sp, k);
  }
}

int batch_load(rgbdflow_ctx* c, const rgbdflow_params* p, int n, const rgbdflow_pair_input* in) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!p || n < 1 || !in) return fail(c, RGBDFLOW_ERR_ARG, "bad batch arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  c->loaded = false;
  c->ran = false;
  c->gen++;
  c->dev_sig.clear();
  c->pairs.assign(n, PairSlot());
  for (int i = 0; i < n; i++) TRY(load_pair(c, p, i, in[i], c->pairs[i]));
  c->load_ratio = p->pyramid_ratio;
  c->load_min_width = p->min_width;
  c->loaded = true;
  return RGBDFLOW_OK;
}

int batch_run(rgbdflow_ctx* c, const rgbdflow_params* p, rgbdflow_stats* st) {
  if (!c) return RGBDFLOW_ERR_ARG;
  TRY(check_run_params(c, p));
  if (!c->loaded) return fail(c, RGBDFLOW_ERR_STATE, "run before load");
  if (p->pyramid_ratio != c->load_ratio || p->min_width != c->load_min_width)
    return fail(c, RGBDFLOW_ERR_STATE, "pyramid shape differs from the load's params");
  for (size_t i = 0; i < c->pairs.size(); i++) {
    const PairSlot& ps = c->pairs[i];
    const rgbdflow_camera& cam = ps.cam_given ? ps.cam : p->cam;
    if ((int)cam.xRes != ps.W || (int)cam.yRes != ps.H)
      return fail(c, RGBDFLOW_ERR_ARG, "camera resolution differs from the image size of pair " + std::to_string(i));
    if (p->regularization_type == RGBDFLOW_REG_HPB && !ps.hpb_given && !ps.hpb_device)
      return fail(c, RGBDFLOW_ERR_ARG, "HPB regularisation needs boundary maps: none loaded, none made (the pair "
                                       "was loaded under other params)");
  }
  HIP_TRY(c, hipSetDevice(c->device));
  TRY(batch_plan(c, p));
  const BatchPlan& bp = c->plan;
  hipStream_t s = c->stream;
  const SolveScalars k = scalars_from(p, p->color_data_weight, p->alpha, p->gamma, p->n_outer, p->n_inner, p->n_sor);
  c->relax_ev_used = 0;
  c->relax_n = 0;
  BatchPlan& bpm = c->plan;
  const bool graph = p->graph && !c->time_relax;
  if (graph && !bpm.exec && bpm.eager_runs > 0) {
    // capture the schedule (every launch configuration has run once, so no
    // launcher queries the device inside the capture)
    HIP_TRY(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue_batch(c, bp, k);
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamEndCapture(s, &g));
    bpm.graph = g;
    HIP_TRY(c, hipGraphInstantiate(&bpm.exec, g, nullptr, nullptr, 0));
  }
  HIP_TRY(c, hipEventRecord(c->ev0, s));
  if (graph && bpm.exec) {
    HIP_TRY(c, hipGraphLaunch(bpm.exec, s));
  } else {
    enqueue_batch(c, bp, k);
    bpm.eager_runs++;
  }
  HIP_TRY(c, hipEventRecord(c->ev1, s));
  HIP_TRY(c, hipGetLastError());
  c->ran = true;
  if (!st) return RGBDFLOW_OK;
  std::vector<std::vector<SolveCtl>> ctl(c->pairs.size());
  for (size_t i = 0; i < c->pairs.size(); i++) {
    ctl[i].resize(c->pairs[i].L.size());
    TRY(download(c, ctl[i].data(), c->pairs[i].wk.ctl, ctl[i].size() * sizeof(SolveCtl)));
  }
  TRY(sync(c));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  for (size_t i = 0; i < c->pairs.size(); i++) {
    const PairSlot& ps = c->pairs[i];
    rgbdflow_stats& o = st[i];
    std::memset(&o, 0, sizeof(o));
    const int nL = (int)ps.L.size();
    o.n_levels = nL;
    for (int l = 0; l < nL; l++) {
      o.width[l] = ps.L[l].W;
      o.height[l] = ps.L[l].H;
      ctl_to_stats(ctl[i][l], k.n_outer, l < nL - 1, o.outer_exec[l], o.sweeps_exec[l], o.last_stat[l],
                   o.stat_margin[l]);
      ctl_residual_stats(ctl[i][l], &o, l);
    }
    o.device_ms = ms;
    collect_relax_timing(c, &o);
  }
  return RGBDFLOW_OK;
}

int batch_read(rgbdflow_ctx* c, int i, double* u, double* v, double* w) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!c->ran) return fail(c, RGBDFLOW_ERR_STATE, "read before run");
  if (i < 0 || i >= (int)c->pairs.size()) return fail(c, RGBDFLOW_ERR_ARG, "pair index out of range");
  if (!u || !v || !w) return fail(c, RGBDFLOW_ERR_ARG, "null output");
  HIP_TRY(c, hipSetDevice(c->device));
  const PairSlot& ps = c->pairs[i];
  const size_t N = (size_t)ps.W * ps.H;
  const double* uu = ps.wk.u[((int)ps.L.size() - 1) & 1];
  TRY(download(c, u, uu, N * 8));
  TRY(download(c, v, uu + N, N * 8));
  TRY(download(c, w, uu + 2 * N, N * 8));
  return sync(c);
}

void fill_stats(rgbdflow_ctx* c, const SolveCtl& ctl, const LevelView& view, int n_outer, rgbdflow_stats* st) {
  if (!st) return;
  std::memset(st, 0, sizeof(*st));
  st->n_levels = 1;
  st->width[0] = view.W;
  st->height[0] = view.H;
  ctl_to_stats(ctl, n_outer, view.allow_early != 0, st->outer_exec[0], st->sweeps_exec[0], st->last_stat[0],
               st->stat_margin[0]);
  ctl_residual_stats(ctl, st, 0);
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  st->device_ms = ms;
  collect_relax_timing(c, st);
}

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
  p->storage = RGBDFLOW_STORE_F32;
  p->omega = .7f;
  p->coarse_levels = 0;
  p->coarse_sweeps = 2;
  p->coarse_at = 9;
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
  scratch_setup(device);
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
  c->plan.drop_graph();
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
  if (opts && (opts->inner_exit < RGBDFLOW_EXIT_NONE || opts->inner_exit > RGBDFLOW_EXIT_RESIDUAL))
    return fail(c, RGBDFLOW_ERR_ARG, "bad inner_exit");
  if (opts && (opts->pixel_update < RGBDFLOW_PIXEL_JACOBI || opts->pixel_update > RGBDFLOW_PIXEL_BLOCK))
    return fail(c, RGBDFLOW_ERR_ARG, "bad pixel_update");
  if (opts) TRY(check_block_storage(c, opts->pixel_update, opts->storage));
  if (opts && opts->image_storage != RGBDFLOW_IMAGES_F64)
    return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "image_storage: the single-level entry points take f64 images");
  if (opts) TRY(check_coarse(c, opts));
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  Work wk;
  LevelView view;
  TRY(stage_level_inputs(c, W, H, focal_level, Im1, dm1, Im2, dm2, warpIm2, warpDm2, u, v, w, dW, sW, mW, rW, wk,
                         view));
  if (opts && opts->coarse_levels > 0) TRY(get_cg_work(c, "d.", W, H, wk));
  view.allow_early = allowEarly ? 1 : 0;
  if (opts) {
    view.storage = opts->storage;
    set_omega(view, opts->omega, opts->pixel_update);
  }
  const SolveScalars k = scalars_from(opts, cw, alpha, gamma, nOuter, nInner, nSOR);
  std::vector<LevelView> items{view};
  StepPlan sp;
  TRY(one_step(c, "d.desc", items, opts, sp, false, &wk));
  hipStream_t s = c->stream;
  c->relax_ev_used = 0;
  c->relax_n = 0;
  HIP_TRY(c, hipEventRecord(c->ev0, s));
  launch_level_weights(s, sp.vpix, sp.pix);
  enqueue_level_solve(c, sp, k);
  HIP_TRY(c, hipEventRecord(c->ev1, s));
  SolveCtl ctl;
  TRY(download(c, &ctl, view.ctl, sizeof(SolveCtl)));
  TRY(download(c, u, wk.u[0], N * 8));
  TRY(download(c, v, wk.u[0] + N, N * 8));
  TRY(download(c, w, wk.u[0] + 2 * N, N * 8));
  TRY(sync(c));
  fill_stats(c, ctl, view, nOuter, st);
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
  LevelView view;
  TRY(stage_level_inputs(c, W, H, focal_level, Im1, dm1, Im2, dm2, warpIm2, warpDm2, u, v, w, dW, sW, mW, rW, wk,
                         view));
  const SolveScalars k = scalars_from(nullptr, cw, alpha, gamma, 1, 1, 0);
  std::vector<LevelView> items{view};
  StepPlan sp;
  TRY(one_step(c, "d.desc", items, nullptr, sp));
  hipStream_t s = c->stream;
  launch_level_weights(s, sp.vpix, sp.pix);
  launch_ctl_reset(s, sp.vpix);
  launch_derivs(s, sp.vpix, sp.pix, 0, sp.img);
  launch_phi(s, sp.vpix, sp.pix, sp.img);
  launch_system(s, sp.vpix, sp.pix, k, sp.img);
  std::vector<float> rec((size_t)kRec * 4 * N);
  TRY(download(c, rec.data(), wk.coef, rec.size() * sizeof(float)));
  TRY(sync(c));
  // record order (eRxy, eRz, eDxy, eDz, Buu, Bvv, Bww, Buv, Buw, Bvw, ru, rv, rw, ...) is the plane order;
  // value k of pixel l is component k % 4 of part k / 4, parts being planes of N float4
  for (size_t l = 0; l < N; l++)
    for (int k = 0; k < kCoef; k++) coef13[(size_t)k * N + l] = rec[((k / 4) * N + l) * 4 + k % 4];
  return RGBDFLOW_OK;
}

namespace {

// Uploads a compact system (13 planes, f32 values widened to f64) as f32
// records and x ([3][N] planes) as the increment, and plans one relaxation
// step over it.
int stage_relax(rgbdflow_ctx* c, const double* coef13, int W, int H, const double* x, const rgbdflow_params* sched,
                Work& wk, StepPlan& sp, LevelView& view) {
  const size_t N = (size_t)W * H;
  TRY(get_work(c, "d.", N, wk));
  // planes -> records (the reciprocals are IEEE single-precision divisions, as on the device)
  std::vector<float> rec((size_t)kRec * 4 * N), hx(4 * N, 0.f);
  // (parts of the records are planes of N float4: value k of pixel l is component k % 4 of part k / 4)
  for (size_t l = 0; l < N; l++) {
    float r[16];
    for (int k = 0; k < kCoef; k++) r[k] = (float)coef13[(size_t)k * N + l];
    r[13] = 1.0f / r[kBuu];
    r[14] = 1.0f / r[kBvv];
    r[15] = 1.0f / r[kBww];
    for (int k = 0; k < 16; k++) rec[((k / 4) * N + l) * 4 + k % 4] = r[k];
    for (int k = 0; k < 3; k++) hx[l * 4 + k] = (float)x[(size_t)k * N + l];
  }
  const int storage = sched ? sched->storage : RGBDFLOW_STORE_F32;
  if (storage < RGBDFLOW_STORE_F32 || storage > RGBDFLOW_STORE_BF16)
    return fail(c, RGBDFLOW_ERR_ARG, "bad system storage");
  if (sched) TRY(check_block_storage(c, sched->pixel_update, storage));
  if (storage == RGBDFLOW_STORE_F32) {
    TRY(upload(c, wk.coef, rec.data(), rec.size() * 4));
  } else {
    // the records as the level solve stores them in that storage
    float4* rec32;
    TRY(get(c, "d.rec32", (size_t)kRec * N, &rec32));
    TRY(upload(c, rec32, rec.data(), rec.size() * 4));
    launch_encode_records(c->stream, rec32, wk.coef, (int)N, storage);
  }
  TRY(upload(c, wk.du, hx.data(), hx.size() * 4));
  view = make_view(wk, W, H, nullptr, nullptr, wk.u[0], 0);
  view.storage = storage;
  if (sched) set_omega(view, sched->omega, sched->pixel_update);
  std::vector<LevelView> items{view};
  TRY(one_step(c, "d.desc", items, sched, sp));
  launch_ctl_reset(c->stream, sp.vpix);
  return RGBDFLOW_OK;
}

int unstage_relax(rgbdflow_ctx* c, const Work& wk, int W, int H, double* x) {
  const size_t N = (size_t)W * H;
  std::vector<float> hx(4 * N);
  TRY(download(c, hx.data(), wk.du, hx.size() * 4));
  TRY(sync(c));
  for (size_t l = 0; l < N; l++)
    for (int k = 0; k < 3; k++) x[(size_t)k * N + l] = hx[l * 4 + k];
  return RGBDFLOW_OK;
}

}  // namespace

int rgbdflow_relax(rgbdflow_ctx* c, const double* coef13, int W, int H, int n_iter, const rgbdflow_params* sched,
                   double* x) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !coef13 || !x || n_iter < 0) return fail(c, RGBDFLOW_ERR_ARG, "bad relax arguments");
  if (n_iter / kCheckEvery + 1 > kMaxChecks) return fail(c, RGBDFLOW_ERR_ARG, "n_iter too large");
  HIP_TRY(c, hipSetDevice(c->device));
  Work wk;
  StepPlan sp;
  LevelView view;
  TRY(stage_relax(c, coef13, W, H, x, sched, wk, sp, view));
  SolveScalars k{};
  k.n_sor = n_iter;
  k.inner_exit = 0;  // checks are recorded but never stop this call
  enqueue_relax(c, sp, k, 0);
  return unstage_relax(c, wk, W, H, x);
}

int rgbdflow_relax_residual(rgbdflow_ctx* c, const double* coef13, int W, int H, int max_iter, double tol,
                            const rgbdflow_params* sched, double* x, int* iters, double* residual, double* energy) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !coef13 || !x || max_iter < 0 || !(tol >= 0))
    return fail(c, RGBDFLOW_ERR_ARG, "bad relax_residual arguments");
  if (max_iter / kCheckEvery + 1 > kMaxChecks) return fail(c, RGBDFLOW_ERR_ARG, "max_iter too large");
  HIP_TRY(c, hipSetDevice(c->device));
  Work wk;
  StepPlan sp;
  LevelView view;
  TRY(stage_relax(c, coef13, W, H, x, sched, wk, sp, view));
  SolveScalars k{};
  k.n_sor = max_iter;
  k.inner_exit = RGBDFLOW_EXIT_RESIDUAL;
  k.residual_tol = tol;
  enqueue_relax(c, sp, k, 0);
  SolveCtl ctl;
  TRY(download(c, &ctl, view.ctl, sizeof(SolveCtl)));
  TRY(unstage_relax(c, wk, W, H, x));
  if (iters) *iters = ctl.sweeps_exec;
  if (residual) *residual = ctl.resid_checks ? ctl.resid : -1.0;
  if (energy) *energy = ctl.energy;
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

int rgbdflow_fill_depth_holes(rgbdflow_ctx* c, const double* depth, int W, int H, double* out, int* unfilled) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !depth || !out) return fail(c, RGBDFLOW_ERR_ARG, "bad hole-fill arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  double* d;
  int *st, *todo, *left;
  TRY(get(c, "fill.d", N, &d));
  TRY(get(c, "fill.st", N, &st));
  TRY(get(c, "fill.todo", 2 * N, &todo));
  TRY(get(c, "fill.left", 1, &left));
  TRY(upload(c, d, depth, N * 8));
  launch_median_fill(c->stream, d, W, H, st, todo, left);
  HIP_TRY(c, hipGetLastError());
  TRY(download(c, out, d, N * 8));
  int l = 0;
  TRY(download(c, &l, left, sizeof(int)));
  TRY(sync(c));
  if (unfilled) *unfilled = l;
  return RGBDFLOW_OK;
}

int rgbdflow_fill_depth_holes_noalloc(rgbdflow_ctx* c, void* stream, const double* depth, int W, int H, double* out,
                                      int* unfilled) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !depth || !out) return fail(c, RGBDFLOW_ERR_ARG, "bad hole-fill arguments");
  const size_t N = (size_t)W * H;
  if (N > kNoAllocMaxPixels)
    return fail(c, RGBDFLOW_ERR_ARG, "frame of " + std::to_string(N) + " pixels exceeds the no-allocation limit of " +
                                         std::to_string(kNoAllocMaxPixels) + " pixels");
  hipError_t e = hipSuccess;
  int* sc = scratch_of(c->device, &e);
  if (!sc)
    return fail(c, RGBDFLOW_ERR_HIP, std::string("hipGetSymbolAddress of the fill scratch: ") + hipGetErrorString(e));
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  if (out != depth) launch_deinterleave(s, depth, (int)N, 1, out);  // one plane: a copy
  launch_median_fill(s, out, W, H, sc, sc + N, unfilled ? unfilled : sc + kNoAllocLeft);
  HIP_TRY(c, hipGetLastError());
  return RGBDFLOW_OK;
}

int rgbdflow_fill_small_holes_colour(rgbdflow_ctx* c, const double* depth, const double* bgr, int W, int H,
                                     int max_region, double* out, int* unfilled) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !depth || !bgr || !out) return fail(c, RGBDFLOW_ERR_ARG, "bad hole-fill arguments");
  if (max_region < 1 || max_region > 63)
    return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "max_region must be in [1, 63]");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  double *d, *b;
  int *label, *st, *todo, *cmax, *left;
  TRY(get(c, "fill.d", N, &d));
  TRY(get(c, "fill.bgr", 3 * N, &b));
  TRY(get(c, "fill.label", N, &label));
  TRY(get(c, "fill.st", N, &st));
  TRY(get(c, "fill.todo", 2 * N, &todo));
  TRY(get(c, "fill.cmax", N, &cmax));
  TRY(get(c, "fill.left", 1, &left));
  TRY(upload(c, d, depth, N * 8));
  TRY(upload_planes(c, bgr, N, 3, b));
  launch_colour_fill(c->stream, d, b, W, H, max_region, label, st, todo, cmax, left);
  HIP_TRY(c, hipGetLastError());
  TRY(download(c, out, d, N * 8));
  int l = 0;
  TRY(download(c, &l, left, sizeof(int)));
  TRY(sync(c));
  if (unfilled) *unfilled = l;
  return RGBDFLOW_OK;
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
  // one launch per wave (a wave's levels read only level 0 or earlier waves)
  std::vector<std::vector<PyrJob>> waves;
  for (int i = 1; i < (int)L.size(); i++) {
    PyrJob j;
    if (!pyramid_job(L, ns, r, C, lvl, i, &j)) return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "filter too wide");
    const int w = (i - 1) / ns;
    if ((int)waves.size() <= w) waves.resize(w + 1);
    waves[w].push_back(j);
  }
  for (const std::vector<PyrJob>& jobs : waves) {
    std::vector<char> host;
    std::vector<int> start(jobs.size() + 1, 0);
    for (size_t j = 0; j < jobs.size(); j++) start[j + 1] = start[j] + pyramid_job_blocks(jobs[j]);
    const size_t io = pack(host, jobs.data(), jobs.size() * sizeof(PyrJob));
    const size_t so = pack(host, start.data(), start.size() * sizeof(int));
    char* dev;
    TRY(upload_desc(c, "gp_desc", host, &dev));
    const WorkList<PyrJob> wl{reinterpret_cast<const PyrJob*>(dev + io), reinterpret_cast<const int*>(dev + so),
                              (int)jobs.size()};
    launch_pyramid_jobs(c->stream, wl, start.back());
  }
  size_t hoff = 0;
  for (size_t i = 0; i < L.size(); i++) {
    const size_t n = (size_t)L[i].W * L[i].H;
    TRY(download_planes(c, lvl[i], n, C, out + hoff));
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

int rgbdflow_heuristic_pb(rgbdflow_ctx* c, const double* bgr, const double* depth, int W, int H, float* out) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !bgr || !depth || !out) return fail(c, RGBDFLOW_ERR_ARG, "bad heuristic_pb arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  double *b, *d;
  float *o, *scratch;
  TRY(get(c, "h.bgr", 3 * N, &b));
  TRY(get(c, "h.depth", N, &d));
  TRY(get(c, "h.out", 2 * N, &o));
  TRY(get(c, "h.scratch", hpb_scratch_floats(W, H), &scratch));
  TRY(upload_planes(c, bgr, N, 3, b));
  TRY(upload(c, d, depth, N * 8));
  launch_hpb(c->stream, hpb_level(W, H, b, d, o, scratch));
  HIP_TRY(c, hipGetLastError());
  TRY(download(c, out, o, 2 * N * 4));
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
  launch_features_planes(c->stream, in, W, H, f);
  return download_planes(c, f, N, kFeat, out5);
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
  TRY(get_work(c, "d.", N, wk));
  double *d1, *d2;
  TRY(get(c, "d.D1", N, &d1));
  TRY(get(c, "d.D2", N, &d2));
  TRY(upload_planes(c, I1, N, kFeat, wk.I1));
  TRY(upload_planes(c, I2, N, kFeat, wk.I2));
  TRY(upload(c, d1, D1, N * 8));
  TRY(upload(c, d2, D2, N * 8));
  TRY(upload(c, wk.u[1], u, N * 8));
  TRY(upload(c, wk.u[1] + N, v, N * 8));
  TRY(upload(c, wk.u[1] + 2 * N, w, N * 8));
  // the level entry with a flow of the level's own size: u taken as it is, then the warp
  LevelView view = make_view(wk, W, H, d1, d2, wk.u[0], 0);
  view.u_prev = wk.u[1];
  view.Wp = W;
  view.Hp = H;
  view.uv_scale = 1;
  std::vector<LevelView> items{view};
  StepPlan sp;
  TRY(one_step(c, "d.desc", items, nullptr, sp));
  launch_level_entry(c->stream, sp.vpix, sp.pix, sp.img);
  TRY(download_planes(c, wk.Iw, N, kFeat, wI));
  TRY(download(c, wD, wk.Dw, N * 8));
  return sync(c);
}

int rgbdflow_upsample_flow(rgbdflow_ctx* c, const double* u, const double* v, const double* w, int Ws, int Hs, int Wd,
                           int Hd, double uv_scale, double* uo, double* vo, double* wo) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(Ws, Hs) || bad_dims(Wd, Hd) || !u || !v || !w || !uo || !vo || !wo)
    return fail(c, RGBDFLOW_ERR_ARG, "bad upsample arguments");
  if (Ws == Wd && Hs == Hd) return fail(c, RGBDFLOW_ERR_ARG, "upsample_flow: source and target sizes are equal");
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t Ns = (size_t)Ws * Hs, Nd = (size_t)Wd * Hd;
  // the level entry of a Wd x Hd level from a Ws x Hs flow; its warp writes
  // scratch planes (frames of zeros)
  Work wk;
  TRY(get_work(c, "d.", Nd, wk));
  double *a, *d1, *d2;
  TRY(get(c, "ups_in", 3 * Ns, &a));
  TRY(get(c, "d.D1", Nd, &d1));
  TRY(get(c, "d.D2", Nd, &d2));
  HIP_TRY(c, hipMemsetAsync(wk.I1, 0, kFeat * Nd * 8, c->stream));
  HIP_TRY(c, hipMemsetAsync(wk.I2, 0, kFeat * Nd * 8, c->stream));
  HIP_TRY(c, hipMemsetAsync(d1, 0, Nd * 8, c->stream));
  HIP_TRY(c, hipMemsetAsync(d2, 0, Nd * 8, c->stream));
  TRY(upload(c, a, u, Ns * 8));
  TRY(upload(c, a + Ns, v, Ns * 8));
  TRY(upload(c, a + 2 * Ns, w, Ns * 8));
  LevelView view = make_view(wk, Wd, Hd, d1, d2, wk.u[0], 0);
  view.u_prev = a;
  view.Wp = Ws;
  view.Hp = Hs;
  view.uv_scale = uv_scale;
  std::vector<LevelView> items{view};
  StepPlan sp;
  TRY(one_step(c, "d.desc", items, nullptr, sp));
  launch_level_entry(c->stream, sp.vpix, sp.pix, sp.img);
  TRY(download(c, uo, wk.u[0], Nd * 8));
  TRY(download(c, vo, wk.u[0] + Nd, Nd * 8));
  TRY(download(c, wo, wk.u[0] + 2 * Nd, Nd * 8));
  return sync(c);
}

// ---------------------------------------------------------------- batches
int rgbdflow_batch_load(rgbdflow_ctx* c, const rgbdflow_params* p, int n, const rgbdflow_pair_input* pairs) {
  return batch_load(c, p, n, pairs);
}
int rgbdflow_batch_run(rgbdflow_ctx* c, const rgbdflow_params* p, rgbdflow_stats* stats) {
  return batch_run(c, p, stats);
}
int rgbdflow_batch_read(rgbdflow_ctx* c, int i, double* u, double* v, double* w) { return batch_read(c, i, u, v, w); }

// ---------------------------------------------------------------- pair path
int rgbdflow_pair_load(rgbdflow_ctx* c, const rgbdflow_params* p, int W, int H, const double* im1, const double* dm1,
                       const double* im2, const double* dm2, const float* sigma1, const float* hpb) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!p) return fail(c, RGBDFLOW_ERR_ARG, "bad pair arguments");
  rgbdflow_pair_input in;
  std::memset(&in, 0, sizeof(in));
  in.W = W;
  in.H = H;
  in.im1 = im1;
  in.dm1 = dm1;
  in.im2 = im2;
  in.dm2 = dm2;
  in.sigma1 = sigma1;
  in.hpb = hpb;
  return batch_load(c, p, 1, &in);
}

int rgbdflow_pair_run(rgbdflow_ctx* c, const rgbdflow_params* p, rgbdflow_stats* st) {
  if (c && c->loaded && c->pairs.size() != 1) return fail(c, RGBDFLOW_ERR_STATE, "a batch is loaded: use batch_run");
  return batch_run(c, p, st);
}

int rgbdflow_pair_read(rgbdflow_ctx* c, double* u, double* v, double* w) { return batch_read(c, 0, u, v, w); }

int rgbdflow_multiscale(rgbdflow_ctx* c, int W, int H, const double* im1, const double* dm1, const double* im2,
                        const double* dm2, const float* sigma1, const float* hpb, const rgbdflow_params* p, double* u,
                        double* v, double* w, rgbdflow_stats* st) {
  TRY(rgbdflow_pair_load(c, p, W, H, im1, dm1, im2, dm2, sigma1, hpb));
  rgbdflow_stats local;
  TRY(rgbdflow_pair_run(c, p, st ? st : &local));
  return rgbdflow_pair_read(c, u, v, w);
}

}  // extern "C"

namespace {

// `batch` independent W x H test systems (own coefficients and increments)
// for timing relaxation schedules.
int timing_items(rgbdflow_ctx* c, int W, int H, int batch, int storage, std::vector<LevelView>& items) {
  const size_t N = (size_t)W * H;
  float4 *coef, *du, *du_alt;
  SolveCtl* ctl;
  TRY(get(c, "at.coef", (size_t)batch * kRec * N, &coef));
  TRY(get(c, "at.du", (size_t)batch * N, &du));
  TRY(get(c, "at.du_alt", (size_t)batch * N, &du_alt));
  TRY(get(c, "at.ctl", batch, &ctl));
  items.assign(batch, LevelView{});
  for (int b = 0; b < batch; b++) {
    LevelView& v = items[b];
    std::memset(&v, 0, sizeof(v));
    v.W = W;
    v.H = H;
    v.N = (int)N;
    v.coef = coef + (size_t)b * kRec * N;
    v.du = du + (size_t)b * N;
    v.du_alt = du_alt + (size_t)b * N;
    v.ctl = ctl + b;
    set_omega(v, .7f);
    v.storage = storage;
    launch_fill_test_system(c->stream, v.coef, v.du, W, H, storage);
  }
  return RGBDFLOW_OK;
}

// Device time per red-black iteration of one check interval (25 iterations)
// of a schedule on the timing items: best of 3 after a warm-up.
int time_schedule(rgbdflow_ctx* c, std::vector<LevelView>& items, const TileSched& ts, hipEvent_t e0, hipEvent_t e1,
                  double* us) {
  const int W = items[0].W, H = items[0].H;
  for (LevelView& v : items) v.ts = ts;
  StepPlan sp;
  sp.tiled_ok = true;
  sp.cfg = ts.config;
  sp.slots = tiled_slots(ts, W, H, ts.tmax);
  sp.T = ts.tmax;
  TRY(one_step(c, "at.desc", items, nullptr, sp, true));
  SolveScalars k{};
  k.n_sor = kCheckEvery;  // one check interval: the unit the schedule repeats
  k.inner_exit = 0;
  hipStream_t s = c->stream;
  const bool timing = c->time_relax;
  c->time_relax = false;
  launch_ctl_reset(s, sp.vpix);
  enqueue_relax(c, sp, k, 0);  // warm-up
  double t_best = 1e300;
  for (int rep = 0; rep < 3; rep++) {
    hipEventRecord(e0, s);
    enqueue_relax(c, sp, k, 0);
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    t_best = std::min(t_best, (double)ms);
  }
  c->time_relax = timing;
  *us = t_best * 1e3 / kCheckEvery;
  return sync(c);
}

// The tile schedule a caller's choice describes, validated for a W x H level.
int schedule_from_choice(rgbdflow_ctx* c, int W, int H, const rgbdflow_tile_choice* ch, TileSched* out) {
  if (ch->tile_config < 0 || ch->tile_config >= tiled_config_count())
    return fail(c, RGBDFLOW_ERR_ARG, "bad workgroup shape index");
  if (ch->tile_iters < 1 || ch->tile_iters > kCheckEvery) return fail(c, RGBDFLOW_ERR_ARG, "bad iterations per launch");
  int nt, pp;
  tiled_config_shape(ch->tile_config, &nt, &pp);
  const int tr = tiled_config_stream(ch->tile_config);
  if (tr > 0) {
    const int TW = ch->tile_w, TH = ch->tile_h;
    if (ch->single || ch->tile_iters > tr) return fail(c, RGBDFLOW_ERR_ARG, "bad iterations for a streamed schedule");
    if (TW < 2 || (TW & 1) || TH < 1) return fail(c, RGBDFLOW_ERR_ARG, "bad tile shape");
    if (!stream_width_ok(TW, W, ch->tile_iters))
      return fail(c, RGBDFLOW_ERR_ARG, "tile and halo exceed a streamed region row");
    *out = TileSched{TW, TH, (W + TW - 1) / TW, (H + TH - 1) / TH, 0, stream_pairs(), ch->tile_config, ch->tile_iters};
    return RGBDFLOW_OK;
  }
  const int cap = nt * pp;
  TileSched t{};
  if (ch->single) {
    if ((long long)((W + 1) / 2) * H > cap) return fail(c, RGBDFLOW_ERR_ARG, "level does not fit one workgroup");
    t = TileSched{W + (W & 1), H, 1, 1, 1, cap, ch->tile_config, ch->tile_iters};
  } else {
    const int TW = ch->tile_w, TH = ch->tile_h, h = 2 * ch->tile_iters;
    if (TW < 2 || (TW & 1) || TH < 1) return fail(c, RGBDFLOW_ERR_ARG, "bad tile shape");
    const int rw = std::min(TW + 2 * h, W + (W & 1)), rh = std::min(TH + 2 * h, H);
    if ((long long)((rw + 1) / 2) * rh > cap) return fail(c, RGBDFLOW_ERR_ARG, "tile and halo exceed the workgroup");
    t = TileSched{TW, TH, (W + TW - 1) / TW, (H + TH - 1) / TH, 0, cap, ch->tile_config, ch->tile_iters};
  }
  if (tiled_lds_bytes(tiled_slots(t, W, H, t.tmax)) > c->lds_per_block)
    return fail(c, RGBDFLOW_ERR_ARG, "tile exceeds the workgroup's LDS");
  *out = t;
  return RGBDFLOW_OK;
}

void choice_from_schedule(const TileSched& t, int candidates, double us, rgbdflow_tile_choice* out) {
  out->tile_w = t.TW;
  out->tile_h = t.TH;
  out->tile_iters = t.tmax;
  out->tile_config = t.config;
  tiled_config_shape(t.config, &out->threads, &out->pairs_per_thread);
  out->single = t.single;
  out->candidates = candidates;
  out->us_per_iter = us;
  out->streamed = tiled_config_stream(t.config);
}

}  // namespace

extern "C" {

int rgbdflow_relax_shape(int index, int* threads, int* pairs_per_thread, int* streamed) {
  if (index < 0 || index >= tiled_config_count()) return RGBDFLOW_ERR_ARG;
  int nt, pp;
  tiled_config_shape(index, &nt, &pp);
  if (threads) *threads = nt;
  if (pairs_per_thread) *pairs_per_thread = pp;
  if (streamed) *streamed = tiled_config_stream(index);
  return RGBDFLOW_OK;
}

int rgbdflow_autotune_relax(rgbdflow_ctx* c, int W, int H, int batch, const rgbdflow_params* p,
                            rgbdflow_tile_choice* out) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || batch < 1 || batch > 256) return fail(c, RGBDFLOW_ERR_ARG, "bad autotune size");
  HIP_TRY(c, hipSetDevice(c->device));
  std::vector<LevelView> items;
  TRY(timing_items(c, W, H, batch, p ? p->storage : RGBDFLOW_STORE_F32, items));
  // candidate schedules
  std::vector<TileSched> cand;
  const long long pairs_img = (long long)((W + 1) / 2) * H;
  const int Ts[] = {1, 2, 3, 4, 5, 6, 8, 12};
  const int TWs[] = {8, 16, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320};
  for (int cfg = 0; cfg < tiled_config_count(); cfg++) {
    int nt, pp;
    tiled_config_shape(cfg, &nt, &pp);
    const int tr = tiled_config_stream(cfg);
    if (tr > 0) {
      // streamed: the fewest columns of tiles whose region rows fit, and the
      // rows of tiles that give the batch about 1..4 rounds of the resident
      // workgroups (or a few rows more or less)
      const int T = tr;
      int TW = 0;
      for (int k = 1; k <= 64 && TW == 0; k++) {
        const int tw = k == 1 ? W + (W & 1) : ((W + k - 1) / k + 1) & ~1;
        if (stream_width_ok(tw, W, T)) TW = tw;
      }
      if (TW == 0) continue;
      const int tx = (W + TW - 1) / TW;
      // resident workgroups per CU (the grid of a launch with many tiles)
      const int occ = std::max(1, tiled_prepare(cfg, p ? p->storage : RGBDFLOW_STORE_F32, kRelaxFixed,
                                                stream_slots(T), 1 << 20) / c->cu_count);
      std::vector<int> ths;
      for (int rounds = 1; rounds <= 4; rounds++) {
        const int rows = std::max(1, (rounds * c->cu_count * occ + batch * tx - 1) / (batch * tx));
        for (int d = -1; d <= 1; d++) {
          const int r = std::max(1, rows + d);
          const int th = (H + r - 1) / r;
          if (th >= 2 && std::find(ths.begin(), ths.end(), th) == ths.end()) ths.push_back(th);
        }
      }
      for (int TH : ths) cand.push_back(TileSched{TW, TH, tx, (H + TH - 1) / TH, 0, stream_pairs(), cfg, T});
      continue;
    }
    const int cap = nt * pp;
    if (pairs_img <= cap) {
      const TileSched t{W + (W & 1), H, 1, 1, 1, cap, cfg, kCheckEvery};
      if (tiled_lds_bytes(tiled_slots(t, W, H, 0)) <= c->lds_per_block) cand.push_back(t);
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
          const TileSched t{TW, TH, (W + TW - 1) / TW, (H + TH - 1) / TH, 0, cap, cfg, T};
          if (tiled_lds_bytes(tiled_slots(t, W, H, T)) > c->lds_per_block) continue;
          cand.push_back(t);
        }
      }
    }
  }
  if (cand.empty()) return fail(c, RGBDFLOW_ERR_UNSUPPORTED, "no tiled schedule fits this level");
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  double best_us = 1e300, best_tiled_us = 1e300;
  TileSched best{}, best_tiled{};
  int rc = RGBDFLOW_OK;
  for (const TileSched& ts : cand) {
    double us = 0;
    rc = time_schedule(c, items, ts, e0, e1, &us);
    if (rc != RGBDFLOW_OK) break;
    if (us < best_us) {
      best_us = us;
      best = ts;
    }
    if (tiled_config_stream(ts.config) == 0 && us < best_tiled_us) {
      best_tiled_us = us;
      best_tiled = ts;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != RGBDFLOW_OK) return rc;
  // the shape's schedule, and the one of this record storage (a batch of
  // another storage takes it only if that storage has none of its own)
  const int storage = p ? p->storage : RGBDFLOW_STORE_F32;
  c->plans[plan_key(W, H, 0, 0, 0, -1)] = best;
  c->plans[storage_key(plan_key(W, H, 0, 0, 0, -1), storage)] = best;
  // the best tiled schedule too, for steps whose items have several sizes
  if (best_tiled.cap != 0) {
    c->plans[plan_key(W, H, 0, 0, 0, kAnyTiledConfig)] = best_tiled;
    c->plans[storage_key(plan_key(W, H, 0, 0, 0, kAnyTiledConfig), storage)] = best_tiled;
  }
  c->plans_gen++;
  if (out) choice_from_schedule(best, (int)cand.size(), best_us, out);
  return RGBDFLOW_OK;
}

int rgbdflow_time_relax_schedule(rgbdflow_ctx* c, int W, int H, int batch, const rgbdflow_params* p,
                                 const rgbdflow_tile_choice* ch, double* us_per_iter) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || batch < 1 || batch > 256 || !ch || !us_per_iter)
    return fail(c, RGBDFLOW_ERR_ARG, "bad timing arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  TileSched t;
  TRY(schedule_from_choice(c, W, H, ch, &t));
  std::vector<LevelView> items;
  TRY(timing_items(c, W, H, batch, p ? p->storage : RGBDFLOW_STORE_F32, items));
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  const int rc = time_schedule(c, items, t, e0, e1, us_per_iter);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return rc;
}

int rgbdflow_set_relax_schedule(rgbdflow_ctx* c, int W, int H, const rgbdflow_tile_choice* ch) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (bad_dims(W, H) || !ch) return fail(c, RGBDFLOW_ERR_ARG, "bad schedule arguments");
  TileSched t;
  TRY(schedule_from_choice(c, W, H, ch, &t));
  c->plans[plan_key(W, H, 0, 0, 0, -1)] = t;
  if (tiled_config_stream(t.config) == 0) c->plans[plan_key(W, H, 0, 0, 0, kAnyTiledConfig)] = t;
  // an installed schedule holds for every record storage
  for (int s = RGBDFLOW_STORE_F32; s <= RGBDFLOW_STORE_BF16; s++) {
    c->plans.erase(storage_key(plan_key(W, H, 0, 0, 0, -1), s));
    c->plans.erase(storage_key(plan_key(W, H, 0, 0, 0, kAnyTiledConfig), s));
  }
  c->plans_gen++;
  return RGBDFLOW_OK;
}

int rgbdflow_with_init(rgbdflow_ctx* c, int W, int H, const double* im1, const double* dm1, const double* im2,
                       const double* dm2, const float* sigma1, const float* hpb0, const double* est,
                       const rgbdflow_params* p, double* u, double* v, double* w, rgbdflow_stats* st) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!p || bad_dims(W, H) || !im1 || !dm1 || !im2 || !dm2 || !est || !u || !v || !w)
    return fail(c, RGBDFLOW_ERR_ARG, "bad with_init arguments");
  if ((int)p->cam.xRes != W || (int)p->cam.yRes != H)
    return fail(c, RGBDFLOW_ERR_ARG, "params.cam resolution differs from the image size");
  TRY(check_run_params(c, p));
  HIP_TRY(c, hipSetDevice(c->device));
  const size_t N = (size_t)W * H;
  Work wk;
  TRY(get_work(c, "d.", N, wk));
  if (p->coarse_levels > 0) TRY(get_cg_work(c, "d.", W, H, wk));
  double *i1, *i2, *d1, *d2, *del0, *sm0;
  float *s1, *h = nullptr;
  TRY(get(c, "wi.im1", 3 * N, &i1));
  TRY(get(c, "wi.im2", 3 * N, &i2));
  TRY(get(c, "wi.dm1", N, &d1));
  TRY(get(c, "wi.dm2", N, &d2));
  TRY(get(c, "wi.sig", N, &s1));
  TRY(get(c, "wi.del0", N, &del0));
  TRY(get(c, "wi.sm0", N, &sm0));
  const bool hpb_dev = !hpb0 && p->regularization_type == RGBDFLOW_REG_HPB;
  if (hpb0 || hpb_dev) TRY(get(c, "wi.hpb", N * 2, &h));
  float* hs = nullptr;
  if (hpb_dev) TRY(get(c, "hpb.scratch", hpb_scratch_floats(W, H), &hs));
  TRY(upload_planes(c, im1, N, 3, i1));
  TRY(upload_planes(c, im2, N, 3, i2));
  TRY(upload(c, d1, dm1, N * 8));
  TRY(upload(c, d2, dm2, N * 8));
  TRY(upload_planes(c, est, N, 3, wk.u[1]));
  if (hpb0) TRY(upload(c, h, hpb0, N * 2 * 4));
  if (sigma1) TRY(upload(c, s1, sigma1, N * 4));
  // per-pair weights, as the batch path computes them
  PairPrep pp;
  std::memset(&pp, 0, sizeof(pp));
  pp.dm1 = d1;
  pp.sigma = s1;
  pp.del0 = del0;
  pp.sm0 = sm0;
  pp.W = W;
  pp.H = H;
  pp.compute_sigma = sigma1 ? 0 : 1;
  pp.dvc = p->depth_vs_color_data_weight;
  const double f0 = p->cam.focalLength;
  pp.f2 = f0 * f0;
  std::vector<char> ph;
  const int pstart[2] = {0, pair_prep_blocks(W, H)};
  const size_t pio = pack(ph, &pp, sizeof(pp));
  const size_t pso = pack(ph, pstart, sizeof(pstart));
  char* pdev;
  TRY(upload_desc(c, "wi.desc", ph, &pdev));
  const WorkList<PairPrep> pwl{reinterpret_cast<const PairPrep*>(pdev + pio), reinterpret_cast<const int*>(pdev + pso),
                               1};
  // sceneFlowWithInit.cpp:7-81: one level at full size from est, no early exit
  LevelView view = make_view(wk, W, H, d1, d2, wk.u[0], 0);
  view.allow_early = 0;
  view.P1k = i1;
  view.P2k = i2;
  view.delk = del0;
  view.smk = sm0;
  view.hpbk = h;
  view.u_prev = wk.u[1];
  view.Wp = W;
  view.Hp = H;
  view.uv_scale = 1;
  view.reg_type = p->regularization_type;
  view.storage = p->storage;
  view.img = p->image_storage;
  set_omega(view, p->omega, p->pixel_update);
  view.W_full = (int)p->cam.xRes;
  view.f_full = p->cam.focalLength;
  view.f_level = p->cam.focalLength;
  std::vector<LevelView> items{view};
  StepPlan sp;
  TRY(one_step(c, "d.desc", items, p, sp, false, &wk));
  const SolveScalars k = scalars_from(p, p->color_data_weight, p->alpha, p->gamma, p->n_outer, p->n_inner, p->n_sor);
  hipStream_t s = c->stream;
  c->relax_ev_used = 0;
  c->relax_n = 0;
  HIP_TRY(c, hipEventRecord(c->ev0, s));
  launch_pair_prep(s, pwl, pstart[1]);
  // sceneFlowWithInit.cpp:51: computeHeuristicPb(Im1, dm1) when no map is given
  if (hpb_dev) launch_hpb(s, hpb_level(W, H, i1, d1, h, hs));
  launch_features(s, sp.vpix, sp.pix, sp.img);
  launch_level_entry(s, sp.vpix, sp.pix, sp.img);
  launch_level_weights(s, sp.vpix, sp.pix);
  enqueue_level_solve(c, sp, k);
  HIP_TRY(c, hipEventRecord(c->ev1, s));
  SolveCtl ctl;
  TRY(download(c, &ctl, view.ctl, sizeof(SolveCtl)));
  TRY(download(c, u, wk.u[0], N * 8));
  TRY(download(c, v, wk.u[0] + N, N * 8));
  TRY(download(c, w, wk.u[0] + 2 * N, N * 8));
  TRY(sync(c));
  fill_stats(c, ctl, view, p->n_outer, st);
  return RGBDFLOW_OK;
}

}  // extern "C"

namespace rgbdflow {

int batch_load_device(rgbdflow_ctx* c, const rgbdflow_params* p, int n, const PairDev* in) {
  if (!c) return RGBDFLOW_ERR_ARG;
  if (!p || n < 1 || !in) return fail(c, RGBDFLOW_ERR_ARG, "bad batch arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  // The shapes of every slot; a batch with the same shapes as the loaded one
  // gets the same slot buffers back from the pool, so its batch plan (and a
  // captured graph of it) stays valid: only the frames are copied in.
  std::string sig;
  for (int i = 0; i < n; i++) {
    const int f[8] = {in[i].W, in[i].H, in[i].min_width, in[i].f1.hpb != nullptr, in[i].init != nullptr,
                      in[i].f1.pim != nullptr && in[i].f1.pd != nullptr, in[i].f1.sigma != nullptr,
                      in[i].f1.hpb_dm != nullptr};
    sig.append(reinterpret_cast<const char*>(f), sizeof(f));
    sig.append(reinterpret_cast<const char*>(&in[i].cam), sizeof(in[i].cam));
  }
  const bool same = c->loaded && !c->dev_sig.empty() && sig == c->dev_sig && (int)c->pairs.size() == n &&
                    p->pyramid_ratio == c->load_ratio && p->min_width == c->load_min_width;
  c->loaded = false;
  c->ran = false;
  c->dev_sig.clear();
  if (!same) {
    c->gen++;
    c->pairs.assign(n, PairSlot());
  }
  for (int i = 0; i < n; i++) TRY(load_pair_device(c, p, i, in[i], c->pairs[i]));
  c->load_ratio = p->pyramid_ratio;
  c->load_min_width = p->min_width;
  c->dev_sig = sig;
  c->loaded = true;
  return RGBDFLOW_OK;
}

int pair_outputs(rgbdflow_ctx* c, int i, FrameDev* f2, const double** flow, size_t* pyr_elems) {
  if (!c || !c->ran || i < 0 || i >= (int)c->pairs.size()) return RGBDFLOW_ERR_STATE;
  const PairSlot& ps = c->pairs[i];
  *f2 = FrameDev{};
  f2->im = ps.im2;
  f2->dm = ps.dm2;
  f2->pim = ps.pim2;
  f2->pd = ps.pd2;
  f2->hpb = nullptr;
  *flow = ps.wk.u[((int)ps.L.size() - 1) & 1];
  *pyr_elems = pyramid_elems(ps.L, 1);
  return RGBDFLOW_OK;
}

hipStream_t ctx_stream(rgbdflow_ctx* c) { return c->stream; }

size_t pyramid_rest_elems(int W, int H, double ratio, int min_width) {
  return pyramid_elems(pyramid_shape(W, H, ratio, min_width, nullptr, nullptr), 1);
}
size_t pyramid_all_elems(int W, int H, double ratio, int min_width) {
  return pyramid_elems(pyramid_shape(W, H, ratio, min_width, nullptr, nullptr), 0);
}

}  // namespace rgbdflow
