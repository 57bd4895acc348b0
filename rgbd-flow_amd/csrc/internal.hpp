// internal.hpp -- device-side data layouts and kernel launchers of librgbdflow.
//
// Every image lives in HBM as structure-of-arrays planes: channel c of an
// image with N pixels is the contiguous range [c*N, (c+1)*N), row-major.
//
// Kernels run over work-lists: one launch serves every item of a list (the
// same pyramid level of all frame pairs of a batch, say), and each block finds
// its item by a binary search over the list's block offsets.  A single pair is
// a list of one.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace rgbdflow {

constexpr int kFeat = 5;            // feature channels: gray, d/dx, d/dy, G-B, G-R
constexpr int kCoef = 13;           // compact per-pixel system, see Coef below
constexpr int kMaxLevels = 16;
constexpr int kCheckEvery = 25;     // relaxation convergence check period (.cu:905)
constexpr int kMaxChecks = 4096;    // check slots per outer iteration
constexpr int kMaxOuter = 1024;
constexpr int kMaxCg = 8;          // coarse levels of the coarse-grid correction
constexpr int kCgFusePx = 2048;    // pixels of the coarse levels one workgroup relaxes in LDS
constexpr int kCgBottomThreads = 1024;
constexpr int kBlock = 256;         // threads of the per-pixel kernels

// The compact system of one pixel p, in the order of the 13-plane export
// (rgbdflow_level_system): right/down edge weights alpha*phi_p*r_p (xy rows)
// and alpha*phi_p*r_p*s_p (z row), the symmetric 3x3 block B (diagonal
// includes gamma terms) and the rhs.
enum Coef { kERxy = 0, kERz, kEDxy, kEDz, kBuu, kBvv, kBww, kBuv, kBuw, kBvw, kRu, kRv, kRw };

// In HBM the system is one 64-byte record per pixel (kRec float4), so the
// relaxation loads a pixel with four 16-byte loads:
//   rec[0] = (eRxy, eRz, eDxy, eDz)   rec[1] = (Buu, Bvv, Bww, Buv)
//   rec[2] = (Buw, Bvw, ru, rv)       rec[3] = (rw, 1/Buu, 1/Bvv, 1/Bww)
// The reciprocals are IEEE single-precision divisions, made once per system.
constexpr int kRec = 4;

// Device control block of one level solve.  Kernels read it at entry to skip
// work after an early exit, so the host never waits inside a level.
struct SolveCtl {
  int outer_done;                     // outer loop left early (stat < threshold)
  int outer_exit_at;                  // outer iteration that set outer_done
  int sweeps_exec;                    // red-black iterations executed (all outers)
  int du_in_alt;                      // the last tiled launch that ran wrote du_alt
  int resid_checks;                   // residual evaluations of this level solve
  double resid;                       // last relative residual ||r - M du|| / ||r|| (f64)
  double energy;                      // last energy r'du - 1/2 du'M du of the system (f64)
  unsigned int maxdiff[kMaxChecks];   // float bits of max|du change| per check of the
                                      // current outer, [inner][check]; zeroed per outer
  double stat[kMaxOuter];             // outer convergence statistic per outer iteration
};

// Tile schedule of the tiled relaxation (kernels_relax_tiled.hip).
struct TileSched {
  int TW, TH;            // interior tile size (TW even)
  int tiles_x, tiles_y;
  int single;            // 1: one tile covers the level (no halo, up to 25 iterations per launch)
  int cap;               // pair capacity of a workgroup (LDS plane length)
  int config;            // kernel configuration index
  int tmax;              // iterations per launch
};

// Everything a level kernel needs about one pair at one pyramid level.
struct LevelView {
  int W, H, N;
  const double* I1;     // [5][N] features of frame 1
  const double* I2;     // [5][N] features of frame 2
  double* Iw;           // [5][N] frame-2 features warped by u
  const double* D1;     // [N]
  const double* D2;     // [N]
  double* Dw;           // [N] frame-2 depth warped by u, minus w
  double* u;            // [3][N] flow (u px, v px, w m)
  float4* du;           // [N] increment (du, dv, dw, unused) solved by the relaxation
  float4* du_alt;       // [N] second buffer of the tiled relaxation (reads one, writes the other)
  float* dW;            // [N] depth-data weights (delta)
  float* sW;            // [N] depth-smoothness weights
  float* mW;            // [N] depth-magnitude weights
  float* rW;            // [2][N] regularisation weights (right, down)
  float* cW;            // [N] convergence weights f/D1
  float* q;             // [N] occlusion weights
  double* Ix;           // [5][N]
  double* Iy;           // [5][N]
  double* It;           // [5][N]
  double* Dx;           // [N]
  double* Dy;           // [N]
  double* Dt;           // [N]
  double* phi;          // [N]
  float4* coef;         // [N][kRec] system records
  double* partials;     // per-block partial sums of the outer statistic
  SolveCtl* ctl;
  int allow_early;      // outer early exit allowed on this level
  // level entry (features, upsample, warp, weights)
  const double* P1k;    // [3][N] colour pyramid level of frame 1 (BGR planes)
  const double* P2k;    // [3][N] frame 2
  const double* delk;   // [N] depth-data weight pyramid level (f64)
  const double* smk;    // [N] depth-smoothness weight pyramid level (f64)
  const float* hpbk;    // [N][2] boundary probabilities, or null
  const double* mag64;  // [N] given depth-magnitude weights (single-level entry), or null
  const double* reg64;  // [N][2] given regularisation weights (single-level entry), or null
  const double* u_prev; // [3][Np] coarser level's flow, or null at the pair's coarsest level
  int Wp, Hp;
  double uv_scale;      // 1 / pyramid ratio
  int reg_type;
  int storage;          // system record storage (records.hpp: f32, f16, bf16)
  float omega;          // relaxation factor (.cu:461: 0.7)
  float omega_c;        // 1 - omega, rounded to float as the update computes it
  int img;              // feature-image storage of I1, I2, Iw (ImgStorage; uniform over a launch)
  int pixel_gs;         // 0: a pixel's unknowns update from each other's old values (.cu:431-459);
                        // 1: u, then v from the new u, then w from both (point SOR order);
                        // 2: the pixel's 3x3 block solved as a whole (block SOR)
  int W_full;           // camera xRes (full size)
  float f_full;         // full-size focal length
  float f_level;        // focal length of the level
  TileSched ts;
  // coarse-grid correction: on the view of a coarse level (2x2 aggregates of
  // the next finer level), that finer level's system and unknowns
  struct Finer {
    int W, H, storage;
    const float4* coef;
    float4* du;
    float4* du_alt;
    const SolveCtl* ctl;      // the finer level's control (coarse levels: outer_done = skip)
    const SolveCtl* top_ctl;  // the relaxed level's control
    int first;                // 1: the finer level is the relaxed level itself
    int level;                // the finer level's index (0: the relaxed level, l: coarse level l)
    int fused;                // on a view of k_cg_bottom: the coarse levels it covers (this and below)
  } finer;
};

// Device code reaches the level buffers through global-address-space
// pointers: the compiler then emits global loads and stores.  Generic
// pointers read from a view give flat accesses, whose completion shares the
// LDS counter, so every LDS wait would also wait for outstanding HBM loads.
#define RGBDFLOW_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ RGBDFLOW_GLOBAL T* gp(T* p) {
  return (RGBDFLOW_GLOBAL T*)p;
}

// 16-byte loads and stores through global pointers (float4 and uint4 are
// classes whose copies do not take address-space-qualified operands; the
// native vector types do).
typedef float gvec4f __attribute__((ext_vector_type(4)));
typedef float gvec3f __attribute__((ext_vector_type(3)));
typedef unsigned gvec4u __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gload(const float4* p) {
  const gvec4f x = *(const RGBDFLOW_GLOBAL gvec4f*)p;
  return make_float4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ void gstore(float4* p, const float4& v) {
  const gvec4f x = {v.x, v.y, v.z, v.w};
  *(RGBDFLOW_GLOBAL gvec4f*)p = x;
}
__device__ __forceinline__ uint4 gload(const uint4* p) {
  const gvec4u x = *(const RGBDFLOW_GLOBAL gvec4u*)p;
  return make_uint4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ void gstore(uint4* p, const uint4& v) {
  const gvec4u x = {v.x, v.y, v.z, v.w};
  *(RGBDFLOW_GLOBAL gvec4u*)p = x;
}

// Sets a view's relaxation factor (<= 0 selects the reference's 0.7) and the
// order of a pixel's unknowns (pixel_gs, see LevelView).
inline void set_omega(LevelView& v, float omega, int pixel_gs = 0) {
  v.omega = omega > 0 ? omega : .7f;
  v.omega_c = 1.f - v.omega;
  v.pixel_gs = pixel_gs;
}

// Feature-image storage (records.hpp: img_ld / img_st).
enum ImgStorage { kImgF64 = 0, kImgF16 = 1, kImgBF16 = 2 };

// Tiled relaxation kernel variants by update rule.
enum RelaxMode { kRelaxFree = 0, kRelaxFixed = 1, kRelaxGS = 2, kRelaxBlock = 3 };
constexpr int kRelaxModes = 4;

// A launch's work-list: items plus the first block of each item (n + 1 entries).
template <class T>
struct WorkList {
  const T* items;
  const int* start;
  int n;
};

template <class T>
__device__ __forceinline__ int work_item(const WorkList<T>& w, int blk) {
  int lo = 0, hi = w.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (w.start[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

typedef WorkList<LevelView> Views;

// The dispatcher hands workgroup b of a launch to XCD b % 8.  xcd_block(b, nb)
// renumbers the nb workgroups so that each XCD takes one contiguous range of
// the launch's blocks, in order: neighbouring tiles and rows then meet in
// the same L2 instead of being fetched from HBM once per XCD.  A bijection on
// [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// One Gaussian-pyramid level of one image (GaussianPyramid.cpp:47-61).
constexpr int kMaxTaps = 31;
struct PyrJob {
  const double* src;
  double* dst;
  int Ws, Hs, Wd, Hd, C, fs;
  double ratio;
  double taps[kMaxTaps];
};

// Per-pair preparation at the start of a solve.
struct PairPrep {
  const double* dm1;    // [N] hole-filled depth of frame 1
  float* sigma;         // [N] depth stdevs (computed here when compute_sigma)
  double* del0;         // [N] depth-data weights, full size
  double* sm0;          // [N] depth-smoothness weights, full size
  int W, H, compute_sigma;
  double dvc;           // depthVsColorDataWeight
  double f2;            // focal length squared
};

// One level of the heuristic boundary probabilities (kernels_hpb.hip):
// inputs, output and the scratch arrays of its stages.
struct HpbLevel {
  int W, H;
  const double* bgr;    // [3][N] colour planes (B, G, R) in [0, 1]
  const double* depth;  // [N] metres, <= 0 invalid
  float* out;           // [N][2] boundary probabilities (right, down)
  float *pts, *col, *nrm;  // [3][N] points, colour, normals
  unsigned* nval;          // [N] normal validity
  int* rad;                // [N] half-disk radius
  float* rowsum;           // [6][N] running row sums before each pixel
  unsigned* vrowsum;       // [N]
  float* ii;               // [6][(W+1)(H+1)] integral images (colour, normals)
  unsigned* iv;            // [(W+1)(H+1)] integral image of the normal validity
  float* feat;             // [4][N] colour / normal half-disk features (horizontal, vertical)
  float* bp;               // [N][2] probabilities before non-maximum suppression
};
size_t hpb_scratch_floats(int W, int H);
HpbLevel hpb_level(int W, int H, const double* bgr, const double* depth, float* out, float* scratch);
void launch_hpb(hipStream_t s, const HpbLevel& L);
// Several levels (of any pairs, each with its own scratch) in one launch per
// stage, passed by value as the kernels' argument (3.5 KB, under the 4 KB
// kernel-argument limit): per stage the first workgroup of each item.
constexpr int kHpbMaxItems = 24;
struct HpbBatch {
  int n;
  int s_pix[kHpbMaxItems + 1];  // per-pixel stages, 256 pixels a workgroup
  int s_row[kHpbMaxItems + 1];  // row sums, 64 rows of one channel a workgroup
  int s_col[kHpbMaxItems + 1];  // integral images, 256 columns x channels a workgroup
  int s_tile[kHpbMaxItems + 1]; // normals, 16x16 pixels a workgroup
  HpbLevel items[kHpbMaxItems];
};
bool hpb_batch_add(HpbBatch& b, const HpbLevel& L);  // false when full
void launch_hpb_batch(hipStream_t s, const HpbBatch& b);

struct SolveScalars {
  double alpha, gamma, color_weight;
  int n_outer, n_inner, n_sor;
  int inner_exit;           // relaxation exits: 0 none, 1 the reference's (max change), 2 FP64 residual
  double residual_tol;      // inner_exit 2: stop once the relative residual is <= this
  int outer_exit;           // outer early exits enabled
  int cg_sweeps, cg_at;     // coarse-grid correction (when the step has coarse levels)
};

// Block counts of one item for each kernel family.
__host__ __device__ inline int pix_blocks(int N) { return (N + kBlock - 1) / kBlock; }
__host__ __device__ inline int plain_blocks(int W, int H) { return (((W + 1) / 2 + 63) / 64) * ((H + 3) / 4); }
__host__ __device__ inline int tiled_blocks(const TileSched& t) { return t.tiles_x * t.tiles_y; }
// tiles of the per-pair preparation (kBlock threads, one pixel each) and the
// radius of the depth stdevs' boundary search (staticDepthNoiseModeling.cpp's
// maxDist, 4 in runFlowTwoFrames.cpp:90)
constexpr int kPrepW = 32, kPrepH = kBlock / kPrepW, kPrepRadius = 4;
__host__ __device__ inline int pair_prep_blocks(int W, int H) {
  return ((W + kPrepW - 1) / kPrepW) * ((H + kPrepH - 1) / kPrepH);
}
// tiles of the fused first-inner kernel (kBlock threads, one pixel each)
constexpr int kOtW = 32, kOtH = kBlock / kOtW;
__host__ __device__ inline int otile_blocks(int W, int H) { return ((W + kOtW - 1) / kOtW) * ((H + kOtH - 1) / kOtH); }

// ---------------------------------------------------------------- launchers
// kernels_solve.hip (blocks = total blocks of the list for that family)
void launch_ctl_reset(hipStream_t s, const Views& v);
void launch_derivs(hipStream_t s, const Views& v, int blocks, int n_slots, int img);
void launch_phi(hipStream_t s, const Views& v, int blocks, int img);
void launch_system(hipStream_t s, const Views& v, int blocks, const SolveScalars& k, int img);
// derivatives + phi + system of the first inner iteration in one pass
// (store: keep derivatives, occlusion weights and phi for later inner iterations)
void launch_outer_first(hipStream_t s, const Views& v, int blocks, int n_slots, const SolveScalars& k, int store,
                        int img);
void launch_relax_plain(hipStream_t s, const Views& v, int blocks, int colour, int slot_base, int n_prev_checks,
                        int check_slot);
void launch_update_warp(hipStream_t s, const Views& v, int blocks, int want_stat, int img);
void launch_stat_final(hipStream_t s, const Views& v, int outer, int decide);
// FP64 residual and energy of every item's system at its current increment
// (alt: the increment is in du_alt), recorded in SolveCtl; check_slot >= 0
// also records the stopping decision (residual <= tol) in that check slot.
void launch_residual(hipStream_t s, const Views& v, int blocks, int alt, int slot_base, int n_prev_checks,
                     int check_slot, double tol);

// kernels_relax_tiled.hip
int tiled_config_count();
void tiled_config_shape(int idx, int* nt, int* p);
// 0 for a k_relax_tiled configuration; TR for a streamed one (k_relax_stream:
// a ring of region rows for up to TR iterations per launch, stream_pairs()
// pixel pairs per region row, any number of rows)
int tiled_config_stream(int idx);
int stream_slots(int TR);
int stream_pairs();
size_t tiled_lds_bytes(int slots);
// LDS slots per colour plane for a tile schedule run T iterations per launch:
// the largest region (tile + halo of 2T, clipped to the level, in pixel pairs)
// padded by one pair on every side; for a streamed schedule, its ring.
inline int tiled_slots(const TileSched& t, int W, int H, int T) {
  const int tr = tiled_config_stream(t.config);
  if (tr > 0) return stream_slots(tr);
  int PR, RH;
  if (t.single) {
    PR = (W + 1) / 2;
    RH = H;
  } else {
    const int h = 2 * T;
    PR = ((t.TW + 2 * h < W ? t.TW + 2 * h : W) + 1) / 2;
    RH = t.TH + 2 * h < H ? t.TH + 2 * h : H;
  }
  return (PR + 2) * (RH + 2);
}
// mode (RelaxMode): kRelaxFixed when every item relaxes as the reference (factor 0.7, old
// values within the pixel), kRelaxGS / kRelaxBlock when every item uses the point SOR order /
// the block update, else kRelaxFree.
// grid: tiled_prepare's for the same (config, storage, mode, slots, blocks),
// obtained when the launch was planned (outside any stream capture)
int tiled_prepare(int config, int storage, int mode, int slots, int blocks);
void launch_relax_tiled(hipStream_t s, const Views& v, int blocks, int grid, int config, int storage, int mode,
                        int slots, int flip, int n_iter, int slot_base, int n_prev_checks, int check_slot);
void launch_copy_du(hipStream_t s, const Views& v, int blocks);
// kernels_coarse.hip: the coarse-grid correction (v: views of one coarse
// level; blocks: pix_blocks of the coarse sizes)
void launch_cg_operator(hipStream_t s, const Views& v, int blocks);
void launch_cg_restrict(hipStream_t s, const Views& v, int blocks, int finer_alt, int slot_base, int n_prev_checks);
void launch_cg_prolong(hipStream_t s, const Views& v, int blocks, int alt, int finer_alt);
// the coarse levels from the first whose remaining levels fit kCgFusePx down,
// one workgroup per item (v.n items); alt_mask bit l: level l's unknowns are
// in du_alt
void launch_cg_bottom(hipStream_t s, const Views& v, int alt_mask, int slot_base, int n_prev_checks, int sweeps);
__host__ __device__ inline int cg_coarse_ok(int W, int H) { return W >= 4 && H >= 4; }
void launch_fill_test_system(hipStream_t s, float4* coef, float4* du, int W, int H, int storage);
// f32 records -> records in `storage` (src and dst distinct)
void launch_encode_records(hipStream_t s, const float4* src, float4* dst, int N, int storage);

// kernels_pyramid.hip
void launch_deinterleave(hipStream_t s, const double* src, int N, int C, double* dst);
void launch_interleave(hipStream_t s, const double* src, int N, int C, double* dst);
void launch_pair_prep(hipStream_t s, const WorkList<PairPrep>& w, int blocks);
void launch_pyramid_jobs(hipStream_t s, const WorkList<PyrJob>& w, int blocks);
// a ratio-0.5 level is made in output tiles of kPyrTW x kPyrTH (one per
// block and channel), other ratios one output sample per thread
constexpr int kPyrTW = 32, kPyrTH = kBlock / kPyrTW;
__host__ __device__ inline bool pyr_halving(const PyrJob& j) {
  return j.ratio == 0.5 && 2 * j.Hd <= j.Hs + 1 && 2 * j.Wd <= j.Ws + 1;
}
inline int pyramid_job_blocks(const PyrJob& j) {
  return pyr_halving(j) ? ((j.Wd + kPyrTW - 1) / kPyrTW) * ((j.Hd + kPyrTH - 1) / kPyrTH) * j.C
                        : pix_blocks(j.Wd * j.Hd * j.C);
}
void launch_features(hipStream_t s, const Views& v, int blocks, int img);
void launch_level_entry(hipStream_t s, const Views& v, int blocks, int img);
void launch_level_weights(hipStream_t s, const Views& v, int blocks);
void launch_depth_sigma(hipStream_t s, const float* depth, int W, int H, int max_dist, float* sigma);
void launch_depth_map_level(hipStream_t s, const double* larger, int LW, int LH, double* dst, int W, int H);
void launch_features_planes(hipStream_t s, const double* bgr, int W, int H, double* out5);
void launch_ingest(hipStream_t s, const unsigned char* bgr, const float* depth, int N, double* im, double* dm);
// depth hole fill in place (the rounds variant); st: [N], todo: [2N] int scratch;
// left: [1] the pixels left invalid (none is known around them)
void launch_median_fill(hipStream_t s, double* depth, int W, int H, int* st, int* todo, int* left);
// the colour-guided fill of invalid regions of at most max_region (<= 63)
// pixels, in place; bgr: [3][N] planes in [0, 1]; label, st, cmax: [N],
// todo: [2N] int scratch; left: [1] the pixels of small regions left invalid
void launch_colour_fill(hipStream_t s, double* depth, const double* bgr, int W, int H, int max_region, int* label,
                        int* st, int* todo, int* cmax, int* left);

}  // namespace rgbdflow
