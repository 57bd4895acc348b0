// internal.hpp -- device-side data layouts and kernel launchers of librgbdflow.
//
// Every image lives in HBM as structure-of-arrays planes: channel c of an
// image with N pixels is the contiguous range [c*N, (c+1)*N), row-major.
// Level buffers are sized for the finest level and reused by coarser ones.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace rgbdflow {

constexpr int kFeat = 5;            // feature channels: gray, d/dx, d/dy, G-B, G-R
constexpr int kCoef = 13;           // compact per-pixel system, see Coef below
constexpr int kMaxLevels = 16;
constexpr int kCheckEvery = 25;     // relaxation convergence check period (.cu:905)
constexpr int kMaxChecks = 4096;    // check slots per level solve
constexpr int kMaxOuter = 1024;

// Plane indices of the compact system of one pixel p:
//   right/down edge weights alpha*phi_p*r_p (xy rows) and alpha*phi_p*r_p*s_p (z row),
//   the symmetric 3x3 block B (diagonal includes gamma terms) and the rhs.
enum Coef { kERxy = 0, kERz, kEDxy, kEDz, kBuu, kBvv, kBww, kBuv, kBuw, kBvw, kRu, kRv, kRw };

// Device control block of one level solve.  Kernels read it at entry to skip
// work after an early exit, so the host never waits inside a level.
struct SolveCtl {
  int outer_done;                     // outer loop left early (stat < threshold)
  int outer_exit_at;                  // outer iteration that set outer_done
  int sweeps_exec;                    // red-black iterations executed (all outers)
  int pad;
  unsigned int maxdiff[kMaxChecks];   // float bits of max|du change| per check of the
                                      // current outer, [inner][check]; zeroed per outer
  double stat[kMaxOuter];             // outer convergence statistic per outer iteration
};

// Everything one level-solve kernel needs (passed by value).
struct LevelView {
  int W, H, N;
  const double* I1;     // [5][N] features of frame 1
  const double* I2;     // [5][N] features of frame 2
  double* Iw;           // [5][N] frame-2 features warped by u
  const double* D1;     // [N]
  const double* D2;     // [N]
  double* Dw;           // [N] frame-2 depth warped by u, minus w
  double* u;            // [3][N] flow (u px, v px, w m)
  float* du;            // [3][N] increment solved by the relaxation
  float* du_alt;        // [3][N] second buffer of the tiled relaxation (reads one, writes the other)
  const float* dW;      // [N] depth-data weights (delta)
  const float* sW;      // [N] depth-smoothness weights
  const float* mW;      // [N] depth-magnitude weights
  const float* rW;      // [2][N] regularisation weights (right, down)
  const float* cW;      // [N] convergence weights f/D1
  float* q;             // [N] occlusion weights
  double* Ix;           // [5][N]
  double* Iy;           // [5][N]
  double* It;           // [5][N]
  double* Dx;           // [N]
  double* Dy;           // [N]
  double* Dt;           // [N]
  double* phi;          // [N]
  float* coef;          // [13][N]
  double* partials;     // per-block partial sums of the outer statistic
  SolveCtl* ctl;
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

struct SolveScalars {
  double alpha, gamma, color_weight;
  int n_outer, n_inner, n_sor;
  int allow_early;          // outer early exit allowed on this level
  int inner_exit;           // relaxation early exits enabled
  int outer_exit;           // outer early exits enabled
};

// ---------------------------------------------------------------- launchers
// kernels_solve.hip
void launch_ctl_reset(hipStream_t s, SolveCtl* ctl);
void launch_derivs(hipStream_t s, const LevelView& v, int n_slots);
void launch_phi(hipStream_t s, const LevelView& v, int outer);
void launch_system(hipStream_t s, const LevelView& v, const SolveScalars& k, int outer);
void launch_relax_plain(hipStream_t s, const LevelView& v, int colour, int outer, int slot_base,
                        int n_prev_checks, int check_slot);
void launch_update_warp(hipStream_t s, const LevelView& v, const SolveScalars& k, int outer);
int stat_blocks(int N);

// kernels_relax_tiled.hip
int tiled_config_count();
void tiled_config_shape(int idx, int* nt, int* p);
size_t tiled_config_lds(int idx);
void launch_relax_tiled(hipStream_t s, const LevelView& v, const TileSched& ts, const float* src, float* dst,
                        int n_iter, int slot_base, int n_prev_checks, int check_slot);
void launch_copy_f32(hipStream_t s, const float* src, size_t n, float* dst);
void launch_fill_test_system(hipStream_t s, float* coef, float* du, int W, int H);

// kernels_pyramid.hip
void launch_deinterleave(hipStream_t s, const double* src, int N, int C, double* dst);
void launch_interleave(hipStream_t s, const double* src, int N, int C, double* dst);
void launch_fill_f64(hipStream_t s, double* dst, size_t n, double v);
void launch_to_f32(hipStream_t s, const double* src, size_t n, float* dst);
void launch_pyramid_level(hipStream_t s, const double* src, int Ws, int Hs, int C, const double* taps,
                          int fsize, double ratio, double* dst, int Wd, int Hd);
void launch_features(hipStream_t s, const double* bgr, int W, int H, double* out5);
void launch_upsample(hipStream_t s, const double* src, int Ws, int Hs, double* dst, int Wd, int Hd,
                     double uv_scale);
void launch_warp_cpu_style(hipStream_t s, const LevelView& v);
void launch_level_weights(hipStream_t s, const double* delta, const double* smooth, const float* hpb,
                          int reg_type, const double* D1, int W, int H, int W_full, float f_full,
                          float f_level, float* dW, float* sW, float* mW, float* rW, float* cW);
void launch_level_weights_f64(hipStream_t s, const double* delta, const double* smooth, const double* mag,
                              const double* reg2, const double* D1, int W, int H, float f_level, float* dW,
                              float* sW, float* mW, float* rW, float* cW);
void launch_depth_sigma(hipStream_t s, const float* depth, int W, int H, int max_dist, float* sigma);
void launch_depth_data_weights(hipStream_t s, const float* sigma, int N, double dvc, double* delta);
void launch_depth_map_level(hipStream_t s, const double* larger, int LW, int LH, double* dst, int W, int H);

}  // namespace rgbdflow
