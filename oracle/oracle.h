/*
 * oracle.h -- C ABI of the CPU restatement of rgbd-flow's scene-flow path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * Layout conventions follow the reference's DImage: row-major, channels
 * interleaved per pixel ([N][C]).  Every function cites the reference
 * file:line whose behaviour it restates.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision modes of the restatement. */
enum {
  ORC_PREC_REF = 0, /* the reference's storage: f64 images/weights-on-host, f32 A/du/weights-on-device, f32 sqr quirk */
  ORC_PREC_F64 = 1  /* every buffer and product in f64 (the FP64 oracle) */
};

enum { ORC_REG_ISOTROPIC = 0, ORC_REG_LETOUZEY = 1, ORC_REG_HPB = 2 };

typedef struct {
  unsigned int xRes, yRes;
  float centerX, centerY, focalLength, stereoBaseline, fovX, fovY;
} orc_camera;

typedef struct {
  int regularization_type;
  double color_data_weight, depth_vs_color_data_weight, alpha, gamma;
  double pyramid_ratio;
  int min_width, n_outer, n_inner, n_sor;
  orc_camera cam;
  int precision;      /* ORC_PREC_* */
  int disable_inner_exit; /* 1: never leave the relaxation loop early (parity runs) */
  int disable_outer_exit; /* 1: never leave the outer loop early */
  float omega;            /* relaxation factor (.cu:461: 0.7; <= 0 means 0.7) */
  int residual_exit;      /* 1: the relaxation's checks (every 25th iteration) stop on the f64
                             relative residual <= residual_tol instead of the max change */
  double residual_tol;
  int pixel_gs;           /* 0: a pixel's unknowns are updated from each other's old values
                             (.cu:431-459); 1: u, then v from the new u, then w from both
                             (point SOR order; convergent for any factor in (0, 2)); 2: the
                             pixel's 3x3 block solved as a whole (block SOR, likewise) */
  int image_storage;      /* ORC_IMG_*: the feature images (frame 1, frame 2, warped frame 2)
                             rounded, wherever they are produced, to f32 and then to 16 bits */
  int record_storage;     /* ORC_REC_*: the level solve's per-pixel systems rounded, right after
                             assembly, as 16-bit records hold them (reference precision only) */
  int cg_levels;          /* coarse-grid correction (not in the reference): 0 off; else a V-cycle
                             over up to this many 2x2-aggregation levels of the relaxed system
                             after iteration cg_at of every 25 (Solver::cg_vcycle) */
  int cg_sweeps;          /* red-black iterations on each coarse level before and after the
                             correction from the next coarser one */
  int cg_at;              /* iteration (0..23) of each 25 after which the V-cycle runs */
} orc_params;

enum { ORC_REC_F32 = 0, ORC_REC_F16 = 1, ORC_REC_BF16 = 2 };
/* one pixel's 13 system coefficients rounded as that record storage holds them */
void orc_record_round(float* c13, int storage);

enum { ORC_IMG_F64 = 0, ORC_IMG_F16 = 1, ORC_IMG_BF16 = 2 };
/* x rounded as a feature image of that storage holds it (nearest-even at each step) */
double orc_img_round(double x, int storage);

typedef struct {
  int n_levels;
  int width[16], height[16];
  int outer_exec[16];      /* outer iterations executed per level */
  int sweeps_exec[16];     /* red-black iterations executed per level (summed over outers) */
  double last_stat[16];    /* last outer convergence statistic per level */
  int residual_checks[16]; /* residual evaluations per level (residual_exit) */
  double residual[16];     /* last relative residual per level (-1: none) */
  double energy[16];       /* energy r'x - 1/2 x'Mx of the last system at that evaluation */
} orc_stats;

/* ---- pyramids (src/GaussianPyramid.cpp:20-63) ---- */
int orc_pyramid_levels(int W, int H, double ratio, int min_width, int* widths, int* heights);
void orc_gaussian_smooth(const double* img, int W, int H, int C, double sigma, int fsize, double* out);
void orc_resize_ratio(const double* src, int W, int H, int C, double ratio, double* dst);
void orc_resize_dims(const double* src, int Ws, int Hs, int C, int Wd, int Hd, double* dst);
/* out: all levels concatenated, level 0 first */
void orc_gaussian_pyramid(const double* img, int W, int H, int C, double ratio, int min_width, double* out);
/* src/depthMapPyramid.cpp:15-140; out: all levels concatenated */
void orc_depth_map_pyramid(const double* depth, int W, int H, double ratio, int min_width, double* out);

/* ---- features, warps, weights ---- */
void orc_im2feature(const double* bgr, int W, int H, double* out5);
void orc_warp_cpu(const double* I1, const double* D1, const double* I2, const double* D2,
                  const double* u, const double* v, const double* w, int W, int H, int C,
                  double* warpI2, double* warpD2);
void orc_depth_stdevs_pointwise(const float* depth, int W, int H, int max_dist, float* stdevs);
void orc_l1_distance_to_boundary(const float* depth, int W, int H, int cap, float* dist);
void orc_depth_data_weights(const float* sigma, int W, int H, double depth_vs_color, double* out);
void orc_depth_mag_weights(const double* depth_level, int Wl, int Hl, int Wfull, int Hfull,
                           float f_full, double* out);
void orc_reg_weights(const float* hpb, const double* depth, int W, int H, int type, double* out);
void orc_median_fill(double* depth, int W, int H, int halfwidth);
/* the device's parallel rounds variant of it (halfwidth 1); returns the pixels left invalid */
int orc_median_fill_rounds(double* depth, int W, int H);
/* the device's rounds variant of the colour-guided fill of small holes; bgr [3][N] planes */
int orc_colour_fill_rounds(double* depth, const double* bgr, int W, int H, int max_region);
double orc_fill_exp(double x);

/* ---- one level: sceneFlowSingleScale (src/sceneFlowSingleScaleJacobiCUDA.cu:660-969) ---- */
typedef struct {
  int W, H;
  float focal_level;                 /* camParams.focalLength of the level */
  const double *I1, *I2, *warpI2;    /* [N][5] */
  const double *D1, *D2, *warpD2;    /* [N] */
  const double *delta, *s, *m;       /* [N] depth-data, depth-smoothness, depth-magnitude weights */
  const double *r;                   /* [N][2] regularisation weights (right, down) */
} orc_level_in;

typedef struct {
  int outer_exec, sweeps_exec;
  double last_stat;
  float last_maxdiff;
  int residual_checks;
  double last_residual, last_energy;
} orc_level_stats;

int orc_level_solve(const orc_level_in* in, const orc_params* p, int allow_early,
                    double* u, double* v, double* w, orc_level_stats* st);

/* The system of one outer iteration, in the compact per-pixel form
 * (eRxy, eRz, eDxy, eDz, Buu, Bvv, Bww, Buv, Buw, Bvw, ru, rv, rw) as 13 planes
 * of N values, computed from (u, du=0, warpI2, warpD2).  Precision REF -> the
 * values are f32 stored in doubles. */
void orc_level_system(const orc_level_in* in, const orc_params* p, const double* u,
                      const double* v, const double* w, double* coef13);
/* The per-outer-iteration warp of the level solve (.cu:579-635): floor base,
 * explicit clamps, f32 weights in ORC_PREC_REF (f64 in ORC_PREC_F64). */
void orc_warp_gpu_style(const double* I1, const double* D1, const double* I2, const double* D2, const double* u,
                        const double* v, const double* w, int W, int H, int precision, double* warpI2,
                        double* warpD2);
/* The same warp in the f32 stage of the 16-bit image variants (image_storage
 * ORC_IMG_F16 or ORC_IMG_BF16: f32 sums, the warped features rounded to it). */
void orc_warp_gpu_style_f32(const double* I1, const double* D1, const double* I2, const double* D2, const double* u,
                            const double* v, const double* w, int W, int H, int image_storage, double* warpI2,
                            double* warpD2);
/* The robust weights at du = 0 (.cu:340-371): phi [N], psi [N][5], pi [N]. */
void orc_level_weights(const orc_level_in* in, const orc_params* p, const double* u, const double* v,
                       const double* w, double* phi, double* psi, double* pi);
/* nIter red-black iterations on a compact system; x = [3][N] planes (in/out). */
void orc_relax(const double* coef13, int W, int H, int precision, int n_iter, double* x);
/* n_iter red-black iterations (factor, pixel order and precision of p) with
 * p's coarse-grid correction, from x; per iteration (may be NULL): the
 * largest |change| of its two colour passes and the f64 relative residual
 * after it (and after the V-cycle that follows it, if one does). */
void orc_relax_cg(const double* coef13, int W, int H, const orc_params* p, int n_iter, double* x, double* maxd,
                  double* res);
/* The coarse-grid correction's parts (Solver::cg_*; not in the reference), on
 * a W x H compact system; the coarse level is ((W+1)/2) x ((H+1)/2):
 * the coarse operator (13 planes, rhs zero); the coarse rhs (3 planes) from
 * the residual at x ([3][N]); x += P xc. */
void orc_cg_operator(const double* coef13, int W, int H, int precision, double* out13);
void orc_cg_restrict(const double* coef13, int W, int H, int precision, const double* x, double* rhs3);
void orc_cg_prolong(double* x, int W, int H, int precision, const double* xc);
/* The residual of a compact system at x in f64: out[0] = ||r - M x||_2 / ||r||_2
 * (||r - M x||_2 when r = 0), out[1] = r'x - 1/2 x'Mx, where M is the operator
 * the relaxation inverts (.cu:404-459): (M x)_k = sum_n e_nk x_nk + sum_k' B_kk' x_k'. */
void orc_system_residual(const double* coef13, int W, int H, const double* x, double* out);
/* Relaxation (factor omega, pixel order as orc_params.pixel_gs, precision as
 * orc_relax) from x with the residual
 * test at iterations 24, 49, ...: stops once out[0] of orc_system_residual is
 * <= tol, or after max_iter iterations.  Out: x, iterations run, last
 * residual (-1: none) and energy. */
void orc_relax_residual(const double* coef13, int W, int H, int precision, float omega, int pixel_gs, int max_iter,
                        double tol, double* x, int* iters, double* residual, double* energy);

/* ---- whole pair: OpticalFlow::sceneFlowMultiscale (src/sceneFlowMultiscale.cpp:30-201) ----
 * im*: [N][3] BGR in [0,1]; dm*: [N] metres; sigma1: [N] depth stdevs of frame 1;
 * hpb: per level [H_k][W_k][2] concatenated (level 0 first); NULL with HPB
 * regularisation computes them (orc_heuristic_pb of each level's colour and
 * depth-map pyramid levels); returns -2 if those pyramids' sizes differ. */
int orc_multiscale(const double* im1, const double* dm1, const double* im2, const double* dm2,
                   const float* sigma1, const float* hpb, int W, int H, const orc_params* p,
                   double* u, double* v, double* w, orc_stats* st);

/* OpticalFlow::sceneFlowWithInit (src/sceneFlowWithInit.cpp:7-81); est: [N][3];
 * hpb0 NULL with HPB: orc_heuristic_pb(im1, dm1) */
int orc_with_init(const double* im1, const double* dm1, const double* im2, const double* dm2,
                  const float* sigma1, const float* hpb0, const double* est, int W, int H,
                  const orc_params* p, double* u, double* v, double* w, orc_stats* st);

/* ---- boundary probabilities: computeHeuristicPb (src/heuristicProbBoundary.cpp:21-41,
 * src/heuristicBoundaries.cpp:40-358), hpb.cpp ----
 * bgr: [N][3] level colour in [0,1]; depth: [N] metres (<= 0 invalid);
 * out: [N][2] f32 (right, down). */
void orc_heuristic_pb(const double* bgr, const double* depth, int W, int H, float* out);
/* the normals it uses: [N][3] (0 where none) and their validity flags [N] */
void orc_hpb_normals(const double* depth, int W, int H, float* normals, unsigned* valid);
float orc_hpb_exp(float x);
void orc_hpb_smallest_eigvec(const double* a9, double* n3);

/* Threads of the level solve's row partitions (1: the scalar restatement,
 * the default).  Results are identical for every count. */
void orc_set_threads(int n);
int orc_threads(void);

#ifdef __cplusplus
}
#endif
