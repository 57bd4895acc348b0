/*
 * rgbdflow.h -- C ABI of librgbdflow, the MI355X (gfx950) scene-flow solver.
 *
 * The drop-in C++ layer (include/rgbdflow/OpticalFlow.h) calls HIP only
 * through these functions.  Plain pointers and sizes; no exceptions cross the
 * ABI; every function returns RGBDFLOW_OK (0) or a negative status, with a
 * message available from rgbdflow_last_error().
 *
 * Layouts follow the reference's DImage: row-major, channels interleaved per
 * pixel ([N][C], N = W*H).  Colour is BGR in [0,1] (ceLiuOpticalFlow.cpp:19-34),
 * depth in metres with <= 0 invalid.
 *
 * Threading: one context per host thread (calls on one context are not
 * re-entrant).  All device work of a context runs on its own HIP stream.
 */
#ifndef RGBDFLOW_H
#define RGBDFLOW_H

#ifdef __cplusplus
extern "C" {
#endif

#define RGBDFLOW_OK 0
#define RGBDFLOW_ERR_ARG (-1)        /* invalid argument (null pointer, bad size, bad enum) */
#define RGBDFLOW_ERR_HIP (-2)        /* a HIP runtime call failed */
#define RGBDFLOW_ERR_BUDGET (-3)     /* the request exceeds the context's HBM budget */
#define RGBDFLOW_ERR_STATE (-4)      /* call order violated (e.g. run before load) */
#define RGBDFLOW_ERR_UNSUPPORTED (-5)

/* OpticalFlow::sceneFlowRegularizationType (include/celiu_flow/OpticalFlow.h:99-104) */
#define RGBDFLOW_REG_ISOTROPIC 0
#define RGBDFLOW_REG_LETOUZEY 1
#define RGBDFLOW_REG_HPB 2

/* OpticalFlow::sceneFlowVariant (OpticalFlow.h:93-98); only FULL is solved
 * (sceneFlowMultiscale.cpp:169-185). */
#define RGBDFLOW_VARIANT_NONROBUST 0
#define RGBDFLOW_VARIANT_LINEARIZED 1
#define RGBDFLOW_VARIANT_FULL 2

/* Relaxation kernels (same iteration, different schedules). */
#define RGBDFLOW_RELAX_PLAIN 0   /* one launch per colour per iteration */
#define RGBDFLOW_RELAX_TILED 1   /* several iterations per launch in LDS tiles */

/* Relaxation exits (params.inner_exit), tested at iterations 24, 49, ... of
 * every (outer, inner) as the reference checks (.cu:905-922). */
#define RGBDFLOW_EXIT_NONE 0      /* run all nSOR iterations */
#define RGBDFLOW_EXIT_REFERENCE 1 /* max |change| over the iteration < 1e-5 or >= 0.3 (.cu:914-921) */
#define RGBDFLOW_EXIT_RESIDUAL 2  /* relative residual ||r - M du||_2 / ||r||_2 <= residual_tol, reduced
                                     in f64 on the device (not in the reference: a precision-study
                                     stopping test that is the same for every relaxation factor and
                                     storage) */

/* Order of a pixel's unknowns within a colour pass (params.pixel_update).
 * JACOBI is the reference's (.cu:431-459): u, v and w are each computed from
 * the others' old values; with a factor above 1 that iteration diverges on
 * these systems.  GS updates u, then v from the new u, then w from the new u
 * and v: with the red-black pixel order this is point SOR, which converges
 * for any factor in (0, 2) (the systems are symmetric definite) to the same
 * fixed point.  BLOCK solves the pixel's three unknowns together from its
 * 3x3 block (x = B^-1 (rhs - neighbour terms), the inverse from the adjugate
 * in f32), then relaxes and clamps as the reference: block SOR, convergent
 * for any factor in (0, 2), to the same fixed point.  BLOCK needs f32 system
 * records (RGBDFLOW_ERR_UNSUPPORTED otherwise: rounded to 16 bits many of
 * these nearly singular blocks are no longer definite); a pixel whose stored
 * block is not negative definite with a finite inverse keeps its values in
 * that pass.  GS and BLOCK are not in the reference. */
#define RGBDFLOW_PIXEL_JACOBI 0
#define RGBDFLOW_PIXEL_GS 1
#define RGBDFLOW_PIXEL_BLOCK 2

/* Storage of the per-pixel systems the relaxation reads.  F32 is the
 * reference's precision; F16 (with a per-pixel power-of-two scale) and BF16
 * halve the bytes the relaxation loads and are precision variants: the
 * arithmetic stays f32, the results differ from the reference's. */
#define RGBDFLOW_STORE_F32 0
#define RGBDFLOW_STORE_F16 1
#define RGBDFLOW_STORE_BF16 2

/* Storage of the feature images of a level (frame 1, frame 2 and the warped
 * frame 2; 5 channels each: gray, its x/y derivatives, G-B, G-R), which the
 * per-outer stage reads.  F64 is the reference's precision; F16 and BF16
 * (the value rounded to f32, then to 16 bits, nearest-even) are precision
 * variants that quarter the bytes: the stage arithmetic stays f64 on the
 * loaded values.  Depth, flow and weights keep their precision. */
#define RGBDFLOW_IMAGES_F64 0
#define RGBDFLOW_IMAGES_F16 1
#define RGBDFLOW_IMAGES_BF16 2

/* rgbd::CameraParams (include/rgbd_util/CameraParams.h:17-28) */
typedef struct rgbdflow_camera {
  unsigned int xRes, yRes;
  float centerX, centerY, focalLength;
  float stereoBaseline;
  float fovX, fovY;
} rgbdflow_camera;

/* OpticalFlow::sceneFlowParams (OpticalFlow.h:105-129) minus its cv::Mat
 * members (depth maps and sigmas are passed as arguments), plus schedule
 * options that do not change the iteration. */
typedef struct rgbdflow_params {
  int variant;                   /* RGBDFLOW_VARIANT_* (FULL) */
  int regularization_type;       /* RGBDFLOW_REG_* */
  rgbdflow_camera cam;
  double color_data_weight;
  double depth_vs_color_data_weight;
  double alpha;                  /* smoothness weight */
  double gamma;                  /* zero-flow prior weight */
  double pyramid_ratio;
  int min_width;
  int n_outer, n_inner, n_sor;
  /* extensions */
  int relax_variant;             /* RGBDFLOW_RELAX_* */
  int inner_exit;                /* RGBDFLOW_EXIT_*: relaxation exits at every 25th iteration */
  int outer_exit;                /* 1: outer early exit as the reference (.cu:943-951) */
  int tile_iters;                /* iterations per launch of the tiled relaxation (0 = planner) */
  int tile_w, tile_h;            /* interior tile of the tiled relaxation (0 = planner) */
  int tile_config;               /* workgroup shape of the tiled relaxation (-1 = planner) */
  int graph;                     /* 1: batch runs replay their schedule as a captured HIP graph
                                    (captured at the second run with the same batch and params) */
  int storage;                   /* RGBDFLOW_STORE_*: storage of the per-pixel systems */
  float omega;                   /* relaxation factor (<= 0: the reference's 0.7, .cu:461) */
  float residual_tol;            /* RGBDFLOW_EXIT_RESIDUAL: stop once the relative residual is <= this
                                    (<= 0: 1e-4) */
  int pixel_update;              /* RGBDFLOW_PIXEL_*: order of a pixel's three unknowns */
  int image_storage;             /* RGBDFLOW_IMAGES_*: storage of the feature images (pairs solved
                                    from frames; the single-level entry points, which take f64
                                    images, accept F64 only) */
  /* Coarse-grid correction of the relaxation (not in the reference; 0 levels: off, the
   * default).  After iteration coarse_at of every 25 red-black iterations, one V-cycle
   * over up to coarse_levels levels of 2x2 aggregates of the level's system (a level is
   * made while the finer one is at least 4x4): the finer level's residual summed into
   * the coarse rhs, coarse_sweeps iterations of the same relaxation before and after the
   * correction from the next coarser level, the coarse unknowns added to their 2x2 fine
   * pixels.  Coarse edges are the mean of the fine edges crossing between aggregates;
   * coarse blocks sum the fine data terms.  The relaxation's exits and step limit are
   * unchanged. */
  int coarse_levels;             /* 0..8 */
  int coarse_sweeps;             /* >= 1 (default 2) */
  int coarse_at;                 /* 0..23 (default 9) */
} rgbdflow_params;

typedef struct rgbdflow_stats {
  int n_levels;
  int width[16], height[16];
  int outer_exec[16];            /* outer iterations run per level */
  int sweeps_exec[16];           /* red-black iterations run per level (all outers) */
  double last_stat[16];          /* last outer convergence statistic per level (-1: none) */
  double device_ms;              /* device time of the call (HIP events) */
  double relax_ms;               /* device time of the relaxation launches, if timed: HIP events
                                    around each run of back-to-back launches (one inner
                                    iteration's; each launch where residual checks interleave) */
  int relax_launches;            /* relaxation launches timed */
  int residual_checks[16];       /* FP64 residual evaluations per level (RGBDFLOW_EXIT_RESIDUAL) */
  double residual[16];           /* last relative residual per level (-1: none evaluated) */
  double energy[16];             /* energy r'du - 1/2 du'M du of the last system at that evaluation */
  double stat_margin[16];        /* min over the level's outer iterations of |stat - 4e-4| / 4e-4:
                                    how near the early-exit statistic (.cu:948-951, summed in f64
                                    here, in f32 by the reference) came to the threshold (-1: no test) */
} rgbdflow_stats;

typedef struct rgbdflow_ctx rgbdflow_ctx;

/* Defaults of OpticalFlow::sceneFlowParams() (OpticalFlow.h:107-108). */
int rgbdflow_default_params(rgbdflow_params* p);

/* Context: owns a HIP stream and pooled device buffers (reused across calls).
 * hbm_budget_bytes = 0 means no limit. */
int rgbdflow_create(rgbdflow_ctx** out, int device, unsigned long long hbm_budget_bytes);
int rgbdflow_destroy(rgbdflow_ctx* ctx);
const char* rgbdflow_last_error(const rgbdflow_ctx* ctx);
unsigned long long rgbdflow_bytes_in_use(const rgbdflow_ctx* ctx);
int rgbdflow_synchronize(rgbdflow_ctx* ctx);
/* Time the relaxation launches with HIP events (stats.relax_ms / relax_launches); the
 * timed runs launch every kernel (no graph replay). */
int rgbdflow_set_relax_timing(rgbdflow_ctx* ctx, int on);

/* Level count and sizes of GaussianPyramid::ConstructPyramid (GaussianPyramid.cpp:20-63);
 * widths/heights hold 16 entries.  Returns the level count (>= 1) or < 0. */
int rgbdflow_pyramid_levels(int W, int H, double ratio, int min_width, int* widths, int* heights);

/* sceneFlowSingleScale (OpticalFlow.h:17-22, src/sceneFlowSingleScaleJacobiCUDA.cu:660-969):
 * one pyramid level.  Im1/Im2/warpIm2: [N][5] features; dm1/dm2/warpDm2,
 * u/v/w (in/out), depthData/depthSmoothness/depthMag weights: [N];
 * regularizationWeights: [N][2].  focal_level = camParams.focalLength of the level.
 * opts may be NULL (plain relaxation, reference exits). */
int rgbdflow_solve_level(rgbdflow_ctx* ctx, int W, int H, float focal_level, const double* Im1, const double* dm1,
                         const double* Im2, const double* dm2, const double* warpIm2, const double* warpDm2,
                         double* u, double* v, double* w, double colorDataTermWeight,
                         const double* depthDataTermWeights, double alpha, const double* depthSmoothnessTermWeights,
                         double gamma, const double* depthMagTermWeights, const double* regularizationWeights,
                         int nOuter, int nInner, int nSOR, int allowEarlyTermination, const rgbdflow_params* opts,
                         rgbdflow_stats* stats);

/* Depth hole fill on the device: the parallel variant of recursiveMedianFilter
 * (recursiveMedianFiltering.cpp:34-100, halfwidth 1; sceneFlowMultiscale.cpp:54,61
 * applies it to both depth maps).  The reference fills one invalid pixel
 * (depth <= 0) at a time, always one with the most known 3x3 neighbours,
 * with the upper median of their values rounded to float, in an order its
 * hash container decides; here every pixel with the most known neighbours
 * is filled at once, round after round.  Equal to the reference wherever
 * no two pixels filled in one round are neighbours (isolated holes),
 * otherwise a deterministic order of its own.  depth, out: [N] metres
 * (may be the same array); unfilled (may be NULL): the pixels left invalid
 * because no known pixel surrounds them (the reference asserts there). */
int rgbdflow_fill_depth_holes(rgbdflow_ctx* ctx, const double* depth, int W, int H, double* out, int* unfilled);

/* rgbdflow_fill_depth_holes on device memory, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  depth, out: [N] metres in device memory (may be the
 * same array); unfilled: one device int that receives the count of pixels
 * left invalid, or NULL.  stream: a hipStream_t, or NULL for the context's
 * stream; the call only enqueues (no host synchronisation), and the result
 * is the same as rgbdflow_fill_depth_holes's, bit for bit.  The fill's
 * scratch (per-pixel state and the lists of unfilled pixels) is one
 * fixed array per device that the library holds until the process ends,
 * so N = W*H is limited to 1920*1088 (larger frames: RGBDFLOW_ERR_ARG).
 * Every call on a device shares that array: calls that overlap on one
 * device (on different streams, or from several threads) race. */
int rgbdflow_fill_depth_holes_noalloc(rgbdflow_ctx* ctx, void* stream, const double* depth, int W, int H,
                                      double* out, int* unfilled);

/* Colour-guided fill of small holes on the device: the parallel variant of
 * recursiveMedianFilterSmallInvalidRegionsUsingColor
 * (recursiveMedianFiltering.cpp:127-222, halfwidth 1;
 * sceneFlowMultiscale.cpp:78-83 applies it, max_region 60, to the depth the
 * boundary maps' depth pyramid is built from).  Only 4-connected invalid
 * regions of at most max_region (<= 63) pixels are filled, each pixel with
 * the mean of its known 3x3 neighbours weighted by
 * exp(-|colour difference|^2 / 0.6); the reference goes region by region,
 * pixel by pixel in hash-container order, here every region's pixels with the
 * region's most known neighbours are filled at once, round after round.
 * bgr: [N][3] in [0, 1]; depth, out: [N] metres; unfilled (may be NULL): the
 * pixels of small regions left invalid (no known neighbour). */
int rgbdflow_fill_small_holes_colour(rgbdflow_ctx* ctx, const double* depth, const double* bgr, int W, int H,
                                     int max_region, double* out, int* unfilled);

/* computeDepthMapStdevsPointwise (staticDepthNoiseModeling.h:40, .cpp:225-258)
 * with sigma(1 m) = stereoError(1).  depth, stdevs: [N] float. */
int rgbdflow_depth_stdevs_pointwise(rgbdflow_ctx* ctx, const float* depth, int W, int H, int max_dist,
                                    float* stdevs);

/* OpticalFlow::sceneFlowMultiscale (OpticalFlow.h:133, sceneFlowMultiscale.cpp:30-201),
 * split so that inputs can stay resident in HBM between calls:
 *   pair_load  copies one frame pair to the device: im1/im2 [N][3] BGR,
 *              dm1/dm2 [N] hole-filled depth (recursiveMedianFilter applied),
 *              sigma1 [N] depth stdevs of frame 1 (params.depthSigma1), or
 *              NULL: computed on the device from dm1 by
 *              computeDepthMapStdevsPointwise(maxDist 4) at every run,
 *              hpb: boundary probabilities of every level, [H_k][W_k][2]
 *              concatenated finest first; NULL with RGBDFLOW_REG_HPB: made on
 *              the device by rgbdflow_heuristic_pb of each level's colour and
 *              depth-map pyramid levels (sceneFlowMultiscale.cpp:116; dm1 as
 *              given stands for the reference's colour-filled raw depth, :78-83).
 *              W, H must equal params->cam.xRes, yRes; the pyramid shape
 *              (pyramid_ratio, min_width) is fixed by the params given here.
 *   pair_run   solves the loaded pair on the device.  stats != NULL waits for
 *              completion and fills them; stats == NULL returns after enqueueing.
 *   pair_read  copies the flow (u, v in pixels, w in metres; [N] each) back. */
int rgbdflow_pair_load(rgbdflow_ctx* ctx, const rgbdflow_params* params, int W, int H, const double* im1,
                       const double* dm1, const double* im2, const double* dm2, const float* sigma1,
                       const float* hpb);
int rgbdflow_pair_run(rgbdflow_ctx* ctx, const rgbdflow_params* params, rgbdflow_stats* stats);
int rgbdflow_pair_read(rgbdflow_ctx* ctx, double* u, double* v, double* w);
/* load + run + read */
int rgbdflow_multiscale(rgbdflow_ctx* ctx, int W, int H, const double* im1, const double* dm1, const double* im2,
                        const double* dm2, const float* sigma1, const float* hpb, const rgbdflow_params* params,
                        double* u, double* v, double* w, rgbdflow_stats* stats);

/* Batched sceneFlowMultiscale: n independent frame pairs, of any sizes,
 * solved together.  Every kernel launch of a batch serves all of its pairs
 * (the same step of each pair's coarse-to-fine schedule; pyramids are aligned
 * on their finest level), and each pair keeps its own early exits, so the
 * results equal n separate pair_run calls bit for bit.  The reference solves
 * one pair per call (runFlowTwoFrames.cpp:114-126); this entry has no
 * counterpart there and is the throughput path for callers with many pairs.
 *   batch_load  copies the pairs to the device (replaces any loaded pair or
 *               batch).  Per pair, as pair_load; cam.xRes == 0 takes
 *               params->cam, min_width <= 0 takes params->min_width.
 *   batch_run   solves every loaded pair.  stats: NULL (returns after
 *               enqueueing) or n entries (waits; device_ms is the batch's).
 *   batch_read  copies pair i's flow back. */
typedef struct rgbdflow_pair_input {
  int W, H;
  const double* im1;   /* [N][3] BGR in [0,1] */
  const double* dm1;   /* [N] hole-filled depth, metres */
  const double* im2;
  const double* dm2;
  const float* sigma1; /* [N] depth stdevs of frame 1, or NULL (computed on the device) */
  const float* hpb;    /* boundary maps of every level, or NULL (HPB: made on the device) */
  rgbdflow_camera cam; /* camera of this pair (xRes == 0: params->cam) */
  int min_width;       /* pyramid min width (<= 0: params->min_width) */
  int reserved;
  const double* est;   /* [N][3] flow estimate, or NULL.  Given: the pair is solved as
                          OpticalFlow::sceneFlowWithInit (sceneFlowWithInit.cpp:7-81, as
                          rgbdflow_with_init) -- full resolution only, from est, no outer
                          early exit; hpb then needs only its level-0 map */
  const double* hpb_dm1; /* [N] the depth that boundary maps made on the device start from (the
                            reference's colour-filled raw depth, sceneFlowMultiscale.cpp:78-83, e.g.
                            rgbdflow_fill_small_holes_colour of the depth as measured), or NULL: dm1 */
} rgbdflow_pair_input;
int rgbdflow_batch_load(rgbdflow_ctx* ctx, const rgbdflow_params* params, int n, const rgbdflow_pair_input* pairs);
int rgbdflow_batch_run(rgbdflow_ctx* ctx, const rgbdflow_params* params, rgbdflow_stats* stats);
int rgbdflow_batch_read(rgbdflow_ctx* ctx, int i, double* u, double* v, double* w);

/* OpticalFlow::sceneFlowWithInit (OpticalFlow.h:145, sceneFlowWithInit.cpp:7-81):
 * full resolution only, from the estimate est [N][3], no early termination.
 * hpb0: level-0 map [H][W][2] (NULL: made on the device from im1, dm1 as
 * sceneFlowWithInit.cpp:51 does). */
int rgbdflow_with_init(rgbdflow_ctx* ctx, int W, int H, const double* im1, const double* dm1, const double* im2,
                       const double* dm2, const float* sigma1, const float* hpb0, const double* est,
                       const rgbdflow_params* params, double* u, double* v, double* w, rgbdflow_stats* stats);

/* ---- stage entry points: each runs one kernel of the path on the device
 * (used by the per-stage parity tests and available to callers that build
 * their own pipeline) ---- */

/* GaussianPyramid::ConstructPyramid of a [N][C] image; out: levels concatenated
 * finest first, each [N_k][C]. */
int rgbdflow_gaussian_pyramid(rgbdflow_ctx* ctx, const double* img, int W, int H, int C, double ratio, int min_width,
                              double* out);
/* DepthMapPyramid::ConstructPyramid (depthMapPyramid.cpp:15-140); out concatenated. */
int rgbdflow_depth_map_pyramid(rgbdflow_ctx* ctx, const double* depth, int W, int H, double ratio, int min_width,
                               double* out);
/* computeHeuristicPb (src/heuristicProbBoundary.cpp:21-41 -> src/heuristicBoundaries.cpp:40-358)
 * of one image: bgr [N][3] in [0,1], depth [N] metres (<= 0 invalid) -> out
 * [N][2] f32 boundary probabilities (right, down neighbour), the map
 * OpticalFlow::calculateRegularizationWeights turns into HPB weights.
 * Normals use an f64 Jacobi eigen-solver, not PCL's: see the kernels. */
int rgbdflow_heuristic_pb(rgbdflow_ctx* ctx, const double* bgr, const double* depth, int W, int H, float* out);
/* OpticalFlow::im2feature RGB_WITH_GRADIENT (OpticalFlow.cpp:148-220): [N][3] -> [N][5] */
int rgbdflow_im2feature(rgbdflow_ctx* ctx, const double* bgr, int W, int H, double* out5);
/* warpBy3DFlow (opticalFlowUtils.cpp:76-99): I [N][C], D/u/v/w [N] */
int rgbdflow_warp_by_3d_flow(rgbdflow_ctx* ctx, const double* I1, const double* D1, const double* I2,
                             const double* D2, const double* u, const double* v, const double* w, int W, int H, int C,
                             double* warpI2, double* warpD2);
/* Image::imresize(w,h) of u, v, w, then u, v *= uv_scale (sceneFlowMultiscale.cpp:130-134) */
int rgbdflow_upsample_flow(rgbdflow_ctx* ctx, const double* u, const double* v, const double* w, int Ws, int Hs,
                           int Wd, int Hd, double uv_scale, double* uo, double* vo, double* wo);
/* The compact system of one outer iteration (13 planes of N values, f32 values
 * widened to f64): eRxy, eRz, eDxy, eDz, Buu, Bvv, Bww, Buv, Buw, Bvw, ru, rv, rw.
 * Arguments as rgbdflow_solve_level. */
int rgbdflow_level_system(rgbdflow_ctx* ctx, int W, int H, float focal_level, const double* Im1, const double* dm1,
                          const double* Im2, const double* dm2, const double* warpIm2, const double* warpDm2,
                          const double* u, const double* v, const double* w, double colorDataTermWeight,
                          const double* depthDataTermWeights, double alpha, const double* depthSmoothnessTermWeights,
                          double gamma, const double* depthMagTermWeights, const double* regularizationWeights,
                          double* coef13);
/* n_iter red-black iterations (red then black) on a compact system; x: [3][N]
 * planes in/out.  schedule: relax_variant, tile_*, omega, pixel_update and
 * storage (the records re-stored as the level solve stores them) of params
 * (NULL: plain, 0.7, the reference's pixel order, f32). */
int rgbdflow_relax(rgbdflow_ctx* ctx, const double* coef13, int W, int H, int n_iter,
                   const rgbdflow_params* schedule, double* x);
/* Relaxes a compact system from x until the relative residual
 * ||r - M x||_2 / ||r||_2, reduced in f64 on the device at iterations 24, 49,
 * ... (RGBDFLOW_EXIT_RESIDUAL), is <= tol, or for max_iter iterations.  Out:
 * x, the iterations run, the last residual evaluated (-1: none) and the
 * system's energy r'x - 1/2 x'Mx at that evaluation (convex: -M is positive
 * definite, and the solution minimises it).  No counterpart in the
 * reference (whose check is the max change, .cu:914-921): the precision
 * study's measure of convergence, the same for every relaxation factor. */
int rgbdflow_relax_residual(rgbdflow_ctx* ctx, const double* coef13, int W, int H, int max_iter, double tol,
                            const rgbdflow_params* schedule, double* x, int* iters, double* residual,
                            double* energy);

/* ---- autotuning ---- */
typedef struct rgbdflow_tile_choice {
  int tile_w, tile_h;            /* interior tile */
  int tile_iters;                /* red-black iterations per launch */
  int tile_config;               /* workgroup shape index */
  int threads, pairs_per_thread; /* the workgroup shape */
  int single;                    /* 1: one tile covers the level */
  int candidates;                /* schedules timed */
  double us_per_iter;            /* measured device time per red-black iteration */
  int streamed;                  /* out: > 0 for a streamed workgroup shape (the tile swept through a ring
                                    of rows; up to this many iterations per launch), 0 otherwise */
} rgbdflow_tile_choice;

/* Workgroup shape `index` of the tiled relaxation (the tile_config values):
 * threads, pixel pairs per thread, and for a streamed shape (the tile swept
 * top to bottom through a ring of region rows in LDS, one wave per ring-row
 * pair) the most iterations per launch, else 0.  RGBDFLOW_ERR_ARG past the
 * last shape.  No counterpart in the reference (which has one schedule). */
int rgbdflow_relax_shape(int index, int* threads, int* pairs_per_thread, int* streamed);

/* Times candidate schedules of the tiled relaxation (workgroup shape, tile
 * shape, iterations per launch) on `batch` W x H levels relaxed by one launch
 * each, with systems stored as params->storage (params may be NULL: f32), and
 * makes the fastest the context's schedule for that shape whenever tile_* are
 * left to the planner (tile_iters = tile_w = tile_h = 0, tile_config = -1):
 * for batches of that record storage, and for the others until they have a
 * schedule tuned with their own storage.  out may be NULL. */
int rgbdflow_autotune_relax(rgbdflow_ctx* ctx, int W, int H, int batch, const rgbdflow_params* params,
                            rgbdflow_tile_choice* out);
/* Makes a schedule (as autotune_relax reports it: tile_w, tile_h, tile_iters,
 * tile_config, single) the context's schedule for W x H levels, every record
 * storage, e.g. one tuned earlier and stored by the caller. */
int rgbdflow_set_relax_schedule(rgbdflow_ctx* ctx, int W, int H, const rgbdflow_tile_choice* choice);
/* Device time per red-black iteration of a schedule on `batch` W x H levels,
 * measured as autotune_relax measures its candidates. */
int rgbdflow_time_relax_schedule(rgbdflow_ctx* ctx, int W, int H, int batch, const rgbdflow_params* params,
                                 const rgbdflow_tile_choice* choice, double* us_per_iter);

/* ---- serving: continuous batching of RGB-D frame streams ----
 * No counterpart in the reference, which solves one pair per call
 * (runFlowTwoFrames.cpp:114-126).  A server owns `lanes` executor contexts
 * and a worker thread.  Each stream keeps its last frame on the device; every
 * submitted frame after the first forms the pair (last frame, frame), solved
 * by sceneFlowMultiscale exactly as rgbdflow_multiscale would (depth stdevs
 * computed on the device; holes filled only with fill_holes), in whatever
 * batch the worker puts it:
 * the result does not depend on the batch.  With warm_start, every pair after
 * a stream's first is instead solved as rgbdflow_with_init would solve it
 * with the stream's previous result as the estimate. */
typedef struct rgbdflow_server rgbdflow_server;
typedef struct rgbdflow_server_config {
  int device;
  int max_batch;                  /* pairs per batch (<= 0: 16) */
  int lanes;                      /* batches in flight at once, each on its own HIP stream (<= 0: 2) */
  int warm_start;                 /* 1: a stream's pairs after its first are solved as
                                     sceneFlowWithInit from the stream's last flow (kept on the
                                     device), as ceLiuSceneFlow does when given an estimate
                                     (ceLiuOpticalFlow.cpp:200-204) */
  unsigned long long hbm_budget;  /* bytes for stream state and in-flight pairs (0: no limit) */
  int warm_n_outer;               /* outer iterations of warm-started pairs (<= 0: params' n_outer) */
  int fill_holes;                 /* 1: each frame's depth is hole-filled on the device by
                                     rgbdflow_fill_depth_holes's fill before it is used, as
                                     sceneFlowMultiscale.cpp:54,61 fills both depth maps; the depth
                                     stdevs still come from the depth as submitted
                                     (runFlowTwoFrames.cpp:89-91) */
} rgbdflow_server_config;
typedef struct rgbdflow_request_info {
  int has_flow;       /* 0 for a stream's first frame */
  int status;         /* RGBDFLOW_OK or the error of the request's batch */
  int batch_pairs;    /* pairs in that batch */
  int stream;
  double queue_ms;    /* submit -> batch start */
  double batch_ms;    /* batch start -> batch done */
  double latency_ms;  /* submit -> result available */
} rgbdflow_request_info;
int rgbdflow_server_create(rgbdflow_server** out, const rgbdflow_server_config* cfg, const rgbdflow_params* params);
int rgbdflow_server_destroy(rgbdflow_server* srv);
const char* rgbdflow_server_last_error(const rgbdflow_server* srv);
/* cam NULL: params->cam; min_width <= 0: params->min_width */
int rgbdflow_stream_open(rgbdflow_server* srv, int W, int H, const rgbdflow_camera* cam, int min_width,
                         int* stream_id);
/* frame: bgr [N][3] u8 (cv2liuBGR: c / 255.0, no channel swap), depth [N] f32 metres
 * (cv2liuDepth, no fill), hpb: the frame's boundary maps of every level, or
 * NULL (with RGBDFLOW_REG_HPB they are then made on the device when the frame
 * is a pair's frame 1).  Copies the arrays; returns a ticket. */
int rgbdflow_stream_submit(rgbdflow_server* srv, int stream_id, const unsigned char* bgr, const float* depth,
                           const float* hpb, unsigned long long* ticket);
/* 1: done (u, v, w [N] copied when non-NULL and info->has_flow; the ticket is
 * released), 0: not done within wait_ms (< 0: wait for it), < 0: error. */
int rgbdflow_stream_poll(rgbdflow_server* srv, unsigned long long ticket, int wait_ms, double* u, double* v,
                         double* w, rgbdflow_request_info* info);
/* waits for the stream's requests, then frees its state */
int rgbdflow_stream_close(rgbdflow_server* srv, int stream_id);
/* rgbdflow_set_relax_schedule on every lane (waits until no batch is in flight) */
int rgbdflow_server_set_relax_schedule(rgbdflow_server* srv, int W, int H, const rgbdflow_tile_choice* choice);

#ifdef __cplusplus
}
#endif
#endif
