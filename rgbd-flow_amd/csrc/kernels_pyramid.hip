// kernels_pyramid.hip -- per-pair and per-level preparation kernels on gfx950.
//
// Pyramids (src/GaussianPyramid.cpp, src/depthMapPyramid.cpp), features
// (src/OpticalFlow.cpp:148-220), the flow upsample (Image.h:764-769 +
// sceneFlowMultiscale.cpp:130-134), the level-entry warp (opticalFlowUtils.cpp:76-99)
// and the static weight maps (OpticalFlow.cpp:71-141, staticDepthNoiseModeling.cpp).
// Each kernel evaluates the reference's sums in the reference's order, so its
// f64 results are the same numbers the reference's host code computes.
#include "internal.hpp"
#include "records.hpp"

namespace rgbdflow {

namespace {


__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
// ImageProcessing.h:25 EnforceRange
__device__ __forceinline__ int enforce_range(int x, int n) { return clampi(x, 0, n - 1); }

__device__ __forceinline__ double dsq(double x) { return x * x; }
// primesensorUtils.cpp:53-64
__device__ __forceinline__ double stereo_error(double z) { return dsq(z) / (1 / .00285 + z); }
__device__ __forceinline__ double stereo_error_ratio(double z) { return stereo_error(z) / stereo_error(1); }

// ImageProcessing.h:128-148 BilinearInterpolate: truncating base, f64 weights
// clamped to [0,1]; taps (xx,yy), (xx,yy+1), (xx+1,yy), (xx+1,yy+1).
struct Bilerp {
  int o[4];
  double s[4];
};
__device__ __forceinline__ Bilerp bilerp_taps(double x, double y, int W, int H) {
  Bilerp b;
  const int xx = (int)x, yy = (int)y;
  double dx = x - xx, dy = y - yy;
  dx = dx > 1 ? 1.0 : dx;
  dx = dx < 0 ? 0.0 : dx;
  dy = dy > 1 ? 1.0 : dy;
  dy = dy < 0 ? 0.0 : dy;
  int t = 0;
  for (int m = 0; m <= 1; m++)
    for (int n = 0; n <= 1; n++) {
      const int uu = enforce_range(xx + m, W), vv = enforce_range(yy + n, H);
      b.o[t] = vv * W + uu;
      b.s[t] = fabs(1 - m - dx) * fabs(1 - n - dy);
      t++;
    }
  return b;
}

__global__ void k_deinterleave(const double* src, int N, int C, double* dst) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= N) return;
  for (int c = 0; c < C; c++) dst[(size_t)c * N + l] = src[(size_t)l * C + c];
}

__global__ void k_interleave(const double* src, int N, int C, double* dst) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= N) return;
  for (int c = 0; c < C; c++) dst[(size_t)l * C + c] = src[(size_t)c * N + l];
}

// One pyramid level (GaussianPyramid.cpp:47-61): Gaussian smoothing of the
// source (Image.h:1179-1199 -> imfilter_hv: h-pass then v-pass, clamp-to-edge,
// ImageProcessing.h:247-269,339-358) followed by ResizeImage(ratio)
// (ImageProcessing.h:203-222).  Only the smoothed samples the resize reads are
// evaluated; taps of zero bilinear weight are skipped (adding x*0 leaves the
// sum unchanged).
// At ratio 0.5 the resize reads one smoothed sample per output, (2i+1, 2j+1)
// with weight 1 (the others' weights are 0), so a block makes a 32 x 8 output
// tile of one channel in two passes: the horizontal sums at the outputs'
// columns for every source row the tile's vertical sums reach (into LDS),
// then the vertical sums -- the same sums in the same order as per sample.
__global__ void __launch_bounds__(kBlock) k_pyramid_jobs(WorkList<PyrJob> w) {
  __shared__ double th[(2 * kPyrTH + kMaxTaps - 1) * kPyrTW];
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int item = work_item(w, blk);
  const PyrJob& J = w.items[item];
  const int Ws = J.Ws, Hs = J.Hs, Wd = J.Wd, Hd = J.Hd, fs = J.fs;
  if (pyr_halving(J)) {
    const int lb = blk - w.start[item];
    const int tx = (Wd + kPyrTW - 1) / kPyrTW, nt = tx * ((Hd + kPyrTH - 1) / kPyrTH);
    const int c = lb / nt, tt = lb - c * nt;
    const int j0 = (tt % tx) * kPyrTW, i0 = (tt / tx) * kPyrTH;
    const double* s = J.src + (size_t)c * Ws * Hs;
    // rows 2 i0 - fs .. 2 (i0 + kPyrTH) - 1 + fs (each clamped as the per-sample sums clamp it)
    const int rbase = 2 * i0 - fs, nr = 2 * kPyrTH + 2 * fs;
    for (int t = threadIdx.x; t < nr * kPyrTW; t += kBlock) {
      const int rr = t / kPyrTW, jj = t - rr * kPyrTW, j = j0 + jj;
      double h = 0;
      if (j < Wd) {
        const int ii = enforce_range(rbase + rr, Hs), uu = enforce_range(2 * j + 1, Ws);
        for (int lh = -fs; lh <= fs; lh++) h += s[(size_t)ii * Ws + enforce_range(uu + lh, Ws)] * J.taps[lh + fs];
      }
      th[t] = h;
    }
    __syncthreads();
    const int ty = threadIdx.x / kPyrTW, jj = threadIdx.x - ty * kPyrTW;
    const int i = i0 + ty, j = j0 + jj;
    if (i >= Hd || j >= Wd) return;
    const int vv = enforce_range(2 * i + 1, Hs);
    double sm = 0;
    for (int lv = -fs; lv <= fs; lv++) sm += th[(vv + lv - rbase) * kPyrTW + jj] * J.taps[lv + fs];
    double acc = 0;
    acc += sm * 1.0;
    J.dst[(size_t)c * Wd * Hd + (size_t)i * Wd + j] = acc;
    return;
  }
  const int l = (blk - w.start[item]) * kBlock + threadIdx.x;
  const int Nd = Wd * Hd;
  if (l >= Nd * J.C) return;
  const int c = l / Nd;
  const int p = l - c * Nd;
  const int i = p / Wd, j = p - i * Wd;
  const double x = (double)(j + 1) / J.ratio - 1;
  const double y = (double)(i + 1) / J.ratio - 1;
  const Bilerp b = bilerp_taps(x, y, Ws, Hs);
  const double* s = J.src + (size_t)c * Ws * Hs;
  double acc = 0;
  for (int t = 0; t < 4; t++) {
    if (b.s[t] == 0.0) continue;
    const int vv = b.o[t] / Ws, uu = b.o[t] - vv * Ws;
    double sm = 0;
    for (int lv = -fs; lv <= fs; lv++) {
      const int ii = enforce_range(vv + lv, Hs);
      double th = 0;
      for (int lh = -fs; lh <= fs; lh++) th += s[(size_t)ii * Ws + enforce_range(uu + lh, Ws)] * J.taps[lh + fs];
      sm += th * J.taps[lv + fs];
    }
    acc += sm * b.s[t];
  }
  J.dst[(size_t)c * Nd + p] = acc;
}

// OpticalFlow.cpp:189-214: gray = .114 B + .587 G + .299 R (Image.h:1464),
// d/dx and d/dy of gray with [1,-8,0,8,-1]/12 (Image.h:955-1016), G-B, G-R.
__device__ __forceinline__ double gray_at(const double* bgr, int N, int o) {
  return (double)bgr[o] * .114 + bgr[N + o] * .587 + bgr[2 * N + o] * .299;
}
template <int S>
__device__ __forceinline__ void features_at(const double* bgr, int W, int H, int l, double* out) {
  const int N = W * H;
  const int i = l / W, j = l - i * W;
  double f[5] = {1, -8, 0, 8, -1};
  for (int t = 0; t < 5; t++) f[t] /= 12;
  double gx = 0, gy = 0;
  for (int t = -2; t <= 2; t++) gx += gray_at(bgr, N, i * W + enforce_range(j + t, W)) * f[t + 2];
  for (int t = -2; t <= 2; t++) gy += gray_at(bgr, N, enforce_range(i + t, H) * W + j) * f[t + 2];
  img_st<S>(out, l, gray_at(bgr, N, l));
  img_st<S>(out, N + l, gx);
  img_st<S>(out, 2 * (size_t)N + l, gy);
  img_st<S>(out, 3 * (size_t)N + l, bgr[N + l] - bgr[l]);
  img_st<S>(out, 4 * (size_t)N + l, bgr[N + l] - bgr[2 * N + l]);
}

__global__ void __launch_bounds__(kBlock) k_features_planes(const double* bgr, int W, int H, double* out) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l < W * H) features_at<kImgF64>(bgr, W, H, l, out);
}

// Features of both frames at a level (sceneFlowMultiscale.cpp:102-103).
template <int S>
__global__ void __launch_bounds__(kBlock) k_features(Views w) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int item = work_item(w, blk);
  const LevelView& v = w.items[item];
  const int l = (blk - w.start[item]) * kBlock + threadIdx.x;
  if (l >= v.N) return;
  features_at<S>(v.P1k, v.W, v.H, l, const_cast<double*>(v.I1));
  features_at<S>(v.P2k, v.W, v.H, l, const_cast<double*>(v.I2));
}

// Level entry (sceneFlowMultiscale.cpp:120-137).  At the pair's coarsest
// level: u = 0 and the warped frame 2 is frame 2.  Otherwise: the coarser
// flow resized to this level (Image.h:764-769 -> ResizeImage by dimensions,
// ImageProcessing.h:224-242), u and v scaled by 1/ratio (:133-134), then
// warpBy3DFlow (opticalFlowUtils.cpp:76-99 -> warpImage, ImageProcessing.h:471-491:
// out of bounds -> frame 1; then depth -= w).  The warp of pixel p reads only
// u(p), so resize and warp share one pass.
template <int S>
__global__ void __launch_bounds__(kBlock) k_level_entry(Views wl) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int item = work_item(wl, blk);
  const LevelView& v = wl.items[item];
  const int l = (blk - wl.start[item]) * kBlock + threadIdx.x;
  const int W = v.W, H = v.H, N = v.N;
  if (l >= N) return;
  if (!v.u_prev) {
    v.u[l] = 0;
    v.u[N + l] = 0;
    v.u[2 * N + l] = 0;
    for (int k = 0; k < kFeat; k++) img_st<S>(v.Iw, (size_t)k * N + l, img_ld<S>(v.I2, (size_t)k * N + l));
    v.Dw[l] = v.D2[l];
    return;
  }
  const int i = l / W, j = l - i * W;
  double uu[3];
  if (v.Wp == W && v.Hp == H) {  // a flow of this level's size (sceneFlowWithInit): taken as it is
    for (int c = 0; c < 3; c++) {
      double a = v.u_prev[(size_t)c * N + l];
      if (c < 2) a *= v.uv_scale;
      uu[c] = a;
      v.u[(size_t)c * N + l] = a;
    }
  } else {
    const int Ws = v.Wp, Hs = v.Hp, Ns = Ws * Hs;
    const double xr = (double)W / Ws, yr = (double)H / Hs;
    const double x = (double)(j + 1) / xr - 1;
    const double y = (double)(i + 1) / yr - 1;
    const Bilerp b = bilerp_taps(x, y, Ws, Hs);
    for (int c = 0; c < 3; c++) {
      const double* s = v.u_prev + (size_t)c * Ns;
      double acc = 0;
      for (int t = 0; t < 4; t++) acc += s[b.o[t]] * b.s[t];
      if (c < 2) acc *= v.uv_scale;
      uu[c] = acc;
      v.u[(size_t)c * N + l] = acc;
    }
  }
  const double y = i + uu[1], x = j + uu[0];
  double d;
  if (x < 0 || x > W - 1 || y < 0 || y > H - 1) {
    for (int k = 0; k < kFeat; k++) img_st<S>(v.Iw, (size_t)k * N + l, img_ld<S>(v.I1, (size_t)k * N + l));
    d = v.D1[l];
  } else {
    const Bilerp b = bilerp_taps(x, y, W, H);
    for (int k = 0; k < kFeat; k++) {
      double acc = 0;
      for (int t = 0; t < 4; t++) acc += img_ld<S>(v.I2, (size_t)k * N + b.o[t]) * b.s[t];
      img_st<S>(v.Iw, (size_t)k * N + l, acc);
    }
    d = 0;
    for (int t = 0; t < 4; t++) d += v.D2[b.o[t]] * b.s[t];
  }
  v.Dw[l] = d - uu[2];
}

// Level weights, converted to f32 as the solver's device copies are (.cu:680-687):
//   delta (pyramid of OpticalFlow.cpp:71-81), s (pyramid of f^2, :90-97),
//   m (OpticalFlow.cpp:104-111 evaluated with the FULL-size camera over the
//      level's depth, as sceneFlowMultiscale.cpp:144 calls it: level pixel l
//      reads D1[(l / W_full) * W + l % W_full]),
//   r (OpticalFlow.cpp:113-141), c = f_level / D1 (.cu:693-700).
__global__ void __launch_bounds__(kBlock) k_level_weights(Views wl) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int item = work_item(wl, blk);
  const LevelView& v = wl.items[item];
  const int l = (blk - wl.start[item]) * kBlock + threadIdx.x;
  const int W = v.W, H = v.H, N = v.N;
  if (l >= N) return;
  const int i = l / W, j = l - i * W;
  v.dW[l] = (float)v.delk[l];
  v.sW[l] = (float)v.smk[l];
  v.cW[l] = (float)(v.f_level / v.D1[l]);
  if (v.mag64) {  // the single-level entry: m and r given by the caller
    v.mW[l] = (float)v.mag64[l];
    v.rW[l] = (float)v.reg64[(size_t)l * 2 + 0];
    v.rW[N + l] = (float)v.reg64[(size_t)l * 2 + 1];
    return;
  }
  const size_t r = (size_t)l / v.W_full, c = (size_t)l % v.W_full;
  const size_t idx = r * W + c;
  const double d = idx < (size_t)N ? v.D1[idx] : 0.0;
  v.mW[l] = (float)(d > 0 ? dsq(v.f_full / d) : dsq(v.f_full / 2));
  double wr = 0, wd = 0;
  if (j < W - 1) {
    if (v.reg_type == 0) wr = 1;
    else if (v.reg_type == 1) wr = exp(-dsq((v.D1[l] - v.D1[l + 1]) / .02));
    else wr = 1 - v.hpbk[(size_t)l * 2 + 0] * .99;
  }
  if (i < H - 1) {
    if (v.reg_type == 0) wd = 1;
    else if (v.reg_type == 1) wd = exp(-dsq((v.D1[l] - v.D1[l + W]) / .02));
    else wd = 1 - v.hpbk[(size_t)l * 2 + 1] * .99;
  }
  v.rW[l] = (float)wr;
  v.rW[N + l] = (float)wd;
}

// staticDepthNoiseModeling.cpp:95-97: depth-boundary pixel (a valid lower or
// right neighbour more than 0.02 * stereoErrorRatio away).  Depths are read as
// float, the type of the reference's depth map (T = double input is cast).
template <class T>
__device__ __forceinline__ bool is_depth_boundary(const T* depth, int W, int H, int i, int j) {
  const float d = (float)depth[(size_t)i * W + j];
  if (i < H - 1) {
    const float dn = (float)depth[(size_t)(i + 1) * W + j];
    if (dn > 0 && fabsf(d - dn) > .02 * stereo_error_ratio(fminf(d, dn))) return true;
  }
  if (j < W - 1) {
    const float dn = (float)depth[(size_t)i * W + j + 1];
    if (dn > 0 && fabsf(d - dn) > .02 * stereo_error_ratio(fminf(d, dn))) return true;
  }
  return false;
}

// computeDepthMapStdevsPointwise (staticDepthNoiseModeling.cpp:225-258): sigma
// = max(1e-4, sigma(1m) * stereoErrorRatio(z)) (1 where invalid), inflated by
// exp((1-a) log(10 sigma) + a log(sigma)), a = d/maxDist, within L1 distance
// d <= maxDist of a depth boundary (:112-124; the exact L1 distance, searched
// over the diamond of radius maxDist).
template <class T>
__device__ __forceinline__ float depth_sigma_at(const T* depth, int W, int H, int l, int max_dist) {
  const int i = l / W, j = l - i * W;
  const float sigmaAt1m = (float)stereo_error(1);
  const float z = (float)depth[l];
  float st = z > 0 ? (float)fmax(1e-4, sigmaAt1m * stereo_error_ratio(z)) : 1.0f;
  if (max_dist > 0) {
    int best = max_dist + 1;
    for (int di = -max_dist; di <= max_dist; di++) {
      const int ii = i + di;
      if (ii < 0 || ii >= H) continue;
      const int rem = max_dist - abs(di);
      for (int dj = -rem; dj <= rem; dj++) {
        const int jj = j + dj;
        if (jj < 0 || jj >= W) continue;
        const int dd = abs(di) + abs(dj);
        if (dd < best && is_depth_boundary(depth, W, H, ii, jj)) best = dd;
      }
    }
    if (best <= max_dist) {
      const float a = (float)best / max_dist;
      st = (float)exp((1 - a) * log((double)(10 * st)) + a * log((double)st));
    }
  }
  return st;
}

__global__ void __launch_bounds__(kBlock) k_depth_sigma(const float* depth, int W, int H, int max_dist,
                                                        float* sigma) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l < W * H) sigma[l] = depth_sigma_at(depth, W, H, l, max_dist);
}

// Per pair at the start of a solve: depth stdevs of frame 1 if not given
// (runFlowTwoFrames.cpp:90, from the float depth), delta = stereoError(1) /
// sigma * depthVsColorDataWeight (OpticalFlow.cpp:71-81), s = f^2 (:90-97).
// One workgroup per kPrepW x kPrepH tile: the depth-boundary test of every
// pixel within the search radius of the tile is made once, into LDS, and each
// pixel searches its diamond there (the values and the minimum distance of
// depth_sigma_at).
__global__ void __launch_bounds__(kBlock) k_pair_prep(WorkList<PairPrep> wl) {
  constexpr int R = kPrepRadius, BW = kPrepW + 2 * R, BH = kPrepH + 2 * R;
  __shared__ float dz[(BH + 1) * (BW + 1)];  // the flags' region, one more column and row for their tests
  __shared__ unsigned char bnd[BH * BW];
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int item = work_item(wl, blk);
  const PairPrep& p = wl.items[item];
  const int W = p.W, H = p.H;
  const int lb = blk - wl.start[item];
  const int tx = (W + kPrepW - 1) / kPrepW;
  const int x0 = (lb % tx) * kPrepW, y0 = (lb / tx) * kPrepH;
  const int i = y0 + threadIdx.x / kPrepW, j = x0 + threadIdx.x % kPrepW;
  const bool inside = i < H && j < W;
  const int l = inside ? i * W + j : 0;
  float sg = 0.f;
  if (p.compute_sigma) {
    for (int t = threadIdx.x; t < (BH + 1) * (BW + 1); t += kBlock) {
      const int ri = y0 - R + t / (BW + 1), rj = x0 - R + t % (BW + 1);
      dz[t] = ri >= 0 && rj >= 0 && ri < H && rj < W ? (float)p.dm1[(size_t)ri * W + rj] : 0.f;
    }
    __syncthreads();
    // is_depth_boundary on the staged values
    for (int t = threadIdx.x; t < BH * BW; t += kBlock) {
      const int bi = t / BW, bj = t - bi * BW;
      const int ri = y0 - R + bi, rj = x0 - R + bj;
      bool b = false;
      if (ri >= 0 && rj >= 0 && ri < H && rj < W) {
        const float d = dz[bi * (BW + 1) + bj];
        if (ri < H - 1) {
          const float dn = dz[(bi + 1) * (BW + 1) + bj];
          if (dn > 0 && fabsf(d - dn) > .02 * stereo_error_ratio(fminf(d, dn))) b = true;
        }
        if (rj < W - 1) {
          const float dn = dz[bi * (BW + 1) + bj + 1];
          if (dn > 0 && fabsf(d - dn) > .02 * stereo_error_ratio(fminf(d, dn))) b = true;
        }
      }
      bnd[t] = b ? 1 : 0;
    }
    __syncthreads();
    if (inside) {
      const float sigmaAt1m = (float)stereo_error(1);
      const float z = dz[(i - y0 + R) * (BW + 1) + (j - x0 + R)];
      sg = z > 0 ? (float)fmax(1e-4, sigmaAt1m * stereo_error_ratio(z)) : 1.0f;
      int best = R + 1;
      for (int di = -R; di <= R; di++) {
        const int rem = R - abs(di);
        for (int dj = -rem; dj <= rem; dj++) {
          const int dd = abs(di) + abs(dj);
          if (dd < best && bnd[(i - y0 + R + di) * BW + (j - x0 + R + dj)]) best = dd;
        }
      }
      if (best <= R) {
        const float a = (float)best / R;
        sg = (float)exp((1 - a) * log((double)(10 * sg)) + a * log((double)sg));
      }
      p.sigma[l] = sg;
    }
  } else if (inside) {
    sg = p.sigma[l];
  }
  if (!inside) return;
  p.del0[l] = stereo_error(1) / sg * p.dvc;
  p.sm0[l] = p.f2;
}

// depthMapPyramid.cpp:33-138: one level from the next larger one; a valid
// centre sample is replaced by the mean of its 2x2 depth-consistent component
// (threshold 5 mm x stereoErrorRatio), an invalid one by -1.
__global__ void __launch_bounds__(kBlock) k_depth_map_level(const double* L, int LW, int LH, double* dst, int W,
                                                            int H) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= W * H) return;
  const int i = l / W, j = l - i * W;
  const float factor = (float)LH / H;
  const float maxDZ = .005f;
  const int x = max(0, min(LW - 1, (int)rintf(factor * j)));
  const int y = max(0, min(LH - 1, (int)rintf(factor * i)));
  const double c = L[(size_t)y * LW + x];
  if (!(c > 0)) { dst[l] = -1; return; }
  int xs[2], ys[2];
  xs[0] = max(0, min(LW - 1, (int)floorf(factor * j)));
  xs[1] = min(LW - 1, xs[0] + 1);
  ys[0] = max(0, min(LH - 1, (int)floorf(factor * i)));
  ys[1] = min(LH - 1, ys[0] + 1);
  const int xi = (x == xs[0]) ? 0 : 1, yi = (y == ys[0]) ? 0 : 1;
  const double a = L[(size_t)ys[yi] * LW + xs[!xi]];
  const double b = L[(size_t)ys[!yi] * LW + xs[xi]];
  const double d = L[(size_t)ys[!yi] * LW + xs[!xi]];
  float sum = (float)c;
  int n = 1;
  const double tc = maxDZ * stereo_error_ratio(c);
  if (fabs(a - c) < tc) { sum += (float)a; n++; }
  if (fabs(b - c) < tc) { sum += (float)b; n++; }
  if (fabs(d - c) < tc) { sum += (float)d; n++; }
  if (n > 1) { dst[l] = sum / (float)n; return; }
  float sum2 = (float)d;
  int n2 = 1;
  const double td = maxDZ * stereo_error_ratio(d);
  if (fabs(a - d) < td) { sum2 += (float)a; n2++; }
  if (fabs(b - d) < td) { sum2 += (float)b; n2++; }
  dst[l] = n2 > 1 ? (double)(sum2 / (float)n2) : c;
}

// Frame ingestion (ceLiuOpticalFlow.cpp:19-47 cv2liuBGR / cv2liuDepth): u8
// BGR [N][3] -> planes of c / 255.0 (no channel swap), f32 depth -> f64.
__global__ void __launch_bounds__(kBlock) k_ingest(const unsigned char* bgr, const float* depth, int N, double* im,
                                                   double* dm) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= N) return;
  for (int c = 0; c < 3; c++) im[(size_t)c * N + l] = bgr[(size_t)l * 3 + c] / 255.0;
  dm[l] = depth[l];
}

// Depth hole fill, the parallel variant of recursiveMedianFilter
// (recursiveMedianFiltering.cpp:34-100, halfwidth 1).  The reference fills
// one pixel at a time, always one with the most known 3x3 neighbours, with
// the upper median of their values (as float), in an order its hash
// container decides.  Here, in rounds r = 1, 2, ...: M = the most known
// neighbours (valid, or filled in an earlier round) of any pixel still
// invalid; every such pixel with M known neighbours is filled at once.  Equal
// to the reference wherever no two pixels filled in one round touch
// (isolated holes).  One workgroup per frame: the invalid pixels are listed
// once, then each round is two passes over the list of the pixels still
// unfilled, with a barrier between; the second pass moves the pixels it
// leaves to the next round's list (st[l]: 0 valid, r filled in round r,
// kUnfilled).  Every round with M > 0 fills at least the pixel that has M, so
// the rounds end.
constexpr int kFillThreads = 1024;
constexpr int kUnfilled = 0x7fffffff;

__device__ __forceinline__ int fill_known(const double* d, const int* st, int W, int H, int l, int r, float* vals) {
  const int i = l / W, j = l - i * W;
  int n = 0;
  for (int ii = max(0, i - 1); ii <= min(H - 1, i + 1); ii++)
    for (int jj = max(0, j - 1); jj <= min(W - 1, j + 1); jj++) {
      const int q = ii * W + jj;
      if (gp(st)[q] < r) {
        if (vals) vals[n] = (float)gp(d)[q];
        n++;
      }
    }
  return n;
}

__global__ void __launch_bounds__(kFillThreads) k_median_fill(double* d, int W, int H, int* st, int* todo,
                                                              int* left_out) {
  __shared__ int s_n, s_m, s_max;
  const int N = W * H;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int l = threadIdx.x; l < N; l += kFillThreads) {
    // known: depth > 0; filled: depth <= 0 (a NaN is neither, as in the reference)
    const double x = gp(d)[l];
    gp(st)[l] = x > 0 ? 0 : kUnfilled;
    if (x <= 0) gp(todo)[atomicAdd(&s_n, 1)] = l;
  }
  __syncthreads();
  // todo holds two lists of N: the pixels still unfilled, and the next round's
  int n = s_n, cur = 0;
  const int n0 = n;
  for (int r = 1; n > 0 && r <= n0; r++) {
    const int* list = todo + (size_t)cur * N;
    int* next = todo + (size_t)(1 - cur) * N;
    if (threadIdx.x == 0) {
      s_max = 0;
      s_m = 0;
    }
    __syncthreads();
    int mx = 0;
    for (int t = threadIdx.x; t < n; t += kFillThreads) mx = max(mx, fill_known(d, st, W, H, gp(list)[t], r, nullptr));
    atomicMax(&s_max, mx);
    __syncthreads();
    const int M = s_max;
    if (M == 0) break;
    for (int t = threadIdx.x; t < n; t += kFillThreads) {
      const int l = gp(list)[t];
      float vals[9];
      const int k = fill_known(d, st, W, H, l, r, vals);
      if (k != M) {
        gp(next)[atomicAdd(&s_m, 1)] = l;  // still unfilled: into the next round's list
        continue;
      }
      for (int a = 1; a < k; a++) {  // insertion sort, ascending
        const float x = vals[a];
        int b = a - 1;
        while (b >= 0 && vals[b] > x) {
          vals[b + 1] = vals[b];
          b--;
        }
        vals[b + 1] = x;
      }
      gp(d)[l] = (double)vals[k / 2];
      gp(st)[l] = r;
    }
    __syncthreads();
    n = s_m;
    cur = 1 - cur;
    __syncthreads();
  }
  if (threadIdx.x == 0) *gp(left_out) = n;
}

// The colour-guided fill of small holes, the parallel variant of
// recursiveMedianFilterSmallInvalidRegionsUsingColor
// (recursiveMedianFiltering.cpp:127-222, halfwidth 1), which fills only the
// 4-connected invalid regions of at most max_region pixels, each pixel with
// the mean of its known 3x3 neighbours weighted by
// exp(-|colour difference|^2 / 0.6), region by region in an order its hash
// containers decide.  Here all small regions at once, in rounds: a region's
// pixels with the region's largest number of known neighbours (valid, or
// filled in an earlier round) are filled together.
constexpr int kMaxRegion = 63;

// A region's label: its smallest pixel index (max_region or fewer pixels),
// -1 (larger), -2 (the pixel is not invalid).  Each invalid pixel walks its
// own region breadth-first, at most max_region + 1 pixels: the queue and an
// open-addressing set of the pixels seen (128 slots, at most 64 used) are
// rows of LDS (odd pitches: lanes at the same position hit distinct banks).
constexpr int kRegionThreads = 128;
constexpr int kQPitch = kMaxRegion + 2, kSeen = 128, kSeenPitch = kSeen + 1;
__global__ void __launch_bounds__(kRegionThreads) k_small_regions(const double* d, int W, int H, int max_region,
                                                                  int* label) {
  __shared__ int qs[kRegionThreads * kQPitch];
  __shared__ int ss[kRegionThreads * kSeenPitch];
  const int l = blockIdx.x * kRegionThreads + threadIdx.x;
  if (l >= W * H) return;
  if (!(gp(d)[l] <= 0)) {
    gp(label)[l] = -2;
    return;
  }
  int* q = qs + threadIdx.x * kQPitch;
  int* seen = ss + threadIdx.x * kSeenPitch;
  for (int t = 0; t < kSeen; t++) seen[t] = -1;
  // inserts m into the set; false when it was there already
  auto insert = [&](int m) {
    unsigned h = ((unsigned)m * 2654435761u) >> 25;
    for (;;) {
      const int v = seen[h];
      if (v == m) return false;
      if (v < 0) {
        seen[h] = m;
        return true;
      }
      h = (h + 1) & (kSeen - 1);
    }
  };
  int head = 0, tail = 1, lo = l;
  q[0] = l;
  insert(l);
  bool large = false;
  while (head < tail && !large) {
    const int p = q[head++];
    const int i = p / W, j = p - i * W;
    const int nb[4] = {i > 0 ? p - W : -1, j > 0 ? p - 1 : -1, j < W - 1 ? p + 1 : -1, i < H - 1 ? p + W : -1};
    for (int k = 0; k < 4 && !large; k++) {
      const int m = nb[k];
      if (m < 0 || !(gp(d)[m] <= 0) || !insert(m)) continue;
      if (tail == max_region) {
        large = true;  // a (max_region + 1)-th pixel
        break;
      }
      q[tail++] = m;
      lo = min(lo, m);
    }
  }
  gp(label)[l] = large ? -1 : lo;
}

// exp(x) for the colour weights, in f64: x = n ln 2 + r (Cody-Waite, ln 2 in
// two parts), e^r by its Taylor series to degree 13 (|r| <= 0.35: error below
// 1e-17 relative), times 2^n.  The oracle evaluates the same operations.
__device__ __forceinline__ double fill_exp(double x) {
  if (!(x > -700.0)) return 0.0;
  const double n = rint(x * 1.4426950408889634);
  const double r = (x - n * 0.693145751953125) - n * 1.4286068203094173e-06;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)n);
}

// One workgroup per frame, as k_median_fill; cmax[label]: the region's
// largest number of known neighbours this round.
__global__ void __launch_bounds__(kFillThreads) k_colour_fill(double* d, const double* bgr, int W, int H,
                                                              const int* label, int* st, int* todo, int* cmax,
                                                              int* left_out) {
  __shared__ int s_n, s_m, s_filled;
  const int N = W * H;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int l = threadIdx.x; l < N; l += kFillThreads) {
    gp(st)[l] = gp(d)[l] > 0 ? 0 : kUnfilled;
    if (gp(label)[l] >= 0) gp(todo)[atomicAdd(&s_n, 1)] = l;
  }
  __syncthreads();
  int n = s_n, cur = 0;
  const int n0 = n;
  for (int r = 1; n > 0 && r <= n0; r++) {
    const int* list = todo + (size_t)cur * N;
    int* next = todo + (size_t)(1 - cur) * N;
    if (threadIdx.x == 0) {
      s_m = 0;
      s_filled = 0;
    }
    for (int t = threadIdx.x; t < n; t += kFillThreads) gp(cmax)[gp(label)[gp(list)[t]]] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += kFillThreads) {
      const int l = gp(list)[t];
      atomicMax(cmax + gp(label)[l], fill_known(d, st, W, H, l, r, nullptr));
    }
    __syncthreads();
    int filled = 0;
    for (int t = threadIdx.x; t < n; t += kFillThreads) {
      const int l = gp(list)[t];
      const int k = fill_known(d, st, W, H, l, r, nullptr);
      if (k == 0 || k != gp(cmax)[gp(label)[l]]) {
        gp(next)[atomicAdd(&s_m, 1)] = l;
        continue;
      }
      const int i = l / W, j = l - i * W;
      const double c0 = gp(bgr)[l], c1 = gp(bgr)[N + l], c2 = gp(bgr)[2 * (size_t)N + l];
      float zs = 0.f, ws = 0.f;
      for (int ii = max(0, i - 1); ii <= min(H - 1, i + 1); ii++)
        for (int jj = max(0, j - 1); jj <= min(W - 1, j + 1); jj++) {
          const int q = ii * W + jj;
          if (gp(st)[q] < r) {
            const double e0 = c0 - gp(bgr)[q], e1 = c1 - gp(bgr)[N + q], e2 = c2 - gp(bgr)[2 * (size_t)N + q];
            const float w = (float)fill_exp(-(e0 * e0 + e1 * e1 + e2 * e2) / .6);
            zs = (float)((double)zs + gp(d)[q] * w);
            ws += w;
          }
        }
      gp(d)[l] = (double)(zs / ws);
      gp(st)[l] = r;
      filled++;
    }
    atomicAdd(&s_filled, filled);
    __syncthreads();
    n = s_m;
    cur = 1 - cur;
    const bool stuck = s_filled == 0;
    __syncthreads();
    if (stuck) break;  // regions with no known neighbour (the reference asserts there)
  }
  if (threadIdx.x == 0) *gp(left_out) = n;
}

inline int blocks_for(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_deinterleave(hipStream_t s, const double* src, int N, int C, double* dst) {
  hipLaunchKernelGGL(k_deinterleave, dim3(blocks_for(N)), dim3(kBlock), 0, s, src, N, C, dst);
}
void launch_interleave(hipStream_t s, const double* src, int N, int C, double* dst) {
  hipLaunchKernelGGL(k_interleave, dim3(blocks_for(N)), dim3(kBlock), 0, s, src, N, C, dst);
}
void launch_ingest(hipStream_t s, const unsigned char* bgr, const float* depth, int N, double* im, double* dm) {
  hipLaunchKernelGGL(k_ingest, dim3(blocks_for(N)), dim3(kBlock), 0, s, bgr, depth, N, im, dm);
}
void launch_pair_prep(hipStream_t s, const WorkList<PairPrep>& w, int blocks) {
  hipLaunchKernelGGL(k_pair_prep, dim3(blocks), dim3(kBlock), 0, s, w);
}
void launch_pyramid_jobs(hipStream_t s, const WorkList<PyrJob>& w, int blocks) {
  hipLaunchKernelGGL(k_pyramid_jobs, dim3(blocks), dim3(kBlock), 0, s, w);
}
void launch_features(hipStream_t s, const Views& v, int blocks, int img) {
  auto fn = img == kImgF16 ? k_features<kImgF16> : img == kImgBF16 ? k_features<kImgBF16> : k_features<kImgF64>;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, s, v);
}
void launch_features_planes(hipStream_t s, const double* bgr, int W, int H, double* out5) {
  hipLaunchKernelGGL(k_features_planes, dim3(blocks_for((size_t)W * H)), dim3(kBlock), 0, s, bgr, W, H, out5);
}
void launch_level_entry(hipStream_t s, const Views& v, int blocks, int img) {
  auto fn = img == kImgF16 ? k_level_entry<kImgF16> : img == kImgBF16 ? k_level_entry<kImgBF16> : k_level_entry<kImgF64>;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, s, v);
}
void launch_level_weights(hipStream_t s, const Views& v, int blocks) {
  hipLaunchKernelGGL(k_level_weights, dim3(blocks), dim3(kBlock), 0, s, v);
}
void launch_depth_sigma(hipStream_t s, const float* depth, int W, int H, int max_dist, float* sigma) {
  hipLaunchKernelGGL(k_depth_sigma, dim3(blocks_for((size_t)W * H)), dim3(kBlock), 0, s, depth, W, H, max_dist, sigma);
}
void launch_colour_fill(hipStream_t s, double* depth, const double* bgr, int W, int H, int max_region, int* label,
                        int* st, int* todo, int* cmax, int* left) {
  hipLaunchKernelGGL(k_small_regions, dim3((W * H + kRegionThreads - 1) / kRegionThreads), dim3(kRegionThreads), 0, s,
                     depth, W, H, max_region, label);
  hipLaunchKernelGGL(k_colour_fill, dim3(1), dim3(kFillThreads), 0, s, depth, bgr, W, H, label, st, todo, cmax, left);
}
void launch_median_fill(hipStream_t s, double* depth, int W, int H, int* st, int* todo, int* left) {
  hipLaunchKernelGGL(k_median_fill, dim3(1), dim3(kFillThreads), 0, s, depth, W, H, st, todo, left);
}
void launch_depth_map_level(hipStream_t s, const double* larger, int LW, int LH, double* dst, int W, int H) {
  hipLaunchKernelGGL(k_depth_map_level, dim3(blocks_for((size_t)W * H)), dim3(kBlock), 0, s, larger, LW, LH, dst, W,
                     H);
}

}  // namespace rgbdflow
