// kernels_pyramid.hip -- per-pair and per-level preparation kernels on gfx950.
//
// Pyramids (src/GaussianPyramid.cpp, src/depthMapPyramid.cpp), features
// (src/OpticalFlow.cpp:148-220), the flow upsample (Image.h:764-769 +
// sceneFlowMultiscale.cpp:130-134), the level-entry warp (opticalFlowUtils.cpp:76-99)
// and the static weight maps (OpticalFlow.cpp:71-141, staticDepthNoiseModeling.cpp).
// Each kernel evaluates the reference's sums in the reference's order, so its
// f64 results are the same numbers the reference's host code computes.
#include "internal.hpp"

namespace rgbdflow {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxTaps = 64;

struct Taps { double g[kMaxTaps]; };

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

__global__ void k_to_f32(const double* src, size_t n, float* dst) {
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) dst[i] = (float)src[i];
}

__global__ void k_fill(double* dst, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) dst[i] = v;
}

// One pyramid level (GaussianPyramid.cpp:47-61): Gaussian smoothing of the
// source (Image.h:1179-1199 -> imfilter_hv: h-pass then v-pass, clamp-to-edge,
// ImageProcessing.h:247-269,339-358) followed by ResizeImage(ratio)
// (ImageProcessing.h:203-222).  Only the smoothed samples the resize reads are
// evaluated; taps of zero bilinear weight are skipped (adding x*0 leaves the
// sum unchanged).
__global__ void __launch_bounds__(kBlock) k_pyramid_level(const double* src, int Ws, int Hs, int C, Taps taps,
                                                          int fs, double ratio, double* dst, int Wd, int Hd) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  const int Nd = Wd * Hd;
  if (l >= Nd * C) return;
  const int c = l / Nd;
  const int p = l - c * Nd;
  const int i = p / Wd, j = p - i * Wd;
  const double x = (double)(j + 1) / ratio - 1;
  const double y = (double)(i + 1) / ratio - 1;
  const Bilerp b = bilerp_taps(x, y, Ws, Hs);
  const double* s = src + (size_t)c * Ws * Hs;
  double acc = 0;
  for (int t = 0; t < 4; t++) {
    if (b.s[t] == 0.0) continue;
    const int vv = b.o[t] / Ws, uu = b.o[t] - vv * Ws;
    double sm = 0;
    for (int lv = -fs; lv <= fs; lv++) {
      const int ii = enforce_range(vv + lv, Hs);
      double th = 0;
      for (int lh = -fs; lh <= fs; lh++) th += s[(size_t)ii * Ws + enforce_range(uu + lh, Ws)] * taps.g[lh + fs];
      sm += th * taps.g[lv + fs];
    }
    acc += sm * b.s[t];
  }
  dst[(size_t)c * Nd + p] = acc;
}

// OpticalFlow.cpp:189-214: gray = .114 B + .587 G + .299 R (Image.h:1464),
// d/dx and d/dy of gray with [1,-8,0,8,-1]/12 (Image.h:955-1016), G-B, G-R.
__device__ __forceinline__ double gray_at(const double* bgr, int N, int o) {
  return (double)bgr[o] * .114 + bgr[N + o] * .587 + bgr[2 * N + o] * .299;
}
__global__ void __launch_bounds__(kBlock) k_features(const double* bgr, int W, int H, double* out) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  const int N = W * H;
  if (l >= N) return;
  const int i = l / W, j = l - i * W;
  double f[5] = {1, -8, 0, 8, -1};
  for (int t = 0; t < 5; t++) f[t] /= 12;
  double gx = 0, gy = 0;
  for (int t = -2; t <= 2; t++) gx += gray_at(bgr, N, i * W + enforce_range(j + t, W)) * f[t + 2];
  for (int t = -2; t <= 2; t++) gy += gray_at(bgr, N, enforce_range(i + t, H) * W + j) * f[t + 2];
  out[l] = gray_at(bgr, N, l);
  out[N + l] = gx;
  out[2 * N + l] = gy;
  out[3 * (size_t)N + l] = bgr[N + l] - bgr[l];
  out[4 * (size_t)N + l] = bgr[N + l] - bgr[2 * N + l];
}

// Image.h:764-769 imresize(w,h) -> ResizeImage(dst dims) (ImageProcessing.h:224-242)
// of u, v, w, then u, v scaled by 1/ratio (sceneFlowMultiscale.cpp:130-134).
__global__ void __launch_bounds__(kBlock) k_upsample(const double* src, int Ws, int Hs, double* dst, int Wd, int Hd,
                                                     double uv_scale) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  const int Nd = Wd * Hd, Ns = Ws * Hs;
  if (l >= Nd) return;
  const int i = l / Wd, j = l - i * Wd;
  const double xr = (double)Wd / Ws, yr = (double)Hd / Hs;
  const double x = (double)(j + 1) / xr - 1;
  const double y = (double)(i + 1) / yr - 1;
  const Bilerp b = bilerp_taps(x, y, Ws, Hs);
  for (int c = 0; c < 3; c++) {
    const double* s = src + (size_t)c * Ns;
    double acc = 0;
    for (int t = 0; t < 4; t++) acc += s[b.o[t]] * b.s[t];
    if (c < 2) acc *= uv_scale;
    dst[(size_t)c * Nd + l] = acc;
  }
}

// warpBy3DFlow (opticalFlowUtils.cpp:76-99) through ImageProcessing::warpImage
// (ImageProcessing.h:471-491): out of bounds -> frame 1; then depth -= w.
__global__ void __launch_bounds__(kBlock) k_warp_cpu_style(LevelView v) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  const int W = v.W, H = v.H, N = v.N;
  if (l >= N) return;
  const int i = l / W, j = l - i * W;
  const double y = i + v.u[N + l], x = j + v.u[l];
  double d;
  if (x < 0 || x > W - 1 || y < 0 || y > H - 1) {
    for (int k = 0; k < kFeat; k++) v.Iw[(size_t)k * N + l] = v.I1[(size_t)k * N + l];
    d = v.D1[l];
  } else {
    const Bilerp b = bilerp_taps(x, y, W, H);
    for (int k = 0; k < kFeat; k++) {
      const double* s = v.I2 + (size_t)k * N;
      double acc = 0;
      for (int t = 0; t < 4; t++) acc += s[b.o[t]] * b.s[t];
      v.Iw[(size_t)k * N + l] = acc;
    }
    d = 0;
    for (int t = 0; t < 4; t++) d += v.D2[b.o[t]] * b.s[t];
  }
  v.Dw[l] = d - v.u[2 * N + l];
}

// Level weights, converted to f32 as the solver's device copies are (.cu:680-687):
//   delta (pyramid of OpticalFlow.cpp:71-81), s (pyramid of f^2, :90-97),
//   m (OpticalFlow.cpp:104-111 evaluated with the FULL-size camera over the
//      level's depth, as sceneFlowMultiscale.cpp:144 calls it: level pixel l
//      reads D1[(l / W_full) * W + l % W_full]),
//   r (OpticalFlow.cpp:113-141), c = f_level / D1 (.cu:693-700).
__global__ void __launch_bounds__(kBlock) k_level_weights(const double* delta, const double* smooth,
                                                          const float* hpb, int reg_type, const double* D1, int W,
                                                          int H, int W_full, float f_full, float f_level, float* dW,
                                                          float* sW, float* mW, float* rW, float* cW) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  const int N = W * H;
  if (l >= N) return;
  const int i = l / W, j = l - i * W;
  dW[l] = (float)delta[l];
  sW[l] = (float)smooth[l];
  const size_t r = (size_t)l / W_full, c = (size_t)l % W_full;
  const size_t idx = r * W + c;
  const double d = idx < (size_t)N ? D1[idx] : 0.0;
  mW[l] = (float)(d > 0 ? dsq(f_full / d) : dsq(f_full / 2));
  double wr = 0, wd = 0;
  if (j < W - 1) {
    if (reg_type == 0) wr = 1;
    else if (reg_type == 1) wr = exp(-dsq((D1[l] - D1[l + 1]) / .02));
    else wr = 1 - hpb[(size_t)l * 2 + 0] * .99;
  }
  if (i < H - 1) {
    if (reg_type == 0) wd = 1;
    else if (reg_type == 1) wd = exp(-dsq((D1[l] - D1[l + W]) / .02));
    else wd = 1 - hpb[(size_t)l * 2 + 1] * .99;
  }
  rW[l] = (float)wr;
  rW[N + l] = (float)wd;
  cW[l] = (float)(f_level / D1[l]);
}

// The same conversion for the single-level entry point, whose caller supplies
// m and r directly (sceneFlowSingleScale's arguments).
__global__ void __launch_bounds__(kBlock) k_level_weights_f64(const double* delta, const double* smooth,
                                                              const double* mag, const double* reg2, const double* D1,
                                                              int W, int H, float f_level, float* dW, float* sW,
                                                              float* mW, float* rW, float* cW) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  const int N = W * H;
  if (l >= N) return;
  dW[l] = (float)delta[l];
  sW[l] = (float)smooth[l];
  mW[l] = (float)mag[l];
  rW[l] = (float)reg2[(size_t)l * 2 + 0];
  rW[N + l] = (float)reg2[(size_t)l * 2 + 1];
  cW[l] = (float)(f_level / D1[l]);
}

// staticDepthNoiseModeling.cpp:95-97: depth-boundary pixel (a valid lower or
// right neighbour more than 0.02 * stereoErrorRatio away).
__device__ __forceinline__ bool is_depth_boundary(const float* depth, int W, int H, int i, int j) {
  const float d = depth[(size_t)i * W + j];
  if (i < H - 1) {
    const float dn = depth[(size_t)(i + 1) * W + j];
    if (dn > 0 && fabsf(d - dn) > .02 * stereo_error_ratio(fminf(d, dn))) return true;
  }
  if (j < W - 1) {
    const float dn = depth[(size_t)i * W + j + 1];
    if (dn > 0 && fabsf(d - dn) > .02 * stereo_error_ratio(fminf(d, dn))) return true;
  }
  return false;
}

// computeDepthMapStdevsPointwise (staticDepthNoiseModeling.cpp:225-258): sigma
// = max(1e-4, sigma(1m) * stereoErrorRatio(z)) (1 where invalid), inflated by
// exp((1-a) log(10 sigma) + a log(sigma)), a = d/maxDist, within L1 distance
// d <= maxDist of a depth boundary (:112-124; the exact L1 distance, searched
// over the diamond of radius maxDist).
__global__ void __launch_bounds__(kBlock) k_depth_sigma(const float* depth, int W, int H, int max_dist,
                                                        float* sigma) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= W * H) return;
  const int i = l / W, j = l - i * W;
  const float sigmaAt1m = (float)stereo_error(1);
  const float z = depth[l];
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
  sigma[l] = st;
}

// OpticalFlow.cpp:71-81: delta = stereoError(1) / sigma * depthVsColorDataWeight
__global__ void k_depth_data_weights(const float* sigma, int N, double dvc, double* delta) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= N) return;
  delta[l] = stereo_error(1) / sigma[l] * dvc;
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

inline int blocks_for(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_deinterleave(hipStream_t s, const double* src, int N, int C, double* dst) {
  hipLaunchKernelGGL(k_deinterleave, dim3(blocks_for(N)), dim3(kBlock), 0, s, src, N, C, dst);
}
void launch_interleave(hipStream_t s, const double* src, int N, int C, double* dst) {
  hipLaunchKernelGGL(k_interleave, dim3(blocks_for(N)), dim3(kBlock), 0, s, src, N, C, dst);
}
void launch_fill_f64(hipStream_t s, double* dst, size_t n, double v) {
  int nb = blocks_for(n);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_fill, dim3(nb), dim3(kBlock), 0, s, dst, n, v);
}
void launch_to_f32(hipStream_t s, const double* src, size_t n, float* dst) {
  int nb = blocks_for(n);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(kBlock), 0, s, src, n, dst);
}
void launch_pyramid_level(hipStream_t s, const double* src, int Ws, int Hs, int C, const double* taps, int fsize,
                          double ratio, double* dst, int Wd, int Hd) {
  if (fsize < 0 || 2 * fsize + 1 > kMaxTaps) return;  // callers keep sigma*3 <= 31
  Taps t{};
  for (int k = 0; k < 2 * fsize + 1; k++) t.g[k] = taps[k];
  hipLaunchKernelGGL(k_pyramid_level, dim3(blocks_for((size_t)Wd * Hd * C)), dim3(kBlock), 0, s, src, Ws, Hs, C, t,
                     fsize, ratio, dst, Wd, Hd);
}
void launch_features(hipStream_t s, const double* bgr, int W, int H, double* out5) {
  hipLaunchKernelGGL(k_features, dim3(blocks_for((size_t)W * H)), dim3(kBlock), 0, s, bgr, W, H, out5);
}
void launch_upsample(hipStream_t s, const double* src, int Ws, int Hs, double* dst, int Wd, int Hd, double uv_scale) {
  hipLaunchKernelGGL(k_upsample, dim3(blocks_for((size_t)Wd * Hd)), dim3(kBlock), 0, s, src, Ws, Hs, dst, Wd, Hd,
                     uv_scale);
}
void launch_warp_cpu_style(hipStream_t s, const LevelView& v) {
  hipLaunchKernelGGL(k_warp_cpu_style, dim3(blocks_for(v.N)), dim3(kBlock), 0, s, v);
}
void launch_level_weights(hipStream_t s, const double* delta, const double* smooth, const float* hpb, int reg_type,
                          const double* D1, int W, int H, int W_full, float f_full, float f_level, float* dW,
                          float* sW, float* mW, float* rW, float* cW) {
  hipLaunchKernelGGL(k_level_weights, dim3(blocks_for((size_t)W * H)), dim3(kBlock), 0, s, delta, smooth, hpb,
                     reg_type, D1, W, H, W_full, f_full, f_level, dW, sW, mW, rW, cW);
}
void launch_level_weights_f64(hipStream_t s, const double* delta, const double* smooth, const double* mag,
                              const double* reg2, const double* D1, int W, int H, float f_level, float* dW, float* sW,
                              float* mW, float* rW, float* cW) {
  hipLaunchKernelGGL(k_level_weights_f64, dim3(blocks_for((size_t)W * H)), dim3(kBlock), 0, s, delta, smooth, mag,
                     reg2, D1, W, H, f_level, dW, sW, mW, rW, cW);
}
void launch_depth_sigma(hipStream_t s, const float* depth, int W, int H, int max_dist, float* sigma) {
  hipLaunchKernelGGL(k_depth_sigma, dim3(blocks_for((size_t)W * H)), dim3(kBlock), 0, s, depth, W, H, max_dist, sigma);
}
void launch_depth_data_weights(hipStream_t s, const float* sigma, int N, double dvc, double* delta) {
  hipLaunchKernelGGL(k_depth_data_weights, dim3(blocks_for(N)), dim3(kBlock), 0, s, sigma, N, dvc, delta);
}
void launch_depth_map_level(hipStream_t s, const double* larger, int LW, int LH, double* dst, int W, int H) {
  hipLaunchKernelGGL(k_depth_map_level, dim3(blocks_for((size_t)W * H)), dim3(kBlock), 0, s, larger, LW, LH, dst, W,
                     H);
}

}  // namespace rgbdflow
