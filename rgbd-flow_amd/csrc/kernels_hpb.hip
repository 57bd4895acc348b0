// kernels_hpb.hip -- the heuristic boundary probabilities of one pyramid level
// on gfx950: computeHeuristicPb (src/heuristicProbBoundary.cpp:21-41 ->
// src/heuristicBoundaries.cpp:40-358), the map the HPB regularisation turns
// into edge weights 1 - 0.99 p (OpticalFlow.cpp:113-141).
//
// Stages, one kernel each, every pixel independent except the two integral
// image passes:
//   k_hpb_prep      colour to u8 by truncation and back to float, the
//                   organised point cloud (f = 525 W / 640, c = (W/2, H/2)),
//                   the per-pixel half-disk radius round(f 0.02 / z) in [1, 5]
//   k_hpb_normals   normal of each valid point: the points of its 9x9 window
//                   (staged in LDS per 16x16 pixels)
//                   within 0.05 x stereoErrorRatio(z) in depth (at least 27),
//                   f64 two-pass covariance, smallest eigenvector by power
//                   iteration on its adjugate (PCL's solver is not
//                   restated), flipped towards the camera
//   k_hpb_rows      running row sums of the 7 channels (colour, normal,
//                   normal validity), left to right in f32 / u32, through
//                   64x64 LDS tiles
//   k_hpb_cols      the integral images, top to bottom per column with the
//                   reference's recurrence ii = (up + row sum) + pixel
//                   (halfDiskFeatures.ipp:139-150), so every sum rounds as
//                   the reference's sequential loop rounds it
//   k_hpb_halfdisk  half-disk colour distance and normal dissimilarity with
//                   the convexity factor
//   k_hpb_calc      the per-edge probability (CALC) with its invalid-run
//                   walk and the image-edge hack
//   k_hpb_nms       non-maximum suppression along each direction
// exp is hpb_exp below (Cody-Waite reduction by ln 2, Taylor to degree 11 in
// f64, one rounding to f32); sqrt and division are IEEE, and the library is
// built with -ffp-contract=off, so each step rounds where the source says.
#include "internal.hpp"

namespace rgbdflow {

namespace {

constexpr int kHB = 256;

// The item of a batch that workgroup blk works on: the last item whose first
// block is <= blk (start holds n + 1 ascending entries).
__device__ __forceinline__ int hpb_item(const int* start, int n, int blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double stereo_error(double z) { return z * z / (1 / .00285 + z); }
__device__ __forceinline__ double stereo_error_ratio(double z) { return stereo_error(z) / stereo_error(1.0); }

__device__ __forceinline__ float hpb_exp(float xf) {
  const double x = xf;
  if (!(x > -104.0)) return 0.f;
  const double n = rint(x * 1.4426950408889634);
  const double r = (x - n * 0.693145751953125) - n * 1.4286068203094173e-06;
  double p = 1.0 / 39916800.0;
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
  return (float)ldexp(p, (int)n);
}

__device__ __forceinline__ float sqrf(float x) { return x * x; }
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return (ax * bx + ay * by) + az * bz;
}
// Eigen's normalize(): v / sqrt(|v|^2) when |v|^2 > 0
__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
  const float s = dot3(x, y, z, x, y, z);
  if (s > 0.f) {
    const float n = sqrtf(s);
    x /= n;
    y /= n;
    z /= n;
  }
}

// Eigenvector of the smallest eigenvalue of a symmetric positive
// semi-definite 3x3 matrix (f64).  The adjugate adj(A) has A's eigenvectors,
// with eigenvalues l1 l2, l0 l2, l0 l1 (l0 <= l1 <= l2), so A's smallest
// eigenvector is adj(A)'s dominant one: start from adj(A)'s column of largest
// norm (the null vector itself when A has rank 2) and apply 8 normalised
// power iterations of adj(A), each shrinking the error by l0 / l1.
__device__ __forceinline__ void smallest_eigvec(double a[3][3], double n[3]) {
  double c[3][3];
  c[0][0] = a[1][1] * a[2][2] - a[1][2] * a[1][2];
  c[1][1] = a[0][0] * a[2][2] - a[0][2] * a[0][2];
  c[2][2] = a[0][0] * a[1][1] - a[0][1] * a[0][1];
  c[0][1] = c[1][0] = a[0][2] * a[1][2] - a[0][1] * a[2][2];
  c[0][2] = c[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
  c[1][2] = c[2][1] = a[0][1] * a[0][2] - a[0][0] * a[1][2];
  int k = 0;
  double best = -1.0;
  for (int j = 0; j < 3; j++) {
    const double s = (c[0][j] * c[0][j] + c[1][j] * c[1][j]) + c[2][j] * c[2][j];
    if (s > best) {
      best = s;
      k = j;
    }
  }
  double x[3] = {c[0][k], c[1][k], c[2][k]};
  for (int it = 0; it < 8; it++) {
    double y[3];
    for (int i = 0; i < 3; i++) y[i] = (c[i][0] * x[0] + c[i][1] * x[1]) + c[i][2] * x[2];
    const double s = sqrt((y[0] * y[0] + y[1] * y[1]) + y[2] * y[2]);
    if (!(s > 0)) break;
    for (int i = 0; i < 3; i++) x[i] = y[i] / s;
  }
  for (int i = 0; i < 3; i++) n[i] = x[i];
}

struct Cam {
  float cx, cy, f;
};
__device__ __forceinline__ Cam hpb_cam(int W, int H) {
  Cam c;
  c.cx = (float)(unsigned)(W / 2);
  c.cy = (float)(unsigned)(H / 2);
  c.f = (float)(525.0 * (unsigned)W / 640);
  return c;
}

__global__ void __launch_bounds__(kHB) k_hpb_prep(const HpbBatch B) {
  const int it = hpb_item(B.s_pix, B.n, blockIdx.x);
  const HpbLevel& L = B.items[it];
  const int l = (blockIdx.x - B.s_pix[it]) * kHB + threadIdx.x;
  const int W = L.W, H = L.H, N = W * H;
  if (l >= N) return;
  const int i = l / W, j = l - i * W;
  const Cam c = hpb_cam(W, H);
  const float d = (float)gp(L.depth)[l];
  gp(L.pts)[l] = ((float)j - c.cx) * d / c.f;
  gp(L.pts)[N + l] = ((float)i - c.cy) * d / c.f;
  gp(L.pts)[2 * N + l] = d;
  // cv::Vec3b(255 * c2, 255 * c1, 255 * c0) (heuristicProbBoundary.cpp:34), then / 255.0 as float
  for (int k = 0; k < 3; k++) {
    const unsigned char u8 = (unsigned char)(255 * gp(L.bgr)[(size_t)(2 - k) * N + l]);
    gp(L.col)[(size_t)k * N + l] = (float)(u8 / 255.0);
  }
  gp(L.rad)[l] = d > 0 ? max(1, min(5, (int)rintf(c.f * .02f / d))) : 5;
}

// A workgroup takes 16x16 pixels and stages the points of their windows
// (24x24, the image's outside never read) in LDS once.
constexpr int kNT = 16, kNR = kNT + 8, kNP = kNR * kNR;
__global__ void __launch_bounds__(kHB) k_hpb_normals(const HpbBatch B) {
  __shared__ float PX[kNP], PY[kNP], PZ[kNP];
  const int it = hpb_item(B.s_tile, B.n, blockIdx.x);
  const HpbLevel& L = B.items[it];
  const int W = L.W, H = L.H, N = W * H;
  const int ntx = (W + kNT - 1) / kNT, b = blockIdx.x - B.s_tile[it];
  const int oy = (b / ntx) * kNT, ox = (b - (b / ntx) * ntx) * kNT;
  const int ry0 = oy - 4, rx0 = ox - 4;
  const RGBDFLOW_GLOBAL float* X = gp(L.pts);
  for (int e = threadIdx.x; e < kNP; e += kHB) {
    const int gy = ry0 + e / kNR, gx = rx0 + e % kNR;
    const bool in = gy >= 0 && gy < H && gx >= 0 && gx < W;
    const int q = gy * W + gx;
    PX[e] = in ? X[q] : 0.f;
    PY[e] = in ? X[N + q] : 0.f;
    PZ[e] = in ? X[2 * N + q] : 0.f;
  }
  __syncthreads();
  const int i = oy + threadIdx.x / kNT, j = ox + threadIdx.x % kNT;
  if (i >= H || j >= W) return;
  const int l = i * W + j;
  float n0 = 0.f, n1 = 0.f, n2 = 0.f;
  unsigned valid = 0;
  const float z = PZ[(i - ry0) * kNR + (j - rx0)];
  if (z > 0.f) {
    const double thr = (double).05f * stereo_error_ratio((double)z);
    const int y0 = max(0, i - 4), y1 = min(H - 1, i + 4), x0 = max(0, j - 4), x1 = min(W - 1, j + 4);
    int cnt = 0;
    double m[3] = {0, 0, 0};
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) {
        const int q = (y - ry0) * kNR + (x - rx0);
        if ((double)fabsf(z - PZ[q]) < thr) {
          cnt++;
          m[0] += PX[q];
          m[1] += PY[q];
          m[2] += PZ[q];
        }
      }
    if (cnt >= 27) {  // max(6, 9 * 9 / 3) neighbours
      for (int k = 0; k < 3; k++) m[k] /= (double)cnt;
      double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) {
          const int q = (y - ry0) * kNR + (x - rx0);
          if ((double)fabsf(z - PZ[q]) < thr) {
            const double d0 = PX[q] - m[0], d1 = PY[q] - m[1], d2 = PZ[q] - m[2];
            a[0][0] += d0 * d0;
            a[0][1] += d0 * d1;
            a[0][2] += d0 * d2;
            a[1][1] += d1 * d1;
            a[1][2] += d1 * d2;
            a[2][2] += d2 * d2;
          }
        }
      a[1][0] = a[0][1];
      a[2][0] = a[0][2];
      a[2][1] = a[1][2];
      double n[3];
      smallest_eigvec(a, n);
      const double nn = sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
      if (nn > 0) {
        for (int k = 0; k < 3; k++) n[k] /= nn;
        const double vp = ((-(double)X[l]) * n[0] + (-(double)X[N + l]) * n[1]) + (-(double)X[2 * N + l]) * n[2];
        if (vp < 0)
          for (int k = 0; k < 3; k++) n[k] = -n[k];
        n0 = (float)n[0];
        n1 = (float)n[1];
        n2 = (float)n[2];
        valid = (double)fabsf(dot3(n0, n1, n2, n0, n1, n2) - 1.f) < 1e-4 ? 1u : 0u;
      }
    }
  }
  gp(L.nrm)[l] = n0;
  gp(L.nrm)[N + l] = n1;
  gp(L.nrm)[2 * N + l] = n2;
  gp(L.nval)[l] = valid;
}

// Running row sums before each pixel: rs[c][l] = sum of the row's values
// left of l, added one by one from the left (every f32 partial sum rounds as
// a sequential loop rounds it).  One wave per 64 rows of a channel: 64x64
// tiles are loaded coalesced (lane = column) into LDS, lane r then scans row
// r of the tile in order, and the tile goes back out coalesced; the next
// tile's loads are issued before the scan.
constexpr int kRowTile = 64, kRowPitch = 65;  // odd pitch: the lanes scanning rows hit distinct banks
template <class V>
__device__ __forceinline__ void row_scan(const RGBDFLOW_GLOBAL V* src, RGBDFLOW_GLOBAL V* dst, int W, int nr,
                                         V* T) {
  const int lane = threadIdx.x;
  const int nt = (W + kRowTile - 1) / kRowTile;
  V x[kRowTile];
  auto load = [&](int t) {
    const int j = t * kRowTile + lane;
    const RGBDFLOW_GLOBAL V* q = src + j;
#pragma unroll
    for (int r = 0; r < kRowTile; r++) {
      x[r] = (r < nr && j < W) ? *q : V(0);
      q += W;
    }
  };
  load(0);
  V s = V(0);
  for (int t = 0; t < nt; t++) {
    V* B = T + (t & 1) * kRowTile * kRowPitch;
#pragma unroll
    for (int r = 0; r < kRowTile; r++) B[r * kRowPitch + lane] = x[r];
    __syncthreads();
    if (t + 1 < nt) load(t + 1);
    // columns past W hold zeros: adding them leaves s as it is, and they are not stored
    V* row = B + lane * kRowPitch;
    V v[kRowTile];
#pragma unroll
    for (int k = 0; k < kRowTile; k++) v[k] = row[k];
#pragma unroll
    for (int k = 0; k < kRowTile; k++) {
      row[k] = s;
      s += v[k];
    }
    __syncthreads();
    const int j = t * kRowTile + lane;
    if (j < W)
      for (int r = 0; r < nr; r++) dst[(size_t)r * W + j] = B[r * kRowPitch + lane];
  }
}

__global__ void __launch_bounds__(kRowTile) k_hpb_rows(const HpbBatch B) {
  __shared__ float T[2 * kRowTile * kRowPitch];
  const int it = hpb_item(B.s_row, B.n, blockIdx.x);
  const HpbLevel& L = B.items[it];
  const int W = L.W, H = L.H, N = W * H;
  const int ng = (H + kRowTile - 1) / kRowTile, b = blockIdx.x - B.s_row[it];
  const int c = b / ng, i0 = (b - c * ng) * kRowTile;
  const int nr = min(kRowTile, H - i0);
  const size_t o = (size_t)i0 * W;
  if (c < 6) {
    const RGBDFLOW_GLOBAL float* src = c < 3 ? gp(L.col) + (size_t)c * N : gp(L.nrm) + (size_t)(c - 3) * N;
    row_scan<float>(src + o, gp(L.rowsum) + (size_t)c * N + o, W, nr, T);
  } else {
    row_scan<unsigned>(gp(L.nval) + o, gp(L.vrowsum) + o, W, nr, reinterpret_cast<unsigned*>(T));
  }
}

// The integral images with a zero row and column in front,
// ii[(i+1)(W+1) + j+1] = (ii above + row sum) + pixel, top to bottom per
// column (one lane per column and channel, consecutive lanes on consecutive
// columns); the next kColChunk rows are loaded before the current ones are
// accumulated and stored.
constexpr int kColChunk = 32;
template <class V>
__device__ __forceinline__ void col_scan(const RGBDFLOW_GLOBAL V* rs, const RGBDFLOW_GLOBAL V* px,
                                         RGBDFLOW_GLOBAL V* ii, int W, int H, int j) {
  const int IW = W + 1;
  ii[j + 1] = V(0);
  if (j == 0)
    for (int i = 0; i <= H; i++) ii[(size_t)i * IW] = V(0);
  V a[kColChunk], b[kColChunk];
  auto load = [&](int i0) {
#pragma unroll
    for (int k = 0; k < kColChunk; k++) {
      const bool in = i0 + k < H;
      a[k] = in ? rs[(size_t)(i0 + k) * W + j] : V(0);
      b[k] = in ? px[(size_t)(i0 + k) * W + j] : V(0);
    }
  };
  load(0);
  V up = V(0);
  for (int i0 = 0; i0 < H; i0 += kColChunk) {
    V ca[kColChunk], cb[kColChunk];
#pragma unroll
    for (int k = 0; k < kColChunk; k++) {
      ca[k] = a[k];
      cb[k] = b[k];
    }
    if (i0 + kColChunk < H) load(i0 + kColChunk);
    const int n = min(kColChunk, H - i0);
#pragma unroll
    for (int k = 0; k < kColChunk; k++)
      if (k < n) {
        const V v = (up + ca[k]) + cb[k];
        ii[(size_t)(i0 + k + 1) * IW + j + 1] = v;
        up = v;
      }
  }
}

__global__ void __launch_bounds__(kHB) k_hpb_cols(const HpbBatch B) {
  const int it = hpb_item(B.s_col, B.n, blockIdx.x);
  const HpbLevel& L = B.items[it];
  const int t = (blockIdx.x - B.s_col[it]) * kHB + threadIdx.x;
  const int W = L.W, H = L.H, N = W * H;
  const size_t IN = (size_t)(W + 1) * (H + 1);
  if (t >= 7 * W) return;
  const int c = t / W, j = t - c * W;
  if (c < 6) {
    const RGBDFLOW_GLOBAL float* px = c < 3 ? gp(L.col) + (size_t)c * N : gp(L.nrm) + (size_t)(c - 3) * N;
    col_scan<float>(gp(L.rowsum) + (size_t)c * N, px, gp(L.ii) + (size_t)c * IN, W, H, j);
  } else {
    col_scan<unsigned>(gp(L.vrowsum), gp(L.nval), gp(L.iv), W, H, j);
  }
}

__global__ void __launch_bounds__(kHB) k_hpb_halfdisk(const HpbBatch B) {
  const int it = hpb_item(B.s_pix, B.n, blockIdx.x);
  const HpbLevel& L = B.items[it];
  const int l = (blockIdx.x - B.s_pix[it]) * kHB + threadIdx.x;
  const int W = L.W, H = L.H, N = W * H, IW = W + 1;
  if (l >= N) return;
  const size_t IN = (size_t)IW * (H + 1);
  const int i = l / W, j = l - i * W;
  const int hw = gp(L.rad)[l];
  const int xmin = max(0, j - hw), xmax = min(W - 1, j + hw - 1), ymin = max(0, i - hw), ymax = min(H - 1, i + hw - 1);
  const float hs0 = (float)((j - xmin) * (ymax - ymin + 1)), hs1 = (float)((xmax - j + 1) * (ymax - ymin + 1));
  const float vs0 = (float)((xmax - xmin + 1) * (i - ymin)), vs1 = (float)((xmax - xmin + 1) * (ymax - i + 1));
  const RGBDFLOW_GLOBAL float* II = gp(L.ii);
  const RGBDFLOW_GLOBAL unsigned* IV = gp(L.iv);
  const RGBDFLOW_GLOBAL float* X = gp(L.pts);
  // element (y, x) of integral image c, y and x from -1
  auto at = [&](int c, int y, int x) { return II[(size_t)c * IN + (size_t)(y + 1) * IW + (x + 1)]; };
  auto atv = [&](int y, int x) { return IV[(size_t)(y + 1) * IW + (x + 1)]; };
  float ch = 0.f, nh = 0.f, cv = 0.f, nv = 0.f;
  if (j > 0) {
    float h0[6], h1[6];
    for (int c = 0; c < 6; c++) {
      h1[c] = ((at(c, ymax, xmax) - at(c, ymax, j - 1)) - at(c, ymin - 1, xmax)) + at(c, ymin - 1, j - 1);
      h0[c] = ((at(c, ymax, j - 1) - at(c, ymin - 1, j - 1)) - at(c, ymax, xmin - 1)) + at(c, ymin - 1, xmin - 1);
    }
    float s = 0.f;
    for (int c = 0; c < 3; c++) s = s + sqrf(h1[c] / hs1 - h0[c] / hs0);
    ch = sqrtf(s);
    float a0 = h0[3], a1 = h0[4], a2 = h0[5], b0 = h1[3], b1 = h1[4], b2 = h1[5];
    normalize3(a0, a1, a2);
    normalize3(b0, b1, b2);
    const unsigned v1 = ((atv(ymax, xmax) - atv(ymax, j - 1)) - atv(ymin - 1, xmax)) + atv(ymin - 1, j - 1);
    const unsigned v0 = ((atv(ymax, j - 1) - atv(ymin - 1, j - 1)) - atv(ymax, xmin - 1)) + atv(ymin - 1, xmin - 1);
    nh = min(v0, v1) < 3 ? 0.f : (float)(.5 * (double)(1.f - dot3(a0, a1, a2, b0, b1, b2)));
    if (min(v0, v1) > 3) {
      float dx = X[l - 1] - X[l], dy = X[N + l - 1] - X[N + l], dz = X[2 * N + l - 1] - X[2 * N + l];
      normalize3(dx, dy, dz);
      const double convexity = (double)((dot3(a0, a1, a2, dx, dy, dz) + dot3(b0, b1, b2, -dx, -dy, -dz)) / 4) + .5;
      nh = (float)((double)nh * (1 - convexity));
    }
  }
  if (i > 0) {
    float h0[6], h1[6];
    for (int c = 0; c < 6; c++) {
      h1[c] = ((at(c, ymax, xmax) - at(c, i - 1, xmax)) - at(c, ymax, xmin - 1)) + at(c, i - 1, xmin - 1);
      h0[c] = ((at(c, i - 1, xmax) - at(c, i - 1, xmin - 1)) - at(c, ymin - 1, xmax)) + at(c, ymin - 1, xmin - 1);
    }
    float s = 0.f;
    for (int c = 0; c < 3; c++) s = s + sqrf(h1[c] / vs1 - h0[c] / vs0);
    cv = sqrtf(s);
    float a0 = h0[3], a1 = h0[4], a2 = h0[5], b0 = h1[3], b1 = h1[4], b2 = h1[5];
    normalize3(a0, a1, a2);
    normalize3(b0, b1, b2);
    const unsigned v1 = ((atv(ymax, xmax) - atv(i - 1, xmax)) - atv(ymax, xmin - 1)) + atv(i - 1, xmin - 1);
    const unsigned v0 = ((atv(i - 1, xmax) - atv(i - 1, xmin - 1)) - atv(ymin - 1, xmax)) + atv(ymin - 1, xmin - 1);
    nv = min(v0, v1) < 3 ? 0.f : (float)(.5 * (double)(1.f - dot3(a0, a1, a2, b0, b1, b2)));
    if (min(v0, v1) > 3) {
      float dx = X[l - W] - X[l], dy = X[N + l - W] - X[N + l], dz = X[2 * N + l - W] - X[2 * N + l];
      normalize3(dx, dy, dz);
      const double convexity = (double)((dot3(a0, a1, a2, dx, dy, dz) + dot3(b0, b1, b2, -dx, -dy, -dz)) / 4) + .5;
      nv = (float)((double)nv * (1 - convexity));
    }
  }
  gp(L.feat)[l] = ch;
  gp(L.feat)[N + l] = nh;
  gp(L.feat)[2 * N + l] = cv;
  gp(L.feat)[3 * N + l] = nv;
}

__device__ float hpb_calc(const HpbLevel& L, int i, int j, int k, float z1, float z2, float dcol, float dnorm) {
  const int W = L.W, H = L.H;
  const float depthSigma = .02f, colSigma = .7f, normSigma = .4f;
  float colProb, depthProb;
  if (fmaxf(z1, z2) <= 0) {
    depthProb = 0;
    colProb = 1 - hpb_exp(-sqrf(dcol / colSigma));
  } else if (fminf(z1, z2) <= 0) {
    // the reference's walk: start and direction codes as written (they do
    // not follow the edge's direction)
    const int di[4] = {0, -1, 0, 1}, dj[4] = {-1, 0, 1, 0};
    const int dir = (z1 <= 0) ? (k + 2) % 4 : k;
    int i0 = (z1 <= 0) ? i : i + di[k], j0 = (z1 <= 0) ? j : j + dj[k];
    unsigned size = 0;
    while (j0 >= 0 && j0 < W && i0 >= 0 && i0 < H && (float)gp(L.depth)[(size_t)i0 * W + j0] <= 0) {
      size++;
      i0 += di[dir];
      j0 += dj[dir];
    }
    const bool oob = j0 < 0 || j0 >= W || i0 < 0 || i0 >= H;
    if (oob) depthProb = 0;
    else {
      const double a = fminf(1.0f, (float)size / 30);
      depthProb = (float)((1 - a) * .2 + a * 1);  // linterp (mathUtils.cpp:23-26)
    }
    colProb = 1 - hpb_exp(-sqrf(dcol / colSigma));
  } else {
    const float dz = (float)(fabsf(z1 - z2) / stereo_error_ratio(fminf(z1, z2)));
    depthProb = 1 - hpb_exp(-sqrf(dz / depthSigma));
    colProb = 0;
  }
  const float normProb = 1 - hpb_exp(-sqrf(dnorm / normSigma));
  auto mx = [](float a, float b) { return (a < b) ? b : a; };  // std::max
  float p = mx(depthProb, mx(colProb, mx(0.0f, normProb)));
  if (((i == 0 || i == H - 2) && k == 1) || ((j == 0 || j == W - 2) && k == 0)) p = 0;
  return p;
}

__global__ void __launch_bounds__(kHB) k_hpb_calc(const HpbBatch B) {
  const int it = hpb_item(B.s_pix, B.n, blockIdx.x);
  const HpbLevel& L = B.items[it];
  const int l = (blockIdx.x - B.s_pix[it]) * kHB + threadIdx.x;
  const int W = L.W, H = L.H, N = W * H;
  if (l >= N) return;
  const int i = l / W, j = l - i * W;
  const RGBDFLOW_GLOBAL float* z = gp(L.pts) + 2 * (size_t)N;
  const RGBDFLOW_GLOBAL float* F = gp(L.feat);
  float p0 = 0.f, p1 = 0.f;
  if (j < W - 1) p0 = hpb_calc(L, i, j, 0, z[l], z[l + 1], F[l + 1], F[N + l + 1]);
  if (i < H - 1) p1 = hpb_calc(L, i, j, 1, z[l], z[l + W], F[2 * N + l + W], F[3 * N + l + W]);
  gp(L.bp)[2 * l] = p0;
  gp(L.bp)[2 * l + 1] = p1;
}

__global__ void __launch_bounds__(kHB) k_hpb_nms(const HpbBatch B) {
  const int it = hpb_item(B.s_pix, B.n, blockIdx.x);
  const HpbLevel& L = B.items[it];
  const int l = (blockIdx.x - B.s_pix[it]) * kHB + threadIdx.x;
  const int W = L.W, H = L.H, N = W * H;
  if (l >= N) return;
  const int i = l / W, j = l - i * W;
  const int hw = gp(L.rad)[l];
  const RGBDFLOW_GLOBAL float* bp = gp(L.bp);
  float o0 = bp[2 * l], o1 = bp[2 * l + 1];
  int mi = -1;
  for (int k = max(0, i - hw); k <= min(H - 1, i + hw); k++)
    if (mi < 0 || bp[((size_t)k * W + j) * 2 + 1] > bp[((size_t)mi * W + j) * 2 + 1]) mi = k;
  if (mi != i) o1 = 0;
  mi = -1;
  for (int k = max(0, j - hw); k <= min(W - 1, j + hw); k++)
    if (mi < 0 || bp[((size_t)i * W + k) * 2] > bp[((size_t)i * W + mi) * 2]) mi = k;
  if (mi != j) o0 = 0;
  gp(L.out)[2 * l] = o0;
  gp(L.out)[2 * l + 1] = o1;
}

}  // namespace

size_t hpb_scratch_floats(int W, int H) {
  const size_t N = (size_t)W * H, IN = (size_t)(W + 1) * (H + 1);
  // pts 3N, col 3N, nrm 3N, nval N, rad N, rowsum 6N, vrowsum N, ii 6 IN, iv IN, feat 4N, bp 2N
  return 3 * N + 3 * N + 3 * N + N + N + 6 * N + N + 7 * IN + 4 * N + 2 * N + 64 * 11;
}

HpbLevel hpb_level(int W, int H, const double* bgr, const double* depth, float* out, float* scratch) {
  const size_t N = (size_t)W * H, IN = (size_t)(W + 1) * (H + 1);
  HpbLevel L;
  L.W = W;
  L.H = H;
  L.bgr = bgr;
  L.depth = depth;
  L.out = out;
  float* p = scratch;
  auto take = [&](size_t n) {
    float* q = p;
    p += (n + 63) & ~size_t(63);
    return q;
  };
  L.pts = take(3 * N);
  L.col = take(3 * N);
  L.nrm = take(3 * N);
  L.nval = reinterpret_cast<unsigned*>(take(N));
  L.rad = reinterpret_cast<int*>(take(N));
  L.rowsum = take(6 * N);
  L.vrowsum = reinterpret_cast<unsigned*>(take(N));
  L.ii = take(6 * IN);
  L.iv = reinterpret_cast<unsigned*>(take(IN));
  L.feat = take(4 * N);
  L.bp = take(2 * N);
  return L;
}

namespace {
int row_blocks(const HpbLevel& L) { return 7 * ((L.H + kRowTile - 1) / kRowTile); }
int col_blocks(const HpbLevel& L) { return (7 * L.W + kHB - 1) / kHB; }
int pix_blocks(const HpbLevel& L) { return (L.W * L.H + kHB - 1) / kHB; }
int tile_blocks(const HpbLevel& L) { return ((L.W + kNT - 1) / kNT) * ((L.H + kNT - 1) / kNT); }
}  // namespace

bool hpb_batch_add(HpbBatch& b, const HpbLevel& L) {
  if (b.n >= kHpbMaxItems) return false;
  if (b.n == 0) b.s_pix[0] = b.s_row[0] = b.s_col[0] = b.s_tile[0] = 0;
  b.items[b.n] = L;
  b.s_pix[b.n + 1] = b.s_pix[b.n] + pix_blocks(L);
  b.s_row[b.n + 1] = b.s_row[b.n] + row_blocks(L);
  b.s_col[b.n + 1] = b.s_col[b.n] + col_blocks(L);
  b.s_tile[b.n + 1] = b.s_tile[b.n] + tile_blocks(L);
  b.n++;
  return true;
}

void launch_hpb_batch(hipStream_t s, const HpbBatch& b) {
  if (b.n == 0) return;
  const dim3 pix(b.s_pix[b.n]);
  hipLaunchKernelGGL(k_hpb_prep, pix, dim3(kHB), 0, s, b);
  hipLaunchKernelGGL(k_hpb_normals, dim3(b.s_tile[b.n]), dim3(kHB), 0, s, b);
  hipLaunchKernelGGL(k_hpb_rows, dim3(b.s_row[b.n]), dim3(kRowTile), 0, s, b);
  hipLaunchKernelGGL(k_hpb_cols, dim3(b.s_col[b.n]), dim3(kHB), 0, s, b);
  hipLaunchKernelGGL(k_hpb_halfdisk, pix, dim3(kHB), 0, s, b);
  hipLaunchKernelGGL(k_hpb_calc, pix, dim3(kHB), 0, s, b);
  hipLaunchKernelGGL(k_hpb_nms, pix, dim3(kHB), 0, s, b);
}

void launch_hpb(hipStream_t s, const HpbLevel& L) {
  HpbBatch b;
  b.n = 0;
  hpb_batch_add(b, L);
  launch_hpb_batch(s, b);
}

}  // namespace rgbdflow
