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
//                   within 0.05 x stereoErrorRatio(z) in depth (at least 27),
//                   f64 two-pass covariance, cyclic Jacobi eigen-solver
//                   (Numerical Recipes 3rd ed. 11.1; PCL's solver is not
//                   restated), flipped towards the camera
//   k_hpb_rows      running row sums of the 7 channels (colour, normal,
//                   normal validity), left to right in f32 / u32
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

__device__ __forceinline__ void smallest_eigvec(double a[3][3], double n[3]) {
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 12; sweep++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const int p = r < 2 ? 0 : 1, q = r == 0 ? 1 : 2;
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
      const double at = fabs(theta);
      double t = 1.0 / (at + sqrt(theta * theta + 1.0));
      if (theta < 0) t = -t;
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      const double tau = s / (1.0 + c);
      a[p][p] = a[p][p] - t * apq;
      a[q][q] = a[q][q] + t * apq;
      a[p][q] = a[q][p] = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (k != p && k != q) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = a[p][k] = akp - s * (akq + tau * akp);
          a[k][q] = a[q][k] = akq + s * (akp - tau * akq);
        }
        const double vkp = v[k][p], vkq = v[k][q];
        v[k][p] = vkp - s * (vkq + tau * vkp);
        v[k][q] = vkq + s * (vkp - tau * vkq);
      }
    }
  int m = 0;
  if (a[1][1] < a[m][m]) m = 1;
  if (a[2][2] < a[m][m]) m = 2;
  for (int k = 0; k < 3; k++) n[k] = v[k][m];
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

__global__ void __launch_bounds__(kHB) k_hpb_prep(HpbLevel L) {
  const int l = blockIdx.x * kHB + threadIdx.x;
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

__global__ void __launch_bounds__(kHB) k_hpb_normals(HpbLevel L) {
  const int l = blockIdx.x * kHB + threadIdx.x;
  const int W = L.W, H = L.H, N = W * H;
  if (l >= N) return;
  const int i = l / W, j = l - i * W;
  const RGBDFLOW_GLOBAL float* X = gp(L.pts);
  float n0 = 0.f, n1 = 0.f, n2 = 0.f;
  unsigned valid = 0;
  const float z = X[2 * N + l];
  if (z > 0.f) {
    const double thr = (double).05f * stereo_error_ratio((double)z);
    const int y0 = max(0, i - 4), y1 = min(H - 1, i + 4), x0 = max(0, j - 4), x1 = min(W - 1, j + 4);
    int cnt = 0;
    double m[3] = {0, 0, 0};
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) {
        const int q = y * W + x;
        if ((double)fabsf(z - X[2 * N + q]) < thr) {
          cnt++;
          m[0] += X[q];
          m[1] += X[N + q];
          m[2] += X[2 * N + q];
        }
      }
    if (cnt >= 27) {  // max(6, 9 * 9 / 3) neighbours
      for (int k = 0; k < 3; k++) m[k] /= (double)cnt;
      double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) {
          const int q = y * W + x;
          if ((double)fabsf(z - X[2 * N + q]) < thr) {
            const double d0 = X[q] - m[0], d1 = X[N + q] - m[1], d2 = X[2 * N + q] - m[2];
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

// channel c of the 7 integral-image inputs at pixel l
__device__ __forceinline__ float chan_f(const HpbLevel& L, int c, int l, int N) {
  return c < 3 ? gp(L.col)[(size_t)c * N + l] : gp(L.nrm)[(size_t)(c - 3) * N + l];
}

// running row sums before each pixel: rs[c][l] = sum of row values left of l
__global__ void __launch_bounds__(kHB) k_hpb_rows(HpbLevel L) {
  const int t = blockIdx.x * kHB + threadIdx.x;
  const int W = L.W, H = L.H, N = W * H;
  if (t >= 7 * H) return;
  const int c = t / H, i = t - c * H;
  if (c < 6) {
    float s = 0.f;
    for (int j = 0; j < W; j++) {
      const int l = i * W + j;
      gp(L.rowsum)[(size_t)c * N + l] = s;
      s += chan_f(L, c, l, N);
    }
  } else {
    unsigned s = 0;
    for (int j = 0; j < W; j++) {
      const int l = i * W + j;
      gp(L.vrowsum)[l] = s;
      s += gp(L.nval)[l];
    }
  }
}

// integral images with a zero row and column in front: ii[(i+1)(W+1) + j+1]
__global__ void __launch_bounds__(kHB) k_hpb_cols(HpbLevel L) {
  const int t = blockIdx.x * kHB + threadIdx.x;
  const int W = L.W, H = L.H, N = W * H, IW = W + 1;
  const size_t IN = (size_t)IW * (H + 1);
  if (t >= 7 * W) return;
  const int c = t / W, j = t - c * W;
  if (c < 6) {
    RGBDFLOW_GLOBAL float* ii = gp(L.ii) + (size_t)c * IN;
    ii[j + 1] = 0.f;
    if (j == 0)
      for (int i = 0; i <= H; i++) ii[(size_t)i * IW] = 0.f;
    float up = 0.f;
    for (int i = 0; i < H; i++) {
      const int l = i * W + j;
      const float v = (up + gp(L.rowsum)[(size_t)c * N + l]) + chan_f(L, c, l, N);
      ii[(size_t)(i + 1) * IW + j + 1] = v;
      up = v;
    }
  } else {
    RGBDFLOW_GLOBAL unsigned* ii = gp(L.iv);
    ii[j + 1] = 0u;
    if (j == 0)
      for (int i = 0; i <= H; i++) ii[(size_t)i * IW] = 0u;
    unsigned up = 0u;
    for (int i = 0; i < H; i++) {
      const int l = i * W + j;
      const unsigned v = (up + gp(L.vrowsum)[l]) + gp(L.nval)[l];
      ii[(size_t)(i + 1) * IW + j + 1] = v;
      up = v;
    }
  }
}

__global__ void __launch_bounds__(kHB) k_hpb_halfdisk(HpbLevel L) {
  const int l = blockIdx.x * kHB + threadIdx.x;
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

__global__ void __launch_bounds__(kHB) k_hpb_calc(HpbLevel L) {
  const int l = blockIdx.x * kHB + threadIdx.x;
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

__global__ void __launch_bounds__(kHB) k_hpb_nms(HpbLevel L) {
  const int l = blockIdx.x * kHB + threadIdx.x;
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

void launch_hpb(hipStream_t s, const HpbLevel& L) {
  const int N = L.W * L.H;
  const int nb = (N + kHB - 1) / kHB;
  hipLaunchKernelGGL(k_hpb_prep, dim3(nb), dim3(kHB), 0, s, L);
  hipLaunchKernelGGL(k_hpb_normals, dim3(nb), dim3(kHB), 0, s, L);
  hipLaunchKernelGGL(k_hpb_rows, dim3((7 * L.H + kHB - 1) / kHB), dim3(kHB), 0, s, L);
  hipLaunchKernelGGL(k_hpb_cols, dim3((7 * L.W + kHB - 1) / kHB), dim3(kHB), 0, s, L);
  hipLaunchKernelGGL(k_hpb_halfdisk, dim3(nb), dim3(kHB), 0, s, L);
  hipLaunchKernelGGL(k_hpb_calc, dim3(nb), dim3(kHB), 0, s, L);
  hipLaunchKernelGGL(k_hpb_nms, dim3(nb), dim3(kHB), 0, s, L);
}

}  // namespace rgbdflow
