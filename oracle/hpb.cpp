/*
 * hpb.cpp -- CPU restatement of the heuristic boundary probabilities the
 * HPB regularisation uses (TEST INFRASTRUCTURE ONLY: the checker of the HIP
 * path, never linked into the product).
 *
 * computeHeuristicPb(img, depth) (src/heuristicProbBoundary.cpp:21-41 ->
 * src/heuristicBoundaries.cpp:40-358) on one pyramid level:
 *   - colour to u8 by truncation, camera f = 525 W / 640, c = (W/2, H/2) with
 *     integer halves (heuristicProbBoundary.cpp:24-34);
 *   - the organised point cloud (depth_to_cloud_lib.ipp:12-60) and its
 *     normals from 9x9 windows of points within 0.05 x stereoErrorRatio(z)
 *     in depth, at least 27 of them, flipped towards the camera
 *     (cloudNormals.ipp:132-207, 395-507);
 *   - integral images of colour, normals and normal validity with the
 *     reference's recurrence (halfDiskFeatures.ipp:101-152), half-disk
 *     colour and normal differences at a per-pixel radius round(f .02 / z)
 *     clamped to [1, 5], with the convexity factor
 *     (heuristicBoundaries.cpp:95-223);
 *   - the per-edge probability (CALC, :232-275) including its invalid-run
 *     walk, whose start and direction codes are the reference's (they do not
 *     match the edge's direction; kept), and the image-edge hack (:274);
 *   - non-maximum suppression along each direction (:309-353).
 * Not restated: the principal curvatures (:52-64), whose value CALC ignores.
 *
 * Third-party arithmetic restated, not copied: the normal of a window is the
 * eigenvector of the smallest eigenvalue of its point covariance
 * (pcl::NormalEstimation::computePointNormal, PCL unpinned version) -- here a
 * two-pass f64 covariance and the smallest eigenvector by power iteration on
 * the adjugate (PCL's eigen33 takes cross products of rows of A - l0 I, the
 * adjugate's columns at l0), so normals agree with PCL's to rounding, not
 * bit for bit.  exp() is evaluated by hpb_exp (Cody-Waite reduction by ln 2
 * and a degree-11 Taylor polynomial in f64, rounded to f32 once), which the
 * HIP path restates with the same operations.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

double stereo_error(double z) { return z * z / (1 / .00285 + z); }  // primesensorUtils.cpp:53-56
double stereo_error_ratio(double z) {                                // primesensorUtils.cpp:60-64
  static const double err1 = stereo_error(1);
  return stereo_error(z) / err1;
}

// exp(x) for x <= 0 as float: x = n ln2 + r with |r| <= ln2/2 (ln2 split in
// two parts, the first exact in 32 bits), e^r by Taylor to degree 11 in f64,
// times 2^n, rounded to float once.
float hpb_exp(float xf) {
  const double x = xf;
  if (!(x > -104.0)) return 0.f;
  const double n = std::nearbyint(x * 1.4426950408889634);
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
  return (float)std::ldexp(p, (int)n);
}

float sqrf(float x) { return x * x; }

struct F3 {
  float x, y, z;
};
float dot3(const F3& a, const F3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
F3 normalized(F3 v) {  // Eigen's normalize(): v / sqrt(|v|^2) when |v|^2 > 0
  const float s = dot3(v, v);
  if (s > 0.f) {
    const float n = std::sqrt(s);
    v.x /= n;
    v.y /= n;
    v.z /= n;
  }
  return v;
}

// Eigenvector of the smallest eigenvalue of a symmetric positive
// semi-definite 3x3 matrix (f64).  The adjugate adj(A) has A's eigenvectors,
// with eigenvalues l1 l2, l0 l2, l0 l1 (l0 <= l1 <= l2), so A's smallest
// eigenvector is adj(A)'s dominant one: start from adj(A)'s column of largest
// norm (the null vector itself when A has rank 2) and apply 8 normalised
// power iterations of adj(A), each shrinking the error by l0 / l1.
void smallest_eigvec(double a[3][3], double n[3]) {
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
    const double s = std::sqrt((y[0] * y[0] + y[1] * y[1]) + y[2] * y[2]);
    if (!(s > 0)) break;
    for (int i = 0; i < 3; i++) x[i] = y[i] / s;
  }
  for (int i = 0; i < 3; i++) n[i] = x[i];
}

}  // namespace

extern "C" {

void orc_hpb_normals(const double* depth_in, int W, int H, float* normals, unsigned* valid) {
  const size_t N = (size_t)W * H;
  const float cX = (float)(unsigned)(W / 2), cY = (float)(unsigned)(H / 2);
  const float f = (float)(525.0 * (unsigned)W / 640);
  std::vector<F3> X(N);
  std::vector<float> z(N);
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      const size_t l = (size_t)i * W + j;
      const float d = (float)depth_in[l];
      z[l] = d;
      X[l] = F3{((float)j - cX) * d / f, ((float)i - cY) * d / f, d};
    }
  const double maxz = (double).05f;
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      const size_t l = (size_t)i * W + j;
      float* nout = normals + 3 * l;
      nout[0] = nout[1] = nout[2] = 0.f;
      valid[l] = 0;
      if (!(z[l] > 0.f)) continue;
      const double thr = maxz * stereo_error_ratio((double)z[l]);
      std::vector<size_t> nb;
      for (int y = std::max(0, i - 4); y <= std::min(H - 1, i + 4); y++)
        for (int x = std::max(0, j - 4); x <= std::min(W - 1, j + 4); x++) {
          const size_t l2 = (size_t)y * W + x;
          if ((double)std::fabs(z[l] - z[l2]) < thr) nb.push_back(l2);
        }
      if ((int)nb.size() < 27) continue;  // max(6, 9 * 9 / 3) neighbours
      double m[3] = {0, 0, 0};
      for (size_t q : nb) {
        m[0] += X[q].x;
        m[1] += X[q].y;
        m[2] += X[q].z;
      }
      for (int k = 0; k < 3; k++) m[k] /= (double)nb.size();
      double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (size_t q : nb) {
        const double d0 = X[q].x - m[0], d1 = X[q].y - m[1], d2 = X[q].z - m[2];
        a[0][0] += d0 * d0;
        a[0][1] += d0 * d1;
        a[0][2] += d0 * d2;
        a[1][1] += d1 * d1;
        a[1][2] += d1 * d2;
        a[2][2] += d2 * d2;
      }
      a[1][0] = a[0][1];
      a[2][0] = a[0][2];
      a[2][1] = a[1][2];
      double n[3];
      smallest_eigvec(a, n);
      const double nn = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
      if (!(nn > 0)) continue;
      for (int k = 0; k < 3; k++) n[k] /= nn;
      // towards the viewpoint (0, 0, 0): flip when (0 - p) . n < 0
      const double vp = ((-(double)X[l].x) * n[0] + (-(double)X[l].y) * n[1]) + (-(double)X[l].z) * n[2];
      if (vp < 0)
        for (int k = 0; k < 3; k++) n[k] = -n[k];
      const F3 nf{(float)n[0], (float)n[1], (float)n[2]};
      nout[0] = nf.x;
      nout[1] = nf.y;
      nout[2] = nf.z;
      valid[l] = (double)std::fabs(dot3(nf, nf) - 1.f) < 1e-4 ? 1u : 0u;
    }
}

void orc_heuristic_pb(const double* bgr, const double* depth_in, int W, int H, float* out) {
  const size_t N = (size_t)W * H;
  const float cX = (float)(unsigned)(W / 2), cY = (float)(unsigned)(H / 2);
  const float f = (float)(525.0 * (unsigned)W / 640);
  std::vector<float> z(N), col(3 * N), nrm(3 * N);
  std::vector<unsigned> nval(N);
  std::vector<F3> X(N);
  std::vector<int> rad(N);
  orc_hpb_normals(depth_in, W, H, nrm.data(), nval.data());
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      const size_t l = (size_t)i * W + j;
      const float d = (float)depth_in[l];
      z[l] = d;
      X[l] = F3{((float)j - cX) * d / f, ((float)i - cY) * d / f, d};
      // cv::Vec3b(255 * c2, 255 * c1, 255 * c0), then / 255.0 as float
      for (int k = 0; k < 3; k++) {
        const unsigned char c = (unsigned char)(255 * bgr[l * 3 + (2 - k)]);
        col[3 * l + k] = (float)(c / 255.0);
      }
      rad[l] = d > 0 ? std::max(1, std::min(5, (int)std::rint(f * .02f / d))) : 5;
    }
  // integral images with a zero row and column in front (index +1)
  const int IW = W + 1;
  std::vector<float> ic((size_t)IW * (H + 1) * 3, 0.f), in((size_t)IW * (H + 1) * 3, 0.f);
  std::vector<unsigned> iv((size_t)IW * (H + 1), 0u);
  for (int i = 0; i < H; i++) {
    float rc[3] = {0, 0, 0}, rn[3] = {0, 0, 0};
    unsigned rv = 0;
    for (int j = 0; j < W; j++) {
      const size_t l = (size_t)i * W + j, o = (size_t)(i + 1) * IW + (j + 1), u = (size_t)i * IW + (j + 1);
      for (int k = 0; k < 3; k++) {
        ic[3 * o + k] = (ic[3 * u + k] + rc[k]) + col[3 * l + k];
        rc[k] += col[3 * l + k];
        in[3 * o + k] = (in[3 * u + k] + rn[k]) + nrm[3 * l + k];
        rn[k] += nrm[3 * l + k];
      }
      iv[o] = (iv[u] + rv) + nval[l];
      rv += nval[l];
    }
  }
  auto II = [&](const std::vector<float>& a, int y, int x, int k) { return a[3 * ((size_t)(y + 1) * IW + (x + 1)) + k]; };
  auto IV = [&](int y, int x) { return iv[(size_t)(y + 1) * IW + (x + 1)]; };
  std::vector<float> cH(N, 0.f), cV(N, 0.f), nH(N, 0.f), nV(N, 0.f);
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      const size_t l = (size_t)i * W + j;
      const int hw = rad[l];
      const int xmin = std::max(0, j - hw), xmax = std::min(W - 1, j + hw - 1), ymin = std::max(0, i - hw),
                ymax = std::min(H - 1, i + hw - 1);
      const float hs0 = (float)((j - xmin) * (ymax - ymin + 1)), hs1 = (float)((xmax - j + 1) * (ymax - ymin + 1));
      const float vs0 = (float)((xmax - xmin + 1) * (i - ymin)), vs1 = (float)((xmax - xmin + 1) * (ymax - i + 1));
      if (j > 0) {
        auto h1 = [&](const std::vector<float>& a, int k) {
          return ((II(a, ymax, xmax, k) - II(a, ymax, j - 1, k)) - II(a, ymin - 1, xmax, k)) + II(a, ymin - 1, j - 1, k);
        };
        auto h0 = [&](const std::vector<float>& a, int k) {
          return ((II(a, ymax, j - 1, k) - II(a, ymin - 1, j - 1, k)) - II(a, ymax, xmin - 1, k)) +
                 II(a, ymin - 1, xmin - 1, k);
        };
        float s = 0.f;
        for (int k = 0; k < 3; k++) s = s + sqrf(h1(ic, k) / hs1 - h0(ic, k) / hs0);
        cH[l] = std::sqrt(s);
        F3 n0{h0(in, 0), h0(in, 1), h0(in, 2)}, n1{h1(in, 0), h1(in, 1), h1(in, 2)};
        n0 = normalized(n0);
        n1 = normalized(n1);
        const unsigned v1 = ((IV(ymax, xmax) - IV(ymax, j - 1)) - IV(ymin - 1, xmax)) + IV(ymin - 1, j - 1);
        const unsigned v0 = ((IV(ymax, j - 1) - IV(ymin - 1, j - 1)) - IV(ymax, xmin - 1)) + IV(ymin - 1, xmin - 1);
        float nh = std::min(v0, v1) < 3 ? 0.f : (float)(.5 * (double)(1.f - dot3(n0, n1)));
        if (std::min(v0, v1) > 3) {
          const F3& a = X[l - 1];
          const F3& b = X[l];
          const F3 dx = normalized(F3{a.x - b.x, a.y - b.y, a.z - b.z});
          const F3 mdx{-dx.x, -dx.y, -dx.z};
          const double convexity = (double)((dot3(n0, dx) + dot3(n1, mdx)) / 4) + .5;
          nh = (float)((double)nh * (1 - convexity));
        }
        nH[l] = nh;
      }
      if (i > 0) {
        auto v1f = [&](const std::vector<float>& a, int k) {
          return ((II(a, ymax, xmax, k) - II(a, i - 1, xmax, k)) - II(a, ymax, xmin - 1, k)) + II(a, i - 1, xmin - 1, k);
        };
        auto v0f = [&](const std::vector<float>& a, int k) {
          return ((II(a, i - 1, xmax, k) - II(a, i - 1, xmin - 1, k)) - II(a, ymin - 1, xmax, k)) +
                 II(a, ymin - 1, xmin - 1, k);
        };
        float s = 0.f;
        for (int k = 0; k < 3; k++) s = s + sqrf(v1f(ic, k) / vs1 - v0f(ic, k) / vs0);
        cV[l] = std::sqrt(s);
        F3 n0{v0f(in, 0), v0f(in, 1), v0f(in, 2)}, n1{v1f(in, 0), v1f(in, 1), v1f(in, 2)};
        n0 = normalized(n0);
        n1 = normalized(n1);
        const unsigned v1 = ((IV(ymax, xmax) - IV(i - 1, xmax)) - IV(ymax, xmin - 1)) + IV(i - 1, xmin - 1);
        const unsigned v0 = ((IV(i - 1, xmax) - IV(i - 1, xmin - 1)) - IV(ymin - 1, xmax)) + IV(ymin - 1, xmin - 1);
        float nv = std::min(v0, v1) < 3 ? 0.f : (float)(.5 * (double)(1.f - dot3(n0, n1)));
        if (std::min(v0, v1) > 3) {
          const F3& a = X[l - W];
          const F3& b = X[l];
          const F3 dx = normalized(F3{a.x - b.x, a.y - b.y, a.z - b.z});
          const F3 mdx{-dx.x, -dx.y, -dx.z};
          const double convexity = (double)((dot3(n0, dx) + dot3(n1, mdx)) / 4) + .5;
          nv = (float)((double)nv * (1 - convexity));
        }
        nV[l] = nv;
      }
    }
  // CALC (heuristicBoundaries.cpp:232-275)
  const float depthSigma = .02f, colSigma = .7f, normSigma = .4f;
  std::vector<float> bp(2 * N, 0.f);
  auto calc = [&](int i, int j, int k, float z1, float z2, float dcol, float dnorm) {
    float colProb, depthProb;
    if (std::max(z1, z2) <= 0) {
      depthProb = 0;
      colProb = 1 - hpb_exp(-sqrf(dcol / colSigma));
    } else if (std::min(z1, z2) <= 0) {
      const int di[4] = {0, -1, 0, 1}, dj[4] = {-1, 0, 1, 0};
      const int dir = (z1 <= 0) ? (k + 2) % 4 : k;
      int i0 = (z1 <= 0) ? i : i + di[k], j0 = (z1 <= 0) ? j : j + dj[k];
      unsigned size = 0;
      while (j0 >= 0 && j0 < W && i0 >= 0 && i0 < H && (float)depth_in[(size_t)i0 * W + j0] <= 0) {
        size++;
        i0 += di[dir];
        j0 += dj[dir];
      }
      const bool oob = j0 < 0 || j0 >= W || i0 < 0 || i0 >= H;
      if (oob) depthProb = 0;
      else {
        const double a = std::min(1.0f, (float)size / 30);
        depthProb = (float)((1 - a) * .2 + a * 1);  // linterp (mathUtils.cpp:23-26)
      }
      colProb = 1 - hpb_exp(-sqrf(dcol / colSigma));
    } else {
      const float dz = (float)(std::fabs(z1 - z2) / stereo_error_ratio(std::min(z1, z2)));
      depthProb = 1 - hpb_exp(-sqrf(dz / depthSigma));
      colProb = 0;
    }
    const float normProb = 1 - hpb_exp(-sqrf(dnorm / normSigma));
    auto mx = [](float a, float b) { return (a < b) ? b : a; };  // std::max
    float p = mx(depthProb, mx(colProb, mx(0.0f, normProb)));
    if (((i == 0 || i == H - 2) && k == 1) || ((j == 0 || j == W - 2) && k == 0)) p = 0;
    bp[((size_t)i * W + j) * 2 + k] = p;
  };
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      const size_t l = (size_t)i * W + j;
      if (j < W - 1) calc(i, j, 0, z[l], z[l + 1], cH[l + 1], nH[l + 1]);
      if (i < H - 1) calc(i, j, 1, z[l], z[l + W], cV[l + W], nV[l + W]);
    }
  // non-maximum suppression (heuristicBoundaries.cpp:309-353)
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      const size_t l = (size_t)i * W + j;
      const int hw = rad[l];
      float* o = out + 2 * l;
      o[0] = bp[2 * l];
      o[1] = bp[2 * l + 1];
      int mi = -1;
      for (int k = std::max(0, i - hw); k <= std::min(H - 1, i + hw); k++)
        if (mi < 0 || bp[((size_t)k * W + j) * 2 + 1] > bp[((size_t)mi * W + j) * 2 + 1]) mi = k;
      if (mi != i) o[1] = 0;
      mi = -1;
      for (int k = std::max(0, j - hw); k <= std::min(W - 1, j + hw); k++)
        if (mi < 0 || bp[((size_t)i * W + k) * 2] > bp[((size_t)i * W + mi) * 2]) mi = k;
      if (mi != j) o[0] = 0;
    }
}

float orc_hpb_exp(float x) { return hpb_exp(x); }
void orc_hpb_smallest_eigvec(const double* a9, double* n3) {
  double a[3][3];
  for (int i = 0; i < 9; i++) a[i / 3][i % 3] = a9[i];
  smallest_eigvec(a, n3);
}

}  // extern "C"
