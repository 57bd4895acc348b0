/*
 * oracle.cpp -- CPU restatement of rgbd-flow's coarse-to-fine scene-flow path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product library never links,
 * loads or calls this file.  Built by oracle/Makefile with g++ and
 * -ffp-contract=off so every product and sum is rounded exactly where the
 * C++ expression says.
 *
 * Two precisions:
 *   ORC_PREC_REF  the reference's storage: images, derivatives, products,
 *                 Laplacians and u in f64; the system, du, q, the device
 *                 weight copies and the "sqrCU" squares in f32
 *                 (src/sceneFlowSingleScaleJacobiCUDA.cu:44-47,680-747).
 *   ORC_PREC_F64  the same algorithm with every buffer and product in f64.
 *
 * Parity status of each function is recorded in tests/test_oracle_*.py
 * (closed forms, invariants, brute force).  The reference has no tests and
 * no golden vectors, so nothing here is compared with reference output.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

using std::vector;

/* Row partitions on a persistent pool of std::threads, for the CPU baseline
 * on all cores.  Ranges follow src/parallelism.cpp:24-33 partitionEvenly
 * (n rows into k contiguous parts, the first n mod k one row longer).  The
 * level solve's loops write disjoint rows and its reductions are maxima, so
 * results do not depend on the thread count.  Workers spin (yielding) on a
 * generation counter between jobs: the solve dispatches thousands of short
 * loops per pair.  One caller at a time. */
class RowPool {
 public:
  ~RowPool() { set_threads(1); }
  int threads() const { return nt_; }
  void set_threads(int n) {
    n = std::max(1, std::min(n, 256));
    if (n == nt_) return;
    quit_.store(true);
    for (auto& t : th_) t.join();
    th_.clear();
    quit_.store(false);
    gen_.store(0);
    nt_ = n;
    for (int k = 1; k < nt_; k++) th_.emplace_back([this, k] { worker(k); });
  }
  /* f(r0, r1, part) over the partitions of [0, n) */
  void rows(int n, const std::function<void(int, int, int)>& f) {
    if (nt_ == 1 || n < 2 * nt_) {
      f(0, n, 0);
      return;
    }
    n_ = n;
    job_ = &f;
    left_.store(nt_ - 1);
    gen_.fetch_add(1, std::memory_order_release);
    run_part(0);
    while (left_.load(std::memory_order_acquire) != 0) std::this_thread::yield();
  }

 private:
  void run_part(int k) const {
    const int q = n_ / nt_, r = n_ % nt_;
    const int r0 = k * q + std::min(k, r), r1 = r0 + q + (k < r ? 1 : 0);
    (*job_)(r0, r1, k);
  }
  void worker(int k) {
    int seen = 0;
    for (;;) {
      int g;
      while ((g = gen_.load(std::memory_order_acquire)) == seen) {
        if (quit_.load()) return;
        std::this_thread::yield();
      }
      seen = g;
      run_part(k);
      left_.fetch_sub(1, std::memory_order_acq_rel);
    }
  }
  int nt_ = 1, n_ = 0;
  const std::function<void(int, int, int)>* job_ = nullptr;
  std::vector<std::thread> th_;
  std::atomic<int> gen_{0}, left_{0};
  std::atomic<bool> quit_{false};
};

RowPool& pool() {
  static RowPool p;
  return p;
}

/* project.h:29-39 -- __min/__max return the FIRST argument's type */
template <class A, class B> A liu_min(A a, B b) { return (a > b) ? (A)b : a; }
template <class A, class B> A liu_max(A a, B b) { return (a < b) ? (A)b : a; }
/* ImageProcessing.h:25 */
inline int enforce_range(int x, int maxv) { return liu_min(liu_max(x, 0), maxv - 1); }

inline double dsq(double x) { return x * x; }

/* primesensorUtils.cpp:53-64 */
double stereo_error(double z) { return dsq(z) / (1 / .00285 + z); }
double stereo_error_ratio(double z) {
  static const double err1 = stereo_error(1);
  return stereo_error(z) / err1;
}

/* --------------------------------------------------------------------------
 * ImageProcessing.h:128-148 BilinearInterpolate (accumulating into result)
 * Truncating base, f64 weights clamped to [0,1], taps in (m,n) order
 * (0,0),(0,1),(1,0),(1,1).
 */
void bilerp_acc(const double* img, int W, int H, int C, double x, double y, double* result) {
  int xx = (int)x, yy = (int)y;
  double dx = liu_max(liu_min(x - xx, 1), 0);
  double dy = liu_max(liu_min(y - yy, 1), 0);
  for (int m = 0; m <= 1; m++)
    for (int n = 0; n <= 1; n++) {
      int uu = enforce_range(xx + m, W);
      int vv = enforce_range(yy + n, H);
      int off = (vv * W + uu) * C;
      double s = std::fabs(1 - m - dx) * std::fabs(1 - n - dy);
      for (int l = 0; l < C; l++) result[l] += img[off + l] * s;
    }
}

/* ImageProcessing.h:203-222 ResizeImage(ratio) */
void resize_ratio(const double* src, int W, int H, int C, double ratio, double* dst) {
  int DW = (double)W * ratio, DH = (double)H * ratio;
  std::fill(dst, dst + (size_t)DW * DH * C, 0.0);
  for (int i = 0; i < DH; i++)
    for (int j = 0; j < DW; j++) {
      double x = (double)(j + 1) / ratio - 1;
      double y = (double)(i + 1) / ratio - 1;
      bilerp_acc(src, W, H, C, x, y, dst + ((size_t)i * DW + j) * C);
    }
}

/* ImageProcessing.h:224-242 ResizeImage(dst dims) */
void resize_dims(const double* src, int Ws, int Hs, int C, int Wd, int Hd, double* dst) {
  double xr = (double)Wd / Ws, yr = (double)Hd / Hs;
  std::fill(dst, dst + (size_t)Wd * Hd * C, 0.0);
  for (int i = 0; i < Hd; i++)
    for (int j = 0; j < Wd; j++) {
      double x = (double)(j + 1) / xr - 1;
      double y = (double)(i + 1) / yr - 1;
      bilerp_acc(src, Ws, Hs, C, x, y, dst + ((size_t)i * Wd + j) * C);
    }
}

/* ImageProcessing.h:247-269 hfiltering (clamp-to-edge, taps l=-f..f in order) */
void hfilter(const double* src, double* dst, int W, int H, int C, const double* f, int fs) {
  std::fill(dst, dst + (size_t)W * H * C, 0.0);
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      size_t off = (size_t)i * W * C;
      double* b = dst + off + (size_t)j * C;
      for (int l = -fs; l <= fs; l++) {
        double wgt = f[l + fs];
        int jj = enforce_range(j + l, W);
        for (int k = 0; k < C; k++) b[k] += src[off + (size_t)jj * C + k] * wgt;
      }
    }
}

/* ImageProcessing.h:339-358 vfiltering */
void vfilter(const double* src, double* dst, int W, int H, int C, const double* f, int fs) {
  std::fill(dst, dst + (size_t)W * H * C, 0.0);
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      double* b = dst + ((size_t)i * W + j) * C;
      for (int l = -fs; l <= fs; l++) {
        double wgt = f[l + fs];
        int ii = enforce_range(i + l, H);
        for (int k = 0; k < C; k++) b[k] += src[((size_t)ii * W + j) * C + k] * wgt;
      }
    }
}

/* Image.h:1179-1199 GaussianSmoothing + Image.h:1321-1330 imfilter_hv */
void gaussian_smooth(const double* img, int W, int H, int C, double sigma, int fsize, double* out) {
  vector<double> g(2 * fsize + 1);
  double sum = 0;
  double s2 = sigma * sigma * 2;
  for (int i = -fsize; i <= fsize; i++) {
    g[i + fsize] = std::exp(-(double)(i * i) / s2);
    sum += g[i + fsize];
  }
  for (int i = 0; i < 2 * fsize + 1; i++) g[i] /= sum;
  vector<double> tmp((size_t)W * H * C);
  hfilter(img, tmp.data(), W, H, C, g.data(), fsize);
  vfilter(tmp.data(), out, W, H, C, g.data(), fsize);
}

struct Level { int W, H; vector<double> d; };

/* GaussianPyramid.cpp:20-63 */
vector<Level> gaussian_pyramid(const double* img, int W, int H, int C, double ratio, int min_width) {
  if (ratio > 0.98 || ratio < 0.4) ratio = 0.75;
  double w = W * ratio;
  int nLevels = 1;
  while (w >= min_width) { nLevels++; w *= ratio; }
  vector<Level> P(nLevels);
  P[0].W = W; P[0].H = H; P[0].d.assign(img, img + (size_t)W * H * C);
  double baseSigma = (1 / ratio - 1);
  int n = std::log(0.25) / std::log(ratio);
  double nSigma = baseSigma * n;
  for (int i = 1; i < nLevels; i++) {
    if (i <= n) {
      double sigma = baseSigma * i;
      vector<double> foo((size_t)W * H * C);
      gaussian_smooth(img, W, H, C, sigma, (int)(sigma * 3), foo.data());
      double r = std::pow(ratio, i);
      P[i].W = (double)W * r; P[i].H = (double)H * r;
      P[i].d.resize((size_t)P[i].W * P[i].H * C);
      resize_ratio(foo.data(), W, H, C, r, P[i].d.data());
    } else {
      const Level& src = P[i - n];
      vector<double> foo((size_t)src.W * src.H * C);
      gaussian_smooth(src.d.data(), src.W, src.H, C, nSigma, (int)(nSigma * 3), foo.data());
      double rate = (double)std::pow(ratio, i) * W / src.W;
      P[i].W = (double)src.W * rate; P[i].H = (double)src.H * rate;
      P[i].d.resize((size_t)P[i].W * P[i].H * C);
      resize_ratio(foo.data(), src.W, src.H, C, rate, P[i].d.data());
    }
  }
  return P;
}

/* depthMapPyramid.cpp:15-140 (validity-aware 2x2 component average) */
vector<Level> depth_map_pyramid(const double* img, int W, int H, double ratio, int min_width) {
  if (ratio > 0.98 || ratio < 0.4) ratio = 0.75;
  double wd = W * ratio;
  int nLevels = 1;
  while (wd >= min_width) { nLevels++; wd *= ratio; }
  vector<Level> P(nLevels);
  P[0].W = W; P[0].H = H; P[0].d.assign(img, img + (size_t)W * H);
  for (int m = 1; m < nLevels; m++) {
    const Level& L = P[m - 1];
    const int LW = L.W, LH = L.H;
    auto at = [&](int y, int x) { return L.d[(size_t)y * LW + x]; };
    double lr = std::pow(ratio, m);
    int FW = W * lr, FH = H * lr;
    vector<double> foo((size_t)FW * FH, 0.0);
    const float factor = (float)LH / FH;
    const float maxDZ = .005;
    for (int i = 0; i < FH; i++)
      for (int j = 0; j < FW; j++) {
        const int x = std::max(0, std::min(LW - 1, (int)std::rint(factor * j)));
        const int y = std::max(0, std::min(LH - 1, (int)std::rint(factor * i)));
        double& o = foo[(size_t)i * FW + j];
        if (at(y, x) > 0) {
          int xs[2], ys[2];
          xs[0] = std::max(0, std::min(LW - 1, (int)std::floor(factor * j)));
          xs[1] = std::min(LW - 1, xs[0] + 1);
          ys[0] = std::max(0, std::min(LH - 1, (int)std::floor(factor * i)));
          ys[1] = std::min(LH - 1, ys[0] + 1);
          const int xi = (x == xs[0]) ? 0 : 1, yi = (y == ys[0]) ? 0 : 1;
          const double c = at(y, x);
          vector<float> comp{(float)c};
          if (std::fabs(at(ys[yi], xs[!xi]) - c) < maxDZ * stereo_error_ratio(c)) comp.push_back(at(ys[yi], xs[!xi]));
          if (std::fabs(at(ys[!yi], xs[xi]) - c) < maxDZ * stereo_error_ratio(c)) comp.push_back(at(ys[!yi], xs[xi]));
          if (std::fabs(at(ys[!yi], xs[!xi]) - c) < maxDZ * stereo_error_ratio(c)) comp.push_back(at(ys[!yi], xs[!xi]));
          if (comp.size() > 1) {
            float avg = 0;
            for (float z : comp) avg += z;
            o = avg / comp.size();
          } else {
            const double c2 = at(ys[!yi], xs[!xi]);
            vector<float> other{(float)c2};
            if (std::fabs(at(ys[yi], xs[!xi]) - c2) < maxDZ * stereo_error_ratio(c2)) other.push_back(at(ys[yi], xs[!xi]));
            if (std::fabs(at(ys[!yi], xs[xi]) - c2) < maxDZ * stereo_error_ratio(c2)) other.push_back(at(ys[!yi], xs[xi]));
            if (other.size() > 1) {
              float avg = 0;
              for (float z : other) avg += z;
              o = avg / other.size();
            } else {
              o = c;
            }
          }
        } else {
          o = -1;
        }
      }
    P[m].W = FW; P[m].H = FH; P[m].d.swap(foo);
  }
  return P;
}

/* The feature-image storage variants: the value rounded to f32, then to FP16
 * (1 sign, 5 exponent, 10 fraction bits; subnormals below 2^-14, quantum
 * 2^-24; overflow to infinity) or BF16 (the f32's top 16 bits), each step
 * to nearest, ties to even.  Scaling by a power of two is exact, so
 * nearbyint (round-half-even in the default mode) on x / quantum rounds the
 * f32 value to the grid of FP16 numbers around it. */
static double img_round(double xd, int storage) {
  if (storage == ORC_IMG_F64) return xd;
  const float x = (float)xd;
  if (storage == ORC_IMG_BF16) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return xd;  // NaN
    u = ((u + 0x7fffu + ((u >> 16) & 1u)) >> 16) << 16;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
  }
  if (!(std::fabs(x) <= 3.4e38f) || x == 0.f) return x;  // zero, infinity, NaN as they are
  int e;
  std::frexp((double)x, &e);                               // |x| in [2^(e-1), 2^e)
  const double q = std::ldexp(1.0, std::max(e - 1, -14) - 10);
  const double r = std::nearbyint((double)x / q) * q;
  return std::fabs(r) > 65504.0 ? std::copysign(INFINITY, r) : r;
}
/* The relaxation's per-pixel systems stored in 16 bits (a precision variant,
 * not in the reference): BF16 of each f32 coefficient, or FP16 with a
 * per-pixel power-of-two scale -- the 13 coefficients times 2^-e with
 * e = ilogb(max |c|) - 14 (clamped to [-100, 100]; 0 when the largest is 0
 * or not finite), so the largest lands in [2^14, 2^15) inside FP16's range,
 * rounded to FP16 and decoded as value times 2^e.  Both round to nearest,
 * ties to even.  c: the 13 coefficients of one pixel, rounded in place. */
static void record_round(float* c, int storage) {
  if (storage == ORC_REC_BF16) {
    for (int k = 0; k < 13; k++) c[k] = (float)img_round(c[k], ORC_IMG_BF16);
    return;
  }
  float m = 0.f;
  for (int k = 0; k < 13; k++) m = std::max(m, std::fabs(c[k]));
  int e = 0;
  if (m > 0.f && m <= 3.4e38f) e = std::min(std::max(std::ilogb(m) - 14, -100), 100);
  const float down = std::ldexp(1.0f, -e), up = std::ldexp(1.0f, e);
  for (int k = 0; k < 13; k++) c[k] = (float)img_round((double)(c[k] * down), ORC_IMG_F16) * up;
}

static void img_round_all(double* a, size_t n, int storage) {
  if (storage != ORC_IMG_F64)
    for (size_t i = 0; i < n; i++) a[i] = img_round(a[i], storage);
}

/* OpticalFlow.cpp:148-220 im2feature(RGB_WITH_GRADIENT); Image.h:955-1016 dx/dy
 * with the 5-tap [1,-8,0,8,-1]/12 filter; Image.h:1435-1468 desaturate. */
void im2feature(const double* im, int W, int H, double* out) {
  size_t N = (size_t)W * H;
  vector<double> gray(N), gx(N), gy(N);
  for (size_t i = 0; i < N; i++) {
    size_t o = i * 3;
    gray[i] = (double)im[o] * .114 + im[o + 1] * .587 + im[o + 2] * .299;
  }
  double f[5] = {1, -8, 0, 8, -1};
  for (int i = 0; i < 5; i++) f[i] /= 12;
  hfilter(gray.data(), gx.data(), W, H, 1, f, 2);
  vfilter(gray.data(), gy.data(), W, H, 1, f, 2);
  for (size_t i = 0; i < N; i++) {
    out[i * 5 + 0] = gray[i];
    out[i * 5 + 1] = gx[i];
    out[i * 5 + 2] = gy[i];
    out[i * 5 + 3] = im[i * 3 + 1] - im[i * 3];
    out[i * 5 + 4] = im[i * 3 + 1] - im[i * 3 + 2];
  }
}

/* opticalFlowUtils.cpp:76-99 warpBy3DFlow (bilinear branch) via
 * ImageProcessing.h:471-491 warpImage: out-of-bounds -> frame-1 value. */
void warp_cpu(const double* I1, const double* D1, const double* I2, const double* D2,
              const double* u, const double* v, const double* w, int W, int H, int C,
              double* wI, double* wD) {
  std::fill(wI, wI + (size_t)W * H * C, 0.0);
  std::fill(wD, wD + (size_t)W * H, 0.0);
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      size_t l = (size_t)i * W + j;
      double y = i + v[l], x = j + u[l];
      if (x < 0 || x > W - 1 || y < 0 || y > H - 1) {
        for (int k = 0; k < C; k++) wI[l * C + k] = I1[l * C + k];
        wD[l] = D1[l];
      } else {
        bilerp_acc(I2, W, H, C, x, y, wI + l * C);
        bilerp_acc(D2, W, H, 1, x, y, wD + l);
      }
    }
  for (size_t l = 0; l < (size_t)W * H; l++) wD[l] -= w[l];
}

/* recursiveMedianFiltering.cpp:34-99 -- greedy fill in order of most known
 * neighbours; the visiting order is std::unordered_set's (same container, same
 * insert/erase sequence as the reference); values pass through a float buffer. */
void median_fill(double* data, int W, int H, int hw) {
  const int win = (2 * hw + 1) * (2 * hw + 1);
  vector<std::unordered_set<uint32_t>> byCount(win);
  vector<int32_t> countOf((size_t)W * H, -1);
  uint32_t left = 0;
  for (int i = 0, l = 0; i < H; i++)
    for (int j = 0; j < W; j++, l++)
      if (data[l] <= 0) {
        uint32_t c = 0;
        for (int ii = std::max(0, i - hw); ii <= std::min(H - 1, i + hw); ii++)
          for (int jj = std::max(0, j - hw); jj <= std::min(W - 1, j + hw); jj++)
            if (data[ii * W + jj] > 0) c++;
        byCount[c].insert(l);
        countOf[l] = c;
        left++;
      }
  vector<float> vals(win);
  while (left > 0) {
    int mx = (int)byCount.size() - 1;
    while (mx >= 0 && byCount[mx].empty()) mx--;
    if (mx <= 0) return; /* reference asserts maxNumNbrs > 0 (an all-invalid region) */
    const uint32_t l = *byCount[mx].begin();
    const int i = l / W, j = l % W;
    byCount[mx].erase(l);
    countOf[l] = -1;
    uint32_t idx = 0;
    for (int ii = std::max(0, i - hw); ii <= std::min(H - 1, i + hw); ii++)
      for (int jj = std::max(0, j - hw); jj <= std::min(W - 1, j + hw); jj++) {
        const int ll = ii * W + jj;
        if (data[ll] > 0) {
          vals[idx++] = data[ll];
        } else if (countOf[ll] >= 0) {
          byCount[countOf[ll]].erase(ll);
          countOf[ll]++;
          byCount[countOf[ll]].insert(ll);
        }
      }
    std::sort(vals.begin(), vals.begin() + idx);
    data[l] = vals[(int)std::floor(idx / 2)];
    left--;
  }
}

/* The device's parallel variant of median_fill (halfwidth 1), restated:
 * rounds r = 1, 2, ...; in each, M = the largest number of known 8-neighbours
 * (valid, or filled in an earlier round) over the pixels still invalid, and
 * every such pixel with M known neighbours is filled at once with the upper
 * median of their values (sorted as float, vals[n / 2], as median_fill);
 * stops when none is left or M is 0.  Equal to median_fill wherever no two
 * pixels filled in one round are neighbours (isolated holes); elsewhere the
 * sequential fill's order (which depends on a hash container) is replaced by
 * whole rounds.  Returns the number of pixels left invalid. */
int median_fill_rounds(double* d, int W, int H) {
  const size_t N = (size_t)W * H;
  const int kUnfilled = 0x7fffffff;
  vector<int> st(N);
  vector<size_t> todo;
  for (size_t l = 0; l < N; l++) {
    st[l] = d[l] > 0 ? 0 : kUnfilled;
    if (d[l] <= 0) todo.push_back(l);
  }
  auto known = [&](int i, int j, int r, float* vals) {
    int n = 0;
    for (int ii = std::max(0, i - 1); ii <= std::min(H - 1, i + 1); ii++)
      for (int jj = std::max(0, j - 1); jj <= std::min(W - 1, j + 1); jj++) {
        const size_t q = (size_t)ii * W + jj;
        if (st[q] < r) {
          if (vals) vals[n] = (float)d[q];
          n++;
        }
      }
    return n;
  };
  size_t left = todo.size();
  for (int r = 1; left > 0; r++) {
    int M = 0;
    for (size_t l : todo)
      if (st[l] == kUnfilled) M = std::max(M, known((int)(l / W), (int)(l % W), r, nullptr));
    if (M == 0) break;
    vector<std::pair<size_t, double>> fills;
    for (size_t l : todo) {
      if (st[l] != kUnfilled) continue;
      float vals[9];
      const int n = known((int)(l / W), (int)(l % W), r, vals);
      if (n != M) continue;
      std::sort(vals, vals + n);
      fills.push_back({l, (double)vals[n / 2]});
    }
    for (const auto& f : fills) {
      d[f.first] = f.second;
      st[f.first] = r;
    }
    left -= fills.size();
  }
  return (int)left;
}

/* exp(x) of the device's colour fill, the same operations: Cody-Waite
 * reduction by ln 2, Taylor series to degree 13 in f64, times 2^n. */
static double fill_exp(double x) {
  if (!(x > -700.0)) return 0.0;
  const double n = std::rint(x * 1.4426950408889634);
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
  return std::ldexp(p, (int)n);
}

/* The device's parallel variant of the colour-guided fill of small holes
 * (recursiveMedianFiltering.cpp:127-222, halfwidth 1), restated: the
 * 4-connected invalid regions of at most max_region pixels are filled
 * together in rounds; in each, a region's pixels with the region's largest
 * number of known 3x3 neighbours (valid, or filled in an earlier round; at
 * least one) are filled at once with the mean of those neighbours' depths
 * weighted by exp(-|colour difference|^2 / 0.6) (float sums, in window
 * order, as the reference); ends when a round fills nothing.  bgr: [3][N]
 * planes.  Returns the pixels of small regions left invalid. */
int colour_fill_rounds(double* d, const double* bgr, int W, int H, int max_region) {
  const size_t N = (size_t)W * H;
  const int kUnfilled = 0x7fffffff;
  // regions: 4-connected invalid pixels, labelled by their smallest index
  vector<int> label(N, -2);
  vector<int> stack;
  for (size_t l0 = 0; l0 < N; l0++) {
    if (!(d[l0] <= 0) || label[l0] != -2) continue;
    vector<int> comp;
    stack.assign(1, (int)l0);
    label[l0] = (int)l0;
    while (!stack.empty()) {
      const int p = stack.back();
      stack.pop_back();
      comp.push_back(p);
      const int i = p / W, j = p % W;
      const int nb[4] = {i > 0 ? p - W : -1, j > 0 ? p - 1 : -1, j < W - 1 ? p + 1 : -1, i < H - 1 ? p + W : -1};
      for (int m : nb)
        if (m >= 0 && d[m] <= 0 && label[m] == -2) {
          label[m] = (int)l0;
          stack.push_back(m);
        }
    }
    if ((int)comp.size() > max_region)
      for (int p : comp) label[p] = -1;
  }
  vector<int> st(N);
  vector<size_t> todo;
  for (size_t l = 0; l < N; l++) {
    st[l] = d[l] > 0 ? 0 : kUnfilled;
    if (label[l] >= 0) todo.push_back(l);
  }
  auto known = [&](size_t l, int r) {
    const int i = (int)(l / W), j = (int)(l % W);
    int n = 0;
    for (int ii = std::max(0, i - 1); ii <= std::min(H - 1, i + 1); ii++)
      for (int jj = std::max(0, j - 1); jj <= std::min(W - 1, j + 1); jj++) n += st[(size_t)ii * W + jj] < r;
    return n;
  };
  size_t left = todo.size();
  for (int r = 1; left > 0; r++) {
    std::unordered_map<int, int> cmax;
    for (size_t l : todo)
      if (st[l] == kUnfilled) cmax[label[l]] = std::max(cmax[label[l]], known(l, r));
    vector<std::pair<size_t, double>> fills;
    for (size_t l : todo) {
      if (st[l] != kUnfilled) continue;
      const int k = known(l, r);
      if (k == 0 || k != cmax[label[l]]) continue;
      const int i = (int)(l / W), j = (int)(l % W);
      float zs = 0.f, ws = 0.f;
      for (int ii = std::max(0, i - 1); ii <= std::min(H - 1, i + 1); ii++)
        for (int jj = std::max(0, j - 1); jj <= std::min(W - 1, j + 1); jj++) {
          const size_t q = (size_t)ii * W + jj;
          if (st[q] < r) {
            const double e0 = bgr[l] - bgr[q], e1 = bgr[N + l] - bgr[N + q], e2 = bgr[2 * N + l] - bgr[2 * N + q];
            const float w = (float)fill_exp(-(e0 * e0 + e1 * e1 + e2 * e2) / .6);
            zs = (float)((double)zs + d[q] * w);
            ws += w;
          }
        }
      fills.push_back({l, (double)(zs / ws)});
    }
    if (fills.empty()) break;
    for (const auto& f : fills) {
      d[f.first] = f.second;
      st[f.first] = r;
    }
    left -= fills.size();
  }
  return (int)left;
}

/* staticDepthNoiseModeling.cpp:85-113: depth-boundary mask, then the exact L1
 * distance to the nearest boundary pixel (what cv::distanceTransform(L1)
 * returns), capped: distances above `cap` are reported as cap+1. */
void l1_distance_to_boundary(const float* depth, int W, int H, int cap, float* dist) {
  vector<uint8_t> boundary((size_t)W * H, 0);
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      const float d = depth[(size_t)i * W + j];
      bool b = false;
      if (i < H - 1) {
        const float dn = depth[(size_t)(i + 1) * W + j];
        if (dn > 0 && std::fabs(d - dn) > .02 * stereo_error_ratio(std::min(d, dn))) b = true;
      }
      if (!b && j < W - 1) {
        const float dn = depth[(size_t)i * W + j + 1];
        if (dn > 0 && std::fabs(d - dn) > .02 * stereo_error_ratio(std::min(d, dn))) b = true;
      }
      boundary[(size_t)i * W + j] = b;
    }
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      int best = cap + 1;
      for (int di = -cap; di <= cap; di++) {
        const int ii = i + di;
        if (ii < 0 || ii >= H) continue;
        const int rem = cap - std::abs(di);
        for (int dj = -rem; dj <= rem; dj++) {
          const int jj = j + dj;
          if (jj < 0 || jj >= W) continue;
          if (boundary[(size_t)ii * W + jj]) best = std::min(best, std::abs(di) + std::abs(dj));
        }
      }
      dist[(size_t)i * W + j] = (float)best;
    }
}

/* staticDepthNoiseModeling.cpp:225-258 computeDepthMapStdevsPointwise with the
 * sigma at 1 m passed as float (:227,232), then :115-124 inflation. */
void depth_stdevs_pointwise(const float* depth, int W, int H, int max_dist, float* st) {
  const float sigmaAt1m = stereo_error(1);
  const float invalidSigma = 1;
  for (size_t l = 0; l < (size_t)W * H; l++)
    st[l] = depth[l] > 0 ? (float)std::max(1e-4, sigmaAt1m * stereo_error_ratio(depth[l])) : invalidSigma;
  if (max_dist > 0) {
    vector<float> dist((size_t)W * H);
    l1_distance_to_boundary(depth, W, H, max_dist, dist.data());
    for (size_t l = 0; l < (size_t)W * H; l++)
      if (dist[l] <= max_dist) {
        const float a = dist[l] / max_dist;
        st[l] = std::exp((1 - a) * std::log((double)(10 * st[l])) + a * std::log((double)st[l]));
      }
  }
}

/* ==========================================================================
 * The level solve.  CT is the storage type of the system and of du, WT the
 * type of the device weight copies (float in the reference); FSQ selects the
 * reference's float squares (sqrCU, .cu:44-47).
 */
/* ST: the arithmetic of the per-outer stage (derivatives, occlusion, penalty
 * derivatives, Laplacians, assembly and the in-solve warp).  double: the
 * reference's (f64 with its f32 roundings); float (with REF, the 16-bit image
 * variants): every stage formula in f32 on f32-rounded inputs, in the same
 * order -- not in the reference. */
template <class CT, bool REF, class ST = double>
struct Solver {
  typedef typename std::conditional<REF, float, double>::type WT;
  static constexpr bool F32 = std::is_same<ST, float>::value;

  static inline double sq(double x) {
    if (REF) { float f = (float)x; return (double)(float)(f * f); }
    return x * x;
  }

  int W, H;
  size_t N;
  static const int C = 5;
  const orc_level_in* in;
  const orc_params* p;

  vector<double> Iw, Dw;                 /* warped frame 2 */
  vector<double> u;                      /* [3][N] planes */
  vector<CT> du;                         /* [3][N] planes */
  vector<WT> dW, sW, mW, rW;             /* device weight copies */
  vector<WT> cW;                         /* convergence weights f/D1 */
  vector<double> Ix, Iy, It, Dx, Dy, Dt; /* derivatives */
  vector<WT> q;
  vector<double> phi, psi, pi;
  vector<CT> coef;                       /* 13 planes */

  Solver(const orc_level_in* in_, const orc_params* p_) : in(in_), p(p_) {
    W = in->W; H = in->H; N = (size_t)W * H;
    Iw.assign(in->warpI2, in->warpI2 + N * C);
    Dw.assign(in->warpD2, in->warpD2 + N);
    u.assign(3 * N, 0.0);
    du.assign(3 * N, (CT)0);
    dW.resize(N); sW.resize(N); mW.resize(N); rW.resize(2 * N); cW.resize(N);
    for (size_t l = 0; l < N; l++) {
      dW[l] = (WT)in->delta[l]; sW[l] = (WT)in->s[l]; mW[l] = (WT)in->m[l];
      rW[2 * l] = (WT)in->r[2 * l]; rW[2 * l + 1] = (WT)in->r[2 * l + 1];
      cW[l] = (WT)(in->focal_level / in->D1[l]);
    }
    Ix.resize(N * C); Iy.resize(N * C); It.resize(N * C);
    Dx.resize(N); Dy.resize(N); Dt.resize(N); q.resize(N);
    phi.resize(N); psi.resize(N * C); pi.resize(N);
    coef.assign(13 * N, (CT)0);
  }

  /* .cu:84-147 imgDerivativesKernel / ...WithClamp */
  void derivatives() {
    if (F32) { derivatives_f32(); return; }
    static const double g[5] = {1, -8, 0, 8, -1};
    const double* I1 = in->I1;
    pool().rows(H, [&](int i0, int i1, int) {
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < W; j++) {
        const size_t l = (size_t)i * W + j;
        for (int k = 0; k < C; k++) {
          double sx = 0, sy = 0;
          for (int c = j - 2, m = 0; c <= j + 2; c++, m++) {
            const int cc = std::max(0, std::min(W - 1, c));
            const size_t o = ((size_t)i * W + cc) * C + k;
            sx += g[m] * (.4 * I1[o] + .6 * Iw[o]);
          }
          for (int r = i - 2, m = 0; r <= i + 2; r++, m++) {
            const int rr = std::max(0, std::min(H - 1, r));
            const size_t o = ((size_t)rr * W + j) * C + k;
            sy += g[m] * (.4 * I1[o] + .6 * Iw[o]);
          }
          Ix[l * C + k] = sx / 12;
          Iy[l * C + k] = sy / 12;
          It[l * C + k] = Iw[l * C + k] - I1[l * C + k];
        }
        const double mg = .04;
        double sx = 0, sy = 0;
        for (int c = j - 2, m = 0; c <= j + 2; c++, m++) {
          const int cc = std::max(0, std::min(W - 1, c));
          sx += g[m] * (.4 * in->D1[(size_t)i * W + cc] + .6 * Dw[(size_t)i * W + cc]);
        }
        for (int r = i - 2, m = 0; r <= i + 2; r++, m++) {
          const int rr = std::max(0, std::min(H - 1, r));
          sy += g[m] * (.4 * in->D1[(size_t)rr * W + j] + .6 * Dw[(size_t)rr * W + j]);
        }
        Dx[l] = std::min(std::max(sx / 12, -mg), mg);
        Dy[l] = std::min(std::max(sy / 12, -mg), mg);
        Dt[l] = std::min(std::max(Dw[l] - in->D1[l], -mg), mg);
      }
    });
  }

  /* derivatives in f32: stencil inputs .4f I1 + .6f Iw of the f32-rounded
   * values, sums in the stencil order, / 12 and the depth clamps in f32 */
  void derivatives_f32() {
    static const float g[5] = {1, -8, 0, 8, -1};
    const double* I1 = in->I1;
    const double* D1 = in->D1;
    pool().rows(H, [&](int i0, int i1, int) {
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < W; j++) {
        const size_t l = (size_t)i * W + j;
        for (int k = 0; k < C; k++) {
          float sx = 0, sy = 0;
          for (int c = j - 2, m = 0; c <= j + 2; c++, m++) {
            const size_t o = ((size_t)i * W + std::max(0, std::min(W - 1, c))) * C + k;
            sx += g[m] * (.4f * (float)I1[o] + .6f * (float)Iw[o]);
          }
          for (int r = i - 2, m = 0; r <= i + 2; r++, m++) {
            const size_t o = ((size_t)std::max(0, std::min(H - 1, r)) * W + j) * C + k;
            sy += g[m] * (.4f * (float)I1[o] + .6f * (float)Iw[o]);
          }
          Ix[l * C + k] = sx / 12.f;
          Iy[l * C + k] = sy / 12.f;
          It[l * C + k] = (float)Iw[l * C + k] - (float)I1[l * C + k];
        }
        const float mg = .04f;
        float sx = 0, sy = 0;
        for (int c = j - 2, m = 0; c <= j + 2; c++, m++) {
          const size_t o = (size_t)i * W + std::max(0, std::min(W - 1, c));
          sx += g[m] * (.4f * (float)D1[o] + .6f * (float)Dw[o]);
        }
        for (int r = i - 2, m = 0; r <= i + 2; r++, m++) {
          const size_t o = (size_t)std::max(0, std::min(H - 1, r)) * W + j;
          sy += g[m] * (.4f * (float)D1[o] + .6f * (float)Dw[o]);
        }
        Dx[l] = std::min(std::max(sx / 12.f, -mg), mg);
        Dy[l] = std::min(std::max(sy / 12.f, -mg), mg);
        Dt[l] = std::min(std::max((float)Dw[l] - (float)D1[l], -mg), mg);
      }
    });
  }

  /* .cu:640-648 fillOcclusionPenaltyKernel (unwarped D2) */
  void occlusion() {
    pool().rows(H, [&](int i0, int i1, int) {
      for (size_t l = (size_t)i0 * W; l < (size_t)i1 * W; l++) {
        if (F32) {
          const float d1 = (float)in->D1[l];
          q[l] = (d1 + (float)u[2 * N + l] < (float)in->D2[l] + .02f * (d1 * d1)) ? (WT)1 : (WT).01;
          continue;
        }
        const double d1 = in->D1[l];
        q[l] = (d1 + u[2 * N + l] < in->D2[l] + .02 * sq(d1)) ? (WT)1 : (WT).01;
      }
    });
  }

  /* .cu:156-182 forward differences of u+du for component c, pixel (i,j) */
  inline void flow_grad(int c, int i, int j, double& gx, double& gy) const {
    const size_t l = (size_t)i * W + j;
    const double* uc = u.data() + c * N;
    const CT* dc = du.data() + c * N;
    if (F32) {
      gy = (i == H - 1) ? 0.f : ((float)uc[l + W] + (float)dc[l + W] - (float)uc[l] - (float)dc[l]);
      gx = (j == W - 1) ? 0.f : ((float)uc[l + 1] + (float)dc[l + 1] - (float)uc[l] - (float)dc[l]);
      return;
    }
    gy = (i == H - 1) ? 0.0 : (uc[l + W] + (double)dc[l + W] - uc[l] - (double)dc[l]);
    gx = (j == W - 1) ? 0.0 : (uc[l + 1] + (double)dc[l + 1] - uc[l] - (double)dc[l]);
  }

  /* .cu:340-371 penaltyDerivativesKernel */
  void penalties() {
    const double epsD = 1e-6, epsZ = 1e-10, epsS = 1e-6;
    const double cw = p->color_data_weight;
    pool().rows(H, [&](int i0, int i1, int) {
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < W; j++) {
        const size_t l = (size_t)i * W + j;
        double ux, uy, vx, vy, wx, wy;
        flow_grad(0, i, j, ux, uy);
        flow_grad(1, i, j, vx, vy);
        flow_grad(2, i, j, wx, wy);
        if (F32) {
          const float fux = (float)ux, fuy = (float)uy, fvx = (float)vx, fvy = (float)vy, fwx = (float)wx,
                      fwy = (float)wy;
          const float a = fux * fux + fuy * fuy + fvx * fvx + fvy * fvy + (float)sW[l] * (fwx * fwx + fwy * fwy);
          phi[l] = .5f / std::sqrt(a + 1e-6f);
          const float d0 = (float)du[l], d1 = (float)du[N + l], d2 = (float)du[2 * N + l];
          for (int k = 0; k < C; k++) {
            const size_t o = l * C + k;
            const float tt = ((float)It[o] + (float)Ix[o] * d0 + (float)Iy[o] * d1) * (float)q[l];
            psi[o] = .5f / std::sqrt(tt * tt + 1e-6f) * (float)cw;
          }
          const float t2 = (float)Dt[l] + (float)Dx[l] * d0 + (float)Dy[l] * d1 - d2;
          pi[l] = .5f / std::sqrt(t2 * t2 + 1e-10f) * (float)dW[l];
          continue;
        }
        double t;
        if (REF) {
          const float a = (float)sq(ux) + (float)sq(uy) + (float)sq(vx) + (float)sq(vy) +
                          (float)sW[l] * ((float)sq(wx) + (float)sq(wy));
          t = a;
        } else {
          t = sq(ux) + sq(uy) + sq(vx) + sq(vy) + (double)sW[l] * (sq(wx) + sq(wy));
        }
        phi[l] = 0.5 / std::sqrt(t + epsS);
        const double d0 = (double)du[l], d1 = (double)du[N + l], d2 = (double)du[2 * N + l];
        for (int k = 0; k < C; k++) {
          const size_t o = l * C + k;
          const double tt = (It[o] + Ix[o] * d0 + Iy[o] * d1) * (double)q[l];
          psi[o] = .5 / std::sqrt(sq(tt) + epsD) * cw;
        }
        const double t2 = Dt[l] + Dx[l] * d0 + Dy[l] * d1 - d2;
        pi[l] = .5 / std::sqrt(sq(t2) + epsZ) * (double)dW[l];
      }
    });
  }

  /* .cu:240-333 uv/w flow Laplacians (on u only, normalised by the mean
   * regularisation weight of the existing edges) */
  double laplacian(int c, int i, int j) const {
    if (F32) return laplacian_f32(c, i, j);
    const size_t l = (size_t)i * W + j;
    const double* uc = u.data() + c * N;
    double lap = 0;
    if (j < W - 1) lap -= (uc[l + 1] - uc[l]) * phi[l] * (double)rW[l * 2 + 0];
    if (j > 0) lap += (uc[l] - uc[l - 1]) * phi[l - 1] * (double)rW[(l - 1) * 2 + 0];
    if (i < H - 1) lap -= (uc[l + W] - uc[l]) * phi[l] * (double)rW[l * 2 + 1];
    if (i > 0) lap += (uc[l] - uc[l - W]) * phi[l - W] * (double)rW[(l - W) * 2 + 1];
    double norm = 0;
    int count = 0;
    if (j < W - 1) { norm += rW[l * 2 + 0]; count++; }
    if (j > 0) { norm += rW[(l - 1) * 2 + 0]; count++; }
    if (i < H - 1) { norm += rW[l * 2 + 1]; count++; }
    if (i > 0) { norm += rW[(l - W) * 2 + 1]; count++; }
    if (std::fabs(norm) < 1e-5) lap = 0;
    else lap /= (norm / count);
    return lap;
  }

  float laplacian_f32(int c, int i, int j) const {
    const size_t l = (size_t)i * W + j;
    const double* uc = u.data() + c * N;
    const float ul = (float)uc[l];
    float lap = 0;
    if (j < W - 1) lap -= ((float)uc[l + 1] - ul) * (float)phi[l] * (float)rW[l * 2 + 0];
    if (j > 0) lap += (ul - (float)uc[l - 1]) * (float)phi[l - 1] * (float)rW[(l - 1) * 2 + 0];
    if (i < H - 1) lap -= ((float)uc[l + W] - ul) * (float)phi[l] * (float)rW[l * 2 + 1];
    if (i > 0) lap += (ul - (float)uc[l - W]) * (float)phi[l - W] * (float)rW[(l - W) * 2 + 1];
    float norm = 0;
    int count = 0;
    if (j < W - 1) { norm += (float)rW[l * 2 + 0]; count++; }
    if (j > 0) { norm += (float)rW[(l - 1) * 2 + 0]; count++; }
    if (i < H - 1) { norm += (float)rW[l * 2 + 1]; count++; }
    if (i > 0) { norm += (float)rW[(l - W) * 2 + 1]; count++; }
    if (std::fabs(norm) < 1e-5f) return 0.f;
    return lap / (norm / (float)count);
  }

  /* .cu:189-230 products, .cu:477-575 fillSystemMtxKernel, written into the
   * compact form: own right/down edge weights (xy and z), the symmetric 3x3
   * block and the rhs. */
  void assemble() {
    const double alpha = p->alpha, gamma = p->gamma;
    CT* eRxy = coef.data() + 0 * N; CT* eRz = coef.data() + 1 * N;
    CT* eDxy = coef.data() + 2 * N; CT* eDz = coef.data() + 3 * N;
    CT* Buu = coef.data() + 4 * N; CT* Bvv = coef.data() + 5 * N; CT* Bww = coef.data() + 6 * N;
    CT* Buv = coef.data() + 7 * N; CT* Buw = coef.data() + 8 * N; CT* Bvw = coef.data() + 9 * N;
    CT* ru = coef.data() + 10 * N; CT* rv = coef.data() + 11 * N; CT* rw = coef.data() + 12 * N;
    if (F32) { assemble_f32(); return; }
    pool().rows(H, [&](int i0, int i1, int) {
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < W; j++) {
        const size_t l = (size_t)i * W + j;
        /* products (multiply3CollapseKernel / ...Collapse2Kernel / multiply3Kernel) */
        double sxy = 0, sxx = 0, syy = 0, stx = 0, sty = 0;
        for (int k = 0; k < C; k++) {
          const size_t o = l * C + k;
          sxy += psi[o] * Ix[o] * Iy[o];
          sxx += psi[o] * Ix[o] * Ix[o];
          syy += psi[o] * Iy[o] * Iy[o];
        }
        for (int k = 0; k < C; k++) { const size_t o = l * C + k; stx += psi[o] * Ix[o] * It[o]; }
        for (int k = 0; k < C; k++) { const size_t o = l * C + k; sty += psi[o] * Iy[o] * It[o]; }
        const double ql = (double)q[l];
        const double ixy = sxy / C, ix2 = sxx / C, iy2 = syy / C;
        const double itx = stx * ql * ql / C, ity = sty * ql * ql / C;
        const double P = pi[l];
        const double dxy = P * Dx[l] * Dy[l], dx2 = P * Dx[l] * Dx[l], dy2 = P * Dy[l] * Dy[l];
        const double dtx = P * Dx[l] * Dt[l], dty = P * Dy[l] * Dt[l];
        const double uL = laplacian(0, i, j), vL = laplacian(1, i, j);
        const double wL = (double)sW[l] * laplacian(2, i, j);

        double wsXY = 0, wsZ = 0;
        if (j > 0) { const double a = phi[l - 1] * (double)rW[(l - 1) * 2 + 0], b = a * (double)sW[l - 1]; wsXY += a; wsZ += b; }
        if (i > 0) { const double a = phi[l - W] * (double)rW[(l - W) * 2 + 1], b = a * (double)sW[l - W]; wsXY += a; wsZ += b; }
        if (j < W - 1) {
          const double a = phi[l] * (double)rW[l * 2 + 0], b = a * (double)sW[l];
          eRxy[l] = (CT)(alpha * a); eRz[l] = (CT)(alpha * b); wsXY += a; wsZ += b;
        }
        if (i < H - 1) {
          const double a = phi[l] * (double)rW[l * 2 + 1], b = a * (double)sW[l];
          eDxy[l] = (CT)(alpha * a); eDz[l] = (CT)(alpha * b); wsXY += a; wsZ += b;
        }
        const double u0 = u[l], u1 = u[N + l], u2 = u[2 * N + l];
        CT t;
        t = (CT)(-(ix2 + dx2 + alpha * wsXY)); Buu[l] = (CT)((double)t + -(gamma));
        Buv[l] = (CT)(-ixy - dxy);
        Buw[l] = (CT)(P * Dx[l]);
        t = (CT)(-(-itx - dtx + alpha * -uL)); ru[l] = (CT)((double)t + -(-gamma * u0));
        t = (CT)(-(iy2 + dy2 + alpha * wsXY)); Bvv[l] = (CT)((double)t + -(gamma));
        Bvw[l] = (CT)(P * Dy[l]);
        t = (CT)(-(-ity - dty + alpha * -vL)); rv[l] = (CT)((double)t + -(-gamma * u1));
        t = (CT)(-(P + alpha * wsZ)); Bww[l] = (CT)((double)t + -(gamma * (double)mW[l]));
        t = (CT)(-(P * Dt[l] + alpha * -wL)); rw[l] = (CT)((double)t + -(-gamma * (double)mW[l] * u2));
      }
    });
  }

  /* assemble in f32: the same sums and products, alpha and gamma rounded to
   * f32, the channel means as f32 divisions by 5 */
  void assemble_f32() {
    const float alpha = (float)p->alpha, gamma = (float)p->gamma;
    CT* eRxy = coef.data() + 0 * N; CT* eRz = coef.data() + 1 * N;
    CT* eDxy = coef.data() + 2 * N; CT* eDz = coef.data() + 3 * N;
    CT* Buu = coef.data() + 4 * N; CT* Bvv = coef.data() + 5 * N; CT* Bww = coef.data() + 6 * N;
    CT* Buv = coef.data() + 7 * N; CT* Buw = coef.data() + 8 * N; CT* Bvw = coef.data() + 9 * N;
    CT* ru = coef.data() + 10 * N; CT* rv = coef.data() + 11 * N; CT* rw = coef.data() + 12 * N;
    pool().rows(H, [&](int i0, int i1, int) {
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < W; j++) {
        const size_t l = (size_t)i * W + j;
        float sxy = 0, sxx = 0, syy = 0, stx = 0, sty = 0;
        for (int k = 0; k < C; k++) {
          const size_t o = l * C + k;
          const float ps = (float)psi[o], ix = (float)Ix[o], iy = (float)Iy[o];
          sxy += ps * ix * iy;
          sxx += ps * ix * ix;
          syy += ps * iy * iy;
        }
        for (int k = 0; k < C; k++) { const size_t o = l * C + k; stx += (float)psi[o] * (float)Ix[o] * (float)It[o]; }
        for (int k = 0; k < C; k++) { const size_t o = l * C + k; sty += (float)psi[o] * (float)Iy[o] * (float)It[o]; }
        const float ql = (float)q[l];
        const float ixy = sxy / 5.f, ix2 = sxx / 5.f, iy2 = syy / 5.f;
        const float itx = stx * ql * ql / 5.f, ity = sty * ql * ql / 5.f;
        const float P = (float)pi[l], dx = (float)Dx[l], dy = (float)Dy[l], dt = (float)Dt[l];
        const float dxy = P * dx * dy, dx2 = P * dx * dx, dy2 = P * dy * dy;
        const float dtx = P * dx * dt, dty = P * dy * dt;
        const float uL = laplacian_f32(0, i, j), vL = laplacian_f32(1, i, j);
        const float wL = (float)sW[l] * laplacian_f32(2, i, j);
        float wsXY = 0, wsZ = 0;
        if (j > 0) { const float a = (float)phi[l - 1] * (float)rW[(l - 1) * 2 + 0], b = a * (float)sW[l - 1]; wsXY += a; wsZ += b; }
        if (i > 0) { const float a = (float)phi[l - W] * (float)rW[(l - W) * 2 + 1], b = a * (float)sW[l - W]; wsXY += a; wsZ += b; }
        if (j < W - 1) {
          const float a = (float)phi[l] * (float)rW[l * 2 + 0], b = a * (float)sW[l];
          eRxy[l] = alpha * a; eRz[l] = alpha * b; wsXY += a; wsZ += b;
        }
        if (i < H - 1) {
          const float a = (float)phi[l] * (float)rW[l * 2 + 1], b = a * (float)sW[l];
          eDxy[l] = alpha * a; eDz[l] = alpha * b; wsXY += a; wsZ += b;
        }
        const float u0 = (float)u[l], u1 = (float)u[N + l], u2 = (float)u[2 * N + l];
        const float m = (float)mW[l];
        float t;
        t = -(ix2 + dx2 + alpha * wsXY); Buu[l] = t + -gamma;
        Buv[l] = -ixy - dxy;
        Buw[l] = P * dx;
        t = -(-itx - dtx + alpha * -uL); ru[l] = t + -(-gamma * u0);
        t = -(iy2 + dy2 + alpha * wsXY); Bvv[l] = t + -gamma;
        Bvw[l] = P * dy;
        t = -(-ity - dty + alpha * -vL); rv[l] = t + -(-gamma * u1);
        t = -(P + alpha * wsZ); Bww[l] = t + -(gamma * m);
        t = -(P * dt + alpha * -wL); rw[l] = t + -(-gamma * m * u2);
      }
    });
  }

  /* the records as the 16-bit storage holds them (record_round) */
  void store_records() {
    for (size_t l = 0; l < N; l++) {
      float c[13];
      for (int k = 0; k < 13; k++) c[k] = (float)coef[k * N + l];
      record_round(c, p->record_storage);
      for (int k = 0; k < 13; k++) coef[k * N + l] = (CT)c[k];
    }
  }

  void build_system() {
    derivatives();
    occlusion();
    penalties();
    assemble();
  }

  /* The inverse of a pixel's symmetric 3x3 block from its adjugate, in the
   * working precision and this order: cofactors, det = Buu c_uu + Buv c_uv +
   * Buw c_uw, 1 / det, then each cofactor times it.  inv: (uu, vv, ww, uv,
   * uw, vw), the block's own order.  Returns whether the block is negative
   * definite with a finite inverse (Buu < 0, c_ww > 0, det < 0, 1 / det
   * finite); a pixel whose block fails keeps its old values as x^. */
  static bool block_inverse(CT Buu, CT Bvv, CT Bww, CT Buv, CT Buw, CT Bvw, CT* inv) {
    const CT cuu = Bvv * Bww - Bvw * Bvw, cvv = Buu * Bww - Buw * Buw, cww = Buu * Bvv - Buv * Buv;
    const CT cuv = Buw * Bvw - Buv * Bww, cuw = Buv * Bvw - Buw * Bvv, cvw = Buv * Buw - Buu * Bvw;
    const CT det = Buu * cuu + Buv * cuv + Buw * cuw;
    const CT id = (CT)1 / det;
    inv[0] = cuu * id; inv[1] = cvv * id; inv[2] = cww * id;
    inv[3] = cuv * id; inv[4] = cuw * id; inv[5] = cvw * id;
    return Buu < 0 && cww > 0 && det < 0 && id >= -std::numeric_limits<CT>::max();
  }

  /* .cu:463-467: the step limit, r kept within xo +- 0.3, as max then min:
   * the reference's two comparisons for every r but a NaN (the reference's
   * CHECK_FLOATS error), which this clamps into the band instead of passing
   * on; the device's two-instruction form. */
  static CT clamp_step(CT r, CT xo, CT step) { return std::fmin(std::fmax(r, xo - step), xo + step); }

  /* .cu:382-470: one colour of the red-black relaxation; returns the largest
   * |change| over the updated pixels (the check of .cu:914-918 splits exactly
   * into the two colours' changes). */
  /* mode 0: the reference's order inside a pixel (each unknown from the
   * others' old values, .cu:431-459); 1: point SOR (u, then v from the new
   * u, then w from both); 2: the pixel's 3x3 block solved as a whole
   * (block SOR: x^ = B^-1 (rhs - sum_n A_n x_n), B^-1 from the adjugate in
   * the working precision, block_inverse) -- not in the reference. */
  static CT relax_colour(const CT* coef, int W, int H, int colour, CT* x, float omega_f = .7f, int mode = 0) {
    const bool gs = mode == 1;
    const size_t N = (size_t)W * H;
    const CT* eRxy = coef + 0 * N; const CT* eRz = coef + 1 * N;
    const CT* eDxy = coef + 2 * N; const CT* eDz = coef + 3 * N;
    const CT* Buu = coef + 4 * N; const CT* Bvv = coef + 5 * N; const CT* Bww = coef + 6 * N;
    const CT* Buv = coef + 7 * N; const CT* Buw = coef + 8 * N; const CT* Bvw = coef + 9 * N;
    const CT* rhs[3] = {coef + 10 * N, coef + 11 * N, coef + 12 * N};
    // .cu:461 `const float omega = .7`; a variant factor is given as a float too
    const CT omega = (CT)omega_f;
    const CT step = (CT).3;
    CT part_max[256] = {};
    pool().rows(H, [&](int i0, int i1, int part) {
    CT maxd = 0;
    for (int i = i0; i < i1; i++)
      for (int j = (i + colour) & 1; j < W; j += 2) {
        const size_t l = (size_t)i * W + j;
        CT xo[3] = {x[l], x[N + l], x[2 * N + l]};
        CT xn[3];
        if (mode == 2) {
          CT a[3];
          for (int k = 0; k < 3; k++) {
            const CT* xk = x + k * N;
            CT nv = 0;
            if (j > 0) nv -= (k < 2 ? eRxy[l - 1] : eRz[l - 1]) * xk[l - 1];
            if (i > 0) nv -= (k < 2 ? eDxy[l - W] : eDz[l - W]) * xk[l - W];
            if (j < W - 1) nv -= (k < 2 ? eRxy[l] : eRz[l]) * xk[l + 1];
            if (i < H - 1) nv -= (k < 2 ? eDxy[l] : eDz[l]) * xk[l + W];
            nv += rhs[k][l];
            a[k] = nv;
          }
          CT inv[6];
          const bool ok = block_inverse(Buu[l], Bvv[l], Bww[l], Buv[l], Buw[l], Bvw[l], inv);
          CT xh[3] = {inv[0] * a[0] + inv[3] * a[1] + inv[4] * a[2],
                      inv[3] * a[0] + inv[1] * a[1] + inv[5] * a[2],
                      inv[4] * a[0] + inv[5] * a[1] + inv[2] * a[2]};
          if (!ok) for (int k = 0; k < 3; k++) xh[k] = xo[k];
          for (int k = 0; k < 3; k++) {
            CT r = omega * xh[k] + (1 - omega) * xo[k];
            r = clamp_step(r, xo[k], step);
            xn[k] = r;
          }
        } else
        for (int k = 0; k < 3; k++) {
          const CT* xk = x + k * N;
          CT nv = 0;
          if (j > 0) nv -= (k < 2 ? eRxy[l - 1] : eRz[l - 1]) * xk[l - 1];
          if (i > 0) nv -= (k < 2 ? eDxy[l - W] : eDz[l - W]) * xk[l - W];
          if (j < W - 1) nv -= (k < 2 ? eRxy[l] : eRz[l]) * xk[l + 1];
          if (i < H - 1) nv -= (k < 2 ? eDxy[l] : eDz[l]) * xk[l + W];
          nv += rhs[k][l];
          /* the pixel's other unknowns: their old values (.cu:431-459), or
           * with gs (point SOR order) the values this pass already gave them */
          const CT x0 = (gs && k > 0) ? xn[0] : xo[0], x1 = (gs && k > 1) ? xn[1] : xo[1];
          if (k == 0) { nv -= Buv[l] * xo[1]; nv -= Buw[l] * xo[2]; nv /= Buu[l]; }
          else if (k == 1) { nv -= Buv[l] * x0; nv -= Bvw[l] * xo[2]; nv /= Bvv[l]; }
          else { nv -= Buw[l] * x0; nv -= Bvw[l] * x1; nv /= Bww[l]; }
          CT r = omega * nv + (1 - omega) * xo[k];
          r = clamp_step(r, xo[k], step);
          xn[k] = r;
        }
        for (int k = 0; k < 3; k++) {
          x[k * N + l] = xn[k];
          const CT d = std::fabs(xn[k] - xo[k]);
          if (d > maxd) maxd = d;
        }
      }
    part_max[part] = maxd;
    });
    CT maxd = 0;
    for (int k = 0; k < pool().threads(); k++)
      if (part_max[k] > maxd) maxd = part_max[k];
    return maxd;
  }

  /* The f64 residual of the system at x (orc_system_residual): per pixel and
   * unknown, res = r - (sum_n e x_n + sum_k' B x_k'), terms summed in f64 in
   * the order written; sums of res^2, r^2 and x (r + res) per row, then the
   * rows in order (the same value for any thread count). */
  static void residual(const CT* coef, int W, int H, const CT* x, double* rel, double* energy) {
    const size_t N = (size_t)W * H;
    const CT* eRxy = coef + 0 * N; const CT* eRz = coef + 1 * N;
    const CT* eDxy = coef + 2 * N; const CT* eDz = coef + 3 * N;
    const CT* Buu = coef + 4 * N; const CT* Bvv = coef + 5 * N; const CT* Bww = coef + 6 * N;
    const CT* Buv = coef + 7 * N; const CT* Buw = coef + 8 * N; const CT* Bvw = coef + 9 * N;
    vector<double> row(3 * (size_t)H);
    pool().rows(H, [&](int i0, int i1, int) {
      for (int i = i0; i < i1; i++) {
        double rr = 0, bb = 0, ex = 0;
        for (int j = 0; j < W; j++) {
          const size_t l = (size_t)i * W + j;
          const double xo[3] = {(double)x[l], (double)x[N + l], (double)x[2 * N + l]};
          const double B[3][3] = {{(double)Buu[l], (double)Buv[l], (double)Buw[l]},
                                  {(double)Buv[l], (double)Bvv[l], (double)Bvw[l]},
                                  {(double)Buw[l], (double)Bvw[l], (double)Bww[l]}};
          for (int k = 0; k < 3; k++) {
            const CT* xk = x + k * N;
            double mx = 0;
            if (j > 0) mx += (double)(k < 2 ? eRxy[l - 1] : eRz[l - 1]) * (double)xk[l - 1];
            if (i > 0) mx += (double)(k < 2 ? eDxy[l - W] : eDz[l - W]) * (double)xk[l - W];
            if (j < W - 1) mx += (double)(k < 2 ? eRxy[l] : eRz[l]) * (double)xk[l + 1];
            if (i < H - 1) mx += (double)(k < 2 ? eDxy[l] : eDz[l]) * (double)xk[l + W];
            mx += B[k][0] * xo[0];
            mx += B[k][1] * xo[1];
            mx += B[k][2] * xo[2];
            const double rk = (double)coef[(10 + k) * N + l];
            const double res = rk - mx;
            rr += res * res;
            bb += rk * rk;
            ex += xo[k] * (rk + res);
          }
        }
        row[3 * i] = rr; row[3 * i + 1] = bb; row[3 * i + 2] = ex;
      }
    });
    double rr = 0, bb = 0, ex = 0;
    for (int i = 0; i < H; i++) { rr += row[3 * i]; bb += row[3 * i + 1]; ex += row[3 * i + 2]; }
    *rel = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    *energy = 0.5 * ex;
  }

  /* ---- coarse-grid correction: not in the reference (opt-in, cg_levels) ----
   * Coarse level l+1 aggregates the 2x2 pixels (2I..2I+1, 2J..2J+1) of
   * level l (one row or column where a size is odd): Wc = (W + 1) / 2.  P
   * copies a coarse pixel's three unknowns to its fine pixels.  The coarse
   * system has the fine system's form (5-point, a 3x3 block per pixel), so
   * the relaxation runs on it unchanged:
   *   g_right(I,J), g_down(I,J): the sums of the fine edge weights (xy and z
   *     alike) that cross to the right / lower aggregate (P^T M P's edges);
   *   edges: g / 2 -- the crossing weights' mean, the weight a smooth
   *     field's coarse Laplacian has (P^T M P's are twice that: a piecewise
   *     constant field pays for its jumps);
   *   block: the children's blocks summed, plus 2 e for every edge inside
   *     the aggregate on the diagonal (P^T M P's block), plus half the sum of
   *     the four g around it (so that the diagonal keeps the children's
   *     data terms and loses exactly the coarse edges).
   * Sums run in the working precision, from zero: children in the order
   * (2I,2J), (2I,2J+1), (2I+1,2J), (2I+1,2J+1); then the inner edges right
   * of (2I,2J), right of (2I+1,2J), down of (2I,2J), down of (2I,2J+1), each
   * added as 2 e; g over rows 2I, 2I+1 (right) or columns 2J, 2J+1 (down);
   * the diagonal's last term as 0.5 * (((0 + g_left) + g_up) + g_right) +
   * g_down), absent neighbours skipped.  A coarse level is made while the
   * finer one is at least 4x4. */
  static bool cg_has_coarse(int W, int H) { return W >= 4 && H >= 4; }
  static bool cg_when_p(const orc_params* p, int it) { return it % 25 == p->cg_at; }
  bool cg_when(int it) const { return cg_when_p(p, it); }

  /* g_right (down = 0) or g_down (down = 1) of coarse pixel (I, J), edge plane z (0: xy, 1: z) */
  static CT cg_cross(const CT* f, int W, int H, int I, int J, int down, int z) {
    const size_t N = (size_t)W * H;
    const int Wc = (W + 1) / 2, Hc = (H + 1) / 2;
    CT s = 0;
    if (!down) {
      if (J + 1 >= Wc) return s;
      const size_t a = (size_t)(2 * I) * W + 2 * J + 1;
      s += f[z * N + a];
      if (2 * I + 1 < H) s += f[z * N + a + W];
    } else {
      if (I + 1 >= Hc) return s;
      const size_t a = (size_t)(2 * I + 1) * W + 2 * J;
      s += f[(2 + z) * N + a];
      if (2 * J + 1 < W) s += f[(2 + z) * N + a + 1];
    }
    return s;
  }

  /* the coarse operator (planes 0-9); the rhs planes are set to zero */
  static void cg_operator(const CT* f, int W, int H, CT* c) {
    const size_t N = (size_t)W * H;
    const int Wc = (W + 1) / 2, Hc = (H + 1) / 2;
    const size_t Nc = (size_t)Wc * Hc;
    for (int I = 0; I < Hc; I++)
      for (int J = 0; J < Wc; J++) {
        const size_t L = (size_t)I * Wc + J;
        const int i0 = 2 * I, j0 = 2 * J;
        const bool r1 = i0 + 1 < H, c1 = j0 + 1 < W;
        const size_t p00 = (size_t)i0 * W + j0;
        size_t ch[4];
        int n = 0;
        ch[n++] = p00;
        if (c1) ch[n++] = p00 + 1;
        if (r1) ch[n++] = p00 + W;
        if (r1 && c1) ch[n++] = p00 + W + 1;
        CT around[2];
        for (int z = 0; z < 2; z++) {
          const CT gr = cg_cross(f, W, H, I, J, 0, z), gd = cg_cross(f, W, H, I, J, 1, z);
          CT t = 0;
          if (J > 0) t += cg_cross(f, W, H, I, J - 1, 0, z);
          if (I > 0) t += cg_cross(f, W, H, I - 1, J, 1, z);
          t += gr;
          t += gd;
          around[z] = t;
          c[z * Nc + L] = (CT).5 * gr;
          c[(2 + z) * Nc + L] = (CT).5 * gd;
        }
        for (int k = 4; k < 10; k++) {
          CT s = 0;
          for (int a = 0; a < n; a++) s += f[k * N + ch[a]];
          if (k <= 6) {
            /* Buu, Bvv: the xy edges (planes 0, 2); Bww: the z edges (1, 3) */
            const int z = k < 6 ? 0 : 1;
            if (c1) s += (CT)2 * f[z * N + p00];
            if (c1 && r1) s += (CT)2 * f[z * N + p00 + W];
            if (r1) s += (CT)2 * f[(2 + z) * N + p00];
            if (r1 && c1) s += (CT)2 * f[(2 + z) * N + p00 + 1];
            s += (CT).5 * around[z];
          }
          c[k * Nc + L] = s;
        }
        for (int k = 10; k < 13; k++) c[k * Nc + L] = 0;
      }
  }

  /* The residual of pixel l's unknown k at x in the working precision:
   * rhs - e_left x_left - e_up x_up - e_right x_right - e_down x_down -
   * B_k0 x_0 - B_k1 x_1 - B_k2 x_2, in that order (missing neighbours
   * skipped). */
  static CT cg_residual(const CT* f, int W, int H, const CT* x, int i, int j, int k) {
    const size_t N = (size_t)W * H, l = (size_t)i * W + j;
    const CT* xk = x + k * N;
    const int er = k < 2 ? 0 : 1, ed = k < 2 ? 2 : 3;
    CT a = f[(10 + k) * N + l];
    if (j > 0) a -= f[er * N + l - 1] * xk[l - 1];
    if (i > 0) a -= f[ed * N + l - W] * xk[l - W];
    if (j < W - 1) a -= f[er * N + l] * xk[l + 1];
    if (i < H - 1) a -= f[ed * N + l] * xk[l + W];
    static const int row[3][3] = {{4, 7, 8}, {7, 5, 9}, {8, 9, 6}}; /* (Buu Buv Buw) (Buv Bvv Bvw) (Buw Bvw Bww) */
    for (int m = 0; m < 3; m++) a -= f[row[k][m] * N + l] * x[m * N + l];
    return a;
  }

  /* the coarse rhs (planes 10-12): sums of the children's residuals at x */
  static void cg_restrict(const CT* f, int W, int H, const CT* x, CT* c) {
    const int Wc = (W + 1) / 2, Hc = (H + 1) / 2;
    const size_t Nc = (size_t)Wc * Hc;
    for (int I = 0; I < Hc; I++)
      for (int J = 0; J < Wc; J++)
        for (int k = 0; k < 3; k++) {
          CT s = 0;
          for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++) {
              const int i = 2 * I + di, j = 2 * J + dj;
              if (i < H && j < W) s += cg_residual(f, W, H, x, i, j, k);
            }
          c[(10 + k) * Nc + (size_t)I * Wc + J] = s;
        }
  }

  /* x += P xc */
  static void cg_prolong(CT* x, int W, int H, const CT* xc) {
    const size_t N = (size_t)W * H;
    const int Wc = (W + 1) / 2, Hc = (H + 1) / 2;
    const size_t Nc = (size_t)Wc * Hc;
    for (int i = 0; i < H; i++)
      for (int j = 0; j < W; j++)
        for (int k = 0; k < 3; k++) x[k * N + (size_t)i * W + j] += xc[k * Nc + (size_t)(i / 2) * Wc + j / 2];
  }

  /* The coarse levels of a system: sizes and operators (rhs zero). */
  struct CgHier {
    vector<int> Ws, Hs;
    vector<vector<CT>> coef, x;
    void build(const CT* f, int W, int H, int levels) {
      Ws.assign(1, W); Hs.assign(1, H);
      coef.clear(); x.clear();
      coef.emplace_back(); x.emplace_back();
      const CT* cur = f;
      while ((int)Ws.size() <= levels && cg_has_coarse(Ws.back(), Hs.back())) {
        const int w = Ws.back(), h = Hs.back(), wc = (w + 1) / 2, hc = (h + 1) / 2;
        coef.emplace_back(13 * (size_t)wc * hc);
        x.emplace_back(3 * (size_t)wc * hc);
        cg_operator(cur, w, h, coef.back().data());
        cur = coef.back().data();
        Ws.push_back(wc); Hs.push_back(hc);
      }
    }
    int levels() const { return (int)Ws.size() - 1; }
  };

  /* One V-cycle on the fine system (f, x): for l = 1..L, the rhs of level l
   * from the residual of level l-1, x_l = 0, then `sweeps` red-black
   * iterations at level l; for l = L..1, `sweeps` iterations at level l, then
   * x_{l-1} += P x_l.  (The coarsest level gets 2 x sweeps.)  The
   * relaxation is the level solve's own: factor, pixel order, step limit. */
  static void cg_vcycle(CgHier& h, const CT* f, CT* x, int sweeps, float omega, int mode) {
    const int L = h.levels();
    auto relax = [&](int l) {
      for (int s = 0; s < sweeps; s++) {
        relax_colour(h.coef[l].data(), h.Ws[l], h.Hs[l], 0, h.x[l].data(), omega, mode);
        relax_colour(h.coef[l].data(), h.Ws[l], h.Hs[l], 1, h.x[l].data(), omega, mode);
      }
    };
    for (int l = 1; l <= L; l++) {
      cg_restrict(l == 1 ? f : h.coef[l - 1].data(), h.Ws[l - 1], h.Hs[l - 1], l == 1 ? x : h.x[l - 1].data(),
                  h.coef[l].data());
      std::fill(h.x[l].begin(), h.x[l].end(), (CT)0);
      relax(l);
    }
    for (int l = L; l >= 1; l--) {
      relax(l);
      cg_prolong(l == 1 ? x : h.x[l - 1].data(), h.Ws[l - 1], h.Hs[l - 1], h.x[l].data());
    }
  }

  /* .cu:579-635 GPU warps: floor base, explicit clamps, f32 weights (f64 in the FP64 oracle) */
  void warp_gpu() {
    typedef typename std::conditional<REF, float, double>::type AT;
    const double* I1 = in->I1; const double* I2 = in->I2;
    pool().rows(H, [&](int i0, int i1, int) {
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < W; j++) {
        const size_t l = (size_t)i * W + j;
        const double x = j + u[l], y = i + u[N + l];
        if (x < 0 || x > W - 1 || y < 0 || y > H - 1) {
          for (int k = 0; k < C; k++) Iw[l * C + k] = I1[l * C + k];
          Dw[l] = F32 ? (double)((float)in->D1[l] - (float)u[2 * N + l]) : in->D1[l] - u[2 * N + l];
        } else {
          const int x0 = std::max(0, std::min(W - 1, (int)std::floor(x)));
          const int y0 = std::max(0, std::min(H - 1, (int)std::floor(y)));
          const int x1 = std::min(W - 1, x0 + 1), y1 = std::min(H - 1, y0 + 1);
          const AT ax = x - x0, ay = y - y0;
          const AT w00 = (1 - ax) * (1 - ay), w01 = (1 - ax) * ay, w10 = ax * (1 - ay), w11 = ax * ay;
          if (F32) {
            for (int k = 0; k < C; k++) {
              const float s = w00 * (float)I2[((size_t)y0 * W + x0) * C + k] +
                              w01 * (float)I2[((size_t)y1 * W + x0) * C + k] +
                              w10 * (float)I2[((size_t)y0 * W + x1) * C + k] +
                              w11 * (float)I2[((size_t)y1 * W + x1) * C + k];
              Iw[l * C + k] = img_round(s, p->image_storage);
            }
            const double* D2 = in->D2;
            Dw[l] = w00 * (float)D2[(size_t)y0 * W + x0] + w01 * (float)D2[(size_t)y1 * W + x0] +
                    w10 * (float)D2[(size_t)y0 * W + x1] + w11 * (float)D2[(size_t)y1 * W + x1] - (float)u[2 * N + l];
            continue;
          }
          for (int k = 0; k < C; k++)
            Iw[l * C + k] = img_round(w00 * I2[((size_t)y0 * W + x0) * C + k] + w01 * I2[((size_t)y1 * W + x0) * C + k] +
                                      w10 * I2[((size_t)y0 * W + x1) * C + k] + w11 * I2[((size_t)y1 * W + x1) * C + k],
                                      p->image_storage);
          const double* D2 = in->D2;
          Dw[l] = w00 * D2[(size_t)y0 * W + x0] + w01 * D2[(size_t)y1 * W + x0] + w10 * D2[(size_t)y0 * W + x1] +
                  w11 * D2[(size_t)y1 * W + x1] - u[2 * N + l];
        }
      }
    });
  }

  /* .cu:755-954: the outer/inner loops with their exits */
  void run(int allow_early, orc_level_stats* st) {
    const double convThreshold = 4e-4;
    const int checkEvery = 25;
    int outer = 0, sweeps = 0, rchecks = 0;
    double lastStat = -1, lastRes = -1, lastEnergy = 0;
    float lastMax = -1;
    CgHier cg;
    for (int count = 0; count < p->n_outer; count++) {
      outer++;
      derivatives();
      std::fill(du.begin(), du.end(), (CT)0);
      occlusion();
      for (int hh = 0; hh < p->n_inner; hh++) {
        penalties();
        assemble();
        if (REF && p->record_storage != ORC_REC_F32) store_records();
        const float om = p->omega > 0 ? p->omega : .7f;
        if (p->cg_levels > 0) cg.build(coef.data(), W, H, p->cg_levels);
        for (int it = 0; it < p->n_sor; it++) {
          const CT m0 = relax_colour(coef.data(), W, H, 0, du.data(), om, p->pixel_gs);
          const CT m1 = relax_colour(coef.data(), W, H, 1, du.data(), om, p->pixel_gs);
          sweeps++;
          if (p->cg_levels > 0 && cg.levels() > 0 && cg_when(it))
            cg_vcycle(cg, coef.data(), du.data(), p->cg_sweeps, om, p->pixel_gs);
          if (it % checkEvery == checkEvery - 1) {
            const CT md = std::max(m0, m1);
            lastMax = (float)md;
            if (p->residual_exit) {
              /* not in the reference: the FP64 residual test (rgbdflow.h RGBDFLOW_EXIT_RESIDUAL) */
              residual(coef.data(), W, H, du.data(), &lastRes, &lastEnergy);
              rchecks++;
              if (lastRes <= p->residual_tol) break;
            } else if (!p->disable_inner_exit) {
              if (md < 1e-5) break;
              if (md >= (CT).3) break;
            }
          }
        }
      }
      for (size_t l = 0; l < 3 * N; l++) u[l] += (double)du[l];
      warp_gpu();
      if (allow_early) {
        double acc = 0;
        for (int c = 0; c < 3; c++)
          for (size_t l = 0; l < N; l++) {
            const CT d = du[c * N + l];
            const CT wgt = (c < 2) ? (CT)1 : (CT)cW[l];
            acc += (double)(CT)((CT)(d * d) * wgt);
          }
        const double stat = acc / (double)(W * H * 3);
        lastStat = stat;
        if (!p->disable_outer_exit && stat < convThreshold) break;
      }
    }
    if (st) {
      st->outer_exec = outer; st->sweeps_exec = sweeps; st->last_stat = lastStat; st->last_maxdiff = lastMax;
      st->residual_checks = rchecks; st->last_residual = lastRes; st->last_energy = lastEnergy;
    }
  }
};

template <class CT, bool REF, class ST = double>
void level_solve_t(const orc_level_in* in, const orc_params* p, int allow_early, double* u, double* v,
                   double* w, orc_level_stats* st) {
  Solver<CT, REF, ST> S(in, p);
  const size_t N = S.N;
  for (size_t l = 0; l < N; l++) { S.u[l] = u[l]; S.u[N + l] = v[l]; S.u[2 * N + l] = w[l]; }
  S.run(allow_early, st);
  for (size_t l = 0; l < N; l++) { u[l] = S.u[l]; v[l] = S.u[N + l]; w[l] = S.u[2 * N + l]; }
}

/* sceneFlowMultiscale.cpp:16-25 getDownsizedCamParams */
orc_camera downsized(orc_camera c, int width) {
  const double f = (double)width / c.xRes;
  c.xRes = std::rint(c.xRes * f);
  c.yRes = std::rint(c.yRes * f);
  c.centerX = std::rint(c.centerX * f);
  c.centerY = std::rint(c.centerY * f);
  c.focalLength *= f;
  return c;
}

/* OpticalFlow.cpp:104-111 with the full-size parameters of
 * sceneFlowMultiscale.cpp:144: the weight of level pixel l is computed from
 * D1_level[(l / W_full) * W_level + l % W_full] (only the first N_level
 * entries reach the solver, .cu:686). */
void depth_mag(const double* D, int Wl, int Hl, int Wf, int Hf, float f, double* out) {
  (void)Hf;
  const size_t Nl = (size_t)Wl * Hl;
  for (size_t l = 0; l < Nl; l++) {
    const size_t r = l / (size_t)Wf, c = l % (size_t)Wf;
    const size_t idx = r * (size_t)Wl + c;
    const double d = idx < Nl ? D[idx] : 0.0; /* the reference reads past the level buffer here */
    out[l] = d > 0 ? dsq(f / d) : dsq(f / 2);
  }
}

/* OpticalFlow.cpp:113-141 */
void reg_weights(const float* hpb, const double* depth, int W, int H, int type, double* out) {
  std::fill(out, out + (size_t)W * H * 2, 0.0);
  for (int i = 0, l = 0; i < H; i++)
    for (int j = 0; j < W; j++, l++) {
      if (j < W - 1) {
        double v = 1;
        if (type == ORC_REG_LETOUZEY) v = std::exp(-dsq((depth[l] - depth[l + 1]) / .02));
        else if (type == ORC_REG_HPB) v = 1 - hpb[((size_t)i * W + j) * 2 + 0] * .99;
        out[(size_t)l * 2 + 0] = v;
      }
      if (i < H - 1) {
        double v = 1;
        if (type == ORC_REG_LETOUZEY) v = std::exp(-dsq((depth[l] - depth[l + W]) / .02));
        else if (type == ORC_REG_HPB) v = 1 - hpb[((size_t)i * W + j) * 2 + 1] * .99;
        out[(size_t)l * 2 + 1] = v;
      }
    }
}

int level_solve_dispatch(const orc_level_in* in, const orc_params* p, int allow_early, double* u, double* v,
                         double* w, orc_level_stats* st) {
  if (p->precision == ORC_PREC_REF && p->image_storage != ORC_IMG_F64)
    level_solve_t<float, true, float>(in, p, allow_early, u, v, w, st);
  else if (p->precision == ORC_PREC_REF) level_solve_t<float, true>(in, p, allow_early, u, v, w, st);
  else level_solve_t<double, false>(in, p, allow_early, u, v, w, st);
  return 0;
}

}  // namespace

/* ========================================================================== */
template <class CT, bool REF>
static void relax_residual_t(const double* coef13, int W, int H, float omega, int gs, int max_iter, double tol,
                             double* x, int* iters, double* residual, double* energy) {
  const size_t N = (size_t)W * H;
  vector<CT> c(coef13, coef13 + 13 * N), xx(x, x + 3 * N);
  int it = 0;
  double res = -1, en = 0;
  while (it < max_iter) {
    Solver<CT, REF>::relax_colour(c.data(), W, H, 0, xx.data(), omega, gs);
    Solver<CT, REF>::relax_colour(c.data(), W, H, 1, xx.data(), omega, gs);
    it++;
    if ((it - 1) % 25 == 24) {
      Solver<CT, REF>::residual(c.data(), W, H, xx.data(), &res, &en);
      if (res <= tol) break;
    }
  }
  for (size_t k = 0; k < 3 * N; k++) x[k] = xx[k];
  *iters = it;
  *residual = res;
  *energy = en;
}

template <class CT, bool REF>
static void relax_cg_t(const double* coef13, int W, int H, const orc_params* p, int n_iter, double* x,
                       double* maxd, double* res) {
  typedef Solver<CT, REF> S;
  const size_t N = (size_t)W * H;
  vector<CT> c(coef13, coef13 + 13 * N), xx(x, x + 3 * N);
  const float om = p->omega > 0 ? p->omega : .7f;
  typename S::CgHier h;
  if (p->cg_levels > 0) h.build(c.data(), W, H, p->cg_levels);
  for (int it = 0; it < n_iter; it++) {
    const CT m0 = S::relax_colour(c.data(), W, H, 0, xx.data(), om, p->pixel_gs);
    const CT m1 = S::relax_colour(c.data(), W, H, 1, xx.data(), om, p->pixel_gs);
    if (p->cg_levels > 0 && h.levels() > 0 && S::cg_when_p(p, it))
      S::cg_vcycle(h, c.data(), xx.data(), p->cg_sweeps, om, p->pixel_gs);
    if (maxd) maxd[it] = (double)std::max(m0, m1);
    if (res) {
      double e;
      S::residual(c.data(), W, H, xx.data(), &res[it], &e);
    }
  }
  for (size_t k = 0; k < 3 * N; k++) x[k] = xx[k];
}

extern "C" {

int orc_pyramid_levels(int W, int H, double ratio, int min_width, int* widths, int* heights) {
  vector<double> one((size_t)W * H, 0.0);
  /* sizes only: run the size recursion of GaussianPyramid.cpp:23-60 */
  if (ratio > 0.98 || ratio < 0.4) ratio = 0.75;
  double w = W * ratio;
  int n = 1;
  while (w >= min_width) { n++; w *= ratio; }
  int nn = std::log(0.25) / std::log(ratio);
  vector<int> ws(n), hs(n);
  ws[0] = W; hs[0] = H;
  for (int i = 1; i < n; i++) {
    if (i <= nn) { double r = std::pow(ratio, i); ws[i] = (double)W * r; hs[i] = (double)H * r; }
    else { double rate = (double)std::pow(ratio, i) * W / ws[i - nn]; ws[i] = (double)ws[i - nn] * rate; hs[i] = (double)hs[i - nn] * rate; }
  }
  for (int i = 0; i < n && i < 16; i++) { if (widths) widths[i] = ws[i]; if (heights) heights[i] = hs[i]; }
  return n;
}

void orc_gaussian_smooth(const double* img, int W, int H, int C, double sigma, int fsize, double* out) {
  gaussian_smooth(img, W, H, C, sigma, fsize, out);
}
void orc_resize_ratio(const double* src, int W, int H, int C, double ratio, double* dst) {
  resize_ratio(src, W, H, C, ratio, dst);
}
void orc_resize_dims(const double* src, int Ws, int Hs, int C, int Wd, int Hd, double* dst) {
  resize_dims(src, Ws, Hs, C, Wd, Hd, dst);
}
void orc_gaussian_pyramid(const double* img, int W, int H, int C, double ratio, int min_width, double* out) {
  vector<Level> P = gaussian_pyramid(img, W, H, C, ratio, min_width);
  size_t off = 0;
  for (auto& L : P) { std::memcpy(out + off, L.d.data(), L.d.size() * sizeof(double)); off += L.d.size(); }
}
void orc_depth_map_pyramid(const double* depth, int W, int H, double ratio, int min_width, double* out) {
  vector<Level> P = depth_map_pyramid(depth, W, H, ratio, min_width);
  size_t off = 0;
  for (auto& L : P) { std::memcpy(out + off, L.d.data(), L.d.size() * sizeof(double)); off += L.d.size(); }
}
void orc_im2feature(const double* bgr, int W, int H, double* out5) { im2feature(bgr, W, H, out5); }
void orc_record_round(float* c13, int storage) {
  if (storage != ORC_REC_F32) record_round(c13, storage);
}

double orc_img_round(double x, int storage) { return img_round(x, storage); }
void orc_warp_cpu(const double* I1, const double* D1, const double* I2, const double* D2, const double* u,
                  const double* v, const double* w, int W, int H, int C, double* wI, double* wD) {
  warp_cpu(I1, D1, I2, D2, u, v, w, W, H, C, wI, wD);
}
void orc_depth_stdevs_pointwise(const float* depth, int W, int H, int max_dist, float* st) {
  depth_stdevs_pointwise(depth, W, H, max_dist, st);
}
void orc_l1_distance_to_boundary(const float* depth, int W, int H, int cap, float* dist) {
  l1_distance_to_boundary(depth, W, H, cap, dist);
}
/* OpticalFlow.cpp:71-81 */
void orc_depth_data_weights(const float* sigma, int W, int H, double dvc, double* out) {
  for (size_t l = 0; l < (size_t)W * H; l++) out[l] = stereo_error(1) / sigma[l] * dvc;
}
void orc_depth_mag_weights(const double* D, int Wl, int Hl, int Wf, int Hf, float f, double* out) {
  depth_mag(D, Wl, Hl, Wf, Hf, f, out);
}
void orc_reg_weights(const float* hpb, const double* depth, int W, int H, int type, double* out) {
  reg_weights(hpb, depth, W, H, type, out);
}
void orc_median_fill(double* depth, int W, int H, int hw) { median_fill(depth, W, H, hw); }
int orc_median_fill_rounds(double* depth, int W, int H) { return median_fill_rounds(depth, W, H); }
int orc_colour_fill_rounds(double* depth, const double* bgr, int W, int H, int max_region) {
  return colour_fill_rounds(depth, bgr, W, H, max_region);
}
double orc_fill_exp(double x) { return fill_exp(x); }

int orc_level_solve(const orc_level_in* in, const orc_params* p, int allow_early, double* u, double* v, double* w,
                    orc_level_stats* st) {
  return level_solve_dispatch(in, p, allow_early, u, v, w, st);
}

void orc_level_system(const orc_level_in* in, const orc_params* p, const double* u, const double* v,
                      const double* w, double* coef13) {
  const size_t N = (size_t)in->W * in->H;
  if (p->precision == ORC_PREC_REF && p->image_storage != ORC_IMG_F64) {
    Solver<float, true, float> S(in, p);
    for (size_t l = 0; l < N; l++) { S.u[l] = u[l]; S.u[N + l] = v[l]; S.u[2 * N + l] = w[l]; }
    S.build_system();
    for (size_t k = 0; k < 13 * N; k++) coef13[k] = S.coef[k];
  } else if (p->precision == ORC_PREC_REF) {
    Solver<float, true> S(in, p);
    for (size_t l = 0; l < N; l++) { S.u[l] = u[l]; S.u[N + l] = v[l]; S.u[2 * N + l] = w[l]; }
    S.build_system();
    for (size_t k = 0; k < 13 * N; k++) coef13[k] = S.coef[k];
  } else {
    Solver<double, false> S(in, p);
    for (size_t l = 0; l < N; l++) { S.u[l] = u[l]; S.u[N + l] = v[l]; S.u[2 * N + l] = w[l]; }
    S.build_system();
    for (size_t k = 0; k < 13 * N; k++) coef13[k] = S.coef[k];
  }
}

void orc_warp_gpu_style(const double* I1, const double* D1, const double* I2, const double* D2, const double* u,
                        const double* v, const double* w, int W, int H, int precision, double* wI, double* wD) {
  const size_t N = (size_t)W * H;
  std::vector<double> z(N, 0.0), one(N * 2, 1.0);
  orc_level_in in;
  in.W = W; in.H = H; in.focal_level = 1;
  in.I1 = I1; in.I2 = I2; in.warpI2 = I2; in.D1 = D1; in.D2 = D2; in.warpD2 = D2;
  in.delta = z.data(); in.s = z.data(); in.m = z.data(); in.r = one.data();
  orc_params p{};
  auto run = [&](auto& S) {
    for (size_t l = 0; l < N; l++) { S.u[l] = u[l]; S.u[N + l] = v[l]; S.u[2 * N + l] = w[l]; }
    S.warp_gpu();
    std::memcpy(wI, S.Iw.data(), N * 5 * sizeof(double));
    std::memcpy(wD, S.Dw.data(), N * sizeof(double));
  };
  if (precision == ORC_PREC_REF) { Solver<float, true> S(&in, &p); run(S); }
  else { Solver<double, false> S(&in, &p); run(S); }
}

void orc_warp_gpu_style_f32(const double* I1, const double* D1, const double* I2, const double* D2, const double* u,
                            const double* v, const double* w, int W, int H, int image_storage, double* wI,
                            double* wD) {
  const size_t N = (size_t)W * H;
  std::vector<double> z(N, 0.0), one(N * 2, 1.0);
  orc_level_in in;
  in.W = W; in.H = H; in.focal_level = 1;
  in.I1 = I1; in.I2 = I2; in.warpI2 = I2; in.D1 = D1; in.D2 = D2; in.warpD2 = D2;
  in.delta = z.data(); in.s = z.data(); in.m = z.data(); in.r = one.data();
  orc_params p{};
  p.image_storage = image_storage;
  Solver<float, true, float> S(&in, &p);
  for (size_t l = 0; l < N; l++) { S.u[l] = u[l]; S.u[N + l] = v[l]; S.u[2 * N + l] = w[l]; }
  S.warp_gpu();
  std::memcpy(wI, S.Iw.data(), N * 5 * sizeof(double));
  std::memcpy(wD, S.Dw.data(), N * sizeof(double));
}

void orc_level_weights(const orc_level_in* in, const orc_params* p, const double* u, const double* v,
                       const double* w, double* phi, double* psi, double* pi) {
  const size_t N = (size_t)in->W * in->H;
  auto run = [&](auto& S) {
    for (size_t l = 0; l < N; l++) { S.u[l] = u[l]; S.u[N + l] = v[l]; S.u[2 * N + l] = w[l]; }
    S.derivatives();
    S.occlusion();
    S.penalties();
    std::memcpy(phi, S.phi.data(), N * sizeof(double));
    std::memcpy(psi, S.psi.data(), N * 5 * sizeof(double));
    std::memcpy(pi, S.pi.data(), N * sizeof(double));
  };
  if (p->precision == ORC_PREC_REF && p->image_storage != ORC_IMG_F64) { Solver<float, true, float> S(in, p); run(S); }
  else if (p->precision == ORC_PREC_REF) { Solver<float, true> S(in, p); run(S); }
  else { Solver<double, false> S(in, p); run(S); }
}

void orc_relax(const double* coef13, int W, int H, int precision, int n_iter, double* x) {
  const size_t N = (size_t)W * H;
  if (precision == ORC_PREC_REF) {
    vector<float> c(coef13, coef13 + 13 * N), xx(x, x + 3 * N);
    for (int it = 0; it < n_iter; it++) {
      Solver<float, true>::relax_colour(c.data(), W, H, 0, xx.data());
      Solver<float, true>::relax_colour(c.data(), W, H, 1, xx.data());
    }
    for (size_t k = 0; k < 3 * N; k++) x[k] = xx[k];
  } else {
    vector<double> c(coef13, coef13 + 13 * N), xx(x, x + 3 * N);
    for (int it = 0; it < n_iter; it++) {
      Solver<double, false>::relax_colour(c.data(), W, H, 0, xx.data());
      Solver<double, false>::relax_colour(c.data(), W, H, 1, xx.data());
    }
    for (size_t k = 0; k < 3 * N; k++) x[k] = xx[k];
  }
}

void orc_relax_cg(const double* coef13, int W, int H, const orc_params* p, int n_iter, double* x, double* maxd,
                  double* res) {
  if (p->precision == ORC_PREC_REF) relax_cg_t<float, true>(coef13, W, H, p, n_iter, x, maxd, res);
  else relax_cg_t<double, false>(coef13, W, H, p, n_iter, x, maxd, res);
}

void orc_cg_operator(const double* coef13, int W, int H, int precision, double* out13) {
  const int Wc = (W + 1) / 2, Hc = (H + 1) / 2;
  const size_t N = (size_t)W * H, Nc = (size_t)Wc * Hc;
  if (precision == ORC_PREC_REF) {
    vector<float> f(coef13, coef13 + 13 * N), c(13 * Nc);
    Solver<float, true>::cg_operator(f.data(), W, H, c.data());
    for (size_t k = 0; k < 13 * Nc; k++) out13[k] = c[k];
  } else {
    Solver<double, false>::cg_operator(coef13, W, H, out13);
  }
}

void orc_cg_restrict(const double* coef13, int W, int H, int precision, const double* x, double* rhs3) {
  const int Wc = (W + 1) / 2, Hc = (H + 1) / 2;
  const size_t N = (size_t)W * H, Nc = (size_t)Wc * Hc;
  if (precision == ORC_PREC_REF) {
    vector<float> f(coef13, coef13 + 13 * N), xx(x, x + 3 * N), c(13 * Nc);
    Solver<float, true>::cg_restrict(f.data(), W, H, xx.data(), c.data());
    for (size_t k = 0; k < 3 * Nc; k++) rhs3[k] = c[10 * Nc + k];
  } else {
    vector<double> c(13 * Nc);
    Solver<double, false>::cg_restrict(coef13, W, H, x, c.data());
    for (size_t k = 0; k < 3 * Nc; k++) rhs3[k] = c[10 * Nc + k];
  }
}

void orc_cg_prolong(double* x, int W, int H, int precision, const double* xc) {
  const int Wc = (W + 1) / 2, Hc = (H + 1) / 2;
  const size_t N = (size_t)W * H, Nc = (size_t)Wc * Hc;
  if (precision == ORC_PREC_REF) {
    vector<float> xx(x, x + 3 * N), cc(xc, xc + 3 * Nc);
    Solver<float, true>::cg_prolong(xx.data(), W, H, cc.data());
    for (size_t k = 0; k < 3 * N; k++) x[k] = xx[k];
  } else {
    Solver<double, false>::cg_prolong(x, W, H, xc);
  }
}

void orc_system_residual(const double* coef13, int W, int H, const double* x, double* out) {
  Solver<double, false>::residual(coef13, W, H, x, &out[0], &out[1]);
}

void orc_relax_residual(const double* coef13, int W, int H, int precision, float omega, int pixel_gs, int max_iter,
                        double tol, double* x, int* iters, double* residual, double* energy) {
  if (omega <= 0) omega = .7f;
  if (precision == ORC_PREC_REF)
    relax_residual_t<float, true>(coef13, W, H, omega, pixel_gs, max_iter, tol, x, iters, residual, energy);
  else
    relax_residual_t<double, false>(coef13, W, H, omega, pixel_gs, max_iter, tol, x, iters, residual, energy);
}

/* sceneFlowMultiscale.cpp:30-201 */
int orc_multiscale(const double* im1, const double* dm1, const double* im2, const double* dm2, const float* sigma1,
                   const float* hpb, int W, int H, const orc_params* p, double* uo, double* vo, double* wo,
                   orc_stats* st) {
  const double ratio = p->pyramid_ratio;
  const size_t N = (size_t)W * H;
  vector<double> d1(dm1, dm1 + N), d2(dm2, dm2 + N);
  median_fill(d1.data(), W, H, 1);
  median_fill(d2.data(), W, H, 1);
  vector<double> delta(N), sm(N);
  orc_depth_data_weights(sigma1, W, H, p->depth_vs_color_data_weight, delta.data());
  for (size_t l = 0; l < N; l++) sm[l] = dsq(p->cam.focalLength);
  // the six pyramids, as concurrent tasks when threads are enabled
  // (sceneFlowMultiscale.cpp:40-77 builds them in a thread group)
  vector<Level> P1, P2, PD1, PD2, Pdelta, Ps;
  {
    const double* src[6] = {im1, im2, d1.data(), d2.data(), delta.data(), sm.data()};
    const int nc[6] = {3, 3, 1, 1, 1, 1};
    vector<Level>* dst[6] = {&P1, &P2, &PD1, &PD2, &Pdelta, &Ps};
    auto task = [&](int t) { *dst[t] = gaussian_pyramid(src[t], W, H, nc[t], ratio, p->min_width); };
    if (pool().threads() > 1) {
      vector<std::thread> th;
      for (int t = 1; t < 6; t++) th.emplace_back(task, t);
      task(0);
      for (auto& x : th) x.join();
    } else {
      for (int t = 0; t < 6; t++) task(t);
    }
  }
  const int L = (int)P1.size();
  vector<size_t> hpbOff(L, 0);
  for (int k = 1; k < L; k++) hpbOff[k] = hpbOff[k - 1] + (size_t)P1[k - 1].W * P1[k - 1].H * 2;
  // no maps given: computeHeuristicPb of each level's colour and raw-depth
  // pyramid levels, as sceneFlowMultiscale.cpp:116 computes them (the raw
  // depth is dm1 as given: its colour-guided fill, :78-83, is the caller's)
  vector<float> hpbOwn;
  if (p->regularization_type == ORC_REG_HPB && !hpb) {
    const vector<Level> DM = depth_map_pyramid(dm1, W, H, ratio, p->min_width);
    if (DM.size() != P1.size()) return -2;
    hpbOwn.resize(hpbOff[L - 1] + (size_t)P1[L - 1].W * P1[L - 1].H * 2);
    for (int k = 0; k < L; k++) {
      if (DM[k].W != P1[k].W || DM[k].H != P1[k].H) return -2;
      orc_heuristic_pb(P1[k].d.data(), DM[k].d.data(), P1[k].W, P1[k].H, hpbOwn.data() + hpbOff[k]);
    }
    hpb = hpbOwn.data();
  }

  vector<double> u, v, w;
  if (st) { std::memset(st, 0, sizeof(*st)); st->n_levels = L; }
  for (int k = L - 1; k >= 0; k--) {
    const int Wk = P1[k].W, Hk = P1[k].H;
    const size_t Nk = (size_t)Wk * Hk;
    vector<double> I1(Nk * 5), I2(Nk * 5), Iw(Nk * 5), Dw(Nk);
    im2feature(P1[k].d.data(), Wk, Hk, I1.data());
    im2feature(P2[k].d.data(), Wk, Hk, I2.data());
    img_round_all(I1.data(), Nk * 5, p->image_storage);
    img_round_all(I2.data(), Nk * 5, p->image_storage);
    const vector<double>& D1 = PD1[k].d;
    const vector<double>& D2 = PD2[k].d;
    const orc_camera camk = downsized(p->cam, Wk);
    if (k == L - 1) {
      u.assign(Nk, 0.0); v.assign(Nk, 0.0); w.assign(Nk, 0.0);
      Iw = I2; Dw = D2;
    } else {
      const int Wp = P1[k + 1].W, Hp = P1[k + 1].H;
      vector<double> nu(Nk), nv(Nk), nw(Nk);
      resize_dims(u.data(), Wp, Hp, 1, Wk, Hk, nu.data());
      resize_dims(v.data(), Wp, Hp, 1, Wk, Hk, nv.data());
      resize_dims(w.data(), Wp, Hp, 1, Wk, Hk, nw.data());
      const double s = 1 / ratio;
      for (size_t l = 0; l < Nk; l++) { nu[l] *= s; nv[l] *= s; }
      u.swap(nu); v.swap(nv); w.swap(nw);
      warp_cpu(I1.data(), D1.data(), I2.data(), D2.data(), u.data(), v.data(), w.data(), Wk, Hk, 5, Iw.data(), Dw.data());
      img_round_all(Iw.data(), Nk * 5, p->image_storage);
    }
    vector<double> mag(Nk), reg(Nk * 2);
    depth_mag(D1.data(), Wk, Hk, (int)p->cam.xRes, (int)p->cam.yRes, p->cam.focalLength, mag.data());
    reg_weights(hpb ? hpb + hpbOff[k] : nullptr, D1.data(), Wk, Hk, p->regularization_type, reg.data());
    orc_level_in in;
    in.W = Wk; in.H = Hk; in.focal_level = camk.focalLength;
    in.I1 = I1.data(); in.I2 = I2.data(); in.warpI2 = Iw.data();
    in.D1 = D1.data(); in.D2 = D2.data(); in.warpD2 = Dw.data();
    in.delta = Pdelta[k].d.data(); in.s = Ps[k].d.data(); in.m = mag.data(); in.r = reg.data();
    orc_level_stats ls;
    level_solve_dispatch(&in, p, k < L - 1, u.data(), v.data(), w.data(), &ls);
    if (st) {
      st->width[k] = Wk; st->height[k] = Hk;
      st->outer_exec[k] = ls.outer_exec; st->sweeps_exec[k] = ls.sweeps_exec; st->last_stat[k] = ls.last_stat;
      st->residual_checks[k] = ls.residual_checks; st->residual[k] = ls.last_residual;
      st->energy[k] = ls.last_energy;
    }
  }
  std::memcpy(uo, u.data(), N * sizeof(double));
  std::memcpy(vo, v.data(), N * sizeof(double));
  std::memcpy(wo, w.data(), N * sizeof(double));
  return 0;
}

/* sceneFlowWithInit.cpp:7-81 (full resolution only, no early exit) */
int orc_with_init(const double* im1, const double* dm1, const double* im2, const double* dm2, const float* sigma1,
                  const float* hpb0, const double* est, int W, int H, const orc_params* p, double* uo, double* vo,
                  double* wo, orc_stats* st) {
  const size_t N = (size_t)W * H;
  // no map given: computeHeuristicPb(Im1, dm1) (sceneFlowWithInit.cpp:51)
  vector<float> hpbOwn;
  if (p->regularization_type == ORC_REG_HPB && !hpb0) {
    hpbOwn.resize(2 * N);
    orc_heuristic_pb(im1, dm1, W, H, hpbOwn.data());
    hpb0 = hpbOwn.data();
  }
  vector<double> u(N), v(N), w(N);
  for (size_t l = 0; l < N; l++) { u[l] = est[l * 3]; v[l] = est[l * 3 + 1]; w[l] = est[l * 3 + 2]; }
  vector<double> d1(dm1, dm1 + N), d2(dm2, dm2 + N);
  median_fill(d1.data(), W, H, 1);
  median_fill(d2.data(), W, H, 1);
  vector<double> delta(N), sm(N);
  orc_depth_data_weights(sigma1, W, H, p->depth_vs_color_data_weight, delta.data());
  for (size_t l = 0; l < N; l++) sm[l] = dsq(p->cam.focalLength);
  vector<double> I1(N * 5), I2(N * 5), Iw(N * 5), Dw(N);
  im2feature(im1, W, H, I1.data());
  im2feature(im2, W, H, I2.data());
  img_round_all(I1.data(), N * 5, p->image_storage);
  img_round_all(I2.data(), N * 5, p->image_storage);
  warp_cpu(I1.data(), d1.data(), I2.data(), d2.data(), u.data(), v.data(), w.data(), W, H, 5, Iw.data(), Dw.data());
  img_round_all(Iw.data(), N * 5, p->image_storage);
  vector<double> reg(N * 2), mag(N);
  reg_weights(hpb0, d1.data(), W, H, p->regularization_type, reg.data());
  depth_mag(d1.data(), W, H, (int)p->cam.xRes, (int)p->cam.yRes, p->cam.focalLength, mag.data());
  orc_level_in in;
  in.W = W; in.H = H; in.focal_level = p->cam.focalLength;
  in.I1 = I1.data(); in.I2 = I2.data(); in.warpI2 = Iw.data();
  in.D1 = d1.data(); in.D2 = d2.data(); in.warpD2 = Dw.data();
  in.delta = delta.data(); in.s = sm.data(); in.m = mag.data(); in.r = reg.data();
  orc_level_stats ls;
  level_solve_dispatch(&in, p, 0, u.data(), v.data(), w.data(), &ls);
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->n_levels = 1; st->width[0] = W; st->height[0] = H;
    st->outer_exec[0] = ls.outer_exec; st->sweeps_exec[0] = ls.sweeps_exec; st->last_stat[0] = ls.last_stat;
    st->residual_checks[0] = ls.residual_checks; st->residual[0] = ls.last_residual;
    st->energy[0] = ls.last_energy;
  }
  std::memcpy(uo, u.data(), N * sizeof(double));
  std::memcpy(vo, v.data(), N * sizeof(double));
  std::memcpy(wo, w.data(), N * sizeof(double));
  return 0;
}

void orc_set_threads(int n) { pool().set_threads(n); }
int orc_threads(void) { return pool().threads(); }

}  // extern "C"
