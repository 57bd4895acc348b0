// kernels_solve.hip -- one pyramid level of the variational scene-flow solve on gfx950.
//
// Restates sceneFlowSingleScale (src/sceneFlowSingleScaleJacobiCUDA.cu:660-969):
// derivatives + occlusion weights once per outer iteration, robust weights and the
// compact system per inner iteration, the red-black relaxation, then u += du and
// the re-warp.  Arithmetic follows the reference's precision exactly: images,
// derivatives, products and u in f64; the system, du, q and the device weight
// copies in f32; the squares inside the robust weights in f32 (sqrCU, .cu:44-47).
// With the feature images in 16 bits (a precision variant, not in the
// reference) every per-outer stage formula -- derivatives, occlusion weight,
// penalty derivatives, Laplacians, assembly, the in-solve warp -- is evaluated
// in f32 on f32-rounded inputs, in the same order (StageOf).
// Built with -ffp-contract=off so every rounding happens where the source says.
//
// Every kernel runs over a work-list of (pair, level) items: block b belongs
// to the item whose block range holds b.  Early exits are decided on the
// device: a kernel reads its item's SolveCtl at entry and returns when that
// pair's outer loop or current relaxation has stopped, so the host enqueues
// whole levels of whole batches without waiting.
#include <type_traits>

#include "internal.hpp"
#include "records.hpp"
#include "relax_pixel.hpp"

namespace rgbdflow {

namespace {

// sqrCU (.cu:44-47): the argument is rounded to f32 and squared in f32.
__device__ __forceinline__ float sqr_f32(double x) {
  const float f = (float)x;
  return f * f;
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// The stage arithmetic of a feature-image storage S: f64 for the reference's
// f64 images, f32 for the 16-bit ones.
template <int S> struct StageOf { typedef double T; };
template <> struct StageOf<kImgF16> { typedef float T; };
template <> struct StageOf<kImgBF16> { typedef float T; };
template <class T> constexpr bool kF32 = std::is_same<T, float>::value;

// 0.5 / sqrt(x^2 + eps) (.cu:356-369): f64 from the f32 square (sqrCU), or all f32.
template <class T>
__device__ __forceinline__ T half_rsqrt_sq(T x, double eps) {
  if constexpr (kF32<T>) return .5f / sqrtf(x * x + (float)eps);
  else return .5 / sqrt((double)sqr_f32(x) + eps);
}

// One block per item: zero the item's control block.
__global__ void k_ctl_reset(Views w) {
  const LevelView& v = w.items[blockIdx.x];
  unsigned int* p = reinterpret_cast<unsigned int*>(v.ctl);
  const int n = (int)(sizeof(SolveCtl) / sizeof(unsigned int));
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// x / D in f64 for a small integer constant D, bit-equal to the IEEE
// division: Markstein's correction with the correctly rounded reciprocal of D
// (IBM J. Res. Dev. 34(1), 1990) away from zero, underflow and overflow; the
// division itself elsewhere (the CPU check of this recipe for D = 12, 5 and 3
// is oracle/divcheck.cpp).
template <int D>
__device__ __forceinline__ double div_const(double x) {
  const double ax = fabs(x);
  if (__builtin_expect(ax >= 0x1p-960 && ax <= 0x1p960, 1)) {
    constexpr double y = 1.0 / D;
    const double q0 = x * y;
    return __builtin_fma(__builtin_fma(-(double)D, q0, x), y, q0);
  }
  return x / D;
}
// The same in f32: Markstein's correction for |x| in [2^-100, 2^100] (the
// quotient, the product and the remainder stay normal or exact there), the
// IEEE division elsewhere; oracle/divcheck.cpp checks it bit for bit over a
// whole binade, which the power-of-two scaling carries to the domain.
template <int D>
__device__ __forceinline__ float div_const(float x) {
  const float ax = fabsf(x);
  if (__builtin_expect(ax >= 0x1p-100f && ax <= 0x1p100f, 1)) {
    constexpr float y = 1.0f / D;
    const float q0 = x * y;
    return __builtin_fmaf(__builtin_fmaf(-(float)D, q0, x), y, q0);
  }
  return x / (float)D;
}
template <class T>
__device__ __forceinline__ T div12(T x) { return div_const<12>(x); }

// Image and depth derivatives of one pixel (.cu:84-147): x/y derivatives of
// 0.4*I1 + 0.6*warped I2 with [1,-8,0,8,-1]/12, clamp-to-edge; It = Iw - I1;
// depth derivatives clamped to +-0.04.
template <class T>
struct PixDerivs {
  T Ix[kFeat], Iy[kFeat], It[kFeat];
  T Dx, Dy, Dt;
};

// The stencil input 0.4 I1 + 0.6 Iw of one value pair.
template <class T>
__device__ __forceinline__ T blend(double a, double b) { return (T).4 * (T)a + (T).6 * (T)b; }
template <class T>
__device__ __forceinline__ T clamp_mg(T x) {
  const T mg = (T).04;
  return fmin(fmax(x, -mg), mg);
}

template <int S, class T = typename StageOf<S>::T>
__device__ __forceinline__ void derivs_at(const LevelView& v, int l, int i, int j, PixDerivs<T>& d) {
  const int W = v.W, H = v.H, N = v.N;
  const T g[5] = {1, -8, 0, 8, -1};
  for (int k = 0; k < kFeat; k++) {
    const size_t kN = (size_t)k * N;
    T sx = 0, sy = 0;
    for (int m = 0; m < 5; m++) {
      const size_t o = kN + i * W + clampi(j - 2 + m, 0, W - 1);
      sx += g[m] * blend<T>(img_ld<S>(v.I1, o), img_ld<S>(v.Iw, o));
    }
    for (int m = 0; m < 5; m++) {
      const size_t o = kN + clampi(i - 2 + m, 0, H - 1) * W + j;
      sy += g[m] * blend<T>(img_ld<S>(v.I1, o), img_ld<S>(v.Iw, o));
    }
    d.Ix[k] = div12(sx);
    d.Iy[k] = div12(sy);
    d.It[k] = (T)img_ld<S>(v.Iw, kN + l) - (T)img_ld<S>(v.I1, kN + l);
  }
  T sx = 0, sy = 0;
  for (int m = 0; m < 5; m++) {
    const int o = i * W + clampi(j - 2 + m, 0, W - 1);
    sx += g[m] * blend<T>(gp(v.D1)[o], gp(v.Dw)[o]);
  }
  for (int m = 0; m < 5; m++) {
    const int o = clampi(i - 2 + m, 0, H - 1) * W + j;
    sy += g[m] * blend<T>(gp(v.D1)[o], gp(v.Dw)[o]);
  }
  d.Dx = clamp_mg(div12(sx));
  d.Dy = clamp_mg(div12(sy));
  d.Dt = clamp_mg((T)gp(v.Dw)[l] - (T)gp(v.D1)[l]);
}

// .cu:640-648: occlusion weight from the UNWARPED D2 (d1, u2, d2 at the pixel).
template <class T>
__device__ __forceinline__ float occlusion_of(double d1, double u2, double d2) {
  if constexpr (kF32<T>) {
    const float f = (float)d1;
    return (f + (float)u2 < (float)d2 + .02f * (f * f)) ? 1.f : (float).01;
  } else {
    return (d1 + u2 < d2 + .02 * (double)sqr_f32(d1)) ? 1.f : (float).01;
  }
}
template <class T>
__device__ __forceinline__ float occlusion_at(const LevelView& v, int l) {
  return occlusion_of<T>(gp(v.D1)[l], gp(v.u)[2 * v.N + l], gp(v.D2)[l]);
}

// Derivatives + occlusion stored for the inner iterations, du = 0 (.cu:776);
// also zeroes the item's check slots of this outer iteration.
template <int S>
__global__ void __launch_bounds__(kBlock) k_derivs(Views w, int n_slots) {
  typedef typename StageOf<S>::T T;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int it = work_item(w, blk);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  const int l = (blk - w.start[it]) * kBlock + threadIdx.x;
  const int stride = (w.start[it + 1] - w.start[it]) * kBlock;
  for (int t = l; t < n_slots; t += stride) v.ctl->maxdiff[t] = 0u;
  if (l >= v.N) return;
  const int N = v.N;
  const int i = l / v.W, j = l - i * v.W;
  PixDerivs<T> d;
  derivs_at<S>(v, l, i, j, d);
  for (int k = 0; k < kFeat; k++) {
    gp(v.Ix)[(size_t)k * N + l] = d.Ix[k];
    gp(v.Iy)[(size_t)k * N + l] = d.Iy[k];
    gp(v.It)[(size_t)k * N + l] = d.It[k];
  }
  gp(v.Dx)[l] = d.Dx;
  gp(v.Dy)[l] = d.Dy;
  gp(v.Dt)[l] = d.Dt;
  gstore(v.du + l, make_float4(0.f, 0.f, 0.f, 0.f));
  gp(v.q)[l] = occlusion_at<T>(v, l);
}

__device__ __forceinline__ float du_comp(const float4& d, int c) { return c == 0 ? d.x : (c == 1 ? d.y : d.z); }

// .cu:156-182: forward differences of u + du (zero on the last row/column).
// zero_du: du is known to be zero (first inner iteration), the same sums
// with du's terms written as 0.
template <class T, bool zero_du>
__device__ __forceinline__ void flow_grad(const LevelView& v, int c, int l, int i, int j, T& gx, T& gy) {
  const RGBDFLOW_GLOBAL double* uc = gp(v.u) + (size_t)c * v.N;
  const T dl = zero_du ? (T)0 : (T)du_comp(gload(v.du + l), c);
  const T dd = zero_du || i == v.H - 1 ? (T)0 : (T)du_comp(gload(v.du + l + v.W), c);
  const T dr = zero_du || j == v.W - 1 ? (T)0 : (T)du_comp(gload(v.du + l + 1), c);
  gy = (i == v.H - 1) ? (T)0 : ((T)uc[l + v.W] + dd - (T)uc[l] - dl);
  gx = (j == v.W - 1) ? (T)0 : ((T)uc[l + 1] + dr - (T)uc[l] - dl);
}

// .cu:354-355: smoothness weight phi' of the Charbonnier penalty from the
// forward differences (x, y) of the three flow components and the pixel's
// depth-smoothness weight.
template <class T>
__device__ __forceinline__ T phi_from_grads(T ux, T uy, T vx, T vy, T wx, T wy, float sw) {
  if constexpr (kF32<T>) {
    const float t = ux * ux + uy * uy + vx * vx + vy * vy + sw * (wx * wx + wy * wy);
    return .5f / sqrtf(t + 1e-6f);
  } else {
    const float t = sqr_f32(ux) + sqr_f32(uy) + sqr_f32(vx) + sqr_f32(vy) + sw * (sqr_f32(wx) + sqr_f32(wy));
    return 0.5 / sqrt((double)t + 1e-6);
  }
}

// phi' at pixel l.
template <class T, bool zero_du>
__device__ __forceinline__ T phi_at(const LevelView& v, int l, int i, int j) {
  T ux, uy, vx, vy, wx, wy;
  flow_grad<T, zero_du>(v, 0, l, i, j, ux, uy);
  flow_grad<T, zero_du>(v, 1, l, i, j, vx, vy);
  flow_grad<T, zero_du>(v, 2, l, i, j, wx, wy);
  return phi_from_grads<T>(ux, uy, vx, vy, wx, wy, gp(v.sW)[l]);
}

template <int S>
__global__ void __launch_bounds__(kBlock) k_phi(Views w) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int it = work_item(w, blk);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  const int l = (blk - w.start[it]) * kBlock + threadIdx.x;
  if (l >= v.N) return;
  const int i = l / v.W, j = l - i * v.W;
  gp(v.phi)[l] = phi_at<typename StageOf<S>::T, false>(v, l, i, j);
}

// The values around one pixel that the Laplacians and the assembly read
// (zeros for neighbours outside the image, which are never used).
template <class T>
struct PixNbr {
  T uC[3], uL[3], uR[3], uU[3], uD[3];  // u at the pixel and its left, right, upper, lower neighbours
  float rR, rRl, rD, rDu;  // right weight at the pixel and its left neighbour; down weight at the pixel and above
  float sC, sL, sU;        // depth-smoothness weights at the pixel, its left and its upper neighbour
  float dW, mW;            // depth-data and depth-magnitude weights at the pixel
};

template <class T>
__device__ __forceinline__ void load_nbr(const LevelView& v, int l, int i, int j, PixNbr<T>& nb) {
  const int W = v.W, H = v.H, N = v.N;
  const bool hl = j > 0, hr = j < W - 1, hu = i > 0, hd = i < H - 1;
  const RGBDFLOW_GLOBAL double* u = gp(v.u);
  for (int c = 0; c < 3; c++) {
    const size_t o = (size_t)c * N + l;
    nb.uC[c] = (T)u[o];
    nb.uL[c] = hl ? (T)u[o - 1] : (T)0;
    nb.uR[c] = hr ? (T)u[o + 1] : (T)0;
    nb.uU[c] = hu ? (T)u[o - W] : (T)0;
    nb.uD[c] = hd ? (T)u[o + W] : (T)0;
  }
  const RGBDFLOW_GLOBAL float* rR = gp(v.rW);
  const RGBDFLOW_GLOBAL float* rD = gp(v.rW) + N;
  const RGBDFLOW_GLOBAL float* sW = gp(v.sW);
  nb.rR = rR[l];
  nb.rRl = hl ? rR[l - 1] : 0.f;
  nb.rD = rD[l];
  nb.rDu = hu ? rD[l - W] : 0.f;
  nb.sC = sW[l];
  nb.sL = hl ? sW[l - 1] : 0.f;
  nb.sU = hu ? sW[l - W] : 0.f;
  nb.dW = gp(v.dW)[l];
  nb.mW = gp(v.mW)[l];
}

// .cu:240-333: phi- and r-weighted Laplacian of component c of u (not u+du),
// normalised by the mean regularisation weight of the existing edges.  pC,
// pL, pU: phi at the pixel, its left and its upper neighbour.  The mean
// (norm / count, the same for the three components) comes from
// laplacian_mean: 0 when |norm| < 1e-5 (the Laplacian is then 0).
template <class T>
__device__ __forceinline__ T laplacian_mean(const PixNbr<T>& nb, int W, int H, int i, int j) {
  T norm = 0;
  int count = 0;
  if (j < W - 1) { norm += nb.rR; count++; }
  if (j > 0) { norm += nb.rRl; count++; }
  if (i < H - 1) { norm += nb.rD; count++; }
  if (i > 0) { norm += nb.rDu; count++; }
  if (fabs(norm) < (T)1e-5) return (T)0;
  // norm / count: exact scalings for 2 and 4 (norm >= 1e-5: no underflow),
  // the exact constant division for 3
  return count == 3 ? div_const<3>(norm) : count == 2 ? norm * (T)0.5 : count == 4 ? norm * (T)0.25 : norm;
}
template <class T>
__device__ __forceinline__ T laplacian(const PixNbr<T>& nb, int c, int W, int H, int i, int j, T pC, T pL, T pU,
                                       T mean) {
  const T uc = nb.uC[c];
  T lap = 0;
  if (j < W - 1) lap -= (nb.uR[c] - uc) * pC * (T)nb.rR;
  if (j > 0) lap += (uc - nb.uL[c]) * pL * (T)nb.rRl;
  if (i < H - 1) lap -= (nb.uD[c] - uc) * pC * (T)nb.rD;
  if (i > 0) lap += (uc - nb.uU[c]) * pU * (T)nb.rDu;
  return mean == (T)0 ? (T)0 : lap / mean;
}

// .cu:356-369 (psi', pi'), .cu:189-230 (products), .cu:477-575 (assembly),
// written as the compact per-pixel system (Coef).  d0, d1, d2: du at the
// pixel; pC, pL, pU: phi at the pixel, its left and upper neighbours (pL, pU
// unused where the neighbour does not exist).
template <class T>
__device__ __forceinline__ void assemble(const LevelView& v, const SolveScalars& s, int l, int i, int j, T d0, T d1,
                                         T d2, const PixDerivs<T>& dv, T ql, T pC, T pL, T pU, const PixNbr<T>& nb) {
  const int W = v.W, H = v.H;
  const T alpha = (T)s.alpha, gamma = (T)s.gamma;
  T sxy = 0, sxx = 0, syy = 0, stx = 0, sty = 0;
  T psi[kFeat];
  for (int k = 0; k < kFeat; k++) {
    const T tt = (dv.It[k] + dv.Ix[k] * d0 + dv.Iy[k] * d1) * ql;
    psi[k] = half_rsqrt_sq<T>(tt, 1e-6) * (T)s.color_weight;
  }
  for (int k = 0; k < kFeat; k++) {
    sxy += psi[k] * dv.Ix[k] * dv.Iy[k];
    sxx += psi[k] * dv.Ix[k] * dv.Ix[k];
    syy += psi[k] * dv.Iy[k] * dv.Iy[k];
  }
  for (int k = 0; k < kFeat; k++) stx += psi[k] * dv.Ix[k] * dv.It[k];
  for (int k = 0; k < kFeat; k++) sty += psi[k] * dv.Iy[k] * dv.It[k];
  static_assert(kFeat == 5, "channel means divide by 5");
  const T ixy = div_const<5>(sxy), ix2 = div_const<5>(sxx), iy2 = div_const<5>(syy);
  const T itx = div_const<5>(stx * ql * ql), ity = div_const<5>(sty * ql * ql);

  const T Dx = dv.Dx, Dy = dv.Dy, Dt = dv.Dt;
  const T t2 = Dt + Dx * d0 + Dy * d1 - d2;
  const T P = half_rsqrt_sq<T>(t2, 1e-10) * (T)nb.dW;
  const T dxy = P * Dx * Dy, dx2 = P * Dx * Dx, dy2 = P * Dy * Dy;
  const T dtx = P * Dx * Dt, dty = P * Dy * Dt;

  const T mean = laplacian_mean(nb, W, H, i, j);
  const T uL = laplacian(nb, 0, W, H, i, j, pC, pL, pU, mean), vL = laplacian(nb, 1, W, H, i, j, pC, pL, pU, mean);
  const T wL = (T)nb.sC * laplacian(nb, 2, W, H, i, j, pC, pL, pU, mean);

  T wsXY = 0, wsZ = 0;
  float eRxy = 0.f, eRz = 0.f, eDxy = 0.f, eDz = 0.f;
  if (j > 0) { const T a = pL * (T)nb.rRl, b = a * (T)nb.sL; wsXY += a; wsZ += b; }
  if (i > 0) { const T a = pU * (T)nb.rDu, b = a * (T)nb.sU; wsXY += a; wsZ += b; }
  if (j < W - 1) {
    const T a = pC * (T)nb.rR, b = a * (T)nb.sC;
    eRxy = (float)(alpha * a); eRz = (float)(alpha * b); wsXY += a; wsZ += b;
  }
  if (i < H - 1) {
    const T a = pC * (T)nb.rD, b = a * (T)nb.sC;
    eDxy = (float)(alpha * a); eDz = (float)(alpha * b); wsXY += a; wsZ += b;
  }
  const T u0 = nb.uC[0], u1 = nb.uC[1], u2 = nb.uC[2];
  const T m = (T)nb.mW;
  float t;
  t = (float)(-(ix2 + dx2 + alpha * wsXY));
  const float Buu = (float)((T)t + -(gamma));
  const float Buv = (float)(-ixy - dxy);
  const float Buw = (float)(P * Dx);
  t = (float)(-(-itx - dtx + alpha * -uL));
  const float ru = (float)((T)t + -(-gamma * u0));
  t = (float)(-(iy2 + dy2 + alpha * wsXY));
  const float Bvv = (float)((T)t + -(gamma));
  const float Bvw = (float)(P * Dy);
  t = (float)(-(-ity - dty + alpha * -vL));
  const float rv = (float)((T)t + -(-gamma * u1));
  t = (float)(-(P + alpha * wsZ));
  const float Bww = (float)((T)t + -(gamma * m));
  t = (float)(-(P * Dt + alpha * -wL));
  const float rw = (float)((T)t + -(-gamma * m * u2));
  const float rec[kCoef] = {eRxy, eRz, eDxy, eDz, Buu, Bvv, Bww, Buv, Buw, Bvw, ru, rv, rw};
  store_record(v.coef, v.storage, (size_t)v.N, (size_t)l, rec);
}

// The system of an inner iteration after the first, from the stored
// derivatives, occlusion weights and phi (k_phi) and the current du.
template <int S>
__global__ void __launch_bounds__(kBlock) k_system(Views w, SolveScalars s) {
  typedef typename StageOf<S>::T T;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int it = work_item(w, blk);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  const int l = (blk - w.start[it]) * kBlock + threadIdx.x;
  if (l >= v.N) return;
  const int W = v.W, N = v.N;
  const int i = l / W, j = l - i * W;
  PixDerivs<T> d;
  for (int k = 0; k < kFeat; k++) {
    d.Ix[k] = (T)gp(v.Ix)[(size_t)k * N + l];
    d.Iy[k] = (T)gp(v.Iy)[(size_t)k * N + l];
    d.It[k] = (T)gp(v.It)[(size_t)k * N + l];
  }
  d.Dx = (T)gp(v.Dx)[l];
  d.Dy = (T)gp(v.Dy)[l];
  d.Dt = (T)gp(v.Dt)[l];
  const T pC = (T)gp(v.phi)[l], pL = j > 0 ? (T)gp(v.phi)[l - 1] : (T)0, pU = i > 0 ? (T)gp(v.phi)[l - W] : (T)0;
  const float4 du = gload(v.du + l);
  PixNbr<T> nb;
  load_nbr(v, l, i, j, nb);
  assemble<T>(v, s, l, i, j, (T)du.x, (T)du.y, (T)du.z, d, (T)gp(v.q)[l], pC, pL, pU, nb);
}

// The first inner iteration of an outer iteration in one pass: derivatives,
// occlusion weight, du = 0, phi at the pixel and at its left and upper
// neighbours (du is zero), and the system -- the values k_derivs, k_phi and
// k_system compute, without storing the intermediates.  Each block owns a
// kOtW x kOtH tile and reads everything it needs in one load phase, all
// loads in flight together: its own pixels' centre values into registers,
// and into LDS the stencil inputs 0.4*I1 + 0.6*Iw of the tile and a 2-pixel
// border (clamped to the image, as the stencils clamp), and u, the
// regularisation and depth-smoothness weights of the tile and a 1-pixel
// border (zeros outside the image).  phi of the tile, its left column and
// upper row then comes from LDS, and so does every neighbour value of the
// Laplacians and the assembly: after the load phase a pixel touches HBM only
// to store its record.  store: also store derivatives, occlusion weights and
// phi for later inner iterations.  Zeroes the item's check slots of this
// outer iteration.
constexpr int kRw = kOtW + 4, kRh = kOtH + 4;  // stencil region (2-pixel border)
constexpr int kUw = kOtW + 2, kUh = kOtH + 2;  // neighbourhood region (1-pixel border)
constexpr int kPw = kOtW + 1, kPh = kOtH + 1;  // phi region (tile, left column, upper row)

// (the f32 stage asks for 6 waves per SIMD: its registers fit without spills,
// and its halved LDS tiles allow them; the f64 stage is held to 4 by its LDS)
template <int S>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(S == kImgF64 ? 1 : 6)))
k_outer_first(Views w, int n_slots, SolveScalars s, int store) {
  typedef typename StageOf<S>::T T;
  __shared__ T Ms[kFeat + 1][kRh * kRw];  // 0.4*I1 + 0.6*Iw (5 features, depth)
  __shared__ T Us[3][kUh * kUw];          // u
  __shared__ float Ss[kUh * kUw], RRs[kUh * kUw], RDs[kUh * kUw];  // sW, rW right, rW down
  __shared__ T Ps[kPh * kPw];             // phi
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int it = work_item(w, blk);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  const int lb = blk - w.start[it];
  {
    const int stride = (w.start[it + 1] - w.start[it]) * kBlock;
    for (int t = lb * kBlock + threadIdx.x; t < n_slots; t += stride) v.ctl->maxdiff[t] = 0u;
  }
  const int W = v.W, H = v.H, N = v.N;
  const int tx = (W + kOtW - 1) / kOtW;
  const int x0 = (lb % tx) * kOtW, y0 = (lb / tx) * kOtH;
  const int ty = threadIdx.x / kOtW, txx = threadIdx.x - ty * kOtW;
  const int i = y0 + ty, j = x0 + txx;
  const bool inside = i < H && j < W;
  const int l = inside ? i * W + j : 0;
  const RGBDFLOW_GLOBAL double* D1 = gp(v.D1);
  const RGBDFLOW_GLOBAL double* Dw = gp(v.Dw);
  // the pixel's own values
  T i1c[kFeat], iwc[kFeat];
  double d1c = 0, dwc = 0, d2c = 0;
  float dwt = 0.f, mwt = 0.f;
  if (inside) {
    for (int k = 0; k < kFeat; k++) {
      i1c[k] = (T)img_ld<S>(v.I1, (size_t)k * N + l);
      iwc[k] = (T)img_ld<S>(v.Iw, (size_t)k * N + l);
    }
    d1c = D1[l];
    dwc = Dw[l];
    d2c = gp(v.D2)[l];
    dwt = gp(v.dW)[l];
    mwt = gp(v.mW)[l];
  }
  for (int t = threadIdx.x; t < kRh * kRw; t += kBlock) {
    const int ry = t / kRw, rx = t - ry * kRw;
    const int o = clampi(y0 - 2 + ry, 0, H - 1) * W + clampi(x0 - 2 + rx, 0, W - 1);
    for (int k = 0; k < kFeat; k++)
      Ms[k][t] = blend<T>(img_ld<S>(v.I1, (size_t)k * N + o), img_ld<S>(v.Iw, (size_t)k * N + o));
    Ms[kFeat][t] = blend<T>(D1[o], Dw[o]);
  }
  for (int t = threadIdx.x; t < kUh * kUw; t += kBlock) {
    const int ry = t / kUw, rx = t - ry * kUw;
    const int gy = y0 - 1 + ry, gx = x0 - 1 + rx;
    if (gy >= 0 && gx >= 0 && gy < H && gx < W) {
      const int o = gy * W + gx;
      for (int c = 0; c < 3; c++) Us[c][t] = (T)gp(v.u)[(size_t)c * N + o];
      Ss[t] = gp(v.sW)[o];
      RRs[t] = gp(v.rW)[o];
      RDs[t] = gp(v.rW)[(size_t)N + o];
    } else {
      for (int c = 0; c < 3; c++) Us[c][t] = (T)0;
      Ss[t] = RRs[t] = RDs[t] = 0.f;
    }
  }
  __syncthreads();
  // phi with du = 0 (flow_grad<true>, phi_at<true>) from the staged u and sW
  for (int t = threadIdx.x; t < kPh * kPw; t += kBlock) {
    const int py = t / kPw, px = t - py * kPw;
    const int pi = y0 - 1 + py, pj = x0 - 1 + px;
    T r = 0;
    if (pi >= 0 && pj >= 0 && pi < H && pj < W) {
      const int c = py * kUw + px;  // (pi, pj) in the neighbourhood region
      const bool last_row = pi == H - 1, last_col = pj == W - 1;
      T gx[3], gy[3];
      for (int k = 0; k < 3; k++) {
        gy[k] = last_row ? (T)0 : (Us[k][c + kUw] + (T)0 - Us[k][c] - (T)0);
        gx[k] = last_col ? (T)0 : (Us[k][c + 1] + (T)0 - Us[k][c] - (T)0);
      }
      r = phi_from_grads<T>(gx[0], gy[0], gx[1], gy[1], gx[2], gy[2], Ss[c]);
    }
    Ps[t] = r;
  }
  __syncthreads();
  if (!inside) return;
  // derivs_at, with the stencil inputs from LDS (same values, same order)
  PixDerivs<T> d;
  const T g[5] = {1, -8, 0, 8, -1};
  const int c0 = (ty + 2) * kRw + txx + 2;
  for (int k = 0; k < kFeat; k++) {
    T sx = 0, sy = 0;
    for (int m = 0; m < 5; m++) sx += g[m] * Ms[k][c0 - 2 + m];
    for (int m = 0; m < 5; m++) sy += g[m] * Ms[k][c0 + (m - 2) * kRw];
    d.Ix[k] = div12(sx);
    d.Iy[k] = div12(sy);
    d.It[k] = iwc[k] - i1c[k];
  }
  {
    T sx = 0, sy = 0;
    for (int m = 0; m < 5; m++) sx += g[m] * Ms[kFeat][c0 - 2 + m];
    for (int m = 0; m < 5; m++) sy += g[m] * Ms[kFeat][c0 + (m - 2) * kRw];
    d.Dx = clamp_mg(div12(sx));
    d.Dy = clamp_mg(div12(sy));
    d.Dt = clamp_mg((T)dwc - (T)d1c);
  }
  PixNbr<T> nb;
  const int cu = (ty + 1) * kUw + txx + 1;
  for (int k = 0; k < 3; k++) {
    nb.uC[k] = Us[k][cu];
    nb.uL[k] = Us[k][cu - 1];
    nb.uR[k] = Us[k][cu + 1];
    nb.uU[k] = Us[k][cu - kUw];
    nb.uD[k] = Us[k][cu + kUw];
  }
  nb.rR = RRs[cu];
  nb.rRl = RRs[cu - 1];
  nb.rD = RDs[cu];
  nb.rDu = RDs[cu - kUw];
  nb.sC = Ss[cu];
  nb.sL = Ss[cu - 1];
  nb.sU = Ss[cu - kUw];
  nb.dW = dwt;
  nb.mW = mwt;
  // occlusion_at (.cu:640-648): the unwarped D2
  const float q = occlusion_of<T>(d1c, (double)nb.uC[2], d2c);
  gstore(v.du + l, make_float4(0.f, 0.f, 0.f, 0.f));
  const int p0 = (ty + 1) * kPw + txx + 1;
  const T pC = Ps[p0], pL = Ps[p0 - 1], pU = Ps[p0 - kPw];
  if (store) {
    for (int k = 0; k < kFeat; k++) {
      gp(v.Ix)[(size_t)k * N + l] = d.Ix[k];
      gp(v.Iy)[(size_t)k * N + l] = d.Iy[k];
      gp(v.It)[(size_t)k * N + l] = d.It[k];
    }
    gp(v.Dx)[l] = d.Dx;
    gp(v.Dy)[l] = d.Dy;
    gp(v.Dt)[l] = d.Dt;
    gp(v.q)[l] = q;
    gp(v.phi)[l] = pC;
  }
  assemble<T>(v, s, l, i, j, (T)0, (T)0, (T)0, d, (T)q, pC, pL, pU, nb);
}

// .cu:382-470: one colour of the red-black relaxation, one thread per pixel of
// the colour, neighbours read from global memory (the plain schedule: one
// launch per colour per iteration, as the reference launches it).
__global__ void __launch_bounds__(256) k_relax_plain(Views w, int colour, int slot_base, int n_prev_checks,
                                                     int check_slot) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int it = work_item(w, blk);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  if (relax_stopped(v.ctl, slot_base, n_prev_checks)) return;
  const int W = v.W, H = v.H, N = v.N;
  const int lb = blk - w.start[it];
  const int gx = ((W + 1) / 2 + 63) / 64;
  const int bx = lb % gx, by = lb / gx;
  const int xc = bx * 64 + (threadIdx.x & 63);
  const int i = by * 4 + (threadIdx.x >> 6);
  const int j = 2 * xc + ((i + colour) & 1);
  float change = 0.f;
  if (i < H && j < W) {
    const int l = i * W + j;
    const bool hl = j > 0, hu = i > 0, hr = j < W - 1, hd = i < H - 1;
    const PixCoef c = load_coef(v.coef, v.storage, W, v.N, l, hl, hu);
    float4* x = v.du;
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 o = gload(x + l), L = hl ? gload(x + l - 1) : zero, U = hu ? gload(x + l - W) : zero,
                 R = hr ? gload(x + l + 1) : zero, D = hd ? gload(x + l + W) : zero;
    const float xo[3] = {o.x, o.y, o.z}, xl[3] = {L.x, L.y, L.z}, xu[3] = {U.x, U.y, U.z}, xr[3] = {R.x, R.y, R.z},
                xd[3] = {D.x, D.y, D.z};
    float xn[3];
    if (v.pixel_gs == 2) relax_update_block(c, v.omega, v.omega_c, hl, hu, hr, hd, xl, xu, xr, xd, xo, xn);
    else relax_update(c, v.omega, v.omega_c, hl, hu, hr, hd, xl, xu, xr, xd, xo, xn, v.pixel_gs == 1);
    gstore(x + l, make_float4(xn[0], xn[1], xn[2], 0.f));
    change = fmaxf(fmaxf(fabsf(xn[0] - xo[0]), fabsf(xn[1] - xo[1])), fabsf(xn[2] - xo[2]));
  }
  if (check_slot >= 0) {
    const float m = wave_max(change);
    if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(&v.ctl->maxdiff[check_slot], __float_as_uint(m));
  }
  if (colour == 1 && lb == 0 && threadIdx.x == 0) v.ctl->sweeps_exec += 1;
}

// .cu:934-941 (u += du, GPU warps: floor base, explicit clamps, f32 weights,
// out-of-bounds -> frame 1) and the per-block part of the outer statistic
// sum(du_k^2 * c_k) (.cu:948), accumulated in f64 in a fixed order.
template <int S>
__global__ void __launch_bounds__(kBlock) k_update_warp(Views wl, int want_stat) {
  typedef typename StageOf<S>::T T;
  __shared__ double red[kBlock];
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int item = work_item(wl, blk);
  const LevelView& v = wl.items[item];
  if (v.ctl->outer_done) return;
  const int lb = blk - wl.start[item];
  const int l = lb * kBlock + threadIdx.x;
  const int W = v.W, H = v.H, N = v.N;
  const bool stat = want_stat && v.allow_early;
  double term = 0;
  if (l < N) {
    const int i = l / W, j = l - i * W;
    // the relaxation's result: du, or du_alt if a tiled launch left it there
    const float4 e = gload((v.ctl->du_in_alt ? v.du_alt : v.du) + l);
    const float e0 = e.x, e1 = e.y, e2 = e.z;
    const double u0 = gp(v.u)[l] + (double)e0;
    const double u1 = gp(v.u)[N + l] + (double)e1;
    const double u2 = gp(v.u)[2 * N + l] + (double)e2;
    gp(v.u)[l] = u0;
    gp(v.u)[N + l] = u1;
    gp(v.u)[2 * N + l] = u2;
    const double x = j + u0, y = i + u1;
    if (x < 0 || x > W - 1 || y < 0 || y > H - 1) {
      for (int k = 0; k < kFeat; k++) img_st<S>(v.Iw, (size_t)k * N + l, img_ld<S>(v.I1, (size_t)k * N + l));
      gp(v.Dw)[l] = (T)gp(v.D1)[l] - (T)u2;
    } else {
      const int x0 = clampi((int)floor(x), 0, W - 1), y0 = clampi((int)floor(y), 0, H - 1);
      const int x1 = min(W - 1, x0 + 1), y1 = min(H - 1, y0 + 1);
      const float ax = x - x0, ay = y - y0;
      const float w00 = (1 - ax) * (1 - ay), w01 = (1 - ax) * ay, w10 = ax * (1 - ay), w11 = ax * ay;
      const int o00 = y0 * W + x0, o01 = y1 * W + x0, o10 = y0 * W + x1, o11 = y1 * W + x1;
      for (int k = 0; k < kFeat; k++) {
        const size_t kN = (size_t)k * N;
        img_st<S>(v.Iw, kN + l,
                  w00 * (T)img_ld<S>(v.I2, kN + o00) + w01 * (T)img_ld<S>(v.I2, kN + o01) +
                      w10 * (T)img_ld<S>(v.I2, kN + o10) + w11 * (T)img_ld<S>(v.I2, kN + o11));
      }
      gp(v.Dw)[l] = w00 * (T)gp(v.D2)[o00] + w01 * (T)gp(v.D2)[o01] + w10 * (T)gp(v.D2)[o10] +
                    w11 * (T)gp(v.D2)[o11] - (T)u2;
    }
    if (stat) {
      // sqrCU(du_k) * c_k in f32 (thrustSqrMultFunctor), c = (1, 1, f/D1)
      term = (double)(e0 * e0);
      term += (double)(e1 * e1);
      term += (double)((e2 * e2) * gp(v.cW)[l]);
    }
  }
  if (!stat) return;
  red[threadIdx.x] = term;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) gp(v.partials)[lb] = red[0];
}

// FP64 residual of the relaxation's system at the current increment x, the
// stopping test of RGBDFLOW_EXIT_RESIDUAL.  Per pixel and unknown k:
//   res_k = r_k - (sum_n e_nk x_nk + sum_k' B_kk' x_k')
// (the operator relax_update inverts: left/up edge weights are the
// neighbours' right/down ones, missing neighbours contribute nothing).  Each
// product of two f32 values is exact in f64; the terms are summed in f64 in
// the order written.  Per block: sum res^2, sum r^2 and sum x.(r + res) (the
// energy r'x - 1/2 x'Mx -- convex, -M being positive definite, and
// minimised by the solution -- equals 1/2 sum x.(r + res)), reduced with wave
// shuffles and then across the block's waves in a fixed order.
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(kBlock) k_residual(Views wl, int alt, int slot_base, int n_prev_checks) {
  __shared__ double red[3][kBlock / 64];
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int item = work_item(wl, blk);
  const LevelView& v = wl.items[item];
  if (v.ctl->outer_done || relax_stopped(v.ctl, slot_base, n_prev_checks)) return;
  const int lb = blk - wl.start[item];
  const int l = lb * kBlock + threadIdx.x;
  const int W = v.W, H = v.H, N = v.N;
  const float4* x = alt ? v.du_alt : v.du;
  double rr = 0, bb = 0, ex = 0;
  if (l < N) {
    const int i = l / W, j = l - i * W;
    const bool hl = j > 0, hu = i > 0, hr = j < W - 1, hd = i < H - 1;
    const PixCoef c = load_coef(v.coef, v.storage, W, v.N, l, hl, hu);
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 o = gload(x + l), L = hl ? gload(x + l - 1) : zero, U = hu ? gload(x + l - W) : zero,
                 R = hr ? gload(x + l + 1) : zero, D = hd ? gload(x + l + W) : zero;
    const double xo[3] = {o.x, o.y, o.z}, xl[3] = {L.x, L.y, L.z}, xu[3] = {U.x, U.y, U.z},
                 xr[3] = {R.x, R.y, R.z}, xd[3] = {D.x, D.y, D.z};
    const double B[3][3] = {{c.Buu, c.Buv, c.Buw}, {c.Buv, c.Bvv, c.Bvw}, {c.Buw, c.Bvw, c.Bww}};
    const double rhs[3] = {c.ru, c.rv, c.rw};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double mx = 0;
      if (hl) mx += (double)(k < 2 ? c.lxy : c.lz) * xl[k];
      if (hu) mx += (double)(k < 2 ? c.uxy : c.uz) * xu[k];
      if (hr) mx += (double)(k < 2 ? c.rxy : c.rz) * xr[k];
      if (hd) mx += (double)(k < 2 ? c.dxy : c.dz) * xd[k];
      mx += B[k][0] * xo[0];
      mx += B[k][1] * xo[1];
      mx += B[k][2] * xo[2];
      const double res = rhs[k] - mx;
      rr += res * res;
      bb += rhs[k] * rhs[k];
      ex += xo[k] * (rhs[k] + res);
    }
  }
  rr = wave_sum(rr);
  bb = wave_sum(bb);
  ex = wave_sum(ex);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wv] = rr;
    red[1][wv] = bb;
    red[2][wv] = ex;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double a = 0;
    for (int k = 0; k < kBlock / 64; k++) a += red[threadIdx.x][k];
    gp(v.partials)[3 * lb + threadIdx.x] = a;
  }
}

// One block per item: fixed-order sums of the item's residual partials ->
// SolveCtl::resid / energy; with a check slot, the slot records whether the
// relaxation stops here (0 reads as "stop" and 0.01 as "go on" to every later
// relaxation launch's relax_stopped test).
__global__ void __launch_bounds__(kBlock) k_residual_final(Views w, int slot_base, int n_prev_checks, int check_slot,
                                                           double tol) {
  __shared__ double red[3][kBlock];
  const LevelView& v = w.items[blockIdx.x];
  if (v.ctl->outer_done || relax_stopped(v.ctl, slot_base, n_prev_checks)) return;
  const int nblocks = pix_blocks(v.N);
  double a[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += kBlock)
    for (int k = 0; k < 3; k++) a[k] += gp(v.partials)[3 * b + k];
  for (int k = 0; k < 3; k++) red[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rr = red[0][0], bb = red[1][0];
    const double rel = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
    v.ctl->resid = rel;
    v.ctl->energy = 0.5 * red[2][0];
    v.ctl->resid_checks += 1;
    if (check_slot >= 0) v.ctl->maxdiff[check_slot] = rel <= tol ? 0u : __float_as_uint(.01f);
  }
}

// One block per item: fixed-order sum of the item's partials -> outer
// statistic; sets the item's outer_done when early termination applies
// (.cu:943-951).
__global__ void __launch_bounds__(kBlock) k_stat_final(Views w, int outer, int decide) {
  __shared__ double red[kBlock];
  const LevelView& v = w.items[blockIdx.x];
  if (v.ctl->outer_done || !v.allow_early) return;
  const int nblocks = pix_blocks(v.N);
  double acc = 0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock) acc += gp(v.partials)[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double stat = red[0] / (double)(v.W * v.H * 3);
    v.ctl->stat[outer] = stat;
    if (decide && stat < 4e-4) {
      v.ctl->outer_done = 1;
      v.ctl->outer_exit_at = outer;
    }
  }
}

}  // namespace

void launch_ctl_reset(hipStream_t s, const Views& v) {
  hipLaunchKernelGGL(k_ctl_reset, dim3(v.n), dim3(256), 0, s, v);
}
void launch_derivs(hipStream_t s, const Views& v, int blocks, int n_slots, int img) {
  auto fn = img == kImgF16 ? k_derivs<kImgF16> : img == kImgBF16 ? k_derivs<kImgBF16> : k_derivs<kImgF64>;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, s, v, n_slots);
}
void launch_phi(hipStream_t s, const Views& v, int blocks, int img) {
  auto fn = img == kImgF16 ? k_phi<kImgF16> : img == kImgBF16 ? k_phi<kImgBF16> : k_phi<kImgF64>;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, s, v);
}
void launch_system(hipStream_t s, const Views& v, int blocks, const SolveScalars& k, int img) {
  auto fn = img == kImgF16 ? k_system<kImgF16> : img == kImgBF16 ? k_system<kImgBF16> : k_system<kImgF64>;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, s, v, k);
}
void launch_outer_first(hipStream_t s, const Views& v, int blocks, int n_slots, const SolveScalars& k, int store,
                        int img) {
  auto fn = img == kImgF16    ? k_outer_first<kImgF16>
            : img == kImgBF16 ? k_outer_first<kImgBF16>
                              : k_outer_first<kImgF64>;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, s, v, n_slots, k, store);
}
void launch_relax_plain(hipStream_t s, const Views& v, int blocks, int colour, int slot_base, int n_prev_checks,
                        int check_slot) {
  hipLaunchKernelGGL(k_relax_plain, dim3(blocks), dim3(256), 0, s, v, colour, slot_base, n_prev_checks, check_slot);
}
void launch_update_warp(hipStream_t s, const Views& v, int blocks, int want_stat, int img) {
  auto fn = img == kImgF16 ? k_update_warp<kImgF16> : img == kImgBF16 ? k_update_warp<kImgBF16> : k_update_warp<kImgF64>;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, s, v, want_stat);
}
void launch_stat_final(hipStream_t s, const Views& v, int outer, int decide) {
  hipLaunchKernelGGL(k_stat_final, dim3(v.n), dim3(kBlock), 0, s, v, outer, decide);
}
void launch_residual(hipStream_t s, const Views& v, int blocks, int alt, int slot_base, int n_prev_checks,
                     int check_slot, double tol) {
  hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(kBlock), 0, s, v, alt, slot_base, n_prev_checks);
  hipLaunchKernelGGL(k_residual_final, dim3(v.n), dim3(kBlock), 0, s, v, slot_base, n_prev_checks, check_slot, tol);
}

}  // namespace rgbdflow
