// kernels_solve.hip -- one pyramid level of the variational scene-flow solve on gfx950.
//
// Restates sceneFlowSingleScale (src/sceneFlowSingleScaleJacobiCUDA.cu:660-969):
// derivatives + occlusion weights once per outer iteration, robust weights and the
// compact system per inner iteration, the red-black relaxation, then u += du and
// the re-warp.  Arithmetic follows the reference's precision exactly: images,
// derivatives, products and u in f64; the system, du, q and the device weight
// copies in f32; the squares inside the robust weights in f32 (sqrCU, .cu:44-47).
// Built with -ffp-contract=off so every rounding happens where the source says.
//
// Every kernel runs over a work-list of (pair, level) items: block b belongs
// to the item whose block range holds b.  Early exits are decided on the
// device: a kernel reads its item's SolveCtl at entry and returns when that
// pair's outer loop or current relaxation has stopped, so the host enqueues
// whole levels of whole batches without waiting.
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

// One block per item: zero the item's control block.
__global__ void k_ctl_reset(Views w) {
  const LevelView& v = w.items[blockIdx.x];
  unsigned int* p = reinterpret_cast<unsigned int*>(v.ctl);
  const int n = (int)(sizeof(SolveCtl) / sizeof(unsigned int));
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// x / 12 in f64, bit-equal to the IEEE division: Markstein's correction with
// the correctly rounded reciprocal of 12 (IBM J. Res. Dev. 34(1), 1990) away
// from zero, underflow and overflow; the division itself elsewhere.
__device__ __forceinline__ double div12(double x) {
  const double ax = fabs(x);
  if (__builtin_expect(ax >= 0x1p-960 && ax <= 0x1p960, 1)) {
    const double y = 1.0 / 12.0;
    const double q0 = x * y;
    return __builtin_fma(__builtin_fma(-12.0, q0, x), y, q0);
  }
  return x / 12;
}

// Image and depth derivatives of one pixel (.cu:84-147): x/y derivatives of
// 0.4*I1 + 0.6*warped I2 with [1,-8,0,8,-1]/12, clamp-to-edge; It = Iw - I1;
// depth derivatives clamped to +-0.04.
struct PixDerivs {
  double Ix[kFeat], Iy[kFeat], It[kFeat];
  double Dx, Dy, Dt;
};

__device__ __forceinline__ void derivs_at(const LevelView& v, int l, int i, int j, PixDerivs& d) {
  const int W = v.W, H = v.H, N = v.N;
  const double g[5] = {1, -8, 0, 8, -1};
  for (int k = 0; k < kFeat; k++) {
    const RGBDFLOW_GLOBAL double* a = gp(v.I1) + (size_t)k * N;
    const RGBDFLOW_GLOBAL double* b = gp(v.Iw) + (size_t)k * N;
    double sx = 0, sy = 0;
    for (int m = 0; m < 5; m++) {
      const int o = i * W + clampi(j - 2 + m, 0, W - 1);
      sx += g[m] * (.4 * a[o] + .6 * b[o]);
    }
    for (int m = 0; m < 5; m++) {
      const int o = clampi(i - 2 + m, 0, H - 1) * W + j;
      sy += g[m] * (.4 * a[o] + .6 * b[o]);
    }
    d.Ix[k] = div12(sx);
    d.Iy[k] = div12(sy);
    d.It[k] = b[l] - a[l];
  }
  const double mg = .04;
  double sx = 0, sy = 0;
  for (int m = 0; m < 5; m++) {
    const int o = i * W + clampi(j - 2 + m, 0, W - 1);
    sx += g[m] * (.4 * gp(v.D1)[o] + .6 * gp(v.Dw)[o]);
  }
  for (int m = 0; m < 5; m++) {
    const int o = clampi(i - 2 + m, 0, H - 1) * W + j;
    sy += g[m] * (.4 * gp(v.D1)[o] + .6 * gp(v.Dw)[o]);
  }
  d.Dx = fmin(fmax(div12(sx), -mg), mg);
  d.Dy = fmin(fmax(div12(sy), -mg), mg);
  d.Dt = fmin(fmax(gp(v.Dw)[l] - gp(v.D1)[l], -mg), mg);
}

// .cu:640-648: occlusion weight from the UNWARPED D2.
__device__ __forceinline__ float occlusion_at(const LevelView& v, int l) {
  const double d1 = gp(v.D1)[l];
  return (d1 + gp(v.u)[2 * v.N + l] < gp(v.D2)[l] + .02 * (double)sqr_f32(d1)) ? 1.f : (float).01;
}

// Derivatives + occlusion stored for the inner iterations, du = 0 (.cu:776);
// also zeroes the item's check slots of this outer iteration.
__global__ void __launch_bounds__(kBlock) k_derivs(Views w, int n_slots) {
  const int it = work_item(w, blockIdx.x);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  const int l = (blockIdx.x - w.start[it]) * kBlock + threadIdx.x;
  const int stride = (w.start[it + 1] - w.start[it]) * kBlock;
  for (int t = l; t < n_slots; t += stride) v.ctl->maxdiff[t] = 0u;
  if (l >= v.N) return;
  const int N = v.N;
  const int i = l / v.W, j = l - i * v.W;
  PixDerivs d;
  derivs_at(v, l, i, j, d);
  for (int k = 0; k < kFeat; k++) {
    gp(v.Ix)[(size_t)k * N + l] = d.Ix[k];
    gp(v.Iy)[(size_t)k * N + l] = d.Iy[k];
    gp(v.It)[(size_t)k * N + l] = d.It[k];
  }
  gp(v.Dx)[l] = d.Dx;
  gp(v.Dy)[l] = d.Dy;
  gp(v.Dt)[l] = d.Dt;
  gstore(v.du + l, make_float4(0.f, 0.f, 0.f, 0.f));
  gp(v.q)[l] = occlusion_at(v, l);
}

__device__ __forceinline__ float du_comp(const float4& d, int c) { return c == 0 ? d.x : (c == 1 ? d.y : d.z); }

// .cu:156-182: forward differences of u + du (zero on the last row/column).
// zero_du: du is known to be zero (first inner iteration), the same sums
// with du's terms written as 0.
template <bool zero_du>
__device__ __forceinline__ void flow_grad(const LevelView& v, int c, int l, int i, int j, double& gx, double& gy) {
  const RGBDFLOW_GLOBAL double* uc = gp(v.u) + (size_t)c * v.N;
  const double dl = zero_du ? 0.0 : (double)du_comp(gload(v.du + l), c);
  const double dd = zero_du || i == v.H - 1 ? 0.0 : (double)du_comp(gload(v.du + l + v.W), c);
  const double dr = zero_du || j == v.W - 1 ? 0.0 : (double)du_comp(gload(v.du + l + 1), c);
  gy = (i == v.H - 1) ? 0.0 : (uc[l + v.W] + dd - uc[l] - dl);
  gx = (j == v.W - 1) ? 0.0 : (uc[l + 1] + dr - uc[l] - dl);
}

// .cu:354-355: smoothness weight phi' of the Charbonnier penalty at pixel l.
template <bool zero_du>
__device__ __forceinline__ double phi_at(const LevelView& v, int l, int i, int j) {
  double ux, uy, vx, vy, wx, wy;
  flow_grad<zero_du>(v, 0, l, i, j, ux, uy);
  flow_grad<zero_du>(v, 1, l, i, j, vx, vy);
  flow_grad<zero_du>(v, 2, l, i, j, wx, wy);
  const float t = sqr_f32(ux) + sqr_f32(uy) + sqr_f32(vx) + sqr_f32(vy) + gp(v.sW)[l] * (sqr_f32(wx) + sqr_f32(wy));
  return 0.5 / sqrt((double)t + 1e-6);
}

__global__ void __launch_bounds__(kBlock) k_phi(Views w) {
  const int it = work_item(w, blockIdx.x);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  const int l = (blockIdx.x - w.start[it]) * kBlock + threadIdx.x;
  if (l >= v.N) return;
  const int i = l / v.W, j = l - i * v.W;
  gp(v.phi)[l] = phi_at<false>(v, l, i, j);
}

// .cu:240-333: phi- and r-weighted Laplacian of u (not u+du), normalised by the
// mean regularisation weight of the existing edges.  pC, pL, pU: phi at the
// pixel, its left and its upper neighbour.
__device__ __forceinline__ double laplacian(const LevelView& v, int c, int l, int i, int j, double pC, double pL,
                                            double pU) {
  const RGBDFLOW_GLOBAL double* uc = gp(v.u) + (size_t)c * v.N;
  const RGBDFLOW_GLOBAL float* rR = gp(v.rW);
  const RGBDFLOW_GLOBAL float* rD = gp(v.rW) + v.N;
  const int W = v.W;
  double lap = 0;
  if (j < W - 1) lap -= (uc[l + 1] - uc[l]) * pC * (double)rR[l];
  if (j > 0) lap += (uc[l] - uc[l - 1]) * pL * (double)rR[l - 1];
  if (i < v.H - 1) lap -= (uc[l + W] - uc[l]) * pC * (double)rD[l];
  if (i > 0) lap += (uc[l] - uc[l - W]) * pU * (double)rD[l - W];
  double norm = 0;
  int count = 0;
  if (j < W - 1) { norm += rR[l]; count++; }
  if (j > 0) { norm += rR[l - 1]; count++; }
  if (i < v.H - 1) { norm += rD[l]; count++; }
  if (i > 0) { norm += rD[l - W]; count++; }
  if (fabs(norm) < 1e-5) lap = 0;
  else lap /= (norm / count);
  return lap;
}

// .cu:356-369 (psi', pi'), .cu:189-230 (products), .cu:477-575 (assembly),
// written as the compact per-pixel system (Coef).  d0, d1, d2: du at the
// pixel; pC, pL, pU: phi at the pixel, its left and upper neighbours (pL, pU
// unused where the neighbour does not exist).
__device__ __forceinline__ void assemble(const LevelView& v, const SolveScalars& s, int l, int i, int j, double d0,
                                         double d1, double d2, const PixDerivs& dv, double ql, double pC, double pL,
                                         double pU) {
  const int W = v.W, H = v.H, N = v.N;
  const double alpha = s.alpha, gamma = s.gamma;
  double sxy = 0, sxx = 0, syy = 0, stx = 0, sty = 0;
  double psi[kFeat];
  for (int k = 0; k < kFeat; k++) {
    const double tt = (dv.It[k] + dv.Ix[k] * d0 + dv.Iy[k] * d1) * ql;
    psi[k] = .5 / sqrt((double)sqr_f32(tt) + 1e-6) * s.color_weight;
  }
  for (int k = 0; k < kFeat; k++) {
    sxy += psi[k] * dv.Ix[k] * dv.Iy[k];
    sxx += psi[k] * dv.Ix[k] * dv.Ix[k];
    syy += psi[k] * dv.Iy[k] * dv.Iy[k];
  }
  for (int k = 0; k < kFeat; k++) stx += psi[k] * dv.Ix[k] * dv.It[k];
  for (int k = 0; k < kFeat; k++) sty += psi[k] * dv.Iy[k] * dv.It[k];
  const double ixy = sxy / kFeat, ix2 = sxx / kFeat, iy2 = syy / kFeat;
  const double itx = stx * ql * ql / kFeat, ity = sty * ql * ql / kFeat;

  const double Dx = dv.Dx, Dy = dv.Dy, Dt = dv.Dt;
  const double t2 = Dt + Dx * d0 + Dy * d1 - d2;
  const double P = .5 / sqrt((double)sqr_f32(t2) + 1e-10) * (double)gp(v.dW)[l];
  const double dxy = P * Dx * Dy, dx2 = P * Dx * Dx, dy2 = P * Dy * Dy;
  const double dtx = P * Dx * Dt, dty = P * Dy * Dt;

  const double uL = laplacian(v, 0, l, i, j, pC, pL, pU), vL = laplacian(v, 1, l, i, j, pC, pL, pU);
  const double wL = (double)gp(v.sW)[l] * laplacian(v, 2, l, i, j, pC, pL, pU);

  const RGBDFLOW_GLOBAL float* rR = gp(v.rW);
  const RGBDFLOW_GLOBAL float* rD = gp(v.rW) + N;
  double wsXY = 0, wsZ = 0;
  float eRxy = 0.f, eRz = 0.f, eDxy = 0.f, eDz = 0.f;
  if (j > 0) { const double a = pL * (double)rR[l - 1], b = a * (double)gp(v.sW)[l - 1]; wsXY += a; wsZ += b; }
  if (i > 0) { const double a = pU * (double)rD[l - W], b = a * (double)gp(v.sW)[l - W]; wsXY += a; wsZ += b; }
  if (j < W - 1) {
    const double a = pC * (double)rR[l], b = a * (double)gp(v.sW)[l];
    eRxy = (float)(alpha * a); eRz = (float)(alpha * b); wsXY += a; wsZ += b;
  }
  if (i < H - 1) {
    const double a = pC * (double)rD[l], b = a * (double)gp(v.sW)[l];
    eDxy = (float)(alpha * a); eDz = (float)(alpha * b); wsXY += a; wsZ += b;
  }
  const double u0 = gp(v.u)[l], u1 = gp(v.u)[N + l], u2 = gp(v.u)[2 * N + l];
  const double m = (double)gp(v.mW)[l];
  float t;
  t = (float)(-(ix2 + dx2 + alpha * wsXY));
  const float Buu = (float)((double)t + -(gamma));
  const float Buv = (float)(-ixy - dxy);
  const float Buw = (float)(P * Dx);
  t = (float)(-(-itx - dtx + alpha * -uL));
  const float ru = (float)((double)t + -(-gamma * u0));
  t = (float)(-(iy2 + dy2 + alpha * wsXY));
  const float Bvv = (float)((double)t + -(gamma));
  const float Bvw = (float)(P * Dy);
  t = (float)(-(-ity - dty + alpha * -vL));
  const float rv = (float)((double)t + -(-gamma * u1));
  t = (float)(-(P + alpha * wsZ));
  const float Bww = (float)((double)t + -(gamma * m));
  t = (float)(-(P * Dt + alpha * -wL));
  const float rw = (float)((double)t + -(-gamma * m * u2));
  const float rec[kCoef] = {eRxy, eRz, eDxy, eDz, Buu, Bvv, Bww, Buv, Buw, Bvw, ru, rv, rw};
  store_record(v.coef, v.storage, (size_t)l, rec);
}

// The system of an inner iteration after the first, from the stored
// derivatives, occlusion weights and phi (k_phi) and the current du.
__global__ void __launch_bounds__(kBlock) k_system(Views w, SolveScalars s) {
  const int it = work_item(w, blockIdx.x);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  const int l = (blockIdx.x - w.start[it]) * kBlock + threadIdx.x;
  if (l >= v.N) return;
  const int W = v.W, N = v.N;
  const int i = l / W, j = l - i * W;
  PixDerivs d;
  for (int k = 0; k < kFeat; k++) {
    d.Ix[k] = gp(v.Ix)[(size_t)k * N + l];
    d.Iy[k] = gp(v.Iy)[(size_t)k * N + l];
    d.It[k] = gp(v.It)[(size_t)k * N + l];
  }
  d.Dx = gp(v.Dx)[l];
  d.Dy = gp(v.Dy)[l];
  d.Dt = gp(v.Dt)[l];
  const double pC = gp(v.phi)[l], pL = j > 0 ? gp(v.phi)[l - 1] : 0.0, pU = i > 0 ? gp(v.phi)[l - W] : 0.0;
  const float4 du = gload(v.du + l);
  assemble(v, s, l, i, j, (double)du.x, (double)du.y, (double)du.z, d, (double)gp(v.q)[l], pC, pL, pU);
}

// The first inner iteration of an outer iteration in one pass: derivatives,
// occlusion weight, du = 0, phi at the pixel and at its left and upper
// neighbours (recomputed from u: du is zero), and the system -- the values
// k_derivs, k_phi and k_system compute, without storing the intermediates.
// Each block owns a kOtW x kOtH tile: the stencil inputs 0.4*I1 + 0.6*Iw of
// the tile and a 2-pixel border (clamped to the image, as the stencils clamp)
// and phi of the tile and its upper row / left column are staged in LDS once.
// store: also store derivatives, occlusion weights and phi for later inner
// iterations.  Zeroes the item's check slots of this outer iteration.
constexpr int kRw = kOtW + 4, kRh = kOtH + 4;  // stencil region
constexpr int kPw = kOtW + 1, kPh = kOtH + 1;  // phi region

__global__ void __launch_bounds__(kBlock) k_outer_first(Views w, int n_slots, SolveScalars s, int store) {
  __shared__ double Ms[kFeat + 1][kRh * kRw];  // 0.4*I1 + 0.6*Iw (5 features, depth)
  __shared__ double Ps[kPh * kPw];            // phi
  const int it = work_item(w, blockIdx.x);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  const int lb = blockIdx.x - w.start[it];
  {
    const int stride = (w.start[it + 1] - w.start[it]) * kBlock;
    for (int t = lb * kBlock + threadIdx.x; t < n_slots; t += stride) v.ctl->maxdiff[t] = 0u;
  }
  const int W = v.W, H = v.H, N = v.N;
  const int tx = (W + kOtW - 1) / kOtW;
  const int x0 = (lb % tx) * kOtW, y0 = (lb / tx) * kOtH;
  for (int t = threadIdx.x; t < kRh * kRw; t += kBlock) {
    const int ry = t / kRw, rx = t - ry * kRw;
    const int o = clampi(y0 - 2 + ry, 0, H - 1) * W + clampi(x0 - 2 + rx, 0, W - 1);
    for (int k = 0; k < kFeat; k++) Ms[k][t] = .4 * gp(v.I1)[(size_t)k * N + o] + .6 * gp(v.Iw)[(size_t)k * N + o];
    Ms[kFeat][t] = .4 * gp(v.D1)[o] + .6 * gp(v.Dw)[o];
  }
  for (int t = threadIdx.x; t < kPh * kPw; t += kBlock) {
    const int py = t / kPw, px = t - py * kPw;
    const int i = y0 - 1 + py, j = x0 - 1 + px;
    Ps[t] = (i >= 0 && j >= 0 && i < H && j < W) ? phi_at<true>(v, i * W + j, i, j) : 0.0;
  }
  __syncthreads();
  const int ty = threadIdx.x / kOtW, txx = threadIdx.x - ty * kOtW;
  const int i = y0 + ty, j = x0 + txx;
  if (i >= H || j >= W) return;
  const int l = i * W + j;
  // derivs_at, with the stencil inputs from LDS (same values, same order)
  PixDerivs d;
  const double g[5] = {1, -8, 0, 8, -1};
  const int c0 = (ty + 2) * kRw + txx + 2;
  for (int k = 0; k < kFeat; k++) {
    double sx = 0, sy = 0;
    for (int m = 0; m < 5; m++) sx += g[m] * Ms[k][c0 - 2 + m];
    for (int m = 0; m < 5; m++) sy += g[m] * Ms[k][c0 + (m - 2) * kRw];
    d.Ix[k] = div12(sx);
    d.Iy[k] = div12(sy);
    d.It[k] = gp(v.Iw)[(size_t)k * N + l] - gp(v.I1)[(size_t)k * N + l];
  }
  {
    const double mg = .04;
    double sx = 0, sy = 0;
    for (int m = 0; m < 5; m++) sx += g[m] * Ms[kFeat][c0 - 2 + m];
    for (int m = 0; m < 5; m++) sy += g[m] * Ms[kFeat][c0 + (m - 2) * kRw];
    d.Dx = fmin(fmax(div12(sx), -mg), mg);
    d.Dy = fmin(fmax(div12(sy), -mg), mg);
    d.Dt = fmin(fmax(gp(v.Dw)[l] - gp(v.D1)[l], -mg), mg);
  }
  const float q = occlusion_at(v, l);
  gstore(v.du + l, make_float4(0.f, 0.f, 0.f, 0.f));
  const int p0 = (ty + 1) * kPw + txx + 1;
  const double pC = Ps[p0], pL = Ps[p0 - 1], pU = Ps[p0 - kPw];
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
  assemble(v, s, l, i, j, 0.0, 0.0, 0.0, d, (double)q, pC, pL, pU);
}

// .cu:382-470: one colour of the red-black relaxation, one thread per pixel of
// the colour, neighbours read from global memory (the plain schedule: one
// launch per colour per iteration, as the reference launches it).
__global__ void __launch_bounds__(256) k_relax_plain(Views w, int colour, int slot_base, int n_prev_checks,
                                                     int check_slot) {
  const int it = work_item(w, blockIdx.x);
  const LevelView& v = w.items[it];
  if (v.ctl->outer_done) return;
  if (relax_stopped(v.ctl, slot_base, n_prev_checks)) return;
  const int W = v.W, H = v.H, N = v.N;
  const int lb = blockIdx.x - w.start[it];
  const int gx = ((W + 1) / 2 + 63) / 64;
  const int bx = lb % gx, by = lb / gx;
  const int xc = bx * 64 + (threadIdx.x & 63);
  const int i = by * 4 + (threadIdx.x >> 6);
  const int j = 2 * xc + ((i + colour) & 1);
  float change = 0.f;
  if (i < H && j < W) {
    const int l = i * W + j;
    const bool hl = j > 0, hu = i > 0, hr = j < W - 1, hd = i < H - 1;
    const PixCoef c = load_coef(v.coef, v.storage, W, l, hl, hu);
    float4* x = v.du;
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 o = gload(x + l), L = hl ? gload(x + l - 1) : zero, U = hu ? gload(x + l - W) : zero,
                 R = hr ? gload(x + l + 1) : zero, D = hd ? gload(x + l + W) : zero;
    const float xo[3] = {o.x, o.y, o.z}, xl[3] = {L.x, L.y, L.z}, xu[3] = {U.x, U.y, U.z}, xr[3] = {R.x, R.y, R.z},
                xd[3] = {D.x, D.y, D.z};
    float xn[3];
    relax_update(c, v.omega, v.omega_c, hl, hu, hr, hd, xl, xu, xr, xd, xo, xn);
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
__global__ void __launch_bounds__(kBlock) k_update_warp(Views wl, int want_stat) {
  __shared__ double red[kBlock];
  const int item = work_item(wl, blockIdx.x);
  const LevelView& v = wl.items[item];
  if (v.ctl->outer_done) return;
  const int lb = blockIdx.x - wl.start[item];
  const int l = lb * kBlock + threadIdx.x;
  const int W = v.W, H = v.H, N = v.N;
  const bool stat = want_stat && v.allow_early;
  double term = 0;
  if (l < N) {
    const int i = l / W, j = l - i * W;
    const float4 e = gload(v.du + l);
    const float e0 = e.x, e1 = e.y, e2 = e.z;
    const double u0 = gp(v.u)[l] + (double)e0;
    const double u1 = gp(v.u)[N + l] + (double)e1;
    const double u2 = gp(v.u)[2 * N + l] + (double)e2;
    gp(v.u)[l] = u0;
    gp(v.u)[N + l] = u1;
    gp(v.u)[2 * N + l] = u2;
    const double x = j + u0, y = i + u1;
    if (x < 0 || x > W - 1 || y < 0 || y > H - 1) {
      for (int k = 0; k < kFeat; k++) gp(v.Iw)[(size_t)k * N + l] = gp(v.I1)[(size_t)k * N + l];
      gp(v.Dw)[l] = gp(v.D1)[l] - u2;
    } else {
      const int x0 = clampi((int)floor(x), 0, W - 1), y0 = clampi((int)floor(y), 0, H - 1);
      const int x1 = min(W - 1, x0 + 1), y1 = min(H - 1, y0 + 1);
      const float ax = x - x0, ay = y - y0;
      const float w00 = (1 - ax) * (1 - ay), w01 = (1 - ax) * ay, w10 = ax * (1 - ay), w11 = ax * ay;
      const int o00 = y0 * W + x0, o01 = y1 * W + x0, o10 = y0 * W + x1, o11 = y1 * W + x1;
      for (int k = 0; k < kFeat; k++) {
        const RGBDFLOW_GLOBAL double* b = gp(v.I2) + (size_t)k * N;
        gp(v.Iw)[(size_t)k * N + l] = w00 * b[o00] + w01 * b[o01] + w10 * b[o10] + w11 * b[o11];
      }
      gp(v.Dw)[l] = w00 * gp(v.D2)[o00] + w01 * gp(v.D2)[o01] + w10 * gp(v.D2)[o10] + w11 * gp(v.D2)[o11] - u2;
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
void launch_derivs(hipStream_t s, const Views& v, int blocks, int n_slots) {
  hipLaunchKernelGGL(k_derivs, dim3(blocks), dim3(kBlock), 0, s, v, n_slots);
}
void launch_phi(hipStream_t s, const Views& v, int blocks) {
  hipLaunchKernelGGL(k_phi, dim3(blocks), dim3(kBlock), 0, s, v);
}
void launch_system(hipStream_t s, const Views& v, int blocks, const SolveScalars& k) {
  hipLaunchKernelGGL(k_system, dim3(blocks), dim3(kBlock), 0, s, v, k);
}
void launch_outer_first(hipStream_t s, const Views& v, int blocks, int n_slots, const SolveScalars& k, int store) {
  hipLaunchKernelGGL(k_outer_first, dim3(blocks), dim3(kBlock), 0, s, v, n_slots, k, store);
}
void launch_relax_plain(hipStream_t s, const Views& v, int blocks, int colour, int slot_base, int n_prev_checks,
                        int check_slot) {
  hipLaunchKernelGGL(k_relax_plain, dim3(blocks), dim3(256), 0, s, v, colour, slot_base, n_prev_checks, check_slot);
}
void launch_update_warp(hipStream_t s, const Views& v, int blocks, int want_stat) {
  hipLaunchKernelGGL(k_update_warp, dim3(blocks), dim3(kBlock), 0, s, v, want_stat);
}
void launch_stat_final(hipStream_t s, const Views& v, int outer, int decide) {
  hipLaunchKernelGGL(k_stat_final, dim3(v.n), dim3(kBlock), 0, s, v, outer, decide);
}

}  // namespace rgbdflow
