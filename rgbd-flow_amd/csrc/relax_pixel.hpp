// relax_pixel.hpp -- the per-pixel update of the red-black relaxation, shared
// by every relaxation kernel so that all schedules compute identical values.
//
// src/sceneFlowSingleScaleJacobiCUDA.cu:404-468: for each unknown k of the
// pixel, newval = rhs_k - sum_n A_nk x_nk (left, up, right, down, in that
// order) - sum_{k' != k} B_kk' x_k'(old), divided by B_kk; then
// x_k = 0.7 newval + 0.3 x_k(old), clamped to x_k(old) +- 0.3.  All f32.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace rgbdflow {

// Coefficients of one pixel in the order the update consumes them.
struct PixCoef {
  float lxy, lz, uxy, uz, rxy, rz, dxy, dz;  // edge weights to left/up/right/down (xy rows, z row)
  float Buu, Bvv, Bww, Buv, Buw, Bvw;        // symmetric 3x3 block
  float ru, rv, rw;                          // rhs
};

__device__ __forceinline__ PixCoef load_coef(const float* C, int N, int W, int l, bool hl, bool hu) {
  PixCoef c;
  c.lxy = hl ? C[kERxy * (size_t)N + l - 1] : 0.f;
  c.lz = hl ? C[kERz * (size_t)N + l - 1] : 0.f;
  c.uxy = hu ? C[kEDxy * (size_t)N + l - W] : 0.f;
  c.uz = hu ? C[kEDz * (size_t)N + l - W] : 0.f;
  c.rxy = C[kERxy * (size_t)N + l];
  c.rz = C[kERz * (size_t)N + l];
  c.dxy = C[kEDxy * (size_t)N + l];
  c.dz = C[kEDz * (size_t)N + l];
  c.Buu = C[kBuu * (size_t)N + l];
  c.Bvv = C[kBvv * (size_t)N + l];
  c.Bww = C[kBww * (size_t)N + l];
  c.Buv = C[kBuv * (size_t)N + l];
  c.Buw = C[kBuw * (size_t)N + l];
  c.Bvw = C[kBvw * (size_t)N + l];
  c.ru = C[kRu * (size_t)N + l];
  c.rv = C[kRv * (size_t)N + l];
  c.rw = C[kRw * (size_t)N + l];
  return c;
}

// One pixel's update.  h* say which neighbours exist in the image; x* hold the
// neighbours' (other colour) values, xo the pixel's old values.
__device__ __forceinline__ void relax_update(const PixCoef& c, bool hl, bool hu, bool hr, bool hd, const float xl[3],
                                             const float xu[3], const float xr[3], const float xd[3],
                                             const float xo[3], float xn[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float el = k < 2 ? c.lxy : c.lz;
    const float eu = k < 2 ? c.uxy : c.uz;
    const float er = k < 2 ? c.rxy : c.rz;
    const float ed = k < 2 ? c.dxy : c.dz;
    float nv = 0.f;
    if (hl) nv -= el * xl[k];
    if (hu) nv -= eu * xu[k];
    if (hr) nv -= er * xr[k];
    if (hd) nv -= ed * xd[k];
    if (k == 0) {
      nv += c.ru;
      nv -= c.Buv * xo[1];
      nv -= c.Buw * xo[2];
      nv /= c.Buu;
    } else if (k == 1) {
      nv += c.rv;
      nv -= c.Buv * xo[0];
      nv -= c.Bvw * xo[2];
      nv /= c.Bvv;
    } else {
      nv += c.rw;
      nv -= c.Buw * xo[0];
      nv -= c.Bvw * xo[1];
      nv /= c.Bww;
    }
    const float omega = .7f;
    float r = omega * nv + (1 - omega) * xo[k];
    if (r < xo[k] - .3f) r = xo[k] - .3f;
    else if (r > xo[k] + .3f) r = xo[k] + .3f;
    xn[k] = r;
  }
}

// The relaxation has stopped for this (outer, inner) if an earlier check
// iteration saw max|change| < 1e-5 or >= 0.3 (.cu:914-922).
__device__ __forceinline__ bool relax_stopped(const SolveCtl* ctl, int slot_base, int n_prev_checks) {
  for (int c = 0; c < n_prev_checks; c++) {
    const float md = __uint_as_float(ctl->maxdiff[slot_base + c]);
    if ((double)md < 1e-5 || md >= .3f) return true;
  }
  return false;
}

__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace rgbdflow
