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
#include "records.hpp"

namespace rgbdflow {

// Coefficients of one pixel in the order the update consumes them.
struct PixCoef {
  float lxy, lz, uxy, uz, rxy, rz, dxy, dz;  // edge weights to left/up/right/down (xy rows, z row)
  float Buu, Bvv, Bww, Buv, Buw, Bvw;        // symmetric 3x3 block
  float ru, rv, rw;                          // rhs
};

// A pixel's coefficients from the system records; the left and up edge
// weights are the left neighbour's right and the upper neighbour's down ones
// (zero where the neighbour does not exist).
__device__ __forceinline__ PixCoef load_coef(const float4* C, int storage, int W, int N, int l, bool hl, bool hu) {
  PixCoef c;
  float r[kCoef], y[3];
  load_record(C, storage, (size_t)N, (size_t)l, r, y);
  float L[kCoef] = {}, U[kCoef] = {};
  if (hl) load_record(C, storage, (size_t)N, (size_t)(l - 1), L, y);
  if (hu) load_record(C, storage, (size_t)N, (size_t)(l - W), U, y);
  c.lxy = hl ? L[kERxy] : 0.f;
  c.lz = hl ? L[kERz] : 0.f;
  c.uxy = hu ? U[kEDxy] : 0.f;
  c.uz = hu ? U[kEDz] : 0.f;
  c.rxy = r[kERxy];
  c.rz = r[kERz];
  c.dxy = r[kEDxy];
  c.dz = r[kEDz];
  c.Buu = r[kBuu];
  c.Bvv = r[kBvv];
  c.Bww = r[kBww];
  c.Buv = r[kBuv];
  c.Buw = r[kBuw];
  c.Bvw = r[kBvw];
  c.ru = r[kRu];
  c.rv = r[kRv];
  c.rw = r[kRw];
  return c;
}

// The reference's step limit (.cu:463-467): x clamped to xo +- 0.3, as max
// then min (two instructions): the reference's two comparisons for every x
// but a NaN -- the reference's CHECK_FLOATS error -- which this clamps into
// the band (max returns the non-NaN operand) instead of passing on.
__device__ __forceinline__ float clamp_step(float r, float xo) {
  const float lo = xo - .3f, hi = xo + .3f;
  return fminf(fmaxf(r, lo), hi);
}

// One pixel's update.  h* say which neighbours exist in the image; x* hold the
// neighbours' (other colour) values, xo the pixel's old values.
// omega_c = 1 - omega in float.  gs: the pixel's other unknowns enter with
// the values this update already gave them (u, then v, then w: point SOR
// order) instead of their old values (.cu:431-459).
__device__ __forceinline__ void relax_update(const PixCoef& c, float omega, float omega_c, bool hl, bool hu, bool hr,
                                             bool hd, const float xl[3], const float xu[3], const float xr[3],
                                             const float xd[3], const float xo[3], float xn[3], bool gs = false) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float x0 = (gs && k > 0) ? xn[0] : xo[0], x1 = (gs && k > 1) ? xn[1] : xo[1];
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
      nv -= c.Buv * x0;
      nv -= c.Bvw * xo[2];
      nv /= c.Bvv;
    } else {
      nv += c.rw;
      nv -= c.Buw * x0;
      nv -= c.Bvw * x1;
      nv /= c.Bww;
    }
    xn[k] = clamp_step(omega * nv + omega_c * xo[k], xo[k]);
  }
}

// The inverse of a pixel's symmetric 3x3 block from its adjugate, in f32 and
// this order: the cofactors, det = Buu c_uu + Buv c_uv + Buw c_uw, 1 / det
// (IEEE), then each cofactor times it.  inv: (uu, vv, ww, uv, uw, vw).
// Returns whether the stored block is negative definite with a finite
// inverse (Buu < 0, c_ww > 0, det < 0, 1 / det finite): the assembled block
// is, but rounding a nearly singular one can break it.
__device__ __forceinline__ bool block_inverse(float Buu, float Bvv, float Bww, float Buv, float Buw, float Bvw,
                                              float (&inv)[6]) {
  const float cuu = Bvv * Bww - Bvw * Bvw, cvv = Buu * Bww - Buw * Buw, cww = Buu * Bvv - Buv * Buv;
  const float cuv = Buw * Bvw - Buv * Bww, cuw = Buv * Bvw - Buw * Bvv, cvw = Buv * Buw - Buu * Bvw;
  const float det = Buu * cuu + Buv * cuv + Buw * cuw;
  const float id = 1.0f / det;
  inv[0] = cuu * id;
  inv[1] = cvv * id;
  inv[2] = cww * id;
  inv[3] = cuv * id;
  inv[4] = cuw * id;
  inv[5] = cvw * id;
  return (Buu < 0.f) & (cww > 0.f) & (det < 0.f) & (id >= -0x1.fffffep127f);
}

// The block update's last half: x^ = B^-1 a (ok: block_inverse's test; a
// pixel whose block fails it keeps x^ = its old values), relaxed and clamped.
__device__ __forceinline__ void block_step(const float (&inv)[6], const float (&a)[3], bool ok, float omega,
                                           float omega_c, const float xo[3], float xn[3]) {
  float xh[3] = {inv[0] * a[0] + inv[3] * a[1] + inv[4] * a[2],
                 inv[3] * a[0] + inv[1] * a[1] + inv[5] * a[2],
                 inv[4] * a[0] + inv[5] * a[1] + inv[2] * a[2]};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    xh[k] = ok ? xh[k] : xo[k];
    xn[k] = clamp_step(omega * xh[k] + omega_c * xo[k], xo[k]);
  }
}

// Block SOR (not in the reference): the pixel's three unknowns solved together
// from its 3x3 block, x^ = B^-1 (rhs - sum_n A_n x_n), then over-relaxed and
// clamped as the reference's update.  Convergent for any factor in (0, 2) on a
// symmetric positive definite system.
__device__ __forceinline__ void relax_update_block(const PixCoef& c, float omega, float omega_c, bool hl, bool hu,
                                                   bool hr, bool hd, const float xl[3], const float xu[3],
                                                   const float xr[3], const float xd[3], const float xo[3],
                                                   float xn[3]) {
  float a[3];
  const float rhs[3] = {c.ru, c.rv, c.rw};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float nv = 0.f;
    if (hl) nv -= (k < 2 ? c.lxy : c.lz) * xl[k];
    if (hu) nv -= (k < 2 ? c.uxy : c.uz) * xu[k];
    if (hr) nv -= (k < 2 ? c.rxy : c.rz) * xr[k];
    if (hd) nv -= (k < 2 ? c.dxy : c.dz) * xd[k];
    a[k] = nv + rhs[k];
  }
  float inv[6];
  const bool ok = block_inverse(c.Buu, c.Bvv, c.Bww, c.Buv, c.Buw, c.Bvw, inv);
  block_step(inv, a, ok, omega, omega_c, xo, xn);
}

// relax_update_block with every neighbour term evaluated (as
// relax_update_dense; the same bits) and the block's inverse and test given.
__device__ __forceinline__ void relax_update_dense_block(const PixCoef& c, const float (&inv)[6], bool ok,
                                                         float omega, float omega_c, const float4& L, const float4& U,
                                                         const float4& R, const float4& D, const float4& o,
                                                         float xn[3]) {
  const float xl[3] = {L.x, L.y, L.z}, xu[3] = {U.x, U.y, U.z}, xr[3] = {R.x, R.y, R.z}, xd[3] = {D.x, D.y, D.z};
  const float xo[3] = {o.x, o.y, o.z};
  const float rhs[3] = {c.ru, c.rv, c.rw};
  float a[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float nv = 0.f;
    nv -= (k < 2 ? c.lxy : c.lz) * xl[k];
    nv -= (k < 2 ? c.uxy : c.uz) * xu[k];
    nv -= (k < 2 ? c.rxy : c.rz) * xr[k];
    nv -= (k < 2 ? c.dxy : c.dz) * xd[k];
    a[k] = nv + rhs[k];
  }
  block_step(inv, a, ok, omega, omega_c, xo, xn);
}

// Markstein's division (IBM J. Res. Dev. 34(1), 1990): with y = RN(1/b),
// q0 = RN(a y), r = a - b q0 (exact by FMA), RN(q0 + r y) = RN(a/b) whenever
// no intermediate underflows or overflows, which holds for |b| in
// [2^-60, 2^60] and a = 0 or |a| in [2^-60, 2^60].
// (on the bit pattern: positive floats order like their bits, so |a| in
// [2^-60, 2^60] is one unsigned range test; NaN and infinity fall outside it)
__device__ __forceinline__ bool recip_div_ok(float a) {
  const unsigned ua = __float_as_uint(a) & 0x7fffffffu;
  return (ua == 0u) | (ua - 0x21800000u <= 0x5d800000u - 0x21800000u);
}
__device__ __forceinline__ bool recip_den_ok(float b) {
  const float ab = fabsf(b);
  return ab >= 0x1p-60f && ab <= 0x1p60f;
}
// recip_div_ok of three numerators at once, with the diagonals' test: a zero
// numerator is mapped into the range, then the smallest and largest bit
// patterns are tested (vector arithmetic; one lane condition in the end).
__device__ __forceinline__ bool recip_div3_ok(const float (&a)[3], bool den_ok) {
  unsigned lo = 0xffffffffu, hi = 0u;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const unsigned ua = __float_as_uint(a[k]) & 0x7fffffffu;
    const unsigned z = ua == 0u ? 0x21800000u : ua;
    lo = min(lo, z);
    hi = max(hi, z);
  }
  return den_ok & (lo >= 0x21800000u) & (hi <= 0x5d800000u);
}
__device__ __forceinline__ float recip_div(float a, float b, float y) {
  const float q0 = a * y;
  const float r = __builtin_fmaf(-b, q0, a);
  return __builtin_fmaf(r, y, q0);
}

// relax_update with every neighbour term evaluated: a missing neighbour
// carries a zero coefficient and a zero value, so its term subtracts +0 from
// a sum that is never -0 (it starts at +0 and only exact cancellation makes
// zero, which rounds to +0) and leaves it unchanged; the result is the same
// bits as relax_update's.  L, U, R, D, o: neighbour and old values (x, y, z
// in .x .y .z).  y3: 1/Buu, 1/Bvv, 1/Bww (IEEE divisions); den_ok: all three
// diagonals are in the reciprocal division's domain.
__device__ __forceinline__ void relax_update_dense(const PixCoef& c, float omega, float omega_c,
                                                   const float (&y3)[3], bool den_ok,
                                                   const float4& L, const float4& U, const float4& R,
                                                   const float4& D, const float4& o, float xn[3]) {
  float nv[3];
  const float xl[3] = {L.x, L.y, L.z}, xu[3] = {U.x, U.y, U.z}, xr[3] = {R.x, R.y, R.z}, xd[3] = {D.x, D.y, D.z};
  const float xo[3] = {o.x, o.y, o.z};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float el = k < 2 ? c.lxy : c.lz;
    const float eu = k < 2 ? c.uxy : c.uz;
    const float er = k < 2 ? c.rxy : c.rz;
    const float ed = k < 2 ? c.dxy : c.dz;
    float a = 0.f;
    a -= el * xl[k];
    a -= eu * xu[k];
    a -= er * xr[k];
    a -= ed * xd[k];
    if (k == 0) {
      a += c.ru;
      a -= c.Buv * xo[1];
      a -= c.Buw * xo[2];
    } else if (k == 1) {
      a += c.rv;
      a -= c.Buv * xo[0];
      a -= c.Bvw * xo[2];
    } else {
      a += c.rw;
      a -= c.Buw * xo[0];
      a -= c.Bvw * xo[1];
    }
    nv[k] = a;
  }
  // every lane divides by the reciprocal; only a wave with a lane outside
  // its domain (rare) divides again, the IEEE way, for those lanes
  const float den[3] = {c.Buu, c.Bvv, c.Bww};
  float q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) q[k] = recip_div(nv[k], den[k], y3[k]);
  const bool ok = recip_div3_ok(nv, den_ok);
  if (__builtin_expect(__any(!ok), 0)) {
#pragma unroll
    for (int k = 0; k < 3; k++) q[k] = ok ? q[k] : nv[k] / den[k];
  }
#pragma unroll
  for (int k = 0; k < 3; k++) xn[k] = clamp_step(omega * q[k] + omega_c * xo[k], xo[k]);
}

// relax_update_dense in the point SOR order (relax_update with gs): each
// unknown is divided, relaxed and clamped before the next one reads it.  The
// reciprocal division is chosen per unknown; in its domain it equals the IEEE
// quotient, so the choice does not change the value.
__device__ __forceinline__ void relax_update_dense_gs(const PixCoef& c, float omega, float omega_c,
                                                      const float (&y3)[3], bool den_ok, const float4& L,
                                                      const float4& U, const float4& R, const float4& D,
                                                      const float4& o, float xn[3]) {
  const float xl[3] = {L.x, L.y, L.z}, xu[3] = {U.x, U.y, U.z}, xr[3] = {R.x, R.y, R.z}, xd[3] = {D.x, D.y, D.z};
  const float xo[3] = {o.x, o.y, o.z};
  const float den[3] = {c.Buu, c.Bvv, c.Bww};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float el = k < 2 ? c.lxy : c.lz;
    const float eu = k < 2 ? c.uxy : c.uz;
    const float er = k < 2 ? c.rxy : c.rz;
    const float ed = k < 2 ? c.dxy : c.dz;
    float a = 0.f;
    a -= el * xl[k];
    a -= eu * xu[k];
    a -= er * xr[k];
    a -= ed * xd[k];
    if (k == 0) {
      a += c.ru;
      a -= c.Buv * xo[1];
      a -= c.Buw * xo[2];
    } else if (k == 1) {
      a += c.rv;
      a -= c.Buv * xn[0];
      a -= c.Bvw * xo[2];
    } else {
      a += c.rw;
      a -= c.Buw * xn[0];
      a -= c.Bvw * xn[1];
    }
    float q = recip_div(a, den[k], y3[k]);
    const bool ok = den_ok & recip_div_ok(a);
    if (__builtin_expect(__any(!ok), 0)) q = ok ? q : a / den[k];
    xn[k] = clamp_step(omega * q + omega_c * xo[k], xo[k]);
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
