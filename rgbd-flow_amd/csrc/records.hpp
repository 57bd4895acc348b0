// records.hpp -- storage of the per-pixel system records.
//
// F32 (the reference's precision): four float4 per pixel,
//   (eRxy, eRz, eDxy, eDz) (Buu, Bvv, Bww, Buv) (Buw, Bvw, ru, rv) (rw, 1/Buu, 1/Bvv, 1/Bww).
// F16 / BF16 (precision variants): sixteen 16-bit values per pixel (32 bytes),
//   the 13 coefficients in the same order, then for F16 a per-pixel power-of-
//   two exponent e (values are stored as c * 2^-e so the largest is in
//   [2^14, 2^15), inside FP16's range); the reciprocals are recomputed from the
//   decoded diagonal when a record is loaded.  Arithmetic stays f32.
// Layout: the 16-byte parts of the records are planes of N (structure of
// arrays): part q of pixel l is at base[q * N + l] (float4 units for F32,
// uint4 units for F16 / BF16), so the lanes of a wave that load part q of
// neighbouring pixels read contiguous memory.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace rgbdflow {

enum Storage { kStoreF32 = 0, kStoreF16 = 1, kStoreBF16 = 2 };

__device__ __forceinline__ unsigned short f32_to_bf16(float x) {
  const unsigned u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // NaN stays NaN
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);                  // round to nearest even
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ unsigned short f32_to_f16_bits(float x) {
  return __half_as_ushort(__float2half_rn(x));
}
__device__ __forceinline__ float f16_bits_to_f32(unsigned short h) { return __half2float(__ushort_as_half(h)); }

// Feature images ([kFeat][N] planes of I1, I2, Iw) in their storage S: f64,
// or 16 bits (FP16 / BF16 of the value rounded to f32) in the first quarter
// of the same buffer.  Loads return f64.  S is a compile-time constant, so
// the f64 kernels are the reference's loads and stores.
template <int S>
__device__ __forceinline__ double img_ld(const double* base, size_t i) {
  if (S == kImgF64) return gp(base)[i];
  const unsigned short h = gp(reinterpret_cast<const unsigned short*>(base))[i];
  return S == kImgF16 ? (double)f16_bits_to_f32(h) : (double)bf16_to_f32(h);
}
template <int S>
__device__ __forceinline__ void img_st(double* base, size_t i, double x) {
  if (S == kImgF64) {
    gp(base)[i] = x;
    return;
  }
  const float f = (float)x;
  gp(reinterpret_cast<unsigned short*>(base))[i] = S == kImgF16 ? f32_to_f16_bits(f) : f32_to_bf16(f);
}

// Power-of-two scale 2^k as a float (k within the normal range).
__device__ __forceinline__ float pow2f(int k) { return __uint_as_float((unsigned)(k + 127) << 23); }

// Writes pixel l's record (N: pixels of the level); c: the 13 values in Coef order.
__device__ __forceinline__ void store_record(float4* base, int storage, size_t N, size_t l,
                                             const float (&c)[kCoef]) {
  if (storage == kStoreF32) {
    gstore(base + l, make_float4(c[0], c[1], c[2], c[3]));
    gstore(base + N + l, make_float4(c[4], c[5], c[6], c[7]));
    gstore(base + 2 * N + l, make_float4(c[8], c[9], c[10], c[11]));
    gstore(base + 3 * N + l, make_float4(c[12], 1.0f / c[kBuu], 1.0f / c[kBvv], 1.0f / c[kBww]));
    return;
  }
  unsigned short h[16];
  if (storage == kStoreF16) {
    float m = 0.f;
    for (int k = 0; k < kCoef; k++) m = fmaxf(m, fabsf(c[k]));
    int e = 0;
    if (m > 0.f && m <= 3.4e38f) e = min(max(ilogbf(m) - 14, -100), 100);
    const float s = pow2f(-e);
    for (int k = 0; k < kCoef; k++) h[k] = f32_to_f16_bits(c[k] * s);
    h[13] = (unsigned short)(short)e;
  } else {
    for (int k = 0; k < kCoef; k++) h[k] = f32_to_bf16(c[k]);
    h[13] = 0;
  }
  h[14] = h[15] = 0;
  uint4* R = reinterpret_cast<uint4*>(base);
  gstore(R + l, make_uint4(h[0] | (unsigned)h[1] << 16, h[2] | (unsigned)h[3] << 16, h[4] | (unsigned)h[5] << 16,
                           h[6] | (unsigned)h[7] << 16));
  gstore(R + N + l, make_uint4(h[8] | (unsigned)h[9] << 16, h[10] | (unsigned)h[11] << 16,
                               h[12] | (unsigned)h[13] << 16, h[14] | (unsigned)h[15] << 16));
}

// 1.0f / c for a c of at most 11 significant bits (a 16-bit record's
// decoded value): v_rcp_f32 (within one ulp of 1/c) refined by one Newton
// step with FMAs.  1/c lies at least 2^-12 ulp from a rounding midpoint
// (its significand ratio is 2^k / H, H an integer below 2^11), farther than
// the refined value's error (under 2^-21 ulp), so the result is the IEEE
// quotient; checked for every such significand and every start within one
// ulp (oracle/divcheck.cpp).  It needs c and 1/c normal (exponent field 1 ..
// 253); a wave with a lane outside that (zero, huge, infinite or NaN
// diagonals: rare) divides again, the IEEE way, for those lanes.
__device__ __forceinline__ float recip_short(float c) {
  const float y0 = __builtin_amdgcn_rcpf(c);
  float y = __builtin_fmaf(__builtin_fmaf(-c, y0, 1.0f), y0, y0);
  const unsigned ex = (__float_as_uint(c) >> 23) & 255u;
  const bool ok = ex - 1u <= 252u;
  if (__builtin_expect(__any(!ok), 0)) y = ok ? y : 1.0f / c;
  return y;
}

// A record as it comes from memory, before decoding: the four f32 part
// planes, or the two 16-byte parts of a 16-bit record.  Fetching and decoding
// are separate steps so that a kernel can issue the loads well before it
// needs the values.
template <int ST>
struct RawRecord {
  uint4 q[2];
};
template <>
struct RawRecord<kStoreF32> {
  float4 q[4];
};

template <int ST>
__device__ __forceinline__ void record_fetch(RawRecord<ST>& r, const float4* base, size_t N, size_t l) {
  if constexpr (ST == kStoreF32) {
#pragma unroll
    for (int i = 0; i < 4; i++) r.q[i] = gload(base + i * N + l);
  } else {
    const uint4* R = reinterpret_cast<const uint4*>(base);
    r.q[0] = gload(R + l);
    r.q[1] = gload(R + N + l);
  }
}

// The 13 values and the reciprocals of the diagonal of a fetched record.
template <int ST>
__device__ __forceinline__ void record_decode(const RawRecord<ST>& r, float (&c)[kCoef], float (&y)[3]) {
  if constexpr (ST == kStoreF32) {
    const float4 r0 = r.q[0], r1 = r.q[1], r2 = r.q[2], r3 = r.q[3];
    c[0] = r0.x; c[1] = r0.y; c[2] = r0.z; c[3] = r0.w;
    c[4] = r1.x; c[5] = r1.y; c[6] = r1.z; c[7] = r1.w;
    c[8] = r2.x; c[9] = r2.y; c[10] = r2.z; c[11] = r2.w;
    c[12] = r3.x;
    y[0] = r3.y; y[1] = r3.z; y[2] = r3.w;
  } else {
    const unsigned w[8] = {r.q[0].x, r.q[0].y, r.q[0].z, r.q[0].w, r.q[1].x, r.q[1].y, r.q[1].z, r.q[1].w};
    if constexpr (ST == kStoreF16) {
      const int e = (short)(unsigned short)(w[6] >> 16);
      const float s = pow2f(e);
#pragma unroll
      for (int k = 0; k < kCoef; k++) c[k] = f16_bits_to_f32((unsigned short)(w[k >> 1] >> ((k & 1) * 16))) * s;
    } else {
#pragma unroll
      for (int k = 0; k < kCoef; k++) c[k] = bf16_to_f32((unsigned short)(w[k >> 1] >> ((k & 1) * 16)));
    }
    y[0] = recip_short(c[kBuu]);
    y[1] = recip_short(c[kBvv]);
    y[2] = recip_short(c[kBww]);
  }
}

// Reads pixel l's record (N: pixels of the level): the 13 values and the
// reciprocals of the diagonal.
__device__ __forceinline__ void load_record(const float4* base, int storage, size_t N, size_t l, float (&c)[kCoef],
                                            float (&y)[3]) {
  if (storage == kStoreF32) {
    RawRecord<kStoreF32> r;
    record_fetch<kStoreF32>(r, base, N, l);
    record_decode<kStoreF32>(r, c, y);
  } else if (storage == kStoreF16) {
    RawRecord<kStoreF16> r;
    record_fetch<kStoreF16>(r, base, N, l);
    record_decode<kStoreF16>(r, c, y);
  } else {
    RawRecord<kStoreBF16> r;
    record_fetch<kStoreBF16>(r, base, N, l);
    record_decode<kStoreBF16>(r, c, y);
  }
}

}  // namespace rgbdflow
