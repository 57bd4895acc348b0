// divcheck.cpp -- checker (test infrastructure) for the reciprocal division
// of the tiled relaxation kernel.
//
// The kernel replaces each IEEE single-precision division a / b of the
// relaxation update (src/sceneFlowSingleScaleJacobiCUDA.cu:440,449,458) by
//   y = RN(1 / b)            (an IEEE division, once per pixel and launch)
//   q0 = RN(a * y)
//   r  = fma(-b, q0, a)      (exact: the remainder of q0)
//   q  = fma(r, y, q0)
// which is RN(a / b) by Markstein's theorem (IBM J. Res. Dev. 34(1), 1990)
// when no intermediate underflows or overflows; the kernel takes this path
// only for |b| in [2^-60, 2^60] and a = 0 or |a| in [2^-60, 2^60] and uses the
// IEEE division otherwise.  This file samples (a, b) in that domain and counts
// the cases where the two differ in any bit.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

float make(uint32_t sign, int e, uint32_t mant) {
  const uint32_t bits = (sign << 31) | ((uint32_t)(e + 127) << 23) | (mant & 0x7FFFFFu);
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

bool in_domain(float a, float b) {
  const float lo = std::ldexp(1.0f, -60), hi = std::ldexp(1.0f, 60);
  const float aa = std::fabs(a), ab = std::fabs(b);
  return ab >= lo && ab <= hi && (a == 0.f || (aa >= lo && aa <= hi));
}

float fast_div(float a, float b, int recipe) {
  const float y = 1.0f / b;
  const float q0 = a * y;
  if (recipe == 1) return q0;  // negative control: no correction step
  const float r = std::fmaf(-b, q0, a);
  return std::fmaf(r, y, q0);
}

}  // namespace

extern "C" {

// mode 0: uniform significands and exponents in the domain; mode 1: b's
// significand near the ends of its binade; mode 2: a's significand near b's
// (quotients near 1 and near the ends of a binade); mode 3: small integer
// multiples (quotients exactly representable and halfway-adjacent).
// recipe 0 is the kernel's; recipe 1 drops the correction step (a negative
// control the check must catch).  Returns the number of mismatches; the first
// is stored in *ea, *eb.
uint64_t orc_check_reciprocal_division(uint64_t n, uint64_t seed, int mode, int recipe, float* ea, float* eb) {
  uint64_t s = seed * 0x2545F4914F6CDD1Dull + 1, bad = 0;
  for (uint64_t i = 0; i < n; i++) {
    const uint64_t r1 = splitmix(s), r2 = splitmix(s);
    const int eA = (int)(r1 % 121) - 60, eB = (int)((r1 >> 8) % 121) - 60;
    uint32_t mA = (uint32_t)(r2 & 0x7FFFFF), mB = (uint32_t)((r2 >> 23) & 0x7FFFFF);
    if (mode == 1) mB = (r1 >> 20) & 1 ? (uint32_t)((r2 >> 23) & 0xFF) : 0x7FFFFFu - (uint32_t)((r2 >> 23) & 0xFF);
    if (mode == 2) mA = (mB + (uint32_t)((r2 >> 50) & 0x3FF) - 512u) & 0x7FFFFFu;
    float a = make((uint32_t)(r1 >> 40) & 1, eA, mA), b = make((uint32_t)(r1 >> 41) & 1, eB, mB);
    if (mode == 3) a = b * (float)((int)((r2 >> 40) % 2001) - 1000);
    if ((r1 >> 42) % 64 == 0) a = 0.f;
    if (!in_domain(a, b)) continue;
    const float q = fast_div(a, b, recipe), ref = a / b;
    uint32_t x, y;
    std::memcpy(&x, &q, 4);
    std::memcpy(&y, &ref, 4);
    if (x != y) {
      if (bad == 0 && ea && eb) {
        *ea = a;
        *eb = b;
      }
      bad++;
    }
  }
  return bad;
}

}  // extern "C"

extern "C" {

// The kernels divide f64 values by small integer constants d: the derivative
// stencils' sums by 12 (src/sceneFlowSingleScaleJacobiCUDA.cu:84-147), the
// channel means of the data-term products by 5 (.cu:211) and the sums of the
// Laplacians' existing regularisation weights by their count 3 (.cu:274-282).
// They take q0 = RN(x * RN(1/d)), r = fma(-d, q0, x), q = fma(r, RN(1/d), q0)
// -- Markstein's correction with the constant's correctly rounded reciprocal
// -- for x != 0 with |x| in [2^-960, 2^960], and the IEEE division otherwise.
// Samples x in that domain (random bit patterns, and multiples of d, whose
// quotients are exact) and counts the cases that differ from x / d in any bit
// (recipe 1: the uncorrected product, a negative control).
uint64_t orc_check_div_const(uint64_t n, uint64_t seed, int d, int recipe, double* ex) {
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 7, bad = 0;
  const double dd = (double)d;
  const double y = 1.0 / dd;
  for (uint64_t i = 0; i < n; i++) {
    const uint64_t r1 = splitmix(s), r2 = splitmix(s);
    double x;
    if ((r1 & 3) == 0) {
      x = (double)(int64_t)(r2 % 2000001 - 1000000) * dd * std::ldexp(1.0, (int)((r1 >> 8) % 200) - 100);
    } else {
      const int e = (int)((r1 >> 8) % 1921) - 960;
      const uint64_t bits = ((uint64_t)((r1 >> 40) & 1) << 63) | ((uint64_t)(e + 1023) << 52) | (r2 & 0xFFFFFFFFFFFFFull);
      std::memcpy(&x, &bits, 8);
    }
    if (x == 0.0 || std::fabs(x) < std::ldexp(1.0, -960) || std::fabs(x) > std::ldexp(1.0, 960)) continue;
    const double q0 = x * y;
    const double q = recipe == 1 ? q0 : std::fma(std::fma(-dd, q0, x), y, q0);
    const double ref = x / dd;
    if (std::memcmp(&q, &ref, 8) != 0) {
      if (bad == 0 && ex) *ex = x;
      bad++;
    }
  }
  return bad;
}

}  // extern "C"

extern "C" {

// The f32 stage of the 16-bit image variants divides f32 values by the same
// constants: q0 = RN(x * RN(1/d)), r = fma(-d, q0, x), q = fma(r, RN(1/d), q0)
// for |x| in [2^-100, 2^100], the IEEE division otherwise.  In that domain the
// recipe commutes with scaling x by a power of two (q0, r and q scale with it
// and stay normal or exact), so every binade behaves as [1, 2): this checks
// each of its 2^23 significands with both signs, bit for bit against x / d
// (recipe 1: the uncorrected product, a negative control).  Returns the
// mismatches; the first is stored in *ex.
uint64_t orc_check_div_const_f32(int d, int recipe, float* ex) {
  const float df = (float)d;
  const float y = 1.0f / df;
  uint64_t bad = 0;
  for (uint32_t m = 0; m < (1u << 23); m++)
    for (uint32_t sgn = 0; sgn < 2; sgn++) {
      const uint32_t bits = (sgn << 31) | (127u << 23) | m;
      float x;
      std::memcpy(&x, &bits, 4);
      const float q0 = x * y;
      const float q = recipe == 1 ? q0 : std::fmaf(std::fmaf(-df, q0, x), y, q0);
      const float ref = x / df;
      if (std::memcmp(&q, &ref, 4) != 0) {
        if (bad == 0 && ex) *ex = x;
        bad++;
      }
    }
  return bad;
}

}  // extern "C"

extern "C" {

// The relaxation kernels take the reciprocals of a 16-bit record's decoded
// diagonal (at most 11 significant bits: FP16 times a power of two; BF16: 8)
// as y0 = v_rcp_f32(c), which is within one ulp of 1/c, refined by one Newton
// step e = fma(-c, y0, 1), y = fma(e, y0, y0), instead of the IEEE division
// 1.0f / c, for c normal with 1/c normal.  1/c then lies at least 2^-12 ulp
// from a rounding midpoint (the significand ratio is 2^k / H with H an
// integer below 2^11), farther than the refined value's error, so y is the
// IEEE quotient.  Scaling c by a power of two scales every step exactly, so
// each binade behaves as [1, 2): this checks every significand of `bits`
// significant bits, both signs, with y0 each float within one ulp of the
// IEEE quotient r (its neighbours and r itself), bit for bit against r.
// recipe 1 returns y0 = the upper neighbour unrefined (a negative control);
// recipe 2 refines a full 24-bit significand the same way (the method's
// premise does not hold there).  Returns the mismatches; the first c is
// stored in *ex.
uint64_t orc_check_recip_short(int bits, int recipe, float* ex) {
  uint64_t bad = 0;
  const int sig = recipe == 2 ? 24 : bits;
  const uint32_t step = 1u << (24 - sig);
  for (uint32_t m = 0; m < (1u << 23); m += step)
    for (uint32_t sgn = 0; sgn < 2; sgn++) {
      const uint32_t cb = (sgn << 31) | (127u << 23) | m;
      float c;
      std::memcpy(&c, &cb, 4);
      const float r = 1.0f / c;
      const float cand[3] = {std::nextafter(r, -INFINITY), r, std::nextafter(r, INFINITY)};
      for (int k = 0; k < 3; k++) {
        const float y0 = cand[k];
        float y;
        if (recipe == 1) {
          y = cand[2];
        } else {
          const float e = std::fmaf(-c, y0, 1.0f);
          y = std::fmaf(e, y0, y0);
        }
        if (std::memcmp(&y, &r, 4) != 0) {
          if (bad == 0 && ex) *ex = c;
          bad++;
        }
      }
    }
  return bad;
}

}  // extern "C"
