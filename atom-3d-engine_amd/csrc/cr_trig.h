// cr_trig.h — sin and cos of an f32 argument evaluated in f64 and rounded once to f32, for the
// twist / bend modifiers (sdf.rs:220-231) and the shadow-jitter hash (march.rs:242-244).
//
// The reference's f32 arithmetic takes sin/cos from the host libm, which returns the correctly
// rounded f32 for nearly every input; twist, bend and the jitter hash amplify a different
// last bit into visible pixels, so the device needs the correctly rounded value too.  The
// argument is reduced by pi/2 in f64 — |x| < 2^20 with a three-part Cody-Waite reduction (k * P1
// and k * P2 exact for |k| < 2^20), larger finite |x| exactly in integer arithmetic against 96
// bits of 2/pi (Payne-Hanek) — and sin / cos on [-pi/4, pi/4] come from their Taylor series
// through x^17 / x^18 (truncation error below 2^-62), Horner steps as separate f64 multiplies
// and adds (with fused multiply-adds the c2 generated kernel spilled 368 instead of 148 bytes
// per lane and ran 17 % slower).  Everything is inline: no library call, whose calling
// convention made every kernel that inlines a field evaluation save registers around it (c2
// generated kernel 1.358 -> 1.333 ms without it).
//
// Plain C++ for the host compiler as well: every step is one IEEE f64 operation or exact
// integer arithmetic, in source order, so host and device results are identical and a
// host program can check the function against libm over all f32 inputs.
#pragma once

#include <stdint.h>

#if defined(__HIP__)
#define MM3E_HD __host__ __device__ inline
#else
#include <cmath>
#define MM3E_HD inline
#endif

namespace mm3e {

// pi/2 = P1 + P2 + P3 with P1, P2 holding 33 significant bits each.
constexpr double kPio2Hi = 0x1.921fb544p+0;
constexpr double kPio2Mid = 0x1.0b4611a6p-34;
constexpr double kPio2Lo = 0x1.3198a2e037073p-69;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
// pi/2 * 2^-64: scales a 64-bit two's-complement fraction of a quarter turn to radians.
constexpr double kPio2Over2p64 = 0x1.921fb54442d18p-64;

// sin / cos of r in [-pi/4, pi/4]: Taylor series, coefficients +-1/n! rounded to f64.
MM3E_HD void sincos_reduced(double r, double& s, double& c) {
  const double r2 = r * r;
  double ps = 0x1.952c77030ad4ap-49;        //  1/17!
  ps = -0x1.ae7f3e733b81fp-41 + r2 * ps;    // -1/15!
  ps = 0x1.6124613a86d09p-33 + r2 * ps;     //  1/13!
  ps = -0x1.ae64567f544e4p-26 + r2 * ps;    // -1/11!
  ps = 0x1.71de3a556c734p-19 + r2 * ps;     //  1/9!
  ps = -0x1.a01a01a01a01ap-13 + r2 * ps;    // -1/7!
  ps = 0x1.1111111111111p-7 + r2 * ps;      //  1/5!
  ps = -0x1.5555555555555p-3 + r2 * ps;     // -1/3!
  s = r + (r * r2) * ps;
  double pc = -0x1.6827863b97d97p-53;       // -1/18!
  pc = 0x1.ae7f3e733b81fp-45 + r2 * pc;     //  1/16!
  pc = -0x1.93974a8c07c9dp-37 + r2 * pc;    // -1/14!
  pc = 0x1.1eed8eff8d898p-29 + r2 * pc;     //  1/12!
  pc = -0x1.27e4fb7789f5cp-22 + r2 * pc;    // -1/10!
  pc = 0x1.a01a01a01a01ap-16 + r2 * pc;     //  1/8!
  pc = -0x1.6c16c16c16c17p-10 + r2 * pc;    // -1/6!
  pc = 0x1.5555555555555p-5 + r2 * pc;      //  1/4!
  pc = -0.5 + r2 * pc;                      // -1/2!
  c = 1.0 + r2 * pc;
}

// Quarter turns of |x| for a finite |x| >= 2^20 (biased exponent be in 147..254, mantissa m with
// its implicit bit): |x| * 2/pi = q + f exactly to 2^-70, q mod 4 returned, f in [-1/2, 1/2) as a
// 64-bit two's-complement fraction.  |x| = m * 2^(be-150); the bits of 2/pi whose products with
// m are multiples of 4 are skipped, the next 96 bits multiply m (24 x 96 -> 120 bits) and the
// binary point of the product sits 94 bits from its low end.
MM3E_HD uint32_t reduce_large(uint32_t be, uint32_t m, int64_t& frac) {
  // 2/pi, big-endian bits, after five zero bits (so bit j of the table is bit j - 4 of 2/pi's
  // binary expansion 0.101000101111...).
  constexpr uint32_t kT[8] = {0x0517cc1bu, 0x727220a9u, 0x4fe13abeu, 0x8fa9a6eeu,
                              0x06db14acu, 0xc9e21c82u, 0x0ff28b1du, 0x5ef5de2bu};
  const uint32_t j0 = be - 147u;  // first table bit used: (be - 150) + 3, in 0..107
  const uint32_t k = j0 >> 5, sh = j0 & 31u;
  const uint64_t a01 = ((uint64_t)kT[k] << 32) | kT[k + 1];
  const uint64_t a12 = ((uint64_t)kT[k + 1] << 32) | kT[k + 2];
  const uint64_t a23 = ((uint64_t)kT[k + 2] << 32) | kT[k + 3];
  const uint32_t w2 = (uint32_t)(a01 >> (32u - sh));  // 96 bits of 2/pi, most significant first
  const uint32_t w1 = (uint32_t)(a12 >> (32u - sh));
  const uint32_t w0 = (uint32_t)(a23 >> (32u - sh));
  const uint64_t p0 = (uint64_t)m * w0;
  const uint64_t t1 = (uint64_t)m * w1 + (p0 >> 32);
  const uint64_t t2 = (uint64_t)m * w2 + (t1 >> 32);
  // product = t2 * 2^64 + (t1 mod 2^32) * 2^32 + (p0 mod 2^32); integer part = product >> 94
  const uint64_t f64 = ((t2 & 0x3fffffffull) << 34) | ((t1 & 0xffffffffull) << 2) | ((p0 & 0xffffffffull) >> 30);
  uint32_t q = (uint32_t)(t2 >> 30) & 3u;
  q = (q + (uint32_t)(f64 >> 63)) & 3u;  // fraction >= 1/2: one more quarter turn, fraction - 1
  frac = (int64_t)f64;
  return q;
}

// sin(x) and cos(x) of an f32, each rounded once from f64.
MM3E_HD void cr_sincos(float x, float& s, float& c) {
  const double xd = x;
  double ps, pc;
  uint32_t q;
  bool neg = false;
  if (xd < 1048576.0 && xd > -1048576.0) {
    const double k = __builtin_rint(xd * kTwoOverPi);
    const double r = ((xd - k * kPio2Hi) - k * kPio2Mid) - k * kPio2Lo;
    sincos_reduced(r, ps, pc);
    q = (uint32_t)(int32_t)k & 3u;
  } else {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    const uint32_t be = (u >> 23) & 0xffu;
    if (be == 0xffu) {  // inf or NaN: NaN, as libm
      s = x - x;
      c = x - x;
      return;
    }
    neg = (u >> 31) != 0u;
    int64_t frac;
    q = reduce_large(be, (u & 0x7fffffu) | 0x800000u, frac);
    sincos_reduced((double)frac * kPio2Over2p64, ps, pc);
  }
  double sd, cd;
  switch (q) {
    case 0: sd = ps; cd = pc; break;
    case 1: sd = pc; cd = -ps; break;
    case 2: sd = -ps; cd = -pc; break;
    default: sd = -pc; cd = ps; break;
  }
  if (neg) sd = -sd;  // reduce_large took |x|: sin is odd, cos even
  s = (float)sd;
  c = (float)cd;
}
MM3E_HD float cr_sin(float x) {
  float s, c;
  cr_sincos(x, s, c);
  return s;
}

}  // namespace mm3e
