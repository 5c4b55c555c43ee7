// cr_trig.h — sin and cos of an f32 argument evaluated in f64 and rounded once to f32, for the
// twist / bend modifiers (sdf.rs:220-231) and the shadow-jitter hash (march.rs:242-244).
//
// The reference's f32 arithmetic takes sin/cos from the host libm, which returns the correctly
// rounded f32 for nearly every input; twist, bend and the jitter hash amplify a different
// last bit into visible pixels, so the device needs the correctly rounded value too.  The
// device library's own f64 sin/cos gives it but carries a large-argument reduction that costs
// registers in every kernel that inlines a field evaluation; this version covers |x| < 2^20
// with a three-part Cody-Waite reduction by pi/2 (k * P1 and k * P2 exact for |k| < 2^20) and
// the Taylor series of sin / cos on [-pi/4, pi/4] through x^17 / x^18 (truncation error below
// 2^-62), and sends larger or non-finite arguments to the library routines.
//
// Plain C++ for the host compiler as well: every operation is one IEEE f64 operation in source
// order (the library builds with -ffp-contract=off), so host and device results are identical
// and a host program can check the function against libm over all f32 inputs.
#pragma once

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

#if defined(__HIP_DEVICE_COMPILE__)
// Kept out of line: the library's large-argument reduction stays off the common path.
static __device__ __noinline__ void sincos_wide(double x, double* s, double* c) { sincos(x, s, c); }
#elif !defined(__HIP__)
inline void sincos_wide(double x, double* s, double* c) {
  *s = std::sin(x);
  *c = std::cos(x);
}
#else
static inline void sincos_wide(double x, double* s, double* c) {
  *s = ::sin(x);
  *c = ::cos(x);
}
#endif

// sin(x) and cos(x) of an f32, each rounded once from f64.
MM3E_HD void cr_sincos(float x, float& s, float& c) {
  const double xd = x;
  double sd, cd;
  if (xd < 1048576.0 && xd > -1048576.0) {
    const double k = __builtin_rint(xd * kTwoOverPi);
    const double r = ((xd - k * kPio2Hi) - k * kPio2Mid) - k * kPio2Lo;
    double ps, pc;
    sincos_reduced(r, ps, pc);
    switch ((int)k & 3) {
      case 0: sd = ps; cd = pc; break;
      case 1: sd = pc; cd = -ps; break;
      case 2: sd = -ps; cd = -pc; break;
      default: sd = -pc; cd = ps; break;
    }
  } else {
    sincos_wide(xd, &sd, &cd);
  }
  s = (float)sd;
  c = (float)cd;
}
MM3E_HD float cr_sin(float x) {
  float s, c;
  cr_sincos(x, s, c);
  return s;
}

}  // namespace mm3e
