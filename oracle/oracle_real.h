// oracle_real.h — scalar types for the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// The oracle is written once over a scalar type R and instantiated twice:
//   R = float  : the renderer used as the parity reference and the CPU baseline;
//   R = CountF : the same code path with every FP32 operation counted, which yields the
//                algorithmic FLOPs per frame used for the roofline figure.
// Counting rules: add, sub, mul, div, sqrt, min, max, compare = 1 FLOP each; abs and negate
// are sign-bit modifiers (0); floor/trunc/round = 1; sin, cos, exp, pow, tan are counted
// separately as transcendental ops.  Host code is compiled with -ffp-contract=off, so no
// FMA is ever formed and the float instantiation rounds after every operation, exactly like
// the reference's Rust f32 arithmetic.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

struct OpCounts {
  uint64_t flops = 0;   // add/sub/mul/div/sqrt/min/max/cmp/floor
  uint64_t trans = 0;   // sin/cos/exp/pow/tan
  uint64_t sqrt = 0;    // sqrt (also included in flops)
  uint64_t div = 0;     // div (also included in flops)
};

extern thread_local OpCounts* g_ops;

struct CountF {
  float v;
  CountF() : v(0.0f) {}
  CountF(float x) : v(x) {}  // NOLINT: implicit so literals mix in
  static inline void tick(uint64_t n = 1) {
    if (g_ops) g_ops->flops += n;
  }
  friend inline CountF operator+(CountF a, CountF b) { tick(); return CountF(a.v + b.v); }
  friend inline CountF operator-(CountF a, CountF b) { tick(); return CountF(a.v - b.v); }
  friend inline CountF operator*(CountF a, CountF b) { tick(); return CountF(a.v * b.v); }
  friend inline CountF operator/(CountF a, CountF b) {
    tick();
    if (g_ops) g_ops->div++;
    return CountF(a.v / b.v);
  }
  friend inline CountF operator-(CountF a) { return CountF(-a.v); }
  friend inline bool operator<(CountF a, CountF b) { tick(); return a.v < b.v; }
  friend inline bool operator>(CountF a, CountF b) { tick(); return a.v > b.v; }
  friend inline bool operator<=(CountF a, CountF b) { tick(); return a.v <= b.v; }
  friend inline bool operator>=(CountF a, CountF b) { tick(); return a.v >= b.v; }
  friend inline bool operator==(CountF a, CountF b) { tick(); return a.v == b.v; }
  friend inline bool operator!=(CountF a, CountF b) { tick(); return a.v != b.v; }
  CountF& operator+=(CountF o) { *this = *this + o; return *this; }
  CountF& operator*=(CountF o) { *this = *this * o; return *this; }
};

// ---- scalar functions (Rust f32 semantics) ---------------------------------------------
inline float val(float x) { return x; }
inline float val(CountF x) { return x.v; }

// f32::min / f32::max are IEEE minNum/maxNum.
inline float rmin(float a, float b) { return std::fmin(a, b); }
inline float rmax(float a, float b) { return std::fmax(a, b); }
inline CountF rmin(CountF a, CountF b) { CountF::tick(); return CountF(std::fmin(a.v, b.v)); }
inline CountF rmax(CountF a, CountF b) { CountF::tick(); return CountF(std::fmax(a.v, b.v)); }

// f32::clamp: `if x < lo {lo} else if x > hi {hi} else x` (NaN passes through).
template <class R> inline R rclamp(R x, R lo, R hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

inline float rabs(float x) { return std::fabs(x); }
inline CountF rabs(CountF x) { return CountF(std::fabs(x.v)); }

inline float rsqrt_(float x) { return std::sqrt(x); }
inline CountF rsqrt_(CountF x) {
  CountF::tick();
  if (g_ops) g_ops->sqrt++;
  return CountF(std::sqrt(x.v));
}

inline void tick_trans() {
  if (g_ops) g_ops->trans++;
}
inline float rsin(float x) { return sinf(x); }
inline float rcos(float x) { return cosf(x); }
inline float rexp(float x) { return expf(x); }
inline float rpow(float x, float y) { return powf(x, y); }
inline float rtan(float x) { return tanf(x); }
inline CountF rsin(CountF x) { tick_trans(); return CountF(sinf(x.v)); }
inline CountF rcos(CountF x) { tick_trans(); return CountF(cosf(x.v)); }
inline CountF rexp(CountF x) { tick_trans(); return CountF(expf(x.v)); }
inline CountF rpow(CountF x, CountF y) { tick_trans(); return CountF(powf(x.v, y.v)); }
inline CountF rtan(CountF x) { tick_trans(); return CountF(tanf(x.v)); }

inline float rfloor(float x) { return std::floor(x); }
inline CountF rfloor(CountF x) { CountF::tick(); return CountF(std::floor(x.v)); }
inline float rtrunc(float x) { return std::trunc(x); }
inline CountF rtrunc(CountF x) { CountF::tick(); return CountF(std::trunc(x.v)); }
// Rust f32::round: half away from zero (C roundf).
inline float rround(float x) { return std::round(x); }
inline CountF rround(CountF x) { CountF::tick(); return CountF(std::round(x.v)); }

inline bool risfinite(float x) { return std::isfinite(x); }
inline bool risfinite(CountF x) { return std::isfinite(x.v); }

// f32::powi with a constant exponent: LLVM expands it by repeated squaring
// (Res = first set bit's square; x^5 = x * ((x*x)*(x*x))).
template <class R> inline R rpowi(R x, int n) {
  bool have = false;
  R res = R(1.0f);
  R cur = x;
  unsigned v = (unsigned)n;
  while (v) {
    if (v & 1u) {
      res = have ? res * cur : cur;
      have = true;
    }
    v >>= 1;
    if (v) cur = cur * cur;
  }
  return res;
}

// `as i32` from f32: saturating, NaN -> 0.
inline int32_t to_i32(float x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x < -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}
// `as usize` from f32 (non-negative after max(0.0)): saturating, NaN -> 0.
inline uint64_t to_usize(float x) {
  if (std::isnan(x) || x <= 0.0f) return 0;
  if (x >= 18446744073709551616.0f) return UINT64_MAX;
  return (uint64_t)x;
}
// `as u8` from f32 (used after clamp*255+0.5): truncating, saturating.
inline uint8_t to_u8(float x) {
  if (std::isnan(x) || x <= 0.0f) return 0;
  if (x >= 255.0f) return 255;
  return (uint8_t)x;
}

}  // namespace orc
