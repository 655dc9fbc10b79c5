// FP64 sine and cosine of one argument, for the plain FP64 path.
//
// Cody-Waite reduction (W. J. Cody and W. Waite, "Software Manual for the
// Elementary Functions", 1980): x = n pi/2 + r with n = rint(2x/pi) and
// pi/2 split into P1 + P2 + P3, P1 and P2 of 33 significant bits, so that
// n P1 and n P2 are exact for |n| < 2^20 and r carries the full FP64
// precision of x - n pi/2.  Then the Taylor series of sin r and cos r on
// |r| <= pi/4, to r^15 and r^16 (the first omitted terms are below 7e-17 and
// 2e-18 relative), and the quadrant n mod 4 selects and signs the pair.
// Valid for |x| < 2^20 pi/2 (about 1.6e6); the force step's phases stay
// below 1e4.  The three split constants are pi/2 rounded to 33 bits, the
// rest rounded to 33 bits, and the remainder rounded to FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

namespace glst {

// x with its sign bit XORed with bit 63 of m
__host__ __device__ inline double flip_sign(double x, uint64_t m) {
  uint64_t u;
  memcpy(&u, &x, sizeof u);
  u ^= m;
  memcpy(&x, &u, sizeof x);
  return x;
}

__host__ __device__ inline void sincos_cw(double x, double *s, double *c) {
  constexpr double kTwoOverPi = 0.6366197723675814;
  constexpr double kP1 = 1.5707963267341256;
  constexpr double kP2 = 6.077100506303966e-11;
  constexpr double kP3 = 2.0222662487959506e-21;
  const double n = rint(x * kTwoOverPi);
  double r = fma(-n, kP1, x);
  r = fma(-n, kP2, r);
  r = fma(-n, kP3, r);
  const double r2 = r * r;
  // sin r = r + r^3 (-1/3! + r^2 (1/5! - r^2 (1/7! - ...)))
  double ps = -7.647163731819816e-13;  // -1/15!
  ps = fma(ps, r2, 1.6059043836821613e-10);   // 1/13!
  ps = fma(ps, r2, -2.505210838544172e-08);   // -1/11!
  ps = fma(ps, r2, 2.7557319223985893e-06);   // 1/9!
  ps = fma(ps, r2, -0.0001984126984126984);   // -1/7!
  ps = fma(ps, r2, 0.008333333333333333);     // 1/5!
  ps = fma(ps, r2, -0.16666666666666666);     // -1/3!
  const double sn = fma(r * r2, ps, r);
  // cos r = 1 + r^2 (-1/2! + r^2 (1/4! - ...))
  double pc = 4.779477332387385e-14;          // 1/16!
  pc = fma(pc, r2, -1.1470745597729725e-11);  // -1/14!
  pc = fma(pc, r2, 2.08767569878681e-09);     // 1/12!
  pc = fma(pc, r2, -2.755731922398589e-07);   // -1/10!
  pc = fma(pc, r2, 2.48015873015873e-05);     // 1/8!
  pc = fma(pc, r2, -0.001388888888888889);    // -1/6!
  pc = fma(pc, r2, 0.041666666666666664);     // 1/4!
  pc = fma(pc, r2, -0.5);                     // -1/2!
  const double cs = fma(r2, pc, 1.0);
  // x = n pi/2 + r: quadrant q = n mod 4 (two's complement for n < 0;
  // |n| < 2^20 fits a 32-bit conversion)
  const uint32_t q = static_cast<uint32_t>(static_cast<int32_t>(n));
  const double a = (q & 1u) ? cs : sn;  // sin x up to sign
  const double b = (q & 1u) ? sn : cs;  // cos x up to sign
  // signs: bit 1 of q for sin x, bit 1 of q + 1 for cos x, moved to bit 63
  *s = flip_sign(a, static_cast<uint64_t>(q & 2u) << 62);
  *c = flip_sign(b, static_cast<uint64_t>((q + 1u) & 2u) << 62);
}

}  // namespace glst
