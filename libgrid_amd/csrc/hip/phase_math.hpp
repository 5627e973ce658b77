// Compact FP64 exp and sin/cos for the phase factors exp(c * V) of the potential step.
//
// The general-purpose device math library handles every argument range (Payne-Hanek
// reduction, special cases) and its inlined code costs more registers than the rest of a
// fused FFT pass, halving that pass's occupancy.  These versions use textbook methods:
//   exp:    x = k ln2 + r with a two-part (Cody-Waite) ln2, |r| <= ln2/2, Taylor series to
//           r^13 (truncation < 1e-17 relative), then 2^k by ldexp.
//   sincos: x = n pi/2 + r with a three-part pi/2 whose leading parts have 33 significant
//           bits, so n * part is exact for |n| < 2^20 (|x| < 1.6e6); Taylor series of sin to
//           r^17 and cos to r^18 on |r| <= pi/4 (truncation < 1e-17); quadrant from n mod 4.
// Both are within a few ulp of the correctly rounded result over their ranges (checked
// against glibc on the host by tests/test_phase_math.py).  |x| >= 1.6e6 in sincos loses
// accuracy gradually; a potential step with |t V / hbar| beyond that is meaningless anyway.
// Constants were derived to 50+ digits (pi by Machin's formula, ln 2 by atanh series).
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define LG_HD __host__ __device__ __forceinline__
#else
#define LG_HD inline
#endif

namespace lg {

// Branch-free, so that a thread's many evaluations interleave: arguments beyond the range are
// clamped by selects (a NaN fails both comparisons and stays NaN), and exp(709.79) overflows
// to inf and exp(-746) underflows to 0 in the final ldexp as the exact values would.
LG_HD double exp_f64(double x) {
  x = x > 709.79 ? 709.79 : x;
  x = x < -746.0 ? -746.0 : x;
  const double inv_ln2 = 1.4426950408889634;
  const double ln2_hi = 0.6931471806019545;      // 32 significant bits
  const double ln2_lo = -4.2009150726810846e-11; // ln 2 - ln2_hi
  const double k = rint(x * inv_ln2);
  const int ki = k == k ? (int)k : 0;  // |k| <= 1077 here; a NaN k is not converted
  const double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = 1.6059043836821613e-10;            // 1/13!
  p = p * r + 2.08767569878681e-09;             // 1/12!
  p = p * r + 2.505210838544172e-08;            // 1/11!
  p = p * r + 2.755731922398589e-07;            // 1/10!
  p = p * r + 2.7557319223985893e-06;           // 1/9!
  p = p * r + 2.48015873015873e-05;             // 1/8!
  p = p * r + 0.0001984126984126984;            // 1/7!
  p = p * r + 0.001388888888888889;             // 1/6!
  p = p * r + 0.008333333333333333;             // 1/5!
  p = p * r + 0.041666666666666664;             // 1/4!
  p = p * r + 0.16666666666666666;              // 1/3!
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, ki);
}

// Branch-free like exp_f64: the quadrant is applied by selects, and an inf or NaN argument is
// replaced by 0 for the evaluation and gives NaN at the end.  That test is a class test: the
// form x - x != 0 breaks when x is a product inlined from the caller, since the compiler may
// contract (a b) - (a b) into fma(a, b, -(a b)), which is the product's rounding error.
LG_HD void sincos_f64(double x_in, double *s, double *c) {
  const bool finite = __builtin_isfinite(x_in);
  const double x = finite ? x_in : 0.0;
  const double two_over_pi = 0.6366197723675814;
  const double pio2_hi = 1.5707963267341256;      // 33 significant bits
  const double pio2_mid = 6.077100506303966e-11;  // 33 significant bits
  const double pio2_lo = 2.0222662487959506e-21;
  const double n = rint(x * two_over_pi);
  const double r = ((x - n * pio2_hi) - n * pio2_mid) - n * pio2_lo;
  const double r2 = r * r;
  double ps = 2.8114572543455206e-15;   // 1/17!
  ps = ps * r2 - 7.647163731819816e-13;  // -1/15!
  ps = ps * r2 + 1.6059043836821613e-10; // 1/13!
  ps = ps * r2 - 2.505210838544172e-08;  // -1/11!
  ps = ps * r2 + 2.7557319223985893e-06; // 1/9!
  ps = ps * r2 - 0.0001984126984126984;  // -1/7!
  ps = ps * r2 + 0.008333333333333333;   // 1/5!
  ps = ps * r2 - 0.16666666666666666;    // -1/3!
  const double sr = r + r * r2 * ps;
  double pc = -1.5619206968586225e-16;   // -1/18!
  pc = pc * r2 + 4.779477332387385e-14;  // 1/16!
  pc = pc * r2 - 1.1470745597729725e-11; // -1/14!
  pc = pc * r2 + 2.08767569878681e-09;   // 1/12!
  pc = pc * r2 - 2.755731922398589e-07;  // -1/10!
  pc = pc * r2 + 2.48015873015873e-05;   // 1/8!
  pc = pc * r2 - 0.001388888888888889;   // -1/6!
  pc = pc * r2 + 0.041666666666666664;   // 1/4!
  pc = pc * r2 - 0.5;
  const double cr = 1.0 + r2 * pc;
  const int q = (int)(n - 4.0 * floor(0.25 * n));  // n mod 4 in 0..3, exact for |n| < 2^52
  const double s0 = (q & 1) ? cr : sr, c0 = (q & 1) ? sr : cr;
  const double nan = __builtin_nan("");
  *s = finite ? ((q & 2) ? -s0 : s0) : nan;
  *c = finite ? (((q + 1) & 2) ? -c0 : c0) : nan;
}

// sin and cos in FP32 of an FP64 argument (the phase of an FP32 potential step): x = n pi/2 + r
// reduced in FP64 with the two leading parts of pi/2 above (exact products for |n| < 2^20;
// the dropped third part is below 1e-14 absolute there), only |r| <= pi/4 rounded to FP32, then
// Taylor series of sin to r^9 and cos to r^10 in FP32 (truncation < 3e-9 and 1.2e-10), quadrant
// from n mod 4 taken in FP64.  Within 2 FP32 ulp of the correctly rounded result
// (tests/test_phase_math.py), and more accurate than rounding x itself to FP32 first.
LG_HD void sincos_f32(double x, float *s, float *c) {
  const double two_over_pi = 0.6366197723675814;
  const double pio2_hi = 1.5707963267341256;
  const double pio2_mid = 6.077100506303966e-11;
  const double n = rint(x * two_over_pi);
  const float r = (float)((x - n * pio2_hi) - n * pio2_mid);
  const double nq = n - 4.0 * floor(0.25 * n);  // n mod 4 in 0..3, NaN for an inf or NaN x
  const int q = nq == nq ? (int)nq : 0;          // (r is NaN then, and so are sin and cos)
  const float r2 = r * r;
  float ps = 2.7557319e-06f;   // 1/9!
  ps = ps * r2 - 1.9841270e-04f;  // -1/7!
  ps = ps * r2 + 8.3333333e-03f;  // 1/5!
  ps = ps * r2 - 1.6666667e-01f;  // -1/3!
  const float sr = r + (r * r2) * ps;
  float pc = -2.7557319e-07f;  // -1/10!
  pc = pc * r2 + 2.4801587e-05f;  // 1/8!
  pc = pc * r2 - 1.3888889e-03f;  // -1/6!
  pc = pc * r2 + 4.1666667e-02f;  // 1/4!
  pc = pc * r2 - 0.5f;
  const float cr = 1.0f + r2 * pc;
  const float s0 = (q & 1) ? cr : sr, c0 = (q & 1) ? sr : cr;
  *s = (q & 2) ? -s0 : s0;
  *c = ((q + 1) & 2) ? -c0 : c0;
}

}  // namespace lg
