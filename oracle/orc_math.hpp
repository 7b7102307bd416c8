// Oracle math routines -- TEST INFRASTRUCTURE ONLY (results are compared
// with them; never on the product path).
//
// exp and log written from their textbook definitions so that the CPU oracle
// and the device code (which carries its own copy of the same algorithm)
// produce bit-identical results; glibc and the device math library may
// differ by an ulp, which the flux divergence would amplify.
//   exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor series of
//        e^r through r^13 (truncation < 5e-18 relative) in Horner form, 2^k
//        by ldexp (exact).
//   log: x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
//        s = (m-1)/(m+1), odd series through s^23; e ln2 added in two parts.
// Accuracy is checked against numpy to <= 2 ulp in tests/test_oracle_pins.py.
#pragma once
#include <cmath>

namespace orc {

inline double cg_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.2) return 0.0;
  const double inv_ln2 = 1.4426950408889634;
  const double ln2_hi = 0.6931471803691238;    // 32 significant bits
  const double ln2_lo = 1.9082149292705877e-10;
  double kd = std::floor(x * inv_ln2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double p = 1.6059043836821613e-10;
  p = p * r + 2.08767569878681e-09;
  p = p * r + 2.505210838544172e-08;
  p = p * r + 2.755731922398589e-07;
  p = p * r + 2.7557319223985893e-06;
  p = p * r + 2.48015873015873e-05;
  p = p * r + 0.0001984126984126984;
  p = p * r + 0.001388888888888889;
  p = p * r + 0.008333333333333333;
  p = p * r + 0.041666666666666664;
  p = p * r + 0.16666666666666666;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return std::ldexp(p, k);
}

inline double cg_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int e;
  double m = std::frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  if (m < 0.7071067811865476) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double q = 0.043478260869565216;
  q = q * s2 + 0.047619047619047616;
  q = q * s2 + 0.05263157894736842;
  q = q * s2 + 0.058823529411764705;
  q = q * s2 + 0.06666666666666667;
  q = q * s2 + 0.07692307692307693;
  q = q * s2 + 0.09090909090909091;
  q = q * s2 + 0.1111111111111111;
  q = q * s2 + 0.14285714285714285;
  q = q * s2 + 0.2;
  q = q * s2 + 0.3333333333333333;
  double lm = 2.0 * s + 2.0 * s * (s2 * q);
  double ed = (double)e;
  return ed * 0.6931471803691238 + (lm + ed * 1.9082149292705877e-10);
}

inline double cg_pow(double a, double b) { return cg_exp(b * cg_log(a)); }

inline double sqr(double x) { return x * x; }

}  // namespace orc
