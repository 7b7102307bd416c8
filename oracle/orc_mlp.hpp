// Oracle elementwise routines of the saturation-guess network -- TEST
// INFRASTRUCTURE ONLY (the results are compared with it; never the product).
//
// The device library carries its own copy of the same algorithms, so that
// everything outside the matrix products is evaluated identically on both
// sides:
//   bf16:  round-to-nearest-even of an IEEE binary32 to its upper 16 bits
//          (NaN kept quiet).
//   exp_f: Cody-Waite reduction x = k ln2 + r with ln2 split as
//          0.693359375 - 2.12194440e-4 (the split of Cephes' expf, exact in
//          k * hi for |k| < 2^15), Taylor series of e^r through r^6 in Horner
//          form, 2^k by ldexp (exact).  Relative error < 2e-7 on |x| < 80.
//   gelu:  the tanh form of Hendrycks & Gimpel (2016),
//          0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))), with
//          tanh(u) = 1 - 2 / (e^{2u} + 1), and its derivative.
// All of it is single precision with no fused multiply-add (the oracle is
// built with -ffp-contract=off).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

inline uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

inline float bf16_float(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

inline float bf16_round(float f) { return bf16_float(bf16_bits(f)); }

inline float exp_f(float x) {
  if (x != x) return x;
  if (x > 88.0f) return HUGE_VALF;
  if (x < -87.0f) return 0.0f;
  float kf = std::floor(x * 1.44269504f + 0.5f);
  float r = (x - kf * 0.693359375f) - kf * -2.12194440e-4f;
  float p = 1.38888889e-3f;
  p = p * r + 8.33333333e-3f;
  p = p * r + 4.16666667e-2f;
  p = p * r + 1.66666667e-1f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return std::ldexp(p, (int)kf);
}

inline float gelu_tanh_f(float z) {
  float z3 = (z * z) * z;
  float u = 0.797884583f * (z + 0.044715f * z3);
  float e = exp_f(2.0f * u);
  return 1.0f - 2.0f / (e + 1.0f);
}

inline float gelu_f(float z) {
  float t = gelu_tanh_f(z);
  return (0.5f * z) * (1.0f + t);
}

inline float gelu_grad_f(float z) {
  float t = gelu_tanh_f(z);
  float dudz = 0.797884583f * (1.0f + 0.134145f * (z * z));
  return 0.5f * (1.0f + t) + (0.5f * z) * ((1.0f - t * t) * dudz);
}

// Double-precision GELU (same tanh form, libm tanh) for the exact mode.
inline double gelu_d(double z) {
  double u = 0.7978845608028654 * (z + 0.044715 * z * z * z);
  return 0.5 * z * (1.0 + std::tanh(u));
}

inline double gelu_grad_d(double z) {
  double t = std::tanh(0.7978845608028654 * (z + 0.044715 * z * z * z));
  double dudz = 0.7978845608028654 * (1.0 + 3.0 * 0.044715 * z * z);
  return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * dudz;
}

}  // namespace orc
