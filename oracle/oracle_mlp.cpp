// oracle_mlp.cpp -- CPU oracle of the saturation-guess network.  TEST
// INFRASTRUCTURE ONLY: it checks the device forward / backward kernels and
// is never on the product path.
//
// The network has no reference counterpart (the reference's adjustment,
// kintera ThermoY::forward called at src/tasklist/implicit_hydro_tasks.cpp:308,
// iterates from the all-vapour state).  It is 4 -> 64 -> 64 -> 64 -> 3 with
// tanh-GELU after each hidden layer, parameters in one fp32 vector:
//   W1 (64x4), b1 (64), W2 (64x64), b2, W3 (64x64), b3, W4 (3x64), b4 (3),
// every W row-major (out, in).  Loss = mean over samples and outputs of
// (y - t)^2.
//
// Two modes:
//   emulate = 1: the device's number formats.  Weights, layer inputs, and
//     the output / hidden gradients that feed matrix products are rounded to
//     bf16; products accumulate in double here (fp32 inside MFMA on the
//     device) and are rounded to fp32; GELU, its derivative and the loss
//     gradient are the single-precision routines of orc_mlp.hpp.
//   emulate = 0: everything in double, no rounding (for finite differences).
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "orc_mlp.hpp"

namespace {

constexpr int NL = 4;
constexpr int DIN[NL] = {4, 64, 64, 64};
constexpr int DOUT[NL] = {64, 64, 64, 3};

struct Offs {
  int w[NL], b[NL], total;
  Offs() {
    int o = 0;
    for (int l = 0; l < NL; ++l) {
      w[l] = o;
      o += DIN[l] * DOUT[l];
      b[l] = o;
      o += DOUT[l];
    }
    total = o;
  }
};

struct Net {
  Offs off;
  bool em;
  std::vector<double> p;  // parameters, bf16-rounded weights in emulate mode
  Net(const float* params, bool emulate) : em(emulate) {
    p.resize(off.total);
    for (int l = 0; l < NL; ++l) {
      for (int e = 0; e < DIN[l] * DOUT[l]; ++e) {
        float v = params[off.w[l] + e];
        p[off.w[l] + e] = em ? (double)orc::bf16_round(v) : (double)v;
      }
      for (int e = 0; e < DOUT[l]; ++e) p[off.b[l] + e] = params[off.b[l] + e];
    }
  }
  double rb(double v) const { return em ? (double)orc::bf16_round((float)v) : v; }
  double r32(double v) const { return em ? (double)(float)v : v; }
  double act(double z) const {
    return em ? (double)orc::gelu_f((float)z) : orc::gelu_d(z);
  }
  double dact(double z) const {
    return em ? (double)orc::gelu_grad_f((float)z) : orc::gelu_grad_d(z);
  }
  // h[l] = input of layer l (rounded as the products see it), z[l] its output
  void forward(const float* x, std::vector<double> h[NL + 1],
               std::vector<double> z[NL]) const {
    h[0].resize(DIN[0]);
    for (int i = 0; i < DIN[0]; ++i) h[0][i] = rb((double)x[i]);
    for (int l = 0; l < NL; ++l) {
      z[l].resize(DOUT[l]);
      for (int o = 0; o < DOUT[l]; ++o) {
        double acc = p[off.b[l] + o];
        const double* w = &p[off.w[l] + o * DIN[l]];
        for (int i = 0; i < DIN[l]; ++i) acc += w[i] * h[l][i];
        z[l][o] = r32(acc);
      }
      if (l + 1 < NL) {
        h[l + 1].resize(DOUT[l]);
        for (int o = 0; o < DOUT[l]; ++o) h[l + 1][o] = rb(act(z[l][o]));
      }
    }
  }
};

}  // namespace

extern "C" {

int or_mlp_nparam(void) { return Offs().total; }

float or_mlp_gelu(float z) { return orc::gelu_f(z); }
float or_mlp_gelu_grad(float z) { return orc::gelu_grad_f(z); }
float or_mlp_exp(float x) { return orc::exp_f(x); }
unsigned short or_mlp_bf16(float x) { return orc::bf16_bits(x); }

int or_mlp_forward(const float* params, int n, const float* x, double* y,
                   int emulate) {
  if (!params || n < 0 || (n > 0 && (!x || !y))) return CG_ERR_ARG;
  Net net(params, emulate != 0);
  std::vector<double> h[NL + 1], z[NL];
  for (int s = 0; s < n; ++s) {
    net.forward(x + 4 * (size_t)s, h, z);
    for (int o = 0; o < DOUT[NL - 1]; ++o) y[3 * (size_t)s + o] = z[NL - 1][o];
  }
  return CG_OK;
}

// Gradient of the mean squared error over n samples; grad has the layout of
// the parameter vector.  Backward pass (emulate mode as on the device):
//   dY = fp32((y - t) * 2/(3n)); dH_{l} = W_l^T bf16(dZ_l); dZ_{l-1} =
//   fp32(dH_l * gelu'(Z_{l-1})); dW_l = sum_s bf16(dZ_l) h_l^T,
//   db_l = sum_s dZ_l.
int or_mlp_loss_grad(const float* params, int n, const float* x,
                     const float* t, double* grad, double* loss, int emulate) {
  if (!params || n <= 0 || !x || !t || !grad) return CG_ERR_ARG;
  Net net(params, emulate != 0);
  const Offs& off = net.off;
  std::fill(grad, grad + off.total, 0.0);
  double sse = 0.0;
  const double scale = emulate ? (double)(2.0f / (3.0f * (float)n)) : 2.0 / (3.0 * n);
  std::vector<double> h[NL + 1], z[NL], dz[NL], dzb[NL];
  for (int s = 0; s < n; ++s) {
    net.forward(x + 4 * (size_t)s, h, z);
    dz[NL - 1].resize(DOUT[NL - 1]);
    for (int o = 0; o < DOUT[NL - 1]; ++o) {
      double d = z[NL - 1][o] - (double)t[3 * (size_t)s + o];
      sse += d * d;
      dz[NL - 1][o] = emulate ? (double)((float)((float)z[NL - 1][o] -
                                                 t[3 * (size_t)s + o]) *
                                         (float)scale)
                              : d * scale;
    }
    for (int l = NL - 1; l >= 0; --l) {
      dzb[l].resize(DOUT[l]);
      for (int o = 0; o < DOUT[l]; ++o) dzb[l][o] = net.rb(dz[l][o]);
      for (int o = 0; o < DOUT[l]; ++o) {
        grad[off.b[l] + o] += dz[l][o];
        double* gw = grad + off.w[l] + o * DIN[l];
        for (int i = 0; i < DIN[l]; ++i) gw[i] += dzb[l][o] * h[l][i];
      }
      if (l == 0) break;
      dz[l - 1].resize(DIN[l]);
      for (int i = 0; i < DIN[l]; ++i) {
        double acc = 0.0;
        for (int o = 0; o < DOUT[l]; ++o)
          acc += net.p[off.w[l] + o * DIN[l] + i] * dzb[l][o];
        double dh = net.r32(acc);
        dz[l - 1][i] = emulate ? (double)((float)dh * orc::gelu_grad_f((float)z[l - 1][i]))
                               : dh * net.dact(z[l - 1][i]);
      }
    }
  }
  if (emulate)
    for (int e = 0; e < off.total; ++e) grad[e] = (double)(float)grad[e];
  if (loss) *loss = sse / (3.0 * n);
  return CG_OK;
}

}  // extern "C"
