// mlp.hip -- C-ABI of the saturation-guess network and of the saturation
// adjustment (plain and guided): parameter handling, inference, training on
// the GPU (forward / backward / Adam kernels of mlp.hpp), training-pair
// generation from the plain adjustment.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/canoe_gpu.h"
#include "capi_internal.hpp"
#include "kernels.hpp"
#include "mlp.hpp"

using namespace cg;
using namespace cg_host;

static_assert(CG_MLP_NPARAM == mlp::NPARAM, "parameter count");

struct cg_mlp {
  cg_ctx* ctx = nullptr;
  float* P = nullptr;      // fp32 master parameters
  float* m = nullptr;      // Adam first moment
  float* v = nullptr;      // Adam second moment
  float* grad = nullptr;   // reduced gradient
  float* loss = nullptr;   // device losses, one per training step
  uint16_t* tiles = nullptr;
  long long adam_t = 0;
  // training scratch, sized for nbp batch positions
  long long nbp = 0;
  float* zbuf = nullptr;
  uint16_t* hbuf = nullptr;
  float* dzbuf = nullptr;
  float* lossp = nullptr;
  float* part = nullptr;
  int loss_cap = 0;
};

namespace {

void free_scratch(cg_mlp* m) {
  hipFree(m->zbuf);
  hipFree(m->hbuf);
  hipFree(m->dzbuf);
  hipFree(m->lossp);
  hipFree(m->part);
  m->zbuf = nullptr;
  m->hbuf = nullptr;
  m->dzbuf = nullptr;
  m->lossp = nullptr;
  m->part = nullptr;
  m->nbp = 0;
}

int ensure_scratch(cg_mlp* m, long long nb) {
  cg_ctx* ctx = m->ctx;
  long long nbp = (nb + mlp::CHUNK - 1) / mlp::CHUNK * mlp::CHUNK;
  if (nbp <= m->nbp) return CG_OK;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_scratch(m);
  const long long nwave = nbp / 64, nchunk = nbp / mlp::CHUNK;
  HIPCHK(ctx, hipMalloc((void**)&m->zbuf, sizeof(float) * nwave * 3 * 4096));
  HIPCHK(ctx, hipMalloc((void**)&m->hbuf, sizeof(uint16_t) * mlp::HROWS * nbp));
  HIPCHK(ctx, hipMalloc((void**)&m->dzbuf, sizeof(float) * mlp::DZROWS * nbp));
  HIPCHK(ctx, hipMalloc((void**)&m->lossp, sizeof(float) * nwave));
  HIPCHK(ctx, hipMalloc((void**)&m->part, sizeof(float) * nchunk * mlp::NPARAM));
  m->nbp = nbp;
  return CG_OK;
}

int pack(cg_mlp* m) {
  mlp::k_mlp_pack<<<nblk(mlp::NTILE * mlp::TILE, 256), 256, 0, m->ctx->stream>>>(
      m->P, m->tiles);
  HIPCHK(m->ctx, hipGetLastError());
  return CG_OK;
}

// One gradient evaluation over nb batch positions (random = 0: samples
// 0..nb-1; random = 1: seeded draws from npairs).  Leaves the gradient in
// m->grad and the loss in loss_dev.
int gradient(cg_mlp* m, long long nb, long long npairs, const float* x,
             const float* t, uint64_t seed, int random, float* loss_dev) {
  cg_ctx* ctx = m->ctx;
  int e = ensure_scratch(m, nb);
  if (e) return e;
  const long long nbp = (nb + mlp::CHUNK - 1) / mlp::CHUNK * mlp::CHUNK;
  const long long nwave = nbp / 64, nchunk = nbp / mlp::CHUNK;
  const float scale = 2.0f / (3.0f * (float)nb);
  hipStream_t s = ctx->stream;
  mlp::k_mlp_fb<<<(unsigned)(nbp / 256), 256, 0, s>>>(
      m->tiles, m->P, nb, nbp, npairs, x, t, seed, random, scale, m->zbuf,
      m->hbuf, m->dzbuf, m->lossp);
  mlp::k_mlp_dw<0><<<(unsigned)nchunk, 64, 0, s>>>(m->hbuf, m->dzbuf, nbp, m->part);
  mlp::k_mlp_dw<1><<<(unsigned)nchunk, 64, 0, s>>>(m->hbuf, m->dzbuf, nbp, m->part);
  mlp::k_mlp_dw<2><<<(unsigned)nchunk, 64, 0, s>>>(m->hbuf, m->dzbuf, nbp, m->part);
  mlp::k_mlp_dw<3><<<(unsigned)nchunk, 64, 0, s>>>(m->hbuf, m->dzbuf, nbp, m->part);
  mlp::k_mlp_reduce<<<nblk(mlp::NPARAM, 256), 256, 0, s>>>(
      m->part, (int)nchunk, m->lossp, (int)nwave, 1.0f / (3.0f * (float)nb),
      m->grad, loss_dev);
  HIPCHK(ctx, hipGetLastError());
  return CG_OK;
}

// Glorot-uniform weights from splitmix64, zero biases.
void init_params(uint64_t seed, std::vector<float>& p) {
  static const int din[4] = {4, 64, 64, 64}, dout[4] = {64, 64, 64, 3};
  p.assign(mlp::NPARAM, 0.f);
  uint64_t st = seed;
  int o = 0;
  for (int l = 0; l < 4; ++l) {
    double a = std::sqrt(6.0 / (din[l] + dout[l]));
    for (int e = 0; e < din[l] * dout[l]; ++e) {
      double u = (double)(mlp::sm64_next(st) >> 11) * 0x1.0p-53;
      p[o++] = (float)((2.0 * u - 1.0) * a);
    }
    o += dout[l];
  }
}

}  // namespace

namespace cg_host {

int launch_satadj_run_mlp(int nmass, const KGrid& g, const KThermo& th,
                          const cg_mlp* net, double* u, double* w, int* status,
                          hipStream_t s) {
  if (nmass != 3) return CG_ERR_ARG;
  const long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
  mlp::k_satadj_run_mlp<3><<<nblk(ncell, 256), 256, 0, s>>>(g, th, net->tiles, net->P, u,
                                                            w, status);
  return CG_OK;
}

}  // namespace cg_host

extern "C" {

int cg_mlp_create(cg_ctx* ctx, unsigned long long seed, cg_mlp** out) {
  if (!ctx || !out) return CG_ERR_ARG;
  *out = nullptr;
  cg_mlp* m = new cg_mlp();
  m->ctx = ctx;
  auto fail = [&](int code) {
    cg_mlp_destroy(m);
    return code;
  };
  const size_t pb = sizeof(float) * mlp::NPARAM;
  if (hipMalloc((void**)&m->P, pb) != hipSuccess ||
      hipMalloc((void**)&m->m, pb) != hipSuccess ||
      hipMalloc((void**)&m->v, pb) != hipSuccess ||
      hipMalloc((void**)&m->grad, pb) != hipSuccess ||
      hipMalloc((void**)&m->loss, sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&m->tiles, sizeof(uint16_t) * mlp::NTILE * mlp::TILE) !=
          hipSuccess)
    return fail(set_err(ctx, CG_ERR_HIP, "mlp allocation failed"));
  m->loss_cap = 1;
  std::vector<float> p;
  init_params(seed, p);
  int e = cg_mlp_set_params(m, p.data());
  if (e) return fail(e);
  *out = m;
  return CG_OK;
}

int cg_mlp_destroy(cg_mlp* m) {
  if (!m) return CG_ERR_ARG;
  if (m->ctx) hipStreamSynchronize(m->ctx->stream);
  free_scratch(m);
  hipFree(m->P);
  hipFree(m->m);
  hipFree(m->v);
  hipFree(m->grad);
  hipFree(m->loss);
  hipFree(m->tiles);
  delete m;
  return CG_OK;
}

int cg_mlp_set_params(cg_mlp* m, const float* host_params) {
  if (!m || !host_params) return CG_ERR_ARG;
  cg_ctx* ctx = m->ctx;
  const size_t pb = sizeof(float) * mlp::NPARAM;
  HIPCHK(ctx, hipMemcpyAsync(m->P, host_params, pb, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(m->m, 0, pb, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(m->v, 0, pb, ctx->stream));
  m->adam_t = 0;
  int e = pack(m);
  if (e) return e;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_mlp_get_params(cg_mlp* m, float* host_params) {
  if (!m || !host_params) return CG_ERR_ARG;
  cg_ctx* ctx = m->ctx;
  HIPCHK(ctx, hipMemcpyAsync(host_params, m->P, sizeof(float) * mlp::NPARAM,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_mlp_forward(cg_mlp* m, long long n, const float* x, float* y) {
  if (!m || n < 0) return CG_ERR_ARG;
  if (n == 0) return CG_OK;
  if (!x || !y) return set_err(m->ctx, CG_ERR_ARG, "null x/y");
  cg_ctx* ctx = m->ctx;
  mlp::k_mlp_infer<<<nblk(n, 256), 256, 0, ctx->stream>>>(m->tiles, m->P, n, x, y);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_mlp_loss_grad(cg_mlp* m, long long n, const float* x, const float* t,
                     float* grad, double* loss_host) {
  if (!m || n <= 0 || !x || !t) return CG_ERR_ARG;
  cg_ctx* ctx = m->ctx;
  int e = gradient(m, n, n, x, t, 0, 0, m->loss);
  if (e) return e;
  if (grad)
    HIPCHK(ctx, hipMemcpyAsync(grad, m->grad, sizeof(float) * mlp::NPARAM,
                               hipMemcpyDeviceToDevice, ctx->stream));
  float l = 0.f;
  HIPCHK(ctx, hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (loss_host) *loss_host = l;
  return CG_OK;
}

int cg_mlp_train(cg_mlp* m, const cg_mlp_train_cfg* cfg, long long npairs,
                 const float* x, const float* t, float* loss_host) {
  if (!m || !cfg || npairs <= 0 || !x || !t) return CG_ERR_ARG;
  cg_ctx* ctx = m->ctx;
  if (cfg->batch <= 0 || cfg->batch > (1 << 20) || cfg->steps < 0)
    return set_err(ctx, CG_ERR_ARG, "batch must be in [1, 2^20], steps >= 0");
  if (cfg->steps == 0) return CG_OK;
  if (cfg->steps > m->loss_cap) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(m->loss);
    m->loss = nullptr;
    HIPCHK(ctx, hipMalloc((void**)&m->loss, sizeof(float) * cfg->steps));
    m->loss_cap = cfg->steps;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const double pi = 3.141592653589793;
  hipStream_t s = ctx->stream;
  for (int k = 0; k < cfg->steps; ++k) {
    uint64_t st = cfg->seed + (uint64_t)k * 0xD1B54A32D192ED03ull;
    uint64_t seed = mlp::sm64_next(st);
    int e = gradient(m, cfg->batch, npairs, x, t, seed, 1, m->loss + k);
    if (e) return e;
    m->adam_t += 1;
    double frac = cfg->steps > 1 ? (double)k / (cfg->steps - 1) : 1.0;
    double lr = cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(pi * frac));
    float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)m->adam_t)));
    float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)m->adam_t)));
    mlp::k_mlp_adam<<<nblk(mlp::NPARAM, 256), 256, 0, s>>>(
        m->P, m->m, m->v, m->grad, (float)lr, b1, b2, eps, c1, c2);
    e = pack(m);
    if (e) return e;
  }
  HIPCHK(ctx, hipGetLastError());
  if (loss_host)
    HIPCHK(ctx, hipMemcpyAsync(loss_host, m->loss, sizeof(float) * cfg->steps,
                               hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  return CG_OK;
}

int cg_mlp_make_pairs(cg_ctx* ctx, const cg_thermo* th, int nbase,
                      const double* rho, const double* ie, const double* y,
                      long long npairs, unsigned long long seed, float* x,
                      float* t, int* iters) {
  if (!ctx || !th || nbase <= 0 || npairs < 0) return CG_ERR_ARG;
  if (th->nmass != 3 || th->ivap <= 0 || th->icld <= 0)
    return set_err(ctx, CG_ERR_ARG, "pairs need nmass 3 and a vapour/cloud pair");
  if (npairs == 0) return CG_OK;
  KThermo kt = make_thermo(*th);
  mlp::k_mlp_pairs<3><<<nblk(npairs, 256), 256, 0, ctx->stream>>>(
      kt, nbase, rho, ie, y, npairs, seed, x, t, iters, ctx->d_status);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return check_status(ctx);
}

int cg_saturation_adjust(cg_ctx* ctx, const cg_thermo* th, int ncell,
                         const double* rho, double* ie, double* y,
                         const cg_mlp* net, int* iters) {
  if (!ctx || !th) return CG_ERR_ARG;
  if (th->nmass != 3 || th->ivap <= 0 || th->icld <= 0)
    return set_err(ctx, CG_ERR_ARG,
                   "saturation adjustment needs nmass 3 and a vapour/cloud pair");
  if (ncell <= 0) return CG_OK;
  KThermo kt = make_thermo(*th);
  if (net)
    mlp::k_satadj_mlp<3><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(
        kt, ncell, rho, ie, y, net->tiles, net->P, iters, ctx->d_status);
  else
    k_satadj<3><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(
        kt, ncell, rho, ie, y, nullptr, iters, ctx->d_status);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return check_status(ctx);
}

}  // extern "C"
