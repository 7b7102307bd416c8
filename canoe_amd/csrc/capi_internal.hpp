// capi_internal.hpp -- host-side pieces shared by the translation units that
// implement include/canoe_gpu.h: the context object, error reporting, the
// conversion of cg_thermo into the kernels' KThermo, and the device status
// word check.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/canoe_gpu.h"
#include "cg_physics.hpp"

struct cg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int* d_status = nullptr;  // device error flags (bit 0 pivot, bit 1 Newton)
  hipEvent_t t0 = nullptr, t1 = nullptr;  // cg_timer_start / cg_timer_stop
  // Device address of this device's fixed scratch array (the implicit
  // solve's entry point that allocates nothing); null if the lookup failed.
  double* fixed_scratch = nullptr;
  char err[512] = {0};
};

namespace cg_host {

inline int set_err(cg_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIPCHK(ctx, call)                                                    \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return set_err(ctx, CG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// Stream-ordered scratch of one entry point, released on every return path
// (a failed HIP call returns early through HIPCHK).
struct AsyncScratch {
  hipStream_t s;
  void* p = nullptr;
  explicit AsyncScratch(hipStream_t st) : s(st) {}
  ~AsyncScratch() {
    if (p) (void)hipFreeAsync(p, s);
  }
  AsyncScratch(const AsyncScratch&) = delete;
  AsyncScratch& operator=(const AsyncScratch&) = delete;
};

inline cg::KThermo make_thermo(const cg_thermo& t) {
  cg::KThermo k;
  memset(&k, 0, sizeof(k));
  k.nmass = t.nmass;
  k.nvapor = t.nvapor;
  k.gammad = t.gammad;
  k.rd = t.rd;
  for (int n = 0; n < 3; ++n) {
    k.cv_ratio_m1[n] = t.cv_ratio_m1[n];
    k.inv_mu_ratio_m1[n] = t.inv_mu_ratio_m1[n];
  }
  k.iv = (t.ivap > 0) ? t.ivap - 1 : -1;
  k.ic = (t.icld > 0) ? t.icld - 1 : -1;
  k.cvd = t.rd / (t.gammad - 1.);
  if (k.iv >= 0 && k.ic >= 0) {
    k.cvv = k.cvd * (1. + t.cv_ratio_m1[k.iv]);
    k.cvc = k.cvd * (1. + t.cv_ratio_m1[k.ic]);
    k.rv = t.rd * (1. + t.inv_mu_ratio_m1[k.iv]);
  }
  k.L = t.latent;
  k.t3 = t.t3;
  k.p3 = t.p3;
  k.beta = t.beta;
  k.delta = t.delta;
  k.tol = t.tol;
  k.max_iter = t.max_iter;
  return k;
}

inline unsigned nblk(long long n, int bs) {
  return (unsigned)((n + bs - 1) / bs);
}

// Reads and clears the device status word written by the kernels.
inline int check_status(cg_ctx* ctx) {
  int st = 0;
  HIPCHK(ctx, hipMemcpyAsync(&st, ctx->d_status, sizeof(int),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (st) {
    HIPCHK(ctx, hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (st & 4)
      return set_err(ctx, CG_ERR_NEGATIVE,
                     "negative reconstructed face density or pressure");
    if (st & 1) return set_err(ctx, CG_ERR_PIVOT, "zero pivot in implicit solve");
    if (st & 2)
      return set_err(ctx, CG_ERR_CONVERGE, "saturation Newton did not converge");
  }
  return CG_OK;
}

// Guided saturation adjustment of a run's resident interior cells (mlp.hip).
int launch_satadj_run_mlp(int nmass, const cg::KGrid& g, const cg::KThermo& th,
                          const cg_mlp* net, double* u, double* w, int* status,
                          hipStream_t s);

}  // namespace cg_host
