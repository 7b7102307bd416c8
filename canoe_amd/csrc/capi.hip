// capi.hip -- implementation of the C-ABI declared in include/canoe_gpu.h.
//
// Host side of the library: argument checks, kernel dispatch over the
// species count (NM = 0: dry, NM = 3: vapour/cloud/precipitation), the
// per-block run object with its device buffers, and the hipGraph capture of
// one full SSPRK(3,3) step.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/canoe_gpu.h"
#include "kernels.hpp"
#include "cg_timer.hpp"
#include "flux32.hpp"
#include "capi_internal.hpp"

using namespace cg;
using namespace cg_host;

namespace {

int check_block(cg_ctx* ctx, const cg_block* b, const cg_thermo* th) {
  if (!b || !th) return set_err(ctx, CG_ERR_ARG, "null block/thermo");
  if (b->nx1 < 1 || b->nx2 < 1 || b->nx3 < 1)
    return set_err(ctx, CG_ERR_ARG, "bad block size");
  if (b->nmass != th->nmass)
    return set_err(ctx, CG_ERR_ARG, "block nmass %d != thermo nmass %d",
                   b->nmass, th->nmass);
  if (b->nmass != 0 && b->nmass != 3)
    return set_err(ctx, CG_ERR_ARG, "nmass must be 0 or 3 (got %d)", b->nmass);
  if (b->nx3 > 1 && b->nx2 == 1)
    return set_err(ctx, CG_ERR_ARG, "nx3 > 1 requires nx2 > 1");
  if (b->dx1 <= 0 || b->dx2 <= 0 || b->dx3 <= 0)
    return set_err(ctx, CG_ERR_ARG, "non-positive spacing");
  for (int s = 0; s < 6; ++s)
    if (b->bc[s] < 0 || b->bc[s] > 2)
      return set_err(ctx, CG_ERR_ARG, "bad boundary flag");
  return CG_OK;
}

// Grid of a block in the AthenaArray layout (lvl = 0, the C-ABI's caller
// arrays) or in the level-major layout of a run's resident state (lvl = 1).
KGrid make_grid(const cg_block& b, int lvl) {
  KGrid g;
  g.nx1 = b.nx1;
  g.nx2 = b.nx2;
  g.nx3 = b.nx3;
  g.f2 = b.nx2 > 1;
  g.f3 = b.nx3 > 1;
  g.nc1 = b.nx1 + 2 * NG;
  g.nc2 = g.f2 ? b.nx2 + 2 * NG : 1;
  g.nc3 = g.f3 ? b.nx3 + 2 * NG : 1;
  g.is = NG;
  g.ie = NG + b.nx1 - 1;
  g.js = g.f2 ? NG : 0;
  g.je = g.js + b.nx2 - 1;
  g.ks = g.f3 ? NG : 0;
  g.ke = g.ks + b.nx3 - 1;
  g.lvl = lvl;
  const long long c1 = g.nc1, c2 = g.nc2, c3 = g.nc3;
  g.sn = c1 * c2 * c3;
  long long ncol = (long long)b.nx2 * b.nx3;
  // extents of cells (e1, e2, e3) -> strides
  auto mk = [&](long long e1, long long e2, long long e3) {
    Strides st;
    if (lvl) {
      st.j = 1;
      st.k = e2;
      st.i = e2 * e3;
    } else {
      st.i = 1;
      st.j = e1;
      st.k = e1 * e2;
    }
    st.n = e1 * e2 * e3;
    return st;
  };
  g.c = mk(c1, c2, c3);
  g.f[0] = mk(c1 + 1, c2, c3);
  g.f[1] = mk(c1, c2 + 1, c3);
  g.f[2] = mk(c1, c2, c3 + 1);
  if (lvl) {
    g.ps_i = ncol;
    g.ps_c = 1;
  } else {
    g.ps_i = 1;
    g.ps_c = c1 + 1;
  }
  g.dx1 = b.dx1;
  g.dx2 = b.dx2;
  g.dx3 = b.dx3;
  g.grav = b.grav;
  for (int s = 0; s < 6; ++s) g.bc[s] = b.bc[s];
  return g;
}

size_t psf_doubles(const KGrid& g) {
  return (size_t)(g.nc1 + 1) * g.nx2 * g.nx3;
}

}  // namespace

namespace {

size_t cell_doubles(const KGrid& g, int N) { return (size_t)N * g.sn; }
size_t f1_doubles(const KGrid& g, int N) {
  return (size_t)N * g.nc3 * g.nc2 * (g.nc1 + 1);
}
size_t f2_doubles(const KGrid& g, int N) {
  return (size_t)N * g.nc3 * (g.nc2 + 1) * g.nc1;
}
size_t f3_doubles(const KGrid& g, int N) {
  return (size_t)N * (g.nc3 + 1) * g.nc2 * g.nc1;
}

// ---- launch helpers (templated on the species count) ----
template <int NM>
void launch_fluxes(const KGrid& g, const KThermo& th, const cg_config& cfg,
                   const double* w, double* psf, double* psv, double* f1,
                   double* f2, double* f3, hipStream_t s, KTimer* tm = nullptr,
                   bool scan = true) {
  const long long ncol = (long long)g.nx2 * g.nx3;
  const int decomp = g.grav != 0.0;
  if (decomp && scan)
    TIMED(tm, KID_SCAN,
          (k_hydrostatic_scan<NM><<<nblk(ncol, 256), 256, 0, s>>>(g, th, w, psf, psv)));
  long long n1 = ncol * (g.nx1 + 1);
  if (cfg.order_v == 5)
    TIMED(tm, KID_FLUX_X1,
          (k_flux_x1<NM, 5><<<nblk(n1, 256), 256, 0, s>>>(g, th, cfg.rsolver, decomp,
                                                           w, psf, psv, f1)));
  else
    TIMED(tm, KID_FLUX_X1,
          (k_flux_x1<NM, 2><<<nblk(n1, 256), 256, 0, s>>>(g, th, cfg.rsolver, decomp,
                                                           w, psf, psv, f1)));
  if (g.f2) {
    long long n2 = (long long)g.nx1 * (g.nx2 + 1) * g.nx3;
    if (cfg.order_h == 5)
      TIMED(tm, KID_FLUX_X2,
            (k_flux_h<NM, 5, 1><<<nblk(n2, 256), 256, 0, s>>>(g, th, cfg.rsolver, w, f2)));
    else
      TIMED(tm, KID_FLUX_X2,
            (k_flux_h<NM, 2, 1><<<nblk(n2, 256), 256, 0, s>>>(g, th, cfg.rsolver, w, f2)));
  }
  if (g.f3) {
    long long n3 = (long long)g.nx1 * g.nx2 * (g.nx3 + 1);
    if (cfg.order_h == 5)
      TIMED(tm, KID_FLUX_X3,
            (k_flux_h<NM, 5, 2><<<nblk(n3, 256), 256, 0, s>>>(g, th, cfg.rsolver, w, f3)));
    else
      TIMED(tm, KID_FLUX_X3,
            (k_flux_h<NM, 2, 2><<<nblk(n3, 256), 256, 0, s>>>(g, th, cfg.rsolver, w, f3)));
  }
}

template <int NM>
void launch_implicit(const KGrid& g, const KThermo& th, double dt, int flag,
                     double* du, const double* w, double* arch, int* status,
                     hipStream_t s, KTimer* tm = nullptr) {
  KImplicit ip;
  ip.dt = dt;
  ip.ncol = (long long)g.nx2 * g.nx3;
  unsigned nb = nblk(ip.ncol, 64);
  if (flag & 8) {
    TIMED(tm, KID_IMPLICIT_FWD,
          (k_implicit_fwd<NM, 5><<<nb, 64, 0, s>>>(g, th, ip, w, du, arch, status)));
    TIMED(tm, KID_IMPLICIT_BWD,
          (k_implicit_bwd<NM, 5><<<nb, 64, 0, s>>>(g, ip, w, arch, du)));
  } else {
    TIMED(tm, KID_IMPLICIT_FWD,
          (k_implicit_fwd<NM, 3><<<nb, 64, 0, s>>>(g, th, ip, w, du, arch, status)));
    TIMED(tm, KID_IMPLICIT_BWD,
          (k_implicit_bwd<NM, 3><<<nb, 64, 0, s>>>(g, ip, w, arch, du)));
  }
}

template <int NM>
void launch_bc(const KGrid& g, double* w, hipStream_t s, KTimer* tm = nullptr) {
  TIMED(tm, KID_BC_X1,
        (k_bc_x1<NM><<<nblk((long long)g.nx2 * g.nx3 * NG, 256), 256, 0, s>>>(g, w)));
  if (g.f2)
    TIMED(tm, KID_BC_X2,
          (k_bc_x2<NM><<<nblk((long long)g.nc1 * g.nx3 * NG, 256), 256, 0, s>>>(g, w)));
  if (g.f3)
    TIMED(tm, KID_BC_X3,
          (k_bc_x3<NM><<<nblk((long long)g.nc1 * g.nc2 * NG, 256), 256, 0, s>>>(g, w)));
}

// Forward sweep + fused downward sweep of a run (back substitution, stage
// update, x1 ghosts, next stage's hydrostatic pressures).
template <int NM>
void launch_implicit_update(const KGrid& g, const KThermo& th, double dt,
                            int flag, const KUpdate& up, const double* du,
                            double* u, double* u1, double* w, double* arch,
                            double* psf, double* psv, int* status,
                            hipStream_t s, KTimer* tm) {
  KImplicit ip;
  ip.dt = dt;
  ip.ncol = (long long)g.nx2 * g.nx3;
  unsigned nb = nblk(ip.ncol, 64);
  if (flag & 8) {
    TIMED(tm, KID_IMPLICIT_FWD,
          (k_implicit_fwd<NM, 5><<<nb, 64, 0, s>>>(g, th, ip, w, du, arch, status)));
    TIMED(tm, KID_IMPLICIT_BWD,
          (k_implicit_bwd_update<NM, 5><<<nb, 64, 0, s>>>(
              g, th, ip, up, arch, du, u, u1, w, psf, psv, status)));
  } else {
    TIMED(tm, KID_IMPLICIT_FWD,
          (k_implicit_fwd<NM, 3><<<nb, 64, 0, s>>>(g, th, ip, w, du, arch, status)));
    TIMED(tm, KID_IMPLICIT_BWD,
          (k_implicit_bwd_update<NM, 3><<<nb, 64, 0, s>>>(
              g, th, ip, up, arch, du, u, u1, w, psf, psv, status)));
  }
}

size_t arch_doubles(const KGrid& g) {
  return (size_t)g.nx1 * 30 * (size_t)g.nx2 * g.nx3;
}

}  // namespace

// ===========================================================================
struct cg_run {
  cg_ctx* ctx = nullptr;
  cg_block blk;
  cg_thermo th;
  cg_config cfg;
  KGrid g;   // level-major resident state
  KGrid ga;  // AthenaArray layout of host transfers
  KThermo kt;
  int N = 5;
  double *u = nullptr, *w = nullptr, *u1 = nullptr, *du = nullptr;
  double *f1 = nullptr, *f2 = nullptr, *f3 = nullptr;
  double *psf = nullptr, *psv = nullptr, *arch = nullptr;
  float* ff = nullptr;     // FP32 face fluxes of the mixed form
  double* pref = nullptr;  // per-level reference pressure of the mixed form
  double* part = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double graph_dt = 0.0;
  KTimer tm;
};

namespace {

template <int NM>
int run_stage(cg_run* r, double dt, int s) {
  const double sbeta[3] = {1.0, 0.25, 2. / 3.};
  const double sg1[3] = {0.0, 0.25, 2. / 3.};
  const double sg2[3] = {1.0, 0.75, 1. / 3.};
  hipStream_t st = r->ctx->stream;
  const KGrid& g = r->g;
  const long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
  const double wght = sbeta[s] * dt;
  KSource src;
  src.wght = wght;
  src.visc = r->cfg.visc;
  src.p0 = r->cfg.p0;
  src.cp = r->th.gammad / (r->th.gammad - 1.) * r->th.rd;
  src.chi = (r->th.gammad - 1.) / r->th.gammad;
  src.heat_flux = r->cfg.heat_flux;
  KTimer* tm = r->tm.on ? &r->tm : nullptr;
  const bool imp = r->cfg.implicit_flag != 0;
  KUpdate up;
  up.stage = s;
  up.g1 = sg1[s];
  up.g2 = sg2[s];
  up.satadj = (s == 2) && (r->cfg.satadj != CG_SATADJ_OFF);
  if (r->cfg.precision == CG_PREC_FP32) {
    launch_fluxes32<NM>(g, r->kt, r->cfg, src, r->w, r->ff, r->psf, r->psv,
                        r->pref, r->du, st, tm, !imp);
  } else {
    // implicit runs get psf / psv from the previous downward sweep
    launch_fluxes<NM>(g, r->kt, r->cfg, r->w, r->psf, r->psv, r->f1, r->f2,
                      r->f3, st, tm, !imp);
    TIMED(tm, KID_DIVERGENCE,
          (k_divergence<NM><<<nblk(ncell, 256), 256, 0, st>>>(
              g, r->kt, src, r->f1, r->f2, r->f3, r->w, r->du)));
  }
  if (imp) {
    launch_implicit_update<NM>(g, r->kt, wght, r->cfg.implicit_flag, up, r->du,
                               r->u, r->u1, r->w, r->arch, r->psf, r->psv,
                               r->ctx->d_status, st, tm);
    if (g.f2)
      TIMED(tm, KID_BC_X2,
            (k_bc_x2<NM><<<nblk((long long)g.nc1 * g.nx3 * NG, 256), 256, 0, st>>>(g, r->w)));
    if (g.f3)
      TIMED(tm, KID_BC_X3,
            (k_bc_x3<NM><<<nblk((long long)g.nc1 * g.nc2 * NG, 256), 256, 0, st>>>(g, r->w)));
  } else {
    TIMED(tm, KID_UPDATE,
          (k_update<NM><<<nblk(ncell, 256), 256, 0, st>>>(g, r->kt, up, r->u, r->u1,
                                                          r->du, r->w, r->ctx->d_status)));
    launch_bc<NM>(g, r->w, st, tm);
  }
  return CG_OK;
}

int run_one_step(cg_run* r, double dt) {
  for (int s = 0; s < 3; ++s) {
    int e = (r->N == 5) ? run_stage<0>(r, dt, s) : run_stage<3>(r, dt, s);
    if (e) return e;
  }
  return CG_OK;
}

}  // namespace

extern "C" {

const char* cg_version(void) {
  return "canoe_gpu gfx950 fp64/fp32 -ffp-contract=off";
}

int cg_create(int device, cg_ctx** out) {
  if (!out) return CG_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CG_ERR_NODEV;
  if (device < 0 || device >= n) return CG_ERR_ARG;
  cg_ctx* c = new cg_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->d_status, sizeof(int)) != hipSuccess ||
      hipMemset(c->d_status, 0, sizeof(int)) != hipSuccess) {
    delete c;
    return CG_ERR_HIP;
  }
  *out = c;
  return CG_OK;
}

int cg_destroy(cg_ctx* ctx) {
  if (!ctx) return CG_ERR_ARG;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->d_status) hipFree(ctx->d_status);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return CG_OK;
}

int cg_synchronize(cg_ctx* ctx) {
  if (!ctx) return CG_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return check_status(ctx);
}

const char* cg_last_error(cg_ctx* ctx) { return ctx ? ctx->err : "no context"; }

int cg_malloc(cg_ctx* ctx, size_t bytes, void** ptr) {
  if (!ctx || !ptr) return CG_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipMalloc(ptr, bytes));
  return CG_OK;
}
int cg_free(cg_ctx* ctx, void* ptr) {
  if (!ctx) return CG_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipFree(ptr));
  return CG_OK;
}
int cg_memcpy_h2d(cg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return CG_ERR_ARG;
  HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}
int cg_memcpy_d2h(cg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return CG_ERR_ARG;
  HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_riemann(cg_ctx* ctx, const cg_thermo* th, int solver, int ivx,
               int n_faces, const double* wl, const double* wr, ptrdiff_t ld,
               double* flx, ptrdiff_t flx_ld) {
  if (!ctx || !th) return CG_ERR_ARG;
  if (th->nmass != 0 && th->nmass != 3)
    return set_err(ctx, CG_ERR_ARG, "nmass must be 0 or 3");
  int IVX = 1 + th->nmass;
  int dir = ivx - IVX;
  if (dir < 0 || dir > 2) return set_err(ctx, CG_ERR_ARG, "bad ivx %d", ivx);
  if (solver != CG_RS_LMARS && solver != CG_RS_HLLC)
    return set_err(ctx, CG_ERR_ARG, "bad solver");
  if (n_faces <= 0) return CG_OK;
  if (ld < n_faces || flx_ld < n_faces) return set_err(ctx, CG_ERR_ARG, "ld < n_faces");
  KThermo kt = make_thermo(*th);
  unsigned nb = nblk(n_faces, 256);
  hipStream_t s = ctx->stream;
#define RL(NM, D) \
  k_riemann_faces<NM, D><<<nb, 256, 0, s>>>(kt, solver, n_faces, wl, wr, ld, flx, flx_ld)
  if (th->nmass == 0) {
    if (dir == 0) RL(0, 0); else if (dir == 1) RL(0, 1); else RL(0, 2);
  } else {
    if (dir == 0) RL(3, 0); else if (dir == 1) RL(3, 1); else RL(3, 2);
  }
#undef RL
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(s));
  return CG_OK;
}

int cg_calculate_fluxes(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                        const cg_config* cfg, const double* w, double* flux1,
                        double* flux2, double* flux3) {
  if (!ctx || !cfg) return CG_ERR_ARG;
  int e = check_block(ctx, blk, th);
  if (e) return e;
  if ((cfg->order_h != 2 && cfg->order_h != 5) ||
      (cfg->order_v != 2 && cfg->order_v != 5))
    return set_err(ctx, CG_ERR_ARG, "order must be 2 or 5");
  KGrid g = make_grid(*blk, 0);
  KThermo kt = make_thermo(*th);
  double *psf = nullptr, *psv = nullptr;
  long long ncol = (long long)g.nx2 * g.nx3;
  HIPCHK(ctx, hipMallocAsync((void**)&psf, sizeof(double) * ncol * (g.nc1 + 1), ctx->stream));
  HIPCHK(ctx, hipMallocAsync((void**)&psv, sizeof(double) * ncol * (g.nc1 + 1), ctx->stream));
  if (blk->nmass == 0)
    launch_fluxes<0>(g, kt, *cfg, w, psf, psv, flux1, flux2, flux3, ctx->stream);
  else
    launch_fluxes<3>(g, kt, *cfg, w, psf, psv, flux1, flux2, flux3, ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipFreeAsync(psf, ctx->stream));
  HIPCHK(ctx, hipFreeAsync(psv, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_add_flux_divergence(cg_ctx* ctx, const cg_block* blk, double wght,
                           const double* flux1, const double* flux2,
                           const double* flux3, double* du) {
  if (!ctx || !blk) return CG_ERR_ARG;
  KGrid g = make_grid(*blk, 0);
  long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
  if (blk->nmass == 0)
    k_add_divergence<0><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(g, wght, flux1, flux2, flux3, du);
  else if (blk->nmass == 3)
    k_add_divergence<3><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(g, wght, flux1, flux2, flux3, du);
  else
    return set_err(ctx, CG_ERR_ARG, "nmass must be 0 or 3");
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_implicit_solve(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                      double* du, const double* w, double dt, int flag) {
  if (!ctx) return CG_ERR_ARG;
  int e = check_block(ctx, blk, th);
  if (e) return e;
  if (flag == 0) return CG_OK;
  if (flag & 6) return set_err(ctx, CG_ERR_ARG, "only x1 implicit (flag 1 or 9) is supported");
  if (blk->bc[0] == CG_BC_PERIODIC || blk->bc[1] == CG_BC_PERIODIC)
    return set_err(ctx, CG_ERR_ARG, "periodic x1 implicit solve not supported");
  if (!(dt > 0.0)) return set_err(ctx, CG_ERR_ARG, "dt must be positive");
  KGrid g = make_grid(*blk, 0);
  KThermo kt = make_thermo(*th);
  double* arch = nullptr;
  HIPCHK(ctx, hipMallocAsync((void**)&arch, sizeof(double) * arch_doubles(g), ctx->stream));
  if (blk->nmass == 0)
    launch_implicit<0>(g, kt, dt, flag, du, w, arch, ctx->d_status, ctx->stream);
  else
    launch_implicit<3>(g, kt, dt, flag, du, w, arch, ctx->d_status, ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipFreeAsync(arch, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return check_status(ctx);
}

int cg_cons_to_prim(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                    const double* u, double* w) {
  if (!ctx) return CG_ERR_ARG;
  int e = check_block(ctx, blk, th);
  if (e) return e;
  KGrid g = make_grid(*blk, 0);
  KThermo kt = make_thermo(*th);
  long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
  if (blk->nmass == 0)
    k_eos<0, 1><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(g, kt, u, w);
  else
    k_eos<3, 1><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(g, kt, u, w);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_prim_to_cons(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                    const double* w, double* u) {
  if (!ctx) return CG_ERR_ARG;
  int e = check_block(ctx, blk, th);
  if (e) return e;
  KGrid g = make_grid(*blk, 0);
  KThermo kt = make_thermo(*th);
  long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
  if (blk->nmass == 0)
    k_eos<0, 0><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(g, kt, w, u);
  else
    k_eos<3, 0><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(g, kt, w, u);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_apply_bc(cg_ctx* ctx, const cg_block* blk, double* w) {
  if (!ctx || !blk) return CG_ERR_ARG;
  KGrid g = make_grid(*blk, 0);
  if (blk->nmass == 0)
    launch_bc<0>(g, w, ctx->stream);
  else if (blk->nmass == 3)
    launch_bc<3>(g, w, ctx->stream);
  else
    return set_err(ctx, CG_ERR_ARG, "nmass must be 0 or 3");
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

// ---------------------------------------------------------------------------
int cg_run_create(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                  const cg_config* cfg, cg_run** out) {
  if (!ctx || !cfg || !out) return CG_ERR_ARG;
  *out = nullptr;
  int e = check_block(ctx, blk, th);
  if (e) return e;
  if ((cfg->order_h != 2 && cfg->order_h != 5) ||
      (cfg->order_v != 2 && cfg->order_v != 5))
    return set_err(ctx, CG_ERR_ARG, "order must be 2 or 5");
  if (cfg->implicit_flag != 0 && cfg->implicit_flag != 1 && cfg->implicit_flag != 9)
    return set_err(ctx, CG_ERR_ARG, "implicit_flag must be 0, 1 or 9");
  if (cfg->implicit_flag && (blk->bc[0] == CG_BC_PERIODIC || blk->bc[1] == CG_BC_PERIODIC))
    return set_err(ctx, CG_ERR_ARG, "periodic x1 implicit solve not supported");
  if (cfg->satadj != CG_SATADJ_OFF && (th->nmass != 3 || th->ivap <= 0 || th->icld <= 0))
    return set_err(ctx, CG_ERR_ARG, "saturation adjustment needs a vapour/cloud pair");
  if (cfg->precision != CG_PREC_FP64 && cfg->precision != CG_PREC_FP32)
    return set_err(ctx, CG_ERR_ARG, "bad precision");
  cg_run* r = new cg_run();
  r->ctx = ctx;
  r->blk = *blk;
  r->th = *th;
  r->cfg = *cfg;
  r->g = make_grid(*blk, 1);
  r->ga = make_grid(*blk, 0);
  r->kt = make_thermo(*th);
  r->N = 5 + blk->nmass;
  const KGrid& g = r->g;
  size_t nc = cell_doubles(g, r->N);
  long long ncol = (long long)g.nx2 * g.nx3;
  auto alloc = [&](double** p, size_t n) {
    return hipMalloc((void**)p, sizeof(double) * (n > 0 ? n : 1)) == hipSuccess;
  };
  bool ok = alloc(&r->u, nc) && alloc(&r->w, nc) && alloc(&r->u1, nc) &&
            alloc(&r->du, nc) && alloc(&r->psf, psf_doubles(g)) &&
            alloc(&r->psv, psf_doubles(g)) && alloc(&r->part, 1024 * 8);
  if (ok && cfg->precision == CG_PREC_FP64)
    ok = alloc(&r->f1, f1_doubles(g, r->N)) && alloc(&r->f2, f2_doubles(g, r->N)) &&
         alloc(&r->f3, f3_doubles(g, r->N));
  if (ok && cfg->precision == CG_PREC_FP32)
    ok = hipMalloc((void**)&r->ff, sizeof(float) * (f1_doubles(g, r->N) +
                                                    f2_doubles(g, r->N) +
                                                    f3_doubles(g, r->N))) ==
             hipSuccess &&
         alloc(&r->pref, g.nc1);
  if (ok && cfg->implicit_flag) ok = alloc(&r->arch, arch_doubles(g));
  if (ok) {
    ok = hipMemset(r->u, 0, sizeof(double) * nc) == hipSuccess &&
         hipMemset(r->w, 0, sizeof(double) * nc) == hipSuccess &&
         hipMemset(r->u1, 0, sizeof(double) * nc) == hipSuccess &&
         hipMemset(r->du, 0, sizeof(double) * nc) == hipSuccess;
  }
  r->tm.s = ctx->stream;
  if (ok) ok = hipEventCreate(&r->tm.e0) == hipSuccess && hipEventCreate(&r->tm.e1) == hipSuccess;
  if (!ok) {
    cg_run_destroy(r);
    return set_err(ctx, CG_ERR_HIP, "allocation failed");
  }
  *out = r;
  return CG_OK;
}

int cg_run_destroy(cg_run* r) {
  if (!r) return CG_ERR_ARG;
  hipStreamSynchronize(r->ctx->stream);
  if (r->gexec) hipGraphExecDestroy(r->gexec);
  if (r->graph) hipGraphDestroy(r->graph);
  double* ps[] = {r->u, r->w, r->u1, r->du, r->f1, r->f2, r->f3, r->psf, r->psv,
                  r->arch, r->part};
  for (double* p : ps)
    if (p) hipFree(p);
  if (r->ff) hipFree(r->ff);
  if (r->pref) hipFree(r->pref);
  if (r->tm.e0) hipEventDestroy(r->tm.e0);
  if (r->tm.e1) hipEventDestroy(r->tm.e1);
  delete r;
  return CG_OK;
}

int cg_run_set_prim_host(cg_run* r, const double* w_host) {
  if (!r || !w_host) return CG_ERR_ARG;
  cg_ctx* ctx = r->ctx;
  size_t nc = cell_doubles(r->g, r->N);
  HIPCHK(ctx, hipMemcpyAsync(r->du, w_host, sizeof(double) * nc,
                             hipMemcpyHostToDevice, ctx->stream));
  k_relayout<<<nblk((long long)nc, 256), 256, 0, ctx->stream>>>(r->ga, r->g, r->N,
                                                               r->du, r->w);
  long long ncell = (long long)r->g.nx1 * r->g.nx2 * r->g.nx3;
  long long ncol = (long long)r->g.nx2 * r->g.nx3;
  if (r->N == 5) {
    k_eos<0, 0><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(r->g, r->kt, r->w, r->u);
    launch_bc<0>(r->g, r->w, ctx->stream);
    if (r->pref)
      k_level_mean_p<0><<<r->g.nc1, 256, 0, ctx->stream>>>(r->g, r->w, r->pref);
    if (r->g.grav != 0.0)
      k_hydrostatic_scan<0><<<nblk(ncol, 256), 256, 0, ctx->stream>>>(r->g, r->kt, r->w,
                                                                     r->psf, r->psv);
  } else {
    k_eos<3, 0><<<nblk(ncell, 256), 256, 0, ctx->stream>>>(r->g, r->kt, r->w, r->u);
    launch_bc<3>(r->g, r->w, ctx->stream);
    if (r->pref)
      k_level_mean_p<3><<<r->g.nc1, 256, 0, ctx->stream>>>(r->g, r->w, r->pref);
    if (r->g.grav != 0.0)
      k_hydrostatic_scan<3><<<nblk(ncol, 256), 256, 0, ctx->stream>>>(r->g, r->kt, r->w,
                                                                     r->psf, r->psv);
  }
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

static int run_download(cg_run* r, const double* src, double* host) {
  cg_ctx* ctx = r->ctx;
  size_t nc = cell_doubles(r->g, r->N);
  double* tmp = nullptr;
  HIPCHK(ctx, hipMallocAsync((void**)&tmp, sizeof(double) * nc, ctx->stream));
  k_relayout<<<nblk((long long)nc, 256), 256, 0, ctx->stream>>>(r->g, r->ga, r->N,
                                                               src, tmp);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(host, tmp, sizeof(double) * nc, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipFreeAsync(tmp, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_run_get_cons_host(cg_run* r, double* u_host) {
  if (!r || !u_host) return CG_ERR_ARG;
  return run_download(r, r->u, u_host);
}

int cg_run_get_prim_host(cg_run* r, double* w_host) {
  if (!r || !w_host) return CG_ERR_ARG;
  return run_download(r, r->w, w_host);
}

int cg_run_device_state(cg_run* r, double** u, double** w) {
  if (!r) return CG_ERR_ARG;
  if (u) *u = r->u;
  if (w) *w = r->w;
  return CG_OK;
}

int cg_run_set_timing(cg_run* r, int on) {
  if (!r) return CG_ERR_ARG;
  r->tm.on = on != 0;
  for (int q = 0; q < KID_COUNT; ++q) {
    r->tm.ms[q] = 0.0;
    r->tm.cnt[q] = 0;
  }
  return CG_OK;
}

int cg_run_kernel_stats(cg_run* r, int kernel_id, double* total_ms,
                        long long* launches, const char** name) {
  if (!r || kernel_id < 0 || kernel_id >= KID_COUNT) return CG_ERR_ARG;
  if (total_ms) *total_ms = r->tm.ms[kernel_id];
  if (launches) *launches = r->tm.cnt[kernel_id];
  if (name) *name = kKernelNames[kernel_id];
  return CG_OK;
}

int cg_run_step(cg_run* r, double dt, int nsteps, int use_graph) {
  if (!r) return CG_ERR_ARG;
  if (!(dt > 0.0) || nsteps < 0) return set_err(r->ctx, CG_ERR_ARG, "bad dt/nsteps");
  cg_ctx* ctx = r->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (use_graph && !r->tm.on) {
    if (!r->gexec || r->graph_dt != dt) {
      if (r->gexec) {
        hipGraphExecDestroy(r->gexec);
        r->gexec = nullptr;
      }
      if (r->graph) {
        hipGraphDestroy(r->graph);
        r->graph = nullptr;
      }
      HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int e = run_one_step(r, dt);
      hipGraph_t gr = nullptr;
      hipError_t ce = hipStreamEndCapture(ctx->stream, &gr);
      if (e) {
        if (gr) hipGraphDestroy(gr);
        return e;
      }
      if (ce != hipSuccess)
        return set_err(ctx, CG_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
      r->graph = gr;
      HIPCHK(ctx, hipGraphInstantiate(&r->gexec, r->graph, nullptr, nullptr, 0));
      r->graph_dt = dt;
    }
    for (int n = 0; n < nsteps; ++n) HIPCHK(ctx, hipGraphLaunch(r->gexec, ctx->stream));
  } else {
    for (int n = 0; n < nsteps; ++n) {
      int e = run_one_step(r, dt);
      if (e) return e;
    }
  }
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return check_status(ctx);
}

int cg_run_sums(cg_run* r, double* out_host) {
  if (!r || !out_host) return CG_ERR_ARG;
  cg_ctx* ctx = r->ctx;
  const int nb = 1024;
  if (r->N == 5)
    k_sums<0><<<nb, 256, 0, ctx->stream>>>(r->g, r->u, r->part);
  else
    k_sums<3><<<nb, 256, 0, ctx->stream>>>(r->g, r->u, r->part);
  HIPCHK(ctx, hipGetLastError());
  std::vector<double> h((size_t)r->N * nb);
  HIPCHK(ctx, hipMemcpyAsync(h.data(), r->part, sizeof(double) * h.size(),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int n = 0; n < r->N; ++n) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += h[(size_t)n * nb + b];
    out_host[n] = s;
  }
  return CG_OK;
}

}  // extern "C"
