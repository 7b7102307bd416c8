// kessler.hpp -- Kessler warm-rain microphysics and sedimentation of the
// precipitation tracer, one column per lane, swept top down.
//
// No counterpart in the reference's live stage (its sedimentation hook is
// commented out at src/tasklist/implicit_hydro_tasks.cpp:292; the rates live
// in kintera, which is not vendored).  Reactions and rate constants:
// examples/2019-Li-snap/earth-precip.yaml:44-53, water-kessler.yaml:4-18;
// the sedimentation mass flux (velocity times upwind density) of
// src/snap/riemann/noflux.cpp:61-71.  Per cell, with the full step dt:
//   a = min(rho_c, dt (k_auto + k_accr q_p) rho_c)        cloud -> precip
//   b = min(rho_p, rho_s - rho_v, dt k_evap rho_p (1 - rho_v / rho_s))
//                                  precip -> vapour while sub-saturated,
// E -= L b (the moist internal energy ie - rho L (q_c + q_p) is kept), then
// a fraction c = min(1, dt v_fall / dx1) of the cell's precipitation falls
// into the cell below with the cell's velocity and its own sensible energy
// cv_p T + |v|^2 / 2; nothing crosses the walls.  The per-cell step,
// kessler_cell, lives in kernels.hpp: the implicit downward sweep applies it
// as well, in the last stage of runs whose x1 sides are walls or outflow.
#pragma once
#include "cg_physics.hpp"
#include "kernels.hpp"

namespace cg {

// One column (k, j): u (conserved) in / out, w (primitive) out, interior
// cells only.  dt is the full step.  FUSE (reflecting or outflow x1 sides):
// the column's x1 ghosts of w (those of k_bc_x1) and its hydrostatic face
// and cell pressures psf / psv at column offset pc (those of
// k_hydrostatic_scan) as each new cell state appears, top down.
template <int NM, bool FUSE = false>
__device__ __forceinline__ void kessler_column(const KGrid& g, const KThermo& th,
                                               const KKessler& kz, double dt,
                                               double* __restrict__ u,
                                               double* __restrict__ w, int k, int j,
                                               double* __restrict__ psf = nullptr,
                                               double* __restrict__ psv = nullptr,
                                               long long pc = 0) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IEN = 4 + NM;
  const double cvp = th.cvd * (1. + th.cv_ratio_m1[kz.ip]);
  double cf = dt * kz.v_fall / g.dx1;
  if (cf > 1.) cf = 1.;
  const long long sn = g.c.n, si = g.c.i;
  const long long cb = g.at(0, k, j, 0);
  const int is = g.is, ie = g.ie;
  const double grav = g.grav, dx = g.dx1;
  double pf_next = 0., pf_top = 0.;
  double fin[5] = {0., 0., 0., 0., 0.};
  // The next level's u is loaded a level ahead, and a level's stores are
  // issued at the next level, together with that load, once the level's own
  // data have been used: a wait for loaded data also waits for every store
  // issued before it, so stores at the end of a level would expose their
  // write latency at the start of the next.
  struct Out {
    int i;
    double uc[N], wc[N];
    double pfi, pvi, ptop;
  };
  Out pend;
  auto flush = [&](const Out& o) {
    const long long c0 = g.at(0, k, j, o.i);
#pragma unroll
    for (int n = 0; n < N; ++n) {
      u[c0 + n * sn] = o.uc[n];
      w[c0 + n * sn] = o.wc[n];
    }
    if (FUSE) {
      // x1 ghosts (bc_pair: reflecting sides mirror, normal velocity
      // negated; outflow sides copy the boundary cell)
      const int mt = ie - o.i, mb = o.i - is;
      if (mt < NG && g.bc[1] == 0) {
#pragma unroll
        for (int n = 0; n < N; ++n)
          w[cb + n * sn + (ie + 1 + mt) * si] = ((n == IVX) ? -1.0 : 1.0) * o.wc[n];
      }
      if (mt == 0 && g.bc[1] == 2) {
        for (int gh = 1; gh <= NG; ++gh)
#pragma unroll
          for (int n = 0; n < N; ++n) w[cb + n * sn + (ie + gh) * si] = o.wc[n];
      }
      if (mb < NG && g.bc[0] == 0) {
#pragma unroll
        for (int n = 0; n < N; ++n)
          w[cb + n * sn + (is - 1 - mb) * si] = ((n == IVX) ? -1.0 : 1.0) * o.wc[n];
      }
      if (mb == 0 && g.bc[0] == 2) {
        for (int gh = 1; gh <= NG; ++gh)
#pragma unroll
          for (int n = 0; n < N; ++n) w[cb + n * sn + (is - gh) * si] = o.wc[n];
      }
      if (grav != 0.0) {
        if (o.i == ie) psf[pc + (ie + 1) * g.ps_i] = o.ptop;
        psf[pc + o.i * g.ps_i] = o.pfi;
        psv[pc + o.i * g.ps_i] = o.pvi;
      }
    }
  };
  double un[N];
  auto load_u = [&](int i) {
    const long long c0 = g.at(0, k, j, i);
#pragma unroll
    for (int n = 0; n < N; ++n) un[n] = u[c0 + n * sn];
  };
  load_u(ie);
  for (int i = g.ie; i >= g.is; --i) {
    double uc[N], wc[N];
#pragma unroll
    for (int n = 0; n < N; ++n) uc[n] = un[n];
    double rho = uc[0];
#pragma unroll
    for (int n = 1; n <= NM; ++n) rho += uc[n];
    cons_to_prim_cell<NM>(th, uc, wc);
    // the level above's stores and the level below's load (at the last
    // level a reload of level is, unused)
    if (i < ie) flush(pend);
    load_u(max(i - 1, is));
    double fout[5];
    kessler_cell<NM>(th, kz, dt, cf, cvp, i == is, rho, uc, wc, fin, fout);
    pend.i = i;
#pragma unroll
    for (int n = 0; n < N; ++n) {
      pend.uc[n] = uc[n];
      pend.wc[n] = wc[n];
    }
    // hydrostatic face and cell pressures (k_hydrostatic_scan)
    if (FUSE && grav != 0.0) {
      if (i == ie) {
        const double RdTv = wc[IEN] / wc[0];
        pf_next = wc[IEN] * exp_d(grav * dx / (2 * RdTv));
        pf_top = pf_next;
        pend.ptop = pf_next;
      }
      const double pfi = pf_next - grav * wc[0] * dx;
      pend.pfi = pfi;
      pend.pvi = log_mean(pfi, pf_next);
      pf_next = pfi;
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) fin[n] = fout[n];
  }
  flush(pend);
  if (FUSE && grav != 0.0) {
    const double gb = (g.bc[0] == 0) ? -grav : grav;
    const double gt = (g.bc[1] == 0) ? -grav : grav;
    double cur = pf_next;
    for (int gh = 1; gh <= NG; ++gh) {
      const double lo = cur - gb * w[cb + (is - gh) * si] * dx;
      psf[pc + (is - gh) * g.ps_i] = lo;
      psv[pc + (is - gh) * g.ps_i] = log_mean(lo, cur);
      cur = lo;
    }
    cur = pf_top;
    for (int gh = 1; gh <= NG; ++gh) {
      const double hi = cur + gt * w[cb + (ie + gh) * si] * dx;
      psf[pc + (ie + gh + 1) * g.ps_i] = hi;
      psv[pc + (ie + gh) * g.ps_i] = log_mean(cur, hi);
      cur = hi;
    }
  }
}

// Every column of a block; the step is dt, or dt * dtp[0] when dtp is given.
template <int NM>
__global__ void __launch_bounds__(64)
    k_kessler(KGrid g, KThermo th, KKessler kz, double dt, const double* __restrict__ dtp,
              double* __restrict__ u, double* __restrict__ w) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx2 * g.nx3) return;
  const int j = g.js + (int)(t % g.nx2);
  const int k = g.ks + (int)(t / g.nx2);
  kessler_column<NM>(g, th, kz, dtp ? dt * dtp[0] : dt, u, w, k, j);
}

// k_kessler followed, in the same column sweep, by the column's x1 ghosts
// and hydrostatic pressures (runs whose x1 sides are walls or outflow).
template <int NM>
__global__ void __launch_bounds__(64)
    k_kessler_fill(KGrid g, KThermo th, KKessler kz, double dt,
                   const double* __restrict__ dtp, double* __restrict__ u,
                   double* __restrict__ w, double* __restrict__ psf,
                   double* __restrict__ psv) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx2 * g.nx3) return;
  const int j = g.js + (int)(t % g.nx2);
  const int k = g.ks + (int)(t / g.nx2);
  kessler_column<NM, true>(g, th, kz, dtp ? dt * dtp[0] : dt, u, w, k, j, psf, psv,
                           t * g.ps_c);
}

}  // namespace cg
