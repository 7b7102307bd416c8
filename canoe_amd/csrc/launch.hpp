// launch.hpp -- host-side launch helpers shared by the stateless entry points
// (capi.hip) and the resident runs (run.hip): argument checks, the grid of a
// block in either layout, buffer sizes and the kernel launch sequences of the
// flux sweep, the implicit column solve and the ghost-cell fill.
#pragma once
#include <hip/hip_runtime.h>

#include "capi_internal.hpp"
#include "cg_timer.hpp"
#include "flux32.hpp"
#include "kernels.hpp"
#include "kessler.hpp"
#include "lines.hpp"

namespace cg_host {
using namespace cg;

inline int check_block(cg_ctx* ctx, const cg_block* b, const cg_thermo* th) {
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
  // each side's NGHOST ghost cells are filled from NGHOST interior cells
  if (b->nx1 < CG_NGHOST || (b->nx2 > 1 && b->nx2 < CG_NGHOST) ||
      (b->nx3 > 1 && b->nx3 < CG_NGHOST))
    return set_err(ctx, CG_ERR_ARG, "a direction with cells needs at least %d of them",
                   CG_NGHOST);
  if (b->dx1 <= 0 || b->dx2 <= 0 || b->dx3 <= 0)
    return set_err(ctx, CG_ERR_ARG, "non-positive spacing");
  for (int s = 0; s < 6; ++s)
    if (b->bc[s] < 0 || b->bc[s] > 2)
      return set_err(ctx, CG_ERR_ARG, "bad boundary flag");
  return CG_OK;
}

// Grid of a block in the AthenaArray layout (lvl = 0, the C-ABI's caller
// arrays) or in the level-major layout of a run's resident state (lvl = 1).
inline KGrid make_grid(const cg_block& b, int lvl) {
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

inline size_t psf_doubles(const KGrid& g) {
  return (size_t)(g.nc1 + 1) * g.nx2 * g.nx3;
}


inline size_t cell_doubles(const KGrid& g, int N) { return (size_t)N * g.sn; }
inline size_t f1_doubles(const KGrid& g, int N) {
  return (size_t)N * g.nc3 * g.nc2 * (g.nc1 + 1);
}
inline size_t f2_doubles(const KGrid& g, int N) {
  return (size_t)N * g.nc3 * (g.nc2 + 1) * g.nc1;
}
inline size_t f3_doubles(const KGrid& g, int N) {
  return (size_t)N * (g.nc3 + 1) * g.nc2 * g.nc1;
}

// ---- launch helpers (templated on the species count) ----
template <int NM>
void launch_fluxes(const KGrid& g, const KThermo& th, const cg_config& cfg,
                   const double* w, double* psf, double* psv, double* f1,
                   double* f2, double* f3, hipStream_t s, KTimer* tm = nullptr,
                   bool scan = true, int bs = 256) {
  const long long ncol = (long long)g.nx2 * g.nx3;
  const int decomp = g.grav != 0.0;
  if (decomp && scan)
    TIMED(tm, KID_SCAN,
          (k_hydrostatic_scan<NM><<<nblk(ncol, 256), 256, 0, s>>>(g, th, w, psf, psv)));
  const long long n1 = ncol * (g.nx1 + 1);
  if (cfg.order_v == 5)
    TIMED(tm, KID_FLUX_X1,
          (k_flux_x1<NM, 5><<<nblk(n1, bs), bs, 0, s>>>(g, th, cfg.rsolver, decomp,
                                                           w, psf, psv, f1)));
  else
    TIMED(tm, KID_FLUX_X1,
          (k_flux_x1<NM, 2><<<nblk(n1, bs), bs, 0, s>>>(g, th, cfg.rsolver, decomp,
                                                           w, psf, psv, f1)));
  if (g.f2) {
    long long n2 = (long long)g.nx1 * (g.nx2 + 1) * g.nx3;
    if (cfg.order_h == 5)
      TIMED(tm, KID_FLUX_X2,
            (k_flux_h<NM, 5, 1><<<nblk(n2, bs), bs, 0, s>>>(g, th, cfg.rsolver, w, f2)));
    else
      TIMED(tm, KID_FLUX_X2,
            (k_flux_h<NM, 2, 1><<<nblk(n2, bs), bs, 0, s>>>(g, th, cfg.rsolver, w, f2)));
  }
  if (g.f3) {
    long long n3 = (long long)g.nx1 * g.nx2 * (g.nx3 + 1);
    if (cfg.order_h == 5)
      TIMED(tm, KID_FLUX_X3,
            (k_flux_h<NM, 5, 2><<<nblk(n3, bs), bs, 0, s>>>(g, th, cfg.rsolver, w, f3)));
    else
      TIMED(tm, KID_FLUX_X3,
            (k_flux_h<NM, 2, 2><<<nblk(n3, bs), bs, 0, s>>>(g, th, cfg.rsolver, w, f3)));
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

inline size_t arch_doubles(const KGrid& g) {
  return (size_t)g.nx1 * 30 * (size_t)g.nx2 * g.nx3;
}

// The lines of direction dir (0 = x1, 1 = x2, 2 = x3) of a block: lanes
// along the transverse index of smaller stride; Cartesian face areas and
// volume (x1 faces dx2 dx3, x2 faces dx1 dx3, x3 faces dx1 dx2).
inline KLines make_lines(const KGrid& g, int dir) {
  const int st[3] = {g.is, g.js, g.ks};
  const int ex[3] = {g.nx1, g.nx2, g.nx3};
  const long long sd[3] = {g.c.i, g.c.j, g.c.k};
  int t0 = (dir + 1) % 3, t1 = (dir + 2) % 3;
  if (sd[t1] < sd[t0]) std::swap(t0, t1);
  KLines L;
  L.nf = ex[t0];
  L.nline = (long long)ex[t0] * ex[t1];
  L.n = ex[dir];
  L.lo = st[dir];
  L.org0 = (long long)st[t0] * sd[t0] + (long long)st[t1] * sd[t1];
  L.sf = sd[t0];
  L.ss = sd[t1];
  L.sl = sd[dir];
  L.sn = g.c.n;
  const double area = (dir == 0) ? g.dx2 * g.dx3 : (dir == 1) ? g.dx1 * g.dx3 : g.dx1 * g.dx2;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  const double s2v = 1. / (2. * vol);
  L.sla = area * s2v;
  L.sra = area * s2v;
  L.sda = (area - area) * s2v;
  L.grav = (dir == 0) ? g.grav : 0.0;
  L.td0 = t0;
  L.td1 = t1;
  L.t0s = st[t0];
  L.t1s = st[t1];
  L.nc1 = g.nc1;
  L.nc2 = g.nc2;
  L.nc3 = g.nc3;
  L.c = g.c;
  return L;
}

// Scratch doubles of the line solve of direction dir (B x B blocks).
inline size_t line_doubles(const KGrid& g, int dir, int B, bool per) {
  KLines L = make_lines(g, dir);
  const size_t ls = B == 5 ? line_level_size<5>(per) : line_level_size<3>(per);
  return (size_t)L.n * ls * (size_t)L.nline;
}

template <int NM, int B, int DIR>
void launch_lines_d(const KLines& L, const KThermo& th, double dt, const double* dtp,
                    bool per, double* du, const double* w, double* lv, int* status,
                    hipStream_t s, KTimer* tm) {
  const unsigned nb = nblk(L.nline, 64);
  if (per)
    TIMED(tm, KID_LINES,
          (k_line_periodic<NM, B, DIR><<<nb, 64, 0, s>>>(L, th, dt, dtp, w, du, lv, status)));
  else
    TIMED(tm, KID_LINES,
          (k_line_walls<NM, B, DIR><<<nb, 64, 0, s>>>(L, th, dt, dtp, w, du, lv, status)));
}

// SolveImplicit3D's lines of one direction in place on du (x2 / x3 lines,
// or periodic lines of any direction); lv holds line_doubles(g, dir, B, per).
// The step is dt, or dt * dtp[0] when dtp is given.
template <int NM>
void launch_lines(const KGrid& g, const KThermo& th, double dt, const double* dtp,
                  int flag, int dir, bool per, double* du, const double* w, double* lv,
                  int* status, hipStream_t s, KTimer* tm = nullptr) {
  const KLines L = make_lines(g, dir);
#define CG_LINES(B, D) \
  launch_lines_d<NM, B, D>(L, th, dt, dtp, per, du, w, lv, status, s, tm)
  if (flag & 8) {
    if (dir == 0) CG_LINES(5, 0);
    if (dir == 1) CG_LINES(5, 1);
    if (dir == 2) CG_LINES(5, 2);
  } else {
    if (dir == 0) CG_LINES(3, 0);
    if (dir == 1) CG_LINES(3, 1);
    if (dir == 2) CG_LINES(3, 2);
  }
#undef CG_LINES
}

// Periodic directions of a block (both boundaries periodic,
// implicit_solver.cpp:127-134) and the checks of SolveImplicit3D's flags.
inline void periodic_dirs(const cg_block& b, bool per[3]) {
  for (int d = 0; d < 3; ++d)
    per[d] = b.bc[2 * d] == CG_BC_PERIODIC && b.bc[2 * d + 1] == CG_BC_PERIODIC;
}

inline int check_implicit_flag(cg_ctx* ctx, const cg_block& b, int flag) {
  if (flag < 0 || flag > 15)
    return set_err(ctx, CG_ERR_ARG, "implicit_flag must be in 0..15 (got %d)", flag);
  bool per[3];
  periodic_dirs(b, per);
  // x2 / x3 lines between walls with species weight the species by w at the
  // shuffled index (w_shuffled, forward_backward.hpp:109-115): the largest
  // such entry must lie inside w (the reference reads past its end otherwise)
  if (b.nmass > 0) {
    const long long nc1 = b.nx1 + 2 * NG, nc2 = b.nx2 > 1 ? b.nx2 + 2 * NG : 1,
                    nc3 = b.nx3 > 1 ? b.nx3 + 2 * NG : 1;
    const long long ie = NG + b.nx1 - 1, je = (b.nx2 > 1 ? NG : 0) + b.nx2 - 1,
                    ke = (b.nx3 > 1 ? NG : 0) + b.nx3 - 1;
    const long long size = (long long)(5 + b.nmass) * nc3 * nc2 * nc1;
    auto flat = [&](long long K, long long J, long long I) {
      return ((b.nmass * nc3 + K) * nc2 + J) * nc1 + I;
    };
    if (((flag & 2) && b.nx2 > 1 && !per[1] && flat(ie, ke, je) >= size) ||
        ((flag & 4) && b.nx3 > 1 && !per[2] && flat(je, ie, ke) >= size))
      return set_err(ctx, CG_ERR_ARG,
                     "x2/x3 implicit lines with species between walls: the "
                     "reference's species weights would read past the end of w "
                     "for this block");
  }
  return CG_OK;
}

// Kessler microphysics of a configuration: on when any rate is non-zero;
// the precipitation is the condensate that is not the adjusted cloud.
inline bool kessler_on(const cg_config& c) {
  return c.k_auto != 0. || c.k_accr != 0. || c.k_evap != 0. || c.v_fall != 0.;
}

inline int make_kessler(cg_ctx* ctx, const cg_config& c, const cg_thermo& th,
                        KKessler* out) {
  out->k_auto = c.k_auto;
  out->k_accr = c.k_accr;
  out->k_evap = c.k_evap;
  out->v_fall = c.v_fall;
  out->ip = -1;
  if (!(c.k_auto >= 0. && c.k_accr >= 0. && c.k_evap >= 0. && c.v_fall >= 0.))
    return set_err(ctx, CG_ERR_ARG, "Kessler rates and fall speed must be >= 0");
  if (th.ivap > 0 && th.icld > 0)
    for (int n = th.nvapor; n < th.nmass; ++n)
      if (n != th.ivap - 1 && n != th.icld - 1) {
        out->ip = n;
        break;
      }
  if (out->ip < 0)
    return set_err(ctx, CG_ERR_ARG,
                   "Kessler microphysics needs vapour, cloud and a precipitation species");
  return CG_OK;
}

template <int NM>
void launch_kessler(const KGrid& g, const KThermo& th, const KKessler& kz, double dt,
                    const double* dtp, double* u, double* w, hipStream_t s,
                    KTimer* tm = nullptr) {
  TIMED(tm, KID_KESSLER,
        (k_kessler<NM><<<nblk((long long)g.nx2 * g.nx3, 64), 64, 0, s>>>(g, th, kz, dt, dtp,
                                                                          u, w)));
}

// launch_kessler with the column's x1 ghosts and hydrostatic pressures in
// the same sweep (x1 sides that are walls or outflow)
template <int NM>
void launch_kessler_fill(const KGrid& g, const KThermo& th, const KKessler& kz, double dt,
                         const double* dtp, double* u, double* w, double* psf,
                         double* psv, hipStream_t s, KTimer* tm = nullptr) {
  TIMED(tm, KID_KESSLER,
        (k_kessler_fill<NM><<<nblk((long long)g.nx2 * g.nx3, 64), 64, 0, s>>>(
            g, th, kz, dt, dtp, u, w, psf, psv)));
}

}  // namespace cg_host
