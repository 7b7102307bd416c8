// cg_physics.hpp -- per-cell / per-face physics of the canoe step, written
// for the device (templated on the number of mass species NM and on the
// arithmetic type R: double for the plain path, float for the FP32 fluxes).
//
// Expression order follows the reference text cited at each function so
// that the FP64 instantiation reproduces a CPU evaluation of the same
// formulas bit for bit (the library is built with -ffp-contract=off).
#pragma once

#include "cg_math.hpp"

namespace cg {

constexpr int NG = 3;

// Element strides of an array indexed (n, k, j, i).
struct Strides {
  long long n, k, j, i;
  CG_HD long long at(int nn, int kk, int jj, int ii) const {
    return (long long)nn * n + (long long)kk * k + (long long)jj * j +
           (long long)ii * i;
  }
};

// Grid of one MeshBlock.  Cell arrays use the strides `c`: the AthenaArray
// layout (i unit stride, the C-ABI's caller arrays) or the level-major
// layout of a run's resident state (j unit stride, so that column-per-lane
// kernels read whole rows per wavefront).  Face arrays use f[0..2];
// hydrostatic face/cell pressures use psf[i * ps_i + col * ps_c] with col =
// (k - ks) nx2 + (j - js).
struct KGrid {
  int nx1, nx2, nx3, nc1, nc2, nc3;
  int is, ie, js, je, ks, ke;
  int f2, f3;
  int lvl;  // 1: level-major mapping of lanes (j fastest), 0: i fastest
  long long sn;
  Strides c, f[3];
  long long ps_i, ps_c;
  double dx1, dx2, dx3, grav;
  int bc[6];
  // status word of the face checks (null: not checked)
  int* status = nullptr;
  CG_HD long long at(int n, int k, int j, int i) const { return c.at(n, k, j, i); }
  CG_HD long long pcol(int k, int j) const {
    return (long long)(k - ks) * nx2 + (j - js);
  }
  // (i, j, k) offsets of item t of a box of extents (n1, n2, n3), lanes
  // running along the unit-stride index
  CG_HD void decode(long long t, int n1, int n2, int n3, int& i, int& j,
                    int& k) const {
    if (lvl) {
      j = (int)(t % n2);
      long long r = t / n2;
      k = (int)(r % n3);
      i = (int)(r / n3);
    } else {
      i = (int)(t % n1);
      long long r = t / n1;
      j = (int)(r % n2);
      k = (int)(r / n2);
    }
  }
};

// Thermodynamic constants (cg_thermo plus derived saturation constants).
struct KThermo {
  int nmass, nvapor;
  double gammad, rd;
  double cv_ratio_m1[3], inv_mu_ratio_m1[3];
  int iv, ic;  // 0-based species of the vapour / condensate pair, -1 = none
  double cvd, cvv, cvc, rv, L, t3, p3, beta, delta, tol;
  int max_iter;
};

// get_gamma: src/interface/eos.hpp:47-67 (rho = 1: w holds mass fractions)
template <int NM, class R>
CG_HD R get_gamma(const KThermo& th, const R* u) {
  R fsig = R(1.), feps = R(1.);
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n <= th.nvapor) {
      fsig += u[n] / R(1.) * R(th.cv_ratio_m1[n - 1]);
      feps += u[n] / R(1.) * R(th.inv_mu_ratio_m1[n - 1]);
    }
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n > th.nvapor) {
      fsig += u[n] / R(1.) * R(th.cv_ratio_m1[n - 1]);
      feps -= u[n] / R(1.);
    }
  return R(1.) + (R(th.gammad) - R(1.)) * feps / fsig;
}

// SoundSpeed: src/snap/eos/ideal_moist_hydro.cpp:87-91
template <int NM, class R>
CG_HD R sound_speed(const KThermo& th, const R* prim) {
  constexpr int IPR = 4 + NM;
  R gamma = get_gamma<NM>(th, prim);
  return sqrt(gamma * prim[IPR] / prim[0]);
}

// LMARS: src/snap/riemann/lmars.cpp:19-85.  wl, wr hold N primitives in the
// global index order; ivx selects the normal velocity.
template <int NM, int DIR, class R>
CG_HD void lmars(const KThermo& th, const R* wli, const R* wri, R* flx) {
  constexpr int IVX = 1 + NM, IPR = IVX + 3, IDN = 0, IEN = IPR;
  constexpr int ivx = IVX + DIR;
  constexpr int ivy = IVX + ((ivx - IVX) + 1) % 3;
  constexpr int ivz = IVX + ((ivx - IVX) + 2) % 3;
  R kappal = R(1.) / (get_gamma<NM>(th, wli) - R(1.));
  R kappar = R(1.) / (get_gamma<NM>(th, wri) - R(1.));
  R hl = wli[IPR] / wli[IDN] * (kappal + R(1.)) +
         R(0.5) * (wli[ivx] * wli[ivx] + wli[ivy] * wli[ivy] +
                   wli[ivz] * wli[ivz]);
  R hr = wri[IPR] / wri[IDN] * (kappar + R(1.)) +
         R(0.5) * (wri[ivx] * wri[ivx] + wri[ivy] * wri[ivy] +
                   wri[ivz] * wri[ivz]);
  R rhobar = R(0.5) * (wli[IDN] + wri[IDN]);
  R cbar = sqrt(R(0.5) * (R(1.) + (R(1.) / kappar + R(1.) / kappal) / R(2.)) *
                (wli[IPR] + wri[IPR]) / rhobar);
  R pbar = R(0.5) * (wli[IPR] + wri[IPR]) +
           R(0.5) * (rhobar * cbar) * (wli[ivx] - wri[ivx]);
  R ubar = R(0.5) * (wli[ivx] + wri[ivx]) +
           R(0.5) / (rhobar * cbar) * (wli[IPR] - wri[IPR]);
  R rdl = R(1.), rdr = R(1.);
#pragma unroll
  for (int n = 1; n < IVX; ++n) {
    rdl -= wli[n];
    rdr -= wri[n];
  }
  // upwind on sign(ubar) (:69-83); values selected, not pointers, so the
  // states stay in registers
  const bool up = ubar > R(0.);
  R ws[5 + NM];
#pragma unroll
  for (int n = 0; n < 5 + NM; ++n) ws[n] = up ? wli[n] : wri[n];
  R rd = up ? rdl : rdr;
  R h = up ? hl : hr;
  flx[IDN] = ubar * ws[IDN] * rd;
#pragma unroll
  for (int n = 1; n < IVX; ++n) flx[n] = ubar * ws[IDN] * ws[n];
  flx[ivx] = ubar * ws[IDN] * ws[ivx] + pbar;
  flx[ivy] = ubar * ws[IDN] * ws[ivy];
  flx[ivz] = ubar * ws[IDN] * ws[ivz];
  flx[IEN] = ubar * ws[IDN] * h;
}

// HLLC: src/snap/riemann/hllc_transform.cpp:25-233 (GENERAL_EOS false).
template <int NM, int DIR, class R>
CG_HD void hllc(const KThermo& th, const R* wl, const R* wr, R* flx) {
  constexpr int IVX = 1 + NM, IVY = IVX + 1, IVZ = IVX + 2, IPR = IVX + 3;
  constexpr int IDN = 0, IEN = IPR, N = 5 + NM;
  const R TINY = R(1.0e-20);
  constexpr int ivx = IVX + DIR;
  constexpr int ivy = IVX + ((ivx - IVX) + 1) % 3;
  constexpr int ivz = IVX + ((ivx - IVX) + 2) % 3;
  R wli[N], wri[N], fl[N], fr[N], flxi[N];
  wli[IDN] = wl[IDN];
  wli[IVX] = wl[ivx];
  wli[IVY] = wl[ivy];
  wli[IVZ] = wl[ivz];
  wli[IPR] = wl[IPR];
  wri[IDN] = wr[IDN];
  wri[IVX] = wr[ivx];
  wri[IVY] = wr[ivy];
  wri[IVZ] = wr[ivz];
  wri[IPR] = wr[IPR];
#pragma unroll
  for (int n = 1; n < IVX; ++n) {
    wli[n] = wl[n];
    wri[n] = wr[n];
  }
  R gammal = get_gamma<NM>(th, wli);
  R kappal = R(1.) / (gammal - R(1.));
  R gammar = get_gamma<NM>(th, wri);
  R kappar = R(1.) / (gammar - R(1.));
  R cl = sound_speed<NM>(th, wli);
  R cr = sound_speed<NM>(th, wri);
  R KE_l = R(.5) * wli[IDN] *
           (sqr(wli[IVX]) + sqr(wli[IVY]) + sqr(wli[IVZ]));
  R KE_r = R(.5) * wri[IDN] *
           (sqr(wri[IVX]) + sqr(wri[IVY]) + sqr(wri[IVZ]));
  R el = wli[IPR] * kappal + KE_l;
  R er = wri[IPR] * kappar + KE_r;
  R rhoa = R(.5) * (wli[IDN] + wri[IDN]);
  R ca = R(.5) * (cl + cr);
  R pmid = R(.5) * (wli[IPR] + wri[IPR] + (wli[IVX] - wri[IVX]) * rhoa * ca);
  R ql = (pmid <= wli[IPR])
             ? R(1.0)
             : sqrt(R(1.0) + (gammal + R(1)) / (R(2) * gammal) *
                                 (pmid / wli[IPR] - R(1.0)));
  R qr = (pmid <= wri[IPR])
             ? R(1.0)
             : sqrt(R(1.0) + (gammar + R(1)) / (R(2) * gammar) *
                                 (pmid / wri[IPR] - R(1.0)));
  R al = wli[IVX] - cl * ql;
  R ar = wri[IVX] + cr * qr;
  R bp = ar > R(0.0) ? ar : TINY;
  R bm = al < R(0.0) ? al : -TINY;
  R vxl = wli[IVX] - al;
  R vxr = wri[IVX] - ar;
  R tl = wli[IPR] + vxl * wli[IDN] * wli[IVX];
  R tr = wri[IPR] + vxr * wri[IDN] * wri[IVX];
  R ml = wli[IDN] * vxl;
  R mr = -(wri[IDN] * vxr);
  R am = (tl - tr) / (ml + mr);
  R cp = (ml * tr + mr * tl) / (ml + mr);
  cp = cp > R(0.0) ? cp : R(0.0);
  vxl = wli[IVX] - bm;
  vxr = wri[IVX] - bp;
  R rdl = R(1.), rdr = R(1.);
#pragma unroll
  for (int n = 1; n < IVX; ++n) {
    rdl -= wli[n];
    rdr -= wri[n];
  }
  fl[IDN] = wli[IDN] * vxl * rdl;
  fr[IDN] = wri[IDN] * vxr * rdr;
#pragma unroll
  for (int n = 1; n < IVX; ++n) {
    fl[n] = wli[IDN] * wli[n] * vxl;
    fr[n] = wri[IDN] * wri[n] * vxr;
  }
  fl[IVX] = wli[IDN] * wli[IVX] * vxl + wli[IPR];
  fr[IVX] = wri[IDN] * wri[IVX] * vxr + wri[IPR];
  fl[IVY] = wli[IDN] * wli[IVY] * vxl;
  fr[IVY] = wri[IDN] * wri[IVY] * vxr;
  fl[IVZ] = wli[IDN] * wli[IVZ] * vxl;
  fr[IVZ] = wri[IDN] * wri[IVZ] * vxr;
  fl[IEN] = el * vxl + wli[IPR] * wli[IVX];
  fr[IEN] = er * vxr + wri[IPR] * wri[IVX];
  R sl, sr, sm;
  if (am >= R(0.0)) {
    sl = am / (am - bm);
    sr = R(0.0);
    sm = -bm / (am - bm);
  } else {
    sl = R(0.0);
    sr = -am / (bp - am);
    sm = bp / (bp - am);
  }
#pragma unroll
  for (int n = 0; n < IVX; ++n) flxi[n] = sl * fl[n] + sr * fr[n];
  flxi[IVX] = sl * fl[IVX] + sr * fr[IVX] + sm * cp;
  flxi[IVY] = sl * fl[IVY] + sr * fr[IVY];
  flxi[IVZ] = sl * fl[IVZ] + sr * fr[IVZ];
  flxi[IEN] = sl * fl[IEN] + sr * fr[IEN] + sm * cp * am;
#pragma unroll
  for (int n = 0; n < IVX; ++n) flx[n] = flxi[n];
  flx[ivx] = flxi[IVX];
  flx[ivy] = flxi[IVY];
  flx[ivz] = flxi[IVZ];
  flx[IEN] = flxi[IEN];
}

template <int NM, int DIR, class R>
CG_HD void riemann(const KThermo& th, int solver, const R* wl, const R* wr,
                   R* flx) {
  if (solver == 1)
    hllc<NM, DIR, R>(th, wl, wr, flx);
  else
    lmars<NM, DIR, R>(th, wl, wr, flx);
}

// Reconstruction of one cell (PLM: van Leer on a uniform mesh; WENO5:
// Jiang & Shu 1996, eps = 1e-6, shared smoothness indicators).  q[0..4] is
// the 5-cell window centred on the cell; qm / qp the lower / upper face.
template <class R>
CG_HD void plm_cell(R qm1, R q0, R qp1, R* qminus, R* qplus) {
  R dql = q0 - qm1;
  R dqr = qp1 - q0;
  R dq2 = dql * dqr;
  R dqm = R(2.0) * dq2 / (dql + dqr);
  if (dq2 <= R(0.0)) dqm = R(0.0);
  *qplus = q0 + R(0.5) * dqm;
  *qminus = q0 - R(0.5) * dqm;
}

template <class R>
CG_HD void weno5_cell(R a, R b, R c, R d, R e, R* qminus, R* qplus) {
  const R eps = R(1.0e-6);
  R b0 = R(13.0 / 12.0) * sqr(a - R(2.0) * b + c) +
         R(0.25) * sqr(a - R(4.0) * b + R(3.0) * c);
  R b1 = R(13.0 / 12.0) * sqr(b - R(2.0) * c + d) + R(0.25) * sqr(b - d);
  R b2 = R(13.0 / 12.0) * sqr(c - R(2.0) * d + e) +
         R(0.25) * sqr(R(3.0) * c - R(4.0) * d + e);
  R g0 = R(1.0) / sqr(eps + b0);
  R g1 = R(1.0) / sqr(eps + b1);
  R g2 = R(1.0) / sqr(eps + b2);
  R p0 = (R(2.0) * a - R(7.0) * b + R(11.0) * c) * R(1.0 / 6.0);
  R p1 = (-b + R(5.0) * c + R(2.0) * d) * R(1.0 / 6.0);
  R p2 = (R(2.0) * c + R(5.0) * d - e) * R(1.0 / 6.0);
  R a0 = R(0.1) * g0, a1 = R(0.6) * g1, a2 = R(0.3) * g2;
  *qplus = (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
  R m0 = (R(2.0) * c + R(5.0) * b - a) * R(1.0 / 6.0);
  R m1 = (-d + R(5.0) * c + R(2.0) * b) * R(1.0 / 6.0);
  R m2 = (R(2.0) * e - R(7.0) * d + R(11.0) * c) * R(1.0 / 6.0);
  R c0 = R(0.3) * g0, c1 = R(0.6) * g1, c2 = R(0.1) * g2;
  *qminus = (c0 * m0 + c1 * m1 + c2 * m2) / (c0 + c1 + c2);
}

// cons <-> prim, closed form (thermodynamics_athena.hpp:153-160)
template <int NM>
CG_HD void cons_to_prim_cell(const KThermo& th, const double* u, double* w) {
  constexpr int IVX = 1 + NM, IPR = IVX + 3;
  double rho = u[0];
#pragma unroll
  for (int n = 1; n <= NM; ++n) rho += u[n];
#pragma unroll
  for (int n = 1; n <= NM; ++n) w[n] = u[n] / rho;
  w[0] = rho;
  double v1 = u[IVX] / rho, v2 = u[IVX + 1] / rho, v3 = u[IVX + 2] / rho;
  double ke = 0.5 * (u[IVX] * v1 + u[IVX + 1] * v2 + u[IVX + 2] * v3);
  double gamma = get_gamma<NM>(th, w);
  w[IVX] = v1;
  w[IVX + 1] = v2;
  w[IVX + 2] = v3;
  w[IPR] = (gamma - 1.) * (u[IPR] - ke);
}

template <int NM>
CG_HD void prim_to_cons_cell(const KThermo& th, const double* w, double* u) {
  constexpr int IVX = 1 + NM, IPR = IVX + 3;
  double rho = w[0];
  double qd = 1.;
#pragma unroll
  for (int n = 1; n <= NM; ++n) qd -= w[n];
  u[0] = rho * qd;
#pragma unroll
  for (int n = 1; n <= NM; ++n) u[n] = rho * w[n];
  double v1 = w[IVX], v2 = w[IVX + 1], v3 = w[IVX + 2];
  u[IVX] = rho * v1;
  u[IVX + 1] = rho * v2;
  u[IVX + 2] = rho * v3;
  double gamma = get_gamma<NM>(th, w);
  u[IPR] = w[IPR] / (gamma - 1.) + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
}

// Saturation adjustment of one cell (definition in the library docs: moist
// internal energy ie - rho L q_c held fixed at fixed rho; Newton on T).
CG_HD double qsat_d(const KThermo& s, double T, double rho) {
  double es =
      s.p3 * exp_d((1. - s.t3 / T) * s.beta - s.delta * log_d(T / s.t3));
  return es / (rho * s.rv * T);
}

// q: NM mass fractions (0-based), ie in/out.  Returns Newton iterations, or
// -1 if not converged within max_iter.  tguess < 0: start from the
// all-vapour temperature.
template <int NM>
CG_HD int adjust_cell(const KThermo& th, double rho, double* ie, double* q,
                      double tguess, double* tout = nullptr) {
  double cvd = th.cvd;
  double qd = 1.;
#pragma unroll
  for (int n = 0; n < NM; ++n) qd -= q[n];
  double cfix = cvd * qd;
#pragma unroll
  for (int n = 0; n < NM; ++n) {
    if (n == th.iv || n == th.ic) continue;
    cfix += cvd * (1. + th.cv_ratio_m1[n]) * q[n];
  }
  double qiv = 0., qic = 0.;
#pragma unroll
  for (int n = 0; n < NM; ++n) {
    if (n == th.iv) qiv = q[n];
    if (n == th.ic) qic = q[n];
  }
  double qt = qiv + qic;
  double em = *ie / rho - th.L * qic;
  double tu = em / (cfix + th.cvv * qt);
  double qc, T;
  int it = 0;
  if (qt <= qsat_d(th, tu, rho)) {
    qc = 0.;
    T = tu;
  } else {
    T = (tguess > 0.) ? tguess : tu;
    bool conv = false;
    for (it = 1; it <= th.max_iter; ++it) {
      double qs = qsat_d(th, T, rho);
      double dqs = qs * ((th.beta * th.t3 / T - th.delta) / T - 1. / T);
      double qcn = qt - qs;
      double cvm = cfix + th.cvv * qs + th.cvc * qcn;
      double gT = cvm * T - th.L * qcn - em;
      double gp = cvm + dqs * (th.L - (th.cvc - th.cvv) * T);
      double dT = gT / gp;
      T = T - dT;
      if (fabs(dT) < th.tol) {
        conv = true;
        break;
      }
    }
    if (!conv) return -1;
    qc = qt - qsat_d(th, T, rho);
  }
  *ie = *ie + rho * th.L * (qc - qic);
#pragma unroll
  for (int n = 0; n < NM; ++n) {
    if (n == th.iv) q[n] = qt - qc;
    if (n == th.ic) q[n] = qc;
  }
  if (tout) *tout = T;
  return it;
}

}  // namespace cg
