// oracle.cpp -- CPU FP64 oracle of the canoe dynamical-core step.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg, never by the product library.
//
// Restates, in the reference's own loop structure (per-block (k,j) rows,
// per-column block Thomas with a 5x5 partial-pivot inverse), the algorithms
// of chengcli/canoe cited at each function.  Where the reference's arithmetic
// lives in code that is not in /root/reference (athenapp's reconstruction and
// flux divergence, snap's EOS, kintera's saturation adjustment) the function
// states the definition adopted and why.
//
// Parity pins (tests/test_oracle_*.py): closed forms, invariants, brute-force
// dense solves and the reference's WENO5 known answer.  Built with
// -O2 -ffp-contract=off so that no multiply-add is fused.

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_math.hpp"

namespace orc {

constexpr int NG = CG_NGHOST;
constexpr double TINY_NUMBER = 1.0e-20;  // Athena++ athena.hpp

static int g_threads = 1;

struct Grid {
  int nx1, nx2, nx3, nc1, nc2, nc3;
  int is, ie, js, je, ks, ke;
  int nmass, N, IVX, IVY, IVZ, IPR;
  size_t ncell;
  bool f2, f3;
  explicit Grid(const cg_block& b) {
    nx1 = b.nx1;
    nx2 = b.nx2;
    nx3 = b.nx3;
    f2 = nx2 > 1;
    f3 = nx3 > 1;
    nc1 = nx1 + 2 * NG;
    nc2 = f2 ? nx2 + 2 * NG : 1;
    nc3 = f3 ? nx3 + 2 * NG : 1;
    is = NG;
    ie = NG + nx1 - 1;
    js = f2 ? NG : 0;
    je = js + nx2 - 1;
    ks = f3 ? NG : 0;
    ke = ks + nx3 - 1;
    nmass = b.nmass;
    N = 5 + nmass;
    IVX = 1 + nmass;
    IVY = IVX + 1;
    IVZ = IVX + 2;
    IPR = IVX + 3;
    ncell = (size_t)nc1 * nc2 * nc3;
  }
  size_t at(int n, int k, int j, int i) const {
    return (((size_t)n * nc3 + k) * nc2 + j) * nc1 + i;
  }
};

// ---------------------------------------------------------------------------
// EOS helpers
// get_gamma: src/interface/eos.hpp:47-67 (Li 2019 eq. 63)
double get_gamma(const cg_thermo& th, const double* u, double rho = 1.) {
  double fsig = 1., feps = 1.;
  for (int n = 1; n <= th.nvapor; ++n) {
    fsig += u[n] / rho * th.cv_ratio_m1[n - 1];
    feps += u[n] / rho * th.inv_mu_ratio_m1[n - 1];
  }
  for (int n = 1 + th.nvapor; n <= th.nmass; ++n) {
    fsig += u[n] / rho * th.cv_ratio_m1[n - 1];
    feps -= u[n] / rho;
  }
  return 1. + (th.gammad - 1.) * feps / fsig;
}

// EquationOfState::SoundSpeed: src/snap/eos/ideal_moist_hydro.cpp:87-91
double sound_speed(const cg_thermo& th, const double* prim, int IPR) {
  double gamma = get_gamma(th, prim);
  return std::sqrt(gamma * prim[IPR] / prim[0]);
}

// ---------------------------------------------------------------------------
// Riemann solvers

// LMARS: src/snap/riemann/lmars.cpp:19-85 (Cartesian metric: vec_lower = id)
void lmars(const cg_thermo& th, int ivx, int il, int iu, const double* wl,
           const double* wr, ptrdiff_t ld, double* flx, ptrdiff_t fld) {
  const int nmass = th.nmass, IVX = 1 + nmass, IPR = IVX + 3, N = 5 + nmass;
  const int IDN = 0, IEN = IPR;
  int ivy = IVX + ((ivx - IVX) + 1) % 3;
  int ivz = IVX + ((ivx - IVX) + 2) % 3;
  double wli[5 + CG_MAX_NMASS], wri[5 + CG_MAX_NMASS];
  for (int i = il; i <= iu; ++i) {
    for (int n = 0; n < N; ++n) {
      wli[n] = wl[n * ld + i];
      wri[n] = wr[n * ld + i];
    }
    double kappal = 1. / (get_gamma(th, wli) - 1.);
    double kappar = 1. / (get_gamma(th, wri) - 1.);
    double hl = wli[IPR] / wli[IDN] * (kappal + 1.) +
                0.5 * (wli[ivx] * wli[ivx] + wli[ivy] * wli[ivy] +
                       wli[ivz] * wli[ivz]);
    double hr = wri[IPR] / wri[IDN] * (kappar + 1.) +
                0.5 * (wri[ivx] * wri[ivx] + wri[ivy] * wri[ivy] +
                       wri[ivz] * wri[ivz]);
    double rhobar = 0.5 * (wli[IDN] + wri[IDN]);
    double cbar = std::sqrt(0.5 * (1. + (1. / kappar + 1. / kappal) / 2.) *
                            (wli[IPR] + wri[IPR]) / rhobar);
    double pbar = 0.5 * (wli[IPR] + wri[IPR]) +
                  0.5 * (rhobar * cbar) * (wli[ivx] - wri[ivx]);
    double ubar = 0.5 * (wli[ivx] + wri[ivx]) +
                  0.5 / (rhobar * cbar) * (wli[IPR] - wri[IPR]);
    double rdl = 1., rdr = 1.;
    for (int n = 1; n < IVX; ++n) {
      rdl -= wli[n];
      rdr -= wri[n];
    }
    if (ubar > 0.) {
      flx[IDN * fld + i] = ubar * wli[IDN] * rdl;
      for (int n = 1; n < IVX; ++n) flx[n * fld + i] = ubar * wli[IDN] * wli[n];
      flx[ivx * fld + i] = ubar * wli[IDN] * wli[ivx] + pbar;
      flx[ivy * fld + i] = ubar * wli[IDN] * wli[ivy];
      flx[ivz * fld + i] = ubar * wli[IDN] * wli[ivz];
      flx[IEN * fld + i] = ubar * wli[IDN] * hl;
    } else {
      flx[IDN * fld + i] = ubar * wri[IDN] * rdr;
      for (int n = 1; n < IVX; ++n) flx[n * fld + i] = ubar * wri[IDN] * wri[n];
      flx[ivx * fld + i] = ubar * wri[IDN] * wri[ivx] + pbar;
      flx[ivy * fld + i] = ubar * wri[IDN] * wri[ivy];
      flx[ivz * fld + i] = ubar * wri[IDN] * wri[ivz];
      flx[IEN * fld + i] = ubar * wri[IDN] * hr;
    }
  }
}

// HLLC: src/snap/riemann/hllc_transform.cpp:25-233 (GENERAL_EOS false,
// no AFFINE / CUBED_SPHERE projection)
void hllc(const cg_thermo& th, int ivx, int il, int iu, const double* wl,
          const double* wr, ptrdiff_t ld, double* flx, ptrdiff_t fld) {
  const int nmass = th.nmass, IVX = 1 + nmass, IVY = IVX + 1, IVZ = IVX + 2;
  const int IPR = IVX + 3, IDN = 0, IEN = IPR;
  int ivy = IVX + ((ivx - IVX) + 1) % 3;
  int ivz = IVX + ((ivx - IVX) + 2) % 3;
  double wli[5 + CG_MAX_NMASS], wri[5 + CG_MAX_NMASS];
  double flxi[5 + CG_MAX_NMASS], fl[5 + CG_MAX_NMASS], fr[5 + CG_MAX_NMASS];
  for (int i = il; i <= iu; ++i) {
    wli[IDN] = wl[IDN * ld + i];
    wli[IVX] = wl[ivx * ld + i];
    wli[IVY] = wl[ivy * ld + i];
    wli[IVZ] = wl[ivz * ld + i];
    wli[IPR] = wl[IPR * ld + i];
    wri[IDN] = wr[IDN * ld + i];
    wri[IVX] = wr[ivx * ld + i];
    wri[IVY] = wr[ivy * ld + i];
    wri[IVZ] = wr[ivz * ld + i];
    wri[IPR] = wr[IPR * ld + i];
    for (int n = 1; n < IVX; ++n) {
      wli[n] = wl[n * ld + i];
      wri[n] = wr[n * ld + i];
    }
    double gammal = get_gamma(th, wli);
    double kappal = 1. / (gammal - 1.);
    double gammar = get_gamma(th, wri);
    double kappar = 1. / (gammar - 1.);

    double cl = sound_speed(th, wli, IPR);
    double cr = sound_speed(th, wri, IPR);
    double KE_l = 0.5 * wli[IDN] *
                  (sqr(wli[IVX]) + sqr(wli[IVY]) + sqr(wli[IVZ]));
    double KE_r = 0.5 * wri[IDN] *
                  (sqr(wri[IVX]) + sqr(wri[IVY]) + sqr(wri[IVZ]));
    double el = wli[IPR] * kappal + KE_l;
    double er = wri[IPR] * kappar + KE_r;
    double rhoa = .5 * (wli[IDN] + wri[IDN]);
    double ca = .5 * (cl + cr);
    double pmid = .5 * (wli[IPR] + wri[IPR] + (wli[IVX] - wri[IVX]) * rhoa * ca);
    // the PVRS velocity (:121-124) feeds only the GENERAL_EOS branch of the
    // shock factors (:128-136); the ideal-gas branch below does not use it
    double umid =
        .5 * (wli[IVX] + wri[IVX] + (wli[IPR] - wri[IPR]) / (rhoa * ca));
    (void)umid;

    double ql = (pmid <= wli[IPR])
                    ? 1.0
                    : std::sqrt(1.0 + (gammal + 1) / (2 * gammal) *
                                          (pmid / wli[IPR] - 1.0));
    double qr = (pmid <= wri[IPR])
                    ? 1.0
                    : std::sqrt(1.0 + (gammar + 1) / (2 * gammar) *
                                          (pmid / wri[IPR] - 1.0));
    double al = wli[IVX] - cl * ql;
    double ar = wri[IVX] + cr * qr;
    double bp = ar > 0.0 ? ar : (TINY_NUMBER);
    double bm = al < 0.0 ? al : -(TINY_NUMBER);

    double vxl = wli[IVX] - al;
    double vxr = wri[IVX] - ar;
    double tl = wli[IPR] + vxl * wli[IDN] * wli[IVX];
    double tr = wri[IPR] + vxr * wri[IDN] * wri[IVX];
    double ml = wli[IDN] * vxl;
    double mr = -(wri[IDN] * vxr);
    double am = (tl - tr) / (ml + mr);
    double cp = (ml * tr + mr * tl) / (ml + mr);
    cp = cp > 0.0 ? cp : 0.0;

    vxl = wli[IVX] - bm;
    vxr = wri[IVX] - bp;
    double rdl = 1., rdr = 1.;
    for (int n = 1; n < IVX; ++n) {
      rdl -= wli[n];
      rdr -= wri[n];
    }
    fl[IDN] = wli[IDN] * vxl * rdl;
    fr[IDN] = wri[IDN] * vxr * rdr;
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

    double sl, sr, sm;
    if (am >= 0.0) {
      sl = am / (am - bm);
      sr = 0.0;
      sm = -bm / (am - bm);
    } else {
      sl = 0.0;
      sr = -am / (bp - am);
      sm = bp / (bp - am);
    }
    for (int n = 0; n < IVX; ++n) flxi[n] = sl * fl[n] + sr * fr[n];
    flxi[IVX] = sl * fl[IVX] + sr * fr[IVX] + sm * cp;
    flxi[IVY] = sl * fl[IVY] + sr * fr[IVY];
    flxi[IVZ] = sl * fl[IVZ] + sr * fr[IVZ];
    flxi[IEN] = sl * fl[IEN] + sr * fr[IEN] + sm * cp * am;

    for (int n = 0; n < IVX; ++n) flx[n * fld + i] = flxi[n];
    flx[ivx * fld + i] = flxi[IVX];
    flx[ivy * fld + i] = flxi[IVY];
    flx[ivz * fld + i] = flxi[IVZ];
    flx[IEN * fld + i] = flxi[IEN];
  }
}

void riemann(const cg_thermo& th, int solver, int ivx, int il, int iu,
             const double* wl, const double* wr, ptrdiff_t ld, double* flx,
             ptrdiff_t fld) {
  if (solver == CG_RS_HLLC)
    hllc(th, ivx, il, iu, wl, wr, ld, flx, fld);
  else
    lmars(th, ivx, il, iu, wl, wr, ld, flx, fld);
}

// ---------------------------------------------------------------------------
// Reconstruction.  athenapp's PiecewiseLinearX* / Weno5X* are not in the
// container; the definitions adopted (call sites patches/19:29-39):
//  PLM: van Leer limited slope on a uniform mesh (Athena++ plm.cpp's
//       uniform-Cartesian branch), face offset exactly 1/2.
//  WENO5: Jiang & Shu (1996) on primitive variables, eps = 1e-6, both faces
//       of a cell from the same three smoothness indicators.  Known answer:
//       tests/test_glog.cpp:28-37 (1,2,3,4,5 -> 2.5 at the face between 2, 3).
void plm_cell(double qm1, double q0, double qp1, double* qminus,
              double* qplus) {
  double dql = q0 - qm1;
  double dqr = qp1 - q0;
  double dq2 = dql * dqr;
  double dqm = 2.0 * dq2 / (dql + dqr);
  if (dq2 <= 0.0) dqm = 0.0;
  *qplus = q0 + 0.5 * dqm;
  *qminus = q0 - 0.5 * dqm;
}

void weno5_cell(double a, double b, double c, double d, double e,
                double* qminus, double* qplus) {
  const double eps = 1.0e-6;
  double b0 = (13.0 / 12.0) * sqr(a - 2.0 * b + c) +
              0.25 * sqr(a - 4.0 * b + 3.0 * c);
  double b1 = (13.0 / 12.0) * sqr(b - 2.0 * c + d) + 0.25 * sqr(b - d);
  double b2 = (13.0 / 12.0) * sqr(c - 2.0 * d + e) +
              0.25 * sqr(3.0 * c - 4.0 * d + e);
  double g0 = 1.0 / sqr(eps + b0);
  double g1 = 1.0 / sqr(eps + b1);
  double g2 = 1.0 / sqr(eps + b2);
  // upper face c+1/2: linear weights (0.1, 0.6, 0.3)
  double p0 = (2.0 * a - 7.0 * b + 11.0 * c) * (1.0 / 6.0);
  double p1 = (-b + 5.0 * c + 2.0 * d) * (1.0 / 6.0);
  double p2 = (2.0 * c + 5.0 * d - e) * (1.0 / 6.0);
  double a0 = 0.1 * g0, a1 = 0.6 * g1, a2 = 0.3 * g2;
  *qplus = (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
  // lower face c-1/2: mirrored weights (0.3, 0.6, 0.1)
  double m0 = (2.0 * c + 5.0 * b - a) * (1.0 / 6.0);
  double m1 = (-d + 5.0 * c + 2.0 * b) * (1.0 / 6.0);
  double m2 = (2.0 * e - 7.0 * d + 11.0 * c) * (1.0 / 6.0);
  double c0 = 0.3 * g0, c1 = 0.6 * g1, c2 = 0.1 * g2;
  *qminus = (c0 * m0 + c1 * m1 + c2 * m2) / (c0 + c1 + c2);
}

void recon_cell(int order, const double* q, double* qm, double* qp) {
  if (order == 5)
    weno5_cell(q[0], q[1], q[2], q[3], q[4], qm, qp);
  else
    plm_cell(q[1], q[2], q[3], qm, qp);
}

// Reconstruct along a line: base points at cell 0 of the line, stride s.
// For cells c in [cl, cu]: wl[c+1] = upper-face value, wr[c] = lower-face.
void recon_line(int order, const double* base, ptrdiff_t s, int cl, int cu,
                double* wl, double* wr) {
  for (int c = cl; c <= cu; ++c) {
    double q[5];
    for (int m = -2; m <= 2; ++m) {
      if (order != 5 && (m == -2 || m == 2)) {
        q[m + 2] = 0.0;
        continue;
      }
      q[m + 2] = base[(ptrdiff_t)(c + m) * s];
    }
    double qm, qp;
    recon_cell(order, q, &qm, &qp);
    wl[c + 1] = qp;
    wr[c] = qm;
  }
}

// ---------------------------------------------------------------------------
// Hydrostatic decomposition of one column (k,j):
// Decomposition::ChangeToTemperature
//   src/snap/decomposition/change_to_temperature.cpp:40-136
// (single block: no neighbours; reflecting walls per block bcs)
// Writes psf (nc1+1), psv (nc1), rdv (nc1) and, if wcol != nullptr,
// overwrites wcol[IDN], wcol[IPR] (stride 1 columns) with T and p - psv.
void decompose_column(const Grid& g, const cg_block& blk, const cg_thermo& th,
                      const double* w, int k, int j, double* psf, double* psv,
                      double* rdv) {
  const double grav = blk.grav;
  const double Rd = th.rd;
  const double dx = blk.dx1;
  const int is = g.is, ie = g.ie;
  auto W = [&](int n, int i) { return w[g.at(n, k, j, i)]; };
  // isothermal extrapolation to the top face (:58-63)
  {
    double RdTv = W(g.IPR, ie) / W(0, ie);
    psf[ie + 1] = W(g.IPR, ie) * cg_exp(grav * dx / (2 * RdTv));
  }
  // IntegrateDownwards(psf, w, grav, is, ie) (:30-38, :65)
  for (int i = ie; i >= is; --i) psf[i] = psf[i + 1] - grav * W(0, i) * dx;
  // bottom ghosts (:71-76)
  double gb = (blk.bc[0] == CG_BC_REFLECT) ? -grav : grav;
  for (int i = is - 1; i >= is - NG; --i)
    psf[i] = psf[i + 1] - gb * W(0, i) * dx;
  // top ghosts (:78-82): IntegrateUpwards
  double gt = (blk.bc[1] == CG_BC_REFLECT) ? -grav : grav;
  for (int i = ie + 1; i <= ie + NG; ++i)
    psf[i + 1] = psf[i] + gt * W(0, i) * dx;
  // (:92-111)
  for (int i = is - NG; i <= ie + NG; ++i) {
    if (std::fabs(psf[i] - psf[i + 1]) < 1.E-6)
      psv[i] = 0.5 * (psf[i] + psf[i + 1]);
    else
      psv[i] = (psf[i] - psf[i + 1]) / cg_log(psf[i] / psf[i + 1]);
    double feps = 1.;
    for (int n = 1; n <= th.nvapor; ++n)
      feps += W(n, i) * th.inv_mu_ratio_m1[n - 1];
    rdv[i] = Rd * feps;
  }
}

// The x1 ghosts of w that the reference's CalculateFluxes leaves behind on
// an outflow side.  ChangeToTemperature gives those ghosts T and p' of the
// boundary cell b (change_to_temperature.cpp:113-127) and
// RestoreFromTemperature (:150-153) puts them back over the ghost's own
// hydrostatic pressure, so until the next boundary fill the stage's sources
// and implicit solve read
//   p_g = (p_b - psv_b) + psv_g,  rho_g = p_g / (T_b R_g),
//   T_b = p_b / (rho_b R_b).
// (On every other cell the restore undoes the change up to rounding; the
// oracle's pure flux sweep leaves those cells alone.)  Interior (k, j)
// rows, the ones later readers use.
void restore_outflow_ghosts(const Grid& g, const cg_block& blk,
                            const cg_thermo& th, double* w) {
  if (blk.grav == 0. ||
      (blk.bc[0] != CG_BC_OUTFLOW && blk.bc[1] != CG_BC_OUTFLOW))
    return;
  const int IPR = g.IPR;
  std::vector<double> psf(g.nc1 + 1), psv(g.nc1), rdv(g.nc1);
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j) {
      decompose_column(g, blk, th, w, k, j, psf.data(), psv.data(), rdv.data());
      auto side = [&](int b, int lo, int hi) {
        double Tb = w[g.at(IPR, k, j, b)] / (w[g.at(0, k, j, b)] * rdv[b]);
        double pb = w[g.at(IPR, k, j, b)] - psv[b];
        for (int i = lo; i <= hi; ++i) {
          double pg = pb + psv[i];
          w[g.at(IPR, k, j, i)] = pg;
          w[g.at(0, k, j, i)] = pg / (Tb * rdv[i]);
        }
      };
      if (blk.bc[0] == CG_BC_OUTFLOW) side(g.is, g.is - NG, g.is - 1);
      if (blk.bc[1] == CG_BC_OUTFLOW) side(g.ie, g.ie + 1, g.ie + NG);
    }
}

// ---------------------------------------------------------------------------
// check_reconstruct (src/checks.cpp:79-97), called after the face states of
// each direction (patches/19.calculate_fluxes.patch:45, 73, 117): the faces
// il..iu of wl / wr (entry n * ld + i) with a negative density or pressure,
// counted once per face into g_face_failures.
std::atomic<long long> g_face_failures{0};
void check_reconstruct(int IPR, const double* wl, const double* wr, size_t ld,
                       int il, int iu) {
  long long n = 0;
  for (int i = il; i <= iu; ++i)
    if (wl[i] < 0. || wr[i] < 0. || wl[IPR * ld + i] < 0. || wr[IPR * ld + i] < 0.)
      ++n;
  if (n) g_face_failures += n;
}

// Hydro::CalculateFluxes (patches/19.calculate_fluxes.patch:19-185), pure
// apart from the face-check count.
void calculate_fluxes(const Grid& g, const cg_block& blk, const cg_thermo& th,
                      const cg_config& cfg, const double* w, double* f1,
                      double* f2, double* f3) {
  const int N = g.N, IPR = g.IPR, nc1 = g.nc1;
  const bool decomp = blk.grav != 0.;
  // ---- x1 ----
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j) {
      std::vector<double> col((size_t)N * nc1), wl((size_t)N * (nc1 + 1)),
          wr((size_t)N * (nc1 + 1)), psf(nc1 + 1), psv(nc1), rdv(nc1);
      for (int n = 0; n < N; ++n)
        for (int i = 0; i < nc1; ++i) col[n * nc1 + i] = w[g.at(n, k, j, i)];
      if (decomp) {
        decompose_column(g, blk, th, w, k, j, psf.data(), psv.data(),
                         rdv.data());
        // change density to temperature, pressure to perturbation (:107-110)
        for (int i = g.is - NG; i <= g.ie + NG; ++i) {
          col[0 * nc1 + i] = col[IPR * nc1 + i] / (col[0 * nc1 + i] * rdv[i]);
          col[IPR * nc1 + i] -= psv[i];
        }
        // outflow sides: the ghosts take T and p' of the boundary cell
        // (:113-127)
        if (blk.bc[0] == CG_BC_OUTFLOW)
          for (int i = g.is - NG; i < g.is; ++i) {
            col[0 * nc1 + i] = col[0 * nc1 + g.is];
            col[IPR * nc1 + i] = col[IPR * nc1 + g.is];
          }
        if (blk.bc[1] == CG_BC_OUTFLOW)
          for (int i = g.ie + 1; i <= g.ie + NG; ++i) {
            col[0 * nc1 + i] = col[0 * nc1 + g.ie];
            col[IPR * nc1 + i] = col[IPR * nc1 + g.ie];
          }
      }
      for (int n = 0; n < N; ++n)
        recon_line(cfg.order_v, &col[n * nc1], 1, g.is - 1, g.ie + 1,
                   &wl[n * (nc1 + 1)], &wr[n * (nc1 + 1)]);
      if (decomp) {
        // RestoreFromTemperature faces (:155-164)
        for (int i = g.is; i <= g.ie + 1; ++i) {
          double* wrp = &wr[IPR * (nc1 + 1) + i];
          double* wlp = &wl[IPR * (nc1 + 1) + i];
          *wrp += psf[i];
          if (*wrp < 0.) *wrp = psf[i];
          double gamma = std::sqrt(rdv[i] * rdv[i - 1]);
          wr[0 * (nc1 + 1) + i] = *wrp / (wr[0 * (nc1 + 1) + i] * gamma);
          *wlp += psf[i];
          if (*wlp < 0.) *wlp = psf[i];
          wl[0 * (nc1 + 1) + i] = *wlp / (wl[0 * (nc1 + 1) + i] * gamma);
        }
      }
      check_reconstruct(IPR, wl.data(), wr.data(), nc1 + 1, g.is, g.ie + 1);
      std::vector<double> fx((size_t)N * (nc1 + 1));
      riemann(th, cfg.rsolver, g.IVX, g.is, g.ie + 1, wl.data(), wr.data(),
              nc1 + 1, fx.data(), nc1 + 1);
      for (int n = 0; n < N; ++n)
        for (int i = g.is; i <= g.ie + 1; ++i)
          f1[(((size_t)n * g.nc3 + k) * g.nc2 + j) * (nc1 + 1) + i] =
              fx[n * (nc1 + 1) + i];
    }
  // ---- x2 ----
  if (g.f2) {
    const ptrdiff_t s2 = nc1;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int k = g.ks; k <= g.ke; ++k) {
      int nf = g.nx2 + 1;
      std::vector<double> wl((size_t)N * nf), wr((size_t)N * nf),
          fx((size_t)N * nf), tl(g.nc2 + 1), tr(g.nc2 + 1);
      for (int i = g.is; i <= g.ie; ++i) {
        for (int n = 0; n < N; ++n) {
          recon_line(cfg.order_h, &w[g.at(n, k, 0, i)], s2, g.js - 1, g.je + 1,
                     tl.data(), tr.data());
          for (int j = g.js; j <= g.je + 1; ++j) {
            wl[n * nf + (j - g.js)] = tl[j];
            wr[n * nf + (j - g.js)] = tr[j];
          }
        }
        check_reconstruct(IPR, wl.data(), wr.data(), nf, 0, nf - 1);
        riemann(th, cfg.rsolver, g.IVY, 0, nf - 1, wl.data(), wr.data(), nf,
                fx.data(), nf);
        for (int n = 0; n < N; ++n)
          for (int j = g.js; j <= g.je + 1; ++j)
            f2[(((size_t)n * g.nc3 + k) * (g.nc2 + 1) + j) * nc1 + i] =
                fx[n * nf + (j - g.js)];
      }
    }
  }
  // ---- x3 ----
  if (g.f3) {
    const ptrdiff_t s3 = (ptrdiff_t)nc1 * g.nc2;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int j = g.js; j <= g.je; ++j) {
      int nf = g.nx3 + 1;
      std::vector<double> wl((size_t)N * nf), wr((size_t)N * nf),
          fx((size_t)N * nf), tl(g.nc3 + 1), tr(g.nc3 + 1);
      for (int i = g.is; i <= g.ie; ++i) {
        for (int n = 0; n < N; ++n) {
          recon_line(cfg.order_h, &w[g.at(n, 0, j, i)], s3, g.ks - 1, g.ke + 1,
                     tl.data(), tr.data());
          for (int k = g.ks; k <= g.ke + 1; ++k) {
            wl[n * nf + (k - g.ks)] = tl[k];
            wr[n * nf + (k - g.ks)] = tr[k];
          }
        }
        check_reconstruct(IPR, wl.data(), wr.data(), nf, 0, nf - 1);
        riemann(th, cfg.rsolver, g.IVZ, 0, nf - 1, wl.data(), wr.data(), nf,
                fx.data(), nf);
        for (int n = 0; n < N; ++n)
          for (int k = g.ks; k <= g.ke + 1; ++k)
            f3[(((size_t)n * (g.nc3 + 1) + k) * g.nc2 + j) * nc1 + i] =
                fx[n * nf + (k - g.ks)];
      }
    }
  }
}

// Hydro::AddFluxDivergence (upstream Athena++ add_flux_divergence.cpp, as
// called at src/tasklist/implicit_hydro_tasks.cpp:163-164), Cartesian:
//   dflx = A1 F1(i+1) - A1 F1(i) [+ (A2 F2(j+1) - A2 F2(j))] [+ (A3 ..)]
//   du  -= wght * dflx / vol
void add_flux_divergence(const Grid& g, const cg_block& blk, double wght,
                         const double* f1, const double* f2, const double* f3,
                         double* du) {
  const double A1 = blk.dx2 * blk.dx3, A2 = blk.dx1 * blk.dx3,
               A3 = blk.dx1 * blk.dx2;
  const double vol = blk.dx1 * blk.dx2 * blk.dx3;
  const int nc1 = g.nc1;
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      for (int n = 0; n < g.N; ++n)
        for (int i = g.is; i <= g.ie; ++i) {
          size_t b1 = (((size_t)n * g.nc3 + k) * g.nc2 + j) * (nc1 + 1);
          double dflx = A1 * f1[b1 + i + 1] - A1 * f1[b1 + i];
          if (g.f2) {
            size_t b2 = (((size_t)n * g.nc3 + k) * (g.nc2 + 1) + j) * nc1;
            dflx += A2 * f2[b2 + nc1 + i] - A2 * f2[b2 + i];
          }
          if (g.f3) {
            size_t b3 = (((size_t)n * (g.nc3 + 1) + k) * g.nc2 + j) * nc1;
            dflx += A3 * f3[b3 + (size_t)g.nc2 * nc1 + i] - A3 * f3[b3 + i];
          }
          du[g.at(n, k, j, i)] -= wght * dflx / vol;
        }
}

// temperature of the ideal-moist gas, T = p / (rho Rd feps); snap "W->T"
// (src/interface/eos.hpp:23-27), closed form of Li 2019.
double temperature(const cg_thermo& th, const double* w, int IPR) {
  double feps = 1.;
  for (int n = 1; n <= th.nvapor; ++n) feps += w[n] * th.inv_mu_ratio_m1[n - 1];
  for (int n = 1 + th.nvapor; n <= th.nmass; ++n) feps -= w[n];
  return w[IPR] / (w[0] * th.rd * feps);
}

// Source terms (HydroSourceTerms::AddSourceTerms order: constant
// acceleration, then the user source).
//  gravity: upstream ConstantAcceleration as seen in the context lines of
//    patches/25.constant_acc.patch (the patch itself is disabled,
//    cmake/athena.cmake:4-19): src = dt rho g1; m1 += src; E += src v1.
//  diffusion: examples/2019-Li-snap/straka.cpp:126-190 (x1-x2 plane).
//  heating: config-2 bottom heat flux into the lowest interior layer.
void add_sources(const Grid& g, const cg_block& blk, const cg_thermo& th,
                 const cg_config& cfg, double dt, const double* w,
                 double* du) {
  const int IVX = g.IVX, IVY = g.IVY, IEN = g.IPR;
  const double g1 = blk.grav;
  if (g1 != 0.0) {
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
    for (int k = g.ks; k <= g.ke; ++k)
      for (int j = g.js; j <= g.je; ++j)
        for (int i = g.is; i <= g.ie; ++i) {
          double src = dt * w[g.at(0, k, j, i)] * g1;
          du[g.at(IVX, k, j, i)] += src;
          du[g.at(IEN, k, j, i)] += src * w[g.at(IVX, k, j, i)];
        }
  }
  if (cfg.visc != 0.0 && g.f2) {
    const double K = cfg.visc, p0 = cfg.p0;
    const double dx = blk.dx1, dy = blk.dx2;
    const double cp = th.gammad / (th.gammad - 1.) * th.rd;
    const double chi = (th.gammad - 1.) / th.gammad;
    std::vector<double> theta(g.ncell), temp(g.ncell);
    double wc[5 + CG_MAX_NMASS];
    for (int k = g.ks; k <= g.ke; ++k)
      for (int j = g.js - 1; j <= g.je + 1; ++j)
        for (int i = g.is - 1; i <= g.ie + 1; ++i) {
          for (int n = 0; n < g.N; ++n) wc[n] = w[g.at(n, k, j, i)];
          size_t c = g.at(0, k, j, i);
          temp[c] = temperature(th, wc, g.IPR);
          theta[c] = temp[c] * cg_pow(p0 / wc[g.IPR], chi);
        }
    for (int k = g.ks; k <= g.ke; ++k)
      for (int j = g.js; j <= g.je; ++j)
        for (int i = g.is; i <= g.ie; ++i) {
          auto Wv = [&](int n, int jj, int ii) { return w[g.at(n, k, jj, ii)]; };
          auto T = [&](int jj, int ii) { return theta[g.at(0, k, jj, ii)]; };
          double rho = Wv(0, j, i);
          du[g.at(IVX, k, j, i)] +=
              dt * K * rho / (dx * dy) *
              (Wv(IVX, j, i - 1) + Wv(IVX, j, i + 1) + Wv(IVX, j - 1, i) +
               Wv(IVX, j + 1, i) - 4. * Wv(IVX, j, i));
          du[g.at(IVY, k, j, i)] +=
              dt * K * rho / (dx * dy) *
              (Wv(IVY, j, i - 1) + Wv(IVY, j, i + 1) + Wv(IVY, j - 1, i) +
               Wv(IVY, j + 1, i) - 4. * Wv(IVY, j, i));
          double tc = temp[g.at(0, k, j, i)], thc = T(j, i);
          du[g.at(IEN, k, j, i)] += dt * K * rho / (dx * dy) * cp * tc / thc *
                                    (T(j + 1, i) + T(j - 1, i) + T(j, i + 1) +
                                     T(j, i - 1) - 4. * thc);
        }
  }
  if (cfg.heat_flux != 0.0) {
    for (int k = g.ks; k <= g.ke; ++k)
      for (int j = g.js; j <= g.je; ++j)
        du[g.at(IEN, k, j, g.is)] += dt * cfg.heat_flux / blk.dx1;
  }
}

// ---------------------------------------------------------------------------
// Implicit vertical solver (x1).  Small fixed-size matrices replace Eigen.
template <int B>
struct Mat {
  double a[B][B];
  void zero() {
    for (int r = 0; r < B; ++r)
      for (int c = 0; c < B; ++c) a[r][c] = 0.;
  }
  void ident() {
    zero();
    for (int r = 0; r < B; ++r) a[r][r] = 1.;
  }
};
template <int B>
struct Vec {
  double v[B];
};

template <int B>
Mat<B> mul(const Mat<B>& x, const Mat<B>& y) {
  Mat<B> z;
  for (int r = 0; r < B; ++r)
    for (int c = 0; c < B; ++c) {
      double s = x.a[r][0] * y.a[0][c];
      for (int k = 1; k < B; ++k) s += x.a[r][k] * y.a[k][c];
      z.a[r][c] = s;
    }
  return z;
}
template <int B>
Vec<B> mul(const Mat<B>& x, const Vec<B>& y) {
  Vec<B> z;
  for (int r = 0; r < B; ++r) {
    double s = x.a[r][0] * y.v[0];
    for (int k = 1; k < B; ++k) s += x.a[r][k] * y.v[k];
    z.v[r] = s;
  }
  return z;
}

// Inverse by Gauss-Jordan elimination with partial (row) pivoting: the
// pivot of column c is the first row r >= c of largest |a[r][c]|; the pivot
// row is scaled by the reciprocal of the pivot.  Eigen's .inverse() of the
// reference (forward_backward.hpp:167,171,198) is a partial-pivot LU; the
// two agree to round-off.  Returns false if a pivot is exactly zero.
template <int B>
bool inverse(Mat<B> m, Mat<B>* out) {
  Mat<B> x;
  x.ident();
  for (int c = 0; c < B; ++c) {
    int p = c;
    for (int r = c + 1; r < B; ++r)
      if (std::fabs(m.a[r][c]) > std::fabs(m.a[p][c])) p = r;
    if (m.a[p][c] == 0.0) return false;
    if (p != c)
      for (int q = 0; q < B; ++q) {
        std::swap(m.a[c][q], m.a[p][q]);
        std::swap(x.a[c][q], x.a[p][q]);
      }
    double inv = 1.0 / m.a[c][c];
    for (int q = 0; q < B; ++q) {
      m.a[c][q] = m.a[c][q] * inv;
      x.a[c][q] = x.a[c][q] * inv;
    }
    for (int r = 0; r < B; ++r) {
      if (r == c) continue;
      double f = m.a[r][c];
      for (int q = 0; q < B; ++q) {
        m.a[r][q] = m.a[r][q] - f * m.a[c][q];
        x.a[r][q] = x.a[r][q] - f * x.a[c][q];
      }
    }
  }
  *out = x;
  return true;
}

// Gauss-Jordan elimination of [M | C | t] with partial pivoting (first row
// of largest |pivot|), the pivot row scaled by the pivot's reciprocal and
// the other rows updated by fused multiply-adds: on return C = M^-1 C and
// t = M^-1 t.  Returns false if a pivot is exactly zero.
template <int B>
bool gj_solve(Mat<B> M, Mat<B>& C, Vec<B>& t) {
  for (int cc = 0; cc < B; ++cc) {
    int p = cc;
    double best = std::fabs(M.a[cc][cc]);
    for (int r = cc + 1; r < B; ++r) {
      double v = std::fabs(M.a[r][cc]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (best == 0.0) return false;
    if (p != cc) {
      for (int q = 0; q < B; ++q) {
        std::swap(M.a[cc][q], M.a[p][q]);
        std::swap(C.a[cc][q], C.a[p][q]);
      }
      std::swap(t.v[cc], t.v[p]);
    }
    double inv = 1.0 / M.a[cc][cc];
    for (int q = cc + 1; q < B; ++q) M.a[cc][q] *= inv;
    for (int q = 0; q < B; ++q) C.a[cc][q] *= inv;
    t.v[cc] *= inv;
    for (int r = 0; r < B; ++r) {
      if (r == cc) continue;
      double f = M.a[r][cc];
      for (int q = cc + 1; q < B; ++q)
        M.a[r][q] = std::fma(-f, M.a[cc][q], M.a[r][q]);
      for (int q = 0; q < B; ++q) C.a[r][q] = std::fma(-f, C.a[cc][q], C.a[r][q]);
      t.v[r] = std::fma(-f, t.v[cc], t.v[r]);
    }
  }
  return true;
}

// RoeAverage: src/snap/implicit/flux_decomposition.hpp:34-67
void roe_average(const Grid& g, double* prim, double gm1, const double* wl,
                 const double* wr) {
  const int IVX = g.IVX, IVY = g.IVY, IVZ = g.IVZ, IPR = g.IPR;
  double sqrtdl = std::sqrt(wl[0]);
  double sqrtdr = std::sqrt(wr[0]);
  double isdlpdr = 1.0 / (sqrtdl + sqrtdr);
  prim[0] = sqrtdl * sqrtdr;
  prim[IVX] = (sqrtdl * wl[IVX] + sqrtdr * wr[IVX]) * isdlpdr;
  prim[IVY] = (sqrtdl * wl[IVY] + sqrtdr * wr[IVY]) * isdlpdr;
  prim[IVZ] = (sqrtdl * wl[IVZ] + sqrtdr * wr[IVZ]) * isdlpdr;
  for (int i = 1; i < IVX; ++i)
    prim[i] = (sqrtdl * wl[i] + sqrtdr * wr[i]) * isdlpdr;
  double el = wl[IPR] / gm1 +
              0.5 * wl[0] * (sqr(wl[IVX]) + sqr(wl[IVY]) + sqr(wl[IVZ]));
  double er = wr[IPR] / gm1 +
              0.5 * wr[0] * (sqr(wr[IVX]) + sqr(wr[IVY]) + sqr(wr[IVZ]));
  double hbar = ((el + wl[IPR]) / sqrtdl + (er + wr[IPR]) / sqrtdr) * isdlpdr;
  prim[IPR] =
      (hbar - 0.5 * (sqr(prim[IVX]) + sqr(prim[IVY]) + sqr(prim[IVZ]))) * gm1 /
      (gm1 + 1.) * prim[0];
}

// |A| = R |Lambda| R^-1 at a face: Eigenvalue / Eigenvector
// (flux_decomposition.hpp:69-99).  The product (R |Lambda|) R^-1 is
// evaluated with the zero/one pattern of R and R^-1 spelled out (terms with
// an exact zero factor dropped, unit factors taken exactly), reciprocals of
// 2 hp, hp and 2 cs hp, and fused multiply-adds; equal to the reference's
// dense Eigen product up to round-off.  dir selects the normal velocity
// u = prim[IVX + dir] and the tangential v, w in cyclic order (:80-83).
Mat<5> abs_roe_matrix(const Grid& g, const cg_thermo& th, const double* prim,
                      double gm1, int dir = 0) {
  const int IPR = g.IPR;
  double cs = sound_speed(th, prim, IPR);
  double u = prim[g.IVX + dir], v = prim[g.IVX + (1 + dir) % 3],
         w = prim[g.IVX + (2 + dir) % 3];
  double hp = (gm1 + 1.) / gm1 * prim[IPR] / prim[0];
  double ke = 0.5 * (u * u + v * v + w * w);
  double h = hp + ke;
  double l0 = std::fabs(u - cs), l1 = std::fabs(u), l2 = std::fabs(u + cs);
  double r2h = 1. / (2. * hp), rhp = 1. / hp, rch = 1. / (2. * cs * hp);
  double Ri0[5] = {(cs * ke + u * hp) * rch, (-hp - cs * u) * rch, -v * r2h,
                   -w * r2h, r2h};
  double Ri1[5] = {(hp - ke) * rhp, u * rhp, v * rhp, w * rhp, -rhp};
  double Ri2[5] = {(cs * ke - u * hp) * rch, (hp - cs * u) * rch, -v * r2h,
                   -w * r2h, r2h};
  double RL[5][5] = {{l0, l1, l2, 0., 0.},
                     {(u - cs) * l0, u * l1, (u + cs) * l2, 0., 0.},
                     {v * l0, v * l1, v * l2, l1, 0.},
                     {w * l0, w * l1, w * l2, 0., l1},
                     {(h - u * cs) * l0, ke * l1, (h + u * cs) * l2, v * l1,
                      w * l1}};
  Mat<5> out;
  for (int a = 0; a < 5; ++a)
    for (int c = 0; c < 5; ++c) {
      double s = RL[a][0] * Ri0[c];
      s = std::fma(RL[a][1], Ri1[c], s);
      s = std::fma(RL[a][2], Ri2[c], s);
      if (a >= 2) {
        if (c == 0) s = std::fma(RL[a][3], -v, std::fma(RL[a][4], -w, s));
        if (c == 2) s = s + RL[a][3];
        if (c == 3) s = s + RL[a][4];
      }
      out.a[a][c] = s;
    }
  return out;
}

// FluxJacobian: flux_decomposition.hpp:101-120, in the frame of direction
// dir (v1 normal, v2 and v3 the next components in cyclic order)
Mat<5> flux_jacobian(const Grid& g, double gm1, const double* w, int dir = 0) {
  double v1 = w[g.IVX + dir];
  double v2 = w[g.IVX + (1 + dir) % 3];
  double v3 = w[g.IVX + (2 + dir) % 3];
  double rho = w[0];
  double pres = w[g.IPR];
  double s2 = v1 * v1 + v2 * v2 + v3 * v3;
  double c1 = ((gm1 - 1) * s2 / 2 - (gm1 + 1) / gm1 * pres / rho) * v1;
  double c2 = (gm1 + 1) / gm1 * pres / rho + s2 / 2 - gm1 * v1 * v1;
  double d[25] = {0,
                  1.,
                  0.,
                  0.,
                  0.,
                  gm1 * s2 / 2 - v1 * v1,
                  (2. - gm1) * v1,
                  -gm1 * v2,
                  -gm1 * v3,
                  gm1,
                  -v1 * v2,
                  v2,
                  v1,
                  0.,
                  0.,
                  -v1 * v3,
                  v3,
                  0.,
                  v1,
                  0.,
                  c1,
                  c2,
                  -gm1 * v2 * v1,
                  -gm1 * v3 * v1,
                  (gm1 + 1) * v1};
  Mat<5> m;
  for (int q = 0; q < 25; ++q) m.a[q / 5][q % 5] = d[q];
  return m;
}

template <int B>
Mat<B> sub5(const Mat<5>& m) {
  Mat<B> s;
  if (B == 5) {
    for (int r = 0; r < B; ++r)
      for (int c = 0; c < B; ++c) s.a[r][c] = m.a[r][c];
  } else {
    // PartialCorrection picks rows/cols (idn, ivx, ien)
    // (src/snap/implicit/partial_correction.cpp:80-82, 96-97)
    const int id[3] = {0, 1, 4};
    for (int r = 0; r < B; ++r)
      for (int c = 0; c < B; ++c) s.a[r][c] = m.a[id[r]][id[c]];
  }
  return s;
}

// Cyclic block-tridiagonal solve of PeriodicForwardSweep /
// PeriodicBackwardSubstitution (periodic_forward_backward.hpp:14-232,
// El-Mikkawy 2005) for one line of n >= 3 levels held by a single block
// (first_block = last_block = true):
//   L_p x_{p-1} + d_p x_p + U_p x_{p+1} = r_p,  x_{-1} = x_{n-1}, x_n = x_0,
// with the lower corner U_{-1} (coupling row n-1 to x_0) passed as um1.
// Recurrences (:75-148): ainv_0 = d_0^-1, gamma_0 = L_0, beta_0 =
// um1 ainv_0, zeta_0 = r_0; for 1 <= p <= n-2 with T = L_p ainv_{p-1}
// (the values the reference's loop also forms at p = n-1 are overwritten
// before use):
// ainv_p = (d_p - T U_{p-1})^-1, gamma_p = -T gamma_{p-1},
// beta_p = -(beta_{p-1} U_{p-1}) ainv_p, zeta_p = r_p - T zeta_{p-1};
// then gamma_{n-2} = U_{n-2} - L_{n-2} ainv_{n-3} gamma_{n-3},
// beta_{n-2} = (L_{n-1} - beta_{n-3} U_{n-3}) ainv_{n-2},
// ainv_{n-1} = (d_{n-1} - sum_{p<n-1} beta_p gamma_p)^-1,
// zeta_{n-1} = r_{n-1} - sum_{p<n-1} beta_p zeta_p.  Back substitution
// (:196-213): U_{n-2} = 0, x_{n-1} = ainv_{n-1} zeta_{n-1},
// x_p = ainv_p (zeta_p - U_p x_{p+1} - gamma_p x_{n-1}).
// Products are evaluated left to right as written; inverses by the
// partial-pivot Gauss-Jordan above.  Returns false on a zero pivot.
template <int B>
bool periodic_sweep(int n, const Mat<B>* d, const Mat<B>* L, const Mat<B>* U,
                    const Mat<B>& um1, const Vec<B>* r, Vec<B>* x) {
  std::vector<Mat<B>> ainv(n), gam(n), bet(n);
  std::vector<Vec<B>> zeta(n);
  auto sub = [](const Mat<B>& p, const Mat<B>& q) {
    Mat<B> z;
    for (int a = 0; a < B; ++a)
      for (int c = 0; c < B; ++c) z.a[a][c] = p.a[a][c] - q.a[a][c];
    return z;
  };
  auto neg = [](Mat<B> p) {
    for (int a = 0; a < B; ++a)
      for (int c = 0; c < B; ++c) p.a[a][c] = -p.a[a][c];
    return p;
  };
  auto vsub = [](const Vec<B>& p, const Vec<B>& q) {
    Vec<B> z;
    for (int a = 0; a < B; ++a) z.v[a] = p.v[a] - q.v[a];
    return z;
  };
  if (!inverse(d[0], &ainv[0])) return false;
  gam[0] = L[0];
  bet[0] = mul(um1, ainv[0]);
  zeta[0] = r[0];
  for (int p = 1; p < n - 1; ++p) {
    Mat<B> T = mul(L[p], ainv[p - 1]);
    if (!inverse(sub(d[p], mul(T, U[p - 1])), &ainv[p])) return false;
    gam[p] = neg(mul(T, gam[p - 1]));
    bet[p] = neg(mul(mul(bet[p - 1], U[p - 1]), ainv[p]));
    zeta[p] = vsub(r[p], mul(T, zeta[p - 1]));
  }
  gam[n - 2] = sub(U[n - 2], mul(mul(L[n - 2], ainv[n - 3]), gam[n - 3]));
  bet[n - 2] = mul(sub(L[n - 1], mul(bet[n - 3], U[n - 3])), ainv[n - 2]);
  Mat<B> sbg;
  sbg.zero();
  Vec<B> sbz;
  for (int a = 0; a < B; ++a) sbz.v[a] = 0.0;
  for (int p = 0; p < n - 1; ++p) {
    Mat<B> bg = mul(bet[p], gam[p]);
    Vec<B> bz = mul(bet[p], zeta[p]);
    for (int a = 0; a < B; ++a) {
      for (int c = 0; c < B; ++c) sbg.a[a][c] = sbg.a[a][c] + bg.a[a][c];
      sbz.v[a] = sbz.v[a] + bz.v[a];
    }
  }
  if (!inverse(sub(d[n - 1], sbg), &ainv[n - 1])) return false;
  zeta[n - 1] = vsub(r[n - 1], sbz);
  x[n - 1] = mul(ainv[n - 1], zeta[n - 1]);
  Mat<B> zero;
  zero.zero();
  for (int p = n - 2; p >= 0; --p) {
    const Mat<B>& up = (p == n - 2) ? zero : U[p];
    Vec<B> t = vsub(vsub(zeta[p], mul(up, x[p + 1])), mul(gam[p], x[n - 1]));
    x[p] = mul(ainv[p], t);
  }
  return true;
}

// FullCorrection (src/snap/implicit/full_correction.cpp:32-161) and
// PartialCorrection (partial_correction.cpp:26-170) for one line of
// direction dir through the transverse pair (k, j) in the reference's
// shuffled order (solve_implicit_3d.cpp:28-31, 60-63): x1 lines k = x3,
// j = x2; x2 lines k = x1, j = x3; x3 lines k = x2, j = x1.  The velocities
// enter in the frame of dir (CopyPrimitives / FluxJacobian / Eigenvector,
// flux_decomposition.hpp:14-32, 76-83, 101-110); gravity (Phi) only on x1.
// Non-periodic lines fold the walls and run ForwardSweep /
// BackwardSubstitution (forward_backward.hpp:14-137); lines with both ends
// periodic run periodic_sweep.  Single block, no pole.  On x2 / x3 lines
// only du of IDN, the normal momentum and IEN is written back: the
// reference shuffles only those three out of its line buffer
// (solve_implicit_3d.cpp:51-56, 82-87).
// Geometry of one x1 column of a cubed-sphere panel (GnomonicEquiangle,
// src/exo3/gnomonic_equiangle.cpp:221-231, 407-416): face areas
// dx2f_ang dx3f_ang x1f^2 sin(theta), volumes dx1f dx2f_ang dx3f_ang x1v^2
// sin(theta); the column's tangential velocities are projected onto the
// local orthonormal frame around the solve (solve_implicit_3d.cpp:103-136:
// v2 += v3 cos(theta), v3 *= sin(theta); the projection is undone on w
// afterwards, so a projected copy is used here).
struct LineGeo {
  const double* x1f;  // face radii, nc1 + 1
  const double* x1v;  // cell radii, nc1
  double dx1, d2a, d3a, cosc, sinc;
};

template <int B>
bool correction_line(const Grid& g, const cg_block& blk, const cg_thermo& th,
                     double* du, const double* w, double dt, int dir, int k,
                     int j, bool periodic, const LineGeo* lg = nullptr) {
  const int N = g.N, IVX = g.IVX, IEN = g.IPR;
  int is, ie, nc;
  if (dir == 0) {
    is = g.is; ie = g.ie; nc = g.nc1;
  } else if (dir == 1) {
    is = g.js; ie = g.je; nc = g.nc2;
  } else {
    is = g.ks; ie = g.ke; nc = g.nc3;
  }
  auto at = [&](int n, int i) -> size_t {
    if (dir == 0) return g.at(n, k, j, i);
    if (dir == 1) return g.at(n, j, i, k);
    return g.at(n, i, k, j);
  };
  const int vn = IVX + dir, vt1 = IVX + (1 + dir) % 3, vt2 = IVX + (2 + dir) % 3;
  const double grav = (dir == 0) ? blk.grav : 0.0;
  Mat<B> Phi;
  Phi.zero();
  if (dir == 0) {
    if (B == 5) {
      Phi.a[1][0] = grav;
      Phi.a[4][1] = grav;
    } else {
      Phi.a[1][0] = grav;
      Phi.a[2][1] = grav;
    }
  }
  std::vector<Mat<B>> a(nc), b(nc), c(nc), dfdq(nc);
  std::vector<Vec<B>> delta(nc);
  const double gm1 = th.gammad - 1.;
  double wl[5 + CG_MAX_NMASS], wr[5 + CG_MAX_NMASS], prim[5 + CG_MAX_NMASS];
  auto load = [&](double* dst, int i) {
    for (int n = 0; n < N; ++n) dst[n] = w[at(n, i)];
    if (lg) {
      dst[g.IVY] = dst[g.IVY] + dst[g.IVZ] * lg->cosc;
      dst[g.IVZ] = dst[g.IVZ] * lg->sinc;
    }
  };
  for (int i = is - 2; i <= ie + 1; ++i) {
    load(wr, i);
    dfdq[i] = sub5<B>(flux_jacobian(g, gm1, wr, dir));
  }
  load(wl, is - 2);
  load(wr, is - 1);
  double gm = 0.5 * (gm1 + gm1);
  roe_average(g, prim, gm, wl, wr);
  Mat<B> Am = sub5<B>(abs_roe_matrix(g, th, prim, gm, dir));
  // Cartesian face areas and volume (Athena++ cartesian.cpp Face{1,2,3}Area,
  // CellVolume): x1 faces dx2 dx3, x2 faces dx1 dx3, x3 faces dx1 dx2
  const double area = (dir == 0) ? blk.dx2 * blk.dx3
                                 : (dir == 1) ? blk.dx1 * blk.dx3 : blk.dx1 * blk.dx2;
  const double aleft = area, aright = area;
  const double vol = blk.dx1 * blk.dx2 * blk.dx3;
  // a_i = (A- aleft + A+ aright + (aright - aleft) dfdq_i) / (2 vol)
  //       + I/dt - Phi,  b_i = -(A- + dfdq_{i-1}) aleft / (2 vol),
  //       c_i = -(A+ - dfdq_{i+1}) aright / (2 vol)   (:133-137),
  // with the division by 2 vol taken as a product with its reciprocal.
  const double s2v = 1. / (2. * vol);
  const double sl0 = aleft * s2v, sr0 = aright * s2v, sd0 = (aright - aleft) * s2v;
  const double idt = 1. / dt;
  for (int i = is - 1; i <= ie; ++i) {
    load(wl, i);
    load(wr, i + 1);
    gm = 0.5 * (gm1 + gm1);
    roe_average(g, prim, gm, wl, wr);
    Mat<B> Ap = sub5<B>(abs_roe_matrix(g, th, prim, gm, dir));
    double sl = sl0, sr = sr0, sd = sd0;
    if (lg) {  // GetFace1Area(k, j, i), (k, j, i + 1), GetCellVolume(k, j, i)
      const double al = lg->d2a * lg->d3a * lg->x1f[i] * lg->x1f[i] * lg->sinc;
      const double ar = lg->d2a * lg->d3a * lg->x1f[i + 1] * lg->x1f[i + 1] * lg->sinc;
      const double v = lg->dx1 * lg->d2a * lg->d3a * lg->x1v[i] * lg->x1v[i] * lg->sinc;
      const double s2 = 1. / (2. * v);
      sl = al * s2;
      sr = ar * s2;
      sd = (ar - al) * s2;
    }
    for (int r = 0; r < B; ++r)
      for (int q = 0; q < B; ++q) {
        double av = std::fma(Am.a[r][q], sl,
                             std::fma(Ap.a[r][q], sr, dfdq[i].a[r][q] * sd));
        if (r == q) av += idt;
        a[i].a[r][q] = av - Phi.a[r][q];
        b[i].a[r][q] = -(Am.a[r][q] + dfdq[i - 1].a[r][q]) * sl;
        c[i].a[r][q] = -(Ap.a[r][q] - dfdq[i + 1].a[r][q]) * sr;
      }
    Am = Ap;
  }
  auto rhs_at = [&](int i) {
    Vec<B> rhs;
    double r0 = du[at(0, i)];
    for (int n = 1; n < IVX; ++n) r0 += du[at(n, i)];
    rhs.v[0] = r0 * idt;
    if (B == 5) {
      rhs.v[1] = du[at(vn, i)] * idt;
      rhs.v[2] = du[at(vt1, i)] * idt;
      rhs.v[3] = du[at(vt2, i)] * idt;
      rhs.v[4] = du[at(IEN, i)] * idt;
    } else {
      rhs.v[1] = du[at(vn, i)] * idt;
      rhs.v[2] = du[at(IEN, i)] * idt;
    }
    return rhs;
  };
  // the line buffer's values after the write-back; stored into du below
  std::vector<double> out((size_t)(ie - is + 1) * N);
  if (!periodic) {
    // a_is += b_is Bnd, a_ie += c_ie Bnd with Bnd = diag(1, -1, 1, ...)
    for (int r = 0; r < B; ++r)
      for (int q = 0; q < B; ++q)
        a[is].a[r][q] += (q == 1) ? -b[is].a[r][q] : b[is].a[r][q];
    for (int r = 0; r < B; ++r)
      for (int q = 0; q < B; ++q)
        a[ie].a[r][q] += (q == 1) ? -c[ie].a[r][q] : c[ie].a[r][q];
    // ForwardSweep (forward_backward.hpp:14-84): with M = a_i - b_i a'_{i-1}
    // and t = rhs_i - b_i delta_{i-1}, a'_i = M^-1 c_i and delta_i = M^-1 t.
    // The reference forms M^-1 (Eigen .inverse()) and multiplies; here one
    // Gauss-Jordan elimination of [M | c_i | t] with partial pivoting gives
    // both (equal in exact arithmetic).
    for (int i = is; i <= ie; ++i) {
      Vec<B> t = rhs_at(i);
      Mat<B> M = a[i], C = c[i];
      if (i > is) {
        for (int r = 0; r < B; ++r) {
          for (int q = 0; q < B; ++q) {
            double s = M.a[r][q];
            for (int kk = 0; kk < B; ++kk)
              s = std::fma(-b[i].a[r][kk], a[i - 1].a[kk][q], s);
            M.a[r][q] = s;
          }
          double s = t.v[r];
          for (int kk = 0; kk < B; ++kk)
            s = std::fma(-b[i].a[r][kk], delta[i - 1].v[kk], s);
          t.v[r] = s;
        }
      }
      if (!gj_solve(M, C, t)) return false;
      a[i] = C;
      delta[i] = t;
    }
    // BackwardSubstitution (forward_backward.hpp:86-137)
    for (int i = ie - 1; i >= is; --i) {
      Vec<B> ad = mul(a[i], delta[i + 1]);
      for (int r = 0; r < B; ++r) delta[i].v[r] = delta[i].v[r] - ad.v[r];
    }
    for (int i = is; i <= ie; ++i) {
      double* o = &out[(size_t)(i - is) * N];
      for (int n = 0; n < N; ++n) o[n] = du[at(n, i)];
      double dens = o[0];
      for (int n = 1; n < IVX; ++n) dens += o[n];
      dens = delta[i].v[0] - dens;
      o[0] = delta[i].v[0];
      // the species weights: BackwardSubstitution reads w(n, k, j, i) with
      // the loop indices of the line buffer du_ (forward_backward.hpp:
      // 109-115), which for x2 / x3 lines are shuffled (solve_implicit_3d.cpp:
      // 28-31, 59-62) while w is not: the AthenaArray entry
      // ((n nc3 + k) nc2 + j) nc1 + i of this line's (k, j) and cell i
      // (in range: implicit_solve checks the largest such index)
      for (int n = 1; n < IVX; ++n)
        o[n] += dens * (dir == 0 ? w[at(n, i)] : w[g.at(n, k, j, i)]);
      if (B == 5) {
        o[vn] = delta[i].v[1];
        o[vt1] = delta[i].v[2];
        o[vt2] = delta[i].v[3];
        o[IEN] = delta[i].v[4];
      } else {
        o[vn] = delta[i].v[1];
        o[IEN] = delta[i].v[2];
      }
      for (int n = 1; n < IVX; ++n) o[0] -= o[n];
    }
  } else {
    const int n = ie - is + 1;
    std::vector<Vec<B>> rhs(n), x(n);
    for (int i = is; i <= ie; ++i) rhs[i - is] = rhs_at(i);
    if (!periodic_sweep<B>(n, &a[is], &b[is], &c[is], c[is - 1], rhs.data(),
                           x.data()))
      return false;
    // PeriodicBackwardSubstitution's write-back (:216-229): no species split
    for (int i = is; i <= ie; ++i) {
      double* o = &out[(size_t)(i - is) * N];
      const Vec<B>& d = x[i - is];
      for (int q = 0; q < N; ++q) o[q] = du[at(q, i)];
      o[0] = d.v[0];
      if (B == 5) {
        o[vn] = d.v[1];
        o[vt1] = d.v[2];
        o[vt2] = d.v[3];
        o[IEN] = d.v[4];
      } else {
        o[vn] = d.v[1];
        o[IEN] = d.v[2];
      }
      for (int q = 1; q < IVX; ++q) o[0] -= o[q];
    }
  }
  for (int i = is; i <= ie; ++i) {
    const double* o = &out[(size_t)(i - is) * N];
    if (dir == 0) {
      for (int q = 0; q < N; ++q) du[at(q, i)] = o[q];
    } else {
      du[at(0, i)] = o[0];
      du[at(vn, i)] = o[vn];
      du[at(IEN, i)] = o[IEN];
    }
  }
  return true;
}

// ImplicitSolver::SolveImplicit3D (solve_implicit_3d.cpp:19-143): x3 lines
// (flag bit 2, nx3 > 1), then x2 lines (bit 1, nx2 > 1), then x1 columns
// (bit 0); bit 3 selects the 5 x 5 FullCorrection, else the 3 x 3
// PartialCorrection.  A direction is periodic when both of its boundaries
// are (implicit_solver.cpp:127-134).
int implicit_solve(const Grid& g, const cg_block& blk, const cg_thermo& th,
                   double* du, const double* w, double dt, int flag) {
  if (flag == 0) return CG_OK;
  if (flag & ~15) return CG_ERR_ARG;
  const bool per[3] = {blk.bc[0] == CG_BC_PERIODIC && blk.bc[1] == CG_BC_PERIODIC,
                       blk.bc[2] == CG_BC_PERIODIC && blk.bc[3] == CG_BC_PERIODIC,
                       blk.bc[4] == CG_BC_PERIODIC && blk.bc[5] == CG_BC_PERIODIC};
  // x2 / x3 walls with species: the reference weights the species by w at
  // the shuffled index (forward_backward.hpp:109-115, see correction_line);
  // a block where that index runs past the end of w is rejected
  if (g.nmass > 0 && (((flag & 2) && g.f2 && !per[1] &&
                       g.at(g.nmass, g.ie, g.ke, g.je) >= (size_t)g.N * g.ncell) ||
                      ((flag & 4) && g.f3 && !per[2] &&
                       g.at(g.nmass, g.je, g.ie, g.ke) >= (size_t)g.N * g.ncell)))
    return CG_ERR_ARG;
  int err = CG_OK;
  const int dirs[3] = {2, 1, 0};
  for (int dd = 0; dd < 3; ++dd) {
    const int dir = dirs[dd];
    if (!(flag & (1 << dir))) continue;
    if ((dir == 1 && !g.f2) || (dir == 2 && !g.f3)) continue;
    int k0, k1, j0, j1;
    if (dir == 0) {
      k0 = g.ks; k1 = g.ke; j0 = g.js; j1 = g.je;
    } else if (dir == 1) {
      k0 = g.is; k1 = g.ie; j0 = g.ks; j1 = g.ke;
    } else {
      k0 = g.js; k1 = g.je; j0 = g.is; j1 = g.ie;
    }
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
    for (int k = k0; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j) {
        bool ok = (flag & 8)
                      ? correction_line<5>(g, blk, th, du, w, dt, dir, k, j, per[dir])
                      : correction_line<3>(g, blk, th, du, w, dt, dir, k, j, per[dir]);
        if (!ok) {
#pragma omp atomic write
          err = CG_ERR_PIVOT;
        }
      }
    if (err) return err;
  }
  return err;
}

// ---------------------------------------------------------------------------
// EOS, closed form (src/snap/thermodynamics/thermodynamics_athena.hpp:153-160
// IntEngToPres / PresToIntEng; snap's U->W / W->U are not in the container).
void cons_to_prim_cell(const cg_thermo& th, int nmass, const double* u,
                       double* w) {
  const int IVX = 1 + nmass, IPR = IVX + 3;
  double rho = u[0];
  for (int n = 1; n <= nmass; ++n) rho += u[n];
  for (int n = 1; n <= nmass; ++n) w[n] = u[n] / rho;
  w[0] = rho;
  double v1 = u[IVX] / rho, v2 = u[IVX + 1] / rho, v3 = u[IVX + 2] / rho;
  double ke = 0.5 * (u[IVX] * v1 + u[IVX + 1] * v2 + u[IVX + 2] * v3);
  double gamma = get_gamma(th, w);
  w[IVX] = v1;
  w[IVX + 1] = v2;
  w[IVX + 2] = v3;
  w[IPR] = (gamma - 1.) * (u[IPR] - ke);
}

void prim_to_cons_cell(const cg_thermo& th, int nmass, const double* w,
                       double* u) {
  const int IVX = 1 + nmass, IPR = IVX + 3;
  double rho = w[0];
  double qd = 1.;
  for (int n = 1; n <= nmass; ++n) qd -= w[n];
  u[0] = rho * qd;
  for (int n = 1; n <= nmass; ++n) u[n] = rho * w[n];
  double v1 = w[IVX], v2 = w[IVX + 1], v3 = w[IVX + 2];
  u[IVX] = rho * v1;
  u[IVX + 1] = rho * v2;
  u[IVX + 2] = rho * v3;
  double gamma = get_gamma(th, w);
  u[IPR] = w[IPR] / (gamma - 1.) + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
}

void cons_to_prim(const Grid& g, const cg_thermo& th, const double* u,
                  double* w) {
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      for (int i = g.is; i <= g.ie; ++i) {
        double uc[5 + CG_MAX_NMASS], wc[5 + CG_MAX_NMASS];
        for (int n = 0; n < g.N; ++n) uc[n] = u[g.at(n, k, j, i)];
        cons_to_prim_cell(th, g.nmass, uc, wc);
        for (int n = 0; n < g.N; ++n) w[g.at(n, k, j, i)] = wc[n];
      }
}

void prim_to_cons(const Grid& g, const cg_thermo& th, const double* w,
                  double* u) {
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      for (int i = g.is; i <= g.ie; ++i) {
        double uc[5 + CG_MAX_NMASS], wc[5 + CG_MAX_NMASS];
        for (int n = 0; n < g.N; ++n) wc[n] = w[g.at(n, k, j, i)];
        prim_to_cons_cell(th, g.nmass, wc, uc);
        for (int n = 0; n < g.N; ++n) u[g.at(n, k, j, i)] = uc[n];
      }
}

// Ghost cells of primitives (upstream BoundaryValues: reflecting flips the
// normal velocity; x1 first over interior (j,k), then x2 over all i, then x3
// over all i, j).
void apply_bc(const Grid& g, const cg_block& blk, double* w) {
  const int N = g.N;
  auto fill = [&](int dir, int bc_in, int bc_out, int lo, int hi, auto idx,
                  int vn) {
    // idx(n, c) -> address along direction `dir` at coordinate c
    for (int n = 0; n < N; ++n) {
      double sgn = (n == vn) ? -1.0 : 1.0;
      for (int gh = 1; gh <= NG; ++gh) {
        if (bc_in == CG_BC_REFLECT)
          w[idx(n, lo - gh)] = sgn * w[idx(n, lo + gh - 1)];
        else if (bc_in == CG_BC_PERIODIC)
          w[idx(n, lo - gh)] = w[idx(n, hi - gh + 1)];
        else
          w[idx(n, lo - gh)] = w[idx(n, lo)];
        if (bc_out == CG_BC_REFLECT)
          w[idx(n, hi + gh)] = sgn * w[idx(n, hi - gh + 1)];
        else if (bc_out == CG_BC_PERIODIC)
          w[idx(n, hi + gh)] = w[idx(n, lo + gh - 1)];
        else
          w[idx(n, hi + gh)] = w[idx(n, hi)];
      }
    }
    (void)dir;
  };
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      fill(0, blk.bc[0], blk.bc[1], g.is, g.ie,
           [&](int n, int c) { return g.at(n, k, j, c); }, g.IVX);
  if (g.f2)
    for (int k = g.ks; k <= g.ke; ++k)
      for (int i = 0; i < g.nc1; ++i)
        fill(1, blk.bc[2], blk.bc[3], g.js, g.je,
             [&](int n, int c) { return g.at(n, k, c, i); }, g.IVY);
  if (g.f3)
    for (int j = 0; j < g.nc2; ++j)
      for (int i = 0; i < g.nc1; ++i)
        fill(2, blk.bc[4], blk.bc[5], g.ks, g.ke,
             [&](int n, int c) { return g.at(n, c, j, i); }, g.IVZ);
}

// ---------------------------------------------------------------------------
// Saturation adjustment.  The live one is kintera's ThermoY::forward
// (external, src/tasklist/implicit_hydro_tasks.cpp:300-312); the legacy
// EquilibrateUV (src/snap/thermodynamics/equilibrate_uv.cpp:12-73) fixes
// its semantics: at fixed density and internal energy, move mass between
// vapour and condensate until the vapour is saturated (or no condensate
// is left).  Definition adopted (kintera's exact model is unavailable):
//   sensible internal energy  ie = rho cv_mix(q) T            (p = (g-1) ie)
//   moist internal energy     ie - rho L q_c  is held fixed
//   saturation                q_sat = e_s(T) / (rho R_v T),
//   e_s = p3 exp((1 - t3/T) beta - delta ln(T/t3))  (the "ideal" formula,
//   earth-moist.yaml:37; delta = 0 is Clausius-Clapeyron with L = beta R_v t3)
// Newton on T to |dT| < tol; ie is returned with the latent release added.
struct SatParams {
  double cvd, cvv, cvc, rv, L, t3, p3, beta, delta;
  int iv, ic;
};

SatParams sat_params(const cg_thermo& th) {
  SatParams s;
  s.iv = th.ivap - 1;
  s.ic = th.icld - 1;
  s.cvd = th.rd / (th.gammad - 1.);
  s.cvv = s.cvd * (1. + th.cv_ratio_m1[s.iv]);
  s.cvc = s.cvd * (1. + th.cv_ratio_m1[s.ic]);
  s.rv = th.rd * (1. + th.inv_mu_ratio_m1[s.iv]);
  s.L = th.latent;
  s.t3 = th.t3;
  s.p3 = th.p3;
  s.beta = th.beta;
  s.delta = th.delta;
  return s;
}

double qsat(const SatParams& s, double T, double rho) {
  double es = s.p3 * cg_exp((1. - s.t3 / T) * s.beta - s.delta * cg_log(T / s.t3));
  return es / (rho * s.rv * T);
}

// one cell; q (nmass) in/out, returns iterations or -1 if not converged
int adjust_cell(const cg_thermo& th, const SatParams& s, double rho,
                double* ie, double* q, const double* tguess,
                double* tout = nullptr) {
  double cvd = s.cvd;
  double qd = 1.;
  for (int n = 0; n < th.nmass; ++n) qd -= q[n];
  double cfix = cvd * qd;
  for (int n = 0; n < th.nmass; ++n) {
    if (n == s.iv || n == s.ic) continue;
    cfix += cvd * (1. + th.cv_ratio_m1[n]) * q[n];
  }
  double qt = q[s.iv] + q[s.ic];
  double em = *ie / rho - s.L * q[s.ic];
  double tu = em / (cfix + s.cvv * qt);
  double qc, T;
  int it = 0;
  if (qt <= qsat(s, tu, rho)) {
    qc = 0.;
    T = tu;
  } else {
    T = (tguess != nullptr) ? *tguess : tu;
    bool conv = false;
    for (it = 1; it <= th.max_iter; ++it) {
      double qs = qsat(s, T, rho);
      double dqs = qs * ((s.beta * s.t3 / T - s.delta) / T - 1. / T);
      double qcn = qt - qs;
      double cvm = cfix + s.cvv * qs + s.cvc * qcn;
      double gT = cvm * T - s.L * qcn - em;
      double gp = cvm + dqs * (s.L - (s.cvc - s.cvv) * T);
      double dT = gT / gp;
      T = T - dT;
      if (std::fabs(dT) < th.tol) {
        conv = true;
        break;
      }
    }
    if (!conv) return -1;
    qc = qt - qsat(s, T, rho);
  }
  *ie = *ie + rho * s.L * (qc - q[s.ic]);
  q[s.iv] = qt - qc;
  q[s.ic] = qc;
  if (tout) *tout = T;
  return it;
}

int saturation_adjust(const cg_thermo& th, int ncell, const double* rho,
                      double* ie, double* y, const double* tguess,
                      int* iters) {
  if (th.ivap <= 0 || th.icld <= 0) return CG_OK;
  SatParams s = sat_params(th);
  int err = CG_OK;
  for (int c = 0; c < ncell; ++c) {
    double q[CG_MAX_NMASS];
    for (int n = 0; n < th.nmass; ++n) q[n] = y[(size_t)n * ncell + c];
    int it = adjust_cell(th, s, rho[c], &ie[c], q,
                         tguess ? &tguess[c] : nullptr);
    if (it < 0) err = CG_ERR_CONVERGE;
    for (int n = 0; n < th.nmass; ++n) y[(size_t)n * ncell + c] = q[n];
    if (iters) iters[c] = it;
  }
  return err;
}

// ---------------------------------------------------------------------------
// Training pairs of the saturation-guess network (no reference counterpart:
// the learned initial guess is this project's addition to the adjustment).
//
// Network inputs of a cell (rho, ie, q): the all-vapour temperature T_u (the
// plain Newton start), p_u = rho T_u (R_d q_d + R_v q_t), the total water
// q_t = q_v + q_c and the remaining (precipitating) mass fraction q_p:
//   x = ((T_u - 260) / 40, ln(p_u / 1e5), 40 q_t, 40 q_p).
// Targets: ((T - T_u) / 10, q_v / q_t, q_c / q_t) of the adjusted state.
struct CellFeat {
  double cfix, qt, qp, em, tu;
};

CellFeat cell_feat(const cg_thermo& th, const SatParams& s, double rho,
                   double ie, const double* q, float* x) {
  CellFeat f;
  double qd = 1.;
  for (int n = 0; n < th.nmass; ++n) qd -= q[n];
  f.cfix = s.cvd * qd;
  f.qp = 0.;
  for (int n = 0; n < th.nmass; ++n) {
    if (n == s.iv || n == s.ic) continue;
    f.cfix += s.cvd * (1. + th.cv_ratio_m1[n]) * q[n];
    f.qp += q[n];
  }
  f.qt = q[s.iv] + q[s.ic];
  f.em = ie / rho - s.L * q[s.ic];
  f.tu = f.em / (f.cfix + s.cvv * f.qt);
  double pu = rho * f.tu * (th.rd * qd + s.rv * f.qt);
  x[0] = (float)((f.tu - 260.) / 40.);
  x[1] = (float)cg_log(pu / 1e5);
  x[2] = (float)(40. * f.qt);
  x[3] = (float)(40. * f.qp);
  return f;
}

// splitmix64 (Steele, Lea & Flood 2014): stream of pair p starts at
// seed + 4 p * 0x9E3779B97F4A7C15; uniform doubles from the top 53 bits.
inline uint64_t sm64_next(uint64_t& st) {
  st += 0x9E3779B97F4A7C15ull;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double sm64_unit(uint64_t& st) {
  return (double)(sm64_next(st) >> 11) * 0x1.0p-53;
}

// Pair p: a base cell b = floor(u0 nbase) of the given states, its total
// water scaled by f = 0.8 + 1.7 u1 and put into vapour, its all-vapour
// temperature shifted by 6 u2 - 3 K at fixed density; the plain Newton
// adjustment of that state gives the targets.
int mlp_pairs(const cg_thermo& th, int nbase, const double* rho,
              const double* ie, const double* y, int npairs, uint64_t seed,
              float* x, float* t, int* iters) {
  if (th.ivap <= 0 || th.icld <= 0 || th.nmass > CG_MAX_NMASS) return CG_ERR_ARG;
  SatParams s = sat_params(th);
  int err = CG_OK;
  for (int p = 0; p < npairs; ++p) {
    uint64_t st = seed + 4ull * (uint64_t)p * 0x9E3779B97F4A7C15ull;
    double u0 = sm64_unit(st), u1 = sm64_unit(st), u2 = sm64_unit(st);
    int b = (int)(u0 * nbase);
    if (b >= nbase) b = nbase - 1;
    double q[CG_MAX_NMASS];
    for (int n = 0; n < th.nmass; ++n) q[n] = y[(size_t)n * nbase + b];
    float xb[4];
    CellFeat f0 = cell_feat(th, s, rho[b], ie[b], q, xb);
    double fac = 0.8 + 1.7 * u1;
    double tn = f0.tu + (6. * u2 - 3.);
    q[s.iv] = f0.qt * fac;
    q[s.ic] = 0.;
    double qd = 1.;
    for (int n = 0; n < th.nmass; ++n) qd -= q[n];
    double cfix = s.cvd * qd;
    for (int n = 0; n < th.nmass; ++n) {
      if (n == s.iv || n == s.ic) continue;
      cfix += s.cvd * (1. + th.cv_ratio_m1[n]) * q[n];
    }
    double e = rho[b] * ((cfix + s.cvv * q[s.iv]) * tn);
    CellFeat f = cell_feat(th, s, rho[b], e, q, x + 4 * (size_t)p);
    double T = 0.;
    int it = adjust_cell(th, s, rho[b], &e, q, nullptr, &T);
    if (it < 0) err = CG_ERR_CONVERGE;
    double qt = f.qt;
    t[3 * (size_t)p + 0] = (float)((T - f.tu) / 10.);
    t[3 * (size_t)p + 1] = (float)(qt > 0. ? q[s.iv] / qt : 1.);
    t[3 * (size_t)p + 2] = (float)(qt > 0. ? q[s.ic] / qt : 0.);
    if (iters) iters[p] = it;
  }
  return err;
}

// UpdateAllConserved (implicit_hydro_tasks.cpp:290-313) on interior cells.
int update_all_conserved(const Grid& g, const cg_thermo& th, double* u) {
  if (th.ivap <= 0 || th.icld <= 0) return CG_OK;
  SatParams s = sat_params(th);
  int err = CG_OK;
  const int IVX = g.IVX, IEN = g.IPR;
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      for (int i = g.is; i <= g.ie; ++i) {
        double rho = u[g.at(0, k, j, i)];
        for (int n = 1; n <= g.nmass; ++n) rho += u[g.at(n, k, j, i)];
        double m1 = u[g.at(IVX, k, j, i)], m2 = u[g.at(IVX + 1, k, j, i)],
               m3 = u[g.at(IVX + 2, k, j, i)];
        double ke = 0.5 * (m1 * (m1 / rho) + m2 * (m2 / rho) + m3 * (m3 / rho));
        double ie = u[g.at(IEN, k, j, i)] - ke;
        double q[CG_MAX_NMASS];
        for (int n = 0; n < g.nmass; ++n) q[n] = u[g.at(n + 1, k, j, i)] / rho;
        int it = adjust_cell(th, s, rho, &ie, q, nullptr);
        if (it < 0) {
#pragma omp atomic write
          err = CG_ERR_CONVERGE;
        }
        for (int n = 0; n < g.nmass; ++n) u[g.at(n + 1, k, j, i)] = q[n] * rho;
        u[g.at(IEN, k, j, i)] = ie + ke;
      }
  return err;
}

// ---------------------------------------------------------------------------
// Kessler warm-rain microphysics and sedimentation of the precipitation
// tracer (no counterpart in the reference's live stage: its sedimentation
// hook is commented out at src/tasklist/implicit_hydro_tasks.cpp:292 and the
// rates live in kintera, which is not vendored).  The reactions and rate
// constants are those of examples/2019-Li-snap/earth-precip.yaml:44-53 and
// water-kessler.yaml:4-18 (species dry, H2O, H2O(l), H2O(p)); the rate laws
// and the integration over one step are this project's definition:
//   autoconversion  H2O(l) -> H2O(p):        rate k_auto rho_c
//   accretion       H2O(l) + H2O(p) -> 2 H2O(p): rate k_accr rho_c q_p
//   evaporation     H2O(p) -> H2O:           rate k_evap rho_p (1 - rho_v / rho_s(T))
//                                            while sub-saturated
// applied once per step with the full dt after the saturation adjustment,
// each conversion capped by what is there (and evaporation by the deficit);
// evaporation takes the latent heat L per kg from the total energy (the
// moist internal energy ie - rho L (q_c + q_p) is conserved).  Precipitation
// then falls with speed v_fall (toward -x1) by first-order upwinding
// (the mass flux of src/snap/riemann/noflux.cpp:61-71: velocity times the
// upwind density): a fraction c = min(1, dt v_fall / dx1) of each cell's
// precipitation moves to the cell below, carrying the cell's velocity and
// its own sensible energy cv_p T + |v|^2 / 2, so that the donor keeps its
// temperature and velocity; nothing crosses the walls (the column's water,
// mass, momentum and energy are conserved).  Columns are swept top down.
void kessler_column(const Grid& g, const cg_block& blk, const cg_thermo& th,
                    const cg_config& cfg, const SatParams& s, int ip,
                    double dt, double* u, int k, int j) {
  const int N = g.N, IVX = g.IVX, IEN = g.IPR;
  const int rv_ = 1 + s.iv, rc_ = 1 + s.ic, rp_ = 1 + ip;
  const double cvp = s.cvd * (1. + th.cv_ratio_m1[ip]);
  double cf = dt * cfg.v_fall / blk.dx1;
  if (cf > 1.) cf = 1.;
  double in_p = 0., in_m[3] = {0., 0., 0.}, in_e = 0.;
  for (int i = g.ie; i >= g.is; --i) {
    double uc[5 + CG_MAX_NMASS], wc[5 + CG_MAX_NMASS];
    for (int n = 0; n < N; ++n) uc[n] = u[g.at(n, k, j, i)];
    double rho = uc[0];
    for (int n = 1; n <= g.nmass; ++n) rho += uc[n];
    cons_to_prim_cell(th, g.nmass, uc, wc);
    double T = temperature(th, wc, IEN);
    // conversions (densities of the species; rho is unchanged)
    double rc = uc[rc_], rp = uc[rp_], rv = uc[rv_];
    double a = dt * (cfg.k_auto + cfg.k_accr * (rp / rho)) * rc;
    if (a > rc) a = rc;
    rc = rc - a;
    rp = rp + a;
    double rs = qsat(s, T, rho) * rho;
    double def = rs - rv;
    double b = 0.;
    if (def > 0. && rp > 0.) {
      b = dt * cfg.k_evap * rp * (def / rs);
      if (b > rp) b = rp;
      if (b > def) b = def;
    }
    rp = rp - b;
    rv = rv + b;
    uc[rc_] = rc;
    uc[rp_] = rp;
    uc[rv_] = rv;
    uc[IEN] = uc[IEN] - s.L * b;
    // sedimentation: what leaves this cell downward, then what arrives
    double out_p = 0., out_m[3] = {0., 0., 0.}, out_e = 0.;
    if (i > g.is && cf > 0.) {
      cons_to_prim_cell(th, g.nmass, uc, wc);
      double T2 = temperature(th, wc, IEN);
      double v1 = wc[IVX], v2 = wc[IVX + 1], v3 = wc[IVX + 2];
      out_p = cf * rp;
      out_m[0] = out_p * v1;
      out_m[1] = out_p * v2;
      out_m[2] = out_p * v3;
      out_e = out_p * (cvp * T2 + 0.5 * (v1 * v1 + v2 * v2 + v3 * v3));
    }
    uc[rp_] = (uc[rp_] - out_p) + in_p;
    for (int d = 0; d < 3; ++d) uc[IVX + d] = (uc[IVX + d] - out_m[d]) + in_m[d];
    uc[IEN] = (uc[IEN] - out_e) + in_e;
    for (int n = 0; n < N; ++n) u[g.at(n, k, j, i)] = uc[n];
    in_p = out_p;
    for (int d = 0; d < 3; ++d) in_m[d] = out_m[d];
    in_e = out_e;
  }
}

bool kessler_on(const cg_config& cfg) {
  return cfg.k_auto != 0. || cfg.k_accr != 0. || cfg.k_evap != 0. || cfg.v_fall != 0.;
}

// Species index (0-based) of the precipitation: the species that is
// neither the adjusted vapour nor the adjusted cloud; -1 if none.
int precip_species(const cg_thermo& th) {
  if (th.ivap <= 0 || th.icld <= 0) return -1;
  for (int n = 0; n < th.nmass; ++n)
    if (n != th.ivap - 1 && n != th.icld - 1 && n >= th.nvapor) return n;
  return -1;
}

int kessler(const Grid& g, const cg_block& blk, const cg_thermo& th,
            const cg_config& cfg, double dt, double* u) {
  if (!kessler_on(cfg)) return CG_OK;
  const int ip = precip_species(th);
  if (ip < 0) return CG_ERR_ARG;
  SatParams s = sat_params(th);
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j) kessler_column(g, blk, th, cfg, s, ip, dt, u, k, j);
  return CG_OK;
}

// ---------------------------------------------------------------------------
// One SSPRK(3,3) step of ImplicitHydroTasks
// (src/tasklist/implicit_hydro_tasks.cpp:134-313; weights
//  src/integrator/integrators.cpp:14-44):
//   stage s: F = F(w); u1 = u1 + delta u; u = g1 u + g2 u1;
//            du = -beta dt div F + S(w); du = Implicit(du, w, beta dt);
//            u += du; [last stage: saturation adjustment]; w = W(u) + bcs.
int step(const cg_block& blk, const cg_thermo& th, const cg_config& cfg,
         double dt, int nsteps, double* u, double* w) {
  Grid g(blk);
  const size_t nt = g.N * g.ncell;
  std::vector<double> u1(nt, 0.0), du(nt, 0.0);
  std::vector<double> f1((size_t)g.N * g.nc3 * g.nc2 * (g.nc1 + 1), 0.0);
  std::vector<double> f2((size_t)g.N * g.nc3 * (g.nc2 + 1) * g.nc1, 0.0);
  std::vector<double> f3((size_t)g.N * (g.nc3 + 1) * g.nc2 * g.nc1, 0.0);
  const double sdelta[3] = {1.0, 0.0, 0.0};
  const double sg1[3] = {0.0, 0.25, 2. / 3.};
  const double sg2[3] = {1.0, 0.75, 1. / 3.};
  const double sbeta[3] = {1.0, 0.25, 2. / 3.};
  for (int step = 0; step < nsteps; ++step) {
    std::fill(u1.begin(), u1.end(), 0.0);
    for (int s = 0; s < 3; ++s) {
      calculate_fluxes(g, blk, th, cfg, w, f1.data(), f2.data(), f3.data());
      restore_outflow_ghosts(g, blk, th, w);
      for (int n = 0; n < g.N; ++n)
        for (int k = g.ks; k <= g.ke; ++k)
          for (int j = g.js; j <= g.je; ++j)
            for (int i = g.is; i <= g.ie; ++i) {
              size_t c = g.at(n, k, j, i);
              u1[c] = u1[c] + sdelta[s] * u[c];
              if (s == 0)
                u[c] = u1[c];
              else
                u[c] = sg1[s] * u[c] + sg2[s] * u1[c];
              du[c] = 0.0;
            }
      double wght = sbeta[s] * dt;
      add_flux_divergence(g, blk, wght, f1.data(), f2.data(), f3.data(),
                          du.data());
      add_sources(g, blk, th, cfg, wght, w, du.data());
      int err = implicit_solve(g, blk, th, du.data(), w, wght,
                               cfg.implicit_flag);
      if (err) return err;
      for (int n = 0; n < g.N; ++n)
        for (int k = g.ks; k <= g.ke; ++k)
          for (int j = g.js; j <= g.je; ++j)
            for (int i = g.is; i <= g.ie; ++i) {
              size_t c = g.at(n, k, j, i);
              u[c] = u[c] + du[c];
            }
      if (s == 2 && cfg.satadj != CG_SATADJ_OFF) {
        err = update_all_conserved(g, th, u);
        if (err) return err;
      }
      if (s == 2) {
        err = kessler(g, blk, th, cfg, dt, u);
        if (err) return err;
      }
      cons_to_prim(g, th, u, w);
      apply_bc(g, blk, w);
    }
  }
  return CG_OK;
}

// ---------------------------------------------------------------------------
// Hydro::NewBlockTimeStep (upstream athenapp hydro/new_blockdt.cpp as patched
// by patches/21.new_blockdt.patch:16-45): per interior cell and active
// direction d, dt_d = dx_d / (|v_d| + c_s) with the moist sound speed of
// ideal_moist_hydro.cpp:87-91; a direction solved implicitly (implicit_flag
// bit 0 for x1 when nx2 > 1, bit 1 for x2, bit 2 for x3) uses
// dt_d = dx_d / (cfl |v_d|) instead (patch:31-45).  The block's step is cfl
// times the minimum over cells and active directions (upstream multiplies
// the hyperbolic minimum by cfl_number).  The `cspeed` floor of the patched
// max(cspeed, |v| + c_s) belongs to the athenapp fork, which is not vendored:
// taken as 0 (no configuration sets a floor).
double new_block_dt(const Grid& g, const cg_block& blk, const cg_thermo& th,
                    const cg_config& cfg, const double* w, double cfl) {
  double mind = std::numeric_limits<double>::max();
  const double dx[3] = {blk.dx1, blk.dx2, blk.dx3};
  const bool act[3] = {true, g.f2, g.f3};
  const bool imp[3] = {(cfg.implicit_flag & 1) && g.f2,
                       (cfg.implicit_flag & 2) != 0,
                       (cfg.implicit_flag & 4) != 0};
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      for (int i = g.is; i <= g.ie; ++i) {
        double wi[CG_MAX_NMASS + 5];
        for (int n = 0; n < g.N; ++n) wi[n] = w[g.at(n, k, j, i)];
        const double cs = sound_speed(th, wi, g.IPR);
        for (int d = 0; d < 3; ++d) {
          if (!act[d]) continue;
          const double v = std::fabs(wi[g.IVX + d]);
          const double dtd = imp[d] ? dx[d] / (cfl * v) : dx[d] / (v + cs);
          mind = std::fmin(mind, dtd);
        }
      }
  return cfl * mind;
}

// Mesh::NewTimeStep (upstream athenapp mesh/mesh.cpp): the new step is the
// minimum of twice the old step and every block's step, cut to end exactly at
// tlim; the caller passes the blocks' minimum.
double new_mesh_dt(double dt_old, double blocks_dt, double time, double tlim) {
  double dt = std::fmin(2.0 * dt_old, blocks_dt);
  if (time < tlim && tlim - time < dt) dt = tlim - time;
  return dt;
}

#include "oracle_sphere.inc"

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI of the oracle
using namespace orc;

extern "C" {

double or_exp(double x) { return cg_exp(x); }
double or_log(double x) { return cg_log(x); }
double or_get_gamma(const cg_thermo* th, const double* w) {
  return get_gamma(*th, w);
}

void or_recon_cell(int order, const double* q, double* qm, double* qp) {
  recon_cell(order, q, qm, qp);
}

int or_riemann(const cg_thermo* th, int solver, int ivx, int n,
               const double* wl, const double* wr, ptrdiff_t ld, double* flx,
               ptrdiff_t fld) {
  riemann(*th, solver, ivx, 0, n - 1, wl, wr, ld, flx, fld);
  return CG_OK;
}

int or_decompose_column(const cg_block* blk, const cg_thermo* th,
                        const double* w, int k, int j, double* psf,
                        double* psv, double* rdv) {
  Grid g(*blk);
  decompose_column(g, *blk, *th, w, k, j, psf, psv, rdv);
  return CG_OK;
}

int or_calculate_fluxes(const cg_block* blk, const cg_thermo* th,
                        const cg_config* cfg, const double* w, double* flux1,
                        double* flux2, double* flux3) {
  Grid g(*blk);
  g_face_failures = 0;
  calculate_fluxes(g, *blk, *th, *cfg, w, flux1, flux2, flux3);
  return CG_OK;
}

long long or_face_failures(void) { return g_face_failures.load(); }

int or_restore_outflow_ghosts(const cg_block* blk, const cg_thermo* th,
                              double* w) {
  Grid g(*blk);
  restore_outflow_ghosts(g, *blk, *th, w);
  return CG_OK;
}

int or_add_flux_divergence(const cg_block* blk, double wght,
                           const double* flux1, const double* flux2,
                           const double* flux3, double* du) {
  Grid g(*blk);
  add_flux_divergence(g, *blk, wght, flux1, flux2, flux3, du);
  return CG_OK;
}

int or_add_sources(const cg_block* blk, const cg_thermo* th,
                   const cg_config* cfg, double dt, const double* w,
                   double* du) {
  Grid g(*blk);
  add_sources(g, *blk, *th, *cfg, dt, w, du);
  return CG_OK;
}

int or_implicit_solve(const cg_block* blk, const cg_thermo* th, double* du,
                      const double* w, double dt, int flag) {
  Grid g(*blk);
  return implicit_solve(g, *blk, *th, du, w, dt, flag);
}

int or_cons_to_prim(const cg_block* blk, const cg_thermo* th, const double* u,
                    double* w) {
  Grid g(*blk);
  cons_to_prim(g, *th, u, w);
  return CG_OK;
}

int or_prim_to_cons(const cg_block* blk, const cg_thermo* th, const double* w,
                    double* u) {
  Grid g(*blk);
  prim_to_cons(g, *th, w, u);
  return CG_OK;
}

int or_apply_bc(const cg_block* blk, double* w) {
  Grid g(*blk);
  apply_bc(g, *blk, w);
  return CG_OK;
}

int or_saturation_adjust(const cg_thermo* th, int ncell, const double* rho,
                         double* ie, double* y, const double* t_guess,
                         int* iters) {
  return saturation_adjust(*th, ncell, rho, ie, y, t_guess, iters);
}

int or_mlp_pairs(const cg_thermo* th, int nbase, const double* rho,
                 const double* ie, const double* y, int npairs,
                 unsigned long long seed, float* x, float* t, int* iters) {
  return mlp_pairs(*th, nbase, rho, ie, y, npairs, seed, x, t, iters);
}

int or_mlp_features(const cg_thermo* th, int ncell, const double* rho,
                    const double* ie, const double* y, float* x, double* tu) {
  if (th->ivap <= 0 || th->icld <= 0 || th->nmass > CG_MAX_NMASS) return CG_ERR_ARG;
  SatParams s = sat_params(*th);
  for (int c = 0; c < ncell; ++c) {
    double q[CG_MAX_NMASS];
    for (int n = 0; n < th->nmass; ++n) q[n] = y[(size_t)n * ncell + c];
    CellFeat f = cell_feat(*th, s, rho[c], ie[c], q, x + 4 * (size_t)c);
    if (tu) tu[c] = f.tu;
  }
  return CG_OK;
}

int or_kessler(const cg_block* blk, const cg_thermo* th, const cg_config* cfg,
               double dt, double* u) {
  orc::Grid g(*blk);
  return orc::kessler(g, *blk, *th, *cfg, dt, u);
}

int or_step(const cg_block* blk, const cg_thermo* th, const cg_config* cfg,
            double dt, int nsteps, double* u, double* w) {
  g_threads = 1;
  return step(*blk, *th, *cfg, dt, nsteps, u, w);
}

int or_step_omp(const cg_block* blk, const cg_thermo* th,
                const cg_config* cfg, double dt, int nsteps, double* u,
                double* w, int nthreads) {
#ifdef _OPENMP
  g_threads = nthreads > 0 ? nthreads : omp_get_num_procs();
#else
  g_threads = 1;
#endif
  int r = step(*blk, *th, *cfg, dt, nsteps, u, w);
  g_threads = 1;
  return r;
}

// Block-tridiagonal solve of n levels with B x B blocks (B = 3 or 5):
// b_i x_{i-1} + a_i x_i + c_i x_{i+1} = r_i, row-major blocks.  The same
// forward sweep / back substitution as the column solve.
int or_block_thomas(int B, int n, const double* a, const double* b,
                    const double* c, const double* r, double* x) {
  auto run = [&](auto tag) {
    constexpr int BB = decltype(tag)::value;
    std::vector<Mat<BB>> ap(n);
    std::vector<Vec<BB>> d(n);
    for (int i = 0; i < n; ++i) {
      Mat<BB> M, Bi, C;
      Vec<BB> t;
      for (int q = 0; q < BB * BB; ++q) {
        M.a[q / BB][q % BB] = a[(size_t)i * BB * BB + q];
        Bi.a[q / BB][q % BB] = b[(size_t)i * BB * BB + q];
        C.a[q / BB][q % BB] = c[(size_t)i * BB * BB + q];
      }
      for (int q = 0; q < BB; ++q) t.v[q] = r[(size_t)i * BB + q];
      if (i > 0) {
        for (int rr = 0; rr < BB; ++rr) {
          for (int q = 0; q < BB; ++q) {
            double s = M.a[rr][q];
            for (int kk = 0; kk < BB; ++kk)
              s = std::fma(-Bi.a[rr][kk], ap[i - 1].a[kk][q], s);
            M.a[rr][q] = s;
          }
          double s = t.v[rr];
          for (int kk = 0; kk < BB; ++kk)
            s = std::fma(-Bi.a[rr][kk], d[i - 1].v[kk], s);
          t.v[rr] = s;
        }
      }
      if (!gj_solve(M, C, t)) return CG_ERR_PIVOT;
      ap[i] = C;
      d[i] = t;
    }
    for (int i = n - 2; i >= 0; --i) {
      Vec<BB> ad = mul(ap[i], d[i + 1]);
      for (int q = 0; q < BB; ++q) d[i].v[q] = d[i].v[q] - ad.v[q];
    }
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < BB; ++q) x[(size_t)i * BB + q] = d[i].v[q];
    return CG_OK;
  };
  if (B == 5) return run(std::integral_constant<int, 5>());
  if (B == 3) return run(std::integral_constant<int, 3>());
  return CG_ERR_ARG;
}

// Cyclic block-tridiagonal solve of n >= 3 levels (periodic_sweep):
// b_i x_{i-1} + a_i x_i + c_i x_{i+1} = r_i with x_{-1} = x_{n-1} and
// x_n = x_0; corner couples row n-1 to x_0.  Row-major blocks.
int or_block_thomas_periodic(int B, int n, const double* a, const double* b,
                             const double* c, const double* corner,
                             const double* r, double* x) {
  if (n < 3) return CG_ERR_ARG;
  auto run = [&](auto tag) {
    constexpr int BB = decltype(tag)::value;
    std::vector<Mat<BB>> A(n), L(n), U(n);
    std::vector<Vec<BB>> R(n), X(n);
    Mat<BB> um1;
    for (int i = 0; i < n; ++i) {
      for (int q = 0; q < BB * BB; ++q) {
        A[i].a[q / BB][q % BB] = a[(size_t)i * BB * BB + q];
        L[i].a[q / BB][q % BB] = b[(size_t)i * BB * BB + q];
        U[i].a[q / BB][q % BB] = c[(size_t)i * BB * BB + q];
      }
      for (int q = 0; q < BB; ++q) R[i].v[q] = r[(size_t)i * BB + q];
    }
    for (int q = 0; q < BB * BB; ++q) um1.a[q / BB][q % BB] = corner[q];
    if (!periodic_sweep<BB>(n, A.data(), L.data(), U.data(), um1, R.data(), X.data()))
      return CG_ERR_PIVOT;
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < BB; ++q) x[(size_t)i * BB + q] = X[i].v[q];
    return CG_OK;
  };
  if (B == 5) return run(std::integral_constant<int, 5>());
  if (B == 3) return run(std::integral_constant<int, 3>());
  return CG_ERR_ARG;
}

// Flux Jacobian dF/dU (x1) at primitive state w, and |A| / the signed Roe
// matrix R Lambda R^-1 at the Roe average of (wl, wr), 5 x 5 row-major.
int or_flux_jacobian(const cg_block* blk, double gm1, const double* w,
                     double* out) {
  Grid g(*blk);
  Mat<5> m = flux_jacobian(g, gm1, w);
  for (int q = 0; q < 25; ++q) out[q] = m.a[q / 5][q % 5];
  return CG_OK;
}

int or_roe_matrices(const cg_block* blk, const cg_thermo* th, double gm1,
                    const double* wl, const double* wr, double* abs_a,
                    double* prim_out) {
  Grid g(*blk);
  double prim[5 + CG_MAX_NMASS];
  roe_average(g, prim, gm1, wl, wr);
  Mat<5> m = abs_roe_matrix(g, *th, prim, gm1);
  for (int q = 0; q < 25; ++q) abs_a[q] = m.a[q / 5][q % 5];
  for (int n = 0; n < g.N; ++n) prim_out[n] = prim[n];
  return CG_OK;
}

double or_new_block_dt(const cg_block* blk, const cg_thermo* th,
                       const cg_config* cfg, const double* w, double cfl) {
  Grid g(*blk);
  return new_block_dt(g, *blk, *th, *cfg, w, cfl);
}

double or_new_mesh_dt(double dt_old, double blocks_dt, double time,
                      double tlim) {
  return new_mesh_dt(dt_old, blocks_dt, time, tlim);
}

// Athena++'s default history columns (upstream outputs/history.cpp, not
// vendored): volume-weighted sums of u(IDN), the three momenta, the three
// kinetic energies 0.5 m_d^2 / u(IDN) and u(IEN), followed here by the sums
// of the species densities u(1..nmass).  out: 8 + nmass doubles.
int or_history(const cg_block* blk, const double* u, double* out) {
  Grid g(*blk);
  const double vol = blk->dx1 * blk->dx2 * blk->dx3;
  const int nq = 8 + g.nmass;
  for (int q = 0; q < nq; ++q) out[q] = 0.0;
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      for (int i = g.is; i <= g.ie; ++i) {
        const double d = u[g.at(0, k, j, i)];
        const double m1 = u[g.at(g.IVX, k, j, i)], m2 = u[g.at(g.IVY, k, j, i)],
                     m3 = u[g.at(g.IVZ, k, j, i)];
        out[0] += vol * d;
        out[1] += vol * m1;
        out[2] += vol * m2;
        out[3] += vol * m3;
        out[4] += vol * 0.5 * (m1 * m1) / d;
        out[5] += vol * 0.5 * (m2 * m2) / d;
        out[6] += vol * 0.5 * (m3 * m3) / d;
        out[7] += vol * u[g.at(g.IPR, k, j, i)];
        for (int n = 1; n <= g.nmass; ++n) out[7 + n] += vol * u[g.at(n, k, j, i)];
      }
  return CG_OK;
}

// DebugOutput::WriteOutputFile (src/outputs/debug.cpp:30-84): per variable,
// the minimum and maximum of w over interior cells (initial values FLT_MAX
// and FLT_MIN, strict comparisons, so the first cell in (k, j, i) order wins
// ties) and the cell-centre coordinates (x1, x2, x3) where they occur (left
// at FLT_MAX / FLT_MIN when no cell qualifies).  xmin, xmax: 3 N doubles.
int or_debug(const cg_block* blk, const double* w, double* vmin, double* vmax,
             double* xmin, double* xmax) {
  Grid g(*blk);
  const double fmax = 3.40282346638528859812e+38, fmin = 1.17549435082228750797e-38;
  for (int n = 0; n < g.N; ++n) {
    vmin[n] = fmax;
    vmax[n] = fmin;
    for (int d = 0; d < 3; ++d) {
      xmin[3 * n + d] = fmax;
      xmax[3 * n + d] = fmin;
    }
    for (int k = g.ks; k <= g.ke; ++k)
      for (int j = g.js; j <= g.je; ++j)
        for (int i = g.is; i <= g.ie; ++i) {
          const double v = w[g.at(n, k, j, i)];
          const double x1 = blk->x1min + (i - g.is + 0.5) * blk->dx1;
          const double x2 = blk->x2min + (j - g.js + 0.5) * blk->dx2;
          const double x3 = blk->x3min + (k - g.ks + 0.5) * blk->dx3;
          if (v < vmin[n]) {
            vmin[n] = v;
            xmin[3 * n] = x1;
            xmin[3 * n + 1] = x2;
            xmin[3 * n + 2] = x3;
          }
          if (v > vmax[n]) {
            vmax[n] = v;
            xmax[3 * n] = x1;
            xmax[3 * n + 1] = x2;
            xmax[3 * n + 2] = x3;
          }
        }
  }
  return CG_OK;
}

// State checks (check_hydro_variables, src/snap/eos/check_hydro_variables.
// cpp:21-64; check_eos_cons2prim, src/checks.cpp:52-75) over interior cells
// in (k, j, i) order: returns 0 for the first cell holding a NaN, 1 for a
// negative density, mass fraction or pressure, 2 for a temperature below
// |g| dx1 / Rd (with gravity), 3 when every cell passes; where = (k, j, i)
// interior indices of that cell.
int or_check(const cg_block* blk, const cg_thermo* th, const double* w,
             int* where) {
  Grid g(*blk);
  where[0] = where[1] = where[2] = -1;
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      for (int i = g.is; i <= g.ie; ++i) {
        double wi[CG_MAX_NMASS + 5];
        bool nan = false, neg = false;
        for (int n = 0; n < g.N; ++n) {
          wi[n] = w[g.at(n, k, j, i)];
          if (std::isnan(wi[n])) nan = true;
          if ((n <= g.nmass || n == g.IPR) && wi[n] < 0.0) neg = true;
        }
        int code = 3;
        if (nan)
          code = 0;
        else if (neg)
          code = 1;
        else if (blk->grav != 0.0 &&
                 temperature(*th, wi, g.IPR) < std::fabs(blk->grav) * blk->dx1 / th->rd)
          code = 2;
        if (code < 3) {
          where[0] = k - g.ks;
          where[1] = j - g.js;
          where[2] = i - g.is;
          return code;
        }
      }
  return 3;
}

int or_sums(const cg_block* blk, const double* u, double* out) {
  Grid g(*blk);
  const double vol = blk->dx1 * blk->dx2 * blk->dx3;
  for (int n = 0; n < g.N; ++n) {
    double s = 0.0;
    for (int k = g.ks; k <= g.ke; ++k)
      for (int j = g.js; j <= g.je; ++j)
        for (int i = g.is; i <= g.ie; ++i) s += u[g.at(n, k, j, i)] * vol;
    out[n] = s;
  }
  return CG_OK;
}

// ---- cubed sphere (oracle_sphere.inc) ------------------------------------
int or_sph_step(const cg_block* blk, const cg_thermo* th, const cg_config* cfg,
                const cg_sphere* sp, double dt, int nsteps, double* u, double* w) {
  g_threads = 1;
  return sph_step(*blk, *th, *cfg, *sp, dt, nsteps, u, w);
}

int or_sph_fill_ghosts(const cg_block* blk, const cg_thermo* th, const double* u,
                       double* w) {
  Grid g(*blk);
  if (g.nmass != 0 || g.nx2 != g.nx3) return CG_ERR_ARG;
  std::vector<SphPanel> P;
  for (int p = 1; p <= 6; ++p) P.push_back(sph_panel(*blk, p));
  sph_fill_ghosts(g, P, th->gammad - 1., u, w);
  sph_x1_walls(g, *blk, w);
  return CG_OK;
}

int or_sph_prim_to_cons(const cg_block* blk, const cg_thermo* th, const double* w,
                        double* u) {
  Grid g(*blk);
  if (g.nmass != 0 || g.nx2 != g.nx3) return CG_ERR_ARG;
  const size_t ps = (size_t)g.N * g.ncell;
  for (int p = 1; p <= 6; ++p) {
    SphPanel P = sph_panel(*blk, p);
    for (int k = g.ks; k <= g.ke; ++k)
      for (int j = g.js; j <= g.je; ++j)
        for (int i = g.is; i <= g.ie; ++i) {
          double wc[5], uc[5];
          for (int v = 0; v < 5; ++v) wc[v] = w[(p - 1) * ps + g.at(v, k, j, i)];
          sph_prim_to_cons_cell(th->gammad - 1., P.cosc[P.kj(k, j)], wc, uc);
          for (int v = 0; v < 5; ++v) u[(p - 1) * ps + g.at(v, k, j, i)] = uc[v];
        }
  }
  return CG_OK;
}

/* metric array `what` of panel p over the (k, j) range of the block with
 * ghosts (nc x nc): 0 cos, 1 sin at cells, 2 cos, 3 sin at x2 faces, 4 cos,
 * 5 sin at x3 faces, 6 dx2f_ang, 7 dx3f_ang, 8 dx2f_ang at x3 faces, 9
 * dx3f_ang at x2 faces, 10 latitude, 11 longitude; 12 cell volumes and 13
 * x1-face areas (lower faces) of the interior, (nx3, nx2, nx1). */
int or_sph_metric(const cg_block* blk, int p, int what, double* out) {
  Grid g(*blk);
  SphPanel P = sph_panel(*blk, p);
  const std::vector<double>* a[12] = {&P.cosc, &P.sinc, &P.cosf2, &P.sinf2, &P.cosf3,
                                      &P.sinf3, &P.d2a, &P.d3a, &P.d2af3, &P.d3af2,
                                      &P.lat, &P.lon};
  if (what >= 0 && what < 12) {
    std::copy(a[what]->begin(), a[what]->end(), out);
    return CG_OK;
  }
  size_t o = 0;
  for (int k = g.ks; k <= g.ke; ++k)
    for (int j = g.js; j <= g.je; ++j)
      for (int i = g.is; i <= g.ie; ++i)
        out[o++] = (what == 12) ? sph_vol(P, k, j, i) : sph_area1(P, k, j, i);
  return (what == 12 || what == 13) ? CG_OK : CG_ERR_ARG;
}

int or_sph_forcing(const cg_block* blk, const cg_thermo* th, const cg_sphere* sp, double wght,
                   const double* w6, double* du6) {
  Grid g(*blk);
  if (g.nmass != 0 || g.nx2 != g.nx3) return CG_ERR_ARG;
  const size_t ps = (size_t)g.N * g.ncell;
  for (int p = 1; p <= 6; ++p) {
    SphPanel P = sph_panel(*blk, p);
    sph_forcing(g, P, *th, *sp, p, wght, w6 + (p - 1) * ps, du6 + (p - 1) * ps);
  }
  return CG_OK;
}

int or_sph_link(int p, int s, int* q, int* sq, int* inv) {
  SphLink L = sph_link(p, s);
  *q = L.q;
  *sq = L.sq;
  *inv = L.inv;
  return CG_OK;
}

void or_sph_latlon(int p, double X, double Y, double* lat, double* lon) {
  sph_latlon(p, X, Y, lat, lon);
}

void or_sph_rll_from_abp(double X, double Y, int block, double v2, double v3, double* U,
                         double* V) {
  sph_rll_from_abp(X, Y, block, v2, v3, U, V);
}

void or_sph_abp_from_rll(double X, double Y, int block, double U, double V, double* v2,
                         double* v3) {
  sph_abp_from_rll(X, Y, block, U, V, v2, v3);
}

}  // extern "C"
