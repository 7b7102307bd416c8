// kernels.hpp -- device kernels of the canoe dynamical-core step (gfx950).
//
// FP64 plain path.  Cell- and face-parallel kernels put the unit-stride
// index on adjacent lanes (KGrid::decode), so every wavefront load or store
// of a variable is one contiguous segment in either layout.  Column kernels
// (hydrostatic scan, block Thomas) put one column per lane, j fastest: in
// the level-major layout of a run a wavefront then reads one contiguous row
// per level.
#pragma once

#include "cg_physics.hpp"

namespace cg {

// check_reconstruct (src/checks.cpp:79-97), run after the face states of
// every direction are reconstructed (patches/19.calculate_fluxes.patch:45,
// 73, 117): a negative face density or pressure (pb + the stored pressure,
// pb the base pressure of a perturbation form) sets bit 4 of the status
// word, which the host reports as CG_ERR_NEGATIVE.  NaN passes, as in the
// reference.
constexpr int CG_STATUS_FACE = 4;
template <int NM, class R>
__device__ __forceinline__ void face_check(int* status, const R* wl, const R* wr,
                                           R pb) {
  constexpr int IPR = 4 + NM;
  if (status && (wl[0] < R(0) || wr[0] < R(0) || wl[IPR] + pb < R(0) ||
                 wr[IPR] + pb < R(0)))
    atomicOr(status, CG_STATUS_FACE);
}

// R = Rd (1 + sum_vap q inv_mu_ratio_m1) of the decomposition
// (src/snap/decomposition/change_to_temperature.cpp:100-104)
template <int NM>
CG_HD double decomp_r(const KThermo& th, const double* wc) {
  double feps = 1.;
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n <= th.nvapor) feps += wc[n] * th.inv_mu_ratio_m1[n - 1];
  return th.rd * feps;
}

// log-mean of two face pressures (change_to_temperature.cpp:94-98)
CG_HD double log_mean(double lo, double hi) {
  if (fabs(lo - hi) < 1.E-6) return 0.5 * (lo + hi);
  return (lo - hi) / log_d(lo / hi);
}

// Cell whose temperature and pressure perturbation the decomposition gives
// cell i of a column: a ghost beyond an outflow x1 side takes those of the
// boundary cell (change_to_temperature.cpp:113-127), every other cell its
// own.
CG_HD int decomp_cell(const KGrid& g, int i) {
  if (i < g.is && g.bc[0] == 2) return g.is;
  if (i > g.ie && g.bc[1] == 2) return g.ie;
  return i;
}

// An outflow x1 ghost as the reference's stage reads it after
// CalculateFluxes: RestoreFromTemperature (:150-153) puts the boundary
// cell's T and p' back over the ghost's own hydrostatic pressure,
//   p_g = (p_b - psv_b) + psv_g,  rho_g = p_g / (T_b R_g),
//   T_b = p_b / (rho_b R_b),
// and those values stay in w until the next boundary fill, so the sources
// and the implicit solve of the stage see them.  wb: boundary cell, wg: the
// ghost (density and pressure replaced in place).
template <int NM>
CG_HD void outflow_ghost_w(const KThermo& th, const double* wb, double psv_b,
                           double psv_g, double* wg) {
  constexpr int IPR = 4 + NM;
  const double Tb = wb[IPR] / (wb[0] * decomp_r<NM>(th, wb));
  const double pg = (wb[IPR] - psv_b) + psv_g;
  wg[IPR] = pg;
  wg[0] = pg / (Tb * decomp_r<NM>(th, wg));
}

// ---------------------------------------------------------------------------
// K2: hydrostatic face pressure psf and log-mean cell pressure psv per
// column (Decomposition::ChangeToTemperature, :40-111).  One column per
// lane.
template <int NM>
__device__ __forceinline__ void scan_column(const KGrid& g, const KThermo& th,
                                            const double* __restrict__ w,
                                            double* __restrict__ psf,
                                            double* __restrict__ psv, long long t) {
  constexpr int IPR = 4 + NM;
  int j = g.js + (int)(t % g.nx2);
  int k = g.ks + (int)(t / g.nx2);
  const double grav = g.grav, dx = g.dx1;
  const int is = g.is, ie = g.ie;
  const long long pc = t * g.ps_c;
  auto PF = [&](int i) -> double& { return psf[pc + i * g.ps_i]; };
  {
    double RdTv = w[g.at(IPR, k, j, ie)] / w[g.at(0, k, j, ie)];
    PF(ie + 1) = w[g.at(IPR, k, j, ie)] * exp_d(grav * dx / (2 * RdTv));
  }
  double cur = PF(ie + 1);
  for (int i = ie; i >= is; --i) {
    cur = cur - grav * w[g.at(0, k, j, i)] * dx;
    PF(i) = cur;
  }
  double gb = (g.bc[0] == 0) ? -grav : grav;
  for (int i = is - 1; i >= is - NG; --i) {
    cur = cur - gb * w[g.at(0, k, j, i)] * dx;
    PF(i) = cur;
  }
  double gt = (g.bc[1] == 0) ? -grav : grav;
  cur = PF(ie + 1);
  for (int i = ie + 1; i <= ie + NG; ++i) {
    cur = cur + gt * w[g.at(0, k, j, i)] * dx;
    PF(i + 1) = cur;
  }
  double lo = PF(0);
  for (int i = 0; i < g.nc1; ++i) {
    double hi = PF(i + 1);
    psv[pc + i * g.ps_i] = log_mean(lo, hi);
    lo = hi;
  }
}

template <int NM>
__global__ void __launch_bounds__(256)
    k_hydrostatic_scan(KGrid g, KThermo th, const double* __restrict__ w,
                       double* __restrict__ psf, double* __restrict__ psv) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx2 * g.nx3) return;
  scan_column<NM>(g, th, w, psf, psv, t);
}

// ---------------------------------------------------------------------------
// K1a: x1 face fluxes with the hydrostatic decomposition around the
// reconstruction (patches/19.calculate_fluxes.patch:23-47;
// change_to_temperature.cpp:107-110 and :155-164).  One lane per face.
// One x1 face of item t; a non-null hc is the (k, j) table of cos(theta)
// with which the tangential momentum fluxes are projected to covariant
// components afterwards (cubed-sphere panels, hllc_transform.cpp:235-272).
template <int NM, int ORDER>
__device__ __forceinline__ void flux_x1_face(const KGrid& g, const KThermo& th, int rsolver,
                                             int decomp, const double* __restrict__ w,
                                             const double* __restrict__ psf,
                                             const double* __restrict__ psv,
                                             double* __restrict__ f1, long long t,
                                             const double* __restrict__ hc = nullptr,
                                             int hnc = 0) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
  constexpr int LO = (ORDER == 5) ? 0 : 1, HI = (ORDER == 5) ? 5 : 4;
  int i, j, k;
  g.decode(t, g.nx1 + 1, g.nx2, g.nx3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  const long long pc = g.pcol(k, j) * g.ps_c;
  double q[N][6];
  double rdv[6];
#pragma unroll
  for (int m = LO; m <= HI; ++m) {
    double wc[N];
#pragma unroll
    for (int n = 0; n < N; ++n) wc[n] = w[g.at(n, k, j, i - 3 + m)];
    if (decomp) {
      rdv[m] = decomp_r<NM>(th, wc);
      const int ib = decomp_cell(g, i - 3 + m);
      double pb = wc[IPR], db = wc[0], rb = rdv[m];
      if (ib != i - 3 + m) {  // outflow ghost: the boundary cell's T, p'
        double wb[N];
#pragma unroll
        for (int n = 0; n < N; ++n) wb[n] = w[g.at(n, k, j, ib)];
        pb = wb[IPR];
        db = wb[0];
        rb = decomp_r<NM>(th, wb);
      }
      wc[0] = pb / (db * rb);
      wc[IPR] = pb - psv[pc + ib * g.ps_i];
    }
#pragma unroll
    for (int n = 0; n < N; ++n) q[n][m] = wc[n];
  }
  double wl[N], wr[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double qm, qp;
    if (ORDER == 5) {
      weno5_cell<double>(q[n][0], q[n][1], q[n][2], q[n][3], q[n][4], &qm, &qp);
      wl[n] = qp;
      weno5_cell<double>(q[n][1], q[n][2], q[n][3], q[n][4], q[n][5], &qm, &qp);
      wr[n] = qm;
    } else {
      plm_cell<double>(q[n][1], q[n][2], q[n][3], &qm, &qp);
      wl[n] = qp;
      plm_cell<double>(q[n][2], q[n][3], q[n][4], &qm, &qp);
      wr[n] = qm;
    }
  }
  if (decomp) {
    double ps = psf[pc + i * g.ps_i];
    wr[IPR] += ps;
    if (wr[IPR] < 0.) wr[IPR] = ps;
    double gamma = sqrt(rdv[3] * rdv[2]);
    wr[0] = wr[IPR] / (wr[0] * gamma);
    wl[IPR] += ps;
    if (wl[IPR] < 0.) wl[IPR] = ps;
    wl[0] = wl[IPR] / (wl[0] * gamma);
  }
  face_check<NM, double>(g.status, wl, wr, 0.0);
  double fx[N];
  riemann<NM, 0, double>(th, rsolver, wl, wr, fx);
  if (hc) {
    constexpr int IVY = 2 + NM, IVZ = 3 + NM;
    const double cth = hc[(long long)k * hnc + j];
    const double ty = fx[IVY], tz = fx[IVZ];
    fx[IVY] = ty + tz * cth;
    fx[IVZ] = tz + ty * cth;
  }
#pragma unroll
  for (int n = 0; n < N; ++n) f1[g.f[0].at(n, k, j, i)] = fx[n];
}

template <int NM, int ORDER>
__global__ void __launch_bounds__(256)
    k_flux_x1(KGrid g, KThermo th, int rsolver, int decomp,
              const double* __restrict__ w, const double* __restrict__ psf,
              const double* __restrict__ psv, double* __restrict__ f1) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)(g.nx1 + 1) * g.nx2 * g.nx3) return;
  flux_x1_face<NM, ORDER>(g, th, rsolver, decomp, w, psf, psv, f1, t);
}

// K1b/K1c: x2 and x3 face fluxes.  One lane per face.
template <int NM, int ORDER, int DIR>
__global__ void __launch_bounds__(256)
    k_flux_h(KGrid g, KThermo th, int rsolver, const double* __restrict__ w,
             double* __restrict__ fh) {
  constexpr int N = 5 + NM;
  constexpr int LO = (ORDER == 5) ? 0 : 1, HI = (ORDER == 5) ? 5 : 4;
  const int n2 = (DIR == 1) ? g.nx2 + 1 : g.nx2;
  const int n3 = (DIR == 2) ? g.nx3 + 1 : g.nx3;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * n2 * n3) return;
  int i, j, k;
  g.decode(t, g.nx1, n2, n3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  const long long s = (DIR == 1) ? g.c.j : g.c.k;
  const long long base = g.at(0, k, j, i);
  double q[N][6];
#pragma unroll
  for (int m = LO; m <= HI; ++m)
#pragma unroll
    for (int n = 0; n < N; ++n) q[n][m] = w[base + n * g.c.n + (m - 3) * s];
  double wl[N], wr[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double qm, qp;
    if (ORDER == 5) {
      weno5_cell<double>(q[n][0], q[n][1], q[n][2], q[n][3], q[n][4], &qm, &qp);
      wl[n] = qp;
      weno5_cell<double>(q[n][1], q[n][2], q[n][3], q[n][4], q[n][5], &qm, &qp);
      wr[n] = qm;
    } else {
      plm_cell<double>(q[n][1], q[n][2], q[n][3], &qm, &qp);
      wl[n] = qp;
      plm_cell<double>(q[n][2], q[n][3], q[n][4], &qm, &qp);
      wr[n] = qm;
    }
  }
  face_check<NM, double>(g.status, wl, wr, 0.0);
  double fx[N];
  riemann<NM, DIR, double>(th, rsolver, wl, wr, fx);
#pragma unroll
  for (int n = 0; n < N; ++n) fh[g.f[DIR].at(n, k, j, i)] = fx[n];
}

// Riemann solver on caller-owned face arrays (cg_riemann).
template <int NM, int DIR>
__global__ void __launch_bounds__(256)
    k_riemann_faces(KThermo th, int rsolver, int nfaces,
                    const double* __restrict__ wl, const double* __restrict__ wr,
                    long long ld, double* __restrict__ flx, long long fld) {
  constexpr int N = 5 + NM;
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nfaces) return;
  double a[N], b[N], fx[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    a[n] = wl[n * ld + f];
    b[n] = wr[n * ld + f];
  }
  riemann<NM, DIR, double>(th, rsolver, a, b, fx);
#pragma unroll
  for (int n = 0; n < N; ++n) flx[n * fld + f] = fx[n];
}

// ---------------------------------------------------------------------------
// Temperature of the ideal-moist gas (snap W->T, closed form)
template <int NM>
CG_HD double temperature_d(const KThermo& th, const double* wc) {
  constexpr int IPR = 4 + NM;
  double feps = 1.;
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n <= th.nvapor) feps += wc[n] * th.inv_mu_ratio_m1[n - 1];
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n > th.nvapor) feps -= wc[n];
  return wc[IPR] / (wc[0] * th.rd * feps);
}

struct KSource {
  double wght;                  // beta dt, or beta when dtp is set
  const double* dtp = nullptr;  // device-resident time step of a run
  double visc, p0, cp, chi, heat_flux;
  __device__ double weight() const { return dtp ? wght * dtp[0] : wght; }
};

// Source terms of one interior cell added to d[] in HydroSourceTerms order:
// constant acceleration (upstream form, see patches/25.constant_acc.patch
// context lines), the straka diffusion (examples/2019-Li-snap/straka.cpp:
// 126-190), the bottom heating.  psv: the stage's cell pressures (null
// without gravity); the diffusion then reads ghosts beyond an outflow x1
// side as CalculateFluxes left them (outflow_ghost_w).
template <int NM>
__device__ inline void add_sources_cell(const KGrid& g, const KThermo& th,
                                        const KSource& src,
                                        const double* __restrict__ w,
                                        const double* __restrict__ psv, int k,
                                        int j, int i, double* d) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IVY = IVX + 1, IEN = 4 + NM;
  const long long c0 = g.at(0, k, j, i);
  if (g.grav != 0.0) {
    double s = src.weight() * w[c0] * g.grav;
    d[IVX] += s;
    d[IEN] += s * w[c0 + IVX * g.c.n];
  }
  if (src.visc != 0.0 && g.f2) {
    const double K = src.visc, dx = g.dx1, dy = g.dx2, dt = src.weight();
    auto W = [&](int n, int jj, int ii) { return w[g.at(n, k, jj, ii)]; };
    auto theta = [&](int jj, int ii, double* tmp) {
      double wc[N];
#pragma unroll
      for (int n = 0; n < N; ++n) wc[n] = W(n, jj, ii);
      const int ib = decomp_cell(g, ii);
      if (psv && g.grav != 0.0 && ib != ii) {
        double wb[N];
#pragma unroll
        for (int n = 0; n < N; ++n) wb[n] = W(n, jj, ib);
        const long long pc = g.pcol(k, jj) * g.ps_c;
        outflow_ghost_w<NM>(th, wb, psv[pc + ib * g.ps_i],
                            psv[pc + ii * g.ps_i], wc);
      }
      double T = temperature_d<NM>(th, wc);
      if (tmp) *tmp = T;
      return T * pow_d(src.p0 / wc[4 + NM], src.chi);
    };
    double rho = W(0, j, i);
    d[IVX] += dt * K * rho / (dx * dy) *
              (W(IVX, j, i - 1) + W(IVX, j, i + 1) + W(IVX, j - 1, i) +
               W(IVX, j + 1, i) - 4. * W(IVX, j, i));
    d[IVY] += dt * K * rho / (dx * dy) *
              (W(IVY, j, i - 1) + W(IVY, j, i + 1) + W(IVY, j - 1, i) +
               W(IVY, j + 1, i) - 4. * W(IVY, j, i));
    double tc;
    double thc = theta(j, i, &tc);
    double tjp = theta(j + 1, i, nullptr), tjm = theta(j - 1, i, nullptr);
    double tip = theta(j, i + 1, nullptr), tim = theta(j, i - 1, nullptr);
    d[IEN] += dt * K * rho / (dx * dy) * src.cp * tc / thc *
              (tjp + tjm + tip + tim - 4. * thc);
  }
  if (src.heat_flux != 0.0 && i == g.is)
    d[IEN] += src.weight() * src.heat_flux / g.dx1;
}

// K4a: du = 0 - wght div F (Hydro::AddFluxDivergence) plus the sources.
template <int NM>
__global__ void __launch_bounds__(256)
    k_divergence(KGrid g, KThermo th, KSource src,
                 const double* __restrict__ f1, const double* __restrict__ f2,
                 const double* __restrict__ f3, const double* __restrict__ w,
                 const double* __restrict__ psv, double* __restrict__ du) {
  constexpr int N = 5 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return;
  int i, j, k;
  g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  const double A1 = g.dx2 * g.dx3, A2 = g.dx1 * g.dx3, A3 = g.dx1 * g.dx2;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  const double wght = src.weight();
  double d[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double dflx = A1 * f1[g.f[0].at(n, k, j, i + 1)] - A1 * f1[g.f[0].at(n, k, j, i)];
    if (g.f2)
      dflx += A2 * f2[g.f[1].at(n, k, j + 1, i)] - A2 * f2[g.f[1].at(n, k, j, i)];
    if (g.f3)
      dflx += A3 * f3[g.f[2].at(n, k + 1, j, i)] - A3 * f3[g.f[2].at(n, k, j, i)];
    d[n] = 0.0 - wght * dflx / vol;
  }
  add_sources_cell<NM>(g, th, src, w, psv, k, j, i, d);
#pragma unroll
  for (int n = 0; n < N; ++n) du[g.at(n, k, j, i)] = d[n];
}

// Flux divergence only, accumulated into a caller du (cg_add_flux_divergence)
template <int NM>
__global__ void __launch_bounds__(256)
    k_add_divergence(KGrid g, double wght, const double* __restrict__ f1,
                     const double* __restrict__ f2,
                     const double* __restrict__ f3, double* __restrict__ du) {
  constexpr int N = 5 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return;
  int i, j, k;
  g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  const double A1 = g.dx2 * g.dx3, A2 = g.dx1 * g.dx3, A3 = g.dx1 * g.dx2;
  const double vol = g.dx1 * g.dx2 * g.dx3;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double dflx = A1 * f1[g.f[0].at(n, k, j, i + 1)] - A1 * f1[g.f[0].at(n, k, j, i)];
    if (g.f2)
      dflx += A2 * f2[g.f[1].at(n, k, j + 1, i)] - A2 * f2[g.f[1].at(n, k, j, i)];
    if (g.f3)
      dflx += A3 * f3[g.f[2].at(n, k + 1, j, i)] - A3 * f3[g.f[2].at(n, k, j, i)];
    du[g.at(n, k, j, i)] -= wght * dflx / vol;
  }
}

// ---------------------------------------------------------------------------
// K3: implicit x1 column solve (FullCorrection B = 5 / PartialCorrection
// B = 3), block Thomas with a partial-pivot Gauss-Jordan inverse.
template <int B>
struct Mt {
  double a[B][B];
};
template <int B>
struct Vt {
  double v[B];
};

template <int B>
__device__ inline void mmul(const Mt<B>& x, const Mt<B>& y, Mt<B>& z) {
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int c = 0; c < B; ++c) {
      double s = x.a[r][0] * y.a[0][c];
#pragma unroll
      for (int k = 1; k < B; ++k) s += x.a[r][k] * y.a[k][c];
      z.a[r][c] = s;
    }
}
template <int B>
__device__ inline void mvec(const Mt<B>& x, const Vt<B>& y, Vt<B>& z) {
#pragma unroll
  for (int r = 0; r < B; ++r) {
    double s = x.a[r][0] * y.v[0];
#pragma unroll
    for (int k = 1; k < B; ++k) s += x.a[r][k] * y.v[k];
    z.v[r] = s;
  }
}

// Gauss-Jordan with partial pivoting (first row of largest |pivot|), pivot
// row scaled by the reciprocal; row exchanges by register selects, run only
// when a lane of the wavefront pivots (a uniform branch).  Columns of m left
// of the pivot column are not updated: neither x nor any later pivot or
// factor reads them, so x equals the full-row elimination bit for bit.
template <int B>
__device__ inline bool minv(Mt<B> m, Mt<B>& x) {
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int c = 0; c < B; ++c) x.a[r][c] = (r == c) ? 1.0 : 0.0;
  bool ok = true;
#pragma unroll
  for (int c = 0; c < B; ++c) {
    int p = c;
    double best = fabs(m.a[c][c]);
#pragma unroll
    for (int r = c + 1; r < B; ++r) {
      double v = fabs(m.a[r][c]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (best == 0.0) ok = false;
    if (__any(p != c)) {
#pragma unroll
      for (int r = c + 1; r < B; ++r) {
        bool sw = (p == r);
#pragma unroll
        for (int q = 0; q < B; ++q) {
          if (q >= c) {
            double mc = m.a[c][q], mr = m.a[r][q];
            m.a[c][q] = sw ? mr : mc;
            m.a[r][q] = sw ? mc : mr;
          }
          double xc = x.a[c][q], xr = x.a[r][q];
          x.a[c][q] = sw ? xr : xc;
          x.a[r][q] = sw ? xc : xr;
        }
      }
    }
    double inv = 1.0 / m.a[c][c];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      if (q >= c) m.a[c][q] = m.a[c][q] * inv;
      x.a[c][q] = x.a[c][q] * inv;
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
      if (r == c) continue;
      double f = m.a[r][c];
#pragma unroll
      for (int q = 0; q < B; ++q) {
        if (q >= c) m.a[r][q] = m.a[r][q] - f * m.a[c][q];
        x.a[r][q] = x.a[r][q] - f * x.a[c][q];
      }
    }
  }
  return ok;
}

// RoeAverage (src/snap/implicit/flux_decomposition.hpp:34-67)
template <int NM>
__device__ inline void roe_average(double* prim, double gm1, const double* wl,
                                   const double* wr) {
  constexpr int IVX = 1 + NM, IVY = IVX + 1, IVZ = IVX + 2, IPR = IVX + 3;
  double sqrtdl = sqrt(wl[0]);
  double sqrtdr = sqrt(wr[0]);
  double isdlpdr = 1.0 / (sqrtdl + sqrtdr);
  prim[0] = sqrtdl * sqrtdr;
  prim[IVX] = (sqrtdl * wl[IVX] + sqrtdr * wr[IVX]) * isdlpdr;
  prim[IVY] = (sqrtdl * wl[IVY] + sqrtdr * wr[IVY]) * isdlpdr;
  prim[IVZ] = (sqrtdl * wl[IVZ] + sqrtdr * wr[IVZ]) * isdlpdr;
#pragma unroll
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

// |A| = (R |Lambda|) R^-1 restricted to the B variables of the correction
// (flux_decomposition.hpp:69-99; rows/cols (idn, ivx, ien) for B = 3,
// partial_correction.cpp:94-97).  R |Lambda| is R with column c scaled by
// |lambda_c|: the diagonal product of the reference adds only exact zeros,
// so every nonzero entry is the same number.
template <int NM, int B>
__device__ inline void abs_roe(const KThermo& th, const double* prim,
                               double gm1, Mt<B>& out) {
  constexpr int IVX = 1 + NM, IPR = IVX + 3;
  double cs = sound_speed<NM, double>(th, prim);
  double r = prim[0];
  double u = prim[IVX];
  double v = prim[IVX + 1];
  double w = prim[IVX + 2];
  double p = prim[IPR];
  double lam[5] = {fabs(u - cs), fabs(u), fabs(u + cs), fabs(u), fabs(u)};
  double ke = 0.5 * (u * u + v * v + w * w);
  double hp = (gm1 + 1.) / gm1 * p / r;
  double h = hp + ke;
  double R[5][5] = {{1., 1., 1., 0., 0.},
                    {u - cs, u, u + cs, 0., 0.},
                    {v, v, v, 1., 0.},
                    {w, w, w, 0., 1.},
                    {h - u * cs, ke, h + u * cs, v, w}};
  double Ri[5][5] = {{(cs * ke + u * hp) / (2. * cs * hp),
                      (-hp - cs * u) / (2. * cs * hp), -v / (2. * hp),
                      -w / (2. * hp), 1. / (2. * hp)},
                     {(hp - ke) / hp, u / hp, v / hp, w / hp, -1. / hp},
                     {(cs * ke - u * hp) / (2. * cs * hp),
                      (hp - cs * u) / (2. * cs * hp), -v / (2. * hp),
                      -w / (2. * hp), 1. / (2. * hp)},
                     {-v, 0., 1., 0., 0.},
                     {-w, 0., 0., 1., 0.}};
  double RL[5][5];
#pragma unroll
  for (int rr = 0; rr < 5; ++rr)
#pragma unroll
    for (int c = 0; c < 5; ++c) RL[rr][c] = R[rr][c] * lam[c];
  const int id3[3] = {0, 1, 4};
#pragma unroll
  for (int a = 0; a < B; ++a)
#pragma unroll
    for (int c = 0; c < B; ++c) {
      int ra = (B == 5) ? a : id3[a];
      int cc = (B == 5) ? c : id3[c];
      double s = RL[ra][0] * Ri[0][cc];
#pragma unroll
      for (int k = 1; k < 5; ++k) s += RL[ra][k] * Ri[k][cc];
      out.a[a][c] = s;
    }
}

// FluxJacobian (flux_decomposition.hpp:101-120), dir = X1
template <int NM, int B>
__device__ inline void flux_jac(double gm1, const double* w, Mt<B>& out) {
  constexpr int IVX = 1 + NM, IPR = IVX + 3;
  double v1 = w[IVX];
  double v2 = w[IVX + 1];
  double v3 = w[IVX + 2];
  double rho = w[0];
  double pres = w[IPR];
  double s2 = v1 * v1 + v2 * v2 + v3 * v3;
  double c1 = ((gm1 - 1) * s2 / 2 - (gm1 + 1) / gm1 * pres / rho) * v1;
  double c2 = (gm1 + 1) / gm1 * pres / rho + s2 / 2 - gm1 * v1 * v1;
  double d[5][5] = {{0, 1., 0., 0., 0.},
                    {gm1 * s2 / 2 - v1 * v1, (2. - gm1) * v1, -gm1 * v2,
                     -gm1 * v3, gm1},
                    {-v1 * v2, v2, v1, 0., 0.},
                    {-v1 * v3, v3, 0., v1, 0.},
                    {c1, c2, -gm1 * v2 * v1, -gm1 * v3 * v1, (gm1 + 1) * v1}};
  const int id3[3] = {0, 1, 4};
#pragma unroll
  for (int a = 0; a < B; ++a)
#pragma unroll
    for (int c = 0; c < B; ++c)
      out.a[a][c] = d[(B == 5) ? a : id3[a]][(B == 5) ? c : id3[c]];
}

struct KImplicit {
  double dt;
  long long ncol;  // interior columns nx2 nx3
};

// |A| at a face, evaluated with the zero/one structure of R and R^-1 of
// flux_decomposition.hpp:91-98 spelled out (terms whose factor is an exact
// zero are dropped, unit factors taken exactly) and fused multiply-adds.
template <int NM, int B>
__device__ inline void abs_roe_fma(const KThermo& th, const double* prim,
                                   double gm1, Mt<B>& out) {
  constexpr int IVX = 1 + NM, IPR = IVX + 3;
  double cs = sound_speed<NM, double>(th, prim);
  double u = prim[IVX], v = prim[IVX + 1], w = prim[IVX + 2];
  double hp = (gm1 + 1.) / gm1 * prim[IPR] / prim[0];
  double ke = 0.5 * (u * u + v * v + w * w);
  double h = hp + ke;
  double l0 = fabs(u - cs), l1 = fabs(u), l2 = fabs(u + cs);
  double r2h = 1. / (2. * hp), rhp = 1. / hp, rch = 1. / (2. * cs * hp);
  // rows of R^-1 (0..2 full; 3 = (-v, 0, 1, 0, 0); 4 = (-w, 0, 0, 1, 0))
  double Ri0[5] = {(cs * ke + u * hp) * rch, (-hp - cs * u) * rch, -v * r2h,
                   -w * r2h, r2h};
  double Ri1[5] = {(hp - ke) * rhp, u * rhp, v * rhp, w * rhp, -rhp};
  double Ri2[5] = {(cs * ke - u * hp) * rch, (hp - cs * u) * rch, -v * r2h,
                   -w * r2h, r2h};
  // rows of R |Lambda|
  double RL[5][5] = {{l0, l1, l2, 0., 0.},
                     {(u - cs) * l0, u * l1, (u + cs) * l2, 0., 0.},
                     {v * l0, v * l1, v * l2, l1, 0.},
                     {w * l0, w * l1, w * l2, 0., l1},
                     {(h - u * cs) * l0, ke * l1, (h + u * cs) * l2, v * l1,
                      w * l1}};
  const int id3[3] = {0, 1, 4};
#pragma unroll
  for (int a = 0; a < B; ++a) {
    const int ra = (B == 5) ? a : id3[a];
#pragma unroll
    for (int c = 0; c < B; ++c) {
      const int cc = (B == 5) ? c : id3[c];
      double s = RL[ra][0] * Ri0[cc];
      s = fma(RL[ra][1], Ri1[cc], s);
      s = fma(RL[ra][2], Ri2[cc], s);
      if (ra >= 2) {  // RL[.][3], RL[.][4] vanish for rows 0, 1
        if (cc == 0) s = fma(RL[ra][3], -v, fma(RL[ra][4], -w, s));
        if (cc == 2) s = s + RL[ra][3];
        if (cc == 3) s = s + RL[ra][4];
      }
      out.a[a][c] = s;
    }
  }
}

// Forward sweep of one column per lane: assembles a_i, b_i, c_i
// (full_correction.cpp:221-288), folds the reflecting walls (:286-288) and
// runs ForwardSweep (forward_backward.hpp:14-84), archiving a'_i (B x B)
// and delta_i (B) per level, lane-contiguous: arch[(lvl*(B*B+B)+e)*ncol+col].
// a'_i = M^-1 c_i and delta_i = M^-1 (rhs_i - b_i delta_{i-1}) with
// M = a_i - b_i a'_{i-1} come from one Gauss-Jordan elimination of
// [M | c_i | rhs] with partial pivoting (the reference forms M^-1 and then
// multiplies; the two agree to round-off).
// One column: c0 = offset of (n = 0, k, j, i = 0) in w / du; the archive of
// level lvl, entry e goes to arch[(lvl (B*B+B) + e) as + ao].  psvr (runs
// with gravity, else null) and pc: the stage's cell pressures psv of the
// column, with which ghosts beyond an outflow x1 side are read as the
// reference's CalculateFluxes left them (outflow_ghost_w).  Returns false on
// a zero pivot.
// Column of a cubed-sphere panel (GEO): face areas dx2f_ang dx3f_ang x1f^2
// sin(theta) and volumes dx1 dx2f_ang dx3f_ang x1v^2 sin(theta) per level,
// tangential velocities projected onto the orthonormal frame
// (partial_correction.cpp:113-137, solve_implicit_3d.cpp:103-136).
struct KColGeo {
  const double* x1f;
  const double* x1v;
  double dx1, d2a, d3a, cosc, sinc;
};

template <int NM, int B, bool GEO = false>
__device__ __forceinline__ bool fwd_column(const KGrid& g, const KThermo& th,
                                           double dt, const double* __restrict__ w,
                                           const double* __restrict__ du,
                                           long long c0, double* __restrict__ arch,
                                           long long as, long long ao,
                                           const double* __restrict__ psvr = nullptr,
                                           long long pc = 0,
                                           const KColGeo* geo = nullptr) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IEN = 4 + NM;
  constexpr int AS = B * B + B;
  const int is = g.is, ie = g.ie;
  const double grav = g.grav;
  const double gm1 = th.gammad - 1.;
  const double gm = 0.5 * (gm1 + gm1);
  const double aleft = g.dx2 * g.dx3, aright = g.dx2 * g.dx3;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  const double s2v = 1. / (2. * vol);
  const double sl0 = aleft * s2v, sr0 = aright * s2v, sd0 = (aright - aleft) * s2v;
  const double idt = 1. / dt;
  const long long sn = g.c.n, si = g.c.i;
  auto load = [&](double* dst, int i) {
#pragma unroll
    for (int n = 0; n < N; ++n) dst[n] = w[c0 + n * sn + i * si];
    if (GEO) {
      dst[IVX + 1] = dst[IVX + 1] + dst[IVX + 2] * geo->cosc;
      dst[IVX + 2] = dst[IVX + 2] * geo->sinc;
    }
  };
  double wm[N], wc[N], wp[N], prim[N];
  load(wm, is - 1);
  load(wc, is);
  const bool rw = psvr != nullptr && grav != 0.0;
  if (rw && g.bc[0] == 2)
    outflow_ghost_w<NM>(th, wc, psvr[pc + is * g.ps_i],
                        psvr[pc + (is - 1) * g.ps_i], wm);
  Mt<B> Am, Ap, dm, dc, dp;
  roe_average<NM>(prim, gm, wm, wc);
  abs_roe_fma<NM, B>(th, prim, gm, Am);
  flux_jac<NM, B>(gm1, wm, dm);
  flux_jac<NM, B>(gm1, wc, dc);
  Mt<B> ap;  // a'_{i-1}
  Vt<B> dl;  // delta_{i-1}
  bool ok = true;
  // the next level's w and this level's du (and, on a panel, the next
  // level's radii) are loaded one level ahead, so that their latency
  // overlaps the block algebra of the current level
  double wn[N], dun[N];
  auto load_du = [&](double* dst, int i) {
#pragma unroll
    for (int n = 0; n < N; ++n) dst[n] = du[c0 + n * sn + i * si];
  };
  load(wn, is + 1);
  load_du(dun, is);
  double xf0 = 0.0, xf1 = 0.0, xv = 0.0, nxf = 0.0, nxv = 0.0;
  if (GEO) {
    xf0 = geo->x1f[is];
    xf1 = geo->x1f[is + 1];
    xv = geo->x1v[is];
  }
  // archive of level lvl: a'_lvl and delta_lvl
  auto archive = [&](int lvl) {
    const long long lb = (long long)(lvl - is) * AS;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int q = 0; q < B; ++q)
        __builtin_nontemporal_store(ap.a[r][q], &arch[(lb + r * B + q) * as + ao]);
#pragma unroll
    for (int r = 0; r < B; ++r)
      __builtin_nontemporal_store(dl.v[r], &arch[(lb + B * B + r) * as + ao]);
  };
  for (int i = is; i <= ie; ++i) {
    double dui[N];
#pragma unroll
    for (int n = 0; n < N; ++n) {
      wp[n] = wn[n];
      dui[n] = dun[n];
    }
    // The previous level's archive and the next level's loads are issued
    // together, right after this level's data have arrived: a wait for
    // loaded data also waits for every store issued before it, so each
    // gets a whole level of block algebra to complete.  The loads run at
    // every level (the last reloads levels ie + 1 and ie, unused) so that no
    // branch joins before their values are read.
    if (i > is) archive(i - 1);
    load(wn, min(i + 2, ie + 1));
    load_du(dun, min(i + 1, ie));
    if (GEO) {
      nxf = geo->x1f[min(i + 2, ie + 1)];
      nxv = geo->x1v[min(i + 1, ie)];
    }
    if (i == ie && rw && g.bc[1] == 2)
      outflow_ghost_w<NM>(th, wc, psvr[pc + ie * g.ps_i],
                          psvr[pc + (ie + 1) * g.ps_i], wp);
    roe_average<NM>(prim, gm, wc, wp);
    abs_roe_fma<NM, B>(th, prim, gm, Ap);
    flux_jac<NM, B>(gm1, wp, dp);
    double sl = sl0, sr = sr0, sd = sd0;
    if (GEO) {
      const double al = geo->d2a * geo->d3a * xf0 * xf0 * geo->sinc;
      const double ar = geo->d2a * geo->d3a * xf1 * xf1 * geo->sinc;
      const double v = geo->dx1 * geo->d2a * geo->d3a * xv * xv * geo->sinc;
      const double s2 = 1. / (2. * v);
      sl = al * s2;
      sr = ar * s2;
      sd = (ar - al) * s2;
    }
    Mt<B> M, b, c;
    double t[B];
    // rhs (forward_backward.hpp:30-36)
    {
      double r0 = dui[0];
#pragma unroll
      for (int n = 1; n < IVX; ++n) r0 += dui[n];
      t[0] = r0 * idt;
      if (B == 5) {
        t[1] = dui[IVX] * idt;
        t[2] = dui[IVX + 1] * idt;
        t[3] = dui[IVX + 2] * idt;
        t[4] = dui[IEN] * idt;
      } else {
        t[1] = dui[IVX] * idt;
        t[2] = dui[IEN] * idt;
      }
    }
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int q = 0; q < B; ++q) {
        double phi = 0.0;
        if (B == 5) {
          if ((r == 1 && q == 0) || (r == 4 && q == 1)) phi = grav;
        } else {
          if ((r == 1 && q == 0) || (r == 2 && q == 1)) phi = grav;
        }
        double a = fma(Am.a[r][q], sl, fma(Ap.a[r][q], sr, dc.a[r][q] * sd));
        if (r == q) a += idt;
        M.a[r][q] = a - phi;
        b.a[r][q] = -(Am.a[r][q] + dm.a[r][q]) * sl;
        c.a[r][q] = -(Ap.a[r][q] - dp.a[r][q]) * sr;
      }
    // reflecting walls: a_is += b_is diag(1,-1,1,..), a_ie += c_ie diag(..)
    if (i == is && g.bc[0] != 1) {
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q)
          M.a[r][q] += (q == 1) ? -b.a[r][q] : b.a[r][q];
    }
    if (i == ie && g.bc[1] != 1) {
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q)
          M.a[r][q] += (q == 1) ? -c.a[r][q] : c.a[r][q];
    }
    if (i > is) {
      // M -= b a'_{i-1};  t -= b delta_{i-1}
#pragma unroll
      for (int r = 0; r < B; ++r) {
#pragma unroll
        for (int q = 0; q < B; ++q) {
          double s = M.a[r][q];
#pragma unroll
          for (int kk = 0; kk < B; ++kk) s = fma(-b.a[r][kk], ap.a[kk][q], s);
          M.a[r][q] = s;
        }
        double s = t[r];
#pragma unroll
        for (int kk = 0; kk < B; ++kk) s = fma(-b.a[r][kk], dl.v[kk], s);
        t[r] = s;
      }
    }
    // Gauss-Jordan on [M | c | t]
#pragma unroll
    for (int cc = 0; cc < B; ++cc) {
      int p = cc;
      double best = fabs(M.a[cc][cc]);
#pragma unroll
      for (int r = cc + 1; r < B; ++r) {
        double v = fabs(M.a[r][cc]);
        if (v > best) {
          best = v;
          p = r;
        }
      }
      if (best == 0.0) ok = false;
      // the row exchange runs only when a lane of the wavefront pivots (a
      // uniform branch: columns 2..4, and every column at the small step
      // of the second stage, keep their diagonal pivot); columns left of cc
      // are not read again
      if (__any(p != cc)) {
#pragma unroll
        for (int r = cc + 1; r < B; ++r) {
          bool sw = (p == r);
#pragma unroll
          for (int q = cc; q < B; ++q) {
            double x0 = M.a[cc][q], x1 = M.a[r][q];
            M.a[cc][q] = sw ? x1 : x0;
            M.a[r][q] = sw ? x0 : x1;
          }
#pragma unroll
          for (int q = 0; q < B; ++q) {
            double y0 = c.a[cc][q], y1 = c.a[r][q];
            c.a[cc][q] = sw ? y1 : y0;
            c.a[r][q] = sw ? y0 : y1;
          }
          double z0 = t[cc], z1 = t[r];
          t[cc] = sw ? z1 : z0;
          t[r] = sw ? z0 : z1;
        }
      }
      double inv = 1.0 / M.a[cc][cc];
#pragma unroll
      for (int q = cc + 1; q < B; ++q) M.a[cc][q] *= inv;
#pragma unroll
      for (int q = 0; q < B; ++q) c.a[cc][q] *= inv;
      t[cc] *= inv;
#pragma unroll
      for (int r = 0; r < B; ++r) {
        if (r == cc) continue;
        double f = M.a[r][cc];
#pragma unroll
        for (int q = cc + 1; q < B; ++q) M.a[r][q] = fma(-f, M.a[cc][q], M.a[r][q]);
#pragma unroll
        for (int q = 0; q < B; ++q) c.a[r][q] = fma(-f, c.a[cc][q], c.a[r][q]);
        t[r] = fma(-f, t[cc], t[r]);
      }
    }
    ap = c;
#pragma unroll
    for (int r = 0; r < B; ++r) dl.v[r] = t[r];
    Am = Ap;
    dm = dc;
    dc = dp;
#pragma unroll
    for (int n = 0; n < N; ++n) {
      wm[n] = wc[n];
      wc[n] = wp[n];
    }
    if (GEO) {
      xf0 = xf1;
      xf1 = nxf;
      xv = nxv;
    }
  }
  archive(ie);
  return ok;
}

template <int NM, int B>
__global__ void __launch_bounds__(64)
    k_implicit_fwd(KGrid g, KThermo th, KImplicit ip,
                   const double* __restrict__ w, const double* __restrict__ du,
                   double* __restrict__ arch, int* __restrict__ status) {
  long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ip.ncol) return;
  int j = g.js + (int)(col % g.nx2);
  int k = g.ks + (int)(col / g.nx2);
  if (!fwd_column<NM, B>(g, th, ip.dt, w, du, g.at(0, k, j, 0), arch, ip.ncol, col))
    atomicOr(status, 1);
}

// delta_i of the back substitution (forward_backward.hpp:86-97): level i
// from the archive and delta_{i+1}.
template <int B>
__device__ inline void back_step(const double* __restrict__ arch,
                                 long long lb, long long as, long long ao,
                                 bool top, const Vt<B>& dn, Vt<B>& d) {
#pragma unroll
  for (int r = 0; r < B; ++r) d.v[r] = __builtin_nontemporal_load(&arch[(lb + B * B + r) * as + ao]);
  if (!top) {
    Mt<B> a;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int q = 0; q < B; ++q)
        a.a[r][q] = __builtin_nontemporal_load(&arch[(lb + r * B + q) * as + ao]);
    Vt<B> ad;
    mvec<B>(a, dn, ad);
#pragma unroll
    for (int r = 0; r < B; ++r) d.v[r] = d.v[r] - ad.v[r];
  }
}

// Write-back of the solution into a cell's du (forward_backward.hpp:221-246:
// the total-density increment is split over the species by w_n).
template <int NM, int B>
__device__ inline void write_back(const Vt<B>& d, const double* duo,
                                  const double* wsp, double* dun) {
  constexpr int IVX = 1 + NM, IEN = 4 + NM;
  double dens = duo[0];
#pragma unroll
  for (int n = 1; n < IVX; ++n) dens += duo[n];
  dens = d.v[0] - dens;
  double d0 = d.v[0];
#pragma unroll
  for (int n = 1; n < IVX; ++n) dun[n] = duo[n] + dens * wsp[n];
  if (B == 5) {
    dun[IVX] = d.v[1];
    dun[IVX + 1] = d.v[2];
    dun[IVX + 2] = d.v[3];
    dun[IEN] = d.v[4];
  } else {
    dun[IVX] = d.v[1];
    dun[IVX + 1] = duo[IVX + 1];
    dun[IVX + 2] = duo[IVX + 2];
    dun[IEN] = d.v[2];
  }
#pragma unroll
  for (int n = 1; n < IVX; ++n) d0 -= dun[n];
  dun[0] = d0;
}

// Back substitution writing du in place (cg_implicit_solve).
template <int NM, int B>
__global__ void __launch_bounds__(64)
    k_implicit_bwd(KGrid g, KImplicit ip, const double* __restrict__ w,
                   const double* __restrict__ arch, double* __restrict__ du) {
  constexpr int N = 5 + NM;
  constexpr int AS = B * B + B;
  long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ip.ncol) return;
  int j = g.js + (int)(col % g.nx2);
  int k = g.ks + (int)(col / g.nx2);
  const int is = g.is, ie = g.ie;
  const long long c0 = g.at(0, k, j, 0);
  const long long sn = g.c.n, si = g.c.i;
  Vt<B> dn;
  for (int i = ie; i >= is; --i) {
    Vt<B> d;
    back_step<B>(arch, (long long)(i - is) * AS, ip.ncol, col, i == ie, dn, d);
    dn = d;
    const long long ci = c0 + i * si;
    double duo[N], wsp[N], dun[N];
#pragma unroll
    for (int n = 0; n < N; ++n) {
      duo[n] = (NM > 0 || n >= 1 + NM) ? du[ci + n * sn] : 0.0;
      wsp[n] = (n >= 1 && n <= NM) ? w[ci + n * sn] : 0.0;
    }
    if (NM == 0) duo[0] = du[ci];
    write_back<NM, B>(d, duo, wsp, dun);
#pragma unroll
    for (int n = 0; n < N; ++n) du[ci + n * sn] = dun[n];
  }
}

// ---------------------------------------------------------------------------
// End of stage per cell: u1 / u averaging (WeightedAve,
// implicit_hydro_tasks.cpp:146-160), u += du (:277-283), saturation
// adjustment on the last stage (UpdateAllConserved, :290-313),
// ConservedToPrimitive.  Returns the new u (uc) and w (wc).
// Kessler warm rain and precipitation sedimentation: rates and fall speed
// (examples/2019-Li-snap/earth-precip.yaml:44-53, water-kessler.yaml:4-18),
// ip the 0-based species of the precipitation
struct KKessler {
  double k_auto, k_accr, k_evap, v_fall;
  int ip;
};

// One cell of the top-down Kessler sweep: uc (conserved state), wc (its
// primitive state) and rho (total density) in; the cloud -> precipitation
// conversion, the evaporation of precipitation into sub-saturated air and
// the sedimentation exchange applied to uc, wc.  fin: the precipitation
// mass, momenta and energy falling in from the cell above; fout: what falls
// out of this cell (nothing at the bottom level: it does not cross the
// wall).  dt the full step, cf = min(1, dt v_fall / dx1), cvp the
// precipitation's cv.
template <int NM>
__device__ __forceinline__ void kessler_cell(const KThermo& th, const KKessler& kz,
                                             double dt, double cf, double cvp, bool bottom,
                                             double rho, double* uc, double* wc,
                                             const double* fin, double* fout) {
  constexpr int IVX = 1 + NM, IEN = 4 + NM;
  const int rv_ = 1 + th.iv, rc_ = 1 + th.ic, rp_ = 1 + kz.ip;
  const double T = temperature_d<NM>(th, wc);
  // species picked by runtime index from registers: select chains
  double rc = 0., rp = 0., rv = 0.;
#pragma unroll
  for (int n = 1; n <= NM; ++n) {
    if (n == rc_) rc = uc[n];
    if (n == rp_) rp = uc[n];
    if (n == rv_) rv = uc[n];
  }
  double a = dt * (kz.k_auto + kz.k_accr * (rp / rho)) * rc;
  if (a > rc) a = rc;
  rc = rc - a;
  rp = rp + a;
  const double rs = qsat_d(th, T, rho) * rho;
  const double def = rs - rv;
  double b = 0.;
  if (def > 0. && rp > 0.) {
    b = dt * kz.k_evap * rp * (def / rs);
    if (b > rp) b = rp;
    if (b > def) b = def;
  }
  rp = rp - b;
  rv = rv + b;
#pragma unroll
  for (int n = 1; n <= NM; ++n) {
    if (n == rc_) uc[n] = rc;
    if (n == rp_) uc[n] = rp;
    if (n == rv_) uc[n] = rv;
  }
  uc[IEN] = uc[IEN] - th.L * b;
  double out_p = 0., out_m0 = 0., out_m1 = 0., out_m2 = 0., out_e = 0.;
  if (!bottom && cf > 0.) {
    cons_to_prim_cell<NM>(th, uc, wc);
    const double T2 = temperature_d<NM>(th, wc);
    const double v1 = wc[IVX], v2 = wc[IVX + 1], v3 = wc[IVX + 2];
    out_p = cf * rp;
    out_m0 = out_p * v1;
    out_m1 = out_p * v2;
    out_m2 = out_p * v3;
    out_e = out_p * (cvp * T2 + 0.5 * (v1 * v1 + v2 * v2 + v3 * v3));
  }
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n == rp_) uc[n] = (uc[n] - out_p) + fin[0];
  uc[IVX] = (uc[IVX] - out_m0) + fin[1];
  uc[IVX + 1] = (uc[IVX + 1] - out_m1) + fin[2];
  uc[IVX + 2] = (uc[IVX + 2] - out_m2) + fin[3];
  uc[IEN] = (uc[IEN] - out_e) + fin[4];
  cons_to_prim_cell<NM>(th, uc, wc);
  fout[0] = out_p;
  fout[1] = out_m0;
  fout[2] = out_m1;
  fout[3] = out_m2;
  fout[4] = out_e;
}

struct KUpdate {
  int stage;  // 0, 1, 2
  double g1, g2;
  int satadj;
  // Kessler microphysics in the downward sweep (its KES instances): the
  // rates and the full step, kdt or kdt * kdtp[0] when kdtp is given
  KKessler kz{};
  double kdt = 0.0;
  const double* kdtp = nullptr;
};

// uo / u1o: the cell's u and u1 (u1o unused on stage 0, which writes u1).
template <int NM>
__device__ inline void update_cell_pre(const KThermo& th, const KUpdate& up,
                                       const double* uo, const double* u1o,
                                       double* __restrict__ u1, long long ci,
                                       long long sn, const double* dun,
                                       double* uc, double* wc,
                                       int* __restrict__ status) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IEN = 4 + NM;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double un = uo[n];
    double v;
    if (up.stage == 0) {
      v = 0.0 + 1.0 * un;
      __builtin_nontemporal_store(v, &u1[ci + n * sn]);
    } else {
      v = up.g1 * un + up.g2 * u1o[n];
    }
    uc[n] = v + dun[n];
  }
  if (up.satadj && NM > 0 && th.iv >= 0) {
    double rho = uc[0];
#pragma unroll
    for (int n = 1; n <= NM; ++n) rho += uc[n];
    double m1 = uc[IVX], m2 = uc[IVX + 1], m3 = uc[IVX + 2];
    double ke = 0.5 * (m1 * (m1 / rho) + m2 * (m2 / rho) + m3 * (m3 / rho));
    double ie = uc[IEN] - ke;
    double q[NM > 0 ? NM : 1];
#pragma unroll
    for (int n = 0; n < NM; ++n) q[n] = uc[n + 1] / rho;
    int it = adjust_cell<NM>(th, rho, &ie, q, -1.0);
    if (it < 0) atomicOr(status, 2);
#pragma unroll
    for (int n = 0; n < NM; ++n) uc[n + 1] = q[n] * rho;
    uc[IEN] = ie + ke;
  }
  cons_to_prim_cell<NM>(th, uc, wc);
}

template <int NM>
__device__ inline void update_cell(const KThermo& th, const KUpdate& up,
                                   double* __restrict__ u,
                                   double* __restrict__ u1, long long ci,
                                   long long sn, const double* dun, double* uc,
                                   double* wc, int* __restrict__ status) {
  constexpr int N = 5 + NM;
  double uo[N], u1o[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    uo[n] = u[ci + n * sn];
    u1o[n] = (up.stage == 0) ? 0.0 : u1[ci + n * sn];
  }
  update_cell_pre<NM>(th, up, uo, u1o, u1, ci, sn, dun, uc, wc, status);
}

// Cell-parallel end of stage (explicit runs).
template <int NM>
__global__ void __launch_bounds__(256)
    k_update(KGrid g, KThermo th, KUpdate up, double* __restrict__ u,
             double* __restrict__ u1, const double* __restrict__ du,
             double* __restrict__ w, int* __restrict__ status) {
  constexpr int N = 5 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return;
  int i, j, k;
  g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  const long long c0 = g.at(0, k, j, i);
  double dun[N], uc[N], wc[N];
#pragma unroll
  for (int n = 0; n < N; ++n) dun[n] = du[c0 + n * g.c.n];
  update_cell<NM>(th, up, u, u1, c0, g.c.n, dun, uc, wc, status);
#pragma unroll
  for (int n = 0; n < N; ++n) {
    u[c0 + n * g.c.n] = uc[n];
    w[c0 + n * g.c.n] = wc[n];
  }
}

// Implicit runs: one downward sweep per column does the back substitution,
// the write-back into du, the end-of-stage update of u and w, the x1 ghost
// cells of w (reflecting / outflow) and the hydrostatic face and cell
// pressures of the next stage (the same recurrences as k_hydrostatic_scan,
// evaluated as the new state appears top-down).  KES (last stage of a run
// with Kessler microphysics, x1 sides walls or outflow): each new cell
// state passes through kessler_cell before it is stored, which is the top-
// down sweep of kessler_column run as the update reaches each level.
// One column: c0 as in fwd_column, pc = offset of the column in psf / psv.
template <int NM, int B, bool KES = false>
__device__ __forceinline__ void bwd_update_column(
    const KGrid& g, const KThermo& th, const KUpdate& up,
    const double* __restrict__ arch, long long as, long long ao,
    const double* __restrict__ du, double* __restrict__ u,
    double* __restrict__ u1, double* __restrict__ w, double* __restrict__ psf,
    double* __restrict__ psv, long long c0, long long pc,
    int* __restrict__ status) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IPR = 4 + NM;
  constexpr int AS = B * B + B;
  const int is = g.is, ie = g.ie;
  const long long sn = g.c.n, si = g.c.i;
  const double grav = g.grav, dx = g.dx1;
  const bool hs = grav != 0.0;
  const double gb = (g.bc[0] == 0) ? -grav : grav;
  const double gt = (g.bc[1] == 0) ? -grav : grav;
  double rtop[NG];    // density of the top three cells (ie, ie-1, ie-2)
  double pf_top[NG + 1];  // psf(ie+1) .. psf(ie+4)
  double pf_next = 0.0;   // psf(i+1)
  // Kessler: the full step, the fall fraction, the precipitation's cv and
  // what falls into the current level from the level above
  double kdt = 0.0, kcf = 0.0, kcvp = 0.0;
  double kin[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if constexpr (KES) {
    kdt = up.kdtp ? up.kdt * up.kdtp[0] : up.kdt;
    kcf = kdt * up.kz.v_fall / dx;
    if (kcf > 1.) kcf = 1.;
    kcvp = th.cvd * (1. + th.cv_ratio_m1[up.kz.ip]);
  }
  // the archive and the state of the levels below are loaded two levels
  // ahead, so that their latency overlaps the current levels' updates
  struct Level {
    Mt<B> a;
    Vt<B> dl;
    double duo[N], wsp[N], uo[N], u1o[N];
  };
  // the archive is read once: non-temporal loads keep it from displacing
  // the state in the caches
  auto fetch = [&](Level& L, int i) {
    const long long lb = (long long)(i - is) * AS;
#pragma unroll
    for (int r = 0; r < B; ++r)
      L.dl.v[r] = __builtin_nontemporal_load(&arch[(lb + B * B + r) * as + ao]);
    if (i != ie) {
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q)
          L.a.a[r][q] = __builtin_nontemporal_load(&arch[(lb + r * B + q) * as + ao]);
    }
    const long long ci = c0 + i * si;
#pragma unroll
    for (int n = 0; n < N; ++n) {
      L.duo[n] = (n <= NM || B == 3) ? du[ci + n * sn] : 0.0;
      L.wsp[n] = (n >= 1 && n <= NM) ? w[ci + n * sn] : 0.0;
      L.uo[n] = u[ci + n * sn];
      L.u1o[n] = (up.stage == 0) ? 0.0 : u1[ci + n * sn];
    }
  };
  Vt<B> dn;
  // level i of the sweep from the buffer cur (which then is free)
  auto level = [&](Level& cur, int i) {
      // back substitution delta_i = d_i - a'_i delta_{i+1}
      // (forward_backward.hpp:86-97)
      Vt<B> d;
      if (i == ie) {
        d = cur.dl;
      } else {
        Vt<B> ad;
        mvec<B>(cur.a, dn, ad);
#pragma unroll
        for (int r = 0; r < B; ++r) d.v[r] = cur.dl.v[r] - ad.v[r];
      }
      dn = d;
      const long long ci = c0 + i * si;
      double dun[N], uc[N], wc[N];
      write_back<NM, B>(d, cur.duo, cur.wsp, dun);
      update_cell_pre<NM>(th, up, cur.uo, cur.u1o, u1, ci, sn, dun, uc, wc, status);
      if constexpr (KES) {
        double rho = uc[0];
#pragma unroll
        for (int n = 1; n <= NM; ++n) rho += uc[n];
        double kout[5];
        kessler_cell<NM>(th, up.kz, kdt, kcf, kcvp, i == is, rho, uc, wc, kin, kout);
#pragma unroll
        for (int n = 0; n < 5; ++n) kin[n] = kout[n];
      }
#pragma unroll
      for (int n = 0; n < N; ++n) {
        // streamed: read again only by the next stage's kernels
        __builtin_nontemporal_store(uc[n], &u[ci + n * sn]);
        __builtin_nontemporal_store(wc[n], &w[ci + n * sn]);
      }
      // x1 ghosts mirrored from the cells next to the walls
      const int mt = ie - i, mb = i - is;  // distance from top / bottom wall
      if (mt < NG) {
        int gi = ie + 1 + mt;
#pragma unroll
        for (int n = 0; n < N; ++n) {
          double v = (g.bc[1] == 0) ? ((n == IVX) ? -wc[n] : wc[n])
                                    : w[c0 + n * sn + ie * si];
          if (g.bc[1] == 2 && mt == 0) v = wc[n];
          w[c0 + n * sn + gi * si] = v;
        }
      }
      if (mb < NG) {
        int gi = is - 1 - mb;
#pragma unroll
        for (int n = 0; n < N; ++n) {
          double v = (g.bc[0] == 0) ? ((n == IVX) ? -wc[n] : wc[n]) : 0.0;
          w[c0 + n * sn + gi * si] = v;
        }
      }
      if (hs) {
        if (i == ie) {
          double RdTv = wc[IPR] / wc[0];
          pf_next = wc[IPR] * exp_d(grav * dx / (2 * RdTv));
          pf_top[0] = pf_next;
          psf[pc + (ie + 1) * g.ps_i] = pf_next;
        }
        if (mt < NG) rtop[mt] = wc[0];
        double pfi = pf_next - grav * wc[0] * dx;
        psf[pc + i * g.ps_i] = pfi;
        psv[pc + i * g.ps_i] = log_mean(pfi, pf_next);
        pf_next = pfi;
      }
  };
  // three buffers: the archive and state of level i - 2 are requested
  // while level i is processed
  Level L0, L1, L2;
  fetch(L0, ie);
  if (ie - 1 >= is) fetch(L1, ie - 1);
  for (int i = ie; i >= is; i -= 3) {
    if (i - 2 >= is) fetch(L2, i - 2);
    level(L0, i);
    if (i - 1 < is) break;
    if (i - 3 >= is) fetch(L0, i - 3);
    level(L1, i - 1);
    if (i - 2 < is) break;
    if (i - 4 >= is) fetch(L1, i - 4);
    level(L2, i - 2);
  }
  // outflow bottom ghosts copy the first interior cell
  if (g.bc[0] == 2) {
#pragma unroll
    for (int n = 0; n < N; ++n) {
      double v = w[c0 + n * sn + is * si];
      for (int gh = 1; gh <= NG; ++gh) w[c0 + n * sn + (is - gh) * si] = v;
    }
  }
  if (g.bc[1] == 2) {
#pragma unroll
    for (int n = 0; n < N; ++n) {
      double v = w[c0 + n * sn + ie * si];
      for (int gh = 1; gh <= NG; ++gh) w[c0 + n * sn + (ie + gh) * si] = v;
    }
  }
  if (hs) {
    // bottom ghosts: psf(is-1..is-3), psv of those cells
    double cur = pf_next;  // psf(is)
    for (int gh = 1; gh <= NG; ++gh) {
      double rho = w[c0 + (is - gh) * si];
      double lo = cur - gb * rho * dx;
      psf[pc + (is - gh) * g.ps_i] = lo;
      psv[pc + (is - gh) * g.ps_i] = log_mean(lo, cur);
      cur = lo;
    }
    // top ghosts: psf(ie+2..ie+4) upward from psf(ie+1)
    cur = pf_top[0];
    for (int gh = 1; gh <= NG; ++gh) {
      double rho = w[c0 + (ie + gh) * si];
      double hi = cur + gt * rho * dx;
      psf[pc + (ie + gh + 1) * g.ps_i] = hi;
      psv[pc + (ie + gh) * g.ps_i] = log_mean(cur, hi);
      cur = hi;
    }
  }
  (void)rtop;
}

template <int NM, int B>
__global__ void __launch_bounds__(64)
    k_implicit_bwd_update(KGrid g, KThermo th, KImplicit ip, KUpdate up,
                          const double* __restrict__ arch,
                          const double* __restrict__ du, double* __restrict__ u,
                          double* __restrict__ u1, double* __restrict__ w,
                          double* __restrict__ psf, double* __restrict__ psv,
                          int* __restrict__ status) {
  long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ip.ncol) return;
  int j = g.js + (int)(col % g.nx2);
  int k = g.ks + (int)(col / g.nx2);
  bwd_update_column<NM, B>(g, th, up, arch, ip.ncol, col, du, u, u1, w, psf, psv,
                           g.at(0, k, j, 0), col * g.ps_c, status);
}

// ---------------------------------------------------------------------------
// Batched column solve: every column of several blocks (of any heights) in
// one launch.  Columns are taken in groups of 64 of one block (one group per
// wavefront); a workgroup holds WPB wavefronts and walks the group table
// with a stride of the whole grid, so a grid smaller than the table makes
// the launch persistent.  The host sorts the groups by column height, tallest
// first.
struct KBlockPtr {
  KGrid g;
  double *u, *w, *u1, *du, *psf, *psv;
  const double* pref;  // per-level reference pressure (FP32-flux form)
};

// A pointer read from a block table, marked as global memory: kernels then
// use global (not flat) loads and stores, which neither wait on nor hold up
// LDS and scalar-memory counters.
template <class T>
__device__ __forceinline__ T* gptr(T* p) {
  typedef __attribute__((address_space(1))) T gT;
  return (T*)(gT*)(uintptr_t)p;
}

// Group of up to 64 columns col0 .. col0+63 of block b.  aoff is the start
// of block b's archive, laid out level-major like the state: entry e of
// level lvl of column c at aoff + (lvl (B*B+B) + e) nx2 nx3 + c.  A
// wavefront's access of one entry is one 512-byte segment, and at any time
// every wavefront works in the same few pages of the archive.
struct KGroup {
  int b, col0;
  long long aoff;
};

template <int NM, int B, int WPB>
__global__ void __launch_bounds__(64 * WPB)
    k_implicit_fwd_batch(const KBlockPtr* __restrict__ bp,
                         const KGroup* __restrict__ grp, int ngroups,
                         KThermo th, double dt, const double* __restrict__ dtp,
                         double* __restrict__ arch, int* __restrict__ status) {
  // dt is beta dt, or beta times the run's device time step dtp[0]
  if (dtp) dt = dt * dtp[0];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (int q = blockIdx.x * WPB + wv; q < ngroups; q += gridDim.x * WPB) {
    const int b = __builtin_amdgcn_readfirstlane(grp[q].b);
    const int col0 = __builtin_amdgcn_readfirstlane(grp[q].col0);
    const long long aoff = grp[q].aoff;
    const KGrid g = bp[b].g;
    const long long col = (long long)col0 + lane;
    if (col < (long long)g.nx2 * g.nx3) {
      const int j = g.js + (int)(col % g.nx2), k = g.ks + (int)(col / g.nx2);
      if (!fwd_column<NM, B>(g, th, dt, gptr(bp[b].w), gptr(bp[b].du),
                             g.at(0, k, j, 0), arch + aoff,
                             (long long)g.nx2 * g.nx3, col, gptr(bp[b].psv),
                             col * g.ps_c))
        atomicOr(status, 1);
    }
  }
}

// The downward sweeps of the groups taken by one wave (KES: with the
// Kessler microphysics of the last stage).
template <int NM, int B, int WPB, bool KES>
__device__ __forceinline__ void bwd_update_groups(const KBlockPtr* __restrict__ bp,
                                                  const KGroup* __restrict__ grp,
                                                  int ngroups, const KThermo& th,
                                                  const KUpdate& up,
                                                  const double* __restrict__ arch,
                                                  int* __restrict__ status) {
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (int q = blockIdx.x * WPB + wv; q < ngroups; q += gridDim.x * WPB) {
    const int b = __builtin_amdgcn_readfirstlane(grp[q].b);
    const int col0 = __builtin_amdgcn_readfirstlane(grp[q].col0);
    const long long aoff = grp[q].aoff;
    const KGrid g = bp[b].g;
    const long long col = (long long)col0 + lane;
    if (col < (long long)g.nx2 * g.nx3) {
      const int j = g.js + (int)(col % g.nx2), k = g.ks + (int)(col / g.nx2);
      const KBlockPtr& p = bp[b];
      bwd_update_column<NM, B, KES>(g, th, up, arch + aoff, (long long)g.nx2 * g.nx3,
                                    col, gptr(p.du),
                                    gptr(p.u), gptr(p.u1), gptr(p.w), gptr(p.psf),
                                    gptr(p.psv), g.at(0, k, j, 0), col * g.ps_c,
                                    status);
    }
  }
}

template <int NM, int B, int WPB>
__global__ void __launch_bounds__(64 * WPB)
    k_implicit_bwd_update_batch(const KBlockPtr* __restrict__ bp,
                                const KGroup* __restrict__ grp, int ngroups,
                                KThermo th, KUpdate up,
                                const double* __restrict__ arch,
                                int* __restrict__ status) {
  bwd_update_groups<NM, B, WPB, false>(bp, grp, ngroups, th, up, arch, status);
}

template <int NM, int B, int WPB>
__global__ void __launch_bounds__(64 * WPB)
    k_implicit_bwd_update_kessler_batch(const KBlockPtr* __restrict__ bp,
                                        const KGroup* __restrict__ grp, int ngroups,
                                        KThermo th, KUpdate up,
                                        const double* __restrict__ arch,
                                        int* __restrict__ status) {
  bwd_update_groups<NM, B, WPB, true>(bp, grp, ngroups, th, up, arch, status);
}

// Forward sweep and downward sweep of each group back to back in one wave:
// the columns of a group depend on no other group within a stage, and the
// top levels' archive, written last by the forward sweep, is read first by
// the downward sweep while still in the caches.
template <int NM, int B, int WPB, bool KES>
__device__ __forceinline__ void fused_groups(const KBlockPtr* __restrict__ bp,
                                             const KGroup* __restrict__ grp, int ngroups,
                                             const KThermo& th, double dt,
                                             const double* __restrict__ dtp,
                                             const KUpdate& up, double* __restrict__ arch,
                                             int* __restrict__ status) {
  if (dtp) dt = dt * dtp[0];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (int q = blockIdx.x * WPB + wv; q < ngroups; q += gridDim.x * WPB) {
    const int b = __builtin_amdgcn_readfirstlane(grp[q].b);
    const int col0 = __builtin_amdgcn_readfirstlane(grp[q].col0);
    const long long aoff = grp[q].aoff;
    const KGrid g = bp[b].g;
    const long long col = (long long)col0 + lane;
    if (col < (long long)g.nx2 * g.nx3) {
      const int j = g.js + (int)(col % g.nx2), k = g.ks + (int)(col / g.nx2);
      const KBlockPtr& p = bp[b];
      const long long ncol = (long long)g.nx2 * g.nx3;
      if (!fwd_column<NM, B>(g, th, dt, gptr(p.w), gptr(p.du), g.at(0, k, j, 0),
                             arch + aoff, ncol, col, gptr(p.psv), col * g.ps_c))
        atomicOr(status, 1);
      bwd_update_column<NM, B, KES>(g, th, up, arch + aoff, ncol, col, gptr(p.du),
                                    gptr(p.u), gptr(p.u1), gptr(p.w), gptr(p.psf),
                                    gptr(p.psv), g.at(0, k, j, 0), col * g.ps_c,
                                    status);
    }
  }
}

template <int NM, int B, int WPB>
__global__ void __launch_bounds__(64 * WPB)
    k_implicit_fused_batch(const KBlockPtr* __restrict__ bp,
                           const KGroup* __restrict__ grp, int ngroups,
                           KThermo th, double dt, const double* __restrict__ dtp,
                           KUpdate up, double* __restrict__ arch,
                           int* __restrict__ status) {
  fused_groups<NM, B, WPB, false>(bp, grp, ngroups, th, dt, dtp, up, arch, status);
}

// The same with the Kessler microphysics in the downward sweep (last stage).
template <int NM, int B, int WPB>
__global__ void __launch_bounds__(64 * WPB)
    k_implicit_fused_kessler_batch(const KBlockPtr* __restrict__ bp,
                                   const KGroup* __restrict__ grp, int ngroups,
                                   KThermo th, double dt, const double* __restrict__ dtp,
                                   KUpdate up, double* __restrict__ arch,
                                   int* __restrict__ status) {
  fused_groups<NM, B, WPB, true>(bp, grp, ngroups, th, dt, dtp, up, arch, status);
}

// Streaming copy of n 16-byte words (cg_memcpy_d2d): grid-stride loop.
static __global__ void __launch_bounds__(256)
    k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
             long long n) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    dst[t] = src[t];
}

static __global__ void __launch_bounds__(256)
    k_copy1(const unsigned char* __restrict__ src,
            unsigned char* __restrict__ dst, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

// cons <-> prim on interior cells
template <int NM, int TO_PRIM>
__global__ void __launch_bounds__(256)
    k_eos(KGrid g, KThermo th, const double* __restrict__ src,
          double* __restrict__ dst) {
  constexpr int N = 5 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return;
  int i, j, k;
  g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  const long long c0 = g.at(0, k, j, i);
  double a[N], b[N];
#pragma unroll
  for (int n = 0; n < N; ++n) a[n] = src[c0 + n * g.c.n];
  if (TO_PRIM)
    cons_to_prim_cell<NM>(th, a, b);
  else
    prim_to_cons_cell<NM>(th, a, b);
#pragma unroll
  for (int n = 0; n < N; ++n) dst[c0 + n * g.c.n] = b[n];
}

// Copy between two layouts of the same cell array (all cells incl. ghosts);
// lanes follow the destination's unit-stride index.
static __global__ void __launch_bounds__(256)
    k_relayout(KGrid gs, KGrid gd, int N, const double* __restrict__ src,
               double* __restrict__ dst) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long per = (long long)gd.nc1 * gd.nc2 * gd.nc3;
  if (t >= per * N) return;
  int n = (int)(t / per);
  long long r = t % per;
  int i, j, k;
  gd.decode(r, gd.nc1, gd.nc2, gd.nc3, i, j, k);
  dst[gd.at(n, k, j, i)] = src[gs.at(n, k, j, i)];
}

// ---------------------------------------------------------------------------
// Ghost cells (upstream BoundaryValues order: x1 over interior (j,k), then x2
// over all i and interior k, then x3 over all i, j).  One lane per
// (ghost layer, side, line); variables looped.
__device__ inline void bc_pair(double* w, long long sn, int N, int vn, int bci,
                               int bco, long long a_lo, long long a_hi,
                               long long stride, int gh) {
  for (int n = 0; n < N; ++n) {
    double sgn = (n == vn) ? -1.0 : 1.0;
    long long o = n * sn;
    double vin, vout;
    if (bci == 0)
      vin = sgn * w[o + a_lo + (gh - 1) * stride];
    else if (bci == 1)
      vin = w[o + a_hi - (gh - 1) * stride];
    else
      vin = w[o + a_lo];
    if (bco == 0)
      vout = sgn * w[o + a_hi - (gh - 1) * stride];
    else if (bco == 1)
      vout = w[o + a_lo + (gh - 1) * stride];
    else
      vout = w[o + a_hi];
    w[o + a_lo - gh * stride] = vin;
    w[o + a_hi + gh * stride] = vout;
  }
}

template <int NM>
__global__ void __launch_bounds__(256) k_bc_x1(KGrid g, double* w) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx2 * g.nx3 * NG) return;
  int gh, j, k;
  g.decode(t, NG, g.nx2, g.nx3, gh, j, k);
  bc_pair(w, g.c.n, 5 + NM, 1 + NM, g.bc[0], g.bc[1],
          g.at(0, k + g.ks, j + g.js, g.is), g.at(0, k + g.ks, j + g.js, g.ie),
          g.c.i, gh + 1);
}

template <int NM>
__global__ void __launch_bounds__(256) k_bc_x2(KGrid g, double* w) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nc1 * g.nx3 * NG) return;
  int i, gh, k;
  g.decode(t, g.nc1, NG, g.nx3, i, gh, k);
  bc_pair(w, g.c.n, 5 + NM, 2 + NM, g.bc[2], g.bc[3],
          g.at(0, k + g.ks, g.js, i), g.at(0, k + g.ks, g.je, i), g.c.j,
          gh + 1);
}

template <int NM>
__global__ void __launch_bounds__(256) k_bc_x3(KGrid g, double* w) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nc1 * g.nc2 * NG) return;
  int i, j, gh;
  g.decode(t, g.nc1, g.nc2, NG, i, j, gh);
  bc_pair(w, g.c.n, 5 + NM, 3 + NM, g.bc[4], g.bc[5], g.at(0, g.ks, j, i),
          g.at(0, g.ke, j, i), g.c.k, gh + 1);
}

// x2 / x3 ghosts of every block of a run in one launch each (blockIdx.y =
// block; blocks without the direction are skipped).
template <int NM>
__global__ void __launch_bounds__(256)
    k_bc_x2_batch(const KBlockPtr* __restrict__ bp) {
  const KGrid& g = bp[blockIdx.y].g;
  if (!g.f2) return;
  double* w = gptr(bp[blockIdx.y].w);
  const long long n = (long long)g.nc1 * g.nx3 * NG;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    int i, gh, k;
    g.decode(t, g.nc1, NG, g.nx3, i, gh, k);
    bc_pair(w, g.c.n, 5 + NM, 2 + NM, g.bc[2], g.bc[3],
            g.at(0, k + g.ks, g.js, i), g.at(0, k + g.ks, g.je, i), g.c.j,
            gh + 1);
  }
}

template <int NM>
__global__ void __launch_bounds__(256)
    k_bc_x3_batch(const KBlockPtr* __restrict__ bp) {
  const KGrid& g = bp[blockIdx.y].g;
  if (!g.f3) return;
  double* w = gptr(bp[blockIdx.y].w);
  const long long n = (long long)g.nc1 * g.nc2 * NG;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, gh;
    g.decode(t, g.nc1, g.nc2, NG, i, j, gh);
    bc_pair(w, g.c.n, 5 + NM, 3 + NM, g.bc[4], g.bc[5], g.at(0, g.ks, j, i),
            g.at(0, g.ke, j, i), g.c.k, gh + 1);
  }
}

// ---------------------------------------------------------------------------
// Saturation adjustment on flat arrays (cg_saturation_adjust).
template <int NM>
__global__ void __launch_bounds__(256)
    k_satadj(KThermo th, int ncell, const double* __restrict__ rho,
             double* __restrict__ ie, double* __restrict__ y,
             const double* __restrict__ tguess, int* __restrict__ iters,
             int* __restrict__ status) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double q[NM > 0 ? NM : 1];
#pragma unroll
  for (int n = 0; n < NM; ++n) q[n] = y[(long long)n * ncell + c];
  double e = ie[c];
  int it = adjust_cell<NM>(th, rho[c], &e, q, tguess ? tguess[c] : -1.0);
  if (it < 0) atomicOr(status, 2);
  ie[c] = e;
#pragma unroll
  for (int n = 0; n < NM; ++n) y[(long long)n * ncell + c] = q[n];
  if (iters) iters[c] = it;
}

// FP64 sums of u times the cell volume over interior cells, per variable:
// one partial per workgroup, finished on the host in a fixed order.
template <int NM>
__global__ void __launch_bounds__(256)
    k_sums(KGrid g, const double* __restrict__ u, double* __restrict__ part) {
  constexpr int N = 5 + NM;
  __shared__ double sh[256];
  long long total = (long long)g.nx1 * g.nx2 * g.nx3;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  for (int n = 0; n < N; ++n) {
    double s = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < total; t += (long long)gridDim.x * blockDim.x) {
      int i, j, k;
      g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
      s += u[g.at(n, k + g.ks, j + g.js, i + g.is)] * vol;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)n * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
}

// Hydro::NewBlockTimeStep (patched by patches/21.new_blockdt.patch:16-45):
// the step of one interior cell, min over the active directions of
// dx / (|v| + c_s), or dx / (cfl |v|) in a direction solved implicitly
// (x1 only when nx2 > 1).  The minimum over cells times cfl is the block's
// step; minima are exact, so any reduction order gives the same double.
template <int NM>
__device__ inline double cell_new_dt(const KGrid& g, const KThermo& th,
                                     const double* __restrict__ w, long long c0,
                                     double cfl, int iflag) {
  constexpr int N = 5 + NM, IVX = 1 + NM;
  double wi[N];
#pragma unroll
  for (int n = 0; n < N; ++n) wi[n] = w[c0 + n * g.c.n];
  const double cs = sound_speed<NM>(th, wi);
  const double dxs[3] = {g.dx1, g.dx2, g.dx3};
  const bool act[3] = {true, g.f2 != 0, g.f3 != 0};
  const bool imp[3] = {(iflag & 1) && g.f2, (iflag & 2) != 0, (iflag & 4) != 0};
  double m = __DBL_MAX__;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!act[d]) continue;
    const double v = fabs(wi[IVX + d]);
    const double dtd = imp[d] ? dxs[d] / (cfl * v) : dxs[d] / (v + cs);
    m = fmin(m, dtd);
  }
  return m;
}

// Minimum of block-wide minima through LDS: one value per workgroup.
__device__ inline double wg_min256(double m, double* sh) {
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  return sh[0];
}

// State checks of one interior cell (check_hydro_variables,
// src/snap/eos/check_hydro_variables.cpp:21-64, and check_eos_cons2prim,
// src/checks.cpp:52-75): 0 a NaN anywhere in w, 1 a negative density,
// mass fraction or pressure, 2 a temperature below |g| dx1 / Rd (with
// gravity), 3 none.  (The reference's negativity loop also runs over the
// velocities, which its message and any flow contradict; it is restated for
// the densities and mass fractions.)
template <int NM>
__device__ inline int cell_check(const KGrid& g, const KThermo& th,
                                 const double* __restrict__ w, long long c0) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
  double wi[N];
  bool nan = false, neg = false;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    wi[n] = w[c0 + n * g.c.n];
    nan = nan || (wi[n] != wi[n]);
    if (n <= NM || n == IPR) neg = neg || (wi[n] < 0.0);
  }
  if (nan) return 0;
  if (neg) return 1;
  if (g.grav != 0.0 && temperature_d<NM>(th, wi) < fabs(g.grav) * g.dx1 / th.rd) return 2;
  return 3;
}

// Per-workgroup minima of the cell steps.  Blocks of a run come from the
// table bp (blockIdx.y = block, part[blockIdx.y gridDim.x + blockIdx.x]);
// with bp null, the single block (g0, w0) of a caller array.  With bad
// set, the state checks of every cell: the first failing cell in (block, k,
// j, i) order and its check, as ((block 2^40 + index) 4 + check), by an
// atomic minimum (bad starts at all ones).
template <int NM>
__global__ void __launch_bounds__(256)
    k_new_dt(const KBlockPtr* __restrict__ bp, KGrid g0,
             const double* __restrict__ w0, KThermo th, double cfl, int iflag,
             double* __restrict__ part,
             unsigned long long* __restrict__ bad = nullptr) {
  const KGrid g = bp ? bp[blockIdx.y].g : g0;
  const double* w = gptr(bp ? bp[blockIdx.y].w : w0);
  const long long total = (long long)g.nx1 * g.nx2 * g.nx3;
  double m = __DBL_MAX__;
  unsigned long long first = ~0ULL;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
    const long long c0 = g.at(0, k + g.ks, j + g.js, i + g.is);
    m = fmin(m, cell_new_dt<NM>(g, th, w, c0, cfl, iflag));
    if (bad) {
      const int code = cell_check<NM>(g, th, w, c0);
      if (code < 3) {
        const unsigned long long x = ((unsigned long long)k * g.nx2 + j) * g.nx1 + i;
        const unsigned long long key =
            ((((unsigned long long)blockIdx.y << 40) | x) << 2) | (unsigned long long)code;
        first = key < first ? key : first;
      }
    }
  }
  if (bad && first != ~0ULL) atomicMin(bad, first);
  __shared__ double sh[256];
  m = wg_min256(m, sh);
  if (threadIdx.x == 0) part[(long long)blockIdx.y * gridDim.x + blockIdx.x] = m;
}

// out[0] = cfl times the minimum of part[0 .. n-1] (one workgroup of 256).
static __global__ void __launch_bounds__(256)
    k_new_dt_finish(const double* __restrict__ part, int n, double cfl,
                    double* __restrict__ out) {
  __shared__ double sh[256];
  double m = __DBL_MAX__;
  for (int q = threadIdx.x; q < n; q += blockDim.x) m = fmin(m, part[q]);
  m = wg_min256(m, sh);
  if (threadIdx.x == 0) out[0] = cfl * m;
}

// Athena++'s default history sums of every block of a run (blockIdx.y =
// block): volume-weighted u(IDN), momenta, kinetic energies 0.5 m^2 / u(IDN),
// u(IEN), then the species densities; 8 + NM partials per workgroup at
// part[((blockIdx.y gridDim.x + blockIdx.x) (8 + NM) + q], summed on the host
// in a fixed order.
template <int NM>
__global__ void __launch_bounds__(256)
    k_history(const KBlockPtr* __restrict__ bp, double* __restrict__ part) {
  constexpr int NQ = 8 + NM, IVX = 1 + NM, IEN = 4 + NM;
  __shared__ double sh[256];
  const KGrid& g = bp[blockIdx.y].g;
  const double* u = gptr(bp[blockIdx.y].u);
  const long long total = (long long)g.nx1 * g.nx2 * g.nx3;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
    const long long c0 = g.at(0, k + g.ks, j + g.js, i + g.is);
    const double d = u[c0];
    const double m1 = u[c0 + IVX * g.c.n], m2 = u[c0 + (IVX + 1) * g.c.n],
                 m3 = u[c0 + (IVX + 2) * g.c.n];
    acc[0] += vol * d;
    acc[1] += vol * m1;
    acc[2] += vol * m2;
    acc[3] += vol * m3;
    acc[4] += vol * 0.5 * (m1 * m1) / d;
    acc[5] += vol * 0.5 * (m2 * m2) / d;
    acc[6] += vol * 0.5 * (m3 * m3) / d;
    acc[7] += vol * u[c0 + IEN * g.c.n];
#pragma unroll
    for (int n = 1; n <= NM; ++n) acc[7 + n] += vol * u[c0 + n * g.c.n];
  }
  const long long pb = ((long long)blockIdx.y * gridDim.x + blockIdx.x) * NQ;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    sh[threadIdx.x] = acc[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[pb + q] = sh[0];
    __syncthreads();
  }
}

// Minimum and maximum of each variable of w over the interior cells of every
// block of a run (DebugOutput::WriteOutputFile, src/outputs/debug.cpp:30-84):
// a cell is a candidate for the minimum when w < FLT_MAX and for the maximum
// when w > FLT_MIN (the reference's initial values under strict
// comparisons); among candidates the smallest (largest) value wins, ties to
// the first cell in (k, j, i) order.  Per workgroup and variable: value and
// interior index ((k - ks) nx2 + j - js) nx1 + i - is (-1: no candidate).
__device__ inline void better_min(double v, long long x, double& bv, long long& bx) {
  if (x >= 0 && (bx < 0 || v < bv || (v == bv && x < bx))) {
    bv = v;
    bx = x;
  }
}
__device__ inline void better_max(double v, long long x, double& bv, long long& bx) {
  if (x >= 0 && (bx < 0 || v > bv || (v == bv && x < bx))) {
    bv = v;
    bx = x;
  }
}

template <int NM>
__global__ void __launch_bounds__(256)
    k_minmax(const KBlockPtr* __restrict__ bp, double* __restrict__ pv,
             long long* __restrict__ px) {
  constexpr int N = 5 + NM;
  constexpr double FMAXV = 3.40282346638528859812e+38, FMINV = 1.17549435082228750797e-38;
  __shared__ double sv[256];
  __shared__ long long sx[256];
  const KGrid& g = bp[blockIdx.y].g;
  const double* w = gptr(bp[blockIdx.y].w);
  const long long total = (long long)g.nx1 * g.nx2 * g.nx3;
  double mn[N], mx[N];
  long long imn[N], imx[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    mn[n] = mx[n] = 0.0;
    imn[n] = imx[n] = -1;
  }
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
    const long long c0 = g.at(0, k + g.ks, j + g.js, i + g.is);
    const long long x = ((long long)k * g.nx2 + j) * g.nx1 + i;
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const double v = w[c0 + n * g.c.n];
      if (v < FMAXV) better_min(v, x, mn[n], imn[n]);
      if (v > FMINV) better_max(v, x, mx[n], imx[n]);
    }
  }
  const long long pb = ((long long)blockIdx.y * gridDim.x + blockIdx.x) * N;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    for (int side = 0; side < 2; ++side) {
      sv[threadIdx.x] = side ? mx[n] : mn[n];
      sx[threadIdx.x] = side ? imx[n] : imn[n];
      __syncthreads();
      for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
          double bv = sv[threadIdx.x];
          long long bx = sx[threadIdx.x];
          if (side)
            better_max(sv[threadIdx.x + o], sx[threadIdx.x + o], bv, bx);
          else
            better_min(sv[threadIdx.x + o], sx[threadIdx.x + o], bv, bx);
          sv[threadIdx.x] = bv;
          sx[threadIdx.x] = bx;
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        pv[2 * (pb + n) + side] = sv[0];
        px[2 * (pb + n) + side] = sx[0];
      }
      __syncthreads();
    }
  }
}

// out[0] = v (a device scalar written in stream order).
static __global__ void k_set_u64(unsigned long long* __restrict__ out,
                                 unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = v;
}


static __global__ void k_set_scalar(double* __restrict__ out, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = v;
}

}  // namespace cg
