// kernels.hpp -- device kernels of the canoe dynamical-core step (gfx950).
//
// FP64 plain path.  Thread mappings put the unit-stride index i (x1, the
// vertical) on adjacent lanes wherever a kernel is cell- or face-parallel,
// so every global load and store of a wavefront is one contiguous segment
// per variable.  The column-sequential kernels (hydrostatic scan, block
// Thomas) put one column per lane.
#pragma once

#include "cg_physics.hpp"

namespace cg {

// ---------------------------------------------------------------------------
// K2: hydrostatic face pressure psf and log-mean cell pressure psv per
// column (Decomposition::ChangeToTemperature,
// src/snap/decomposition/change_to_temperature.cpp:40-111).
template <int NM>
__global__ void __launch_bounds__(256)
    k_hydrostatic_scan(KGrid g, KThermo th, const double* __restrict__ w,
                       double* __restrict__ psf, double* __restrict__ psv) {
  constexpr int IPR = 4 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long ncol = (long long)g.nx2 * g.nx3;
  if (t >= ncol) return;
  int j = g.js + (int)(t % g.nx2);
  int k = g.ks + (int)(t / g.nx2);
  const double grav = g.grav, dx = g.dx1;
  const int is = g.is, ie = g.ie;
  double* pf = psf + t * (g.nc1 + 1);
  double* pv = psv + t * g.nc1;
  {
    double RdTv = w[g.at(IPR, k, j, ie)] / w[g.at(0, k, j, ie)];
    pf[ie + 1] = w[g.at(IPR, k, j, ie)] * exp_d(grav * dx / (2 * RdTv));
  }
  double cur = pf[ie + 1];
  for (int i = ie; i >= is; --i) {
    cur = cur - grav * w[g.at(0, k, j, i)] * dx;
    pf[i] = cur;
  }
  double gb = (g.bc[0] == 0) ? -grav : grav;
  for (int i = is - 1; i >= is - NG; --i) {
    cur = cur - gb * w[g.at(0, k, j, i)] * dx;
    pf[i] = cur;
  }
  double gt = (g.bc[1] == 0) ? -grav : grav;
  cur = pf[ie + 1];
  for (int i = ie + 1; i <= ie + NG; ++i) {
    cur = cur + gt * w[g.at(0, k, j, i)] * dx;
    pf[i + 1] = cur;
  }
  double lo = pf[0];
  for (int i = 0; i < g.nc1; ++i) {
    double hi = pf[i + 1];
    double v;
    if (fabs(lo - hi) < 1.E-6)
      v = 0.5 * (lo + hi);
    else
      v = (lo - hi) / log_d(lo / hi);
    pv[i] = v;
    lo = hi;
  }
}

// R = Rd (1 + sum_vap q inv_mu_ratio_m1) of the decomposition (:100-104)
template <int NM>
CG_HD double decomp_r(const KThermo& th, const double* wc) {
  double feps = 1.;
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n <= th.nvapor) feps += wc[n] * th.inv_mu_ratio_m1[n - 1];
  return th.rd * feps;
}

// ---------------------------------------------------------------------------
// K1a: x1 face fluxes with the hydrostatic decomposition around the
// reconstruction (patches/19.calculate_fluxes.patch:23-47;
// change_to_temperature.cpp:107-110 and :155-164).  One lane per face,
// faces of a column on adjacent lanes.
template <int NM, int ORDER>
__global__ void __launch_bounds__(256)
    k_flux_x1(KGrid g, KThermo th, int rsolver, int decomp,
              const double* __restrict__ w, const double* __restrict__ psf,
              const double* __restrict__ psv, double* __restrict__ f1) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
  constexpr int LO = (ORDER == 5) ? 0 : 1, HI = (ORDER == 5) ? 5 : 4;
  const int nf = g.nx1 + 1;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long ncol = (long long)g.nx2 * g.nx3;
  if (t >= ncol * nf) return;
  int fi = (int)(t % nf);
  long long cidx = t / nf;
  int j = g.js + (int)(cidx % g.nx2);
  int k = g.ks + (int)(cidx / g.nx2);
  int i = g.is + fi;
  // window cells i-3 .. i+2 -> q[n][0..5]
  double q[N][6];
  double rdv[6];
#pragma unroll
  for (int m = LO; m <= HI; ++m) {
    double wc[N];
#pragma unroll
    for (int n = 0; n < N; ++n) wc[n] = w[g.at(n, k, j, i - 3 + m)];
    if (decomp) {
      rdv[m] = decomp_r<NM>(th, wc);
      double T = wc[IPR] / (wc[0] * rdv[m]);
      wc[0] = T;
      wc[IPR] = wc[IPR] - psv[cidx * g.nc1 + (i - 3 + m)];
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
    double ps = psf[cidx * (g.nc1 + 1) + i];
    wr[IPR] += ps;
    if (wr[IPR] < 0.) wr[IPR] = ps;
    double gamma = sqrt(rdv[3] * rdv[2]);
    wr[0] = wr[IPR] / (wr[0] * gamma);
    wl[IPR] += ps;
    if (wl[IPR] < 0.) wl[IPR] = ps;
    wl[0] = wl[IPR] / (wl[0] * gamma);
  }
  double fx[N];
  riemann<NM, 0, double>(th, rsolver, wl, wr, fx);
  const long long fs = (long long)g.nc3 * g.nc2 * (g.nc1 + 1);
  const long long b = ((long long)k * g.nc2 + j) * (g.nc1 + 1) + i;
#pragma unroll
  for (int n = 0; n < N; ++n) f1[n * fs + b] = fx[n];
}

// K1b/K1c: x2 and x3 face fluxes.  One lane per face, i on adjacent lanes.
template <int NM, int ORDER, int DIR>
__global__ void __launch_bounds__(256)
    k_flux_h(KGrid g, KThermo th, int rsolver, const double* __restrict__ w,
             double* __restrict__ fh) {
  constexpr int N = 5 + NM;
  constexpr int LO = (ORDER == 5) ? 0 : 1, HI = (ORDER == 5) ? 5 : 4;
  const int n1 = g.nx1;
  const int na = (DIR == 1) ? g.nx2 + 1 : g.nx2;  // j range
  const int nb = (DIR == 2) ? g.nx3 + 1 : g.nx3;  // k range
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n1 * na * nb) return;
  int i = g.is + (int)(t % n1);
  long long r = t / n1;
  int j = g.js + (int)(r % na);
  int k = g.ks + (int)(r / na);
  const long long s = (DIR == 1) ? (long long)g.nc1 : (long long)g.nc1 * g.nc2;
  const long long base = g.at(0, k, j, i);
  double q[N][6];
#pragma unroll
  for (int m = LO; m <= HI; ++m)
#pragma unroll
    for (int n = 0; n < N; ++n) q[n][m] = w[base + n * g.sn + (m - 3) * s];
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
  double fx[N];
  riemann<NM, DIR, double>(th, rsolver, wl, wr, fx);
  long long fs, b;
  if (DIR == 1) {
    fs = (long long)g.nc3 * (g.nc2 + 1) * g.nc1;
    b = ((long long)k * (g.nc2 + 1) + j) * g.nc1 + i;
  } else {
    fs = (long long)(g.nc3 + 1) * g.nc2 * g.nc1;
    b = ((long long)k * g.nc2 + j) * g.nc1 + i;
  }
#pragma unroll
  for (int n = 0; n < N; ++n) fh[n * fs + b] = fx[n];
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
  double wght;       // beta dt
  double visc, p0, cp, chi, heat_flux;
};

// K4a: du = 0 - wght div F (Hydro::AddFluxDivergence) then the source terms
// in HydroSourceTerms order: constant acceleration (upstream form, see
// patches/25.constant_acc.patch context), the straka diffusion
// (examples/2019-Li-snap/straka.cpp:126-190), the bottom heating.
template <int NM>
__global__ void __launch_bounds__(256)
    k_divergence(KGrid g, KThermo th, KSource src,
                 const double* __restrict__ f1, const double* __restrict__ f2,
                 const double* __restrict__ f3, const double* __restrict__ w,
                 double* __restrict__ du) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IVY = IVX + 1, IEN = 4 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return;
  int i = g.is + (int)(t % g.nx1);
  long long r = t / g.nx1;
  int j = g.js + (int)(r % g.nx2);
  int k = g.ks + (int)(r / g.nx2);
  const double A1 = g.dx2 * g.dx3, A2 = g.dx1 * g.dx3, A3 = g.dx1 * g.dx2;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  const long long fs1 = (long long)g.nc3 * g.nc2 * (g.nc1 + 1);
  const long long fs2 = (long long)g.nc3 * (g.nc2 + 1) * g.nc1;
  const long long fs3 = (long long)(g.nc3 + 1) * g.nc2 * g.nc1;
  const long long b1 = ((long long)k * g.nc2 + j) * (g.nc1 + 1) + i;
  const long long b2 = ((long long)k * (g.nc2 + 1) + j) * g.nc1 + i;
  const long long b3 = ((long long)k * g.nc2 + j) * g.nc1 + i;
  double d[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double dflx = A1 * f1[n * fs1 + b1 + 1] - A1 * f1[n * fs1 + b1];
    if (g.f2) dflx += A2 * f2[n * fs2 + b2 + g.nc1] - A2 * f2[n * fs2 + b2];
    if (g.f3)
      dflx += A3 * f3[n * fs3 + b3 + (long long)g.nc2 * g.nc1] -
              A3 * f3[n * fs3 + b3];
    d[n] = 0.0 - src.wght * dflx / vol;
  }
  const long long c0 = g.at(0, k, j, i);
  if (g.grav != 0.0) {
    double s = src.wght * w[c0] * g.grav;
    d[IVX] += s;
    d[IEN] += s * w[c0 + IVX * g.sn];
  }
  if (src.visc != 0.0 && g.f2) {
    const double K = src.visc, dx = g.dx1, dy = g.dx2, dt = src.wght;
    auto W = [&](int n, int jj, int ii) { return w[g.at(n, k, jj, ii)]; };
    auto theta = [&](int jj, int ii, double* tmp) {
      double wc[N];
#pragma unroll
      for (int n = 0; n < N; ++n) wc[n] = W(n, jj, ii);
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
    d[IEN] += src.wght * src.heat_flux / g.dx1;
#pragma unroll
  for (int n = 0; n < N; ++n) du[c0 + n * g.sn] = d[n];
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
  int i = g.is + (int)(t % g.nx1);
  long long r = t / g.nx1;
  int j = g.js + (int)(r % g.nx2);
  int k = g.ks + (int)(r / g.nx2);
  const double A1 = g.dx2 * g.dx3, A2 = g.dx1 * g.dx3, A3 = g.dx1 * g.dx2;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  const long long fs1 = (long long)g.nc3 * g.nc2 * (g.nc1 + 1);
  const long long fs2 = (long long)g.nc3 * (g.nc2 + 1) * g.nc1;
  const long long fs3 = (long long)(g.nc3 + 1) * g.nc2 * g.nc1;
  const long long b1 = ((long long)k * g.nc2 + j) * (g.nc1 + 1) + i;
  const long long b2 = ((long long)k * (g.nc2 + 1) + j) * g.nc1 + i;
  const long long b3 = ((long long)k * g.nc2 + j) * g.nc1 + i;
  const long long c0 = g.at(0, k, j, i);
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double dflx = A1 * f1[n * fs1 + b1 + 1] - A1 * f1[n * fs1 + b1];
    if (g.f2) dflx += A2 * f2[n * fs2 + b2 + g.nc1] - A2 * f2[n * fs2 + b2];
    if (g.f3)
      dflx += A3 * f3[n * fs3 + b3 + (long long)g.nc2 * g.nc1] -
              A3 * f3[n * fs3 + b3];
    du[c0 + n * g.sn] -= wght * dflx / vol;
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
// row scaled by the reciprocal; row exchanges by register selects.
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
    // pivot value is m[p][c]; zero pivot -> singular
#pragma unroll
    for (int r = c; r < B; ++r)
      if (p == r && m.a[r][c] == 0.0) ok = false;
#pragma unroll
    for (int r = c + 1; r < B; ++r) {
      bool sw = (p == r);
#pragma unroll
      for (int q = 0; q < B; ++q) {
        double mc = m.a[c][q], mr = m.a[r][q];
        m.a[c][q] = sw ? mr : mc;
        m.a[r][q] = sw ? mc : mr;
        double xc = x.a[c][q], xr = x.a[r][q];
        x.a[c][q] = sw ? xr : xc;
        x.a[r][q] = sw ? xc : xr;
      }
    }
    double inv = 1.0 / m.a[c][c];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      m.a[c][q] = m.a[c][q] * inv;
      x.a[c][q] = x.a[c][q] * inv;
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
      if (r == c) continue;
      double f = m.a[r][c];
#pragma unroll
      for (int q = 0; q < B; ++q) {
        m.a[r][q] = m.a[r][q] - f * m.a[c][q];
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
// |lambda_c| (the diagonal product adds only exact zeros).
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
  // RL = R * diag(lam) as a full product: sum over k of R[r][k] L[k][c]
  // with L[k][c] = 0 for k != c; computed as the same sum.
  double RL[5][5];
#pragma unroll
  for (int rr = 0; rr < 5; ++rr)
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double s = R[rr][0] * (c == 0 ? lam[0] : 0.0);
#pragma unroll
      for (int k = 1; k < 5; ++k) s += R[rr][k] * (c == k ? lam[k] : 0.0);
      RL[rr][c] = s;
    }
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
  int nlev;        // nx1
};

// Forward sweep of one column per lane: assembles a_i, b_i, c_i
// (full_correction.cpp:221-288), folds the reflecting walls (:286-288) and
// runs ForwardSweep (forward_backward.hpp:14-84), archiving a'_i (B x B)
// and delta_i (B) per level, lane-contiguous: arch[(lvl*(B*B+B)+e)*ncol+col].
template <int NM, int B>
__global__ void __launch_bounds__(64)
    k_implicit_fwd(KGrid g, KThermo th, KImplicit ip,
                   const double* __restrict__ w, const double* __restrict__ du,
                   double* __restrict__ arch, int* __restrict__ status) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IEN = 4 + NM;
  constexpr int AS = B * B + B;
  long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ip.ncol) return;
  int j = g.js + (int)(col % g.nx2);
  int k = g.ks + (int)(col / g.nx2);
  const int is = g.is, ie = g.ie;
  const double dt = ip.dt, grav = g.grav;
  const double gm1 = th.gammad - 1.;
  const double gm = 0.5 * (gm1 + gm1);
  const double aleft = g.dx2 * g.dx3, aright = g.dx2 * g.dx3;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  const long long c0 = g.at(0, k, j, 0);
  auto load = [&](double* dst, int i) {
#pragma unroll
    for (int n = 0; n < N; ++n) dst[n] = w[c0 + n * g.sn + i];
  };
  double wm[N], wc[N], wp[N], prim[N];
  load(wm, is - 1);
  load(wc, is);
  Mt<B> Am, Ap, dm, dc, dp;
  roe_average<NM>(prim, gm, wm, wc);
  abs_roe<NM, B>(th, prim, gm, Am);
  flux_jac<NM, B>(gm1, wm, dm);
  flux_jac<NM, B>(gm1, wc, dc);
  Mt<B> ap;   // a'_{i-1}
  Vt<B> dl;   // delta_{i-1}
  bool ok = true;
  for (int i = is; i <= ie; ++i) {
    load(wp, i + 1);
    roe_average<NM>(prim, gm, wc, wp);
    abs_roe<NM, B>(th, prim, gm, Ap);
    flux_jac<NM, B>(gm1, wp, dp);
    Mt<B> a, b, c;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int q = 0; q < B; ++q) {
        double dtq = (r == q) ? 1. / dt : 0.0;
        double phi = 0.0;
        if (B == 5) {
          if ((r == 1 && q == 0) || (r == 4 && q == 1)) phi = grav;
        } else {
          if ((r == 1 && q == 0) || (r == 2 && q == 1)) phi = grav;
        }
        a.a[r][q] = (Am.a[r][q] * aleft + Ap.a[r][q] * aright +
                     (aright - aleft) * dc.a[r][q]) /
                        (2. * vol) +
                    dtq - phi;
        b.a[r][q] = -(Am.a[r][q] + dm.a[r][q]) * aleft / (2. * vol);
        c.a[r][q] = -(Ap.a[r][q] - dp.a[r][q]) * aright / (2. * vol);
      }
    // reflecting walls: a_is += b_is Bnd, a_ie += c_ie Bnd
    if ((i == is && g.bc[0] != 1) || (i == ie && g.bc[1] != 1)) {
      Mt<B> bnd, t;
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q)
          bnd.a[r][q] = (r == q) ? ((r == 1) ? -1. : 1.) : 0.0;
      if (i == is && g.bc[0] != 1) {
        mmul<B>(b, bnd, t);
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
          for (int q = 0; q < B; ++q) a.a[r][q] += t.a[r][q];
      }
      if (i == ie && g.bc[1] != 1) {
        mmul<B>(c, bnd, t);
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
          for (int q = 0; q < B; ++q) a.a[r][q] += t.a[r][q];
      }
    }
    // rhs (forward_backward.hpp:30-36)
    Vt<B> rhs;
    {
      double r0 = du[c0 + i];
#pragma unroll
      for (int n = 1; n < IVX; ++n) r0 += du[c0 + n * g.sn + i];
      r0 /= dt;
      rhs.v[0] = r0;
      if (B == 5) {
        rhs.v[1] = du[c0 + IVX * g.sn + i] / dt;
        rhs.v[2] = du[c0 + (IVX + 1) * g.sn + i] / dt;
        rhs.v[3] = du[c0 + (IVX + 2) * g.sn + i] / dt;
        rhs.v[4] = du[c0 + IEN * g.sn + i] / dt;
      } else {
        rhs.v[1] = du[c0 + IVX * g.sn + i] / dt;
        rhs.v[2] = du[c0 + IEN * g.sn + i] / dt;
      }
    }
    Mt<B> inv;
    Vt<B> dn;
    if (i == is) {
      ok = minv<B>(a, inv) && ok;
      mvec<B>(inv, rhs, dn);
    } else {
      Mt<B> ba;
      mmul<B>(b, ap, ba);
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q) a.a[r][q] = a.a[r][q] - ba.a[r][q];
      ok = minv<B>(a, inv) && ok;
      Vt<B> bd, tv;
      mvec<B>(b, dl, bd);
#pragma unroll
      for (int r = 0; r < B; ++r) tv.v[r] = rhs.v[r] - bd.v[r];
      mvec<B>(inv, tv, dn);
    }
    mmul<B>(inv, c, ap);
    dl = dn;
    const long long lb = (long long)(i - is) * AS;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int q = 0; q < B; ++q)
        arch[(lb + r * B + q) * ip.ncol + col] = ap.a[r][q];
#pragma unroll
    for (int r = 0; r < B; ++r) arch[(lb + B * B + r) * ip.ncol + col] = dl.v[r];
    // shift the window
    Am = Ap;
    dm = dc;
    dc = dp;
#pragma unroll
    for (int n = 0; n < N; ++n) {
      wm[n] = wc[n];
      wc[n] = wp[n];
    }
  }
  if (!ok) atomicOr(status, 1);
}

// Back substitution (forward_backward.hpp:86-124) and the write-back of
// du (species split of the density increment, :221-246).
template <int NM, int B>
__global__ void __launch_bounds__(64)
    k_implicit_bwd(KGrid g, KImplicit ip, const double* __restrict__ w,
                   const double* __restrict__ arch, double* __restrict__ du) {
  constexpr int IVX = 1 + NM, IEN = 4 + NM;
  constexpr int AS = B * B + B;
  long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ip.ncol) return;
  int j = g.js + (int)(col % g.nx2);
  int k = g.ks + (int)(col / g.nx2);
  const int is = g.is, ie = g.ie;
  const long long c0 = g.at(0, k, j, 0);
  Vt<B> dn;  // delta_{i+1}
  for (int i = ie; i >= is; --i) {
    const long long lb = (long long)(i - is) * AS;
    Vt<B> d;
#pragma unroll
    for (int r = 0; r < B; ++r) d.v[r] = arch[(lb + B * B + r) * ip.ncol + col];
    if (i < ie) {
      Mt<B> a;
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q) a.a[r][q] = arch[(lb + r * B + q) * ip.ncol + col];
      Vt<B> ad;
      mvec<B>(a, dn, ad);
#pragma unroll
      for (int r = 0; r < B; ++r) d.v[r] = d.v[r] - ad.v[r];
    }
    dn = d;
    // write-back
    double dens = du[c0 + i];
#pragma unroll
    for (int n = 1; n < IVX; ++n) dens += du[c0 + n * g.sn + i];
    dens = d.v[0] - dens;
    double d0 = d.v[0];
    double dsp[NM > 0 ? NM : 1];
#pragma unroll
    for (int n = 1; n < IVX; ++n) {
      dsp[n - 1] = du[c0 + n * g.sn + i] + dens * w[c0 + n * g.sn + i];
      du[c0 + n * g.sn + i] = dsp[n - 1];
    }
    if (B == 5) {
      du[c0 + IVX * g.sn + i] = d.v[1];
      du[c0 + (IVX + 1) * g.sn + i] = d.v[2];
      du[c0 + (IVX + 2) * g.sn + i] = d.v[3];
      du[c0 + IEN * g.sn + i] = d.v[4];
    } else {
      du[c0 + IVX * g.sn + i] = d.v[1];
      du[c0 + IEN * g.sn + i] = d.v[2];
    }
#pragma unroll
    for (int n = 1; n < IVX; ++n) d0 -= dsp[n - 1];
    du[c0 + i] = d0;
  }
}

// ---------------------------------------------------------------------------
// K4b: end of stage per cell: u1 / u averaging (WeightedAve,
// implicit_hydro_tasks.cpp:146-160), u += du (:277-283), saturation
// adjustment on the last stage (UpdateAllConserved, :290-313), and
// ConservedToPrimitive into w.
struct KUpdate {
  int stage;  // 0, 1, 2
  double g1, g2;
  int satadj;
};

template <int NM>
__global__ void __launch_bounds__(256)
    k_update(KGrid g, KThermo th, KUpdate up, double* __restrict__ u,
             double* __restrict__ u1, const double* __restrict__ du,
             double* __restrict__ w, int* __restrict__ status) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IEN = 4 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return;
  int i = g.is + (int)(t % g.nx1);
  long long r = t / g.nx1;
  int j = g.js + (int)(r % g.nx2);
  int k = g.ks + (int)(r / g.nx2);
  const long long c0 = g.at(0, k, j, i);
  double uc[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double un = u[c0 + n * g.sn];
    double v;
    if (up.stage == 0) {
      v = 0.0 + 1.0 * un;
      u1[c0 + n * g.sn] = v;
    } else {
      v = up.g1 * un + up.g2 * u1[c0 + n * g.sn];
    }
    uc[n] = v + du[c0 + n * g.sn];
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
  double wc[N];
  cons_to_prim_cell<NM>(th, uc, wc);
#pragma unroll
  for (int n = 0; n < N; ++n) {
    u[c0 + n * g.sn] = uc[n];
    w[c0 + n * g.sn] = wc[n];
  }
}

// cons <-> prim on interior cells
template <int NM, int TO_PRIM>
__global__ void __launch_bounds__(256)
    k_eos(KGrid g, KThermo th, const double* __restrict__ src,
          double* __restrict__ dst) {
  constexpr int N = 5 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return;
  int i = g.is + (int)(t % g.nx1);
  long long r = t / g.nx1;
  int j = g.js + (int)(r % g.nx2);
  int k = g.ks + (int)(r / g.nx2);
  const long long c0 = g.at(0, k, j, i);
  double a[N], b[N];
#pragma unroll
  for (int n = 0; n < N; ++n) a[n] = src[c0 + n * g.sn];
  if (TO_PRIM)
    cons_to_prim_cell<NM>(th, a, b);
  else
    prim_to_cons_cell<NM>(th, a, b);
#pragma unroll
  for (int n = 0; n < N; ++n) dst[c0 + n * g.sn] = b[n];
}

// ---------------------------------------------------------------------------
// Ghost cells (upstream BoundaryValues order: x1 over interior (j,k), then x2
// over all i and interior k, then x3 over all i, j).  One lane per
// (ghost layer, side, line); variables looped.
__device__ inline void bc_pair(double* w, long long sn, int N, int vn, int bci,
                               int bco, long long a_lo, long long a_hi,
                               long long stride, int gh) {
  // a_lo: address of the first interior cell, a_hi: of the last, stride
  // between cells along the direction; gh = 1..3
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
  long long nl = (long long)g.nx2 * g.nx3;
  if (t >= nl * NG) return;
  int gh = 1 + (int)(t % NG);
  long long l = t / NG;
  int j = g.js + (int)(l % g.nx2), k = g.ks + (int)(l / g.nx2);
  bc_pair(w, g.sn, 5 + NM, 1 + NM, g.bc[0], g.bc[1], g.at(0, k, j, g.is),
          g.at(0, k, j, g.ie), 1, gh);
}

template <int NM>
__global__ void __launch_bounds__(256) k_bc_x2(KGrid g, double* w) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long nl = (long long)g.nc1 * g.nx3;
  if (t >= nl * NG) return;
  int i = (int)(t % g.nc1);
  long long l = t / g.nc1;
  int gh = 1 + (int)(l % NG);
  int k = g.ks + (int)(l / NG);
  bc_pair(w, g.sn, 5 + NM, 2 + NM, g.bc[2], g.bc[3], g.at(0, k, g.js, i),
          g.at(0, k, g.je, i), g.nc1, gh);
}

template <int NM>
__global__ void __launch_bounds__(256) k_bc_x3(KGrid g, double* w) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long nl = (long long)g.nc1 * g.nc2;
  if (t >= nl * NG) return;
  int i = (int)(t % g.nc1);
  long long l = t / g.nc1;
  int j = (int)(l % g.nc2);
  int gh = 1 + (int)(l / g.nc2);
  bc_pair(w, g.sn, 5 + NM, 3 + NM, g.bc[4], g.bc[5], g.at(0, g.ks, j, i),
          g.at(0, g.ke, j, i), (long long)g.nc1 * g.nc2, gh);
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
      int i = g.is + (int)(t % g.nx1);
      long long r = t / g.nx1;
      int j = g.js + (int)(r % g.nx2);
      int k = g.ks + (int)(r / g.nx2);
      s += u[g.at(n, k, j, i)] * vol;
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

}  // namespace cg
