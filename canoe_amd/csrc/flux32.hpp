// flux32.hpp -- FP32 reconstruction + Riemann fluxes, FP64 accumulation.
//
// The mixed form of the step: face states and fluxes in float, the flux
// divergence, sources and the state in double.  Pressure enters the float
// path only as a perturbation p' = p - pb about a base pb that is the same
// on both sides of a face:
//   x1 faces: pb = psf, the hydrostatic face pressure of the decomposition
//             (FP64, change_to_temperature.cpp:58-82); the reconstruction
//             already works on p - psv and T, as in the reference;
//   x2/x3:    pb = pref(i), a per-level reference pressure (FP64, horizontal
//             mean of the initial state); it is uniform across every x2/x3
//             face of a level, so its flux divergence vanishes exactly.
// The Riemann solvers are written in perturbation form: every difference of
// pressures uses p' (exact), absolute pressures (sound speed, enthalpy) use
// pb + p' in float, and the normal-momentum flux is returned without pb.
// The divergence kernel adds A1 (psf(i+1) - psf(i)) in FP64 to the x1
// normal-momentum flux difference, which is the part that balances gravity.
#pragma once

#include <type_traits>

#include "cg_timer.hpp"
#include "kernels.hpp"

namespace cg {

// Single-precision division and square root of the mixed form on the
// hardware reciprocal / square-root instructions (1 ulp) instead of the
// correctly rounded sequences; the FP32 faces are an approximation to the
// FP64 ones anyway, checked against them by relative error.
// The FP32 face arithmetic (reconstruction, Riemann fluxes) lets the compiler
// fuse multiplies and adds; FP64 code keeps -ffp-contract=off.
#define CG_FP32_CONTRACT _Pragma("clang fp contract(fast)")

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * frcp(b); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Upper-face WENO5 value (Jiang & Shu 1996, eps = 1e-6) of the cell c of the
// window (a, b, c, d, e); the lower face is weno5_up_f(e, d, c, b, a).
__device__ __forceinline__ float weno5_up_f(float a, float b, float c, float d,
                                            float e) {
  CG_FP32_CONTRACT
  const float eps = 1.0e-6f;
  float s0 = a - 2.f * b + c, t0 = a - 4.f * b + 3.f * c;
  float s1 = b - 2.f * c + d, t1 = b - d;
  float s2 = c - 2.f * d + e, t2 = 3.f * c - 4.f * d + e;
  float b0 = (13.f / 12.f) * (s0 * s0) + 0.25f * (t0 * t0);
  float b1 = (13.f / 12.f) * (s1 * s1) + 0.25f * (t1 * t1);
  float b2 = (13.f / 12.f) * (s2 * s2) + 0.25f * (t2 * t2);
  float e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  float a0 = 0.1f * frcp(e0 * e0), a1 = 0.6f * frcp(e1 * e1),
        a2 = 0.3f * frcp(e2 * e2);
  float p0 = (2.f * a - 7.f * b + 11.f * c) * (1.f / 6.f);
  float p1 = (-b + 5.f * c + 2.f * d) * (1.f / 6.f);
  float p2 = (2.f * c + 5.f * d - e) * (1.f / 6.f);
  return (a0 * p0 + a1 * p1 + a2 * p2) * frcp(a0 + a1 + a2);
}

// Both faces of cell c of the window (a, b, c, d, e) from one set of
// smoothness indicators: qm = weno5_up_f(e, d, c, b, a), qp = weno5_up_f(a,
// b, c, d, e).
__device__ __forceinline__ void weno5_both_f(float a, float b, float c,
                                             float d, float e, float* qm,
                                             float* qp) {
  CG_FP32_CONTRACT
  const float eps = 1.0e-6f;
  float s0 = a - 2.f * b + c, t0 = a - 4.f * b + 3.f * c;
  float s1 = b - 2.f * c + d, t1 = b - d;
  float s2 = c - 2.f * d + e, t2 = 3.f * c - 4.f * d + e;
  float b0 = (13.f / 12.f) * (s0 * s0) + 0.25f * (t0 * t0);
  float b1 = (13.f / 12.f) * (s1 * s1) + 0.25f * (t1 * t1);
  float b2 = (13.f / 12.f) * (s2 * s2) + 0.25f * (t2 * t2);
  float e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  float g0 = frcp(e0 * e0), g1 = frcp(e1 * e1), g2 = frcp(e2 * e2);
  float p0 = (2.f * a - 7.f * b + 11.f * c) * (1.f / 6.f);
  float p1 = (-b + 5.f * c + 2.f * d) * (1.f / 6.f);
  float p2 = (2.f * c + 5.f * d - e) * (1.f / 6.f);
  float m0 = (2.f * c + 5.f * b - a) * (1.f / 6.f);
  float m1 = (-d + 5.f * c + 2.f * b) * (1.f / 6.f);
  float m2 = (2.f * e - 7.f * d + 11.f * c) * (1.f / 6.f);
  float a0 = 0.1f * g0, a1 = 0.6f * g1, a2 = 0.3f * g2;
  *qp = (a0 * p0 + a1 * p1 + a2 * p2) * frcp(a0 + a1 + a2);
  float c0 = 0.3f * g0, c1 = 0.6f * g1, c2 = 0.1f * g2;
  *qm = (c0 * m0 + c1 * m1 + c2 * m2) * frcp(c0 + c1 + c2);
}

// PLM (van Leer) upper face of cell b of (a, b, c); lower face plm_up_f(c, b, a).
__device__ __forceinline__ float plm_up_f(float a, float b, float c) {
  CG_FP32_CONTRACT
  float dql = b - a, dqr = c - b;
  float dq2 = dql * dqr;
  float dqm = dq2 > 0.f ? 2.f * dq2 * frcp(dql + dqr) : 0.f;
  return b + 0.5f * dqm;
}

// get_gamma in float (src/interface/eos.hpp:47-67)
template <int NM>
__device__ inline float gamma_f(const KThermo& th, const float* u) {
  CG_FP32_CONTRACT
  float fsig = 1.f, feps = 1.f;
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n <= th.nvapor) {
      fsig += u[n] * (float)th.cv_ratio_m1[n - 1];
      feps += u[n] * (float)th.inv_mu_ratio_m1[n - 1];
    }
#pragma unroll
  for (int n = 1; n <= NM; ++n)
    if (n > th.nvapor) {
      fsig += u[n] * (float)th.cv_ratio_m1[n - 1];
      feps -= u[n];
    }
  return 1.f + ((float)th.gammad - 1.f) * fdiv(feps, fsig);
}

// LMARS (src/snap/riemann/lmars.cpp:19-85) in perturbation form: wl[IPR],
// wr[IPR] hold p - pb.  Normal-momentum flux excludes pb.
template <int NM, int DIR>
__device__ inline void lmars_p(const KThermo& th, const float* wli,
                               const float* wri, float pb, float* flx) {
  CG_FP32_CONTRACT
  constexpr int IVX = 1 + NM, IPR = IVX + 3, IDN = 0, IEN = IPR;
  constexpr int ivx = IVX + DIR;
  constexpr int ivy = IVX + ((ivx - IVX) + 1) % 3;
  constexpr int ivz = IVX + ((ivx - IVX) + 2) % 3;
  float pl = wli[IPR] + pb, pr = wri[IPR] + pb;
  float kappal = frcp(gamma_f<NM>(th, wli) - 1.f);
  float kappar = frcp(gamma_f<NM>(th, wri) - 1.f);
  float hl = fdiv(pl, wli[IDN]) * (kappal + 1.f) +
             0.5f * (wli[ivx] * wli[ivx] + wli[ivy] * wli[ivy] +
                     wli[ivz] * wli[ivz]);
  float hr = fdiv(pr, wri[IDN]) * (kappar + 1.f) +
             0.5f * (wri[ivx] * wri[ivx] + wri[ivy] * wri[ivy] +
                     wri[ivz] * wri[ivz]);
  float rhobar = 0.5f * (wli[IDN] + wri[IDN]);
  float cbar = fsqrt(fdiv(0.5f * (1.f + (frcp(kappar) + frcp(kappal)) * 0.5f) *
                              (pl + pr), rhobar));
  float rc = rhobar * cbar;
  float pbar = 0.5f * (wli[IPR] + wri[IPR]) + 0.5f * rc * (wli[ivx] - wri[ivx]);
  float ubar = 0.5f * (wli[ivx] + wri[ivx]) +
               0.5f * frcp(rc) * (wli[IPR] - wri[IPR]);
  float rdl = 1.f, rdr = 1.f;
#pragma unroll
  for (int n = 1; n < IVX; ++n) {
    rdl -= wli[n];
    rdr -= wri[n];
  }
  const bool up = ubar > 0.f;
  float ws[5 + NM];
#pragma unroll
  for (int n = 0; n < 5 + NM; ++n) ws[n] = up ? wli[n] : wri[n];
  float rd = up ? rdl : rdr;
  float h = up ? hl : hr;
  float m = ubar * ws[IDN];
  flx[IDN] = m * rd;
#pragma unroll
  for (int n = 1; n < IVX; ++n) flx[n] = m * ws[n];
  flx[ivx] = m * ws[ivx] + pbar;
  flx[ivy] = m * ws[ivy];
  flx[ivz] = m * ws[ivz];
  flx[IEN] = m * h;
}

// HLLC (src/snap/riemann/hllc_transform.cpp:25-233) in perturbation form.
template <int NM, int DIR>
__device__ inline void hllc_p(const KThermo& th, const float* wl,
                              const float* wr, float pb, float* flx) {
  CG_FP32_CONTRACT
  constexpr int IVX = 1 + NM, IVY = IVX + 1, IVZ = IVX + 2, IPR = IVX + 3;
  constexpr int IDN = 0, IEN = IPR, N = 5 + NM;
  constexpr int ivx = IVX + DIR;
  constexpr int ivy = IVX + ((ivx - IVX) + 1) % 3;
  constexpr int ivz = IVX + ((ivx - IVX) + 2) % 3;
  const float TINY = 1.0e-20f;
  float wli[N], wri[N], fl[N], fr[N], flxi[N];
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
  float pl = wli[IPR] + pb, pr = wri[IPR] + pb;
  float gammal = gamma_f<NM>(th, wli), gammar = gamma_f<NM>(th, wri);
  float kappal = frcp(gammal - 1.f), kappar = frcp(gammar - 1.f);
  float cl = fsqrt(fdiv(gammal * pl, wli[IDN]));
  float cr = fsqrt(fdiv(gammar * pr, wri[IDN]));
  float KE_l = 0.5f * wli[IDN] *
               (wli[IVX] * wli[IVX] + wli[IVY] * wli[IVY] + wli[IVZ] * wli[IVZ]);
  float KE_r = 0.5f * wri[IDN] *
               (wri[IVX] * wri[IVX] + wri[IVY] * wri[IVY] + wri[IVZ] * wri[IVZ]);
  float el = pl * kappal + KE_l;
  float er = pr * kappar + KE_r;
  float rhoa = 0.5f * (wli[IDN] + wri[IDN]);
  float ca = 0.5f * (cl + cr);
  float pmid = 0.5f * (pl + pr + (wli[IVX] - wri[IVX]) * rhoa * ca);
  float ql = (pmid <= pl) ? 1.f
                          : fsqrt(1.f + fdiv(gammal + 1.f, 2.f * gammal) *
                                            (fdiv(pmid, pl) - 1.f));
  float qr = (pmid <= pr) ? 1.f
                          : fsqrt(1.f + fdiv(gammar + 1.f, 2.f * gammar) *
                                            (fdiv(pmid, pr) - 1.f));
  float al = wli[IVX] - cl * ql;
  float ar = wri[IVX] + cr * qr;
  float bp = ar > 0.f ? ar : TINY;
  float bm = al < 0.f ? al : -TINY;
  float vxl = wli[IVX] - al;
  float vxr = wri[IVX] - ar;
  float tl = wli[IPR] + vxl * wli[IDN] * wli[IVX];  // t - pb
  float tr = wri[IPR] + vxr * wri[IDN] * wri[IVX];
  float ml = wli[IDN] * vxl;
  float mr = -(wri[IDN] * vxr);
  const float imm = frcp(ml + mr);
  float am = (tl - tr) * imm;
  float cp = (ml * tr + mr * tl) * imm;  // contact pressure - pb
  cp = cp > -pb ? cp : -pb;
  vxl = wli[IVX] - bm;
  vxr = wri[IVX] - bp;
  float rdl = 1.f, rdr = 1.f;
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
  fl[IEN] = el * vxl + pl * wli[IVX];
  fr[IEN] = er * vxr + pr * wri[IVX];
  float sl, sr, sm;
  if (am >= 0.f) {
    const float r = frcp(am - bm);
    sl = am * r;
    sr = 0.f;
    sm = -bm * r;
  } else {
    const float r = frcp(bp - am);
    sl = 0.f;
    sr = -am * r;
    sm = bp * r;
  }
#pragma unroll
  for (int n = 0; n < IVX; ++n) flxi[n] = sl * fl[n] + sr * fr[n];
  flxi[IVX] = sl * fl[IVX] + sr * fr[IVX] + sm * cp;
  flxi[IVY] = sl * fl[IVY] + sr * fr[IVY];
  flxi[IVZ] = sl * fl[IVZ] + sr * fr[IVZ];
  flxi[IEN] = sl * fl[IEN] + sr * fr[IEN] + sm * (cp + pb) * am;
#pragma unroll
  for (int n = 0; n < IVX; ++n) flx[n] = flxi[n];
  flx[ivx] = flxi[IVX];
  flx[ivy] = flxi[IVY];
  flx[ivz] = flxi[IVZ];
  flx[IEN] = flxi[IEN];
}

template <int NM, int DIR>
__device__ inline void riemann_p(const KThermo& th, int solver, const float* wl,
                                 const float* wr, float pb, float* flx) {
  if (solver == 1)
    hllc_p<NM, DIR>(th, wl, wr, pb, flx);
  else
    lmars_p<NM, DIR>(th, wl, wr, pb, flx);
}

// Face states from the 6-cell window q[0..5] around the face between q[2]
// and q[3]: the upper face of q[2] and the lower face of q[3].
template <int ORDER>
__device__ __forceinline__ void recon_pair_f(const float* q, float* l, float* r) {
  if (ORDER == 5) {
    *l = weno5_up_f(q[0], q[1], q[2], q[3], q[4]);
    *r = weno5_up_f(q[5], q[4], q[3], q[2], q[1]);
  } else {
    *l = plm_up_f(q[1], q[2], q[3]);
    *r = plm_up_f(q[4], q[3], q[2]);
  }
}

// Decomposition of x1 stencil cell i of the column at colb (offset of
// n = 0, i = 0): T in float and the perturbation p - psv as an FP64
// difference rounded to float; a cell beyond an outflow x1 side takes the
// boundary cell's (decomp_cell, change_to_temperature.cpp:113-127).  wc
// holds cell i.  Returns R of cell i itself (its gamma_ in the face
// restore, :158).
template <int NM>
__device__ __forceinline__ double decomp32(const KGrid& g, const KThermo& th,
                                           const double* __restrict__ w,
                                           long long colb,
                                           const double* __restrict__ psv,
                                           long long pc, int i, const double* wc,
                                           float& T, float& pp) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
  const double r = decomp_r<NM>(th, wc);
  const int ib = decomp_cell(g, i);
  double pb = wc[IPR], db = wc[0], rb = r;
  if (ib != i) {
    double wb[N];
#pragma unroll
    for (int n = 0; n < N; ++n) wb[n] = w[colb + n * g.c.n + (long long)ib * g.c.i];
    pb = wb[IPR];
    db = wb[0];
    rb = decomp_r<NM>(th, wb);
  }
  T = fdiv((float)pb, (float)db * (float)rb);
  pp = (float)(pb - psv[pc + ib * g.ps_i]);
  return r;
}

// Window slot m of an x1 march: cell i of column (k, j) as floats, with T
// and the perturbation p - psv (an FP64 difference rounded to float) in
// place of density and pressure when decomp; rdv[m] = its R (decomp32).
template <int NM>
__device__ __forceinline__ void x1_window_cell(const KGrid& g, const KThermo& th,
                                               int decomp, const double* __restrict__ w,
                                               const double* __restrict__ psv,
                                               long long pc, int k, int j, int i,
                                               float (*q)[6], double* rdv, int m) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
  double wc[N];
#pragma unroll
  for (int n = 0; n < N; ++n) wc[n] = w[g.at(n, k, j, i)];
#pragma unroll
  for (int n = 0; n < N; ++n) q[n][m] = (float)wc[n];
  if (decomp)  // T in float; the perturbation p - psv in FP64 before rounding
    rdv[m] = decomp32<NM>(g, th, w, g.at(0, k, j, 0), psv, pc, i, wc, q[0][m], q[IPR][m]);
}

// x1 face i from the window q / rdv of cells i-3 .. i+2 (x1_window_cell):
// reconstruction and LMARS/HLLC in float about pb = psf(i), pb left out of
// the normal momentum flux.
template <int NM, int ORDER>
__device__ __forceinline__ void x1_face_from_window(const KGrid& g, const KThermo& th,
                                                    int rsolver, int decomp,
                                                    const double* __restrict__ psf,
                                                    long long pc, int i, float (*q)[6],
                                                    const double* rdv, float* fx) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
  float wl[N], wr[N];
#pragma unroll
  for (int n = 0; n < N; ++n) recon_pair_f<ORDER>(q[n], &wl[n], &wr[n]);
  float pb = 0.f;
  if (decomp) {
    pb = (float)psf[pc + i * g.ps_i];
    float gam = fsqrt((float)(rdv[3] * rdv[2]));
    // RestoreFromTemperature (:155-164): p = psf + p', clamped at psf
    if (wr[IPR] + pb < 0.f) wr[IPR] = 0.f;
    if (wl[IPR] + pb < 0.f) wl[IPR] = 0.f;
    wr[0] = fdiv(pb + wr[IPR], wr[0] * gam);
    wl[0] = fdiv(pb + wl[IPR], wl[0] * gam);
  }
  face_check<NM, float>(g.status, wl, wr, pb);
  riemann_p<NM, 0>(th, rsolver, wl, wr, pb, fx);
}

// x1 face t (of the nx1 + 1 faces of each interior column) in float:
// decomposition in FP64 (T, p - psv), reconstruction and LMARS/HLLC in
// float about pb = psf(i); the normal momentum flux excludes pb (the caller
// adds the face pressures in FP64).  Sets the face's (k, j, i).
template <int NM, int ORDER>
__device__ __forceinline__ void flux32_x1_face(const KGrid& g, const KThermo& th,
                                               int rsolver, int decomp,
                                               const double* __restrict__ w,
                                               const double* __restrict__ psf,
                                               const double* __restrict__ psv,
                                               long long t, int& i, int& j, int& k,
                                               float* fx) {
  constexpr int N = 5 + NM;
  constexpr int LO = (ORDER == 5) ? 0 : 1, HI = (ORDER == 5) ? 5 : 4;
  g.decode(t, g.nx1 + 1, g.nx2, g.nx3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  const long long pc = g.pcol(k, j) * g.ps_c;
  float q[N][6];
  double rdv[6];
#pragma unroll
  for (int m = LO; m <= HI; ++m)
    x1_window_cell<NM>(g, th, decomp, w, psv, pc, k, j, i - 3 + m, q, rdv, m);
  x1_face_from_window<NM, ORDER>(g, th, rsolver, decomp, psf, pc, i, q, rdv, fx);
}

template <int NM, int ORDER>
__global__ void __launch_bounds__(256)
    k_flux32_x1(KGrid g, KThermo th, int rsolver, int decomp,
                const double* __restrict__ w, const double* __restrict__ psf,
                const double* __restrict__ psv, float* __restrict__ f1) {
  constexpr int N = 5 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)(g.nx1 + 1) * g.nx2 * g.nx3) return;
  int i, j, k;
  float fx[N];
  flux32_x1_face<NM, ORDER>(g, th, rsolver, decomp, w, psf, psv, t, i, j, k, fx);
#pragma unroll
  for (int n = 0; n < N; ++n) f1[g.f[0].at(n, k, j, i)] = fx[n];
}

// x2 / x3 faces in float about pb = pref(i).
template <int NM, int ORDER, int DIR>
__global__ void __launch_bounds__(256)
    k_flux32_h(KGrid g, KThermo th, int rsolver, const double* __restrict__ w,
               const double* __restrict__ pref, float* __restrict__ fh) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
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
  const double pr0 = pref[i];
  float q[N][6];
#pragma unroll
  for (int m = LO; m <= HI; ++m)
#pragma unroll
    for (int n = 0; n < N; ++n) {
      double v = w[base + n * g.c.n + (m - 3) * s];
      if (n == IPR) v -= pr0;
      q[n][m] = (float)v;
    }
  float wl[N], wr[N];
#pragma unroll
  for (int n = 0; n < N; ++n) recon_pair_f<ORDER>(q[n], &wl[n], &wr[n]);
  float fx[N];
  face_check<NM, float>(g.status, wl, wr, (float)pr0);
  riemann_p<NM, DIR>(th, rsolver, wl, wr, (float)pr0, fx);
#pragma unroll
  for (int n = 0; n < N; ++n) fh[g.f[DIR].at(n, k, j, i)] = fx[n];
}

// du = 0 - wght div F with float fluxes accumulated in FP64; the hydrostatic
// face pressure difference enters the x1 normal-momentum flux in FP64.
template <int NM>
__global__ void __launch_bounds__(256)
    k_divergence32(KGrid g, KThermo th, KSource src,
                   const float* __restrict__ f1, const float* __restrict__ f2,
                   const float* __restrict__ f3, const double* __restrict__ psf,
                   const double* __restrict__ psv, int decomp,
                   const double* __restrict__ w, double* __restrict__ du) {
  constexpr int N = 5 + NM, IVX = 1 + NM;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return;
  int i, j, k;
  g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  const double A1 = g.dx2 * g.dx3, A2 = g.dx1 * g.dx3, A3 = g.dx1 * g.dx2;
  const double vol = g.dx1 * g.dx2 * g.dx3;
  const double wv = src.weight() / vol;
  double d[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double dflx = A1 * (double)f1[g.f[0].at(n, k, j, i + 1)] -
                  A1 * (double)f1[g.f[0].at(n, k, j, i)];
    if (n == IVX && decomp) {
      const long long pc = g.pcol(k, j) * g.ps_c;
      dflx += A1 * (psf[pc + (i + 1) * g.ps_i] - psf[pc + i * g.ps_i]);
    }
    if (g.f2)
      dflx += A2 * (double)f2[g.f[1].at(n, k, j + 1, i)] -
              A2 * (double)f2[g.f[1].at(n, k, j, i)];
    if (g.f3)
      dflx += A3 * (double)f3[g.f[2].at(n, k + 1, j, i)] -
              A3 * (double)f3[g.f[2].at(n, k, j, i)];
    d[n] = 0.0 - wv * dflx;
  }
  add_sources_cell<NM>(g, th, src, w, decomp ? psv : nullptr, k, j, i, d);
#pragma unroll
  for (int n = 0; n < N; ++n) du[g.at(n, k, j, i)] = d[n];
}

// x2 faces from one WENO evaluation per cell: lane l of a wavefront holds
// cell cb + l of an x2 row and gets its 5-cell window from lanes l-2 .. l+2
// by cross-lane shuffles; the face between lanes l-1 and l takes the upper
// state of lane l-1 and the lower state of lane l.  Lanes 3..61 produce the
// faces cb+3 .. cb+61 (59 per wavefront); wavefronts of a row overlap by 5
// cells.
template <int NM, int ORDER>
__global__ void __launch_bounds__(256)
    k_mix_x2(KGrid g, KThermo th, int rsolver, const double* __restrict__ w,
             const double* __restrict__ pref, float* __restrict__ f2) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
  constexpr int FPW = 59;  // faces per wavefront
  const int nf = g.nx2 + 1;
  const int wpr = (nf + FPW - 1) / FPW;  // wavefronts per (i, k) row
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const long long nwave = (long long)g.nx1 * g.nx3 * wpr;
  if (wave >= nwave) return;  // whole wavefronts leave together
  const int wr = (int)(wave % wpr);
  const long long row = wave / wpr;
  const int i = g.is + (int)(row % g.nx1);
  const int k = g.ks + (int)(row / g.nx1);
  const int cb = g.js - 3 + FPW * wr;
  const int c = cb + lane;
  const int cc = min(max(c, 0), g.nc2 - 1);  // lanes past the row load a copy
  const double prd = pref[i];
  const long long base = g.at(0, k, cc, i);
  float v[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double x = w[base + n * g.c.n];
    if (n == IPR) x -= prd;
    v[n] = (float)x;
  }
  float qm[N], qp[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    float a = __shfl(v[n], lane - 2 < 0 ? 0 : lane - 2);
    float b = __shfl(v[n], lane - 1 < 0 ? 0 : lane - 1);
    float d = __shfl(v[n], lane + 1 > 63 ? 63 : lane + 1);
    float e = __shfl(v[n], lane + 2 > 63 ? 63 : lane + 2);
    if (ORDER == 5) {
      weno5_both_f(a, b, v[n], d, e, &qm[n], &qp[n]);
    } else {
      qp[n] = plm_up_f(b, v[n], d);
      qm[n] = plm_up_f(d, v[n], b);
    }
  }
  float ql[N];
#pragma unroll
  for (int n = 0; n < N; ++n) ql[n] = __shfl(qp[n], lane - 1 < 0 ? 0 : lane - 1);
  const int f = c;  // face between cells c-1 and c
  if (lane < 3 || lane > 61 || f > g.je + 1) return;
  float fx[N];
  face_check<NM, float>(g.status, ql, qm, (float)prd);
  riemann_p<NM, 1>(th, rsolver, ql, qm, (float)prd, fx);
#pragma unroll
  for (int n = 0; n < N; ++n) f2[g.f[1].at(n, k, f, i)] = fx[n];
}

// x3 faces marched along k over a segment of faces, one lane per (i, j,
// segment), lanes along j: every cell's two face states come from one WENO
// evaluation (weno5_both_f) on a sliding 5-cell window, and each face takes
// the upper state of the cell below it (carried) and the lower state of the
// cell above.  Same face fluxes as k_flux32_h<x3> up to float rounding.
template <int NM, int ORDER>
__global__ void __launch_bounds__(256)
    k_mix_x3(KGrid g, KThermo th, int rsolver, int seg,
             const double* __restrict__ w, const double* __restrict__ pref,
             float* __restrict__ f3) {
  constexpr int N = 5 + NM, IPR = 4 + NM;
  const int nf = g.nx3 + 1;
  const int nseg = (nf + seg - 1) / seg;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx1 * g.nx2 * nseg) return;
  const int j = g.js + (int)(t % g.nx2);
  const long long r = t / g.nx2;
  const int i = g.is + (int)(r % g.nx1);
  const int sg = (int)(r / g.nx1);
  const int f0 = g.ks + sg * seg, f1 = min(f0 + seg, g.ks + nf);  // faces f0..f1-1
  const double prd = pref[i];
  float q[N][5];
  auto load = [&](int m, int k) {
    const long long c = g.at(0, k, j, i);
#pragma unroll
    for (int n = 0; n < N; ++n) {
      double v = w[c + n * g.c.n];
      if (n == IPR) v -= prd;
      q[n][m] = (float)v;
    }
  };
  auto cell = [&](float* qm, float* qp) {
#pragma unroll
    for (int n = 0; n < N; ++n) {
      if (ORDER == 5) {
        weno5_both_f(q[n][0], q[n][1], q[n][2], q[n][3], q[n][4], &qm[n], &qp[n]);
      } else {
        qp[n] = plm_up_f(q[n][1], q[n][2], q[n][3]);
        qm[n] = plm_up_f(q[n][3], q[n][2], q[n][1]);
      }
    }
  };
  // window of cell f0 - 1: cells f0 - 3 .. f0 + 1
#pragma unroll
  for (int m = 0; m < 5; ++m) load(m, f0 - 3 + m);
  float qm[N], qpp[N];
  cell(qm, qpp);
  for (int f = f0; f < f1; ++f) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = 0; n < N; ++n) q[n][m] = q[n][m + 1];
    load(4, f + 2);
    float qp[N];
    cell(qm, qp);
    float fx[N];
    face_check<NM, float>(g.status, qpp, qm, (float)prd);
    riemann_p<NM, 2>(th, rsolver, qpp, qm, (float)prd, fx);
#pragma unroll
    for (int n = 0; n < N; ++n) f3[g.f[2].at(n, f, j, i)] = fx[n];
#pragma unroll
    for (int n = 0; n < N; ++n) qpp[n] = qp[n];
  }
}

// x1 faces marched up a column segment, fused with the flux divergence of
// every cell of the segment: one lane per (column, segment), lanes along the
// unit-stride index.  The face states come from a sliding 6-cell window
// (the decomposition of k_flux32_x1: T and p - psv with FP64 differences);
// each cell then takes its two x1 face fluxes from the march, its x2 / x3
// face fluxes from f2 / f3, the hydrostatic term A1 (psf(i+1) - psf(i)) in
// FP64, and the sources, and writes du once.
template <int NM, int ORDER>
__global__ void __launch_bounds__(256)
    k_mix_x1_div(KGrid g, KThermo th, KSource src, int rsolver, int decomp,
                 int seg, const double* __restrict__ w,
                 const double* __restrict__ psf, const double* __restrict__ psv,
                 const float* __restrict__ f2, const float* __restrict__ f3,
                 double* __restrict__ du) {
  constexpr int N = 5 + NM, IPR = 4 + NM, IVX = 1 + NM;
  constexpr int LO = (ORDER == 5) ? 0 : 1, HI = (ORDER == 5) ? 5 : 4;
  const long long ncol = (long long)g.nx2 * g.nx3;
  const int nseg = (g.nx1 + seg - 1) / seg;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * nseg) return;
  const long long col = t % ncol;
  const int sg = (int)(t / ncol);
  const int j = g.js + (int)(col % g.nx2), k = g.ks + (int)(col / g.nx2);
  const int c0 = g.is + sg * seg;
  const int c1 = min(c0 + seg, g.ie + 1);  // cells c0 .. c1-1
  const long long pc = g.pcol(k, j) * g.ps_c;
  const double A1 = g.dx2 * g.dx3, A2 = g.dx1 * g.dx3, A3 = g.dx1 * g.dx2;
  const double wv = src.weight() / (g.dx1 * g.dx2 * g.dx3);
  float q[N][6];
  float rdv[6];
  auto load = [&](int m, int i) {
    double wc[N];
#pragma unroll
    for (int n = 0; n < N; ++n) wc[n] = w[g.at(n, k, j, i)];
#pragma unroll
    for (int n = 0; n < N; ++n) q[n][m] = (float)wc[n];
    if (decomp)
      rdv[m] = (float)decomp32<NM>(g, th, w, g.at(0, k, j, 0), psv, pc, i, wc,
                                   q[0][m], q[IPR][m]);
  };
  // face f (between cells f-1 and f) uses window cells f-3 .. f+2
  auto face = [&](int f, float* fx) {
    float wl[N], wr[N];
#pragma unroll
    for (int n = 0; n < N; ++n) recon_pair_f<ORDER>(q[n], &wl[n], &wr[n]);
    float pb = 0.f;
    if (decomp) {
      pb = (float)psf[pc + f * g.ps_i];
      float gam = fsqrt(rdv[3] * rdv[2]);
      if (wr[IPR] + pb < 0.f) wr[IPR] = 0.f;
      if (wl[IPR] + pb < 0.f) wl[IPR] = 0.f;
      wr[0] = fdiv(pb + wr[IPR], wr[0] * gam);
      wl[0] = fdiv(pb + wl[IPR], wl[0] * gam);
    }
    face_check<NM, float>(g.status, wl, wr, pb);
    riemann_p<NM, 0>(th, rsolver, wl, wr, pb, fx);
  };
#pragma unroll
  for (int m = LO; m <= HI; ++m) load(m, c0 - 3 + m);
  float fm[N];
  face(c0, fm);
  double pfm = psf[pc + c0 * g.ps_i];
  for (int i = c0; i < c1; ++i) {
    // slide the window up one cell: cells i-2 .. i+3 for face i+1
#pragma unroll
    for (int m = LO; m < HI; ++m) {
#pragma unroll
      for (int n = 0; n < N; ++n) q[n][m] = q[n][m + 1];
      rdv[m] = rdv[m + 1];
    }
    load(HI, i + 1 - 3 + HI);
    float fp[N];
    face(i + 1, fp);
    const double pfp = psf[pc + (i + 1) * g.ps_i];
    double d[N];
#pragma unroll
    for (int n = 0; n < N; ++n) {
      double dflx = A1 * ((double)fp[n] - (double)fm[n]);
      if (n == IVX && decomp) dflx += A1 * (pfp - pfm);
      if (g.f2)
        dflx += A2 * ((double)f2[g.f[1].at(n, k, j + 1, i)] -
                      (double)f2[g.f[1].at(n, k, j, i)]);
      if (g.f3)
        dflx += A3 * ((double)f3[g.f[2].at(n, k + 1, j, i)] -
                      (double)f3[g.f[2].at(n, k, j, i)]);
      d[n] = 0.0 - wv * dflx;
    }
    add_sources_cell<NM>(g, th, src, w, decomp ? psv : nullptr, k, j, i, d);
#pragma unroll
    for (int n = 0; n < N; ++n) du[g.at(n, k, j, i)] = d[n];
#pragma unroll
    for (int n = 0; n < N; ++n) fm[n] = fp[n];
    pfm = pfp;
  }
}

// All faces of a tile fused with the flux divergence (mixed form, 3D
// blocks): a workgroup owns a TJ x TK tile of columns and marches up a
// segment of levels.  Per level, the tile's plane with a 3-cell halo is
// staged in LDS (float, p - pref(i)); the x2 and x3 face states come from
// one WENO5/PLM evaluation per cell in LDS, the x2 / x3 Riemann fluxes go to
// LDS, and each lane takes its cell's two x2 and two x3 face fluxes from
// there.  The x1 faces are marched in registers as in k_mix_x1_div (a
// sliding window of T, p - psv with FP64 differences).  Every cell's du is
// written once: w is read from HBM about once, the face fluxes never leave
// the chip.
// The tile is FT_J x FT_K columns (32 x 8 or 16 x 16, both 256 lanes).

// LDS of one tile of the fused face kernel.
template <int NM, int FT_J, int FT_K>
struct FusedLds {
  static constexpr int H = 3, PJ = FT_J + 2 * H, PK = FT_K + 2 * H, N = 5 + NM;
  static constexpr int C2 = FT_J + 2, C3 = FT_K + 2;  // cells with face states
  // x2 items first (FT_K rows of C2 cells / FT_J + 1 faces), then x3 items
  // (C3 cells / FT_K + 1 faces by FT_J columns)
  static constexpr int NQ2 = FT_K * C2, NQ = NQ2 + C3 * FT_J;
  static constexpr int NF2 = FT_K * (FT_J + 1), NF = NF2 + (FT_K + 1) * FT_J;
  float plane[N][PK][PJ];
  float qm[N][NQ], qp[N][NQ];  // face states, velocities rotated
  float fx[N][NF];             // face fluxes, momenta rotated
};

// One tile: columns j0 .. j0+FT_J-1, k0 .. k0+FT_K-1 of block g, levels
// c0 .. c1-1.
template <int NM, int OH, int OV, int FT_J, int FT_K>
__device__ __forceinline__ void mix_fused_tile(
    const KGrid& g, const KThermo& th, const KSource& src, int rsolver, int j0,
    int k0, int c0, int c1, const double* __restrict__ w,
    const double* __restrict__ psf, const double* __restrict__ psv,
    const double* __restrict__ pref, double* __restrict__ du,
    FusedLds<NM, FT_J, FT_K>& L) {
  constexpr int FT_H = 3;
  constexpr int FT_PJ = FT_J + 2 * FT_H, FT_PK = FT_K + 2 * FT_H;
  constexpr int NT = FT_J * FT_K, NP = FT_PJ * FT_PK;
  constexpr int N = 5 + NM, IPR = 4 + NM, IVX = 1 + NM;
  constexpr int LO = (OV == 5) ? 0 : 1, HI = (OV == 5) ? 5 : 4;
  constexpr int C2 = FT_J + 2, C3 = FT_K + 2;  // cells with face states
  using LDS = FusedLds<NM, FT_J, FT_K>;
  constexpr int NQ2 = LDS::NQ2, NQ = LDS::NQ, NF2 = LDS::NF2, NF = LDS::NF;
  auto& plane = L.plane;
  const int decomp = g.grav != 0.0;
  const int t = threadIdx.x;
  const int tj = t % FT_J, tk = t / FT_J;
  const int j = j0 + tj, k = k0 + tk;
  const bool own = (j <= g.je) && (k <= g.ke);
  const double A1 = g.dx2 * g.dx3, A2 = g.dx1 * g.dx3, A3 = g.dx1 * g.dx2;
  const double wv = src.weight() / (g.dx1 * g.dx2 * g.dx3);
  const int jc = min(j, g.je), kc = min(k, g.ke);  // lanes past the block
  const long long pc = g.pcol(kc, jc) * g.ps_c;
  // column offsets, fixed over the march: the own column and the plane cells
  // this lane stages (the level adds i times the level stride)
  const long long colb = (long long)kc * g.c.k + (long long)jc * g.c.j;
  constexpr int NPL = (NP + NT - 1) / NT;
  long long pbase[NPL];
#pragma unroll
  for (int m = 0; m < NPL; ++m) {
    const int e = min(t + m * NT, NP - 1);
    const int ej = e % FT_PJ, ek = e / FT_PJ;
    const int gj = min(max(j0 - FT_H + ej, 0), g.nc2 - 1);
    const int gk = min(max(k0 - FT_H + ek, 0), g.nc3 - 1);
    pbase[m] = (long long)gk * g.c.k + (long long)gj * g.c.j;
  }
  // x1 window (own column)
  float q[N][6];
  float rdv[6];
  auto load1 = [&](int m, int i) {
    double wc[N];
#pragma unroll
    for (int n = 0; n < N; ++n) wc[n] = w[colb + n * g.c.n + (long long)i * g.c.i];
#pragma unroll
    for (int n = 0; n < N; ++n) q[n][m] = (float)wc[n];
    if (decomp)
      rdv[m] = (float)decomp32<NM>(g, th, w, colb, psv, pc, i, wc, q[0][m],
                                   q[IPR][m]);
  };
  auto face1 = [&](int f, float* fx) {
    float wl[N], wr[N];
#pragma unroll
    for (int n = 0; n < N; ++n) recon_pair_f<OV>(q[n], &wl[n], &wr[n]);
    float pb = 0.f;
    if (decomp) {
      pb = (float)psf[pc + f * g.ps_i];
      float gam = fsqrt(rdv[3] * rdv[2]);
      if (wr[IPR] + pb < 0.f) wr[IPR] = 0.f;
      if (wl[IPR] + pb < 0.f) wl[IPR] = 0.f;
      wr[0] = fdiv(pb + wr[IPR], wr[0] * gam);
      wl[0] = fdiv(pb + wl[IPR], wl[0] * gam);
    }
    face_check<NM, float>(g.status, wl, wr, pb);
    riemann_p<NM, 0>(th, rsolver, wl, wr, pb, fx);
  };
#pragma unroll
  for (int m = LO; m <= HI; ++m) load1(m, c0 - 3 + m);
  float fm[N];
  face1(c0, fm);
  double pfm = psf[pc + c0 * g.ps_i];

  auto recon = [&](float a, float b, float c, float d, float e, float* qm,
                   float* qp) {
    if (OH == 5) {
      weno5_both_f(a, b, c, d, e, qm, qp);
    } else {
      *qp = plm_up_f(b, c, d);
      *qm = plm_up_f(d, c, b);
    }
  };
  for (int i = c0; i < c1; ++i) {
    const double prd = pref[i];
    // the own column's next x1 window cell: its loads are issued here so
    // that they complete during the plane phases
    const int inew = i + 1 - 3 + HI;
    double wn[N], psvn = 0.0;
#pragma unroll
    for (int n = 0; n < N; ++n) wn[n] = w[colb + n * g.c.n + (long long)inew * g.c.i];
    if (decomp) psvn = psv[pc + inew * g.ps_i];
    const double pfp = psf[pc + (i + 1) * g.ps_i];
    // (1) the level-i plane of the tile with its halo
#pragma unroll
    for (int m = 0; m < NPL; ++m) {
      const int e = t + m * NT;
      if (e < NP) {
        const long long c = pbase[m] + (long long)i * g.c.i;
        float* pl = &plane[0][0][0] + e;
#pragma unroll
        for (int n = 0; n < N; ++n) {
          double v = w[c + n * g.c.n];
          if (n == IPR) v -= prd;
          pl[n * NP] = (float)v;
        }
      }
    }
    __syncthreads();
    // (2) face states: x2 for cells j0-1 .. j0+FT_J of the interior rows,
    //     x3 for cells k0-1 .. k0+FT_K of the interior columns, as two item
    //     loops with the direction's strides known at compile time; the x3
    //     items start one wavefront further on, so that the wavefront that
    //     took the x2 remainder is not also given the x3 remainder (no
    //     wavefront does more than one item beyond the even share).  The
    //     velocity slots are rotated so that the normal component comes
    //     first (x2: v2, v3, v1; x3: v3, v1, v2), the order in which the
    //     Riemann solver of that direction reads them: one code path serves
    //     both directions.
    auto states = [&](auto dirc, int e, int pj, int pk) {
      constexpr int D = decltype(dirc)::value;  // 1: x2, 2: x3
      constexpr int st = (D == 1) ? 1 : FT_PJ;
      const float* pc0 = &plane[0][pk][pj];
#pragma unroll
      for (int n = 0; n < N; ++n) {
        const int sv = (n >= IVX && n < IVX + 3) ? IVX + (n - IVX + D) % 3 : n;
        const float* pv = pc0 + sv * (FT_PK * FT_PJ);
        float qm, qp;
        recon(pv[-2 * st], pv[-st], pv[0], pv[st], pv[2 * st], &qm, &qp);
        L.qm[n][e] = qm;
        L.qp[n][e] = qp;
      }
    };
    for (int e = t; e < NQ2; e += NT)
      states(std::integral_constant<int, 1>(), e, (e % C2) + FT_H - 1, (e / C2) + FT_H);
    {
      constexpr int R2 = ((NQ2 % NT) + 63) / 64 * 64;
      for (int e3 = (t + NT - R2) % NT; e3 < NQ - NQ2; e3 += NT)
        states(std::integral_constant<int, 2>(), NQ2 + e3, (e3 % FT_J) + FT_H,
               (e3 / FT_J) + FT_H - 1);
    }
    __syncthreads();
    // (3) x2 / x3 Riemann fluxes of the tile's faces, one item list
    for (int e = t; e < NF; e += NT) {
      const bool x2 = e < NF2;
      const int e3 = e - NF2;
      const int lq = x2 ? (e / (FT_J + 1)) * C2 + e % (FT_J + 1) : NQ2 + e3;
      const int rq = lq + (x2 ? 1 : FT_J);
      float wl[N], wr[N], fx[N];
#pragma unroll
      for (int n = 0; n < N; ++n) {
        wl[n] = L.qp[n][lq];
        wr[n] = L.qm[n][rq];
      }
      face_check<NM, float>(g.status, wl, wr, (float)prd);
      riemann_p<NM, 0>(th, rsolver, wl, wr, (float)prd, fx);
#pragma unroll
      for (int n = 0; n < N; ++n) L.fx[n][e] = fx[n];
    }
    __syncthreads();
    // (4) x1 face i+1 in registers, the cell's divergence and sources
#pragma unroll
    for (int m = LO; m < HI; ++m) {
#pragma unroll
      for (int n = 0; n < N; ++n) q[n][m] = q[n][m + 1];
      rdv[m] = rdv[m + 1];
    }
#pragma unroll
    for (int n = 0; n < N; ++n) q[n][HI] = (float)wn[n];
    if (decomp) {
      if (decomp_cell(g, inew) != inew) {  // outflow ghost
        rdv[HI] = (float)decomp32<NM>(g, th, w, colb, psv, pc, inew, wn,
                                      q[0][HI], q[IPR][HI]);
      } else {
        double r = decomp_r<NM>(th, wn);
        rdv[HI] = (float)r;
        q[0][HI] = fdiv((float)wn[IPR], (float)wn[0] * (float)r);
        q[IPR][HI] = (float)(wn[IPR] - psvn);
      }
    }
    float fp[N];
    face1(i + 1, fp);
    if (own) {
      double d[N];
#pragma unroll
      for (int n = 0; n < N; ++n) {
        double dflx = A1 * ((double)fp[n] - (double)fm[n]);
        if (n == IVX && decomp) dflx += A1 * (pfp - pfm);
        // un-rotate: global component IVX + c sits in slot IVX + (c - rot) % 3
        const int s2 = (n >= IVX && n < IVX + 3) ? IVX + (n - IVX + 2) % 3 : n;
        const int s3 = (n >= IVX && n < IVX + 3) ? IVX + (n - IVX + 1) % 3 : n;
        const int f2 = tk * (FT_J + 1) + tj, f3 = NF2 + tk * FT_J + tj;
        dflx += A2 * ((double)L.fx[s2][f2 + 1] - (double)L.fx[s2][f2]);
        dflx += A3 * ((double)L.fx[s3][f3 + FT_J] - (double)L.fx[s3][f3]);
        d[n] = 0.0 - wv * dflx;
      }
      add_sources_cell<NM>(g, th, src, w, decomp ? psv : nullptr, k, j, i, d);
#pragma unroll
      for (int n = 0; n < N; ++n)
        __builtin_nontemporal_store(
            d[n], &du[colb + n * g.c.n + (long long)i * g.c.i]);
    }
#pragma unroll
    for (int n = 0; n < N; ++n) fm[n] = fp[n];
    pfm = pfp;
  }
}

// Tile of a batch: block b, first column (j0, k0), levels c0 .. c1-1.
struct KTile {
  int b, j0, k0, c0, c1;
};

// All faces and the divergence of the mixed form, over the tiles of every 3D
// block of a run in one launch: a workgroup owns a TJ x TK tile of columns
// and marches up a segment of levels (mix_fused_tile).
template <int NM, int OH, int OV, int FT_J, int FT_K>
__global__ void __launch_bounds__(FT_J * FT_K)
    k_mix_fused_batch(const KBlockPtr* __restrict__ bp,
                      const KTile* __restrict__ tiles, KThermo th, KSource src,
                      int rsolver) {
  __shared__ FusedLds<NM, FT_J, FT_K> L;
  const KTile tt = tiles[blockIdx.x];
  const int b = __builtin_amdgcn_readfirstlane(tt.b);
  if (b < 0) return;  // padding of the table
  const KBlockPtr& p = bp[b];
  mix_fused_tile<NM, OH, OV, FT_J, FT_K>(
      p.g, th, src, rsolver, __builtin_amdgcn_readfirstlane(tt.j0),
      __builtin_amdgcn_readfirstlane(tt.k0), __builtin_amdgcn_readfirstlane(tt.c0),
      __builtin_amdgcn_readfirstlane(tt.c1), gptr(p.w), gptr(p.psf), gptr(p.psv),
      gptr(p.pref), gptr(p.du), L);
}

// Face kernels of the mixed form.  variant 0: one lane per face in every
// direction (k_flux32_h, k_flux32_x1) and a separate divergence kernel
// (k_divergence32).  variants 1 and 2: x2 faces by cross-lane WENO
// (k_mix_x2), x3 faces marched over segments of `seg` faces (k_mix_x3);
// x1 faces then either marched up column segments of `seg` cells fused with
// the divergence (k_mix_x1_div, variant 1) or one lane per face followed by
// the divergence kernel (variant 2).  Variants 3 and 4 (every face and the
// divergence in one tile kernel of 32 x 8 or 16 x 16 columns) run over a
// run's tile table (launch_fused_batch); here, for 2D blocks, they act as
// variant 2.
template <int NM>
void launch_fluxes32(const KGrid& g, const KThermo& th, const cg_config& cfg,
                     const KSource& src, const double* w, float* ff,
                     double* psf, double* psv, const double* pref, double* du,
                     hipStream_t s, KTimer* tm, bool scan, int bs = 256,
                     int variant = 2, int seg = 16) {
  const long long ncol = (long long)g.nx2 * g.nx3;
  const long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
  const int decomp = g.grav != 0.0;
  const bool w5h = cfg.order_h == 5, w5v = cfg.order_v == 5;
  float* f1 = ff;
  float* f2 = f1 + (size_t)(5 + NM) * g.f[0].n;
  float* f3 = f2 + (size_t)(5 + NM) * g.f[1].n;
  if (decomp && scan)
    TIMED(tm, KID_SCAN,
          (k_hydrostatic_scan<NM><<<(unsigned)((ncol + 255) / 256), 256, 0, s>>>(
              g, th, w, psf, psv)));
  if (variant == 0) {
    const unsigned b1 = (unsigned)((ncol * (g.nx1 + 1) + bs - 1) / bs);
    if (w5v)
      TIMED(tm, KID_FLUX_X1, (k_flux32_x1<NM, 5><<<b1, bs, 0, s>>>(
                                 g, th, cfg.rsolver, decomp, w, psf, psv, f1)));
    else
      TIMED(tm, KID_FLUX_X1, (k_flux32_x1<NM, 2><<<b1, bs, 0, s>>>(
                                 g, th, cfg.rsolver, decomp, w, psf, psv, f1)));
    if (g.f2) {
      const unsigned b2 =
          (unsigned)(((long long)g.nx1 * (g.nx2 + 1) * g.nx3 + bs - 1) / bs);
      if (w5h)
        TIMED(tm, KID_FLUX_X2, (k_flux32_h<NM, 5, 1><<<b2, bs, 0, s>>>(
                                   g, th, cfg.rsolver, w, pref, f2)));
      else
        TIMED(tm, KID_FLUX_X2, (k_flux32_h<NM, 2, 1><<<b2, bs, 0, s>>>(
                                   g, th, cfg.rsolver, w, pref, f2)));
    }
    if (g.f3) {
      const unsigned b3 =
          (unsigned)(((long long)g.nx1 * g.nx2 * (g.nx3 + 1) + bs - 1) / bs);
      if (w5h)
        TIMED(tm, KID_FLUX_X3, (k_flux32_h<NM, 5, 2><<<b3, bs, 0, s>>>(
                                   g, th, cfg.rsolver, w, pref, f3)));
      else
        TIMED(tm, KID_FLUX_X3, (k_flux32_h<NM, 2, 2><<<b3, bs, 0, s>>>(
                                   g, th, cfg.rsolver, w, pref, f3)));
    }
    TIMED(tm, KID_DIVERGENCE,
          (k_divergence32<NM><<<(unsigned)((ncell + bs - 1) / bs), bs, 0, s>>>(
              g, th, src, f1, f2, f3, psf, psv, decomp, w, du)));
    return;
  }
  if (g.f2) {
    // one wavefront per 59 faces of an x2 row
    const long long nw = (long long)g.nx1 * g.nx3 * ((g.nx2 + 1 + 58) / 59);
    const unsigned b2 = (unsigned)((nw * 64 + bs - 1) / bs);
    if (w5h)
      TIMED(tm, KID_FLUX_X2, (k_mix_x2<NM, 5><<<b2, bs, 0, s>>>(
                                 g, th, cfg.rsolver, w, pref, f2)));
    else
      TIMED(tm, KID_FLUX_X2, (k_mix_x2<NM, 2><<<b2, bs, 0, s>>>(
                                 g, th, cfg.rsolver, w, pref, f2)));
  }
  if (g.f3) {
    const long long n3 = (long long)g.nx1 * g.nx2 * ((g.nx3 + 1 + seg - 1) / seg);
    const unsigned b3 = (unsigned)((n3 + bs - 1) / bs);
    if (w5h)
      TIMED(tm, KID_FLUX_X3, (k_mix_x3<NM, 5><<<b3, bs, 0, s>>>(
                                 g, th, cfg.rsolver, seg, w, pref, f3)));
    else
      TIMED(tm, KID_FLUX_X3, (k_mix_x3<NM, 2><<<b3, bs, 0, s>>>(
                                 g, th, cfg.rsolver, seg, w, pref, f3)));
  }
  if (variant >= 2) {
    const unsigned b1 = (unsigned)((ncol * (g.nx1 + 1) + bs - 1) / bs);
    if (w5v)
      TIMED(tm, KID_FLUX_X1, (k_flux32_x1<NM, 5><<<b1, bs, 0, s>>>(
                                 g, th, cfg.rsolver, decomp, w, psf, psv, f1)));
    else
      TIMED(tm, KID_FLUX_X1, (k_flux32_x1<NM, 2><<<b1, bs, 0, s>>>(
                                 g, th, cfg.rsolver, decomp, w, psf, psv, f1)));
    TIMED(tm, KID_DIVERGENCE,
          (k_divergence32<NM><<<(unsigned)((ncell + bs - 1) / bs), bs, 0, s>>>(
              g, th, src, f1, f2, f3, psf, psv, decomp, w, du)));
    return;
  }
  const long long nx1t = ncol * ((g.nx1 + seg - 1) / seg);
  const unsigned b1 = (unsigned)((nx1t + bs - 1) / bs);
  if (w5v)
    TIMED(tm, KID_FLUX_X1, (k_mix_x1_div<NM, 5><<<b1, bs, 0, s>>>(
                               g, th, src, cfg.rsolver, decomp, seg, w, psf, psv,
                               f2, f3, du)));
  else
    TIMED(tm, KID_FLUX_X1, (k_mix_x1_div<NM, 2><<<b1, bs, 0, s>>>(
                               g, th, src, cfg.rsolver, decomp, seg, w, psf, psv,
                               f2, f3, du)));
}

// The fused tile kernel over a batch's tile table (face variants 3 / 4).
template <int NM>
void launch_fused_batch(const KBlockPtr* bp, const KTile* tiles, int ntiles,
                        const KThermo& th, const cg_config& cfg,
                        const KSource& src, int variant, hipStream_t s,
                        KTimer* tm) {
  const bool w5h = cfg.order_h == 5, w5v = cfg.order_v == 5;
#define CG_FUSEDB(OH, OV, TJ, TK)                                                \
  TIMED(tm, KID_MIX_FUSED,                                                      \
        (k_mix_fused_batch<NM, OH, OV, TJ, TK><<<(unsigned)ntiles, 256, 0, s>>>( \
            bp, tiles, th, src, cfg.rsolver)))
#define CG_FUSEDB_T(TJ, TK)                      \
  do {                                           \
    if (w5h && w5v) CG_FUSEDB(5, 5, TJ, TK);     \
    else if (w5h) CG_FUSEDB(5, 2, TJ, TK);       \
    else if (w5v) CG_FUSEDB(2, 5, TJ, TK);       \
    else CG_FUSEDB(2, 2, TJ, TK);                \
  } while (0)
  if (variant == 3)
    CG_FUSEDB_T(32, 8);
  else
    CG_FUSEDB_T(16, 16);
#undef CG_FUSEDB_T
#undef CG_FUSEDB
}

// per-level reference pressure: horizontal mean of p over interior columns
template <int NM>
__global__ void __launch_bounds__(256)
    k_level_mean_p(KGrid g, const double* __restrict__ w,
                   double* __restrict__ pref) {
  constexpr int IPR = 4 + NM;
  __shared__ double sh[256];
  int i = blockIdx.x;  // 0 .. nc1-1
  double s = 0.0;
  long long ncol = (long long)g.nx2 * g.nx3;
  for (long long c = threadIdx.x; c < ncol; c += blockDim.x) {
    int j = g.js + (int)(c % g.nx2), k = g.ks + (int)(c / g.nx2);
    s += w[g.at(IPR, k, j, i)];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) pref[i] = sh[0] / (double)ncol;
}

}  // namespace cg
