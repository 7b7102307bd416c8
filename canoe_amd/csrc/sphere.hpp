// sphere.hpp -- device kernels of the cubed-sphere run (config 3): six
// gnomonic-equiangle panels (src/exo3/gnomonic_equiangle.cpp), each a block
// of the run in the level-major layout, dry gas, FP64 state.
//
// Per stage: x1 fluxes (k_sph_flux_x1: decomposition, WENO5, HLLC) with the
// covariant projection of the tangential momentum fluxes
// (hllc_transform.cpp:235-272); the panel-edge exchange of face states
// (k_sph_edge_states, cubed_sphere.cpp:250-565: the neighbours' inner states
// at the edge faces); x2 / x3 face states and HLLC in the face's
// orthonormal frame (k_sph_face_flux: PrimToLocal2/3, FluxToGlobal2/3 of
// gnomonic_equiangle.cpp:559-828, then the covariant projection); flux
// divergence with the panel's areas and volumes, the geometric source terms
// (gnomonic_equiangle.cpp:831-877), gravity and the Held-Suarez / Coriolis
// forcing of examples/2023-Chen-exo3/hs94.cpp:56-151 (k_sph_divergence,
// sph_cell_du); the x1 column solve with the panel's areas
// (k_sph_implicit_fwd, then k_sph_bwd_update, which also applies the stage,
// the x1 walls and the scan); the stage update with ConservedToPrimitive
// for covariant momenta (k_sph_update, patches/26.adiabatic_hydro.patch:
// 21-46); the ghost zones of x2 / x3 interpolated from the neighbour panels
// (k_sph_ghosts, cubed_sphere_utility.cpp:80-400); x1 walls and the
// hydrostatic scan.  The FP32-flux form replaces the face kernels and the
// divergence with k_sph_flux32_x1m (x1 faces in float, marched up column
// segments) and k_sph_faces32 (x2 / x3 faces of column tiles through LDS,
// fused with the divergence); its state, sources and column solve stay
// FP64.  The FP64 form's arithmetic is written in the order of the oracle's
// restatement so that the two agree to rounding.
#pragma once
#include "cg_physics.hpp"
#include "flux32.hpp"
#include "kernels.hpp"
#include "sphere_geom.hpp"

namespace cg {

using cg_host::SphEdge;
using cg_host::SphGhost;

// Arguments of the sphere kernels; arrays are indexed by the panel (0-5),
// which is wavefront uniform (blockIdx.y) except in the table kernels,
// which read the panel pointers from ptr (device memory).
struct KSph {
  KGrid g;  // level-major grid of a panel (the same for all six)
  const double* kj[6];
  double* u[6];
  double* w[6];
  double* u1[6];
  double* du[6];
  // face fluxes: doubles, or floats in the FP32-flux form (a run keeps one
  // precision, so each buffer is only ever read as the type it was written)
  double* f1[6];
  double* f2[6];
  double* f3[6];
  double* psf[6];
  double* psv[6];
  double* lr;  // edge buffer: neighbours' inner states at the panel-edge faces
  long long lrn;
  const double* x1f;
  const double* x1v;
  int n, nc, nr;
  double dx1, gm1;
  double omega, ts, dt_eq, dtheta, p0, kf, ka, ks, sigmab, cp, rd, kappa;
  int hs94;
};

struct KSphPtrs {
  double* u[6];
  double* w[6];
  const double* kj[6];
};

__device__ __forceinline__ double skj(const KSph& s, int p, int a, int k, int j) {
  return s.kj[p][(long long)a * s.nc * s.nc + (long long)k * s.nc + j];
}

// ConservedToPrimitive / PrimitiveToConserved with covariant momenta
CG_HD void sph_cons_to_prim(double gm1, double cth, const double* u, double* w) {
  double di = 1.0 / u[0];
  w[0] = u[0];
  w[1] = u[1] * di;
  double v = u[2] * di, x = u[3] * di;
  double sth2 = 1. - cth * cth;
  w[2] = v / sth2 - x * cth / sth2;
  w[3] = -v * cth / sth2 + x / sth2;
  double e_k = 0.5 * (u[1] * w[1] + u[2] * w[2] + u[3] * w[3]);
  w[4] = gm1 * (u[4] - e_k);
}

CG_HD void sph_prim_to_cons(double gm1, double cth, const double* w, double* u) {
  double igm1 = 1.0 / gm1;
  u[1] = w[1] * w[0];
  double m2 = w[2] * w[0], m3 = w[3] * w[0];
  u[2] = m2 + m3 * cth;
  u[3] = m3 + m2 * cth;
  u[4] = w[4] * igm1 + 0.5 * (u[1] * w[1] + u[2] * w[2] + u[3] * w[3]);
  u[0] = w[0];
}

// interior cell (k, j, i) of item t, panel blockIdx.y
__device__ __forceinline__ bool sph_cell(const KGrid& g, long long t, int& i, int& j, int& k) {
  if (t >= (long long)g.nx1 * g.nx2 * g.nx3) return false;
  g.decode(t, g.nx1, g.nx2, g.nx3, i, j, k);
  i += g.is;
  j += g.js;
  k += g.ks;
  return true;
}

__global__ void __launch_bounds__(256) k_sph_prim_to_cons(KSph s, int p) {
  int i, j, k;
  if (!sph_cell(s.g, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const KGrid& g = s.g;
  const long long c0 = g.at(0, k, j, i);
  double w[5], u[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) w[v] = s.w[p][c0 + v * g.c.n];
  sph_prim_to_cons(s.gm1, skj(s, p, cg_host::SKJ_COSC, k, j), w, u);
#pragma unroll
  for (int v = 0; v < 5; ++v) s.u[p][c0 + v * g.c.n] = u[v];
}

// x1 face fluxes of the six panels (decomposition, WENO5, HLLC) with the
// covariant projection of the tangential momentum fluxes
__global__ void __launch_bounds__(256) k_sph_flux_x1(KSph s, KThermo th) {
  const int p = blockIdx.y;
  const KGrid& g = s.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)(g.nx1 + 1) * g.nx2 * g.nx3) return;
  flux_x1_face<0, 5>(g, th, CG_RS_HLLC, g.grav != 0.0 ? 1 : 0, s.w[p], s.psf[p], s.psv[p],
                     s.f1[p], t, s.kj[p] + (long long)cg_host::SKJ_HC * s.nc * s.nc, s.nc);
}

// The same x1 faces in float (the FP32-flux form): decomposition in FP64,
// WENO5 and HLLC in float about the face's hydrostatic pressure psf, which
// the normal momentum flux leaves out (k_sph_divergence adds the face
// pressures times the face areas in FP64), then the covariant projection;
// stored as float.  Marched up column segments of `seg` faces (one lane per
// column segment, lanes along j): each cell's float window entry and
// decomposition are formed once per segment instead of once for every face
// whose stencil holds it (the faces equal the per-face flux32_x1_face's).
__global__ void __launch_bounds__(256) k_sph_flux32_x1m(KSph s, KThermo th, int seg) {
  const int p = blockIdx.y;
  const KGrid& g = s.g;
  const long long ncol = (long long)g.nx2 * g.nx3;
  const int nseg = (g.nx1 + 1 + seg - 1) / seg;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * nseg) return;
  const long long col = t % ncol;
  const int sg = (int)(t / ncol);
  const int j = g.js + (int)(col % g.nx2), k = g.ks + (int)(col / g.nx2);
  const int f0 = g.is + sg * seg;
  const int f1 = min(f0 + seg, g.ie + 2);  // faces f0 .. f1-1
  const int decomp = g.grav != 0.0 ? 1 : 0;
  const double* w = s.w[p];
  const double* psv = s.psv[p];
  const long long pc = g.pcol(k, j) * g.ps_c;
  const float hc = (float)skj(s, p, cg_host::SKJ_HC, k, j);
  float* F1 = reinterpret_cast<float*>(s.f1[p]);
  // window: cells f-3 .. f+2 of face f
  float q[5][6];
  double rdv[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) x1_window_cell<0>(g, th, decomp, w, psv, pc, k, j, f0 - 3 + m, q, rdv, m);
  for (int f = f0; f < f1; ++f) {
    float fx[5];
    x1_face_from_window<0, 5>(g, th, CG_RS_HLLC, decomp, s.psf[p], pc, f, q, rdv, fx);
    const float ty = fx[2], tz = fx[3];
    fx[2] = ty + tz * hc;
    fx[3] = tz + ty * hc;
#pragma unroll
    for (int v = 0; v < 5; ++v) F1[g.f[0].at(v, k, j, f)] = fx[v];
    if (f + 1 < f1) {
#pragma unroll
      for (int m = 0; m < 5; ++m) {
#pragma unroll
        for (int v = 0; v < 5; ++v) q[v][m] = q[v][m + 1];
        rdv[m] = rdv[m + 1];
      }
      x1_window_cell<0>(g, th, decomp, w, psv, pc, k, j, f + 3, q, rdv, 5);
    }
  }
}

// hydrostatic face and cell pressures of the six panels' columns
__global__ void __launch_bounds__(256) k_sph_scan(KSph s, KThermo th) {
  const int p = blockIdx.y;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)s.g.nx2 * s.g.nx3) return;
  scan_column<0>(s.g, th, s.w[p], s.psf[p], s.psv[p], t);
}

// reflecting x1 walls of the six panels' columns
__global__ void __launch_bounds__(256) k_sph_bc_x1(KSph s) {
  const int p = blockIdx.y;
  const KGrid& g = s.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx2 * g.nx3 * NG) return;
  int gh, j, k;
  g.decode(t, NG, g.nx2, g.nx3, gh, j, k);
  bc_pair(s.w[p], g.c.n, 5, 1, g.bc[0], g.bc[1], g.at(0, k + g.ks, j + g.js, g.is),
          g.at(0, k + g.ks, j + g.js, g.ie), g.c.i, gh + 1);
}

// One face state of panel p at level ii: x2 (DIR = 1) face f between cells
// j = js + f - 1 and js + f of row k = ks + t, or x3 (DIR = 2) face f between
// cells k = ks + f - 1 and ks + f of column j = js + t; RIGHT false: the left
// state (upper face of the cell below, patches/19.calculate_fluxes.patch
// wl_), true: the right state (lower face of the cell above, wr_).  ORDER 5
// (WENO5) or 2 (PLM), in the form's arithmetic type R.
template <int ORDER, int DIR, bool RIGHT, class R>
__device__ __forceinline__ void sph_face_state(const KSph& s, int p, int f, int t, int ii,
                                               R* out) {
  const KGrid& g = s.g;
  const int i = g.is + ii;
  const int j = (DIR == 1) ? g.js + f : g.js + t;
  const int k = (DIR == 1) ? g.ks + t : g.ks + f;
  const long long st = (DIR == 1) ? g.c.j : g.c.k;
  const long long base = g.at(0, k, j, i);
  const double* w = s.w[p];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double* c = w + base + v * g.c.n;
    if constexpr (std::is_same<R, float>::value) {
      // the FP32 form: only the needed side, with the fast reciprocals of
      // the box's FP32 faces
      const float a = (float)c[-2 * st], b = (float)c[-st], m = (float)c[0],
                  d = (float)c[st];
      if (ORDER == 5)
        out[v] = RIGHT ? weno5_up_f((float)c[2 * st], d, m, b, a)
                       : weno5_up_f((float)c[-3 * st], a, b, m, d);
      else
        out[v] = RIGHT ? plm_up_f(d, m, b) : plm_up_f(a, b, m);
      continue;
    }
    R qm, qp;
    if (ORDER == 5) {
      if (!RIGHT) {
        weno5_cell<R>((R)c[-3 * st], (R)c[-2 * st], (R)c[-st], (R)c[0], (R)c[st], &qm, &qp);
        out[v] = qp;
      } else {
        weno5_cell<R>((R)c[-2 * st], (R)c[-st], (R)c[0], (R)c[st], (R)c[2 * st], &qm, &qp);
        out[v] = qm;
      }
    } else {
      if (!RIGHT) {
        plm_cell<R>((R)c[-2 * st], (R)c[-st], (R)c[0], &qm, &qp);
        out[v] = qp;
      } else {
        plm_cell<R>((R)c[-st], (R)c[0], (R)c[st], &qm, &qp);
        out[v] = qm;
      }
    }
  }
}

// edge buffer entry of panel-edge item e (= (p 4 + side) n + t), variable v,
// level ii: the neighbour's inner state at that face in this panel's frame
__device__ __forceinline__ long long sph_eb(const KSph& s, long long e, int v, int ii) {
  return (e * 5 + v) * s.nr + ii;
}

// Panel-edge exchange (SaveLR3DValues / SynchronizeFluxes*, cubed_sphere.cpp:
// 250-565): for every panel-edge face, the neighbour q's inner state there
// (its own reconstruction next to the shared edge), its contravariant
// velocity carried into this panel's frame through (U, V) at the face, into
// the edge buffer.  Only the edge faces are reconstructed here; every other
// face state is formed where its flux is.
template <int ORDER, class R = double>
__global__ void __launch_bounds__(256)
    k_sph_edge_states(KSph s, const SphEdge* __restrict__ edges, int nedge) {
  const long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= (long long)nedge * s.nr) return;
  const long long ei = item / s.nr;
  const SphEdge e = edges[ei];
  const int ii = (int)(item % s.nr);
  const int n = s.n;
  // the inner state of q at its side sq: the right state of its first face
  // on a lower side, the left state of its last face on an upper side
  R q[5];
  switch (e.sq) {
    case 0: sph_face_state<ORDER, 1, true, R>(s, e.q, 0, e.tq, ii, q); break;
    case 1: sph_face_state<ORDER, 1, false, R>(s, e.q, n, e.tq, ii, q); break;
    case 2: sph_face_state<ORDER, 2, true, R>(s, e.q, 0, e.tq, ii, q); break;
    default: sph_face_state<ORDER, 2, false, R>(s, e.q, n, e.tq, ii, q); break;
  }
  const R U = (R)e.r00 * q[2] + (R)e.r01 * q[3];
  const R V = (R)e.r10 * q[2] + (R)e.r11 * q[3];
  q[2] = (R)e.a00 * U + (R)e.a01 * V;
  q[3] = (R)e.a10 * U + (R)e.a11 * V;
  R* eb = reinterpret_cast<R*>(s.lr);
#pragma unroll
  for (int v = 0; v < 5; ++v) eb[sph_eb(s, ei, v, ii)] = q[v];
}

// The x2 (DIR = 1) / x3 (DIR = 2) face f of panel p (row / column t, level
// ii) from its left and right states: HLLC in the face's orthonormal frame
// (PrimToLocal2/3, FluxToGlobal2/3 of gnomonic_equiangle.cpp:559-828) and the
// covariant projection of the tangential momentum fluxes
// (hllc_transform.cpp:235-272), stored in f2 / f3.
template <int DIR, class R>
__device__ __forceinline__ void sph_face_riemann(const KSph& s, const KThermo& th, int p,
                                                 int f, int t, int ii, R* a, R* b,
                                                 R* out = nullptr) {
  const KGrid& g = s.g;
  const int i = g.is + ii;
  const int j = (DIR == 1) ? g.js + f : g.js + t;
  const int k = (DIR == 1) ? g.ks + t : g.ks + f;
  // the face metric, the frame transforms and their products in the form's
  // arithmetic type (the metric itself is FP64 on the host)
  const R cf = (R)skj(s, p, DIR == 1 ? cg_host::SKJ_COSF2 : cg_host::SKJ_COSF3, k, j);
  const R sf = (R)skj(s, p, DIR == 1 ? cg_host::SKJ_SINF2 : cg_host::SKJ_SINF3, k, j);
  const R hc = (R)skj(s, p, DIR == 1 ? cg_host::SKJ_HF2 : cg_host::SKJ_HF3, k, j);
  const R one = (R)1.0;
  // gi = 1 / sin^2 of the face angle, rg its square root, irg 1 / rg (the
  // FP32 form with the fast reciprocal and square root)
  R gi, rg, irg;
  if constexpr (std::is_same<R, float>::value) {
    gi = frcp(sf * sf);
    rg = fsqrt(gi);
    irg = frcp(rg);
  } else {
    gi = one / (sf * sf);
    rg = sqrt(gi);
    irg = one / sqrt(gi);
  }
  if (DIR == 1) {  // PrimToLocal2
    const R T22 = irg, T32 = cf / sqrt(one), T33 = sqrt(one);
    R uu2 = a[2], uu3 = a[3];
    a[2] = T22 * uu2;
    a[3] = T32 * uu2 + T33 * uu3;
    uu2 = b[2];
    uu3 = b[3];
    b[2] = T22 * uu2;
    b[3] = T32 * uu2 + T33 * uu3;
  } else {  // PrimToLocal3
    const R T22 = sqrt(one), T23 = cf / sqrt(one), T33 = irg;
    R uu2 = a[2], uu3 = a[3];
    a[2] = T22 * uu2 + T23 * uu3;
    a[3] = T33 * uu3;
    uu2 = b[2];
    uu3 = b[3];
    b[2] = T22 * uu2 + T23 * uu3;
    b[3] = T33 * uu3;
  }
  face_check<0, R>(g.status, a, b, R(0));
  R fx[5];
  if constexpr (std::is_same<R, float>::value)
    riemann_p<0, DIR>(th, CG_RS_HLLC, a, b, 0.f, fx);  // the FP32 faces' HLLC
  else
    riemann<0, DIR, R>(th, CG_RS_HLLC, a, b, fx);
  const R txx = fx[1], txy = fx[2], txz = fx[3];
  if (DIR == 1) {  // FluxToGlobal2
    const R T22 = rg, T32 = -rg * cf / one, T33 = one / sqrt(one);
    fx[1] = one * txx;
    fx[2] = T22 * txy;
    fx[3] = T32 * txy + T33 * txz;
  } else {  // FluxToGlobal3
    const R T22 = one / sqrt(one), T23 = -cf / one * rg, T33 = rg;
    fx[1] = one * txx;
    fx[2] = T22 * txy + T23 * txz;
    fx[3] = T33 * txz;
  }
  const R ty = fx[2], tz = fx[3];
  fx[2] = ty + tz * hc;
  fx[3] = tz + ty * hc;
  if (out) {  // to the caller (k_sph_faces32 with the divergence fused)
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = fx[v];
    return;
  }
  // stored in the form's type (the FP32 form's fluxes take half the bytes)
  R* F = reinterpret_cast<R*>((DIR == 1) ? s.f2[p] : s.f3[p]);
#pragma unroll
  for (int v = 0; v < 5; ++v) F[g.f[DIR].at(v, k, j, i)] = fx[v];
}

// x2 (DIR = 1) / x3 (DIR = 2) faces of every panel, one lane per face: the
// two face states reconstructed here (at a panel edge the outer one from the
// edge buffer), then sph_face_riemann
template <int ORDER, int DIR, class R = double>
__global__ void __launch_bounds__(256) k_sph_face_flux(KSph s, KThermo th) {
  const int p = blockIdx.y;
  const int n = s.n;
  const long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= (long long)s.nr * n * (n + 1)) return;
  int t, f, ii;
  if (DIR == 1) {
    f = (int)(item % (n + 1));
    const long long r = item / (n + 1);
    t = (int)(r % n);
    ii = (int)(r / n);
  } else {
    t = (int)(item % n);
    const long long r = item / n;
    f = (int)(r % (n + 1));
    ii = (int)(r / (n + 1));
  }
  const R* eb = reinterpret_cast<const R*>(s.lr);
  R a[5], b[5];
  if (f == 0) {  // lower panel edge: the outer (left) state is the neighbour's
    const long long e = ((long long)p * 4 + (DIR == 1 ? 0 : 2)) * n + t;
#pragma unroll
    for (int v = 0; v < 5; ++v) a[v] = eb[sph_eb(s, e, v, ii)];
  } else {
    sph_face_state<ORDER, DIR, false, R>(s, p, f, t, ii, a);
  }
  if (f == n) {  // upper panel edge: the outer (right) state is the neighbour's
    const long long e = ((long long)p * 4 + (DIR == 1 ? 1 : 3)) * n + t;
#pragma unroll
    for (int v = 0; v < 5; ++v) b[v] = eb[sph_eb(s, e, v, ii)];
  } else {
    sph_face_state<ORDER, DIR, true, R>(s, p, f, t, ii, b);
  }
  sph_face_riemann<DIR, R>(s, th, p, f, t, ii, a, b);
}

// areas and volume of a panel cell (gnomonic_equiangle.cpp:221-260, 407-416)
__device__ __forceinline__ double sph_a1(const KSph& s, int p, int k, int j, int i) {
  return skj(s, p, cg_host::SKJ_D2A, k, j) * skj(s, p, cg_host::SKJ_D3A, k, j) * s.x1f[i] *
         s.x1f[i] * skj(s, p, cg_host::SKJ_SINC, k, j);
}
__device__ __forceinline__ double sph_a2(const KSph& s, int p, int k, int j, int i) {
  return s.x1v[i] * skj(s, p, cg_host::SKJ_D3AF2, k, j) * s.dx1;
}
__device__ __forceinline__ double sph_a3(const KSph& s, int p, int k, int j, int i) {
  return s.x1v[i] * skj(s, p, cg_host::SKJ_D2AF3, k, j) * s.dx1;
}
__device__ __forceinline__ double sph_v(const KSph& s, int p, int k, int j, int i) {
  return s.dx1 * skj(s, p, cg_host::SKJ_D2A, k, j) * skj(s, p, cg_host::SKJ_D3A, k, j) *
         s.x1v[i] * s.x1v[i] * skj(s, p, cg_host::SKJ_SINC, k, j);
}

// du of cell (k, j, i) of panel p = -wght div F (areas, volume) + geometric
// terms + gravity + forcing, from its six face fluxes (f1l / f1r the lower /
// upper x1 face, then x2, x3); psplit: the x1 fluxes leave the face
// pressures psf out of the normal momentum flux (k_sph_flux32_x1m), added
// here in FP64
__device__ __forceinline__ void sph_cell_du(const KSph& s, int p, int k, int j, int i,
                                            double wght, const double* f1l,
                                            const double* f1r, const double* f2l,
                                            const double* f2r, const double* f3l,
                                            const double* f3r, int psplit) {
  const KGrid& g = s.g;
  const double a1l = sph_a1(s, p, k, j, i), a1r = sph_a1(s, p, k, j, i + 1);
  const double a2l = sph_a2(s, p, k, j, i), a2r = sph_a2(s, p, k, j + 1, i);
  const double a3l = sph_a3(s, p, k, j, i), a3r = sph_a3(s, p, k + 1, j, i);
  const double vol = sph_v(s, p, k, j, i);
  double d[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double dflx = a1r * f1r[v] - a1l * f1l[v];
    if (v == 1 && psplit) {
      const double* pf = s.psf[p] + g.pcol(k, j) * g.ps_c;
      dflx += a1r * pf[(i + 1) * g.ps_i] - a1l * pf[i * g.ps_i];
    }
    dflx += a2r * f2r[v] - a2l * f2l[v];
    dflx += a3r * f3r[v] - a3l * f3l[v];
    d[v] = 0.0 - wght * dflx / vol;
  }
  const long long c0 = g.at(0, k, j, i);
  const double* w = s.w[p];
  const double rho = w[c0], v1 = w[c0 + g.c.n], v2 = w[c0 + 2 * g.c.n],
               v3 = w[c0 + 3 * g.c.n], pr = w[c0 + 4 * g.c.n];
  // AddCoordTermsDivergence
  {
    const double cosine = skj(s, p, cg_host::SKJ_COSC, k, j);
    const double sinc = skj(s, p, cg_host::SKJ_SINC, k, j);
    const double sine2 = sinc * sinc;
    const double radius = s.x1v[i];
    const double v_2 = v2 + v3 * cosine;
    const double v_3 = v3 + v2 * cosine;
    const double src1 = (2.0 * pr + rho * (v2 * v_2 + v3 * v_3)) / radius;
    d[1] += wght * src1;
    const double xr = (a2l * skj(s, p, cg_host::SKJ_SINF2, k, j) -
                       a2r * skj(s, p, cg_host::SKJ_SINF2, k, j + 1)) /
                      vol;
    const double yr = (a3l * skj(s, p, cg_host::SKJ_SINF3, k, j) -
                       a3r * skj(s, p, cg_host::SKJ_SINF3, k + 1, j)) /
                      vol;
    const double src2 = -xr * (pr + rho * v3 * v3 * sine2) - rho * v1 * v_2 / radius;
    d[2] += wght * src2;
    const double src3 = -yr * (pr + rho * v2 * v2 * sine2) - rho * v1 * v_3 / radius;
    d[3] += wght * src3;
  }
  // gravity (ConstantAcceleration)
  if (g.grav != 0.0) {
    const double src = wght * rho * g.grav;
    d[1] += src;
    d[4] += src * v1;
  }
  const double cth = skj(s, p, cg_host::SKJ_COSC, k, j);
  const double sl = skj(s, p, cg_host::SKJ_SINLAT, k, j);
  const double cl = skj(s, p, cg_host::SKJ_COSLAT, k, j);
  if (s.omega != 0.) {  // Coriolis in (U, V) (hs94.cpp:77-96)
    const double f = 2. * s.omega * sl;
    double U = skj(s, p, cg_host::SKJ_R00, k, j) * v2 + skj(s, p, cg_host::SKJ_R01, k, j) * v3;
    const double V =
        skj(s, p, cg_host::SKJ_R10, k, j) * v2 + skj(s, p, cg_host::SKJ_R11, k, j) * v3;
    U = -U;
    const double m2 = rho * U, m3 = rho * V;
    const double accU = f * m3, accV = -f * m2;
    const double a2 = skj(s, p, cg_host::SKJ_A00, k, j) * (-accU) +
                      skj(s, p, cg_host::SKJ_A01, k, j) * accV;
    const double a3 = skj(s, p, cg_host::SKJ_A10, k, j) * (-accU) +
                      skj(s, p, cg_host::SKJ_A11, k, j) * accV;
    const double b2 = a2 + a3 * cth, b3 = a3 + a2 * cth;
    d[2] += wght * b2;
    d[3] += wght * b3;
  }
  if (s.hs94) {  // Newtonian relaxation and Rayleigh drag (hs94.cpp:103-150)
    const double scaled_z = pr / s.p0;
    const double sigma = pr / s.p0;
    double sigma_p = (sigma - s.sigmab) / (1. - s.sigmab);
    const double Kv = (sigma_p < 0.0) ? 0.0 : sigma_p * s.kf;
    double Teq_p = s.ts - s.dt_eq * (sl * sl) - s.dtheta * log_d(scaled_z) * (cl * cl);
    Teq_p *= exp_d(s.kappa * log_d(scaled_z));
    const double Teq = (Teq_p > 200.) ? Teq_p : 200.;
    const double temp = pr / rho / s.rd;
    sigma_p = (sigma_p < 0.0) ? 0.0 : sigma_p * (cl * cl * cl * cl);
    const double Kt = s.ka + (s.ks - s.ka) * sigma_p;
    const double m1 = rho * v1, m2 = rho * v2, m3 = rho * v3;
    const double n2 = m2 + m3 * cth, n3 = m3 + m2 * cth;
    d[1] += -wght * Kv * m1;
    d[2] += -wght * Kv * n2;
    d[3] += -wght * Kv * n3;
    d[4] += -wght * (s.cp - s.rd) * rho * Kt * (temp - Teq);
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) s.du[p][c0 + v * g.c.n] = d[v];
}

// sph_cell_du of every cell, the face fluxes read from f1 / f2 / f3, stored
// as FT (double, or float in the FP32-flux form)
template <class FT>
__global__ void __launch_bounds__(256)
    k_sph_divergence(KSph s, double beta, const double* __restrict__ dtp, int psplit) {
  const int p = blockIdx.y;
  int i, j, k;
  if (!sph_cell(s.g, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const KGrid& g = s.g;
  const FT* F1 = reinterpret_cast<const FT*>(s.f1[p]);
  const FT* F2 = reinterpret_cast<const FT*>(s.f2[p]);
  const FT* F3 = reinterpret_cast<const FT*>(s.f3[p]);
  double f1l[5], f1r[5], f2l[5], f2r[5], f3l[5], f3r[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    f1l[v] = (double)F1[g.f[0].at(v, k, j, i)];
    f1r[v] = (double)F1[g.f[0].at(v, k, j, i + 1)];
    f2l[v] = (double)F2[g.f[1].at(v, k, j, i)];
    f2r[v] = (double)F2[g.f[1].at(v, k, j + 1, i)];
    f3l[v] = (double)F3[g.f[2].at(v, k, j, i)];
    f3r[v] = (double)F3[g.f[2].at(v, k + 1, j, i)];
  }
  sph_cell_du(s, p, k, j, i, beta * dtp[0], f1l, f1r, f2l, f2r, f3l, f3r, psplit);
}

// The FP32 form's x2 and x3 faces of a tile of SF_J x SF_K columns of one
// panel at one level: the tile's plane with a 3-cell halo is staged in LDS
// as floats, each cell's two face values per direction come from one WENO5
// (or PLM) evaluation (weno5_both_f: the smoothness indicators are shared),
// and each face's HLLC in its orthonormal frame runs from LDS
// (sph_face_riemann); at a panel edge the outer state is the neighbour's
// from the edge buffer.  DIV false: the fluxes are stored in f2 / f3, a face
// on the boundary between two tiles by the tile it opens (the panel's last
// face by the last tile).  DIV true: each tile keeps all its faces' fluxes
// in LDS and each lane finishes its cell with sph_cell_du (the x1 fluxes
// from f1, stage weight beta dtp[0]): the x2 / x3 fluxes never leave the
// chip.
constexpr int SF_J = 32, SF_K = 8;
template <int ORDER, bool DIV>
__global__ void __launch_bounds__(SF_J * SF_K)
    k_sph_faces32(KSph s, KThermo th, double beta, const double* __restrict__ dtp,
                  int psplit) {
  constexpr int H = 3, PJ = SF_J + 2 * H, PK = SF_K + 2 * H, NT = SF_J * SF_K;
  constexpr int C2 = SF_J + 2, C3 = SF_K + 2;  // cells with face values
  constexpr int NQ2 = SF_K * C2, NQ3 = C3 * SF_J;
  constexpr int NF2 = SF_K * (SF_J + 1), NF = NF2 + (SF_K + 1) * SF_J;
  __shared__ float plane[5][PK][PJ];
  __shared__ float ql[5][NF], qr[5][NF];  // left / right state of each face
  __shared__ float fxs[DIV ? 5 : 1][DIV ? NF : 1];  // face fluxes (DIV)
  const int p = blockIdx.y;
  const KGrid& g = s.g;
  const int n = s.n;
  const int ntj = (n + SF_J - 1) / SF_J, ntk = (n + SF_K - 1) / SF_K;
  const int tj = blockIdx.x % ntj, tk = (blockIdx.x / ntj) % ntk;
  const int ii = blockIdx.x / (ntj * ntk);
  const int j0 = tj * SF_J, k0 = tk * SF_K;  // first face / cell of the tile
  const int i = g.is + ii;
  const int t = threadIdx.x;
  const double* w = s.w[p];
  // (1) the plane of cells j0-3 .. j0+SF_J+2, k0-3 .. k0+SF_K+2
  for (int e = t; e < PJ * PK; e += NT) {
    const int ej = e % PJ, ek = e / PJ;
    const int gj = min(g.js + j0 - H + ej, g.nc2 - 1);
    const int gk = min(g.ks + k0 - H + ek, g.nc3 - 1);
#pragma unroll
    for (int v = 0; v < 5; ++v) plane[v][ek][ej] = (float)w[g.at(v, gk, gj, i)];
  }
  __syncthreads();
  // (2) face values: cell c of a row (column) puts its upper value into the
  // left state of tile face c and its lower value into the right state of
  // tile face c - 1
  auto values = [&](int st, int pk, int pj, int fl, int fr) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const float* c = &plane[v][pk][pj];
      float qm, qp;
      if (ORDER == 5) {
        weno5_both_f(c[-2 * st], c[-st], c[0], c[st], c[2 * st], &qm, &qp);
      } else {
        qp = plm_up_f(c[-st], c[0], c[st]);
        qm = plm_up_f(c[st], c[0], c[-st]);
      }
      if (fl >= 0) ql[v][fl] = qp;
      if (fr >= 0) qr[v][fr] = qm;
    }
  };
  for (int e = t; e < NQ2; e += NT) {  // x2: rows k0 .. k0+SF_K-1
    const int c = e % C2, r = e / C2, f0 = r * (SF_J + 1);
    values(1, r + H, c + H - 1, c <= SF_J ? f0 + c : -1, c >= 1 ? f0 + c - 1 : -1);
  }
  for (int e = t; e < NQ3; e += NT) {  // x3: columns j0 .. j0+SF_J-1
    const int q = e % SF_J, c = e / SF_J;
    values(PJ, c + H - 1, q + H, c <= SF_K ? NF2 + c * SF_J + q : -1,
           c >= 1 ? NF2 + (c - 1) * SF_J + q : -1);
  }
  __syncthreads();
  // (3) the faces
  const float* eb = reinterpret_cast<const float*>(s.lr);
  for (int e = t; e < NF; e += NT) {
    const bool x2 = e < NF2;
    int f, tt, fc, last;
    if (x2) {
      fc = e % (SF_J + 1);
      f = j0 + fc;
      tt = k0 + e / (SF_J + 1);
      last = SF_J;
    } else {
      const int e3 = e - NF2;
      fc = e3 / SF_J;
      f = k0 + fc;
      tt = j0 + e3 % SF_J;
      last = SF_K;
    }
    if (tt >= n || f > n || (!DIV && fc == last && f != n)) continue;
    float a[5], b[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      a[v] = ql[v][e];
      b[v] = qr[v][e];
    }
    const int side = x2 ? 0 : 2;
    if (f == 0) {  // lower panel edge: the outer (left) state is the neighbour's
      const long long ei = ((long long)p * 4 + side) * n + tt;
#pragma unroll
      for (int v = 0; v < 5; ++v) a[v] = eb[sph_eb(s, ei, v, ii)];
    }
    if (f == n) {  // upper panel edge: the outer (right) state is the neighbour's
      const long long ei = ((long long)p * 4 + side + 1) * n + tt;
#pragma unroll
      for (int v = 0; v < 5; ++v) b[v] = eb[sph_eb(s, ei, v, ii)];
    }
    float fo[5];
    float* out = DIV ? fo : nullptr;
    if (x2)
      sph_face_riemann<1, float>(s, th, p, f, tt, ii, a, b, out);
    else
      sph_face_riemann<2, float>(s, th, p, f, tt, ii, a, b, out);
    if (DIV) {
#pragma unroll
      for (int v = 0; v < 5; ++v) fxs[v][e] = fo[v];
    }
  }
  if (!DIV) return;
  __syncthreads();
  // (4) the cell of this lane: its x2 / x3 face fluxes from LDS, its x1 ones
  // from f1
  const int cj = t % SF_J, ck = t / SF_J;
  if (j0 + cj >= n || k0 + ck >= n) return;
  const int J = g.js + j0 + cj, K = g.ks + k0 + ck;
  const float* F1 = reinterpret_cast<const float*>(s.f1[p]);
  double f1l[5], f1r[5], f2l[5], f2r[5], f3l[5], f3r[5];
  const int e2 = ck * (SF_J + 1) + cj, e3 = NF2 + ck * SF_J + cj;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    f1l[v] = (double)F1[g.f[0].at(v, K, J, i)];
    f1r[v] = (double)F1[g.f[0].at(v, K, J, i + 1)];
    f2l[v] = (double)fxs[v][e2];
    f2r[v] = (double)fxs[v][e2 + 1];
    f3l[v] = (double)fxs[v][e3];
    f3r[v] = (double)fxs[v][e3 + SF_J];
  }
  sph_cell_du(s, p, K, J, i, beta * dtp[0], f1l, f1r, f2l, f2r, f3l, f3r, psplit);
}

// x1 columns of panel blockIdx.y: forward sweep with the panel geometry
template <int B>
__global__ void __launch_bounds__(64)
    k_sph_implicit_fwd(KSph s, KThermo th, double beta, const double* __restrict__ dtp,
                       double* __restrict__ arch, long long arch_panel,
                       int* __restrict__ status) {
  const int p = blockIdx.y;
  const KGrid& g = s.g;
  const long long ncol = (long long)g.nx2 * g.nx3;
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int j = g.js + (int)(col % g.nx2);
  const int k = g.ks + (int)(col / g.nx2);
  KColGeo cgeo;
  cgeo.x1f = s.x1f;
  cgeo.x1v = s.x1v;
  cgeo.dx1 = s.dx1;
  cgeo.d2a = skj(s, p, cg_host::SKJ_D2A, k, j);
  cgeo.d3a = skj(s, p, cg_host::SKJ_D3A, k, j);
  cgeo.cosc = skj(s, p, cg_host::SKJ_COSC, k, j);
  cgeo.sinc = skj(s, p, cg_host::SKJ_SINC, k, j);
  if (!fwd_column<0, B, true>(g, th, beta * dtp[0], s.w[p], s.du[p], g.at(0, k, j, 0),
                              arch + p * arch_panel, ncol, col, nullptr, 0, &cgeo))
    atomicOr(status, 1);
}

// x1 columns of panel blockIdx.y after the forward sweep, top down: the
// back substitution and write-back (forward_backward.hpp:86-124), the stage
// update and ConservedToPrimitive of k_sph_update, the reflecting x1 ghosts
// of k_sph_bc_x1 and the next stage's hydrostatic face and cell pressures of
// k_sph_scan, each cell as its new state appears (one pass over the column;
// each step's arithmetic is that of the separate kernel, which the explicit
// step still uses).
template <int B>
__global__ void __launch_bounds__(64)
    k_sph_bwd_update(KSph s, KUpdate up, const double* __restrict__ arch,
                     long long arch_panel) {
  constexpr int AS = B * B + B;
  const int p = blockIdx.y;
  const KGrid& g = s.g;
  const long long ncol = (long long)g.nx2 * g.nx3;
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int j = g.js + (int)(col % g.nx2);
  const int k = g.ks + (int)(col / g.nx2);
  const double* ap = arch + p * arch_panel;
  const double* du = s.du[p];
  double* u = s.u[p];
  double* u1 = s.u1[p];
  double* w = s.w[p];
  double* psf = s.psf[p];
  double* psv = s.psv[p];
  const long long sn = g.c.n, si = g.c.i;
  const long long c0 = g.at(0, k, j, 0);
  const long long pc = col * g.ps_c;
  const double cth = skj(s, p, cg_host::SKJ_COSC, k, j);
  const int is = g.is, ie = g.ie;
  const double grav = g.grav, dx = g.dx1;
  const bool hs = grav != 0.0;
  double pf_next = 0.0, pf_top = 0.0;
  // the archive and state of a level are requested two levels ahead of its
  // use (three buffers), so that their latency overlaps the levels above
  struct Level {
    Mt<B> a;
    Vt<B> dl;
    double duo[5], uo[5], u1o[5];
  };
  auto fetch = [&](Level& L, int i) {
    const long long lb = (long long)(i - is) * AS;
#pragma unroll
    for (int r = 0; r < B; ++r)
      L.dl.v[r] = __builtin_nontemporal_load(&ap[(lb + B * B + r) * ncol + col]);
    if (i != ie) {
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q)
          L.a.a[r][q] = __builtin_nontemporal_load(&ap[(lb + r * B + q) * ncol + col]);
    }
    const long long ci = c0 + i * si;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      L.duo[v] = du[ci + v * sn];
      L.uo[v] = u[ci + v * sn];
      L.u1o[v] = (up.stage == 0) ? 0.0 : u1[ci + v * sn];
    }
  };
  Vt<B> dn;
  auto level = [&](const Level& L, int i) {
    // back substitution (forward_backward.hpp:86-97)
    Vt<B> d;
    if (i == ie) {
      d = L.dl;
    } else {
      Vt<B> ad;
      mvec<B>(L.a, dn, ad);
#pragma unroll
      for (int r = 0; r < B; ++r) d.v[r] = L.dl.v[r] - ad.v[r];
    }
    dn = d;
    const long long ci = c0 + i * si;
    double wsp[5], dun[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) wsp[v] = 0.0;
    write_back<0, B>(d, L.duo, wsp, dun);
    double uc[5], wc[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      double x;
      if (up.stage == 0) {
        x = 0.0 + 1.0 * L.uo[v];
        u1[ci + v * sn] = x;
      } else {
        x = up.g1 * L.uo[v] + up.g2 * L.u1o[v];
      }
      uc[v] = x + dun[v];
    }
    sph_cons_to_prim(s.gm1, cth, uc, wc);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      u[ci + v * sn] = uc[v];
      w[ci + v * sn] = wc[v];
    }
    // reflecting x1 walls: ghost m mirrors the cell m from the wall
    const int mt = ie - i, mb = i - is;
    if (mt < NG) {
#pragma unroll
      for (int v = 0; v < 5; ++v)
        w[c0 + v * sn + (ie + 1 + mt) * si] = (v == 1) ? -1.0 * wc[v] : wc[v];
    }
    if (mb < NG) {
#pragma unroll
      for (int v = 0; v < 5; ++v)
        w[c0 + v * sn + (is - 1 - mb) * si] = (v == 1) ? -1.0 * wc[v] : wc[v];
    }
    if (hs) {
      if (i == ie) {
        const double RdTv = wc[4] / wc[0];
        pf_next = wc[4] * exp_d(grav * dx / (2 * RdTv));
        pf_top = pf_next;
        psf[pc + (ie + 1) * g.ps_i] = pf_next;
      }
      const double pfi = pf_next - grav * wc[0] * dx;
      psf[pc + i * g.ps_i] = pfi;
      psv[pc + i * g.ps_i] = log_mean(pfi, pf_next);
      pf_next = pfi;
    }
  };
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
  if (hs) {
    // ghost levels: below from psf(is) down, above from psf(ie + 1) up
    const double gb = (g.bc[0] == 0) ? -grav : grav;
    const double gt = (g.bc[1] == 0) ? -grav : grav;
    double cur = pf_next;
    for (int gh = 1; gh <= NG; ++gh) {
      const double lo = cur - gb * w[c0 + (is - gh) * si] * dx;
      psf[pc + (is - gh) * g.ps_i] = lo;
      psv[pc + (is - gh) * g.ps_i] = log_mean(lo, cur);
      cur = lo;
    }
    cur = pf_top;
    for (int gh = 1; gh <= NG; ++gh) {
      const double hi = cur + gt * w[c0 + (ie + gh) * si] * dx;
      psf[pc + (ie + gh + 1) * g.ps_i] = hi;
      psv[pc + (ie + gh) * g.ps_i] = log_mean(cur, hi);
      cur = hi;
    }
  }
}

// end of stage: u1 / u averaging, u += du, ConservedToPrimitive
__global__ void __launch_bounds__(256) k_sph_update(KSph s, KUpdate up) {
  const int p = blockIdx.y;
  int i, j, k;
  if (!sph_cell(s.g, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const KGrid& g = s.g;
  const long long c0 = g.at(0, k, j, i);
  double uc[5], wc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const long long c = c0 + v * g.c.n;
    const double un = s.u[p][c];
    double x;
    if (up.stage == 0) {
      x = 0.0 + 1.0 * un;
      s.u1[p][c] = x;
    } else {
      x = up.g1 * un + up.g2 * s.u1[p][c];
    }
    uc[v] = x + s.du[p][c];
  }
  sph_cons_to_prim(s.gm1, skj(s, p, cg_host::SKJ_COSC, k, j), uc, wc);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    s.u[p][c0 + v * g.c.n] = uc[v];
    s.w[p][c0 + v * g.c.n] = wc[v];
  }
}

// x2 / x3 ghost zones of every panel from the neighbours' conserved state
__global__ void __launch_bounds__(256)
    k_sph_ghosts(KSph s, const KSphPtrs* __restrict__ ptr, const SphGhost* __restrict__ gh,
                 int nghost) {
  const long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= (long long)nghost * s.nr) return;
  const SphGhost h = gh[item / s.nr];
  const int i = s.g.is + (int)(item % s.nr);
  const KGrid& g = s.g;
  const double* uq = ptr->u[h.q];
  double* wp = ptr->w[h.p];
  const long long a1 = g.at(0, h.k1, h.j1, i), a2 = g.at(0, h.k2, h.j2, i);
  double ug[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double x1 = uq[a1 + v * g.c.n], x2 = uq[a2 + v * g.c.n];
    ug[v] = ((h.y2 - h.yg) * x1 + (h.yg - h.y1) * x2) / (h.y2 - h.y1);
  }
  const double vy = ug[2], vz = ug[3];
  const double c2 = vy / h.sq / h.sq - vz * h.cq / h.sq / h.sq;
  const double c3 = vz / h.sq / h.sq - vy * h.cq / h.sq / h.sq;
  const double U = h.r00 * c2 + h.r01 * c3;
  const double V = h.r10 * c2 + h.r11 * c3;
  const double d2 = h.a00 * U + h.a01 * V;
  const double d3 = h.a10 * U + h.a11 * V;
  ug[2] = d2 + d3 * h.cp;
  ug[3] = d3 + d2 * h.cp;
  double wg[5];
  sph_cons_to_prim(s.gm1, ptr->kj[h.p][(long long)cg_host::SKJ_COSC * s.nc * s.nc +
                                       (long long)h.kg * s.nc + h.jg],
                   ug, wg);
  const long long c0 = g.at(0, h.kg, h.jg, i);
#pragma unroll
  for (int v = 0; v < 5; ++v) wp[c0 + v * g.c.n] = wg[v];
}

}  // namespace cg
