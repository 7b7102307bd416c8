// lines.hpp -- the lines of ImplicitSolver::SolveImplicit3D other than the
// batched non-periodic x1 columns: x2 and x3 lines (implicit_flag bits 1
// and 2) and lines whose two ends are periodic, in any direction
// (src/snap/implicit/solve_implicit_3d.cpp:19-143,
// periodic_forward_backward.hpp:14-232).
//
// One line per lane, one launch per direction: the blocks a_i, b_i, c_i
// and the right-hand side of each level are assembled in registers as the
// lane climbs the line and fed straight into the recurrence; only what the
// way back needs goes to a lane-contiguous scratch (walls: a'_i and
// delta_i; periodic ends: ainv_i, gamma_i, zeta_i and c_i).
#pragma once

#include "kernels.hpp"

namespace cg {

// A family of lines of one direction.  Line t has transverse interior
// indices (f, s) = (t % nf, t / nf) along the transverse index of smaller
// stride (f, on consecutive lanes) and the other one (s); level p of the
// line (p = 0 .. n-1) is the cell at org(t) + (lo + p) * sl.
struct KLines {
  long long nline;
  int n, lo, nf;
  long long org0, sf, ss, sl, sn;
  double sla, sra, sda;  // aleft / (2 vol), aright / (2 vol), their difference
  double grav;           // the x1 lines' gravity, 0 on x2 / x3 lines
  // cell indices of line t: direction td0 runs from t0s with t % nf,
  // td1 from t1s with t / nf; nc* and c describe w for the shuffled read
  int td0, td1, t0s, t1s;
  int nc1, nc2, nc3;
  Strides c;
  CG_HD long long org(long long t) const {
    return org0 + (t % nf) * sf + (t / nf) * ss;
  }
};

// w_n at the entry the reference's BackwardSubstitution reads for a cell
// of an x2 / x3 line (forward_backward.hpp:109-115): it indexes w with the
// loop indices (k, j, i) of the shuffled line buffer (solve_implicit_3d.cpp:
// 28-31, 59-62) while w itself is not shuffled, so the read is the
// AthenaArray entry ((n nc3 + K) nc2 + J) nc1 + I with (K, J, I) = (x1, x3,
// x2) of the cell on x2 lines and (x2, x1, x3) on x3 lines.  That entry,
// decoded into (n, k, j, i) of w's own layout; the host checks that it lies
// inside w.
template <int DIR>
__device__ inline double w_shuffled(const KLines& L, const double* __restrict__ w,
                                    int n, long long t, int i) {
  const int a0 = L.t0s + (int)(t % L.nf), a1 = L.t1s + (int)(t / L.nf);
  int x[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) x[d] = (d == DIR) ? i : (d == L.td0 ? a0 : a1);
  long long K, J, I;
  if (DIR == 1) {
    K = x[0]; J = x[2]; I = x[1];
  } else {
    K = x[1]; J = x[0]; I = x[2];
  }
  const long long f = ((n * (long long)L.nc3 + K) * L.nc2 + J) * L.nc1 + I;
  const long long r = f / L.nc1, r2 = r / L.nc2;
  return w[L.c.at((int)(r2 / L.nc3), (int)(r2 % L.nc3), (int)(r % L.nc2),
                  (int)(f % L.nc1))];
}

// Scratch doubles of one level: walls a'_i, delta_i (B x B + B); periodic
// ends ainv_i, gamma_i, c_i (B x B each) and zeta_i (B).
template <int B>
CG_HD constexpr int line_level_size(bool per) {
  return per ? 4 * B * B + B : B * B + B;
}

// The velocities of a state in the frame of direction DIR: slot IVX holds
// the normal component, IVX + 1 and IVX + 2 the next ones in cyclic order
// (flux_decomposition.hpp:80-83, 105-107).
template <int NM, int DIR>
__device__ inline void to_frame(const double* src, double* dst) {
  constexpr int N = 5 + NM, IVX = 1 + NM;
#pragma unroll
  for (int n = 0; n < N; ++n) dst[n] = src[n];
#pragma unroll
  for (int s = 0; s < 3; ++s) dst[IVX + s] = src[IVX + (s + DIR) % 3];
}

template <int B>
__device__ inline void mstore(const Mt<B>& m, double* p, long long st) {
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int q = 0; q < B; ++q) p[(r * B + q) * st] = m.a[r][q];
}
template <int B>
__device__ inline void mload(Mt<B>& m, const double* p, long long st) {
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int q = 0; q < B; ++q) m.a[r][q] = p[(r * B + q) * st];
}

// Assembly of FullCorrection / PartialCorrection (full_correction.cpp:
// 77-146, partial_correction.cpp:69-155) along one line, one level per
// call: level(i) returns a_i (unfolded), b_i, c_i and rhs_i of cell i and
// moves the window up.  Velocities enter in the frame of DIR; the Roe
// average is formed from the unrotated states (RoeAverage precedes the
// choice of frame, :98-104).  start(i0) primes the window so that the
// first level() is cell i0.  The w of cell i + 2 and the du of cell i + 1
// are requested at the top of level(i), one level ahead of their use, up to
// the line's last cell.
template <int NM, int B, int DIR>
struct LineAsm {
  static constexpr int N = 5 + NM, IVX = 1 + NM, IEN = 4 + NM;
  const KLines& L;
  const KThermo& th;
  const double* __restrict__ w;
  const double* __restrict__ du;
  long long o;
  double gm1, gm, idt;
  int hi;  // last cell of the line
  double wc[N], wn[N], dn[N];
  Mt<B> Am, dm, dc;
  __device__ LineAsm(const KLines& L_, const KThermo& th_, const double* w_,
                     const double* du_, long long o_, double dt)
      : L(L_), th(th_), w(w_), du(du_), o(o_) {
    gm1 = th.gammad - 1.;
    gm = 0.5 * (gm1 + gm1);
    idt = 1. / dt;
    hi = L.lo + L.n - 1;
  }
  __device__ void load(double* dst, int i) const {
#pragma unroll
    for (int n = 0; n < N; ++n) dst[n] = w[o + n * L.sn + (long long)i * L.sl];
  }
  __device__ void load_du(double* dst, int i) const {
#pragma unroll
    for (int n = 0; n < N; ++n) dst[n] = du[o + n * L.sn + (long long)i * L.sl];
  }
  __device__ void jac(const double* wv, Mt<B>& out) const {
    double f[N];
    to_frame<NM, DIR>(wv, f);
    flux_jac<NM, B>(gm1, f, out);
  }
  __device__ void absA(const double* wl, const double* wr, Mt<B>& out) const {
    double prim[N], pf[N];
    roe_average<NM>(prim, gm, wl, wr);
    to_frame<NM, DIR>(prim, pf);
    abs_roe_fma<NM, B>(th, pf, gm, out);
  }
  __device__ void start(int i0) {
    double wm[N];
    load(wm, i0 - 1);
    load(wc, i0);
    load(wn, i0 + 1);
    if (i0 >= L.lo) {
      load_du(dn, i0);
    } else {
#pragma unroll
      for (int n = 0; n < N; ++n) dn[n] = 0.0;
    }
    absA(wm, wc, Am);
    jac(wm, dm);
    jac(wc, dc);
  }
  __device__ void level(int i, Mt<B>& a, Mt<B>& b, Mt<B>& c, double* rhs,
                        bool want_rhs) {
    double wp[N], d[N];
#pragma unroll
    for (int n = 0; n < N; ++n) {
      wp[n] = wn[n];
      d[n] = dn[n];
    }
    if (i + 1 <= hi) {
      load(wn, i + 2);
      load_du(dn, i + 1);
    }
    Mt<B> Ap, dp;
    absA(wc, wp, Ap);
    jac(wp, dp);
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int q = 0; q < B; ++q) {
        double phi = 0.0;
        if (B == 5) {
          if ((r == 1 && q == 0) || (r == 4 && q == 1)) phi = L.grav;
        } else {
          if ((r == 1 && q == 0) || (r == 2 && q == 1)) phi = L.grav;
        }
        double av = fma(Am.a[r][q], L.sla, fma(Ap.a[r][q], L.sra, dc.a[r][q] * L.sda));
        if (r == q) av += idt;
        a.a[r][q] = av - phi;
        b.a[r][q] = -(Am.a[r][q] + dm.a[r][q]) * L.sla;
        c.a[r][q] = -(Ap.a[r][q] - dp.a[r][q]) * L.sra;
      }
    if (want_rhs) {  // rhs in the frame of DIR (forward_backward.hpp:24-36)
      double r0 = d[0];
#pragma unroll
      for (int n = 1; n < IVX; ++n) r0 += d[n];
      rhs[0] = r0 * idt;
      if (B == 5) {
        rhs[1] = d[IVX + DIR] * idt;
        rhs[2] = d[IVX + (1 + DIR) % 3] * idt;
        rhs[3] = d[IVX + (2 + DIR) % 3] * idt;
        rhs[4] = d[IEN] * idt;
      } else {
        rhs[1] = d[IVX + DIR] * idt;
        rhs[2] = d[IEN] * idt;
      }
    }
    // the next level's b: -(A_{i+1/2} + dfdq_i) aleft / (2 vol)
    Am = Ap;
    dm = dc;
    dc = dp;
#pragma unroll
    for (int n = 0; n < N; ++n) wc[n] = wp[n];
  }
  // b of the next level without moving the window: -(Am + dm) sla after
  // the last level() call
  __device__ void next_b(Mt<B>& b) const {
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int q = 0; q < B; ++q) b.a[r][q] = -(Am.a[r][q] + dm.a[r][q]) * L.sla;
  }
};

// du of a cell from the line's solution d (frame of DIR).  Walls: the
// write-back of BackwardSubstitution (forward_backward.hpp:100-124), the
// total-density increment split over the species by w_n (the cell's own
// on x1 lines, w_shuffled on x2 / x3 lines; cell i of line t).  Periodic: the
// write-back of PeriodicBackwardSubstitution (periodic_forward_backward.hpp:
// 216-229), species unchanged.  x2 / x3 lines store IDN, the normal momentum
// and IEN only, the variables the reference copies out of its line buffer
// (solve_implicit_3d.cpp:51-56, 82-87).
template <int NM, int B, int DIR, bool PER>
__device__ inline void line_write_back(const Vt<B>& d, const double* __restrict__ w,
                                       double* __restrict__ du, long long ci,
                                       long long sn, const KLines& L, long long t,
                                       int i) {
  constexpr int N = 5 + NM, IVX = 1 + NM, IEN = 4 + NM;
  constexpr int vn = IVX + DIR, vt1 = IVX + (1 + DIR) % 3, vt2 = IVX + (2 + DIR) % 3;
  double o[N];
#pragma unroll
  for (int n = 0; n < N; ++n) o[n] = du[ci + n * sn];
  if (!PER) {
    double dens = o[0];
#pragma unroll
    for (int n = 1; n < IVX; ++n) dens += o[n];
    dens = d.v[0] - dens;
#pragma unroll
    for (int n = 1; n < IVX; ++n)
      o[n] += dens * (DIR == 0 ? w[ci + n * sn] : w_shuffled<DIR>(L, w, n, t, i));
  }
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
#pragma unroll
  for (int n = 1; n < IVX; ++n) o[0] -= o[n];
  if (DIR == 0) {
#pragma unroll
    for (int n = 0; n < N; ++n) du[ci + n * sn] = o[n];
  } else {
    du[ci] = o[0];
    du[ci + vn * sn] = o[vn];
    du[ci + IEN * sn] = o[IEN];
  }
}

// Gauss-Jordan elimination of [M | C | t] with partial pivoting (first row
// of largest |pivot|): C = M^-1 C, t = M^-1 t.  Row exchanges run only when
// a lane of the wavefront pivots, over the columns of M still read.
// Returns false on a zero pivot.
template <int B>
__device__ inline bool gj_solve_d(Mt<B>& M, Mt<B>& C, double* t) {
  bool ok = true;
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
    if (__any(p != cc)) {
#pragma unroll
      for (int r = cc + 1; r < B; ++r) {
        bool sw = (p == r);
#pragma unroll
        for (int q = 0; q < B; ++q) {
          if (q >= cc) {
            double x0 = M.a[cc][q], x1 = M.a[r][q];
            M.a[cc][q] = sw ? x1 : x0;
            M.a[r][q] = sw ? x0 : x1;
          }
          double y0 = C.a[cc][q], y1 = C.a[r][q];
          C.a[cc][q] = sw ? y1 : y0;
          C.a[r][q] = sw ? y0 : y1;
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
    for (int q = 0; q < B; ++q) C.a[cc][q] *= inv;
    t[cc] *= inv;
#pragma unroll
    for (int r = 0; r < B; ++r) {
      if (r == cc) continue;
      double f = M.a[r][cc];
#pragma unroll
      for (int q = cc + 1; q < B; ++q) M.a[r][q] = fma(-f, M.a[cc][q], M.a[r][q]);
#pragma unroll
      for (int q = 0; q < B; ++q) C.a[r][q] = fma(-f, C.a[cc][q], C.a[r][q]);
      t[r] = fma(-f, t[cc], t[r]);
    }
  }
  return ok;
}

template <int B>
__device__ inline void msub(const Mt<B>& x, const Mt<B>& y, Mt<B>& z) {
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int q = 0; q < B; ++q) z.a[r][q] = x.a[r][q] - y.a[r][q];
}

// ForwardSweep / BackwardSubstitution (forward_backward.hpp:14-137) of a
// line with walls (full_correction.cpp:144-145 folds them into a_lo and
// a_hi): M = a_p - b_p a'_{p-1}, a'_p = M^-1 c_p and
// delta_p = M^-1 (rhs_p - b_p delta_{p-1}) from one Gauss-Jordan elimination
// of [M | c_p | rhs_p], archived; then delta_p -= a'_p delta_{p+1} down the
// line and the write-back.  The step is dt, or dt * dtp[0] with a device
// step (runs: dt is the stage weight).
template <int NM, int B, int DIR>
__global__ void __launch_bounds__(64)
    k_line_walls(KLines L, KThermo th, double dt, const double* __restrict__ dtp,
                 const double* __restrict__ w, double* __restrict__ du,
                 double* __restrict__ lv, int* __restrict__ status) {
  constexpr int AS = line_level_size<B>(false);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.nline) return;
  const long long o = L.org(t);
  const long long st = L.nline;
  const int lo = L.lo, hi = L.lo + L.n - 1;
  LineAsm<NM, B, DIR> as(L, th, w, du, o, dtp ? dt * dtp[0] : dt);
  as.start(lo);
  bool ok = true;
  Mt<B> ap;
  Vt<B> dl;
  for (int i = lo; i <= hi; ++i) {
    Mt<B> M, b, c;
    double tv[B];
    as.level(i, M, b, c, tv, true);
    if (i == lo) {
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q) M.a[r][q] += (q == 1) ? -b.a[r][q] : b.a[r][q];
    }
    if (i == hi) {
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q) M.a[r][q] += (q == 1) ? -c.a[r][q] : c.a[r][q];
    }
    if (i > lo) {
#pragma unroll
      for (int r = 0; r < B; ++r) {
#pragma unroll
        for (int q = 0; q < B; ++q) {
          double s = M.a[r][q];
#pragma unroll
          for (int kk = 0; kk < B; ++kk) s = fma(-b.a[r][kk], ap.a[kk][q], s);
          M.a[r][q] = s;
        }
        double s = tv[r];
#pragma unroll
        for (int kk = 0; kk < B; ++kk) s = fma(-b.a[r][kk], dl.v[kk], s);
        tv[r] = s;
      }
    }
    if (!gj_solve_d<B>(M, c, tv)) ok = false;
    ap = c;
#pragma unroll
    for (int r = 0; r < B; ++r) dl.v[r] = tv[r];
    double* base = lv + (long long)(i - lo) * AS * st + t;
    mstore<B>(ap, base, st);
#pragma unroll
    for (int r = 0; r < B; ++r) base[(B * B + r) * st] = dl.v[r];
  }
  if (!ok) atomicOr(status, 1);
  Vt<B> dn;
  for (int p = L.n - 1; p >= 0; --p) {
    const double* base = lv + (long long)p * AS * st + t;
    Vt<B> d;
#pragma unroll
    for (int r = 0; r < B; ++r) d.v[r] = base[(B * B + r) * st];
    if (p < L.n - 1) {
      Mt<B> a;
      mload<B>(a, base, st);
      Vt<B> ad;
      mvec<B>(a, dn, ad);
#pragma unroll
      for (int r = 0; r < B; ++r) d.v[r] = d.v[r] - ad.v[r];
    }
    dn = d;
    line_write_back<NM, B, DIR, false>(d, w, du, o + (long long)(lo + p) * L.sl, L.sn,
                                       L, t, lo + p);
  }
}

// PeriodicForwardSweep / PeriodicBackwardSubstitution (periodic_forward_
// backward.hpp:14-232) for a line held by one block (first_block =
// last_block = true), with L_p = b_p, d_p = a_p, U_p = c_p and the corner
// U_{-1} = c_{lo-1}:
//   ainv_0 = d_0^-1, gamma_0 = L_0, beta_0 = U_{-1} ainv_0, zeta_0 = r_0;
//   p = 1 .. n-2, T = L_p ainv_{p-1}: ainv_p = (d_p - T U_{p-1})^-1,
//     zeta_p = r_p - T zeta_{p-1}, gamma_p = -T gamma_{p-1},
//     beta_p = -(beta_{p-1} U_{p-1}) ainv_p, except at p = n-2:
//     gamma_{n-2} = U_{n-2} - T gamma_{n-3},
//     beta_{n-2} = (L_{n-1} - beta_{n-3} U_{n-3}) ainv_{n-2};
//   ainv_{n-1} = (d_{n-1} - sum beta_p gamma_p)^-1,
//   zeta_{n-1} = r_{n-1} - sum beta_p zeta_p (p < n-1);
//   x_{n-1} = ainv_{n-1} zeta_{n-1},
//   x_p = ainv_p (zeta_p - U_p x_{p+1} - gamma_p x_{n-1}), U_{n-2} = 0.
// Each level forms zeta_p and gamma_p first and parks them in the scratch,
// so that only the pivoted inverse and beta are live across it; beta_p is
// parked too, and the two sums run as their own pass over the scratch
// (p = 0 .. n-2, the order of the recurrence); L_{n-1} comes from the
// assembly window at level n-2.
// Products left to right as written, inverses by Gauss-Jordan with partial
// pivoting.
template <int NM, int B, int DIR>
__global__ void __launch_bounds__(64)
    k_line_periodic(KLines L, KThermo th, double dt, const double* __restrict__ dtp,
                    const double* __restrict__ w, double* __restrict__ du,
                    double* __restrict__ lv, int* __restrict__ status) {
  constexpr int AS = line_level_size<B>(true);
  // gamma, c, zeta, beta after ainv
  constexpr int OG = B * B, OU = 2 * B * B, OZ = 3 * B * B, OB = 3 * B * B + B;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.nline) return;
  const long long o = L.org(t);
  const long long st = L.nline;
  const int n = L.n, lo = L.lo;
  auto lvl = [&](int p) { return lv + (long long)p * AS * st + t; };
  LineAsm<NM, B, DIR> as(L, th, w, du, o, dtp ? dt * dtp[0] : dt);
  as.start(lo - 1);
  bool ok = true;
  Mt<B> bet;
  {  // the corner c_{lo-1}, then level 0
    Mt<B> a, b, um1;
    double rhs[B];
    as.level(lo - 1, a, b, um1, rhs, false);
    Mt<B> c, ai;
    as.level(lo, a, b, c, rhs, true);
    double* cur = lvl(0);
    mstore<B>(c, cur + OU * st, st);
    mstore<B>(b, cur + OG * st, st);  // gamma_0 = L_0
#pragma unroll
    for (int r = 0; r < B; ++r) cur[(OZ + r) * st] = rhs[r];
    if (!minv<B>(a, ai)) ok = false;
    mstore<B>(ai, cur, st);
    mmul<B>(um1, ai, bet);
    mstore<B>(bet, cur + OB * st, st);
  }
  // level p reads level p - 1's scratch and writes level p's: disjoint
  // entries, declared so that the reads may be issued before the writes
  auto step = [&](int p, double* __restrict__ cur, const double* __restrict__ prv) {
    Mt<B> d;
    {
      Mt<B> b, c;
      double rhs[B];
      as.level(lo + p, d, b, c, rhs, true);
      mstore<B>(c, cur + OU * st, st);
      Mt<B> T;
      {
        Mt<B> ap;
        mload<B>(ap, prv, st);  // ainv_{p-1}
        mmul<B>(b, ap, T);
      }
      {  // zeta_p = r_p - T zeta_{p-1}
        Vt<B> zp, Tz;
#pragma unroll
        for (int r = 0; r < B; ++r) zp.v[r] = prv[(OZ + r) * st];
        mvec<B>(T, zp, Tz);
#pragma unroll
        for (int r = 0; r < B; ++r) cur[(OZ + r) * st] = rhs[r] - Tz.v[r];
      }
      {  // gamma_p = -T gamma_{p-1}, or U_{n-2} - T gamma_{n-3}
        Mt<B> gp, Tg, g;
        mload<B>(gp, prv + OG * st, st);
        mmul<B>(T, gp, Tg);
        if (p < n - 2) {
#pragma unroll
          for (int r = 0; r < B; ++r)
#pragma unroll
            for (int q = 0; q < B; ++q) g.a[r][q] = -Tg.a[r][q];
        } else {
          msub<B>(c, Tg, g);
        }
        mstore<B>(g, cur + OG * st, st);
      }
      {  // d_p - T U_{p-1}
        Mt<B> Up, TU;
        mload<B>(Up, prv + OU * st, st);
        mmul<B>(T, Up, TU);
        msub<B>(d, TU, d);
      }
    }
    Mt<B> ai;
    if (!minv<B>(d, ai)) ok = false;
    mstore<B>(ai, cur, st);
    {  // beta_p = -(beta_{p-1} U_{p-1}) ainv_p, or
       // (L_{n-1} - beta_{n-3} U_{n-3}) ainv_{n-2}
      Mt<B> Up, bu, bnew;
      mload<B>(Up, prv + OU * st, st);
      mmul<B>(bet, Up, bu);
      if (p < n - 2) {
        mmul<B>(bu, ai, bnew);
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
          for (int q = 0; q < B; ++q) bnew.a[r][q] = -bnew.a[r][q];
      } else {
        Mt<B> Ln, x;
        as.next_b(Ln);
        msub<B>(Ln, bu, x);
        mmul<B>(x, ai, bnew);
      }
      bet = bnew;
      mstore<B>(bet, cur + OB * st, st);
    }
  };
  for (int p = 1; p <= n - 2; ++p) step(p, lvl(p), lvl(p - 1));
  // sums of beta_p gamma_p and beta_p zeta_p
  Mt<B> sbg;
  Vt<B> sbz;
#pragma unroll
  for (int r = 0; r < B; ++r) {
    sbz.v[r] = 0.0;
#pragma unroll
    for (int q = 0; q < B; ++q) sbg.a[r][q] = 0.0;
  }
  for (int p = 0; p <= n - 2; ++p) {
    const double* cur = lvl(p);
    Mt<B> bp, g, bg;
    Vt<B> z, bz;
    mload<B>(bp, cur + OB * st, st);
    mload<B>(g, cur + OG * st, st);
#pragma unroll
    for (int r = 0; r < B; ++r) z.v[r] = cur[(OZ + r) * st];
    mmul<B>(bp, g, bg);
    mvec<B>(bp, z, bz);
#pragma unroll
    for (int r = 0; r < B; ++r) {
      sbz.v[r] = sbz.v[r] + bz.v[r];
#pragma unroll
      for (int q = 0; q < B; ++q) sbg.a[r][q] = sbg.a[r][q] + bg.a[r][q];
    }
  }
  // last level
  Vt<B> xl;
  {
    Mt<B> d, b, c, ai;
    double rhs[B];
    as.level(lo + n - 1, d, b, c, rhs, true);
    msub<B>(d, sbg, d);
    if (!minv<B>(d, ai)) ok = false;
    Vt<B> z;
#pragma unroll
    for (int r = 0; r < B; ++r) z.v[r] = rhs[r] - sbz.v[r];
    mvec<B>(ai, z, xl);
  }
  if (!ok) atomicOr(status, 1);
  line_write_back<NM, B, DIR, true>(xl, w, du, o + (long long)(lo + n - 1) * L.sl, L.sn,
                                    L, t, lo + n - 1);
  Vt<B> xn = xl;
  Mt<B> zero;
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int q = 0; q < B; ++q) zero.a[r][q] = 0.0;
  for (int p = n - 2; p >= 0; --p) {
    const double* cur = lvl(p);
    Mt<B> U, ai, g;
    if (p == n - 2)
      U = zero;
    else
      mload<B>(U, cur + OU * st, st);
    mload<B>(ai, cur, st);
    mload<B>(g, cur + OG * st, st);
    Vt<B> ux, gx, tv, x;
    mvec<B>(U, xn, ux);
    mvec<B>(g, xl, gx);
#pragma unroll
    for (int r = 0; r < B; ++r) tv.v[r] = (cur[(OZ + r) * st] - ux.v[r]) - gx.v[r];
    mvec<B>(ai, tv, x);
    xn = x;
    line_write_back<NM, B, DIR, true>(x, w, du, o + (long long)(lo + p) * L.sl, L.sn, L,
                                      t, lo + p);
  }
}

}  // namespace cg
