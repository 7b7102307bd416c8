// sphere_geom.cpp -- host-side geometry of the cubed-sphere run; see
// sphere_geom.hpp.  The metric formulas are the reference's
// (src/exo3/gnomonic_equiangle.cpp:136-196); neighbours, ghost points and
// vector maps follow from the panel positions of GetLatLon.
#include "sphere_geom.hpp"

#include <algorithm>
#include <cmath>

namespace cg_host {

namespace {

constexpr int NGH = 3;
constexpr double PI_ = 3.14159265358979323846;

// GetLatLon (cubed_sphere.cpp:78-89) of panel p (1-based) at (X, Y)
void latlon(int p, double X, double Y, double* lat, double* lon) {
  const double dX = Y, dY = -X;
  switch (p - 1) {
    case 1:
      *lon = std::atan(dX);
      *lat = std::atan(dY / std::sqrt(1.0 + dX * dX));
      break;
    case 3:
      *lon = std::atan(dX) + 0.5 * M_PI;
      *lat = std::atan(dY / std::sqrt(1.0 + dX * dX));
      break;
    case 5:
      *lon = std::atan(dX) + M_PI;
      *lat = std::atan(dY / std::sqrt(1.0 + dX * dX));
      break;
    case 2:
      *lon = std::atan(dX) + 1.5 * M_PI;
      *lat = std::atan(dY / std::sqrt(1.0 + dX * dX));
      break;
    case 0:
      if (std::fabs(dX) > 1.0e-10)
        *lon = std::atan2(dX, -dY);
      else if (dY <= 0.0)
        *lon = 0.0;
      else
        *lon = M_PI;
      *lat = 0.5 * M_PI - std::atan(std::sqrt(dX * dX + dY * dY));
      break;
    default:
      if (std::fabs(dX) > 1.0e-10)
        *lon = std::atan2(dX, dY);
      else if (dY > 0.0)
        *lon = 0.0;
      else
        *lon = M_PI;
      *lat = -0.5 * M_PI + std::atan(std::sqrt(dX * dX + dY * dY));
      break;
  }
  if (*lon < 0.0) *lon += 2.0 * M_PI;
  *lon = 2.0 * PI_ - *lon;
}

// row angle of CalculateInterpLocations (cubed_sphere_utility.cpp:80-100)
double row_angle(double tn, double ta) {
  double temp0 = 1.0 + tn * tn;
  double s = std::acos(std::sqrt(temp0 / (temp0 + ta * ta)));
  return (ta < 0) ? -s : s;
}

// (X, Y) of the point (lat, lon) on panel q (1-based), inverting latlon();
// false on the far side of q's cube face
bool xy_on_panel(int q, double lat, double lon, double* X, double* Y) {
  const double lp = 2.0 * PI_ - lon;
  const int nP = q - 1;
  double a, b;
  if (nP == 0) {
    if (lat <= 0.) return false;
    const double r = 1.0 / std::tan(lat);
    a = r * std::sin(lp);
    b = -r * std::cos(lp);
  } else if (nP == 4) {
    if (lat >= 0.) return false;
    const double r = -1.0 / std::tan(lat);
    a = r * std::sin(lp);
    b = r * std::cos(lp);
  } else {
    const double l0 = (nP == 1) ? 0. : (nP == 3) ? 0.5 * PI_ : (nP == 5) ? PI_ : 1.5 * PI_;
    const double d = lp - l0;
    if (std::cos(d) <= 0.) return false;
    a = std::tan(d);
    b = std::tan(lat) * std::sqrt(1.0 + a * a);
  }
  *Y = a;
  *X = -b;
  return true;
}

// the panel other than p (1-based) whose face holds (lat, lon)
int find_panel(int p, double lat, double lon, double* X, double* Y) {
  int best = -1;
  double bm = 1e30;
  for (int q = 1; q <= 6; ++q) {
    if (q == p) continue;
    double x, y;
    if (!xy_on_panel(q, lat, lon, &x, &y)) continue;
    const double m = std::max(std::fabs(x), std::fabs(y));
    if (m < bm) {
      bm = m;
      best = q;
      *X = x;
      *Y = y;
    }
  }
  return best;
}

// VecTransRLLFromABP / VecTransABPFromRLL (cubed_sphere_transform.cpp:11-57)
// at (X, Y) of panel `block` (1-based) as coefficients of the two components:
// U = r00 V2 + r01 V3, V = r10 V2 + r11 V3 and V2 = a00 U + a01 V,
// V3 = a10 U + a11 V, each evaluated in the reference's expression order (a
// difference a - b as the sum with the negated product, a missing term as a
// zero coefficient).
void vectrans(double X, double Y, int block, double* r, double* a) {
  const double C = std::sqrt(1 + X * X), D = std::sqrt(1 + Y * Y);
  const double E = std::sqrt(X * X + Y * Y), delta = std::sqrt(1 + X * X + Y * Y);
  if (block == 1) {
    r[0] = -delta * Y / (D * E);
    r[1] = delta * X / (C * E);
    r[2] = -X / (D * E);
    r[3] = -(Y / (C * E));
    a[0] = -D * Y / (delta * E);
    a[1] = -(D * X / E);
    a[2] = C * X / (delta * E);
    a[3] = -(C * Y / E);
  } else if (block == 5) {
    r[0] = delta * Y / (D * E);
    r[1] = -(delta * X / (C * E));
    r[2] = X / (D * E);
    r[3] = Y / (C * E);
    a[0] = D * Y / (delta * E);
    a[1] = D * X / E;
    a[2] = -C * X / (delta * E);
    a[3] = C * Y / E;
  } else {
    r[0] = 0.0;
    r[1] = delta / (C * D);
    r[2] = -1.0;
    r[3] = X * Y / (C * D);
    a[0] = X * Y / delta;
    a[1] = -1.0;
    a[2] = C * D / delta;
    a[3] = 0.0;
  }
}

}  // namespace

SphGeom sphere_geometry(int n, int nr, double r0, double dx1) {
  SphGeom G;
  G.n = n;
  G.nc = n + 2 * NGH;
  G.nr = nr;
  G.nc1 = nr + 2 * NGH;
  const int nc = G.nc;
  std::vector<double> af(nc + 1), av(nc);
  for (int j = 0; j <= nc; ++j) af[j] = (double)(2 * (j - NGH) - n) * (PI_ / (4.0 * n));
  for (int j = 0; j < nc; ++j) av[j] = 0.5 * (af[j + 1] + af[j]);
  G.x1f.resize(G.nc1 + 1);
  G.x1v.resize(G.nc1);
  for (int i = 0; i <= G.nc1; ++i) G.x1f[i] = r0 + (double)(i - NGH) * dx1;
  for (int i = 0; i < G.nc1; ++i) G.x1v[i] = 0.5 * (G.x1f[i + 1] + G.x1f[i]);
  const size_t m = (size_t)nc * nc;
  for (int p = 0; p < 6; ++p) {
    std::vector<double>& T = G.kj[p];
    T.assign(SKJ_COUNT * m, 0.0);
    auto at = [&](int a, int k, int j) -> double& { return T[a * m + (size_t)k * nc + j]; };
    for (int k = 0; k < nc; ++k)
      for (int j = 0; j < nc; ++j) {
        double x = std::tan(av[j]), xf = std::tan(af[j]);
        double y = std::tan(av[k]), yf = std::tan(af[k]);
        double C = std::sqrt(1.0 + x * x), Cf = std::sqrt(1.0 + xf * xf);
        double D = std::sqrt(1.0 + y * y), Df = std::sqrt(1.0 + yf * yf);
        at(SKJ_COSC, k, j) = -x * y / (C * D);
        at(SKJ_SINC, k, j) = std::sqrt(1.0 + x * x + y * y) / (C * D);
        at(SKJ_COSF2, k, j) = -xf * y / (Cf * D);
        at(SKJ_SINF2, k, j) = std::sqrt(1.0 + xf * xf + y * y) / (Cf * D);
        at(SKJ_COSF3, k, j) = -x * yf / (C * Df);
        at(SKJ_SINF3, k, j) = std::sqrt(1.0 + x * x + yf * yf) / (C * Df);
        double x1 = std::tan(af[j]), x2 = std::tan(af[j + 1]);
        double d1 = std::sqrt(1.0 + x1 * x1 + y * y), d2 = std::sqrt(1.0 + x2 * x2 + y * y);
        double d1f = std::sqrt(1.0 + x1 * x1 + yf * yf), d2f = std::sqrt(1.0 + x2 * x2 + yf * yf);
        at(SKJ_D2A, k, j) = std::acos((1 + x1 * x2 + y * y) / (d1 * d2));
        at(SKJ_D2AF3, k, j) = std::acos((1 + x1 * x2 + yf * yf) / (d1f * d2f));
        double y1 = std::tan(af[k]), y2 = std::tan(af[k + 1]);
        d1 = std::sqrt(1.0 + x * x + y1 * y1);
        d2 = std::sqrt(1.0 + x * x + y2 * y2);
        d1f = std::sqrt(1.0 + xf * xf + y1 * y1);
        d2f = std::sqrt(1.0 + xf * xf + y2 * y2);
        at(SKJ_D3A, k, j) = std::acos((1 + x * x + y1 * y2) / (d1 * d2));
        at(SKJ_D3AF2, k, j) = std::acos((1 + xf * xf + y1 * y2) / (d1f * d2f));
        at(SKJ_HC, k, j) = -x * y / C / D;
        at(SKJ_HF2, k, j) = -xf * y / Cf / D;
        at(SKJ_HF3, k, j) = -x * yf / C / Df;
        double lat, lon;
        latlon(p + 1, x, y, &lat, &lon);
        at(SKJ_SINLAT, k, j) = std::sin(lat);
        at(SKJ_COSLAT, k, j) = std::cos(lat);
        double rr[4], aa[4];
        vectrans(x, y, p + 1, rr, aa);
        const double r00 = rr[0], r01 = rr[1], r10 = rr[2], r11 = rr[3];
        const double a00 = aa[0], a01 = aa[1], a10 = aa[2], a11 = aa[3];
        at(SKJ_R00, k, j) = r00;
        at(SKJ_R01, k, j) = r01;
        at(SKJ_R10, k, j) = r10;
        at(SKJ_R11, k, j) = r11;
        at(SKJ_A00, k, j) = a00;
        at(SKJ_A01, k, j) = a01;
        at(SKJ_A10, k, j) = a10;
        at(SKJ_A11, k, j) = a11;
      }
  }
  // neighbours across each side: two points just outside the side, located
  // on the neighbour; its side is the coordinate nearest +-1 there, the
  // orientation the order of the two points along it
  const int js = NGH, je = NGH + n - 1, ks = NGH, ke = NGH + n - 1;
  for (int p = 0; p < 6; ++p)
    for (int s = 0; s < 4; ++s) {
      const double e = PI_ / 4, o = PI_ / 64;
      double Xq[2], Yq[2];
      int q = -1;
      for (int t = 0; t < 2; ++t) {
        const double along = (t == 0 ? -0.5 : 0.5) * e;
        const double a2 = (s < 2) ? ((s == 0) ? -(e + o) : (e + o)) : along;
        const double a3 = (s < 2) ? along : ((s == 2) ? -(e + o) : (e + o));
        double lat, lon;
        latlon(p + 1, std::tan(a2), std::tan(a3), &lat, &lon);
        q = find_panel(p + 1, lat, lon, &Xq[t], &Yq[t]);
      }
      G.link_q[p][s] = q - 1;
      if (0.5 * (std::fabs(Xq[0]) + std::fabs(Xq[1])) > 0.5 * (std::fabs(Yq[0]) + std::fabs(Yq[1]))) {
        G.link_sq[p][s] = (Xq[0] < 0) ? 0 : 1;
        G.link_inv[p][s] = (Yq[1] < Yq[0]) ? 1 : 0;
      } else {
        G.link_sq[p][s] = (Yq[0] < 0) ? 2 : 3;
        G.link_inv[p][s] = (Xq[1] < Xq[0]) ? 1 : 0;
      }
    }
  // ghost cells of the x2 / x3 sides (InteprolateX2 / X3,
  // cubed_sphere_utility.cpp:102-380): layer m beyond side s lies on row m of
  // the neighbour next to the shared edge
  for (int p = 0; p < 6; ++p)
    for (int s = 0; s < 4; ++s) {
      const int q = G.link_q[p][s], sq = G.link_sq[p][s];
      for (int mm = 0; mm < NGH; ++mm)
        for (int t = 0; t < n; ++t) {
          SphGhost h;
          h.p = p;
          h.q = q;
          if (s == 0) { h.jg = js - 1 - mm; h.kg = ks + t; }
          else if (s == 1) { h.jg = je + 1 + mm; h.kg = ks + t; }
          else if (s == 2) { h.kg = ks - 1 - mm; h.jg = js + t; }
          else { h.kg = ke + 1 + mm; h.jg = js + t; }
          const double Xp = std::tan(av[h.jg]), Yp = std::tan(av[h.kg]);
          double lat, lon, X = 0.0, Y = 0.0;
          latlon(p + 1, Xp, Yp, &lat, &lon);
          xy_on_panel(q + 1, lat, lon, &X, &Y);
          const bool along_k = sq < 2;
          int row;
          if (sq == 0) row = js + mm;
          else if (sq == 1) row = je - mm;
          else if (sq == 2) row = ks + mm;
          else row = ke - mm;
          const double tn = std::tan(av[row]);
          h.yg = row_angle(tn, along_k ? Y : X);
          const int a0 = along_k ? ks : js;
          int lo = 0;
          while (lo < n - 2 && h.yg > row_angle(tn, std::tan(av[a0 + lo + 1]))) ++lo;
          h.y1 = row_angle(tn, std::tan(av[a0 + lo]));
          h.y2 = row_angle(tn, std::tan(av[a0 + lo + 1]));
          if (along_k) {
            h.j1 = h.j2 = row;
            h.k1 = a0 + lo;
            h.k2 = a0 + lo + 1;
          } else {
            h.k1 = h.k2 = row;
            h.j1 = a0 + lo;
            h.j2 = a0 + lo + 1;
          }
          // covariant momenta raised in q's frame (:160-171), carried through
          // (U, V), lowered in p's frame (:183-192)
          const double Cq = std::sqrt(1 + X * X), Dq = std::sqrt(1 + Y * Y);
          h.cq = -X * Y / (Cq * Dq);
          h.sq = std::sqrt(1 + X * X + Y * Y) / (Cq * Dq);
          const double Cp = std::sqrt(1 + Xp * Xp), Dp = std::sqrt(1 + Yp * Yp);
          h.cp = -Xp * Yp / (Cp * Dp);
          double rq[4], aq[4], rp[4], ap[4];
          vectrans(X, Y, q + 1, rq, aq);
          vectrans(Xp, Yp, p + 1, rp, ap);
          h.r00 = rq[0];
          h.r01 = rq[1];
          h.r10 = rq[2];
          h.r11 = rq[3];
          h.a00 = ap[0];
          h.a01 = ap[1];
          h.a10 = ap[2];
          h.a11 = ap[3];
          G.ghosts.push_back(h);
        }
    }
  // panel-edge faces (SynchronizeFluxes, cubed_sphere.cpp:250-565): the
  // neighbour's face point in its own grid, this panel's in its grid
  for (int p = 0; p < 6; ++p)
    for (int s = 0; s < 4; ++s) {
      const int q = G.link_q[p][s];
      for (int t = 0; t < n; ++t) {
        SphEdge ed;
        ed.p = p;
        ed.s = s;
        ed.t = t;
        ed.q = q;
        ed.sq = G.link_sq[p][s];
        ed.tq = G.link_inv[p][s] ? n - 1 - t : t;
        const int fp = (s % 2 == 0) ? 0 : n, fq = (ed.sq % 2 == 0) ? 0 : n;
        double Xp, Yp, Xq, Yq;
        if (s < 2) {
          Xp = std::tan(af[js + fp]);
          Yp = std::tan(av[ks + t]);
        } else {
          Yp = std::tan(af[ks + fp]);
          Xp = std::tan(av[js + t]);
        }
        if (ed.sq < 2) {
          Xq = std::tan(af[js + fq]);
          Yq = std::tan(av[ks + ed.tq]);
        } else {
          Yq = std::tan(af[ks + fq]);
          Xq = std::tan(av[js + ed.tq]);
        }
        double rq[4], aq[4], rp[4], ap[4];
        vectrans(Xq, Yq, q + 1, rq, aq);
        vectrans(Xp, Yp, p + 1, rp, ap);
        ed.r00 = rq[0];
        ed.r01 = rq[1];
        ed.r10 = rq[2];
        ed.r11 = rq[3];
        ed.a00 = ap[0];
        ed.a01 = ap[1];
        ed.a10 = ap[2];
        ed.a11 = ap[3];
        G.edges.push_back(ed);
      }
    }
  return G;
}

}  // namespace cg_host
