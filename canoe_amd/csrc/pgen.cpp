// pgen.cpp -- problem generators of the benchmark configurations (host).
//
// The configurations (cg_pgen_setup) and their seeded initial states
// (cg_pgen_init) play the role of the reference's problem generators
// (examples/2019-Li-snap/straka.cpp:220-268, robert.cpp:72-114) and input
// files (straka.inp, robert3d.inp).  Random numbers come from a counter-based
// splitmix64 hash of (seed, global cell index), so a state does not depend
// on traversal order or thread count.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/canoe_gpu.h"

namespace {

constexpr int NG = CG_NGHOST;

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// uniform in [0, 1)
double urand(uint64_t seed, uint64_t idx) {
  uint64_t h = splitmix64(seed ^ splitmix64(idx + 0x632BE59BD9B4E019ull));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

struct Dims {
  int nc1, nc2, nc3, is, js, ks, N, IVX, IPR;
  size_t sn;
  explicit Dims(const cg_block& b) {
    nc1 = b.nx1 + 2 * NG;
    nc2 = b.nx2 > 1 ? b.nx2 + 2 * NG : 1;
    nc3 = b.nx3 > 1 ? b.nx3 + 2 * NG : 1;
    is = NG;
    js = b.nx2 > 1 ? NG : 0;
    ks = b.nx3 > 1 ? NG : 0;
    N = 5 + b.nmass;
    IVX = 1 + b.nmass;
    IPR = IVX + 3;
    sn = (size_t)nc1 * nc2 * nc3;
  }
  size_t at(int n, int k, int j, int i) const {
    return ((size_t)n * nc3 + k) * nc2 * nc1 + (size_t)j * nc1 + i;
  }
};

void dry_thermo(cg_thermo* th, double gamma, double rd) {
  memset(th, 0, sizeof(*th));
  th->nmass = 0;
  th->nvapor = 0;
  th->gammad = gamma;
  th->rd = rd;
  th->tol = 1e-10;
  th->max_iter = 100;
}

// Config-4 moist thermodynamics: vapour, cloud liquid, precipitation.
// Constant-L Clausius-Clapeyron (beta = L / (Rv T3), delta = 0) with the
// condensate heat capacity equal to cp of vapour, which is what keeps the
// latent heat temperature independent (d L / d T = cp_v - c_l = 0).
void moist_thermo(cg_thermo* th) {
  const double rd = 287.0, rv = 461.5, gamma = 1.4;
  const double cvd = rd / (gamma - 1.0);
  const double cpv = 1870.0, cvv = cpv - rv, cl = cpv;
  memset(th, 0, sizeof(*th));
  th->nmass = 3;
  th->nvapor = 1;
  th->gammad = gamma;
  th->rd = rd;
  th->cv_ratio_m1[0] = cvv / cvd - 1.0;
  th->cv_ratio_m1[1] = cl / cvd - 1.0;
  th->cv_ratio_m1[2] = cl / cvd - 1.0;
  th->inv_mu_ratio_m1[0] = rv / rd - 1.0;
  th->inv_mu_ratio_m1[1] = 0.0;
  th->inv_mu_ratio_m1[2] = 0.0;
  th->ivap = 1;
  th->icld = 2;
  th->t3 = 273.16;
  th->p3 = 611.7;
  th->latent = 2.5e6;
  th->beta = th->latent / (rv * th->t3);
  th->delta = 0.0;
  th->tol = 1e-10;
  th->max_iter = 100;
}

// Hydrostatic profile of the stratified boxes: T falls `lapse` K/m from Ts
// to an isothermal cap Tt (config 2 / 4: 6.5 K/km, 300 K, 220 K); p from the
// analytic integral.
void box_profile(double z, double g, double rd, double* T, double* p,
                 double Ts = 300.0, double lapse = 6.5e-3, double Tt = 220.0) {
  const double ps = 1e5;
  const double zt = (Ts - Tt) / lapse;
  if (z <= zt) {
    *T = Ts - lapse * z;
    *p = ps * std::pow(*T / Ts, g / (rd * lapse));
  } else {
    double pt = ps * std::pow(Tt / Ts, g / (rd * lapse));
    *T = Tt;
    *p = pt * std::exp(-g * (z - zt) / (rd * Tt));
  }
}

double qsat_host(const cg_thermo& th, double T, double rho) {
  double rv = th.rd * (1.0 + th.inv_mu_ratio_m1[th.ivap - 1]);
  double es = th.p3 * std::exp((1.0 - th.t3 / T) * th.beta -
                               th.delta * std::log(T / th.t3));
  return es / (rho * rv * T);
}

}  // namespace

extern "C" {

// Ghost cells of a host primitive array (reflecting / periodic / outflow),
// same order as the device boundary kernels: x1, then x2, then x3.
int cg_apply_bc_host(const cg_block* blk, double* w) {
  if (!blk || !w) return CG_ERR_ARG;
  Dims d(*blk);
  const int ie = d.is + blk->nx1 - 1, je = d.js + blk->nx2 - 1,
            ke = d.ks + blk->nx3 - 1;
  auto fill = [&](auto idx, int lo, int hi, int bci, int bco, int vn) {
    for (int n = 0; n < d.N; ++n) {
      double sgn = (n == vn) ? -1.0 : 1.0;
      for (int gh = 1; gh <= NG; ++gh) {
        double vin, vout;
        if (bci == CG_BC_REFLECT)
          vin = sgn * w[idx(n, lo + gh - 1)];
        else if (bci == CG_BC_PERIODIC)
          vin = w[idx(n, hi - gh + 1)];
        else
          vin = w[idx(n, lo)];
        if (bco == CG_BC_REFLECT)
          vout = sgn * w[idx(n, hi - gh + 1)];
        else if (bco == CG_BC_PERIODIC)
          vout = w[idx(n, lo + gh - 1)];
        else
          vout = w[idx(n, hi)];
        w[idx(n, lo - gh)] = vin;
        w[idx(n, hi + gh)] = vout;
      }
    }
  };
  for (int k = d.ks; k <= ke; ++k)
    for (int j = d.js; j <= je; ++j)
      fill([&](int n, int c) { return d.at(n, k, j, c); }, d.is, ie, blk->bc[0],
           blk->bc[1], d.IVX);
  if (blk->nx2 > 1)
    for (int k = d.ks; k <= ke; ++k)
      for (int i = 0; i < d.nc1; ++i)
        fill([&](int n, int c) { return d.at(n, k, c, i); }, d.js, je,
             blk->bc[2], blk->bc[3], d.IVX + 1);
  if (blk->nx3 > 1)
    for (int j = 0; j < d.nc2; ++j)
      for (int i = 0; i < d.nc1; ++i)
        fill([&](int n, int c) { return d.at(n, c, j, i); }, d.ks, ke,
             blk->bc[4], blk->bc[5], d.IVX + 2);
  return CG_OK;
}

// config: 1 = 2D dry density current (straka), 2 = 3D dry stratified box,
// 4 = moist box, 5 = one block of the uneven-column batch (the dry box's
// numerics; its profile is drawn from the seed at cg_pgen_init).  nx* = 0
// selects the configuration's own size; other sizes keep the grid spacing
// (the domain shrinks or grows).
int cg_pgen_setup(int config, int nx1, int nx2, int nx3, cg_block* blk,
                  cg_thermo* th, cg_config* cfg) {
  if (!blk || !th || !cfg) return CG_ERR_ARG;
  memset(blk, 0, sizeof(*blk));
  memset(cfg, 0, sizeof(*cfg));
  if (config == 1) {
    // straka.inp / straka.cpp, 512 x 64 over 51.2 km x 6.4 km
    blk->nx1 = nx1 > 0 ? nx1 : 64;
    blk->nx2 = nx2 > 0 ? nx2 : 512;
    blk->nx3 = 1;
    blk->dx1 = 100.0;
    blk->dx2 = 100.0;
    blk->dx3 = 1.0;
    blk->x1min = 0.0;
    blk->x2min = -0.5 * blk->nx2 * blk->dx2;
    blk->x3min = -0.5;
    blk->grav = -9.8;
    for (int s = 0; s < 4; ++s) blk->bc[s] = CG_BC_REFLECT;
    blk->bc[4] = blk->bc[5] = CG_BC_PERIODIC;
    dry_thermo(th, 1.4, 287.0);
    cfg->order_h = 2;
    cfg->order_v = 2;
    cfg->rsolver = CG_RS_LMARS;
    cfg->implicit_flag = 0;
    cfg->visc = 75.0;
    cfg->p0 = 1e5;
  } else if (config == 2 || config == 4 || config == 5) {
    blk->nx1 = nx1 > 0 ? nx1 : 128;
    blk->nx2 = nx2 > 0 ? nx2 : 256;
    blk->nx3 = nx3 > 0 ? nx3 : 256;
    blk->dx1 = 100.0;
    blk->dx2 = 200.0;
    blk->dx3 = 200.0;
    blk->grav = -9.8;
    blk->bc[0] = blk->bc[1] = CG_BC_REFLECT;
    blk->bc[2] = blk->bc[3] = blk->bc[4] = blk->bc[5] = CG_BC_PERIODIC;
    if (config == 4)
      moist_thermo(th);
    else
      dry_thermo(th, 1.4, 287.0);
    cfg->order_h = 5;
    cfg->order_v = 2;
    cfg->rsolver = CG_RS_HLLC;
    cfg->implicit_flag = 9;
    cfg->heat_flux = 50.0;
    cfg->p0 = 1e5;
    cfg->satadj = (config == 4) ? CG_SATADJ_NEWTON : CG_SATADJ_OFF;
  } else if (config == 3) {
    // one panel of the cubed sphere (hs94.inp: R = 6.371e6 m, 25 km deep,
    // g = 9.81; six panels of 96 x 96 x 64)
    blk->nx1 = nx1 > 0 ? nx1 : 64;
    blk->nx2 = nx2 > 0 ? nx2 : 96;
    blk->nx3 = nx3 > 0 ? nx3 : blk->nx2;
    blk->x1min = 6.371e6;
    blk->dx1 = 25000.0 / blk->nx1;
    blk->dx2 = M_PI / (2.0 * blk->nx2);
    blk->dx3 = M_PI / (2.0 * blk->nx3);
    blk->x2min = -0.25 * M_PI;
    blk->x3min = -0.25 * M_PI;
    blk->grav = -9.81;
    blk->bc[0] = blk->bc[1] = CG_BC_REFLECT;
    blk->bc[2] = blk->bc[3] = blk->bc[4] = blk->bc[5] = CG_BC_PANEL;
    dry_thermo(th, 1.4, 287.0);
    cfg->order_h = 5;
    cfg->order_v = 5;
    cfg->rsolver = CG_RS_HLLC;
    cfg->implicit_flag = 1;  // hs94.inp
    cfg->p0 = 1e5;
  } else {
    return CG_ERR_ARG;
  }
  blk->nmass = th->nmass;
  cfg->precision = CG_PREC_FP64;
  return CG_OK;
}

// Rotation and Held-Suarez parameters of examples/2023-Chen-exo3/hs94.inp
// (Omega, Ts, dT, dtheta, p0, Kf, Ka, Ks in 1/day, sigmab).
int cg_pgen_sphere(cg_sphere* sp) {
  if (!sp) return CG_ERR_ARG;
  const double day = 8.64e4;
  sp->omega = 7.292e-5;
  sp->hs94 = 1;
  sp->ts = 315.0;
  sp->dt_eq = 60.0;
  sp->dtheta = 10.0;
  sp->p0 = 1e5;
  sp->kf = 1.0 / day;
  sp->ka = 0.025 / day;
  sp->ks = 0.25 / day;
  sp->sigmab = 0.7;
  return CG_OK;
}

// Isothermal 300 K atmosphere at rest, p = p0 exp(-g z / (Rd T0)) at the
// cell centres, plus seeded Gaussian 0.1 K temperature noise (Box-Muller of
// two uniforms per cell and panel).
int cg_pgen_sphere_init(uint64_t seed, const cg_block* blk, const cg_thermo* th,
                        double* w6) {
  if (!blk || !th || !w6) return CG_ERR_ARG;
  Dims d(*blk);
  const size_t ps = (size_t)d.N * d.sn;
  memset(w6, 0, sizeof(double) * ps * 6);
  const double T0 = 300.0, p0 = 1e5, g = -blk->grav, Rd = th->rd;
  for (int p = 0; p < 6; ++p) {
    double* w = w6 + p * ps;
    for (int k = d.ks; k < d.ks + blk->nx3; ++k)
      for (int j = d.js; j < d.js + blk->nx2; ++j)
        for (int i = d.is; i < d.is + blk->nx1; ++i) {
          const double z = (i - d.is + 0.5) * blk->dx1;
          const double pr = p0 * std::exp(-g * z / (Rd * T0));
          const uint64_t idx =
              ((((uint64_t)p * blk->nx3 + (k - d.ks)) * blk->nx2 + (j - d.js)) * blk->nx1 +
               (i - d.is)) * 2;
          const double u1 = urand(seed, idx), u2 = urand(seed, idx + 1);
          const double gz = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
          const double T = T0 + 0.1 * gz;
          w[d.at(0, k, j, i)] = pr / (Rd * T);
          w[d.at(d.IPR, k, j, i)] = pr;
        }
  }
  return CG_OK;
}

// Seeded primitive state (interior and ghosts) for a configuration.
int cg_pgen_init(int config, uint64_t seed, const cg_block* blk,
                 const cg_thermo* th, const cg_config* cfg, double* w) {
  if (!blk || !th || !cfg || !w) return CG_ERR_ARG;
  Dims d(*blk);
  memset(w, 0, sizeof(double) * d.N * d.sn);
  const double g = -blk->grav;
  const int ie = d.is + blk->nx1 - 1, je = d.js + blk->nx2 - 1,
            ke = d.ks + blk->nx3 - 1;
  if (config == 1) {
    // straka.cpp:220-268: Ts = 300, p0 = 1e5, cold bubble dT = -15 K at
    // (xc, zc) = (0, 3 km), radii (4 km, 2 km)
    const double Ts = 300.0, p0 = 1e5, dT = -15.0;
    const double xc = 0.0, xr = 4e3, zc = 3e3, zr = 2e3;
    const double Rd = th->rd, cp = th->gammad / (th->gammad - 1.0) * Rd;
    for (int k = d.ks; k <= ke; ++k)
      for (int j = d.js; j <= je; ++j)
        for (int i = d.is; i <= ie; ++i) {
          double x1 = blk->x1min + (i - d.is + 0.5) * blk->dx1;
          double x2 = blk->x2min + (j - d.js + 0.5) * blk->dx2;
          double L = std::sqrt(((x2 - xc) / xr) * ((x2 - xc) / xr) +
                               ((x1 - zc) / zr) * ((x1 - zc) / zr));
          double temp = Ts - g * x1 / cp;
          double p = p0 * std::pow(temp / Ts, cp / Rd);
          if (L <= 1.) temp += dT * (std::cos(M_PI * L) + 1.) / 2.;
          w[d.at(d.IPR, k, j, i)] = p;
          w[d.at(0, k, j, i)] = p / (Rd * temp);
        }
  } else if (config == 2 || config == 4 || config == 5) {
    const double Rd = th->rd, p0 = cfg->p0 > 0 ? cfg->p0 : 1e5;
    const double chi = (th->gammad - 1.0) / th->gammad;
    // config 5: surface temperature 250..320 K, lapse rate 4..8 K/km,
    // isothermal above 200 K
    double Ts = 300.0, lapse = 6.5e-3, Tt = 220.0;
    if (config == 5) {
      Ts = 250.0 + 70.0 * urand(seed, 0xFFFFFFFF00000001ull);
      lapse = 4e-3 + 4e-3 * urand(seed, 0xFFFFFFFF00000002ull);
      Tt = 200.0;
    }
    for (int k = d.ks; k <= ke; ++k)
      for (int j = d.js; j <= je; ++j)
        for (int i = d.is; i <= ie; ++i) {
          double z = blk->x1min + (i - d.is + 0.5) * blk->dx1;
          double T, p;
          box_profile(z, g, Rd, &T, &p, Ts, lapse, Tt);
          if (z < 1000.0) {
            uint64_t idx = ((uint64_t)(k - d.ks) * blk->nx2 + (j - d.js)) *
                               blk->nx1 + (i - d.is);
            double dth = urand(seed, idx) - 0.5;  // +-0.5 K potential temp.
            T += dth * std::pow(p / p0, chi);
          }
          double rho = p / (Rd * T);
          if (config == 4) {
            double rv = Rd * (1.0 + th->inv_mu_ratio_m1[0]);
            double q = 0.016 * std::exp(-z / 2500.0);
            for (int it = 0; it < 4; ++it) {
              double qs = qsat_host(*th, T, rho);
              double qv = std::fmin(0.016 * std::exp(-z / 2500.0), 0.95 * qs);
              q = qv;
              rho = p / (T * (Rd * (1.0 - q) + rv * q));
            }
            w[d.at(1, k, j, i)] = q;
          }
          w[d.at(0, k, j, i)] = rho;
          w[d.at(d.IPR, k, j, i)] = p;
        }
  } else {
    return CG_ERR_ARG;
  }
  return cg_apply_bc_host(blk, w);
}

}  // extern "C"
