// canoe_step.cpp -- C++ host driver over the C-ABI (include/canoe_gpu.h).
//
// The main integration loop of the reference (main.cpp:114-152): cycles of
// the task-list stages (here one SSPRK(3,3) step of the resident block,
// replayed from its hipGraph), time += dt, a new time step, cycle
// diagnostics; at the end the history totals and the zone-cycles report of
// src/mesh_destroy.cpp:38-57.  The block and its initial state come from
// the benchmark configurations (1: dry density current, 2: dry stratified
// box, 3: the cubed sphere of six panels with the Held-Suarez forcing,
// 4: moist box, 5: one seeded block of the uneven-column batch).
//
//   canoe_step <config> [--nx nx1 nx2 nx3] [--nlim N] [--tlim T] [--cfl C]
//              [--fp32] [--no-graph] [--seed S] [--out K] [--hst] [--dbg]
//              [--nc] [--basename NAME] [--implicit-flag F] [--kessler]
//              [--dt DT]
//
// --kessler turns on Kessler warm rain and precipitation sedimentation with
// the rates of examples/2019-Li-snap/earth-precip.yaml:44-53 and
// water-kessler.yaml:4-18 (moist configurations).  Config 3 advances the
// six panels with the fixed step of hs94.inp (--dt, default 60 s) and
// reports the panels' total mass; the other options that need a single
// block (--hst, --dbg, --nc, --cfl) do not apply to it.
//
// The cycles run on the device through cg_run_advance: each step is
// followed by the patched Hydro::NewBlockTimeStep reduced over the resident
// state (patches/21.new_blockdt.patch:16-45) and Mesh::NewTimeStep's limits;
// the host sees one double per cycle and prints the cycle diagnostics every
// K cycles (--out).  At the same points --hst appends Athena++'s history
// columns to NAME.hst and --dbg the per-variable extremes of DebugOutput
// (src/outputs/debug.cpp:137-182) to NAME.dbg, in the reference's text
// layouts with its %15.12e data format, and --nc one record of the
// primitive variables to NAME.prim.nc and of temperature and potential
// temperature (the user output variables of straka.cpp:88-115) to
// NAME.uov.nc, in the layout of the reference's netcdf output.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/canoe_gpu.h"


namespace {

struct Dims {
  int nc1, nc2, nc3, is, js, ks, N;
  size_t sn;
  explicit Dims(const cg_block& b) {
    nc1 = b.nx1 + 2 * CG_NGHOST;
    nc2 = b.nx2 > 1 ? b.nx2 + 2 * CG_NGHOST : 1;
    nc3 = b.nx3 > 1 ? b.nx3 + 2 * CG_NGHOST : 1;
    is = CG_NGHOST;
    js = b.nx2 > 1 ? CG_NGHOST : 0;
    ks = b.nx3 > 1 ? CG_NGHOST : 0;
    N = 5 + b.nmass;
    sn = (size_t)nc1 * nc2 * nc3;
  }
  size_t at(int n, int k, int j, int i) const {
    return (size_t)n * sn + ((size_t)k * nc2 + j) * nc1 + i;
  }
};

void usage() {
  std::fprintf(stderr,
               "usage: canoe_step <config 1|2|4|5> [--nx nx1 nx2 nx3] "
               "[--nlim N] [--tlim T] [--cfl C] [--fp32] "
               "[--no-graph] [--seed S] [--out K] [--hst] [--dbg] "
               "[--nc] [--basename NAME] [--implicit-flag F] [--kessler] "
               "[--dt DT]\n");
}

int check(cg_ctx* ctx, int rc, const char* what) {
  if (rc != CG_OK)
    std::fprintf(stderr, "canoe_step: %s failed (%d): %s\n", what, rc,
                 ctx ? cg_last_error(ctx) : "");
  return rc;
}

// Config 3: the six panels of the cubed sphere (cg_sphere_create) advanced by
// nlim fixed steps of dt, replayed from the run's graph; every out_every
// cycles the density summed over the panels' interior cells.
int run_sphere(const cg_block& blk, const cg_thermo& th, const cg_config& cfg,
               uint64_t seed, long long nlim, long long out_every, double dt,
               bool graph) {
  Dims d(blk);
  cg_sphere sp;
  std::vector<double> w6((size_t)6 * d.N * d.sn);
  if (cg_pgen_sphere(&sp) != CG_OK ||
      cg_pgen_sphere_init(seed, &blk, &th, w6.data()) != CG_OK)
    return 1;
  cg_ctx* ctx = nullptr;
  int rc = cg_create(0, &ctx);
  if (rc != CG_OK) {
    std::fprintf(stderr, "canoe_step: no HIP device (%d)\n", rc);
    return 1;
  }
  cg_run* run = nullptr;
  int err = check(ctx, cg_sphere_create(ctx, &blk, &th, &cfg, &sp, &run),
                  "cg_sphere_create");
  for (int p = 0; p < 6 && !err; ++p)
    err = check(ctx, cg_run_block_set_prim_host(run, p, &w6[(size_t)p * d.N * d.sn]),
                "set_prim");
  if (err) {
    if (run) cg_run_destroy(run);
    cg_destroy(ctx);
    return 1;
  }
  std::printf("canoe_step: config 3, 6 panels of %d x %d cells and %d levels, %s, %s\n",
              blk.nx2, blk.nx3, blk.nx1,
              cfg.precision == CG_PREC_FP32 ? "FP32 x2/x3 faces" : "FP64", cg_version());
  std::vector<double> u((size_t)d.N * d.sn);
  auto mass = [&](double* out) -> int {
    double m = 0.0;
    for (int p = 0; p < 6; ++p) {
      if (check(ctx, cg_run_block_get_cons_host(run, p, u.data()), "get_cons")) return 1;
      for (int k = d.ks; k < d.ks + blk.nx3; ++k)
        for (int j = d.js; j < d.js + blk.nx2; ++j)
          for (int i = d.is; i < d.is + blk.nx1; ++i) m += u[d.at(0, k, j, i)];
    }
    *out = m;
    return 0;
  };
  double m0 = 0.0, m1 = 0.0;
  err = mass(&m0);
  long long ncycle = 0;
  auto t0 = std::chrono::steady_clock::now();
  while (!err && ncycle < nlim) {
    std::printf("cycle=%lld time=%.6e dt=%.6e\n", ncycle, ncycle * dt, dt);
    const long long n = std::min(out_every, nlim - ncycle);
    if ((err = check(ctx, cg_run_step(run, dt, (int)n, graph ? 1 : 0), "cg_run_step"))) break;
    ncycle += n;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (!err) err = mass(&m1);
  const double wall = std::chrono::duration<double>(t1 - t0).count();
  const double zc = 6.0 * blk.nx1 * blk.nx2 * blk.nx3 * (double)ncycle;
  std::printf("cycle=%lld time=%.6e dt=%.6e\n", ncycle, ncycle * dt, dt);
  std::printf("panel density sum (initial -> final): %.16e -> %.16e\n", m0, m1);
  std::printf("zone-cycles = %.0f\n", zc);
  std::printf("wall time = %.6f s\n", wall);
  std::printf("zone-cycles/wall_second = %.6e\n", wall > 0 ? zc / wall : 0.0);
  cg_run_destroy(run);
  cg_destroy(ctx);
  return err ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    usage();
    return 2;
  }
  int config = std::atoi(argv[1]);
  int nx[3] = {0, 0, 0};
  long long nlim = 10, out_every = 1;
  double tlim = 1e300, cfl = 0.9;
  bool fp32 = false, graph = true, hst = false, dbg = false, nc = false;
  int iflag = -1;  // <hydro> implicit_flag (implicit_solver.cpp:47)
  bool kessler = false;
  double fixed_dt = 60.0;  // config 3: hs94.inp's step
  std::string base = "canoe";
  uint64_t seed = 1;
  for (int a = 2; a < argc; ++a) {
    std::string s = argv[a];
    auto need = [&](int n) {
      if (a + n >= argc) {
        usage();
        std::exit(2);
      }
    };
    if (s == "--nx") {
      need(3);
      for (int q = 0; q < 3; ++q) nx[q] = std::atoi(argv[++a]);
    } else if (s == "--nlim") {
      need(1);
      nlim = std::atoll(argv[++a]);
    } else if (s == "--tlim") {
      need(1);
      tlim = std::atof(argv[++a]);
    } else if (s == "--cfl") {
      need(1);
      cfl = std::atof(argv[++a]);
    } else if (s == "--out") {
      need(1);
      out_every = std::max(1LL, std::atoll(argv[++a]));
    } else if (s == "--seed") {
      need(1);
      seed = std::strtoull(argv[++a], nullptr, 10);
    } else if (s == "--hst") {
      hst = true;
    } else if (s == "--dbg") {
      dbg = true;
    } else if (s == "--nc") {
      nc = true;
    } else if (s == "--basename") {
      need(1);
      base = argv[++a];
    } else if (s == "--fp32") {
      fp32 = true;
    } else if (s == "--no-graph") {
      graph = false;
    } else if (s == "--implicit-flag") {
      need(1);
      iflag = std::atoi(argv[++a]);
    } else if (s == "--kessler") {
      kessler = true;
    } else if (s == "--dt") {
      need(1);
      fixed_dt = std::atof(argv[++a]);
    } else {
      usage();
      return 2;
    }
  }
  cg_block blk;
  cg_thermo th;
  cg_config cfg;
  if (cg_pgen_setup(config, nx[0], nx[1], nx[2], &blk, &th, &cfg) != CG_OK) {
    std::fprintf(stderr, "canoe_step: unknown configuration %d\n", config);
    return 2;
  }
  if (fp32) cfg.precision = CG_PREC_FP32;
  if (iflag >= 0) cfg.implicit_flag = iflag;
  if (kessler) {
    cfg.k_auto = 0.01;
    cfg.k_accr = 0.02;
    cfg.k_evap = 0.1;
    cfg.v_fall = 8.0;
  }
  if (config == 3) return run_sphere(blk, th, cfg, seed, nlim, out_every, fixed_dt, graph);
  Dims d(blk);
  std::vector<double> w((size_t)d.N * d.sn);
  if (cg_pgen_init(config, seed, &blk, &th, &cfg, w.data()) != CG_OK) return 1;

  cg_ctx* ctx = nullptr;
  int rc = cg_create(0, &ctx);
  if (rc != CG_OK) {
    std::fprintf(stderr, "canoe_step: no HIP device (%d)\n", rc);
    return 1;
  }
  cg_run* run = nullptr;
  if (check(ctx, cg_run_create(ctx, &blk, &th, &cfg, &run), "cg_run_create") ||
      check(ctx, cg_run_set_prim_host(run, w.data()), "set_prim")) {
    cg_destroy(ctx);
    return 1;
  }
  const long long cells = (long long)blk.nx1 * blk.nx2 * blk.nx3;
  std::printf("canoe_step: config %d, %d x %d x %d cells (x1 vertical), %s, %s\n",
              config, blk.nx1, blk.nx2, blk.nx3,
              fp32 ? "FP32 fluxes + FP64 state" : "FP64", cg_version());
  double time = 0.0, dt = 0.0;
  long long ncycle = 0;
  std::vector<double> sums(d.N), sums0(d.N);
  check(ctx, cg_run_sums(run, sums0.data()), "sums");
  // the first step: the blocks' NewBlockTimeStep, cut to end at tlim
  int err = check(ctx, cg_run_new_dt(run, cfl, &dt), "cg_run_new_dt");
  if (!err && tlim - time < dt) dt = tlim - time;
  FILE* fh = hst ? std::fopen((base + ".hst").c_str(), "w") : nullptr;
  FILE* fd = dbg ? std::fopen((base + ".dbg").c_str(), "w") : nullptr;
  if ((hst && !fh) || (dbg && !fd)) {
    std::fprintf(stderr, "canoe_step: cannot open the output files\n");
    return 1;
  }
  const char* fmt = " %15.12e";
  if (fh) {
    const char* cols[] = {"time", "dt", "mass", "1-mom", "2-mom", "3-mom",
                          "1-KE", "2-KE", "3-KE", "tot-E"};
    std::fprintf(fh, "# Athena++ history data\n#");
    int c = 1;
    for (const char* nm : cols) std::fprintf(fh, " [%d]=%-8s", c++, nm);
    for (int n = 1; n <= blk.nmass; ++n) std::fprintf(fh, " [%d]=%d-mass  ", c++, n);
    std::fprintf(fh, "\n");
  }
  if (fd) {
    std::fprintf(fd, "# Athena++ debug data\n# ");
    const char* cols[] = {"time", "id", "dmin", "dmax", "x1min", "x1max",
                          "x2min", "x2max", "x3min", "x3max"};
    int c = 1;
    for (const char* nm : cols) std::fprintf(fd, "[%d]=%-8s", c++, nm);
    std::fprintf(fd, "\n");
  }
  std::vector<double> hv(8 + blk.nmass), vmin(d.N), vmax(d.N), xmin(3 * d.N),
      xmax(3 * d.N);
  // NetCDF dumps: prim (rho, press, vel1..3, species) and uov (temp, theta)
  std::vector<double> wp, tt, th_;
  int nc_records = 0;
  auto nc_dump = [&]() -> int {
    wp.resize((size_t)d.N * d.sn);
    tt.assign(d.sn, 0.0);
    th_.assign(d.sn, 0.0);
    if (check(ctx, cg_run_get_prim_host(run, wp.data()), "get_prim")) return 1;
    const int IVX = 1 + blk.nmass, IPR = IVX + 3;
    const double chi = (th.gammad - 1.) / th.gammad;
    for (int k = d.ks; k < d.ks + blk.nx3; ++k)
      for (int j = d.js; j < d.js + blk.nx2; ++j)
        for (int i = d.is; i < d.is + blk.nx1; ++i) {
          // T = p / (rho Rd feps), feps = 1 + sum_vap q inv_mu_ratio_m1
          // - sum_cld q; theta = T (p0 / p)^((gamma_d - 1) / gamma_d)
          double feps = 1.;
          for (int n = 1; n <= blk.nmass; ++n) {
            double q = wp[d.at(n, k, j, i)];
            feps += (n <= th.nvapor) ? q * th.inv_mu_ratio_m1[n - 1] : -q;
          }
          const double p = wp[d.at(IPR, k, j, i)], rho = wp[d.at(0, k, j, i)];
          const double T = p / (rho * th.rd * feps);
          const size_t c = d.at(0, k, j, i);
          tt[c] = T;
          th_[c] = T * std::pow(cfg.p0 / p, chi);
        }
    std::vector<std::string> nm = {"rho", "press", "vel1", "vel2", "vel3"};
    std::vector<std::string> un = {"kg/m^3", "pa", "m/s", "m/s", "m/s"};
    std::vector<std::string> ln = {"density", "pressure", "vertical velocity",
                                   "horizontal velocity", "horizontal velocity"};
    std::vector<const double*> fp = {&wp[d.at(0, 0, 0, 0)], &wp[d.at(IPR, 0, 0, 0)],
                                     &wp[d.at(IVX, 0, 0, 0)], &wp[d.at(IVX + 1, 0, 0, 0)],
                                     &wp[d.at(IVX + 2, 0, 0, 0)]};
    for (int n = 1; n <= blk.nmass; ++n) {
      nm.push_back("q" + std::to_string(n));
      un.push_back("kg/kg");
      ln.push_back("mass mixing ratio of species " + std::to_string(n));
      fp.push_back(&wp[d.at(n, 0, 0, 0)]);
    }
    std::vector<const char*> a, b2, c2;
    for (size_t q = 0; q < nm.size(); ++q) {
      a.push_back(nm[q].c_str());
      b2.push_back(un[q].c_str());
      c2.push_back(ln[q].c_str());
    }
    const char* un_names[2] = {"temp", "theta"};
    const char* un_units[2] = {"K", "K"};
    const char* un_long[2] = {"temperature", "potential temperature"};
    const double* un_f[2] = {tt.data(), th_.data()};
    int e = cg_netcdf_write((base + ".prim.nc").c_str(), &blk, time, (int)nm.size(),
                            a.data(), b2.data(), c2.data(), fp.data(), nc_records > 0);
    if (!e)
      e = cg_netcdf_write((base + ".uov.nc").c_str(), &blk, time, 2, un_names, un_units,
                          un_long, un_f, nc_records > 0);
    if (e) {
      std::fprintf(stderr, "canoe_step: netcdf output failed (%d)\n", e);
      return e;
    }
    ++nc_records;
    return 0;
  };
  auto outputs = [&]() {
    if (nc && nc_dump()) return;
    if (fh && !check(ctx, cg_run_history(run, hv.data()), "cg_run_history")) {
      std::fprintf(fh, fmt, time);
      std::fprintf(fh, fmt, dt);
      for (double v : hv) std::fprintf(fh, fmt, v);
      std::fprintf(fh, "\n");
    }
    if (fd && !check(ctx, cg_run_debug(run, vmin.data(), vmax.data(), xmin.data(),
                                       xmax.data()),
                     "cg_run_debug")) {
      for (int n = 0; n < d.N; ++n) {
        if (n == 0)
          std::fprintf(fd, fmt, time);
        else
          std::fprintf(fd, "                ");
        std::fprintf(fd, " --  %2d  -- ", n);
        std::fprintf(fd, fmt, vmin[n]);
        std::fprintf(fd, fmt, vmax[n]);
        for (int q = 0; q < 3; ++q) {
          std::fprintf(fd, fmt, xmin[3 * n + q]);
          std::fprintf(fd, fmt, xmax[3 * n + q]);
        }
        std::fprintf(fd, "\n");
      }
    }
  };
  auto t0 = std::chrono::steady_clock::now();
  while (!err && time < tlim && ncycle < nlim) {
    std::printf("cycle=%lld time=%.6e dt=%.6e\n", ncycle, time, dt);
    outputs();
    const long long upto = std::min(nlim, ncycle + out_every);
    if ((err = check(ctx, cg_run_advance(run, cfl, tlim, upto, graph ? 1 : 0, &time,
                                         &dt, &ncycle),
                     "cg_run_advance")))
      break;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (!err) outputs();
  if (fh) std::fclose(fh);
  if (fd) std::fclose(fd);
  double wall = std::chrono::duration<double>(t1 - t0).count();
  check(ctx, cg_run_sums(run, sums.data()), "sums");
  std::printf("cycle=%lld time=%.6e dt=%.6e\n", ncycle, time, dt);
  std::printf("history totals (initial -> final):\n");
  for (int n = 0; n < d.N; ++n)
    std::printf("  var %d: %.16e -> %.16e\n", n, sums0[n], sums[n]);
  const double zc = (double)cells * ncycle;
  std::printf("zone-cycles = %.0f\n", zc);
  std::printf("wall time = %.6f s\n", wall);
  std::printf("zone-cycles/wall_second = %.6e\n", wall > 0 ? zc / wall : 0.0);
  cg_run_destroy(run);
  cg_destroy(ctx);
  return err ? 1 : 0;
}
