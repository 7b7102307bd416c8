// run.hip -- resident runs: one MeshBlock (cg_run_create) or a batch of
// independent MeshBlocks of any sizes (cg_batch_create) kept on the device
// and advanced by whole SSPRK(3,3) steps of ImplicitHydroTasks
// (src/tasklist/implicit_hydro_tasks.cpp:134-313,
//  src/integrator/integrators.cpp:14-44), replayed from one hipGraph.
//
// Per stage: each block's flux sweep and flux divergence (per-block launches;
// the mixed form computes the faces in FP32), then ONE batched launch of the
// implicit forward sweep over every column of every block and ONE batched
// launch of the downward sweep (back substitution + stage update + x1 ghosts
// + next stage's hydrostatic pressures), then each block's x2 / x3 ghosts.
// Launch shapes (waves per workgroup and grid of the batched column solve,
// workgroup size of the face kernels) come from cg_run_autotune.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/canoe_gpu.h"
#include "launch.hpp"
#include "sphere.hpp"
#include "sphere_geom.hpp"

using namespace cg;
using namespace cg_host;

namespace {

struct RunBlock {
  cg_block blk;
  KGrid g;   // level-major resident state
  KGrid ga;  // AthenaArray layout of host transfers
  double *u = nullptr, *w = nullptr, *u1 = nullptr, *du = nullptr;
  double *f1 = nullptr, *f2 = nullptr, *f3 = nullptr;
  double *psf = nullptr, *psv = nullptr;
  float* ff = nullptr;     // FP32 face fluxes of the mixed form
  double* pref = nullptr;  // per-level reference pressure of the mixed form
};

}  // namespace

// Device state of a cubed-sphere run (six panel blocks; sphere.hpp).
struct SphRun {
  cg_sphere sp;
  KSph k;                        // kernel arguments (panel pointers, tables)
  KSphPtrs* d_ptr = nullptr;     // panel pointers for the table kernels
  double* kj = nullptr;          // six panels' (k, j) tables
  double* x1 = nullptr;          // x1f (nc1 + 1) then x1v (nc1)
  double* lr = nullptr;          // face states
  double* arch = nullptr;        // column archive, arch_panel doubles per panel
  long long arch_panel = 0;
  SphGhost* d_ghosts = nullptr;
  int nghost = 0;
  SphEdge* d_edges = nullptr;
  int nedge = 0;
  bool dirty = true;             // ghosts and hydrostatic pressures to refresh
  int x1seg = 8;                 // faces per lane of the FP32 x1 march
};

struct cg_run {
  cg_ctx* ctx = nullptr;
  cg_thermo th;
  cg_config cfg;
  KThermo kt;
  int N = 5;
  int AS = 30;  // archive doubles per level (B*B + B)
  std::vector<RunBlock> blocks;
  double* arch = nullptr;
  double* lines = nullptr;  // scratch of the line solves (x2 / x3, periodic x1)
  bool x1_periodic = false; // x1 columns solved as periodic lines, not fused
  KBlockPtr* d_bp = nullptr;
  KGroup* d_grp = nullptr;
  int ngroups = 0;
  cg_tune tune;
  int ncu = 256;
  double* part = nullptr;
  const cg_mlp* mlp = nullptr;  // network of the guided adjustment
  bool kes = false;             // Kessler microphysics after the last stage
  KKessler kz;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // d_dt[0]: the step every stage kernel reads (a graph replays with any
  // step); d_dt[1]: the last NewBlockTimeStep minimum; d_dt[2]: the state
  // checks' first failing cell (64-bit key, all ones when none); h_dt:
  // pinned copy of d_dt[1..2].
  double* d_dt = nullptr;
  double* h_dt = nullptr;
  double* dt_part = nullptr;  // per-workgroup minima, dt_parts per block
  int dt_parts = 1;
  // tiles of the fused face kernel over every 3D block (FP32-flux form,
  // face variants 3 / 4), built for (tile_variant, tile_seg)
  KTile* d_tiles = nullptr;
  int ntiles = 0, tile_variant = -1, tile_seg = -1;
  unsigned bc2_grid = 1, bc3_grid = 1;  // workgroups per block of the ghost kernels
  KTimer tm;
  SphRun* sph = nullptr;  // cubed-sphere run (cg_sphere_create)
};

namespace {

void drop_graph(cg_run* r) {
  if (r->gexec) (void)hipGraphExecDestroy(r->gexec);
  if (r->graph) (void)hipGraphDestroy(r->graph);
  r->gexec = nullptr;
  r->graph = nullptr;
}

void free_block(RunBlock& b) {
  double* ps[] = {b.u, b.w, b.u1, b.du, b.f1, b.f2, b.f3, b.psf, b.psv, b.pref};
  for (double* p : ps)
    if (p) (void)hipFree(p);
  if (b.ff) (void)hipFree(b.ff);
  b = RunBlock();
}

// Face variants 3 / 4 of the FP32-flux form: one tile launch per stage.
bool fused_tiles(const cg_run* r) {
  return r->cfg.precision == CG_PREC_FP32 &&
         (r->tune.face_variant == 3 || r->tune.face_variant == 4);
}

unsigned implicit_grid(const cg_run* r) {
  const int wpb = r->tune.implicit_wpb;
  long long full = (r->ngroups + wpb - 1) / wpb;
  long long g = r->tune.implicit_grid > 0 ? r->tune.implicit_grid : full;
  return (unsigned)std::max(1LL, std::min(g, full));
}

// beta: the stage weight; the sweep's step is beta times the device step.
template <int NM, int B, int WPB, bool KES>
void launch_col_batch(cg_run* r, double beta, const KUpdate& up, KTimer* tm) {
  hipStream_t s = r->ctx->stream;
  const unsigned gr = implicit_grid(r);
  if (r->tune.implicit_fused) {
    if constexpr (KES)
      TIMED(tm, KID_IMPLICIT_FUSED,
            (k_implicit_fused_kessler_batch<NM, B, WPB><<<gr, 64 * WPB, 0, s>>>(
                r->d_bp, r->d_grp, r->ngroups, r->kt, beta, r->d_dt, up, r->arch,
                r->ctx->d_status)));
    else
      TIMED(tm, KID_IMPLICIT_FUSED,
            (k_implicit_fused_batch<NM, B, WPB><<<gr, 64 * WPB, 0, s>>>(
                r->d_bp, r->d_grp, r->ngroups, r->kt, beta, r->d_dt, up, r->arch,
                r->ctx->d_status)));
    return;
  }
  TIMED(tm, KID_IMPLICIT_FWD,
        (k_implicit_fwd_batch<NM, B, WPB><<<gr, 64 * WPB, 0, s>>>(
            r->d_bp, r->d_grp, r->ngroups, r->kt, beta, r->d_dt, r->arch,
            r->ctx->d_status)));
  if constexpr (KES)
    TIMED(tm, KID_IMPLICIT_BWD,
          (k_implicit_bwd_update_kessler_batch<NM, B, WPB><<<gr, 64 * WPB, 0, s>>>(
              r->d_bp, r->d_grp, r->ngroups, r->kt, up, r->arch, r->ctx->d_status)));
  else
    TIMED(tm, KID_IMPLICIT_BWD,
          (k_implicit_bwd_update_batch<NM, B, WPB><<<gr, 64 * WPB, 0, s>>>(
              r->d_bp, r->d_grp, r->ngroups, r->kt, up, r->arch, r->ctx->d_status)));
}

template <int NM, bool KES>
void launch_columns_k(cg_run* r, double beta, const KUpdate& up, KTimer* tm) {
  if (r->cfg.implicit_flag & 8) {
    if (r->tune.implicit_wpb == 4)
      launch_col_batch<NM, 5, 4, KES>(r, beta, up, tm);
    else if (r->tune.implicit_wpb == 2)
      launch_col_batch<NM, 5, 2, KES>(r, beta, up, tm);
    else
      launch_col_batch<NM, 5, 1, KES>(r, beta, up, tm);
  } else {
    launch_col_batch<NM, 3, 1, KES>(r, beta, up, tm);
  }
}

// kes: the sweep also applies the Kessler microphysics (up.kz, up.kdt,
// up.kdtp set; NM = 3 only)
template <int NM>
void launch_columns(cg_run* r, double beta, const KUpdate& up, KTimer* tm, bool kes) {
  if constexpr (NM == 3) {
    if (kes) {
      launch_columns_k<NM, true>(r, beta, up, tm);
      return;
    }
  }
  launch_columns_k<NM, false>(r, beta, up, tm);
}

// Stage s of a step whose size the kernels read from r->d_dt[0].
template <int NM>
int run_stage(cg_run* r, int s) {
  const double sbeta[3] = {1.0, 0.25, 2. / 3.};
  const double sg1[3] = {0.0, 0.25, 2. / 3.};
  const double sg2[3] = {1.0, 0.75, 1. / 3.};
  hipStream_t st = r->ctx->stream;
  KSource src;
  src.wght = sbeta[s];
  src.dtp = r->d_dt;
  src.visc = r->cfg.visc;
  src.p0 = r->cfg.p0;
  src.cp = r->th.gammad / (r->th.gammad - 1.) * r->th.rd;
  src.chi = (r->th.gammad - 1.) / r->th.gammad;
  src.heat_flux = r->cfg.heat_flux;
  KTimer* tm = r->tm.on ? &r->tm : nullptr;
  // x1 columns in the batched solve fused with the update; x2 / x3 lines
  // (implicit_flag bits 1, 2), and periodic x1 columns, on du before it
  const bool imp = (r->cfg.implicit_flag & 1) != 0 && !r->x1_periodic;
  const int bs = r->tune.flux_block;
  KUpdate up;
  up.stage = s;
  up.g1 = sg1[s];
  up.g2 = sg2[s];
  // the guided adjustment (network set, satadj = MLP) runs as its own pass
  // after the update; otherwise the plain Newton adjustment is fused into it
  const bool guided = (s == 2) && r->cfg.satadj == CG_SATADJ_MLP && r->mlp;
  up.satadj = (s == 2) && (r->cfg.satadj != CG_SATADJ_OFF) && !guided;
  const bool tiled = fused_tiles(r) && r->ntiles > 0;
  for (RunBlock& b : r->blocks) {
    const KGrid& g = b.g;
    const long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
    // implicit runs get psf / psv from the previous downward sweep
    if (r->cfg.precision == CG_PREC_FP32) {
      if (tiled && g.f2 && g.f3) {  // faces in the batched tile launch below
        if (!imp && g.grav != 0.0)
          TIMED(tm, KID_SCAN,
                (k_hydrostatic_scan<NM><<<nblk((long long)g.nx2 * g.nx3, 256), 256, 0, st>>>(
                    g, r->kt, b.w, b.psf, b.psv)));
        continue;
      }
      launch_fluxes32<NM>(g, r->kt, r->cfg, src, b.w, b.ff, b.psf, b.psv, b.pref,
                          b.du, st, tm, !imp, bs, r->tune.face_variant,
                          r->tune.march_seg);
    } else {
      launch_fluxes<NM>(g, r->kt, r->cfg, b.w, b.psf, b.psv, b.f1, b.f2, b.f3, st,
                        tm, !imp, bs);
      TIMED(tm, KID_DIVERGENCE,
            (k_divergence<NM><<<nblk(ncell, bs), bs, 0, st>>>(
                g, r->kt, src, b.f1, b.f2, b.f3, b.w,
                g.grav != 0.0 ? b.psv : nullptr, b.du)));
    }
  }
  if (tiled)
    launch_fused_batch<NM>(r->d_bp, r->d_tiles, r->ntiles, r->kt, r->cfg, src,
                           r->tune.face_variant, st, tm);
  // SolveImplicit3D's x3 then x2 lines (solve_implicit_3d.cpp:25-88), with
  // the stage's step beta * dt read from the device
  for (int d = 2; d >= 1; --d) {
    if (!(r->cfg.implicit_flag & (1 << d))) continue;
    for (RunBlock& b : r->blocks) {
      if ((d == 1 && !b.g.f2) || (d == 2 && !b.g.f3)) continue;
      const bool per = b.blk.bc[2 * d] == CG_BC_PERIODIC && b.blk.bc[2 * d + 1] == CG_BC_PERIODIC;
      launch_lines<NM>(b.g, r->kt, sbeta[s], r->d_dt, r->cfg.implicit_flag, d, per, b.du,
                       b.w, r->lines, r->ctx->d_status, st, tm);
    }
  }
  // periodic x1 columns by the cyclic recurrences, then the plain update
  if ((r->cfg.implicit_flag & 1) && r->x1_periodic)
    for (RunBlock& b : r->blocks)
      launch_lines<NM>(b.g, r->kt, sbeta[s], r->d_dt, r->cfg.implicit_flag, 0, true, b.du,
                       b.w, r->lines, r->ctx->d_status, st, tm);
  // Kessler microphysics and sedimentation after the last stage's update
  // and adjustment, with the full step (NM = 3 only: it needs the
  // precipitation species)
  const bool kes = (s == 2) && r->kes && NM == 3;
  if (imp) {
    // with the adjustment in the update (not the guided pass) and every
    // block's x1 sides walls or outflow, the downward sweep applies the
    // microphysics to each new cell state before it fills the x1 ghosts and
    // the hydrostatic pressures
    bool kcol = kes && !guided;
    for (const RunBlock& b : r->blocks)
      kcol = kcol && b.g.bc[0] != CG_BC_PERIODIC && b.g.bc[1] != CG_BC_PERIODIC;
    if (kcol) {
      up.kz = r->kz;
      up.kdt = 1.0;
      up.kdtp = r->d_dt;
    }
    launch_columns<NM>(r, sbeta[s], up, tm, kcol);
    for (RunBlock& b : r->blocks) {
      const KGrid& g = b.g;
      if (guided)
        TIMED(tm, KID_SATADJ,
              (launch_satadj_run_mlp(NM, g, r->kt, r->mlp, b.u, b.w, r->ctx->d_status, st)));
      if (kcol) continue;
      // the Kessler sweep also refills the x1 ghosts and the hydrostatic
      // pressures where the x1 sides are walls or outflow
      const bool kfill = kes && g.bc[0] != CG_BC_PERIODIC && g.bc[1] != CG_BC_PERIODIC;
      if (kfill)
        launch_kessler_fill<NM>(g, r->kt, r->kz, 1.0, r->d_dt, b.u, b.w, b.psf, b.psv, st,
                                tm);
      else if (kes)
        launch_kessler<NM>(g, r->kt, r->kz, 1.0, r->d_dt, b.u, b.w, st, tm);
      if ((guided || kes) && !kfill) {
        // changed cells change w: x1 ghosts and hydrostatic pressures again
        TIMED(tm, KID_BC_X1,
              (k_bc_x1<NM><<<nblk((long long)g.nx2 * g.nx3 * NG, 256), 256, 0, st>>>(g, b.w)));
        if (g.grav != 0.0)
          TIMED(tm, KID_SCAN,
                (k_hydrostatic_scan<NM><<<nblk((long long)g.nx2 * g.nx3, 256), 256, 0, st>>>(
                    g, r->kt, b.w, b.psf, b.psv)));
      }
    }
    // x2 then x3 ghosts of every block (x3 copies the x2 ghosts' corners)
    const unsigned nb = (unsigned)r->blocks.size();
    TIMED(tm, KID_BC_X2,
          (k_bc_x2_batch<NM><<<dim3(r->bc2_grid, nb), 256, 0, st>>>(r->d_bp)));
    TIMED(tm, KID_BC_X3,
          (k_bc_x3_batch<NM><<<dim3(r->bc3_grid, nb), 256, 0, st>>>(r->d_bp)));
  } else {
    for (RunBlock& b : r->blocks) {
      const KGrid& g = b.g;
      const long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
      TIMED(tm, KID_UPDATE,
            (k_update<NM><<<nblk(ncell, 256), 256, 0, st>>>(
                g, r->kt, up, b.u, b.u1, b.du, b.w, r->ctx->d_status)));
      if (guided)
        TIMED(tm, KID_SATADJ,
              (launch_satadj_run_mlp(NM, g, r->kt, r->mlp, b.u, b.w, r->ctx->d_status, st)));
      if (kes) launch_kessler<NM>(g, r->kt, r->kz, 1.0, r->d_dt, b.u, b.w, st, tm);
      launch_bc<NM>(g, b.w, st, tm);
    }
  }
  return CG_OK;
}

// Ghost zones of the six panels from the neighbours, x1 walls, and the
// hydrostatic pressures the next x1 fluxes read.
void sphere_fill(cg_run* r) {
  SphRun& S = *r->sph;
  hipStream_t st = r->ctx->stream;
  KTimer* tm = r->tm.on ? &r->tm : nullptr;
  const KGrid& g = S.k.g;
  const long long ncol = (long long)g.nx2 * g.nx3;
  TIMED(tm, KID_BC_X2,
        (k_sph_ghosts<<<nblk((long long)S.nghost * S.k.nr, 256), 256, 0, st>>>(
            S.k, S.d_ptr, S.d_ghosts, S.nghost)));
  TIMED(tm, KID_BC_X1, (k_sph_bc_x1<<<dim3(nblk(ncol * NG, 256), 6), 256, 0, st>>>(S.k)));
  if (g.grav != 0.0)
    TIMED(tm, KID_SCAN, (k_sph_scan<<<dim3(nblk(ncol, 256), 6), 256, 0, st>>>(S.k, r->kt)));
  S.dirty = false;
}

// One SSPRK(3,3) stage of the cubed sphere (FP64, dry)
int sphere_stage(cg_run* r, int s) {
  const double sbeta[3] = {1.0, 0.25, 2. / 3.};
  const double sg1[3] = {0.0, 0.25, 2. / 3.};
  const double sg2[3] = {1.0, 0.75, 1. / 3.};
  SphRun& S = *r->sph;
  const KSph& ks = S.k;
  const KGrid& g = ks.g;
  hipStream_t st = r->ctx->stream;
  KTimer* tm = r->tm.on ? &r->tm : nullptr;
  const long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
  const long long nf1 = (long long)(g.nx1 + 1) * g.nx2 * g.nx3;
  const long long nface = (long long)ks.nr * ks.n * (ks.n + 1);
  // faces: FP64, or FP32 reconstruction and Riemann fluxes (the FP32-flux
  // form; the x1 decomposition, the face pressures of the x1 momentum flux,
  // the divergence, sources and the column solve stay FP64)
  const bool f32 = r->cfg.precision == CG_PREC_FP32;
  const int psplit = (f32 && g.grav != 0.0) ? 1 : 0;
  if (f32) {
    const long long nx1l = (long long)g.nx2 * g.nx3 * ((g.nx1 + 1 + S.x1seg - 1) / S.x1seg);
    TIMED(tm, KID_FLUX_X1, (k_sph_flux32_x1m<<<dim3(nblk(nx1l, 256), 6), 256, 0, st>>>(
                               ks, r->kt, S.x1seg)));
  }
  else
    TIMED(tm, KID_FLUX_X1, (k_sph_flux_x1<<<dim3(nblk(nf1, 256), 6), 256, 0, st>>>(ks, r->kt)));
  const dim3 gf(nblk(nface, 256), 6);
  const unsigned ge = nblk((long long)S.nedge * ks.nr, 256);
#define CG_SPH_FACES(O, R)                                                                  \
  do {                                                                                    \
    TIMED(tm, KID_FLUX_X2,                                                                \
          (k_sph_edge_states<O, R><<<ge, 256, 0, st>>>(ks, S.d_edges, S.nedge)));         \
    TIMED(tm, KID_FLUX_X2, (k_sph_face_flux<O, 1, R><<<gf, 256, 0, st>>>(ks, r->kt)));    \
    TIMED(tm, KID_FLUX_X3, (k_sph_face_flux<O, 2, R><<<gf, 256, 0, st>>>(ks, r->kt)));    \
  } while (0)
  if (f32) {
    // FP32 form: the x2 and x3 faces of 32 x 8 column tiles at one level,
    // the cells' face values shared through LDS
    const unsigned nt = (unsigned)(((ks.n + SF_J - 1) / SF_J) * ((ks.n + SF_K - 1) / SF_K) *
                                   ks.nr);
    // with the flux divergence, sources and forcing of the tile's cells fused
    if (r->cfg.order_h == 5) {
      TIMED(tm, KID_FLUX_X2,
            (k_sph_edge_states<5, float><<<ge, 256, 0, st>>>(ks, S.d_edges, S.nedge)));
      TIMED(tm, KID_DIVERGENCE, (k_sph_faces32<5, true><<<dim3(nt, 6), SF_J * SF_K, 0, st>>>(
                                    ks, r->kt, sbeta[s], r->d_dt, psplit)));
    } else {
      TIMED(tm, KID_FLUX_X2,
            (k_sph_edge_states<2, float><<<ge, 256, 0, st>>>(ks, S.d_edges, S.nedge)));
      TIMED(tm, KID_DIVERGENCE, (k_sph_faces32<2, true><<<dim3(nt, 6), SF_J * SF_K, 0, st>>>(
                                    ks, r->kt, sbeta[s], r->d_dt, psplit)));
    }
  } else {
    if (r->cfg.order_h == 5)
      CG_SPH_FACES(5, double);
    else
      CG_SPH_FACES(2, double);
  }
#undef CG_SPH_FACES
  if (!f32)
    TIMED(tm, KID_DIVERGENCE, (k_sph_divergence<double><<<dim3(nblk(ncell, 256), 6), 256, 0, st>>>(
                                  ks, sbeta[s], r->d_dt, psplit)));
  KUpdate up;
  up.stage = s;
  up.g1 = sg1[s];
  up.g2 = sg2[s];
  up.satadj = 0;
  if (r->cfg.implicit_flag & 1) {
    // forward sweep, then one downward pass per column: back substitution,
    // stage update, ConservedToPrimitive, x1 walls, next stage's psf / psv;
    // then the x2 / x3 ghosts from the neighbour panels
    const long long ncol = (long long)g.nx2 * g.nx3;
    const dim3 gc(nblk(ncol, 64), 6);
    if (r->cfg.implicit_flag & 8) {
      TIMED(tm, KID_IMPLICIT_FWD,
            (k_sph_implicit_fwd<5><<<gc, 64, 0, st>>>(ks, r->kt, sbeta[s], r->d_dt, S.arch,
                                                      S.arch_panel, r->ctx->d_status)));
      TIMED(tm, KID_IMPLICIT_BWD,
            (k_sph_bwd_update<5><<<gc, 64, 0, st>>>(ks, up, S.arch, S.arch_panel)));
    } else {
      TIMED(tm, KID_IMPLICIT_FWD,
            (k_sph_implicit_fwd<3><<<gc, 64, 0, st>>>(ks, r->kt, sbeta[s], r->d_dt, S.arch,
                                                      S.arch_panel, r->ctx->d_status)));
      TIMED(tm, KID_IMPLICIT_BWD,
            (k_sph_bwd_update<3><<<gc, 64, 0, st>>>(ks, up, S.arch, S.arch_panel)));
    }
    TIMED(tm, KID_BC_X2,
          (k_sph_ghosts<<<nblk((long long)S.nghost * ks.nr, 256), 256, 0, st>>>(
              ks, S.d_ptr, S.d_ghosts, S.nghost)));
    S.dirty = false;
    return CG_OK;
  }
  TIMED(tm, KID_UPDATE, (k_sph_update<<<dim3(nblk(ncell, 256), 6), 256, 0, st>>>(ks, up)));
  sphere_fill(r);
  return CG_OK;
}

int run_one_step(cg_run* r) {
  if (r->sph) {
    for (int s = 0; s < 3; ++s) {
      int e = sphere_stage(r, s);
      if (e) return e;
    }
    return CG_OK;
  }
  for (int s = 0; s < 3; ++s) {
    int e = (r->N == 5) ? run_stage<0>(r, s) : run_stage<3>(r, s);
    if (e) return e;
  }
  return CG_OK;
}

int check_config(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                 const cg_config* cfg) {
  int e = check_block(ctx, blk, th);
  if (e) return e;
  if ((cfg->order_h != 2 && cfg->order_h != 5) ||
      (cfg->order_v != 2 && cfg->order_v != 5))
    return set_err(ctx, CG_ERR_ARG, "order must be 2 or 5");
  e = check_implicit_flag(ctx, *blk, cfg->implicit_flag);
  if (e) return e;
  // x1 periodic at one end only: no such mesh in the reference
  if ((blk->bc[0] == CG_BC_PERIODIC) != (blk->bc[1] == CG_BC_PERIODIC))
    return set_err(ctx, CG_ERR_ARG, "x1 must be periodic at both ends or at neither");
  if (cfg->satadj != CG_SATADJ_OFF && (th->nmass != 3 || th->ivap <= 0 || th->icld <= 0))
    return set_err(ctx, CG_ERR_ARG, "saturation adjustment needs a vapour/cloud pair");
  if (cfg->precision != CG_PREC_FP64 && cfg->precision != CG_PREC_FP32)
    return set_err(ctx, CG_ERR_ARG, "bad precision");
  if (kessler_on(*cfg)) {
    KKessler kz;
    e = make_kessler(ctx, *cfg, *th, &kz);
    if (e) return e;
  }
  return CG_OK;
}

bool alloc_block(RunBlock& b, int N, const cg_config& cfg) {
  auto alloc = [&](double** p, size_t n) {
    return hipMalloc((void**)p, sizeof(double) * (n > 0 ? n : 1)) == hipSuccess;
  };
  const KGrid& g = b.g;
  size_t nc = cell_doubles(g, N);
  bool ok = alloc(&b.u, nc) && alloc(&b.w, nc) && alloc(&b.u1, nc) &&
            alloc(&b.du, nc) && alloc(&b.psf, psf_doubles(g)) &&
            alloc(&b.psv, psf_doubles(g));
  if (ok && cfg.precision == CG_PREC_FP64)
    ok = alloc(&b.f1, f1_doubles(g, N)) && alloc(&b.f2, f2_doubles(g, N)) &&
         alloc(&b.f3, f3_doubles(g, N));
  if (ok && cfg.precision == CG_PREC_FP32)
    ok = hipMalloc((void**)&b.ff, sizeof(float) * (f1_doubles(g, N) + f2_doubles(g, N) +
                                                   f3_doubles(g, N))) == hipSuccess &&
         alloc(&b.pref, g.nc1);
  // Every buffer starts as NaN (all bytes 0xFF), so that a kernel reading
  // an entry no earlier kernel wrote spreads NaN into the state, where the
  // step's checks and the tests see it, instead of a plausible zero.
  auto poison = [&](void* p, size_t bytes) {
    return !p || hipMemset(p, 0xFF, bytes) == hipSuccess;
  };
  if (ok)
    ok = poison(b.u, sizeof(double) * nc) && poison(b.w, sizeof(double) * nc) &&
         poison(b.u1, sizeof(double) * nc) && poison(b.du, sizeof(double) * nc) &&
         poison(b.psf, sizeof(double) * psf_doubles(g)) &&
         poison(b.psv, sizeof(double) * psf_doubles(g)) &&
         poison(b.f1, sizeof(double) * f1_doubles(g, N)) &&
         poison(b.f2, sizeof(double) * f2_doubles(g, N)) &&
         poison(b.f3, sizeof(double) * f3_doubles(g, N)) &&
         poison(b.ff, sizeof(float) * (f1_doubles(g, N) + f2_doubles(g, N) +
                                       f3_doubles(g, N))) &&
         poison(b.pref, sizeof(double) * g.nc1);
  return ok;
}

// Column groups of 64, blocks taken tallest first; each block's archive
// (nx1 AS nx2 nx3 doubles, level-major) starts at its group's aoff.
void build_groups(const std::vector<RunBlock>& blocks, int AS,
                  std::vector<KGroup>& groups, size_t& arch_n) {
  std::vector<int> order(blocks.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return blocks[a].g.nx1 > blocks[b].g.nx1;
  });
  groups.clear();
  arch_n = 0;
  for (int b : order) {
    const KGrid& g = blocks[b].g;
    const long long ncol = (long long)g.nx2 * g.nx3;
    for (long long c0 = 0; c0 < ncol; c0 += 64) {
      KGroup q;
      q.b = b;
      q.col0 = (int)c0;
      q.aoff = (long long)arch_n;
      groups.push_back(q);
    }
    arch_n += (size_t)g.nx1 * AS * ncol;
  }
}

int run_download(cg_run* r, const RunBlock& b, const double* src, double* host) {
  cg_ctx* ctx = r->ctx;
  size_t nc = cell_doubles(b.g, r->N);
  AsyncScratch tmp_s(ctx->stream);
  HIPCHK(ctx, hipMallocAsync(&tmp_s.p, sizeof(double) * nc, ctx->stream));
  double* tmp = (double*)tmp_s.p;
  k_relayout<<<nblk((long long)nc, 256), 256, 0, ctx->stream>>>(b.g, b.ga, r->N, src,
                                                               tmp);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(host, tmp, sizeof(double) * nc, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

template <int NM>
void prime_block(cg_run* r, RunBlock& b) {
  hipStream_t s = r->ctx->stream;
  const KGrid& g = b.g;
  const long long ncell = (long long)g.nx1 * g.nx2 * g.nx3;
  const long long ncol = (long long)g.nx2 * g.nx3;
  k_eos<NM, 0><<<nblk(ncell, 256), 256, 0, s>>>(g, r->kt, b.w, b.u);
  launch_bc<NM>(g, b.w, s);
  if (b.pref) k_level_mean_p<NM><<<g.nc1, 256, 0, s>>>(g, b.w, b.pref);
  if (g.grav != 0.0)
    k_hydrostatic_scan<NM><<<nblk(ncol, 256), 256, 0, s>>>(g, r->kt, b.w, b.psf, b.psv);
}

int ensure_tiles(cg_run* r);

int capture(cg_run* r) {
  cg_ctx* ctx = r->ctx;
  drop_graph(r);
  int te = ensure_tiles(r);
  if (te) return te;
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int e = run_one_step(r);
  hipGraph_t gr = nullptr;
  hipError_t ce = hipStreamEndCapture(ctx->stream, &gr);
  if (e) {
    if (gr) (void)hipGraphDestroy(gr);
    return e;
  }
  if (ce != hipSuccess)
    return set_err(ctx, CG_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
  r->graph = gr;
  HIPCHK(ctx, hipGraphInstantiate(&r->gexec, r->graph, nullptr, nullptr, 0));
  return CG_OK;
}

// The tile table of the run's launch shape (host work and a synchronous
// copy: called before capture and before direct steps, never while
// capturing).  Tiles of the tallest blocks come first.
int ensure_tiles(cg_run* r) {
  if (!fused_tiles(r)) return CG_OK;
  const int v = r->tune.face_variant, seg = r->tune.march_seg;
  if (r->d_tiles && r->tile_variant == v && r->tile_seg == seg) return CG_OK;
  const int tj = (v == 3) ? 32 : 16, tk = (v == 3) ? 8 : 16;
  std::vector<int> order(r->blocks.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return r->blocks[a].g.nx1 > r->blocks[b].g.nx1;
  });
  // Workgroup w runs on XCD w % 8: each XCD gets a band of tile rows (k)
  // of every block and level segment, so that a tile's halo neighbours
  // mostly run on the same XCD and share its L2.  Queues of unequal length
  // are padded with empty tiles (block -1).
  std::vector<KTile> q[8];
  for (int b : order) {
    const KGrid& g = r->blocks[b].g;
    if (!(g.f2 && g.f3)) continue;
    const int nbk = (g.nx3 + tk - 1) / tk;
    for (int c0 = g.is; c0 <= g.ie; c0 += seg)
      for (int rk = 0; rk < nbk; ++rk) {
        const int x = (int)((long long)rk * 8 / nbk);
        for (int j0 = g.js; j0 <= g.je; j0 += tj)
          q[x].push_back(KTile{b, j0, g.ks + rk * tk, c0, std::min(c0 + seg, g.ie + 1)});
      }
  }
  size_t qmax = 0;
  for (int x = 0; x < 8; ++x) qmax = std::max(qmax, q[x].size());
  std::vector<KTile> tiles;
  for (size_t slot = 0; slot < qmax; ++slot)
    for (int x = 0; x < 8; ++x)
      tiles.push_back(slot < q[x].size() ? q[x][slot] : KTile{-1, 0, 0, 0, 0});
  if (r->d_tiles) (void)hipFree(r->d_tiles);
  r->d_tiles = nullptr;
  r->ntiles = (int)tiles.size();
  r->tile_variant = v;
  r->tile_seg = seg;
  if (tiles.empty()) return CG_OK;
  HIPCHK(r->ctx, hipMalloc((void**)&r->d_tiles, sizeof(KTile) * tiles.size()));
  HIPCHK(r->ctx, hipMemcpy(r->d_tiles, tiles.data(), sizeof(KTile) * tiles.size(),
                           hipMemcpyHostToDevice));
  // the copy runs on the null stream: done before the next launch reads it
  HIPCHK(r->ctx, hipDeviceSynchronize());
  return CG_OK;
}

// Mesh::NewTimeStep (upstream athenapp mesh/mesh.cpp): at most twice the
// old step, at most the blocks' minimum, cut to end exactly at tlim.
double mesh_new_dt(double dt_old, double blocks_dt, double time, double tlim) {
  double dt = std::fmin(2.0 * dt_old, blocks_dt);
  if (time < tlim && tlim - time < dt) dt = tlim - time;
  return dt;
}

// Enqueues the step size every stage kernel reads.
int set_step(cg_run* r, double dt) {
  k_set_scalar<<<1, 64, 0, r->ctx->stream>>>(r->d_dt, dt);
  HIPCHK(r->ctx, hipGetLastError());
  return CG_OK;
}

// Enqueues Hydro::NewBlockTimeStep over every block of the run (one launch
// over the block table) and the minimum over blocks into d_dt[1].
int enqueue_new_dt(cg_run* r, double cfl) {
  hipStream_t s = r->ctx->stream;
  const int nb = (int)r->blocks.size();
  dim3 grid(r->dt_parts, nb);
  KGrid g0 = r->blocks[0].g;
  KTimer* tm = r->tm.on ? &r->tm : nullptr;
  unsigned long long* bad = reinterpret_cast<unsigned long long*>(r->d_dt + 2);
  k_set_u64<<<1, 64, 0, s>>>(bad, ~0ULL);
  if (r->N == 5)
    TIMED(tm, KID_NEW_DT,
          (k_new_dt<0><<<grid, 256, 0, s>>>(r->d_bp, g0, nullptr, r->kt, cfl,
                                            r->cfg.implicit_flag, r->dt_part, bad)));
  else
    TIMED(tm, KID_NEW_DT,
          (k_new_dt<3><<<grid, 256, 0, s>>>(r->d_bp, g0, nullptr, r->kt, cfl,
                                            r->cfg.implicit_flag, r->dt_part, bad)));
  k_new_dt_finish<<<1, 256, 0, s>>>(r->dt_part, r->dt_parts * nb, cfl, r->d_dt + 1);
  HIPCHK(r->ctx, hipGetLastError());
  return CG_OK;
}

// Reads d_dt[1] and the state checks' key after the queued work; with
// report, a failing cell becomes CG_ERR_NAN / CG_ERR_NEGATIVE / CG_ERR_COLD
// with its place (where: block, k, j, i, or -1).
int read_new_dt(cg_run* r, double* out, bool report, int* where = nullptr) {
  cg_ctx* ctx = r->ctx;
  HIPCHK(ctx, hipMemcpyAsync(r->h_dt, r->d_dt + 1, 2 * sizeof(double),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *out = r->h_dt[0];
  unsigned long long key;
  std::memcpy(&key, &r->h_dt[1], sizeof(key));
  if (where)
    for (int q = 0; q < 4; ++q) where[q] = -1;
  if (key == ~0ULL || !report) return CG_OK;
  const int code = (int)(key & 3ULL);
  const unsigned long long cell = key >> 2;
  const int b = (int)(cell >> 40);
  const unsigned long long x = cell & ((1ULL << 40) - 1);
  const KGrid& g = r->blocks[b].g;
  const int i = (int)(x % g.nx1), j = (int)((x / g.nx1) % g.nx2),
            k = (int)(x / ((unsigned long long)g.nx1 * g.nx2));
  if (where) {
    where[0] = b;
    where[1] = k;
    where[2] = j;
    where[3] = i;
  }
  static const char* what[3] = {"NaN in the state", "negative density or pressure",
                                "temperature below |g| dx1 / Rd"};
  const int err[3] = {CG_ERR_NAN, CG_ERR_NEGATIVE, CG_ERR_COLD};
  return set_err(ctx, err[code], "%s at block %d, interior cell (k, j, i) = (%d, %d, %d)",
                 what[code], b, k, j, i);
}

}  // namespace

extern "C" {

int cg_batch_create(cg_ctx* ctx, int nblocks, const cg_block* blks,
                    const cg_thermo* th, const cg_config* cfg, cg_run** out) {
  if (!ctx || !cfg || !out || !blks || !th) return CG_ERR_ARG;
  *out = nullptr;
  if (nblocks < 1) return set_err(ctx, CG_ERR_ARG, "nblocks must be >= 1");
  for (int b = 0; b < nblocks; ++b) {
    int e = check_config(ctx, &blks[b], th, cfg);
    if (e) return e;
    // one x1 treatment for every block of a run
    if ((blks[b].bc[0] == CG_BC_PERIODIC) != (blks[0].bc[0] == CG_BC_PERIODIC))
      return set_err(ctx, CG_ERR_ARG, "the blocks of a run must agree on periodic x1");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  cg_run* r = new cg_run();
  r->x1_periodic = blks[0].bc[0] == CG_BC_PERIODIC;
  r->ctx = ctx;
  r->th = *th;
  r->cfg = *cfg;
  r->kt = make_thermo(*th);
  r->kes = kessler_on(*cfg) && make_kessler(ctx, *cfg, *th, &r->kz) == CG_OK;
  r->N = 5 + th->nmass;
  r->AS = (cfg->implicit_flag & 8) ? 30 : 12;
  r->tune.implicit_wpb = 1;
  r->tune.implicit_grid = 0;
  r->tune.flux_block = 256;
  r->tune.face_variant = 3;
  r->tune.march_seg = 16;
  r->tune.implicit_fused = 1;
  r->tune.step_ms = 0.0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
    r->ncu = prop.multiProcessorCount;
  r->blocks.resize(nblocks);
  bool ok = true;
  for (int b = 0; b < nblocks && ok; ++b) {
    RunBlock& rb = r->blocks[b];
    rb.blk = blks[b];
    rb.g = make_grid(blks[b], 1);
    rb.g.status = ctx->d_status;  // face checks of the flux kernels
    rb.ga = make_grid(blks[b], 0);
    ok = alloc_block(rb, r->N, *cfg);
  }
  // workgroups per block of the time-step reduction: about 16 cells per
  // thread in the largest block, at most 1024
  long long maxcells = 1;
  for (const RunBlock& rb : r->blocks)
    maxcells = std::max(maxcells, (long long)rb.g.nx1 * rb.g.nx2 * rb.g.nx3);
  r->dt_parts = (int)std::min(1024LL, std::max(1LL, maxcells / (256 * 16)));
  ok = ok && hipMalloc((void**)&r->part, sizeof(double) * 1024 * 8) == hipSuccess &&
       hipMalloc((void**)&r->d_dt, sizeof(double) * 3) == hipSuccess &&
       hipMemset(r->d_dt, 0, sizeof(double) * 3) == hipSuccess &&
       hipHostMalloc((void**)&r->h_dt, sizeof(double) * 2, hipHostMallocDefault) ==
           hipSuccess &&
       hipMalloc((void**)&r->dt_part, sizeof(double) * r->dt_parts * nblocks) == hipSuccess;
  if (ok) {
    std::vector<KBlockPtr> bp(nblocks);
    for (int b = 0; b < nblocks; ++b) {
      const RunBlock& rb = r->blocks[b];
      bp[b].g = rb.g;
      bp[b].u = rb.u;
      bp[b].w = rb.w;
      bp[b].u1 = rb.u1;
      bp[b].du = rb.du;
      bp[b].psf = rb.psf;
      bp[b].psv = rb.psv;
      bp[b].pref = rb.pref;
      const KGrid& g = rb.g;
      r->bc2_grid = std::max(r->bc2_grid, nblk((long long)g.nc1 * g.nx3 * NG, 256));
      r->bc3_grid = std::max(r->bc3_grid, nblk((long long)g.nc1 * g.nc2 * NG, 256));
    }
    ok = hipMalloc((void**)&r->d_bp, sizeof(KBlockPtr) * nblocks) == hipSuccess &&
         hipMemcpy(r->d_bp, bp.data(), sizeof(KBlockPtr) * nblocks,
                   hipMemcpyHostToDevice) == hipSuccess;
  }
  const bool x1_lines = (cfg->implicit_flag & 1) && r->x1_periodic;
  if (ok && ((cfg->implicit_flag & 6) || x1_lines)) {
    size_t n = 0;
    const int B = (cfg->implicit_flag & 8) ? 5 : 3;
    for (const RunBlock& rb : r->blocks) {
      for (int d = 1; d <= 2; ++d)
        if ((cfg->implicit_flag & (1 << d)) && (d == 1 ? rb.g.f2 : rb.g.f3))
          n = std::max(n, line_doubles(rb.g, d, B,
                                       rb.blk.bc[2 * d] == CG_BC_PERIODIC &&
                                           rb.blk.bc[2 * d + 1] == CG_BC_PERIODIC));
      if (x1_lines) n = std::max(n, line_doubles(rb.g, 0, B, true));
    }
    ok = hipMalloc((void**)&r->lines, sizeof(double) * std::max<size_t>(n, 1)) == hipSuccess &&
         hipMemset(r->lines, 0xFF, sizeof(double) * std::max<size_t>(n, 1)) == hipSuccess;
  }
  if (ok && (cfg->implicit_flag & 1) && !r->x1_periodic) {
    std::vector<KGroup> groups;
    size_t arch_n = 0;
    build_groups(r->blocks, r->AS, groups, arch_n);
    r->ngroups = (int)groups.size();
    ok = hipMalloc((void**)&r->arch, sizeof(double) * std::max<size_t>(arch_n, 1)) ==
             hipSuccess &&
         hipMemset(r->arch, 0xFF, sizeof(double) * std::max<size_t>(arch_n, 1)) ==
             hipSuccess &&
         hipMalloc((void**)&r->d_grp, sizeof(KGroup) * std::max(r->ngroups, 1)) ==
             hipSuccess &&
         hipMemcpy(r->d_grp, groups.data(), sizeof(KGroup) * groups.size(),
                   hipMemcpyHostToDevice) == hipSuccess;
  }
  r->tm.s = ctx->stream;
  if (ok) ok = hipEventCreate(&r->tm.e0) == hipSuccess && hipEventCreate(&r->tm.e1) == hipSuccess;
  // the fills and table copies above are queued on the null stream, which
  // the context's non-blocking stream does not wait for: finish them before
  // any later work (set_prim's relayout, the first step) can be overwritten
  if (ok) ok = hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    cg_run_destroy(r);
    return set_err(ctx, CG_ERR_HIP, "allocation failed");
  }
  *out = r;
  return CG_OK;
}

int cg_run_create(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                  const cg_config* cfg, cg_run** out) {
  if (!blk) return CG_ERR_ARG;
  return cg_batch_create(ctx, 1, blk, th, cfg, out);
}

int cg_sphere_create(cg_ctx* ctx, const cg_block* panel, const cg_thermo* th,
                     const cg_config* cfg, const cg_sphere* sp, cg_run** out) {
  if (!ctx || !panel || !th || !cfg || !sp || !out) return CG_ERR_ARG;
  *out = nullptr;
  if (th->nmass != 0 || panel->nmass != 0)
    return set_err(ctx, CG_ERR_ARG, "the cubed sphere runs dry gas (nmass 0)");
  if (panel->nx2 != panel->nx3 || panel->nx2 < 2 * NG || panel->nx2 % 2)
    return set_err(ctx, CG_ERR_ARG, "panels must be square with an even side of >= 6 cells");
  if (panel->bc[0] != CG_BC_REFLECT || panel->bc[1] != CG_BC_REFLECT)
    return set_err(ctx, CG_ERR_ARG, "the cubed sphere needs reflecting x1 walls");
  for (int s = 2; s < 6; ++s)
    if (panel->bc[s] != CG_BC_PANEL)
      return set_err(ctx, CG_ERR_ARG, "x2 / x3 sides of a panel must be CG_BC_PANEL");
  if ((cfg->precision != CG_PREC_FP64 && cfg->precision != CG_PREC_FP32) ||
      cfg->rsolver != CG_RS_HLLC ||
      (cfg->implicit_flag != 0 && cfg->implicit_flag != 1 && cfg->implicit_flag != 9) ||
      cfg->satadj != CG_SATADJ_OFF || cfg->visc != 0.0 || cfg->heat_flux != 0.0 ||
      kessler_on(*cfg))
    return set_err(ctx, CG_ERR_ARG,
                   "cubed sphere: HLLC, implicit_flag 0 / 1 / 9, no adjustment, "
                   "diffusion, heating or microphysics");
  if (!(panel->x1min > 0.0)) return set_err(ctx, CG_ERR_ARG, "x1min is the inner radius");
  // six blocks; their x2 / x3 flags are never read by the sphere's kernels
  cg_block b = *panel;
  for (int s = 2; s < 6; ++s) b.bc[s] = CG_BC_PERIODIC;
  std::vector<cg_block> blks(6, b);
  cg_config c = *cfg;
  c.order_v = 5;  // Weno5X1 under CUBED_SPHERE (patches/19.calculate_fluxes.patch:29-31)
  c.implicit_flag = 0;  // the batch's own column solve is not used
  c.precision = CG_PREC_FP64;  // FP64 flux arrays; the FP32 form only changes x2 / x3 faces
  cg_run* r = nullptr;
  int e = cg_batch_create(ctx, 6, blks.data(), th, &c, &r);
  if (e) return e;
  r->cfg.implicit_flag = cfg->implicit_flag;
  r->cfg.precision = cfg->precision;
  r->sph = new SphRun();
  SphRun& S = *r->sph;
  S.sp = *sp;
  const int n = panel->nx2, nr = panel->nx1;
  SphGeom G = sphere_geometry(n, nr, panel->x1min, panel->dx1);
  const size_t nkj = (size_t)SKJ_COUNT * G.nc * G.nc;
  const long long ncol = (long long)n * n;
  const int B = (cfg->implicit_flag & 8) ? 5 : 3;
  S.arch_panel = (long long)(B * B + B) * nr * ncol;
  S.nghost = (int)G.ghosts.size();
  S.nedge = (int)G.edges.size();
  KSph& k = S.k;
  k.g = r->blocks[0].g;
  k.lrn = (long long)5 * nr * n * (n + 1);
  bool ok = hipMalloc((void**)&S.kj, sizeof(double) * nkj * 6) == hipSuccess &&
            hipMalloc((void**)&S.x1, sizeof(double) * (2 * G.nc1 + 1)) == hipSuccess &&
            hipMalloc((void**)&S.lr, sizeof(double) * 24 * k.lrn) == hipSuccess &&
            hipMemset(S.lr, 0xFF, sizeof(double) * 24 * k.lrn) == hipSuccess &&
            hipMalloc((void**)&S.d_ghosts, sizeof(SphGhost) * S.nghost) == hipSuccess &&
            hipMalloc((void**)&S.d_edges, sizeof(SphEdge) * S.nedge) == hipSuccess &&
            hipMalloc((void**)&S.d_ptr, sizeof(KSphPtrs)) == hipSuccess;
  if (ok && (cfg->implicit_flag & 1))
    ok = hipMalloc((void**)&S.arch, sizeof(double) * 6 * S.arch_panel) == hipSuccess &&
         hipMemset(S.arch, 0xFF, sizeof(double) * 6 * S.arch_panel) == hipSuccess;
  if (ok) {
    for (int p = 0; p < 6; ++p)
      ok = ok && hipMemcpy(S.kj + p * nkj, G.kj[p].data(), sizeof(double) * nkj,
                           hipMemcpyHostToDevice) == hipSuccess;
    ok = ok &&
         hipMemcpy(S.x1, G.x1f.data(), sizeof(double) * (G.nc1 + 1), hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(S.x1 + G.nc1 + 1, G.x1v.data(), sizeof(double) * G.nc1,
                   hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(S.d_ghosts, G.ghosts.data(), sizeof(SphGhost) * S.nghost,
                   hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(S.d_edges, G.edges.data(), sizeof(SphEdge) * S.nedge,
                   hipMemcpyHostToDevice) == hipSuccess;
  }
  if (ok) {
    KSphPtrs hp;
    for (int p = 0; p < 6; ++p) {
      const RunBlock& rb = r->blocks[p];
      k.kj[p] = S.kj + p * nkj;
      k.u[p] = rb.u;
      k.w[p] = rb.w;
      k.u1[p] = rb.u1;
      k.du[p] = rb.du;
      k.f1[p] = rb.f1;
      k.f2[p] = rb.f2;
      k.f3[p] = rb.f3;
      k.psf[p] = rb.psf;
      k.psv[p] = rb.psv;
      hp.u[p] = rb.u;
      hp.w[p] = rb.w;
      hp.kj[p] = k.kj[p];
    }
    ok = hipMemcpy(S.d_ptr, &hp, sizeof(KSphPtrs), hipMemcpyHostToDevice) == hipSuccess;
  }
  // null-stream fills and copies finish before work on the context's stream
  if (ok) ok = hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    cg_run_destroy(r);
    return set_err(ctx, CG_ERR_HIP, "allocation failed");
  }
  k.lr = S.lr;
  k.x1f = S.x1;
  k.x1v = S.x1 + G.nc1 + 1;
  k.n = n;
  k.nc = G.nc;
  k.nr = nr;
  k.dx1 = panel->dx1;
  k.gm1 = th->gammad - 1.;
  k.omega = sp->omega;
  k.hs94 = sp->hs94;
  k.ts = sp->ts;
  k.dt_eq = sp->dt_eq;
  k.dtheta = sp->dtheta;
  k.p0 = sp->p0;
  k.kf = sp->kf;
  k.ka = sp->ka;
  k.ks = sp->ks;
  k.sigmab = sp->sigmab;
  k.cp = th->gammad / (th->gammad - 1.) * th->rd;
  k.rd = th->rd;
  k.kappa = th->rd / k.cp;
  *out = r;
  return CG_OK;
}

int cg_run_destroy(cg_run* r) {
  if (!r) return CG_ERR_ARG;
  (void)hipStreamSynchronize(r->ctx->stream);
  drop_graph(r);
  for (RunBlock& b : r->blocks) free_block(b);
  if (r->arch) (void)hipFree(r->arch);
  if (r->lines) (void)hipFree(r->lines);
  if (r->d_bp) (void)hipFree(r->d_bp);
  if (r->d_grp) (void)hipFree(r->d_grp);
  if (r->part) (void)hipFree(r->part);
  if (r->d_dt) (void)hipFree(r->d_dt);
  if (r->dt_part) (void)hipFree(r->dt_part);
  if (r->d_tiles) (void)hipFree(r->d_tiles);
  if (r->h_dt) (void)hipHostFree(r->h_dt);
  if (r->tm.e0) (void)hipEventDestroy(r->tm.e0);
  if (r->tm.e1) (void)hipEventDestroy(r->tm.e1);
  if (r->sph) {
    SphRun& S = *r->sph;
    for (void* q : {(void*)S.d_ptr, (void*)S.kj, (void*)S.x1, (void*)S.lr, (void*)S.arch,
                    (void*)S.d_ghosts, (void*)S.d_edges})
      if (q) (void)hipFree(q);
    delete r->sph;
  }
  delete r;
  return CG_OK;
}

int cg_run_nblocks(cg_run* r, int* n) {
  if (!r || !n) return CG_ERR_ARG;
  *n = (int)r->blocks.size();
  return CG_OK;
}

int cg_run_block_set_prim_host(cg_run* r, int b, const double* w_host) {
  if (!r || !w_host || b < 0 || b >= (int)r->blocks.size()) return CG_ERR_ARG;
  cg_ctx* ctx = r->ctx;
  RunBlock& rb = r->blocks[b];
  size_t nc = cell_doubles(rb.g, r->N);
  HIPCHK(ctx, hipMemcpyAsync(rb.du, w_host, sizeof(double) * nc, hipMemcpyHostToDevice,
                             ctx->stream));
  // the step never reads the conserved ghosts (only interior u is set from
  // w below); they are handed back as zeros
  HIPCHK(ctx, hipMemsetAsync(rb.u, 0, sizeof(double) * nc, ctx->stream));
  k_relayout<<<nblk((long long)nc, 256), 256, 0, ctx->stream>>>(rb.ga, rb.g, r->N, rb.du,
                                                               rb.w);
  if (r->sph) {
    // interior u with covariant momenta; the panels' ghosts and hydrostatic
    // pressures are refreshed before the next step
    const KGrid& g = rb.g;
    k_sph_prim_to_cons<<<nblk((long long)g.nx1 * g.nx2 * g.nx3, 256), 256, 0, ctx->stream>>>(
        r->sph->k, b);
    r->sph->dirty = true;
  } else if (r->N == 5) {
    prime_block<0>(r, rb);
  } else {
    prime_block<3>(r, rb);
  }
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CG_OK;
}

int cg_run_block_get_cons_host(cg_run* r, int b, double* u_host) {
  if (!r || !u_host || b < 0 || b >= (int)r->blocks.size()) return CG_ERR_ARG;
  return run_download(r, r->blocks[b], r->blocks[b].u, u_host);
}

int cg_run_block_get_prim_host(cg_run* r, int b, double* w_host) {
  if (!r || !w_host || b < 0 || b >= (int)r->blocks.size()) return CG_ERR_ARG;
  return run_download(r, r->blocks[b], r->blocks[b].w, w_host);
}

int cg_run_set_prim_host(cg_run* r, const double* w_host) {
  return cg_run_block_set_prim_host(r, 0, w_host);
}

int cg_run_get_cons_host(cg_run* r, double* u_host) {
  return cg_run_block_get_cons_host(r, 0, u_host);
}

int cg_run_get_prim_host(cg_run* r, double* w_host) {
  return cg_run_block_get_prim_host(r, 0, w_host);
}

int cg_run_device_state(cg_run* r, double** u, double** w) {
  if (!r) return CG_ERR_ARG;
  if (u) *u = r->blocks[0].u;
  if (w) *w = r->blocks[0].w;
  return CG_OK;
}

int cg_run_set_timing(cg_run* r, int on) {
  if (!r) return CG_ERR_ARG;
  r->tm.on = on != 0;
  for (int q = 0; q < KID_COUNT; ++q) {
    r->tm.ms[q] = 0.0;
    r->tm.cnt[q] = 0;
  }
  return CG_OK;
}

int cg_run_kernel_stats(cg_run* r, int kernel_id, double* total_ms,
                        long long* launches, const char** name) {
  if (!r || kernel_id < 0 || kernel_id >= KID_COUNT) return CG_ERR_ARG;
  if (total_ms) *total_ms = r->tm.ms[kernel_id];
  if (launches) *launches = r->tm.cnt[kernel_id];
  if (name) *name = kKernelNames[kernel_id];
  return CG_OK;
}

int cg_run_step(cg_run* r, double dt, int nsteps, int use_graph) {
  if (!r) return CG_ERR_ARG;
  if (!(dt > 0.0) || nsteps < 0) return set_err(r->ctx, CG_ERR_ARG, "bad dt/nsteps");
  cg_ctx* ctx = r->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int e = set_step(r, dt);
  if (!e) e = ensure_tiles(r);
  if (e) return e;
  if (r->sph && r->sph->dirty) sphere_fill(r);
  if (use_graph && !r->tm.on) {
    if (!r->gexec && (e = capture(r))) return e;
    for (int n = 0; n < nsteps; ++n) HIPCHK(ctx, hipGraphLaunch(r->gexec, ctx->stream));
  } else {
    for (int n = 0; n < nsteps; ++n)
      if ((e = run_one_step(r))) return e;
  }
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return check_status(ctx);
}

static __global__ void k_copy_scalar(double* __restrict__ out,
                                     const double* __restrict__ in) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = in[0];
}

int cg_run_stage(cg_run* r, int stage, double dt, const double* dt_dev) {
  if (!r || stage < 0 || stage > 2) return CG_ERR_ARG;
  if (!dt_dev && !(dt > 0.0)) return set_err(r->ctx, CG_ERR_ARG, "dt must be positive");
  cg_ctx* ctx = r->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int e = ensure_tiles(r);
  if (e) return e;
  if (dt_dev)
    k_copy_scalar<<<1, 64, 0, ctx->stream>>>(r->d_dt, dt_dev);
  else if ((e = set_step(r, dt)))
    return e;
  if (r->sph && r->sph->dirty) sphere_fill(r);
  e = r->sph ? sphere_stage(r, stage)
             : (r->N == 5) ? run_stage<0>(r, stage) : run_stage<3>(r, stage);
  if (e) return e;
  HIPCHK(ctx, hipGetLastError());
  return CG_OK;
}

int cg_run_new_dt(cg_run* r, double cfl, double* dt_out) {
  if (!r || !dt_out || !(cfl > 0.0)) return CG_ERR_ARG;
  if (r->sph) return set_err(r->ctx, CG_ERR_ARG, "not available for cubed-sphere runs");
  HIPCHK(r->ctx, hipSetDevice(r->ctx->device));
  int e = enqueue_new_dt(r, cfl);
  if (!e) e = read_new_dt(r, dt_out, false);
  return e;
}

int cg_run_check(cg_run* r, int* where) {
  if (!r) return CG_ERR_ARG;
  if (r->sph) return set_err(r->ctx, CG_ERR_ARG, "not available for cubed-sphere runs");
  HIPCHK(r->ctx, hipSetDevice(r->ctx->device));
  double dt = 0.0;
  int e = enqueue_new_dt(r, 1.0);
  if (!e) e = read_new_dt(r, &dt, true, where);
  return e;
}

int cg_run_advance(cg_run* r, double cfl, double tlim, long long nlim,
                   int use_graph, double* time, double* dt, long long* ncycle) {
  if (!r || !time || !dt || !ncycle || !(cfl > 0.0)) return CG_ERR_ARG;
  if (r->sph) return set_err(r->ctx, CG_ERR_ARG, "not available for cubed-sphere runs");
  cg_ctx* ctx = r->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int e = CG_OK;
  double bdt = 0.0;
  if (!(*dt > 0.0)) {  // first step: the blocks' minimum, cut at tlim
    if ((e = enqueue_new_dt(r, cfl)) || (e = read_new_dt(r, &bdt, true))) return e;
    *dt = mesh_new_dt(__DBL_MAX__, bdt, *time, tlim);
  }
  if ((e = ensure_tiles(r))) return e;
  if (use_graph && !r->tm.on && !r->gexec && (e = capture(r))) return e;
  while (*time < tlim && (nlim < 0 || *ncycle < nlim)) {
    if (!(*dt > 0.0)) return set_err(ctx, CG_ERR_ARG, "time step %g is not positive", *dt);
    if ((e = set_step(r, *dt))) return e;
    if (use_graph && !r->tm.on)
      HIPCHK(ctx, hipGraphLaunch(r->gexec, ctx->stream));
    else if ((e = run_one_step(r)))
      return e;
    if ((e = enqueue_new_dt(r, cfl)) || (e = read_new_dt(r, &bdt, true)) ||
        (e = check_status(ctx)))
      return e;
    *ncycle += 1;
    *time += *dt;
    *dt = mesh_new_dt(*dt, bdt, *time, tlim);
  }
  return CG_OK;
}

int cg_run_block_sums(cg_run* r, int b, double* out_host) {
  if (!r || !out_host || b < 0 || b >= (int)r->blocks.size()) return CG_ERR_ARG;
  cg_ctx* ctx = r->ctx;
  const RunBlock& rb = r->blocks[b];
  if (r->sph) return set_err(r->ctx, CG_ERR_ARG, "not available for cubed-sphere runs");
  const int nb = 1024;
  if (r->N == 5)
    k_sums<0><<<nb, 256, 0, ctx->stream>>>(rb.g, rb.u, r->part);
  else
    k_sums<3><<<nb, 256, 0, ctx->stream>>>(rb.g, rb.u, r->part);
  HIPCHK(ctx, hipGetLastError());
  std::vector<double> h((size_t)r->N * nb);
  HIPCHK(ctx, hipMemcpyAsync(h.data(), r->part, sizeof(double) * h.size(),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int n = 0; n < r->N; ++n) {
    double s = 0.0;
    for (int q = 0; q < nb; ++q) s += h[(size_t)n * nb + q];
    out_host[n] = s;
  }
  return CG_OK;
}

int cg_run_sums(cg_run* r, double* out_host) { return cg_run_block_sums(r, 0, out_host); }

int cg_run_history(cg_run* r, double* out_host) {
  if (!r || !out_host) return CG_ERR_ARG;
  if (r->sph) return set_err(r->ctx, CG_ERR_ARG, "not available for cubed-sphere runs");
  cg_ctx* ctx = r->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int nb = (int)r->blocks.size(), P = 64, NQ = 8 + r->th.nmass;
  double* part = nullptr;
  HIPCHK(ctx, hipMallocAsync((void**)&part, sizeof(double) * nb * P * NQ, ctx->stream));
  if (r->N == 5)
    k_history<0><<<dim3(P, nb), 256, 0, ctx->stream>>>(r->d_bp, part);
  else
    k_history<3><<<dim3(P, nb), 256, 0, ctx->stream>>>(r->d_bp, part);
  std::vector<double> h((size_t)nb * P * NQ);
  hipError_t le = hipGetLastError();
  hipError_t ce = hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(),
                                 hipMemcpyDeviceToHost, ctx->stream);
  (void)hipFreeAsync(part, ctx->stream);
  HIPCHK(ctx, le);
  HIPCHK(ctx, ce);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < NQ; ++q) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b)
      for (int p = 0; p < P; ++p) s += h[((size_t)b * P + p) * NQ + q];
    out_host[q] = s;
  }
  return CG_OK;
}

int cg_run_debug(cg_run* r, double* vmin, double* vmax, double* xmin, double* xmax) {
  if (!r || !vmin || !vmax || !xmin || !xmax) return CG_ERR_ARG;
  if (r->sph) return set_err(r->ctx, CG_ERR_ARG, "not available for cubed-sphere runs");
  cg_ctx* ctx = r->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int nb = (int)r->blocks.size(), P = 64, N = r->N;
  const size_t np = (size_t)nb * P * N * 2;
  AsyncScratch pv_s(ctx->stream), px_s(ctx->stream);
  HIPCHK(ctx, hipMallocAsync(&pv_s.p, sizeof(double) * np, ctx->stream));
  HIPCHK(ctx, hipMallocAsync(&px_s.p, sizeof(long long) * np, ctx->stream));
  double* pv = (double*)pv_s.p;
  long long* px = (long long*)px_s.p;
  if (N == 5)
    k_minmax<0><<<dim3(P, nb), 256, 0, ctx->stream>>>(r->d_bp, pv, px);
  else
    k_minmax<3><<<dim3(P, nb), 256, 0, ctx->stream>>>(r->d_bp, pv, px);
  std::vector<double> hv(np);
  std::vector<long long> hx(np);
  hipError_t le = hipGetLastError();
  hipError_t c1 = hipMemcpyAsync(hv.data(), pv, sizeof(double) * np, hipMemcpyDeviceToHost,
                                 ctx->stream);
  hipError_t c2 = hipMemcpyAsync(hx.data(), px, sizeof(long long) * np,
                                 hipMemcpyDeviceToHost, ctx->stream);
  HIPCHK(ctx, le);
  HIPCHK(ctx, c1);
  HIPCHK(ctx, c2);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  // DebugOutput's initial values and strict comparisons over the blocks in
  // order; within a block the device candidates tie to the first cell
  const double FMAXV = 3.40282346638528859812e+38, FMINV = 1.17549435082228750797e-38;
  for (int n = 0; n < N; ++n) {
    vmin[n] = FMAXV;
    vmax[n] = FMINV;
    for (int d = 0; d < 3; ++d) {
      xmin[3 * n + d] = FMAXV;
      xmax[3 * n + d] = FMINV;
    }
    for (int b = 0; b < nb; ++b) {
      const cg_block& bk = r->blocks[b].blk;
      double bv[2] = {0.0, 0.0};
      long long bx[2] = {-1, -1};
      for (int p = 0; p < P; ++p)
        for (int side = 0; side < 2; ++side) {
          const size_t e = 2 * (((size_t)b * P + p) * N + n) + side;
          const double v = hv[e];
          const long long x = hx[e];
          if (x < 0) continue;
          const bool take = bx[side] < 0 ||
                            (side ? v > bv[side] : v < bv[side]) ||
                            (v == bv[side] && x < bx[side]);
          if (take) {
            bv[side] = v;
            bx[side] = x;
          }
        }
      for (int side = 0; side < 2; ++side) {
        if (bx[side] < 0) continue;
        double& best = side ? vmax[n] : vmin[n];
        if (side ? !(bv[side] > best) : !(bv[side] < best)) continue;
        best = bv[side];
        const long long x = bx[side];
        const int i = (int)(x % bk.nx1), j = (int)((x / bk.nx1) % bk.nx2),
                  k = (int)(x / ((long long)bk.nx1 * bk.nx2));
        double* xo = side ? xmax : xmin;
        xo[3 * n] = bk.x1min + (i + 0.5) * bk.dx1;
        xo[3 * n + 1] = bk.x2min + (j + 0.5) * bk.dx2;
        xo[3 * n + 2] = bk.x3min + (k + 0.5) * bk.dx3;
      }
    }
  }
  return CG_OK;
}

int cg_run_set_mlp(cg_run* r, const cg_mlp* net) {
  if (!r) return CG_ERR_ARG;
  if (r->sph) return set_err(r->ctx, CG_ERR_ARG, "not available for cubed-sphere runs");
  if (net && r->cfg.satadj != CG_SATADJ_MLP)
    return set_err(r->ctx, CG_ERR_ARG, "the run's config does not use the guided adjustment");
  r->mlp = net;
  drop_graph(r);
  return CG_OK;
}

int cg_run_get_tune(cg_run* r, cg_tune* out) {
  if (!r || !out) return CG_ERR_ARG;
  *out = r->tune;
  return CG_OK;
}

int cg_run_set_tune(cg_run* r, const cg_tune* t) {
  if (!r || !t) return CG_ERR_ARG;
  if ((t->implicit_wpb != 1 && t->implicit_wpb != 2 && t->implicit_wpb != 4) ||
      t->implicit_grid < 0 ||
      (t->flux_block != 64 && t->flux_block != 128 && t->flux_block != 256) ||
      t->face_variant < 0 || t->face_variant > 4 || t->march_seg < 1 ||
      t->march_seg > 4096 || (t->implicit_fused != 0 && t->implicit_fused != 1))
    return set_err(r->ctx, CG_ERR_ARG,
                   "implicit_wpb in {1,2,4}, implicit_grid >= 0, flux_block in "
                   "{64,128,256}, face_variant in 0..4, march_seg in [1,4096], "
                   "implicit_fused in {0,1}");
  if (!(r->cfg.implicit_flag & 8) && t->implicit_wpb != 1)
    return set_err(r->ctx, CG_ERR_ARG, "3x3 solve runs with implicit_wpb 1");
  r->tune = *t;
  drop_graph(r);
  return CG_OK;
}

// Times one graph-replayed step for each candidate launch shape, restoring
// the state after every trial, and keeps the fastest: first the batched
// column solve (waves per workgroup x grid: one workgroup per group set, or
// a persistent grid of 1, 2 or 4 workgroups per compute unit; sweeps fused
// per group or in two launches), then the workgroup size of the face
// kernels.
int cg_run_autotune(cg_run* r, double dt, cg_tune* out) {
  if (!r || !(dt > 0.0)) return CG_ERR_ARG;
  cg_ctx* ctx = r->ctx;
  if (r->sph) return set_err(r->ctx, CG_ERR_ARG, "not available for cubed-sphere runs");
  if (r->tm.on) return set_err(ctx, CG_ERR_ARG, "switch timing off before autotune");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // save the state a step reads
  struct Saved {
    double *u, *w, *psf, *psv;
  };
  std::vector<Saved> sv(r->blocks.size(), Saved{nullptr, nullptr, nullptr, nullptr});
  auto release = [&]() {
    for (Saved& s : sv) {
      double* ps[] = {s.u, s.w, s.psf, s.psv};
      for (double* p : ps)
        if (p) (void)hipFree(p);
    }
  };
  auto copy_all = [&](bool save) -> hipError_t {
    for (size_t b = 0; b < r->blocks.size(); ++b) {
      RunBlock& rb = r->blocks[b];
      size_t nc = sizeof(double) * cell_doubles(rb.g, r->N);
      size_t np = sizeof(double) * psf_doubles(rb.g);
      double* dev[] = {rb.u, rb.w, rb.psf, rb.psv};
      double* bak[] = {sv[b].u, sv[b].w, sv[b].psf, sv[b].psv};
      size_t sz[] = {nc, nc, np, np};
      for (int q = 0; q < 4; ++q) {
        hipError_t e = save ? hipMemcpyAsync(bak[q], dev[q], sz[q], hipMemcpyDeviceToDevice,
                                             ctx->stream)
                            : hipMemcpyAsync(dev[q], bak[q], sz[q], hipMemcpyDeviceToDevice,
                                             ctx->stream);
        if (e != hipSuccess) return e;
      }
    }
    return hipStreamSynchronize(ctx->stream);
  };
  for (size_t b = 0; b < r->blocks.size(); ++b) {
    RunBlock& rb = r->blocks[b];
    size_t nc = sizeof(double) * cell_doubles(rb.g, r->N);
    size_t np = sizeof(double) * psf_doubles(rb.g);
    if (hipMalloc((void**)&sv[b].u, nc) != hipSuccess ||
        hipMalloc((void**)&sv[b].w, nc) != hipSuccess ||
        hipMalloc((void**)&sv[b].psf, np) != hipSuccess ||
        hipMalloc((void**)&sv[b].psv, np) != hipSuccess) {
      release();
      return set_err(ctx, CG_ERR_HIP, "autotune: no memory for the state copy");
    }
  }
  if (copy_all(true) != hipSuccess) {
    release();
    return set_err(ctx, CG_ERR_HIP, "autotune: state copy failed");
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    release();
    return set_err(ctx, CG_ERR_HIP, "autotune: events");
  }
  int err = set_step(r, dt);
  auto trial = [&](const cg_tune& t, double* ms) -> int {
    r->tune = t;
    int e = capture(r);
    if (e) return e;
    if (hipGraphLaunch(r->gexec, ctx->stream) != hipSuccess) return CG_ERR_HIP;
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) return CG_ERR_HIP;
    const int reps = 2;
    for (int q = 0; q < reps; ++q)
      if (hipGraphLaunch(r->gexec, ctx->stream) != hipSuccess) return CG_ERR_HIP;
    if (hipEventRecord(e1, ctx->stream) != hipSuccess) return CG_ERR_HIP;
    if (hipEventSynchronize(e1) != hipSuccess) return CG_ERR_HIP;
    float f = 0.f;
    (void)hipEventElapsedTime(&f, e0, e1);
    *ms = f / reps;
    e = check_status(ctx);
    if (e) return e;
    return copy_all(false) == hipSuccess ? CG_OK : CG_ERR_HIP;
  };
  cg_tune best = r->tune;
  double best_ms = 1e300;
  std::vector<cg_tune> cands;
  const int wpbs[3] = {1, 2, 4};
  for (int wi = 0; wi < 3; ++wi) {
    int wpb = wpbs[wi];
    if (!(r->cfg.implicit_flag & 8) && wpb != 1) continue;
    const int grids[4] = {0, r->ncu * 4 / wpb, r->ncu * 2 / wpb, r->ncu / wpb};
    const bool cols = (r->cfg.implicit_flag & 1) != 0 && !r->x1_periodic;
    for (int gi = 0; gi < (cols ? 4 : 1); ++gi)
      for (int fu = 1; fu >= (cols ? 0 : 1); --fu) {
        cg_tune t = best;
        t.implicit_wpb = wpb;
        t.implicit_grid = grids[gi];
        t.implicit_fused = fu;
        t.flux_block = 256;
        cands.push_back(t);
      }
  }
  for (const cg_tune& t : cands) {
    if (err) break;
    double ms = 0.0;
    err = trial(t, &ms);
    if (err) break;
    if (ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  const int fbs[2] = {128, 64};
  for (int fi = 0; fi < 2 && !err; ++fi) {
    cg_tune t = best;
    t.flux_block = fbs[fi];
    double ms = 0.0;
    err = trial(t, &ms);
    if (err) break;
    if (ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  // face kernels of the mixed form: per-face lanes, or the cell-centric
  // kernels with march segments of 8, 16, 32 cells
  if (r->cfg.precision == CG_PREC_FP32) {
    const int vs[13][2] = {{0, 16}, {1, 8},  {1, 16}, {1, 32}, {2, 8},
                           {2, 16}, {2, 32}, {3, 16}, {3, 32}, {3, 64},
                           {4, 16}, {4, 32}, {4, 64}};
    for (int vi = 0; vi < 13 && !err; ++vi) {
      cg_tune t = best;
      t.face_variant = vs[vi][0];
      t.march_seg = vs[vi][1];
      if (t.face_variant == best.face_variant && t.march_seg == best.march_seg)
        continue;
      double ms = 0.0;
      err = trial(t, &ms);
      if (err) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!err && copy_all(false) != hipSuccess) err = CG_ERR_HIP;
  release();
  drop_graph(r);
  best.step_ms = best_ms;
  r->tune = best;
  if (err) return err == CG_ERR_HIP ? set_err(ctx, CG_ERR_HIP, "autotune trial failed") : err;
  if (out) *out = best;
  return CG_OK;
}

}  // extern "C"
