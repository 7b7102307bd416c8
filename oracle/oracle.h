/* oracle.h -- CPU FP64 oracle of the canoe dynamical-core step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the reference
 * the results are compared with or the timed CPU baseline -- never on the product path.
 *
 * Every function restates the reference algorithm on host memory laid out
 * as in include/canoe_gpu.h (AthenaArray (n,k,j,i), NGHOST = 3) and cites the
 * reference file:line it follows.  Arithmetic is plain IEEE double without
 * FMA contraction (built with -ffp-contract=off).
 */
#ifndef CANOE_ORACLE_H_
#define CANOE_ORACLE_H_

#include <stddef.h>

#include "../include/canoe_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* math */
double or_exp(double x);
double or_log(double x);
double or_get_gamma(const cg_thermo* th, const double* w);

/* reconstruction of one cell from its stencil: order 2 uses q[1..3] of a
 * 5-value window q[0..4] centred on q[2]; order 5 uses all five.  qm is the
 * value at the cell's lower face, qp at its upper face. */
void or_recon_cell(int order, const double* q, double* qm, double* qp);

/* Riemann solvers on n faces: wl, wr (N, ld), flx (N, fld). */
int or_riemann(const cg_thermo* th, int solver, int ivx, int n,
               const double* wl, const double* wr, ptrdiff_t ld, double* flx,
               ptrdiff_t fld);

/* Hydrostatic decomposition of one column of nc1 cells (ChangeToTemperature):
 * outputs psf (nc1 + 1), psv (nc1), rdv (nc1, the R used for T). */
int or_decompose_column(const cg_block* blk, const cg_thermo* th,
                        const double* w, int k, int j, double* psf,
                        double* psv, double* rdv);

int or_calculate_fluxes(const cg_block* blk, const cg_thermo* th,
                        const cg_config* cfg, const double* w, double* flux1,
                        double* flux2, double* flux3);
/* faces of the last or_calculate_fluxes with a negative density or pressure
 * (check_reconstruct, src/checks.cpp:79-97) */
long long or_face_failures(void);
int or_restore_outflow_ghosts(const cg_block* blk, const cg_thermo* th,
                              double* w);
int or_add_flux_divergence(const cg_block* blk, double wght,
                           const double* flux1, const double* flux2,
                           const double* flux3, double* du);
int or_add_sources(const cg_block* blk, const cg_thermo* th,
                   const cg_config* cfg, double dt, const double* w,
                   double* du);
int or_implicit_solve(const cg_block* blk, const cg_thermo* th, double* du,
                      const double* w, double dt, int flag);
int or_cons_to_prim(const cg_block* blk, const cg_thermo* th, const double* u,
                    double* w);
int or_prim_to_cons(const cg_block* blk, const cg_thermo* th, const double* w,
                    double* u);
int or_apply_bc(const cg_block* blk, double* w);
int or_saturation_adjust(const cg_thermo* th, int ncell, const double* rho,
                         double* ie, double* y, const double* t_guess,
                         int* iters);
/* Kessler microphysics + precipitation sedimentation over dt on the
 * conserved state u (interior cells), as applied after the last stage. */
int or_kessler(const cg_block* blk, const cg_thermo* th, const cg_config* cfg,
               double dt, double* u);
/* nsteps full SSPRK(3,3) steps of ImplicitHydroTasks on (u, w): u interior is
 * the state, w (with ghosts) is recomputed from u at the end of each stage. */
int or_step(const cg_block* blk, const cg_thermo* th, const cg_config* cfg,
            double dt, int nsteps, double* u, double* w);
/* Same step with OpenMP over rows/columns (nthreads <= 0: all cores). */
int or_step_omp(const cg_block* blk, const cg_thermo* th,
                const cg_config* cfg, double dt, int nsteps, double* u,
                double* w, int nthreads);
/* Hydro::NewBlockTimeStep of one block (cfl times the minimum signal-speed
 * step over interior cells) and Mesh::NewTimeStep's limits. */
double or_new_block_dt(const cg_block* blk, const cg_thermo* th,
                       const cg_config* cfg, const double* w, double cfl);
double or_new_mesh_dt(double dt_old, double blocks_dt, double time,
                      double tlim);
int or_sums(const cg_block* blk, const double* u, double* out);
/* State checks: 0 NaN, 1 negative, 2 too cold, 3 pass; where = (k, j, i). */
int or_check(const cg_block* blk, const cg_thermo* th, const double* w,
             int* where);
/* Athena++ default history sums (+ species) of one block: 8 + nmass. */
int or_history(const cg_block* blk, const double* u, double* out);
/* DebugOutput: per-variable min / max of w and their cell centres. */
int or_debug(const cg_block* blk, const double* w, double* vmin, double* vmax,
             double* xmin, double* xmax);
int or_block_thomas(int B, int n, const double* a, const double* b,
                    const double* c, const double* r, double* x);
int or_block_thomas_periodic(int B, int n, const double* a, const double* b,
                             const double* c, const double* corner,
                             const double* r, double* x);
int or_flux_jacobian(const cg_block* blk, double gm1, const double* w,
                     double* out);
int or_roe_matrices(const cg_block* blk, const cg_thermo* th, double gm1,
                    const double* wl, const double* wr, double* abs_a,
                    double* prim_out);

/* Cubed sphere (oracle_sphere.inc): six panels of the block `blk` (nx1
 * levels from radius x1min, nx2 = nx3 cells per panel side, dry), u and w
 * six consecutive AthenaArrays.  or_sph_step: nsteps SSPRK(3,3) steps;
 * or_sph_fill_ghosts: w's x2 / x3 ghosts from the neighbours' u and the x1
 * walls; or_sph_prim_to_cons: interior u from w. */
int or_sph_step(const cg_block* blk, const cg_thermo* th, const cg_config* cfg,
                const cg_sphere* sp, double dt, int nsteps, double* u, double* w);
int or_sph_fill_ghosts(const cg_block* blk, const cg_thermo* th, const double* u,
                       double* w);
int or_sph_prim_to_cons(const cg_block* blk, const cg_thermo* th, const double* w,
                        double* u);
/* Coriolis and Held-Suarez terms of the six panels added to du6 over wght */
int or_sph_forcing(const cg_block* blk, const cg_thermo* th, const cg_sphere* sp, double wght,
                   const double* w6, double* du6);
int or_sph_metric(const cg_block* blk, int p, int what, double* out);
int or_sph_link(int p, int s, int* q, int* sq, int* inv);
void or_sph_latlon(int p, double X, double Y, double* lat, double* lon);
void or_sph_rll_from_abp(double X, double Y, int block, double v2, double v3, double* U,
                         double* V);
void or_sph_abp_from_rll(double X, double Y, int block, double U, double V, double* v2,
                         double* v3);

/* Saturation-guess network (oracle_mlp.cpp).  Parameters: or_mlp_nparam()
 * floats, layout W1 (64x4), b1, W2 (64x64), b2, W3 (64x64), b3, W4 (3x64),
 * b4.  emulate = 1 rounds as the device does (bf16 operands, fp32 sums). */
int or_mlp_nparam(void);
float or_mlp_gelu(float z);
float or_mlp_gelu_grad(float z);
float or_mlp_exp(float x);
unsigned short or_mlp_bf16(float x);
int or_mlp_forward(const float* params, int n, const float* x, double* y,
                   int emulate);
int or_mlp_loss_grad(const float* params, int n, const float* x,
                     const float* t, double* grad, double* loss, int emulate);
/* Training pairs from base states (rho, ie, y as in or_saturation_adjust) and
 * network inputs of cells; x is (n, 4), t is (n, 3), tu the all-vapour T. */
int or_mlp_pairs(const cg_thermo* th, int nbase, const double* rho,
                 const double* ie, const double* y, int npairs,
                 unsigned long long seed, float* x, float* t, int* iters);
int or_mlp_features(const cg_thermo* th, int ncell, const double* rho,
                    const double* ie, const double* y, float* x, double* tu);

#ifdef __cplusplus
}
#endif

#endif /* CANOE_ORACLE_H_ */
