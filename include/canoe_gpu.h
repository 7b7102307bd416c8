/* canoe_gpu.h -- C-ABI of the MI355X-native canoe dynamical-core step.
 *
 * Plain C types only (no torch, no HIP types): a host program written
 * against Athena++/canoe binds these through cgo/ctypes/a C++ shim.
 *
 * Conventions (mirroring the reference's AthenaArray contract):
 *   - A cell-centred array is laid out (n, k, j, i) with i (x1, the vertical)
 *     fastest, including NGHOST=3 ghost layers in every active direction:
 *     nc1 = nx1 + 6, nc2 = nx2 > 1 ? nx2 + 6 : 1, nc3 = nx3 > 1 ? nx3 + 6 : 1.
 *   - Variable index: IDN = 0 (total density in primitives, dry density in
 *     conserved variables), 1..nmass the mass species (mass fractions in
 *     primitives, partial densities in conserved), IVX = 1 + nmass, IVY, IVZ,
 *     IPR = IEN = IVX + 3.  N = 5 + nmass.
 *   - Face fluxes: flux1 (N, nc3, nc2, nc1 + 1), flux2 (N, nc3, nc2 + 1, nc1),
 *     flux3 (N, nc3 + 1, nc2, nc1); face i lies between cells i-1 and i.
 *   - Every pointer argument is a DEVICE pointer unless the function name ends
 *     in _host.  Every function returns 0 (CG_OK) or a CG_ERR_* code and never
 *     aborts the process.  A context owns one HIP stream; calls on different
 *     contexts may run concurrently, calls on one context are serialised.
 */
#ifndef CANOE_GPU_H_
#define CANOE_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CG_NGHOST 3
#define CG_MAX_NMASS 3

enum cg_status {
  CG_OK = 0,
  CG_ERR_ARG = 1,       /* bad argument / unsupported combination */
  CG_ERR_HIP = 2,       /* HIP runtime error (launch, allocation, copy) */
  CG_ERR_NAN = 3,       /* NaN produced (flux, state, solve) */
  CG_ERR_NEGATIVE = 4,  /* negative density or pressure */
  CG_ERR_PIVOT = 5,     /* singular 5x5 block in the implicit solve */
  CG_ERR_CONVERGE = 6,  /* saturation-adjustment Newton did not converge */
  CG_ERR_SIZE = 7,      /* size limit exceeded */
  CG_ERR_NODEV = 8,     /* no HIP device / library built without one */
  CG_ERR_COLD = 9,      /* temperature below |g| dx1 / Rd */
  CG_ERR_IO = 10        /* a file could not be opened, read or written */
};

enum cg_bc {
  CG_BC_REFLECT = 0,
  CG_BC_PERIODIC = 1,
  CG_BC_OUTFLOW = 2,
  CG_BC_PANEL = 3 /* x2 / x3 side of a cubed-sphere panel: the neighbour panel */
};
enum cg_rsolver { CG_RS_LMARS = 0, CG_RS_HLLC = 1 };
enum cg_precision { CG_PREC_FP64 = 0, CG_PREC_FP32 = 1 };
enum cg_satadj { CG_SATADJ_OFF = 0, CG_SATADJ_NEWTON = 1, CG_SATADJ_MLP = 2 };

/* Species thermodynamics of the ideal-moist EOS (Li 2019 eq. 63).  Replaces
 * the snap::IdealMoist / kintera tensors that get_gamma() reads
 * (src/interface/eos.hpp:47-67) and the svp parameters of the condensation
 * reaction (examples/2019-Li-snap/earth-moist.yaml:37). */
typedef struct cg_thermo {
  int nmass;   /* number of mass species (vapours first, then condensates) */
  int nvapor;  /* species 1..nvapor are vapours, nvapor+1..nmass condensates */
  double gammad;                          /* dry cp/cv */
  double rd;                              /* dry gas constant, J/kg/K */
  double cv_ratio_m1[CG_MAX_NMASS];       /* cv_n / cv_d - 1 */
  double inv_mu_ratio_m1[CG_MAX_NMASS];   /* mu_d / mu_n - 1 (vapours) */
  /* saturation adjustment: species ivap (vapour) <-> icld (condensate),
   * 1-based species indices, 0 = no adjustment */
  int ivap, icld;
  double t3, p3, beta, delta; /* svp = p3 exp((1 - t3/T) beta - delta ln(T/t3)) */
  double latent;              /* internal-energy offset of condensates, J/kg */
  double tol;                 /* Newton tolerance on T, K */
  int max_iter;
} cg_thermo;

/* One MeshBlock: uniform Cartesian spacing, constant gravity along x1. */
typedef struct cg_block {
  int nx1, nx2, nx3;         /* interior cells (x1 = vertical) */
  int nmass;                 /* must equal cg_thermo.nmass */
  double x1min, x2min, x3min;
  double dx1, dx2, dx3;      /* for nx3 == 1, dx3 is the slab thickness */
  double grav;               /* g1, acceleration along +x1 (negative = down) */
  int bc[6];                 /* ix1, ox1, ix2, ox2, ix3, ox3 (enum cg_bc) */
} cg_block;

/* Numerics of one step (the reference's <time>/<hydro> input blocks). */
typedef struct cg_config {
  int order_h;        /* horizontal reconstruction: 2 = PLM, 5 = WENO5 */
  int order_v;        /* vertical (x1) reconstruction: 2 = PLM, 5 = WENO5 */
  int rsolver;        /* enum cg_rsolver (the RSOLVER link-time choice) */
  int implicit_flag;  /* 0 = explicit, 9 = 5x5 full x1, 1 = 3x3 partial x1 */
  int precision;      /* enum cg_precision: FP64, or FP32 fluxes + FP64 state */
  int satadj;         /* enum cg_satadj (last RK stage) */
  double visc;        /* constant viscosity/diffusivity K, m^2/s (0 = off) */
  double p0;          /* reference pressure for potential temperature */
  double heat_flux;   /* W/m^2 deposited in the lowest interior layer */
  /* Kessler warm-rain microphysics and precipitation sedimentation, applied
   * after the last stage (all zero = off; needs vapour, cloud and a third
   * condensate, the precipitation): rates of autoconversion cloud -> precip,
   * accretion, evaporation of precip (1/s; examples/2019-Li-snap/
   * earth-precip.yaml:44-53, water-kessler.yaml:4-18) and the fall speed of
   * precipitation along -x1 (m/s). */
  double k_auto, k_accr, k_evap, v_fall;
} cg_config;

/* Cubed-sphere run (config 3; src/exo3, examples/2023-Chen-exo3/hs94.cpp):
 * rotation and the Held-Suarez (1994) forcing of hs94.cpp:56-151 (Newtonian
 * relaxation toward T_eq and Rayleigh drag below sigma_b). */
typedef struct cg_sphere {
  double omega;   /* planetary rotation, 1/s (Coriolis force; 0 = off) */
  int hs94;       /* 1: Held-Suarez relaxation and drag, 0: off */
  double ts;      /* T_eq at the surface and equator, K (Ts) */
  double dt_eq;   /* equator-pole temperature difference, K (dT) */
  double dtheta;  /* vertical potential temperature difference, K */
  double p0;      /* reference surface pressure, Pa */
  double kf, ka, ks; /* drag, free-atmosphere and surface relaxation rates, 1/s */
  double sigmab;  /* top of the frictional boundary layer, p / p0 */
} cg_sphere;

typedef struct cg_ctx cg_ctx;

/* ---- context ------------------------------------------------------------ */
/* New context on HIP device `device` with its own stream.  No reference
 * counterpart (Athena++ state lives in MeshBlock; this owns the stream). */
int cg_create(int device, cg_ctx** out);
int cg_destroy(cg_ctx* ctx);
int cg_synchronize(cg_ctx* ctx);
/* Last error text of this context (static storage owned by the context). */
const char* cg_last_error(cg_ctx* ctx);
/* Device allocation helpers so a C host needs no HIP headers. */
int cg_malloc(cg_ctx* ctx, size_t bytes, void** ptr);
int cg_free(cg_ctx* ctx, void* ptr);
int cg_memcpy_h2d(cg_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy through a streaming kernel (the copy used to
 * calibrate HBM traffic counters). */
int cg_memcpy_d2d(cg_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Event timing on the context's stream (no reference counterpart): start
 * records an event, stop records a second one, waits for it and returns the
 * device time between the two in milliseconds; the work enqueued between the
 * calls is what is timed, without host launch or copy overheads. */
int cg_timer_start(cg_ctx* ctx);
int cg_timer_stop(cg_ctx* ctx, double* ms);
int cg_memcpy_d2h(cg_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- Riemann plugin ----------------------------------------------------- *
 * Replaces void Hydro::RiemannSolver(int k, int j, int il, int iu, int ivx,
 *   AthenaArray<Real>& wl, AthenaArray<Real>& wr, AthenaArray<Real>& flx,
 *   const AthenaArray<Real>& dxw)          (src/snap/riemann/lmars.cpp:19-22,
 *                                           src/snap/riemann/hllc_transform.cpp:25-28)
 * wl, wr: (N, ld) left/right primitive states of n_faces faces; flx: (N,
 * flx_ld).  ivx in {IVX, IVY, IVZ} selects the normal velocity.  FP64. */
int cg_riemann(cg_ctx* ctx, const cg_thermo* th, int solver, int ivx,
               int n_faces, const double* wl, const double* wr, ptrdiff_t ld,
               double* flx, ptrdiff_t flx_ld);

/* ---- flux sweep --------------------------------------------------------- *
 * Replaces Hydro::CalculateFluxes(w, b, bcc, order) as patched by canoe
 * (patches/19.calculate_fluxes.patch:19-185): hydrostatic decomposition
 * (Decomposition::ChangeToTemperature / RestoreFromTemperature,
 * src/snap/decomposition/change_to_temperature.cpp:40-167) around the x1
 * reconstruction, then the Riemann solver along x1, x2, x3.  `w` must have
 * its ghost cells filled.  Pure: `w` is not modified (the reference mutates
 * and restores it, which leaves the ghosts of an outflow x1 side rewritten
 * with the boundary cell's temperature and pressure perturbation,
 * change_to_temperature.cpp:113-127 and :150-153; runs apply that rewrite
 * where the stage reads those ghosts).  Fluxes are computed on every face
 * of interior cells.  flux2 / flux3 are not read or written for a direction
 * without cells (nx2 = 1 / nx3 = 1) and may then be null.  Every face state
 * is checked as check_reconstruct does (src/checks.cpp:79-97, called at
 * patches/19.calculate_fluxes.patch:45, 73, 117): a negative reconstructed
 * density or pressure returns CG_ERR_NEGATIVE (the fluxes are still
 * written), where the reference built with glog aborts. */
int cg_calculate_fluxes(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                        const cg_config* cfg, const double* w, double* flux1,
                        double* flux2, double* flux3);

/* Replaces Hydro::AddFluxDivergence(wght, du) (upstream; called from
 * src/tasklist/implicit_hydro_tasks.cpp:163-164): du -= wght * div(F). */
int cg_add_flux_divergence(cg_ctx* ctx, const cg_block* blk, double wght,
                           const double* flux1, const double* flux2,
                           const double* flux3, double* du);

/* ---- implicit solve ----------------------------------------------------- *
 * Replaces void ImplicitSolver::SolveImplicit3D(AthenaArray<Real>& du,
 *   AthenaArray<Real>& w, Real dt)     (src/snap/implicit/solve_implicit_3d.cpp:19-20)
 * with implicit_flag semantics (implicit_solver.cpp:47): bit 0 = x1 columns,
 * bit 1 = x2 lines, bit 2 = x3 lines (solved in the order x3, x2, x1, each
 * on the du left by the previous one), bit 3 = FullCorrection 5x5
 * (full_correction.cpp) instead of PartialCorrection 3x3
 * (partial_correction.cpp).  A direction whose two boundaries are periodic
 * is solved by the cyclic recurrences of PeriodicForwardSweep /
 * PeriodicBackwardSubstitution (periodic_forward_backward.hpp:14-232), the
 * others with reflecting walls folded in.  x2 / x3 lines write back du of
 * IDN, the normal momentum and IEN only, as the reference does.  Between
 * walls their density increment is split over the species with w read where
 * the reference reads it: at the loop indices of its shuffled line buffer
 * (forward_backward.hpp:109-115 with solve_implicit_3d.cpp:28-31, 59-62),
 * i.e. the AthenaArray entry ((n nc3 + K) nc2 + J) nc1 + I, (K, J, I) =
 * (x1, x3, x2) of the cell on x2 lines and (x2, x1, x3) on x3 lines;
 * CG_ERR_ARG where that entry lies past the end of w.  `w` must have
 * the ghost cells of every solved direction filled.  du is updated in
 * place; all lines of a direction in one launch. */
int cg_implicit_solve(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                      double* du, const double* w, double dt, int flag);

/* No counterpart in the reference.  cg_implicit_solve without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the
 * archive / line scratch is a fixed array of CG_NOALLOC_SCRATCH_BYTES per
 * device, held from the first cg_create on that device until the process
 * ends.  A solve that needs more scratch (30 doubles per cell for x1
 * columns between walls; 30 or 105 doubles per cell of the largest line
 * direction otherwise, e.g. 64 MiB = 279,620 cells of 5x5 x1 columns) is
 * rejected with CG_ERR_ARG before any launch.  Every call on one device
 * shares that array, which keeps the last call's contents (nothing in it
 * needs to start at zero: each level is written before it is read): calls
 * that overlap on one device, from any contexts or threads, race.  Results
 * equal cg_implicit_solve's bit for bit. */
#define CG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int cg_implicit_solve_noalloc(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                              double* du, const double* w, double dt, int flag);

/* ---- host-array variants ------------------------------------------------ *
 * The same calls on host arrays of the same layouts, for a caller that keeps
 * its AthenaArrays in host memory (the reference's own buffers: wl_ / wr_
 * rows and flx of lmars.cpp:19-22, w and the flux arrays of
 * Hydro::CalculateFluxes, du and w of solve_implicit_3d.cpp:19-20).  Each
 * copies its inputs to the device, runs the device call and copies the
 * outputs back before returning; entries the device call leaves alone
 * (faces past n_faces, faces of ghost cells) keep the caller's values. */
int cg_riemann_host(cg_ctx* ctx, const cg_thermo* th, int solver, int ivx,
                    int n_faces, const double* wl, const double* wr, ptrdiff_t ld,
                    double* flx, ptrdiff_t flx_ld);
int cg_calculate_fluxes_host(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                             const cg_config* cfg, const double* w, double* flux1,
                             double* flux2, double* flux3);
int cg_implicit_solve_host(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                           double* du, const double* w, double dt, int flag);

/* ---- EOS and boundaries ------------------------------------------------- *
 * Replace EquationOfState::ConservedToPrimitive / PrimitiveToConserved
 * (src/snap/eos/ideal_moist_hydro.cpp:47-81) on interior cells, in the
 * closed form of src/snap/thermodynamics/thermodynamics_athena.hpp:153-160. */
int cg_cons_to_prim(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                    const double* u, double* w);
int cg_prim_to_cons(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                    const double* w, double* u);
/* Kessler microphysics + precipitation sedimentation over the step dt on
 * caller arrays: u (conserved, interior cells) updated in place, w
 * (primitive) rewritten from it on interior cells.  No counterpart in the
 * reference's live stage (the sedimentation hook of UpdateAllConserved is
 * commented out, src/tasklist/implicit_hydro_tasks.cpp:292, and the
 * reaction rates are kintera's, src/snap/riemann/noflux.cpp:61-71 gives the
 * sedimentation mass flux); runs apply it after the last stage when a rate
 * of cfg is non-zero.  CG_ERR_ARG without a precipitation species. */
int cg_microphysics(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                    const cg_config* cfg, double dt, double* u, double* w);
/* Fills the ghost cells of a primitive array (reflecting / periodic /
 * outflow), as the upstream BoundaryValues do after CONS2PRIM. */
int cg_apply_bc(cg_ctx* ctx, const cg_block* blk, double* w);

/* ---- time step ------------------------------------------------------------ *
 * Replaces void Hydro::NewBlockTimeStep() (upstream athenapp
 * hydro/new_blockdt.cpp as patched by patches/21.new_blockdt.patch:16-45)
 * on a caller primitive array w: cfl times the minimum over interior cells
 * and active directions of dx / (|v| + c_s) (moist sound speed,
 * src/snap/eos/ideal_moist_hydro.cpp:87-91), or dx / (cfl |v|) in a
 * direction cfg->implicit_flag solves implicitly (x1 only when nx2 > 1).
 * dt_out: host double. */
int cg_new_block_dt(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                    const cg_config* cfg, const double* w, double cfl,
                    double* dt_out);

/* ---- saturation-guess network -------------------------------------------- *
 * No counterpart in the reference (its adjustment, kintera ThermoY::forward
 * called at src/tasklist/implicit_hydro_tasks.cpp:308, always starts from the
 * all-vapour state).  A 4 -> 64 -> 64 -> 64 -> 3 perceptron with tanh-GELU,
 * bf16 weights and activations with fp32 accumulation on the matrix cores,
 * maps a cell's inputs
 *   x = ((T_u - 260)/40, ln(p_u/1e5), 40 q_t, 40 q_p)
 * (T_u the all-vapour temperature, p_u = rho T_u (R_d q_d + R_v q_t), q_t
 * total water, q_p the remaining species) to
 *   y = ((T - T_u)/10, q_v/q_t, q_c/q_t)
 * of the adjusted state; T_u + 10 y0 starts the Newton iteration of the
 * guided adjustment.  Parameters: CG_MLP_NPARAM fp32 values, layout W1 (64x4),
 * b1 (64), W2 (64x64), b2, W3 (64x64), b3, W4 (3x64), b4 (3), W row-major
 * (out, in).  A network belongs to one context. */
#define CG_MLP_NPARAM 8835
typedef struct cg_mlp cg_mlp;
typedef struct cg_mlp_train_cfg {
  int batch;           /* samples per Adam step (<= 2^20) */
  int steps;           /* Adam steps */
  float lr, lr_final;  /* cosine schedule from lr to lr_final */
  unsigned long long seed;  /* batch draws */
} cg_mlp_train_cfg;
/* Glorot-uniform weights from seed, zero biases. */
int cg_mlp_create(cg_ctx* ctx, unsigned long long seed, cg_mlp** out);
int cg_mlp_destroy(cg_mlp* net);
/* Host fp32 parameter vectors; set resets the Adam moments. */
int cg_mlp_set_params(cg_mlp* net, const float* host_params);
int cg_mlp_get_params(cg_mlp* net, float* host_params);
/* y (n, 3) = net(x (n, 4)), device arrays. */
int cg_mlp_forward(cg_mlp* net, long long n, const float* x, float* y);
/* Mean squared error over the n pairs (x, t) and its gradient (device,
 * CG_MLP_NPARAM floats, nullable); loss_host nullable. */
int cg_mlp_loss_grad(cg_mlp* net, long long n, const float* x, const float* t,
                     float* grad, double* loss_host);
/* cfg->steps Adam steps on batches drawn from the npairs device pairs;
 * loss_host (nullable) receives cfg->steps batch losses. */
int cg_mlp_train(cg_mlp* net, const cg_mlp_train_cfg* cfg, long long npairs,
                 const float* x, const float* t, float* loss_host);
/* Training pairs from nbase device states (rho, ie, y laid out as in
 * cg_saturation_adjust): pair p takes base cell floor(u0 nbase), scales its
 * total water by 0.8 + 1.7 u1 into vapour, shifts its all-vapour temperature
 * by 6 u2 - 3 K at fixed density (u0..u2 from a splitmix64 stream of seed
 * and p) and records the plain adjustment of that state.  x (npairs, 4),
 * t (npairs, 3), iters (nullable) the plain Newton counts. */
int cg_mlp_make_pairs(cg_ctx* ctx, const cg_thermo* th, int nbase,
                      const double* rho, const double* ie, const double* y,
                      long long npairs, unsigned long long seed, float* x,
                      float* t, int* iters);

/* ---- saturation adjustment ---------------------------------------------- *
 * Replaces kintera pthermo->forward(rho, ie, yfrac)
 * (src/tasklist/implicit_hydro_tasks.cpp:300-312) on ncell cells: rho total
 * density, ie internal energy density (in/out), y (nmass, ncell) mass
 * fractions (in/out).  Newton on T to th->tol.  With a network (net non-NULL)
 * a wavefront holding a saturated cell starts Newton from the network's
 * guess; iteration and tolerance are unchanged.  iters (nullable) receives
 * the Newton iteration count per cell (0 for sub-saturated cells). */
int cg_saturation_adjust(cg_ctx* ctx, const cg_thermo* th, int ncell,
                         const double* rho, double* ie, double* y,
                         const cg_mlp* net, int* iters);

/* ---- whole step on a device-resident block ------------------------------ *
 * A run owns u, w, u1, du and the scratch of one MeshBlock.  One step is
 * the three SSPRK(3,3) stages of ImplicitHydroTasks
 * (src/tasklist/implicit_hydro_tasks.cpp:134-313,
 *  src/integrator/integrators.cpp:14-44), captured once as a hipGraph. */
typedef struct cg_run cg_run;
int cg_run_create(cg_ctx* ctx, const cg_block* blk, const cg_thermo* th,
                  const cg_config* cfg, cg_run** out);
int cg_run_destroy(cg_run* run);
/* Primitive state (with or without ghosts filled) from host memory. */
int cg_run_set_prim_host(cg_run* run, const double* w_host);
int cg_run_get_cons_host(cg_run* run, double* u_host);
int cg_run_get_prim_host(cg_run* run, double* w_host);
/* nsteps steps of size dt (use_graph != 0: replay the captured graph).  The
 * stage kernels read the step from device memory, so one captured graph
 * serves every dt.  CG_ERR_NEGATIVE: a negative reconstructed face density
 * or pressure in some stage (the face checks of cg_calculate_fluxes). */
int cg_run_step(cg_run* run, double dt, int nsteps, int use_graph);
/* One SSPRK(3,3) stage (0, 1 or 2) of the run's blocks: the per-stage tasks
 * of ImplicitHydroTasks (src/tasklist/implicit_hydro_tasks.cpp:134-313) --
 * fluxes, AddFluxDivergence with weight beta_s dt and the sources, the
 * implicit x3 / x2 / x1 solves, WeightedAve and the update, the saturation
 * adjustment on stage 2, ghost cells.  The step is read from device memory
 * at dt_dev (one double) when dt_dev is not null, else it is dt.  Stages go
 * in order 0, 1, 2 per step.  Asynchronous on the run's stream: kernel
 * status (zero pivots) is reported by the next cg_synchronize or
 * cg_run_step.  After the first call on a run (which builds its launch
 * tables) the call launches kernels only and may be captured into a
 * caller's hipGraph. */
int cg_run_stage(cg_run* run, int stage, double dt, const double* dt_dev);
/* The minimum of cg_new_block_dt over the run's blocks, reduced on the
 * device from the resident state (Mesh::NewTimeStep's min over blocks). */
int cg_run_new_dt(cg_run* run, double cfl, double* dt_out);
/* State checks of the resident w (check_hydro_variables,
 * src/snap/eos/check_hydro_variables.cpp:21-64; check_eos_cons2prim,
 * src/checks.cpp:52-75): CG_ERR_NAN for a NaN, CG_ERR_NEGATIVE for a
 * negative density, mass fraction or pressure, CG_ERR_COLD for a temperature
 * below |g| dx1 / Rd, else CG_OK.  where (nullable, 4 ints): block, and the
 * interior k, j, i of the first failing cell in (block, k, j, i) order, or
 * -1.  cg_run_advance runs the same checks after every step. */
int cg_run_check(cg_run* run, int* where);
/* The main integration loop of main.cpp:114-152 on the device: while
 * *time < tlim and (nlim < 0 or *ncycle < nlim), one step of size *dt, then
 * ++*ncycle, *time += *dt, and the next step as in Mesh::NewTimeStep
 * (upstream athenapp): min(2 *dt, the blocks' NewBlockTimeStep), cut to end
 * at tlim.  *dt <= 0 on entry starts from the blocks' step.  Per cycle the
 * host reads back one double (and the status word); the state stays on the
 * device. */
int cg_run_advance(cg_run* run, double cfl, double tlim, long long nlim,
                   int use_graph, double* time, double* dt, long long* ncycle);
/* Device pointers of the run's state (for in-place use by a host code). */
int cg_run_device_state(cg_run* run, double** u, double** w);
/* FP64 global sums of the conserved variables over interior cells times the
 * cell volume (the reference's `hst` history totals).  out: N doubles. */
int cg_run_sums(cg_run* run, double* out_host);
/* Athena++'s default history columns (upstream outputs/history.cpp), summed
 * over every block of the run: volume-weighted u(IDN), the three momenta,
 * the three kinetic energies 0.5 m_d^2 / u(IDN) and u(IEN), then the species
 * densities u(1..nmass).  out: 8 + nmass doubles. */
int cg_run_history(cg_run* run, double* out_host);
/* Replaces DebugOutput::WriteOutputFile (src/outputs/debug.cpp:30-84): per
 * variable n of w, the minimum and maximum over the interior cells of every
 * block and the cell-centre coordinates (x1, x2, x3) where they occur
 * (xmin / xmax: 3 N doubles); the reference's initial values FLT_MAX /
 * FLT_MIN and strict comparisons in block, k, j, i order. */
int cg_run_debug(cg_run* run, double* vmin, double* vmax, double* xmin,
                 double* xmax);
/* Replaces the netcdf output (NetCDFOutput::WriteOutputFile,
 * src/outputs/netcdf.cpp:100-420) for one block, written in NetCDF's classic
 * 64-bit-offset format: dimensions time (record), x1, x1f, x2, x2f, x3, x3f
 * (a face dimension only for a direction with more than one cell),
 * coordinate variables with axis / units / long_name / domain_decomposition,
 * and one float variable (time, x1, x2, x3) per field with its units and
 * long_name.  fields[f] is a host array (k, j, i) of the block with ghost
 * cells; its interior is written.  append = 0 creates the file with one
 * record; append != 0 adds a record to a file this function created with
 * the same block and fields.  Host code only: no device is used. */
int cg_netcdf_write(const char* path, const cg_block* blk, double time, int nfield,
                    const char* const* names, const char* const* units,
                    const char* const* long_names, const double* const* fields,
                    int append);
/* Per-kernel timing: with timing on, cg_run_step launches without the graph
 * and records HIP events around every kernel on the run's stream.
 * cg_run_kernel_stats returns the accumulated milliseconds, the number of
 * launches and the kernel name of kernel id 0..CG_NUM_KERNEL_IDS-1 (ids
 * with no launches report 0).  cg_run_set_timing resets the counters. */
#define CG_NUM_KERNEL_IDS 18
int cg_run_set_timing(cg_run* run, int on);
int cg_run_kernel_stats(cg_run* run, int kernel_id, double* total_ms,
                        long long* launches, const char** name);

/* ---- batches of MeshBlocks ---------------------------------------------- *
 * A batch is a run over nblocks independent blocks of any sizes (each with
 * its own boundary flags), sharing thermo and config.  Replaces the loop
 * over the MeshBlocks of a rank in Mesh::UpdateOneStep / the task list
 * (src/tasklist/implicit_hydro_tasks.cpp:134-313, one ImplicitSolver per
 * block): every column of every block is solved in one batched launch per
 * sweep.  cg_run_create is a batch of one block; the cg_run_* calls without
 * a block index act on block 0 (cg_run_step on all blocks). */
int cg_batch_create(cg_ctx* ctx, int nblocks, const cg_block* blks,
                    const cg_thermo* th, const cg_config* cfg, cg_run** out);
int cg_run_nblocks(cg_run* run, int* nblocks);
int cg_run_block_set_prim_host(cg_run* run, int block, const double* w_host);
int cg_run_block_get_cons_host(cg_run* run, int block, double* u_host);
int cg_run_block_get_prim_host(cg_run* run, int block, double* w_host);
int cg_run_block_sums(cg_run* run, int block, double* out_host);

/* Launch shape of a run: waves per workgroup (1, 2, 4) and grid (0 = one
 * workgroup per set of column groups, else a persistent grid of that many
 * workgroups) of the batched column solve, whether each wave runs a group's
 * forward and downward sweeps back to back in one launch (implicit_fused =
 * 1) or all forward sweeps then all downward sweeps (0), workgroup size
 * (64, 128, 256) of
 * the face and divergence kernels, and for the FP32-flux form the face
 * kernel variant (0: one lane per face in every direction; 1: one WENO
 * evaluation per cell, x2 across lanes, x3 and x1 marched over segments of
 * march_seg cells, x1 fused with the divergence; 2: as 1 for x2 / x3, one
 * lane per x1 face and a divergence kernel; 3 / 4: all faces and the
 * divergence in one LDS tile kernel of 32 x 8 / 16 x 16 columns marching up
 * segments of march_seg levels, 3D blocks only, else as 2); step_ms is the
 * step time
 * measured by the autotuner.  No counterpart in the reference (CPU loop
 * nests). */
typedef struct cg_tune {
  int implicit_wpb;
  int implicit_grid;
  int flux_block;
  int face_variant;
  int march_seg;
  int implicit_fused;
  double step_ms;
} cg_tune;
/* Network of a run whose config has satadj = CG_SATADJ_MLP (NULL: plain
 * Newton).  The guided adjustment then runs as its own pass at the end of
 * the last stage.  No counterpart in the reference. */
int cg_run_set_mlp(cg_run* run, const cg_mlp* net);
int cg_run_get_tune(cg_run* run, cg_tune* out);
int cg_run_set_tune(cg_run* run, const cg_tune* tune);
/* Times one graph-replayed step of size dt per candidate shape (the state is
 * restored after every trial) and keeps the fastest. */
int cg_run_autotune(cg_run* run, double dt, cg_tune* out);

/* ---- cubed sphere (config 3) ------------------------------------------- *
 * Six gnomonic-equiangle panels (src/exo3/gnomonic_equiangle.cpp, panel
 * numbering of src/exo3/cs_find_block_id.cpp:11-70 and positions of
 * cubed_sphere.cpp:78-89) as the six blocks of one run: `panel` describes
 * each (nx1 levels from the inner radius x1min with spacing dx1; nx2 = nx3
 * cells of pi / (2 nx2) per side; reflecting x1 walls; x2 / x3 sides
 * CG_BC_PANEL).  The step is the reference's CUBED_SPHERE path: WENO5 x1
 * reconstruction, HLLC with the panel frames (hllc_transform.cpp), the x2 /
 * x3 face states exchanged across panel edges (cubed_sphere.cpp:250-565),
 * ghost zones interpolated from the neighbour panels
 * (cubed_sphere_utility.cpp:80-400), areas, volumes and geometric source
 * terms of the panel metric, covariant momenta in u and contravariant
 * velocities in w (patches/26.adiabatic_hydro.patch), the x1 column solve
 * with the panel areas (implicit_flag 0, 1 or 9), gravity and the Coriolis /
 * Held-Suarez forcing of cg_sphere.  Dry gas; cfg->precision CG_PREC_FP32
 * computes the face states and Riemann fluxes of all three directions in
 * FP32 and keeps the fluxes in FP32 (the x1 decomposition, the face
 * pressures of the x1 momentum flux, divergence, column solve and state in
 * FP64).  The cg_run_* calls
 * taking a block index address panel b (0-5: panels 1-6); the time-step,
 * check, history, debug, sums and autotune calls return CG_ERR_ARG. */
int cg_sphere_create(cg_ctx* ctx, const cg_block* panel, const cg_thermo* th,
                     const cg_config* cfg, const cg_sphere* sp, cg_run** out);

/* ---- problem generators and host boundaries (host code) ----------------- *
 * The configurations' MeshBlock::ProblemGenerator and input-file settings
 * (examples/2019-Li-snap/straka.cpp:220-268 and straka.inp for config 1;
 * the dry and moist boxes and one block of the uneven-column batch for
 * configs 2, 4 and 5).  cg_pgen_setup fills the block, species constants and
 * run settings (nx* = 0: the configuration's own size; other sizes keep the
 * grid spacing); cg_pgen_init writes the seeded primitive state, interior and
 * ghosts, into a host array.  cg_apply_bc_host fills the ghost cells of a
 * host primitive array in the order of the device boundary kernels (x1, x2,
 * x3). */
int cg_pgen_setup(int config, int nx1, int nx2, int nx3, cg_block* blk,
                  cg_thermo* th, cg_config* cfg);
int cg_pgen_init(int config, uint64_t seed, const cg_block* blk,
                 const cg_thermo* th, const cg_config* cfg, double* w);
int cg_apply_bc_host(const cg_block* blk, double* w);
/* Config 3 (the cubed-sphere Held-Suarez GCM step; examples/2023-Chen-exo3/
 * hs94.cpp, hs94.inp): cg_pgen_setup(3, ...) gives the panel block (nx1 = 64
 * levels over 25 km from R = 6.371e6 m, 96 x 96 cells per panel, g = 9.81),
 * dry air and the run settings; cg_pgen_sphere fills the rotation and
 * Held-Suarez parameters of hs94.inp; cg_pgen_sphere_init writes the six
 * panels' interior primitives (isothermal 300 K hydrostatic atmosphere at
 * rest plus seeded 0.1 K white noise), six consecutive arrays. */
int cg_pgen_sphere(cg_sphere* sp);
int cg_pgen_sphere_init(uint64_t seed, const cg_block* panel, const cg_thermo* th,
                        double* w6);

/* Library build string (arch, flags). */
const char* cg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CANOE_GPU_H_ */
