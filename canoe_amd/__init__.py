"""canoe_amd -- MI355X-native canoe dynamical-core step.

Python host mirror of the reference's operator interface for this path,
over the C-ABI of libcanoe_gpu.so (include/canoe_gpu.h):

  Hydro.calculate_fluxes        <- Hydro::CalculateFluxes (patches/19.calculate_fluxes.patch)
  Hydro.riemann_solver          <- Hydro::RiemannSolver  (src/snap/riemann/lmars.cpp:19-22)
  Hydro.add_flux_divergence     <- Hydro::AddFluxDivergence
  ImplicitSolver.solve_implicit <- ImplicitSolver::SolveImplicit3D (src/snap/implicit/solve_implicit_3d.cpp:19)
  Thermo.forward                <- kintera ThermoY::forward (src/tasklist/implicit_hydro_tasks.cpp:308)
  Run.step                      <- ImplicitHydroTasks RK3 step (src/tasklist/implicit_hydro_tasks.cpp:134-313)

The HIP library is required: every entry point raises if it is missing or if
no GPU is present; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcanoe_gpu.so")

NGHOST = 3
MAX_NMASS = 3
NUM_KERNEL_IDS = 18
# bytes of the fixed device scratch of cg_implicit_solve_noalloc
NOALLOC_SCRATCH_BYTES = 64 << 20

OK, ERR_ARG, ERR_HIP, ERR_NAN, ERR_NEGATIVE, ERR_PIVOT, ERR_CONVERGE, \
    ERR_SIZE, ERR_NODEV, ERR_COLD, ERR_IO = range(11)
BC_REFLECT, BC_PERIODIC, BC_OUTFLOW, BC_PANEL = 0, 1, 2, 3
RS_LMARS, RS_HLLC = 0, 1
PREC_FP64, PREC_FP32 = 0, 1
SATADJ_OFF, SATADJ_NEWTON, SATADJ_MLP = 0, 1, 2


class Thermo(ctypes.Structure):
    _fields_ = [
        ("nmass", ctypes.c_int), ("nvapor", ctypes.c_int),
        ("gammad", ctypes.c_double), ("rd", ctypes.c_double),
        ("cv_ratio_m1", ctypes.c_double * MAX_NMASS),
        ("inv_mu_ratio_m1", ctypes.c_double * MAX_NMASS),
        ("ivap", ctypes.c_int), ("icld", ctypes.c_int),
        ("t3", ctypes.c_double), ("p3", ctypes.c_double),
        ("beta", ctypes.c_double), ("delta", ctypes.c_double),
        ("latent", ctypes.c_double), ("tol", ctypes.c_double),
        ("max_iter", ctypes.c_int),
    ]


class Block(ctypes.Structure):
    _fields_ = [
        ("nx1", ctypes.c_int), ("nx2", ctypes.c_int), ("nx3", ctypes.c_int),
        ("nmass", ctypes.c_int),
        ("x1min", ctypes.c_double), ("x2min", ctypes.c_double),
        ("x3min", ctypes.c_double),
        ("dx1", ctypes.c_double), ("dx2", ctypes.c_double),
        ("dx3", ctypes.c_double),
        ("grav", ctypes.c_double),
        ("bc", ctypes.c_int * 6),
    ]

    # AthenaArray geometry
    @property
    def nc1(self):
        return self.nx1 + 2 * NGHOST

    @property
    def nc2(self):
        return self.nx2 + 2 * NGHOST if self.nx2 > 1 else 1

    @property
    def nc3(self):
        return self.nx3 + 2 * NGHOST if self.nx3 > 1 else 1

    @property
    def N(self):
        return 5 + self.nmass

    def shape(self):
        """(N, nc3, nc2, nc1) of a cell-centred array."""
        return (self.N, self.nc3, self.nc2, self.nc1)

    def flux_shapes(self):
        N, c3, c2, c1 = self.shape()
        return ((N, c3, c2, c1 + 1), (N, c3, c2 + 1, c1), (N, c3 + 1, c2, c1))

    def interior(self):
        """Slices (k, j, i) of the interior cells."""
        g = NGHOST
        sj = slice(g, g + self.nx2) if self.nx2 > 1 else slice(0, 1)
        sk = slice(g, g + self.nx3) if self.nx3 > 1 else slice(0, 1)
        return (sk, sj, slice(g, g + self.nx1))

    def ncells(self):
        return self.nx1 * self.nx2 * self.nx3


class Config(ctypes.Structure):
    _fields_ = [
        ("order_h", ctypes.c_int), ("order_v", ctypes.c_int),
        ("rsolver", ctypes.c_int), ("implicit_flag", ctypes.c_int),
        ("precision", ctypes.c_int), ("satadj", ctypes.c_int),
        ("visc", ctypes.c_double), ("p0", ctypes.c_double),
        ("heat_flux", ctypes.c_double),
        ("k_auto", ctypes.c_double), ("k_accr", ctypes.c_double),
        ("k_evap", ctypes.c_double), ("v_fall", ctypes.c_double),
    ]


class Sphere(ctypes.Structure):
    """cg_sphere: rotation and Held-Suarez forcing of a cubed-sphere run."""
    _fields_ = [
        ("omega", ctypes.c_double), ("hs94", ctypes.c_int),
        ("ts", ctypes.c_double), ("dt_eq", ctypes.c_double),
        ("dtheta", ctypes.c_double), ("p0", ctypes.c_double),
        ("kf", ctypes.c_double), ("ka", ctypes.c_double),
        ("ks", ctypes.c_double), ("sigmab", ctypes.c_double),
    ]


class CanoeGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("canoe_gpu error %d: %s" % (code, msg))
        self.code = code


_lib = None

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_SIGS = {
    "cg_create": (_I, [_I, ctypes.POINTER(_P)]),
    "cg_destroy": (_I, [_P]),
    "cg_synchronize": (_I, [_P]),
    "cg_last_error": (ctypes.c_char_p, [_P]),
    "cg_malloc": (_I, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "cg_free": (_I, [_P, _P]),
    "cg_memcpy_h2d": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "cg_memcpy_d2h": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "cg_memcpy_d2d": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "cg_timer_start": (_I, [_P]),
    "cg_timer_stop": (_I, [_P, ctypes.POINTER(_D)]),
    "cg_riemann": (_I, [_P, _P, _I, _I, _I, _P, _P, ctypes.c_ssize_t, _P,
                        ctypes.c_ssize_t]),
    "cg_calculate_fluxes": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "cg_add_flux_divergence": (_I, [_P, _P, _D, _P, _P, _P, _P]),
    "cg_implicit_solve": (_I, [_P, _P, _P, _P, _P, _D, _I]),
    "cg_implicit_solve_noalloc": (_I, [_P, _P, _P, _P, _P, _D, _I]),
    "cg_riemann_host": (_I, [_P, _P, _I, _I, _I, _P, _P, ctypes.c_ssize_t, _P,
                             ctypes.c_ssize_t]),
    "cg_calculate_fluxes_host": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "cg_implicit_solve_host": (_I, [_P, _P, _P, _P, _P, _D, _I]),
    "cg_cons_to_prim": (_I, [_P, _P, _P, _P, _P]),
    "cg_microphysics": (_I, [_P, _P, _P, _P, _D, _P, _P]),
    "cg_sphere_create": (_I, [_P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "cg_pgen_sphere": (_I, [_P]),
    "cg_pgen_sphere_init": (_I, [ctypes.c_uint64, _P, _P, _P]),
    "cg_prim_to_cons": (_I, [_P, _P, _P, _P, _P]),
    "cg_apply_bc": (_I, [_P, _P, _P]),
    "cg_new_block_dt": (_I, [_P, _P, _P, _P, _P, _D, ctypes.POINTER(_D)]),
    "cg_saturation_adjust": (_I, [_P, _P, _I, _P, _P, _P, _P, _P]),
    "cg_run_create": (_I, [_P, _P, _P, _P, ctypes.POINTER(_P)]),
    "cg_run_destroy": (_I, [_P]),
    "cg_run_set_prim_host": (_I, [_P, _P]),
    "cg_run_get_cons_host": (_I, [_P, _P]),
    "cg_run_get_prim_host": (_I, [_P, _P]),
    "cg_run_step": (_I, [_P, _D, _I, _I]),
    "cg_run_stage": (_I, [_P, _I, _D, _P]),
    "cg_run_new_dt": (_I, [_P, _D, ctypes.POINTER(_D)]),
    "cg_run_check": (_I, [_P, _P]),
    "cg_run_advance": (_I, [_P, _D, _D, ctypes.c_longlong, _I,
                            ctypes.POINTER(_D), ctypes.POINTER(_D),
                            ctypes.POINTER(ctypes.c_longlong)]),
    "cg_run_device_state": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "cg_run_sums": (_I, [_P, _P]),
    "cg_run_history": (_I, [_P, _P]),
    "cg_run_debug": (_I, [_P, _P, _P, _P, _P]),
    "cg_run_set_timing": (_I, [_P, _I]),
    "cg_run_kernel_stats": (_I, [_P, _I, ctypes.POINTER(_D),
                                 ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(ctypes.c_char_p)]),
    "cg_batch_create": (_I, [_P, _I, _P, _P, _P, ctypes.POINTER(_P)]),
    "cg_run_nblocks": (_I, [_P, ctypes.POINTER(_I)]),
    "cg_run_block_set_prim_host": (_I, [_P, _I, _P]),
    "cg_run_block_get_cons_host": (_I, [_P, _I, _P]),
    "cg_run_block_get_prim_host": (_I, [_P, _I, _P]),
    "cg_run_block_sums": (_I, [_P, _I, _P]),
    "cg_run_get_tune": (_I, [_P, _P]),
    "cg_run_set_tune": (_I, [_P, _P]),
    "cg_run_autotune": (_I, [_P, _D, _P]),
    "cg_run_set_mlp": (_I, [_P, _P]),
    "cg_version": (ctypes.c_char_p, []),
    "cg_apply_bc_host": (_I, [_P, _P]),
    "cg_pgen_setup": (_I, [_I, _I, _I, _I, _P, _P, _P]),
    "cg_pgen_init": (_I, [_I, ctypes.c_uint64, _P, _P, _P, _P]),
    "cg_netcdf_write": (_I, [ctypes.c_char_p, _P, _D, _I, _P, _P, _P, _P, _I]),
}


def lib():
    """The loaded libcanoe_gpu.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CanoeGpuError(
                ERR_NODEV, "%s not built (run python -m canoe_amd.build)"
                % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _ref(s):
    return ctypes.cast(ctypes.pointer(s), ctypes.c_void_p)


# ---------------------------------------------------------------------------
def pgen_setup(config, nx1=0, nx2=0, nx3=0):
    """(Block, Thermo, Config) of configuration `config` (1, 2, 4 or 5)."""
    b, t, c = Block(), Thermo(), Config()
    r = lib().cg_pgen_setup(config, nx1, nx2, nx3, _ref(b), _ref(t), _ref(c))
    if r:
        raise CanoeGpuError(r, "cg_pgen_setup(%d)" % config)
    return b, t, c


def pgen_init(config, seed, blk, th, cfg):
    """Seeded primitive state (N, nc3, nc2, nc1), ghosts filled."""
    w = np.zeros(blk.shape(), dtype=np.float64)
    r = lib().cg_pgen_init(config, seed, _ref(blk), _ref(th), _ref(cfg),
                           _ptr(w))
    if r:
        raise CanoeGpuError(r, "cg_pgen_init(%d)" % config)
    return w


def pgen_sphere():
    """Rotation and Held-Suarez parameters of hs94.inp (cg_pgen_sphere)."""
    sp = Sphere()
    r = lib().cg_pgen_sphere(_ref(sp))
    if r:
        raise CanoeGpuError(r, "cg_pgen_sphere")
    return sp


def pgen_sphere_init(seed, blk, th):
    """Six panels' primitive states (6, N, nc3, nc2, nc1): isothermal 300 K
    at rest plus 0.1 K noise on the interior, ghosts zero."""
    w = np.zeros((6,) + blk.shape(), dtype=np.float64)
    r = lib().cg_pgen_sphere_init(seed, _ref(blk), _ref(th), _ptr(w))
    if r:
        raise CanoeGpuError(r, "cg_pgen_sphere_init")
    return w


def batch_blocks(seed, nblocks=32, heights=(40, 64, 96, 128, 200, 256),
                 footprints=(32, 64, 96, 128)):
    """Seeded blocks of the uneven-column batch: heights and footprint edges
    drawn from the given sets, each block with its own seeded profile.
    Returns [(Block, Thermo, Config, w0)]."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(nblocks):
        nx1 = int(rng.choice(heights))
        nx2 = int(rng.choice(footprints))
        nx3 = int(rng.choice(footprints))
        blk, th, cfg = pgen_setup(5, nx1, nx2, nx3)
        w = pgen_init(5, int(seed) * 1000 + b, blk, th, cfg)
        out.append((blk, th, cfg, w))
    return out


def netcdf_write(path, blk, time, fields, append=False):
    """NetCDF dump of a block in the layout of the reference's netcdf output
    (cg_netcdf_write).  fields: list of (name, units, long_name, array) with
    arrays (k, j, i) of the block including ghost cells."""
    n = len(fields)
    arrs = [np.ascontiguousarray(f[3], dtype=np.float64) for f in fields]
    for a in arrs:
        assert a.shape == blk.shape()[1:], (a.shape, blk.shape())
    enc = [[f[i].encode() for f in fields] for i in range(3)]
    cs = [(ctypes.c_char_p * max(n, 1))(*e) for e in enc]
    ps = (ctypes.c_void_p * max(n, 1))(*[a.ctypes.data for a in arrs])
    r = lib().cg_netcdf_write(path.encode(), _ref(blk), float(time), n,
                              ctypes.cast(cs[0], ctypes.c_void_p),
                              ctypes.cast(cs[1], ctypes.c_void_p),
                              ctypes.cast(cs[2], ctypes.c_void_p),
                              ctypes.cast(ps, ctypes.c_void_p), int(append))
    if r:
        raise CanoeGpuError(r, "cg_netcdf_write(%s)" % path)


def apply_bc_host(blk, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    r = lib().cg_apply_bc_host(_ref(blk), _ptr(w))
    if r:
        raise CanoeGpuError(r, "cg_apply_bc_host")
    return w


# ---------------------------------------------------------------------------
class DeviceArray:
    """A device buffer owned by a Context, with the shape of a numpy array."""

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        ctx._check(lib().cg_malloc(ctx.h, max(self.nbytes, 8),
                                   ctypes.byref(p)), "cg_malloc")
        self.ptr = p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        self.ctx._check(lib().cg_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a),
                                            self.nbytes), "h2d")
        return self

    def download(self):
        a = np.empty(self.shape, dtype=self.dtype)
        self.ctx._check(lib().cg_memcpy_d2h(self.ctx.h, _ptr(a), self.ptr,
                                            self.nbytes), "d2h")
        return a

    def free(self):
        if self.ptr is not None and self.ctx.h is not None:
            lib().cg_free(self.ctx.h, self.ptr)
        self.ptr = None


class Context:
    """One HIP device + stream (cg_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        r = lib().cg_create(device, ctypes.byref(h))
        if r:
            raise CanoeGpuError(r, "cg_create(device=%d)" % device)
        self.h = h

    def _check(self, r, what):
        if r:
            msg = lib().cg_last_error(self.h)
            raise CanoeGpuError(r, "%s: %s" % (what, msg.decode()
                                               if msg else ""))

    def synchronize(self):
        """Wait for the context's stream; reports kernel status (zero
        pivots, launch errors) of asynchronous calls."""
        self._check(lib().cg_synchronize(self.h), "cg_synchronize")

    def timer_start(self):
        """Record a HIP event on the context's stream (cg_timer_start)."""
        self._check(lib().cg_timer_start(self.h), "cg_timer_start")

    def timer_stop(self):
        """Device milliseconds since timer_start (cg_timer_stop)."""
        ms = ctypes.c_double()
        self._check(lib().cg_timer_stop(self.h, ctypes.byref(ms)),
                    "cg_timer_stop")
        return ms.value

    def array(self, a=None, shape=None, dtype=np.float64):
        if a is not None:
            a = np.ascontiguousarray(a, dtype=dtype)
            return DeviceArray(self, a.shape, dtype).upload(a)
        return DeviceArray(self, shape, dtype)

    def close(self):
        if self.h is not None:
            lib().cg_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---------------------------------------------------------------------------
class Hydro:
    """Flux sweep of one MeshBlock (host buffers in, host buffers out)."""

    def __init__(self, ctx, blk, th, cfg):
        self.ctx, self.blk, self.th, self.cfg = ctx, blk, th, cfg

    def riemann_solver(self, ivx, wl, wr):
        """Hydro::RiemannSolver on wl, wr (N, n_faces) -> flx (N, n_faces)."""
        wl = np.ascontiguousarray(wl, np.float64)
        wr = np.ascontiguousarray(wr, np.float64)
        N, nf = wl.shape
        dl, dr = self.ctx.array(wl), self.ctx.array(wr)
        df = self.ctx.array(shape=(N, nf))
        self.ctx._check(lib().cg_riemann(
            self.ctx.h, _ref(self.th), self.cfg.rsolver, ivx, nf, dl.ptr,
            dr.ptr, nf, df.ptr, nf), "cg_riemann")
        out = df.download()
        for a in (dl, dr, df):
            a.free()
        return out

    def calculate_fluxes(self, w):
        """Hydro::CalculateFluxes(w) -> (flux1, flux2, flux3)."""
        dw = self.ctx.array(w)
        fs = [self.ctx.array(np.zeros(s)) for s in self.blk.flux_shapes()]
        self.ctx._check(lib().cg_calculate_fluxes(
            self.ctx.h, _ref(self.blk), _ref(self.th), _ref(self.cfg), dw.ptr,
            fs[0].ptr, fs[1].ptr, fs[2].ptr), "cg_calculate_fluxes")
        out = tuple(f.download() for f in fs)
        for a in [dw] + fs:
            a.free()
        return out

    def add_flux_divergence(self, wght, fluxes, du):
        dfs = [self.ctx.array(f) for f in fluxes]
        ddu = self.ctx.array(du)
        self.ctx._check(lib().cg_add_flux_divergence(
            self.ctx.h, _ref(self.blk), wght, dfs[0].ptr, dfs[1].ptr,
            dfs[2].ptr, ddu.ptr), "cg_add_flux_divergence")
        out = ddu.download()
        for a in dfs + [ddu]:
            a.free()
        return out

    def new_block_dt(self, w, cfl):
        """Hydro::NewBlockTimeStep on the primitive array w (the patched
        form of patches/21.new_blockdt.patch): cfl times the minimum
        signal-speed step over interior cells."""
        dw = self.ctx.array(w)
        out = ctypes.c_double()
        self.ctx._check(lib().cg_new_block_dt(
            self.ctx.h, _ref(self.blk), _ref(self.th), _ref(self.cfg), dw.ptr,
            cfl, ctypes.byref(out)), "cg_new_block_dt")
        dw.free()
        return out.value


class ImplicitSolver:
    def __init__(self, ctx, blk, th, flag=9):
        self.ctx, self.blk, self.th, self.flag = ctx, blk, th, flag

    def solve_implicit_3d(self, du, w, dt):
        """ImplicitSolver::SolveImplicit3D(du, w, dt): returns the new du."""
        ddu, dw = self.ctx.array(du), self.ctx.array(w)
        try:
            self.ctx._check(lib().cg_implicit_solve(
                self.ctx.h, _ref(self.blk), _ref(self.th), ddu.ptr, dw.ptr, dt,
                self.flag), "cg_implicit_solve")
            return ddu.download()
        finally:
            ddu.free()
            dw.free()


class Tune(ctypes.Structure):
    """cg_tune: launch shape of a run."""
    _fields_ = [("implicit_wpb", ctypes.c_int), ("implicit_grid", ctypes.c_int),
                ("flux_block", ctypes.c_int), ("face_variant", ctypes.c_int),
                ("march_seg", ctypes.c_int), ("implicit_fused", ctypes.c_int),
                ("step_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Run:
    """Device-resident MeshBlocks advanced by full SSPRK(3,3) steps: one
    block (blk) or a batch of independent blocks (blocks = [Block, ...]);
    methods taking `b` address block b of a batch."""

    def __init__(self, ctx, blk, th, cfg, blocks=None):
        self.ctx, self.th, self.cfg = ctx, th, cfg
        self.blocks = list(blocks) if blocks is not None else [blk]
        self.blk = self.blocks[0]
        h = ctypes.c_void_p()
        arr = (Block * len(self.blocks))(*self.blocks)
        ctx._check(lib().cg_batch_create(
            ctx.h, len(self.blocks), ctypes.cast(arr, ctypes.c_void_p),
            _ref(th), _ref(cfg), ctypes.byref(h)), "cg_batch_create")
        self.h = h

    def set_prim(self, w, b=0):
        w = np.ascontiguousarray(w, np.float64)
        assert w.shape == self.blocks[b].shape()
        self.ctx._check(lib().cg_run_block_set_prim_host(self.h, b, _ptr(w)),
                        "set_prim")

    def cons(self, b=0):
        u = np.zeros(self.blocks[b].shape())
        self.ctx._check(lib().cg_run_block_get_cons_host(self.h, b, _ptr(u)),
                        "get_cons")
        return u

    def prim(self, b=0):
        w = np.zeros(self.blocks[b].shape())
        self.ctx._check(lib().cg_run_block_get_prim_host(self.h, b, _ptr(w)),
                        "get_prim")
        return w

    def step(self, dt, nsteps=1, graph=True):
        self.ctx._check(lib().cg_run_step(self.h, dt, nsteps, int(graph)),
                        "cg_run_step")

    def stage(self, s, dt=0.0, dt_dev=None):
        """One SSPRK(3,3) stage s (0, 1, 2), asynchronous; the step is dt or
        the double at device pointer dt_dev."""
        self.ctx._check(lib().cg_run_stage(self.h, s, dt, dt_dev),
                        "cg_run_stage")

    def new_dt(self, cfl):
        """Minimum over the run's blocks of NewBlockTimeStep, reduced on the
        device from the resident state."""
        out = ctypes.c_double()
        self.ctx._check(lib().cg_run_new_dt(self.h, cfl, ctypes.byref(out)),
                        "cg_run_new_dt")
        return out.value

    def check(self):
        """State checks of the resident w (check_hydro_variables): returns
        (status, (block, k, j, i)) with status 0 or ERR_NAN / ERR_NEGATIVE /
        ERR_COLD and the first failing interior cell."""
        where = np.zeros(4, np.int32)
        r = lib().cg_run_check(self.h, _ptr(where))
        if r not in (0, ERR_NAN, ERR_NEGATIVE, ERR_COLD):
            self.ctx._check(r, "cg_run_check")
        return r, tuple(int(x) for x in where)

    def advance(self, cfl, tlim=float("inf"), nlim=-1, time=0.0, dt=0.0,
                ncycle=0, graph=True):
        """The main loop of main.cpp:114-152 on the device (cycles until
        time >= tlim or ncycle >= nlim, each followed by Mesh::NewTimeStep).
        Returns (time, dt, ncycle) to continue from."""
        t, d, n = ctypes.c_double(time), ctypes.c_double(dt), \
            ctypes.c_longlong(ncycle)
        self.ctx._check(lib().cg_run_advance(
            self.h, cfl, tlim, nlim, int(graph), ctypes.byref(t),
            ctypes.byref(d), ctypes.byref(n)), "cg_run_advance")
        return t.value, d.value, n.value

    def sums(self, b=0):
        out = np.zeros(self.blocks[b].N)
        self.ctx._check(lib().cg_run_block_sums(self.h, b, _ptr(out)),
                        "cg_run_sums")
        return out

    def history(self):
        """Athena++ default history sums over all blocks (8 + nmass)."""
        out = np.zeros(8 + self.th.nmass)
        self.ctx._check(lib().cg_run_history(self.h, _ptr(out)), "history")
        return out

    def debug(self):
        """DebugOutput: (vmin, vmax, xmin (N, 3), xmax (N, 3)) of w."""
        N = self.blocks[0].N
        vmin, vmax = np.zeros(N), np.zeros(N)
        xmin, xmax = np.zeros((N, 3)), np.zeros((N, 3))
        self.ctx._check(lib().cg_run_debug(self.h, _ptr(vmin), _ptr(vmax),
                                           _ptr(xmin), _ptr(xmax)), "debug")
        return vmin, vmax, xmin, xmax

    def set_mlp(self, net):
        """Network of the guided saturation adjustment (config satadj =
        SATADJ_MLP); None restores the plain Newton adjustment."""
        self._mlp = net
        self.ctx._check(lib().cg_run_set_mlp(
            self.h, net.h if net is not None else None), "set_mlp")

    def tune(self):
        t = Tune()
        self.ctx._check(lib().cg_run_get_tune(self.h, _ref(t)), "get_tune")
        return t

    def set_tune(self, implicit_wpb=1, implicit_grid=0, flux_block=256,
                 face_variant=3, march_seg=16, implicit_fused=1):
        t = Tune(implicit_wpb, implicit_grid, flux_block, face_variant,
                 march_seg, implicit_fused, 0.0)
        self.ctx._check(lib().cg_run_set_tune(self.h, _ref(t)), "set_tune")

    def autotune(self, dt):
        t = Tune()
        self.ctx._check(lib().cg_run_autotune(self.h, dt, _ref(t)), "autotune")
        return t

    def set_timing(self, on):
        self.ctx._check(lib().cg_run_set_timing(self.h, int(on)), "timing")

    def kernel_stats(self):
        """{kernel name: (total ms, launches)} of the timed steps."""
        out = {}
        for kid in range(NUM_KERNEL_IDS):
            ms, n = ctypes.c_double(), ctypes.c_longlong()
            name = ctypes.c_char_p()
            self.ctx._check(lib().cg_run_kernel_stats(
                self.h, kid, ctypes.byref(ms), ctypes.byref(n),
                ctypes.byref(name)), "kernel_stats")
            if n.value:
                out[name.value.decode()] = (ms.value, n.value)
        return out

    def close(self):
        if self.h is not None:
            lib().cg_run_destroy(self.h)
            self.h = None


class SphereRun(Run):
    """The six panels of a cubed-sphere run (cg_sphere_create); panel b
    (0-5) is block b.  set_prim / prim take contravariant velocities, cons
    holds covariant momenta."""

    def __init__(self, ctx, panel, th, cfg, sp):
        self.ctx, self.th, self.cfg, self.sp = ctx, th, cfg, sp
        self.blocks = [panel] * 6
        self.blk = panel
        h = ctypes.c_void_p()
        ctx._check(lib().cg_sphere_create(ctx.h, _ref(panel), _ref(th),
                                          _ref(cfg), _ref(sp),
                                          ctypes.byref(h)), "cg_sphere_create")
        self.h = h

    def set_all(self, w6):
        for b in range(6):
            self.set_prim(w6[b], b)

    def cons_all(self):
        return np.stack([self.cons(b) for b in range(6)])

    def prim_all(self):
        return np.stack([self.prim(b) for b in range(6)])

