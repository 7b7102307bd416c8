"""ctypes loader of the CPU oracle (oracle/lib/liboracle.so).

Test infrastructure: results are compared with the oracle; it is never the
thing measured.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "lib", "liboracle.so")

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_SIGS = {
    "or_exp": (_D, [_D]),
    "or_log": (_D, [_D]),
    "or_get_gamma": (_D, [_P, _P]),
    "or_recon_cell": (None, [_I, _P, _P, _P]),
    "or_riemann": (_I, [_P, _I, _I, _I, _P, _P, ctypes.c_ssize_t, _P,
                        ctypes.c_ssize_t]),
    "or_decompose_column": (_I, [_P, _P, _P, _I, _I, _P, _P, _P]),
    "or_calculate_fluxes": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "or_face_failures": (ctypes.c_longlong, []),
    "or_restore_outflow_ghosts": (_I, [_P, _P, _P]),
    "or_add_flux_divergence": (_I, [_P, _D, _P, _P, _P, _P]),
    "or_add_sources": (_I, [_P, _P, _P, _D, _P, _P]),
    "or_implicit_solve": (_I, [_P, _P, _P, _P, _D, _I]),
    "or_cons_to_prim": (_I, [_P, _P, _P, _P]),
    "or_prim_to_cons": (_I, [_P, _P, _P, _P]),
    "or_apply_bc": (_I, [_P, _P]),
    "or_saturation_adjust": (_I, [_P, _I, _P, _P, _P, _P, _P]),
    "or_step": (_I, [_P, _P, _P, _D, _I, _P, _P]),
    "or_kessler": (_I, [_P, _P, _P, _D, _P]),
    "or_sph_step": (_I, [_P, _P, _P, _P, _D, _I, _P, _P]),
    "or_sph_fill_ghosts": (_I, [_P, _P, _P, _P]),
    "or_sph_prim_to_cons": (_I, [_P, _P, _P, _P]),
    "or_sph_forcing": (_I, [_P, _P, _P, _D, _P, _P]),
    "or_sph_metric": (_I, [_P, _I, _I, _P]),
    "or_sph_link": (_I, [_I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                         ctypes.POINTER(_I)]),
    "or_sph_latlon": (None, [_I, _D, _D, ctypes.POINTER(_D),
                             ctypes.POINTER(_D)]),
    "or_sph_rll_from_abp": (None, [_D, _D, _I, _D, _D, ctypes.POINTER(_D),
                                   ctypes.POINTER(_D)]),
    "or_sph_abp_from_rll": (None, [_D, _D, _I, _D, _D, ctypes.POINTER(_D),
                                   ctypes.POINTER(_D)]),
    "or_step_omp": (_I, [_P, _P, _P, _D, _I, _P, _P, _I]),
    "or_sums": (_I, [_P, _P, _P]),
    "or_new_block_dt": (_D, [_P, _P, _P, _P, _D]),
    "or_history": (_I, [_P, _P, _P]),
    "or_check": (_I, [_P, _P, _P, _P]),
    "or_debug": (_I, [_P, _P, _P, _P, _P, _P]),
    "or_new_mesh_dt": (_D, [_D, _D, _D, _D]),
    "or_block_thomas": (_I, [_I, _I, _P, _P, _P, _P, _P]),
    "or_block_thomas_periodic": (_I, [_I, _I, _P, _P, _P, _P, _P, _P]),
    "or_flux_jacobian": (_I, [_P, _D, _P, _P]),
    "or_roe_matrices": (_I, [_P, _P, _D, _P, _P, _P, _P]),
    "or_mlp_nparam": (_I, []),
    "or_mlp_gelu": (ctypes.c_float, [ctypes.c_float]),
    "or_mlp_gelu_grad": (ctypes.c_float, [ctypes.c_float]),
    "or_mlp_exp": (ctypes.c_float, [ctypes.c_float]),
    "or_mlp_bf16": (ctypes.c_ushort, [ctypes.c_float]),
    "or_mlp_forward": (_I, [_P, _I, _P, _P, _I]),
    "or_mlp_loss_grad": (_I, [_P, _I, _P, _P, _P, _P, _I]),
    "or_mlp_pairs": (_I, [_P, _I, _P, _P, _P, _I, ctypes.c_ulonglong, _P,
                          _P, _P]),
    "or_mlp_features": (_I, [_P, _I, _P, _P, _P, _P, _P]),
}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        for name, (res, args) in _SIGS.items():
            try:
                f = getattr(L, name)
            except AttributeError:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def ref(s):
    return ctypes.cast(ctypes.pointer(s), ctypes.c_void_p)


def riemann(th, solver, ivx, wl, wr):
    wl = np.ascontiguousarray(wl, np.float64)
    wr = np.ascontiguousarray(wr, np.float64)
    N, nf = wl.shape
    out = np.zeros_like(wl)
    lib().or_riemann(ref(th), solver, ivx, nf, ptr(wl), ptr(wr), nf, ptr(out),
                     nf)
    return out


def calculate_fluxes(blk, th, cfg, w):
    w = np.ascontiguousarray(w, np.float64)
    fs = [np.zeros(s) for s in blk.flux_shapes()]
    lib().or_calculate_fluxes(ref(blk), ref(th), ref(cfg), ptr(w), ptr(fs[0]),
                              ptr(fs[1]), ptr(fs[2]))
    return fs


def face_failures():
    """Faces of the last calculate_fluxes with a negative reconstructed
    density or pressure (check_reconstruct)."""
    return int(lib().or_face_failures())


def restore_outflow_ghosts(blk, th, w):
    """The x1 ghosts of w that the reference's CalculateFluxes leaves on an
    outflow side (returns a copy of w)."""
    w = np.array(w, dtype=np.float64, copy=True)
    lib().or_restore_outflow_ghosts(ref(blk), ref(th), ptr(w))
    return w


def add_flux_divergence(blk, wght, fs, du):
    du = np.array(du, dtype=np.float64, copy=True)
    fs = [np.ascontiguousarray(f, np.float64) for f in fs]
    lib().or_add_flux_divergence(ref(blk), wght, ptr(fs[0]), ptr(fs[1]),
                                 ptr(fs[2]), ptr(du))
    return du


def add_sources(blk, th, cfg, dt, w, du):
    du = np.array(du, dtype=np.float64, copy=True)
    w = np.ascontiguousarray(w, np.float64)
    lib().or_add_sources(ref(blk), ref(th), ref(cfg), dt, ptr(w), ptr(du))
    return du


def implicit_solve(blk, th, du, w, dt, flag):
    du = np.array(du, dtype=np.float64, copy=True)
    w = np.ascontiguousarray(w, np.float64)
    r = lib().or_implicit_solve(ref(blk), ref(th), ptr(du), ptr(w), dt, flag)
    return du, r


def cons_to_prim(blk, th, u, w=None):
    u = np.ascontiguousarray(u, np.float64)
    w = np.zeros_like(u) if w is None else np.array(w, copy=True)
    lib().or_cons_to_prim(ref(blk), ref(th), ptr(u), ptr(w))
    return w


def prim_to_cons(blk, th, w):
    w = np.ascontiguousarray(w, np.float64)
    u = np.zeros_like(w)
    lib().or_prim_to_cons(ref(blk), ref(th), ptr(w), ptr(u))
    return u


def apply_bc(blk, w):
    w = np.array(w, dtype=np.float64, copy=True)
    lib().or_apply_bc(ref(blk), ptr(w))
    return w


def step(blk, th, cfg, dt, nsteps, u, w, threads=1):
    u = np.array(u, dtype=np.float64, copy=True)
    w = np.array(w, dtype=np.float64, copy=True)
    if threads == 1:
        r = lib().or_step(ref(blk), ref(th), ref(cfg), dt, nsteps, ptr(u),
                          ptr(w))
    else:
        r = lib().or_step_omp(ref(blk), ref(th), ref(cfg), dt, nsteps, ptr(u),
                              ptr(w), threads)
    assert r == 0, r
    return u, w


def kessler(blk, th, cfg, dt, u):
    """Kessler microphysics + sedimentation of the oracle on conserved u."""
    u = np.array(u, dtype=np.float64, copy=True)
    r = lib().or_kessler(ref(blk), ref(th), ref(cfg), dt, ptr(u))
    assert r == 0, r
    return u


def new_block_dt(blk, th, cfg, w, cfl):
    """Hydro::NewBlockTimeStep of the oracle (cfl times the minimum step)."""
    w = np.ascontiguousarray(w, np.float64)
    return lib().or_new_block_dt(ref(blk), ref(th), ref(cfg), ptr(w), cfl)


def check(blk, th, w):
    """(code, (k, j, i)): 0 NaN, 1 negative, 2 too cold, 3 pass."""
    w = np.ascontiguousarray(w, np.float64)
    where = np.zeros(3, np.int32)
    code = lib().or_check(ref(blk), ref(th), ptr(w), ptr(where))
    return code, tuple(int(x) for x in where)


def history(blk, u):
    u = np.ascontiguousarray(u, np.float64)
    out = np.zeros(8 + blk.nmass)
    lib().or_history(ref(blk), ptr(u), ptr(out))
    return out


def debug(blk, w):
    """(vmin, vmax, xmin (N, 3), xmax (N, 3)) of DebugOutput."""
    w = np.ascontiguousarray(w, np.float64)
    N = blk.N
    vmin, vmax = np.zeros(N), np.zeros(N)
    xmin, xmax = np.zeros((N, 3)), np.zeros((N, 3))
    lib().or_debug(ref(blk), ptr(w), ptr(vmin), ptr(vmax), ptr(xmin), ptr(xmax))
    return vmin, vmax, xmin, xmax


def new_mesh_dt(dt_old, blocks_dt, time, tlim):
    return lib().or_new_mesh_dt(dt_old, blocks_dt, time, tlim)


def saturation_adjust(th, rho, ie, y, tguess=None):
    rho = np.ascontiguousarray(rho, np.float64)
    ie = np.array(ie, dtype=np.float64, copy=True)
    y = np.array(y, dtype=np.float64, copy=True)
    n = rho.size
    it = np.zeros(n, dtype=np.int32)
    tg = None if tguess is None else ptr(np.ascontiguousarray(tguess,
                                                              np.float64))
    r = lib().or_saturation_adjust(ref(th), n, ptr(rho), ptr(ie), ptr(y), tg,
                                   ptr(it))
    return ie, y, it, r


# ------------------------------------------------------- saturation guess
def mlp_forward(params, x, emulate=1):
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros((x.shape[0], 3), np.float64)
    r = lib().or_mlp_forward(ptr(p), x.shape[0], ptr(x), ptr(y), emulate)
    assert r == 0
    return y


def mlp_loss_grad(params, x, t, emulate=1):
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    g = np.zeros(p.size, np.float64)
    loss = ctypes.c_double()
    r = lib().or_mlp_loss_grad(ptr(p), x.shape[0], ptr(x), ptr(t), ptr(g),
                               ctypes.cast(ctypes.byref(loss), ctypes.c_void_p),
                               emulate)
    assert r == 0
    return g, loss.value


def mlp_features(th, rho, ie, y):
    rho = np.ascontiguousarray(rho, np.float64)
    ie = np.ascontiguousarray(ie, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = rho.shape[0]
    x = np.zeros((n, 4), np.float32)
    tu = np.zeros(n, np.float64)
    r = lib().or_mlp_features(ref(th), n, ptr(rho), ptr(ie), ptr(y), ptr(x),
                              ptr(tu))
    assert r == 0
    return x, tu


def mlp_pairs(th, rho, ie, y, npairs, seed):
    rho = np.ascontiguousarray(rho, np.float64)
    ie = np.ascontiguousarray(ie, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    x = np.zeros((npairs, 4), np.float32)
    t = np.zeros((npairs, 3), np.float32)
    it = np.zeros(npairs, np.int32)
    r = lib().or_mlp_pairs(ref(th), rho.shape[0], ptr(rho), ptr(ie), ptr(y),
                           npairs, seed, ptr(x), ptr(t), ptr(it))
    assert r == 0
    return x, t, it


# ------------------------------------------------------------ cubed sphere
def sph_prim_to_cons(blk, th, w6):
    w6 = np.ascontiguousarray(w6, np.float64)
    u6 = np.zeros_like(w6)
    assert lib().or_sph_prim_to_cons(ref(blk), ref(th), ptr(w6), ptr(u6)) == 0
    return u6


def sph_fill_ghosts(blk, th, u6, w6):
    """w6 with its x2 / x3 ghosts filled from the neighbours' u6 and the x1
    walls."""
    u6 = np.ascontiguousarray(u6, np.float64)
    w6 = np.array(w6, dtype=np.float64, copy=True)
    assert lib().or_sph_fill_ghosts(ref(blk), ref(th), ptr(u6), ptr(w6)) == 0
    return w6


def sph_step(blk, th, cfg, sp, dt, nsteps, u6, w6):
    u6 = np.array(u6, dtype=np.float64, copy=True)
    w6 = np.array(w6, dtype=np.float64, copy=True)
    r = lib().or_sph_step(ref(blk), ref(th), ref(cfg), ref(sp), dt, nsteps,
                          ptr(u6), ptr(w6))
    assert r == 0, r
    return u6, w6


def sph_metric(blk, p, what):
    """Metric array `what` of panel p (1-6): (nc, nc) tables for 0-11, the
    interior (nx3, nx2, nx1) volumes (12) or lower x1-face areas (13)."""
    if what < 12:
        out = np.zeros((blk.nx2 + 6, blk.nx2 + 6))
    else:
        out = np.zeros((blk.nx3, blk.nx2, blk.nx1))
    assert lib().or_sph_metric(ref(blk), p, what, ptr(out)) == 0
    return out


def sph_link(p, s):
    q, sq, inv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().or_sph_link(p, s, ctypes.byref(q), ctypes.byref(sq), ctypes.byref(inv))
    return q.value, sq.value, inv.value


def sph_latlon(p, X, Y):
    lat, lon = ctypes.c_double(), ctypes.c_double()
    lib().or_sph_latlon(p, X, Y, ctypes.byref(lat), ctypes.byref(lon))
    return lat.value, lon.value


def sph_rll_from_abp(X, Y, block, v2, v3):
    U, V = ctypes.c_double(), ctypes.c_double()
    lib().or_sph_rll_from_abp(X, Y, block, v2, v3, ctypes.byref(U), ctypes.byref(V))
    return U.value, V.value


def sph_abp_from_rll(X, Y, block, U, V):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().or_sph_abp_from_rll(X, Y, block, U, V, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def sph_forcing(blk, th, sp, wght, w6):
    """Coriolis + Held-Suarez terms of du over the step wght."""
    w6 = np.ascontiguousarray(w6, np.float64)
    du = np.zeros_like(w6)
    assert lib().or_sph_forcing(ref(blk), ref(th), ref(sp), wght, ptr(w6), ptr(du)) == 0
    return du

