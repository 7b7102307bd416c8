"""ctypes binding of the saturation-guess network and of the saturation
adjustment (include/canoe_gpu.h: cg_mlp_*, cg_saturation_adjust).

    net = Mlp(ctx, seed=1)
    x, t, it = make_pairs(ctx, th, rho, ie, y, npairs=1 << 20, seed=3)
    losses = net.train(x, t, steps=2000, batch=8192)
    ie2, y2, iters = saturation_adjust(ctx, th, rho, ie, y, net)

Array arguments are numpy arrays (copied to and from the device) or
DeviceArray objects (used in place).
"""
import ctypes

import numpy as np

from . import (CanoeGpuError, DeviceArray, _ptr, _ref, lib)

NPARAM = 8835
_P = ctypes.c_void_p
_I = ctypes.c_int
_LL = ctypes.c_longlong
_SIGS = {
    "cg_mlp_create": (_I, [_P, ctypes.c_ulonglong, ctypes.POINTER(_P)]),
    "cg_mlp_destroy": (_I, [_P]),
    "cg_mlp_set_params": (_I, [_P, _P]),
    "cg_mlp_get_params": (_I, [_P, _P]),
    "cg_mlp_forward": (_I, [_P, _LL, _P, _P]),
    "cg_mlp_loss_grad": (_I, [_P, _LL, _P, _P, _P,
                              ctypes.POINTER(ctypes.c_double)]),
    "cg_mlp_train": (_I, [_P, _P, _LL, _P, _P, _P]),
    "cg_mlp_make_pairs": (_I, [_P, _P, _I, _P, _P, _P, _LL,
                               ctypes.c_ulonglong, _P, _P, _P]),
    "cg_saturation_adjust": (_I, [_P, _P, _I, _P, _P, _P, _P, _P]),
}
_done = False


def _lib():
    global _done
    L = lib()
    if not _done:
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _done = True
    return L


class TrainCfg(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int), ("steps", ctypes.c_int),
                ("lr", ctypes.c_float), ("lr_final", ctypes.c_float),
                ("seed", ctypes.c_ulonglong)]


def _dev(ctx, a, dtype):
    """(DeviceArray, owned) for a numpy array or an existing DeviceArray."""
    if isinstance(a, DeviceArray):
        return a, False
    return ctx.array(np.ascontiguousarray(a, dtype=dtype), dtype=dtype), True


class Mlp:
    """cg_mlp: parameters, Adam state and training scratch on one context."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._check(_lib().cg_mlp_create(ctx.h, seed, ctypes.byref(h)),
                   "cg_mlp_create")
        self.h = h

    def close(self):
        if self.h is not None:
            _lib().cg_mlp_destroy(self.h)
            self.h = None

    def params(self):
        p = np.empty(NPARAM, np.float32)
        self.ctx._check(_lib().cg_mlp_get_params(self.h, _ptr(p)),
                        "cg_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == NPARAM
        self.ctx._check(_lib().cg_mlp_set_params(self.h, _ptr(p)),
                        "cg_mlp_set_params")

    def forward(self, x):
        dx, own = _dev(self.ctx, x, np.float32)
        n = dx.shape[0]
        dy = self.ctx.array(shape=(n, 3), dtype=np.float32)
        try:
            self.ctx._check(_lib().cg_mlp_forward(self.h, n, dx.ptr, dy.ptr),
                            "cg_mlp_forward")
            return dy.download()
        finally:
            dy.free()
            if own:
                dx.free()

    def loss_grad(self, x, t):
        dx, ox = _dev(self.ctx, x, np.float32)
        dt, ot = _dev(self.ctx, t, np.float32)
        dg = self.ctx.array(shape=(NPARAM,), dtype=np.float32)
        loss = ctypes.c_double()
        try:
            self.ctx._check(_lib().cg_mlp_loss_grad(
                self.h, dx.shape[0], dx.ptr, dt.ptr, dg.ptr,
                ctypes.byref(loss)), "cg_mlp_loss_grad")
            return dg.download(), loss.value
        finally:
            dg.free()
            if ox:
                dx.free()
            if ot:
                dt.free()

    def train(self, x, t, steps, batch=8192, lr=2e-3, lr_final=2e-5,
              seed=11):
        dx, ox = _dev(self.ctx, x, np.float32)
        dt, ot = _dev(self.ctx, t, np.float32)
        cfg = TrainCfg(batch, steps, lr, lr_final, seed)
        losses = np.zeros(max(steps, 1), np.float32)
        try:
            self.ctx._check(_lib().cg_mlp_train(
                self.h, _ref(cfg), dx.shape[0], dx.ptr, dt.ptr,
                _ptr(losses)), "cg_mlp_train")
            return losses[:steps]
        finally:
            if ox:
                dx.free()
            if ot:
                dt.free()


def make_pairs(ctx, th, rho, ie, y, npairs, seed, device=False):
    """Training pairs (x (npairs, 4), t (npairs, 3), plain iterations) from
    base states rho (n,), ie (n,), y (nmass, n).  device=True returns the
    DeviceArrays of x and t instead of host copies (iterations stay host)."""
    drho, o1 = _dev(ctx, rho, np.float64)
    die, o2 = _dev(ctx, ie, np.float64)
    dy, o3 = _dev(ctx, y, np.float64)
    n = drho.shape[0]
    dx = ctx.array(shape=(npairs, 4), dtype=np.float32)
    dtt = ctx.array(shape=(npairs, 3), dtype=np.float32)
    dit = ctx.array(shape=(npairs,), dtype=np.int32)
    try:
        ctx._check(_lib().cg_mlp_make_pairs(
            ctx.h, _ref(th), n, drho.ptr, die.ptr, dy.ptr, npairs, seed,
            dx.ptr, dtt.ptr, dit.ptr), "cg_mlp_make_pairs")
        it = dit.download()
        if device:
            return dx, dtt, it
        x, t = dx.download(), dtt.download()
        dx.free()
        dtt.free()
        return x, t, it
    finally:
        dit.free()
        for d, o in ((drho, o1), (die, o2), (dy, o3)):
            if o:
                d.free()


def saturation_adjust(ctx, th, rho, ie, y, net=None):
    """cg_saturation_adjust on host arrays: returns (ie, y, iterations)."""
    rho = np.ascontiguousarray(rho, np.float64)
    n = rho.shape[0]
    drho = ctx.array(rho)
    die = ctx.array(np.ascontiguousarray(ie, np.float64))
    dy = ctx.array(np.ascontiguousarray(y, np.float64))
    dit = ctx.array(shape=(n,), dtype=np.int32)
    try:
        ctx._check(_lib().cg_saturation_adjust(
            ctx.h, _ref(th), n, drho.ptr, die.ptr, dy.ptr,
            net.h if net is not None else None, dit.ptr),
            "cg_saturation_adjust")
        return die.download(), dy.download(), dit.download()
    finally:
        for d in (drho, die, dy, dit):
            d.free()
