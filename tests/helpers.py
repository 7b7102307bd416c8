"""Shared helpers of the parity tests."""
import numpy as np

import canoe_amd as ca


def relerr(a, b):
    """Max over entries of |a - b| / max(|b|, rms_n(b)), the floor being the
    root-mean-square of variable n (first axis) over the whole array: a
    per-cell relative error that stays meaningful where a component crosses
    zero."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    if b.ndim == 1:
        b2, a2 = b[None], a[None]
    else:
        b2 = b.reshape(b.shape[0], -1)
        a2 = a.reshape(a.shape[0], -1)
    rms = np.sqrt(np.mean(b2 * b2, axis=1, keepdims=True))
    rms = np.where(rms > 0, rms, 1.0)
    den = np.maximum(np.abs(b2), rms)
    return float(np.max(np.abs(a2 - b2) / den))


def perturbed_state(config, dims, seed=7, vel=2.0, dp=5e-3, cloud=True):
    """Configuration `config` at size dims with seeded perturbations of
    velocity, density and pressure (and cloud/precip for moist runs) so that
    every branch of the fluxes and the solvers is exercised.  Ghosts filled."""
    blk, th, cfg = ca.pgen_setup(config, *dims)
    w = ca.pgen_init(config, seed, blk, th, cfg)
    rng = np.random.default_rng(seed)
    sk, sj, si = blk.interior()
    ivx = 1 + blk.nmass
    ipr = ivx + 3
    shp = w[0][sk, sj, si].shape
    for d in range(3):
        if d == 0 or (d == 1 and blk.nx2 > 1) or (d == 2 and blk.nx3 > 1):
            w[ivx + d][sk, sj, si] = vel * rng.uniform(-1, 1, shp)
    w[0][sk, sj, si] *= 1 + dp * rng.uniform(-1, 1, shp)
    w[ipr][sk, sj, si] *= 1 + dp * rng.uniform(-1, 1, shp)
    if blk.nmass == 3 and cloud:
        w[2][sk, sj, si] = 2e-3 * rng.uniform(0, 1, shp) * \
            (rng.uniform(0, 1, shp) < 0.3)
        w[3][sk, sj, si] = 1e-3 * rng.uniform(0, 1, shp)
    w = ca.apply_bc_host(blk, w)
    return blk, th, cfg, w


def cfl_dt(blk, th, cfg, w, cfl=0.9):
    """Time step from the signal speeds of the patched
    Hydro::NewBlockTimeStep (patches/21.new_blockdt.patch:31-45: the
    implicit x1 direction drops the sound speed), summed over directions as
    the unsplit update requires.  Sound speed with the dry gamma."""
    sk, sj, si = blk.interior()
    ivx = 1 + blk.nmass
    rho = w[0][sk, sj, si]
    p = w[ivx + 3][sk, sj, si]
    cs = np.sqrt(th.gammad * p / rho)
    inv = 0.0
    dxs = (blk.dx1, blk.dx2, blk.dx3)
    for d in range(3):
        if (d == 1 and blk.nx2 == 1) or (d == 2 and blk.nx3 == 1):
            continue
        v = np.abs(w[ivx + d][sk, sj, si])
        s = v if (d == 0 and cfg.implicit_flag) else v + cs
        inv = inv + s / dxs[d]
    return float(cfl / np.max(inv))
