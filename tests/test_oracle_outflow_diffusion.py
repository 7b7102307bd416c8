"""Pins of two oracle rules that closed forms fix (CPU, no GPU).

1. ChangeToTemperature's outflow rule (src/snap/decomposition/
   change_to_temperature.cpp:113-127): ghosts beyond an outflow x1 side take
   the temperature T and the pressure perturbation p' = p - psv of the
   boundary cell.  With PLM the boundary face then sees a flat stencil on
   the ghost side, so its state is known in closed form: p_f = p'_b + psf_f,
   T_f = T_b, v_f = v_b, and the flux is the analytic Euler flux of that
   state (the Riemann solvers are consistent).  RestoreFromTemperature
   (:150-153) leaves those ghosts with p_g = p'_b + psv_g and
   T_g = T_b, the values the stage's sources and implicit solve read.
2. The Straka diffusion source (examples/2019-Li-snap/straka.cpp:178-188):
   du_v = dt K rho / (dx dy) * L5(v) for the x1 and x2 velocities and
   du_E = dt K rho / (dx dy) * cp T / theta * L5(theta), with L5 the
   five-point sum f(i-1) + f(i+1) + f(j-1) + f(j+1) - 4 f.  L5 vanishes on
   linear fields and equals 2 a dx^2 + 2 b dy^2 on a x^2 + b y^2.
"""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import perturbed_state


def _decompose(blk, th, w, k, j):
    nc1 = blk.nc1
    psf, psv, rdv = np.zeros(nc1 + 1), np.zeros(nc1), np.zeros(nc1)
    orl.lib().or_decompose_column(orl.ref(blk), orl.ref(th), orl.ptr(w), k, j,
                                  orl.ptr(psf), orl.ptr(psv), orl.ptr(rdv))
    return psf, psv, rdv


def _euler_flux(th, w):
    """Analytic x1 flux of a dry primitive state w (5,)."""
    rho, v1, v2, v3, p = w
    g = th.gammad
    h = g / (g - 1) * p / rho + 0.5 * (v1 * v1 + v2 * v2 + v3 * v3)
    return np.array([rho * v1, rho * v1 * v1 + p, rho * v1 * v2,
                     rho * v1 * v3, rho * v1 * h])


def _outflow_case(bcs, seed=3):
    blk, th, cfg, w = perturbed_state(2, (16, 6, 5), seed=seed, vel=3.0,
                                      dp=2e-3)
    for s, b in enumerate(bcs):
        blk.bc[s] = b
    cfg.order_v = 2  # PLM: flat ghost-side stencils give a closed form
    w = orl.apply_bc(blk, w)
    return blk, th, cfg, w


@pytest.mark.parametrize("solver", [ca.RS_LMARS, ca.RS_HLLC])
@pytest.mark.parametrize("side", ["top", "bottom"])
def test_outflow_face_closed_form(solver, side):
    """The boundary face of an outflow x1 side carries the state built from
    the boundary cell's T and p' over the face's hydrostatic pressure."""
    top = side == "top"
    bcs = (0 if top else 2, 2 if top else 0, 1, 1, 1, 1)
    blk, th, cfg, w = _outflow_case(bcs)
    cfg.rsolver = solver
    f1, _, _ = orl.calculate_fluxes(blk, th, cfg, w)
    is_, ie = 3, 3 + blk.nx1 - 1
    b, fc = (ie, ie + 1) if top else (is_, is_)
    worst, moved = 0.0, 0.0
    for k in range(3, 3 + blk.nx3):
        for j in range(3, 3 + blk.nx2):
            psf, psv, rdv = _decompose(blk, th, w, k, j)
            rho_b, v1, v2, v3, p_b = w[:, k, j, b]
            Tb = p_b / (rho_b * rdv[b])
            pp = p_b - psv[b]
            p_f = pp + psf[fc]
            gam = np.sqrt(rdv[fc] * rdv[fc - 1])
            wf = np.array([p_f / (Tb * gam), v1, v2, v3, p_f])
            ref = _euler_flux(th, wf)
            got = f1[:, k, j, fc]
            scale = np.array([abs(ref[0]) + 1e-3, p_f, p_f, p_f,
                              abs(ref[4]) + p_f * 10.0])
            worst = max(worst, float(np.max(np.abs(got - ref) / scale)))
            # the rule a dropped override would give instead: p' of the
            # ghost's own hydrostatic pressure (off by about rho g dz)
            g = ie + 1 if top else is_ - 1
            moved = max(moved, abs((p_b - psv[g]) - pp) / p_f)
    assert worst < 5e-12, worst  # round-off of the solver
    assert moved > 1e-3, moved  # the test sees the difference


def test_outflow_ghosts_after_calculate_fluxes():
    """Ghosts of an outflow x1 side after CalculateFluxes: T_g = T_b and
    p_g - psv_g = p_b - psv_b; interior cells and other sides untouched."""
    blk, th, cfg, w = _outflow_case((2, 2, 1, 1, 1, 1), seed=5)
    w2 = orl.restore_outflow_ghosts(blk, th, w)
    is_, ie = 3, 3 + blk.nx1 - 1
    for k in range(3, 3 + blk.nx3):
        for j in range(3, 3 + blk.nx2):
            psf, psv, rdv = _decompose(blk, th, w, k, j)
            for b, ghosts in ((is_, range(0, 3)), (ie, range(ie + 1, ie + 4))):
                Tb = w[4, k, j, b] / (w[0, k, j, b] * rdv[b])
                for g in ghosts:
                    Tg = w2[4, k, j, g] / (w2[0, k, j, g] * rdv[g])
                    assert abs(Tg / Tb - 1) < 4e-16 * 4
                    assert abs((w2[4, k, j, g] - psv[g]) -
                               (w[4, k, j, b] - psv[b])) <= 1e-11 * w[4, k, j, b]
                    # velocities are not touched
                    assert np.array_equal(w2[1:4, k, j, g], w[1:4, k, j, g])
    sk, sj, si = blk.interior()
    assert np.array_equal(w2[:, sk, sj, si], w[:, sk, sj, si])
    # a reflecting side is left alone
    blk2, th2, cfg2, wr = _outflow_case((0, 0, 1, 1, 1, 1), seed=5)
    assert np.array_equal(orl.restore_outflow_ghosts(blk2, th2, wr), wr)


# ---------------------------------------------------------------- diffusion
def _straka_block(dx1=100.0, dx2=160.0, nx1=12, nx2=10):
    blk, th, cfg = ca.pgen_setup(1, nx1, nx2, 1)
    blk.dx1, blk.dx2 = dx1, dx2
    cfg.heat_flux = 0.0
    return blk, th, cfg


def _diffusion(blk, th, cfg, w, dt):
    """The diffusion term alone: sources with K minus sources without."""
    K = cfg.visc
    du1 = orl.add_sources(blk, th, cfg, dt, w, np.zeros(blk.shape()))
    cfg.visc = 0.0
    du0 = orl.add_sources(blk, th, cfg, dt, w, np.zeros(blk.shape()))
    cfg.visc = K
    return du1 - du0


def _fields(blk, th, cfg, rng, vel, theta):
    """State with the given velocity and theta fields of the cell-centre
    coordinates (x: x1, y: x2), random density and pressure; ghosts hold
    the same fields so every stencil sees them."""
    w = np.zeros(blk.shape())
    x = (np.arange(blk.nc1) - 2.5) * blk.dx1
    y = (np.arange(blk.nc2) - 2.5) * blk.dx2
    X, Y = np.meshgrid(x, y)  # (nc2, nc1)
    w[0, 0] = rng.uniform(0.6, 1.3, X.shape)
    p = rng.uniform(5e4, 1.05e5, X.shape)
    w[4, 0] = p
    w[1, 0], w[2, 0] = vel(X, Y)
    chi = (th.gammad - 1) / th.gammad
    T = theta(X, Y) * (p / cfg.p0) ** chi
    w[0, 0] = p / (th.rd * T)  # dry: T = p / (rho Rd)
    return w, X, Y


def test_diffusion_vanishes_on_linear_fields():
    blk, th, cfg = _straka_block()
    rng = np.random.default_rng(11)
    w, X, Y = _fields(blk, th, cfg, rng,
                      lambda X, Y: (3.0 + 2e-3 * X - 1e-3 * Y,
                                    -1.0 + 5e-4 * X + 3e-3 * Y),
                      lambda X, Y: 300.0 + 1e-3 * X - 2e-3 * Y)
    dt = 0.7
    d = _diffusion(blk, th, cfg, w, dt)
    sk, sj, si = blk.interior()
    K, dx, dy = cfg.visc, blk.dx1, blk.dx2
    rho = w[0][sk, sj, si]
    cp = th.gammad / (th.gammad - 1) * th.rd
    s_v = dt * K * rho / (dx * dy) * 1.0            # per m/s of L5
    s_e = dt * K * rho / (dx * dy) * cp * 300.0     # per theta scale
    assert np.max(np.abs(d[1][sk, sj, si]) / s_v) < 1e-12
    assert np.max(np.abs(d[2][sk, sj, si]) / s_v) < 1e-12
    assert np.max(np.abs(d[4][sk, sj, si]) / s_e) < 1e-13
    assert np.all(d[0] == 0.0) and np.all(d[3] == 0.0)


def test_diffusion_quadratic_fields_closed_form():
    """v1 = a1 x^2 + b1 y^2, v2 = a2 x^2 + b2 y^2, theta = 290 + a x^2 +
    b y^2 give du_v = dt K rho (2 a dx^2 + 2 b dy^2) / (dx dy) and
    du_E = dt K rho cp (T / theta) (2 a dx^2 + 2 b dy^2) / (dx dy) with
    T / theta = (p / p0)^chi: a dropped rho, cp in place of cp T / theta or a
    dx^2 denominator each fail (dx != dy, p != p0, rho != 1)."""
    blk, th, cfg = _straka_block()
    rng = np.random.default_rng(12)
    a1, b1, a2, b2, a, b = 2e-7, -3e-7, -1e-7, 4e-7, 3e-8, -5e-8
    w, X, Y = _fields(blk, th, cfg, rng,
                      lambda X, Y: (a1 * X * X + b1 * Y * Y,
                                    a2 * X * X + b2 * Y * Y),
                      lambda X, Y: 290.0 + a * X * X + b * Y * Y)
    dt = 0.4
    d = _diffusion(blk, th, cfg, w, dt)
    sk, sj, si = blk.interior()
    K, dx, dy = cfg.visc, blk.dx1, blk.dx2
    rho = w[0][sk, sj, si]
    p = w[4][sk, sj, si]
    chi = (th.gammad - 1) / th.gammad
    cp = th.gammad / (th.gammad - 1) * th.rd
    c = dt * K * rho / (dx * dy)
    ref1 = c * (2 * a1 * dx * dx + 2 * b1 * dy * dy)
    ref2 = c * (2 * a2 * dx * dx + 2 * b2 * dy * dy)
    refe = c * cp * (p / cfg.p0) ** chi * (2 * a * dx * dx + 2 * b * dy * dy)
    assert np.max(np.abs(d[1][sk, sj, si] / ref1 - 1)) < 1e-9
    assert np.max(np.abs(d[2][sk, sj, si] / ref2 - 1)) < 1e-9
    assert np.max(np.abs(d[4][sk, sj, si] / refe - 1)) < 1e-9
    # each factor matters here
    assert np.ptp(rho) > 0.1 and np.ptp((p / cfg.p0) ** chi) > 0.05
    assert abs(dx - dy) > 10.0
