"""GPU parity of the FP64 plain path against the CPU oracle.

Every comparison is on the same seeded inputs; the bar is <= 1e-12 per
entry (helpers.relerr).  The FP64 device code evaluates the same
expressions in the same order without contraction, so the errors observed
are expected to be exactly 0.
"""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import cfl_dt, perturbed_state, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12
CASES = [(1, (24, 40, 1)), (2, (20, 12, 10)), (4, (20, 12, 10))]


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("nmass", [0, 3])
@pytest.mark.parametrize("solver", [ca.RS_LMARS, ca.RS_HLLC])
def test_riemann(ctx, nmass, solver):
    cfgid = 2 if nmass == 0 else 4
    blk, th, cfg = ca.pgen_setup(cfgid, 8, 8, 8)
    cfg.rsolver = solver
    rng = np.random.default_rng(11 + nmass + 7 * solver)
    N = 5 + nmass
    nf = 4099
    ivx = 1 + nmass

    def states():
        w = np.zeros((N, nf))
        w[0] = rng.uniform(0.2, 1.5, nf)
        for n in range(1, 1 + nmass):
            w[n] = rng.uniform(0, 0.01, nf)
        for d in range(3):
            w[ivx + d] = rng.uniform(-60, 60, nf)
        w[ivx + 3] = rng.uniform(2e4, 1.1e5, nf)
        return w
    wl, wr = states(), states()
    h = ca.Hydro(ctx, blk, th, cfg)
    for d in range(3):
        _say("riemann nmass=%d solver=%d dir=%d" % (nmass, solver, d))
        got = h.riemann_solver(ivx + d, wl, wr)
        ref = orl.riemann(th, solver, ivx + d, wl, wr)
        assert relerr(got, ref) <= TOL


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("combo", [(2, 2, ca.RS_LMARS), (5, 2, ca.RS_HLLC),
                                   (5, 5, ca.RS_LMARS), (2, 5, ca.RS_HLLC)])
def test_calculate_fluxes(ctx, case, combo):
    config, dims = case
    blk, th, cfg, w = perturbed_state(config, dims)
    cfg.order_h, cfg.order_v, cfg.rsolver = combo
    _say("fluxes config=%d dims=%s combo=%s" % (config, dims, combo))
    got = ca.Hydro(ctx, blk, th, cfg).calculate_fluxes(w)
    ref = orl.calculate_fluxes(blk, th, cfg, w)
    for d in range(3):
        e = relerr(got[d], ref[d])
        _say("  dir %d relerr %.3e" % (d, e))
        assert e <= TOL


@pytest.mark.parametrize("case", CASES)
def test_flux_divergence(ctx, case):
    config, dims = case
    blk, th, cfg, w = perturbed_state(config, dims)
    _say("divergence config=%d" % config)
    fs = orl.calculate_fluxes(blk, th, cfg, w)
    du0 = np.random.default_rng(3).uniform(-1, 1, blk.shape())
    got = ca.Hydro(ctx, blk, th, cfg).add_flux_divergence(0.37, fs, du0)
    ref = orl.add_flux_divergence(blk, 0.37, fs, du0)
    assert relerr(got, ref) <= TOL


@pytest.mark.parametrize("case", [CASES[1], CASES[2]])
@pytest.mark.parametrize("flag", [9, 1])
def test_implicit_solve(ctx, case, flag):
    config, dims = case
    blk, th, cfg, w = perturbed_state(config, dims)
    fs = orl.calculate_fluxes(blk, th, cfg, w)
    dt = 0.8
    du = orl.add_flux_divergence(blk, dt, fs, np.zeros(blk.shape()))
    du = orl.add_sources(blk, th, cfg, dt, w, du)
    _say("implicit config=%d flag=%d" % (config, flag))
    got = ca.ImplicitSolver(ctx, blk, th, flag).solve_implicit_3d(du, w, dt)
    ref, r = orl.implicit_solve(blk, th, du, w, dt, flag)
    assert r == 0
    sk, sj, si = blk.interior()
    e = relerr(got[:, sk, sj, si], ref[:, sk, sj, si])
    _say("  relerr %.3e" % e)
    assert e <= TOL


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("graph", [False, True])
def test_step(ctx, case, graph):
    config, dims = case
    blk, th, cfg, w = perturbed_state(config, dims, vel=1.0, dp=1e-3,
                                      cloud=True)
    u = orl.prim_to_cons(blk, th, w)
    dt = cfl_dt(blk, th, cfg, w, 0.6)
    nsteps = 3
    _say("step config=%d graph=%s" % (config, graph))
    run = ca.Run(ctx, blk, th, cfg)
    run.set_prim(w)
    run.step(dt, nsteps, graph=graph)
    ug, wg = run.cons(), run.prim()
    run.close()
    uo, wo = orl.step(blk, th, cfg, dt, nsteps, u, w)
    sk, sj, si = blk.interior()
    eu = relerr(ug[:, sk, sj, si], uo[:, sk, sj, si])
    ew = relerr(wg[:, sk, sj, si], wo[:, sk, sj, si])
    _say("  relerr u %.3e w %.3e" % (eu, ew))
    assert eu <= TOL and ew <= TOL


def test_graph_equals_direct_full_box(ctx):
    """The full 256x256x128 box: two graph-replayed steps equal two directly
    launched steps bit for bit, also for a second run in the same process
    (whose buffers reuse freed device memory)."""
    blk, th, cfg = ca.pgen_setup(2)
    w = ca.pgen_init(2, 1, blk, th, cfg)
    dt = cfl_dt(blk, th, cfg, w, 0.9)
    sk, sj, si = blk.interior()
    out = []
    for graph in (False, True, True):
        _say("full box graph=%s" % graph)
        run = ca.Run(ctx, blk, th, cfg)
        try:
            run.set_prim(w)
            run.step(dt, 2, graph=graph)
            out.append(run.cons()[:, sk, sj, si])
        finally:
            run.close()
    assert np.all(np.isfinite(out[0]))
    np.testing.assert_array_equal(out[1], out[0])
    np.testing.assert_array_equal(out[2], out[0])


@pytest.mark.parametrize("bcs", [(0, 0, 0, 2, 2, 0), (0, 2, 2, 1, 1, 0),
                                 (0, 0, 1, 1, 0, 2)])
@pytest.mark.parametrize("dims", [(20, 12, 10), (20, 5, 3), (16, 3, 6)])
def test_step_x2_x3_boundaries(ctx, bcs, dims):
    """Reflecting / periodic / outflow x2 and x3 sides: graph-replayed steps
    of a run equal the oracle, interior and x2 / x3 ghosts and corners."""
    blk, th, cfg, w = perturbed_state(2, dims, vel=1.0, dp=1e-3)
    for q in range(6):
        blk.bc[q] = bcs[q]
    w = ca.apply_bc_host(blk, w)
    u = orl.prim_to_cons(blk, th, w)
    dt = cfl_dt(blk, th, cfg, w, 0.5)
    _say("x2/x3 boundaries %s dims %s" % (bcs, dims))
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        run.step(dt, 2, graph=True)
        ug, wg = run.cons(), run.prim()
    finally:
        run.close()
    uo, wo = orl.step(blk, th, cfg, dt, 2, u, w)
    sk, sj, si = blk.interior()
    assert relerr(ug[:, sk, sj, si], uo[:, sk, sj, si]) <= TOL
    # every ghost and corner of w as the oracle's boundary fill leaves them
    # (the x1 ghosts are the run's own: compare the x2 / x3 ghost columns of
    # the interior levels)
    np.testing.assert_array_equal(wg[:, :, :, si], wo[:, :, :, si])


def test_blocks_thinner_than_the_ghost_zone_are_rejected(ctx):
    """A direction with cells needs NGHOST of them (its ghosts copy that
    many interior cells)."""
    blk, th, cfg = ca.pgen_setup(2, 16, 2, 6)
    with pytest.raises(ca.CanoeGpuError):
        ca.Run(ctx, blk, th, cfg)


@pytest.mark.parametrize("flag", [1, 0])
def test_step_partial_correction_and_explicit_3d(ctx, flag):
    """The 3x3 partial correction (flag 1) and the explicit path (flag 0) on
    a 3D block equal the oracle."""
    blk, th, cfg, w = perturbed_state(2, (20, 12, 10), vel=1.0, dp=1e-3)
    cfg.implicit_flag = flag
    u = orl.prim_to_cons(blk, th, w)
    dt = cfl_dt(blk, th, cfg, w, 0.3 if flag == 0 else 0.5)
    if flag == 0:
        dt = min(dt, 0.3 * blk.dx1 / 350.0)  # explicit x1: the sound speed sets it
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        run.step(dt, 2, graph=True)
        ug = run.cons()
    finally:
        run.close()
    uo, _ = orl.step(blk, th, cfg, dt, 2, u, w)
    sk, sj, si = blk.interior()
    assert relerr(ug[:, sk, sj, si], uo[:, sk, sj, si]) <= TOL
