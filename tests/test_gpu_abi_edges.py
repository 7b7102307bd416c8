"""Edge cases at the C-ABI boundary.

* Empty inputs are no-ops that return CG_OK, as the reference's loops over
  il .. iu simply do not run (lmars.cpp:19-85): cg_riemann with no faces,
  cg_saturation_adjust with no cells, cg_mlp_forward with no samples, and
  cg_run_step with no steps (the state is unchanged).
* Host arrays: the cg_*_host variants stage the caller's arrays through
  the device and equal the device calls.
* Ragged rows: the reference's wl_, wr_ are (N, nc1) rows of which faces
  il .. iu are used and its flx is the caller's (N, k, j, i) array
  (lmars.cpp:19-22), so cg_riemann takes leading dimensions larger than the
  face count; the faces past n_faces of the output are left alone and the
  others equal the oracle.
"""
import numpy as np
import pytest

import canoe_amd as ca
import canoe_amd.mlp as cm
import oracle_lib as orl
from helpers import perturbed_state, relerr

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_empty_inputs_are_no_ops(ctx):
    L = ca.lib()
    blk, th, cfg, w = perturbed_state(2, (8, 6, 5))
    _say("empty: riemann, run step")
    for solver in (ca.RS_LMARS, ca.RS_HLLC):
        assert L.cg_riemann(ctx.h, ca._ref(th), solver, 1, 0, None, None, 0,
                            None, 0) == ca.OK
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        u0 = run.cons()
        run.step(0.1, 0, graph=True)
        run.step(0.1, 0, graph=False)
        np.testing.assert_array_equal(run.cons(), u0)
    finally:
        run.close()
    _say("empty: saturation adjustment, network forward")
    blk, th, cfg, w = perturbed_state(4, (8, 6, 5))
    assert L.cg_saturation_adjust(ctx.h, ca._ref(th), 0, None, None, None,
                                  None, None) == ca.OK
    net = cm.Mlp(ctx, seed=3)
    try:
        assert L.cg_mlp_forward(net.h, 0, None, None) == ca.OK
    finally:
        net.close()
    ctx.synchronize()


@pytest.mark.parametrize("solver", [ca.RS_LMARS, ca.RS_HLLC])
def test_riemann_ragged_rows(ctx, solver):
    blk, th, cfg, w = perturbed_state(2, (8, 6, 5))
    rng = np.random.default_rng(17 + solver)
    N, nf, ld, fld = 5, 37, 64, 50

    def states():
        s = np.zeros((N, ld))
        s[0] = rng.uniform(0.2, 1.5, ld)
        s[1:4] = rng.uniform(-60, 60, (3, ld))
        s[4] = rng.uniform(2e4, 1.1e5, ld)
        return s
    wl, wr = states(), states()
    sentinel = np.full((N, fld), 7.25)
    for ivx in (1, 2, 3):
        _say("ragged riemann solver=%d ivx=%d nf=%d ld=%d flx_ld=%d"
             % (solver, ivx, nf, ld, fld))
        dl, dr, df = ctx.array(wl), ctx.array(wr), ctx.array(sentinel)
        try:
            ctx._check(ca.lib().cg_riemann(ctx.h, ca._ref(th), solver, ivx, nf,
                                           dl.ptr, dr.ptr, ld, df.ptr, fld),
                       "cg_riemann")
            got = df.download()
        finally:
            for a in (dl, dr, df):
                a.free()
        ref = orl.riemann(th, solver, ivx, wl[:, :nf], wr[:, :nf])
        assert relerr(got[:, :nf], ref) <= 1e-12
        np.testing.assert_array_equal(got[:, nf:], sentinel[:, nf:])


def test_host_variants(ctx):
    """cg_riemann_host, cg_calculate_fluxes_host and cg_implicit_solve_host
    on host arrays equal the oracle and the device calls; entries the device
    call leaves alone keep the caller's values."""
    L = ca.lib()
    blk, th, cfg, w = perturbed_state(2, (8, 6, 5))
    rng = np.random.default_rng(29)
    N, nf, ld, fld = 5, 41, 48, 45
    wl = np.zeros((N, ld))
    wr = np.zeros((N, ld))
    for s in (wl, wr):
        s[0] = rng.uniform(0.2, 1.5, ld)
        s[1:4] = rng.uniform(-60, 60, (3, ld))
        s[4] = rng.uniform(2e4, 1.1e5, ld)
    flx = np.full((N, fld), -3.5)
    _say("host riemann nf=%d ld=%d flx_ld=%d" % (nf, ld, fld))
    assert L.cg_riemann_host(ctx.h, ca._ref(th), ca.RS_HLLC, 2, nf, ca._ptr(wl),
                             ca._ptr(wr), ld, ca._ptr(flx), fld) == ca.OK
    ref = orl.riemann(th, ca.RS_HLLC, 2, wl[:, :nf], wr[:, :nf])
    assert relerr(flx[:, :nf], ref) <= 1e-12
    assert np.all(flx[:, nf:] == -3.5)

    _say("host calculate_fluxes")
    cfg.order_h, cfg.order_v, cfg.rsolver = 5, 2, ca.RS_HLLC
    w = np.ascontiguousarray(w, np.float64)  # held while the call reads it
    dev = ca.Hydro(ctx, blk, th, cfg).calculate_fluxes(w)
    host = [np.zeros(s) for s in blk.flux_shapes()]
    assert L.cg_calculate_fluxes_host(ctx.h, ca._ref(blk), ca._ref(th), ca._ref(cfg),
                                      ca._ptr(w), *[ca._ptr(f) for f in host]) == ca.OK
    for a, b in zip(host, dev):
        np.testing.assert_array_equal(a, b)

    _say("host implicit_solve flag 9")
    du = np.zeros(blk.shape())
    sk, sj, si = blk.interior()
    du[:, sk, sj, si] = rng.normal(size=du[:, sk, sj, si].shape) * np.array(
        [1e-3, 1e-2, 1e-2, 1e-2, 3e2])[:, None, None, None]
    got = du.copy()
    assert L.cg_implicit_solve_host(ctx.h, ca._ref(blk), ca._ref(th), ca._ptr(got),
                                    ca._ptr(w), 0.5, 9) == ca.OK
    ref, r = orl.implicit_solve(blk, th, du, w, 0.5, 9)
    assert r == 0
    assert relerr(got[:, sk, sj, si], ref[:, sk, sj, si]) <= 1e-12
    mask = np.ones(blk.shape(), bool)
    mask[:, sk, sj, si] = False
    np.testing.assert_array_equal(got[mask], du[mask])


def test_host_fluxes_2d_without_x3_array(ctx):
    """A 2D block (config 1, nx3 = 1) has no x3 flux array: the host flux
    sweep takes a null flux3 and equals the oracle on flux1 / flux2."""
    L = ca.lib()
    blk, th, cfg, w = perturbed_state(1, (24, 40, 1))
    w = np.ascontiguousarray(w, np.float64)
    f1 = np.zeros(blk.flux_shapes()[0])
    f2 = np.zeros(blk.flux_shapes()[1])
    _say("host calculate_fluxes 2D, flux3 = NULL")
    assert L.cg_calculate_fluxes_host(ctx.h, ca._ref(blk), ca._ref(th), ca._ref(cfg),
                                      ca._ptr(w), ca._ptr(f1), ca._ptr(f2),
                                      None) == ca.OK
    ref = orl.calculate_fluxes(blk, th, cfg, w)
    assert relerr(f1, ref[0]) <= 1e-12 and relerr(f2, ref[1]) <= 1e-12
