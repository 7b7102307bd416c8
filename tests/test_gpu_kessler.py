"""GPU parity of the Kessler microphysics and precipitation sedimentation
(k_kessler, and the last stage's downward sweep of implicit runs):
cg_microphysics on caller arrays and moist runs that apply it after the last
stage equal the oracle (FP64 form, graph and direct, with and without the
implicit x1 solve); the sweep's operator equals the column kernel's bit for
bit; the FP32-flux form stays within 1e-5 of the FP64 form; water is
conserved by the operator."""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import cfl_dt, perturbed_state, relerr

pytestmark = pytest.mark.gpu

RATES = dict(k_auto=0.01, k_accr=0.02, k_evap=0.1, v_fall=8.0)


def _say(*a):
    print(*a, flush=True)


def _moist(dims=(20, 12, 10), seed=7, implicit_flag=9):
    blk, th, cfg, w = perturbed_state(4, dims, seed=seed, vel=1.0, dp=1e-3,
                                      cloud=True)
    sk, sj, si = blk.interior()
    rng = np.random.default_rng(seed + 1)
    # precipitation everywhere, cloud in patches, some sub-saturated cells
    w[3][sk, sj, si] = rng.uniform(0, 1e-3, w[3][sk, sj, si].shape)
    w[1][sk, sj, si] *= rng.uniform(0.5, 1.05, w[1][sk, sj, si].shape)
    w = ca.apply_bc_host(blk, w)
    for k, v in RATES.items():
        setattr(cfg, k, v)
    cfg.implicit_flag = implicit_flag
    return blk, th, cfg, w


def test_microphysics_entry_equals_oracle(ctx):
    blk, th, cfg, w = _moist()
    u = orl.prim_to_cons(blk, th, w)
    dt = 4.0
    _say("cg_microphysics 20x12x10 dt=%g" % dt)
    du, dw = ctx.array(u), ctx.array(w)
    try:
        assert ca.lib().cg_microphysics(ctx.h, ca._ref(blk), ca._ref(th),
                                        ca._ref(cfg), dt, du.ptr, dw.ptr) == ca.OK
        ug, wg = du.download(), dw.download()
    finally:
        du.free()
        dw.free()
    uo = orl.kessler(blk, th, cfg, dt, u)
    wo = orl.cons_to_prim(blk, th, uo, w)
    sk, sj, si = blk.interior()
    np.testing.assert_array_equal(ug[:, sk, sj, si], uo[:, sk, sj, si])
    np.testing.assert_array_equal(wg[:, sk, sj, si], wo[:, sk, sj, si])
    assert not np.array_equal(uo, u)
    water = lambda a: (a[1] + a[2] + a[3])[sk, sj, si].sum(axis=-1)
    np.testing.assert_allclose(water(ug), water(u), rtol=1e-14)


def test_microphysics_rejects_dry_thermo(ctx):
    blk, th, cfg, w = perturbed_state(2, (8, 6, 5))
    for k, v in RATES.items():
        setattr(cfg, k, v)
    u = orl.prim_to_cons(blk, th, w)
    du, dw = ctx.array(u), ctx.array(w)
    try:
        assert ca.lib().cg_microphysics(ctx.h, ca._ref(blk), ca._ref(th),
                                        ca._ref(cfg), 1.0, du.ptr,
                                        dw.ptr) == ca.ERR_ARG
    finally:
        du.free()
        dw.free()


@pytest.mark.parametrize("implicit_flag", [9, 0])
@pytest.mark.parametrize("graph", [True, False])
def test_run_with_kessler_equals_oracle(ctx, implicit_flag, graph):
    blk, th, cfg, w = _moist(implicit_flag=implicit_flag)
    dt = cfl_dt(blk, th, cfg, w, 0.5)
    nsteps = 3
    _say("kessler run flag=%d graph=%s dt=%.4f" % (implicit_flag, graph, dt))
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        run.step(dt, nsteps, graph=graph)
        ug, wg = run.cons(), run.prim()
    finally:
        run.close()
    u = orl.prim_to_cons(blk, th, w)
    uo, wo = orl.step(blk, th, cfg, dt, nsteps, u, w)
    sk, sj, si = blk.interior()
    eu = relerr(ug[:, sk, sj, si], uo[:, sk, sj, si])
    ew = relerr(wg[:, sk, sj, si], wo[:, sk, sj, si])
    _say("  relerr u %.3e w %.3e" % (eu, ew))
    assert eu <= 1e-12 and ew <= 1e-12
    # the microphysics did something: precipitation differs from a run
    # without it
    cfg0 = _moist(implicit_flag=implicit_flag)[2]
    for k in RATES:
        setattr(cfg0, k, 0.0)
    un, _ = orl.step(blk, th, cfg0, dt, nsteps, u, w)
    assert relerr(uo[3][sk, sj, si], un[3][sk, sj, si]) > 1e-6


def test_run_with_kessler_mixed_close_to_fp64(ctx):
    blk, th, cfg, w = _moist(dims=(24, 16, 12))
    dt = cfl_dt(blk, th, cfg, w, 0.5)
    out = []
    for prec in (ca.PREC_FP64, ca.PREC_FP32):
        cfg.precision = prec
        run = ca.Run(ctx, blk, th, cfg)
        try:
            run.set_prim(w)
            run.step(dt, 5, graph=True)
            out.append(run.cons())
        finally:
            run.close()
    sk, sj, si = blk.interior()
    e = relerr(out[1][:, sk, sj, si], out[0][:, sk, sj, si])
    _say("kessler mixed vs fp64 relerr %.3e" % e)
    assert e < 1e-5


@pytest.mark.parametrize("prec", [ca.PREC_FP64, ca.PREC_FP32])
@pytest.mark.parametrize("implicit_fused", [1, 0])
def test_sweep_microphysics_equals_the_column_operator(ctx, prec, implicit_fused):
    """Implicit runs with wall x1 sides apply the operator per cell inside
    the last stage's downward sweep: one step with it equals, bit for bit,
    one step without it followed by cg_microphysics (the separate column
    kernel) on that step's state, in both flux forms and with the forward
    and downward sweeps in one launch or in two."""
    blk, th, cfg, w = _moist()
    cfg.precision = prec
    dt = cfl_dt(blk, th, cfg, w, 0.5)
    _say("sweep microphysics prec=%d fused=%d dt=%.4f" % (prec, implicit_fused, dt))
    outs = []
    for on in (True, False):
        for k, v in RATES.items():
            setattr(cfg, k, v if on else 0.0)
        run = ca.Run(ctx, blk, th, cfg)
        try:
            run.set_tune(implicit_fused=implicit_fused)
            run.set_prim(w)
            run.step(dt, 1, graph=False)
            outs.append((run.cons(), run.prim()))
        finally:
            run.close()
    for k, v in RATES.items():
        setattr(cfg, k, v)
    du, dw = ctx.array(outs[1][0]), ctx.array(outs[1][1])
    try:
        assert ca.lib().cg_microphysics(ctx.h, ca._ref(blk), ca._ref(th),
                                        ca._ref(cfg), dt, du.ptr, dw.ptr) == ca.OK
        ug, wg = du.download(), dw.download()
    finally:
        du.free()
        dw.free()
    sk, sj, si = blk.interior()
    np.testing.assert_array_equal(outs[0][0][:, sk, sj, si], ug[:, sk, sj, si])
    np.testing.assert_array_equal(outs[0][1][:, sk, sj, si], wg[:, sk, sj, si])
    # the operator changed the state
    assert not np.array_equal(outs[0][0], outs[1][0])


@pytest.mark.parametrize("bcs", [(ca.BC_REFLECT, ca.BC_OUTFLOW), (ca.BC_OUTFLOW, ca.BC_REFLECT)])
def test_run_with_kessler_outflow_x1_equals_oracle(ctx, bcs):
    """The Kessler sweep of an implicit run also writes the column's x1
    ghosts (here an outflow side: copies of the boundary cell) and the next
    stage's hydrostatic pressures; the steps equal the oracle's."""
    blk, th, cfg, w = _moist()
    blk.bc[0], blk.bc[1] = bcs
    w = ca.apply_bc_host(blk, w)
    dt = cfl_dt(blk, th, cfg, w, 0.5)
    _say("kessler run, x1 bcs %s" % (bcs,))
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        run.step(dt, 3, graph=True)
        ug, wg = run.cons(), run.prim()
    finally:
        run.close()
    u = orl.prim_to_cons(blk, th, w)
    uo, wo = orl.step(blk, th, cfg, dt, 3, u, w)
    sk, sj, si = blk.interior()
    eu = relerr(ug[:, sk, sj, si], uo[:, sk, sj, si])
    ew = relerr(wg[:, sk, sj, si], wo[:, sk, sj, si])
    _say("  relerr u %.3e w %.3e" % (eu, ew))
    assert eu <= 1e-12 and ew <= 1e-12
