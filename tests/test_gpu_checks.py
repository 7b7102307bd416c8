"""GPU face checks (check_reconstruct, src/checks.cpp:79-97): the flux
kernels flag a negative reconstructed face density or pressure and the call
returns CG_ERR_NEGATIVE, where the reference built with glog aborts.  The
states are the oracle pins' (tests/test_oracle_checks.py): a positive state
passes; a negative cell fails on the GPU exactly when the oracle counts
failing faces, and the fluxes still equal the oracle's (NaN where the
Riemann solver takes a negative pressure).  A run holding such a cell
reports the same code from its step, in the FP64 and the FP32-flux form."""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl

pytestmark = pytest.mark.gpu

CG_ERR_NEGATIVE = 4


def _say(*a):
    print(*a, flush=True)


def _state(neg):
    blk, th, cfg = ca.pgen_setup(2, 16, 12, 10)
    blk.grav = 0.0
    cfg.order_h = cfg.order_v = 2
    w = np.zeros(blk.shape())
    w[0] = 1.2
    w[4] = 1e5
    w[1], w[2], w[3] = 0.3, -0.2, 0.1
    for (v, k, j, i, x) in neg:
        w[v, k, j, i] = x
    return blk, th, cfg, w


def _gpu_fluxes(ctx, blk, th, cfg, w):
    dw = ctx.array(w)
    fs = [ctx.array(np.zeros(s)) for s in blk.flux_shapes()]
    try:
        rc = ca.lib().cg_calculate_fluxes(ctx.h, ca._ref(blk), ca._ref(th),
                                          ca._ref(cfg), dw.ptr, fs[0].ptr,
                                          fs[1].ptr, fs[2].ptr)
        return rc, [f.download() for f in fs]
    finally:
        for a in [dw] + fs:
            a.free()


@pytest.mark.parametrize("neg", [[], [(4, 6, 7, 9, -1.0)], [(0, 6, 7, 9, -1.0)],
                                 [(4, 6, 7, 9, -1.0), (4, 6, 7, 10, -2.0)]])
@pytest.mark.parametrize("rs", [ca.RS_LMARS, ca.RS_HLLC])
def test_calculate_fluxes_reports_negative_faces(ctx, neg, rs):
    blk, th, cfg, w = _state(neg)
    cfg.rsolver = rs
    _say("face check: negative cells %s, solver %d" % (neg, rs))
    rc, fg = _gpu_fluxes(ctx, blk, th, cfg, w)
    fo = orl.calculate_fluxes(blk, th, cfg, w)
    nf = orl.face_failures()
    _say("  rc %d, oracle failing faces %d" % (rc, nf))
    assert nf == (0 if not neg else (11 if len(neg) == 2 else 6))
    assert rc == (CG_ERR_NEGATIVE if nf else 0)
    sk, sj, si = blk.interior()
    for d, (a, b) in enumerate(zip(fg, fo)):
        s = [slice(None), sk, sj, si]
        s[3 - d] = slice(s[3 - d].start, s[3 - d].stop + 1)
        np.testing.assert_array_equal(a[tuple(s)], b[tuple(s)])


@pytest.mark.parametrize("prec", [ca.PREC_FP64, ca.PREC_FP32])
def test_run_step_reports_negative_faces(ctx, prec):
    blk, th, cfg, w = _state([])
    cfg.precision = prec
    _say("run face check, precision %d" % prec)
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        run.step(0.01, 1, graph=True)  # a positive state steps
        w2 = w.copy()
        w2[4, 6, 7, 9] = -1.0
        run.set_prim(w2)
        with pytest.raises(ca.CanoeGpuError) as e:
            run.step(0.01, 1, graph=True)
        assert e.value.code == CG_ERR_NEGATIVE
        run.set_prim(w)  # the status word was cleared: the next step passes
        run.step(0.01, 1, graph=True)
    finally:
        run.close()


def test_host_variant_reports_and_returns_fluxes(ctx):
    blk, th, cfg, w = _state([(4, 6, 7, 9, -1.0)])
    fs = [np.full(s, 7.0) for s in blk.flux_shapes()]
    rc = ca.lib().cg_calculate_fluxes_host(
        ctx.h, ca._ref(blk), ca._ref(th), ca._ref(cfg), ca._ptr(w), ca._ptr(fs[0]),
        ca._ptr(fs[1]), ca._ptr(fs[2]))
    assert rc == CG_ERR_NEGATIVE
    fo = orl.calculate_fluxes(blk, th, cfg, w)
    sk, sj, si = blk.interior()
    for d, (a, b) in enumerate(zip(fs, fo)):
        s = [slice(None), sk, sj, si]
        s[3 - d] = slice(s[3 - d].start, s[3 - d].stop + 1)
        np.testing.assert_array_equal(a[tuple(s)], b[tuple(s)])
