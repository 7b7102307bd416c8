"""Edge cases at the C-ABI boundary.

* Empty inputs are no-ops that return CG_OK, as the reference's loops over
  il .. iu simply do not run (lmars.cpp:19-85): cg_riemann with no faces,
  cg_saturation_adjust with no cells, cg_mlp_forward with no samples, and
  cg_run_step with no steps (the state is unchanged).
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
