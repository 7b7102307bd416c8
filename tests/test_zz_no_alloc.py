"""cg_implicit_solve_noalloc: the implicit solve with the fixed per-device
scratch (no device allocation) equals cg_implicit_solve bit for bit, on two
calls in a row with the second smaller than the first (the fixed scratch
keeps the first call's contents), for x1 columns between walls (flag 9)
and for periodic x3 / x2 lines plus x1 columns (flag 15); a solve whose
scratch exceeds the fixed size is rejected before any launch and leaves du
alone."""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import perturbed_state

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _inputs(config, dims, seed):
    blk, th, cfg, w = perturbed_state(config, dims, seed=seed)
    fs = orl.calculate_fluxes(blk, th, cfg, w)
    dt = 0.8
    du = orl.add_flux_divergence(blk, dt, fs, np.zeros(blk.shape()))
    du = orl.add_sources(blk, th, cfg, dt, w, du)
    return blk, th, w, du, dt


def _solve(ctx, fn, blk, th, w, du, dt, flag):
    ddu, dw = ctx.array(du), ctx.array(w)
    try:
        rc = fn(ctx.h, ca._ref(blk), ca._ref(th), ddu.ptr, dw.ptr, dt, flag)
        return rc, ddu.download()
    finally:
        ddu.free()
        dw.free()


@pytest.mark.parametrize("flag,sizes", [
    (9, [(64, 32, 24), (24, 12, 10)]),
    (15, [(24, 20, 16), (12, 8, 6)]),
])
def test_noalloc_equals_allocating_entry(ctx, flag, sizes):
    L = ca.lib()
    for n, dims in enumerate(sizes):
        _say("noalloc flag=%d call %d dims=%s" % (flag, n, dims))
        blk, th, w, du, dt = _inputs(2, dims, seed=11 + n)
        rc0, ref = _solve(ctx, L.cg_implicit_solve, blk, th, w, du, dt, flag)
        rc1, got = _solve(ctx, L.cg_implicit_solve_noalloc, blk, th, w, du, dt,
                          flag)
        assert rc0 == ca.OK and rc1 == ca.OK, (rc0, rc1, ca.lib().cg_last_error(ctx.h))
        assert not np.array_equal(ref, du)
        np.testing.assert_array_equal(got, ref)


def test_noalloc_rejects_oversized_solve(ctx):
    L = ca.lib()
    blk, th, cfg, w = perturbed_state(2, (128, 64, 40))
    cells = blk.nx1 * blk.nx2 * blk.nx3
    assert 30 * 8 * cells > ca.NOALLOC_SCRATCH_BYTES
    du = np.full(blk.shape(), 0.25)
    _say("noalloc oversized: %d cells" % cells)
    rc, got = _solve(ctx, L.cg_implicit_solve_noalloc, blk, th, w, du, 0.8, 9)
    assert rc == ca.ERR_ARG
    assert b"scratch" in ca.lib().cg_last_error(ctx.h)
    np.testing.assert_array_equal(got, du)
