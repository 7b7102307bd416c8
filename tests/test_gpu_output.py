"""GPU parity of the output reductions over the resident state of runs and
batches: Athena++'s default history sums (relative 1e-12: device partial
sums in a fixed order against the oracle's serial sum) and DebugOutput's
extremes and their coordinates (exact, including ties and the reference's
initial values)."""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import perturbed_state

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _oracle_debug(bl):
    """DebugOutput over blocks in order: strict comparisons across blocks."""
    N = bl[0][0].N
    fmax, fmin = float(np.finfo(np.float32).max), float(np.finfo(np.float32).tiny)
    vmin, vmax = np.full(N, fmax), np.full(N, fmin)
    xmin, xmax = np.full((N, 3), fmax), np.full((N, 3), fmin)
    for blk, w in bl:
        a, b, xa, xb = orl.debug(blk, w)
        for n in range(N):
            if a[n] < vmin[n]:
                vmin[n], xmin[n] = a[n], xa[n]
            if b[n] > vmax[n]:
                vmax[n], xmax[n] = b[n], xb[n]
    return vmin, vmax, xmin, xmax


@pytest.mark.parametrize("config,dims", [(1, (24, 40, 1)), (2, (20, 12, 10)),
                                         (4, (20, 12, 10))])
def test_history_and_debug_single_block(ctx, config, dims):
    blk, th, cfg, w = perturbed_state(config, dims, vel=3.0)
    blk.x1min, blk.x2min = 10.0, -400.0
    _say("history/debug config=%d" % config)
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        h = run.history()
        dbg = run.debug()
        u, wr = run.cons(), run.prim()
    finally:
        run.close()
    np.testing.assert_allclose(h, orl.history(blk, u), rtol=1e-12, atol=0)
    want = orl.debug(blk, wr)
    for a, b in zip(dbg, want):
        np.testing.assert_array_equal(a, b)


def test_debug_ties_and_batch_order(ctx):
    """Equal extremes planted in two blocks and twice inside one block: the
    first block in order and the first cell in (k, j, i) order win."""
    specs = [(40, 8, 8), (24, 10, 10), (32, 16, 4)]
    bl = []
    for b, dims in enumerate(specs):
        blk, th, cfg = ca.pgen_setup(5, *dims)
        blk.x2min = 1.0e5 * b
        w = ca.pgen_init(5, 70 + b, blk, th, cfg)
        bl.append([blk, th, cfg, w])
    ivx = 1
    for b in (1, 2):
        bl[b][3][ivx + 1, 3 + 1, 3 + 2, 3 + 5] = 55.0   # v2 maximum in blocks 1, 2
        bl[b][3][ivx + 1, 3 + 0, 3 + 7, 3 + 9] = 55.0   # and earlier in k
        bl[b][3] = ca.apply_bc_host(bl[b][0], bl[b][3])
    th, cfg = bl[0][1], bl[0][2]
    run = ca.Run(ctx, None, th, cfg, blocks=[b[0] for b in bl])
    try:
        for i, b in enumerate(bl):
            run.set_prim(b[3], i)
        dbg = run.debug()
        h = run.history()
        us = [run.cons(i) for i in range(len(bl))]
    finally:
        run.close()
    want = _oracle_debug([(b[0], b[3]) for b in bl])
    for a, b in zip(dbg, want):
        np.testing.assert_array_equal(a, b)
    # block 1's cell (i, j, k) = (9, 7, 0), the first of its two maxima
    b1 = bl[1][0]
    assert dbg[1][ivx + 1] == 55.0
    np.testing.assert_array_equal(
        dbg[3][ivx + 1], [b1.x1min + 9.5 * b1.dx1, 1.0e5 + 7.5 * b1.dx2,
                          b1.x3min + 0.5 * b1.dx3])
    hs = sum(orl.history(b[0], u) for b, u in zip(bl, us))
    np.testing.assert_allclose(h, hs, rtol=1e-12, atol=0)


@pytest.mark.parametrize("n,value,code", [(0, -1e-3, ca.ERR_NEGATIVE),
                                          (7, -5.0, ca.ERR_NEGATIVE),
                                          (4, np.nan, ca.ERR_NAN),
                                          (7, 1e-3, ca.ERR_COLD)])
def test_state_checks_match_oracle(ctx, n, value, code):
    """The run's state checks report the first failing cell of the first
    failing block, as the oracle's serial scan; the device main loop stops
    with the same status before stepping."""
    specs = [(24, 8, 8), (20, 10, 6), (24, 6, 12)]
    bl = []
    for b, dims in enumerate(specs):
        blk, th, cfg = ca.pgen_setup(4, *dims)
        w = ca.pgen_init(4, 90 + b, blk, th, cfg)
        bl.append([blk, th, cfg, w])
    for b in (1, 2):  # bad cells in blocks 1 and 2, two in block 1
        for (k, j, i) in ((3, 1, 2), (1, 4, 7)):
            if b == 2 and k == 3:
                continue
            bl[b][3][n, 3 + k, 3 + j, 3 + i] = value
            if code == ca.ERR_COLD:
                bl[b][3][1:4, 3 + k, 3 + j, 3 + i] = 0.0
    th, cfg = bl[0][1], bl[0][2]
    run = ca.Run(ctx, None, th, cfg, blocks=[b[0] for b in bl])
    try:
        for i, b in enumerate(bl):
            run.set_prim(b[3], i)
        st, where = run.check()
        with pytest.raises(ca.CanoeGpuError) as ex:
            run.advance(0.45, nlim=2)
    finally:
        run.close()
    want = [orl.check(b[0], b[1], b[3]) for b in bl]
    first = next(i for i, c in enumerate(want) if c[0] < 3)
    assert first == 1
    assert st == code and where == (first,) + want[first][1], (st, where, want)
    assert ex.value.code == code


def test_state_checks_pass_on_a_physical_run(ctx):
    blk, th, cfg, w = perturbed_state(4, (20, 12, 10), vel=3.0)
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        assert run.check() == (0, (-1, -1, -1, -1))
        t, dt, n = run.advance(0.45, nlim=2)
        assert n == 2 and run.check()[0] == 0
    finally:
        run.close()
