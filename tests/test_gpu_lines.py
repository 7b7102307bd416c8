"""GPU parity of SolveImplicit3D's x2 / x3 lines and periodic lines
(cg_implicit_solve with implicit_flag bits 1 and 2, and directions whose two
boundaries are periodic) against the oracle on the same seeded inputs.

Both evaluate the same expressions in FP64 without contraction, so the bar
is 1e-12 relative per cell (in practice equality).
"""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import perturbed_state, relerr
from test_oracle_lines import _case, _set_bc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a, flush=True)


def _check(ctx, blk, th, w, du, dt, flag):
    got = ca.ImplicitSolver(ctx, blk, th, flag).solve_implicit_3d(du, w, dt)
    ref, r = orl.implicit_solve(blk, th, du, w, dt, flag)
    assert r == 0
    sk, sj, si = blk.interior()
    e = relerr(got[:, sk, sj, si], ref[:, sk, sj, si])
    _say("  relerr %.3e" % e)
    assert e <= TOL
    # cells outside the interior are left alone
    mask = np.ones(blk.shape(), bool)
    mask[:, sk, sj, si] = False
    np.testing.assert_array_equal(got[mask], du[mask])
    return got


@pytest.mark.parametrize("flag", [10, 12, 14, 15, 2, 4, 7])
@pytest.mark.parametrize("bc", [ca.BC_PERIODIC, ca.BC_REFLECT])
def test_x2_x3_lines(ctx, flag, bc):
    blk, th, w, du = _case(bc, bc)
    _say("lines flag=%d bc=%d" % (flag, bc))
    _check(ctx, blk, th, w, du, 0.7, flag)


@pytest.mark.parametrize("flag", [9, 1, 15])
def test_periodic_x1_columns(ctx, flag):
    blk, th, w, du = _case(ca.BC_PERIODIC, ca.BC_PERIODIC)
    bc = list(blk.bc)
    bc[0] = bc[1] = ca.BC_PERIODIC
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    _say("periodic x1 flag=%d" % flag)
    _check(ctx, blk, th, w, du, 0.7, flag)


def test_uneven_extents_and_2d_block(ctx):
    # lanes run along the transverse index of smaller stride; extents that
    # are not multiples of the wavefront and a block with nx3 = 1
    blk, th, w, du = _case(ca.BC_PERIODIC, ca.BC_REFLECT, dims=(13, 7, 5), seed=3)
    _say("lines uneven 13x7x5")
    _check(ctx, blk, th, w, du, 0.4, 15)
    blk, th, cfg, w = perturbed_state(1, (12, 20, 1), vel=20.0)
    du = np.zeros(blk.shape())
    sk, sj, si = blk.interior()
    du[:, sk, sj, si] = np.random.default_rng(5).normal(
        size=du[:, sk, sj, si].shape) * np.array(
            [1e-3, 1e-2, 1e-2, 1e-2, 3e2])[:, None, None, None]
    _say("lines 2D 12x20x1")
    for flag in (10, 14, 11):  # x3 is skipped: nx3 = 1
        _check(ctx, blk, th, w, du, 0.4, flag)


@pytest.mark.parametrize("bc", [ca.BC_PERIODIC, ca.BC_REFLECT])
def test_shortest_lines(ctx, bc):
    # three cells, the smallest extent a block takes: the first level, the
    # cyclic corner terms and the last level of every line meet
    blk, th, w, du = _case(bc, bc, dims=(3, 3, 3), seed=21)
    for flag in (15, 7):
        _say("lines 3x3x3 flag=%d bc=%d" % (flag, bc))
        _check(ctx, blk, th, w, du, 0.5, flag)
    bcs = list(blk.bc)
    bcs[0] = bcs[1] = ca.BC_PERIODIC
    _set_bc(blk, bcs)
    w = ca.apply_bc_host(blk, w)
    _say("periodic x1 3x3x3 bc=%d" % bc)
    _check(ctx, blk, th, w, du, 0.5, 15)


def test_moist_periodic_lines(ctx):
    blk, th, cfg, w = perturbed_state(4, (8, 10, 6), vel=10.0)
    du = np.zeros(blk.shape())
    sk, sj, si = blk.interior()
    du[:, sk, sj, si] = np.random.default_rng(8).normal(
        size=du[:, sk, sj, si].shape) * 1e-3
    _say("moist periodic lines")
    _check(ctx, blk, th, w, du, 0.5, 15)


@pytest.mark.parametrize("flag", [10, 12, 14, 2, 4, 6])
def test_species_with_walled_lines(ctx, flag):
    """Moist block, walls at both x2 and x3 ends: the species weights read w
    at the shuffled index of the reference's line buffer (w_shuffled)."""
    blk, th, cfg, w = perturbed_state(4, (8, 10, 6), vel=10.0)
    bc = list(blk.bc)
    bc[2] = bc[3] = bc[4] = bc[5] = ca.BC_REFLECT
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    du = np.zeros(blk.shape())
    sk, sj, si = blk.interior()
    du[:, sk, sj, si] = np.random.default_rng(5).normal(
        size=du[:, sk, sj, si].shape) * 1e-3
    _say("moist walled lines flag=%d" % flag)
    _check(ctx, blk, th, w, du, 0.5, flag)


def test_species_lines_past_the_end_of_w_rejected(ctx):
    blk, th, cfg, w = perturbed_state(4, (8, 10, 1), vel=10.0)
    bc = list(blk.bc)
    bc[2] = bc[3] = ca.BC_REFLECT
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    with pytest.raises(ca.CanoeGpuError):
        ca.ImplicitSolver(ctx, blk, th, 10).solve_implicit_3d(
            np.zeros(blk.shape()), w, 0.5)


@pytest.mark.parametrize("flag", [15, 14, 11, 7])
@pytest.mark.parametrize("graph", [False, True])
def test_run_with_lines_matches_oracle(ctx, flag, graph):
    """A run with x2 / x3 lines (SolveImplicit3D's x3, x2, then x1 solves in
    every stage, the step read from the device) against the oracle's step."""
    blk, th, cfg, w = perturbed_state(2, (20, 12, 10), vel=1.0, dp=1e-3)
    cfg.implicit_flag = flag
    dt = orl.new_block_dt(blk, th, cfg, w, 0.5)
    _say("run lines flag=%d graph=%s" % (flag, graph))
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        run.step(dt, 3, graph=graph)
        ug = run.cons()
        dtg = run.new_dt(0.5)
        wg = run.prim()
    finally:
        run.close()
    u0 = orl.prim_to_cons(blk, th, w)
    uo, _ = orl.step(blk, th, cfg, dt, 3, u0, w)
    sk, sj, si = blk.interior()
    e = relerr(ug[:, sk, sj, si], uo[:, sk, sj, si])
    _say("  relerr %.3e" % e)
    assert e <= TOL
    assert dtg == orl.new_block_dt(blk, th, cfg, wg, 0.5)


def test_moist_run_with_walled_lines(ctx):
    """A moist run with walls at the x2 ends and x2 lines + x1 columns in
    every stage (the species weights of the x2 lines read the run's resident
    w at the shuffled index) against the oracle's step."""
    blk, th, cfg, w = perturbed_state(4, (12, 10, 8), vel=1.0, dp=1e-3)
    bc = list(blk.bc)
    bc[2] = bc[3] = ca.BC_REFLECT
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    cfg.implicit_flag = 11
    dt = orl.new_block_dt(blk, th, cfg, w, 0.5)
    _say("moist run, walled x2 lines, flag 11")
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        run.step(dt, 3, graph=True)
        ug = run.cons()
    finally:
        run.close()
    u0 = orl.prim_to_cons(blk, th, w)
    uo, _ = orl.step(blk, th, cfg, dt, 3, u0, w)
    sk, sj, si = blk.interior()
    e = relerr(ug[:, sk, sj, si], uo[:, sk, sj, si])
    _say("  relerr %.3e" % e)
    assert e <= TOL


def test_mixed_run_with_lines(ctx):
    blk, th, cfg, w = perturbed_state(2, (24, 16, 12), vel=1.0, dp=1e-3)
    cfg.implicit_flag = 15
    dt = orl.new_block_dt(blk, th, cfg, w, 0.5)
    us = {}
    for prec in (ca.PREC_FP64, ca.PREC_FP32):
        cfg.precision = prec
        _say("mixed run lines prec=%d" % prec)
        run = ca.Run(ctx, blk, th, cfg)
        try:
            run.set_prim(w)
            run.step(dt, 10, graph=True)
            us[prec] = run.cons()
        finally:
            run.close()
    sk, sj, si = blk.interior()
    a, b = us[ca.PREC_FP32][:, sk, sj, si], us[ca.PREC_FP64][:, sk, sj, si]
    for idx in ([0], [1, 2, 3], [4]):
        e = np.sqrt(np.sum((a[idx] - b[idx]) ** 2) / np.sum(b[idx] ** 2))
        assert e <= 1e-5, (idx, e)


@pytest.mark.parametrize("flag", [9, 1, 11])
def test_run_with_periodic_x1_matches_oracle(ctx, flag):
    """A run whose x1 ends are periodic solves its x1 columns by the cyclic
    recurrences (no gravity: a vertically periodic box) and equals the
    oracle's steps."""
    blk, th, cfg, w = perturbed_state(2, (20, 12, 10), vel=1.0, dp=1e-3)
    blk.grav = 0.0
    bc = list(blk.bc)
    bc[0] = bc[1] = ca.BC_PERIODIC
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    cfg.implicit_flag = flag
    dt = orl.new_block_dt(blk, th, cfg, w, 0.5)
    _say("run periodic x1 flag=%d" % flag)
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        run.step(dt, 3, graph=True)
        ug = run.cons()
    finally:
        run.close()
    uo, _ = orl.step(blk, th, cfg, dt, 3, orl.prim_to_cons(blk, th, w), w)
    sk, sj, si = blk.interior()
    e = relerr(ug[:, sk, sj, si], uo[:, sk, sj, si])
    _say("  relerr %.3e" % e)
    assert e <= TOL


def test_run_rejects_one_sided_periodic_x1(ctx):
    blk, th, cfg, w = perturbed_state(2, (20, 12, 10))
    bc = list(blk.bc)
    bc[0] = ca.BC_PERIODIC
    _set_bc(blk, bc)
    with pytest.raises(ca.CanoeGpuError):
        ca.Run(ctx, blk, th, cfg)


def test_batch_with_lines_matches_oracle(ctx):
    """Every block of an uneven batch (partial 64-column groups, one shared
    line scratch) with x3, x2 and x1 solves equals its own oracle run."""
    from test_gpu_batch import _blocks, _oracle, _run_batch
    bl = _blocks(seed=21)
    for b in bl:
        b[2].implicit_flag = 15
    dt = min(orl.new_block_dt(b[0], b[1], b[2], b[3], 0.5) for b in bl)
    _say("batch lines dt=%.4f" % dt)
    ug = _run_batch(ctx, bl, dt, 2)
    uo = _oracle(bl, dt, 2)
    for i, (blk, th, cfg, w) in enumerate(bl):
        sk, sj, si = blk.interior()
        np.testing.assert_array_equal(ug[i][:, sk, sj, si], uo[i][:, sk, sj, si])


def test_implicit_flag_out_of_range_rejected(ctx):
    blk, th, w, du = _case(ca.BC_PERIODIC, ca.BC_PERIODIC)
    for flag in (16, -1):
        with pytest.raises(ca.CanoeGpuError):
            ca.ImplicitSolver(ctx, blk, th, flag).solve_implicit_3d(du, w, 0.5)
    # only directions without cells: nothing to do, du unchanged
    blk2, th2, cfg2, w2 = perturbed_state(1, (12, 20, 1), vel=20.0)
    du2 = np.random.default_rng(1).normal(size=blk2.shape())
    got = ca.ImplicitSolver(ctx, blk2, th2, 12).solve_implicit_3d(du2, w2, 0.5)
    np.testing.assert_array_equal(got, du2)
