"""GPU parity of the time step: Hydro::NewBlockTimeStep on caller arrays and
on the resident state of runs and batches, and the device main loop
(cg_run_advance: step, NewBlockTimeStep reduction, Mesh::NewTimeStep) against
the same loop driven by the oracle.

The step is a minimum of per-cell values computed with the same expressions
in FP64, so any reduction order gives the same double: the bar is equality.
"""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import perturbed_state, relerr

pytestmark = pytest.mark.gpu

CASES = [(1, (24, 40, 1)), (2, (20, 12, 10)), (4, (20, 12, 10))]
DBL_MAX = np.finfo(np.float64).max


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("cfl", [0.9, 0.35])
def test_new_block_dt_matches_oracle(ctx, case, cfl):
    config, dims = case
    blk, th, cfg, w = perturbed_state(config, dims, vel=30.0)
    _say("new_block_dt config=%d cfl=%g" % (config, cfl))
    got = ca.Hydro(ctx, blk, th, cfg).new_block_dt(w, cfl)
    want = orl.new_block_dt(blk, th, cfg, w, cfl)
    assert got == want, (got, want)


def test_new_block_dt_explicit_flag(ctx):
    # the same state with the x1 direction explicit: c_s enters x1 as well
    blk, th, cfg, w = perturbed_state(2, (20, 12, 10), vel=30.0)
    cfg.implicit_flag = 0
    got = ca.Hydro(ctx, blk, th, cfg).new_block_dt(w, 0.9)
    assert got == orl.new_block_dt(blk, th, cfg, w, 0.9)


@pytest.mark.parametrize("case", CASES)
def test_run_new_dt_matches_oracle(ctx, case):
    config, dims = case
    blk, th, cfg, w = perturbed_state(config, dims, vel=30.0)
    _say("run new_dt config=%d" % config)
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        got = run.new_dt(0.9)
        wr = run.prim()
    finally:
        run.close()
    assert got == orl.new_block_dt(blk, th, cfg, wr, 0.9)
    assert got == orl.new_block_dt(blk, th, cfg, w, 0.9)


def test_batch_new_dt_is_min_over_blocks(ctx):
    specs = [(40, 8, 8), (64, 12, 6), (24, 10, 10), (32, 16, 4)]
    bl = []
    for b, dims in enumerate(specs):
        blk, th, cfg = ca.pgen_setup(5, *dims)
        w = ca.pgen_init(5, 300 + b, blk, th, cfg)
        sk, sj, si = blk.interior()
        w[2][sk, sj, si] += (b + 1) * 3.0  # a different fastest block each
        w = ca.apply_bc_host(blk, w)
        bl.append((blk, th, cfg, w))
    th, cfg = bl[0][1], bl[0][2]
    run = ca.Run(ctx, None, th, cfg, blocks=[b[0] for b in bl])
    try:
        for i, b in enumerate(bl):
            run.set_prim(b[3], i)
        got = run.new_dt(0.7)
    finally:
        run.close()
    per = [orl.new_block_dt(b[0], b[1], b[2], b[3], 0.7) for b in bl]
    assert got == min(per)


def _oracle_loop(blk, th, cfg, w, cfl, tlim, nlim):
    """main.cpp:114-152 driven by the oracle: the reference's step, then
    Mesh::NewTimeStep from the blocks' NewBlockTimeStep."""
    u = orl.prim_to_cons(blk, th, w)
    time, ncycle = 0.0, 0
    dt = orl.new_mesh_dt(DBL_MAX, orl.new_block_dt(blk, th, cfg, w, cfl),
                         time, tlim)
    dts = []
    while time < tlim and ncycle < nlim:
        u, w = orl.step(blk, th, cfg, dt, 1, u, w)
        dts.append(dt)
        ncycle += 1
        time += dt
        dt = orl.new_mesh_dt(dt, orl.new_block_dt(blk, th, cfg, w, cfl),
                             time, tlim)
    return u, time, dt, ncycle, dts


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("graph", [False, True])
def test_advance_matches_oracle_loop(ctx, case, graph):
    config, dims = case
    blk, th, cfg, w = perturbed_state(config, dims, vel=1.0, dp=1e-3)
    cfl = 0.6
    _say("advance config=%d graph=%s" % (config, graph))
    uo, to, dto, no, _ = _oracle_loop(blk, th, cfg, w, cfl, 1e30, 4)
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        t, dt, n = run.advance(cfl, nlim=4, graph=graph)
        ug = run.cons()
    finally:
        run.close()
    assert n == no == 4
    assert t == to and dt == dto, (t, to, dt, dto)
    sk, sj, si = blk.interior()
    e = relerr(ug[:, sk, sj, si], uo[:, sk, sj, si])
    _say("  relerr u %.3e" % e)
    assert e <= 1e-12


def test_advance_stops_at_tlim(ctx):
    blk, th, cfg, w = perturbed_state(2, (20, 12, 10), vel=1.0, dp=1e-3)
    cfl = 0.6
    dt0 = orl.new_block_dt(blk, th, cfg, w, cfl)
    tlim = 2.5 * dt0  # the third step is cut short
    uo, to, dto, no, dts = _oracle_loop(blk, th, cfg, w, cfl, tlim, 100)
    assert no == 3 and dts[-1] < dts[0]
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        t, dt, n = run.advance(cfl, tlim=tlim, nlim=100)
        # a second call past tlim runs no cycle
        t2, dt2, n2 = run.advance(cfl, tlim=tlim, nlim=100, time=t, dt=dt,
                                  ncycle=n)
        ug = run.cons()
    finally:
        run.close()
    assert (n, t, dt) == (no, to, dto)
    assert (n2, t2, dt2) == (n, t, dt)
    sk, sj, si = blk.interior()
    assert relerr(ug[:, sk, sj, si], uo[:, sk, sj, si]) <= 1e-12


def test_one_graph_serves_every_dt(ctx):
    """Steps of different sizes replayed from one captured graph equal the
    directly launched steps bit for bit."""
    blk, th, cfg, w = perturbed_state(2, (20, 12, 10), vel=1.0, dp=1e-3)
    dt0 = orl.new_block_dt(blk, th, cfg, w, 0.6)
    out = []
    for graph in (False, True):
        run = ca.Run(ctx, blk, th, cfg)
        try:
            run.set_prim(w)
            for f in (1.0, 0.5, 0.8):
                run.step(dt0 * f, 1, graph=graph)
            out.append(run.cons())
        finally:
            run.close()
    np.testing.assert_array_equal(out[1], out[0])


@pytest.mark.parametrize("case", [CASES[1], CASES[2]])
def test_run_stage_by_stage_equals_steps(ctx, case):
    """cg_run_stage driven stage by stage, with the step given by value or
    read from device memory, equals cg_run_step bit for bit."""
    config, dims = case
    blk, th, cfg, w = perturbed_state(config, dims, vel=1.0, dp=1e-3)
    cfg.implicit_flag = 15 if config == 2 else 9
    dt = orl.new_block_dt(blk, th, cfg, w, 0.5)
    out = []
    ddt = ctx.array(np.array([dt]))
    try:
        for mode in ("step", "value", "device"):
            run = ca.Run(ctx, blk, th, cfg)
            try:
                run.set_prim(w)
                if mode == "step":
                    run.step(dt, 2, graph=False)
                else:
                    for _ in range(2):
                        for s in range(3):
                            if mode == "value":
                                run.stage(s, dt)
                            else:
                                run.stage(s, dt_dev=ddt.ptr)
                    ctx.synchronize()
                out.append(run.cons())
            finally:
                run.close()
    finally:
        ddt.free()
    np.testing.assert_array_equal(out[1], out[0])
    np.testing.assert_array_equal(out[2], out[0])


def test_run_stage_rejects_bad_arguments(ctx):
    blk, th, cfg, w = perturbed_state(2, (20, 12, 10), vel=1.0, dp=1e-3)
    run = ca.Run(ctx, blk, th, cfg)
    try:
        run.set_prim(w)
        for s, dt in ((3, 0.1), (-1, 0.1), (0, 0.0), (0, -1.0)):
            with pytest.raises(ca.CanoeGpuError):
                run.stage(s, dt)
    finally:
        run.close()
