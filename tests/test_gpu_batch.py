"""GPU tests of batched runs: several independent MeshBlocks of uneven
column heights and footprints advanced together, every column of every
block in one batched launch per implicit sweep.  Each block must reproduce
its own single-block oracle run bit for bit (FP64 form), for every launch
shape the autotuner may pick."""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import cfl_dt, relerr

pytestmark = pytest.mark.gpu

# heights and footprints that leave partial 64-column groups
SPECS = [(40, 8, 8), (64, 12, 6), (24, 10, 10), (64, 8, 8), (32, 16, 4)]


def _say(msg):
    print(msg, flush=True)


def _blocks(seed=100):
    out = []
    for b, (nx1, nx2, nx3) in enumerate(SPECS):
        blk, th, cfg = ca.pgen_setup(5, nx1, nx2, nx3)
        w = ca.pgen_init(5, seed + b, blk, th, cfg)
        out.append((blk, th, cfg, w))
    return out


def _run_batch(ctx, bl, dt, nsteps, prec=ca.PREC_FP64, tune=None, graph=True):
    th, cfg = bl[0][1], bl[0][2]
    cfg.precision = prec
    run = ca.Run(ctx, None, th, cfg, blocks=[b[0] for b in bl])
    try:
        if tune is not None:
            run.set_tune(*tune)
        for i, b in enumerate(bl):
            run.set_prim(b[3], i)
        run.step(dt, nsteps, graph=graph)
        return [run.cons(i) for i in range(len(bl))]
    finally:
        run.close()


def _oracle(bl, dt, nsteps):
    out = []
    for blk, th, cfg, w in bl:
        cfg.precision = ca.PREC_FP64
        u = orl.prim_to_cons(blk, th, w)
        uo, _ = orl.step(blk, th, cfg, dt, nsteps, u, w)
        out.append(uo)
    return out


@pytest.mark.parametrize("tune", [None, (1, 3, 64), (2, 0, 128), (4, 1, 256),
                                  (1, 3, 64, 3, 16, 0), (4, 0, 256, 3, 16, 0)])
def test_batch_fp64_matches_oracle_per_block(ctx, tune):
    bl = _blocks()
    dt = min(cfl_dt(b[0], b[1], b[2], b[3], 0.6) for b in bl)
    nsteps = 3
    _say("batch fp64 tune=%s dt=%.4f" % (tune, dt))
    ug = _run_batch(ctx, bl, dt, nsteps, tune=tune)
    uo = _oracle(bl, dt, nsteps)
    for i, (blk, th, cfg, w) in enumerate(bl):
        sk, sj, si = blk.interior()
        np.testing.assert_array_equal(ug[i][:, sk, sj, si], uo[i][:, sk, sj, si])


def test_batch_equals_single_block_runs(ctx):
    bl = _blocks(seed=7)
    dt = min(cfl_dt(b[0], b[1], b[2], b[3], 0.6) for b in bl)
    ub = _run_batch(ctx, bl, dt, 2)
    for i, b in enumerate(bl):
        us = _run_batch(ctx, [b], dt, 2)[0]
        np.testing.assert_array_equal(ub[i], us)


def _moist_kessler_blocks(seed=300):
    """Moist blocks of uneven heights with Kessler rates, cloud and rain in
    the lower half of each column."""
    rates = dict(k_auto=0.01, k_accr=0.02, k_evap=0.1, v_fall=8.0)
    out = []
    for b, (nx1, nx2, nx3) in enumerate(SPECS[:3]):
        blk, th, cfg = ca.pgen_setup(4, nx1, nx2, nx3)
        for k, v in rates.items():
            setattr(cfg, k, v)
        w = ca.pgen_init(4, seed + b, blk, th, cfg)
        sk, sj, si = blk.interior()
        low = slice(si.start, si.start + nx1 // 2)
        w[2][sk, sj, low] += 1.0e-3
        w[3][sk, sj, low] += 5.0e-4
        out.append((blk, th, cfg, ca.apply_bc_host(blk, w)))
    return out


@pytest.mark.parametrize("prec", [ca.PREC_FP64, ca.PREC_FP32])
def test_moist_kessler_batch_equals_single_block_runs(ctx, prec):
    """Moist blocks with Kessler microphysics, which the last stage's
    batched column sweep applies: each block of the batch equals its own
    single-block run bit for bit."""
    bl = _moist_kessler_blocks()
    dt = min(cfl_dt(b[0], b[1], b[2], b[3], 0.6) for b in bl)
    _say("moist kessler batch prec=%d dt=%.4f" % (prec, dt))
    ub = _run_batch(ctx, bl, dt, 2, prec=prec)
    for i, b in enumerate(bl):
        us = _run_batch(ctx, [b], dt, 2, prec=prec)[0]
        np.testing.assert_array_equal(ub[i], us)
        sk, sj, si = b[0].interior()
        assert np.all(np.isfinite(us[:, sk, sj, si]))


@pytest.mark.parametrize("variant", [(1, 256, 0, 16), (1, 256, 1, 16),
                                     (1, 128, 1, 3), (1, 64, 1, 40),
                                     (1, 256, 2, 5), (1, 256, 3, 7),
                                     (1, 256, 3, 64), (1, 256, 4, 9)])
def test_batch_mixed_close_to_fp64(ctx, variant):
    """Both face-kernel variants of the FP32-flux form, several march
    segments (shorter than, and longer than, the columns)."""
    bl = _blocks(seed=11)
    dt = min(cfl_dt(b[0], b[1], b[2], b[3], 0.6) for b in bl)
    u64 = _run_batch(ctx, bl, dt, 5)
    wpb, fb, fv, seg = variant
    u32 = _run_batch(ctx, bl, dt, 5, prec=ca.PREC_FP32,
                     tune=(wpb, 0, fb, fv, seg))
    for i, b in enumerate(bl):
        sk, sj, si = b[0].interior()
        e = relerr(u32[i][:, sk, sj, si], u64[i][:, sk, sj, si])
        _say("  block %d mixed %s vs fp64 relerr %.3e" % (i, variant, e))
        assert e < 1e-5


def test_autotune_keeps_state_and_picks_a_valid_shape(ctx):
    bl = _blocks(seed=13)
    th, cfg = bl[0][1], bl[0][2]
    cfg.precision = ca.PREC_FP64
    dt = min(cfl_dt(b[0], b[1], b[2], b[3], 0.6) for b in bl)
    run = ca.Run(ctx, None, th, cfg, blocks=[b[0] for b in bl])
    try:
        for i, b in enumerate(bl):
            run.set_prim(b[3], i)
        u0 = [run.cons(i) for i in range(len(bl))]
        t = run.autotune(dt)
        _say("  autotune: %s" % t.as_dict())
        assert t.implicit_wpb in (1, 2, 4) and t.flux_block in (64, 128, 256)
        assert t.face_variant in range(5) and t.march_seg >= 1
        assert t.implicit_fused in (0, 1)
        assert t.step_ms > 0
        for i in range(len(bl)):
            np.testing.assert_array_equal(run.cons(i), u0[i])
        run.step(dt, 2)
        ua = [run.cons(i) for i in range(len(bl))]
    finally:
        run.close()
    ub = _run_batch(ctx, bl, dt, 2)
    for i in range(len(bl)):
        np.testing.assert_array_equal(ua[i], ub[i])


@pytest.mark.parametrize("variant", [(3, 16), (4, 9)])
def test_batch_tiles_equal_single_block_runs_mixed(ctx, variant):
    """The FP32-flux form's tile kernel runs over one table of tiles of
    every block: a batch reproduces each block's own run bit for bit."""
    bl = _blocks(seed=17)
    dt = min(cfl_dt(b[0], b[1], b[2], b[3], 0.6) for b in bl)
    tune = (1, 0, 256) + variant
    _say("batch tiles mixed variant=%s" % (variant,))
    ub = _run_batch(ctx, bl, dt, 2, prec=ca.PREC_FP32, tune=tune)
    for i, b in enumerate(bl):
        us = _run_batch(ctx, [b], dt, 2, prec=ca.PREC_FP32, tune=tune)[0]
        np.testing.assert_array_equal(ub[i], us)
