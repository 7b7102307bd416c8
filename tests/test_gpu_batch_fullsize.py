"""The uneven-column batch at its full workload: 32 blocks with seeded
heights from {40, 64, 96, 128, 200, 256} levels and footprints from 32x32
to 128x128 (24.7 M cells, 189,440 columns), advanced in one batched run.
Each block equals its own single-block GPU run bit for bit (the batched
column launch, the tile table of the FP32-flux form and the ghost kernels
treat every block alone), and each block's total mass is conserved to
1e-13 (reflecting x1 walls, periodic x2 / x3: the fluxes telescope)."""
import numpy as np
import pytest

import canoe_amd as ca
from helpers import cfl_dt

pytestmark = pytest.mark.gpu

SEED = 2024
NSTEPS = 2


def _say(msg):
    print(msg, flush=True)


def _run(ctx, blocks, dt, prec):
    th, cfg = blocks[0][1], blocks[0][2]
    cfg.precision = prec
    run = ca.Run(ctx, None, th, cfg, blocks=[b[0] for b in blocks])
    try:
        for i, b in enumerate(blocks):
            run.set_prim(b[3], i)
        m0 = [run.sums(i)[0] for i in range(len(blocks))]
        run.step(dt, NSTEPS, graph=True)
        m1 = [run.sums(i)[0] for i in range(len(blocks))]
        return [run.cons(i) for i in range(len(blocks))], m0, m1
    finally:
        run.close()


@pytest.mark.parametrize("prec", [ca.PREC_FP64, ca.PREC_FP32])
def test_full_batch_equals_single_block_runs(ctx, prec):
    bl = ca.batch_blocks(SEED, 32)
    cells = sum(b[0].nx1 * b[0].nx2 * b[0].nx3 for b in bl)
    assert max(b[0].nx1 for b in bl) == 256 and max(b[0].nx2 for b in bl) == 128
    dt = min(cfl_dt(b[0], b[1], b[2], b[3], 0.6) for b in bl)
    _say("full batch prec=%d: %d blocks, %d cells, dt %.4f" %
         (prec, len(bl), cells, dt))
    ub, m0, m1 = _run(ctx, bl, dt, prec)
    for i, b in enumerate(bl):
        assert abs(m1[i] - m0[i]) <= 1e-13 * abs(m0[i]), (i, m0[i], m1[i])
    _say("  mass conserved per block; single-block runs follow")
    for i, b in enumerate(bl):
        us, _, _ = _run(ctx, [b], dt, prec)
        np.testing.assert_array_equal(ub[i], us[0])
        del us
    _say("  every block equals its own run")
