"""Pins of the oracle's output reductions: Athena++'s default history sums
(upstream outputs/history.cpp) and DebugOutput's per-variable extremes
(src/outputs/debug.cpp:30-84).

The history sums of a uniform state are closed forms (total volume times the
cell value); the extremes are checked on fields with planted values, ties
(the first cell in k, j, i order wins), an all-negative field (the maximum
stays at the reference's initial FLT_MIN, with its FLT_MIN coordinates) and
NaN cells (never selected by the strict comparisons)."""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl

FLT_MAX = float(np.finfo(np.float32).max)
FLT_MIN = float(np.finfo(np.float32).tiny)


def _block(config=4, dims=(8, 6, 5)):
    blk, th, cfg = ca.pgen_setup(config, *dims)
    blk.x1min, blk.x2min, blk.x3min = 100.0, -300.0, 50.0
    return blk, th, cfg


def test_history_uniform_state_closed_form():
    blk, th, cfg = _block()
    u = np.zeros(blk.shape())
    ivx = 1 + blk.nmass
    vals = {0: 1.1, 1: 0.02, 2: 0.003, 3: 0.001, ivx: 2.0, ivx + 1: -3.0,
            ivx + 2: 0.5, ivx + 3: 2.5e5}
    for n, v in vals.items():
        u[n] = v
    V = blk.nx1 * blk.nx2 * blk.nx3 * blk.dx1 * blk.dx2 * blk.dx3
    h = orl.history(blk, u)
    want = [1.1 * V, 2.0 * V, -3.0 * V, 0.5 * V, 0.5 * 4.0 / 1.1 * V,
            0.5 * 9.0 / 1.1 * V, 0.5 * 0.25 / 1.1 * V, 2.5e5 * V,
            0.02 * V, 0.003 * V, 0.001 * V]
    np.testing.assert_allclose(h, want, rtol=1e-13)


def test_history_ignores_ghosts():
    blk, th, cfg = _block(2)
    u = np.zeros(blk.shape())
    u[0] = 1.0
    u[4] = 1.0
    sk, sj, si = blk.interior()
    g = u.copy()
    g[:, :, :, 0] = 1e9  # x1 ghosts only
    np.testing.assert_array_equal(orl.history(blk, g), orl.history(blk, u))


def test_debug_planted_extremes_and_coordinates():
    blk, th, cfg = _block(2)
    w = np.ones(blk.shape())
    # variable 3: minimum at interior (i, j, k) = (2, 4, 1), maximum at (7, 0, 3)
    w[3, 3 + 1, 3 + 4, 3 + 2] = -5.0
    w[3, 3 + 3, 3 + 0, 3 + 7] = 9.0
    vmin, vmax, xmin, xmax = orl.debug(blk, w)
    assert vmin[3] == -5.0 and vmax[3] == 9.0
    np.testing.assert_array_equal(
        xmin[3], [100.0 + 2.5 * blk.dx1, -300.0 + 4.5 * blk.dx2, 50.0 + 1.5 * blk.dx3])
    np.testing.assert_array_equal(
        xmax[3], [100.0 + 7.5 * blk.dx1, -300.0 + 0.5 * blk.dx2, 50.0 + 3.5 * blk.dx3])
    # a uniform variable: both extremes at the first interior cell
    np.testing.assert_array_equal(xmin[0], [100.0 + 0.5 * blk.dx1, -300.0 + 0.5 * blk.dx2,
                                            50.0 + 0.5 * blk.dx3])
    np.testing.assert_array_equal(xmax[0], xmin[0])


def test_debug_ties_first_in_kji_order():
    blk, th, cfg = _block(2)
    w = np.zeros(blk.shape())
    w[1, 3 + 2, 3 + 0, 3 + 5] = 7.0   # k = 2
    w[1, 3 + 1, 3 + 5, 3 + 0] = 7.0   # k = 1: first in (k, j, i) order
    _, vmax, _, xmax = orl.debug(blk, w)
    assert vmax[1] == 7.0
    assert xmax[1][2] == 50.0 + 1.5 * blk.dx3


def test_debug_reference_initial_values_and_nan():
    blk, th, cfg = _block(2)
    w = -np.ones(blk.shape())
    w[2, 3, 3, 3] = np.nan
    vmin, vmax, xmin, xmax = orl.debug(blk, w)
    # every value is below FLT_MIN: the maximum keeps its initial value
    assert vmax[0] == FLT_MIN and np.all(xmax[0] == FLT_MIN)
    assert vmin[2] == -1.0 and not np.isnan(vmin[2])
    assert xmin[2][0] == 100.0 + 1.5 * blk.dx1  # the NaN cell (first) skipped


def _rest_state():
    blk, th, cfg = ca.pgen_setup(4, 8, 6, 5)
    w = ca.pgen_init(4, 3, blk, th, cfg)
    return blk, th, cfg, w


def test_check_passes_a_physical_state():
    blk, th, cfg, w = _rest_state()
    assert orl.check(blk, th, w) == (3, (-1, -1, -1))


@pytest.mark.parametrize("n,value,code", [(0, -1e-3, 1), (2, -1e-9, 1),
                                          (7, -5.0, 1), (4, np.nan, 0),
                                          (1, np.nan, 0)])
def test_check_finds_the_first_bad_cell(n, value, code):
    # two bad cells: the one first in (k, j, i) order is reported; a
    # negative velocity is not an error
    blk, th, cfg, w = _rest_state()
    w[4, 3 + 2, 3 + 1, 3 + 4] = -30.0  # v2 < 0: fine
    w[n, 3 + 3, 3 + 0, 3 + 1] = value
    w[n, 3 + 1, 3 + 5, 3 + 6] = value
    assert orl.check(blk, th, w) == (code, (1, 5, 6))


def test_check_temperature_floor():
    # T below |g| dx1 / Rd (the reference's minimum temperature)
    blk, th, cfg, w = _rest_state()
    tmin = abs(blk.grav) * blk.dx1 / th.rd
    rho = w[0, 3 + 2, 3 + 2, 3 + 2]
    w[7, 3 + 2, 3 + 2, 3 + 2] = 0.5 * tmin * rho * th.rd  # dry-ish cell, T < tmin
    w[1:4, 3 + 2, 3 + 2, 3 + 2] = 0.0
    assert orl.check(blk, th, w) == (2, (2, 2, 2))
    w[7, 3 + 2, 3 + 2, 3 + 2] = 2.0 * tmin * rho * th.rd
    assert orl.check(blk, th, w)[0] == 3
