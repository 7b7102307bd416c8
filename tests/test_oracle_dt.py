"""Pins of the oracle's time step (Hydro::NewBlockTimeStep as patched by
patches/21.new_blockdt.patch:16-45, and Mesh::NewTimeStep's limits).

Each case fixes the answer by something other than the oracle's own
formula: a uniform state reduces the step to the textbook CFL bound of one
direction, chosen by making every direction's bound distinct (a transposed
dx / velocity pairing, a sum instead of a minimum, or a missing cfl factor
all change the number); a single fast cell placed anywhere must set the
step, while the same speed in a ghost cell must not; the moist sound speed
follows from gamma = 1 + R_mix / cv_mix of the mixture.
"""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl


def _uniform(config, dims, rho=1.2, p=9.0e4, vel=(3.0, -5.0, 7.0)):
    blk, th, cfg = ca.pgen_setup(config, *dims)
    blk.dx1, blk.dx2, blk.dx3 = 50.0, 400.0, 250.0
    w = np.zeros(blk.shape())
    ivx = 1 + blk.nmass
    w[0] = rho
    w[ivx + 3] = p
    for d in range(3):
        w[ivx + d] = vel[d]
    return blk, th, cfg, w


def _cs_dry(th, rho, p):
    return np.sqrt(th.gammad * p / rho)


@pytest.mark.parametrize("cfl", [0.9, 0.3])
def test_explicit_uniform_state_is_min_over_directions(cfl):
    blk, th, cfg, w = _uniform(1, (8, 6, 1))
    cfg.implicit_flag = 0
    # config 1 is two-dimensional (x1, x2): x3 does not enter
    cs = _cs_dry(th, 1.2, 9.0e4)
    want = cfl * min(50.0 / (3.0 + cs), 400.0 / (5.0 + cs))
    got = orl.new_block_dt(blk, th, cfg, w, cfl)
    assert got == pytest.approx(want, rel=1e-15)


def test_explicit_3d_uniform_state():
    blk, th, cfg, w = _uniform(2, (8, 6, 5))
    cfg.implicit_flag = 0
    cs = _cs_dry(th, 1.2, 9.0e4)
    want = 0.5 * min(50.0 / (3.0 + cs), 400.0 / (5.0 + cs), 250.0 / (7.0 + cs))
    assert orl.new_block_dt(blk, th, cfg, w, 0.5) == pytest.approx(want, rel=1e-15)


def test_implicit_x1_drops_sound_speed_and_cfl():
    # x1 solved implicitly (flag 9, nx2 > 1): its bound is dx1 / |v1| with no
    # cfl factor; the horizontal bounds keep c_s and the factor.
    blk, th, cfg, w = _uniform(2, (8, 6, 5), vel=(400.0, 0.0, 0.0))
    assert cfg.implicit_flag == 9
    cs = _cs_dry(th, 1.2, 9.0e4)
    cfl = 0.8
    want = min(50.0 / 400.0, cfl * 400.0 / cs, cfl * 250.0 / cs)
    got = orl.new_block_dt(blk, th, cfg, w, cfl)
    assert got == pytest.approx(want, rel=1e-15)
    assert got == pytest.approx(50.0 / 400.0, rel=1e-15)  # x1 sets it here
    # at rest the implicit direction sets no bound at all
    w[1 + blk.nmass] = 0.0
    got0 = orl.new_block_dt(blk, th, cfg, w, cfl)
    assert got0 == pytest.approx(cfl * 250.0 / cs, rel=1e-15)


def test_implicit_flag_ignored_in_a_single_column():
    # the patch applies the implicit form to x1 only when nx2 > 1
    blk, th, cfg, w = _uniform(2, (8, 1, 1), vel=(40.0, 0.0, 0.0))
    blk.nx2 = blk.nx3 = 1
    w = np.zeros(blk.shape())
    ivx = 1 + blk.nmass
    w[0], w[ivx], w[ivx + 3] = 1.2, 40.0, 9.0e4
    cs = _cs_dry(th, 1.2, 9.0e4)
    got = orl.new_block_dt(blk, th, cfg, w, 0.8)
    assert got == pytest.approx(0.8 * 50.0 / (40.0 + cs), rel=1e-15)


@pytest.mark.parametrize("cell", [(3, 3, 3), (10, 8, 7), (6, 4, 5)])
def test_fastest_interior_cell_sets_the_step(cell):
    blk, th, cfg, w = _uniform(2, (8, 6, 5), vel=(0.0, 0.0, 0.0))
    cfg.implicit_flag = 0
    ivx = 1 + blk.nmass
    i, j, k = cell
    w[ivx + 1, k, j, i] = 300.0  # x2 velocity in one cell
    cs = _cs_dry(th, 1.2, 9.0e4)
    want = 0.9 * min(50.0 / cs, 400.0 / (300.0 + cs))
    assert orl.new_block_dt(blk, th, cfg, w, 0.9) == pytest.approx(want, rel=1e-15)


def test_ghost_cells_do_not_enter():
    blk, th, cfg, w = _uniform(2, (8, 6, 5), vel=(0.0, 0.0, 0.0))
    cfg.implicit_flag = 0
    ivx = 1 + blk.nmass
    w[ivx, 3, 3, 0] = 1e6  # x1 ghost
    w[ivx + 1, 0, 4, 5] = 1e6  # x3 ghost
    cs = _cs_dry(th, 1.2, 9.0e4)
    assert orl.new_block_dt(blk, th, cfg, w, 0.9) == pytest.approx(
        0.9 * 50.0 / cs, rel=1e-15)


def test_moist_sound_speed_from_mixture_heat_capacities():
    blk, th, cfg, w = _uniform(4, (8, 6, 5), vel=(0.0, 0.0, 0.0))
    cfg.implicit_flag = 0
    qv, qc, qp = 0.012, 0.002, 0.001
    w[1], w[2], w[3] = qv, qc, qp
    # mixture: cv = sum q_i cv_i, R = sum over gases q_i R_i (condensates
    # carry no pressure); gamma = 1 + R / cv
    rd = th.rd
    cvd = rd / (th.gammad - 1.0)
    cv = (1 - qv - qc - qp) * cvd + sum(
        q * cvd * (1 + th.cv_ratio_m1[n]) for n, q in enumerate((qv, qc, qp)))
    R = (1 - qv - qc - qp) * rd + qv * rd * (1 + th.inv_mu_ratio_m1[0])
    gamma = 1.0 + R / cv
    cs = np.sqrt(gamma * 9.0e4 / 1.2)
    assert abs(gamma - th.gammad) > 1e-3  # the moist value differs
    want = 0.9 * 50.0 / cs
    assert orl.new_block_dt(blk, th, cfg, w, 0.9) == pytest.approx(want, rel=1e-13)


def test_mesh_new_time_step_limits():
    # growth at most 2x, the blocks' minimum otherwise, cut to reach tlim
    assert orl.new_mesh_dt(1.0, 5.0, 0.0, 100.0) == 2.0
    assert orl.new_mesh_dt(1.0, 1.5, 0.0, 100.0) == 1.5
    assert orl.new_mesh_dt(1.0, 1.5, 99.25, 100.0) == 0.75
    assert orl.new_mesh_dt(1.0, 1.5, 100.0, 100.0) == 1.5  # past tlim: no cut
