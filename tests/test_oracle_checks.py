"""Pins of the oracle's check_reconstruct (src/checks.cpp:79-97, called after
the face states of every direction, patches/19.calculate_fluxes.patch:45,
73, 117): on a block without gravity (no decomposition) and PLM
reconstruction, a cell whose density or pressure is negative in a uniform
state is a local minimum, so van Leer's limited slope is zero there and both
faces of that cell in each direction carry the negative value while every
other face stays positive.  One such cell gives exactly 6 failing faces; two
neighbours along x1 share a face, which counts once: 3 + 4 + 4 = 11."""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl


def _uniform():
    blk, th, cfg = ca.pgen_setup(2, 16, 12, 10)
    blk.grav = 0.0
    cfg.order_h = cfg.order_v = 2
    w = np.zeros(blk.shape())
    w[0] = 1.2
    w[4] = 1e5
    w[1] = 0.3
    w[2] = -0.2
    w[3] = 0.1
    return blk, th, cfg, w


def test_positive_state_has_no_failing_face():
    blk, th, cfg, w = _uniform()
    rng = np.random.default_rng(3)
    w[0] *= 1 + 0.1 * rng.uniform(-1, 1, w[0].shape)
    w[4] *= 1 + 0.1 * rng.uniform(-1, 1, w[4].shape)
    orl.calculate_fluxes(blk, th, cfg, w)
    assert orl.face_failures() == 0


@pytest.mark.parametrize("var", [0, 4])
def test_one_negative_cell_fails_its_six_faces(var):
    blk, th, cfg, w = _uniform()
    w[var, 6, 7, 9] = -1.0
    orl.calculate_fluxes(blk, th, cfg, w)
    assert orl.face_failures() == 6


def test_a_shared_face_counts_once():
    blk, th, cfg, w = _uniform()
    w[4, 6, 7, 9] = -1.0
    w[4, 6, 7, 10] = -2.0
    orl.calculate_fluxes(blk, th, cfg, w)
    assert orl.face_failures() == 11
