"""Pins of the oracle's implicit lines beyond the non-periodic x1 columns
(SolveImplicit3D's x2 / x3 lines and PeriodicForwardSweep /
PeriodicBackwardSubstitution, src/snap/implicit/solve_implicit_3d.cpp:19-143,
periodic_forward_backward.hpp:14-232).

* The cyclic block recurrences solve the dense cyclic block system built and
  solved by numpy.
* An x2 (x3) line solve equals the x1 column solve of the block transposed so
  that x2 (x3) becomes x1, with its velocity components carried along and no
  gravity (the x2 / x3 matrices have no Phi), for walls and periodic ends.
  A wrong stride, frame rotation, face area or missing fold fails it.
* On a periodic line without gravity the block columns of the system sum to
  I / dt (the fluxes telescope around the ring), so the line sums of the new
  du equal those of the old one for every variable solved (to round-off of
  the solve, 1e-10 of the summed magnitudes; a wrong corner block or
  recurrence misses by O(1)).  With gravity (x1) the mass and tangential
  momentum rows still telescope.
* x2 / x3 lines change only IDN, the normal momentum and IEN; the directions
  run in the order x3, x2, x1.
* Walled x2 / x3 lines of a moist block weight the species by w at the
  shuffled index of the reference's line buffer: a change of w_n at that
  entry (an x1 ghost that nothing else reads) moves exactly one cell's
  density increment, by the amount the write-back's algebra predicts; a
  block where the entry runs past the end of w is rejected.
"""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl
from helpers import perturbed_state
from test_oracle_pins import _rand_blocks


def _set_bc(blk, bc):
    for i, v in enumerate(bc):
        blk.bc[i] = v


@pytest.mark.parametrize("B", [3, 5])
@pytest.mark.parametrize("n", [3, 4, 7, 40])
def test_periodic_block_thomas_vs_dense(B, n):
    rng = np.random.default_rng(B * 31 + n)
    a, b, c = _rand_blocks(B, n, rng)
    corner = _rand_blocks(B, 1, rng)[2][0]
    r = rng.normal(size=(n, B))
    x = np.zeros((n, B))
    rc = orl.lib().or_block_thomas_periodic(
        B, n, *[orl.ptr(np.ascontiguousarray(v)) for v in (a, b, c, corner, r)],
        orl.ptr(x))
    assert rc == 0
    A = np.zeros((n * B, n * B))
    for i in range(n):
        rows = slice(i * B, (i + 1) * B)
        A[rows, i * B:(i + 1) * B] += a[i]
        im, ip = (i - 1) % n, (i + 1) % n
        A[rows, im * B:(im + 1) * B] += b[i]
        A[rows, ip * B:(ip + 1) * B] += c[i] if i < n - 1 else corner
    xd = np.linalg.solve(A, r.reshape(-1)).reshape(n, B)
    assert np.max(np.abs(x - xd)) <= 1e-12 * np.abs(xd).max()


def test_periodic_block_thomas_needs_three_levels():
    z = np.zeros(50)
    assert orl.lib().or_block_thomas_periodic(5, 2, *[orl.ptr(z)] * 7) == ca.ERR_ARG


def _case(bc_x2, bc_x3, dims=(6, 10, 8), seed=11):
    blk, th, cfg, w = perturbed_state(2, dims, seed=seed, vel=30.0, dp=2e-3)
    blk.dx1, blk.dx2, blk.dx3 = 90.0, 230.0, 170.0
    bc = list(blk.bc)
    bc[2], bc[3], bc[4], bc[5] = bc_x2, bc_x2, bc_x3, bc_x3
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    rng = np.random.default_rng(seed)
    du = np.zeros(blk.shape())
    sk, sj, si = blk.interior()
    du[:, sk, sj, si] = rng.normal(size=du[:, sk, sj, si].shape) * \
        np.array([1e-3, 1e-2, 1e-2, 1e-2, 3e2])[:, None, None, None]
    return blk, th, w, du


def _transpose(blk, arr, dir):
    """The block and a cell array with direction dir moved to x1."""
    t = ca.Block()
    for f, _ in ca.Block._fields_:
        setattr(t, f, getattr(blk, f))
    nx = [blk.nx1, blk.nx2, blk.nx3]
    dx = [blk.dx1, blk.dx2, blk.dx3]
    bc = list(blk.bc)
    if dir == 1:  # axes (n, k, j, i) -> (n, k, i, j); x1 <-> x2
        perm, axes = [1, 0, 2], (0, 1, 3, 2)
    else:         # axes (n, k, j, i) -> (n, i, j, k); x1 <-> x3
        perm, axes = [2, 1, 0], (0, 3, 2, 1)
    t.nx1, t.nx2, t.nx3 = [nx[p] for p in perm]
    t.dx1, t.dx2, t.dx3 = [dx[p] for p in perm]
    _set_bc(t, sum(([bc[2 * p], bc[2 * p + 1]] for p in perm), []))
    t.grav = 0.0
    out = np.ascontiguousarray(arr.transpose(axes))
    ivx = 1 + blk.nmass
    out[ivx:ivx + 3] = out[[ivx + p for p in perm]]
    return t, out


def _back(blk, arr, dir):
    ivx = 1 + blk.nmass
    perm = [1, 0, 2] if dir == 1 else [2, 1, 0]
    a = arr.copy()
    a[[ivx + p for p in perm]] = arr[ivx:ivx + 3]
    axes = (0, 1, 3, 2) if dir == 1 else (0, 3, 2, 1)
    return a.transpose(axes)


@pytest.mark.parametrize("dir", [1, 2])
@pytest.mark.parametrize("bc", [ca.BC_REFLECT, ca.BC_PERIODIC])
@pytest.mark.parametrize("full", [True, False])
def test_line_equals_x1_column_of_transposed_block(dir, bc, full):
    blk, th, w, du = _case(bc, bc)
    flag = (1 << dir) | (8 if full else 0)
    got, r = orl.implicit_solve(blk, th, du, w, 0.7, flag)
    assert r == 0
    tb, tw = _transpose(blk, w, dir)
    _, tdu = _transpose(blk, du, dir)
    ref, r = orl.implicit_solve(tb, th, tdu, tw, 0.7, 9 if full else 1)
    assert r == 0
    ref = _back(blk, ref, dir)
    sk, sj, si = blk.interior()
    ivx = 1 + blk.nmass
    for n in (0, ivx + dir, ivx + 3):
        g, e = got[n][sk, sj, si], ref[n][sk, sj, si]
        assert np.max(np.abs(g - e)) <= 1e-11 * np.abs(e).max(), n
    # the line changed the solved variables and nothing else
    for n in range(blk.nmass + 5):
        moved = not np.array_equal(got[n], du[n])
        assert moved == (n in (0, ivx + dir, ivx + 3)), n


@pytest.mark.parametrize("dir", [1, 2])
@pytest.mark.parametrize("full", [True, False])
def test_periodic_line_sums_are_conserved(dir, full):
    blk, th, w, du = _case(ca.BC_PERIODIC, ca.BC_PERIODIC)
    flag = (1 << dir) | (8 if full else 0)
    got, r = orl.implicit_solve(blk, th, du, w, 0.7, flag)
    assert r == 0
    sk, sj, si = blk.interior()
    ivx = 1 + blk.nmass
    ax = 1 if dir == 1 else 0  # line axis of the (k, j, i) interior array
    for n in (0, ivx + dir, ivx + 3):
        s0 = du[n][sk, sj, si].sum(axis=ax)
        s1 = got[n][sk, sj, si].sum(axis=ax)
        scale = np.abs(du[n][sk, sj, si]).sum(axis=ax)
        assert np.max(np.abs(s1 - s0) / scale) < 1e-10, n
        # and the solve did something
        assert np.max(np.abs(got[n] - du[n])) > 1e-6 * np.abs(du[n]).max()


def test_periodic_x1_conserves_column_mass():
    blk, th, w, du = _case(ca.BC_PERIODIC, ca.BC_PERIODIC)
    bc = list(blk.bc)
    bc[0] = bc[1] = ca.BC_PERIODIC
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    got, r = orl.implicit_solve(blk, th, du, w, 0.7, 9)
    assert r == 0
    sk, sj, si = blk.interior()
    s0 = du[0][sk, sj, si].sum(axis=2)
    s1 = got[0][sk, sj, si].sum(axis=2)
    assert np.max(np.abs(s1 - s0) / np.abs(du[0][sk, sj, si]).sum(axis=2)) < 1e-10
    # so are the tangential momenta (gravity acts on the normal one and E)
    for n in (2, 3):
        s0 = du[n][sk, sj, si].sum(axis=2)
        s1 = got[n][sk, sj, si].sum(axis=2)
        assert np.max(np.abs(s1 - s0) / np.abs(du[n][sk, sj, si]).sum(axis=2)) < 1e-10
    # the periodic column differs from the walled one
    bc[0] = bc[1] = ca.BC_REFLECT
    _set_bc(blk, bc)
    walled, _ = orl.implicit_solve(blk, th, du, w, 0.7, 9)
    assert np.max(np.abs(walled[0] - got[0])) > 1e-8 * np.abs(got[0]).max()


def test_directions_run_x3_then_x2_then_x1():
    blk, th, w, du = _case(ca.BC_PERIODIC, ca.BC_REFLECT)
    all3, r = orl.implicit_solve(blk, th, du, w, 0.7, 15)
    assert r == 0
    seq = du
    for f in (12, 10, 9):
        seq, r = orl.implicit_solve(blk, th, seq, w, 0.7, f)
        assert r == 0
    np.testing.assert_array_equal(all3, seq)
    other = du
    for f in (9, 10, 12):
        other, _ = orl.implicit_solve(blk, th, other, w, 0.7, f)
    assert not np.array_equal(all3, other)


def _moist_walls(dims, seed=2):
    blk, th, cfg, w = perturbed_state(4, dims, vel=10.0)
    bc = list(blk.bc)
    bc[2] = bc[3] = ca.BC_REFLECT
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    du = np.zeros(blk.shape())
    sk, sj, si = blk.interior()
    du[:, sk, sj, si] = np.random.default_rng(seed).normal(
        size=du[:, sk, sj, si].shape) * 1e-3
    return blk, th, w, du


def test_walled_species_lines_read_w_at_the_shuffled_index():
    """BackwardSubstitution splits the density increment over the species
    with w(n, k, j, i) indexed by the shuffled line buffer's loop indices
    (forward_backward.hpp:109-115, solve_implicit_3d.cpp:59-62): for the x2
    line cell (x1, x2, x3) = (5, 9, 6) that is the entry w[n, 5, 6, 9]
    (x3 = 5, x2 = 6, x1 = 9), an x1 ghost no x2 line reads otherwise.
    Raising w_n there by D changes only that cell's density increment, by
    -dens D, with dens = (du_IDN' - du_IDN) / (1 - sum_n w_n(entry)) from the
    write-back's own algebra (du_IDN' = dens + du_IDN - dens sum_n w_n)."""
    blk, th, w, du = _moist_walls((6, 8, 7))
    d0, r = orl.implicit_solve(blk, th, du, w, 0.5, 10)
    assert r == 0
    cell = (0, 6, 9, 5)
    wT = np.array([w[n, 5, 6, 9] for n in range(1, 4)])
    dens = (d0[cell] - du[cell]) / (1.0 - wT.sum())
    for n, D in ((1, 1e-3), (3, -2e-3)):
        w1 = w.copy()
        w1[n, 5, 6, 9] += D
        d1, r = orl.implicit_solve(blk, th, du, w1, 0.5, 10)
        assert r == 0
        diff = d1 - d0
        changed = np.argwhere(diff != 0)
        assert [tuple(c) for c in changed] == [cell]
        np.testing.assert_allclose(diff[cell], -dens * D, rtol=1e-9)


def test_walled_species_lines_past_the_end_of_w_rejected():
    # 2D x1-x2 block: the shuffled index of the x2 lines' species weights
    # runs past the end of w (the reference reads out of bounds there)
    blk, th, cfg, w = perturbed_state(4, (8, 10, 1), vel=10.0)
    bc = list(blk.bc)
    bc[2] = bc[3] = ca.BC_REFLECT
    _set_bc(blk, bc)
    w = ca.apply_bc_host(blk, w)
    du = np.zeros(blk.shape())
    _, r = orl.implicit_solve(blk, th, du, w, 0.5, 10)
    assert r == ca.ERR_ARG
    bc[2] = bc[3] = ca.BC_PERIODIC  # periodic lines weight no species
    _set_bc(blk, bc)
    _, r = orl.implicit_solve(blk, th, du, ca.apply_bc_host(blk, w), 0.5, 10)
    assert r == 0
