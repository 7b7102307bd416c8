"""GPU tests of the saturation-guess network and the guided adjustment.

The device forward and gradient are checked against the oracle's bf16
emulation (tolerances below: the device accumulates in fp32 inside MFMA in
an order the oracle does not reproduce, so a bf16 rounding of an activation
can land on the other side of a tie), the training pairs bit for bit, and
the guided adjustment against the plain one (same tolerance, same result)
and against its iteration counts.
"""
import numpy as np
import pytest

import canoe_amd as ca
from canoe_amd import mlp as cm
import oracle_lib as orl

pytestmark = pytest.mark.gpu


def moist_states(dims=(32, 8, 8), seed=3):
    """(th, rho, ie, q) of the interior cells of the moist box: ie is the
    sensible internal energy density rho c_v T, q (3, n) mass fractions."""
    blk, th, cfg = ca.pgen_setup(4, *dims)
    w = ca.pgen_init(4, seed, blk, th, cfg)
    sk, sj, si = blk.interior()
    ivx = 1 + blk.nmass
    rho = w[0][sk, sj, si].ravel().copy()
    q = np.stack([w[n][sk, sj, si].ravel() for n in (1, 2, 3)]).copy()
    p = w[ivx + 3][sk, sj, si].ravel()
    cvd = th.rd / (th.gammad - 1)
    rv = th.rd * (1 + th.inv_mu_ratio_m1[0])
    qd = 1 - q.sum(0)
    R = th.rd * qd + rv * q[0]
    cv = cvd * qd + sum(cvd * (1 + th.cv_ratio_m1[n]) * q[n] for n in range(3))
    ie = rho * cv * p / (rho * R)
    return th, rho, ie, q


def supersaturated(th, rho, ie, q, n, seed):
    """n states drawn from the given cells with their water scaled by
    1.1..2.2 (all vapour) and the temperature shifted by -2..2 K (part of
    them stays sub-saturated)."""
    rng = np.random.default_rng(seed)
    b = rng.integers(0, rho.size, n)
    cvd = th.rd / (th.gammad - 1)
    cvv = cvd * (1 + th.cv_ratio_m1[0])
    cvp = cvd * (1 + th.cv_ratio_m1[2])
    r, qb = rho[b], q[:, b].copy()
    qd = 1 - qb.sum(0)
    cv = cvd * qd + cvv * qb[0] + cvd * (1 + th.cv_ratio_m1[1]) * qb[1] + cvp * qb[2]
    T = ie[b] / (r * cv) + rng.uniform(-2, 2, n)
    qn = qb.copy()
    qn[0] = (qb[0] + qb[1]) * rng.uniform(1.1, 2.2, n)
    qn[1] = 0.0
    qdn = 1 - qn.sum(0)
    cvn = cvd * qdn + cvv * qn[0] + cvp * qn[2]
    return r.copy(), r * cvn * T, np.ascontiguousarray(qn)


def test_params_round_trip_and_forward_matches_oracle(ctx):
    th, rho, ie, q = moist_states()
    x, t, _ = orl.mlp_pairs(th, rho, ie, q, 3000, seed=2)
    net = cm.Mlp(ctx, seed=5)
    try:
        p = net.params()
        assert p.size == cm.NPARAM and np.all(np.isfinite(p))
        p = p + np.float32(0.01) * np.random.default_rng(1).standard_normal(
            p.size).astype(np.float32)
        net.set_params(p)
        np.testing.assert_array_equal(net.params(), p)
        y = net.forward(x)
        yo = orl.mlp_forward(p, x, emulate=1)
        err = np.linalg.norm(y - yo) / np.linalg.norm(yo)
        print("forward rel L2 vs oracle %.3e, max abs %.3e"
              % (err, np.max(np.abs(y - yo))))
        assert err < 1e-3
        assert np.max(np.abs(y - yo)) < 1e-2 * max(1.0, np.max(np.abs(yo)))
    finally:
        net.close()


def test_gradient_matches_oracle(ctx):
    th, rho, ie, q = moist_states()
    x, t, _ = orl.mlp_pairs(th, rho, ie, q, 1000, seed=4)
    net = cm.Mlp(ctx, seed=6)
    try:
        p = net.params()
        g, loss = net.loss_grad(x, t)
        go, lo = orl.mlp_loss_grad(p, x, t, emulate=1)
        g0, l0 = orl.mlp_loss_grad(p, x, t, emulate=0)
        print("loss gpu %.6e oracle %.6e exact %.6e" % (loss, lo, l0))
        assert abs(loss - lo) <= 1e-4 * lo
        o = 0
        for din, dout in ((4, 64), (64, 64), (64, 64), (64, 3)):
            for n in (din * dout, dout):
                a, b = g[o:o + n], go[o:o + n]
                e = np.linalg.norm(a - b) / np.linalg.norm(b)
                print("  block %d..%d rel L2 %.3e" % (o, o + n, e))
                assert e < 1e-2
                o += n
    finally:
        net.close()


def test_pairs_bit_exact(ctx):
    th, rho, ie, q = moist_states()
    x, t, it = cm.make_pairs(ctx, th, rho, ie, q, 5000, seed=8)
    xo, to, ito = orl.mlp_pairs(th, rho, ie, q, 5000, seed=8)
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_array_equal(t, to)
    np.testing.assert_array_equal(it, ito)


def test_train_and_guided_adjustment(ctx):
    th, rho, ie, q = moist_states(dims=(48, 16, 16), seed=9)
    dx, dt, it = cm.make_pairs(ctx, th, rho, ie, q, 1 << 18, seed=12,
                               device=True)
    net = cm.Mlp(ctx, seed=13)
    try:
        losses = net.train(dx, dt, steps=20000, batch=8192, lr=3e-3,
                           lr_final=1e-5, seed=14)
        l0, l1 = losses[:20].mean(), losses[-100:].mean()
        print("training loss %.4e -> %.4e" % (l0, l1))
        assert np.all(np.isfinite(losses))
        assert l1 < 0.02 * l0
        # held-out pairs: temperature guess error
        xh, th_, ith = orl.mlp_pairs(th, rho, ie, q, 20000, seed=99)
        yh = net.forward(xh)
        sat = th_[:, 2] > 0
        rms = np.sqrt(np.mean((yh[sat, 0] - th_[sat, 0]) ** 2)) * 10.0
        print("held-out T guess rms error %.4f K (%d saturated)"
              % (rms, sat.sum()))
        assert rms < 0.5
        # guided vs plain adjustment on fresh supersaturated states
        r, e, qq = supersaturated(th, rho, ie, q, 50000, seed=21)
        ie_p, y_p, it_p = cm.saturation_adjust(ctx, th, r, e, qq)
        ie_g, y_g, it_g = cm.saturation_adjust(ctx, th, r, e, qq, net)
        ie_o, y_o, it_o, rc = orl.saturation_adjust(th, r, e, qq)
        assert rc == 0
        np.testing.assert_array_equal(ie_p, ie_o)
        np.testing.assert_array_equal(y_p, y_o)
        np.testing.assert_array_equal(it_p, it_o)
        np.testing.assert_allclose(ie_g, ie_p, rtol=1e-11, atol=0)
        np.testing.assert_allclose(y_g, y_p, rtol=0, atol=1e-13)
        s = it_p > 0
        mp, mg = it_p[s].mean(), it_g[s].mean()
        print("Newton iterations: plain %.3f guided %.3f (%d cells)"
              % (mp, mg, s.sum()))
        assert s.sum() > 10000
        assert np.all(it_g[~s] == 0)
        assert mg < 0.85 * mp
    finally:
        net.close()
        for d in (dx, dt):
            d.free()


def test_run_guided_adjustment_matches_newton(ctx):
    """A moist run with the guided adjustment (satadj = MLP) reproduces the
    plain Newton run and the oracle to the Newton tolerance."""
    from helpers import cfl_dt, perturbed_state, relerr
    blk, th, cfg, w = perturbed_state(4, (24, 8, 8), vel=1.0, dp=1e-3,
                                      cloud=True)
    sk, sj, si = blk.interior()
    # supersaturate the lower levels so that the adjustment condenses
    w[1][sk, sj, si][..., :12] *= 1.6
    w = ca.apply_bc_host(blk, w)
    dt = cfl_dt(blk, th, cfg, w, 0.6)
    u0 = orl.prim_to_cons(blk, th, w)
    ths, rho, ie, q = moist_states(dims=(24, 8, 8), seed=3)
    dx, dt_, _ = cm.make_pairs(ctx, th, rho, ie, q, 1 << 16, seed=5,
                               device=True)
    net = cm.Mlp(ctx, seed=6)
    try:
        net.train(dx, dt_, steps=2000, batch=4096)
        res = {}
        for mode in (ca.SATADJ_NEWTON, ca.SATADJ_MLP):
            cfg.satadj = mode
            run = ca.Run(ctx, blk, th, cfg)
            try:
                if mode == ca.SATADJ_MLP:
                    run.set_mlp(net)
                run.set_prim(w)
                run.step(dt, 2)
                res[mode] = run.cons()
            finally:
                run.close()
        cfg.satadj = ca.SATADJ_NEWTON
        uo, _ = orl.step(blk, th, cfg, dt, 2, u0, w)
        a = res[ca.SATADJ_MLP][:, sk, sj, si]
        b = res[ca.SATADJ_NEWTON][:, sk, sj, si]
        np.testing.assert_array_equal(b, uo[:, sk, sj, si])
        e = relerr(a, b)
        print("guided vs Newton run relerr %.3e" % e)
        # The two Newton iterations end within round-off of the same root;
        # the next stages amplify those ulp-level differences (the
        # hydrostatic decomposition subtracts nearly equal pressures) to
        # ~1e-11 after two steps, as any reordering of FP64 arithmetic does.
        assert e < 1e-9
        # the adjustment did condense
        assert np.max(b[2]) > 0
    finally:
        net.close()
        dx.free()
        dt_.free()
