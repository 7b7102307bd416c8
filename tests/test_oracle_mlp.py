"""Pins of the oracle's saturation-guess network and training pairs (CPU).

Each check uses something the oracle does not share code with: numpy's
exp / tanh and float64 rounding (bf16, exp, GELU), an independent numpy
forward pass, central finite differences of the loss (backward pass), and
the ideal-gas closed forms of the network inputs.
"""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl


def _bf16_ref(x):
    """Nearest bf16 (8 significant bits) of float32 values, ties to even,
    by float64 scaling and numpy's round-half-even."""
    x = np.asarray(x, np.float64)
    out = np.empty_like(x)
    for k, v in enumerate(x):
        if not np.isfinite(v) or v == 0.0:
            out[k] = v
            continue
        e = np.floor(np.log2(abs(v)))
        q = 2.0 ** (e - 7)
        r = np.round(v / q) * q
        out[k] = r
    return out


def _bits_to_float(b):
    return (np.asarray(b, np.uint32) << 16).view(np.float32).astype(np.float64)


def test_bf16_round_to_nearest_even():
    rng = np.random.default_rng(3)
    xs = np.concatenate([
        rng.standard_normal(4000).astype(np.float32),
        (rng.standard_normal(2000) * 1e-20).astype(np.float32),
        (rng.standard_normal(2000) * 1e20).astype(np.float32),
        np.float32([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8,
                    0.0, 1.0, 2.0 ** -126 * 1.5])])
    got = _bits_to_float([orl.lib().or_mlp_bf16(float(x)) for x in xs])
    np.testing.assert_array_equal(got, _bf16_ref(xs))
    # ties: 1 + 2^-8 -> 1 (even), 1 + 3 2^-8 -> 1 + 2^-6
    assert orl.lib().or_mlp_bf16(1.0 + 2.0 ** -8) == 0x3F80
    assert orl.lib().or_mlp_bf16(1.0 + 3 * 2.0 ** -8) == 0x3F82
    assert orl.lib().or_mlp_bf16(float("inf")) == 0x7F80
    assert np.isnan(_bits_to_float([orl.lib().or_mlp_bf16(float("nan"))])[0])


def test_exp_f_accuracy():
    xs = np.linspace(-80, 80, 20001).astype(np.float32)
    got = np.array([orl.lib().or_mlp_exp(float(x)) for x in xs], np.float64)
    ref = np.exp(xs.astype(np.float64))
    assert np.max(np.abs(got / ref - 1)) < 3e-7
    assert orl.lib().or_mlp_exp(0.0) == 1.0
    assert orl.lib().or_mlp_exp(100.0) == np.inf
    assert orl.lib().or_mlp_exp(-100.0) == 0.0


def _gelu64(z):
    z = np.asarray(z, np.float64)
    return 0.5 * z * (1 + np.tanh(np.sqrt(2 / np.pi) * (z + 0.044715 * z ** 3)))


def test_gelu_and_derivative():
    zs = np.linspace(-12, 12, 4801).astype(np.float32)
    g = np.array([orl.lib().or_mlp_gelu(float(z)) for z in zs])
    assert np.max(np.abs(g - _gelu64(zs)) / np.maximum(1, np.abs(zs))) < 2e-6
    # Hendrycks & Gimpel tanh form at +-1
    assert abs(orl.lib().or_mlp_gelu(1.0) - 0.8411919906082768) < 2e-6
    assert abs(orl.lib().or_mlp_gelu(-1.0) + 0.15880800939172324) < 2e-6
    assert orl.lib().or_mlp_gelu(0.0) == 0.0
    assert orl.lib().or_mlp_gelu(30.0) == 30.0
    assert orl.lib().or_mlp_gelu(-30.0) == 0.0
    d = np.array([orl.lib().or_mlp_gelu_grad(float(z)) for z in zs])
    h = 1e-5
    fd = (_gelu64(zs.astype(np.float64) + h) - _gelu64(zs.astype(np.float64) - h)) / (2 * h)
    assert np.max(np.abs(d - fd)) < 2e-6


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(orl.lib().or_mlp_nparam(), np.float32)
    o = 0
    for din, dout in ((4, 64), (64, 64), (64, 64), (64, 3)):
        a = np.sqrt(6.0 / (din + dout))
        p[o:o + din * dout] = rng.uniform(-a, a, din * dout)
        o += din * dout
        p[o:o + dout] = rng.uniform(-0.1, 0.1, dout)
        o += dout
    return p


def _unpack(p):
    p = np.asarray(p, np.float64)
    out, o = [], 0
    for din, dout in ((4, 64), (64, 64), (64, 64), (64, 3)):
        W = p[o:o + din * dout].reshape(dout, din)
        o += din * dout
        b = p[o:o + dout]
        o += dout
        out.append((W, b))
    return out


def _np_forward(p, x):
    h = np.asarray(x, np.float64)
    layers = _unpack(p)
    for k, (W, b) in enumerate(layers):
        h = h @ W.T + b
        if k < 3:
            h = _gelu64(h)
    return h


def test_forward_exact_matches_numpy():
    import canoe_amd.mlp
    assert orl.lib().or_mlp_nparam() == canoe_amd.mlp.NPARAM
    p = _params(1)
    x = np.random.default_rng(2).standard_normal((50, 4)).astype(np.float32)
    y = orl.mlp_forward(p, x, emulate=0)
    np.testing.assert_allclose(y, _np_forward(p, x), rtol=1e-12, atol=1e-13)


def test_forward_emulated_close_to_exact():
    p = _params(4)
    x = np.random.default_rng(5).standard_normal((200, 4)).astype(np.float32)
    ye = orl.mlp_forward(p, x, emulate=1)
    y0 = orl.mlp_forward(p, x, emulate=0)
    assert np.linalg.norm(ye - y0) / np.linalg.norm(y0) < 2e-2
    assert np.max(np.abs(ye - y0)) < 0.05 * np.max(np.abs(y0))


def test_backward_exact_matches_finite_differences():
    p = _params(7)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((6, 4)).astype(np.float32)
    t = rng.standard_normal((6, 3)).astype(np.float32)
    g, loss = orl.mlp_loss_grad(p, x, t, emulate=0)
    y = _np_forward(p, x)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-12 * max(1.0, loss)
    # every bias and 40 weights of every layer
    idx, o = [], 0
    for din, dout in ((4, 64), (64, 64), (64, 64), (64, 3)):
        idx += list(rng.choice(din * dout, size=min(40, din * dout),
                               replace=False) + o)
        o += din * dout
        idx += list(range(o, o + dout))
        o += dout
    h = np.float32(2.0 ** -9)
    worst = 0.0
    scale = np.max(np.abs(g))
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] = p[k] + h
        pm[k] = p[k] - h
        delta = float(pp[k]) - float(pm[k])
        lp = orl.mlp_loss_grad(pp, x, t, emulate=0)[1]
        lm = orl.mlp_loss_grad(pm, x, t, emulate=0)[1]
        fd = (lp - lm) / delta
        worst = max(worst, abs(fd - g[k]) / scale)
    assert worst < 2e-4


def test_backward_emulated_close_to_exact():
    p = _params(9)
    rng = np.random.default_rng(10)
    x = rng.standard_normal((256, 4)).astype(np.float32)
    t = rng.standard_normal((256, 3)).astype(np.float32)
    ge, le = orl.mlp_loss_grad(p, x, t, emulate=1)
    g0, l0 = orl.mlp_loss_grad(p, x, t, emulate=0)
    assert abs(le - l0) < 2e-2 * l0
    o = 0
    for din, dout in ((4, 64), (64, 64), (64, 64), (64, 3)):
        for n in (din * dout, dout):
            a, b = ge[o:o + n], g0[o:o + n]
            assert np.linalg.norm(a - b) < 5e-2 * np.linalg.norm(b) + 1e-12
            o += n


def test_features_closed_forms():
    """x0 and x1 of an all-vapour cell are its temperature p/(rho R) and
    ln(p/1e5) (ideal gas with R = R_d q_d + R_v q_v)."""
    blk, th, cfg = ca.pgen_setup(4, 16, 4, 4)
    w = ca.pgen_init(4, 3, blk, th, cfg)
    sk, sj, si = blk.interior()
    ivx = 1 + blk.nmass
    rho = w[0][sk, sj, si].ravel().copy()
    q = np.stack([w[n][sk, sj, si].ravel() for n in (1, 2, 3)])
    p = w[ivx + 3][sk, sj, si].ravel()
    assert np.all(q[1] == 0.0)
    rv = th.rd * (1 + th.inv_mu_ratio_m1[0])
    R = th.rd * (1 - q.sum(0)) + rv * q[0]
    T = p / (rho * R)
    cvd = th.rd / (th.gammad - 1)
    cv = cvd * (1 - q.sum(0)) + cvd * (1 + th.cv_ratio_m1[0]) * q[0] \
        + cvd * (1 + th.cv_ratio_m1[2]) * q[2]
    ie = rho * cv * T
    x, tu = orl.mlp_features(th, rho, ie, q)
    np.testing.assert_allclose(tu, T, rtol=1e-13)
    np.testing.assert_allclose(x[:, 0], (T - 260) / 40, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(x[:, 1], np.log(p / 1e5), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(x[:, 2], 40 * q[0], rtol=1e-6)
    np.testing.assert_allclose(x[:, 3], 40 * q[2], rtol=1e-6, atol=1e-12)


def test_pairs_properties():
    blk, th, cfg = ca.pgen_setup(4, 16, 4, 4)
    w = ca.pgen_init(4, 3, blk, th, cfg)
    sk, sj, si = blk.interior()
    rho = w[0][sk, sj, si].ravel().copy()
    q = np.stack([w[n][sk, sj, si].ravel() for n in (1, 2, 3)])
    ivx = 1 + blk.nmass
    p = w[ivx + 3][sk, sj, si].ravel()
    cvd = th.rd / (th.gammad - 1)
    rv = th.rd * (1 + th.inv_mu_ratio_m1[0])
    R = th.rd * (1 - q.sum(0)) + rv * q[0]
    cv = cvd * (1 - q.sum(0)) + cvd * (1 + th.cv_ratio_m1[0]) * q[0] \
        + cvd * (1 + th.cv_ratio_m1[2]) * q[2]
    ie = rho * cv * p / (rho * R)
    x, t, it = orl.mlp_pairs(th, rho, ie, q, 4000, seed=5)
    x2, t2, _ = orl.mlp_pairs(th, rho, ie, q, 4000, seed=5)
    x3, _, _ = orl.mlp_pairs(th, rho, ie, q, 4000, seed=6)
    np.testing.assert_array_equal(x, x2)
    np.testing.assert_array_equal(t, t2)
    assert not np.array_equal(x, x3)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(t))
    np.testing.assert_allclose(t[:, 1] + t[:, 2], 1.0, atol=1e-6)
    sat = t[:, 2] > 0
    assert 0.5 < sat.mean() < 1.0
    assert np.all(t[sat, 0] > 0)          # latent heating warms
    assert np.all(it[sat] >= 1)
    assert np.all(t[~sat, 0] == 0) and np.all(it[~sat] == 0)
    # temperatures inside the base range +- 3 K, water scaled by 0.8..2.5
    tb = p / (rho * R)
    assert x[:, 0].min() * 40 + 260 > tb.min() - 3.01
    assert x[:, 0].max() * 40 + 260 < tb.max() + 3.01
    assert x[:, 2].max() <= 40 * 2.5 * q[0].max() * (1 + 1e-6)
