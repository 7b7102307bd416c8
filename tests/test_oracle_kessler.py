"""Pins of the oracle's Kessler microphysics and precipitation sedimentation
(or_kessler, the operator applied after the last stage when any of the
cg_config rates k_auto, k_accr, k_evap, v_fall is non-zero).

The rates are those of examples/2019-Li-snap/earth-precip.yaml:44-53 and
water-kessler.yaml:4-18; their rate laws are this project's definition
(kintera is not vendored), so the pins are closed forms of that definition
and the conservation laws it must keep:
* nothing to convert (no cloud, no precipitation): the state is unchanged
  bit for bit;
* autoconversion and accretion alone: the cloud density drops by
  dt (k_auto + k_accr q_p) rho_c, the precipitation gains it, energy and
  vapour untouched; a step longer than 1 / k converts all the cloud;
* evaporation in a sub-saturated cell: b = dt k_evap rho_p (1 - rho_v /
  rho_s(T)) with rho_s from the saturation vapour pressure written out here
  in numpy, the total energy drops by L b, total water is kept; a
  saturated cell does not evaporate;
* sedimentation of a uniform column: the top cell keeps (1 - c) of its
  precipitation, the bottom cell gains c, the cells between are unchanged
  (c = dt v_fall / dx1), velocities and the donor's temperature are kept;
* random columns: water, total mass, momentum and energy of every column
  are conserved (the fluxes telescope, nothing crosses the walls).
"""
import numpy as np

import canoe_amd as ca
import oracle_lib as orl


def _block(nx1=12, nx2=3, nx3=2):
    blk, th, cfg = ca.pgen_setup(4, nx1, nx2, nx3)
    return blk, th, cfg


def _state(blk, th, rng, qv, qc, qp, rho=1.0, T=285.0, vel=(0.0, 0.0, 0.0)):
    """Interior prim state of uniform (or per-cell array) fields."""
    w = np.zeros(blk.shape())
    sk, sj, si = blk.interior()
    w[0][sk, sj, si] = rho
    w[1][sk, sj, si] = qv
    w[2][sk, sj, si] = qc
    w[3][sk, sj, si] = qp
    for d in range(3):
        w[4 + d][sk, sj, si] = vel[d]
    # p from T: p = rho Rd T feps (feps = 1 + qv (mu_d/mu_v - 1) - qc - qp)
    feps = 1 + w[1] * th.inv_mu_ratio_m1[0] - w[2] - w[3]
    w[7][sk, sj, si] = (rho * th.rd * T * feps)[sk, sj, si]
    u = orl.prim_to_cons(blk, th, w)
    return u


def _kessler(blk, th, cfg, dt, u):
    return orl.kessler(blk, th, cfg, dt, u)


def _cfg(cfg, k_auto=0.0, k_accr=0.0, k_evap=0.0, v_fall=0.0):
    cfg.k_auto, cfg.k_accr, cfg.k_evap, cfg.v_fall = k_auto, k_accr, k_evap, v_fall
    return cfg


def _esat_rho(th, T):
    """Saturation vapour density e_s(T) / (R_v T) of the svp formula
    p3 exp((1 - t3/T) beta - delta ln(T/t3))."""
    es = th.p3 * np.exp((1 - th.t3 / T) * th.beta - th.delta * np.log(T / th.t3))
    rv = th.rd * (1 + th.inv_mu_ratio_m1[0])
    return es / (rv * T)


def test_nothing_to_convert_is_a_no_op():
    blk, th, cfg = _block()
    rng = np.random.default_rng(1)
    u = _state(blk, th, rng, qv=0.004, qc=0.0, qp=0.0, vel=(1.0, -2.0, 0.5))
    cfg = _cfg(cfg, 0.01, 0.01, 0.1, 5.0)
    np.testing.assert_array_equal(_kessler(blk, th, cfg, 0.7, u), u)


def test_autoconversion_and_accretion_closed_form():
    blk, th, cfg = _block()
    rng = np.random.default_rng(2)
    # saturated vapour so that nothing evaporates
    u = _state(blk, th, rng, qv=0.02, qc=1e-3, qp=4e-4)
    sk, sj, si = blk.interior()
    I = (slice(None), sk, sj, si)
    ui = u[I]
    rho = ui[0] + ui[1] + ui[2] + ui[3]
    dt = 3.0
    for ka, kc in ((0.01, 0.0), (0.0, 0.02), (0.01, 0.02)):
        out = _kessler(blk, th, _cfg(cfg, ka, kc), dt, u)
        a = dt * (ka + kc * (ui[3] / rho)) * ui[2]
        np.testing.assert_allclose(out[I][2], ui[2] - a, rtol=1e-14)
        np.testing.assert_allclose(out[I][3], ui[3] + a, rtol=1e-14)
        np.testing.assert_array_equal(out[1], u[1])
        np.testing.assert_array_equal(out[7], u[7])
    # a step longer than 1 / k_auto converts all of the cloud
    out = _kessler(blk, th, _cfg(cfg, 0.5), 3.0, u)
    assert np.all(out[2][sk, sj, si] == 0.0)
    np.testing.assert_allclose(out[3][sk, sj, si], (u[3] + u[2])[sk, sj, si],
                               rtol=1e-15)


def test_evaporation_closed_form_and_energy():
    blk, th, cfg = _block()
    rng = np.random.default_rng(3)
    T = 290.0
    u = _state(blk, th, rng, qv=0.004, qc=0.0, qp=6e-4, T=T)
    sk, sj, si = blk.interior()
    rho = (u[0] + u[1] + u[2] + u[3])[sk, sj, si]
    rs = _esat_rho(th, T)
    assert np.all(u[1][sk, sj, si] < rs)  # sub-saturated
    dt, ke = 2.0, 0.1
    out = _kessler(blk, th, _cfg(cfg, k_evap=ke), dt, u)
    b = dt * ke * u[3][sk, sj, si] * (1 - u[1][sk, sj, si] / rs)
    np.testing.assert_allclose(out[1][sk, sj, si] - u[1][sk, sj, si], b,
                               rtol=1e-9)
    np.testing.assert_allclose(u[3][sk, sj, si] - out[3][sk, sj, si], b,
                               rtol=1e-9)
    np.testing.assert_allclose(u[7][sk, sj, si] - out[7][sk, sj, si],
                               th.latent * b, rtol=1e-9)
    water0 = (u[1] + u[2] + u[3])[sk, sj, si]
    water1 = (out[1] + out[2] + out[3])[sk, sj, si]
    np.testing.assert_allclose(water1, water0, rtol=1e-15)
    # the cell cooled: T drops by L b / (rho cv_mix) to first order
    w1 = orl.cons_to_prim(blk, th, out)[:, sk, sj, si]
    feps = 1 + w1[1] * th.inv_mu_ratio_m1[0] - w1[2] - w1[3]
    T1 = w1[7] / (w1[0] * th.rd * feps)
    assert np.all(T1 < T)
    # a saturated cell keeps its precipitation
    us = _state(blk, th, rng, qv=1.2 * rs / rho.mean(), qc=0.0, qp=6e-4, T=T)
    outs = _kessler(blk, th, _cfg(cfg, k_evap=ke), dt, us)
    np.testing.assert_array_equal(outs, us)


def test_sedimentation_of_a_uniform_column():
    blk, th, cfg = _block(nx1=10, nx2=2, nx3=1)
    rng = np.random.default_rng(4)
    vel = (0.0, 3.0, -1.0)
    u = _state(blk, th, rng, qv=0.02, qc=0.0, qp=5e-4, vel=vel)
    dt, vf = 4.0, 2.5
    c = dt * vf / blk.dx1
    out = _kessler(blk, th, _cfg(cfg, v_fall=vf), dt, u)
    sk, sj, si = blk.interior()
    rp = u[3][sk, sj, si]
    got = out[3][sk, sj, si]
    np.testing.assert_allclose(got[..., -1], rp[..., -1] * (1 - c), rtol=1e-14)
    np.testing.assert_allclose(got[..., 0], rp[..., 0] * (1 + c), rtol=1e-14)
    np.testing.assert_allclose(got[..., 1:-1], rp[..., 1:-1], rtol=1e-14)
    w0, w1 = orl.cons_to_prim(blk, th, u), orl.cons_to_prim(blk, th, out)
    for d in range(3):  # velocities are kept
        np.testing.assert_allclose(w1[4 + d][sk, sj, si],
                                   w0[4 + d][sk, sj, si], atol=1e-13)

    def temp(w):
        w = w[:, sk, sj, si]
        feps = 1 + w[1] * th.inv_mu_ratio_m1[0] - w[2] - w[3]
        return w[7] / (w[0] * th.rd * feps)
    # every cell had the same temperature: mixing leaves it
    np.testing.assert_allclose(temp(w1), temp(w0), rtol=1e-13)


def test_random_columns_conserve_water_mass_momentum_energy():
    blk, th, cfg = _block(nx1=16, nx2=4, nx3=3)
    rng = np.random.default_rng(5)
    sk, sj, si = blk.interior()
    shp = (blk.nx3, blk.nx2, blk.nx1)
    qv = rng.uniform(0.002, 0.02, shp)
    qc = rng.uniform(0, 2e-3, shp) * (rng.uniform(0, 1, shp) < 0.5)
    qp = rng.uniform(0, 1e-3, shp)
    w = np.zeros(blk.shape())
    w[0][sk, sj, si] = rng.uniform(0.8, 1.2, shp)
    w[1][sk, sj, si], w[2][sk, sj, si], w[3][sk, sj, si] = qv, qc, qp
    for d in range(3):
        w[4 + d][sk, sj, si] = rng.uniform(-5, 5, shp)
    w[7][sk, sj, si] = rng.uniform(7e4, 1e5, shp)
    u = orl.prim_to_cons(blk, th, w)
    out = _kessler(blk, th, _cfg(cfg, 0.01, 0.01, 0.1, 6.0), 5.0, u)
    b = lambda a: a[sk, sj, si].sum(axis=-1)  # column sums
    np.testing.assert_allclose(b(out[1] + out[2] + out[3]),
                               b(u[1] + u[2] + u[3]), rtol=1e-14)
    np.testing.assert_allclose(b(out[0] + out[1] + out[2] + out[3]),
                               b(u[0] + u[1] + u[2] + u[3]), rtol=1e-15)
    np.testing.assert_array_equal(out[0], u[0])
    for n in (4, 5, 6):
        np.testing.assert_allclose(b(out[n]), b(u[n]), atol=1e-12)
    # energy: sensible + kinetic kept up to the latent heat of evaporation
    evap = b(out[1]) - b(u[1])
    np.testing.assert_allclose(b(out[7]) + th.latent * evap, b(u[7]),
                               rtol=1e-14)
