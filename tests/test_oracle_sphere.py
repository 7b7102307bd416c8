"""Pins of the oracle's cubed-sphere step (config 3; oracle_sphere.inc).

* Panel topology: the neighbour of every panel side, the side it meets and
  the orientation along the edge are derived from the panel positions of
  GetLatLon (src/exo3/cubed_sphere.cpp:78-89); they must equal the
  reference's hard-coded tables tgbid, tgside and dinv
  (src/exo3/cubed_sphere_utility.cpp:382-395).
* Metric (gnomonic_equiangle.cpp:136-196, 221-260, 407-416): cos^2 + sin^2
  = 1; the x1-face areas of the six panels sum to 4 pi r^2 and the volumes
  to the shell volume, to second order in the angular spacing.
* (U, V) maps (cubed_sphere_transform.cpp:11-57): each inverts the other and
  preserves the length of a vector; (U, V) are the components along the
  local west and north directions of the GetLatLon positions (finite
  differences of the position map give the panel's tangent vectors).
* Ghost zones across panel edges: a constant field is copied exactly; a
  smooth scalar and a solid-body rotation are reproduced at the ghost
  points to second order (the error drops about 4x when the panels are
  refined 2x) -- a wrong neighbour, orientation or vector map gives O(1).
* Forcing (hs94.cpp:56-151): the Coriolis force does no work, has magnitude
  |f| rho |v| and turns a northern-hemisphere eastward wind southward; the
  Held-Suarez relaxation and drag equal their closed forms written here.
* A whole step conserves the total mass of the six panels to rounding (the
  exchanged face states make the fluxes of neighbouring panels agree), and
  an isothermal atmosphere at rest stays near rest.
"""
import numpy as np
import pytest

import canoe_amd as ca
import oracle_lib as orl

# cubed_sphere_utility.cpp:382-395 (rows: panels 1-6; columns: Up (x2
# lower), Down (x2 upper), Left (x3 lower), Right (x3 upper))
TGBID = [[6, 2, 3, 4], [1, 5, 3, 4], [1, 5, 6, 2], [1, 5, 2, 6], [2, 6, 3, 4],
         [1, 5, 4, 3]]
DINV = [[1, 0, 0, 1], [0, 0, 0, 0], [0, 1, 0, 0], [1, 0, 0, 0], [0, 1, 1, 0],
        [1, 1, 0, 0]]
TGSIDE = [[-1, -1, -1, -1], [1, -1, 1, -1], [-1, -1, 1, -1], [1, 1, 1, -1],
          [1, 1, 1, 1], [-1, 1, 1, -1]]


def _xyz(p, a2, a3):
    lat, lon = orl.sph_latlon(p, np.tan(a2), np.tan(a3))
    return np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon),
                     np.sin(lat)])


def _tangents(p, a2, a3, h=1e-6):
    t2 = (_xyz(p, a2 + h, a3) - _xyz(p, a2 - h, a3)) / (2 * h)
    t3 = (_xyz(p, a2, a3 + h) - _xyz(p, a2, a3 - h)) / (2 * h)
    return t2 / np.linalg.norm(t2), t3 / np.linalg.norm(t3)


def _west_north(P):
    lat = np.arcsin(P[2])
    lon = np.arctan2(P[1], P[0])
    east = np.array([-np.sin(lon), np.cos(lon), 0.0])
    north = np.array([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon),
                      np.cos(lat)])
    return -east, north


def _setup(n, nr=3):
    blk, th, cfg = ca.pgen_setup(3, nr, n, n)
    return blk, th, cfg


def test_panel_links_match_the_reference_tables():
    for p in range(1, 7):
        for s in range(4):
            q, sq, inv = orl.sph_link(p, s)
            assert q == TGBID[p - 1][s], (p, s, q)
            assert inv == DINV[p - 1][s], (p, s, inv)
            side = -1 if sq % 2 == 0 else 1  # lower sides at -pi/4
            assert side == TGSIDE[p - 1][s], (p, s, sq)


def test_metric_identities_and_second_order_areas():
    errs = []
    for n in (8, 16, 32):
        blk, th, cfg = _setup(n)
        R = blk.x1min
        area = sum(orl.sph_metric(blk, p, 13)[..., 0].sum() for p in range(1, 7))
        vol = sum(orl.sph_metric(blk, p, 12).sum() for p in range(1, 7))
        r1 = R + blk.nx1 * blk.dx1
        errs.append((abs(area / (4 * np.pi * R * R) - 1),
                     abs(vol / (4 * np.pi / 3 * (r1 ** 3 - R ** 3)) - 1)))
        for p in range(1, 7):
            for c, s in ((0, 1), (2, 3), (4, 5)):
                cc, ss = orl.sph_metric(blk, p, c), orl.sph_metric(blk, p, s)
                np.testing.assert_allclose(cc * cc + ss * ss, 1.0, rtol=1e-14)
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a[0] / b[0] < 4.5, errs  # second order in the spacing
        assert 3.5 < a[1] / b[1] < 4.5, errs
    assert errs[-1][0] < 2e-3


def test_vector_maps_invert_preserve_length_and_follow_the_positions():
    rng = np.random.default_rng(3)
    for _ in range(60):
        p = int(rng.integers(1, 7))
        a2, a3 = rng.uniform(-np.pi / 4, np.pi / 4, 2)
        X, Y = np.tan(a2), np.tan(a3)
        v2, v3 = rng.normal(size=2)
        U, V = orl.sph_rll_from_abp(X, Y, p, v2, v3)
        b2, b3 = orl.sph_abp_from_rll(X, Y, p, U, V)
        np.testing.assert_allclose([b2, b3], [v2, v3], rtol=1e-12, atol=1e-12)
        cth = -X * Y / np.sqrt((1 + X * X) * (1 + Y * Y))
        np.testing.assert_allclose(U * U + V * V,
                                   v2 * v2 + v3 * v3 + 2 * cth * v2 * v3,
                                   rtol=1e-12)
        t2, t3 = _tangents(p, a2, a3)
        np.testing.assert_allclose(t2 @ t3, cth, atol=1e-8)
        vec = v2 * t2 + v3 * t3
        west, north = _west_north(_xyz(p, a2, a3))
        np.testing.assert_allclose([U, V], [vec @ west, vec @ north], atol=1e-8)


def _smooth_state(blk, th, omega_sb=1e-5):
    """rho = 1 + 0.1 x + 0.2 y + 0.3 z and a solid-body rotation about z,
    contravariant velocities, on the interior of the six panels."""
    w = np.zeros((6,) + blk.shape())
    sk, sj, si = blk.interior()
    R = blk.x1min
    x2v = (np.arange(blk.nx2 + 6) - 3 + 0.5) * (np.pi / 2 / blk.nx2) - np.pi / 4
    for p in range(1, 7):
        for k in range(3, 3 + blk.nx3):
            for j in range(3, 3 + blk.nx2):
                P = _xyz(p, x2v[j], x2v[k])
                t2, t3 = _tangents(p, x2v[j], x2v[k])
                vec = omega_sb * R * np.cross([0, 0, 1.0], P)
                c = t2 @ t3
                g = np.linalg.inv([[1, c], [c, 1]])
                vc = g @ np.array([vec @ t2, vec @ t3])
                w[p - 1, 0, k, j, si] = 1 + 0.1 * P[0] + 0.2 * P[1] + 0.3 * P[2]
                w[p - 1, 2, k, j, si] = vc[0]
                w[p - 1, 3, k, j, si] = vc[1]
                w[p - 1, 4, k, j, si] = 1e5
    return w, x2v


def _ghost_errors(n):
    blk, th, cfg = _setup(n)
    w, x2v = _smooth_state(blk, th)
    u = orl.sph_prim_to_cons(blk, th, w)
    wg = orl.sph_fill_ghosts(blk, th, u, w)
    R = blk.x1min
    erho, ev, vmax = 0.0, 0.0, 0.0
    i = 3
    for p in range(1, 7):
        for m in range(3):
            for t in range(3, 3 + n):
                for (k, j) in ((t, 2 - m), (t, 3 + n + m), (2 - m, t), (3 + n + m, t)):
                    P = _xyz(p, x2v[j], x2v[k])
                    t2, t3 = _tangents(p, x2v[j], x2v[k])
                    rho = 1 + 0.1 * P[0] + 0.2 * P[1] + 0.3 * P[2]
                    vec = 1e-5 * R * np.cross([0, 0, 1.0], P)
                    got = wg[p - 1, 2, k, j, i] * t2 + wg[p - 1, 3, k, j, i] * t3
                    erho = max(erho, abs(wg[p - 1, 0, k, j, i] - rho))
                    ev = max(ev, np.linalg.norm(got - vec))
                    vmax = max(vmax, np.linalg.norm(vec))
    return erho, ev / vmax


def test_ghosts_copy_constants_and_follow_smooth_fields_to_second_order():
    blk, th, cfg = _setup(8)
    w = np.zeros((6,) + blk.shape())
    sk, sj, si = blk.interior()
    w[:, 0][:, sk, sj, si] = 1.25
    w[:, 4][:, sk, sj, si] = 9e4
    wg = orl.sph_fill_ghosts(blk, th, orl.sph_prim_to_cons(blk, th, w), w)
    for p in range(6):
        np.testing.assert_allclose(wg[p, 0, sk, 0:3, si], 1.25, rtol=1e-14)
        np.testing.assert_allclose(wg[p, 0, 0:3, sj, si], 1.25, rtol=1e-14)
        np.testing.assert_allclose(wg[p, 4, sk, -3:, si], 9e4, rtol=1e-14)
        np.testing.assert_allclose(wg[p, 2:4, sk, -3:, si], 0.0, atol=1e-12)
    e8, e16 = _ghost_errors(8), _ghost_errors(16)
    assert e16[0] < 0.3 * e8[0] and e16[1] < 0.3 * e8[1], (e8, e16)
    assert e16[1] < 2e-2, e16


def test_forcing_closed_forms():
    blk, th, cfg = _setup(8)
    rng = np.random.default_rng(9)
    sk, sj, si = blk.interior()
    w = np.zeros((6,) + blk.shape())
    shp = w[:, 0][:, sk, sj, si].shape
    w[:, 0][:, sk, sj, si] = rng.uniform(0.5, 1.2, shp)
    w[:, 1][:, sk, sj, si] = rng.uniform(-1, 1, shp)
    w[:, 2][:, sk, sj, si] = rng.uniform(-20, 20, shp)
    w[:, 3][:, sk, sj, si] = rng.uniform(-20, 20, shp)
    w[:, 4][:, sk, sj, si] = rng.uniform(2e4, 1.05e5, shp)
    sp = ca.pgen_sphere()
    wght = 50.0
    # Coriolis alone
    sp.hs94 = 0
    du = orl.sph_forcing(blk, th, sp, wght, w)
    for p in range(6):
        cth = orl.sph_metric(blk, p + 1, 0)[sk, sj][..., None]
        lat = orl.sph_metric(blk, p + 1, 10)[sk, sj][..., None]
        a2, a3 = du[p, 2][sk, sj, si] / wght, du[p, 3][sk, sj, si] / wght
        v2, v3 = w[p, 2][sk, sj, si], w[p, 3][sk, sj, si]
        rho = w[p, 0][sk, sj, si]
        vv = v2 * v2 + v3 * v3 + 2 * cth * v2 * v3
        s2 = 1 - cth * cth  # contravariant acceleration from covariant
        ac2, ac3 = (a2 - cth * a3) / s2, (a3 - cth * a2) / s2
        aa = a2 * ac2 + a3 * ac3
        f = 2 * sp.omega * np.sin(lat)
        np.testing.assert_allclose(a2 * v2 + a3 * v3, 0.0, atol=1e-12 * np.sqrt(aa * vv).max())
        np.testing.assert_allclose(np.sqrt(aa), np.abs(f) * rho * np.sqrt(vv), rtol=1e-10)
    # an eastward wind in the northern hemisphere is turned southward
    p, k, j = 2, 3 + 1, 3 + 2  # an equatorial panel, cell off the equator
    x2v = (np.arange(blk.nx2 + 6) - 3 + 0.5) * (np.pi / 2 / blk.nx2) - np.pi / 4
    P = _xyz(p, x2v[j], x2v[k])
    if P[2] < 0:
        k = 3 + blk.nx3 - 2
        P = _xyz(p, x2v[j], x2v[k])
    t2, t3 = _tangents(p, x2v[j], x2v[k])
    west, north = _west_north(P)
    c = t2 @ t3
    ve = np.linalg.inv([[1, c], [c, 1]]) @ np.array([-west @ t2, -west @ t3])
    w1 = w.copy()
    w1[p - 1, 2, k, j, si] = 10 * ve[0]
    w1[p - 1, 3, k, j, si] = 10 * ve[1]
    du = orl.sph_forcing(blk, th, sp, wght, w1)
    a = np.array([du[p - 1, 2, k, j, 3], du[p - 1, 3, k, j, 3]]) / wght
    ac = np.linalg.inv([[1, c], [c, 1]]) @ a
    assert (ac[0] * t2 + ac[1] * t3) @ north < 0
    # Held-Suarez relaxation and drag alone
    sp = ca.pgen_sphere()
    sp.omega = 0.0
    du = orl.sph_forcing(blk, th, sp, wght, w)
    cp = th.gammad / (th.gammad - 1) * th.rd
    for p in range(6):
        lat = orl.sph_metric(blk, p + 1, 10)[sk, sj][..., None]
        rho, pr = w[p, 0][sk, sj, si], w[p, 4][sk, sj, si]
        sig = pr / sp.p0
        sp_ = (sig - sp.sigmab) / (1 - sp.sigmab)
        kv = np.where(sp_ < 0, 0.0, sp_ * sp.kf)
        teq = (sp.ts - sp.dt_eq * np.sin(lat) ** 2 -
               sp.dtheta * np.log(sig) * np.cos(lat) ** 2) * sig ** (th.rd / cp)
        teq = np.maximum(teq, 200.0)
        kt = sp.ka + (sp.ks - sp.ka) * np.where(sp_ < 0, 0.0, sp_) * np.cos(lat) ** 4
        T = pr / rho / th.rd
        np.testing.assert_allclose(du[p, 4][sk, sj, si] / wght,
                                   -(cp - th.rd) * rho * kt * (T - teq), rtol=1e-12,
                                   atol=1e-12)
        np.testing.assert_allclose(du[p, 1][sk, sj, si] / wght,
                                   -kv * rho * w[p, 1][sk, sj, si], rtol=1e-12,
                                   atol=1e-16)


@pytest.mark.parametrize("flag", [1, 9, 0])
def test_step_conserves_mass_and_keeps_a_rest_state_near_rest(flag):
    blk, th, cfg = ca.pgen_setup(3, 12, 8, 8)
    cfg.implicit_flag = flag
    sp = ca.pgen_sphere()
    sp.omega, sp.hs94 = 0.0, 0
    w = ca.pgen_sphere_init(5, blk, th)
    sk, sj, si = blk.interior()
    # without the noise: an isothermal atmosphere at rest
    T0 = 300.0
    pr = w[:, 4]
    w[:, 0][:, sk, sj, si] = pr[:, sk, sj, si] / (th.rd * T0)
    u = orl.sph_prim_to_cons(blk, th, w)
    w = orl.sph_fill_ghosts(blk, th, u, w)
    vol = np.stack([orl.sph_metric(blk, p, 12) for p in range(1, 7)])
    mass = lambda a: (a[:, 0][:, sk, sj, si] * vol).sum()
    dt = 100.0 if flag else 5.0
    u1, w1 = orl.sph_step(blk, th, cfg, sp, dt, 3, u, w)
    assert abs(mass(u1) / mass(u) - 1) < 1e-14
    vmax = max(np.abs(w1[:, n][:, sk, sj, si]).max() for n in (1, 2, 3))
    assert vmax < 0.05, vmax
