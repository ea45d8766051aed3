"""The seeded workload generators against the distributions they claim.

* Plummer (Aarseth, Henon & Wielen 1974): half-mass radius, virial ratio
  Q = T / |W| = 1/2, centre of mass and momentum at the origin.
* Cube + Plummer core: the 50/50 split, log-normal masses normalised to 1,
  the cube's extent.
* Two galaxies: the bulge fraction, the exponential disk's scale length and
  sech^2 height, the flat rotation speed, the Hernquist bulge's half-mass
  radius, and the parabolic orbit (zero orbital energy at the stated
  separation).
* Ensemble: sizes log-uniform in [4,096, 524,288], virial ratios in
  [0.3, 0.7] (kinetic energy scaled from the Q = 1/2 sphere).
"""
import math

import numpy as np

from gravity_amd import initial_conditions as ic


def _potential(pos, mass):
    """Exact (unsoftened) potential energy sum_{i<j} -m_i m_j / r_ij, float64."""
    pos = pos.astype(np.float64)
    mass = mass.astype(np.float64)
    w = 0.0
    for i in range(0, pos.shape[0], 1024):
        d = pos[i:i + 1024, None, :] - pos[None, :, :]
        r = np.sqrt((d * d).sum(-1))
        mm = mass[i:i + 1024, None] * mass[None, :]
        ii = np.arange(i, min(i + 1024, pos.shape[0]))
        r[np.arange(ii.size), ii] = np.inf
        w -= 0.5 * (mm / r).sum()
    return w


def test_plummer_half_mass_radius_and_centre():
    b = ic.plummer(200_000, 3)
    assert b.dtype == np.float32 and b.shape == (200_000, 8)
    np.testing.assert_allclose(b[:, 3].astype(np.float64).sum(), 1.0, rtol=1e-6)
    r = np.linalg.norm(b[:, 0:3].astype(np.float64), axis=1)
    # X = r^3 / (r^2 + 1)^(3/2) uniform in (0, 0.999]: the median radius is
    # 1 / sqrt(X^(-2/3) - 1) at X = 0.4995, about 1.3035 (1.305 without the cut)
    x = 0.999 / 2
    want = 1.0 / math.sqrt(x ** (-2.0 / 3.0) - 1.0)
    assert abs(np.median(r) - want) < 0.01 * want
    # the radial cut at 99.9 % of the mass: r < 1 / sqrt(0.999^(-2/3) - 1) = 38.7
    assert r.max() < 38.72
    assert np.abs(b[:, 0:3].astype(np.float64).mean(axis=0)).max() < 1e-6
    assert np.abs(b[:, 4:7].astype(np.float64).mean(axis=0)).max() < 1e-6


def test_plummer_virial_ratio():
    b = ic.plummer(8192, 5)
    m = b[:, 3].astype(np.float64)
    v = b[:, 4:7].astype(np.float64)
    T = 0.5 * (m * (v * v).sum(1)).sum()
    W = _potential(b[:, 0:3], m)
    # Q = 1/2 for the distribution function (3 pi / 64 * T = ... equilibrium);
    # the 0.1 % mass cut and the 8,192-body sample move it by ~2 %
    assert abs(T / abs(W) - 0.5) < 0.03, T / abs(W)


def test_cube_plus_core():
    n = 100_000
    b = ic.cube_plus_core(n, 9, density=64.0)
    m = b[:, 3].astype(np.float64)
    np.testing.assert_allclose(m.sum(), 1.0, rtol=1e-6)
    # log-normal(0, 0.5) before normalisation: the log-masses have sigma 0.5
    assert abs(np.log(m).std() - 0.5) < 0.01
    cube, core = b[: n // 2], b[n // 2:]
    assert np.abs(cube[:, 0:3]).max() <= 1.0 and np.abs(cube[:, 4:7]).max() == 0.0
    assert abs(np.abs(cube[:, 0:3]).mean() - 0.5) < 0.01   # uniform in [-1, 1]
    rc = np.linalg.norm(core[:, 0:3].astype(np.float64), axis=1)
    assert 0.1 < np.median(rc) < 0.16                       # Plummer core, a = 0.1
    assert np.all(b[:, 7] == 64.0)


def _galaxies(n_per, seed=23):
    b = ic.two_galaxy(n_per, seed).astype(np.float64)
    return b[:n_per], b[n_per:]


def test_two_galaxy_components():
    n_per = 100_000
    g0, g1 = _galaxies(n_per)
    n_bulge = int(round(0.2 * n_per))
    n_disk = n_per - n_bulge
    for g in (g0, g1):
        np.testing.assert_allclose(g[:, 3].sum(), 1.0, rtol=1e-6)     # each galaxy has mass 1
        assert np.all(g[:, 3] == g[0, 3])                             # equal masses: bulge 20 % of the mass
    # galaxy 0 is not tilted: its disk lies in the xy plane about its centre
    centre = np.average(g0[:, 0:3], axis=0, weights=g0[:, 3])
    disk, bulge = g0[:n_disk] - 0.0, g0[n_disk:]
    dp = disk[:, 0:3] - centre
    R = np.hypot(dp[:, 0], dp[:, 1])
    assert abs(R.mean() - 2.0) < 0.03                   # R ~ Gamma(2, R_d = 1): mean 2 R_d
    # sech^2(z / z0): the median |z| is z0 artanh(1/2)
    assert abs(np.median(np.abs(dp[:, 2])) - 0.1 * math.atanh(0.5)) < 0.003
    # flat rotation curve v = 1 about the galaxy's own motion
    vcom = np.average(g0[:, 4:7], axis=0, weights=g0[:, 3])
    dv = disk[:, 4:7] - vcom
    vt = (-dp[:, 1] * dv[:, 0] + dp[:, 0] * dv[:, 1]) / R
    assert abs(np.median(vt) - 1.0) < 0.02
    # Hernquist bulge a = 0.2, mass cut at 99 %: M(<r)/M = r^2/(r+a)^2 gives
    # the median radius a sqrt(x) / (1 - sqrt(x)) at x = 0.495
    rb = np.linalg.norm(bulge[:, 0:3] - centre, axis=1)
    s = math.sqrt(0.99 / 2)
    assert abs(np.median(rb) - 0.2 * s / (1 - s)) < 0.01


def test_two_galaxy_parabolic_orbit():
    g0, g1 = _galaxies(50_000, seed=4)
    c0 = np.average(g0[:, 0:3], axis=0, weights=g0[:, 3])
    c1 = np.average(g1[:, 0:3], axis=0, weights=g1[:, 3])
    v0 = np.average(g0[:, 4:7], axis=0, weights=g0[:, 3])
    v1 = np.average(g1[:, 4:7], axis=0, weights=g1[:, 3])
    d = np.linalg.norm(c1 - c0)
    assert abs(d - 10.0) < 1e-4
    # point-mass orbital energy of the pair: mu v^2 / 2 - G M1 M2 / d = 0
    mu = 0.5
    e = 0.5 * mu * float(np.dot(v1 - v0, v1 - v0)) - 1.0 / d
    assert abs(e) < 1e-4 * (1.0 / d), e
    # the centre of mass of the pair is at rest at the origin
    assert np.abs(v0 + v1).max() < 1e-5 and np.abs(c0 + c1).max() < 1e-4


def test_ensemble_sizes_log_uniform():
    s = ic.ensemble_sizes(20_000, 8)
    assert s.min() >= 4096 and s.max() <= 524288
    u = (np.log(s) - math.log(4096)) / (math.log(524289) - math.log(4096))
    # uniform on [0, 1]: mean 1/2, variance 1/12, and every decile filled
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1.0 / 12.0) < 0.005
    h, _ = np.histogram(u, bins=10, range=(0, 1))
    assert h.min() > 0.9 * 2000 and h.max() < 1.1 * 2000
    # the 256-simulation ensemble's total is near its expectation
    # E[N] = (hi - lo) / ln(hi / lo) = 107,311 per simulation
    tot = int(ic.ensemble_sizes(256, 2024).sum())
    assert 0.8 * 256 * 107311 < tot < 1.2 * 256 * 107311


def test_ensemble_virial_ratios():
    seed = 3
    sims = ic.ensemble(4, seed)
    sizes = ic.ensemble_sizes(4, seed)
    q = np.random.default_rng(seed + 1).uniform(0.3, 0.7, 4)
    for i, b in enumerate(sims):
        assert b.shape[0] == sizes[i]
        base = ic.plummer(int(sizes[i]), seed * 1000 + i)    # the same sphere at Q = 1/2
        np.testing.assert_array_equal(b[:, 0:4], base[:, 0:4])
        t = (b[:, 3].astype(np.float64) * (b[:, 4:7].astype(np.float64) ** 2).sum(1)).sum()
        t0 = (base[:, 3].astype(np.float64) * (base[:, 4:7].astype(np.float64) ** 2).sum(1)).sum()
        np.testing.assert_allclose(t / t0, q[i] / 0.5, rtol=1e-5)
        assert 0.3 <= q[i] <= 0.7
