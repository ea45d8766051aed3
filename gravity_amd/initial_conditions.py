"""Seeded initial-condition generators for the Barnes-Hut step.

The reference draws its scenes from unseeded ``Math.random``
(simulation/PlanetGenerator.java:330,441,565) and has no Plummer, Hernquist or
exponential-disk generator, so every workload here is generated by these
functions from a fixed seed. All of them return a float32 array of shape
``(n, 8)`` laid out exactly as the reference's ``Body`` std430 struct
(common/common.glsl:25-29, gpu/Body.java:74-86): ``pos.x, pos.y, pos.z, mass,
vel.x, vel.y, vel.z, density``.

Units: G = 1 is obtained by the reference's unit derivation
(common/common.glsl:346-352) with ``len = time = 1`` and
``mass = 1 / 6.67430e-11`` (see :func:`g_one_units`).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

GRAVITATIONAL_CONSTANT_SI = 6.67430e-11


@dataclass(frozen=True)
class Units:
    """The reference's ``UnitSet`` uniforms (mass, density, len, time)."""

    mass: float
    density: float
    len: float
    time: float


def g_one_units() -> Units:
    """Units for which the reference's in-shader G evaluates to 1.

    ``G = 6.67430e-11 * len^-3 * time^2 * mass`` (common/common.glsl:346-348);
    the density unit equals the mass unit so that the body-length constant
    ``0.6203504909 * (mass/density)^(1/3) / len`` is 0.6203504909.
    """
    m = float(np.float32(1.0 / GRAVITATIONAL_CONSTANT_SI))
    return Units(mass=m, density=m, len=1.0, time=1.0)


def _pack(pos: np.ndarray, mass: np.ndarray, vel: np.ndarray, density: np.ndarray) -> np.ndarray:
    n = pos.shape[0]
    out = np.empty((n, 8), dtype=np.float32)
    out[:, 0:3] = pos
    out[:, 3] = mass
    out[:, 4:7] = vel
    out[:, 7] = density
    return out


def _isotropic(rng: np.random.Generator, r: np.ndarray) -> np.ndarray:
    """Vectors of length r in uniformly random directions."""
    n = r.shape[0]
    cos_t = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    return np.stack([r * sin_t * np.cos(phi), r * sin_t * np.sin(phi), r * cos_t], axis=1)


def plummer_pos_vel(n: int, rng: np.random.Generator, a: float = 1.0, total_mass: float = 1.0,
                    G: float = 1.0, mass_cut: float = 0.999) -> tuple[np.ndarray, np.ndarray]:
    """Plummer sphere by Aarseth, Henon & Wielen (1974), float64.

    Radii follow the cumulative mass ``X = r^3/(r^2+a^2)^(3/2)`` with X drawn
    uniformly in (0, mass_cut] (the radial cutoff: 99.9% of the mass, r < 38.7a).
    Speeds are ``q * v_esc`` with q drawn by rejection from
    ``g(q) = q^2 (1-q^2)^(7/2)``, ``v_esc = sqrt(2GM) (r^2+a^2)^(-1/4)``.
    The centre of mass and its velocity are moved to the origin.
    """
    x = rng.uniform(0.0, mass_cut, n)
    x = np.maximum(x, 1e-12)
    r = a / np.sqrt(x ** (-2.0 / 3.0) - 1.0)
    pos = _isotropic(rng, r)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        cand = rng.uniform(0.0, 1.0, todo.size)
        y = rng.uniform(0.0, 0.1, todo.size)
        ok = y < cand * cand * (1.0 - cand * cand) ** 3.5
        q[todo[ok]] = cand[ok]
        todo = todo[~ok]
    v_esc = math.sqrt(2.0 * G * total_mass) * (r * r + a * a) ** (-0.25)
    vel = _isotropic(rng, q * v_esc)
    pos -= pos.mean(axis=0)
    vel -= vel.mean(axis=0)
    return pos, vel


def plummer(n: int, seed: int, a: float = 1.0, density: float = 1.0) -> np.ndarray:
    """Plummer sphere, equal masses 1/n, G = 1, scale radius ``a``.

    The Plummer workloads: N=65,536 for the reference's CPU-runnable case and
    N=16,777,216 for the headline throughput case.
    """
    rng = np.random.default_rng(seed)
    pos, vel = plummer_pos_vel(n, rng, a=a)
    mass = np.full(n, 1.0 / n)
    return _pack(pos, mass, vel, np.full(n, density))


def cube_plus_core(n: int, seed: int, core_a: float = 0.1, sigma: float = 0.5,
                   density: float = 64.0) -> np.ndarray:
    """Uniform cube [-1,1]^3 (n/2 bodies, at rest) plus a Plummer core (n/2).

    Masses are log-normal(mu=0, sigma) normalised to a total of 1; the core
    is a Plummer sphere of scale radius ``core_a`` whose velocities follow its
    own mass. Every body has the same density, so the reference's body radius
    ``0.62035 * (m/rho)^(1/3)`` (common/common.glsl:228-230) is proportional
    to m^(1/3) and overlapping bodies merge.
    """
    rng = np.random.default_rng(seed)
    n_cube = n // 2
    n_core = n - n_cube
    mass = rng.lognormal(0.0, sigma, n)
    mass /= mass.sum()
    cube = rng.uniform(-1.0, 1.0, (n_cube, 3))
    core_mass = float(mass[n_cube:].sum())
    cpos, cvel = plummer_pos_vel(n_core, rng, a=core_a, total_mass=core_mass)
    pos = np.concatenate([cube, cpos])
    vel = np.concatenate([np.zeros((n_cube, 3)), cvel])
    return _pack(pos, mass, vel, np.full(n, density))


def _exp_disk(n: int, rng: np.random.Generator, rd: float, z0: float, vc: float) -> tuple[np.ndarray, np.ndarray]:
    # surface density ~ exp(-R/rd): R ~ Gamma(2, rd); vertical sech^2(z/z0)
    R = rng.gamma(2.0, rd, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    u = rng.uniform(-1.0 + 1e-12, 1.0 - 1e-12, n)
    z = z0 * np.arctanh(u)
    pos = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    # flat rotation curve with a small isotropic dispersion
    disp = 0.05 * vc
    vt = vc + rng.normal(0.0, disp, n)
    vel = np.stack([-vt * np.sin(phi), vt * np.cos(phi), np.zeros(n)], axis=1)
    vel += rng.normal(0.0, disp, (n, 3)) * np.array([1.0, 1.0, 0.5])
    return pos, vel


def _hernquist(n: int, rng: np.random.Generator, a: float, mass: float, G: float = 1.0,
               mass_cut: float = 0.99) -> tuple[np.ndarray, np.ndarray]:
    # M(<r) = M r^2/(r+a)^2  ->  r = a sqrt(X)/(1-sqrt(X))
    x = rng.uniform(1e-12, mass_cut, n)
    s = np.sqrt(x)
    r = a * s / (1.0 - s)
    pos = _isotropic(rng, r)
    # isotropic Gaussian velocities with the local circular-speed scale
    sigma = np.sqrt(G * mass * r / (r + a) ** 2 / np.maximum(r, 1e-9)) / math.sqrt(3.0)
    vel = rng.normal(0.0, 1.0, (n, 3)) * sigma[:, None]
    return pos, vel


def two_galaxy(n_per_galaxy: int, seed: int, separation: float = 10.0, bulge_fraction: float = 0.2,
               density: float = 1.0) -> np.ndarray:
    """Two exponential disks (R_d=1, z0=0.1, v_c=1) with Hernquist bulges (a=0.2).

    Each galaxy has mass 1 (disk 0.8, bulge 0.2 by body count and equal
    masses). The centres start ``separation`` apart with an impact parameter
    of 2 (the offset across the initial relative velocity), on a parabolic
    orbit: relative speed sqrt(2 G M_total / separation), zero orbital
    energy. Total 2*n_per_galaxy bodies.
    """
    rng = np.random.default_rng(seed)
    n_bulge = int(round(n_per_galaxy * bulge_fraction))
    n_disk = n_per_galaxy - n_bulge
    parts = []
    v_rel = math.sqrt(2.0 * 2.0 / separation)
    impact = 2.0
    along = math.sqrt(separation * separation - impact * impact)
    for k, sign in enumerate((-1.0, 1.0)):
        dp, dv = _exp_disk(n_disk, rng, 1.0, 0.1, 1.0)
        bp, bv = _hernquist(n_bulge, rng, 0.2, bulge_fraction)
        pos = np.concatenate([dp, bp])
        vel = np.concatenate([dv, bv])
        # the galaxy's own centre of mass and mean motion (equal masses) at
        # rest at its origin, so that the pair's orbit is the stated one
        pos -= pos.mean(axis=0)
        vel -= vel.mean(axis=0)
        if k == 1:
            # tilt the second disk by 30 degrees about x
            c, s = math.cos(math.pi / 6), math.sin(math.pi / 6)
            rot = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
            pos = pos @ rot.T
            vel = vel @ rot.T
        pos[:, 0] += sign * along / 2.0
        pos[:, 1] += sign * impact / 2.0
        vel[:, 1] += -sign * v_rel / 2.0
        parts.append((pos, vel))
    pos = np.concatenate([p for p, _ in parts])
    vel = np.concatenate([v for _, v in parts])
    n = pos.shape[0]
    return _pack(pos, np.full(n, 1.0 / n_per_galaxy), vel, np.full(n, density))


def ensemble_sizes(n_sims: int, seed: int, lo: int = 4096, hi: int = 524288) -> np.ndarray:
    """Per-simulation body counts drawn log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi + 1), n_sims))).astype(np.int64).clip(lo, hi)


def ensemble(n_sims: int, seed: int) -> list[np.ndarray]:
    """Independent Plummer spheres of uneven N, virial ratio Q in U[0.3, 0.7].

    A Plummer sphere drawn from its distribution function has Q = 0.5; the
    velocities are scaled by sqrt(Q/0.5).
    """
    sizes = ensemble_sizes(n_sims, seed)
    rng = np.random.default_rng(seed + 1)
    q = rng.uniform(0.3, 0.7, n_sims)
    sims = []
    for i, n in enumerate(sizes):
        b = plummer(int(n), seed * 1000 + i)
        b[:, 4:7] *= np.float32(math.sqrt(q[i] / 0.5))
        sims.append(b)
    return sims


def tight_bounds(bodies: np.ndarray) -> np.ndarray:
    """Tight AABB (min xyz, max xyz) of the bodies with non-zero mass, float32."""
    alive = bodies[:, 3] != 0
    p = bodies[alive, 0:3]
    return np.concatenate([p.min(axis=0), p.max(axis=0)]).astype(np.float32)
