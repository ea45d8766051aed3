"""Input wire format and units of the reference's host side.

Mirrors, on the host, what the reference's Java does before the first
dispatch: the unit system (``UnitSet`` / ``Unit``,
simulation/UnitSet.java:11-251), the planet record and its unit conversion
(``Planet``, simulation/Planet.java:9-211), the 32-byte body packing
(``Body.packPlanets``, gpu/Body.java:74-86), the chunked planet stream and its
scene generators (``PlanetGenerator``, simulation/PlanetGenerator.java:14-597)
and the chunked upload into the bodies buffer (``GPU.uploadPlanetsData``,
gpu/GPU.java:1166-1196), here through ``bh_upload`` at a byte offset.

Arithmetic follows the Java: ``float`` fields and ``float`` vector
operations are IEEE binary32 (numpy float32), ``double`` expressions are
binary64, casts round to nearest. The vector operations are JOML 1.10.5's
``Vector3f`` (``pom.xml``; not part of the reference sources), restated from
its published code with fused multiply-add off (JOML's default):
``mul(s)`` multiplies each component by ``s``, ``div(s)`` multiplies by
``1.0f / s``, ``cross`` forms ``y*b.z - z*b.y`` etc., ``normalize`` scales by
``1.0f / (float) sqrt(x*x + (y*y + z*z))``. Nothing in the reference's tests
or fixtures pins these (it has none), so the conversions are checked here by
invariants (round trips, dimensional consistency) only.

``Math.random()`` in the scene generators is unseeded in the reference; here
every generator takes a seeded ``numpy.random.Generator`` and draws the same
distributions in the same order.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass
from typing import Callable, Iterable, Iterator

import numpy as np

F32 = np.float32


def _f(x) -> np.float32:
    """Java's (float) cast of a double: round to nearest."""
    return F32(x)


# ---------------------------------------------------------------- units
SI_PREFIX = {"YOTTA": 24, "ZETTA": 21, "EXA": 18, "PETA": 15, "TERA": 12, "GIGA": 9, "MEGA": 6, "KILO": 3,
             "HECTO": 2, "DEKA": 1, "NONE": 0, "DECI": -1, "CENTI": -2, "MILLI": -3, "MICRO": -6, "NANO": -9,
             "PICO": -12, "FEMTO": -15, "ATTO": -18, "ZEPTO": -21, "YOCTO": -24}

# dimensions as (length, mass, time) exponents (UnitSet.java:20-24)
LENGTH, MASS, TIME, DENSITY, GRAV = (1, 0, 0), (0, 1, 0), (0, 0, 1), (-3, 1, 0), (3, -1, -2)


@dataclass(frozen=True)
class Unit:
    """A unit: value in SI base units, its dimensions, an SI prefix (UnitSet.java:68-153)."""

    value: float
    dimensions: tuple
    prefix: str = "NONE"

    def get_value(self) -> float:
        # value * Math.pow(10, exponent) (UnitSet.java:142-144)
        return self.value * math.pow(10.0, SI_PREFIX[self.prefix])

    def with_prefix(self, prefix: str) -> "Unit":
        return Unit(self.get_value(), self.dimensions, prefix)


ASTRONOMICAL_UNIT = Unit(1.496e11, LENGTH)
STELLAR_RADIUS = Unit(6.96e8, LENGTH)
SOLAR_MASS = Unit(1.989e30, MASS)
STELLAR_DENSITY = Unit(1.408e3, DENSITY)
GRAVITATIONAL_CONSTANT = Unit(6.67430e-11, GRAV)
SECOND = Unit(1.0, TIME)
HOUR = Unit(3600.0, TIME)
KILOGRAM = Unit(1.0, MASS)
METER = Unit(1.0, LENGTH)
KILOGRAM_PER_CUBIC_METER = Unit(1.0, DENSITY)


class UnitSet:
    """Units of body mass, body density, space length and time (UnitSet.java:161-251)."""

    THREE_OVER_FOUR_PI_TO_THE_THIRD = 0.6203504909

    def __init__(self, mass: Unit, density: Unit, length: Unit, time: Unit, name: str = ""):
        for u, d, what in ((mass, MASS, "Body mass must be in kg"), (density, DENSITY, "Body length must be in m"),
                           (length, LENGTH, "Space length must be in m"), (time, TIME, "Time must be in s")):
            if u.dimensions != d:
                raise ValueError(what)
        self._mass, self._density, self._len, self._time = mass, density, length, time
        self.name = name

    def mass(self) -> float:
        return self._mass.get_value()

    def density(self) -> float:
        return self._density.get_value()

    def len(self) -> float:
        return self._len.get_value()

    def time(self) -> float:
        return self._time.get_value()

    def gravitational_constant(self) -> float:
        # G * len^-3 * time^2 * mass^1, left to right in double (UnitSet.java:242-245)
        return (GRAVITATIONAL_CONSTANT.get_value() * math.pow(self.len(), -3) * math.pow(self.time(), 2)
                * math.pow(self.mass(), 1))

    def uniforms(self) -> tuple:
        """The (mass, density, len, time) uniforms the compute programs receive, as floats."""
        return (float(_f(self.mass())), float(_f(self.density())), float(_f(self.len())), float(_f(self.time())))

    def __repr__(self) -> str:
        return f"UnitSet({self.name or 'custom'})"


SOLAR_SYSTEM_SECOND = UnitSet(SOLAR_MASS, STELLAR_DENSITY, ASTRONOMICAL_UNIT, SECOND, "SOLAR_SYSTEM_SECOND")
SOLAR_SYSTEM_HOUR = UnitSet(SOLAR_MASS, STELLAR_DENSITY, ASTRONOMICAL_UNIT, HOUR, "SOLAR_SYSTEM_HOUR")
ASTRONOMICAL = UnitSet(SOLAR_MASS, STELLAR_DENSITY, ASTRONOMICAL_UNIT, SECOND, "ASTRONOMICAL")
METRIC = UnitSet(KILOGRAM, KILOGRAM_PER_CUBIC_METER, METER, SECOND, "METRIC")
GALACTIC_MERGE = UnitSet(SOLAR_MASS.with_prefix("KILO"), STELLAR_DENSITY, ASTRONOMICAL_UNIT.with_prefix("KILO"), HOUR,
                         "GALACTIC_MERGE")


# ---------------------------------------------------------------- JOML Vector3f (binary32)
def _v(x, y, z) -> np.ndarray:
    return np.array([x, y, z], dtype=F32)


def _vmul(a: np.ndarray, s) -> np.ndarray:
    return (a * F32(s)).astype(F32)


def _vdiv(a: np.ndarray, s) -> np.ndarray:
    inv = F32(1.0) / F32(s)
    return (a * inv).astype(F32)


def _vcross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return _v(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _vnormalize(a: np.ndarray) -> np.ndarray:
    l2 = a[0] * a[0] + (a[1] * a[1] + a[2] * a[2])   # fma(x, x, fma(y, y, z * z)) with fma off
    inv = F32(1.0) / _f(math.sqrt(float(l2)))
    return (a * inv).astype(F32)


# ---------------------------------------------------------------- planets
class Planet:
    """One body before packing (simulation/Planet.java:14-211); float fields."""

    __slots__ = ("mass", "position", "velocity", "density", "unit_set")

    def __init__(self, position, velocity, mass, density=1.0, unit_set: UnitSet | None = None):
        self.position = np.asarray(position, dtype=F32).copy()
        self.velocity = np.asarray(velocity, dtype=F32).copy()
        self.mass = _f(mass)
        self.density = _f(density)
        # a planet without units is in SOLAR_SYSTEM_SECOND (Planet.java:95-104)
        self.unit_set = unit_set if unit_set is not None else SOLAR_SYSTEM_SECOND

    @staticmethod
    def from_json_node(node: dict) -> "Planet":
        # Planet.fromJson (Planet.java:141-156): metric, density 100
        pos = [_f(node["position"][k]) for k in range(3)]
        vel = [_f(node["velocity"][k]) for k in range(3)]
        return Planet(pos, vel, _f(node["mass"]), 100.0, METRIC)

    @staticmethod
    def dead_body() -> "Planet":
        return Planet((0, 0, 0), (0, 0, 0), 0.0, 0.0, None)

    def change_unit_set(self, new: UnitSet) -> None:
        """Planet.changeUnitSet (Planet.java:170-187): to SI through the old
        units, then to the new ones; mass and density in double, vectors in
        float."""
        if self.unit_set is new:
            return
        old = self.unit_set
        self.mass = _f(old.mass() * float(self.mass))
        self.density = _f(old.density() * float(self.density))
        self.position = _vmul(self.position, _f(old.len()))
        self.velocity = _vdiv(_vmul(self.velocity, _f(old.len())), _f(old.time()))
        self.mass = _f(float(self.mass) / new.mass())
        self.density = _f(float(self.density) / new.density())
        self.position = _vdiv(self.position, _f(new.len()))
        self.velocity = _vmul(_vdiv(self.velocity, _f(new.len())), _f(new.time()))
        self.unit_set = new

    def __repr__(self) -> str:
        return (f"Planet(mass={self.mass}, position={self.position.tolist()}, velocity={self.velocity.tolist()}, "
                f"density={self.density})")


def pack_planets(planets: Iterable[Planet]) -> np.ndarray:
    """Body.packPlanets (gpu/Body.java:74-86): per body x, y, z, mass,
    vx, vy, vz, density as binary32 -- the (n, 8) float32 layout of the
    bodies buffer."""
    ps = list(planets)
    out = np.empty((len(ps), 8), dtype=F32)
    for i, p in enumerate(ps):
        out[i, 0:3] = p.position
        out[i, 3] = p.mass
        out[i, 4:7] = p.velocity
        out[i, 7] = p.density
    return out


# ---------------------------------------------------------------- the planet stream
DEFAULT_CHUNK_SIZE = 100_000


class PlanetGenerator:
    """A stream of planets handed out in chunks, converted to the stream's
    unit set as they leave it (simulation/PlanetGenerator.java:14-308)."""

    def __init__(self, source: Callable[[int], Planet] | None = None, num_planets: int = 0,
                 chunk_size: int = DEFAULT_CHUNK_SIZE, unit_set: UnitSet | None = None):
        self._source = source
        self.num_planets = int(num_planets)
        self.chunk_size = int(chunk_size)
        self.planets_generated = 0
        self.unit_set = unit_set if unit_set is not None else SOLAR_SYSTEM_SECOND

    # -- constructors of the reference
    @classmethod
    def from_list(cls, planets: list[Planet], unit_set: UnitSet | None = None) -> "PlanetGenerator":
        ps = list(planets)
        return cls(lambda i: ps[i], len(ps), unit_set=unit_set)

    @classmethod
    def single(cls, planet: Planet) -> "PlanetGenerator":
        return cls(lambda i: planet, 1)

    @classmethod
    def concat(cls, first: "PlanetGenerator", second: "PlanetGenerator") -> "PlanetGenerator":
        """PlanetGenerator(pg1, pg2): pg1's planets, then pg2's, in pg1's units."""
        n1 = first.num_planets

        def src(i: int) -> Planet:
            return first.next_planet() if i < n1 else second.next_planet()

        return cls(src, n1 + second.num_planets, unit_set=first.unit_set)

    def add(self, other: "PlanetGenerator") -> None:
        prev, prev_n = self._source, self.num_planets

        def src(i: int) -> Planet:
            return prev(i) if i < prev_n else other.next_planet()

        self._source = src
        self.num_planets = prev_n + other.num_planets

    def change_unit_set(self, unit_set: UnitSet) -> None:
        self.unit_set = unit_set

    # -- iteration
    def has_next(self) -> bool:
        return self.planets_generated < self.num_planets

    def next_planet(self) -> Planet:
        if not self.has_next():
            raise RuntimeError("No more planets")
        i = self.planets_generated
        self.planets_generated += 1
        p = self._source(i)
        if p.unit_set is not self.unit_set:
            p.change_unit_set(self.unit_set)
        return p

    def next_chunk(self) -> list[Planet]:
        if not self.has_next():
            raise RuntimeError("No more planets")
        k = min(self.chunk_size, self.num_planets - self.planets_generated)
        return [self.next_planet() for _ in range(k)]

    def chunks(self) -> Iterator[np.ndarray]:
        """Packed chunks, as GPU.uploadPlanetsData packs them."""
        while self.has_next():
            yield pack_planets(self.next_chunk())

    def pack_all(self) -> np.ndarray:
        parts = list(self.chunks())
        return np.concatenate(parts) if parts else np.empty((0, 8), dtype=F32)


def upload_planets(bh, generator: PlanetGenerator) -> int:
    """GPU.uploadPlanetsData (gpu/GPU.java:1166-1196): every chunk of the
    stream into the bodies buffer at the running byte offset, through
    ``bh_upload``. Returns the number of bodies written."""
    from gravity_amd.barnes_hut import BODIES_IN
    offset = 0
    count = 0
    for chunk in generator.chunks():
        bh.upload(BODIES_IN, chunk, offset)
        offset += chunk.nbytes
        count += chunk.shape[0]
    return count


# ---------------------------------------------------------------- scenes
def from_json(text: str) -> PlanetGenerator:
    """PlanetGenerator.fromJson (PlanetGenerator.java:542-556): the "bodies"
    array of a planet_data file, metric units."""
    doc = json.loads(text)
    return PlanetGenerator.from_list([Planet.from_json_node(b) for b in doc["bodies"]], METRIC)


def make_solar_system(path: str) -> PlanetGenerator:
    """The "Solar System" scene (SimulationSetup.java:227-233): solar_system.json
    converted to SOLAR_SYSTEM_HOUR."""
    with open(path) as f:
        pg = from_json(f.read())
    pg.change_unit_set(SOLAR_SYSTEM_HOUR)
    return pg


def _rand_range(rng: np.random.Generator, lo_hi, density: float = 1.0) -> np.float32:
    # (float)(Math.pow(Math.random(), density) * (range[1] - range[0]) + range[0]):
    # the range difference is a float subtraction, the rest double
    lo, hi = F32(lo_hi[0]), F32(lo_hi[1])
    return _f(math.pow(rng.random(), density) * float(hi - lo) + float(lo))


def _rand_int(rng: np.random.Generator, lo_hi) -> int:
    return int(_rand_range(rng, lo_hi))


def make_new_random_box(num: int, x, y, z, xv, yv, zv, m, density, rng: np.random.Generator) -> PlanetGenerator:
    """PlanetGenerator.makeNewRandomBox (PlanetGenerator.java:310-322): every
    quantity uniform in its range, drawn in the order x, y, z, vx, vy, vz, m, density."""

    def src(i: int) -> Planet:
        vals = [_rand_range(rng, r) for r in (x, y, z, xv, yv, zv, m, density)]
        return Planet(vals[0:3], vals[3:6], vals[6], vals[7])

    return PlanetGenerator(src, num)


def _adherence(rng: np.random.Generator, adherence_to_plane: float) -> np.float32:
    # (float)(PI * (1 - Math.random()^(1 - adherence))) (PlanetGenerator.java:328-332)
    factor = 1.0 - math.pow(rng.random(), 1.0 - float(F32(adherence_to_plane)))
    return _f(math.pi * factor)


def _polar_to_planar(r, theta, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    # u * (r cos theta) + v * (r sin theta), the factors as floats (PlanetGenerator.java:342-348)
    cu = _vmul(u, F32(r) * _f(math.cos(float(F32(theta)))))
    cv = _vmul(v, F32(r) * _f(math.sin(float(F32(theta)))))
    return (cu + cv).astype(F32)


def make_new_random_disk(num: int, radius, mass, density, center, relative_velocity, phi: float, center_mass: float,
                         center_density: float, adherence_to_plane: float, orbital_factor: float, ccw: bool,
                         give_orbital_velocity: bool, unit_set: UnitSet, rng: np.random.Generator
                         ) -> PlanetGenerator:
    """PlanetGenerator.makeNewRandomDisk (PlanetGenerator.java:418-492): num
    planets in a disk tilted by phi about x, then the centre body; the stream
    is in SOLAR_SYSTEM_SECOND (its planets are converted from ``unit_set``),
    the centre body has no units of its own."""
    phi = F32(phi)
    normal = _v(0.0, _f(math.sin(float(phi))), _f(math.cos(float(phi))))
    a = _v(1, 0, 0) if abs(normal[0]) < F32(0.9) else _v(0, 1, 0)
    u = _vnormalize(_vcross(normal, a))
    v = _vcross(normal, u)
    radius = [F32(radius[0]), F32(radius[1])]
    mass = [F32(mass[0]), F32(mass[1])]
    center = _v(*center)
    rel = _v(*relative_velocity)
    G = unit_set.gravitational_constant()

    def src(i: int) -> Planet:
        r = _rand_range(rng, radius, 1.0)
        theta = _f(rng.random() * 2.0 * math.pi)
        dev = _adherence(rng, adherence_to_plane)
        above = rng.random() < 0.5
        dev = dev if above else -dev
        planar = _polar_to_planar(1.0, theta, u, v)
        with_dev = _polar_to_planar(1.0, dev, planar, normal)
        new_normal = _polar_to_planar(1.0, dev, normal, planar)
        approx_mass = ((mass[1] + mass[0]) / F32(2) * F32(num)
                       * _f(math.pow(float(r / radius[1]), float(F32(center_density) + F32(1)))))
        position = _vmul(with_dev, r)
        velocity = _v(0, 0, 0)
        if give_orbital_velocity:
            direction = 1 if ccw else -1
            speed = _f(math.sqrt(float(F32(center_mass) + approx_mass) * G / float(r))) * F32(orbital_factor)
            # Vector3f.mul scales in place, so the tangent is taken against
            # the already scaled direction (PlanetGenerator.java:458-467)
            tangent = _vnormalize(_vcross(new_normal, position))
            velocity = _vmul(tangent, speed * F32(direction))
        position = (position + center).astype(F32)
        velocity = (velocity + rel).astype(F32)
        return Planet(position, velocity, _rand_range(rng, mass), _rand_range(rng, density), unit_set)

    disk = PlanetGenerator(src, num)
    centre = PlanetGenerator.single(Planet(center, rel, center_mass, center_density))
    return PlanetGenerator.concat(disk, centre)


def create_several_disks(num_disks: int, num_planets_range, radius_range_low, stellar_density_range, m_range,
                         density_range, cx_range, cy_range, cz_range, rvx, rvy, rvz, phi_range, center_mass_range,
                         center_density_range, adherence_range, orbital_factor: float, give_orbital_velocity: bool,
                         unit_set: UnitSet, rng: np.random.Generator) -> PlanetGenerator:
    """PlanetGenerator.createSeveralDisks (PlanetGenerator.java:372-416)."""
    pg = PlanetGenerator(lambda i: None, 0)
    for _ in range(num_disks):
        num = _rand_int(rng, num_planets_range)
        radius = [_rand_range(rng, radius_range_low), F32(num) / _rand_range(rng, stellar_density_range)]
        mass = [_rand_range(rng, m_range), _rand_range(rng, m_range)]
        density = [_rand_range(rng, density_range), _rand_range(rng, density_range)]
        center = [_rand_range(rng, cx_range), _rand_range(rng, cy_range), _rand_range(rng, cz_range)]
        rel = [_rand_range(rng, rvx), _rand_range(rng, rvy), _rand_range(rng, rvz)]
        phi = _rand_range(rng, phi_range)
        cm = _rand_range(rng, center_mass_range)
        cd = _rand_range(rng, center_density_range)
        adh = _rand_range(rng, adherence_range)
        disk = make_new_random_disk(num, radius, mass, density, center, rel, phi, cm, cd, adh, orbital_factor, False,
                                    give_orbital_velocity, unit_set, rng)
        pg = PlanetGenerator.concat(pg, disk)
    return pg


def make_new_in_orbit(radius, mass, density, center: Planet, rng: np.random.Generator) -> Planet:
    """PlanetGenerator.makeNewInOrbit (PlanetGenerator.java:494-510): a body on
    a circular orbit in the z = 0 plane at 1.1 x the circular speed."""
    r = _f(rng.random() * float(F32(radius[1]) - F32(radius[0])) + float(F32(radius[0])))
    speed = _f(1.1 * math.sqrt(float(center.mass / r)))            # a float division
    theta = _f(rng.random() * 2.0 * math.pi)
    phi = _f(math.pi / 2.0)
    st, ct, sp, cp = (math.sin(float(theta)), math.cos(float(theta)), math.sin(float(phi)), math.cos(float(phi)))
    pos = [_f(float(r) * ct * sp), _f(float(r) * st * sp), _f(float(r) * cp)]
    vel = [_f(-float(speed) * st * sp), _f(float(speed) * ct * sp), _f(-float(speed) * cp)]
    return Planet(pos, vel, _rand_range(rng, mass), _rand_range(rng, density))


def suggested_solar_dt(unit_set: UnitSet) -> float:
    """The solar-only scene's time step: a hundredth of 2 pi sqrt(1/G)
    (SimulationSetup.java:289-293)."""
    period = 2.0 * math.pi * math.sqrt(1.0 / unit_set.gravitational_constant())
    return float(_f(min(max(period / 100.0, 1e-6), 1e12)))
