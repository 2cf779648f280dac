"""Synthetic workloads for the benchmark and the serving loop (numpy, seeded).

* ``plummer`` -- the reference's own seeded sample
  (initial_conditions/plummer.cpp:91-128 over core/random.cpp:12-47): the
  engine's host-side ``orr_plummer_sample`` draws the same mt19937_64 stream
  in the same order, so seed s gives the particles the reference's CPU path
  gives for seed s, bit for bit. Radii by the inverse mass CDF
  a / sqrt(u^(-2/3) - 1) truncated at a mass fraction, speeds q v_esc with q
  drawn by rejection under the envelope 0.1 on q^2 (1 - q^2)^(7/2), isotropic
  directions, then the centre-of-mass frame.
* ``disc_collision`` -- two exponential discs (sech^2 vertical profile,
  truncated at ``truncation`` scale lengths) with Hernquist bulges, circular
  speeds from the enclosed mass plus Gaussian dispersion, a fraction of the
  disc particles inclined, the pair on a parabolic orbit of given pericentre
  and initial separation.
* ``batched_spheres`` -- the uneven batch: sizes log-uniform in [lo, hi]
  (numpy), each sphere the reference's seeded Plummer sample.

The disc model is this project's own (the two-disc configuration differs from
the reference's galaxy generator) and draws from numpy's PCG64.

These are workload generators only: the engine consumes whatever positions and
masses it is given.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi as C

PLUMMER_CUTOFF = (100.0 / 101.0) ** 1.5  # r < 10 a for a = 1

DISC_DEFAULTS = dict(mass_ratio=3.0, scale_length=(1.0, 0.7), scale_height=0.1, truncation=5.0,
                     bulge_fraction=0.2, bulge_scale=0.2, dispersion=0.1, pericentre=2.0,
                     separation=20.0, inclined_fraction=0.3, inclination=np.deg2rad(30.0),
                     total_mass=1.0)


@dataclass
class Workload:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    m: np.ndarray

    @property
    def n(self) -> int:
        return len(self.m)

    def arrays(self):
        return self.x, self.y, self.z, self.vx, self.vy, self.vz, self.m


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(int(seed)))


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    h = rng.uniform(-1.0, 1.0, n)
    az = rng.uniform(0.0, 2.0 * np.pi, n)
    r = np.sqrt(1.0 - h * h)
    return np.stack([r * np.cos(az), r * np.sin(az), h])


def _com_frame(pos: np.ndarray, vel: np.ndarray, m: np.ndarray) -> Workload:
    mt = m.sum()
    pos = pos - (pos * m).sum(axis=1, keepdims=True) / mt
    vel = vel - (vel * m).sum(axis=1, keepdims=True) / mt
    return Workload(*(np.ascontiguousarray(a) for a in (*pos, *vel)), np.ascontiguousarray(m))


def plummer(n: int, seed: int, total_mass: float = 1.0, scale_radius: float = 1.0,
            cutoff: float = PLUMMER_CUTOFF) -> Workload:
    """The reference's Plummer sphere for ``seed`` (a = 1, r < 10 a by default)."""
    n = int(n)
    arrays = [np.zeros(n) for _ in range(7)]
    ptr = ctypes.POINTER(ctypes.c_double)
    C.check(C.load().orr_plummer_sample(n, float(total_mass), float(scale_radius), float(cutoff),
                                        int(seed), *(a.ctypes.data_as(ptr) for a in arrays)))
    return Workload(*arrays)


class _Disc:
    def __init__(self, mass, rd, zd, rmax, bulge_mass, bulge_a):
        self.mass, self.rd, self.zd, self.rmax = mass, rd, zd, rmax
        self.bulge_mass, self.bulge_a = bulge_mass, bulge_a
        self.norm = 1.0 - (1.0 + rmax / rd) * np.exp(-rmax / rd)

    def disc_fraction(self, R):
        u = R / self.rd
        return np.minimum(1.0, (1.0 - (1.0 + u) * np.exp(-u)) / self.norm)

    def bulge_fraction(self, r):
        return np.minimum(1.0, r * r / ((r + self.bulge_a) ** 2))

    def circular_speed(self, r):
        enclosed = (self.mass - self.bulge_mass) * self.disc_fraction(r) + \
            self.bulge_mass * self.bulge_fraction(r)
        return np.where(r > 0, np.sqrt(enclosed / np.where(r > 0, r, 1.0)), 0.0)

    def disc_radius(self, u):
        lo, hi = np.zeros_like(u), np.full_like(u, self.rmax)
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            below = self.disc_fraction(mid) < u
            lo = np.where(below, mid, lo)
            hi = np.where(below, hi, mid)
        return 0.5 * (lo + hi)


def _galaxy(rng, count, g: _Disc, bulge_fraction, dispersion, inclined_fraction, inclination):
    nb = int(round(bulge_fraction * count))
    nd = count - nb
    u = rng.uniform(0.0, 0.99, nb)
    su = np.sqrt(u)
    rb = g.bulge_a * su / (1.0 - su)
    pos_b = rb * _unit_vectors(rng, nb)
    vel_b = (g.circular_speed(rb) / np.sqrt(3.0)) * rng.standard_normal((3, nb))
    R = g.disc_radius(rng.uniform(size=nd))
    phi = rng.uniform(0.0, 2.0 * np.pi, nd)
    w = np.clip(rng.uniform(-1.0, 1.0, nd), -1 + 1e-15, 1 - 1e-15)
    zz = g.zd * np.arctanh(w)
    vc = g.circular_speed(R)
    nrm = rng.standard_normal((3, nd))
    c, s = np.cos(phi), np.sin(phi)
    pos_d = np.stack([R * c, R * s, zz])
    vel_d = np.stack([-vc * s + dispersion * vc * nrm[0], vc * c + dispersion * vc * nrm[1],
                      dispersion * vc * nrm[2]])
    tilt = rng.uniform(size=nd) < inclined_fraction
    ci, si = np.cos(inclination), np.sin(inclination)
    for v in (pos_d, vel_d):
        y, z = v[1].copy(), v[2].copy()
        v[1] = np.where(tilt, ci * y - si * z, y)
        v[2] = np.where(tilt, si * y + ci * z, z)
    return np.concatenate([pos_b, pos_d], axis=1), np.concatenate([vel_b, vel_d], axis=1)


def disc_collision(n: int, seed: int, **overrides) -> Workload:
    cfg = dict(DISC_DEFAULTS)
    cfg.update(overrides)
    n = int(n)
    rng = _rng(seed)
    Mt = float(cfg["total_mass"])
    M1 = Mt * cfg["mass_ratio"] / (1.0 + cfg["mass_ratio"])
    M2 = Mt - M1
    n1 = int(round(n * M1 / Mt))
    n2 = n - n1
    parts = []
    for count, M, rd in ((n1, M1, cfg["scale_length"][0]), (n2, M2, cfg["scale_length"][1])):
        g = _Disc(M, rd, cfg["scale_height"], cfg["truncation"] * rd, cfg["bulge_fraction"] * M,
                  cfg["bulge_scale"])
        pos, vel = _galaxy(rng, count, g, cfg["bulge_fraction"], cfg["dispersion"],
                           cfg["inclined_fraction"], cfg["inclination"])
        parts.append((pos, vel, np.full(count, M / max(count, 1))))
    d, q = float(cfg["separation"]), float(cfg["pericentre"])
    v = np.sqrt(2.0 * Mt / d)
    vt = np.sqrt(2.0 * Mt * q) / d
    vr = -np.sqrt(max(0.0, v * v - vt * vt))
    rrel = np.array([[d], [0.0], [0.0]])
    vrel = np.array([[vr], [vt], [0.0]])
    (p1, v1, m1), (p2, v2, m2) = parts
    pos = np.concatenate([p1 - (M2 / Mt) * rrel, p2 + (M1 / Mt) * rrel], axis=1)
    vel = np.concatenate([v1 - (M2 / Mt) * vrel, v2 + (M1 / Mt) * vrel], axis=1)
    return _com_frame(pos, vel, np.concatenate([m1, m2]))


def batched_sizes(count: int, lo: int, hi: int, seed: int) -> list[int]:
    """Sizes log-uniform in [lo, hi] (rounded down, at least lo)."""
    rng = _rng(seed)
    return [max(int(lo), int(v)) for v in np.exp(rng.uniform(np.log(lo), np.log(hi), int(count)))]


def batched_spheres(count: int, lo: int, hi: int, seed: int) -> list[Workload]:
    """Independent Plummer spheres (M = 1, a = 1), system k seeded seed * 100003 + k."""
    return [plummer(n, seed * 100003 + k) for k, n in enumerate(batched_sizes(count, lo, hi, seed))]
