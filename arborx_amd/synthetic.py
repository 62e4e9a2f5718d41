"""Seeded synthetic inputs for the BVH workloads (numpy, host side).

Counter-based: value i of stream `seed` is splitmix64(seed * 2^32 + i), so
every consumer (GPU path, CPU oracle, tests, benchmark) that asks for the
same (seed, n) gets bit-identical float32 arrays.  Uniform floats are
(u >> 40) * 2^-24 in [0, 1); normals use Box-Muller in float64 and are then
rounded to float32.
"""
import math

import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n uint64 values of counter stream `seed` starting at counter `offset`."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) << np.uint64(32)) + np.arange(offset, offset + n, dtype=np.uint64)
        z = z * _M1 + _M1
        z = (z ^ (z >> np.uint64(30))) * _M2
        z = (z ^ (z >> np.uint64(27))) * _M3
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n, offset=0):
    """float32 in [0, 1) with 24 random bits."""
    u = splitmix64(seed, n, offset) >> np.uint64(40)
    return (u.astype(np.float64) * 2.0 ** -24).astype(np.float32)


def _uniform01_f64(seed, n, offset=0):
    u = splitmix64(seed, n, offset) >> np.uint64(11)
    return (u.astype(np.float64) + 0.5) * 2.0 ** -53  # in (0, 1)


def normal(seed, n, offset=0):
    """float64 standard normals (Box-Muller)."""
    m = (n + 1) // 2
    u1 = _uniform01_f64(seed, m, 2 * offset)
    u2 = _uniform01_f64(seed, m, 2 * offset + m)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def uniform_points(n, seed):
    """n points i.i.d. uniform in [0,1)^3, float32 (n, 3)."""
    return uniform01(seed, 3 * n).reshape(n, 3)


def gaussian_mixture_points(n, seed, n_centres=1000, sigma=0.01, centre_seed=None):
    """n points from an isotropic Gaussian mixture clamped to [0,1]^3.

    Centres are uniform in the cube (stream `centre_seed`, default seed+1000);
    point i belongs to centre i % n_centres.
    """
    cs = seed + 1000 if centre_seed is None else centre_seed
    centres = uniform01(cs, 3 * n_centres).astype(np.float64).reshape(n_centres, 3)
    z = normal(seed, 3 * n).reshape(n, 3)
    p = centres[np.arange(n) % n_centres] + sigma * z
    return np.clip(p, 0.0, 1.0).astype(np.float32)


def lognormal_boxes(n, seed, mu=math.log(0.002), sigma=0.5):
    """n boxes (n, 6) {lo.xyz, hi.xyz}: centres uniform in [0,1)^3 and
    per-axis side lengths lognormal(mu, sigma)."""
    c = uniform01(seed, 3 * n).astype(np.float64).reshape(n, 3)
    side = np.exp(mu + sigma * normal(seed + 500, 3 * n).reshape(n, 3))
    lo = (c - 0.5 * side).astype(np.float32)
    hi = (c + 0.5 * side).astype(np.float32)
    return np.concatenate([lo, hi], axis=1)


def rays(n, seed):
    """n rays as (origins, directions), float32 (n, 3) each: origins uniform
    in [-0.2, 1.2)^3, Gaussian directions (not normalised), every fifth ray
    along a coordinate axis (the two other components zero)."""
    o = (uniform01(seed, 3 * n).reshape(n, 3) * np.float32(1.4) - np.float32(0.2))
    d = normal(seed + 700, 3 * n).reshape(n, 3).astype(np.float32)
    along = np.arange(n) % 5 == 1
    axis = (np.arange(n) // 5) % 3
    keep = d[along, axis[along]].copy()
    d[along] = 0.0
    d[along, axis[along]] = keep
    return o.astype(np.float32), d


def fcc_lattice(cells, seed, a=1.0, jitter=0.05):
    """FCC lattice with nearest-neighbour distance `a` (cubic cell edge a*sqrt(2)),
    `cells` = (nx, ny, nz) cubic cells, 4 atoms per cell, each coordinate jittered
    by U(-jitter*a, +jitter*a).  float32 (4*nx*ny*nz, 3)."""
    nx, ny, nz = cells
    edge = a * math.sqrt(2.0)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], np.float64)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cell = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
    p = (cell[:, None, :] + basis[None, :, :]).reshape(-1, 3) * edge
    n = p.shape[0]
    u = uniform01(seed, 3 * n).astype(np.float64).reshape(n, 3)
    p = p + (2.0 * u - 1.0) * jitter * a
    return p.astype(np.float32)


def radius_for_neighbours(n, k, volume=1.0):
    """Radius whose ball holds k expected points of n uniform in `volume`
    (self included): r = (3 k V / (4 pi n))^(1/3)."""
    return (3.0 * k * volume / (4.0 * math.pi * n)) ** (1.0 / 3.0)
