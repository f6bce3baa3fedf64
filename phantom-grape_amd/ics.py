"""Seeded synthetic initial conditions for the benchmark configurations.

* ``plummer``   Plummer sphere in Henon (virial) units, G = M = 1,
                E = -1/4, equal masses 1/N, cut at ``rcut`` scale radii
                (Aarseth, Henon & Wielen 1974 sampling), centre-of-mass frame.
* ``hernquist`` Hernquist halo, a = 1, M = 1, truncated at ``rtrunc`` a,
                inverse-CDF radii, isotropic directions (positions only).
* ``clumpy_box`` periodic unit box: half uniform, half in Gaussian clumps.
* ``cluster_sizes`` log-uniform cluster sizes for the many-cluster workload.

All generators take an explicit seed and use ``numpy.random.default_rng``.
"""
from __future__ import annotations

import numpy as np


def _iso(rng, n):
    mu = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(1.0 - mu * mu)
    return np.stack([s * np.cos(phi), s * np.sin(phi), mu], axis=1)


def plummer(n: int, seed: int, rcut: float = 20.0):
    """Return (m, x, v) with x, v of shape (n, 3) in Henon units."""
    rng = np.random.default_rng(seed)
    r = np.empty(n)
    filled = 0
    while filled < n:
        k = n - filled
        x1 = rng.uniform(0.0, 1.0, 2 * k + 16)
        x1 = x1[x1 > 0]
        rr = 1.0 / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr <= rcut][:k]
        r[filled:filled + len(rr)] = rr
        filled += len(rr)
    # speed in units of the local escape speed: q^2 (1 - q^2)^(7/2) rejection
    q = np.empty(n)
    filled = 0
    while filled < n:
        k = n - filled
        x4 = rng.uniform(0.0, 1.0, 3 * k + 16)
        x5 = rng.uniform(0.0, 0.1, 3 * k + 16)
        ok = x5 < x4 * x4 * (1.0 - x4 * x4) ** 3.5
        sel = x4[ok][:k]
        q[filled:filled + len(sel)] = sel
        filled += len(sel)
    vesc = np.sqrt(2.0) * (1.0 + r * r) ** (-0.25)
    x = _iso(rng, n) * r[:, None]
    v = _iso(rng, n) * (q * vesc)[:, None]
    m = np.full(n, 1.0 / n)
    x -= (m[:, None] * x).sum(0)
    v -= (m[:, None] * v).sum(0)
    # structural units (a = 1) -> Henon units (E = -1/4)
    sx = 3.0 * np.pi / 16.0
    x *= sx
    v /= np.sqrt(sx)
    return m, np.ascontiguousarray(x), np.ascontiguousarray(v)


def hernquist(n: int, seed: int, rtrunc: float = 50.0):
    """Return (m, x) for a Hernquist halo with a = M = 1 truncated at rtrunc."""
    rng = np.random.default_rng(seed)
    mmax = rtrunc * rtrunc / (rtrunc + 1.0) ** 2
    u = rng.uniform(0.0, mmax, n)
    su = np.sqrt(u)
    r = su / (1.0 - su)
    x = _iso(rng, n) * r[:, None]
    m = np.full(n, 1.0 / n)
    return m, np.ascontiguousarray(x)


def clumpy_box(n: int, seed: int, nclump: int = 64, sigma: float = 0.01):
    """Periodic unit box: n/2 uniform + n/2 in nclump Gaussian clumps, wrapped."""
    rng = np.random.default_rng(seed)
    nu = n // 2
    nc = n - nu
    xu = rng.uniform(0.0, 1.0, (nu, 3))
    centres = rng.uniform(0.0, 1.0, (nclump, 3))
    which = rng.integers(0, nclump, nc)
    xc = centres[which] + rng.normal(0.0, sigma, (nc, 3))
    x = np.concatenate([xu, xc]) % 1.0
    m = np.full(n, 1.0 / n)
    return m, np.ascontiguousarray(x)


def cluster_sizes(ncl: int, seed: int, nmin: int = 64, nmax: int = 4096):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), ncl))).astype(np.int64)
