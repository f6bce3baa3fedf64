"""Particle-mesh long-range force of the TreePM split — TEST INFRASTRUCTURE ONLY.

The reference ships only the short-range side of TreePM: the table force
f(r) = R(r, eps_PP) - R(r, eps_PM) (PG5/table/libpg5/pg5_table.c:8-34), with
R(r, a) the force between two S2 spheres of diameter a (Hockney & Eastwood,
"Computer Simulation Using Particles", sec. 8-3), and the mesh parameters of
PG5/table/libpg5/pppm.h:4-6.  The long-range side is therefore fixed by the
mathematics: the periodic force of R(r, eps_PM).  This module restates it two
ways, in FP64 numpy:

* pm_force  - the mesh method the GPU path implements (TSC assignment on an
  M^3 mesh, FFT, Green's function -4 pi / k^2 S(k a)^2 / W(k)^2 with the TSC
  window W, spectral gradient, TSC interpolation);
* pm_exact  - the same force without a mesh: the Ewald sum of point masses
  minus the short-range difference 1/r^3 - R(r, eps_PM) summed over nearby
  images.

Pinning: ewald_force against a lattice (zero force by symmetry) and the
Newtonian limit at small separation; R against its S2 closed form
(or_s2_reff); pm_force against pm_exact (the mesh's discretisation error).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc


def reff(r, a):
    """R(r, a): force per unit mass and unit separation between two S2
    spheres of diameter a (pg5_table.c:8-22), vectorised."""
    r = np.asarray(r, dtype=np.float64)
    eta = 2.0 * r / a
    out = np.empty_like(r)
    m1 = eta < 1.0
    m2 = (eta >= 1.0) & (eta < 2.0)
    m3 = eta >= 2.0
    e = eta[m1]
    out[m1] = (224. + e * e * (-224. + e * (70. + e * (48. - e * 21.)))) * 2.0 / (35.0 * a ** 3)
    e = eta[m2]
    out[m2] = ((12. / (e * e) - 224. + e * (896. + e * (-840. + e * (224. + e * (70. + e * (-48. + e * 7.))))))
               * 2.0 / (35.0 * e * a ** 3))
    out[m3] = 1.0 / r[m3] ** 3
    return out


def s2_form(k, a):
    """Fourier transform of an S2 sphere of diameter a (unit mass)."""
    u = np.asarray(k, dtype=np.float64) * a / 2.0
    out = np.ones_like(u)
    big = u > 1e-3
    ub = u[big]
    out[big] = 12.0 / ub ** 4 * (2.0 - 2.0 * np.cos(ub) - ub * np.sin(ub))
    us = u[~big]
    out[~big] = 1.0 - us * us / 15.0
    return out


def pm_green(mesh, a):
    """-4 pi S(k a)^2 / (k^2 W(k)^2) on the rfft grid of an M^3 mesh of the
    unit box (0 at k = 0), and the wave vectors."""
    n = np.fft.fftfreq(mesh, d=1.0 / mesh)
    nz = np.arange(mesh // 2 + 1)
    kx, ky, kz = np.meshgrid(2 * np.pi * n, 2 * np.pi * n, 2 * np.pi * nz, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    h = 1.0 / mesh

    def sinc(x):
        return np.where(np.abs(x) < 1e-12, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))
    w = (sinc(kx * h / 2) * sinc(ky * h / 2) * sinc(kz * h / 2)) ** 3
    s = s2_form(np.sqrt(k2), a)
    g = np.zeros_like(k2)
    nzero = k2 > 0
    g[nzero] = -4.0 * np.pi * s[nzero] ** 2 / (k2[nzero] * w[nzero] ** 2)
    # the gradient of a Nyquist mode has no real counterpart: dropped
    nyq = mesh // 2
    g[nyq, :, :] = 0.0
    g[:, nyq, :] = 0.0
    g[:, :, nyq] = 0.0
    return g, kx, ky, kz


def tsc_weights(u):
    """Nearest mesh point and the three TSC weights along one axis (u in
    mesh units, mesh points at integers)."""
    g = np.floor(u + 0.5)
    d = u - g
    w = np.stack([0.5 * (0.5 - d) ** 2, 0.75 - d * d, 0.5 * (0.5 + d) ** 2], axis=-1)
    return g.astype(np.int64), w


def pm_force(x, m, mesh, a):
    """Long-range accelerations of particles at x (unit periodic box) by the
    mesh method: TSC mass assignment, FFT, Green's function, spectral
    gradient, TSC interpolation.  FP64."""
    x = np.asarray(x, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    h = 1.0 / mesh
    rho = np.zeros((mesh, mesh, mesh))
    gi, wi = [], []
    for d in range(3):
        g, w = tsc_weights(x[:, d] / h)
        gi.append(g)
        wi.append(w)
    for a_ in range(3):
        for b_ in range(3):
            for c_ in range(3):
                idx = ((gi[0] + a_ - 1) % mesh, (gi[1] + b_ - 1) % mesh, (gi[2] + c_ - 1) % mesh)
                np.add.at(rho, idx, m * wi[0][:, a_] * wi[1][:, b_] * wi[2][:, c_] / h ** 3)
    rk = np.fft.rfftn(rho)
    g, kx, ky, kz = pm_green(mesh, a)
    phik = rk * g
    acc = np.zeros((len(m), 3))
    for d, k in enumerate((kx, ky, kz)):
        f = np.fft.irfftn(-1j * k * phik, s=(mesh, mesh, mesh), axes=(0, 1, 2))
        for a_ in range(3):
            for b_ in range(3):
                for c_ in range(3):
                    idx = ((gi[0] + a_ - 1) % mesh, (gi[1] + b_ - 1) % mesh,
                           (gi[2] + c_ - 1) % mesh)
                    acc[:, d] += f[idx] * wi[0][:, a_] * wi[1][:, b_] * wi[2][:, c_]
    return acc


def ewald_force(x, m, alpha=None, nreal=2, kmax=6):
    """Periodic Newtonian accelerations (unit box, G = 1, uniform neutralising
    background) by Ewald summation.  O(N^2 (nreal^3 + kmax^3)); small N."""
    x = np.asarray(x, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    alpha = alpha or 2.0
    n = len(m)
    acc = np.zeros((n, 3))
    shifts = np.array([(i, j, k) for i in range(-nreal, nreal + 1)
                       for j in range(-nreal, nreal + 1) for k in range(-nreal, nreal + 1)],
                      dtype=np.float64)
    kv = np.array([(i, j, k) for i in range(-kmax, kmax + 1) for j in range(-kmax, kmax + 1)
                   for k in range(-kmax, kmax + 1) if (i, j, k) != (0, 0, 0)], dtype=np.float64)
    kv *= 2 * np.pi
    k2 = (kv * kv).sum(1)
    fk = 4 * np.pi * np.exp(-k2 / (4 * alpha * alpha)) / k2
    for i in range(n):
        d = x - x[i]  # vector from i to j
        # real space
        dd = d[:, None, :] + shifts[None, :, :]
        r = np.sqrt((dd * dd).sum(-1))
        with np.errstate(divide="ignore", invalid="ignore"):
            f = (erfc(alpha * r) + 2 * alpha / math.sqrt(math.pi) * r * np.exp(-(alpha * r) ** 2)) / r ** 3
        f[~np.isfinite(f)] = 0.0
        acc[i] += (m[:, None, None] * f[:, :, None] * dd).sum((0, 1))
        # reciprocal space
        kd = d @ kv.T  # [n, nk]
        acc[i] += (m[:, None, None] * (fk * np.sin(kd))[:, :, None] * kv[None, :, :]).sum((0, 1))
    return acc


def pm_exact(x, m, a, nimg=1):
    """The long-range force of R(r, a) without a mesh: Ewald (point masses)
    minus sum over nearby images of m d (1/r^3 - R(r, a)) for r < a."""
    x = np.asarray(x, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    acc = ewald_force(x, m)
    shifts = np.array([(i, j, k) for i in range(-nimg, nimg + 1)
                       for j in range(-nimg, nimg + 1) for k in range(-nimg, nimg + 1)],
                      dtype=np.float64)
    for i in range(len(m)):
        dd = (x - x[i])[:, None, :] + shifts[None, :, :]
        r = np.sqrt((dd * dd).sum(-1))
        near = (r < a) & (r > 0)
        corr = np.zeros_like(r)
        corr[near] = 1.0 / r[near] ** 3 - reff(r[near], a)
        acc[i] -= (m[:, None, None] * corr[:, :, None] * dd).sum((0, 1))
    return acc
