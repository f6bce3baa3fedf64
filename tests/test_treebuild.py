"""The host list builder (libpgtree.so) against the definitions of its
outputs, by brute force on small inputs (CPU only)."""
import numpy as np

import ics
import pgamd


def test_treepm_ghosts_are_exact_periodic_images():
    rcut = 3.0 / 32
    m, x = ics.clumpy_box(1 << 13, 13)
    L = pgamd.treepm_lists(x, rcut, 64)
    n = len(x)
    P, src = L["pos"], L["src"]
    np.testing.assert_array_equal(P[:n], x[src[:n]])
    shift = np.round(P[n:] - x[src[n:]])
    np.testing.assert_array_equal(P[n:], x[src[n:]] + shift)  # whole box lengths, exactly
    assert np.all(np.abs(shift) <= 1) and np.all(np.abs(shift).sum(1) > 0)
    # every image within rcut of the box is present, and nothing else
    lo = x >= 1 - rcut
    hi = x < rcut
    want = ((1 + lo + hi).prod(1) - 1).sum()
    assert len(P) - n == want


def test_treepm_jlists_are_the_particles_in_reach():
    rcut = 0.12
    m, x = ics.clumpy_box(3000, 14)
    L = pgamd.treepm_lists(x, rcut, 64)
    P = L["pos"]
    glo, gL = -rcut, 1.0 + 2 * rcut
    nc = max(1, int(np.floor(gL / rcut)))
    h = gL / nc
    for g in range(0, len(L["ioff"]), 7):
        ii = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
        jj = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
        bmin, bmax = P[ii].min(0), P[ii].max(0)
        e = np.maximum(np.maximum(bmin - P, P - bmax), 0.0)
        want = np.nonzero((e * e).sum(1) < rcut * rcut)[0]
        np.testing.assert_array_equal(np.sort(jj), want)
        # listed in grid-cell order (cells of side >= rcut over [-rcut, 1 + rcut),
        # by x, y, z index), a cell's particles by index
        c = np.clip(np.floor((P[jj] - glo) / h).astype(np.int64), 0, nc - 1)
        key = ((c[:, 0] * nc + c[:, 1]) * nc + c[:, 2]) * len(P) + jj
        assert np.all(np.diff(key) > 0)
    # groups partition the particles in Morton order
    np.testing.assert_array_equal(np.sort(L["ilist"]), np.arange(len(x)))


def test_bh_cells_hold_the_moments_of_their_particles():
    m, x = ics.hernquist(4000, 15)
    L = pgamd.bh_lists(x, m, 0.5, 64, 4)
    root = L["cells"][0]  # the root is the first cell
    M = m.sum()
    com = (m[:, None] * x).sum(0) / M
    d = x - com
    r2 = (d * d).sum(1)
    q = np.array([(m * (3 * d[:, a] * d[:, b] - (r2 if a == b else 0))).sum()
                  for a, b in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2))])
    np.testing.assert_allclose(root[3], M, rtol=1e-13)
    np.testing.assert_allclose(root[:3], com, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(root[4:9], q, rtol=1e-9, atol=1e-12 * np.abs(q).max())
    np.testing.assert_allclose(root[9], -(q[0] + q[3]), rtol=1e-9, atol=1e-12 * np.abs(q).max())


def test_unsplittable_nodes_keep_every_particle():
    """More than ncrit (TreePM) or nleaf (Barnes-Hut) coincident particles
    cannot be split by the 21-level key.  They form one i-group, and a
    Barnes-Hut walk that opens their node lists them as a leaf: every
    particle has a group and every group reaches the whole mass exactly once
    (listed particles + accepted cells)."""
    rng = np.random.default_rng(5)
    x = np.vstack([rng.uniform(0.1, 0.9, (2000, 3)), np.full((300, 3), 0.5)])
    L = pgamd.treepm_lists(x, 0.05, 128)
    np.testing.assert_array_equal(np.sort(L["ilist"]), np.arange(len(x)))
    assert L["icnt"].max() == 300  # the coincident particles: one group
    x = np.vstack([rng.uniform(-1, 1, (2000, 3)), np.full((40, 3), 0.25)])
    m = rng.uniform(0.5, 1.5, len(x)) / len(x)
    L = pgamd.bh_lists(x, m, 0.5, 64, 4)
    n = len(x)
    np.testing.assert_array_equal(np.sort(L["ilist"]), np.arange(n))
    assert L["icnt"].max() <= 64  # 40 coincident < ncrit: a group holds them with others
    mp = m[L["src"]]  # masses in the builder's particle order
    for g in range(len(L["ioff"])):
        pj = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
        cj = L["clist"][L["coff"][g]:L["coff"][g] + L["ccnt"][g]]
        assert len(np.unique(pj)) == len(pj)
        np.testing.assert_allclose(mp[pj].sum() + L["cells"][cj, 3].sum(), m.sum(), rtol=1e-12)
