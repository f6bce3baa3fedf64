"""Pins the CPU oracle (oracle/pg_oracle.c) before it is trusted as the
checker of the GPU path: against the reference's own outputs (golden fixtures
made by running the compiled reference on its own input files), against the
compiled reference live where it is built, and against what the mathematics
fixes (closed forms, invariants, finite differences, brute force).
"""
import os
from fractions import Fraction

import numpy as np
import pytest

import ics
import oracle as O
from conftest import golden, rel_err


# ------------------------------------------------------------------- GRAPE-5
@pytest.mark.parametrize("name", ["pl1k", "pl1k_1", "pl1k_3"])
def test_pg5_oracle_matches_reference_fixture(name):
    g = golden("pg5_direct.npz")
    m, x, eps = g[f"{name}_m"], g[f"{name}_x"], float(g[f"{name}_eps"])
    a, p = O.pg5_force(x, eps * eps, x, m)
    # the reference uses the 12-bit rsqrtps estimate plus a bias correction
    # (PG5/newton/libpg5/rsqrt.c:42-81, phantom_g5.c:45-60): its own error
    # against FP64 is about 1e-4 per particle
    e = rel_err(g[f"{name}_acc"], a)
    assert np.median(e) < 1e-4 and e.max() < 3e-3
    ep = rel_err(g[f"{name}_pot"], p)
    assert np.median(ep) < 5e-5 and ep.max() < 1e-3


def test_pg5_symmetric_oracle_matches_reference_fixture():
    g = golden("pg5_symmetric.npz")
    a, p = O.pg5_force(g["x"], g["eps2"], g["x"], g["m"], eps2j=g["eps2"])
    assert np.median(rel_err(g["acc"], a)) < 1e-4
    assert np.median(rel_err(g["pot"], p)) < 5e-5
    # a wrong softening (eps_i^2 only) is detected
    a2, _ = O.pg5_force(g["x"], g["eps2"], g["x"], g["m"])
    assert np.median(rel_err(g["acc"], a2)) > 1e-3


def test_pg5_two_body_closed_form():
    xj = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 1.2]])
    mj = np.array([2.0, 0.5])
    eps = 0.1
    a, p = O.pg5_force(xj[:1], eps * eps, xj, mj)
    d = xj[1] - xj[0]
    r2 = d @ d + eps * eps
    np.testing.assert_allclose(a[0], 0.5 * d / r2 ** 1.5, rtol=1e-15)
    np.testing.assert_allclose(p[0], 2.0 / eps + 0.5 / np.sqrt(r2), rtol=1e-15)


def test_pg5_momentum_and_gradient():
    m, x, _ = ics.plummer(512, 3)
    eps = 1.0 / 64
    a, p = O.pg5_force(x, eps * eps, x, m)
    # Newton's third law: sum m a = 0 when the i-set is the j-set
    assert np.abs((m[:, None] * a).sum(0)).max() < 1e-13 * (m[:, None] * np.abs(a)).sum()
    # a = +grad_x(p) (p = + sum m/r)
    h = 1e-6
    for i in (0, 17, 300):
        for k in range(3):
            xp = x[i:i + 1].copy()
            xm = x[i:i + 1].copy()
            xp[0, k] += h
            xm[0, k] -= h
            _, pp = O.pg5_force(xp, eps * eps, x, m)
            _, pm = O.pg5_force(xm, eps * eps, x, m)
            np.testing.assert_allclose((pp - pm) / (2 * h), a[i, k], rtol=1e-6,
                                       atol=1e-8 * np.abs(a[i]).max())


def test_pg5_against_compiled_reference():
    if not O.ref_available("pg5"):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    m, x, _ = ics.plummer(2048, 8)
    ra, rp = O.RefPG5().force(x, x, m, 1.0 / 64)
    a, p = O.pg5_force(x, (1.0 / 64) ** 2, x, m)
    assert np.percentile(rel_err(ra, a), 99) < 1e-3
    assert np.median(rel_err(rp, p)) < 5e-5


# ------------------------------------------------------------------- GRAPE-6
def test_pg6_predictor_is_the_exact_polynomial():
    """With binary-fraction inputs every operation is exact, so the result must
    equal the rational polynomial x + v dt + a2 dt^2 + j6 dt^3."""
    rng = np.random.default_rng(1)
    n = 64
    q = lambda: rng.integers(-64, 64, (n, 3)) / 64.0  # noqa: E731
    x, v, a2, j6 = q(), q(), q(), q()
    tj = -rng.integers(0, 8, n) / 16.0
    ti = 0.125
    xp, vp = O.pg6_predict(ti, x, v, a2, j6, tj)
    for j in range(0, n, 7):
        dt = Fraction(ti) - Fraction(tj[j])
        for k in range(3):
            X = (Fraction(x[j, k]) + Fraction(v[j, k]) * dt + Fraction(a2[j, k]) * dt ** 2
                 + Fraction(j6[j, k]) * dt ** 3)
            V = Fraction(v[j, k]) + 2 * Fraction(a2[j, k]) * dt + 3 * Fraction(j6[j, k]) * dt ** 2
            assert xp[j, k] == float(X) and vp[j, k] == float(V)


def test_pg6_oracle_matches_reference_fixture():
    """Reference run on its own banana input (PG6/banana/init/pl001k.init),
    j-particles predicted from t_j to t_i: pins predictor and force together."""
    g = golden("pg6_hermite.npz")
    xp, vp = O.pg6_predict(float(g["ti"]), g["x"], g["v"], g["acc2"], g["jrk6"], g["tj"])
    idx = np.arange(len(g["m"]), dtype=np.int32)
    isel = g["isel"]
    np.testing.assert_array_equal(xp[isel], g["xi"])
    o = O.pg6_force(g["xi"], g["vi"], idx[isel], float(g["eps2"]), xp, vp, g["m"], idx)
    ea = rel_err(g["acc"], o["acc"])
    ej = rel_err(g["jrk"], o["jrk"])
    ep = rel_err(g["pot"], o["pot"])
    # reference: FP32 pair math with one Newton step on rsqrtps, FP32 jerk sums
    assert np.percentile(ea, 99) < 1e-6 and ea.max() < 5e-6
    assert np.percentile(ej, 99) < 1e-5
    assert ep.max() < 1e-6  # pot is returned through a float field
    assert np.mean(g["nnb"] == o["nnb"]) == 1.0
    assert rel_err(g["rnnb"], o["rnnb"]).max() < 1e-6
    # a predictor without the jerk term is detected
    xq, vq = O.pg6_predict(float(g["ti"]), g["x"], g["v"], g["acc2"], 0 * g["jrk6"], g["tj"])
    o2 = O.pg6_force(g["xi"], g["vi"], idx[isel], float(g["eps2"]), xq, vq, g["m"], idx)
    assert np.percentile(rel_err(g["jrk"], o2["jrk"]), 99) > 1e-5


def test_pg6_neighbour_lists_match_reference_fixture():
    g = golden("pg6_hermite.npz")
    xp, _ = O.pg6_predict(float(g["ti"]), g["x"], g["v"], g["acc2"], g["jrk6"], g["tj"])
    idx = np.arange(len(g["m"]), dtype=np.int32)
    lists, nblen = O.pg6_nblist(g["xi"], idx[g["isel"]], g["h2"], float(g["eps2"]), xp, idx)
    np.testing.assert_array_equal(nblen, g["nblen"])
    np.testing.assert_array_equal(np.concatenate(lists), g["nbl"])


def test_pg6_two_body_closed_form():
    xi = np.array([[0.1, 0.2, 0.3]])
    vi = np.array([[0.0, 0.5, 0.0]])
    xj = np.array([[0.1, 0.2, 0.3], [1.1, 0.2, -0.7]])
    vj = np.array([[9.0, 9.0, 9.0], [0.3, 0.0, 0.1]])
    mj = np.array([5.0, 0.25])
    eps2 = 0.01
    o = O.pg6_force(xi, vi, [0], eps2, xj, vj, mj, [0, 1])
    d = xj[1] - xi[0]
    dv = vj[1] - vi[0]
    r2 = d @ d + eps2
    a = 0.25 * d / r2 ** 1.5
    jk = 0.25 * (dv / r2 ** 1.5 - 3 * (d @ dv) * d / r2 ** 2.5)
    np.testing.assert_allclose(o["acc"][0], a, rtol=1e-14)
    np.testing.assert_allclose(o["jrk"][0], jk, rtol=1e-13)
    np.testing.assert_allclose(o["pot"][0], -0.25 / np.sqrt(r2), rtol=1e-15)
    assert o["nnb"][0] == 1 and abs(o["rnnb"][0] - np.sqrt(r2)) < 1e-15


def test_pg6_jerk_is_the_time_derivative_of_acc():
    m, x, v = ics.plummer(256, 4)
    idx = np.arange(256, dtype=np.int32)
    eps2 = (1.0 / 64) ** 2
    o = O.pg6_force(x, v, idx, eps2, x, v, m, idx)
    h = 1e-6
    op = O.pg6_force(x + h * v, v, idx, eps2, x + h * v, v, m, idx)
    om = O.pg6_force(x - h * v, v, idx, eps2, x - h * v, v, m, idx)
    fd = (op["acc"] - om["acc"]) / (2 * h)
    assert np.percentile(rel_err(fd, o["jrk"]), 99) < 1e-5


def test_pg6_against_compiled_reference():
    if not O.ref_available("pg6"):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    m, x, v = ics.plummer(1024, 9)
    n = len(m)
    idx = np.arange(n, dtype=np.int32)
    eps2 = (1.0 / 256) ** 2
    R = O.RefPG6()
    z = np.zeros((n, 3))
    R.set_j(x, v, z, z, m, np.zeros(n), idx)
    r = R.calc(n, 0.0, x, v, idx, eps2)
    o = O.pg6_force(x, v, idx, eps2, x, v, m, idx)
    assert np.percentile(rel_err(r["acc"], o["acc"]), 99) < 1e-6
    assert np.percentile(rel_err(r["jrk"], o["jrk"]), 99) < 2e-5
    assert np.mean(r["nnb"] == o["nnb"]) == 1.0


# --------------------------------------------------------------------- table
def test_s2_shape_is_continuous_and_cut():
    pp, pm = 0.1 / 128, 3.0 / 128
    for sft in (pp, pm):
        for eta in (1.0, 2.0):
            r = eta * sft / 2
            lo = O.lib().or_s2_reff(sft, r * (1 - 1e-12))
            hi = O.lib().or_s2_reff(sft, r * (1 + 1e-12))
            assert abs(lo - hi) < 1e-9 * abs(hi)
    r = np.array([pm * 1.0001, pm * 2, 0.5])
    np.testing.assert_allclose(O.s2_force(pp, pm, r), 0.0, atol=1e-9)
    # unsoftened Newtonian limit between the two softening lengths
    r = np.array([pp * 3, pp * 5])
    f = O.s2_force(pp, pm, r)
    assert np.all(f > 0) and np.all(f < 1 / r ** 3)


@pytest.mark.parametrize("key,eps", [("tab64", (0.1 / 64, 3.0 / 64)),
                                     ("tab128", (0.1 / 128, 3.0 / 128))])
def test_table_build_is_bit_exact_with_reference(key, eps):
    g = golden("pg5_table.npz")
    tab, _, _ = O.table_build(*eps)
    np.testing.assert_array_equal(tab.view(np.uint32), g[key].astype(np.float32).view(np.uint32))


def test_table_lookup_hits_entries_and_interpolates():
    tab, r2s, xs = O.table_build(0.1 / 128, 3.0 / 128)
    # at the float-bit ticks s_k the lookup returns the entry exactly
    ticks = (np.arange(0, 512, 3, dtype=np.uint32) << 18) + np.uint32(0x40000000)
    s = ticks.view(np.float32)
    np.testing.assert_array_equal(O.table_lookup(tab, s), tab[0::2][np.arange(0, 512, 3)])
    # beyond the cut-off the last entry (about 0) is used
    assert O.table_lookup(tab, [1e9])[0] == tab[2 * 511]


def test_table_force_emulation_matches_reference_fixture():
    g = golden("pg5_table.npz")
    tab, r2s, xs = O.table_build(0.1 / 128, 3.0 / 128)
    a = O.table_force_emul(g["xi"], g["xj"], g["mj"], tab, xs)
    ref = g["acc"]
    scale = np.linalg.norm(ref, axis=1).max()
    # identical per-pair terms; the reference sums them in FP32
    e = rel_err(ref, a, floor=1e-3 * scale)
    assert np.percentile(e, 99) < 1e-5, np.percentile(e, 99)
    # and the closed-form S2 force is close, within the table's own error
    f = O.table_force_f64(g["xi"], g["xj"], g["mj"], 0.1 / 128, 3.0 / 128)
    ef = rel_err(f, a, floor=1e-3 * scale)
    assert np.median(ef) < 5e-3


def test_table_against_compiled_reference():
    if not O.ref_available("pg5table"):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    T = O.RefTable()
    T.gen_s2(0.1 / 128, 3.0 / 128)
    tab, _, xs = O.table_build(0.1 / 128, 3.0 / 128)
    np.testing.assert_array_equal(T.table().view(np.uint32), tab.view(np.uint32))
    m, x = ics.clumpy_box(4096, 2)
    sel = x[:, 0] < 0.05
    xj, mj = x[sel], m[sel]
    if len(mj) % 2:
        xj, mj = xj[:-1], mj[:-1]
    xi = xj[::5]
    ra = T.force(xi, xj, mj)
    a = O.table_force_emul(xi, xj, mj, tab, xs)
    scale = np.linalg.norm(ra, axis=1).max()
    assert np.percentile(rel_err(ra, a, floor=1e-3 * scale), 99) < 1e-5


# ---------------------------------------------------------------- quadrupole
def test_quad_moments_traceless_and_definition():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(20, 3))
    m = rng.uniform(0.5, 1.5, 20)
    c, M, q = O.quad_moments(x, m)
    assert abs(q[0] + q[3] + q[5]) < 1e-12
    d = x - c
    Q = 3 * np.einsum("n,ni,nj->ij", m, d, d) - np.eye(3) * (m * (d * d).sum(1)).sum()
    np.testing.assert_allclose(q, [Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2], Q[2, 2]],
                               atol=1e-12)


def test_quad_far_field_converges_to_direct_sum():
    """Monopole + quadrupole of a particle cloud reproduces the direct sum at
    distance D with residual O((size/D)^4) (the dipole vanishes about the
    centre of mass), while monopole-only leaves O((size/D)^2)."""
    rng = np.random.default_rng(3)
    x = rng.normal(size=(50, 3)) * [1.0, 0.6, 0.3]
    m = rng.uniform(0.5, 1.5, 50)
    c, M, q = O.quad_moments(x, m)
    prev = None
    for D in (20.0, 40.0):
        xi = c + D * np.array([[0.6, 0.64, 0.48], [-0.8, 0.0, 0.6]])
        ad, pd = O.pg5_force(xi, 0.0, x, m)
        aq, pq = O.quad_force(xi, 0.0, c[None], [M], q[None])
        am, pm = O.pg5_force(xi, 0.0, c[None], np.array([M]))
        eq = rel_err(aq, ad).max()
        em = rel_err(am, ad).max()
        assert eq < 0.05 * em
        np.testing.assert_allclose(pq, pd, rtol=5 * (2.0 / D) ** 3)
        if prev is not None:
            assert eq < prev / 8  # at least third order
        prev = eq


def test_quad_oracle_matches_reference_fixture():
    g = golden("pg5_quad.npz")
    a, p = O.quad_force(g["xi"], float(g["eps"]) ** 2, g["xc"], g["mc"], g["qc"])
    # reference: raw 12-bit rsqrtps, FP32 sums; the g5c potential is negative
    assert np.median(rel_err(g["acc"], a)) < 3e-4
    np.testing.assert_allclose(g["pot"], -p, rtol=3e-3)


# ------------------------------------------------------------ rsqrt bias
def test_rsqrt_bias_restatement():
    """or_rsqrt_bias_of restates rsqrt.c:42-81.  With the reference's own
    rsqrt (SSE rsqrtss) it gives the compiled reference's rsqrt_bias() (that
    build sums under -ffast-math, so to 1e-11 relative, not bit for bit); with
    correctly rounded FP32 values the bias vanishes; and values planted 1 + d
    too large give b = d to first order (a dropped factor of the rule fails)."""
    import ctypes as C
    y0, y1 = O.rsqrtss_table()
    b = O.rsqrt_bias_of(y0, y1)
    if O.ref_available("pg5"):
        R = C.CDLL(O._refpath("pg5").decode(), mode=C.RTLD_LOCAL)
        R.rsqrt_bias.restype = C.c_double
        br = R.rsqrt_bias()
        assert abs(b - br) < 1e-11 * abs(br), (b, br)
    assert 0 < abs(b) < 1e-5
    x0, x1 = O.rsqrt_args()
    e0 = (1 / np.sqrt(x0.astype(np.float64))).astype(np.float32)
    e1 = (1 / np.sqrt(x1.astype(np.float64))).astype(np.float32)
    assert abs(O.rsqrt_bias_of(e0, e1)) < 1e-10
    d = 3e-4
    bp = O.rsqrt_bias_of((e0 * (1 + d)).astype(np.float32), (e1 * (1 + d)).astype(np.float32))
    assert abs(bp - d) < 2 * d * d, (bp, d)


# --------------------------------------------------------- learned cut-off
def test_mlp_oracle_gradient_matches_finite_differences():
    import mlp_oracle as M
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.7, M.NP)
    v = rng.uniform(0, 1, 64)
    t = rng.normal(0, 1, 64)
    wt = rng.uniform(0.5, 2, 64)
    _, g = M.loss_and_grad(p, v, t, wt)
    h = 1e-6
    for k in list(range(0, 32, 5)) + [40, 100, 287, 290, 305, 320]:
        e = np.zeros(M.NP)
        e[k] = h
        fd = (M.loss_and_grad(p + e, v, t, wt)[0] - M.loss_and_grad(p - e, v, t, wt)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(g[k])), (k, fd, g[k])
    # forward agrees with the loss' own forward
    y = M.forward(p, v)
    assert abs(M.loss_and_grad(p, v, y, wt)[0]) < 1e-28


def test_mlp_oracle_jacobian_matches_finite_differences():
    """Every column of the oracle Jacobian against central differences of the
    forward pass, and the normal equations against their definition (the
    gradient of the loss is 2 J^T r / n)."""
    import mlp_oracle as M
    rng = np.random.default_rng(6)
    p = rng.normal(0, 0.7, M.NP)
    v = rng.uniform(0, 1, 48)
    t = rng.normal(0, 1, 48)
    wt = rng.uniform(0.5, 2, 48)
    y, J = M.jacobian(p, v)
    np.testing.assert_allclose(y, M.forward(p, v), rtol=0, atol=1e-14)
    h = 1e-6
    for k in range(M.NP):
        e = np.zeros(M.NP)
        e[k] = h
        fd = (M.forward(p + e, v) - M.forward(p - e, v)) / (2 * h)
        assert np.max(np.abs(fd - J[:, k])) < 1e-7 * max(1.0, np.abs(J[:, k]).max()), k
    jtj, jtr, rtr = M.normal_eq(p, v, t, wt)
    loss, g = M.loss_and_grad(p, v, t, wt)
    np.testing.assert_allclose(2 * jtr / len(v), g, rtol=1e-10, atol=1e-14)
    assert abs(rtr / len(v) - loss) < 1e-12 * loss
    np.testing.assert_allclose(jtj, jtj.T, rtol=0, atol=0)


def test_s2_unit_shape_gives_the_cutoff_force():
    """The S2 form of the learned target: with H(u) = R(u, 1), the short-range
    shape is f(r) = [r >= eps_PP] / r^3 + [r < eps_PP] H(r / eps_PP) / eps_PP^3
    - H(r / eps_PM) / eps_PM^3 below eps_PM.  Checked against the oracle's
    closed form of pg5_table.c:27-29 on a grid through the core, both joints
    of the polynomial pieces and the cut-off; H(0) = 12.8, H(1) = 1."""
    import mlp_oracle as M
    pp, pm = 0.1 / 128, 3.0 / 128
    H = lambda u: M.s2_unit(u, O.s2_reff)
    np.testing.assert_allclose(H([0.0, 1.0]), [12.8, 1.0], rtol=1e-15)
    r = np.concatenate([np.geomspace(1e-3 * pp, pm, 3000)[:-1], [pp / 2, pp, pm / 2]])
    core = r < pp
    f = np.where(core, H(np.minimum(r / pp, 1.0)) / pp ** 3, 1.0 / r ** 3) - H(r / pm) / pm ** 3
    # relative to the pair's Newtonian 1/r^3 (f itself vanishes at the cut-off)
    assert np.max(np.abs(f - O.s2_force(pp, pm, r)) * r ** 3) < 1e-12


def test_mlp_target_is_bounded_and_smooth():
    import mlp_oracle as M
    pp, pm = 0.1 / 128, 3.0 / 128
    r = np.linspace(0, pm, 20001)[:-1]
    v, y = M.target(lambda rr: O.s2_force(pp, pm, rr), r, pm, pp)
    assert y.max() < 13 and y.min() > -1e-6
    assert np.max(np.abs(np.diff(y))) < 0.05  # no jumps on a fine grid


# ------------------------------------------ CPU baseline drivers of the batches
def _need_ref(*names):
    for n in names:
        if not O.ref_available(n):
            pytest.skip(f"reference library {n} not built")


def test_cpu_treepm_driver_matches_per_group_calls():
    _need_ref("pg5table")
    import pgamd
    pp, pm = 0.1 / 16, 3.0 / 16
    m, x = ics.clumpy_box(4096, 5)
    L = pgamd.treepm_lists(x, pm, 64)
    P, mm = L["pos"], m[L["src"]]
    groups = np.arange(0, len(L["ioff"]), 3)
    t, acc = O.ref_treepm(L, P, mm, pp, pm, 4, groups)
    assert t > 0
    R = O.RefTable()
    R.gen_s2(pp, pm)
    for g in groups[:12]:
        ii = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
        jj = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
        xj, mj = P[jj], mm[jj]
        if len(jj) % 2:  # ref_treepm pads odd lists with one massless j
            xj, mj = np.vstack([xj, np.zeros((1, 3))]), np.append(mj, 0.0)
        a = R.force(P[ii], xj, mj, devid=0)
        np.testing.assert_array_equal(acc[ii], a)
    # and it is the short-range force: compare with the FP64 closed form
    sel = L["ilist"][L["ioff"][groups[0]]:L["ioff"][groups[0]] + L["icnt"][groups[0]]]
    f = O.table_force_f64(P[sel], P, mm, pp, pm)
    assert np.median(rel_err(acc[sel], f)) < 1e-2


def test_cpu_bh_driver_matches_per_group_calls():
    _need_ref("pg5quad")
    import pgamd
    eps = 1.0 / 64
    m, x = ics.hernquist(4096, 9)
    L = pgamd.bh_lists(x, m, 0.5, 128, 8)
    P, mm = L["pos"], m[L["src"]]
    groups = np.arange(len(L["ioff"]))
    t, acc, pot = O.ref_bh(L, P, mm, L["cells"], eps, 4, groups)
    assert t > 0
    R = O.RefPG5(quad=True)
    for g in groups[:6]:
        ii = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
        jj = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
        cc = L["clist"][L["coff"][g]:L["coff"][g] + L["ccnt"][g]]
        a1, p1 = R.force(P[ii], P[jj], mm[jj], eps)
        a2, p2 = (R.cell_force(P[ii], L["cells"][cc, :3], L["cells"][cc, 3], L["cells"][cc, 4:],
                               eps) if len(cc) else (0.0, 0.0))
        np.testing.assert_allclose(acc[ii], a1 + a2, rtol=1e-12, atol=0)
        np.testing.assert_allclose(pot[ii], p1 + p2, rtol=1e-12, atol=0)
    # the tree force approximates the direct sum
    ad, _ = O.pg5_force(P, eps * eps, P, mm)
    assert np.median(rel_err(acc, ad)) < 5e-3


def test_cpu_cluster_driver_matches_reference_sessions():
    _need_ref("pg6")
    sizes = np.array([64, 100, 77])
    parts = [ics.plummer(int(n), 70 + k) for k, n in enumerate(sizes)]
    M = np.concatenate([p[0] for p in parts])
    X = np.concatenate([p[1] for p in parts])
    V = np.concatenate([p[2] for p in parts])
    coff = np.concatenate([[0], np.cumsum(sizes)])
    e2 = (4.0 / sizes) ** 2
    t, acc, jrk, pot = O.ref_clusters([0, 2], coff, X, V, M, e2)
    assert t > 0
    for c in (0, 2):
        s = slice(coff[c], coff[c + 1])
        idx = np.arange(sizes[c], dtype=np.int32)
        o = O.pg6_force(X[s], V[s], idx, e2[c], X[s], V[s], M[s], idx)
        assert np.percentile(rel_err(acc[s], o["acc"]), 99) < 1e-6
        assert np.percentile(rel_err(jrk[s], o["jrk"]), 99) < 1e-4
    assert np.all(acc[coff[1]:coff[2]] == 0)


# --------------------------------------------- Hermite integrator restatement
def test_calc_modulus_is_zero_exactly_on_the_block_grid():
    dtmax = 0.125
    for k in range(1, 9):
        mod = dtmax / 2 ** k
        for m in range(1, 2 ** k + 1):
            assert O.calc_modulus(m * mod, mod, dtmax) == 0.0
            if m * mod + mod / 2 <= dtmax:
                assert O.calc_modulus(m * mod + mod / 2, mod, dtmax) != 0.0
    assert O.calc_modulus(0.0, 1e-3, dtmax) == 0.0


def test_hermite_predictor_is_the_taylor_cubic():
    rng = np.random.default_rng(1)
    x, v, a, j = (rng.normal(size=(50, 3)) for _ in range(4))
    t = rng.uniform(0, 0.1, 50)
    tg = 0.125
    xp, vp = O.hermite_predict(tg, t, x, v, a, j)
    h = (tg - t)[:, None]
    np.testing.assert_allclose(xp, x + v * h + a * h ** 2 / 2 + j * h ** 3 / 6, rtol=1e-14,
                               atol=1e-15)
    np.testing.assert_allclose(vp, v + a * h + j * h ** 2 / 2, rtol=1e-14, atol=1e-15)


def test_hermite_corrector_is_exact_for_quintic_motion():
    """The 4th-order Hermite corrector reproduces x(t), v(t) exactly when the
    acceleration is a cubic in t (x a quintic): a closed-form check."""
    rng = np.random.default_rng(2)
    c = rng.normal(size=(6, 20, 3))  # x(t) = sum c_k t^k

    def deriv(k, t):
        out = 0.0
        for p in range(k, 6):
            f = np.prod(np.arange(p - k + 1, p + 1)) if k else 1.0
            out = out + f * c[p] * t ** (p - k)
        return out
    dt = 0.03125
    x0, v0, a0, j0 = (deriv(k, 0.0) for k in range(4))
    a1, j1 = deriv(2, dt), deriv(3, dt)
    t = np.zeros(20)
    xp, vp = O.hermite_predict(dt, t, x0, v0, a0, j0)
    x1, v1, _ = O.hermite_correct(0.2, 0.125, t, np.full(20, dt), a0, j0, xp, vp, a1, j1)
    np.testing.assert_allclose(x1, deriv(0, dt), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(v1, deriv(1, dt), rtol=1e-12, atol=1e-14)


def test_hermite_steps_follow_the_criteria():
    """dt_startup: the largest block step <= eta_s |a|/|j| on the grid;
    dt_criterion: halve below the Aarseth step, double only when commensurate."""
    rng = np.random.default_rng(3)
    a, j = rng.normal(size=(200, 3)), rng.normal(size=(200, 3)) * 10 ** rng.uniform(-2, 3, (200, 1))
    dtmax = 0.125
    dt = O.dt_startup(0.01, dtmax, 0.0, a, j)
    crit = 0.01 * np.linalg.norm(a, axis=1) / np.linalg.norm(j, axis=1)
    k = np.ceil(np.log2(dtmax / crit)).clip(0)
    np.testing.assert_array_equal(dt, dtmax / 2 ** k)
    # corrector step rule, with a0 = a1 and j0 = j1 (s0 and c0 then follow from them)
    t = np.zeros(200)
    dts = np.full(200, dtmax / 64)
    xp = vp = np.zeros((200, 3))
    _, _, dtn = O.hermite_correct(0.2, dtmax, t, dts, a, j, xp, vp, a, j)
    s0 = 2.0 * (-(3 * j) * dts[:, None]) / dts[:, None] ** 2
    c0 = 6.0 * (2 * j * dts[:, None]) / dts[:, None] ** 3
    s1 = s0 + c0 * dts[:, None]
    A, J, S, Cc = (np.linalg.norm(q, axis=1) for q in (a, j, s1, c0))
    dti = 0.2 * np.sqrt((A * S + J * J) / (J * Cc + S * S))
    for i in range(200):
        if dti[i] < dts[i]:
            assert dtn[i] <= dti[i] or dtn[i] == dts[i] / 2 ** 51
            assert 2 * dtn[i] > dti[i]
        elif dti[i] > 2 * dts[i]:
            # t + dt = dtmax/64 is a multiple of 2 dt only for dt <= dtmax/128
            assert dtn[i] == dts[i]
        else:
            assert dtn[i] == dts[i]


def test_reference_integrator_conserves_energy():
    """The CPU baseline integrator (reference library + restated banana loop)
    on a Plummer sphere: relative energy drift over two windows."""
    if not O.ref_available("pg6"):
        pytest.skip("reference library pg6 not built")
    m, x, v = ics.plummer(512, 21)
    t, xe, ve, info, en = O.ref_hermite(x, v, m, (1 / 64) ** 2, 0.01, 0.2, 0.125, 0.25)
    e = en.sum(1)
    assert info["windows"] == 2 and info["blocks"] > 10
    # potential from the library excludes self: E0 from the initial state
    o = O.pg5_force(x, (1 / 64) ** 2, x, m)
    e0 = 0.5 * (m * (v * v).sum(1)).sum() - 0.5 * (m * (o[1] - m / (1 / 64))).sum()
    assert abs(e[-1] - e0) / abs(e0) < 1e-4, (e, e0)


# ------------------------------------------- long-range (mesh) side of TreePM
def test_ewald_sum_pins():
    import pm_oracle as P
    g = np.array([(i, j, k) for i in (0.25, 0.75) for j in (0.25, 0.75) for k in (0.25, 0.75)])
    assert np.abs(P.ewald_force(g, np.ones(8))).max() < 1e-9  # lattice: no net force
    x = np.array([[0.5, 0.5, 0.5], [0.51, 0.5, 0.5]])
    a = P.ewald_force(x, np.ones(2))
    np.testing.assert_allclose(a[0], [1e4, 0, 0], rtol=1e-3, atol=1e-9)  # Newtonian up close
    r = np.linspace(1e-4, 0.2, 64)
    np.testing.assert_allclose(P.reff(r, 0.05), [O.s2_reff(0.05, v) for v in r], rtol=1e-14)


def test_mesh_force_converges_to_the_meshless_long_range_force():
    import pm_oracle as P
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, (48, 3))
    m = np.full(48, 1 / 48)
    ex = P.pm_exact(x, m, 3 / 32)
    e32 = np.median(rel_err(P.pm_force(x, m, 32, 3 / 32), ex))
    e64 = np.median(rel_err(P.pm_force(x, m, 64, 3 / 32), ex))
    assert e32 < 2e-2 and e64 < 1e-3 and e64 < e32 / 10


def test_short_plus_long_range_is_the_periodic_force():
    import pgamd
    import pm_oracle as P
    mesh = 32
    pp, pm = 0.1 / mesh, 3.0 / mesh
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, (96, 3))
    m = rng.uniform(0.5, 1.5, 96) / 96
    L = pgamd.treepm_lists(x, pm, 32)
    tot = np.zeros((96, 3))
    tot[L["src"][:96]] = O.table_force_f64(L["pos"][:96], L["pos"], m[L["src"]], pp, pm)
    tot += P.pm_force(x, m, mesh, pm)
    e = rel_err(tot, P.pm_exact(x, m, pp))
    assert np.median(e) < 3e-2 and np.percentile(e, 90) < 8e-2


# ------------------------------------------ the reference's own sample programs
def test_reference_sample_programs_run(tmp_path):
    """direct.c and banana01 built against the reference library (the other
    half of the drop-in check runs them against libpg5 / libpg6 on a GPU)."""
    import dropin
    if not (os.path.exists(dropin.program("direct_ref")) and
            os.path.exists(dropin.program("banana_ref"))):
        pytest.skip("reference sample programs not built (oracle/ref.mk)")
    m, x, v = ics.plummer(512, 5)
    dropin.write_direct_input(tmp_path / "pl", m, x, v)
    de = dropin.direct_energy_errors(dropin.run("direct_ref", ["pl", "pl.out"], tmp_path))
    assert len(de) == 9 and np.abs(de).max() < 1e-2
    dropin.write_banana_input(tmp_path, m, x, v, 0.25)
    out = dropin.run("banana_ref", [], tmp_path)
    de = dropin.banana_energy_errors(out)
    assert len(de) == 3 and np.abs(de).max() < 1e-5
