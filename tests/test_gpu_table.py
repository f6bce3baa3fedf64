"""Tabulated-cutoff (TreePM short-range) parity on the GPU: libpg5_table.so
against the oracle's bit-exact table and lookup, the reference's outputs
(golden fixture), and the closed-form S2 force."""
import ctypes as C

import numpy as np
import pytest

import ics
import oracle as O
from conftest import golden, rel_err

pytestmark = pytest.mark.gpu

PP, PM = 0.1 / 128, 3.0 / 128


@pytest.fixture(scope="module")
def tl():
    import pgamd
    L = pgamd.load("pg5_table")
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    sig = pgamd._sig
    sig(L, "g5_open", None)
    sig(L, "pg5_gen_s2_force_table", None, C.c_double, C.c_double)
    sig(L, "pg5_set_table_frc_bit", C.c_int, C.c_int)
    sig(L, "pg5_get_table", C.c_int, fp, dp, fp)
    sig(L, "g5_set_xmj", None, C.c_int, C.c_int, dp, dp)
    sig(L, "g5_set_xj", None, C.c_int, C.c_int, dp)
    sig(L, "g5_set_mj", None, C.c_int, C.c_int, dp)
    sig(L, "g5_set_n", None, C.c_int)
    sig(L, "g5_calculate_force_on_x", None, dp, dp, dp, C.c_int)
    sig(L, "g5_set_xmjMC", None, C.c_int, C.c_int, C.c_int, dp, dp)
    sig(L, "g5_set_nMC", None, C.c_int, C.c_int)
    sig(L, "g5_calculate_force_on_xMC", None, C.c_int, dp, dp, dp, C.c_int)
    sig(L, "pg5t_dev_eval", C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p)
    L.pg5t_dev_eval.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
    L.g5_open()
    return L


def table(L):
    import pgamd
    n = L.pg5_get_table(None, None, None)
    t = np.zeros(2 * n, np.float32)
    xs = np.zeros(1)
    L.pg5_get_table(pgamd.fptr(t), pgamd.dptr(xs), None)
    return t, float(xs[0])


def force(L, xi, xj, mj):
    import pgamd
    xj, mj, xi = pgamd.f64(xj), pgamd.f64(mj), pgamd.f64(xi)
    L.g5_set_xmj(0, len(mj), pgamd.dptr(xj), pgamd.dptr(mj))
    L.g5_set_n(len(mj))
    a = np.zeros((len(xi), 3))
    p = np.ones(len(xi))
    L.g5_calculate_force_on_x(pgamd.dptr(xi), pgamd.dptr(a), pgamd.dptr(p), len(xi))
    assert np.all(p == 0.0)
    return a


def test_default_table_is_the_reference_table(tl):
    g = golden("pg5_table.npz")
    t, _ = table(tl)
    np.testing.assert_array_equal(t.view(np.uint32), g["tab64"].view(np.uint32))
    exported = np.ctypeslib.as_array((C.c_float * 1024).in_dll(tl, "Force_table"))
    np.testing.assert_array_equal(exported, t)


def test_lookup_bit_exact_on_device(tl):
    import torch
    tl.pg5_gen_s2_force_table(PP, PM)
    t, _ = table(tl)
    rng = np.random.default_rng(0)
    s = np.concatenate([2 + rng.uniform(0, 1, 4000) ** 4 * 1.3e5, [2.0, 129024.0, 1e9, 5e5]])
    s = s.astype(np.float32)
    ds = torch.from_numpy(s).cuda()
    out = torch.empty_like(ds)
    assert tl.pg5t_dev_eval(len(s), ds.data_ptr(), out.data_ptr()) == 0
    got = out.cpu().numpy()
    want = O.table_lookup(t, s)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_reference_fixture_and_emulation(tl):
    g = golden("pg5_table.npz")
    tl.pg5_gen_s2_force_table(PP, PM)
    a = force(tl, g["xi"], g["xj"], g["mj"])
    scale = np.linalg.norm(g["acc"], axis=1).max()
    assert np.percentile(rel_err(a, g["acc"], floor=1e-3 * scale), 99) < 1e-5
    t, xs = table(tl)
    e = O.table_force_emul(g["xi"], g["xj"], g["mj"], t, xs)
    # same per-pair terms: only FP32 tile partial sums differ
    assert np.percentile(rel_err(a, e, floor=1e-3 * scale), 99) < 2e-6


@pytest.mark.parametrize("frc", [5, 6])
def test_clumpy_box_vs_emulation_and_closed_form(tl, frc):
    assert tl.pg5_set_table_frc_bit(frc) == 0
    tl.pg5_gen_s2_force_table(PP, PM)
    t, xs = table(tl)
    tab_o, _, xs_o = O.table_build(PP, PM, frc_bit=frc)
    np.testing.assert_array_equal(t.view(np.uint32), tab_o.view(np.uint32))
    m, x = ics.clumpy_box(1 << 15, 31)
    xc = x[32768 // 2:]  # clumped half: dense neighbourhoods
    xi = xc[::97]
    a = force(tl, xi, x, m)
    e = O.table_force_emul(xi, x, m, t, xs, frc_bit=frc)
    scale = np.linalg.norm(e, axis=1).max()
    assert np.percentile(rel_err(a, e, floor=1e-3 * scale), 99) < 2e-6
    f = O.table_force_f64(xi, x, m, PP, PM)
    err = rel_err(a, f, floor=1e-3 * scale)
    # the table's own interpolation error against the closed form
    assert np.median(err) < (2e-3 if frc == 5 else 6e-4)
    tl.pg5_set_table_frc_bit(5)


def test_set_xj_and_set_mj(tl):
    import pgamd
    tl.pg5_gen_s2_force_table(PP, PM)
    m, x = ics.clumpy_box(4096, 32)
    a0 = force(tl, x[:50], x, m)
    # overwrite positions then masses separately, then restore
    xx = pgamd.f64(x + 0.5)
    tl.g5_set_xj(0, len(m), pgamd.dptr(xx))
    mm = pgamd.f64(m * 2)
    tl.g5_set_mj(0, len(m), pgamd.dptr(mm))
    tl.g5_set_n(len(m))
    ai = np.zeros((50, 3))
    p = np.zeros(50)
    xi = pgamd.f64(x[:50] + 0.5)
    tl.g5_calculate_force_on_x(pgamd.dptr(xi), pgamd.dptr(ai), pgamd.dptr(p), 50)
    scale = np.linalg.norm(a0, axis=1).max()
    assert np.percentile(rel_err(ai, 2 * a0, floor=1e-3 * scale), 99) < 1e-4


def test_n_beyond_the_addresses_written(tl):
    """g5_set_n past the addresses set adds zero records (no force).  On a
    board no other test writes (devid 37): addresses another call wrote keep
    their records, as in the reference's static j-memory."""
    import pgamd
    dev = 37
    tl.pg5_gen_s2_force_table(PP, PM)
    m, x = ics.clumpy_box(1000, 33)
    xi = pgamd.f64(x[:50])

    def calc(n):
        tl.g5_set_nMC(dev, n)
        a = np.zeros((50, 3))
        p = np.zeros(50)
        tl.g5_calculate_force_on_xMC(dev, pgamd.dptr(xi), pgamd.dptr(a), pgamd.dptr(p), 50)
        return a

    tl.g5_set_xmjMC(dev, 0, len(m), pgamd.dptr(pgamd.f64(x)), pgamd.dptr(pgamd.f64(m)))
    a0 = calc(len(m))
    np.testing.assert_allclose(a0, force(tl, x[:50], x, m), rtol=1e-6, atol=1e-12)
    scale = np.linalg.norm(a0, axis=1).max()
    for n in (3000, 20000):
        a = calc(n)
        assert np.all(np.isfinite(a))
        assert rel_err(a, a0, floor=1e-3 * scale).max() < 1e-6
