"""GRAPE-5 Newtonian, symmetric-softening and quadrupole parity on the GPU:
libpg5.so / libpg5_sym.so through the C ABI against the CPU oracle and the
reference's own outputs (golden fixtures)."""
import threading

import numpy as np
import pytest

import ics
import oracle as O
from conftest import golden, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g5():
    import pgamd
    return pgamd.G5()


@pytest.fixture(scope="module")
def g5s():
    import pgamd
    return pgamd.G5(symmetric=True)


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("name", ["pl1k", "pl1k_1", "pl1k_3"])
def test_reference_fixture(g5, name, variant):
    """Reference direct.c inputs (N = 1024, 1023, 1021: the tail cases), both
    particle kernels (LDS tiles, scalar broadcast)."""
    assert g5.lib.pg5_dev_set_kernel(variant) == 0
    g = golden("pg5_direct.npz")
    m, x, eps = g[f"{name}_m"], g[f"{name}_x"], float(g[f"{name}_eps"])
    g5.set_xmj(0, x, m)
    g5.set_eps_to_all(eps)
    g5.set_n(len(m))
    a, p = g5.calculate_force_on_x(x)
    # within the reference's own error (12-bit rsqrtps, FP32 sums)
    assert np.median(rel_err(a, g[f"{name}_acc"])) < 1e-4
    assert np.median(rel_err(p, g[f"{name}_pot"])) < 5e-5
    # against FP64 on the FP32-rounded inputs the library computes with
    oa, op = O.pg5_force(f32(x), np.float32(eps * eps), f32(x), f32(m))
    assert np.percentile(rel_err(a, oa), 99) < 1e-5
    assert rel_err(p, op).max() < 1e-6
    g5.lib.pg5_dev_set_kernel(0)


@pytest.mark.parametrize("variant", [0, 1])
def test_symmetric_softening_both_kernels(g5s, variant):
    assert g5s.lib.pg5_dev_set_kernel(variant) == 0
    g = golden("pg5_symmetric.npz")
    x, m, e2 = g["x"], g["m"], g["eps2"]
    g5s.set_xmj(0, x, m, epsj2=e2)
    g5s.set_n(len(m))
    a, p = g5s.calculate_force_on_x(x, eps2=e2)
    oa, op = O.pg5_force(f32(x), f32(e2), f32(x), f32(m), eps2j=f32(e2))
    assert np.percentile(rel_err(a, oa), 99) < 1e-5
    assert rel_err(p, op).max() < 1e-6
    g5s.lib.pg5_dev_set_kernel(0)


def test_plummer_16k_vs_fp64(g5):
    m, x, _ = ics.plummer(16384, 21)
    eps = 1.0 / 64
    g5.set_xmj(0, x, m)
    g5.set_eps_to_all(eps)
    g5.set_n(len(m))
    a, p = g5.calculate_force_on_x(x)
    sel = np.arange(0, 16384, 13)
    oa, op = O.pg5_force(f32(x[sel]), np.float32(eps * eps), f32(x), f32(m))
    assert np.percentile(rel_err(a[sel], oa), 99) < 1e-5
    assert rel_err(p[sel], op).max() < 1e-6
    # and against exact FP64 inputs: dominated by the FP32 positions
    ea, _ = O.pg5_force(x[sel], eps * eps, x, m)
    assert np.percentile(rel_err(a[sel], ea), 99) < 1e-4


def test_pipeline_api_matches_calculate(g5):
    m, x, _ = ics.plummer(1000, 22)
    g5.set_xmj(0, x, m)
    g5.set_eps_to_all(0.02)
    g5.set_n(len(m))
    a, p = g5.calculate_force_on_x(x[:4])
    L = g5.lib
    import pgamd
    np_ = L.g5_get_number_of_pipelines()
    assert np_ == 4
    xi = np.ascontiguousarray(x[:4])
    L.g5_set_xiMC(0, 4, pgamd.dptr(xi))
    L.g5_runMC(0)
    a2 = np.zeros((4, 3))
    p2 = np.zeros(4)
    L.g5_get_forceMC(0, 4, pgamd.dptr(a2), pgamd.dptr(p2))
    np.testing.assert_array_equal(a2, a)
    np.testing.assert_array_equal(p2, p)
    # per-i softening through g5_set_xei / g5_calculate_force_on_xe
    e2 = np.full(4, 0.02 ** 2)
    a3, p3 = g5.calculate_force_on_x(x[:4], eps2=e2)
    np.testing.assert_array_equal(a3, a)


def test_partial_j_update_and_devids(g5):
    m, x, _ = ics.plummer(2048, 23)
    eps = 0.01
    g5.set_eps_to_all(eps)
    # devid 1 holds the first half, devid 2 the full set; addresses from adr
    g5.set_xmj(0, x[:1024], m[:1024], devid=1)
    g5.set_n(1024, devid=1)
    g5.set_xmj(0, x, m, devid=2)
    g5.set_n(2048, devid=2)
    a1, _ = g5.calculate_force_on_x(x[:100], devid=1)
    a2, _ = g5.calculate_force_on_x(x[:100], devid=2)
    o1, _ = O.pg5_force(f32(x[:100]), np.float32(eps * eps), f32(x[:1024]), f32(m[:1024]))
    o2, _ = O.pg5_force(f32(x[:100]), np.float32(eps * eps), f32(x), f32(m))
    assert np.percentile(rel_err(a1, o1), 99) < 1e-5
    assert np.percentile(rel_err(a2, o2), 99) < 1e-5
    # rewrite addresses [1024, 2048) of devid 1 and extend n: now equals devid 2
    g5.set_xmj(1024, x, m, devid=1)
    g5.set_n(2048, devid=1)
    a1b, _ = g5.calculate_force_on_x(x[:100], devid=1)
    np.testing.assert_array_equal(a1b, a2)


def test_threads_on_separate_devids(g5):
    """The quad_tree caller drives one devid per OpenMP thread (source.cpp:813-860)."""
    m, x, _ = ics.plummer(4096, 24)
    g5.set_eps_to_all(0.01)
    res = {}

    def work(d):
        lo = 512 * d
        g5.set_xmj(0, x[lo:lo + 512], m[lo:lo + 512], devid=8 + d)
        g5.set_n(512, devid=8 + d)
        res[d] = g5.calculate_force_on_x(x[:256], devid=8 + d)[0]

    th = [threading.Thread(target=work, args=(d,)) for d in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for d in range(8):
        lo = 512 * d
        o, _ = O.pg5_force(f32(x[:256]), np.float32(1e-4), f32(x[lo:lo + 512]),
                           f32(m[lo:lo + 512]))
        assert np.percentile(rel_err(res[d], o), 99) < 1e-5


def test_symmetric_softening_fixture(g5s):
    g = golden("pg5_symmetric.npz")
    x, m, e2 = g["x"], g["m"], g["eps2"]
    g5s.set_xmj(0, x, m, epsj2=e2)
    g5s.set_n(len(m))
    a, p = g5s.calculate_force_on_x(x, eps2=e2)
    assert np.median(rel_err(a, g["acc"])) < 1e-4
    oa, op = O.pg5_force(f32(x), f32(e2), f32(x), f32(m), eps2j=f32(e2))
    assert np.percentile(rel_err(a, oa), 99) < 1e-5


def test_quadrupole_fixture_and_oracle(g5):
    g = golden("pg5_quad.npz")
    eps = float(g["eps"])
    g5.set_eps_to_all(eps)
    for mem in (0, 1):
        g5.set_cells(g["xc"], g["mc"], g["qc"], mem=mem)
        a, p = g5.cell_force_on_x(g["xi"], mem=mem)
        assert np.median(rel_err(a, g["acc"])) < 3e-4
        np.testing.assert_allclose(p, g["pot"], rtol=3e-3)
        oa, op = O.quad_force(f32(g["xi"]), np.float32(eps * eps), f32(g["xc"]), f32(g["mc"]),
                              f32(g["qc"]))
        assert np.percentile(rel_err(a, oa), 99) < 1e-5
        np.testing.assert_allclose(-p, op, rtol=1e-6)


def test_quadrupole_near_cells(g5):
    """Cells at distances comparable to their size (theta ~ 0.5)."""
    mh, xh = ics.hernquist(8192, 25)
    cells = []
    for k in range(0, 8192, 32):
        cells.append(O.quad_moments(xh[k:k + 32], mh[k:k + 32]))
    xc = np.array([c[0] for c in cells])
    mc = np.array([c[1] for c in cells])
    qc = np.array([c[2] for c in cells])
    xi = xh[::97]
    eps = 0.05
    g5.set_eps_to_all(eps)
    g5.set_cells(xc, mc, qc)
    a, p = g5.cell_force_on_x(xi)
    oa, op = O.quad_force(f32(xi), np.float32(eps * eps), f32(xc), f32(mc), f32(qc))
    assert np.percentile(rel_err(a, oa), 99) < 1e-5
    # phi_p and phi_q cancel partly for near cells: FP32 sums, 5e-5 relative
    np.testing.assert_allclose(-p, op, rtol=5e-5)


def test_n_beyond_the_addresses_written(g5):
    """g5_set_n larger than the highest address set reads zero records, as the
    reference's zero-initialised j-memory does (PG5/newton/libpg5/
    phantom_g5.c:25-28): the force equals the one over the addresses set."""
    m, x, _ = ics.plummer(1000, 25)
    g5.set_eps_to_all(0.01)
    dev = 30  # a board no other test has written
    g5.set_xmj(0, x, m, devid=dev)
    g5.set_n(1000, devid=dev)
    a0, p0 = g5.calculate_force_on_x(x[:64], devid=dev)
    for n in (3000, 9000):  # within the first device buffer, then past it
        g5.set_n(n, devid=dev)
        a, p = g5.calculate_force_on_x(x[:64], devid=dev)
        assert np.all(np.isfinite(a)) and np.all(np.isfinite(p))
        assert rel_err(a, a0).max() < 1e-6
        assert rel_err(p, p0).max() < 1e-6


def test_rsqrt_bias_of_the_device_and_its_correction(g5):
    """The reference's rsqrt-bias rule (rsqrt.c:42-81) applied to the device's
    own rsqrt: the library's b equals the oracle's restatement of the rule on
    the v_rsq_f32 values the force kernel produces (read back as the
    potential of one unit mass at distance 0 with eps^2 = x); mode 1 scales
    the g5 outputs by float(1 - 3b) and float(1 - b) as phantom_g5.c:45-60,
    244-269 do."""
    import ctypes as C

    import torch
    L = g5.lib
    L.pg5_rsqrt_bias.restype = C.c_double
    b = L.pg5_rsqrt_bias()
    ys = []
    for x in O.rsqrt_args():
        n = len(x)
        xi = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        xi[:, 3] = torch.from_numpy(x).cuda()
        xj = torch.tensor([[0.0, 0.0, 0.0, 1.0]], dtype=torch.float32, device="cuda")
        acc = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        pot = torch.empty(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert L.pg5_dev_force(0, n, xi.data_ptr(), 1, xj.data_ptr(), None, acc.data_ptr(),
                               pot.data_ptr(), None) == 0
        torch.cuda.synchronize()
        ys.append(pot.cpu().numpy())
        del xi, acc, pot
    y0, y1 = ys
    assert np.all(y0 == y0.astype(np.float32))  # FP32 values, carried exactly
    bo = O.rsqrt_bias_of(y0, y1)
    print(f"device rsqrt bias {b:.3e} (oracle on the device's values {bo:.3e})")
    assert abs(b - bo) <= 1e-9 * abs(bo) + 1e-18, (b, bo)
    assert abs(b) < 1e-6
    # the correction on the g5 outputs
    m, x, _ = ics.plummer(512, 27)
    g5.set_eps_to_all(0.01)
    g5.set_xmj(0, x, m)
    g5.set_n(512)
    a0, p0 = g5.calculate_force_on_x(x[:64])
    assert L.pg5_set_rsqrt_bias_mode(1) == 0
    try:
        a1, p1 = g5.calculate_force_on_x(x[:64])
    finally:
        L.pg5_set_rsqrt_bias_mode(0)
    np.testing.assert_array_equal(a1, a0 * float(np.float32(1 - 3 * b)))
    np.testing.assert_array_equal(p1, p0 * float(np.float32(1 - b)))
