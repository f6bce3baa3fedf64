"""Long-range (particle-mesh) side of TreePM on the GPU (pg5t_pm_force)
against the FP64 mesh restatement and, with the short-range table force,
against the Ewald sum of the S2-softened periodic force."""
import numpy as np
import pytest

import ics
import pm_oracle as P
from conftest import rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    torch.cuda.synchronize()
    return t


@pytest.fixture(scope="module")
def tab():
    import pgamd
    return pgamd.G5Table()


def pm_gpu(tab, x, m, mesh, a):
    import torch
    d_x, d_m = dev(x), dev(m)
    acc = torch.zeros((len(m), 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert tab.lib.pg5t_pm_force(len(m), d_x.data_ptr(), d_m.data_ptr(), mesh, a,
                                 acc.data_ptr(), None) == 0
    torch.cuda.synchronize()
    return acc.cpu().numpy()


@pytest.mark.parametrize("mesh", [32, 64])
def test_pm_matches_the_mesh_restatement(tab, mesh):
    m, x = ics.clumpy_box(3000, 40 + mesh)
    a = 3.0 / mesh
    g = pm_gpu(tab, x, m, mesh, a)
    o = P.pm_force(x, m, mesh, a)
    assert rel_err(g, o).max() < 1e-9
    # momentum: the TSC pair assignment/interpolation conserves it closely
    assert np.abs((m[:, None] * g).sum(0)).max() < 1e-6 * np.abs(m[:, None] * g).sum()


def test_mesh_force_does_not_depend_on_particle_order(tab):
    """The mass assignment sums in 64-bit fixed point, so the order in which
    particles (and their atomics) arrive cannot change a bit: the same
    particles in another order get the same accelerations, bit for bit, and
    a second call repeats the first."""
    m, x = ics.clumpy_box(20000, 77)
    mesh, a = 64, 3.0 / 64
    g = pm_gpu(tab, x, m, mesh, a)
    np.testing.assert_array_equal(pm_gpu(tab, x, m, mesh, a), g)
    perm = np.random.default_rng(5).permutation(len(m))
    gp = pm_gpu(tab, x[perm], m[perm], mesh, a)
    np.testing.assert_array_equal(gp, g[perm])


def test_treepm_total_force_is_the_periodic_force(tab):
    """Table (short range) + mesh (long range) = Ewald sum of the softened
    force, to the mesh's discretisation error at r_cut = 3 cells."""
    import torch
    import pgamd
    mesh = 32
    pp, pm = 0.1 / mesh, 3.0 / mesh
    rng = np.random.default_rng(7)
    n = 192
    x = rng.uniform(0, 1, (n, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    # short range through the batched table path on GPU-built lists
    tab.lib.pg5_gen_s2_force_table(pp, pm)
    _, xs = tab.table()
    d_x, d_m = dev(x), dev(m)
    L = pgamd.DevLists.treepm(d_x.data_ptr(), n, pm, 32)
    i = L.info()
    d_pos4 = torch.empty((i["np"], 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    L.pack(d_m.data_ptr(), xs, d_pos4.data_ptr())
    torch.cuda.synchronize()
    acc = torch.zeros((i["np"], 3), dtype=torch.float64, device="cuda")
    keys = ("ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "order")
    assert tab.lib.pg5t_batch_force(0, i["ngroups"], *[L.ptr(k) for k in keys],
                                    d_pos4.data_ptr(), None, acc.data_ptr(), 0, None) == 0
    torch.cuda.synchronize()
    src = L.host()["src"]
    L.free()
    short = np.zeros((n, 3))
    short[src[:n]] = acc.cpu().numpy()[:n]
    total = short + pm_gpu(tab, x, m, mesh, pm)
    # exact: Ewald of point masses minus the S2 softening inside eps_PP
    exact = P.pm_exact(x, m, pp)
    e = rel_err(total, exact)
    assert np.median(e) < 3e-2 and np.percentile(e, 90) < 8e-2, (np.median(e), e.max())
    tab.lib.pg5_gen_s2_force_table(0.1 / 64, 3.0 / 64)  # back to the default table


def test_bad_arguments_are_reported(tab):
    import pgamd
    tab.lib.pgamd_set_abort_on_error(0)
    try:
        assert tab.lib.pg5t_pm_force(10, None, None, 33, 0.1, None, None) == -1
        assert pgamd.last_error(tab.lib)[0] != 0
    finally:
        tab.lib.pgamd_set_abort_on_error(1)
