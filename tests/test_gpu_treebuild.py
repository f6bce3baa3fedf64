"""The GPU list builder (libpgtree_dev.so) against the host builder
(libpgtree.so): identical arrays, and identical forces through the batched
calls."""
import numpy as np
import pytest

import ics

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "src", "ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "coff", "ccnt", "order")


def dev(a):
    """Device copy, complete before return: the libraries run on their own
    streams, which do not wait for torch's."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    torch.cuda.synchronize()
    return t


def same_lists(h, d, cells):
    for k in FIELDS:
        np.testing.assert_array_equal(d[k], h[k], err_msg=k)
    assert d["work_pp"] == h["work_pp"] and d["work_pc"] == h["work_pc"]
    if cells:
        # cell numbering differs; the lists hold the same cell contents in order
        np.testing.assert_array_equal(d["cells"][d["clist"]], h["cells"][h["clist"]])
        assert len(d["cells"]) == len(h["cells"])
        np.testing.assert_array_equal(np.sort(d["cells"], axis=0), np.sort(h["cells"], axis=0))


@pytest.mark.parametrize("n,ncrit", [(1 << 16, 256), (5000, 64), (10, 256)])
def test_treepm_lists_match_host_builder(n, ncrit):
    import pgamd
    rcut = 3.0 / 64
    m, x = ics.clumpy_box(n, 11)
    h = pgamd.treepm_lists(x, rcut, ncrit)
    d_x = dev(x)
    L = pgamd.DevLists.treepm(d_x.data_ptr(), n, rcut, ncrit)
    d = L.host()
    L.free()
    same_lists(h, d, cells=False)


@pytest.mark.parametrize("n,ncrit,nleaf", [(1 << 15, 512, 8), (3000, 64, 4)])
def test_bh_lists_match_host_builder(n, ncrit, nleaf):
    import pgamd
    m, x = ics.hernquist(n, 12)
    h = pgamd.bh_lists(x, m, 0.5, ncrit, nleaf)
    d_x, d_m = dev(x), dev(m)
    L = pgamd.DevLists.bh(d_x.data_ptr(), d_m.data_ptr(), n, 0.5, ncrit, nleaf)
    d = L.host()
    L.free()
    same_lists(h, d, cells=True)


def test_coincident_particles_behave_as_host_builder():
    """More than ncrit particles at one point cannot be split by the 21-level
    key: both builders make that node one group of all its particles."""
    import pgamd
    rng = np.random.default_rng(5)
    x = np.vstack([rng.uniform(0.1, 0.9, (2000, 3)), np.full((300, 3), 0.5)])
    h = pgamd.treepm_lists(x, 0.05, 128)
    d_x = dev(x)
    L = pgamd.DevLists.treepm(d_x.data_ptr(), len(x), 0.05, 128)
    d = L.host()
    L.free()
    same_lists(h, d, cells=False)
    assert d["icnt"].sum() == 2300 and d["icnt"].max() == 300


@pytest.mark.parametrize("case", ["coincident", "n2", "n9"])
def test_bh_edge_cases_behave_as_host_builder(case):
    """Barnes-Hut walk on a node that the key cannot split (more than nleaf
    coincident particles: a walk that opens it lists its particles as a leaf,
    in both builders) and on tiny systems (the root a leaf, or one level of
    children)."""
    import pgamd
    rng = np.random.default_rng(6)
    if case == "coincident":
        x = np.vstack([rng.uniform(-1, 1, (2000, 3)), np.full((40, 3), 0.25)])
    else:
        x = rng.uniform(-1, 1, (int(case[1:]), 3))
    m = rng.uniform(0.5, 1.5, len(x)) / len(x)
    h = pgamd.bh_lists(x, m, 0.5, 64, 4)
    d_x, d_m = dev(x), dev(m)
    L = pgamd.DevLists.bh(d_x.data_ptr(), d_m.data_ptr(), len(x), 0.5, 64, 4)
    d = L.host()
    L.free()
    same_lists(h, d, cells=True)


def test_gpu_lists_give_identical_batch_forces():
    import torch
    import pgamd
    pp, pm = 0.1 / 32, 3.0 / 32
    n = 1 << 15
    m, x = ics.clumpy_box(n, 13)
    T = pgamd.G5Table()
    T.lib.pg5_gen_s2_force_table(pp, pm)
    _, xs = T.table()
    # host lists, packed on the host as bench.py does
    h = pgamd.treepm_lists(x, pm, 256)
    P = h["pos"]
    pos4 = np.zeros((len(P), 4), np.float32)
    pos4[:, :3] = P.astype(np.float32) * np.float32(xs)
    pos4[:, 3] = m[h["src"]]
    keys = ("ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "order")
    dh = {k: dev(h[k]) for k in keys}
    acc_h = torch.zeros((len(P), 3), dtype=torch.float64, device="cuda")
    h_pos4 = dev(pos4)
    ng = len(h["ioff"])
    assert T.lib.pg5t_batch_force(0, ng, *[dh[k].data_ptr() for k in keys], h_pos4.data_ptr(),
                                  None, acc_h.data_ptr(), 0, None) == 0
    torch.cuda.synchronize()
    # GPU lists, packed on the GPU
    d_x = dev(x)
    L = pgamd.DevLists.treepm(d_x.data_ptr(), n, pm, 256)
    i = L.info()
    d_m = dev(m)
    d_pos4 = torch.empty((i["np"], 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    L.pack(d_m.data_ptr(), xs, d_pos4.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_pos4.cpu().numpy(), pos4)
    acc_d = torch.zeros((i["np"], 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert T.lib.pg5t_batch_force(0, i["ngroups"], *[L.ptr(k) for k in keys], d_pos4.data_ptr(),
                                  None, acc_d.data_ptr(), 0, None) == 0
    torch.cuda.synchronize()
    L.free()
    np.testing.assert_array_equal(acc_d.cpu().numpy(), acc_h.cpu().numpy())


def test_bad_arguments_are_reported():
    import pgamd
    lib = pgamd.DevLists.lib()
    lib.pgamd_set_abort_on_error(0)
    try:
        assert lib.pgt_dev_treepm_build(0, None, 0.1, 64, None) is None
        assert lib.pgt_dev_treepm_build(10, None, 0.6, 64, None) is None
        assert lib.pgt_dev_bh_build(1, None, None, 0.5, 64, 8, None) is None
        assert pgamd.last_error(lib)[0] != 0
    finally:
        lib.pgamd_set_abort_on_error(1)
