"""Parity at the benchmark configurations' full sizes, through properties that
do not depend on size (total momentum of a closed system, the direct-sum
oracle on a seeded sample, run-to-run bit equality):

* PG6 Hermite step, Plummer N = 65536, eps = 1/256 (all i against all j);
* TreePM short range, clumpy periodic box N = 2^20, r_cut = 3/128, lists
  built on the GPU;
* Barnes-Hut with quadrupole cells, Hernquist N = 2^21, theta = 0.5, lists
  built on the GPU;
* 4096 independent clusters, N log-uniform in [64, 4096].
Each case prints its name first, so a failure names itself."""
import numpy as np
import pytest

import ics
import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    torch.cuda.synchronize()
    return t


def momentum_ratio(m, a):
    """|sum m a| / sum m |a|: zero for exact pairwise-antisymmetric forces."""
    return np.linalg.norm((m[:, None] * a).sum(0)) / (m * np.linalg.norm(a, axis=1)).sum()


def test_pg6_step_n65536():
    print("case: PG6 N=65536 eps=1/256", flush=True)
    import pgamd
    g6 = pgamd.G6()
    n, eps = 65536, 1.0 / 256
    m, x, v = ics.plummer(n, 65536)
    idx = np.arange(n, dtype=np.int32)
    z = np.zeros((n, 3))
    g6.set_j_bulk(x, v, z, z, m, np.zeros(n), idx)
    g6.set_ti(0.0)
    r = g6.calc(n, x, v, idx, eps * eps, variant="2p")
    again = g6.calc(n, x, v, idx, eps * eps, variant="2p")
    for k in ("acc", "jrk", "pot", "nnb", "rnnb"):
        np.testing.assert_array_equal(r[k], again[k], err_msg=k)
    # a closed system: total force and total jerk vanish
    assert momentum_ratio(m, r["acc"]) < 1e-6
    assert momentum_ratio(m, r["jrk"]) < 1e-5
    sel = np.sort(np.random.default_rng(7).choice(n, 1024, replace=False))
    vf = v.astype(np.float32).astype(np.float64)
    o = O.pg6_force(x[sel], vf[sel], idx[sel], eps * eps, x, vf, m, idx)
    assert np.percentile(rel_err(r["acc"][sel], o["acc"]), 99) < 1e-6
    assert np.percentile(rel_err(r["jrk"][sel], o["jrk"]), 99) < 1e-5
    assert rel_err(r["pot"][sel], o["pot"]).max() < 1e-6
    # each large-i kernel at full size, the sorted chunk-origin one included
    for k in (2, 4, 5):
        assert g6.lib.pg6_dev_set_kernel(k) == 0
        rk = g6.calc(n, x, v, idx, eps * eps, variant="2p")
        g6.lib.pg6_dev_set_kernel(-1)
        e = rel_err(rk["acc"][sel], o["acc"])
        print(f"kernel {k}: acc p99 {np.percentile(e, 99):.3e} max {e.max():.3e}")
        assert np.percentile(e, 99) < 1e-6, (k, np.percentile(e, 99))
        assert np.percentile(rel_err(rk["jrk"][sel], o["jrk"]), 99) < 1e-5
        assert rel_err(rk["pot"][sel], o["pot"]).max() < 1e-6
        assert momentum_ratio(m, rk["acc"]) < 1e-6


def test_treepm_short_range_n2p20():
    print("case: TreePM N=2^20 r_cut=3/128", flush=True)
    import torch

    import pgamd
    pp, pm = 0.1 / 128, 3.0 / 128
    m, x = ics.clumpy_box(1 << 20, 20)
    d_x, d_m = dev(x), dev(m)
    T = pgamd.G5Table()
    T.lib.pg5_set_table_frc_bit(6)
    T.lib.pg5_gen_s2_force_table(pp, pm)
    _, xs = T.table()
    DL = pgamd.DevLists.treepm(d_x.data_ptr(), len(x), pm, 256)
    try:
        L = DL.host()
        P = L["pos"]
        mm = m[L["src"]]
        d_pos = torch.empty((len(P), 4), dtype=torch.float32, device="cuda")
        DL.pack(d_m.data_ptr(), xs, d_pos.data_ptr(), None, None)
        torch.cuda.synchronize()
        d = {k: DL.ptr(k) for k in ("ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "order")}
        acc = torch.zeros((len(P), 3), dtype=torch.float64, device="cuda")
        ng = len(L["ioff"])
        lists = (d["ioff"], d["icnt"], d["ilist"], d["joff"], d["jcnt"], d["jlist"], d["order"])
        args = (0, ng, *lists, d_pos.data_ptr(), None, acc.data_ptr(), 0, None)
        assert T.lib.pg5t_batch_force(*args) == 0
        torch.cuda.synchronize()
        a = acc.cpu().numpy()[: len(x)]
        assert T.lib.pg5t_batch_force(*args) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(acc.cpu().numpy()[: len(x)], a)
        # FP64 positions relative to each group's origin (the list builder's
        # positions in HBM), table and learned cut-off
        acc64 = torch.zeros((len(P), 3), dtype=torch.float64, device="cuda")
        assert T.lib.pg5t_batch_force(0, ng, *lists, d_pos.data_ptr(), DL.ptr("pos"),
                                      acc64.data_ptr(), 0, None) == 0
        torch.cuda.synchronize()
        a64 = acc64.cpu().numpy()
        tab, _ = T.table()
        T.mlp_train(iters=150, starts=3, refine=2500, seed=7)
        assert T.lib.pg5t_batch_force(0, ng, *lists, d_pos.data_ptr(), DL.ptr("pos"),
                                      acc64.data_ptr(), 1, None) == 0
        torch.cuda.synchronize()
        a_net = acc64.cpu().numpy()
    finally:
        DL.free()
        T.lib.pg5_set_table_frc_bit(5)
    # every pair inside r_cut sits in both particles' lists, periodic images
    # included: the total short-range force of the box vanishes (up to the
    # FP32 rounding of image positions, which differs for the two partners)
    assert momentum_ratio(m, a) < 3e-5
    sel = np.sort(np.random.default_rng(8).choice(len(x), 128, replace=False))
    f = O.table_force_f64(P[sel], P, mm, pp, pm)
    e = rel_err(a[sel], f, floor=1e-3 * np.linalg.norm(f, axis=1).max())
    assert np.median(e) < 3e-4 and np.percentile(e, 99) < 3e-3
    # the group-origin path: the reference's pair terms (the oracle's
    # bit-exact emulation of gravity_kernel_avx2.c:22-137 on x - o over each
    # sampled particle's own list) up to summation order, and against the
    # closed form the 1024-entry table's own interpolation error
    pos_in = np.full(len(P), -1, np.int64)  # place of each particle in the i-lists
    pos_in[L["ilist"]] = np.arange(len(L["ilist"]))
    emul = np.zeros((len(sel), 3))
    for q, i in enumerate(sel):
        g = np.searchsorted(L["ioff"], pos_in[i], side="right") - 1
        il = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
        jl = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
        o = P[il[0]]
        emul[q] = O.table_force_emul(P[i:i + 1] - o, P[jl] - o, mm[jl], tab, xs, 4, 6)[0]
    fl = 1e-3 * np.linalg.norm(f, axis=1).max()  # floor for near-zero short-range forces
    e = rel_err(a64[sel], emul, floor=fl)
    assert np.percentile(e, 99) < 2e-6 and e.max() < 1e-5, (np.percentile(e, 99), e.max())
    e = rel_err(a64[sel], f, floor=fl)
    assert np.percentile(e, 99) < 3e-3
    # the learned cut-off at full size
    e = rel_err(a_net[sel], f, floor=fl)
    print(f"learned cut-off N=2^20: median {np.median(e):.3e} p99 {np.percentile(e, 99):.3e}")
    assert np.median(e) < 2e-6 and np.percentile(e, 99) < 1e-5, (np.median(e),
                                                                 np.percentile(e, 99))


def test_barnes_hut_n2p21():
    print("case: Barnes-Hut N=2^21 theta=0.5", flush=True)
    import torch

    import pgamd
    n, eps = 1 << 21, 1.0 / 128
    m, x = ics.hernquist(n, 21)
    d_x, d_m = dev(x), dev(m)
    g5 = pgamd.G5()
    DL = pgamd.DevLists.bh(d_x.data_ptr(), d_m.data_ptr(), n, 0.5, 512, 8)
    try:
        L = DL.host()
        P = L["pos"]
        mm = m[L["src"]]
        keys = ("ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "coff", "ccnt", "clist",
                "order")
        d = {k: DL.ptr(k) for k in keys}
        d_pos = torch.empty((len(P), 4), dtype=torch.float32, device="cuda")
        d_cells = torch.empty((len(L["cells"]), 12), dtype=torch.float32, device="cuda")
        DL.pack(d_m.data_ptr(), 1.0, d_pos.data_ptr(), d_cells.data_ptr(), None)
        torch.cuda.synchronize()
        acc = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        pot = torch.zeros(n, dtype=torch.float64, device="cuda")
        args = (0, len(L["ioff"]), *[d[k] for k in keys], d_pos.data_ptr(), d_cells.data_ptr(),
                eps * eps, acc.data_ptr(), pot.data_ptr(), None)
        assert g5.lib.pg5_bh_batch(*args) == 0
        torch.cuda.synchronize()
        a = acc.cpu().numpy()
        p = pot.cpu().numpy()
        assert g5.lib.pg5_bh_batch(*args) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(acc.cpu().numpy(), a)
        P4 = d_pos.cpu().numpy().astype(np.float64)
        Cc = d_cells.cpu().numpy().astype(np.float64)
    finally:
        DL.free()
    # against the FP64 oracle over the same lists (the caller contract of
    # PG5/newton_quad/source.cpp:845-867: particle part through the g5 calls,
    # cells through the g5c calls) on the FP32 inputs the kernel reads
    e2 = float(np.float32(eps * eps))
    for g in np.random.default_rng(10).choice(len(L["ioff"]), 4, replace=False):
        il = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
        pl = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
        cl = L["clist"][L["coff"][g]:L["coff"][g] + L["ccnt"][g]]
        ap, pp = O.pg5_force(P4[il, :3], e2, P4[pl, :3], P4[pl, 3])
        aq, pq = O.quad_force(P4[il, :3], e2, Cc[cl, :3], Cc[cl, 3], Cc[cl, 4:10])
        assert np.percentile(rel_err(a[il], ap + aq), 99) < 1e-5
        assert rel_err(p[il], pp + pq).max() < 1e-5
    # against the FP64 direct sum over all 2^21 particles: the tree's own
    # error at theta = 0.5 with quadrupoles
    sel = np.sort(np.random.default_rng(9).choice(n, 96, replace=False))
    ad, _ = O.pg5_force(P[sel], eps * eps, P, mm)
    e = rel_err(a[sel], ad)
    assert np.median(e) < 1e-3 and np.percentile(e, 99) < 5e-3


def test_clusters_4096():
    print("case: 4096 clusters N in [64, 4096]", flush=True)
    import torch

    import pgamd
    sizes = ics.cluster_sizes(4096, 4096)
    xs, vs, ms = [], [], []
    for c, nc in enumerate(sizes):
        mc, xc, vc = ics.plummer(int(nc), 50000 + c)
        xs.append(xc)
        vs.append(vc)
        ms.append(mc)
    X, V, Mm = np.concatenate(xs), np.concatenate(vs), np.concatenate(ms)
    coff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    e2 = ((4.0 / sizes) ** 2).astype(np.float32)
    g6 = pgamd.G6()
    items = g6.cluster_items(coff)
    d = dict(items=dev(items), coff=dev(coff), x=dev(X), v=dev(V), m=dev(Mm), e2=dev(e2))
    nt = len(Mm)
    acc = torch.zeros((nt, 3), dtype=torch.float64, device="cuda")
    jrk = torch.zeros((nt, 3), dtype=torch.float64, device="cuda")
    pot = torch.zeros(nt, dtype=torch.float64, device="cuda")
    args = (len(items) // 2, d["items"].data_ptr(), d["coff"].data_ptr(), d["x"].data_ptr(),
            d["v"].data_ptr(), d["m"].data_ptr(), d["e2"].data_ptr(), acc.data_ptr(),
            jrk.data_ptr(), pot.data_ptr(), None)
    assert g6.lib.pg6_cluster_force(*args) == 0
    torch.cuda.synchronize()
    a, j, p = acc.cpu().numpy(), jrk.cpu().numpy(), pot.cpu().numpy()
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(j)) and np.all(p < 0)
    # each cluster is closed: its total force vanishes
    worst = max(momentum_ratio(Mm[coff[c]:coff[c + 1]], a[coff[c]:coff[c + 1]])
                for c in range(len(sizes)))
    assert worst < 1e-5
    # whole clusters against the FP64 oracle, small to largest
    order = np.argsort(sizes)
    for c in (order[0], order[len(order) // 2], order[-1]):
        s = slice(coff[c], coff[c + 1])
        idx = np.arange(coff[c + 1] - coff[c], dtype=np.int32)
        vf = V[s].astype(np.float32).astype(np.float64)
        o = O.pg6_force(X[s], vf, idx, float(e2[c]), X[s], vf, Mm[s], idx)
        assert np.percentile(rel_err(a[s], o["acc"]), 99) < 1e-6
        assert np.percentile(rel_err(j[s], o["jrk"]), 99) < 1e-5
        assert rel_err(p[s], o["pot"]).max() < 1e-6
