"""Batched GPU paths (one persistent launch over many groups / clusters),
their hipGraph replays and autotuners, and the learned cut-off: parity with
the per-group drop-in calls, the CPU oracle and the closed forms."""
import numpy as np
import pytest

import ics
import mlp_oracle as M
import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

PP, PM = 0.1 / 128, 3.0 / 128


def dev(a, dtype=None):
    """Device copy, complete before return: the libraries run on their own
    streams, which do not wait for torch's."""
    import torch
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    t = torch.from_numpy(a).cuda()
    torch.cuda.synchronize()
    return t


def ptr(t):
    return t.data_ptr()


@pytest.fixture(scope="module")
def tab():
    import pgamd
    t = pgamd.G5Table()
    t.lib.pg5_gen_s2_force_table(PP, PM)
    return t


@pytest.fixture(scope="module")
def treepm_case(tab):
    import pgamd
    m, x = ics.clumpy_box(1 << 15, 41)
    L = pgamd.treepm_lists(x, PM, 256)
    _, xs = tab.table()
    P = L["pos"]
    mm = m[L["src"]]
    pos4 = np.zeros((len(P), 4), np.float32)
    pos4[:, :3] = (P.astype(np.float32) * np.float32(xs))
    pos4[:, 3] = mm
    d = {k: dev(L[k]) for k in ("ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "order")}
    d["pos"] = dev(pos4)
    d["pos64"] = dev(P)
    return dict(L=L, m=mm, pos4=pos4, xs=xs, d=d, n=len(x))


def run_treepm(tab, c, mode, acc=None, p64=False):
    """Batched short-range force; p64: separations from the FP64 positions
    relative to each group's origin."""
    import torch
    d = c["d"]
    if acc is None:
        acc = torch.zeros((len(c["pos4"]), 3), dtype=torch.float64, device="cuda")
    ng = len(c["L"]["ioff"])
    rc = tab.lib.pg5t_batch_force(0, ng, ptr(d["ioff"]), ptr(d["icnt"]), ptr(d["ilist"]),
                                  ptr(d["joff"]), ptr(d["jcnt"]), ptr(d["jlist"]), ptr(d["order"]),
                                  ptr(d["pos"]), ptr(d["pos64"]) if p64 else None, ptr(acc),
                                  mode, None)
    assert rc == 0
    torch.cuda.synchronize()
    return acc.cpu().numpy()


def p99(a, f):
    """99th percentile of the per-particle relative error, relative to
    max(|f|, 1e-3 max |f|): in this 2^15-particle box the uniform background
    has about one neighbour inside r_cut, so some short-range forces are
    near zero (or exactly zero) and a bare relative error says nothing."""
    return np.percentile(rel_err(a, f, floor=1e-3 * np.linalg.norm(f, axis=1).max()), 99)


def group_lists(L, g):
    il = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
    jl = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
    return il, jl


@pytest.mark.parametrize("frc", [5, 9])
def test_treepm_group_origin_matches_emulation(tab, treepm_case, frc):
    """FP64 positions taken relative to each group's origin (its first i):
    the pair terms are the reference's (gravity_kernel_avx2.c:22-137 through
    the oracle's bit-exact emulation) on x - o, so the batch equals the
    emulation up to summation order; and the separations keep their digits,
    so the error against the closed-form S2 drops to the table's own."""
    c = treepm_case
    L, P = c["L"], c["L"]["pos"]
    assert tab.lib.pg5_set_table_frc_bit(frc) == 0
    try:
        t, xs = tab.table()
        a = run_treepm(tab, c, 0, p64=True)
        rng = np.random.default_rng(frc)
        for g in rng.choice(len(L["ioff"]), 8, replace=False):
            il, jl = group_lists(L, g)
            o = P[il[0]]
            ref = O.table_force_emul(P[il] - o, P[jl] - o, c["m"][jl], t, xs, 4, frc)
            scale = np.linalg.norm(ref, axis=1).max()
            assert rel_err(a[il], ref, floor=1e-3 * scale).max() < 1e-6
        sel = rng.choice(c["n"], 400, replace=False)
        f = O.table_force_f64(P[sel], P, c["m"], PP, PM)
        e = p99(a[sel], f)
        print(f"frc {frc} group origin vs closed form: p99 {e:.3e}")
        assert e < (3e-3 if frc == 5 else 4e-5), e
    finally:
        tab.lib.pg5_set_table_frc_bit(5)


def test_treepm_batch_table_matches_dropin_and_closed_form(tab, treepm_case):
    import torch
    c = treepm_case
    L = c["L"]
    a = run_treepm(tab, c, 0)
    # the drop-in kernel on the same scaled positions, group by group
    rng = np.random.default_rng(0)
    for g in rng.choice(len(L["ioff"]), 6, replace=False):
        il = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
        jl = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
        xi = dev(c["pos4"][il] * np.array([1, 1, 1, 0], np.float32))
        xj = dev(c["pos4"][jl])
        out = torch.zeros((len(il), 3), dtype=torch.float64, device="cuda")
        assert tab.lib.pg5t_dev_force(0, len(il), ptr(xi), len(jl), ptr(xj), ptr(out), None) == 0
        torch.cuda.synchronize()
        ref = out.cpu().numpy()
        scale = np.linalg.norm(ref, axis=1).max()
        assert rel_err(a[il], ref, floor=1e-3 * scale).max() < 1e-6
    # against the closed-form S2 force (FP64) on a sample of i
    sel = rng.choice(c["n"], 300, replace=False)
    f = O.table_force_f64(L["pos"][sel], L["pos"], c["m"], PP, PM)
    scale = np.linalg.norm(f, axis=1).max()
    e = rel_err(a[sel], f, floor=1e-3 * scale)
    assert np.median(e) < 3e-3


def test_treepm_graph_replay(tab, treepm_case):
    import torch
    c = treepm_case
    d = c["d"]
    ng = len(c["L"]["ioff"])
    eager = run_treepm(tab, c, 0)
    acc = torch.zeros((len(c["pos4"]), 3), dtype=torch.float64, device="cuda")
    args = (0, ng, ptr(d["ioff"]), ptr(d["icnt"]), ptr(d["ilist"]), ptr(d["joff"]),
            ptr(d["jcnt"]), ptr(d["jlist"]), ptr(d["order"]), ptr(d["pos"]), None, ptr(acc), 0)
    h = tab.lib.pg5t_batch_graph(*args)
    assert h
    try:
        for rep in range(3):
            acc.zero_()
            assert tab.lib.pg5t_batch_graph_launch(h, None) == 0
            torch.cuda.synchronize()
            np.testing.assert_array_equal(acc.cpu().numpy(), eager)
            run_treepm(tab, c, 0)  # eager work between replays
    finally:
        tab.lib.pg5t_batch_graph_destroy(h)


def test_treepm_block_sizes_and_autotune(tab, treepm_case):
    """Each block size of the table kernel (256, 512 or 1024 threads, each
    also the j-tile staged in LDS) matches the oracle's bit-exact emulation
    of the reference's pair terms up to summation order; the autotuner
    searches block x blocks-per-CU, reports the pair it kept, and a call at
    that shape equals a call at the same shape set by hand."""
    import pgamd
    c = treepm_case
    L, P = c["L"], c["L"]["pos"]
    t, xs = tab.table()
    out = {}
    shape = np.zeros(2, np.int32)
    try:
        for blk in (256, 512, 1024):
            assert tab.lib.pg5t_batch_set_block(blk) == 0
            assert tab.lib.pg5t_batch_launch_shape(0, pgamd.iptr(shape)) == 0
            assert shape[0] == blk
            out[blk] = a = run_treepm(tab, c, 0, p64=True)
            for g in np.random.default_rng(blk).choice(len(L["ioff"]), 6, replace=False):
                il, jl = group_lists(L, g)
                o = P[il[0]]
                ref = O.table_force_emul(P[il] - o, P[jl] - o, c["m"][jl], t, xs, 4, 5)
                scale = np.linalg.norm(ref, axis=1).max()
                if scale == 0:  # no j inside the cut-off
                    assert np.all(a[il] == 0), (blk, int(g))
                    continue
                e = rel_err(a[il], ref, floor=1e-3 * scale).max()
                assert e < 1e-6, (blk, int(g), len(il), e)
        d = c["d"]
        import torch
        acc = torch.zeros((len(c["pos4"]), 3), dtype=torch.float64, device="cuda")
        args = (0, len(L["ioff"]), ptr(d["ioff"]), ptr(d["icnt"]), ptr(d["ilist"]),
                ptr(d["joff"]), ptr(d["jcnt"]), ptr(d["jlist"]), ptr(d["order"]), ptr(d["pos"]),
                ptr(d["pos64"]), ptr(acc), 0)
        best = np.zeros(1, np.int32)
        ms = np.zeros(8, np.float32)
        assert tab.lib.pg5t_batch_autotune(*args, 2, pgamd.iptr(best), pgamd.fptr(ms)) == 0
        assert tab.lib.pg5t_batch_launch_shape(0, pgamd.iptr(shape)) == 0
        print(f"treepm batch autotune: block {shape[0]}, {shape[1]} blocks per CU, "
              f"ms {np.round(ms, 3).tolist()}")
        assert shape[0] in out and shape[1] == best[0] and np.all(ms > 0)
        np.testing.assert_array_equal(run_treepm(tab, c, 0, p64=True), out[int(shape[0])])
        tab.lib.pgamd_set_abort_on_error(0)
        assert tab.lib.pg5t_batch_set_block(300) == -1
        assert tab.lib.pg5t_batch_launch_shape(2, pgamd.iptr(shape)) == -1
    finally:
        tab.lib.pgamd_set_abort_on_error(1)
        tab.lib.pg5t_batch_set_block(0)
        tab.lib.pg5t_batch_set_blocks_per_cu(0, 0)


def test_mlp_normal_equations_match_oracle(tab):
    """The FP64 training kernels (per-sample forward + backward, J^T J in a
    fixed order) against the oracle's Jacobian, itself pinned by finite
    differences."""
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.5, M.NP)
    v = rng.uniform(0, 1, 3000)
    t = rng.normal(0, 1, 3000)
    wt = rng.uniform(0.5, 2, 3000)
    jtj, jtr, rtr = tab.mlp_normal_eq(v, t, wt, p)
    ojtj, ojtr, ortr = M.normal_eq(p, v, t, wt)
    assert abs(rtr - ortr) < 1e-12 * ortr
    assert np.max(np.abs(jtr - ojtr)) < 1e-11 * np.max(np.abs(ojtr))
    assert np.max(np.abs(jtj - ojtj)) < 1e-11 * np.max(np.abs(ojtj))
    # deterministic: the same bits on a second call
    jtj2, jtr2, rtr2 = tab.mlp_normal_eq(v, t, wt, p)
    np.testing.assert_array_equal(jtj2, jtj)
    assert rtr2 == rtr


def learned_f(tab, r):
    """The learned shape at physical radii through the GPU evaluator."""
    import torch
    dr = dev(np.asarray(r, np.float32))
    out = torch.empty_like(dr)
    assert tab.lib.pg5_mlp_eval_dev(len(r), ptr(dr), ptr(out)) == 0
    return out.cpu().numpy().astype(np.float64)


@pytest.fixture(scope="module")
def trained(tab):
    """The S2 network trained on the GPU (Levenberg-Marquardt, FP64)."""
    loss = tab.mlp_train(iters=150, starts=3, refine=2500, seed=7)
    return loss, tab.mlp_params()


def test_mlp_training_and_learned_force(tab, treepm_case, trained):
    loss, (p, form, rc, r0) = trained
    assert form == 1 and abs(rc - PM) < 1e-15 and abs(r0 - PP) < 1e-15
    # the fit: the network reproduces H(u) = R(u, 1) of the S2 shape (the
    # oracle's forward pass of the trained FP64 parameters)
    u = np.linspace(0, 1, 4001)
    H = M.s2_unit(u, O.s2_reff)
    # in the fit's weighting u^3 + 0.2 (a pair's error relative to its
    # Newtonian force is u^3 times the network's error outside the core)
    e = np.max(np.abs(M.forward(p, u) - H) * (u ** 3 + 0.2))
    print(f"trained: loss {loss:.3e} max |y - H| (u^3 + 0.2) {e:.3e}")
    assert e < 3e-6 and loss < 1e-12
    # the GPU evaluator decodes it into the cut-off shape f(r): compared with
    # the closed form relative to the pair's Newtonian 1/r^3
    r = np.geomspace(PP / 20, 0.999 * PM, 3000)
    f = learned_f(tab, r)
    fo = O.s2_force(PP, PM, r)
    e = np.abs(f - fo) * r ** 3
    print(f"learned f vs closed form (x r^3): max {np.max(e):.3e} median {np.median(e):.3e}")
    assert np.max(e) < 3e-4 and np.median(e) < 2e-5, (np.max(e), np.median(e))
    assert np.all(learned_f(tab, [0.0, 1.001 * PM, 1.01 * PM]) == 0.0)
    # forces through the batch kernel with the learned cut-off, FP64 group
    # origins: per-particle error against the closed-form S2
    c = treepm_case
    a = run_treepm(tab, c, 1, p64=True)
    L = c["L"]
    sel = np.random.default_rng(1).choice(c["n"], 400, replace=False)
    fa = O.table_force_f64(L["pos"][sel], L["pos"], c["m"], PP, PM)
    e = p99(a[sel], fa)
    print(f"learned force vs closed form: p99 {e:.3e}")
    assert e < 2e-5, e


def test_trained_network_gains_over_the_table_and_its_start(tab, treepm_case, trained):
    """What the learned cut-off gains: with FP64 group origins, the trained
    network (321 parameters) is more accurate than a table of the same
    parameter budget (frc_bit 3: 128 entries x {value, slope} = 256 floats)
    and than the network before training (one damped step from its seeded
    start), by the p99 of the per-particle force error against the closed-form
    S2."""
    c = treepm_case
    L = c["L"]
    sel = np.random.default_rng(2).choice(c["n"], 400, replace=False)
    fa = O.table_force_f64(L["pos"][sel], L["pos"], c["m"], PP, PM)
    _, (p, form, rc, r0) = trained
    e_net = p99(run_treepm(tab, c, 1, p64=True)[sel], fa)
    assert tab.lib.pg5_set_table_frc_bit(3) == 0
    try:
        e_tab = p99(run_treepm(tab, c, 0, p64=True)[sel], fa)
    finally:
        tab.lib.pg5_set_table_frc_bit(5)
    tab.mlp_train(iters=1, starts=1, refine=0, seed=7)
    try:
        e_start = p99(run_treepm(tab, c, 1, p64=True)[sel], fa)
    finally:
        tab.mlp_set_params(p, form, rc, r0)
    print(f"p99: network {e_net:.3e}, 128-entry table {e_tab:.3e}, untrained {e_start:.3e}")
    assert e_net < e_tab / 10 and e_net < e_start / 100, (e_net, e_tab, e_start)


@pytest.fixture(scope="module")
def bh_case():
    import pgamd
    g5 = pgamd.G5()
    mh, xh = ics.hernquist(1 << 15, 51)
    L = pgamd.bh_lists(xh, mh, 0.5, 512, 8)
    P = L["pos"]
    mm = mh[L["src"]]
    pos4 = np.zeros((len(P), 4), np.float32)
    pos4[:, :3] = P
    pos4[:, 3] = mm
    cells = np.zeros((len(L["cells"]), 12), np.float32)
    cells[:, :10] = L["cells"]
    d = {k: dev(L[k]) for k in ("ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "coff",
                                "ccnt", "clist", "order")}
    d["pos"] = dev(pos4)
    d["cells"] = dev(cells)
    return dict(g5=g5, L=L, m=mm, pos4=pos4, cells=cells, d=d)


def run_bh(c, eps):
    import torch
    d = c["d"]
    n = len(c["pos4"])
    acc = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    pot = torch.zeros(n, dtype=torch.float64, device="cuda")
    args = (0, len(c["L"]["ioff"]), ptr(d["ioff"]), ptr(d["icnt"]), ptr(d["ilist"]),
            ptr(d["joff"]), ptr(d["jcnt"]), ptr(d["jlist"]), ptr(d["coff"]), ptr(d["ccnt"]),
            ptr(d["clist"]), ptr(d["order"]), ptr(d["pos"]), ptr(d["cells"]), eps * eps,
            ptr(acc), ptr(pot))
    return args, acc, pot


def test_bh_batch_matches_oracle_and_graph(bh_case):
    import torch
    c = bh_case
    L = c["L"]
    eps = 1.0 / 128
    args, acc, pot = run_bh(c, eps)
    g5 = c["g5"]
    assert g5.lib.pg5_bh_batch(*args, None) == 0
    torch.cuda.synchronize()
    a = acc.cpu().numpy()
    p = pot.cpu().numpy()
    f32 = lambda z: np.asarray(z, np.float32).astype(np.float64)  # noqa: E731
    P4 = c["pos4"].astype(np.float64)
    Cc = c["cells"].astype(np.float64)
    for g in np.random.default_rng(2).choice(len(L["ioff"]), 5, replace=False):
        il = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
        pl = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
        cl = L["clist"][L["coff"][g]:L["coff"][g] + L["ccnt"][g]]
        ap, pp = O.pg5_force(P4[il, :3], f32(eps * eps), P4[pl, :3], P4[pl, 3])
        aq, pq = O.quad_force(P4[il, :3], f32(eps * eps), Cc[cl, :3], Cc[cl, 3], Cc[cl, 4:10])
        assert np.percentile(rel_err(a[il], ap + aq), 99) < 1e-5
        assert rel_err(p[il], pp + pq).max() < 1e-5
        # and the tree force is close to the direct sum (theta = 0.5 + quadrupole)
        ad, _ = O.pg5_force(L["pos"][il], eps * eps, L["pos"], c["m"])
        assert np.median(rel_err(a[il], ad)) < 1e-3
    h = g5.lib.pg5_bh_graph(*args)
    try:
        for rep in range(3):
            acc.zero_()
            assert g5.lib.pg5_bh_graph_launch(h, None) == 0
            torch.cuda.synchronize()
            np.testing.assert_array_equal(acc.cpu().numpy(), a)
            assert g5.lib.pg5_bh_batch(*args, None) == 0
    finally:
        g5.lib.pg5_bh_graph_destroy(h)
    import pgamd
    best = np.zeros(1, np.int32)
    assert g5.lib.pg5_bh_autotune(*args, 2, pgamd.iptr(best), None) == 0
    assert best[0] in (1, 2, 3, 4, 6, 8, 12, 16)
    g5.lib.pg5_bh_set_launch(0, 0)


def test_bh_block_sizes_and_autotune(bh_case):
    """Each block size of the tree-list kernel (256, 512 or 1024 threads; an
    LDS tile holds that many particles or half as many cells; groups of up to
    512 i run in passes at 256) matches the oracle over the same lists; the
    autotuner searches block x blocks-per-CU, reports the pair it kept, and a
    call at that shape equals a call at the same shape set by hand."""
    import torch

    import pgamd
    c = bh_case
    L = c["L"]
    g5 = c["g5"]
    eps = 1.0 / 128
    args, acc, pot = run_bh(c, eps)
    f32 = lambda z: np.asarray(z, np.float32).astype(np.float64)  # noqa: E731
    P4 = c["pos4"].astype(np.float64)
    Cc = c["cells"].astype(np.float64)
    groups = np.random.default_rng(4).choice(len(L["ioff"]), 4, replace=False)
    big = np.argmax(L["icnt"])
    assert L["icnt"][big] > 256  # a group that takes two passes at block 256
    groups = np.append(groups, big)
    out = {}
    shape = np.zeros(2, np.int32)
    try:
        for blk in (256, 512, 1024):
            assert g5.lib.pg5_bh_set_launch(blk, -1) == 0
            assert g5.lib.pg5_bh_launch_shape(pgamd.iptr(shape)) == 0
            assert shape[0] == blk
            assert g5.lib.pg5_bh_batch(*args, None) == 0
            torch.cuda.synchronize()
            a, p = acc.cpu().numpy(), pot.cpu().numpy()
            out[blk] = (a, p)
            for g in groups:
                il = L["ilist"][L["ioff"][g]:L["ioff"][g] + L["icnt"][g]]
                pl = L["jlist"][L["joff"][g]:L["joff"][g] + L["jcnt"][g]]
                cl = L["clist"][L["coff"][g]:L["coff"][g] + L["ccnt"][g]]
                ap, pp = O.pg5_force(P4[il, :3], f32(eps * eps), P4[pl, :3], P4[pl, 3])
                aq, pq = O.quad_force(P4[il, :3], f32(eps * eps), Cc[cl, :3], Cc[cl, 3],
                                      Cc[cl, 4:10])
                assert np.percentile(rel_err(a[il], ap + aq), 99) < 1e-5, blk
                assert rel_err(p[il], pp + pq).max() < 1e-5, blk
        best = np.zeros(1, np.int32)
        ms = np.zeros(8, np.float32)
        assert g5.lib.pg5_bh_autotune(*args, 2, pgamd.iptr(best), pgamd.fptr(ms)) == 0
        assert g5.lib.pg5_bh_launch_shape(pgamd.iptr(shape)) == 0
        print(f"tree-list batch autotune: block {shape[0]}, {shape[1]} blocks per CU, "
              f"ms {np.round(ms, 3).tolist()}")
        assert shape[0] in out and shape[1] == best[0] and np.all(ms > 0)
        assert g5.lib.pg5_bh_batch(*args, None) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(acc.cpu().numpy(), out[int(shape[0])][0])
        np.testing.assert_array_equal(pot.cpu().numpy(), out[int(shape[0])][1])
        g5.lib.pgamd_set_abort_on_error(0)
        assert g5.lib.pg5_bh_set_launch(300, -1) == -1
        assert g5.lib.pg5_bh_set_launch(-1, 65) == -1
    finally:
        g5.lib.pgamd_set_abort_on_error(1)
        g5.lib.pg5_bh_set_launch(0, 0)


def test_clusters_match_oracle_and_graph():
    import torch

    import pgamd
    g6 = pgamd.G6()
    sizes = ics.cluster_sizes(48, 61, 64, 1024)
    xs, vs, ms, e2 = [], [], [], []
    for c, n in enumerate(sizes):
        m, x, v = ics.plummer(int(n), 1000 + c)
        xs.append(x + 10.0 * c)  # far apart (they never interact anyway)
        vs.append(v)
        ms.append(m)
        e2.append((4.0 / n) ** 2)
    coff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    X, V, Mm = np.concatenate(xs), np.concatenate(vs), np.concatenate(ms)
    items = g6.cluster_items(coff)
    n = len(Mm)
    d = dict(items=dev(items), coff=dev(coff), pos=dev(X), vel=dev(V), mass=dev(Mm),
             eps2=dev(np.array(e2, np.float32)))
    acc = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    jrk = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    pot = torch.zeros(n, dtype=torch.float64, device="cuda")
    args = (len(items) // 2, ptr(d["items"]), ptr(d["coff"]), ptr(d["pos"]), ptr(d["vel"]),
            ptr(d["mass"]), ptr(d["eps2"]), ptr(acc), ptr(jrk), ptr(pot))
    assert g6.lib.pg6_cluster_force(*args, None) == 0
    torch.cuda.synchronize()
    a, j, p = acc.cpu().numpy(), jrk.cpu().numpy(), pot.cpu().numpy()
    for c in range(0, len(sizes), 5):
        s = slice(coff[c], coff[c + 1])
        idx = np.arange(coff[c + 1] - coff[c], dtype=np.int32)
        vf = V[s].astype(np.float32).astype(np.float64)
        o = O.pg6_force(X[s], vf, idx, float(np.float32(e2[c])), X[s], vf, Mm[s], idx)
        assert np.percentile(rel_err(a[s], o["acc"]), 99) < 1e-6
        assert np.percentile(rel_err(j[s], o["jrk"]), 99) < 1e-5
        assert rel_err(p[s], o["pot"]).max() < 1e-6
    h = g6.lib.pg6_cluster_graph(*args)
    try:
        for rep in range(3):
            acc.zero_()
            assert g6.lib.pg6_cluster_graph_launch(h, None) == 0
            torch.cuda.synchronize()
            np.testing.assert_array_equal(acc.cpu().numpy(), a)
            assert g6.lib.pg6_cluster_force(*args, None) == 0
    finally:
        g6.lib.pg6_cluster_graph_destroy(h)
    best = np.zeros(1, np.int32)
    assert g6.lib.pg6_cluster_autotune(*args, 2, pgamd.iptr(best), None) == 0
    assert best[0] in (1, 2, 3, 4, 6, 8, 12, 16)
    # every j-tile against the oracle (clusters of 48 to 1024 particles: one
    # partial tile up to eight whole ones), then the autotuner's pick
    out = {}
    shape = np.zeros(2, np.int32)
    try:
        for tile in (128, 256, 512):
            assert g6.lib.pg6_cluster_set_launch(tile, -1) == 0
            assert g6.lib.pg6_cluster_launch_shape(pgamd.iptr(shape)) == 0
            assert shape[0] == tile
            assert g6.lib.pg6_cluster_force(*args, None) == 0
            torch.cuda.synchronize()
            out[tile] = [t.cpu().numpy() for t in (acc, jrk, pot)]
            a, j, p = out[tile]
            for c in list(range(1, len(sizes), 5)) + [int(np.argmax(sizes))]:
                s = slice(coff[c], coff[c + 1])
                idx = np.arange(coff[c + 1] - coff[c], dtype=np.int32)
                vf = V[s].astype(np.float32).astype(np.float64)
                o = O.pg6_force(X[s], vf, idx, float(np.float32(e2[c])), X[s], vf, Mm[s], idx)
                assert np.percentile(rel_err(a[s], o["acc"]), 99) < 1e-6, tile
                assert np.percentile(rel_err(j[s], o["jrk"]), 99) < 1e-5, tile
                assert rel_err(p[s], o["pot"]).max() < 1e-6, tile
        # FP32 sums fold into FP64 every 256 j, so tiles of 256 and 512 agree
        for r256, r512 in zip(out[256], out[512]):
            np.testing.assert_array_equal(r512, r256)
        ms = np.zeros(8, np.float32)
        assert g6.lib.pg6_cluster_autotune(*args, 2, pgamd.iptr(best), pgamd.fptr(ms)) == 0
        assert g6.lib.pg6_cluster_launch_shape(pgamd.iptr(shape)) == 0
        print(f"cluster batch autotune: tile {shape[0]}, {shape[1]} blocks per CU, "
              f"ms {np.round(ms, 3).tolist()}")
        assert shape[0] in out and shape[1] == best[0] and np.all(ms > 0)
        assert g6.lib.pg6_cluster_force(*args, None) == 0
        torch.cuda.synchronize()
        for t, r in zip((acc, jrk, pot), out[int(shape[0])]):
            np.testing.assert_array_equal(t.cpu().numpy(), r)
        g6.lib.pgamd_set_abort_on_error(0)
        assert g6.lib.pg6_cluster_set_launch(64, -1) == -1
    finally:
        g6.lib.pgamd_set_abort_on_error(1)
        g6.lib.pg6_cluster_set_launch(0, 0)


def test_quadratic_table_with_group_origins(tab, treepm_case):
    """The table configuration at a force error of 1e-5: a quadratic
    through each interval of the 1024-entry table (order 2) and FP64 group
    origins keep the per-particle error against the closed-form S2 below
    1e-5 at the 99th percentile; the linear table of the same size does not."""
    c = treepm_case
    L = c["L"]
    sel = np.random.default_rng(3).choice(c["n"], 400, replace=False)
    fa = O.table_force_f64(L["pos"][sel], L["pos"], c["m"], PP, PM)
    assert tab.lib.pg5_set_table_frc_bit(6) == 0
    try:
        e_lin = p99(run_treepm(tab, c, 0, p64=True)[sel], fa)
        assert tab.lib.pg5_set_table_order(2) == 0
        e_quad = p99(run_treepm(tab, c, 0, p64=True)[sel], fa)
    finally:
        tab.lib.pg5_set_table_order(1)
        tab.lib.pg5_set_table_frc_bit(5)
    print(f"1024 entries, group origins: linear p99 {e_lin:.3e}, quadratic p99 {e_quad:.3e}")
    assert e_quad < 1e-5 and e_quad < e_lin / 10, (e_quad, e_lin)
