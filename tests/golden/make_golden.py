"""Writes the golden fixtures in tests/golden/*.npz by running the reference
libraries, compiled from their own sources by oracle/ref.mk, on the
reference's own input files and on small seeded inputs.

    python tests/golden/make_golden.py [/root/reference]

Each fixture holds inputs and the reference's outputs only.  Provenance: the
reference PG5/quad kernels use the x86 rsqrtps estimate, whose bits depend on
the CPU vendor; the CPU model is recorded in each fixture ("cpu").
"""
from __future__ import annotations

import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "phantom-grape_amd"))

import ics  # noqa: E402
import oracle as O  # noqa: E402


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def read_direct(path):
    """PG5/newton/direct/direct.c:35-65 format: n, 3, t, masses, positions, velocities."""
    tok = open(path).read().split()
    n = int(tok[0])
    vals = np.array(tok[3:3 + 7 * n], dtype=np.float64)
    m = vals[:n]
    x = vals[n:4 * n].reshape(n, 3)
    v = vals[4 * n:7 * n].reshape(n, 3)
    return m, x, v


def read_banana(path):
    """PG6/banana/io.c:12-38 format: t, n, then 'i m x y z vx vy vz' per particle."""
    tok = open(path).read().split()
    n = int(tok[1])
    rows = np.array(tok[2:2 + 8 * n], dtype=np.float64).reshape(n, 8)
    # banana reads the mass with "%f" into a float (io.c:21)
    return rows[:, 1].astype(np.float32).astype(np.float64), rows[:, 2:5].copy(), rows[:, 5:8].copy()


def main(ref: str) -> None:
    cpu = cpu_model()
    # ---------------------------------------------------------------- PG5
    R = O.RefPG5()
    out = {}
    for name in ("pl1k", "pl1k-1", "pl1k-3"):
        m, x, _ = read_direct(os.path.join(ref, "PG5/newton/direct/init", name))
        eps = 4.0 / len(m)  # ERRORTEST setting, direct.c:252-254
        a, p = R.force(x, x, m, eps)
        key = name.replace("-", "_")
        out[f"{key}_m"], out[f"{key}_x"] = m, x
        out[f"{key}_eps"] = np.float64(eps)
        out[f"{key}_acc"], out[f"{key}_pot"] = a, p
    np.savez_compressed(os.path.join(HERE, "pg5_direct.npz"), cpu=cpu, **out)

    # symmetric softening, eps_j^2 as direct.c:255-260
    m, x, _ = read_direct(os.path.join(ref, "PG5/newton/direct/init", "pl1k"))
    n = len(m)
    e2 = (0.01 + 0.01 * np.arange(n) / n) ** 2
    import ctypes as C
    L = O._ref("pg5sym")
    dp = C.POINTER(C.c_double)
    for nm, args in (("g5_open", []), ("g5_set_n", [C.c_int]),
                     ("g5_set_xmj", [C.c_int, C.c_int, dp, dp, dp]),
                     ("g5_calculate_force_on_xe", [dp, dp, dp, dp, C.c_int])):
        getattr(L, nm).argtypes = args
        getattr(L, nm).restype = None
    L.g5_open()
    xc, mc, ec = (np.ascontiguousarray(v) for v in (x, m, e2))
    L.g5_set_xmj(0, n, xc.ctypes.data_as(dp), mc.ctypes.data_as(dp), ec.ctypes.data_as(dp))
    L.g5_set_n(n)
    a = np.zeros((n, 3))
    p = np.zeros(n)
    L.g5_calculate_force_on_xe(xc.ctypes.data_as(dp), ec.ctypes.data_as(dp),
                               a.ctypes.data_as(dp), p.ctypes.data_as(dp), n)
    np.savez_compressed(os.path.join(HERE, "pg5_symmetric.npz"), cpu=cpu, m=m, x=x, eps2=e2,
                        acc=a, pot=p)

    # ---------------------------------------------------------------- PG6
    m, x, v = read_banana(os.path.join(ref, "PG6/banana/init/pl001k.init"))
    n = len(m)
    idx = np.arange(n, dtype=np.int32)
    eps2 = (1.0 / 64) ** 2
    o0 = O.pg6_force(x, v, idx, eps2, x, v, m, idx)  # acc/jerk for the predictor inputs
    rng = np.random.default_rng(7)
    tj = -rng.integers(0, 4, n) / 256.0  # last-update times before the block time 0
    ti = 1.0 / 128
    acc2, jrk6 = o0["acc"] / 2, o0["jrk"] / 6
    P = O.RefPG6()
    P.set_j(x, v, acc2, jrk6, m, tj, idx)
    xp, vp = O.pg6_predict(ti, x, v, acc2, jrk6, tj)
    isel = rng.choice(n, 200, replace=False).astype(np.int32)
    r = P.calc(n, ti, xp[isel], vp[isel], idx[isel], eps2)
    h2 = np.full(len(isel), 0.05 ** 2)
    h2[::7] = 0.0
    rn = P.calc(n, ti, xp[isel], vp[isel], idx[isel], eps2, predict=False, h2=h2)
    nbl = [l if l is not None else np.zeros(0, np.int32) for l in rn["nbl"]]
    nblen = np.array([len(l) for l in nbl], np.int32)
    np.savez_compressed(os.path.join(HERE, "pg6_hermite.npz"), cpu=cpu, m=m, x=x, v=v, acc2=acc2,
                        jrk6=jrk6, tj=tj, ti=ti, eps2=eps2, isel=isel, xi=xp[isel],
                        vi=vp[isel], acc=r["acc"], jrk=r["jrk"], pot=r["pot"], nnb=r["nnb"],
                        rnnb=r["rnnb"], h2=h2, nblen=nblen,
                        nbl=np.concatenate(nbl) if nbl else np.zeros(0, np.int32))

    # -------------------------------------------------------------- table
    T = O.RefTable()
    tab64 = T.table()
    T.gen_s2(0.1 / 128, 3.0 / 128)
    tab128 = T.table()
    mb, xb = ics.clumpy_box(8192, 11)
    xb = xb[np.argsort(xb[:, 0])]
    # a slab, so most pairs fall inside r_cut = 3/128
    sel = xb[:, 0] < 0.08
    xj, mj = xb[sel][:2048], mb[sel][:2048]
    xi = xj[::8].copy()
    acc = T.force(xi, xj, mj)
    np.savez_compressed(os.path.join(HERE, "pg5_table.npz"), cpu=cpu, tab64=tab64,
                        tab128=tab128, xi=xi, xj=xj, mj=mj, acc=acc)

    # ---------------------------------------------------------- quadrupole
    Q = O.RefPG5(quad=True)
    mh, xh = ics.hernquist(4096, 5)
    order = np.lexsort((xh[:, 2], xh[:, 1], xh[:, 0]))
    xh, mh = xh[order], mh[order]
    cx, cm, cq = [], [], []
    for k in range(0, 4096, 16):
        c, M, q = O.quad_moments(xh[k:k + 16], mh[k:k + 16])
        cx.append(c)
        cm.append(M)
        cq.append(q)
    cx, cm, cq = np.array(cx), np.array(cm), np.array(cq)
    xi = xh[::37][:64] * 0.01 + 60.0  # far away from every cell
    eps = 0.01
    a, p = Q.cell_force(xi, cx, cm, cq, eps)
    np.savez_compressed(os.path.join(HERE, "pg5_quad.npz"), cpu=cpu, xc=cx, mc=cm, qc=cq,
                        xi=xi, eps=eps, acc=a, pot=p)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
