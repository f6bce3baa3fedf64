"""Loader for the CPU oracle — TEST INFRASTRUCTURE ONLY.

Two kinds of checker live here:

* ``liboracle.so`` — the plain-C FP64 restatement of the reference algorithms
  (``pg_oracle.c``; every function cites the reference lines it follows);
* the reference libraries themselves, compiled from their own sources by
  ``oracle/ref.mk`` into ``oracle/_ref/`` (``RefPG5``, ``RefPG6``,
  ``RefTable``, ``RefQuad``), used to pin the restatement and as the timed
  CPU baseline.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFDIR = os.path.join(HERE, "_ref")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _P(a, t=_dp):
    return a.ctypes.data_as(t)


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -f oracle/Makefile)")
        L = C.CDLL(path, mode=os.RTLD_LOCAL)
        _sig(L, "or_pg5_force", None, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp)
        _sig(L, "or_pg6_predict", None, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp)
        _sig(L, "or_pg6_force", None, C.c_int, _dp, _dp, _ip, C.c_double, C.c_int, _dp, _dp,
             _dp, _ip, _dp, _dp, _dp, _ip, _dp)
        _sig(L, "or_pg6_nblist", None, C.c_int, _dp, _ip, _dp, C.c_double, C.c_int, _dp, _ip,
             C.c_int, _ip, _ip)
        _sig(L, "or_rsqrt_bias_of", C.c_double, _fp, _fp)
        _sig(L, "or_rsqrtss_table", None, _fp, _fp)
        _sig(L, "or_s2_reff", C.c_double, C.c_double, C.c_double)
        _sig(L, "or_s2_force", C.c_double, C.c_double, C.c_double, C.c_double)
        _sig(L, "or_table_build", None, C.c_double, C.c_double, C.c_int, C.c_int, _fp, _fp, _dp)
        _sig(L, "or_table_lookup", C.c_float, _fp, C.c_int, C.c_int, C.c_float)
        _sig(L, "or_table_force_emul", None, C.c_int, _dp, C.c_int, _dp, _dp, _fp, C.c_int,
             C.c_int, C.c_double, _dp)
        _sig(L, "or_table_force_f64", None, C.c_int, _dp, C.c_int, _dp, _dp, C.c_double,
             C.c_double, _dp)
        _sig(L, "or_quad_force", None, C.c_int, _dp, C.c_double, C.c_int, _dp, _dp, _dp, _dp,
             _dp)
        _sig(L, "or_quad_moments", None, C.c_int, _dp, _dp, _dp, _dp, _dp)
        _sig(L, "or_calc_modulus", C.c_double, C.c_double, C.c_double, C.c_double)
        _sig(L, "or_dt_startup", None, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp)
        _sig(L, "or_hermite_predict", None, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp,
             _dp)
        _sig(L, "or_hermite_correct", None, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
             _dp, _dp, _dp, _dp, _dp, _dp, _dp)
        _lib = L
    return _lib


# ------------------------------------------------------------- restatement
def pg5_force(xi, eps2i, xj, mj, eps2j=None):
    xi, xj, mj = _d(xi), _d(xj), _d(mj)
    ni, nj = len(xi), len(mj)
    e = _d(np.broadcast_to(eps2i, (ni,)))
    ej = None if eps2j is None else _d(eps2j)
    acc = np.zeros((ni, 3))
    pot = np.zeros(ni)
    lib().or_pg5_force(ni, _P(xi), _P(e), nj, _P(xj), _P(mj), None if ej is None else _P(ej),
                       _P(acc), _P(pot))
    return acc, pot


def pg6_predict(ti, x, v, a2, j6, tj):
    x, v, a2, j6, tj = map(_d, (x, v, a2, j6, tj))
    xp = np.zeros_like(x)
    vp = np.zeros_like(v)
    lib().or_pg6_predict(len(tj), float(ti), _P(x), _P(v), _P(a2), _P(j6), _P(tj), _P(xp),
                         _P(vp))
    return xp, vp


def pg6_force(xi, vi, idi, eps2, xj, vj, mj, idj):
    xi, vi, xj, vj, mj = map(_d, (xi, vi, xj, vj, mj))
    idi, idj = _i(idi), _i(idj)
    ni, nj = len(idi), len(idj)
    acc = np.zeros((ni, 3))
    jrk = np.zeros((ni, 3))
    pot = np.zeros(ni)
    nnb = np.zeros(ni, np.int32)
    rnnb = np.zeros(ni)
    lib().or_pg6_force(ni, _P(xi), _P(vi), _P(idi, _ip), float(eps2), nj, _P(xj), _P(vj),
                       _P(mj), _P(idj, _ip), _P(acc), _P(jrk), _P(pot), _P(nnb, _ip), _P(rnnb))
    return dict(acc=acc, jrk=jrk, pot=pot, nnb=nnb, rnnb=rnnb)


def pg6_nblist(xi, idi, h2, eps2, xj, idj, maxlen=1024):
    xi, xj, h2 = _d(xi), _d(xj), _d(h2)
    idi, idj = _i(idi), _i(idj)
    ni = len(idi)
    nbl = np.zeros((ni, maxlen), np.int32)
    nblen = np.zeros(ni, np.int32)
    lib().or_pg6_nblist(ni, _P(xi), _P(idi, _ip), _P(h2), float(eps2), len(idj), _P(xj),
                        _P(idj, _ip), maxlen, _P(nbl, _ip), _P(nblen, _ip))
    return [nbl[i, : min(nblen[i], maxlen)].copy() for i in range(ni)], nblen


RSQRT_N = 1 << 23


def rsqrt_args():
    """The arguments of rsqrt.c:42-81: floats with bits (127 << 23) + i and
    (128 << 23) + i, i < 2^23 (x0 in [1, 2), x1 = 2 x0 in [2, 4))."""
    i = np.arange(RSQRT_N, dtype=np.int32)
    return ((127 << 23) + i).view(np.float32), ((128 << 23) + i).view(np.float32)


def rsqrt_bias_of(y0, y1):
    """rsqrt_bias() of PG5/newton/libpg5/rsqrt.c:42-81 for the rsqrt values
    y0, y1 at rsqrt_args() (or_rsqrt_bias_of)."""
    y0 = np.ascontiguousarray(y0, np.float32)
    y1 = np.ascontiguousarray(y1, np.float32)
    assert len(y0) == len(y1) == RSQRT_N
    return lib().or_rsqrt_bias_of(_P(y0, _fp), _P(y1, _fp))


def rsqrtss_table():
    """The reference's rsqrt() (SSE rsqrtss) at rsqrt_args()."""
    y0 = np.zeros(RSQRT_N, np.float32)
    y1 = np.zeros(RSQRT_N, np.float32)
    lib().or_rsqrtss_table(_P(y0, _fp), _P(y1, _fp))
    return y0, y1


def s2_reff(a, r):
    """R(r, a) of pg5_table.c:8-22 (or_s2_reff)."""
    return lib().or_s2_reff(float(a), float(r))


def s2_force(eps_pp, eps_pm, r):
    return np.array([lib().or_s2_force(eps_pp, eps_pm, float(x)) for x in np.atleast_1d(r)])


def table_build(eps_pp, eps_pm, exp_bit=4, frc_bit=5):
    n = 1 << (exp_bit + frc_bit)
    tab = np.zeros(2 * n, np.float32)
    r2s = np.zeros(1, np.float32)
    xs = np.zeros(1)
    lib().or_table_build(eps_pp, eps_pm, exp_bit, frc_bit, _P(tab, _fp), _P(r2s, _fp), _P(xs))
    return tab, float(r2s[0]), float(xs[0])


def table_lookup(tab, s, exp_bit=4, frc_bit=5):
    tab = np.ascontiguousarray(tab, np.float32)
    return np.array([lib().or_table_lookup(_P(tab, _fp), exp_bit, frc_bit, float(v))
                     for v in np.atleast_1d(s)], np.float32)


def table_force_emul(xi, xj, mj, tab, xscale, exp_bit=4, frc_bit=5):
    xi, xj, mj = _d(xi), _d(xj), _d(mj)
    tab = np.ascontiguousarray(tab, np.float32)
    acc = np.zeros((len(xi), 3))
    lib().or_table_force_emul(len(xi), _P(xi), len(mj), _P(xj), _P(mj), _P(tab, _fp), exp_bit,
                              frc_bit, float(xscale), _P(acc))
    return acc


def table_force_f64(xi, xj, mj, eps_pp, eps_pm):
    xi, xj, mj = _d(xi), _d(xj), _d(mj)
    acc = np.zeros((len(xi), 3))
    lib().or_table_force_f64(len(xi), _P(xi), len(mj), _P(xj), _P(mj), float(eps_pp),
                             float(eps_pm), _P(acc))
    return acc


def quad_force(xi, eps2, xc, mc, qc):
    xi, xc, mc, qc = map(_d, (xi, xc, mc, qc))
    acc = np.zeros((len(xi), 3))
    pot = np.zeros(len(xi))
    lib().or_quad_force(len(xi), _P(xi), float(eps2), len(mc), _P(xc), _P(mc), _P(qc), _P(acc),
                        _P(pot))
    return acc, pot


def quad_moments(x, m):
    x, m = _d(x), _d(m)
    com = np.zeros(3)
    mass = np.zeros(1)
    q = np.zeros(6)
    lib().or_quad_moments(len(m), _P(x), _P(m), _P(com), _P(mass), _P(q))
    return com, float(mass[0]), q


# ------------------------------------------------------- Hermite integrator
def calc_modulus(num, mod, modmax):
    """timestep.c:176-190."""
    return lib().or_calc_modulus(float(num), float(mod), float(modmax))


def dt_startup(eta_s, dtmax, t_gl, a, j):
    """timestep.c:24-48 for every particle (0 where no step fits)."""
    a, j = _d(a), _d(j)
    dt = np.zeros(len(a))
    lib().or_dt_startup(len(a), float(eta_s), float(dtmax), float(t_gl), _P(a), _P(j), _P(dt))
    return dt


def hermite_predict(t_gl, t, x, v, a, j):
    """predict_exmotion, hermite.c:39-62."""
    t, x, v, a, j = map(_d, (t, x, v, a, j))
    xp, vp = np.zeros_like(x), np.zeros_like(v)
    lib().or_hermite_predict(len(t), float(t_gl), _P(t), _P(x), _P(v), _P(a), _P(j), _P(xp),
                             _P(vp))
    return xp, vp


def hermite_correct(eta, dtmax, t, dt, a0, j0, xp, vp, a1, j1):
    """correct (hermite.c:66-110) + dt_criterion (timestep.c:50-79): returns
    corrected x, v and the new steps."""
    t, dt, a0, j0, xp, vp, a1, j1 = map(_d, (t, dt, a0, j0, xp, vp, a1, j1))
    x1, v1, dtn = np.zeros_like(xp), np.zeros_like(vp), np.zeros(len(t))
    lib().or_hermite_correct(len(t), float(eta), float(dtmax), _P(t), _P(dt), _P(a0), _P(j0),
                             _P(xp), _P(vp), _P(a1), _P(j1), _P(x1), _P(v1), _P(dtn))
    return x1, v1, dtn


# ------------------------------------------------ the compiled reference
def ref_available(name: str) -> bool:
    return os.path.exists(os.path.join(REFDIR, f"libref_{name}.so"))


def _ref(name):
    path = os.path.join(REFDIR, f"libref_{name}.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built (make -f oracle/ref.mk)")
    return C.CDLL(path, mode=os.RTLD_LOCAL)


class RefPG5:
    """Reference GRAPE-5 library (AVX2, PG5/newton/libpg5).  Its j-memory holds
    65536 particles (gp5util.h:8)."""

    def __init__(self, omp=False, quad=False):
        self.L = L = _ref("pg5quad" if quad else ("pg5_omp" if omp else "pg5"))
        _sig(L, "g5_open", None)
        _sig(L, "g5_set_eps_to_all", None, C.c_double)
        _sig(L, "g5_set_n", None, C.c_int)
        _sig(L, "g5_set_xmj", None, C.c_int, C.c_int, _dp, _dp)
        _sig(L, "g5_calculate_force_on_x", None, _dp, _dp, _dp, C.c_int)
        _sig(L, "g5_calculate_force_on_xe", None, _dp, _dp, _dp, _dp, C.c_int)
        if quad:
            _sig(L, "g5c_set_xmjMC", None, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp)
            _sig(L, "g5c_set_nMC", None, C.c_int, C.c_int)
            _sig(L, "g5c_calculate_force_on_xMC", None, C.c_int, _dp, _dp, _dp, C.c_int)
        L.g5_open()

    def force(self, xi, xj, mj, eps):
        xi, xj, mj = _d(xi), _d(xj), _d(mj)
        self.L.g5_set_xmj(0, len(mj), _P(xj), _P(mj))
        self.L.g5_set_eps_to_all(float(eps))
        self.L.g5_set_n(len(mj))
        a = np.zeros((len(xi), 3))
        p = np.zeros(len(xi))
        self.L.g5_calculate_force_on_x(_P(xi), _P(a), _P(p), len(xi))
        return a, p

    def cell_force(self, xi, xc, mc, qc, eps):
        xi, xc, mc, qc = map(_d, (xi, xc, mc, qc))
        nc = len(mc)
        if nc % 2:  # g5c_set_xmjMC stores cells in pairs and reads xj[nc]
            xc = np.concatenate([xc, np.zeros((1, 3))])
            mc = np.concatenate([mc, [0.0]])
            qc = np.concatenate([qc, np.zeros((1, 6))])
        self.L.g5_set_eps_to_all(float(eps))
        self.L.g5c_set_xmjMC(0, 0, len(mc), _P(xc), _P(mc), _P(qc))
        self.L.g5c_set_nMC(0, len(mc))
        a = np.zeros((len(xi), 3))
        p = np.zeros(len(xi))
        self.L.g5c_calculate_force_on_xMC(0, _P(xi), _P(a), _P(p), len(xi))
        return a, p


class RefTable:
    """Reference tabulated-cutoff library (PG5/table/libpg5, AVX2)."""

    def __init__(self, omp=False):
        self.L = L = _ref("pg5table_omp" if omp else "pg5table")
        _sig(L, "g5_open", None)
        _sig(L, "pg5_gen_s2_force_table", None, C.c_double, C.c_double)
        _sig(L, "g5_set_n", None, C.c_int)
        _sig(L, "g5_set_nMC", None, C.c_int, C.c_int)
        _sig(L, "g5_set_xmj", None, C.c_int, C.c_int, _dp, _dp)
        _sig(L, "g5_set_xmjMC", None, C.c_int, C.c_int, C.c_int, _dp, _dp)
        _sig(L, "g5_calculate_force_on_x", None, _dp, _dp, _dp, C.c_int)
        _sig(L, "g5_calculate_force_on_xMC", None, C.c_int, _dp, _dp, _dp, C.c_int)
        L.g5_open()

    def gen_s2(self, eps_pp, eps_pm):
        self.L.pg5_gen_s2_force_table(float(eps_pp), float(eps_pm))

    def table(self) -> np.ndarray:
        arr = (C.c_float * 1024).in_dll(self.L, "Force_table")
        return np.ctypeslib.as_array(arr).copy()

    def force(self, xi, xj, mj, devid=None):
        xi, xj, mj = _d(xi), _d(xj), _d(mj)
        nj = len(mj)
        a = np.zeros((len(xi), 3))
        p = np.zeros(len(xi))
        if devid is None:
            self.L.g5_set_xmj(0, nj, _P(xj), _P(mj))
            self.L.g5_set_n(nj)
            self.L.g5_calculate_force_on_x(_P(xi), _P(a), _P(p), len(xi))
        else:
            self.L.g5_set_xmjMC(devid, 0, nj, _P(xj), _P(mj))
            self.L.g5_set_nMC(devid, nj)
            self.L.g5_calculate_force_on_xMC(devid, _P(xi), _P(a), _P(p), len(xi))
        return a


class RefPG6:
    """Reference GRAPE-6 library (PG6/libpg6, AVX2 + OpenMP).  Each lasthalf
    call takes at most g6_npipes() = 48 i-particles (avx_type.h:1)."""

    NPIPES = 48

    def __init__(self):
        self.L = L = _ref("pg6")
        _sig(L, "g6_open", None, C.c_int)
        _sig(L, "g6_set_j_particle", None, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
             C.c_double, _dp, _dp, _dp, _dp, _dp)
        _sig(L, "g6_set_ti", None, C.c_int, C.c_double)
        _sig(L, "g6calc_firsthalf", None, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp,
             _dp, C.c_double, _dp)
        _sig(L, "g6calc_lasthalf2p", C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp,
             C.c_double, _dp, _dp, _dp, _dp, _ip, _fp)
        _sig(L, "g6calc_lasthalfn", C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp,
             C.c_double, _dp, _dp, _dp, _dp)
        _sig(L, "g6_get_neighbour_list", C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip)
        _sig(L, "g6_read_neighbour_list", C.c_int, C.c_int)
        L.g6_open(0)

    def set_j(self, pos, vel, acc2, jrk6, mass, tj, idx):
        pos, vel, acc2, jrk6 = map(_d, (pos, vel, acc2, jrk6))
        z = np.zeros(3)
        for k in range(len(mass)):
            self.L.g6_set_j_particle(0, k, int(idx[k]), float(tj[k]), 0.0, float(mass[k]),
                                     _P(z), _P(jrk6[k]), _P(acc2[k]), _P(vel[k]), _P(pos[k]))

    def calc(self, nj, ti, pos, vel, idx, eps2, predict=True, h2=None):
        pos, vel = _d(pos), _d(vel)
        idx = _i(idx)
        ni = len(idx)
        acc = np.zeros((ni, 3))
        jrk = np.zeros((ni, 3))
        pot = np.zeros(ni)
        nnb = np.zeros(ni, np.int32)
        rnnb = np.zeros(ni, np.float32)
        h2a = _d(np.zeros(ni) if h2 is None else h2)
        if predict:
            self.L.g6_set_ti(0, float(ti))
        nbls = []
        for o in range(0, ni, self.NPIPES):
            k = min(self.NPIPES, ni - o)
            sl = slice(o, o + k)
            self.L.g6calc_firsthalf(0, nj, k, _P(idx[sl], _ip), _P(pos[sl]), _P(vel[sl]),
                                    _P(acc[sl]), _P(jrk[sl]), _P(pot[sl]), float(eps2),
                                    _P(h2a[sl]))
            if h2 is None:
                self.L.g6calc_lasthalf2p(0, nj, k, _P(idx[sl], _ip), _P(pos[sl]), _P(vel[sl]),
                                         float(eps2), _P(h2a[sl]), _P(acc[sl]), _P(jrk[sl]),
                                         _P(pot[sl]), _P(nnb[sl], _ip), _P(rnnb[sl], _fp))
            else:
                self.L.g6calc_lasthalfn(0, nj, k, _P(idx[sl], _ip), _P(pos[sl]), _P(vel[sl]),
                                        float(eps2), _P(h2a[sl]), _P(acc[sl]), _P(jrk[sl]),
                                        _P(pot[sl]))
                for q in range(k):
                    n = np.zeros(1, np.int32)
                    buf = np.zeros(1024, np.int32)
                    rc = self.L.g6_get_neighbour_list(0, q, 1024, _P(n, _ip), _P(buf, _ip))
                    nbls.append(buf[: n[0]].copy() if rc == 0 else None)
        out = dict(acc=acc, jrk=jrk, pot=pot, nnb=nnb, rnnb=rnnb)
        if h2 is not None:
            out["nbl"] = nbls
        return out


# ------------------------------------------- CPU baselines of the batches
_cb = None


def _cpubatch() -> C.CDLL:
    global _cb
    if _cb is None:
        path = os.path.join(HERE, "libcpubatch.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -f oracle/Makefile)")
        L = C.CDLL(path, mode=os.RTLD_LOCAL)
        _sig(L, "cb_treepm", C.c_double, C.c_char_p, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip,
             _dp, _dp, C.c_double, C.c_double, C.c_int, _dp)
        _sig(L, "cb_bh", C.c_double, C.c_char_p, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip,
             _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, _dp, _dp)
        _sig(L, "cb_clusters", C.c_double, C.c_char_p, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp,
             _dp, _dp, _dp)
        _sig(L, "cb_hermite", C.c_double, C.c_char_p, C.c_int, _dp, _dp, _dp, C.c_double,
             C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_longlong), _dp,
             C.c_int)
        _cb = L
    return _cb


def _refpath(name: str) -> bytes:
    path = os.path.join(REFDIR, f"libref_{name}.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built (make -f oracle/ref.mk)")
    return path.encode()


def _csr_subset(off, cnt, groups):
    """Sub-CSR (offsets, counts) of the selected groups, same lists."""
    return _i(np.asarray(off)[groups]), _i(np.asarray(cnt)[groups])


def ref_treepm(lists, pos, mass, eps_pp, eps_pm, nthreads, groups):
    """Reference table library (PG5/table/libpg5) over the selected groups, in
    an OpenMP loop with devid = thread.  Returns (seconds, acc[n][3])."""
    pos, mass = _d(pos), _d(mass)
    g = np.asarray(groups)
    ioff, icnt = _csr_subset(lists["ioff"], lists["icnt"], g)
    joff, jcnt = _csr_subset(lists["joff"], lists["jcnt"], g)
    ilist, jlist = _i(lists["ilist"]), _i(lists["jlist"])
    acc = np.zeros((len(pos), 3))
    t = _cpubatch().cb_treepm(_refpath("pg5table"), len(g), _P(ioff, _ip), _P(icnt, _ip),
                              _P(ilist, _ip), _P(joff, _ip), _P(jcnt, _ip), _P(jlist, _ip),
                              _P(pos), _P(mass), float(eps_pp), float(eps_pm), int(nthreads),
                              _P(acc))
    if t < 0:
        raise RuntimeError("cb_treepm failed")
    return t, acc


def ref_bh(lists, pos, mass, cells10, eps, nthreads, groups):
    """Reference GRAPE-5 + quad_tree libraries (PG5/newton_quad/quad) over the
    selected groups, OpenMP loop with devid = thread.  Returns (seconds, acc,
    pot) with the libraries' sign conventions."""
    pos, mass, cells10 = _d(pos), _d(mass), _d(cells10)
    g = np.asarray(groups)
    ioff, icnt = _csr_subset(lists["ioff"], lists["icnt"], g)
    joff, jcnt = _csr_subset(lists["joff"], lists["jcnt"], g)
    coff, ccnt = _csr_subset(lists["coff"], lists["ccnt"], g)
    ilist, jlist, clist = _i(lists["ilist"]), _i(lists["jlist"]), _i(lists["clist"])
    acc = np.zeros((len(pos), 3))
    pot = np.zeros(len(pos))
    t = _cpubatch().cb_bh(_refpath("pg5quad"), len(g), _P(ioff, _ip), _P(icnt, _ip),
                          _P(ilist, _ip), _P(joff, _ip), _P(jcnt, _ip), _P(jlist, _ip),
                          _P(coff, _ip), _P(ccnt, _ip), _P(clist, _ip), _P(pos), _P(mass),
                          _P(cells10), float(eps), int(nthreads), _P(acc), _P(pot))
    if t < 0:
        raise RuntimeError("cb_bh failed")
    return t, acc, pot


def ref_clusters(sel, coff, pos, vel, mass, eps2):
    """Reference GRAPE-6 library, one session per selected cluster (members at
    t = 0).  Returns (seconds, acc, jrk, pot) for the members of `sel`."""
    pos, vel, mass, eps2 = _d(pos), _d(vel), _d(mass), _d(eps2)
    sel, coff = _i(sel), _i(coff)
    acc = np.zeros((len(mass), 3))
    jrk = np.zeros((len(mass), 3))
    pot = np.zeros(len(mass))
    t = _cpubatch().cb_clusters(_refpath("pg6"), len(sel), _P(sel, _ip), _P(coff, _ip),
                                _P(pos), _P(vel), _P(mass), _P(eps2), _P(acc), _P(jrk), _P(pot))
    if t < 0:
        raise RuntimeError("cb_clusters failed")
    return t, acc, jrk, pot


def ref_hermite(x, v, m, eps2, eta_s, eta, dtmax, t_end, maxw=64):
    """The reference's sample Hermite integration (PG6/banana) on the compiled
    reference library.  Returns (seconds, x, v, info{steps, blocks, windows},
    energies[w] = (ek, ep) at each window end)."""
    x, v, m = _d(x).copy(), _d(v).copy(), _d(m)
    info = (C.c_longlong * 3)()
    en = np.zeros(2 * maxw)
    t = _cpubatch().cb_hermite(_refpath("pg6"), len(m), _P(x), _P(v), _P(m), float(eps2),
                               float(eta_s), float(eta), float(dtmax), float(t_end), info,
                               _P(en), int(maxw))
    if t < 0:
        raise RuntimeError("cb_hermite failed")
    w = int(info[2])
    return t, x, v, dict(steps=int(info[0]), blocks=int(info[1]), windows=w), \
        en[:2 * min(w, maxw)].reshape(-1, 2)
