"""Python mirror of the MI355X Phantom-GRAPE C ABI.

The product is the three shared libraries built from ``csrc/``:

* ``libpg5.so``       GRAPE-5 API ``g5_*`` + quad_tree ``g5c_*`` + ``pg5_dev_*``
  (reference ``PG5/newton/libpg5/gp5util.h``, ``PG5/newton_quad/quad/quad_util.h``)
* ``libpg5_sym.so``   the same, built with symmetric softening (reference
  ``-DSYMMETRIC``, ``PG5/newton/setting.mk:6,17-19``)
* ``libpg5_table.so`` GRAPE-5 tabulated-cutoff API + ``pg5_*`` table calls +
  learned-cutoff and batched TreePM entry points
  (reference ``PG5/table/libpg5/gp5util.h``, ``pg5_table.h``)
* ``libpg6.so``       GRAPE-6 API ``g6_*`` / ``g6calc_*`` + ``pg6_dev_*``
  (reference ``PG6/libpg6/gp6util.h``)
* ``libpgtree.so`` / ``libpgtree_dev.so``  interaction lists for the batched
  calls, built on the host (OpenMP) or on the GPU from positions in HBM

This module loads them with ctypes (one private handle each, so the ``g5_*``
names of ``libpg5.so`` and ``libpg5_table.so`` never clash) and wraps the calls
for numpy arrays.  Function names, argument meaning and order follow the C API.
There is no CPU fallback: if a library is missing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
LIBS = ("pg5", "pg5_sym", "pg5_table", "pg6", "pgtree", "pgtree_dev")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


def lib_path(name: str) -> str:
    return os.path.join(LIBDIR, f"lib{name}.so")


_handles: dict[str, C.CDLL] = {}


def _one_hip_runtime():
    """PyTorch-ROCm ships its own libamdhip64.so.7 with the same soname as the
    system one these libraries link.  Whichever is loaded first serves the
    whole process, and torch cannot use the system runtime; so when torch is
    importable it is imported before the first library loads."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(name: str) -> C.CDLL:
    """Load ``lib<name>.so`` from the package's lib/ directory (built in-tree)."""
    if name in _handles:
        return _handles[name]
    _one_hip_runtime()
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run __graft_entry__.build() (hipcc for gfx950); "
            "this package has no CPU fallback")
    lib = C.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
    _handles[name] = lib
    return lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_fp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_ip)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def last_error(lib: C.CDLL) -> tuple[int, str]:
    buf = C.create_string_buffer(512)
    code = lib.pgamd_last_error(buf, 512)
    return code, buf.value.decode()


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


# ----------------------------------------------------------------- GRAPE-5
class G5:
    """GRAPE-5 Newtonian board (libpg5.so).  Potential is returned with the
    API's sign: ``+sum m/r`` including the self term (``phantom_g5.c:45-60``);
    callers negate it and remove ``m/eps`` as ``direct.c:165-172`` does."""

    def __init__(self, symmetric: bool = False):
        self.lib = load("pg5_sym" if symmetric else "pg5")
        self.symmetric = symmetric
        L = self.lib
        _sig(L, "g5_open", None)
        _sig(L, "g5_close", None)
        _sig(L, "g5_get_number_of_pipelines", C.c_int)
        _sig(L, "g5_get_jmemsize", C.c_int)
        _sig(L, "g5_set_eps_to_all", None, C.c_double)
        _sig(L, "g5_set_n", None, C.c_int)
        _sig(L, "g5_set_nMC", None, C.c_int, C.c_int)
        if symmetric:
            _sig(L, "g5_set_xmjMC", None, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp)
        else:
            _sig(L, "g5_set_xmjMC", None, C.c_int, C.c_int, C.c_int, _dp, _dp)
        _sig(L, "g5_set_xiMC", None, C.c_int, C.c_int, _dp)
        _sig(L, "g5_set_xeiMC", None, C.c_int, C.c_int, _dp, _dp)
        _sig(L, "g5_set_epsMC", None, C.c_int, C.c_int, _dp)
        _sig(L, "g5_runMC", None, C.c_int)
        _sig(L, "g5_get_forceMC", None, C.c_int, C.c_int, _dp, _dp)
        _sig(L, "g5_calculate_force_on_xMC", None, C.c_int, _dp, _dp, _dp, C.c_int)
        _sig(L, "g5_calculate_force_on_xeMC", None, C.c_int, _dp, _dp, _dp, _dp, C.c_int)
        _sig(L, "g5c_set_nMC", None, C.c_int, C.c_int)
        _sig(L, "g5c_set_xmjMC", None, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp)
        _sig(L, "g5c_set_xmjMC0", None, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp)
        _sig(L, "g5c_calculate_force_on_xMC", None, C.c_int, _dp, _dp, _dp, C.c_int)
        _sig(L, "g5c_calculate_force_on_xMC0", None, C.c_int, _dp, _dp, _dp, C.c_int)
        _sig(L, "g5c_runMC", None, C.c_int)
        _sig(L, "g5c_runMC0", None, C.c_int)
        _sig(L, "pg5_dev_force", C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp,
             _vp)
        _sig(L, "pg5_dev_quad", C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp)
        _sig(L, "pg5_dev_stream", _vp, C.c_int)
        _sig(L, "pg5_set_rsqrt_bias_mode", C.c_int, C.c_int)
        _sig(L, "pg5_rsqrt_bias", C.c_double)
        _sig(L, "pg5_dev_set_launch", C.c_int, C.c_int, C.c_int, C.c_int)
        _sig(L, "pg5_dev_set_kernel", C.c_int, C.c_int)
        bh = [C.c_int] * 1 + [_vp] * 10 + [_vp, _vp, C.c_double, _vp, _vp]
        _sig(L, "pg5_bh_batch", C.c_int, C.c_int, *bh, _vp)
        _sig(L, "pg5_bh_graph", _vp, C.c_int, *bh)
        _sig(L, "pg5_bh_graph_launch", C.c_int, _vp, _vp)
        _sig(L, "pg5_bh_graph_destroy", C.c_int, _vp)
        _sig(L, "pg5_bh_autotune", C.c_int, C.c_int, *bh, C.c_int, _ip, _fp)
        _sig(L, "pg5_bh_set_launch", C.c_int, C.c_int, C.c_int)
        _sig(L, "pg5_bh_launch_shape", C.c_int, _ip)
        _sig(L, "pgamd_set_abort_on_error", None, C.c_int)
        _sig(L, "pgamd_last_error", C.c_int, C.c_char_p, C.c_int)
        L.g5_open()

    def set_xmj(self, adr, x, m, epsj2=None, devid=0):
        """``g5_set_xmj(adr, nj, xj, mj)``: x is indexed from ``adr`` (the
        reference reads ``xj[j]`` for ``j`` in ``[adr, adr+nj)``)."""
        x = f64(x)
        m = f64(m)
        nj = len(m) - adr
        if self.symmetric:
            e = f64(epsj2)
            self.lib.g5_set_xmjMC(devid, adr, nj, dptr(x), dptr(m), dptr(e))
        else:
            self.lib.g5_set_xmjMC(devid, adr, nj, dptr(x), dptr(m))

    def set_eps_to_all(self, eps):
        self.lib.g5_set_eps_to_all(float(eps))

    def set_n(self, n, devid=0):
        self.lib.g5_set_nMC(devid, int(n))

    def calculate_force_on_x(self, x, devid=0, eps2=None):
        x = f64(x)
        ni = x.shape[0]
        a = np.zeros((ni, 3))
        p = np.zeros(ni)
        if eps2 is None:
            self.lib.g5_calculate_force_on_xMC(devid, dptr(x), dptr(a), dptr(p), ni)
        else:
            e = f64(eps2)
            self.lib.g5_calculate_force_on_xeMC(devid, dptr(x), dptr(e), dptr(a), dptr(p), ni)
        return a, p

    def set_cells(self, x, m, q, devid=0, mem=0):
        x, m, q = f64(x), f64(m), f64(q)
        f = self.lib.g5c_set_xmjMC if mem == 0 else self.lib.g5c_set_xmjMC0
        f(devid, 0, len(m), dptr(x), dptr(m), dptr(q))
        self.lib.g5c_set_nMC(devid, len(m))

    def cell_force_on_x(self, x, devid=0, mem=0):
        x = f64(x)
        ni = x.shape[0]
        a = np.zeros((ni, 3))
        p = np.zeros(ni)
        f = self.lib.g5c_calculate_force_on_xMC if mem == 0 else self.lib.g5c_calculate_force_on_xMC0
        f(devid, dptr(x), dptr(a), dptr(p), ni)
        return a, p


# ----------------------------------------------------------------- GRAPE-6
class G6:
    """GRAPE-6 board (libpg6.so).  j-particle inputs carry acc/2 and jerk/6
    (``PG6/banana/banana01.c:199-212``); outputs are acc, jerk and
    ``pot = -sum m/r`` with the self pair excluded by id."""

    def __init__(self):
        self.lib = L = load("pg6")
        _sig(L, "g6_open", None, C.c_int)
        _sig(L, "g6_close", None, C.c_int)
        _sig(L, "g6_npipes", C.c_int)
        _sig(L, "g6_set_j_particle", None, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
             C.c_double, _dp, _dp, _dp, _dp, _dp)
        _sig(L, "g6_set_ti", None, C.c_int, C.c_double)
        _sig(L, "g6calc_firsthalf", None, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp,
             _dp, C.c_double, _dp)
        for nm in ("g6calc_lasthalf", "g6calc_lasthalfn"):
            _sig(L, nm, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, C.c_double, _dp, _dp,
                 _dp, _dp)
        for nm in ("g6calc_lasthalf2", "g6calc_lasthalf2n"):
            _sig(L, nm, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, C.c_double, _dp, _dp,
                 _dp, _dp, _ip)
        for nm in ("g6calc_lasthalf2p", "g6calc_lasthalf2np"):
            _sig(L, nm, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, C.c_double, _dp, _dp,
                 _dp, _dp, _ip, _fp)
        _sig(L, "g6_read_neighbour_list", C.c_int, C.c_int)
        _sig(L, "g6_get_neighbour_list", C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip)
        _sig(L, "pg6_dev_set_j", C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip)
        _sig(L, "pg6_dev_predict", C.c_int, C.c_int, C.c_double, _vp)
        _sig(L, "pg6_dev_force", C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.c_double, _vp,
             _vp, _vp, _vp, _vp, _vp)
        _sig(L, "pg6_dev_sync", C.c_int)
        _sig(L, "pg6_dev_stream", _vp)
        _sig(L, "pg6_dev_set_launch", C.c_int, C.c_int, C.c_int, C.c_int)
        _sig(L, "pg6_dev_plan", C.c_int, C.c_int, C.c_int, _ip)
        _sig(L, "pg6_dev_read_predicted", C.c_int, C.c_int, _dp, _fp, _fp)
        _sig(L, "pg6_dev_set_kernel", C.c_int, C.c_int)
        _sig(L, "pg6_dev_autotune", C.c_int, C.c_int, C.c_double, C.c_int, _ip)
        _sig(L, "pg6_dev_autotune_reset", C.c_int)
        _sig(L, "pg6_hermite_init", C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double,
             C.c_double, C.c_double)
        _sig(L, "pg6_hermite_step", C.c_int, _dp, _ip)
        _sig(L, "pg6_hermite_run", C.c_longlong, C.c_double)
        _sig(L, "pg6_hermite_state", C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
             C.POINTER(C.c_longlong))
        _sig(L, "pg6_hermite_energy", C.c_int, _dp, _dp)
        _sig(L, "pg6_hermite_end", C.c_int)
        _sig(L, "pg6_dev_timing_start", C.c_int, C.c_int)
        _sig(L, "pg6_dev_timing_read", C.c_int, _dp, _ip)
        _sig(L, "pg6_cluster_items", C.c_int, C.c_int, _ip, _ip, C.c_int)
        cl = [C.c_int] + [_vp] * 9
        _sig(L, "pg6_cluster_force", C.c_int, *cl, _vp)
        _sig(L, "pg6_cluster_graph", _vp, *cl)
        _sig(L, "pg6_cluster_graph_launch", C.c_int, _vp, _vp)
        _sig(L, "pg6_cluster_graph_destroy", C.c_int, _vp)
        _sig(L, "pg6_cluster_autotune", C.c_int, *cl, C.c_int, _ip, _fp)
        _sig(L, "pg6_cluster_set_launch", C.c_int, C.c_int, C.c_int)
        _sig(L, "pg6_cluster_launch_shape", C.c_int, _ip)
        _sig(L, "pgamd_set_abort_on_error", None, C.c_int)
        _sig(L, "pgamd_last_error", C.c_int, C.c_char_p, C.c_int)
        L.g6_open(0)

    def close(self):
        self.lib.g6_close(0)

    def set_j_particles(self, pos, vel, acc2, jrk6, mass, tj, idx, adr0=0):
        """One ``g6_set_j_particle`` call per particle, addresses adr0.."""
        pos, vel, acc2, jrk6 = f64(pos), f64(vel), f64(acc2), f64(jrk6)
        z = np.zeros(3)
        for k in range(len(mass)):
            self.lib.g6_set_j_particle(0, adr0 + k, int(idx[k]), float(tj[k]), 0.0,
                                       float(mass[k]), dptr(z), dptr(jrk6[k]), dptr(acc2[k]),
                                       dptr(vel[k]), dptr(pos[k]))

    def set_j_bulk(self, pos, vel, acc2, jrk6, mass, tj, idx):
        return self.lib.pg6_dev_set_j(len(mass), dptr(f64(pos)), dptr(f64(vel)), dptr(f64(acc2)),
                                      dptr(f64(jrk6)), dptr(f64(mass)), dptr(f64(tj)),
                                      iptr(i32(idx)))

    def set_ti(self, t):
        self.lib.g6_set_ti(0, float(t))

    def calc(self, nj, pos, vel, idx, eps2, h2=None, variant="2p"):
        """g6calc_firsthalf followed by g6calc_lasthalf<variant> over all i in
        one call.  Returns dict of outputs."""
        pos, vel, idx = f64(pos), f64(vel), i32(idx)
        ni = len(idx)
        h2a = f64(h2 if h2 is not None else np.zeros(ni))
        acc = np.zeros((ni, 3))
        jrk = np.zeros((ni, 3))
        pot = np.zeros(ni)
        nnb = np.zeros(ni, np.int32)
        rnnb = np.zeros(ni, np.float32)
        dummy = np.zeros((ni, 3))
        self.lib.g6calc_firsthalf(0, nj, ni, iptr(idx), dptr(pos), dptr(vel), dptr(dummy),
                                  dptr(dummy), dptr(pot), float(eps2), dptr(h2a))
        L = self.lib
        args = (0, nj, ni, iptr(idx), dptr(pos), dptr(vel), float(eps2), dptr(h2a), dptr(acc),
                dptr(jrk), dptr(pot))
        if variant == "":
            rc = L.g6calc_lasthalf(*args)
        elif variant == "2":
            rc = L.g6calc_lasthalf2(*args, iptr(nnb))
        elif variant == "2p":
            rc = L.g6calc_lasthalf2p(*args, iptr(nnb), fptr(rnnb))
        elif variant == "n":
            rc = L.g6calc_lasthalfn(*args)
        elif variant == "2n":
            rc = L.g6calc_lasthalf2n(*args, iptr(nnb))
        elif variant == "2np":
            rc = L.g6calc_lasthalf2np(*args, iptr(nnb), fptr(rnnb))
        else:
            raise ValueError(variant)
        return dict(rc=rc, acc=acc, jrk=jrk, pot=pot, nnb=nnb, rnnb=rnnb)

    def read_predicted(self, nj):
        x = np.zeros((nj, 3))
        v = np.zeros((nj, 3), np.float32)
        m = np.zeros(nj, np.float32)
        rc = self.lib.pg6_dev_read_predicted(nj, dptr(x), fptr(v), fptr(m))
        assert rc == 0
        return x, v, m

    def plan(self, ni, nj):
        """Launch shape for (ni, nj): dict of variant, bi, bpc, nsplit, jper."""
        o = np.zeros(5, np.int32)
        self.lib.pg6_dev_plan(ni, nj, iptr(o))
        return dict(zip(("variant", "bi", "bpc", "nsplit", "jper"), (int(v) for v in o)))

    # ------------------------------------------------ Hermite integration
    def hermite_init(self, x, v, m, eps2, eta_s, eta, dtmax):
        x, v, m = f64(x), f64(v), f64(m)
        self._hn = len(m)
        rc = self.lib.pg6_hermite_init(len(m), dptr(x), dptr(v), dptr(m), float(eps2),
                                       float(eta_s), float(eta), float(dtmax))
        if rc != 0:
            raise RuntimeError(last_error(self.lib))

    def hermite_step(self):
        t = np.zeros(1)
        k = np.zeros(1, np.int32)
        if self.lib.pg6_hermite_step(dptr(t), iptr(k)) != 0:
            raise RuntimeError(last_error(self.lib))
        return float(t[0]), int(k[0])

    def hermite_run(self, t_end):
        k = self.lib.pg6_hermite_run(float(t_end))
        if k < 0:
            raise RuntimeError(last_error(self.lib))
        return int(k)

    def hermite_state(self):
        n = self._hn
        out = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), a=np.zeros((n, 3)),
                   j=np.zeros((n, 3)), pot=np.zeros(n), t=np.zeros(n), dt=np.zeros(n))
        info = (C.c_longlong * 4)()
        assert self.lib.pg6_hermite_state(*[dptr(out[k]) for k in ("x", "v", "a", "j", "pot", "t",
                                                                    "dt")], info) == 0
        out.update(t_mod=int(info[0]), steps=int(info[1]), blocks=int(info[2]))
        return out

    def hermite_energy(self):
        ek, ep = np.zeros(1), np.zeros(1)
        assert self.lib.pg6_hermite_energy(dptr(ek), dptr(ep)) == 0
        return float(ek[0]), float(ep[0])

    def autotune(self, nj, eps2, reps=3):
        """Runs pg6_dev_autotune; returns {ni: (variant, bi, bpc, us)}."""
        o = np.zeros(4 * 11, np.int32)
        k = self.lib.pg6_dev_autotune(nj, eps2, reps, iptr(o))
        if k < 0:
            raise RuntimeError(last_error(self.lib))
        return {64 << b: tuple(int(v) for v in o[4 * b:4 * b + 4]) for b in range(k)}

    def cluster_items(self, coff):
        coff = i32(coff)
        n = self.lib.pg6_cluster_items(len(coff) - 1, iptr(coff), None, 0)
        items = np.zeros(2 * n, np.int32)
        self.lib.pg6_cluster_items(len(coff) - 1, iptr(coff), iptr(items), n)
        return items

    def neighbour_list(self, ipipe, maxlen=1024):
        n = np.zeros(1, np.int32)
        buf = np.zeros(max(maxlen, 1), np.int32)
        rc = self.lib.g6_get_neighbour_list(0, ipipe, maxlen, iptr(n), iptr(buf))
        return rc, buf[: n[0]].copy()


# --------------------------------------------------------- table (TreePM)
class G5Table:
    """Tabulated-cutoff GRAPE-5 board (libpg5_table.so) with the learned
    cut-off and the batched TreePM entry points."""

    def __init__(self):
        self.lib = L = load("pg5_table")
        _sig(L, "g5_open", None)
        _sig(L, "pg5_gen_s2_force_table", None, C.c_double, C.c_double)
        _sig(L, "pg5_set_table_frc_bit", C.c_int, C.c_int)
        _sig(L, "pg5_set_table_order", C.c_int, C.c_int)
        _sig(L, "pg5_get_table", C.c_int, _fp, _dp, _fp)
        _sig(L, "g5_set_xmj", None, C.c_int, C.c_int, _dp, _dp)
        _sig(L, "g5_set_xmjMC", None, C.c_int, C.c_int, C.c_int, _dp, _dp)
        _sig(L, "g5_set_n", None, C.c_int)
        _sig(L, "g5_set_nMC", None, C.c_int, C.c_int)
        _sig(L, "g5_calculate_force_on_x", None, _dp, _dp, _dp, C.c_int)
        _sig(L, "g5_calculate_force_on_xMC", None, C.c_int, _dp, _dp, _dp, C.c_int)
        _sig(L, "pg5_set_cutoff_mode", C.c_int, C.c_int)
        _sig(L, "pg5_get_cutoff_mode", C.c_int)
        _sig(L, "pg5_mlp_set_params", C.c_int, _dp, C.c_int, C.c_double, C.c_double)
        _sig(L, "pg5_mlp_get_params", C.c_int, _dp, _ip, _dp, _dp)
        _sig(L, "pg5_mlp_normal_eq", C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp)
        _sig(L, "pg5_mlp_train", C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, _dp)
        _sig(L, "pg5_mlp_eval_dev", C.c_int, C.c_int, _vp, _vp)
        _sig(L, "pg5t_dev_eval", C.c_int, C.c_int, _vp, _vp)
        _sig(L, "pg5t_dev_force", C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, _vp)
        bt = [C.c_int] + [_vp] * 10 + [C.c_int]  # ngroups, 7 lists, pos, pos64, acc, mode
        _sig(L, "pg5t_batch_force", C.c_int, C.c_int, *bt, _vp)
        _sig(L, "pg5t_batch_graph", _vp, C.c_int, *bt)
        _sig(L, "pg5t_batch_graph_launch", C.c_int, _vp, _vp)
        _sig(L, "pg5t_batch_graph_destroy", C.c_int, _vp)
        _sig(L, "pg5t_batch_autotune", C.c_int, C.c_int, *bt, C.c_int, _ip, _fp)
        _sig(L, "pg5t_batch_set_blocks_per_cu", C.c_int, C.c_int, C.c_int)
        _sig(L, "pg5t_batch_set_block", C.c_int, C.c_int)
        _sig(L, "pg5t_batch_launch_shape", C.c_int, C.c_int, _ip)
        _sig(L, "pg5t_pm_force", C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_double, _vp, _vp)
        _sig(L, "pgamd_set_abort_on_error", None, C.c_int)
        _sig(L, "pgamd_last_error", C.c_int, C.c_char_p, C.c_int)
        L.g5_open()

    def table(self):
        n = self.lib.pg5_get_table(None, None, None)
        t = np.zeros(2 * n, np.float32)
        xs = np.zeros(1)
        self.lib.pg5_get_table(fptr(t), dptr(xs), None)
        return t, float(xs[0])

    def force(self, xi, xj, mj, devid=0):
        xi, xj, mj = f64(xi), f64(xj), f64(mj)
        self.lib.g5_set_xmjMC(devid, 0, len(mj), dptr(xj), dptr(mj))
        self.lib.g5_set_nMC(devid, len(mj))
        a = np.zeros((len(xi), 3))
        p = np.zeros(len(xi))
        self.lib.g5_calculate_force_on_xMC(devid, dptr(xi), dptr(a), dptr(p), len(xi))
        return a

    def mlp_params(self):
        """(p[321] tanh form FP64, form, rc, r0) of the loaded network, or None."""
        p = np.zeros(321)
        form = np.zeros(1, np.int32)
        rc = np.zeros(1)
        r0 = np.zeros(1)
        ok = self.lib.pg5_mlp_get_params(dptr(p), iptr(form), dptr(rc), dptr(r0))
        return (p, int(form[0]), float(rc[0]), float(r0[0])) if ok else None

    def mlp_set_params(self, p, form, rc, r0):
        assert self.lib.pg5_mlp_set_params(dptr(f64(p)), int(form), float(rc), float(r0)) == 0

    def mlp_train(self, iters=200, starts=4, refine=3000, seed=1):
        """pg5_mlp_train: Levenberg-Marquardt from `starts` seeded starts of
        `iters` trial steps, the best refined by `refine` more; returns the
        final weighted MSE."""
        loss = np.zeros(1)
        rc = self.lib.pg5_mlp_train(int(iters), int(starts), int(refine), int(seed), dptr(loss))
        assert rc == 0
        return float(loss[0])

    def mlp_normal_eq(self, v, t, wt, p):
        """(J^T J, J^T r, r^T r) of the weighted fit at p, from the GPU kernels."""
        v, t, wt, p = (f64(a) for a in (v, t, wt, p))
        jtj = np.zeros((321, 321))
        jtr = np.zeros(321)
        rtr = np.zeros(1)
        assert self.lib.pg5_mlp_normal_eq(len(v), dptr(v), dptr(t), dptr(wt), dptr(p),
                                          dptr(jtj), dptr(jtr), dptr(rtr)) == 0
        return jtj, jtr, float(rtr[0])


# ------------------------------------------------------- interaction lists
_TREE_FIELDS = ("pos", "src", "ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "coff", "ccnt",
                "clist", "order", "cells")


def _tree_result(L, h):
    info = np.zeros(10, np.int64)
    L.pgt_info(h, info.ctypes.data_as(C.c_void_p))
    np_, ng, ni, nj, nc, ncell = (int(v) for v in info[:6])
    sizes = dict(pos=(np_, 3), src=np_, ioff=ng, icnt=ng, ilist=ni, joff=ng, jcnt=ng, jlist=nj,
                 coff=ng, ccnt=ng, clist=nc, order=ng, cells=(ncell, 10))
    out = {}
    for k, name in enumerate(_TREE_FIELDS):
        dt = np.float64 if name in ("pos", "cells") else np.int32
        a = np.zeros(sizes[name], dt)
        if a.size:
            L.pgt_get(h, k, a.ctypes.data_as(C.c_void_p))
        out[name] = a
    out["work_pp"] = int(info[6])
    out["work_pc"] = int(info[7])
    L.pgt_free(h)
    return out


def _tree_lib():
    L = load("pgtree")
    _sig(L, "pgt_treepm_build", _vp, C.c_int, _dp, C.c_double, C.c_int)
    _sig(L, "pgt_bh_build", _vp, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int)
    _sig(L, "pgt_info", C.c_int, _vp, _vp)
    _sig(L, "pgt_get", C.c_int, _vp, C.c_int, _vp)
    _sig(L, "pgt_free", None, _vp)
    return L


def treepm_lists(x, rcut, ncrit=256):
    """TreePM i-groups and j-lists (libpgtree.so, host): periodic unit box,
    ghosts appended after the n originals (``src`` maps back)."""
    L = _tree_lib()
    x = f64(x)
    return _tree_result(L, L.pgt_treepm_build(len(x), dptr(x), float(rcut), int(ncrit)))


def bh_lists(x, m, theta=0.5, ncrit=512, nleaf=8):
    """Barnes-Hut i-groups, particle lists and quadrupole cell lists."""
    L = _tree_lib()
    x, m = f64(x), f64(m)
    return _tree_result(L, L.pgt_bh_build(len(x), dptr(x), dptr(m), float(theta), int(ncrit),
                                          int(nleaf)))


class DevLists:
    """Interaction lists built on the GPU (libpgtree_dev.so) from positions
    already in HBM; the arrays stay on the device.  ``ptr(name)`` is a device
    pointer for the batched calls, ``host()`` copies every array into the dict
    layout of :func:`treepm_lists` / :func:`bh_lists`."""

    FIELDS = _TREE_FIELDS

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("pgt_dev build failed: " + str(last_error(self.lib())))
        self.h = handle

    @staticmethod
    def lib():
        L = load("pgtree_dev")
        _sig(L, "pgt_dev_treepm_build", _vp, C.c_int, _vp, C.c_double, C.c_int, _vp)
        _sig(L, "pgt_dev_bh_build", _vp, C.c_int, _vp, _vp, C.c_double, C.c_int, C.c_int, _vp)
        _sig(L, "pgt_dev_info", C.c_int, _vp, _vp)
        _sig(L, "pgt_dev_ptr", _vp, _vp, C.c_int)
        _sig(L, "pgt_dev_get", C.c_int, _vp, C.c_int, _vp)
        _sig(L, "pgt_dev_pack", C.c_int, _vp, _vp, C.c_double, _vp, _vp, _vp)
        _sig(L, "pgt_dev_free", None, _vp)
        _sig(L, "pgamd_set_abort_on_error", None, C.c_int)
        _sig(L, "pgamd_last_error", C.c_int, C.c_char_p, C.c_int)
        return L

    @classmethod
    def treepm(cls, d_x, n, rcut, ncrit=256, stream=None):
        """d_x: device pointer to [n][3] FP64 positions in [0, 1)."""
        return cls(cls.lib().pgt_dev_treepm_build(int(n), d_x, float(rcut), int(ncrit), stream))

    @classmethod
    def bh(cls, d_x, d_m, n, theta=0.5, ncrit=512, nleaf=8, stream=None):
        return cls(cls.lib().pgt_dev_bh_build(int(n), d_x, d_m, float(theta), int(ncrit),
                                              int(nleaf), stream))

    def info(self):
        a = np.zeros(8, np.int64)
        self.lib().pgt_dev_info(self.h, a.ctypes.data_as(C.c_void_p))
        return dict(zip(("np", "ngroups", "ni", "nj", "nc", "ncells", "work_pp", "work_pc"),
                        (int(v) for v in a)))

    def ptr(self, name):
        return self.lib().pgt_dev_ptr(self.h, self.FIELDS.index(name))

    def host(self):
        i = self.info()
        np_, ng = i["np"], i["ngroups"]
        sizes = dict(pos=(np_, 3), src=np_, ioff=ng, icnt=ng, ilist=i["ni"], joff=ng, jcnt=ng,
                     jlist=i["nj"], coff=ng, ccnt=ng, clist=i["nc"], order=ng,
                     cells=(i["ncells"], 10))
        out = {}
        for k, name in enumerate(self.FIELDS):
            dt = np.float64 if name in ("pos", "cells") else np.int32
            a = np.zeros(sizes[name], dt)
            if a.size:
                self.lib().pgt_dev_get(self.h, k, a.ctypes.data_as(C.c_void_p))
            out[name] = a
        out["work_pp"], out["work_pc"] = i["work_pp"], i["work_pc"]
        return out

    def pack(self, d_mass, scale, d_pos4=None, d_cells12=None, stream=None):
        """float4 {x*scale, y*scale, z*scale, m} per list particle and 12-float cells."""
        assert self.lib().pgt_dev_pack(self.h, d_mass, float(scale), d_pos4, d_cells12,
                                       stream) == 0

    def free(self):
        if self.h:
            self.lib().pgt_dev_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
