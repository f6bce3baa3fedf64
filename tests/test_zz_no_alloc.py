"""The allocation-free entry point pg5_dev_force_noalloc (partial sums in a
static device array) against pg5_dev_force on the same inputs, through the
C ABI: two calls in a row, the second with a smaller input (the array keeps
the first call's contents), results identical (the same j-range plan fits)."""
import ctypes as C

import numpy as np
import pytest

import ics

pytestmark = pytest.mark.gpu


def test_noalloc_matches_dev_force_twice():
    import torch

    import pgamd
    g5 = pgamd.G5()
    L = g5.lib
    vp = C.c_void_p
    L.pg5_dev_force_noalloc.restype = C.c_int
    L.pg5_dev_force_noalloc.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, vp]
    s = torch.cuda.Stream()
    for n_i, n_j, seed in ((4096, 16384, 61), (1000, 3000, 62)):
        print(f"case: ni={n_i} nj={n_j}", flush=True)
        m, x, _ = ics.plummer(n_j, seed)
        eps = 1.0 / 64
        xi = np.zeros((n_i, 4), np.float32)
        xi[:, :3] = x[:n_i]
        xi[:, 3] = eps * eps
        xj = np.zeros((n_j, 4), np.float32)
        xj[:, :3] = x
        xj[:, 3] = m
        d_xi = torch.from_numpy(xi).cuda()
        d_xj = torch.from_numpy(xj).cuda()
        out = {k: (torch.full((n_i, 3), np.nan, dtype=torch.float64, device="cuda"),
                   torch.full((n_i,), np.nan, dtype=torch.float64, device="cuda"))
               for k in ("alloc", "noalloc")}
        torch.cuda.synchronize()
        sp = s.cuda_stream
        assert L.pg5_dev_force(0, n_i, d_xi.data_ptr(), n_j, d_xj.data_ptr(), None,
                               out["alloc"][0].data_ptr(), out["alloc"][1].data_ptr(), sp) == 0
        assert L.pg5_dev_force_noalloc(n_i, d_xi.data_ptr(), n_j, d_xj.data_ptr(), None,
                                       out["noalloc"][0].data_ptr(),
                                       out["noalloc"][1].data_ptr(), sp) == 0
        s.synchronize()
        a0, p0 = (t.cpu().numpy() for t in out["alloc"])
        a1, p1 = (t.cpu().numpy() for t in out["noalloc"])
        assert np.all(np.isfinite(a1)) and np.all(np.isfinite(p1))
        np.testing.assert_array_equal(a1, a0)
        np.testing.assert_array_equal(p1, p0)
