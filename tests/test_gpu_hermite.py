"""Block-time-step Hermite integration on the GPU (pg6_hermite_*) against the
CPU restatement of the reference's integrator (oracle/hermite_ref.h) and the
reference library's own integration (oracle/cpu_batch.c: cb_hermite)."""
import numpy as np
import pytest

import ics
import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

ETA_S, ETA, DTMAX = 0.01, 0.2, 0.125


@pytest.fixture(scope="module")
def g6():
    import pgamd
    g = pgamd.G6()
    yield g
    g.lib.pg6_hermite_end()


def test_startup_forces_and_steps(g6):
    n, eps2 = 1024, (1 / 64) ** 2
    m, x, v = ics.plummer(n, 31)
    g6.hermite_init(x, v, m, eps2, ETA_S, ETA, DTMAX)
    s = g6.hermite_state()
    np.testing.assert_array_equal(s["x"], x)
    np.testing.assert_array_equal(s["t"], 0.0)
    idx = np.arange(n, dtype=np.int32)
    o = O.pg6_force(x, v, idx, eps2, x, v.astype(np.float32).astype(np.float64), m, idx)
    assert np.percentile(rel_err(s["a"], o["acc"]), 99) < 1e-6
    assert np.percentile(rel_err(s["j"], o["jrk"]), 99) < 1e-5
    assert rel_err(s["pot"], o["pot"]).max() < 1e-6
    # first steps: dt_startup of the GPU's own forces, bit for bit
    np.testing.assert_array_equal(s["dt"], O.dt_startup(ETA_S, DTMAX, 0.0, s["a"], s["j"]))


def test_block_steps_match_the_restated_integrator(g6):
    n, eps2 = 2048, (1 / 128) ** 2
    m, x, v = ics.plummer(n, 32)
    g6.hermite_init(x, v, m, eps2, ETA_S, ETA, DTMAX)
    s0 = g6.hermite_state()
    idx = np.arange(n, dtype=np.int32)
    seen_reset = False
    for step in range(60):
        tn = (s0["t"] + s0["dt"]).min()
        act = s0["t"] + s0["dt"] == tn
        t, nact = g6.hermite_step()
        s1 = g6.hermite_state()
        assert nact == act.sum(), step
        # the active particles: corrector + step criterion on the GPU's forces
        xp, vp = O.hermite_predict(tn, s0["t"][act], s0["x"][act], s0["v"][act], s0["a"][act],
                                   s0["j"][act])
        x1, v1, dtn = O.hermite_correct(ETA, DTMAX, s0["t"][act], s0["dt"][act], s0["a"][act],
                                        s0["j"][act], xp, vp, s1["a"][act], s1["j"][act])
        np.testing.assert_array_equal(s1["x"][act], x1, err_msg=f"step {step}")
        np.testing.assert_array_equal(s1["v"][act], v1, err_msg=f"step {step}")
        np.testing.assert_array_equal(s1["dt"][act], dtn, err_msg=f"step {step}")
        # their forces, against FP64 at the predicted positions (j predicted too)
        if step % 15 == 0:
            xj, vj = O.pg6_predict(tn, s0["x"], s0["v"], s0["a"] / 2, s0["j"] * (1.0 / 6.0),
                                   s0["t"])
            o = O.pg6_force(xp, vp, idx[act], eps2, xj, vj.astype(np.float32).astype(np.float64),
                            m, idx)
            assert np.percentile(rel_err(s1["a"][act], o["acc"]), 99) < 1e-6
        # the others are untouched
        for k in ("x", "v", "a", "j", "dt"):
            np.testing.assert_array_equal(s1[k][~act], s0[k][~act], err_msg=k)
        if tn == DTMAX:  # window end: every time restarts at 0
            seen_reset = True
            assert act.all() and np.all(s1["t"] == 0) and s1["t_mod"] == s0["t_mod"] + 1
        else:
            np.testing.assert_array_equal(s1["t"][act], tn)
        # block invariants: every step a power-of-two fraction of dtmax, t a multiple of it
        q = DTMAX / s1["dt"]
        assert np.all(q == 2.0 ** np.round(np.log2(q)))
        assert np.all(np.mod(s1["t"], s1["dt"]) == 0)
        s0 = s1
        if seen_reset and step > 5:
            break


def test_integration_conserves_energy_like_the_reference(g6):
    n, eps2 = 1024, (1 / 64) ** 2
    m, x, v = ics.plummer(n, 33)
    g6.hermite_init(x, v, m, eps2, ETA_S, ETA, DTMAX)
    ek0, ep0 = g6.hermite_energy()
    e0 = ek0 + ep0
    blocks = g6.hermite_run(0.25)
    s = g6.hermite_state()
    assert s["t_mod"] == 2 and blocks > 10
    ek, ep = g6.hermite_energy()
    de_gpu = abs(ek + ep - e0) / abs(e0)
    assert de_gpu < 1e-4, de_gpu
    if O.ref_available("pg6"):
        _, _, _, info, en = O.ref_hermite(x, v, m, eps2, ETA_S, ETA, DTMAX, 0.25)
        de_ref = abs(en[-1].sum() - e0) / abs(e0)
        # same integrator, FP32 pair forces on both sides: comparable drift
        assert de_gpu < max(10 * de_ref, 1e-6), (de_gpu, de_ref)
        assert abs(info["blocks"] - (blocks + 1)) < 0.2 * info["blocks"]


def test_run_in_graphs_equals_single_steps(g6):
    """pg6_hermite_run queues captured graphs of steps (replayed many times,
    with host reads between batches, and steps past t_end skipped on the
    device); stepping the same integration one call at a time ends in the
    same state, bit for bit."""
    n, eps2 = 1024, (1 / 64) ** 2
    m, x, v = ics.plummer(n, 34)
    t_end = 3.0 / 16
    g6.hermite_init(x, v, m, eps2, ETA_S, ETA, DTMAX)
    blocks = g6.hermite_run(t_end)
    a = g6.hermite_state()
    assert blocks > 100  # several graphs of steps
    g6.hermite_init(x, v, m, eps2, ETA_S, ETA, DTMAX)
    t, k = 0.0, 0
    while t < t_end:
        t, _ = g6.hermite_step()
        k += 1
    b = g6.hermite_state()
    assert k == blocks
    for key in ("x", "v", "a", "j", "pot", "t", "dt", "t_mod", "steps", "blocks"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    # a second run continues where the first stopped
    assert g6.hermite_run(t_end) == 0
    assert g6.hermite_run(t_end + 1.0 / 64) > 0


@pytest.mark.parametrize("n", [2, 3, 100])
def test_small_systems(g6, n):
    """Particle counts below one wavefront and not a multiple of 64: the
    graph-driven run and single steps agree bit for bit, the two-body orbit
    keeps its energy."""
    m, x, v = ics.plummer(n, 40 + n)
    eps2 = (1 / 64) ** 2
    t_end = 1.0 / 16
    g6.hermite_init(x, v, m, eps2, ETA_S, ETA, DTMAX)
    ek0, ep0 = g6.hermite_energy()
    blocks = g6.hermite_run(t_end)
    a = g6.hermite_state()
    g6.hermite_init(x, v, m, eps2, ETA_S, ETA, DTMAX)
    t, k = 0.0, 0
    while t < t_end:
        t, nact = g6.hermite_step()
        assert 1 <= nact <= n
        k += 1
    b = g6.hermite_state()
    assert k == blocks > 0
    for key in ("x", "v", "a", "j", "pot", "t", "dt", "t_mod", "steps", "blocks"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    if t_end % DTMAX == 0 or np.all(b["t"] == b["t"][0]):
        ek, ep = g6.hermite_energy()
        assert abs(ek + ep - ek0 - ep0) < 1e-3 * abs(ek0 + ep0)


def test_errors_are_reported(g6):
    import pgamd
    g6.lib.pgamd_set_abort_on_error(0)
    try:
        assert g6.lib.pg6_hermite_init(1, None, None, None, 0.0, 0.01, 0.2, 0.125) == -1
        assert pgamd.last_error(g6.lib)[0] != 0
    finally:
        g6.lib.pgamd_set_abort_on_error(1)
