#!/usr/bin/env python3
"""Benchmark of the GRAPE-6 Hermite force path on one MI355X.

Workload (the configuration the headline metric is quoted on): Plummer sphere,
N = 65536, G = M = 1, equal masses, eps = 1/256, FP64 positions/velocities,
j-particles carrying acc/2 and jerk/6 from a start-up force pass and block
time steps 2^-3..2^-12 (Aarseth criterion), predicted to the block time.
One step = one Hermite predictor pass over all j plus one force + jerk +
potential + nearest-neighbour evaluation of all N i-particles against all N
j-particles (N (N-1) interactions), inputs resident in HBM.

Prints one JSON line.  Fields beyond the driver's contract:
  roofline      dominant kernel (pg6_force) against the FP32 vector peak,
                kernel time from HIP events recorded around each launch on
                its stream; 60 flops per interaction (PG6/banana/banana01.h:9)
  accuracy      median / p99 relative error of acc, jerk, pot against the
                FP64 oracle on a seeded sample of i-particles
  cpu_baseline  the reference library (PG6/libpg6, AVX2 + OpenMP, compiled
                from its sources into oracle/_ref) on this host, same inputs
  sweep         i-batch sizes 64..65536 against all j (block-step sizes)
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "phantom-grape_amd"))

METRIC = "Ginteractions/s + % FP32 vector peak; p99 rel force err vs FP64"
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector (packed) peak
FLOPS_PER_INTERACTION = 60   # force + jerk convention of the reference
N = 65536
EPS = 1.0 / 256
SEED = 65536
TBLOCK = 1.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def block_steps(acc, jrk, eta=0.02):
    """Aarseth-type dt = eta |a|/|j| rounded down to a power of two in [2^-12, 2^-3]."""
    import numpy as np
    a = np.linalg.norm(acc, axis=1)
    j = np.linalg.norm(jrk, axis=1)
    dt = eta * a / np.maximum(j, 1e-300)
    lev = np.clip(np.ceil(-np.log2(np.maximum(dt, 1e-300))), 3, 12)
    return 2.0 ** (-lev)


def read_traffic():
    """HBM bytes per pg6_force launch from the committed rocprofv3 PMC summary."""
    p = os.path.join(ROOT, "profiles", "pg6_force_n65536_traffic.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ics
    import pgamd

    dev = torch.device("cuda:0")
    log(f"device {torch.cuda.get_device_name(0)}")
    g6 = pgamd.G6()
    L = g6.lib
    n = N
    m, x, v = ics.plummer(n, SEED)
    idx = np.arange(n, dtype=np.int32)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def dev_arrays(ni):
        return dict(acc=torch.empty((ni, 3), dtype=torch.float64, device=dev),
                    jrk=torch.empty((ni, 3), dtype=torch.float64, device=dev),
                    pot=torch.empty(ni, dtype=torch.float64, device=dev),
                    nnb=torch.empty(ni, dtype=torch.int32, device=dev),
                    rnnb=torch.empty(ni, dtype=torch.float32, device=dev))

    def force(ni, ip, iv, ii, out):
        rc = L.pg6_dev_force(n, ni, ip.data_ptr(), iv.data_ptr(), ii.data_ptr(), EPS * EPS,
                             out["acc"].data_ptr(), out["jrk"].data_ptr(), out["pot"].data_ptr(),
                             out["nnb"].data_ptr(), out["rnnb"].data_ptr(), sp)
        assert rc == 0

    # start-up force pass at t = 0 (all j at t_j = 0, so prediction is the identity)
    z = np.zeros((n, 3))
    assert g6.set_j_bulk(x, v, z, z, m, np.zeros(n), idx) == 0
    d_x = torch.from_numpy(x).to(dev)
    d_v = torch.from_numpy(v).to(dev)
    d_id = torch.from_numpy(idx).to(dev)
    out = dev_arrays(n)
    assert L.pg6_dev_predict(n, 0.0, sp) == 0
    force(n, d_x, d_v, d_id, out)
    torch.cuda.synchronize()
    acc0 = out["acc"].cpu().numpy()
    jrk0 = out["jrk"].cpu().numpy()

    # block time steps; every j last updated one step before the block time
    dt = block_steps(acc0, jrk0)
    tj = TBLOCK - dt
    acc2, jrk6 = acc0 / 2, jrk0 / 6
    assert g6.set_j_bulk(x, v, acc2, jrk6, m, tj, idx) == 0
    levels, counts = np.unique(-np.log2(dt).astype(int), return_counts=True)
    log("block levels " + ", ".join(f"2^-{l}:{c}" for l, c in zip(levels, counts)))
    # i-particles = all particles predicted to the block time (host FP64,
    # as the reference caller does, PG6/banana/hermite.c:39-60)
    h = (TBLOCK - tj)[:, None]
    xi = x + v * h + acc2 * h * h + jrk6 * h * h * h
    vi = v + 2 * acc2 * h + 3 * jrk6 * h * h
    d_xi = torch.from_numpy(np.ascontiguousarray(xi)).to(dev)
    d_vi = torch.from_numpy(np.ascontiguousarray(vi)).to(dev)

    def step():
        assert L.pg6_dev_predict(n, TBLOCK, sp) == 0
        force(n, d_xi, d_vi, d_id, out)

    plan = g6.plan(n, n)
    log(f"launch plan ni=nj={n}: bi={plan[0]} nsplit={plan[1]} jper={plan[2]}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    assert L.pg6_dev_timing_start(args.steps) == 0
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    total_ms = e0.elapsed_time(e1)
    kms = np.zeros(1)
    kcount = np.zeros(1, np.int32)
    L.pg6_dev_timing_read(pgamd.dptr(kms), pgamd.iptr(kcount))
    ms_step = total_ms / args.steps
    inter = float(n) * (n - 1)
    value = inter / (ms_step * 1e-3) / 1e9
    k_ms = float(kms[0]) / max(int(kcount[0]), 1)
    achieved = FLOPS_PER_INTERACTION * inter / (k_ms * 1e-3) / 1e12
    log(f"step {ms_step:.3f} ms ({value:.1f} Ginteractions/s); pg6_force {k_ms:.3f} ms "
        f"avg over {int(kcount[0])} launches = {achieved:.2f} TFLOP/s")
    gpu_acc = out["acc"].cpu().numpy()
    gpu_jrk = out["jrk"].cpu().numpy()
    gpu_pot = out["pot"].cpu().numpy()

    # i-batch sweep (block-step sizes) against all j
    sweep = []
    rng = np.random.default_rng(SEED + 1)
    for ni in [64 << k for k in range(11)]:
        sel = np.sort(rng.choice(n, ni, replace=False))
        ip = d_xi[torch.from_numpy(sel).to(dev)].contiguous()
        iv = d_vi[torch.from_numpy(sel).to(dev)].contiguous()
        ii = d_id[torch.from_numpy(sel).to(dev)].contiguous()
        o = dev_arrays(ni)
        reps = max(3, min(200, int(2e10 / (ni * n))))
        for _ in range(2):
            force(ni, ip, iv, ii, o)
        L.pg6_dev_timing_start(reps)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            assert L.pg6_dev_predict(n, TBLOCK, sp) == 0
            force(ni, ip, iv, ii, o)
        e1.record(stream)
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / reps
        L.pg6_dev_timing_read(pgamd.dptr(kms), pgamd.iptr(kcount))
        tk = float(kms[0]) / max(int(kcount[0]), 1)
        p = g6.plan(ni, n)
        sweep.append(dict(ni=ni, ms_step=round(t, 5), ms_force_kernel=round(tk, 5),
                          ginteractions_s=round(ni * (n - 1) / (t * 1e-3) / 1e9, 2),
                          kernel_pct_fp32_peak=round(100 * FLOPS_PER_INTERACTION * ni * (n - 1)
                                                     / (tk * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 2),
                          plan=dict(bi=p[0], nsplit=p[1], jper=p[2])))
        log(f"sweep ni={ni}: {sweep[-1]['ginteractions_s']} Gint/s, kernel "
            f"{sweep[-1]['kernel_pct_fp32_peak']}% of FP32 peak")

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "pg6_hermite_plummer", "n": n, "eps": EPS,
                   "i_batch": n, "j": n, "block_levels": "2^-3..2^-12",
                   "step": "predict all j + force/jerk/pot/nn of all i",
                   "parallelism": "single GPU, j-range split"},
        "pct_fp32_peak": round(100 * achieved / FP32_PEAK_TFLOPS, 2),
        "roofline": {"bound": "fp32-valu", "kernel": "pg6_force",
                     "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "flops_per_interaction": FLOPS_PER_INTERACTION,
                     "interactions_per_launch": inter, "kernel_ms": round(k_ms, 5),
                     "traffic": read_traffic()},
        "sweep": sweep,
    }

    if not args.no_cpu:
        result.update(cpu_leg(g6, x, v, xi, vi, m, idx, acc2, jrk6, tj, gpu_acc, gpu_jrk,
                              gpu_pot, value))
    print(json.dumps(result), flush=True)


def cpu_leg(g6, x, v, xi, vi, m, idx, acc2, jrk6, tj, gpu_acc, gpu_jrk, gpu_pot, gpu_value):
    """Accuracy against the FP64 oracle and the reference library timed on this
    host's cores, on a seeded sample of 2048 i-particles against all N j."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    n = len(m)
    rng = np.random.default_rng(SEED + 2)
    sel = np.sort(rng.choice(n, 2048, replace=False))
    eps2 = EPS * EPS
    t0 = time.time()
    xp, vp = O.pg6_predict(TBLOCK, x, v, acc2, jrk6, tj)
    o = O.pg6_force(xi[sel], vi[sel], idx[sel], eps2, xp, vp, m, idx)
    log(f"FP64 oracle on {len(sel)} i: {time.time() - t0:.1f} s")

    def stats(a, b):
        e = (np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1) if a.ndim == 2
             else np.abs(a - b) / np.abs(b))
        return {"median": float(f"{np.median(e):.3e}"), "p99": float(f"{np.percentile(e, 99):.3e}"),
                "max": float(f"{e.max():.3e}"),
                "normwise": float(f"{np.abs(a - b).sum() / np.abs(b).sum():.3e}")}

    acc_stats = {"acc": stats(gpu_acc[sel], o["acc"]), "jerk": stats(gpu_jrk[sel], o["jrk"]),
                 "pot": stats(gpu_pot[sel], o["pot"]), "sample_i": int(len(sel)),
                 "reference": "FP64 oracle (oracle/pg_oracle.c) on the same inputs"}
    out = {"accuracy": acc_stats}
    if not O.ref_available("pg6"):
        out["cpu_baseline"] = None
        return out
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    R = O.RefPG6()
    R.set_j(x, v, acc2, jrk6, m, tj, idx)
    times = []
    r = None
    for k in range(6):
        t0 = time.perf_counter()
        r = R.calc(n, TBLOCK, xi[sel], vi[sel], idx[sel], eps2, predict=True)
        times.append(time.perf_counter() - t0)
    tmed = float(np.median(times[1:]))
    cpu_value = len(sel) * (n - 1) / tmed / 1e9
    log(f"reference CPU: {tmed * 1e3:.1f} ms for {len(sel)} i ({cpu_value:.3f} Gint/s, "
        f"{threads} threads)")
    out["cpu_baseline"] = {
        "value": round(cpu_value, 4), "unit": "Ginteractions/s", "cores": threads,
        "kind": "reference",
        "sample": f"{len(sel)} i x {n} j, g6_set_ti + g6calc_firsthalf + g6calc_lasthalf2p in "
                  f"chunks of 48 i (PG6/libpg6 AVX2+OpenMP), median of 5 after 1 warm-up",
        "cpu": cpu_model(), "gpu_over_cpu": round(gpu_value / cpu_value, 1),
        "reference_accuracy": {"acc": stats(r["acc"], o["acc"]), "jerk": stats(r["jrk"], o["jrk"])},
    }
    return out


if __name__ == "__main__":
    main()
