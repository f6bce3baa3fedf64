#!/usr/bin/env python3
"""Benchmark of the GRAPE-6 Hermite force path on one MI355X.

Workload (the configuration the headline metric is quoted on): Plummer sphere,
N = 65536, G = M = 1, equal masses, eps = 1/256, FP64 positions/velocities,
j-particles carrying acc/2 and jerk/6 from a start-up force pass and block
time steps 2^-3..2^-12 (Aarseth criterion), predicted to the block time.
One step = one Hermite predictor pass over all j plus one force + jerk +
potential + nearest-neighbour evaluation of all N i-particles against all N
j-particles (N (N-1) interactions), inputs resident in HBM.

Prints one JSON line.  Fields beyond the metric, value and configuration:
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


def cpu_info():
    """Host CPU as the CPU baselines see it: model name (lscpu, else
    /proc/cpuinfo), the host's physical cores (sockets x cores per socket),
    and the logical CPUs this process may run on (its affinity mask, the
    box's share of the host)."""
    info = {"model": platform.processor() or "unknown", "host_physical_cores": None,
            "process_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        f = {l.split(":", 1)[0].strip(): l.split(":", 1)[1].strip()
             for l in out.splitlines() if ":" in l}
        if f.get("Model name"):
            info["model"] = f["Model name"]
        if f.get("Socket(s)", "").isdigit() and f.get("Core(s) per socket", "").isdigit():
            info["host_physical_cores"] = int(f["Socket(s)"]) * int(f["Core(s) per socket"])
    except OSError:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                if len(name) > len(info["model"]):
                    info["model"] = name
                break
    except OSError:
        pass
    return info


def block_steps(acc, jrk, eta=0.02):
    """Aarseth-type dt = eta |a|/|j| rounded down to a power of two in [2^-12, 2^-3]."""
    import numpy as np
    a = np.linalg.norm(acc, axis=1)
    j = np.linalg.norm(jrk, axis=1)
    dt = eta * a / np.maximum(j, 1e-300)
    lev = np.clip(np.ceil(-np.log2(np.maximum(dt, 1e-300))), 3, 12)
    return 2.0 ** (-lev)


def read_traffic(plan):
    """HBM bytes per pg6_force launch from the committed rocprofv3 PMC summary,
    if it was measured on the launch shape this run timed (kernel variant, i
    per block, j-range count); else None.  Returns (bytes, note)."""
    p = os.path.join(ROOT, "profiles", "pg6_force_n65536_traffic.txt")
    if not os.path.exists(p):
        return None, "no traffic profile"
    val, shape = None, None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            val = float(line.split(":", 1)[1])
        elif line.startswith("launch_shape:"):
            shape = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
    here = {"variant": str(plan.get("variant")), "bi": str(plan.get("bi")),
            "nsplit": str(plan.get("nsplit"))}
    if val is None or shape != here:
        return None, f"profiled launch {shape} differs from the timed one {here}"
    return val, f"rocprofv3 FETCH_SIZE + WRITE_SIZE of the timed launch shape {here}"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel", type=int, default=-1,
                    help="PG6 force kernel: 0 LDS j-tiles, 1 SGPR broadcast, 2 SGPR broadcast "
                         "with two-float positions, 3 one launch (prediction and reduction "
                         "inside), 4 Morton-sorted chunk-origin positions, 5 as 4 with the "
                         "j-range of each 64-i block split over its wavefronts (default: "
                         "library's)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the library's default launch shapes instead of autotuning them")
    ap.add_argument("--launch", default="",
                    help="fixed PG6 launch 'variant,bi,blocks_per_cu' (no autotuning), e.g. 1,256,16")
    ap.add_argument("--no-sweep", action="store_true", help="skip the i-batch size sweep")
    ap.add_argument("--sweep-sizes", default="",
                    help="comma list of i-batch sizes for the sweep (default 64, 128, ..., 65536)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy 5 times (reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--configs", default="c1,c2i,c3,c4,c5",
                    help="extra configurations to measure (comma list, '' for none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: 5 copies of {a.numel() * 4} bytes")
        return

    import ics
    import pgamd

    dev = torch.device("cuda:0")
    log(f"device {torch.cuda.get_device_name(0)}")
    g6 = pgamd.G6()
    L = g6.lib
    if args.kernel >= 0:
        assert L.pg6_dev_set_kernel(args.kernel) == 0
    if args.launch:
        lv, lbi, lbpc = (int(t) for t in args.launch.split(","))
        assert L.pg6_dev_set_kernel(lv) == 0 and L.pg6_dev_set_launch(lbi, lbpc, 0) == 0
        args.no_autotune = True
    n = N
    m, x, v = ics.plummer(n, SEED)
    idx = np.arange(n, dtype=np.int32)
    # A dedicated stream at the default priority: torch's default stream has
    # the null handle, which the library reads as "the board's own stream".
    # Every library call and every timing event below uses this stream.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream
    assert sp, "expected a non-null HIP stream handle"

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

    tuned = None
    if not args.no_autotune:
        # untimed: pick kernel and launch shape per i-count on these j-particles
        assert L.pg6_dev_predict(n, TBLOCK, sp) == 0
        t0 = time.time()
        tuned = g6.autotune(n, EPS * EPS)
        log(f"autotune ({time.time() - t0:.1f} s): " + ", ".join(
            f"{ni}:v{c[0]}/bi{c[1]}/bpc{c[2]}" for ni, c in tuned.items()))
    plan = g6.plan(n, n)
    log(f"launch plan ni=nj={n}: {plan}")
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
    sizes = ([int(t) for t in args.sweep_sizes.split(",") if t] if args.sweep_sizes
             else [64 << k for k in range(11)])
    for ni in ([] if args.no_sweep else sizes):
        sel = np.sort(rng.choice(n, ni, replace=False))
        ip = d_xi[torch.from_numpy(sel).to(dev)].contiguous()
        iv = d_vi[torch.from_numpy(sel).to(dev)].contiguous()
        ii = d_id[torch.from_numpy(sel).to(dev)].contiguous()
        o = dev_arrays(ni)
        reps = max(3, min(200, int(2e10 / (ni * n))))
        for _ in range(2):
            force(ni, ip, iv, ii, o)
        # the step time without the kernel's own timing events (at small ni
        # they would add several microseconds), then the kernel time
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            assert L.pg6_dev_predict(n, TBLOCK, sp) == 0
            force(ni, ip, iv, ii, o)
        e1.record(stream)
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / reps
        L.pg6_dev_timing_start(reps)
        for _ in range(reps):
            assert L.pg6_dev_predict(n, TBLOCK, sp) == 0
            force(ni, ip, iv, ii, o)
        torch.cuda.synchronize()
        L.pg6_dev_timing_read(pgamd.dptr(kms), pgamd.iptr(kcount))
        tk = float(kms[0]) / max(int(kcount[0]), 1)
        p = g6.plan(ni, n)
        sweep.append(dict(ni=ni, ms_step=round(t, 5), ms_force_kernel=round(tk, 5),
                          ginteractions_s=round(ni * (n - 1) / (t * 1e-3) / 1e9, 2),
                          kernel_pct_fp32_peak=round(100 * FLOPS_PER_INTERACTION * ni * (n - 1)
                                                     / (tk * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 2),
                          plan=p))
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
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pg6_hermite_plummer", "n": n, "eps": EPS,
                   "i_batch": n, "j": n, "block_levels": "2^-3..2^-12",
                   "step": "predict all j + force/jerk/pot/nn of all i",
                   "parallelism": "single GPU, j-range split",
                   "launch": plan, "autotuned": tuned is not None},
        "pct_fp32_peak": round(100 * achieved / FP32_PEAK_TFLOPS, 2),
        "roofline": {"bound": "mfma",
                     "bound_note": "compute-bound (HBM traffic < 1% of its roofline): the pair "
                                   "math runs on the FP32 vector ALU, not on MFMA, against "
                                   "MI355X's FP32 peak of 157.3 TFLOP/s (the dense FP32 figure)",
                     "kernel": "pg6_force",
                     "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "flops_per_interaction": FLOPS_PER_INTERACTION,
                     "interactions_per_launch": inter, "kernel_ms": round(k_ms, 5),
                     "traffic": read_traffic(plan)[0], "traffic_note": read_traffic(plan)[1]},
        "sweep": sweep,
    }

    if not args.no_cpu:
        result.update(cpu_leg(g6, x, v, xi, vi, m, idx, acc2, jrk6, tj, gpu_acc, gpu_jrk,
                              gpu_pot, value))
    extra = {}
    for c in [c for c in args.configs.split(",") if c]:
        t0 = time.time()
        extra[c] = EXTRA[c](args, stream, not args.no_cpu)
        log(f"{c}: {json.dumps(extra[c])[:400]} ({time.time() - t0:.0f} s)")
    if extra:
        result["configs"] = extra
    print(json.dumps(result), flush=True)


def cpu_leg(g6, x, v, xi, vi, m, idx, acc2, jrk6, tj, gpu_acc, gpu_jrk, gpu_pot, gpu_value):
    """Accuracy against the FP64 oracle on a seeded sample of 2048 i-particles,
    and the reference library timed on this host's cores over the whole step
    (all N i-particles against all N j)."""
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
    threads = _threads()
    R = O.RefPG6()
    R.set_j(x, v, acc2, jrk6, m, tj, idx)
    times = []
    r = None
    for k in range(6):
        t0 = time.perf_counter()
        r = R.calc(n, TBLOCK, xi, vi, idx, eps2, predict=True)
        times.append(time.perf_counter() - t0)
    tmed = float(np.median(times[1:]))
    cpu_value = n * (n - 1) / tmed / 1e9
    log(f"reference CPU: {tmed * 1e3:.1f} ms per step ({cpu_value:.3f} Gint/s, "
        f"{threads} threads)")
    out["cpu_baseline"] = {
        "value": round(cpu_value, 4), "unit": "Ginteractions/s", "cores": threads,
        "kind": "reference",
        "sample": f"the whole step, {n} i x {n} j: g6_set_ti + g6calc_firsthalf + "
                  f"g6calc_lasthalf2p in chunks of 48 i (PG6/libpg6 AVX2+OpenMP), median of 5 "
                  f"after 1 warm-up, {tmed:.2f} s per step",
        "cpu": cpu_info(), "gpu_over_cpu": round(gpu_value / cpu_value, 1),
        "reference_accuracy": {"acc": stats(r["acc"][sel], o["acc"]),
                               "jerk": stats(r["jrk"][sel], o["jrk"])},
    }
    return out


# ------------------------------------------------------------ other configs
def _errs(a, b):
    import numpy as np
    e = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    return {"median": float(f"{np.median(e):.3e}"), "p99": float(f"{np.percentile(e, 99):.3e}"),
            "normwise": float(f"{np.abs(a - b).sum() / np.abs(b).sum():.3e}")}


def _time(fn, stream, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _threads():
    """OpenMP threads of the CPU baselines: OMP_NUM_THREADS (the box sets it to
    its CPU share), else the CPUs this process may use."""
    return int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def bench_c1(args, stream, cpu):
    """PG5 direct sum: Plummer N = 16384, eps = 1/64, one N x N force+pot."""
    import numpy as np
    import torch

    import ics
    import pgamd
    n, eps = 16384, 1.0 / 64
    m, x, _ = ics.plummer(n, 16384)
    g5 = pgamd.G5()
    xi = np.zeros((n, 4), np.float32)
    xi[:, :3] = x
    xi[:, 3] = eps * eps
    xj = np.zeros((n, 4), np.float32)
    xj[:, :3] = x
    xj[:, 3] = m
    d_xi = torch.from_numpy(xi).cuda()
    d_xj = torch.from_numpy(xj).cuda()
    acc = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    pot = torch.empty(n, dtype=torch.float64, device="cuda")
    sp = stream.cuda_stream

    def run():
        assert g5.lib.pg5_dev_force(0, n, d_xi.data_ptr(), n, d_xj.data_ptr(), None,
                                    acc.data_ptr(), pot.data_ptr(), sp) == 0
    # a call is 0.1 ms: half a second of calls first, so that the timed ones
    # run at the clock a sustained load holds (this config follows CPU work)
    t0 = time.time()
    while time.time() - t0 < 0.5:
        for _ in range(50):
            run()
        torch.cuda.synchronize()
    kernels = {}
    for variant, name in ((0, "lds_tiles"), (1, "scalar_broadcast")):
        g5.lib.pg5_dev_set_kernel(variant)
        kernels[name] = round(_time(run, stream, 200), 4)
    g5.lib.pg5_dev_set_kernel(0)
    ms = _time(run, stream, 200)
    inter = float(n) * n
    out = {"workload": "pg5_plummer_n16384_eps1/64", "interactions": inter,
           "ms": round(ms, 4), "ginteractions_s": round(inter / ms / 1e6, 1),
           "pct_fp32_peak": round(100 * 38 * inter / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 2),
           "flops_per_interaction": 38, "kernel_ms": kernels}
    O = _oracle()
    sel = np.sort(np.random.default_rng(3).choice(n, 1024, replace=False))
    oa, _ = O.pg5_force(x[sel], eps * eps, x, m)
    out["accuracy_vs_fp64"] = _errs(acc.cpu().numpy()[sel], oa)
    if cpu and O.ref_available("pg5_omp"):
        R = O.RefPG5(omp=True)
        ts = []
        for _ in range(6):
            t0 = time.perf_counter()
            ra, _ = R.force(x, x, m, eps)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts[1:]))
        out["cpu_baseline"] = {"ginteractions_s": round(n * n / t / 1e9, 3),
                               "kind": "reference",
                               "sample": f"the whole N x N = {n} x {n} g5_calculate_force_on_x "
                                         "(OpenMP build), median of 5 after 1 warm-up",
                               "seconds": round(t, 4), "threads": _threads(),
                               "accuracy_vs_fp64": _errs(ra[sel], oa)}
    if cpu:
        out["dropin_direct"] = _direct_dropin()
    return out


def bench_c3(args, stream, cpu):
    """TreePM short range: N = 2^20 clumpy periodic box, mesh 128^3,
    r_cut = 3/128, eps_PP = 0.1/128; 1024-entry table vs learned MLP cut-off,
    all groups in one graph-captured persistent launch."""
    import numpy as np
    import torch

    import ics
    import pgamd
    pp, pm = 0.1 / 128, 3.0 / 128
    m, x = ics.clumpy_box(1 << 20, 20)
    sp = stream.cuda_stream
    t0 = time.time()
    Lh = pgamd.treepm_lists(x, pm, 256)
    tb = time.time() - t0
    # the same lists built on the GPU from positions in HBM (identical arrays)
    d_x = torch.from_numpy(x).cuda()
    d_m = torch.from_numpy(m).cuda()
    torch.cuda.synchronize()
    tgpu = []
    for _ in range(4):
        t0 = time.perf_counter()
        DL = pgamd.DevLists.treepm(d_x.data_ptr(), len(x), pm, 256, stream=sp)
        tgpu.append(time.perf_counter() - t0)
        if len(tgpu) < 4:
            DL.free()
    L = DL.host()
    assert all(np.array_equal(L[k], Lh[k]) for k in ("pos", "src", "ioff", "jlist", "order"))
    T = pgamd.G5Table()
    T.lib.pg5_gen_s2_force_table(pp, pm)
    t0 = time.time()
    loss = T.mlp_train(iters=200, starts=4, refine=3000, seed=7)
    ttrain = time.time() - t0
    P = L["pos"]
    mm = m[L["src"]]
    d_pos = torch.empty((len(P), 4), dtype=torch.float32, device="cuda")
    d = {k: DL.ptr(k) for k in ("ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "order")}
    d_pos64 = DL.ptr("pos")  # the list builder's FP64 positions (ghosts included)
    acc = torch.zeros((len(P), 3), dtype=torch.float64, device="cuda")
    ng = len(L["ioff"])
    res = {"workload": "treepm_clumpy_box_n2^20_mesh128", "groups": ng,
           "interactions": L["work_pp"], "list_build_host_s": round(tb, 3),
           "list_build_gpu_s": round(float(np.median(tgpu[1:])), 4),
           "mlp_train": {"s": round(ttrain, 2), "loss": loss,
                         "method": "Levenberg-Marquardt, FP64 GPU kernels, 4 starts x 200 trial steps, best + 3000"}}
    O = _oracle()
    sel = np.sort(np.random.default_rng(4).choice(len(x), 512, replace=False))
    f = O.table_force_f64(P[sel], P, mm, pp, pm)
    tab5, _, xs5 = O.table_build(pp, pm, 4, 5)
    # the reference's own output on these particles: its 512-entry table on
    # FP32 positions (the oracle's bit-exact emulation of gravity_kernel_avx2.c)
    fref = O.table_force_emul(P[sel], P, mm, tab5, xs5, 4, 5)
    res["reference_table_vs_closed_form_s2"] = _errs(fref, f)
    # (name, cut-off mode, table frc_bit, table order, FP64 group origins)
    runs = [("table1024", 0, 6, 1, False), ("table1024_origin", 0, 6, 1, True),
            ("table1024_quadratic_origin", 0, 6, 2, True),
            ("table8192_origin", 0, 9, 1, True), ("mlp_origin", 1, 6, 1, True),
            ("mlp", 1, 6, 1, False)]
    for name, mode, frc, order, p64 in runs:
        T.lib.pg5_set_table_order(1)
        T.lib.pg5_set_table_frc_bit(frc)
        T.lib.pg5_set_table_order(order)
        _, xs = T.table()
        DL.pack(d_m.data_ptr(), xs, d_pos.data_ptr(), None, sp)
        torch.cuda.synchronize()
        args_ = (0, ng, d["ioff"], d["icnt"], d["ilist"], d["joff"], d["jcnt"], d["jlist"],
                 d["order"], d_pos.data_ptr(), d_pos64 if p64 else None, acc.data_ptr(), mode)
        best = np.zeros(1, np.int32)
        T.lib.pg5t_batch_autotune(*args_, 2, pgamd.iptr(best), None)
        shape = np.zeros(2, np.int32)
        T.lib.pg5t_batch_launch_shape(mode, pgamd.iptr(shape))
        h = T.lib.pg5t_batch_graph(*args_)
        ms = _time(lambda: T.lib.pg5t_batch_graph_launch(h, sp), stream, 5)
        T.lib.pg5t_batch_graph_destroy(h)
        a = acc.cpu().numpy()[sel]
        res[name] = {"ms": round(ms, 3), "ginteractions_s": round(L["work_pp"] / ms / 1e6, 1),
                     "pct_fp32_peak": round(100 * 38 * L["work_pp"] / (ms * 1e-3) / 1e12
                                            / FP32_PEAK_TFLOPS, 2),
                     "block": int(shape[0]), "blocks_per_cu": int(shape[1]),
                     "table": f"frc_bit {frc}, order {order}",
                     "positions": "FP64 group origins" if p64 else "FP32 scaled",
                     "accuracy_vs_closed_form_s2": _errs(a, f),
                     "accuracy_vs_reference_table": _errs(a, fref)}
        log(f"c3 {name}: {ms:.3f} ms, p99 {res[name]['accuracy_vs_closed_form_s2']['p99']}")
    T.lib.pg5_set_table_order(1)
    T.lib.pg5_set_table_frc_bit(6)
    _, xs = T.table()
    # the learned path against the table at equal accuracy: the cheapest table
    # configuration whose p99 is at most the network's
    net = res["mlp_origin"]
    eq = [k for k in ("table1024_quadratic_origin", "table8192_origin", "table1024_origin")
          if res[k]["accuracy_vs_closed_form_s2"]["p99"] <= net["accuracy_vs_closed_form_s2"]["p99"]]
    best_tab = min(eq, key=lambda k: res[k]["ms"]) if eq else None
    res["learned_vs_table_at_equal_error"] = {
        "table": best_tab, "mlp_p99": net["accuracy_vs_closed_form_s2"]["p99"],
        "table_p99": res[best_tab]["accuracy_vs_closed_form_s2"]["p99"] if best_tab else None,
        "mlp_over_table_speed": round(res[best_tab]["ms"] / net["ms"], 4) if best_tab else None}
    res["mlp_over_table_speed"] = round(res["table1024"]["ms"] / res["mlp"]["ms"], 3)
    # long-range side on the 128^3 mesh, and the whole TreePM force step
    acc_pm = torch.zeros((len(x), 3), dtype=torch.float64, device="cuda")
    pm_ms = _time(lambda: T.lib.pg5t_pm_force(len(x), d_x.data_ptr(), d_m.data_ptr(), 128, pm,
                                              acc_pm.data_ptr(), sp), stream, 5)
    res["pm_mesh128_ms"] = round(pm_ms, 3)
    res["treepm_step_ms"] = {"gpu_lists": round(1e3 * res["list_build_gpu_s"], 2),
                             "short_range_table": res["table1024"]["ms"],
                             "long_range_pm": round(pm_ms, 3),
                             "total": round(1e3 * res["list_build_gpu_s"]
                                            + res["table1024"]["ms"] + pm_ms, 2)}
    T.lib.pg5_set_table_frc_bit(5)
    DL.free()
    if cpu and O.ref_available("pg5table"):
        thr = _threads()
        groups = np.arange(ng)
        t, ra = O.ref_treepm(L, P, mm, pp, pm, thr, groups)
        res["cpu_baseline"] = {
            "ginteractions_s": round(L["work_pp"] / t / 1e9, 3), "kind": "reference",
            "sample": f"all {ng} groups: g5_set_xmjMC + g5_calculate_force_on_xMC per group, "
                      "OpenMP over groups with devid = thread (reference 512-entry table)",
            "threads": min(thr, 24), "seconds": round(t, 3),
            "accuracy_vs_closed_form_s2": _errs(ra[sel], f)}
    if cpu:
        res["dropin_force_prof"] = _force_prof_dropin()
    return res


def bench_c4(args, stream, cpu):
    """Barnes-Hut: Hernquist N = 2^21 (a = 1, truncated at 50a), theta = 0.5,
    groups <= 512, leaves <= 8, eps = 1/128, particle + quadrupole cell lists,
    all groups in one graph-captured persistent launch."""
    import numpy as np
    import torch

    import ics
    import pgamd
    n, eps = 1 << 21, 1.0 / 128
    m, x = ics.hernquist(n, 21)
    sp = stream.cuda_stream
    t0 = time.time()
    Lh = pgamd.bh_lists(x, m, 0.5, 512, 8)
    tb = time.time() - t0
    # the same lists built on the GPU (cell numbering aside, identical arrays)
    d_x = torch.from_numpy(x).cuda()
    d_m = torch.from_numpy(m).cuda()
    torch.cuda.synchronize()
    tgpu = []
    for _ in range(4):
        t0 = time.perf_counter()
        DL = pgamd.DevLists.bh(d_x.data_ptr(), d_m.data_ptr(), n, 0.5, 512, 8, stream=sp)
        tgpu.append(time.perf_counter() - t0)
        if len(tgpu) < 4:
            DL.free()
    L = DL.host()
    assert all(np.array_equal(L[k], Lh[k]) for k in ("pos", "src", "ioff", "jlist", "order"))
    g5 = pgamd.G5()
    P = L["pos"]
    mm = m[L["src"]]
    keys = ("ioff", "icnt", "ilist", "joff", "jcnt", "jlist", "coff", "ccnt", "clist", "order")
    d = {k: DL.ptr(k) for k in keys}
    d_pos = torch.empty((len(P), 4), dtype=torch.float32, device="cuda")
    d_cells = torch.empty((len(L["cells"]), 12), dtype=torch.float32, device="cuda")
    DL.pack(d_m.data_ptr(), 1.0, d_pos.data_ptr(), d_cells.data_ptr(), sp)
    torch.cuda.synchronize()
    acc = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    pot = torch.zeros(n, dtype=torch.float64, device="cuda")
    ng = len(L["ioff"])
    args_ = (0, ng, *[d[k] for k in keys], d_pos.data_ptr(), d_cells.data_ptr(),
             eps * eps, acc.data_ptr(), pot.data_ptr())
    best = np.zeros(1, np.int32)
    g5.lib.pg5_bh_autotune(*args_, 2, pgamd.iptr(best), None)
    shape = np.zeros(2, np.int32)
    g5.lib.pg5_bh_launch_shape(pgamd.iptr(shape))
    h = g5.lib.pg5_bh_graph(*args_)
    ms = _time(lambda: g5.lib.pg5_bh_graph_launch(h, stream.cuda_stream), stream, 5)
    g5.lib.pg5_bh_graph_destroy(h)
    flops = 38 * L["work_pp"] + 71 * L["work_pc"]
    res = {"workload": "bh_hernquist_n2^21_theta0.5", "groups": ng,
           "particle_interactions": L["work_pp"], "cell_interactions": L["work_pc"],
           "list_build_host_s": round(tb, 3),
           "list_build_gpu_s": round(float(np.median(tgpu[1:])), 4), "ms": round(ms, 3),
           "ginteractions_s": round((L["work_pp"] + L["work_pc"]) / ms / 1e6, 1),
           "pct_fp32_peak": round(100 * flops / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 2),
           "flops": "38/particle + 71/cell interaction", "block": int(shape[0]),
           "blocks_per_cu": int(shape[1])}
    O = _oracle()
    sel = np.sort(np.random.default_rng(5).choice(n, 256, replace=False))
    a = acc.cpu().numpy()[sel]
    ad, _ = O.pg5_force(P[sel], eps * eps, P, mm)
    res["accuracy_vs_fp64_direct"] = _errs(a, ad)
    DL.free()
    if cpu and O.ref_available("pg5quad"):
        thr = _threads()
        t, ra, _ = O.ref_bh(L, P, mm, L["cells"], eps, thr, np.arange(ng))
        res["cpu_baseline"] = {
            "ginteractions_s": round((L["work_pp"] + L["work_pc"]) / t / 1e9, 3),
            "kind": "reference",
            "sample": f"all {ng} groups: g5 particle + g5c cell calls per group, OpenMP over "
                      "groups with devid = thread",
            "threads": min(thr, 24), "seconds": round(t, 3),
            "accuracy_vs_fp64_direct": _errs(ra[sel], ad)}
    return res


def bench_c5(args, stream, cpu):
    """4096 independent Plummer clusters, N log-uniform in [64, 4096], eps = 4/N,
    force + jerk + pot of every member, one graph-captured persistent launch."""
    import numpy as np
    import torch

    import ics
    import pgamd
    sizes = ics.cluster_sizes(4096, 4096)
    xs, vs, ms_ = [], [], []
    for c, nc in enumerate(sizes):
        mc, xc, vc = ics.plummer(int(nc), 50000 + c)
        xs.append(xc)
        vs.append(vc)
        ms_.append(mc)
    X, V, Mm = np.concatenate(xs), np.concatenate(vs), np.concatenate(ms_)
    coff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    e2 = ((4.0 / sizes) ** 2).astype(np.float32)
    g6 = pgamd.G6()
    items = g6.cluster_items(coff)
    nt = len(Mm)
    dd = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda()
          for k, v in (("items", items), ("coff", coff), ("x", X), ("v", V), ("m", Mm),
                       ("e2", e2))}
    acc = torch.zeros((nt, 3), dtype=torch.float64, device="cuda")
    jrk = torch.zeros((nt, 3), dtype=torch.float64, device="cuda")
    pot = torch.zeros(nt, dtype=torch.float64, device="cuda")
    args_ = (len(items) // 2, dd["items"].data_ptr(), dd["coff"].data_ptr(), dd["x"].data_ptr(),
             dd["v"].data_ptr(), dd["m"].data_ptr(), dd["e2"].data_ptr(), acc.data_ptr(),
             jrk.data_ptr(), pot.data_ptr())
    best = np.zeros(1, np.int32)
    g6.lib.pg6_cluster_autotune(*args_, 2, pgamd.iptr(best), None)
    shape = np.zeros(2, np.int32)
    g6.lib.pg6_cluster_launch_shape(pgamd.iptr(shape))
    h = g6.lib.pg6_cluster_graph(*args_)
    ms = _time(lambda: g6.lib.pg6_cluster_graph_launch(h, stream.cuda_stream), stream, 5)
    g6.lib.pg6_cluster_graph_destroy(h)
    inter = float((sizes * (sizes - 1)).sum())
    res = {"workload": "4096_plummer_clusters_n64-4096", "particles": int(nt),
           "interactions": inter, "ms": round(ms, 3),
           "ginteractions_s": round(inter / ms / 1e6, 1),
           "pct_fp32_peak": round(100 * 60 * inter / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 2),
           "j_tile": int(shape[0]), "blocks_per_cu": int(shape[1])}
    O = _oracle()
    a, j = acc.cpu().numpy(), jrk.cpu().numpy()
    ea, ej, oa, oj = [], [], [], []
    for c in np.random.default_rng(6).choice(len(sizes), 24, replace=False):
        s = slice(coff[c], coff[c + 1])
        idx = np.arange(sizes[c], dtype=np.int32)
        o = O.pg6_force(X[s], V[s], idx, (4.0 / sizes[c]) ** 2, X[s], V[s], Mm[s], idx)
        ea.append(a[s]); oa.append(o["acc"]); ej.append(j[s]); oj.append(o["jrk"])
    res["accuracy_vs_fp64"] = {"acc": _errs(np.concatenate(ea), np.concatenate(oa)),
                               "jerk": _errs(np.concatenate(ej), np.concatenate(oj)),
                               "sample": "24 whole clusters"}
    if cpu and O.ref_available("pg6"):
        sel = np.arange(len(sizes))
        t, *_ = O.ref_clusters(sel, coff, X, V, Mm, (4.0 / sizes) ** 2)
        res["cpu_baseline"] = {
            "ginteractions_s": round(inter / t / 1e9, 3), "kind": "reference",
            "sample": "all 4096 clusters, one PG6 session each (set_j_particle per member, "
                      "firsthalf + lasthalf2 in chunks of 48)",
            "threads": _threads(), "seconds": round(t, 3)}
    return res


def bench_c2i(args, stream, cpu):
    """Block-time-step Hermite integration on the GPU (the reference's sample
    program PG6/banana with its published setup, user guide 5.3): Plummer
    N = 16384, eps = 0, eta_s = 0.01, eta = 0.2, dtmax = 1/8, t = 0 -> 1/8."""
    import numpy as np
    import torch

    import ics
    import pgamd
    n, t_end = 16384, 0.125
    g6 = pgamd.G6()
    # warm-up: a short integration loads and runs every kernel once
    mw, xw, vw = ics.plummer(1024, 5)
    g6.hermite_init(xw, vw, mw, 0.0, 0.01, 0.2, 0.125)
    g6.hermite_run(1.0 / 64)
    m, x, v = ics.plummer(n, 16391)
    t0 = time.perf_counter()
    g6.hermite_init(x, v, m, 0.0, 0.01, 0.2, 0.125)
    t_init = time.perf_counter() - t0
    ek0, ep0 = g6.hermite_energy()
    t0 = time.perf_counter()
    blocks = g6.hermite_run(t_end)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = g6.hermite_state()
    ek, ep = g6.hermite_energy()
    g6.lib.pg6_hermite_end()
    steps = st["steps"] - n  # particle steps after the start-up pass
    inter = float(steps) * (n - 1)
    res = {"workload": "hermite_block_steps_plummer_n16384_eps0_t0.125", "wall_s": round(wall, 3),
           "init_s": round(t_init, 3), "block_steps": blocks, "particle_steps": steps,
           "mean_block": round(steps / max(blocks, 1), 1),
           "ginteractions_s": round(inter / wall / 1e9, 2),
           "pct_fp32_peak": round(100 * 60 * inter / wall / 1e12 / FP32_PEAK_TFLOPS, 2),
           "energy_error": float(f"{abs(ek + ep - ek0 - ep0) / abs(ek0 + ep0):.3e}"),
           "published_reference": {"tcalc_s": 23.8, "gflops": 18.4, "mean_block": 232,
                                   "source": "doc/user_guide.pdf 5.3 (pl016k.init, CPU not stated)"}}
    O = _oracle()
    if cpu and O.ref_available("pg6"):
        t, _, _, info, en = O.ref_hermite(x, v, m, 0.0, 0.01, 0.2, 0.125, t_end)
        csteps = info["steps"] - n
        res["cpu_baseline"] = {
            "ginteractions_s": round(csteps * (n - 1) / t / 1e9, 3), "kind": "reference",
            "sample": "the whole integration: reference library PG6/libpg6 (AVX2 + OpenMP) "
                      "driven by the banana loop (oracle/cpu_batch.c cb_hermite)",
            "threads": _threads(), "seconds": round(t, 2), "block_steps": info["blocks"] - 1,
            "particle_steps": csteps,
            "energy_error": float(f"{abs(en[-1].sum() - ek0 - ep0) / abs(ek0 + ep0):.3e}"),
            "gpu_speedup": round(t / wall, 1)}
    if cpu:
        res["dropin_banana"] = _banana_dropin(m, x, v, t_end)
    return res


def _direct_dropin():
    """The reference's own PG5 benchmark (user guide 5.1, Listing 6: direct
    with init/pl4k, 2.06e9 interactions/s on its CPU): PG5/newton/direct, built
    from its sources by oracle/ref.mk, on a pl4k-style Plummer N = 4096
    (eps = 0.02 and 1000 leapfrog steps, as direct.c fixes them), linked
    against the reference library and against libpg5.so.  Its own report,
    interactions per second over each 100 steps (direct.c:306-316): the
    median of the nine reports."""
    import re
    import tempfile

    import numpy as np

    import ics
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import dropin
    if not all(os.path.exists(dropin.program(p)) for p in ("direct_ref", "direct_pgamd")):
        return None
    m, x, v = ics.plummer(4096, 4096)
    out = {"setup": "pl4k-style Plummer N = 4096, eps = 0.02, dt = 0.01, 1000 steps",
           "threads": _threads(), "published_reference_interactions_s": 2.060791e9}
    for key, prog in (("reference_library", "direct_ref"), ("libpg5", "direct_pgamd")):
        with tempfile.TemporaryDirectory() as d:
            dropin.write_direct_input(os.path.join(d, "pl"), m, x, v)
            o = dropin.run(prog, ["pl", "out"], d, timeout=600, threads=_threads())
        r = [float(a) for a in re.findall(r"^(\S+) interaction per sec", o, re.M)]
        out[key + "_interactions_s"] = float(f"{np.median(r):.4e}") if r else None
    return out


def _force_prof_dropin():
    """The reference's own table benchmark (user guide 5.2, Listing 11:
    force_prof 32768 0.1, 8.5e8 interactions/s on its CPU): PG5/table/
    force_prof, built from its sources by oracle/ref.mk, linked against the
    reference's serial table library (its setting.mk default) and against
    libpg5_table.so.  It times one g5_calculate_force_on_xMC over all N and
    prints interactions per second on stderr."""
    import re
    import subprocess

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import dropin
    if not all(os.path.exists(dropin.program(p)) for p in ("force_prof_ref", "force_prof_pgamd")):
        return None
    out = {"setup": "force_prof 32768 0.1 (uniform cube + one heavy particle, one call)",
           "published_reference_interactions_s": 8.508592e8}
    for key, prog in (("reference_library", "force_prof_ref"), ("libpg5_table", "force_prof_pgamd")):
        r = subprocess.run([dropin.program(prog), "32768", "0.1"], capture_output=True, text=True,
                           timeout=600, env=dict(os.environ, OMP_NUM_THREADS=str(_threads())))
        v = re.findall(r"(\S+) interactions / sec", r.stderr) if r.returncode == 0 else []
        out[key + "_interactions_s"] = float(v[-1]) if v else None
    return out


def _banana_dropin(m, x, v, t_end):
    """The reference's own Hermite program (PG6/banana, built from its sources
    by oracle/ref.mk) on the same initial conditions, linked against the
    reference library and against libpg6.so: its own reported integration
    time (tcalc).  libpg6 runs once with the reference's 48 i per
    g6calc_lasthalf call and once with g6_npipes() raised to 128, the size of
    banana's own i buffers (PG6/banana/hermite.c:14,218-219)."""
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import dropin
    if not all(os.path.exists(dropin.program(p)) for p in ("banana_ref", "banana_pgamd")):
        return None
    out = {"setup": "pl016k-style Plummer N = 16384, eps = 0, eta = 0.2, eta_s = 0.01, "
                    "dtmax = 1/8, t = 0 -> 1/8", "threads": _threads()}
    for key, prog, env in (("reference_library_s", "banana_ref", None),
                           ("libpg6_s", "banana_pgamd", None),
                           ("libpg6_npipes128_s", "banana_pgamd", {"PGAMD_G6_NPIPES": "128"})):
        with tempfile.TemporaryDirectory() as d:
            dropin.write_banana_input(d, m, x, v, t_end)
            t = dropin.banana_tcalc(dropin.run(prog, [], d, timeout=600, threads=_threads(),
                                               env_extra=env))
            out[key] = None if t is None else round(t, 3)
    return out


EXTRA = {"c1": bench_c1, "c2i": bench_c2i, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5}


if __name__ == "__main__":
    main()
