"""Helpers for running the reference's own sample programs (built by
oracle/ref.mk against the reference library and against this repository's
libraries) on generated inputs."""
import os
import re
import subprocess

import numpy as np

from conftest import ROOT

REFDIR = os.path.join(ROOT, "oracle", "_ref")


def program(name):
    return os.path.join(REFDIR, name)


def write_direct_input(path, m, x, v):
    """PG5/newton/direct/direct.c:34-65 format."""
    with open(path, "w") as f:
        f.write(f"{len(m)}\n3\n0.0\n")
        f.writelines(f"{q:.17e}\n" for q in m)
        f.writelines(f"{r[0]:.17e} {r[1]:.17e} {r[2]:.17e}\n" for r in x)
        f.writelines(f"{r[0]:.17e} {r[1]:.17e} {r[2]:.17e}\n" for r in v)


def write_banana_input(dirname, m, x, v, tend, epsinv=0.0, eta=0.2, eta_s=0.01, dtinv=8.0):
    """PG6/banana/io.c:12-38 particle file and the input.list read by
    banana01.c:42-48."""
    with open(os.path.join(dirname, "pl.init"), "w") as f:
        f.write(f"0.0\n{len(m)}\n")
        f.writelines(f"{i} {m[i]:.9e} {x[i, 0]:.17e} {x[i, 1]:.17e} {x[i, 2]:.17e} "
                     f"{v[i, 0]:.17e} {v[i, 1]:.17e} {v[i, 2]:.17e}\n" for i in range(len(m)))
    with open(os.path.join(dirname, "input.list"), "w") as f:
        f.write(f"0 {tend} 100.0\n{epsinv}\n{eta} {eta_s} {dtinv}\npl.init out\n")


def run(name, args, cwd, timeout=300, threads=4, env_extra=None):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), **(env_extra or {}))
    r = subprocess.run([program(name), *args], cwd=cwd, capture_output=True, text=True,
                       timeout=timeout, env=env)
    assert r.returncode == 0, (name, r.returncode, r.stderr[-2000:])
    return r.stdout


def direct_energy_errors(out):
    """Relative energy errors printed by direct.c every 100 steps."""
    es = [float(a) for a in re.findall(r"^e: (\S+) de:", out, re.M)]
    des = [float(a) for a in re.findall(r"^e: \S+ de: (\S+)", out, re.M)]
    return np.array(des) / np.abs(np.array(es))


def banana_energy_errors(out):
    """de/|e| at every dtmax window printed by banana01.c (io.c:70-90)."""
    return np.array([float(a) for a in re.findall(r"de/\|e\|: (\S+)", out)])


def banana_gflops(out):
    return np.array([float(a) for a in re.findall(r"tcalc: \S+ (\S+) Gflops", out)])


def banana_tcalc(out):
    """The integration time banana01.c reports at its last output (io.c:82)."""
    t = re.findall(r"tcalc: (\S+)", out)
    return float(t[-1]) if t else None
