"""Build recipe for the gfx950 shared libraries (in-tree, into lib/).

    python phantom-grape_amd/build.py

Every HIP source is compiled with ``--offload-arch=gfx950`` and with packed
FP32 instructions disabled (``-Xclang -target-feature -Xclang
-packed-fp32-ops``).  Only ``extern "C"`` entry points are exported
(``-fvisibility=hidden``), so each library keeps private copies of the shared
helpers.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PGAMD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-fvisibility=hidden", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

CXX = os.environ.get("CXX", shutil.which("g++") or "g++")
# host-only C++ libraries (no GPU code): name -> sources
HOST_TARGETS = {
    "pgtree": ["treebuild.cpp"],
}

# library name -> (sources, extra flags)
TARGETS = {
    "pg5": (["pg5.hip", "pgamd_common.cpp"], []),
    "pg5_sym": (["pg5.hip", "pgamd_common.cpp"], ["-DPG5_SYMMETRIC"]),
    "pg5_table": (["pg5_table.hip", "pgamd_common.cpp"], ["-lhipfft"]),
    "pg6": (["pg6.hip", "pgamd_common.cpp"], []),
    "pgtree_dev": (["treebuild_dev.hip", "pgamd_common.cpp"], []),
}


def _cmd(name: str) -> list[str]:
    if name in HOST_TARGETS:
        out = os.path.join(LIBDIR, f"lib{name}.so")
        return [CXX, "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-fvisibility=hidden",
                "-o", out + ".tmp", *[os.path.join(CSRC, s) for s in HOST_TARGETS[name]]]
    srcs, extra = TARGETS[name]
    out = os.path.join(LIBDIR, f"lib{name}.so")
    return [HIPCC, *FLAGS, *extra, "-I", CSRC, "-o", out + ".tmp",
            *[os.path.join(CSRC, s) for s in srcs]]


def build(names=None, verbose=False) -> None:
    os.makedirs(LIBDIR, exist_ok=True)
    names = list(names or (list(TARGETS) + list(HOST_TARGETS)))
    procs = []
    for n in names:
        srcs = HOST_TARGETS[n] if n in HOST_TARGETS else TARGETS[n][0]
        if not all(os.path.exists(os.path.join(CSRC, s)) for s in srcs):
            raise FileNotFoundError(f"sources of lib{n}.so missing")
        cmd = _cmd(n)
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((n, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                          stderr=subprocess.STDOUT, text=True)))
    failed = []
    for n, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append(n)
            sys.stderr.write(f"--- lib{n}.so failed ---\n{out}\n")
        else:
            final = os.path.join(LIBDIR, f"lib{n}.so")
            os.replace(final + ".tmp", final)
            if verbose and out.strip():
                print(out)
    if failed:
        raise RuntimeError(f"HIP build failed for {failed}")


if __name__ == "__main__":
    build(sys.argv[1:] or None, verbose=True)
