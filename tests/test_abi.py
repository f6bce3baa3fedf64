"""The C-ABI libraries load and export every symbol include/*.h declares.

No GPU work: loading a library registers its code objects but launches
nothing, so this runs on a machine without a GPU.
"""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

import pgamd

INCLUDE = os.path.join(ROOT, "include")


def declared_symbols():
    """{library: set(symbol)} parsed from the headers' 'pgamd-library:' markers."""
    out = {}
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        # drop comments but keep the marker lines they contain
        text = re.sub(r"/\*.*?\*/",
                      lambda m: "\n".join(re.findall(r"pgamd-library:[^\n]*", m.group(0))) + "\n",
                      text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        libs = None
        for part in re.split(r"(pgamd-library:[^\n]*)", text):
            if part.startswith("pgamd-library:"):
                libs = part.split(":", 1)[1].split()
                continue
            if libs is None:
                continue
            body = "\n".join(l for l in part.splitlines() if not l.strip().startswith("#"))
            body = body.replace('extern "C" {', " ").replace("}", " ")
            for stmt in body.split(";"):
                stmt = " ".join(stmt.split())
                md = re.match(r"extern\s+\w+\s+(\w+)\s*\[", stmt)
                mf = re.match(r"(?:const\s+)?(?:\w+\s+)*\w+[\s\*]+(\w+)\s*\(", stmt)
                name = md.group(1) if md else (mf.group(1) if mf else None)
                if name:
                    for lib in libs:
                        out.setdefault(lib, set()).add(name)
    return out


def test_headers_declare_the_reference_api():
    syms = declared_symbols()
    assert {"g5_open", "g5_calculate_force_on_x", "g5_set_xmj", "g5_calculate_force_on_x_"} <= syms["pg5"]
    assert {"g5c_set_xmjMC", "g5c_calculate_force_on_xMC", "g5c_runMC0"} <= syms["pg5"]
    assert {"Force_table", "pg5_gen_force_table", "g5_set_range"} <= syms["pg5_table"]
    assert {"g6calc_lasthalf2p", "g6_get_neighbour_list", "g6_set_j_particle_"} <= syms["pg6"]


@pytest.mark.parametrize("lib", pgamd.LIBS)
def test_library_exports_declared_symbols(lib):
    path = pgamd.lib_path(lib)
    assert os.path.exists(path), f"{path} not built"
    handle = ctypes.CDLL(path, mode=os.RTLD_LOCAL)
    want = declared_symbols()[lib]
    missing = sorted(s for s in want if not hasattr(handle, s))
    assert not missing, f"lib{lib}.so lacks {missing}"


@pytest.mark.parametrize("lib", pgamd.LIBS)
def test_library_exports_nothing_private(lib):
    """Only C entry points (and the kernels' host-side launch stubs) are
    exported; the shared helpers in namespace pgamd stay hidden."""
    out = subprocess.run(["nm", "-D", "--defined-only", pgamd.lib_path(lib)], capture_output=True,
                         text=True, check=True).stdout
    names = [l.split()[-1] for l in out.splitlines() if l.split() and l.split()[-2] in "TDB"]
    helpers = [n for n in names if n.startswith("_ZN5pgamd")]
    assert not helpers, helpers[:5]
    want = declared_symbols()[lib]
    extra = sorted(n for n in names if n not in want and not n.startswith("_"))
    assert not extra, f"lib{lib}.so exports undeclared {extra}"
