"""The C ABI as a C++ caller sees it: tests/abi/abi_check.cpp, compiled with
g++ against include/stzmg.h and include/stz.h and linked to the in-tree
library (built by __graft_entry__.build()).

* layout: sizeof and offsetof of every struct of the two headers against the
  ctypes mirror (stzpp_amd/_lib.py) the Python side passes by pointer;
* run (GPU): one GMRES-IR projection solve, one QS step and two error
  returns through the C++ binding, with the program's own checks."""
import ctypes
import json
import os
import subprocess

import pytest

from stzpp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "abi", "abi_check")

MIRRORS = {
    "stzmg_desc": _lib.StzmgDesc,
    "stzmg_solve_opts": _lib.StzmgSolveOpts,
    "stzmg_solve_stats": _lib.StzmgSolveStats,
    "stz_params": _lib.StzParams,
    "stz_step_stats": _lib.StzStepStats,
    "stz_adjoint_opts": _lib.StzAdjointOpts,
    "stz_grf_params": _lib.StzGrfParams,
}


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.fail("tests/abi/abi_check is not built (run __graft_entry__.build())")
    return EXE


def test_struct_layout_matches_ctypes_mirror(exe):
    out = subprocess.run([exe, "layout"], check=True, capture_output=True, text=True).stdout
    layout = json.loads(out)
    assert set(layout) == set(MIRRORS)
    for name, cls in MIRRORS.items():
        c = layout[name]
        assert c.pop("__size__")[1] == ctypes.sizeof(cls), name
        fields = [f[0] for f in cls._fields_]
        # every field the C++ side lays out is mirrored, at the same offset
        # and size, and the mirror holds nothing else
        assert sorted(c) == sorted(fields), name
        for f in fields:
            d = getattr(cls, f)
            assert c[f] == [d.offset, d.size], (name, f, c[f], d.offset, d.size)


@pytest.mark.gpu
def test_cpp_program_solves_and_steps_on_gpu(exe):
    p = subprocess.run([exe, "run"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.count("OK") == 8 and "FAIL" not in p.stdout
