"""Host-side C ABI: the library loads and exports every declared symbol;
the host pack logic is correct and bounds-safe (address sanitizer)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "acehip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(ace_\w+)\s*\(", src, re.M)))


def test_library_exports_all_declared_symbols():
    from acehip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    lib = C.CDLL(_lib.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTS)
    assert lib.ace_abi_version() == _lib.ABI_VERSION


def test_host_pack_logic_under_asan(tmp_path):
    so = tmp_path / "libacehost_asan.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-omit-frame-pointer", "-fPIC", "-shared", "-o", str(so),
                    os.path.join(ROOT, "ace.jl_amd", "csrc", "pack_host_check.cpp")], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host_check_driver.py"), str(so)],
                       env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "host pack checks passed" in r.stdout
    assert "host neighbour-grid checks passed" in r.stdout


def test_cpu_port_matches_oracle():
    import numpy as np

    import ace_oracle as O
    from ace_cpu import CpuModel
    from acehip import synth
    from helpers import basis_for, rel_err
    for args, kw, wid in [((3, 6, 4), {}, 1), ((4, 12, 6), {}, 2),
                          ((3, 10, 5), {"prop": "EuclideanVector", "species": 2}, 4)]:
        basis, T = basis_for(*args, **kw)
        c = synth.coefficients(len(basis), 1)
        off, R, sp = synth.workload_envs(wid, nenv=6)
        E, g = CpuModel(T, c).eval(off, R, species=sp)
        assert rel_err(E, O.energy(T, c, off, R, sp)) < 1e-13
        assert rel_err(g, O.grad_config(T, c, off, R, sp)) < 1e-12
