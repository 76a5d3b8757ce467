"""Driver run in a subprocess with the address sanitizer preloaded: checks
the host-side pack logic (pack_host.hpp via pack_host_check.cpp)."""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(HERE, "..", "ace.jl_amd"), os.path.join(HERE, "..", "oracle"), HERE):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import ace_oracle as O  # noqa: E402
from acehip import _lib, programs as PR, synth  # noqa: E402
from acehip.pack import Pack  # noqa: E402
from helpers import basis_for  # noqa: E402

lib = C.CDLL(sys.argv[1])
f64p = C.POINTER(C.c_double)
fails = 0
for args in [(3, 6, 4, "Invariant", 0), (3, 6, 4, "EuclideanVector", 0), (3, 6, 3, "Invariant", 2)]:
    basis, T = basis_for(*args)
    pk = Pack(T)
    s, keep = _lib.make_pack(pk)
    msg = C.create_string_buffer(256)
    assert lib.acehost_validate(C.byref(s), msg, 256) == 0, msg.value
    c = synth.coefficients(len(basis), 3)
    nc = 1 if pk.prop == 0 else 6
    leaf = np.zeros(len(pk.leaf_fid) * nc)
    task = np.zeros(len(pk.tasks) * (2 if nc == 1 else 6))
    e0 = np.zeros(6)
    assert lib.acehost_coeffs(C.byref(s), c.ctypes.data_as(f64p), leaf.ctypes.data_as(f64p),
                              task.ctypes.data_as(f64p), e0.ctypes.data_as(f64p)) == 0
    ct = O.ctilde(T, c)
    ref = [PR.eval_terms(t, ct) for t in pk.prog.leaf_terms]
    if nc == 1:
        assert np.max(np.abs(leaf - np.asarray(ref))) < 1e-15
    else:
        L = np.asarray(ref).reshape(-1, 3)
        assert np.max(np.abs(leaf - np.stack([L.real, L.imag], -1).reshape(-1))) < 1e-15
    ns = max(1, pk.nspecies)
    kidx = np.zeros(ns * (T["L"] + 1) ** 2 * len(T["rn"]["A"]), dtype=np.int16)
    a4 = np.zeros(4 * len(pk.aspec_fid), dtype=np.int32)
    assert lib.acehost_lookup(C.byref(s), kidx.ctypes.data_as(C.c_void_p), a4.ctypes.data_as(C.c_void_p)) == 0
    assert np.all(a4.reshape(-1, 4) == np.asarray(T["aspec"]))
    # corrupted packs are rejected with a message, never read out of bounds
    for field, bad, expect in [("npool", 0, b"npool"), ("L", 9, b"L must"), ("maxorder", 5, b"maxorder"),
                               ("ngslots", 0, b"task out of range")]:
        s2, keep2 = _lib.make_pack(pk)
        setattr(s2, field, bad)
        assert lib.acehost_validate(C.byref(s2), msg, 256) != 0 and expect in msg.value, (field, msg.value)
    bad_runs = pk.runs.copy()
    bad_runs[0, 5] = 10 ** 6
    s3, keep3 = _lib.make_pack(pk)
    keep3.append(np.ascontiguousarray(bad_runs, dtype=np.int32))
    s3.runs = keep3[-1].ctypes.data_as(C.POINTER(C.c_int32))
    assert lib.acehost_validate(C.byref(s3), msg, 256) != 0 and b"run out of range" in msg.value
print("host pack checks passed")
