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
    nc = 1 if pk.prop == 0 else 2 if pk.classed else 6
    leaf = np.zeros(len(pk.leaf_fid) * nc)
    task = np.zeros(len(pk.tasks) * (6 if nc == 6 else 2))
    e0 = np.zeros(6)
    assert lib.acehost_coeffs(C.byref(s), c.ctypes.data_as(f64p), leaf.ctypes.data_as(f64p),
                              task.ctypes.data_as(f64p), e0.ctypes.data_as(f64p)) == 0
    ct = O.ctilde(T, c)
    ref = [PR.eval_terms(t, ct) for t in pk.prog.leaf_terms]
    if nc == 1:
        assert np.max(np.abs(leaf - np.asarray(ref))) < 1e-15
    elif nc == 2:
        # class-split: leaf scalar times its class vector = the three complex
        # coefficients of the componentwise program
        L = np.asarray(ref).reshape(-1, 3)
        ok, u, _ = PR.class_vectors(pk.a2b_cols, pk.a2b_vals.reshape(-1, 6).view(np.complex128), pk.row_class)
        assert ok
        cls = [PR.term_class(t[0], pk.row_class) for t in pk.prog.leaf_terms]
        got = (leaf[0::2] + 1j * leaf[1::2])[:, None] * u[cls]
        assert np.max(np.abs(got - L)) < 1e-15 * max(1.0, np.abs(L).max())
        # task constants: a task's class is its energy slot's or its G slot's
        t5 = pk.tasks.reshape(-1, 5)
        tcls = [-1 - sl if sl < 0 else pk.gslot_class[sl] for sl in t5[:, 2]]
        Tt = np.asarray([PR.eval_terms(t, ct) if t else np.zeros(3, complex) for t in pk.prog.task_terms])
        got = (task[0::2] + 1j * task[1::2])[:, None] * u[tcls]
        assert np.max(np.abs(got - Tt.reshape(-1, 3))) < 1e-15 * max(1.0, np.abs(Tt).max())
    else:
        L = np.asarray(ref).reshape(-1, 3)
        assert np.max(np.abs(leaf - np.stack([L.real, L.imag], -1).reshape(-1))) < 1e-15
    ns = max(1, pk.nspecies)
    kidx = np.zeros(ns * (T["L"] + 1) ** 2 * len(T["rn"]["A"]), dtype=np.int16)
    a4 = np.zeros(4 * len(pk.aspec_fid), dtype=np.int32)
    assert lib.acehost_lookup(C.byref(s), kidx.ctypes.data_as(C.c_void_p), a4.ctypes.data_as(C.c_void_p)) == 0
    assert np.all(a4.reshape(-1, 4) == np.asarray(T["aspec"]))
    # pooling lane sets: every pooled function on exactly one (lane, pass);
    # the functions of a lane share (species, l, m); sets fit 64 lanes or none
    P = np.ascontiguousarray(pk.pool, dtype=np.int32).reshape(-1, 4)
    for kp in (1, 2, 3, 8):
        slot = np.full(64 * kp, -7, dtype=np.int32)
        ok = lib.acehost_pool_sets(P.ctypes.data_as(C.c_void_p), len(P), kp, slot.ctypes.data_as(C.c_void_p))
        cls = {}
        for mu, n, l, m in P.tolist():
            cls[(mu, l, m)] = cls.get((mu, l, m), 0) + 1
        need = sum(-(-v // kp) for v in cls.values())
        assert ok == (need <= 64), (kp, ok, need)
        if ok:
            S = slot.reshape(kp, 64)
            assert sorted(S[S >= 0].tolist()) == list(range(len(P)))
            for g in range(64):
                ks = S[:, g][S[:, g] >= 0]
                assert len({(P[k, 0], P[k, 2], P[k, 3]) for k in ks}) <= 1
                # members fill the first passes of a lane
                assert np.all(S[:len(ks), g] >= 0)
    # corrupted packs are rejected with a message, never read out of bounds
    for field, bad, expect in [("npool", 0, b"npool"), ("L", 11, b"L must"), ("maxorder", 9, b"maxorder"),
                               ("ngslots", 0, b"task out of range")]:
        s2, keep2 = _lib.make_pack(pk)
        setattr(s2, field, bad)
        assert lib.acehost_validate(C.byref(s2), msg, 256) != 0 and expect in msg.value, (field, msg.value)
    bad_runs = pk.runs.copy()
    bad_runs[0, 2] = 10 ** 6
    s3, keep3 = _lib.make_pack(pk)
    keep3.append(np.ascontiguousarray(bad_runs, dtype=np.int32))
    s3.runs = keep3[-1].ctypes.data_as(C.POINTER(C.c_int32))
    assert lib.acehost_validate(C.byref(s3), msg, 256) != 0 and b"run out of range" in msg.value
# the program compiler (compile_host.hpp) under the sanitizers: the tables of
# small bases equal the Python packer's
lib.acehost_compile.restype = C.c_void_p
lib.acehost_compile.argtypes = [C.POINTER(_lib.AceJlBasis), C.c_char_p, C.c_int]
lib.acehost_compiled_array.restype = C.c_int64
lib.acehost_compiled_array.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
lib.acehost_compiled_free.argtypes = [C.c_void_p]
for args in [(3, 6, 4, "Invariant", 0), (3, 6, 3, "EuclideanVector", 2), (5, 5, 2, "Invariant", 0)]:
    basis, T = basis_for(*args)
    pk = Pack(T)
    s, keep = _lib.make_jl_basis(T)
    msg = C.create_string_buffer(256)
    h = lib.acehost_compile(C.byref(s), msg, 256)
    assert h, msg.value
    for name, dt in [("runs", np.int32), ("leaf_fid", np.int32), ("term_scale", np.float64), ("tasks", np.int32),
                     ("rec_hdr", np.uint32), ("wave_task_list", np.int32), ("fgrad_f", np.int32)]:
        n = lib.acehost_compiled_array(h, name.encode(), None)
        a = np.zeros(n, dtype=dt)
        lib.acehost_compiled_array(h, name.encode(), a.ctypes.data_as(C.c_void_p))
        assert np.array_equal(a, np.asarray(getattr(pk, name), dtype=dt).reshape(-1)), (args, name)
    lib.acehost_compiled_free(h)
print("host pack checks passed")
# neighbour-list binning: bins are consistent and cover every pair within rcut
import nbr_oracle as NO  # noqa: E402

i32p = C.POINTER(C.c_int32)


def _grid(x, cell, pbc, rcut):
    n = len(x)
    pos = np.ascontiguousarray(np.asarray(x, dtype=np.float64).T)
    nb = (C.c_int * 3)()
    R = (C.c_int * 3)()
    bin_ = np.zeros(max(n, 1), np.int32)
    bstart = np.zeros(2 * n + 66, np.int32)
    batoms = np.zeros(max(n, 1), np.int32)
    wrap = np.zeros(3 * max(n, 1), np.int32)
    pw = np.zeros(3 * max(n, 1))
    cellc = np.ascontiguousarray(cell, dtype=np.float64)
    pbcc = np.asarray(pbc, dtype=np.int32)
    m = C.create_string_buffer(256)
    rc = lib.acehost_nbr_grid(C.c_int64(n), pos.ctypes.data_as(f64p), cellc.ctypes.data_as(f64p),
                              pbcc.ctypes.data_as(i32p), C.c_double(rcut), nb, R, bin_.ctypes.data_as(i32p),
                              bstart.ctypes.data_as(i32p), batoms.ctypes.data_as(i32p), wrap.ctypes.data_as(i32p),
                              pw.ctypes.data_as(f64p), m, 256)
    return rc, m.value, list(nb), list(R), bin_[:n], bstart, batoms[:n], wrap[:3 * n].reshape(3, n).T, \
        pw[:3 * n].reshape(3, n).T


lib.acehost_nbr_grid.argtypes = [C.c_int64, f64p, f64p, i32p, C.c_double, C.c_void_p, C.c_void_p, i32p, i32p,
                                 i32p, i32p, f64p, C.c_char_p, C.c_int]
tri = np.array([[4.0, 0.0, 0.0], [1.3, 3.7, 0.0], [0.6, -0.9, 4.4]])
cases = [(tri, [1, 1, 1], 200, 3.1, True), (tri, [1, 1, 0], 150, 2.5, True), (20 * np.eye(3), [0, 0, 0], 60, 4.0, False),
         (NO.fcc_cell(3.6), [1, 1, 1], 1, 5.0, True), (np.diag([3.0, 3.0, 30.0]), [1, 1, 0], 5, 7.5, True),
         (tri * 3, [1, 1, 1], 500, 2.2, False)]
for cell, pbc, n, rcut, outside in cases:
    x = NO.random_system(n, cell, pbc, seed=n, outside=outside)
    rc, m, nb, R, bin_, bstart, batoms, wrap, pw = _grid(x, cell, pbc, rcut)
    assert rc == 0, m
    nbins = nb[0] * nb[1] * nb[2]
    assert nbins <= 2 * n + 64
    assert bstart[0] == 0 and bstart[nbins] == n and np.all(np.diff(bstart[:nbins + 1]) >= 0)
    assert sorted(batoms.tolist()) == list(range(n))
    for b in range(nbins):
        assert np.all(bin_[batoms[bstart[b]:bstart[b + 1]]] == b)
    assert np.allclose(pw, x - wrap @ cell, atol=1e-12)
    sw = pw @ np.linalg.inv(cell)
    per = np.asarray(pbc, bool)
    assert np.all(sw[:, per] >= -1e-12) and np.all(sw[:, per] < 1 + 1e-12)
    i, j, S, r = NO.neighbour_pairs(x, cell, pbc, rcut)
    Sw = S + wrap[j] - wrap[i]                       # shift between wrapped positions
    nbv = np.asarray(nb)
    bc = np.stack([bin_ % nb[0], (bin_ // nb[0]) % nb[1], bin_ // (nb[0] * nb[1])], -1)
    reach = bc[j] + Sw * nbv - bc[i]
    assert np.all(np.abs(reach) <= np.asarray(R)), (cell, pbc, nb, R)
    assert np.all(Sw[:, ~per] == 0)
for cell, pbc, rcut, x, expect in [(np.zeros((3, 3)), [1, 1, 1], 1.0, np.zeros((2, 3)), b"singular"),
                                   (np.eye(3), [1, 1, 1], -1.0, np.zeros((2, 3)), b"rcut"),
                                   (np.eye(3), [1, 2, 1], 1.0, np.zeros((2, 3)), b"pbc"),
                                   (np.eye(3), [1, 1, 1], 1.0, np.array([[0.0, np.nan, 0.0]]), b"finite"),
                                   (0.01 * np.eye(3), [1, 1, 1], 5.0, np.zeros((1, 3)), b"too small")]:
    rc, m, *_ = _grid(x, cell, pbc, rcut)
    assert rc != 0 and expect in m, (expect, m)
print("host neighbour-grid checks passed")

