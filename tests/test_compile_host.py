"""The C++ program compiler (csrc/compile_host.hpp) against the Python host's
packer (acehip/pack.py): from the same basis, one in ACE.jl's own layout
(ace_jl_basis: Product1pBasis.indices, PIBasisSpec, A2Bmap as CSC) and one
from the basis tables, every device table must be identical, entry for entry.
That is what makes a context created from ACE.jl's tables
(ace_ctx_create_jl) evaluate bit-identically to one created from the Python
pack.  The shim is the CPU build of the compiler (pack_host_check.cpp)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from acehip import _lib
from acehip.pack import Pack
from helpers import basis_for

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# ace_pack field -> (Pack attribute, dtype)
FIELDS = {
    "pool": ("pool", np.int32), "aspec_fid": ("aspec_fid", np.int32), "aspec_sgn": ("aspec_sgn", np.float64),
    "orders": ("orders", np.int32), "iAA2iA": ("iAA2iA", np.int32), "a2b_rowptr": ("a2b_rowptr", np.int32),
    "a2b_cols": ("a2b_cols", np.int32), "a2b_vals": ("a2b_vals", np.float64), "runs": ("runs", np.int32),
    "leaf_fid": ("leaf_fid", np.int32), "leaf_term_ptr": ("leaf_term_ptr", np.int32),
    "tasks": ("tasks", np.int32), "task_term_ptr": ("task_term_ptr", np.int32), "term_row": ("term_row", np.int32),
    "term_scale": ("term_scale", np.float64), "term_conj": ("term_conj", np.int32),
    "gslot_ptr": ("gslot_ptr", np.int32), "gslot_list": ("gslot_list", np.int32),
    "gslot_class": ("gslot_class", np.int32),
    "rec_hdr": ("rec_hdr", np.uint32), "rec_leaf": ("rec_leaf", np.int32), "feat_rows": ("feat_rows", np.int32),
    "feat_ptr": ("feat_ptr", np.int32), "feat_f": ("feat_f", np.int32), "feat_coef": ("feat_coef", np.float64),
    "fgrad_ptr": ("fgrad_ptr", np.int32), "fgrad_feat": ("fgrad_feat", np.int32), "fgrad_f": ("fgrad_f", np.int32),
    "fgrad_coef": ("fgrad_coef", np.float64), "wave_cfg_W": ("wave_cfg_W", np.int32),
    "wave_task_ptr": ("wave_task_ptr", np.int32), "wave_task_list": ("wave_task_list", np.int32),
    "wave_etask_ptr": ("wave_etask_ptr", np.int32), "wave_etask_list": ("wave_etask_list", np.int32),
}


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    so = tmp_path_factory.mktemp("shim") / "libacehost.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-o", str(so),
                    os.path.join(ROOT, "ace.jl_amd", "csrc", "pack_host_check.cpp")], check=True)
    lib = C.CDLL(str(so))
    lib.acehost_compile.restype = C.c_void_p
    lib.acehost_compile.argtypes = [C.POINTER(_lib.AceJlBasis), C.c_char_p, C.c_int]
    lib.acehost_compiled_array.restype = C.c_int64
    lib.acehost_compiled_array.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
    lib.acehost_compiled_free.argtypes = [C.c_void_p]
    return lib


def compiled(lib, T):
    s, keep = _lib.make_jl_basis(T)
    msg = C.create_string_buffer(256)
    h = lib.acehost_compile(C.byref(s), msg, 256)
    assert h, msg.value
    out = {}
    for name, (_, dt) in FIELDS.items():
        n = lib.acehost_compiled_array(h, name.encode(), None)
        assert n >= 0, name
        a = np.zeros(n, dtype=dt)
        lib.acehost_compiled_array(h, name.encode(), a.ctypes.data_as(C.c_void_p))
        out[name] = a
    lib.acehost_compiled_free(h)
    return out


CASES = [
    ((3, 6, 4), {}),                                        # config 1
    ((3, 6, 4), {"prop": "EuclideanVector"}),               # vector, degree 6
    ((3, 6, 3), {"species": 2}),                            # two species
    ((2, 4, 2), {"species": 2}),                            # 32-or-fewer pooled values
    ((5, 6, 3), {}),                                        # correlation order 5 (no records)
    ((4, 12, 6), {}),                                       # config 2
    ((3, 10, 5), {"prop": "EuclideanVector", "species": 2}),  # config 4
]


@pytest.mark.parametrize("args,kw", CASES)
def test_compiled_tables_equal_python_pack(shim, args, kw):
    basis, T = basis_for(*args, **kw)
    pk = Pack(T)
    got = compiled(shim, T)
    for name, (attr, dt) in FIELDS.items():
        ref = np.asarray(getattr(pk, attr), dtype=dt).reshape(-1)
        assert got[name].shape == ref.shape, (name, got[name].shape, ref.shape)
        assert np.array_equal(got[name], ref), name


def test_order5_has_no_records_and_deep_runs(shim):
    _, T = basis_for(5, 6, 3)
    got = compiled(shim, T)
    assert got["rec_hdr"].size == 0
    runs = got["runs"].reshape(-1, _lib.RUN_W if hasattr(_lib, "RUN_W") else 11)
    assert runs[:, 0].max() == 4              # order-5 monomials: prefixes of four factors


def test_bad_tables_are_rejected(shim):
    _, T = basis_for(3, 6, 4)
    msg = C.create_string_buffer(256)
    for field, bad, expect in [("maxorder", 9, b"correlation order"), ("nr", 0, b"radial"),
                               ("ncomp", 1, b"Product1pBasis"), ("maxL", 11, b"maxL")]:
        s, keep = _lib.make_jl_basis(T)
        setattr(s, field, bad)
        assert not shim.acehost_compile(C.byref(s), msg, 256)
        assert expect in msg.value, (field, msg.value)
    J = _lib.jl_arrays(T)
    for key, idx, val, expect in [("indices", (0, 0), 99, b"Rn index"), ("iAA2iA", (0, 5), 10 ** 6, b"iAA2iA"),
                                  ("rowval", (0,), 10 ** 6, b"A2B row"), ("colptr", (0,), 0, b"colptr")]:
        s, keep = _lib.make_jl_basis(T)
        a = np.ascontiguousarray(J[key].copy())
        a[idx] = val
        keep.append(a)
        setattr(s, key, a.ctypes.data_as(C.POINTER(C.c_int64)))
        assert not shim.acehost_compile(C.byref(s), msg, 256)
        assert expect in msg.value, (key, msg.value)
