"""ctypes binding of libacehip.so (the C ABI declared in include/acehip.h).

The library is loaded from this package directory.  There is no fallback:
if the library is missing or fails to load, every evaluation raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libacehip.so")

ABI_VERSION = 3

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)

COMP_CATEGORICAL, COMP_RN, COMP_YLM = 0, 1, 2  # ACE_COMP_* of include/acehip.h


class AceJlBasis(C.Structure):
    """``ace_jl_basis``: a basis in ACE.jl's own memory layout."""
    _fields_ = [
        ("abi_version", C.c_int32), ("prop", C.c_int32),
        ("nr", C.c_int64), ("A", _f64p), ("B", _f64p), ("C", _f64p),
        ("pl", C.c_int64), ("pr", C.c_int64), ("tl", C.c_double), ("tr", C.c_double),
        ("tp", C.c_int64), ("r0", C.c_double), ("maxL", C.c_int64), ("ncategories", C.c_int64),
        ("ncomp", C.c_int64), ("comp_kind", _i32p), ("nA", C.c_int64), ("indices", _i64p),
        ("nAA", C.c_int64), ("maxorder", C.c_int64), ("orders", _i64p), ("iAA2iA", _i64p),
        ("nB", C.c_int64), ("colptr", _i64p), ("rowval", _i64p), ("nzval", _f64p),
    ]


class AcePack(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("prop", C.c_int32), ("L", C.c_int32), ("nr", C.c_int32),
        ("nspecies", C.c_int32), ("tp", C.c_int32),
        ("r0", C.c_double), ("tl", C.c_double), ("tr", C.c_double),
        ("pl", C.c_int32), ("pr", C.c_int32),
        ("rA", _f64p), ("rB", _f64p), ("rC", _f64p),
        ("npool", C.c_int32), ("pool", _i32p),
        ("nA", C.c_int32), ("aspec_fid", _i32p), ("aspec_sgn", _f64p),
        ("nAA", C.c_int32), ("maxorder", C.c_int32), ("orders", _i32p), ("iAA2iA", _i32p),
        ("nB", C.c_int32), ("a2b_nnz", C.c_int32), ("a2b_rowptr", _i32p), ("a2b_cols", _i32p),
        ("a2b_vals", _f64p),
        ("nruns", C.c_int32), ("runs", _i32p),
        ("nleaves", C.c_int32), ("leaf_fid", _i32p), ("leaf_term_ptr", _i32p),
        ("ntasks", C.c_int32), ("tasks", _i32p), ("task_term_ptr", _i32p),
        ("nterms", C.c_int32), ("term_row", _i32p), ("term_scale", _f64p), ("term_conj", _i32p),
        ("ngslots", C.c_int32), ("gslot_ptr", _i32p), ("gslot_list", _i32p),
        ("classed", C.c_int32), ("gslot_class", _i32p),
        ("nrec", C.c_int32), ("rec_hdr", C.POINTER(C.c_uint32)), ("rec_leaf", _i32p),
        ("nfeat", C.c_int32), ("feat_rows", _i32p), ("feat_ptr", _i32p), ("feat_f", _i32p),
        ("feat_coef", _f64p), ("fgrad_ptr", _i32p), ("fgrad_feat", _i32p), ("fgrad_f", _i32p),
        ("fgrad_coef", _f64p),
        ("nwave_cfg", C.c_int32), ("wave_cfg_W", _i32p), ("wave_task_ptr", _i32p),
        ("wave_task_list", _i32p), ("wave_etask_ptr", _i32p), ("wave_etask_list", _i32p),
    ]


# every symbol declared in include/acehip.h
EXPORTS = ("ace_ctx_create", "ace_ctx_create_jl", "ace_compile_check", "ace_last_dag_shape", "ace_set_kernel", "ace_set_pool_shape", "ace_set_params", "ace_set_tile", "ace_set_parts", "ace_set_tail_parts", "ace_energy",
           "ace_energy_forces", "ace_basis_A", "ace_basis_AA", "ace_basis_B",
           "ace_basis_dB", "ace_surr_features", "ace_surr_energy_forces",
           "ace_last_timing", "ace_set_validation", "ace_last_error",
           "ace_ctx_destroy", "ace_abi_version", "ace_nbrlist_create", "ace_nbrlist_destroy",
           "ace_nbrlist_build", "ace_nbrlist_update", "ace_nbrlist_forces", "ace_nbrlist_forces_multi",
           "ace_adjoint_EVAL_D", "ace_energy_forces_noalloc", "ace_energy_host", "ace_energy_forces_host",
           "ace_basis_B_host", "ace_basis_dB_host", "ace_adjoint_EVAL_D_host")

ERR_CAPACITY = -2  # ACE_ERR_CAPACITY

_lib = None


class AceError(RuntimeError):
    pass


def load(path: str | None = None):
    """Load libacehip.so (raises if it is missing: no CPU fallback exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise AceError(f"libacehip.so not built ({p}); run __graft_entry__.build()")
    lib = C.CDLL(p)
    vp = C.c_void_p
    lib.ace_ctx_create.argtypes = [C.POINTER(AcePack), C.c_int, _f64p, C.c_int64, C.POINTER(vp)]
    lib.ace_ctx_create_jl.argtypes = [C.POINTER(AceJlBasis), C.c_int, _f64p, C.c_int64, C.POINTER(vp)]
    lib.ace_compile_check.argtypes = [C.POINTER(AceJlBasis), _i64p]
    lib.ace_last_dag_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int)]
    lib.ace_set_params.argtypes = [vp, _f64p, C.c_int64]
    lib.ace_set_tile.argtypes = [vp, C.c_int]
    lib.ace_set_parts.argtypes = [vp, C.c_int]
    lib.ace_set_tail_parts.argtypes = [vp, C.c_int]
    ev = [vp, C.c_int64, vp, vp, vp, vp, vp]
    lib.ace_energy.argtypes = ev
    lib.ace_energy_forces.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, vp]
    lib.ace_energy_forces_noalloc.argtypes = [vp, C.c_int64, vp, C.c_int64, vp, vp, vp, vp, vp]
    for name in ("ace_basis_A", "ace_basis_AA", "ace_basis_B", "ace_basis_dB"):
        getattr(lib, name).argtypes = ev
    lib.ace_surr_features.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.ace_surr_energy_forces.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_int, vp, vp, C.c_double,
                                           C.c_double, vp, C.c_double, vp, vp, vp, vp, vp, vp]
    lib.ace_last_timing.argtypes = [vp, _f64p, _f64p, _f64p]
    lib.ace_set_validation.argtypes = [vp, C.c_int]
    lib.ace_set_kernel.argtypes = [vp, C.c_int]
    lib.ace_set_pool_shape.argtypes = [vp, C.c_int, C.c_int]
    lib.ace_last_error.restype = C.c_char_p
    lib.ace_last_error.argtypes = []
    lib.ace_ctx_destroy.argtypes = [vp]
    lib.ace_ctx_destroy.restype = None
    lib.ace_abi_version.restype = C.c_int
    lib.ace_nbrlist_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.ace_nbrlist_destroy.argtypes = [vp]
    lib.ace_nbrlist_destroy.restype = None
    lib.ace_nbrlist_build.argtypes = [vp, C.c_int64, vp, _f64p, _i32p, C.c_double, vp, C.c_int64, vp, vp, vp, vp,
                                      vp, vp, C.POINTER(C.c_int64), vp]
    lib.ace_nbrlist_update.argtypes = [vp, C.c_int64, vp, _f64p, C.c_int64, vp, vp, vp, vp, vp]
    lib.ace_nbrlist_forces.argtypes = [vp, C.c_int64, vp, vp, C.c_int64, vp, vp, vp, vp, vp]
    lib.ace_nbrlist_forces_multi.argtypes = [vp, C.c_int64, vp, vp, C.c_int64, C.c_int64, vp, vp, vp]
    lib.ace_adjoint_EVAL_D.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, vp]
    lib.ace_energy_host.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    lib.ace_energy_forces_host.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp]
    lib.ace_basis_B_host.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    lib.ace_basis_dB_host.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    lib.ace_adjoint_EVAL_D_host.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp]
    if lib.ace_abi_version() != ABI_VERSION:
        raise AceError("libacehip.so ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


def check(rc: int):
    if rc != 0:
        msg = load().ace_last_error().decode()
        raise AceError(msg)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def make_pack(pk):
    """Fill an AcePack from a :class:`acehip.pack.Pack`; returns (struct,
    keepalive list)."""
    T = pk.T
    rn = T["rn"]
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(x, dtype=dt)
        keep.append(a)
        return a

    s = AcePack()
    s.abi_version = ABI_VERSION
    s.prop = pk.prop
    s.L = int(T["L"])
    s.nr = len(rn["A"])
    s.nspecies = pk.nspecies
    s.tp = int(rn["p"])
    s.r0, s.tl, s.tr = float(rn["r0"]), float(rn["tl"]), float(rn["tr"])
    s.pl, s.pr = int(rn["pl"]), int(rn["pr"])
    s.rA = _ptr(arr(rn["A"], np.float64), C.c_double)
    s.rB = _ptr(arr(rn["B"], np.float64), C.c_double)
    s.rC = _ptr(arr(rn["C"], np.float64), C.c_double)
    s.npool = len(pk.pool)
    s.pool = _ptr(arr(pk.pool, np.int32), C.c_int32)
    s.nA = len(pk.aspec_fid)
    s.aspec_fid = _ptr(arr(pk.aspec_fid, np.int32), C.c_int32)
    s.aspec_sgn = _ptr(arr(pk.aspec_sgn, np.float64), C.c_double)
    s.nAA = len(pk.orders)
    s.maxorder = pk.maxorder
    s.orders = _ptr(arr(pk.orders, np.int32), C.c_int32)
    s.iAA2iA = _ptr(arr(pk.iAA2iA, np.int32), C.c_int32)
    s.nB = pk.nB
    s.a2b_nnz = len(pk.a2b_cols)
    s.a2b_rowptr = _ptr(arr(pk.a2b_rowptr, np.int32), C.c_int32)
    s.a2b_cols = _ptr(arr(pk.a2b_cols, np.int32), C.c_int32)
    s.a2b_vals = _ptr(arr(pk.a2b_vals, np.float64), C.c_double)
    s.nruns = len(pk.runs)
    s.runs = _ptr(arr(pk.runs, np.int32), C.c_int32)
    s.nleaves = len(pk.leaf_fid)
    s.leaf_fid = _ptr(arr(pk.leaf_fid, np.int32), C.c_int32)
    s.leaf_term_ptr = _ptr(arr(pk.leaf_term_ptr, np.int32), C.c_int32)
    s.ntasks = len(pk.tasks)
    s.tasks = _ptr(arr(pk.tasks, np.int32), C.c_int32)
    s.task_term_ptr = _ptr(arr(pk.task_term_ptr, np.int32), C.c_int32)
    s.nterms = len(pk.term_row)
    s.term_row = _ptr(arr(pk.term_row, np.int32), C.c_int32)
    s.term_scale = _ptr(arr(pk.term_scale, np.float64), C.c_double)
    s.term_conj = _ptr(arr(pk.term_conj, np.int32), C.c_int32)
    s.ngslots = pk.ngslots
    s.gslot_ptr = _ptr(arr(pk.gslot_ptr, np.int32), C.c_int32)
    s.gslot_list = _ptr(arr(pk.gslot_list, np.int32), C.c_int32)
    s.classed = int(pk.classed)
    s.gslot_class = _ptr(arr(pk.gslot_class, np.int32), C.c_int32)
    s.nrec = len(pk.rec_hdr)
    s.rec_hdr = _ptr(arr(pk.rec_hdr, np.uint32), C.c_uint32)
    s.rec_leaf = _ptr(arr(pk.rec_leaf, np.int32), C.c_int32)
    s.nfeat = len(pk.feat_rows)
    s.feat_rows = _ptr(arr(pk.feat_rows, np.int32), C.c_int32)
    s.feat_ptr = _ptr(arr(pk.feat_ptr, np.int32), C.c_int32)
    s.feat_f = _ptr(arr(pk.feat_f, np.int32), C.c_int32)
    s.feat_coef = _ptr(arr(pk.feat_coef, np.float64), C.c_double)
    s.fgrad_ptr = _ptr(arr(pk.fgrad_ptr, np.int32), C.c_int32)
    s.fgrad_feat = _ptr(arr(pk.fgrad_feat, np.int32), C.c_int32)
    s.fgrad_f = _ptr(arr(pk.fgrad_f, np.int32), C.c_int32)
    s.fgrad_coef = _ptr(arr(pk.fgrad_coef, np.float64), C.c_double)
    s.nwave_cfg = len(pk.wave_cfg_W)
    s.wave_cfg_W = _ptr(arr(pk.wave_cfg_W, np.int32), C.c_int32)
    s.wave_task_ptr = _ptr(arr(pk.wave_task_ptr, np.int32), C.c_int32)
    s.wave_task_list = _ptr(arr(pk.wave_task_list, np.int32), C.c_int32)
    s.wave_etask_ptr = _ptr(arr(pk.wave_etask_ptr, np.int32), C.c_int32)
    s.wave_etask_list = _ptr(arr(pk.wave_etask_list, np.int32), C.c_int32)
    return s, keep


def jl_arrays(T: dict) -> dict:
    """The basis tables ``T`` (acehip.tables.basis_tables) in the layout ACE.jl
    keeps them in (what a Julia caller passes to ``ace_ctx_create_jl``):
    ``indices[nA][ncomp]`` int64 of Product1pBasis (category index, n,
    index_y(l, m) = m + l + l^2 + 1; 1-based), ``iAA2iA`` column-major and
    1-based with 0 padding, the A2B map as SparseMatrixCSC (colptr, rowval
    1-based; nzval real, or 3 complex values re/im interleaved)."""
    aspec = np.asarray(T["aspec"], dtype=np.int64).reshape(-1, 4)
    cat = T["categories"] is not None
    kinds = ([COMP_CATEGORICAL] if cat else []) + [COMP_RN, COMP_YLM]
    cols = ([aspec[:, 0] + 1] if cat else []) + [aspec[:, 1], aspec[:, 3] + aspec[:, 2] + aspec[:, 2] ** 2 + 1]
    indices = np.stack(cols, axis=1).astype(np.int64)
    orders = np.asarray(T["orders"], dtype=np.int64)
    iAA2iA = np.asarray(T["iAA2iA"], dtype=np.int64) + 1          # -1 padding -> 0
    a2b = T["a2b"]
    nB, nAA = (int(x) for x in a2b["shape"])
    rows = np.asarray(a2b["rows"], dtype=np.int64)
    ccol = np.asarray(a2b["cols"], dtype=np.int64)
    order = np.lexsort((rows, ccol))                              # column-major, rows ascending
    rows, ccol = rows[order], ccol[order]
    vals = np.asarray(a2b["vals"])[order]
    colptr = np.zeros(nAA + 1, dtype=np.int64)
    np.add.at(colptr, ccol + 1, 1)
    colptr = np.cumsum(colptr) + 1
    if T["prop"] == "Invariant":
        nz = np.asarray(vals, dtype=np.float64)
    else:
        v = np.asarray(vals, dtype=np.complex128).reshape(-1, 3)
        nz = np.stack([v.real, v.imag], axis=-1).reshape(-1)
    rn = T["rn"]
    return {"prop": 0 if T["prop"] == "Invariant" else 1, "A": np.asarray(rn["A"], dtype=np.float64),
            "B": np.asarray(rn["B"], dtype=np.float64), "C": np.asarray(rn["C"], dtype=np.float64),
            "pl": int(rn["pl"]), "pr": int(rn["pr"]), "tl": float(rn["tl"]), "tr": float(rn["tr"]),
            "tp": int(rn["p"]), "r0": float(rn["r0"]), "maxL": int(T["L"]),
            "ncategories": len(T["categories"]) if cat else 0, "comp_kind": np.asarray(kinds, dtype=np.int32),
            "indices": indices, "orders": orders, "iAA2iA": np.asfortranarray(iAA2iA).T.copy(), "nB": nB,
            "colptr": colptr, "rowval": rows + 1, "nzval": nz}


def make_jl_basis(T: dict):
    """Fill an AceJlBasis from the basis tables; returns (struct, keepalive)."""
    J = jl_arrays(T)
    keep = []

    def arr(x, dt, ct):
        a = np.ascontiguousarray(x, dtype=dt)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(ct))

    s = AceJlBasis()
    s.abi_version = ABI_VERSION
    s.prop = J["prop"]
    s.nr = len(J["A"])
    s.A, s.B, s.C = (arr(J[k], np.float64, C.c_double) for k in ("A", "B", "C"))
    s.pl, s.pr, s.tl, s.tr, s.tp, s.r0 = J["pl"], J["pr"], J["tl"], J["tr"], J["tp"], J["r0"]
    s.maxL, s.ncategories = J["maxL"], J["ncategories"]
    s.ncomp = len(J["comp_kind"])
    s.comp_kind = arr(J["comp_kind"], np.int32, C.c_int32)
    s.nA = J["indices"].shape[0]
    s.indices = arr(J["indices"], np.int64, C.c_int64)
    s.nAA = len(J["orders"])
    s.maxorder = J["iAA2iA"].shape[0]
    s.orders = arr(J["orders"], np.int64, C.c_int64)
    s.iAA2iA = arr(J["iAA2iA"], np.int64, C.c_int64)        # [maxorder][nAA]: Julia's column-major matrix
    s.nB = J["nB"]
    s.colptr = arr(J["colptr"], np.int64, C.c_int64)
    s.rowval = arr(J["rowval"], np.int64, C.c_int64)
    s.nzval = arr(J["nzval"], np.float64, C.c_double)
    return s, keep
