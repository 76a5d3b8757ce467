"""ACE.jl model files: ``write_dict`` / ``read_dict`` of a LinearACEModel and
the basis types on the evaluation path, as JSON.

The dictionary layouts (``__id__`` strings and keys) are the reference's:
  * LinearACEModel ``src/linearmodel.jl:80-93``; ProductEvaluator / Naive
    ``src/evaluator.jl:20-22``, ``src/linearmodel.jl:95-99``
  * SymmetricBasis ``src/symmbasis.jl:50-62``; O3 ``src/symmetrygroups.jl:82-88``
  * PIBasis / PIBasisSpec ``src/pibasis.jl:235-252``
  * Product1pBasis ``src/product_1pbasis.jl:57-66`` (``indices``: per 1p
    function, 1-based indices into each component's spec)
  * B1pComponent ``src/b1pcomponent.jl:159-177``; SChain ``src/chain.jl:75-83``;
    Lambda ``src/transforms/lambdas.jl:46-51``; StaticGet
    ``src/transforms/statetransforms.jl:85-88``; polytransform string
    ``src/transforms/distancetransforms.jl:16``
  * OrthPolyBasis ``src/polynomials/orthpolys.jl:119-140``; SHBasis
    ``src/polynomials/sphericalharmonics.jl:318-324``
  * Invariant / EuclideanVector ``src/properties.jl:106-112,207-213``
  * SList / Categorical1pBasis ``src/discrete1pbasis.jl:42-52,134-144``
  * Rn component's rin / rcut from tl / tr through the inverse transform
    (``src/b1pcomponents/Rn.jl:17-29``; not stored in the file).

The encodings of Julia types, dense arrays and sparse matrices belong to
ACEbase.FIO (ACEbase 0.2.4 in ACE.jl's ``[compat]``), which is not part of the
reference and not available here.  The reader therefore recognises them by
structure: a type is ``{"T": "<name>"}``; a dense array is a plain (nested)
JSON list (a matrix as a list of columns, as JSON.jl writes Julia matrices) or
a dict with ``size`` and column-major data (``data`` / ``vals``, or real and
imaginary parts as ``real``/``imag``, ``re``/``im`` or ``realdata``/
``imagdata``); a sparse matrix is a dict with the SparseMatrixCSC fields
``m, n, colptr, rowval, nzval`` (1-based) or COO ``I, J, V``.  The writer emits
one of these forms (types ``{"__id__": "Type", "T": name}``, dense arrays as
``{"__id__": "ACE_Array", "T", "size", "data" | "real" + "imag"}``, sparse as
CSC).  Whether ACEbase's own files use exactly these spellings is not
checkable here.
"""
from __future__ import annotations

import json
import re

import numpy as np

from .coupling import EuclideanVector, Invariant
from .onep import Categorical1pBasis, Product1pBasis, Rn1pBasis, Ylm1pBasis
from .orthpolys import OrthPolyBasis, PolyTransform
from .pibasis import PIBasis, PIBasisSpec
from .symmbasis import A2BMap, LinearACEModel, SymmetricBasis

# ---------------------------------------------------------------------------
# primitive encodings (ACEbase.FIO level)


def _type_dict(name):
    return {"__id__": "Type", "T": name}


def _array_dict(a):
    a = np.asarray(a)
    flat = a.reshape(-1, order="F")
    if np.iscomplexobj(a):
        return {"__id__": "ACE_Array", "T": _type_dict("ComplexF64"), "size": list(a.shape),
                "real": flat.real.tolist(), "imag": flat.imag.tolist()}
    T = "Int64" if np.issubdtype(a.dtype, np.integer) else "Float64"
    return {"__id__": "ACE_Array", "T": _type_dict(T), "size": list(a.shape), "data": flat.tolist()}


def _read_array(x, dtype=None):
    """Dense array from a plain list (nested lists = columns) or an array dict."""
    if isinstance(x, dict):
        size = tuple(int(s) for s in x.get("size", ()))
        if "data" in x or "vals" in x:
            flat = np.asarray(x["data"] if "data" in x else x["vals"])
        else:
            for re_k, im_k in (("real", "imag"), ("re", "im"), ("realdata", "imagdata")):
                if re_k in x:
                    flat = np.asarray(x[re_k], dtype=np.float64) + 1j * np.asarray(x.get(im_k, 0.0))
                    break
            else:
                raise ValueError(f"unrecognised array encoding: keys {sorted(x)}")
        if not size:
            size = flat.shape
        a = flat.reshape(size, order="F")
    else:
        a = np.asarray(x)
        if a.ndim == 2:       # JSON.jl writes a matrix as its list of columns
            a = a.T
    return a.astype(dtype) if dtype is not None else a


def _sparse_dict(nrows, ncols, rows, cols, nzvals):
    """CSC (1-based) of COO input; ``nzvals`` already encoded."""
    order = np.lexsort((rows, cols))
    rows, cols = np.asarray(rows)[order], np.asarray(cols)[order]
    nz = [nzvals[i] for i in order]
    colptr = np.zeros(ncols + 1, dtype=np.int64)
    np.add.at(colptr, cols + 1, 1)
    colptr = np.cumsum(colptr) + 1
    return {"__id__": "SparseMatrixCSC", "m": int(nrows), "n": int(ncols), "colptr": colptr.tolist(),
            "rowval": (rows + 1).tolist(), "nzval": nz}


def _read_sparse(D):
    """(nrows, ncols, rows0, cols0, nzval list) from CSC or COO dicts."""
    if "colptr" in D:
        colptr = np.asarray(_read_array(D["colptr"]), dtype=np.int64)
        rows = np.asarray(_read_array(D["rowval"]), dtype=np.int64) - 1
        nz = D["nzval"]
        if isinstance(nz, dict) and "__id__" in nz and nz["__id__"] not in ("ACE_Invariant", "ACE_EuclideanVector"):
            nz = list(_read_array(nz))
        ncols = len(colptr) - 1
        cols = np.repeat(np.arange(ncols), np.diff(colptr))
        nrows = int(D.get("m", rows.max() + 1 if len(rows) else 0))
        ncols = int(D.get("n", ncols))
        return nrows, ncols, rows, cols, list(nz)
    if "I" in D:
        rows = np.asarray(D["I"], dtype=np.int64) - 1
        cols = np.asarray(D["J"], dtype=np.int64) - 1
        return int(D["m"]), int(D["n"]), rows, cols, list(D["V"])
    raise ValueError(f"unrecognised sparse matrix encoding: keys {sorted(D)}")


def _read_property_value(v):
    """An A2Bmap entry: Invariant -> float, EuclideanVector -> complex (3,)."""
    if isinstance(v, dict):
        if v.get("__id__") == "ACE_Invariant":
            return float(v["val"])
        if v.get("__id__") == "ACE_EuclideanVector":
            return np.asarray(_read_array(v["val"]), dtype=np.complex128).reshape(3)
        raise ValueError(f"unsupported A2Bmap entry {v.get('__id__')}")
    return float(v)


# ---------------------------------------------------------------------------
# writer


def _component_dict(B):
    if isinstance(B, Rn1pBasis):
        basis = {"__id__": "ACE_SChain",
                 "F": [{"__id__": "ACE_Lambda", "exstr": "rr -> norm(rr)"}, B.trans.to_dict(), B.J.to_dict()]}
        syms, spec, degrees = [B.nsym], [{B.nsym: n} for n in range(1, len(B) + 1)], list(range(len(B)))
    elif isinstance(B, Ylm1pBasis):
        basis = {"__id__": "ACE_SHBasis", "T": _type_dict("Float64"), "maxL": B.L}
        spec = [{B.lsym: l, B.msym: m} for l in range(B.L + 1) for m in range(-l, l + 1)]
        syms, degrees = [B.lsym, B.msym], [b[B.lsym] for b in spec]
    elif isinstance(B, Categorical1pBasis):
        T = "Int64" if all(isinstance(c, (int, np.integer)) for c in B.categories) else "Symbol"
        return {"__id__": "ACE_Categorical1pBasis",
                "categories": {"__id__": "ACE_SList", "T": _type_dict(T),
                               "list": [c if T == "Int64" else str(c) for c in B.categories]},
                "VSYM": B.varsym, "ISYM": B.idxsym, "label": B.label}
    else:
        raise TypeError(f"no write_dict for {type(B).__name__}")
    return {"__id__": "ACE_B1pComponent", "syms": syms, "basis": basis,
            "fval": {"__id__": "ACE_StaticGet", "expr": "ACE.Transforms.GetVal{:rr}"},
            "spec": spec, "degrees": degrees, "label": B.label}


def _component_specs(B):
    """Component spec as a list of {sym: value} in the component's order."""
    if isinstance(B, Rn1pBasis):
        return [{B.nsym: n} for n in range(1, len(B) + 1)]
    if isinstance(B, Ylm1pBasis):
        return [{B.lsym: l, B.msym: m} for l in range(B.L + 1) for m in range(-l, l + 1)]
    return [{B.idxsym: c} for c in B.categories]


def write_dict(obj):
    """ACE.jl ``write_dict`` of a LinearACEModel or SymmetricBasis."""
    if isinstance(obj, LinearACEModel):
        return {"__id__": "ACE_LinearACEModel", "basis": write_dict(obj.basis), "c": _array_dict(obj.c),
                "evaluator": {"__id__": "ACE_ProductEvaluator"}}
    if isinstance(obj, SymmetricBasis):
        pib = obj.pibasis
        B1p = pib.basis1p
        # 1p indices: position of each function's component values in the component specs
        comp = []
        for B in B1p.bases:
            cs = _component_specs(B)
            comp.append({tuple(sorted(d.items())): i + 1 for i, d in enumerate(cs)})
        indices = []
        for b in B1p.get_spec():
            d = dict(zip(B1p.syms, b))
            idx = []
            for B, lut in zip(B1p.bases, comp):
                idx.append(lut[tuple(sorted((s, d[s]) for s in B.syms))])
            indices.append(idx)
        it = pib.spec.iAA2iA
        iAA2iA = np.where(it >= 0, it + 1, 0)
        a2b = obj.A2Bmap
        if a2b.vals.ndim == 1:
            nz = [{"__id__": "ACE_Invariant", "val": float(v), "T": _type_dict("Float64")} for v in a2b.vals]
        else:
            nz = [{"__id__": "ACE_EuclideanVector", "val": _array_dict(np.asarray(v, dtype=np.complex128))}
                  for v in a2b.vals]
        return {
            "__id__": "ACE_SymmetricBasis",
            "pibasis": {"__id__": "ACE_PIBasis",
                        "basis1p": {"__id__": "ACE_Product1pBasis",
                                    "bases": [_component_dict(B) for B in B1p.bases],
                                    "indices": indices},
                        "spec": {"__id__": "ACE_PIBasisSpec", "orders": pib.spec.orders.tolist(),
                                 "iAA2iA": _array_dict(iAA2iA)},
                        "real": bool(pib.isreal)},
            "A2Bmap": _sparse_dict(a2b.shape[0], a2b.shape[1], a2b.rows, a2b.cols, nz),
            "symgrp": {"__id__": "ACE_O3", "lsym": "l", "msym": "m"},
            "isreal": bool(obj.isreal),
        }
    raise TypeError(f"no write_dict for {type(obj).__name__}")


# ---------------------------------------------------------------------------
# reader

_POLY = re.compile(r"^\s*r\s*->\s*\(\(\s*1\s*\+\s*([-+0-9.eE]+)\s*\)\s*/\s*\(\s*1\s*\+\s*r\s*\)\s*\)\s*\^\s*([0-9]+)\s*$")


def _read_component(D):
    cid = D.get("__id__")
    if cid == "ACE_Categorical1pBasis":
        lst = D["categories"]["list"]
        return Categorical1pBasis(lst, varsym=D.get("VSYM", "mu"), idxsym=D.get("ISYM", "mu"),
                                  label=D.get("label"))
    if cid != "ACE_B1pComponent":
        raise ValueError(f"unsupported 1p component {cid}")
    basis = D["basis"]
    syms = list(D["syms"])
    label = D.get("label", "")
    if basis.get("__id__") == "ACE_SHBasis":
        if len(syms) != 2:
            raise ValueError("SHBasis component needs (l, m) symbols")
        B = Ylm1pBasis(int(basis["maxL"]), lsym=syms[0], msym=syms[1], label=label or "Ylm")
        want = _component_specs(B)
    elif basis.get("__id__") == "ACE_SChain":
        F = basis["F"]
        if len(F) != 3 or F[0].get("exstr", "").replace(" ", "") != "rr->norm(rr)" \
                or F[2].get("__id__") != "ACE_OrthPolyBasis":
            raise ValueError("unsupported radial chain (need norm -> polytransform -> OrthPolyBasis)")
        m = _POLY.match(F[1].get("exstr", ""))
        if not m:
            raise ValueError(f"unsupported distance transform {F[1].get('exstr')!r} (polytransform only)")
        trans = PolyTransform(int(m.group(2)), float(m.group(1)))
        J = OrthPolyBasis.from_dict({k: (_read_array(v) if k in ("A", "B", "C", "tdf", "ww") else v)
                                     for k, v in F[2].items()})
        B = Rn1pBasis(J, trans, nsym=syms[0], label=label or "Rn")
        want = _component_specs(B)
    else:
        raise ValueError(f"unsupported 1p component basis {basis.get('__id__')}")
    got = [{s: int(b[s]) for s in syms} for b in D["spec"]]
    if got != want:
        raise ValueError(f"component {label!r}: spec order differs from the evaluator's")
    return B


def _read_symmetric_basis(D):
    if D.get("__id__") != "ACE_SymmetricBasis":
        raise ValueError("expected ACE_SymmetricBasis")
    sg = D.get("symgrp", {"__id__": "ACE_O3"})
    if sg.get("__id__") != "ACE_O3":
        raise ValueError(f"unsupported symmetry group {sg.get('__id__')}")
    P = D["pibasis"]
    B1pD = P["basis1p"]
    bases = [_read_component(c) for c in B1pD["bases"]]
    B1p = Product1pBasis(bases)
    comp_specs = [_component_specs(B) for B in bases]
    spec = []
    for idx in B1pD["indices"]:
        d = {}
        for B, cs, i in zip(bases, comp_specs, idx):
            d.update(cs[int(i) - 1])
        spec.append(tuple(d[s] for s in B1p.syms))
    B1p.set_spec(spec)
    S = P["spec"]
    orders = np.asarray(S["orders"], dtype=np.int64)
    it = np.asarray(_read_array(S["iAA2iA"]), dtype=np.int64)
    if it.ndim == 1:
        it = it.reshape(len(orders), -1)
    if it.shape[0] != len(orders) and it.shape[1] == len(orders):
        it = it.T
    iAA2iA = np.where(it > 0, it - 1, -1)
    pib = PIBasis(B1p, PIBasisSpec(orders, iAA2iA), bool(P.get("real", True)))
    nrows, ncols, rows, cols, nz = _read_sparse(D["A2Bmap"])
    vals = [_read_property_value(v) for v in nz]
    vec = len(vals) > 0 and np.ndim(vals[0]) == 1
    if ncols != len(orders):
        raise ValueError("A2Bmap columns != length of the PIBasis")
    vals_arr = np.asarray(vals, dtype=np.complex128 if vec else np.float64).reshape((-1, 3) if vec else (-1,))
    prop = EuclideanVector(bool(D.get("isreal", True))) if vec else Invariant()
    return SymmetricBasis(pib, A2BMap(nrows, ncols, rows, cols, vals_arr), prop, bool(D.get("isreal", True)))


def read_dict(D):
    """Rebuild a LinearACEModel or SymmetricBasis from its ACE.jl dict."""
    cid = D.get("__id__")
    if cid == "ACE_LinearACEModel":
        basis = _read_symmetric_basis(D["basis"])
        c = np.asarray(_read_array(D["c"]), dtype=np.float64).reshape(-1)
        ev = D.get("evaluator", {}).get("__id__", "ACE_ProductEvaluator")
        kind = {"ACE_ProductEvaluator": "hip", "ACE_NaiveEvaluator": "naive"}.get(ev)
        if kind is None:
            raise ValueError(f"unsupported evaluator {ev}")
        return LinearACEModel(basis, c, evaluator=kind)
    if cid == "ACE_SymmetricBasis":
        return _read_symmetric_basis(D)
    raise ValueError(f"unsupported object {cid}")


def save_json(path, obj):
    with open(path, "w") as f:
        json.dump(write_dict(obj), f)


def load_json(path):
    with open(path) as f:
        return read_dict(json.load(f))
