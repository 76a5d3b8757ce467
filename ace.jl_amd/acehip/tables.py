"""Flat basis tables: the shared definition of a built basis.

These plain arrays are what the device pack is made from and what a CPU
checker consumes, so both evaluate the identical basis.  The JSON form uses
the same keys.
"""
from __future__ import annotations

import json

import numpy as np


def basis_tables(basis) -> dict:
    """Export a ``SymmetricBasis`` as flat tables.

    ``aspec[iA] = (mu, n, l, m)`` with ``mu`` the category index (-1 when the
    1p basis has no categorical component), ``n`` 1-based radial index.
    """
    B1p = basis.basis1p
    Rn = B1p.component("Rn")
    Yl = B1p.component("Ylm")
    Cat = B1p.component("Categorical")
    syms = B1p.syms
    aspec = []
    for b in B1p.get_spec():
        d = dict(zip(syms, b))
        mu = Cat.val2i(d[Cat.idxsym]) if Cat is not None else -1
        aspec.append((mu, d[Rn.nsym], d[Yl.lsym], d[Yl.msym]))
    J = Rn.J
    a2b = basis.A2Bmap
    return {
        "prop": basis.prop.name,
        "rn": {"A": J.A.copy(), "B": J.B.copy(), "C": J.C.copy(), "pl": J.pl,
               "tl": J.tl, "pr": J.pr, "tr": J.tr, "p": Rn.trans.p, "r0": Rn.trans.r0,
               "rcut": Rn.rcut, "rin": Rn.rin},
        "L": Yl.L,
        "categories": list(Cat.categories) if Cat is not None else None,
        "aspec": np.asarray(aspec, dtype=np.int64).reshape(-1, 4),
        "orders": basis.pibasis.spec.orders.copy(),
        "iAA2iA": basis.pibasis.spec.iAA2iA.copy(),
        "a2b": {"rows": a2b.rows.copy(), "cols": a2b.cols.copy(),
                "vals": a2b.vals.copy(), "shape": a2b.shape},
    }


def _enc(x):
    if isinstance(x, np.ndarray):
        if np.iscomplexobj(x):
            return {"__complex__": True, "re": x.real.tolist(), "im": x.imag.tolist()}
        return x.tolist()
    if isinstance(x, dict):
        return {k: _enc(v) for k, v in x.items()}
    if isinstance(x, tuple):
        return list(x)
    return x


def _dec(x, key=None):
    if isinstance(x, dict):
        if x.get("__complex__"):
            return np.asarray(x["re"]) + 1j * np.asarray(x["im"])
        return {k: _dec(v, k) for k, v in x.items()}
    if isinstance(x, list) and key in ("A", "B", "C", "vals", "aspec", "orders",
                                       "iAA2iA", "rows", "cols"):
        return np.asarray(x)
    return x


def dump_tables(T: dict, path: str):
    with open(path, "w") as f:
        json.dump(_enc(T), f)


def load_tables(path: str) -> dict:
    with open(path) as f:
        T = _dec(json.load(f))
    T["a2b"]["shape"] = tuple(T["a2b"]["shape"])
    return T
