"""Basis selection, sparse-grid enumeration and the PIBasis specification.

Reference semantics:
  * ``SimpleSparseBasis(maxorder, maxdeg)``: level = total degree
    (ACE.jl ``src/basisselectors.jl:106-124``).
  * ``gensparse`` / ``_gensparse`` with ``ordered = true``
    (``src/sparsegrids.jl:55-148``): ascending index tuples with ``0`` marking
    an empty slot, enumerated by incrementing the last slot and carrying on
    overshoot.  The order of this enumeration fixes the AA row order.
  * ``PIBasisSpec(AAspec)`` stores each tuple reversed (descending) with the
    correlation order alongside (``src/pibasis.jl:35-101``); row 1 is the
    constant when the property admits it.

Indices here are 0-based; an empty slot is ``-1``.
"""
from __future__ import annotations

import numpy as np


class SimpleSparseBasis:
    def __init__(self, maxorder: int, maxdeg: float):
        self.maxorder = int(maxorder)
        self.maxlevel = float(maxdeg)

    # 1p level (basisselectors.jl:111-112)
    def level1(self, b: dict, B1p) -> float:
        return B1p.degree(b)

    def filter1(self, b: dict, B1p) -> bool:
        return True

    # product level (basisselectors.jl:114-115)
    def level(self, bb, B1p) -> float:
        return 0 if len(bb) == 0 else sum(B1p.degree(b) for b in bb)

    def filter(self, bb, B1p) -> bool:
        return True


class NoConstant:
    """basisselectors.jl:91-94."""

    def __call__(self, bb):
        return len(bb) > 0


def gensparse_ordered(NU: int, maxvv: int, admissible, filt):
    """sparsegrids.jl:69-148 for ``ordered = true``; ``vv`` entries are 1-based
    with 0 meaning "empty".  Returns the list of accepted tuples."""
    vv = [0] * NU
    spec = []
    if NU == 0:
        if admissible(vv) and filt(vv):
            spec.append(tuple(vv))
        return spec
    lastidx = 0
    while True:
        if any(v > maxvv for v in vv):
            isadm = False
        else:
            isadm = admissible(vv)
        if isadm:
            if filt(vv):
                spec.append(tuple(vv))
            lastidx = NU
            vv[lastidx - 1] += 1
        else:
            if lastidx == 0:
                raise ValueError("gensparse: the smallest basis function is "
                                 "already inadmissible; the basis is empty")
            if lastidx == 1:
                break
            vv[lastidx - 2] += 1
            for i in range(lastidx - 1, NU):
                vv[i] = vv[lastidx - 2]
            lastidx -= 1
    return spec


class PIBasisSpec:
    """``orders[i]`` and ``iAA2iA[i, :]`` (descending 0-based A indices, -1 pad)."""

    def __init__(self, orders, iAA2iA):
        self.orders = np.asarray(orders, dtype=np.int64)
        self.iAA2iA = np.asarray(iAA2iA, dtype=np.int64)

    def __len__(self):
        return len(self.orders)

    @property
    def maxorder(self):
        return self.iAA2iA.shape[1]

    def get(self, i):
        return tuple(int(x) for x in self.iAA2iA[i, : self.orders[i]])

    @classmethod
    def from_tuples(cls, AAspec, NU):
        orders = np.zeros(len(AAspec), dtype=np.int64)
        iAA2iA = -np.ones((len(AAspec), NU), dtype=np.int64)
        for i, vv in enumerate(AAspec):
            nz = [v - 1 for v in reversed(vv) if v != 0]
            orders[i] = len(nz)
            iAA2iA[i, : len(nz)] = nz
        return cls(orders, iAA2iA)

    def sparsify(self, keep):
        keep = np.asarray(keep, dtype=np.int64)
        return PIBasisSpec(self.orders[keep], self.iAA2iA[keep, :])

    def fix_A_indices(self, new_inds):
        new_inds = np.asarray(new_inds, dtype=np.int64)
        out = self.iAA2iA.copy()
        for i in range(len(self.orders)):
            for a in range(self.orders[i]):
                v = new_inds[out[i, a]]
                assert v >= 0
                out[i, a] = v
        self.iAA2iA = out


def build_pispec(B1p, Bsel, prop=None, filterfun=None, init1pbasis=True):
    """pibasis.jl:35-87."""
    from .onep import init1pspec

    if init1pbasis:
        init1pspec(B1p, Bsel)
    Aspec = B1p.get_spec()
    Adicts = [B1p.as_dict(b) for b in Aspec]
    levels = [Bsel.level1(b, B1p) for b in Adicts]
    if any(levels[i] > levels[i + 1] for i in range(len(levels) - 1)):
        raise ValueError("PIBasisSpec: Aspec is not sorted by degree")
    deg = np.asarray([B1p.degree(b) for b in Adicts], dtype=np.float64)

    def tup2b(vv):
        return [Adicts[v - 1] for v in vv if v != 0]

    def admissible(vv):
        return sum(deg[v - 1] for v in vv if v != 0) <= Bsel.maxlevel

    def filt(vv):
        bb = tup2b(vv)
        if filterfun is not None and not filterfun(bb):
            return False
        if not Bsel.filter(bb, B1p):
            return False
        if prop is not None and not prop.filter(bb):
            return False
        return True

    AAspec = gensparse_ordered(Bsel.maxorder, len(Aspec), admissible, filt)
    return PIBasisSpec.from_tuples(AAspec, Bsel.maxorder)


class PIBasis:
    """pibasis.jl:145-149: 1p basis + product spec + whether AA is real."""

    def __init__(self, basis1p, spec: PIBasisSpec, isreal: bool):
        self.basis1p = basis1p
        self.spec = spec
        self.isreal = bool(isreal)

    def __len__(self):
        return len(self.spec)

    def get_spec(self, i=None):
        A = self.basis1p.get_spec()
        if i is None:
            return [[A[j] for j in self.spec.get(k)] for k in range(len(self.spec))]
        return [A[j] for j in self.spec.get(i)]

    def clean_1pbasis(self):
        """pibasis.jl:194-207."""
        used = []
        seen = set()
        for bb in self.get_spec():
            for b in bb:
                if b not in seen:
                    seen.add(b)
                    used.append(b)
        new_inds = self.basis1p.sparsify(used)
        self.spec.fix_A_indices(new_inds)
