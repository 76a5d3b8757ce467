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


class SparseBasis:
    """``SparseBasis(; maxorder, p, weight, default_maxdeg | maxlevels)``
    (basisselectors.jl:175-198, 133-158): the level of a 1p function is its
    weighted degree (``weight[:n] (n - 1) + weight[:l] l``, b1pcomponent.jl:
    145-147; a categorical component has degree 0, discrete1pbasis.jl:123),
    the level of a product the p-norm of its factors' levels, and the bound
    depends on the correlation order: ``maxlevels[ord]``, else
    ``maxlevels["default"]``."""

    def __init__(self, maxorder: int, p: float = 1.0, weight=None, default_maxdeg=None, maxlevels=None):
        if (default_maxdeg is None) == (maxlevels is None):
            raise ValueError("SparseBasis: give exactly one of default_maxdeg and maxlevels")
        self.maxorder = int(maxorder)
        self.p = float(p)
        self.weight = {"l": 1.0, "n": 1.0} if weight is None else {k: float(v) for k, v in weight.items()}
        self.maxlevels = {"default": float(default_maxdeg)} if maxlevels is None else \
            {k: float(v) for k, v in maxlevels.items()}
        # maxlevel(Bsel, basis): the largest bound over orders 1..maxorder (the 1p cut)
        self.maxlevel = max(self.maxlevel_order(o) for o in range(1, self.maxorder + 1))

    def maxlevel_order(self, order: int) -> float:
        return self.maxlevels.get(order, self.maxlevels["default"])

    def level1(self, b: dict, B1p) -> float:
        return B1p.degree(b, self.weight)

    def filter1(self, b: dict, B1p) -> bool:
        return True

    def level(self, bb, B1p) -> float:
        return 0.0 if len(bb) == 0 else pnorm([self.level1(b, B1p) for b in bb], self.p)

    def filter(self, bb, B1p) -> bool:
        return True


def pnorm(x, p: float) -> float:
    """``norm(x, p)`` of non-negative levels as Julia's LinearAlgebra computes
    it: a running sum for p = 1, the maximum for p = Inf, else the sum of
    powers of x / max(x), rooted and rescaled by max(x)."""
    x = [float(v) for v in x]
    if not x:
        return 0.0
    if p == 1.0:
        s = 0.0
        for v in x:
            s += abs(v)
        return s
    m = max(abs(v) for v in x)
    if p == float("inf") or m == 0.0:
        return m
    s = 0.0
    for v in x:
        s += (abs(v) / m) ** p
    return m * s ** (1.0 / p)


class NoConstant:
    """basisselectors.jl:91-94."""

    def __call__(self, bb):
        return len(bb) > 0


def gensparse_levels(NU: int, lev, maxlevel_of, p: float = 1.0):
    """All tuples of ``gensparse_ordered`` for the downset of a level that is
    the p-norm of 1p levels ``lev`` (non-decreasing in the 1p index) under a
    bound ``maxlevel_of(order)``; int32 [ntuples, NU], 1-based, 0 = empty, in
    the traversal's order, without the per-tuple filter.

    For such a level the traversal of sparsegrids.jl:85-148 visits, order by
    order (every order-k tuple precedes the order-(k+1) ones, since the empty
    slots lead), exactly the admissible non-decreasing tuples in
    lexicographic order: within a row the level grows with the last index,
    so the row ends at its first inadmissible tuple and nothing admissible is
    skipped, and a row whose first tuple is inadmissible ends its parent row
    for the same reason.  Each order is therefore built by extending the
    admissible prefixes with contiguous index ranges; the final test
    recomputes each tuple's level as ``pnorm`` does."""
    lev = np.asarray(lev, dtype=np.float64)
    nA = len(lev)
    if nA and np.any(np.diff(lev) < 0):
        raise ValueError("gensparse_levels: 1p levels must be non-decreasing")
    if np.any(lev < 0):
        raise ValueError("gensparse_levels: 1p levels must be non-negative")
    out = [np.zeros((1, NU), dtype=np.int32)]                 # the empty tuple (level 0)
    pinf = p == float("inf")
    lp = lev if (pinf or p == 1.0) else lev ** p
    for k in range(1, NU + 1):
        bound = maxlevel_of(k)
        # the prefix search uses a slightly widened bound; the exact test follows
        bp = (bound if (pinf or p == 1.0) else bound ** p) * (1 + 1e-12) + 1e-12
        T = np.arange(nA, dtype=np.int32)[:, None]
        acc = lp.copy()
        keep = acc <= bp
        T, acc = T[keep], acc[keep]
        for _ in range(1, k):
            last = T[:, -1].astype(np.int64)
            if pinf:
                hi = np.full(len(T), np.searchsorted(lp, bp, side="right"), dtype=np.int64)
            else:
                hi = np.searchsorted(lp, bp - acc, side="right").astype(np.int64)
            cnt = np.maximum(hi - last, 0)
            rows = np.repeat(np.arange(len(T)), cnt)
            start = np.repeat(np.cumsum(cnt) - cnt, cnt)
            nxt = (last[rows] + (np.arange(len(rows)) - start)).astype(np.int32)
            acc = np.maximum(acc[rows], lp[nxt]) if pinf else acc[rows] + lp[nxt]
            T = np.concatenate([T[rows], nxt[:, None]], axis=1)
        if len(T) == 0:
            continue
        # exact level of every tuple (left to right, as pnorm)
        L = lev[T]
        if p == 1.0:
            lv = np.zeros(len(T))
            for j in range(k):
                lv = lv + np.abs(L[:, j])
        else:
            m = np.abs(L).max(axis=1)
            if pinf:
                lv = m
            else:
                sm = np.zeros(len(T))
                safe = np.where(m > 0, m, 1.0)
                for j in range(k):
                    sm = sm + (np.abs(L[:, j]) / safe) ** p
                lv = np.where(m > 0, m * sm ** (1.0 / p), 0.0)
        T = T[lv <= bound]
        full = np.zeros((len(T), NU), dtype=np.int32)
        full[:, NU - k:] = T + 1
        out.append(full)
    return np.concatenate(out, axis=0)


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
    def from_array(cls, V):
        """From gensparse tuples V [n, NU] (1-based, 0 = empty, ascending):
        iAA2iA holds each reversed (descending), 0-based, -1 padded."""
        V = np.asarray(V, dtype=np.int64)
        orders = (V > 0).sum(axis=1)
        iAA2iA = (V[:, ::-1] - 1).astype(np.int64)  # empty slots (0) -> -1, last after reversal
        return cls(orders, iAA2iA)

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


def build_pispec(B1p, Bsel, prop=None, filterfun=None, init1pbasis=True, fast=True):
    """pibasis.jl:35-87.  ``fast``: enumerate with :func:`gensparse_levels`
    and a vectorised filter when the selector and filters allow it (same
    tuples, same order); otherwise, or with ``fast=False``, the literal
    traversal :func:`gensparse_ordered` with per-tuple filters."""
    from .onep import init1pspec

    if init1pbasis:
        init1pspec(B1p, Bsel)
    Aspec = B1p.get_spec()
    Adicts = [B1p.as_dict(b) for b in Aspec]
    levels = [Bsel.level1(b, B1p) for b in Adicts]
    if any(levels[i] > levels[i + 1] for i in range(len(levels) - 1)):
        raise ValueError("PIBasisSpec: Aspec is not sorted by degree")
    deg = np.asarray([B1p.degree(b) for b in Adicts], dtype=np.float64)
    fast_mask = _fast_filter(prop, filterfun, Bsel)
    if fast and fast_mask is not None:
        p = Bsel.p if isinstance(Bsel, SparseBasis) else 1.0
        maxlev = Bsel.maxlevel_order if isinstance(Bsel, SparseBasis) else (lambda k: Bsel.maxlevel)
        V = gensparse_levels(Bsel.maxorder, levels, maxlev, p)
        ll = np.asarray([b.get("l", 0) for b in Adicts], dtype=np.int64)
        mm = np.asarray([b.get("m", 0) for b in Adicts], dtype=np.int64)
        V = V[fast_mask(V, ll, mm)]
        return PIBasisSpec.from_array(V)

    def tup2b(vv):
        return [Adicts[v - 1] for v in vv if v != 0]

    def admissible(vv):
        if isinstance(Bsel, SparseBasis):
            bb = [v for v in vv if v != 0]
            return pnorm([levels[v - 1] for v in bb], Bsel.p) <= Bsel.maxlevel_order(len(bb))
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


def _fast_filter(prop, filterfun, Bsel):
    """The tuple filter of build_pispec as a vectorised mask over gensparse
    tuples V (1-based, 0 = empty) when every part of it is known: no
    filterfun or NoConstant, a selector without a filter of its own, and the
    Invariant (properties.jl:115-125) or EuclideanVector (:189-203)
    selection rules on the sums of l and m.  None otherwise."""
    from .coupling import EuclideanVector, Invariant

    if filterfun is not None and not isinstance(filterfun, NoConstant):
        return None
    if prop is not None and not isinstance(prop, (Invariant, EuclideanVector)):
        return None
    if type(Bsel) not in (SimpleSparseBasis, SparseBasis):
        return None

    def mask(V, ll, mm):
        nz = V > 0
        order = nz.sum(axis=1)
        idx = np.where(nz, V - 1, 0)
        suml = np.where(nz, ll[idx], 0).sum(axis=1)
        summ = np.where(nz, mm[idx], 0).sum(axis=1)
        keep = np.ones(len(V), dtype=bool)
        if isinstance(filterfun, NoConstant):
            keep &= order > 0
        if isinstance(prop, Invariant):
            keep &= (order <= 1) | ((suml % 2 == 0) & (summ == 0))
        elif isinstance(prop, EuclideanVector):
            keep &= (order == 1) | ((order >= 2) & (suml % 2 == 1) & (np.abs(summ) <= 1))
        return keep

    return mask


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
