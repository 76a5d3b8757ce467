"""Compile the AA contraction into trie "run programs" for the device.

The energy of an environment is a polynomial in the 1p values,
``P(A) = c0 + sum_r c~_r prod_t A[i_rt]`` (evaluator.jl:121-147), with every
1p value ``A[iA] = s * conj?(A+[k])`` expressed through the pooled ``m >= 0``
values ``A+`` (``Y_l^{-m} = (-1)^m conj(Y_l^m)``, sphericalharmonics.jl:
394-398, and R_n real).

Monomials are grouped by their prefix (all factors but the last, factors in
ascending order of the pooled index, which is sorted by polynomial degree);
one *run* is a prefix plus its *leaves* (last factor + coefficient), so
``P = c0 + sum_runs Pprefix * sum_leaves c * A[leaf]``.  Runs are sorted
lexicographically so consecutive runs share leading prefix factors, which the
kernel keeps as running products in registers (depth-first evaluation of the
product DAG).

The forces need ``G[k] = dP/dA+[k] + conj(dP/dconj(A+[k]))`` for every
pooled value k: the adjoint of ``E = Re P`` (evaluator.jl:174-192) restated
in terms of A+, because ``dE = Re sum_k G[k] dA+[k]``.  Each ``G[k]`` is again
a polynomial (one order less) and is compiled the same way.

Coefficients are kept symbolically as *terms* ``(AA row, scale, conj)`` with
value ``scale * c~[row]`` (conjugated when ``conj``), so that new model
parameters only need ``c~ = A2B^T c`` and a gather (``ace_set_params``).

Internal factor id: ``2k + conj``; device factor id: ``k | 0x8000*conj``.
"""
from __future__ import annotations

from collections import defaultdict

import numpy as np

DEV_CONJ = 0x8000
RUN_W = 11        # int32 words per run (ACE_RUN_W of the C ABI)
MAX_ORDER = 8     # largest correlation order the programs encode


def pool_map(aspec):
    """Pooled ``m >= 0`` table and the map A index -> (k, conj, sign).

    Returns ``(pool, kk, cj, sgn)``: ``pool[k] = (mu, n, l, m>=0)`` sorted by
    (degree n-1+l, mu, l, m, n); for every 1p function ``iA``: pooled index
    ``kk[iA]``, ``cj[iA] = m < 0`` and ``sgn[iA] = (-1)^m`` for ``m < 0``.
    """
    keys = sorted({(int(mu), int(n), int(l), abs(int(m))) for (mu, n, l, m) in aspec},
                  key=lambda p: (p[1] - 1 + p[2], p[0], p[2], p[3], p[1]))
    index = {p: k for k, p in enumerate(keys)}
    kk = np.zeros(len(aspec), dtype=np.int64)
    cj = np.zeros(len(aspec), dtype=np.int64)
    sgn = np.ones(len(aspec))
    for i, (mu, n, l, m) in enumerate(aspec):
        kk[i] = index[(int(mu), int(n), int(l), abs(int(m)))]
        if m < 0:
            cj[i] = 1
            sgn[i] = -1.0 if (m % 2) else 1.0
    return np.asarray(keys, dtype=np.int64).reshape(-1, 4), kk, cj, sgn


def energy_polynomial(orders, iAA2iA, kk, cj, sgn):
    """Monomials of P with symbolic coefficients.

    Returns ``(poly, const_terms)``: ``poly[mono] = [(row, scale, conj)]``
    with ``mono`` a tuple of internal factor ids in ascending order."""
    poly = {}
    const_terms = []
    for i in range(len(orders)):
        o = int(orders[i])
        if o == 0:
            const_terms.append((i, 1.0, 0))
            continue
        idx = iAA2iA[i, :o]
        s = float(np.prod(sgn[idx]))
        mono = tuple(sorted(int(2 * kk[a] + cj[a]) for a in idx))
        poly.setdefault(mono, []).append((i, s, 0))
    return poly, const_terms


def merge_conjugates(poly, is_m0):
    """Merge every monomial with its mirror image.

    The property is ``Re P`` (Invariant: ``E = Re sum_AA c~ AA``; the
    EuclideanVector components likewise), and ``Re(c x) = Re(conj(c)
    conj(x))``.  A monomial and its conjugate (every factor ``A+[k]`` swapped
    with ``conj(A+[k])``; m = 0 pooled values are real and map to themselves)
    therefore contribute through one monomial whose coefficient is the sum of
    the first coefficient and the conjugate of the second.  The mirror pairs
    are the AA terms of m and -m, related by Y_l^{-m} = (-1)^m conj(Y_l^m)
    (sphericalharmonics.jl:394-398): the config-2 energy shrinks from 35.7k
    to 20.2k monomials.  ``Re P`` is unchanged, and so is the adjoint G[k] of
    every m != 0 pooled value (it is fixed by dE = Re sum_k G[k] dA+[k] with
    complex dA+[k]); for m = 0 pooled values only Re G[k] is fixed, and only
    Re G[k] enters the forces."""

    def conj_id(x):
        k = x >> 1
        return 2 * k if is_m0[k] else (x ^ 1)

    out = {}
    for mono, terms in poly.items():
        c = tuple(sorted(conj_id(f) for f in mono))
        if c < mono:
            out.setdefault(c, []).extend((row, s, 1 - cj) for (row, s, cj) in terms)
        else:
            out.setdefault(mono, []).extend(terms)
    return out


def gradient_polynomials(poly, npool, is_m0):
    """``G[k]`` polynomials (list over k) with symbolic coefficients.

    The derivative w.r.t. ``A+[k]`` (factor ``2k``) is taken directly; the
    derivative w.r.t. ``conj(A+[k])`` (factor ``2k+1``) is conjugated, which
    maps every remaining factor to its conjugate partner and conjugates the
    coefficient."""

    def conj_id(x):
        k = x >> 1
        return 2 * k if is_m0[k] else (x ^ 1)

    G = [defaultdict(list) for _ in range(npool)]
    for mono, terms in poly.items():
        seen = set()
        for f in mono:
            if f in seen:
                continue
            seen.add(f)
            mult = mono.count(f)
            rest = list(mono)
            rest.remove(f)
            k = f >> 1
            if (f & 1) == 0:
                key = tuple(sorted(rest))
                for (row, s, c) in terms:
                    G[k][key].append((row, s * mult, c))
            else:
                key = tuple(sorted(conj_id(x) for x in rest))
                for (row, s, c) in terms:
                    G[k][key].append((row, s * mult, 1 - c))
    return [dict(g) for g in G]


def row_classes(orders, iAA2iA, aspec):
    """Class index M + 1 of every AA row, M = the sum of its factors' m (the
    total m of the product); -1 for the order-0 row."""
    out = np.full(len(orders), -1, dtype=np.int64)
    m = np.asarray(aspec)[:, 3]
    for i in range(len(orders)):
        o = int(orders[i])
        if o:
            out[i] = int(m[iAA2iA[i, :o]].sum()) + 1
    return out


def class_vectors(cols, vals3, row_class, tol=1e-12):
    """Class structure of a EuclideanVector A2B map: ``(ok, u, sig)``.

    Every value of an AA column of class c (= M + 1, M the column's total m)
    should be one complex number times a unit vector u[c], with conj(u[c]) =
    sig[c] u[2 - c] (the l = 1 coupling of symmbasis.jl:267-285).  u[c] is the
    value of largest norm of the class (the first such), normalised; a class
    without entries takes conj(u) of its mirror.  ``ok``: every entry within
    tol * (largest norm of the map) of its class line and |sig[c]| >= 1 - tol
    where both classes occur.  ``vals3`` [nnz, 3] complex, in CSR order."""
    cls = np.asarray(row_class)[np.asarray(cols)]
    u = np.zeros((3, 3), dtype=np.complex128)
    sig = np.zeros(3, dtype=np.complex128)
    if np.any((cls < 0) | (cls > 2)):
        return False, u, sig
    V = np.asarray(vals3, dtype=np.complex128).reshape(-1, 3)
    nrm = np.sqrt(np.sum(V.real ** 2 + V.imag ** 2, axis=1))
    maxn = float(nrm.max()) if len(nrm) else 0.0
    have = [False] * 3
    for c in range(3):
        idx = np.nonzero(cls == c)[0]
        if len(idx) and nrm[idx].max() > 0:
            z = idx[int(np.argmax(nrm[idx]))]
            u[c] = V[z] / nrm[z]
            have[c] = True
    for c in range(3):
        if not have[c] and have[2 - c]:
            u[c] = np.conj(u[2 - c])
    U = u[cls]
    d = np.sum(V * np.conj(U), axis=1)
    R = V - d[:, None] * U
    if len(R) and np.sqrt(np.sum(R.real ** 2 + R.imag ** 2, axis=1)).max() > tol * maxn:
        return False, u, sig
    for c in range(3):
        sig[c] = np.conj(np.sum(u[c] * u[2 - c]))
        if have[c] and have[2 - c] and abs(sig[c]) < 1.0 - tol:
            return False, u, sig
    return True, u, sig


def a2b_is_classed(cols, vals3, row_class, tol=1e-12):
    """True when the EuclideanVector A2B map has the class structure of
    :func:`class_vectors`: then the three Cartesian outputs are fixed
    combinations of three scalar-coefficient polynomials Q_M."""
    return class_vectors(cols, vals3, row_class, tol)[0]


def term_class(term, row_class):
    """Class of a symbolic coefficient term (row, scale, conj): the row's
    class, mirrored (M -> -M) when the term is conjugated."""
    row, _, cj = term
    c = int(row_class[row])
    return 2 - c if cj else c


def split_by_class(poly, row_class):
    """The polynomial as three polynomials, one per term class (insertion
    order kept)."""
    out = [dict(), dict(), dict()]
    for mono, terms in poly.items():
        for t in terms:
            out[term_class(t, row_class)].setdefault(mono, []).append(t)
    return out


def eval_terms(terms, ct):
    """Numeric value of a symbolic coefficient for c~ (real or (nAA,3))."""
    v = 0.0
    for (row, s, c) in terms:
        x = ct[row] * s
        v = v + (np.conj(x) if c else x)
    return v


class RunProgram:
    """Flat encoding of a set of polynomials ("tasks").

    run  = (depth, jstart, nleaf, leaf_off, f1, ..., f7) (RUN_W words): depth =
           prefix length (0..7); jstart = first prefix level (1-based) that
           differs from the previous run of the same task (depth+1 when none).
    leaf = device factor id + symbolic coefficient terms.
    task = (run_begin, run_end, out_slot) + constant terms; out_slot -1 is
           the (real part of the) energy, otherwise a G slot.
    """

    def __init__(self):
        self.runs = []
        self.leaf_fid = []
        self.leaf_terms = []
        self.tasks = []
        self.task_terms = []

    @staticmethod
    def dev(f):
        return (f >> 1) | (DEV_CONJ if (f & 1) else 0)

    def add_task(self, poly, out_slot, const_terms=()):
        const_terms = list(const_terms)
        groups = defaultdict(list)
        for mono, terms in poly.items():
            if len(mono) == 0:
                const_terms.extend(terms)
                continue
            if len(mono) > MAX_ORDER:
                raise ValueError("correlation order above 8 is not supported by the device programs")
            groups[mono[:-1]].append((mono[-1], terms))
        r0 = len(self.runs)
        prev = ()
        for pre in sorted(groups):
            leaves = sorted(groups[pre], key=lambda x: x[0])
            d = len(pre)
            j = 1
            while j <= d and j <= len(prev) and pre[j - 1] == prev[j - 1]:
                j += 1
            f = [self.dev(x) for x in pre] + [0] * (RUN_W - 4 - d)
            self.runs.append((d, j, len(leaves), len(self.leaf_fid), *f))
            for fl, terms in leaves:
                self.leaf_fid.append(self.dev(fl))
                self.leaf_terms.append(terms)
            prev = pre
        self.tasks.append((r0, len(self.runs), out_slot))
        self.task_terms.append(const_terms)
        return len(self.tasks) - 1

    def cost(self, t, leaf_w=1.0, run_w=2.0):
        r0, r1 = self.tasks[t][0], self.tasks[t][1]
        return leaf_w * sum(self.runs[r][2] for r in range(r0, r1)) + run_w * (r1 - r0)


def split_poly(poly, nparts):
    """Split a polynomial into ``nparts`` with balanced cost, keeping whole
    prefix groups (runs) together."""
    groups = defaultdict(list)
    for mono in poly:
        groups[mono[:-1]].append(mono)
    keys = sorted(groups, key=lambda p: (-len(groups[p]), p))
    parts = [dict() for _ in range(nparts)]
    load = [0.0] * nparts
    for p in keys:
        i = int(np.argmin(load))
        for mono in groups[p]:
            parts[i][mono] = poly[mono]
        load[i] += len(groups[p]) + 2.0
    return [p for p in parts if p]


def lpt_assign(costs, W):
    """Longest-processing-time assignment of tasks to W waves."""
    order = sorted(range(len(costs)), key=lambda t: -costs[t])
    load = [0.0] * W
    lists = [[] for _ in range(W)]
    for t in order:
        w = int(np.argmin(load))
        lists[w].append(t)
        load[w] += costs[t]
    for l in lists:
        l.sort()
    return lists, load
