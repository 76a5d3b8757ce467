"""O(3) coupling coefficients: Clebsch-Gordan, generalized CG, and the
rotation-equivariant + permutation-invariant reduction.  Set-up only (CPU).

Reference semantics (ACE.jl):
  * ``clebschgordan(j1,m1,j2,m2,J,M)`` exact formula in big arithmetic
    (``src/rotations3d.jl:108-137``); here exact rationals + 50-digit sqrt.
  * generalized CG ``A(ll, mm, kk)`` by recursion on the last two entries
    (``rotations3d.jl:191-248``), with ``coco_init`` at order 1
    (``src/properties.jl:139-142`` Invariant, ``:219-222`` EuclideanVector).
    Unrolled, the recursion is ``A(ll,mm,kk) = sum_p U_p(mm) U_p(kk) I(p)``
    over intermediate couplings ``p``; this module builds ``U`` once per ``ll``
    and forms the sums as matrix products (same values, different rounding).
  * ``re_basis`` Gram matrix + SVD + rank(rtol=1e-7) (``rotations3d.jl:257-273``),
    ``compute_Al`` (``:277-329``), ``rpe_basis`` and ``_gramian``
    (``src/symmetrygroups.jl:119-146``).
"""
from __future__ import annotations

import itertools
from decimal import Decimal, getcontext
from fractions import Fraction
from functools import lru_cache
from math import comb, factorial

import numpy as np

getcontext().prec = 50


def _cg_conditions(j1, m1, j2, m2, J, M):
    return (abs(j1 - j2) <= J <= j1 + j2) and (M == m1 + m2) and \
        abs(m1) <= j1 and abs(m2) <= j2 and abs(M) <= J


@lru_cache(maxsize=None)
def clebschgordan(j1, m1, j2, m2, J, M) -> float:
    """``C_{j1 m1 j2 m2}^{J M}`` (rotations3d.jl:108-137)."""
    if not _cg_conditions(j1, m1, j2, m2, J, M):
        return 0.0
    N = Fraction((2 * J + 1) * factorial(j1 + m1) * factorial(j1 - m1)
                 * factorial(j2 + m2) * factorial(j2 - m2)
                 * factorial(J + M) * factorial(J - M),
                 factorial(j1 + j2 - J) * factorial(j1 - j2 + J)
                 * factorial(-j1 + j2 + J) * factorial(j1 + j2 + J + 1))
    lb = max(0, j2 - J - m1, j1 - J + m2)
    ub = min(j1 + j2 - J, j1 - m1, j2 + m2)
    G = 0
    for k in range(lb, ub + 1):
        G += (-1) ** k * comb(j1 + j2 - J, k) * comb(j1 - j2 + J, j1 - m1 - k) \
            * comb(-j1 + j2 + J, j2 + m2 - k)
    if G == 0:
        return 0.0
    sq = (Decimal(N.numerator) / Decimal(N.denominator)).sqrt()
    return float(sq * G)


# ---------------------------------------------------------------------------
# properties: filters and order-1 initial coupling


# crmatrices (ACE.jl src/eucl/cov_coeffs_dict.jl:1-29): data table for the
# l = 1 EuclideanVector coupling, keyed by (m, mu, i) -> complex 3-vector.
_a = 0.1666666666666667
_b = 0.2357022603955158
_c = 0.3333333333333333
_CR = {
    (-1, -1, 1): (_a, 1j * _a, 0), (-1, 1, 1): (-_a, -1j * _a, 0),
    (0, -1, 1): (0, 0, -_b), (0, 1, 1): (0, 0, _b),
    (1, -1, 1): (-_a, 1j * _a, 0), (1, 1, 1): (_a, -1j * _a, 0),
    (-1, -1, 2): (-1j * _a, _a, 0), (-1, 1, 2): (-1j * _a, _a, 0),
    (0, -1, 2): (0, 0, 1j * _b), (0, 1, 2): (0, 0, 1j * _b),
    (1, -1, 2): (1j * _a, _a, 0), (1, 1, 2): (1j * _a, _a, 0),
    (-1, 0, 3): (-_b, -1j * _b, 0), (0, 0, 3): (0, 0, _c),
    (1, 0, 3): (_b, -1j * _b, 0),
}


def crmatrix(m, mu, i):
    return np.asarray(_CR.get((m, mu, i), (0, 0, 0)), dtype=np.complex128)


class Invariant:
    """Scalar invariant output (properties.jl:77-152)."""
    name = "Invariant"
    isrealAA = True
    isrealB = True
    numcc = 1
    final_l = 0

    def filter(self, bb) -> bool:
        """properties.jl:115-125 (with the m symbol present)."""
        if len(bb) <= 1:
            return True
        suml = sum(b["l"] for b in bb)
        summ = sum(b["m"] for b in bb)
        return suml % 2 == 0 and summ == 0

    def coco_filter(self, ll, mm) -> bool:
        return sum(ll) % 2 == 0 and sum(mm) == 0

    def to_dict(self):
        return {"__id__": "ACE_Invariant", "val": 0.0, "T": {"__id__": "Float64"}}


class EuclideanVector:
    """Equivariant 3-vector output (properties.jl:165-248)."""
    name = "EuclideanVector"
    isrealAA = False
    numcc = 3
    final_l = 1

    def __init__(self, isrealB: bool = True):
        self.isrealB = bool(isrealB)

    def filter(self, bb) -> bool:
        """properties.jl:189-203."""
        if len(bb) == 0:
            return False
        if len(bb) == 1:
            return True
        suml = sum(b["l"] for b in bb)
        summ = sum(b["m"] for b in bb)
        return suml % 2 == 1 and abs(summ) <= 1

    def coco_filter(self, ll, mm) -> bool:
        return sum(ll) % 2 == 1 and abs(sum(mm)) <= 1

    def to_dict(self):
        return {"__id__": "ACE_EuclideanVector",
                "val": {"__id__": "ACE_SVector", "vals": [0.0, 0.0, 0.0]}}


# ---------------------------------------------------------------------------
# generalized CG via coupling paths


def mrange(phi, ll):
    """``_mrange`` (rotations3d.jl:49-66): all m-tuples, first index fastest,
    filtered by ``coco_filter``."""
    ranges = [range(-l, l + 1) for l in ll]
    out = []
    for mm_rev in itertools.product(*reversed(ranges)):
        mm = tuple(reversed(mm_rev))
        if phi.coco_filter(ll, mm):
            out.append(mm)
    return out


def _paths(ll, final_l):
    """Intermediate couplings: couple the last two entries into J1, then the
    next-to-last with J1 into J2, ...; the last coupling must give final_l."""
    N = len(ll)
    if N == 1:
        return [()] if ll[0] == final_l else []
    out = []

    def rec(pos, Jcur, acc):
        # couple ll[pos] with Jcur
        lo, hi = abs(ll[pos] - Jcur), ll[pos] + Jcur
        if pos == 0:
            if lo <= final_l <= hi:
                out.append(tuple(acc) + (final_l,))
            return
        for J in range(lo, hi + 1):
            rec(pos - 1, J, acc + [J])

    l1, l2 = ll[N - 2], ll[N - 1]
    for J1 in range(abs(l1 - l2), l1 + l2 + 1):
        if N == 2:
            if J1 == final_l:
                out.append((J1,))
        else:
            rec(N - 3, J1, [J1])
    return out


def _path_values(ll, paths, mlist):
    """U[p, i] = prod of CG along path p for m-tuple mlist[i]; also the final
    total M for each m-tuple."""
    N = len(ll)
    U = np.zeros((len(paths), len(mlist)))
    for ip, p in enumerate(paths):
        for im, mm in enumerate(mlist):
            if N == 1:
                U[ip, im] = 1.0
                continue
            M = mm[N - 2] + mm[N - 1]
            v = clebschgordan(ll[N - 2], mm[N - 2], ll[N - 1], mm[N - 1], p[0], M)
            Jprev = p[0]
            for s, pos in enumerate(range(N - 3, -1, -1)):
                if v == 0.0:
                    break
                Mn = mm[pos] + M
                v *= clebschgordan(ll[pos], mm[pos], Jprev, M, p[s + 1], Mn)
                Jprev, M = p[s + 1], Mn
            U[ip, im] = v
    return U


@lru_cache(maxsize=None)
def compute_CC(phi_name: str, ll: tuple):
    """``compute_Al`` (rotations3d.jl:277-329): rows of generalized CG
    coefficients.  Returns (CC, Mll) where CC has shape
    ``(len(Mll) * numcc, len(Mll) [, 3])``: row ``ik*numcc + p`` holds
    ``A(ll, mm, Mll[ik])[p]`` for every ``mm`` in ``Mll``."""
    phi = _PHI[phi_name]
    Mll = mrange(phi, ll)
    if len(Mll) == 0:
        return None, Mll
    paths = _paths(ll, phi.final_l)
    U = _path_values(ll, paths, Mll)              # (npath, nM)
    if phi.numcc == 1:
        CC = U.T @ U                                # CC[ik, im]
        return CC, Mll
    # EuclideanVector: A(ll,mm,kk)[i] = sum_p U[p,mm] U[p,kk] conj(cr(-M,-K,i))
    nM = len(Mll)
    Ms = np.array([sum(mm) for mm in Mll])
    S = U.T @ U                                     # S[ik, im]
    CC = np.zeros((nM * 3, nM, 3), dtype=np.complex128)
    for ik in range(nM):
        for i in range(3):
            for im in range(nM):
                if S[ik, im] == 0.0:
                    continue
                CC[ik * 3 + i, im, :] = S[ik, im] * np.conj(crmatrix(-Ms[im], -Ms[ik], i + 1))
    return CC, Mll


def _svd_rank(G):
    U, S, _ = np.linalg.svd(G)
    if len(S) == 0 or S[0] == 0.0:
        return U, S, 0
    rk = int(np.sum(S > 1e-7 * S[0]))
    return U, S, rk


@lru_cache(maxsize=None)
def re_basis(phi_name: str, ll: tuple):
    """rotations3d.jl:257-273: rotation-equivariant basis for ``ll``."""
    phi = _PHI[phi_name]
    CC, Mll = compute_CC(phi_name, ll)
    if CC is None:
        return None, Mll
    if phi.numcc == 1:
        G = CC @ CC.T
        U, S, rk = _svd_rank(G)
        Ured = np.sqrt(S[:rk])[:, None] * U[:, :rk].T
        return Ured @ CC, Mll                      # (rk, nM)
    flat = CC.reshape(CC.shape[0], -1)
    G = np.conj(flat) @ flat.T
    U, S, rk = _svd_rank(G)
    Ured = np.sqrt(S[:rk])[:, None] * np.conj(U[:, :rk].T)
    Ure = (Ured @ flat).reshape(rk, CC.shape[1], 3)
    return Ure, Mll


def rpe_basis(phi_name: str, nn: tuple, ll: tuple):
    """symmetrygroups.jl:119-146: rotation-equivariant, permutation-invariant
    coupling coefficients for the product with radial/species labels ``nn``
    and angular labels ``ll``."""
    phi = _PHI[phi_name]
    Ure, Mre = re_basis(phi_name, ll)
    if Ure is None or Ure.shape[0] == 0:
        return np.zeros((0, len(Mre))), Mre
    N = len(ll)
    index = {mm: i for i, mm in enumerate(Mre)}
    nre = Ure.shape[0]
    if phi.numcc == 1:
        G = np.zeros((nre, nre))
    else:
        G = np.zeros((nre, nre), dtype=np.complex128)
    for sigma in itertools.permutations(range(N)):
        if any(nn[s] != nn[i] for i, s in enumerate(sigma)) or \
           any(ll[s] != ll[i] for i, s in enumerate(sigma)):
            continue
        # G[i1,i2] += sum_{iU1} Ure[i1,iU1] Ure[i2, idx(mm1[sigma])]
        perm = np.array([index[tuple(mm[s] for s in sigma)] for mm in Mre])
        if phi.numcc == 1:
            G += Ure @ Ure[:, perm].T
        else:
            a = Ure.reshape(nre, -1)
            b = Ure[:, perm, :].reshape(nre, -1)
            G += np.conj(a) @ b.T
    U, S, rk = _svd_rank(G)
    if phi.numcc == 1:
        Urpe = U[:, :rk].T
        return (np.sqrt(S[:rk])[:, None] * Urpe) @ Ure, Mre
    Urpe = np.conj(U[:, :rk].T)
    out = (np.sqrt(S[:rk])[:, None] * Urpe) @ Ure.reshape(nre, -1)
    return out.reshape(rk, len(Mre), 3), Mre


_PHI = {"Invariant": Invariant(), "EuclideanVector": EuclideanVector()}
