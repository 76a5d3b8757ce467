"""Basis construction (set-up, CPU): Clebsch-Gordan, generalized CG, 1p and
AA specification order, cleaning."""
import itertools

import numpy as np
import pytest
from sympy import S
from sympy.physics.quantum.cg import CG

import acehip as ah
from acehip import coupling
from helpers import basis_for


def test_clebsch_gordan_vs_sympy():
    # test/test_cg.jl:16-43 restated against the local sympy
    rng = np.random.default_rng(0)
    for _ in range(200):
        j1, j2 = rng.integers(0, 5, 2)
        J = rng.integers(abs(j1 - j2), j1 + j2 + 1)
        m1 = rng.integers(-j1, j1 + 1)
        m2 = rng.integers(-j2, j2 + 1)
        M = m1 + m2
        ref = float(CG(S(int(j1)), S(int(m1)), S(int(j2)), S(int(m2)), S(int(J)), S(int(M))).doit()) \
            if abs(M) <= J else 0.0
        assert coupling.clebschgordan(int(j1), int(m1), int(j2), int(m2), int(J), int(M)) == pytest.approx(ref, abs=1e-14)


def test_cg_orthogonality():
    for j1, j2 in [(1, 1), (2, 3), (4, 2)]:
        for J1 in range(abs(j1 - j2), j1 + j2 + 1):
            for J2 in range(abs(j1 - j2), j1 + j2 + 1):
                s = sum(coupling.clebschgordan(j1, m1, j2, M - m1, J1, M) *
                        coupling.clebschgordan(j1, m1, j2, M - m1, J2, M)
                        for M in range(-min(J1, J2), min(J1, J2) + 1)
                        for m1 in range(-j1, j1 + 1) if abs(M - m1) <= j2)
                assert s == pytest.approx((2 * J1 + 1) if J1 == J2 else 0.0, abs=1e-12)


def test_generalized_cg_matches_recursion():
    """CC via coupling paths == the literal recursion (rotations3d.jl:216-248)."""
    phi = coupling.Invariant()

    def rec(ll, mm, kk):
        N = len(ll)
        if N == 1:
            return 1.0 if (ll[0] == 0 and mm[0] == 0 and kk[0] == 0) else 0.0
        val = 0.0
        jmin = max(abs(ll[-2] - ll[-1]), abs(kk[-2] + kk[-1]), abs(mm[-2] + mm[-1]))
        for j in range(jmin, ll[-2] + ll[-1] + 1):
            cgk = coupling.clebschgordan(ll[-2], kk[-2], ll[-1], kk[-1], j, kk[-2] + kk[-1])
            cgm = coupling.clebschgordan(ll[-2], mm[-2], ll[-1], mm[-1], j, mm[-2] + mm[-1])
            if cgk * cgm != 0:
                val += cgk * cgm * rec(ll[:-2] + (j,), mm[:-2] + (mm[-2] + mm[-1],), kk[:-2] + (kk[-2] + kk[-1],))
        return val

    for ll in [(1, 1), (2, 1, 1), (1, 2, 3), (2, 2, 1, 1)]:
        CC, Mll = coupling.compute_CC("Invariant", ll)
        for ik, kk in enumerate(Mll):
            for im, mm in enumerate(Mll):
                assert CC[ik, im] == pytest.approx(rec(ll, mm, kk), abs=1e-13)


def test_onep_spec_order_and_sizes():
    # init1pspec!: n fastest, then l, then m; stable sort by degree n-1+l
    B1p = ah.RnYlm_1pbasis(maxdeg=12, maxL=6, r0=1.0, rcut=5.0)
    ah.init1pspec(B1p, ah.SimpleSparseBasis(4, 12))
    spec = B1p.get_spec()
    assert len(spec) == 433
    deg = [n - 1 + l for (n, l, m) in spec]
    assert deg == sorted(deg)
    for d in set(deg):
        block = [b for b in spec if b[0] - 1 + b[1] == d]
        keys = [(m, l, n) for (n, l, m) in block]
        assert keys == sorted(keys)


def test_aa_spec_properties():
    basis, T = basis_for(4, 12, 6)
    o = np.bincount(T["orders"])
    assert list(o) == [1, 12, 475, 6761, 28479]
    assert T["orders"][0] == 0
    aspec = T["aspec"]
    deg = aspec[:, 1] - 1 + aspec[:, 2]
    seen = set()
    for i in range(len(T["orders"])):
        idx = tuple(T["iAA2iA"][i, : T["orders"][i]])
        assert list(idx) == sorted(idx, reverse=True)        # descending (pibasis.jl:97)
        assert idx not in seen
        seen.add(idx)
        assert deg[list(idx)].sum() <= 12 if idx else True
        if len(idx) >= 2:
            assert aspec[list(idx), 2].sum() % 2 == 0 and aspec[list(idx), 3].sum() == 0
    # every remaining AA column is used by the A2B map (clean_pibasis!)
    assert set(np.unique(T["a2b"]["cols"])) == set(range(len(T["orders"])))


def test_config1_sizes():
    basis, T = basis_for(3, 6, 4)
    assert len(T["aspec"]) == 43
    assert list(np.bincount(T["orders"])) == [1, 6, 59, 181]
    assert len(basis) == 97


def test_gensparse_order_matches_bruteforce():
    """AA rows in the enumeration order of _gensparse (sparsegrids.jl:69-148)."""
    from acehip.pibasis import gensparse_ordered
    deg = np.array([0, 1, 1, 2, 2, 3])
    adm = lambda vv: sum(deg[v - 1] for v in vv if v) <= 3  # noqa: E731
    spec = gensparse_ordered(3, len(deg), adm, lambda vv: True)
    brute = [vv for vv in itertools.combinations_with_replacement(range(len(deg) + 1), 3) if adm(vv)]
    assert sorted(spec) == sorted(brute)
    assert spec == sorted(spec)          # ascending lexicographic enumeration


def _b1p(maxdeg, L, species=0):
    RnYlm = ah.RnYlm_1pbasis(maxdeg=maxdeg, maxL=L, r0=1.0, rcut=5.0)
    if not species:
        return RnYlm
    cat = ah.Categorical1pBasis(list(range(species)), varsym="mu", idxsym="mu")
    return ah.Product1pBasis((cat,) + RnYlm.bases)


FAST_CASES = [
    ("simple-inv", lambda: ah.SimpleSparseBasis(3, 6), 6, 4, 0, "Invariant", None),
    ("simple-vec", lambda: ah.SimpleSparseBasis(3, 6), 6, 4, 0, "EuclideanVector", None),
    ("simple-species", lambda: ah.SimpleSparseBasis(3, 6), 6, 3, 2, "Invariant", None),
    ("simple-noconst", lambda: ah.SimpleSparseBasis(4, 7), 7, 4, 0, "Invariant", ah.NoConstant()),
    ("simple-order5", lambda: ah.SimpleSparseBasis(5, 6), 6, 3, 0, "Invariant", None),
    ("sparse-weights", lambda: ah.SparseBasis(3, weight={"n": 1.0, "l": 1.5}, default_maxdeg=8), 8, 5, 0,
     "Invariant", None),
    ("sparse-p2", lambda: ah.SparseBasis(3, p=2, default_maxdeg=6), 6, 4, 0, "Invariant", None),
    ("sparse-pinf", lambda: ah.SparseBasis(3, p=float("inf"), default_maxdeg=4), 4, 3, 0, "Invariant", None),
    # order-dependent bounds that are not monotone in the order
    ("sparse-orders", lambda: ah.SparseBasis(4, maxlevels={1: 9, 2: 5, 3: 7, "default": 4}), 9, 4, 0,
     "EuclideanVector", None),
]


@pytest.mark.parametrize("name,mk,maxdeg,L,species,prop,ff", FAST_CASES, ids=[c[0] for c in FAST_CASES])
def test_fast_gensparse_equals_literal_traversal(name, mk, maxdeg, L, species, prop, ff):
    """gensparse_levels + the vectorised filter give the tuples of the literal
    traversal (sparsegrids.jl:85-148 restated in gensparse_ordered) in the same
    order, for SimpleSparseBasis and SparseBasis (basisselectors.jl:175-198)
    with weights, p-norms and order-dependent bounds."""
    from acehip.pibasis import build_pispec
    P = ah.Invariant() if prop == "Invariant" else ah.EuclideanVector(True)
    Bsel = mk()
    fast = build_pispec(_b1p(maxdeg, L, species), Bsel, prop=P, filterfun=ff, fast=True)
    slow = build_pispec(_b1p(maxdeg, L, species), Bsel, prop=P, filterfun=ff, fast=False)
    assert len(fast) > 10
    assert np.array_equal(fast.orders, slow.orders)
    assert np.array_equal(fast.iAA2iA, slow.iAA2iA)


def test_sparse_basis_levels():
    """SparseBasis: weighted 1p degree, p-norm product level, order bounds."""
    from acehip.pibasis import pnorm
    B1p = _b1p(6, 4)
    Bsel = ah.SparseBasis(3, p=2, weight={"n": 2.0, "l": 1.0}, maxlevels={2: 5.0, "default": 6.0})
    b = {"n": 3, "l": 2, "m": 1}
    assert Bsel.level1(b, B1p) == 2.0 * 2 + 1.0 * 2
    assert Bsel.level([b, b], B1p) == pytest.approx(np.sqrt(2) * 6)
    assert Bsel.maxlevel_order(2) == 5.0 and Bsel.maxlevel_order(3) == 6.0 and Bsel.maxlevel == 6.0
    assert pnorm([3.0, 4.0], 2.0) == 5.0 and pnorm([1.0, 2.0, 3.0], float("inf")) == 3.0
    with pytest.raises(ValueError):
        ah.SparseBasis(3)
