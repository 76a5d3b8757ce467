"""One-particle basis components and the product 1p basis (set-up side).

A 1p basis function is specified by a tuple of index values in the product
basis' symbol order, e.g. ``(n, l, m)`` for ``Rn x Ylm`` or ``(mu, n, l, m)``
for ``Categorical x Rn x Ylm``.

Reference semantics:
  * ``Rn1pBasis`` spec ``n = 1..N`` with degree ``n-1``
    (ACE.jl ``src/b1pcomponents/Rn.jl:17-29``).
  * ``Ylm1pBasis`` spec ``(l, m)`` in ``idx2lm`` order with degree ``l``
    (``src/b1pcomponents/Ylm.jl:47-55``,
    ``src/polynomials/sphericalharmonics.jl:102-108``).
  * ``Categorical1pBasis`` one-hot with degree 0
    (``src/discrete1pbasis.jl:78-131``).
  * ``Product1pBasis`` (``src/product_1pbasis.jl:5-8,283-315``) and
    ``init1pspec!`` (``src/sparsegrids.jl:9-31``): enumerate the Cartesian
    product of index ranges with the first symbol varying fastest, keep the
    admissible tuples of level at most ``maxlevel``, stable-sort by level.
"""
from __future__ import annotations

from . import orthpolys
from .orthpolys import PolyTransform, discrete_jacobi


class Rn1pBasis:
    kind = "Rn"

    def __init__(self, J: orthpolys.OrthPolyBasis, trans: PolyTransform,
                 nsym: str = "n", label: str = "Rn"):
        self.J = J
        self.trans = trans
        self.nsym = nsym
        self.label = label
        rl = trans.inverse(J.tl)
        rr = trans.inverse(J.tr)
        self.rin, self.rcut = min(rl, rr), max(rl, rr)

    @property
    def syms(self):
        return (self.nsym,)

    def __len__(self):
        return len(self.J)

    def indexrange(self):
        return {self.nsym: list(range(1, len(self.J) + 1))}

    def isadmissible(self, b):
        return 1 <= b[self.nsym] <= len(self.J)

    def degree(self, b, weight=None):
        """n - 1 (Rn.jl:20); with a weight dict, weight[n-symbol] (n - 1)
        (b1pcomponent.jl:145-147)."""
        d = b[self.nsym] - 1
        return d if weight is None else weight[self.nsym] * d


class Ylm1pBasis:
    kind = "Ylm"

    def __init__(self, L: int, lsym: str = "l", msym: str = "m", label: str = "Ylm"):
        self.L = int(L)
        self.lsym, self.msym = lsym, msym
        self.label = label

    @property
    def syms(self):
        return (self.lsym, self.msym)

    def __len__(self):
        return (self.L + 1) ** 2

    def indexrange(self):
        return {self.lsym: list(range(0, self.L + 1)),
                self.msym: list(range(-self.L, self.L + 1))}

    def isadmissible(self, b):
        l, m = b[self.lsym], b[self.msym]
        return 0 <= l <= self.L and abs(m) <= l

    def degree(self, b, weight=None):
        """l (Ylm.jl:24); with a weight dict, weight[l-symbol] l."""
        d = b[self.lsym]
        return d if weight is None else weight[self.lsym] * d


class Categorical1pBasis:
    kind = "Categorical"

    def __init__(self, categories, varsym: str = "mu", idxsym: str = "mu",
                 label: str | None = None):
        self.categories = list(categories)
        if len(set(self.categories)) != len(self.categories):
            raise ValueError("Categorical1pBasis: categories must be unique")
        self.varsym, self.idxsym = varsym, idxsym
        self.label = label if label is not None else f"C{idxsym}"

    @property
    def syms(self):
        return (self.idxsym,)

    def __len__(self):
        return len(self.categories)

    def indexrange(self):
        return {self.idxsym: list(self.categories)}

    def isadmissible(self, b):
        return b[self.idxsym] in self.categories

    def degree(self, b, weight=None):
        return 0  # discrete1pbasis.jl:123, with or without weights

    def val2i(self, val):
        """discrete1pbasis.jl:33-40: unknown categories are an error."""
        try:
            return self.categories.index(val)
        except ValueError:
            raise KeyError(f"val = {val} not found in this list") from None


class Product1pBasis:
    """Product of 1p components; ``spec`` lists the 1p basis functions as
    tuples over ``self.syms``."""

    def __init__(self, bases):
        self.bases = tuple(bases)
        syms = []
        for B in self.bases:
            for s in B.syms:
                if s not in syms:
                    syms.append(s)
        self.syms = tuple(syms)
        self.spec: list[tuple] = []

    def __len__(self):
        return len(self.spec)

    def __getitem__(self, label):
        hits = [B for B in self.bases if B.label == label]
        if len(hits) != 1:
            raise KeyError(label)
        return hits[0]

    def component(self, kind):
        for B in self.bases:
            if B.kind == kind:
                return B
        return None

    def as_dict(self, b):
        return dict(zip(self.syms, b))

    def indexrange(self):
        rg = {s: [] for s in self.syms}
        for B in self.bases:
            for s, vals in B.indexrange().items():
                for v in vals:
                    if v not in rg[s]:
                        rg[s].append(v)
        # the m range follows the maximal l (product_1pbasis.jl:294-300)
        if "m" in rg and "l" in rg:
            maxl = max(rg["l"])
            rg["m"] = list(range(-maxl, maxl + 1))
        return rg

    def isadmissible(self, b: dict):
        return all(B.isadmissible(b) for B in self.bases)

    def degree(self, b: dict, weight=None):
        """product_1pbasis.jl:328-331: the sum over the components."""
        if weight is None:
            return sum(B.degree(b) for B in self.bases)
        return sum(B.degree(b, weight) for B in self.bases)

    def set_spec(self, spec):
        self.spec = [tuple(b) for b in spec]
        return self

    def get_spec(self):
        return list(self.spec)

    def sparsify(self, keep):
        """product_1pbasis.jl:345-374: keep listed functions, preserving order;
        returns the old -> new index map (-1 for removed)."""
        keep = set(keep)
        new_spec, new_inds = [], []
        for b in self.spec:
            if b in keep:
                new_spec.append(b)
                new_inds.append(len(new_spec) - 1)
            else:
                new_inds.append(-1)
        self.spec = new_spec
        return new_inds


def init1pspec(B1p: Product1pBasis, Bsel) -> Product1pBasis:
    """sparsegrids.jl:9-31."""
    syms = B1p.syms
    rgs = B1p.indexrange()
    lens = [len(rgs[s]) for s in syms]
    maxlev = Bsel.maxlevel
    spec = []
    # CartesianIndices: first index fastest
    total = 1
    for n in lens:
        total *= n
    for flat in range(total):
        idx = []
        rem = flat
        for n in lens:
            idx.append(rem % n)
            rem //= n
        b = tuple(rgs[s][i] for s, i in zip(syms, idx))
        bd = dict(zip(syms, b))
        if B1p.isadmissible(bd):
            if not Bsel.filter1(bd, B1p):
                continue
            if Bsel.level1(bd, B1p) <= maxlev:
                spec.append(b)
    spec.sort(key=lambda b: Bsel.level1(dict(zip(syms, b)), B1p))  # stable
    return B1p.set_spec(spec)


def Rn_basis(r0=1.0, trans=None, maxdeg=6, rcut=2.5, rin=None, pcut=2, pin=0,
             nsym="n", label=None):
    """utils/utils.jl:28-47."""
    if trans is None:
        trans = PolyTransform(2, r0)
    if rin is None:
        rin = 0.5 * r0
    J = discrete_jacobi(maxdeg, pcut=pcut, xcut=rcut, pin=pin, xin=rin, trans=trans)
    return Rn1pBasis(J, trans, nsym=nsym, label=label or f"R{nsym}")


def RnYlm_1pbasis(maxdeg=6, maxL=None, Bsel=None, **kwargs):
    """utils/utils.jl:54-64."""
    if maxL is None:
        maxL = maxdeg
    Rn = Rn_basis(maxdeg=maxdeg, **kwargs)
    Ylm = Ylm1pBasis(maxL)
    B1p = Product1pBasis((Rn, Ylm))
    if Bsel is not None:
        init1pspec(B1p, Bsel)
    return B1p
