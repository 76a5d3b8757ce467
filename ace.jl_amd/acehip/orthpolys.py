"""Radial basis set-up: distance transform and envelope orthogonal polynomials.

Set-up only (CPU, once per model).  The recursion coefficients built here are
copied into the device basis pack; evaluation happens in the HIP kernels.

Reference semantics:
  * ``polytransform(p, r0)`` = ``r -> ((1+r0)/(1+r))^p``
    (ACE.jl ``src/transforms/distancetransforms.jl:73``).
  * ``discrete_jacobi`` (``src/polynomials/orthpolys.jl:340-349``): a uniform
    midpoint measure of ``Nquad = max(300, 3N)`` points on ``[tl, tr]``.
  * ``OrthPolyBasis(N, pcut, tcut, pin, tin, tdf)`` discrete Stieltjes
    procedure (``orthpolys.jl:153-218``) for polynomials with the envelope
    ``(t-tl)^pl (t-tr)^pr`` that are orthonormal w.r.t. the measure.
"""
from __future__ import annotations

import numpy as np


class PolyTransform:
    """``t(r) = ((1 + r0) / (1 + r))^p`` (distancetransforms.jl:73)."""

    def __init__(self, p: int = 2, r0: float = 1.0):
        self.p = int(p)
        self.r0 = float(r0)

    def __call__(self, r):
        return ((1.0 + self.r0) / (1.0 + r)) ** self.p

    def inverse(self, t: float) -> float:
        """Solve ``t(r) = t`` (closed form; the reference uses a root finder,
        distancetransforms.jl:89-96)."""
        return (1.0 + self.r0) / t ** (1.0 / self.p) - 1.0

    def to_dict(self):
        return {"__id__": "ACE_Lambda",
                "exstr": f"r -> ((1+{self.r0})/(1+r))^{self.p}"}


def fcut(pl, tl, pr, tr, t):
    """Envelope ``(t-tl)^pl (t-tr)^pr``, zero outside ``[tl, tr]`` where the
    corresponding power is positive (orthpolys.jl:32,41-46)."""
    t = np.asarray(t, dtype=np.float64)
    out = (t - tl) ** pl * (t - tr) ** pr
    mask = np.zeros(t.shape, dtype=bool)
    if pl > 0:
        mask |= t < tl
    if pr > 0:
        mask |= t > tr
    return np.where(mask, 0.0, out)


class OrthPolyBasis:
    """Recursion coefficients of an envelope orthonormal polynomial basis.

    ``P1 = A1 fcut(t)``, ``P2 = (A2 t + B2) P1``,
    ``Pn = (An t + Bn) P(n-1) + Cn P(n-2)`` (orthpolys.jl:65-99,253-263).
    """

    def __init__(self, pl, tl, pr, tr, A, B, C, tdf, ww):
        self.pl, self.tl, self.pr, self.tr = int(pl), float(tl), int(pr), float(tr)
        self.A = np.asarray(A, dtype=np.float64)
        self.B = np.asarray(B, dtype=np.float64)
        self.C = np.asarray(C, dtype=np.float64)
        self.tdf = np.asarray(tdf, dtype=np.float64)
        self.ww = np.asarray(ww, dtype=np.float64)

    def __len__(self):
        return len(self.A)

    @classmethod
    def stieltjes(cls, N: int, pcut: int, tcut: float, pin: int, tin: float,
                  tdf, ww=None) -> "OrthPolyBasis":
        """Discrete Stieltjes procedure, orthpolys.jl:153-218."""
        assert pcut >= 0 and pin >= 0 and N > 0
        tdf = np.asarray(tdf, dtype=np.float64)
        if ww is None:
            ww = np.ones_like(tdf)
        if tcut < tin:
            tl, tr, pl, pr = tcut, tin, pcut, pin
        else:
            tl, tr, pl, pr = tin, tcut, pin, pcut
        A = np.zeros(N)
        B = np.zeros(N)
        C = np.zeros(N)
        ww = np.asarray(ww, dtype=np.float64)
        ww = ww / ww.sum()

        def dotw(f1, f2):
            return float(np.dot(f1, ww * f2))

        _J1 = fcut(pl, tl, pr, tr, tdf)
        a = np.sqrt(dotw(_J1, _J1))
        A[0] = 1.0 / a
        J1 = A[0] * _J1
        if N > 1:
            b = dotw(tdf * J1, J1)
            _J2 = (tdf - b) * J1
            a = np.sqrt(dotw(_J2, _J2))
            A[1] = 1.0 / a
            B[1] = -b / a
            J2 = (A[1] * tdf + B[1]) * J1
            Jprev, Jpprev = J2, J1
        for n in range(2, N):
            b = dotw(tdf * Jprev, Jprev)
            c = dotw(tdf * Jprev, Jpprev)
            _J = (tdf - b) * Jprev - c * Jpprev
            a = np.sqrt(dotw(_J, _J))
            A[n] = 1.0 / a
            B[n] = -b / a
            C[n] = -c / a
            Jprev, Jpprev = _J / a, Jprev
        return cls(pl, tl, pr, tr, A, B, C, tdf, ww)

    def to_dict(self):
        """ACE.jl ``write_dict`` layout (orthpolys.jl:119-131)."""
        return {"__id__": "ACE_OrthPolyBasis", "T": {"__id__": "Float64"},
                "pr": self.pr, "tr": self.tr, "pl": self.pl, "tl": self.tl,
                "A": self.A.tolist(), "B": self.B.tolist(), "C": self.C.tolist(),
                "tdf": self.tdf.tolist(), "ww": self.ww.tolist()}

    @classmethod
    def from_dict(cls, D):
        return cls(D["pl"], D["tl"], D["pr"], D["tr"], D["A"], D["B"], D["C"],
                   D["tdf"], D["ww"])


def discrete_jacobi(N: int, pcut: int = 0, xcut: float = 1.0, pin: int = 0,
                    xin: float = -1.0, Nquad: int | None = None, trans=None):
    """orthpolys.jl:340-349: midpoint grid of ``Nquad`` points on the
    transformed interval and the Stieltjes basis on it."""
    if Nquad is None:
        Nquad = max(300, 3 * N)
    if trans is None:
        trans = lambda x: x  # noqa: E731
    tcut = float(trans(xcut))
    tin = float(trans(xin))
    tl, tr = min(tin, tcut), max(tin, tcut)
    dt = (tr - tl) / Nquad
    tdf = np.linspace(tl + dt / 2, tr - dt / 2, Nquad)
    return OrthPolyBasis.stieltjes(N, pcut, tcut, pin, tin, tdf)
