"""SymmetricBasis and LinearACEModel construction (set-up side, CPU).

Reference semantics (ACE.jl):
  * ``SymmetricBasis(phi, symgrp, pibasis, real)`` (``src/symmbasis.jl:89-162``):
    for every AA function whose factors all have ``m = 0`` compute the
    rotation/permutation coupling rows, scatter each row over the AA columns
    reached by the m-tuples (sorted into AA order, ``:174-177``) and sum
    duplicates (``sparse``, ``:156``); then ``clean_pibasis!`` (``:225-236``)
    drops AA columns that are entirely zero and ``clean_1pbasis!``
    (``src/pibasis.jl:194-207``) renumbers the 1p basis.
  * ``ProductEvaluator`` coefficients ``c~ = A2Bmap^T c``
    (``src/evaluator.jl:30-31,59-67``; ``symmbasis.jl:267-285``).
  * ``LinearACEModel(basis, c; evaluator)`` (``src/linearmodel.jl:36-71``).
"""
from __future__ import annotations

import numpy as np

from . import coupling
from .coupling import EuclideanVector, Invariant
from .pibasis import PIBasis, build_pispec


class A2BMap:
    """Sparse ``[len(B), len(AA)]`` coupling matrix in COO form, rows sorted.

    Invariant: real values ``vals[nnz]``.  EuclideanVector: complex
    3-vectors ``vals[nnz, 3]``.
    """

    def __init__(self, nrows, ncols, rows, cols, vals):
        self.shape = (int(nrows), int(ncols))
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals)
        order = np.lexsort((cols, rows))
        rows, cols, vals = rows[order], cols[order], vals[order]
        # sum duplicates (sparse(I,J,V) semantics: numerical zeros are kept)
        if len(rows):
            key = rows * (self.shape[1] + 1) + cols
            uniq, start = np.unique(key, return_index=True)
            rows_u = rows[start]
            cols_u = cols[start]
            if vals.ndim == 1:
                vals_u = np.add.reduceat(vals, start)
            else:
                vals_u = np.add.reduceat(vals, start, axis=0)
            rows, cols, vals = rows_u, cols_u, vals_u
        self.rows, self.cols, self.vals = rows, cols, vals

    @property
    def nnz(self):
        return len(self.rows)

    def col_norm_sum(self):
        if self.vals.ndim == 1:
            w = np.abs(self.vals)
        else:
            w = np.sqrt(np.sum(np.abs(self.vals) ** 2, axis=1))
        out = np.zeros(self.shape[1])
        np.add.at(out, self.cols, w)
        return out

    def select_cols(self, keep):
        keep = np.asarray(keep, dtype=np.int64)
        newidx = -np.ones(self.shape[1], dtype=np.int64)
        newidx[keep] = np.arange(len(keep))
        m = newidx[self.cols] >= 0
        return A2BMap(self.shape[0], len(keep), self.rows[m], newidx[self.cols[m]],
                      self.vals[m])

    def select_rows(self, keep):
        keep = np.asarray(keep, dtype=np.int64)
        newidx = -np.ones(self.shape[0], dtype=np.int64)
        newidx[keep] = np.arange(len(keep))
        m = newidx[self.rows] >= 0
        return A2BMap(len(keep), self.shape[1], newidx[self.rows[m]], self.cols[m],
                      self.vals[m])

    def csr(self):
        """(rowptr, cols, vals) with rows in order."""
        rowptr = np.zeros(self.shape[0] + 1, dtype=np.int64)
        np.add.at(rowptr, self.rows + 1, 1)
        return np.cumsum(rowptr), self.cols.copy(), self.vals.copy()

    def transpose_matvec(self, c):
        """``c~[col] = sum_row A2B[row, col] * c[row]`` (evaluator.jl:59-60)."""
        c = np.asarray(c, dtype=np.float64)
        if self.vals.ndim == 1:
            out = np.zeros(self.shape[1])
            np.add.at(out, self.cols, self.vals * c[self.rows])
        else:
            out = np.zeros((self.shape[1], 3), dtype=np.complex128)
            np.add.at(out, self.cols, self.vals * c[self.rows][:, None])
        return out


class SymmetricBasis:
    def __init__(self, pibasis: PIBasis, A2Bmap: A2BMap, prop, isreal: bool):
        self.pibasis = pibasis
        self.A2Bmap = A2Bmap
        self.prop = prop
        self.isreal = bool(isreal)

    def __len__(self):
        return self.A2Bmap.shape[0]

    @property
    def basis1p(self):
        return self.pibasis.basis1p

    def clean_pibasis(self):
        """symmbasis.jl:225-236."""
        Inz = np.nonzero(self.A2Bmap.col_norm_sum() > 0.0)[0]
        if len(Inz) < self.A2Bmap.shape[1]:
            self.pibasis.spec = self.pibasis.spec.sparsify(Inz)
            self.A2Bmap = self.A2Bmap.select_cols(Inz)
        self.pibasis.clean_1pbasis()
        return self

    def sparsify(self, keep=None, delete=None):
        """symmbasis.jl:204-216."""
        if (keep is None) == (delete is None):
            raise ValueError("sparsify: provide exactly one of keep / delete")
        if delete is not None:
            keep = sorted(set(range(len(self))) - set(delete))
        self.A2Bmap = self.A2Bmap.select_rows(keep)
        return self.clean_pibasis()


def _nlm_parts(B1p, b):
    d = B1p.as_dict(b)
    nn = tuple((k, v) for k, v in d.items() if k not in ("l", "m"))
    return d["l"], d["m"], nn


def symmetric_basis(prop, B1p, Bsel, isreal=None, filterfun=None) -> SymmetricBasis:
    """``SymmetricBasis(phi, B1p, O3(), Bsel; isreal)`` (symmbasis.jl:74-82,
    89-162)."""
    if isreal is None:
        isreal = prop.isrealB
    spec = build_pispec(B1p, Bsel, prop=prop, filterfun=filterfun)
    pib = PIBasis(B1p, spec, prop.isrealAA)
    Aspec = B1p.get_spec()
    invAspec = {b: i for i, b in enumerate(Aspec)}
    invAA = {spec.get(i): i for i in range(len(spec))}
    syms = B1p.syms
    mpos = syms.index("m")

    rows, cols, vals = [], [], []
    idxB = 0
    vec = prop.numcc > 1
    for iAA in range(len(spec)):
        fac = spec.get(iAA)
        bb = [Aspec[i] for i in fac]
        if any(b[mpos] != 0 for b in bb):
            continue           # not a reference function (symmetrygroups.jl:91)
        if len(bb) == 0:
            # coco_init(phi): a single row with value 1 (properties.jl:139)
            idxB += 1
            rows.append(idxB - 1)
            cols.append(iAA)
            vals.append(1.0 if not vec else np.zeros(3, dtype=np.complex128))
            continue
        parts = [_nlm_parts(B1p, b) for b in bb]
        ll = tuple(p[0] for p in parts)
        nn = tuple(p[2] for p in parts)
        U, Ms = coupling.rpe_basis(prop.name, nn, ll)
        if U.shape[0] == 0:
            continue
        # AA column for every m-tuple
        colidx = np.empty(len(Ms), dtype=np.int64)
        for icol, mm in enumerate(Ms):
            iAs = []
            for t, b in enumerate(bb):
                bl = list(b)
                bl[mpos] = mm[t]
                iAs.append(invAspec[tuple(bl)])
            key = tuple(sorted(iAs, reverse=True))
            if key not in invAA:
                raise RuntimeError(f"bcol_ordered not in AA-spec: {key}")
            colidx[icol] = invAA[key]
        for irow in range(U.shape[0]):
            idxB += 1
            rows.extend([idxB - 1] * len(Ms))
            cols.extend(colidx.tolist())
            if vec:
                vals.extend(list(U[irow]))
            else:
                vals.extend(U[irow].tolist())
    if vec:
        vals_arr = np.asarray(vals, dtype=np.complex128).reshape(-1, 3)
    else:
        vals_arr = np.asarray(vals, dtype=np.float64)
    A2B = A2BMap(idxB, len(spec), rows, cols, vals_arr)
    basis = SymmetricBasis(pib, A2B, prop, isreal)
    basis.clean_pibasis()
    return basis


class LinearACEModel:
    """linearmodel.jl:36-71.  ``evaluator`` is ``"hip"`` (the device
    evaluator of this library) or ``"naive"`` (basis-level evaluation)."""

    def __init__(self, basis: SymmetricBasis, c=None, evaluator: str = "hip"):
        self.basis = basis
        self.c = np.zeros(len(basis)) if c is None else np.array(c, dtype=np.float64)
        if self.c.shape != (len(basis),):
            raise ValueError("LinearACEModel: len(c) must equal len(basis)")
        if evaluator not in ("hip", "naive", "standard"):
            raise ValueError("unknown evaluator")
        self.evaluator_kind = evaluator
        self._ev = None

    @property
    def ctilde(self):
        return self.basis.A2Bmap.transpose_matvec(self.c)

    def nparams(self):
        return len(self.c)

    def params(self):
        return self.c.copy()

    def set_params(self, c):
        c = np.asarray(c, dtype=np.float64)
        if c.shape != self.c.shape:
            raise ValueError("set_params: wrong length")
        self.c[:] = c
        if self._ev is not None:
            self._ev.set_params(self.c)
        return self


def invariant():
    return Invariant()


def euclidean_vector(isreal=True):
    return EuclideanVector(isreal)
