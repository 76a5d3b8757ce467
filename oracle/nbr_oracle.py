"""CPU oracle for environment extraction and force assembly — TEST
INFRASTRUCTURE ONLY (imported by ``tests/`` and nothing in the product).

ACE.jl evaluates one environment at a time; its atoms-level callers build the
environments of a configuration and sum the site energies and forces:
``energy(V, at)`` / ``forces(V, at)`` in ``src/testing/testlsq.jl:27-35,42-69``
and ``src/testing/testmodel.jl:64-81`` (forces are ``-dE/dx``, assembled from
pair derivatives that enter both atoms with opposite signs,
``testmodel.jl:49-62``).  The neighbour list itself is JuLIP's, which ACE.jl
v0.12.46 does not depend on and which is absent here, so nothing of its code
is restated: the oracle restates the definition.

  * ``neighbour_pairs``: all (i, j, S') with 0 < |r| < rcut,
    r = (x_j - x_i) + S' . cell, (i, i, 0) excluded, found by brute force over
    every atom pair and every periodic image in a range that covers rcut.
  * ``assemble_forces``: F_k = -dE/dx_k = sum_{m: i(m)=k} g_m -
    sum_{m: j(m)=k} g_m for per-neighbour gradients g_m = +dE_i/dr_m, by
    scatter-add; the virial W = -sum_m g_m r_m^T.

Pinned by: closed forms (dimer; simple-cubic and fcc one-atom cells, whose
neighbour shells 6/12/8 and 12/6/24/12 fix the count of self images; a
cell smaller than rcut in every direction), the symmetry of the list, and
finite differences of a pair-potential total energy through the assembly
(tests/test_nbr_oracle.py).
"""
from __future__ import annotations

import numpy as np


def _plane_spacings(cell):
    cell = np.asarray(cell, dtype=np.float64)
    vol = abs(np.linalg.det(cell))
    a, b, c = cell
    return np.array([vol / np.linalg.norm(np.cross(b, c)), vol / np.linalg.norm(np.cross(c, a)),
                     vol / np.linalg.norm(np.cross(a, b))])


def neighbour_pairs(pos, cell, pbc, rcut):
    """Brute-force neighbour list.

    ``pos`` (N, 3), ``cell`` (3, 3) rows = lattice vectors, ``pbc`` (3,) bool.
    Returns (i, j, S (M, 3) int, r (M, 3)) sorted by (i, j, S).
    """
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    cell = np.asarray(cell, dtype=np.float64)
    pbc = np.asarray(pbc, dtype=bool)
    n = len(pos)
    if n == 0:
        z = np.zeros(0, dtype=np.int64)
        return z, z, np.zeros((0, 3), np.int64), np.zeros((0, 3))
    inv = np.linalg.inv(cell)
    s = pos @ inv
    w = np.where(pbc, np.floor(s), 0.0)
    pw = pos - w @ cell                                  # wrapped positions
    h = _plane_spacings(cell)
    nimg = np.where(pbc, np.ceil(rcut / h).astype(int) + 1, 0)
    rng = [np.arange(-k, k + 1) for k in nimg]
    S = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3)
    out_i, out_j, out_S, out_r = [], [], [], []
    d0 = pw[None, :, :] - pw[:, None, :]                 # [i, j] = p_j - p_i
    for sv, Sv in zip(S @ cell, S):
        d = d0 + sv
        r2 = np.einsum("ijk,ijk->ij", d, d)
        ok = r2 < rcut * rcut
        if not Sv.any():
            np.fill_diagonal(ok, False)
        ii, jj = np.nonzero(ok)
        if len(ii) == 0:
            continue
        out_i.append(ii)
        out_j.append(jj)
        # shift relative to the given (unwrapped) positions
        out_S.append(Sv[None, :] - w[jj].astype(np.int64) + w[ii].astype(np.int64))
        out_r.append(d[ii, jj])
    if not out_i:
        z = np.zeros(0, dtype=np.int64)
        return z, z, np.zeros((0, 3), np.int64), np.zeros((0, 3))
    i = np.concatenate(out_i)
    j = np.concatenate(out_j)
    Sx = np.concatenate(out_S).astype(np.int64)
    # displacement from the given positions: (x_j - x_i) + S' . cell
    r = (pos[j] - pos[i]) + Sx @ cell
    order = np.lexsort((Sx[:, 2], Sx[:, 1], Sx[:, 0], j, i))
    return i[order], j[order], Sx[order], r[order]


def offsets_of(i, natoms):
    """CSR offsets of a list sorted by centre."""
    return np.concatenate([[0], np.cumsum(np.bincount(i, minlength=natoms))]).astype(np.int64)


def assemble_forces(natoms, i, j, g, r=None):
    """F (natoms, 3) = -dE/dx from g (M, 3) = +dE_i/dr_m; with ``r`` also the
    virial W = -sum_m g_m r_m^T (3, 3)."""
    g = np.asarray(g, dtype=np.float64).reshape(-1, 3)
    F = np.zeros((natoms, 3))
    np.add.at(F, i, g)
    np.add.at(F, j, -g)
    if r is None:
        return F
    W = -np.einsum("ma,mb->ab", g, np.asarray(r).reshape(-1, 3))
    return F, W


def fcc_cell(a):
    """Primitive one-atom fcc cell of cubic lattice constant ``a``."""
    return 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


def random_system(natoms, cell, pbc, seed, jitter=0.0, outside=False):
    """Uniform random positions in the cell (fractional), optionally moved by
    whole lattice vectors along periodic axes (``outside``)."""
    rng = np.random.default_rng(seed)
    s = rng.random((natoms, 3))
    if outside:
        s = s + np.where(np.asarray(pbc, bool), rng.integers(-3, 4, (natoms, 3)), 0)
    x = s @ np.asarray(cell, dtype=np.float64)
    if jitter:
        x = x + jitter * rng.standard_normal(x.shape)
    return x
