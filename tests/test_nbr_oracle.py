"""Neighbour-list / force-assembly oracle pinned by closed forms and finite
differences (CPU)."""
import numpy as np
import pytest

import nbr_oracle as N


def _count_per_atom(i, n):
    return np.bincount(i, minlength=n)


def test_dimer_open():
    cell = 20.0 * np.eye(3)
    x = np.array([[1.0, 2.0, 3.0], [2.5, 2.0, 3.5]])
    i, j, S, r = N.neighbour_pairs(x, cell, [False] * 3, 2.0)
    assert list(i) == [0, 1] and list(j) == [1, 0]
    assert np.all(S == 0)
    assert np.allclose(r[0], x[1] - x[0]) and np.allclose(r[1], x[0] - x[1])
    i, _, _, _ = N.neighbour_pairs(x, cell, [False] * 3, 1.5)   # |d| = 1.58 > 1.5
    assert len(i) == 0


@pytest.mark.parametrize("rcut,count", [(1.5, 6 + 12), (1.8, 6 + 12 + 8), (2.05, 6 + 12 + 8 + 6)])
def test_simple_cubic_shells(rcut, count):
    i, j, S, r = N.neighbour_pairs(np.zeros((1, 3)), np.eye(3), [True] * 3, rcut)
    assert len(i) == count and np.all(j == 0)
    d = np.linalg.norm(r, axis=1)
    assert np.all(d > 0.99) and np.all(d < rcut)


@pytest.mark.parametrize("rcut,count", [(0.8, 12), (1.1, 18), (1.3, 42), (1.45, 54)])
def test_fcc_shells(rcut, count):
    # fcc shells at a/sqrt2 (12), a (6), a sqrt(3/2) (24), a sqrt2 (12)
    cell = N.fcc_cell(1.0)
    x = np.array([[0.3, -0.2, 0.7]])
    i, _, _, r = N.neighbour_pairs(x, cell, [True] * 3, rcut)
    assert len(i) == count
    # translation by a lattice vector changes shifts, not vectors
    i2, _, S2, r2 = N.neighbour_pairs(x + cell[0] * 3, cell, [True] * 3, rcut)
    key = lambda rr: np.round(rr, 9).tolist()
    assert sorted(key(r)) == sorted(key(r2))


def test_list_is_symmetric_triclinic_unwrapped():
    cell = np.array([[4.0, 0.0, 0.0], [1.3, 3.7, 0.0], [0.6, -0.9, 4.4]])
    pbc = [True, True, False]
    x = N.random_system(40, cell, pbc, seed=3, outside=True)
    i, j, S, r = N.neighbour_pairs(x, cell, pbc, 3.1)
    fwd = {(a, b, *s): rr for a, b, s, rr in zip(i, j, S.tolist(), r)}
    for (a, b, s0, s1, s2), rr in fwd.items():
        back = fwd[(b, a, -s0, -s1, -s2)]
        assert np.allclose(back, -rr, atol=1e-12)
    assert np.all(S[:, 2] == 0)   # open axis: no images
    d = np.linalg.norm(r, axis=1)
    assert np.all(d < 3.1) and np.all(d > 0)


def _site_energies(i, r, natoms, kind):
    rc = 3.0
    d = np.linalg.norm(r, axis=1)
    phi = np.where(d < rc, (d - rc) ** 2 * np.exp(-d), 0.0)
    dphi = np.where(d < rc, (2 * (d - rc) - (d - rc) ** 2) * np.exp(-d), 0.0)
    s = np.bincount(i, weights=phi, minlength=natoms)
    if kind == "pair":
        E = s
        g = dphi[:, None] * r / d[:, None]
    else:   # non-additive site energy: g_m depends on the whole environment
        E = s ** 2
        g = (2 * s[i] * dphi)[:, None] * r / d[:, None]
    return E, g


def _total(x, cell, pbc, kind):
    i, j, S, r = N.neighbour_pairs(x, cell, pbc, 3.0)
    E, _ = _site_energies(i, r, len(x), kind)
    return E.sum()


@pytest.mark.parametrize("kind", ["pair", "squared"])
def test_assembly_is_minus_gradient_of_total_energy(kind):
    cell = np.array([[5.0, 0.0, 0.0], [0.8, 4.6, 0.0], [0.0, 0.5, 4.9]])
    pbc = [True, True, True]
    x = N.random_system(24, cell, pbc, seed=5)
    i, j, S, r = N.neighbour_pairs(x, cell, pbc, 3.0)
    _, g = _site_energies(i, r, len(x), kind)
    F, W = N.assemble_forces(len(x), i, j, g, r)
    h = 1e-6
    Ffd = np.zeros_like(x)
    for k in range(len(x)):
        for a in range(3):
            xp = x.copy()
            xp[k, a] += h
            xm = x.copy()
            xm[k, a] -= h
            Ffd[k, a] = -(_total(xp, cell, pbc, kind) - _total(xm, cell, pbc, kind)) / (2 * h)
    assert np.max(np.abs(F - Ffd)) < 1e-6 * np.max(np.abs(F))
    assert np.max(np.abs(F.sum(0))) < 1e-10 * np.max(np.abs(F))
    # virial: W = -dE/d(strain) under x -> (I + e) x, cell -> cell (I + e)^T
    Wfd = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            E = []
            for sgn in (1, -1):
                eps = np.eye(3)
                eps[a, b] += sgn * h
                E.append(_total(x @ eps.T, cell @ eps.T, pbc, kind))
            Wfd[a, b] = -(E[0] - E[1]) / (2 * h)
    assert np.max(np.abs(W - Wfd)) < 1e-6 * np.max(np.abs(W))
