"""Environment extraction and force assembly (gpu): neighbour lists against
the brute-force oracle, the reverse index, position updates, force/virial
assembly, and whole-configuration ACE energies/forces against the oracle
pipeline and finite differences."""
import ctypes as C

import numpy as np
import pytest
import torch

import ace_oracle as O
import nbr_oracle as N
from acehip import _lib, synth
from acehip.atoms import ACECalculator, Atoms, NeighbourList
from acehip.evaluator import HIPEvaluator
from helpers import basis_for, rel_err

pytestmark = pytest.mark.gpu

TRI = np.array([[4.0, 0.0, 0.0], [1.3, 3.7, 0.0], [0.6, -0.9, 4.4]])

CASES = {
    "cubic_pbc": (8.0 * np.eye(3), [1, 1, 1], 150, 3.5, False),
    "triclinic_unwrapped": (2.0 * TRI, [1, 1, 1], 200, 3.0, True),
    "slab_unwrapped": (TRI, [1, 1, 0], 120, 2.5, True),
    "open_cluster": (12.0 * np.eye(3), [0, 0, 0], 80, 4.0, False),
    "one_atom_fcc_self_images": (N.fcc_cell(3.6), [1, 1, 1], 1, 5.0, False),
    "two_atoms_thin_cell": (np.diag([2.0, 2.5, 3.0]), [1, 1, 1], 2, 4.5, True),
    "single_open": (np.eye(3), [0, 0, 0], 1, 2.0, False),
    "empty": (np.eye(3), [1, 1, 1], 0, 2.0, False),
}


def _gpu_list(x, cell, pbc, rcut, species=None):
    nl = NeighbourList(0)
    nl.build(x, cell, pbc, rcut, species)
    torch.cuda.synchronize()
    return nl


def _entries(nl):
    off = nl.offsets.cpu().numpy()
    i = np.repeat(np.arange(nl.natoms), np.diff(off))
    return off, i, nl.nbr_j.cpu().numpy(), nl.shift.cpu().numpy().T, nl.xyz.cpu().numpy().T


@pytest.mark.parametrize("name", list(CASES))
def test_list_matches_brute_force(cuda, name):
    cell, pbc, n, rcut, outside = CASES[name]
    print(f"[gpu] neighbour list {name}: {n} atoms, rcut {rcut}", flush=True)
    x = N.random_system(n, cell, pbc, seed=n + 7, outside=outside)
    nl = _gpu_list(x, cell, pbc, rcut)
    i0, j0, S0, r0 = N.neighbour_pairs(x, cell, pbc, rcut)
    off, i, j, S, r = _entries(nl)
    assert nl.nnbr == len(i0) and off[0] == 0 and off[-1] == len(i0)
    assert np.array_equal(off, N.offsets_of(i0, n))
    order = np.lexsort((S[:, 2], S[:, 1], S[:, 0], j, i))
    assert np.array_equal(j[order], j0) and np.array_equal(S[order], S0)
    assert np.max(np.abs(r[order] - r0), initial=0.0) < 1e-12
    # reverse index: an involution onto the mirrored entry, vectors exactly negated
    rev = nl.rev.cpu().numpy()
    if len(rev):
        assert np.array_equal(rev[rev], np.arange(len(rev)))
        assert np.array_equal(j[rev], i) and np.array_equal(S[rev], -S)
        assert np.array_equal(r[rev], -r)


def test_species_capacity_and_errors(cuda):
    print("[gpu] neighbour species, capacity protocol, argument errors", flush=True)
    x = N.random_system(60, TRI, [1, 1, 1], seed=11)
    z = (np.arange(60) % 3).astype(np.int8)
    nl = _gpu_list(x, TRI, [1, 1, 1], 3.0, z)
    _, _, j, _, _ = _entries(nl)
    assert np.array_equal(nl.species.cpu().numpy(), z[j])
    # too small a capacity: ACE_ERR_CAPACITY and the size needed, nothing written
    lib = _lib.load()
    h = C.c_void_p()
    _lib.check(lib.ace_nbrlist_create(0, C.byref(h)))
    pos = torch.as_tensor(x.T.copy(), device="cuda:0")
    cell = np.ascontiguousarray(TRI)
    pbc = np.ones(3, np.int32)
    off = torch.full((61,), -7, dtype=torch.int32, device="cuda:0")
    small = torch.empty(30, dtype=torch.int32, device="cuda:0")
    xyz = torch.empty(90, dtype=torch.float64, device="cuda:0")
    need = C.c_int64(0)
    rc = lib.ace_nbrlist_build(h, 60, C.c_void_p(pos.data_ptr()), cell.ctypes.data_as(C.POINTER(C.c_double)),
                               pbc.ctypes.data_as(C.POINTER(C.c_int32)), 3.0, None, 10, C.c_void_p(off.data_ptr()),
                               C.c_void_p(small.data_ptr()), C.c_void_p(small.data_ptr()),
                               C.c_void_p(xyz.data_ptr()), None, None, C.byref(need), None)
    torch.cuda.synchronize()
    assert rc == _lib.ERR_CAPACITY and need.value == nl.nnbr
    assert b"capacity" in lib.ace_last_error()
    assert torch.all(off == -7)
    bad = np.zeros((3, 3))
    rc = lib.ace_nbrlist_build(h, 60, C.c_void_p(pos.data_ptr()), bad.ctypes.data_as(C.POINTER(C.c_double)),
                               pbc.ctypes.data_as(C.POINTER(C.c_int32)), 3.0, None, 10, C.c_void_p(off.data_ptr()),
                               C.c_void_p(small.data_ptr()), C.c_void_p(small.data_ptr()),
                               C.c_void_p(xyz.data_ptr()), None, None, C.byref(need), None)
    assert rc < 0 and b"singular" in lib.ace_last_error()
    lib.ace_nbrlist_destroy(h)


def test_update_follows_positions(cuda):
    print("[gpu] neighbour-list update after displacements", flush=True)
    cell, pbc = 2.0 * TRI, [1, 1, 1]
    x = N.random_system(150, cell, pbc, seed=21, outside=True)
    nl = _gpu_list(x, cell, pbc, 3.2)
    off, i, j, S, r = _entries(nl)
    x2 = x + 0.05 * np.random.default_rng(2).standard_normal(x.shape)
    nl.update(x2)
    torch.cuda.synchronize()
    r2 = nl.xyz.cpu().numpy().T
    assert np.max(np.abs(r2 - ((x2[j] - x2[i]) + S @ cell))) < 1e-12
    # update at the build positions reproduces the build exactly
    nl.update(x)
    torch.cuda.synchronize()
    assert np.array_equal(nl.xyz.cpu().numpy().T, r)


def test_force_and_virial_assembly(cuda):
    print("[gpu] force and virial assembly from per-neighbour gradients", flush=True)
    cell, pbc = 2.0 * TRI, [1, 1, 1]
    x = N.random_system(180, cell, pbc, seed=31, outside=True)
    nl = _gpu_list(x, cell, pbc, 3.5)
    off, i, j, S, r = _entries(nl)
    g = np.random.default_rng(5).standard_normal((nl.nnbr, 3))   # any per-neighbour gradients
    F, W = nl.forces(torch.as_tensor(g.T.copy(), device="cuda:0"), virial=True)
    Fr, Wr = N.assemble_forces(nl.natoms, i, j, g, r)
    assert rel_err(F.cpu().numpy().T, Fr) < 1e-13
    assert rel_err(W.cpu().numpy(), Wr) < 1e-13
    # deterministic: a second call is bit-identical
    F2, W2 = nl.forces(torch.as_tensor(g.T.copy(), device="cuda:0"), virial=True)
    assert torch.equal(F, F2) and torch.equal(W, W2)


def _system(ncell=2, a=3.6, jitter=0.12, seed=3):
    base = 0.5 * a * np.array([[0, 0, 0], [0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=np.float64)
    grid = np.array([[p, q, s] for p in range(ncell) for q in range(ncell) for s in range(ncell)], dtype=np.float64)
    x = (grid[:, None, :] * a + base[None]).reshape(-1, 3)
    x = x + jitter * np.random.default_rng(seed).standard_normal(x.shape)
    cell = ncell * a * np.eye(3) + np.array([[0, 0, 0], [0.3, 0, 0], [0, -0.2, 0]])
    return x, cell


def _oracle_total(T, c, x, cell, pbc, rcut, z=None):
    i, j, S, r = N.neighbour_pairs(x, cell, pbc, rcut)
    off = N.offsets_of(i, len(x)).astype(np.int32)
    sp = None if z is None else z[j]
    E = O.energy(T, c, off, r, sp)
    g = O.grad_config(T, c, off, r, sp)
    F, W = N.assemble_forces(len(x), i, j, g, r)
    return E.sum(), F, W


@pytest.mark.parametrize("species", [0, 2])
def test_ace_configuration_energy_forces(cuda, species):
    print(f"[gpu] ACE on a periodic fcc configuration (species={species})", flush=True)
    basis, T = basis_for(3, 6, 4, species=species)
    c = synth.coefficients(len(basis), 4)
    ev = HIPEvaluator(basis, c, device=0)
    x, cell = _system()
    z = (np.arange(len(x)) % 2).astype(np.int8) if species else None
    at = Atoms(x, cell, (True, True, True), z)
    calc = ACECalculator(ev)
    assert calc.rcut == 5.0
    E, F, W = calc.energy_forces(at, virial=True)
    Er, Fr, Wr = _oracle_total(T, c, x, cell, [1, 1, 1], 5.0, z)
    assert abs(E - Er) <= 1e-12 * abs(Er)
    assert rel_err(F, Fr) < 1e-10
    assert rel_err(W, Wr) < 1e-10
    assert np.max(np.abs(F.sum(0))) < 1e-11 * np.max(np.abs(F))
    # forces are -dE/dx of the GPU total energy (central differences)
    h = 1e-5
    for k, a in [(0, 0), (5, 1), (17, 2), (31, 0)]:
        xp = x.copy()
        xp[k, a] += h
        xm = x.copy()
        xm[k, a] -= h
        fd = -(calc.energy(Atoms(xp, cell, at.pbc, z)) - calc.energy(Atoms(xm, cell, at.pbc, z))) / (2 * h)
        assert abs(fd - F[k, a]) < 1e-6 * np.max(np.abs(F)), (k, a, fd, F[k, a])


def test_skin_reuses_the_list(cuda):
    print("[gpu] calculator with a skin: list reuse and rebuild", flush=True)
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 4)
    ev = HIPEvaluator(basis, c, device=0)
    x, cell = _system(seed=8)
    skin = ACECalculator(ev, skin=0.6)
    plain = ACECalculator(ev)
    rng = np.random.default_rng(1)
    for step in range(4):
        xs = x + 0.02 * step * rng.standard_normal(x.shape)
        at = Atoms(xs, cell)
        E1, F1 = skin.energy_forces(at)
        E2, F2 = plain.energy_forces(at)
        assert abs(E1 - E2) <= 1e-12 * abs(E2) and rel_err(F1, F2) < 1e-11
    assert skin.rebuilds == 1 and plain.rebuilds == 4
    at = Atoms(x + 0.5, cell)          # rigid shift beyond skin / 2: rebuild
    E1, _ = skin.energy_forces(at)
    assert skin.rebuilds == 2
    assert abs(E1 - plain.energy(at)) <= 1e-12 * abs(E1)


def test_bad_positions_are_rejected(cuda):
    print("[gpu] neighbour list: non-finite and far-away positions", flush=True)
    from acehip._lib import AceError
    x = N.random_system(20, TRI, [1, 1, 1], seed=4)
    nl = NeighbourList(0)
    bad = x.copy()
    bad[7, 1] = np.nan
    with pytest.raises(AceError, match="finite"):
        nl.build(bad, TRI, [1, 1, 1], 3.0)
    far = x.copy()
    far[3] += 5e6 * TRI[0]
    with pytest.raises(AceError, match="too far"):
        nl.build(far, TRI, [1, 1, 1], 3.0)
    nl.build(x, TRI, [1, 1, 1], 3.0)         # the handle stays usable
    torch.cuda.synchronize()
    assert nl.nnbr == len(N.neighbour_pairs(x, TRI, [1, 1, 1], 3.0)[0])
