"""Training-side calls (gpu): adjoint_EVAL_D, per-basis-function forces of
configurations and the least-squares fit, against the oracle and against the
model-level energy/forces."""
import numpy as np
import pytest
import torch

import ace_oracle as O
import nbr_oracle as N
from acehip import _lib, synth
from acehip.atoms import ACECalculator, Atoms
from acehip.evaluator import EnvBatch, HIPEvaluator
from acehip.fit import get_V0, lsq, training_set
from helpers import basis_for, random_batch, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("species", [0, 2])
def test_adjoint_eval_d_matches_oracle(cuda, species):
    print(f"[gpu] adjoint_EVAL_D vs oracle (species={species})", flush=True)
    basis, T = basis_for(3, 6, 4, species=species)
    c = synth.coefficients(len(basis), 6)
    ev = HIPEvaluator(basis, c, device=0)
    off, R, sp = random_batch(24, 40 + species, nspecies=species)
    off = off.copy()
    # one empty environment in the middle
    off = np.concatenate([off[:5], [off[5]], off[5:]]).astype(np.int32)
    w = np.random.default_rng(3).standard_normal(R.shape)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    got = ev.adjoint_EVAL_D(b, torch.as_tensor(w.T.copy())).cpu().numpy()
    ref = np.stack([O.adjoint_EVAL_D(T, c, off[e:e + 2] - off[e], R[off[e]:off[e + 1]], w[off[e]:off[e + 1]],
                                     None if sp is None else sp[off[e]:off[e + 1]])
                    for e in range(len(off) - 1)])
    assert rel_err(got, ref) < 1e-10
    assert np.all(got[5] == 0.0)
    # equals the contraction of the full dB with w
    dB = ev.basis_dB(b).cpu().numpy()
    con = np.stack([np.einsum("jbd,jd->b", dB[off[e]:off[e + 1]], w[off[e]:off[e + 1]])
                    for e in range(len(off) - 1)])
    assert rel_err(got, con) < 1e-12


def test_adjoint_eval_d_rejects_vector_property(cuda):
    basis, T = basis_for(3, 6, 3, prop="EuclideanVector")
    ev = HIPEvaluator(basis, synth.coefficients(len(basis), 1), device=0)
    off, R, _ = random_batch(3, 2)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    with pytest.raises(_lib.AceError, match="Invariant"):
        ev.adjoint_EVAL_D(b, torch.zeros(3, b.nnbr))


def _config(seed, ncell=2, a=3.6):
    x, cell = synth.fcc_crystal((ncell, ncell, ncell), a=a, sigma=0.15, seed=seed)
    cell = cell + np.array([[0, 0, 0], [0.25, 0, 0], [0, -0.15, 0]])
    return Atoms(x, cell, (True, True, True))


def test_basis_energy_forces(cuda):
    print("[gpu] per-basis-function energies and forces of a configuration", flush=True)
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 7)
    ev = HIPEvaluator(basis, c, device=0)
    calc = ACECalculator(ev)
    at = _config(1)
    EB, FB = calc.basis_energy_forces(at)
    E, F = calc.energy_forces(at)
    ct = torch.as_tensor(c, device="cuda:0")
    # linearity: the model is c . basis
    assert abs(float(ct @ EB) - E) <= 1e-12 * abs(E)
    assert rel_err(torch.einsum("b,bnd->nd", ct, FB).cpu().numpy(), F) < 1e-10
    # oracle: dB per neighbour from the oracle, assembled per basis function
    i, j, S, r = N.neighbour_pairs(at.positions, at.cell, at.pbc, 5.0)
    off = N.offsets_of(i, len(at)).astype(np.int32)
    AA = O.eval_AA(T, O.eval_A(T, off, r))
    EBr = O.eval_B(T, AA).sum(0)
    _, dAA = O.eval_AA_dAA(T, off, r)
    dB = O.eval_dB(T, dAA)                       # (nnbr, nB, 3)
    FBr = np.stack([N.assemble_forces(len(at), i, j, dB[:, k, :]) for k in range(len(basis))])
    assert rel_err(EB.cpu().numpy(), EBr) < 1e-12
    assert rel_err(FB.cpu().numpy(), FBr) < 1e-10


def test_lsq_recovers_the_model(cuda):
    print("[gpu] least-squares fit on device-labelled configurations", flush=True)
    basis, T = basis_for(2, 8, 4)
    c_true = synth.coefficients(len(basis), 9)
    calc = ACECalculator(HIPEvaluator(basis, c_true, device=0))
    train = training_set(calc, [_config(s) for s in range(6)])
    E0 = get_V0(train)
    assert E0.shape == (1,)
    c, relrmse = lsq(train, calc, wE=1.0, wF=1.0)
    assert relrmse < 1e-9, relrmse
    # the fitted model reproduces held-out energies and forces
    test = _config(99)
    fit = ACECalculator(HIPEvaluator(basis, c, device=0))
    E1, F1 = fit.energy_forces(test)
    E2, F2 = calc.energy_forces(test)
    assert abs(E1 - E2) < 1e-6 * abs(E2) and rel_err(F1, F2) < 1e-6
