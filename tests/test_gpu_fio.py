"""A model read back from its ACE.jl JSON file evaluates bit-identically on
the device (gpu), and its device energies and forces match the oracle run on
the original model's tables (energies 1e-12, forces 1e-10 relative to the
batch's largest value)."""
import json

import pytest
import torch

import ace_oracle as O
from acehip import fio, synth
from acehip.evaluator import EnvBatch, HIPEvaluator
from acehip.symmbasis import LinearACEModel
from helpers import basis_for, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("args", [(4, 12, 6, "Invariant", 0), (3, 6, 3, "Invariant", 2),
                                  (3, 6, 3, "EuclideanVector", 0)])
def test_loaded_model_matches_original(cuda, args):
    order, deg, L, prop, ns = args
    print(f"[gpu] JSON round trip then device evaluation {args}", flush=True)
    basis, T = basis_for(order, deg, L, prop=prop, species=ns)
    c = synth.coefficients(len(basis), 8)
    m2 = fio.read_dict(json.loads(json.dumps(fio.write_dict(LinearACEModel(basis, c)))))
    ev1 = HIPEvaluator(basis, c, device=0)
    ev2 = HIPEvaluator(m2.basis, m2.c, device=0)
    off, R, sp = synth.random_envs(50, "poisson30", 12, nspecies=ns)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    E1, dE1 = ev1.energy_forces(b)
    E2, dE2 = ev2.energy_forces(b)
    torch.cuda.synchronize()
    assert torch.equal(E1, E2) and torch.equal(dE1, dE2)
    # the loaded model on the device against the oracle on the original tables
    n = 8
    o, Rs = off[: n + 1], R[: off[n]]
    ss = None if sp is None else sp[: off[n]]
    E2, dE2 = E2.cpu().numpy(), dE2.cpu().numpy()
    assert rel_err(E2[:n], O.energy(T, c, o, Rs, ss)) < 1e-12
    g = dE2.T[: off[n]] if prop == "Invariant" else dE2.reshape(3, 3, -1).transpose(2, 0, 1)[: off[n]]
    assert rel_err(g, O.grad_config(T, c, o, Rs, ss)) < 1e-10
