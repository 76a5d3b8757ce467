"""ace_energy_forces_noalloc (the entry point that allocates nothing; its
scratch is a fixed device array) against ace_energy_forces through the C ABI:
bit-identical energies and gradients on two calls in a row, the second on a
smaller batch, for the record kernel (order-4 degree-12 Invariant basis) and
the run interpreter (config-4 EuclideanVector basis, 32-environment tiles);
a batch larger than the scratch array is rejected with an error."""
import pytest
import torch

from acehip import synth
from acehip._lib import AceError
from acehip.evaluator import EnvBatch, HIPEvaluator
from helpers import basis_for

pytestmark = pytest.mark.gpu

CASES = [
    ("order4-degree12", (4, 12, 6), {}, 2, (3000, 700)),
    ("config4", (3, 10, 5), {"prop": "EuclideanVector", "species": 2}, 4, (400, 90)),
]


@pytest.mark.parametrize("name,args,kw,wid,sizes", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_energy_forces(cuda, name, args, kw, wid, sizes):
    print(f"[gpu] no-alloc entry point vs ace_energy_forces: {name} {sizes}", flush=True)
    basis, T = basis_for(*args, **kw)
    c = synth.coefficients(len(basis), wid)
    ev = HIPEvaluator(basis, c, device=0, waves=16 if wid == 2 else 8)
    nc = 1 if ev.prop == 0 else 3
    for i, n in enumerate(sizes):
        off, R, sp = synth.workload_envs(wid, nenv=n, seed=100 + i)
        b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
        E0, dE0 = ev.energy_forces(b)
        E = torch.full_like(E0, float("nan"))
        dE = torch.full_like(dE0, float("nan"))
        ev.energy_forces_noalloc(b, E, dE)
        torch.cuda.synchronize()
        assert torch.equal(E, E0), (name, n)
        assert torch.equal(dE, dE0), (name, n)
        assert E.shape[0] == n and dE.shape == (3 * nc, b.nnbr)


def test_noalloc_rejects_batches_beyond_the_scratch(cuda):
    print("[gpu] no-alloc entry point: too large a batch", flush=True)
    basis, T = basis_for(4, 12, 6)
    c = synth.coefficients(len(basis), 2)
    ev = HIPEvaluator(basis, c, device=0, waves=16)
    off, R, _ = synth.workload_envs(2, nenv=20000)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E = torch.zeros(b.nenv, dtype=torch.float64, device="cuda:0")
    dE = torch.zeros((3, b.nnbr), dtype=torch.float64, device="cuda:0")
    with pytest.raises(AceError, match="scratch"):
        ev.energy_forces_noalloc(b, E, dE)
    torch.cuda.synchronize()
    assert not torch.any(E != 0)  # nothing was launched
