"""hipGraph capture (gpu): one energy+forces call captured with torch's graph
API (HIP stream capture) replays to the eager result, replay after replay,
with eager work and synchronisations between the replays."""
import pytest
import torch

from acehip import synth
from acehip.evaluator import EnvBatch, HIPEvaluator
from helpers import basis_for

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(cuda):
    print("[gpu] graph capture of energy+forces, order-4 degree-12 basis, 300 envs", flush=True)
    basis, T = basis_for(4, 12, 6)
    c = synth.coefficients(len(basis), 2)
    off, R, _ = synth.workload_envs(2, nenv=300)
    off2, R2, _ = synth.workload_envs(2, nenv=90, seed=77)
    ev = HIPEvaluator(basis, c, device=0, waves=16)
    ev.set_validation(False)  # host-side checks copy to the host: not capturable
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    b2 = EnvBatch(off2, R2.T.copy(), None, device="cuda:0")
    E0, dE0 = ev.energy_forces(b)  # eager reference; sizes every buffer
    torch.cuda.synchronize()
    E = torch.empty_like(E0)
    dE = torch.empty_like(dE0)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ev.energy_forces(b, E, dE)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ev.energy_forces(b, E, dE)
    for rep in range(4):
        E.zero_()
        dE.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(E, E0), rep
        assert torch.equal(dE, dE0), rep
        Eb, dEb = ev.energy_forces(b2)  # eager work on another batch between replays
        torch.cuda.synchronize()
        assert torch.isfinite(Eb).all()
