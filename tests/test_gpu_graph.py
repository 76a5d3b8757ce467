"""hipGraph capture (gpu): one energy+forces call captured with torch's graph
API (HIP stream capture) replays to the eager result, replay after replay,
with eager work and synchronisations between the replays."""
import pytest
import torch

from acehip import synth
from acehip.evaluator import EnvBatch, HIPEvaluator
from helpers import basis_for

pytestmark = pytest.mark.gpu


GRAPH_CASES = [
    ("order4-degree12", (4, 12, 6), {}, 2, 300, 0),
    # class-split EuclideanVector, 32-environment record tiles, 2 parts per
    # tile (the class shares summed by k_esum)
    ("config4-small", (3, 10, 5), {"prop": "EuclideanVector", "species": 2}, 4, 150, 2),
]


@pytest.mark.parametrize("name,args,kw,wid,nenv,parts", GRAPH_CASES, ids=[c[0] for c in GRAPH_CASES])
def test_graph_replays_match_eager(cuda, name, args, kw, wid, nenv, parts):
    print(f"[gpu] graph capture of energy+forces, {name}, {nenv} envs", flush=True)
    basis, T = basis_for(*args, **kw)
    c = synth.coefficients(len(basis), 2)
    off, R, sp = synth.workload_envs(wid, nenv=nenv)
    off2, R2, sp2 = synth.workload_envs(wid, nenv=90, seed=77)
    ev = HIPEvaluator(basis, c, device=0, waves=16)
    ev.set_parts(parts)
    ev.set_validation(False)  # host-side checks copy to the host: not capturable
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    b2 = EnvBatch(off2, R2.T.copy(), sp2, device="cuda:0")
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


def test_graph_replay_after_set_params(cuda):
    """set_params between replays: the captured kernels read the coefficients,
    program records and the energy constant from device memory, which
    set_params rewrites in place, so a replay after set_params equals an
    eager call with the new coefficients (the constant term included)."""
    print("[gpu] graph replay after set_params", flush=True)
    import numpy as np
    basis, T = basis_for(4, 12, 6)
    c1 = synth.coefficients(len(basis), 2)
    c2 = np.random.default_rng(5).normal(size=len(basis)) / len(basis)
    c2[T["a2b"]["rows"][np.asarray(T["orders"])[T["a2b"]["cols"]] == 0]] += 3.0  # move the constant
    off, R, _ = synth.workload_envs(2, nenv=150)
    ev = HIPEvaluator(basis, c1, device=0, waves=16)
    ev.set_validation(False)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E0, dE0 = ev.energy_forces(b)
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
    for rep, c in enumerate((c2, c1, c2)):
        ev.set_params(c)
        Ee, dEe = ev.energy_forces(b)  # eager reference with these coefficients
        torch.cuda.synchronize()
        E.zero_()
        dE.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(E, Ee), rep
        assert torch.equal(dE, dEe), rep
    assert not torch.equal(Ee, E0)


def test_offsets_written_on_a_side_stream(cuda):
    """The call reads env_offsets[nenv] (the neighbour count) in the order of
    the caller's stream: offsets produced by a kernel queued on a side stream
    just before the call are the ones the call sees."""
    print("[gpu] env_offsets produced on a side stream", flush=True)
    import numpy as np
    import ace_oracle as O
    from helpers import rel_err
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 1)
    off, R, _ = synth.workload_envs(1, nenv=500)
    ev = HIPEvaluator(basis, c, device=0)
    s = torch.cuda.Stream()
    counts = torch.tensor(np.diff(off), dtype=torch.int32, device="cuda:0")
    xyz = torch.tensor(R.T.copy(), device="cuda:0")
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        offsets = torch.zeros(len(off), dtype=torch.int32, device="cuda:0")
        if hasattr(torch.cuda, "_sleep"):
            torch.cuda._sleep(2_000_000)  # keep the side stream busy: a read not ordered on it would see zeros
        offsets[1:] = torch.cumsum(counts, 0)
        b = EnvBatch(offsets, xyz, None, device="cuda:0")
        E, dE = ev.energy_forces(b, stream=s)
    torch.cuda.synchronize()
    assert rel_err(E.cpu().numpy(), O.energy(T, c, off, R)) < 1e-12
    assert rel_err(dE.cpu().numpy().T, O.grad_config(T, c, off, R)) < 1e-10
