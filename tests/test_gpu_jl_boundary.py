"""The context built from ACE.jl's own tables (ace_ctx_create_jl: the library
compiles the device programs from Product1pBasis.indices, PIBasisSpec and
A2Bmap in CSC) against the one built from the Python packer's tables
(ace_ctx_create): energies, gradients and basis values must be bit-identical.
Bases beyond the record kernel (correlation order 5, more than 160 pooled
values) run the run interpreter; they are checked against the oracle.

Tolerances: energies / basis values within 1e-12 of the batch's largest
value, forces within 1e-10 (as every parity test here)."""
import numpy as np
import pytest
import torch

import ace_oracle as O
from acehip import _lib, synth
from acehip.evaluator import EnvBatch, HIPEvaluator
from helpers import basis_for, random_batch, rel_err

pytestmark = pytest.mark.gpu

E_TOL = 1e-12
F_TOL = 1e-10

BOUNDARY_CASES = [
    ("config1", (3, 6, 4), {}, 1, 300),
    ("config2-small", (4, 12, 6), {}, 2, 200),
    ("config4-small", (3, 10, 5), {"prop": "EuclideanVector", "species": 2}, 4, 70),
    ("two-species", (3, 6, 3), {"species": 2}, None, 90),
    ("order5", (5, 6, 3), {}, None, 130),
]


@pytest.mark.parametrize("name,args,kw,wid,nenv", BOUNDARY_CASES, ids=[c[0] for c in BOUNDARY_CASES])
def test_jl_tables_bit_identical_to_python_pack(cuda, name, args, kw, wid, nenv):
    print(f"[gpu] ace_ctx_create_jl vs ace_ctx_create: {name}", flush=True)
    basis, T = basis_for(*args, **kw)
    c = synth.coefficients(len(basis), 7)
    ns = kw.get("species", 0)
    if wid is not None:
        off, R, sp = synth.workload_envs(wid, nenv=nenv)
    else:
        off, R, sp = random_batch(nenv, 31, nspecies=ns)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    out = {}
    for compiler in ("native", "python"):
        ev = HIPEvaluator(basis, c, device=0, waves=8, compiler=compiler)
        E, dE = ev.energy_forces(b)
        E1 = ev.energy(b)
        B = ev.basis_B(b)
        torch.cuda.synchronize()
        out[compiler] = (E.clone(), dE.clone(), E1.clone(), B.clone())
    for a, r in zip(out["native"], out["python"]):
        assert torch.equal(a, r)
    # and both against the oracle on the first environments
    n = min(nenv, 12)
    o = off[: n + 1]
    Rs = R[: off[n]]
    ss = None if sp is None else sp[: off[n]]
    E, dE = out["native"][0].cpu().numpy(), out["native"][1].cpu().numpy()
    assert rel_err(E[:n], O.energy(T, c, o, Rs, ss)) < E_TOL
    g_ref = O.grad_config(T, c, o, Rs, ss)
    g = dE.T[: off[n]] if T["prop"] == "Invariant" else dE.reshape(3, 3, -1).transpose(2, 0, 1)[: off[n]]
    assert rel_err(g, g_ref) < F_TOL


def test_order5_basis_vs_oracle(cuda):
    """Correlation order 5 (degree 6, L 3): prefixes of four factors, run by
    the segment kernel (records with deep prefixes) and by the interpreter
    with seven prefix levels; 200 environments = 4 tiles, with the tiles'
    tasks split over 2 workgroups too.  The per-record dispatch kernel has no
    records for such bases and says so."""
    print("[gpu] order-5 basis vs oracle", flush=True)
    basis, T = basis_for(5, 6, 3)
    assert max(T["orders"]) == 5
    c = synth.coefficients(len(basis), 11)
    off, R, _ = random_batch(200, 41, nmin=4, nmax=30)
    ev = HIPEvaluator(basis, c, device=0, waves=8)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E_ref, g_ref = O.energy(T, c, off, R), O.grad_config(T, c, off, R)
    for mode, parts, kern in ((0, 0, 0), (0, 2, 0), (1, 0, 64)):
        ev.set_kernel(mode)
        ev.set_parts(parts)
        E, dE = ev.energy_forces(b)
        E1 = ev.energy(b)
        torch.cuda.synchronize()
        assert ev.last_dag_shape()[2] == kern
        assert rel_err(E.cpu().numpy(), E_ref) < E_TOL
        assert rel_err(E1.cpu().numpy(), E_ref) < E_TOL
        assert rel_err(dE.cpu().numpy().T, g_ref) < F_TOL
    ev.set_kernel(2)
    ev.set_parts(0)
    with pytest.raises(_lib.AceError, match="correlation order"):
        ev.energy_forces(b)
    ev.set_kernel(0)
    Bv = ev.basis_B(b).cpu().numpy()
    B_ref = O.eval_B(T, O.eval_AA(T, O.eval_A(T, off, R)))
    assert rel_err(Bv, B_ref) < E_TOL


@pytest.mark.parametrize("deg,order,nenv", [(14, 3, 150), (14, 4, 70)])
def test_invariant_above_160_pooled_values(cuda, deg, order, nenv):
    """One species, degree 14, L 6: 195 pooled values, too many for a
    64-environment LDS tile.  The record kernel runs 32-environment tiles (the
    two half-waves on one record stream, two leaves each); the run interpreter
    (forced) runs 32-environment tiles with two task streams per wave; mode 2
    runs the records with per-record dispatch instead of kind-sorted
    segments."""
    print(f"[gpu] npool 195 invariant, order {order}, degree {deg}", flush=True)
    basis, T = basis_for(order, deg, 6)
    c = synth.coefficients(len(basis), 12)
    off, R, _ = synth.workload_envs(2, nenv=nenv, seed=43)
    ev = HIPEvaluator(basis, c, device=0, waves=8)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E_ref = O.energy(T, c, off, R)
    n = 40 if order == 4 else nenv
    g_ref = O.grad_config(T, c, off[: n + 1], R[: off[n]])
    for mode, kern in ((0, 1), (1, 32), (2, 1)):
        ev.set_kernel(mode)
        E, dE = ev.energy_forces(b)
        E1 = ev.energy(b)
        torch.cuda.synchronize()
        assert ev.last_dag_shape()[2] == kern
        assert rel_err(E.cpu().numpy(), E_ref) < E_TOL
        # the energy-only call sums the energy tasks over another wave assignment
        assert rel_err(E1.cpu().numpy(), E_ref) < E_TOL
        assert rel_err(dE.cpu().numpy().T[: off[n]], g_ref) < F_TOL


def test_two_species_order4_degree12_vs_oracle(cuda):
    """Two species, order 4, degree 12, L 6 (Invariant): 278 pooled values,
    491k AA functions, 38.8k basis functions.  The record kernel runs
    32-environment tiles; 96 environments are three tiles.  The oracle costs
    about 0.7 s per environment for the energy and 4.6 s for the gradient, so
    it checks 16 energies and 3 gradients spread over the tiles."""
    print("[gpu] two species, order 4, degree 12: npool 278", flush=True)
    basis, T = basis_for(4, 12, 6, species=2)
    c = synth.coefficients(len(basis), 13)
    off, R, sp = synth.random_envs(96, "poisson30", 17, nspecies=2)
    ev = HIPEvaluator(basis, c, device=0, waves=8)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    E, dE = ev.energy_forces(b)
    torch.cuda.synchronize()
    assert ev.last_dag_shape()[2] == 1
    E, dE = E.cpu().numpy(), dE.cpu().numpy()
    picks = list(range(0, 96, 6))
    E_ref, g, g_ref = [], [], []
    for e in picks:
        oe = np.array([0, off[e + 1] - off[e]], dtype=np.int32)
        Re, se = R[off[e]:off[e + 1]], sp[off[e]:off[e + 1]]
        E_ref.append(O.energy(T, c, oe, Re, se)[0])
        if e in (0, 42, 90):
            g.append(dE[:, off[e]:off[e + 1]].T)
            g_ref.append(O.grad_config(T, c, oe, Re, se))
    assert rel_err(E[picks], np.asarray(E_ref)) < E_TOL
    assert rel_err(np.concatenate(g), np.concatenate(g_ref)) < F_TOL


@pytest.mark.parametrize("args,kw", [((3, 6, 4), {}), ((3, 6, 3), {"species": 2}),
                                     ((3, 6, 4), {"prop": "EuclideanVector"})])
def test_host_memory_entry_points(cuda, args, kw):
    """The host-memory calls a Julia caller uses (ace_*_host: host arrays in and
    out) return exactly what the device entry points return."""
    import ctypes as C
    print(f"[gpu] host-memory entry points {args} {kw}", flush=True)
    basis, T = basis_for(*args, **kw)
    c = synth.coefficients(len(basis), 21)
    ns = kw.get("species", 0)
    off, R, sp = random_batch(70, 22, nspecies=ns)
    ev = HIPEvaluator(basis, c, device=0)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    nc = 1 if ev.prop == 0 else 3
    nnbr = int(off[-1])
    E_d, dE_d = ev.energy_forces(b)
    B_d = ev.basis_B(b)
    dB_d = ev.basis_dB(b)
    torch.cuda.synchronize()
    lib = ev.lib
    p = lambda a: C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p(0)  # noqa: E731
    off32 = np.ascontiguousarray(off, dtype=np.int32)
    xyz = np.ascontiguousarray(R.T, dtype=np.float64)
    spa = None if sp is None else np.ascontiguousarray(sp, dtype=np.int8)
    E = np.zeros(70 * nc)
    g = np.zeros(3 * nc * nnbr)
    from acehip._lib import check
    check(lib.ace_energy_forces_host(ev.ctx, 70, p(off32), p(xyz), p(spa), p(E), p(g)))
    assert np.array_equal(E, E_d.cpu().numpy().reshape(-1))
    assert np.array_equal(g, dE_d.cpu().numpy().reshape(-1))
    E1 = np.zeros(70 * nc)
    check(lib.ace_energy_host(ev.ctx, 70, p(off32), p(xyz), p(spa), p(E1)))
    assert rel_err(E1, E) < E_TOL
    B = np.zeros(70 * ev.nB * nc)
    check(lib.ace_basis_B_host(ev.ctx, 70, p(off32), p(xyz), p(spa), p(B)))
    assert np.array_equal(B, B_d.cpu().numpy().reshape(-1))
    dB = np.zeros(nnbr * ev.nB * 3 * nc)
    check(lib.ace_basis_dB_host(ev.ctx, 70, p(off32), p(xyz), p(spa), p(dB)))
    assert np.array_equal(dB, dB_d.cpu().numpy().reshape(-1))
    if ev.prop == 0:
        w = np.random.default_rng(3).normal(size=(3, nnbr))
        out = np.zeros(70 * ev.nB)
        check(lib.ace_adjoint_EVAL_D_host(ev.ctx, 70, p(off32), p(xyz), p(spa), p(np.ascontiguousarray(w)), p(out)))
        ref = ev.adjoint_EVAL_D(b, torch.tensor(w)).cpu().numpy().reshape(-1)
        torch.cuda.synchronize()
        assert np.array_equal(out, ref)
    bad = off32.copy()
    bad[3] = bad[2] - 1
    from acehip._lib import AceError
    with pytest.raises(AceError, match="non-decreasing"):
        check(lib.ace_energy_host(ev.ctx, 70, p(bad), p(xyz), p(spa), p(E1)))


@pytest.mark.parametrize("wid,nenv", [(6, 70)], ids=["reference-profiled-model"])
def test_lmax10_reference_profiled_model(cuda, wid, nenv):
    """The model of profile/profile_linearmodel.jl:11-25: SparseBasis order
    4, degree 14 with weights n 1 / l 1.5, maxL = 10 (per-neighbour kernels at
    Lmax 10), on 10 neighbours; 70 environments (two tiles) against the
    oracle, gradients on the first 12."""
    print("[gpu] reference profiled model, Lmax 10", flush=True)
    w = synth.WORKLOADS[wid]
    basis = synth.make_basis(w)
    from acehip.tables import basis_tables
    T = basis_tables(basis)
    assert T["L"] == 10
    c = synth.coefficients(len(basis), 31)
    off, R, sp = synth.workload_envs(wid, nenv=nenv, seed=32)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    for compiler in ("native", "python"):
        ev = HIPEvaluator(basis, c, device=0, waves=8, compiler=compiler)
        E, dE = ev.energy_forces(b)
        E1 = ev.energy(b)
        torch.cuda.synchronize()
        E, dE, E1 = E.cpu().numpy(), dE.cpu().numpy(), E1.cpu().numpy()
        E_ref = O.energy(T, c, off, R)
        assert rel_err(E, E_ref) < E_TOL
        assert rel_err(E1, E_ref) < E_TOL
        n = 12
        assert rel_err(dE.T[: off[n]], O.grad_config(T, c, off[: n + 1], R[: off[n]])) < F_TOL


def test_lmax9_basis(cuda):
    """Order 2, degree 9, Lmax 9: the Lmax-9 kernels against the oracle."""
    print("[gpu] order 2, degree 9, Lmax 9", flush=True)
    basis, T = basis_for(2, 9, 9)
    c = synth.coefficients(len(basis), 33)
    off, R, _ = random_batch(40, 34)
    ev = HIPEvaluator(basis, c, device=0)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E, dE = ev.energy_forces(b)
    torch.cuda.synchronize()
    assert rel_err(E.cpu().numpy(), O.energy(T, c, off, R)) < E_TOL
    assert rel_err(dE.cpu().numpy().T, O.grad_config(T, c, off, R)) < F_TOL


@pytest.mark.parametrize("prop", ["Invariant", "EuclideanVector"])
def test_three_species_beyond_320_pooled_values(cuda, prop):
    """Three species, order 2, degree 12, L 6: 417 (Invariant) and 498
    (EuclideanVector, class-split) pooled values: the segment kernel on tiles
    of 24 and 20 environments (as many as fit the LDS), and the run
    interpreter (forced) on 16-environment tiles with four task streams per
    wave.  50 environments are several tiles; every energy and gradient
    against the oracle, energy-only calls included, both compilers."""
    print(f"[gpu] three species, npool > 320, {prop}", flush=True)
    basis, T = basis_for(2, 12, 6, species=3, prop=prop)
    c = synth.coefficients(len(basis), 35)
    off, R, sp = random_batch(50, 36, nspecies=3)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    E_ref = O.energy(T, c, off, R, sp)
    g_ref = O.grad_config(T, c, off, R, sp)
    for compiler, mode, kern in (("native", 0, 1), ("native", 1, 16), ("python", 0, 1)):
        ev = HIPEvaluator(basis, c, device=0, waves=8, compiler=compiler)
        ev.set_kernel(mode)
        E, dE = ev.energy_forces(b)
        E1 = ev.energy(b)
        torch.cuda.synchronize()
        assert ev.last_dag_shape()[2] == kern
        E, dE, E1 = E.cpu().numpy(), dE.cpu().numpy(), E1.cpu().numpy()
        assert rel_err(E, E_ref) < E_TOL
        assert rel_err(E1, E_ref) < E_TOL
        g = dE.T if prop == "Invariant" else dE.reshape(3, 3, -1).transpose(2, 0, 1)
        assert rel_err(g, g_ref) < F_TOL
