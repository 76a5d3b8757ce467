"""GPU parity: libacehip.so through the C ABI vs the CPU oracle.

Tolerances (written here, as the bar): energies / basis values within
1e-12 relative to the largest value of the batch; forces within 1e-10.
"""
import numpy as np
import pytest
import torch

import ace_oracle as O
from acehip import synth
from acehip.evaluator import EnvBatch, HIPEvaluator
from acehip.pack import Pack
from helpers import basis_for, random_batch, rel_err

pytestmark = pytest.mark.gpu

E_TOL = 1e-12
F_TOL = 1e-10


def _run(basis, T, c, off, R, sp=None, waves=8, forces=True, parts=0):
    ev = HIPEvaluator(basis, c, device=0, waves=waves)
    ev.set_parts(parts)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    if forces:
        E, dE = ev.energy_forces(b)
        torch.cuda.synchronize()
        return ev, E.cpu().numpy(), dE.cpu().numpy()
    E = ev.energy(b)
    torch.cuda.synchronize()
    return ev, E.cpu().numpy(), None


@pytest.mark.parametrize("waves", [4, 8, 16])
def test_config1_energy_forces(cuda, waves):
    print(f"[gpu] config1 energy+forces waves={waves}", flush=True)
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 1)
    off, R, sp = synth.workload_envs(1, nenv=1000)
    _, E, dE = _run(basis, T, c, off, R, waves=waves)
    E_ref = O.energy(T, c, off, R)
    assert rel_err(E, E_ref) < E_TOL
    g_ref = O.grad_config(T, c, off, R)
    assert rel_err(dE.T, g_ref) < F_TOL


@pytest.mark.parametrize("parts", [1, 2, 3, 4])
def test_config2_tile_parts(cuda, parts):
    """Each tile's tasks split over 1 to 4 workgroups: energies summed from
    the parts' shares, gradient tasks each in exactly one part."""
    print(f"[gpu] order-4 degree-12 basis, parts={parts}", flush=True)
    basis, T = basis_for(4, 12, 6)
    c = synth.coefficients(len(basis), 2)
    off, R, sp = synth.workload_envs(2, nenv=130)
    for forces in (True, False):
        _, E, dE = _run(basis, T, c, off, R, waves=16, forces=forces, parts=parts)
        assert rel_err(E, O.energy(T, c, off, R)) < E_TOL
        if forces:
            assert rel_err(dE.T, O.grad_config(T, c, off, R)) < F_TOL


@pytest.mark.parametrize("args,wid,nenv,kern", [((4, 12, 6), 2, 130, 0)], ids=["config2-64env"])
def test_segment_and_per_record_kernels(cuda, args, wid, nenv, kern):
    """The kind-sorted segment kernel (default) and the per-record dispatch
    kernel (mode 2) run the same records in different orders: both match the
    oracle (64-environment tiles, 139 pooled values; the 32-environment tiles
    are covered by the 195-pooled-value test)."""
    print(f"[gpu] segment vs per-record kernel, basis {args}, {nenv} envs", flush=True)
    basis, T = basis_for(*args)
    c = synth.coefficients(len(basis), 2)
    off, R, sp = synth.workload_envs(wid, nenv=nenv)
    ev = HIPEvaluator(basis, c, device=0, waves=16)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    Eref, gref = O.energy(T, c, off, R), O.grad_config(T, c, off, R)
    for mode in (0, 2):
        ev.set_kernel(mode)
        E, dE = ev.energy_forces(b)
        torch.cuda.synchronize()
        assert ev.last_dag_shape()[2] == kern
        assert rel_err(E.cpu().numpy(), Eref) < E_TOL
        assert rel_err(dE.cpu().numpy().T, gref) < F_TOL


@pytest.mark.parametrize("deg,nlong", [(6, 30), (12, 30), (12, 300)])
def test_force_blocks_of_short_environments(cuda, deg, nlong):
    """The flat force kernel stages the G of a block's environments in LDS (up
    to 32 of them for the small basis, 16 for the order-4 one) and stages a
    block that spans more in several rounds: one batch holds blocks of both
    kinds.  150 environments of 1-3 neighbours then nlong of 20-40: with 30
    long ones the order-4 basis runs 128-neighbour blocks (short on average),
    with 300 256-neighbour blocks."""
    print(f"[gpu] force blocks of short environments deg={deg} nlong={nlong}", flush=True)
    basis, T = basis_for(3, 6, 4) if deg == 6 else basis_for(4, 12, 6)
    c = synth.coefficients(len(basis), 3)
    o1, R1, _ = random_batch(150, 60, nmin=1, nmax=3)
    o2, R2, _ = random_batch(nlong, 61, nmin=20, nmax=40)
    off = np.concatenate([o1, o2[1:] + o1[-1]]).astype(np.int32)
    R = np.concatenate([R1, R2])
    _, E, dE = _run(basis, T, c, off, R, waves=16 if deg == 12 else 8)
    assert rel_err(E, O.energy(T, c, off, R)) < E_TOL
    assert rel_err(dE.T, O.grad_config(T, c, off, R)) < F_TOL


_TAIL_REF = {}


@pytest.mark.parametrize("parts,tail", [(2, 1), (3, 2), (2, 3), (3, 9)])
def test_tail_tiles_with_one_more_part(cuda, parts, tail):
    """The last `tail` tiles of the product-DAG grid run parts + 1 workgroups
    with the (parts + 1)-way task split; their energies sum parts + 1 shares.
    200 environments are 4 tiles (the last one partial); tail 9 covers all."""
    print(f"[gpu] tail tiles parts={parts} tail={tail}", flush=True)
    basis, T = basis_for(4, 12, 6)
    c = synth.coefficients(len(basis), 2)
    off, R, sp = synth.workload_envs(2, nenv=200)
    if not _TAIL_REF:
        _TAIL_REF["E"] = O.energy(T, c, off, R)
        _TAIL_REF["g"] = O.grad_config(T, c, off, R)
    ev = HIPEvaluator(basis, c, device=0, waves=16)
    ev.set_parts(parts)
    ev.set_tail_parts(tail)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E, dE = ev.energy_forces(b)
    E2 = ev.energy(b)
    torch.cuda.synchronize()
    assert rel_err(E.cpu().numpy(), _TAIL_REF["E"]) < E_TOL
    assert rel_err(E2.cpu().numpy(), _TAIL_REF["E"]) < E_TOL
    assert rel_err(dE.cpu().numpy().T, _TAIL_REF["g"]) < F_TOL


def test_tail_tiles_full_size(cuda):
    """At 100k environments (1,563 tiles) the automatic choice gives the tiles
    of the grid's last round one more part (checked through the launch the
    library reports).  Energies and gradients equal the run without tail
    tiles, and the oracle on the last tile's environments and on 48 more
    spread over the whole batch (every XCD's range, full and tail tiles)."""
    print("[gpu] tail tiles at 100k environments", flush=True)
    basis, T = basis_for(4, 12, 6)
    c = synth.coefficients(len(basis), 2)
    off, R, _ = synth.workload_envs(2, nenv=100000)
    ev = HIPEvaluator(basis, c, device=0, waves=16)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    runs = {}
    for parts, tail in ((0, 0), (0, -1), (3, 128)):
        ev.set_parts(parts)
        ev.set_tail_parts(tail)
        E, dE = ev.energy_forces(b)
        torch.cuda.synchronize()
        runs[(parts, tail)] = (E.cpu().numpy(), dE.cpu().numpy())
        P, ntail, kern = ev.last_dag_shape()
        assert kern == 0
        if tail == -1:
            assert P >= 2 and ntail > 0, (P, ntail)  # the automatic choice really runs tail tiles
        if tail == 0:
            assert ntail == 0
    E0, g0 = runs[(0, 0)]
    for key, (E, g) in runs.items():
        assert rel_err(E, E0) < 1e-14, key
        assert rel_err(g, g0) < 1e-14, key
    e0 = 100000 - 24
    o = (off[e0:] - off[e0]).astype(np.int32)
    Rs = R[off[e0]:]
    Et, gt = runs[(3, 128)]
    assert rel_err(Et[e0:], O.energy(T, c, o, Rs)) < E_TOL
    assert rel_err(gt[:, off[e0]:].T, O.grad_config(T, c, o, Rs)) < F_TOL
    Ea, ga = runs[(0, -1)]
    for e in np.linspace(0, 100000 - 25, 48).astype(int):
        oe = np.array([0, off[e + 1] - off[e]], dtype=np.int32)
        Re = R[off[e]:off[e + 1]]
        assert rel_err(Ea[e:e + 1], O.energy(T, c, oe, Re)) < E_TOL, e
        assert rel_err(ga[:, off[e]:off[e + 1]].T, O.grad_config(T, c, oe, Re)) < F_TOL, e


def test_tail_parts_rejects_bad_settings(cuda):
    from acehip._lib import AceError
    basis, T = basis_for(3, 6, 4)
    ev = HIPEvaluator(basis, synth.coefficients(len(basis), 1), device=0)
    with pytest.raises(AceError):
        ev.set_tail_parts(-2)
    off, R, _ = random_batch(70, 3)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    for parts in (1, 4):
        ev.set_parts(parts)
        ev.set_tail_parts(1)
        with pytest.raises(AceError, match="tail"):
            ev.energy(b)
    ev.set_tail_parts(0)
    E = ev.energy(b)
    torch.cuda.synchronize()
    assert rel_err(E.cpu().numpy(), O.energy(T, synth.coefficients(len(basis), 1), off, R)) < E_TOL


def test_set_parts_rejects_bad_values(cuda):
    basis, T = basis_for(3, 6, 4)
    ev = HIPEvaluator(basis, synth.coefficients(len(basis), 1), device=0)
    for bad in (-1, 5):
        with pytest.raises(Exception):
            ev.set_parts(bad)


def test_config1_energy_only(cuda):
    print("[gpu] config1 energy only", flush=True)
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 1)
    off, R, sp = synth.workload_envs(1, nenv=300)
    _, E, _ = _run(basis, T, c, off, R, forces=False)
    assert rel_err(E, O.energy(T, c, off, R)) < E_TOL


def test_config2_basis_small_batch(cuda):
    print("[gpu] order-4 degree-12 L6 basis, 48 envs", flush=True)
    basis, T = basis_for(4, 12, 6)
    c = synth.coefficients(len(basis), 2)
    off, R, sp = synth.workload_envs(2, nenv=48)
    _, E, dE = _run(basis, T, c, off, R)
    E_ref = O.energy(T, c, off, R)
    assert rel_err(E, E_ref) < E_TOL
    g_ref = O.grad_config(T, c, off, R[: off[8]]) if False else None
    # forces of the first 8 environments (oracle cost)
    g_ref = O.grad_config(T, c, off[:9], R[: off[8]])
    assert rel_err(dE.T[: off[8]], g_ref) < F_TOL


def test_config3_fcc_forces_fd(cuda):
    print("[gpu] config3 fcc energy+forces + central FD", flush=True)
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 3)
    off, R, _ = synth.workload_envs(3, nenv=64)
    ev, E, dE = _run(basis, T, c, off, R)
    assert rel_err(E, O.energy(T, c, off, R)) < E_TOL
    assert rel_err(dE.T, O.grad_config(T, c, off, R)) < F_TOL
    # central finite differences of the GPU energy, h = 1e-5
    h = 1e-5
    e = 0
    j0, j1 = off[e], off[e + 1]
    fd = np.zeros((j1 - j0, 3))
    for j in range(j0, j1):
        for d in range(3):
            Rp = R[j0:j1].copy(); Rp[j - j0, d] += h
            Rm = R[j0:j1].copy(); Rm[j - j0, d] -= h
            bp = EnvBatch([0, j1 - j0], Rp.T.copy(), None, device="cuda:0")
            bm = EnvBatch([0, j1 - j0], Rm.T.copy(), None, device="cuda:0")
            fd[j - j0, d] = (ev.energy(bp).item() - ev.energy(bm).item()) / (2 * h)
    assert rel_err(dE.T[j0:j1], fd) < 1e-6


_VEC_REF = {}


@pytest.mark.parametrize("mode,parts,tail", [(0, 0, -1), (0, 2, -1), (0, 3, 1), (1, 0, -1), (2, 2, -1)],
                         ids=["records", "records-2parts", "records-3parts-tail", "interpreter",
                              "per-record-2parts"])
def test_vector_property(cuda, mode, parts, tail):
    """EuclideanVector, order 3, degree 6 (class-split programs): 64-environment
    tiles; 150 environments are two full tiles and a partial one.  The record
    kernel with complex leaf coefficients (kind-sorted segments; mode 2 the
    per-record dispatch kernel), with the tiles split over 2 and 3 workgroups
    (and one tail tile of 4), and the run interpreter; energies with and
    without forces."""
    print(f"[gpu] EuclideanVector order 3 degree 6, 150 envs, kernel mode {mode}, parts {parts}", flush=True)
    basis, T = basis_for(3, 6, 4, prop="EuclideanVector")
    c = synth.coefficients(len(basis), 4)
    off, R, sp = random_batch(150, 11)
    ev = HIPEvaluator(basis, c, device=0, waves=8)
    ev.set_kernel(mode)
    ev.set_parts(parts)
    ev.set_tail_parts(tail)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E, dE = ev.energy_forces(b)
    torch.cuda.synchronize()
    P, ntail, kern = ev.last_dag_shape()
    assert kern == (64 if mode == 1 else 0)
    if mode != 1 and parts:
        assert P == parts and ntail == max(0, tail)
    E1 = ev.energy(b)
    E, dE, E1 = E.cpu().numpy(), dE.cpu().numpy(), E1.cpu().numpy()
    if not _VEC_REF:
        _VEC_REF["E"] = O.energy(T, c, off, R)
        _VEC_REF["g"] = O.grad_config(T, c, off, R)   # [nbr, p, d]
    assert rel_err(E, _VEC_REF["E"]) < E_TOL
    assert rel_err(E1, _VEC_REF["E"]) < E_TOL
    g = dE.reshape(3, 3, -1).transpose(2, 0, 1)
    assert rel_err(g, _VEC_REF["g"]) < F_TOL


def test_two_species_invariant(cuda):
    print("[gpu] two species invariant order 3 degree 6", flush=True)
    basis, T = basis_for(3, 6, 3, species=2)
    c = synth.coefficients(len(basis), 5)
    off, R, sp = random_batch(40, 12, nspecies=2)
    _, E, dE = _run(basis, T, c, off, R, sp)
    assert rel_err(E, O.energy(T, c, off, R, sp)) < E_TOL
    assert rel_err(dE.T, O.grad_config(T, c, off, R, sp)) < F_TOL


def test_basis_level(cuda):
    print("[gpu] basis-level A, AA, B, dB", flush=True)
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 1)
    off, R, sp = random_batch(12, 13)
    ev = HIPEvaluator(basis, c, device=0)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    A = ev.basis_A(b).cpu().numpy()
    AA = ev.basis_AA(b).cpu().numpy()
    B = ev.basis_B(b).cpu().numpy()
    dB = ev.basis_dB(b).cpu().numpy()
    A_ref = O.eval_A(T, off, R)
    AA_ref = O.eval_AA(T, A_ref)
    assert rel_err(A, A_ref) < E_TOL
    assert rel_err(AA, AA_ref) < E_TOL
    assert rel_err(B, O.eval_B(T, AA_ref)) < E_TOL
    _, dAA = O.eval_AA_dAA(T, off, R)
    assert rel_err(dB, O.eval_dB(T, dAA)) < F_TOL



@pytest.mark.parametrize("args,kw", [((3, 10, 5), {"prop": "EuclideanVector", "species": 2}), ((2, 16, 6), {})],
                         ids=["vector-two-species-npool220", "npool251"])
def test_basis_B_32_environment_tiles(cuda, args, kw):
    """Basis values of bases whose A+ tile of 64 environments exceeds the LDS
    (more than 160 pooled values) run the fused kernel on 32-environment
    tiles, the two half-waves taking alternate rows; over three tiles and a
    partial one, against the oracle's B = A2B . AA."""
    print(f"[gpu] basis B on 32-environment tiles {args} {kw}", flush=True)
    basis, T = basis_for(*args, **kw)
    c = synth.coefficients(len(basis), 2)
    ns = kw.get("species", 0)
    off, R, sp = random_batch(100, 41, nmin=3, nmax=30, nspecies=ns)
    ev = HIPEvaluator(basis, c, device=0)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    B = ev.basis_B(b).cpu().numpy()
    ref = O.eval_B(T, O.eval_AA(T, O.eval_A(T, off, R, sp)))
    assert B.size == ref.size
    assert rel_err(B.reshape(ref.shape), ref) < E_TOL


@pytest.mark.parametrize("nenv,species", [(1, 0), (3, 0), (65, 0), (33, 2)])
def test_two_environments_per_wave(cuda, nenv, species):
    """Bases with at most 32 pooled values pool two environments per wave;
    odd batch sizes leave the last half-wave without an environment."""
    print(f"[gpu] two environments per wave nenv={nenv} species={species}", flush=True)
    from acehip.pack import Pack
    basis, T = basis_for(3, 6, 4) if species == 0 else basis_for(2, 4, 2, species=2)
    assert len(Pack(T).pool) <= 32
    c = synth.coefficients(len(basis), 9)
    off, R, sp = random_batch(nenv, 50 + nenv, nmin=1, nmax=40, nspecies=species)
    ev, E, dE = _run(basis, T, c, off, R, sp)
    assert rel_err(E, O.energy(T, c, off, R, sp)) < E_TOL
    assert rel_err(dE.T, O.grad_config(T, c, off, R, sp)) < F_TOL
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    assert rel_err(ev.basis_A(b).cpu().numpy(), O.eval_A(T, off, R, sp)) < E_TOL


def test_empty_and_ragged(cuda):
    print("[gpu] empty environments", flush=True)
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 1)
    off = np.array([0, 0, 3, 3, 10], dtype=np.int32)
    _, _, Rr = None, None, random_batch(1, 14, 10, 10)[1]
    _, E, dE = _run(basis, T, c, off, Rr)
    ct = O.ctilde(T, c)
    e0 = ct[0] if T["orders"][0] == 0 else 0.0
    assert E[0] == pytest.approx(e0, rel=1e-14)
    assert E[2] == pytest.approx(e0, rel=1e-14)
    nz = [1, 3]
    offs = np.array([0, 3, 10], dtype=np.int32)
    assert rel_err(E[nz], O.energy(T, c, offs, Rr)) < E_TOL
    assert rel_err(dE.T, O.grad_config(T, c, offs, Rr)) < F_TOL


def test_set_params(cuda):
    print("[gpu] set_params", flush=True)
    basis, T = basis_for(3, 6, 4)
    c1 = synth.coefficients(len(basis), 1)
    c2 = np.random.default_rng(7).normal(size=len(basis))
    off, R, _ = random_batch(20, 15)
    ev = HIPEvaluator(basis, c1, device=0)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    ev.set_params(c2)
    E, dE = ev.energy_forces(b)
    assert rel_err(E.cpu().numpy(), O.energy(T, c2, off, R)) < E_TOL
    assert rel_err(dE.cpu().numpy().T, O.grad_config(T, c2, off, R)) < F_TOL


def test_invalid_inputs_raise(cuda):
    print("[gpu] invalid inputs", flush=True)
    from acehip._lib import AceError
    basis, T = basis_for(3, 6, 3, species=2)
    c = synth.coefficients(len(basis), 5)
    ev = HIPEvaluator(basis, c, device=0)
    off, R, sp = random_batch(4, 16, nspecies=2)
    bad = sp.copy(); bad[0] = 5
    with pytest.raises(AceError, match="species"):
        ev.energy(EnvBatch(off, R.T.copy(), bad, device="cuda:0"))
    with pytest.raises(AceError, match="species array required"):
        ev.energy(EnvBatch(off, R.T.copy(), None, device="cuda:0"))
    off2 = off.copy(); off2[2] = off2[1] - 1
    with pytest.raises(AceError, match="non-decreasing"):
        ev.energy(EnvBatch(off2, R.T.copy(), sp, device="cuda:0"))


_C4_REF = {}


@pytest.mark.parametrize("waves,mode", [(8, 0), (16, 0), (8, 1), (8, 2)],
                         ids=["records-w8", "records-w16", "interpreter-w8", "per-record-w8"])
def test_config4_two_species_vector(cuda, waves, mode):
    """Config 4 (2 species, EuclideanVector, order 3, degree 10, L 5): 220
    pooled values, 32-environment tiles: the record kernel with the two
    half-waves on one record stream (kind-sorted segments; mode 2 per-record
    dispatch), or the run interpreter with two task streams per wave.  100 environments are three full tiles and a partial
    one; every energy and every per-neighbour Jacobian is compared with the
    oracle."""
    print(f"[gpu] config4: 100 envs, waves={waves}, kernel mode {mode} (32-env tiles)", flush=True)
    basis, T = basis_for(3, 10, 5, prop="EuclideanVector", species=2)
    c = synth.coefficients(len(basis), 4)
    off, R, sp = synth.workload_envs(4, nenv=100)
    ev = HIPEvaluator(basis, c, device=0, waves=waves)
    ev.set_kernel(mode)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    E, dE = ev.energy_forces(b)
    torch.cuda.synchronize()
    assert ev.last_dag_shape()[2] == (32 if mode == 1 else 1)
    E, dE = E.cpu().numpy(), dE.cpu().numpy()
    if not _C4_REF:
        _C4_REF["E"] = O.energy(T, c, off, R, sp)
        _C4_REF["g"] = O.grad_config(T, c, off, R, sp)
    assert rel_err(E, _C4_REF["E"]) < E_TOL
    g = dE.reshape(3, 3, -1).transpose(2, 0, 1)
    assert rel_err(g, _C4_REF["g"]) < F_TOL


@pytest.mark.parametrize("args,wid,nenv", [((3, 6, 4), 1, 300), ((4, 12, 6), 2, 200)], ids=["config1", "config2-small"])
def test_pool_shapes_bit_identical(cuda, args, wid, nenv):
    """Every pooling launch shape (8, 16 or 32 neighbours per chunk; one or
    two environments per wave for bases of at most 32 pooled values) sums the
    neighbours in the same order: energies and forces are bit-identical to the
    default shape, and match the oracle."""
    print(f"[gpu] pooling shapes {args}", flush=True)
    basis, T = basis_for(*args)
    c = synth.coefficients(len(basis), 9)
    off, R, _ = synth.workload_envs(wid, nenv=nenv, seed=17)
    ev = HIPEvaluator(basis, c, device=0, waves=8)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E0, dE0 = ev.energy_forces(b)
    E0, dE0 = E0.clone(), dE0.clone()
    shapes = [(ch, 0) for ch in (8, 16, 32)] + ([(ch, 1) for ch in (8, 16, 32)] if ev.npool <= 32 else [])
    for ch, epw in shapes:
        ev.set_pool_shape(ch, epw)
        E, dE = ev.energy_forces(b)
        torch.cuda.synchronize()
        assert torch.equal(E, E0) and torch.equal(dE, dE0), (ch, epw)
    n = 16
    assert rel_err(E0.cpu().numpy()[:n], O.energy(T, c, off[: n + 1], R[: off[n]])) < E_TOL



@pytest.mark.parametrize("args,shared", [((2, 18, 6), True), ((2, 16, 6), False)],
                         ids=["npool307-shared-Y", "npool251-per-pass-Y"])
def test_pool_lane_sets_large_bases(cuda, args, shared):
    """One species, more than 192 pooled values: 307 pooled functions take 8
    passes whose lane sets (functions of one (l, m) per lane) fit 64 lanes, so
    each lane reads Y once per neighbour; 251 take 4 passes whose sets need 77
    lanes, so lanes read Y per pass.  Energies and forces match the oracle and
    every chunk width gives bit-identical results."""
    print(f"[gpu] pooling lane sets {args}", flush=True)
    basis, T = basis_for(*args)
    P = np.asarray(Pack(T).pool).reshape(-1, 4)
    kp = 8 if len(P) > 256 else 4
    sets = {}
    for mu, n, l, m in P.tolist():
        sets[(mu, l, m)] = sets.get((mu, l, m), 0) + 1
    assert (sum(-(-v // kp) for v in sets.values()) <= 64) == shared
    c = synth.coefficients(len(basis), 5)
    off, R, _ = synth.workload_envs(2, nenv=70, seed=23)
    ev = HIPEvaluator(basis, c, device=0)
    b = EnvBatch(off, R.T.copy(), None, device="cuda:0")
    E0, dE0 = ev.energy_forces(b)
    E0, dE0 = E0.clone(), dE0.clone()
    assert rel_err(E0.cpu().numpy(), O.energy(T, c, off, R)) < E_TOL
    assert rel_err(dE0.cpu().numpy().T, O.grad_config(T, c, off, R)) < F_TOL
    for ch in (8, 32):
        ev.set_pool_shape(ch, 0)
        E, dE = ev.energy_forces(b)
        torch.cuda.synchronize()
        assert torch.equal(E, E0) and torch.equal(dE, dE0), ch



FORCE_TABLE_CASES = [
    ("two-species-nr14", (2, 14, 4), {"species": 2}),        # 16 radial registers, Invariant with species
    ("two-species-nr10", (2, 10, 4), {"species": 2}),        # 12 radial registers
    ("vector-nr13", (3, 13, 3), {"prop": "EuclideanVector"}),  # 16 radial registers, three components
]


@pytest.mark.parametrize("name,args,kw", FORCE_TABLE_CASES, ids=[c[0] for c in FORCE_TABLE_CASES])
def test_force_table_radial_registers(cuda, name, args, kw):
    """The per-environment force kernel (species or vector property) with its
    compact adjoint table (each (l, m) holds only the n its degree admits) and
    12 or 16 radial functions in registers: energies and every per-neighbour
    gradient against the oracle."""
    print(f"[gpu] per-environment force table: {name}", flush=True)
    basis, T = basis_for(*args, **kw)
    c = synth.coefficients(len(basis), 6)
    ns = kw.get("species", 0)
    off, R, sp = random_batch(40, 90 + ns, nmin=3, nmax=40, nspecies=ns)
    ev = HIPEvaluator(basis, c, device=0)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    E, dE = ev.energy_forces(b)
    torch.cuda.synchronize()
    E, dE = E.cpu().numpy(), dE.cpu().numpy()
    assert rel_err(E, O.energy(T, c, off, R, sp)) < E_TOL
    g_ref = O.grad_config(T, c, off, R, sp)
    if kw.get("prop") == "EuclideanVector":
        assert rel_err(dE.reshape(3, 3, -1).transpose(2, 0, 1), g_ref) < F_TOL
    else:
        assert rel_err(dE.T, g_ref) < F_TOL


DB_CASES = [
    ("config2", (4, 12, 6), {}, 3),                                            # 3235 B: 13 chunks of 256
    ("two-species", (3, 6, 3), {"species": 2}, 6),
    ("order5", (5, 6, 3), {}, 6),                                              # products of up to 5 factors
    ("config4", (3, 10, 5), {"prop": "EuclideanVector", "species": 2}, 2),     # 3 components per slot
]


@pytest.mark.parametrize("name,args,kw,nenv", DB_CASES, ids=[c[0] for c in DB_CASES])
def test_basis_dB_weight_kernel(cuda, name, args, kw, nenv):
    """ace_basis_dB through per-environment adjoint weights (one workgroup per
    environment and chunk of basis functions) against the oracle's dB =
    A2B . dAA (per neighbour, all basis functions, all directions)."""
    print(f"[gpu] basis dB weight kernel: {name}", flush=True)
    basis, T = basis_for(*args, **kw)
    c = synth.coefficients(len(basis), 3)
    ns = kw.get("species", 0)
    off, R, sp = random_batch(nenv, 60 + nenv, nmin=3, nmax=25, nspecies=ns)
    ev = HIPEvaluator(basis, c, device=0)
    b = EnvBatch(off, R.T.copy(), sp, device="cuda:0")
    dB = ev.basis_dB(b).cpu().numpy()
    _, dAA = O.eval_AA_dAA(T, off, R, sp)
    ref = O.eval_dB(T, dAA)
    assert dB.size == ref.size
    assert rel_err(dB.reshape(ref.shape), ref) < F_TOL
