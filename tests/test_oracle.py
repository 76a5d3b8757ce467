"""The CPU oracle pinned by closed forms, identities and finite differences
(the reference's own test properties; it stores no golden vectors)."""
import numpy as np
import pytest
import scipy.special as sps

import ace_oracle as O
from acehip import synth
from helpers import basis_for, random_batch, rel_err

RNG = np.random.default_rng(2024)


def explicit_shs(th, ph):
    """Closed-form complex Ylm for l <= 3 (test/polynomials/test_ylm.jl:14-35)."""
    s, c, e = np.sin(th), np.cos(th), np.exp(1j * ph)
    return np.array([
        0.5 * np.sqrt(1 / np.pi),
        0.5 * np.sqrt(3 / (2 * np.pi)) * s / e, 0.5 * np.sqrt(3 / np.pi) * c,
        -0.5 * np.sqrt(3 / (2 * np.pi)) * s * e,
        0.25 * np.sqrt(15 / (2 * np.pi)) * s ** 2 / e ** 2,
        0.5 * np.sqrt(15 / (2 * np.pi)) * s * c / e, 0.25 * np.sqrt(5 / np.pi) * (3 * c ** 2 - 1),
        -0.5 * np.sqrt(15 / (2 * np.pi)) * s * c * e, 0.25 * np.sqrt(15 / (2 * np.pi)) * s ** 2 * e ** 2,
        1 / 8 * np.sqrt(35 / np.pi) * s ** 3 / e ** 3,
        1 / 4 * np.sqrt(105 / (2 * np.pi)) * c * s ** 2 / e ** 2,
        1 / 8 * np.sqrt(21 / np.pi) * (-1 + 5 * c ** 2) * s / e,
        1 / 4 * np.sqrt(7 / np.pi) * (-3 * c + 5 * c ** 3),
        -(1 / 8) * np.sqrt(21 / np.pi) * (-1 + 5 * c ** 2) * s * e,
        1 / 4 * np.sqrt(105 / (2 * np.pi)) * c * s ** 2 * e ** 2,
        -(1 / 8) * np.sqrt(35 / np.pi) * s ** 3 * e ** 3])


def _pts(n, pole=False):
    th = RNG.uniform(0, np.pi, n) if not pole else RNG.uniform(0, 1e-9, n)
    ph = RNG.uniform(-np.pi, np.pi, n)
    r = 0.1 + RNG.uniform(size=n)
    R = np.stack([r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph), r * np.cos(th)], -1)
    return th, ph, R


@pytest.mark.parametrize("pole", [False, True])
def test_ylm_closed_form(pole):
    th, ph, R = _pts(30, pole)
    Y = O.ylm_eval(3, R)
    Yex = np.stack([explicit_shs(a, b) for a, b in zip(th, ph)])
    assert np.max(np.abs(Y - Yex)) < 1e-13


def test_ylm_matches_scipy_l6():
    th, ph, R = _pts(40)
    Y = O.ylm_eval(6, R)
    Yed, _ = O.ylm_eval_ed(6, R)
    for l in range(7):
        for m in range(-l, l + 1):
            ref = sps.sph_harm_y(l, m, th, ph)
            assert np.max(np.abs(Y[:, l * l + l + m] - ref)) < 1e-13
            assert np.max(np.abs(Yed[:, l * l + l + m] - ref)) < 1e-13


def test_ylm_gradient_fd():
    _, _, R = _pts(10)
    _, dY = O.ylm_eval_ed(5, R)
    h = 1e-6
    for d in range(3):
        Rp, Rm = R.copy(), R.copy()
        Rp[:, d] += h
        Rm[:, d] -= h
        fd = (O.ylm_eval(5, Rp) - O.ylm_eval(5, Rm)) / (2 * h)
        assert np.max(np.abs(fd - dY[:, :, d])) < 1e-7


def test_radial_orthonormal_and_derivative():
    basis, T = basis_for(4, 12, 6)
    rn = T["rn"]
    # orthonormal on the 300-point midpoint measure the basis was built on
    N = 300
    dt = (rn["tr"] - rn["tl"]) / N
    tq = np.linspace(rn["tl"] + dt / 2, rn["tr"] - dt / 2, N)
    P = O.rn_eval(rn, tq)
    G = P.T @ P / N
    assert np.max(np.abs(G - np.eye(len(rn["A"])))) < 1e-12
    assert np.all(np.asarray(rn["A"]) > 0)
    # transform closed form: t(r) = (2/(1+r))^2, tl = t(5) = 1/9, tr = t(0.5) = 16/9
    assert rn["tl"] == pytest.approx(1 / 9, rel=1e-15)
    assert rn["tr"] == pytest.approx(16 / 9, rel=1e-15)
    r = RNG.uniform(0.8, 5.5, 50)
    assert np.max(np.abs(O.transform(rn, r) - (2 / (1 + r)) ** 2)) < 1e-15
    # derivatives vs central differences; zero beyond rcut
    R3 = np.stack([r, 0 * r, 0 * r], -1)
    _, dR = O.radial_ed(rn, R3)
    h = 1e-6
    fd = (O.radial(rn, R3 + [h, 0, 0]) - O.radial(rn, R3 - [h, 0, 0])) / (2 * h)
    inside = r < 5.0 - 1e-3
    assert np.max(np.abs(fd[inside] - dR[inside, :, 0])) < 1e-6
    assert np.all(O.radial(rn, R3[r > 5.0]) == 0.0)


def test_A_brute_force_and_permutation():
    basis, T = basis_for(3, 6, 4)
    off, R, _ = random_batch(3, 1)
    A = O.eval_A(T, off, R)
    Rn = O.radial(T["rn"], R)
    Y = O.ylm_eval(T["L"], R)
    for e in range(3):
        for iA, (mu, n, l, m) in enumerate(T["aspec"]):
            ref = sum(Rn[j, n - 1] * Y[j, l * l + l + m] for j in range(off[e], off[e + 1]))
            assert abs(A[e, iA] - ref) < 1e-13
    perm = np.concatenate([off[0] + RNG.permutation(off[1] - off[0]), np.arange(off[1], off[-1])])
    assert rel_err(O.eval_A(T, off, R[perm]), A) < 1e-14


def test_AA_naive_products():
    basis, T = basis_for(3, 6, 4)
    off, R, _ = random_batch(2, 2)
    A = O.eval_A(T, off, R)
    AA = O.eval_AA(T, A)
    for i in range(len(T["orders"])):
        ref = np.prod([A[:, T["iAA2iA"][i, t]] for t in range(T["orders"][i])], axis=0) \
            if T["orders"][i] else np.ones(2)
        assert np.max(np.abs(AA[:, i] - ref.real)) < 1e-13


@pytest.mark.parametrize("prop", ["Invariant", "EuclideanVector"])
def test_symmetry_rank_and_model(prop):
    basis, T = basis_for(3, 6, 4, prop=prop)
    off, R, _ = random_batch(1, 3, 10, 10)
    B = O.eval_B(T, O.eval_AA(T, O.eval_A(T, off, R)))[0]
    for _ in range(5):
        Q, _ = np.linalg.qr(RNG.normal(size=(3, 3)))
        if RNG.random() < 0.5:
            Q = -Q
        R2 = (R @ Q.T)[RNG.permutation(len(R))]
        B2 = O.eval_B(T, O.eval_AA(T, O.eval_A(T, off, R2)))[0]
        ref = B if prop == "Invariant" else B @ Q.T
        assert rel_err(B2, ref) < 1e-10           # test_symmbasis.jl:63-67
    if prop == "Invariant":
        M = []
        for s in range(3 * len(basis)):
            o2, R2, _ = random_batch(1, 100 + s, 10, 10)
            M.append(O.eval_B(T, O.eval_AA(T, O.eval_A(T, o2, R2)))[0])
        assert np.linalg.matrix_rank(np.array(M)) == len(basis)   # test_symmbasis.jl:102-110
    c = RNG.normal(size=len(basis)) / len(basis)
    E = O.energy(T, c, off, R)[0]
    assert rel_err(E, np.tensordot(c, B, axes=(0, 0))) < 1e-12   # naive vs product evaluator


@pytest.mark.parametrize("prop", ["Invariant", "EuclideanVector"])
def test_grad_config_fd_and_naive(prop):
    basis, T = basis_for(3, 6, 4, prop=prop)
    off, R, _ = random_batch(1, 4, 8, 8)
    c = RNG.normal(size=len(basis)) / len(basis)
    g = O.grad_config(T, c, off, R)
    h = 1e-5
    for j in range(len(R)):
        for d in range(3):
            Rp, Rm = R.copy(), R.copy()
            Rp[j, d] += h
            Rm[j, d] -= h
            fd = (O.energy(T, c, off, Rp) - O.energy(T, c, off, Rm))[0] / (2 * h)
            got = g[j, d] if prop == "Invariant" else g[j, :, d]
            assert np.max(np.abs(got - fd)) < 1e-7 * max(1.0, np.max(np.abs(g)))
    _, dAA = O.eval_AA_dAA(T, off, R)
    dB = O.eval_dB(T, dAA)
    gn = np.tensordot(c, dB, axes=(0, 1))
    assert rel_err(gn, g) < 1e-12               # naive vs product grad_config


def test_species_basis_invariance():
    basis, T = basis_for(3, 6, 3, species=2)
    off, R, sp = random_batch(1, 5, 10, 10, nspecies=2)
    B = O.eval_B(T, O.eval_AA(T, O.eval_A(T, off, R, sp)))
    Q, _ = np.linalg.qr(RNG.normal(size=(3, 3)))
    p = RNG.permutation(len(R))
    B2 = O.eval_B(T, O.eval_AA(T, O.eval_A(T, off, (R @ Q.T)[p], sp[p])))
    assert rel_err(B2, B) < 1e-10                  # test_discrete.jl:82-87


def test_adjoint_eval_d():
    basis, T = basis_for(3, 6, 4)
    off, R, _ = random_batch(1, 6, 7, 7)
    w = RNG.normal(size=(7, 3))
    c = np.zeros(len(basis))
    dB = O.adjoint_EVAL_D(T, c, off, R, w)
    _, dAA = O.eval_AA_dAA(T, off, R)
    ref = np.einsum("jkd,jd->k", O.eval_dB(T, dAA), w)
    assert rel_err(dB, ref) < 1e-12
