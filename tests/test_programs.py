"""Device run programs (the compiled product DAG) emulated on the CPU and
checked against the oracle's energy and adjoint."""
import numpy as np
import pytest

import ace_oracle as O
from acehip import programs as PR, synth
from acehip.pack import Pack
from helpers import basis_for, random_batch, rel_err


def emulate(pk, Aplus, ct):
    vec = pk.prop == 1
    nenv = Aplus.shape[0]

    def fac(f):
        v = Aplus[:, f & 0x7fff]
        return np.conj(v) if f & 0x8000 else v

    lc = [PR.eval_terms(t, ct) for t in pk.prog.leaf_terms]
    tc = [PR.eval_terms(t, ct) for t in pk.prog.task_terms]
    NC = 3 if vec else 1
    E = np.zeros((nenv, NC))
    G = np.zeros((nenv, max(1, pk.ngslots), NC), dtype=complex)
    for ti, (r0, r1, slot) in enumerate(pk.tasks[:, :3]):
        acc = np.zeros((nenv, NC), dtype=complex) + np.asarray(tc[ti]).reshape(1, -1)
        P1 = P2 = P3 = np.ones(nenv, dtype=complex)
        for r in range(r0, r1):
            d, j, nl, lo, f1, f2, f3 = pk.runs[r][:7]
            if j <= 1 and d >= 1:
                P1 = fac(f1)
            if j <= 2 and d >= 2:
                P2 = P1 * fac(f2)
            if j <= 3 and d >= 3:
                P3 = P2 * fac(f3)
            Pd = [np.ones(nenv), P1, P2, P3][d]
            S = np.zeros((nenv, NC), dtype=complex)
            for i in range(nl):
                S += np.asarray(lc[lo + i]).reshape(1, -1) * fac(pk.leaf_fid[lo + i])[:, None]
            acc += Pd[:, None] * S
        if slot < 0:
            E += acc.real
        else:
            G[:, slot] = acc
    Gk = np.zeros((nenv, len(pk.pool), NC), dtype=complex)
    for k in range(len(pk.pool)):
        for s in pk.gslot_list[pk.gslot_ptr[k]:pk.gslot_ptr[k + 1]]:
            Gk[:, k] += G[:, s]
    return E, Gk


@pytest.mark.parametrize("args", [(3, 6, 4, "Invariant", 0), (3, 6, 4, "EuclideanVector", 0),
                                  (3, 6, 3, "Invariant", 2), (4, 8, 4, "Invariant", 0)])
def test_programs_match_oracle(args):
    basis, T = basis_for(*args)
    nsp = args[4]
    c = synth.coefficients(len(basis), 7)
    pk = Pack(T)
    ct = O.ctilde(T, c)
    off, R, sp = random_batch(4, 9, nspecies=nsp)
    A, dphi = O.eval_A_dA(T, off, R, sp)
    Aplus = np.zeros((4, len(pk.pool)), dtype=complex)
    for iA in range(len(T["aspec"])):
        f = pk.aspec_fid[iA]
        if not f & 0x8000:
            Aplus[:, f & 0x7fff] = A[:, iA]
    E, Gk = emulate(pk, Aplus, ct)
    vec = args[3] != "Invariant"
    e0 = 0.0 if vec else sum(ct[r] for r in range(len(T["orders"])) if T["orders"][r] == 0)
    Eref = O.energy(T, c, off, R, sp)
    assert rel_err(E if vec else E[:, 0] + e0, Eref) < 1e-13
    dAco = O._dAco(T, ct, A)
    Gref = np.zeros_like(Gk)
    for iA in range(len(T["aspec"])):
        f = pk.aspec_fid[iA]
        k, s = f & 0x7fff, pk.aspec_sgn[iA]
        D = dAco[:, iA, :] if vec else dAco[:, iA][:, None]
        Gref[:, k] += s * np.conj(D) if f & 0x8000 else D
    # G[k] is fixed by dE = Re sum_k G[k] dA+[k]: completely where A+[k] is
    # complex (m != 0), only its real part where A+[k] is real (m = 0)
    m0 = pk.pool[:, 3] == 0
    assert rel_err(Gk[:, ~m0], Gref[:, ~m0]) < 1e-13
    assert rel_err(Gk[:, m0].real, Gref[:, m0].real) < 1e-13
    # forces assembled from G equal grad_config: g_j = Re sum_k G[k] dA+[k]/dr_j
    env_of = np.repeat(np.arange(4), np.diff(off))
    g = np.zeros((len(R), 3) if not vec else (len(R), 3, 3))
    for iA in range(len(T["aspec"])):
        f = pk.aspec_fid[iA]
        if f & 0x8000:
            continue
        k = f & 0x7fff
        if vec:
            g += np.einsum("jp,jd->jpd", Gk[env_of, k, :], dphi[:, iA, :]).real
        else:
            g += (Gk[env_of, k, 0][:, None] * dphi[:, iA, :]).real
    assert rel_err(g, O.grad_config(T, c, off, R, sp)) < 1e-12


def test_task_tables_consistent():
    basis, T = basis_for(4, 8, 4)
    pk = Pack(T)
    assert pk.task_term_ptr[0] == pk.leaf_term_ptr[-1]
    assert pk.task_term_ptr[-1] == len(pk.term_row)
    for W in pk.wave_cfg_W:
        assert pk.wave_load[int(W)][0] >= pk.wave_load[int(W)][1] / W
    # every G slot is produced by exactly one task and owned by one pooled index
    slots = sorted(int(s) for s in pk.tasks[:, 2] if s >= 0)
    assert slots == list(range(pk.ngslots))
    assert sorted(pk.gslot_list.tolist()) == list(range(pk.ngslots))
