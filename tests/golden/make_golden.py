"""Writes the golden fixtures for the config-1 basis (order 3, degree 6,
Lmax 4, r0 = 1, rcut = 5, Invariant) and a small EuclideanVector basis:
shared basis tables, 16 environments, coefficients and the oracle's
energies, forces (+dE/dr), B and dB.  The oracle itself is pinned by closed
forms, symmetry and finite differences (tests/test_oracle.py); these vectors
pin regressions of the oracle, the tables and the device.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ace.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import ace_oracle as O  # noqa: E402
from acehip import synth  # noqa: E402
from acehip.tables import basis_tables, dump_tables  # noqa: E402


def make(name, w, nenv, wid):
    basis = synth.make_basis(w)
    T = basis_tables(basis)
    c = synth.coefficients(len(basis), 1)
    off, R, sp = synth.workload_envs(wid, nenv=nenv)
    E = O.energy(T, c, off, R, sp)
    g = O.grad_config(T, c, off, R, sp)
    A = O.eval_A(T, off, R, sp)
    AA = O.eval_AA(T, A)
    B = O.eval_B(T, AA)
    o4 = off[:5]
    _, dAA = O.eval_AA_dAA(T, o4, R[: o4[-1]], None if sp is None else sp[: o4[-1]])
    dB = O.eval_dB(T, dAA)
    dump_tables(T, os.path.join(HERE, f"{name}_tables.json"))
    np.savez_compressed(os.path.join(HERE, f"{name}_vectors.npz"), offsets=off, R=R, c=c, E=E, g=g,
                        B=B, dB=dB, A=A)


if __name__ == "__main__":
    make("config1", synth.WORKLOADS[1], 16, 1)
    wv = dict(order=3, maxdeg=6, L=4, prop="EuclideanVector")
    make("vector_deg6", wv, 8, 1)
    print("written")
