"""Least-squares fitting of a linear ACE model on the device.

Mirrors ACE.jl's fitting helpers: ``get_V0`` (one-body energies per species
from energies and species counts, ``src/testing/testlsq.jl:3-24``) and
``lsq`` (``src/testing/testlsq.jl:42-69``).  The design matrix holds, per
configuration, the energy row ``wE * energy(basis, at) / N`` and the force rows
``wF * forces(basis, at)``, in atom-major order (x, y, z of atom 1, then atom 2,
as ``mat(F)[:]``).  Here the rows come from :class:`~acehip.atoms.ACECalculator`
(basis values and per-basis-function forces on the device).  The system is
solved by QR in FP64 on the device.
"""
from __future__ import annotations

import numpy as np
import torch

from .atoms import ACECalculator, Atoms


def get_V0(train, nspecies: int = 1):
    """E0 per species category minimising sum_t (E_t - sum_z n_tz E0_z)^2
    (testlsq.jl:3-24); ``train`` is a list of (Atoms, E, F)."""
    A = np.zeros((len(train), nspecies))
    y = np.zeros(len(train))
    for it, (at, E, _) in enumerate(train):
        y[it] = E
        z = np.zeros(len(at), dtype=np.int64) if at.species is None else np.asarray(at.species, dtype=np.int64)
        A[it] = np.bincount(z, minlength=nspecies)
    return np.linalg.lstsq(A, y, rcond=None)[0]


def design_matrix(train, calc: ACECalculator, wE: float = 1.0, wF: float = 1.0):
    """(A, y) on the device: energy row and 3N force rows per configuration
    (testlsq.jl:44-65)."""
    dev = calc.nl.dev
    nB = calc.model.pack.nB
    nobs = sum(1 + 3 * len(at) for at, _, _ in train)
    A = torch.zeros((nobs, nB), dtype=torch.float64, device=dev)
    y = torch.zeros(nobs, dtype=torch.float64, device=dev)
    row = 0
    for at, E, F in train:
        n = len(at)
        EB, FB = calc.basis_energy_forces(at)
        A[row] = wE * EB / n
        y[row] = wE * float(E) / n
        row += 1
        A[row:row + 3 * n] = wF * FB.permute(1, 2, 0).reshape(3 * n, nB)
        y[row:row + 3 * n] = wF * torch.as_tensor(np.asarray(F, dtype=np.float64).reshape(-1), device=dev)
        row += 3 * n
    return A, y


def lsq(train, calc: ACECalculator, wE: float = 1.0, wF: float = 1.0):
    """Coefficients c (NumPy) and the relative RMSE |A c - y| / |y| of the QR
    least-squares solution (testlsq.jl:42-69)."""
    A, y = design_matrix(train, calc, wE, wF)
    Q, R = torch.linalg.qr(A)
    c = torch.linalg.solve_triangular(R, (Q.T @ y)[:, None], upper=True)[:, 0]
    relrmse = float(torch.linalg.norm(A @ c - y) / torch.linalg.norm(y))
    return c.cpu().numpy(), relrmse


def training_set(calc: ACECalculator, configs):
    """(Atoms, E, F) triples labelled by the calculator's model (testlsq.jl:27-35)."""
    out = []
    for at in configs:
        E, F = calc.energy_forces(at)
        out.append((at, E, F))
    return out


__all__ = ["Atoms", "get_V0", "design_matrix", "lsq", "training_set"]
