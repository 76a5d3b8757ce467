"""Synthetic environments and the workload definitions.

Directions are normalized Gaussians (ACE.jl ``rand_sphere``,
src/utils/auxiliary.jl:6-9); radii uniform in [rmin, rmax] (``rand_radial``,
src/utils/random.jl:22-23).  All generators use ``numpy.random.default_rng``.
"""
from __future__ import annotations

import numpy as np

from .coupling import EuclideanVector, Invariant
from .onep import Categorical1pBasis, Product1pBasis, RnYlm_1pbasis
from .pibasis import SimpleSparseBasis, SparseBasis
from .symmbasis import symmetric_basis

WORKLOADS = {
    1: dict(name="plain-ace-cpu-runnable", order=3, maxdeg=6, L=4, prop="Invariant",
            nenv=1000, nbr="fixed12", seed=1),
    2: dict(name="plain-ace-batched", order=4, maxdeg=12, L=6, prop="Invariant",
            nenv=100_000, nbr="poisson30", seed=2),
    3: dict(name="energy-forces-fcc", order=3, maxdeg=6, L=4, prop="Invariant",
            nenv=32_768, nbr="fcc", seed=3),
    4: dict(name="two-species-equivariant", order=3, maxdeg=10, L=5, prop="EuclideanVector",
            nenv=65_536, nbr="poisson40", seed=4, species=2),
    5: dict(name="learned-surrogate", order=4, maxdeg=12, L=6, prop="Invariant",
            nenv=1_000_000, nbr="poisson30", seed=5),
    # size workloads beyond the configurations above
    # the model of profile/profile_linearmodel.jl:11-25 (SparseBasis, p = 1,
    # weights n 1, l 1.5; maxL = ceil(14 / 1.5)) on 10 neighbours
    6: dict(name="reference-profiled-model", order=4, maxdeg=14, L=10, prop="Invariant",
            nenv=100_000, nbr="fixed10", seed=6, weight={"n": 1.0, "l": 1.5}),
    # one species, SimpleSparseBasis: 195 pooled values (32-environment tiles)
    7: dict(name="large-pool-invariant", order=4, maxdeg=14, L=6, prop="Invariant",
            nenv=100_000, nbr="poisson30", seed=7),
    # two species: 278 pooled values, 491k AA functions
    8: dict(name="two-species-order4", order=4, maxdeg=12, L=6, prop="Invariant",
            nenv=32_768, nbr="poisson30", seed=8, species=2),
    # correlation order 5 (run interpreter, prefixes of four factors)
    9: dict(name="order-5", order=5, maxdeg=10, L=4, prop="Invariant",
            nenv=100_000, nbr="poisson30", seed=9),
}


def make_basis(w: dict):
    """Basis of a workload: RnYlm_1pbasis(maxdeg, maxL, r0=1, rcut=5) (with a
    leading Categorical1pBasis for two species), SimpleSparseBasis(order,
    maxdeg) -- or SparseBasis(order, p=1, default_maxdeg=maxdeg, weight) when
    the workload gives weights, the 1p spec then selected by the same
    weighted degree --, Invariant or EuclideanVector(isreal=true)."""
    if w.get("weight"):
        Bsel = SparseBasis(w["order"], p=1, default_maxdeg=w["maxdeg"], weight=w["weight"])
        RnYlm = RnYlm_1pbasis(maxdeg=w["maxdeg"], maxL=w["L"], Bsel=Bsel, r0=1.0, rcut=5.0)
    else:
        Bsel = SimpleSparseBasis(w["order"], w["maxdeg"])
        RnYlm = RnYlm_1pbasis(maxdeg=w["maxdeg"], maxL=w["L"], r0=1.0, rcut=5.0)
    if w.get("species"):
        cat = Categorical1pBasis(list(range(w["species"])), varsym="mu", idxsym="mu")
        B1p = Product1pBasis((cat,) + RnYlm.bases)
    else:
        B1p = RnYlm
    prop = Invariant() if w["prop"] == "Invariant" else EuclideanVector(True)
    return symmetric_basis(prop, B1p, Bsel)


def coefficients(nB: int, seed: int):
    """c ~ N(0,1)/len(B)."""
    return np.random.default_rng(seed + 1000).normal(size=nB) / nB


def rand_dirs(rng, n):
    g = rng.normal(size=(n, 3))
    return g / np.linalg.norm(g, axis=1)[:, None]


def random_envs(nenv, nbr, seed, rmin=0.8, rmax=5.0, nspecies=0):
    """Returns (offsets[nenv+1] int32, R[nnbr,3], species[nnbr] int8 | None)."""
    rng = np.random.default_rng(seed)
    if nbr.startswith("fixed"):
        nX = np.full(nenv, int(nbr[len("fixed"):]), dtype=np.int64)
    elif nbr.startswith("poisson"):
        lam = float(nbr[len("poisson"):])
        nX = np.clip(rng.poisson(lam, size=nenv), 5, 80)
    else:
        raise ValueError(nbr)
    off = np.concatenate([[0], np.cumsum(nX)]).astype(np.int32)
    n = int(off[-1])
    R = rand_dirs(rng, n) * rng.uniform(rmin, rmax, size=(n, 1))
    sp = None
    if nspecies:
        sp = rng.integers(0, nspecies, size=n).astype(np.int8)
    return off, R, sp


def fcc_envs(nenv, seed, a=3.6, ncell=7, sigma=0.1, rcut=5.0):
    """Environments cut from displaced FCC snapshots (periodic, minimum image)."""
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(ncell) for j in range(ncell) for k in range(ncell)])
    X0 = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a
    L = ncell * a
    offs, Rs = [0], []
    count = 0
    while count < nenv:
        X = X0 + rng.normal(scale=sigma, size=X0.shape)
        for i in range(len(X)):
            if count >= nenv:
                break
            d = X - X[i]
            d -= L * np.round(d / L)
            r = np.linalg.norm(d, axis=1)
            m = (r > 0) & (r < rcut)
            Rs.append(d[m])
            offs.append(offs[-1] + int(m.sum()))
            count += 1
    return np.asarray(offs, dtype=np.int32), np.concatenate(Rs, axis=0), None


def fcc_crystal(ncells=(30, 30, 28), a=3.6, sigma=0.1, seed=0):
    """Positions (N, 3) and cubic cell (3, 3) of a periodic FCC crystal of
    ``ncells`` conventional cells with Gaussian displacements ``sigma``."""
    basis = 0.5 * a * np.array([[0, 0, 0], [0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=np.float64)
    g = np.stack(np.meshgrid(*[np.arange(k) for k in ncells], indexing="ij"), -1).reshape(-1, 3)
    x = (g[:, None, :] * a + basis[None]).reshape(-1, 3)
    x = x + sigma * np.random.default_rng(seed).standard_normal(x.shape)
    return x, np.diag(np.asarray(ncells, dtype=np.float64) * a)


def workload_envs(wid: int, nenv: int | None = None, seed: int | None = None):
    w = WORKLOADS[wid]
    nenv = w["nenv"] if nenv is None else nenv
    seed = w["seed"] if seed is None else seed
    if w["nbr"] == "fcc":
        return fcc_envs(nenv, seed)
    return random_envs(nenv, w["nbr"], seed, nspecies=w.get("species", 0))
