"""Shared test helpers: basis construction cache and tolerance checks."""
import functools

import numpy as np

import acehip as ah
from acehip import synth
from acehip.tables import basis_tables


@functools.lru_cache(maxsize=None)
def basis_for(order, maxdeg, L, prop="Invariant", species=0):
    w = dict(order=order, maxdeg=maxdeg, L=L, prop=prop, species=species)
    b = synth.make_basis(w)
    return b, basis_tables(b)


def rel_err(x, ref):
    x = np.asarray(x)
    ref = np.asarray(ref)
    scale = np.max(np.abs(ref))
    return float(np.max(np.abs(x - ref)) / scale) if scale > 0 else float(np.max(np.abs(x)))


def random_batch(nenv, seed, nmin=5, nmax=20, rmin=0.8, rmax=5.0, nspecies=0):
    rng = np.random.default_rng(seed)
    nX = rng.integers(nmin, nmax + 1, size=nenv)
    off = np.concatenate([[0], np.cumsum(nX)]).astype(np.int32)
    n = int(off[-1])
    g = rng.normal(size=(n, 3))
    R = g / np.linalg.norm(g, axis=1)[:, None] * rng.uniform(rmin, rmax, size=(n, 1))
    sp = rng.integers(0, nspecies, size=n).astype(np.int8) if nspecies else None
    return off, R, sp
