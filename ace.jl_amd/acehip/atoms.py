"""Atoms-level layer around the evaluator: environments from positions and
atomic forces from per-neighbour gradients, both on the device.

ACE.jl evaluates one environment at a time; its callers evaluate site-energy
models on whole configurations: ``energy(V, at)`` and ``forces(V, at)`` in
``src/testing/testlsq.jl:27-35,42-69``, with forces ``-dE/dx`` assembled
from pair terms that enter both atoms with opposite signs
(``src/testing/testmodel.jl:49-81``).  :class:`ACECalculator` offers that
interface over :class:`~acehip.evaluator.HIPEvaluator`:
:class:`NeighbourList` extracts the environments (``ace_nbrlist_build``,
``ace_nbrlist_update``) and assembles forces and the virial
(``ace_nbrlist_forces``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .evaluator import EnvBatch, HIPEvaluator


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


@dataclass
class Atoms:
    """A configuration: ``positions`` (N, 3), ``cell`` (3, 3) with the
    lattice vectors as rows, ``pbc`` (3,) periodicity, ``species`` (N,)
    category indices of a categorical 1p basis (or None)."""
    positions: np.ndarray
    cell: np.ndarray
    pbc: tuple = (True, True, True)
    species: np.ndarray | None = None

    def __len__(self):
        return len(self.positions)


def _pos_soa(pos, device):
    """(N, 3) array or tensor, or (3, N) device tensor -> (3, N) float64 device."""
    if torch.is_tensor(pos) and pos.is_cuda and pos.dim() == 2 and pos.shape[0] == 3 and pos.is_contiguous() \
            and pos.dtype == torch.float64:
        return pos
    t = pos if torch.is_tensor(pos) else torch.as_tensor(np.asarray(pos, dtype=np.float64))
    if t.dim() != 2 or t.shape[-1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    return t.to(device=device, dtype=torch.float64).t().contiguous()


class NeighbourList:
    """Device-resident neighbour lists of one configuration.

    After :meth:`build`: ``offsets`` [natoms+1] int32, ``nbr_j`` [nnbr] int32,
    ``shift`` (3, nnbr) int32, ``xyz`` (3, nnbr) float64, ``rev`` [nnbr]
    int32 and ``species`` [nnbr] int8 (or None), all views into buffers that
    grow as needed and are reused by later builds.
    """

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.device = device
        self.dev = torch.device("cuda", device)
        h = C.c_void_p()
        _lib.check(self.lib.ace_nbrlist_create(device, C.byref(h)))
        self.h = h
        self.cap = 0
        self.natoms = 0
        self.nnbr = 0
        self.cell = np.eye(3)
        self._alloc(1024)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ace_nbrlist_destroy(self.h)
            self.h = None

    def _alloc(self, cap):
        d = self.dev
        self.cap = cap
        self._j = torch.empty(cap, dtype=torch.int32, device=d)
        self._shift = torch.empty(3 * cap, dtype=torch.int32, device=d)
        self._xyz = torch.empty(3 * cap, dtype=torch.float64, device=d)
        self._rev = torch.empty(cap, dtype=torch.int32, device=d)
        self._sp = torch.empty(cap, dtype=torch.int8, device=d)

    def _stream(self, stream):
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        return C.c_void_p(stream.cuda_stream)

    def build(self, positions, cell, pbc, rcut: float, species=None, stream=None):
        pos = _pos_soa(positions, self.dev)
        n = pos.shape[1]
        cellc = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3))
        pbcc = np.ascontiguousarray(np.asarray(pbc, dtype=bool).astype(np.int32).reshape(3))
        zat = None
        if species is not None:
            zat = (species if torch.is_tensor(species) else torch.as_tensor(np.asarray(species, dtype=np.int8)))
            zat = zat.to(device=self.dev, dtype=torch.int8).contiguous()
            if zat.numel() != n:
                raise ValueError("species must have one entry per atom")
        self.offsets = torch.empty(n + 1, dtype=torch.int32, device=self.dev)
        need = C.c_int64(0)
        for _ in range(2):
            rc = self.lib.ace_nbrlist_build(
                self.h, n, _vp(pos), cellc.ctypes.data_as(C.POINTER(C.c_double)),
                pbcc.ctypes.data_as(C.POINTER(C.c_int32)), float(rcut), _vp(zat), self.cap, _vp(self.offsets),
                _vp(self._j), _vp(self._shift), _vp(self._xyz), _vp(self._rev),
                _vp(self._sp) if zat is not None else C.c_void_p(0), C.byref(need), self._stream(stream))
            if rc != _lib.ERR_CAPACITY:
                break
            self._alloc(int(need.value * 1.25) + 1024)
        _lib.check(rc)
        m = int(need.value)
        self.natoms, self.nnbr, self.rcut = n, m, float(rcut)
        self.cell, self.pbc = cellc, pbcc
        self.nbr_j = self._j[:m]
        self.shift = self._shift[:3 * m].view(3, m)
        self.xyz = self._xyz[:3 * m].view(3, m)
        self.rev = self._rev[:m]
        self.species = self._sp[:m] if zat is not None else None
        return self

    def update(self, positions, cell=None, stream=None):
        """New relative positions for the same pairs (x_j - x_i + S' . cell)."""
        pos = _pos_soa(positions, self.dev)
        if pos.shape[1] != self.natoms:
            raise ValueError("update needs the atoms of the last build")
        if cell is not None:
            self.cell = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3))
        _lib.check(self.lib.ace_nbrlist_update(
            self.h, self.natoms, _vp(pos), self.cell.ctypes.data_as(C.POINTER(C.c_double)), self.nnbr,
            _vp(self.offsets), _vp(self.nbr_j), _vp(self.shift), _vp(self.xyz), self._stream(stream)))
        return self

    def batch(self) -> EnvBatch:
        """The environments as an evaluator batch (no copy)."""
        return EnvBatch(self.offsets, self.xyz, self.species, device=self.dev)

    def forces(self, dE_dr, F=None, virial=False, stream=None):
        """F (3, natoms) = -dE/dx from dE_dr (3, nnbr); with ``virial`` also
        W (3, 3) = -sum_m dE_dr_m r_m^T."""
        if tuple(dE_dr.shape) != (3, self.nnbr):
            raise ValueError("dE_dr must have shape (3, nnbr) of this list")
        if F is None:
            F = torch.empty((3, self.natoms), dtype=torch.float64, device=self.dev)
        W = torch.empty((3, 3), dtype=torch.float64, device=self.dev) if virial else None
        _lib.check(self.lib.ace_nbrlist_forces(self.h, self.natoms, _vp(self.offsets), _vp(self.rev), self.nnbr,
                                               _vp(self.xyz), _vp(dE_dr), _vp(F), _vp(W), self._stream(stream)))
        return (F, W) if virial else F

    def forces_multi(self, g, stream=None):
        """Forces of many fields at once: g (nnbr, nvec, 3) -> F (natoms, nvec, 3)."""
        if g.dim() != 3 or g.shape[0] != self.nnbr or g.shape[2] != 3:
            raise ValueError("g must have shape (nnbr, nvec, 3) of this list")
        g = g.contiguous()
        F = torch.empty((self.natoms, g.shape[1], 3), dtype=torch.float64, device=self.dev)
        _lib.check(self.lib.ace_nbrlist_forces_multi(self.h, self.natoms, _vp(self.offsets), _vp(self.rev),
                                                     self.nnbr, g.shape[1], _vp(g), _vp(F),
                                                     self._stream(stream)))
        return F


class ACECalculator:
    """Site-energy model on configurations: total energy, forces, virial.

    ``skin`` > 0 keeps the neighbour list between calls and only updates the
    relative positions until some atom has moved more than skin / 2 since the
    last build (or the cell, periodicity or atom count changes).
    """

    def __init__(self, model: HIPEvaluator, rcut: float | None = None, skin: float = 0.0):
        if model.prop != 0:
            raise ValueError("ACECalculator needs an Invariant (site-energy) model")
        self.model = model
        self.rcut = float(rcut if rcut is not None else model.pack.T["rn"]["rcut"])
        self.skin = float(skin)
        self.nl = NeighbourList(model.device)
        self._ref = None
        self.rebuilds = 0

    def _list(self, at: Atoms):
        pos = _pos_soa(at.positions, self.nl.dev)
        cell = np.asarray(at.cell, dtype=np.float64).reshape(3, 3)
        pbc = np.asarray(at.pbc, dtype=bool).astype(np.int32)
        fresh = (self._ref is None or self.skin <= 0 or pos.shape[1] != self.nl.natoms
                 or not np.array_equal(cell, self.nl.cell) or not np.array_equal(pbc, self.nl.pbc))
        if not fresh:
            disp = (pos - self._ref).square().sum(0).max().sqrt().item() if pos.shape[1] else 0.0
            fresh = disp > 0.5 * self.skin
        if fresh:
            self.nl.build(pos, cell, pbc, self.rcut + self.skin, at.species)
            self._ref = pos.clone()
            self.rebuilds += 1
        else:
            self.nl.update(pos)
        return self.nl

    def site_energies(self, at: Atoms):
        nl = self._list(at)
        return self.model.energy(nl.batch())

    def energy(self, at: Atoms) -> float:
        return float(self.site_energies(at).sum().item())

    def energy_forces(self, at: Atoms, virial: bool = False):
        """(E total, F (N, 3)[, W (3, 3)]) as NumPy."""
        nl = self._list(at)
        E, dE = self.model.energy_forces(nl.batch())
        out = nl.forces(dE, virial=virial)
        F, W = out if virial else (out, None)
        res = (float(E.sum().item()), F.t().cpu().numpy())
        return res + (W.cpu().numpy(),) if virial else res

    def forces(self, at: Atoms):
        return self.energy_forces(at)[1]

    # -- basis level: energy(basis, at) and forces(basis, at) of least-squares
    #    fitting (src/testing/testlsq.jl:50-61)
    def basis_energy(self, at: Atoms):
        """sum_i B(env_i): (nB,) device tensor."""
        nl = self._list(at)
        return self.model.basis_B(nl.batch()).sum(0)

    def basis_energy_forces(self, at: Atoms):
        """(sum_i B(env_i) (nB,), per-basis-function forces (nB, N, 3)) as
        device tensors; the forces are -d/dx of the summed basis."""
        nl = self._list(at)
        b = nl.batch()
        EB = self.model.basis_B(b).sum(0)
        if nl.nnbr == 0:
            return EB, torch.zeros((len(EB), nl.natoms, 3), dtype=torch.float64, device=nl.dev)
        dB = self.model.basis_dB(b)
        return EB, nl.forces_multi(dB).permute(1, 0, 2)

    def virial(self, at: Atoms):
        return self.energy_forces(at, virial=True)[2]
