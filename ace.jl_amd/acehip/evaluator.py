"""Device evaluator and the ACE.jl-shaped host API.

``HIPEvaluator`` occupies the evaluator slot of ACE.jl's ``LinearACEModel``
(linearmodel.jl:44-56, 107-111): ``evaluate(model, cfg)`` and
``grad_config(model, cfg)`` dispatch to it, ``set_params!`` forwards new
coefficients (linearmodel.jl:67-71).  Basis-level ``evaluate`` /
``evaluate_d`` of Product1pBasis, PIBasis and SymmetricBasis run on the same
context.  Batches of environments are the native unit; a single
configuration is a batch of one.

Device memory and streams come from PyTorch; every computation runs in
libacehip.so.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .pack import Pack
from .tables import basis_tables


class EnvBatch:
    """Ragged batch of environments resident on the device.

    ``offsets[nenv+1]`` int32, ``xyz[3, nnbr]`` float64 SoA, ``species[nnbr]``
    int8 (category index) or None.
    """

    def __init__(self, offsets, xyz, species=None, device="cuda"):
        self.offsets = torch.as_tensor(np.asarray(offsets, dtype=np.int32) if not torch.is_tensor(offsets) else offsets,
                                       dtype=torch.int32, device=device).contiguous()
        x = xyz if torch.is_tensor(xyz) else torch.as_tensor(np.asarray(xyz, dtype=np.float64))
        x = x.to(device=device, dtype=torch.float64)
        if x.dim() != 2 or x.shape[0] != 3:
            raise ValueError("xyz must be SoA with shape (3, nnbr)")
        self.xyz = x.contiguous()
        self.species = None
        if species is not None:
            s = species if torch.is_tensor(species) else torch.as_tensor(np.asarray(species, dtype=np.int8))
            self.species = s.to(device=device, dtype=torch.int8).contiguous()
        self.nenv = self.offsets.numel() - 1
        self.nnbr = self.xyz.shape[1]

    @classmethod
    def from_configs(cls, cfgs, species=None, device="cuda"):
        """``cfgs``: list of (nX, 3) arrays; ``species``: list of (nX,) ints."""
        lens = [len(c) for c in cfgs]
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        R = np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 3) for c in cfgs], axis=0) \
            if cfgs else np.zeros((0, 3))
        sp = None
        if species is not None:
            sp = np.concatenate([np.asarray(s, dtype=np.int8) for s in species]) if cfgs else np.zeros(0, np.int8)
        return cls(off, R.T.copy(), sp, device=device)


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class HIPEvaluator:
    """Context of libacehip.so for one SymmetricBasis and coefficients.

    ``compiler="native"`` (default) hands the library the basis in ACE.jl's
    own layout (``ace_ctx_create_jl``: Product1pBasis indices, PIBasisSpec,
    A2Bmap as CSC), and the library compiles the device programs itself, as
    it does for a Julia caller.  ``compiler="python"`` builds the same tables
    with the Python packer (:class:`acehip.pack.Pack`) and passes them to
    ``ace_ctx_create``; both give identical tables."""

    def __init__(self, basis, c, device: int = 0, waves: int = 8, compiler: str = "native"):
        self.lib = _lib.load()
        self.basis = basis
        self.T = basis_tables(basis)
        self._pack = None
        self.prop = 0 if self.T["prop"] == "Invariant" else 1
        self.nB = int(self.T["a2b"]["shape"][0])
        self.nA = len(self.T["aspec"])
        self.nAA = len(self.T["orders"])
        # pooled (m >= 0) 1p functions: distinct (mu, n, l, |m|)
        self.npool = len({(int(a[0]), int(a[1]), int(a[2]), abs(int(a[3]))) for a in self.T["aspec"]})
        self.device = device
        c = np.ascontiguousarray(c, dtype=np.float64)
        ctx = C.c_void_p()
        if compiler == "native":
            struct, keep = _lib.make_jl_basis(self.T)
            _lib.check(self.lib.ace_ctx_create_jl(C.byref(struct), device, c.ctypes.data_as(_lib._f64p),
                                                  len(c), C.byref(ctx)))
        elif compiler == "python":
            struct, keep = _lib.make_pack(self.pack)
            _lib.check(self.lib.ace_ctx_create(C.byref(struct), device, c.ctypes.data_as(_lib._f64p),
                                               len(c), C.byref(ctx)))
        else:
            raise ValueError("compiler must be 'native' or 'python'")
        self.ctx = ctx
        self.set_tile(waves)

    @property
    def pack(self):
        """The Python packer's tables of this basis (built on first use)."""
        if self._pack is None:
            self._pack = Pack(self.T)
        return self._pack

    def last_dag_shape(self):
        """Product-DAG launch of the last call: (parts per tile, tail tiles,
        kernel: 0 / 1 the record kernel on 64 / 32-environment tiles, else the
        run interpreter's environments per tile)."""
        p, t, k = C.c_int(), C.c_int64(), C.c_int()
        _lib.check(self.lib.ace_last_dag_shape(self.ctx, C.byref(p), C.byref(t), C.byref(k)))
        return p.value, t.value, k.value

    def __del__(self):
        ctx = getattr(self, "ctx", None)
        if ctx:
            self.lib.ace_ctx_destroy(ctx)
            self.ctx = None

    def set_tile(self, waves: int):
        _lib.check(self.lib.ace_set_tile(self.ctx, int(waves)))
        self.waves = int(waves)

    def set_parts(self, parts: int = 0):
        """Workgroups per product-DAG tile: 0 chooses per call (default), else 1 to 4."""
        _lib.check(self.lib.ace_set_parts(self.ctx, int(parts)))

    def set_pool_shape(self, chunk: int = 0, envs_per_wave: int = 0):
        """Pooling launch: neighbours per chunk (8, 16, 32; 0 = 16) and
        environments per wave (0 automatic, 1 or 2; two only for bases of at
        most 32 pooled values)."""
        _lib.check(self.lib.ace_set_pool_shape(self.ctx, int(chunk), int(envs_per_wave)))

    def set_kernel(self, mode: int = 0):
        """Product-DAG kernel: 0 the record-segment kernel where the basis
        allows it (default), 1 the run interpreter always, 2 the per-record
        dispatch kernel over the same records (comparison)."""
        _lib.check(self.lib.ace_set_kernel(self.ctx, int(mode)))

    def set_tail_parts(self, ntail: int = -1):
        """Tiles at the end of the product-DAG grid that run one more part:
        -1 chooses per call (automatic parts only), else that many."""
        _lib.check(self.lib.ace_set_tail_parts(self.ctx, int(ntail)))

    def set_validation(self, on: bool):
        _lib.check(self.lib.ace_set_validation(self.ctx, int(bool(on))))

    def set_params(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        _lib.check(self.lib.ace_set_params(self.ctx, c.ctypes.data_as(_lib._f64p), len(c)))
        self._autograd_c = None  # the coefficient cache of acehip.autograd no longer matches

    def _stream(self, stream):
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        return C.c_void_p(stream.cuda_stream)

    # -- model level
    def energy(self, b: EnvBatch, out=None, stream=None):
        nc = 1 if self.prop == 0 else 3
        if out is None:
            out = torch.empty((b.nenv,) if nc == 1 else (b.nenv, 3), dtype=torch.float64,
                              device=b.xyz.device)
        _lib.check(self.lib.ace_energy(self.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz), _vp(b.species),
                                       _vp(out), self._stream(stream)))
        return out

    def energy_forces(self, b: EnvBatch, E=None, dE=None, stream=None):
        nc = 1 if self.prop == 0 else 3
        if E is None:
            E = torch.empty((b.nenv,) if nc == 1 else (b.nenv, 3), dtype=torch.float64,
                            device=b.xyz.device)
        if dE is None:
            dE = torch.empty((3 * nc, b.nnbr), dtype=torch.float64, device=b.xyz.device)
        _lib.check(self.lib.ace_energy_forces(self.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz),
                                              _vp(b.species), _vp(E), _vp(dE), self._stream(stream)))
        return E, dE

    def energy_forces_noalloc(self, b: EnvBatch, E, dE, stream=None):
        """``ace_energy_forces_noalloc``: the same results as energy_forces,
        with the library allocating nothing (its scratch is a fixed device
        array; too large a batch raises).  E and dE are the caller's."""
        _lib.check(self.lib.ace_energy_forces_noalloc(self.ctx, b.nenv, _vp(b.offsets), b.nnbr, _vp(b.xyz),
                                                      _vp(b.species), _vp(E), _vp(dE), self._stream(stream)))
        return E, dE

    def last_timing(self):
        a, b_, c = C.c_double(), C.c_double(), C.c_double()
        _lib.check(self.lib.ace_last_timing(self.ctx, C.byref(a), C.byref(b_), C.byref(c)))
        return {"dag_ms": a.value, "pool_ms": b_.value, "force_ms": c.value}

    # -- basis level
    def basis_A(self, b: EnvBatch, stream=None):
        out = torch.empty((b.nenv, self.nA, 2), dtype=torch.float64, device=b.xyz.device)
        _lib.check(self.lib.ace_basis_A(self.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz), _vp(b.species),
                                        _vp(out), self._stream(stream)))
        return torch.view_as_complex(out)

    def basis_AA(self, b: EnvBatch, stream=None):
        nAA = self.nAA
        if self.prop == 0:
            out = torch.empty((b.nenv, nAA), dtype=torch.float64, device=b.xyz.device)
        else:
            out = torch.empty((b.nenv, nAA, 2), dtype=torch.float64, device=b.xyz.device)
        _lib.check(self.lib.ace_basis_AA(self.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz), _vp(b.species),
                                         _vp(out), self._stream(stream)))
        return out if self.prop == 0 else torch.view_as_complex(out)

    def basis_B(self, b: EnvBatch, stream=None):
        shape = (b.nenv, self.nB) if self.prop == 0 else (b.nenv, self.nB, 3)
        out = torch.empty(shape, dtype=torch.float64, device=b.xyz.device)
        _lib.check(self.lib.ace_basis_B(self.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz), _vp(b.species),
                                        _vp(out), self._stream(stream)))
        return out

    def basis_dB(self, b: EnvBatch, stream=None):
        shape = (b.nnbr, self.nB, 3) if self.prop == 0 else (b.nnbr, self.nB, 3, 3)
        out = torch.empty(shape, dtype=torch.float64, device=b.xyz.device)  # every entry is written
        _lib.check(self.lib.ace_basis_dB(self.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz), _vp(b.species),
                                         _vp(out), self._stream(stream)))
        return out

    def adjoint_EVAL_D(self, b: EnvBatch, w, stream=None):
        """dB[nenv, nB] = sum_j w_j . dB/dr_j for weights w (3, nnbr)
        (evaluator.jl:204-244), Invariant bases."""
        w = w.to(device=b.xyz.device, dtype=torch.float64).contiguous()
        if tuple(w.shape) != (3, b.nnbr):
            raise ValueError("w must have shape (3, nnbr)")
        out = torch.empty((b.nenv, self.nB), dtype=torch.float64, device=b.xyz.device)
        _lib.check(self.lib.ace_adjoint_EVAL_D(self.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz), _vp(b.species),
                                               _vp(w), _vp(out), self._stream(stream)))
        return out


# ---------------------------------------------------------------------------
# ACE.jl-shaped functional API


def _as_batch(cfg, species=None):
    if isinstance(cfg, EnvBatch):
        return cfg, False
    arr = np.asarray(cfg, dtype=np.float64)
    if arr.ndim == 2 and arr.shape[1] == 3:
        return EnvBatch.from_configs([arr], None if species is None else [species]), True
    raise TypeError("cfg must be an (nX, 3) array of neighbour positions or an EnvBatch")


def _ev(model):
    if getattr(model, "_ev", None) is None:
        model._ev = HIPEvaluator(model.basis, model.c)
    return model._ev


def evaluate(obj, cfg, species=None):
    """``evaluate(model, cfg)`` -> site energy (linearmodel.jl:107-108);
    ``evaluate(basis::SymmetricBasis, cfg)`` -> B (symmbasis.jl:297-302);
    PIBasis -> AA; Product1pBasis -> A."""
    from .symmbasis import LinearACEModel, SymmetricBasis
    from .pibasis import PIBasis

    b, single = _as_batch(cfg, species)
    if isinstance(obj, LinearACEModel):
        E = _ev(obj).energy(b).cpu().numpy()
        return E[0] if single else E
    if isinstance(obj, SymmetricBasis):
        ev = _basis_ev(obj)
        B = ev.basis_B(b).cpu().numpy()
        return B[0] if single else B
    if isinstance(obj, PIBasis):
        ev = _basis_ev(obj._symm)
        AA = ev.basis_AA(b).cpu().numpy()
        return AA[0] if single else AA
    raise TypeError("evaluate: unsupported object")


def _basis_ev(basis):
    if getattr(basis, "_ev", None) is None:
        basis._ev = HIPEvaluator(basis, np.zeros(len(basis)))
    return basis._ev


def grad_config(model, cfg, species=None):
    """``grad_config(model, cfg)``: +dE/dr_j per neighbour (evaluator.jl:150-200)."""
    b, single = _as_batch(cfg, species)
    _, dE = _ev(model).energy_forces(b)
    g = dE.cpu().numpy()
    if model._ev.prop == 0:
        return g.T.copy()
    return g.reshape(3, 3, -1).transpose(2, 0, 1).copy()


def grad_params(model, cfg, species=None):
    """``grad_params`` = B (linearmodel.jl:114-123)."""
    return evaluate(model.basis, cfg, species)


def adjoint_EVAL_D(model, cfg, w, species=None):
    """``adjoint_EVAL_D(m, cfg, w)`` (linearmodel.jl:133-134 ->
    evaluator.jl:204-244): sum_j w_j . dB/dr_j for one configuration (w: nX x 3)."""
    b, single = _as_batch(cfg, species)
    wt = torch.as_tensor(np.asarray(w, dtype=np.float64).reshape(-1, 3).T.copy())
    out = _ev(model).adjoint_EVAL_D(b, wt).cpu().numpy()
    return out[0] if single else out


def evaluate_d(basis, cfg, species=None):
    """``evaluate_d(basis::SymmetricBasis, cfg)``: dB[nX, nB, 3]
    (symmbasis.jl:322-336)."""
    b, _ = _as_batch(cfg, species)
    return _basis_ev(basis).basis_dB(b).cpu().numpy()
