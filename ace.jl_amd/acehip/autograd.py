"""Differentiable site energies for PyTorch callers.

ACE.jl makes ``evaluate(model, cfg)`` differentiable through a
``ChainRules.rrule`` (linearmodel.jl:185-189) whose pullback returns the
parameter gradient ``grad_params(model, cfg) * dp`` and the configuration
gradient ``_rrule_evaluate(dp, model, cfg)`` (linearmodel.jl:176-181,
evaluator.jl:154-200).  A second rrule, on that pullback
(linearmodel.jl:196-226), supplies the mixed derivatives that force training
needs: for a cotangent ``dq`` of the configuration gradient it returns
``adjoint_EVAL_D(model, cfg, dq) * dp`` for the parameters and
``sum_k c_k adjoint_EVAL_D(...)[k]`` for ``dp``, and no tangent for the
configuration.

``site_energies`` is the same pair of rules as ``torch.autograd.Function``s:

* forward: ``ace_energy_forces`` gives the site energies and the per-neighbour
  gradients of a batch in one call;
* first backward: the configuration gradient is ``gE[env(j)] * dE_j``
  (Invariant) or ``sum_p gE[env(j), p] * dE_j[p]`` (EuclideanVector), and the
  parameter gradient is ``sum_env gE[env] . B[env]`` from ``ace_basis_B``;
* second backward (Invariant bases): the parameter part is
  ``sum_env gE[env] * ace_adjoint_EVAL_D(w)[env]`` and the ``gE`` part is
  ``sum_{j in env} w_j . dE_j``, which equals the reference's
  ``sum_k c_k adjoint_EVAL_D(...)[k]``.  Positions get no second-order
  tangent (the position Hessian is not formed), as in the reference.

All evaluation runs in libacehip.so; the contractions over environments are
device tensor operations.
"""
from __future__ import annotations

import numpy as np
import torch

from .evaluator import EnvBatch, HIPEvaluator


def _env_of(offsets: torch.Tensor, nnbr: int) -> torch.Tensor:
    counts = (offsets[1:] - offsets[:-1]).long()
    return torch.repeat_interleave(torch.arange(counts.numel(), device=offsets.device), counts,
                                   output_size=nnbr)


class _Pullback(torch.autograd.Function):
    """Configuration gradient as a function of (c, gE); differentiable once
    more in c and gE (linearmodel.jl:196-226)."""

    @staticmethod
    def forward(ctx, c, gE, ev, b, dE, env_of):
        ctx.ev, ctx.b, ctx.env_of = ev, b, env_of
        ctx.save_for_backward(gE, dE)
        if ev.prop == 0:
            return gE[env_of].unsqueeze(0) * dE
        # dE[3p + d, j] = d phi_p / d r_jd
        return torch.einsum("jp,pdj->dj", gE[env_of], dE.view(3, 3, -1))

    @staticmethod
    def backward(ctx, w):
        gE, dE = ctx.saved_tensors
        ev, b, env_of = ctx.ev, ctx.b, ctx.env_of
        w = w.contiguous()
        g_c = g_gE = None
        if ctx.needs_input_grad[1]:
            if ev.prop == 0:
                per = (w * dE).sum(0)
                g_gE = torch.zeros_like(gE).index_add_(0, env_of, per)
            else:
                per = torch.einsum("dj,pdj->jp", w, dE.view(3, 3, -1))
                g_gE = torch.zeros_like(gE).index_add_(0, env_of, per)
        if ctx.needs_input_grad[0]:
            if ev.prop != 0:
                raise NotImplementedError("mixed second derivatives are provided for Invariant bases only")
            adj = ev.adjoint_EVAL_D(b, w)
            g_c = gE @ adj
        return g_c, g_gE, None, None, None, None


class _SiteEnergy(torch.autograd.Function):
    """Site energies of a batch as a function of (c, xyz)
    (linearmodel.jl:185-189)."""

    @staticmethod
    def forward(ctx, c, xyz, ev, b):
        E, dE = ev.energy_forces(b)
        ctx.ev, ctx.b = ev, b
        ctx.save_for_backward(c, dE)
        return E

    @staticmethod
    def backward(ctx, gE):
        c, dE = ctx.saved_tensors
        ev, b = ctx.ev, ctx.b
        gE = gE.contiguous()
        g_c = g_xyz = None
        if ctx.needs_input_grad[0]:
            B = ev.basis_B(b)
            g_c = gE @ B if ev.prop == 0 else torch.einsum("ep,ekp->k", gE, B)
        if ctx.needs_input_grad[1]:
            g_xyz = _Pullback.apply(c, gE, ev, b, dE, _env_of(b.offsets, b.nnbr))
        return g_c, g_xyz, None, None


def site_energies(ev: HIPEvaluator, c: torch.Tensor, offsets, xyz: torch.Tensor, species=None):
    """Site energies E[nenv] (Invariant) or E[nenv, 3] (EuclideanVector) of the
    environments ``offsets[nenv+1]`` with neighbour vectors ``xyz[3, nnbr]``
    (float64 on the evaluator's device), differentiable in ``c`` and ``xyz``.

    ``c`` is the coefficient vector; when it differs from the evaluator's
    current coefficients it is uploaded (``set_params``) before evaluation.
    """
    if not torch.is_tensor(xyz) or xyz.dtype != torch.float64 or xyz.dim() != 2 or xyz.shape[0] != 3:
        raise ValueError("xyz must be a float64 tensor of shape (3, nnbr)")
    if not torch.is_tensor(c) or c.dtype != torch.float64 or c.shape != (ev.nB,):
        raise ValueError(f"c must be a float64 tensor of shape ({ev.nB},)")
    # upload c when it differs from the coefficients on the device: compared on
    # the device, so the usual case (unchanged c) needs no host round trip
    last = getattr(ev, "_autograd_c", None)
    if last is None or last.shape != c.shape or last.device != c.device or not bool(torch.equal(last, c.detach())):
        ev.set_params(c.detach().cpu().numpy())
        ev._autograd_c = c.detach().clone()
    b = EnvBatch(offsets, xyz.detach(), species, device=xyz.device)
    return _SiteEnergy.apply(c, xyz, ev, b)


class SiteEnergy(torch.nn.Module):
    """A ``LinearACEModel`` as a torch module: the coefficients are a
    parameter, the input is a batch of environments."""

    def __init__(self, model, device: int = 0):
        super().__init__()
        self.ev = HIPEvaluator(model.basis, model.c, device=device)
        self.c = torch.nn.Parameter(torch.as_tensor(np.asarray(model.c, dtype=np.float64),
                                                     device=f"cuda:{device}").clone())

    def forward(self, offsets, xyz, species=None):
        return site_energies(self.ev, self.c, offsets, xyz, species)
