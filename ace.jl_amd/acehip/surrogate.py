"""Learned surrogate for the correlation-order-3/4 share of the site energy.

No reference counterpart (ACE.jl is a linear model).  Inputs are the
order-1 and order-2 invariants of the same SymmetricBasis (the B functions
whose AA columns have order 1 or 2); the exact order <= 2 energy share
``E2 = c0 + sum_f c_f X_f`` (c0 the order-0 constant) is kept in FP64.  An ensemble of MLPs
(libacemlp.so, BF16 MFMA, FP32 accumulate) predicts the normalized remainder
``E34 = E - E2``; plain FP64 ACE (libacehip.so) produces the training labels
and is the fallback wherever the ensemble spread exceeds ``tau``.

Forces of accepted environments follow by the chain rule:
dE/dX_f = c_f + d E34_ML / d X_f, pushed through the feature gradient into
the same adjoint G[k] the plain path uses, then the same force kernel.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from .evaluator import EnvBatch, HIPEvaluator, _vp
from .mlp import INP, MLPEnsemble


class Surrogate:
    def __init__(self, ev: HIPEvaluator, nmemb: int = 4, seed: int = 0):
        if ev.prop != 0 or len(ev.pack.feat_rows) == 0:
            raise ValueError("the surrogate needs an Invariant basis with order-1 or order-2 functions")
        self.ev = ev
        self.nfeat = len(ev.pack.feat_rows)
        self.nmemb = nmemb
        self.mlp = MLPEnsemble(nmemb=nmemb, nin=self.nfeat, device=ev.device, seed=seed)
        dev = f"cuda:{ev.device}"
        self.dev = dev
        self.xmean = torch.zeros(self.nfeat, dtype=torch.float64, device=dev)
        self.xinvstd = torch.ones(self.nfeat, dtype=torch.float64, device=dev)
        self.ymean, self.ystd = 0.0, 1.0
        self.tau = float("inf")

    # -- device calls
    def features(self, b: EnvBatch, want_X: bool = False):
        """(X [nfeat, nenv] FP64 or None, Xn [nenv, 160] BF16, E2 [nenv] FP64)."""
        X = torch.empty((self.nfeat, b.nenv), dtype=torch.float64, device=self.dev) if want_X else None
        Xn = torch.empty((b.nenv, INP), dtype=torch.bfloat16, device=self.dev)
        E2 = torch.empty(b.nenv, dtype=torch.float64, device=self.dev)
        _lib.check(self.ev.lib.ace_surr_features(self.ev.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz),
                                                 _vp(b.species), _vp(self.xmean), _vp(self.xinvstd),
                                                 _vp(X) if X is not None else C.c_void_p(0), _vp(Xn), _vp(E2),
                                                 self.ev._stream(None)))
        return X, Xn, E2

    def energy_forces(self, b: EnvBatch, tau: float | None = None, forces: bool = True):
        """Surrogate E (and +dE/dr) with plain-ACE fallback; returns
        (E, dE or None, fallback_count tensor, fallback_list tensor)."""
        tau = self.tau if tau is None else tau
        _, Xn, E2 = self.features(b)
        if forces:
            y, dX = self.mlp.predict_grad(Xn)
        else:
            y, dX = self.mlp.predict(Xn), None
        E = torch.empty(b.nenv, dtype=torch.float64, device=self.dev)
        dE = torch.empty((3, b.nnbr), dtype=torch.float64, device=self.dev) if forces else None
        cnt = torch.zeros(1, dtype=torch.int32, device=self.dev)
        lst = torch.empty(max(1, b.nenv), dtype=torch.int32, device=self.dev)
        _lib.check(self.ev.lib.ace_surr_energy_forces(
            self.ev.ctx, b.nenv, _vp(b.offsets), _vp(b.xyz), _vp(b.species), self.nmemb, _vp(y),
            _vp(dX) if dX is not None else C.c_void_p(0), float(self.ymean), float(self.ystd),
            _vp(self.xinvstd), float(tau), _vp(E2), _vp(E), _vp(dE) if dE is not None else C.c_void_p(0),
            _vp(cnt), _vp(lst), self.ev._stream(None)))
        return E, dE, cnt, lst

    # -- training
    def fit(self, batches, epochs: int = 30, batch_rows: int = 4096, lr: float = 2e-3,
            weight_decay: float = 0.0, seed: int = 0, log=None):
        """Fit on a list of EnvBatch (labels from the plain FP64 path)."""
        t0 = time.time()
        Xs, E2s, Es = [], [], []
        for b in batches:
            X, _, E2 = self.features(b, want_X=True)
            E = self.ev.energy(b)
            Xs.append(X)
            E2s.append(E2)
            Es.append(E)
        X = torch.cat(Xs, dim=1)
        E34 = torch.cat(Es) - torch.cat(E2s)
        mean = X.mean(dim=1)
        std = X.std(dim=1)
        inv = torch.where(std > 1e-300, 1.0 / std, torch.zeros_like(std))
        self.xmean.copy_(mean)
        self.xinvstd.copy_(inv)
        self.ymean = float(E34.mean())
        self.ystd = float(E34.std())
        # normalized BF16 inputs (recomputed by the feature kernel) and targets
        Xn = torch.cat([self.features(b)[1] for b in batches], dim=0)
        t = ((E34 - self.ymean) / self.ystd).to(torch.float32)
        n = Xn.shape[0]
        g = torch.Generator(device=self.dev)
        g.manual_seed(seed)
        steps = 0
        hist = []
        nb = n // batch_rows
        for ep in range(epochs):
            perm = torch.randperm(n, device=self.dev, generator=g)
            lr_ep = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))      # cosine schedule
            for i in range(nb):
                idx = perm[i * batch_rows:(i + 1) * batch_rows]
                want = i == nb - 1
                loss = self.mlp.train_step(Xn.index_select(0, idx), t.index_select(0, idx), lr_ep,
                                           weight_decay, want_loss=want)
                steps += 1
            hist.append(float(np.mean(loss)))
            if log:
                log(f"epoch {ep} lr {lr_ep:.2e} loss {hist[-1]:.4e} ({time.time() - t0:.1f} s)")
        return {"steps": steps, "loss": hist, "seconds": time.time() - t0, "ymean": self.ymean,
                "ystd": self.ystd, "rows": n}

    def calibrate_tau(self, b: EnvBatch, mae_target: float, E_ref=None):
        """Smallest fallback fraction whose energy MAE against plain FP64 ACE
        stays below ``mae_target``: environments are accepted in order of
        increasing ensemble spread.  Returns (tau, fallback fraction, MAE)."""
        _, Xn, E2 = self.features(b)
        y = self.mlp.predict(Xn).double()
        E_ml = E2 + self.ymean + self.ystd * y.mean(0)
        spread = self.ystd * y.std(0, unbiased=False)
        if E_ref is None:
            E_ref = self.ev.energy(b)
        err = (E_ml - E_ref).abs()
        order = torch.argsort(spread)
        cum = torch.cumsum(err[order], 0) / b.nenv          # MAE if the first i are accepted
        ok = torch.nonzero(cum <= mae_target).flatten()
        nacc = int(ok[-1]) + 1 if len(ok) else 0
        tau = float(spread[order[nacc - 1]]) if nacc > 0 else -1.0
        mae = float(cum[nacc - 1]) if nacc > 0 else 0.0
        self.tau = tau
        return tau, 1.0 - nacc / b.nenv, mae
