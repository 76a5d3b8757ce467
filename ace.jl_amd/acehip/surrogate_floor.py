"""How far can an MLP on the order-1/2 invariants go?  The accuracy floor of
the learned surrogate, measured with plain PyTorch FP32 networks.

The surrogate (``surrogate.py``) predicts the correlation-order-3/4 share
``E34 = E - E2`` of the config-2 site energy from the 151 order-1/2
invariants ``X`` and falls back to plain FP64 ACE where its ensemble spread
is large.  This study separates what the inputs allow from what the BF16
MFMA kernels deliver: ensembles of FP32 networks of several sizes are
trained with Adam on labels from the plain GPU path, and for each the
held-out environments are accepted in order of increasing ensemble spread.
Reported per model:

* ``rmse``: RMS error of the ensemble mean on all held-out environments;
* ``accept_at_overall_mae``: the largest accepted fraction whose MAE, taken
  over *all* environments with the fallback ones exact, stays <= the target
  (what the production calibration does);
* ``accept_at_accepted_mae``: the largest accepted fraction whose MAE over
  the *accepted* environments alone stays <= the target;
* the MAE-versus-acceptance curve at 10 % steps.

A ``bf16in`` variant rounds the normalised inputs to BF16 (8-bit mantissa)
before the FP32 network, as the production kernels do.

Run on a GPU (data from libacehip.so, training in torch FP32):
``python -m acehip.surrogate_floor --train 1000000 --test 200000``; prints
one JSON line.  No reference counterpart (ACE.jl is a linear model).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import synth
from .evaluator import EnvBatch, HIPEvaluator
from .surrogate import Surrogate


def make_data(ev: HIPEvaluator, n: int, seed: int, chunk: int = 100_000):
    """FP64 features X [n, nfeat] and labels E34 = E - E2 [n] (torch, device)."""
    s = Surrogate(ev, nmemb=1, seed=0)
    Xs, ys = [], []
    for i in range(0, n, chunk):
        off, R, _ = synth.random_envs(min(chunk, n - i), "poisson30", seed * 1000 + i // chunk)
        b = EnvBatch(off, R.T.copy(), None, device=f"cuda:{ev.device}")
        X, _, E2 = s.features(b, want_X=True)
        E = ev.energy(b)
        Xs.append(X.T.clone())
        ys.append(E - E2)
    return torch.cat(Xs), torch.cat(ys)


def mlp(nin: int, width: int, depth: int):
    layers, d = [], nin
    for _ in range(depth):
        layers += [torch.nn.Linear(d, width), torch.nn.SiLU()]
        d = width
    layers.append(torch.nn.Linear(d, 1))
    return torch.nn.Sequential(*layers)


def train_member(Xn, t, width, depth, epochs, batch, lr, seed, log=None):
    torch.manual_seed(seed)
    net = mlp(Xn.shape[1], width, depth).to(Xn.device)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    n = Xn.shape[0]
    nb = n // batch
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=epochs * nb)
    g = torch.Generator(device=Xn.device)
    g.manual_seed(seed)
    for ep in range(epochs):
        perm = torch.randperm(n, device=Xn.device, generator=g)
        tot = 0.0
        for i in range(nb):
            idx = perm[i * batch:(i + 1) * batch]
            loss = torch.mean((net(Xn[idx]).squeeze(1) - t[idx]) ** 2)
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            sched.step()
            tot += float(loss.detach()) if i == nb - 1 else 0.0
        if log and (ep % 10 == 9 or ep == epochs - 1):
            log(f"  w{width} d{depth} seed {seed} epoch {ep + 1}: last-batch loss {tot:.3e}")
    return net


def acceptance(err: np.ndarray, spread: np.ndarray, target: float):
    """Accept in order of increasing spread; returns the curves' summary."""
    n = len(err)
    order = np.argsort(spread, kind="stable")
    e = err[order]
    cum = np.cumsum(e)
    overall = cum / n                               # fallback environments exact
    accepted = cum / np.arange(1, n + 1)            # MAE over the accepted ones
    ok_o = np.nonzero(overall <= target)[0]
    ok_a = np.nonzero(accepted <= target)[0]
    curve = {}
    for q in np.arange(0.1, 1.0001, 0.1):
        k = max(1, int(round(q * n)))
        curve[f"{q:.1f}"] = {"mae_overall": float(overall[k - 1]), "mae_accepted": float(accepted[k - 1])}
    return {"accept_at_overall_mae": (int(ok_o[-1]) + 1) / n if len(ok_o) else 0.0,
            "accept_at_accepted_mae": (int(ok_a[-1]) + 1) / n if len(ok_a) else 0.0,
            "curve": curve}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=1_000_000)
    ap.add_argument("--test", type=int, default=200_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--target", type=float, default=1e-3)
    ap.add_argument("--models", default="128x3,128x3bf16in,256x3,512x4")
    args = ap.parse_args(argv)
    log = lambda m: print("[floor]", m, flush=True)  # noqa: E731
    w = synth.WORKLOADS[2]
    basis = synth.make_basis(w)
    c = synth.coefficients(len(basis), w["seed"])
    ev = HIPEvaluator(basis, c, device=0)
    t0 = time.time()
    Xtr, ytr = make_data(ev, args.train, 5)
    Xte, yte = make_data(ev, args.test, 6)
    log(f"data: {args.train} + {args.test} environments, {Xtr.shape[1]} features ({time.time() - t0:.1f} s)")
    mean, std = Xtr.mean(0), Xtr.std(0)
    inv = torch.where(std > 1e-300, 1.0 / std, torch.zeros_like(std))
    ym, ys = float(ytr.mean()), float(ytr.std())
    Xn_tr = ((Xtr - mean) * inv).float()
    Xn_te = ((Xte - mean) * inv).float()
    t_tr = ((ytr - ym) / ys).float()
    out = {"workload": "config 2 basis (order 4, degree 12, Lmax 6), Poisson(30) environments",
           "features": int(Xtr.shape[1]), "train_envs": args.train, "test_envs": args.test,
           "label_std": ys, "target_mae": args.target, "models": {}}
    for spec in args.models.split(","):
        bf = spec.endswith("bf16in")
        wd = spec.replace("bf16in", "")
        width, depth = (int(v) for v in wd.split("x"))
        xtr = Xn_tr.bfloat16().float() if bf else Xn_tr
        xte = Xn_te.bfloat16().float() if bf else Xn_te
        t1 = time.time()
        preds = []
        for m in range(args.members):
            net = train_member(xtr, t_tr, width, depth, args.epochs, 4096, 2e-3, 100 + m, log)
            with torch.no_grad():
                preds.append(torch.cat([net(xte[i:i + 65536]).squeeze(1) for i in range(0, len(xte), 65536)]))
            del net
        P = torch.stack(preds).double()
        E_ml = ym + ys * P.mean(0)
        spread = ys * P.std(0, unbiased=False)
        err = (E_ml - yte).abs().cpu().numpy()
        res = acceptance(err, spread.cpu().numpy(), args.target)
        res.update({"rmse": float(torch.sqrt(torch.mean((E_ml - yte) ** 2))), "mae_all": float(err.mean()),
                    "train_seconds": round(time.time() - t1, 1), "epochs": args.epochs, "members": args.members,
                    "inputs": "BF16-rounded" if bf else "FP32"})
        out["models"][spec] = res
        log(f"{spec}: rmse {res['rmse']:.4e} accept(overall MAE<=target) {res['accept_at_overall_mae']:.3f} "
            f"accept(accepted MAE<=target) {res['accept_at_accepted_mae']:.3f}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
