"""Learned surrogate (gpu): features, chain-rule forces, fallback, training."""
import numpy as np
import pytest
import torch

import ace_oracle as O
from acehip import synth
from acehip.evaluator import EnvBatch, HIPEvaluator
from helpers import basis_for, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(cuda):
    basis, T = basis_for(4, 12, 6)
    c = synth.coefficients(len(basis), 2)
    ev = HIPEvaluator(basis, c, device=0)
    return basis, T, c, ev


def _batch(nenv, seed):
    off, R, _ = synth.random_envs(nenv, "poisson30", seed)
    return off, R, EnvBatch(off, R.T.copy(), None, device="cuda:0")


def test_features_and_E2(setup):
    print("[gpu] surrogate features and exact order<=2 energy share", flush=True)
    from acehip.surrogate import Surrogate
    basis, T, c, ev = setup
    s = Surrogate(ev, nmemb=2, seed=0)
    off, R, b = _batch(12, 100)
    X, Xn, E2 = s.features(b, want_X=True)
    A = O.eval_A(T, off, R)
    B = O.eval_B(T, O.eval_AA(T, A))
    rows = ev.pack.feat_rows
    assert rel_err(X.cpu().numpy().T, B[:, rows]) < 1e-12
    c2 = np.zeros_like(c)
    keep = np.concatenate([rows, ev.pack.const_rows])   # E2 carries the constant term
    c2[keep] = c[keep]
    assert rel_err(E2.cpu().numpy(), O.energy(T, c2, off, R)) < 1e-12
    # BF16 normalized copy, zero padding
    assert torch.all(Xn[:, len(rows):] == 0)


def test_chain_rule_forces(setup):
    print("[gpu] surrogate forces through the feature gradient", flush=True)
    from acehip.surrogate import Surrogate
    basis, T, c, ev = setup
    s = Surrogate(ev, nmemb=2, seed=1)
    off, R, b = _batch(6, 101)
    X, _, _ = s.features(b, want_X=True)
    s.xmean.copy_(X.mean(1))
    s.xinvstd.copy_(1.0 / (X.std(1) + 1e-3))
    s.ymean, s.ystd = 0.3, 0.7
    E, dE, cnt, _ = s.energy_forces(b, tau=float("inf"))
    assert int(cnt.item()) == 0
    _, Xn, _ = s.features(b)
    y, dX = s.mlp.predict_grad(Xn)
    rows = ev.pack.feat_rows
    g = dE.cpu().numpy().T
    dXm = dX.double().mean(0).cpu().numpy()[:, : len(rows)]
    inv = s.xinvstd.cpu().numpy()
    for e in range(6):
        w = c[rows] + s.ystd * inv * dXm[e]
        ce = np.zeros_like(c)
        ce[rows] = w
        ref = O.grad_config(T, ce, off[e:e + 2] - off[e], R[off[e]:off[e + 1]])
        assert rel_err(g[off[e]:off[e + 1]], ref) < 1e-10
    keep = np.concatenate([rows, ev.pack.const_rows])
    E2 = O.energy(T, np.where(np.isin(np.arange(len(c)), keep), c, 0.0), off, R)
    Eml = E2 + s.ymean + s.ystd * y.double().mean(0).cpu().numpy()
    assert rel_err(E.cpu().numpy(), Eml) < 1e-12


def test_full_fallback_equals_plain(setup):
    print("[gpu] surrogate with every environment falling back", flush=True)
    from acehip.surrogate import Surrogate
    basis, T, c, ev = setup
    s = Surrogate(ev, nmemb=2, seed=2)
    off, R, b = _batch(130, 102)
    E, dE, cnt, lst = s.energy_forces(b, tau=-1.0)
    assert int(cnt.item()) == b.nenv
    assert sorted(lst[: b.nenv].cpu().tolist()) == list(range(b.nenv))
    Ep, dEp = ev.energy_forces(b)
    assert rel_err(E.cpu().numpy(), Ep.cpu().numpy()) < 1e-14
    assert rel_err(dE.cpu().numpy(), dEp.cpu().numpy()) < 1e-14


def test_training_learns_and_fallback_meets_target(setup):
    print("[gpu] surrogate training on 40k environments", flush=True)
    from acehip.surrogate import Surrogate
    basis, T, c, ev = setup
    s = Surrogate(ev, nmemb=4, seed=3)
    train = [_batch(20000, 200 + i)[2] for i in range(2)]
    info = s.fit(train, epochs=8, batch_rows=2048, lr=2e-3)
    assert info["loss"][-1] < 0.5 * info["loss"][0]
    _, _, test = _batch(5000, 300)
    Eref = ev.energy(test)
    _, Xn, E2 = s.features(test)
    y = s.mlp.predict(Xn).double().mean(0)
    mae_ml = float((E2 + s.ymean + s.ystd * y - Eref).abs().mean())
    mae_mean = float((E2 + s.ymean - Eref).abs().mean())
    assert mae_ml < 0.7 * mae_mean, (mae_ml, mae_mean)
    tau, frac, mae = s.calibrate_tau(test, 1e-3, Eref)
    E, _, cnt, _ = s.energy_forces(test, forces=False)
    got = float((E - Eref).abs().mean())
    assert got <= 1e-3 * 1.0001
    assert abs(int(cnt.item()) / test.nenv - frac) < 0.02
