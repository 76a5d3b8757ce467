"""The surrogate's BF16 MFMA MLP kernels vs the NumPy MLP oracle (gpu)."""
import numpy as np
import pytest
import torch

import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _bf16_tensor(X):
    return torch.as_tensor(M.bf16(X).astype(np.float32)).to(torch.bfloat16).cuda()


def test_mfma_lane_layout(cuda):
    print("[gpu] mfma 16x16x32 bf16 lane layout", flush=True)
    from acehip.mlp import layout_probe
    D = layout_probe()
    i = np.arange(16)[:, None]
    k = np.arange(32)[None, :]
    A = ((3 * i + k) % 7 - 3).astype(np.float64)
    kk = np.arange(32)[:, None]
    j = np.arange(16)[None, :]
    B = ((5 * kk + 2 * j) % 9 - 4).astype(np.float64)
    assert np.array_equal(D, (A @ B).astype(np.float32))


def test_forward_backward_vs_oracle(cuda):
    print("[gpu] mlp forward/backward vs oracle", flush=True)
    from acehip.mlp import MLPEnsemble
    rng = np.random.default_rng(0)
    ens = MLPEnsemble(nmemb=2, nin=152, seed=3)
    P = ens.get_params()
    rows = 200   # not a multiple of the 64-row tile
    X = np.zeros((rows, 160))
    X[:, :152] = rng.normal(size=(rows, 152))
    Xt = _bf16_tensor(X)
    t = rng.normal(size=rows)
    tt = torch.as_tensor(t, dtype=torch.float32).cuda()
    y = ens.predict(Xt).cpu().numpy()
    g, loss = ens.loss_grad(Xt, tt)
    y2, dX = ens.predict_grad(Xt)
    for m in range(2):
        yr, cache = M.forward(P[m], X)
        assert np.max(np.abs(y[m] - yr)) < 2e-3 * max(1.0, np.max(np.abs(yr)))
        assert np.max(np.abs(y2[m].cpu().numpy() - yr)) < 2e-3 * max(1.0, np.max(np.abs(yr)))
        dy = 2.0 * (yr - t) / rows
        gr, _ = M.backward(P[m], cache, dy)
        assert loss[m] == pytest.approx(np.mean((yr - t) ** 2), rel=1e-2)
        rel = np.linalg.norm(g[m] - gr) / np.linalg.norm(gr)
        assert rel < 2e-2, rel
        _, dXr = M.backward(P[m], cache, np.ones(rows))
        relx = np.linalg.norm(dX[m].cpu().numpy()[:, :152] - dXr[:, :152]) / np.linalg.norm(dXr[:, :152])
        assert relx < 2e-2, relx


def test_training_reduces_loss(cuda):
    print("[gpu] mlp training on a smooth target", flush=True)
    from acehip.mlp import MLPEnsemble
    rng = np.random.default_rng(1)
    rows = 4096
    X = np.zeros((rows, 160))
    X[:, :8] = rng.normal(size=(rows, 8))
    t = np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]
    Xt = _bf16_tensor(X)
    tt = torch.as_tensor(t, dtype=torch.float32).cuda()
    ens = MLPEnsemble(nmemb=4, nin=8, seed=2)
    first = ens.train_step(Xt, tt, 1e-3, want_loss=True)
    for _ in range(300):
        ens.train_step(Xt, tt, 1e-3)
    last = ens.train_step(Xt, tt, 1e-3, want_loss=True)
    assert np.all(last < 0.1 * first), (first, last)
