"""The NumPy MLP oracle pinned by finite differences (BF16 rounding off)."""
import numpy as np

import mlp_oracle as M


def test_mlp_oracle_gradients_fd(monkeypatch):
    monkeypatch.setattr(M, "bf16", lambda x: np.asarray(x, dtype=np.float64))
    rng = np.random.default_rng(0)
    p = rng.normal(size=M.NPARAM) * 0.1
    X = np.zeros((5, 160))
    X[:, :20] = rng.normal(size=(5, 20))
    t = rng.normal(size=5)

    def loss(pp, XX):
        y, _ = M.forward(pp, XX)
        return np.mean((y - t) ** 2)

    y, cache = M.forward(p, X)
    g, dX = M.backward(p, cache, 2 * (y - t) / 5)
    h = 1e-6
    for idx in list(rng.choice(M.OFF["W1"][1], 5)) + [M.OFF["b2"][0] + 3, M.OFF["W3"][0] + 77,
                                                     M.OFF["w4"][0] + 9, M.OFF["b4"][0]]:
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        assert abs((loss(pp, X) - loss(pm, X)) / (2 * h) - g[idx]) < 1e-6
    for (r, c) in [(0, 3), (4, 17)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[r, c] += h
        Xm[r, c] -= h
        assert abs((loss(p, Xp) - loss(p, Xm)) / (2 * h) - dX[r, c]) < 1e-6


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 65504.0])
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12
