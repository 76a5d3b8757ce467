"""Host logic of the torch.autograd binding, on the CPU: a stand-in evaluator
returns the oracle's energies, gradients, basis values and adjoints, and the
binding's first and mixed second derivatives are checked against central
finite differences of those energies."""
import numpy as np
import pytest
import torch

import ace_oracle as O
from acehip import synth
from acehip.autograd import site_energies
from helpers import basis_for, random_batch, rel_err


class _OracleEvaluator:
    """The evaluator interface that the binding calls, answered by the oracle."""

    def __init__(self, T, c, nB):
        self.T, self.c = T, np.array(c, dtype=np.float64)
        self.prop = 0 if T["prop"] == "Invariant" else 1
        self.nB = nB

    def set_params(self, c):
        self.c = np.array(c, dtype=np.float64)
        self._autograd_c = None

    def _args(self, b):
        off = b.offsets.numpy()
        R = b.xyz.numpy().T.copy()
        sp = None if b.species is None else b.species.numpy()
        return off, R, sp

    def energy_forces(self, b):
        off, R, sp = self._args(b)
        E = O.energy(self.T, self.c, off, R, sp)
        g = O.grad_config(self.T, self.c, off, R, sp)
        dE = g.T.copy() if self.prop == 0 else g.transpose(1, 2, 0).reshape(9, -1).copy()
        return torch.tensor(E), torch.tensor(dE)

    def basis_B(self, b):
        off, R, sp = self._args(b)
        return torch.tensor(O.eval_B(self.T, O.eval_AA(self.T, O.eval_A(self.T, off, R, sp))))

    def adjoint_EVAL_D(self, b, w):
        off, R, sp = self._args(b)
        wn = w.detach().numpy().T
        return torch.tensor(np.stack([
            O.adjoint_EVAL_D(self.T, self.c, off[e:e + 2] - off[e], R[off[e]:off[e + 1]], wn[off[e]:off[e + 1]],
                             None if sp is None else sp[off[e]:off[e + 1]])
            for e in range(len(off) - 1)]))


def _case(prop, nenv, seed):
    basis, T = basis_for(3, 6, 4, prop=prop)
    c = synth.coefficients(len(basis), seed)
    off, R, _ = random_batch(nenv, seed, nmin=3, nmax=6)
    return _OracleEvaluator(T, c, len(basis)), c, off, R


@pytest.mark.parametrize("prop", ["Invariant", "EuclideanVector"])
def test_first_order_central_differences(prop):
    ev, c, off, R = _case(prop, 3, 2)
    ct = torch.tensor(c, requires_grad=True)
    x = torch.tensor(R.T.copy(), requires_grad=True)
    assert torch.autograd.gradcheck(lambda cc, xx: site_energies(ev, cc, off, xx), (ct, x),
                                    eps=1e-6, atol=1e-7, rtol=1e-6)


def test_mixed_second_order_central_differences():
    """d/dc and d/dgE of w . dE/dx (gE-weighted) against central differences
    of the first-order gradient."""
    ev, c, off, R = _case("Invariant", 4, 6)
    rng = np.random.default_rng(1)
    nenv = len(off) - 1
    x = torch.tensor(R.T.copy(), requires_grad=True)
    w = torch.tensor(rng.standard_normal(R.T.shape))

    def L(cc, gE, create_graph):
        E = site_energies(ev, cc, off, x)
        (g,) = torch.autograd.grad(E, x, gE, create_graph=create_graph)
        return (g * w).sum()

    ct = torch.tensor(c, requires_grad=True)
    gE = torch.tensor(rng.standard_normal(nenv), requires_grad=True)
    dc, dgE = torch.autograd.grad(L(ct, gE, True), [ct, gE])
    h = 1e-4
    for k in (0, 5, len(c) - 1):
        e = np.zeros(len(c))
        e[k] = h
        fd = (L(torch.tensor(c + e), gE.detach(), False) - L(torch.tensor(c - e), gE.detach(), False)) / (2 * h)
        assert abs(fd.item() - dc[k].item()) <= 1e-9 * max(1.0, abs(fd.item()))
    # L is linear in gE: one difference per environment is exact up to rounding
    base = gE.detach()
    for e_ in range(nenv):
        d = torch.zeros(nenv, dtype=torch.float64)
        d[e_] = 1.0
        fd = (L(ct.detach(), base + d, False) - L(ct.detach(), base - d, False)) / 2
        assert abs(fd.item() - dgE[e_].item()) <= 1e-10 * max(1.0, abs(fd.item()))


def test_argument_checks():
    ev, c, off, R = _case("Invariant", 2, 3)
    x = torch.tensor(R.T.copy())
    with pytest.raises(ValueError):
        site_energies(ev, torch.tensor(c), off, x.T.contiguous())
    with pytest.raises(ValueError):
        site_energies(ev, torch.tensor(c[:-1]), off, x)
    with pytest.raises(ValueError):
        site_energies(ev, torch.tensor(c, dtype=torch.float32), off, x)
    E = site_energies(ev, torch.tensor(c), off, x)
    assert rel_err(E.numpy(), O.energy(ev.T, c, off, R)) == 0.0
