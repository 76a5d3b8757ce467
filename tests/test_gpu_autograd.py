"""torch.autograd binding of the site energies (gpu): first derivatives in
positions and coefficients, and the mixed second derivative used by force
training, against the oracle and against central finite differences."""
import numpy as np
import pytest
import torch

import ace_oracle as O
from acehip import synth
from acehip.autograd import SiteEnergy, site_energies
from acehip.evaluator import HIPEvaluator
from acehip.symmbasis import LinearACEModel
from helpers import basis_for, random_batch, rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _case(prop, species, nenv, seed, nmin=5, nmax=20):
    basis, T = basis_for(3, 6, 4, prop=prop, species=species)
    c = synth.coefficients(len(basis), seed)
    ev = HIPEvaluator(basis, c, device=0)
    off, R, sp = random_batch(nenv, seed, nmin=nmin, nmax=nmax, nspecies=species)
    return basis, T, c, ev, off, R, sp


def _leaves(c, R):
    ct = torch.tensor(c, device=DEV, requires_grad=True)
    x = torch.tensor(R.T.copy(), device=DEV, requires_grad=True)
    return ct, x


@pytest.mark.parametrize("prop,species", [("Invariant", 0), ("Invariant", 2), ("EuclideanVector", 0)])
def test_first_order_matches_oracle(cuda, prop, species):
    print(f"[gpu] autograd first order {prop} species={species}", flush=True)
    basis, T, c, ev, off, R, sp = _case(prop, species, 20, 11)
    ct, x = _leaves(c, R)
    E = site_energies(ev, ct, off, x, sp)
    assert rel_err(E.detach().cpu().numpy(), O.energy(T, c, off, R, sp)) < 1e-12
    rng = np.random.default_rng(4)
    gE = rng.standard_normal(tuple(E.shape))
    gc, gx = torch.autograd.grad(E, [ct, x], torch.tensor(gE, device=DEV))
    env_of = np.repeat(np.arange(len(off) - 1), np.diff(off))
    g = O.grad_config(T, c, off, R, sp)
    B = O.eval_B(T, O.eval_AA(T, O.eval_A(T, off, R, sp)))
    if prop == "Invariant":
        ref_x = gE[env_of, None] * g
        ref_c = gE @ B
    else:
        ref_x = np.einsum("jp,jpd->jd", gE[env_of], g)
        ref_c = np.einsum("ep,ekp->k", gE, B)
    assert rel_err(gx.cpu().numpy().T, ref_x) < 1e-10
    assert rel_err(gc.cpu().numpy(), ref_c) < 1e-12


def test_gradcheck_central_differences(cuda):
    print("[gpu] autograd gradcheck", flush=True)
    basis, T, c, ev, off, R, sp = _case("Invariant", 0, 3, 5, nmin=4, nmax=7)
    ct, x = _leaves(c, R)
    assert torch.autograd.gradcheck(lambda cc, xx: site_energies(ev, cc, off, xx), (ct, x),
                                    eps=1e-6, atol=1e-7, rtol=1e-6)


@pytest.mark.parametrize("species", [0, 2])
def test_mixed_second_order_matches_oracle(cuda, species):
    print(f"[gpu] autograd mixed second order species={species}", flush=True)
    basis, T, c, ev, off, R, sp = _case("Invariant", species, 16, 21 + species)
    ct, x = _leaves(c, R)
    rng = np.random.default_rng(8)
    nenv = len(off) - 1
    gE = torch.tensor(rng.standard_normal(nenv), device=DEV, requires_grad=True)
    w = rng.standard_normal(R.shape)
    E = site_energies(ev, ct, off, x, sp)
    (gx,) = torch.autograd.grad(E, x, gE, create_graph=True)
    L = (gx * torch.tensor(w.T.copy(), device=DEV)).sum()
    dc, dgE = torch.autograd.grad(L, [ct, gE])
    adj = np.stack([O.adjoint_EVAL_D(T, c, off[e:e + 2] - off[e], R[off[e]:off[e + 1]], w[off[e]:off[e + 1]],
                                     None if sp is None else sp[off[e]:off[e + 1]])
                    for e in range(nenv)])
    gEn = gE.detach().cpu().numpy()
    assert rel_err(dc.cpu().numpy(), gEn @ adj) < 1e-10
    g = O.grad_config(T, c, off, R, sp)
    ref_gE = np.array([np.sum(w[off[e]:off[e + 1]] * g[off[e]:off[e + 1]]) for e in range(nenv)])
    assert rel_err(dgE.cpu().numpy(), ref_gE) < 1e-10
    # the reference's dp tangent: sum_k c_k adjoint_EVAL_D[k] (linearmodel.jl:223)
    assert rel_err(dgE.cpu().numpy(), adj @ c) < 1e-10


def test_force_loss_gradient_central_differences(cuda):
    """d/dc of a force loss, the quantity force training needs, against central
    differences of the loss along random coefficient directions."""
    print("[gpu] autograd force-loss gradient vs central differences", flush=True)
    basis, T, c, ev, off, R, sp = _case("Invariant", 0, 10, 31)
    rng = np.random.default_rng(12)
    Fref = torch.tensor(rng.standard_normal(R.T.shape) * 0.01, device=DEV)
    x = torch.tensor(R.T.copy(), device=DEV, requires_grad=True)

    def loss(cc, create_graph):
        E = site_energies(ev, cc, off, x)
        (g,) = torch.autograd.grad(E.sum(), x, create_graph=create_graph)
        return ((-g - Fref) ** 2).sum()

    ct = torch.tensor(c, device=DEV, requires_grad=True)
    (dc,) = torch.autograd.grad(loss(ct, True), ct)
    dc = dc.cpu().numpy()
    h = 1e-4
    for k in range(3):
        v = rng.standard_normal(len(c))
        lp = loss(torch.tensor(c + h * v, device=DEV), False).item()
        lm = loss(torch.tensor(c - h * v, device=DEV), False).item()
        fd = (lp - lm) / (2 * h)
        # the loss is quadratic in c, so central differences are exact up to rounding
        assert abs(fd - dc @ v) <= 1e-8 * max(1.0, abs(fd)), (k, fd, dc @ v)


def test_vector_mixed_second_order_not_provided(cuda):
    print("[gpu] autograd EuclideanVector second order raises", flush=True)
    basis, T, c, ev, off, R, sp = _case("EuclideanVector", 0, 4, 3)
    ct, x = _leaves(c, R)
    E = site_energies(ev, ct, off, x)
    (gx,) = torch.autograd.grad(E.sum(), x, create_graph=True)
    with pytest.raises(NotImplementedError):
        torch.autograd.grad(gx.sum(), ct)


def test_module_and_validation(cuda):
    print("[gpu] autograd SiteEnergy module", flush=True)
    basis, T, c, ev, off, R, sp = _case("Invariant", 0, 8, 41)
    m = SiteEnergy(LinearACEModel(basis, c))
    x = torch.tensor(R.T.copy(), device=DEV)
    E = m(off, x)
    assert rel_err(E.detach().cpu().numpy(), O.energy(T, c, off, R)) < 1e-12
    E.sum().backward()
    B = O.eval_B(T, O.eval_AA(T, O.eval_A(T, off, R)))
    assert rel_err(m.c.grad.cpu().numpy(), B.sum(0)) < 1e-12
    # new coefficients are uploaded before the next evaluation
    with torch.no_grad():
        m.c.mul_(2.0)
    assert rel_err(m(off, x).detach().cpu().numpy(), O.energy(T, 2.0 * c, off, R)) < 1e-12
    with pytest.raises(ValueError):
        site_energies(m.ev, m.c, off, x.T)
    with pytest.raises(ValueError):
        site_energies(m.ev, m.c[:-1], off, x)
