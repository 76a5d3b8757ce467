"""CPU oracle for the ACE basis-evaluation path — TEST INFRASTRUCTURE ONLY.

This module restates, in NumPy FP64, the evaluation algorithm of ACE.jl
(v0.12.46) for the Rn x Ylm Product1pBasis -> PIBasis -> SymmetricBasis ->
LinearACEModel path.  It is the checker for the HIP implementation and must
never be imported by the product package (``ace.jl_amd``); only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it.

Inputs are *shared basis tables* (radial recursion coefficients, 1p spec,
AA spec, A2B map, coefficients) so the oracle and the GPU consume identical
basis definitions; the oracle recomputes everything downstream of them.

Restated reference functions (file:line in ACE.jl ``src/``):
  * ``orthpolys.jl:250-263``  OrthPolyBasis evaluate!          -> ``rn_eval``
  * ``orthpolys.jl:287-306``  evaluate_ed!                      -> ``rn_eval_ed``
  * ``distancetransforms.jl:73``, ``Rn.jl:21``, ``chain.jl:40-52``  (chain rule)
  * ``sphericalharmonics.jl:43-51``   cart2spher                 -> ``cart2spher``
  * ``sphericalharmonics.jl:167-197`` ALP evaluate!             -> ``alp_eval``
  * ``sphericalharmonics.jl:212-267`` ALP _evaluate_ed!         -> ``alp_eval_ed``
  * ``sphericalharmonics.jl:379-403`` cYlm!                     -> ``ylm_eval``
  * ``sphericalharmonics.jl:410-443,60-65`` cYlm_ed!, dspher_to_dcart -> ``ylm_eval_ed``
  * ``product_1pbasis.jl:99-134``  add_into_A!, evaluate(UConfig) -> ``eval_A``
  * ``product_1pbasis.jl:169-244`` _add_into_A_dA!, evaluate_ed   -> ``eval_A_dA``
  * ``pibasis.jl:265-275``  PIBasis evaluate!                     -> ``eval_AA``
  * ``pibasis.jl:338-390``  _AA_local_adjoints!                   -> ``_local_adjoints``
  * ``pibasis.jl:402-432``  _evaluate_ed!                         -> ``eval_AA_dAA``
  * ``symmbasis.jl:248-292,312-334`` genmul!, evaluate, evaluate_ed -> ``eval_B``, ``eval_B_dB``
  * ``evaluator.jl:59-67`` + ``symmbasis.jl:267-285`` c~ = A2B^T c   -> ``ctilde``
  * ``evaluator.jl:121-147`` ProductEvaluator energy               -> ``energy``
  * ``evaluator.jl:161-200`` _rrule_evaluate (grad_config)          -> ``grad_config``
  * ``evaluator.jl:204-244`` adjoint_EVAL_D                         -> ``adjoint_EVAL_D``
  * ``properties.jl:44-59`` coco_o_daa (Invariant / EuclideanVector gradients)

Pinned by: closed-form Ylm (test/polynomials/test_ylm.jl:14-48) and SciPy
``sph_harm_y``; orthonormality of the radial basis on its measure; central
finite differences of every derivative; brute-force A = sum_j phi(X_j) and
AA = prod A on small inputs; B invariance (O(3) + permutation) and
equivariance (EuclideanVector); ``E = c . B`` (naive vs product evaluator,
test/test_linearmodel.jl:50-81).  See tests/test_oracle.py.

Environments are given as ragged batches: ``offsets[nenv+1]`` into neighbour
arrays ``R[nnbr, 3]`` (and ``species[nnbr]`` for a categorical 1p basis).
"""
from __future__ import annotations

import numpy as np

EPS = np.finfo(np.float64).eps


# ---------------------------------------------------------------- radial


def _fcut(pl, tl, pr, tr, t):
    out = (t - tl) ** pl * (t - tr) ** pr
    mask = np.zeros(np.shape(t), dtype=bool)
    if pl > 0:
        mask |= t < tl
    if pr > 0:
        mask |= t > tr
    return np.where(mask, 0.0, out)


def _fcut_d(pl, tl, pr, tr, t):
    # d/dt (t-tl)^pl (t-tr)^pr  (orthpolys.jl:34-35,48-53)
    d = np.zeros_like(t)
    if pl > 0:
        d = d + pl * (t - tl) ** (pl - 1) * (t - tr) ** pr
    if pr > 0:
        d = d + (t - tl) ** pl * pr * (t - tr) ** (pr - 1)
    mask = np.zeros(np.shape(t), dtype=bool)
    if pl > 0:
        mask |= t < tl
    if pr > 0:
        mask |= t > tr
    return np.where(mask, 0.0, d)


def transform(rn, r):
    return ((1.0 + rn["r0"]) / (1.0 + r)) ** rn["p"]


def transform_d(rn, r):
    p, r0 = rn["p"], rn["r0"]
    u = (1.0 + r0) / (1.0 + r)
    du = -(1.0 + r0) / (1.0 + r) ** 2
    return p * u ** (p - 1) * du


def rn_eval(rn, t):
    """orthpolys.jl:253-263; t: array (...) -> P (..., N)."""
    A, B, C = rn["A"], rn["B"], rn["C"]
    N = len(A)
    t = np.asarray(t, dtype=np.float64)
    P = np.zeros(t.shape + (N,))
    P[..., 0] = A[0] * _fcut(rn["pl"], rn["tl"], rn["pr"], rn["tr"], t)
    if N > 1:
        P[..., 1] = (A[1] * t + B[1]) * P[..., 0]
    for n in range(2, N):
        P[..., n] = (A[n] * t + B[n]) * P[..., n - 1] + C[n] * P[..., n - 2]
    return P


def rn_eval_ed(rn, t):
    """orthpolys.jl:287-306: (P, dP/dt)."""
    A, B, C = rn["A"], rn["B"], rn["C"]
    N = len(A)
    t = np.asarray(t, dtype=np.float64)
    P = np.zeros(t.shape + (N,))
    dP = np.zeros(t.shape + (N,))
    P[..., 0] = A[0] * _fcut(rn["pl"], rn["tl"], rn["pr"], rn["tr"], t)
    dP[..., 0] = A[0] * _fcut_d(rn["pl"], rn["tl"], rn["pr"], rn["tr"], t)
    if N > 1:
        al = A[1] * t + B[1]
        P[..., 1] = al * P[..., 0]
        dP[..., 1] = al * dP[..., 0] + A[1] * P[..., 0]
    for n in range(2, N):
        al = A[n] * t + B[n]
        P[..., n] = al * P[..., n - 1] + C[n] * P[..., n - 2]
        dP[..., n] = al * dP[..., n - 1] + C[n] * dP[..., n - 2] + A[n] * P[..., n - 1]
    return P, dP


def radial(rn, R):
    """R_n(|r|) for positions R (..., 3)."""
    r = np.sqrt(np.sum(R * R, axis=-1))
    return rn_eval(rn, transform(rn, r))


def radial_ed(rn, R):
    """(R_n, grad R_n): grad = dP/dt * dt/dr * rr/r (chain.jl:40-52)."""
    r = np.sqrt(np.sum(R * R, axis=-1))
    P, dP = rn_eval_ed(rn, transform(rn, r))
    g = transform_d(rn, r)[..., None] * (R / r[..., None])      # (..., 3)
    return P, dP[..., :, None] * g[..., None, :]                 # (..., N, 3)


# ---------------------------------------------------------------- Ylm


def index_p(l, m):
    return m + l * (l + 1) // 2


def index_y(l, m):
    return m + l + l * l


def sizeP(L):
    return (L + 1) * (L + 2) // 2


def alp_coeffs(L):
    """sphericalharmonics.jl:146-159."""
    A = np.zeros(sizeP(L))
    B = np.zeros(sizeP(L))
    for l in range(2, L + 1):
        ls, lm1s = l * l, (l - 1) * (l - 1)
        for m in range(0, l - 1):
            ms = m * m
            A[index_p(l, m)] = np.sqrt((4 * ls - 1.0) / (ls - ms))
            B[index_p(l, m)] = -np.sqrt((lm1s - ms) / (4 * lm1s - 1.0))
    return A, B


def cart2spher(R):
    """sphericalharmonics.jl:43-51."""
    x, y, z = R[..., 0], R[..., 1], R[..., 2]
    r = np.sqrt(x * x + y * y + z * z)
    phi = np.arctan2(y, x)
    return dict(r=r, cosphi=np.cos(phi), sinphi=np.sin(phi), costh=z / r,
                sinth=np.sqrt(x ** 2 + y ** 2) / r)


def alp_eval(L, S):
    A, B = alp_coeffs(L)
    P = np.zeros(S["r"].shape + (sizeP(L),))
    temp = np.full(S["r"].shape, np.sqrt(0.5 / np.pi))
    P[..., 0] = temp
    if L == 0:
        return P
    P[..., index_p(1, 0)] = S["costh"] * np.sqrt(3) * temp
    temp = -np.sqrt(1.5) * S["sinth"] * temp
    P[..., index_p(1, 1)] = temp
    for l in range(2, L + 1):
        il = l * (l + 1) // 2
        ilm1 = il - l
        ilm2 = ilm1 - l + 1
        for m in range(0, l - 1):
            P[..., il + m] = A[il + m] * (S["costh"] * P[..., ilm1 + m] + B[il + m] * P[..., ilm2 + m])
        P[..., il + l - 1] = S["costh"] * np.sqrt(2 * (l - 1) + 3) * temp
        temp = -np.sqrt(1.0 + 0.5 / l) * S["sinth"] * temp
        P[..., il + l] = temp
    return P


def alp_eval_ed(L, S):
    """sphericalharmonics.jl:212-267 (P/sin for m >= 1; dP = dP/dtheta)."""
    A, B = alp_coeffs(L)
    ct, st = S["costh"], S["sinth"]
    shp = S["r"].shape + (sizeP(L),)
    P = np.zeros(shp)
    dP = np.zeros(shp)
    temp = np.full(S["r"].shape, np.sqrt(0.5 / np.pi))
    P[..., 0] = temp
    temp_d = np.zeros(S["r"].shape)
    if L == 0:
        return P, dP
    P[..., index_p(1, 0)] = ct * np.sqrt(3) * temp
    dP[..., index_p(1, 0)] = -st * np.sqrt(3) * temp + ct * np.sqrt(3) * temp_d
    temp1, temp_d = (-np.sqrt(1.5) * temp, -np.sqrt(1.5) * (ct * temp + st * temp_d))
    P[..., index_p(1, 1)] = temp1
    dP[..., index_p(1, 1)] = temp_d
    for l in range(2, L + 1):
        i, i1, i2 = index_p(l, 0), index_p(l - 1, 0), index_p(l - 2, 0)
        P[..., i] = A[i] * (ct * P[..., i1] + B[i] * P[..., i2])
        dP[..., i] = A[i] * (-st * P[..., i1] + ct * dP[..., i1] + B[i] * dP[..., i2])
        for m in range(1, l - 1):
            i, i1, i2 = index_p(l, m), index_p(l - 1, m), index_p(l - 2, m)
            P[..., i] = A[i] * (ct * P[..., i1] + B[i] * P[..., i2])
            dP[..., i] = A[i] * (-st ** 2 * P[..., i1] + ct * dP[..., i1] + B[i] * dP[..., i2])
        c = np.sqrt(2 * (l - 1) + 3)
        P[..., index_p(l, l - 1)] = c * ct * temp1
        dP[..., index_p(l, l - 1)] = c * (-st ** 2 * temp1 + ct * temp_d)
        f = np.sqrt(1.0 + 0.5 / l)
        temp1, temp_d = (-f * st * temp1, -f * (ct * temp1 * st + st * temp_d))
        P[..., index_p(l, l)] = temp1
        dP[..., index_p(l, l)] = temp_d
    return P, dP


def ylm_eval(L, R):
    """cYlm! (sphericalharmonics.jl:379-403) for positions R (..., 3)."""
    S = cart2spher(R)
    P = alp_eval(L, S)
    Y = np.zeros(S["r"].shape + ((L + 1) ** 2,), dtype=np.complex128)
    ep = np.full(S["r"].shape, 1.0 / np.sqrt(2) + 0j)
    for l in range(L + 1):
        Y[..., index_y(l, 0)] = P[..., index_p(l, 0)] * ep
    sig = 1
    ep_fact = S["cosphi"] + 1j * S["sinphi"]
    for m in range(1, L + 1):
        sig *= -1
        ep = ep * ep_fact
        em = sig * np.conj(ep)
        for l in range(m, L + 1):
            p = P[..., index_p(l, m)]
            Y[..., index_y(l, -m)] = em * p
            Y[..., index_y(l, m)] = ep * p
    return Y


def _dspher_to_dcart(S, f_phi_div_sinth, f_th):
    r = S["r"] + EPS
    return np.stack([-(S["sinphi"] * f_phi_div_sinth) + (S["cosphi"] * S["costh"] * f_th),
                     (S["cosphi"] * f_phi_div_sinth) + (S["sinphi"] * S["costh"] * f_th),
                     -(S["sinth"] * f_th)], axis=-1) / r[..., None]


def ylm_eval_ed(L, R):
    """cYlm_ed! (sphericalharmonics.jl:410-443): (Y, dY (...,nY,3))."""
    S = cart2spher(R)
    P, dP = alp_eval_ed(L, S)
    nY = (L + 1) ** 2
    Y = np.zeros(S["r"].shape + (nY,), dtype=np.complex128)
    dY = np.zeros(S["r"].shape + (nY, 3), dtype=np.complex128)
    ep = 1.0 / np.sqrt(2)
    zero = np.zeros(S["r"].shape)
    for l in range(L + 1):
        Y[..., index_y(l, 0)] = P[..., index_p(l, 0)] * ep
        dY[..., index_y(l, 0), :] = _dspher_to_dcart(S, zero, dP[..., index_p(l, 0)] * ep)
    sig = 1
    ep_fact = S["cosphi"] + 1j * S["sinphi"]
    ep = np.full(S["r"].shape, ep + 0j)
    for m in range(1, L + 1):
        sig *= -1
        ep = ep * ep_fact
        em = sig * np.conj(ep)
        dep = 1j * m * ep
        dem = 1j * (-m) * em
        for l in range(m, L + 1):
            pds = P[..., index_p(l, m)]
            Y[..., index_y(l, -m)] = em * pds * S["sinth"]
            Y[..., index_y(l, m)] = ep * pds * S["sinth"]
            dpt = dP[..., index_p(l, m)]
            dY[..., index_y(l, -m), :] = _dspher_to_dcart(S, dem * pds, em * dpt)
            dY[..., index_y(l, m), :] = _dspher_to_dcart(S, dep * pds, ep * dpt)
    return Y, dY


# ---------------------------------------------------------------- 1p basis


def _phi_parts(T, R, species):
    """Per neighbour: R_n values, Ylm values and the species mask for every
    1p function.  aspec rows: (mu, n, l, m) with mu = -1 for no categorical."""
    aspec = np.asarray(T["aspec"])
    Rn = radial(T["rn"], R)                       # (nnbr, N)
    Y = ylm_eval(T["L"], R)                       # (nnbr, nY)
    iR = aspec[:, 1] - 1
    iY = aspec[:, 2] + aspec[:, 2] ** 2 + aspec[:, 3]
    return Rn[:, iR], Y[:, iY], _mask(aspec, species, len(R))


def _mask(aspec, species, nnbr):
    mu = aspec[:, 0]
    if np.all(mu < 0):
        return None
    return (np.asarray(species)[:, None] == mu[None, :])


def eval_A(T, offsets, R, species=None):
    """Product1pBasis evaluate(cfg) for each environment: A[env, iA]
    accumulated neighbour by neighbour (product_1pbasis.jl:99-134)."""
    R = np.asarray(R, dtype=np.float64)
    nenv = len(offsets) - 1
    nA = len(T["aspec"])
    Rv, Yv, mask = _phi_parts(T, R, species)
    phi = Rv * Yv
    if mask is not None:
        phi = np.where(mask, phi, 0.0)
    A = np.zeros((nenv, nA), dtype=np.complex128)
    for e in range(nenv):
        for j in range(offsets[e], offsets[e + 1]):
            A[e] += phi[j]
    return A


def eval_A_dA(T, offsets, R, species=None):
    """_add_into_A_dA! per neighbour (product_1pbasis.jl:169-244):
    A[env, iA] and dA[nbr, iA, 3] (dA of neighbour j w.r.t. r_j)."""
    R = np.asarray(R, dtype=np.float64)
    aspec = np.asarray(T["aspec"])
    nenv = len(offsets) - 1
    Pn, dPn = radial_ed(T["rn"], R)
    Y, dY = ylm_eval_ed(T["L"], R)
    iR = aspec[:, 1] - 1
    iY = aspec[:, 2] + aspec[:, 2] ** 2 + aspec[:, 3]
    Rv, dRv = Pn[:, iR], dPn[:, iR, :]
    Yv, dYv = Y[:, iY], dY[:, iY, :]
    phi = Rv * Yv
    dphi = dRv * Yv[..., None] + dYv * Rv[..., None]
    mask = _mask(aspec, species, len(R))
    if mask is not None:
        phi = np.where(mask, phi, 0.0)
        dphi = np.where(mask[..., None], dphi, 0.0)
    A = np.zeros((nenv, len(aspec)), dtype=np.complex128)
    for e in range(nenv):
        for j in range(offsets[e], offsets[e + 1]):
            A[e] += phi[j]
    return A, dphi


# ---------------------------------------------------------------- AA, B


def eval_AA(T, A):
    """PIBasis evaluate! (pibasis.jl:265-275): left-to-right products;
    real part for isrealAA."""
    orders, iAA2iA = np.asarray(T["orders"]), np.asarray(T["iAA2iA"])
    AA = np.ones((A.shape[0], len(orders)), dtype=np.complex128)
    for t in range(iAA2iA.shape[1]):
        act = orders > t
        idx = np.where(act, iAA2iA[:, t], 0)
        AA = np.where(act[None, :], AA * A[:, idx], AA)
    return AA.real.copy() if T["prop"] == "Invariant" else AA


def _local_adjoints(Af):
    """_AA_local_adjoints! (pibasis.jl:338-390) for factor values
    Af[..., ord]; returns (aa, dAAdA[..., ord])."""
    ord_ = Af.shape[-1]
    d = np.empty_like(Af)
    if ord_ == 1:
        d[..., 0] = 1.0
        return Af[..., 0], d
    if ord_ == 2:
        d[..., 0] = Af[..., 1]
        d[..., 1] = Af[..., 0]
        return Af[..., 0] * Af[..., 1], d
    A1, A2 = Af[..., 0], Af[..., 1]
    d[..., 0] = 1.0
    d[..., 1] = A1
    fwd = A1 * A2
    for a in range(2, ord_ - 1):
        d[..., a] = fwd
        fwd = fwd * Af[..., a]
    d[..., ord_ - 1] = fwd
    Aend = Af[..., ord_ - 1]
    aa = fwd * Aend
    bwd = Aend
    for a in range(ord_ - 2, 1, -1):
        d[..., a] = d[..., a] * bwd
        bwd = bwd * Af[..., a]
    d[..., 1] = d[..., 1] * bwd
    bwd = bwd * A2
    d[..., 0] = d[..., 0] * bwd
    return aa, d


def eval_AA_dAA(T, offsets, R, species=None):
    """PIBasis evaluate_ed (pibasis.jl:402-432): AA[env, iAA] and
    dAA[nbr, iAA, 3] (real part taken for Invariant)."""
    A, dphi = eval_A_dA(T, offsets, R, species)
    orders, iAA2iA = np.asarray(T["orders"]), np.asarray(T["iAA2iA"])
    nenv, nAA = A.shape[0], len(orders)
    real = T["prop"] == "Invariant"
    AA = np.zeros((nenv, nAA), dtype=np.complex128)
    dAA = np.zeros((len(R), nAA, 3), dtype=np.complex128)
    env_of = np.repeat(np.arange(nenv), np.diff(offsets))
    for o in np.unique(orders):
        rows = np.nonzero(orders == o)[0]
        if o == 0:
            AA[:, rows] = 1.0
            continue
        idx = iAA2iA[rows, :o]                              # (nr, o)
        aa, d = _local_adjoints(A[:, idx])                  # (nenv, nr), (nenv, nr, o)
        AA[:, rows] = aa
        # dAA[j, iAA] = sum_a dAAdA[a] * dA[v_a, j]
        acc = np.zeros((len(R), len(rows), 3), dtype=np.complex128)
        for a in range(o):
            acc += d[env_of][:, :, a, None] * dphi[:, idx[:, a], :]
        dAA[:, rows, :] = acc
    if real:
        return AA.real.copy(), dAA.real.copy()
    return AA, dAA


def _a2b_csc(T):
    a2b = T["a2b"]
    rows, cols, vals = (np.asarray(a2b["rows"]), np.asarray(a2b["cols"]),
                        np.asarray(a2b["vals"]))
    order = np.lexsort((rows, cols))
    return rows[order], cols[order], vals[order], tuple(a2b["shape"])


def eval_B(T, AA):
    """genmul!(C, A2Bmap, AA, (a,b)->real(a*b)) (symmbasis.jl:248-264):
    column-major scatter in the CSC order."""
    rows, cols, vals, (nB, _) = _a2b_csc(T)
    if T["prop"] == "Invariant":
        B = np.zeros((AA.shape[0], nB))
        for r, c, v in zip(rows, cols, vals):
            B[:, r] += v * AA[:, c]
        return B
    B = np.zeros((AA.shape[0], nB, 3))
    for r, c, v in zip(rows, cols, vals):
        B[:, r, :] += (np.asarray(v)[None, :] * AA[:, c, None]).real
    return B


def eval_dB(T, dAA):
    """genmul(A2Bmap, dAA, (a,b)->real(coco_o_daa(a,b))) (symmbasis.jl:332):
    dB[nbr, iB, 3] (Invariant) or dB[nbr, iB, 3(out), 3(space)]."""
    rows, cols, vals, (nB, _) = _a2b_csc(T)
    if T["prop"] == "Invariant":
        dB = np.zeros((dAA.shape[0], nB, 3))
        for r, c, v in zip(rows, cols, vals):
            dB[:, r, :] += v * dAA[:, c, :]
        return dB
    dB = np.zeros((dAA.shape[0], nB, 3, 3))
    for r, c, v in zip(rows, cols, vals):
        dB[:, r, :, :] += (np.asarray(v)[None, :, None] * dAA[:, c, None, :]).real
    return dB


# ---------------------------------------------------------------- model


def ctilde(T, c):
    """c~ = A2Bmap^T c (genmul! transposed, symmbasis.jl:267-285)."""
    rows, cols, vals, (nB, nAA) = _a2b_csc(T)
    c = np.asarray(c, dtype=np.float64)
    vec = T["prop"] != "Invariant"
    out = np.zeros((nAA, 3), dtype=np.complex128) if vec else np.zeros(nAA)
    for r, cc, v in zip(rows, cols, vals):
        out[cc] = out[cc] + np.asarray(v) * c[r]
    return out


def energy(T, c, offsets, R, species=None):
    """ProductEvaluator evaluate (evaluator.jl:121-147).  Invariant: E[env];
    EuclideanVector: phi[env, 3]."""
    ct = ctilde(T, c)
    A = eval_A(T, offsets, R, species)
    orders, iAA2iA = np.asarray(T["orders"]), np.asarray(T["iAA2iA"])
    vec = T["prop"] != "Invariant"
    nenv = A.shape[0]
    val = np.zeros((nenv, 3)) if vec else np.zeros(nenv)
    i0 = 0
    if orders[0] == 0:
        val = val + (ct[0].real if vec else ct[0])
        i0 = 1
    for iAA in range(i0, len(orders)):
        aa = A[:, iAA2iA[iAA, 0]]
        for t in range(1, orders[iAA]):
            aa = aa * A[:, iAA2iA[iAA, t]]
        if vec:
            val = val + (aa[:, None] * ct[iAA][None, :]).real
        else:
            val = val + aa.real * ct[iAA]
    return val


def _dAco(T, ct, A):
    """Stage 2 of _rrule_evaluate (evaluator.jl:174-185)."""
    orders, iAA2iA = np.asarray(T["orders"]), np.asarray(T["iAA2iA"])
    vec = T["prop"] != "Invariant"
    nenv, nA = A.shape
    dAco = np.zeros((nenv, nA, 3) if vec else (nenv, nA), dtype=np.complex128)
    i0 = 1 if orders[0] == 0 else 0
    for iAA in range(i0, len(orders)):
        o = orders[iAA]
        idx = iAA2iA[iAA, :o]
        _, d = _local_adjoints(A[:, idx])
        for t in range(o):
            if vec:
                dAco[:, idx[t], :] += d[:, t, None] * ct[iAA][None, :]
            else:
                dAco[:, idx[t]] += d[:, t] * ct[iAA]
    return dAco


def grad_config(T, c, offsets, R, species=None):
    """grad_config via _rrule_evaluate(_One(), ...) (evaluator.jl:150-200):
    +grad E w.r.t. each neighbour vector: g[nbr, 3] (Invariant) or the
    Jacobian g[nbr, 3(out), 3(space)] (EuclideanVector)."""
    ct = ctilde(T, c)
    A, dphi = eval_A_dA(T, offsets, R, species)
    dAco = _dAco(T, ct, A)
    env_of = np.repeat(np.arange(A.shape[0]), np.diff(offsets))
    if T["prop"] == "Invariant":
        return np.einsum("ja,jad->jd", dAco[env_of], dphi).real
    return np.einsum("jap,jad->jpd", dAco[env_of], dphi).real


def adjoint_EVAL_D(T, c, offsets, R, w, species=None):
    """adjoint_EVAL_D (evaluator.jl:204-244) for one environment (Invariant):
    dB_k = sum_j w_j . dB_k/dr_j via the AA adjoint."""
    assert len(offsets) == 2
    A, dphi = eval_A_dA(T, offsets, R, species)
    dAw = np.einsum("jad,jd->a", dphi, np.asarray(w))
    orders, iAA2iA = np.asarray(T["orders"]), np.asarray(T["iAA2iA"])
    dAAw = np.zeros(len(orders))
    i0 = 1 if orders[0] == 0 else 0
    for iAA in range(i0, len(orders)):
        o = orders[iAA]
        idx = iAA2iA[iAA, :o]
        _, d = _local_adjoints(A[0, idx][None, :])
        for t in range(o):
            dAAw[iAA] += (dAw[idx[t]] * d[0, t]).real
    rows, cols, vals, (nB, _) = _a2b_csc(T)
    dB = np.zeros(nB)
    for r, cc, v in zip(rows, cols, vals):
        dB[r] += v * dAAw[cc]
    return dB
