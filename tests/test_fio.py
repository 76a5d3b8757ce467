"""ACE.jl model files: the reference's dict layout, round trips, the
primitive encodings the reader accepts, and index conventions pinned by a
hand-written file (CPU)."""
import json

import numpy as np
import pytest

import ace_oracle as O
from acehip import fio, synth
from acehip.symmbasis import LinearACEModel
from acehip.tables import basis_tables
from helpers import basis_for, random_batch, rel_err


def _tables_equal(T1, T2):
    for k in ("aspec", "orders", "iAA2iA"):
        assert np.array_equal(np.asarray(T1[k]), np.asarray(T2[k])), k
    for k in ("A", "B", "C"):
        assert np.array_equal(T1["rn"][k], T2["rn"][k]), k
    for k in ("pl", "pr", "tl", "tr", "p", "r0"):
        assert T1["rn"][k] == T2["rn"][k], k
    assert abs(T1["rn"]["rcut"] - T2["rn"]["rcut"]) < 1e-12
    assert T1["L"] == T2["L"] and T1["prop"] == T2["prop"] and T1["categories"] == T2["categories"]
    a, b = T1["a2b"], T2["a2b"]
    assert tuple(a["shape"]) == tuple(b["shape"])
    for k in ("rows", "cols", "vals"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.parametrize("args", [dict(order=3, maxdeg=6, L=4), dict(order=3, maxdeg=6, L=3, species=2),
                                  dict(order=3, maxdeg=6, L=3, prop="EuclideanVector")])
def test_round_trip(args):
    basis, T = basis_for(args["order"], args["maxdeg"], args["L"], prop=args.get("prop", "Invariant"),
                         species=args.get("species", 0))
    c = synth.coefficients(len(basis), 3)
    model = LinearACEModel(basis, c)
    text = json.dumps(fio.write_dict(model))
    m2 = fio.read_dict(json.loads(text))
    assert isinstance(m2, LinearACEModel) and np.array_equal(m2.c, c)
    T2 = basis_tables(m2.basis)
    _tables_equal(T, T2)
    if T["prop"] == "Invariant":
        ns = args.get("species", 0)
        off, R, sp = random_batch(6, 5, nspecies=ns)
        assert np.array_equal(O.energy(T, c, off, R, sp), O.energy(T2, m2.c, off, R, sp))


def test_layout_follows_the_reference():
    basis, _ = basis_for(3, 6, 4)
    D = fio.write_dict(LinearACEModel(basis, np.zeros(len(basis))))
    assert D["__id__"] == "ACE_LinearACEModel" and set(D) == {"__id__", "basis", "c", "evaluator"}
    assert D["evaluator"] == {"__id__": "ACE_ProductEvaluator"}
    S = D["basis"]
    assert S["__id__"] == "ACE_SymmetricBasis" and set(S) == {"__id__", "pibasis", "A2Bmap", "symgrp", "isreal"}
    assert S["symgrp"] == {"__id__": "ACE_O3", "lsym": "l", "msym": "m"}
    P = S["pibasis"]
    assert P["__id__"] == "ACE_PIBasis" and set(P) == {"__id__", "basis1p", "spec", "real"}
    assert P["spec"]["__id__"] == "ACE_PIBasisSpec" and set(P["spec"]) == {"__id__", "orders", "iAA2iA"}
    B1 = P["basis1p"]
    assert B1["__id__"] == "ACE_Product1pBasis" and set(B1) == {"__id__", "bases", "indices"}
    Rn, Y = B1["bases"]
    keys = {"__id__", "syms", "basis", "fval", "spec", "degrees", "label"}
    assert set(Rn) == keys and set(Y) == keys
    assert [f["__id__"] for f in Rn["basis"]["F"]] == ["ACE_Lambda", "ACE_Lambda", "ACE_OrthPolyBasis"]
    assert Rn["basis"]["F"][1]["exstr"] == "r -> ((1+1.0)/(1+r))^2"
    assert set(Rn["basis"]["F"][2]) == {"__id__", "T", "pr", "tr", "pl", "tl", "A", "B", "C", "tdf", "ww"}
    assert Y["basis"]["__id__"] == "ACE_SHBasis" and Y["basis"]["maxL"] == 4
    assert Y["spec"][:4] == [{"l": 0, "m": 0}, {"l": 1, "m": -1}, {"l": 1, "m": 0}, {"l": 1, "m": 1}]
    nz = S["A2Bmap"]["nzval"][0]
    assert nz["__id__"] == "ACE_Invariant" and set(nz) == {"__id__", "val", "T"}


def test_reader_accepts_other_primitive_encodings():
    basis, T = basis_for(3, 6, 4)
    c = synth.coefficients(len(basis), 2)
    D = json.loads(json.dumps(fio.write_dict(LinearACEModel(basis, c))))
    P = D["basis"]["pibasis"]["spec"]
    it = fio._read_array(P["iAA2iA"])
    P["iAA2iA"] = it.T.tolist()                       # plain JSON: list of columns
    A = D["basis"]["A2Bmap"]
    colptr = np.asarray(A["colptr"])
    cols = np.repeat(np.arange(len(colptr) - 1), np.diff(colptr)) + 1
    D["basis"]["A2Bmap"] = {"m": A["m"], "n": A["n"], "I": A["rowval"], "J": cols.tolist(),
                            "V": [v["val"] for v in A["nzval"]]}       # COO, plain values
    D["c"] = c.tolist()
    m2 = fio.read_dict(D)
    _tables_equal(T, basis_tables(m2.basis))
    assert np.array_equal(m2.c, c)


def test_reader_rejects_unsupported_parts():
    basis, _ = basis_for(2, 4, 2)
    D = fio.write_dict(LinearACEModel(basis, np.zeros(len(basis))))
    bad = json.loads(json.dumps(D))
    bad["basis"]["pibasis"]["basis1p"]["bases"][0]["basis"]["F"][1]["exstr"] = "r -> exp(- 2.0 * (r / 1.0 - 1))"
    with pytest.raises(ValueError, match="polytransform"):
        fio.read_dict(bad)
    bad = json.loads(json.dumps(D))
    bad["basis"]["pibasis"]["basis1p"]["bases"][1]["spec"][1] = {"l": 1, "m": 1}
    with pytest.raises(ValueError, match="spec order"):
        fio.read_dict(bad)
    bad = json.loads(json.dumps(D))
    bad["basis"]["symgrp"] = {"__id__": "ACE_O3O3"}
    with pytest.raises(ValueError, match="symmetry group"):
        fio.read_dict(bad)


def test_hand_written_file_pins_index_conventions():
    """A file written by hand in the reference's conventions (1-based
    indices, iAA2iA descending with zero padding, Julia column-major
    matrices, CSC A2Bmap): two B functions with closed-form energies."""
    basis, T0 = basis_for(2, 4, 2)     # borrow a valid radial basis and its JSON component
    D0 = fio.write_dict(LinearACEModel(basis, np.zeros(len(basis))))
    Rn = D0["basis"]["pibasis"]["basis1p"]["bases"][0]
    Y = {"__id__": "ACE_B1pComponent", "syms": ["l", "m"],
         "basis": {"__id__": "ACE_SHBasis", "T": {"__id__": "Type", "T": "Float64"}, "maxL": 1},
         "fval": {"__id__": "ACE_StaticGet", "expr": "ACE.Transforms.GetVal{:rr}"},
         "spec": [{"l": 0, "m": 0}, {"l": 1, "m": -1}, {"l": 1, "m": 0}, {"l": 1, "m": 1}],
         "degrees": [0, 1, 1, 1], "label": "Ylm"}
    # 1p functions: A1 = R1 Y00, A2 = R2 Y00, A3 = R1 Y1-1, A4 = R1 Y10, A5 = R1 Y11 ([iRn, iY], 1-based)
    indices = [[1, 1], [2, 1], [1, 2], [1, 3], [1, 4]]
    # AA rows (descending factors): 1: A2 (order 1); 2: A2*A1; 3: A5*A3 (m sum 0)
    orders = [1, 2, 2]
    cols = [[2, 2, 5], [0, 1, 3]]      # Julia matrix iAA2iA[3x2] as its list of columns
    # B1 = 0.5 AA1 + 2 AA2 ; B2 = -1.5 AA3   (CSC over the 3 AA columns)
    A2B = {"m": 2, "n": 3, "colptr": [1, 2, 3, 4], "rowval": [1, 1, 2],
           "nzval": [{"__id__": "ACE_Invariant", "val": v, "T": {"__id__": "Type", "T": "Float64"}}
                     for v in (0.5, 2.0, -1.5)]}
    D = {"__id__": "ACE_LinearACEModel",
         "basis": {"__id__": "ACE_SymmetricBasis",
                   "pibasis": {"__id__": "ACE_PIBasis",
                               "basis1p": {"__id__": "ACE_Product1pBasis", "bases": [Rn, Y], "indices": indices},
                               "spec": {"__id__": "ACE_PIBasisSpec", "orders": orders, "iAA2iA": cols},
                               "real": True},
                   "A2Bmap": A2B, "symgrp": {"__id__": "ACE_O3", "lsym": "l", "msym": "m"}, "isreal": True},
         "c": [1.25, -0.75], "evaluator": {"__id__": "ACE_ProductEvaluator"}}
    m = fio.read_dict(json.loads(json.dumps(D)))
    T = basis_tables(m.basis)
    off, R, _ = random_batch(5, 9)
    E = O.energy(T, m.c, off, R)
    # closed form from the radial basis and Ylm directly
    rn = T0["rn"]
    for e in range(5):
        Rj = R[off[e]:off[e + 1]]
        r = np.linalg.norm(Rj, axis=1)
        P = O.rn_eval(rn, O.transform(rn, r))            # (nX, nr)
        Yv = O.ylm_eval(1, Rj)                           # (nX, 4) in index_y order
        A1 = np.sum(P[:, 0] * Yv[:, 0])
        A2 = np.sum(P[:, 1] * Yv[:, 0])
        A3, A5 = np.sum(P[:, 0] * Yv[:, 1]), np.sum(P[:, 0] * Yv[:, 3])
        B1 = 0.5 * A2.real + 2.0 * (A2 * A1).real
        B2 = -1.5 * (A5 * A3).real
        assert abs(E[e] - (1.25 * B1 - 0.75 * B2)) < 1e-12 * max(1.0, abs(E[e]))
