"""acehip: host side of the MI355X basis-evaluation path for the Atomic
Cluster Expansion (ACE.jl semantics).  Set-up (basis construction) runs on the
CPU; every evaluation runs in ``libacehip.so`` on the GPU."""
from .onep import (Categorical1pBasis, Product1pBasis, Rn1pBasis, Rn_basis,
                   RnYlm_1pbasis, Ylm1pBasis, init1pspec)
from .pibasis import NoConstant, PIBasis, PIBasisSpec, SimpleSparseBasis, SparseBasis
from .coupling import EuclideanVector, Invariant
from .symmbasis import LinearACEModel, SymmetricBasis, symmetric_basis
from .fio import load_json, read_dict, save_json, write_dict

__all__ = ["Categorical1pBasis", "Product1pBasis", "Rn1pBasis", "Rn_basis",
           "RnYlm_1pbasis", "Ylm1pBasis", "init1pspec", "NoConstant", "PIBasis",
           "PIBasisSpec", "SimpleSparseBasis", "SparseBasis", "EuclideanVector", "Invariant",
           "LinearACEModel", "SymmetricBasis", "symmetric_basis", "load_json", "read_dict",
           "save_json", "write_dict"]
