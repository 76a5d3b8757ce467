"""ctypes wrapper of the C++ ProductEvaluator port (ace_cpu.cpp) —
TEST INFRASTRUCTURE ONLY (second oracle and timed CPU baseline)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import ace_oracle as O

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libace_cpu.so")
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class Tables(C.Structure):
    _fields_ = [("nr", C.c_int32), ("tp", C.c_int32), ("pl", C.c_int32), ("pr", C.c_int32),
                ("L", C.c_int32), ("r0", C.c_double), ("tl", C.c_double), ("tr", C.c_double),
                ("rA", _f64p), ("rB", _f64p), ("rC", _f64p), ("nA", C.c_int32), ("aspec", _i32p),
                ("nAA", C.c_int32), ("maxorder", C.c_int32), ("orders", _i32p), ("iAA2iA", _i32p),
                ("ctilde", _f64p), ("ncomp", C.c_int32)]


def build():
    import subprocess
    subprocess.run(["g++", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
                    "-o", LIB, os.path.join(_HERE, "ace_cpu.cpp")], check=True)


class CpuModel:
    """Invariant or EuclideanVector LinearACEModel on the CPU port; ``T`` =
    shared basis tables.  EuclideanVector: E [nenv, 3], g [nnbr, 3 (p), 3 (d)]."""

    def __init__(self, T, c):
        if not os.path.exists(LIB):
            build()
        self.lib = C.CDLL(LIB)
        self.lib.acecpu_eval.argtypes = [C.POINTER(Tables), C.c_int64, _i32p, _f64p, C.c_void_p,
                                         _f64p, _f64p, C.c_int, C.c_int]
        self.lib.acecpu_max_threads.restype = C.c_int
        rn = T["rn"]
        self._keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_int32 if dt == np.int32 else C.c_double))

        t = Tables()
        t.nr, t.tp, t.pl, t.pr, t.L = len(rn["A"]), int(rn["p"]), int(rn["pl"]), int(rn["pr"]), int(T["L"])
        t.r0, t.tl, t.tr = float(rn["r0"]), float(rn["tl"]), float(rn["tr"])
        t.rA, t.rB, t.rC = arr(rn["A"], np.float64), arr(rn["B"], np.float64), arr(rn["C"], np.float64)
        t.nA = len(T["aspec"])
        t.aspec = arr(T["aspec"], np.int32)
        t.nAA = len(T["orders"])
        t.maxorder = int(np.asarray(T["iAA2iA"]).shape[1])
        t.orders = arr(T["orders"], np.int32)
        t.iAA2iA = arr(T["iAA2iA"], np.int32)
        ct = O.ctilde(T, c)
        self.ncomp = 1 if T["prop"] == "Invariant" else 3
        if self.ncomp == 3:
            ct = np.stack([ct.real, ct.imag], axis=-1).reshape(-1)
        t.ctilde = arr(ct, np.float64)
        t.ncomp = self.ncomp
        self.t = t

    def max_threads(self):
        return self.lib.acecpu_max_threads()

    def eval(self, off, R, species=None, forces=True, nthreads=0):
        off = np.ascontiguousarray(off, dtype=np.int32)
        R = np.ascontiguousarray(R, dtype=np.float64)
        nenv = len(off) - 1
        nc = self.ncomp
        E = np.zeros(nenv) if nc == 1 else np.zeros((nenv, 3))
        g = (np.zeros_like(R) if nc == 1 else np.zeros((len(R), 3, 3))) if forces else np.zeros(1)
        sp = None if species is None else np.ascontiguousarray(species, dtype=np.int8)
        rc = self.lib.acecpu_eval(C.byref(self.t), nenv, off.ctypes.data_as(_i32p), R.ctypes.data_as(_f64p),
                                  None if sp is None else sp.ctypes.data_as(C.c_void_p),
                                  E.ctypes.data_as(_f64p), g.ctypes.data_as(_f64p), int(forces), int(nthreads))
        if rc != 0:
            raise RuntimeError("acecpu_eval failed")
        return (E, g) if forces else E
