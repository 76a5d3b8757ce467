"""ctypes binding of libacemlp.so (include/acemlp.h): the surrogate's MLP
ensemble with hand-written BF16 MFMA kernels."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libacemlp.so")
INP = 160
HID = 128
_lib = None


class MlpError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MlpError(f"libacemlp.so not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp, f32p = C.c_void_p, C.POINTER(C.c_float)
    lib.acemlp_nparam.restype = C.c_int
    lib.acemlp_input_width.restype = C.c_int
    lib.acemlp_last_error.restype = C.c_char_p
    lib.acemlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]
    lib.acemlp_destroy.argtypes = [vp]
    lib.acemlp_destroy.restype = None
    lib.acemlp_get_params.argtypes = [vp, f32p, C.c_int64]
    lib.acemlp_set_params.argtypes = [vp, f32p, C.c_int64]
    lib.acemlp_predict.argtypes = [vp, C.c_int64, vp, vp, vp]
    lib.acemlp_predict_grad.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    lib.acemlp_train_step.argtypes = [vp, C.c_int64, vp, vp, C.c_float, C.c_float, f32p, vp]
    lib.acemlp_loss_grad.argtypes = [vp, C.c_int64, vp, vp, f32p, f32p]
    lib.acemlp_layout_probe.argtypes = [f32p]
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise MlpError(load().acemlp_last_error().decode())


def _vp(t):
    return C.c_void_p(t.data_ptr())


class MLPEnsemble:
    """nmemb MLPs (nin -> 128 -> 128 -> 128 -> 1, SiLU) on the device."""

    def __init__(self, nmemb=4, nin=151, device=0, seed=0):
        self.lib = load()
        self.nmemb, self.nin, self.device = nmemb, nin, device
        self.nparam = self.lib.acemlp_nparam()
        ctx = C.c_void_p()
        check(self.lib.acemlp_create(nmemb, nin, device, seed, C.byref(ctx)))
        self.ctx = ctx

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib.acemlp_destroy(self.ctx)
            self.ctx = None

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def get_params(self):
        out = np.zeros(self.nmemb * self.nparam, dtype=np.float32)
        check(self.lib.acemlp_get_params(self.ctx, out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out.reshape(self.nmemb, self.nparam)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1)
        check(self.lib.acemlp_set_params(self.ctx, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))

    def predict(self, Xn):
        """Xn: bf16 tensor [rows, 160] on the device -> y [nmemb, rows] fp32."""
        y = torch.empty((self.nmemb, Xn.shape[0]), dtype=torch.float32, device=Xn.device)
        check(self.lib.acemlp_predict(self.ctx, Xn.shape[0], _vp(Xn), _vp(y), self._stream()))
        return y

    def predict_grad(self, Xn):
        y = torch.empty((self.nmemb, Xn.shape[0]), dtype=torch.float32, device=Xn.device)
        dX = torch.empty((self.nmemb, Xn.shape[0], INP), dtype=torch.float32, device=Xn.device)
        check(self.lib.acemlp_predict_grad(self.ctx, Xn.shape[0], _vp(Xn), _vp(y), _vp(dX), self._stream()))
        return y, dX

    def train_step(self, Xn, t, lr, weight_decay=0.0, want_loss=False):
        loss = np.zeros(self.nmemb, dtype=np.float32)
        check(self.lib.acemlp_train_step(self.ctx, Xn.shape[0], _vp(Xn), _vp(t), float(lr), float(weight_decay),
                                         loss.ctypes.data_as(C.POINTER(C.c_float)) if want_loss else None,
                                         self._stream()))
        return loss if want_loss else None

    def loss_grad(self, Xn, t):
        g = np.zeros(self.nmemb * self.nparam, dtype=np.float32)
        loss = np.zeros(self.nmemb, dtype=np.float32)
        check(self.lib.acemlp_loss_grad(self.ctx, Xn.shape[0], _vp(Xn), _vp(t),
                                        g.ctypes.data_as(C.POINTER(C.c_float)),
                                        loss.ctypes.data_as(C.POINTER(C.c_float))))
        return g.reshape(self.nmemb, self.nparam), loss


def layout_probe():
    out = np.zeros(256, dtype=np.float32)
    check(load().acemlp_layout_probe(out.ctypes.data_as(C.POINTER(C.c_float))))
    return out.reshape(16, 16)
