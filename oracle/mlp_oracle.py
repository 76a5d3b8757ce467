"""NumPy reference of the surrogate MLP — TEST INFRASTRUCTURE ONLY.

Forward and backward of IN(160, zero-padded) -> 128 -> 128 -> 128 -> 1 with
SiLU, in float64 arithmetic, rounding to bfloat16 exactly where the device
kernels do (inputs, weights of layers 1-3, hidden activations, and the
backpropagated pre-activation gradients), so device results agree up to
FP32 accumulation order.  Parameter layout of one member matches
libacemlp.so: W1[128][160], b1, W2[128][128], b2, W3[128][128], b3, w4[128], b4.
"""
import numpy as np

INP, HID = 160, 128
OFF = {}
_o = 0
for name, size in (("W1", HID * INP), ("b1", HID), ("W2", HID * HID), ("b2", HID), ("W3", HID * HID),
                   ("b3", HID), ("w4", HID), ("b4", 1)):
    OFF[name] = (_o, size)
    _o += size
NPARAM = ((_o + 15) // 16) * 16


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    x = np.asarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b >> 16) & 1) + 0x7FFF
    return ((b + r) & 0xFFFF0000).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    g = {k: p[o:o + s] for k, (o, s) in OFF.items()}
    g["W1"] = g["W1"].reshape(HID, INP)
    g["W2"] = g["W2"].reshape(HID, HID)
    g["W3"] = g["W3"].reshape(HID, HID)
    return g


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_d(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X):
    """X: [rows, 160] (already bf16 values).  Returns y and the cache."""
    g = unpack(p)
    h = bf16(X)
    cache = {"h0": h}
    for i, (W, b) in enumerate((("W1", "b1"), ("W2", "b2"), ("W3", "b3")), start=1):
        z = h @ bf16(g[W]).T + g[b]
        h = bf16(silu(z))
        cache[f"z{i}"] = z
        cache[f"h{i}"] = h
    y = h @ g["w4"] + g["b4"][0]
    return y, cache


def backward(p, cache, dy):
    """dL/dparams and dL/dX for dL/dy = dy [rows]."""
    g = unpack(p)
    grads = np.zeros(NPARAM)
    h3 = cache["h3"]
    grads[OFF["w4"][0]:OFF["w4"][0] + HID] = dy @ h3
    grads[OFF["b4"][0]] = dy.sum()
    dz = bf16(dy[:, None] * g["w4"][None, :] * silu_d(cache["z3"]))
    for i, (W, b) in ((3, ("W3", "b3")), (2, ("W2", "b2")), (1, ("W1", "b1"))):
        hin = cache[f"h{i-1}"]
        o, s = OFF[W]
        grads[o:o + s] = (dz.T @ hin).reshape(-1)
        ob, _ = OFF[b]
        grads[ob:ob + HID] = dz.sum(0)
        dh = dz @ bf16(g[W])
        if i > 1:
            dz = bf16(dh * silu_d(cache[f"z{i-1}"]))
        else:
            dX = dh
    return grads, dX
