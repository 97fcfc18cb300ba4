"""FP64 CPU restatement of the control-variate MLP (test infrastructure only).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
MLP kernels.  The network and its parameter packing follow
include/mbs_hip.h (mbs_mlp_*): layers n_in -> depth x width (GELU, exact erf
form) -> n_out, parameters per layer W [in][out] row-major then b [out];
loss = mean of squared errors over rows x outputs.  With bf16=True every GEMM
operand is rounded to BF16 (round to nearest even) exactly as the device
loads it, and activations are held in FP32 as on the device, so the only
remaining difference is the order of FP32 accumulation.
"""
import numpy as np
from scipy.special import erf


def layer_shapes(n_in, width, depth, n_out):
    return [(n_in if l == 0 else width, n_out if l == depth else width) for l in range(depth + 1)]


def unpack(params, shapes):
    out, off = [], 0
    for i, o in shapes:
        W = params[off:off + i * o].reshape(i, o)
        off += i * o
        b = params[off:off + o]
        off += o
        out.append((W, b))
    assert off == len(params)
    return out


def bf16(x):
    """Rounds to BF16 (nearest even) and returns float64 values."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def _f32(x, on):
    return x.astype(np.float32).astype(np.float64) if on else x


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(params, shapes, x, bf=False):
    """Returns (y, cache)."""
    layers = unpack(np.asarray(params, dtype=np.float64), shapes)
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    a = [_f32(np.asarray(x, dtype=np.float64), bf)]
    z = []
    for l, (W, b) in enumerate(layers):
        zl = _f32(q(a[-1]) @ q(W) + b, bf)
        if l < len(layers) - 1:
            z.append(zl)
            a.append(_f32(gelu(zl), bf))
        else:
            y = zl
    return y, (layers, a, z)


def loss_and_grad(params, shapes, x, t, bf=False):
    y, (layers, a, z) = forward(params, shapes, x, bf)
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    rows, n_out = y.shape
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(d * d) / (rows * n_out))
    dout = _f32(2.0 * d / (rows * n_out), bf)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        if l < len(layers) - 1:
            dout = _f32(dout * gelu_grad(z[l]), bf)
        dW = q(a[l]).T @ q(dout)
        db = dout.sum(axis=0)
        grads[l] = (dW, db)
        if l > 0:
            dout = _f32(q(dout) @ q(W).T, bf)
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat
