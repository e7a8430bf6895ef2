"""CPU reference of the chemistry surrogate's network (test infrastructure).

Forward pass, loss and backward pass of the MLP that the HIP training and
inference kernels implement, in numpy: hidden layers z = x W^T + b, h = GELU(z)
(tanh form), a linear output layer, and the weighted mean-square loss on
normalised targets.  With bf16=True every quantity the device stores in BF16
(normalised inputs, weights, hidden activations, the output gradient and the
back-propagated pre-activation gradients) is rounded to BF16 here as well
(round to nearest even), with FP64 accumulation, so the device gradients can
be compared entry by entry.  The surrogate has no counterpart in the reference
code base; this file only checks the device kernels.
"""
from __future__ import annotations

import numpy as np

C_GELU = 0.7978845608028654


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * z * z)


def unpack(params, nin, nout, width, depth):
    """Per-layer (W [out][in], b [out]) views of the flat parameter vector."""
    out, off = [], 0
    for l in range(depth + 1):
        i = nin if l == 0 else width
        o = width if l < depth else nout
        W = params[off:off + o * i].reshape(o, i)
        off += o * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    assert off == params.size
    return out


def forward_backward(params, X, Yt, norm, nin, nout, width, depth, bf=True):
    """Loss and flat gradient for raw features X [n][nin] and targets Yt [n][nout].

    norm = (in_mean, in_std, out_mean, out_std, weight)."""
    in_mean, in_std, out_mean, out_std, w = norm
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float64), nin, nout, width, depth)
    n = X.shape[0]
    x = r((X - in_mean) * np.float32(1.0) / in_std)
    acts, zs = [x], []
    for l in range(depth):
        W, b = layers[l]
        z = acts[-1] @ r(W).T + b
        zs.append(z)
        acts.append(r(gelu(z)))
    W, b = layers[depth]
    out = acts[-1] @ r(W).T + b
    t = (Yt - out_mean) / out_std
    res = out - t
    loss = float((w * res * res).sum() / n)
    d = r(2.0 * w * res / n)
    grads = [None] * (depth + 1)
    grads[depth] = (d.T @ acts[depth], d.sum(axis=0))
    dz = r((d @ r(W)) * gelu_grad(zs[depth - 1]))
    for l in range(depth - 1, -1, -1):
        grads[l] = (dz.T @ acts[l], dz.sum(axis=0))
        if l > 0:
            dz = r((dz @ r(layers[l][0])) * gelu_grad(zs[l - 1]))
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat, out


def predict(params, X, norm, nin, nout, width, depth, bf=True):
    """Denormalised outputs (raw increments) for raw features X."""
    in_mean, in_std, out_mean, out_std, _ = norm
    _, _, out = forward_backward(params, X, np.zeros((X.shape[0], nout)), norm, nin, nout, width,
                                 depth, bf)
    return out * out_std + out_mean
