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


# ---- FP8 E4M3 (OCP: bias 7, largest finite 448, no infinities) ----
def _e4m3_table():
    b = np.arange(256)
    e, m = (b >> 3) & 15, b & 7
    v = np.where(e == 0, m * 2.0 ** -9, (1.0 + m / 8.0) * 2.0 ** (e - 7.0))
    v = np.where(b & 0x80, -v, v)
    v[(e == 15) & (m == 7)] = np.nan
    return v


E4M3 = _e4m3_table()
_POS = E4M3[:0x7F]  # codes 0x00..0x7e, increasing


def fp8(x):
    """Round to the nearest E4M3 value (ties to the even code), saturating at
    +-448; returns float64."""
    a = np.abs(np.asarray(x, dtype=np.float64))
    a = np.minimum(a, 448.0)
    hi = np.clip(np.searchsorted(_POS, a, side="left"), 0, _POS.size - 1)
    lo = np.clip(hi - 1, 0, _POS.size - 1)
    dl, dh = a - _POS[lo], _POS[hi] - a
    pick = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(lo % 2 == 0, lo, hi)))
    return np.copysign(_POS[pick], x)


def predict_fp8(params, X, norm, nin, nout, width, depth):
    """Denormalised outputs of the FP8 inference: BF16 first and last layers,
    hidden-to-hidden layers on FP8 weights (per-row scale max|w| / 448) and FP8
    activations, FP64 sums."""
    in_mean, in_std, out_mean, out_std, _ = norm
    layers = unpack(np.asarray(params, dtype=np.float64), nin, nout, width, depth)
    # the device computes 1 / std in FP32 and multiplies
    rstd = (np.float32(1.0) / np.asarray(in_std, dtype=np.float32)).astype(np.float64)
    a = bf16(((X - in_mean) * rstd).astype(np.float32))
    for l in range(depth):
        W, b = layers[l]
        last_hidden = l == depth - 1
        if l == 0:
            z = a @ bf16(W).T + b
        else:
            sc = (np.abs(W.astype(np.float32)).max(axis=1) / np.float32(448.0)).astype(np.float32)
            sc = np.where(sc > 0, sc, np.float32(1.0))
            q = fp8((W.astype(np.float32) / sc[:, None]).astype(np.float32))
            z = (a @ q.T) * sc.astype(np.float64) + b
        h = gelu(z)
        a = bf16(h) if last_hidden else fp8(h)
    W, b = layers[depth]
    out = a @ bf16(W).T + b
    return out * out_std + out_mean


def inverse_target(z, delta):
    """Increment d from the transformed target z = sign(d) ln(1 + |d| / delta)."""
    return np.sign(z) * np.expm1(np.abs(z)) * delta
