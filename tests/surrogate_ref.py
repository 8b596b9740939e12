"""CPU reference of the surrogate network's forward and backward passes (test infrastructure).

FP64 arithmetic with the GPU's roundings made explicit: every GEMM operand is rounded to BF16
(round to nearest even), accumulation is exact-ish (FP64 here, FP32 on the GPU), SiLU and the loss
follow kqed_surrogate.hip.  Layer shapes: K = (16, 128, 128, 128, 128), N = (128, 128, 128, 128, 48).
"""
import numpy as np


def bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def exact(a):
    """No operand rounding (for checking the reference itself against finite differences)."""
    return np.asarray(a, np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward_backward(layers, X3, tgt, wgt, rnd=bf16):
    """layers: [(W [K][N], b [N])]; X3 [B][3]; tgt, wgt [B][34].  Returns loss, [(dW, db)].
    rnd: operand rounding of the GEMMs (bf16 as on the GPU, or exact)."""
    B = X3.shape[0]
    X = np.zeros((B, 16))
    X[:, :3] = X3
    A = [X]
    Z = []
    for l, (W, b) in enumerate(layers):
        z = rnd(A[-1]) @ rnd(W) + np.asarray(b, np.float64)
        Z.append(z)
        if l < 4:
            A.append(silu(z if rnd is exact else np.float32(z).astype(np.float64)))
    nout = tgt.shape[1]
    d = Z[4][:, :nout] - tgt
    loss = float((wgt * d * d).sum() / (B * nout))
    dZ = np.zeros_like(Z[4])
    dZ[:, :nout] = 2.0 * wgt * d / (B * nout)
    grads = [None] * 5
    for l in range(4, -1, -1):
        W, _ = layers[l]
        dW = rnd(A[l]).T @ rnd(dZ)
        db = dZ.sum(axis=0)
        grads[l] = (dW, db)
        if l == 0:
            break
        dA = rnd(dZ) @ rnd(W).T
        dZ = dA * dsilu(Z[l - 1])
    return loss, grads
