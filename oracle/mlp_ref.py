"""float64 numpy restatement of the learned-form MLP -- TEST INFRASTRUCTURE ONLY.

No counterpart in the reference.  The network is in -> hidden -> hidden -> C + 1
with ReLU; loss = softmax cross-entropy over the first C outputs + stop_w * squared
error of output C.  With emulate_bf16 the values the GPU rounds to bfloat16
(weights, hidden activations, output and hidden gradients) are rounded at the same
points, so the GPU's gradients differ from these only by float32 summation order.
The analytic backward is pinned separately against central finite differences.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split(flat, width, hidden, n_classes):
    shapes = [(hidden, width), (hidden,), (hidden, hidden), (hidden,), (n_classes + 1, hidden), (n_classes + 1,)]
    out, o = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(np.asarray(flat[o:o + n], dtype=np.float64).reshape(sh))
        o += n
    return out


def forward(params, X, emulate_bf16=True):
    W1, b1, W2, b2, W3, b3 = params
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    H1 = q(np.maximum(X @ q(W1).T + b1, 0.0))
    H2 = q(np.maximum(H1 @ q(W2).T + b2, 0.0))
    Z = H2 @ q(W3).T + b3
    return H1, H2, Z


def loss_grad(params, X, labels, stop_t, stop_w, emulate_bf16=True):
    W1, b1, W2, b2, W3, b3 = params
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    H1, H2, Z = forward(params, X, emulate_bf16)
    B, NO = Z.shape
    C = NO - 1
    logits = Z[:, :C]
    m = logits.max(1, keepdims=True)
    e = np.exp(logits - m)
    se = e.sum(1, keepdims=True)
    p = e / se
    xent = float(np.mean(np.log(se[:, 0]) + m[:, 0] - logits[np.arange(B), labels]))
    diff = Z[:, C] - stop_t
    mse = float(np.mean(diff ** 2))
    dZ = np.zeros_like(Z)
    dZ[:, :C] = p
    dZ[np.arange(B), labels] -= 1.0
    dZ[:, :C] /= B
    dZ[:, C] = 2.0 * stop_w * diff / B
    dZ = q(dZ)
    dW3 = dZ.T @ H2
    db3 = dZ.sum(0)
    dH2 = q((dZ @ q(W3)) * (H2 > 0))
    dW2 = dH2.T @ H1
    db2 = dH2.sum(0)
    dH1 = q((dH2 @ q(W2)) * (H1 > 0))
    dW1 = dH1.T @ X
    db1 = dH1.sum(0)
    return xent, mse, [dW1, db1, dW2, db2, dW3, db3]


def total_loss(params, X, labels, stop_t, stop_w):
    xent, mse, _ = loss_grad(params, X, labels, stop_t, stop_w, emulate_bf16=False)
    return xent + stop_w * mse


def top2_conf(Z):
    """Two largest logits of the first C outputs (ties to the lower index), their summed
    softmax probability, and the stop output."""
    C = Z.shape[1] - 1
    logits = Z[:, :C]
    order = np.lexsort((np.broadcast_to(np.arange(C), logits.shape), -logits), axis=1)[:, :2]
    m = logits.max(1, keepdims=True)
    p = np.exp(logits - m) / np.exp(logits - m).sum(1, keepdims=True)
    conf = np.take_along_axis(p, order, 1).sum(1)
    return order, conf, Z[:, C]
