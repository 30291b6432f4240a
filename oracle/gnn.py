"""Numpy restatement of the warm-start network's layers (forward and hand-derived
backward) for checking the device kernels' gradients. There is no reference
counterpart (the reference has no learned component); the restatement is pinned by a
central finite-difference check in the tests.

Layer: pre = h16 Ws16 + agg(h16) Wn16 + b, out = relu(pre), agg(h)_i = mean over the
neighbours j != i of h_j; inputs and weights are rounded to float16 first, arithmetic in
float64 here.
"""
import numpy as np


def aggregate(indptr, indices, X, adjoint=False):
    n = indptr.shape[0] - 1
    deg = np.diff(indptr) - 1
    Y = np.zeros_like(X, dtype=np.float64)
    for i in range(n):
        js = indices[indptr[i]:indptr[i + 1]]
        js = js[js != i]
        if adjoint:
            w = 1.0 / np.maximum(deg[js], 1)
            Y[i] = (w[:, None] * X[js]).sum(0)
        elif deg[i] > 0:
            Y[i] = X[js].sum(0) / deg[i]
    return Y


def f16(a):
    return np.asarray(a, dtype=np.float16).astype(np.float64)


def layer_forward(indptr, indices, h, Ws, Wn, b):
    h16 = f16(h)
    agg = f16(aggregate(indptr, indices, h16))
    pre = h16 @ f16(Ws) + agg @ f16(Wn) + b
    return np.maximum(pre, 0.0), (h16, agg, pre)


def layer_backward(indptr, indices, cache, Ws, Wn, dout):
    h16, agg, pre = cache
    dpre = dout * (pre > 0)
    dWs = h16.T @ dpre
    dWn = agg.T @ dpre
    db = dpre.sum(0)
    dh = dpre @ f16(Ws).T + aggregate(indptr, indices, dpre @ f16(Wn).T, adjoint=True)
    return dh, dWs, dWn, db
