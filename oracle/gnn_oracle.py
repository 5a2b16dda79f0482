"""CPU restatement of the learned surrogate's network -- TEST INFRASTRUCTURE.

Only the tests may use this module, as the checker of the HIP kernels in
sdem_amd/csrc/gnn_kernels.hip and gnn_api.cpp; nothing in the product
imports it.  The network has no counterpart in the reference.

It restates, in float64 numpy:
  * the forward pass: node / edge encoders, `layers` message-passing layers
    (e' = e + MLP_e([e, h_src, h_dst]); agg = sum of e' over each receiver's
    incoming edges; h' = h + MLP_n([h, agg])), decoder; every MLP is
    Linear-SiLU-Linear-SiLU-Linear;
  * the backward pass, written out by reverse-mode differentiation of the
    above, for the gradient of sum(out * dout) with respect to every
    parameter.
With bf16=True it rounds to bfloat16 (nearest even) exactly where the GPU
path hands values to its matrix cores: the GEMM inputs (features, packed
activations, weights, incoming gradients and SiLU-backward outputs); sums
stay in float64, so the GPU's fp32 accumulation differs from it only by
rounding.  With bf16=False the function is smooth and its backward pass is
pinned by finite differences (tests/test_gnn_oracle.py).

The parameter vector and its layout (per MLP the offsets of W1, b1, W2, b2,
W3, b3 and the dimensions in, hid, out; encoders take 16 padded inputs, the
decoder gives 16 outputs of which the first `out` are used) are the
library's interface (include/sdem_gnn.h).
"""
from __future__ import annotations

import numpy as np

PAD = 16


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def silu(x):
    return x / (1.0 + np.exp(-x))


def silu_grad(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


class Net:
    def __init__(self, params: np.ndarray, offsets: np.ndarray, dims: np.ndarray, layers: int, out: int,
                 bf16_inputs: bool = True):
        self.p = np.asarray(params, np.float64)
        self.off = np.asarray(offsets).reshape(-1, 6)
        self.dims = np.asarray(dims).reshape(-1, 3)
        self.L = layers
        self.out = out
        self.q = bf16 if bf16_inputs else (lambda x: np.asarray(x, np.float64))

    def _w(self, k):
        o = self.off[k]
        i, h, u = self.dims[k]
        p = self.p
        return (p[o[0]:o[0] + i * h].reshape(i, h), p[o[1]:o[1] + h],
                p[o[2]:o[2] + h * h].reshape(h, h), p[o[3]:o[3] + h],
                p[o[4]:o[4] + h * u].reshape(h, u), p[o[5]:o[5] + u])

    # one MLP: returns output and the cache for the backward pass
    def _mlp(self, k, x, resid=None):
        q = self.q
        W1, b1, W2, b2, W3, b3 = self._w(k)
        xq = q(x)
        pre1 = xq @ q(W1) + b1
        a1 = q(silu(pre1))
        pre2 = a1 @ q(W2) + b2
        a2 = q(silu(pre2))
        y = a2 @ q(W3) + b3
        if resid is not None:
            y = y + resid
        return y, (k, xq, pre1, a1, pre2, a2)

    def _mlp_bwd(self, cache, dy, grad, want_dx=True):
        q = self.q
        k, xq, pre1, a1, pre2, a2 = cache
        W1, b1, W2, b2, W3, b3 = self._w(k)
        o = self.off[k]
        i, h, u = self.dims[k]
        dyq = q(dy)
        grad[o[4]:o[4] + h * u] += (a2.T @ dyq).ravel()
        grad[o[5]:o[5] + u] += dy.sum(0)
        dpre2 = (dyq @ q(W3).T) * silu_grad(pre2)
        dq2 = q(dpre2)
        grad[o[2]:o[2] + h * h] += (a1.T @ dq2).ravel()
        grad[o[3]:o[3] + h] += dpre2.sum(0)
        dpre1 = (dq2 @ q(W2).T) * silu_grad(pre1)
        dq1 = q(dpre1)
        grad[o[0]:o[0] + i * h] += (xq.T @ dq1).ravel()
        grad[o[1]:o[1] + h] += dpre1.sum(0)
        return dq1 @ q(W1).T if want_dx else None

    def forward(self, node_feat, src, dst, edge_feat):
        n = node_feat.shape[0]
        fn = np.zeros((n, PAD))
        fn[:, :node_feat.shape[1]] = node_feat
        fe = np.zeros((len(src), PAD))
        fe[:, :edge_feat.shape[1]] = edge_feat
        caches = []
        h, c = self._mlp(0, fn)
        caches.append(c)
        e, c = self._mlp(1, fe)
        caches.append(c)
        for l in range(self.L):
            xe = np.concatenate([e, h[src], h[dst]], 1)
            e, c = self._mlp(2 + 2 * l, xe, resid=e)
            caches.append(c)
            agg = np.zeros_like(h)
            np.add.at(agg, dst, e)
            xn = np.concatenate([h, agg], 1)
            h, c = self._mlp(3 + 2 * l, xn, resid=h)
            caches.append(c)
        y, c = self._mlp(2 + 2 * self.L, h)
        caches.append(c)
        self._cache = (caches, src, dst, n, len(src))
        return y[:, :self.out]

    def backward(self, dout):
        caches, src, dst, n, ne = self._cache
        grad = np.zeros_like(self.p)
        H = self.dims[0][1]
        dy = np.zeros((n, PAD))
        dy[:, :self.out] = dout
        dh = self._mlp_bwd(caches[-1], dy, grad)
        de = np.zeros((ne, H))
        for l in range(self.L - 1, -1, -1):
            dxn = self._mlp_bwd(caches[3 + 2 * l], dh, grad)
            dh = dh + dxn[:, :H]
            dagg = dxn[:, H:]
            de = de + dagg[dst]
            dxe = self._mlp_bwd(caches[2 + 2 * l], de, grad)
            de = de + dxe[:, :H]
            np.add.at(dh, src, dxe[:, H:2 * H])
            np.add.at(dh, dst, dxe[:, 2 * H:])
        self._mlp_bwd(caches[1], de, grad, want_dx=False)
        self._mlp_bwd(caches[0], dh, grad, want_dx=False)
        return grad


def layout(hidden: int, layers: int):
    """The parameter layout the library uses (64-element aligned blocks),
    restated for tests that build a network without the library."""
    def a64(x):
        return (x + 63) // 64 * 64
    offs, dims, off = [], [], 0
    shapes = [(PAD, hidden, hidden), (PAD, hidden, hidden)]
    for _ in range(layers):
        shapes += [(3 * hidden, hidden, hidden), (2 * hidden, hidden, hidden)]
    shapes.append((hidden, hidden, PAD))
    for i, h, u in shapes:
        o = []
        for size in (i * h, h, h * h, h, h * u, u):
            o.append(off)
            off = a64(off + size)
        offs.append(o)
        dims.append((i, h, u))
    return np.array(offs), np.array(dims), off
