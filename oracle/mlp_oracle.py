"""CPU restatement of the warm-start MLP (test infrastructure only).

The network and its arithmetic are the library's (include/usg_mlp.h): the
reference has no learned component, so this oracle is pinned by the
mathematics instead -- its exact mode (float64, no BF16 rounding) against
central finite differences, and its rounding helper against known BF16 bit
patterns (tests/test_mlp_cpu.py).  With ``bf16=True`` it rounds the same
operands to BF16 (round to nearest even) as the GPU does, accumulating in
float64 where the GPU accumulates in FP32.

Parameters use the library layout: per layer W[k][n] (fan-in major) and b[n].
"""
from __future__ import annotations

import numpy as np


def to_bf16(a) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def split_params(sizes, w, b):
    """Flat parameter vectors -> per-layer (W[k][n], b[n]) for layer sizes [(k, n), ...]."""
    Ws, bs, ow, ob = [], [], 0, 0
    for k, n in sizes:
        Ws.append(np.asarray(w[ow:ow + k * n], np.float64).reshape(k, n))
        bs.append(np.asarray(b[ob:ob + n], np.float64))
        ow += k * n
        ob += n
    return Ws, bs


def layer_sizes(nin, width, nlayer, nout):
    return [(nin if l == 0 else width, nout if l == nlayer - 1 else width) for l in range(nlayer)]


def init_params(nin, width, nlayer, nout, seed):
    """Uniform +-sqrt(6/fan_in) weights (+-sqrt(3/fan_in) on the linear output
    layer), zero biases."""
    rng = np.random.default_rng(seed)
    w, b = [], []
    for l, (k, n) in enumerate(layer_sizes(nin, width, nlayer, nout)):
        lim = np.sqrt((3.0 if l == nlayer - 1 else 6.0) / k)
        w.append(rng.uniform(-lim, lim, size=k * n))
        b.append(np.zeros(n))
    return np.concatenate(w).astype(np.float32), np.concatenate(b).astype(np.float32)


class MLP:
    def __init__(self, nin, width, nlayer, nout, w, b, bf16=True):
        self.sizes = layer_sizes(nin, width, nlayer, nout)
        self.Ws, self.bs = split_params(self.sizes, w, b)
        self.bf16 = bf16

    def _r(self, a):
        """Operand rounding of the GPU: float32 then BF16 (or nothing, exact mode)."""
        if not self.bf16:
            return np.asarray(a, np.float64)
        return to_bf16(np.asarray(a, np.float32)).astype(np.float64)

    def _acc(self, a):
        """Accumulator precision: FP32 (GPU) or float64 (exact mode)."""
        return np.asarray(a, np.float32).astype(np.float64) if self.bf16 else a

    def forward(self, x):
        """Returns (out, cache): cache = per layer (input operand H, pre-activation Z)."""
        H = self._r(x)
        cache = []
        L = len(self.Ws)
        for l in range(L):
            Z = self._acc(H @ self._r(self.Ws[l]))
            Z = self._acc(Z + self.bs[l])
            cache.append((H, Z))
            if l < L - 1:
                H = self._r(self._acc(Z * sigmoid(Z)))
        return Z, cache

    def loss_grads(self, x, y):
        out, cache = self.forward(x)
        nv, nout = y.shape
        e = out - y
        loss = float(np.mean(e * e))
        dZ = self._r(self._acc(2.0 * e / (nv * nout)))
        gW, gb = [None] * len(self.Ws), [None] * len(self.Ws)
        for l in range(len(self.Ws) - 1, -1, -1):
            H, _ = cache[l]
            gW[l] = H.T @ dZ
            gb[l] = dZ.sum(axis=0)
            if l == 0:
                break
            dH = self._acc(dZ @ self._r(self.Ws[l]).T)
            Zp = cache[l - 1][1]
            s = sigmoid(Zp)
            dZ = self._r(self._acc(dH * (s * (1.0 + Zp * (1.0 - s)))))
        return loss, gW, gb

    def flat(self, gW, gb):
        return (np.concatenate([g.reshape(-1) for g in gW]), np.concatenate(gb))
