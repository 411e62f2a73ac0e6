"""Warm-start MLP kernels (BF16 MFMA, gfx950) against the CPU oracle
(oracle/mlp_oracle.py, itself pinned by finite differences).

Bars: the single MFMA layout check to FP32 rounding (1e-5 relative); the
network's loss within 1e-3 and each layer's weight / bias gradient within
2e-2 (relative L2) of the oracle, which rounds the same operands to BF16 but
accumulates in float64 -- FP32 accumulation order can flip single BF16
roundings of intermediate activations; predictions within 1e-2 of the
largest output.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from tmp_amd import mlp

pytestmark = pytest.mark.gpu

NIN, WIDTH, NLAYER, NOUT = 72, 256, 8, 4


def test_mfma_operand_layout():
    rng = np.random.default_rng(11)
    a = rng.normal(size=(16, 32)).astype(np.float32)
    b = rng.normal(size=(32, 16)).astype(np.float32)
    d = mlp.mfma_check(a, b)
    ref = MO.to_bf16(a).astype(np.float64) @ MO.to_bf16(b).astype(np.float64)
    err = np.max(np.abs(d - ref)) / np.max(np.abs(ref))
    print(f"mfma layout rel err {err:.2e}")
    assert err <= 1e-5


def _net(seed=5006):
    w, b = MO.init_params(NIN, WIDTH, NLAYER, NOUT, seed)
    rng = np.random.default_rng(seed + 1)
    b = (b + rng.normal(0, 0.05, b.size)).astype(np.float32)
    g = mlp.MLP(NIN, WIDTH, NLAYER, NOUT)
    g.set_params(w, b)
    return g, w, b


def test_gradients_match_oracle():
    g, w, b = _net()
    rng = np.random.default_rng(3)
    x = rng.normal(size=(700, NIN)).astype(np.float32)      # not a multiple of 64: padding rows
    y = rng.normal(size=(700, NOUT)).astype(np.float32)
    loss, dw, db = g.gradients(x, y)
    o = MO.MLP(NIN, WIDTH, NLAYER, NOUT, w, b, bf16=True)
    lo, gW, gb = o.loss_grads(x.astype(np.float64), y.astype(np.float64))
    print(f"loss gpu {loss:.6e} oracle {lo:.6e}")
    assert abs(loss - lo) <= 1e-3 * lo
    ow = ob = 0
    for l, (k, n) in enumerate(MO.layer_sizes(NIN, WIDTH, NLAYER, NOUT)):
        gw_l = dw[ow:ow + k * n].reshape(k, n)
        gb_l = db[ob:ob + n]
        ew = np.linalg.norm(gw_l - gW[l]) / np.linalg.norm(gW[l])
        eb = np.linalg.norm(gb_l - gb[l]) / np.linalg.norm(gb[l])
        print(f"layer {l}: dW rel {ew:.2e}  db rel {eb:.2e}")
        assert ew <= 2e-2 and eb <= 2e-2, l
        ow += k * n
        ob += n


def test_predict_matches_oracle():
    g, w, b = _net(7)
    x = np.random.default_rng(4).normal(size=(1000, NIN)).astype(np.float32)
    y = g.predict(x)
    ref, _ = MO.MLP(NIN, WIDTH, NLAYER, NOUT, w, b, bf16=True).forward(x.astype(np.float64))
    err = np.max(np.abs(y - ref), axis=0) / np.max(np.abs(ref), axis=0)
    print(f"predict rel err per output {err}")
    assert np.all(err <= 1e-2)


def test_training_is_reproducible_and_learns():
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, size=(8192, NIN)).astype(np.float32)
    y = np.stack([np.sin(2 * x[:, 0]) + np.cos(x[:, 1]), x[:, 2] * x[:, 3], np.tanh(x[:, 4] - x[:, 5]),
                  x[:, 6] ** 2], axis=1).astype(np.float32)
    runs = []
    for _ in range(2):
        g, w, b = _net(5006)
        g.set_data(x, y)
        losses = []
        prng = np.random.default_rng(5005)
        for ep in range(25):
            losses.append(g.train_epoch(prng.permutation(len(x)), 1024, 1e-3))
        runs.append((losses, g.get_params()))
    losses = runs[0][0]
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] <= 0.1 * losses[0]
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1][0], runs[1][1][0]) and np.array_equal(runs[0][1][1], runs[1][1][1])
