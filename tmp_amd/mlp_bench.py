"""Throughput of the warm-start MLP kernels (BF16 MFMA): one training epoch
and one inference pass of the 8 x 256 network on synthetic points.

    python -m tmp_amd.mlp_bench [--points N] [--batch B]

Prints one JSON line: points/s and TFLOP/s (algorithmic: forward 2*sum of
fan_in*fan_out per point, training 3x that) for training and inference.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import mlp
from . import warmstart as W


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--batch", type=int, default=16384)
    a = ap.parse_args()
    rng = np.random.default_rng(0)
    n = a.points
    X = rng.uniform(-1, 1, size=(n, W.NIN)).astype(np.float32)
    Y = np.stack([np.sin(3 * X[:, 0]), X[:, 1] * X[:, 2], np.cos(X[:, 3]), X[:, 4]], 1).astype(np.float32)
    net = mlp.MLP(W.NIN, 256, 8, 4)
    net.set_params(*W._init(W.NIN, 256, 8, 4, 5006))
    net.set_data(X, Y)
    fwd_flop = 2 * (W.NIN * 256 + 6 * 256 * 256 + 256 * 4)
    net.train_epoch(rng.permutation(n), a.batch, 1e-4)
    t0 = time.perf_counter()
    loss = net.train_epoch(rng.permutation(n), a.batch, 1e-4)
    tt = time.perf_counter() - t0
    net.predict(X[:65536])
    t0 = time.perf_counter()
    net.predict(X)
    tp = time.perf_counter() - t0
    print(json.dumps({"points": n, "batch": a.batch, "train_points_per_s": n / tt,
                      "train_tflops": 3 * fwd_flop * n / tt / 1e12, "predict_points_per_s": n / tp,
                      "predict_tflops_incl_host_copies": fwd_flop * n / tp / 1e12, "loss": loss}), flush=True)


if __name__ == "__main__":
    main()
