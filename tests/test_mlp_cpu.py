"""The MLP oracle (oracle/mlp_oracle.py) pinned without a GPU: BF16 rounding
against known bit patterns, exact-mode gradients against central finite
differences, and the library's C ABI exports."""
import numpy as np

from oracle import mlp_oracle as MO
from tmp_amd import mlp, usg


def test_bf16_rounding_known_values():
    e = 2.0 ** -7
    vals = np.array([1.0, 1 + e / 2, 1 + 1.5 * e, 1 + 0.6 * e, -2.0 - e / 2, 3.0e-3], np.float32)
    r = MO.to_bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # tie -> even mantissa
    assert r[2] == np.float32(1 + 2 * e)   # tie -> even (k = 2)
    assert r[3] == np.float32(1 + e)       # nearest
    assert r[4] == -2.0                    # tie -> even, sign kept
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8
    # bf16 values are exactly representable: rounding again changes nothing
    assert np.array_equal(MO.to_bf16(r), r)


def _fd_check(nin, width, nlayer, nout, seed):
    rng = np.random.default_rng(seed)
    w, b = MO.init_params(nin, width, nlayer, nout, seed)
    b = rng.normal(0, 0.1, b.size).astype(np.float64)
    w = w.astype(np.float64)
    x = rng.normal(size=(7, nin))
    y = rng.normal(size=(7, nout))
    net = MO.MLP(nin, width, nlayer, nout, w, b, bf16=False)
    _, gW, gb = net.loss_grads(x, y)
    gw_flat, gb_flat = net.flat(gW, gb)
    h = 1e-6
    for vec, g, is_w in ((w, gw_flat, True), (b, gb_flat, False)):
        idx = rng.choice(vec.size, size=min(25, vec.size), replace=False)
        for i in idx:
            vp, vm = vec.copy(), vec.copy()
            vp[i] += h
            vm[i] -= h
            lp = MO.MLP(nin, width, nlayer, nout, vp if is_w else w, b if is_w else vp, bf16=False).loss_grads(x, y)[0]
            lm = MO.MLP(nin, width, nlayer, nout, vm if is_w else w, b if is_w else vm, bf16=False).loss_grads(x, y)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (is_w, i, fd, g[i])


def test_exact_gradients_match_finite_differences():
    _fd_check(3, 8, 3, 2, seed=1)
    _fd_check(5, 6, 4, 3, seed=2)


def test_bf16_mode_close_to_exact():
    nin, width, nlayer, nout = 12, 64, 4, 4
    w, b = MO.init_params(nin, width, nlayer, nout, 5006)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(64, nin))
    y = rng.normal(size=(64, nout))
    a = MO.MLP(nin, width, nlayer, nout, w, b, bf16=True).loss_grads(x, y)
    e = MO.MLP(nin, width, nlayer, nout, w, b, bf16=False).loss_grads(x, y)
    assert abs(a[0] - e[0]) <= 2e-2 * e[0]
    for ga, ge in zip(a[1], e[1]):
        assert np.linalg.norm(ga - ge) <= 5e-2 * np.linalg.norm(ge)


def test_mlp_exports_and_shape_checks():
    L = usg.lib()
    for n in mlp.EXPORTS:
        assert hasattr(L, n), n
    assert mlp.sizes(72, 256, 8, 4) == (72 * 256 + 6 * 256 * 256 + 256 * 4, 7 * 256 + 4)
    for bad in ((200, 256, 8, 4), (72, 100, 8, 4), (72, 256, 1, 4), (72, 256, 8, 17)):
        try:
            mlp.sizes(*bad)
        except mlp.MlpError:
            continue
        raise AssertionError(bad)


def test_warm_start_screen():
    """The screen of a predicted field (tmp_amd.warmstart):
    finite, positive density and pressure everywhere."""
    from tmp_amd import warmstart
    U = np.tile([1.0, 1.0, 0.3, 0.0], (10, 1))
    assert warmstart.screen_prediction(U)
    for j, v in ((0, -1e-3), (1, 0.0), (2, np.nan), (3, np.inf)):
        bad = U.copy()
        bad[4, j] = v
        assert not warmstart.screen_prediction(bad)
