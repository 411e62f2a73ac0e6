"""Overset donor weights (computeNodalWeight2D, mathTools/math.cpp:229-312)
without a GPU: the oracle against the reference's own known answers
(mul_test.cpp:200-263: vertices get unit weights, interior points are
reproduced, the weight-sign inclusion test agrees with a direct one) and
against linear-field reproduction on jittered quadrilaterals listed in any
vertex order; the product's host routine against the oracle bit for bit."""
import numpy as np

from oracle import oracle as O
from tmp_amd import usg

TRI = np.array([[0.0, 1.0], [1.0, 1.0], [1.0, 0.0]])   # mul_test.cpp barycentric_test triangle


def _inside_tri(p, t):
    def s(a, b, c):
        return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
    d = [s(t[0], t[1], p), s(t[1], t[2], p), s(t[2], t[0], p)]
    return (min(d) >= 0) or (max(d) <= 0)


def test_triangle_known_answers():
    for i in range(3):
        w = O.nodal_weights(3, TRI, TRI[i])
        assert np.allclose(w, np.eye(3)[i], atol=1e-15)
    rng = np.random.default_rng(0)
    for p in rng.uniform(-4, 4, size=(500, 2)):       # mul_test draws from U(-4, 4)
        w = O.nodal_weights(3, TRI, p)
        assert abs(w.sum() - 1) <= 1e-14
        assert np.allclose(w @ TRI, p, atol=1e-12)
        inner = bool(np.all(w >= 0) and np.all(w <= 1))
        assert inner == _inside_tri(p, TRI)


def _quad(rng):
    base = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    return base + rng.uniform(-0.2, 0.2, size=(4, 2))


def test_quadrilateral_reproduces_points_and_linear_fields():
    rng = np.random.default_rng(1)
    orders = ([0, 1, 2, 3], [0, 2, 1, 3], [0, 1, 3, 2], [0, 3, 2, 1], [1, 3, 0, 2])
    for _ in range(200):
        q = _quad(rng)
        u, v = rng.uniform(0.02, 0.98, size=2)
        p = (1 - u) * (1 - v) * q[0] + u * (1 - v) * q[1] + u * v * q[2] + (1 - u) * v * q[3]
        for o in orders:       # vertex lists in any order (findClockWise)
            xv = q[o]
            w = O.nodal_weights(4, xv, p)
            assert abs(w.sum() - 1) <= 1e-12
            assert np.all(w >= -1e-12) and np.all(w <= 1 + 1e-12)
            assert np.allclose(w @ xv, p, atol=1e-12)
            a, b, c = rng.normal(size=3)
            assert abs(w @ (a + b * xv[:, 0] + c * xv[:, 1]) - (a + b * p[0] + c * p[1])) <= 1e-12
        for i in range(4):
            assert np.allclose(O.nodal_weights(4, q, q[i]), np.eye(4)[i], atol=1e-12)


def test_degenerate_triangle_is_an_error():
    assert O.nodal_weights(3, [[0, 0], [1, 1], [2, 2]], [0.5, 0.5]) is None
    import pytest
    with pytest.raises(usg.UsgError):
        usg.nodal_weights([3], [[[0, 0], [1, 1], [2, 2], [0, 0]]], [[0.5, 0.5]])


def test_product_weights_equal_oracle_bitwise():
    rng = np.random.default_rng(2)
    n = 400
    nv = rng.choice([3, 4], size=n)
    xv = np.zeros((n, 4, 2))
    xp = np.zeros((n, 2))
    for r in range(n):
        q = _quad(rng)
        xv[r] = q[rng.permutation(4)] if nv[r] == 4 else np.vstack([q[:3], [0, 0]])
        xp[r] = q[:nv[r]].mean(axis=0) + rng.uniform(-0.3, 0.3, size=2)
    w = usg.nodal_weights(nv, xv, xp)
    for r in range(n):
        wo = O.nodal_weights(nv[r], xv[r, :nv[r]], xp[r])
        assert np.array_equal(w[r, :nv[r]], wo), r


def test_oracle_apply_hand_example():
    src = np.array([[1.0, 10.0], [2.0, 20.0], [4.0, 40.0]])
    dst = np.zeros((2, 2))
    O.stencil_apply(src, dst, recv=[1], nw=[3], donor=[0, 1, 2, 0], w=[0.5, 0.25, 0.25, 0.0])
    assert dst.tolist() == [[0.0, 0.0], [2.0, 20.0]]   # 0.5*1 + 0.25*2 + 0.25*4
