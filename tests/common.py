"""Shared fixtures of the test suite: meshes, perturbed states, comparisons."""
from __future__ import annotations

import functools
import os

import numpy as np

from tmp_amd import mesh

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@functools.lru_cache(maxsize=None)
def part2():
    """Shipped overset component fixture (2222 quads): wall tag -2, overset
    tag -2147483647 treated as far field (Orchestra overTag)."""
    return mesh.read_grd(os.path.join(GOLDEN, "part2.grd"), -2, -2147483647)


@functools.lru_cache(maxsize=None)
def bump():
    """Channel with a 10 % circular-arc bump, ~20k mixed cells, seed 1001,
    colour-blocked Hilbert numbering."""
    return mesh.prepare(mesh.channel_bump(seed=1001))[0]


@functools.lru_cache(maxsize=None)
def airfoil_small():
    """Hybrid NACA 0012 O-mesh, ~8k cells, colour-blocked numbering."""
    return mesh.prepare(mesh.airfoil_hybrid(ni=64, dy1=1e-4, rfar=20.0))[0]


@functools.lru_cache(maxsize=None)
def delaunay_small():
    """Irregular Delaunay triangulation around a cylinder (node valences
    3-8), ~8k cells, colour-blocked numbering."""
    return mesh.prepare(mesh.cylinder_delaunay(8000, r_far=15.0))[0]


def two_cells() -> mesh.Mesh:
    """Quad (0,1,2,3) on the unit square and triangle (1,4,2) to its right.
    Faces in file order: (0,1) far field, (1,2) interior, (2,3) far field,
    (3,0) wall, (1,4) far field, (4,2) wall."""
    xy = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [2, 0.5]], np.float64)
    fnode = np.array([[0, 1], [1, 2], [2, 3], [3, 0], [1, 4], [4, 2]], np.int32)
    flc = np.array([0, 0, 0, 0, 1, 1], np.int32)
    frc = np.array([mesh.FARFIELD, 1, mesh.FARFIELD, mesh.WALL, mesh.FARFIELD, mesh.WALL], np.int32)
    return mesh.Mesh(xy, fnode, flc, frc, 2, "two_cells")


def one_triangle() -> mesh.Mesh:
    """A single triangle: wall face (0,1), far-field faces (1,2), (2,0)."""
    xy = np.array([[0.0, 0.0], [1.0, 0.0], [0.2, 0.9]])
    fnode = np.array([[0, 1], [1, 2], [2, 0]], np.int32)
    frc = np.array([mesh.WALL, mesh.FARFIELD, mesh.FARFIELD], np.int32)
    return mesh.Mesh(xy, fnode, np.zeros(3, np.int32), frc, 1, "one_triangle")


def one_quad() -> mesh.Mesh:
    """A single skewed quadrilateral: wall face (0,1), three far-field faces."""
    xy = np.array([[0.0, 0.0], [1.0, 0.1], [1.1, 1.0], [-0.1, 0.8]])
    fnode = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    frc = np.array([mesh.WALL, mesh.FARFIELD, mesh.FARFIELD, mesh.FARFIELD], np.int32)
    return mesh.Mesh(xy, fnode, np.zeros(4, np.int32), frc, 1, "one_quad")


def perturbed(m, cc, flow_fs, amp=0.05, seed=0):
    """Smooth perturbation of the free stream (so gradients and the limiter
    are exercised), as AoS U[nc,4]."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 2.0, size=(4, 2))
    ph = rng.uniform(0, 2 * np.pi, size=4)
    x, y = cc[:, 0], cc[:, 1]
    U = np.tile(np.asarray(flow_fs, dtype=np.float64), (m.nc, 1))
    for j in range(4):
        scale = 1.0 if j < 2 else max(abs(flow_fs[2]), 0.1)
        U[:, j] = U[:, j] + amp * scale * np.sin(a[j, 0] * x + a[j, 1] * y + ph[j])
    return U


def relerr(a, b):
    """max |a-b| / max |b| per column (whole array for 1-D)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.ndim == 1:
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    return np.max(np.abs(a - b), axis=0) / np.maximum(np.max(np.abs(b), axis=0), 1e-300)


def exact_fraction(a, b):
    return float(np.mean(np.asarray(a) == np.asarray(b)))
