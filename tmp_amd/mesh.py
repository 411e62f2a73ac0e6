"""Unstructured 2-D meshes: Cobalt ``.grd`` I/O, seeded generators and the
colour-blocked Hilbert renumbering used by the GPU solver.

Conventions kept from the reference reader (reader/UnstructReader.h:124-285):

* header ``ndim nzone npatch`` then ``nnode nface ncell maxNodesPerFace
  maxFacesPerCell``; node coordinates; one line per face ``k n1 .. nk lC rC``
  with 1-based node and cell ids and a negative patch tag for boundary faces;
* the face normal ``(dy, -dx)/|d|`` of the segment n1 -> n2 points out of the
  left cell (toolBox/edge3d_int.h:200-216), so the left cell lies to the left
  of n1 -> n2;
* boundary types are the reference's enum (toolBox/edge3d_int.h:33-40):
  ``WALL = -2``, ``FARFIELD = -3``; the file's wall / far-field tags are
  mapped to them by the caller's ``wall_tag`` / ``far_tag`` (the reference's
  ``setWallSign`` / ``setOverSign``).

Everything here is host-side preparation.  The numbering chosen by
:func:`colour_hilbert_order` decides the LU-SGS sweep order, so the CPU
oracle and the GPU solver are always run on the same renumbered mesh.
"""
from __future__ import annotations

import dataclasses
import gzip
import heapq
import math
import os

import numpy as np

WALL = -2
FARFIELD = -3


@dataclasses.dataclass
class Mesh:
    """Face-based 2-D mesh (0-based ids)."""

    xy: np.ndarray      # (nn, 2) float64 node coordinates
    fnode: np.ndarray   # (nf, 2) int32 face nodes, normal (dy,-dx) out of flc
    flc: np.ndarray     # (nf,) int32 left cell
    frc: np.ndarray     # (nf,) int32 right cell, or WALL / FARFIELD
    nc: int
    name: str = "mesh"

    @property
    def nn(self) -> int:
        return int(self.xy.shape[0])

    @property
    def nf(self) -> int:
        return int(self.fnode.shape[0])

    def copy(self) -> "Mesh":
        return Mesh(self.xy.copy(), self.fnode.copy(), self.flc.copy(), self.frc.copy(), self.nc, self.name)

    def neighbours(self) -> np.ndarray:
        """(nc, 4) neighbour cells in ascending face order, -1 padded."""
        nb = np.full((self.nc, 4), -1, dtype=np.int64)
        cnt = np.zeros(self.nc, dtype=np.int64)
        for f in range(self.nf):
            l, r = int(self.flc[f]), int(self.frc[f])
            if r >= 0:
                nb[l, cnt[l]] = r
                cnt[l] += 1
                nb[r, cnt[r]] = l
                cnt[r] += 1
        return nb


# ---------------------------------------------------------------------------
# Cobalt .grd I/O
# ---------------------------------------------------------------------------

def read_grd(path: str, wall_tag: int, far_tag: int) -> Mesh:
    """Read a 2-D Cobalt ``.grd`` file (reader/UnstructReader.h:124-262).

    ``wall_tag`` and ``far_tag`` are the file's negative patch tags for wall
    and far-field faces; any other negative tag raises ``ValueError`` as the
    reference throws "Not a valid boundary type".  A path ending in ``.gz``
    is read through gzip.
    """
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as fh:
        tok = fh.read().split()
    ndim = int(tok[0])
    if ndim != 2:
        raise ValueError("only 2-D meshes are supported")
    nn, nf, nc = int(tok[3]), int(tok[4]), int(tok[5])
    pos = 8
    xy = np.array(tok[pos:pos + 2 * nn], dtype=np.float64).reshape(nn, 2)
    pos += 2 * nn
    rec = np.array(tok[pos:pos + 5 * nf], dtype=np.int64).reshape(nf, 5)
    if not np.all(rec[:, 0] == 2):
        raise ValueError("2-D faces must have two nodes")
    fnode = (rec[:, 1:3] - 1).astype(np.int32)
    lc = rec[:, 3]
    rc = rec[:, 4]
    if np.any(lc <= 0):
        raise ValueError("Left cell can't be negative")
    frc = np.empty(nf, dtype=np.int32)
    inner = rc > 0
    frc[inner] = rc[inner] - 1
    wall = rc == wall_tag
    far = rc == far_tag
    if np.any(~(inner | wall | far)):
        raise ValueError("Not a valid boundary type")
    frc[wall] = WALL
    frc[far] = FARFIELD
    return Mesh(xy, fnode, (lc - 1).astype(np.int32), frc, nc, os.path.basename(path).removesuffix(".gz"))


def write_grd(path: str, m: Mesh, wall_tag: int = -1, far_tag: int = -2) -> None:
    """Write ``m`` in the Cobalt format read by :func:`read_grd`."""
    rc = m.frc.astype(np.int64) + 1
    rc[m.frc == WALL] = wall_tag
    rc[m.frc == FARFIELD] = far_tag
    npatch = int(np.any(m.frc == WALL)) + int(np.any(m.frc == FARFIELD))
    with open(path, "w") as fh:
        fh.write(f"2 1 {npatch}\n{m.nn} {m.nf} {m.nc} 2 4\n")
        np.savetxt(fh, m.xy, fmt="%.17e")
        rec = np.column_stack([np.full(m.nf, 2), m.fnode + 1, m.flc + 1, rc])
        np.savetxt(fh, rec, fmt="%d")


# ---------------------------------------------------------------------------
# Building a face mesh from cells
# ---------------------------------------------------------------------------

def _signed_area(xy: np.ndarray, cells: np.ndarray, nv: np.ndarray) -> np.ndarray:
    a = np.zeros(cells.shape[0])
    for k in range(4):
        i0 = cells[:, k]
        i1 = np.where(k + 1 < nv, cells[:, (k + 1) % 4], cells[:, 0])
        valid = k < nv
        x0, y0 = xy[i0, 0], xy[i0, 1]
        x1, y1 = xy[i1, 0], xy[i1, 1]
        a += np.where(valid, x0 * y1 - x1 * y0, 0.0)
    return 0.5 * a


def faces_from_cells(xy: np.ndarray, cells: np.ndarray, nv: np.ndarray,
                     bc_of_edge) -> Mesh:
    """Build the face list of a conforming tri/quad mesh.

    ``cells`` is (nc, 4) node ids (triangles padded with -1), counter-clockwise
    or clockwise; ``bc_of_edge(x0, y0, x1, y1)`` returns WALL or FARFIELD for
    the midpoints of boundary edges (vectorised, arrays in, array out).
    """
    nc = cells.shape[0]
    cells = cells.copy()
    area = _signed_area(xy, cells, nv)
    # orient every cell counter-clockwise
    cw = area < 0
    if np.any(cw):
        tri = cw & (nv == 3)
        quad = cw & (nv == 4)
        cells[tri, :3] = cells[tri][:, [0, 2, 1]]
        cells[quad] = cells[quad][:, [0, 3, 2, 1]]
    e0, e1, ec = [], [], []
    for k in range(4):
        valid = nv > k
        a = cells[valid, k]
        b = np.where(k + 1 < nv[valid], cells[valid, (k + 1) % 4], cells[valid, 0])
        e0.append(a)
        e1.append(b)
        ec.append(np.nonzero(valid)[0])
    e0 = np.concatenate(e0)
    e1 = np.concatenate(e1)
    ec = np.concatenate(ec)
    lo = np.minimum(e0, e1)
    hi = np.maximum(e0, e1)
    key = lo.astype(np.int64) * (xy.shape[0] + 1) + hi
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.ones(len(ks), dtype=bool)
    first[1:] = ks[1:] != ks[:-1]
    grp = np.cumsum(first) - 1
    ngrp = grp[-1] + 1
    cnt = np.bincount(grp, minlength=ngrp)
    if np.any(cnt > 2):
        raise ValueError("non-manifold edge")
    # for a counter-clockwise cell, travelling a->b has the cell on the left,
    # so the normal (dy,-dx) of a->b points out of that cell.
    idx_first = order[first]
    fa = e0[idx_first]
    fb = e1[idx_first]
    flc = ec[idx_first]
    frc = np.full(ngrp, -1, dtype=np.int64)
    second = order[~first]
    frc[grp[~first]] = ec[second]
    bnd = frc < 0
    mx = 0.5 * (xy[fa[bnd], 0] + xy[fb[bnd], 0])
    my = 0.5 * (xy[fa[bnd], 1] + xy[fb[bnd], 1])
    frc[bnd] = bc_of_edge(xy[fa[bnd], 0], xy[fa[bnd], 1], xy[fb[bnd], 0], xy[fb[bnd], 1]) if mx.size else frc[bnd]
    fnode = np.column_stack([fa, fb]).astype(np.int32)
    return Mesh(xy.astype(np.float64), fnode, flc.astype(np.int32), frc.astype(np.int32), nc)


# ---------------------------------------------------------------------------
# Renumbering: colour blocks in Hilbert order
# ---------------------------------------------------------------------------

def _hilbert_keys(x: np.ndarray, y: np.ndarray, order: int = 16) -> np.ndarray:
    n = 1 << order
    xmin, xmax = float(x.min()), float(x.max())
    ymin, ymax = float(y.min()), float(y.max())
    span = max(xmax - xmin, ymax - ymin, 1e-300)
    xi = np.clip(((x - xmin) / span * (n - 1)).astype(np.int64), 0, n - 1)
    yi = np.clip(((y - ymin) / span * (n - 1)).astype(np.int64), 0, n - 1)
    d = np.zeros_like(xi)
    s = n >> 1
    while s > 0:
        rx = ((xi & s) > 0).astype(np.int64)
        ry = ((yi & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate
        flip = ry == 0
        swap_x = flip & (rx == 1)
        xi = np.where(swap_x, s - 1 - xi, xi)
        yi = np.where(swap_x, s - 1 - yi, yi)
        xi, yi = np.where(flip, yi, xi), np.where(flip, xi, yi)
        s >>= 1
    return d


def cell_centres(m: Mesh) -> np.ndarray:
    """Mean of face midpoints per cell (only used to sort; not the solver's centre)."""
    fx = 0.5 * (m.xy[m.fnode[:, 0]] + m.xy[m.fnode[:, 1]])
    acc = np.zeros((m.nc, 2))
    cnt = np.zeros(m.nc)
    np.add.at(acc, m.flc, fx)
    np.add.at(cnt, m.flc, 1)
    inner = m.frc >= 0
    np.add.at(acc, m.frc[inner], fx[inner])
    np.add.at(cnt, m.frc[inner], 1)
    return acc / cnt[:, None]


def _dsatur_ball(nbl: list, col: list, ball: list, ncolours: int) -> bool:
    """Colours the (uncoloured) cells of ``ball`` with at most ``ncolours``
    colours by DSatur (Brelaz 1979), the colours outside the ball fixed: the
    next cell is the one with the most distinct colours around it (ties: more
    neighbours), and it takes the smallest free colour.  False, with the ball
    partly coloured, when some cell has no free colour."""
    inb = set(ball)
    sat = {}
    for v in ball:
        m = 0
        for o in nbl[v]:
            if o not in inb and col[o] >= 0:
                m |= 1 << col[o]
        sat[v] = m
    nsat = {v: bin(sat[v]).count("1") for v in ball}
    heap = [(-nsat[v], -len(nbl[v]), v) for v in ball]
    heapq.heapify(heap)
    while heap:
        s, _, v = heapq.heappop(heap)
        if col[v] >= 0 or -s != nsat[v]:
            continue
        k = 0
        while k < ncolours and sat[v] & (1 << k):
            k += 1
        if k == ncolours:
            return False
        col[v] = k
        for o in nbl[v]:
            if o in inb and col[o] < 0 and not sat[o] & (1 << k):
                sat[o] |= 1 << k
                nsat[o] += 1
                heapq.heappush(heap, (-nsat[o], -len(nbl[o]), o))
    return True


def _three_colours(nbl: list, visit: list, max_radius: int = 8) -> list:
    """Proper colouring with colours 0-2 for all but a few cells.

    Greedy over ``visit`` (the smallest colour no coloured neighbour uses)
    leaves the cells whose neighbours already hold all three colours.  Each
    is repaired by recolouring the ball of cells within 1, 2, ...
    ``max_radius`` faces of it with DSatur, everything outside held fixed.
    A 3-colouring of a quadrilateral grid can hold defects no local change
    removes; the cells still left are repaired the same way with a fourth
    colour allowed, which then holds only a few cells.
    """
    col = [-1] * len(nbl)
    left = []
    for v in visit:
        used = 0
        for o in nbl[v]:
            if col[o] >= 0:
                used |= 1 << col[o]
        k = 0
        while k < 3 and used & (1 << k):
            k += 1
        if k == 3:
            left.append(v)
        else:
            col[v] = k
    def repair(v, ncolours):
        for r in range(1, max_radius + 1):
            ball, seen, front = [v], {v}, [v]
            for _ in range(r):
                nxt = []
                for x in front:
                    for o in nbl[x]:
                        if o not in seen:
                            seen.add(o)
                            nxt.append(o)
                ball += nxt
                front = nxt
            saved = [col[x] for x in ball]
            for x in ball:
                col[x] = -1
            if _dsatur_ball(nbl, col, ball, ncolours):
                return True
            for x, c in zip(ball, saved):
                col[x] = c
        return False

    still = [v for v in left if col[v] < 0 and not repair(v, 3)]
    # the rest: the same repair with a fourth colour, which DSatur gives only
    # where colours 0-2 are all taken
    for v in still:
        if col[v] < 0 and not repair(v, 4):
            used = {col[o] for o in nbl[v]}
            col[v] = min(k for k in range(6) if k not in used)
    return col


def colour_hilbert_order(m: Mesh, parts: int = 8) -> tuple[np.ndarray, int]:
    """Permutation ``perm`` (new -> old cell) grouping cells by colour.

    Cells are coloured in Hilbert order of their centres with three colours
    (``_three_colours``: greedy, local repair, a fourth colour only for the
    few cells no local repair fixes).  The new numbering sorts by (colour,
    spatial part, Hilbert key), the parts being ``parts`` equal slices of
    the Hilbert curve.  No two cells of one colour are neighbours, so a
    forward LU-SGS sweep over this numbering can process each colour block
    as one parallel wave (depth = number of colours), while the Hilbert
    order inside each block keeps neighbour gathers local.  Few colours mean
    few waves: a wave's launch costs its threads' dependent chain of loads
    and divisions however few cells it holds.  (A plain greedy colouring
    needs five, the last two holding ~1% of the cells; a DSatur colouring of
    the whole mesh gets three, but its near two-colour, red-black ordering
    slows the LU-SGS convergence.)  Returns ``(perm, ncolours)``.
    """
    c = cell_centres(m)
    key = _hilbert_keys(c[:, 0], c[:, 1])
    visit = np.argsort(key, kind="stable")
    nbl = [[o for o in row if o >= 0] for row in m.neighbours().tolist()]
    colour = np.array(_three_colours(nbl, visit.tolist()), dtype=np.int64)
    # 8 equal spatial parts along the Hilbert curve (one per XCD of the GPU):
    # inside a colour block the cells are grouped by part, then Hilbert order
    rank = np.empty(m.nc, dtype=np.int64)
    rank[visit] = np.arange(m.nc)
    part = rank * parts // m.nc
    perm = np.lexsort((key, part, colour))
    return perm, int(colour.max()) + 1


def renumber(m: Mesh, perm: np.ndarray, part=None) -> Mesh:
    """Apply a cell permutation (new -> old) and rebuild the face order.

    Interior faces get ``lC < rC`` (the convention of the shipped meshes,
    flipping node order so the normal still points out of the left cell);
    faces are sorted by (lower cell, upper cell) with boundary faces of a cell
    after its interior faces, so each cell's ascending face list starts with
    the faces it owns; with ``part`` (old cell -> spatial part) the faces are
    first grouped by the part of their left cell.
    """
    inv = np.empty(m.nc, dtype=np.int64)
    inv[perm] = np.arange(m.nc)
    lc = inv[m.flc]
    rc = np.where(m.frc >= 0, inv[np.maximum(m.frc, 0)], m.frc)
    fnode = m.fnode.copy()
    swap = (rc >= 0) & (rc < lc)
    lc2 = np.where(swap, rc, lc)
    rc2 = np.where(swap, lc, rc)
    fnode[swap] = fnode[swap][:, ::-1]
    sort_r = np.where(rc2 >= 0, rc2, m.nc + (-rc2))
    if part is None:
        order = np.lexsort((sort_r, lc2))
    else:
        pl = np.asarray(part)[perm][lc2]
        order = np.lexsort((sort_r, lc2, pl))
    return Mesh(m.xy.copy(), fnode[order].astype(np.int32), lc2[order].astype(np.int32),
                rc2[order].astype(np.int32), m.nc, m.name)


def prepare(m: Mesh, parts: int = 8) -> tuple[Mesh, int]:
    """Colour-blocked, part-grouped Hilbert renumbering; returns (mesh, ncolours)."""
    perm, ncol = colour_hilbert_order(m, parts)
    c = cell_centres(m)
    key = _hilbert_keys(c[:, 0], c[:, 1])
    rank = np.empty(m.nc, dtype=np.int64)
    rank[np.argsort(key, kind="stable")] = np.arange(m.nc)
    part = rank * parts // m.nc
    return renumber(m, perm, part), ncol


# ---------------------------------------------------------------------------
# Generators
# ---------------------------------------------------------------------------

def _split_quads(xy: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Split quads (n,4 ccw) into two triangles along the Delaunay diagonal:
    the diagonal whose opposite angles sum to less than pi."""
    p = xy[q]  # (n,4,2)

    def angle(a, b, c):
        u = a - b
        v = c - b
        cosv = (u * v).sum(-1) / (np.linalg.norm(u, axis=-1) * np.linalg.norm(v, axis=-1))
        return np.arccos(np.clip(cosv, -1, 1))

    # diagonal 0-2: opposite angles at 1 and 3
    s02 = angle(p[:, 0], p[:, 1], p[:, 2]) + angle(p[:, 2], p[:, 3], p[:, 0])
    use02 = s02 <= np.pi
    t1 = np.where(use02[:, None], q[:, [0, 1, 2]], q[:, [0, 1, 3]])
    t2 = np.where(use02[:, None], q[:, [0, 2, 3]], q[:, [1, 2, 3]])
    return np.concatenate([t1, t2])


def _assemble(xy, quads, tris):
    cells = np.full((len(quads) + len(tris), 4), -1, dtype=np.int64)
    nv = np.empty(len(cells), dtype=np.int64)
    cells[:len(quads)] = quads
    nv[:len(quads)] = 4
    cells[len(quads):, :3] = tris
    nv[len(quads):] = 3
    return cells, nv


def channel_bump(ni: int = 160, nj: int = 70, nq: int = 8, seed: int = 1001,
                 jitter: float = 0.1, height: float = 0.1) -> Mesh:
    """Channel x in [-1,2], y in [0,1] with a 10 %-height circular-arc bump of
    chord 1 on y=0 (x in [0,1]).  Quads in the ``nq`` layers next to each
    wall, Delaunay-split triangles elsewhere; interior nodes jittered by
    U(-jitter, jitter) x local spacing (seeded).  Walls at y=0 and y=1,
    far field at x=-1 and x=2.  Default size ~19.8k cells."""
    rng = np.random.default_rng(seed)
    h = height
    R = (0.25 + h * h) / (2 * h) if h > 0 else 1.0
    xs = np.linspace(-1.0, 2.0, ni + 1)
    eta = np.linspace(0.0, 1.0, nj + 1)
    # cluster toward both walls
    eta = 0.5 * (1 - np.cos(np.pi * eta)) * 0.6 + 0.4 * eta
    ylo = np.where((xs > 0) & (xs < 1) & (h > 0), np.sqrt(np.maximum(R * R - (xs - 0.5) ** 2, 0)) - (R - h), 0.0)
    X = np.repeat(xs[:, None], nj + 1, axis=1)
    Y = ylo[:, None] + (1.0 - ylo[:, None]) * eta[None, :]
    dx = np.gradient(xs)[:, None] * np.ones_like(X)
    dy = np.gradient(Y, axis=1)
    jx = rng.uniform(-jitter, jitter, X.shape) * dx
    jy = rng.uniform(-jitter, jitter, X.shape) * dy
    interior = np.zeros_like(X, dtype=bool)
    interior[1:-1, 1:-1] = True
    X = X + np.where(interior, jx, 0)
    Y = Y + np.where(interior, jy, 0)
    xy = np.column_stack([X.ravel(), Y.ravel()])
    nid = np.arange((ni + 1) * (nj + 1)).reshape(ni + 1, nj + 1)
    q = np.stack([nid[:-1, :-1], nid[1:, :-1], nid[1:, 1:], nid[:-1, 1:]], axis=-1)  # ccw
    jj = np.arange(nj)[None, :].repeat(ni, 0)
    isq = (jj < nq) | (jj >= nj - nq)
    quads = q[isq]
    tris = _split_quads(xy, q[~isq])
    cells, nv = _assemble(xy, quads, tris)

    def bc(x0, y0, x1, y1):
        mx = 0.5 * (x0 + x1)
        return np.where((mx < -1 + 1e-9) | (mx > 2 - 1e-9), FARFIELD, WALL)

    m = faces_from_cells(xy, cells, nv, bc)
    m.name = f"channel_bump_{m.nc}"
    return m


def naca4(t: float, m_camber: float = 0.0, p: float = 0.4, n: int = 256, closed_te: bool = False) -> np.ndarray:
    """Closed contour of a 4-digit airfoil (chord 1), counter-clockwise from
    the upper trailing-edge point.  Thickness y_t = 5t(0.2969 sqrt(x) -
    0.1260x - 0.3516x^2 + 0.2843x^3 - 0.1015x^4): a blunt trailing edge whose
    base closes the contour; with ``closed_te`` the last coefficient is
    -0.1036, the variant of the formula with a sharp trailing edge."""
    beta = np.linspace(0, np.pi, n // 2 + 1)
    x = 0.5 * (1 - np.cos(beta))
    a4 = -0.1036 if closed_te else -0.1015
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 + a4 * x ** 4)
    if m_camber > 0:
        yc = np.where(x < p, m_camber / p ** 2 * (2 * p * x - x ** 2),
                      m_camber / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * x - x ** 2))
        dyc = np.where(x < p, 2 * m_camber / p ** 2 * (p - x), 2 * m_camber / (1 - p) ** 2 * (p - x))
    else:
        yc = np.zeros_like(x)
        dyc = np.zeros_like(x)
    th = np.arctan(dyc)
    xu, yu = x - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = x + yt * np.sin(th), yc - yt * np.cos(th)
    # upper surface TE -> LE, then lower surface LE -> TE: counter-clockwise;
    # the blunt base closes the contour from the lower to the upper TE point.
    cx = np.concatenate([xu[1:][::-1], xl])
    cy = np.concatenate([yu[1:][::-1], yl])
    return np.column_stack([cx, cy])


def airfoil_hybrid(ni: int = 192, t: float = 0.12, camber: float = 0.0, dy1: float = 4.5e-6,
                   rfar: float = 50.0, growth: float = 1.15, seed: int | None = None,
                   jitter: float = 0.0, quad_only: bool = False, blend_r: float = 1.0,
                   closed_te: bool = False) -> Mesh:
    """Hybrid O-mesh around a 4-digit airfoil (chord 1, far field at ``rfar``).

    Layers leave the wall with first spacing ``dy1`` growing by ``growth``
    until the cells become isotropic (layer step equal to the local
    circumferential spacing 2*pi*(0.5+eta)/ni); from there the step follows
    that spacing, so the outer cells have aspect ratio ~1 out to the far field.
    The anisotropic wall layers are quadrilaterals, the isotropic region is
    split into Delaunay triangles (kept as quadrilaterals with
    ``quad_only``).  ``ni`` points around the contour.  Layers are marched
    out to ``blend_r`` chords; beyond, each point is blended linearly from
    its position on that (nearly convex) layer to the far-field circle at the
    same angle, which cannot fold the way long marches from a concave
    (cambered) surface do.
    """
    c = naca4(t, camber, n=ni, closed_te=closed_te)
    c = c[:-1] if np.allclose(c[0], c[-1]) else c
    # resample to exactly ni points around (contour is closed through the TE base)
    seg = np.linalg.norm(np.diff(np.vstack([c, c[:1]]), axis=0), axis=1)
    s = np.concatenate([[0], np.cumsum(seg)])
    tot = s[-1]
    # cluster at LE and TE via cosine in arc length of each half
    u = np.linspace(0, 1, ni, endpoint=False)
    target = tot * (u - np.sin(2 * np.pi * u) / (2 * np.pi) * 0.8)
    cc = np.vstack([c, c[:1]])
    sx = np.interp(target, s, cc[:, 0])
    sy = np.interp(target, s, cc[:, 1])
    surf = np.column_stack([sx, sy])
    centre = np.array([0.5, 0.0])
    eta = [0.0]
    step = dy1
    nq = None
    while eta[-1] < rfar:
        iso = 2 * np.pi * (0.5 + eta[-1]) / ni
        if step < iso:
            d = step
            step *= growth
        else:
            if nq is None:
                nq = len(eta) - 1
            d = iso
        eta.append(eta[-1] + d)
    eta = np.array(eta)
    eta *= rfar / eta[-1]
    nj = len(eta) - 1
    if nq is None or quad_only:
        nq = nj
    nt = nj - nq
    # marching with smoothed normals: each layer steps along its own normals,
    # averaged with neighbours (more passes further out) so that the lines
    # fan out smoothly behind the trailing edge and the far layers approach
    # circles; a light tangential Laplacian evens the spacing far away.
    pts = np.empty((ni, nj + 1, 2))
    pts[:, 0] = surf
    cur = surf.copy()
    for j in range(1, nj + 1):
        d = np.roll(cur, -1, 0) - np.roll(cur, 1, 0)
        n = np.column_stack([d[:, 1], -d[:, 0]])
        n /= np.linalg.norm(n, axis=1)[:, None]
        passes = 2 + int(60 * min(1.0, (eta[j - 1] / 0.5)) ** 0.5)
        for _ in range(passes):
            n = n + 0.25 * (np.roll(n, 1, 0) + np.roll(n, -1, 0) - 2 * n)
            n /= np.linalg.norm(n, axis=1)[:, None]
        step = eta[j] - eta[j - 1]
        cur = cur + n * step
        mu = 0.2 * min(1.0, eta[j] / 2.0)
        if mu > 0:
            for _ in range(4):
                lap = np.roll(cur, 1, 0) + np.roll(cur, -1, 0) - 2 * cur
                # tangential part only, so the layer distance is kept
                tng = np.roll(cur, -1, 0) - np.roll(cur, 1, 0)
                tng /= np.linalg.norm(tng, axis=1)[:, None]
                cur = cur + mu * (lap * tng).sum(1)[:, None] * tng
        pts[:, j] = cur
        if eta[j] >= blend_r and j < nj:
            jb = j
            break
    else:
        jb = nj
    if jb < nj:
        base = pts[:, jb]
        cm = base.mean(axis=0)
        th = np.unwrap(np.arctan2(base[:, 1] - cm[1], base[:, 0] - cm[0]))
        if not (np.all(np.diff(th) > 0) or np.all(np.diff(th) < 0)):
            raise ValueError("airfoil_hybrid: blend layer is not star-shaped; raise blend_r")
        r_out = rfar + 0.5
        outer = cm + r_out * np.column_stack([np.cos(th), np.sin(th)])
        for j in range(jb + 1, nj + 1):
            w = (eta[j] - eta[jb]) / (eta[nj] - eta[jb])
            pts[:, j] = (1 - w) * base + w * outer
    if seed is not None and jitter > 0:
        rng = np.random.default_rng(seed)
        dr = np.gradient(eta)[None, :, None]
        jit = rng.uniform(-jitter, jitter, pts.shape) * dr
        jit[:, :nq + 1] = 0
        jit[:, -1] = 0
        pts += jit
    xy = pts.reshape(-1, 2)
    nid = np.arange(ni * (nj + 1)).reshape(ni, nj + 1)
    ip = (np.arange(ni) + 1) % ni
    # i increases counter-clockwise around the contour, j outward: (i,j),(i,j+1),(i+1,j+1),(i+1,j)
    # is counter-clockwise for an outward j and a contour traversed ccw? orientation fixed later.
    q = np.stack([nid[:, :-1], nid[ip, :-1], nid[ip, 1:], nid[:, 1:]], axis=-1)
    quads = q[:, :nq].reshape(-1, 4)
    tq = q[:, nq:].reshape(-1, 4)
    # orient for the Delaunay split (needs ccw)
    area = _signed_area(xy, tq, np.full(len(tq), 4))
    tq = np.where((area < 0)[:, None], tq[:, [0, 3, 2, 1]], tq)
    tris = _split_quads(xy, tq)
    cells, nv = _assemble(xy, quads, tris)
    rwall = np.linalg.norm(surf - centre, axis=1).max() + 1e-9

    def bc(x0, y0, x1, y1):
        mx = 0.5 * (x0 + x1) - centre[0]
        my = 0.5 * (y0 + y1) - centre[1]
        return np.where(np.hypot(mx, my) < rwall + 1.0, WALL, FARFIELD)

    m = faces_from_cells(xy, cells, nv, bc)
    m.name = f"airfoil_hybrid_{m.nc}"
    return m


def airfoil_1m() -> Mesh:
    """~1.0M-cell hybrid mesh around NACA 0012: first spacing 4.5e-6 chords
    (y+ ~ 1 at Re 6e6), quad wall layers, isotropic triangles out to a far
    field at 50 chords, 810 points around the contour."""
    return airfoil_hybrid(ni=810, t=0.12)


def batch_cases(n: int = 64, seed: int = 4004, nc_min: float = 2e4, nc_max: float = 4e5) -> list[dict]:
    """Parameters of a batch of airfoil cases: cell counts log-uniform in
    [nc_min, nc_max] (through the contour resolution ni, ~1.52 ni^2 cells for
    the hybrid y+~1 mesh), 4-digit thickness U(0.06, 0.18), camber U(0, 0.04)
    at p = 0.4, Mach U(0.3, 0.8), angle of attack U(-4, 8) degrees."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nc = float(np.exp(rng.uniform(np.log(nc_min), np.log(nc_max))))
        out.append({"ni": max(32, int(round(np.sqrt(nc / 1.52)))), "t": float(rng.uniform(0.06, 0.18)),
                    "camber": float(rng.uniform(0.0, 0.04)), "mach": float(rng.uniform(0.3, 0.8)),
                    "aoa": float(rng.uniform(-4.0, 8.0))})
    return out


def batch_mesh(case: dict) -> Mesh:
    """The hybrid y+~1 mesh of one batch case (first spacing 4.5e-6 c, far field 50 c)."""
    return airfoil_hybrid(ni=case["ni"], t=case["t"], camber=case["camber"])


def rect(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0, bottom: int = FARFIELD, top: int = FARFIELD,
         sides: int = FARFIELD, jitter: float = 0.0, seed: int = 0, triangles: bool = False) -> Mesh:
    """Rectangle [0,lx]x[0,ly] of nx*ny quads (or Delaunay-split triangles),
    interior nodes jittered by U(-jitter, jitter) x spacing."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(0, lx, nx + 1), np.linspace(0, ly, ny + 1), indexing="ij")
    if jitter > 0:
        jx = rng.uniform(-jitter, jitter, X.shape) * lx / nx
        jy = rng.uniform(-jitter, jitter, X.shape) * ly / ny
        inner = np.zeros_like(X, dtype=bool)
        inner[1:-1, 1:-1] = True
        X = X + np.where(inner, jx, 0)
        Y = Y + np.where(inner, jy, 0)
    xy = np.column_stack([X.ravel(), Y.ravel()])
    nid = np.arange((nx + 1) * (ny + 1)).reshape(nx + 1, ny + 1)
    q = np.stack([nid[:-1, :-1], nid[1:, :-1], nid[1:, 1:], nid[:-1, 1:]], axis=-1).reshape(-1, 4)
    if triangles:
        cells, nv = _assemble(xy, np.zeros((0, 4), np.int64), _split_quads(xy, q))
    else:
        cells, nv = _assemble(xy, q, np.zeros((0, 3), np.int64))
    eps = 1e-9 * max(lx, ly)

    def bc(x0, y0, x1, y1):
        my = 0.5 * (y0 + y1)
        out = np.full(x0.shape, sides)
        out = np.where(my < eps, bottom, out)
        out = np.where(my > ly - eps, top, out)
        return out

    m = faces_from_cells(xy, cells, nv, bc)
    m.name = f"rect_{nx}x{ny}{'_tri' if triangles else ''}"
    return m


# ---------------------------------------------------------------------------
# Overset pair: cylinder O-grid over a Cartesian background
# ---------------------------------------------------------------------------

def background(nx: int = 400, ny: int = 200, x0: float = -10.0, x1: float = 30.0, y0: float = -10.0,
               y1: float = 10.0) -> Mesh:
    """Cartesian background of nx*ny quads on [x0,x1]x[y0,y1], far field all
    round."""
    m = rect(nx, ny, x1 - x0, y1 - y0)
    m.xy = m.xy + np.array([x0, y0])
    m.name = f"background_{nx}x{ny}"
    return m


def ogrid_radii(nj: int, r0: float, dr1: float, ratio: float, dr_max: float) -> np.ndarray:
    """nj+1 radii from r0: spacing dr1 growing by ``ratio`` and capped at dr_max."""
    r = [r0]
    dr = dr1
    for _ in range(nj):
        r.append(r[-1] + dr)
        dr = min(dr * ratio, dr_max)
    return np.array(r)


def cylinder_ogrid(ni: int = 256, nj: int = 96, d: float = 1.0, dr1: float = 1e-3, ratio: float = 1.08,
                   dr_max: float = 0.05) -> tuple[Mesh, np.ndarray]:
    """O-grid of ni*nj quads around a cylinder of diameter d at the origin:
    wall on the cylinder, overset boundary outside.  The radial spacing
    starts at dr1 and grows by ``ratio`` up to dr_max (uncapped, 96 layers at
    1.08 would reach r = 20.7, beyond a background of half-height 10; capped
    at 0.05 they end at r = 3.4).  Returns the mesh and its overset-face
    flags (1 on the outer boundary, where the face is far field)."""
    r = ogrid_radii(nj, 0.5 * d, dr1, ratio, dr_max)
    th = 2 * np.pi * np.arange(ni) / ni
    X = r[None, :] * np.cos(th)[:, None]
    Y = r[None, :] * np.sin(th)[:, None]
    xy = np.column_stack([X.ravel(), Y.ravel()])
    nid = np.arange(ni * (nj + 1)).reshape(ni, nj + 1)
    ip = np.roll(np.arange(ni), -1)
    q = np.stack([nid[:, :-1], nid[ip, :-1], nid[ip, 1:], nid[:, 1:]], axis=-1).reshape(-1, 4)
    cells, nv = _assemble(xy, q, np.zeros((0, 3), np.int64))
    rmid = 0.5 * (r[0] + r[-1])

    def bc(x0, y0, x1, y1):
        rr = np.hypot(0.5 * (x0 + x1), 0.5 * (y0 + y1))
        return np.where(rr < rmid, WALL, FARFIELD)

    m = faces_from_cells(xy, cells, nv, bc)
    m.name = f"cylinder_ogrid_{ni}x{nj}"
    return m, (m.frc == FARFIELD).astype(np.int32)


def cylinder_delaunay(n_cells: int = 1_000_000, d: float = 1.0, r_far: float = 30.0, alpha: float = 0.6,
                      seed: int = 7) -> Mesh:
    """Irregular triangulation around a cylinder of diameter d: points on
    rings whose spacing grows as h(r) = h_w (2r/d)^alpha, each ring with its
    own point count round(2 pi r / h) and a random angular phase, interior
    points jittered by U(-0.2, 0.2) h radially and along the ring, then the
    Delaunay triangulation (scipy) with the triangles inside the cylinder
    removed.  Ring counts that change from ring to ring and the random
    phases give node valences from 4 to 8, unlike the structured and
    split-quad meshes.  h_w is scaled so that the mesh has about n_cells
    triangles.  Wall on the cylinder, far field on the circle r = r_far."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    r0 = 0.5 * d

    def rings(hw):
        rs = [r0]
        while True:
            h = hw * (rs[-1] / r0) ** alpha
            r = rs[-1] + 0.866 * h
            if r >= r_far:
                break
            rs.append(r)
        rs.append(r_far)
        rs = np.array(rs)
        h = hw * (rs / r0) ** alpha
        return rs, h, np.maximum(8, np.round(2 * np.pi * rs / h).astype(np.int64))

    hw = 0.01
    for _ in range(6):           # triangles ~ 2 x points
        _, _, n = rings(hw)
        hw *= np.sqrt(2.0 * n.sum() / n_cells)
    rs, h, n = rings(hw)
    pts = []
    for k, (r, hk, nk) in enumerate(zip(rs, h, n)):
        th = 2 * np.pi * (np.arange(nk) + rng.uniform()) / nk
        rr = np.full(nk, r)
        if 0 < k < len(rs) - 1:
            rr = rr + rng.uniform(-0.2, 0.2, nk) * hk
            th = th + rng.uniform(-0.2, 0.2, nk) * hk / r
        pts.append(np.column_stack([rr * np.cos(th), rr * np.sin(th)]))
    xy = np.concatenate(pts)
    tri = Delaunay(xy).simplices.astype(np.int64)
    c = xy[tri].mean(axis=1)
    tri = tri[np.hypot(c[:, 0], c[:, 1]) > r0]
    cells, nv = _assemble(xy, np.zeros((0, 4), np.int64), tri)
    rmid = 0.5 * (r0 + r_far)

    def bc(x0, y0, x1, y1):
        return np.where(np.hypot(0.5 * (x0 + x1), 0.5 * (y0 + y1)) < rmid, WALL, FARFIELD)

    m = faces_from_cells(xy, cells, nv, bc)
    m.name = f"cylinder_delaunay_{m.nc}"
    return m

