"""Cell-centred overset coupling restated stage by stage in numpy.

TEST INFRASTRUCTURE ONLY: the checker that tests/ compare the product's
coupling (usg_overset_cell_stencil + the receptor kernel) against; the
product never imports it.

The three stages the reference strings together around tioga::dataUpdate,
each applied in turn to the current cell states of every block:

1. cell -> node: each node of a block takes the mean of the field cells
   (cell iblank 1) around it, the averaging of Euler2D::outPutCp
   (UnstructSolver/Euler2D.cpp:396-422) restricted to field cells;
2. node -> node: each receptor node takes sum_m w_m q[donor node m] over its
   donor cell's vertices (MeshBlock::getInterpolatedSolution,
   UniTioga/dataUpdate.cpp:6-113), its donor being the first one the
   connectivity lists;
3. node -> cell: each receptor cell (cell iblank -1) takes the mean of its
   vertices (Orchestra.cpp:269-292).

A receptor cell with a vertex that has no donor, or with a donor node that
touches no field cell, stays a field cell.
"""
from __future__ import annotations

import numpy as np


class OversetCoupling:
    """Stages 1-3 over blocks given as OracleMesh objects (their cell
    vertices in the reference's order) and the connectivity of
    tmp_amd.usg.overset_coupling / overset_connect (per block: node and cell
    iblank, and the receptors whose donor cell lies in that block)."""

    def __init__(self, omeshes, conn):
        self.nb = len(omeshes)
        self.cnode = [om.cnode.copy() for om in omeshes]
        self.cnv = [om.cnv.copy() for om in omeshes]
        self.nn = [int(om.cnode.max()) + 1 for om in omeshes]
        cib = [np.asarray(c["cell_iblank"]).copy() for c in conn]
        self.field = [c == 1 for c in cib]
        # node stencils of receptor nodes: (donor block, nodes, weights)
        self.stencil = [dict() for _ in range(self.nb)]
        for a, c in enumerate(conn):
            for i in range(len(c["recv_node"])):
                b, v = int(c["recv_block"][i]), int(c["recv_node"][i])
                if v in self.stencil[b]:
                    continue
                nw = int(c["nw"][i])
                self.stencil[b][v] = (a, c["donor_node"][i, :nw].astype(np.int64), c["w"][i, :nw].copy())
        # field cells around each node
        self.nfield = []
        for b in range(self.nb):
            cnt = np.zeros(self.nn[b], np.int64)
            for c in np.nonzero(self.field[b])[0]:
                for v in self.cnode[b][c, :self.cnv[b][c]]:
                    cnt[v] += 1
            self.nfield.append(cnt)
        # receptor cells that can be interpolated; the others stay field cells
        self.blank = []
        self.recv = []
        for b in range(self.nb):
            bl = cib[b].copy()
            rc = []
            for c in np.nonzero(cib[b] == -1)[0]:
                ok = True
                for v in self.cnode[b][c, :self.cnv[b][c]]:
                    st = self.stencil[b].get(int(v))
                    if st is None or len(st[1]) == 0 or np.any(self.nfield[st[0]][st[1]] == 0):
                        ok = False
                        break
                if ok:
                    rc.append(c)
                else:
                    bl[c] = 1
            self.blank.append(bl)
            self.recv.append(np.array(rc, np.int64))

    def node_means(self, U):
        """Stage 1: per block, the mean of the field cells around each node."""
        out = []
        for b in range(self.nb):
            s = np.zeros((self.nn[b], U[b].shape[1]))
            cells = np.nonzero(self.field[b])[0]
            for k in range(4):
                sel = cells[self.cnv[b][cells] > k]
                np.add.at(s, self.cnode[b][sel, k], U[b][sel])
            with np.errstate(invalid="ignore", divide="ignore"):
                out.append(s / self.nfield[b][:, None])
        return out

    def receptor_states(self, U):
        """Stages 2-3: the receptor cells' states of every block from the
        current states U[b][nc_b, nvar]: a list of (cells, states)."""
        node = self.node_means(U)
        out = []
        for b in range(self.nb):
            vals = np.zeros((len(self.recv[b]), U[b].shape[1]))
            for k, c in enumerate(self.recv[b]):
                acc = np.zeros(U[b].shape[1])
                vs = self.cnode[b][c, :self.cnv[b][c]]
                for v in vs:
                    a, dn, w = self.stencil[b][int(v)]
                    q = np.zeros(U[b].shape[1])
                    for m in range(len(dn)):
                        q = q + w[m] * node[a][dn[m]]
                    acc = acc + q
                vals[k] = acc / len(vs)
            out.append((self.recv[b], vals))
        return out

    def apply(self, solvers):
        """Set the receptor cells of oracle solvers (one per block)."""
        res = self.receptor_states([s.U.copy() for s in solvers])
        for s, (idx, vals) in zip(solvers, res):
            if len(idx):
                s.set_cells(idx, vals)
