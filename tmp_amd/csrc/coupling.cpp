// Cell-centred overset coupling: the donor stencil of each receptor cell as
// one linear map of donor field cells, built on the host from the
// connectivity (usg_overset_*) and applied on the GPU by usg_set_overset.
//
// The map composes the three stages the reference strings together around
// tioga::dataUpdate:
//   1. cell -> node: a donor mesh's node takes the mean of the cells around
//      it, as Euler2D::outPutCp averages cell values to nodes
//      (Euler2D.cpp:396-422) -- here over field cells only (cell iblank 1),
//      so that holes and receptors never feed a donor value;
//   2. node -> node: a receptor node takes sum_m w_m q[donor node m] of its
//      donor cell's vertices (MeshBlock::getInterpolatedSolution,
//      UniTioga/dataUpdate.cpp:6-113, the weights of findInterpData);
//   3. node -> cell: a receptor cell (every vertex a receptor, cell iblank
//      -1) takes the mean of its vertices (Orchestra.cpp:269-292).
// A receptor cell with a vertex that has no donor, or whose donor nodes
// touch no field cell, cannot be interpolated and is left a field cell.
// Plain C++17, host only.
#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "../../include/usg.h"
#include "geometry.h"

using usg::set_error;

extern "C" int usg_overset_cell_stencil(usg_overset *ov, const usg_mesh_desc *meshes, int nblocks, int *blank,
                                        int *nrecv, int *recv, int *off, int *ndonor, int *donor, double *w) {
  if (!ov || !meshes || nblocks <= 0 || !nrecv || !ndonor) return set_error(USG_ERR_ARG, "null argument");
  std::vector<long> cell0(nblocks + 1, 0);
  for (int b = 0; b < nblocks; b++) cell0[b + 1] = cell0[b] + meshes[b].nc;
  if (cell0[nblocks] > 2147483647L) return set_error(USG_ERR_ARG, "too many cells");
  // per block: cell vertices (reference order), iblanks
  std::vector<std::vector<int>> cnode(nblocks), cnv(nblocks), cib(nblocks), nib(nblocks);
  for (int b = 0; b < nblocks; b++) {
    const usg_mesh_desc &m = meshes[b];
    cnode[b].assign(4 * (size_t)m.nc, -1);
    cnv[b].assign(m.nc, 0);
    cib[b].assign(m.nc, 1);
    nib[b].assign(m.nn, 1);
    if (int rc = usg_mesh_cells(&m, cnode[b].data(), cnv[b].data())) return rc;
    if (int rc = usg_overset_iblank(ov, b, nib[b].data(), cib[b].data())) return rc;
  }
  // stage 1: the field cells around each node of each block (ascending)
  std::vector<std::vector<int>> nstart(nblocks), ncell(nblocks);
  for (int b = 0; b < nblocks; b++) {
    const int nn = meshes[b].nn, nc = meshes[b].nc;
    std::vector<int> cnt(nn + 1, 0);
    for (int c = 0; c < nc; c++)
      if (cib[b][c] == 1)
        for (int v = 0; v < cnv[b][c]; v++) cnt[cnode[b][4 * c + v] + 1]++;
    for (int i = 0; i < nn; i++) cnt[i + 1] += cnt[i];
    ncell[b].assign(cnt[nn], 0);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int c = 0; c < nc; c++)
      if (cib[b][c] == 1)
        for (int v = 0; v < cnv[b][c]; v++) ncell[b][fill[cnode[b][4 * c + v]]++] = c;
    nstart[b] = std::move(cnt);
  }
  // stage 2: the donor stencil of each receptor node: (donor block, nw, nodes, weights)
  struct NodeStencil {
    int block = -1, nw = 0;
    int node[4] = {0, 0, 0, 0};
    double w[4] = {0, 0, 0, 0};
  };
  std::vector<std::vector<NodeStencil>> ns(nblocks);
  for (int b = 0; b < nblocks; b++) ns[b].assign(meshes[b].nn, NodeStencil());
  for (int a = 0; a < nblocks; a++) {
    int n = 0;
    if (int rc = usg_overset_receptors(ov, a, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr)) return rc;
    std::vector<int> rb(n), rn(n), dc(n), nw(n), dn(4 * (size_t)n);
    std::vector<double> ww(4 * (size_t)n);
    if (int rc = usg_overset_receptors(ov, a, &n, rb.data(), rn.data(), dc.data(), nw.data(), dn.data(), ww.data()))
      return rc;
    for (int i = 0; i < n; i++) {
      if (rb[i] < 0 || rb[i] >= nblocks || rn[i] < 0 || rn[i] >= meshes[rb[i]].nn) continue;
      NodeStencil &t = ns[rb[i]][rn[i]];
      if (t.block >= 0) continue;  // one donor per receptor node: the first listed
      t.block = a;
      t.nw = std::min(std::max(nw[i], 0), 4);
      for (int m = 0; m < t.nw; m++) {
        t.node[m] = dn[4 * (size_t)i + m];
        t.w[m] = ww[4 * (size_t)i + m];
      }
    }
  }
  // stage 3 and the composition, receptor cells in union order
  std::vector<int> o_recv, o_off{0}, o_donor;
  std::vector<double> o_w;
  std::vector<int> o_blank(cell0[nblocks], 1);
  std::vector<std::pair<int, double>> terms;
  for (int b = 0; b < nblocks; b++) {
    for (int c = 0; c < meshes[b].nc; c++) {
      const long uc = cell0[b] + c;
      o_blank[uc] = cib[b][c];
      if (cib[b][c] != -1) continue;
      terms.clear();
      const int nv = cnv[b][c];
      bool ok = nv > 0;
      for (int v = 0; v < nv && ok; v++) {
        const NodeStencil &t = ns[b][cnode[b][4 * c + v]];
        if (t.block < 0 || t.nw == 0) {
          ok = false;
          break;
        }
        for (int m = 0; m < t.nw && ok; m++) {
          const int d = t.node[m];
          if (d < 0 || d >= meshes[t.block].nn) {
            ok = false;
            break;
          }
          const int i0 = nstart[t.block][d], i1 = nstart[t.block][d + 1];
          if (i1 == i0) {
            ok = false;
            break;
          }
          const double f = t.w[m] / (double(nv) * double(i1 - i0));
          for (int i = i0; i < i1; i++) terms.push_back({int(cell0[t.block] + ncell[t.block][i]), f});
        }
      }
      if (!ok) {
        o_blank[uc] = 1;  // cannot be interpolated: solved as a field cell
        continue;
      }
      // merge equal donors, ascending donor cell
      std::stable_sort(terms.begin(), terms.end(),
                       [](const std::pair<int, double> &x, const std::pair<int, double> &y) { return x.first < y.first; });
      o_recv.push_back(int(uc));
      for (size_t i = 0; i < terms.size();) {
        size_t j = i;
        double acc = 0.0;
        while (j < terms.size() && terms[j].first == terms[i].first) acc += terms[j++].second;
        o_donor.push_back(terms[i].first);
        o_w.push_back(acc);
        i = j;
      }
      o_off.push_back(int(o_donor.size()));
    }
  }
  *nrecv = int(o_recv.size());
  *ndonor = int(o_donor.size());
  if (blank) std::copy(o_blank.begin(), o_blank.end(), blank);
  if (recv) std::copy(o_recv.begin(), o_recv.end(), recv);
  if (off) std::copy(o_off.begin(), o_off.end(), off);
  if (donor) std::copy(o_donor.begin(), o_donor.end(), donor);
  if (w) std::copy(o_w.begin(), o_w.end(), w);
  return USG_OK;
}
