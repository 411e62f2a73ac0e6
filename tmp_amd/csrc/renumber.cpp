// Colour-blocked cell renumbering for the LU-SGS wave schedule (host only).
//
// The sweeps run one launch per dependency level of the mesh's cell
// numbering (LUSGS_Integrator.cpp:492-561 sweeps cells in index order); the
// reference's file numbering gives hundreds of levels.  This numbering puts
// the cells of each colour of a proper colouring of the cell graph in one
// contiguous block, so each block is one wave:
//   1. order the cells along a Hilbert curve through their centres (mean of
//      their face midpoints);
//   2. greedy colouring over that order with three colours; each cell whose
//      neighbours already hold all three is repaired by recolouring the
//      ball of cells within 1..8 faces of it with DSatur (Brelaz 1979), the
//      rest held fixed; cells no local repair fixes are repaired the same
//      way with a fourth colour allowed (a 3-colouring of a quadrilateral
//      grid can hold defects that only a global change removes);
//   3. sort cells by (colour, spatial part, Hilbert key), the parts being 8
//      equal slices of the curve; renumber the faces with lC < rC (node
//      order flipped where the left and right cells swap, so the normal still
//      points out of the left cell) and sort them by (part of the left cell,
//      lC, rC), boundary faces of a cell after its interior ones.
// Every step, including the tie-breaking of each sort and of the DSatur
// queue, is the one tmp_amd/mesh.py `prepare` performs, so both give the same
// permutation.
#include <algorithm>
#include <cstdint>
#include <queue>
#include <tuple>
#include <vector>

#include "../../include/usg.h"
#include "geometry.h"

namespace usg {
namespace {

constexpr int kRepairRadius = 8;

struct Graph {
  std::vector<std::vector<int>> nb;  // neighbours in ascending face order
};

uint64_t hilbert_key(int64_t x, int64_t y, int order) {
  uint64_t d = 0;
  for (int64_t s = int64_t(1) << (order - 1); s > 0; s >>= 1) {
    const int64_t rx = (x & s) > 0, ry = (y & s) > 0;
    d += (uint64_t)(s * s * ((3 * rx) ^ ry));
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

// DSatur over the uncoloured cells of `ball`, colours outside held fixed.
// Next: the most distinct colours around, then more neighbours, then the
// lower index (a min-heap of (-saturation, -degree, cell), each tuple
// pushed once).  False when a cell finds all `ncol` colours taken.
bool dsatur_ball(const Graph &g, std::vector<int> &col, const std::vector<int> &ball, int ncol,
                 std::vector<char> &inball) {
  for (int v : ball) inball[v] = 1;
  std::vector<int> sat(ball.size()), nsat(ball.size());
  // position of a cell in the ball, by binary search (balls are small)
  std::vector<std::pair<int, int>> pos(ball.size());
  for (size_t i = 0; i < ball.size(); i++) pos[i] = {ball[i], (int)i};
  std::sort(pos.begin(), pos.end());
  auto where = [&](int v) {
    auto it = std::lower_bound(pos.begin(), pos.end(), std::make_pair(v, -1));
    return it->second;
  };
  using Item = std::tuple<int, int, int>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  for (size_t i = 0; i < ball.size(); i++) {
    const int v = ball[i];
    int m = 0;
    for (int o : g.nb[v])
      if (!inball[o] && col[o] >= 0) m |= 1 << col[o];
    sat[i] = m;
    nsat[i] = __builtin_popcount(m);
    heap.emplace(-nsat[i], -(int)g.nb[v].size(), v);
  }
  bool ok = true;
  while (!heap.empty()) {
    const auto [s, d, v] = heap.top();
    heap.pop();
    (void)d;
    const int i = where(v);
    if (col[v] >= 0 || -s != nsat[i]) continue;
    int k = 0;
    while (k < ncol && (sat[i] & (1 << k))) k++;
    if (k == ncol) {
      ok = false;
      break;
    }
    col[v] = k;
    for (int o : g.nb[v]) {
      if (!inball[o] || col[o] >= 0) continue;
      const int j = where(o);
      if (!(sat[j] & (1 << k))) {
        sat[j] |= 1 << k;
        nsat[j]++;
        heap.emplace(-nsat[j], -(int)g.nb[o].size(), o);
      }
    }
  }
  for (int v : ball) inball[v] = 0;
  return ok;
}

bool repair(const Graph &g, std::vector<int> &col, int v, int ncol, std::vector<char> &seen,
            std::vector<char> &inball) {
  for (int r = 1; r <= kRepairRadius; r++) {
    std::vector<int> ball{v}, front{v};
    seen[v] = 1;
    for (int step = 0; step < r; step++) {
      std::vector<int> nxt;
      for (int x : front)
        for (int o : g.nb[x])
          if (!seen[o]) {
            seen[o] = 1;
            nxt.push_back(o);
          }
      ball.insert(ball.end(), nxt.begin(), nxt.end());
      front.swap(nxt);
    }
    for (int x : ball) seen[x] = 0;
    std::vector<int> saved(ball.size());
    for (size_t i = 0; i < ball.size(); i++) {
      saved[i] = col[ball[i]];
      col[ball[i]] = -1;
    }
    if (dsatur_ball(g, col, ball, ncol, inball)) return true;
    for (size_t i = 0; i < ball.size(); i++) col[ball[i]] = saved[i];
  }
  return false;
}

std::vector<int> three_colours(const Graph &g, const std::vector<int> &visit) {
  const int n = (int)g.nb.size();
  std::vector<int> col(n, -1), left;
  for (int v : visit) {
    int used = 0;
    for (int o : g.nb[v])
      if (col[o] >= 0) used |= 1 << col[o];
    int k = 0;
    while (k < 3 && (used & (1 << k))) k++;
    if (k == 3)
      left.push_back(v);
    else
      col[v] = k;
  }
  std::vector<char> seen(n, 0), inball(n, 0);
  std::vector<int> still;
  for (int v : left)
    if (col[v] < 0 && !repair(g, col, v, 3, seen, inball)) still.push_back(v);
  for (int v : still) {
    if (col[v] < 0 && !repair(g, col, v, 4, seen, inball)) {
      int used = 0;
      for (int o : g.nb[v])
        if (col[o] >= 0) used |= 1 << col[o];
      int k = 0;
      while (used & (1 << k)) k++;
      col[v] = k;
    }
  }
  return col;
}

}  // namespace
}  // namespace usg

extern "C" int usg_mesh_renumber(const usg_mesh_desc *mesh, int *perm, int *fnode, int *flc, int *frc,
                                 int *ncolours) {
  using namespace usg;
  if (!mesh || !perm || !fnode || !flc || !frc || !ncolours) return set_error(USG_ERR_ARG, "null argument");
  const int nc = mesh->nc, nf = mesh->nf;
  if (nc <= 0 || nf <= 0 || !mesh->xy || !mesh->fnode || !mesh->flc || !mesh->frc)
    return set_error(USG_ERR_MESH, "empty mesh");
  for (int f = 0; f < nf; f++) {
    const int l = mesh->flc[f], r = mesh->frc[f];
    if (l < 0 || l >= nc || r >= nc) return set_error(USG_ERR_MESH, "face cell index out of range");
    const int a = mesh->fnode[2 * f], b = mesh->fnode[2 * f + 1];
    if (a < 0 || a >= mesh->nn || b < 0 || b >= mesh->nn) return set_error(USG_ERR_MESH, "face node out of range");
  }
  const double *xy = mesh->xy;
  // 1. centres: face midpoints summed over left faces, then right faces
  std::vector<double> ax(nc, 0.0), ay(nc, 0.0), cnt(nc, 0.0);
  std::vector<double> fx(nf), fy(nf);
  for (int f = 0; f < nf; f++) {
    const int a = mesh->fnode[2 * f], b = mesh->fnode[2 * f + 1];
    fx[f] = 0.5 * (xy[2 * a] + xy[2 * b]);
    fy[f] = 0.5 * (xy[2 * a + 1] + xy[2 * b + 1]);
  }
  for (int f = 0; f < nf; f++) {
    ax[mesh->flc[f]] += fx[f];
    ay[mesh->flc[f]] += fy[f];
    cnt[mesh->flc[f]] += 1;
  }
  for (int f = 0; f < nf; f++)
    if (mesh->frc[f] >= 0) {
      ax[mesh->frc[f]] += fx[f];
      ay[mesh->frc[f]] += fy[f];
      cnt[mesh->frc[f]] += 1;
    }
  std::vector<double> cx(nc), cy(nc);
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
  for (int i = 0; i < nc; i++) {
    if (cnt[i] == 0) return set_error(USG_ERR_MESH, "cell without faces");
    cx[i] = ax[i] / cnt[i];
    cy[i] = ay[i] / cnt[i];
    xmin = std::min(xmin, cx[i]);
    xmax = std::max(xmax, cx[i]);
    ymin = std::min(ymin, cy[i]);
    ymax = std::max(ymax, cy[i]);
  }
  const int order = 16;
  const int64_t n = int64_t(1) << order;
  const double span = std::max(std::max(xmax - xmin, ymax - ymin), 1e-300);
  std::vector<uint64_t> key(nc);
  for (int i = 0; i < nc; i++) {
    int64_t xi = (int64_t)((cx[i] - xmin) / span * (double)(n - 1));
    int64_t yi = (int64_t)((cy[i] - ymin) / span * (double)(n - 1));
    xi = std::min(std::max(xi, int64_t(0)), n - 1);
    yi = std::min(std::max(yi, int64_t(0)), n - 1);
    key[i] = hilbert_key(xi, yi, order);
  }
  std::vector<int> visit(nc);
  for (int i = 0; i < nc; i++) visit[i] = i;
  std::stable_sort(visit.begin(), visit.end(), [&](int a, int b) { return key[a] < key[b]; });
  // 2. colours
  Graph g;
  g.nb.assign(nc, {});
  for (int f = 0; f < nf; f++) {
    const int l = mesh->flc[f], r = mesh->frc[f];
    if (r >= 0) {
      if (g.nb[l].size() >= 4 || g.nb[r].size() >= 4) return set_error(USG_ERR_MESH, "cell with more than 4 faces");
      g.nb[l].push_back(r);
      g.nb[r].push_back(l);
    }
  }
  const std::vector<int> col = three_colours(g, visit);
  // 3. cells by (colour, part, key); faces by (part of left, lC, rC)
  std::vector<int> part(nc);
  for (int r = 0; r < nc; r++) part[visit[r]] = (int)((int64_t)r * kParts / nc);
  std::vector<int> order_c(nc);
  for (int i = 0; i < nc; i++) order_c[i] = i;
  std::stable_sort(order_c.begin(), order_c.end(), [&](int a, int b) {
    if (col[a] != col[b]) return col[a] < col[b];
    if (part[a] != part[b]) return part[a] < part[b];
    return key[a] < key[b];
  });
  std::vector<int> inv(nc);
  for (int i = 0; i < nc; i++) {
    perm[i] = order_c[i];
    inv[order_c[i]] = i;
  }
  std::vector<int> lc2(nf), rc2(nf), n0(nf), n1(nf);
  std::vector<int64_t> sort_r(nf);
  for (int f = 0; f < nf; f++) {
    int l = inv[mesh->flc[f]], r = mesh->frc[f] >= 0 ? inv[mesh->frc[f]] : mesh->frc[f];
    int a = mesh->fnode[2 * f], b = mesh->fnode[2 * f + 1];
    if (r >= 0 && r < l) {
      std::swap(l, r);
      std::swap(a, b);
    }
    lc2[f] = l;
    rc2[f] = r;
    n0[f] = a;
    n1[f] = b;
    sort_r[f] = r >= 0 ? r : (int64_t)nc - r;
  }
  std::vector<int> order_f(nf);
  for (int f = 0; f < nf; f++) order_f[f] = f;
  std::stable_sort(order_f.begin(), order_f.end(), [&](int a, int b) {
    const int pa = part[perm[lc2[a]]], pb = part[perm[lc2[b]]];
    if (pa != pb) return pa < pb;
    if (lc2[a] != lc2[b]) return lc2[a] < lc2[b];
    return sort_r[a] < sort_r[b];
  });
  for (int i = 0; i < nf; i++) {
    const int f = order_f[i];
    fnode[2 * i] = n0[f];
    fnode[2 * i + 1] = n1[f];
    flc[i] = lc2[f];
    frc[i] = rc2[f];
  }
  *ncolours = *std::max_element(col.begin(), col.end()) + 1;
  return USG_OK;
}
