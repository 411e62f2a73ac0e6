// Overset connectivity of 2-D unstructured blocks in one process: UniTioga's
// tioga::performConnectivity (UniTioga/tioga.cpp:152-212) with one mesh
// block per mesh tag and every "processor exchange" a loop over the blocks.
// Each step follows the reference routine named beside it:
//
//   preprocess     MeshBlock::preprocess / tagBoundary (MeshBlock.cpp:135-429):
//                  oriented bounding box, cell and node resolutions, the
//                  node bins of the box, overset-boundary exclusions
//   hole map       tioga::getHoleMap (holeMap.cpp:7-179), markWallBoundary
//                  (MeshBlock.cpp:687-798), fillHoleMap2D (utils.cpp:782-859)
//   boxes          tioga::exchangeBoxes (exchangeboxes.cpp), obbIntersectCheck2D
//                  (utils.cpp:294-352), getReducedOBB (MeshBlock.cpp:800-913)
//   query points   tioga::exchangeSearchData, getQueryPoints2 (MeshBlock.cpp:961-1161)
//   donor search   MeshBlock::search (search.cpp:81-270), ADT::buildADT /
//                  searchADT (ADT.cpp, recursiveADT.cpp), find_median
//                  (mathTools/median.cpp), checkContainment
//   donors         tioga::exchangeDonors (exchangeDonor.cpp), processDonors,
//                  findInterpData, cancellation, setIblanks (bookkeeping.cpp)
//   fringes        tioga::reduce_fringes (tioga.cpp:474-619),
//                  MeshBlock::reduce_fringes (bookkeeping.cpp:572-652)
//   cell iblank    MeshBlock::getCellIblanks (bookkeeping.cpp:654-705)
//
// Query points are distinct nodes of one block per mesh tag, so the
// duplicate-point merge of uniquenodes_octree and resetCoincident leave
// everything unchanged and are not restated.  Plain C++17.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../../include/usg.h"
#include "geometry.h"

using usg::set_error;

namespace {

constexpr int BASE = 1;
constexpr double BIGVALUE = 1.0e+15;
constexpr int BIGINT = 2147483647;
constexpr double TOL = 1.0e-10;
constexpr int HOLEMAPSIZE = 192;
constexpr int D = 2;  // dimension
// tioga::isym (tioga.h:69); mexclude and nfringe are parameters
// (tioga::setMexclude / setNfringe)
constexpr int kIsym = 3;
constexpr double kResolutionScale = 1.0;

struct Obb {
  double xc[3] = {0, 0, 0}, dxc[3] = {0, 0, 0}, vec[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
};

// Jacobi rotations of a symmetric n x n matrix (mathTools/kaiser.cpp);
// columns of a become the eigenvectors
void kaiser(double a[3][3], int n) {
  double trace = 0, ss = 0;
  for (int i = 0; i < n; i++) {
    trace += a[i][i];
    for (int j = 0; j < n; j++) ss += a[i][j] * a[i][j];
  }
  const double eps = 1.0e-12 * ss / n;
  for (int iter = 0; iter < 100; iter++) {
    int skip = 0;
    for (int i = 0; i < n - 1; i++)
      for (int j = i + 1; j < n; j++) {
        double halfP = 0, q = 0;
        for (int k = 0; k < n; k++) {
          halfP += a[k][i] * a[k][j];
          q += (a[k][i] + a[k][j]) * (a[k][i] - a[k][j]);
        }
        const double p = 2 * halfP, absP = std::fabs(p);
        if (absP < eps && q >= 0) {
          skip += 1;
          continue;
        }
        const double absQ = std::fabs(q);
        double c, s;
        if (absP <= absQ) {
          const double t = absP / absQ;
          c = 1 / std::sqrt(1 + t * t);
          s = t * c;
        } else {
          const double ct = absQ / absP;
          s = 1 / std::sqrt(1 + ct * ct);
          c = s * ct;
        }
        c = std::sqrt((1 + c) / 2);
        s = s / (2 * c);
        if (q < 0) std::swap(c, s);
        if (p < 0) s = -s;
        for (int k = 0; k < n; k++) {
          const double t = a[k][i];
          a[k][i] = t * c + a[k][j] * s;
          a[k][j] = -t * s + a[k][j] * c;
        }
      }
    if (skip == 0.5 * n * (n - 1)) break;
  }
  for (int i = 0; i < n; i++) {
    double t = 0;
    for (int j = 0; j < n; j++) t += a[j][i] * a[j][i];
    const double ev = std::sqrt(t);
    for (int j = 0; j < n; j++) a[j][i] = ev > 0 ? a[j][i] / ev : 0;
  }
}

// findOBB (utils.cpp:43-292) in 2-D
void find_obb(const double *x, int n, Obb &o) {
  for (int i = 0; i < D; i++) o.xc[i] = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < D; j++) o.xc[j] += x[D * i + j];
  for (int i = 0; i < D; i++) o.xc[i] = o.xc[i] / n;
  if (n <= D) {
    for (int i = 0; i < D; i++)
      for (int j = 0; j < D; j++) o.vec[i][j] = i == j ? 1 : 0;
    for (int i = 0; i < D; i++) {
      if (n == 1) {
        o.dxc[i] = 1e-3;
      } else if (n == 2) {
        // TIOGA_MAX(1e-3, d) * 0.5 expands to (1e-3 > d ? 1e-3 : d * 0.5)
        const double d = std::fabs(x[D + i] - x[i]);
        o.dxc[i] = 1e-3 > d ? 1e-3 : d * 0.5;
      } else {
        o.dxc[i] = 0;
      }
    }
    return;
  }
  double aa[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < n; i++)
    for (int j = 0; j < D; j++)
      for (int k = j; k < D; k++) aa[j][k] += (x[D * i + j] - o.xc[j]) * (x[D * i + k] - o.xc[k]);
  for (int i = 0; i < D; i++)
    for (int j = 0; j < i; j++) aa[i][j] = aa[j][i];
  kaiser(aa, D);
  for (int i = 0; i < D; i++)
    for (int j = 0; j < D; j++) o.vec[i][j] = aa[i][j];
  double xmin[3], xmax[3], xd[3];
  for (int i = 0; i < D; i++) {
    xmax[i] = -BIGVALUE;
    xmin[i] = BIGVALUE;
  }
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < D; j++) xd[j] = 0;
    for (int j = 0; j < D; j++)
      for (int k = 0; k < D; k++) xd[j] += (x[D * i + k] - o.xc[k]) * o.vec[j][k];
    for (int j = 0; j < D; j++) {
      xmax[j] = xmax[j] > xd[j] ? xmax[j] : xd[j];
      xmin[j] = xmin[j] < xd[j] ? xmin[j] : xd[j];
    }
  }
  for (int i = 0; i < D; i++) {
    o.dxc[i] = (xmax[i] - xmin[i]) * 0.5 * 1.01;
    xd[i] = (xmax[i] + xmin[i]) * 0.5;
  }
  for (int i = 0; i < D; i++)
    for (int j = 0; j < D; j++) o.xc[i] += (xd[j] * o.vec[j][i]);
}

// obbIntersectCheck2D (utils.cpp:294-352), including its |v_A| . D test
int obb_intersect(const Obb &A, const Obb &B) {
  const double eps = 1e-12;
  double Dv[3], c[3][3];
  for (int i = 0; i < 2; i++) Dv[i] = B.xc[i] - A.xc[i];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      c[i][j] = 0;
      for (int k = 0; k < 2; k++) c[i][j] = c[i][j] + A.vec[i][k] * B.vec[j][k];
    }
  for (int i = 0; i < 2; i++) {
    const double r0 = A.dxc[i];
    double r1 = 0, r = 0;
    for (int j = 0; j < 2; j++) {
      r1 += B.dxc[j] * std::fabs(c[i][j]);
      r += std::fabs(A.vec[i][j]) * Dv[j];
    }
    if (r > (r0 + r1 + eps)) return 0;
  }
  for (int i = 0; i < 2; i++) {
    const double r1 = B.dxc[i];
    double r0 = 0, r = 0;
    for (int j = 0; j < 2; j++) {
      r0 += A.dxc[j] * std::fabs(c[j][i]);
      r += std::fabs(B.vec[i][j]) * Dv[j];
    }
    if (r > (r0 + r1 + eps)) return 0;
  }
  return 1;
}

void to_obb(const double *x, const double *xc, const double vec[3][3], double *xd) {
  for (int j = 0; j < D; j++) {
    xd[j] = 0;
    for (int k = 0; k < D; k++) xd[j] += (x[k] - xc[k]) * vec[j][k];
  }
}

// getobbcoords (utils.cpp:454-469)
void obb_corners(const double *xc, const double *dxc, const double vec[3][3], double xv[4][3]) {
  for (int i = 0; i < 4; i++) {
    for (int k = 0; k < D; k++) xv[i][k] = xc[k];
    for (int k = 0; k < D; k++) {
      const int ik = (2 * ((i & (1 << k)) >> k) - 1);
      for (int j = 0; j < D; j++) xv[i][j] += ik * dxc[k] * vec[k][j];
    }
  }
}

// find_median (mathTools/median.cpp:5-186): partial partition of x (and the
// companion ix) around the median
void find_median(int *ix, double *x, int n, int &ilow, int &ihigh) {
  int low = ilow, high = ihigh;
  const int mid = (low + high) / 2;
  const int nby2 = (n - 1) / 2, nby2p1 = (n - 1) / 2 + 1;
  const bool is_odd = n % 2 == 1;
  double xlow = x[low], xhigh = x[high], xmed = x[mid];
  if (xlow > xhigh) std::swap(xlow, xhigh);
  if (xmed > xhigh)
    xmed = xhigh;
  else if (xmed < xlow)
    xmed = xlow;
  int i = low, j = high;
  bool again = true;
  while (again) {
    again = false;
    while (!(x[i] >= xmed)) ++i;
    while (!(x[j] <= xmed)) --j;
    if (i < j) {
      std::swap(x[i], x[j]);
      std::swap(ix[i], ix[j]);
      i = i + 1;
      j = j - 1;
      if (i <= j) again = true;
    }
  }
  if (!is_odd) {
    if (j == nby2 && i == nby2p1) {
      int idx = j;
      for (int k = 0; k < nby2; k++)
        if (x[k] > x[idx]) idx = k;
      if (idx != j) {
        std::swap(x[idx], x[j]);
        std::swap(ix[idx], ix[j]);
      }
      return;
    }
    if (j < nby2) low = i;
    if (i > nby2p1) high = j;
    if (i != j) {
      if (low < high - 1) find_median(ix, x, n, low, high);
      return;
    }
    if (i == nby2) low = nby2;
    if (j == nby2p1) high = nby2p1;
  } else {
    if (j < nby2) low = i;
    if (i > nby2) high = j;
    if (i != j) {
      if (low < high - 1) find_median(ix, x, n, low, high);
      ilow = low;
      ihigh = high;
      return;
    }
    if (i == nby2) {
      ilow = low;
      ihigh = high;
      return;
    }
  }
  if (low < high - 1) {
    ilow = low;
    ihigh = high;
    find_median(ix, x, n, low, high);
  }
}

void find_median_wrap(int *ix, double *x, int n) {
  if (n <= 2) {
    if (n == 2 && x[0] > x[1]) {
      std::swap(x[0], x[1]);
      std::swap(ix[0], ix[1]);
    }
    return;
  }
  int ilow = 0, ihigh = n - 1;
  const int nby2 = (n - 1) / 2, nby2p1 = nby2 + 1;
  find_median(ix, x, n, ilow, ihigh);
  if (n % 2 == 0) {
    if (x[nby2] > x[nby2p1]) {
      std::swap(x[nby2], x[nby2p1]);
      std::swap(ix[nby2], ix[nby2p1]);
    }
  } else if (x[ilow] > x[ihigh]) {
    std::swap(x[ilow], x[ihigh]);
    std::swap(ix[ilow], ix[ihigh]);
  }
}

struct Block;

// Alternating digital tree over element bounding boxes (ADT.cpp,
// recursiveADT.cpp): coord[ndim*e] = (xmin, ymin, xmax, ymax)
struct Adt {
  int ndim = 4, nelem = 0;
  const double *coord = nullptr;
  std::vector<double> ext, reals, work;
  std::vector<int> ints;

  void build_rec(int *avail, int &count, int side, int parent, int level, int nav) {
    const int nd = ndim / 2;
    if (nav > 1) {
      const int dimcut = level % ndim;
      for (int i = 0; i < nav; i++) work[i] = coord[ndim * avail[i] + dimcut];
      find_median_wrap(avail, work.data(), nav);
      const int nleft = (nav + 1) / 2;
      count++;
      int ii = count * 4;
      ints[ii] = avail[nleft - 1];
      ints[ii + 1] = ints[ii + 2] = ints[ii + 3] = -1;
      for (int i = 0; i < nd; i++) {
        reals[ndim * count + i] = BIGVALUE;
        reals[ndim * count + i + nd] = -BIGVALUE;
      }
      for (int i = 0; i < nav; i++)
        for (int j = 0; j < nd; j++) {
          const int a = ndim * count + j, b = ndim * avail[i] + j;
          reals[a] = reals[a] < coord[b] ? reals[a] : coord[b];
          reals[a + nd] = reals[a + nd] > coord[b + nd] ? reals[a + nd] : coord[b + nd];
        }
      if (side > 0) ints[4 * parent + side] = avail[nleft - 1];
      const int me = count;
      if (nleft > 1) build_rec(avail, count, 1, me, level + 1, nleft - 1);
      build_rec(avail + nleft, count, 2, me, level + 1, nav - nleft);
    } else if (nav == 1) {
      count++;
      const int ii = 4 * count, jj = ndim * count;
      ints[ii] = avail[0];
      ints[ii + 1] = ints[ii + 2] = ints[ii + 3] = -1;
      for (int j = 0; j < ndim; j++) reals[jj + j] = coord[ndim * avail[0] + j];
      if (side > 0) ints[4 * parent + side] = avail[0];
    }
  }

  void build(int n, const double *bbox) {
    nelem = n;
    coord = bbox;
    ext.assign(ndim, 0);
    ints.assign(4 * (size_t)std::max(n, 1), -1);
    reals.assign((size_t)ndim * std::max(n, 1), 0);
    work.assign(std::max(n, 1), 0);
    for (int i = 0; i < ndim / 2; i++) {
      ext[2 * i] = BIGVALUE;
      ext[2 * i + 1] = -BIGVALUE;
    }
    for (int j = 0; j < n; j++)
      for (int i = 0; i < ndim / 2; i++) {
        ext[2 * i] = ext[2 * i] < coord[ndim * j + i] ? ext[2 * i] : coord[ndim * j + i];
        const double v = coord[ndim * j + i + ndim / 2];
        ext[2 * i + 1] = ext[2 * i + 1] > v ? ext[2 * i + 1] : v;
      }
    for (int i = 0; i < ndim / 2; i++) {
      const double delta = 0.01 * (ext[2 * i + 1] - ext[2 * i]);
      ext[2 * i] -= delta;
      ext[2 * i + 1] += delta;
    }
    std::vector<int> avail(std::max(n, 1));
    for (int i = 0; i < n; i++) avail[i] = i;
    int count = -1;
    build_rec(avail.data(), count, 0, 0, 0, n);
    for (int i = 0; i < n; i++) ints[4 * ints[4 * i] + 3] = i;
  }

  void search_rec(const Block &mb, int *ci, int node, const double *xs) const;
  void search(const Block &mb, int *ci, const double *xs) const {
    ci[0] = -1;
    ci[1] = 0;
    if (nelem == 0) return;
    bool flag = true;
    for (int i = 0; i < ndim / 2; i++) flag = flag && (xs[i] >= ext[2 * i] - TOL);
    for (int i = 0; i < ndim / 2; i++) flag = flag && (xs[i] <= ext[2 * i + 1] + TOL);
    if (flag) search_rec(mb, ci, 0, xs);
  }
};

struct Donor {
  int meshtag, remoteid;
  double donorRes, receptorRes;
};

struct Interp {
  int cancel = 0, nweights = 0, pointid = 0, blockid = 0;
  int inode[5] = {0, 0, 0, 0, 0};
  double w[5] = {0, 0, 0, 0, 0};
  int donor_cell = -1;  // type-major index
};

struct Block {
  int tag = 0, nnodes = 0, ncells = 0;
  int mexclude = 3, nfringe = 1;  // MeshBlock defaults (MeshBlock.h)
  std::vector<double> x;                  // [2*nnodes]
  std::vector<int> nv, nc;                // per type (ascending vertex count)
  std::vector<std::vector<int>> vconn;    // per type, 1-based node ids
  std::vector<int> cell_of;               // type-major index -> input cell id
  std::vector<int> wbc, obc;              // 1-based, ascending
  Obb obb;
  std::vector<double> nodeRes, cellRes;
  double mapdx[3] = {0, 0, 0};
  int mapdims[3] = {10, 10, 1};
  std::vector<int> icft, invmap;
  std::vector<int> iblank, iblank_reduced, iblank_cell;
  bool have_reduced = false;
  // search
  std::vector<double> xsearch, res_search;
  std::vector<int> isearch, tagsearch, donorId;
  std::vector<double> elementBbox;
  std::vector<int> elementList;
  Adt adt;
  // donors
  std::vector<std::vector<Donor>> donorList;
  std::vector<Interp> interp;
  std::vector<int> interp2donor;
  std::vector<int> cancelList, cancelArray;

  int nsearch() const { return (int)tagsearch.size(); }

  void cell_nodes(int icell, int &nvert, int *inode) const {
    int isum = 0, n = 0, i = 0;
    for (n = 0; n < (int)nc.size(); n++) {
      isum += nc[n];
      if (icell < isum) {
        i = icell - (isum - nc[n]);
        break;
      }
    }
    nvert = nv[n];
    for (int m = 0; m < nvert; m++) inode[m] = vconn[n][(size_t)nvert * i + m] - BASE;
  }

  // checkContainment (checkContainment.cpp)
  void contain(int *ci, int adtElement, const double *xs) const {
    const int icell = elementList[adtElement];
    int nvert, inode[4];
    cell_nodes(icell, nvert, inode);
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0}, frac[4];
    for (int m = 0; m < nvert; m++)
      for (int j = 0; j < D; j++) v[D * m + j] = x[D * (size_t)inode[m] + j];
    const int rc = usg::nodal_weight_2d(v, nvert, xs, frac);
    ci[0] = icell;
    ci[1] = 0;
    for (int m = 0; m < nvert; m++) {
      if (rc != 0 || (frac[m] + TOL) * (frac[m] - 1.0 - TOL) > 0) {
        ci[0] = -1;
        return;
      }
      if (std::fabs(frac[m]) < TOL && cellRes[icell] == BIGVALUE) ci[1] = 1;
    }
  }
};

void Adt::search_rec(const Block &mb, int *ci, int node, const double *xs) const {
  double el[4];
  for (int i = 0; i < ndim; i++) el[i] = coord[ndim * ints[4 * node] + i];
  bool flag = true;
  for (int i = 0; i < ndim / 2; i++) flag = flag && (xs[i] >= el[i] - TOL);
  for (int i = ndim / 2; i < ndim; i++) flag = flag && (xs[i - ndim / 2] <= el[i] + TOL);
  if (flag) {
    mb.contain(ci, ints[4 * node], xs);
    if (ci[0] > -1 && ci[1] == 0) return;
  }
  for (int d = 1; d < 3; d++) {
    int child = ints[4 * node + d];
    if (child > -1) {
      child = ints[4 * child + 3];
      for (int i = 0; i < ndim; i++) el[i] = reals[ndim * child + i];
      flag = true;
      for (int i = 0; i < ndim / 2; i++) flag = flag && (xs[i] >= el[i] - TOL);
      for (int i = ndim / 2; i < ndim; i++) flag = flag && (xs[i - ndim / 2] <= el[i] + TOL);
      if (flag) {
        search_rec(mb, ci, child, xs);
        if (ci[0] > -1 && ci[1] == 0) return;
      }
    }
  }
}

struct HoleMap {
  int existWall = 0;
  int nx[3] = {0, 0, 0};
  double extents[6] = {0, 0, 0, 0, 0, 0};
  std::vector<int> sam;
};

// MeshBlock::preprocess + tagBoundary (MeshBlock.cpp:135-429)
void preprocess(Block &b) {
  b.iblank.assign(b.nnodes, 1);
  find_obb(b.x.data(), b.nnodes, b.obb);
  const int dim_vol = b.mapdims[0] * b.mapdims[1];
  for (int i = 0; i < D; i++) b.mapdx[i] = 2 * b.obb.dxc[i] / b.mapdims[i];
  const int frac[2] = {1, b.mapdims[0]};
  b.cellRes.assign(b.ncells, 0.0);
  b.nodeRes.assign(b.nnodes, 0.0);
  b.icft.assign(dim_vol + 1, -1);
  b.icft[0] = 0;
  b.invmap.assign(b.nnodes, 0);
  std::vector<int> iflag(b.nnodes, 0), iext(b.nnodes, 0), iext1(b.nnodes, 0), iptr(b.nnodes, 0);
  int k = 0;
  for (size_t n = 0; n < b.nc.size(); n++) {
    const int nvert = b.nv[n];
    for (int i = 0; i < b.nc[n]; i++) {
      int inode[4];
      double xv[4][2];
      for (int m = 0; m < nvert; m++) {
        inode[m] = b.vconn[n][(size_t)nvert * i + m] - BASE;
        for (int j = 0; j < D; j++) xv[m][j] = b.x[D * (size_t)inode[m] + j];
      }
      const double vol = usg::cell_volume(xv, nvert);
      b.cellRes[k] = (vol * kResolutionScale);
      for (int m = 0; m < nvert; m++) {
        ++iflag[inode[m]];
        b.nodeRes[inode[m]] += b.cellRes[k];
      }
      ++k;
    }
  }
  for (int i = 0; i < b.nnodes; i++) {
    if (iflag[i] != 0) b.nodeRes[i] = b.nodeRes[i] / iflag[i];
    iflag[i] = 0;
    double xd[3];
    int idx[3];
    for (int j = 0; j < D; j++) {
      xd[j] = b.obb.dxc[j];
      for (int kk = 0; kk < D; kk++) xd[j] += (b.x[D * (size_t)i + kk] - b.obb.xc[kk]) * b.obb.vec[j][kk];
      idx[j] = static_cast<int>(xd[j] / b.mapdx[j]);
    }
    int indx = 0;
    for (int j = 0; j < D; j++) indx += idx[j] * frac[j];
    indx = std::min(std::max(indx, 0), dim_vol - 1);  // in range for a box of nonzero size
    iptr[i] = b.icft[indx + 1];
    b.icft[indx + 1] = i;
  }
  int kc = 0;
  for (int i = 0; i < dim_vol; i++) {
    int ip = b.icft[i + 1], m = 0;
    while (ip != -1) {
      b.invmap[kc++] = ip;
      ip = iptr[ip];
      m++;
    }
    b.icft[i + 1] = b.icft[i] + m;
  }
  for (int v : b.obc) iflag[v - BASE] = 1;
  for (size_t n = 0; n < b.nc.size(); n++) {
    const int nvert = b.nv[n];
    for (int i = 0; i < b.nc[n]; i++) {
      bool itag = false;
      for (int m = 0; m < nvert; m++)
        if (iflag[b.vconn[n][(size_t)nvert * i + m] - BASE]) itag = true;
      if (itag)
        for (int m = 0; m < nvert; m++) {
          const int v = b.vconn[n][(size_t)nvert * i + m] - BASE;
          b.nodeRes[v] = BIGVALUE;
          iext[v] = iext1[v] = 1;
        }
    }
  }
  for (int iex = 0; iex < b.mexclude; iex++) {
    int kk = 0;
    for (size_t n = 0; n < b.nc.size(); n++) {
      const int nvert = b.nv[n];
      for (int i = 0; i < b.nc[n]; i++) {
        for (int m = 0; m < nvert; m++)
          if (iext[b.vconn[n][(size_t)nvert * i + m] - BASE] == 1) {
            b.cellRes[kk] = BIGVALUE;
            break;
          }
        if (b.cellRes[kk] == BIGVALUE)
          for (int m = 0; m < nvert; m++) {
            const int v = b.vconn[n][(size_t)nvert * i + m] - BASE;
            if (iext[v] != 1) iext1[v] = 1;
          }
        kk++;
      }
    }
    iext = iext1;
  }
}

// getWallBounds + markWallBoundary + fillHoleMap2D (one block per tag)
void hole_map(const Block &b, HoleMap &h) {
  if (b.wbc.empty()) {
    h.existWall = 0;
    return;
  }
  h.existWall = 1;
  double wbox[4];
  for (int i = 0; i < D; i++) {
    wbox[i] = BIGVALUE;
    wbox[D + i] = -BIGVALUE;
  }
  for (int v : b.wbc)
    for (int j = 0; j < D; j++) {
      const double xv = b.x[D * (size_t)(v - BASE) + j];
      wbox[j] = wbox[j] < xv ? wbox[j] : xv;
      wbox[D + j] = wbox[D + j] > xv ? wbox[D + j] : xv;
    }
  double ds[3] = {0, 0, 0};  // the reference leaves ds[2] unset in 2-D
  for (int j = 0; j < D; j++) {
    h.extents[j] = wbox[j];
    h.extents[j + D] = wbox[j + D];
    ds[j] = h.extents[j + D] - h.extents[j];
  }
  double dsmax = ds[0] > ds[1] ? ds[0] : ds[1];
  dsmax = dsmax > ds[2] ? dsmax : ds[2];
  const double dsbox = dsmax / HOLEMAPSIZE;
  for (int j = 0; j < D; j++) {
    h.extents[j] -= (2 * dsbox);
    h.extents[j + D] += (2 * dsbox);
    const double r = (h.extents[j + D] - h.extents[j]) / dsbox;
    h.nx[j] = (int)std::floor(r > 1 ? r : 1);
  }
  const int nx = h.nx[0], ny = h.nx[1];
  h.sam.assign((size_t)nx * ny, 0);
  // markWallBoundary: bins covered by cells touching a wall node
  std::vector<char> wall(b.nnodes, 0);
  for (int v : b.wbc) wall[v - BASE] = 1;
  double dsx[2];
  for (int kk = 0; kk < D; kk++) dsx[kk] = (h.extents[kk + D] - h.extents[kk]) / h.nx[kk];
  for (size_t n = 0; n < b.nc.size(); n++) {
    const int nvert = b.nv[n];
    for (int i = 0; i < b.nc[n]; i++) {
      bool touch = false;
      for (int j = 0; j < nvert && !touch; j++) touch = wall[b.vconn[n][(size_t)nvert * i + j] - BASE] == 1;
      if (!touch) continue;
      int imin[2] = {BIGINT, BIGINT}, imax[2] = {-BIGINT, -BIGINT};
      for (int j = 0; j < nvert; j++) {
        const int id = D * (b.vconn[n][(size_t)nvert * i + j] - BASE);
        for (int kk = 0; kk < D; kk++) {
          const int iv = (int)std::floor((b.x[id + kk] - h.extents[kk]) / dsx[kk]);
          imin[kk] = std::min(imin[kk], iv);
          imax[kk] = std::max(imax[kk], iv);
        }
      }
      for (int j = 0; j < D; j++) {
        imin[j] = std::max(imin[j], 0);
        imax[j] = std::min(imax[j], h.nx[j] - 1);
      }
      for (int jj = imin[1]; jj < imax[1] + 1; jj++)
        for (int ii = imin[0]; ii < imax[0] + 1; ii++) h.sam[(size_t)jj * nx + ii] = 2;
    }
  }
  // fillHoleMap2D: paint from the box boundary inwards, 4-neighbour flood
  int *hm = h.sam.data();
  for (int jj = 0; jj < ny; jj += std::max(ny - 1, 1))
    for (int ii = 0; ii < nx; ii++) hm[jj * nx + ii] = 1;
  for (int jj = 0; jj < ny; jj++)
    for (int ii = 0; ii < nx; ii += std::max(nx - 1, 1)) hm[jj * nx + ii] = 1;
  int npaint = nx * ny;
  while (npaint > 0) {
    npaint = 0;
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        const int m = j * nx + i;
        if (hm[m] != 0) continue;
        const int mk[4] = {m - nx, m + nx, m - 1, m + 1};
        const int nneig = kIsym == 1 || kIsym == 2 ? 2 : 4;
        int ipaint = 0;
        for (int kk = 0; kk < nneig && ipaint == 0; kk++) ipaint = (ipaint || hm[mk[kk]] == 1);
        if (ipaint > 0) {
          hm[m] = 1;
          npaint++;
        }
      }
  }
  for (int i = 0; i < nx * ny; i++) hm[i] = (hm[i] == 0 || hm[i] == 2);
}

// checkHoleMap (utils.cpp:645-667)
int check_hole_map(const double *x, const HoleMap &h) {
  double dx[2];
  int ix[2];
  for (int i = 0; i < D; i++) dx[i] = (h.extents[D + i] - h.extents[i]) / h.nx[i];
  for (int i = 0; i < D; i++) {
    ix[i] = (int)((x[i] - h.extents[i]) / dx[i]);
    if (ix[i] < 0 || ix[i] > h.nx[i] - 1) return 0;
  }
  return h.sam[(size_t)ix[1] * h.nx[0] + ix[0]];
}

// getReducedOBB (MeshBlock.cpp:800-913): the part of block b's cells that
// overlaps box obc, as a box in b's own frame
void reduced_obb(const Block &b, const Obb &obc, Obb &out) {
  double rd[4];
  for (int i = 0; i < D; i++) {
    rd[i] = BIGVALUE;
    rd[D + i] = -BIGVALUE;
  }
  int involved = 0;
  for (size_t n = 0; n < b.nc.size(); n++) {
    const int nvert = b.nv[n];
    for (int i = 0; i < b.nc[n]; i++) {
      double bbox[4] = {BIGVALUE, BIGVALUE, -BIGVALUE, -BIGVALUE}, xd[3];
      for (int m = 0; m < nvert; m++) {
        to_obb(&b.x[D * (size_t)(b.vconn[n][(size_t)nvert * i + m] - BASE)], obc.xc, obc.vec, xd);
        for (int j = 0; j < D; j++) {
          bbox[j] = bbox[j] < xd[j] ? bbox[j] : xd[j];
          bbox[D + j] = bbox[D + j] > xd[j] ? bbox[D + j] : xd[j];
        }
      }
      bool out_of = false;
      for (int j = 0; j < D; j++) out_of = out_of || (bbox[j] > obc.dxc[j]);
      if (out_of) continue;
      out_of = false;
      for (int j = 0; j < D; j++) out_of = out_of || (bbox[D + j] < -obc.dxc[j]);
      if (out_of) continue;
      involved++;
      for (int m = 0; m < nvert; m++) {
        to_obb(&b.x[D * (size_t)(b.vconn[n][(size_t)nvert * i + m] - BASE)], b.obb.xc, b.obb.vec, xd);
        for (int j = 0; j < D; j++) {
          rd[j] = rd[j] < xd[j] ? rd[j] : xd[j];
          rd[D + j] = rd[D + j] > xd[j] ? rd[D + j] : xd[j];
        }
      }
    }
  }
  double bb[4] = {rd[0], rd[1], rd[2], rd[3]};
  for (int i = 0; i < D; i++) {
    out.xc[i] = b.obb.xc[i];
    for (int j = 0; j < D; j++) out.xc[i] += ((bb[j] + bb[D + j]) * 0.5) * b.obb.vec[j][i];
    out.dxc[i] = (bb[D + i] - bb[i]) * 0.51;
  }
  if (involved == 0)
    for (int i = 0; i < D; i++) out.xc[i] = out.dxc[i] = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) out.vec[i][j] = b.obb.vec[i][j];
}

// getQueryPoints2 (MeshBlock.cpp:961-1161): nodes of b inside box obc,
// visited through b's node bins
void query_points(const Block &b, const Obb &obc, std::vector<int> &nodes) {
  nodes.clear();
  double xv[4][3], xd[3], xc[3], xmin[3], xmax[3], mdx[3];
  int imin[3], imax[3];
  obb_corners(obc.xc, obc.dxc, obc.vec, xv);
  for (int i = 0; i < D; i++) {
    xmin[i] = BIGVALUE;
    xmax[i] = -BIGVALUE;
  }
  for (int n = 0; n < 4; n++) {
    to_obb(xv[n], b.obb.xc, b.obb.vec, xd);
    for (int j = 0; j < D; j++) {
      xmin[j] = xmin[j] < xd[j] + b.obb.dxc[j] ? xmin[j] : xd[j] + b.obb.dxc[j];
      xmax[j] = xmax[j] > xd[j] + b.obb.dxc[j] ? xmax[j] : xd[j] + b.obb.dxc[j];
    }
  }
  for (int i = 0; i < D; i++) {
    const double delta = 0.01 * (xmax[i] - xmin[i]);
    xmin[i] -= delta;
    xmax[i] += delta;
    imin[i] = std::max(static_cast<int>(xmin[i] / b.mapdx[i]), 0);
    imax[i] = std::min(static_cast<int>(xmax[i] / b.mapdx[i]), b.mapdims[i] - 1);
    mdx[i] = 0.5 * b.mapdx[i];
  }
  obb_corners(obc.xc, mdx, b.obb.vec, xv);
  for (int i = 0; i < D; i++) {
    xmin[i] = BIGVALUE;
    xmax[i] = -BIGVALUE;
  }
  for (int n = 0; n < 4; n++) {
    to_obb(xv[n], obc.xc, obc.vec, xd);
    for (int i = 0; i < D; i++) {
      xmin[i] = xmin[i] < xd[i] ? xmin[i] : xd[i];
      xmax[i] = xmax[i] > xd[i] ? xmax[i] : xd[i];
    }
  }
  for (int k = imin[1]; k < imax[1] + 1; k++)
    for (int j = imin[0]; j < imax[0] + 1; j++) {
      xd[0] = -b.obb.dxc[0] + j * b.mapdx[0] + b.mapdx[0] * 0.5;
      xd[1] = -b.obb.dxc[1] + k * b.mapdx[1] + b.mapdx[1] * 0.5;
      for (int n = 0; n < 2; n++) {
        xc[n] = b.obb.xc[n];
        for (int ij = 0; ij < 2; ij++) xc[n] += (xd[ij] * b.obb.vec[ij][n]);
      }
      to_obb(xc, obc.xc, obc.vec, xd);
      bool skip = false;
      for (int ij = 0; ij < 2 && !skip; ij++) skip = skip || (xmin[ij] + xd[ij] > obc.dxc[ij]);
      if (skip) continue;
      for (int ij = 0; ij < 2 && !skip; ij++) skip = skip || (xmax[ij] + xd[ij] < -obc.dxc[ij]);
      if (skip) continue;
      const int indx = k * b.mapdims[0] + j;
      for (int m = b.icft[indx]; m < b.icft[indx + 1]; m++) {
        const int v = b.invmap[m];
        to_obb(&b.x[D * (size_t)v], obc.xc, obc.vec, xd);
        if (std::fabs(xd[0]) <= obc.dxc[0] && std::fabs(xd[1]) <= obc.dxc[1]) nodes.push_back(v);
      }
    }
}

// MeshBlock::search (search.cpp:81-270)
void search(Block &b) {
  const int ns = b.nsearch();
  b.donorId.assign(ns, -1);
  if (ns == 0) return;
  Obb obq;
  find_obb(b.xsearch.data(), ns, obq);
  std::vector<int> icell(b.ncells, -1);
  int iptr = -1, count = 0, p = 0;
  for (size_t n = 0; n < b.nc.size(); n++) {
    const int nvert = b.nv[n];
    for (int i = 0; i < b.nc[n]; i++) {
      double xmin[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, xmax[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
      double xd[3] = {0, 0, 0}, dxc[3] = {0, 0, 0};
      for (int m = 0; m < nvert; m++) {
        const int id = D * (b.vconn[n][(size_t)nvert * i + m] - BASE);
        for (int j = 0; j < D; j++) {
          xd[j] = 0;
          for (int k = 0; k < D; k++) xd[j] += (b.x[id + k] - obq.xc[k]) * obq.vec[j][k];
          xmin[j] = xmin[j] < xd[j] ? xmin[j] : xd[j];
          xmax[j] = xmax[j] > xd[j] ? xmax[j] : xd[j];
        }
        for (int j = 0; j < D; j++) {
          xd[j] = (xmax[j] + xmin[j]) * 0.5;
          dxc[j] = (xmax[j] - xmin[j]) * 0.5;
        }
      }
      bool cross = true;
      for (int j = 0; j < D; j++) cross = cross && (std::fabs(xd[j]) <= dxc[j] + obq.dxc[j]);
      if (cross) {
        icell[p] = iptr;
        iptr = p;
        count++;
      }
      p++;
    }
  }
  b.elementBbox.assign(2 * D * (size_t)std::max(count, 1), 0.0);
  b.elementList.assign(std::max(count, 1), 0);
  int l = 0;
  for (int k = iptr; k != -1; k = icell[k]) {
    int nvert, inode[4];
    b.cell_nodes(k, nvert, inode);
    double xmin[2] = {BIGVALUE, BIGVALUE}, xmax[2] = {-BIGVALUE, -BIGVALUE};
    for (int m = 0; m < nvert; m++)
      for (int j = 0; j < D; j++) {
        const double v = b.x[D * (size_t)inode[m] + j];
        xmin[j] = xmin[j] < v ? xmin[j] : v;
        xmax[j] = xmax[j] > v ? xmax[j] : v;
      }
    for (int q = 0; q < D; q++) {
      b.elementBbox[(2 * l) * D + q] = xmin[q];
      b.elementBbox[(2 * l + 1) * D + q] = xmax[q];
    }
    b.elementList[l] = k;
    l++;
  }
  b.adt.build(count, b.elementBbox.data());
  for (int i = 0; i < ns; i++) {
    int ci[2];
    b.adt.search(b, ci, &b.xsearch[D * (size_t)i]);
    b.donorId[i] = ci[0];
  }
}

// MeshBlock::processDonors (bookkeeping.cpp:75-277); receptor records as
// (donor block tag, remote id, receptor resolution)
void process_donors(Block &b, const std::vector<HoleMap> &holemap, std::vector<std::array<double, 3>> &rec) {
  const int nmesh = (int)holemap.size();
  std::vector<int> iflag(nmesh);
  for (int i = 0; i < b.nnodes; i++) {
    b.iblank[i] = 1;
    if (b.donorList[i].empty()) {
      for (int j = 0; j < nmesh; j++)
        if (j != (b.tag - BASE) && holemap[j].existWall && check_hole_map(&b.x[D * (size_t)i], holemap[j])) {
          b.iblank[i] = 0;
          break;
        }
    } else {
      std::fill(iflag.begin(), iflag.end(), 0);
      for (const Donor &d : b.donorList[i]) {
        iflag[d.meshtag - BASE] = 1;
        b.nodeRes[i] = b.nodeRes[i] > d.receptorRes ? b.nodeRes[i] : d.receptorRes;
      }
      for (int j = 0; j < nmesh; j++)
        if (j != (b.tag - BASE) && holemap[j].existWall && !iflag[j] &&
            check_hole_map(&b.x[D * (size_t)i], holemap[j])) {
          b.iblank[i] = 0;
          break;
        }
    }
  }
  std::vector<int> mtag(b.nnodes, 0), mtag1(b.nnodes, 0);
  for (int i = 0; i < b.nnodes; i++)
    if (b.iblank[i] == 0) mtag[i] = mtag1[i] = 1;
  for (int iter = 0; iter < b.nfringe; iter++) {
    for (size_t n = 0; n < b.nc.size(); n++) {
      const int nvert = b.nv[n];
      for (int i = 0; i < b.nc[n]; i++) {
        const int *cv = &b.vconn[n][(size_t)nvert * i];
        for (int m = 0; m < nvert; m++)
          if (mtag[cv[m] - BASE] == 1)
            for (int mm = 0; mm < nvert; mm++)
              if (m != mm && mtag[cv[mm] - BASE] != 1) mtag1[cv[mm] - BASE] = 1;
      }
    }
    mtag = mtag1;
  }
  for (int i = 0; i < b.nnodes; i++)
    if (mtag1[i] && b.iblank[i]) b.nodeRes[i] = BIGVALUE;
  for (int i = 0; i < b.nnodes; i++) {
    if (b.donorList[i].empty() || b.iblank[i] == 0) continue;
    for (const Donor &d : b.donorList[i])
      if (d.donorRes < b.nodeRes[i]) {
        b.iblank[i] = -d.meshtag;
        rec.push_back({(double)d.meshtag, (double)d.remoteid,
                       kResolutionScale > 1.0 ? -b.nodeRes[i] : b.nodeRes[i]});
        break;
      }
  }
}

// MeshBlock::findInterpData (bookkeeping.cpp:310-428) on the donor block
int find_interp(Block &b, int &recid, int irecord, double receptorRes2) {
  const double receptorRes = std::fabs(receptorRes2);
  const int meshtagrecv = b.tagsearch[irecord];
  const double *xp = &b.xsearch[D * (size_t)irecord];
  int nvert, inode[4];
  b.cell_nodes(b.donorId[irecord], nvert, inode);
  int accept = 1;
  double v[8];
  for (int m = 0; m < nvert; m++) {
    if (b.iblank[inode[m]] <= 0 && receptorRes2 > 0.0) {
      if (b.nodeRes[inode[m]] == BIGVALUE) accept = 0;
      if (std::abs(b.iblank[inode[m]]) == meshtagrecv) accept = 0;
    }
    for (int j = 0; j < D; j++) v[D * m + j] = b.x[D * (size_t)inode[m] + j];
  }
  if (receptorRes == BIGVALUE && kResolutionScale == 1.0)
    for (int m = 0; m < nvert; m++)
      if (b.iblank[inode[m]] <= 0 && b.nodeRes[inode[m]] != BIGVALUE) {
        if (b.iblank[inode[m]] < 0) b.iblank[inode[m]] = 1;
        b.cancelList.push_back(inode[m]);
      }
  double frac[4];
  if (usg::nodal_weight_2d(v, nvert, xp, frac) != 0) return -1;
  b.interp2donor[irecord] = recid;
  Interp &it = b.interp[recid];
  it.cancel = 0;
  it.nweights = nvert;
  it.pointid = b.isearch[3 * (size_t)irecord + 1];
  it.blockid = b.isearch[3 * (size_t)irecord + 2];
  for (int m = 0; m < nvert; m++) {
    it.inode[m] = inode[m];
    it.w[m] = frac[m];
    if (frac[m] < -0.2 || frac[m] > 1.2) return -2;  // "Frac Problem Again?"
  }
  it.inode[nvert] = b.donorId[irecord];
  it.w[nvert] = 0.0;
  it.donor_cell = b.donorId[irecord];
  if (accept == 0 && receptorRes != BIGVALUE) it.cancel = 1;
  recid++;
  return 0;
}

struct Overset {
  std::vector<Block> blocks;
  std::vector<HoleMap> holemap;
};

// the receptor-side cancellation records of a list of nodes (getCancellationData /
// getCancellationReduce) applied to the donor blocks (cancelDonor)
void cancel_nodes(Overset &o, Block &b, const std::vector<int> &nodes) {
  for (int inode : nodes) {
    if (b.donorList[inode].empty()) continue;
    const Donor &d = b.donorList[inode].front();
    Block &db = o.blocks[d.meshtag - BASE];
    const int iptr = db.interp2donor[d.remoteid];
    if (iptr > -1) db.interp[iptr].cancel = 1;
  }
}

// getInterpData + clearIblanks + setIblanks
void apply_interp(Overset &o) {
  for (Block &b : o.blocks) {
    for (int i = 0; i < b.nnodes; i++)
      if (b.iblank[i] < 0) b.iblank[i] = 1;
    if (b.have_reduced) {
      for (int i = 0; i < b.nnodes; i++)
        if (b.iblank_reduced[i] == 0) b.iblank[i] = 0;
      b.have_reduced = false;
    }
  }
  for (Block &b : o.blocks)
    for (const Interp &it : b.interp)
      if (!it.cancel) o.blocks[it.blockid].iblank[it.pointid] = -1;
}

// MeshBlock::reduce_fringes (bookkeeping.cpp:572-652)
void reduce_fringes(Block &b) {
  b.iblank_reduced.assign(b.nnodes, 0);
  for (int i = 0; i < b.nnodes; i++) b.iblank_reduced[i] = b.iblank[i] > 0 ? b.iblank[i] : 0;
  std::vector<int> tmp = b.iblank_reduced;
  for (int iter = 0; iter < b.nfringe + 1; iter++) {
    for (size_t n = 0; n < b.nc.size(); n++) {
      const int nvert = b.nv[n];
      for (int i = 0; i < b.nc[n]; i++) {
        const int *cv = &b.vconn[n][(size_t)nvert * i];
        int ncount = 0;
        for (int m = 0; m < nvert; m++)
          ncount += (b.iblank_reduced[cv[m] - BASE] == 1 || b.iblank_reduced[cv[m] - BASE] < 0);
        if (ncount > 0 && ncount < nvert)
          for (int m = 0; m < nvert; m++) {
            const int v = cv[m] - BASE;
            if (tmp[v] == 0 && b.iblank[v] < 0) tmp[v] = b.iblank[v];
          }
      }
    }
    b.iblank_reduced = tmp;
  }
  b.have_reduced = true;
  b.cancelArray.clear();
  for (int i = 0; i < b.nnodes; i++)
    if (b.iblank[i] < 0 && b.iblank_reduced[i] == 0) b.cancelArray.push_back(i);
}

// MeshBlock::getCellIblanks (bookkeeping.cpp:654-705)
void cell_iblanks(Block &b) {
  b.iblank_cell.assign(b.ncells, 1);
  int icell = 0;
  for (size_t n = 0; n < b.nc.size(); n++) {
    const int nvert = b.nv[n];
    for (int i = 0; i < b.nc[n]; i++) {
      const int *cv = &b.vconn[n][(size_t)nvert * i];
      bool flag = true;
      int ncount = 0;
      for (int m = 0; m < nvert && flag; m++) {
        if (b.iblank[cv[m] - BASE] == 0) {
          b.iblank_cell[icell] = 0;
          flag = false;
        }
        ncount = ncount + (b.iblank[cv[m] - BASE] == -1);
      }
      if (flag && ncount == nvert) b.iblank_cell[icell] = -1;
      icell++;
    }
  }
}

int connect(Overset &o) {
  const int nb = (int)o.blocks.size();
  for (Block &b : o.blocks) preprocess(b);
  o.holemap.assign(nb, HoleMap());
  for (int i = 0; i < nb; i++) hole_map(o.blocks[i], o.holemap[i]);
  // exchangeBoxes: intersecting pairs (outer loop over the remote boxes)
  struct Pair {
    int ib, ob;
    Obb box;
  };
  std::vector<Pair> pairs;
  for (int ob = 0; ob < nb; ob++)
    for (int ib = 0; ib < nb; ib++) {
      if (o.blocks[ob].tag == o.blocks[ib].tag) continue;
      if (obb_intersect(o.blocks[ib].obb, o.blocks[ob].obb) == 1 ||
          obb_intersect(o.blocks[ob].obb, o.blocks[ib].obb) == 1)
        pairs.push_back({ib, ob, Obb()});
    }
  // each pair's box: the remote block's cells that overlap the local box,
  // in the remote frame (the reduced box the remote side sends)
  for (Pair &p : pairs) reduced_obb(o.blocks[p.ob], o.blocks[p.ib].obb, p.box);
  // exchangeSearchData: the local block's nodes inside the box become
  // search points of the remote block, in pair order
  for (Block &b : o.blocks) {
    b.xsearch.clear();
    b.res_search.clear();
    b.isearch.clear();
    b.tagsearch.clear();
  }
  std::vector<int> nodes;
  for (const Pair &p : pairs) {
    const Block &src = o.blocks[p.ib];
    Block &dst = o.blocks[p.ob];
    query_points(src, p.box, nodes);
    for (int v : nodes) {
      dst.isearch.insert(dst.isearch.end(), {0, v, p.ib});
      dst.tagsearch.push_back(src.tag);
      dst.xsearch.push_back(src.x[D * (size_t)v]);
      dst.xsearch.push_back(src.x[D * (size_t)v + 1]);
      dst.res_search.push_back(src.nodeRes[v]);
    }
  }
  for (Block &b : o.blocks) search(b);
  // exchangeDonors
  for (Block &b : o.blocks) b.donorList.assign(b.nnodes, {});
  for (Block &b : o.blocks)
    for (int i = 0; i < b.nsearch(); i++) {
      if (b.donorId[i] < 0) continue;
      Block &r = o.blocks[b.isearch[3 * (size_t)i + 2]];
      Donor d{b.tag, i, b.cellRes[b.donorId[i]], b.res_search[i]};
      auto &L = r.donorList[b.isearch[3 * (size_t)i + 1]];
      auto at = L.begin();
      while (at != L.end() && !(std::fabs(at->donorRes) > d.donorRes)) ++at;  // insertInList
      L.insert(at, d);
    }
  std::vector<std::vector<std::array<double, 3>>> recs(nb);
  for (int i = 0; i < nb; i++) process_donors(o.blocks[i], o.holemap, recs[i]);
  std::vector<int> ninterp(nb, 0);
  for (int i = 0; i < nb; i++)
    for (const auto &r : recs[i]) ninterp[(int)r[0] - BASE]++;
  for (int i = 0; i < nb; i++) {
    Block &b = o.blocks[i];
    b.interp.assign(ninterp[i], Interp());
    b.cancelList.clear();
    b.interp2donor.assign(b.nsearch(), -1);
    ninterp[i] = 0;
  }
  for (int i = 0; i < nb; i++)
    for (const auto &r : recs[i]) {
      const int ib = (int)r[0] - BASE;
      int rc = find_interp(o.blocks[ib], ninterp[ib], (int)r[1], r[2]);
      if (rc == -2) return set_error(USG_ERR_NUMERIC, "donor weight outside [-0.2, 1.2] (Frac Problem Again?)");
      if (rc) return set_error(USG_ERR_MESH, "degenerate donor cell");
    }
  for (int i = 0; i < nb; i++) o.blocks[i].interp.resize(ninterp[i]);
  for (Block &b : o.blocks) cancel_nodes(o, b, b.cancelList);
  apply_interp(o);
  // tioga::reduce_fringes
  for (Block &b : o.blocks) reduce_fringes(b);
  for (Block &b : o.blocks) cancel_nodes(o, b, b.cancelArray);
  apply_interp(o);
  for (Block &b : o.blocks) cell_iblanks(b);
  return USG_OK;
}

}  // namespace

struct usg_overset {
  Overset o;
};

extern "C" {

int usg_overset_create(const usg_mesh_desc *meshes, const int *const *overset_faces, int nblocks, int mexclude,
                       int nfringe, usg_overset **out) {
  if (!meshes || !out || nblocks < 1) return set_error(USG_ERR_ARG, "null argument");
  if (mexclude < 0 || nfringe < 0) return set_error(USG_ERR_ARG, "negative mexclude or nfringe");
  *out = nullptr;
  std::unique_ptr<usg_overset> h(new usg_overset);
  for (int ib = 0; ib < nblocks; ib++) {
    const usg_mesh_desc &m = meshes[ib];
    if (!m.xy || !m.fnode || !m.flc || !m.frc) return set_error(USG_ERR_ARG, "null mesh array");
    Block b;
    b.tag = ib + 1;
    b.mexclude = mexclude;
    b.nfringe = nfringe;
    b.nnodes = m.nn;
    b.ncells = m.nc;
    b.x.assign(m.xy, m.xy + 2 * (size_t)m.nn);
    std::vector<int> cnode, cnv;
    std::string e = usg::cell_vertices(m.nn, m.nf, m.fnode, m.flc, m.frc, m.nc, cnode, cnv);
    if (!e.empty()) return set_error(USG_ERR_MESH, e);
    // cell types in ascending vertex count, cells in input order inside a
    // type (CobaltReader.cpp:1025-1104)
    for (int nvert = 3; nvert <= 4; nvert++) {
      std::vector<int> conn;
      int count = 0;
      for (int c = 0; c < m.nc; c++) {
        if (cnv[c] != nvert) continue;
        for (int k = 0; k < nvert; k++) conn.push_back(cnode[4 * (size_t)c + k] + BASE);
        b.cell_of.push_back(c);
        count++;
      }
      if (count) {
        b.nv.push_back(nvert);
        b.nc.push_back(count);
        b.vconn.push_back(std::move(conn));
      }
    }
    std::vector<char> w(m.nn, 0), ov(m.nn, 0);
    const int *of = overset_faces ? overset_faces[ib] : nullptr;
    for (int f = 0; f < m.nf; f++) {
      const bool isw = m.frc[f] == USG_WALL, iso = of && of[f];
      for (int j = 0; j < 2; j++) {
        if (isw) w[m.fnode[2 * f + j]] = 1;
        if (iso) ov[m.fnode[2 * f + j]] = 1;
      }
    }
    for (int v = 0; v < m.nn; v++) {
      if (w[v]) b.wbc.push_back(v + BASE);
      if (ov[v]) b.obc.push_back(v + BASE);
    }
    h->o.blocks.push_back(std::move(b));
  }
  int rc = connect(h->o);
  if (rc) return rc;
  *out = h.release();
  return USG_OK;
}

int usg_overset_destroy(usg_overset *o) {
  delete o;
  return USG_OK;
}

int usg_overset_iblank(usg_overset *o, int block, int *node_iblank, int *cell_iblank) {
  if (!o || block < 0 || block >= (int)o->o.blocks.size()) return set_error(USG_ERR_ARG, "bad block");
  const Block &b = o->o.blocks[block];
  if (node_iblank) std::copy(b.iblank.begin(), b.iblank.end(), node_iblank);
  if (cell_iblank)
    for (int k = 0; k < b.ncells; k++) cell_iblank[b.cell_of[k]] = b.iblank_cell[k];
  return USG_OK;
}

int usg_overset_receptors(usg_overset *o, int block, int *n, int *recv_block, int *recv_node, int *donor_cell,
                          int *nw, int *donor_node, double *w) {
  if (!o || !n || block < 0 || block >= (int)o->o.blocks.size()) return set_error(USG_ERR_ARG, "bad block");
  const Block &b = o->o.blocks[block];
  int k = 0;
  for (const Interp &it : b.interp) {
    if (it.cancel) continue;
    if (recv_block) recv_block[k] = it.blockid;
    if (recv_node) recv_node[k] = it.pointid;
    if (donor_cell) donor_cell[k] = b.cell_of[it.donor_cell];
    if (nw) nw[k] = it.nweights;
    for (int m = 0; m < 4; m++) {
      if (donor_node) donor_node[4 * k + m] = m < it.nweights ? it.inode[m] : it.inode[0];
      if (w) w[4 * k + m] = m < it.nweights ? it.w[m] : 0.0;
    }
    k++;
  }
  *n = k;
  return USG_OK;
}

int usg_overset_median(int *ix, double *x, int n) {
  if (n < 0 || (n > 0 && (!ix || !x))) return set_error(USG_ERR_ARG, "null argument");
  find_median_wrap(ix, x, n);
  return USG_OK;
}

int usg_overset_obb(const double *x, int n, double *xc, double *dxc, double *vec) {
  if (n < 1 || !x || !xc || !dxc || !vec) return set_error(USG_ERR_ARG, "null argument");
  Obb o;
  find_obb(x, n, o);
  for (int i = 0; i < 2; i++) {
    xc[i] = o.xc[i];
    dxc[i] = o.dxc[i];
    for (int j = 0; j < 2; j++) vec[2 * i + j] = o.vec[i][j];
  }
  return USG_OK;
}

int usg_overset_resolution(usg_overset *o, int block, double *node_res, double *cell_res) {
  if (!o || block < 0 || block >= (int)o->o.blocks.size()) return set_error(USG_ERR_ARG, "bad block");
  const Block &b = o->o.blocks[block];
  if (node_res) std::copy(b.nodeRes.begin(), b.nodeRes.end(), node_res);
  if (cell_res)
    for (int k = 0; k < b.ncells; k++) cell_res[b.cell_of[k]] = b.cellRes[k];
  return USG_OK;
}

}  // extern "C"
