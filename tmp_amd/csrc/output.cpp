// Host-side outputs of a solution: the Tecplot cell-data file of
// TopologyHolderStrategy::writeCellData (TopologyHolderStrategyOutput.cpp:
// 466-554) and the wall pressure coefficient of Euler2D::outPutCp
// (Euler2D.cpp:392-623; ViscousFlow2D::outPutCp is the same routine).
// Numbers go through std::ostream's default formatting (6 significant
// digits, %g style), as in the reference, so the text matches byte for
// byte.  Plain C++17, no HIP.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <set>
#include <string>
#include <vector>

#include "../../include/usg.h"
#include "geometry.h"

using usg::set_error;

namespace {

bool mesh_args(const usg_mesh_desc *m) {
  return m && m->xy && m->fnode && m->flc && m->frc && m->nn > 0 && m->nf > 0 && m->nc > 0;
}

std::string vertices(const usg_mesh_desc *m, std::vector<int> &cnode, std::vector<int> &cnv) {
  return usg::cell_vertices(m->nn, m->nf, m->fnode, m->flc, m->frc, m->nc, cnode, cnv);
}

// One wall node of outPutCp: position, signed cell count (negative when an
// adjacent wall face has a normal with a positive y component) and C_p.
struct WallNode {
  double x, y, count, cp;
};

// rankWithXY (Euler2D.cpp:39-70): non-negative counts first in ascending
// x, then negative counts in descending x.  The reference's comparator is
// not a strict ordering for equal x; here equal keys keep ascending node id.
bool before(const WallNode &a, const WallNode &b) {
  bool an = a.count < 0, bn = b.count < 0;
  if (an != bn) return !an;
  return an ? a.x > b.x : a.x < b.x;
}

}  // namespace

extern "C" {

int usg_mesh_cells(const usg_mesh_desc *mesh, int *cnode, int *cnv) {
  if (!mesh_args(mesh) || !cnode || !cnv) return set_error(USG_ERR_ARG, "null mesh or output");
  std::vector<int> cn, nv;
  std::string e = vertices(mesh, cn, nv);
  if (!e.empty()) return set_error(USG_ERR_MESH, e);
  std::copy(cn.begin(), cn.end(), cnode);
  std::copy(nv.begin(), nv.end(), cnv);
  return USG_OK;
}

int usg_freestream_state(double mach, double aoa_deg, double *fs) {
  if (!fs) return set_error(USG_ERR_ARG, "null output");
  usg::freestream_state(mach, aoa_deg, fs, nullptr);
  return USG_OK;
}

int usg_write_cell_data(const usg_mesh_desc *mesh, const double *U, const double *fs, const char *filename,
                        int proc, int mesh_tag) {
  if (!mesh_args(mesh) || !U || !fs || !filename) return set_error(USG_ERR_ARG, "null argument");
  std::vector<int> cnode, cnv;
  std::string e = vertices(mesh, cnode, cnv);
  if (!e.empty()) return set_error(USG_ERR_MESH, e);
  const int nn = mesh->nn, nc = mesh->nc, nequ = 4, dim = 2;
  std::string name = std::string(filename) + "_" + std::to_string(mesh_tag + 1) + "_" + std::to_string(proc + 1);
  std::ofstream out(name, std::ios::out);
  if (!out.is_open()) return set_error(USG_ERR_ARG, "cannot open " + name);
  out << "TITLE =\"Tioga output\"\n";
  out << "VARIABLES=\"X\",\"Y\",";
  out << "\"Density\",\"Pressure\",\"U\",\"V\",\n";
  out << "\"Cp\",\"Ma\"";  // no line break before ZONE, as written by the reference
  out << "ZONE T=\"VOL_MIXED\",N=" << nn << " E=" << nc;
  out << " ET = QUADRILATERAL, F = FEBLOCK\n";
  out << "VARLOCATION =  (";
  for (int i = 0; i < dim; i++) out << i + 1 << "=NODAL,";
  for (int i = 0; i < nequ + 1; i++) out << dim + 1 + i << "=CELLCENTERED,";
  out << dim + nequ + 2 << "=CELLCENTERED)\n";
  for (int j = 0; j < dim; j++)
    for (int i = 0; i < nn; i++) out << mesh->xy[2 * (size_t)i + j] << '\n';
  for (int q = 0; q < nequ; q++)
    for (int j = 0; j < nc; j++) out << U[(size_t)nequ * j + q] << '\n';
  double vmag2 = (fs[2] * fs[2] + fs[3] * fs[3]);
  for (int j = 0; j < nc; j++) out << 2 * (U[(size_t)nequ * j + 1] - fs[1]) / vmag2 << '\n';
  // Mach column: |u| / sqrt(p / rho) (no gamma), as the reference writes it
  for (int j = 0; j < nc; j++) {
    const double *u = U + (size_t)nequ * j;
    double c = std::sqrt(u[1] / u[0]);
    double vm = 0;
    for (int k = 0; k < dim; k++) vm += u[2 + k] * u[2 + k];
    vm = std::sqrt(vm);
    out << vm / c << '\n';
  }
  for (int i = 0; i < nc; i++) {
    int nv = cnv[i];
    for (int j = 0; j < nv; j++) out << cnode[4 * (size_t)i + j] + 1 << '\t';
    for (int j = 0; j < 4 - nv; j++) out << cnode[4 * (size_t)i + nv - 1] + 1 << '\t';
    out << '\n';
  }
  out.close();
  if (!out) return set_error(USG_ERR_ARG, "write failed: " + name);
  return USG_OK;
}

int usg_surface_cp(const usg_mesh_desc *mesh, const double *U, const double *fs, int cap, double *table,
                   double *curve, int *nwall, int *ncurve) {
  if (!mesh_args(mesh) || !U || !fs || !nwall || !ncurve || cap < 0)
    return set_error(USG_ERR_ARG, "null argument");
  std::vector<int> cnode, cnv;
  std::string e = vertices(mesh, cnode, cnv);
  if (!e.empty()) return set_error(USG_ERR_MESH, e);
  const int nn = mesh->nn, nc = mesh->nc, nf = mesh->nf;
  // cell pressure averaged to the nodes, then C_p (Euler2D.cpp:396-423)
  std::vector<double> pav(nn, 0.0);
  std::vector<int> count(nn, 0);
  for (int i = 0; i < nc; i++)
    for (int k = 0; k < cnv[i]; k++) {
      int v = cnode[4 * (size_t)i + k];
      ++count[v];
      pav[v] += U[1 + 4 * (size_t)i];
    }
  double rhoV2 = fs[0] * (fs[2] * fs[2] + fs[3] * fs[3]);
  for (int i = 0; i < nn; i++) {
    pav[i] = pav[i] / count[i];
    pav[i] = (pav[i] - fs[1]) * 2 / rhoV2;
  }
  // wall nodes in ascending id; a node of a wall face whose normal points
  // up gets a negative count (Euler2D.cpp:425-444)
  std::set<int> wall;
  for (int k = 0; k < nf; k++) {
    if (mesh->frc[k] != USG_WALL) continue;
    const int *fv = mesh->fnode + 2 * (size_t)k;
    double dx = mesh->xy[2 * (size_t)fv[1]] - mesh->xy[2 * (size_t)fv[0]];
    double dy = mesh->xy[2 * (size_t)fv[1] + 1] - mesh->xy[2 * (size_t)fv[0] + 1];
    double len = std::sqrt(dx * dx + dy * dy);
    double ny = -dx / len;
    for (int j = 0; j < 2; j++) {
      if (ny > 0) count[fv[j]] = -std::abs(count[fv[j]]);
      wall.insert(fv[j]);
    }
  }
  std::vector<WallNode> rows;
  rows.reserve(wall.size());
  for (int v : wall) rows.push_back({mesh->xy[2 * (size_t)v], mesh->xy[2 * (size_t)v + 1], (double)count[v], pav[v]});
  std::stable_sort(rows.begin(), rows.end(), before);
  // merge neighbours closer than 1e-4 in x, keeping the C_p of the one with
  // the largest |count| at the first one's position; the last node is also
  // compared with the first (Euler2D.cpp:517-552)
  std::vector<WallNode> merged;
  const size_t S = rows.size();
  for (size_t i = 0; i < S; i++) {
    size_t cur = i;
    WallNode w = rows[cur];
    size_t next = (i + 1) % S;
    // the reference reads past the end when the run of equal x reaches it;
    // the run stops there instead
    while (next < S && !(std::fabs(rows[cur].x - rows[next].x) > 0.0001)) {
      if (std::fabs(rows[next].count) > std::fabs(rows[cur].count)) {
        w.cp = rows[next].cp;
        w.count = rows[next].count;
        cur = next;
      }
      next = next + 1;
      i++;
    }
    merged.push_back(w);
  }
  *nwall = (int)S;
  *ncurve = (int)merged.size();
  if ((table || curve) && (size_t)cap < S) return set_error(USG_ERR_ARG, "capacity below the wall node count");
  for (size_t i = 0; table && i < S; i++) {
    table[4 * i] = rows[i].x;
    table[4 * i + 1] = rows[i].y;
    table[4 * i + 2] = rows[i].count;
    table[4 * i + 3] = rows[i].cp;
  }
  for (size_t i = 0; curve && i < merged.size(); i++) {
    curve[4 * i] = merged[i].x;
    curve[4 * i + 1] = merged[i].y;
    curve[4 * i + 2] = merged[i].count;
    curve[4 * i + 3] = merged[i].cp;
  }
  return USG_OK;
}

int usg_write_cp(const char *path, const double *curve, int ncurve) {
  if (!path || !curve || ncurve <= 0) return set_error(USG_ERR_ARG, "null path or empty curve");
  std::ofstream out(path);
  if (!out.is_open()) return set_error(USG_ERR_ARG, std::string("cannot open ") + path);
  out << "Title = \"Dussin_exp\"\n";
  out << "VARIABLES = \"X\",\"Cp\"\n";
  out << "ZONE T=\"Dussin_exp\", I = " << ncurve + 1 << ", F=POINT\n";
  for (int i = 0; i < ncurve; i++) out << curve[4 * i] << "\t" << curve[4 * i + 3] << '\n';
  out << curve[0] << "\t" << curve[3] << '\n';
  out.close();
  if (!out) return set_error(USG_ERR_ARG, std::string("write failed: ") + path);
  return USG_OK;
}

}  // extern "C"
