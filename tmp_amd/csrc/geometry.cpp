// Host geometry precompute, restated from the reference's serial reader and
// geometry types so that every derived value has the reference's bits:
//   cell vertex order   UniTioga/CellBrick.cpp:13-136 (Brick4Cobalt2D::insert)
//   per-cell face list  reader/UnstructReader.h:233-247 (push_edge in face order)
//   cell volume         mathTools/cellVolume.cpp:6-78
//   cell centre         toolBox/cell3d_int.h:100-108 (vertex mean)
//   face n, A, centre   toolBox/edge3d_int.h:106-115,200-239
//   wall distance       UnstructSolver/UnstructBlock.h:265-298 (projected)
#include "geometry.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace usg {
namespace {

struct Brick {
  int nodes[4] = {-1, -1, -1, -1};
  int size = 0;
  bool closed = false;
  void insert(const int *start) {
    if (closed) return;
    if (size == 0) {
      nodes[0] = start[0];
      nodes[1] = start[1];
      size = 2;
      return;
    }
    int idx[8];
    int k = 0;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 2; j++)
        if (nodes[i] == start[j]) {
          idx[k++] = i;
          idx[k++] = j;
        }
    if (k == 0) return;  // opposite side of a quadrilateral
    if (k == 2) {
      if (idx[0] < 2)
        nodes[3 - idx[0]] = start[1 - idx[1]];
      else
        nodes[5 - idx[0]] = start[1 - idx[1]];
      size += 1;
      if (size == 4) closed = true;
      return;
    }
    if (size == 3 && nodes[2] == -1 && nodes[3] != -1) {
      nodes[2] = nodes[3];
      nodes[3] = -1;
    }
    closed = true;
  }
};

double tri_area(const double *p1, const double *p2, const double *p3) {
  double a = 0.5 * ((p1[0] - p2[0]) * (p1[1] + p2[1]) + (p2[0] - p3[0]) * (p2[1] + p3[1]) +
                    (p3[0] - p1[0]) * (p3[1] + p1[1]));
  return a < 0 ? -a : a;
}

double cross_z(double ax, double ay, double bx, double by) { return ax * by - ay * bx; }

double volume(double xv[4][2], int nv) {
  if (nv == 3) return tri_area(xv[0], xv[1], xv[2]);
  if (nv != 4) return -1;
  double ABx = xv[1][0] - xv[0][0], ABy = xv[1][1] - xv[0][1];
  double ACx = xv[2][0] - xv[0][0], ACy = xv[2][1] - xv[0][1];
  double ADx = xv[3][0] - xv[0][0], ADy = xv[3][1] - xv[0][1];
  double ABcAC = cross_z(ABx, ABy, ACx, ACy);
  double ABcAD = cross_z(ABx, ABy, ADx, ADy);
  double area = 0;
  if (ABcAC * ABcAD > 0) {
    double ACcAD = cross_z(ACx, ACy, ADx, ADy);
    if (ACcAD * ABcAD > 0) {
      area += tri_area(xv[0], xv[2], xv[1]);
      area += tri_area(xv[0], xv[2], xv[3]);
    } else {
      area += tri_area(xv[0], xv[3], xv[1]);
      area += tri_area(xv[0], xv[3], xv[2]);
    }
  } else {
    area += tri_area(xv[0], xv[1], xv[2]);
    area += tri_area(xv[0], xv[1], xv[3]);
  }
  return area;
}

inline double dot2(double a0, double a1, double b0, double b1) {
  double r = 0;
  r += a0 * b0;
  r += a1 * b1;
  return r;
}

void waves(HostMesh &m) {
  const int nc = m.nc;
  auto other = [&](int k, int face) { return m.flc[face] == k ? m.frc[face] : m.flc[face]; };
  std::vector<int> lev(nc, 0);
  int nl = 0;
  for (int k = 0; k < nc; k++) {
    int lv = 0;
    for (int f = 0; f < m.cnf[k]; f++) {
      int o = other(k, m.cface[4 * k + f]);
      if (o >= 0 && o < k) lv = std::max(lv, lev[o] + 1);
    }
    lev[k] = lv;
    nl = std::max(nl, lv + 1);
  }
  m.fwd_off.assign(nl + 1, 0);
  for (int k = 0; k < nc; k++) m.fwd_off[lev[k] + 1]++;
  for (int l = 0; l < nl; l++) m.fwd_off[l + 1] += m.fwd_off[l];
  m.fwd_cells.assign(nc, 0);
  {
    std::vector<int> pos(m.fwd_off.begin(), m.fwd_off.end() - 1);
    for (int k = 0; k < nc; k++) m.fwd_cells[pos[lev[k]]++] = k;
  }
  std::fill(lev.begin(), lev.end(), 0);
  nl = 0;
  for (int k = nc - 1; k >= 0; k--) {
    int lv = 0;
    for (int f = 0; f < m.cnf[k]; f++) {
      int o = other(k, m.cface[4 * k + f]);
      if (o > k) lv = std::max(lv, lev[o] + 1);
    }
    lev[k] = lv;
    nl = std::max(nl, lv + 1);
  }
  m.bwd_off.assign(nl + 1, 0);
  for (int k = 0; k < nc; k++) m.bwd_off[lev[k] + 1]++;
  for (int l = 0; l < nl; l++) m.bwd_off[l + 1] += m.bwd_off[l];
  m.bwd_cells.assign(nc, 0);
  {
    std::vector<int> pos(m.bwd_off.begin(), m.bwd_off.end() - 1);
    for (int k = 0; k < nc; k++) m.bwd_cells[pos[lev[k]]++] = k;
  }
  m.ranges.assign(1, 0);
  for (int k = 0; k < nc; k++) {
    int s0 = m.ranges.back();
    for (int f = 0; f < m.cnf[k]; f++) {
      int o = other(k, m.cface[4 * k + f]);
      if (o >= s0 && o < k) {
        m.ranges.push_back(k);
        break;
      }
    }
  }
  m.ranges.push_back(nc);
  m.use_ranges = m.ranges.size() - 1 == m.fwd_off.size() - 1;
}

uint64_t hilbert(uint32_t x, uint32_t y, int order) {
  uint64_t d = 0;
  for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
    uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
    d += (uint64_t)s * s * ((3 * rx) ^ ry);
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

// Group `items` (ascending) by the part of each item, keeping the order.
void by_part(const std::vector<int> &items, const std::vector<int> &item_part, std::vector<int> &out,
             int *off) {
  int cnt[kParts + 1] = {0};
  for (int it : items) cnt[item_part[it] + 1]++;
  for (int p = 0; p < kParts; p++) cnt[p + 1] += cnt[p];
  for (int p = 0; p <= kParts; p++) off[p] = cnt[p];
  std::vector<int> pos(cnt, cnt + kParts);
  out.resize(items.size());
  for (int it : items) out[pos[item_part[it]]++] = it;
}

void xcd_lists(HostMesh &m) {
  const int nc = m.nc, nf = m.nf;
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
  for (int i = 0; i < nc; i++) {
    xmin = std::min(xmin, m.cc[2 * i]);
    xmax = std::max(xmax, m.cc[2 * i]);
    ymin = std::min(ymin, m.cc[2 * i + 1]);
    ymax = std::max(ymax, m.cc[2 * i + 1]);
  }
  double span = std::max(std::max(xmax - xmin, ymax - ymin), 1e-300);
  const int order = 16;
  const double n = (double)((1u << order) - 1);
  std::vector<std::pair<uint64_t, int>> key(nc);
  for (int i = 0; i < nc; i++) {
    uint32_t xi = (uint32_t)std::min(n, std::max(0.0, (m.cc[2 * i] - xmin) / span * n));
    uint32_t yi = (uint32_t)std::min(n, std::max(0.0, (m.cc[2 * i + 1] - ymin) / span * n));
    key[i] = {hilbert(xi, yi, order), i};
  }
  std::stable_sort(key.begin(), key.end());
  m.part.assign(nc, 0);
  for (int r = 0; r < nc; r++) m.part[key[r].second] = (int)((long long)r * kParts / nc);
  // cell kernels (gradient, cell gather) walk the cells in Hilbert order of
  // their centres, across colours: a wave's own cells and its neighbours are
  // then the same few cache lines (the numbering stays colour-blocked)
  std::vector<int> all(nc);
  for (int r = 0; r < nc; r++) all[r] = key[r].second;
  m.cell_off.assign(kParts + 1, 0);
  by_part(all, m.part, m.cell_list, m.cell_off.data());
  std::vector<int> fpart(nf), fall(nf);
  for (int f = 0; f < nf; f++) {
    fpart[f] = m.part[m.flc[f]];
    fall[f] = f;
  }
  m.face_off.assign(kParts + 1, 0);
  by_part(fall, fpart, m.face_list, m.face_off.data());
  auto waves_by_part = [&](const std::vector<int> &cells, const std::vector<int> &woff, std::vector<int> &list,
                           std::vector<int> &poff) {
    const int nw = (int)woff.size() - 1;
    list.assign(cells.size(), 0);
    poff.assign((size_t)nw * (kParts + 1), 0);
    for (int w = 0; w < nw; w++) {
      std::vector<int> sub(cells.begin() + woff[w], cells.begin() + woff[w + 1]), out;
      int off[kParts + 1];
      by_part(sub, m.part, out, off);
      std::copy(out.begin(), out.end(), list.begin() + woff[w]);
      for (int p = 0; p <= kParts; p++) poff[(size_t)w * (kParts + 1) + p] = woff[w] + off[p];
    }
  };
  if (m.use_ranges) {
    std::vector<int> ident(nc);
    for (int i = 0; i < nc; i++) ident[i] = i;
    waves_by_part(ident, m.ranges, m.fwd_list, m.fwd_poff);
    m.bwd_list = m.fwd_list;
    // backward: the same ranges in reverse order
    const int nw = (int)m.ranges.size() - 1;
    m.bwd_poff.assign(m.fwd_poff.size(), 0);
    for (int w = 0; w < nw; w++)
      for (int p = 0; p <= kParts; p++)
        m.bwd_poff[(size_t)w * (kParts + 1) + p] = m.fwd_poff[(size_t)(nw - 1 - w) * (kParts + 1) + p];
  } else {
    waves_by_part(m.fwd_cells, m.fwd_off, m.fwd_list, m.fwd_poff);
    waves_by_part(m.bwd_cells, m.bwd_off, m.bwd_list, m.bwd_poff);
  }
}

}  // namespace

double cell_volume(double xv[4][2], int nv) { return volume(xv, nv); }

std::string cell_vertices(int nn, int nf, const int *fnode, const int *flc, const int *frc, int nc,
                          std::vector<int> &cnode, std::vector<int> &cnv) {
  if (nn <= 0 || nf <= 0 || nc <= 0 || !fnode || !flc || !frc) return "empty mesh";
  for (int k = 0; k < nf; k++) {
    if (flc[k] < 0 || flc[k] >= nc) return "Left cell can't be negative";
    if (frc[k] >= nc || (frc[k] < 0 && frc[k] != kWall && frc[k] != kFarfield))
      return "Not a valid boundary type";
    if (fnode[2 * k] < 0 || fnode[2 * k] >= nn || fnode[2 * k + 1] < 0 || fnode[2 * k + 1] >= nn)
      return "face node out of range";
  }
  std::vector<Brick> br(nc);
  for (int k = 0; k < nf; k++) {
    br[flc[k]].insert(&fnode[2 * k]);
    if (frc[k] >= 0) br[frc[k]].insert(&fnode[2 * k]);
  }
  cnode.assign(4 * (size_t)nc, -1);
  cnv.assign(nc, 0);
  for (int i = 0; i < nc; i++) {
    int nv = br[i].size;
    if (nv < 3 || nv > 4) return "cell is not a triangle or quadrilateral";
    cnv[i] = nv;
    for (int v = 0; v < nv; v++) cnode[4 * i + v] = br[i].nodes[v];
  }
  return "";
}

void freestream_state(double mach, double aoa_deg, double fs[4], double *aoa_rad) {
  const double Gamma = 1.4, PI = 3.141593;
  double fs_AOA = PI * aoa_deg / 180;
  double fs_density = 1.225, fs_pressure = 101325;
  double fs_soundSpeed = std::sqrt(Gamma * fs_pressure / fs_density);
  double fs_velocity_magnitude = fs_soundSpeed * mach;
  double u = fs_velocity_magnitude * std::cos(fs_AOA);
  double v = fs_velocity_magnitude * std::sin(fs_AOA);
  fs[0] = fs_density / fs_density;
  fs[1] = fs_pressure / fs_pressure;
  fs[2] = u / (fs_soundSpeed / std::sqrt(Gamma));
  fs[3] = v / (fs_soundSpeed / std::sqrt(Gamma));
  if (aoa_rad) *aoa_rad = fs_AOA;
}

std::string build_mesh(HostMesh &m, int nn, const double *xy, int nf, const int *fnode,
                       const int *flc, const int *frc, int nc) {
  if (nn <= 0 || nf <= 0 || nc <= 0 || !xy || !fnode || !flc || !frc) return "empty mesh";
  m.nn = nn;
  m.nf = nf;
  m.nc = nc;
  m.xy.assign(xy, xy + 2 * (size_t)nn);
  m.fnode.assign(fnode, fnode + 2 * (size_t)nf);
  m.flc.assign(flc, flc + nf);
  m.frc.assign(frc, frc + nf);
  std::string e = cell_vertices(nn, nf, fnode, flc, frc, nc, m.cnode, m.cnv);
  if (!e.empty()) return e;
  m.cface.assign(4 * (size_t)nc, kNoFace);
  m.cnf.assign(nc, 0);
  for (int k = 0; k < nf; k++) {
    int l = flc[k], r = frc[k];
    if (m.cnf[l] >= 4 || (r >= 0 && m.cnf[r] >= 4)) return "cell with more than 4 faces";
    m.cface[4 * l + m.cnf[l]++] = k;
    if (r >= 0) m.cface[4 * r + m.cnf[r]++] = k;
  }
  m.vol.assign(nc, 0.0);
  m.cc.assign(2 * (size_t)nc, 0.0);
  for (int i = 0; i < nc; i++) {
    int nv = m.cnv[i];
    if (m.cnf[i] != nv) return "cell is not a triangle or quadrilateral";
    double xv[4][2];
    for (int v = 0; v < nv; v++) {
      xv[v][0] = xy[2 * m.cnode[4 * i + v]];
      xv[v][1] = xy[2 * m.cnode[4 * i + v] + 1];
    }
    m.vol[i] = volume(xv, nv);
    double sx = 0, sy = 0;
    for (int v = 0; v < nv; v++) {
      sx = sx + xv[v][0];
      sy = sy + xv[v][1];
    }
    m.cc[2 * i] = sx / nv;
    m.cc[2 * i + 1] = sy / nv;
  }
  m.fn.assign(2 * (size_t)nf, 0.0);
  m.fa.assign(nf, 0.0);
  m.fc.assign(2 * (size_t)nf, 0.0);
  m.wall_faces.clear();
  for (int k = 0; k < nf; k++) {
    const double *p0 = &xy[2 * fnode[2 * k]];
    const double *p1 = &xy[2 * fnode[2 * k + 1]];
    double dx = p1[0] - p0[0], dy = p1[1] - p0[1];
    double L2 = 0;
    L2 += dx * dx;
    L2 += dy * dy;
    L2 = std::sqrt(L2);
    m.fa[k] = L2;
    m.fn[2 * k] = dy / L2;
    m.fn[2 * k + 1] = -dx / L2;
    double sx = 0, sy = 0;
    sx = sx + p0[0];
    sy = sy + p0[1];
    sx = sx + p1[0];
    sy = sy + p1[1];
    m.fc[2 * k] = sx / 2;
    m.fc[2 * k + 1] = sy / 2;
    if (frc[k] == kWall) m.wall_faces.push_back(k);
  }
  m.wdist.assign(nc, 0.0);
  for (int i = 0; i < nc; i++) {
    double d = 100000000000.0;
    for (int k : m.wall_faces) {
      double cx = m.cc[2 * i] - m.fc[2 * k];
      double cy = m.cc[2 * i + 1] - m.fc[2 * k + 1];
      double dist = std::fabs(dot2(cx, cy, m.fn[2 * k], m.fn[2 * k + 1]));
      d = (dist < d) ? dist : d;
    }
    m.wdist[i] = d;
  }
  waves(m);
  xcd_lists(m);
  m.case_cell = {0, nc};
  m.case_face = {0, nf};
  m.case_wall = {0, (int)m.wall_faces.size()};
  return std::string();
}

void merge_meshes(const std::vector<HostMesh> &cases, HostMesh &u) {
  u = HostMesh();
  const int nk = (int)cases.size();
  u.case_cell.assign(nk + 1, 0);
  u.case_face.assign(nk + 1, 0);
  u.case_wall.assign(nk + 1, 0);
  std::vector<int> noff(nk + 1, 0);
  for (int i = 0; i < nk; i++) {
    noff[i + 1] = noff[i] + cases[i].nn;
    u.case_cell[i + 1] = u.case_cell[i] + cases[i].nc;
    u.case_face[i + 1] = u.case_face[i] + cases[i].nf;
    u.case_wall[i + 1] = u.case_wall[i] + (int)cases[i].wall_faces.size();
  }
  u.nn = noff[nk];
  u.nc = u.case_cell[nk];
  u.nf = u.case_face[nk];
  auto shift = [](std::vector<int> &dst, const std::vector<int> &src, int off, bool keep_negative) {
    for (int v : src) dst.push_back(keep_negative && v < 0 ? v : v + off);
  };
  for (int i = 0; i < nk; i++) {
    const HostMesh &c = cases[i];
    const int co = u.case_cell[i], fo = u.case_face[i];
    u.xy.insert(u.xy.end(), c.xy.begin(), c.xy.end());
    shift(u.fnode, c.fnode, noff[i], false);
    shift(u.flc, c.flc, co, false);
    shift(u.frc, c.frc, co, true);
    for (int v : c.cnode) u.cnode.push_back(v < 0 ? v : v + noff[i]);
    u.cnv.insert(u.cnv.end(), c.cnv.begin(), c.cnv.end());
    for (int v : c.cface) u.cface.push_back(v == kNoFace ? v : v + fo);
    u.cnf.insert(u.cnf.end(), c.cnf.begin(), c.cnf.end());
    u.vol.insert(u.vol.end(), c.vol.begin(), c.vol.end());
    u.cc.insert(u.cc.end(), c.cc.begin(), c.cc.end());
    u.fn.insert(u.fn.end(), c.fn.begin(), c.fn.end());
    u.fa.insert(u.fa.end(), c.fa.begin(), c.fa.end());
    u.fc.insert(u.fc.end(), c.fc.begin(), c.fc.end());
    u.wdist.insert(u.wdist.end(), c.wdist.begin(), c.wdist.end());
    shift(u.wall_faces, c.wall_faces, fo, false);
    u.part.insert(u.part.end(), c.part.begin(), c.part.end());
  }
  // part lists: part p of the union = the cases' part p lists in case order
  auto merge_parts = [&](auto get_list, auto get_off, const std::vector<int> &base, std::vector<int> &list,
                         std::vector<int> &off) {
    list.clear();
    off.assign(kParts + 1, 0);
    for (int p = 0; p < kParts; p++) {
      for (int i = 0; i < nk; i++) {
        const std::vector<int> &l = get_list(cases[i]);
        const std::vector<int> &o = get_off(cases[i]);
        for (int j = o[p]; j < o[p + 1]; j++) list.push_back(l[j] + base[i]);
      }
      off[p + 1] = (int)list.size();
    }
  };
  merge_parts([](const HostMesh &c) -> const std::vector<int> & { return c.cell_list; },
              [](const HostMesh &c) -> const std::vector<int> & { return c.cell_off; }, u.case_cell, u.cell_list,
              u.cell_off);
  merge_parts([](const HostMesh &c) -> const std::vector<int> & { return c.face_list; },
              [](const HostMesh &c) -> const std::vector<int> & { return c.face_off; }, u.case_face, u.face_list,
              u.face_off);
  // waves: wave w, part p of the union = the cases' (w, p) slices in case order
  auto merge_waves = [&](bool fwd, std::vector<int> &list, std::vector<int> &poff) {
    int nw = 0;
    for (const HostMesh &c : cases)
      nw = std::max(nw, (int)(fwd ? c.fwd_poff : c.bwd_poff).size() / (kParts + 1));
    list.clear();
    poff.assign((size_t)nw * (kParts + 1), 0);
    for (int w = 0; w < nw; w++) {
      for (int p = 0; p < kParts; p++) {
        poff[(size_t)w * (kParts + 1) + p] = (int)list.size();
        for (int i = 0; i < nk; i++) {
          const HostMesh &c = cases[i];
          const std::vector<int> &l = fwd ? c.fwd_list : c.bwd_list;
          const std::vector<int> &o = fwd ? c.fwd_poff : c.bwd_poff;
          if ((size_t)(w + 1) * (kParts + 1) > o.size()) continue;
          for (int j = o[(size_t)w * (kParts + 1) + p]; j < o[(size_t)w * (kParts + 1) + p + 1]; j++)
            list.push_back(l[j] + u.case_cell[i]);
        }
      }
      poff[(size_t)w * (kParts + 1) + kParts] = (int)list.size();
    }
  };
  merge_waves(true, u.fwd_list, u.fwd_poff);
  merge_waves(false, u.bwd_list, u.bwd_poff);
  u.use_ranges = false;
}

}  // namespace usg
