// Donor weights of the overset interpolation stencil (host side): the
// receptor point's weights on the vertices of its donor cell, as
// computeNodalWeight2D (mathTools/math.cpp:229-312) computes them --
// barycentric coordinates for a triangle (2x2 elimination with row swaps,
// solvec, math.cpp:4-58), bilinear coordinates for a quadrilateral (vertex
// order fixed by findClockWise, math.cpp:189-227; Newton iteration from
// (0.5, 0.5) to |F| <= 1e-14 within 500 steps, newtonSolve2D,
// math.cpp:126-188; a failed solve gives u = 2, v = 0 as there).
#include <cmath>
#include <string>

#include "../../include/usg.h"
#include "geometry.h"

namespace {

// Gaussian elimination of a 2x2 system in place; false when no usable pivot
bool solve2(double a[2][2], double b[2]) {
  const double eps = 1e-8;
  for (int i = 0; i < 2; i++) {
    if (std::fabs(a[i][i]) < eps) {
      bool found = false;
      for (int k = i + 1; k < 2 && !found; k++) {
        if (!(std::fabs(a[k][i]) < eps)) {
          found = true;
          for (int l = 0; l < 2; l++) std::swap(a[k][l], a[i][l]);
          std::swap(b[k], b[i]);
        }
      }
      if (!found) return false;
    }
    for (int k = i + 1; k < 2; k++) {
      const double fact = -a[k][i] / a[i][i];
      for (int j = 0; j < 2; j++) a[k][j] += fact * a[i][j];
      b[k] += fact * b[i];
    }
  }
  for (int i = 1; i >= 0; i--) {
    double sum = 0;
    for (int j = i + 1; j < 2; j++) sum += a[i][j] * b[j];
    b[i] = (b[i] - sum) / a[i][i];
  }
  return true;
}

// z of the cross product of (b - a) and (c - a)
double cross(const double *a, const double *b, const double *c) {
  return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
}

// vertex order around the quadrilateral (findClockWise)
void quad_order(const double xv[4][2], int map[4]) {
  for (int i = 0; i < 4; i++) map[i] = i;
  const double abac = cross(xv[0], xv[1], xv[2]), abad = cross(xv[0], xv[1], xv[3]);
  if (abac * abad > 0) {
    const double acad = cross(xv[0], xv[2], xv[3]);
    if (acad * abad > 0) return;  // AB and AD are edges, AC the diagonal
    map[2] = 3;                   // AD is the diagonal
    map[3] = 2;
  } else {  // AB is the diagonal
    map[1] = 2;
    map[2] = 1;
  }
}

}  // namespace

namespace usg {

int nodal_weight_2d(const double *v, int nvert, const double *p, double *f) {
  {
    f[0] = f[1] = f[2] = f[3] = 0.0;
    if (nvert == 3) {
      double a[2][2], b[2];
      for (int i = 0; i < 2; i++) {
        for (int j = 0; j < 2; j++) a[i][j] = v[2 * j + i] - v[4 + i];
        b[i] = p[i] - v[4 + i];
      }
      if (!solve2(a, b)) return -4;  // degenerate triangle ("Error matrix degenerate")
      f[0] = b[0];
      f[1] = b[1];
      f[2] = 1. - f[0] - f[1];
    } else if (nvert == 4) {
      double q[4][2];
      for (int i = 0; i < 4; i++) {
        q[i][0] = v[2 * i];
        q[i][1] = v[2 * i + 1];
      }
      int mp[4];
      quad_order(q, mp);
      double F[4][2];
      for (int i = 0; i < 2; i++) {
        F[0][i] = q[mp[0]][i] - p[i];
        F[1][i] = q[mp[1]][i] - q[mp[0]][i];
        F[2][i] = q[mp[3]][i] - q[mp[0]][i];
        F[3][i] = q[mp[0]][i] - q[mp[1]][i] + q[mp[2]][i] - q[mp[3]][i];
      }
      double u = 0.5, vv = 0.5;
      int it;
      bool ok = true;
      for (it = 0; it < 500; it++) {
        const double uv = u * vv;
        double rhs[2];
        for (int i = 0; i < 2; i++) rhs[i] = F[0][i] + F[1][i] * u + F[2][i] * vv + F[3][i] * uv;
        const double norm = rhs[0] * rhs[0] + rhs[1] * rhs[1];
        if (std::sqrt(norm) <= 1e-14) break;
        double lhs[2][2];
        for (int i = 0; i < 2; i++) {
          lhs[i][0] = F[1][i] + F[3][i] * vv;
          lhs[i][1] = F[2][i] + F[3][i] * u;
        }
        if (!solve2(lhs, rhs)) {
          ok = false;
          break;
        }
        u -= rhs[0] * 1.0;
        vv -= rhs[1] * 1.0;
      }
      if (it == 500 || !ok) {
        u = 2.0;
        vv = 0.0;
      }
      f[mp[0]] = (1 - u) * (1 - vv);
      f[mp[1]] = u * (1 - vv);
      f[mp[2]] = u * vv;
      f[mp[3]] = (1 - u) * vv;
    } else {
      return -1;  // interpolation is defined for triangles and quadrilaterals
    }
  }
  return 0;
}

}  // namespace usg

extern "C" int usg_nodal_weights(int n, const int *nvert, const double *xv, const double *xp, double *w) {
  if (n < 0 || (n && (!nvert || !xv || !xp || !w))) return -1;
  for (int r = 0; r < n; r++) {
    int rc = usg::nodal_weight_2d(xv + 8 * (size_t)r, nvert[r], xp + 2 * (size_t)r, w + 4 * (size_t)r);
    if (rc) return rc;
  }
  return 0;
}
