// Host-side mesh preparation for the gfx950 solver: the reference's one-time
// geometry precompute (cell vertex order, volume, vertex-mean centre, face
// normal/area/midpoint, projected wall distance) and the LU-SGS wave
// schedule.  Plain C++17, no HIP.
#pragma once
#include <string>
#include <vector>

namespace usg {

constexpr int kWall = -2;      // edge3d<2>::BoundaryType::WALL
constexpr int kFarfield = -3;  // edge3d<2>::BoundaryType::FARFIELD
constexpr int kNoFace = -2147483647 - 1;

struct HostMesh {
  int nn = 0, nf = 0, nc = 0;
  std::vector<double> xy;     // [nn*2]
  std::vector<int> fnode;     // [nf*2]
  std::vector<int> flc, frc;  // [nf]
  // derived
  std::vector<int> cnode, cnv;       // [nc*4], [nc]
  std::vector<int> cface, cnf;       // [nc*4] ascending, [nc]
  std::vector<double> vol, cc;       // [nc], [nc*2]
  std::vector<double> fn, fa, fc;    // [nf*2], [nf], [nf*2]
  std::vector<double> wdist;         // [nc]
  // LU-SGS waves: cells grouped by dependency level (ascending cell id inside)
  std::vector<int> fwd_cells, fwd_off;
  std::vector<int> bwd_cells, bwd_off;
  // maximal contiguous ranges of mutually non-adjacent cells; when there are
  // as many as forward levels (colour-blocked numbering) they are used as the
  // waves of both sweeps (forward in order, backward in reverse: a cell's
  // upper neighbours always lie in later ranges)
  std::vector<int> ranges;
  bool use_ranges = false;
  std::vector<int> wall_faces;
  // XCD-aware work lists: cells / faces / wave cells grouped by spatial part
  // (kParts equal slices of the Hilbert curve of the cell centres; a face
  // belongs to its left cell's part); cells in Hilbert order, faces and wave
  // cells in ascending id inside a part.  Launches
  // send part p's blocks to XCD p (blockIdx % kParts) so that neighbour
  // gathers hit the XCD's own L2.
  std::vector<int> part;                      // [nc]
  std::vector<int> cell_list, cell_off;       // [nc], [kParts+1]
  std::vector<int> face_list, face_off;       // [nf], [kParts+1]
  std::vector<int> fwd_list, fwd_poff;        // [nc], [nwaves*(kParts+1)]
  std::vector<int> bwd_list, bwd_poff;        // [nc], [nwaves*(kParts+1)]
  // batch of independent cases (one case when built by build_mesh): case i
  // owns cells [case_cell[i], case_cell[i+1]), faces [case_face[i],
  // case_face[i+1]) and wall faces wall_faces[case_wall[i] .. case_wall[i+1])
  std::vector<int> case_cell, case_face, case_wall;
};

constexpr int kParts = 8;  // XCDs of an MI355X

// Validates the faces and returns each cell's vertices in the reference's
// order (Brick4Cobalt2D::insert, CellBrick.cpp:13-136): cnode[nc*4] (-1
// padded), cnv[nc] in {3, 4}.  Returns an empty string or the error text.
std::string cell_vertices(int nn, int nf, const int *fnode, const int *flc, const int *frc, int nc,
                          std::vector<int> &cnode, std::vector<int> &cnv);

// Non-dimensional free stream (rho, p, u, v) of Mach `mach` at `aoa_deg`,
// as TopologyHolderStrategy::set_fs_variable2 computes it (sea-level
// density and pressure, pi = 3.141593); aoa_rad (may be null) gets the
// angle in radians.
void freestream_state(double mach, double aoa_deg, double fs[4], double *aoa_rad);

// Area of a triangle or quadrilateral (computeCellVolume2D,
// mathTools/cellVolume.cpp:21-78).
double cell_volume(double xv[4][2], int nv);

// Weights f[0..nvert) of point p on the vertices v[2*nvert] of a triangle or
// quadrilateral (computeNodalWeight2D, mathTools/math.cpp:229-312); 0, or -1
// for another vertex count, -4 for a degenerate triangle.
int nodal_weight_2d(const double *v, int nvert, const double *p, double *f);

// Sets the calling thread's error text (usg_last_error) and returns code.
int set_error(int code, const std::string &msg);

// Validates and copies the mesh, then computes the derived geometry.
// Returns an empty string on success, else the error text.
std::string build_mesh(HostMesh &m, int nn, const double *xy, int nf, const int *fnode,
                       const int *flc, const int *frc, int nc);

// Block-diagonal union of independently built cases, case-major: every
// index is offset by the cases before it, wave w (and part p) of the union
// is the concatenation of the cases' wave w (part p) lists.  Each case keeps
// its own geometry, wall distance and numbering, so its arithmetic is
// unchanged by batching.
void merge_meshes(const std::vector<HostMesh> &cases, HostMesh &u);

}  // namespace usg
