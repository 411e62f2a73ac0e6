// Command-line driver over the C++ adapters, in the shape of the reference's
// serial scenarios (src/main.cpp:588-697, unstruct_serial): read a Cobalt
// mesh, set the flow, run the integrator, print the residual history and
// the wall force.  With outputEvery > 0 the loop writes the reference's
// Tecplot cell data and Cp files into the working directory every
// outputEvery steps (the reference: 5000).
//
//   usg_run [--renumber] <mesh.grd> <wallTag> <farTag> <Euler2D|ViscousFlow2D|SST2D>
//           <LUSGS|RungeKutta5> <Ma> <Re> <AoA> <maxStep> [minRes] [outputEvery]
//
// --renumber applies the colour-blocked renumbering (usg_mesh_renumber)
// first: the sweeps then run a few waves instead of the file numbering's
// levels.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "usg_strategy.h"

// Overset mode, the sequence of Orchestra::init (Orchestra.cpp:200-264):
//   usg_run --overset <mesh1.grd> <mesh2.grd> <wallTag> <overTag> <dx> <dy> <mexclude> <nfringe>
// reads both meshes, moves the second by (dx, dy), runs the connectivity,
// sets every node value of mesh i to i+1 (4 variables), runs dataUpdate and
// prints per mesh the iblank counts and how many receptor nodes took the
// other mesh's value.
static int overset_main(int argc, char **argv) {
  if (argc < 10) {
    std::fprintf(stderr, "usage: %s --overset m1.grd m2.grd wallTag overTag dx dy mexclude nfringe\n", argv[0]);
    return 2;
  }
  usg::UnstructReader rd;
  rd.setDim(2);
  rd.setWallSign(std::atoi(argv[4]));
  rd.setOverSign(std::atoi(argv[5]));
  rd.addFile(argv[2]);
  rd.addFile(argv[3]);
  rd.readAll();
  rd.make_movement(1, std::atof(argv[6]), std::atof(argv[7]), 0);
  usg::Tioga tg;
  int mexclude = std::atoi(argv[8]), nfringe = std::atoi(argv[9]);
  tg.setMexclude(&mexclude);
  tg.setNfringe(&nfringe);
  for (int i = 0; i < rd.nMesh(); i++) tg.registerGridData(i + 1, rd.mesh(i));
  tg.performConnectivity();
  const int nequ = 4;
  std::vector<std::vector<double>> q(rd.nMesh());
  for (int i = 0; i < rd.nMesh(); i++) {
    q[i].assign((size_t)nequ * rd.mesh(i).nn, (double)(i + 1));
    tg.registerSolution(i + 1, q[i].data());
  }
  tg.dataUpdate(nequ, 0, 0);
  for (int i = 0; i < rd.nMesh(); i++) {
    std::vector<int> ib = tg.iblank(i + 1), ic = tg.iblankCell(i + 1);
    int field = 0, hole = 0, recv = 0, took = 0, kept = 0, cell_recv = 0;
    for (int v = 0; v < rd.mesh(i).nn; v++) {
      field += ib[v] == 1;
      hole += ib[v] == 0;
      recv += ib[v] == -1;
      bool all_other = true, all_own = true;
      for (int k = 0; k < nequ; k++) {
        const double x = q[i][(size_t)nequ * v + k];
        all_other = all_other && std::fabs(x - (double)(2 - i)) < 1e-12;
        all_own = all_own && x == (double)(i + 1);
      }
      took += ib[v] == -1 && all_other;
      kept += ib[v] != -1 && all_own;
    }
    for (int c : ic) cell_recv += c == -1;
    std::printf("mesh %d nodes %d field %d hole %d receptor %d took_other %d unchanged %d cell_receptors %d\n", i + 1,
                rd.mesh(i).nn, field, hole, recv, took, kept, cell_recv);
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::string(argv[1]) == "--overset") {
    try {
      return overset_main(argc, argv);
    } catch (const std::exception &e) {
      std::fprintf(stderr, "error: %s\n", e.what());
      return 1;
    }
  }
  // --renumber: colour-blocked renumbering of the mesh before the run; the
  // printed state is then that of the file's cell 0
  const bool renumber = argc > 1 && std::string(argv[1]) == "--renumber";
  if (renumber) {
    argv++;
    argc--;
  }
  if (argc < 10) {
    std::fprintf(stderr,
                 "usage: %s [--renumber] mesh.grd wallTag farTag Euler2D|ViscousFlow2D|SST2D LUSGS|RungeKutta5 Ma "
                 "Re AoA maxStep [minRes] [outputEvery]\n",
                 argv[0]);
    return 2;
  }
  try {
    usg::UnstructReader rd;
    rd.setDim(2);
    rd.setWallSign(std::atoi(argv[2]));
    rd.setOverSign(std::atoi(argv[3]));
    rd.addFile(argv[1]);
    rd.readAll();
    int c0 = 0;
    if (renumber) {
      const std::vector<int> perm = rd.renumber(0);
      c0 = (int)(std::find(perm.begin(), perm.end(), 0) - perm.begin());
    }
    double aoa = std::atof(argv[8]);
    double minres = argc > 10 ? std::atof(argv[10]) : 1e-8;
    usg::WrapSolver w;
    w.setTopology(&rd);
    w.setParams(argv[4], argv[5], "STABLE", std::atoi(argv[9]), minres, std::atof(argv[6]), std::atof(argv[7]),
                &aoa);
    w.setOutputEvery(argc > 11 ? std::atoi(argv[11]) : 0);
    w.init();
    int n = w.Loop(std::max(1, std::atoi(argv[9]) / 10));
    double cl, cd;
    w.flow()->preprocessAdvanceSerial(0);
    w.flow()->forces(0, cl, cd);
    double **U = w.flow()->getU();  // U[mesh][4 * cell + var]
    std::printf("iterations %d  residual %.15e  CL %.15e  CD %.15e  U[0] = (%.17e %.17e %.17e %.17e)\n", n,
                w.advance()->lastResidual(), cl, cd, U[0][4 * c0], U[0][4 * c0 + 1], U[0][4 * c0 + 2],
                U[0][4 * c0 + 3]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
