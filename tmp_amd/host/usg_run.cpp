// Command-line driver over the C++ adapters, in the shape of the reference's
// serial scenarios (src/main.cpp:588-697, unstruct_serial): read a Cobalt
// mesh, set the flow, run the integrator, print the residual history and
// the wall force.  With outputEvery > 0 the loop writes the reference's
// Tecplot cell data and Cp files into the working directory every
// outputEvery steps (the reference: 5000).
//
//   usg_run <mesh.grd> <wallTag> <farTag> <Euler2D|ViscousFlow2D|SST2D>
//           <LUSGS|RungeKutta5> <Ma> <Re> <AoA> <maxStep> [minRes] [outputEvery]
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <string>

#include "usg_strategy.h"

int main(int argc, char **argv) {
  if (argc < 10) {
    std::fprintf(stderr,
                 "usage: %s mesh.grd wallTag farTag Euler2D|ViscousFlow2D|SST2D LUSGS|RungeKutta5 Ma Re AoA "
                 "maxStep [minRes] [outputEvery]\n",
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
    double **U = w.flow()->getU();
    std::printf("iterations %d  residual %.15e  CL %.15e  CD %.15e  U[0] = (%.17e %.17e %.17e %.17e)\n", n,
                w.advance()->lastResidual(), cl, cd, U[0][0], U[0][1], U[0][2], U[0][3]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
