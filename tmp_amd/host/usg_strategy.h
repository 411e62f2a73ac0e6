// C++ host adapters over the C ABI (include/usg.h), mirroring the reference's
// solver-facing interfaces so a driver written against them switches to the
// gfx950 path by changing the constructed types:
//   UnstructReader           reader/UnstructReader.h:88-285 (serial Cobalt reader)
//   TopologyHolderStrategy   UnstructSolver/TopologyHolderStrategy.h:15-336
//   TimeIntegrator           UnstructSolver/TimeIntegrator.h:5-27
//   LUSGSIntegrator          UnstructSolver/LUSGS_Integrator.h:8-82
//   RungeKuttaStrategy       UnstructSolver/RungeKuttaStrategy.h:5-221
//   WrapSolver               UnstructSolver/SolverWrapper.h:7-95
// Getters return host mirrors in the reference's AoS layout (double** per
// mesh), refreshed from the device on each call.  Errors throw
// std::runtime_error with the library's message, as the reference throws.
// Header-only C++17; link against libusg.so.
#pragma once

#include <cmath>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/usg.h"

namespace usg {

inline void check(int rc) {
  if (rc != USG_OK) throw std::runtime_error(std::string("usg: ") + usg_last_error());
}

// ---------------------------------------------------------------------------
// Mesh input: the serial Cobalt reader's conventions (UnstructReader.h:124-262)
// ---------------------------------------------------------------------------
struct MeshData {
  int nn = 0, nf = 0, nc = 0;
  std::vector<double> xy;
  std::vector<int> fnode, flc, frc;
  usg_mesh_desc desc() const {
    return usg_mesh_desc{nn, nf, nc, xy.data(), fnode.data(), flc.data(), frc.data()};
  }
};

class UnstructReader {
 public:
  void setDim(int ndim) {
    if (ndim != 2) throw std::runtime_error("Not implemented yet\n");
  }
  void setOverSign(int sign) { signOver_ = sign; }
  void setWallSign(int sign) { signWall_ = sign; }
  void addFile(const std::string &f) { files_.push_back(f); }
  int nMesh() const { return (int)meshes_.size(); }
  const MeshData &mesh(int i) const { return meshes_.at(i); }

  void readAll() {
    meshes_.clear();
    for (const auto &f : files_) meshes_.push_back(read(f));
  }
  // Colour-blocked renumbering of mesh i (usg_mesh_renumber), so that the
  // LU-SGS sweeps run one wave per colour instead of one per level of the
  // file numbering; returns perm[new cell] = file cell.  No counterpart in
  // the reference, which sweeps its file numbering.
  std::vector<int> renumber(int i) {
    MeshData &m = meshes_.at(i);
    const usg_mesh_desc d = m.desc();
    std::vector<int> perm(m.nc), fn(2 * (size_t)m.nf), l(m.nf), r(m.nf);
    int ncolours = 0;
    if (usg_mesh_renumber(&d, perm.data(), fn.data(), l.data(), r.data(), &ncolours) != USG_OK)
      throw std::runtime_error(usg_last_error());
    m.fnode.swap(fn);
    m.flc.swap(l);
    m.frc.swap(r);
    return perm;
  }
  // CobaltReader::make_movement (reader/CobaltReader.h:87-93): translate mesh i
  void make_movement(int i, double dx, double dy, double /*dz*/) {
    MeshData &m = meshes_.at(i);
    for (int k = 0; k < m.nn; k++) {
      m.xy[2 * (size_t)k] += dx;
      m.xy[2 * (size_t)k + 1] += dy;
    }
  }

 private:
  MeshData read(const std::string &filename) {
    std::ifstream fin(filename);
    if (!fin.is_open()) throw std::runtime_error("Open file Failure\n");
    MeshData m;
    int skip;
    fin >> skip >> skip >> skip;
    fin >> m.nn >> m.nf >> m.nc >> skip >> skip;
    m.xy.resize(2 * (size_t)m.nn);
    for (auto &v : m.xy) fin >> v;
    m.fnode.resize(2 * (size_t)m.nf);
    m.flc.resize(m.nf);
    m.frc.resize(m.nf);
    for (int i = 0; i < m.nf; i++) {
      int cnt, lC, rC;
      fin >> cnt;
      if (cnt != 2) throw std::runtime_error("2-D faces must have two nodes\n");
      fin >> m.fnode[2 * i] >> m.fnode[2 * i + 1] >> lC >> rC;
      m.fnode[2 * i] -= 1;
      m.fnode[2 * i + 1] -= 1;
      if (lC <= 0) throw std::runtime_error("Left cell can't be negative\n");
      m.flc[i] = lC - 1;
      if (rC > 0)
        m.frc[i] = rC - 1;
      else if (rC == signWall_)
        m.frc[i] = USG_WALL;
      else if (rC == signOver_)
        m.frc[i] = USG_FARFIELD;
      else
        throw std::runtime_error(" Not a valid boundary type\n");
    }
    if (!fin) throw std::runtime_error("truncated .grd file\n");
    return m;
  }
  int signWall_ = -2, signOver_ = -3;
  std::vector<std::string> files_;
  std::vector<MeshData> meshes_;
};

// ---------------------------------------------------------------------------
// Flow strategy: one GPU context per mesh (the reference's d_nmesh blocks)
// ---------------------------------------------------------------------------
class GpuFlowStrategy {
 public:
  // flow: USG_EULER ("Euler2D") / USG_LAMINAR ("ViscousFlow2D") / USG_SST
  explicit GpuFlowStrategy(int flow, int precision = USG_FP64) : flow_(flow), precision_(precision) {}
  ~GpuFlowStrategy() {
    for (auto *c : ctx_) usg_ctx_destroy(c);
  }
  GpuFlowStrategy(const GpuFlowStrategy &) = delete;
  GpuFlowStrategy &operator=(const GpuFlowStrategy &) = delete;

  // TopologyHolderStrategy::set_fs_variable2 (TopologyHolderStrategy.cpp:325-351);
  // density and pressure are fixed at 1.225 / 101325 as WrapSolver passes them.
  void set_fs_variable2(double Ma, double AOA, double /*density*/, double /*pressure*/, double Re) {
    ma_ = Ma;
    aoa_ = AOA;
    re_ = Re;
  }
  // TopologyHolderStrategy::initializeStrategy
  void initializeStrategy(const UnstructReader &rd) {
    for (int i = 0; i < rd.nMesh(); i++) {
      usg_mesh_desc d = rd.mesh(i).desc();
      usg_params p{flow_, precision_, ma_, aoa_, re_, cfl_lusgs_, cfl_rk_};
      usg_ctx *c = nullptr;
      check(usg_ctx_create(&d, &p, &c));
      ctx_.push_back(c);
      nc_.push_back(d.nc);
      nf_.push_back(d.nf);
      meshes_.push_back(rd.mesh(i));
    }
    mirrors_.assign(kFields, std::vector<std::vector<double>>(ctx_.size()));
    ptrs_.assign(kFields, std::vector<double *>(ctx_.size(), nullptr));
  }
  void setCFL(double lusgs, double rk) {
    cfl_lusgs_ = lusgs;
    cfl_rk_ = rk;
  }
  int getNMesh() const { return (int)ctx_.size(); }
  int getNEquation() const { return 4; }
  int nCells(int m) const { return nc_.at(m); }
  bool isInvicid() const { return flow_ == USG_EULER; }
  bool isLaminar() const { return flow_ != USG_SST; }

  // Euler2D::initializeData (Euler2D.cpp:153-171)
  void initializeData() {
    for (auto *c : ctx_) check(usg_init_freestream(c));
  }
  // warm start from AoS primitives of mesh m (and (k, omega) for SST)
  void setPrimitive(int m, const double *U, const double *turb = nullptr) {
    check(usg_set_prim(ctx_.at(m), U, turb));
  }
  // preprocessAdvanceSerial (Euler2D.h:133-140 / ViscousFlow2D.h:192-212)
  void preprocessAdvanceSerial(int /*istage*/) {
    for (auto *c : ctx_) check(usg_residual(c));
  }
  // preprocessAdvance (Euler2D.h:115-131): the MPI variant exchanges halo
  // cells first; one GPU holds the whole mesh, so it is the serial one
  void preprocessAdvance(int istage) { preprocessAdvanceSerial(istage); }
  // postprocessAdvanceSerial (Euler2D.h:142-146) does nothing;
  // postprocessAdvance (Euler2D.h:165-) exchanges halo cells, which one GPU
  // does not have
  void postprocessAdvance(int /*istage*/) {}
  void postprocessAdvanceSerial(int /*istage*/) {}
  // SolveTime(CFL) (Euler2D.h:67-77): the local time step is formed inside
  // each LU-SGS iteration; this sets the CFL number it uses from the next
  // iteration on
  void SolveTime(double CFL) {
    cfl_lusgs_ = CFL;
    for (auto *c : ctx_) check(usg_set_cfl(c, CFL));
  }
  // cleaning (Euler2D.h:150-163) zeroes MPI receive cells; there are none
  void cleaning(double ** /*W0*/) {}

  // getters of TopologyHolderStrategy.h:242-303 (host mirrors, AoS)
  double **getU() { return mirror(USG_F_PRIM, 4); }
  double **getUEdge() { return mirror(USG_F_UEDGE, 4, true); }
  double **getResidual() { return mirror(USG_F_RES, 4); }
  void **getGradientPrimitive() { return (void **)mirror(USG_F_GRAD, 8); }
  double **getSpectrumConvective() { return mirror(USG_F_LAMC, 1); }
  double **getSpectrumViscous() { return isInvicid() ? nullptr : mirror(USG_F_LAMV, 1); }
  void **getGradientT() { return isInvicid() ? nullptr : (void **)mirror(USG_F_GRADT, 2); }
  double **getStableTime() { return mirror(USG_F_DT, 1); }
  double **getLimiter() { return mirror(USG_F_PHI, 4); }
  double **getTurbulence() { return flow_ == USG_SST ? mirror(USG_F_TURB, 2) : nullptr; }

  // integrated wall force of mesh m
  void forces(int m, double &cl, double &cd) { check(usg_forces(ctx_.at(m), &cl, &cd)); }
  double residualNorm(int m) {
    double r;
    check(usg_res_l2(ctx_.at(m), &r));
    return r;
  }
  usg_ctx *context(int m) { return ctx_.at(m); }

  // TopologyHolderStrategy::writeCellData (TopologyHolderStrategyOutput.cpp:
  // 466-554): Tecplot file "<filename>_<meshTag+1>_<proc+1>" of the current
  // primitive state of mesh meshTag
  void writeCellData(const std::string &filename, int proc, int meshTag) {
    double **U = getU();
    usg_mesh_desc d = meshes_.at(meshTag).desc();
    double fs[4];
    check(usg_freestream_state(ma_, aoa_, fs));
    check(usg_write_cell_data(&d, U[meshTag], fs, filename.c_str(), proc, meshTag));
  }
  // Wall C_p of mesh mesh_ind (Euler2D::outPutCp, Euler2D.cpp:392-623):
  // table = every wall node (x, y, signed count, C_p) in the reference's
  // order, curve = the merged rows its text file plots
  void surfaceCp(int mesh_ind, std::vector<double> &table, std::vector<double> &curve) {
    double **U = getU();
    usg_mesh_desc d = meshes_.at(mesh_ind).desc();
    double fs[4];
    check(usg_freestream_state(ma_, aoa_, fs));
    int nw = 0, ncv = 0;
    check(usg_surface_cp(&d, U[mesh_ind], fs, 0, nullptr, nullptr, &nw, &ncv));
    table.assign(4 * (size_t)nw, 0.0);
    curve.assign(4 * (size_t)nw, 0.0);
    check(usg_surface_cp(&d, U[mesh_ind], fs, nw, table.data(), curve.data(), &nw, &ncv));
    curve.resize(4 * (size_t)ncv);
  }
  // Euler2D::outPutCp: writes the text file "Cp_<step>" in the working
  // directory.  The reference also stores the sorted table in the HDF5 file
  // `filename`; there is no HDF5 library here, surfaceCp returns that table.
  void outPutCp(const std::string & /*filename*/, int mesh_ind, int step) {
    std::vector<double> table, curve;
    surfaceCp(mesh_ind, table, curve);
    check(usg_write_cp(("Cp_" + std::to_string(step)).c_str(), curve.data(), (int)(curve.size() / 4)));
  }

 private:
  static constexpr int kFields = 32;
  double **mirror(int field, int width, bool faces = false) {
    for (size_t m = 0; m < ctx_.size(); m++) {
      auto &buf = mirrors_[field][m];
      buf.resize((size_t)width * (faces ? nf_[m] : nc_[m]));
      check(usg_get_field(ctx_[m], field, buf.data(), (long)buf.size()));
      ptrs_[field][m] = buf.data();
    }
    return ptrs_[field].data();
  }
  int flow_, precision_;
  double ma_ = 0.5, aoa_ = 0.0, re_ = 1e6, cfl_lusgs_ = 400, cfl_rk_ = 0.8;
  std::vector<usg_ctx *> ctx_;
  std::vector<int> nc_, nf_;
  std::vector<MeshData> meshes_;
  std::vector<std::vector<std::vector<double>>> mirrors_;
  std::vector<std::vector<double *>> ptrs_;
};

// ---------------------------------------------------------------------------
// Overset assembly (UniTioga/tioga.h): the calls Orchestra makes
// (Orchestra.cpp:100-264) -- registerGridData, setMexclude / setNfringe,
// performConnectivity, registerSolution, dataUpdate -- over usg_overset_*
// and the GPU stencil apply.  Node solutions are AoS (nvar per node), as
// registerSolution takes them.
// ---------------------------------------------------------------------------
class Tioga {
 public:
  ~Tioga() {
    if (ov_) usg_overset_destroy(ov_);
  }
  // mesh tags are 1, 2, ... in registration order; overset faces are those
  // with the reader's overset tag (here: the far-field faces of the mesh)
  void registerGridData(int btag, const MeshData &m) {
    if (btag != (int)meshes_.size() + 1) throw std::runtime_error("tioga: mesh tags must be 1, 2, ...");
    meshes_.push_back(m);
    std::vector<int> f(m.nf, 0);
    for (int i = 0; i < m.nf; i++) f[i] = m.frc[i] == USG_FARFIELD;
    overset_.push_back(std::move(f));
    q_.push_back(nullptr);
  }
  void setMexclude(int *m) { mexclude_ = *m; }
  void setNfringe(int *n) { nfringe_ = *n; }
  void performConnectivity() {
    std::vector<usg_mesh_desc> d;
    std::vector<const int *> f;
    for (size_t i = 0; i < meshes_.size(); i++) {
      d.push_back(meshes_[i].desc());
      f.push_back(overset_[i].data());
    }
    if (ov_) usg_overset_destroy(ov_);
    ov_ = nullptr;
    check(usg_overset_create(d.data(), f.data(), (int)d.size(), mexclude_, nfringe_, &ov_));
  }
  // MeshBlock::iblank / iblank_cell of mesh `btag`
  std::vector<int> iblank(int btag) const { return ib(btag, true); }
  std::vector<int> iblankCell(int btag) const { return ib(btag, false); }
  void registerSolution(int btag, double *q) { q_.at(btag - 1) = q; }
  // tioga::dataUpdate (tioga.cpp:326-471): every receptor node takes the
  // weighted donor-cell vertex values of the other mesh.  As there, all
  // donor values are taken before any receptor is written.
  void dataUpdate(int nvar, int /*interptype*/, int /*at_points*/) {
    if (!ov_) throw std::runtime_error("tioga: dataUpdate before performConnectivity");
    std::vector<std::vector<double>> src(meshes_.size());
    for (size_t b = 0; b < meshes_.size(); b++) {
      if (!q_[b]) throw std::runtime_error("tioga: Solution data not set, cannot update");
      src[b].assign(q_[b], q_[b] + (size_t)nvar * meshes_[b].nn);
    }
    for (size_t b = 0; b < meshes_.size(); b++) {
      int n = 0;
      check(usg_overset_receptors(ov_, (int)b, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
      if (n == 0) continue;
      std::vector<int> rb(n), rn(n), dc(n), nw(n), dn(4 * (size_t)n);
      std::vector<double> w(4 * (size_t)n);
      check(usg_overset_receptors(ov_, (int)b, &n, rb.data(), rn.data(), dc.data(), nw.data(), dn.data(), w.data()));
      // receptors grouped by their mesh
      for (size_t r = 0; r < meshes_.size(); r++) {
        std::vector<int> sr, snw, sdn;
        std::vector<double> sw;
        for (int k = 0; k < n; k++) {
          if (rb[k] != (int)r) continue;
          sr.push_back(rn[k]);
          snw.push_back(nw[k]);
          sdn.insert(sdn.end(), &dn[4 * (size_t)k], &dn[4 * (size_t)k + 4]);
          sw.insert(sw.end(), &w[4 * (size_t)k], &w[4 * (size_t)k + 4]);
        }
        if (sr.empty()) continue;
        check(usg_stencil_apply_host(src[b].data(), q_[r], nvar, USG_ROW, meshes_[b].nn, meshes_[r].nn,
                                     (int)sr.size(), sr.data(), snw.data(), sdn.data(), sw.data()));
      }
    }
  }

 private:
  std::vector<int> ib(int btag, bool node) const {
    const MeshData &m = meshes_.at(btag - 1);
    std::vector<int> out(node ? m.nn : m.nc);
    check(usg_overset_iblank(ov_, btag - 1, node ? out.data() : nullptr, node ? nullptr : out.data()));
    return out;
  }
  std::vector<MeshData> meshes_;
  std::vector<std::vector<int>> overset_;
  std::vector<double *> q_;
  int mexclude_ = 3, nfringe_ = 1;
  usg_overset *ov_ = nullptr;
};

// ---------------------------------------------------------------------------
// Time integrators (TimeIntegrator.h:5-27)
// ---------------------------------------------------------------------------
class GpuTimeIntegrator {
 public:
  virtual ~GpuTimeIntegrator() = default;
  void initializeStrategy(GpuFlowStrategy *s) { strategy_ = s; }
  virtual void initializeData() { strategy_->initializeData(); }
  virtual void singleStep(int curStep) = 0;
  // The reference's converged() falls off the end (LUSGS_Integrator.cpp:156-161);
  // here: ||R_rho/V||_2 of the last step below tolerance * first step's.
  virtual bool converged() {
    if (first_ <= 0) return false;
    return last_ <= tol_ * first_;
  }
  void setTolerance(double t) { tol_ = t; }
  double lastResidual() const { return last_; }

 protected:
  void record(double r) {
    if (first_ <= 0) first_ = r;
    last_ = r;
  }
  GpuFlowStrategy *strategy_ = nullptr;
  double first_ = -1, last_ = -1, tol_ = 1e-8;
};

// LUSGSIntegrator::singleStepSerial -> UpdateSerial (LUSGS_Integrator.cpp:124-129,228-255)
class GpuLUSGSIntegrator : public GpuTimeIntegrator {
 public:
  void singleStep(int /*curStep*/) override { run(1); }
  // n iterations in one call (one launch sequence, one synchronisation)
  void run(int n) {
    std::vector<double> h(n > 0 ? n : 1);
    for (int m = 0; m < strategy_->getNMesh(); m++) check(usg_lusgs_step(strategy_->context(m), n, h.data()));
    if (n > 0) record(h[n - 1]);
  }
};

// RungeKuttaStrategy(5)::singleStepSerial (RungeKuttaStrategy.h:28-72)
class GpuRK5Integrator : public GpuTimeIntegrator {
 public:
  void singleStep(int /*curStep*/) override {
    double h = 0;
    for (int m = 0; m < strategy_->getNMesh(); m++) check(usg_rk5_step(strategy_->context(m), 1, &h));
    record(h);
  }
};

// ---------------------------------------------------------------------------
// WrapSolver (SolverWrapper.h:7-95): string dispatch of flow and integrator
// ---------------------------------------------------------------------------
class WrapSolver {
 public:
  void setTopology(UnstructReader *rd) { reader_ = rd; }
  void setParams(const std::string &flowT, const std::string &advanceT, const std::string & /*solveT*/,
                 int MaxStep, double MinRes, double ma, double re, double *aoa) {
    int flow;
    if (flowT == "Euler2D")
      flow = USG_EULER;
    else if (flowT == "ViscousFlow2D")
      flow = USG_LAMINAR;
    else if (flowT == "SST2D")
      flow = USG_SST;
    else
      throw std::runtime_error("Undefined Flow Type\n");
    flow_.reset(new GpuFlowStrategy(flow));
    if (advanceT == "RungeKutta5")
      advance_.reset(new GpuRK5Integrator());
    else if (advanceT == "LUSGS")
      advance_.reset(new GpuLUSGSIntegrator());
    else
      throw std::runtime_error("Undefined Advance Type\n");
    flow_->set_fs_variable2(ma, *aoa, 1.225, 101325, re);
    maxStep_ = MaxStep;  // stored (the reference leaves d_maxStep uninitialised)
    minRes_ = MinRes;
    advance_->setTolerance(MinRes);
  }
  void init() {
    flow_->initializeStrategy(*reader_);
    advance_->initializeStrategy(flow_.get());
    advance_->initializeData();
  }
  // Loop (SolverWrapper.h:71-90): after step i with i % outputEvery == 0
  // every mesh j writes "CellDataParallelLUSGS<i+1>_<j+1>_1" (writeCellData)
  // and "Cp_<i>" (outPutCp), as the reference does every 5000 steps; stops
  // at MaximumTimeStep or when the relative residual reaches MinimumResidual.
  int Loop(int report_every = 0) {
    int i = 0;
    for (; i < maxStep_; i++) {
      advance_->singleStep(i);
      if (report_every > 0 && i % report_every == 0)
        std::cout << i << " " << advance_->lastResidual() << '\n';
      if (outputEvery_ > 0 && i % outputEvery_ == 0) {
        for (int j = 0; j < flow_->getNMesh(); j++) {
          flow_->writeCellData("CellDataParallelLUSGS" + std::to_string(i + 1), 0, j);
          flow_->outPutCp("OneAndOnlyLegendaryCp_" + std::to_string(j + 1), j, i);
        }
      }
      if (advance_->converged()) {
        i++;
        break;
      }
    }
    return i;
  }
  GpuFlowStrategy *flow() { return flow_.get(); }
  GpuTimeIntegrator *advance() { return advance_.get(); }
  // steps between output dumps in Loop (the reference's 5000); 0 = none
  void setOutputEvery(int n) { outputEvery_ = n; }

 private:
  UnstructReader *reader_ = nullptr;
  std::unique_ptr<GpuFlowStrategy> flow_;
  std::unique_ptr<GpuTimeIntegrator> advance_;
  int maxStep_ = 0;
  double minRes_ = 0;
  int outputEvery_ = 5000;
};

}  // namespace usg
