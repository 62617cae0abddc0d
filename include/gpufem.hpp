// gpufem.hpp -- C++ host layer over the C-ABI (gpufem.h), shaped like the
// reference's L2/L3 interfaces so code written against them reads the same:
//
//   reference                                   here
//   FemData (src/FEM/FemData.h:47-105)          gfhost::FemData
//   StiffAlgo::CalcGlobalStiffness              gfhost::StiffAlgoHip::CalcGlobalStiffness
//     (src/FEM/StiffAlgo.h:46)
//   FEM::ApplyConstraint(PEN)                   gfhost::FEM::ApplyConstraint
//     (src/FEM/fem.cc:112-145)
//   ELLmatrix::GetElem/Axy/CG                   gfhost::EllMatrixHip::GetElem/Axy/CG
//     (src/SPRmatrix/ELLmatrix.cc:142-310)
//   FEM::CalcStiffnessMat (fem.cc:90-95)        gfhost::FEM::CalcStiffnessMat
//
// Header-only: it only calls the extern "C" entry points, so it adds no C++
// ABI to libgpufem.so.  Errors: the reference asserts or prints; here every
// failing call throws gfhost::Error carrying the gf_* status code and message.
#ifndef GPUFEM_HPP_
#define GPUFEM_HPP_

#include <chrono>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpufem.h"

namespace gfhost {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int code, const gf_ctx* ctx) {
  if (code != GF_OK) throw Error(code, gf_last_error(ctx));
}

// One device context (stream, events) shared by the objects built on it.
class Device {
 public:
  explicit Device(int device = 0) { check(gf_create(&ctx_, device), nullptr); }
  ~Device() { gf_destroy(ctx_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  gf_ctx* ctx() const { return ctx_; }

 private:
  gf_ctx* ctx_ = nullptr;
};

// Model arrays in the reference conventions: 1-based connectivity in the
// reference node ordering, coordinates with stride dim, supports rows
// [node, fix_x, fix_y(, fix_z)], load rows [node, Fx, Fy(, Fz)].
struct FemData {
  int elem_type = GF_Q4;
  int ngauss = 2;
  int dim = 2;
  int nen = 4;
  int64_t nnodes = 0, nelem = 0;
  std::vector<int32_t> conn;
  std::vector<double> coords;
  std::vector<double> elem_E;  // empty: uniform E
  double E = 210000.0, nu = 0.3;
  std::vector<int32_t> supports;  // nsup * (dim + 1)
  std::vector<double> loads;      // nload * (dim + 1)

  int64_t GetNumDof() const { return nnodes * dim; }

  // FemData::CalcForceVector (src/FEM/FemData.cc:370-385): assignment.
  std::vector<double> GetForceVector() const {
    std::vector<double> f(size_t(GetNumDof()), 0.0);
    const size_t row = size_t(dim) + 1;
    for (size_t i = 0; i + row <= loads.size(); i += row)
      for (int j = 1; j <= dim; ++j)
        if (loads[i + size_t(j)] != 0.0)
          f[size_t(dim) * (size_t(loads[i]) - 1) + size_t(j - 1)] = loads[i + size_t(j)];
    return f;
  }

  // Fixed DOFs (0-based) in the order ApplyConstraint visits them.
  std::vector<int32_t> GetFixedDofs() const {
    std::vector<int32_t> d;
    const size_t row = size_t(dim) + 1;
    for (size_t i = 0; i + row <= supports.size(); i += row)
      for (int j = 1; j <= dim; ++j)
        if (supports[i + size_t(j)] == 1) d.push_back(dim * (supports[i] - 1) + (j - 1));
    return d;
  }

  // A structured grid from the library's generator (plate: left edge
  // clamped, unit traction in -y on the right edge; cube: z = 0 clamped,
  // unit pressure on the top face).
  static FemData Structured(int type, int nx, int ny, int nz = 0, int ngauss = 0) {
    FemData d;
    gf_grid_desc g{type, nx, ny, nz, 1.0};
    int32_t nen = 0, dim = 0;
    check(gf_grid_sizes(&g, &d.nnodes, &d.nelem, &nen, &dim), nullptr);
    d.elem_type = type;
    d.nen = nen;
    d.dim = dim;
    d.ngauss = ngauss ? ngauss : ((type == GF_Q8 || type == GF_BRICK20) ? 3 : 2);
    d.conn.resize(size_t(d.nelem) * size_t(nen));
    d.coords.resize(size_t(d.nnodes) * size_t(dim));
    check(gf_grid_fill(&g, d.conn.data(), d.coords.data()), nullptr);
    const int fixed_face = dim == 2 ? 0 : 4, load_face = dim == 2 ? 1 : 5;
    const int64_t nf = gf_grid_face_nodes(&g, fixed_face, nullptr);
    std::vector<int32_t> fixed(static_cast<size_t>(nf));
    gf_grid_face_nodes(&g, fixed_face, fixed.data());
    for (int32_t node : fixed) {
      d.supports.push_back(node);
      for (int c = 0; c < dim; ++c) d.supports.push_back(1);
    }
    std::vector<double> f(size_t(d.GetNumDof()), 0.0);
    check(gf_grid_face_load(&g, load_face, dim - 1, -1.0, f.data()), nullptr);
    for (int64_t v = 0; v < d.nnodes; ++v) {
      bool any = false;
      for (int c = 0; c < dim; ++c) any = any || f[size_t(v * dim + c)] != 0.0;
      if (!any) continue;
      d.loads.push_back(double(v + 1));
      for (int c = 0; c < dim; ++c) d.loads.push_back(f[size_t(v * dim + c)]);
    }
    return d;
  }
};

// Device ELL matrix with the reference's read-side interface.
class EllMatrixHip {
 public:
  EllMatrixHip(Device& dev, gf_ell* m) : dev_(dev), m_(m) {
    check(gf_ell_info_get(m_, &info_), dev_.ctx());
  }
  ~EllMatrixHip() { gf_ell_free(m_); }
  EllMatrixHip(const EllMatrixHip&) = delete;
  EllMatrixHip& operator=(const EllMatrixHip&) = delete;

  gf_ell* handle() const { return m_; }
  int64_t GetMatDim() const { return info_.n; }
  int GetMaxRowLen() const { return info_.width; }
  int64_t GetNNZ() const { return info_.nnz; }

  // ELLmatrix::GetElem (ELLmatrix.cc:142-159): 0 when the entry is absent.
  double GetElem(int64_t row, int64_t col) {
    if (row < 0 || row >= info_.n || col < 0 || col >= info_.n) return 0.0;
    Sync();
    for (int32_t s = 0; s < rownnz_[size_t(row)]; ++s) {
      const size_t idx = size_t(row + int64_t(s) * info_.n);
      if (col_[idx] == col) return val_[idx];
    }
    return 0.0;
  }

  // ELLmatrix::Axy (ELLmatrix.cc:194-204)
  void Axy(const double* x, double* y) {
    check(gf_spmv(dev_.ctx(), m_, x, y, 1, nullptr), dev_.ctx());
  }

  // SPRmatrix::CG(x, b, niter, eps) semantics (SPRmatrix.cc:405-480):
  // x0 = 0, stop when ||r|| <= eps ||b||; unpreconditioned like the
  // reference unless jacobi is set.  Returns the iteration count.
  int CG(double* x, const double* b, int niter, double eps, bool jacobi = false) {
    gf_solve_opts o;
    gf_solve_opts_default(&o);
    o.precond = jacobi ? GF_PRECOND_JACOBI : GF_PRECOND_NONE;
    o.maxit = niter;
    o.rtol = eps;
    gf_solve_stats st;
    const int code = gf_pcg(dev_.ctx(), m_, b, x, 1, &o, &st);
    if (code != GF_OK && code != GF_ERR_NOCONV) check(code, dev_.ctx());
    last_ = st;
    return st.iterations;
  }
  const gf_solve_stats& LastStats() const { return last_; }

  void Invalidate() { synced_ = false; }

 private:
  void Sync() {
    if (synced_) return;
    const size_t slots = size_t(info_.n) * size_t(info_.width);
    val_.resize(slots);
    col_.resize(slots);
    rownnz_.resize(size_t(info_.n));
    check(gf_ell_download(dev_.ctx(), m_, val_.data(), col_.data(), rownnz_.data()), dev_.ctx());
    synced_ = true;
  }

  Device& dev_;
  gf_ell* m_ = nullptr;
  gf_ell_info info_{};
  bool synced_ = false;
  std::vector<double> val_;
  std::vector<int32_t> col_, rownnz_;
  gf_solve_stats last_{};
};

// StiffAlgo strategy on the GPU: CalcGlobalStiffness returns the elapsed
// seconds (as StiffAlgoCPU::CalcStiffnessColoring does) and leaves the
// assembled matrix in Stiffness().  Coloring is always used (the scatter is
// atomic-free only per color).
class StiffAlgoHip {
 public:
  explicit StiffAlgoHip(Device& dev) : dev_(dev) {}
  ~StiffAlgoHip() {
    if (mesh_) gf_mesh_free(mesh_);
  }
  StiffAlgoHip(const StiffAlgoHip&) = delete;
  StiffAlgoHip& operator=(const StiffAlgoHip&) = delete;

  double CalcGlobalStiffness(FemData* d) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!mesh_) {
      gf_mesh_desc md{d->elem_type, d->ngauss, d->nnodes, d->nelem, d->conn.data(),
                      d->coords.data(), d->elem_E.empty() ? nullptr : d->elem_E.data(),
                      d->E, d->nu};
      check(gf_mesh_upload(dev_.ctx(), &md, &mesh_), dev_.ctx());
      gf_ell* m = nullptr;
      check(gf_ell_symbolic(dev_.ctx(), mesh_, &m, nullptr), dev_.ctx());
      K_.reset(new EllMatrixHip(dev_, m));
    }
    check(gf_assemble(dev_.ctx(), mesh_, K_->handle(), nullptr), dev_.ctx());
    K_->Invalidate();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  EllMatrixHip* Stiffness() { return K_.get(); }

 private:
  Device& dev_;
  gf_mesh* mesh_ = nullptr;
  std::unique_ptr<EllMatrixHip> K_;
};

// FEM facade (src/FEM/fem.h:57-76) for the GPU path.
class FEM {
 public:
  explicit FEM(Device& dev) : dev_(dev), algo_(dev) {}
  void Init(FemData data) { data_ = std::move(data); }
  FemData* GetFemData() { return &data_; }
  double CalcStiffnessMat() { return algo_.CalcGlobalStiffness(&data_); }
  EllMatrixHip* GetStiffnessMatrix() { return algo_.Stiffness(); }
  // FEM::ApplyConstraint(PEN): returns the penalty factor.
  double ApplyConstraint(double coef = 1e4) {
    const std::vector<int32_t> dofs = data_.GetFixedDofs();
    double pen = 0.0;
    check(gf_apply_penalty(dev_.ctx(), algo_.Stiffness()->handle(), dofs.data(),
                           int64_t(dofs.size()), coef, &pen),
          dev_.ctx());
    algo_.Stiffness()->Invalidate();
    return pen;
  }

 private:
  Device& dev_;
  StiffAlgoHip algo_;
  FemData data_;
};

}  // namespace gfhost

#endif  // GPUFEM_HPP_
