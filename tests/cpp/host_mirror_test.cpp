// Tests of the C++ host layer (include/gpufem.hpp), written the way the
// reference's own suites are: FemTest (src/Test/FemTest.cpp:117-187) and
// AxyGPUTest (src/Test/AxyGPUTest.cpp:44-78, 201-239).  Prints one line per
// test and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpufem.hpp"

#define EXPECT(cond)                                                        \
  do {                                                                      \
    if (!(cond)) {                                                          \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);           \
      std::exit(1);                                                         \
    }                                                                       \
  } while (0)

using gfhost::Device;
using gfhost::EllMatrixHip;
using gfhost::FemData;
using gfhost::FEM;

static void stiffness_diag_positive(Device& dev, int type, int nx, int ny, int nz) {
  FEM fem(dev);
  fem.Init(FemData::Structured(type, nx, ny, nz));
  double t = fem.CalcStiffnessMat();
  EXPECT(t >= 0.0);
  EllMatrixHip* K = fem.GetStiffnessMatrix();
  for (int64_t i = 0; i < K->GetMatDim(); ++i) EXPECT(K->GetElem(i, i) > 0.0);
}

static void penalty_increases_k00(Device& dev) {
  FEM fem(dev);
  fem.Init(FemData::Structured(GF_Q4, 8, 4));
  fem.CalcStiffnessMat();
  const double k00 = fem.GetStiffnessMatrix()->GetElem(0, 0);
  const double pen = fem.ApplyConstraint();
  const double k00c = fem.GetStiffnessMatrix()->GetElem(0, 0);
  EXPECT(k00 <= k00c);
  EXPECT(std::fabs(k00c - pen * k00) <= 1e-12 * k00c);
}

static void solve_plate(Device& dev) {
  FEM fem(dev);
  fem.Init(FemData::Structured(GF_Q8, 16, 8));
  fem.CalcStiffnessMat();
  fem.ApplyConstraint();
  EllMatrixHip* K = fem.GetStiffnessMatrix();
  const std::vector<double> f = fem.GetFemData()->GetForceVector();
  std::vector<double> u(f.size(), 0.0), r(f.size(), 0.0);
  const int it = K->CG(u.data(), f.data(), 10000, 1e-8, /*jacobi=*/true);
  EXPECT(it > 0 && K->LastStats().converged);
  K->Axy(u.data(), r.data());
  double rn = 0, bn = 0;
  for (size_t i = 0; i < f.size(); ++i) {
    rn += (f[i] - r[i]) * (f[i] - r[i]);
    bn += f[i] * f[i];
  }
  EXPECT(std::sqrt(rn / bn) <= 1.2e-8);
}

static EllMatrixHip* upload(Device& dev, int64_t n, int32_t w, const std::vector<double>& val,
                            const std::vector<int32_t>& col, const std::vector<int32_t>& rn) {
  gf_ell* m = nullptr;
  gfhost::check(gf_ell_upload(dev.ctx(), n, w, val.data(), col.data(), rn.data(), &m), dev.ctx());
  return new EllMatrixHip(dev, m);
}

static void axy_diagonal_kat(Device& dev) {
  const int64_t n = 4096;
  const size_t un = static_cast<size_t>(n);
  std::vector<double> val(un), x(un), y(un);
  std::vector<int32_t> col(un), rn(un, 1);
  for (int64_t i = 0; i < n; ++i) {
    val[size_t(i)] = double(i + 1);
    col[size_t(i)] = int32_t(i);
    x[size_t(i)] = double(i + 1);
  }
  std::unique_ptr<EllMatrixHip> K(upload(dev, n, 1, val, col, rn));
  K->Axy(x.data(), y.data());
  for (int64_t i = 0; i < n; ++i) EXPECT(std::fabs(y[size_t(i)] - double((i + 1) * (i + 1))) < 1e-3);
}

static void cg_diag_corner_kat(Device& dev) {
  const int64_t n = 32;
  // slot 0: diagonal; slot 1: the (0,1) / (1,0) coupling
  const size_t un = static_cast<size_t>(n);
  std::vector<double> val(2 * un, 0.0), b(un, 0.0), x(un, 0.0);
  std::vector<int32_t> col(2 * un, 0), rn(un, 1);
  for (int64_t i = 0; i < n; ++i) {
    val[size_t(i)] = double((i + 1) * (i + 1));
    col[size_t(i)] = int32_t(i);
    col[size_t(n + i)] = int32_t(i);
    b[size_t(i)] = double((i + 1) * (i + 1));
  }
  val[0] = 2.0;
  val[size_t(n)] = 2.0;      // row 0, col 1
  col[size_t(n)] = 1;
  val[size_t(n + 1)] = 2.0;  // row 1: cols 0 (slot 0) and 1 (slot 1)
  col[1] = 0;
  col[size_t(n + 1)] = 1;
  val[1] = 2.0;
  val[size_t(n + 1)] = 4.0;
  rn[0] = 2;
  rn[1] = 2;
  std::unique_ptr<EllMatrixHip> K(upload(dev, n, 2, val, col, rn));
  K->CG(x.data(), b.data(), 100, 1e-6);
  EXPECT(std::fabs(x[0] + 1.0) < 1e-3);
  EXPECT(std::fabs(x[1] - 1.5) < 1e-3);
  for (int64_t i = 2; i < n; ++i) EXPECT(std::fabs(x[size_t(i)] - 1.0) < 1e-3);
}

int main() {
  Device dev(0);
  stiffness_diag_positive(dev, GF_Q4, 8, 4, 0);
  std::printf("PASS macro_test_2dstiffness_diag_nonzero\n");
  stiffness_diag_positive(dev, GF_BRICK20, 3, 3, 2);
  std::printf("PASS macro_test_3dstiffness_diag_nonzero\n");
  penalty_increases_k00(dev);
  std::printf("PASS macro_test_stiffness_constraint\n");
  solve_plate(dev);
  std::printf("PASS solve_q8_plate_jacobi\n");
  axy_diagonal_kat(dev);
  std::printf("PASS Ell_diagonal_Ax_y_GPU\n");
  cg_diag_corner_kat(dev);
  std::printf("PASS Ell_CGGPU\n");
  return 0;
}
