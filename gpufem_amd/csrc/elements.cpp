// Element catalogue: node orderings, shape-function derivatives, Gauss rules
// and elastic coefficients used to build the device tables.
//
// The node orderings are the reference's (they follow from
// src/FEM/StiffAlgo.cc:47-202).  The derivatives are derived here from the
// textbook Lagrange / serendipity shape functions written in terms of the
// nodal natural coordinates, rather than spelled out per node.
#include <cmath>
#include <cstring>

#include "gf_internal.hpp"

namespace gf {
namespace {

const double kQ4[4 * 2] = {-1, -1, 1, -1, 1, 1, -1, 1};
const double kQ8[8 * 2] = {-1, -1, 0, -1, 1, -1, 1, 0, 1, 1, 0, 1, -1, 1, -1, 0};
const double kB8[8 * 3] = {1,  -1, -1, 1,  1, -1, 1,  1, 1, 1,  -1, 1,
                           -1, -1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1};
// Brick20: corners at even positions 0,2,4,6 (r=+1 face) and 12,14,16,18
// (r=-1 face); mid-edge nodes elsewhere.
const double kB20[20 * 3] = {
    1,  -1, -1, 1,  0,  -1, 1,  1,  -1, 1,  1,  0,  1,  1,  1,  1,  0,  1,  1,  -1,
    1,  1,  -1, 0,  0,  -1, -1, 0,  1,  -1, 0,  1,  1,  0,  -1, 1,  -1, -1, -1, -1,
    0,  -1, -1, 1,  -1, -1, 1,  0,  -1, 1,  1,  -1, 0,  1,  -1, -1, 1,  -1, -1, 0};

}  // namespace

bool elem_info(int type, ElemInfo* out) {
  switch (type) {
    case GF_Q4: *out = {GF_Q4, 2, 4}; return true;
    case GF_Q8: *out = {GF_Q8, 2, 8}; return true;
    case GF_BRICK8: *out = {GF_BRICK8, 3, 8}; return true;
    case GF_BRICK20: *out = {GF_BRICK20, 3, 20}; return true;
    default: return false;
  }
}

const double* elem_natural_coords(int type) {
  switch (type) {
    case GF_Q4: return kQ4;
    case GF_Q8: return kQ8;
    case GF_BRICK8: return kB8;
    case GF_BRICK20: return kB20;
    default: return nullptr;
  }
}

// Multilinear:   N_a = prod_i (1 + x_i a_i) / 2^d
// Serendipity corner (all |a_i| = 1):
//   2D: N_a = (1/4) prod_i (1 + x_i a_i) (sum_i x_i a_i - 1)
//   3D: N_a = (1/8) prod_i (1 + x_i a_i) (sum_i x_i a_i - 2)
// Serendipity mid-side / mid-edge (a_m = 0 for one m):
//   N_a = (1/2^(d-1)) (1 - x_m^2) prod_{i != m} (1 + x_i a_i)
void shape_derivs(int type, const double* x, double* dN) {
  ElemInfo ei;
  if (!elem_info(type, &ei)) return;
  const int d = ei.dim, nen = ei.nen;
  const double* nat = elem_natural_coords(type);
  const bool serendipity = (type == GF_Q8 || type == GF_BRICK20);
  for (int a = 0; a < nen; ++a) {
    const double* an = nat + a * d;
    double lin[3];
    for (int i = 0; i < d; ++i) lin[i] = 1.0 + x[i] * an[i];
    int mid = -1;
    for (int i = 0; i < d; ++i)
      if (an[i] == 0.0) mid = i;
    for (int k = 0; k < d; ++k) {
      double v;
      if (!serendipity) {
        double p = an[k];
        for (int i = 0; i < d; ++i)
          if (i != k) p *= lin[i];
        v = p / double(1 << d);
      } else if (mid < 0) {
        // d/dx_k [prod (1+x_i a_i) (S - (d-1))] with S = sum x_i a_i
        //   = a_k prod_{i!=k}(1+x_i a_i) (S - (d-1) + 1 + x_k a_k)
        double S = 0.0;
        for (int i = 0; i < d; ++i) S += x[i] * an[i];
        double p = an[k];
        for (int i = 0; i < d; ++i)
          if (i != k) p *= lin[i];
        v = p * (S + x[k] * an[k] - double(d - 2)) / double(1 << d);
      } else {
        double scale = 1.0 / double(1 << (d - 1));
        if (k == mid) {
          double p = -2.0 * x[mid];
          for (int i = 0; i < d; ++i)
            if (i != mid) p *= lin[i];
          v = p * scale;
        } else {
          double p = (1.0 - x[mid] * x[mid]) * an[k];
          for (int i = 0; i < d; ++i)
            if (i != mid && i != k) p *= lin[i];
          v = p * scale;
        }
      }
      dN[k * nen + a] = v;
    }
  }
}

int gauss_rule(int dim, int ngauss, double* pts, double* w) {
  double x[3], wt[3];
  if (ngauss == 2) {
    x[0] = -1.0 / std::sqrt(3.0);
    x[1] = -x[0];
    wt[0] = wt[1] = 1.0;
  } else if (ngauss == 3) {
    x[0] = -std::sqrt(0.6);
    x[1] = 0.0;
    x[2] = -x[0];
    wt[0] = wt[2] = 5.0 / 9.0;
    wt[1] = 8.0 / 9.0;
  } else {
    return 0;
  }
  int n = 0;
  if (dim == 2) {
    for (int i = 0; i < ngauss; ++i)
      for (int j = 0; j < ngauss; ++j) {
        pts[2 * n] = x[i];
        pts[2 * n + 1] = x[j];
        w[n++] = wt[i] * wt[j];
      }
  } else {
    for (int i = 0; i < ngauss; ++i)
      for (int j = 0; j < ngauss; ++j)
        for (int k = 0; k < ngauss; ++k) {
          pts[3 * n] = x[i];
          pts[3 * n + 1] = x[j];
          pts[3 * n + 2] = x[k];
          w[n++] = wt[i] * wt[j] * wt[k];
        }
  }
  return n;
}

void elastic_coeffs(int dim, double nu, double c[3]) {
  if (dim == 2) {
    double a = 1.0 / (1.0 - nu * nu);
    c[0] = a;
    c[1] = a * nu;
    c[2] = a * (1.0 - nu) / 2.0;
  } else {
    double c1 = (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
    c[0] = c1;
    c[1] = c1 * (nu / (1.0 - nu));
    c[2] = c1 * ((1.0 - 2.0 * nu) / (2.0 * (1.0 - nu)));
  }
}

int host_threads() {
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 4;
  if (hc > 16) hc = 16;
  return int(hc);
}

}  // namespace gf
