/*
 * oracle.c -- FP64 CPU restatement of fpaboim/gpufem's assembly-and-solve
 * path (Gauss-quadrature Ke for Q4/Q8/Brick8/Brick20, colored scatter into
 * slot-major ELL, penalty constraints, CG and Jacobi-PCG).
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP product path is tested
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * What pins each function (checks that do not restate the formula):
 *   orc_gauss_points      closed forms: sum w = 2^dim; integrates x^k exactly
 *                         for k <= 2n-1 and fails at k = 2n.
 *   orc_constitutive      closed forms (plane stress / isotropic Lame form).
 *   orc_dshape            finite differences of the textbook Lagrange /
 *                         serendipity N_a at the nodal coordinates of
 *                         orc_node_natural_coords; partition of unity
 *                         (sum_a dN_a = 0) and linear (and for Q8/Brick20
 *                         quadratic) completeness.
 *   orc_jacobian/det/inverse  analytic J = (h/2) I on axis-aligned elements;
 *                         J Jinv = I on random affine maps.
 *   orc_element_stiffness closed-form rectangular Q4 plane-stress Ke
 *                         (textbook k1..k8 coefficients); symmetry; rigid
 *                         body null space (3 modes 2D, 6 modes 3D); rank;
 *                         linearity in E; agreement with a 5-point rule
 *                         on affine elements; constant-strain patch test.
 *   orc_ell_*             the reference's own known answers
 *                         (src/Test/SPRmatrixTest.cpp:139-436,
 *                         src/Test/SPRSearchTest.cpp:160-197) and brute
 *                         force against a dense matrix.
 *   orc_greedy_coloring   no two same-color elements share a node; brute
 *                         force first-fit on tiny meshes; structured grids
 *                         give 4 (2D) / 8 (3D) colors.
 *   orc_assemble          dense brute-force sum of element matrices.
 *   orc_apply_penalty     closed form on a hand-built matrix.
 *   orc_apply_elimination the solution of the eliminated system equals the
 *                         dense solve of the free-free block (scipy), and
 *                         the penalty solution to ~1e-9.
 *   orc_cg/orc_pcg_jacobi the reference's CG known answers
 *                         (src/Test/AxyGPUTest.cpp:201-367,
 *                         src/Test/LAopsTest.cpp:108-138) and scipy direct
 *                         solves of SPD systems.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Gauss rule: src/FEM/FemData.cc:226-316                                   */
/* ------------------------------------------------------------------------ */
int orc_gauss_points(int dim, int ngauss, double* pts, double* w) {
  double x[3], wt[3];
  if (ngauss == 2) {
    x[0] = -1.0 / sqrt(3.0);
    x[1] = -x[0];
    wt[0] = 1.0;
    wt[1] = 1.0;
  } else if (ngauss == 3) {
    x[0] = -sqrt(3.0 / 5.0);
    x[1] = 0.0;
    x[2] = -x[0];
    wt[0] = 5.0 / 9.0;
    wt[1] = 8.0 / 9.0;
    wt[2] = wt[0];
  } else {
    return 0;
  }
  int s = 0;
  if (dim == 3) {
    for (int i = 0; i < ngauss; ++i)
      for (int j = 0; j < ngauss; ++j)
        for (int k = 0; k < ngauss; ++k) {
          pts[3 * s + 0] = x[i];
          pts[3 * s + 1] = x[j];
          pts[3 * s + 2] = x[k];
          w[s] = (ngauss == 2) ? 1.0 : wt[i] * wt[j] * wt[k];
          ++s;
        }
  } else if (dim == 2) {
    for (int i = 0; i < ngauss; ++i)
      for (int j = 0; j < ngauss; ++j) {
        pts[2 * s + 0] = x[i];
        pts[2 * s + 1] = x[j];
        w[s] = (ngauss == 2) ? 1.0 : wt[i] * wt[j];
        ++s;
      }
  } else {
    return 0;
  }
  return s;
}

/* ------------------------------------------------------------------------ */
/* Constitutive matrix: src/FEM/FemData.cc:322-365                          */
/* The 2D branch actually executed by the reference is plane stress (its    */
/* analysis-type comparison never matches the quoted token, FemData.cc:330).*/
/* ------------------------------------------------------------------------ */
void orc_constitutive(int dim, double E, double nu, double* D) {
  if (dim == 2) {
    memset(D, 0, 9 * sizeof(double));
    double aux = E / (1.0 - nu * nu);
    double c1 = aux;
    double c2 = aux * ((1.0 - nu) / 2.0);
    double c3 = aux * nu;
    D[0 * 3 + 0] = c1;
    D[1 * 3 + 1] = c1;
    D[2 * 3 + 2] = c2;
    D[0 * 3 + 1] = c3;
    D[1 * 3 + 0] = c3;
  } else {
    memset(D, 0, 36 * sizeof(double));
    double c1 = E * (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double c2 = nu / (1.0 - nu);
    double c3 = (1.0 - 2.0 * nu) / (2.0 * (1.0 - nu));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) D[i * 6 + j] = (i == j) ? c1 : c1 * c2;
    D[3 * 6 + 3] = c1 * c3;
    D[4 * 6 + 4] = c1 * c3;
    D[5 * 6 + 5] = c1 * c3;
  }
}

/* ------------------------------------------------------------------------ */
/* Nodal natural coordinates in the reference node ordering.  These are    */
/* the points where the shape functions whose derivatives appear in        */
/* src/FEM/StiffAlgo.cc:47-202 equal one.                                   */
/* ------------------------------------------------------------------------ */
int orc_node_natural_coords(int dim, int nen, double* xi) {
  static const double q4[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  static const double q8[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {1, 0},
                                  {1, 1},   {0, 1},  {-1, 1}, {-1, 0}};
  static const double b8[8][3] = {{1, -1, -1},  {1, 1, -1},  {1, 1, 1},  {1, -1, 1},
                                  {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1}, {-1, -1, 1}};
  static const double b20[20][3] = {
      {1, -1, -1}, {1, 0, -1},  {1, 1, -1},   {1, 1, 0},    {1, 1, 1},
      {1, 0, 1},   {1, -1, 1},  {1, -1, 0},   {0, -1, -1},  {0, 1, -1},
      {0, 1, 1},   {0, -1, 1},  {-1, -1, -1}, {-1, 0, -1},  {-1, 1, -1},
      {-1, 1, 0},  {-1, 1, 1},  {-1, 0, 1},   {-1, -1, 1},  {-1, -1, 0}};
  const double* src = NULL;
  if (dim == 2 && nen == 4) src = &q4[0][0];
  else if (dim == 2 && nen == 8) src = &q8[0][0];
  else if (dim == 3 && nen == 8) src = &b8[0][0];
  else if (dim == 3 && nen == 20) src = &b20[0][0];
  else return 0;
  memcpy(xi, src, (size_t)dim * nen * sizeof(double));
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Shape-function derivatives: src/FEM/StiffAlgo.cc:47-202                  */
/* dN[i*nen + a] = dN_a / dxi_i                                             */
/* ------------------------------------------------------------------------ */
int orc_dshape(int dim, int nen, const double* rst, double* dN) {
  if (dim == 2 && nen == 4) {
    double r = rst[0], s = rst[1];
    double* dr = dN;
    double* ds = dN + 4;
    dr[0] = 0.25 * (s - 1);
    dr[1] = -0.25 * (s - 1);
    dr[2] = 0.25 * (s + 1);
    dr[3] = -0.25 * (s + 1);
    ds[0] = 0.25 * (r - 1);
    ds[1] = -0.25 * (r + 1);
    ds[2] = 0.25 * (r + 1);
    ds[3] = -0.25 * (r - 1);
    return 1;
  }
  if (dim == 2 && nen == 8) {
    double r = rst[0], s = rst[1];
    double* dr = dN;
    double* ds = dN + 8;
    dr[0] = -((2 * r + s) * (s - 1)) / 4;
    dr[1] = r * (s - 1);
    dr[2] = -((2 * r - s) * (s - 1)) / 4;
    dr[3] = 0.5 * (1 - s * s);
    dr[4] = ((2 * r + s) * (s + 1)) / 4;
    dr[5] = -r * (s + 1);
    dr[6] = ((2 * r - s) * (s + 1)) / 4;
    dr[7] = 0.5 * (s * s - 1);
    ds[0] = -((r + 2 * s) * (r - 1)) / 4;
    ds[1] = 0.5 * (r * r - 1);
    ds[2] = -((r - 2 * s) * (r + 1)) / 4;
    ds[3] = -s * (r + 1);
    ds[4] = ((r + 2 * s) * (r + 1)) / 4;
    ds[5] = 0.5 * (1 - r * r);
    ds[6] = ((r - 2 * s) * (r - 1)) / 4;
    ds[7] = s * (r - 1);
    return 1;
  }
  if (dim == 3 && nen == 8) {
    double r = rst[0], s = rst[1], t = rst[2];
    double* dr = dN;
    double* ds = dN + 8;
    double* dt = dN + 16;
    dr[0] = 0.125 * (s - 1) * (t - 1);
    dr[1] = -0.125 * (s + 1) * (t - 1);
    dr[2] = 0.125 * (s + 1) * (t + 1);
    dr[3] = -0.125 * (s - 1) * (t + 1);
    dr[4] = -0.125 * (s - 1) * (t - 1);
    dr[5] = 0.125 * (s + 1) * (t - 1);
    dr[6] = -0.125 * (s + 1) * (t + 1);
    dr[7] = 0.125 * (s - 1) * (t + 1);
    ds[0] = 0.125 * (r + 1) * (t - 1);
    ds[1] = -0.125 * (r + 1) * (t - 1);
    ds[2] = 0.125 * (r + 1) * (t + 1);
    ds[3] = -0.125 * (r + 1) * (t + 1);
    ds[4] = -0.125 * (r - 1) * (t - 1);
    ds[5] = 0.125 * (r - 1) * (t - 1);
    ds[6] = -0.125 * (r - 1) * (t + 1);
    ds[7] = 0.125 * (r - 1) * (t + 1);
    dt[0] = 0.125 * (r + 1) * (s - 1);
    dt[1] = -0.125 * (r + 1) * (s + 1);
    dt[2] = 0.125 * (r + 1) * (s + 1);
    dt[3] = -0.125 * (r + 1) * (s - 1);
    dt[4] = -0.125 * (r - 1) * (s - 1);
    dt[5] = 0.125 * (r - 1) * (s + 1);
    dt[6] = -0.125 * (r - 1) * (s + 1);
    dt[7] = 0.125 * (r - 1) * (s - 1);
    return 1;
  }
  if (dim == 3 && nen == 20) {
    double r = rst[0], s = rst[1], t = rst[2];
    double* dr = dN;
    double* ds = dN + 20;
    double* dt = dN + 40;
    /* corners */
    dr[0] = 0.125 * (1 - s) * (1 - t) * (2 * r - s - t - 1);
    dr[2] = 0.125 * (1 + s) * (1 - t) * (2 * r + s - t - 1);
    dr[4] = 0.125 * (1 + s) * (1 + t) * (2 * r + s + t - 1);
    dr[6] = 0.125 * (1 - s) * (1 + t) * (2 * r - s + t - 1);
    dr[12] = 0.125 * (1 - s) * (1 - t) * (2 * r + s + t + 1);
    dr[14] = 0.125 * (1 + s) * (1 - t) * (2 * r - s + t + 1);
    dr[16] = 0.125 * (1 + s) * (1 + t) * (2 * r - s - t + 1);
    dr[18] = 0.125 * (1 - s) * (1 + t) * (2 * r + s - t + 1);
    /* mid-edge */
    dr[1] = 0.25 * (1 - s * s) * (1 - t);
    dr[3] = 0.25 * (1 + s) * (1 - t * t);
    dr[5] = 0.25 * (1 - s * s) * (1 + t);
    dr[7] = 0.25 * (1 - s) * (1 - t * t);
    dr[8] = 0.25 * (-2 * r) * (1 - s) * (1 - t);
    dr[9] = 0.25 * (-2 * r) * (1 + s) * (1 - t);
    dr[10] = 0.25 * (-2 * r) * (1 + s) * (1 + t);
    dr[11] = 0.25 * (-2 * r) * (1 - s) * (1 + t);
    dr[13] = -0.25 * (1 - s * s) * (1 - t);
    dr[15] = -0.25 * (1 + s) * (1 - t * t);
    dr[17] = -0.25 * (1 - s * s) * (1 + t);
    dr[19] = -0.25 * (1 - s) * (1 - t * t);

    ds[0] = 0.125 * (1 + r) * (1 - t) * (-r + 2 * s + t + 1);
    ds[2] = 0.125 * (1 + r) * (1 - t) * (r + 2 * s - t - 1);
    ds[4] = 0.125 * (1 + r) * (1 + t) * (r + 2 * s + t - 1);
    ds[6] = 0.125 * (1 + r) * (1 + t) * (-r + 2 * s - t + 1);
    ds[12] = 0.125 * (1 - r) * (1 - t) * (r + 2 * s + t + 1);
    ds[14] = 0.125 * (1 - r) * (1 - t) * (-r + 2 * s - t - 1);
    ds[16] = 0.125 * (1 - r) * (1 + t) * (-r + 2 * s + t - 1);
    ds[18] = 0.125 * (1 - r) * (1 + t) * (r + 2 * s - t + 1);
    ds[1] = 0.25 * (1 + r) * (-2 * s) * (1 - t);
    ds[3] = 0.25 * (1 + r) * (1 - t * t);
    ds[5] = 0.25 * (1 + r) * (-2 * s) * (1 + t);
    ds[7] = -0.25 * (1 + r) * (1 - t * t);
    ds[8] = -0.25 * (1 - r * r) * (1 - t);
    ds[9] = 0.25 * (1 - r * r) * (1 - t);
    ds[10] = 0.25 * (1 - r * r) * (1 + t);
    ds[11] = -0.25 * (1 - r * r) * (1 + t);
    ds[13] = 0.25 * (1 - r) * (-2 * s) * (1 - t);
    ds[15] = 0.25 * (1 - r) * (1 - t * t);
    ds[17] = 0.25 * (1 - r) * (-2 * s) * (1 + t);
    ds[19] = -0.25 * (1 - r) * (1 - t * t);

    dt[0] = 0.125 * (1 + r) * (1 - s) * (-r + s + 2 * t + 1);
    dt[2] = 0.125 * (1 + r) * (1 + s) * (-r - s + 2 * t + 1);
    dt[4] = 0.125 * (1 + r) * (1 + s) * (r + s + 2 * t - 1);
    dt[6] = 0.125 * (1 + r) * (1 - s) * (r - s + 2 * t - 1);
    dt[12] = 0.125 * (1 - r) * (1 - s) * (r + s + 2 * t + 1);
    dt[14] = 0.125 * (1 - r) * (1 + s) * (r - s + 2 * t + 1);
    dt[16] = 0.125 * (1 - r) * (1 + s) * (-r + s + 2 * t - 1);
    dt[18] = 0.125 * (1 - r) * (1 - s) * (-r - s + 2 * t - 1);
    dt[1] = -0.25 * (1 + r) * (1 - s * s);
    dt[3] = 0.25 * (1 + r) * (1 + s) * (-2 * t);
    dt[5] = 0.25 * (1 + r) * (1 - s * s);
    dt[7] = 0.25 * (1 + r) * (1 - s) * (-2 * t);
    dt[8] = -0.25 * (1 - r * r) * (1 - s);
    dt[9] = -0.25 * (1 - r * r) * (1 + s);
    dt[10] = 0.25 * (1 - r * r) * (1 + s);
    dt[11] = 0.25 * (1 - r * r) * (1 - s);
    dt[13] = -0.25 * (1 - r) * (1 - s * s);
    dt[15] = 0.25 * (1 - r) * (1 + s) * (-2 * t);
    dt[17] = 0.25 * (1 - r) * (1 - s * s);
    dt[19] = 0.25 * (1 - r) * (1 - s) * (-2 * t);
    return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Jacobian, determinant, inverse                                            */
/* ------------------------------------------------------------------------ */
static double dot_seq(int n, const double* a, const double* b) {
  /* src/LAops/LAops.cc:45-52 */
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

void orc_jacobian(int dim, int nen, const double* xe, const double* dN, double* J) {
  /* src/FEM/StiffAlgo.cc:207-244: J[i][j] = dot(x_j, dN/dxi_i) */
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) J[i * dim + j] = dot_seq(nen, xe + j * nen, dN + i * nen);
}

double orc_det(int dim, const double* M) {
  /* src/LAops/LAops.cc:80-94 */
#define A(i, j) M[(i) * dim + (j)]
  if (dim == 2) return A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
  return A(0, 0) * A(1, 1) * A(2, 2) + A(1, 0) * A(2, 1) * A(0, 2) +
         A(2, 0) * A(0, 1) * A(1, 2) - A(0, 0) * A(2, 1) * A(1, 2) -
         A(1, 0) * A(0, 1) * A(2, 2) - A(2, 0) * A(1, 1) * A(0, 2);
#undef A
}

void orc_inverse(int dim, const double* M, double det, double* R) {
  /* src/LAops/LAops.cc:99-117 (adjugate / det) */
#define A(i, j) M[(i) * dim + (j)]
#define I(i, j) R[(i) * dim + (j)]
  if (dim == 2) {
    I(0, 0) = A(1, 1) / det;
    I(0, 1) = -A(0, 1) / det;
    I(1, 0) = -A(1, 0) / det;
    I(1, 1) = A(0, 0) / det;
    return;
  }
  I(0, 0) = (A(2, 2) * A(1, 1) - A(2, 1) * A(1, 2)) / det;
  I(0, 1) = -(A(2, 2) * A(0, 1) - A(2, 1) * A(0, 2)) / det;
  I(0, 2) = (A(1, 2) * A(0, 1) - A(1, 1) * A(0, 2)) / det;
  I(1, 0) = -(A(2, 2) * A(1, 0) - A(2, 0) * A(1, 2)) / det;
  I(1, 1) = (A(2, 2) * A(0, 0) - A(2, 0) * A(0, 2)) / det;
  I(1, 2) = -(A(1, 2) * A(0, 0) - A(1, 0) * A(0, 2)) / det;
  I(2, 0) = (A(2, 1) * A(1, 0) - A(2, 0) * A(1, 1)) / det;
  I(2, 1) = -(A(2, 1) * A(0, 0) - A(2, 0) * A(0, 1)) / det;
  I(2, 2) = (A(1, 1) * A(0, 0) - A(1, 0) * A(0, 1)) / det;
#undef A
#undef I
}

/* ------------------------------------------------------------------------ */
/* B matrix: src/FEM/StiffAlgo.cc:249-278.                                  */
/* 2D strain order (xx, yy, xy); 3D (xx, yy, zz, xy, yz, zx), engineering   */
/* shear.                                                                   */
/* ------------------------------------------------------------------------ */
void orc_bmatrix(int dim, int nen, const double* dNdx, double* B) {
  int ne = dim * nen;
  int s = (dim - 1) * 3;
  memset(B, 0, (size_t)s * ne * sizeof(double));
  const double* dx = dNdx;
  const double* dy = dNdx + nen;
  if (dim == 2) {
    for (int a = 0; a < nen; ++a) {
      B[0 * ne + 2 * a] = dx[a];
      B[1 * ne + 2 * a + 1] = dy[a];
      B[2 * ne + 2 * a] = dy[a];
      B[2 * ne + 2 * a + 1] = dx[a];
    }
  } else {
    const double* dz = dNdx + 2 * nen;
    for (int a = 0; a < nen; ++a) {
      B[0 * ne + 3 * a] = dx[a];
      B[1 * ne + 3 * a + 1] = dy[a];
      B[2 * ne + 3 * a + 2] = dz[a];
      B[3 * ne + 3 * a] = dy[a];
      B[3 * ne + 3 * a + 1] = dx[a];
      B[4 * ne + 3 * a + 1] = dz[a];
      B[4 * ne + 3 * a + 2] = dy[a];
      B[5 * ne + 3 * a] = dz[a];
      B[5 * ne + 3 * a + 2] = dx[a];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Element stiffness: the Gauss loop of src/FEM/StiffAlgoCPU.cc:233-255,    */
/* with dense matMult (src/LAops/LAops.cc:122-133) over the full B.         */
/* No thickness factor (the reference has none).                            */
/* ------------------------------------------------------------------------ */
#define ORC_MAXNE 60
#define ORC_MAXS 6
int orc_element_stiffness(int dim, int nen, int ngauss, const double* xe, double E,
                          double nu, double* Ke) {
  double pts[27 * 3], w[27];
  int ngp = orc_gauss_points(dim, ngauss, pts, w);
  if (ngp == 0) return 0;
  int ne = dim * nen;
  int s = (dim - 1) * 3;
  double D[36];
  orc_constitutive(dim, E, nu, D);
  double dN[3 * 20], J[9], Jinv[9], dNdx[3 * 20];
  static const int kB = ORC_MAXS * ORC_MAXNE;
  double B[kB], CxB[kB];
  double ktemp;
  memset(Ke, 0, (size_t)ne * ne * sizeof(double));
  int ok = 1;
  for (int g = 0; g < ngp; ++g) {
    if (!orc_dshape(dim, nen, pts + dim * g, dN)) return 0;
    orc_jacobian(dim, nen, xe, dN, J);
    double detJ = orc_det(dim, J);
    if (!(detJ > 0.0)) ok = -1;
    orc_inverse(dim, J, detJ, Jinv);
    /* dNdx = Jinv * dN (matMult, dim x dim x nen) */
    for (int i = 0; i < dim; ++i)
      for (int a = 0; a < nen; ++a) {
        double acc = 0.0;
        for (int k = 0; k < dim; ++k) acc += Jinv[i * dim + k] * dN[k * nen + a];
        dNdx[i * nen + a] = acc;
      }
    orc_bmatrix(dim, nen, dNdx, B);
    /* CxB = D * B  (s x s x ne) */
    for (int i = 0; i < s; ++i)
      for (int j = 0; j < ne; ++j) {
        double acc = 0.0;
        for (int k = 0; k < s; ++k) acc += D[i * s + k] * B[k * ne + j];
        CxB[i * ne + j] = acc;
      }
    /* k_temp = Bt * CxB (ne x s x ne); k_temp *= detJ*w; k_local += k_temp */
    double scale = detJ * w[g];
    for (int i = 0; i < ne; ++i)
      for (int j = 0; j < ne; ++j) {
        double acc = 0.0;
        for (int k = 0; k < s; ++k) acc += B[k * ne + i] * CxB[k * ne + j];
        ktemp = acc * scale;
        Ke[i * ne + j] = Ke[i * ne + j] + ktemp;
      }
  }
  return ok;
}

/* ------------------------------------------------------------------------ */
/* ELL storage: src/SPRmatrix/ELLmatrix.cc                                   */
/* ------------------------------------------------------------------------ */
orc_ell* orc_ell_create(int64_t n, int skip_zero) {
  /* constructor ELLmatrix.cc:44-54 + AllocateMatrix :533-545 (64 slots) */
  orc_ell* m = (orc_ell*)calloc(1, sizeof(orc_ell));
  if (!m) return NULL;
  m->n = n;
  m->maxrowlen = 64;
  m->growstep = 64;
  m->skip_zero = skip_zero;
  size_t cap = (size_t)m->maxrowlen * (size_t)n;
  m->val = (double*)calloc(cap ? cap : 1, sizeof(double));
  m->col = (int32_t*)malloc((cap ? cap : 1) * sizeof(int32_t));
  m->rownnz = (int32_t*)calloc(n ? (size_t)n : 1, sizeof(int32_t));
  if (!m->val || !m->col || !m->rownnz) {
    orc_ell_free(m);
    return NULL;
  }
  for (int64_t s = 0; s < m->maxrowlen; ++s)
    for (int64_t r = 0; r < n; ++r) m->col[r + s * n] = (int32_t)r;
  return m;
}

void orc_ell_free(orc_ell* m) {
  if (!m) return;
  free(m->val);
  free(m->col);
  free(m->rownnz);
  free(m);
}

static void ell_grow(orc_ell* m) {
  /* GrowMatrix, ELLmatrix.cc:513-529: slot-major, so growth appends slabs. */
  int32_t old = m->maxrowlen;
  m->maxrowlen += m->growstep;
  size_t cap = (size_t)m->maxrowlen * (size_t)m->n;
  m->val = (double*)realloc(m->val, cap * sizeof(double));
  m->col = (int32_t*)realloc(m->col, cap * sizeof(int32_t));
  for (int64_t s = old; s < m->maxrowlen; ++s)
    for (int64_t r = 0; r < m->n; ++r) {
      m->val[r + s * m->n] = 0.0;
      m->col[r + s * m->n] = (int32_t)r;
    }
}

void orc_ell_reserve(orc_ell* m, int32_t extra) {
  int32_t mx = orc_ell_max_rownnz(m);
  while (mx + extra >= m->maxrowlen) ell_grow(m);
}

int64_t orc_ell_binsearch_step(const int32_t* v, int32_t val, int32_t steppedlen,
                               int64_t startpos, int64_t step) {
  /* BinSearchIntStep, ELLmatrix.cc:575-599: lower bound over a strided view;
   * returns the absolute position of the first entry >= val, or -1 if all
   * entries are smaller (or the view is empty). */
  int32_t lo = 0, hi = steppedlen;
  if (hi == 0) return -1;
  const int32_t* base = v + startpos;
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    if (val > base[(int64_t)mid * step]) lo = mid + 1;
    else hi = mid;
  }
  if (lo == steppedlen) return -1;
  return startpos + (int64_t)lo * step;
}

static int input_ok(const orc_ell* m, int64_t row, int64_t col, double val) {
  /* SPRmatrix::InputIsOK / BoundsOK, src/SPRmatrix/SPRmatrix.cc:97-114 */
  if (m->skip_zero && val == 0.0) return 0;
  if (row < 0 || row >= m->n || col < 0 || col >= m->n) return 0;
  return 1;
}

static void ell_insert(orc_ell* m, int32_t rownnz, int64_t pos, double val, int64_t col,
                       int64_t row) {
  /* InsertElem, ELLmatrix.cc:116-138: shift the tail of the row one slot. */
  int64_t n = m->n;
  int64_t zeropos = row + (int64_t)rownnz * n;
  for (int64_t i = zeropos; i > pos; i -= n) {
    m->val[i] = m->val[i - n];
    m->col[i] = m->col[i - n];
  }
  m->val[pos] = val;
  m->col[pos] = (int32_t)col;
  m->rownnz[row]++;
  if (m->rownnz[row] == m->maxrowlen) ell_grow(m);
}

void orc_ell_set(orc_ell* m, int64_t row, int64_t col, double val) {
  if (!input_ok(m, row, col, val)) return;
  int32_t rn = m->rownnz[row];
  int64_t pos = orc_ell_binsearch_step(m->col, (int32_t)col, rn, row, m->n);
  if (pos == -1) {
    ell_insert(m, rn, row + m->n * rn, val, col, row);
    return;
  }
  if (m->col[pos] == col) m->val[pos] = val;
  else ell_insert(m, rn, pos, val, col, row);
}

void orc_ell_add(orc_ell* m, int64_t row, int64_t col, double val) {
  if (!input_ok(m, row, col, val)) return;
  int32_t rn = m->rownnz[row];
  int64_t pos = orc_ell_binsearch_step(m->col, (int32_t)col, rn, row, m->n);
  if (pos == -1) {
    ell_insert(m, rn, row + m->n * rn, val, col, row);
  } else if (m->col[pos] == col) {
    m->val[pos] += val;
  } else {
    ell_insert(m, rn, pos, val, col, row);
  }
}

double orc_ell_get(const orc_ell* m, int64_t row, int64_t col) {
  /* GetElem, ELLmatrix.cc:142-159 (columns are sorted, so a lower-bound
   * search finds the same entry the reference's linear scan finds). */
  if (row < 0 || row >= m->n || col < 0 || col >= m->n) return 0.0;
  int64_t pos = orc_ell_binsearch_step(m->col, (int32_t)col, m->rownnz[row], row, m->n);
  if (pos == -1 || m->col[pos] != col) return 0.0;
  return m->val[pos];
}

int32_t orc_ell_max_rownnz(const orc_ell* m) {
  int32_t mx = 0;
  for (int64_t r = 0; r < m->n; ++r)
    if (m->rownnz[r] > mx) mx = m->rownnz[r];
  return mx;
}

int64_t orc_ell_nnz(const orc_ell* m) {
  int64_t s = 0;
  for (int64_t r = 0; r < m->n; ++r) s += m->rownnz[r];
  return s;
}

void orc_ell_zero_values(orc_ell* m) {
  size_t cap = (size_t)m->maxrowlen * (size_t)m->n;
  memset(m->val, 0, cap * sizeof(double));
}

void orc_ell_axy(const orc_ell* m, const double* x, double* y, int nthreads) {
  /* Axy, ELLmatrix.cc:194-204: row-parallel, slots in order. */
  int64_t n = m->n;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double acc = 0.0;
    int32_t rn = m->rownnz[i];
    for (int32_t j = 0; j < rn; ++j) {
      int64_t idx = i + (int64_t)j * n;
      acc += m->val[idx] * x[m->col[idx]];
    }
    y[i] = acc;
  }
}

int orc_ell_export(const orc_ell* m, int32_t width, double* val, int32_t* col,
                   int32_t* rownnz) {
  int64_t n = m->n;
  if (width < orc_ell_max_rownnz(m)) return 0;
  for (int64_t r = 0; r < n; ++r) {
    int32_t rn = m->rownnz[r];
    rownnz[r] = rn;
    for (int32_t s = 0; s < width; ++s) {
      int64_t dst = r + (int64_t)s * n;
      if (s < rn) {
        val[dst] = m->val[r + (int64_t)s * n];
        col[dst] = m->col[r + (int64_t)s * n];
      } else {
        val[dst] = 0.0;
        col[dst] = (int32_t)r;
      }
    }
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Greedy coloring: src/FEM/femColor.cc:79-191                              */
/* ------------------------------------------------------------------------ */
int orc_greedy_coloring(int64_t nelem, int nen, const int32_t* conn, int64_t nnodes,
                        int32_t* elem_color) {
  /* node -> incident elements (CSR), the dual graph with ncommon = 1 */
  int64_t* ptr = (int64_t*)calloc((size_t)nnodes + 1, sizeof(int64_t));
  int64_t* adj = (int64_t*)malloc(((size_t)nelem * nen + 1) * sizeof(int64_t));
  int64_t* fill = (int64_t*)calloc((size_t)nnodes + 1, sizeof(int64_t));
  int64_t* nb = (int64_t*)malloc(((size_t)nen * 64 + 64) * sizeof(int64_t));
  int ncolors = 0;
  if (!ptr || !adj || !fill || !nb) { ncolors = -2; goto done; }
  for (int64_t e = 0; e < nelem; ++e)
    for (int a = 0; a < nen; ++a) ptr[conn[e * nen + a]]++;  /* 1-based -> slot node */
  for (int64_t v = 0; v < nnodes; ++v) ptr[v + 1] += ptr[v];
  /* ptr[v] now = count of nodes <= v (1-based index shifts by one) */
  for (int64_t e = 0; e < nelem; ++e)
    for (int a = 0; a < nen; ++a) {
      int64_t v = conn[e * nen + a] - 1;
      adj[ptr[v] + fill[v]++] = e;
    }
  for (int64_t e = 0; e < nelem; ++e) elem_color[e] = 0;
  for (int64_t e = 0; e < nelem; ++e) {
    int nn = 0;
    for (int a = 0; a < nen; ++a) {
      int64_t v = conn[e * nen + a] - 1;
      for (int64_t k = ptr[v]; k < ptr[v + 1]; ++k) {
        int64_t f = adj[k];
        if (f == e) continue;
        int dup = 0;
        for (int q = 0; q < nn; ++q)
          if (nb[q] == f) { dup = 1; break; }
        if (!dup) nb[nn++] = f;
      }
    }
    if (nn > 64) { ncolors = -1; goto done; }
    /* smallest color >= 1 not used by any neighbour (uncolored = 0) */
    int color = 1;
    for (;;) {
      int used = 0;
      for (int q = 0; q < nn; ++q)
        if (elem_color[nb[q]] == color) { used = 1; break; }
      if (!used) break;
      ++color;
    }
    elem_color[e] = color;
    if (color > ncolors) ncolors = color;
  }
done:
  free(ptr);
  free(adj);
  free(fill);
  free(nb);
  return ncolors;
}

/* ------------------------------------------------------------------------ */
/* Colored assembly: src/FEM/StiffAlgoCPU.cc:169-284, scatter              */
/* src/FEM/StiffAlgo.cc:342-401                                             */
/* ------------------------------------------------------------------------ */
int orc_assemble(orc_ell* K, int dim, int nen, int ngauss, int64_t nelem,
                 const int32_t* conn, const double* coords, const double* elem_E,
                 double E, double nu, const int32_t* elem_color, int ncolors,
                 int nthreads, double* seconds) {
  int ne = dim * nen;
  /* color buckets in increasing element id (femColor.cc:177-183) */
  int64_t* cptr = (int64_t*)calloc((size_t)ncolors + 2, sizeof(int64_t));
  int64_t* celem = (int64_t*)malloc(((size_t)nelem + 1) * sizeof(int64_t));
  int64_t* cfill = (int64_t*)calloc((size_t)ncolors + 2, sizeof(int64_t));
  if (!cptr || !celem || !cfill) { free(cptr); free(celem); free(cfill); return -2; }
  for (int64_t e = 0; e < nelem; ++e) cptr[elem_color[e]]++;
  for (int c = 1; c <= ncolors + 1; ++c) cptr[c] += cptr[c - 1];
  for (int64_t e = 0; e < nelem; ++e) {
    int c = elem_color[e];
    celem[cptr[c - 1] + cfill[c]++] = e;
  }
  int status = 1;
  double t0 = omp_get_wtime();
  for (int c = 1; c <= ncolors; ++c) {
    /* growth happens only here, at a serial point (see GrowMatrix) */
    orc_ell_reserve(K, ne);
    int64_t b = cptr[c - 1], en = cptr[c];
#pragma omp parallel num_threads(nthreads)
    {
      double* Ke = (double*)malloc((size_t)ne * ne * sizeof(double));
      double xe[3 * 20];
#pragma omp for schedule(static)
      for (int64_t p = b; p < en; ++p) {
        int64_t e = celem[p];
        for (int a = 0; a < nen; ++a)
          for (int i = 0; i < dim; ++i)
            xe[i * nen + a] = coords[(int64_t)(conn[e * nen + a] - 1) * dim + i];
        double Ee = elem_E ? elem_E[e] : E;
        int st = orc_element_stiffness(dim, nen, ngauss, xe, Ee, nu, Ke);
        if (st != 1) {
#pragma omp atomic write
          status = st;
        }
        for (int a = 0; a < nen; ++a) {
          int64_t gi = (int64_t)(conn[e * nen + a] - 1) * dim;
          for (int bb = 0; bb < nen; ++bb) {
            int64_t gj = (int64_t)(conn[e * nen + bb] - 1) * dim;
            for (int cj = 0; cj < dim; ++cj)
              for (int ci = 0; ci < dim; ++ci)
                orc_ell_add(K, gi + ci, gj + cj, Ke[(dim * a + ci) * ne + dim * bb + cj]);
          }
        }
      }
      free(Ke);
    }
  }
  double t1 = omp_get_wtime();
  if (seconds) *seconds = t1 - t0;
  free(cptr);
  free(celem);
  free(cfill);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Penalty: src/FEM/fem.cc:112-145                                            */
/* ------------------------------------------------------------------------ */
double orc_apply_penalty(orc_ell* K, int dim, int64_t nsup, const int32_t* supports,
                         double coef) {
  double maxdiag = orc_ell_get(K, 0, 0);
  for (int64_t i = 0; i < K->n; ++i) {
    double v = orc_ell_get(K, i, i);
    if (v > maxdiag) maxdiag = v;
  }
  double pen = maxdiag * coef;
  for (int64_t i = 0; i < nsup; ++i) {
    const int32_t* row = supports + i * (dim + 1);
    for (int j = 1; j < 1 + dim; ++j) {
      if (row[j] == 1) {
        int64_t dof = (int64_t)dim * (row[0] - 1) + (j - 1);
        orc_ell_set(K, dof, dof, pen * orc_ell_get(K, dof, dof));
      }
    }
  }
  return pen;
}

/* ------------------------------------------------------------------------ */
/* Zero substitution: the intent of FEM::ApplyConstraint(SUB)               */
/* (src/FEM/fem.cc:146-163): row and column of every fixed DOF are zeroed   */
/* and its diagonal set to 1.  (The reference's SetElem(.., 0) calls are    */
/* no-ops because InputIsOK rejects zeros, SPRmatrix.cc:97-102, so its SUB  */
/* only sets the diagonal; this is the elimination the comment describes.)  */
/* Stored entries stay stored (their values become 0).                      */
/* ------------------------------------------------------------------------ */
void orc_apply_elimination(orc_ell* K, int dim, int64_t nsup, const int32_t* supports) {
  unsigned char* fixed = (unsigned char*)calloc((size_t)K->n, 1);
  for (int64_t i = 0; i < nsup; ++i) {
    const int32_t* row = supports + i * (dim + 1);
    for (int j = 1; j < 1 + dim; ++j)
      if (row[j] == 1) fixed[(int64_t)dim * (row[0] - 1) + (j - 1)] = 1;
  }
  for (int64_t r = 0; r < K->n; ++r)
    for (int32_t s = 0; s < K->rownnz[r]; ++s) {
      const int64_t idx = r + (int64_t)s * K->n;
      const int64_t c = K->col[idx];
      if (fixed[r] || fixed[c]) K->val[idx] = (r == c) ? 1.0 : 0.0;
    }
  free(fixed);
}

/* ------------------------------------------------------------------------ */
/* Force vector: src/FEM/FemData.cc:370-385 (assignment, not accumulation)  */
/* ------------------------------------------------------------------------ */
void orc_force_vector(int dim, int64_t nnodes, int64_t nload, const double* loads,
                      double* f) {
  memset(f, 0, (size_t)nnodes * dim * sizeof(double));
  for (int64_t i = 0; i < nload; ++i) {
    const double* row = loads + i * (dim + 1);
    for (int j = 1; j < dim + 1; ++j) {
      if (row[j] != 0.0) {
        int64_t dof = (int64_t)dim * ((int64_t)row[0] - 1) + (j - 1);
        f[dof] = row[j];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* BLAS-1 and solvers                                                        */
/* ------------------------------------------------------------------------ */
#define ORC_DOT_BLOCK 4096
double orc_dot(int64_t n, const double* a, const double* b, int nthreads) {
  int64_t nb = (n + ORC_DOT_BLOCK - 1) / ORC_DOT_BLOCK;
  if (nb == 0) return 0.0;
  double* part = (double*)malloc((size_t)nb * sizeof(double));
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t k = 0; k < nb; ++k) {
    int64_t lo = k * ORC_DOT_BLOCK;
    int64_t hi = lo + ORC_DOT_BLOCK < n ? lo + ORC_DOT_BLOCK : n;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s += a[i] * b[i];
    part[k] = s;
  }
  double s = 0.0;
  for (int64_t k = 0; k < nb; ++k) s += part[k];
  free(part);
  return s;
}

int orc_cg(const orc_ell* K, const double* b, double* x, int maxit, double eps,
           int nthreads, double* seconds) {
  /* SPRmatrix::CPU_CG, src/SPRmatrix/SPRmatrix.cc:405-480 */
  int64_t n = K->n;
  double* d = (double*)malloc((size_t)n * sizeof(double));
  double* r = (double*)malloc((size_t)n * sizeof(double));
  double* q = (double*)malloc((size_t)n * sizeof(double));
  double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    x[i] = 0.0;
    d[i] = b[i];
    r[i] = b[i];
    q[i] = 0.0;
  }
  double dnew = orc_dot(n, r, r, nthreads);
  double bound = eps * eps * dnew;
  int it;
  for (it = 0; it < maxit && dnew > bound; ++it) {
    orc_ell_axy(K, d, q, nthreads);
    double alpha = dnew / orc_dot(n, d, q, nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] = alpha * d[i] + x[i];
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = (-alpha) * q[i] + r[i];
    double dold = dnew;
    dnew = orc_dot(n, r, r, nthreads);
    double beta = dnew / dold;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) d[i] = r[i] + beta * d[i];
  }
  if (seconds) *seconds = omp_get_wtime() - t0;
  free(d);
  free(r);
  free(q);
  return it;
}

int orc_pcg_jacobi(const orc_ell* K, const double* b, double* x, int maxit, double rtol,
                   int use_x0, int nthreads, double* true_relres, double* seconds,
                   int progress_every) {
  /* Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 9.1,
   * with M = diag(K).  The vector passes mirror CPU_CG's separate BLAS-1
   * loops (src/SPRmatrix/SPRmatrix.cc:432-451). */
  int64_t n = K->n;
  double* dinv = (double*)malloc((size_t)n * sizeof(double));
  double* r = (double*)malloc((size_t)n * sizeof(double));
  double* z = (double*)malloc((size_t)n * sizeof(double));
  double* p = (double*)malloc((size_t)n * sizeof(double));
  double* q = (double*)malloc((size_t)n * sizeof(double));
  double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) dinv[i] = 1.0 / orc_ell_get(K, i, i);
  if (use_x0) {
    orc_ell_axy(K, x, q, nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  } else {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] = 0.0;
      r[i] = b[i];
    }
  }
  double bnorm = sqrt(orc_dot(n, b, b, nthreads));
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
  }
  double rho = orc_dot(n, r, z, nthreads);
  double rr = orc_dot(n, r, r, nthreads);
  int it = 0;
  while (it < maxit && sqrt(rr) > rtol * bnorm) {
    orc_ell_axy(K, p, q, nthreads);
    double alpha = rho / orc_dot(n, p, q, nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] = x[i] + alpha * p[i];
      r[i] = r[i] - alpha * q[i];
      z[i] = dinv[i] * r[i];
    }
    double rho_new = orc_dot(n, r, z, nthreads);
    rr = orc_dot(n, r, r, nthreads);
    double beta = rho_new / rho;
    rho = rho_new;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
    if (progress_every > 0 && it % progress_every == 0) {
      /* long host solves report as they go */
      fprintf(stderr, "[oracle] pcg iteration %d, %.1f s, residual %.3e of |b|\n", it,
              omp_get_wtime() - t0, bnorm > 0 ? sqrt(rr) / bnorm : sqrt(rr));
      fflush(stderr);
    }
  }
  if (seconds) *seconds = omp_get_wtime() - t0;
  if (true_relres) {
    orc_ell_axy(K, x, q, nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) q[i] = b[i] - q[i];
    double rn = sqrt(orc_dot(n, q, q, nthreads));
    *true_relres = bnorm > 0 ? rn / bnorm : rn;
  }
  free(dinv);
  free(r);
  free(z);
  free(p);
  free(q);
  return it;
}
