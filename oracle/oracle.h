/*
 * oracle.h -- FP64 CPU restatement of fpaboim/gpufem's assembly-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the HIP product path
 * is tested against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  It is never linked into, or called by, the
 * product library.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic is IEEE double throughout; the
 * reference computes in float (src/Utils/util.h:11), so this is the reference
 * algorithm in double precision.
 *
 * Pinning (which independent check fixes each function) is listed in the
 * header comment of oracle.c.
 */
#ifndef GF_ORACLE_H_
#define GF_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- element kinematics -------------------------------------------------- */

/* Tensor Gauss rule, src/FEM/FemData.cc:226-316.  Point s = i*n^2 + j*n + k
 * (first coordinate slowest).  pts is [ngp][dim], w is [ngp].  Returns ngp,
 * or 0 for an unsupported rule (ngauss not 2 or 3). */
int orc_gauss_points(int dim, int ngauss, double* pts, double* w);

/* Constitutive matrix, src/FEM/FemData.cc:322-365.  2D: plane stress (the
 * branch the reference actually executes).  3D: isotropic.  D is s x s
 * row-major with s = 3 (2D) or 6 (3D). */
void orc_constitutive(int dim, double E, double nu, double* D);

/* dN_a/dxi_i in natural coordinates, src/FEM/StiffAlgo.cc:47-202.
 * dN is [dim][nen] row-major.  Returns 0 for an unsupported (dim, nen). */
int orc_dshape(int dim, int nen, const double* rst, double* dN);

/* Natural coordinates of the element nodes in the reference ordering
 * (derived from the formulas of src/FEM/StiffAlgo.cc:47-202). xi is [nen][dim]. */
int orc_node_natural_coords(int dim, int nen, double* xi);

/* J_ij = sum_a dN_a/dxi_i x_a,j; src/FEM/StiffAlgo.cc:207-244.
 * xe is [dim][nen] (coordinate-major, as m_elem_coords). */
void orc_jacobian(int dim, int nen, const double* xe, const double* dN, double* J);

/* src/LAops/LAops.cc:80-94 and :99-117 */
double orc_det(int dim, const double* J);
void orc_inverse(int dim, const double* J, double det, double* Jinv);

/* B matrix (strain-displacement), src/FEM/StiffAlgo.cc:249-278.
 * dNdx is [dim][nen]; B is [s][dim*nen], fully overwritten. */
void orc_bmatrix(int dim, int nen, const double* dNdx, double* B);

/* Element stiffness Ke = sum_gp B^T (D B) detJ w, dense products in the
 * reference order, src/FEM/StiffAlgoCPU.cc:210-255.  xe is [dim][nen].
 * Ke is [ne][ne] row-major, ne = dim*nen.  Returns 0 on unsupported
 * element, -1 if some detJ <= 0, else 1. */
int orc_element_stiffness(int dim, int nen, int ngauss, const double* xe,
                          double E, double nu, double* Ke);

/* ---- ELL storage with the reference's insertion semantics ---------------- */

/* Slot-major ELLPACK: val[row + slot*n], col[row + slot*n] sorted ascending
 * within a row, rownnz[n]; src/SPRmatrix/ELLmatrix.h:35-98.  Padding slots
 * hold val 0 and col = row. */
typedef struct orc_ell {
  int64_t n;
  int32_t maxrowlen;  /* allocated slots per row (grows by growstep) */
  int32_t growstep;
  int32_t skip_zero;  /* 1: drop val == 0 inputs as InputIsOK does */
  int32_t pad_;
  double* val;
  int32_t* col;
  int32_t* rownnz;
} orc_ell;

orc_ell* orc_ell_create(int64_t n, int skip_zero);
void     orc_ell_free(orc_ell* m);
/* src/SPRmatrix/ELLmatrix.cc:575-599 */
int64_t  orc_ell_binsearch_step(const int32_t* v, int32_t val, int32_t steppedlen,
                                int64_t startpos, int64_t step);
/* src/SPRmatrix/ELLmatrix.cc:68-112 (Set/Add), :116-138 (Insert), :142-159 (Get) */
void     orc_ell_set(orc_ell* m, int64_t row, int64_t col, double val);
void     orc_ell_add(orc_ell* m, int64_t row, int64_t col, double val);
double   orc_ell_get(const orc_ell* m, int64_t row, int64_t col);
/* Ensure room for `extra` more entries in every row (GrowMatrix,
 * src/SPRmatrix/ELLmatrix.cc:513-529, applied at a serial point). */
void     orc_ell_reserve(orc_ell* m, int32_t extra);
int32_t  orc_ell_max_rownnz(const orc_ell* m);
int64_t  orc_ell_nnz(const orc_ell* m);
/* Zero every stored value, keep the structure (numeric re-assembly). */
void     orc_ell_zero_values(orc_ell* m);
/* y = A x, src/SPRmatrix/ELLmatrix.cc:194-204 */
void     orc_ell_axy(const orc_ell* m, const double* x, double* y, int nthreads);
/* Copy out compacted to width `width` (>= max rownnz), padding col = row. */
int      orc_ell_export(const orc_ell* m, int32_t width, double* val, int32_t* col,
                        int32_t* rownnz);

/* ---- mesh coloring -------------------------------------------------------- */

/* Greedy first-fit coloring of the element dual graph (elements sharing >= 1
 * node), in element order; src/FEM/femColor.cc:79-191 (adjacency restated
 * without METIS).  conn is 1-based [nelem][nen].  elem_color gets colors
 * 1..ncolors.  Returns ncolors, or -1 if an element has > 64 neighbours
 * (MAXNEIGHBORS, src/FEM/femColor.cc:40). */
int orc_greedy_coloring(int64_t nelem, int nen, const int32_t* conn, int64_t nnodes,
                        int32_t* elem_color);

/* ---- assembly, constraints, loads ---------------------------------------- */

/* Colored global assembly, src/FEM/StiffAlgoCPU.cc:169-284 with the scatter
 * of src/FEM/StiffAlgo.cc:342-401.  elem_E may be NULL (uniform E).
 * elem_color from orc_greedy_coloring.  Elements of one color are processed
 * in increasing id by `nthreads` OpenMP threads.  Returns 1, or <= 0 on
 * failure (see orc_element_stiffness).  *seconds gets omp_get_wtime span. */
int orc_assemble(orc_ell* K, int dim, int nen, int ngauss, int64_t nelem,
                 const int32_t* conn, const double* coords, const double* elem_E,
                 double E, double nu, const int32_t* elem_color, int ncolors,
                 int nthreads, double* seconds);

/* Penalty constraint, src/FEM/fem.cc:112-145: pen = coef * max_i K_ii;
 * K_dd <- pen * K_dd for each fixed dof.  supports is [nsup][dim+1] with
 * (node 1-based, fix_x, fix_y[, fix_z]).  Returns the penalty factor. */
double orc_apply_penalty(orc_ell* K, int dim, int64_t nsup, const int32_t* supports,
                         double coef);

/* Force vector by assignment, src/FEM/FemData.cc:370-385.  loads is
 * [nload][dim+1] with (node 1-based, Fx, Fy[, Fz]).  f is zeroed first. */
void orc_apply_elimination(orc_ell* K, int dim, int64_t nsup, const int32_t* supports);
void orc_force_vector(int dim, int64_t nnodes, int64_t nload, const double* loads,
                      double* f);

/* ---- solvers --------------------------------------------------------------- */

/* Fixed-order dot product (blocked sums of 4096, blocks summed in order);
 * replaces the order-dependent OpenMP reduction of src/LAops/LAops.cc:54-61. */
double orc_dot(int64_t n, const double* a, const double* b, int nthreads);

/* Unpreconditioned CG, src/SPRmatrix/SPRmatrix.cc:405-480: x0 = 0, stop when
 * r.r <= eps^2 r0.r0 or after maxit iterations.  Returns iterations. */
int orc_cg(const orc_ell* K, const double* b, double* x, int maxit, double eps,
           int nthreads, double* seconds);

/* Jacobi-preconditioned CG (Saad, Iterative Methods, Alg. 9.1).  If
 * use_x0 is 0, x starts at zero.  Stops when ||r||_2 <= rtol ||b||_2
 * (recursive residual) or after maxit iterations.  Returns iterations;
 * *true_relres gets ||b - K x|| / ||b|| at exit.  progress_every > 0: a line
 * on stderr every that many iterations. */
int orc_pcg_jacobi(const orc_ell* K, const double* b, double* x, int maxit, double rtol,
                   int use_x0, int nthreads, double* true_relres, double* seconds,
                   int progress_every);

#ifdef __cplusplus
}
#endif
#endif /* GF_ORACLE_H_ */
