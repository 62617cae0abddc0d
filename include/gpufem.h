/*
 * gpufem.h -- C-ABI of the MI355X (gfx950) assembly-and-solve engine.
 *
 * Plain C types only: handles are opaque, arrays are caller-owned host
 * pointers (copied in or out) unless a function says it takes device
 * pointers.  Every function returns GF_OK (0) or a negative GF_ERR_* code;
 * gf_last_error() gives the message.  One gf_ctx owns one HIP stream and is
 * not thread-safe; use one context per host thread.  No function changes
 * global OpenMP or HIP device state beyond selecting its own device.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   - StiffAlgo::CalcGlobalStiffness(FemData*)   src/FEM/StiffAlgo.h:46
 *     (GPU strategies StiffAlgoGPU / StiffAlgoGpuOmp,
 *      src/FEM/StiffAlgoGPU.cc:49-178, src/FEM/StiffAlgoGpuOmp.cc:49-204)
 *       -> gf_mesh_upload + gf_ell_symbolic + gf_assemble
 *   - FEM::ApplyConstraint(PEN)                   src/FEM/fem.cc:112-145
 *       -> gf_apply_penalty
 *   - ELLmatrix::AxyGPU                           src/SPRmatrix/ELLmatrix.cc:209-285
 *       -> gf_spmv
 *   - ELLmatrix::CG / SolveCgGpu                  src/SPRmatrix/ELLmatrix.cc:289-459
 *     (SPRmatrix::CPU_CG semantics, src/SPRmatrix/SPRmatrix.cc:405-480)
 *       -> gf_pcg
 *   - ELL arrays m_matdata/m_colidx/m_rownnz      src/SPRmatrix/ELLmatrix.h:82-90
 *       -> gf_ell_download / gf_ell_upload
 *   - femColor::MakeGreedyColoring                src/FEM/femColor.cc:132-191
 *       -> computed inside gf_ell_symbolic, read back by gf_ell_colors
 */
#ifndef GPUFEM_H_
#define GPUFEM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define GF_OK 0
#define GF_ERR_ARG (-1)         /* bad argument (null, out of range, mismatch) */
#define GF_ERR_HIP (-2)         /* a HIP runtime call failed */
#define GF_ERR_NOMEM (-3)       /* host or device allocation failed */
#define GF_ERR_UNSUPPORTED (-4) /* element type / rule / size not supported */
#define GF_ERR_DEGENERATE (-5)  /* an element has detJ <= 0 at a Gauss point */
#define GF_ERR_BREAKDOWN (-6)   /* solver breakdown (p.q <= 0 or NaN) */
#define GF_ERR_NOCONV (-7)      /* maxit reached before the tolerance */

/* ---- element types (reference node orderings, src/FEM/StiffAlgo.cc:47-202) */
#define GF_Q4 1
#define GF_Q8 2
#define GF_BRICK8 3
#define GF_BRICK20 4

/* ---- opaque handles ------------------------------------------------------ */
typedef struct gf_ctx gf_ctx;
typedef struct gf_mesh gf_mesh;
typedef struct gf_ell gf_ell;

/* Device-side timing of one call, measured with HIP events recorded on the
 * context's stream around the call's kernels. */
typedef struct gf_timing {
  double ms_device;  /* event-to-event time on the stream (kernels only) */
  double ms_host;    /* host wall time of the call (includes copies) */
  int64_t launches;  /* kernel launches issued */
} gf_timing;

/* ======================================================================== */
/* Structured synthetic meshes (host only; no GPU needed).                  */
/* Nodes are numbered lexicographically (x fastest, then y, then z) over    */
/* the points the element type uses; connectivity is 1-based in the         */
/* reference node ordering; DOF = dim*(node-1)+c.                           */
/* No reference counterpart: the reference reads neutral files              */
/* (src/Utils/fileIO.cc:73-404); these grids stand in for them.             */
/* ======================================================================== */
typedef struct gf_grid_desc {
  int32_t elem_type; /* GF_Q4 .. GF_BRICK20 */
  int32_t nx, ny, nz; /* elements per axis (nz ignored in 2D) */
  double h;           /* element edge length */
} gf_grid_desc;

int gf_grid_sizes(const gf_grid_desc* g, int64_t* nnodes, int64_t* nelem, int32_t* nen,
                  int32_t* dim);
/* conn: [nelem][nen] 1-based; coords: [nnodes][dim] */
int gf_grid_fill(const gf_grid_desc* g, int32_t* conn, double* coords);
/* Nodes (1-based, ascending) on a boundary face: 0:x=0 1:x=max 2:y=0 3:y=max
 * 4:z=0 5:z=max.  Pass nodes = NULL to get the count.  Returns the count or
 * a negative error. */
int64_t gf_grid_face_nodes(const gf_grid_desc* g, int32_t face, int32_t* nodes);
/* Adds the consistent nodal loads of a uniform traction q (force per unit
 * edge length in 2D, per unit area in 3D) in direction comp (0:x 1:y 2:z)
 * on a boundary face into f[nnodes*dim]. */
int gf_grid_face_load(const gf_grid_desc* g, int32_t face, int32_t comp, double q,
                      double* f);
/* Deterministic streams (SplitMix64 + Box-Muller). */
int gf_rng_normal(uint64_t seed, int64_t n, double* out);
int gf_rng_uniform_int(uint64_t seed, int64_t n, int32_t lo, int32_t hi, int32_t* out);
/* Unit-variance Gaussian random field on an nx x ny x nz cell grid (x
 * fastest; nz = 1 for 2D): white noise from gf_rng_normal(seed) on the grid
 * extended by ceil(3 lengthscale) cells per side of every axis longer than 1,
 * smoothed by a separable Gaussian of standard deviation `lengthscale` cells
 * whose per-axis weights have unit sum of squares.  No reference counterpart
 * (random modulus and load fields of the learned warm-start family). */
int gf_grf(uint64_t seed, int64_t nx, int64_t ny, int64_t nz, double lengthscale, double* out);

/* ======================================================================== */
/* Context                                                                   */
/* ======================================================================== */
int gf_create(gf_ctx** out, int32_t device);
int gf_destroy(gf_ctx* ctx);
/* Message of the last failure on ctx (ctx may be NULL: last global error). */
const char* gf_last_error(const gf_ctx* ctx);
/* Compute units and device name of the context's GPU. */
int gf_device_info(gf_ctx* ctx, int32_t* num_cu, char* name, int32_t name_len);

/* ======================================================================== */
/* Mesh on the device.  Replaces the FemData arrays the GPU strategies read  */
/* (src/FEM/FemData.h:79-105; H2D at src/FEM/StiffAlgoGpuOmp.cc:105-125).    */
/* ======================================================================== */
typedef struct gf_mesh_desc {
  int32_t elem_type;    /* GF_Q4 .. GF_BRICK20 */
  int32_t ngauss;       /* Gauss points per direction: 2 or 3 */
  int64_t nnodes, nelem;
  const int32_t* conn;  /* [nelem][nen], 1-based */
  const double* coords; /* [nnodes][dim] */
  const double* elem_E; /* [nelem] Young's modulus per element, or NULL */
  double E, nu;         /* E used when elem_E is NULL */
} gf_mesh_desc;

int gf_mesh_upload(gf_ctx* ctx, const gf_mesh_desc* desc, gf_mesh** out);
int gf_mesh_free(gf_mesh* mesh);
/* Replaces the per-element moduli (nelem doubles, host) of an uploaded mesh;
 * the next gf_assemble uses them.  Stands in for re-reading the material
 * table of FemData (src/FEM/FemData.cc:322-365) without a new upload. */
int gf_mesh_set_E(gf_ctx* ctx, gf_mesh* mesh, const double* elem_E);

/* ======================================================================== */
/* ELL matrix.  Slot-major ("column-wise") ELLPACK as the reference's       */
/* ELLmatrix (src/SPRmatrix/ELLmatrix.h:35-44): val[row + slot*n],          */
/* col[row + slot*n] ascending within a row, rownnz[n]; width = max rownnz. */
/* Padding slots hold val 0 and col = row.                                  */
/* ======================================================================== */
typedef struct gf_ell_info {
  int64_t n;       /* rows (= DOF) */
  int32_t width;   /* slots per row */
  int32_t ncolors; /* element colors (0 for a matrix made by gf_ell_upload) */
  int64_t nnz;     /* sum of rownnz */
  int64_t bytes;   /* device bytes held by the matrix */
} gf_ell_info;

/* Symbolic structure (node adjacency x dim^2, sorted), greedy element
 * coloring and the element -> slot map, built once per mesh. */
int gf_ell_symbolic(gf_ctx* ctx, const gf_mesh* mesh, gf_ell** out, gf_timing* t);
/* The same matrix built on the device from the uploaded connectivity: node
 * adjacency by a node -> element incidence and one wavefront per node,
 * first-fit coloring in element order in rounds (each element colored once
 * its lower-numbered neighbours are), the slot map by binary search.  The
 * structure, the colors (identical to gf_ell_symbolic's, i.e. the
 * reference's femColor, src/FEM/femColor.cc:79-191) and the slot map equal
 * the host phase's; replaces the ELL growth of AddElem
 * (src/SPRmatrix/ELLmatrix.cc:91-138) and femColor without METIS.
 * GF_ERR_UNSUPPORTED beyond 64 colors, 255 neighbours per node or 2048
 * (incident elements x nen) per node.  t->ms_device covers the whole phase. */
int gf_ell_symbolic_device(gf_ctx* ctx, const gf_mesh* mesh, gf_ell** out, gf_timing* t);
int gf_ell_info_get(const gf_ell* m, gf_ell_info* info);
/* elem_color: [nelem], colors 1..ncolors */
int gf_ell_colors(gf_ctx* ctx, const gf_ell* m, int32_t* elem_color);
int gf_ell_free(gf_ell* m);
/* Matrix from host arrays (slot-major, width slots). */
int gf_ell_upload(gf_ctx* ctx, int64_t n, int32_t width, const double* val,
                  const int32_t* col, const int32_t* rownnz, gf_ell** out);
/* val, col: [width*n]; rownnz: [n] (any may be NULL to skip). */
int gf_ell_download(gf_ctx* ctx, const gf_ell* m, double* val, int32_t* col,
                    int32_t* rownnz);

/* Numeric assembly: Ke by Gauss quadrature fused with the colored,
 * atomic-free scatter (one launch per color), or with a node gather
 * (variants 2 and 3); every stored value is recomputed.  Replaces
 * StiffAlgo::CalcGlobalStiffness (src/FEM/StiffAlgo.h:46). */
int gf_assemble(gf_ctx* ctx, const gf_mesh* mesh, gf_ell* m, gf_timing* t);
/* Assembly kernel of this matrix: 0 = one thread per (element, node pair)
 * with the element outermost; 1 = pair-major, consecutive threads on the same
 * node pair of consecutive elements of a color (neighbouring rows in the
 * scatter); 2 (Brick8/Brick20 with at most 27 / 81 neighbours per node) =
 * element tensors T_pq[a][b] = sum_g w detJ E dNa/dx_p dNb/dx_q formed on
 * the FP64 matrix cores (one wavefront per element, into a buffer of
 * nelem x 9 x nen^2 doubles the matrix keeps while 2 is its variant), then
 * the node gather over them; 3 (at most 9 / 21 / 27 / 81 neighbours per
 * node for Q4 / Q8 / Brick8 / Brick20) = node gather, the lanes of a node
 * (or node row) forming its rows of every surrounding element's Ke and
 * writing them once (no coloring, no read-modify-write); 4 (forced only,
 * where 3 applies) = the Gauss-point geometry of every element (J^-1 and
 * w detJ E) computed once into a buffer the matrix keeps while 4 is its
 * variant (nelem x ngp x (dim^2 + 1) doubles), then the node gather of 3
 * reading it instead of forming it per node.  0 and 1 add the same terms in
 * the same order (bitwise-identical values), as do 3 and 4; 2 and 3 sum in
 * other orders (equal to rounding).
 * force_variant 0..4 fixes it (GF_ERR_UNSUPPORTED where 2, 3 or 4 does not
 * fit); -1 assembles twice with each of 0, 1, 3 and, for Brick20 when its
 * tensors take at most half the free device memory, 2, and keeps the
 * fastest (gf_assemble does this on its first call for meshes of 65536
 * elements or more).  When the solver's block-sliced copy of the matrix
 * exists, 2, 3 and 4 write it with only the diagonal of the ELL values
 * (the other ELL values are rebuilt from the copy by the first call that
 * reads them) and the others are followed by the copy, inside the
 * timing.  ms_per_variant (may be NULL): GF_ASM_VARIANTS entries, the better
 * of two timings of each variant timed, -1 otherwise. */
#define GF_ASM_VARIANTS 5
int gf_assemble_tune(gf_ctx* ctx, const gf_mesh* mesh, gf_ell* m, int32_t force_variant,
                     double* ms_per_variant, int32_t* chosen);

/* Penalty constraint (FEM::ApplyConstraint PEN): pen = coef * max_i K_ii,
 * then K_dd <- pen * K_dd for each listed DOF (0-based; a DOF listed k
 * times is multiplied k times, as repeated SetElem calls would).  *pen_out
 * may be NULL. */
int gf_apply_penalty(gf_ctx* ctx, gf_ell* m, const int32_t* dofs, int64_t ndofs,
                     double coef, double* pen_out);

/* Zero substitution, the elimination alternative to the penalty (the intent
 * of FEM::ApplyConstraint(SUB), src/FEM/fem.cc:146-163): for each listed
 * DOF d (0-based), every stored entry of row d and of column d becomes 0
 * and K_dd = 1; stored entries stay stored.  The solution then has
 * u_d = b_d: clear the right-hand sides at those DOFs with gf_zero_dofs
 * for fixed supports.  (The reference's own SUB reaches SetElem(.., 0),
 * which InputIsOK turns into a no-op, SPRmatrix.cc:97-102, so it only sets
 * K_dd = 1; this entry point does the elimination its comment describes.) */
int gf_apply_elimination(gf_ctx* ctx, gf_ell* m, const int32_t* dofs, int64_t ndofs);

/* v[k*n + d] = 0 for every listed DOF d (host list, 0-based) and
 * right-hand side k < nrhs; v is a device pointer.  No reference
 * counterpart (the right-hand-side half of the elimination). */
int gf_zero_dofs(gf_ctx* ctx, double* v, int64_t n, int32_t nrhs, const int32_t* dofs,
                 int64_t ndofs);

/* Force vector on the device (FemData::CalcForceVector,
 * src/FEM/FemData.cc:370-385): f = 0, then for each load row
 * [node (1-based), F_1..F_dim] in order, its nonzero components are
 * assigned (a later row overwrites an earlier row's nonzero component; a
 * zero component overwrites nothing).  loads: host, nload x (dim+1);
 * f: device pointer of nnodes*dim doubles.  GF_ERR_ARG for a node outside
 * 1..nnodes. */
int gf_load_vector(gf_ctx* ctx, int32_t dim, int64_t nnodes, const double* loads, int64_t nload,
                   double* f);

/* Y = K X for nrhs right-hand sides stored column after column
 * (x[k*n + i]).  Host pointers. */
int gf_spmv(gf_ctx* ctx, const gf_ell* m, const double* x, double* y, int32_t nrhs,
            gf_timing* t);

/* ======================================================================== */
/* Solver                                                                    */
/* ======================================================================== */
#define GF_PRECOND_NONE 0   /* reference CG (src/SPRmatrix/SPRmatrix.cc:405-480) */
#define GF_PRECOND_JACOBI 1 /* Jacobi-PCG (Saad Alg. 9.1, M = diag K) */

#define GF_PREC_FP64 0  /* everything in FP64 */
#define GF_PREC_MIXED 1 /* FP32 inner PCG inside FP64 iterative refinement */

#define GF_SCHED_AUTO 0
#define GF_SCHED_KERNELS 1    /* one launch per BLAS step, host-free loop */
#define GF_SCHED_GRAPH 2      /* iteration chunks captured in a hipGraph */
#define GF_SCHED_PERSISTENT 3 /* one cooperative kernel, grid barriers */

typedef struct gf_solve_opts {
  int32_t precond;   /* GF_PRECOND_* */
  int32_t precision; /* GF_PREC_* */
  int32_t schedule;  /* GF_SCHED_* */
  int32_t use_x0;    /* 1: x holds the initial guess; 0: start from zero; 2 (FP64,
                        single right-hand side; elsewhere as 1): x holds a guess that
                        is kept only when ||b - K x|| < ||b||, else the solve starts
                        from zero (zero start is the fallback) */
  int32_t maxit;     /* iteration cap (per right-hand side) */
  int32_t device_ptrs; /* 1: b and x are device pointers */
  double rtol;       /* stop when ||b - K x||_2 <= rtol ||b||_2 */
  double inner_rtol; /* reserved */
  int32_t batched;   /* nrhs > 1: 1 = all right-hand sides advance together with
                        one ELL SpMM per iteration (GF_PREC_MIXED: an FP32 SpMM, each
                        column refreshed in FP64 as the single-RHS mixed solve does);
                        0 = one after another */
  int32_t verify;    /* 1 (default): when the recursive residual meets rtol, recompute
                        ||b - K x|| in FP64 and, while it exceeds rtol, restart PCG from
                        the current x (at most 4 times); 0: stop on the recursive
                        residual alone, as the reference CG does */
} gf_solve_opts;

typedef struct gf_solve_stats {
  int32_t iterations;   /* inner iterations (max over right-hand sides) */
  int32_t refinements;  /* outer FP64 sweeps (GF_PREC_MIXED) */
  int32_t converged;    /* 1 if every right-hand side met rtol */
  int32_t status;       /* GF_OK, GF_ERR_NOCONV or GF_ERR_BREAKDOWN */
  double relres;        /* max over RHS of recursive ||r||/||b|| at exit */
  double true_relres;   /* max over RHS of ||b - K x||/||b|| recomputed in FP64 */
  double ms_device;     /* event time of the solve on the stream */
  double ms_host;       /* host wall time of the call */
  int32_t fallbacks;    /* GF_PREC_MIXED: right-hand sides (batched: solves) finished by
                           FP64 PCG: from zero after the FP32 iteration broke down, from
                           the current x once the FP32 operator proved too ill-conditioned
                           (condition estimate from the CG coefficients above 1/u32) or
                           the refreshes stopped halving the true residual */
  int32_t x0_rejected;  /* 1 if a guarded initial guess (use_x0 = 2) was replaced by
                           zero */
} gf_solve_stats;

void gf_solve_opts_default(gf_solve_opts* o);
/* Solves K X = B for nrhs right-hand sides (column after column). */
int gf_pcg(gf_ctx* ctx, gf_ell* m, const double* b, double* x, int32_t nrhs,
           const gf_solve_opts* opts, gf_solve_stats* stats);

/* SpMV schedule of this matrix (the autotuner).  Variants: 0..2 one thread
 * per row with 1/4/8 slots in flight, 3..4 two/four wavefronts per 64-row
 * tile, 5/6 one thread per row over the 2x2 / 3x3 node blocks of a matrix
 * built by gf_ell_symbolic for a 2D / 3D mesh (one node column per block
 * row, the node's x entries read together), 7/8 one thread per node row of
 * a block-sliced copy the solver keeps beside the ELL arrays (node rows
 * sorted by length in windows of 2048, slices of 64, one column index per
 * 2x2 / 3x3 block; rebuilt from the ELL values after each assembly or
 * penalty), the PCG grid sweeping the slices, 9/10 the same rows with one
 * workgroup per group of four slices (the hardware dispatcher balances
 * slices of unequal length).  Variants 0..2 and 5..10 add each row's entries
 * in slot order and give bitwise-identical products.  force_variant >= 0 fixes a
 * variant; -1 times every variant that fits this matrix and keeps the
 * fastest (gf_pcg does this on first use).  ms_per_variant (may be NULL)
 * receives GF_SPMV_VARIANTS averages (-1 for variants that do not fit).  No
 * reference counterpart; it plays the role of SetOclStrategy /
 * SetOclLocalSize (src/SPRmatrix/SPRmatrix.h:110-121) chosen by measurement
 * instead of by hand. */
#define GF_SPMV_VARIANTS 11
int gf_ell_tune(gf_ctx* ctx, gf_ell* m, int32_t force_variant, double* ms_per_variant,
                int32_t* chosen);

/* Average duration (ms) of the SpMV kernel the solver runs each iteration
 * (q = K p fused with the p.q partial sums), timed with HIP events over nrep
 * back-to-back launches on the context's stream.  No reference counterpart:
 * it mirrors the per-kernel event timing of OCLwrapper
 * (src/OpenCL/OCLwrapper.cc:400-434) for the roofline figure. */
int gf_profile_spmv(gf_ctx* ctx, gf_ell* m, int32_t nrep, double* ms_avg);

/* Runs iters Jacobi-PCG iterations from x = 0 for the right-hand side b
 * (host array of n) with the three-launch schedule, HIP events around each
 * iteration's SpMV kernel: ms_spmv is that kernel's average duration inside
 * the solve (p freshly rewritten by the previous iteration, as in gf_pcg),
 * ms_iteration the average wall time of one iteration including the event
 * records.  GF_ERR_BREAKDOWN when the iteration stops early.  No reference
 * counterpart; the same per-kernel event timing as gf_profile_spmv. */
int gf_profile_pcg(gf_ctx* ctx, gf_ell* m, const double* b, int32_t iters, double* ms_spmv,
                   double* ms_iteration);
/* Cost of one grid barrier of the persistent schedule (GF_SCHED_PERSISTENT,
 * three barriers per iteration): nbar barriers with no work between them on
 * the persistent kernel's co-resident grid for the matrix's SpMV variant
 * (*nblocks workgroups), timed with HIP events, minus an empty launch;
 * *us_per_barrier in microseconds.  No reference counterpart (the reference
 * ends every CG step with a host synchronisation, ELLmatrix.cc:315-459). */
int gf_profile_barrier(gf_ctx* ctx, gf_ell* m, int32_t nbar, double* us_per_barrier,
                       int32_t* nblocks);
/* Bandwidth probe: nrep launches of a plain streaming copy kernel of `bytes`
 * (multiple of 16; 16-byte aligned device pointers) on the context's stream;
 * *ms_avg = average launch time.  Measurement aid with no reference
 * counterpart: the copy's achieved rate is the attainable HBM ceiling, and
 * its counter readings calibrate hardware byte counters. */
int gf_stream_copy(gf_ctx* ctx, void* dst, const void* src, int64_t bytes, int32_t nrep,
                   double* ms_avg);

/* ======================================================================== */
/* Grouped batch: many independent meshes of one element type (same Gauss   */
/* rule and nu; E may differ) assembled and solved together as one          */
/* block-diagonal system.  Assembly is one launch per color over all the    */
/* meshes; the penalty factor, the PCG scalars and the stopping test are    */
/* per mesh.  No reference counterpart: the reference runs its batches one  */
/* file after another (src/interface.cpp:503-605).                          */
/* ======================================================================== */
typedef struct gf_group gf_group;
int gf_group_create(gf_ctx* ctx, int32_t nmesh, const gf_mesh_desc* meshes, gf_group** out);
int gf_group_info(const gf_group* g, int64_t* total_dof, int64_t* nnz, int32_t* ncolors);
int gf_group_assemble(gf_ctx* ctx, gf_group* g, gf_timing* t);
/* New per-element moduli for every mesh (all meshes' elements concatenated,
 * host); the structure, coloring and tiles are kept. */
int gf_group_set_E(gf_ctx* ctx, gf_group* g, const double* elem_E);
/* dofs: fixed DOFs of every mesh (0-based within its mesh) concatenated;
 * ndofs[m] of them belong to mesh m.  pen_m = coef * max diag of mesh m. */
int gf_group_apply_penalty(gf_ctx* ctx, gf_group* g, const int32_t* dofs, const int64_t* ndofs,
                           double coef);
/* b, x: all meshes concatenated (mesh m after meshes 0..m-1).  Jacobi-PCG,
 * FP64.  iterations_per_mesh (may be NULL): [nmesh]. */
int gf_group_pcg(gf_ctx* ctx, gf_group* g, const double* b, double* x, const gf_solve_opts* opts,
                 gf_solve_stats* stats, int32_t* iterations_per_mesh);
int gf_group_free(gf_group* g);

/* ======================================================================== */
/* Learned warm start: a BF16 3D convolutional encoder-decoder on the node  */
/* grid of a structured mesh, trained on displacement fields this solver    */
/* computes, whose prediction is the PCG initial guess (opts.use_x0 = 1;    */
/* zero-start PCG stays the fallback and every answer meets opts.rtol).     */
/* No reference counterpart.  Tensors are device pointers, channels-last:  */
/* input [batch][nz][ny][nx][16] bf16 bit patterns (features, zero-padded), */
/* output and target [batch][nodes][16] / [batch][nodes][4] fp32 (the first */
/* 3 channels are displacement).  Every node-grid axis is 1 mod 8.          */
/* ======================================================================== */
typedef struct gf_nn gf_nn;
int gf_nn_create(gf_ctx* ctx, int32_t nx, int32_t ny, int32_t nz, int32_t batch_max,
                 uint64_t seed, gf_nn** out);
int gf_nn_free(gf_nn* net);
int64_t gf_nn_nparams(const gf_nn* net);
int gf_nn_layer_count(const gf_nn* net);
/* info[8] = cin, cout, stride, activation, input level, output level,
 * weight offset, bias offset (weights W[27][cin][cout], tap = 9(dz+1) +
 * 3(dy+1) + (dx+1), in the flat fp32 parameter vector) */
int gf_nn_layer_info(const gf_nn* net, int32_t layer, int32_t* info);
int gf_nn_params_get(gf_ctx* ctx, gf_nn* net, float* host);
/* sets the parameters and resets the Adam moments */
int gf_nn_params_set(gf_ctx* ctx, gf_nn* net, const float* host);
int gf_nn_grads_get(gf_ctx* ctx, gf_nn* net, float* host);
/* out_dev (may be NULL): [batch][nodes][16] fp32 */
int gf_nn_forward(gf_ctx* ctx, gf_nn* net, const uint16_t* in_dev, int32_t batch, float* out_dev);
/* forward + mean squared error over the 3 displacement channels + backward:
 * leaves the parameter gradient on the device (gf_nn_grads_get) */
int gf_nn_backward(gf_ctx* ctx, gf_nn* net, const uint16_t* in_dev, const float* target_dev,
                   int32_t batch, float* loss_host);
int gf_nn_adam(gf_ctx* ctx, gf_nn* net, float lr, float beta1, float beta2, float eps);
/* gf_nn_backward then Adam (0.9, 0.999, 1e-8); loss_host may be NULL */
int gf_nn_train_step(gf_ctx* ctx, gf_nn* net, const uint16_t* in_dev, const float* target_dev,
                     int32_t batch, float lr, float* loss_host);
/* x_dev[3 i + c] = scale * (pred[sample][i][c] + base_weight *
 * in[sample][i][base_channel + c]), zero at nodes whose input feature
 * channel mask_channel exceeds 1/2; base_channel = -1: no base term (the
 * network then predicts the whole field, otherwise a correction to the
 * base estimate carried in three input channels); x_dev is an FP64 DOF
 * vector */
int gf_nn_to_dofs(gf_ctx* ctx, gf_nn* net, const float* pred_dev, const uint16_t* in_dev,
                  int32_t sample, int32_t mask_channel, int32_t base_channel, double base_weight,
                  double scale, double* x_dev);
/* One 3x3x3 convolution layer through the same kernels (zero padding):
 *   transposed = 0: out[v] = act(bias + sum_t in[s v + t] . W[t])  W: [27][cin][cout]
 *   transposed = 1: out[u] = sum_t in[(u - t)/s] . W[t]^T           W: [27][cout][cin]
 * vin/vout: node grids {nx, ny, nz}; channels multiples of 16; exactly one
 * of out_dev (bf16) / out32_dev (fp32) non-NULL; accumulate adds to it. */
int gf_nn_conv3d(gf_ctx* ctx, const uint16_t* in_dev, int32_t batch, const int32_t* vin,
                 int32_t cin, const float* w_dev, const float* bias_dev, int32_t cout,
                 int32_t stride, int32_t transposed, int32_t act, float slope, const int32_t* vout,
                 uint16_t* out_dev, float* out32_dev, int32_t accumulate);
/* dW[t][ci][co] = sum_v x[s v + t][ci] dz[v][co]; db[co] = sum_v dz[v][co] (db may be NULL) */
int gf_nn_conv3d_wgrad(gf_ctx* ctx, const uint16_t* x_dev, int32_t batch, const int32_t* vin,
                       int32_t cin, const uint16_t* dz_dev, const int32_t* vout, int32_t cout,
                       int32_t stride, float* dw_dev, float* db_dev);

/* ======================================================================== */
/* Allocation-free solve.  No counterpart in the reference: for callers     */
/* that cannot allocate device memory (code under HIP graph capture, tight  */
/* serving loops, harnesses that forbid hipMalloc).  The FP64 (Jacobi-)PCG   */
/* of gf_pcg for one right-hand side, b and x device pointers (x: initial   */
/* guess when opts->use_x0), kernel launches only; the Jacobi inverse and   */
/* the r, z, p, q vectors, partial sums and scalar state live in a static   */
/* device array of GF_NOALLOC_SCRATCH_BYTES that the library holds on each  */
/* GPU it runs on until the process ends.  Matrices of more than            */
/* GF_NOALLOC_MAX_ROWS rows are rejected (GF_ERR_UNSUPPORTED).  Uses the    */
/* matrix's tuned SpMV variant when its copies are current (gf_ell_tune or  */
/* an earlier gf_pcg builds them), else one thread per ELL row; calls no    */
/* allocating routine.  All calls on one GPU share its array: calls that    */
/* overlap on one GPU race.                                                  */
/* ======================================================================== */
#define GF_NOALLOC_MAX_ROWS (1 << 20)
#define GF_NOALLOC_SCRATCH_BYTES 42520832
int gf_pcg_noalloc(gf_ctx* ctx, gf_ell* m, const double* b, double* x,
                   const gf_solve_opts* opts, gf_solve_stats* stats);

/* Device memory helpers for callers that keep vectors resident in HBM. */
int gf_dev_alloc(gf_ctx* ctx, int64_t bytes, void** dptr);
int gf_dev_free(gf_ctx* ctx, void* dptr);
int gf_copy_h2d(gf_ctx* ctx, void* dst, const void* src, int64_t bytes);
int gf_copy_d2h(gf_ctx* ctx, void* dst, const void* src, int64_t bytes);
/* stream-ordered device-to-device copy (returns before it completes) */
int gf_copy_d2d(gf_ctx* ctx, void* dst, const void* src, int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* GPUFEM_H_ */
