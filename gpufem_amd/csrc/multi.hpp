// Batched (multi right-hand-side) Jacobi-PCG with one ELL SpMM per iteration
// (kernels in multi.hip, host driver in multi.cpp).
#pragma once

#include "gf_internal.hpp"
#include "launch.hpp"

struct gf_solve_plan;

namespace gf {

struct MultiBuffers {
  int64_t n;
  int S;      // columns in the row-major layout (8, 16, 32 or 64)
  int ncols;  // columns actually solved (<= S); the rest stay zero
  int use_x0;
  int nblocks;
  const double* val;
  const int32_t* col;
  const int32_t* rownnz;
  const double* dinv;  // null: no preconditioner
  const double* b;     // [n][S]
  double* x;           // [n][S]
  double *r, *p, *q;  // z = D^-1 r is formed where it is used
  double* part1;  // [S][nblocks]
  double* part2;  // [3S][nblocks]
  void* st;       // MultiState<S>
  // block-sliced copy (bsell_d = 2 or 3: val/col are its arrays; 0: ELL)
  int bsell_d;
  BsellView bs;
  // extra LDS bytes per SpMM workgroup (multi_spmm_lds): at most three
  // workgroups share a CU
  int spmm_lds;
};

struct MultiHostState {
  int32_t it, done;
  int32_t col_it[64], col_status[64];
  double rr[64], bnorm2[64];
};

int multi_width(int ncols);       // smallest supported S >= ncols (0: too many)
size_t multi_state_bytes(int S);
int multi_blocks(int64_t n, int S);
// Dynamic LDS that leaves room for three block-sliced SpMM workgroups on a CU
// with lds_per_cu bytes (0: no padding).
int multi_spmm_lds(int lds_per_cu);
hipError_t launch_multi_pack(hipStream_t s, int S, int64_t n, int ncols, const double* cols,
                             double* rows);
hipError_t launch_multi_unpack(hipStream_t s, int S, int64_t n, int ncols, const double* rows,
                               double* cols);
hipError_t launch_multi_setup(hipStream_t s, const MultiBuffers& B, double rtol, int maxit);
hipError_t launch_multi_iteration(hipStream_t s, const MultiBuffers& B);
// x += alpha p of the iteration that finished the loop (once, after it)
hipError_t launch_multi_x_finish(hipStream_t s, const MultiBuffers& B);
hipError_t launch_multi_residual(hipStream_t s, const MultiBuffers& B, double* out);
hipError_t read_multi_state(hipStream_t s, const MultiBuffers& B, MultiHostState* h);

// Solves ncols right-hand sides (column after column at Bd / Xd, device
// pointers) together.  Fills per-column iterations/residuals.
int run_multi(gf_ctx* ctx, gf_ell* m, const gf_solve_opts& opts, const double* Bd, double* Xd,
              int ncols, MultiHostState* h, double* true_res);
void release_multi(gf_solve_plan* P);

}  // namespace gf
