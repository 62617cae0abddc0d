// Batched mixed-precision Jacobi-PCG (kernels in multi_mixed.hip, host
// driver in multi_mixed.cpp).
#pragma once

#include "gf_internal.hpp"
#include "launch.hpp"

struct gf_solve_plan;

namespace gf {

struct MixedMultiBuffers {
  int64_t n;
  int S;      // columns in the row-major layout (8, 16, 32 or 64)
  int ncols;  // columns actually solved (<= S); the rest stay zero
  int nblocks;
  // FP64 values (refresh) and their FP32 copy (inner iteration), in the
  // same layout: ELL (bsell_d = 0) or the block-sliced copy (bsell_d = 2, 3)
  const double* val;
  const float* val32;
  const int32_t* col;
  const int32_t* rownnz;
  int bsell_d;
  BsellView bs;
  const float* dinv32;  // null: no preconditioner
  const double* b;      // [n][S]
  double* x;            // [n][S]
  float *r, *p, *q;  // [n][S]; z = D^-1 r is formed where it is used
  double* part1;  // [S][nblocks]   p.q
  double* part2;  // [3S][nblocks]  r.z, r.r (and b.b at setup)
  double* part3;  // [2S][nblocks]  refreshed r.z, r.r
  void* st;       // MMState<S>
  // alpha, beta of every column for iterations 1..kLanczosIters:
  // hist[2 (S (it - 1) + k) + {0, 1}], for the host's condition estimate
  double* hist;
};

struct MixedMultiHostState {
  int32_t it, done;
  int32_t ill;  // set by the host: some column's FP32 operator too ill-conditioned
  int32_t col_it[64], col_status[64], col_refinements[64];
  double rr_true[64], rr_rec[64], bnorm2[64];
};

size_t mixed_multi_state_bytes(int S);
hipError_t launch_mixed_multi_setup(hipStream_t s, const MixedMultiBuffers& B, double rtol,
                                    double delta, int maxit);
hipError_t launch_mixed_multi_iteration(hipStream_t s, const MixedMultiBuffers& B);
hipError_t read_mixed_multi_state(hipStream_t s, const MixedMultiBuffers& B,
                                  MixedMultiHostState* h);

// Solves ncols right-hand sides (column after column at Bd / Xd, device
// pointers) together in mixed precision; per-column iterations, relative
// true residuals (b - K x in FP64 at the last refresh) and refinements.
int run_multi_mixed(gf_ctx* ctx, gf_ell* m, const gf_solve_opts& opts, const double* Bd,
                    double* Xd, int ncols, MixedMultiHostState* h, double* true_res);
void release_multi_mixed(gf_solve_plan* P);

}  // namespace gf
