// Mixed-precision Jacobi-PCG (FP32 inner iterations inside FP64 refinement).
#pragma once

#include "gf_internal.hpp"
#include "launch.hpp"

struct gf_solve_plan;

namespace gf {

struct MixedState;

struct MixedBuffers {
  int64_t n;
  int variant;
  int nblocks;    // workgroups of every kernel but the FP32 SpMV
  int spmv_grid;  // workgroups of the FP32 SpMV (slice-group variants: one per group)
  // FP64 outer system
  const double* val;
  const int32_t* col;
  const int32_t* rownnz;
  const double* b;
  double* x;
  double* q64;  // K x for the refreshed residual
  // FP32 inner iteration
  const float* val32;
  const float* dinv32;
  float* r32;
  float* z32;
  float* p32;
  float* q32;
  double* part1;  // p.q
  double* part2;  // r.z, r.r of the update
  double* part3;  // r.z, r.r of a refreshed residual
  MixedState* st;
  BsellView bs;   // block-sliced view (kBsell variants: val, val32, col are its arrays)
  // alpha, beta of iterations 1..kLanczosIters (pairs), for the host's
  // condition estimate
  double* hist;
};

size_t mixed_state_bytes();
hipError_t launch_to_f32(hipStream_t s, int64_t count, const double* in, float* out);
// FP32 copy of block-sliced values in difference form (d x d blocks, FP64
// pair layout in, FP32 quad layout out; the diagonal slot of every node row
// holds the block row sum, see bsell_row)
hipError_t launch_bsell_to_f32(hipStream_t s, int d, const BsellView& bs, const int32_t* bcol,
                               const double* in, float* out);
hipError_t launch_mixed_setup(hipStream_t s, const MixedBuffers& B, int use_x0, double rtol,
                              double delta, int maxit);
hipError_t launch_mixed_iteration(hipStream_t s, const MixedBuffers& B, int parity);
// status of a mixed-precision iteration stopped because kStallRefreshes
// consecutive FP64 refreshes each found the true residual less than halved
// (the FP32 operator no longer resolves the remaining error); the solve then
// continues in FP64 from the current x
constexpr int kMixedStalled = 1;
constexpr int kStallRefreshes = 3;

// Condition check of the FP32 inner iteration.  The alpha and beta of CG are
// the Lanczos coefficients of the Jacobi-scaled operator: T_k (diagonal
// 1/alpha_j + beta_{j-1}/alpha_{j-1}, off-diagonal sqrt(beta_j)/alpha_j) has
// extreme eigenvalues that approach those of the operator, lambda_min from
// above.  When u32 lambda_max(T_k) / lambda_min(T_k) approaches 1 the FP32
// rounding of the operator and of the vectors reaches its smallest
// eigenvalues: FP32 iterations can no longer resolve the slowest error
// components and the solve continues in FP64 (the same status as a stall).
// The estimate grows slowly (lambda_min from above), so the hand-over is at
// kKappaU32 = 0.05: a Q8 1024x512 plate (u32 kappa ~ 12) crosses it after
// 1,024 iterations instead of 2,304 at 1; the estimates of the Brick8 128^3
// and Brick20 48^3 cubes stay below 0.018.  Checked every kLanczosCheck
// iterations over the first kLanczosIters.
constexpr int kLanczosIters = 8192;
constexpr int kLanczosCheck = 256;
constexpr double kKappaU32 = 0.05;
constexpr double kU32 = 5.9604644775390625e-8;  // 2^-24
// lambda_max / lambda_min of T_k from k CG steps (alpha[j], beta[j] at
// stride `stride`); 0 when T_k is not positive definite or k < 2
double lanczos_condition(const double* alpha, const double* beta, int k, int stride);

struct MixedHostState {
  int32_t it, done, status, refinements;
  double rr_true, rr_rec, bnorm2;
};
hipError_t read_mixed_state(hipStream_t s, const MixedBuffers& B, MixedHostState* h);

int run_mixed(gf_ctx* ctx, gf_ell* m, const gf_solve_opts& opts, PcgHostState* h,
              int* refinements);
void release_mixed(gf_solve_plan* P);
// the FP32 copy of the values the SpMV variant reads (ELL or block-sliced),
// made or refreshed when stale
int ensure_val32(gf_ctx* ctx, gf_ell* m, int variant, const float** out);

}  // namespace gf
