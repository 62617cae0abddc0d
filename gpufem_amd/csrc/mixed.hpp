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
};

size_t mixed_state_bytes();
hipError_t launch_to_f32(hipStream_t s, int64_t count, const double* in, float* out);
// FP32 copy of block-sliced values (slots block slots of dd values, FP64
// pair layout in, FP32 quad layout out)
hipError_t launch_bsell_to_f32(hipStream_t s, int dd, int64_t slots, const double* in,
                              float* out);
hipError_t launch_mixed_setup(hipStream_t s, const MixedBuffers& B, int use_x0, double rtol,
                              double delta, int maxit);
hipError_t launch_mixed_iteration(hipStream_t s, const MixedBuffers& B, int parity);
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
