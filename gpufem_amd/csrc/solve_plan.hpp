// Solver workspace cached on an ELL matrix (solve.cpp, mixed.cpp).
#pragma once

#include <string>

#include "gf_internal.hpp"
#include "launch.hpp"

struct gf_mixed_ws;  // mixed-precision workspace (mixed.cpp)
struct gf_multi_ws;  // batched multi right-hand-side workspace (multi.cpp)
struct gf_mmulti_ws;  // batched mixed-precision workspace (multi_mixed.cpp)

struct gf_solve_plan {
  int64_t n = 0;
  int num_cu = 0;
  int variant = -1;  // SpMV variant (-1: not tuned yet)
  int nblocks = 0;   // PCG workgroups
  bool persistent_ok = false;
  double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *wb = nullptr, *wx = nullptr;  // right-hand side and solution being solved
  double *part1 = nullptr, *part2 = nullptr;
  double* res = nullptr;  // true residual per right-hand side
  gf::PcgState* st = nullptr;
  void* barrier = nullptr;  // grid barrier of the persistent kernel
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // keyed by preconditioner use
  gf_mixed_ws* mixed = nullptr;
  gf_multi_ws* multi = nullptr;
  gf_mmulti_ws* mmulti = nullptr;
};

namespace gf {
void release_solve_plan(gf_ell* m);
int ensure_plan(gf_ctx* ctx, gf_ell* m);
void set_variant(gf_ell* m, int variant);
// workgroups of the PCG kernels for this variant and size (the grid every
// schedule uses, so all give bitwise-identical results)
int pcg_grid(int variant, int64_t n, int num_cu);
int tune_spmv(gf_ctx* ctx, gf_ell* m, int force, double* ms_out, int* chosen);
PcgBuffers make_buffers(gf_ell* m, const double* dinv);
// arrays the SpMV variant reads: the ELL arrays, the node columns (kNode) or
// the block-sliced copy (kBsell)
struct SpmvOperands {
  const double* val;
  const int32_t* col;
  BsellView bs;
};
SpmvOperands spmv_operands(const gf_ell* m, int variant);
// the block-sliced copy of the matrix (valid after prepare_variant of a kBsell variant)
BsellView bsell_view(const gf_ell* m);
// builds what the variant reads and refreshes copies of stale values
int prepare_variant(gf_ctx* ctx, gf_ell* m, int variant);
// the ELL values of every slot current (from the block-sliced copy when the
// node gather left them stale)
int ensure_ell_vals(gf_ctx* ctx, gf_ell* m);
// the node-block / block-sliced (swept) variant of the matrix's dimension, or -1
int node_variant(const gf_ell* m);
int bsell_variant(const gf_ell* m);
int prepare_jacobi(gf_ctx* ctx, gf_ell* m);
int run_pcg(gf_ctx* ctx, gf_ell* m, const PcgBuffers& B, int sched, int key, int use_x0,
            double rtol, int maxit, int verify, PcgHostState* h);
}  // namespace gf
