// Host driver of the batched mixed-precision PCG (kernels in
// multi_mixed.hip): workspace, FP32 copies, and the chunked iteration loop
// (replayed from a captured hipGraph unless the kernel schedule is asked).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "mixed.hpp"
#include "multi.hpp"
#include "multi_mixed.hpp"
#include "solve_plan.hpp"

struct gf_mmulti_ws {
  int64_t n = 0;
  int S = 0;
  double *b = nullptr, *x = nullptr;
  float *r = nullptr, *p = nullptr, *q = nullptr, *dinv32 = nullptr;
  double *part1 = nullptr, *part2 = nullptr, *part3 = nullptr;
  double* hist = nullptr;  // alpha, beta per column and iteration (kLanczosIters)
  void* st = nullptr;
  uint64_t dinv_version = ~0ull;  // matrix values the FP32 Jacobi copy was made from
  hipGraphExec_t graph = nullptr;  // kChunk iterations, captured for graph_B
  gf::MixedMultiBuffers graph_B{};
};

namespace gf {

int set_error(gf_ctx* ctx, int code, const std::string& msg);

namespace {
constexpr int kChunk = 16;
constexpr double kDelta = 0.1;  // refresh when a column's residual fell 10x

hipError_t ensure_graph(gf_ctx* ctx, gf_mmulti_ws* W, const MixedMultiBuffers& B) {
  if (W->graph && std::memcmp(&W->graph_B, &B, sizeof(B)) == 0) return hipSuccess;
  if (W->graph) {
    (void)hipGraphExecDestroy(W->graph);
    W->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  for (int it = 0; it < kChunk; ++it) {
    e = launch_mixed_multi_iteration(ctx->stream, B);
    if (e != hipSuccess) break;
  }
  hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(&W->graph, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e == hipSuccess) std::memcpy(&W->graph_B, &B, sizeof(B));
  else W->graph = nullptr;
  return e;
}
}  // namespace

void release_multi_mixed(gf_solve_plan* P) {
  gf_mmulti_ws* W = P->mmulti;
  if (!W) return;
  void* v[] = {W->b,      W->x,     W->r,     W->p,     W->q,    W->dinv32,
               W->part1, W->part2, W->part3, W->hist, W->st};
  for (void* d : v)
    if (d) (void)hipFree(d);
  if (W->graph) (void)hipGraphExecDestroy(W->graph);
  delete W;
  P->mmulti = nullptr;
}

int run_multi_mixed(gf_ctx* ctx, gf_ell* m, const gf_solve_opts& opts, const double* Bd,
                    double* Xd, int ncols, MixedMultiHostState* h, double* true_res) {
  gf_solve_plan* P = m->plan;
  const int S = multi_width(ncols);
  if (S == 0) return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_pcg: at most 64 right-hand sides");
  const int64_t n = m->n;
  const int nb = multi_blocks(n, S);
  if (!P->mmulti || P->mmulti->S != S || P->mmulti->n != n) {
    release_multi_mixed(P);
    gf_mmulti_ws* W = new gf_mmulti_ws();
    P->mmulti = W;
    W->n = n;
    W->S = S;
    const size_t cells = size_t(n) * size_t(S);
    hipError_t e = hipMalloc(&W->b, cells * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&W->x, cells * sizeof(double));
    float** f[] = {&W->r, &W->p, &W->q};
    for (float** v : f)
      if (e == hipSuccess) e = hipMalloc(v, cells * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&W->dinv32, size_t(n > 0 ? n : 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&W->part1, size_t(S) * nb * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&W->part2, size_t(3 * S) * nb * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&W->part3, size_t(2 * S) * nb * sizeof(double));
    if (e == hipSuccess)
      e = hipMalloc(&W->hist, 2 * size_t(S) * size_t(kLanczosIters) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&W->st, mixed_multi_state_bytes(S));
    if (e != hipSuccess) {
      release_multi_mixed(P);
      return set_error(ctx, GF_ERR_NOMEM,
                       std::string("gf_pcg (batched mixed): workspace: ") + hipGetErrorString(e));
    }
  }
  gf_mmulti_ws* W = P->mmulti;
  int st = prepare_variant(ctx, m, P->variant);
  if (st) return st;
  const float* val32 = nullptr;
  if ((st = ensure_val32(ctx, m, P->variant, &val32))) return st;
  const bool jacobi = opts.precond == GF_PRECOND_JACOBI;
  if (jacobi && W->dinv_version != m->val_version) {
    GF_HIP_CHECK(ctx, launch_to_f32(ctx->stream, n, m->d_dinv, W->dinv32));
    W->dinv_version = m->val_version;
  }
  const SpmvOperands o = spmv_operands(m, P->variant);
  const bool bsell = is_bsell_variant(P->variant);
  MixedMultiBuffers B{};  // zero padding: the graph cache compares the bytes
  B.n = n;
  B.S = S;
  B.ncols = ncols;
  B.nblocks = nb;
  B.val = bsell ? o.val : m->d_val;
  B.val32 = val32;
  B.col = bsell ? o.col : m->d_col;
  B.rownnz = m->d_rownnz;
  B.bsell_d = bsell ? m->dim : 0;
  B.bs = o.bs;
  B.dinv32 = jacobi ? W->dinv32 : nullptr;
  B.b = W->b;
  B.x = W->x;
  B.r = W->r;
  B.p = W->p;
  B.q = W->q;
  B.part1 = W->part1;
  B.part2 = W->part2;
  B.part3 = W->part3;
  B.st = W->st;
  B.hist = W->hist;
  GF_HIP_CHECK(ctx, launch_multi_pack(ctx->stream, S, n, ncols, Bd, W->b));
  if (opts.use_x0) {
    GF_HIP_CHECK(ctx, launch_multi_pack(ctx->stream, S, n, ncols, Xd, W->x));
  } else {
    GF_HIP_CHECK(ctx, hipMemsetAsync(W->x, 0, size_t(n) * size_t(S) * sizeof(double), ctx->stream));
  }
  GF_HIP_CHECK(ctx, launch_mixed_multi_setup(ctx->stream, B, opts.rtol, kDelta, opts.maxit));
  GF_HIP_CHECK(ctx, read_mixed_multi_state(ctx->stream, B, h));
  const bool graph = opts.schedule != GF_SCHED_KERNELS;
  if (graph && !h->done) GF_HIP_CHECK(ctx, ensure_graph(ctx, W, B));
  // condition check of every column, as the single-RHS mixed solve does
  std::vector<double> hist;
  int checked = 0;
  h->ill = 0;
  while (!h->done) {
    if (graph) {
      GF_HIP_CHECK(ctx, hipGraphLaunch(W->graph, ctx->stream));
    } else {
      for (int it = 0; it < kChunk; ++it)
        GF_HIP_CHECK(ctx, launch_mixed_multi_iteration(ctx->stream, B));
    }
    GF_HIP_CHECK(ctx, read_mixed_multi_state(ctx->stream, B, h));
    const int k = std::min(h->it, kLanczosIters);
    if (!h->done && k >= checked + kLanczosCheck) {
      hist.resize(2 * size_t(S) * size_t(k));
      const size_t o = 2 * size_t(S) * size_t(checked);
      GF_HIP_CHECK(ctx, hipMemcpy(hist.data() + o, W->hist + o,
                                  2 * size_t(S) * size_t(k - checked) * sizeof(double),
                                  hipMemcpyDeviceToHost));
      checked = k;
      for (int c = 0; c < ncols && !h->ill; ++c) {
        const int kc = std::min(h->col_it[c], k);
        const double kappa = lanczos_condition(hist.data() + 2 * c, hist.data() + 2 * c + 1, kc,
                                               2 * S);
        if (kappa * kU32 > kKappaU32) h->ill = 1;
      }
      if (h->ill) break;
    }
  }
  for (int k = 0; k < ncols; ++k) {
    const double bb = h->bnorm2[k];
    true_res[k] = bb > 0.0 ? std::sqrt(h->rr_true[k] / bb) : std::sqrt(h->rr_true[k]);
  }
  GF_HIP_CHECK(ctx, launch_multi_unpack(ctx->stream, S, n, ncols, W->x, Xd));
  return GF_OK;
}

}  // namespace gf
