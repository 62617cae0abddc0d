// Host side of the mixed-precision solve (kernels in mixed.hip): FP32 copies
// of the matrix values and of the Jacobi inverse, the FP32 work vectors, and
// the chunked iteration loop.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mixed.hpp"
#include "solve_plan.hpp"

struct gf_mixed_ws {
  int64_t n = 0;
  float *dinv32 = nullptr, *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr;
  double* q64 = nullptr;
  double* part3 = nullptr;
  double* hist = nullptr;  // alpha, beta per iteration (kLanczosIters pairs)
  gf::MixedState* st = nullptr;
  uint64_t dinv_version = 0;  // matrix values the FP32 Jacobi copy was made from
  hipGraphExec_t graph = nullptr;  // kChunk iterations, captured for graph_B
  gf::MixedBuffers graph_B{};
};

namespace gf {

int set_error(gf_ctx* ctx, int code, const std::string& msg);

namespace {
// eigenvalues of the symmetric tridiagonal (d, e) below x (Sturm count)
int sturm_below(const std::vector<double>& d, const std::vector<double>& e, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t j = 0; j < d.size(); ++j) {
    const double off = j ? e[j - 1] * e[j - 1] : 0.0;
    q = d[j] - x - (j ? off / q : 0.0);
    if (q == 0.0) q = -1e-300;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}
}  // namespace

double lanczos_condition(const double* alpha, const double* beta, int k, int stride) {
  if (k < 2) return 0.0;
  std::vector<double> d(static_cast<size_t>(k)), e(static_cast<size_t>(k - 1));
  for (int j = 0; j < k; ++j) {
    const double a = alpha[size_t(j) * stride];
    if (!(a > 0.0)) return 0.0;
    d[size_t(j)] = 1.0 / a;
    if (j > 0) {
      const double ap = alpha[size_t(j - 1) * stride], bp = beta[size_t(j - 1) * stride];
      if (!(bp >= 0.0)) return 0.0;
      d[size_t(j)] += bp / ap;
      e[size_t(j - 1)] = std::sqrt(bp) / ap;
    }
  }
  double hi = 0.0;  // Gershgorin bound
  for (int j = 0; j < k; ++j)
    hi = std::max(hi, d[size_t(j)] + (j ? e[size_t(j - 1)] : 0.0) + (j + 1 < k ? e[size_t(j)] : 0.0));
  // smallest and largest eigenvalue by bisection, to 1e-3 relative
  double lo = 0.0, up = hi;
  for (int it = 0; it < 200 && up - lo > 1e-3 * up; ++it) {
    const double mid = 0.5 * (lo + up);
    if (sturm_below(d, e, mid) >= 1) up = mid;
    else lo = mid;
  }
  const double lmin = up;
  double lo2 = 0.0, up2 = hi;
  for (int it = 0; it < 200 && up2 - lo2 > 1e-3 * up2; ++it) {
    const double mid = 0.5 * (lo2 + up2);
    if (sturm_below(d, e, mid) >= k) up2 = mid;
    else lo2 = mid;
  }
  if (!(lmin > 0.0)) return 0.0;
  return up2 / lmin;
}

namespace {
constexpr int kChunk = 32;
constexpr double kDelta = 0.1;  // refresh when the recursive residual fell 10x
// kChunk iterations captured once into a graph; recaptured when the
// buffers (arrays, variant, grid) differ from the captured ones
hipError_t ensure_mixed_graph(gf_ctx* ctx, gf_mixed_ws* W, const MixedBuffers& B) {
  if (W->graph && std::memcmp(&W->graph_B, &B, sizeof(B)) == 0) return hipSuccess;
  if (W->graph) {
    (void)hipGraphExecDestroy(W->graph);
    W->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  for (int it = 0; it < kChunk; ++it) {
    e = launch_mixed_iteration(ctx->stream, B, it & 1);
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

int ensure_val32(gf_ctx* ctx, gf_ell* m, int variant, const float** out) {
  if (is_bsell_variant(variant)) {
    // FP32 copy of the block-sliced values
    const size_t cnt = size_t(m->bslots) * 64 * m->dim * m->dim;
    if (!m->d_bval32) GF_HIP_CHECK(ctx, hipMalloc(&m->d_bval32, cnt * sizeof(float)));
    if (m->bval32_version != m->val_version) {
      GF_HIP_CHECK(ctx, launch_bsell_to_f32(ctx->stream, m->dim, bsell_view(m), m->d_bcol,
                                            m->d_bval, m->d_bval32));
      m->bval32_version = m->val_version;
    }
    *out = m->d_bval32;
  } else {
    const size_t slots = size_t(m->n) * size_t(m->width);
    if (!m->d_val32) GF_HIP_CHECK(ctx, hipMalloc(&m->d_val32, slots * sizeof(float)));
    if (!m->val32_valid) {
      GF_HIP_CHECK(ctx, launch_to_f32(ctx->stream, int64_t(slots), m->d_val, m->d_val32));
      m->val32_valid = true;
    }
    *out = m->d_val32;
  }
  return GF_OK;
}

void release_mixed(gf_solve_plan* P) {
  gf_mixed_ws* W = P->mixed;
  if (!W) return;
  float* f[] = {W->dinv32, W->r32, W->z32, W->p32, W->q32};
  for (float* v : f)
    if (v) (void)hipFree(v);
  if (W->q64) (void)hipFree(W->q64);
  if (W->part3) (void)hipFree(W->part3);
  if (W->hist) (void)hipFree(W->hist);
  if (W->st) (void)hipFree(W->st);
  if (W->graph) (void)hipGraphExecDestroy(W->graph);
  delete W;
  P->mixed = nullptr;
}

int run_mixed(gf_ctx* ctx, gf_ell* m, const gf_solve_opts& opts, PcgHostState* h,
              int* refinements) {
  gf_solve_plan* P = m->plan;
  const int64_t n = m->n;
  if (!P->mixed) {
    gf_mixed_ws* W = new gf_mixed_ws();
    P->mixed = W;
    W->n = n;
    const size_t fb = size_t(n) * sizeof(float);
    float** f[] = {&W->dinv32, &W->r32, &W->z32, &W->p32, &W->q32};
    hipError_t e = hipSuccess;
    for (float** v : f)
      if (e == hipSuccess) e = hipMalloc(v, fb);
    if (e == hipSuccess) e = hipMalloc(&W->q64, size_t(n) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&W->part3, 3 * kMaxPcgBlocks * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&W->hist, 2 * size_t(kLanczosIters) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&W->st), mixed_state_bytes());
    if (e != hipSuccess) {
      release_mixed(P);
      return set_error(ctx, GF_ERR_NOMEM, std::string("gf_pcg (mixed): workspace: ") + hipGetErrorString(e));
    }
  }
  gf_mixed_ws* W = P->mixed;
  int st = prepare_variant(ctx, m, P->variant);
  if (st) return st;
  const SpmvOperands o = spmv_operands(m, P->variant);
  const float* val32 = nullptr;
  if ((st = ensure_val32(ctx, m, P->variant, &val32))) return st;
  if (W->dinv_version != m->val_version) {
    GF_HIP_CHECK(ctx, launch_to_f32(ctx->stream, n, m->d_dinv, W->dinv32));
    W->dinv_version = m->val_version;
  }
  MixedBuffers B{};  // zero padding: the graph cache compares the bytes
  B.n = n;
  B.variant = P->variant;
  B.nblocks = P->nblocks;
  B.spmv_grid = spmv_grid(P->variant, o.bs, P->nblocks);
  B.val = o.val;
  B.col = o.col;
  B.bs = o.bs;
  B.rownnz = m->d_rownnz;
  B.b = P->wb;
  B.x = P->wx;
  B.q64 = W->q64;
  B.val32 = val32;
  B.dinv32 = W->dinv32;
  B.r32 = W->r32;
  B.z32 = W->z32;
  B.p32 = W->p32;
  B.q32 = W->q32;
  B.part1 = P->part1;
  B.part2 = P->part2;
  B.part3 = W->part3;
  B.st = W->st;
  B.hist = W->hist;
  GF_HIP_CHECK(ctx, launch_mixed_setup(ctx->stream, B, opts.use_x0, opts.rtol, kDelta, opts.maxit));
  MixedHostState hs{};
  GF_HIP_CHECK(ctx, read_mixed_state(ctx->stream, B, &hs));
  // the graph replays the same kChunk launches (kernel schedule: launched
  // one by one); the results are identical
  const bool graph = opts.schedule != GF_SCHED_KERNELS;
  if (graph && !hs.done) GF_HIP_CHECK(ctx, ensure_mixed_graph(ctx, W, B));
  std::vector<double> hist;
  int checked = 0;
  bool ill = false;
  while (!hs.done) {
    if (graph) {
      GF_HIP_CHECK(ctx, hipGraphLaunch(W->graph, ctx->stream));
    } else {
      for (int it = 0; it < kChunk; ++it)
        GF_HIP_CHECK(ctx, launch_mixed_iteration(ctx->stream, B, it & 1));
    }
    GF_HIP_CHECK(ctx, read_mixed_state(ctx->stream, B, &hs));
    // condition check: continue in FP64 once FP32 cannot resolve the
    // operator's smallest eigenvalues
    const int k = std::min(hs.it, kLanczosIters);
    if (!hs.done && k >= checked + kLanczosCheck) {
      hist.resize(2 * size_t(k));
      GF_HIP_CHECK(ctx, hipMemcpy(hist.data() + 2 * size_t(checked), W->hist + 2 * size_t(checked),
                                  2 * size_t(k - checked) * sizeof(double), hipMemcpyDeviceToHost));
      checked = k;
      const double kappa = lanczos_condition(hist.data(), hist.data() + 1, k, 2);
      if (kappa * kU32 > kKappaU32) {
        ill = true;
        break;
      }
    }
  }
  if (ill) hs.status = kMixedStalled;
  h->it = hs.it;
  h->done = hs.done;
  h->status = hs.status;
  h->rr = hs.rr_true;
  h->bnorm2 = hs.bnorm2;
  h->rho = 0.0;
  *refinements = hs.refinements;
  return GF_OK;
}

}  // namespace gf
