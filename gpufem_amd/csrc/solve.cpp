// Solver driver: gf_pcg.  Chooses the schedule, keeps the workspace on the
// matrix, and runs each right-hand side through the device-resident PCG of
// solver.hip.  Semantics follow SPRmatrix::CPU_CG
// (src/SPRmatrix/SPRmatrix.cc:405-480) extended with the Jacobi
// preconditioner: x0 = 0 (or the caller's guess), stop when
// ||r||_2 <= rtol ||b||_2, report the true residual at exit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gf_internal.hpp"
#include "launch.hpp"
#include "solve_plan.hpp"

struct gf_solve_plan {
  int64_t n = 0;
  double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *wb = nullptr, *wx = nullptr;
  double *part1 = nullptr, *part2 = nullptr;
  double* res = nullptr;  // true residual per call
  gf::PcgState* st = nullptr;
  int nblocks = 0;
  // captured chunk of iterations, keyed by preconditioner use
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_chunk = 0;
};

namespace gf {

int set_error(gf_ctx* ctx, int code, const std::string& msg);

namespace {

constexpr int kMaxBlocks = 2048;
constexpr int kChunk = 32;  // iterations between convergence checks (even)

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int ensure_plan(gf_ctx* ctx, gf_ell* m) {
  if (m->plan && m->plan->n == m->n) return GF_OK;
  release_solve_plan(m);
  gf_solve_plan* P = new gf_solve_plan();
  P->n = m->n;
  const size_t vb = size_t(m->n) * sizeof(double);
  hipError_t e = hipSuccess;
  double** vecs[] = {&P->r, &P->z, &P->p, &P->q, &P->wb, &P->wx};
  for (double** v : vecs)
    if (e == hipSuccess) e = hipMalloc(v, vb);
  if (e == hipSuccess) e = hipMalloc(&P->part1, kMaxBlocks * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&P->part2, 3 * kMaxBlocks * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&P->res, 64 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&P->st), pcg_state_bytes());
  m->plan = P;
  if (e != hipSuccess) {
    release_solve_plan(m);
    return set_error(ctx, GF_ERR_NOMEM, std::string("gf_pcg: workspace: ") + hipGetErrorString(e));
  }
  int64_t nb = (m->n + 255) / 256;
  P->nblocks = int(std::max<int64_t>(1, std::min<int64_t>(nb, kMaxBlocks)));
  return GF_OK;
}

PcgBuffers make_buffers(gf_ell* m, const double* dinv) {
  gf_solve_plan* P = m->plan;
  PcgBuffers B;
  B.n = m->n;
  B.width = m->width;
  B.val = m->d_val;
  B.col = m->d_col;
  B.rownnz = m->d_rownnz;
  B.dinv = dinv;
  B.b = P->wb;
  B.x = P->wx;
  B.r = P->r;
  B.z = P->z;
  B.p = P->p;
  B.q = P->q;
  B.part1 = P->part1;
  B.part2 = P->part2;
  B.nblocks = P->nblocks;
  B.st = P->st;
  return B;
}

hipError_t ensure_graph(gf_ctx* ctx, gf_ell* m, const PcgBuffers& B, int key) {
  gf_solve_plan* P = m->plan;
  if (P->graph[key]) return hipSuccess;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  for (int it = 0; it < kChunk; ++it) {
    e = launch_pcg_iteration(ctx->stream, B, it & 1);
    if (e != hipSuccess) break;
  }
  hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(&P->graph[key], g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  P->graph_chunk = kChunk;
  return e;
}

}  // namespace

void release_solve_plan(gf_ell* m) {
  gf_solve_plan* P = m->plan;
  if (!P) return;
  for (auto& g : P->graph)
    if (g) (void)hipGraphExecDestroy(g);
  double* vecs[] = {P->r, P->z, P->p, P->q, P->wb, P->wx, P->part1, P->part2, P->res};
  for (double* v : vecs)
    if (v) (void)hipFree(v);
  if (P->st) (void)hipFree(P->st);
  delete P;
  m->plan = nullptr;
}

}  // namespace gf

using namespace gf;

extern "C" {

void gf_solve_opts_default(gf_solve_opts* o) {
  if (!o) return;
  o->precond = GF_PRECOND_JACOBI;
  o->precision = GF_PREC_FP64;
  o->schedule = GF_SCHED_AUTO;
  o->use_x0 = 0;
  o->maxit = 100000;
  o->device_ptrs = 0;
  o->rtol = 1e-8;
  o->inner_rtol = 1e-4;
}

int gf_pcg(gf_ctx* ctx, gf_ell* m, const double* b, double* x, int32_t nrhs,
           const gf_solve_opts* opts_in, gf_solve_stats* stats) {
  if (!ctx || !m || !b || !x || nrhs < 1)
    return set_error(ctx, GF_ERR_ARG, "gf_pcg: bad argument");
  gf_solve_opts opts;
  if (opts_in) opts = *opts_in;
  else gf_solve_opts_default(&opts);
  if (!(opts.rtol >= 0.0) || opts.maxit < 0)
    return set_error(ctx, GF_ERR_ARG, "gf_pcg: rtol must be >= 0 and maxit >= 0");
  if (opts.precond != GF_PRECOND_NONE && opts.precond != GF_PRECOND_JACOBI)
    return set_error(ctx, GF_ERR_ARG, "gf_pcg: unknown preconditioner");
  if (opts.precision != GF_PREC_FP64)
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_pcg: precision mode not available");
  if (nrhs > 64) return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_pcg: at most 64 right-hand sides");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const double h0 = now_ms();
  int st = ensure_plan(ctx, m);
  if (st) return st;
  gf_solve_plan* P = m->plan;
  const int64_t n = m->n;
  const size_t vb = size_t(n) * sizeof(double);

  const double* dinv = nullptr;
  if (opts.precond == GF_PRECOND_JACOBI) {
    if (!m->d_dinv) GF_HIP_CHECK(ctx, hipMalloc(&m->d_dinv, vb));
    if (!m->dinv_valid) {
      GF_HIP_CHECK(ctx, launch_jacobi_inv(ctx->stream, n, m->d_val, m->d_diag, m->d_dinv));
      m->dinv_valid = true;
    }
    dinv = m->d_dinv;
  }
  const int key = dinv ? 1 : 0;
  PcgBuffers B = make_buffers(m, dinv);
  int sched = opts.schedule == GF_SCHED_AUTO ? GF_SCHED_GRAPH : opts.schedule;
  if (sched == GF_SCHED_PERSISTENT) sched = GF_SCHED_GRAPH;  // not built yet: same numerics
  if (sched == GF_SCHED_GRAPH) GF_HIP_CHECK(ctx, ensure_graph(ctx, m, B, key));

  // right-hand sides and solutions resident on the device
  const double* Bd = b;
  double* Xd = x;
  double* tmp = nullptr;
  if (!opts.device_ptrs) {
    GF_HIP_CHECK(ctx, hipMalloc(&tmp, 2 * vb * size_t(nrhs)));
    GF_HIP_CHECK(ctx, hipMemcpyAsync(tmp, b, vb * nrhs, hipMemcpyHostToDevice, ctx->stream));
    if (opts.use_x0)
      GF_HIP_CHECK(ctx, hipMemcpyAsync(tmp + n * nrhs, x, vb * nrhs, hipMemcpyHostToDevice, ctx->stream));
    Bd = tmp;
    Xd = tmp + n * nrhs;
  }
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  std::vector<PcgHostState> hs(static_cast<size_t>(nrhs));
  hipError_t e = hipSuccess;
  for (int k = 0; k < nrhs && e == hipSuccess; ++k) {
    e = hipMemcpyAsync(P->wb, Bd + k * n, vb, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess && opts.use_x0)
      e = hipMemcpyAsync(P->wx, Xd + k * n, vb, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_pcg_setup(ctx->stream, B, opts.use_x0, opts.rtol, opts.maxit);
    PcgHostState h{};
    if (e == hipSuccess) e = read_pcg_state(ctx->stream, B, &h);
    while (e == hipSuccess && !h.done) {
      if (sched == GF_SCHED_GRAPH) {
        e = hipGraphLaunch(P->graph[key], ctx->stream);
      } else {
        for (int it = 0; it < kChunk && e == hipSuccess; ++it)
          e = launch_pcg_iteration(ctx->stream, B, it & 1);
      }
      if (e == hipSuccess) e = read_pcg_state(ctx->stream, B, &h);
    }
    if (e == hipSuccess) e = launch_true_residual(ctx->stream, B, P->res + k);
    if (e == hipSuccess)
      e = hipMemcpyAsync(Xd + k * n, P->wx, vb, hipMemcpyDeviceToDevice, ctx->stream);
    hs[size_t(k)] = h;
  }
  if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ctx->ev_stop);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
  std::vector<double> res(size_t(nrhs), 0.0);
  if (e == hipSuccess)
    e = hipMemcpy(res.data(), P->res, size_t(nrhs) * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess && !opts.device_ptrs)
    e = hipMemcpy(x, Xd, vb * nrhs, hipMemcpyDeviceToHost);
  if (tmp) (void)hipFree(tmp);
  if (e != hipSuccess) return set_error(ctx, GF_ERR_HIP, std::string("gf_pcg: ") + hipGetErrorString(e));

  gf_solve_stats S{};
  S.converged = 1;
  S.status = GF_OK;
  for (int k = 0; k < nrhs; ++k) {
    const PcgHostState& h = hs[size_t(k)];
    S.iterations = std::max(S.iterations, h.it);
    const double rel = h.bnorm2 > 0 ? std::sqrt(h.rr / h.bnorm2) : std::sqrt(h.rr);
    S.relres = std::max(S.relres, rel);
    S.true_relres = std::max(S.true_relres, res[size_t(k)]);
    if (h.status != GF_OK) {
      S.status = h.status;
      S.converged = 0;
    } else if (!(h.rr <= opts.rtol * opts.rtol * h.bnorm2)) {
      if (S.status == GF_OK) S.status = GF_ERR_NOCONV;
      S.converged = 0;
    }
  }
  S.ms_device = ms;
  S.ms_host = now_ms() - h0;
  if (stats) *stats = S;
  if (S.status == GF_ERR_BREAKDOWN)
    return set_error(ctx, GF_ERR_BREAKDOWN, "gf_pcg: breakdown (p.Kp <= 0 or NaN)");
  if (S.status == GF_ERR_NOCONV)
    return set_error(ctx, GF_ERR_NOCONV, "gf_pcg: iteration limit reached before rtol");
  return GF_OK;
}

int gf_profile_spmv(gf_ctx* ctx, gf_ell* m, int32_t nrep, double* ms_avg) {
  if (!ctx || !m || nrep < 1 || !ms_avg)
    return set_error(ctx, GF_ERR_ARG, "gf_profile_spmv: bad argument");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int st = ensure_plan(ctx, m);
  if (st) return st;
  PcgBuffers B = make_buffers(m, nullptr);
  // a fixed, finite direction vector
  GF_HIP_CHECK(ctx, hipMemsetAsync(m->plan->p, 0, size_t(m->n) * sizeof(double), ctx->stream));
  GF_HIP_CHECK(ctx, launch_spmv_pq_repeat(ctx->stream, B, 1));  // warm caches / code
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  GF_HIP_CHECK(ctx, launch_spmv_pq_repeat(ctx->stream, B, nrep));
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
  float ms = 0.f;
  GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  *ms_avg = double(ms) / nrep;
  return GF_OK;
}

}  // extern "C"
