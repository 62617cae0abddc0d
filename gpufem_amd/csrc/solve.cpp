// Solver driver: gf_pcg, gf_ell_tune, gf_profile_spmv.
//
// Semantics follow SPRmatrix::CPU_CG (src/SPRmatrix/SPRmatrix.cc:405-480)
// extended with the Jacobi preconditioner: x0 = 0 (or the caller's guess),
// stop when ||r||_2 <= rtol ||b||_2 (recursive r); with opts.verify the true
// residual b - K x is then checked and PCG restarted from x while it exceeds
// rtol.  The true residual at exit is always reported.  The
// workspace (vectors, partials, scalar state, captured graphs) is cached on
// the matrix; the SpMV variant is picked once per matrix by timing every
// variant on the matrix itself.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gf_internal.hpp"
#include "launch.hpp"
#include "mixed.hpp"
#include "multi.hpp"
#include "multi_mixed.hpp"
#include "solve_plan.hpp"

namespace gf {

int set_error(gf_ctx* ctx, int code, const std::string& msg);

namespace {

constexpr int kChunk = 32;  // iterations between convergence checks (even)
constexpr int kMaxRestarts = 4;  // true-residual restarts per right-hand side

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void drop_graphs(gf_solve_plan* P) {
  for (auto& g : P->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
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
  return e;
}

}  // namespace

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
  if (e == hipSuccess) e = hipMalloc(&P->part1, kMaxSpmvBlocks * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&P->part2, 3 * kMaxPcgBlocks * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&P->res, 64 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&P->st), pcg_state_bytes());
  if (e == hipSuccess) e = hipMalloc(&P->barrier, grid_barrier_bytes());
  m->plan = P;
  if (e != hipSuccess) {
    release_solve_plan(m);
    return set_error(ctx, GF_ERR_NOMEM, std::string("gf_pcg: workspace: ") + hipGetErrorString(e));
  }
  P->num_cu = ctx->num_cu;
  return GF_OK;
}

// Fixes the variant and the PCG grid (the same grid for every schedule, so
// all schedules give bitwise-identical results).
int pcg_grid(int variant, int64_t n, int num_cu) {
  int nb = spmv_blocks(n);
  const int pc = persistent_blocks(variant, num_cu);
  if (pc > 0 && nb > pc) nb = pc;
  if (nb >= 8) nb &= ~7;
  return std::max(1, nb);
}

void set_variant(gf_ell* m, int variant) {
  gf_solve_plan* P = m->plan;
  if (P->variant == variant && P->nblocks > 0) return;
  drop_graphs(P);
  P->variant = variant;
  P->persistent_ok = persistent_blocks(variant, P->num_cu) > 0;
  P->nblocks = pcg_grid(variant, m->n, P->num_cu);
}

// The node-block / block-sliced variants of the matrix's dimension, or -1
// (matrices made by gf_ell_upload carry no node structure).
int node_variant(const gf_ell* m) {
  if (m->dim == 2 && m->width % 2 == 0) return kNode2;
  if (m->dim == 3 && m->width % 3 == 0) return kNode3;
  return -1;
}

int bsell_variant(const gf_ell* m) {
  if (m->dim == 2 && m->width % 2 == 0 && m->n % 2 == 0) return kBsell2;
  if (m->dim == 3 && m->width % 3 == 0 && m->n % 3 == 0) return kBsell3;
  return -1;
}

bool variant_fits(const gf_ell* m, int v) {
  if (v < 0 || v >= kNumSpmvVariants) return false;
  if (is_node_variant(v)) return v == node_variant(m);
  if (v == kBsell2G) return bsell_variant(m) == kBsell2;
  if (v == kBsell3G) return bsell_variant(m) == kBsell3;
  if (is_bsell_variant(v)) return v == bsell_variant(m);
  return true;
}

int ensure_coln(gf_ctx* ctx, gf_ell* m) {
  if (m->d_coln) return GF_OK;
  const int d = m->dim;
  GF_HIP_CHECK(ctx, hipMalloc(&m->d_coln, size_t(m->n) * (m->width / d) * sizeof(int32_t)));
  GF_HIP_CHECK(ctx, launch_build_coln(ctx->stream, m->n, m->width, d, m->d_col, m->d_coln));
  return GF_OK;
}

BsellView bsell_view(const gf_ell* m) {
  return BsellView{m->d_bperm, m->d_bsptr, m->bslices, m->d_bcoff};
}

// Structure of the block-sliced copy: node rows sorted by length inside
// windows, slices of 64 positions, one block slot per entry of the longest
// row of the slice; then the node columns.
int ensure_bsell(gf_ctx* ctx, gf_ell* m) {
  if (m->d_bperm) return GF_OK;
  const int d = m->dim;
  const int64_t N = m->n / d;
  std::vector<int32_t> rn(static_cast<size_t>(m->n));
  GF_HIP_CHECK(ctx, hipMemcpy(rn.data(), m->d_rownnz, size_t(m->n) * 4, hipMemcpyDeviceToHost));
  std::vector<int32_t> cnt(static_cast<size_t>(N)), perm;
  for (int64_t R = 0; R < N; ++R) cnt[size_t(R)] = rn[size_t(d * R)] / d;
  // windows of kBsellWindow node rows, restarting at every break
  std::vector<int64_t> brk = m->bsell_breaks;
  if (brk.empty() || brk.front() != 0) brk.insert(brk.begin(), 0);
  if (brk.back() != N) brk.push_back(N);
  std::vector<std::pair<int64_t, int64_t>> wins;
  std::vector<size_t> brk_win;  // first window of each segment
  for (size_t k = 0; k + 1 < brk.size(); ++k) {
    brk_win.push_back(wins.size());
    for (int64_t r = brk[k]; r < brk[k + 1]; r += kBsellWindow)
      wins.emplace_back(r, std::min<int64_t>(brk[k + 1], r + kBsellWindow));
  }
  brk_win.push_back(wins.size());
  std::vector<int64_t> win_slice;
  std::vector<int64_t> sptr(1, 0);
  std::vector<int32_t> order;
  for (size_t w = 0; w < wins.size(); ++w) {
    win_slice.push_back(int64_t(sptr.size()) - 1);
    const int64_t r0 = wins[w].first, r1 = wins[w].second;
    order.resize(size_t(r1 - r0));
    for (int64_t R = r0; R < r1; ++R) order[size_t(R - r0)] = int32_t(R);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return cnt[size_t(a)] > cnt[size_t(b)]; });
    for (size_t s0 = 0; s0 < order.size(); s0 += 64) {
      for (size_t l = 0; l < 64; ++l) perm.push_back(s0 + l < order.size() ? order[s0 + l] : -1);
      sptr.push_back(sptr.back() + cnt[size_t(order[s0])]);
    }
  }
  win_slice.push_back(int64_t(sptr.size()) - 1);
  std::vector<int32_t> inv(static_cast<size_t>(N), -1);
  for (size_t pos = 0; pos < perm.size(); ++pos)
    if (perm[pos] >= 0) inv[size_t(perm[pos])] = int32_t(pos);
  // every array is allocated and filled before the copy is published
  // (m->d_bperm set last): a failure leaves no half-built copy behind
  struct Arrays {
    int32_t *perm = nullptr, *inv = nullptr, *col = nullptr;
    int64_t* sptr = nullptr;
    int16_t* coff = nullptr;
    double* val = nullptr;
    ~Arrays() {
      for (void* p : {static_cast<void*>(perm), static_cast<void*>(inv), static_cast<void*>(col),
                      static_cast<void*>(sptr), static_cast<void*>(coff), static_cast<void*>(val)})
        if (p) (void)hipFree(p);
    }
  } A;
  const int64_t bslices = int64_t(sptr.size()) - 1, bslots = sptr.back();
  GF_HIP_CHECK(ctx, hipMalloc(&A.perm, perm.size() * 4));
  GF_HIP_CHECK(ctx, hipMalloc(&A.inv, std::max<size_t>(inv.size(), 1) * 4));
  GF_HIP_CHECK(ctx, hipMalloc(&A.sptr, sptr.size() * 8));
  GF_HIP_CHECK(ctx, hipMalloc(&A.col, size_t(bslots) * 64 * 4));
  GF_HIP_CHECK(ctx, hipMalloc(&A.val, size_t(bslots) * 64 * d * d * sizeof(double)));
  GF_HIP_CHECK(ctx, hipMalloc(&A.coff, size_t(bslots) * 64 * sizeof(int16_t)));
  GF_HIP_CHECK(ctx, hipMemcpy(A.perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice));
  GF_HIP_CHECK(ctx, hipMemcpy(A.inv, inv.data(), inv.size() * 4, hipMemcpyHostToDevice));
  GF_HIP_CHECK(ctx, hipMemcpy(A.sptr, sptr.data(), sptr.size() * 8, hipMemcpyHostToDevice));
  const BsellView v0{A.perm, A.sptr, bslices, nullptr};
  GF_HIP_CHECK(ctx, launch_bsell_cols(ctx->stream, v0, d, m->n, m->d_col, m->d_rownnz, A.col));
  // 16-bit node-column offsets for the single-RHS kernels when they all fit
  int* d_over = reinterpret_cast<int*>(static_cast<char*>(ctx->d_state) + 160);
  GF_HIP_CHECK(ctx, hipMemsetAsync(d_over, 0, sizeof(int), ctx->stream));
  GF_HIP_CHECK(ctx, launch_bsell_offsets(ctx->stream, v0, A.col, A.coff, d_over));
  int over = 0;
  GF_HIP_CHECK(ctx, hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (over) {
    GF_HIP_CHECK(ctx, hipFree(A.coff));
    A.coff = nullptr;
  }
  m->bslices = bslices;
  m->bslots = bslots;
  m->bsell_break_slice.clear();
  for (size_t k = 0; k < brk_win.size(); ++k) m->bsell_break_slice.push_back(win_slice[brk_win[k]]);
  m->d_binv = A.inv;
  m->d_bsptr = A.sptr;
  m->d_bcol = A.col;
  m->d_bcoff = A.coff;
  m->d_bval = A.val;
  m->d_bperm = A.perm;
  A.perm = A.inv = A.col = nullptr;
  A.sptr = nullptr;
  A.coff = nullptr;
  A.val = nullptr;
  m->bval_version = 0;
  m->bval32_version = 0;
  return GF_OK;
}

int ensure_ell_vals(gf_ctx* ctx, gf_ell* m) {
  if (!m->ell_stale) return GF_OK;
  GF_HIP_CHECK(ctx, launch_bsell_to_ell(ctx->stream, bsell_view(m), m->dim, m->n, m->d_bval,
                                        m->d_rownnz, m->d_val));
  m->ell_stale = false;
  return GF_OK;
}

// Values of the copy follow the ELL values (assembly and penalty bump the version).
int ensure_bsell_vals(gf_ctx* ctx, gf_ell* m) {
  int st = ensure_bsell(ctx, m);
  if (st) return st;
  if (m->bval_version != m->val_version) {
    // stale ELL values come with a current copy; anything else is a bug
    if (m->ell_stale)
      return set_error(ctx, GF_ERR_HIP, "block-sliced copy and ELL values both stale");
    GF_HIP_CHECK(ctx, launch_bsell_vals(ctx->stream, bsell_view(m), m->dim, m->n, m->d_val,
                                        m->d_rownnz, m->d_bval));
    m->bval_version = m->val_version;
  }
  return GF_OK;
}

int prepare_variant(gf_ctx* ctx, gf_ell* m, int v) {
  if (is_bsell_variant(v)) return ensure_bsell_vals(ctx, m);
  int st = ensure_ell_vals(ctx, m);
  if (st) return st;
  if (is_node_variant(v)) return ensure_coln(ctx, m);
  return GF_OK;
}

SpmvOperands spmv_operands(const gf_ell* m, int variant) {
  SpmvOperands o;
  o.val = m->d_val;
  o.col = m->d_col;
  o.bs = BsellView{nullptr, nullptr, 0, nullptr};
  if (is_node_variant(variant)) {
    o.col = m->d_coln;
  } else if (is_bsell_variant(variant)) {
    o.val = m->d_bval;
    o.col = m->d_bcol;
    o.bs = bsell_view(m);
  }
  return o;
}

PcgBuffers make_buffers(gf_ell* m, const double* dinv) {
  gf_solve_plan* P = m->plan;
  const SpmvOperands o = spmv_operands(m, P->variant);
  PcgBuffers B;
  B.n = m->n;
  B.width = m->width;
  B.variant = P->variant;
  B.val = o.val;
  B.col = o.col;
  B.bs = o.bs;
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
  B.spmv_grid = spmv_grid(P->variant, o.bs, P->nblocks);
  B.st = P->st;
  return B;
}

int tune_spmv(gf_ctx* ctx, gf_ell* m, int force, double* ms_out, int* chosen) {
  int st = ensure_plan(ctx, m);
  if (st) return st;
  gf_solve_plan* P = m->plan;
  if (force >= 0) {
    if (force >= kNumSpmvVariants) return set_error(ctx, GF_ERR_ARG, "gf_ell_tune: bad variant");
    if (!variant_fits(m, force))
      return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_ell_tune: the variant does not fit the matrix");
    if ((st = prepare_variant(ctx, m, force))) return st;
    set_variant(m, force);
    if (chosen) *chosen = force;
    return GF_OK;
  }
  GF_HIP_CHECK(ctx, hipMemsetAsync(P->p, 0, size_t(m->n) * sizeof(double), ctx->stream));
  const int reps = m->n > 200000 ? 10 : 30;
  double best = 1e300;
  int best_v = kThreadU8;
  for (int v = 0; v < kNumSpmvVariants; ++v) {
    if (!variant_fits(m, v)) {
      if (ms_out) ms_out[v] = -1.0;
      continue;
    }
    if ((st = prepare_variant(ctx, m, v))) return st;
    set_variant(m, v);
    PcgBuffers B = make_buffers(m, nullptr);
    GF_HIP_CHECK(ctx, launch_spmv_pq_repeat(ctx->stream, B, 2));
    GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    GF_HIP_CHECK(ctx, launch_spmv_pq_repeat(ctx->stream, B, reps));
    GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
    float ms = 0.f;
    GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    const double avg = double(ms) / reps;
    if (ms_out) ms_out[v] = avg;
    if (avg < best) {
      best = avg;
      best_v = v;
    }
  }
  set_variant(m, best_v);
  if (chosen) *chosen = best_v;
  return GF_OK;
}

// Runs one right-hand side already placed in the plan's wb (and wx when
// use_x0) to completion with the given schedule.
int run_pcg(gf_ctx* ctx, gf_ell* m, const PcgBuffers& B, int sched, int key, int use_x0,
            double rtol, int maxit, int verify, PcgHostState* h) {
  gf_solve_plan* P = m->plan;
  int total = 0;
  bool set_up = false;
  if (use_x0 == 2) {
    // guarded initial guess: kept only when ||b - K x0|| < ||b||, otherwise
    // the solve starts from zero (the setup's own SpMV is the check)
    GF_HIP_CHECK(ctx, launch_pcg_setup(ctx->stream, B, 1, rtol, maxit));
    GF_HIP_CHECK(ctx, read_pcg_state(ctx->stream, B, h));
    set_up = h->rr < h->bnorm2;
    use_x0 = set_up ? 1 : 0;
    h->x0_rejected = set_up ? 0 : 1;
  }
  for (int restart = 0;; ++restart) {
    if (!set_up) GF_HIP_CHECK(ctx, launch_pcg_setup(ctx->stream, B, use_x0, rtol, maxit - total));
    set_up = false;
    if (sched == GF_SCHED_PERSISTENT) {
      GF_HIP_CHECK(ctx, launch_pcg_persistent(ctx->stream, B, P->barrier));
      GF_HIP_CHECK(ctx, read_pcg_state(ctx->stream, B, h));
      if (h->status == GF_ERR_HIP)
        return set_error(ctx, GF_ERR_HIP, "gf_pcg: grid barrier timed out in the persistent kernel");
    } else {
      GF_HIP_CHECK(ctx, read_pcg_state(ctx->stream, B, h));
      while (!h->done) {
        if (sched == GF_SCHED_GRAPH) {
          GF_HIP_CHECK(ctx, ensure_graph(ctx, m, B, B.dinv ? 1 : 0));
          GF_HIP_CHECK(ctx, hipGraphLaunch(P->graph[key], ctx->stream));
        } else {
          for (int it = 0; it < kChunk; ++it)
            GF_HIP_CHECK(ctx, launch_pcg_iteration(ctx->stream, B, it & 1));
        }
        GF_HIP_CHECK(ctx, read_pcg_state(ctx->stream, B, h));
      }
    }
    total += h->it;
    h->it = total;
    // The recursive residual drifts from b - K x by rounding accumulated over
    // many iterations; a restart from the current x resets the drift.
    if (!verify || restart >= kMaxRestarts || h->status != GF_OK || total >= maxit ||
        !(h->rr <= rtol * rtol * h->bnorm2))
      return GF_OK;
    double tr = 0.0;
    GF_HIP_CHECK(ctx, launch_true_residual(ctx->stream, B, P->res));
    GF_HIP_CHECK(ctx, hipMemcpyAsync(&tr, P->res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (tr <= rtol) return GF_OK;
    use_x0 = 1;
  }
}

int prepare_jacobi(gf_ctx* ctx, gf_ell* m) {
  if (!m->d_dinv) GF_HIP_CHECK(ctx, hipMalloc(&m->d_dinv, size_t(m->n) * sizeof(double)));
  if (!m->dinv_valid) {
    GF_HIP_CHECK(ctx, launch_jacobi_inv(ctx->stream, m->n, m->d_val, m->d_diag, m->d_dinv));
    m->dinv_valid = true;
  }
  return GF_OK;
}

void release_solve_plan(gf_ell* m) {
  gf_solve_plan* P = m->plan;
  if (!P) return;
  drop_graphs(P);
  double* vecs[] = {P->r, P->z, P->p, P->q, P->wb, P->wx, P->part1, P->part2, P->res};
  for (double* v : vecs)
    if (v) (void)hipFree(v);
  if (P->st) (void)hipFree(P->st);
  if (P->barrier) (void)hipFree(P->barrier);
  release_mixed(P);
  release_multi(P);
  release_multi_mixed(P);
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
  o->batched = 1;
  o->verify = 1;
}

int gf_ell_tune(gf_ctx* ctx, gf_ell* m, int32_t force_variant, double* ms_per_variant,
                int32_t* chosen) {
  if (!ctx || !m) return set_error(ctx, GF_ERR_ARG, "gf_ell_tune: null argument");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int c = -1;
  int st = tune_spmv(ctx, m, force_variant, ms_per_variant, &c);
  if (st) return st;
  if (chosen) *chosen = c;
  return GF_OK;
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
  if (opts.precision != GF_PREC_FP64 && opts.precision != GF_PREC_MIXED)
    return set_error(ctx, GF_ERR_ARG, "gf_pcg: unknown precision mode");
  if (opts.schedule < GF_SCHED_AUTO || opts.schedule > GF_SCHED_PERSISTENT)
    return set_error(ctx, GF_ERR_ARG, "gf_pcg: unknown schedule");
  if (nrhs > 64) return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_pcg: at most 64 right-hand sides");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const double h0 = now_ms();
  int st = ensure_plan(ctx, m);
  if (st) return st;
  gf_solve_plan* P = m->plan;
  if (P->variant < 0 && (st = tune_spmv(ctx, m, -1, nullptr, nullptr))) return st;
  const int64_t n = m->n;
  const size_t vb = size_t(n) * sizeof(double);
  int sched = opts.schedule == GF_SCHED_AUTO ? GF_SCHED_GRAPH : opts.schedule;
  if (sched == GF_SCHED_PERSISTENT && !P->persistent_ok)
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_pcg: persistent kernel cannot be co-resident");

  // right-hand sides and solutions resident on the device
  const double* Bd = b;
  double* Xd = x;
  struct DevTmp {
    double* p = nullptr;
    ~DevTmp() {
      if (p) (void)hipFree(p);
    }
  } tmp;
  if (!opts.device_ptrs) {
    GF_HIP_CHECK(ctx, hipMalloc(&tmp.p, 2 * vb * size_t(nrhs)));
    GF_HIP_CHECK(ctx, hipMemcpyAsync(tmp.p, b, vb * nrhs, hipMemcpyHostToDevice, ctx->stream));
    if (opts.use_x0)
      GF_HIP_CHECK(ctx, hipMemcpyAsync(tmp.p + n * nrhs, x, vb * nrhs, hipMemcpyHostToDevice, ctx->stream));
    Bd = tmp.p;
    Xd = tmp.p + n * nrhs;
  }
  // the solve's time includes every copy it derives from the matrix values
  // (Jacobi inverse, block-sliced copy after a change the assembly or the
  // penalty did not carry into it, FP32 copies)
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  if (opts.precond == GF_PRECOND_JACOBI || opts.precision == GF_PREC_MIXED) {
    st = prepare_jacobi(ctx, m);
    if (st) return st;
  }
  const double* dinv = opts.precond == GF_PRECOND_JACOBI ? m->d_dinv : nullptr;
  const int key = dinv ? 1 : 0;
  if ((st = prepare_variant(ctx, m, P->variant))) return st;
  PcgBuffers B = make_buffers(m, dinv);
  if (nrhs > 1 && opts.batched) {
    MultiHostState mh;
    std::vector<double> res(static_cast<size_t>(nrhs), 0.0);
    int fallbacks_b = 0, refinements_b = 0, mixed_it_b = 0;
    bool done = false, warm_b = false;
    if (opts.precision == GF_PREC_MIXED) {
      MixedMultiHostState xh;
      st = run_multi_mixed(ctx, m, opts, Bd, Xd, nrhs, &xh, res.data());
      if (st) return st;
      bool broke = false, stalled = false;
      stalled = xh.ill != 0;
      for (int k = 0; k < nrhs; ++k) {
        broke |= xh.col_status[k] == GF_ERR_BREAKDOWN;
        stalled |= xh.col_status[k] == kMixedStalled;
      }
      mixed_it_b = xh.it;
      if (!broke && !stalled) {
        // the same per-column record as the FP64 path; the residual is the
        // FP64 one of the last refresh (the only one convergence is declared on)
        mh.it = xh.it;
        mh.done = xh.done;
        for (int k = 0; k < nrhs; ++k) {
          mh.col_it[k] = xh.col_it[k];
          mh.col_status[k] = xh.col_status[k];
          mh.rr[k] = xh.rr_true[k];
          mh.bnorm2[k] = xh.bnorm2[k];
          refinements_b = std::max(refinements_b, xh.col_refinements[k]);
        }
        done = true;
      } else {
        // an FP32 operator that is not positive definite (breakdown):
        // zero-start FP64 batched PCG for every column; an FP32 iteration
        // that stalled: FP64 batched PCG from the current X (the columns
        // that converged are done at its setup)
        fallbacks_b = 1;
        warm_b = !broke;
      }
    }
    if (!done) {
      gf_solve_opts o64 = opts;
      o64.precision = GF_PREC_FP64;
      if (fallbacks_b) {
        o64.use_x0 = warm_b ? 1 : 0;
        o64.maxit = std::max(0, opts.maxit - mixed_it_b);
      }
      st = run_multi(ctx, m, o64, Bd, Xd, nrhs, &mh, res.data());
      if (st) return st;
      if (fallbacks_b)
        for (int k = 0; k < nrhs; ++k) mh.col_it[k] += mixed_it_b;
    }
    GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
    float msb = 0.f;
    GF_HIP_CHECK(ctx, hipEventElapsedTime(&msb, ctx->ev_start, ctx->ev_stop));
    if (!opts.device_ptrs) GF_HIP_CHECK(ctx, hipMemcpy(x, Xd, vb * nrhs, hipMemcpyDeviceToHost));
    gf_solve_stats S{};
    S.converged = 1;
    S.status = GF_OK;
    for (int k = 0; k < nrhs; ++k) {
      S.iterations = std::max(S.iterations, mh.col_it[k]);
      const double bb = mh.bnorm2[k];
      S.relres = std::max(S.relres, bb > 0 ? std::sqrt(mh.rr[k] / bb) : std::sqrt(mh.rr[k]));
      S.true_relres = std::max(S.true_relres, res[size_t(k)]);
      if (mh.col_status[k] != GF_OK) {
        S.status = mh.col_status[k];
        S.converged = 0;
      } else if (!(mh.rr[k] <= opts.rtol * opts.rtol * bb)) {
        if (S.status == GF_OK) S.status = GF_ERR_NOCONV;
        S.converged = 0;
      }
    }
    S.ms_device = msb;
    S.ms_host = now_ms() - h0;
    S.refinements = refinements_b;
    S.fallbacks = fallbacks_b;
    if (stats) *stats = S;
    if (S.status == GF_ERR_BREAKDOWN)
      return set_error(ctx, GF_ERR_BREAKDOWN, "gf_pcg: breakdown (p.Kp <= 0 or NaN)");
    if (S.status == GF_ERR_NOCONV)
      return set_error(ctx, GF_ERR_NOCONV, "gf_pcg: iteration limit reached before rtol");
    return GF_OK;
  }
  std::vector<PcgHostState> hs(static_cast<size_t>(nrhs));
  std::vector<int> refinements(static_cast<size_t>(nrhs), 0);
  int fallbacks = 0;
  for (int k = 0; k < nrhs; ++k) {
    GF_HIP_CHECK(ctx, hipMemcpyAsync(P->wb, Bd + k * n, vb, hipMemcpyDeviceToDevice, ctx->stream));
    if (opts.use_x0)
      GF_HIP_CHECK(ctx, hipMemcpyAsync(P->wx, Xd + k * n, vb, hipMemcpyDeviceToDevice, ctx->stream));
    PcgHostState h{};
    if (opts.precision == GF_PREC_MIXED) {
      st = run_mixed(ctx, m, opts, &h, &refinements[size_t(k)]);
      if (st) return st;
      const bool ok = h.status == GF_OK && h.rr <= opts.rtol * opts.rtol * h.bnorm2;
      const int mixed_it = h.it;
      if (!ok && opts.maxit - mixed_it > 0) {
        // FP32 rounding can leave a badly conditioned operator indefinite
        // (breakdown: re-solve in FP64 from zero) or unable to resolve the
        // remaining error (stall: continue in FP64 from the current x), with
        // the iterations the FP32 part left
        int warm = h.status == kMixedStalled ? 1 : 0;
        PcgBuffers Bj = make_buffers(m, m->d_dinv);
        if (warm) {
          // the FP32 iterate can carry large errors in the lowest modes (a
          // small residual, a large energy error): FP64 continues from it only
          // when its energy functional is below the zero start's
          GF_HIP_CHECK(ctx, launch_energy(ctx->stream, Bj, P->res + k));
          double energy = 0.0;
          GF_HIP_CHECK(ctx, hipMemcpyAsync(&energy, P->res + k, sizeof(double),
                                           hipMemcpyDeviceToHost, ctx->stream));
          GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
          if (!(energy < 0.0)) warm = 0;
        }
        st = run_pcg(ctx, m, Bj, sched == GF_SCHED_PERSISTENT ? sched : GF_SCHED_KERNELS, 1, warm,
                     opts.rtol, opts.maxit - mixed_it, opts.verify, &h);
        if (st) return st;
        h.it += mixed_it;
        ++fallbacks;
      } else if (!ok && h.status == kMixedStalled) {
        h.status = GF_ERR_NOCONV;
      }
    } else {
      st = run_pcg(ctx, m, B, sched, key, opts.use_x0, opts.rtol, opts.maxit, opts.verify, &h);
      if (st) return st;
    }
    GF_HIP_CHECK(ctx, launch_true_residual(ctx->stream, B, P->res + k));
    GF_HIP_CHECK(ctx, hipMemcpyAsync(Xd + k * n, P->wx, vb, hipMemcpyDeviceToDevice, ctx->stream));
    hs[size_t(k)] = h;
  }
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
  float ms = 0.f;
  GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  std::vector<double> res(size_t(nrhs), 0.0);
  GF_HIP_CHECK(ctx, hipMemcpy(res.data(), P->res, size_t(nrhs) * sizeof(double), hipMemcpyDeviceToHost));
  if (!opts.device_ptrs) GF_HIP_CHECK(ctx, hipMemcpy(x, Xd, vb * nrhs, hipMemcpyDeviceToHost));

  gf_solve_stats S{};
  S.converged = 1;
  S.status = GF_OK;
  for (int k = 0; k < nrhs; ++k) {
    const PcgHostState& h = hs[size_t(k)];
    S.iterations = std::max(S.iterations, h.it);
    S.refinements = std::max(S.refinements, refinements[size_t(k)]);
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
  S.fallbacks = fallbacks;
  for (int k = 0; k < nrhs; ++k) S.x0_rejected |= hs[size_t(k)].x0_rejected;
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
  if (m->plan->variant < 0 && (st = tune_spmv(ctx, m, -1, nullptr, nullptr))) return st;
  if ((st = prepare_variant(ctx, m, m->plan->variant))) return st;
  PcgBuffers B = make_buffers(m, nullptr);
  GF_HIP_CHECK(ctx, hipMemsetAsync(m->plan->p, 0, size_t(m->n) * sizeof(double), ctx->stream));
  GF_HIP_CHECK(ctx, launch_spmv_pq_repeat(ctx->stream, B, 1));
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  GF_HIP_CHECK(ctx, launch_spmv_pq_repeat(ctx->stream, B, nrep));
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
  float ms = 0.f;
  GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  *ms_avg = double(ms) / nrep;
  return GF_OK;
}

int gf_profile_barrier(gf_ctx* ctx, gf_ell* m, int32_t nbar, double* us_per_barrier,
                       int32_t* nblocks) {
  if (!ctx || !m || nbar < 1 || !us_per_barrier || !nblocks)
    return set_error(ctx, GF_ERR_ARG, "gf_profile_barrier: bad argument");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int st = ensure_plan(ctx, m);
  if (st) return st;
  gf_solve_plan* P = m->plan;
  if (P->variant < 0 && (st = tune_spmv(ctx, m, -1, nullptr, nullptr))) return st;
  const int nb = persistent_blocks(P->variant, P->num_cu);
  if (nb < 1)
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_profile_barrier: no persistent grid for this matrix");
  float ms[2] = {0.f, 0.f};
  const int counts[2] = {0, nbar};
  for (int k = 0; k < 2; ++k) {
    GF_HIP_CHECK(ctx, launch_barrier_probe(ctx->stream, P->barrier, nb, 1));  // warm
    GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    GF_HIP_CHECK(ctx, launch_barrier_probe(ctx->stream, P->barrier, nb, counts[k]));
    GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
    GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms[k], ctx->ev_start, ctx->ev_stop));
  }
  int aborted = 0;
  GF_HIP_CHECK(ctx, hipError_t(barrier_aborted(ctx->stream, P->barrier, &aborted)));
  if (aborted) return set_error(ctx, GF_ERR_HIP, "gf_profile_barrier: a barrier wait timed out");
  *us_per_barrier = 1e3 * double(ms[1] - ms[0]) / nbar;
  *nblocks = nb;
  return GF_OK;
}

int gf_profile_pcg(gf_ctx* ctx, gf_ell* m, const double* b, int32_t iters, double* ms_spmv,
                   double* ms_iteration) {
  if (!ctx || !m || !b || iters < 1 || !ms_spmv || !ms_iteration)
    return set_error(ctx, GF_ERR_ARG, "gf_profile_pcg: bad argument");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int st = ensure_plan(ctx, m);
  if (st) return st;
  gf_solve_plan* P = m->plan;
  if (P->variant < 0 && (st = tune_spmv(ctx, m, -1, nullptr, nullptr))) return st;
  if ((st = prepare_jacobi(ctx, m))) return st;
  if ((st = prepare_variant(ctx, m, P->variant))) return st;
  GF_HIP_CHECK(ctx, hipMemcpyAsync(P->wb, b, size_t(m->n) * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream));
  const PcgBuffers B = make_buffers(m, m->d_dinv);
  // rtol 0: every one of the iters iterations runs (unless the residual
  // vanishes exactly)
  GF_HIP_CHECK(ctx, launch_pcg_setup(ctx->stream, B, 0, 0.0, iters));
  struct Events {
    std::vector<hipEvent_t> ev;
    ~Events() {
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    }
  } E;
  E.ev.assign(2 * kChunk, nullptr);
  for (hipEvent_t& e : E.ev) GF_HIP_CHECK(ctx, hipEventCreate(&e));
  double spmv = 0.0;
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  for (int done = 0; done < iters;) {
    const int c = std::min(kChunk, iters - done);
    for (int k = 0; k < c; ++k)
      GF_HIP_CHECK(ctx, launch_pcg_iteration_timed(ctx->stream, B, (done + k) & 1, E.ev[2 * k],
                                                   E.ev[2 * k + 1]));
    GF_HIP_CHECK(ctx, hipEventSynchronize(E.ev[2 * c - 1]));
    for (int k = 0; k < c; ++k) {
      float ms = 0.f;
      GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, E.ev[2 * k], E.ev[2 * k + 1]));
      spmv += ms;
    }
    done += c;
  }
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
  float total = 0.f;
  GF_HIP_CHECK(ctx, hipEventElapsedTime(&total, ctx->ev_start, ctx->ev_stop));
  PcgHostState h;
  GF_HIP_CHECK(ctx, read_pcg_state(ctx->stream, B, &h));
  if (h.it != iters)
    return set_error(ctx, GF_ERR_BREAKDOWN, "gf_profile_pcg: the iteration stopped early");
  *ms_spmv = spmv / iters;
  *ms_iteration = double(total) / iters;
  return GF_OK;
}

}  // extern "C"
