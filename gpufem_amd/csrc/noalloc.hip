// gf_pcg_noalloc: the FP64 Jacobi-PCG of gf_pcg for one right-hand side
// without any device allocation, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Every buffer the solve's kernels write besides b and x -- the Jacobi
// inverse, r, z, p, q, the per-workgroup partial sums, the true-residual
// result and the scalar state -- lives in `scratch`, one file-scope device
// array of kNoAllocBytes per GPU (held until the process ends).  Host code
// never passes the array's name to a kernel or a copy: its device address is
// taken once per device by noalloc_setup (on the entry point's first call on
// that device, with the device current) and kept in g_scratch[device].
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU race.  Each call writes every scratch value it reads first (the setup
// kernel initialises the state, x when there is no initial guess, r, z, p).
#include <cmath>
#include <mutex>
#include <string>

#include "gf_internal.hpp"
#include "launch.hpp"
#include "solve_plan.hpp"

namespace gf {
namespace {

constexpr int64_t kRows = GF_NOALLOC_MAX_ROWS;
constexpr size_t kAlign = 256;
constexpr size_t round_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
// layout: dinv, r, z, p, q (kRows doubles each), part1, part2, res, state
constexpr size_t kVecBytes = round_up(size_t(kRows) * sizeof(double));
constexpr size_t kPart1Off = 5 * kVecBytes;
constexpr size_t kPart2Off = kPart1Off + round_up(size_t(kMaxSpmvBlocks) * sizeof(double));
constexpr size_t kResOff = kPart2Off + round_up(size_t(3 * kMaxPcgBlocks) * sizeof(double));
constexpr size_t kStateOff = kResOff + kAlign;
constexpr size_t kStateBytes = 4096;
constexpr size_t kNoAllocBytes = kStateOff + kStateBytes;
static_assert(kNoAllocBytes == GF_NOALLOC_SCRATCH_BYTES,
              "GF_NOALLOC_SCRATCH_BYTES (gpufem.h) must match the scratch layout");

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
char* g_scratch[kMaxDevices] = {};

constexpr int kChunk = 32;       // iterations between convergence checks, as gf_pcg
constexpr int kMaxRestarts = 4;  // true-residual restarts, as gf_pcg

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of this device's scratch array (the current device must be
// `device`), taken once per device; hipSuccess when it is known.
static hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device]) return hipSuccess;
  if (pcg_state_bytes() > kStateBytes) return hipErrorInvalidValue;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = static_cast<char*>(ptr);
  return hipSuccess;
}

}  // namespace gf

using namespace gf;

extern "C" int gf_pcg_noalloc(gf_ctx* ctx, gf_ell* m, const double* b, double* x,
                              const gf_solve_opts* opts_in, gf_solve_stats* stats) {
  if (!ctx || !m || !b || !x) return set_error(ctx, GF_ERR_ARG, "gf_pcg_noalloc: null argument");
  gf_solve_opts opts;
  if (opts_in) opts = *opts_in;
  else gf_solve_opts_default(&opts);
  if (!(opts.rtol >= 0.0) || opts.maxit < 0)
    return set_error(ctx, GF_ERR_ARG, "gf_pcg_noalloc: rtol must be >= 0 and maxit >= 0");
  if (opts.precond != GF_PRECOND_NONE && opts.precond != GF_PRECOND_JACOBI)
    return set_error(ctx, GF_ERR_ARG, "gf_pcg_noalloc: unknown preconditioner");
  if (opts.precision != GF_PREC_FP64)
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_pcg_noalloc: FP64 only");
  if (opts.schedule != GF_SCHED_AUTO && opts.schedule != GF_SCHED_KERNELS)
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_pcg_noalloc: kernel launches only");
  if (m->n > kRows)
    return set_error(ctx, GF_ERR_UNSUPPORTED,
                     "gf_pcg_noalloc: " + std::to_string(m->n) + " rows exceed the scratch array's " +
                         std::to_string(kRows) + " (GF_NOALLOC_MAX_ROWS)");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  {
    const hipError_t e = noalloc_setup(ctx->device);
    if (e != hipSuccess)
      return set_error(ctx, GF_ERR_HIP, std::string("gf_pcg_noalloc: address of the scratch array: ") +
                                            hipGetErrorString(e));
  }
  char* base = g_scratch[ctx->device];
  const int64_t n = m->n;

  // the matrix's tuned SpMV variant when its operands are current (the
  // block-sliced or node-column copy is not built here), else one thread
  // per row over the ELL arrays; the grid is the one gf_pcg uses
  int variant = kThreadU4;
  if (m->plan && m->plan->variant >= 0) {
    const int v = m->plan->variant;
    const bool current = is_bsell_variant(v)  ? (m->d_bval && m->bval_version == m->val_version)
                         : is_node_variant(v) ? m->d_coln != nullptr
                                              : true;
    if (current) variant = v;
  }
  // the ELL values of every slot for the ELL and node-column kernels (a
  // kernel launch, no allocation)
  if (!is_bsell_variant(variant)) {
    const int est = ensure_ell_vals(ctx, m);
    if (est) return est;
  }
  const SpmvOperands o = spmv_operands(m, variant);
  PcgBuffers B;
  B.n = n;
  B.width = m->width;
  B.variant = variant;
  B.val = o.val;
  B.col = o.col;
  B.bs = o.bs;
  B.rownnz = m->d_rownnz;
  double* dinv = reinterpret_cast<double*>(base);
  B.dinv = opts.precond == GF_PRECOND_JACOBI ? dinv : nullptr;
  B.b = b;
  B.x = x;
  B.r = reinterpret_cast<double*>(base + kVecBytes);
  B.z = reinterpret_cast<double*>(base + 2 * kVecBytes);
  B.p = reinterpret_cast<double*>(base + 3 * kVecBytes);
  B.q = reinterpret_cast<double*>(base + 4 * kVecBytes);
  B.part1 = reinterpret_cast<double*>(base + kPart1Off);
  B.part2 = reinterpret_cast<double*>(base + kPart2Off);
  B.nblocks = pcg_grid(variant, n, ctx->num_cu);
  B.spmv_grid = spmv_grid(variant, o.bs, B.nblocks);
  B.st = reinterpret_cast<PcgState*>(base + kStateOff);
  double* res = reinterpret_cast<double*>(base + kResOff);

  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  if (B.dinv) GF_HIP_CHECK(ctx, launch_jacobi_inv(ctx->stream, n, m->d_val, m->d_diag, dinv));
  PcgHostState h{};
  int total = 0, use_x0 = opts.use_x0;
  for (int restart = 0;; ++restart) {
    GF_HIP_CHECK(ctx, launch_pcg_setup(ctx->stream, B, use_x0, opts.rtol, opts.maxit - total));
    GF_HIP_CHECK(ctx, read_pcg_state(ctx->stream, B, &h));
    while (!h.done) {
      for (int it = 0; it < kChunk; ++it)
        GF_HIP_CHECK(ctx, launch_pcg_iteration(ctx->stream, B, it & 1));
      GF_HIP_CHECK(ctx, read_pcg_state(ctx->stream, B, &h));
    }
    total += h.it;
    if (!opts.verify || restart >= kMaxRestarts || h.status != GF_OK || total >= opts.maxit ||
        !(h.rr <= opts.rtol * opts.rtol * h.bnorm2))
      break;
    double tr = 0.0;
    GF_HIP_CHECK(ctx, launch_true_residual(ctx->stream, B, res));
    GF_HIP_CHECK(ctx, hipMemcpyAsync(&tr, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (tr <= opts.rtol) break;
    use_x0 = 1;
  }
  GF_HIP_CHECK(ctx, launch_true_residual(ctx->stream, B, res));
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  double tr = 0.0;
  GF_HIP_CHECK(ctx, hipMemcpyAsync(&tr, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  gf_solve_stats S{};
  S.iterations = total;
  S.status = h.status;
  S.relres = h.bnorm2 > 0 ? std::sqrt(h.rr / h.bnorm2) : std::sqrt(h.rr);
  S.true_relres = tr;
  S.ms_device = ms;
  S.ms_host = ms;
  S.converged = h.status == GF_OK && h.rr <= opts.rtol * opts.rtol * h.bnorm2;
  if (S.status == GF_OK && !S.converged) S.status = GF_ERR_NOCONV;
  if (stats) *stats = S;
  if (S.status == GF_ERR_BREAKDOWN)
    return set_error(ctx, GF_ERR_BREAKDOWN, "gf_pcg_noalloc: breakdown (p.Kp <= 0 or NaN)");
  if (S.status == GF_ERR_NOCONV)
    return set_error(ctx, GF_ERR_NOCONV, "gf_pcg_noalloc: iteration limit reached before rtol");
  return GF_OK;
}
