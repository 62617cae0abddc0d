// ELL SpMV and the Jacobi-preconditioned CG in FP64.
//
// Rows are visited in 256-row tiles in grid-stride order, with workgroup
// ids remapped so that each XCD (workgroups are dispatched round-robin over
// the 8 XCDs) owns a contiguous part of every window of tiles: the
// x-gathers of neighbouring rows then hit the same L2.
//
// SpMV variants (chosen per matrix by the tuner in solve.cpp):
//   kThreadU1/U4/U8  one thread per row, slots streamed 1/4/8 at a time
//                    (more loads in flight per thread); same summation
//                    order as the reference's Axy, so all three are bitwise
//                    identical.
//   kTile2/kTile4    2 or 4 waves share a 64-row tile, each taking every
//                    2nd/4th slot; partial sums combined through LDS.
//   kNode2/kNode3    thread per row over the d x d node blocks of a 2D / 3D
//                    FEM matrix: one node column per d slots and the d
//                    entries of x read together (fewer loads, 4/d instead of
//                    4 column bytes per entry); bitwise equal to kThreadU1.
//   kBsell2/kBsell3  thread per node row of the block-sliced copy, four
//                    slices per workgroup in grid-stride sweeps.
//   kBsell2G/3G      the same with one workgroup per four-slice group (one
//                    sweep; the dispatcher balances unequal slices).
//
// CG schedules: three launches per iteration (k_spmv_pq, k_update,
// k_direction) with all scalars in a device PcgState, so the host can
// enqueue or graph-replay chunks of iterations; or one cooperative
// persistent kernel (k_pcg_persistent) with three grid barriers per
// iteration.  Every block reduces the per-block partials in the same fixed
// order, so all blocks take the same convergence decision and results are
// bitwise reproducible for a given matrix and variant.  The SpMV kernel may
// run on a grid of its own (the slice-group variants): k_update reduces its
// spmv_grid p.q partials, and the persistent kernel plays the same logical
// SpMV workgroups in turn.  No workgroup waits on another inside a kernel
// but the persistent one: on gfx950 an agent-scope release writes back the
// XCD's L2, far too costly once per workgroup per SpMV.
#include "device_common.hpp"
#include "gf_internal.hpp"
#include "launch.hpp"

namespace gf {

struct PcgState {
  double rho[2];  // r.z, double-buffered by iteration parity
  double rr;      // r.r after the last update
  double bnorm2;  // b.b
  double rtol2;   // rtol^2
  double spare;
  int32_t it;
  int32_t done;
  int32_t status;
  int32_t maxit;
};

struct GridBarrier {
  unsigned long long count;  // arrivals since the launch began (monotonic)
  int abort;
  int pad;
};

namespace {

using namespace dev;

// ---------------------------------------------------------------------------
// standalone SpMV (multi-RHS by repeated sweeps)
// ---------------------------------------------------------------------------
template <int V>
__global__ __launch_bounds__(kNT) void k_spmv(int64_t n, const double* __restrict__ val,
                                              const int32_t* __restrict__ col,
                                              const int32_t* __restrict__ rownnz,
                                              const double* __restrict__ x,
                                              double* __restrict__ y, BsellView bs) {
  for_rows_spmv<V, double>(logical_block(), gridDim.x, n, val, col, rownnz, x,
                           [&](int64_t i, double v) { y[i] = v; }, bs);
}

// block-sliced copy: node columns (padding blocks point at the row's own
// node, whose x entries exist) and values (padding blocks are zero)
__global__ void k_bsell_cols(BsellView bs, int d, int64_t n, const int32_t* __restrict__ col,
                             const int32_t* __restrict__ rownnz, int32_t* __restrict__ bcol) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < bs.nslices * 64;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t sl = t >> 6;
    const int lane = int(t & 63);
    const int32_t R = bs.perm[t];
    const int64_t b0 = bs.sptr[sl], w = bs.sptr[sl + 1] - b0;
    const int32_t cnt = R >= 0 ? rownnz[int64_t(d) * R] / d : 0;
    for (int64_t j = 0; j < w; ++j)
      bcol[(b0 + j) * 64 + lane] =
          j < cnt ? col[int64_t(d) * R + j * d * n] / d : (R >= 0 ? R : 0);
  }
}

__global__ void k_bsell_offsets(BsellView bs, const int32_t* __restrict__ bcol,
                                int16_t* __restrict__ coff, int* overflow) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < bs.nslices * 64;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t sl = t >> 6;
    const int lane = int(t & 63);
    const int32_t R = bs.perm[t];
    const int32_t base = R >= 0 ? R : 0;
    const int64_t b0 = bs.sptr[sl], w = bs.sptr[sl + 1] - b0;
    bool bad = false;
    for (int64_t j = 0; j < w; ++j) {
      const int32_t d = bcol[(b0 + j) * 64 + lane] - base;
      bad |= d < -32768 || d > 32767;
      coff[(b0 + j) * 64 + lane] = int16_t(d);
    }
    if (bad) atomicOr(overflow, 1);
  }
}

__global__ void k_bsell_vals(BsellView bs, int d, int64_t n, const double* __restrict__ val,
                             const int32_t* __restrict__ rownnz, double* __restrict__ bval) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < bs.nslices * 64;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t sl = t >> 6;
    const int lane = int(t & 63);
    const int32_t R = bs.perm[t];
    const int64_t b0 = bs.sptr[sl], w = bs.sptr[sl + 1] - b0;
    const int32_t cnt = R >= 0 ? rownnz[int64_t(d) * R] / d : 0;
    for (int64_t j = 0; j < w; ++j)
      for (int r = 0; r < d; ++r)
        for (int e = 0; e < d; ++e)
          bval[bsell_vidx(d * d, b0 + j, r * d + e, lane)] =
              j < cnt ? val[int64_t(d) * R + r + (j * d + e) * n] : 0.0;
  }
}

// The inverse: the ELL values of the live slots from the block-sliced copy
// (padding slots of the ELL rows keep their zeros).
__global__ void k_bsell_to_ell(BsellView bs, int d, int64_t n, const double* __restrict__ bval,
                               const int32_t* __restrict__ rownnz, double* __restrict__ val) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < bs.nslices * 64;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t sl = t >> 6;
    const int lane = int(t & 63);
    const int32_t R = bs.perm[t];
    if (R < 0) continue;
    const int64_t b0 = bs.sptr[sl];
    const int32_t cnt = rownnz[int64_t(d) * R] / d;
    for (int64_t j = 0; j < cnt; ++j)
      for (int e = 0; e < d; ++e)
        for (int r = 0; r < d; ++r)
          val[int64_t(d) * R + r + (j * d + e) * n] = bval[bsell_vidx(d * d, b0 + j, r * d + e, lane)];
  }
}

__global__ void k_build_coln(int64_t n, int32_t nbw, int d, const int32_t* col, int32_t* coln) {
  const int64_t total = n * nbw;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t j = t / n, i = t - j * n;
    coln[t] = col[i + j * d * n] / d;
  }
}

// plain streaming copy, 16 bytes per thread per step (the bandwidth probe)
typedef double v2d __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kNT) void k_stream_copy(const v2d* __restrict__ src,
                                                     v2d* __restrict__ dst, int64_t n16) {
  for (int64_t i = int64_t(blockIdx.x) * kNT + threadIdx.x; i < n16; i += int64_t(gridDim.x) * kNT)
    dst[i] = __builtin_nontemporal_load(src + i);
}

__global__ void k_jacobi_inv(int64_t n, const double* val, const int32_t* diag, double* dinv) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t s = diag[i];
    const double d = s >= 0 ? val[i + int64_t(s) * n] : 0.0;
    dinv[i] = 1.0 / d;
  }
}

// ---------------------------------------------------------------------------
// PCG phases (shared by the multi-launch and the persistent schedules)
// ---------------------------------------------------------------------------
// q = K p over the rows of logical SpMV workgroup lb of nlb; returns this
// thread's part of p.q
template <int V>
__device__ inline double phase_spmv_pq(const PcgBuffers& B, int lb, int nlb) {
  double pq = 0.0;
  for_rows_spmv<V, double>(lb, nlb, B.n, B.val, B.col, B.rownnz, B.p, [&](int64_t i, double qi) {
    B.q[i] = qi;
    pq += B.p[i] * qi;
  }, B.bs);
  return pq;
}

__device__ inline void phase_update(const PcgBuffers& B, double alpha, double* rz_out,
                                    double* rr_out) {
  double rz = 0, rr = 0;
  for_my_rows(B.n, [&](int64_t i) {
    B.x[i] += alpha * B.p[i];
    const double ri = B.r[i] - alpha * B.q[i];
    const double zi = B.dinv ? B.dinv[i] * ri : ri;
    B.r[i] = ri;
    B.z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  });
  *rz_out = rz;
  *rr_out = rr;
}

__device__ inline void phase_direction(const PcgBuffers& B, double beta) {
  for_my_rows(B.n, [&](int64_t i) { B.p[i] = B.z[i] + beta * B.p[i]; });
}

// r = b - q (use_x0) or x = 0, r = b; z = M^-1 r; p = z; partials r.z, r.r, b.b
__global__ __launch_bounds__(kNT) void k_pcg_init(PcgBuffers B, int use_x0) {
  __shared__ double sh[kNT / 64];
  double rz = 0, rr = 0, bb = 0;
  for_my_rows(B.n, [&](int64_t i) {
    const double bi = B.b[i];
    double ri;
    if (use_x0) {
      ri = bi - B.q[i];
    } else {
      B.x[i] = 0.0;
      ri = bi;
    }
    const double zi = B.dinv ? B.dinv[i] * ri : ri;
    B.r[i] = ri;
    B.z[i] = zi;
    B.p[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
    bb += bi * bi;
  });
  rz = block_sum(rz, sh);
  rr = block_sum(rr, sh);
  bb = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    B.part2[blockIdx.x] = rz;
    B.part2[gridDim.x + blockIdx.x] = rr;
    B.part2[2 * gridDim.x + blockIdx.x] = bb;
  }
}

__global__ __launch_bounds__(kNT) void k_pcg_init_finish(PcgBuffers B, int nb, double rtol2,
                                                         int maxit) {
  __shared__ double sh[kNT / 64];
  const double rz = reduce_partials(B.part2, nb, sh);
  const double rr = reduce_partials(B.part2 + nb, nb, sh);
  const double bb = reduce_partials(B.part2 + 2 * nb, nb, sh);
  if (threadIdx.x == 0) {
    PcgState* st = B.st;
    st->rho[0] = rz;
    st->rho[1] = 0.0;
    st->rr = rr;
    st->bnorm2 = bb;
    st->rtol2 = rtol2;
    st->it = 0;
    st->status = GF_OK;
    st->maxit = maxit;
    st->done = (rr <= rtol2 * bb) || maxit <= 0;
  }
}

template <int V>
__global__ __launch_bounds__(kNT) void k_spmv_pq(PcgBuffers B) {
  if (B.st->done) return;
  __shared__ double sh[kNT / 64];
  const double pq = block_sum(phase_spmv_pq<V>(B, logical_block(), gridDim.x), sh);
  if (threadIdx.x == 0) B.part1[blockIdx.x] = pq;
}

__global__ __launch_bounds__(kNT) void k_update(PcgBuffers B, int parity) {
  PcgState* st = B.st;
  if (st->done) return;
  __shared__ double sh[kNT / 64];
  const double pq = reduce_partials(B.part1, B.spmv_grid, sh);
  if (!(pq > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->status = GF_ERR_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  double rz, rr;
  phase_update(B, st->rho[parity] / pq, &rz, &rr);
  rz = block_sum(rz, sh);
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    B.part2[blockIdx.x] = rz;
    B.part2[gridDim.x + blockIdx.x] = rr;
  }
}

__global__ __launch_bounds__(kNT) void k_direction(PcgBuffers B, int parity) {
  PcgState* st = B.st;
  if (st->done) return;
  __shared__ double sh[kNT / 64];
  const int nb = gridDim.x;
  const double rz = reduce_partials(B.part2, nb, sh);
  const double rr = reduce_partials(B.part2 + nb, nb, sh);
  phase_direction(B, rz / st->rho[parity]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rho[parity ^ 1] = rz;
    st->rr = rr;
    const int it = st->it + 1;
    st->it = it;
    if (rr <= st->rtol2 * st->bnorm2 || it >= st->maxit) st->done = 1;
  }
}

// ---------------------------------------------------------------------------
// bounded grid barrier for the cooperative kernel
// ---------------------------------------------------------------------------
constexpr long kSpinLimit = 1L << 22;  // several seconds of polling: a barrier never takes that long

// Sense-free barrier on a monotonic arrival counter: arrival number a (0-based)
// belongs to barrier a / nb and waits until the counter reaches the end of
// that barrier's range.  The poll is a relaxed load (no cache maintenance per
// poll: an acquire load would invalidate the XCD's L2 on every spin, under the
// feet of the waves still gathering through it); one acquire fence follows.
__device__ inline bool grid_barrier(GridBarrier* bar, int* s_flag) {
  // the workgroup barrier makes every wave's stores complete; one agent-scope
  // release (L2 write-back) then publishes them for the other XCDs
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned long long nb = gridDim.x;
    const unsigned long long arrived =
        __hip_atomic_fetch_add(&bar->count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long target = (arrived / nb + 1) * nb;
    long spins = 0;
    while (__hip_atomic_load(&bar->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (++spins > kSpinLimit) {
        __hip_atomic_store(&bar->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    *s_flag = __hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return *s_flag == 0;
}

// All iterations in one launch.  The grid never exceeds the co-resident
// block count (cooperative launch), every block runs the same number of
// barriers, and the convergence test uses values every block computed
// identically, so all blocks leave the loop together.
template <int V>
__global__ __launch_bounds__(kNT) void k_pcg_persistent(PcgBuffers B, GridBarrier* bar) {
  __shared__ double sh[kNT / 64];
  __shared__ int s_flag;
  PcgState* st = B.st;
  if (st->done) return;  // read before any block can change it (it only changes at exit)
  if (threadIdx.x == 0) s_flag = 0;
  __syncthreads();
  int it = st->it;
  double rho = st->rho[it & 1];
  const double rtol2 = st->rtol2, bnorm2 = st->bnorm2;
  const int maxit = st->maxit;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const int nlb = B.spmv_grid;
  for (;;) {
    // the SpMV kernel's logical workgroups j (same rows, same partial slots)
    for (int j = blockIdx.x; j < nlb; j += gridDim.x) {
      const double pq_part = block_sum(phase_spmv_pq<V>(B, xcd_block(j, nlb), nlb), sh);
      if (threadIdx.x == 0) B.part1[j] = pq_part;
    }
    if (!grid_barrier(bar, &s_flag)) break;
    const double pq = reduce_partials(B.part1, nlb, sh);
    if (!(pq > 0.0)) {
      if (lead) {
        st->status = GF_ERR_BREAKDOWN;
        st->done = 1;
      }
      break;
    }
    double rz, rr;
    phase_update(B, rho / pq, &rz, &rr);
    rz = block_sum(rz, sh);
    rr = block_sum(rr, sh);
    if (threadIdx.x == 0) {
      B.part2[blockIdx.x] = rz;
      B.part2[gridDim.x + blockIdx.x] = rr;
    }
    if (!grid_barrier(bar, &s_flag)) break;
    rz = reduce_partials(B.part2, gridDim.x, sh);
    rr = reduce_partials(B.part2 + gridDim.x, gridDim.x, sh);
    phase_direction(B, rz / rho);
    rho = rz;
    ++it;
    const bool finished = rr <= rtol2 * bnorm2 || it >= maxit;
    if (lead) {
      st->rho[it & 1] = rz;
      st->rr = rr;
      st->it = it;
      if (finished) st->done = 1;
    }
    if (finished) break;
    if (!grid_barrier(bar, &s_flag)) break;
  }
  if (lead && s_flag) {
    st->status = GF_ERR_HIP;
    st->done = 1;
  }
}

// nbar grid barriers and nothing else: the synchronisation cost of the
// persistent schedule (three per iteration) on its own grid
__global__ __launch_bounds__(kNT) void k_barrier_probe(GridBarrier* bar, int nbar) {
  __shared__ int s_flag;
  for (int k = 0; k < nbar; ++k)
    if (!grid_barrier(bar, &s_flag)) break;
}

__global__ void k_barrier_reset(GridBarrier* bar) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    bar->count = 0;
    bar->abort = 0;
  }
}

// ---------------------------------------------------------------------------
// residual and helpers
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void k_resid_partial(PcgBuffers B) {
  __shared__ double sh[kNT / 64];
  double s = 0.0, bb = 0.0;
  for_my_rows(B.n, [&](int64_t i) {
    const double bi = B.b[i];
    const double d = bi - B.q[i];
    s += d * d;
    bb += bi * bi;
  });
  s = block_sum(s, sh);
  bb = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    B.part1[blockIdx.x] = s;
    B.part2[blockIdx.x] = bb;
  }
}

// ||b - K x|| / ||b|| (or ||b - K x|| when b = 0) from the partials above
__global__ __launch_bounds__(kNT) void k_resid_finish(PcgBuffers B, int nb, double* out) {
  __shared__ double sh[kNT / 64];
  const double s = reduce_partials(B.part1, nb, sh);
  const double bb = reduce_partials(B.part2, nb, sh);
  if (threadIdx.x == 0) out[0] = bb > 0.0 ? sqrt(s / bb) : sqrt(s);
}

// The energy functional 1/2 x.Kx - b.x (q holds K x): CG's error measure up
// to a constant, 1/2 ||x - x*||_K^2 - 1/2 ||x*||_K^2, so x beats the zero
// start when it is negative.
__global__ __launch_bounds__(kNT) void k_energy_partial(PcgBuffers B) {
  __shared__ double sh[kNT / 64];
  double xq = 0.0, xb = 0.0;
  for_my_rows(B.n, [&](int64_t i) {
    const double xi = B.x[i];
    xq += xi * B.q[i];
    xb += xi * B.b[i];
  });
  xq = block_sum(xq, sh);
  xb = block_sum(xb, sh);
  if (threadIdx.x == 0) {
    B.part1[blockIdx.x] = xq;
    B.part2[blockIdx.x] = xb;
  }
}

__global__ __launch_bounds__(kNT) void k_energy_finish(PcgBuffers B, int nb, double* out) {
  __shared__ double sh[kNT / 64];
  const double xq = reduce_partials(B.part1, nb, sh);
  const double xb = reduce_partials(B.part2, nb, sh);
  if (threadIdx.x == 0) out[0] = 0.5 * xq - xb;
}

__global__ void k_state_live(PcgState* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->done = 0;
    st->status = GF_OK;
  }
}

inline int grid_for(int64_t work, int threads, int cap) {
  int64_t b = (work + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return int(b);
}

}  // namespace

size_t pcg_state_bytes() { return sizeof(PcgState); }
size_t grid_barrier_bytes() { return sizeof(GridBarrier); }

int spmv_blocks(int64_t n) {
  // enough rows per block to amortize the partition, at most kMaxPcgBlocks;
  // a multiple of 8 so the XCD remap applies
  int64_t nb = (n + 1023) / 1024;
  if (nb > kMaxPcgBlocks) nb = kMaxPcgBlocks;
  if (nb >= 8) nb &= ~int64_t(7);
  if (nb < 1) nb = 1;
  return int(nb);
}

hipError_t launch_build_coln(hipStream_t s, int64_t n, int32_t width, int d, const int32_t* col,
                             int32_t* coln) {
  if (d < 2 || width % d) return hipErrorInvalidValue;
  const int64_t total = n * (width / d);
  int64_t nb = (total + kNT - 1) / kNT;
  if (nb > 8192) nb = 8192;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_build_coln, dim3(unsigned(nb)), dim3(kNT), 0, s, n, width / d, d, col, coln);
  return hipGetLastError();
}

hipError_t launch_stream_copy(hipStream_t s, const void* src, void* dst, int64_t bytes) {
  if ((bytes & 15) || (reinterpret_cast<uintptr_t>(src) & 15) || (reinterpret_cast<uintptr_t>(dst) & 15))
    return hipErrorInvalidValue;
  const int64_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_stream_copy, dim3(grid_for(n16, kNT, 8192)), dim3(kNT), 0, s,
                     static_cast<const v2d*>(src), static_cast<v2d*>(dst), n16);
  return hipGetLastError();
}

hipError_t launch_bsell_cols(hipStream_t s, const BsellView& bs, int d, int64_t n,
                             const int32_t* col, const int32_t* rownnz, int32_t* bcol) {
  hipLaunchKernelGGL(k_bsell_cols, dim3(grid_for(bs.nslices * 64, kNT, 8192)), dim3(kNT), 0, s, bs,
                     d, n, col, rownnz, bcol);
  return hipGetLastError();
}

hipError_t launch_bsell_offsets(hipStream_t s, const BsellView& bs, const int32_t* bcol,
                                int16_t* coff, int* overflow) {
  hipLaunchKernelGGL(k_bsell_offsets, dim3(grid_for(bs.nslices * 64, kNT, 8192)), dim3(kNT), 0, s,
                     bs, bcol, coff, overflow);
  return hipGetLastError();
}

hipError_t launch_bsell_vals(hipStream_t s, const BsellView& bs, int d, int64_t n,
                             const double* val, const int32_t* rownnz, double* bval) {
  hipLaunchKernelGGL(k_bsell_vals, dim3(grid_for(bs.nslices * 64, kNT, 8192)), dim3(kNT), 0, s, bs,
                     d, n, val, rownnz, bval);
  return hipGetLastError();
}

hipError_t launch_bsell_to_ell(hipStream_t s, const BsellView& bs, int d, int64_t n,
                               const double* bval, const int32_t* rownnz, double* val) {
  hipLaunchKernelGGL(k_bsell_to_ell, dim3(grid_for(bs.nslices * 64, kNT, 8192)), dim3(kNT), 0, s,
                     bs, d, n, bval, rownnz, val);
  return hipGetLastError();
}

int spmv_grid(int variant, const BsellView& bs, int nblocks) {
  if (!is_bsell_group_variant(variant)) return nblocks;
  int64_t nb = (bs.nslices + kNT / 64 - 1) / (kNT / 64);
  nb = (nb + 7) & ~int64_t(7);  // the XCD remap needs a multiple of 8
  if (nb > kMaxSpmvBlocks) nb = kMaxSpmvBlocks;
  if (nb < 8) nb = 8;
  return int(nb);
}

hipError_t launch_spmv(hipStream_t s, int variant, int64_t n, const double* val,
                       const int32_t* col, const int32_t* rownnz, const double* x, double* y,
                       int nrhs, const BsellView* bs) {
  const BsellView view = bs ? *bs : BsellView{nullptr, nullptr, 0, nullptr};
  if (is_bsell_variant(variant) && !view.perm) return hipErrorInvalidValue;
  const dim3 g(spmv_grid(variant, view, spmv_blocks(n)));
  for (int k = 0; k < nrhs; ++k) {
    const double* xk = x + int64_t(k) * n;
    double* yk = y + int64_t(k) * n;
    GF_DISPATCH_VARIANT(variant, k_spmv, g, s, n, val, col, rownnz, xk, yk, view);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_jacobi_inv(hipStream_t s, int64_t n, const double* val,
                             const int32_t* diag, double* dinv) {
  hipLaunchKernelGGL(k_jacobi_inv, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, val,
                     diag, dinv);
  return hipGetLastError();
}

hipError_t launch_pcg_setup(hipStream_t s, const PcgBuffers& B, int use_x0, double rtol,
                            int maxit) {
  hipError_t e;
  if (use_x0) {
    e = launch_spmv(s, B.variant, B.n, B.val, B.col, B.rownnz, B.x, B.q, 1, &B.bs);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_pcg_init, dim3(B.nblocks), dim3(kNT), 0, s, B, use_x0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_pcg_init_finish, dim3(1), dim3(kNT), 0, s, B, B.nblocks, rtol * rtol,
                     maxit);
  return hipGetLastError();
}

hipError_t launch_pcg_iteration(hipStream_t s, const PcgBuffers& B, int parity) {
  hipError_t e;
  GF_DISPATCH_VARIANT(B.variant, k_spmv_pq, dim3(B.spmv_grid), s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_update, dim3(B.nblocks), dim3(kNT), 0, s, B, parity);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_direction, dim3(B.nblocks), dim3(kNT), 0, s, B, parity);
  return hipGetLastError();
}

hipError_t launch_pcg_iteration_timed(hipStream_t s, const PcgBuffers& B, int parity,
                                      hipEvent_t spmv_start, hipEvent_t spmv_stop) {
  hipError_t e = hipEventRecord(spmv_start, s);
  if (e != hipSuccess) return e;
  GF_DISPATCH_VARIANT(B.variant, k_spmv_pq, dim3(B.spmv_grid), s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipEventRecord(spmv_stop, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_update, dim3(B.nblocks), dim3(kNT), 0, s, B, parity);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_direction, dim3(B.nblocks), dim3(kNT), 0, s, B, parity);
  return hipGetLastError();
}

int persistent_blocks(int variant, int num_cu) {
  int per_cu = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (variant) {
    case kThreadU1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kThreadU1>, kNT, 0); break;
    case kThreadU4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kThreadU4>, kNT, 0); break;
    case kThreadU8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kThreadU8>, kNT, 0); break;
    case kTile2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kTile2>, kNT, 0); break;
    case kTile4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kTile4>, kNT, 0); break;
    case kNode2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kNode2>, kNT, 0); break;
    case kNode3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kNode3>, kNT, 0); break;
    case kBsell2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kBsell2>, kNT, 0); break;
    case kBsell3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kBsell3>, kNT, 0); break;
    case kBsell2G: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kBsell2G>, kNT, 0); break;
    case kBsell3G: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg_persistent<kBsell3G>, kNT, 0); break;
    default: break;
  }
  if (e != hipSuccess || per_cu < 1) return 0;
  int64_t nb = int64_t(per_cu) * num_cu;
  if (nb > kMaxPcgBlocks) nb = kMaxPcgBlocks;
  nb &= ~int64_t(7);
  return int(nb);
}

hipError_t launch_pcg_persistent(hipStream_t s, const PcgBuffers& B, void* barrier) {
  GridBarrier* bar = static_cast<GridBarrier*>(barrier);
  hipLaunchKernelGGL(k_barrier_reset, dim3(1), dim3(64), 0, s, bar);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  PcgBuffers Bc = B;
  void* args[] = {&Bc, &bar};
  const dim3 g(B.nblocks), blk(kNT);
  switch (B.variant) {
    case kThreadU1: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kThreadU1>), g, blk, args, 0, s);
    case kThreadU4: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kThreadU4>), g, blk, args, 0, s);
    case kThreadU8: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kThreadU8>), g, blk, args, 0, s);
    case kTile2: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kTile2>), g, blk, args, 0, s);
    case kTile4: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kTile4>), g, blk, args, 0, s);
    case kNode2: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kNode2>), g, blk, args, 0, s);
    case kNode3: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kNode3>), g, blk, args, 0, s);
    case kBsell2: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kBsell2>), g, blk, args, 0, s);
    case kBsell3: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kBsell3>), g, blk, args, 0, s);
    case kBsell2G: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kBsell2G>), g, blk, args, 0, s);
    case kBsell3G: return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<kBsell3G>), g, blk, args, 0, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_barrier_probe(hipStream_t s, void* barrier, int nblocks, int nbar) {
  GridBarrier* bar = static_cast<GridBarrier*>(barrier);
  hipLaunchKernelGGL(k_barrier_reset, dim3(1), dim3(64), 0, s, bar);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  void* args[] = {&bar, &nbar};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_barrier_probe), dim3(nblocks),
                                    dim3(kNT), args, 0, s);
}

int barrier_aborted(hipStream_t s, const void* barrier, int* aborted) {
  GridBarrier h;
  hipError_t e = hipMemcpyAsync(&h, barrier, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  *aborted = e == hipSuccess ? h.abort : 1;
  return int(e);
}

hipError_t read_pcg_state(hipStream_t s, const PcgBuffers& B, PcgHostState* h) {
  PcgState st;
  hipError_t e = hipMemcpyAsync(&st, B.st, sizeof(st), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  h->it = st.it;
  h->done = st.done;
  h->status = st.status;
  h->rr = st.rr;
  h->bnorm2 = st.bnorm2;
  h->rho = st.rho[st.it & 1];
  return hipSuccess;
}

hipError_t launch_spmv_pq_repeat(hipStream_t s, const PcgBuffers& B, int nrep) {
  hipLaunchKernelGGL(k_state_live, dim3(1), dim3(64), 0, s, B.st);
  hipError_t e = hipGetLastError();
  for (int i = 0; i < nrep && e == hipSuccess; ++i) {
    GF_DISPATCH_VARIANT(B.variant, k_spmv_pq, dim3(B.spmv_grid), s, B);
    e = hipGetLastError();
  }
  return e;
}

hipError_t launch_energy(hipStream_t s, const PcgBuffers& B, double* out_dev) {
  hipError_t e = launch_spmv(s, B.variant, B.n, B.val, B.col, B.rownnz, B.x, B.q, 1, &B.bs);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_energy_partial, dim3(B.nblocks), dim3(kNT), 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(kNT), 0, s, B, B.nblocks, out_dev);
  return hipGetLastError();
}

hipError_t launch_true_residual(hipStream_t s, const PcgBuffers& B, double* out_dev) {
  hipError_t e = launch_spmv(s, B.variant, B.n, B.val, B.col, B.rownnz, B.x, B.q, 1, &B.bs);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_resid_partial, dim3(B.nblocks), dim3(kNT), 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_resid_finish, dim3(1), dim3(kNT), 0, s, B, B.nblocks, out_dev);
  return hipGetLastError();
}

}  // namespace gf
