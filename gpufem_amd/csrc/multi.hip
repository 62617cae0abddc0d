// Batched Jacobi-PCG over S load cases with one ELL SpMM per iteration.
//
// The S right-hand sides advance together, each with its own alpha, beta
// and stopping test (the columns of a block Krylov solve without the
// cross-column Gram terms), so every iteration reads the matrix once for
// all S columns instead of once per column.  Vectors are stored row-major,
// V[i*S + k], so the S entries a gathered column contributes are contiguous.
// In the vector kernels a workgroup is 256/S row groups of S lanes, lane k
// of a group handling load case k of its rows; the SpMM gives each lane
// four load cases of a row instead (fewer, wider loads per matrix entry).  A converged column keeps alpha = beta = 0 from
// then on, freezing its solution.
#include "gf_internal.hpp"
#include "launch.hpp"
#include "multi.hpp"
#include "multi_kernels.hpp"

namespace gf {

template <int S>
struct MultiState {
  double rho[S];
  double alpha[S];
  double beta[S];
  double bnorm2[S];
  double rr[S];
  int32_t col_it[S];    // iterations each column took
  int32_t col_done[S];  // 1: converged (or empty) column
  int32_t col_status[S];
  int32_t it, done, status, maxit;
  uint32_t arrived;  // workgroups of k_b_beta finished in this iteration
  double rtol2;
};

namespace {

using namespace mdev;

__device__ inline SpmmArgs spmm_args(const MultiBuffers& B) {
  return SpmmArgs{B.n, B.rownnz, B.col, B.bs};
}

template <int S>
__global__ __launch_bounds__(kNT) void k_b_init(MultiBuffers B) {
  __shared__ double sh[kNT];
  double rz = 0, rr = 0, bb = 0;
  for_group_rows<S>(B.n, [&](int64_t i, int k) {
    const int64_t o = i * S + k;
    const double bi = B.b[o];
    const double ri = B.use_x0 ? bi - B.q[o] : bi;
    if (!B.use_x0) B.x[o] = 0.0;
    const double zi = B.dinv ? B.dinv[i] * ri : ri;
    B.r[o] = ri;
    B.p[o] = zi;
    rz += ri * zi;
    rr += ri * ri;
    bb += bi * bi;
  });
  rz = column_sum<S>(rz, sh);
  rr = column_sum<S>(rr, sh);
  bb = column_sum<S>(bb, sh);
  if (threadIdx.x < S) {
    const int k = threadIdx.x, nb = gridDim.x;
    B.part2[int64_t(k) * nb + blockIdx.x] = rz;
    B.part2[int64_t(S + k) * nb + blockIdx.x] = rr;
    B.part2[int64_t(2 * S + k) * nb + blockIdx.x] = bb;
  }
}

template <int S>
__global__ __launch_bounds__(kNT) void k_b_init_finish(MultiBuffers B, int nb, int ncols,
                                                       double rtol2, int maxit) {
  MultiState<S>* st = static_cast<MultiState<S>*>(B.st);
  __shared__ double sh[kNT];
  __shared__ double rz[S], rr[S], bb[S];
  reduce_columns<S>(B.part2, nb, 0, sh, rz);
  reduce_columns<S>(B.part2, nb, S, sh, rr);
  reduce_columns<S>(B.part2, nb, 2 * S, sh, bb);
  if (threadIdx.x < S) {
    const int k = threadIdx.x;
    st->rho[k] = rz[k];
    st->alpha[k] = 0.0;
    st->beta[k] = 0.0;
    st->bnorm2[k] = bb[k];
    st->rr[k] = rr[k];
    st->col_it[k] = 0;
    st->col_status[k] = GF_OK;
    st->col_done[k] = (k >= ncols) || (rr[k] <= rtol2 * bb[k]) || maxit <= 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int j = 0; j < S; ++j) all &= st->col_done[j];
    st->it = 0;
    st->done = all;
    st->status = GF_OK;
    st->maxit = maxit;
    st->arrived = 0;
    st->rtol2 = rtol2;
  }
}

// q = K p and the per-column partials of p.q.  Launched at three workgroups
// per CU (padded LDS), i.e. three waves per SIMD, so the compiler may spend up
// to 168 registers keeping every load of a step in flight.
template <int S, int D, int NB = 2>
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_b_spmm_pq_bs(
    MultiBuffers B) {
  const MultiState<S>* st = static_cast<const MultiState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT * kCB];
  double pq[kCB] = {};
  constexpr bool share = Chunks<S, kCB>::L >= NB * (D * D / 2);
  for_chunk_node_rows<S, D, double, double, false, NB, true, kCB, share>(
      spmm_args(B), B.val, B.p, [&](int32_t R, int c0, const double (*acc)[kCB]) {
#pragma unroll
        for (int r = 0; r < D; ++r) {
          const int64_t o = (int64_t(D) * R + r) * S;
          double pv[kCB];
          load_split<S, kCB>(B.p + o, c0, pv);
          store_split<S, kCB>(B.q + o, c0, acc[r]);
#pragma unroll
          for (int j = 0; j < kCB; ++j) pq[j] += pv[j] * acc[r][j];
        }
      });
  chunk_partials<S, true, kCB>(pq, sh, B.part1);
}

template <int S, int D>
__global__ __launch_bounds__(kNT) void k_b_spmm_bs(MultiBuffers B, const double* __restrict__ xin,
                                                   double* __restrict__ yout) {
  for_chunk_node_rows<S, D, double, double, false, 2, true, kCB>(
      spmm_args(B), B.val, xin, [&](int32_t R, int c0, const double (*acc)[kCB]) {
#pragma unroll
        for (int r = 0; r < D; ++r)
          store_split<S, kCB>(yout + (int64_t(D) * R + r) * S, c0, acc[r]);
      });
}

template <int S>
__global__ __launch_bounds__(kNT) void k_b_spmm_pq(MultiBuffers B) {
  const MultiState<S>* st = static_cast<const MultiState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT * kC];
  double pq[kC] = {};
  for_chunk_rows<S>(B.n, [&](int64_t i, int c0) {
    double acc[kC], pv[kC];
    ell_row_chunk<S>(spmm_args(B), B.val, B.p, i, c0, acc);
    const int64_t o = i * S + c0;
    load_chunk(B.p + o, pv);
    store_chunk(B.q + o, acc);
#pragma unroll
    for (int j = 0; j < kC; ++j) pq[j] += pv[j] * acc[j];
  });
  chunk_partials<S>(pq, sh, B.part1);
}

// alpha_k = rho_k / p_k.q_k (0 for finished columns; a column with
// p.q <= 0 breaks down and stops).  One workgroup per column k.
template <int S>
__global__ __launch_bounds__(kNT) void k_b_alpha(MultiBuffers B, int nb) {
  MultiState<S>* st = static_cast<MultiState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT];
  const int k = blockIdx.x;
  const double pq = reduce_column(B.part1, nb, k, sh);
  if (threadIdx.x == 0) {
    double a = 0.0;
    if (!st->col_done[k]) {
      if (pq > 0.0) {
        a = st->rho[k] / pq;
      } else {
        st->col_status[k] = GF_ERR_BREAKDOWN;
        st->col_done[k] = 1;
      }
    }
    st->alpha[k] = a;
  }
}

template <int S>
__global__ __launch_bounds__(kNT) void k_b_update(MultiBuffers B) {
  MultiState<S>* st = static_cast<MultiState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT];
  __shared__ double s_alpha[S];
  if (threadIdx.x < S) s_alpha[threadIdx.x] = st->alpha[threadIdx.x];
  __syncthreads();
  const int nb = gridDim.x;
  double rz = 0, rr = 0;
  // x and p are updated together in k_b_direction; z = D^-1 r is not stored
  for_group_rows<S>(B.n, [&](int64_t i, int k) {
    const int64_t o = i * S + k;
    const double ri = B.r[o] - s_alpha[k] * B.q[o];
    const double zi = B.dinv ? B.dinv[i] * ri : ri;
    B.r[o] = ri;
    rz += ri * zi;
    rr += ri * ri;
  });
  rz = column_sum<S>(rz, sh);
  rr = column_sum<S>(rr, sh);
  if (threadIdx.x < S) {
    B.part2[int64_t(threadIdx.x) * nb + blockIdx.x] = rz;
    B.part2[int64_t(S + threadIdx.x) * nb + blockIdx.x] = rr;
  }
}

// beta_k, the per-column stopping test and the iteration counters.  One
// workgroup per column k; the last of them to finish (an arrival counter)
// sets the iteration count and, when every column is done, the done flag.
template <int S>
__global__ __launch_bounds__(kNT) void k_b_beta(MultiBuffers B, int nb) {
  MultiState<S>* st = static_cast<MultiState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT];
  const int k = blockIdx.x;
  const double rz = reduce_column(B.part2, nb, k, sh);
  const double rr = reduce_column(B.part2, nb, S + k, sh);
  if (threadIdx.x == 0) {
    const int it = st->it + 1;
    double beta = 0.0;
    if (!st->col_done[k]) {
      beta = rz / st->rho[k];
      st->rho[k] = rz;
      st->rr[k] = rr;
      st->col_it[k] = it;
      if (rr <= st->rtol2 * st->bnorm2[k] || it >= st->maxit) {
        st->col_done[k] = 1;
        beta = 0.0;
      }
    }
    st->beta[k] = beta;
    // publish this column's state before arriving; the last arrival reads
    // every column's
    __threadfence();
    if (atomicAdd(&st->arrived, 1u) == S - 1) {
      __threadfence();
      int all = 1;
      for (int j = 0; j < S; ++j)
        all &= __hip_atomic_load(&st->col_done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      st->arrived = 0;
      st->it = it;
      if (all) st->done = 1;
    }
  }
}

// x += alpha p with the p of this iteration, then p = D^-1 r + beta p.  The
// iteration that finishes the solve returns here before the x update (done is
// set); k_b_x_finish applies it once after the loop.
template <int S>
__global__ __launch_bounds__(kNT) void k_b_direction(MultiBuffers B) {
  MultiState<S>* st = static_cast<MultiState<S>*>(B.st);
  if (st->done) return;
  __shared__ double s_alpha[S], s_beta[S];
  if (threadIdx.x < S) {
    s_alpha[threadIdx.x] = st->alpha[threadIdx.x];
    s_beta[threadIdx.x] = st->beta[threadIdx.x];
  }
  __syncthreads();
  for_group_rows<S>(B.n, [&](int64_t i, int k) {
    const int64_t o = i * S + k;
    const double pk = B.p[o];
    B.x[o] += s_alpha[k] * pk;
    const double ri = B.r[o];
    const double zi = B.dinv ? B.dinv[i] * ri : ri;
    B.p[o] = zi + s_beta[k] * pk;
  });
}

// the x update of the last iteration (alpha = 0 for columns finished before
// it, and after the setup)
template <int S>
__global__ __launch_bounds__(kNT) void k_b_x_finish(MultiBuffers B) {
  const MultiState<S>* st = static_cast<const MultiState<S>*>(B.st);
  __shared__ double s_alpha[S];
  if (threadIdx.x < S) s_alpha[threadIdx.x] = st->alpha[threadIdx.x];
  __syncthreads();
  for_group_rows<S>(B.n, [&](int64_t i, int k) {
    const int64_t o = i * S + k;
    B.x[o] += s_alpha[k] * B.p[o];
  });
}

// transposes between column-after-column and row-major (i*S + k) layouts
template <int S>
__global__ void k_b_pack(int64_t n, int ncols, const double* __restrict__ cols,
                         double* __restrict__ rows) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n * S;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = t / S;
    const int k = int(t % S);
    rows[t] = k < ncols ? cols[int64_t(k) * n + i] : 0.0;
  }
}

template <int S>
__global__ void k_b_unpack(int64_t n, int ncols, const double* __restrict__ rows,
                           double* __restrict__ cols) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n * ncols;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int k = int(t / n);
    const int64_t i = t % n;
    cols[t] = rows[i * S + k];
  }
}

// per-column ||b - K x||^2 and ||b||^2 (q holds K x)
template <int S>
__global__ __launch_bounds__(kNT) void k_b_resid(MultiBuffers B) {
  __shared__ double sh[kNT];
  double s = 0, bb = 0;
  for_group_rows<S>(B.n, [&](int64_t i, int k) {
    const int64_t o = i * S + k;
    const double d = B.b[o] - B.q[o];
    s += d * d;
    bb += B.b[o] * B.b[o];
  });
  s = column_sum<S>(s, sh);
  bb = column_sum<S>(bb, sh);
  if (threadIdx.x < S) {
    B.part1[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] = s;
    B.part2[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] = bb;
  }
}

template <int S>
__global__ __launch_bounds__(kNT) void k_b_resid_finish(MultiBuffers B, int nb, double* out) {
  __shared__ double sh[kNT];
  __shared__ double ss[S], bb[S];
  reduce_columns<S>(B.part1, nb, 0, sh, ss);
  reduce_columns<S>(B.part2, nb, 0, sh, bb);
  if (threadIdx.x < S) {
    const int k = threadIdx.x;
    out[k] = bb[k] > 0.0 ? sqrt(ss[k] / bb[k]) : sqrt(ss[k]);
  }
}

// K x for the true residual (and the warm start): the same SpMM on x
template <int S>
__global__ __launch_bounds__(kNT) void k_b_spmm(MultiBuffers B, const double* __restrict__ xin,
                                                double* __restrict__ yout) {
  for_chunk_rows<S>(B.n, [&](int64_t i, int c0) {
    double acc[kC];
    ell_row_chunk<S>(spmm_args(B), B.val, xin, i, c0, acc);
    store_chunk(yout + i * S + c0, acc);
  });
}

inline int grid_for(int64_t work, int threads, int cap) {
  int64_t b = (work + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return int(b);
}

template <int S>
hipError_t launch_spmm_pq(hipStream_t s, const MultiBuffers& B) {
  const dim3 g(B.nblocks), t(kNT);
  // padded LDS: three workgroups per CU (four fit) keep each XCD's rows in
  // flight (and the 32-wide rows of p they gather) closer to what its L2 holds
  if (B.bsell_d == 2) hipLaunchKernelGGL((k_b_spmm_pq_bs<S, 2>), g, t, B.spmm_lds, s, B);
  else if (B.bsell_d == 3) hipLaunchKernelGGL((k_b_spmm_pq_bs<S, 3>), g, t, B.spmm_lds, s, B);
  else hipLaunchKernelGGL(k_b_spmm_pq<S>, g, t, 0, s, B);
  return hipGetLastError();
}

template <int S>
hipError_t launch_spmm(hipStream_t s, const MultiBuffers& B, const double* xin, double* yout) {
  const dim3 g(B.nblocks), t(kNT);
  if (B.bsell_d == 2) hipLaunchKernelGGL((k_b_spmm_bs<S, 2>), g, t, 0, s, B, xin, yout);
  else if (B.bsell_d == 3) hipLaunchKernelGGL((k_b_spmm_bs<S, 3>), g, t, 0, s, B, xin, yout);
  else hipLaunchKernelGGL(k_b_spmm<S>, g, t, 0, s, B, xin, yout);
  return hipGetLastError();
}

template <int S>
hipError_t setup_s(hipStream_t s, const MultiBuffers& B, double rtol, int maxit) {
  hipError_t e;
  if (B.use_x0) {
    if ((e = launch_spmm<S>(s, B, B.x, B.q)) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_b_init<S>, dim3(B.nblocks), dim3(kNT), 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_b_init_finish<S>, dim3(1), dim3(kNT), 0, s, B, B.nblocks, B.ncols,
                     rtol * rtol, maxit);
  return hipGetLastError();
}

template <int S>
hipError_t iteration_s(hipStream_t s, const MultiBuffers& B) {
  hipError_t e;
  const dim3 g(B.nblocks), t(kNT);
  if ((e = launch_spmm_pq<S>(s, B)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_b_alpha<S>, dim3(S), t, 0, s, B, B.nblocks);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_b_update<S>, g, t, 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_b_beta<S>, dim3(S), t, 0, s, B, B.nblocks);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_b_direction<S>, g, t, 0, s, B);
  return hipGetLastError();
}

template <int S>
hipError_t x_finish_s(hipStream_t s, const MultiBuffers& B) {
  hipLaunchKernelGGL(k_b_x_finish<S>, dim3(B.nblocks), dim3(kNT), 0, s, B);
  return hipGetLastError();
}

template <int S>
hipError_t residual_s(hipStream_t s, const MultiBuffers& B, double* out) {
  hipError_t e;
  if ((e = launch_spmm<S>(s, B, B.x, B.q)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_b_resid<S>, dim3(B.nblocks), dim3(kNT), 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_b_resid_finish<S>, dim3(1), dim3(kNT), 0, s, B, B.nblocks, out);
  return hipGetLastError();
}

template <int S>
hipError_t read_s(hipStream_t s, const MultiBuffers& B, MultiHostState* h) {
  MultiState<S> st;
  hipError_t e = hipMemcpyAsync(&st, B.st, sizeof(st), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  h->it = st.it;
  h->done = st.done;
  for (int k = 0; k < B.ncols; ++k) {
    h->col_it[k] = st.col_it[k];
    h->col_status[k] = st.col_status[k];
    h->rr[k] = st.rr[k];
    h->bnorm2[k] = st.bnorm2[k];
  }
  return hipSuccess;
}

}  // namespace

int multi_width(int ncols) {
  if (ncols <= 8) return 8;
  if (ncols <= 16) return 16;
  if (ncols <= 32) return 32;
  if (ncols <= 64) return 64;
  return 0;
}

size_t multi_state_bytes(int S) {
  switch (S) {
    case 8: return sizeof(MultiState<8>);
    case 16: return sizeof(MultiState<16>);
    case 32: return sizeof(MultiState<32>);
    default: return sizeof(MultiState<64>);
  }
}

int multi_spmm_lds(int lds_per_cu) {
  // k_b_spmm_pq_bs holds kNT * kCB doubles of LDS; 512 bytes of allocation
  // slack.  Three per CU measured 2 % faster per block iteration than two
  // (Brick20 48^3, 32 load cases); four made an earlier form of the kernel
  // 1.4x slower per launch than two (more rows of p in flight, L2 misses).
  const int pad = lds_per_cu / 3 - int(kNT * kCB * sizeof(double)) - 512;
  return pad > 0 ? pad : 0;
}

int multi_blocks(int64_t n, int S) {
  const int64_t rt = 8 * (kNT / S);
  int64_t nb = (n + rt - 1) / rt;
  if (nb > kMaxPcgBlocks) nb = kMaxPcgBlocks;
  if (nb >= 8) nb &= ~int64_t(7);
  return int(nb < 1 ? 1 : nb);
}

#define GF_MULTI_SWITCH(S, FN, ...)            \
  switch (S) {                                 \
    case 8: return FN<8>(__VA_ARGS__);         \
    case 16: return FN<16>(__VA_ARGS__);       \
    case 32: return FN<32>(__VA_ARGS__);       \
    case 64: return FN<64>(__VA_ARGS__);       \
    default: return hipErrorInvalidValue;      \
  }

hipError_t launch_multi_pack(hipStream_t s, int S, int64_t n, int ncols, const double* cols,
                             double* rows) {
  const dim3 g(grid_for(n * S, 256, 16384)), b(256);
  switch (S) {
    case 8: hipLaunchKernelGGL(k_b_pack<8>, g, b, 0, s, n, ncols, cols, rows); break;
    case 16: hipLaunchKernelGGL(k_b_pack<16>, g, b, 0, s, n, ncols, cols, rows); break;
    case 32: hipLaunchKernelGGL(k_b_pack<32>, g, b, 0, s, n, ncols, cols, rows); break;
    case 64: hipLaunchKernelGGL(k_b_pack<64>, g, b, 0, s, n, ncols, cols, rows); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_multi_unpack(hipStream_t s, int S, int64_t n, int ncols, const double* rows,
                               double* cols) {
  const dim3 g(grid_for(n * ncols, 256, 16384)), b(256);
  switch (S) {
    case 8: hipLaunchKernelGGL(k_b_unpack<8>, g, b, 0, s, n, ncols, rows, cols); break;
    case 16: hipLaunchKernelGGL(k_b_unpack<16>, g, b, 0, s, n, ncols, rows, cols); break;
    case 32: hipLaunchKernelGGL(k_b_unpack<32>, g, b, 0, s, n, ncols, rows, cols); break;
    case 64: hipLaunchKernelGGL(k_b_unpack<64>, g, b, 0, s, n, ncols, rows, cols); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_multi_setup(hipStream_t s, const MultiBuffers& B, double rtol, int maxit) {
  GF_MULTI_SWITCH(B.S, setup_s, s, B, rtol, maxit)
}

hipError_t launch_multi_iteration(hipStream_t s, const MultiBuffers& B) {
  GF_MULTI_SWITCH(B.S, iteration_s, s, B)
}

hipError_t launch_multi_x_finish(hipStream_t s, const MultiBuffers& B) {
  GF_MULTI_SWITCH(B.S, x_finish_s, s, B)
}

hipError_t launch_multi_residual(hipStream_t s, const MultiBuffers& B, double* out) {
  GF_MULTI_SWITCH(B.S, residual_s, s, B, out)
}

hipError_t read_multi_state(hipStream_t s, const MultiBuffers& B, MultiHostState* h) {
  GF_MULTI_SWITCH(B.S, read_s, s, B, h)
}

}  // namespace gf
