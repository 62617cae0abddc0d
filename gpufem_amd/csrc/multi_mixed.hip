// Batched mixed-precision Jacobi-PCG: S load cases advance together, each
// with FP32 inner iterations inside FP64 refinement (the single-RHS scheme
// of mixed.hip, column by column).
//
// The inner iteration runs one FP32 SpMM per iteration (FP32 values, FP32
// accumulation, FP32 r, z, p, q of S columns); x += alpha p and every dot
// product are FP64.  When a column's recursive residual has dropped by
// delta since its last refresh (or claims convergence), its residual is
// refreshed in FP64, r = b - K x with the FP64 matrix, and its iteration
// continues with the same search direction (reliable update).  A column is
// declared converged only on a refreshed FP64 residual.  One iteration is
// seven launches:
//   k_mm_spmm_pq   q = K32 p, partial p.q per column
//   k_mm_alpha     (one workgroup) alpha per column
//   k_mm_update    x += alpha p; r -= alpha q; z = M^-1 r; partials
//   k_mm_decide    (one workgroup) which columns refresh
//   k_mm_refresh   (only when some column is due) FP64 b - K x for them
//   k_mm_beta      (one workgroup) convergence on refreshed residuals, beta
//   k_mm_direction p = z + beta p
#include "gf_internal.hpp"
#include "launch.hpp"
#include "mixed.hpp"
#include "multi_mixed.hpp"
#include "multi_kernels.hpp"

namespace gf {

template <int S>
struct MMState {
  double rho[S], alpha[S], beta[S], bnorm2[S];
  double rref2[S];   // ||r||^2 at the column's last refresh
  double rr_rec[S];  // recursive ||r||^2 after the last update
  double rr_true[S]; // ||b - K x||^2 at the last refresh
  double rz_new[S];  // r.z after the last update
  int32_t col_it[S], col_done[S], col_status[S], col_refresh[S], col_refinements[S];
  int32_t col_stall[S];  // consecutive refreshes that less than halved the true residual
  int32_t it, done, status, maxit, any_refresh;
  int32_t arrived;  // workgroups of k_mm_beta finished in this iteration
  double rtol2, delta2;
};

namespace {

using namespace mdev;

__device__ inline SpmmArgs spmm_args(const MixedMultiBuffers& B) {
  return SpmmArgs{B.n, B.rownnz, B.col, B.bs};
}

// FP64 residual refresh for the columns that are due (setup: all columns,
// also sets p = z and the b.b partials).  The FP64 SpMM of x covers every
// column; only the refreshed ones are written.
template <int S, int D>
__global__ __launch_bounds__(kNT) void k_mm_refresh(MixedMultiBuffers B, int setup) {
  const MMState<S>* st = static_cast<const MMState<S>*>(B.st);
  if (!setup && (st->done || !st->any_refresh)) return;
  __shared__ double sh[kNT * kC];
  double rz[kC] = {}, rr[kC] = {}, bb[kC] = {};
  auto row = [&](int64_t i, int c0, const double* kx) {
    const int64_t o = i * S + c0;
    double bv[kC];
    load_chunk(B.b + o, bv);
    float r32[kC], z32[kC];
    bool any = false;
#pragma unroll
    for (int j = 0; j < kC; ++j) {
      const bool w = setup || st->col_refresh[c0 + j];
      const double ri = bv[j] - kx[j];
      r32[j] = float(ri);
      z32[j] = B.dinv32 ? B.dinv32[i] * r32[j] : r32[j];
      if (w) {
        rz[j] += double(r32[j]) * double(z32[j]);
        rr[j] += ri * ri;
        any = true;
      }
      if (setup) bb[j] += bv[j] * bv[j];
    }
    if (!any) return;
    if (setup) {
      store_chunk(B.r + o, r32);
      store_chunk(B.p + o, z32);
    } else {
#pragma unroll
      for (int j = 0; j < kC; ++j)
        if (st->col_refresh[c0 + j]) B.r[o + j] = r32[j];
    }
  };
  if constexpr (D == 0) {
    for_chunk_rows<S>(B.n, [&](int64_t i, int c0) {
      double kx[kC];
      ell_row_chunk<S>(spmm_args(B), B.val, B.x, i, c0, kx);
      row(i, c0, kx);
    });
  } else {
    for_chunk_node_rows<S, D>(spmm_args(B), B.val, B.x,
                              [&](int32_t R, int c0, const double (*acc)[kC]) {
#pragma unroll
                                for (int r = 0; r < D; ++r) row(int64_t(D) * R + r, c0, acc[r]);
                              });
  }
  double* out = setup ? B.part2 : B.part3;
  const int nb = gridDim.x;
  chunk_partials<S>(rz, sh, out);
  __syncthreads();
  chunk_partials<S>(rr, sh, out + int64_t(S) * nb);
  if (setup) {
    __syncthreads();
    chunk_partials<S>(bb, sh, out + int64_t(2 * S) * nb);
  }
}

template <int S>
__global__ __launch_bounds__(kNT) void k_mm_setup_finish(MixedMultiBuffers B, int nb, int ncols,
                                                         double rtol2, double delta2, int maxit) {
  MMState<S>* st = static_cast<MMState<S>*>(B.st);
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
    st->rref2[k] = rr[k];
    st->rr_rec[k] = rr[k];
    st->rr_true[k] = rr[k];
    st->rz_new[k] = rz[k];
    st->col_it[k] = 0;
    st->col_status[k] = GF_OK;
    st->col_refresh[k] = 0;
    st->col_refinements[k] = 0;
    st->col_stall[k] = 0;
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
    st->any_refresh = 0;
    st->arrived = 0;
    st->rtol2 = rtol2;
    st->delta2 = delta2;
  }
}

template <int S, int D>
__global__ __launch_bounds__(kNT) void k_mm_spmm_pq(MixedMultiBuffers B) {
  const MMState<S>* st = static_cast<const MMState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT * kC];
  double pq[kC] = {};
  auto row = [&](int64_t i, int c0, const float* acc) {
    const int64_t o = i * S + c0;
    float pv[kC];
    load_chunk(B.p + o, pv);
    store_chunk(B.q + o, acc);
#pragma unroll
    for (int j = 0; j < kC; ++j) pq[j] += double(pv[j]) * double(acc[j]);
  };
  if constexpr (D == 0) {
    for_chunk_rows<S>(B.n, [&](int64_t i, int c0) {
      float acc[kC];
      ell_row_chunk<S>(spmm_args(B), B.val32, B.p, i, c0, acc);
      row(i, c0, acc);
    });
  } else {
    constexpr bool share = Chunks<S>::L >= 2 * (D * D / 4);
    for_chunk_node_rows<S, D, float, float, true, 2, false, kC, share>(
        spmm_args(B), B.val32, B.p, [&](int32_t R, int c0, const float (*acc)[kC]) {
#pragma unroll
          for (int r = 0; r < D; ++r) row(int64_t(D) * R + r, c0, acc[r]);
        });
  }
  chunk_partials<S>(pq, sh, B.part1);
}

// alpha_k = rho_k / p_k.q_k (0 for finished columns).  p.q <= 0 in any
// column (an FP32 operator that is not positive definite) ends the whole
// solve: the host then redoes every column in FP64 (so the columns whose
// workgroups then see the solve done and skip their alpha lose nothing).
// One workgroup per column k.
template <int S>
__global__ __launch_bounds__(kNT) void k_mm_alpha(MixedMultiBuffers B, int nb) {
  MMState<S>* st = static_cast<MMState<S>*>(B.st);
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
        st->done = 1;
      }
    }
    st->alpha[k] = a;
  }
}

template <int S>
__global__ __launch_bounds__(kNT) void k_mm_update(MixedMultiBuffers B) {
  const MMState<S>* st = static_cast<const MMState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT];
  __shared__ double s_alpha[S];
  if (threadIdx.x < S) s_alpha[threadIdx.x] = st->alpha[threadIdx.x];
  __syncthreads();
  const int nb = gridDim.x;
  double rz = 0, rr = 0;
  for_group_rows<S>(B.n, [&](int64_t i, int k) {
    const int64_t o = i * S + k;
    const double a = s_alpha[k];
    B.x[o] += a * double(B.p[o]);
    const float r32 = float(double(B.r[o]) - a * double(B.q[o]));
    const float z32 = B.dinv32 ? B.dinv32[i] * r32 : r32;
    B.r[o] = r32;
    rz += double(r32) * double(z32);
    rr += double(r32) * double(r32);
  });
  rz = column_sum<S>(rz, sh);
  rr = column_sum<S>(rr, sh);
  if (threadIdx.x < S) {
    B.part2[int64_t(threadIdx.x) * nb + blockIdx.x] = rz;
    B.part2[int64_t(S + threadIdx.x) * nb + blockIdx.x] = rr;
  }
}

// which columns refresh: the recursive residual fell by delta since the
// column's last refresh, or claims convergence.  One workgroup per column k;
// any_refresh (0 since the last beta) is set by every column that is due.
template <int S>
__global__ __launch_bounds__(kNT) void k_mm_decide(MixedMultiBuffers B, int nb) {
  MMState<S>* st = static_cast<MMState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT];
  const int k = blockIdx.x;
  const double rz = reduce_column(B.part2, nb, k, sh);
  const double rr = reduce_column(B.part2, nb, S + k, sh);
  if (threadIdx.x == 0) {
    st->rz_new[k] = rz;
    st->rr_rec[k] = rr;
    const int due = !st->col_done[k] &&
                    (rr <= st->delta2 * st->rref2[k] || rr <= st->rtol2 * st->bnorm2[k]);
    st->col_refresh[k] = due;
    if (due) st->any_refresh = 1;
  }
}

// beta, the per-column stopping test (on refreshed FP64 residuals only) and
// the iteration counters.  One workgroup per column k; the last of them to
// finish (an arrival counter) sets the iteration count, clears any_refresh
// and, when every column is done, sets the done flag.
template <int S>
__global__ __launch_bounds__(kNT) void k_mm_beta(MixedMultiBuffers B, int nb) {
  MMState<S>* st = static_cast<MMState<S>*>(B.st);
  if (st->done) return;
  __shared__ double sh[kNT];
  const int k = blockIdx.x;
  const int refreshed = st->any_refresh;
  double rz2 = 0.0, rr2 = 0.0;
  if (refreshed) {  // the same value in every thread of every workgroup
    rz2 = reduce_column(B.part3, nb, k, sh);
    rr2 = reduce_column(B.part3, nb, S + k, sh);
  }
  if (threadIdx.x == 0) {
    const int it = st->it + 1;
    double beta = 0.0;
    if (!st->col_done[k]) {
      double rz = st->rz_new[k];
      bool stalled = false;
      if (st->col_refresh[k]) {
        rz = rz2;
        // as the single-RHS solve: kStallRefreshes refreshes in a row that
        // less than halve the true residual stop this column's FP32 iteration
        const bool met = rr2 <= st->rtol2 * st->bnorm2[k];
        st->col_stall[k] = !met && rr2 > 0.25 * st->rref2[k] ? st->col_stall[k] + 1 : 0;
        stalled = st->col_stall[k] >= kStallRefreshes;
        st->rr_true[k] = rr2;
        st->rref2[k] = rr2;
        st->col_refinements[k] += 1;
      }
      st->col_it[k] = it;
      if (B.hist && it <= kLanczosIters) {
        B.hist[2 * (int64_t(S) * (it - 1) + k)] = st->alpha[k];
        B.hist[2 * (int64_t(S) * (it - 1) + k) + 1] = rz / st->rho[k];
      }
      if (st->col_refresh[k] && st->rr_true[k] <= st->rtol2 * st->bnorm2[k]) {
        st->col_done[k] = 1;
      } else if (stalled) {
        st->col_done[k] = 1;
        st->col_status[k] = kMixedStalled;
      } else if (it >= st->maxit) {
        st->col_done[k] = 1;
        st->col_status[k] = GF_ERR_NOCONV;
      } else {
        beta = rz / st->rho[k];
        st->rho[k] = rz;
      }
    }
    st->beta[k] = beta;
    st->col_refresh[k] = 0;
    // publish this column's state before arriving; the last arrival reads
    // every column's
    __threadfence();
    if (atomicAdd(&st->arrived, 1) == S - 1) {
      __threadfence();
      int all = 1;
      for (int j = 0; j < S; ++j)
        all &= __hip_atomic_load(&st->col_done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      st->arrived = 0;
      st->it = it;
      st->any_refresh = 0;
      if (all) st->done = 1;
    }
  }
}

template <int S>
__global__ __launch_bounds__(kNT) void k_mm_direction(MixedMultiBuffers B) {
  const MMState<S>* st = static_cast<const MMState<S>*>(B.st);
  if (st->done) return;
  __shared__ double s_beta[S];
  if (threadIdx.x < S) s_beta[threadIdx.x] = st->beta[threadIdx.x];
  __syncthreads();
  // z = D^-1 r formed here (the FP32 product the update and refresh use)
  for_group_rows<S>(B.n, [&](int64_t i, int k) {
    const int64_t o = i * S + k;
    const float r32 = B.r[o];
    const float z32 = B.dinv32 ? B.dinv32[i] * r32 : r32;
    B.p[o] = float(double(z32) + s_beta[k] * double(B.p[o]));
  });
}

template <int S>
hipError_t refresh_s(hipStream_t s, const MixedMultiBuffers& B, int setup) {
  const dim3 g(B.nblocks), t(kNT);
  if (B.bsell_d == 2) hipLaunchKernelGGL((k_mm_refresh<S, 2>), g, t, 0, s, B, setup);
  else if (B.bsell_d == 3) hipLaunchKernelGGL((k_mm_refresh<S, 3>), g, t, 0, s, B, setup);
  else hipLaunchKernelGGL((k_mm_refresh<S, 0>), g, t, 0, s, B, setup);
  return hipGetLastError();
}

template <int S>
hipError_t setup_s(hipStream_t s, const MixedMultiBuffers& B, double rtol, double delta,
                   int maxit) {
  hipError_t e = refresh_s<S>(s, B, 1);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mm_setup_finish<S>, dim3(1), dim3(kNT), 0, s, B, B.nblocks, B.ncols,
                     rtol * rtol, delta * delta, maxit);
  return hipGetLastError();
}

template <int S>
hipError_t iteration_s(hipStream_t s, const MixedMultiBuffers& B) {
  hipError_t e;
  const dim3 g(B.nblocks), t(kNT);
  if (B.bsell_d == 2) hipLaunchKernelGGL((k_mm_spmm_pq<S, 2>), g, t, 0, s, B);
  else if (B.bsell_d == 3) hipLaunchKernelGGL((k_mm_spmm_pq<S, 3>), g, t, 0, s, B);
  else hipLaunchKernelGGL((k_mm_spmm_pq<S, 0>), g, t, 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_mm_alpha<S>, dim3(S), t, 0, s, B, B.nblocks);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_mm_update<S>, g, t, 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_mm_decide<S>, dim3(S), t, 0, s, B, B.nblocks);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = refresh_s<S>(s, B, 0)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_mm_beta<S>, dim3(S), t, 0, s, B, B.nblocks);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_mm_direction<S>, g, t, 0, s, B);
  return hipGetLastError();
}

template <int S>
hipError_t read_s(hipStream_t s, const MixedMultiBuffers& B, MixedMultiHostState* h) {
  MMState<S> st;
  hipError_t e = hipMemcpyAsync(&st, B.st, sizeof(st), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  h->it = st.it;
  h->done = st.done;
  for (int k = 0; k < B.ncols; ++k) {
    h->col_it[k] = st.col_it[k];
    h->col_status[k] = st.col_status[k];
    h->col_refinements[k] = st.col_refinements[k];
    h->rr_true[k] = st.rr_true[k];
    h->rr_rec[k] = st.rr_rec[k];
    h->bnorm2[k] = st.bnorm2[k];
  }
  return hipSuccess;
}

}  // namespace

size_t mixed_multi_state_bytes(int S) {
  switch (S) {
    case 8: return sizeof(MMState<8>);
    case 16: return sizeof(MMState<16>);
    case 32: return sizeof(MMState<32>);
    default: return sizeof(MMState<64>);
  }
}

#define GF_MM_SWITCH(S, FN, ...)               \
  switch (S) {                                 \
    case 8: return FN<8>(__VA_ARGS__);         \
    case 16: return FN<16>(__VA_ARGS__);       \
    case 32: return FN<32>(__VA_ARGS__);       \
    case 64: return FN<64>(__VA_ARGS__);       \
    default: return hipErrorInvalidValue;      \
  }

hipError_t launch_mixed_multi_setup(hipStream_t s, const MixedMultiBuffers& B, double rtol,
                                    double delta, int maxit) {
  GF_MM_SWITCH(B.S, setup_s, s, B, rtol, delta, maxit)
}

hipError_t launch_mixed_multi_iteration(hipStream_t s, const MixedMultiBuffers& B) {
  GF_MM_SWITCH(B.S, iteration_s, s, B)
}

hipError_t read_mixed_multi_state(hipStream_t s, const MixedMultiBuffers& B,
                                  MixedMultiHostState* h) {
  GF_MM_SWITCH(B.S, read_s, s, B, h)
}

}  // namespace gf
