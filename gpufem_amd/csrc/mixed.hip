// Mixed-precision Jacobi-PCG: FP32 inner iterations inside FP64 refinement.
//
// The inner iteration runs on an FP32 copy of the ELL values with FP32
// SpMV accumulation and FP32 r, z, p, q vectors; the solution update
// x += alpha p and every dot product are carried in FP64.  Whenever the
// recursive FP32 residual has dropped by delta (or claims convergence), the
// residual is refreshed in FP64, r = b - K x with the FP64 matrix, and the
// iteration continues from it with the same search direction (a reliable
// update: the refinement keeps the Krylov information instead of restarting
// the inner solve).  Convergence is only ever declared on a refreshed, FP64
// residual.  One iteration is four launches:
//   k_m_spmv_pq   q = K32 p, partial p.q
//   k_m_update    alpha; x += alpha p; r -= alpha q; z = M^-1 r; partials
//   k_m_refresh   (only when due) r = b - K64 x; z = M^-1 r; partials
//   k_m_direction beta; p = z + beta p; workgroup 0 advances the state
// Every workgroup takes the refresh decision from the same fixed-order
// reduction, so they agree without extra synchronization.
#include "device_common.hpp"
#include "gf_internal.hpp"
#include "launch.hpp"
#include "mixed.hpp"

namespace gf {

struct MixedState {
  double rho[2];    // r.z by iteration parity
  double rref2[2];  // ||r||^2 at the last refresh, by parity
  double bnorm2, rtol2, delta2;
  double rr_rec;   // recursive ||r||^2 after the last update
  double rr_true;  // ||b - K x||^2 at the last refresh
  int32_t it, done, status, maxit, refinements;
  int32_t stall;  // consecutive refreshes that found the true residual less than halved
  double alpha_last;  // alpha of the current iteration (history)
};

namespace {

using namespace dev;

__global__ void k_to_f32(int64_t count, const double* __restrict__ in, float* __restrict__ out) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = float(in[i]);
}

// FP32 copy of the block-sliced values in difference form (FP64 pairs to
// FP32 quads): the off-diagonal blocks rounded, the diagonal slot (the first
// slot whose node column is the row's node) replaced by the block row sum
// S_R = sum_J K_RJ, summed in FP64 in slot order and rounded once.  One
// thread per slice position (node row).
__global__ void k_bsell_to_f32(int D, BsellView bs, const int32_t* __restrict__ bcol,
                               const double* __restrict__ in, float* __restrict__ out) {
  const int dd = D * D;
  for (int64_t pos = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; pos < bs.nslices * 64;
       pos += int64_t(gridDim.x) * blockDim.x) {
    const int l = int(pos & 63);
    const int32_t R = bs.perm[pos];
    const int64_t b0 = bs.sptr[pos >> 6], b1 = bs.sptr[(pos >> 6) + 1];
    double S[9] = {};
    int64_t bdiag = -1;
    for (int64_t b = b0; b < b1; ++b) {
      const bool diag = R >= 0 && bdiag < 0 && bcol[b * 64 + l] == R;
      if (diag) bdiag = b;
      for (int k = 0; k < dd; ++k) {
        const double v = in[bsell_vidx(dd, b, k, l)];
        S[k] += v;
        out[bsell_vidx32(dd, b, k, l)] = float(v);
      }
    }
    if (bdiag >= 0)
      for (int k = 0; k < dd; ++k) out[bsell_vidx32(dd, bdiag, k, l)] = float(S[k]);
  }
}

__global__ __launch_bounds__(kNT) void k_m_init(MixedBuffers B, int use_x0) {
  __shared__ double sh[kNT / 64];
  double rz = 0, rr = 0, bb = 0;
  for_my_rows(B.n, [&](int64_t i) {
    const double bi = B.b[i];
    double ri;
    if (use_x0) {
      ri = bi - B.q64[i];
    } else {
      B.x[i] = 0.0;
      ri = bi;
    }
    const float r32 = float(ri);
    const float z32 = B.dinv32[i] * r32;
    B.r32[i] = r32;
    B.z32[i] = z32;
    B.p32[i] = z32;
    rz += double(r32) * double(z32);
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

__global__ __launch_bounds__(kNT) void k_m_init_finish(MixedBuffers B, int nb, double rtol2,
                                                       double delta2, int maxit) {
  __shared__ double sh[kNT / 64];
  const double rz = reduce_partials(B.part2, nb, sh);
  const double rr = reduce_partials(B.part2 + nb, nb, sh);
  const double bb = reduce_partials(B.part2 + 2 * nb, nb, sh);
  if (threadIdx.x == 0) {
    MixedState* st = B.st;
    st->rho[0] = rz;
    st->rho[1] = 0.0;
    st->rref2[0] = rr;
    st->rref2[1] = rr;
    st->bnorm2 = bb;
    st->rtol2 = rtol2;
    st->delta2 = delta2;
    st->rr_rec = rr;
    st->rr_true = rr;
    st->it = 0;
    st->status = GF_OK;
    st->maxit = maxit;
    st->refinements = 0;
    st->stall = 0;
    st->done = (rr <= rtol2 * bb) || maxit <= 0;
  }
}

template <int V>
__global__ __launch_bounds__(kNT) void k_m_spmv_pq(MixedBuffers B) {
  if (B.st->done) return;
  __shared__ double sh[kNT / 64];
  double pq = 0.0;
  for_rows_spmv<V, float, true>(logical_block(), gridDim.x, B.n, B.val32, B.col, B.rownnz, B.p32,
                          [&](int64_t i, float qi) {
                            B.q32[i] = qi;
                            pq += double(B.p32[i]) * double(qi);
                          }, B.bs);
  pq = block_sum(pq, sh);
  if (threadIdx.x == 0) B.part1[blockIdx.x] = pq;
}

__global__ __launch_bounds__(kNT) void k_m_update(MixedBuffers B, int parity) {
  MixedState* st = B.st;
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
  const double alpha = st->rho[parity] / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha_last = alpha;
  double rz = 0, rr = 0;
  for_my_rows(B.n, [&](int64_t i) {
    const double pi = double(B.p32[i]);
    B.x[i] += alpha * pi;
    const float r32 = float(double(B.r32[i]) - alpha * double(B.q32[i]));
    const float z32 = B.dinv32[i] * r32;
    B.r32[i] = r32;
    B.z32[i] = z32;
    rz += double(r32) * double(z32);
    rr += double(r32) * double(r32);
  });
  rz = block_sum(rz, sh);
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    B.part2[blockIdx.x] = rz;
    B.part2[gridDim.x + blockIdx.x] = rr;
  }
}

// identical in every workgroup: refresh when the recursive residual fell by
// delta since the last refresh or claims convergence
__device__ inline bool refresh_due(const MixedBuffers& B, int parity, double* rz, double* rr,
                                   double* sh) {
  const int nb = gridDim.x;
  *rz = reduce_partials(B.part2, nb, sh);
  *rr = reduce_partials(B.part2 + nb, nb, sh);
  const MixedState* st = B.st;
  return *rr <= st->delta2 * st->rref2[parity] || *rr <= st->rtol2 * st->bnorm2;
}

template <int V>
__global__ __launch_bounds__(kNT) void k_m_refresh(MixedBuffers B, int parity) {
  if (B.st->done) return;
  __shared__ double sh[kNT / 64];
  double rz, rr;
  if (!refresh_due(B, parity, &rz, &rr, sh)) return;
  double rz2 = 0.0, rr2 = 0.0;
  for_rows_spmv<V, double>(logical_block(), gridDim.x, B.n, B.val, B.col, B.rownnz, B.x,
                           [&](int64_t i, double kx) {
    const double ri = B.b[i] - kx;
    const float r32 = float(ri);
    const float z32 = B.dinv32[i] * r32;
    B.r32[i] = r32;
    B.z32[i] = z32;
    rz2 += double(r32) * double(z32);
    rr2 += ri * ri;
  }, B.bs);
  rz2 = block_sum(rz2, sh);
  rr2 = block_sum(rr2, sh);
  if (threadIdx.x == 0) {
    B.part3[blockIdx.x] = rz2;
    B.part3[gridDim.x + blockIdx.x] = rr2;
  }
}

__global__ __launch_bounds__(kNT) void k_m_direction(MixedBuffers B, int parity) {
  MixedState* st = B.st;
  if (st->done) return;
  __shared__ double sh[kNT / 64];
  double rz, rr;
  const bool rel = refresh_due(B, parity, &rz, &rr, sh);
  double rr_true = st->rr_true;
  bool converged = false;
  if (rel) {
    rz = reduce_partials(B.part3, gridDim.x, sh);
    rr_true = reduce_partials(B.part3 + gridDim.x, gridDim.x, sh);
    converged = rr_true <= st->rtol2 * st->bnorm2;
  }
  if (!converged) {
    const double beta = rz / st->rho[parity];
    for_my_rows(B.n, [&](int64_t i) {
      B.p32[i] = float(double(B.z32[i]) + beta * double(B.p32[i]));
    });
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // a refresh that finds the true residual less than halved since the
    // previous one means the FP32 operator no longer resolves the error
    // that is left; after kStallRefreshes of them in a row the FP32
    // iteration stops (status kMixedStalled) and the host continues in FP64
    int stall = st->stall;
    if (rel && !converged) stall = rr_true > 0.25 * st->rref2[parity] ? stall + 1 : 0;
    st->stall = stall;
    st->rho[parity ^ 1] = rz;
    st->rref2[parity ^ 1] = rel ? rr_true : st->rref2[parity];
    st->rr_rec = rel ? rr_true : rr;
    st->rr_true = rr_true;
    st->refinements += rel ? 1 : 0;
    const int it = st->it + 1;
    if (B.hist && it <= kLanczosIters) {
      B.hist[2 * (it - 1)] = st->alpha_last;
      B.hist[2 * (it - 1) + 1] = converged ? 0.0 : rz / st->rho[parity];
    }
    st->it = it;
    if (stall >= kStallRefreshes && !converged) st->status = kMixedStalled;
    if (converged || it >= st->maxit || stall >= kStallRefreshes) st->done = 1;
  }
}

inline int grid_for(int64_t work, int threads, int cap) {
  int64_t b = (work + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return int(b);
}

}  // namespace

size_t mixed_state_bytes() { return sizeof(MixedState); }

hipError_t launch_to_f32(hipStream_t s, int64_t count, const double* in, float* out) {
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(count, 256, 16384)), dim3(256), 0, s, count, in,
                     out);
  return hipGetLastError();
}

hipError_t launch_bsell_to_f32(hipStream_t s, int d, const BsellView& bs, const int32_t* bcol,
                               const double* in, float* out) {
  hipLaunchKernelGGL(k_bsell_to_f32, dim3(grid_for(bs.nslices * 64, 256, 16384)), dim3(256), 0, s,
                     d, bs, bcol, in, out);
  return hipGetLastError();
}

hipError_t launch_mixed_setup(hipStream_t s, const MixedBuffers& B, int use_x0, double rtol,
                              double delta, int maxit) {
  hipError_t e;
  if (use_x0) {
    e = launch_spmv(s, B.variant, B.n, B.val, B.col, B.rownnz, B.x, B.q64, 1, &B.bs);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_m_init, dim3(B.nblocks), dim3(kNT), 0, s, B, use_x0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_m_init_finish, dim3(1), dim3(kNT), 0, s, B, B.nblocks, rtol * rtol,
                     delta * delta, maxit);
  return hipGetLastError();
}

hipError_t launch_mixed_iteration(hipStream_t s, const MixedBuffers& B, int parity) {
  hipError_t e;
  const dim3 g(B.nblocks);
  GF_DISPATCH_VARIANT(B.variant, k_m_spmv_pq, dim3(B.spmv_grid), s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_m_update, g, dim3(kNT), 0, s, B, parity);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  GF_DISPATCH_VARIANT(B.variant, k_m_refresh, g, s, B, parity);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_m_direction, g, dim3(kNT), 0, s, B, parity);
  return hipGetLastError();
}

hipError_t read_mixed_state(hipStream_t s, const MixedBuffers& B, MixedHostState* h) {
  MixedState st;
  hipError_t e = hipMemcpyAsync(&st, B.st, sizeof(st), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  h->it = st.it;
  h->done = st.done;
  h->status = st.status;
  h->refinements = st.refinements;
  h->rr_true = st.rr_true;
  h->rr_rec = st.rr_rec;
  h->bnorm2 = st.bnorm2;
  return hipSuccess;
}

}  // namespace gf
