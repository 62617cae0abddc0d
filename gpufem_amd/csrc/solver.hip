// ELL SpMV and the Jacobi-preconditioned CG, FP64.
//
// The CG loop runs without host round trips: every scalar (alpha, beta,
// rho, ||r||^2) lives in a device PcgState, and each kernel reduces the
// per-block partials of the previous one in a fixed order, so results are
// bitwise reproducible for a given problem size.  One iteration is three
// launches:
//   k_spmv_pq   q = K p, partial p.q
//   k_update    alpha = rho / p.q; x += alpha p; r -= alpha q; z = M^-1 r;
//               partials r.z, r.r
//   k_direction beta = r.z / rho; p = z + beta p; block 0 advances the state
// Kernels return immediately once the state says the solve is done, so the
// host can enqueue (or replay as a hipGraph) fixed chunks of iterations and
// check convergence between chunks.
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

namespace {

constexpr int kNT = 256;  // threads per block for the vector kernels

// Sum over the block; every thread gets the same (bitwise) result.
__device__ inline double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kNT / 64; ++w) t += sh[w];
  return t;
}

// Sum of np partials, identical in every block.
__device__ inline double reduce_partials(const double* part, int np, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += kNT) v += part[i];
  return block_sum(v, sh);
}

__device__ inline double ell_row(int64_t i, int64_t n, const double* __restrict__ val,
                                 const int32_t* __restrict__ col, int32_t rn,
                                 const double* __restrict__ x) {
  double acc = 0.0;
  for (int32_t s = 0; s < rn; ++s) {
    const int64_t idx = i + int64_t(s) * n;
    acc += __builtin_nontemporal_load(val + idx) * x[__builtin_nontemporal_load(col + idx)];
  }
  return acc;
}

__global__ __launch_bounds__(kNT) void k_spmv(int64_t n, const double* __restrict__ val,
                                              const int32_t* __restrict__ col,
                                              const int32_t* __restrict__ rownnz,
                                              const double* __restrict__ x,
                                              double* __restrict__ y, int nrhs) {
  for (int64_t i = int64_t(blockIdx.x) * kNT + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kNT) {
    const int32_t rn = rownnz[i];
    for (int k = 0; k < nrhs; ++k) y[int64_t(k) * n + i] = ell_row(i, n, val, col, rn, x + int64_t(k) * n);
  }
}

__global__ void k_jacobi_inv(int64_t n, const double* val, const int32_t* diag, double* dinv) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t s = diag[i];
    const double d = s >= 0 ? val[i + int64_t(s) * n] : 0.0;
    dinv[i] = 1.0 / d;
  }
}

// r = b - q (use_x0) or x = 0, r = b; z = M^-1 r; p = z; partials r.z, r.r, b.b
__global__ __launch_bounds__(kNT) void k_pcg_init(PcgBuffers B, int use_x0) {
  __shared__ double sh[kNT / 64];
  double rz = 0, rr = 0, bb = 0;
  for (int64_t i = int64_t(blockIdx.x) * kNT + threadIdx.x; i < B.n;
       i += int64_t(gridDim.x) * kNT) {
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
  }
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

__global__ __launch_bounds__(kNT) void k_spmv_pq(PcgBuffers B) {
  if (B.st->done) return;
  __shared__ double sh[kNT / 64];
  double pq = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kNT + threadIdx.x; i < B.n;
       i += int64_t(gridDim.x) * kNT) {
    const double qi = ell_row(i, B.n, B.val, B.col, B.rownnz[i], B.p);
    B.q[i] = qi;
    pq += B.p[i] * qi;
  }
  pq = block_sum(pq, sh);
  if (threadIdx.x == 0) B.part1[blockIdx.x] = pq;
}

__global__ __launch_bounds__(kNT) void k_update(PcgBuffers B, int parity) {
  PcgState* st = B.st;
  if (st->done) return;
  __shared__ double sh[kNT / 64];
  const double pq = reduce_partials(B.part1, gridDim.x, sh);
  if (!(pq > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->status = GF_ERR_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  const double alpha = st->rho[parity] / pq;
  double rz = 0, rr = 0;
  for (int64_t i = int64_t(blockIdx.x) * kNT + threadIdx.x; i < B.n;
       i += int64_t(gridDim.x) * kNT) {
    B.x[i] += alpha * B.p[i];
    const double ri = B.r[i] - alpha * B.q[i];
    const double zi = B.dinv ? B.dinv[i] * ri : ri;
    B.r[i] = ri;
    B.z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
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
  const double beta = rz / st->rho[parity];
  for (int64_t i = int64_t(blockIdx.x) * kNT + threadIdx.x; i < B.n;
       i += int64_t(gridDim.x) * kNT)
    B.p[i] = B.z[i] + beta * B.p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rho[parity ^ 1] = rz;
    st->rr = rr;
    const int it = st->it + 1;
    st->it = it;
    if (rr <= st->rtol2 * st->bnorm2 || it >= st->maxit) st->done = 1;
  }
}

// partial sums of (b - q)^2 then ||b - q|| / ||b||
__global__ __launch_bounds__(kNT) void k_resid_partial(PcgBuffers B) {
  __shared__ double sh[kNT / 64];
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kNT + threadIdx.x; i < B.n;
       i += int64_t(gridDim.x) * kNT) {
    const double d = B.b[i] - B.q[i];
    s += d * d;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) B.part1[blockIdx.x] = s;
}

__global__ __launch_bounds__(kNT) void k_resid_finish(PcgBuffers B, int nb, double* out) {
  __shared__ double sh[kNT / 64];
  const double s = reduce_partials(B.part1, nb, sh);
  if (threadIdx.x == 0) {
    const double bb = B.st->bnorm2;
    out[0] = bb > 0.0 ? sqrt(s / bb) : sqrt(s);
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

hipError_t launch_spmv(hipStream_t s, int64_t n, int32_t width, const double* val,
                       const int32_t* col, const int32_t* rownnz, const double* x,
                       double* y, int nrhs) {
  (void)width;
  hipLaunchKernelGGL(k_spmv, dim3(grid_for(n, kNT, 8192)), dim3(kNT), 0, s, n, val, col,
                     rownnz, x, y, nrhs);
  return hipGetLastError();
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
    e = launch_spmv(s, B.n, B.width, B.val, B.col, B.rownnz, B.x, B.q, 1);
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
  hipLaunchKernelGGL(k_spmv_pq, dim3(B.nblocks), dim3(kNT), 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_update, dim3(B.nblocks), dim3(kNT), 0, s, B, parity);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_direction, dim3(B.nblocks), dim3(kNT), 0, s, B, parity);
  return hipGetLastError();
}

__global__ void k_state_live(PcgState* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->done = 0;
    st->status = GF_OK;
  }
}

hipError_t launch_spmv_pq_repeat(hipStream_t s, const PcgBuffers& B, int nrep) {
  hipLaunchKernelGGL(k_state_live, dim3(1), dim3(64), 0, s, B.st);
  hipError_t e = hipGetLastError();
  for (int i = 0; i < nrep && e == hipSuccess; ++i) {
    hipLaunchKernelGGL(k_spmv_pq, dim3(B.nblocks), dim3(kNT), 0, s, B);
    e = hipGetLastError();
  }
  return e;
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

hipError_t launch_true_residual(hipStream_t s, const PcgBuffers& B, double* out_dev) {
  hipError_t e = launch_spmv(s, B.n, B.width, B.val, B.col, B.rownnz, B.x, B.q, 1);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_resid_partial, dim3(B.nblocks), dim3(kNT), 0, s, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_resid_finish, dim3(1), dim3(kNT), 0, s, B, B.nblocks, out_dev);
  return hipGetLastError();
}

}  // namespace gf
