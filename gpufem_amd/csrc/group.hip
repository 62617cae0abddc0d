// Grouped batch: many independent meshes solved as one block-diagonal
// system with a segmented Jacobi-PCG.
//
// Rows of mesh m are [off[m], off[m+1]); they are cut into tiles of at most
// 256 rows that never cross a mesh boundary, so every per-tile partial sum
// belongs to one mesh and the per-mesh reductions add a contiguous run of
// tile partials in order (deterministic).  Each mesh has its own alpha,
// beta, stopping test and iteration count; the tiles of a converged mesh
// are skipped.  Workgroups take tiles in grid-stride order with the same
// XCD remap as the single-system kernels.  One iteration is five launches:
// SpMV + p.q partials, per-mesh alpha (a workgroup per mesh), update +
// partials, per-mesh beta and stopping test (a workgroup per mesh), direction.
#include <cfloat>

#include "device_common.hpp"
#include "gf_internal.hpp"
#include "group.hpp"
#include "launch.hpp"

namespace gf {

namespace {

using namespace dev;

// f(t, m) for every tile t owned by this workgroup (m: its mesh); the rows
// of tile t are [tile_r0[t], tile_r1[t]), at most kNT of them
template <class F>
__device__ inline void tiles_loop(const GroupBuffers& G, F&& f) {
  for (int64_t t = logical_block(); t < G.ntiles; t += gridDim.x) f(t, G.tile_mesh[t]);
}

__device__ inline double tile_sum(double v, double* sh) { return block_sum(v, sh); }

// v[a] + ... + v[b - 1] over one workgroup in a fixed order (thread t adds
// a + t, a + t + kNT, ... in turn, then the workgroup sum); in every thread
__device__ inline double run_sum(const double* __restrict__ v, int64_t a, int64_t b, double* sh) {
  double t = 0.0;
  for (int64_t s = a + threadIdx.x; s < b; s += kNT) t += v[s];
  return block_sum(t, sh);
}

// ---- penalty: pen[m] = coef * max diagonal of mesh m; K_dd *= pen[m] -------
__global__ __launch_bounds__(kNT) void k_g_diag_max(GroupBuffers G, double coef) {
  __shared__ double sh[kNT / 64];
  const int m = blockIdx.x;
  double mx = -DBL_MAX;
  for (int64_t i = G.off[m] + threadIdx.x; i < G.off[m + 1]; i += kNT) {
    const int32_t s = G.diag[i];
    mx = fmax(mx, s >= 0 ? G.val[i + int64_t(s) * G.n] : 0.0);
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) G.pen[m] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3])) * coef;
}

__global__ void k_g_penalty_apply(GroupBuffers G, const int64_t* dofs, const int32_t* mesh,
                                  const int32_t* counts, int64_t nd) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nd;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t d = dofs[t];
    const int32_t s = G.diag[d];
    if (s < 0) continue;
    double* v = G.val + d + int64_t(s) * G.n;
    double x = *v;
    const double p = G.pen[mesh[t]];
    for (int c = 0; c < counts[t]; ++c) x = p * x;
    *v = x;
  }
}

// ---- segmented PCG ------------------------------------------------------------
__global__ __launch_bounds__(kNT) void k_g_init(GroupBuffers G, int use_x0) {
  __shared__ double sh[kNT / 64];
  // the arrival counter of k_g_init_finish, which follows
  if (blockIdx.x == 0 && threadIdx.x == 0) G.st->arrived = 0;
  tiles_loop(G, [&](int64_t t, int m) {
    const int64_t i = G.tile_r0[t] + threadIdx.x;
    double rz = 0, rr = 0, bb = 0;
    if (i < G.tile_r1[t]) {
      const double bi = G.b[i];
      const double ri = use_x0 ? bi - G.q[i] : bi;
      if (!use_x0) G.x[i] = 0.0;
      const double zi = G.dinv[i] * ri;
      G.r[i] = ri;
      G.p[i] = zi;
      rz = ri * zi;
      rr = ri * ri;
      bb = bi * bi;
    }
    rz = tile_sum(rz, sh);
    rr = tile_sum(rr, sh);
    bb = tile_sum(bb, sh);
    if (threadIdx.x == 0) {
      G.part[t] = rz;
      G.part[G.ntiles + t] = rr;
      G.part[2 * G.ntiles + t] = bb;
    }
    (void)m;
  });
}

// one workgroup per mesh m; the last to finish (an arrival counter) sets the
// state of the whole group
__global__ __launch_bounds__(kNT) void k_g_init_finish(GroupBuffers G, double rtol2, int maxit) {
  GroupState* st = G.st;
  __shared__ double sh[kNT / 64];
  const int m = blockIdx.x;
  const int64_t t0 = G.tile_ptr[m], t1 = G.tile_ptr[m + 1];
  const double rz = run_sum(G.part, t0, t1, sh);
  const double rr = run_sum(G.part + G.ntiles, t0, t1, sh);
  const double bb = run_sum(G.part + 2 * G.ntiles, t0, t1, sh);
  if (threadIdx.x == 0) {
    G.rho[m] = rz;
    G.rr[m] = rr;
    G.bb[m] = bb;
    G.alpha[m] = 0.0;
    G.beta[m] = 0.0;
    G.it[m] = 0;
    G.status[m] = GF_OK;
    G.done[m] = (rr <= rtol2 * bb) || maxit <= 0;
    __threadfence();
    if (atomicAdd(&st->arrived, 1) == G.nmesh - 1) {
      __threadfence();
      int all = 1;
      for (int j = 0; j < G.nmesh; ++j)
        all &= __hip_atomic_load(&G.done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      st->arrived = 0;
      st->it = 0;
      st->done = all;
      st->maxit = maxit;
      st->rtol2 = rtol2;
    }
  }
}

template <int V>
__global__ __launch_bounds__(kNT) void k_g_spmv_pq(GroupBuffers G) {
  if (G.st->done) return;
  __shared__ double sh[kNT / 64];
  constexpr int U = V == kThreadU1 ? 1 : (V == kThreadU4 ? 4 : 8);
  constexpr int D = V == kNode2 ? 2 : 3;
  tiles_loop(G, [&](int64_t t, int m) {
    if (G.done[m]) {  // uniform across the workgroup
      if (threadIdx.x == 0) G.part[t] = 0.0;
      return;
    }
    const int64_t i = G.tile_r0[t] + threadIdx.x;
    double pq = 0.0;
    if (i < G.tile_r1[t]) {
      double qi;
      if constexpr (V == kNode2 || V == kNode3)
        qi = row_dot_node<D, double>(i, G.n, G.val, G.scol, G.rownnz[i], G.p);
      else
        qi = row_dot<U, double>(i, G.n, G.val, G.scol, G.rownnz[i], G.p);
      G.q[i] = qi;
      pq = G.p[i] * qi;
    }
    pq = tile_sum(pq, sh);
    if (threadIdx.x == 0) G.part[t] = pq;
  });
}

// q = K p over the block-sliced copy: workgroup gi takes slice group gi (at
// most four slices of one mesh, one per wave); partial p.q per group
template <int D>
__global__ __launch_bounds__(kNT) void k_g_spmv_pq_bs(GroupBuffers G) {
  if (G.st->done) return;
  __shared__ double sh[kNT / 64];
  const int lane = threadIdx.x & 63;
  for (int64_t gi = blockIdx.x; gi < G.nsg; gi += gridDim.x) {
    const int m = G.sg_mesh[gi];
    if (G.done[m]) {  // uniform across the workgroup
      if (threadIdx.x == 0) G.part_sg[gi] = 0.0;
      continue;
    }
    const int64_t sl = G.sg_s0[gi] + (threadIdx.x >> 6);
    double pq = 0.0;
    if (sl < G.sg_s1[gi]) {
      const int64_t b0 = G.bs.sptr[sl];
      const int32_t w = int32_t(G.bs.sptr[sl + 1] - b0);
      const int32_t R = G.bs.perm[sl * 64 + lane];
      double acc[D];
      if (G.bs.coff)
        bsell_row<D, double, true>(G.bs, G.bval, G.bcol, b0, w, lane, G.p, acc, R >= 0 ? R : 0);
      else
        bsell_row<D, double, false>(G.bs, G.bval, G.bcol, b0, w, lane, G.p, acc, 0);
      if (R >= 0) {
#pragma unroll
        for (int r = 0; r < D; ++r) {
          const int64_t i = int64_t(D) * R + r;
          G.q[i] = acc[r];
          pq += G.p[i] * acc[r];
        }
      }
    }
    pq = block_sum(pq, sh);
    if (threadIdx.x == 0) G.part_sg[gi] = pq;
  }
}

// one workgroup per mesh m
__global__ __launch_bounds__(kNT) void k_g_alpha(GroupBuffers G) {
  if (G.st->done) return;
  __shared__ double sh[kNT / 64];
  const int m = blockIdx.x;
  if (G.done[m]) {
    if (threadIdx.x == 0) G.alpha[m] = 0.0;
    return;
  }
  const double pq = G.bsell_d ? run_sum(G.part_sg, G.sg_ptr[m], G.sg_ptr[m + 1], sh)
                              : run_sum(G.part, G.tile_ptr[m], G.tile_ptr[m + 1], sh);
  if (threadIdx.x == 0) {
    double a = 0.0;
    if (pq > 0.0) {
      a = G.rho[m] / pq;
    } else {
      G.status[m] = GF_ERR_BREAKDOWN;
      G.done[m] = 1;
    }
    G.alpha[m] = a;
  }
}

__global__ __launch_bounds__(kNT) void k_g_update(GroupBuffers G) {
  if (G.st->done) return;
  __shared__ double sh[kNT / 64];
  tiles_loop(G, [&](int64_t t, int m) {
    if (G.done[m]) {
      if (threadIdx.x == 0) {
        G.part[t] = 0.0;
        G.part[G.ntiles + t] = 0.0;
      }
      return;
    }
    const int64_t i = G.tile_r0[t] + threadIdx.x;
    const double a = G.alpha[m];
    double rz = 0, rr = 0;
    if (i < G.tile_r1[t]) {  // x is updated with p in k_g_direction
      const double ri = G.r[i] - a * G.q[i];
      const double zi = G.dinv[i] * ri;
      G.r[i] = ri;
      rz = ri * zi;
      rr = ri * ri;
    }
    rz = tile_sum(rz, sh);
    rr = tile_sum(rr, sh);
    if (threadIdx.x == 0) {
      G.part[t] = rz;
      G.part[G.ntiles + t] = rr;
    }
  });
}

// one workgroup per mesh m; the last to finish (an arrival counter) sets
// the iteration count and, when every mesh is done, the done flag
__global__ __launch_bounds__(kNT) void k_g_beta(GroupBuffers G) {
  GroupState* st = G.st;
  if (st->done) return;
  __shared__ double sh[kNT / 64];
  const int m = blockIdx.x;
  const int it = st->it + 1;
  double rz = 0.0, rr = 0.0;
  const int live = !G.done[m];
  if (live) {  // the same in every thread of the workgroup
    const int64_t t0 = G.tile_ptr[m], t1 = G.tile_ptr[m + 1];
    rz = run_sum(G.part, t0, t1, sh);
    rr = run_sum(G.part + G.ntiles, t0, t1, sh);
  }
  if (threadIdx.x == 0) {
    double beta = 0.0;
    if (live) {
      beta = rz / G.rho[m];
      G.rho[m] = rz;
      G.rr[m] = rr;
      G.it[m] = it;
      if (rr <= st->rtol2 * G.bb[m] || it >= st->maxit) {
        G.done[m] = 1;
        beta = 0.0;
      }
    }
    G.beta[m] = beta;
    __threadfence();
    if (atomicAdd(&st->arrived, 1) == G.nmesh - 1) {
      __threadfence();
      int all = 1;
      for (int j = 0; j < G.nmesh; ++j)
        all &= __hip_atomic_load(&G.done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      st->arrived = 0;
      st->it = it;
      if (all) st->done = 1;
    }
  }
}

// x += alpha p with this iteration's p (alpha = 0 for meshes done before
// it), then p = D^-1 r + beta p.  The iteration that finishes the group
// returns here before the x update (done is set); k_g_x_finish applies it.
__global__ __launch_bounds__(kNT) void k_g_direction(GroupBuffers G) {
  if (G.st->done) return;
  tiles_loop(G, [&](int64_t t, int m) {
    const int64_t i = G.tile_r0[t] + threadIdx.x;
    if (i < G.tile_r1[t]) {
      const double pk = G.p[i];
      G.x[i] += G.alpha[m] * pk;
      G.p[i] = G.dinv[i] * G.r[i] + G.beta[m] * pk;
    }
  });
}

// the x update of the group's last iteration (alpha = 0 for meshes finished
// before it, and after the setup)
__global__ __launch_bounds__(kNT) void k_g_x_finish(GroupBuffers G) {
  tiles_loop(G, [&](int64_t t, int m) {
    const int64_t i = G.tile_r0[t] + threadIdx.x;
    if (i < G.tile_r1[t]) G.x[i] += G.alpha[m] * G.p[i];
  });
}

// per-mesh ||b - K x|| / ||b||; q holds K x
__global__ __launch_bounds__(kNT) void k_g_resid(GroupBuffers G) {
  __shared__ double sh[kNT / 64];
  tiles_loop(G, [&](int64_t t, int m) {
    const int64_t i = G.tile_r0[t] + threadIdx.x;
    double s = 0, bb = 0;
    if (i < G.tile_r1[t]) {
      const double d = G.b[i] - G.q[i];
      s = d * d;
      bb = G.b[i] * G.b[i];
    }
    s = tile_sum(s, sh);
    bb = tile_sum(bb, sh);
    if (threadIdx.x == 0) {
      G.part[t] = s;
      G.part[G.ntiles + t] = bb;
    }
    (void)m;
  });
}

__global__ __launch_bounds__(kNT) void k_g_resid_finish(GroupBuffers G, double* out) {
  for (int m = threadIdx.x; m < G.nmesh; m += kNT) {
    double s = 0, bb = 0;
    for (int64_t t = G.tile_ptr[m]; t < G.tile_ptr[m + 1]; ++t) {
      s += G.part[t];
      bb += G.part[G.ntiles + t];
    }
    out[m] = bb > 0.0 ? sqrt(s / bb) : sqrt(s);
  }
}

inline int grid_for(int64_t work, int threads, int cap) {
  int64_t b = (work + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return int(b);
}

}  // namespace

size_t group_state_bytes() { return sizeof(GroupState); }

int group_blocks(int64_t ntiles) {
  int64_t nb = ntiles < kMaxPcgBlocks ? ntiles : kMaxPcgBlocks;
  if (nb >= 8) nb &= ~int64_t(7);
  return int(nb < 1 ? 1 : nb);
}

hipError_t launch_group_penalty(hipStream_t s, const GroupBuffers& G, double coef,
                                const int64_t* dofs, const int32_t* mesh, const int32_t* counts,
                                int64_t nd) {
  hipLaunchKernelGGL(k_g_diag_max, dim3(G.nmesh), dim3(kNT), 0, s, G, coef);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nd == 0) return e;
  hipLaunchKernelGGL(k_g_penalty_apply, dim3(grid_for(nd, 256, 4096)), dim3(256), 0, s, G, dofs,
                     mesh, counts, nd);
  return hipGetLastError();
}

hipError_t launch_group_setup(hipStream_t s, const GroupBuffers& G, int use_x0, double rtol,
                              int maxit) {
  hipError_t e;
  if (use_x0) {
    e = launch_spmv(s, kThreadU4, G.n, G.val, G.col, G.rownnz, G.x, G.q, 1);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_g_init, dim3(G.nblocks), dim3(kNT), 0, s, G, use_x0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_g_init_finish, dim3(G.nmesh), dim3(kNT), 0, s, G, rtol * rtol, maxit);
  return hipGetLastError();
}

hipError_t launch_group_iteration(hipStream_t s, const GroupBuffers& G) {
  hipError_t e;
  const dim3 g(G.nblocks), t(kNT);
  const dim3 gs(G.nsg < 65536 ? (G.nsg > 0 ? G.nsg : 1) : 65536);
  if (G.bsell_d == 2) hipLaunchKernelGGL(k_g_spmv_pq_bs<2>, gs, t, 0, s, G);
  else if (G.bsell_d == 3) hipLaunchKernelGGL(k_g_spmv_pq_bs<3>, gs, t, 0, s, G);
  else switch (G.variant) {
    case kThreadU1: hipLaunchKernelGGL(k_g_spmv_pq<kThreadU1>, g, t, 0, s, G); break;
    case kThreadU4: hipLaunchKernelGGL(k_g_spmv_pq<kThreadU4>, g, t, 0, s, G); break;
    case kNode2: hipLaunchKernelGGL(k_g_spmv_pq<kNode2>, g, t, 0, s, G); break;
    case kNode3: hipLaunchKernelGGL(k_g_spmv_pq<kNode3>, g, t, 0, s, G); break;
    default: hipLaunchKernelGGL(k_g_spmv_pq<kThreadU8>, g, t, 0, s, G); break;
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_g_alpha, dim3(G.nmesh), t, 0, s, G);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_g_update, g, t, 0, s, G);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_g_beta, dim3(G.nmesh), t, 0, s, G);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_g_direction, g, t, 0, s, G);
  return hipGetLastError();
}

hipError_t launch_group_x_finish(hipStream_t s, const GroupBuffers& G) {
  hipLaunchKernelGGL(k_g_x_finish, dim3(G.nblocks), dim3(kNT), 0, s, G);
  return hipGetLastError();
}

hipError_t launch_group_residual(hipStream_t s, const GroupBuffers& G, double* out) {
  hipError_t e = launch_spmv(s, kThreadU4, G.n, G.val, G.col, G.rownnz, G.x, G.q, 1);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_g_resid, dim3(G.nblocks), dim3(kNT), 0, s, G);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_g_resid_finish, dim3(1), dim3(kNT), 0, s, G, out);
  return hipGetLastError();
}

}  // namespace gf
