// Device helpers shared by the solver kernels (solver.hip, mixed.hip):
// deterministic block reductions, the XCD-aware contiguous row partition and
// the ELL row kernels of every SpMV variant.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.hpp"

namespace gf {
namespace dev {

constexpr int kNT = 256;  // threads per workgroup of every solver kernel

// Sum over the workgroup; every thread gets the same (bitwise) result.
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

// Sum of np per-block partials, identical in every workgroup (a fixed order:
// four strided accumulators per thread, so four loads are in flight when the
// SpMV grid leaves thousands of partials).
__device__ inline double reduce_partials(const double* part, int np, double* sh) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  int i = threadIdx.x;
  for (; i + 3 * kNT < np; i += 4 * kNT) {
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] += part[i + u * kNT];
  }
  // at most three left
  if (i < np) v[0] += part[i];
  if (i + kNT < np) v[1] += part[i + kNT];
  if (i + 2 * kNT < np) v[2] += part[i + 2 * kNT];
  return block_sum((v[0] + v[1]) + (v[2] + v[3]), sh);
}

// Workgroups are dispatched round-robin over the 8 XCDs; the remap gives
// XCD k the logical blocks [k*nb/8, (k+1)*nb/8), i.e. a contiguous slab of
// rows, so gathers of neighbouring rows share that XCD's L2.
__device__ inline int xcd_block(int b, int nb) {
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}
__device__ inline int logical_block() { return xcd_block(blockIdx.x, gridDim.x); }

// Rows are visited in tiles of kNT consecutive rows; tile t belongs to
// logical workgroup t mod gridDim.x (grid-stride sweeps).  At any moment
// the whole grid works inside one window of gridDim.x tiles, which keeps
// the slot-major ELL streams DRAM-page friendly, and each XCD owns a
// contiguous part of every window.  The PCG vector kernels use this
// ownership (the slice-group SpMV grids own one slab of rows per XCD instead;
// giving the vector kernels the same slabs measured no faster on the Q8
// plate).  for_rows_of: the rows of logical workgroup lb of nlb (a
// persistent kernel plays several logical workgroups in turn).
template <class F>
__device__ inline void for_rows_of(int lb, int nlb, int64_t n, F&& f) {
  const int64_t stride = int64_t(nlb) * kNT;
  for (int64_t base = int64_t(lb) * kNT; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    if (i < n) f(i);
  }
}
template <class F>
__device__ inline void for_my_rows(int64_t n, F&& f) {
  for_rows_of(logical_block(), gridDim.x, n, f);
}

// One row of a slot-major ELL matrix, U slots in flight; slots are summed in
// order, as ELLmatrix::Axy does (src/SPRmatrix/ELLmatrix.cc:194-204).
template <int U, class T>
__device__ inline T row_dot(int64_t i, int64_t n, const T* __restrict__ val,
                            const int32_t* __restrict__ col, int32_t rn,
                            const T* __restrict__ x) {
  T acc = T(0);
  int32_t s = 0;
  if constexpr (U > 1) {
    for (; s + U <= rn; s += U) {
      T v[U];
      int32_t c[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t idx = i + int64_t(s + u) * n;
        v[u] = __builtin_nontemporal_load(val + idx);
        c[u] = __builtin_nontemporal_load(col + idx);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) acc += v[u] * x[c[u]];
    }
  }
  for (; s < rn; ++s) {
    const int64_t idx = i + int64_t(s) * n;
    acc += __builtin_nontemporal_load(val + idx) * x[__builtin_nontemporal_load(col + idx)];
  }
  return acc;
}

// One row over node blocks: slots d j .. d j + d - 1 hold the d DOFs of node
// coln[i + j n]; summed in slot order, so bitwise equal to row_dot.
template <int D, class T>
__device__ inline T row_dot_node(int64_t i, int64_t n, const T* __restrict__ val,
                                 const int32_t* __restrict__ coln, int32_t rn,
                                 const T* __restrict__ x) {
  T acc = T(0);
  const int32_t nb = rn / D;
  int32_t j = 0;
  for (; j + 2 <= nb; j += 2) {
    T v[2][D];
    int64_t c[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      c[u] = int64_t(__builtin_nontemporal_load(coln + i + int64_t(j + u) * n)) * D;
#pragma unroll
      for (int d = 0; d < D; ++d)
        v[u][d] = __builtin_nontemporal_load(val + i + int64_t((j + u) * D + d) * n);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int d = 0; d < D; ++d) acc += v[u][d] * x[c[u] + d];
  }
  for (; j < nb; ++j) {
    const int64_t c = int64_t(__builtin_nontemporal_load(coln + i + int64_t(j) * n)) * D;
#pragma unroll
    for (int d = 0; d < D; ++d)
      acc += __builtin_nontemporal_load(val + i + int64_t(j * D + d) * n) * x[c + d];
  }
  return acc;
}

// the d*d values of block slot b for this lane, 16 bytes per load: FP64
// pairs (bsell_vidx) or FP32 quads (bsell_vidx32)
template <int D, class T>
__device__ inline void load_block(const T* __restrict__ bval, int64_t b, int lane, T* v) {
  constexpr int DD = D * D;
  const T* base = bval + b * DD * 64;
  if constexpr (sizeof(T) == 8) {
    typedef double P2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int kp = 0; kp < DD / 2; ++kp) {
      const P2 pr = __builtin_nontemporal_load(reinterpret_cast<const P2*>(base) + kp * 64 + lane);
      v[2 * kp] = pr.x;
      v[2 * kp + 1] = pr.y;
    }
    if constexpr (DD & 1) v[DD - 1] = __builtin_nontemporal_load(base + (DD - 1) * 64 + lane);
  } else {
    typedef float P4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int kq = 0; kq < DD / 4; ++kq) {
      const P4 q = __builtin_nontemporal_load(reinterpret_cast<const P4*>(base) + kq * 64 + lane);
      v[4 * kq] = q.x;
      v[4 * kq + 1] = q.y;
      v[4 * kq + 2] = q.z;
      v[4 * kq + 3] = q.w;
    }
#pragma unroll
    for (int k = DD & ~3; k < DD; ++k) v[k] = __builtin_nontemporal_load(base + k * 64 + lane);
  }
}

template <class T, int D>
struct VecOf;
template <>
struct VecOf<double, 2> {
  using type = double2;
};
template <>
struct VecOf<float, 2> {
  using type = float2;
};

// the d entries x[c .. c + d - 1] of one node (one vector load when d = 2)
template <int D, class T>
__device__ inline void load_node(const T* __restrict__ x, int64_t c, T* xv) {
  if constexpr (D == 2) {
    const auto v = *reinterpret_cast<const typename VecOf<T, 2>::type*>(x + c);
    xv[0] = v.x;
    xv[1] = v.y;
  } else {
#pragma unroll
    for (int e = 0; e < D; ++e) xv[e] = x[c + e];
  }
}

// Node row R of the block-sliced copy, lane l of slice s: acc[r] = (K x)_{dR+r}
// over cnt block slots from b0 (the row's blocks, then zero padding).  The
// node column of a slot is base + coff[...] (16-bit offsets, OFF) or bcol[...].
// Entries are added in ascending column order, as row_dot adds its slots.
// U blocks per step (measured on the Q8 plate and the Brick8 cube: 4 for
// 2 x 2 blocks, 3 for 3 x 3; fewer lose more in-flight loads than the extra
// occupancy gives back, more cost occupancy; the FP32 copy too, whose blocks
// carry half the bytes: 4 or 6 per step made the cube's mixed solve 1-4 %
// slower).
//
// DIFF: the copy is in difference form (the FP32 copy of the mixed-precision
// solve): an off-diagonal block K_RJ multiplies x_J - x_R and the diagonal
// slot holds the block row sum S_R = sum_J K_RJ, so
//   (K x)_R = sum_{J != R} K_RJ (x_J - x_R) + S_R x_R.
// The stored operator then annihilates rigid translations exactly whatever
// the rounding of its entries (the row sums are rounded once, not the sum of
// rounded entries), which keeps an FP32 elasticity operator positive
// definite: a plain FP32 copy of a structured mesh repeats the same rounding
// in every row, which shifts every eigenvalue by about the same amount.
// Padding slots point at the row's own node and hold zeros (adding 0).
template <int D, class T, bool OFF>
__device__ inline int64_t bsell_col(const int32_t* __restrict__ bcol,
                                    const int16_t* __restrict__ coff, int32_t base, int64_t b,
                                    int lane) {
  if constexpr (OFF)
    return int64_t(base + int32_t(__builtin_nontemporal_load(coff + b * 64 + lane))) * D;
  else
    return int64_t(__builtin_nontemporal_load(bcol + b * 64 + lane)) * D;
}

template <int D, class T, bool OFF, bool DIFF = false>
__device__ inline void bsell_row(const BsellView& bs, const T* __restrict__ bval,
                                 const int32_t* __restrict__ bcol, int64_t b0, int32_t cnt,
                                 int lane, const T* __restrict__ x, T* acc, int32_t base,
                                 int32_t row = 0) {
  constexpr int U = D == 2 ? 4 : 3;
#pragma unroll
  for (int r = 0; r < D; ++r) acc[r] = T(0);
  T xr[D];
  if constexpr (DIFF) load_node<D, T>(x, int64_t(row) * D, xr);
  // x_J - x_R off the diagonal, x_R on it
  auto rel = [&](int64_t c, T* xv) {
    if constexpr (DIFF) {
      if (c != int64_t(row) * D) {
#pragma unroll
        for (int e = 0; e < D; ++e) xv[e] -= xr[e];
      }
    }
  };
  int32_t j = 0;
  for (; j + U <= cnt; j += U) {
    T xv[U][D], v[U][D * D];
    int64_t cu[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t b = b0 + j + u;
      cu[u] = bsell_col<D, T, OFF>(bcol, bs.coff, base, b, lane);
      load_block<D, T>(bval, b, lane, v[u]);
      load_node<D, T>(x, cu[u], xv[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      rel(cu[u], xv[u]);
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int e = 0; e < D; ++e) acc[r] += v[u][r * D + e] * xv[u][e];
    }
  }
  for (; j < cnt; ++j) {
    const int64_t b = b0 + j;
    const int64_t c = bsell_col<D, T, OFF>(bcol, bs.coff, base, b, lane);
    T xv[D], v[D * D];
    load_block<D, T>(bval, b, lane, v);
    load_node<D, T>(x, c, xv);
    rel(c, xv);
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int e = 0; e < D; ++e) acc[r] += v[r * D + e] * xv[e];
  }
}

// Calls f(i, (K x)_i) for every row of logical workgroup lb of nlb, with
// the SpMV variant V (thread per row with 1/4/8 slots in flight, 2/4 waves
// sharing a tile of rows, node blocks, block-sliced slices).  Every thread of
// the workgroup must call it.
template <int V, class T, bool DIFF = false, class F>
__device__ inline void for_rows_spmv(int lb, int nlb, int64_t n, const T* __restrict__ val,
                                     const int32_t* __restrict__ col,
                                     const int32_t* __restrict__ rownnz,
                                     const T* __restrict__ x, F&& f, const BsellView& bs) {
  if constexpr (V == kBsell2 || V == kBsell3 || V == kBsell2G || V == kBsell3G) {
    // four slices (one per wave) per workgroup and sweep, grid-stride (one
    // sweep when the grid has a workgroup per group)
    constexpr int D = (V == kBsell2 || V == kBsell2G) ? 2 : 3;
    const int lane = threadIdx.x & 63;
    const int64_t stride = int64_t(nlb) * (kNT / 64);
    for (int64_t sl = int64_t(lb) * (kNT / 64) + (threadIdx.x >> 6); sl < bs.nslices;
         sl += stride) {
      // every lane runs the slice's length: the loads start without waiting
      // for the row number, and padding blocks add zeros (x finite: the same
      // sums, up to the sign of a zero sum)
      const int64_t b0 = bs.sptr[sl];
      const int32_t w = int32_t(bs.sptr[sl + 1] - b0);
      const int32_t R = bs.perm[sl * 64 + lane];
      T acc[D];
      const int32_t Rn = R >= 0 ? R : 0;
      if (bs.coff) bsell_row<D, T, true, DIFF>(bs, val, col, b0, w, lane, x, acc, Rn, Rn);
      else bsell_row<D, T, false, DIFF>(bs, val, col, b0, w, lane, x, acc, 0, Rn);
      if (R < 0) continue;
#pragma unroll
      for (int r = 0; r < D; ++r) f(int64_t(D) * R + r, acc[r]);
    }
  } else if constexpr (V == kNode2 || V == kNode3) {
    constexpr int D = V == kNode2 ? 2 : 3;
    for_rows_of(lb, nlb, n, [&](int64_t i) { f(i, row_dot_node<D, T>(i, n, val, col, rownnz[i], x)); });
  } else if constexpr (V <= kThreadU8) {
    constexpr int U = V == kThreadU1 ? 1 : (V == kThreadU4 ? 4 : 8);
    for_rows_of(lb, nlb, n, [&](int64_t i) { f(i, row_dot<U, T>(i, n, val, col, rownnz[i], x)); });
  } else {
    // the workgroup's kNT-row tile is covered in SUB passes of RPB rows
    constexpr int SUB = V == kTile2 ? 2 : 4;
    constexpr int RPB = kNT / SUB;  // rows per pass
    __shared__ T sh_part[SUB][RPB];
    const int rl = threadIdx.x % RPB, sub = threadIdx.x / RPB;
    const int64_t stride = int64_t(nlb) * kNT;
    for (int64_t tile = int64_t(lb) * kNT; tile < n; tile += stride) {
      for (int pass = 0; pass < SUB; ++pass) {
        const int64_t i = tile + pass * RPB + rl;
        T acc = T(0);
        if (i < n) {
          const int32_t rn = rownnz[i];
          int32_t s = sub;
          for (; s + SUB < rn; s += 2 * SUB) {
            const int64_t i0 = i + int64_t(s) * n, i1 = i + int64_t(s + SUB) * n;
            const T v0 = __builtin_nontemporal_load(val + i0);
            const T v1 = __builtin_nontemporal_load(val + i1);
            const int32_t c0 = __builtin_nontemporal_load(col + i0);
            const int32_t c1 = __builtin_nontemporal_load(col + i1);
            acc += v0 * x[c0];
            acc += v1 * x[c1];
          }
          for (; s < rn; s += SUB) {
            const int64_t i0 = i + int64_t(s) * n;
            acc += __builtin_nontemporal_load(val + i0) * x[__builtin_nontemporal_load(col + i0)];
          }
        }
        sh_part[sub][rl] = acc;
        __syncthreads();
        if (sub == 0 && i < n) {
          T t = sh_part[0][rl];
#pragma unroll
          for (int k = 1; k < SUB; ++k) t += sh_part[k][rl];
          f(i, t);
        }
        __syncthreads();
      }
    }
  }
}

}  // namespace dev
}  // namespace gf

// Launches KERNEL<variant> (a template over the SpMV variant) or returns
// hipErrorInvalidValue from the enclosing function for an unknown variant.
#define GF_DISPATCH_VARIANT(v, KERNEL, GRID, STREAM, ...)                                  \
  switch (v) {                                                                            \
    case ::gf::kThreadU1:                                                                 \
      hipLaunchKernelGGL(KERNEL<::gf::kThreadU1>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kThreadU4:                                                                 \
      hipLaunchKernelGGL(KERNEL<::gf::kThreadU4>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kThreadU8:                                                                 \
      hipLaunchKernelGGL(KERNEL<::gf::kThreadU8>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kTile2:                                                                    \
      hipLaunchKernelGGL(KERNEL<::gf::kTile2>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kTile4:                                                                    \
      hipLaunchKernelGGL(KERNEL<::gf::kTile4>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kNode2:                                                                    \
      hipLaunchKernelGGL(KERNEL<::gf::kNode2>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kNode3:                                                                    \
      hipLaunchKernelGGL(KERNEL<::gf::kNode3>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kBsell2:                                                                   \
      hipLaunchKernelGGL(KERNEL<::gf::kBsell2>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kBsell3:                                                                   \
      hipLaunchKernelGGL(KERNEL<::gf::kBsell3>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kBsell2G:                                                                  \
      hipLaunchKernelGGL(KERNEL<::gf::kBsell2G>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    case ::gf::kBsell3G:                                                                  \
      hipLaunchKernelGGL(KERNEL<::gf::kBsell3G>, GRID, dim3(::gf::dev::kNT), 0, STREAM, __VA_ARGS__); \
      break;                                                                              \
    default:                                                                              \
      return hipErrorInvalidValue;                                                        \
  }
