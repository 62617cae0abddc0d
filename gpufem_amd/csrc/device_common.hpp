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

// Sum of np per-block partials, identical in every workgroup.
__device__ inline double reduce_partials(const double* part, int np, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += kNT) v += part[i];
  return block_sum(v, sh);
}

// Workgroups are dispatched round-robin over the 8 XCDs; the remap gives
// XCD k the logical blocks [k*nb/8, (k+1)*nb/8), i.e. a contiguous slab of
// rows, so gathers of neighbouring rows share that XCD's L2.
__device__ inline int logical_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// Contiguous rows [r0, r1) of this workgroup (multiples of a wavefront).
__device__ inline void row_range(int64_t n, int64_t* r0, int64_t* r1) {
  const int64_t nb = gridDim.x;
  int64_t R = (n + nb - 1) / nb;
  R = (R + 63) & ~int64_t(63);
  const int64_t lb = logical_block();
  int64_t a = lb * R, e = a + R;
  if (a > n) a = n;
  if (e > n) e = n;
  *r0 = a;
  *r1 = e;
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

// Calls f(i, (K x)_i) for every row of this workgroup's range, with the
// SpMV variant V (thread per row with 1/4/8 slots in flight, or 2/4 waves
// sharing a tile of rows).  Every thread of the workgroup must call it.
template <int V, class T, class F>
__device__ inline void for_rows_spmv(int64_t n, const T* __restrict__ val,
                                     const int32_t* __restrict__ col,
                                     const int32_t* __restrict__ rownnz,
                                     const T* __restrict__ x, F&& f) {
  int64_t r0, r1;
  row_range(n, &r0, &r1);
  if constexpr (V <= kThreadU8) {
    constexpr int U = V == kThreadU1 ? 1 : (V == kThreadU4 ? 4 : 8);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kNT)
      f(i, row_dot<U, T>(i, n, val, col, rownnz[i], x));
  } else {
    constexpr int SUB = V == kTile2 ? 2 : 4;
    constexpr int RPB = kNT / SUB;  // rows per pass
    __shared__ T sh_part[SUB][RPB];
    const int rl = threadIdx.x % RPB, sub = threadIdx.x / RPB;
    for (int64_t base = r0; base < r1; base += RPB) {
      const int64_t i = base + rl;
      T acc = T(0);
      if (i < r1) {
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
      if (sub == 0 && i < r1) {
        T t = sh_part[0][rl];
#pragma unroll
        for (int k = 1; k < SUB; ++k) t += sh_part[k][rl];
        f(i, t);
      }
      __syncthreads();
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
    default:                                                                              \
      return hipErrorInvalidValue;                                                        \
  }
