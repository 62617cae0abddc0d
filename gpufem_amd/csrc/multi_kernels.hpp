// Device helpers shared by the batched solvers (multi.hip: FP64 block
// PCG; multi_mixed.hip: FP32 inner iterations in FP64 refinement).
//
// Vectors of S load cases are stored row-major, V[i*S + k].  The vector
// kernels give each row S lanes (lane k = load case k); the SpMM gives a
// row L = S / kC lanes, lane j owning load cases [j kC, j kC + kC), read
// and written as 16-byte pieces.  The L lanes of a row share each matrix
// entry (one fetch), and two or four slots are loaded before any is added,
// so a wave keeps many independent loads in flight; each output entry still
// adds its slots in ascending column order, as the single-RHS kernels do.
#pragma once

#include "device_common.hpp"
#include "launch.hpp"

namespace gf {
namespace mdev {

using namespace dev;

template <int S>
__device__ inline void groups(int* g, int* k) {
  *k = threadIdx.x % S;
  *g = threadIdx.x / S;
}

// rows of the workgroup: tiles of 256/S*8 rows would change the ownership of
// the single-RHS kernels; here a tile is the group count times 8 rows.
template <int S>
struct Tiles {
  static constexpr int G = kNT / S;   // row groups per workgroup
  static constexpr int RT = 8 * G;    // rows per tile
};

template <int S, class F>
__device__ inline void for_group_rows(int64_t n, F&& f) {
  int g, k;
  groups<S>(&g, &k);
  const int64_t stride = int64_t(gridDim.x) * Tiles<S>::RT;
  for (int64_t base = int64_t(logical_block()) * Tiles<S>::RT; base < n; base += stride) {
    const int64_t end = base + Tiles<S>::RT < n ? base + Tiles<S>::RT : n;
    for (int64_t i = base + g; i < end; i += Tiles<S>::G) f(i, k);
  }
}

// per-column sums over the workgroup: every thread of column k gets the sum
template <int S>
__device__ inline double column_sum(double v, double* sh) {
  int g, k;
  groups<S>(&g, &k);
  __syncthreads();
  sh[g * S + k] = v;
  __syncthreads();
  double t = 0.0;
  for (int j = 0; j < Tiles<S>::G; ++j) t += sh[j * S + k];
  return t;
}

// Single workgroup: sums[k] = sum over b of part[(base + k) * nb + b] for the S
// columns, in a fixed order (row group g takes b = g, g + G, ...; groups are
// then added in order).  Result in LDS out[S], visible after the call.
template <int S>
__device__ inline void reduce_columns(const double* part, int nb, int base, double* sh,
                                      double* out) {
  int g, k;
  groups<S>(&g, &k);
  const double* src = part + int64_t(base + k) * nb;
  constexpr int G = Tiles<S>::G;
  double t = 0.0;
  int bb = g;
  for (; bb + 7 * G < nb; bb += 8 * G) {  // eight loads in flight, added in order
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[bb + u * G];
#pragma unroll
    for (int u = 0; u < 8; ++u) t += v[u];
  }
  for (; bb < nb; bb += G) t += src[bb];
  __syncthreads();
  sh[g * S + k] = t;
  __syncthreads();
  if (g == 0) {
    double u = 0.0;
    for (int j = 0; j < Tiles<S>::G; ++j) u += sh[j * S + k];
    out[k] = u;
  }
  __syncthreads();
}

// One workgroup per column: the sum over b of part[col * nb + b] in a fixed
// order (thread t adds b = t, t + kNT, ... in turn, then a tree over the
// workgroup).  The result is returned to every thread.
__device__ inline double reduce_column(const double* __restrict__ part, int nb, int col,
                                       double* sh) {
  const double* src = part + int64_t(col) * nb;
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += kNT) t += src[b];
  __syncthreads();
  sh[threadIdx.x] = t;
  __syncthreads();
#pragma unroll
  for (int off = kNT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  return sh[0];
}

constexpr int kC = 4;
// load cases per lane in the FP64 block SpMM.  8 would halve the lanes that
// receive each block's 72 bytes of values, but a workgroup pass then covers
// twice the node rows, the rows in flight per XCD gather from twice the rows
// of p, and the block iteration measured 8 % slower (Brick20 48^3, 32 cases).
constexpr int kCB = 4;

// KC load cases per lane (kC, or 8 in the FP64 block SpMM)
template <int S, int KC = kC>
struct Chunks {
  static constexpr int L = S / KC;   // lanes per row
  static constexpr int G = kNT / L;  // rows per workgroup pass
};

template <int S, int KC = kC>
__device__ inline void chunk_lane(int* g, int* c0) {
  *g = threadIdx.x / Chunks<S, KC>::L;
  *c0 = (threadIdx.x % Chunks<S, KC>::L) * KC;
}

using dv2 = double __attribute__((ext_vector_type(2)));
using fv4 = float __attribute__((ext_vector_type(4)));

__device__ inline void load_chunk(const double* __restrict__ p, double* v) {
  const dv2 a = *reinterpret_cast<const dv2*>(p);
  const dv2 b = *reinterpret_cast<const dv2*>(p + 2);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

__device__ inline void load_chunk(const float* __restrict__ p, float* v) {
  const fv4 a = *reinterpret_cast<const fv4*>(p);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}

__device__ inline void store_chunk(double* p, const double* v) {
  dv2 a, b;
  a.x = v[0];
  a.y = v[1];
  b.x = v[2];
  b.y = v[3];
  *reinterpret_cast<dv2*>(p) = a;
  *reinterpret_cast<dv2*>(p + 2) = b;
}

__device__ inline void store_chunk(float* p, const float* v) {
  fv4 a;
  a.x = v[0];
  a.y = v[1];
  a.z = v[2];
  a.w = v[3];
  *reinterpret_cast<fv4*>(p) = a;
}

// FP64 chunks in split order: lane l = c0 / KC of a row owns the KC/2 pairs
// of load cases {2l, 2l + 1} + q S / (KC/2), q = 0 .. KC/2 - 1, so each
// 16-byte load of the L lanes of a row reads one contiguous piece of the row
// (one 128-byte line at S = 32, KC = 4) instead of parts of two lines.
template <int S, int KC>
__device__ inline int split_col(int c0, int j) {
  return (j / 2) * (S / (KC / 2)) + 2 * (c0 / KC) + (j & 1);
}

template <int S, int KC>
__device__ inline void load_split(const double* __restrict__ row, int c0, double* v) {
#pragma unroll
  for (int q = 0; q < KC / 2; ++q) {
    const dv2 a = *reinterpret_cast<const dv2*>(row + split_col<S, KC>(c0, 2 * q));
    v[2 * q] = a.x;
    v[2 * q + 1] = a.y;
  }
}

template <int S, int KC>
__device__ inline void store_split(double* row, int c0, const double* v) {
#pragma unroll
  for (int q = 0; q < KC / 2; ++q) {
    dv2 a;
    a.x = v[2 * q];
    a.y = v[2 * q + 1];
    *reinterpret_cast<dv2*>(row + split_col<S, KC>(c0, 2 * q)) = a;
  }
}

// the arrays an SpMM reads besides the values: ELL columns and row lengths,
// or the block-sliced copy (bs.nslices > 0 with the copy's node columns)
struct SpmmArgs {
  int64_t n;
  const int32_t* rownnz;
  const int32_t* col;
  BsellView bs;
};

// rows of this workgroup in the tiles for_group_rows uses (the same
// ownership as the vector kernels), G rows per pass
template <int S, class F>
__device__ inline void for_chunk_rows(int64_t n, F&& f) {
  int g, c0;
  chunk_lane<S>(&g, &c0);
  const int64_t stride = int64_t(gridDim.x) * Tiles<S>::RT;
  for (int64_t base = int64_t(logical_block()) * Tiles<S>::RT; base < n; base += stride) {
    const int64_t end = base + Tiles<S>::RT < n ? base + Tiles<S>::RT : n;
    for (int64_t i = base + g; i < end; i += Chunks<S>::G) f(i, c0);
  }
}

// row i of the slot-major ELL matrix times load cases [c0, c0 + kC) of xin,
// accumulated in T
template <int S, class T, class VT>
__device__ inline void ell_row_chunk(const SpmmArgs& A, const VT* __restrict__ val,
                                     const T* __restrict__ xin, int64_t i, int c0, T* acc) {
  const int64_t n = A.n;
  const int32_t rn = A.rownnz[i];
#pragma unroll
  for (int j = 0; j < kC; ++j) acc[j] = T(0);
  int32_t s = 0;
  for (; s + 4 <= rn; s += 4) {
    T v[4], xv[4][kC];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t idx = i + int64_t(s + u) * n;
      v[u] = T(val[idx]);
      load_chunk(xin + int64_t(A.col[idx]) * S + c0, xv[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < kC; ++j) acc[j] += v[u] * xv[u][j];
  }
  for (; s < rn; ++s) {
    const int64_t idx = i + int64_t(s) * n;
    T xv[kC];
    load_chunk(xin + int64_t(A.col[idx]) * S + c0, xv);
    const T v = T(val[idx]);
#pragma unroll
    for (int j = 0; j < kC; ++j) acc[j] += v * xv[j];
  }
}

// Block-sliced copy: positions of the node-row permutation in grid-stride
// order, G per workgroup pass; f(R, c0, acc) gets the D rows of node row R.
// DIFF: the values are the difference form of the FP32 copy (see bsell_row):
// off-diagonal blocks multiply x_J - x_R, the diagonal slot (node column R)
// holds the block row sum and multiplies x_R.
// NB: blocks loaded per step (their loads in flight together); the blocks
// are added in slot order whatever NB is.  SPLIT (FP64 only): the lanes hold
// their KC load cases in split order (load_split), acc included.  SHARE: the L
// lanes of a row load a step's block values once between them and exchange
// them (a cross-lane permute through the LDS crossbar) instead of each loading
// all of them; the vector-memory path then carries mostly the gathers of x.
template <int S, int D, class T, class VT, bool DIFF = false, int NB = 2, bool SPLIT = false,
          int KC = kC, bool SHARE = false, class F>
__device__ inline void for_chunk_node_rows(const SpmmArgs& A, const VT* __restrict__ val,
                                           const T* __restrict__ xin, F&& f) {
  static_assert(!(DIFF && SPLIT), "the difference form reads x_R in the contiguous order");
  int g, c0;
  chunk_lane<S, KC>(&g, &c0);
  const int64_t npos = A.bs.nslices * 64;
  const int64_t stride = int64_t(gridDim.x) * Chunks<S, KC>::G;
  for (int64_t pos = int64_t(logical_block()) * Chunks<S, KC>::G + g; pos < npos; pos += stride) {
    const int32_t R = A.bs.perm[pos];
    if (R < 0) continue;
    const int lp = int(pos & 63);
    const int32_t cnt = A.rownnz[int64_t(D) * R] / D;
    const int64_t b0 = A.bs.sptr[pos >> 6];
    T acc[D][KC];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int j = 0; j < KC; ++j) acc[r][j] = T(0);
    auto gather = [&](int64_t row, T* xv) {
      if constexpr (SPLIT) load_split<S, KC>(xin + row * S, c0, xv);
      else load_chunk(xin + row * S + c0, xv);
    };
    T xr[D][KC];
    if constexpr (DIFF) {
#pragma unroll
      for (int e = 0; e < D; ++e) load_chunk(xin + (int64_t(D) * R + e) * S + c0, xr[e]);
    }
    // x_J - x_R off the diagonal (DIFF)
    auto rel = [&](int64_t c, T (*xv)[KC]) {
      if constexpr (DIFF) {
        if (c != int64_t(D) * R) {
#pragma unroll
          for (int e = 0; e < D; ++e)
#pragma unroll
            for (int j = 0; j < KC; ++j) xv[e][j] -= xr[e][j];
        }
      }
    };
    // The node columns of the next NB blocks are loaded one step ahead, so a
    // step issues its gathers of x at once instead of first waiting for its
    // columns; the look-ahead index is clamped to the row's last block (the
    // extra load is never used).
    int32_t cn[NB];
    const int64_t last = b0 + (cnt > 0 ? cnt - 1 : 0);
    auto col_at = [&](int32_t jn) {
      const int64_t b = b0 + jn;
      return A.col[(b < last ? b : last) * 64 + lp];
    };
    if (cnt > 0) {
#pragma unroll
      for (int u = 0; u < NB; ++u) cn[u] = col_at(u);
    }
    int32_t jb = 0;
    for (; jb + NB <= cnt; jb += NB) {
      T v[NB][D * D], xv[NB][D][KC];
      int64_t cu[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int64_t b = b0 + jb + u;
        const int64_t c = int64_t(cn[u]) * D;
        cu[u] = c;
#pragma unroll
        for (int e = 0; e < D; ++e) gather(c + e, xv[u][e]);
        if constexpr (!SHARE) {
#pragma unroll
          for (int k = 0; k < D * D; ++k) v[u][k] = T(val[bsell_index<VT>(D * D, b, k, lp)]);
        }
      }
      if constexpr (SHARE) {
        // lane ln of the row loads vector ln % P (a pair of doubles or a quad
        // of floats) of block ln / P, and the lanes ln < NB the block's odd
        // last value; every lane of the row then takes the step's values from
        // those lanes
        constexpr int L = Chunks<S, KC>::L;
        constexpr int W = sizeof(VT) == 4 ? 4 : 2;
        constexpr int P = D * D / W;
        constexpr bool ODD = D * D - P * W == 1;
        static_assert(NB * P <= L && D * D - P * W <= 1, "a step's value vectors need a lane each");
        using VW = VT __attribute__((ext_vector_type(W)));
        const int ln = c0 / KC;
        VW pr = {};
        if (ln < NB * P)
          pr = *reinterpret_cast<const VW*>(
              val + bsell_index<VT>(D * D, b0 + jb + ln / P, W * (ln % P), lp));
        VT sg = VT(0);
        if constexpr (ODD) {
          if (ln < NB) sg = val[bsell_index<VT>(D * D, b0 + jb + ln, D * D - 1, lp)];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
#pragma unroll
          for (int q = 0; q < P; ++q)
#pragma unroll
            for (int w = 0; w < W; ++w) v[u][W * q + w] = T(__shfl(pr[w], u * P + q, L));
          if constexpr (ODD) v[u][D * D - 1] = T(__shfl(sg, u, L));
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) cn[u] = col_at(jb + NB + u);
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        rel(cu[u], xv[u]);
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
          for (int e = 0; e < D; ++e)
#pragma unroll
            for (int j = 0; j < KC; ++j) acc[r][j] += v[u][r * D + e] * xv[u][e][j];
      }
    }
    // the last cnt - jb < NB blocks, whose columns are already in cn
#pragma unroll
    for (int u = 0; u + 1 < NB; ++u) {
      if (jb + u < cnt) {
        const int64_t b = b0 + jb + u;
        const int64_t c = int64_t(cn[u]) * D;
        T v[D * D], xv[D][KC];
#pragma unroll
        for (int e = 0; e < D; ++e) gather(c + e, xv[e]);
#pragma unroll
        for (int k = 0; k < D * D; ++k) v[k] = T(val[bsell_index<VT>(D * D, b, k, lp)]);
        rel(c, xv);
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
          for (int e = 0; e < D; ++e)
#pragma unroll
            for (int j = 0; j < KC; ++j) acc[r][j] += v[r * D + e] * xv[e][j];
      }
    }
    f(R, c0, acc);
  }
}

// per-column workgroup sums of the lanes' chunk partials into part[k][block]
// (SPLIT: the partials are in split order)
template <int S, bool SPLIT = false, int KC = kC>
__device__ inline void chunk_partials(const double* v, double* sh, double* part) {
  int g, c0;
  chunk_lane<S, KC>(&g, &c0);
#pragma unroll
  for (int j = 0; j < KC; ++j) sh[g * S + (SPLIT ? split_col<S, KC>(c0, j) : c0 + j)] = v[j];
  __syncthreads();
  if (threadIdx.x < S) {
    double t = 0.0;
    for (int q = 0; q < Chunks<S, KC>::G; ++q) t += sh[q * S + threadIdx.x];
    part[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
  }
}

}  // namespace mdev
}  // namespace gf
