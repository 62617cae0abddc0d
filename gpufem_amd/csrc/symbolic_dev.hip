// Device symbolic phase: the same node adjacency, first-fit element coloring
// and element -> slot map as the host phase (symbolic.cpp), built on the GPU
// from the connectivity already resident there.
//
// Adjacency: node -> incident elements by counting, a scan and an atomic
// fill (the order inside a node's list is arbitrary and nothing depends on
// it); then one wavefront per node gathers its candidate nodes into LDS and
// writes the distinct ones in ascending order (rank = number of distinct
// candidates below), which is the set the host phase sorts.
//
// Coloring: first-fit in element order (src/FEM/femColor.cc:132-191) in
// rounds: an uncolored element whose lower-numbered neighbours are all
// colored takes the smallest color none of them uses.  When every
// lower-numbered neighbour is final, this is exactly the choice the
// sequential sweep makes, so the coloring equals the host phase's (and the
// reference's); each round colors at least the lowest uncolored element.
// Colors are kept as bits of a 64-bit mask (at most 64 colors).
//
// Prefix sums (incidence and adjacency pointers) and the grouping of the
// elements by color (femColor.cc:172-179 appends each element to its color's
// list in element order) are done here too: a three-pass tiled exclusive
// scan, and a stable counting sort over tiles of elements whose per-(color,
// tile) counts are scanned color-major.  Both are deterministic.
#include <algorithm>
#include <string>
#include <vector>

#include "gf_internal.hpp"
#include "launch.hpp"
#include "symbolic.hpp"

namespace gf {
namespace {

constexpr int kMaxCand = 2048;  // candidate nodes per node (incident elements x nen)
constexpr int kRoundsPerCheck = 32;

inline int grid_for(int64_t work, int threads, int cap) {
  int64_t b = (work + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return int(b);
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x)
    p[t] = v;
}

// err codes: 1 node id out of range, 2 too many candidates, 3 > 64 colors
__global__ void k_incidence_count(const int32_t* __restrict__ conn, int64_t total, int64_t nnodes,
                                  int32_t* __restrict__ cnt, int32_t* err) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int32_t v = conn[t];
    if (v < 1 || v > nnodes) {
      atomicMax(err, 1);
      continue;
    }
    atomicAdd(cnt + (v - 1), 1);
  }
}

__global__ void k_incidence_fill(const int32_t* __restrict__ conn, int64_t nelem, int nen,
                                 const int64_t* __restrict__ ptr, int32_t* __restrict__ cursor,
                                 int32_t* __restrict__ n2e) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nelem * nen;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = conn[t] - 1;
    const int32_t pos = atomicAdd(cursor + v, 1);
    n2e[ptr[v] + pos] = int32_t(t / nen);
  }
}

// One wavefront (= workgroup) per node.  FILL = false: distinct count per
// node (and the maximum); FILL = true: the distinct nodes in ascending order.
template <bool FILL>
__global__ __launch_bounds__(64) void k_adjacency(const int32_t* __restrict__ conn, int nen,
                                                  const int64_t* __restrict__ n2e_ptr,
                                                  const int32_t* __restrict__ n2e, int64_t nnodes,
                                                  int32_t* __restrict__ cnt,
                                                  const int64_t* __restrict__ adj_ptr,
                                                  int32_t* __restrict__ adj, int32_t* maxcnt,
                                                  int32_t* err) {
  __shared__ int32_t s_c[kMaxCand];
  __shared__ uint8_t s_first[kMaxCand];
  const int lane = threadIdx.x;
  int32_t my_max = 0;  // this workgroup's largest count (one atomic at the end)
  for (int64_t v = blockIdx.x; v < nnodes; v += gridDim.x) {
    const int64_t p0 = n2e_ptr[v];
    const int m = int((n2e_ptr[v + 1] - p0) * nen);
    if (m > kMaxCand) {
      if (lane == 0) atomicMax(err, 2);
      continue;  // the same branch for the whole workgroup
    }
    if (m == 0) {  // isolated node: the diagonal only
      if (lane == 0) {
        if (FILL) adj[adj_ptr[v]] = int32_t(v);
        else {
          cnt[v] = 1;
          my_max = my_max > 1 ? my_max : 1;
        }
      }
      continue;
    }
    for (int i = lane; i < m; i += 64)
      s_c[i] = conn[int64_t(n2e[p0 + i / nen]) * nen + i % nen] - 1;
    __syncthreads();
    int local = 0;
    for (int i = lane; i < m; i += 64) {
      const int32_t c = s_c[i];
      uint8_t first = 1;
      for (int j = 0; j < i; ++j)
        if (s_c[j] == c) {
          first = 0;
          break;
        }
      s_first[i] = first;
      local += first;
    }
    __syncthreads();
    if (FILL) {
      const int64_t base = adj_ptr[v];
      for (int i = lane; i < m; i += 64) {
        if (!s_first[i]) continue;
        const int32_t c = s_c[i];
        int r = 0;
        for (int j = 0; j < m; ++j) r += (s_first[j] && s_c[j] < c) ? 1 : 0;
        adj[base + r] = c;
      }
    } else {
      for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off);
      if (lane == 0) {
        cnt[v] = local;
        my_max = my_max > local ? my_max : local;
      }
    }
    __syncthreads();  // LDS reused by the next node
  }
  if (!FILL && lane == 0 && my_max > 0) atomicMax(maxcnt, my_max);
}

// slot[e][a][b] = rank of node b in the sorted adjacency of node a
__global__ void k_slot_map(const int32_t* __restrict__ conn, int64_t nelem, int nen,
                           const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj,
                           uint8_t* __restrict__ slot) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nelem * nen;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = t / nen;
    const int64_t v = conn[t] - 1;
    const int64_t lo = adj_ptr[v], hi = adj_ptr[v + 1];
    for (int b = 0; b < nen; ++b) {
      const int32_t w = conn[e * nen + b] - 1;
      int64_t l = lo, h = hi;  // first position with adj >= w
      while (l < h) {
        const int64_t mid = (l + h) >> 1;
        if (adj[mid] < w) l = mid + 1;
        else h = mid;
      }
      slot[t * nen + b] = uint8_t(l - lo);
    }
  }
}

// One round of first-fit coloring in element order.
__global__ void k_color_round(const int32_t* __restrict__ conn, int64_t nelem, int nen,
                              const int64_t* __restrict__ n2e_ptr,
                              const int32_t* __restrict__ n2e, int32_t* color, int32_t* pending,
                              int32_t* err) {
  bool pend = false;  // some element of this thread waits for a neighbour
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nelem;
       e += int64_t(gridDim.x) * blockDim.x) {
    if (__hip_atomic_load(color + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) continue;
    uint64_t used = 0;
    bool ready = true;
    for (int a = 0; a < nen && ready; ++a) {
      const int64_t v = conn[e * nen + a] - 1;
      for (int64_t k = n2e_ptr[v]; k < n2e_ptr[v + 1]; ++k) {
        const int32_t f = n2e[k];
        if (f >= e) continue;
        const int32_t c = __hip_atomic_load(color + f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c == 0) {
          ready = false;
          break;
        }
        used |= 1ull << (c - 1);
      }
    }
    if (!ready) {
      pend = true;
      continue;
    }
    if (used == ~0ull) {
      atomicMax(err, 3);
      continue;
    }
    const int32_t c = __ffsll(static_cast<long long>(~used));  // 1-based lowest free
    __hip_atomic_store(color + e, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // one store per wavefront with a waiting element (pending was zeroed
  // before the round; every writer writes 1)
  const int any = __any(pend);
  if (any && (threadIdx.x & 63) == 0) *pending = 1;
}

// ---- exclusive prefix sums of int32 counts into int64 -------------------------
// Tiles of kScanItems consecutive entries, kScanThreads threads of
// kScanItems / kScanThreads consecutive entries each.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanItems = kScanThreads * kScanPer;

// inclusive scan of one value per thread over the workgroup (LDS)
__device__ inline int64_t block_inclusive_scan(int64_t v, int64_t* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int64_t add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  const int64_t r = sh[t];
  __syncthreads();
  return r;
}

// tile_sum[b] = sum of tile b
__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(const int32_t* __restrict__ in,
                                                                 int64_t n,
                                                                 int64_t* __restrict__ tile_sum) {
  __shared__ int64_t sh[kScanThreads];
  const int64_t base = int64_t(blockIdx.x) * kScanItems + int64_t(threadIdx.x) * kScanPer;
  int64_t v = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) v += in[base + k];
  v = block_inclusive_scan(v, sh);
  if (threadIdx.x == kScanThreads - 1) tile_sum[blockIdx.x] = v;
}

// one workgroup: exclusive scan of the ntiles tile sums in place
__global__ __launch_bounds__(kScanThreads) void k_scan_tile_offsets(int64_t* tile_sum,
                                                                    int64_t ntiles) {
  __shared__ int64_t sh[kScanThreads];
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < ntiles; c0 += kScanThreads) {
    const int64_t i = c0 + threadIdx.x;
    const int64_t v = i < ntiles ? tile_sum[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
      const int64_t add = int(threadIdx.x) >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < ntiles) tile_sum[i] = carry + sh[threadIdx.x] - v;
    carry += sh[kScanThreads - 1];
    __syncthreads();
  }
}

// out[i] = tile offset + exclusive sum inside the tile
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int32_t* __restrict__ in,
                                                             int64_t n,
                                                             const int64_t* __restrict__ tile_off,
                                                             int64_t* __restrict__ out) {
  __shared__ int64_t sh[kScanThreads];
  const int64_t base = int64_t(blockIdx.x) * kScanItems + int64_t(threadIdx.x) * kScanPer;
  int32_t v[kScanPer];
  int64_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    sum += v[k];
  }
  int64_t run = tile_off[blockIdx.x] + block_inclusive_scan(sum, sh) - sum;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

// ---- elements grouped by color, ascending element id inside a color ---------
// One thread walks a tile of kColorTile consecutive elements; 64 threads per
// workgroup keep their 64 per-color counters in LDS ([color][thread]).
constexpr int kColorTile = 256;
constexpr int kColorThreads = 64;

// cnt[(c - 1) * ntiles + t] = elements of color c in tile t (colors 1..64)
__global__ __launch_bounds__(kColorThreads) void k_color_tile_counts(
    const int32_t* __restrict__ color, int64_t nelem, int64_t ntiles, int32_t* __restrict__ cnt) {
  __shared__ int32_t c64[64][kColorThreads];
  const int tid = threadIdx.x;
  const int64_t t = int64_t(blockIdx.x) * kColorThreads + tid;
  for (int c = 0; c < 64; ++c) c64[c][tid] = 0;
  if (t < ntiles) {
    const int64_t e1 = std::min<int64_t>(nelem, (t + 1) * kColorTile);
    for (int64_t e = t * kColorTile; e < e1; ++e) c64[color[e] - 1][tid]++;
    for (int c = 0; c < 64; ++c) cnt[int64_t(c) * ntiles + t] = c64[c][tid];
  }
}

// each tile writes its elements in element order from its color-major
// offsets (off = exclusive sums of cnt)
__global__ __launch_bounds__(kColorThreads) void k_color_tile_scatter(
    const int32_t* __restrict__ color, int64_t nelem, int64_t ntiles,
    const int64_t* __restrict__ off, int32_t* __restrict__ out) {
  __shared__ int64_t pos[64][kColorThreads];
  const int tid = threadIdx.x;
  const int64_t t = int64_t(blockIdx.x) * kColorThreads + tid;
  if (t >= ntiles) return;
  for (int c = 0; c < 64; ++c) pos[c][tid] = off[int64_t(c) * ntiles + t];
  const int64_t e1 = std::min<int64_t>(nelem, (t + 1) * kColorTile);
  for (int64_t e = t * kColorTile; e < e1; ++e) out[pos[color[e] - 1][tid]++] = int32_t(e);
}

// device allocations released on every exit path
struct DevScratch {
  std::vector<void*> ptrs;
  template <class T>
  hipError_t alloc(T** p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count > 0 ? count * sizeof(T) : 1);
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
  ~DevScratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

}  // namespace

int build_symbolic_device(hipStream_t s, const int32_t* d_conn, int64_t nelem, int nen,
                          int64_t nnodes, SymbolicDevice* out, std::string* err) {
  DevScratch tmp;
  hipError_t e = hipSuccess;
  auto hip_fail = [&](const char* what) {
    *err = std::string(what) + ": " + hipGetErrorString(e);
    return GF_ERR_HIP;
  };
#define TRY(expr)                              \
  if ((e = (expr)) != hipSuccess) return hip_fail(#expr)
  const int64_t total = nelem * nen;
  int32_t *d_cnt = nullptr, *d_cursor = nullptr, *d_n2e = nullptr, *d_flags = nullptr;
  int64_t* d_n2e_ptr = nullptr;
  TRY(tmp.alloc(&d_flags, 4));  // [0] error code, [1] max count, [2] pending
  TRY(tmp.alloc(&d_cnt, size_t(nnodes) + 1));
  TRY(tmp.alloc(&d_cursor, size_t(nnodes)));
  TRY(tmp.alloc(&d_n2e_ptr, size_t(nnodes) + 1));
  TRY(tmp.alloc(&d_n2e, size_t(total)));
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_flags, int64_t(4), 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nnodes + 1, 256, 16384)), dim3(256), 0, s, d_cnt,
                     nnodes + 1, 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nnodes, 256, 16384)), dim3(256), 0, s, d_cursor,
                     nnodes, 0);
  hipLaunchKernelGGL(k_incidence_count, dim3(grid_for(total, 256, 16384)), dim3(256), 0, s,
                     d_conn, total, nnodes, d_cnt, d_flags);
  TRY(hipGetLastError());
  // exclusive sums (over nnodes + 1 entries, the last count 0: the total)
  auto exclusive_sum = [&](const int32_t* in, int64_t* outp, int64_t n) -> hipError_t {
    if (n <= 0) return hipSuccess;
    const int64_t ntiles = (n + kScanItems - 1) / kScanItems;
    int64_t* tile_sum = nullptr;
    hipError_t r = tmp.alloc(&tile_sum, size_t(ntiles));
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(unsigned(ntiles)), dim3(kScanThreads), 0, s, in, n,
                       tile_sum);
    hipLaunchKernelGGL(k_scan_tile_offsets, dim3(1), dim3(kScanThreads), 0, s, tile_sum, ntiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(ntiles)), dim3(kScanThreads), 0, s, in, n,
                       tile_sum, outp);
    return hipGetLastError();
  };
  TRY(exclusive_sum(d_cnt, d_n2e_ptr, nnodes + 1));
  int32_t h_flags[4] = {0, 0, 0, 0};
  TRY(hipMemcpyAsync(h_flags, d_flags, sizeof(h_flags), hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  if (h_flags[0] == 1) {
    *err = "connectivity references a node outside 1.." + std::to_string(nnodes);
    return GF_ERR_ARG;
  }
  hipLaunchKernelGGL(k_incidence_fill, dim3(grid_for(total, 256, 16384)), dim3(256), 0, s, d_conn,
                     nelem, nen, d_n2e_ptr, d_cursor, d_n2e);
  TRY(hipGetLastError());

  // node adjacency
  const int gnodes = int(std::min<int64_t>(std::max<int64_t>(nnodes, 1), 65536));
  hipLaunchKernelGGL(k_adjacency<false>, dim3(gnodes), dim3(64), 0, s, d_conn, nen, d_n2e_ptr,
                     d_n2e, nnodes, d_cnt, nullptr, nullptr, d_flags + 1, d_flags);
  TRY(hipGetLastError());
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_cnt + nnodes, int64_t(1), 0);
  TRY(hipMalloc(&out->d_adj_ptr, (size_t(nnodes) + 1) * sizeof(int64_t)));
  TRY(exclusive_sum(d_cnt, out->d_adj_ptr, nnodes + 1));
  int64_t nadj = 0;
  TRY(hipMemcpyAsync(&nadj, out->d_adj_ptr + nnodes, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  TRY(hipMemcpyAsync(h_flags, d_flags, sizeof(h_flags), hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  if (h_flags[0] == 2) {
    *err = "a node has more than " + std::to_string(kMaxCand) + " candidate neighbours";
    return GF_ERR_UNSUPPORTED;
  }
  out->max_adj = h_flags[1];
  out->nadj = nadj;
  if (out->max_adj > 255) {
    *err = "a node has " + std::to_string(out->max_adj) +
           " neighbours; the slot map stores ranks in 8 bits";
    return GF_ERR_UNSUPPORTED;
  }
  TRY(hipMalloc(&out->d_adj, size_t(nadj) * sizeof(int32_t)));
  hipLaunchKernelGGL(k_adjacency<true>, dim3(gnodes), dim3(64), 0, s, d_conn, nen, d_n2e_ptr,
                     d_n2e, nnodes, nullptr, out->d_adj_ptr, out->d_adj, d_flags + 1, d_flags);
  TRY(hipGetLastError());
  TRY(hipMalloc(&out->d_slot, size_t(total) * size_t(nen)));
  if (total > 0)
    hipLaunchKernelGGL(k_slot_map, dim3(grid_for(total, 256, 16384)), dim3(256), 0, s, d_conn,
                       nelem, nen, out->d_adj_ptr, out->d_adj, out->d_slot);
  TRY(hipGetLastError());

  // first-fit coloring in rounds
  int32_t* d_color = nullptr;
  TRY(tmp.alloc(&d_color, size_t(nelem)));
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nelem, 256, 16384)), dim3(256), 0, s, d_color,
                     nelem, 0);
  const int gel = grid_for(nelem, 256, 16384);
  int64_t rounds = 0;
  for (;;) {
    for (int r = 0; r < kRoundsPerCheck; ++r) {
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_flags + 2, int64_t(1), 0);
      hipLaunchKernelGGL(k_color_round, dim3(gel), dim3(256), 0, s, d_conn, nelem, nen, d_n2e_ptr,
                         d_n2e, d_color, d_flags + 2, d_flags);
    }
    TRY(hipGetLastError());
    rounds += kRoundsPerCheck;
    TRY(hipMemcpyAsync(h_flags, d_flags, sizeof(h_flags), hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    if (h_flags[0] == 3) {
      *err = "first-fit coloring needs more than 64 colors";
      return GF_ERR_UNSUPPORTED;
    }
    if (h_flags[2] == 0) break;
    if (rounds > nelem + kRoundsPerCheck) {  // cannot happen: every round colors one element
      *err = "coloring did not finish";
      return GF_ERR_HIP;
    }
  }
  out->color_rounds = rounds;
  // elements grouped by color, ascending element id inside a color (stable sort)
  out->elem_color.resize(size_t(nelem));
  if (nelem > 0)
    TRY(hipMemcpyAsync(out->elem_color.data(), d_color, size_t(nelem) * sizeof(int32_t),
                       hipMemcpyDeviceToHost, s));
  TRY(hipMalloc(&out->d_color_elems, size_t(nelem > 0 ? nelem : 1) * sizeof(int32_t)));
  if (nelem > 0) {
    const int64_t ntiles = (nelem + kColorTile - 1) / kColorTile;
    int32_t* d_tcnt = nullptr;
    int64_t* d_toff = nullptr;
    TRY(tmp.alloc(&d_tcnt, size_t(64 * ntiles)));
    TRY(tmp.alloc(&d_toff, size_t(64 * ntiles)));
    const unsigned gt = unsigned((ntiles + kColorThreads - 1) / kColorThreads);
    hipLaunchKernelGGL(k_color_tile_counts, dim3(gt), dim3(kColorThreads), 0, s, d_color, nelem,
                       ntiles, d_tcnt);
    TRY(hipGetLastError());
    TRY(exclusive_sum(d_tcnt, d_toff, 64 * ntiles));
    hipLaunchKernelGGL(k_color_tile_scatter, dim3(gt), dim3(kColorThreads), 0, s, d_color, nelem,
                       ntiles, d_toff, out->d_color_elems);
    TRY(hipGetLastError());
  }
  TRY(hipStreamSynchronize(s));
#undef TRY
  int ncolors = 0;
  for (int32_t c : out->elem_color) ncolors = std::max(ncolors, int(c));
  out->ncolors = ncolors;
  out->color_ptr.assign(size_t(ncolors) + 1, 0);
  for (int32_t c : out->elem_color) out->color_ptr[size_t(c)]++;
  for (int c = 0; c < ncolors; ++c) out->color_ptr[size_t(c) + 1] += out->color_ptr[size_t(c)];
  return GF_OK;
}

}  // namespace gf
