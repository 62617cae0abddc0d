// Internal definitions shared by the host and device translation units of
// libgpufem.  Nothing here is part of the C-ABI (include/gpufem.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gpufem.h"

namespace gf {

// ---- element catalogue ----------------------------------------------------
struct ElemInfo {
  int type;  // GF_Q4 ..
  int dim;
  int nen;
};

bool elem_info(int type, ElemInfo* out);
// Natural coordinates of the element nodes, [nen][dim], in the reference
// ordering (the points where the shape functions of
// src/FEM/StiffAlgo.cc:47-202 equal one).
const double* elem_natural_coords(int type);
// dN_a/dxi_i at natural point xi: dN[i*nen + a].
void shape_derivs(int type, const double* xi, double* dN);
// Tensor Gauss rule (first coordinate slowest, as src/FEM/FemData.cc:255-313).
int gauss_rule(int dim, int ngauss, double* pts, double* w);

// Elastic coefficients per unit E.  2D plane stress: c[0]=D11=D22, c[1]=D12,
// c[2]=D33.  3D isotropic: c[0]=normal diagonal, c[1]=normal off-diagonal,
// c[2]=shear diagonal.
void elastic_coeffs(int dim, double nu, double c[3]);

// Number of host threads used for the symbolic phase.
int host_threads();

}  // namespace gf

struct gf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  int num_cu = 0;
  int lds_per_cu = 0;  // bytes of LDS per compute unit
  std::string name;
  std::string err;
  // reduction scratch: per-block partial sums (doubles)
  double* d_partials = nullptr;
  int64_t partials_cap = 0;
  // small solver state blocks
  void* d_state = nullptr;
  int64_t state_cap = 0;
};

struct gf_mesh {
  int type = 0, dim = 0, nen = 0, ngauss = 0;
  int64_t nnodes = 0, nelem = 0;
  double E = 0, nu = 0;
  std::vector<int32_t> h_conn;  // kept for the symbolic phase
  int32_t* d_conn = nullptr;
  double* d_coords = nullptr;
  double* d_E = nullptr;  // may stay null (uniform E)
  int dev = 0;
};

struct gf_ell {
  int64_t n = 0;
  int32_t width = 0;
  int64_t nnz = 0;
  int dim = 1;
  int dev = 0;
  double* d_val = nullptr;
  int32_t* d_col = nullptr;
  int32_t* d_rownnz = nullptr;
  int32_t* d_diag = nullptr;  // slot of the diagonal entry per row (-1: none)
  int32_t* d_coln = nullptr;  // node columns for the node-block SpMV (lazily built)
  // block-sliced copy for the kBsell SpMV variants (solve.cpp; lazily built)
  int32_t* d_bperm = nullptr;
  int32_t* d_binv = nullptr;  // node row -> its position in d_bperm
  int64_t* d_bsptr = nullptr;
  int64_t bslices = 0, bslots = 0;
  // node rows where block-sliced windows restart (a grouped batch: the first
  // node of every mesh, then the node count), and the first slice of each
  std::vector<int64_t> bsell_breaks, bsell_break_slice;
  // The node gather, once the block-sliced copy exists, writes that copy and
  // only the diagonal of the ELL values: then ell_stale is set, the copy is
  // current (bval_version == val_version), d_val holds a current diagonal
  // and stale off-diagonal slots.  ensure_ell_vals materialises d_val from
  // the copy for every reader of the other slots.
  bool ell_stale = false;
  int32_t* d_bcol = nullptr;
  int16_t* d_bcoff = nullptr;  // 16-bit node-column offsets (null: some did not fit)
  double* d_bval = nullptr;
  float* d_bval32 = nullptr;
  // values change -> val_version++; the copies remember the version they hold
  uint64_t val_version = 1, bval_version = 0, bval32_version = 0;
  // assembly plan
  int32_t ncolors = 0;
  std::vector<int64_t> color_ptr;     // [ncolors+1] into d_color_elems
  int32_t* d_color_elems = nullptr;   // element ids grouped by color
  uint8_t* d_slot = nullptr;          // [nelem][nen][nen] node rank of b in adj(a)
  // node -> (element << 5 | local node), ascending element (node-gather
  // assembly; lazily built)
  int64_t* d_n2e_ptr = nullptr;
  int32_t* d_n2e = nullptr;
  // J^-1 and w detJ E of every element at its Gauss points (assembly
  // variant 4; lazily built)
  double* d_egrad = nullptr;
  double* d_etens = nullptr;  // element tensors of the matrix-core assembly (variant 2)
  std::vector<int32_t> h_elem_color;  // [nelem], 1-based colors
  int64_t nelem = 0;
  int nen = 0;
  // assembly kernel variant (0 element-major, 1 pair-major; -1: time both
  // on the first assembly of a large mesh)
  int asm_variant = -1;
  // mixed precision copy (lazily built)
  float* d_val32 = nullptr;
  double* d_dinv = nullptr;
  bool dinv_valid = false;
  bool val32_valid = false;
  // solver workspace, created on the first solve (solve.cpp)
  struct gf_solve_plan* plan = nullptr;
};

#define GF_HIP_CHECK(ctx, expr)                                              \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      return gf::set_error((ctx), GF_ERR_HIP,                                \
                           std::string(#expr) + ": " + hipGetErrorString(e_)); \
    }                                                                        \
  } while (0)

namespace gf {
int set_error(gf_ctx* ctx, int code, const std::string& msg);
int set_error(const gf_ctx* ctx, int code, const std::string& msg);
}  // namespace gf

#include <thread>
namespace gf {
// Runs f(begin, end) over [0, n) on up to `threads` std::threads.
template <class F>
void parallel_for(int64_t n, int threads, F&& f) {
  if (n <= 0) return;
  if (threads < 1) threads = 1;
  if (n < 4096 || threads == 1) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int64_t b = t * chunk, e = b + chunk < n ? b + chunk : n;
    if (b >= e) break;
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}
}  // namespace gf
