// Host-callable launchers for the kernels in kernels.hip and solver.hip.
// Every launcher enqueues on the given stream and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gf {

// ---- element tables --------------------------------------------------------
// Uploads the per-(type, rule) dN/dxi and weight tables to the current device.
hipError_t upload_element_tables();

// ---- symbolic expansion ------------------------------------------------------
hipError_t launch_ell_expand(hipStream_t s, int64_t n, int32_t width, int dim,
                             const int64_t* adj_ptr, const int32_t* adj, double* val,
                             int32_t* col, int32_t* rownnz, int32_t* diag);

// ---- numeric assembly --------------------------------------------------------
struct AssemblyArgs {
  int type, ngauss;
  const int32_t* conn;
  const double* coords;
  const double* elemE;  // may be null
  double E;
  double c[3];          // elastic coefficients per unit E
  const int32_t* elems; // element ids of one color
  int64_t nce;          // number of elements in the color
  const uint8_t* slot;
  double* val;
  int64_t n;
  int* bad_jac;         // set to 1 when some detJ <= 0
  int variant;          // 0: element-major threads, 1: pair-major threads,
                        // 2: FP64 matrix cores (3D elements), 3: node gather
  // node gather: node -> (element << 5 | local node) lists, ascending element
  const int64_t* n2e_ptr;
  const int32_t* n2e;
  const int32_t* rownnz;
  int64_t nnodes;
  // node gather: also write the block-sliced copy of the values (null: no
  // copy); bsptr = its slice pointers, binv = node row -> slice position
  double* bval;
  const int64_t* bsptr;
  const int32_t* binv;
  // node gather with the copy: ell_diag_only = 1 writes only the diagonal
  // (slot diag[row]) of the ELL values
  const int32_t* diag;
  int ell_diag_only;
  // the element count (the first kernels of variants 2 and 4)
  int64_t nelem;
};
hipError_t launch_assembly_color(hipStream_t s, const AssemblyArgs& a);
// Variant 4: the Gauss-point geometry of every element once into geom (a
// buffer of gauss_geom_doubles(type, ngauss, nelem) doubles), then the node
// gather reading it instead of forming it per node; where the gather fits.
int64_t gauss_geom_doubles(int type, int ngauss, int64_t nelem);
hipError_t launch_assembly_geom(hipStream_t s, const AssemblyArgs& a, double* geom);
// Node-gather assembly (variant 3): one launch for the whole mesh, no
// coloring.  Each thread owns rows of one node, computes the rows of Ke of
// every element around the node and adds them in LDS, then writes its rows
// of the ELL values once (every live slot).  Applicable when the node has at
// most gather_max_slots(type) neighbours (0: not available for the type).
// Variant 2 (3D): element tensors on the matrix cores into tens (a buffer of
// elem_tensor_doubles(type, nelem) doubles), then the node gather over them.
int64_t elem_tensor_doubles(int type, int64_t nelem);
hipError_t launch_assembly_tensors(hipStream_t s, const AssemblyArgs& a, double* tens);
int gather_max_slots(int type);
hipError_t launch_assembly_gather(hipStream_t s, const AssemblyArgs& a);

// ---- penalty -------------------------------------------------------------------
// result[0] <- coef * max_i K_ii (0 for rows with no stored diagonal)
hipError_t launch_penalty_factor(hipStream_t s, const double* val, const int32_t* diag,
                                 int64_t n, double coef, double* partials, int nparts,
                                 double* result);
// K_dd <- pen^count_d * K_dd (applied as count repeated multiplications);
// with bval, the same value is stored into the block-sliced copy (d x d
// blocks, slice pointers bsptr, node row -> slice position binv)
hipError_t launch_penalty_apply(hipStream_t s, double* val, const int32_t* diag, int64_t n,
                                const int32_t* dofs, const int32_t* counts, int64_t nd,
                                const double* pen, double* bval, const int64_t* bsptr,
                                const int32_t* binv, int d);

// ---- zero substitution, right-hand sides, force vector --------------------------
// fixed[d] = 1 for the listed DOFs (fixed zeroed by the caller); then every
// stored entry in a fixed row or column becomes 0, the diagonal 1
hipError_t launch_eliminate(hipStream_t s, int64_t n, int32_t width, const int32_t* dofs,
                            int64_t nd, uint8_t* fixed, const int32_t* col,
                            const int32_t* rownnz, double* val);
// v[k*n + d] = 0 for the listed DOFs, k < nrhs
hipError_t launch_zero_dofs(hipStream_t s, double* v, int64_t n, int32_t nrhs,
                            const int32_t* dofs, int64_t nd);
// f = 0, then the nonzero components of load rows [node, F_1..F_dim] are
// assigned, the last row naming a DOF winning (owner: n int32 scratch)
hipError_t launch_load_vector(hipStream_t s, int dim, int64_t n, const double* loads,
                              int64_t nload, int32_t* owner, double* f);

// ---- SpMV / solver -----------------------------------------------------------------
// SpMV variants (see solver.hip)
// kNode2 / kNode3: thread per row over node blocks of a 2D / 3D FEM matrix,
// reading one node column per d slots (the "col" argument is then the node
// column array of gf_ell::d_coln) and the d entries of x of that node together.
// kBsell2 / kBsell3: thread per node row of the block-sliced copy (BsellView;
// "val"/"col" are then gf_ell::d_bval / d_bcol): d output rows per thread,
// one node column and one gather of the node's d x entries per d x d block;
// the PCG grid sweeps the slices.
// kBsell2G / kBsell3G: the same rows, one group of four slices (one per wave)
// per workgroup and as many workgroups as groups, so that the hardware
// dispatcher balances slices of unequal length.
enum SpmvVariant {
  kThreadU1 = 0, kThreadU4 = 1, kThreadU8 = 2, kTile2 = 3, kTile4 = 4, kNode2 = 5, kNode3 = 6,
  kBsell2 = 7, kBsell3 = 8, kBsell2G = 9, kBsell3G = 10
};
constexpr int kNumSpmvVariants = 11;
inline bool is_node_variant(int v) { return v == kNode2 || v == kNode3; }
inline bool is_bsell_group_variant(int v) { return v == kBsell2G || v == kBsell3G; }
inline bool is_bsell_variant(int v) {
  return v == kBsell2 || v == kBsell3 || is_bsell_group_variant(v);
}
// node columns: coln[i + j n] = col[i + d j n] / d for j < width / d
hipError_t launch_build_coln(hipStream_t s, int64_t n, int32_t width, int d, const int32_t* col,
                             int32_t* coln);

// Block-sliced ELL ("BSELL"), a solver-side copy of a d-dimensional FEM
// matrix.  Node rows are sorted by length (descending, stable) inside windows
// of kBsellWindow node rows and dealt into slices of 64 positions (one
// wavefront); slice s has sptr[s+1] - sptr[s] block slots, its longest row.
// Block slot b, lane l: node column bcol[b*64 + l]; block value k = r d + e,
// K(d R + r, d c + e) for the lane's node row R = perm[s*64 + l] (-1: empty
// position), at bval[bsell_vidx(d*d, b, k, l)]: the d*d*64 values of a slot
// hold the pairs (k, k+1), k even, lane-interleaved (one 16-byte load per
// lane and pair), then the odd last value of 3 x 3 blocks lane by lane.
// Slots are in ascending column order, so every output row sums its entries
// in the order of the ELL slots.
__host__ __device__ inline int64_t bsell_vidx(int dd, int64_t b, int k, int l) {
  const int64_t base = b * dd * 64;
  return k < (dd & ~1) ? base + ((k >> 1) * 64 + l) * 2 + (k & 1) : base + (dd - 1) * 64 + l;
}
// The FP32 copy of the block values: quads (k .. k+3, k a multiple of 4)
// lane-interleaved (one 16-byte load per lane and quad), then the ninth value
// of 3 x 3 blocks lane by lane.
__host__ __device__ inline int64_t bsell_vidx32(int dd, int64_t b, int k, int l) {
  const int64_t base = b * dd * 64;
  return k < (dd & ~3) ? base + ((k >> 2) * 64 + l) * 4 + (k & 3) : base + int64_t(k) * 64 + l;
}
// the layout of block values of type T (double: pairs, float: quads)
template <class T>
__host__ __device__ inline int64_t bsell_index(int dd, int64_t b, int k, int l) {
  if constexpr (sizeof(T) == 4) return bsell_vidx32(dd, b, k, l);
  else return bsell_vidx(dd, b, k, l);
}
struct BsellView {
  const int32_t* perm;
  const int64_t* sptr;
  int64_t nslices;
  // node column - node row per block slot and lane (padding 0), when every
  // offset of the matrix fits 16 bits; null: the single-RHS kernels read bcol
  const int16_t* coff;
};
constexpr int kBsellWindow = 2048;
hipError_t launch_bsell_cols(hipStream_t s, const BsellView& bs, int d, int64_t n,
                             const int32_t* col, const int32_t* rownnz, int32_t* bcol);
// coff[b*64 + l] = bcol[b*64 + l] - max(R, 0); *overflow = 1 when one does not fit
hipError_t launch_bsell_offsets(hipStream_t s, const BsellView& bs, const int32_t* bcol,
                                int16_t* coff, int* overflow);
hipError_t launch_bsell_vals(hipStream_t s, const BsellView& bs, int d, int64_t n,
                             const double* val, const int32_t* rownnz, double* bval);
// ELL values of the live slots from the block-sliced copy (the inverse)
hipError_t launch_bsell_to_ell(hipStream_t s, const BsellView& bs, int d, int64_t n,
                               const double* bval, const int32_t* rownnz, double* val);
constexpr int kMaxPcgBlocks = 2048;
// SpMV workgroups of the slice-group variants (one partial p.q each)
constexpr int kMaxSpmvBlocks = 65536;

// workgroups of the standalone SpMV and the PCG kernels for n rows
int spmv_blocks(int64_t n);
// workgroups of the SpMV kernels of `variant`: one per four-slice group for
// the slice-group variants, else `nblocks`
int spmv_grid(int variant, const BsellView& bs, int nblocks);
// bs: the block-sliced view (used by the kBsell variants only; may be null)
hipError_t launch_spmv(hipStream_t s, int variant, int64_t n, const double* val,
                       const int32_t* col, const int32_t* rownnz, const double* x, double* y,
                       int nrhs, const BsellView* bs = nullptr);
// plain copy of `bytes` (a multiple of 16, 16-byte aligned pointers)
hipError_t launch_stream_copy(hipStream_t s, const void* src, void* dst, int64_t bytes);
hipError_t launch_jacobi_inv(hipStream_t s, int64_t n, const double* val,
                             const int32_t* diag, double* dinv);

struct PcgState;  // device-resident scalars (solver.hip)

struct PcgBuffers {
  int64_t n;
  int32_t width;
  int variant;
  const double* val;
  const int32_t* col;
  const int32_t* rownnz;
  const double* dinv;  // null: no preconditioner (reference CG)
  const double* b;
  double* x;
  double* r;
  double* z;
  double* p;
  double* q;
  double* part1;  // per-SpMV-workgroup partials of p.q
  double* part2;  // per-block partials of (r.z, r.r[, b.b])
  int nblocks;    // workgroups of every PCG kernel but the SpMV (<= kMaxPcgBlocks)
  int spmv_grid;  // workgroups of the SpMV kernel (<= kMaxSpmvBlocks)
  PcgState* st;
  BsellView bs;   // block-sliced view (kBsell variants)
};
size_t pcg_state_bytes();
size_t grid_barrier_bytes();
hipError_t launch_pcg_setup(hipStream_t s, const PcgBuffers& B, int use_x0, double rtol,
                            int maxit);
hipError_t launch_pcg_iteration(hipStream_t s, const PcgBuffers& B, int parity);
// the same, with events recorded around the SpMV kernel
hipError_t launch_pcg_iteration_timed(hipStream_t s, const PcgBuffers& B, int parity,
                                      hipEvent_t spmv_start, hipEvent_t spmv_stop);
// co-resident workgroups for the persistent kernel of this variant (0: unavailable)
int persistent_blocks(int variant, int num_cu);
// nbar grid barriers of the persistent schedule on a cooperative grid of
// nblocks workgroups; barrier_aborted reads whether a wait hit its spin limit
hipError_t launch_barrier_probe(hipStream_t s, void* barrier, int nblocks, int nbar);
int barrier_aborted(hipStream_t s, const void* barrier, int* aborted);
// all remaining iterations in one cooperative launch
hipError_t launch_pcg_persistent(hipStream_t s, const PcgBuffers& B, void* barrier);
// Host view of the scalar state.
struct PcgHostState {
  int32_t it, done, status;
  int32_t x0_rejected;  // set by the host: a guarded initial guess was replaced by zero
  double rr, bnorm2, rho;
};
hipError_t read_pcg_state(hipStream_t s, const PcgBuffers& B, PcgHostState* h);
// nrep launches of the iteration's SpMV kernel with the state forced live
hipError_t launch_spmv_pq_repeat(hipStream_t s, const PcgBuffers& B, int nrep);
// true residual ||b - K x|| / ||b|| into *out (device), deterministic
hipError_t launch_true_residual(hipStream_t s, const PcgBuffers& B, double* out_dev);
// *out_dev = 1/2 x.Kx - b.x (B.q is overwritten with K x)
hipError_t launch_energy(hipStream_t s, const PcgBuffers& B, double* out_dev);

}  // namespace gf
