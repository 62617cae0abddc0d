// Device kernels for the symbolic expansion, the numeric assembly (Ke by
// Gauss quadrature fused with the colored scatter) and the penalty
// constraint.  gfx950 (CDNA4): 64-wide wavefronts, 256-thread workgroups.
#include <cfloat>

#include "gf_internal.hpp"
#include "launch.hpp"

namespace gf {
namespace {

// ---------------------------------------------------------------------------
// Element tables: dN/dxi at every Gauss point and the Gauss weights, per
// (element type, points per direction).  Block (type-1)*2 + (ngauss-2) starts
// at c_off[block]: ngp*dim*nen derivatives (dN[(g*dim + i)*nen + a]) followed
// by ngp weights.
// ---------------------------------------------------------------------------
constexpr int kTabCap = 3400;
__constant__ double c_tab[kTabCap];
int h_off[8];
int h_ngp[8];

inline int tab_block(int type, int ngauss) { return (type - 1) * 2 + (ngauss - 2); }

template <int DIM>
__device__ inline double det_inv(const double (&J)[DIM][DIM], double (&Ji)[DIM][DIM]) {
  if constexpr (DIM == 2) {
    double d = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double id = 1.0 / d;  // one division, four products
    Ji[0][0] = J[1][1] * id;
    Ji[0][1] = -J[0][1] * id;
    Ji[1][0] = -J[1][0] * id;
    Ji[1][1] = J[0][0] * id;
    return d;
  } else {
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double d = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / d;  // one division, nine products
    Ji[0][0] = c00 * id;
    Ji[1][0] = c01 * id;
    Ji[2][0] = c02 * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    return d;
  }
}

struct KArgs {
  const int32_t* conn;
  const double* coords;
  const double* elemE;
  double E;
  double c0, c1, c2;
  const int32_t* elems;
  int64_t nce;
  const uint8_t* slot;
  double* val;
  int64_t n;
  int* bad;
  int tab;  // offset of the (type, rule) block in c_tab
};

// One workgroup handles EPB elements of one color.
//  phase 0: element node ids and coordinates -> LDS
//  phase 1: one thread per (element, Gauss point): J, detJ, J^-1, dN/dx -> LDS
//  phase 2: one thread per (element, node pair a-b): the dim x dim block
//           K_ab = sum_g G_a^T D G_b w_g detJ_g E_e, accumulated in registers
//  phase 3: the same thread adds its block into the ELL values.  Elements of
//           one color share no node, so the read-modify-write is race-free.
template <int DIM, int NEN, int NGP>
__global__ __launch_bounds__(256) void k_assemble(KArgs a) {
  constexpr int PAIRS = NEN * NEN;
  constexpr int EPB = PAIRS >= 256 ? 1 : 256 / PAIRS;
  __shared__ double s_dndx[EPB][NGP][DIM][NEN];
  __shared__ double s_sc[EPB][NGP];
  __shared__ double s_x[EPB][DIM][NEN];
  __shared__ int32_t s_node[EPB][NEN];
  __shared__ int32_t s_elem[EPB];
  const int tid = threadIdx.x;
  const int64_t base = int64_t(blockIdx.x) * EPB;

  for (int t = tid; t < EPB * NEN; t += 256) {
    const int el = t / NEN, an = t % NEN;
    if (base + el < a.nce) {
      const int32_t e = a.elems[base + el];
      if (an == 0) s_elem[el] = e;
      const int32_t node = a.conn[int64_t(e) * NEN + an] - 1;
      s_node[el][an] = node;
#pragma unroll
      for (int d = 0; d < DIM; ++d) s_x[el][d][an] = a.coords[int64_t(node) * DIM + d];
    }
  }
  __syncthreads();

  for (int t = tid; t < EPB * NGP; t += 256) {
    const int el = t / NGP, g = t % NGP;
    if (base + el >= a.nce) continue;
    const double* dN = c_tab + a.tab + g * DIM * NEN;
    double J[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int n = 0; n < NEN; ++n) acc += dN[i * NEN + n] * s_x[el][j][n];
        J[i][j] = acc;
      }
    double Ji[DIM][DIM];
    const double det = det_inv<DIM>(J, Ji);
    if (!(det > 0.0)) *a.bad = 1;
#pragma unroll
    for (int n = 0; n < NEN; ++n)
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) acc += Ji[i][k] * dN[k * NEN + n];
        s_dndx[el][g][i][n] = acc;
      }
    const double w = c_tab[a.tab + NGP * DIM * NEN + g];
    const double Ee = a.elemE ? a.elemE[s_elem[el]] : a.E;
    s_sc[el][g] = w * det * Ee;
  }
  __syncthreads();

  for (int t = tid; t < EPB * PAIRS; t += 256) {
    const int el = t / PAIRS, p = t % PAIRS;
    if (base + el >= a.nce) continue;
    const int an = p / NEN, bn = p % NEN;
    double k[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) k[i][j] = 0.0;
    for (int g = 0; g < NGP; ++g) {
      const double sc = s_sc[el][g];
      double ga[DIM], gb[DIM];
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        ga[i] = s_dndx[el][g][i][an];
        gb[i] = s_dndx[el][g][i][bn];
      }
      if constexpr (DIM == 2) {
        // plane stress, D = [[c0, c1, 0], [c1, c0, 0], [0, 0, c2]]
        k[0][0] += sc * (a.c0 * ga[0] * gb[0] + a.c2 * ga[1] * gb[1]);
        k[0][1] += sc * (a.c1 * ga[0] * gb[1] + a.c2 * ga[1] * gb[0]);
        k[1][0] += sc * (a.c1 * ga[1] * gb[0] + a.c2 * ga[0] * gb[1]);
        k[1][1] += sc * (a.c0 * ga[1] * gb[1] + a.c2 * ga[0] * gb[0]);
      } else {
        // isotropic 3D: normal diagonal c0, normal off-diagonal c1, shear c2
        const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            if (i == j)
              k[i][i] += sc * (a.c0 * ga[i] * gb[i] + a.c2 * (dot - ga[i] * gb[i]));
            else
              k[i][j] += sc * (a.c1 * ga[i] * gb[j] + a.c2 * ga[j] * gb[i]);
          }
      }
    }
    const int32_t e = s_elem[el];
    const int64_t slot = a.slot[(int64_t(e) * NEN + an) * NEN + bn];
    const int64_t row0 = int64_t(s_node[el][an]) * DIM;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double* v = a.val + (row0 + i) + (slot * DIM + j) * a.n;
        *v += k[i][j];
      }
  }
}

// Pair-major variant: EPB elements per workgroup and consecutive threads on
// the same node pair of consecutive elements of the color, so the
// read-modify-write of one pair touches neighbouring rows together; the LDS
// tables are element-fastest ([g][i][node][element]) so those threads also
// read consecutive words.  Same arithmetic, same order: bitwise equal to
// k_assemble.
template <int DIM, int NEN, int NGP, int EPB>
__global__ __launch_bounds__(256) void k_assemble_pm(KArgs a) {
  constexpr int PAIRS = NEN * NEN;
  __shared__ double s_dndx[NGP][DIM][NEN][EPB];
  __shared__ double s_sc[NGP][EPB];
  __shared__ double s_x[DIM][NEN][EPB];
  __shared__ int32_t s_node[NEN][EPB];
  __shared__ int32_t s_elem[EPB];
  const int tid = threadIdx.x;
  const int64_t base = int64_t(blockIdx.x) * EPB;
  const int nel = int(a.nce - base < EPB ? a.nce - base : EPB);

  for (int t = tid; t < EPB * NEN; t += 256) {
    const int el = t % EPB, an = t / EPB;
    if (el < nel) {
      const int32_t e = a.elems[base + el];
      if (an == 0) s_elem[el] = e;
      const int32_t node = a.conn[int64_t(e) * NEN + an] - 1;
      s_node[an][el] = node;
#pragma unroll
      for (int d = 0; d < DIM; ++d) s_x[d][an][el] = a.coords[int64_t(node) * DIM + d];
    }
  }
  __syncthreads();

  for (int t = tid; t < EPB * NGP; t += 256) {
    const int el = t % EPB, g = t / EPB;
    if (el >= nel) continue;
    const double* dN = c_tab + a.tab + g * DIM * NEN;
    double J[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int n = 0; n < NEN; ++n) acc += dN[i * NEN + n] * s_x[j][n][el];
        J[i][j] = acc;
      }
    double Ji[DIM][DIM];
    const double det = det_inv<DIM>(J, Ji);
    if (!(det > 0.0)) *a.bad = 1;
#pragma unroll
    for (int n = 0; n < NEN; ++n)
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) acc += Ji[i][k] * dN[k * NEN + n];
        s_dndx[g][i][n][el] = acc;
      }
    const double w = c_tab[a.tab + NGP * DIM * NEN + g];
    const double Ee = a.elemE ? a.elemE[s_elem[el]] : a.E;
    s_sc[g][el] = w * det * Ee;
  }
  __syncthreads();

  for (int t = tid; t < EPB * PAIRS; t += 256) {
    const int el = t % EPB, p = t / EPB;
    if (el >= nel) continue;
    const int an = p / NEN, bn = p % NEN;
    double k[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) k[i][j] = 0.0;
    for (int g = 0; g < NGP; ++g) {
      const double sc = s_sc[g][el];
      double ga[DIM], gb[DIM];
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        ga[i] = s_dndx[g][i][an][el];
        gb[i] = s_dndx[g][i][bn][el];
      }
      if constexpr (DIM == 2) {
        k[0][0] += sc * (a.c0 * ga[0] * gb[0] + a.c2 * ga[1] * gb[1]);
        k[0][1] += sc * (a.c1 * ga[0] * gb[1] + a.c2 * ga[1] * gb[0]);
        k[1][0] += sc * (a.c1 * ga[1] * gb[0] + a.c2 * ga[0] * gb[1]);
        k[1][1] += sc * (a.c0 * ga[1] * gb[1] + a.c2 * ga[0] * gb[0]);
      } else {
        const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            if (i == j)
              k[i][i] += sc * (a.c0 * ga[i] * gb[i] + a.c2 * (dot - ga[i] * gb[i]));
            else
              k[i][j] += sc * (a.c1 * ga[i] * gb[j] + a.c2 * ga[j] * gb[i]);
          }
      }
    }
    const int32_t e = s_elem[el];
    const int64_t slot = a.slot[(int64_t(e) * NEN + an) * NEN + bn];
    const int64_t row0 = int64_t(s_node[an][el]) * DIM;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double* v = a.val + (row0 + i) + (slot * DIM + j) * a.n;
        *v += k[i][j];
      }
  }
}

// FP64 matrix-core variant for the 3D elements: one wavefront per element
// computes Ke = sum_g B_g^T (sc_g D B_g) as a (3 nen) x (3 nen) GEMM with
// K = 6 ngp on v_mfma_f64_16x16x4f64.  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and receives D[4(l/16)+r][l%16] (r = 0..3).  The operands
// are generated from the dN/dx tables in LDS (B_g has the 6-strain sparsity
// pattern, D is isotropic); the result is scattered like k_assemble.
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Element tensors on the matrix cores (phase 1 of variant 2, 3D elements).
// With s_g = w_g detJ_g E_e and G_g the dN/dx table at Gauss point g, the
// isotropic element block of nodes a, b is
//   K_ab[i][j] = c0 T_ii + c2 (T_jj + T_kk)   (i == j)
//   K_ab[i][j] = c1 T_ij + c2 T_ji            (i != j)
// with T_pq[a][b] = sum_g s_g G_g[p][a] G_g[q][b]: nine nen x nen products
// over the Gauss points, which skip the zeros of the strain matrix B.  One
// wavefront per element forms them as 16x16x4 FP64 MFMA tiles (M = node a,
// N = node b, K = Gauss point) and writes tens[e][a][p][b][q] (the gather
// reads a chunk of neighbours b as three contiguous runs); the node gather
// then reads them instead of recomputing the geometry.
struct TArgs {
  KArgs k;
  double* tens;
  int64_t nelem;
};

template <int NEN, int NGP>
__global__ __launch_bounds__(256) void k_elem_tensors(TArgs ta) {
  constexpr int NT = (NEN + 15) / 16, KS = (NGP + 3) / 4;
  constexpr int EPB = 4;  // one wavefront per element
  __shared__ double s_g[EPB][NGP][3][NEN];
  __shared__ double s_sc[EPB][NGP];
  __shared__ double s_x[EPB][3][NEN];
  const KArgs& a = ta.k;
  const int tid = threadIdx.x;
  const int64_t base = int64_t(blockIdx.x) * EPB;
  for (int t = tid; t < EPB * NEN; t += 256) {
    const int el = t / NEN, an = t % NEN;
    if (base + el < ta.nelem) {
      const int64_t node = a.conn[(base + el) * NEN + an] - 1;
#pragma unroll
      for (int d = 0; d < 3; ++d) s_x[el][d][an] = a.coords[node * 3 + d];
    }
  }
  __syncthreads();
  for (int t = tid; t < EPB * NGP; t += 256) {
    const int el = t / NGP, g = t % NGP;
    if (base + el >= ta.nelem) continue;
    const double* dN = c_tab + a.tab + g * 3 * NEN;
    double J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double acc = 0.0;
#pragma unroll(NEN > 8 ? 4 : NEN)
        for (int n = 0; n < NEN; ++n) acc += dN[i * NEN + n] * s_x[el][j][n];
        J[i][j] = acc;
      }
    double Ji[3][3];
    const double det = det_inv<3>(J, Ji);
    if (!(det > 0.0)) *a.bad = 1;
#pragma unroll(NEN > 8 ? 4 : NEN)
    for (int n = 0; n < NEN; ++n)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) acc += Ji[i][k] * dN[k * NEN + n];
        s_g[el][g][i][n] = acc;
      }
    const double w = c_tab[a.tab + NGP * 3 * NEN + g];
    const double Ee = a.elemE ? a.elemE[base + el] : a.E;
    s_sc[el][g] = w * det * Ee;
  }
  __syncthreads();
  const int el = tid >> 6, lane = tid & 63;
  const int64_t e = base + el;
  if (e >= ta.nelem) return;  // the whole wavefront leaves; no barrier follows
  const int li = lane & 15, lk = lane >> 4;
  double* te = ta.tens + e * (9 * NEN * NEN);
  for (int p = 0; p < 3; ++p) {
    f64x4 acc[3][NT][NT];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[q][i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int ks = 0; ks < KS; ++ks) {
      // lane (li, lk): A[row li][k lk] and B[k lk][column li] of the step
      const int g = ks * 4 + lk;
      const bool gv = g < NGP;
      const double sc = gv ? s_sc[el][g] : 0.0;
      double av[NT], bv[3][NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int n = t * 16 + li;
        const bool v = gv && n < NEN;
        av[t] = v ? sc * s_g[el][g][p][n] : 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) bv[q][t] = v ? s_g[el][g][q][n] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
          for (int tj = 0; tj < NT; ++tj)
            acc[q][ti][tj] =
                __builtin_amdgcn_mfma_f64_16x16x4f64(av[ti], bv[q][tj], acc[q][ti][tj], 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // v_mfma_f64_16x16x4f64 result: lane l, register r holds
            // D[l/16 + 4r][l%16] (lane layout measured on gfx950)
            const int an = ti * 16 + lk + 4 * r, bn = tj * 16 + li;
            if (an < NEN && bn < NEN) te[((an * 3 + p) * NEN + bn) * 3 + q] = acc[q][ti][tj][r];
          }
  }
}

// ---------------------------------------------------------------------------
// Node-gather assembly.  Thread u owns RPT rows of node v = u / (DIM / RPT)
// (2D: both rows, 3D: one row, so consecutive threads own consecutive rows
// and the final writes coalesce).  For every element around v (ascending
// element id) it forms the rows of Ke that belong to v -- J, detJ, J^-1 and
// dN/dx at each Gauss point from the element's coordinates, then the
// node-pair blocks G_a^T D G_b w detJ E as in the scatter kernels -- and
// adds them into per-thread LDS accumulators indexed by ELL node slot (the
// slot map).  Each ELL entry is a sum over elements in element order; the
// values agree with the colored scatter's to rounding.
// ---------------------------------------------------------------------------
struct GArgs {
  KArgs k;
  const int64_t* n2e_ptr;
  const int32_t* n2e;
  const int32_t* rownnz;
  int64_t nnodes;
  // block-sliced copy written beside the ELL values (bval null: none)
  double* bval;
  const int64_t* bsptr;
  const int32_t* binv;
  // ell_diag_only: of the ELL values only the diagonal (slot diag[row])
  const int32_t* diag;
  int ell_diag_only;
  // element tensors (variant 2): tens[e][a][p][b][q]
  const double* tens;
  // Gauss-point geometry (variant 4): per element and Gauss point, J^-1
  // (DIM x DIM, row-major) and w detJ E
  const double* geom;
};

// A unit is RPT rows of one node (UPN units per node).  LPU lanes share a
// unit: lane j of the unit takes the node's incidences j, j + LPU, ... and the
// LPU lanes add their element's rows into the unit's LDS accumulators in LPU
// masked phases, so every accumulator receives the elements in ascending
// incidence order when the neighbours form one chunk (BCH = NEN).  With
// smaller chunks the LPU elements of a trip add chunk by chunk, and a slot
// that is neighbour b of one element and a later-chunk neighbour of an
// earlier one receives them out of order: the sums then depend on LPU in
// their last bits (the same configuration gives the same bits).  UPW units per 64-lane
// workgroup: LPU > 1 divides the LDS per workgroup by LPU and lets more
// wavefronts share a CU.
// With RPT == DIM (every row of the node in one unit) the lane forms the
// products T_pq = sum_g (w detJ E dNa/dx_p)(dNb/dx_q) per neighbour b and the
// elastic combinations once per element; the element's nodes b are taken BCH
// at a time (the Gauss-point geometry is recomputed per chunk), which bounds
// the accumulator registers.  With RPT < DIM (one row per unit) the row's
// entries are accumulated directly and BCH = NEN.  GPU: unroll factor of the
// Gauss-point loop.  SRC: 0 forms the geometry from the coordinates, 1 reads
// the element tensors, 2 reads J^-1 and w detJ E per Gauss point (formed once
// per element by k_gauss_geom with the same arithmetic, so the values equal
// SRC 0's bitwise).
template <int DIM, int NEN, int NGP, int MAXS, int RPT, int LPU, int UPW, int BCH, int GPU = 1,
          int SRC = 0>
__global__ __launch_bounds__(64) void k_assemble_gather(GArgs ga) {
  static_assert(LPU * UPW <= 64, "a workgroup is one wavefront");
  static_assert(SRC != 1 || (RPT == DIM && DIM == 3), "element tensors: 3D node units");
  static_assert(SRC != 2 || RPT == DIM, "stored geometry: node units");
  static_assert(NEN % BCH == 0 && (RPT == DIM || BCH == NEN), "neighbour chunks");
  constexpr int UPN = DIM / RPT;  // units per node
  constexpr int ACC = MAXS * RPT * DIM;
  __shared__ double s_acc[ACC * UPW];
  const KArgs& a = ga.k;
  const int lu = threadIdx.x / LPU;  // unit of this lane inside the workgroup
  const int lj = threadIdx.x % LPU;  // lane of the unit
  const int64_t units = ga.nnodes * UPN;
  // grid-stride over units; every lane runs the same trip count (LDS is per unit)
  for (int64_t u0 = int64_t(blockIdx.x) * UPW; u0 < units; u0 += int64_t(gridDim.x) * UPW) {
    const int64_t u = u0 + lu;
    const bool live = lu < UPW && u < units;
    const int64_t v = live ? u / UPN : 0;
    const int r0 = live ? int(u % UPN) * RPT : 0;
    if (lu < UPW)
      for (int q = lj; q < ACC; q += LPU) s_acc[q * UPW + lu] = 0.0;
    const int64_t k0 = live ? ga.n2e_ptr[v] : 0, k1 = live ? ga.n2e_ptr[v + 1] : 0;
    // trip count: the largest over the wavefront, so the phases below are
    // reached by every lane together
    int trips = int((k1 - k0 + LPU - 1) / LPU);
    if constexpr (LPU > 1) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) trips = max(trips, __shfl_xor(trips, off));
    }
    for (int t = 0; t < trips; ++t) {
      const int64_t kk = k0 + int64_t(t) * LPU + lj;
      if (kk >= k1) continue;  // this lane's share of the node is done
      const int32_t packed = ga.n2e[kk];
      const int64_t e = packed >> 5;
      const int la = packed & 31;
      double x[NEN][DIM];
      double Ee = 0.0;
      if constexpr (SRC == 0) {
#pragma unroll
        for (int n = 0; n < NEN; ++n) {
          const int64_t node = a.conn[e * NEN + n] - 1;
#pragma unroll
          for (int d = 0; d < DIM; ++d) x[n][d] = a.coords[node * DIM + d];
        }
        Ee = a.elemE ? a.elemE[e] : a.E;
      }
      const uint8_t* sl = a.slot + (e * NEN + la) * NEN;
      for (int c0 = 0; c0 < NEN; c0 += BCH) {
        double kl[BCH][RPT][DIM];
#pragma unroll
        for (int b = 0; b < BCH; ++b)
#pragma unroll
          for (int r = 0; r < RPT; ++r)
#pragma unroll
            for (int j = 0; j < DIM; ++j) kl[b][r][j] = 0.0;
        if constexpr (SRC == 1) {
          // the element tensors T_pq[a][b] of node a = la, b in the chunk
          const double* tt = ga.tens + (e * NEN + la) * (9 * NEN) + c0 * 3;
#pragma unroll
          for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int b = 0; b < BCH; ++b)
#pragma unroll
              for (int q = 0; q < 3; ++q) kl[b][p][q] = tt[(p * NEN + b) * 3 + q];
        }
#pragma unroll(GPU)
        for (int g = 0; g < (SRC != 1 ? NGP : 0); ++g) {
          const double* dN = c_tab + a.tab + g * DIM * NEN;
          double Ji[DIM][DIM];
          double sc;
          if constexpr (SRC == 0) {
            double J[DIM][DIM];
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
              for (int j = 0; j < DIM; ++j) {
                double acc = 0.0;
#pragma unroll
                for (int n = 0; n < NEN; ++n) acc += dN[i * NEN + n] * x[n][j];
                J[i][j] = acc;
              }
            const double det = det_inv<DIM>(J, Ji);
            if (!(det > 0.0)) *a.bad = 1;
            sc = c_tab[a.tab + NGP * DIM * NEN + g] * det * Ee;
          } else {
            const double* gp = ga.geom + (e * NGP + g) * (DIM * DIM + 1);
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
              for (int j = 0; j < DIM; ++j) Ji[i][j] = gp[i * DIM + j];
            sc = gp[DIM * DIM];
          }
          // dN_n/dx at this Gauss point
          auto grad = [&](int n, double* gv) {
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
              double acc = 0.0;
#pragma unroll
              for (int k = 0; k < DIM; ++k) acc += Ji[i][k] * dN[k * NEN + n];
              gv[i] = acc;
            }
          };
          double gA[DIM];
          grad(la, gA);
          if constexpr (RPT == DIM) {
            // every row of the node: accumulate the products T_pq; the
            // elastic combinations are formed once per element below
            double sA[DIM];
#pragma unroll
            for (int p = 0; p < DIM; ++p) sA[p] = sc * gA[p];
#pragma unroll
            for (int b = 0; b < BCH; ++b) {
              double gB[DIM];
              grad(c0 + b, gB);
#pragma unroll
              for (int p = 0; p < DIM; ++p)
#pragma unroll
                for (int q = 0; q < DIM; ++q) kl[b][p][q] += sA[p] * gB[q];
            }
          } else {
#pragma unroll
            for (int b = 0; b < BCH; ++b) {
              double gB[DIM];
              grad(c0 + b, gB);
              double dot = 0.0;
#pragma unroll
              for (int i = 0; i < DIM; ++i) dot += gA[i] * gB[i];
#pragma unroll
              for (int r = 0; r < RPT; ++r) {
                const int i = r0 + r;
#pragma unroll
                for (int j = 0; j < DIM; ++j) {
                  double kij;
                  if constexpr (DIM == 2) {
                    if (i == j) kij = a.c0 * gA[i] * gB[i] + a.c2 * gA[1 - i] * gB[1 - i];
                    else kij = a.c1 * gA[i] * gB[j] + a.c2 * gA[j] * gB[i];
                  } else {
                    if (i == j) kij = a.c0 * gA[i] * gB[i] + a.c2 * (dot - gA[i] * gB[i]);
                    else kij = a.c1 * gA[i] * gB[j] + a.c2 * gA[j] * gB[i];
                  }
                  kl[b][r][j] += sc * kij;
                }
              }
            }
          }
        }
        if constexpr (RPT == DIM) {
          // K_ii = c0 T_ii + c2 (sum of T_kk, k != i), K_ij = c1 T_ij + c2 T_ji
#pragma unroll
          for (int b = 0; b < BCH; ++b) {
            double T[DIM][DIM];
#pragma unroll
            for (int p = 0; p < DIM; ++p)
#pragma unroll
              for (int q = 0; q < DIM; ++q) T[p][q] = kl[b][p][q];
#pragma unroll
            for (int p = 0; p < DIM; ++p)
#pragma unroll
              for (int q = 0; q < DIM; ++q) {
                if (p == q) {
                  double o = 0.0;
#pragma unroll
                  for (int k = 0; k < DIM; ++k)
                    if (k != p) o += T[k][k];
                  kl[b][p][p] = a.c0 * T[p][p] + a.c2 * o;
                } else {
                  kl[b][p][q] = a.c1 * T[p][q] + a.c2 * T[q][p];
                }
              }
          }
        }
        int sb[BCH];
#pragma unroll
        for (int b = 0; b < BCH; ++b) sb[b] = sl[c0 + b];
#pragma unroll
        for (int p = 0; p < LPU; ++p) {
          if constexpr (LPU > 1) __builtin_amdgcn_wave_barrier();
          if (lj == p) {
#pragma unroll
            for (int b = 0; b < BCH; ++b)
#pragma unroll
              for (int r = 0; r < RPT; ++r)
#pragma unroll
                for (int j = 0; j < DIM; ++j)
                  s_acc[((sb[b] * RPT + r) * DIM + j) * UPW + lu] += kl[b][r][j];
          }
        }
      }
    }
    if constexpr (LPU > 1) __builtin_amdgcn_wave_barrier();
    if (live) {
      // the block-sliced copy: block s of node v sits at slot bsptr[slice] + s,
      // lane pos & 63 of the slice, value (row, column) of the block at index
      // row * DIM + column
      int64_t bb0 = 0;
      int bl = 0;
      if (ga.bval) {
        const int32_t pos = ga.binv[v];
        bb0 = ga.bsptr[pos >> 6];
        bl = pos & 63;
      }
      if constexpr (RPT == DIM) {
        // every row of the node: its blocks go out as 16-byte pairs
        constexpr int DD = DIM * DIM;
        const int32_t ns = ga.rownnz[v * DIM] / DIM;
        for (int s = lj; s < ns; s += LPU) {
          const double* sa = s_acc + int64_t(s) * DD * UPW + lu;
          if (!ga.ell_diag_only) {
#pragma unroll
            for (int j = 0; j < DIM; ++j)
#pragma unroll
              for (int r = 0; r < DIM; ++r)
                a.val[v * DIM + r + (int64_t(s) * DIM + j) * a.n] = sa[(r * DIM + j) * UPW];
          }
          if (ga.bval) {
            typedef double P2 __attribute__((ext_vector_type(2)));
            double* blk = ga.bval + (bb0 + s) * DD * 64;
#pragma unroll
            for (int kp = 0; kp < DD / 2; ++kp) {
              P2 pr;
              pr.x = sa[(2 * kp) * UPW];
              pr.y = sa[(2 * kp + 1) * UPW];
              reinterpret_cast<P2*>(blk)[kp * 64 + bl] = pr;
            }
            if constexpr (DD & 1) blk[(DD - 1) * 64 + bl] = sa[(DD - 1) * UPW];
          }
        }
        if (ga.ell_diag_only && lj == 0) {
#pragma unroll
          for (int r = 0; r < DIM; ++r) {
            const int64_t row = v * DIM + r;
            const int32_t ds = ga.diag[row];
            if (ds >= 0)
              a.val[row + int64_t(ds) * a.n] =
                  s_acc[(int64_t(ds / DIM) * DD + r * DIM + ds % DIM) * UPW + lu];
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const int64_t row = v * DIM + r0 + r;
          const int32_t ns = ga.rownnz[row] / DIM;
          for (int s = lj; s < ns; s += LPU)
#pragma unroll
            for (int j = 0; j < DIM; ++j) {
              const double kv = s_acc[((s * RPT + r) * DIM + j) * UPW + lu];
              if (!ga.ell_diag_only) a.val[row + (int64_t(s) * DIM + j) * a.n] = kv;
              if (ga.bval) ga.bval[bsell_vidx(DIM * DIM, bb0 + s, (r0 + r) * DIM + j, bl)] = kv;
            }
          const int32_t ds = ga.diag[row];
          if (ga.ell_diag_only && lj == 0 && ds >= 0)
            a.val[row + int64_t(ds) * a.n] =
                s_acc[((int64_t(ds / DIM) * RPT + r) * DIM + ds % DIM) * UPW + lu];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// ELL expansion: row dim*v+c gets columns dim*w+c' for w in adj(v).
// ---------------------------------------------------------------------------
__global__ void k_ell_expand(int64_t n, int32_t width, int dim, const int64_t* adj_ptr,
                             const int32_t* adj, double* val, int32_t* col,
                             int32_t* rownnz) {
  const int64_t total = n * int64_t(width);
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t row = t % n;
    const int32_t s = int32_t(t / n);
    const int64_t v = row / dim;
    const int64_t p0 = adj_ptr[v];
    const int32_t rn = int32_t(adj_ptr[v + 1] - p0) * dim;
    int32_t c = int32_t(row);
    if (s < rn) c = adj[p0 + s / dim] * dim + s % dim;
    col[t] = c;
    val[t] = 0.0;
    if (s == 0) rownnz[row] = rn;
  }
}

__global__ void k_diag_slot(int64_t n, int dim, const int64_t* adj_ptr, const int32_t* adj,
                            int32_t* diag) {
  for (int64_t row = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; row < n;
       row += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = row / dim;
    int64_t lo = adj_ptr[v], hi = adj_ptr[v + 1];
    const int64_t p0 = lo;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (adj[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    diag[row] = (lo < adj_ptr[v + 1] && adj[lo] == v)
                    ? int32_t((lo - p0) * dim + row % dim)
                    : -1;
  }
}

// ---------------------------------------------------------------------------
// Penalty (src/FEM/fem.cc:112-145)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag_max(const double* val, const int32_t* diag,
                                                  int64_t n, double* partials) {
  __shared__ double sh[4];
  double m = -DBL_MAX;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * 256) {
    const int32_t s = diag[i];
    const double v = s >= 0 ? val[i + int64_t(s) * n] : 0.0;
    m = fmax(m, v);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ __launch_bounds__(256) void k_max_finish(const double* partials, int np,
                                                    double coef, double* result) {
  __shared__ double sh[4];
  double m = -DBL_MAX;
  for (int i = threadIdx.x; i < np; i += 256) m = fmax(m, partials[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) result[0] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3])) * coef;
}

__global__ void k_penalty_apply(double* val, const int32_t* diag, int64_t n,
                                const int32_t* dofs, const int32_t* counts, int64_t nd,
                                const double* pen, double* bval, const int64_t* bsptr,
                                const int32_t* binv, int D) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nd;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t d = dofs[t];
    const int32_t s = diag[d];
    if (s < 0) continue;
    double* v = val + d + int64_t(s) * n;
    double x = *v;
    const double p = *pen;
    for (int c = 0; c < counts[t]; ++c) x = p * x;
    *v = x;
    if (bval) {
      // ELL slot s of row d = D R + c is value (c, s % D) of block s / D of node row R
      const int32_t pos = binv[d / D];
      const int c = int(d % D);
      bval[bsell_vidx(D * D, bsptr[pos >> 6] + s / D, c * D + s % D, pos & 63)] = x;
    }
  }
}

// ---------------------------------------------------------------------------
// Zero substitution (the elimination described at src/FEM/fem.cc:146-163):
// every stored entry of a fixed row or column becomes 0, a fixed diagonal 1.
// Slot-major sweep: consecutive threads take consecutive rows of one slot.
// ---------------------------------------------------------------------------
__global__ void k_mark_dofs(const int32_t* __restrict__ dofs, int64_t nd,
                            uint8_t* __restrict__ fixed) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nd;
       t += int64_t(gridDim.x) * blockDim.x)
    fixed[dofs[t]] = 1;
}

__global__ void k_eliminate(int64_t n, int32_t width, const uint8_t* __restrict__ fixed,
                            const int32_t* __restrict__ col, const int32_t* __restrict__ rownnz,
                            double* __restrict__ val) {
  const int64_t total = n * int64_t(width);
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t row = t % n;
    if (t / n >= rownnz[row]) continue;
    const int64_t c = col[t];
    if (fixed[row] | fixed[c]) val[t] = row == c ? 1.0 : 0.0;
  }
}

__global__ void k_zero_dofs(double* __restrict__ v, int64_t n, int32_t nrhs,
                            const int32_t* __restrict__ dofs, int64_t nd) {
  const int64_t total = nd * nrhs;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x)
    v[(t / nd) * n + dofs[t % nd]] = 0.0;
}

// ---------------------------------------------------------------------------
// Force vector (FemData::CalcForceVector, src/FEM/FemData.cc:370-385): rows
// [node, F_1..F_dim] assign their nonzero components in row order.  In
// parallel: each DOF's owner is the last row with a nonzero component for
// it (atomic max of the row index), and only the owner writes.
// ---------------------------------------------------------------------------
__global__ void k_fill_f64_i32(double* __restrict__ f, int32_t* __restrict__ owner, int64_t n) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x) {
    f[t] = 0.0;
    owner[t] = -1;
  }
}

__global__ void k_load_owner(const double* __restrict__ loads, int64_t nload, int dim,
                             int32_t* __restrict__ owner) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nload * dim;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = t / dim;
    const int j = int(t % dim);
    const double* row = loads + i * (dim + 1);
    if (row[1 + j] != 0.0) atomicMax(owner + int64_t(dim) * (int64_t(row[0]) - 1) + j, int32_t(i));
  }
}

__global__ void k_load_assign(const double* __restrict__ loads, int64_t nload, int dim,
                              const int32_t* __restrict__ owner, double* __restrict__ f) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nload * dim;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = t / dim;
    const int j = int(t % dim);
    const double* row = loads + i * (dim + 1);
    const int64_t dof = int64_t(dim) * (int64_t(row[0]) - 1) + j;
    if (row[1 + j] != 0.0 && owner[dof] == int32_t(i)) f[dof] = row[1 + j];
  }
}

inline int grid_for(int64_t work, int threads, int cap) {
  int64_t b = (work + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return int(b);
}

// elements per workgroup of the pair-major kernel: LDS of the dN/dx tables
// about 48 KB or less
template <int DIM, int NEN, int NGP>
constexpr int pm_epb() {
  constexpr int per = NGP * DIM * NEN * 8;
  return per * 32 <= 50 * 1024 ? 32 : (per * 16 <= 64 * 1024 ? 16 : (per * 8 <= 64 * 1024 ? 8 : 4));
}

template <int DIM, int NEN, int NGP>
hipError_t launch_asm(hipStream_t s, const KArgs& k, int variant) {
  if (variant == 1) {
    constexpr int EPB = pm_epb<DIM, NEN, NGP>();
    const int64_t blocks = (k.nce + EPB - 1) / EPB;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_assemble_pm<DIM, NEN, NGP, EPB>), dim3(unsigned(blocks)), dim3(256), 0,
                       s, k);
    return hipGetLastError();
  }
  constexpr int PAIRS = NEN * NEN;
  constexpr int EPB = PAIRS >= 256 ? 1 : 256 / PAIRS;
  const int64_t blocks = (k.nce + EPB - 1) / EPB;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((k_assemble<DIM, NEN, NGP>), dim3(unsigned(blocks)), dim3(256), 0, s, k);
  return hipGetLastError();
}

}  // namespace

hipError_t upload_element_tables() {
  static double host[kTabCap];
  int off = 0;
  for (int type = GF_Q4; type <= GF_BRICK20; ++type) {
    ElemInfo ei;
    elem_info(type, &ei);
    for (int ng = 2; ng <= 3; ++ng) {
      double pts[27 * 3], w[27];
      const int ngp = gauss_rule(ei.dim, ng, pts, w);
      const int blk = tab_block(type, ng);
      h_off[blk] = off;
      h_ngp[blk] = ngp;
      for (int g = 0; g < ngp; ++g) shape_derivs(type, pts + g * ei.dim, host + off + g * ei.dim * ei.nen);
      off += ngp * ei.dim * ei.nen;
      for (int g = 0; g < ngp; ++g) host[off + g] = w[g];
      off += ngp;
    }
  }
  if (off > kTabCap) return hipErrorInvalidValue;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), host, sizeof(double) * size_t(off));
}

hipError_t launch_ell_expand(hipStream_t s, int64_t n, int32_t width, int dim,
                             const int64_t* adj_ptr, const int32_t* adj, double* val,
                             int32_t* col, int32_t* rownnz, int32_t* diag) {
  const int64_t total = n * int64_t(width);
  hipLaunchKernelGGL(k_ell_expand, dim3(grid_for(total, 256, 65536)), dim3(256), 0, s, n,
                     width, dim, adj_ptr, adj, val, col, rownnz);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_slot, dim3(grid_for(n, 256, 65536)), dim3(256), 0, s, n, dim,
                     adj_ptr, adj, diag);
  return hipGetLastError();
}

hipError_t launch_assembly_color(hipStream_t s, const AssemblyArgs& a) {
  ElemInfo ei;
  if (!elem_info(a.type, &ei) || (a.ngauss != 2 && a.ngauss != 3)) return hipErrorInvalidValue;
  KArgs k;
  k.conn = a.conn;
  k.coords = a.coords;
  k.elemE = a.elemE;
  k.E = a.E;
  k.c0 = a.c[0];
  k.c1 = a.c[1];
  k.c2 = a.c[2];
  k.elems = a.elems;
  k.nce = a.nce;
  k.slot = a.slot;
  k.val = a.val;
  k.n = a.n;
  k.bad = a.bad_jac;
  k.tab = h_off[tab_block(a.type, a.ngauss)];
  const bool g3 = a.ngauss == 3;
  const int v = a.variant;
  switch (a.type) {
    case GF_Q4: return g3 ? launch_asm<2, 4, 9>(s, k, v) : launch_asm<2, 4, 4>(s, k, v);
    case GF_Q8: return g3 ? launch_asm<2, 8, 9>(s, k, v) : launch_asm<2, 8, 4>(s, k, v);
    case GF_BRICK8: return g3 ? launch_asm<3, 8, 27>(s, k, v) : launch_asm<3, 8, 8>(s, k, v);
    case GF_BRICK20: return g3 ? launch_asm<3, 20, 27>(s, k, v) : launch_asm<3, 20, 8>(s, k, v);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Gauss-point geometry of every element, once (the first kernel of variant
// 4): one thread per (element, Gauss point) forms J, detJ and J^-1 from the
// coordinates with the node gather's own arithmetic and stores J^-1
// (row-major) and w detJ E; the node gather (SRC 2) then reads them instead
// of forming them again for every node of the element and chunk of
// neighbours.  Measured slower than forming them (Brick8 128^3 4.83 against
// 4.15-4.3 ms, Q8 1024x512 0.70 against 0.58 ms; storing dN/dx of every node
// instead: 5.9-6.3 and 1.02 ms): the gather is bound by its loads, not by
// the FP64 work of the geometry.
// ---------------------------------------------------------------------------
template <int DIM, int NEN, int NGP>
__global__ __launch_bounds__(256) void k_gauss_geom(KArgs a, double* geom, int64_t nelem) {
  constexpr int GS = DIM * DIM + 1;
  const int64_t units = nelem * NGP;
  for (int64_t u = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < units;
       u += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = u / NGP;
    const int g = int(u - e * NGP);
    double x[NEN][DIM];
#pragma unroll
    for (int n = 0; n < NEN; ++n) {
      const int64_t node = a.conn[e * NEN + n] - 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) x[n][d] = a.coords[node * DIM + d];
    }
    const double* dN = c_tab + a.tab + g * DIM * NEN;
    double J[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int n = 0; n < NEN; ++n) acc += dN[i * NEN + n] * x[n][j];
        J[i][j] = acc;
      }
    double Ji[DIM][DIM];
    const double det = det_inv<DIM>(J, Ji);
    if (!(det > 0.0)) *a.bad = 1;
    const double Ee = a.elemE ? a.elemE[e] : a.E;
    double* out = geom + u * GS;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) out[i * DIM + j] = Ji[i][j];
    out[GS - 1] = c_tab[a.tab + NGP * DIM * NEN + g] * det * Ee;
  }
}

namespace {
template <int DIM, int NEN, int NGP, int MAXS, int RPT, int LPU, int UPW, int BCH, int GPU = 1,
          int SRC = 0>
hipError_t launch_gather_cfg(hipStream_t s, const GArgs& g) {
  const int64_t units = g.nnodes * (DIM / RPT);
  if (units == 0) return hipSuccess;
  // one tile of UPW units per workgroup: a wavefront ends after its stores
  // instead of waiting for them before the loads of a next tile
  hipLaunchKernelGGL((k_assemble_gather<DIM, NEN, NGP, MAXS, RPT, LPU, UPW, BCH, GPU, SRC>),
                     dim3(grid_for(units, UPW, 1 << 30)), dim3(64), 0, s, g);
  return hipGetLastError();
}

// Unit shapes (measured on MI355X, gather kernel alone, LDS and registers
// from the compiler's resource report):
//  2D: node units (both rows), 2 lanes per node; Q8 1024x512 0.62 ms (one lane
//      per node: 0.86 ms, 43 KB of LDS per workgroup, 1 wavefront per SIMD).
//  Brick8: node units, 8 lanes per node, 8 nodes per workgroup (15.5 KB of
//      LDS: up to 8 one-wave workgroups per CU, as the registers allow; 4 x 16
//      with 31 KB fits 5: 4.18 ms), neighbours in chunks of 4 (all 8 at once:
//      6.0 ms), Gauss-point loop unrolled by 2 (the table loads of two points
//      batched: 4.38 -> 4.18 ms at 4 x 16; by 4: 4.75 ms; unrolled by 1 at
//      8 x 8: 4.29 ms; the 2D and Brick20 kernels gain nothing or lose):
//      128^3 4.06-4.08 ms; row units (one row per unit, each element's
//      geometry once per row): 6.3-6.6 ms with 2 lanes per row, 9.7 ms with
//      one lane per row.
//  Brick20: row units, 4 lanes per row (256 VGPRs + AGPRs): 48^3 8.1-8.7 ms;
//      node units need chunks of at most 4 neighbours for their registers and
//      recompute the 20-node geometry per chunk (10.8-20.5 ms).
template <int NEN, int NGP, int MAXS>
hipError_t launch_gather2(hipStream_t s, const GArgs& g) {
  if constexpr (NEN == 4) return launch_gather_cfg<2, NEN, NGP, MAXS, 2, 2, 32, NEN>(s, g);
  else return launch_gather_cfg<2, NEN, NGP, MAXS, 2, 2, 30, NEN>(s, g);
}
}  // namespace

int gather_max_slots(int type) {
  switch (type) {
    case GF_Q4: return 9;
    case GF_Q8: return 21;
    case GF_BRICK8: return 27;
    case GF_BRICK20: return 81;
    default: return 0;
  }
}

namespace {
GArgs gather_args(const AssemblyArgs& a) {
  GArgs g{};
  KArgs& k = g.k;
  k.conn = a.conn;
  k.coords = a.coords;
  k.elemE = a.elemE;
  k.E = a.E;
  k.c0 = a.c[0];
  k.c1 = a.c[1];
  k.c2 = a.c[2];
  k.elems = nullptr;
  k.nce = 0;
  k.slot = a.slot;
  k.val = a.val;
  k.n = a.n;
  k.bad = a.bad_jac;
  k.tab = h_off[tab_block(a.type, a.ngauss)];
  g.n2e_ptr = a.n2e_ptr;
  g.n2e = a.n2e;
  g.rownnz = a.rownnz;
  g.nnodes = a.nnodes;
  g.bval = a.bval;
  g.bsptr = a.bsptr;
  g.binv = a.binv;
  g.diag = a.diag;
  g.ell_diag_only = a.bval && a.ell_diag_only ? 1 : 0;
  g.tens = nullptr;
  g.geom = nullptr;
  return g;
}

template <int NEN, int NGP, int MAXS, int LPU, int UPW, int BCH>
hipError_t launch_tensors(hipStream_t s, const AssemblyArgs& a, double* tens) {
  GArgs g = gather_args(a);
  g.tens = tens;
  if (a.nelem > 0) {
    TArgs ta{g.k, tens, a.nelem};
    hipLaunchKernelGGL((k_elem_tensors<NEN, NGP>), dim3(unsigned((a.nelem + 3) / 4)), dim3(256),
                       0, s, ta);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return launch_gather_cfg<3, NEN, NGP, MAXS, 3, LPU, UPW, BCH, 1, 1>(s, g);
}
}  // namespace

int64_t elem_tensor_doubles(int type, int64_t nelem) {
  ElemInfo ei;
  if (!elem_info(type, &ei) || ei.dim != 3) return 0;
  return nelem * 9 * int64_t(ei.nen) * ei.nen;
}

hipError_t launch_assembly_tensors(hipStream_t s, const AssemblyArgs& a, double* tens) {
  if (a.ngauss != 2 && a.ngauss != 3) return hipErrorInvalidValue;
  const bool g3 = a.ngauss == 3;
  switch (a.type) {
    case GF_BRICK8:
      return g3 ? launch_tensors<8, 27, 27, 4, 16, 8>(s, a, tens)
                : launch_tensors<8, 8, 27, 4, 16, 8>(s, a, tens);
    case GF_BRICK20:
      // gather phase: 8 lanes x 8 nodes, neighbours 10 at a time (48^3: 4.19 ms;
      // chunks of 4/5: 4.22-4.23, 16 lanes x 4 nodes: 4.40-4.49, 4 x 8: 4.46)
      return g3 ? launch_tensors<20, 27, 81, 8, 8, 10>(s, a, tens)
                : launch_tensors<20, 8, 81, 8, 8, 10>(s, a, tens);
    default: return hipErrorInvalidValue;
  }
}

int64_t gauss_geom_doubles(int type, int ngauss, int64_t nelem) {
  ElemInfo ei;
  if (!elem_info(type, &ei)) return 0;
  int ngp = 1;
  for (int d = 0; d < ei.dim; ++d) ngp *= ngauss;
  return nelem * ngp * (int64_t(ei.dim) * ei.dim + 1);
}

namespace {
template <int DIM, int NEN, int NGP, int MAXS, int LPU, int UPW, int BCH, int GPU>
hipError_t launch_geom(hipStream_t s, const AssemblyArgs& a, double* geom) {
  GArgs g = gather_args(a);
  g.geom = geom;
  if (a.nelem > 0) {
    hipLaunchKernelGGL((k_gauss_geom<DIM, NEN, NGP>),
                       dim3(grid_for(a.nelem * NGP, 256, 65536)), dim3(256), 0, s, g.k, geom,
                       a.nelem);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return launch_gather_cfg<DIM, NEN, NGP, MAXS, DIM, LPU, UPW, BCH, GPU, 2>(s, g);
}
}  // namespace

hipError_t launch_assembly_geom(hipStream_t s, const AssemblyArgs& a, double* geom) {
  ElemInfo ei;
  if (!elem_info(a.type, &ei) || (a.ngauss != 2 && a.ngauss != 3) || gather_max_slots(a.type) == 0)
    return hipErrorInvalidValue;
  const bool g3 = a.ngauss == 3;
  switch (a.type) {
    case GF_Q4:
      return g3 ? launch_geom<2, 4, 9, 9, 2, 32, 4, 1>(s, a, geom)
                : launch_geom<2, 4, 4, 9, 2, 32, 4, 1>(s, a, geom);
    case GF_Q8:
      return g3 ? launch_geom<2, 8, 9, 21, 2, 30, 8, 1>(s, a, geom)
                : launch_geom<2, 8, 4, 21, 2, 30, 8, 1>(s, a, geom);
    case GF_BRICK8:
      return g3 ? launch_geom<3, 8, 27, 27, 8, 8, 4, 2>(s, a, geom)
                : launch_geom<3, 8, 8, 27, 8, 8, 4, 2>(s, a, geom);
    case GF_BRICK20:
      return g3 ? launch_geom<3, 20, 27, 81, 8, 8, 10, 1>(s, a, geom)
                : launch_geom<3, 20, 8, 81, 8, 8, 10, 1>(s, a, geom);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_assembly_gather(hipStream_t s, const AssemblyArgs& a) {
  ElemInfo ei;
  if (!elem_info(a.type, &ei) || (a.ngauss != 2 && a.ngauss != 3) || gather_max_slots(a.type) == 0)
    return hipErrorInvalidValue;
  const GArgs g = gather_args(a);
  const bool g3 = a.ngauss == 3;
  switch (a.type) {
    case GF_Q4: return g3 ? launch_gather2<4, 9, 9>(s, g) : launch_gather2<4, 4, 9>(s, g);
    case GF_Q8: return g3 ? launch_gather2<8, 9, 21>(s, g) : launch_gather2<8, 4, 21>(s, g);
    case GF_BRICK8:
      return g3 ? launch_gather_cfg<3, 8, 27, 27, 3, 8, 8, 4, 2>(s, g)
                : launch_gather_cfg<3, 8, 8, 27, 3, 8, 8, 4, 2>(s, g);
    case GF_BRICK20:
      return g3 ? launch_gather_cfg<3, 20, 27, 81, 1, 4, 16, 20>(s, g)
                : launch_gather_cfg<3, 20, 8, 81, 1, 4, 16, 20>(s, g);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_penalty_factor(hipStream_t s, const double* val, const int32_t* diag,
                                 int64_t n, double coef, double* partials, int nparts,
                                 double* result) {
  const int nb = grid_for(n, 256, nparts);
  hipLaunchKernelGGL(k_diag_max, dim3(nb), dim3(256), 0, s, val, diag, n, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_max_finish, dim3(1), dim3(256), 0, s, partials, nb, coef, result);
  return hipGetLastError();
}

hipError_t launch_eliminate(hipStream_t s, int64_t n, int32_t width, const int32_t* dofs,
                            int64_t nd, uint8_t* fixed, const int32_t* col,
                            const int32_t* rownnz, double* val) {
  if (nd > 0) {
    hipLaunchKernelGGL(k_mark_dofs, dim3(grid_for(nd, 256, 4096)), dim3(256), 0, s, dofs, nd,
                       fixed);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int64_t total = n * int64_t(width);
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_eliminate, dim3(grid_for(total, 256, 65536)), dim3(256), 0, s, n, width,
                     fixed, col, rownnz, val);
  return hipGetLastError();
}

hipError_t launch_zero_dofs(hipStream_t s, double* v, int64_t n, int32_t nrhs,
                            const int32_t* dofs, int64_t nd) {
  if (nd == 0 || nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_dofs, dim3(grid_for(nd * nrhs, 256, 4096)), dim3(256), 0, s, v, n,
                     nrhs, dofs, nd);
  return hipGetLastError();
}

hipError_t launch_load_vector(hipStream_t s, int dim, int64_t n, const double* loads,
                              int64_t nload, int32_t* owner, double* f) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f64_i32, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, f, owner, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nload == 0) return e;
  const int g = grid_for(nload * dim, 256, 4096);
  hipLaunchKernelGGL(k_load_owner, dim3(g), dim3(256), 0, s, loads, nload, dim, owner);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_load_assign, dim3(g), dim3(256), 0, s, loads, nload, dim, owner, f);
  return hipGetLastError();
}

hipError_t launch_penalty_apply(hipStream_t s, double* val, const int32_t* diag, int64_t n,
                                const int32_t* dofs, const int32_t* counts, int64_t nd,
                                const double* pen, double* bval, const int64_t* bsptr,
                                const int32_t* binv, int d) {
  if (nd == 0) return hipSuccess;
  hipLaunchKernelGGL(k_penalty_apply, dim3(grid_for(nd, 256, 4096)), dim3(256), 0, s, val,
                     diag, n, dofs, counts, nd, pen, bval, bsptr, binv, d);
  return hipGetLastError();
}

}  // namespace gf
