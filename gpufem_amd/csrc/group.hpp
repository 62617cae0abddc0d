// Grouped batch of independent meshes (kernels in group.hip, host side in
// group.cpp).
#pragma once

#include "gf_internal.hpp"
#include "launch.hpp"

namespace gf {

struct GroupState {
  int32_t it, done, maxit;
  int32_t arrived;  // workgroups of k_g_beta finished in this iteration
  double rtol2;
};

struct GroupBuffers {
  int64_t n;
  int nmesh;
  int64_t ntiles;
  int nblocks;
  int variant;
  const int64_t* off;        // [nmesh+1] first row of each mesh
  const int64_t* tile_ptr;   // [nmesh+1] first tile of each mesh
  const int32_t* tile_mesh;  // [ntiles]
  const int64_t* tile_r0;    // [ntiles]
  const int64_t* tile_r1;    // [ntiles]
  double* val;
  const int32_t* col;   // ELL columns
  const int32_t* scol;  // columns of the iteration's SpMV (node columns for kNode2/3)
  const int32_t* rownnz;
  const int32_t* diag;
  const double* dinv;
  const double* b;
  double *x, *r, *p, *q;  // z = D^-1 r is formed where it is used
  double* part;  // [3 * ntiles]
  // block-sliced SpMV (bsell_d = 2, 3; 0: thread per row over col/scol):
  // slice groups of at most four slices of one mesh, one workgroup each
  int bsell_d;
  const double* bval;
  const int32_t* bcol;
  BsellView bs;
  int64_t nsg;
  const int64_t* sg_ptr;   // [nmesh+1] first slice group of each mesh
  const int32_t* sg_mesh;  // [nsg]
  const int64_t* sg_s0;    // [nsg] first slice
  const int64_t* sg_s1;    // [nsg] end slice
  double* part_sg;         // [nsg] p.q partials
  double *rho, *alpha, *beta, *rr, *bb, *pen;  // [nmesh]
  int32_t *it, *done, *status;                 // [nmesh]
  GroupState* st;
};

size_t group_state_bytes();
int group_blocks(int64_t ntiles);
hipError_t launch_group_penalty(hipStream_t s, const GroupBuffers& G, double coef,
                                const int64_t* dofs, const int32_t* mesh, const int32_t* counts,
                                int64_t nd);
hipError_t launch_group_setup(hipStream_t s, const GroupBuffers& G, int use_x0, double rtol,
                              int maxit);
hipError_t launch_group_iteration(hipStream_t s, const GroupBuffers& G);
// x += alpha p of the group's last iteration (once, after the loop)
hipError_t launch_group_x_finish(hipStream_t s, const GroupBuffers& G);
hipError_t launch_group_residual(hipStream_t s, const GroupBuffers& G, double* out);

}  // namespace gf
