// Grouped batch API (gf_group_*): many independent meshes of one element
// type concatenated into one block-diagonal system.  Assembly reuses the
// single-mesh path on the concatenated mesh (one launch per color covers
// every mesh); the penalty factor and the PCG scalars are per mesh.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "group.hpp"
#include "solve_plan.hpp"

struct gf_group {
  int dev = 0;
  int nmesh = 0;
  gf_mesh* mesh = nullptr;
  gf_ell* ell = nullptr;
  std::vector<int64_t> off;       // row offsets
  std::vector<int64_t> tile_ptr;  // tile offsets per mesh
  int64_t ntiles = 0;
  // device arrays
  int64_t *d_off = nullptr, *d_tile_ptr = nullptr, *d_tile_r0 = nullptr, *d_tile_r1 = nullptr;
  int32_t* d_tile_mesh = nullptr;
  double *b = nullptr, *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
  double* part = nullptr;
  // slice groups of the block-sliced SpMV (built on the first solve that uses it)
  int64_t nsg = 0;
  int64_t *d_sg_ptr = nullptr, *d_sg_s0 = nullptr, *d_sg_s1 = nullptr;
  int32_t* d_sg_mesh = nullptr;
  double* part_sg = nullptr;
  double* per_mesh = nullptr;   // rho, alpha, beta, rr, bb, pen, res: 7 * nmesh
  int32_t* per_mesh_i = nullptr;  // it, done, status: 3 * nmesh
  gf::GroupState* st = nullptr;
  hipGraphExec_t graph = nullptr;
};

namespace gf {
int set_error(gf_ctx* ctx, int code, const std::string& msg);
}

using namespace gf;

namespace {

constexpr int kChunk = 16;
constexpr int kMaxRestarts = 4;  // true-residual restarts

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void group_release(gf_group* g) {
  if (g->graph) (void)hipGraphExecDestroy(g->graph);
  g->graph = nullptr;
  dfree(g->d_off);
  dfree(g->d_tile_ptr);
  dfree(g->d_tile_r0);
  dfree(g->d_tile_r1);
  dfree(g->d_tile_mesh);
  double** v[] = {&g->b, &g->x, &g->r, &g->p, &g->q, &g->part, &g->per_mesh};
  for (double** d : v) dfree(*d);
  dfree(g->per_mesh_i);
  dfree(g->d_sg_ptr);
  dfree(g->d_sg_s0);
  dfree(g->d_sg_s1);
  dfree(g->d_sg_mesh);
  dfree(g->part_sg);
  dfree(g->st);
  if (g->ell) gf_ell_free(g->ell);
  if (g->mesh) gf_mesh_free(g->mesh);
  g->ell = nullptr;
  g->mesh = nullptr;
}

// The grouped SpMV runs over the block-sliced copy whose windows restart at
// every mesh (a slice belongs to one mesh) whenever the merged matrix has
// node blocks (on 256 Q4 plates 0.40 s against 0.52 s over node columns and
// 0.59 s over ELL rows); otherwise thread per row with the tuner's unroll.
int group_variant(const gf_ell* m) {
  const int b = bsell_variant(m);
  if (b == kBsell2) return kBsell2G;
  if (b == kBsell3) return kBsell3G;
  const int v = m->plan ? m->plan->variant : -1;
  if (v >= 0 && v <= kThreadU8) return v;
  return kThreadU8;
}

// slice groups per mesh from the block-sliced structure (once)
int ensure_slice_groups(gf_ctx* ctx, gf_group* g) {
  if (g->d_sg_ptr) return GF_OK;
  const std::vector<int64_t>& bs = g->ell->bsell_break_slice;
  if (int(bs.size()) != g->nmesh + 1)
    return set_error(ctx, GF_ERR_HIP, "gf_group_pcg: block-sliced windows do not follow the meshes");
  std::vector<int64_t> ptr(size_t(g->nmesh) + 1, 0), s0, s1;
  std::vector<int32_t> mesh;
  for (int k = 0; k < g->nmesh; ++k) {
    for (int64_t s = bs[size_t(k)]; s < bs[size_t(k) + 1]; s += 4) {
      s0.push_back(s);
      s1.push_back(std::min<int64_t>(s + 4, bs[size_t(k) + 1]));
      mesh.push_back(k);
    }
    ptr[size_t(k) + 1] = int64_t(mesh.size());
  }
  g->nsg = int64_t(mesh.size());
  const size_t ns = std::max<size_t>(mesh.size(), 1);
  hipError_t e = hipMalloc(&g->d_sg_ptr, ptr.size() * 8);
  if (e == hipSuccess) e = hipMalloc(&g->d_sg_s0, ns * 8);
  if (e == hipSuccess) e = hipMalloc(&g->d_sg_s1, ns * 8);
  if (e == hipSuccess) e = hipMalloc(&g->d_sg_mesh, ns * 4);
  if (e == hipSuccess) e = hipMalloc(&g->part_sg, ns * 8);
  if (e == hipSuccess) e = hipMemcpy(g->d_sg_ptr, ptr.data(), ptr.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_sg_s0, s0.data(), s0.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_sg_s1, s1.data(), s1.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_sg_mesh, mesh.data(), mesh.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return set_error(ctx, GF_ERR_NOMEM, std::string("gf_group_pcg: ") + hipGetErrorString(e));
  return GF_OK;
}

GroupBuffers buffers(gf_group* g) {
  GroupBuffers G;
  gf_ell* m = g->ell;
  G.n = m->n;
  G.nmesh = g->nmesh;
  G.ntiles = g->ntiles;
  G.nblocks = group_blocks(g->ntiles);
  G.variant = group_variant(m);
  G.off = g->d_off;
  G.tile_ptr = g->d_tile_ptr;
  G.tile_mesh = g->d_tile_mesh;
  G.tile_r0 = g->d_tile_r0;
  G.tile_r1 = g->d_tile_r1;
  G.val = m->d_val;
  G.col = m->d_col;
  G.scol = is_node_variant(G.variant) ? m->d_coln : m->d_col;
  const bool bsell = is_bsell_variant(G.variant) && g->d_sg_ptr;
  G.bsell_d = bsell ? m->dim : 0;
  G.bval = bsell ? m->d_bval : nullptr;
  G.bcol = bsell ? m->d_bcol : nullptr;
  G.bs = BsellView{m->d_bperm, m->d_bsptr, m->bslices, m->d_bcoff};
  G.nsg = g->nsg;
  G.sg_ptr = g->d_sg_ptr;
  G.sg_mesh = g->d_sg_mesh;
  G.sg_s0 = g->d_sg_s0;
  G.sg_s1 = g->d_sg_s1;
  G.part_sg = g->part_sg;
  G.rownnz = m->d_rownnz;
  G.diag = m->d_diag;
  G.dinv = m->d_dinv;
  G.b = g->b;
  G.x = g->x;
  G.r = g->r;
  G.p = g->p;
  G.q = g->q;
  G.part = g->part;
  const int M = g->nmesh;
  G.rho = g->per_mesh;
  G.alpha = g->per_mesh + M;
  G.beta = g->per_mesh + 2 * M;
  G.rr = g->per_mesh + 3 * M;
  G.bb = g->per_mesh + 4 * M;
  G.pen = g->per_mesh + 5 * M;
  G.it = g->per_mesh_i;
  G.done = g->per_mesh_i + M;
  G.status = g->per_mesh_i + 2 * M;
  G.st = g->st;
  return G;
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int gf_group_create(gf_ctx* ctx, int32_t nmesh, const gf_mesh_desc* meshes, gf_group** out) {
  if (!ctx || !meshes || !out || nmesh < 1)
    return set_error(ctx, GF_ERR_ARG, "gf_group_create: bad argument");
  *out = nullptr;
  const gf_mesh_desc& m0 = meshes[0];
  ElemInfo ei;
  if (!elem_info(m0.elem_type, &ei))
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_group_create: unknown element type");
  int64_t nn = 0, ne = 0;
  for (int k = 0; k < nmesh; ++k) {
    const gf_mesh_desc& d = meshes[k];
    if (d.elem_type != m0.elem_type || d.ngauss != m0.ngauss || d.nu != m0.nu)
      return set_error(ctx, GF_ERR_ARG,
                       "gf_group_create: meshes must share element type, Gauss rule and nu");
    if (!d.conn || !d.coords || d.nnodes <= 0 || d.nelem <= 0)
      return set_error(ctx, GF_ERR_ARG, "gf_group_create: empty mesh");
    nn += d.nnodes;
    ne += d.nelem;
  }
  // concatenated mesh: node ids shifted, E per element
  std::vector<int32_t> conn(size_t(ne) * ei.nen);
  std::vector<double> coords(static_cast<size_t>(nn) * ei.dim);
  std::vector<double> E(static_cast<size_t>(ne));
  std::unique_ptr<gf_group> g(new gf_group());
  g->dev = ctx->device;
  g->nmesh = nmesh;
  g->off.assign(size_t(nmesh) + 1, 0);
  int64_t node0 = 0, el0 = 0;
  for (int k = 0; k < nmesh; ++k) {
    const gf_mesh_desc& d = meshes[k];
    for (int64_t i = 0; i < d.nelem * ei.nen; ++i) {
      if (d.conn[i] < 1 || d.conn[i] > d.nnodes)
        return set_error(ctx, GF_ERR_ARG, "gf_group_create: connectivity out of range");
      conn[size_t(el0 * ei.nen + i)] = int32_t(d.conn[i] + node0);
    }
    std::memcpy(coords.data() + node0 * ei.dim, d.coords, size_t(d.nnodes) * ei.dim * sizeof(double));
    for (int64_t e = 0; e < d.nelem; ++e) E[size_t(el0 + e)] = d.elem_E ? d.elem_E[e] : d.E;
    node0 += d.nnodes;
    el0 += d.nelem;
    g->off[size_t(k) + 1] = node0 * ei.dim;
  }
  gf_mesh_desc all{m0.elem_type, m0.ngauss, nn, ne, conn.data(), coords.data(), E.data(), m0.E,
                   m0.nu};
  int st = gf_mesh_upload(ctx, &all, &g->mesh);
  if (st) return st;
  st = gf_ell_symbolic(ctx, g->mesh, &g->ell, nullptr);
  if (st) {
    group_release(g.get());
    return st;
  }
  // block-sliced windows restart at every mesh, so that a slice belongs to one mesh
  g->ell->bsell_breaks.clear();
  for (int k = 0; k <= nmesh; ++k) g->ell->bsell_breaks.push_back(g->off[size_t(k)] / ei.dim);
  // tiles that never cross a mesh boundary
  std::vector<int32_t> tmesh;
  std::vector<int64_t> tr0, tr1;
  g->tile_ptr.assign(size_t(nmesh) + 1, 0);
  for (int k = 0; k < nmesh; ++k) {
    for (int64_t r = g->off[size_t(k)]; r < g->off[size_t(k) + 1]; r += 256) {
      tmesh.push_back(k);
      tr0.push_back(r);
      tr1.push_back(std::min<int64_t>(r + 256, g->off[size_t(k) + 1]));
    }
    g->tile_ptr[size_t(k) + 1] = int64_t(tmesh.size());
  }
  g->ntiles = int64_t(tmesh.size());
  const int64_t n = g->ell->n;
  const size_t vb = size_t(n) * sizeof(double);
  hipError_t e = hipSuccess;
  auto alloc = [&](auto** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  };
  alloc(&g->d_off, g->off.size() * 8);
  alloc(&g->d_tile_ptr, g->tile_ptr.size() * 8);
  alloc(&g->d_tile_r0, tr0.size() * 8);
  alloc(&g->d_tile_r1, tr1.size() * 8);
  alloc(&g->d_tile_mesh, tmesh.size() * 4);
  for (double** d : {&g->b, &g->x, &g->r, &g->p, &g->q}) alloc(d, vb);
  alloc(&g->part, size_t(3 * g->ntiles) * 8);
  alloc(&g->per_mesh, size_t(7 * nmesh) * 8);
  alloc(&g->per_mesh_i, size_t(3 * nmesh) * 4);
  alloc(&g->st, sizeof(GroupState));
  // the arrival counter of the per-mesh reductions starts at zero (each of
  // those kernels leaves it at zero)
  if (e == hipSuccess) e = hipMemset(g->st, 0, sizeof(GroupState));
  if (e == hipSuccess) e = hipMemcpy(g->d_off, g->off.data(), g->off.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g->d_tile_ptr, g->tile_ptr.data(), g->tile_ptr.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_tile_r0, tr0.data(), tr0.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_tile_r1, tr1.data(), tr1.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g->d_tile_mesh, tmesh.data(), tmesh.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    group_release(g.get());
    return set_error(ctx, GF_ERR_NOMEM, std::string("gf_group_create: ") + hipGetErrorString(e));
  }
  *out = g.release();
  return GF_OK;
}

int gf_group_info(const gf_group* g, int64_t* total_dof, int64_t* nnz, int32_t* ncolors) {
  if (!g) return set_error((gf_ctx*)nullptr, GF_ERR_ARG, "gf_group_info: null group");
  if (total_dof) *total_dof = g->ell->n;
  if (nnz) *nnz = g->ell->nnz;
  if (ncolors) *ncolors = g->ell->ncolors;
  return GF_OK;
}

int gf_group_set_E(gf_ctx* ctx, gf_group* g, const double* elem_E) {
  if (!ctx || !g || !elem_E) return set_error(ctx, GF_ERR_ARG, "gf_group_set_E: null argument");
  return gf_mesh_set_E(ctx, g->mesh, elem_E);
}

int gf_group_assemble(gf_ctx* ctx, gf_group* g, gf_timing* t) {
  if (!ctx || !g) return set_error(ctx, GF_ERR_ARG, "gf_group_assemble: null argument");
  return gf_assemble(ctx, g->mesh, g->ell, t);
}

int gf_group_apply_penalty(gf_ctx* ctx, gf_group* g, const int32_t* dofs, const int64_t* ndofs,
                           double coef) {
  if (!ctx || !g || !ndofs) return set_error(ctx, GF_ERR_ARG, "gf_group_apply_penalty: null argument");
  // global dof -> multiplicity, with its mesh
  std::vector<std::pair<int64_t, int32_t>> all;
  int64_t pos = 0;
  for (int k = 0; k < g->nmesh; ++k) {
    const int64_t nk = g->off[size_t(k) + 1] - g->off[size_t(k)];
    if (ndofs[k] < 0 || (ndofs[k] > 0 && !dofs))
      return set_error(ctx, GF_ERR_ARG, "gf_group_apply_penalty: bad DOF count");
    for (int64_t i = 0; i < ndofs[k]; ++i) {
      const int32_t d = dofs[pos + i];
      if (d < 0 || d >= nk) return set_error(ctx, GF_ERR_ARG, "gf_group_apply_penalty: DOF out of range");
      all.push_back({g->off[size_t(k)] + d, k});
    }
    pos += ndofs[k];
  }
  std::sort(all.begin(), all.end());
  std::vector<int64_t> ud;
  std::vector<int32_t> um, uc;
  for (size_t i = 0; i < all.size();) {
    size_t j = i;
    while (j < all.size() && all[j].first == all[i].first) ++j;
    ud.push_back(all[i].first);
    um.push_back(all[i].second);
    uc.push_back(int32_t(j - i));
    i = j;
  }
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int st = ensure_ell_vals(ctx, g->ell);
  if (st) return st;
  int64_t* d_d = nullptr;
  int32_t* d_mc = nullptr;
  const size_t nd = ud.size();
  if (nd) {
    GF_HIP_CHECK(ctx, hipMalloc(&d_d, nd * 8));
    GF_HIP_CHECK(ctx, hipMalloc(&d_mc, nd * 8));
    GF_HIP_CHECK(ctx, hipMemcpy(d_d, ud.data(), nd * 8, hipMemcpyHostToDevice));
    GF_HIP_CHECK(ctx, hipMemcpy(d_mc, um.data(), nd * 4, hipMemcpyHostToDevice));
    GF_HIP_CHECK(ctx, hipMemcpy(d_mc + nd, uc.data(), nd * 4, hipMemcpyHostToDevice));
  }
  GroupBuffers G = buffers(g);
  hipError_t e = launch_group_penalty(ctx->stream, G, coef, d_d, d_mc, d_mc + nd, int64_t(nd));
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (d_d) (void)hipFree(d_d);
  if (d_mc) (void)hipFree(d_mc);
  if (e != hipSuccess)
    return set_error(ctx, GF_ERR_HIP, std::string("gf_group_apply_penalty: ") + hipGetErrorString(e));
  g->ell->dinv_valid = false;
  g->ell->val32_valid = false;
  ++g->ell->val_version;
  return GF_OK;
}

int gf_group_pcg(gf_ctx* ctx, gf_group* g, const double* b, double* x, const gf_solve_opts* opts_in,
                 gf_solve_stats* stats, int32_t* iterations_per_mesh) {
  if (!ctx || !g || !b || !x) return set_error(ctx, GF_ERR_ARG, "gf_group_pcg: null argument");
  gf_solve_opts opts;
  if (opts_in) opts = *opts_in;
  else gf_solve_opts_default(&opts);
  if (opts.precond != GF_PRECOND_JACOBI || opts.precision != GF_PREC_FP64)
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_group_pcg: FP64 Jacobi-PCG only");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const double h0 = now_ms();
  gf_ell* m = g->ell;
  int st = ensure_plan(ctx, m);
  if (st) return st;
  if (m->plan->variant < 0 && (st = tune_spmv(ctx, m, -1, nullptr, nullptr))) return st;
  if ((st = prepare_jacobi(ctx, m))) return st;
  if ((st = prepare_variant(ctx, m, group_variant(m)))) return st;
  // the ELL values unless the grouped kernels read the block-sliced copy
  if (!(is_bsell_variant(group_variant(m)) && g->d_sg_ptr) && (st = ensure_ell_vals(ctx, m)))
    return st;
  if (is_bsell_variant(group_variant(m)) && (st = ensure_slice_groups(ctx, g))) return st;
  const int64_t n = m->n;
  const size_t vb = size_t(n) * sizeof(double);
  const hipMemcpyKind in = opts.device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  GF_HIP_CHECK(ctx, hipMemcpyAsync(g->b, b, vb, in, ctx->stream));
  if (opts.use_x0) GF_HIP_CHECK(ctx, hipMemcpyAsync(g->x, x, vb, in, ctx->stream));
  GroupBuffers G = buffers(g);
  const bool use_graph = opts.schedule != GF_SCHED_KERNELS;
  if (use_graph && !g->graph) {
    hipGraph_t gr = nullptr;
    GF_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int it = 0; it < kChunk && e == hipSuccess; ++it) e = launch_group_iteration(ctx->stream, G);
    hipError_t e2 = hipStreamEndCapture(ctx->stream, &gr);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&g->graph, gr, nullptr, nullptr, 0);
    if (gr) (void)hipGraphDestroy(gr);
    GF_HIP_CHECK(ctx, e);
  }
  const int M = g->nmesh;
  std::vector<double> pm(size_t(7 * M));
  std::vector<int32_t> pi(size_t(3 * M)), it_total(size_t(M), 0);
  double* d_res = g->per_mesh + 6 * g->nmesh;
  int use_x0 = opts.use_x0, total = 0;
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  for (int restart = 0;; ++restart) {
    GF_HIP_CHECK(ctx, launch_group_setup(ctx->stream, G, use_x0, opts.rtol, opts.maxit - total));
    GroupState hs{};
    for (;;) {
      GF_HIP_CHECK(ctx, hipMemcpyAsync(&hs, g->st, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
      GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      if (hs.done) break;
      if (use_graph) {
        GF_HIP_CHECK(ctx, hipGraphLaunch(g->graph, ctx->stream));
      } else {
        for (int it = 0; it < kChunk; ++it) GF_HIP_CHECK(ctx, launch_group_iteration(ctx->stream, G));
      }
    }
    GF_HIP_CHECK(ctx, launch_group_x_finish(ctx->stream, G));
    GF_HIP_CHECK(ctx, launch_group_residual(ctx->stream, G, d_res));
    GF_HIP_CHECK(ctx, hipMemcpyAsync(pm.data(), g->per_mesh, pm.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP_CHECK(ctx, hipMemcpyAsync(pi.data(), g->per_mesh_i, pi.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    // per-mesh restart from the current x when the recursive residual met
    // rtol while b - K x did not (meshes already below rtol stay done)
    bool again = false, healthy = true;
    for (int k = 0; k < M; ++k) {
      it_total[size_t(k)] += pi[size_t(k)];
      total = std::max(total, it_total[size_t(k)]);
      const double rr = pm[size_t(3 * M + k)], bb = pm[size_t(4 * M + k)];
      healthy &= pi[size_t(2 * M + k)] == GF_OK;
      again |= rr <= opts.rtol * opts.rtol * bb && pm[size_t(6 * M + k)] > opts.rtol;
    }
    if (!opts.verify || !again || !healthy || restart >= kMaxRestarts || total >= opts.maxit) break;
    use_x0 = 1;
  }
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
  float ms = 0.f;
  GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  GF_HIP_CHECK(ctx, hipMemcpy(x, g->x, vb, opts.device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
  gf_solve_stats S{};
  S.converged = 1;
  S.status = GF_OK;
  for (int k = 0; k < M; ++k) {
    const int it = it_total[size_t(k)], status = pi[size_t(2 * M + k)];
    const double rr = pm[size_t(3 * M + k)], bb = pm[size_t(4 * M + k)];
    if (iterations_per_mesh) iterations_per_mesh[k] = it;
    S.iterations = std::max(S.iterations, it);
    S.relres = std::max(S.relres, bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr));
    S.true_relres = std::max(S.true_relres, pm[size_t(6 * M + k)]);
    if (status != GF_OK) {
      S.status = status;
      S.converged = 0;
    } else if (!(rr <= opts.rtol * opts.rtol * bb)) {
      if (S.status == GF_OK) S.status = GF_ERR_NOCONV;
      S.converged = 0;
    }
  }
  S.ms_device = ms;
  S.ms_host = now_ms() - h0;
  if (stats) *stats = S;
  if (S.status == GF_ERR_BREAKDOWN) return set_error(ctx, GF_ERR_BREAKDOWN, "gf_group_pcg: breakdown");
  if (S.status == GF_ERR_NOCONV)
    return set_error(ctx, GF_ERR_NOCONV, "gf_group_pcg: iteration limit reached before rtol");
  return GF_OK;
}

int gf_group_free(gf_group* g) {
  if (!g) return GF_OK;
  (void)hipSetDevice(g->dev);
  group_release(g);
  delete g;
  return GF_OK;
}

}  // extern "C"
