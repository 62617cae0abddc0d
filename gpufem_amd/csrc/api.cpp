// C-ABI implementation: contexts, device meshes, ELL matrices, assembly,
// penalty, SpMV and the solver driver.  Kernels live in kernels.hip and
// solver.hip; this file only orchestrates HIP runtime calls.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "gf_internal.hpp"
#include "launch.hpp"
#include "solve_plan.hpp"
#include "symbolic.hpp"

namespace gf {

namespace {
thread_local std::string g_last_error;
}

int set_error(gf_ctx* ctx, int code, const std::string& msg) {
  g_last_error = msg;
  if (ctx) ctx->err = msg;
  return code;
}
int set_error(const gf_ctx* ctx, int code, const std::string& msg) {
  g_last_error = msg;
  if (ctx) const_cast<gf_ctx*>(ctx)->err = msg;
  return code;
}

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

int ensure_partials(gf_ctx* ctx, int64_t count) {
  if (ctx->partials_cap >= count) return GF_OK;
  dfree(ctx->d_partials);
  GF_HIP_CHECK(ctx, hipMalloc(&ctx->d_partials, size_t(count) * sizeof(double)));
  ctx->partials_cap = count;
  return GF_OK;
}

int record_start(gf_ctx* ctx) {
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  return GF_OK;
}

int record_stop(gf_ctx* ctx, double* ms) {
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
  float f = 0.f;
  GF_HIP_CHECK(ctx, hipEventElapsedTime(&f, ctx->ev_start, ctx->ev_stop));
  if (ms) *ms = double(f);
  return GF_OK;
}

int select_device(gf_ctx* ctx) {
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return GF_OK;
}

}  // namespace gf

using namespace gf;

extern "C" {

const char* gf_last_error(const gf_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_last_error.c_str();
}

int gf_create(gf_ctx** out, int32_t device) {
  if (!out) return set_error((gf_ctx*)nullptr, GF_ERR_ARG, "gf_create: out is null");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return set_error((gf_ctx*)nullptr, GF_ERR_HIP,
                     std::string("gf_create: no HIP device available (") +
                         hipGetErrorString(e) + ")");
  if (device < 0 || device >= ndev)
    return set_error((gf_ctx*)nullptr, GF_ERR_ARG, "gf_create: device index out of range");
  std::unique_ptr<gf_ctx> c(new gf_ctx());
  c->device = device;
  GF_HIP_CHECK(c.get(), hipSetDevice(device));
  hipDeviceProp_t prop;
  GF_HIP_CHECK(c.get(), hipGetDeviceProperties(&prop, device));
  c->num_cu = prop.multiProcessorCount;
  c->lds_per_cu = int(prop.maxSharedMemoryPerMultiProcessor);
  c->name = prop.gcnArchName;
  GF_HIP_CHECK(c.get(), hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  GF_HIP_CHECK(c.get(), hipEventCreate(&c->ev_start));
  GF_HIP_CHECK(c.get(), hipEventCreate(&c->ev_stop));
  GF_HIP_CHECK(c.get(), upload_element_tables());
  int st = ensure_partials(c.get(), 4 * 8192);
  if (st != GF_OK) return st;
  GF_HIP_CHECK(c.get(), hipMalloc(&c->d_state, 256));
  c->state_cap = 256;
  *out = c.release();
  return GF_OK;
}

int gf_destroy(gf_ctx* ctx) {
  if (!ctx) return GF_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_partials) (void)hipFree(ctx->d_partials);
  if (ctx->d_state) (void)hipFree(ctx->d_state);
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return GF_OK;
}

int gf_device_info(gf_ctx* ctx, int32_t* num_cu, char* name, int32_t name_len) {
  if (!ctx) return set_error(ctx, GF_ERR_ARG, "gf_device_info: null context");
  if (num_cu) *num_cu = ctx->num_cu;
  if (name && name_len > 0) {
    std::strncpy(name, ctx->name.c_str(), size_t(name_len) - 1);
    name[name_len - 1] = '\0';
  }
  return GF_OK;
}

// ---------------------------------------------------------------------------
int gf_mesh_upload(gf_ctx* ctx, const gf_mesh_desc* d, gf_mesh** out) {
  if (!ctx || !d || !out) return set_error(ctx, GF_ERR_ARG, "gf_mesh_upload: null argument");
  *out = nullptr;
  ElemInfo ei;
  if (!elem_info(d->elem_type, &ei))
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_mesh_upload: unknown element type");
  if (d->ngauss != 2 && d->ngauss != 3)
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_mesh_upload: ngauss must be 2 or 3");
  if (d->nnodes <= 0 || d->nelem <= 0 || !d->conn || !d->coords)
    return set_error(ctx, GF_ERR_ARG, "gf_mesh_upload: empty mesh or null arrays");
  if (d->nnodes * ei.dim >= (int64_t(1) << 31))
    return set_error(ctx, GF_ERR_UNSUPPORTED, "gf_mesh_upload: DOF count exceeds int32");
  if (!(d->nu > -1.0 && d->nu < 0.5))
    return set_error(ctx, GF_ERR_ARG, "gf_mesh_upload: Poisson ratio outside (-1, 0.5)");
  for (int64_t i = 0; i < d->nelem * ei.nen; ++i)
    if (d->conn[i] < 1 || d->conn[i] > d->nnodes)
      return set_error(ctx, GF_ERR_ARG, "gf_mesh_upload: connectivity entry out of range");
  int st = select_device(ctx);
  if (st) return st;
  std::unique_ptr<gf_mesh> m(new gf_mesh());
  m->type = d->elem_type;
  m->dim = ei.dim;
  m->nen = ei.nen;
  m->ngauss = d->ngauss;
  m->nnodes = d->nnodes;
  m->nelem = d->nelem;
  m->E = d->E;
  m->nu = d->nu;
  m->dev = ctx->device;
  m->h_conn.assign(d->conn, d->conn + d->nelem * ei.nen);
  const size_t cb = size_t(d->nelem) * ei.nen * sizeof(int32_t);
  const size_t xb = size_t(d->nnodes) * ei.dim * sizeof(double);
  auto fail = [&](int code) {
    dfree(m->d_conn);
    dfree(m->d_coords);
    dfree(m->d_E);
    return code;
  };
  hipError_t e;
  if ((e = hipMalloc(&m->d_conn, cb)) != hipSuccess ||
      (e = hipMalloc(&m->d_coords, xb)) != hipSuccess)
    return fail(set_error(ctx, GF_ERR_NOMEM, std::string("gf_mesh_upload: ") + hipGetErrorString(e)));
  if ((e = hipMemcpy(m->d_conn, d->conn, cb, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->d_coords, d->coords, xb, hipMemcpyHostToDevice)) != hipSuccess)
    return fail(set_error(ctx, GF_ERR_HIP, std::string("gf_mesh_upload: ") + hipGetErrorString(e)));
  if (d->elem_E) {
    const size_t eb = size_t(d->nelem) * sizeof(double);
    if ((e = hipMalloc(&m->d_E, eb)) != hipSuccess)
      return fail(set_error(ctx, GF_ERR_NOMEM, "gf_mesh_upload: E allocation"));
    if ((e = hipMemcpy(m->d_E, d->elem_E, eb, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(set_error(ctx, GF_ERR_HIP, "gf_mesh_upload: E copy"));
  }
  *out = m.release();
  return GF_OK;
}

int gf_mesh_set_E(gf_ctx* ctx, gf_mesh* m, const double* elem_E) {
  if (!ctx || !m || !elem_E) return set_error(ctx, GF_ERR_ARG, "gf_mesh_set_E: null argument");
  int st = select_device(ctx);
  if (st) return st;
  const size_t eb = size_t(m->nelem) * sizeof(double);
  if (!m->d_E) GF_HIP_CHECK(ctx, hipMalloc(&m->d_E, eb));
  GF_HIP_CHECK(ctx, hipMemcpyAsync(m->d_E, elem_E, eb, hipMemcpyHostToDevice, ctx->stream));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return GF_OK;
}

int gf_mesh_free(gf_mesh* m) {
  if (!m) return GF_OK;
  (void)hipSetDevice(m->dev);
  dfree(m->d_conn);
  dfree(m->d_coords);
  dfree(m->d_E);
  delete m;
  return GF_OK;
}

// ---------------------------------------------------------------------------
static void ell_release(gf_ell* m) {
  dfree(m->d_val);
  dfree(m->d_col);
  dfree(m->d_coln);
  dfree(m->d_bperm);
  dfree(m->d_binv);
  dfree(m->d_bsptr);
  dfree(m->d_bcol);
  dfree(m->d_bcoff);
  dfree(m->d_bval);
  dfree(m->d_bval32);
  dfree(m->d_rownnz);
  dfree(m->d_diag);
  dfree(m->d_color_elems);
  dfree(m->d_slot);
  dfree(m->d_n2e_ptr);
  dfree(m->d_n2e);
  dfree(m->d_egrad);
  dfree(m->d_etens);
  dfree(m->d_val32);
  dfree(m->d_dinv);
  release_solve_plan(m);
}

int gf_ell_symbolic(gf_ctx* ctx, const gf_mesh* mesh, gf_ell** out, gf_timing* t) {
  if (!ctx || !mesh || !out) return set_error(ctx, GF_ERR_ARG, "gf_ell_symbolic: null argument");
  *out = nullptr;
  int st = select_device(ctx);
  if (st) return st;
  const double h0 = now_ms();
  SymbolicHost S;
  std::string err;
  st = build_symbolic(mesh->h_conn.data(), mesh->nelem, mesh->nen, mesh->nnodes, &S, &err);
  if (st != GF_OK) return set_error(ctx, st, "gf_ell_symbolic: " + err);
  std::unique_ptr<gf_ell> m(new gf_ell());
  m->dev = ctx->device;
  m->dim = mesh->dim;
  m->n = mesh->nnodes * mesh->dim;
  m->width = S.max_adj * mesh->dim;
  m->nnz = int64_t(S.adj.size()) * mesh->dim * mesh->dim;
  m->ncolors = S.ncolors;
  m->color_ptr = S.color_ptr;
  m->h_elem_color = S.elem_color;
  m->nelem = mesh->nelem;
  m->nen = mesh->nen;
  const size_t slots = size_t(m->n) * size_t(m->width);
  int64_t* d_adj_ptr = nullptr;
  int32_t* d_adj = nullptr;
  auto fail = [&](int code, const std::string& msg) {
    dfree(d_adj_ptr);
    dfree(d_adj);
    ell_release(m.get());
    return set_error(ctx, code, "gf_ell_symbolic: " + msg);
  };
  hipError_t e;
#define TRY(expr, code)                                   \
  if ((e = (expr)) != hipSuccess) return fail(code, std::string(#expr) + ": " + hipGetErrorString(e))
  TRY(hipMalloc(&m->d_val, slots * sizeof(double)), GF_ERR_NOMEM);
  TRY(hipMalloc(&m->d_col, slots * sizeof(int32_t)), GF_ERR_NOMEM);
  TRY(hipMalloc(&m->d_rownnz, size_t(m->n) * sizeof(int32_t)), GF_ERR_NOMEM);
  TRY(hipMalloc(&m->d_diag, size_t(m->n) * sizeof(int32_t)), GF_ERR_NOMEM);
  TRY(hipMalloc(&m->d_color_elems, S.color_elems.size() * sizeof(int32_t)), GF_ERR_NOMEM);
  TRY(hipMalloc(&m->d_slot, S.slot.size()), GF_ERR_NOMEM);
  TRY(hipMalloc(&d_adj_ptr, S.adj_ptr.size() * sizeof(int64_t)), GF_ERR_NOMEM);
  TRY(hipMalloc(&d_adj, S.adj.size() * sizeof(int32_t)), GF_ERR_NOMEM);
  TRY(hipMemcpyAsync(d_adj_ptr, S.adj_ptr.data(), S.adj_ptr.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice, ctx->stream), GF_ERR_HIP);
  TRY(hipMemcpyAsync(d_adj, S.adj.data(), S.adj.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice, ctx->stream), GF_ERR_HIP);
  TRY(hipMemcpyAsync(m->d_color_elems, S.color_elems.data(),
                     S.color_elems.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                     ctx->stream), GF_ERR_HIP);
  TRY(hipMemcpyAsync(m->d_slot, S.slot.data(), S.slot.size(), hipMemcpyHostToDevice,
                     ctx->stream), GF_ERR_HIP);
  TRY(hipEventRecord(ctx->ev_start, ctx->stream), GF_ERR_HIP);
  TRY(launch_ell_expand(ctx->stream, m->n, m->width, m->dim, d_adj_ptr, d_adj, m->d_val,
                        m->d_col, m->d_rownnz, m->d_diag), GF_ERR_HIP);
  TRY(hipEventRecord(ctx->ev_stop, ctx->stream), GF_ERR_HIP);
  TRY(hipStreamSynchronize(ctx->stream), GF_ERR_HIP);
#undef TRY
  float f = 0.f;
  (void)hipEventElapsedTime(&f, ctx->ev_start, ctx->ev_stop);
  dfree(d_adj_ptr);
  dfree(d_adj);
  if (t) {
    t->ms_device = f;
    t->ms_host = now_ms() - h0;
    t->launches = 2;
  }
  *out = m.release();
  return GF_OK;
}

int gf_ell_symbolic_device(gf_ctx* ctx, const gf_mesh* mesh, gf_ell** out, gf_timing* t) {
  if (!ctx || !mesh || !out)
    return set_error(ctx, GF_ERR_ARG, "gf_ell_symbolic_device: null argument");
  *out = nullptr;
  int st = select_device(ctx);
  if (st) return st;
  const double h0 = now_ms();
  SymbolicDevice S;
  std::string err;
  hipError_t e = hipEventRecord(ctx->ev_start, ctx->stream);
  if (e != hipSuccess)
    return set_error(ctx, GF_ERR_HIP, std::string("gf_ell_symbolic_device: ") + hipGetErrorString(e));
  st = build_symbolic_device(ctx->stream, mesh->d_conn, mesh->nelem, mesh->nen, mesh->nnodes, &S,
                             &err);
  std::unique_ptr<gf_ell> m(new gf_ell());
  auto fail = [&](int code, const std::string& msg) {
    dfree(S.d_adj_ptr);
    dfree(S.d_adj);
    dfree(S.d_slot);
    dfree(S.d_color_elems);
    m->d_slot = nullptr;
    m->d_color_elems = nullptr;
    ell_release(m.get());
    return set_error(ctx, code, "gf_ell_symbolic_device: " + msg);
  };
  if (st != GF_OK) return fail(st, err);
  m->dev = ctx->device;
  m->dim = mesh->dim;
  m->n = mesh->nnodes * mesh->dim;
  m->width = S.max_adj * mesh->dim;
  m->nnz = S.nadj * mesh->dim * mesh->dim;
  m->ncolors = S.ncolors;
  m->color_ptr = S.color_ptr;
  m->h_elem_color = std::move(S.elem_color);
  m->nelem = mesh->nelem;
  m->nen = mesh->nen;
  m->d_slot = S.d_slot;
  m->d_color_elems = S.d_color_elems;
  const size_t slots = size_t(m->n) * size_t(m->width);
#define TRY(expr, code)                                   \
  if ((e = (expr)) != hipSuccess) return fail(code, std::string(#expr) + ": " + hipGetErrorString(e))
  TRY(hipMalloc(&m->d_val, slots * sizeof(double)), GF_ERR_NOMEM);
  TRY(hipMalloc(&m->d_col, slots * sizeof(int32_t)), GF_ERR_NOMEM);
  TRY(hipMalloc(&m->d_rownnz, size_t(m->n) * sizeof(int32_t)), GF_ERR_NOMEM);
  TRY(hipMalloc(&m->d_diag, size_t(m->n) * sizeof(int32_t)), GF_ERR_NOMEM);
  TRY(launch_ell_expand(ctx->stream, m->n, m->width, m->dim, S.d_adj_ptr, S.d_adj, m->d_val,
                        m->d_col, m->d_rownnz, m->d_diag), GF_ERR_HIP);
  TRY(hipEventRecord(ctx->ev_stop, ctx->stream), GF_ERR_HIP);
  TRY(hipStreamSynchronize(ctx->stream), GF_ERR_HIP);
#undef TRY
  float f = 0.f;
  (void)hipEventElapsedTime(&f, ctx->ev_start, ctx->ev_stop);
  dfree(S.d_adj_ptr);
  dfree(S.d_adj);
  if (t) {
    t->ms_device = f;
    t->ms_host = now_ms() - h0;
    t->launches = S.color_rounds * 2 + 12;
  }
  *out = m.release();
  return GF_OK;
}

int gf_ell_info_get(const gf_ell* m, gf_ell_info* info) {
  if (!m || !info) return set_error((gf_ctx*)nullptr, GF_ERR_ARG, "gf_ell_info_get: null argument");
  info->n = m->n;
  info->width = m->width;
  info->ncolors = m->ncolors;
  info->nnz = m->nnz;
  int64_t slots = m->n * int64_t(m->width);
  info->bytes = slots * int64_t(sizeof(double) + sizeof(int32_t)) + m->n * 8 +
                int64_t(m->h_elem_color.size()) * 4 + m->nelem * m->nen * m->nen;
  return GF_OK;
}

int gf_ell_colors(gf_ctx* ctx, const gf_ell* m, int32_t* elem_color) {
  if (!m || !elem_color) return set_error(ctx, GF_ERR_ARG, "gf_ell_colors: null argument");
  if (m->h_elem_color.empty())
    return set_error(ctx, GF_ERR_ARG, "gf_ell_colors: matrix has no element coloring");
  std::memcpy(elem_color, m->h_elem_color.data(), m->h_elem_color.size() * sizeof(int32_t));
  return GF_OK;
}

int gf_ell_free(gf_ell* m) {
  if (!m) return GF_OK;
  (void)hipSetDevice(m->dev);
  ell_release(m);
  delete m;
  return GF_OK;
}

int gf_ell_upload(gf_ctx* ctx, int64_t n, int32_t width, const double* val,
                  const int32_t* col, const int32_t* rownnz, gf_ell** out) {
  if (!ctx || !out || !val || !col || !rownnz || n <= 0 || width <= 0)
    return set_error(ctx, GF_ERR_ARG, "gf_ell_upload: bad argument");
  *out = nullptr;
  std::vector<int32_t> diag(size_t(n), -1);
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (rownnz[i] < 0 || rownnz[i] > width)
      return set_error(ctx, GF_ERR_ARG, "gf_ell_upload: rownnz outside [0, width]");
    nnz += rownnz[i];
    for (int32_t s = 0; s < rownnz[i]; ++s) {
      const int32_t c = col[i + int64_t(s) * n];
      if (c < 0 || c >= n) return set_error(ctx, GF_ERR_ARG, "gf_ell_upload: column out of range");
      if (c == i) diag[size_t(i)] = s;
    }
  }
  int st = select_device(ctx);
  if (st) return st;
  std::unique_ptr<gf_ell> m(new gf_ell());
  m->dev = ctx->device;
  m->n = n;
  m->width = width;
  m->nnz = nnz;
  m->dim = 1;
  // padding slots are rewritten as (val 0, col = row) so every gather stays in bounds
  std::vector<double> hv(val, val + n * int64_t(width));
  std::vector<int32_t> hc(col, col + n * int64_t(width));
  for (int64_t i = 0; i < n; ++i)
    for (int32_t s = rownnz[i]; s < width; ++s) {
      hv[size_t(i + int64_t(s) * n)] = 0.0;
      hc[size_t(i + int64_t(s) * n)] = int32_t(i);
    }
  const size_t slots = size_t(n) * size_t(width);
  hipError_t e;
  if ((e = hipMalloc(&m->d_val, slots * 8)) != hipSuccess ||
      (e = hipMalloc(&m->d_col, slots * 4)) != hipSuccess ||
      (e = hipMalloc(&m->d_rownnz, size_t(n) * 4)) != hipSuccess ||
      (e = hipMalloc(&m->d_diag, size_t(n) * 4)) != hipSuccess) {
    ell_release(m.get());
    return set_error(ctx, GF_ERR_NOMEM, "gf_ell_upload: allocation failed");
  }
  if ((e = hipMemcpy(m->d_val, hv.data(), slots * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->d_col, hc.data(), slots * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->d_rownnz, rownnz, size_t(n) * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->d_diag, diag.data(), size_t(n) * 4, hipMemcpyHostToDevice)) != hipSuccess) {
    ell_release(m.get());
    return set_error(ctx, GF_ERR_HIP, std::string("gf_ell_upload: ") + hipGetErrorString(e));
  }
  *out = m.release();
  return GF_OK;
}

int gf_ell_download(gf_ctx* ctx, const gf_ell* m, double* val, int32_t* col,
                    int32_t* rownnz) {
  if (!ctx || !m) return set_error(ctx, GF_ERR_ARG, "gf_ell_download: null argument");
  int st = select_device(ctx);
  if (st) return st;
  const size_t slots = size_t(m->n) * size_t(m->width);
  if (val && (st = ensure_ell_vals(ctx, const_cast<gf_ell*>(m)))) return st;
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (val) GF_HIP_CHECK(ctx, hipMemcpy(val, m->d_val, slots * 8, hipMemcpyDeviceToHost));
  if (col) GF_HIP_CHECK(ctx, hipMemcpy(col, m->d_col, slots * 4, hipMemcpyDeviceToHost));
  if (rownnz)
    GF_HIP_CHECK(ctx, hipMemcpy(rownnz, m->d_rownnz, size_t(m->n) * 4, hipMemcpyDeviceToHost));
  return GF_OK;
}

// ---------------------------------------------------------------------------
// node -> (element << 5 | local node) lists for the node-gather assembly,
// ascending element id per node (its order of summation)
static int ensure_n2e(gf_ctx* ctx, const gf_mesh* mesh, gf_ell* m) {
  if (m->d_n2e) return GF_OK;
  const int64_t nn = mesh->nnodes, ne = mesh->nelem;
  const int nen = mesh->nen;
  std::vector<int64_t> ptr(size_t(nn) + 1, 0);
  for (int64_t i = 0; i < ne * nen; ++i) ptr[size_t(mesh->h_conn[size_t(i)])]++;
  for (int64_t v = 0; v < nn; ++v) ptr[size_t(v) + 1] += ptr[size_t(v)];
  std::vector<int32_t> lst(static_cast<size_t>(ptr[size_t(nn)]));
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t e = 0; e < ne; ++e)  // ascending e: each node's list comes out sorted
    for (int a = 0; a < nen; ++a) {
      const int64_t v = mesh->h_conn[size_t(e * nen + a)] - 1;
      lst[size_t(fill[size_t(v)]++)] = int32_t((e << 5) | a);
    }
  GF_HIP_CHECK(ctx, hipMalloc(&m->d_n2e_ptr, ptr.size() * sizeof(int64_t)));
  GF_HIP_CHECK(ctx, hipMalloc(&m->d_n2e, std::max<size_t>(lst.size(), 1) * sizeof(int32_t)));
  GF_HIP_CHECK(ctx, hipMemcpy(m->d_n2e_ptr, ptr.data(), ptr.size() * sizeof(int64_t),
                              hipMemcpyHostToDevice));
  GF_HIP_CHECK(ctx, hipMemcpy(m->d_n2e, lst.data(), lst.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice));
  return GF_OK;
}

// the node-gather kernel exists for the type and every node fits its slots
static bool gather_fits(const gf_mesh* mesh, const gf_ell* m) {
  const int cap = gather_max_slots(mesh->type);
  return cap > 0 && m->width / m->dim <= cap && mesh->nelem < (int64_t(1) << 26);
}


static int assemble_with(gf_ctx* ctx, const gf_mesh* mesh, gf_ell* m, int variant, gf_timing* t) {
  const double h0 = now_ms();
  const bool node_rows = variant == 2 || variant == 3 || variant == 4;  // no scatter
  if (node_rows) {
    int s3 = ensure_n2e(ctx, mesh, m);
    if (s3) return s3;
  }
  if (variant == 2 && !m->d_etens) {
    const int64_t cnt = elem_tensor_doubles(mesh->type, mesh->nelem);
    GF_HIP_CHECK(ctx, hipMalloc(&m->d_etens, size_t(std::max<int64_t>(cnt, 1)) * sizeof(double)));
  }
  if (variant == 4 && !m->d_egrad) {
    const int64_t cnt = gauss_geom_doubles(mesh->type, mesh->ngauss, mesh->nelem);
    GF_HIP_CHECK(ctx, hipMalloc(&m->d_egrad, size_t(std::max<int64_t>(cnt, 1)) * sizeof(double)));
  }
  int* d_bad = reinterpret_cast<int*>(static_cast<char*>(ctx->d_state) + 128);
  GF_HIP_CHECK(ctx, hipMemsetAsync(d_bad, 0, sizeof(int), ctx->stream));
  int st = record_start(ctx);
  if (st) return st;
  const size_t slots = size_t(m->n) * size_t(m->width);
  // the gathers write every live slot; the scatters
  // accumulate from zero
  if (!node_rows)
    GF_HIP_CHECK(ctx, hipMemsetAsync(m->d_val, 0, slots * sizeof(double), ctx->stream));
  AssemblyArgs a{};
  a.type = mesh->type;
  a.ngauss = mesh->ngauss;
  a.conn = mesh->d_conn;
  a.coords = mesh->d_coords;
  a.elemE = mesh->d_E;
  a.E = mesh->E;
  elastic_coeffs(mesh->dim, mesh->nu, a.c);
  a.slot = m->d_slot;
  a.val = m->d_val;
  a.n = m->n;
  a.bad_jac = d_bad;
  a.variant = variant;
  a.n2e_ptr = m->d_n2e_ptr;
  a.n2e = m->d_n2e;
  a.rownnz = m->d_rownnz;
  a.nnodes = mesh->nnodes;
  // a block-sliced copy of the values the solver reads is kept current
  // inside the assembly (and its time): the node gather writes it beside
  // the ELL values, the colored scatters are followed by the copy
  const bool bsell = m->d_bval && m->bval_version != 0;
  a.bval = bsell && node_rows ? m->d_bval : nullptr;
  // with the copy current the gather writes only the diagonal of the ELL
  // values (what the penalty and the Jacobi preconditioner read); the other
  // slots are materialised from the copy when something reads them
  a.diag = m->d_diag;
  a.ell_diag_only = bsell && node_rows;
  a.bsptr = m->d_bsptr;
  a.binv = m->d_binv;
  a.nelem = mesh->nelem;
  int64_t launches = node_rows ? 0 : 1;
  if (variant == 2) {
    GF_HIP_CHECK(ctx, launch_assembly_tensors(ctx->stream, a, m->d_etens));
    launches += 2;
  } else if (variant == 3) {
    GF_HIP_CHECK(ctx, launch_assembly_gather(ctx->stream, a));
    ++launches;
  } else if (variant == 4) {
    GF_HIP_CHECK(ctx, launch_assembly_geom(ctx->stream, a, m->d_egrad));
    launches += 2;
  } else {
    for (int c = 0; c < m->ncolors; ++c) {
      a.elems = m->d_color_elems + m->color_ptr[size_t(c)];
      a.nce = m->color_ptr[size_t(c) + 1] - m->color_ptr[size_t(c)];
      GF_HIP_CHECK(ctx, launch_assembly_color(ctx->stream, a));
      ++launches;
    }
  }
  if (bsell && !node_rows) {
    GF_HIP_CHECK(ctx, launch_bsell_vals(ctx->stream, bsell_view(m), m->dim, m->n, m->d_val,
                                        m->d_rownnz, m->d_bval));
    ++launches;
  }
  double ms = 0;
  if ((st = record_stop(ctx, &ms))) return st;
  m->ell_stale = a.ell_diag_only != 0;
  int bad = 0;
  GF_HIP_CHECK(ctx, hipMemcpy(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost));
  m->dinv_valid = false;
  m->val32_valid = false;
  ++m->val_version;
  if (bsell) m->bval_version = m->val_version;
  if (t) {
    t->ms_device = ms;
    t->ms_host = now_ms() - h0;
    t->launches = launches;
  }
  if (bad)
    return set_error(ctx, GF_ERR_DEGENERATE, "gf_assemble: an element has detJ <= 0");
  return GF_OK;
}

static int check_assembly_args(gf_ctx* ctx, const gf_mesh* mesh, gf_ell* m, const char* fn) {
  if (!ctx || !mesh || !m) return set_error(ctx, GF_ERR_ARG, std::string(fn) + ": null argument");
  if (m->ncolors == 0 || m->nelem != mesh->nelem || m->nen != mesh->nen ||
      m->n != mesh->nnodes * mesh->dim)
    return set_error(ctx, GF_ERR_ARG, std::string(fn) + ": matrix was not built for this mesh");
  return select_device(ctx);
}

int gf_assemble_tune(gf_ctx* ctx, const gf_mesh* mesh, gf_ell* m, int32_t force,
                     double* ms_per_variant, int32_t* chosen) {
  int st = check_assembly_args(ctx, mesh, m, "gf_assemble_tune");
  if (st) return st;
  if (force > 4 || (force == 2 && mesh->dim != 3))
    return set_error(ctx, GF_ERR_ARG, "gf_assemble_tune: variant must be 0, 1, 3, 4 or (3D) 2");
  if (force == 2 && !gather_fits(mesh, m))
    return set_error(ctx, GF_ERR_UNSUPPORTED,
                     "gf_assemble_tune: the matrix-core kernel does not fit this mesh");
  if (force == 3 && !gather_fits(mesh, m))
    return set_error(ctx, GF_ERR_UNSUPPORTED,
                     "gf_assemble_tune: the node-gather kernel does not fit this mesh");
  if (force == 4 && !gather_fits(mesh, m))
    return set_error(ctx, GF_ERR_UNSUPPORTED,
                     "gf_assemble_tune: the stored-geometry gather does not fit this mesh");
  if (ms_per_variant)
    for (int v = 0; v < GF_ASM_VARIANTS; ++v) ms_per_variant[v] = -1.0;
  if (force >= 0) {
    m->asm_variant = force;
    if (force != 4 && m->d_egrad) {  // the stored geometry is no longer needed
      GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      GF_HIP_CHECK(ctx, hipFree(m->d_egrad));
      m->d_egrad = nullptr;
    }
    if (force != 2 && m->d_etens) {  // nor the matrix-core kernel's tensors
      GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      GF_HIP_CHECK(ctx, hipFree(m->d_etens));
      m->d_etens = nullptr;
    }
  } else {
    double best = 1e300;
    int best_v = 0;
    // the matrix-core kernel (2) pays for the 20-node brick only: its element
    // tensors are 9 x 20 x 20 products over 27 Gauss points, where the
    // 8-node brick's 9 x 8 x 8 fill a quarter of a 16 x 16 tile and cost
    // more in tensor traffic than they save
    for (int v : {0, 1, 3, 2}) {
      if ((v == 3 || v == 2) && !gather_fits(mesh, m)) continue;
      if (v == 2 && mesh->type != GF_BRICK20) continue;
      if (v == 2 && !m->d_etens) {
        // the tensors must leave most of the device free: else not tried
        size_t free_b = 0, total_b = 0;
        const double need = double(elem_tensor_doubles(mesh->type, mesh->nelem)) * sizeof(double);
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || need > 0.5 * double(free_b)) continue;
      }
      gf_timing t1{}, t2{};
      if ((st = assemble_with(ctx, mesh, m, v, &t1)) || (st = assemble_with(ctx, mesh, m, v, &t2)))
        return st;
      const double ms = std::min(t1.ms_device, t2.ms_device);
      if (ms_per_variant) ms_per_variant[v] = ms;
      if (ms < best) {
        best = ms;
        best_v = v;
      }
    }
    m->asm_variant = best_v;
    if (best_v != 2 && m->d_etens) {
      GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      GF_HIP_CHECK(ctx, hipFree(m->d_etens));
      m->d_etens = nullptr;
    }
  }
  if (chosen) *chosen = m->asm_variant;
  return GF_OK;
}

int gf_assemble(gf_ctx* ctx, const gf_mesh* mesh, gf_ell* m, gf_timing* t) {
  int st = check_assembly_args(ctx, mesh, m, "gf_assemble");
  if (st) return st;
  if (m->asm_variant < 0) {
    // small meshes: not worth timing; large ones pick the faster kernel once
    if (m->nelem < 65536) m->asm_variant = 0;
    else if ((st = gf_assemble_tune(ctx, mesh, m, -1, nullptr, nullptr))) return st;
  }
  return assemble_with(ctx, mesh, m, m->asm_variant, t);
}

int gf_apply_penalty(gf_ctx* ctx, gf_ell* m, const int32_t* dofs, int64_t nd, double coef,
                     double* pen_out) {
  if (!ctx || !m || (nd > 0 && !dofs) || nd < 0)
    return set_error(ctx, GF_ERR_ARG, "gf_apply_penalty: bad argument");
  std::vector<int32_t> d(dofs, dofs + nd);
  for (int32_t v : d)
    if (v < 0 || v >= m->n) return set_error(ctx, GF_ERR_ARG, "gf_apply_penalty: DOF out of range");
  std::sort(d.begin(), d.end());
  std::vector<int32_t> uniq, counts;
  for (size_t i = 0; i < d.size();) {
    size_t j = i;
    while (j < d.size() && d[j] == d[i]) ++j;
    uniq.push_back(d[i]);
    counts.push_back(int32_t(j - i));
    i = j;
  }
  int st = select_device(ctx);
  if (st) return st;
  // a current block-sliced copy gets the same diagonal values
  const bool bsell = m->d_bval && m->bval_version == m->val_version;
  double* d_pen = reinterpret_cast<double*>(static_cast<char*>(ctx->d_state) + 192);
  GF_HIP_CHECK(ctx, launch_penalty_factor(ctx->stream, m->d_val, m->d_diag, m->n, coef,
                                          ctx->d_partials, 8192, d_pen));
  int32_t* d_buf = nullptr;
  if (!uniq.empty()) {
    GF_HIP_CHECK(ctx, hipMalloc(&d_buf, uniq.size() * 2 * sizeof(int32_t)));
    hipError_t e = hipMemcpyAsync(d_buf, uniq.data(), uniq.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_buf + uniq.size(), counts.data(), counts.size() * 4,
                         hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = launch_penalty_apply(ctx->stream, m->d_val, m->d_diag, m->n, d_buf,
                               d_buf + uniq.size(), int64_t(uniq.size()), d_pen,
                               bsell ? m->d_bval : nullptr, m->d_bsptr, m->d_binv, m->dim);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_buf);
    if (e != hipSuccess)
      return set_error(ctx, GF_ERR_HIP, std::string("gf_apply_penalty: ") + hipGetErrorString(e));
  }
  if (pen_out) {
    GF_HIP_CHECK(ctx, hipMemcpyAsync(pen_out, d_pen, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  }
  m->dinv_valid = false;
  m->val32_valid = false;
  ++m->val_version;
  if (bsell) m->bval_version = m->val_version;
  return GF_OK;
}

namespace {
// device copy of a host DOF list (checked against [0, n)); *out = nullptr
// for an empty list
int upload_dofs(gf_ctx* ctx, const char* what, const int32_t* dofs, int64_t nd, int64_t n,
                int32_t** out) {
  *out = nullptr;
  for (int64_t i = 0; i < nd; ++i)
    if (dofs[i] < 0 || dofs[i] >= n)
      return set_error(ctx, GF_ERR_ARG, std::string(what) + ": DOF out of range");
  if (nd == 0) return GF_OK;
  hipError_t e = hipMalloc(out, size_t(nd) * sizeof(int32_t));
  if (e == hipSuccess)
    e = hipMemcpyAsync(*out, dofs, size_t(nd) * sizeof(int32_t), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e != hipSuccess) {
    if (*out) (void)hipFree(*out);
    *out = nullptr;
    return set_error(ctx, GF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
  return GF_OK;
}
}  // namespace

int gf_apply_elimination(gf_ctx* ctx, gf_ell* m, const int32_t* dofs, int64_t nd) {
  if (!ctx || !m || (nd > 0 && !dofs) || nd < 0)
    return set_error(ctx, GF_ERR_ARG, "gf_apply_elimination: bad argument");
  int st = select_device(ctx);
  if (st) return st;
  if ((st = ensure_ell_vals(ctx, m))) return st;
  int32_t* d_dofs = nullptr;
  if ((st = upload_dofs(ctx, "gf_apply_elimination", dofs, nd, m->n, &d_dofs))) return st;
  uint8_t* d_fixed = nullptr;
  hipError_t e = hipMalloc(&d_fixed, size_t(m->n > 0 ? m->n : 1));
  if (e == hipSuccess) e = hipMemsetAsync(d_fixed, 0, size_t(m->n), ctx->stream);
  if (e == hipSuccess)
    e = launch_eliminate(ctx->stream, m->n, m->width, d_dofs, nd, d_fixed, m->d_col,
                         m->d_rownnz, m->d_val);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (d_fixed) (void)hipFree(d_fixed);
  if (d_dofs) (void)hipFree(d_dofs);
  if (e != hipSuccess)
    return set_error(ctx, GF_ERR_HIP, std::string("gf_apply_elimination: ") + hipGetErrorString(e));
  m->dinv_valid = false;
  m->val32_valid = false;
  ++m->val_version;
  return GF_OK;
}

int gf_zero_dofs(gf_ctx* ctx, double* v, int64_t n, int32_t nrhs, const int32_t* dofs,
                 int64_t nd) {
  if (!ctx || (nd > 0 && (!dofs || !v)) || nd < 0 || n < 0 || nrhs < 0)
    return set_error(ctx, GF_ERR_ARG, "gf_zero_dofs: bad argument");
  int st = select_device(ctx);
  if (st) return st;
  int32_t* d_dofs = nullptr;
  if ((st = upload_dofs(ctx, "gf_zero_dofs", dofs, nd, n, &d_dofs))) return st;
  hipError_t e = launch_zero_dofs(ctx->stream, v, n, nrhs, d_dofs, nd);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (d_dofs) (void)hipFree(d_dofs);
  if (e != hipSuccess)
    return set_error(ctx, GF_ERR_HIP, std::string("gf_zero_dofs: ") + hipGetErrorString(e));
  return GF_OK;
}

int gf_load_vector(gf_ctx* ctx, int32_t dim, int64_t nnodes, const double* loads, int64_t nload,
                   double* f) {
  if (!ctx || (dim != 2 && dim != 3) || nnodes < 0 || nload < 0 || (nload > 0 && !loads) ||
      (nnodes > 0 && !f) || nload > INT32_MAX)
    return set_error(ctx, GF_ERR_ARG, "gf_load_vector: bad argument");
  for (int64_t i = 0; i < nload; ++i) {
    const double node = loads[i * (dim + 1)];
    if (!(node >= 1.0 && node <= double(nnodes)))
      return set_error(ctx, GF_ERR_ARG, "gf_load_vector: load on a node outside 1..nnodes");
  }
  int st = select_device(ctx);
  if (st) return st;
  const int64_t n = nnodes * dim;
  double* d_loads = nullptr;
  int32_t* d_owner = nullptr;
  hipError_t e = hipSuccess;
  if (nload > 0) {
    e = hipMalloc(&d_loads, size_t(nload) * size_t(dim + 1) * sizeof(double));
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_loads, loads, size_t(nload) * size_t(dim + 1) * sizeof(double),
                         hipMemcpyHostToDevice, ctx->stream);
  }
  if (e == hipSuccess && n > 0) e = hipMalloc(&d_owner, size_t(n) * sizeof(int32_t));
  if (e == hipSuccess) e = launch_load_vector(ctx->stream, dim, n, d_loads, nload, d_owner, f);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (d_loads) (void)hipFree(d_loads);
  if (d_owner) (void)hipFree(d_owner);
  if (e != hipSuccess)
    return set_error(ctx, GF_ERR_HIP, std::string("gf_load_vector: ") + hipGetErrorString(e));
  return GF_OK;
}

int gf_spmv(gf_ctx* ctx, const gf_ell* m, const double* x, double* y, int32_t nrhs,
            gf_timing* t) {
  if (!ctx || !m || !x || !y || nrhs < 1)
    return set_error(ctx, GF_ERR_ARG, "gf_spmv: bad argument");
  int st = select_device(ctx);
  if (st) return st;
  const double h0 = now_ms();
  const int variant = m->plan && m->plan->variant >= 0 ? m->plan->variant : kThreadU8;
  // the node-column / block-sliced copies are caches of the matrix: refreshing
  // them does not change the matrix the caller sees
  if ((st = prepare_variant(ctx, const_cast<gf_ell*>(m), variant))) return st;
  const SpmvOperands o = spmv_operands(m, variant);
  const size_t vb = size_t(m->n) * size_t(nrhs) * sizeof(double);
  double *dx = nullptr, *dy = nullptr;
  GF_HIP_CHECK(ctx, hipMalloc(&dx, vb));
  hipError_t e = hipMalloc(&dy, vb);
  double ms = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(dx, x, vb, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev_start, ctx->stream);
  if (e == hipSuccess)
    e = launch_spmv(ctx->stream, variant, m->n, o.val, o.col, m->d_rownnz, dx, dy, nrhs, &o.bs);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y, dy, vb, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) {
    float f = 0.f;
    e = hipEventElapsedTime(&f, ctx->ev_start, ctx->ev_stop);
    ms = f;
  }
  (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (e != hipSuccess) return set_error(ctx, GF_ERR_HIP, std::string("gf_spmv: ") + hipGetErrorString(e));
  if (t) {
    t->ms_device = ms;
    t->ms_host = now_ms() - h0;
    t->launches = 1;
  }
  return GF_OK;
}

int gf_stream_copy(gf_ctx* ctx, void* dst, const void* src, int64_t bytes, int32_t nrep,
                   double* ms_avg) {
  if (!ctx || !dst || !src || bytes <= 0 || nrep < 1)
    return set_error(ctx, GF_ERR_ARG, "gf_stream_copy: bad argument");
  int st = select_device(ctx);
  if (st) return st;
  GF_HIP_CHECK(ctx, launch_stream_copy(ctx->stream, src, dst, bytes));  // warm
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  for (int i = 0; i < nrep; ++i) GF_HIP_CHECK(ctx, launch_stream_copy(ctx->stream, src, dst, bytes));
  GF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  GF_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_stop));
  float ms = 0.f;
  GF_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  if (ms_avg) *ms_avg = double(ms) / nrep;
  return GF_OK;
}

int gf_dev_alloc(gf_ctx* ctx, int64_t bytes, void** dptr) {
  if (!ctx || !dptr || bytes < 0) return set_error(ctx, GF_ERR_ARG, "gf_dev_alloc: bad argument");
  int st = select_device(ctx);
  if (st) return st;
  hipError_t e = hipMalloc(dptr, size_t(bytes > 0 ? bytes : 1));
  if (e != hipSuccess) return set_error(ctx, GF_ERR_NOMEM, std::string("gf_dev_alloc: ") + hipGetErrorString(e));
  return GF_OK;
}

int gf_dev_free(gf_ctx* ctx, void* dptr) {
  if (!ctx) return set_error(ctx, GF_ERR_ARG, "gf_dev_free: null context");
  if (dptr) GF_HIP_CHECK(ctx, hipFree(dptr));
  return GF_OK;
}

int gf_copy_d2d(gf_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || bytes < 0 || (bytes && (!dst || !src)))
    return set_error(ctx, GF_ERR_ARG, "gf_copy_d2d: bad argument");
  GF_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToDevice, ctx->stream));
  return GF_OK;
}

int gf_copy_h2d(gf_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || bytes < 0 || (bytes && (!dst || !src)))
    return set_error(ctx, GF_ERR_ARG, "gf_copy_h2d: bad argument");
  GF_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyHostToDevice, ctx->stream));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return GF_OK;
}

int gf_copy_d2h(gf_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx || bytes < 0 || (bytes && (!dst || !src)))
    return set_error(ctx, GF_ERR_ARG, "gf_copy_d2h: bad argument");
  GF_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToHost, ctx->stream));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return GF_OK;
}

}  // extern "C"
