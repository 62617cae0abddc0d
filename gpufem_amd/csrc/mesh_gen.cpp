// Structured synthetic meshes, consistent boundary loads and deterministic
// random streams (host only).
//
// Node numbering: every grid point the element type uses, in lexicographic
// order (x fastest, then y, then z), 1-based.  Linear types (Q4, Brick8) use
// the (nx+1)(ny+1)(nz+1) vertices; quadratic types (Q8, Brick20) use the
// points of the half-step lattice with at most one odd coordinate.
// Element e = (k*ny + j)*nx + i; its local nodes follow the reference node
// ordering (natural axis r -> x, s -> y, t -> z), which gives detJ > 0.
#include <cmath>
#include <cstring>
#include <vector>

#include "gf_internal.hpp"

namespace {

struct Grid {
  gf::ElemInfo ei;
  bool quadratic = false;
  int64_t nx = 0, ny = 0, nz = 0;
  int64_t px = 0, py = 0, pz = 0;  // half-step lattice extents (points)
  std::vector<int32_t> id;         // lattice point -> node id (0: unused)
  int64_t nnodes = 0, nelem = 0;
  double h = 1.0;

  int64_t lat(int64_t I, int64_t J, int64_t K) const { return (K * py + J) * px + I; }
};

int build_grid(const gf_grid_desc* g, Grid* G, bool with_ids) {
  if (!g || !gf::elem_info(g->elem_type, &G->ei)) return GF_ERR_ARG;
  if (g->nx < 1 || g->ny < 1 || (G->ei.dim == 3 && g->nz < 1)) return GF_ERR_ARG;
  if (!(g->h > 0)) return GF_ERR_ARG;
  G->quadratic = (g->elem_type == GF_Q8 || g->elem_type == GF_BRICK20);
  G->nx = g->nx;
  G->ny = g->ny;
  G->nz = G->ei.dim == 3 ? g->nz : 0;
  G->h = g->h;
  G->px = 2 * G->nx + 1;
  G->py = 2 * G->ny + 1;
  G->pz = G->ei.dim == 3 ? 2 * G->nz + 1 : 1;
  G->nelem = G->nx * G->ny * (G->ei.dim == 3 ? G->nz : 1);
  // count / number the used points
  int64_t total = G->px * G->py * G->pz;
  if (total > (int64_t(1) << 31) - 2) return GF_ERR_UNSUPPORTED;
  if (with_ids) G->id.assign(size_t(total), 0);
  int64_t next = 0;
  for (int64_t K = 0; K < G->pz; ++K)
    for (int64_t J = 0; J < G->py; ++J)
      for (int64_t I = 0; I < G->px; ++I) {
        int odd = int(I & 1) + int(J & 1) + int(K & 1);
        bool used = G->quadratic ? odd <= 1 : odd == 0;
        if (!used) continue;
        ++next;
        if (with_ids) G->id[size_t(G->lat(I, J, K))] = int32_t(next);
      }
  G->nnodes = next;
  return GF_OK;
}

// lattice position (half steps) of local node a of element (i,j,k)
inline void node_lattice(const Grid& G, const double* nat, int a, int64_t i, int64_t j,
                         int64_t k, int64_t* I, int64_t* J, int64_t* K) {
  int d = G.ei.dim;
  *I = 2 * i + 1 + int64_t(nat[a * d + 0]);
  *J = 2 * j + 1 + int64_t(nat[a * d + 1]);
  *K = d == 3 ? 2 * k + 1 + int64_t(nat[a * d + 2]) : 0;
}

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

}  // namespace

extern "C" {

int gf_grid_sizes(const gf_grid_desc* g, int64_t* nnodes, int64_t* nelem, int32_t* nen,
                  int32_t* dim) {
  Grid G;
  int st = build_grid(g, &G, false);
  if (st != GF_OK) return st;
  if (nnodes) *nnodes = G.nnodes;
  if (nelem) *nelem = G.nelem;
  if (nen) *nen = G.ei.nen;
  if (dim) *dim = G.ei.dim;
  return GF_OK;
}

int gf_grid_fill(const gf_grid_desc* g, int32_t* conn, double* coords) {
  Grid G;
  int st = build_grid(g, &G, true);
  if (st != GF_OK) return st;
  if (!conn || !coords) return GF_ERR_ARG;
  const int d = G.ei.dim, nen = G.ei.nen;
  const double half = 0.5 * G.h;
  for (int64_t K = 0; K < G.pz; ++K)
    for (int64_t J = 0; J < G.py; ++J)
      for (int64_t I = 0; I < G.px; ++I) {
        int32_t id = G.id[size_t(G.lat(I, J, K))];
        if (!id) continue;
        double* c = coords + int64_t(id - 1) * d;
        c[0] = double(I) * half;
        c[1] = double(J) * half;
        if (d == 3) c[2] = double(K) * half;
      }
  const double* nat = gf::elem_natural_coords(g->elem_type);
  const int64_t nz = d == 3 ? G.nz : 1;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < G.ny; ++j)
      for (int64_t i = 0; i < G.nx; ++i) {
        int64_t e = (k * G.ny + j) * G.nx + i;
        for (int a = 0; a < nen; ++a) {
          int64_t I, J, K;
          node_lattice(G, nat, a, i, j, k, &I, &J, &K);
          conn[e * nen + a] = G.id[size_t(G.lat(I, J, K))];
        }
      }
  return GF_OK;
}

int64_t gf_grid_face_nodes(const gf_grid_desc* g, int32_t face, int32_t* nodes) {
  Grid G;
  int st = build_grid(g, &G, true);
  if (st != GF_OK) return st;
  const int d = G.ei.dim;
  if (face < 0 || face >= 2 * d) return GF_ERR_ARG;
  int axis = face / 2;
  bool hi = face & 1;
  int64_t ext[3] = {G.px, G.py, G.pz};
  int64_t target = hi ? ext[axis] - 1 : 0;
  int64_t cnt = 0;
  for (int64_t K = 0; K < G.pz; ++K)
    for (int64_t J = 0; J < G.py; ++J)
      for (int64_t I = 0; I < G.px; ++I) {
        int64_t p[3] = {I, J, K};
        if (p[axis] != target) continue;
        int32_t id = G.id[size_t(G.lat(I, J, K))];
        if (!id) continue;
        if (nodes) nodes[cnt] = id;
        ++cnt;
      }
  return cnt;
}

int gf_grid_face_load(const gf_grid_desc* g, int32_t face, int32_t comp, double q,
                      double* f) {
  Grid G;
  int st = build_grid(g, &G, true);
  if (st != GF_OK) return st;
  const int d = G.ei.dim, nen = G.ei.nen;
  if (face < 0 || face >= 2 * d || comp < 0 || comp >= d || !f) return GF_ERR_ARG;
  const int axis = face / 2;
  const bool hi = face & 1;
  const double* nat = gf::elem_natural_coords(g->elem_type);
  const double measure = d == 2 ? G.h : G.h * G.h;
  const int64_t nel_ax[3] = {G.nx, G.ny, d == 3 ? G.nz : 1};
  const int64_t nz = d == 3 ? G.nz : 1;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < G.ny; ++j)
      for (int64_t i = 0; i < G.nx; ++i) {
        int64_t idx[3] = {i, j, k};
        if (idx[axis] != (hi ? nel_ax[axis] - 1 : 0)) continue;
        for (int a = 0; a < nen; ++a) {
          const double* an = nat + a * d;
          if (an[axis] != (hi ? 1.0 : -1.0)) continue;
          // weight of node a in the consistent load of the face
          bool corner = true;
          for (int c = 0; c < d; ++c)
            if (an[c] == 0.0) corner = false;
          double w;
          switch (g->elem_type) {
            case GF_Q4: w = 0.5; break;                        // linear edge
            case GF_Q8: w = corner ? 1.0 / 6.0 : 2.0 / 3.0; break;  // quadratic edge
            case GF_BRICK8: w = 0.25; break;                   // bilinear face
            default: w = corner ? -1.0 / 12.0 : 1.0 / 3.0; break;  // 8-node face
          }
          int64_t I, J, K;
          node_lattice(G, nat, a, i, j, k, &I, &J, &K);
          int32_t id = G.id[size_t(G.lat(I, J, K))];
          f[int64_t(id - 1) * d + comp] += w * q * measure;
        }
      }
  return GF_OK;
}

int gf_rng_normal(uint64_t seed, int64_t n, double* out) {
  if (n < 0 || (n > 0 && !out)) return GF_ERR_ARG;
  SplitMix64 rng(seed);
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; i += 2) {
    double u1 = (double((rng.next() >> 11) + 1)) * 0x1.0p-53;  // (0, 1]
    double u2 = double(rng.next() >> 11) * 0x1.0p-53;          // [0, 1)
    double r = std::sqrt(-2.0 * std::log(u1));
    out[i] = r * std::cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = r * std::sin(two_pi * u2);
  }
  return GF_OK;
}

int gf_grf(uint64_t seed, int64_t nx, int64_t ny, int64_t nz, double lengthscale,
           double* out) {
  if (nx < 1 || ny < 1 || nz < 1 || !(lengthscale > 0.0) || !out) return GF_ERR_ARG;
  // white noise on the grid extended by R cells per side along every axis
  // of extent > 1, smoothed by a separable Gaussian (std = lengthscale)
  // with per-axis weights scaled to unit sum of squares: every output cell
  // is a unit-variance combination of the same number of noise cells
  const int64_t R = int64_t(std::ceil(3.0 * lengthscale));
  const int64_t n_[3] = {nx, ny, nz};
  int64_t e_[3], r_[3];
  for (int a = 0; a < 3; ++a) {
    r_[a] = n_[a] > 1 ? R : 0;
    e_[a] = n_[a] + 2 * r_[a];
  }
  std::vector<double> w(size_t(2 * R + 1));
  double s2 = 0.0;
  for (int64_t k = -R; k <= R; ++k) {
    const double v = std::exp(-0.5 * double(k * k) / (lengthscale * lengthscale));
    w[size_t(k + R)] = v;
    s2 += v * v;
  }
  for (double& v : w) v /= std::sqrt(s2);
  std::vector<double> a(static_cast<size_t>(e_[0] * e_[1] * e_[2]));
  gf_rng_normal(seed, int64_t(a.size()), a.data());
  // pass along x: (e0, e1, e2) -> (n0, e1, e2)
  std::vector<double> b(static_cast<size_t>(nx * e_[1] * e_[2]), 0.0);
  for (int64_t k = 0; k < e_[2]; ++k)
    for (int64_t j = 0; j < e_[1]; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        double s = 0.0;
        for (int64_t t = -r_[0]; t <= r_[0]; ++t)
          s += w[size_t(t + R)] * a[size_t((k * e_[1] + j) * e_[0] + i + r_[0] + t)];
        b[size_t((k * e_[1] + j) * nx + i)] = r_[0] ? s : a[size_t((k * e_[1] + j) * e_[0] + i)];
      }
  // along y: -> (n0, n1, e2)
  std::vector<double> c(static_cast<size_t>(nx * ny * e_[2]), 0.0);
  for (int64_t k = 0; k < e_[2]; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        double s = 0.0;
        for (int64_t t = -r_[1]; t <= r_[1]; ++t)
          s += w[size_t(t + R)] * b[size_t((k * e_[1] + j + r_[1] + t) * nx + i)];
        c[size_t((k * ny + j) * nx + i)] = r_[1] ? s : b[size_t((k * e_[1] + j) * nx + i)];
      }
  // along z: -> (n0, n1, n2)
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        double s = 0.0;
        for (int64_t t = -r_[2]; t <= r_[2]; ++t)
          s += w[size_t(t + R)] * c[size_t(((k + r_[2] + t) * ny + j) * nx + i)];
        out[size_t((k * ny + j) * nx + i)] = r_[2] ? s : c[size_t((k * ny + j) * nx + i)];
      }
  return GF_OK;
}

int gf_rng_uniform_int(uint64_t seed, int64_t n, int32_t lo, int32_t hi, int32_t* out) {
  if (n < 0 || hi < lo || (n > 0 && !out)) return GF_ERR_ARG;
  SplitMix64 rng(seed);
  uint64_t range = uint64_t(int64_t(hi) - int64_t(lo) + 1);
  for (int64_t i = 0; i < n; ++i) {
    uint64_t r = rng.next() >> 32;
    out[i] = int32_t(int64_t(lo) + int64_t((r * range) >> 32));
  }
  return GF_OK;
}

}  // extern "C"
