// Host symbolic phase: node adjacency (the ELL structure is node adjacency
// x dim^2), greedy element coloring, and the element -> ELL slot map used by
// the atomic-free scatter.
//
// Structure: row dim*v+c holds columns dim*w+c' for every node w that shares
// an element with v (v included), sorted ascending -- the pattern the
// reference's AddElem insertion produces (src/SPRmatrix/ELLmatrix.cc:91-138)
// when no contribution is dropped as an exact zero.
//
// Coloring: first-fit in element order over the element dual graph with
// "shares at least one node" adjacency (METIS_MeshToDual with ncommon = 1 in
// src/FEM/femColor.cc:103-107); a color is the smallest positive integer not
// used by an already-colored neighbour (src/FEM/femColor.cc:132-191).
#include <algorithm>
#include <cstring>

#include "symbolic.hpp"

namespace gf {

int build_symbolic(const int32_t* conn, int64_t nelem, int nen, int64_t nnodes,
                   SymbolicHost* S, std::string* err) {
  const int T = host_threads();
  // node -> incident elements
  std::vector<int64_t> n2e_ptr(size_t(nnodes) + 1, 0);
  for (int64_t i = 0; i < nelem * nen; ++i) {
    int32_t v = conn[i];
    if (v < 1 || v > nnodes) {
      *err = "connectivity references node " + std::to_string(v) + " outside 1.." +
             std::to_string(nnodes);
      return GF_ERR_ARG;
    }
    n2e_ptr[size_t(v)]++;
  }
  for (int64_t v = 0; v < nnodes; ++v) n2e_ptr[size_t(v) + 1] += n2e_ptr[size_t(v)];
  std::vector<int32_t> n2e(static_cast<size_t>(n2e_ptr[size_t(nnodes)]));
  {
    std::vector<int64_t> fill(n2e_ptr.begin(), n2e_ptr.end() - 1);
    for (int64_t e = 0; e < nelem; ++e)
      for (int a = 0; a < nen; ++a) {
        int64_t v = conn[e * nen + a] - 1;
        n2e[size_t(fill[size_t(v)]++)] = int32_t(e);
      }
  }

  // node adjacency: counts, then lists
  S->adj_ptr.assign(size_t(nnodes) + 1, 0);
  auto node_list = [&](int64_t v, std::vector<int32_t>& buf) {
    buf.clear();
    for (int64_t k = n2e_ptr[size_t(v)]; k < n2e_ptr[size_t(v) + 1]; ++k) {
      const int32_t* c = conn + int64_t(n2e[size_t(k)]) * nen;
      for (int a = 0; a < nen; ++a) buf.push_back(c[a] - 1);
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  parallel_for(nnodes, T, [&](int64_t b, int64_t e) {
    std::vector<int32_t> buf;
    for (int64_t v = b; v < e; ++v) {
      node_list(v, buf);
      if (buf.empty()) buf.push_back(int32_t(v));  // isolated node: diagonal only
      S->adj_ptr[size_t(v) + 1] = int64_t(buf.size());
    }
  });
  int32_t maxadj = 0;
  for (int64_t v = 0; v < nnodes; ++v) {
    maxadj = std::max<int32_t>(maxadj, int32_t(S->adj_ptr[size_t(v) + 1]));
    S->adj_ptr[size_t(v) + 1] += S->adj_ptr[size_t(v)];
  }
  S->max_adj = maxadj;
  S->adj.resize(size_t(S->adj_ptr[size_t(nnodes)]));
  parallel_for(nnodes, T, [&](int64_t b, int64_t e) {
    std::vector<int32_t> buf;
    for (int64_t v = b; v < e; ++v) {
      node_list(v, buf);
      if (buf.empty()) buf.push_back(int32_t(v));
      std::copy(buf.begin(), buf.end(), S->adj.begin() + S->adj_ptr[size_t(v)]);
    }
  });

  // greedy first-fit coloring in element order
  S->elem_color.assign(size_t(nelem), 0);
  int ncolors = 0;
  {
    std::vector<int32_t> mark(64, -1);  // mark[c] == e: color c used around e
    for (int64_t e = 0; e < nelem; ++e) {
      const int32_t* c = conn + e * nen;
      for (int a = 0; a < nen; ++a) {
        int64_t v = c[a] - 1;
        for (int64_t k = n2e_ptr[size_t(v)]; k < n2e_ptr[size_t(v) + 1]; ++k) {
          int32_t f = n2e[size_t(k)];
          int32_t col = S->elem_color[size_t(f)];
          if (f == e || col == 0) continue;
          if (size_t(col) >= mark.size()) mark.resize(size_t(col) * 2, -1);
          mark[size_t(col)] = int32_t(e);
        }
      }
      int32_t color = 1;
      while (size_t(color) < mark.size() && mark[size_t(color)] == int32_t(e)) ++color;
      S->elem_color[size_t(e)] = color;
      ncolors = std::max(ncolors, int(color));
    }
  }
  S->ncolors = ncolors;
  S->color_ptr.assign(size_t(ncolors) + 1, 0);
  for (int64_t e = 0; e < nelem; ++e) S->color_ptr[size_t(S->elem_color[size_t(e)])]++;
  for (int c = 0; c < ncolors; ++c) S->color_ptr[size_t(c) + 1] += S->color_ptr[size_t(c)];
  S->color_elems.resize(size_t(nelem));
  {
    std::vector<int64_t> fill(S->color_ptr.begin(), S->color_ptr.end() - 1);
    for (int64_t e = 0; e < nelem; ++e) {
      int c = S->elem_color[size_t(e)] - 1;
      S->color_elems[size_t(fill[size_t(c)]++)] = int32_t(e);
    }
  }

  // element -> slot map: rank of node b in the sorted adjacency of node a
  if (maxadj > 255) {
    *err = "a node has " + std::to_string(maxadj) +
           " neighbours; the slot map stores ranks in 8 bits";
    return GF_ERR_UNSUPPORTED;
  }
  S->slot.resize(size_t(nelem) * nen * nen);
  parallel_for(nelem, T, [&](int64_t b, int64_t e_end) {
    for (int64_t e = b; e < e_end; ++e) {
      const int32_t* c = conn + e * nen;
      for (int a = 0; a < nen; ++a) {
        int64_t v = c[a] - 1;
        const int32_t* lo = S->adj.data() + S->adj_ptr[size_t(v)];
        const int32_t* hi = S->adj.data() + S->adj_ptr[size_t(v) + 1];
        for (int bb = 0; bb < nen; ++bb) {
          int32_t w = c[bb] - 1;
          const int32_t* p = std::lower_bound(lo, hi, w);
          S->slot[size_t((e * nen + a) * nen + bb)] = uint8_t(p - lo);
        }
      }
    }
  });
  return GF_OK;
}

}  // namespace gf
