// Host symbolic phase (see symbolic.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gf_internal.hpp"

namespace gf {

struct SymbolicHost {
  std::vector<int64_t> adj_ptr;      // [nnodes+1]
  std::vector<int32_t> adj;          // sorted 0-based node ids per node
  std::vector<int32_t> elem_color;   // [nelem], 1-based
  std::vector<int64_t> color_ptr;    // [ncolors+1]
  std::vector<int32_t> color_elems;  // element ids grouped by color, ascending
  std::vector<uint8_t> slot;         // [nelem][nen][nen]
  int32_t max_adj = 0;
  int ncolors = 0;
};

int build_symbolic(const int32_t* conn, int64_t nelem, int nen, int64_t nnodes,
                   SymbolicHost* out, std::string* err);

}  // namespace gf
