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

// The same structure built on the device (symbolic_dev.hip) from the
// device connectivity.  The device arrays belong to the caller (also on
// failure: free the non-null ones).
struct SymbolicDevice {
  int64_t* d_adj_ptr = nullptr;      // [nnodes+1]
  int32_t* d_adj = nullptr;          // sorted 0-based node ids per node
  uint8_t* d_slot = nullptr;         // [nelem][nen][nen]
  int32_t* d_color_elems = nullptr;  // element ids grouped by color, ascending
  std::vector<int32_t> elem_color;   // [nelem], 1-based (host copy)
  std::vector<int64_t> color_ptr;    // [ncolors+1]
  int64_t nadj = 0;
  int32_t max_adj = 0;
  int ncolors = 0;
  int64_t color_rounds = 0;  // first-fit rounds the coloring took
};

int build_symbolic_device(hipStream_t s, const int32_t* d_conn, int64_t nelem, int nen,
                          int64_t nnodes, SymbolicDevice* out, std::string* err);

}  // namespace gf
