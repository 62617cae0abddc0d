// Solver workspace cached on an ELL matrix (see solve.cpp).
#pragma once

#include "gf_internal.hpp"

namespace gf {
void release_solve_plan(gf_ell* m);
}  // namespace gf
