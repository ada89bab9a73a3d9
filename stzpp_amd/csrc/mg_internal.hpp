// Library-internal access to a multigrid handle's level-0 device fields, so
// the step kernels can write the right-hand side and read the solution in
// place (no copies between the simulation state and the solver).
#pragma once

#include "../../include/stzmg.h"

namespace stzmg_internal {
double *x0(stzmg_t h);  // level-0 solution, SoA, 3 x m*n*o doubles
double *r0(stzmg_t h);  // level-0 source
}  // namespace stzmg_internal
