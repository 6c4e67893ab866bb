// The device memory a solve takes, per part, before it starts ( see memory.hip ).
#pragma once

#include "common.h"

namespace sdot_amd {

/// the bytes of each of the `SDOT_AMD_NB_MEMORY_PARTS` parts of a solve of `n` seeds in `dim` dimensions ( the order of
/// `sdot_amd_solve_bytes` ); `o` the 41 options ( may be null: the defaults ), `precision` the 2D cells' float
void solve_parts( I64 n, int dim, int density_kind, const double *o, int precision, double *parts );

/// their sum
double solve_bytes( I64 n, int dim, int density_kind, const double *o, int precision );

} // namespace sdot_amd
