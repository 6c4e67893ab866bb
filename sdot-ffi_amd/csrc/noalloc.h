// One diagram whose scratch is a fixed device array instead of allocated memory ( `sdot_amd_measures_2d_noalloc` ).
#pragma once

#include "cells.h"

namespace sdot_amd {

/// the largest cloud the fixed scratch serves, and the scratch's size per GPU ( the tree and its build ~165 B a seed, the
/// rank-order masses, overflow lists and zero weights ~24 B a seed, the overflow slots of `NOALLOC_OVF_LANES` lanes ~2.6 MB,
/// with room to spare )
constexpr I64    NOALLOC_MAX_SEEDS = I64( 1 ) << 17;
constexpr size_t NOALLOC_BYTES = size_t( 40 ) << 20;
/// the third pass's lanes ( each with `CAP3` vertex slots in the scratch )
constexpr int    NOALLOC_OVF_LANES = 64;

/// the masses of `sdot_amd_measures_2d` ( no facets ), with every buffer carved from the fixed scratch: nothing allocated
void measures_2d_noalloc( hipStream_t st, I64 n, const double *positions, const double *weights, const double *box, const Density &dn,
                          double *masses, unsigned long long *events );

} // namespace sdot_amd
