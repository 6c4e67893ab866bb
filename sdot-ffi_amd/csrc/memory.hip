// WHAT A SOLVE TAKES FROM THE DEVICE, buffer by buffer, before it starts: the model the server admits requests by and the
// Python side checks against the free memory ( the role of the reference's `src/sdot/CardMemory.py:47-190`,
// `card_solve_bytes`, for this library's own buffers ). It follows the arena's takes of `solve_2d` ( solver.hip ), the cells
// ( cells.hip, cells3.hip ), the trees ( tree_gpu.hip, tree3.hip ), the multigrid ( multigrid.hip ) and the step by the limits
// ( limits.hip ), with the sizes they ask for on a generic cloud; a diagram whose rows overflow their first room takes more.

#include "../../include/sdot_amd.h"
#include "memory.h"
#include "solver.h"
#include "tree.h"
#include "cells3.h"

#include <cmath>

namespace sdot_amd {

namespace {
constexpr double MG_STRIDE_B = 96;        ///< entries of a coarse row ( multigrid.hip )
constexpr double RS_TILE_B = 2048;        ///< the radix sort's tile ( tree_gpu.hip )
double pad( double b ) { return std::ceil( std::max( b, 256.0 ) / 256.0 ) * 256.0; }
}

void solve_parts( I64 n, int dim, int density_kind, const double *o, int precision, double *parts ) {
    for ( int q = 0; q < SDOT_AMD_NB_MEMORY_PARTS; ++q ) parts[ q ] = 0;
    const double N = double( ( I64( 1 ) << max_depth_for( n, 10 ) ) - 1 ), half = ( N + 1 ) / 2, m = double( n );
    const bool d3 = dim == 3;
    // the tree: records, boxes, slices, the level table, the permutation, the seeds, the majorants' moments; and its device
    // build's sort keys, sorted lists, segments and flags
    if ( d3 ) parts[ 0 ] = 96 * N + 48 * N + 12 * N + 8 * m + 32 * m + ( 16 + 3 + 3 ) * 8 * N +
                           ( 16 + 8 + 24 + 4 + 1 + 8 ) * m + 1024 * std::ceil( m / RS_TILE_B ) + 24 * half + 8 * N + 48 * N;
    else      parts[ 0 ] = 64 * N + 32 * N + 32 * N + 12 * N + 8 * m + 32 * m + ( 11 + 2 + 3 ) * 8 * N +
                           ( 16 + 24 + 4 + 1 + 8 ) * m + 1024 * std::ceil( m / RS_TILE_B ) + 24 * half + 8 * N + 32 * N;
    // two diagram slots: masses, row offsets and counts, diagonals, the rows ( 12 n + 4096 entries, 3D: 44 per seed ), the
    // overflow lists and counters
    const double cap = d3 ? std::ceil( m / 64 ) * 64 * 44 : ( ( o && o[ O_AGG_NB_DUPS ] > 0 ) ? 14 : 12 ) * m + 4096;
    parts[ 1 ] = 2 * ( 24 * m + 12 * cap + 8 * m + 32 );
    // the step by the limits ( 2D ): each cell's cuts and their count, per slot
    const bool limits = ! d3 && o && int( o[ O_STEP ] ) == 1;
    if ( limits ) parts[ 2 ] = 2 * ( EDGE_CAP * 4 * m + 4 * m );
    // the FP32 kernel's polygons ( 2D, 'fp32' and 'mixed' ); 3D with a sum of Gaussians: each cell's polytope
    if ( ! d3 && precision != 0 ) parts[ 3 ] = CAPW * 12 * m + 4 * m + 8 * m;   // ( and the FP64 redo's marks and list )
    if ( d3 && ( density_kind == SDOT_AMD_GAUSSIANS || density_kind == SDOT_AMD_IMAGE ) ) parts[ 3 ] = 64 * m * ( 3 * 8 + 2 * 4 + 3 * 8 + 4 ) + 8 * m;
    // the third pass's vertex slots ( a fixed pool )
    parts[ 4 ] = d3 ? 0 : double( OVF_SLOTS ) * CAP3 * 20;
    // Newton's vectors ( targets, weights and trial weights, direction, right-hand side, staging ), the moments' diagram
    parts[ 5 ] = 7 * 8 * m;
    parts[ 6 ] = ( d3 ? 24 : 16 ) * m + 16 * m;
    // the linear solver: the multigrid's levels ( coarse rows of 96 entries, their vectors, the K-cycle's ), its outer
    // vectors; Jacobi-PCG's when asked for
    if ( o && int( o[ O_LIN ] ) == 1 ) {
        const int pack = o[ O_MG_SHIFT ] > 0 ? 1 << int( o[ O_MG_SHIFT ] ) : 8;
        const int kc = o[ O_MG_KCYCLE ] >= 0 ? int( o[ O_MG_KCYCLE ] ) : 2;
        // the coarse levels in `s`-byte values; the FP64 flexible CG's four vectors and the finest level's work vector; with the
        // FP32 preconditioner ( `mg_float` ) its own levels, work vector, FP32 copies of the rows, diagonal, `r` and `z`, beside
        // the FP64 solve it falls back to
        const bool f32 = o[ O_MG_FLOAT ] != 0;
        auto coarse = [&]( double s ) {
            double b = 0;
            double nl = m;
            for ( int l = 1; nl > 64 && l < 24; ++l ) {
                nl = std::ceil( nl / pack );
                b += nl * ( 4 + MG_STRIDE_B * ( 4 + s ) + s ) + 3 * s * nl + ( l <= kc ? 5 * s * nl : 0 );
            }
            return b;
        };
        double lin = 4 * 8 * m + 8 * m + coarse( 8 );
        if ( f32 ) lin += 4 * 8 * m + 4 * m + coarse( 4 ) + 4 * cap + 4 * m + 8 * m;
        // each precision's partial sums, K-cycle partials and dense inverse of the coarsest level
        parts[ 7 ] = lin + ( f32 ? 2 : 1 ) * ( 24 * 5 * 512 * 8 + 64 * 64 * 8 + 4 * 512 * 8 );
    } else
        parts[ 7 ] = 3 * 8 * m;
    // the density's copies: an image's blurred copies ( the width continuation ) and its pyramid of minima ( the limits )
    if ( ! d3 && density_kind == SDOT_AMD_IMAGE && o ) {
        const double px = o[ O_IMG_NX ] * o[ O_IMG_NY ];
        if ( int( o[ O_CONTINUATION ] ) != 0 ) parts[ 8 ] += 2 * 8 * px;
        if ( limits ) parts[ 8 ] += 8 * px * 4.0 / 3.0 + 8 * 64;
    }
    // the reductions' partials ( one per 256 cells for the widest ) and small state
    parts[ 9 ] = 2 * 8 * std::ceil( m / 256 ) + 64 * 1024;
    for ( int q = 0; q < SDOT_AMD_NB_MEMORY_PARTS; ++q ) if ( parts[ q ] > 0 ) parts[ q ] = pad( parts[ q ] );
}

double solve_bytes( I64 n, int dim, int density_kind, const double *o, int precision ) {
    double p[ SDOT_AMD_NB_MEMORY_PARTS ];
    solve_parts( n, dim, density_kind, o, precision, p );
    double s = 0;
    for ( double v : p ) s += v;
    return s;
}

} // namespace sdot_amd
