// The power cells of a 2D diagram on the device: one lane per cell, what Newton needs of a diagram.
#pragma once

#include "tree.h"

namespace sdot_amd {

/// the target density: a constant on the domain box, a piecewise-constant image `img[ j * nx + i ]` on pixels
/// `[ x0 + i hx, x0 + ( i + 1 ) hx ] x [ y0 + j hy, ... ]`, or a sum of isotropic Gaussians truncated to the box
/// THE CELL THE CUTS START FROM when the domain is a convex polygon and not the box ( the reference's start cell,
/// `StartCell.py`, `gpu/Cell2D.cuh:295-333, 465-478` ): its vertices counterclockwise, and the line of each edge ( edge `i` from
/// vertex `i` to `i + 1`, the cut `-5 - i`: `nx x + ny y <= off`, `( nx, ny )` the unit outward normal ). `nb = 0`: the box
constexpr int DOMAIN_MAX = 12;             ///< at most this many vertices ( the first passes' 16 slots keep room for the cuts )
struct Polygon2 {
    int    nb = 0;
    double x[ DOMAIN_MAX ] = {}, y[ DOMAIN_MAX ] = {};
    double nx[ DOMAIN_MAX ] = {}, ny[ DOMAIN_MAX ] = {}, off[ DOMAIN_MAX ] = {};
};

/// entry `j` of one of a `Polygon2`'s arrays, read with constant indices only: the polygon travels in the kernels'
/// arguments, and an index from data into them makes every lane copy the arguments to scratch memory first
__device__ __forceinline__ double dom_at( const double ( &a )[ DOMAIN_MAX ], int j ) {
    double v = 0;
#pragma unroll
    for ( int q = 0; q < DOMAIN_MAX; ++q ) v = q == j ? a[ q ] : v;
    return v;
}

struct Density {
    int           kind = 0;        ///< 0 constant, 1 image, 2 a sum of isotropic Gaussians
    double        rho = 1;
    const double *img = nullptr;   ///< device
    int           nx = 0, ny = 0;
    double        x0 = 0, y0 = 0, hx = 1, hy = 1;
    const double *gauss = nullptr; ///< device, `[ ng ][ 4 ]`: centre x, y, width, mass
    int           ng = 0;
    Polygon2      dom;             ///< the domain when it is a polygon ( the density is zero outside it )
};

/// the polygon domain of `nb` counterclockwise vertices `xy[ nb ][ 2 ]` ( HOST ): its edges' lines; throws when it is not a
/// convex counterclockwise polygon of 3 to `DOMAIN_MAX` vertices
void set_domain( Density &dn, int nb, const double *xy );

/// what one diagram writes, in RANK order: the masses, the rows of the Laplacian ( each cell's
/// own view of its facets: `col[ row_off[ k ] + q * ROW_STEP ]`, `val[ ... ]`, `q < row_cnt[ k ]`, and
/// `dia[ k ]` the row sum: the rows of a wave interleaved, entry `q` of its 64 rows in 64 consecutive words ),
/// and the diagnostics
struct DiagramOut {
    double   *mass = nullptr;
    int      *row_off = nullptr, *row_cnt = nullptr;
    int      *col = nullptr;
    double   *val = nullptr, *dia = nullptr;
    I64       facet_cap = 0;
    // moments ( instead of facets ): barycentres ( absolute, rank order ) and per-cell cost
    double   *bary = nullptr, *cost = nullptr;
    // counters on the device: [ 0 ] facets used, [ 1 ] cells past the first pass ( FP32 kernel: and the cells that failed its
    // conditioning test ), [ 2 ] failures, [ 3 ] the first failing rank, [ 4 ] cells past the second pass, [ 6 ] FP32 kernel:
    // cells redone in FP64 for a facet only one side had, [ 7 ] facets still one-sided after that, dropped
    unsigned *counters = nullptr;
    int      *ovf_list = nullptr;
    unsigned long long *events = nullptr;   ///< side tests, pruning vertex tests, new vertices, node visits, cuts
    // facets mode, when set: each cell's cuts in polygon order ( `edges[ q * n + k ]`: a seed rank, or `-1 - f` side `f`
    // of the box, 0 bottom, 1 right, 2 top, 3 left ) and their count ( 0 for an empty cell, -1 past `EDGE_CAP` or failed )
    int      *edges = nullptr, *nb_edges = nullptr;
    // facets mode: slots left free at the end of every row ( a wave's rows are `64 x ( longest + row_extra )` ), for the
    // links the solver appends after the cells ( the exact duplicates' rows )
    int       row_extra = 0;
    // cells mode: each cell's vertex count ( rank order ), and, when `cell_off` is set, its vertices from `cell_off[ order[ k ] ]`
    // on ( user order ): absolute coordinates in `cell_xy`, and in `cell_cut` the cut of the edge that leaves each vertex ( a
    // seed's user index, or `-1 - f` for side `f` of the box )
    int       *cell_nb = nullptr;
    const long long *cell_off = nullptr;
    double    *cell_xy = nullptr;
    long long *cell_cut = nullptr;
    const int *order = nullptr;
    // vjp mode: the cotangent of the masses in ( rank order ), the gradients of the weights and of the positions out ( rank order )
    const double *vjp_g = nullptr;
    double    *grad_w = nullptr, *grad_pos = nullptr;
};

/// the Laplacian rows of a wave's 64 cells are interleaved: entry `q` of the row of lane `l` at `base + q * 64 + l`
constexpr int ROW_STEP = 64;

/// the vertex slots of the passes ( the first in LDS, the second in global slots )
constexpr int CELL_BLOCK = 128;
constexpr int CAP1 = 16;             ///< first pass, LDS
#ifndef SDOT_AMD_CAPW
#define SDOT_AMD_CAPW 16
#endif
/// the FP32 walk's LDS slots ( even, at most `CAP1` ): 12 bytes a slot, so its count sets the walk's waves per CU
constexpr int CAPW = SDOT_AMD_CAPW;
static_assert( CAPW % 2 == 0 && CAPW >= 6 && CAPW <= CAP1, "the FP32 walk's slots: even, from 6 to CAP1" );
constexpr int CAP2 = 96;             ///< second pass, LDS ( one wave per block )
constexpr int CAP3 = 2048;           ///< third pass, global slots: the vertex limit of a cell
constexpr int PASS2_BLOCKS = 1024;
constexpr int REDO_BLOCKS = 256;     ///< the FP32 kernel's FP64 redo of one-sided facets ( one wave each )
constexpr int OVF_SLOTS = 4096;
constexpr int EDGE_CAP = 16;         ///< the edges a cell keeps for the step by the limits ( `gpu/Cell2D.cuh:97` )

enum CellMode { MODE_MASSES = 0, MODE_FACETS = 1, MODE_MOMENTS = 2, MODE_CELLS = 3, MODE_VJP = 4 };

/// one diagram: every cell of the tree at its current weights ( `t.seeds[ k ].z` ) in the box
/// `lo / hi`. Resets and uses the counters; the overflow pass is launched without a read back.
/// `fp32`: the walk and the cuts in float, every vertex then re-solved in double from its two planes before the measures
/// ( the cells past the first pass run the FP64 kernel either way ).
/// `ovf_slots`: the lanes ( a multiple of 64 ) of the third pass, each with `CAP3` vertex slots in global memory.
void run_cells( const DevTree &t, const double *lo, const double *hi, const Density &dn, int mode, DiagramOut &out, hipStream_t st,
                bool fp32 = false, int ovf_slots = OVF_SLOTS );

/// the device scratch a diagram needs besides its outputs ( overflow list and slots )
void cells_scratch( DiagramOut &out, I64 n, const std::string &tag, hipStream_t st );

} // namespace sdot_amd
