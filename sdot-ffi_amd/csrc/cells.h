// The power cells of a 2D diagram on the device: one lane per cell, what Newton needs of a diagram.
#pragma once

#include "tree.h"

namespace sdot_amd {

/// the target density: a constant on the domain box, or a piecewise-constant image
/// `img[ j * nx + i ]` on pixels `[ x0 + i hx, x0 + ( i + 1 ) hx ] x [ y0 + j hy, ... ]`
struct Density {
    int           kind = 0;        ///< 0 constant, 1 image
    double        rho = 1;
    const double *img = nullptr;   ///< device
    int           nx = 0, ny = 0;
    double        x0 = 0, y0 = 0, hx = 1, hy = 1;
};

/// what one diagram writes, in RANK order: the masses, the rows of the Laplacian ( each cell's
/// own view of its facets: `col[ row_off[ k ] + q ]`, `val[ ... ]`, `q < row_cnt[ k ]`, and
/// `dia[ k ]` the row sum ), and the diagnostics
struct DiagramOut {
    double   *mass = nullptr;
    int      *row_off = nullptr, *row_cnt = nullptr;
    int      *col = nullptr;
    double   *val = nullptr, *dia = nullptr;
    I64       facet_cap = 0;
    // moments ( instead of facets ): barycentres ( absolute, rank order ) and per-cell cost
    double   *bary = nullptr, *cost = nullptr;
    // counters on the device: [ 0 ] facets used, [ 1 ] cells past the first pass, [ 2 ] failures,
    // [ 3 ] the first failing rank, [ 4 ] cells past the second pass
    unsigned *counters = nullptr;
    int      *ovf_list = nullptr;
    unsigned long long *events = nullptr;   ///< side tests, pruning vertex tests, new vertices, node visits, cuts
};

/// the vertex slots of the passes ( the first in LDS, the second in global slots )
constexpr int CELL_BLOCK = 128;
constexpr int CAP1 = 16;             ///< first pass, LDS
constexpr int CAP2 = 96;             ///< second pass, LDS ( one wave per block )
constexpr int CAP3 = 2048;           ///< third pass, global slots: the vertex limit of a cell
constexpr int PASS2_BLOCKS = 1024;
constexpr int OVF_SLOTS = 4096;

enum CellMode { MODE_MASSES = 0, MODE_FACETS = 1, MODE_MOMENTS = 2 };

/// one diagram: every cell of the tree at its current weights ( `t.seeds[ k ].z` ) in the box
/// `lo / hi`. Resets and uses the counters; the overflow pass is launched without a read back.
void run_cells( const DevTree &t, const double *lo, const double *hi, const Density &dn, int mode, DiagramOut &out, hipStream_t st );

/// the device scratch a diagram needs besides its outputs ( overflow list and slots )
void cells_scratch( DiagramOut &out, I64 n, const std::string &tag );

} // namespace sdot_amd
