// The power cells of a 3D diagram on the device: one wave per cell, one lane per vertex of the cell's polytope.
#pragma once

#include "cells.h"
#include "tree3.h"

namespace sdot_amd {

/// the 3D target density: a constant on the domain box, or a sum of isotropic Gaussians truncated to the box
struct Density3 {
    int           kind = 0;        ///< 0 constant, 1 a voxel image, 2 a sum of Gaussians
    double        rho = 1;
    const double *gauss = nullptr; ///< device, `[ ng ][ 5 ]`: centre x, y, z, width, mass
    int           ng = 0;
    /// a voxel image: voxel `( i, j, k )` covers `[ x0 + i hx, x0 + ( i + 1 ) hx ] x ...`, its value at `img[ ( k ny + j ) nx + i ]`
    const double *img = nullptr;   ///< device
    int           nx = 0, ny = 0, nz = 0;
    double        x0 = 0, y0 = 0, z0 = 0, hx = 1, hy = 1, hz = 1;
};

/// the vertex and face slots of a cell in the first pass ( one lane per vertex ); a cell past them goes to the second
/// pass ( one thread, global slots of `CELL3_SLOT_*` ), past which it is a failure ( code 1 )
constexpr int CELL3_VERTICES = 64;
constexpr int CELL3_FACES = 64;
constexpr int CELL3_SLOT_VERTICES = 192;
constexpr int CELL3_SLOT_FACES = 192;

/// one 3D diagram into `out` ( rank order ): masses, and the rows of the Laplacian ( facets mode ) or the moments. The rows
/// of 64 consecutive ranks share a block of `64 room` entries, interleaved as the 2D rows ( entry `q` of rank `k` at
/// `( k / 64 ) 64 room + k % 64 + 64 q` ); `out.counters[ 0 ]` receives the longest row, so that a diagram whose rows
/// outgrew `room` is redone with more. Failures: `counters[ 2 ]`, the first failing rank in `counters[ 3 ]` and its code in
/// `counters[ 5 ]`; cells of the second pass: `counters[ 1 ]`.
void run_cells_3d( const DevTree3 &t, const double *lo, const double *hi, const Density3 &dn, int mode, DiagramOut &out, int room,
                   hipStream_t st );

} // namespace sdot_amd
