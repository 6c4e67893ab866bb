// The piecewise-constant image density as the device integrates it on segments ( Green's theorem: a
// cell's mass is `loop_integral F0 dy`, `F0( x, y ) = int_{x_ref}^x rho( s, y ) ds` along the row of `y`,
// exact piece by piece since `F0` is linear in `x` inside a pixel; `gpu/Density2D.cuh:80-140` of the
// reference integrates the same way ).
#pragma once

#include "cells.h"

#pragma clang fp contract( off )

namespace sdot_amd {

/// the pixel grid as seen from one seed: columns `u_i = x0 + i hx - p0x`, rows `v_j = y0 + j hy - p0y`
struct Grid {
    const double *img;
    int nx, ny;
    double x0, y0, hx, hy;   ///< x0, y0 in the seed's frame
    int iref;                ///< the column the row integrals start from
    /// the row integrals start from the column of `x` ( seed frame ): any column that is the same for the whole cell gives
    /// the same loop integral, and one in the cell keeps `F0`'s loop and terms of the cell's width, where a converged
    /// transport puts cells tens of cells away from their seeds
    __device__ void anchor( double x ) {
        const double c = floor( ( x - x0 ) / hx );
        iref = c < 0 ? 0 : ( c >= double( nx - 1 ) ? nx - 1 : int( c ) );
    }
    __device__ double u( int i ) const { return x0 + double( i ) * hx; }
    __device__ double v( int j ) const { return y0 + double( j ) * hy; }
    __device__ double val( int i, int j ) const { return img[ size_t( j ) * size_t( nx ) + size_t( i ) ]; }
    /// `int_{u_iref}^{u} rho( s, row j ) ds` for `u` in column `i`
    __device__ double F0( int i, int j, double x ) const {
        double acc = 0;
        if ( i >= iref ) for ( int q = iref; q < i; ++q ) acc += val( q, j ) * hx;
        else             for ( int q = i; q < iref; ++q ) acc -= val( q, j ) * hx;
        return acc + val( i, j ) * ( x - u( i ) );
    }
};

/// walks the segment `a -> b` ( seed frame ) across the pixel grid; `f( i, j, ta, tb )` per piece
template<class F>
__device__ __forceinline__ void grid_walk( const Grid &g, double ax, double ay, double bx, double by, F &&f ) {
    const double gax = ( ax - g.x0 ) / g.hx, gay = ( ay - g.y0 ) / g.hy;
    const double gbx = ( bx - g.x0 ) / g.hx, gby = ( by - g.y0 ) / g.hy;
    const double ddx = gbx - gax, ddy = gby - gay;
    const int sx = ddx > 0 ? 1 : ( ddx < 0 ? -1 : 0 ), sy = ddy > 0 ? 1 : ( ddy < 0 ? -1 : 0 );
    const double rx = sx ? 1.0 / ddx : 0.0, ry = sy ? 1.0 / ddy : 0.0;
    // the next vertical / horizontal grid line crossed, as a parameter
    double nxl = sx > 0 ? floor( gax ) + 1 : ceil( gax ) - 1;
    double nyl = sy > 0 ? floor( gay ) + 1 : ceil( gay ) - 1;
    double t0 = 0;
    for ( int guard = 0; guard < 1 << 20; ++guard ) {
        const double tx = sx ? ( nxl - gax ) * rx : 2.0;
        const double ty = sy ? ( nyl - gay ) * ry : 2.0;
        double t1 = fmin( fmin( tx, ty ), 1.0 );
        if ( t1 < t0 ) t1 = t0;
        const double tm = 0.5 * ( t0 + t1 );
        int i = int( floor( gax + tm * ddx ) ), j = int( floor( gay + tm * ddy ) );
        i = i < 0 ? 0 : ( i >= g.nx ? g.nx - 1 : i );
        j = j < 0 ? 0 : ( j >= g.ny ? g.ny - 1 : j );
        if ( t1 > t0 ) f( i, j, t0, t1 );
        if ( t1 >= 1.0 ) break;
        if ( tx <= t1 ) nxl += sx;
        if ( ty <= t1 ) nyl += sy;
        t0 = t1;
    }
}

/// the pixel grid of `dn` seen from the seed `( px, py )`, rows integrated from the seed's column
__device__ __forceinline__ Grid grid_from( const Density &dn, double px, double py ) {
    Grid g;
    g.img = dn.img; g.nx = dn.nx; g.ny = dn.ny; g.hx = dn.hx; g.hy = dn.hy;
    g.x0 = dn.x0 - px; g.y0 = dn.y0 - py;
    const int ir = int( floor( -g.x0 / g.hx ) );
    g.iref = ir < 0 ? 0 : ( ir >= g.nx ? g.nx - 1 : ir );
    return g;
}

/// the term of the segment `a -> b` ( seed frame ) in `mass = loop_integral F0 dy`
__device__ __forceinline__ double segment_mass( const Grid &g, double ax, double ay, double bx, double by ) {
    const double ex = bx - ax, ey = by - ay;
    double m = 0;
    grid_walk( g, ax, ay, bx, by, [&]( int i, int j, double t0, double t1 ) {
        m += g.F0( i, j, fma( ex, 0.5 * ( t0 + t1 ), ax ) ) * ( ey * ( t1 - t0 ) );
    } );
    return m;
}

} // namespace sdot_amd
