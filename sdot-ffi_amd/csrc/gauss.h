// A sum of isotropic Gaussians as the device integrates it on a cell's edges ( 2D ).
//
// Mass: per Gaussian, the signed standard-normal measures of the corners ( centre, a, b ) summed over the edges of the
// polygon, in absolute value ( `gpu/Density2D.cuh:255-275` of the reference; `SumOfGaussians.cxx:210-266` for the corner ).
// The corner's measure is `( 1 / 2 pi ) int ( 1 - exp( -r^2 / 2 ) ) p dt / ( p^2 + t^2 )` along the edge ( `p` the distance
// from the centre to its line, `t` the abscissa ): exact arctangent tails beyond `|t| = 8`, an 8-point Gauss-Legendre on
// panels of at most half a unit in between. Facet integrals are `erf` differences; the moments boundary fluxes
// ( `int ( x - c ) phi = -s^2 circ phi n ds`, `int |x - c|^2 phi = 2 s^2 m - s^2 circ ( ( x - c ) . n ) phi ds` ).
#pragma once

#include "cells.h"

#pragma clang fp contract( off )

namespace sdot_amd {

/// the signed standard normal measure of the triangle `( 0, P, Q )`, counter-clockwise positive
__device__ __forceinline__ double gauss_wedge( double px, double py, double qx, double qy ) {
    const double dx = qx - px, dy = qy - py;
    const double L = sqrt( dx * dx + dy * dy );
    if ( ! ( L > 0 ) ) return 0;
    const double ux = dx / L, uy = dy / L;
    const double p = uy * px - ux * py, ap = fabs( p );
    if ( ! ( ap > 0 ) ) return 0;
    const double t0 = ux * px + uy * py, t1 = ux * qx + uy * qy, cut = 8;
    double acc = 0;
    if ( ap >= cut ) acc = atan( t1 / ap ) - atan( t0 / ap );
    else {
        if ( t0 < -cut ) acc += atan( fmin( t1, -cut ) / ap ) - atan( t0 / ap );
        if ( t1 > cut ) acc += atan( t1 / ap ) - atan( fmax( t0, cut ) / ap );
        const double c0 = fmax( t0, -cut ), c1 = fmin( t1, cut );
        if ( c1 > c0 ) {
            const double gx[ 4 ] = { 0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363 };
            const double gw[ 4 ] = { 0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763 };
            const int np = int( ceil( ( c1 - c0 ) * 2 ) );
            const double h = ( c1 - c0 ) / ( 2 * np );
            for ( int k = 0; k < np; ++k ) {
                const double m = c0 + ( 2 * k + 1 ) * h;
#pragma unroll
                for ( int q = 0; q < 4; ++q )
#pragma unroll
                    for ( int sg = -1; sg <= 1; sg += 2 ) {
                        const double t = m + sg * h * gx[ q ];
                        const double r2 = ap * ap + t * t;
                        acc += h * gw[ q ] * ( 1 - exp( -r2 / 2 ) ) * ap / r2;
                    }
            }
        }
    }
    return ( p < 0 ? -acc : acc ) / 6.283185307179586476925286766559;
}

/// Gaussian `q` of `dn` seen from the seed `( ox, oy )`: its centre in the seed's frame, its width and its mass
struct GaussTerm { double cx, cy, s, w; };
__device__ __forceinline__ GaussTerm gauss_term( const Density &dn, int q, double ox, double oy ) {
    const double *g = dn.gauss + 4 * q;
    return GaussTerm{ g[ 0 ] - ox, g[ 1 ] - oy, g[ 2 ], g[ 3 ] };
}

/// `int phi_q ds` along the segment `a -> b` of the seed frame ( `phi_q` of mass `w` )
__device__ __forceinline__ double gauss_segment_line( const GaussTerm &G, double ax, double ay, double bx, double by ) {
    const double ex = bx - ax, ey = by - ay;
    const double L = sqrt( ex * ex + ey * ey );
    if ( ! ( L > 0 ) ) return 0;
    const double ux = ex / L, uy = ey / L;
    const double px = ax - G.cx, py = ay - G.cy;
    const double d = -px * uy + py * ux;
    const double t0 = px * ux + py * uy, t1 = t0 + L;
    const double qq = d * d / ( 2 * G.s * G.s );
    if ( qq > 700 ) return 0;
    const double is2 = 0.70710678118654752440 / G.s;
    return G.w * exp( -qq ) * ( erf( t1 * is2 ) - erf( t0 * is2 ) ) / ( 2 * G.s * 2.50662827463100050242 );
}

/// the signed corner sum of Gaussian `G` over the polygon whose edges `edges( f )` visits as `f( ax, ay, bx, by )`
template<class Edges>
__device__ __forceinline__ double gauss_corner_sum( const GaussTerm &G, const Edges &edges ) {
    double W = 0;
    edges( [&]( double ax, double ay, double bx, double by ) {
        W += gauss_wedge( ( ax - G.cx ) / G.s, ( ay - G.cy ) / G.s, ( bx - G.cx ) / G.s, ( by - G.cy ) / G.s );
    } );
    return W;
}

} // namespace sdot_amd
