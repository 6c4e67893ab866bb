// The step of Newton by the limits of the cells ( 2D ), and the width continuation's blurred image.
//
// With its edges frozen, a cell along `w + t d` is known at every `t`: a bisector's offset is affine in the
// weights and its normal fixed, so each vertex ( the crossing of two consecutive edge lines ) is affine in `t`
// and the area a polynomial of degree 2 ( `gpu/Newton2D.cuh:505-575` of the reference ). FORWARD, on the
// accepted diagram: the first `t > 0` where a cell's mass reaches the floor `eps`. BACKWARD, on a trial
// diagram's cells under the floor: where they crossed it going back ( `gpu/Newton2D.cuh:594-650` ). An image
// density bisects the frozen cell's mass instead of solving a quadratic ( `gpu/Newton2D.cuh:655-770` ).
#pragma once

#include "cells.h"

namespace sdot_amd {

/// the smallest pixel value over blocks of `2^k x 2^k` pixels, level by level ( level 0: the image itself; none when
/// `nl == 0` ): a lower bound of an image density over any box in two or four reads. Level `l + 1` is `ceil( w / 2 ) x
/// ceil( h / 2 )` for level `l` of `w x h`, the levels one after the other in `v`
struct MinPyramid {
    const double *v = nullptr;
    int nl = 0;
    int nx = 0, ny = 0;                 ///< level 0's size
};

/// the pyramid of the image `img[ ny ][ nx ]` into `buf` ( `min_pyramid_size( nx, ny )` doubles )
I64 min_pyramid_size( int nx, int ny );
MinPyramid build_min_pyramid( const double *img, int nx, int ny, double *buf, hipStream_t st );

/// a diagram's cells with their edges frozen, along `w + t d` ( RANK order )
struct Along {
    const double4 *seeds = nullptr;     ///< the positions ( x, y ) by rank
    const double  *w = nullptr, *d = nullptr;
    const int     *edges = nullptr, *nb_edges = nullptr;
    I64            n = 0;
    double         lo[ 2 ] = { 0, 0 }, hi[ 2 ] = { 1, 1 };
    Density        dn;
    const double  *nue = nullptr;       ///< the targets with the exact duplicates folded in ( a duplicate's 0: no say ); may be null
    MinPyramid     pyr;                 ///< an image density's pyramid of minima ( the forward limit's shortcut ); may be empty
};

/// FORWARD: `out[ 0 ]` = the smallest, over the cells, first `t > 0` where the frozen mass reaches `eps`
/// ( 1e300: none ). `out` is device memory.
void alpha_forward( const Along &al, double eps, double *out, hipStream_t st );

/// BACKWARD: over the cells with `a[ k ] < eps` ( the trial diagram at `t_trial`, whose edges `al` holds ):
/// `out[ 0 ]` their count, `out[ 1 ]` the smallest `t` where one crossed the floor ( half the trial when unknown )
void alpha_backward( const Along &al, const double *a, double eps, double t_trial, double *out, hipStream_t st );

/// the image `src[ ny ][ nx ]` ( pixel sizes `hx`, `hy` ) at width `s` into `dst`: along x then along y, a Gaussian
/// of standard deviation `s / h` pixels truncated at four of them, each value divided by the sum of the weights
/// that fall inside the image ( `sdotplan/convolved/Image.h:56-85` ). `tmp` is as large as the image.
void blur_image( const double *src, int nx, int ny, double hx, double hy, double s, double *dst, double *tmp, hipStream_t st );

} // namespace sdot_amd
