// The step by the limits and the blurred image ( see limits.h ). One lane per cell inside the fixed-shape
// reductions; the expressions are the oracle's, with no contraction, so a constant density gives its limits to
// the bit for the same weights and direction.

#include "limits.h"
#include "gauss.h"
#include "image.h"
#include "reduce.h"

#include <cmath>
#include <vector>

#pragma clang fp contract( off )

namespace sdot_amd {

namespace {

constexpr double DET_MIN = 1e-6;   ///< under this relative determinant two edge lines are parallel: no say

/// the line of cut `c` of cell `k` in its seed's frame at `w + t0 d`: `nx x + ny y = off + s delta` at `w + ( t0 + s ) d`
__device__ __forceinline__ void line_along( const Along &al, I64 k, int c, double t0, double px, double py, double &nx, double &ny, double &off,
                                            double &delta ) {
    if ( c >= 0 ) {
        const double4 q = al.seeds[ c ];
        nx = q.x - px;
        ny = q.y - py;
        off = 0.5 * ( nx * nx + ny * ny ) + 0.5 * ( ( al.w[ k ] + t0 * al.d[ k ] ) - ( al.w[ c ] + t0 * al.d[ c ] ) );
        delta = 0.5 * ( al.d[ k ] - al.d[ c ] );
        return;
    }
    if ( c <= -5 ) {                                     // an edge of the polygon domain: it does not move either
        const Polygon2 &D = al.dn.dom;
        const int j = -5 - c;
        nx = dom_at( D.nx, j );
        ny = dom_at( D.ny, j );
        off = dom_at( D.off, j ) - nx * px - ny * py;
        delta = 0;
        return;
    }
    const int f = -1 - c;                               // 0 bottom, 1 right, 2 top, 3 left: they do not move
    nx = ( f & 1 ) ? 1.0 : 0.0;
    ny = ( f & 1 ) ? 0.0 : 1.0;
    off = f == 0 ? al.lo[ 1 ] - py : f == 1 ? al.hi[ 0 ] - px : f == 2 ? al.hi[ 1 ] - py : al.lo[ 0 ] - px;
    delta = 0;
}

/// the vertices of the frozen cell at `w + t0 d` in order ( vertex `q`: the crossing of the lines of edges `q - 1` and
/// `q` ) with their velocities along `d`, streamed to `f( q, vx, vy, ux, uy )`; `false` if two consecutive lines are
/// nearly parallel ( the cell then has no say )
template<class F>
__device__ __forceinline__ bool for_each_vertex( const Along &al, I64 k, int nb, double t0, F &&f ) {
    const double4 p = al.seeds[ k ];
    const int *e = al.edges + k;
    const size_t n = size_t( al.n );
    double ax, ay, ao, ad;
    line_along( al, k, e[ size_t( nb - 1 ) * n ], t0, p.x, p.y, ax, ay, ao, ad );
    for ( int q = 0; q < nb; ++q ) {
        double bx, by, bo, bd;
        line_along( al, k, e[ size_t( q ) * n ], t0, p.x, p.y, bx, by, bo, bd );
        const double det = ax * by - ay * bx;
        if ( ! ( det * det > DET_MIN * DET_MIN * ( ax * ax + ay * ay ) * ( bx * bx + by * by ) ) ) return false;
        const double inv = 1 / det;
        f( q, ( ao * by - bo * ay ) * inv, ( ax * bo - bx * ao ) * inv, ( ad * by - bd * ay ) * inv, ( ax * bd - bx * ad ) * inv );
        ax = bx; ay = by; ao = bo; ad = bd;
    }
    return true;
}

/// `area( t0 + s ) = a0 + a1 s + a2 s^2` while the edges hold ( the shoelace terms in the order `( 0, 1 ), ..., ( nb - 1, 0 )` )
__device__ bool area_polynomial( const Along &al, I64 k, int nb, double t0, double &a0, double &a1, double &a2 ) {
    double c0 = 0, c1 = 0, c2 = 0;
    double fx = 0, fy = 0, fux = 0, fuy = 0, qx = 0, qy = 0, qux = 0, quy = 0;
    auto term = [&]( double vx, double vy, double ux, double uy, double rx, double ry, double rux, double ruy ) {
        c0 += vx * ry - rx * vy;
        c1 += vx * ruy + ux * ry - rx * uy - rux * vy;
        c2 += ux * ruy - rux * uy;
    };
    const bool ok = for_each_vertex( al, k, nb, t0, [&]( int q, double vx, double vy, double ux, double uy ) {
        if ( q == 0 ) { fx = vx; fy = vy; fux = ux; fuy = uy; }
        else term( qx, qy, qux, quy, vx, vy, ux, uy );
        qx = vx; qy = vy; qux = ux; quy = uy;
    } );
    if ( ! ok ) return false;
    term( qx, qy, qux, quy, fx, fy, fux, fuy );
    const double sg = c0 < 0 ? -0.5 : 0.5;
    a0 = sg * c0; a1 = sg * c1; a2 = sg * c2;
    return true;
}

/// the mass of the frozen cell at `w + t d` against a density that is not a constant ( a cell turned inside out has a
/// negative mass )
__device__ bool frozen_mass( const Along &al, I64 k, int nb, double t, double &m ) {
    const double4 p = al.seeds[ k ];
    if ( al.dn.kind == 2 ) {                            // the signed corner sums, Gaussian by Gaussian
        m = 0;
        for ( int q = 0; q < al.dn.ng; ++q ) {
            const GaussTerm G = gauss_term( al.dn, q, p.x, p.y );
            double W = 0;
            bool ok = true;
            double fx = 0, fy = 0, qx = 0, qy = 0;
            ok = for_each_vertex( al, k, nb, t, [&]( int v, double vx, double vy, double, double ) {
                if ( v == 0 ) { fx = vx; fy = vy; }
                else W += gauss_wedge( ( qx - G.cx ) / G.s, ( qy - G.cy ) / G.s, ( vx - G.cx ) / G.s, ( vy - G.cy ) / G.s );
                qx = vx; qy = vy;
            } );
            if ( ! ok ) return false;
            W += gauss_wedge( ( qx - G.cx ) / G.s, ( qy - G.cy ) / G.s, ( fx - G.cx ) / G.s, ( fy - G.cy ) / G.s );
            m += G.w * W;
        }
        return true;
    }
    Grid g = grid_from( al.dn, p.x, p.y );
    double acc = 0, fx = 0, fy = 0, qx = 0, qy = 0;
    const bool ok = for_each_vertex( al, k, nb, t, [&]( int q, double vx, double vy, double, double ) {
        if ( q == 0 ) { fx = vx; fy = vy; g.anchor( vx ); }
        else acc += segment_mass( g, qx, qy, vx, vy );
        qx = vx; qy = vy;
    } );
    if ( ! ok ) return false;
    m = acc + segment_mass( g, qx, qy, fx, fy );
    return true;
}

/// the real roots of `a2 s^2 + a1 s + c = 0`, sorted ( `sdotplan/Bounds.h:170-185` )
__device__ int quadratic_roots( double a2, double a1, double c, double &r1, double &r2 ) {
    if ( fabs( a2 ) <= 1e-300 ) {
        if ( a1 == 0 ) return 0;
        r1 = r2 = -c / a1;
        return 1;
    }
    const double disc = a1 * a1 - 4 * a2 * c;
    if ( disc < 0 ) return 0;
    const double sq = sqrt( disc );
    const double q = -0.5 * ( a1 + ( a1 >= 0 ? sq : -sq ) );
    double x1 = q / a2, x2 = q != 0 ? c / q : x1;
    if ( x1 > x2 ) { const double s = x1; x1 = x2; x2 = s; }
    r1 = x1; r2 = x2;
    return 2;
}

/// the smallest forward limit found so far by the lanes of a pass ( the bits of a non-negative double, which order as
/// the doubles do ), read coherently
__device__ __forceinline__ double best_so_far( const unsigned long long *best ) {
    return __longlong_as_double( (long long) __hip_atomic_load( best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) );
}

__device__ __forceinline__ void publish_best( unsigned long long *best, double v ) {
    if ( v < best_so_far( best ) ) atomicMin( best, (unsigned long long) __double_as_longlong( v ) );
}

/// the admissible end of `[ lo, hi ]` ( `mass( lo ) >= eps > mass( hi )` ), to 1e-2 relative and above zero. With `best`,
/// the bisection stops once `lo` reaches the smallest limit found so far: the end it would give is at least `lo`, so it
/// cannot be the smallest ( the value returned is then `lo` )
__device__ double mass_bisection( const Along &al, I64 k, int nb, double lo, double hi, double eps, const unsigned long long *best = nullptr ) {
    for ( int round = 0; round < 64; ++round ) {
        if ( hi - lo <= 1e-2 * hi && lo > 0 ) break;
        if ( best && lo > 0 && lo >= best_so_far( best ) ) break;
        const double mid = 0.5 * ( lo + hi );
        double m;
        if ( frozen_mass( al, k, nb, mid, m ) && m >= eps ) lo = mid;
        else                                                hi = mid;
    }
    return lo;
}

/// A CERTIFICATE that the frozen cell's mass stays above `eps` on `t in [ 0, T ]` ( an image density ): every edge keeps
/// its orientation ( its length is affine in `t`, its direction fixed: positive at both ends, it stays positive, and the
/// cell stays convex ), the cell then lies in the box its vertices sweep, and its smallest area on `[ 0, 1 ]` times the
/// image's smallest value over that box passes `eps` with a margin far above the rounding of the exact integral. The
/// exact computation would then find no crossing: the cell has no limit, without its image integrals.
__device__ bool stays_above( const Along &al, I64 k, int nb, double eps, double T = 1.0 ) {
    const double4 p = al.seeds[ k ];
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    double qx = 0, qy = 0, qux = 0, quy = 0, fx = 0, fy = 0, fux = 0, fuy = 0;
    bool kept = true;
    auto edge = [&]( double ax, double ay, double aux, double auy, double bx, double by, double bux, double buy ) {
        const double e0x = bx - ax, e0y = by - ay, e1x = e0x + T * ( bux - aux ), e1y = e0y + T * ( buy - auy );
        kept = kept && e0x * e1x + e0y * e1y > 0;
    };
    const bool ok = for_each_vertex( al, k, nb, 0.0, [&]( int q, double vx, double vy, double ux, double uy ) {
        xmin = fmin( xmin, fmin( vx, vx + T * ux ) ); xmax = fmax( xmax, fmax( vx, vx + T * ux ) );
        ymin = fmin( ymin, fmin( vy, vy + T * uy ) ); ymax = fmax( ymax, fmax( vy, vy + T * uy ) );
        if ( q == 0 ) { fx = vx; fy = vy; fux = ux; fuy = uy; }
        else edge( qx, qy, qux, quy, vx, vy, ux, uy );
        qx = vx; qy = vy; qux = ux; quy = uy;
    } );
    if ( ! ok ) return false;
    edge( qx, qy, qux, quy, fx, fy, fux, fuy );
    if ( ! kept ) return false;
    double a0, a1, a2;
    if ( ! area_polynomial( al, k, nb, 0.0, a0, a1, a2 ) ) return false;
    double amin = fmin( a0, a0 + T * ( a1 + T * a2 ) );
    if ( a2 > 0 ) {
        const double ts = -a1 / ( 2 * a2 );
        if ( ts > 0 && ts < T ) amin = fmin( amin, a0 + ts * ( a1 + ts * a2 ) );
    }
    if ( ! ( amin > 0 ) ) return false;
    // the pixels of the box, a pixel more on each side, and the pyramid's level where it spans two blocks at most
    const Density &d = al.dn;
    auto clampi = []( double v, int hi ) { return v < 0 ? 0 : ( v > double( hi ) ? hi : int( v ) ); };
    const int i0 = clampi( floor( ( p.x + xmin - d.x0 ) / d.hx ) - 1, d.nx - 1 ), i1 = clampi( floor( ( p.x + xmax - d.x0 ) / d.hx ) + 1, d.nx - 1 );
    const int j0 = clampi( floor( ( p.y + ymin - d.y0 ) / d.hy ) - 1, d.ny - 1 ), j1 = clampi( floor( ( p.y + ymax - d.y0 ) / d.hy ) + 1, d.ny - 1 );
    const MinPyramid &P = al.pyr;
    int lv = 0;
    while ( lv + 1 < P.nl && ( ( i1 >> lv ) - ( i0 >> lv ) > 1 || ( j1 >> lv ) - ( j0 >> lv ) > 1 ) ) ++lv;
    if ( ( i1 >> lv ) - ( i0 >> lv ) > 1 || ( j1 >> lv ) - ( j0 >> lv ) > 1 ) return false;
    int w = P.nx, h = P.ny;                             // ( level `lv`'s width and place: sizes, not a table, so that
    I64 off = 0;                                        //   no array of the arguments is indexed per lane )
    for ( int l = 0; l < lv; ++l ) {
        off += I64( w ) * h;
        w = ( w + 1 ) / 2; h = ( h + 1 ) / 2;
    }
    const double *v = P.v + off;
    const double rho = fmin( fmin( v[ ( j0 >> lv ) * w + ( i0 >> lv ) ], v[ ( j0 >> lv ) * w + ( i1 >> lv ) ] ),
                             fmin( v[ ( j1 >> lv ) * w + ( i0 >> lv ) ], v[ ( j1 >> lv ) * w + ( i1 >> lv ) ] ) );
    return amin * rho > eps * ( 1 + 1e-6 );
}

/// the forward limit of cell `k` ( 1e300: none ). With `best`, the smallest limit found so far by other lanes, a cell that
/// cannot go under it returns early ( 1e300, or a value not under it ): a cell certified above `eps` on `[ 0, T ]`,
/// `T = best / 0.99`, has a limit above `best` ( the bisection's end is at least 0.99 times a `t` of mass under `eps`,
/// and every such `t` is past `T` ); the smallest limit over the cells is therefore the same
__device__ double forward_cell( const Along &al, I64 k, double eps, const unsigned long long *best = nullptr ) {
    const int nb = al.nb_edges[ k ];
    if ( nb < 3 ) return 1e300;
    double a0, a1, a2;
    if ( al.dn.kind == 0 ) {
        if ( ! area_polynomial( al, k, nb, 0.0, a0, a1, a2 ) ) return 1e300;
        const double rho = al.dn.rho;
        double r1, r2;
        const int nr = quadratic_roots( rho * a2, rho * a1, fmax( rho * a0 - eps, 0.0 ), r1, r2 );
        if ( nr >= 1 && r1 > 0 ) return r1;
        if ( nr >= 2 && r2 > 0 ) return r2;
        return 1e300;
    }
    // an image: the frozen mass at `t = 1`, and at the bottom of the area's parabola ( a cell that dips, then grows back ),
    // unless the cell is certified to stay above `eps`
    double T = 1;
    if ( best ) {
        const double b = best_so_far( best );
        if ( b < 0.99 ) T = b / 0.99;
    }
    if ( al.dn.kind == 1 && al.pyr.nl > 0 && stays_above( al, k, nb, eps, T ) ) return 1e300;
    double m1, hi = -1;
    if ( ! frozen_mass( al, k, nb, 1.0, m1 ) ) return 1e300;
    if ( m1 < eps ) hi = 1;
    else if ( area_polynomial( al, k, nb, 0.0, a0, a1, a2 ) && a2 > 0 ) {
        const double ts = -a1 / ( 2 * a2 );
        double ms;
        if ( ts > 0 && ts < 1 && frozen_mass( al, k, nb, ts, ms ) && ms < eps ) hi = ts;
    }
    if ( hi < 0 ) return 1e300;
    return mass_bisection( al, k, nb, 0.0, hi, eps, best );
}

__device__ double backward_cell( const Along &al, I64 k, double eps, double t_trial ) {
    double target = 0.5 * t_trial;
    const int nb = al.nb_edges[ k ];
    if ( nb < 3 ) return target;
    if ( al.dn.kind == 0 ) {
        double a0, a1, a2;
        if ( area_polynomial( al, k, nb, t_trial, a0, a1, a2 ) ) {
            const double rho = al.dn.rho;
            double r1, r2, beta = -1e300;
            const int nr = quadratic_roots( rho * a2, rho * a1, rho * a0 - eps, r1, r2 );
            if ( nr >= 1 && r1 < 0 ) beta = r1;
            if ( nr >= 2 && r2 < 0 ) beta = r2;
            const double tg = t_trial + beta;
            if ( tg > 0 && tg < t_trial ) target = tg;
        }
        return target;
    }
    double m0;
    if ( frozen_mass( al, k, nb, 0.0, m0 ) && m0 >= eps ) {
        const double tg = mass_bisection( al, k, nb, 0.0, t_trial, eps );
        if ( tg > 0 && tg < t_trial ) target = tg;
    }
    return target;
}

struct ForwardF {
    Along al;
    double eps;
    unsigned long long *best;
    __device__ void operator()( I64 k, double *v ) const {
        v[ 0 ] = forward_cell( al, k, eps, best );
        if ( v[ 0 ] < 1e300 ) publish_best( best, v[ 0 ] );
    }
};

__global__ void k_best_reset( unsigned long long *best ) {
    if ( threadIdx.x == 0 ) *best = (unsigned long long) __double_as_longlong( 1e300 );
}

struct BackwardF {
    Along al;
    const double *a;
    double eps, t_trial;
    __device__ void operator()( I64 k, double *v ) const {
        const bool below = a[ k ] < eps && ! ( al.nue && ! ( al.nue[ k ] > 0 ) );   // ( a duplicate is empty by construction )
        v[ 0 ] = below ? 1.0 : 0.0;
        v[ 1 ] = below ? backward_cell( al, k, eps, t_trial ) : 1e300;
    }
};

/// one level of the pyramid of minima from the one below ( blocks of 2 x 2, cut at the borders )
__global__ void k_min_down( const double *src, int sw, int sh, double *dst, int dw, int dh ) {
    const I64 at = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( at >= I64( dw ) * dh ) return;
    const int I = int( at % dw ), J = int( at / dw );
    double m = src[ size_t( 2 * J ) * sw + size_t( 2 * I ) ];
    if ( 2 * I + 1 < sw ) m = fmin( m, src[ size_t( 2 * J ) * sw + size_t( 2 * I + 1 ) ] );
    if ( 2 * J + 1 < sh ) {
        m = fmin( m, src[ size_t( 2 * J + 1 ) * sw + size_t( 2 * I ) ] );
        if ( 2 * I + 1 < sw ) m = fmin( m, src[ size_t( 2 * J + 1 ) * sw + size_t( 2 * I + 1 ) ] );
    }
    dst[ at ] = m;
}

/// one axis of the blur: `dst[ at ] = sum_k ker[ k ] src[ at + k stride ] / sum_k ker[ k ]` over the taps inside the image
__global__ void k_blur_axis( int nx, int ny, int axis, int r, const double *ker, const double *src, double *dst ) {
    const I64 at = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( at >= I64( nx ) * ny ) return;
    const int i = int( at % nx ), j = int( at / nx );
    const int len = axis ? ny : nx, pos = axis ? j : i;
    const I64 stride = axis ? nx : 1;
    const int kb = -r > -pos ? -r : -pos, ke = r < len - 1 - pos ? r : len - 1 - pos;
    double sum = 0, wsum = 0;
    for ( int k = kb; k <= ke; ++k ) {
        const double wk = ker[ k + r ];
        sum += wk * src[ at + I64( k ) * stride ];
        wsum += wk;
    }
    dst[ at ] = sum / wsum;
}

} // namespace

void alpha_forward( const Along &al, double eps, double *out, hipStream_t st ) {
    unsigned long long *best = arena().get<unsigned long long>( "lim.best", 1 );
    k_best_reset<<<1, 64, 0, st>>>( best );
    check_launch( "k_best_reset" );
    // ( the lanes publish the limits they complete as they go, and the others read them: a pass over a sample first to
    //   seed the bound cost more than it saved, 2.4 ms of long bisections on a thin grid )
    reduce_wide<1>( al.n, ForwardF{ al, eps, best }, Ops<1>{ { OP_MIN } }, out, st );
}

void alpha_backward( const Along &al, const double *a, double eps, double t_trial, double *out, hipStream_t st ) {
    reduce_wide<2>( al.n, BackwardF{ al, a, eps, t_trial }, Ops<2>{ { OP_SUM, OP_MIN } }, out, st );
}

I64 min_pyramid_size( int nx, int ny ) {
    I64 tot = 0;
    for ( int w = nx, h = ny, l = 0; l < 16; ++l ) {
        tot += I64( w ) * h;
        if ( w == 1 && h == 1 ) break;
        w = ( w + 1 ) / 2; h = ( h + 1 ) / 2;
    }
    return tot;
}

MinPyramid build_min_pyramid( const double *img, int nx, int ny, double *buf, hipStream_t st ) {
    MinPyramid P;
    P.v = buf; P.nx = nx; P.ny = ny;
    int w = nx, h = ny;
    I64 at = 0;
    SDOT_HIP( hipMemcpyAsync( buf, img, sizeof( double ) * size_t( w ) * size_t( h ), hipMemcpyDeviceToDevice, st ) );
    for ( int l = 0; l < 16; ++l ) {
        P.nl = l + 1;
        if ( w == 1 && h == 1 ) break;
        const int dw = ( w + 1 ) / 2, dh = ( h + 1 ) / 2;
        if ( l + 1 < 16 ) {
            k_min_down<<<blocks_for( I64( dw ) * dh, 256 ), 256, 0, st>>>( buf + at, w, h, buf + at + I64( w ) * h, dw, dh );
            check_launch( "k_min_down" );
        }
        at += I64( w ) * h;
        w = dw; h = dh;
    }
    return P;
}

void blur_image( const double *src, int nx, int ny, double hx, double hy, double s, double *dst, double *tmp, hipStream_t st ) {
    const size_t nb = size_t( nx ) * size_t( ny );
    SDOT_HIP( hipMemcpyAsync( dst, src, sizeof( double ) * nb, hipMemcpyDeviceToDevice, st ) );
    if ( ! ( s > 0 ) ) return;
    for ( int axis = 0; axis < 2; ++axis ) {
        const double sp = s / ( axis ? hy : hx );
        if ( ! ( sp > 1e-3 ) ) continue;
        const int len = axis ? ny : nx;
        const double rr = std::ceil( 4 * sp );
        const int r = rr < double( len ) ? int( rr ) : len;
        std::vector<double> ker( size_t( 2 * r + 1 ) );
        for ( int k = -r; k <= r; ++k ) ker[ size_t( k + r ) ] = std::exp( -0.5 * double( I64( k ) * k ) / ( sp * sp ) );
        double *dker = arena().get<double>( "blur.ker", ker.size() );
        SDOT_HIP( hipMemcpyAsync( dker, ker.data(), sizeof( double ) * ker.size(), hipMemcpyHostToDevice, st ) );
        k_blur_axis<<<blocks_for( I64( nb ), 256 ), 256, 0, st>>>( nx, ny, axis, r, dker, dst, tmp );
        check_launch( "k_blur_axis" );
        SDOT_HIP( hipMemcpyAsync( dst, tmp, sizeof( double ) * nb, hipMemcpyDeviceToDevice, st ) );
        SDOT_HIP( hipStreamSynchronize( st ) );          // `ker` is a host vector of this scope
    }
}

} // namespace sdot_amd
