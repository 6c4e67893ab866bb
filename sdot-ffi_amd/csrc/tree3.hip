// The 3D BSP tree's weights and the device-side refresh of its affine weight majorants: centred moments merged up the tree,
// one root -> leaf pass over the seeds for the residual extremes, as the 2D refresh with three coordinates. The tree itself is
// built on the device by `build_tree_device3` ( tree_gpu.hip ).

#include "tree3.h"

#include <algorithm>
#include <cmath>

namespace sdot_amd {

// ---- weights -------------------------------------------------------------------------------------------------------

namespace {

__global__ void k_set_weights_user3( double4 *seeds, const int *order, const double *w_user, int n ) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k < n ) seeds[ k ].w = w_user[ order[ k ] ];
}

__global__ void k_set_weights_rank3( double4 *seeds, const double *w_rank, int n ) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k < n ) seeds[ k ].w = w_rank[ k ];
}

// ---- majorants ------------------------------------------------------------------------------------------------------

/// count; mean x, y, z; centred Sxx, Sxy, Sxz, Syy, Syz, Szz; mean w; centred Sxw, Syw, Szw; min w, max w
enum { T_M = 0, T_MX, T_MY, T_MZ, T_XX, T_XY, T_XZ, T_YY, T_YZ, T_ZZ, T_MW, T_XW, T_YW, T_ZW, T_WMIN, T_WMAX, T_NB };

struct Stat3 { double v[ T_NB ]; };

__device__ __forceinline__ void store3( double *S, int N, int node, const Stat3 &s ) {
#pragma unroll
    for ( int q = 0; q < T_NB; ++q ) S[ q * N + node ] = s.v[ q ];
}

__device__ __forceinline__ Stat3 load3( const double *S, int N, int node ) {
    Stat3 s;
#pragma unroll
    for ( int q = 0; q < T_NB; ++q ) s.v[ q ] = S[ q * N + node ];
    return s;
}

__global__ void k_major3_leaves( const int *lvl_nodes, int first, int count, const int *beg, const int *end, const double4 *seeds, double *S,
                                 int N ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= count ) return;
    const int node = lvl_nodes[ first + j ];
    const int b = beg[ node ], e = end[ node ];
    Stat3 s;
#pragma unroll
    for ( int q = 0; q < T_NB; ++q ) s.v[ q ] = 0;
    if ( e > b ) {
        double sx = 0, sy = 0, sz = 0, sw = 0, wmin = seeds[ b ].w, wmax = wmin;
        for ( int k = b; k < e; ++k ) {
            const double4 p = seeds[ k ];
            sx += p.x; sy += p.y; sz += p.z; sw += p.w;
            wmin = fmin( wmin, p.w ); wmax = fmax( wmax, p.w );
        }
        const double m = double( e - b );
        s.v[ T_M ] = m; s.v[ T_MX ] = sx / m; s.v[ T_MY ] = sy / m; s.v[ T_MZ ] = sz / m; s.v[ T_MW ] = sw / m;
        s.v[ T_WMIN ] = wmin; s.v[ T_WMAX ] = wmax;
        for ( int k = b; k < e; ++k ) {
            const double4 p = seeds[ k ];
            const double qx = p.x - s.v[ T_MX ], qy = p.y - s.v[ T_MY ], qz = p.z - s.v[ T_MZ ], dw = p.w - s.v[ T_MW ];
            s.v[ T_XX ] += qx * qx; s.v[ T_XY ] += qx * qy; s.v[ T_XZ ] += qx * qz;
            s.v[ T_YY ] += qy * qy; s.v[ T_YZ ] += qy * qz; s.v[ T_ZZ ] += qz * qz;
            s.v[ T_XW ] += qx * dw; s.v[ T_YW ] += qy * dw; s.v[ T_ZW ] += qz * dw;
        }
    }
    store3( S, N, node, s );
}

__global__ void k_major3_up( const int *lvl_nodes, int first, int count, int h, double *S, int N ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= count ) return;
    const int node = lvl_nodes[ first + j ];
    const int lc = node + 1, rc = node + ( 1 << ( h - 1 ) );
    const Stat3 a = load3( S, N, lc ), b = load3( S, N, rc );
    Stat3 s;
    if ( a.v[ T_M ] == 0 ) s = b;
    else if ( b.v[ T_M ] == 0 ) s = a;
    else {
        const double m = a.v[ T_M ] + b.v[ T_M ];
        const double fa = a.v[ T_M ] / m, fb = b.v[ T_M ] / m, f = a.v[ T_M ] * b.v[ T_M ] / m;
        const double dx = b.v[ T_MX ] - a.v[ T_MX ], dy = b.v[ T_MY ] - a.v[ T_MY ], dz = b.v[ T_MZ ] - a.v[ T_MZ ], dw = b.v[ T_MW ] - a.v[ T_MW ];
        s.v[ T_M ] = m;
        s.v[ T_MX ] = a.v[ T_MX ] * fa + b.v[ T_MX ] * fb;
        s.v[ T_MY ] = a.v[ T_MY ] * fa + b.v[ T_MY ] * fb;
        s.v[ T_MZ ] = a.v[ T_MZ ] * fa + b.v[ T_MZ ] * fb;
        s.v[ T_MW ] = a.v[ T_MW ] * fa + b.v[ T_MW ] * fb;
        s.v[ T_XX ] = a.v[ T_XX ] + b.v[ T_XX ] + dx * dx * f;
        s.v[ T_XY ] = a.v[ T_XY ] + b.v[ T_XY ] + dx * dy * f;
        s.v[ T_XZ ] = a.v[ T_XZ ] + b.v[ T_XZ ] + dx * dz * f;
        s.v[ T_YY ] = a.v[ T_YY ] + b.v[ T_YY ] + dy * dy * f;
        s.v[ T_YZ ] = a.v[ T_YZ ] + b.v[ T_YZ ] + dy * dz * f;
        s.v[ T_ZZ ] = a.v[ T_ZZ ] + b.v[ T_ZZ ] + dz * dz * f;
        s.v[ T_XW ] = a.v[ T_XW ] + b.v[ T_XW ] + dx * dw * f;
        s.v[ T_YW ] = a.v[ T_YW ] + b.v[ T_YW ] + dy * dw * f;
        s.v[ T_ZW ] = a.v[ T_ZW ] + b.v[ T_ZW ] + dz * dw * f;
        s.v[ T_WMIN ] = fmin( a.v[ T_WMIN ], b.v[ T_WMIN ] );
        s.v[ T_WMAX ] = fmax( a.v[ T_WMAX ], b.v[ T_WMAX ] );
    }
    store3( S, N, node, s );
}

/// rows `i` and `j` of the 3 x 4 system swapped when `|A[ j ][ c ]| > |A[ i ][ c ]|` ( fixed indices: no scratch )
__device__ __forceinline__ void pivot_rows( double ( &A )[ 3 ][ 4 ], int i, int j, int c ) {
    const bool sw = fabs( A[ j ][ c ] ) > fabs( A[ i ][ c ] );
#pragma unroll
    for ( int k = 0; k < 4; ++k ) {
        const double a = A[ i ][ k ], b = A[ j ][ k ];
        A[ i ][ k ] = sw ? b : a;
        A[ j ][ k ] = sw ? a : b;
    }
}

/// the least-squares slope of every node ( Gauss with partial pivoting on the centred normal equations ); NaN when there
/// is none to try ( fewer than `2 ( 3 + 1 )` seeds or equal weights ), and the residual extremes reset
__global__ void k_major3_fit( const double *S, int N, double *fit, unsigned long long *rk ) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if ( m >= N ) return;
    const Stat3 s = load3( S, N, m );
    double f0 = NAN, f1 = NAN, f2 = NAN;
    if ( s.v[ T_M ] >= 8 && s.v[ T_WMAX ] - s.v[ T_WMIN ] > 0 ) {
        double A[ 3 ][ 4 ] = { { s.v[ T_XX ], s.v[ T_XY ], s.v[ T_XZ ], s.v[ T_XW ] },
                               { s.v[ T_XY ], s.v[ T_YY ], s.v[ T_YZ ], s.v[ T_YW ] },
                               { s.v[ T_XZ ], s.v[ T_YZ ], s.v[ T_ZZ ], s.v[ T_ZW ] } };
        pivot_rows( A, 0, 1, 0 );
        pivot_rows( A, 0, 2, 0 );
        bool ok = fabs( A[ 0 ][ 0 ] ) > 0;
        if ( ok ) {
#pragma unroll
            for ( int i = 1; i < 3; ++i ) {
                const double f = A[ i ][ 0 ] / A[ 0 ][ 0 ];
#pragma unroll
                for ( int k = 0; k < 4; ++k ) A[ i ][ k ] -= f * A[ 0 ][ k ];
            }
            pivot_rows( A, 1, 2, 1 );
            ok = fabs( A[ 1 ][ 1 ] ) > 0;
        }
        if ( ok ) {
            const double f = A[ 2 ][ 1 ] / A[ 1 ][ 1 ];
#pragma unroll
            for ( int k = 1; k < 4; ++k ) A[ 2 ][ k ] -= f * A[ 1 ][ k ];
            ok = fabs( A[ 2 ][ 2 ] ) > 0;
        }
        if ( ok ) {
            f2 = A[ 2 ][ 3 ] / A[ 2 ][ 2 ];
            f1 = ( A[ 1 ][ 3 ] - A[ 1 ][ 2 ] * f2 ) / A[ 1 ][ 1 ];
            f0 = ( A[ 0 ][ 3 ] - A[ 0 ][ 1 ] * f1 - A[ 0 ][ 2 ] * f2 ) / A[ 0 ][ 0 ];
        }
    }
    fit[ m ] = f0; fit[ N + m ] = f1; fit[ 2 * N + m ] = f2;
    rk[ m ] = ~0ull;
    rk[ N + m ] = 0ull;
    rk[ 2 * N + m ] = 0ull;
}

/// one thread per seed ( rank order ), root -> leaf: the residual `w - a . y` of every node on the path, reduced over the
/// lanes that share the node, one atomic per segment
__global__ void k_major3_spread( const double4 *seeds, int n, const int *node_end, int depth, const double *fit, int N,
                                 unsigned long long *rk ) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool live = k < n;
    const double4 p = live ? seeds[ k ] : make_double4( 0, 0, 0, 0 );
    int m = 0;
    for ( int h = depth; h >= 1; --h ) {
        const int node = live ? m : -1;
        double rmin = INFINITY, rmax = -INFINITY, amax = 0;
        bool has = false;
        if ( live ) {
            const double f0 = fit[ m ], f1 = fit[ N + m ], f2 = fit[ 2 * N + m ];
            if ( f0 == f0 ) {
                const double ay = f0 * p.x + f1 * p.y + f2 * p.z;
                const double r = p.w - ay;
                rmin = r; rmax = r; amax = fabs( ay );
                has = true;
            }
        }
        for ( int off = 1; off < 64; off <<= 1 ) {
            const int on = __shfl_down( node, off, 64 );
            const double omin = __shfl_down( rmin, off, 64 ), omax = __shfl_down( rmax, off, 64 ), oa = __shfl_down( amax, off, 64 );
            if ( lane + off < 64 && on == node ) { rmin = fmin( rmin, omin ); rmax = fmax( rmax, omax ); amax = fmax( amax, oa ); }
        }
        const int prev = __shfl_up( node, 1, 64 );
        const bool head = live && ( lane == 0 || prev != node );
        if ( head && has ) {
            atomicMin( &rk[ m ], ordered_key( rmin ) );
            atomicMax( &rk[ N + m ], ordered_key( rmax ) );
            atomicMax( &rk[ 2 * N + m ], ordered_key( amax ) );
        }
        if ( h > 1 && live ) {
            const int lc = m + 1, rc = m + ( 1 << ( h - 1 ) );
            m = k < node_end[ lc ] ? lc : rc;
        }
    }
}

/// the decision of every node ( `bsp_build_level.h:176-208`, three coordinates ) and its walk record
__global__ void k_major3_final( const double *S, int N, const double *fit, const unsigned long long *rk, const double *box, const int *beg,
                                const int *end, NodeRec3 *nodes ) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if ( m >= N ) return;
    NodeRec3 r;
#pragma unroll
    for ( int d = 0; d < 3; ++d ) { r.lo[ d ] = box[ 6 * m + d ]; r.hi[ d ] = box[ 6 * m + 3 + d ]; r.wa[ d ] = 0; }
    r.beg = beg[ m ];
    r.end = end[ m ];
    r.wb = 0;
    const double cnt = S[ T_M * N + m ];
    if ( cnt > 0 ) {
        const double wmin = S[ T_WMIN * N + m ], wmax = S[ T_WMAX * N + m ], spread = wmax - wmin;
        const double f[ 3 ] = { fit[ m ], fit[ N + m ], fit[ 2 * N + m ] };
        bool accept = false;
        double bb = wmax, amax = 0;
        if ( f[ 0 ] == f[ 0 ] ) {
            const double rmin = ordered_value( rk[ m ] ), rmax = ordered_value( rk[ N + m ] );
            const double u = 1.0 - 3.0 / ( cnt - 1 );
            const double by_chance = sqrt( u > 0 ? u : 0.0 );
            bool well = true;
#pragma unroll
            for ( int d = 0; d < 3; ++d ) {
                const double reach = fmax( fabs( r.lo[ d ] ), fabs( r.hi[ d ] ) );
                if ( fabs( f[ d ] ) * ( r.hi[ d ] - r.lo[ d ] ) > 8 * spread || fabs( f[ d ] ) * reach > 100 * spread ) well = false;
            }
            if ( well && rmax - rmin < 0.85 * by_chance * spread ) {
                accept = true;
                bb = rmax;
                amax = ordered_value( rk[ 2 * N + m ] );
            }
        }
        if ( accept ) { r.wa[ 0 ] = f[ 0 ]; r.wa[ 1 ] = f[ 1 ]; r.wa[ 2 ] = f[ 2 ]; }
        r.wb = bb + 1e-6 * ( fabs( bb ) + spread + amax );
    }
    nodes[ m ] = r;
}

} // namespace

void set_weights_user3( DevTree3 &t, const double *w_user, hipStream_t st ) {
    k_set_weights_user3<<<blocks_for( t.n, 256 ), 256, 0, st>>>( t.seeds, t.order, w_user, int( t.n ) );
    check_launch( "k_set_weights_user3" );
}

void set_weights_rank3( DevTree3 &t, const double *w_rank, hipStream_t st ) {
    k_set_weights_rank3<<<blocks_for( t.n, 256 ), 256, 0, st>>>( t.seeds, w_rank, int( t.n ) );
    check_launch( "k_set_weights_rank3" );
}

void refresh_majorants3( DevTree3 &t, hipStream_t st ) {
    const int N = int( t.nb_nodes ), depth = t.depth;
    Timing::Pending tp;
    const bool timed = timing().begin( TK_MAJORANTS, st, tp );
    {
        const int L = depth - 1, first = ( 1 << L ) - 1, count = 1 << L;
        k_major3_leaves<<<blocks_for( count, 256 ), 256, 0, st>>>( t.lvl_nodes, first, count, t.node_beg, t.node_end, t.seeds, t.stat, N );
        check_launch( "k_major3_leaves" );
    }
    for ( int L = depth - 2; L >= 0; --L ) {
        const int first = ( 1 << L ) - 1, count = 1 << L, h = depth - L;
        k_major3_up<<<blocks_for( count, 256 ), 256, 0, st>>>( t.lvl_nodes, first, count, h, t.stat, N );
        check_launch( "k_major3_up" );
    }
    k_major3_fit<<<blocks_for( N, 256 ), 256, 0, st>>>( t.stat, N, t.fit, t.rkeys );
    check_launch( "k_major3_fit" );
    k_major3_spread<<<blocks_for( t.n, 256 ), 256, 0, st>>>( t.seeds, int( t.n ), t.node_end, depth, t.fit, N, t.rkeys );
    check_launch( "k_major3_spread" );
    k_major3_final<<<blocks_for( N, 256 ), 256, 0, st>>>( t.stat, N, t.fit, t.rkeys, t.node_box, t.node_beg, t.node_end, t.nodes );
    check_launch( "k_major3_final" );
    if ( timed ) timing().end( tp, st );
}

} // namespace sdot_amd
