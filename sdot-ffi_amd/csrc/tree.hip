// The BSP tree: host build ( median cut on the longest axis, level by level ) and the device-side
// refresh of the affine weight majorants.
//
// Majorant rule ( reference `bsp_build_level.h:84-209` ): least-squares slope `a` of `w` over the
// node's seeds, kept when the residual spread `max r - min r` ( `r = w - a . y` ) is below
// `0.85 sqrt( 1 - 2 / ( m - 1 ) )` times the weight spread and the slope is well behaved over the
// node's extent; then `b = max( w - a . y )` plus `1e-6 ( |b| + spread + max |a . y| )`. A majorant
// only prunes: any `( a, b )` that majorizes gives the same cells, so the device computes it its own
// way -- centred moments merged up the tree ( Chan's parallel formula ), then one pass over the seeds
// that walks root -> leaf and takes the residual extremes of every node on the path with a
// wave-segmented min / max and one ordered-integer atomic per segment.

#include "tree.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>

namespace sdot_amd {

int max_depth_for( I64 n, I64 leaf ) {
    if ( leaf < 1 ) leaf = 1;
    if ( n <= leaf ) return 1;
    return int( std::ceil( std::log2( double( n ) / double( leaf ) ) ) ) + 1;
}

// ---- host build -----------------------------------------------------------------------------------

namespace {

struct HostBuild {
    I64 n;
    int depth;
    std::vector<double> x, y;     ///< permuted positions
    std::vector<int> ord;         ///< permuted user indices
};

void run_parallel( int nt, const std::function<void( int )> &f ) {
    if ( nt <= 1 ) { f( 0 ); return; }
    std::vector<std::thread> ts;
    for ( int i = 0; i < nt; ++i ) ts.emplace_back( f, i );
    for ( auto &t : ts ) t.join();
}

} // namespace

void build_tree( DevTree &t, I64 n, const double *pos, int leaf, hipStream_t st, const std::string &tag ) {
    SDOT_CHECK( n >= 1, "a tree needs at least one seed" );
    SDOT_CHECK( n < ( I64( 1 ) << 31 ) - 8, "more seeds than 32-bit ranks can index" );
    const int depth = max_depth_for( n, leaf );
    const I64 N = ( I64( 1 ) << depth ) - 1;
    t.n = n;
    t.depth = depth;
    t.leaf = leaf;
    t.nb_nodes = N;

    std::vector<double> X( ( size_t( n ) ) ), Y( ( size_t( n ) ) );
    std::vector<int> ord( ( size_t( n ) ) );
    for ( I64 i = 0; i < n; ++i ) { X[ size_t( i ) ] = pos[ 2 * i ]; Y[ size_t( i ) ] = pos[ 2 * i + 1 ]; ord[ size_t( i ) ] = int( i ); }

    // heap numbering while building, preorder at the end
    std::vector<int> hbeg( ( size_t( N ) ) ), hend( ( size_t( N ) ) );
    std::vector<double> hbox( size_t( N ) * 4, 0.0 );
    std::vector<int> beg{ 0 }, end{ int( n ) };
    const int hw = int( std::max( 1u, std::min( 16u, std::thread::hardware_concurrency() ) ) );
    std::vector<int> idx( ( size_t( n ) ) );
    for ( int level = 0; level < depth; ++level ) {
        const I64 first = ( I64( 1 ) << level ) - 1;
        const int nb = int( beg.size() );
        std::vector<int> mid( end );
        // one node: its box, the longest axis, the median of its slice
        auto do_node = [&]( int j ) {
            const int b = beg[ size_t( j ) ], e = end[ size_t( j ) ];
            const I64 node = first + j;
            hbeg[ size_t( node ) ] = b;
            hend[ size_t( node ) ] = e;
            if ( e <= b ) return;
            double lo[ 2 ] = { X[ size_t( b ) ], Y[ size_t( b ) ] }, hi[ 2 ] = { lo[ 0 ], lo[ 1 ] };
            for ( int k = b + 1; k < e; ++k ) {
                lo[ 0 ] = std::min( lo[ 0 ], X[ size_t( k ) ] ); hi[ 0 ] = std::max( hi[ 0 ], X[ size_t( k ) ] );
                lo[ 1 ] = std::min( lo[ 1 ], Y[ size_t( k ) ] ); hi[ 1 ] = std::max( hi[ 1 ], Y[ size_t( k ) ] );
            }
            double *bx = &hbox[ size_t( node ) * 4 ];
            bx[ 0 ] = lo[ 0 ]; bx[ 1 ] = lo[ 1 ]; bx[ 2 ] = hi[ 0 ]; bx[ 3 ] = hi[ 1 ];
            const int ax = ( hi[ 1 ] - lo[ 1 ] > hi[ 0 ] - lo[ 0 ] ) ? 1 : 0;
            if ( e - b > leaf && hi[ ax ] > lo[ ax ] ) {
                const int m = b + ( e - b ) / 2;
                const std::vector<double> &K = ax ? Y : X;
                for ( int k = b; k < e; ++k ) idx[ size_t( k ) ] = k;
                std::nth_element( idx.begin() + b, idx.begin() + m, idx.begin() + e, [&]( int u, int v ) {
                    return K[ size_t( u ) ] < K[ size_t( v ) ] || ( K[ size_t( u ) ] == K[ size_t( v ) ] && u < v );
                } );
                std::vector<double> tx( size_t( e - b ) ), ty( size_t( e - b ) );
                std::vector<int> to( size_t( e - b ) );
                for ( int k = b; k < e; ++k ) {
                    const int s = idx[ size_t( k ) ];
                    tx[ size_t( k - b ) ] = X[ size_t( s ) ]; ty[ size_t( k - b ) ] = Y[ size_t( s ) ]; to[ size_t( k - b ) ] = ord[ size_t( s ) ];
                }
                std::copy( tx.begin(), tx.end(), X.begin() + b );
                std::copy( ty.begin(), ty.end(), Y.begin() + b );
                std::copy( to.begin(), to.end(), ord.begin() + b );
                mid[ size_t( j ) ] = m;
            }
        };
        if ( nb >= 4 * hw ) {
            run_parallel( hw, [&]( int th ) { for ( int j = th; j < nb; j += hw ) do_node( j ); } );
        } else {
            for ( int j = 0; j < nb; ++j ) do_node( j );
        }
        if ( level + 1 < depth ) {
            std::vector<int> nb_( size_t( 2 * nb ) ), ne_( size_t( 2 * nb ) );
            for ( int j = 0; j < nb; ++j ) {
                nb_[ size_t( 2 * j ) ] = beg[ size_t( j ) ]; ne_[ size_t( 2 * j ) ] = mid[ size_t( j ) ];
                nb_[ size_t( 2 * j + 1 ) ] = mid[ size_t( j ) ]; ne_[ size_t( 2 * j + 1 ) ] = end[ size_t( j ) ];
            }
            beg.swap( nb_ );
            end.swap( ne_ );
        }
    }

    // the seeds of each leaf in ascending user index ( a canonical order: the cut order inside a leaf
    // is then the same whatever selection built the tree )
    {
        const I64 first = ( I64( 1 ) << ( depth - 1 ) ) - 1;
        run_parallel( hw, [&]( int th ) {
            std::vector<std::pair<int, int>> tmp;
            std::vector<double> tx, ty;
            for ( I64 h = first + th; h < N; h += hw ) {
                const int b = hbeg[ size_t( h ) ], e = hend[ size_t( h ) ];
                if ( e - b < 2 ) continue;
                tmp.clear();
                for ( int k = b; k < e; ++k ) tmp.emplace_back( ord[ size_t( k ) ], k );
                std::sort( tmp.begin(), tmp.end() );
                tx.resize( tmp.size() ); ty.resize( tmp.size() );
                for ( size_t q = 0; q < tmp.size(); ++q ) { tx[ q ] = X[ size_t( tmp[ q ].second ) ]; ty[ q ] = Y[ size_t( tmp[ q ].second ) ]; }
                for ( size_t q = 0; q < tmp.size(); ++q ) { X[ size_t( b ) + q ] = tx[ q ]; Y[ size_t( b ) + q ] = ty[ q ]; ord[ size_t( b ) + q ] = tmp[ q ].first; }
            }
        } );
    }

    // heap -> preorder ( reference `AaBsp.py:514-532` )
    std::vector<int> pre( ( size_t( N ) ), 0 );
    for ( int level = 0; level + 1 < depth; ++level ) {
        const int h = depth - level;
        for ( I64 i = ( I64( 1 ) << level ) - 1; i < ( I64( 1 ) << ( level + 1 ) ) - 1; ++i ) {
            pre[ size_t( 2 * i + 1 ) ] = pre[ size_t( i ) ] + 1;
            pre[ size_t( 2 * i + 2 ) ] = pre[ size_t( i ) ] + int( I64( 1 ) << ( h - 1 ) );
        }
    }
    std::vector<int> pbeg( ( size_t( N ) ) ), pend( ( size_t( N ) ) );
    std::vector<double> pbox( size_t( N ) * 4 );
    for ( I64 h = 0; h < N; ++h ) {
        const int p = pre[ size_t( h ) ];
        pbeg[ size_t( p ) ] = hbeg[ size_t( h ) ];
        pend[ size_t( p ) ] = hend[ size_t( h ) ];
        for ( int c = 0; c < 4; ++c ) pbox[ size_t( p ) * 4 + c ] = hbox[ size_t( h ) * 4 + c ];
    }
    t.h_lvl_nodes = pre;                                  // heap rank `h` ( level by level ) -> preorder node

    std::vector<int> rank_of( ( size_t( n ) ) );
    for ( I64 k = 0; k < n; ++k ) rank_of[ size_t( ord[ size_t( k ) ] ) ] = int( k );
    std::vector<double4> seeds( ( size_t( n ) ) );
    for ( I64 k = 0; k < n; ++k ) seeds[ size_t( k ) ] = make_double4( X[ size_t( k ) ], Y[ size_t( k ) ], 0.0, 0.0 );

    Arena &A = arena();
    t.nodes = A.get<NodeRec>( tag + ".nodes", size_t( N ) );
    t.nodes32 = A.get<NodeF>( tag + ".nodes32", size_t( N ) );
    t.node_box = A.get<double>( tag + ".node_box", size_t( N ) * 4 );
    t.node_beg = A.get<int>( tag + ".node_beg", size_t( N ) );
    t.node_end = A.get<int>( tag + ".node_end", size_t( N ) );
    t.lvl_nodes = A.get<int>( tag + ".lvl_nodes", size_t( N ) );
    t.order = A.get<int>( tag + ".order", size_t( n ) );
    t.rank_of = A.get<int>( tag + ".rank_of", size_t( n ) );
    t.seeds = A.get<double4>( tag + ".seeds", size_t( n ) );
    t.stat = A.get<double>( tag + ".stat", size_t( N ) * 11 );
    t.fit = A.get<double>( tag + ".fit", size_t( N ) * 2 );
    t.rkeys = A.get<unsigned long long>( tag + ".rkeys", size_t( N ) * 3 );
    SDOT_HIP( hipMemcpyAsync( t.node_box, pbox.data(), sizeof( double ) * size_t( N ) * 4, hipMemcpyHostToDevice, st ) );
    SDOT_HIP( hipMemcpyAsync( t.node_beg, pbeg.data(), sizeof( int ) * size_t( N ), hipMemcpyHostToDevice, st ) );
    SDOT_HIP( hipMemcpyAsync( t.node_end, pend.data(), sizeof( int ) * size_t( N ), hipMemcpyHostToDevice, st ) );
    SDOT_HIP( hipMemcpyAsync( t.lvl_nodes, pre.data(), sizeof( int ) * size_t( N ), hipMemcpyHostToDevice, st ) );
    SDOT_HIP( hipMemcpyAsync( t.order, ord.data(), sizeof( int ) * size_t( n ), hipMemcpyHostToDevice, st ) );
    SDOT_HIP( hipMemcpyAsync( t.rank_of, rank_of.data(), sizeof( int ) * size_t( n ), hipMemcpyHostToDevice, st ) );
    SDOT_HIP( hipMemcpyAsync( t.seeds, seeds.data(), sizeof( double4 ) * size_t( n ), hipMemcpyHostToDevice, st ) );
    // the host vectors die with this scope: the copies must be done before returning
    SDOT_HIP( hipStreamSynchronize( st ) );
}

// ---- weights -------------------------------------------------------------------------------------

__global__ void k_set_weights_user( double4 *seeds, const int *order, const double *w_user, int n ) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k < n ) seeds[ k ].z = w_user[ order[ k ] ];
}

__global__ void k_set_weights_rank( double4 *seeds, const double *w_rank, int n ) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k < n ) seeds[ k ].z = w_rank[ k ];
}

void set_weights_user( DevTree &t, const double *w_user, hipStream_t st ) {
    k_set_weights_user<<<blocks_for( t.n, 256 ), 256, 0, st>>>( t.seeds, t.order, w_user, int( t.n ) );
    check_launch( "k_set_weights_user" );
}

void set_weights_rank( DevTree &t, const double *w_rank, hipStream_t st ) {
    k_set_weights_rank<<<blocks_for( t.n, 256 ), 256, 0, st>>>( t.seeds, w_rank, int( t.n ) );
    check_launch( "k_set_weights_rank" );
}

// ---- majorants -----------------------------------------------------------------------------------

/// statistics of a node, SoA over the nodes: count, mean x, mean y, centred Sxx, Sxy, Syy, mean w,
/// centred Sxw, Syw, min w, max w
enum { S_M = 0, S_MX, S_MY, S_XX, S_XY, S_YY, S_MW, S_XW, S_YW, S_WMIN, S_WMAX, S_NB };

struct Stat { double m, mx, my, xx, xy, yy, mw, xw, yw, wmin, wmax; };

__device__ __forceinline__ void store_stat( double *S, int N, int node, const Stat &s ) {
    S[ S_M * N + node ] = s.m;   S[ S_MX * N + node ] = s.mx;  S[ S_MY * N + node ] = s.my;
    S[ S_XX * N + node ] = s.xx; S[ S_XY * N + node ] = s.xy;  S[ S_YY * N + node ] = s.yy;
    S[ S_MW * N + node ] = s.mw; S[ S_XW * N + node ] = s.xw;  S[ S_YW * N + node ] = s.yw;
    S[ S_WMIN * N + node ] = s.wmin; S[ S_WMAX * N + node ] = s.wmax;
}

__device__ __forceinline__ Stat load_stat( const double *S, int N, int node ) {
    Stat s;
    s.m = S[ S_M * N + node ];   s.mx = S[ S_MX * N + node ];  s.my = S[ S_MY * N + node ];
    s.xx = S[ S_XX * N + node ]; s.xy = S[ S_XY * N + node ];  s.yy = S[ S_YY * N + node ];
    s.mw = S[ S_MW * N + node ]; s.xw = S[ S_XW * N + node ];  s.yw = S[ S_YW * N + node ];
    s.wmin = S[ S_WMIN * N + node ]; s.wmax = S[ S_WMAX * N + node ];
    return s;
}

/// the leaves ( last level ): centred moments of their few seeds, two passes
__global__ void k_major_leaves( const int *lvl_nodes, int first, int count, const int *beg, const int *end, const double4 *seeds,
                                double *S, int N ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= count ) return;
    const int node = lvl_nodes[ first + j ];
    const int b = beg[ node ], e = end[ node ];
    Stat s{};
    if ( e > b ) {
        double sx = 0, sy = 0, sw = 0;
        s.wmin = seeds[ b ].z;
        s.wmax = s.wmin;
        for ( int k = b; k < e; ++k ) {
            const double4 p = seeds[ k ];
            sx += p.x; sy += p.y; sw += p.z;
            s.wmin = fmin( s.wmin, p.z ); s.wmax = fmax( s.wmax, p.z );
        }
        s.m = double( e - b );
        s.mx = sx / s.m; s.my = sy / s.m; s.mw = sw / s.m;
        for ( int k = b; k < e; ++k ) {
            const double4 p = seeds[ k ];
            const double qx = p.x - s.mx, qy = p.y - s.my, dw = p.z - s.mw;
            s.xx += qx * qx; s.xy += qx * qy; s.yy += qy * qy;
            s.xw += qx * dw; s.yw += qy * dw;
        }
    }
    store_stat( S, N, node, s );
}

/// one level up: a node's moments from its two children ( Chan et al.'s pairwise combination )
__global__ void k_major_up( const int *lvl_nodes, int first, int count, int h, double *S, int N ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= count ) return;
    const int node = lvl_nodes[ first + j ];
    const int lc = node + 1, rc = node + ( 1 << ( h - 1 ) );
    const Stat a = load_stat( S, N, lc ), b = load_stat( S, N, rc );
    Stat s;
    if ( a.m == 0 ) s = b;
    else if ( b.m == 0 ) s = a;
    else {
        s.m = a.m + b.m;
        const double fa = a.m / s.m, fb = b.m / s.m, f = a.m * b.m / s.m;
        const double dx = b.mx - a.mx, dy = b.my - a.my, dw = b.mw - a.mw;
        s.mx = a.mx * fa + b.mx * fb;
        s.my = a.my * fa + b.my * fb;
        s.mw = a.mw * fa + b.mw * fb;
        s.xx = a.xx + b.xx + dx * dx * f;
        s.xy = a.xy + b.xy + dx * dy * f;
        s.yy = a.yy + b.yy + dy * dy * f;
        s.xw = a.xw + b.xw + dx * dw * f;
        s.yw = a.yw + b.yw + dy * dw * f;
        s.wmin = fmin( a.wmin, b.wmin );
        s.wmax = fmax( a.wmax, b.wmax );
    }
    store_stat( S, N, node, s );
}

/// the least-squares slope of every node ( Gauss with partial pivoting on the centred normal
/// equations ); NaN when there is none to try, and the residual extremes reset
__global__ void k_major_fit( const double *S, int N, double *fit, unsigned long long *rk ) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if ( m >= N ) return;
    const Stat s = load_stat( S, N, m );
    double fx = NAN, fy = NAN;
    const double spread = s.wmax - s.wmin;
    if ( s.m >= 6 && spread > 0 ) {
        double A[ 2 ][ 3 ] = { { s.xx, s.xy, s.xw }, { s.xy, s.yy, s.yw } };
        bool ok = true;
        if ( fabs( A[ 1 ][ 0 ] ) > fabs( A[ 0 ][ 0 ] ) )
            for ( int c = 0; c < 3; ++c ) { const double t = A[ 0 ][ c ]; A[ 0 ][ c ] = A[ 1 ][ c ]; A[ 1 ][ c ] = t; }
        if ( ! ( fabs( A[ 0 ][ 0 ] ) > 0 ) ) ok = false;
        if ( ok ) {
            const double f = A[ 1 ][ 0 ] / A[ 0 ][ 0 ];
            A[ 1 ][ 1 ] -= f * A[ 0 ][ 1 ];
            A[ 1 ][ 2 ] -= f * A[ 0 ][ 2 ];
            if ( ! ( fabs( A[ 1 ][ 1 ] ) > 0 ) ) ok = false;
        }
        if ( ok ) {
            fy = A[ 1 ][ 2 ] / A[ 1 ][ 1 ];
            fx = ( A[ 0 ][ 2 ] - A[ 0 ][ 1 ] * fy ) / A[ 0 ][ 0 ];
        }
    }
    fit[ m ] = fx;
    fit[ N + m ] = fy;
    rk[ m ] = ~0ull;           // min r
    rk[ N + m ] = 0ull;        // max r
    rk[ 2 * N + m ] = 0ull;    // max | a . y |
}

/// one thread per seed ( rank order ), root -> leaf: the residual `w - a . y` of every node on the
/// path, reduced over the lanes that share the node, one atomic per segment
__global__ void k_major_spread( const double4 *seeds, int n, const int *node_end, int depth, const double *fit, int N,
                                unsigned long long *rk ) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool live = k < n;
    double4 p = live ? seeds[ k ] : make_double4( 0, 0, 0, 0 );
    int m = 0;
    for ( int h = depth; h >= 1; --h ) {
        const int node = live ? m : -1;
        double rmin = INFINITY, rmax = -INFINITY, amax = 0;
        bool has = false;
        if ( live ) {
            const double fx = fit[ m ], fy = fit[ N + m ];
            if ( fx == fx ) {
                const double ay = fx * p.x + fy * p.y;
                const double r = p.z - ay;
                rmin = r; rmax = r; amax = fabs( ay );
                has = true;
            }
        }
        // segmented reduction towards the first lane of each run of equal nodes
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

/// the decision of every node ( `bsp_build_level.h:176-208` ) and its walk record
__global__ void k_major_final( const double *S, int N, const double *fit, const unsigned long long *rk, const double *box, const int *beg,
                               const int *end, NodeRec *nodes, NodeF *nodes32 ) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if ( m >= N ) return;
    NodeRec r;
    r.lo_x = box[ 4 * m ]; r.lo_y = box[ 4 * m + 1 ]; r.hi_x = box[ 4 * m + 2 ]; r.hi_y = box[ 4 * m + 3 ];
    r.beg = beg[ m ];
    r.end = end[ m ];
    r.wa_x = 0; r.wa_y = 0; r.wb = 0;
    const double cnt = S[ S_M * N + m ];
    if ( cnt > 0 ) {
        const double wmin = S[ S_WMIN * N + m ], wmax = S[ S_WMAX * N + m ], spread = wmax - wmin;
        const double fx = fit[ m ], fy = fit[ N + m ];
        bool accept = false;
        double bb = wmax, amax = 0;
        if ( fx == fx ) {
            const double rmin = ordered_value( rk[ m ] ), rmax = ordered_value( rk[ N + m ] );
            const double u = 1.0 - 2.0 / ( cnt - 1 );
            const double by_chance = sqrt( u > 0 ? u : 0.0 );
            bool well = true;
            const double reach_x = fmax( fabs( r.lo_x ), fabs( r.hi_x ) ), reach_y = fmax( fabs( r.lo_y ), fabs( r.hi_y ) );
            if ( fabs( fx ) * ( r.hi_x - r.lo_x ) > 8 * spread || fabs( fx ) * reach_x > 100 * spread ) well = false;
            if ( fabs( fy ) * ( r.hi_y - r.lo_y ) > 8 * spread || fabs( fy ) * reach_y > 100 * spread ) well = false;
            if ( well && rmax - rmin < 0.85 * by_chance * spread ) {
                accept = true;
                bb = rmax;
                amax = ordered_value( rk[ 2 * N + m ] );
            }
        }
        if ( accept ) { r.wa_x = fx; r.wa_y = fy; }
        r.wb = bb + 1e-6 * ( fabs( bb ) + spread + amax );
    }
    nodes[ m ] = r;
    if ( nodes32 ) {
        NodeF f;
        f.lo_x = __double2float_rd( r.lo_x ); f.lo_y = __double2float_rd( r.lo_y );
        f.hi_x = __double2float_ru( r.hi_x ); f.hi_y = __double2float_ru( r.hi_y );
        f.wa_x = float( r.wa_x ); f.wa_y = float( r.wa_y );
        // `( wa - wa_f ) . y` on the box, bounded by the farthest corner, goes into `wb` before it is rounded up
        const double ex = fabs( r.wa_x - double( f.wa_x ) ) * fmax( fabs( r.lo_x ), fabs( r.hi_x ) );
        const double ey = fabs( r.wa_y - double( f.wa_y ) ) * fmax( fabs( r.lo_y ), fabs( r.hi_y ) );
        f.wb = __double2float_ru( r.wb + ex + ey );
        f.beg = r.beg < r.end ? r.beg : -1;
        nodes32[ m ] = f;
    }
}

void refresh_majorants( DevTree &t, hipStream_t st ) {
    const int N = int( t.nb_nodes ), depth = t.depth;
    Timing::Pending tp;
    const bool timed = timing().begin( TK_MAJORANTS, st, tp );
    // the leaves, then each level up
    {
        const int L = depth - 1, first = ( 1 << L ) - 1, count = 1 << L;
        k_major_leaves<<<blocks_for( count, 256 ), 256, 0, st>>>( t.lvl_nodes, first, count, t.node_beg, t.node_end, t.seeds, t.stat, N );
        check_launch( "k_major_leaves" );
    }
    for ( int L = depth - 2; L >= 0; --L ) {
        const int first = ( 1 << L ) - 1, count = 1 << L, h = depth - L;
        k_major_up<<<blocks_for( count, 256 ), 256, 0, st>>>( t.lvl_nodes, first, count, h, t.stat, N );
        check_launch( "k_major_up" );
    }
    k_major_fit<<<blocks_for( N, 256 ), 256, 0, st>>>( t.stat, N, t.fit, t.rkeys );
    check_launch( "k_major_fit" );
    k_major_spread<<<blocks_for( t.n, 256 ), 256, 0, st>>>( t.seeds, int( t.n ), t.node_end, depth, t.fit, N, t.rkeys );
    check_launch( "k_major_spread" );
    k_major_final<<<blocks_for( N, 256 ), 256, 0, st>>>( t.stat, N, t.fit, t.rkeys, t.node_box, t.node_beg, t.node_end, t.nodes, t.nodes32 );
    check_launch( "k_major_final" );
    if ( timed ) timing().end( tp, st );
}

} // namespace sdot_amd
