// ============================================================================================
// THE ORACLE -- test infrastructure only.
//
// A plain FP64 C++ restatement of the reference's CPU semi-discrete OT solver for the 2D path,
// used as the checker of the HIP library: the tests, `__graft_entry__.smoke()` and the CPU
// baseline leg of `bench.py` may load it, nothing in the product does.
//
// What it restates ( reference file:line, paths under `src/sdot/include/sdot/` unless noted ):
//   * the BSP tree: median cut on the longest axis, leaf <= 10, perfect tree, heap -> preorder
//       `bsp_build_level.h:43` ( quickselect ), `:230` ( one level ), `src/sdot/AaBsp.py:264`
//       ( depth ), `src/sdot/AaBsp.py:545` ( levels, children slices ), `src/sdot/AaBsp.py:514`
//       ( preorder numbering ).
//   * the affine weight majorant of a node: `bsp_build_level.h:84-209`.
//   * the walk: `cell/Providers.h:167` ( nearest child first, prune at pop time ), the pruning
//       test `cell/Pruning.h:46` ( exact per-axis clamp ), the bisector `cell/Plane.h:39`.
//   * the in-place convex polygon cut with the several-runs rule: `cell/Local2.h:206`, `:300`.
//   * densities: constant, and `Image` cut into pixel pieces `Image.cxx:106-232`.
//   * one cell's mass and facets `sdotplan/Sweep.h:64-128`, the Laplacian `sdotplan/Laplacian.h:40`.
//   * the Jacobi-preconditioned CG `sdotplan/Linear.cpp:44-88`.
//   * the damped Newton ( trials ), LOG -> LIN residual switch, HopelessDamping
//       `sdotplan/Newton.h:243-420`, `sdotplan/Report.h:37-57`.
//   * the start selection, target rescaling, moments and the stats / history layout
//       `sdotplan/Solve.h:112-371`, `sdotplan/Report.h:15-31`.
//
// Geometry runs in the SEED'S FRAME ( a vertex is stored as `x - p0` ): the half-space of the
// bisector of ( p0, w0 ) and ( q, wq ) is then `d . v <= ( ( w0 - wq ) + |d|^2 ) / 2`, `d = q - p0`,
// which is the reference's `Plane.h:39` half-space translated by `p0` ( the reference card does
// the same, `gpu/Cell2D.cuh:139-160` ). The weights' difference is taken first: `|d|^2` added to a
// weight would be lost for a pair far closer than `sqrt( |w| )`, and the two cells of such a pair
// would then leave a gap of width `|d|` between them. The HIP library uses the same frame, the same
// walk and the same cut, so its cells agree with these to rounding.
//
// Parity pins ( the tests that keep this restatement honest, all in `tests/` ):
//   tree           shape = f( n ) only, every seed once, boxes contain their subtree, majorant
//                  majorates ( reference tests test_PowerDiagram.py:773,801,835,870 )
//   cells          sum of masses = domain mass, one seed takes all, grid seeds give equal cells,
//                  weight shift invariance ( test_PowerDiagram.py:46,58,217,304 ); neighbour sets
//                  and areas against scipy's lifted convex hull / half-space intersection;
//                  brute-force O( n^2 ) cells
//   image          constant image == Lebesgue, one-pixel closed forms, exact polygon/pixel clipping
//   facets         Hessian == finite differences of the masses, zero row sums
//                  ( test_SdotPlanNd.py:184-243 )
//   CG             residual and dense least squares ( test_SdotPlanNd.py:680-727 )
//   Newton         max |m - nu| <= tol at the end, quadratic tail, warm start <= 6 steps
//                  ( test_SdotPlanNd.py:39,245-276 ), similarity start
// ============================================================================================

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

using I64 = long long;

static inline size_t sz( I64 v ) { return size_t( v ); }

static double now() {
    using namespace std::chrono;
    return duration<double>( steady_clock::now().time_since_epoch() ).count();
}

// ---- densities -------------------------------------------------------------------------------

/// kind 0: constant `value` on the box; kind 1: an image, `img[ j * nx + i ]` the value of pixel
/// `( i, j )`, pixel `( i, j )` covering `[ x0 + i hx, x0 + ( i + 1 ) hx ] x [ y0 + j hy, ... ]`;
/// kind 2: a sum of `ng` isotropic Gaussians, `gauss[ 4 q .. ]` = centre x, y, sigma, mass, each widened to
/// `sqrt( sigma^2 + conv_s^2 )` ( the width continuation, `sdotplan/convolved/SumOfGaussians.h:34` )
/// a convex polygon domain, the cell the cuts start from ( `StartCell.py`, `gpu/Cell2D.cuh:295-333, 465-478` ): vertices
/// counterclockwise, edge `i` ( vertex `i` to `i + 1` ) the cut `-5 - i`, its line `nx x + ny y <= off` ( unit outward normal );
/// empty: the box
struct Domain {
    std::vector<double> x, y, nx, ny, off;
    bool empty() const { return x.empty(); }
    void set( int nb, const double *xy ) {
        x.assign( size_t( nb ), 0.0 ); y = x; nx = x; ny = x; off = x;
        for ( int i = 0; i < nb; ++i ) { x[ size_t( i ) ] = xy[ 2 * i ]; y[ size_t( i ) ] = xy[ 2 * i + 1 ]; }
        for ( int i = 0; i < nb; ++i ) {
            const int j = ( i + 1 ) % nb;
            const double ex = x[ size_t( j ) ] - x[ size_t( i ) ], ey = y[ size_t( j ) ] - y[ size_t( i ) ], len = std::sqrt( ex * ex + ey * ey );
            nx[ size_t( i ) ] = ey / len;
            ny[ size_t( i ) ] = -ex / len;
            off[ size_t( i ) ] = nx[ size_t( i ) ] * x[ size_t( i ) ] + ny[ size_t( i ) ] * y[ size_t( i ) ];
        }
    }
};

struct Density {
    Domain        dom;             ///< the polygon domain ( empty: the box )
    int           kind  = 0;
    double        value = 1;
    int           nx = 0, ny = 0;
    double        x0 = 0, y0 = 0, hx = 1, hy = 1;
    const double *img = nullptr;
    int           ng = 0;
    const double *gauss = nullptr;
    double        conv_s = 0;
    double sigma( int q ) const { const double s = gauss[ 4 * q + 2 ]; return conv_s > 0 ? std::sqrt( s * s + conv_s * conv_s ) : s; }
};

// ---- the tree --------------------------------------------------------------------------------

struct Tree {
    I64                 n = 0, nb_nodes = 0;
    int                 depth = 0;
    std::vector<I64>    order;           ///< rank -> user index ( `seed_indices` )
    std::vector<I64>    rank_of;         ///< user index -> rank
    std::vector<I64>    beg, end;        ///< per node, preorder
    std::vector<double> box;             ///< [ N ][ lo/hi ][ 2 ]
    std::vector<double> wa, wb;          ///< [ N ][ 2 ], [ N ]
    std::vector<double> px, py;          ///< positions in rank order
    std::vector<double> w;               ///< weights in rank order
};

/// `AaBsp.py:264`: `1 if n <= leaf else ceil( log2( n / leaf ) ) + 1`
int max_depth_for( I64 n, I64 leaf ) {
    if ( leaf < 1 ) leaf = 1;
    if ( n <= leaf ) return 1;
    return int( std::ceil( std::log2( double( n ) / double( leaf ) ) ) ) + 1;
}

/// Hoare quickselect with a median-of-three pivot that is a value of the slice
/// ( `bsp_build_level.h:43-83` ): `perm[ b .. e )` rearranged so that rank `t - b` sits at `t`
static void select_median( std::vector<I64> &perm, const std::vector<double> &key_of, I64 b, I64 e, I64 t ) {
    auto key = [&]( I64 k ) { return key_of[ size_t( perm[ size_t( k ) ] ) ]; };
    I64 lo = b, hi = e;
    while ( hi - lo > 2 ) {
        const I64 c = lo + ( hi - lo ) / 2;
        const double k0 = key( lo ), k1 = key( c ), k2 = key( hi - 1 );
        const double pivot = k0 < k1 ? ( k1 < k2 ? k1 : ( k0 < k2 ? k2 : k0 ) )
                                     : ( k0 < k2 ? k0 : ( k1 < k2 ? k2 : k1 ) );
        I64 i = lo - 1, j = hi;
        for ( ;; ) {
            do { ++i; } while ( key( i ) < pivot );
            do { --j; } while ( key( j ) > pivot );
            if ( i >= j ) break;
            std::swap( perm[ size_t( i ) ], perm[ size_t( j ) ] );
        }
        if ( t <= j ) hi = j + 1;
        else          lo = j + 1;
    }
    if ( hi - lo == 2 && key( lo ) > key( lo + 1 ) )
        std::swap( perm[ size_t( lo ) ], perm[ size_t( lo + 1 ) ] );
}

/// the affine majorant `w <= a . y + b` of the slice `[ b, e )` of rank-ordered seeds
/// ( `bsp_build_level.h:84-209` ): least-squares slope kept only when it clearly tightens the
/// spread of the residuals and stays well behaved, `b` raised to majorize plus a relative margin
static void weight_majorant( const double *px, const double *py, const double *w, I64 b, I64 e, double *wa, double &wb ) {
    const I64 m = e - b;
    double wmin = w[ b ], wmax = w[ b ], wsum = 0, psx = 0, psy = 0;
    double plo[ 2 ] = { px[ b ], py[ b ] }, phi[ 2 ] = { px[ b ], py[ b ] };
    for ( I64 k = b; k < e; ++k ) {
        wmin = std::min( wmin, w[ k ] );
        wmax = std::max( wmax, w[ k ] );
        wsum += w[ k ];
        psx += px[ k ];
        psy += py[ k ];
        plo[ 0 ] = std::min( plo[ 0 ], px[ k ] ); phi[ 0 ] = std::max( phi[ 0 ], px[ k ] );
        plo[ 1 ] = std::min( plo[ 1 ], py[ k ] ); phi[ 1 ] = std::max( phi[ 1 ], py[ k ] );
    }
    const double spread = wmax - wmin;
    double a[ 2 ] = { 0, 0 };
    if ( m >= 2 * ( 2 + 1 ) && spread > 0 ) {
        const double inv = 1.0 / double( m );
        const double pm[ 2 ] = { psx * inv, psy * inv }, wm = wsum * inv;
        double A[ 2 ][ 3 ] = { { 0, 0, 0 }, { 0, 0, 0 } };
        for ( I64 k = b; k < e; ++k ) {
            const double dw = w[ k ] - wm;
            const double q[ 2 ] = { px[ k ] - pm[ 0 ], py[ k ] - pm[ 1 ] };
            for ( int i = 0; i < 2; ++i ) {
                for ( int j = 0; j < 2; ++j )
                    A[ i ][ j ] += q[ i ] * q[ j ];
                A[ i ][ 2 ] += q[ i ] * dw;
            }
        }
        bool ok = true;
        for ( int c = 0; c < 2 && ok; ++c ) {
            int p = c;
            for ( int i = c + 1; i < 2; ++i )
                if ( std::fabs( A[ i ][ c ] ) > std::fabs( A[ p ][ c ] ) ) p = i;
            if ( ! ( std::fabs( A[ p ][ c ] ) > 0 ) ) { ok = false; break; }
            if ( p != c )
                for ( int j = c; j <= 2; ++j ) std::swap( A[ c ][ j ], A[ p ][ j ] );
            for ( int i = c + 1; i < 2; ++i ) {
                const double f = A[ i ][ c ] / A[ c ][ c ];
                for ( int j = c; j <= 2; ++j ) A[ i ][ j ] -= f * A[ c ][ j ];
            }
        }
        if ( ok ) {
            double fit[ 2 ];
            for ( int i = 1; i >= 0; --i ) {
                double s = A[ i ][ 2 ];
                for ( int j = i + 1; j < 2; ++j ) s -= A[ i ][ j ] * fit[ j ];
                fit[ i ] = s / A[ i ][ i ];
            }
            double rmin = 0, rmax = 0;
            for ( I64 k = b; k < e; ++k ) {
                const double r = w[ k ] - fit[ 0 ] * px[ k ] - fit[ 1 ] * py[ k ];
                if ( k == b ) { rmin = r; rmax = r; }
                else { rmin = std::min( rmin, r ); rmax = std::max( rmax, r ); }
            }
            const double u = 1.0 - 2.0 / double( m - 1 );
            const double by_chance = std::sqrt( u > 0 ? u : 0.0 );
            bool well_behaved = true;
            for ( int d = 0; d < 2; ++d ) {
                const double reach = std::max( std::fabs( plo[ d ] ), std::fabs( phi[ d ] ) );
                if ( std::fabs( fit[ d ] ) * ( phi[ d ] - plo[ d ] ) > 8 * spread || std::fabs( fit[ d ] ) * reach > 100 * spread )
                    well_behaved = false;
            }
            if ( well_behaved && rmax - rmin < 0.85 * by_chance * spread ) { a[ 0 ] = fit[ 0 ]; a[ 1 ] = fit[ 1 ]; }
        }
    }
    double bb = 0, amax = 0;
    for ( I64 k = b; k < e; ++k ) {
        const double ay = a[ 0 ] * px[ k ] + a[ 1 ] * py[ k ];
        const double v = w[ k ] - ay;
        bb = k == b ? v : std::max( bb, v );
        amax = std::max( amax, std::fabs( ay ) );
    }
    wa[ 0 ] = a[ 0 ];
    wa[ 1 ] = a[ 1 ];
    wb = bb + 1e-6 * ( std::fabs( bb ) + spread + amax );
}

/// `AaBsp.py:514-532`: where the node of heap rank `i` lands in preorder
static std::vector<I64> preorder_of_heap( int depth ) {
    const I64 nb = ( I64( 1 ) << depth ) - 1;
    std::vector<I64> pre( sz( nb ), 0 );
    for ( int level = 0; level + 1 < depth; ++level ) {
        const int h = depth - level;
        for ( I64 idx = ( I64( 1 ) << level ) - 1; idx < ( I64( 1 ) << ( level + 1 ) ) - 1; ++idx ) {
            pre[ size_t( 2 * idx + 1 ) ] = pre[ size_t( idx ) ] + 1;
            pre[ size_t( 2 * idx + 2 ) ] = pre[ size_t( idx ) ] + ( I64( 1 ) << ( h - 1 ) );
        }
    }
    return pre;
}

/// the tree on `pos[ n ][ 2 ]`, level by level ( `AaBsp.py:545-680`, `bsp_build_level.h:230-318` )
void build_tree( Tree &t, I64 n, const double *pos, I64 leaf ) {
    const int depth = max_depth_for( n, leaf );
    const I64 nb_nodes = ( I64( 1 ) << depth ) - 1;
    t.n = n;
    t.depth = depth;
    t.nb_nodes = nb_nodes;
    std::vector<double> X( sz( n ) ), Y( sz( n ) ), X2( sz( n ) ), Y2( sz( n ) );
    std::vector<I64> ord( sz( n ) ), ord2( sz( n ) ), perm( sz( n ) );
    for ( I64 i = 0; i < n; ++i ) { X[ size_t( i ) ] = pos[ 2 * i ]; Y[ size_t( i ) ] = pos[ 2 * i + 1 ]; ord[ size_t( i ) ] = i; }
    std::vector<I64> hbeg( sz( nb_nodes ) ), hend( sz( nb_nodes ) );
    std::vector<double> hbox( sz( nb_nodes ) * 4, 0.0 );
    std::vector<I64> beg{ 0 }, end{ n };
    for ( int level = 0; level < depth; ++level ) {
        const I64 first = ( I64( 1 ) << level ) - 1;
        std::vector<I64> mid( end );
        for ( size_t j = 0; j < beg.size(); ++j ) {
            const I64 b = beg[ j ], e = end[ j ], node = first + I64( j );
            hbeg[ size_t( node ) ] = b;
            hend[ size_t( node ) ] = e;
            if ( e <= b ) continue;
            double lo[ 2 ] = { X[ size_t( b ) ], Y[ size_t( b ) ] }, hi[ 2 ] = { lo[ 0 ], lo[ 1 ] };
            for ( I64 k = b + 1; k < e; ++k ) {
                lo[ 0 ] = std::min( lo[ 0 ], X[ size_t( k ) ] ); hi[ 0 ] = std::max( hi[ 0 ], X[ size_t( k ) ] );
                lo[ 1 ] = std::min( lo[ 1 ], Y[ size_t( k ) ] ); hi[ 1 ] = std::max( hi[ 1 ], Y[ size_t( k ) ] );
            }
            double *bx = &hbox[ size_t( node ) * 4 ];
            bx[ 0 ] = lo[ 0 ]; bx[ 1 ] = lo[ 1 ]; bx[ 2 ] = hi[ 0 ]; bx[ 3 ] = hi[ 1 ];
            const int ax = ( hi[ 1 ] - lo[ 1 ] > hi[ 0 ] - lo[ 0 ] ) ? 1 : 0;
            for ( I64 k = b; k < e; ++k ) perm[ size_t( k ) ] = k;
            I64 m = e;
            if ( e - b > leaf && hi[ ax ] > lo[ ax ] ) {
                m = b + ( e - b ) / 2;
                select_median( perm, ax ? Y : X, b, e, m );
            }
            mid[ j ] = m;
            for ( I64 k = b; k < e; ++k ) {
                const I64 s = perm[ size_t( k ) ];
                X2[ size_t( k ) ] = X[ size_t( s ) ];
                Y2[ size_t( k ) ] = Y[ size_t( s ) ];
                ord2[ size_t( k ) ] = ord[ size_t( s ) ];
            }
        }
        X.swap( X2 ); Y.swap( Y2 ); ord.swap( ord2 );
        if ( level + 1 < depth ) {
            std::vector<I64> nb_( 2 * beg.size() ), ne_( 2 * beg.size() );
            for ( size_t j = 0; j < beg.size(); ++j ) {
                nb_[ 2 * j ] = beg[ j ]; ne_[ 2 * j ] = mid[ j ];
                nb_[ 2 * j + 1 ] = mid[ j ]; ne_[ 2 * j + 1 ] = end[ j ];
            }
            beg.swap( nb_ ); end.swap( ne_ );
        }
    }
    // the seeds of each leaf in ascending user index: the in-leaf order of a quickselect is arbitrary
    // ( `AaBsp.py:606-609` ), a canonical one lets two builds of the tree cut cells in the same order
    {
        const I64 first = ( I64( 1 ) << ( depth - 1 ) ) - 1;
        std::vector<I64> idx;
        for ( I64 h = first; h < nb_nodes; ++h ) {
            const I64 b = hbeg[ size_t( h ) ], e = hend[ size_t( h ) ];
            if ( e - b < 2 ) continue;
            idx.resize( size_t( e - b ) );
            for ( I64 k = b; k < e; ++k ) idx[ size_t( k - b ) ] = k;
            std::sort( idx.begin(), idx.end(), [&]( I64 u, I64 v ) { return ord[ size_t( u ) ] < ord[ size_t( v ) ]; } );
            std::vector<double> tx( idx.size() ), ty( idx.size() );
            std::vector<I64> to( idx.size() );
            for ( size_t q = 0; q < idx.size(); ++q ) { tx[ q ] = X[ size_t( idx[ q ] ) ]; ty[ q ] = Y[ size_t( idx[ q ] ) ]; to[ q ] = ord[ size_t( idx[ q ] ) ]; }
            for ( size_t q = 0; q < idx.size(); ++q ) { X[ size_t( b ) + q ] = tx[ q ]; Y[ size_t( b ) + q ] = ty[ q ]; ord[ size_t( b ) + q ] = to[ q ]; }
        }
    }
    const std::vector<I64> pre = preorder_of_heap( depth );
    t.beg.assign( size_t( nb_nodes ), 0 );
    t.end.assign( size_t( nb_nodes ), 0 );
    t.box.assign( size_t( nb_nodes ) * 4, 0.0 );
    for ( I64 h = 0; h < nb_nodes; ++h ) {
        const I64 p = pre[ size_t( h ) ];
        t.beg[ size_t( p ) ] = hbeg[ size_t( h ) ];
        t.end[ size_t( p ) ] = hend[ size_t( h ) ];
        for ( int c = 0; c < 4; ++c ) t.box[ size_t( p ) * 4 + c ] = hbox[ size_t( h ) * 4 + c ];
    }
    t.order = ord;
    t.rank_of.assign( size_t( n ), 0 );
    for ( I64 k = 0; k < n; ++k ) t.rank_of[ size_t( ord[ size_t( k ) ] ) ] = k;
    t.px = X;
    t.py = Y;
    t.w.assign( size_t( n ), 0.0 );
    t.wa.assign( size_t( nb_nodes ) * 2, 0.0 );
    t.wb.assign( size_t( nb_nodes ), 0.0 );
}

// ---- a tiny thread pool: blocks of 256 cells dealt in turn ( `diagram/Ops.h:234-250` ) --------

static int g_threads = 0;

int nb_threads() {
    if ( g_threads > 0 ) return g_threads;
    const unsigned h = std::thread::hardware_concurrency();
    return h ? int( h ) : 1;
}

void parallel_blocks( I64 n, I64 block, const std::function<void( int, I64, I64 )> &body ) {
    const int nt = int( std::max<I64>( 1, std::min<I64>( nb_threads(), ( n + block - 1 ) / block ) ) );
    auto run = [&]( int th ) {
        for ( I64 b = I64( th ) * block; b < n; b += I64( nt ) * block )
            body( th, b, std::min( n, b + block ) );
    };
    if ( nt == 1 ) { run( 0 ); return; }
    std::vector<std::thread> ts;
    for ( int th = 0; th < nt; ++th ) ts.emplace_back( run, th );
    for ( auto &x : ts ) x.join();
}

/// weights ( user order ) into the tree, majorants refreshed node by node ( `sdotplan/Sweep.h:216` )
void set_weights( Tree &t, const double *w_user ) {
    for ( I64 k = 0; k < t.n; ++k ) t.w[ size_t( k ) ] = w_user[ t.order[ size_t( k ) ] ];
    parallel_blocks( t.nb_nodes, 64, [&]( int, I64 b, I64 e ) {
        for ( I64 m = b; m < e; ++m ) {
            const I64 sb = t.beg[ size_t( m ) ], se = t.end[ size_t( m ) ];
            if ( se <= sb ) { t.wa[ size_t( 2 * m ) ] = t.wa[ size_t( 2 * m + 1 ) ] = 0; t.wb[ size_t( m ) ] = 0; continue; }
            weight_majorant( t.px.data(), t.py.data(), t.w.data(), sb, se, &t.wa[ size_t( 2 * m ) ], t.wb[ size_t( m ) ] );
        }
    } );
}

// ---- the polygon ------------------------------------------------------------------------------

enum CutStatus { UNCHANGED = 0, CUT = 1, EMPTY = 2 };

/// cut ids: >= 0 a seed rank, `-1 - f` side `f` of the domain box ( 0 bottom, 1 right, 2 top, 3 left,
/// as `gpu/Newton2D.cuh:529-534` reads them ), `PIECE_ID` a pixel plane
constexpr int PIECE_ID  = -16;

/// events of the cut loop ( side tests, pruning tests, new vertices ), for the roofline's count
struct Events {
    I64 side_tests = 0, prune_vertex_tests = 0, new_vertices = 0, node_visits = 0, cuts = 0;
    void add( const Events &o ) {
        side_tests += o.side_tests; prune_vertex_tests += o.prune_vertex_tests; new_vertices += o.new_vertices;
        node_visits += o.node_visits; cuts += o.cuts;
    }
};

/// a convex polygon in the seed's frame, vertex `i` being the corner of cuts `i - 1` and `i`;
/// cut `i` carries the edge `[ v_i, v_i+1 ]` ( `cell/Local2.h:1-40` )
struct Poly {
    int                 nb = 0;
    std::vector<double> x, y, s;
    std::vector<int>    id;
    void reserve( int c ) { if ( int( x.size() ) < c ) { x.resize( size_t( c ) ); y.resize( size_t( c ) ); s.resize( size_t( c ) ); id.resize( size_t( c ) ); } }
};

/// the box `[ lo, hi ]` seen from `p0`, counter-clockwise; or the polygon domain `dom` when there is one
static void init_box( Poly &c, const double *lo, const double *hi, double p0x, double p0y, const Domain *dom = nullptr ) {
    if ( dom && ! dom->empty() ) {
        const int nb = int( dom->x.size() );
        c.reserve( std::max( 16, 2 * nb ) );
        c.nb = nb;
        for ( int i = 0; i < nb; ++i ) { c.x[ size_t( i ) ] = dom->x[ size_t( i ) ] - p0x; c.y[ size_t( i ) ] = dom->y[ size_t( i ) ] - p0y; c.id[ size_t( i ) ] = -5 - i; }
        return;
    }
    c.reserve( 16 );
    c.nb = 4;
    c.x[ 0 ] = lo[ 0 ] - p0x; c.y[ 0 ] = lo[ 1 ] - p0y;
    c.x[ 1 ] = hi[ 0 ] - p0x; c.y[ 1 ] = lo[ 1 ] - p0y;
    c.x[ 2 ] = hi[ 0 ] - p0x; c.y[ 2 ] = hi[ 1 ] - p0y;
    c.x[ 3 ] = lo[ 0 ] - p0x; c.y[ 3 ] = hi[ 1 ] - p0y;
    for ( int i = 0; i < 4; ++i ) c.id[ size_t( i ) ] = -1 - i;
}

/// the outside run holding the largest `s` when the signs come in several runs
/// ( concurrent bisectors; `cell/Local2.h:300-315` )
static void main_run( const Poly &c, int &i1, int &nb_out ) {
    const int nb = c.nb;
    int a = 0;
    for ( int i = 1; i < nb; ++i ) a = c.s[ size_t( i ) ] > c.s[ size_t( a ) ] ? i : a;
    i1 = a;
    for ( int q = a ? a - 1 : nb - 1; c.s[ size_t( q ) ] > 0; q = q ? q - 1 : nb - 1 ) i1 = q;
    nb_out = 1;
    for ( int q = a + 1 < nb ? a + 1 : 0; c.s[ size_t( q ) ] > 0; q = q + 1 < nb ? q + 1 : 0 ) ++nb_out;
    nb_out += ( a - i1 + nb ) % nb;
}

/// the cut by `dx . v <= off`, in place ( `cell/Local2.h:206-298` )
static int cut( Poly &c, double dx, double dy, double off, int pid, Events *ev ) {
    const int nb = c.nb;
    int nb_out = 0;
    for ( int i = 0; i < nb; ++i ) {
        c.s[ size_t( i ) ] = std::fma( dx, c.x[ size_t( i ) ], std::fma( dy, c.y[ size_t( i ) ], -off ) );
        nb_out += c.s[ size_t( i ) ] > 0;
    }
    if ( ev ) ev->side_tests += nb;
    if ( nb_out == 0 ) return UNCHANGED;
    if ( nb_out == nb ) { c.nb = 0; return EMPTY; }
    int i1 = 0;
    for ( int i = 0; i < nb; ++i ) {
        const int q = i ? i - 1 : nb - 1;
        if ( c.s[ size_t( i ) ] > 0 && ! ( c.s[ size_t( q ) ] > 0 ) ) { i1 = i; break; }
    }
    for ( int o = 1, i = i1 + 1; o < nb_out; ++o, ++i ) {
        if ( i == nb ) i = 0;
        if ( ! ( c.s[ size_t( i ) ] > 0 ) ) { main_run( c, i1, nb_out ); break; }
    }
    const int nb_in = nb - nb_out, new_nb = nb_in + 2;
    c.reserve( new_nb + 1 );
    const int j0 = ( i1 + nb - 1 ) % nb, j2 = ( i1 + nb_out - 1 ) % nb, j3 = ( j2 + 1 ) % nb;
    const double s0 = c.s[ size_t( j0 ) ], s1 = c.s[ size_t( i1 ) ], s2 = c.s[ size_t( j2 ) ], s3 = c.s[ size_t( j3 ) ];
    const double ta = s0 / ( s0 - s1 );
    const double pax = std::fma( c.x[ size_t( i1 ) ] - c.x[ size_t( j0 ) ], ta, c.x[ size_t( j0 ) ] );
    const double pay = std::fma( c.y[ size_t( i1 ) ] - c.y[ size_t( j0 ) ], ta, c.y[ size_t( j0 ) ] );
    const double tb = s3 / ( s3 - s2 );
    const double pbx = std::fma( c.x[ size_t( j2 ) ] - c.x[ size_t( j3 ) ], tb, c.x[ size_t( j3 ) ] );
    const double pby = std::fma( c.y[ size_t( j2 ) ] - c.y[ size_t( j3 ) ], tb, c.y[ size_t( j3 ) ] );
    const int bid = c.id[ size_t( j2 ) ];
    auto move = [&]( int d, int s ) { c.x[ size_t( d ) ] = c.x[ size_t( s ) ]; c.y[ size_t( d ) ] = c.y[ size_t( s ) ]; c.id[ size_t( d ) ] = c.id[ size_t( s ) ]; };
    auto put = [&]( int d, double x, double y, int id ) { c.x[ size_t( d ) ] = x; c.y[ size_t( d ) ] = y; c.id[ size_t( d ) ] = id; };
    if ( i1 <= j2 ) {
        if ( nb_out == 1 ) {
            for ( int i = nb; i > i1 + 1; --i ) move( i, i - 1 );
        } else if ( nb_out > 2 ) {
            const int gap = nb_out - 2;
            for ( int i = j2 + 1; i < nb; ++i ) move( i - gap, i );
        }
        put( i1, pax, pay, pid );
        put( i1 + 1, pbx, pby, bid );
    } else {
        if ( j3 >= 2 ) for ( int o = 0; o < nb_in; ++o ) move( 2 + o, j3 + o );
        else           for ( int o = nb_in - 1; o >= 0; --o ) move( 2 + o, j3 + o );
        put( 0, pax, pay, pid );
        put( 1, pbx, pby, bid );
    }
    c.nb = new_nb;
    if ( ev ) { ev->new_vertices += 2; ev->cuts += 1; }
    return CUT;
}

/// the shoelace area ( `cell/Local2.h:430-440` )
static double area( const Poly &c ) {
    double s = 0;
    for ( int i = 0; i < c.nb; ++i ) {
        const int j = i + 1 < c.nb ? i + 1 : 0;
        s += c.x[ size_t( i ) ] * c.y[ size_t( j ) ] - c.x[ size_t( j ) ] * c.y[ size_t( i ) ];
    }
    return s / 2;
}

// ---- the walk --------------------------------------------------------------------------------

/// can a seed of node `m` still cut the cell of `( p0, w0 )`? ( `cell/Pruning.h:46-76` )
static bool can_cut( const Tree &t, I64 m, const Poly &c, double p0x, double p0y, double w0, Events *ev ) {
    const double *bx = &t.box[ size_t( m ) * 4 ];
    const double ax = t.wa[ size_t( 2 * m ) ], ay = t.wa[ size_t( 2 * m + 1 ) ], b = t.wb[ size_t( m ) ];
    if ( ev ) ev->node_visits += 1;
    for ( int i = 0; i < c.nb; ++i ) {
        if ( ev ) ev->prune_vertex_tests += 1;
        const double X = c.x[ size_t( i ) ] + p0x, Y = c.y[ size_t( i ) ] + p0y;
        double yx = X + ax / 2, yy = Y + ay / 2;
        yx = yx < bx[ 0 ] ? bx[ 0 ] : ( yx > bx[ 2 ] ? bx[ 2 ] : yx );
        yy = yy < bx[ 1 ] ? bx[ 1 ] : ( yy > bx[ 3 ] ? bx[ 3 ] : yy );
        const double ux = yx - X, uy = yy - Y;
        const double s = w0 - b + ux * ux - c.x[ size_t( i ) ] * c.x[ size_t( i ) ] - ax * yx
                                + uy * uy - c.y[ size_t( i ) ] * c.y[ size_t( i ) ] - ay * yy;
        if ( s <= 0 ) return true;
    }
    return false;
}

/// squared distance from `p0` to the box of `m`: the ordering key of the walk ( `cell/Providers.h:148` )
static double proximity( const Tree &t, I64 m, double p0x, double p0y ) {
    const double *bx = &t.box[ size_t( m ) * 4 ];
    double e0 = std::max( bx[ 0 ] - p0x, p0x - bx[ 2 ] ), e1 = std::max( bx[ 1 ] - p0y, p0y - bx[ 3 ] );
    e0 = e0 > 0 ? e0 : 0;
    e1 = e1 > 0 ? e1 : 0;
    return e0 * e0 + e1 * e1;
}

/// the cell of rank `k0` in `c`, from the domain box, cut by the planes the tree proposes
/// ( `diagram/Ops.h:82`, `cell/Providers.h:167-236` ). Returns false if the cell is empty.
bool make_cell( const Tree &t, I64 k0, const double *lo, const double *hi, Poly &c, Events *ev, const Domain *dom = nullptr ) {
    const double p0x = t.px[ size_t( k0 ) ], p0y = t.py[ size_t( k0 ) ], w0 = t.w[ size_t( k0 ) ];
    init_box( c, lo, hi, p0x, p0y, dom );
    I64 stack[ 128 ];
    int top = 0;
    stack[ top++ ] = ( I64( 0 ) << 7 ) | t.depth;
    while ( top ) {
        const I64 en = stack[ --top ];
        const I64 m = en >> 7;
        const int h = int( en & 127 );
        const I64 b = t.beg[ size_t( m ) ], e = t.end[ size_t( m ) ];
        if ( b >= e ) continue;
        if ( ! can_cut( t, m, c, p0x, p0y, w0, ev ) ) continue;
        if ( h <= 1 ) {
            for ( I64 k = b; k < e; ++k ) {
                if ( k == k0 ) continue;
                const double dx = t.px[ size_t( k ) ] - p0x, dy = t.py[ size_t( k ) ] - p0y;
                const double off = ( ( w0 - t.w[ size_t( k ) ] ) + ( dx * dx + dy * dy ) ) / 2;
                if ( cut( c, dx, dy, off, int( k ), ev ) == EMPTY ) return false;
            }
            continue;
        }
        const I64 lc = m + 1, rc = m + ( I64( 1 ) << ( h - 1 ) );
        const int hc = h - 1;
        if ( proximity( t, lc, p0x, p0y ) <= proximity( t, rc, p0x, p0y ) ) {
            stack[ top++ ] = ( rc << 7 ) | hc;
            stack[ top++ ] = ( lc << 7 ) | hc;
        } else {
            stack[ top++ ] = ( lc << 7 ) | hc;
            stack[ top++ ] = ( rc << 7 ) | hc;
        }
    }
    return c.nb > 0;
}

/// the same cell, cut by EVERY other seed in rank order: no tree at all ( `cell/Providers.h:36-64` )
bool make_cell_all( const Tree &t, I64 k0, const double *lo, const double *hi, Poly &c, const Domain *dom = nullptr ) {
    const double p0x = t.px[ size_t( k0 ) ], p0y = t.py[ size_t( k0 ) ], w0 = t.w[ size_t( k0 ) ];
    init_box( c, lo, hi, p0x, p0y, dom );
    for ( I64 k = 0; k < t.n; ++k ) {
        if ( k == k0 ) continue;
        const double dx = t.px[ size_t( k ) ] - p0x, dy = t.py[ size_t( k ) ] - p0y;
        const double off = ( ( w0 - t.w[ size_t( k ) ] ) + ( dx * dx + dy * dy ) ) / 2;
        if ( cut( c, dx, dy, off, int( k ), nullptr ) == EMPTY ) return false;
    }
    return true;
}

// ---- integrals against a density ---------------------------------------------------------------

/// moments of one piece of constant density `rho`, in the seed's frame: `m = int rho`,
/// `mx = int v rho`, `m2 = int |v|^2 rho` by the fan from vertex 0 ( `diagram/Ops.h:187-228` )
static void piece_moments( const Poly &c, double rho, double &m, double *mx, double &m2 ) {
    for ( int i = 1; i + 1 < c.nb; ++i ) {
        const double ax = c.x[ 0 ], ay = c.y[ 0 ];
        const double bx = c.x[ size_t( i ) ], by = c.y[ size_t( i ) ];
        const double cx = c.x[ size_t( i + 1 ) ], cy = c.y[ size_t( i + 1 ) ];
        const double ar = std::fabs( ( bx - ax ) * ( cy - ay ) - ( by - ay ) * ( cx - ax ) ) / 2;
        const double w = rho * ar;
        const double sx = ax + bx + cx, sy = ay + by + cy;
        const double sq = ax * ax + ay * ay + bx * bx + by * by + cx * cx + cy * cy;
        m += w;
        mx[ 0 ] += w * sx / 3;
        mx[ 1 ] += w * sy / 3;
        m2 += w * ( sq + sx * sx + sy * sy ) / 12;
    }
}


// ---- a sum of isotropic Gaussians ( `SumOfGaussians.cxx:210-266`, `gpu/Density2D.cuh:153-280` ) ----------------------
//
// The standard normal measure of the triangle ( 0, P, Q ) is `( 1 / 2 pi ) int ( 1 - exp( -r^2 / 2 ) ) dtheta` over its
// angle, `r` the distance from 0 to the segment along the ray; with `t` the abscissa along PQ and `p` the distance from 0
// to its line, `dtheta = p dt / ( p^2 + t^2 )`. Beyond `|t| = 8` the exponential is under 1e-14: the tails are the exact
// arctangents, the core a composite 8-point Gauss-Legendre on panels of at most half a unit.

static double gauss_wedge( double px, double py, double qx, double qy ) {
    const double dx = qx - px, dy = qy - py;
    const double L = std::sqrt( dx * dx + dy * dy );
    if ( ! ( L > 0 ) ) return 0;
    const double ux = dx / L, uy = dy / L;
    const double p = uy * px - ux * py, ap = std::fabs( p );
    if ( ! ( ap > 0 ) ) return 0;
    const double t0 = ux * px + uy * py, t1 = ux * qx + uy * qy, cut = 8;
    double acc = 0;
    if ( ap >= cut ) acc = std::atan( t1 / ap ) - std::atan( t0 / ap );
    else {
        if ( t0 < -cut ) acc += std::atan( std::min( t1, -cut ) / ap ) - std::atan( t0 / ap );
        if ( t1 > cut ) acc += std::atan( t1 / ap ) - std::atan( std::max( t0, cut ) / ap );
        const double c0 = std::max( t0, -cut ), c1 = std::min( t1, cut );
        if ( c1 > c0 ) {
            static const double gx[ 4 ] = { 0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363 };
            static const double gw[ 4 ] = { 0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763 };
            const int np = int( std::ceil( ( c1 - c0 ) * 2 ) );
            const double h = ( c1 - c0 ) / ( 2 * np );
            for ( int k = 0; k < np; ++k ) {
                const double m = c0 + ( 2 * k + 1 ) * h;
                for ( int q = 0; q < 4; ++q )
                    for ( int sg = -1; sg <= 1; sg += 2 ) {
                        const double t = m + sg * h * gx[ q ];
                        const double r2 = ap * ap + t * t;
                        acc += h * gw[ q ] * ( 1 - std::exp( -r2 / 2 ) ) * ap / r2;
                    }
            }
        }
    }
    return ( p < 0 ? -acc : acc ) / 6.283185307179586476925286766559;
}

/// the mass of the convex polygon `c` ( seed frame, the seed at `( ox, oy )` ): the fan from vertex 0, each triangle's
/// measure the absolute value of its three wedges ( `SumOfGaussians.cxx:282-288`, `:316-343` without the quadrature shortcut )
static double gauss_polygon_mass( const Density &dn, const Poly &c, double ox, double oy ) {
    double m = 0;
    for ( int q = 0; q < dn.ng; ++q ) {
        const double s = dn.sigma( q ), cx = dn.gauss[ 4 * q ] - ox, cy = dn.gauss[ 4 * q + 1 ] - oy;
        auto Y = [&]( int v, double &x, double &y ) { x = ( c.x[ sz( v ) ] - cx ) / s; y = ( c.y[ sz( v ) ] - cy ) / s; };
        double x0, y0;
        Y( 0, x0, y0 );
        double acc = 0;
        for ( int v = 1; v + 1 < c.nb; ++v ) {
            double x1, y1, x2, y2;
            Y( v, x1, y1 );
            Y( v + 1, x2, y2 );
            const double w = gauss_wedge( x0, y0, x1, y1 ) + gauss_wedge( x1, y1, x2, y2 ) + gauss_wedge( x2, y2, x0, y0 );
            acc += std::fabs( w );
        }
        m += dn.gauss[ 4 * q + 3 ] * acc;
    }
    return m;
}

/// the signed mass of a polygon given by its vertices ( a frozen cell: a cell turned inside out is negative )
static double gauss_signed_mass( const Density &dn, const double *vx, const double *vy, int nb, double ox, double oy ) {
    double m = 0;
    for ( int q = 0; q < dn.ng; ++q ) {
        const double s = dn.sigma( q ), cx = dn.gauss[ 4 * q ] - ox, cy = dn.gauss[ 4 * q + 1 ] - oy;
        double W = 0;
        for ( int v = 0; v < nb; ++v ) {
            const int r = v + 1 < nb ? v + 1 : 0;
            W += gauss_wedge( ( vx[ v ] - cx ) / s, ( vy[ v ] - cy ) / s, ( vx[ r ] - cx ) / s, ( vy[ r ] - cy ) / s );
        }
        m += dn.gauss[ 4 * q + 3 ] * W;
    }
    return m;
}

/// `int rho ds` along the segment `a -> b`: per Gaussian `w exp( -d^2 / 2 s^2 ) ( erf( u1 ) - erf( u0 ) ) / ( 2 s sqrt( 2 pi ) )`,
/// `d` the distance from the centre to the line ( `SumOfGaussians.cxx:127-152` )
static double gauss_segment_line( const Density &dn, double ox, double oy, double ax, double ay, double bx, double by ) {
    const double ex = bx - ax, ey = by - ay;
    const double L = std::sqrt( ex * ex + ey * ey );
    if ( ! ( L > 0 ) ) return 0;
    const double ux = ex / L, uy = ey / L;
    double res = 0;
    for ( int q = 0; q < dn.ng; ++q ) {
        const double s = dn.sigma( q ), w = dn.gauss[ 4 * q + 3 ];
        const double px = ax - ( dn.gauss[ 4 * q ] - ox ), py = ay - ( dn.gauss[ 4 * q + 1 ] - oy );
        const double d = -px * uy + py * ux;
        const double t0 = px * ux + py * uy, t1 = t0 + L;
        const double qq = d * d / ( 2 * s * s );
        if ( qq > 700 ) continue;
        const double is2 = 0.70710678118654752440 / s;
        res += w * std::exp( -qq ) * ( std::erf( t1 * is2 ) - std::erf( t0 * is2 ) ) / ( 2 * s * 2.50662827463100050242 );
    }
    return res;
}

/// the moments of the convex polygon `c` about the seed: `m`, `int ( x - p ) rho`, `int |x - p|^2 rho`. Per Gaussian,
/// `int ( x - c ) phi = -s^2 circ phi n ds` and `int |x - c|^2 phi = 2 s^2 m - s^2 circ ( ( x - c ) . n ) phi ds`
/// ( divergence theorem on `grad phi = -( x - c ) phi / s^2` ), each edge's `int phi ds` in closed form
static void gauss_polygon_moments( const Density &dn, const Poly &c, double ox, double oy, double &m, double *mx, double &m2 ) {
    for ( int q = 0; q < dn.ng; ++q ) {
        Density one = dn;
        one.ng = 1;
        one.gauss = dn.gauss + 4 * q;
        const double s = dn.sigma( q ), s2 = s * s;
        const double cx = dn.gauss[ 4 * q ] - ox, cy = dn.gauss[ 4 * q + 1 ] - oy;   // the centre in the seed frame
        const double mq = gauss_polygon_mass( one, c, ox, oy );
        double fx = 0, fy = 0, fr = 0;                  // circ phi n ds, circ ( ( x - c ) . n ) phi ds
        for ( int v = 0; v < c.nb; ++v ) {
            const int r = v + 1 < c.nb ? v + 1 : 0;
            const double ex = c.x[ sz( r ) ] - c.x[ sz( v ) ], ey = c.y[ sz( r ) ] - c.y[ sz( v ) ];
            const double L = std::sqrt( ex * ex + ey * ey );
            if ( ! ( L > 0 ) ) continue;
            const double nx = ey / L, ny = -ex / L;     // outward: the polygon is counter-clockwise
            const double J = gauss_segment_line( one, ox, oy, c.x[ sz( v ) ], c.y[ sz( v ) ], c.x[ sz( r ) ], c.y[ sz( r ) ] );
            fx += nx * J;
            fy += ny * J;
            fr += ( ( c.x[ sz( v ) ] - cx ) * nx + ( c.y[ sz( v ) ] - cy ) * ny ) * J;
        }
        // about the seed: int ( x - p ) = int ( x - c ) + ( c - p ) m, int |x - p|^2 = int |x - c|^2 + 2 ( c - p ) . int ( x - c ) + |c - p|^2 m
        const double ix = -s2 * fx, iy = -s2 * fy, i2 = 2 * s2 * mq - s2 * fr;
        m += mq;
        mx[ 0 ] += ix + cx * mq;
        mx[ 1 ] += iy + cy * mq;
        m2 += i2 + 2 * ( cx * ix + cy * iy ) + ( cx * cx + cy * cy ) * mq;
    }
}

/// `func( piece, value )` for each piece of the cell on which the density is constant
/// ( `Image.cxx:106-232`: the tiles the cell's box meets, each cut by the tile's planes when the
/// cell crosses them ); `p0` is the seed, the cell is in its frame
template<class F>
static void for_each_piece( const Density &dn, const Poly &c, double p0x, double p0y, Poly &piece, F &&func ) {
    if ( c.nb == 0 ) return;
    if ( dn.kind == 0 ) { func( c, dn.value ); return; }
    const double org[ 2 ] = { dn.x0 - p0x, dn.y0 - p0y }, h[ 2 ] = { dn.hx, dn.hy };
    const int nbk[ 2 ] = { dn.nx, dn.ny };
    I64 k0[ 2 ], k1[ 2 ];
    double tlo[ 2 ], thi[ 2 ];
    for ( int a = 0; a < 2; ++a ) {
        double mn = 0, mx = 0;
        for ( int v = 0; v < c.nb; ++v ) {
            const double t = ( ( a ? c.y[ size_t( v ) ] : c.x[ size_t( v ) ] ) - org[ a ] ) / h[ a ];
            if ( v == 0 || t < mn ) mn = t;
            if ( v == 0 || t > mx ) mx = t;
        }
        if ( mx < 0 || mn > nbk[ a ] ) return;
        tlo[ a ] = mn;
        thi[ a ] = mx;
        auto knot_index = [&]( double t ) {
            I64 b = 0, e = nbk[ a ];
            while ( e - b > 1 ) { const I64 m = ( b + e ) / 2; if ( double( m ) <= t ) b = m; else e = m; }
            return b;
        };
        k0[ a ] = knot_index( mn );
        k1[ a ] = knot_index( mx ) + 1;
    }
    for ( I64 j = k0[ 1 ]; j < k1[ 1 ]; ++j ) {
        for ( I64 i = k0[ 0 ]; i < k1[ 0 ]; ++i ) {
            const I64 k[ 2 ] = { i, j };
            bool alive = true, started = false;
            for ( int a = 0; a < 2 && alive; ++a ) {
                const double tile_lo = double( k[ a ] ), tile_hi = double( k[ a ] + 1 );
                const double lo = org[ a ] + tile_lo * h[ a ], hi = org[ a ] + tile_hi * h[ a ];
                const double dx = a == 0, dy = a == 1;
                if ( thi[ a ] > tile_hi ) {
                    if ( ! started ) { piece.reserve( c.nb + 8 ); piece.nb = c.nb; for ( int q = 0; q < c.nb; ++q ) { piece.x[ size_t( q ) ] = c.x[ size_t( q ) ]; piece.y[ size_t( q ) ] = c.y[ size_t( q ) ]; piece.id[ size_t( q ) ] = c.id[ size_t( q ) ]; } }
                    started = true;
                    if ( cut( piece, dx, dy, hi, PIECE_ID, nullptr ) == EMPTY ) alive = false;
                }
                if ( alive && tlo[ a ] < tile_lo ) {
                    if ( ! started ) { piece.reserve( c.nb + 8 ); piece.nb = c.nb; for ( int q = 0; q < c.nb; ++q ) { piece.x[ size_t( q ) ] = c.x[ size_t( q ) ]; piece.y[ size_t( q ) ] = c.y[ size_t( q ) ]; piece.id[ size_t( q ) ] = c.id[ size_t( q ) ]; } }
                    started = true;
                    if ( cut( piece, -dx, -dy, -lo, PIECE_ID, nullptr ) == EMPTY ) alive = false;
                }
            }
            if ( ! alive ) continue;
            const double value = dn.img[ size_t( j ) * size_t( dn.nx ) + size_t( i ) ];
            func( started ? piece : c, value );
        }
    }
}

/// one cell's mass and its facets `c_kj = int_facet rho / ( 2 |p_k - p_j| )` ( `sdotplan/Sweep.h:64-128` );
/// `facet( j_rank, c )` per neighbour, in the cell's edge order
template<class F>
static double mass_and_facets( const Tree &t, const Density &dn, I64 k, const Poly &c, Poly &piece, bool with_facets, F &&facet ) {
    const double p0x = t.px[ size_t( k ) ], p0y = t.py[ size_t( k ) ];
    if ( dn.kind == 2 ) {                                // the whole cell at once: wedges and erf lines
        if ( with_facets )
            for ( int u = 0; u < c.nb; ++u ) {
                const int id = c.id[ size_t( u ) ];
                if ( id < 0 ) continue;
                const int r = u + 1 < c.nb ? u + 1 : 0;
                const double line = gauss_segment_line( dn, p0x, p0y, c.x[ sz( u ) ], c.y[ sz( u ) ], c.x[ sz( r ) ], c.y[ sz( r ) ] );
                const double dx = t.px[ size_t( id ) ] - p0x, dy = t.py[ size_t( id ) ] - p0y;
                const double d2 = dx * dx + dy * dy;
                if ( d2 > 0 ) facet( I64( id ), line / ( 2 * std::sqrt( d2 ) ) );
            }
        return c.nb >= 3 ? gauss_polygon_mass( dn, c, p0x, p0y ) : 0.0;
    }
    double mass = 0;
    std::vector<double> vals( sz( c.nb ), 0.0 );
    for_each_piece( dn, c, p0x, p0y, piece, [&]( const Poly &pc, double value ) {
        mass += value * area( pc );
        if ( ! with_facets ) return;
        for ( int q = 0; q < pc.nb; ++q ) {
            const int id = pc.id[ size_t( q ) ];
            if ( id < 0 ) continue;
            const int r = q + 1 < pc.nb ? q + 1 : 0;
            const double dx = pc.x[ size_t( r ) ] - pc.x[ size_t( q ) ], dy = pc.y[ size_t( r ) ] - pc.y[ size_t( q ) ];
            const double val = value * std::sqrt( dx * dx + dy * dy );
            for ( int u = 0; u < c.nb; ++u )
                if ( c.id[ size_t( u ) ] == id ) { vals[ size_t( u ) ] += val; break; }
        }
    } );
    if ( with_facets ) {
        for ( int u = 0; u < c.nb; ++u ) {
            const int id = c.id[ size_t( u ) ];
            if ( id < 0 ) continue;
            const double dx = t.px[ size_t( id ) ] - p0x, dy = t.py[ size_t( id ) ] - p0y;
            const double d2 = dx * dx + dy * dy;
            if ( d2 > 0 ) facet( I64( id ), vals[ size_t( u ) ] / ( 2 * std::sqrt( d2 ) ) );
        }
    }
    return mass;
}

// ---- the Laplacian and the CG -----------------------------------------------------------------

struct Facet { I64 i, j; double c; };

/// `sdotplan/Laplacian.h:40-72`: each row keeps its own view of a facet; `L_ii = sum_j c_ij`
struct Laplacian {
    I64 n = 0;
    std::vector<I64> row, col;
    std::vector<double> c, dia;
    void assemble( I64 nb, const std::vector<Facet> &fa ) {
        n = nb;
        row.assign( size_t( n + 1 ), 0 );
        for ( const Facet &e : fa ) ++row[ size_t( e.i + 1 ) ];
        for ( I64 i = 0; i < n; ++i ) row[ size_t( i + 1 ) ] += row[ size_t( i ) ];
        col.resize( size_t( row[ size_t( n ) ] ) );
        c.resize( size_t( row[ size_t( n ) ] ) );
        dia.assign( size_t( n ), 0.0 );
        std::vector<I64> at( row.begin(), row.end() - 1 );
        for ( const Facet &e : fa ) {
            const I64 p = at[ size_t( e.i ) ]++;
            col[ size_t( p ) ] = e.j;
            c[ size_t( p ) ] = e.c;
            dia[ size_t( e.i ) ] += e.c;
        }
        for ( I64 i = 0; i < n; ++i ) if ( ! ( dia[ size_t( i ) ] > 0 ) ) dia[ size_t( i ) ] = 1;
    }
    void product( const double *x, double *y ) const {
        for ( I64 i = 0; i < n; ++i ) {
            double s = dia[ size_t( i ) ] * x[ i ];
            for ( I64 e = row[ size_t( i ) ]; e < row[ size_t( i + 1 ) ]; ++e ) s -= c[ size_t( e ) ] * x[ col[ size_t( e ) ] ];
            y[ i ] = s;
        }
    }
};

struct LinStats { int nb_iter = 0; double worst = 0, t = 0; };

/// Jacobi-preconditioned CG, the gauge `d_0 = 0` ( `sdotplan/Linear.cpp:44-88` )
bool cg_solve( const Laplacian &L, const std::vector<double> &b, std::vector<double> &d, double tol, int maxit, LinStats &st ) {
    const double t0 = now();
    const I64 n = L.n;
    std::vector<double> r( sz( n ) ), z( sz( n ) ), p( sz( n ) ), q( sz( n ) );
    d.assign( size_t( n ), 0.0 );
    double nb = 0;
    for ( I64 i = 1; i < n; ++i ) { r[ size_t( i ) ] = b[ size_t( i ) ]; nb += b[ size_t( i ) ] * b[ size_t( i ) ]; }
    r[ 0 ] = 0;
    nb = std::sqrt( nb );
    if ( ! ( nb > 0 ) ) { st.t += now() - t0; return true; }
    double rz = 0;
    for ( I64 i = 1; i < n; ++i ) { z[ size_t( i ) ] = r[ size_t( i ) ] / L.dia[ size_t( i ) ]; p[ size_t( i ) ] = z[ size_t( i ) ]; rz += r[ size_t( i ) ] * z[ size_t( i ) ]; }
    double err = 1;
    int it = 0;
    for ( ; it < maxit; ++it ) {
        L.product( p.data(), q.data() );
        q[ 0 ] = 0;
        double pq = 0;
        for ( I64 i = 1; i < n; ++i ) pq += p[ size_t( i ) ] * q[ size_t( i ) ];
        if ( ! ( pq > 0 ) ) break;
        const double alpha = rz / pq;
        double nr = 0;
        for ( I64 i = 1; i < n; ++i ) { d[ size_t( i ) ] += alpha * p[ size_t( i ) ]; r[ size_t( i ) ] -= alpha * q[ size_t( i ) ]; nr += r[ size_t( i ) ] * r[ size_t( i ) ]; }
        err = std::sqrt( nr ) / nb;
        if ( err <= tol ) break;
        double rz2 = 0;
        for ( I64 i = 1; i < n; ++i ) { z[ size_t( i ) ] = r[ size_t( i ) ] / L.dia[ size_t( i ) ]; rz2 += r[ size_t( i ) ] * z[ size_t( i ) ]; }
        const double beta = rz2 / rz;
        rz = rz2;
        for ( I64 i = 1; i < n; ++i ) p[ size_t( i ) ] = z[ size_t( i ) ] + beta * p[ size_t( i ) ];
    }
    st.nb_iter += it;
    st.worst = std::max( st.worst, err );
    st.t += now() - t0;
    return err < 1;
}

#include "oracle_mg.h"

// ---- the diagram as Newton sees it ( `sdotplan/Sweep.h:142-364` ) ------------------------------

/// the cuts of a cell's edges in polygon order ( user index of the neighbour, or `-1 - f` a side of the box ), at most
/// `EDGE_CAP` ( more: the count is -1 ): what the step by the limits rebuilds a cell from ( `gpu/Cell2D.cuh:1088-1096` )
constexpr int EDGE_CAP = 16;
struct Edges {
    std::vector<int> id;     ///< [ EDGE_CAP ][ n ]
    std::vector<int> nb;     ///< [ n ]
};

struct Sweep {
    static constexpr int DIM = 2;
    Tree   &t;
    const double *lo, *hi;
    Density dn;
    int    nb_diag = 0;
    double t_majorant = 0, t_diag = 0;
    Events ev;
    Edges *edges = nullptr;  ///< filled by `measures` when set

    Sweep( Tree &t, const double *lo, const double *hi, const Density &dn ) : t( t ), lo( lo ), hi( hi ), dn( dn ) {}

    void set_weights( const std::vector<double> &W ) {
        const double t0 = now();
        orc::set_weights( t, W.data() );
        t_majorant += now() - t0;
    }

    /// masses ( user order ) and, if asked, the facets ( user indices )
    void measures( std::vector<double> &a, std::vector<Facet> *fa ) {
        const double t0 = now();
        const I64 n = t.n;
        a.assign( size_t( n ), 0.0 );
        const int nt = nb_threads();
        std::vector<std::vector<Facet>> fth( sz( nt ) );
        std::vector<Events> eth( sz( nt ) );
        if ( edges ) { edges->id.assign( sz( n ) * EDGE_CAP, 0 ); edges->nb.assign( sz( n ), 0 ); }
        parallel_blocks( n, 256, [&]( int th, I64 b, I64 e ) {
            Poly c, piece;
            c.reserve( 64 );
            piece.reserve( 64 );
            for ( I64 k = b; k < e; ++k ) {
                const I64 i = t.order[ size_t( k ) ];
                if ( ! make_cell( t, k, lo, hi, c, &eth[ size_t( th ) ], &dn.dom ) ) { a[ size_t( i ) ] = 0; if ( edges ) edges->nb[ size_t( i ) ] = 0; continue; }
                if ( edges ) {
                    int ne = c.nb < 3 ? 0 : c.nb;
                    if ( ne > EDGE_CAP ) ne = -1;
                    else
                        for ( int q = 0; q < ne; ++q ) {
                            const int id = c.id[ size_t( q ) ];
                            edges->id[ size_t( q ) * size_t( n ) + size_t( i ) ] = id >= 0 ? int( t.order[ size_t( id ) ] ) : id;
                        }
                    edges->nb[ size_t( i ) ] = ne;
                }
                a[ size_t( i ) ] = mass_and_facets( t, dn, k, c, piece, fa != nullptr, [&]( I64 j, double cij ) {
                    fth[ size_t( th ) ].push_back( Facet{ i, t.order[ size_t( j ) ], cij } );
                } );
            }
        } );
        if ( fa ) {
            fa->clear();
            for ( auto &v : fth ) fa->insert( fa->end(), v.begin(), v.end() );
            // a deterministic order whatever the thread count: by row, then in the cell's edge order
            std::stable_sort( fa->begin(), fa->end(), []( const Facet &x, const Facet &y ) { return x.i < y.i; } );
        }
        for ( auto &e : eth ) ev.add( e );
        t_diag += now() - t0;
        ++nb_diag;
    }

    /// barycenters ( the seed for an empty cell ) and the transport cost ( `sdotplan/Sweep.h:305-362` )
    void moments( std::vector<double> &bary, double &cost ) {
        const I64 n = t.n;
        bary.assign( size_t( n ) * 2, 0.0 );
        const int nt = nb_threads();
        std::vector<double> cth( sz( nt ), 0.0 );
        parallel_blocks( n, 256, [&]( int th, I64 b, I64 e ) {
            Poly c, piece;
            c.reserve( 64 );
            piece.reserve( 64 );
            for ( I64 k = b; k < e; ++k ) {
                const I64 i = t.order[ size_t( k ) ];
                const double p0x = t.px[ size_t( k ) ], p0y = t.py[ size_t( k ) ];
                double m = 0, mx[ 2 ] = { 0, 0 }, m2 = 0;
                if ( make_cell( t, k, lo, hi, c, nullptr, &dn.dom ) && c.nb >= 3 ) {
                    if ( dn.kind == 2 ) gauss_polygon_moments( dn, c, p0x, p0y, m, mx, m2 );
                    else for_each_piece( dn, c, p0x, p0y, piece, [&]( const Poly &pc, double value ) { piece_moments( pc, value, m, mx, m2 ); } );
                }
                cth[ size_t( th ) ] += m2;
                bary[ size_t( 2 * i ) ] = m > 0 ? p0x + mx[ 0 ] / m : p0x;
                bary[ size_t( 2 * i + 1 ) ] = m > 0 ? p0y + mx[ 1 ] / m : p0y;
            }
        } );
        cost = 0;
        for ( double v : cth ) cost += v;
    }
};


// ---- the width continuation of an image ( `sdotplan/Continuation.h:31-39`, `sdotplan/convolved/Image.h:40-85` ) -------

/// the widths of the stages: `s0, s0 / ratio, ...` while `>= s_min`, then 0
static std::vector<double> continuation_steps( double s0, double ratio, double s_min ) {
    std::vector<double> res;
    if ( s0 > 0 && ratio > 1 )
        for ( double s = s0; s >= s_min * ( 1 - 1e-12 ) && s > 0; s /= ratio ) res.push_back( s );
    res.push_back( 0 );
    return res;
}

/// the image at width `s`: along x, then along y, a Gaussian of standard deviation `s / h` pixels truncated at four of
/// them, each value divided by the sum of the weights that fall inside the image ( so a constant stays that constant )
static void blur_image( const Density &dn, double s, std::vector<double> &out ) {
    const size_t nb = size_t( dn.nx ) * size_t( dn.ny );
    out.assign( dn.img, dn.img + nb );
    if ( ! ( s > 0 ) ) return;
    std::vector<double> tmp( nb );
    for ( int a = 0; a < 2; ++a ) {
        const double sp = s / ( a ? dn.hy : dn.hx );
        if ( ! ( sp > 1e-3 ) ) continue;
        const I64 len = a ? dn.ny : dn.nx, stride = a ? dn.nx : 1;
        const I64 r = std::min<I64>( I64( std::ceil( 4 * sp ) ), len );
        std::vector<double> ker( sz( 2 * r + 1 ) );
        for ( I64 k = -r; k <= r; ++k ) ker[ sz( k + r ) ] = std::exp( -0.5 * double( k * k ) / ( sp * sp ) );
        for ( I64 j = 0; j < dn.ny; ++j )
            for ( I64 i = 0; i < dn.nx; ++i ) {
                const I64 at = j * dn.nx + i, pos = a ? j : i;
                const I64 kb = std::max<I64>( -r, -pos ), ke = std::min<I64>( r, len - 1 - pos );
                double sum = 0, wsum = 0;
                for ( I64 k = kb; k <= ke; ++k ) { sum += ker[ sz( k + r ) ] * out[ sz( at + k * stride ) ]; wsum += ker[ sz( k + r ) ]; }
                tmp[ sz( at ) ] = sum / wsum;
            }
        out.swap( tmp );
    }
}

// ---- the step by the limits ( the card's, `gpu/Newton2D.cuh:505-770`, `:1615-1670` ) ---------------------------------
//
// With its edges frozen, a cell along `w + t d` is known at every `t`: a bisector's offset is affine in the weights and its
// normal fixed, so each vertex ( the crossing of two consecutive edge lines ) is affine in `t` and the area a polynomial of
// degree 2. FORWARD ( the accepted diagram ): the first `t > 0` where a cell's mass reaches the floor `eps`; BACKWARD ( a
// trial diagram, its cells under the floor ): where they crossed it going back. A density that is not a constant bisects
// the frozen cell's mass instead.

constexpr double DET_MIN = 1e-6;

/// a diagram's cells, edges frozen, along `w + t d`: positions, weights, direction in USER order
struct Along {
    const double *pos, *w, *d, *lo, *hi;
    const Edges  *e;
    I64           n;
    const Density *dn;

    /// the line of cut `c` of cell `k` in its seed's frame at `w + t0 d`: `nx x + ny y = off + s delta` at `w + ( t0 + s ) d`
    void line( I64 k, int c, double t0, double &nx, double &ny, double &off, double &delta ) const {
        const double px = pos[ 2 * k ], py = pos[ 2 * k + 1 ];
        if ( c >= 0 ) {
            nx = pos[ 2 * c ] - px;
            ny = pos[ 2 * c + 1 ] - py;
            off = 0.5 * ( nx * nx + ny * ny ) + 0.5 * ( ( w[ k ] + t0 * d[ k ] ) - ( w[ c ] + t0 * d[ c ] ) );
            delta = 0.5 * ( d[ k ] - d[ c ] );
            return;
        }
        if ( c <= -5 ) {                                 // an edge of the polygon domain
            const Domain &D = dn->dom;
            const size_t j = size_t( -5 - c );
            nx = D.nx[ j ];
            ny = D.ny[ j ];
            off = D.off[ j ] - nx * px - ny * py;
            delta = 0;
            return;
        }
        const int f = -1 - c;
        nx = ( f & 1 ) ? 1.0 : 0.0;
        ny = ( f & 1 ) ? 0.0 : 1.0;
        off = f == 0 ? lo[ 1 ] - py : f == 1 ? hi[ 0 ] - px : f == 2 ? hi[ 1 ] - py : lo[ 0 ] - px;
        delta = 0;
    }
    int nb_edges( I64 k ) const { return e->nb[ sz( k ) ]; }
    int edge( I64 k, int q ) const { return e->id[ size_t( q ) * size_t( n ) + size_t( k ) ]; }

    /// vertex `q` of the frozen cell at `w + t0 d` ( crossing of the lines of edges `q - 1` and `q` ) and its velocity;
    /// `false` if two consecutive lines are nearly parallel
    bool vertices( I64 k, double t0, int nb, double *vx, double *vy, double *ux, double *uy ) const {
        double ax, ay, ao, ad;
        line( k, edge( k, nb - 1 ), t0, ax, ay, ao, ad );
        for ( int q = 0; q < nb; ++q ) {
            double bx, by, bo, bd;
            line( k, edge( k, q ), t0, bx, by, bo, bd );
            const double det = ax * by - ay * bx;
            if ( ! ( det * det > DET_MIN * DET_MIN * ( ax * ax + ay * ay ) * ( bx * bx + by * by ) ) ) return false;
            const double inv = 1 / det;
            vx[ q ] = ( ao * by - bo * ay ) * inv;
            vy[ q ] = ( ax * bo - bx * ao ) * inv;
            ux[ q ] = ( ad * by - bd * ay ) * inv;
            uy[ q ] = ( ax * bd - bx * ad ) * inv;
            ax = bx; ay = by; ao = bo; ad = bd;
        }
        return true;
    }

    /// `area( t0 + s ) = a0 + a1 s + a2 s^2` while the edges hold
    bool area_polynomial( I64 k, double t0, double &a0, double &a1, double &a2 ) const {
        const int nb = nb_edges( k );
        if ( nb < 3 ) return false;
        double vx[ EDGE_CAP ], vy[ EDGE_CAP ], ux[ EDGE_CAP ], uy[ EDGE_CAP ];
        if ( ! vertices( k, t0, nb, vx, vy, ux, uy ) ) return false;
        double c0 = 0, c1 = 0, c2 = 0;
        for ( int q = 0; q < nb; ++q ) {
            const int r = q + 1 < nb ? q + 1 : 0;
            c0 += vx[ q ] * vy[ r ] - vx[ r ] * vy[ q ];
            c1 += vx[ q ] * uy[ r ] + ux[ q ] * vy[ r ] - vx[ r ] * uy[ q ] - ux[ r ] * vy[ q ];
            c2 += ux[ q ] * uy[ r ] - ux[ r ] * uy[ q ];
        }
        const double sg = c0 < 0 ? -0.5 : 0.5;
        a0 = sg * c0; a1 = sg * c1; a2 = sg * c2;
        return true;
    }

    /// the mass of the frozen cell at `w + t d` against the density ( counter-clockwise as the cell it comes from: a cell
    /// turned inside out has a negative mass ); `false` if two consecutive lines are nearly parallel
    bool frozen_mass( I64 k, double t, double &m ) const;
};

/// `int F0 dy` along the segment `a -> b` of the seed frame, `F0( x, y ) = int rho( s, y ) ds` from the seed's column to `x`
/// along the row of `y`: exact per pixel piece ( `F0` linear there: its middle times the piece's `dy` ). The pieces come from
/// the sorted crossings of the segment with the grid lines ( `gpu/Density2D.cuh:95` walks them instead ). A frozen cell may
/// leave the domain: its points off the grid take the nearest pixel, as `image_col` / `image_row` of `gpu/Density2D.cuh:82-83`
/// clamp them.
static double image_edge_mass( const Density &dn, double px, double py, double ax, double ay, double bx, double by ) {
    const double ox = dn.x0 - px, oy = dn.y0 - py;
    int iref = int( std::floor( -ox / dn.hx ) );
    iref = std::min( std::max( iref, 0 ), dn.nx - 1 );
    std::vector<double> ts{ 0.0, 1.0 };
    for ( int a = 0; a < 2; ++a ) {
        const double o = a ? oy : ox, h = a ? dn.hy : dn.hx;
        const double ga = ( ( a ? ay : ax ) - o ) / h, gb = ( ( a ? by : bx ) - o ) / h;
        if ( ga == gb ) continue;
        for ( double g = std::ceil( std::min( ga, gb ) ); g <= std::floor( std::max( ga, gb ) ); g += 1 ) {
            const double t = ( g - ga ) / ( gb - ga );
            if ( t > 0 && t < 1 ) ts.push_back( t );
        }
    }
    std::sort( ts.begin(), ts.end() );
    double m = 0;
    for ( size_t q = 0; q + 1 < ts.size(); ++q ) {
        const double t0 = ts[ q ], t1 = ts[ q + 1 ];
        if ( ! ( t1 > t0 ) ) continue;
        const double tm = 0.5 * ( t0 + t1 );
        const double x = ax + tm * ( bx - ax ), y = ay + tm * ( by - ay );
        int i = int( std::floor( ( x - ox ) / dn.hx ) ), j = int( std::floor( ( y - oy ) / dn.hy ) );
        i = std::min( std::max( i, 0 ), dn.nx - 1 );
        j = std::min( std::max( j, 0 ), dn.ny - 1 );
        const double *row = dn.img + size_t( j ) * size_t( dn.nx );
        double f0 = 0;
        if ( i >= iref ) for ( int c = iref; c < i; ++c ) f0 += row[ c ] * dn.hx;
        else             for ( int c = i; c < iref; ++c ) f0 -= row[ c ] * dn.hx;
        f0 += row[ i ] * ( x - ( ox + double( i ) * dn.hx ) );
        m += f0 * ( ( by - ay ) * ( t1 - t0 ) );
    }
    return m;
}

bool Along::frozen_mass( I64 k, double t, double &m ) const {
    const int nb = nb_edges( k );
    double vx[ EDGE_CAP ], vy[ EDGE_CAP ], ux[ EDGE_CAP ], uy[ EDGE_CAP ];
    if ( nb < 3 || ! vertices( k, t, nb, vx, vy, ux, uy ) ) return false;
    m = 0;
    if ( dn->kind == 0 ) {
        for ( int q = 0; q < nb; ++q ) {
            const int r = q + 1 < nb ? q + 1 : 0;
            m += vx[ q ] * vy[ r ] - vx[ r ] * vy[ q ];
        }
        m *= 0.5 * dn->value;
        return true;
    }
    const double px = pos[ 2 * k ], py = pos[ 2 * k + 1 ];
    if ( dn->kind == 2 ) { m = gauss_signed_mass( *dn, vx, vy, nb, px, py ); return true; }
    for ( int q = 0; q < nb; ++q ) {
        const int r = q + 1 < nb ? q + 1 : 0;
        m += image_edge_mass( *dn, px, py, vx[ q ], vy[ q ], vx[ r ], vy[ r ] );
    }
    return true;
}

/// the real roots of `a2 s^2 + a1 s + c = 0`, sorted ( `sdotplan/Bounds.h:170-185` )
static int quadratic_roots( double a2, double a1, double c, double &r1, double &r2 ) {
    if ( std::fabs( a2 ) <= 1e-300 ) {
        if ( a1 == 0 ) return 0;
        r1 = r2 = -c / a1;
        return 1;
    }
    const double disc = a1 * a1 - 4 * a2 * c;
    if ( disc < 0 ) return 0;
    const double sq = std::sqrt( disc );
    const double q = -0.5 * ( a1 + ( a1 >= 0 ? sq : -sq ) );
    double x1 = q / a2, x2 = q != 0 ? c / q : x1;
    if ( x1 > x2 ) std::swap( x1, x2 );
    r1 = x1; r2 = x2;
    return 2;
}

/// the admissible end of `[ lo, hi ]` ( `mass( lo ) >= eps > mass( hi )` ), to 1e-2 relative and above zero
template<class F>
static double mass_bisection( double lo, double hi, double eps, const F &mass ) {
    for ( int round = 0; round < 64; ++round ) {
        if ( hi - lo <= 1e-2 * hi && lo > 0 ) break;
        const double mid = 0.5 * ( lo + hi );
        double m;
        if ( mass( mid, m ) && m >= eps ) lo = mid;
        else                              hi = mid;
    }
    return lo;
}

/// FORWARD, one cell of the accepted diagram: the first `t > 0` where its mass reaches `eps` ( 1e300: never, or no say )
static double alpha_forward( const Along &al, I64 k, double eps ) {
    double a0, a1, a2;
    if ( al.dn->kind == 0 ) {
        if ( ! al.area_polynomial( k, 0.0, a0, a1, a2 ) ) return 1e300;
        const double rho = al.dn->value;
        double r1, r2;
        const int nr = quadratic_roots( rho * a2, rho * a1, std::max( rho * a0 - eps, 0.0 ), r1, r2 );
        if ( nr >= 1 && r1 > 0 ) return r1;
        if ( nr >= 2 && r2 > 0 ) return r2;
        return 1e300;
    }
    // a density: the frozen mass at `t = 1`, and at the bottom of the area's parabola
    if ( al.nb_edges( k ) < 3 ) return 1e300;
    auto mass = [&]( double t, double &m ) { return al.frozen_mass( k, t, m ); };
    double m1, hi = -1;
    if ( ! mass( 1.0, m1 ) ) return 1e300;
    if ( m1 < eps ) hi = 1;
    else if ( al.area_polynomial( k, 0.0, a0, a1, a2 ) && a2 > 0 ) {
        const double ts = -a1 / ( 2 * a2 );
        double ms;
        if ( ts > 0 && ts < 1 && mass( ts, ms ) && ms < eps ) hi = ts;
    }
    if ( hi < 0 ) return 1e300;
    return mass_bisection( 0.0, hi, eps, mass );
}

/// BACKWARD, one cell under the floor in the trial diagram at `t_trial`: where it crossed it ( half the trial if unknown )
static double alpha_backward( const Along &al, I64 k, double eps, double t_trial ) {
    double target = 0.5 * t_trial;
    const int nb = al.nb_edges( k );
    if ( al.dn->kind == 0 ) {
        double a0, a1, a2;
        if ( nb >= 3 && al.area_polynomial( k, t_trial, a0, a1, a2 ) ) {
            const double rho = al.dn->value;
            double r1, r2, beta = -1e300;
            const int nr = quadratic_roots( rho * a2, rho * a1, rho * a0 - eps, r1, r2 );
            if ( nr >= 1 && r1 < 0 ) beta = r1;
            if ( nr >= 2 && r2 < 0 ) beta = r2;
            const double tg = t_trial + beta;
            if ( tg > 0 && tg < t_trial ) target = tg;
        }
        return target;
    }
    auto mass = [&]( double t, double &m ) { return al.frozen_mass( k, t, m ); };
    double m0;
    if ( nb >= 3 && mass( 0.0, m0 ) && m0 >= eps ) {
        const double tg = mass_bisection( 0.0, t_trial, eps, mass );
        if ( tg > 0 && tg < t_trial ) target = tg;
    }
    return target;
}

// ---- Newton ( `sdotplan/Newton.h` ) ------------------------------------------------------------

enum Stat {
    STATUS = 0, RESIDUAL, RESIDUAL0, NB_ITER, NB_DIAG, NB_BACKTRACKS, T_MAJORANT, T_DIAG, T_ASM, T_LIN, T_LIM, EPS,
    DOMAIN_MASS, NB_OVERFLOWED, NB_CELL_LIM, NB_LIMIT_ROUNDS, LIN_NB_HIERARCHIES, LIN_NB_ITER, LIN_WORST, START, T_TOTAL,
    NB_CONTINUATION_STEPS, MIN_START_MASS, IT_SWITCH, IT_DOUBLE, SCRATCH_BYTES,
    NB_CLUSTERS, NB_AGGREGATED, NB_DUPLICATES, RESIDUAL_FULL, NB_POLISH, NB_STATS
};
enum Hist { H_STEP = 0, H_T, H_RESIDUAL_L2, H_MIN_MASS, H_MAX_RESIDUAL, H_NB_DIAG, H_NB_EVALS, H_S, NB_HIST };
enum Residual { LIN = 0, LOG = 1, POWER = 2 };

enum Step { TRIALS = 0, LIMITS = 1 };

#include "oracle3d.h"

struct Options {
    double tol_abs = 1e-8, tol_rel = 0, t_min = 1e-10, mult_ok = 4, power = 0.5, switch_residual = 2, lin_tol = 1e-12;
    int    maxit = 100, max_backtracks = 60, residual = LOG, lin_maxit = 20000;
    int    keep_start = 0;
    int    step = TRIALS;
    double factor = 0.9;     ///< LIMITS: the trial is `factor` times the limit
    int    linear = 0;       ///< the Newton systems: 0 Jacobi-PCG, 1 the multigrid ( `Linear.cpp:208-233` )
};

/// `Report.h:37-57`: two refused trials whose secant slopes agree and extrapolate under `nr / 4`
struct Hopeless {
    double t_prev = 0, s_prev = 0;
    bool has_prev = false;
    bool refused( double t, double n2r, double nr, bool ok ) {
        if ( ! ok || ! ( nr > 0 ) ) { has_prev = false; return false; }
        const double s = ( nr - n2r ) / t;
        const double as = s_prev > 0 ? s_prev : -s_prev;
        const bool hopeless = has_prev && t_prev == 2 * t && s < nr / 4 && s_prev < nr / 4 &&
                              s - s_prev <= 0.25 * as && s_prev - s <= 0.25 * as && 2 * s - s_prev < nr / 4;
        t_prev = t; s_prev = s; has_prev = true;
        return hopeless;
    }
};

static double g_of( double x, int r, double p ) {
    x = std::max( x, 1e-300 );
    if ( r == POWER ) return p == 0 ? std::log( x ) : ( std::pow( x, p ) - 1 ) / p;
    return r == LOG ? std::log( x ) : x - 1;
}
static double gp_of( double x, int r, double p ) {
    x = std::max( x, 1e-300 );
    if ( r == POWER ) return std::pow( x, p - 1 );
    return r == LOG ? 1 / x : 1;
}

/// the Newton solve over a sweep `SW` ( `Sweep`: 2D; `Sweep3`: 3D, where the step is by trials only )
template<class SW>
struct SolverT {
    SW &sw;
    Options o;
    const double *pos = nullptr;                         ///< user order ( the step by the limits reads it )
    std::vector<double> nu, w, a, d;
    std::vector<Facet> fa;
    Edges ea;                                            ///< the accepted diagram's edges
    // THE AGGREGATION ( user indices, `sdotplan/Aggregation.h:65-320` ). Clusters are the union-find classes of more than one
    // seed, numbered by their smallest index, members increasing; the tests and the merit read each member's share of its
    // cluster's mass, `nu_m sum_C a / sum_C nu`.
    // * EXACT DUPLICATES ( given ): the k-th duplicate of a representative has `w_dup = w_rep - k gap` ( its cell empty ), its
    //   target carried by the representative ( `nue` ), its row of the Laplacian linked to the representative's;
    // * NEAR-COINCIDENT PAIRS ( found on the diagrams, `detect` ): merged when one ulp of their weights moves more than
    //   `tau / margin` of mass between them; latched. Once the aggregated problem converged, the clusters' members are re-split
    //   by local Newton steps on their `k x k` blocks ( `resplit` ).
    std::vector<I64> dup, rep, dup_rep;                  ///< the pairs; per seed its representative, -1 if not a duplicate
    std::vector<double> dup_gap, nue;
    double c_link = 1;                                   ///< the coefficient of the facet linking a duplicate to its representative
    double margin = 0, tau_min = 0;                      ///< the criterion's `kappa` ( 0: no detection ), `tau` of the smallest target
    std::vector<I64> parent, cl_begin, cl_members, cluster_of;
    I64 nb_merged_pairs = 0;
    int nb_polish = 0;
    I64 find( I64 i ) {
        while ( parent[ sz( i ) ] != i ) { parent[ sz( i ) ] = parent[ sz( parent[ sz( i ) ] ) ]; i = parent[ sz( i ) ]; }
        return i;
    }
    bool unite( I64 i, I64 j, I64 n ) {
        if ( parent.empty() ) { parent.resize( sz( n ) ); for ( I64 k = 0; k < n; ++k ) parent[ sz( k ) ] = k; }
        i = find( i ); j = find( j );
        if ( i == j ) return false;
        parent[ sz( std::max( i, j ) ) ] = std::min( i, j );
        return true;
    }
    void rebuild( I64 n ) {
        cl_begin.clear(); cl_members.clear();
        cluster_of.assign( sz( n ), -1 );
        if ( parent.empty() ) return;
        std::vector<I64> size( sz( n ), 0 ), id( sz( n ), -1 );
        for ( I64 i = 0; i < n; ++i ) ++size[ sz( find( i ) ) ];
        cl_begin.push_back( 0 );
        for ( I64 i = 0; i < n; ++i )
            if ( parent[ sz( i ) ] == i && size[ sz( i ) ] > 1 ) { id[ sz( i ) ] = I64( cl_begin.size() ) - 1; cl_begin.push_back( cl_begin.back() + size[ sz( i ) ] ); }
        cl_members.resize( sz( cl_begin.back() ) );
        std::vector<I64> at( cl_begin.begin(), cl_begin.end() - 1 );
        for ( I64 i = 0; i < n; ++i ) {
            const I64 c = id[ sz( find( i ) ) ];
            if ( c < 0 ) continue;
            cluster_of[ sz( i ) ] = c;
            cl_members[ sz( at[ sz( c ) ]++ ) ] = i;
        }
    }
    I64 nb_clusters() const { return cl_begin.empty() ? 0 : I64( cl_begin.size() ) - 1; }
    bool any() const { return nb_clusters() > 0; }
    void set_duplicates( const std::vector<I64> &d, const std::vector<I64> &r, double gap, I64 n ) {
        dup = d; rep = r;
        dup_gap.clear();
        if ( d.empty() ) return;
        dup_rep.assign( sz( n ), -1 );
        std::vector<int> count( sz( n ), 0 );
        for ( size_t q = 0; q < d.size(); ++q ) {
            dup_rep[ sz( d[ q ] ) ] = r[ q ];
            dup_gap.push_back( gap * double( ++count[ sz( r[ q ] ) ] ) );
            unite( d[ q ], r[ q ], n );
        }
        rebuild( n );
    }
    bool is_dup( size_t i ) const { return ! dup_rep.empty() && dup_rep[ i ] >= 0; }
    /// the residual a seed of target `nu_i` must reach ( `Aggregation::tau` )
    double tau( double nu_i ) const { return std::max( { o.tol_abs, o.tol_rel * nu_i, 1e-12 * nu_i } ); }
    /// `nue` and `tau_min` for the current `nu` ( `Aggregation::prepare` )
    void prepare() {
        double nm = nu.empty() ? 0 : nu[ 0 ];
        for ( double v : nu ) nm = std::min( nm, v );
        tau_min = tau( nm );
        if ( dup.empty() ) { nue.clear(); return; }
        nue = nu;
        for ( size_t q = 0; q < dup.size(); ++q ) { nue[ sz( rep[ q ] ) ] += nu[ sz( dup[ q ] ) ]; nue[ sz( dup[ q ] ) ] = 0; }
    }
    const std::vector<double> &targets() const { return dup.empty() ? nu : nue; }
    /// the masses the tests read: each member's share of its cluster's mass ( `Aggregation::shares` )
    const std::vector<double> &shares( const std::vector<double> &A, std::vector<double> &ae ) const {
        if ( ! any() ) return A;
        ae = A;
        for ( I64 c = 0; c < nb_clusters(); ++c ) {
            double sa = 0, sn = 0;
            for ( I64 k = cl_begin[ sz( c ) ]; k < cl_begin[ sz( c + 1 ) ]; ++k ) { sa += A[ sz( cl_members[ sz( k ) ] ) ]; sn += nu[ sz( cl_members[ sz( k ) ] ) ]; }
            const double r = sn > 0 ? sa / sn : 0;
            for ( I64 k = cl_begin[ sz( c ) ]; k < cl_begin[ sz( c + 1 ) ]; ++k ) ae[ sz( cl_members[ sz( k ) ] ) ] = nu[ sz( cl_members[ sz( k ) ] ) ] * r;
        }
        return ae;
    }
    /// the spacing of the doubles at `x`
    static double ulp_of( double x ) { x = std::fabs( x ); return std::nextafter( x, INFINITY ) - x; }
    /// THE DETECTION ( `Aggregation::detect` ) on the Laplacian of the accepted diagram and its weights: `true` if a cluster
    /// was made or grew. The scan is skipped while no diagonal passes the bound `tau_min / ( margin ulp( max |w| ) )`
    bool detect( const Laplacian &L, const std::vector<double> &W ) {
        if ( ! ( margin > 0 ) || ! ( tau_min > 0 ) ) return false;
        double wmax = 0, dmax = 0;
        for ( double x : W ) wmax = std::max( wmax, std::fabs( x ) );
        const double c_star = tau_min / ( margin * ulp_of( wmax ) );
        for ( I64 i = 0; i < L.n; ++i ) dmax = std::max( dmax, L.dia[ sz( i ) ] );
        if ( ! ( dmax > c_star ) ) return false;
        bool changed = false;
        for ( I64 i = 0; i < L.n; ++i ) {
            if ( ! ( L.dia[ sz( i ) ] > c_star ) ) continue;
            for ( I64 e = L.row[ sz( i ) ]; e < L.row[ sz( i + 1 ) ]; ++e ) {
                const I64 j = L.col[ sz( e ) ];
                if ( j <= i || ! ( L.c[ sz( e ) ] > c_star ) ) continue;
                const double m = L.c[ sz( e ) ] * ulp_of( std::max( std::fabs( W[ sz( i ) ] ), std::fabs( W[ sz( j ) ] ) ) );
                if ( m * margin > tau( std::min( nu[ sz( i ) ], nu[ sz( j ) ] ) ) && unite( i, j, L.n ) ) { changed = true; ++nb_merged_pairs; }
            }
        }
        if ( changed ) rebuild( L.n );
        return changed;
    }
    /// the members of cluster `c` the re-splitting moves ( the duplicates left out ), increasing
    void movable_members( I64 c, std::vector<I64> &mem ) const {
        mem.clear();
        for ( I64 k = cl_begin[ sz( c ) ]; k < cl_begin[ sz( c + 1 ) ]; ++k )
            if ( ! is_dup( sz( cl_members[ sz( k ) ] ) ) ) mem.push_back( cl_members[ sz( k ) ] );
    }
    /// the members' largest residual ( the duplicates left out )
    double members_residual( const std::vector<double> &A ) const {
        const std::vector<double> &N = targets();
        double worst = 0;
        for ( I64 i : cl_members ) if ( ! is_dup( sz( i ) ) ) worst = std::max( worst, std::fabs( A[ sz( i ) ] - N[ sz( i ) ] ) );
        return worst;
    }
    /// `M x = r` for a symmetric positive definite `k x k` `M` ( row-major, overwritten ) by Cholesky; `false` if not definite
    static bool cholesky( I64 k, std::vector<double> &M, std::vector<double> &r ) {
        for ( I64 j = 0; j < k; ++j ) {
            double dj = M[ sz( j * k + j ) ];
            for ( I64 q = 0; q < j; ++q ) dj -= M[ sz( j * k + q ) ] * M[ sz( j * k + q ) ];
            if ( ! ( dj > 0 ) ) return false;
            const double lj = std::sqrt( dj );
            M[ sz( j * k + j ) ] = lj;
            for ( I64 i = j + 1; i < k; ++i ) {
                double v = M[ sz( i * k + j ) ];
                for ( I64 q = 0; q < j; ++q ) v -= M[ sz( i * k + q ) ] * M[ sz( j * k + q ) ];
                M[ sz( i * k + j ) ] = v / lj;
            }
        }
        for ( I64 i = 0; i < k; ++i ) {
            double v = r[ sz( i ) ];
            for ( I64 q = 0; q < i; ++q ) v -= M[ sz( i * k + q ) ] * r[ sz( q ) ];
            r[ sz( i ) ] = v / M[ sz( i * k + i ) ];
        }
        for ( I64 i = k - 1; i >= 0; --i ) {
            double v = r[ sz( i ) ];
            for ( I64 q = i + 1; q < k; ++q ) v -= M[ sz( q * k + i ) ] * r[ sz( q ) ];
            r[ sz( i ) ] = v / M[ sz( i * k + i ) ];
        }
        return true;
    }
    /// one local Newton step per cluster ( `Aggregation::local_step` ): `L_CC dw_C = nu_C - a_C` on its movable members, the
    /// rest fixed ( `dw` 0 elsewhere; clusters of one movable member or more than 64 left as they are ); returns the
    /// members' largest residual before the step
    double local_step( const Laplacian &L, const std::vector<double> &A, std::vector<double> &dw ) const {
        const std::vector<double> &N = targets();
        dw.assign( sz( L.n ), 0.0 );
        double worst = 0;
        std::vector<I64> mem;
        std::vector<double> M, r;
        for ( I64 c = 0; c < nb_clusters(); ++c ) {
            movable_members( c, mem );
            const I64 k = I64( mem.size() );
            for ( I64 i : mem ) worst = std::max( worst, std::fabs( A[ sz( i ) ] - N[ sz( i ) ] ) );
            if ( k < 2 || k > 64 ) continue;
            M.assign( sz( k * k ), 0.0 );
            r.resize( sz( k ) );
            for ( I64 u = 0; u < k; ++u ) {
                const I64 i = mem[ sz( u ) ];
                M[ sz( u * k + u ) ] = L.dia[ sz( i ) ];
                r[ sz( u ) ] = N[ sz( i ) ] - A[ sz( i ) ];
                for ( I64 e = L.row[ sz( i ) ]; e < L.row[ sz( i + 1 ) ]; ++e ) {
                    const auto it = std::lower_bound( mem.begin(), mem.end(), L.col[ sz( e ) ] );
                    if ( it != mem.end() && *it == L.col[ sz( e ) ] ) M[ sz( u * k + I64( it - mem.begin() ) ) ] -= L.c[ sz( e ) ];
                }
            }
            if ( ! cholesky( k, M, r ) ) continue;
            for ( I64 u = 0; u < k; ++u ) dw[ sz( mem[ sz( u ) ] ) ] = r[ sz( u ) ];
        }
        return worst;
    }
    int res_cur = LIN;
    double t_last = 1;
    // stats
    int status = 0, nb_iter = 0, nb_backtracks = 0, it_switch = -1;
    I64 nb_cell_lim = 0, nb_limit_rounds = 0;
    double residual = 0, residual0 = 0, eps = 0, t_asm = 0, t_lim = 0;
    LinStats lst;
    std::unique_ptr<MgSolver> mg;                        ///< `o.linear == 1`: kept over the solve ( hierarchy, subspace )
    std::function<void( int, double, int )> after_step;

    SolverT( SW &sw, const Options &o ) : sw( sw ), o( o ) {}

    /// the Newton system, by the solver `o.linear` names
    bool linear_solve( const Laplacian &L, const std::vector<double> &b, std::vector<double> &dd ) {
        if ( o.linear != 1 ) return cg_solve( L, b, dd, o.lin_tol, o.lin_maxit, lst );
        if ( ! mg ) {
            mg = std::make_unique<MgSolver>( nb_threads() );
            mg->nu = SW::DIM >= 3 ? 1 : 3;
            mg->set_ranks( sw.t.rank_of );
        }
        return mg->solve( L, b, dd, o.lin_tol, o.lin_maxit, lst );
    }

    /// `Newton.h:176-203`: on the targets `nue` when there are duplicates ( a duplicate's `b` is 0 )
    void rhs( const std::vector<double> &A, std::vector<double> &b ) const {
        const std::vector<double> &N = dup.empty() ? nu : nue;
        const size_t n = A.size();
        b.assign( n, 0.0 );
        if ( res_cur == LIN ) { for ( size_t i = 0; i < n; ++i ) b[ i ] = N[ i ] > 0 ? N[ i ] - A[ i ] : 0.0; return; }
        double su = 0, sug = 0;
        for ( size_t i = 0; i < n; ++i ) {
            if ( ! ( N[ i ] > 0 ) ) continue;
            const double x = A[ i ] / N[ i ], u = N[ i ] / gp_of( x, res_cur, o.power );
            su += u;
            sug += u * g_of( x, res_cur, o.power );
        }
        const double c = sug / su;
        for ( size_t i = 0; i < n; ++i ) {
            if ( ! ( N[ i ] > 0 ) ) continue;
            const double x = A[ i ] / N[ i ];
            b[ i ] = N[ i ] / gp_of( x, res_cur, o.power ) * ( c - g_of( x, res_cur, o.power ) );
        }
    }
    static void project_on_range( std::vector<double> &b ) {
        if ( b.empty() ) return;
        double s = 0;
        for ( double v : b ) s += v;
        s /= double( b.size() );
        for ( double &v : b ) v -= s;
    }
    double merit( const std::vector<double> &A ) const {
        const size_t n = A.size();
        double s = 0;
        if ( res_cur == LIN ) {
            for ( size_t i = 0; i < n; ++i ) s += ( nu[ i ] - A[ i ] ) * ( nu[ i ] - A[ i ] );
            return std::sqrt( s );
        }
        double m = 0;
        for ( size_t i = 0; i < n; ++i ) m += g_of( A[ i ] / nu[ i ], res_cur, o.power );
        m /= double( n );
        for ( size_t i = 0; i < n; ++i ) { const double e = g_of( A[ i ] / nu[ i ], res_cur, o.power ) - m; s += e * e; }
        return std::sqrt( s );
    }
    /// the smallest mass the floor reads, the duplicates ( empty by construction ) left out ( `Aggregation::floor_min` )
    double floor_min( const std::vector<double> &A ) const {
        double m = 1e300;
        for ( size_t i = 0; i < A.size(); ++i ) if ( ! is_dup( i ) ) m = std::min( m, A[ i ] );
        return A.empty() ? 0 : m;
    }
    /// `max |a - nue|` and its relative form over the seeds that are not duplicates ( `Aggregation::full_residual` )
    void full_residual( const std::vector<double> &A, double &worst, double &worst_rel ) const {
        const std::vector<double> &N = dup.empty() ? nu : nue;
        worst = worst_rel = 0;
        for ( size_t i = 0; i < A.size(); ++i ) {
            if ( is_dup( i ) ) continue;
            const double e = std::fabs( A[ i ] - N[ i ] );
            worst = std::max( worst, e );
            worst_rel = std::max( worst_rel, e / N[ i ] );
        }
    }
    void tie( std::vector<double> &W ) const {
        for ( size_t q = 0; q < dup.size(); ++q ) W[ sz( dup[ q ] ) ] = W[ sz( rep[ q ] ) ] - dup_gap[ q ];
    }
    /// the duplicates' rows linked to their representatives by a facet of coefficient `c_link` ( `Aggregation::link_duplicates` )
    void link( const std::vector<Facet> &f, std::vector<Facet> &out ) const {
        out = f;
        for ( size_t q = 0; q < dup.size(); ++q ) {
            out.push_back( Facet{ dup[ q ], rep[ q ], c_link } );
            out.push_back( Facet{ rep[ q ], dup[ q ], c_link } );
        }
    }
    void measures_and_facets( std::vector<double> &W, std::vector<double> &res, std::vector<Facet> &f, Edges *E = nullptr ) {
        tie( W );
        sw.set_weights( W );
        sw.edges = E;
        sw.measures( res, &f );
        sw.edges = nullptr;
    }

    /// THE RE-SPLITTING once the aggregated problem converged ( `Newton.h:431-470` ): local Newton steps on the clusters'
    /// members, one diagram each, kept while their residual decreases, the floor holds and the aggregated test still
    /// passes; stops as soon as the full problem passes the test. Works on the accepted `w`, `a`, `fa`.
    void resplit() {
        if ( ! any() ) return;
        const I64 n = I64( w.size() );
        Laplacian L;
        std::vector<double> dw, a2, w2, ae;
        std::vector<Facet> fa2;
        for ( int round = 0; round < 4; ++round ) {
            double wf, wfr;
            full_residual( a, wf, wfr );
            if ( wf <= o.tol_abs || ( o.tol_rel > 0 && wfr <= o.tol_rel ) ) return;
            L.assemble( n, fa );
            const double before = local_step( L, a, dw );
            const double d0 = dw[ 0 ];                   // the gauge `w_0 = 0`, kept by a shift of everyone
            w2.resize( sz( n ) );
            for ( I64 i = 0; i < n; ++i ) w2[ sz( i ) ] = w[ sz( i ) ] + dw[ sz( i ) ] - d0;
            measures_and_facets( w2, a2, fa2 );
            ++nb_polish;
            const double after = members_residual( a2 );
            const std::vector<double> &A2 = shares( a2, ae );
            double worst = 0, worst_rel = 0;
            for ( I64 i = 0; i < n; ++i ) {
                worst = std::max( worst, std::fabs( nu[ sz( i ) ] - A2[ sz( i ) ] ) );
                worst_rel = std::max( worst_rel, std::fabs( nu[ sz( i ) ] - A2[ sz( i ) ] ) / nu[ sz( i ) ] );
            }
            if ( ! ( after < before ) || ! ( floor_min( a2 ) > 0 ) || ! ( worst <= o.tol_abs || ( o.tol_rel > 0 && worst_rel <= o.tol_rel ) ) ) {
                sw.set_weights( w );                     // the diagram takes the accepted weights back
                return;
            }
            w.swap( w2 );
            a.swap( a2 );
            fa.swap( fa2 );
            residual = worst;
        }
    }

    /// `Newton.h:243-400` ( no aggregation ), the step by trials or by the limits ( `gpu/Newton2D.cuh:1615-1690` ): `a`, `fa`
    /// and `ea` already measured for `w_init`
    bool solves( const std::vector<double> &w_init ) {
        const I64 n = I64( w_init.size() );
        std::vector<double> a2, b, w2;
        std::vector<Facet> fa2;
        Edges e2;
        const bool limits = SW::DIM == 2 && o.step == LIMITS;
        Laplacian L;
        std::vector<double> ae, ae2;
        std::vector<Facet> fa_lin;
        prepare();
        w = w_init;
        const double gauge = w[ 0 ];
        for ( auto &v : w ) v -= gauge;
        tie( w );
        t_last = 1;
        if ( after_step ) after_step( 0, 0, 1 );
        res_cur = o.residual;
        for ( int it = 0; it < o.maxit; ++it ) {
            const std::vector<double> &A = shares( a, ae );
            double worst = 0, worst_rel = 0;
            for ( I64 i = 0; i < n; ++i ) {
                worst = std::max( worst, std::fabs( nu[ size_t( i ) ] - A[ size_t( i ) ] ) );
                worst_rel = std::max( worst_rel, std::fabs( nu[ size_t( i ) ] - A[ size_t( i ) ] ) / nu[ size_t( i ) ] );
            }
            if ( o.switch_residual > 0 && res_cur != LIN && worst_rel <= o.switch_residual ) { res_cur = LIN; it_switch = it; }
            rhs( a, b );
            project_on_range( b );
            if ( it == 0 ) {
                double nm = nu[ 0 ];
                for ( I64 i = 0; i < n; ++i ) nm = std::min( nm, nu[ size_t( i ) ] );
                eps = 0.5 * std::min( nm, floor_min( a ) );
                residual0 = worst;
            }
            double nr = merit( A );
            residual = worst;
            if ( worst <= o.tol_abs || ( o.tol_rel > 0 && worst_rel <= o.tol_rel ) ) { status = 1; resplit(); return true; }
            ++nb_iter;
            double t0 = now();
            if ( ! dup.empty() ) { link( fa, fa_lin ); L.assemble( n, fa_lin ); }
            else L.assemble( n, fa );
            // the pairs whose plane the doubles can no longer place are merged: the merit reads their clusters from here on
            if ( detect( L, w ) ) nr = merit( shares( a, ae ) );
            t_asm += now() - t0;
            if ( ! linear_solve( L, b, d ) ) { status = 4; return false; }
            double t = std::min( 1.0, o.mult_ok * t_last );
            bool taken = false, already = false;
            Hopeless hope;
            int nb_evals = 0;
            w2.resize( size_t( n ) );
            auto trial_diagram = [&]( double tt ) {
                for ( I64 i = 0; i < n; ++i ) w2[ size_t( i ) ] = w[ size_t( i ) ] + tt * d[ size_t( i ) ];
                w2[ 0 ] = 0;
                measures_and_facets( w2, a2, fa2, limits ? &e2 : nullptr );
                ++nb_evals;
            };
            if constexpr ( SW::DIM == 2 ) if ( limits ) {
                // FORWARD on the accepted diagram, then up to eight rounds of BACKWARD on the trial diagrams
                double th = now();
                const Along fw{ pos, w.data(), d.data(), sw.lo, sw.hi, &ea, n, &sw.dn };
                double am = 1e300;
                for ( I64 k = 0; k < n; ++k ) if ( ! is_dup( sz( k ) ) ) am = std::min( am, alpha_forward( fw, k, eps ) );
                t_lim += now() - th;
                t = am >= 1 ? 1.0 : o.factor * am;
                if ( sw.dn.kind != 0 && ! ( t >= o.t_min ) ) t = std::min( 1.0, o.mult_ok * t_last );
                double t_done = -1;
                for ( int round = 0; round < 8; ++round ) {
                    trial_diagram( t );
                    t_done = t;
                    I64 nb_below = 0;
                    for ( I64 k = 0; k < n; ++k ) nb_below += ! is_dup( sz( k ) ) && a2[ sz( k ) ] < eps;
                    if ( nb_below == 0 ) break;
                    ++nb_limit_rounds;
                    th = now();
                    const Along bw{ pos, w.data(), d.data(), sw.lo, sw.hi, &e2, n, &sw.dn };
                    double mb = 1e300;
                    for ( I64 k = 0; k < n; ++k )
                        if ( ! is_dup( sz( k ) ) && a2[ size_t( k ) ] < eps ) { ++nb_cell_lim; mb = std::min( mb, alpha_backward( bw, k, eps, t ) ); }
                    t_lim += now() - th;
                    t = o.factor * std::min( t, mb );
                    if ( t < o.t_min ) break;
                }
                if ( t < o.t_min ) t = t_done / 2;
                already = t == t_done;
            }
            for ( int trial = 0; trial < o.max_backtracks; ++trial ) {
                if ( ! ( trial == 0 && already ) ) trial_diagram( t );
                const double m2 = floor_min( a2 );
                const double n2r = merit( shares( a2, ae2 ) );
                if ( m2 >= eps && std::isfinite( n2r ) && n2r <= ( 1 - t / 2 ) * nr && n2r < nr ) { taken = true; break; }
                const bool hopeless = hope.refused( t, n2r, nr, m2 >= eps && std::isfinite( n2r ) );
                t /= 2;
                ++nb_backtracks;
                if ( t < o.t_min || hopeless ) break;
            }
            if ( ! taken && res_cur != LIN ) {
                sw.set_weights( w );
                res_cur = LIN;
                it_switch = it;
                --nb_iter;
                --it;
                continue;
            }
            if ( ! taken ) { sw.set_weights( w ); status = 3; return false; }
            t_last = t;
            w.swap( w2 );
            a.swap( a2 );
            fa.swap( fa2 );
            if ( limits ) std::swap( ea, e2 );
            if ( after_step ) after_step( it + 1, t, nb_evals );
        }
        const std::vector<double> &A = shares( a, ae );
        double worst = 0;
        for ( I64 i = 0; i < n; ++i ) worst = std::max( worst, std::fabs( nu[ size_t( i ) ] - A[ size_t( i ) ] ) );
        residual = worst;
        status = 2;
        return false;
    }
};

using Solver = SolverT<Sweep>;

/// the similarity start ( `sdotplan/Solve.h:69-100` )
static void similarity( const Tree &t, const double *pos, const double *lo, const double *hi, std::vector<double> &w, double margin = 0.1 ) {
    const I64 n = t.n;
    double plo[ 2 ] = { 1e300, 1e300 }, phi[ 2 ] = { -1e300, -1e300 };
    for ( I64 k = 0; k < n; ++k )
        for ( int d = 0; d < 2; ++d ) { plo[ d ] = std::min( plo[ d ], pos[ 2 * k + d ] ); phi[ d ] = std::max( phi[ d ], pos[ 2 * k + d ] ); }
    double a = 1;
    for ( int d = 0; d < 2; ++d ) {
        const double span_dom = ( hi[ d ] - lo[ d ] ) * ( 1 - 2 * margin );
        const double span_pts = std::max( phi[ d ] - plo[ d ], 1e-300 );
        a = std::min( a, span_dom / span_pts );
    }
    double b[ 2 ];
    for ( int d = 0; d < 2; ++d ) b[ d ] = ( lo[ d ] + hi[ d ] ) / 2 - a * ( plo[ d ] + phi[ d ] ) / 2;
    w.assign( size_t( n ), 0.0 );
    for ( I64 k = 0; k < n; ++k ) {
        double pp = 0, qq = 0;
        for ( int d = 0; d < 2; ++d ) {
            const double q = a * pos[ 2 * k + d ] + b[ d ];
            pp += pos[ 2 * k + d ] * pos[ 2 * k + d ];
            qq += q * q;
        }
        w[ size_t( k ) ] = pp - qq / a;
    }
}

} // namespace orc

// ============================================================================================
// the C entry points, for ctypes
// ============================================================================================

using orc::I64;
using orc::sz;

/// the polygon domain of the next calls ( `oracle_set_domain`; empty: the box ), set by the caller around them
static orc::Domain g_domain;

extern "C" void oracle_set_domain( int nb, const double *xy ) {
    if ( nb >= 3 && xy ) g_domain.set( nb, xy );
    else g_domain = orc::Domain();
}

static orc::Density make_density( int kind, double value, int nx, int ny, const double *geo, const double *img ) {
    orc::Density d;
    d.dom = g_domain;
    d.kind = kind;
    d.value = value;
    if ( kind == 1 ) {
        d.nx = nx; d.ny = ny;
        d.x0 = geo[ 0 ]; d.y0 = geo[ 1 ]; d.hx = geo[ 2 ]; d.hy = geo[ 3 ];
        d.img = img;
    }
    if ( kind == 2 ) {                                   // `img` holds the Gaussians ( x, y, sigma, mass ), `nx` their count
        d.ng = nx;
        d.gauss = img;
    }
    return d;
}

extern "C" {

void oracle_set_threads( int nt ) { orc::g_threads = nt; }
int  oracle_get_threads() { return orc::nb_threads(); }

int oracle_max_depth_for( I64 n, I64 leaf ) { return orc::max_depth_for( n, leaf ); }

/// the tree of `pos[ n ][ 2 ]` with the majorants of `w` ( user order; null: zeros ). Outputs sized by
/// the caller: `order[ n ]`, `beg/end[ N ]`, `box[ N * 4 ]`, `wa[ N * 2 ]`, `wb[ N ]`, N = 2^depth - 1
int oracle_build_tree( I64 n, const double *pos, const double *w, I64 leaf, I64 *order, I64 *beg, I64 *end, double *box,
                       double *wa, double *wb ) {
    orc::Tree t;
    orc::build_tree( t, n, pos, leaf );
    std::vector<double> W( sz( n ), 0.0 );
    if ( w ) for ( I64 i = 0; i < n; ++i ) W[ size_t( i ) ] = w[ i ];
    orc::set_weights( t, W.data() );
    std::memcpy( order, t.order.data(), sizeof( I64 ) * size_t( n ) );
    std::memcpy( beg, t.beg.data(), sizeof( I64 ) * size_t( t.nb_nodes ) );
    std::memcpy( end, t.end.data(), sizeof( I64 ) * size_t( t.nb_nodes ) );
    std::memcpy( box, t.box.data(), sizeof( double ) * size_t( t.nb_nodes ) * 4 );
    std::memcpy( wa, t.wa.data(), sizeof( double ) * size_t( t.nb_nodes ) * 2 );
    std::memcpy( wb, t.wb.data(), sizeof( double ) * size_t( t.nb_nodes ) );
    return t.depth;
}

/// the majorant of one slice ( testing hook )
void oracle_weight_majorant( I64 m, const double *pos, const double *w, double *wa, double *wb ) {
    std::vector<double> px( sz( m ) ), py( sz( m ) );
    for ( I64 k = 0; k < m; ++k ) { px[ size_t( k ) ] = pos[ 2 * k ]; py[ size_t( k ) ] = pos[ 2 * k + 1 ]; }
    orc::weight_majorant( px.data(), py.data(), w, 0, m, wa, *wb );
}

/// the masses ( user order ) of the diagram of `( pos, w )` in the box `lo/hi` against a density
/// ( kind 0: constant `value`; kind 1: image `img[ ny ][ nx ]`, `geo = x0, y0, hx, hy` ), and if
/// `fi` is not null the facets `( fi, fj, fc )` ( user indices, rows in order ), at most `cap`;
/// returns the number of facets ( -1 if more than `cap` ). `ev[ 5 ]`: the event counts ( or null ).
I64 oracle_measures( I64 n, const double *pos, const double *w, const double *lo, const double *hi, int kind, double value,
                     int nx, int ny, const double *geo, const double *img, double *masses, I64 *fi, I64 *fj, double *fc, I64 cap,
                     I64 *ev, int brute_force ) {
    orc::Tree t;
    orc::build_tree( t, n, pos, 10 );
    std::vector<double> W( sz( n ), 0.0 );
    if ( w ) for ( I64 i = 0; i < n; ++i ) W[ size_t( i ) ] = w[ i ];
    orc::Sweep sw( t, lo, hi, make_density( kind, value, nx, ny, geo, img ) );
    sw.set_weights( W );
    std::vector<double> a;
    std::vector<orc::Facet> fa;
    if ( brute_force ) {
        a.assign( size_t( n ), 0.0 );
        orc::Poly c, piece;
        for ( I64 k = 0; k < n; ++k ) {
            const I64 i = t.order[ size_t( k ) ];
            if ( ! orc::make_cell_all( t, k, lo, hi, c, &g_domain ) ) continue;
            a[ size_t( i ) ] = orc::mass_and_facets( t, sw.dn, k, c, piece, fi != nullptr, [&]( I64 j, double cij ) {
                fa.push_back( orc::Facet{ i, t.order[ size_t( j ) ], cij } );
            } );
        }
        std::stable_sort( fa.begin(), fa.end(), []( const orc::Facet &x, const orc::Facet &y ) { return x.i < y.i; } );
    } else
        sw.measures( a, fi ? &fa : nullptr );
    std::memcpy( masses, a.data(), sizeof( double ) * size_t( n ) );
    if ( ev ) {
        ev[ 0 ] = sw.ev.side_tests; ev[ 1 ] = sw.ev.prune_vertex_tests; ev[ 2 ] = sw.ev.new_vertices;
        ev[ 3 ] = sw.ev.node_visits; ev[ 4 ] = sw.ev.cuts;
    }
    if ( ! fi ) return 0;
    if ( I64( fa.size() ) > cap ) return -1;
    for ( size_t q = 0; q < fa.size(); ++q ) { fi[ q ] = fa[ q ].i; fj[ q ] = fa[ q ].j; fc[ q ] = fa[ q ].c; }
    return I64( fa.size() );
}

/// the event counts of each cell ( testing hook, RANK order: `ev[ 5 k .. ]` side tests, pruning vertex tests, new
/// vertices, node visits, cuts of the cell of rank `k` ), and the user index of each rank in `order`
void oracle_cell_events( I64 n, const double *pos, const double *w, const double *lo, const double *hi, I64 *ev, I64 *order ) {
    orc::Tree t;
    orc::build_tree( t, n, pos, 10 );
    std::vector<double> W( sz( n ), 0.0 );
    if ( w ) for ( I64 i = 0; i < n; ++i ) W[ size_t( i ) ] = w[ i ];
    orc::set_weights( t, W.data() );
    orc::parallel_blocks( n, 256, [&]( int, I64 b, I64 e ) {
        orc::Poly c;
        c.reserve( 64 );
        for ( I64 k = b; k < e; ++k ) {
            orc::Events E;
            orc::make_cell( t, k, lo, hi, c, &E, &g_domain );
            I64 *o = ev + 5 * k;
            o[ 0 ] = E.side_tests; o[ 1 ] = E.prune_vertex_tests; o[ 2 ] = E.new_vertices; o[ 3 ] = E.node_visits; o[ 4 ] = E.cuts;
        }
    } );
    std::memcpy( order, t.order.data(), sizeof( I64 ) * size_t( n ) );
}

/// the vertices of the cell of user seed `i` ( absolute coordinates, counter-clockwise ) and its
/// cut ids ( user indices of the neighbours, -1 for the domain ); returns the vertex count
int oracle_cell( I64 n, const double *pos, const double *w, const double *lo, const double *hi, I64 i, double *xy, I64 *ids,
                 int cap, int brute_force ) {
    orc::Tree t;
    orc::build_tree( t, n, pos, 10 );
    std::vector<double> W( sz( n ), 0.0 );
    if ( w ) for ( I64 q = 0; q < n; ++q ) W[ size_t( q ) ] = w[ q ];
    orc::set_weights( t, W.data() );
    orc::Poly c;
    const I64 k = t.rank_of[ size_t( i ) ];
    const bool ok = brute_force ? orc::make_cell_all( t, k, lo, hi, c, &g_domain ) : orc::make_cell( t, k, lo, hi, c, nullptr, &g_domain );
    if ( ! ok ) return 0;
    if ( c.nb > cap ) return -c.nb;
    for ( int q = 0; q < c.nb; ++q ) {
        xy[ 2 * q ] = c.x[ size_t( q ) ] + t.px[ size_t( k ) ];
        xy[ 2 * q + 1 ] = c.y[ size_t( q ) ] + t.py[ size_t( k ) ];
        ids[ q ] = c.id[ size_t( q ) ] >= 0 ? t.order[ size_t( c.id[ size_t( q ) ] ) ] : c.id[ size_t( q ) ];
    }
    return c.nb;
}

/// the step's cell polynomials ( testing hook ): the diagram of `( pos, w )` gives each cell's edges; with them frozen,
/// `poly[ 3 i .. ]` = the area polynomial at `w + 0 d` ( `ok[ i ]` = 0 when there is none ), `fmass[ i ]` the frozen mass at
/// `w + t d`, `fwd[ i ]` the forward limit at floor `eps`, `bwd[ i ]` the backward target from the trial `t` ( all user order )
void oracle_along( I64 n, const double *pos, const double *w, const double *d, const double *lo, const double *hi, int kind, double value,
                   int nx, int ny, const double *geo, const double *img, double t, double eps, double *poly, int *ok, double *fmass,
                   double *fwd, double *bwd ) {
    orc::Tree tr;
    orc::build_tree( tr, n, pos, 10 );
    std::vector<double> W( w, w + n );
    const orc::Density dn = make_density( kind, value, nx, ny, geo, img );
    orc::Sweep sw( tr, lo, hi, dn );
    sw.set_weights( W );
    orc::Edges E;
    sw.edges = &E;
    std::vector<double> a;
    sw.measures( a, nullptr );
    const orc::Along al{ pos, w, d, lo, hi, &E, n, &sw.dn };
    for ( I64 i = 0; i < n; ++i ) {
        double a0 = 0, a1 = 0, a2 = 0, m = 0;
        ok[ i ] = al.area_polynomial( i, 0.0, a0, a1, a2 ) ? 1 : 0;
        poly[ 3 * i ] = a0; poly[ 3 * i + 1 ] = a1; poly[ 3 * i + 2 ] = a2;
        fmass[ i ] = al.frozen_mass( i, t, m ) ? m : std::nan( "" );
        fwd[ i ] = orc::alpha_forward( al, i, eps );
        bwd[ i ] = orc::alpha_backward( al, i, eps, t );
    }
}

/// the image at width `s` ( testing hook ): `out[ ny * nx ]`
void oracle_blur( int nx, int ny, double hx, double hy, const double *img, double s, double *out ) {
    orc::Density dn;
    dn.kind = 1; dn.nx = nx; dn.ny = ny; dn.hx = hx; dn.hy = hy; dn.img = img;
    std::vector<double> b;
    orc::blur_image( dn, s, b );
    std::memcpy( out, b.data(), sizeof( double ) * b.size() );
}

/// the Jacobi-PCG alone on the Laplacian of facets `( fi, fj, fc )`: `L d = b`, `d[ 0 ] = 0`
int oracle_cg( I64 n, I64 nf, const I64 *fi, const I64 *fj, const double *fc, const double *b, double tol, int maxit, double *d ) {
    std::vector<orc::Facet> fa( sz( nf ) );
    for ( I64 q = 0; q < nf; ++q ) fa[ size_t( q ) ] = orc::Facet{ fi[ q ], fj[ q ], fc[ q ] };
    orc::Laplacian L;
    L.assemble( n, fa );
    std::vector<double> B( b, b + n ), D;
    orc::LinStats st;
    const bool ok = orc::cg_solve( L, B, D, tol, maxit, st );
    std::memcpy( d, D.data(), sizeof( double ) * size_t( n ) );
    return ok ? st.nb_iter : -1;
}

/// the multigrid-preconditioned CG alone on the Laplacian of facets `( fi, fj, fc )`, the first level aggregated by
/// `rank_of[ i ] >> 3` ( null: the index ); `nu` smoothing steps; `L d = b`, `d[ 0 ] = 0`. Returns the CG iterations ( -1:
/// not converged ), `levels[ 0 ]` the depth of the hierarchy.
int oracle_mg( I64 n, I64 nf, const I64 *fi, const I64 *fj, const double *fc, const double *b, const I64 *rank_of, int nu,
               double tol, int maxit, double *d, int *levels ) {
    std::vector<orc::Facet> fa( sz( nf ) );
    for ( I64 q = 0; q < nf; ++q ) fa[ size_t( q ) ] = orc::Facet{ fi[ q ], fj[ q ], fc[ q ] };
    orc::Laplacian L;
    L.assemble( n, fa );
    orc::MgSolver mg( orc::nb_threads() );
    mg.nu = nu;
    if ( rank_of ) mg.set_ranks( std::vector<I64>( rank_of, rank_of + n ) );
    std::vector<double> B( b, b + n ), D;
    orc::LinStats st;
    const bool ok = mg.solve( L, B, D, tol, maxit, st );
    std::memcpy( d, D.data(), sizeof( double ) * size_t( n ) );
    if ( levels ) levels[ 0 ] = mg.nb_levels();
    return ok ? st.nb_iter : -1;
}

/// THE SOLVE ( `sdotplan/Solve.h:111-371`, no aggregation ).
/// `opts` ( 23 doubles ): tol_abs, tol_rel, maxit, max_backtracks, t_min, mult_ok, residual, power, switch, lin_tol,
/// lin_maxit, keep_start, step ( 0 trials, 1 limits ), factor, continuation ( 0 never, 1 auto, 2 always ), conv_threshold,
/// conv_s0, conv_ratio, conv_min, gap, nb_dups, margin, linear ( 0 Jacobi-PCG, 1 multigrid ). `start_box`: flag, lo[ 2 ], hi[ 2 ].
/// Outputs: `weights[ n ]`, `stats[ 31 ]`, `hist[ ( maxit + 1 ) * 8 ]` ( returns the row count ),
/// `masses[ n ]`, `bary[ n * 2 ]`, `cost[ 1 ]`.
int oracle_solve( I64 n, const double *pos, const double *nu_in, const double *w0, const double *lo, const double *hi, int kind,
                  double value, int nx, int ny, const double *geo, const double *img, const double *opts, const double *start_box,
                  const I64 *dups, double *weights, double *stats, double *hist, double *masses, double *bary, double *cost,
                  I64 *clusters ) {
    using namespace orc;
    const double t_begin = now();
    Options o;
    o.tol_abs = opts[ 0 ]; o.tol_rel = opts[ 1 ]; o.maxit = int( opts[ 2 ] ); o.max_backtracks = int( opts[ 3 ] );
    o.t_min = opts[ 4 ]; o.mult_ok = opts[ 5 ]; o.residual = int( opts[ 6 ] ); o.power = opts[ 7 ]; o.switch_residual = opts[ 8 ];
    o.lin_tol = opts[ 9 ]; o.lin_maxit = int( opts[ 10 ] ); o.keep_start = int( opts[ 11 ] );
    o.step = int( opts[ 12 ] ); o.factor = opts[ 13 ] > 0 ? opts[ 13 ] : 0.9;
    o.linear = int( opts[ 22 ] );
    const int continuation = int( opts[ 14 ] );
    const double conv_threshold = opts[ 15 ], conv_s0 = opts[ 16 ], conv_ratio = opts[ 17 ], conv_min = opts[ 18 ];
    Tree t;
    build_tree( t, n, pos, 10 );
    const Density dn0 = make_density( kind, value, nx, ny, geo, img );
    Sweep sw( t, lo, hi, dn0 );
    Solver nt( sw, o );
    nt.pos = pos;
    // the exact duplicates: `opts[ 19 ]` the gap, `opts[ 20 ]` the count of `( dup, rep )` pairs in `dups` ( user ids );
    // `opts[ 21 ]` the aggregation's margin ( 0: near-coincident pairs are not merged )
    const I64 nb_dups = dups ? I64( opts[ 20 ] ) : 0;
    nt.margin = opts[ 21 ];
    {
        std::vector<I64> dd, rr;
        for ( I64 q = 0; q < nb_dups; ++q ) { dd.push_back( dups[ 2 * q ] ); rr.push_back( dups[ 2 * q + 1 ] ); }
        nt.set_duplicates( dd, rr, opts[ 19 ], n );
        // the link's coefficient: the scale of a facet, the density or the mean density of the box ( `gpu/Newton2D.cuh:1273` )
        const double vol = ( hi[ 0 ] - lo[ 0 ] ) * ( hi[ 1 ] - lo[ 1 ] );
        nt.c_link = kind == 0 ? value : 1.0 / std::max( vol, 1e-300 );
    }
    // the stages of the width continuation ( `Solve.h:176-200`; an image only: a constant is positive everywhere )
    const bool possible = kind == 1 || kind == 2;
    const double s0 = conv_s0 > 0 ? conv_s0 : 0.5 * std::sqrt( ( hi[ 0 ] - lo[ 0 ] ) * ( hi[ 0 ] - lo[ 0 ] ) + ( hi[ 1 ] - lo[ 1 ] ) * ( hi[ 1 ] - lo[ 1 ] ) );
    double smallest = 0;                                 // the image's pixel, or the narrowest Gaussian ( `convolved/*.h::min_scale` )
    if ( kind == 1 ) smallest = std::min( dn0.hx, dn0.hy );
    if ( kind == 2 ) { smallest = 1e300; for ( int q = 0; q < dn0.ng; ++q ) smallest = std::min( smallest, dn0.gauss[ 4 * q + 2 ] ); }
    const double s_min = conv_min > 0 ? conv_min : smallest / 4;
    std::vector<double> scales = continuation == 2 && possible ? continuation_steps( s0, conv_ratio, s_min ) : std::vector<double>{ 0.0 };
    std::vector<double> blurred;
    double s_current = 0;
    auto set_stage = [&]( double sc ) {
        s_current = sc;
        if ( ! possible ) return;
        Density dn = dn0;
        if ( sc > 0 && kind == 1 ) { blur_image( dn0, sc, blurred ); dn.img = blurred.data(); }
        if ( kind == 2 ) dn.conv_s = sc;
        sw.dn = dn;
    };
    set_stage( scales[ 0 ] );
    const int cap_steps = o.maxit + 1;
    int nb_steps = 0, nb_steps_before = 0;
    nt.after_step = [&]( int it, double tt, int nb_evals ) {
        if ( nb_steps >= cap_steps ) return;
        std::vector<double> ae;
        const std::vector<double> &A = nt.shares( nt.a, ae );
        const double mn = nt.floor_min( nt.a );
        double mx = 0, l2 = 0;
        for ( I64 i = 0; i < n; ++i ) {
            mx = std::max( mx, std::fabs( A[ size_t( i ) ] - nt.nu[ size_t( i ) ] ) );
            l2 += ( A[ size_t( i ) ] - nt.nu[ size_t( i ) ] ) * ( A[ size_t( i ) ] - nt.nu[ size_t( i ) ] );
        }
        double *r = hist + size_t( nb_steps ) * NB_HIST;
        r[ H_STEP ] = nb_steps_before + it; r[ H_T ] = tt; r[ H_RESIDUAL_L2 ] = std::sqrt( l2 ); r[ H_MIN_MASS ] = mn;
        r[ H_MAX_RESIDUAL ] = mx; r[ H_NB_DIAG ] = sw.nb_diag; r[ H_NB_EVALS ] = nb_evals; r[ H_S ] = s_current;
        ++nb_steps;
    };
    std::vector<double> nu( nu_in, nu_in + n ), w( sz( n ), 0.0 );
    bool given = false;
    if ( w0 ) for ( I64 i = 0; i < n; ++i ) { w[ size_t( i ) ] = w0[ i ]; given |= w[ size_t( i ) ] != 0; }
    int start = given ? 0 : 1;
    const double gauge = w[ 0 ];
    for ( auto &v : w ) v -= gauge;
    std::vector<double> a;
    std::vector<Facet> fa;
    Edges ea;
    nt.measures_and_facets( w, a, fa, &ea );
    double nu_min = nu[ 0 ];
    for ( double v : nu ) nu_min = std::min( nu_min, v );
    auto min_mass = [&]( const std::vector<double> &m ) { return nt.floor_min( m ); };
    if ( given && ! o.keep_start && min_mass( a ) < 1e-3 * nu_min ) {
        std::vector<double> wz( sz( n ), 0.0 ), az;
        std::vector<Facet> fz;
        Edges ez;
        nt.measures_and_facets( wz, az, fz, &ez );
        if ( min_mass( az ) > min_mass( a ) ) { w.swap( wz ); a.swap( az ); fa.swap( fz ); std::swap( ea, ez ); start = 1; }
        else sw.set_weights( w );
    }
    if ( min_mass( a ) <= 0 && ! ( given && o.keep_start ) ) {
        double blo[ 2 ] = { lo[ 0 ], lo[ 1 ] }, bhi[ 2 ] = { hi[ 0 ], hi[ 1 ] };
        if ( start_box && start_box[ 0 ] > 0 ) { blo[ 0 ] = start_box[ 1 ]; blo[ 1 ] = start_box[ 2 ]; bhi[ 0 ] = start_box[ 3 ]; bhi[ 1 ] = start_box[ 4 ]; }
        std::vector<double> w1, a1;
        std::vector<Facet> f1;
        Edges e1;
        similarity( t, pos, blo, bhi, w1 );
        nt.measures_and_facets( w1, a1, f1, &e1 );
        if ( min_mass( a1 ) > min_mass( a ) ) { w.swap( w1 ); a.swap( a1 ); fa.swap( f1 ); std::swap( ea, e1 ); start = 2; }
        else sw.set_weights( w );
    }
    const double min_start_mass = min_mass( a );
    // AUTO: the density is missing where cells are -> the continuation, from the same start ( `Solve.h:228-236` )
    if ( continuation == 1 && possible && scales.size() == 1 && min_start_mass < conv_threshold * nu_min ) {
        scales = continuation_steps( s0, conv_ratio, s_min );
        set_stage( scales[ 0 ] );
        nt.measures_and_facets( w, a, fa, &ea );
    }
    double domain_mass = 0, residual0 = 0, eps0 = 0;
    int it_switch = -1;
    for ( size_t stage = 0; stage < scales.size(); ++stage ) {
        if ( stage > 0 ) {                               // the next density: the measures of the start redone
            set_stage( scales[ stage ] );
            nt.measures_and_facets( w, a, fa, &ea );
        }
        double nu_mass = 0;
        domain_mass = 0;
        for ( I64 i = 0; i < n; ++i ) { domain_mass += a[ size_t( i ) ]; nu_mass += nu[ size_t( i ) ]; }
        std::vector<double> nu_s = nu;
        if ( domain_mass > 0 && nu_mass > 0 && domain_mass != nu_mass )
            for ( auto &v : nu_s ) v *= domain_mass / nu_mass;
        nt.nu = nu_s;
        nt.a = a;
        nt.fa = fa;
        nt.ea = ea;
        nt.it_switch = -1;
        nt.solves( w );
        w = nt.w;
        a = nt.a;
        fa = nt.fa;
        ea = nt.ea;
        if ( stage == 0 ) { residual0 = nt.residual0; eps0 = nt.eps; it_switch = nt.it_switch; }
        nb_steps_before = nb_steps > 0 ? int( hist[ size_t( nb_steps - 1 ) * NB_HIST + H_STEP ] ) + 1 : 0;
        if ( nt.status == 4 ) break;
    }
    for ( I64 i = 0; i < n; ++i ) { weights[ i ] = w[ size_t( i ) ]; masses[ i ] = a[ size_t( i ) ]; }
    // the full problem's residual; with clusters, `converged` only when it passes the test too ( `gpu/Newton2D.cuh:1752-1760` )
    double full = nt.residual;
    int status = nt.status;
    if ( nt.any() ) {
        double wf, wfr;
        nt.full_residual( a, wf, wfr );
        full = wf;
        if ( status == 1 && ! ( wf <= o.tol_abs || ( o.tol_rel > 0 && wfr <= o.tol_rel ) ) ) status = 7;
    }
    const int nb_diag_descent = sw.nb_diag;
    {
        set_stage( 0 );                                  // the moments on the TRUE density
        sw.set_weights( w );
        std::vector<double> bc;
        double c = 0;
        sw.moments( bc, c );
        std::memcpy( bary, bc.data(), sizeof( double ) * size_t( n ) * 2 );
        *cost = c;
    }
    for ( int q = 0; q < NB_STATS; ++q ) stats[ q ] = 0;
    stats[ STATUS ] = status;
    stats[ RESIDUAL ] = nt.residual;
    stats[ RESIDUAL0 ] = residual0;
    stats[ NB_ITER ] = nt.nb_iter;
    stats[ NB_DIAG ] = nb_diag_descent;
    stats[ NB_BACKTRACKS ] = nt.nb_backtracks;
    stats[ T_MAJORANT ] = sw.t_majorant;
    stats[ T_DIAG ] = sw.t_diag;
    stats[ T_ASM ] = nt.t_asm;
    stats[ T_LIN ] = nt.lst.t;
    stats[ EPS ] = eps0;
    stats[ T_LIM ] = nt.t_lim;
    stats[ NB_CELL_LIM ] = double( nt.nb_cell_lim );
    stats[ NB_LIMIT_ROUNDS ] = double( nt.nb_limit_rounds );
    stats[ DOMAIN_MASS ] = domain_mass;
    stats[ LIN_NB_ITER ] = nt.lst.nb_iter;
    stats[ LIN_WORST ] = nt.lst.worst;
    stats[ START ] = start;
    stats[ T_TOTAL ] = now() - t_begin;
    stats[ NB_CONTINUATION_STEPS ] = double( scales.size() );
    stats[ MIN_START_MASS ] = min_start_mass;
    stats[ IT_SWITCH ] = it_switch;
    stats[ RESIDUAL_FULL ] = full;
    stats[ NB_CLUSTERS ] = double( nt.nb_clusters() );
    stats[ NB_AGGREGATED ] = double( nt.cl_members.size() );
    stats[ NB_DUPLICATES ] = double( nb_dups );
    stats[ NB_POLISH ] = nt.nb_polish;
    if ( clusters ) {                                    // every seed its own index, a member its cluster's smallest
        for ( I64 i = 0; i < n; ++i ) clusters[ i ] = i;
        for ( I64 c = 0; c < nt.nb_clusters(); ++c )
            for ( I64 k = nt.cl_begin[ sz( c ) ]; k < nt.cl_begin[ sz( c + 1 ) ]; ++k ) clusters[ nt.cl_members[ sz( k ) ] ] = nt.cl_members[ sz( nt.cl_begin[ sz( c ) ] ) ];
    }
    return nb_steps;
}

// ---- 3D ---------------------------------------------------------------------------------------------------------------

/// the voxel image of the next 3D calls with density kind 1 ( `oracle_set_image3` ), kept by the caller around them
static orc::Density3 g_image3;

extern "C" void oracle_set_image3( int nx, int ny, int nz, const double *geo, const double *img ) {
    g_image3 = orc::Density3();
    if ( ! img ) return;
    g_image3.kind = 1;
    g_image3.img = img; g_image3.nx = nx; g_image3.ny = ny; g_image3.nz = nz;
    g_image3.x0 = geo[ 0 ]; g_image3.y0 = geo[ 1 ]; g_image3.z0 = geo[ 2 ]; g_image3.hx = geo[ 3 ]; g_image3.hy = geo[ 4 ]; g_image3.hz = geo[ 5 ];
}

static orc::Density3 make_density3( int kind, double value, int ng, const double *gauss ) {
    if ( kind == 1 ) return g_image3;
    orc::Density3 d;
    d.kind = kind;
    d.value = value;
    if ( kind == 2 ) { d.ng = ng; d.gauss = gauss; }
    return d;
}

/// the quadrature rules: `tet[ 14 ][ 5 ]` ( four barycentric coordinates, weight ), `tri[ 7 ][ 4 ]`
void oracle_rules( double *tet, double *tri ) {
    const orc::TetRule &T = orc::tet_rule();
    for ( int q = 0; q < 14; ++q ) { for ( int i = 0; i < 4; ++i ) tet[ 5 * q + i ] = T.b[ q ][ i ]; tet[ 5 * q + 4 ] = T.w[ q ]; }
    const orc::TriRule &R = orc::tri_rule();
    for ( int q = 0; q < 7; ++q ) { for ( int i = 0; i < 3; ++i ) tri[ 4 * q + i ] = R.b[ q ][ i ]; tri[ 4 * q + 3 ] = R.w[ q ]; }
}

/// the 3D tree: `order[ n ]`, `beg / end[ N ]`, `box[ N ][ 6 ]`, and with weights the majorants `wa[ N ][ 3 ]`, `wb[ N ]`
int oracle_build_tree_3d( I64 n, const double *pos, const double *w, I64 leaf, I64 *order, I64 *beg, I64 *end, double *box, double *wa,
                          double *wb ) {
    orc::Tree3 t;
    orc::build_tree3( t, n, pos, leaf );
    if ( w ) orc::set_weights3( t, w );
    for ( I64 k = 0; k < n; ++k ) order[ k ] = t.order[ sz( k ) ];
    for ( I64 m = 0; m < t.nb_nodes; ++m ) {
        beg[ m ] = t.beg[ sz( m ) ]; end[ m ] = t.end[ sz( m ) ];
        for ( int c = 0; c < 6; ++c ) box[ 6 * m + c ] = t.box[ sz( 6 * m + c ) ];
        if ( wa ) for ( int d = 0; d < 3; ++d ) wa[ 3 * m + d ] = t.wa[ sz( 3 * m + d ) ];
        if ( wb ) wb[ m ] = t.wb[ sz( m ) ];
    }
    return t.depth;
}

/// the masses ( user order ) and, up to `cap`, the facets ( user indices ); returns the facet count ( all of them ).
/// `info[ 0 ]` the cells whose cut could not be closed, `info[ 1 .. 5 ]` the events ( side tests, pruning tests, new
/// vertices, node visits, cuts )
I64 oracle_measures_3d( I64 n, const double *pos, const double *w, const double *lo, const double *hi, int kind, double value, int ng,
                        const double *gauss, double *masses, I64 *fi, I64 *fj, double *fc, I64 cap, I64 *info ) {
    orc::Tree3 t;
    orc::build_tree3( t, n, pos, 10 );
    orc::Sweep3 sw( t, lo, hi, make_density3( kind, value, ng, gauss ) );
    std::vector<double> W( w, w + n ), a;
    std::vector<orc::Facet> fa;
    sw.set_weights( W );
    sw.measures( a, &fa );
    for ( I64 i = 0; i < n; ++i ) masses[ i ] = a[ sz( i ) ];
    for ( size_t q = 0; q < fa.size() && I64( q ) < cap; ++q ) { fi[ q ] = fa[ q ].i; fj[ q ] = fa[ q ].j; fc[ q ] = fa[ q ].c; }
    if ( info ) {
        info[ 0 ] = sw.nb_broken;
        info[ 1 ] = sw.ev.side_tests; info[ 2 ] = sw.ev.prune_vertex_tests; info[ 3 ] = sw.ev.new_vertices;
        info[ 4 ] = sw.ev.node_visits; info[ 5 ] = sw.ev.cuts;
    }
    return I64( fa.size() );
}

/// one cell ( user index `i` ): its vertices ( global frame, up to `cap` ) and the ids of its faces that carry a polygon
/// ( user index of the seed, or `-1 - f` for face `f` of the box ); `brute`: cut by every seed instead of the tree.
/// Returns the vertex count ( 0: empty ), `-1` if the cell's topology broke
int oracle_cell_3d( I64 n, const double *pos, const double *w, const double *lo, const double *hi, I64 i, int brute, double *xyz, int cap,
                    I64 *face_ids, double *face_areas, int *nb_faces ) {
    orc::Tree3 t;
    orc::build_tree3( t, n, pos, 10 );
    std::vector<double> W( w, w + n );
    orc::set_weights3( t, W.data() );
    const I64 k = t.rank_of[ sz( i ) ];
    orc::Poly3 c;
    const bool ok = brute ? orc::make_cell3_all( t, k, lo, hi, c ) : orc::make_cell3( t, k, lo, hi, c, nullptr );
    *nb_faces = 0;
    if ( ! ok ) return 0;
    if ( c.broken ) return -1;
    const double *p0 = &t.p[ sz( 3 * k ) ];
    for ( int v = 0; v < c.nv && v < cap; ++v ) { xyz[ 3 * v ] = c.x[ sz( v ) ] + p0[ 0 ]; xyz[ 3 * v + 1 ] = c.y[ sz( v ) ] + p0[ 1 ]; xyz[ 3 * v + 2 ] = c.z[ sz( v ) ] + p0[ 2 ]; }
    orc::Density3 one;
    int nf = 0;
    orc::for_each_face_and_tet( c, [&]( int f, const std::vector<int> &loop ) {
        if ( nf >= cap ) return;
        const int id = c.f[ sz( f ) ].id;
        face_ids[ nf ] = id >= 0 ? t.order[ sz( id ) ] : I64( id );
        face_areas[ nf ] = orc::face_integral( one, p0, c, loop );
        ++nf;
    }, []( const double *, const double *, const double *, const double *, double ) {} );
    *nb_faces = nf;
    return c.nv;
}

/// the 3D solve: the same options, statistics and history as `oracle_solve` ( `opts[ 12 ]`, the step, is by trials: the
/// limits are a 2D step ); a constant density ( kind 0 ) or a sum of Gaussians ( kind 2, `gauss[ ng ][ 5 ]` ); no duplicates
int oracle_solve_3d( I64 n, const double *pos, const double *nu_in, const double *w0, const double *lo, const double *hi, int kind,
                     double value, int ng, const double *gauss, const double *opts, double *weights, double *stats, double *hist,
                     double *masses, double *bary, double *cost ) {
    using namespace orc;
    const double t_begin = now();
    Options o;
    o.tol_abs = opts[ 0 ]; o.tol_rel = opts[ 1 ]; o.maxit = int( opts[ 2 ] ); o.max_backtracks = int( opts[ 3 ] );
    o.t_min = opts[ 4 ]; o.mult_ok = opts[ 5 ]; o.residual = int( opts[ 6 ] ); o.power = opts[ 7 ]; o.switch_residual = opts[ 8 ];
    o.lin_tol = opts[ 9 ]; o.lin_maxit = int( opts[ 10 ] ); o.keep_start = int( opts[ 11 ] );
    o.step = TRIALS;
    o.linear = int( opts[ 22 ] );
    const int continuation = int( opts[ 14 ] );
    const double conv_threshold = opts[ 15 ], conv_s0 = opts[ 16 ], conv_ratio = opts[ 17 ], conv_min = opts[ 18 ];
    Tree3 t;
    build_tree3( t, n, pos, 10 );
    const Density3 dn0 = make_density3( kind, value, ng, gauss );
    Sweep3 sw( t, lo, hi, dn0 );
    SolverT<Sweep3> nt( sw, o );
    nt.pos = pos;
    nt.set_duplicates( {}, {}, 0.0, n );
    const bool possible = kind == 2;
    double diam2 = 0;
    for ( int d = 0; d < 3; ++d ) diam2 += ( hi[ d ] - lo[ d ] ) * ( hi[ d ] - lo[ d ] );
    const double s0 = conv_s0 > 0 ? conv_s0 : 0.5 * std::sqrt( diam2 );
    double smallest = 1e300;
    for ( int q = 0; q < dn0.ng; ++q ) smallest = std::min( smallest, dn0.gauss[ 5 * q + 3 ] );
    const double s_min = conv_min > 0 ? conv_min : smallest / 4;
    std::vector<double> scales = continuation == 2 && possible ? continuation_steps( s0, conv_ratio, s_min ) : std::vector<double>{ 0.0 };
    double s_current = 0;
    auto set_stage = [&]( double sc ) {
        s_current = sc;
        Density3 dn = dn0;
        dn.conv_s = possible ? sc : 0;
        sw.dn = dn;
    };
    set_stage( scales[ 0 ] );
    const int cap_steps = o.maxit + 1;
    int nb_steps = 0, nb_steps_before = 0;
    nt.after_step = [&]( int it, double tt, int nb_evals ) {
        if ( nb_steps >= cap_steps ) return;
        const double mn = nt.floor_min( nt.a );
        double mx = 0, l2 = 0;
        for ( I64 i = 0; i < n; ++i ) {
            const double e = nt.a[ sz( i ) ] - nt.nu[ sz( i ) ];
            mx = std::max( mx, std::fabs( e ) );
            l2 += e * e;
        }
        double *r = hist + sz( nb_steps ) * NB_HIST;
        r[ H_STEP ] = nb_steps_before + it; r[ H_T ] = tt; r[ H_RESIDUAL_L2 ] = std::sqrt( l2 ); r[ H_MIN_MASS ] = mn;
        r[ H_MAX_RESIDUAL ] = mx; r[ H_NB_DIAG ] = sw.nb_diag; r[ H_NB_EVALS ] = nb_evals; r[ H_S ] = s_current;
        ++nb_steps;
    };
    std::vector<double> nu( nu_in, nu_in + n ), w( sz( n ), 0.0 );
    bool given = false;
    if ( w0 ) for ( I64 i = 0; i < n; ++i ) { w[ sz( i ) ] = w0[ i ]; given |= w[ sz( i ) ] != 0; }
    int start = given ? 0 : 1;
    const double gauge = w[ 0 ];
    for ( auto &v : w ) v -= gauge;
    std::vector<double> a;
    std::vector<Facet> fa;
    nt.measures_and_facets( w, a, fa );
    double nu_min = nu[ 0 ];
    for ( double v : nu ) nu_min = std::min( nu_min, v );
    auto min_mass = [&]( const std::vector<double> &m ) { return nt.floor_min( m ); };
    if ( given && ! o.keep_start && min_mass( a ) < 1e-3 * nu_min ) {
        std::vector<double> wz( sz( n ), 0.0 ), az;
        std::vector<Facet> fz;
        nt.measures_and_facets( wz, az, fz );
        if ( min_mass( az ) > min_mass( a ) ) { w.swap( wz ); a.swap( az ); fa.swap( fz ); start = 1; }
        else sw.set_weights( w );
    }
    if ( min_mass( a ) <= 0 && ! ( given && o.keep_start ) ) {
        std::vector<double> w1, a1;
        std::vector<Facet> f1;
        similarity3( pos, n, lo, hi, w1 );
        nt.measures_and_facets( w1, a1, f1 );
        if ( min_mass( a1 ) > min_mass( a ) ) { w.swap( w1 ); a.swap( a1 ); fa.swap( f1 ); start = 2; }
        else sw.set_weights( w );
    }
    const double min_start_mass = min_mass( a );
    if ( continuation == 1 && possible && scales.size() == 1 && min_start_mass < conv_threshold * nu_min ) {
        scales = continuation_steps( s0, conv_ratio, s_min );
        set_stage( scales[ 0 ] );
        nt.measures_and_facets( w, a, fa );
    }
    double domain_mass = 0, residual0 = 0, eps0 = 0;
    int it_switch = -1;
    for ( size_t stage = 0; stage < scales.size(); ++stage ) {
        if ( stage > 0 ) {
            set_stage( scales[ stage ] );
            nt.measures_and_facets( w, a, fa );
        }
        double nu_mass = 0;
        domain_mass = 0;
        for ( I64 i = 0; i < n; ++i ) { domain_mass += a[ sz( i ) ]; nu_mass += nu[ sz( i ) ]; }
        std::vector<double> nu_s = nu;
        if ( domain_mass > 0 && nu_mass > 0 && domain_mass != nu_mass )
            for ( auto &v : nu_s ) v *= domain_mass / nu_mass;
        nt.nu = nu_s;
        nt.a = a;
        nt.fa = fa;
        nt.it_switch = -1;
        nt.solves( w );
        w = nt.w;
        a = nt.a;
        fa = nt.fa;
        if ( stage == 0 ) { residual0 = nt.residual0; eps0 = nt.eps; it_switch = nt.it_switch; }
        nb_steps_before = nb_steps > 0 ? int( hist[ sz( nb_steps - 1 ) * NB_HIST + H_STEP ] ) + 1 : 0;
        if ( nt.status == 4 ) break;
    }
    for ( I64 i = 0; i < n; ++i ) { weights[ i ] = w[ sz( i ) ]; masses[ i ] = a[ sz( i ) ]; }
    const int nb_diag_descent = sw.nb_diag;
    {
        set_stage( 0 );
        sw.set_weights( w );
        std::vector<double> bc;
        double c = 0;
        sw.moments( bc, c );
        std::memcpy( bary, bc.data(), sizeof( double ) * sz( n ) * 3 );
        *cost = c;
    }
    for ( int q = 0; q < NB_STATS; ++q ) stats[ q ] = 0;
    stats[ STATUS ] = nt.status;
    stats[ RESIDUAL ] = nt.residual;
    stats[ RESIDUAL0 ] = residual0;
    stats[ NB_ITER ] = nt.nb_iter;
    stats[ NB_DIAG ] = nb_diag_descent;
    stats[ NB_BACKTRACKS ] = nt.nb_backtracks;
    stats[ T_MAJORANT ] = sw.t_majorant;
    stats[ T_DIAG ] = sw.t_diag;
    stats[ T_ASM ] = nt.t_asm;
    stats[ T_LIN ] = nt.lst.t;
    stats[ EPS ] = eps0;
    stats[ DOMAIN_MASS ] = domain_mass;
    stats[ LIN_NB_ITER ] = nt.lst.nb_iter;
    stats[ LIN_WORST ] = nt.lst.worst;
    stats[ START ] = start;
    stats[ T_TOTAL ] = now() - t_begin;
    stats[ NB_CONTINUATION_STEPS ] = double( scales.size() );
    stats[ MIN_START_MASS ] = min_start_mass;
    stats[ IT_SWITCH ] = it_switch;
    stats[ RESIDUAL_FULL ] = nt.residual;
    return nb_steps;
}

} // extern "C"
