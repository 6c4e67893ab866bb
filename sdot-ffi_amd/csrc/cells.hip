// The cells of a 2D power diagram on CDNA4.
//
// ONE LANE PER CELL, the lanes of a wave holding consecutive tree ranks ( neighbours in space, so the
// 64 walks of a wave touch the same nodes ). A cell is a convex polygon kept in LDS as three arrays
// strided by the block ( `x[ i * BLOCK + lane ]` ), in the seed's frame ( a vertex is `x - p0` ).
// Reference semantics followed here ( paths under `src/sdot/include/sdot/` ):
//   * the cut: `cell/Local2.h:206-298` ( in place, the outside a single cyclic run, intersections
//     anchored on the inside vertex ), the several-runs rule `cell/Local2.h:300-315`;
//   * the walk: `cell/Providers.h:167-236` ( depth first, nearest child first, a node tested when it
//     is reached ), the test `cell/Pruning.h:46-76` ( exact per-axis clamp of the free minimum );
//   * the bisector in the seed's frame ( `gpu/Cell2D.cuh:139-160`, `cell/Plane.h:39` );
//   * the facets `c_kj = int_facet rho / ( 2 |p_k - p_j| )` ( `sdotplan/Sweep.h:64-128` ).
// The DFS keeps no stack: the heights on the current path carry two bits each ( a sibling is still
// pending; the path went left ), and a pop climbs to the lowest pending height -- the node indices
// are arithmetic in the preorder tree ( left `m + 1`, right `m + 2^( h - 1 )` ).
// An image density is integrated on the cell's edges ( Green: `mass = loop_integral F dy`,
// `F( x, y ) = int_{x_ref}^x rho( s, y ) ds`, exact piece by piece since `F` is linear in `x` inside a
// pixel ), so the cell is never cut by the pixel grid.
// Cells that outgrow `CAP1` LDS slots are listed and redone by a second pass with global slots.
// Arithmetic is written with explicit `fma` and no contraction, the same expressions as the
// oracle's, so a constant density gives the oracle's masses to the bit on the same tree.

#include "cells.h"
#include "gauss.h"
#include "image.h"

#pragma clang fp contract( off )

namespace sdot_amd {

namespace {

enum { CUT_UNCHANGED = 0, CUT_DONE = 1, CUT_EMPTY = 2, CUT_NO_ROOM = 3 };

struct Ev {
    unsigned side = 0, prune = 0, newv = 0, visits = 0, cuts = 0;
};

/// a polygon's arrays: `STRIDE` lanes apart ( LDS ) or contiguous ( a global slot ), coordinates in `TT` ( double; float
/// for the walk of the FP32 kernel, whose vertices the finish re-solves in double )
template<int STRIDE, class TT = double>
struct Store {
    using T = TT;
    T   *x, *y;
    int *id;
    __device__ __forceinline__ T   &X( int i ) const { return x[ i * STRIDE ]; }
    __device__ __forceinline__ T   &Y( int i ) const { return y[ i * STRIDE ]; }
    __device__ __forceinline__ int &I( int i ) const { return id[ i * STRIDE ]; }
};

template<class S, class T = typename S::T>
__device__ __forceinline__ T side( const S &c, int i, T dx, T dy, T off ) {
    return fma( dx, c.X( i ), fma( dy, c.Y( i ), -off ) );
}

/// the outside run holding the largest `s`, when the signs come in several runs
template<class S, class T = typename S::T>
__device__ void main_run( const S &c, int nb, T dx, T dy, T off, int &i1, int &nb_out ) {
    int a = 0;
    T sa = side( c, 0, dx, dy, off );
    for ( int i = 1; i < nb; ++i ) {
        const T si = side( c, i, dx, dy, off );
        if ( si > sa ) { a = i; sa = si; }
    }
    i1 = a;
    for ( int q = a ? a - 1 : nb - 1; side( c, q, dx, dy, off ) > 0; q = q ? q - 1 : nb - 1 ) i1 = q;
    nb_out = 1;
    for ( int q = a + 1 < nb ? a + 1 : 0; side( c, q, dx, dy, off ) > 0; q = q + 1 < nb ? q + 1 : 0 ) ++nb_out;
    nb_out += a >= i1 ? a - i1 : a - i1 + nb;           // ( `( a - i1 ) mod nb` without a division )
}

/// the cut by `dx . v <= off`, in place; `CUT_NO_ROOM` leaves the cell intact
template<class S, class T = typename S::T>
__device__ int cut( const S &c, int &nb, int cap, T dx, T dy, T off, int pid, Ev &ev ) {
    // the signs, four vertices a step ( their reads issued together: the slots past `nb`, up to the next multiple of
    // four, exist since `cap` is one, and are masked ), with the mask of the outside ones among the first 64
    int nb_out = 0;
    unsigned long long out = 0;
    for ( int i = 0; i < nb; i += 4 ) {
        const T s0 = side( c, i, dx, dy, off ), s1 = side( c, i + 1, dx, dy, off );
        const T s2 = side( c, i + 2, dx, dy, off ), s3 = side( c, i + 3, dx, dy, off );
        const unsigned o = unsigned( s0 > 0 ) | unsigned( i + 1 < nb && s1 > 0 ) << 1 | unsigned( i + 2 < nb && s2 > 0 ) << 2 |
                           unsigned( i + 3 < nb && s3 > 0 ) << 3;
        nb_out += __popc( o );
        if ( i < 64 ) out |= ( unsigned long long )( o ) << i;
    }
    ev.side += nb;
    if ( nb_out == 0 ) return CUT_UNCHANGED;
    if ( nb_out == nb ) { nb = 0; return CUT_EMPTY; }
    int i1 = 0;
    if ( nb <= 64 ) {
        // the first outside vertex after an inside one; the outside vertices make one run iff there is one such
        const unsigned long long full = nb == 64 ? ~0ull : ( 1ull << nb ) - 1;
        const unsigned long long starts = out & ~( ( ( out << 1 ) | ( out >> ( nb - 1 ) ) ) & full );
        i1 = __ffsll( ( long long )( starts ) ) - 1;
        if ( __popcll( starts ) != 1 ) main_run( c, nb, dx, dy, off, i1, nb_out );
    } else {
        bool prev = side( c, nb - 1, dx, dy, off ) > 0;
        for ( int i = 0; i < nb; ++i ) {
            const bool o = side( c, i, dx, dy, off ) > 0;
            if ( o && ! prev ) { i1 = i; break; }
            prev = o;
        }
        for ( int o = 1, i = i1 + 1; o < nb_out; ++o, ++i ) {
            if ( i == nb ) i = 0;
            if ( ! ( side( c, i, dx, dy, off ) > 0 ) ) { main_run( c, nb, dx, dy, off, i1, nb_out ); break; }
        }
    }
    const int nb_in = nb - nb_out, new_nb = nb_in + 2;
    if ( new_nb > cap ) return CUT_NO_ROOM;
    // ( the indices mod `nb` by comparisons: `i1 < nb` and `nb_out <= nb`, so one subtraction at most )
    const int j0 = i1 ? i1 - 1 : nb - 1, j2 = i1 + nb_out - 1 >= nb ? i1 + nb_out - 1 - nb : i1 + nb_out - 1, j3 = j2 + 1 == nb ? 0 : j2 + 1;
    const T s0 = side( c, j0, dx, dy, off ), s1 = side( c, i1, dx, dy, off );
    const T s2 = side( c, j2, dx, dy, off ), s3 = side( c, j3, dx, dy, off );
    const T ta = s0 / ( s0 - s1 );
    const T pax = fma( c.X( i1 ) - c.X( j0 ), ta, c.X( j0 ) );
    const T pay = fma( c.Y( i1 ) - c.Y( j0 ), ta, c.Y( j0 ) );
    const T tb = s3 / ( s3 - s2 );
    const T pbx = fma( c.X( j2 ) - c.X( j3 ), tb, c.X( j3 ) );
    const T pby = fma( c.Y( j2 ) - c.Y( j3 ), tb, c.Y( j3 ) );
    const int bid = c.I( j2 );
    auto move = [&]( int d, int s ) { c.X( d ) = c.X( s ); c.Y( d ) = c.Y( s ); c.I( d ) = c.I( s ); };
    auto put = [&]( int d, T x, T y, int id ) { c.X( d ) = x; c.Y( d ) = y; c.I( d ) = id; };
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
    nb = new_nb;
    ev.newv += 2;
    ev.cuts += 1;
    return CUT_DONE;
}

/// can a seed of node `r` still cut the cell of `( p0, w0 )`? ( exact: per-axis clamp )
template<class S>
__device__ __forceinline__ bool can_cut( const S &c, int nb, const NodeRec &r, double p0x, double p0y, double w0, Ev &ev ) {
    ev.visits += 1;
    const double hax = r.wa_x / 2, hay = r.wa_y / 2, cb = w0 - r.wb;
    auto test = [&]( int i ) {
        const double vx = double( c.X( i ) ), vy = double( c.Y( i ) );
        const double X = vx + p0x, Y = vy + p0y;
        double yx = X + hax, yy = Y + hay;
        yx = yx < r.lo_x ? r.lo_x : ( yx > r.hi_x ? r.hi_x : yx );
        yy = yy < r.lo_y ? r.lo_y : ( yy > r.hi_y ? r.hi_y : yy );
        const double ux = yx - X, uy = yy - Y;
        return cb + ux * ux - vx * vx - r.wa_x * yx + uy * uy - vy * vy - r.wa_y * yy <= 0;
    };
    for ( int i = 0; i < nb; i += 2 ) {                  // ( two vertices a step, their reads issued together; `cap` is even )
        const bool t0 = test( i ), u1 = test( i + 1 ), t1 = i + 1 < nb && u1;
        if ( t0 ) { ev.prune += i + 1; return true; }
        if ( t1 ) { ev.prune += i + 2; return true; }
    }
    ev.prune += nb;
    return false;
}

__device__ __forceinline__ double proximity( const NodeRec &b, double p0x, double p0y ) {
    double e0 = fmax( b.lo_x - p0x, p0x - b.hi_x ), e1 = fmax( b.lo_y - p0y, p0y - b.hi_y );
    e0 = e0 > 0 ? e0 : 0;
    e1 = e1 > 0 ? e1 : 0;
    return e0 * e0 + e1 * e1;
}

/// the seed as the float pruning sees it: each coordinate and the weight split in two floats ( `x = xh + xl` ), so that the
/// box centred on the seed carries an error relative to its distance from the seed, not to the coordinates
struct Frame32 {
    float xh, xl, yh, yl, wh, wl;
};

__device__ __forceinline__ Frame32 frame32( double x, double y, double w ) {
    Frame32 f;
    f.xh = float( x ); f.xl = float( x - double( f.xh ) );
    f.yh = float( y ); f.yl = float( y - double( f.yh ) );
    f.wh = float( w ); f.wl = float( w - double( f.wh ) );
    return f;
}

/// the test of `can_cut` on a float record: the node's outward-rounded box centred on the seed, its rounded-up majorant,
/// in float. It answers "may cut" up to a margin of 1e-6 of the magnitude of its terms ( sixteen float epsilons, far above the
/// roundings of the float vertex, of the centred box and of the majorant's constant ), so that it prunes no node the exact
/// test would keep on the same polygon
template<class S>
__device__ __forceinline__ bool can_cut( const S &c, int nb, const NodeF &r, const Frame32 &f, Ev &ev ) {
    ev.visits += 1;
    const float lx = ( r.lo_x - f.xh ) - f.xl, ly = ( r.lo_y - f.yh ) - f.yl;
    const float hx = ( r.hi_x - f.xh ) - f.xl, hy = ( r.hi_y - f.yh ) - f.yl;
    const float cc = ( ( f.wh - r.wb ) + f.wl ) - ( r.wa_x * f.xh + r.wa_y * f.yh );
    const float hax = 0.5f * r.wa_x, hay = 0.5f * r.wa_y, acc = fabsf( cc );
    auto test = [&]( int i ) {
        const float vx = float( c.X( i ) ), vy = float( c.Y( i ) );
        const float yx = fminf( fmaxf( vx + hax, lx ), hx ), yy = fminf( fmaxf( vy + hay, ly ), hy );
        const float ux = yx - vx, uy = yy - vy;
        const float p = ux * ux + uy * uy, q = vx * vx + vy * vy;
        const float ax = r.wa_x * yx, ay = r.wa_y * yy;
        const float sv = ( p - q ) + ( cc - ( ax + ay ) );
        const float tv = ( p + q ) + ( acc + ( fabsf( ax ) + fabsf( ay ) ) );
        return sv <= 1e-6f * tv;
    };
    for ( int i = 0; i < nb; i += 2 ) {                  // ( two vertices a step, their reads issued together; `cap` is even )
        const bool t0 = test( i ), u1 = test( i + 1 ), t1 = i + 1 < nb && u1;
        if ( t0 ) { ev.prune += i + 1; return true; }
        if ( t1 ) { ev.prune += i + 2; return true; }
    }
    ev.prune += nb;
    return false;
}

/// `can_cut` on a double record ( the seed itself in `p0x, p0y, w0` )
template<class S>
__device__ __forceinline__ bool can_cut( const S &c, int nb, const NodeRec &r, double p0x, double p0y, double w0, const Frame32 &, Ev &ev ) {
    return can_cut( c, nb, r, p0x, p0y, w0, ev );
}
template<class S>
__device__ __forceinline__ bool can_cut( const S &c, int nb, const NodeF &r, double, double, double, const Frame32 &f, Ev &ev ) {
    return can_cut( c, nb, r, f, ev );
}

__device__ __forceinline__ double proximity( const NodeF &b, double p0x, double p0y ) {
    double e0 = fmax( double( b.lo_x ) - p0x, p0x - double( b.hi_x ) ), e1 = fmax( double( b.lo_y ) - p0y, p0y - double( b.hi_y ) );
    e0 = e0 > 0 ? e0 : 0;
    e1 = e1 > 0 ? e1 : 0;
    return e0 * e0 + e1 * e1;
}

struct Walk {
    const NodeRec *nodes;
    const NodeF   *nodes32;                              ///< the float records ( the FP32 walk )
    const int     *node_end;                             ///< each node's end ( a float record's leaf )
    const double4 *seeds;
    int depth;
};

/// the record of node `m` and what the walk needs of it, for either kind of record
__device__ __forceinline__ NodeRec node_at( const Walk &wk, int m, const NodeRec * ) { return wk.nodes[ m ]; }
__device__ __forceinline__ NodeF node_at( const Walk &wk, int m, const NodeF * ) { return wk.nodes32[ m ]; }
__device__ __forceinline__ bool node_live( const NodeRec &r ) { return r.beg < r.end; }
__device__ __forceinline__ bool node_live( const NodeF &r ) { return r.beg >= 0; }
__device__ __forceinline__ int node_stop( const Walk &, int, const NodeRec &r ) { return r.end; }
__device__ __forceinline__ int node_stop( const Walk &wk, int m, const NodeF & ) { return wk.node_end[ m ]; }

/// builds the cell of rank `k` in `c`; returns the vertex count ( 0: empty ), or -1 when the slots
/// were too few ( the cell is then left for the next pass ). One loop iteration does ONE thing of about
/// the same cost -- a cut by one seed of the open leaf, or one node's test ( after a pop if needed ) --
/// so the lanes of a wave, each at its own place in its walk, stay in step. The loads run one step
/// ahead of their use: a descent loads both children's records ( the nearer is the next node tested,
/// from registers ), and a leaf's next seed is loaded while the current one cuts.
template<class S, class ND = NodeRec>
__device__ __forceinline__ int make_cell( const Walk &wk, int k, const double *lo, const double *hi, const Polygon2 &dom, const S &c, int cap,
                                          Ev &ev ) {
    using T = typename S::T;
    const ND *kind = nullptr;                            // ( which records the walk reads: `NodeRec` or the float `NodeF` )
    const double4 p = wk.seeds[ k ];
    const double p0x = p.x, p0y = p.y, w0 = p.z;
    const Frame32 f32 = frame32( p0x, p0y, w0 );
    int nb = 4;
    if ( dom.nb > 0 ) {                                  // the polygon domain, laid down as it is ( its edges: `-5 - i` )
        if ( dom.nb > cap ) return -1;
        for ( int i = 0; i < dom.nb; ++i ) { c.X( i ) = T( dom.x[ i ] - p0x ); c.Y( i ) = T( dom.y[ i ] - p0y ); c.I( i ) = -5 - i; }
        nb = dom.nb;
    } else {
        c.X( 0 ) = T( lo[ 0 ] - p0x ); c.Y( 0 ) = T( lo[ 1 ] - p0y );
        c.X( 1 ) = T( hi[ 0 ] - p0x ); c.Y( 1 ) = T( lo[ 1 ] - p0y );
        c.X( 2 ) = T( hi[ 0 ] - p0x ); c.Y( 2 ) = T( hi[ 1 ] - p0y );
        c.X( 3 ) = T( lo[ 0 ] - p0x ); c.Y( 3 ) = T( hi[ 1 ] - p0y );
        for ( int i = 0; i < 4; ++i ) c.I( i ) = -1 - i; // the box's sides: bottom, right, top, left
    }
    int m = 0, h = wk.depth;
    unsigned long long pend = 0, left = 0;
    int q = 0, qe = 0;                                   // the open leaf's remaining seeds
    bool visit = true;                                   // ( m, h ) is still to be tested
    ND rec;                                              // the record of ( m, h ) when the descent loaded it
    bool have = false;
    double4 nxt = make_double4( 0, 0, 0, 0 );            // the open leaf's seed `q`, loaded ahead
    // two inner loops, "while-while": the open leaf's seeds, then the walk down to the next leaf that may cut. A wave's
    // lanes run one kind of step at a time ( one loop with both kinds of step in it runs both bodies whenever its lanes
    // disagree ); each lane takes the same steps in the same order either way
    for ( ;; ) {
        while ( q < qe ) {
            const int s = q++;
            const double4 sd = nxt;
            if ( q < qe ) nxt = wk.seeds[ q ];
            if ( s == k ) continue;
            const double dx = sd.x - p0x, dy = sd.y - p0y;
            const double off = ( ( w0 - sd.z ) + ( dx * dx + dy * dy ) ) / 2;   // ( `|d|^2` after the weights' difference )
            const int res = cut( c, nb, cap, T( dx ), T( dy ), T( off ), s, ev );
            if ( res == CUT_EMPTY ) return 0;
            if ( res == CUT_NO_ROOM ) return -1;
        }
        for ( bool open = false; ! open; ) {
            if ( ! visit ) {                                  // pop: climb to the lowest pending height, take the sibling
                if ( ! pend ) return nb;
                const int hp = __ffsll( (long long) pend ) - 1;
                if ( h < hp ) {                              // up from `h` to `hp`: a left child's parent is 1 before it, a right
                    const unsigned long long R = ( ( 1ull << hp ) - 1 ) & ~( ( 1ull << h ) - 1 );   // child's `2^height` before
                    m -= __popcll( left & R ) + int( R & ~left );
                    h = hp;
                }
                const unsigned long long bit = 1ull << hp;
                if ( left & bit ) { m = m - 1 + ( 1 << hp ); left &= ~bit; }
                else              { m = m + 1 - ( 1 << hp ); left |= bit; }
                pend &= ~bit;
            }
            visit = false;
            const ND r = have ? rec : node_at( wk, m, kind );
            have = false;
            if ( node_live( r ) && can_cut( c, nb, r, p0x, p0y, w0, f32, ev ) ) {
                if ( h <= 1 ) {
                    q = r.beg;
                    qe = node_stop( wk, m, r );
                    nxt = wk.seeds[ q ];
                    open = true;
                } else {
                    const int lc = m + 1, rc = m + ( 1 << ( h - 1 ) );
                    const ND rl = node_at( wk, lc, kind ), rr = node_at( wk, rc, kind );
                    const bool go_left = proximity( rl, p0x, p0y ) <= proximity( rr, p0x, p0y );
                    const unsigned long long bit = 1ull << ( h - 1 );
                    pend |= bit;
                    left = go_left ? ( left | bit ) : ( left & ~bit );
                    m = go_left ? lc : rc;
                    rec = go_left ? rl : rr;
                    have = true;
                    h -= 1;
                    visit = true;
                }
            }
        }
    }
}

// ---- integrals -------------------------------------------------------------------------------------

template<class S>
__device__ __forceinline__ double shoelace( const S &c, int nb ) {
    double s = 0;
    for ( int i = 0; i < nb; ++i ) {
        const int j = i + 1 < nb ? i + 1 : 0;
        s += c.X( i ) * c.Y( j ) - c.X( j ) * c.Y( i );
    }
    return s / 2;
}

/// edge `e` of the cell against the image, in one walk: its term of `mass = loop_integral F0 dy` and
/// its integral `int_edge rho dl`
template<class S>
__device__ __forceinline__ void image_edge( const Grid &g, const S &c, int nb, int e, double &mass_term, double &line ) {
    const int f = e + 1 < nb ? e + 1 : 0;
    const double ax = c.X( e ), ay = c.Y( e ), bx = c.X( f ), by = c.Y( f );
    const double ex = bx - ax, ey = by - ay;
    double m = 0, l = 0;
    grid_walk( g, ax, ay, bx, by, [&]( int i, int j, double t0, double t1 ) {
        const double dt = t1 - t0;
        m += g.F0( i, j, fma( ex, 0.5 * ( t0 + t1 ), ax ) ) * ( ey * dt );
        l += g.val( i, j ) * dt;
    } );
    mass_term = m;
    line = l * sqrt( ex * ex + ey * ey );
}

/// moments of the cell against the image: `m`, `int v rho`, `int |v|^2 rho` ( seed frame ), by Green
/// with the row integrals of `rho`, `rho s`, `rho s^2`; two Gauss points per piece ( exact: cubic )
template<class S>
__device__ void image_moments( const Grid &g, const S &c, int nb, double &m, double &mx, double &my, double &m2 ) {
    m = mx = my = m2 = 0;
    const double gq = 0.5 / sqrt( 3.0 );
    for ( int e = 0; e < nb; ++e ) {
        const int f = e + 1 < nb ? e + 1 : 0;
        const double ax = c.X( e ), ay = c.Y( e ), bx = c.X( f ), by = c.Y( f );
        if ( ay == by ) continue;
        grid_walk( g, ax, ay, bx, by, [&]( int i, int j, double t0, double t1 ) {
            // row prefixes at the column's left side
            double c0 = 0, c1 = 0, c2 = 0;
            if ( i >= g.iref )
                for ( int q = g.iref; q < i; ++q ) {
                    const double v = g.val( q, j ), a = g.u( q ), b = g.u( q + 1 );
                    c0 += v * ( b - a ); c1 += v * ( b * b - a * a ) / 2; c2 += v * ( b * b * b - a * a * a ) / 3;
                }
            else
                for ( int q = i; q < g.iref; ++q ) {
                    const double v = g.val( q, j ), a = g.u( q ), b = g.u( q + 1 );
                    c0 -= v * ( b - a ); c1 -= v * ( b * b - a * a ) / 2; c2 -= v * ( b * b * b - a * a * a ) / 3;
                }
            const double v = g.val( i, j ), ui = g.u( i );
            const double dt = t1 - t0, tm = 0.5 * ( t0 + t1 );
            const double dy = ( by - ay ) * dt;
            for ( int s = -1; s <= 1; s += 2 ) {
                const double t = tm + s * gq * dt;
                const double x = fma( bx - ax, t, ax ), y = fma( by - ay, t, ay );
                const double f0 = c0 + v * ( x - ui );
                const double f1 = c1 + v * ( x * x - ui * ui ) / 2;
                const double f2 = c2 + v * ( x * x * x - ui * ui * ui ) / 3;
                m += 0.5 * dy * f0;
                mx += 0.5 * dy * f1;
                my += 0.5 * dy * y * f0;
                m2 += 0.5 * dy * ( f2 + y * y * f0 );
            }
        } );
    }
}

/// a sum of Gaussians on the cell ( seed at `( ox, oy )` ): per Gaussian its mass `w | corner sum |`, and for `mom` the
/// moments about the seed from the boundary fluxes of each edge
template<class S>
__device__ void gauss_cell( const Density &dn, double ox, double oy, const S &c, int nb, bool mom, double &m, double &mx, double &my, double &m2 ) {
    m = mx = my = m2 = 0;
    for ( int q = 0; q < dn.ng; ++q ) {
        const GaussTerm G = gauss_term( dn, q, ox, oy );
        const double W = gauss_corner_sum( G, [&]( auto &&f ) {
            for ( int e = 0; e < nb; ++e ) { const int r = e + 1 < nb ? e + 1 : 0; f( c.X( e ), c.Y( e ), c.X( r ), c.Y( r ) ); }
        } );
        const double mq = G.w * fabs( W );
        m += mq;
        if ( ! mom ) continue;
        double fx = 0, fy = 0, fr = 0;
        for ( int e = 0; e < nb; ++e ) {
            const int r = e + 1 < nb ? e + 1 : 0;
            const double ex = c.X( r ) - c.X( e ), ey = c.Y( r ) - c.Y( e );
            const double L = sqrt( ex * ex + ey * ey );
            if ( ! ( L > 0 ) ) continue;
            const double nx = ey / L, ny = -ex / L;      // outward: the cell is counter-clockwise
            const double J = gauss_segment_line( G, c.X( e ), c.Y( e ), c.X( r ), c.Y( r ) );
            fx += nx * J;
            fy += ny * J;
            fr += ( ( c.X( e ) - G.cx ) * nx + ( c.Y( e ) - G.cy ) * ny ) * J;
        }
        const double s2 = G.s * G.s, ix = -s2 * fx, iy = -s2 * fy, i2 = 2 * s2 * mq - s2 * fr;
        mx += ix + G.cx * mq;
        my += iy + G.cy * mq;
        m2 += i2 + 2 * ( G.cx * ix + G.cy * iy ) + ( G.cx * G.cx + G.cy * G.cy ) * mq;
    }
}

/// moments of a constant density: the fan from vertex 0 ( `diagram/Ops.h:187-228` )
template<class S>
__device__ void const_moments( const S &c, int nb, double rho, double &m, double &mx, double &my, double &m2 ) {
    m = mx = my = m2 = 0;
    const double ax = c.X( 0 ), ay = c.Y( 0 );
    for ( int i = 1; i + 1 < nb; ++i ) {
        const double bx = c.X( i ), by = c.Y( i ), cx = c.X( i + 1 ), cy = c.Y( i + 1 );
        const double ar = fabs( ( bx - ax ) * ( cy - ay ) - ( by - ay ) * ( cx - ax ) ) / 2;
        const double w = rho * ar;
        const double sx = ax + bx + cx, sy = ay + by + cy;
        const double sq = ax * ax + ay * ay + bx * bx + by * by + cx * cx + cy * cy;
        m += w;
        mx += w * sx / 3;
        my += w * sy / 3;
        m2 += w * ( sq + sx * sx + sy * sy ) / 12;
    }
}

// ---- wave helpers -------------------------------------------------------------------------------------

/// the largest `v` of the wave
__device__ __forceinline__ unsigned wave_max( unsigned v ) {
    for ( int off = 32; off > 0; off >>= 1 ) { const unsigned o = __shfl_xor( v, off, 64 ); v = o > v ? o : v; }
    return v;
}

/// exclusive prefix sum of `v` over the wave, and the wave's total
__device__ __forceinline__ unsigned wave_scan( unsigned v, unsigned &total ) {
    const int lane = threadIdx.x & 63;
    unsigned x = v;
    for ( int off = 1; off < 64; off <<= 1 ) {
        const unsigned y = __shfl_up( x, off, 64 );
        if ( lane >= off ) x += y;
    }
    total = __shfl( x, 63, 64 );
    return x - v;
}

__device__ __forceinline__ unsigned long long wave_sum64( unsigned long long v ) {
    for ( int off = 32; off > 0; off >>= 1 ) v += __shfl_down( v, off, 64 );
    return v;
}

/// the polygons of the FP32 kernel's walk, for its finish: slot `i` of rank `k` at `i n + k`
struct Poly32 {
    float *x, *y;
    int   *id, *nb;
};

struct Params {
    Walk          wk;
    int           n;
    double        lo[ 2 ], hi[ 2 ];
    Density       dn;
    DiagramOut    out;
    int           mode;
};

/// the end of a cell: its mass, and its facets ( one atomic per wave ) or its moments
template<class S>
__device__ void finish_cell( const Params &P, int k, const S &c, int nb, bool active, bool wave_full ) {
    const double4 p = P.wk.seeds[ k < P.n ? k : 0 ];
    // facets: the live cuts that face a seed
    unsigned cnt = 0;
    if ( active && P.mode == MODE_FACETS )
        for ( int e = 0; e < nb; ++e ) cnt += c.I( e ) >= 0;
    // the wave's rows interleaved: `64 x` its longest row, one atomic for the wave
    const int lane = threadIdx.x & 63;
    unsigned wmax = 0, base = 0;
    if ( P.mode == MODE_FACETS ) {
        wmax = wave_max( cnt ) + unsigned( P.out.row_extra );
        if ( lane == 0 && wmax ) base = atomicAdd( &P.out.counters[ 0 ], unsigned( ROW_STEP ) * wmax );
        base = __shfl( base, 0, 64 );
    }
    if ( ! active ) return;
    Grid g{};
    if ( P.dn.kind == 1 ) {
        g = grid_from( P.dn, p.x, p.y );
        if ( nb > 0 ) g.anchor( c.X( 0 ) );
    }
    if ( P.mode == MODE_CELLS ) {
        P.out.cell_nb[ k ] = nb;
        if ( P.out.cell_off ) {
            const long long o = P.out.cell_off[ P.out.order[ k ] ];
            for ( int e = 0; e < nb; ++e ) {
                P.out.cell_xy[ 2 * ( o + e ) ] = c.X( e ) + p.x;
                P.out.cell_xy[ 2 * ( o + e ) + 1 ] = c.Y( e ) + p.y;
                const int id = c.I( e );
                P.out.cell_cut[ o + e ] = id >= 0 ? (long long) P.out.order[ id ] : (long long) id;
            }
        }
        return;
    }
    if ( P.mode == MODE_VJP ) {
        // the adjoint of the masses: for each facet ( k, j ) of density integral `F` and first moment `M = int rho ( x - p_k )`,
        // `d m_k / d w_j = -F / 2 |d|` and `d m_k / d p_k = sum_j M / |d|` ( the facet moves along `d = p_j - p_k` ), so
        // `grad_w_k = sum_j ( g_k - g_j ) F / 2 |d|` and `grad_p_k = sum_j ( g_k - g_j ) M / |d|`, each from the cell's own facets
        const double gk = P.out.vjp_g[ k ];
        double gw = 0, gx = 0, gy = 0;
        for ( int e = 0; e < nb; ++e ) {
            const int id = c.I( e );
            if ( id < 0 ) continue;
            const int f = e + 1 < nb ? e + 1 : 0;
            const double ax = c.X( e ), ay = c.Y( e ), bx = c.X( f ), by = c.Y( f );
            const double ex = bx - ax, ey = by - ay, L = sqrt( ex * ex + ey * ey );
            double F = 0, Mx = 0, My = 0;
            if ( P.dn.kind == 1 ) {
                grid_walk( g, ax, ay, bx, by, [&]( int i, int j, double t0, double t1 ) {
                    const double v = g.val( i, j ) * L, dt = t1 - t0, h = ( t1 * t1 - t0 * t0 ) / 2;
                    F += v * dt;
                    Mx += v * ( ax * dt + ex * h );
                    My += v * ( ay * dt + ey * h );
                } );
            } else {
                F = P.dn.rho * L;
                Mx = F * ( ax + bx ) / 2;
                My = F * ( ay + by ) / 2;
            }
            const double4 s = P.wk.seeds[ id ];
            const double dx = s.x - p.x, dy = s.y - p.y, dd = sqrt( dx * dx + dy * dy );
            if ( ! ( dd > 0 ) ) continue;
            const double dg = gk - P.out.vjp_g[ id ];
            gw += dg * F / ( 2 * dd );
            gx += dg * Mx / dd;
            gy += dg * My / dd;
        }
        P.out.grad_w[ k ] = gw;
        P.out.grad_pos[ 2 * k ] = gx;
        P.out.grad_pos[ 2 * k + 1 ] = gy;
        return;
    }
    if ( P.mode == MODE_MOMENTS ) {
        double m = 0, mx = 0, my = 0, m2 = 0;
        if ( nb > 0 ) {
            if ( P.dn.kind == 1 )      image_moments( g, c, nb, m, mx, my, m2 );
            else if ( P.dn.kind == 2 ) gauss_cell( P.dn, p.x, p.y, c, nb, true, m, mx, my, m2 );
            else                       const_moments( c, nb, P.dn.rho, m, mx, my, m2 );
        }
        P.out.bary[ 2 * k ] = m > 0 ? p.x + mx / m : p.x;
        P.out.bary[ 2 * k + 1 ] = m > 0 ? p.y + my / m : p.y;
        P.out.cost[ k ] = m2;
        P.out.mass[ k ] = m;
        return;
    }
    if ( P.mode != MODE_FACETS ) {
        double mass = 0;
        if ( nb > 0 ) {
            if ( P.dn.kind == 1 ) {
                for ( int e = 0; e < nb; ++e ) { double mt, ln; image_edge( g, c, nb, e, mt, ln ); mass += mt; }
            } else if ( P.dn.kind == 2 ) {
                double u0, u1, u2;
                gauss_cell( P.dn, p.x, p.y, c, nb, false, mass, u0, u1, u2 );
            } else
                mass = P.dn.rho * shoelace( c, nb );
        }
        P.out.mass[ k ] = mass;
        return;
    }
    if ( P.out.edges ) {                                 // the cuts in polygon order, for the step by the limits
        int ne = nb < 3 ? 0 : nb;
        if ( ne > EDGE_CAP ) ne = -1;
        else
            for ( int e = 0; e < ne; ++e ) P.out.edges[ size_t( e ) * size_t( P.n ) + size_t( k ) ] = c.I( e );
        P.out.nb_edges[ k ] = ne;
    }
    const I64 at = I64( base ) + lane;
    const bool fits = I64( base ) + I64( ROW_STEP ) * wmax <= P.out.facet_cap;
    P.out.row_off[ k ] = int( fits ? at : 0 );
    P.out.row_cnt[ k ] = int( cnt );
    double dia = 0, mass = 0;
    int q = 0;
    for ( int e = 0; e < nb; ++e ) {
        const int id = c.I( e );
        double fm = 0;
        if ( P.dn.kind == 1 ) {
            double mt;
            image_edge( g, c, nb, e, mt, fm );
            mass += mt;
        } else if ( P.dn.kind == 2 ) {
            if ( id >= 0 ) {
                const int f = e + 1 < nb ? e + 1 : 0;
                for ( int u = 0; u < P.dn.ng; ++u ) fm += gauss_segment_line( gauss_term( P.dn, u, p.x, p.y ), c.X( e ), c.Y( e ), c.X( f ), c.Y( f ) );
            }
        } else if ( id >= 0 ) {
            const int f = e + 1 < nb ? e + 1 : 0;
            const double ex = c.X( f ) - c.X( e ), ey = c.Y( f ) - c.Y( e );
            fm = P.dn.rho * sqrt( ex * ex + ey * ey );
        }
        if ( id < 0 ) continue;
        const double4 s = P.wk.seeds[ id ];
        const double dx = s.x - p.x, dy = s.y - p.y;
        const double d2 = dx * dx + dy * dy;
        const double cij = d2 > 0 ? fm / ( 2 * sqrt( d2 ) ) : 0.0;
        if ( fits ) {
            P.out.col[ at + I64( q ) * ROW_STEP ] = id;
            P.out.val[ at + I64( q ) * ROW_STEP ] = cij;
        }
        dia += cij;
        ++q;
    }
    if ( P.dn.kind == 0 && nb > 0 ) mass = P.dn.rho * shoelace( c, nb );
    if ( P.dn.kind == 2 && nb > 0 ) { double u0, u1, u2; gauss_cell( P.dn, p.x, p.y, c, nb, false, mass, u0, u1, u2 ); }
    P.out.mass[ k ] = mass;
    P.out.dia[ k ] = dia > 0 ? dia : 1.0;
}

__device__ __forceinline__ void flush_events( const Params &P, const Ev &ev ) {
    const unsigned long long a = wave_sum64( ev.side ), b = wave_sum64( ev.prune ), c = wave_sum64( ev.newv );
    const unsigned long long d = wave_sum64( ev.visits ), e = wave_sum64( ev.cuts );
    if ( ( threadIdx.x & 63 ) == 0 ) {
        atomicAdd( &P.out.events[ 0 ], a );
        atomicAdd( &P.out.events[ 1 ], b );
        atomicAdd( &P.out.events[ 2 ], c );
        atomicAdd( &P.out.events[ 3 ], d );
        atomicAdd( &P.out.events[ 4 ], e );
    }
}

/// the rank of the calling lane in a first pass: the blocks of one XCD take a contiguous range of ranks ( blocks are
/// dealt round-robin to the 8 XCDs, each with its own L2 )
__device__ __forceinline__ int pass1_rank() {
    const int nblk = int( gridDim.x ), b = int( blockIdx.x );
    const int per = nblk / 8, rem = nblk % 8, xcd = b % 8, slot = b / 8;
    const int lb = xcd < rem ? xcd * ( per + 1 ) + slot : rem * ( per + 1 ) + ( xcd - rem ) * per + slot;
    return lb * CELL_BLOCK + int( threadIdx.x );
}

/// the plane of a cut in double, from its identifier ( `gpu/Cell2D.cuh:946-971` ): the bisector with seed `id`,
/// `n . v = ( ( w0 - w ) + |n|^2 ) / 2` in the seed's frame ( `e = |n|^2` ), side `-1 - id` of the box ( 0 bottom,
/// 1 right, 2 top, 3 left; `e = 1` ), or edge `-5 - id` of the polygon domain ( `e = 1` )
__device__ __forceinline__ void plane_of( const Params &P, int id, double p0x, double p0y, double w0, double &nx, double &ny, double &off,
                                          double &e ) {
    if ( id >= 0 ) {
        const double4 q = P.wk.seeds[ id ];
        nx = q.x - p0x;
        ny = q.y - p0y;
        e = nx * nx + ny * ny;
        off = ( ( w0 - q.z ) + e ) / 2;
        return;
    }
    if ( id <= -5 ) {                                    // an edge of the polygon domain
        const Polygon2 &D = P.dn.dom;
        const int j = -5 - id;
        nx = dom_at( D.nx, j );
        ny = dom_at( D.ny, j );
        e = 1.0;
        off = dom_at( D.off, j ) - nx * p0x - ny * p0y;
        return;
    }
    const int f = -1 - id;
    const bool vert = f & 1;
    nx = vert ? 1.0 : 0.0;
    ny = vert ? 0.0 : 1.0;
    e = 1.0;
    off = f == 0 ? P.lo[ 1 ] - p0y : f == 1 ? P.hi[ 0 ] - p0x : f == 2 ? P.hi[ 1 ] - p0y : P.lo[ 0 ] - p0x;
}

} // namespace

/// the first pass: one lane per cell, `CAP1` vertex slots in LDS
__global__ __launch_bounds__( CELL_BLOCK ) void k_cells_pass1( Params P ) {
    __shared__ double sx[ CAP1 * CELL_BLOCK ], sy[ CAP1 * CELL_BLOCK ];
    __shared__ int sid[ CAP1 * CELL_BLOCK ];
    const int k = pass1_rank();
    const Store<CELL_BLOCK> c{ sx + threadIdx.x, sy + threadIdx.x, sid + threadIdx.x };
    Ev ev;
    int nb = 0;
    bool active = k < P.n;
    if ( active ) {
        nb = make_cell( P.wk, k, P.lo, P.hi, P.dn.dom, c, CAP1, ev );
        if ( nb < 0 ) {
            const unsigned at = atomicAdd( &P.out.counters[ 1 ], 1u );
            P.out.ovf_list[ at ] = k;
            active = false;
        }
    }
    finish_cell( P, k, c, nb, active, true );
    flush_events( P, ev );
}

/// the first pass of the FP32 kernel: the walk and the cuts in float ( the seed's frame; each plane computed in double,
/// then rounded ), `CAPW` float slots in LDS; the polygon ( float vertices, the cut of each edge ) to `po`, slot `i` of
/// rank `k` at `i n + k`, and its vertex count ( -1: listed for the second pass, which runs the FP64 kernel )
__global__ __launch_bounds__( CELL_BLOCK ) void k_cells_walk32( Params P, Poly32 po ) {
    __shared__ float sx[ CAPW * CELL_BLOCK ], sy[ CAPW * CELL_BLOCK ];
    __shared__ int sid[ CAPW * CELL_BLOCK ];
    const int k = pass1_rank();
    const Store<CELL_BLOCK, float> c{ sx + threadIdx.x, sy + threadIdx.x, sid + threadIdx.x };
    Ev ev;
    if ( k < P.n ) {
        const int nb = make_cell<Store<CELL_BLOCK, float>, NodeF>( P.wk, k, P.lo, P.hi, P.dn.dom, c, CAPW, ev );
        if ( nb < 0 ) {
            const unsigned at = atomicAdd( &P.out.counters[ 1 ], 1u );
            P.out.ovf_list[ at ] = k;
        }
        po.nb[ k ] = nb;
        for ( int i = 0; i < nb; ++i ) {
            const size_t at = size_t( i ) * size_t( P.n ) + size_t( k );
            po.x[ at ] = c.X( i );
            po.y[ at ] = c.Y( i );
            po.id[ at ] = c.I( i );
        }
    }
    flush_events( P, ev );
}

/// the end of the FP32 kernel's first pass: each vertex re-solved in double from the planes of the edge that arrives and
/// the edge that leaves ( `gpu/Cell2D.cuh:973-1034` ). The solution is kept when the 2 x 2 system is well conditioned
/// ( `det^2 > 1e-12 |n_a|^2 |n_b|^2` ) and it lies within what the float vertex can be off by ( `|v - v_f|^2 det^2 <=
/// 1e-8 L^2 |n_a|^2 |n_b|^2`, `L` the farthest corner of the box from the seed ); else the float vertex projected on the
/// plane of the edge that leaves it. Then the cell's measures and facets in double, as the FP64 kernel's.
__global__ __launch_bounds__( CELL_BLOCK ) void k_cells_finish32( Params P, Poly32 po ) {
    __shared__ double sx[ CAP1 * CELL_BLOCK ], sy[ CAP1 * CELL_BLOCK ];
    __shared__ int sid[ CAP1 * CELL_BLOCK ];
    const int k = pass1_rank();
    const Store<CELL_BLOCK> c{ sx + threadIdx.x, sy + threadIdx.x, sid + threadIdx.x };
    const bool live = k < P.n;
    const int nb = live ? po.nb[ k ] : 0;
    bool active = live && nb >= 0;
    if ( active && nb > 0 ) {
        const double4 p = P.wk.seeds[ k ];
        const double x0 = P.lo[ 0 ] - p.x, x1 = P.hi[ 0 ] - p.x, y0 = P.lo[ 1 ] - p.y, y1 = P.hi[ 1 ] - p.y;
        const double tol2 = 1e-8 * ( fmax( x0 * x0, x1 * x1 ) + fmax( y0 * y0, y1 * y1 ) );
        const size_t n = size_t( P.n );
        bool bad = false;
        double ax, ay, ao, ea;
        plane_of( P, po.id[ size_t( nb - 1 ) * n + size_t( k ) ], p.x, p.y, p.z, ax, ay, ao, ea );
        for ( int i = 0; i < nb; ++i ) {
            const size_t at = size_t( i ) * n + size_t( k );
            const double x = po.x[ at ], y = po.y[ at ];
            const int id = po.id[ at ];
            double bx, by, bo, eb;
            plane_of( P, id, p.x, p.y, p.z, bx, by, bo, eb );
            const double det = ax * by - ay * bx, det2 = det * det, e12 = ea * eb;
            double vx = 0, vy = 0;
            bool ok = det2 > 1e-12 * e12;
            if ( ok ) {
                vx = ( ao * by - bo * ay ) / det;
                vy = ( ax * bo - bx * ao ) / det;
                ok = ( ( vx - x ) * ( vx - x ) + ( vy - y ) * ( vy - y ) ) * det2 <= tol2 * e12;
            }
            bad |= ! ok;
            c.X( i ) = vx;
            c.Y( i ) = vy;
            c.I( i ) = id;
            ax = bx; ay = by; ao = bo; ea = eb;
        }
        if ( bad ) {                                     // THE CONDITIONING TEST failed: the cell is redone by the FP64 kernel
            const unsigned at = atomicAdd( &P.out.counters[ 1 ], 1u );
            P.out.ovf_list[ at ] = k;
            active = false;
        }
    }
    finish_cell( P, k, c, nb, active, true );
}

/// the FP32 rows checked against each other: a facet `( k, j )` that only one of its two cells has ( a topology the floats
/// decided differently: a vertex of one cell within the float's error of the other's plane ) lists BOTH cells, once, for the
/// FP64 kernel. `mark` holds the stamp of the diagram that listed a rank last
__global__ void k_sym_check( int n, const int *row_off, const int *row_cnt, const int *col, unsigned *mark, unsigned stamp, int *list,
                             unsigned *count, const unsigned *counters, I64 cap ) {
    const int k = int( blockIdx.x ) * int( blockDim.x ) + int( threadIdx.x );
    if ( k >= n || I64( counters[ 0 ] ) > cap ) return;
    const int ok = row_off[ k ], ck = row_cnt[ k ];
    for ( int q = 0; q < ck; ++q ) {
        const int j = col[ ok + q * ROW_STEP ];
        const int oj = row_off[ j ], cj = row_cnt[ j ];
        bool found = false;
        for ( int r = 0; r < cj && ! found; ++r ) found = col[ oj + r * ROW_STEP ] == k;
        if ( found ) continue;
        if ( atomicExch( &mark[ k ], stamp ) != stamp ) list[ atomicAdd( count, 1u ) ] = k;
        if ( atomicExch( &mark[ j ], stamp ) != stamp ) list[ atomicAdd( count, 1u ) ] = j;
    }
}

/// the cells `k_sym_check` listed, redone by the FP64 kernel ( one lane per cell, `CAP2` slots in LDS ); each writes a new row
/// and its measures over the FP32 ones. A cell past `CAP2` keeps its FP32 result. The list is short ( hundreds of cells ) and
/// dealt one cell per wave before any wave takes a second: a wave's time is then one walk, not its lanes' divergent walks
/// one after the other
__global__ __launch_bounds__( 64 ) void k_cells_redo( Params P, const int *list, const unsigned *count_p ) {
    __shared__ double sx[ CAP2 * 64 ], sy[ CAP2 * 64 ];
    __shared__ int sid[ CAP2 * 64 ];
    const Store<64> c{ sx + threadIdx.x, sy + threadIdx.x, sid + threadIdx.x };
    const unsigned count = *count_p;
    const unsigned lanes = gridDim.x * 64;
    const unsigned rounds = ( count + lanes - 1 ) / lanes;
    Ev ev;
    for ( unsigned r = 0; r < rounds; ++r ) {
        const unsigned idx = r * lanes + threadIdx.x * gridDim.x + blockIdx.x;   // consecutive cells to different waves
        const bool live = idx < count;
        const int k = live ? list[ idx ] : 0;
        int nb = 0;
        bool active = live;
        if ( live ) {
            nb = make_cell( P.wk, k, P.lo, P.hi, P.dn.dom, c, CAP2, ev );
            active = nb >= 0;
        }
        finish_cell( P, k, c, nb, active, false );
    }
    flush_events( P, ev );
}

/// the second pass: the cells that outgrew `CAP1`, one lane per cell with `CAP2` slots in LDS, one wave per
/// block and a fixed grid striding over the list ( its length only the device knows ), dealt one cell per wave first; the cells that
/// outgrow `CAP2` too are listed again for the third pass
__global__ __launch_bounds__( 64 ) void k_cells_pass2( Params P ) {
    __shared__ double sx[ CAP2 * 64 ], sy[ CAP2 * 64 ];
    __shared__ int sid[ CAP2 * 64 ];
    const Store<64> c{ sx + threadIdx.x, sy + threadIdx.x, sid + threadIdx.x };
    const unsigned count = P.out.counters[ 1 ];
    const unsigned lanes = gridDim.x * 64;
    const unsigned rounds = ( count + lanes - 1 ) / lanes;
    Ev ev;
    for ( unsigned r = 0; r < rounds; ++r ) {
        const unsigned idx = r * lanes + threadIdx.x * gridDim.x + blockIdx.x;   // consecutive cells to different waves
        const bool live = idx < count;
        const int k = live ? P.out.ovf_list[ idx ] : 0;
        int nb = 0;
        bool active = live;
        if ( live ) {
            nb = make_cell( P.wk, k, P.lo, P.hi, P.dn.dom, c, CAP2, ev );
            if ( nb < 0 ) {
                const unsigned at = atomicAdd( &P.out.counters[ 4 ], 1u );
                P.out.ovf_list[ P.n + int( at ) ] = k;              // ( the second half: the first holds this pass's input )
                active = false;
            }
        }
        finish_cell( P, k, c, nb, active, false );
    }
    flush_events( P, ev );
}

/// the third pass: `CAP3` slots per cell in global memory, `slots` lanes ( the grid ); past that a cell is a failure
__global__ __launch_bounds__( 64 ) void k_cells_pass3( Params P, double *gx, double *gy, int *gid, unsigned slots ) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const Store<1> c{ gx + size_t( slot ) * CAP3, gy + size_t( slot ) * CAP3, gid + size_t( slot ) * CAP3 };
    const unsigned count = P.out.counters[ 4 ];
    const unsigned rounds = ( count + slots - 1 ) / slots;
    Ev ev;
    for ( unsigned r = 0; r < rounds; ++r ) {
        const unsigned idx = r * slots + unsigned( slot );
        const bool live = idx < count;
        const int k = live ? P.out.ovf_list[ P.n + int( idx ) ] : 0;
        int nb = 0;
        bool active = live;
        if ( live ) {
            nb = make_cell( P.wk, k, P.lo, P.hi, P.dn.dom, c, CAP3, ev );
            if ( nb < 0 ) {
                if ( atomicAdd( &P.out.counters[ 2 ], 1u ) == 0 ) P.out.counters[ 3 ] = unsigned( k );
                active = false;
                if ( P.out.mass ) P.out.mass[ k ] = 0;
                if ( P.mode == MODE_FACETS ) {
                    P.out.row_cnt[ k ] = 0; P.out.row_off[ k ] = 0; P.out.dia[ k ] = 1;
                    if ( P.out.edges ) P.out.nb_edges[ k ] = -1;
                }
                if ( P.mode == MODE_CELLS ) P.out.cell_nb[ k ] = 0;
                if ( P.mode == MODE_VJP ) { P.out.grad_w[ k ] = 0; P.out.grad_pos[ 2 * k ] = 0; P.out.grad_pos[ 2 * k + 1 ] = 0; }
            }
        }
        finish_cell( P, k, c, nb, active, false );
    }
    flush_events( P, ev );
}

/// the FP32 kernel's rows made symmetric, as the reference card mirrors its upper view into both rows
/// ( `gpu/Laplacian2D.cuh:139-232` ): facet `( k, j )` takes the value row `min( k, j )` gives it, and a facet that only one
/// of its two cells has ( a topology the floats decided differently: a facet within the float's error, tiny ) is dropped;
/// `dia` is the row's sum again. The value a lane reads in another row is one its owner keeps, so the rows are updated in
/// place. Nothing is done when the rows did not fit ( the diagram is then redone )
__global__ void k_sym_rows( int n, const int *row_off, const int *row_cnt, const int *col, double *val, double *dia, const unsigned *counters,
                            I64 cap, unsigned *dropped ) {
    const int k = int( blockIdx.x ) * int( blockDim.x ) + int( threadIdx.x );
    if ( k >= n || I64( counters[ 0 ] ) > cap ) return;
    const int ok = row_off[ k ], ck = row_cnt[ k ];
    double sum = 0;
    for ( int q = 0; q < ck; ++q ) {
        const int at = ok + q * ROW_STEP;
        const int j = col[ at ];
        const double own = val[ at ];
        const int oj = row_off[ j ], cj = row_cnt[ j ];
        double mirror = -1;
        for ( int r = 0; r < cj; ++r )
            if ( col[ oj + r * ROW_STEP ] == k ) { mirror = val[ oj + r * ROW_STEP ]; break; }
        const double c = mirror < 0 ? 0.0 : ( j < k ? mirror : own );
        if ( c != own ) val[ at ] = c;
        if ( mirror < 0 ) atomicAdd( dropped, 1u );
        sum += c;
    }
    dia[ k ] = sum > 0 ? sum : 1.0;
}

__global__ void k_reset_counters( unsigned *c ) {
    if ( threadIdx.x < 8 ) c[ threadIdx.x ] = 0;
}

__global__ void k_fill_u32( I64 n, unsigned *v, unsigned x ) {
    const I64 k = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( k < n ) v[ k ] = x;
}

void cells_scratch( DiagramOut &out, I64 n, const std::string &tag, hipStream_t st ) {
    Arena &A = arena();
    out.counters = A.get<unsigned>( tag + ".counters", 8 );
    out.ovf_list = A.get<int>( tag + ".ovf_list", 2 * size_t( n ) );
    if ( ! out.events ) {                                // ( the counters run on from solve to solve: zeroed when made, and by the
        bool fresh = false;                              // timing's reset )
        out.events = A.get<unsigned long long>( "events", 8, &fresh );
        if ( fresh ) SDOT_HIP( hipMemsetAsync( out.events, 0, 8 * sizeof( unsigned long long ), st ) );
    }
}

void set_domain( Density &dn, int nb, const double *xy ) {
    SDOT_CHECK( nb >= 3 && nb <= DOMAIN_MAX && xy, "domain: a convex polygon of 3 to 12 vertices" );
    Polygon2 &D = dn.dom;
    D.nb = nb;
    double area = 0;
    for ( int i = 0; i < nb; ++i ) {
        D.x[ i ] = xy[ 2 * i ];
        D.y[ i ] = xy[ 2 * i + 1 ];
        SDOT_CHECK( std::isfinite( D.x[ i ] ) && std::isfinite( D.y[ i ] ), "domain: a vertex is not finite" );
    }
    for ( int i = 0; i < nb; ++i ) {
        const int j = ( i + 1 ) % nb, l = ( i + 2 ) % nb;
        const double ex = D.x[ j ] - D.x[ i ], ey = D.y[ j ] - D.y[ i ];
        const double fx = D.x[ l ] - D.x[ j ], fy = D.y[ l ] - D.y[ j ];
        const double len = std::sqrt( ex * ex + ey * ey );
        SDOT_CHECK( len > 0, "domain: two consecutive vertices coincide" );
        SDOT_CHECK( ex * fy - ey * fx > 0, "domain: the vertices are not those of a convex polygon, counterclockwise" );
        D.nx[ i ] = ey / len;                            // ( the outward normal of a counterclockwise edge )
        D.ny[ i ] = -ex / len;
        D.off[ i ] = D.nx[ i ] * D.x[ i ] + D.ny[ i ] * D.y[ i ];
        area += D.x[ i ] * D.y[ j ] - D.x[ j ] * D.y[ i ];
    }
    SDOT_CHECK( area > 0, "domain: the polygon is not counterclockwise" );
}

void run_cells( const DevTree &t, const double *lo, const double *hi, const Density &dn, int mode, DiagramOut &out, hipStream_t st, bool fp32,
                int ovf_slots ) {
    SDOT_CHECK( ovf_slots >= 64 && ovf_slots % 64 == 0, "cells: the third pass's lanes are a positive multiple of 64" );
    SDOT_CHECK( ! fp32 || t.nodes32, "cells: the FP32 walk needs the tree's float records" );
    Params P;
    P.wk.nodes = t.nodes;
    P.wk.nodes32 = t.nodes32;
    P.wk.node_end = t.node_end;
    P.wk.seeds = t.seeds;
    P.wk.depth = t.depth;
    P.n = int( t.n );
    P.lo[ 0 ] = lo[ 0 ]; P.lo[ 1 ] = lo[ 1 ]; P.hi[ 0 ] = hi[ 0 ]; P.hi[ 1 ] = hi[ 1 ];
    P.dn = dn;
    P.out = out;
    P.mode = mode;
    k_reset_counters<<<1, 64, 0, st>>>( out.counters );
    check_launch( "k_reset_counters" );
    Timing::Pending tp;
    bool timed = timing().begin( TK_CELLS, st, tp );
    Arena &A = arena();
    if ( fp32 ) {
        Poly32 po;
        po.x = A.get<float>( "p32.x", size_t( CAPW ) * size_t( t.n ) );
        po.y = A.get<float>( "p32.y", size_t( CAPW ) * size_t( t.n ) );
        po.id = A.get<int>( "p32.id", size_t( CAPW ) * size_t( t.n ) );
        po.nb = A.get<int>( "p32.nb", size_t( t.n ) );
        k_cells_walk32<<<blocks_for( t.n, CELL_BLOCK ), CELL_BLOCK, 0, st>>>( P, po );
        check_launch( "k_cells_walk32" );
        k_cells_finish32<<<blocks_for( t.n, CELL_BLOCK ), CELL_BLOCK, 0, st>>>( P, po );
        check_launch( "k_cells_finish32" );
    } else {
        k_cells_pass1<<<blocks_for( t.n, CELL_BLOCK ), CELL_BLOCK, 0, st>>>( P );
        check_launch( "k_cells_pass1" );
    }
    if ( timed ) timing().end( tp, st );
    timed = timing().begin( TK_CELLS_OVF, st, tp );
    k_cells_pass2<<<PASS2_BLOCKS, 64, 0, st>>>( P );
    check_launch( "k_cells_pass2" );
    double *gx = A.get<double>( "ovf.x", size_t( ovf_slots ) * CAP3 );
    double *gy = A.get<double>( "ovf.y", size_t( ovf_slots ) * CAP3 );
    int *gid = A.get<int>( "ovf.id", size_t( ovf_slots ) * CAP3 );
    k_cells_pass3<<<unsigned( ovf_slots ) / 64, 64, 0, st>>>( P, gx, gy, gid, unsigned( ovf_slots ) );
    check_launch( "k_cells_pass3" );
    if ( fp32 && mode == MODE_FACETS ) {
        // the cells whose facets the floats decided differently from their neighbours', redone by the FP64 kernel; then the rows
        // made symmetric, a facet still one-sided dropped ( counted )
        static std::atomic<unsigned> stamps{ 0 };
        unsigned stamp = ++stamps;
        if ( stamp == 0 ) stamp = ++stamps;
        bool fresh = false;
        unsigned *mark = A.get<unsigned>( "cells.mark", size_t( t.n ), &fresh );
        int *redo = A.get<int>( "cells.redo", size_t( t.n ) );
        if ( fresh ) {
            k_fill_u32<<<blocks_for( t.n, 256 ), 256, 0, st>>>( t.n, mark, 0u );
            check_launch( "k_fill_u32" );
        }
        k_sym_check<<<blocks_for( t.n, 256 ), 256, 0, st>>>( int( t.n ), out.row_off, out.row_cnt, out.col, mark, stamp, redo, &out.counters[ 6 ],
                                                            out.counters, out.facet_cap );
        check_launch( "k_sym_check" );
        k_cells_redo<<<REDO_BLOCKS, 64, 0, st>>>( P, redo, &out.counters[ 6 ] );
        check_launch( "k_cells_redo" );
        k_sym_rows<<<blocks_for( t.n, 256 ), 256, 0, st>>>( int( t.n ), out.row_off, out.row_cnt, out.col, out.val, out.dia, out.counters,
                                                           out.facet_cap, &out.counters[ 7 ] );
        check_launch( "k_sym_rows" );
    }
    if ( timed ) timing().end( tp, st );
}

} // namespace sdot_amd
