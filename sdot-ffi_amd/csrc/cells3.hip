// The 3D power cells: ONE WAVE PER CELL, one lane per vertex of the cell ( a simple polytope: every vertex on three faces
// with three neighbours, the neighbour across slot `j` at the other end of the edge that leaves face `j` ).
//
// Reference semantics ( paths under `src/sdot/include/sdot/` ): the cut `cell/LocalN.h:320-420` ( inside vertices kept,
// one new vertex per crossing edge interpolated from its inside end, the new face closed by the faces the new vertices
// share; a vertex within rounding of the plane stays inside ), the 3D card cell `gpu/Cell3D.cuh:417-730` ( there a
// register cell per lane; here the lanes of a wave share one cell, so a cut is a ballot, a scan and an LDS exchange, and
// the walk's pruning test is one vertex per lane ), the walk and pruning `cell/Providers.h:167-236` and
// `cell/Pruning.h:46-76` in three coordinates, facets `c_kj = int_face rho / ( 2 |p_j - p_k| )` ( `Laplacian.h:40-72` ).
// Sums of Gaussians: a cell's mass is the flux through its faces of a field whose divergence is the density, so the masses
// of a diagram sum to the box's whatever the weights; face integrals by the 7-point degree-5 rule on the triangles of a
// face's mean vertex, moments by the 14-point degree-5 rule on the tetrahedra of those triangles and the cell's mean vertex.

#include "cells3.h"

#pragma clang fp contract( off )

namespace sdot_amd {

namespace {

constexpr int WAVES3 = 4;                    ///< waves ( cells ) per block
constexpr int STACK3 = 64;

struct Params3 {
    const NodeRec3 *nodes;
    const double4  *seeds;
    int             depth, n;
    double          lo[ 3 ], hi[ 3 ];
    Density3        dn;
    DiagramOut      out;
    int             mode, room;
};

/// what a wave keeps in LDS: its cell's faces, the staging of a cut, the walk's stack
struct WaveLds {
    double fx[ CELL3_FACES ], fy[ CELL3_FACES ], fz[ CELL3_FACES ], fo[ CELL3_FACES ];
    int    fid[ CELL3_FACES ];
    double sx[ CELL3_VERTICES ], sy[ CELL3_VERTICES ], sz[ CELL3_VERTICES ];
    int    spl[ CELL3_VERTICES ], snb[ CELL3_VERTICES ];
    int    np1[ CELL3_VERTICES ], np2[ CELL3_VERTICES ];
    int    stk_m[ STACK3 ], stk_h[ STACK3 ];
    int    used[ CELL3_FACES ];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
}

/// the sum over the wave, the same bits in every lane
__device__ __forceinline__ double wave_total( double v ) {
    for ( int o = 32; o > 0; o >>= 1 ) v += __shfl_down( v, o, 64 );
    return __shfl( v, 0, 64 );
}

__device__ __forceinline__ double wave_maxd( double v ) {
    for ( int o = 32; o > 0; o >>= 1 ) v = fmax( v, __shfl_down( v, o, 64 ) );
    return __shfl( v, 0, 64 );
}

__device__ __forceinline__ int byte_of( int packed, int j ) { return ( packed >> ( 8 * j ) ) & 255; }

// ---- the density -------------------------------------------------------------------------------------------------

__device__ __forceinline__ double gauss_peak( double m, double s ) {
    return m / ( 2.5066282746310002 * 2.5066282746310002 * 2.5066282746310002 * s * s * s );
}

__device__ double rho_at( const Density3 &dn, double x, double y, double z ) {
    if ( dn.kind == 0 ) return dn.rho;
    double r = 0;
    for ( int q = 0; q < dn.ng; ++q ) {
        const double *g = dn.gauss + 5 * q;
        const double s = g[ 3 ], dx = x - g[ 0 ], dy = y - g[ 1 ], dz = z - g[ 2 ];
        r += gauss_peak( g[ 4 ], s ) * exp( -( dx * dx + dy * dy + dz * dz ) / ( 2 * s * s ) );
    }
    return r;
}

/// `int_0^1 exp( -u t^2 ) t^2 dt`: the series below 1, else `( sqrt( pi ) erf( sqrt u ) / 4 - sqrt( u ) exp( -u ) / 2 ) / u^1.5`
__device__ double radial_h( double u ) {
    if ( u < 1 ) {
        double term = 1, r = 1.0 / 3;
        for ( int k = 1; k < 40; ++k ) {
            term *= -u / k;
            const double c = term / ( 2 * k + 3 );
            r += c;
            if ( fabs( c ) < 1e-18 * r ) break;
        }
        return r;
    }
    const double su = sqrt( u );
    return ( 0.44311346272637900 * erf( su ) - 0.5 * su * exp( -u ) ) / ( u * su );
}

/// the field of divergence `rho`: per Gaussian `peak ( x - mu ) h( |x - mu|^2 / ( 2 s^2 ) )`, dotted with `n`
__device__ double flux_dot( const Density3 &dn, double x, double y, double z, double nx, double ny, double nz ) {
    double fx = 0, fy = 0, fz = 0;
    for ( int q = 0; q < dn.ng; ++q ) {
        const double *g = dn.gauss + 5 * q;
        const double s = g[ 3 ], dx = x - g[ 0 ], dy = y - g[ 1 ], dz = z - g[ 2 ];
        const double f = gauss_peak( g[ 4 ], s ) * radial_h( ( dx * dx + dy * dy + dz * dz ) / ( 2 * s * s ) );
        fx += f * dx; fy += f * dy; fz += f * dz;
    }
    return fx * nx + fy * ny + fz * nz;
}

// ---- what a face's edge contributes ----------------------------------------------------------------------------------

struct V3 { double x, y, z; };

constexpr int BOX_SUB = 4;                    ///< a box face's triangles: on `BOX_SUB x BOX_SUB` sub-triangles

/// the sums a cell accumulates over the triangles ( face mean `g`, edge `v o` ) of its faces
struct CellSums {
    double part = 0;                 ///< `int rho` over the current face's triangles ( its facet )
    double mass = 0, mx = 0, my = 0, mz = 0, m2 = 0;
};

/// the triangle `( g, v, o )` of a face of outward unit normal `u`, and the tetrahedron it makes with the cell's mean
/// vertex `A` ( seed frame; the seed at `p0` ): the facet's `int rho`, the mass ( a constant: the tetrahedron's volume; a
/// sum of Gaussians: the flux through the triangle ), the moments ( the tetrahedron's )
__device__ __forceinline__ void edge_terms( const Density3 &dn, const double4 &p0, bool facets, bool moments, bool box_face, V3 A, V3 g, V3 v,
                                           V3 o, V3 u, CellSums &S ) {
    const double ux = v.x - g.x, uy = v.y - g.y, uz = v.z - g.z, tx = o.x - g.x, ty = o.y - g.y, tz = o.z - g.z;
    const double cx = uy * tz - uz * ty, cy = uz * tx - ux * tz, cz = ux * ty - uy * tx;
    const double ar = 0.5 * sqrt( cx * cx + cy * cy + cz * cz );
    const double bx = g.x - A.x, by = g.y - A.y, bz = g.z - A.z, ex = v.x - A.x, ey = v.y - A.y, ez = v.z - A.z;
    const double hx = o.x - A.x, hy = o.y - A.y, hz = o.z - A.z;
    const double det = bx * ( ey * hz - ez * hy ) - by * ( ex * hz - ez * hx ) + bz * ( ex * hy - ey * hx );
    const double vol = fabs( det ) / 6;
    if ( dn.kind != 2 ) {
        if ( facets ) S.part += dn.rho * ar;
        if ( ! moments ) { S.mass += dn.rho * vol; return; }
        const double w = dn.rho * vol;
        const double sx = A.x + g.x + v.x + o.x, sy = A.y + g.y + v.y + o.y, sz = A.z + g.z + v.z + o.z;
        const double sq = A.x * A.x + A.y * A.y + A.z * A.z + g.x * g.x + g.y * g.y + g.z * g.z +
                          v.x * v.x + v.y * v.y + v.z * v.z + o.x * o.x + o.y * o.y + o.z * o.z;
        S.mass += w;
        S.mx += w * sx / 4; S.my += w * sy / 4; S.mz += w * sz / 4;
        S.m2 += w * ( sq + ( sx * sx + sy * sy + sz * sz ) ) / 20;
        return;
    }
    if ( ! moments ) {                                   // the 7-point rule: the facet's density and the mass's flux
        const double s15 = sqrt( 15.0 );
        const double ta1 = ( 6 - s15 ) / 21, tw1 = ( 155 - s15 ) / 1200, ta2 = ( 6 + s15 ) / 21, tw2 = ( 155 + s15 ) / 1200;
        // a box face: its flux on a `BOX_SUB x BOX_SUB` grid of sub-triangles, so that the masses' total ( the box's flux )
        // hardly depends on the weights; an inner face: the triangle itself
        const int sub = box_face ? BOX_SUB : 1;
        const double sar = ar / ( sub * sub );
        for ( int i = 0; i < sub; ++i )
            for ( int jg = 0; i + jg < sub; ++jg )
                for ( int down = 0; down < ( i + jg + 1 < sub ? 2 : 1 ); ++down ) {
                    const int i0 = down ? i + 1 : i, j0 = jg, i1 = down ? i + 1 : i + 1, j1 = down ? jg + 1 : jg, i2 = i, j2 = jg + 1;
                    auto at = [&]( int a, int b, double &x, double &y, double &z ) {
                        x = g.x + ( v.x - g.x ) * a / sub + ( o.x - g.x ) * b / sub;
                        y = g.y + ( v.y - g.y ) * a / sub + ( o.y - g.y ) * b / sub;
                        z = g.z + ( v.z - g.z ) * a / sub + ( o.z - g.z ) * b / sub;
                    };
                    V3 P0, P1, P2;
                    at( i0, j0, P0.x, P0.y, P0.z );
                    at( i1, j1, P1.x, P1.y, P1.z );
                    at( i2, j2, P2.x, P2.y, P2.z );
                    for ( int qq = 0; qq < 7; ++qq ) {
                        double b0, b1, b2, wq;
                        if ( qq == 0 ) { b0 = b1 = b2 = 1.0 / 3; wq = 9.0 / 40; }
                        else {
                            const double a = qq <= 3 ? ta1 : ta2;
                            wq = qq <= 3 ? tw1 : tw2;
                            const int jj = ( qq - 1 ) % 3;
                            b0 = jj == 0 ? 1 - 2 * a : a; b1 = jj == 1 ? 1 - 2 * a : a; b2 = jj == 2 ? 1 - 2 * a : a;
                        }
                        const double X = p0.x + ( b0 * P0.x + b1 * P1.x + b2 * P2.x ), Y = p0.y + ( b0 * P0.y + b1 * P1.y + b2 * P2.y );
                        const double Z = p0.z + ( b0 * P0.z + b1 * P1.z + b2 * P2.z );
                        if ( facets ) S.part += wq * sar * rho_at( dn, X, Y, Z );
                        S.mass += wq * sar * flux_dot( dn, X, Y, Z, u.x, u.y, u.z );
                    }
                }
        return;
    }
    // the 14-point rule on the tetrahedron ( A, g, v, o )
    const double r1 = 0.0927352503108912264, r1w = 0.0734930431163619495;
    const double r2 = 0.3108859192633006098, r2w = 0.1126879257180158507;
    const double r3 = 0.0455037041256496494, r3w = 0.0425460207770814664;
    for ( int qq = 0; qq < 14; ++qq ) {
        double b0, b1, b2, b3, wq;
        if ( qq < 8 ) {
            const double a = qq < 4 ? r1 : r2;
            wq = qq < 4 ? r1w : r2w;
            const int jj = qq & 3;
            b0 = jj == 0 ? 1 - 3 * a : a; b1 = jj == 1 ? 1 - 3 * a : a; b2 = jj == 2 ? 1 - 3 * a : a; b3 = jj == 3 ? 1 - 3 * a : a;
        } else {
            const int pr = qq - 8;                       // pairs ( 0 1 ) ( 0 2 ) ( 0 3 ) ( 1 2 ) ( 1 3 ) ( 2 3 )
            const int i0 = pr < 3 ? 0 : pr < 5 ? 1 : 2, i1 = pr < 3 ? pr + 1 : pr < 5 ? pr - 1 : 3;
            wq = r3w;
            b0 = ( i0 == 0 || i1 == 0 ) ? r3 : 0.5 - r3; b1 = ( i0 == 1 || i1 == 1 ) ? r3 : 0.5 - r3;
            b2 = ( i0 == 2 || i1 == 2 ) ? r3 : 0.5 - r3; b3 = ( i0 == 3 || i1 == 3 ) ? r3 : 0.5 - r3;
        }
        const double x = b0 * A.x + b1 * g.x + b2 * v.x + b3 * o.x;
        const double y = b0 * A.y + b1 * g.y + b2 * v.y + b3 * o.y;
        const double z = b0 * A.z + b1 * g.z + b2 * v.z + b3 * o.z;
        const double w = wq * vol * rho_at( dn, p0.x + x, p0.y + y, p0.z + z );
        S.mass += w;
        S.mx += w * x; S.my += w * y; S.mz += w * z;
        S.m2 += w * ( x * x + y * y + z * z );
    }
}

/// the outputs of a measured cell ( `mass` and moments already totalled )
__device__ void write_cell( const Params3 &P, int k0, const double4 &p0, const CellSums &S, double dia, int q, I64 row0 ) {
    P.out.mass[ k0 ] = S.mass;
    if ( P.mode == MODE_MOMENTS ) {
        P.out.bary[ 3 * k0 ] = S.mass > 0 ? p0.x + S.mx / S.mass : p0.x;
        P.out.bary[ 3 * k0 + 1 ] = S.mass > 0 ? p0.y + S.my / S.mass : p0.y;
        P.out.bary[ 3 * k0 + 2 ] = S.mass > 0 ? p0.z + S.mz / S.mass : p0.z;
        P.out.cost[ k0 ] = S.m2;
    }
    if ( P.mode == MODE_FACETS ) {
        P.out.row_off[ k0 ] = int( row0 );
        P.out.row_cnt[ k0 ] = q <= P.room ? q : 0;
        P.out.dia[ k0 ] = dia > 0 ? dia : 1.0;
        atomicMax( &P.out.counters[ 0 ], unsigned( q ) );
    }
}

/// a cell the passes could not build: `code` 1 past the slots' vertices or faces, 2 a cut whose new face could not be
/// closed ( the reference's error codes, `gpu/Cell2D.cuh:99-102` ); the first failure's rank and code are kept
__device__ void record_failure( const Params3 &P, int k0, I64 row0, int code ) {
    if ( atomicAdd( &P.out.counters[ 2 ], 1u ) == 0 ) { P.out.counters[ 3 ] = unsigned( k0 ); P.out.counters[ 5 ] = unsigned( code ); }
    P.out.mass[ k0 ] = 0;
    if ( P.mode == MODE_FACETS ) { P.out.row_off[ k0 ] = int( row0 ); P.out.row_cnt[ k0 ] = 0; P.out.dia[ k0 ] = 1; }
}

// ---- the cut -------------------------------------------------------------------------------------------------------

enum { C3_UNCHANGED = 0, C3_CUT = 1, C3_EMPTY = 2, C3_FAILED = 3, C3_FULL = 4 };

struct Cell3 {
    double x, y, z;            ///< this lane's vertex ( seed frame )
    int    pl, nb;             ///< its three faces and neighbours, a byte each
    int    nv, nf;             ///< ( uniform ) vertex and face counts
};

/// faces without a vertex dropped, the others renumbered in order ( when the face slots are full )
__device__ void compact_faces( WaveLds &L, Cell3 &c, int lane ) {
    wave_sync();
    L.used[ lane ] = 0;
    wave_sync();
    if ( lane < c.nv ) { L.used[ byte_of( c.pl, 0 ) ] = 1; L.used[ byte_of( c.pl, 1 ) ] = 1; L.used[ byte_of( c.pl, 2 ) ] = 1; }
    wave_sync();
    const bool u = lane < c.nf && L.used[ lane ] != 0;
    const unsigned long long mu = __ballot( u );
    const int to = __popcll( mu & ( ( 1ull << lane ) - 1 ) );
    double fx = 0, fy = 0, fz = 0, fo = 0;
    int fid = 0;
    if ( u ) { fx = L.fx[ lane ]; fy = L.fy[ lane ]; fz = L.fz[ lane ]; fo = L.fo[ lane ]; fid = L.fid[ lane ]; }
    wave_sync();
    L.used[ lane ] = to;                                 // ( now the new index of face `lane` )
    if ( u ) { L.fx[ to ] = fx; L.fy[ to ] = fy; L.fz[ to ] = fz; L.fo[ to ] = fo; L.fid[ to ] = fid; }
    wave_sync();
    if ( lane < c.nv ) c.pl = L.used[ byte_of( c.pl, 0 ) ] | ( L.used[ byte_of( c.pl, 1 ) ] << 8 ) | ( L.used[ byte_of( c.pl, 2 ) ] << 16 );
    c.nf = __popcll( mu );
}

/// the cut by `n . v <= off` ( face id `id` ), the oracle's `cut3` with the lanes of a wave
__device__ int cut_cell( WaveLds &L, Cell3 &c, int lane, double nx, double ny, double nz, double off, int id, unsigned long long &ev_side,
                         unsigned long long &ev_new ) {
    const bool act = lane < c.nv;
    double s = act ? fma( nx, c.x, fma( ny, c.y, fma( nz, c.z, -off ) ) ) : -1.0;
    ev_side += unsigned( c.nv );
    if ( __ballot( s > 0 ) == 0 ) return C3_UNCHANGED;
    // a vertex on the plane up to the rounding of `s` stays inside ( where more than three planes meet )
    const double vm = wave_maxd( act ? fmax( fabs( c.x ), fmax( fabs( c.y ), fabs( c.z ) ) ) : 0.0 );
    const double tie = 1e-13 * ( ( fabs( nx ) + fabs( ny ) + fabs( nz ) ) * vm + fabs( off ) );
    const bool out = act && s > tie;
    if ( ! out ) s = fmin( s, 0.0 );
    const unsigned long long mo = __ballot( out );
    if ( mo == 0 ) return C3_UNCHANGED;
    const unsigned long long mact = c.nv >= 64 ? ~0ull : ( ( 1ull << c.nv ) - 1 );
    const unsigned long long mi = mact & ~mo;
    if ( mi == 0 ) return C3_EMPTY;
    if ( c.nf >= CELL3_FACES ) compact_faces( L, c, lane );
    if ( c.nf >= CELL3_FACES ) return C3_FULL;
    const int n0 = byte_of( c.nb, 0 ), n1 = byte_of( c.nb, 1 ), n2 = byte_of( c.nb, 2 );
    const int cnt = out ? int( ( mi >> n0 ) & 1 ) + int( ( mi >> n1 ) & 1 ) + int( ( mi >> n2 ) & 1 ) : 0;
    int inc = cnt;
    for ( int d = 1; d < 64; d <<= 1 ) {
        const int t = __shfl_up( inc, d, 64 );
        if ( lane >= d ) inc += t;
    }
    const int excl = inc - cnt, T = __shfl( inc, 63, 64 ), nin = __popcll( mi );
    if ( nin + T > CELL3_VERTICES ) return C3_FULL;
    const int fn = c.nf;
    wave_sync();
    if ( lane == 0 ) { L.fx[ fn ] = nx; L.fy[ fn ] = ny; L.fz[ fn ] = nz; L.fo[ fn ] = off; L.fid[ fn ] = id; }
    // the neighbours' positions, signs, neighbour lists and scan offsets ( every lane shuffles: no lane is masked )
    double ax[ 3 ], ay[ 3 ], az[ 3 ], as[ 3 ];
    int bn[ 3 ], bex[ 3 ];
    const int nk[ 3 ] = { n0, n1, n2 };
#pragma unroll
    for ( int k = 0; k < 3; ++k ) {
        ax[ k ] = __shfl( c.x, nk[ k ], 64 ); ay[ k ] = __shfl( c.y, nk[ k ], 64 ); az[ k ] = __shfl( c.z, nk[ k ], 64 );
        as[ k ] = __shfl( s, nk[ k ], 64 );
        bn[ k ] = __shfl( c.nb, nk[ k ], 64 );
        bex[ k ] = __shfl( excl, nk[ k ], 64 );
    }
    if ( out ) {                                         // the new vertices of this outside vertex's crossing edges
        int j = 0;
#pragma unroll
        for ( int k = 0; k < 3; ++k ) {
            const int a = nk[ k ];
            if ( ! ( ( mi >> a ) & 1 ) ) continue;
            const double sa = as[ k ], t = sa / ( sa - s );
            const int g = nin + excl + j++;
            L.sx[ g ] = fma( c.x - ax[ k ], t, ax[ k ] );
            L.sy[ g ] = fma( c.y - ay[ k ], t, ay[ k ] );
            L.sz[ g ] = fma( c.z - az[ k ], t, az[ k ] );
            const int p1 = byte_of( c.pl, ( k + 1 ) % 3 ), p2 = byte_of( c.pl, ( k + 2 ) % 3 );
            L.spl[ g ] = fn | ( p1 << 8 ) | ( p2 << 16 );
            L.snb[ g ] = __popcll( mi & ( ( 1ull << a ) - 1 ) );
            L.np1[ g - nin ] = p1;
            L.np2[ g - nin ] = p2;
        }
    } else if ( act ) {                                  // an inside vertex: its slot and its neighbours renumbered
        const int ia = __popcll( mi & ( ( 1ull << lane ) - 1 ) );
        int nbn = 0;
#pragma unroll
        for ( int k = 0; k < 3; ++k ) {
            const int m = nk[ k ];
            int to;
            if ( ( mi >> m ) & 1 ) to = __popcll( mi & ( ( 1ull << m ) - 1 ) );
            else {                                       // the new vertex on the edge to `m`: its rank among `m`'s crossings
                int j = 0;
                to = -1;
#pragma unroll
                for ( int q = 0; q < 3; ++q ) {
                    const int cq = byte_of( bn[ k ], q );
                    if ( to < 0 && cq == lane ) to = nin + bex[ k ] + j;
                    else if ( to < 0 && ( ( mi >> cq ) & 1 ) ) ++j;
                }
            }
            nbn |= ( to & 255 ) << ( 8 * k );
        }
        L.sx[ ia ] = c.x; L.sy[ ia ] = c.y; L.sz[ ia ] = c.z;
        L.spl[ ia ] = c.pl;
        L.snb[ ia ] = nbn;
    }
    wave_sync();
    // the new face closed: a new vertex's neighbour leaving `p1` is the other new vertex on `p2`, and conversely
    bool broken = false;
    int nbv = 0;
    if ( lane < nin + T ) {
        nbv = L.snb[ lane ];
        if ( lane >= nin ) {
            const int t0 = lane - nin, p1 = L.np1[ t0 ], p2 = L.np2[ t0 ];
            int m1 = 0, m2 = 0, c1 = 0, c2 = 0;
            for ( int t = 0; t < T; ++t ) {
                if ( t == t0 ) continue;
                const int q1 = L.np1[ t ], q2 = L.np2[ t ];
                if ( q1 == p2 || q2 == p2 ) { m1 = nin + t; ++c1; }
                if ( q1 == p1 || q2 == p1 ) { m2 = nin + t; ++c2; }
            }
            broken = c1 != 1 || c2 != 1;
            nbv = ( nbv & 255 ) | ( m1 << 8 ) | ( m2 << 16 );
        }
    }
    if ( __ballot( broken ) ) return C3_FAILED;
    if ( lane < nin + T ) { c.x = L.sx[ lane ]; c.y = L.sy[ lane ]; c.z = L.sz[ lane ]; c.pl = L.spl[ lane ]; c.nb = nbv; }
    else { c.pl = 0; c.nb = 0; }
    c.nv = nin + T;
    c.nf = fn + 1;
    ev_new += unsigned( T );
    return C3_CUT;
}

// ---- the cells -----------------------------------------------------------------------------------------------------

__device__ __forceinline__ double proximity3( const NodeRec3 &r, double px, double py, double pz ) {
    const double p[ 3 ] = { px, py, pz };
    double s = 0;
#pragma unroll
    for ( int d = 0; d < 3; ++d ) {
        double e = fmax( r.lo[ d ] - p[ d ], p[ d ] - r.hi[ d ] );
        e = e > 0 ? e : 0;
        s += e * e;
    }
    return s;
}

__global__ __launch_bounds__( 64 * WAVES3 ) void k_cells3( Params3 P ) {
    __shared__ WaveLds lds[ WAVES3 ];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    WaveLds &L = lds[ wid ];
    const int k0 = int( blockIdx.x ) * WAVES3 + wid;
    if ( k0 >= P.n ) return;                             // ( a whole wave: every lane of it leaves )
    const double4 p0 = P.seeds[ k0 ];
    const double w0 = p0.w;
    unsigned long long ev_side = 0, ev_prune = 0, ev_new = 0, ev_visit = 0, ev_cut = 0;

    // the domain box, seen from the seed: vertex `v` at the corner of bits ( x, y, z ), on faces `2 d + bit_d`
    Cell3 c;
    c.nv = 8; c.nf = 6;
    {
        const int v = lane & 7;
        c.x = ( v & 1 ? P.hi[ 0 ] : P.lo[ 0 ] ) - p0.x;
        c.y = ( v & 2 ? P.hi[ 1 ] : P.lo[ 1 ] ) - p0.y;
        c.z = ( v & 4 ? P.hi[ 2 ] : P.lo[ 2 ] ) - p0.z;
        c.pl = ( 0 + ( v & 1 ) ) | ( ( 2 + ( ( v >> 1 ) & 1 ) ) << 8 ) | ( ( 4 + ( ( v >> 2 ) & 1 ) ) << 16 );
        c.nb = ( v ^ 1 ) | ( ( v ^ 2 ) << 8 ) | ( ( v ^ 4 ) << 16 );
        if ( lane < 6 ) {
            const int d = lane >> 1, hi = lane & 1;
            const double pc = d == 0 ? p0.x : d == 1 ? p0.y : p0.z;
            L.fx[ lane ] = d == 0 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            L.fy[ lane ] = d == 1 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            L.fz[ lane ] = d == 2 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            L.fo[ lane ] = hi ? P.hi[ d ] - pc : pc - P.lo[ d ];
            L.fid[ lane ] = hi ? -2 - 2 * d : -1 - 2 * d;
        }
    }
    int status = C3_CUT;
    // the walk: nearest child first, a node pruned when no vertex can be cut by a seed under its majorant
    int top = 0;
    if ( lane == 0 ) { L.stk_m[ 0 ] = 0; L.stk_h[ 0 ] = P.depth; }
    top = 1;
    wave_sync();
    while ( top > 0 && status == C3_CUT ) {
        --top;
        const int m = __builtin_amdgcn_readfirstlane( L.stk_m[ top ] ), h = __builtin_amdgcn_readfirstlane( L.stk_h[ top ] );
        const NodeRec3 r = P.nodes[ m ];
        if ( r.beg >= r.end ) continue;
        ++ev_visit;
        ev_prune += unsigned( c.nv );
        bool may = false;
        if ( lane < c.nv ) {
            const double v[ 3 ] = { c.x, c.y, c.z }, pp[ 3 ] = { p0.x, p0.y, p0.z };
            double sv = w0 - r.wb;
#pragma unroll
            for ( int d = 0; d < 3; ++d ) {
                const double X = v[ d ] + pp[ d ];
                double y = X + r.wa[ d ] / 2;
                y = y < r.lo[ d ] ? r.lo[ d ] : ( y > r.hi[ d ] ? r.hi[ d ] : y );
                const double u = y - X;
                sv += u * u - v[ d ] * v[ d ] - r.wa[ d ] * y;
            }
            may = sv <= 0;
        }
        if ( __ballot( may ) == 0 ) continue;
        if ( h <= 1 ) {
            for ( int k = r.beg; k < r.end; ++k ) {
                if ( k == k0 ) continue;
                const double4 q = P.seeds[ k ];
                const double dx = q.x - p0.x, dy = q.y - p0.y, dz = q.z - p0.z;
                const double off = ( ( w0 - q.w ) + ( dx * dx + dy * dy + dz * dz ) ) / 2;
                const int res = cut_cell( L, c, lane, dx, dy, dz, off, k, ev_side, ev_new );
                if ( res == C3_CUT ) ++ev_cut;
                if ( res == C3_EMPTY || res == C3_FAILED || res == C3_FULL ) { status = res; break; }
            }
            continue;
        }
        const int lc = m + 1, rc = m + ( 1 << ( h - 1 ) );
        const double dl = proximity3( P.nodes[ lc ], p0.x, p0.y, p0.z ), dr = proximity3( P.nodes[ rc ], p0.x, p0.y, p0.z );
        if ( top + 2 > STACK3 ) { status = C3_FAILED; break; }
        wave_sync();
        if ( lane == 0 ) {
            const int first = dl <= dr ? rc : lc, second = dl <= dr ? lc : rc;
            L.stk_m[ top ] = first; L.stk_h[ top ] = h - 1;
            L.stk_m[ top + 1 ] = second; L.stk_h[ top + 1 ] = h - 1;
        }
        top += 2;
        wave_sync();
    }
    if ( lane == 0 ) {
        atomicAdd( &P.out.events[ 0 ], ev_side );
        atomicAdd( &P.out.events[ 1 ], ev_prune );
        atomicAdd( &P.out.events[ 2 ], ev_new );
        atomicAdd( &P.out.events[ 3 ], ev_visit );
        atomicAdd( &P.out.events[ 4 ], ev_cut );
    }
    const I64 row0 = I64( k0 / 64 ) * 64 * P.room + ( k0 & 63 );
    if ( status == C3_FAILED ) {
        if ( lane == 0 ) record_failure( P, k0, row0, 2 );
        return;
    }
    if ( status == C3_FULL ) {                           // past the lanes: the second pass, one thread in global slots
        if ( lane == 0 ) P.out.ovf_list[ atomicAdd( &P.out.counters[ 1 ], 1u ) ] = k0;
        return;
    }
    if ( status == C3_EMPTY ) c.nv = 0;

    // ---- the measures: face by face ( the faces in cut order, as the oracle lists the facets )
    const bool act = lane < c.nv;
    const double inv_nv = c.nv > 0 ? 1.0 / c.nv : 0.0;
    const double Ax = wave_total( act ? c.x : 0.0 ) * inv_nv, Ay = wave_total( act ? c.y : 0.0 ) * inv_nv, Az = wave_total( act ? c.z : 0.0 ) * inv_nv;
    double wx[ 3 ], wy[ 3 ], wz[ 3 ];
    int nk[ 3 ];
#pragma unroll
    for ( int k = 0; k < 3; ++k ) {
        nk[ k ] = byte_of( c.nb, k );
        wx[ k ] = __shfl( c.x, nk[ k ], 64 ); wy[ k ] = __shfl( c.y, nk[ k ], 64 ); wz[ k ] = __shfl( c.z, nk[ k ], 64 );
    }
    const bool facets = P.mode == MODE_FACETS, moments = P.mode == MODE_MOMENTS;
    const V3 Av{ Ax, Ay, Az };
    CellSums S;
    double dia = 0;
    int q = 0;
    for ( int f = 0; f < ( c.nv > 0 ? c.nf : 0 ); ++f ) {
        int j = -1;
        if ( act ) j = byte_of( c.pl, 0 ) == f ? 0 : byte_of( c.pl, 1 ) == f ? 1 : byte_of( c.pl, 2 ) == f ? 2 : -1;
        const int cntf = __popcll( __ballot( j >= 0 ) );
        if ( cntf < 3 ) continue;
        const V3 g{ wave_total( j >= 0 ? c.x : 0.0 ) / cntf, wave_total( j >= 0 ? c.y : 0.0 ) / cntf, wave_total( j >= 0 ? c.z : 0.0 ) / cntf };
        const double fnx = L.fx[ f ], fny = L.fy[ f ], fnz = L.fz[ f ];
        const int fid = L.fid[ f ];
        const double dn = sqrt( fnx * fnx + fny * fny + fnz * fnz );
        const V3 u{ fnx / dn, fny / dn, fnz / dn };
        S.part = 0;
        if ( j >= 0 ) {
            const V3 v{ c.x, c.y, c.z };
#pragma unroll
            for ( int e = 1; e <= 2; ++e ) {             // the two edges of the face at this vertex, each counted by its lower lane
                const int ks = ( j + e ) % 3;           // ( selects, not an index: no scratch )
                const int w = ks == 0 ? nk[ 0 ] : ks == 1 ? nk[ 1 ] : nk[ 2 ];
                if ( ! ( lane < w ) ) continue;
                const V3 o{ ks == 0 ? wx[ 0 ] : ks == 1 ? wx[ 1 ] : wx[ 2 ], ks == 0 ? wy[ 0 ] : ks == 1 ? wy[ 1 ] : wy[ 2 ],
                            ks == 0 ? wz[ 0 ] : ks == 1 ? wz[ 1 ] : wz[ 2 ] };
                edge_terms( P.dn, p0, facets, moments, fid < 0, Av, g, v, o, u, S );
            }
        }
        if ( facets && fid >= 0 ) {
            const double cf = wave_total( S.part ) / ( 2 * dn );
            if ( lane == 0 && q < P.room ) {
                P.out.col[ row0 + I64( q ) * 64 ] = fid;
                P.out.val[ row0 + I64( q ) * 64 ] = cf;
            }
            dia += cf;
            ++q;
        }
    }
    S.mass = wave_total( S.mass );
    if ( moments ) { S.mx = wave_total( S.mx ); S.my = wave_total( S.my ); S.mz = wave_total( S.mz ); S.m2 = wave_total( S.m2 ); }
    if ( lane == 0 ) write_cell( P, k0, p0, S, dia, q, row0 );
}

// ---- the second pass: cells past the lanes, one thread each in global slots ----------------------------------------

constexpr int SLOT_V = CELL3_SLOT_VERTICES, SLOT_F = CELL3_SLOT_FACES;   ///< a slot's vertices and faces
constexpr int PASS2_THREADS = 2048;

/// one thread's slot: its cell ( vertices with their faces and neighbours, faces ) and a cut's staging
struct Slot3 {
    double x[ SLOT_V ], y[ SLOT_V ], z[ SLOT_V ], s[ SLOT_V ];
    int    pl[ 3 * SLOT_V ], nb[ 3 * SLOT_V ];
    double X[ SLOT_V ], Y[ SLOT_V ], Z[ SLOT_V ];
    int    PL[ 3 * SLOT_V ], NB[ 3 * SLOT_V ], idx[ SLOT_V ], np1[ SLOT_V ], np2[ SLOT_V ];
    double fx[ SLOT_F ], fy[ SLOT_F ], fz[ SLOT_F ], fo[ SLOT_F ];
    int    fid[ SLOT_F ];
    int    stk_m[ STACK3 ], stk_h[ STACK3 ];
    int    nv, nf;
};

/// the cut of `cut_cell`, one thread: the same classification, interpolation and numbering ( the oracle's `cut3` )
__device__ int cut_slot( Slot3 &C, double nx, double ny, double nz, double off, int id, unsigned long long &ev_side, unsigned long long &ev_new ) {
    const int nv = C.nv;
    ev_side += unsigned( nv );
    bool any = false;
    double vm = 0;
    for ( int v = 0; v < nv; ++v ) {
        C.s[ v ] = fma( nx, C.x[ v ], fma( ny, C.y[ v ], fma( nz, C.z[ v ], -off ) ) );
        any |= C.s[ v ] > 0;
        vm = fmax( vm, fmax( fabs( C.x[ v ] ), fmax( fabs( C.y[ v ] ), fabs( C.z[ v ] ) ) ) );
    }
    if ( ! any ) return C3_UNCHANGED;
    const double tie = 1e-13 * ( ( fabs( nx ) + fabs( ny ) + fabs( nz ) ) * vm + fabs( off ) );
    int nin = 0;
    for ( int v = 0; v < nv; ++v ) {
        if ( ! ( C.s[ v ] > tie ) ) { C.s[ v ] = fmin( C.s[ v ], 0.0 ); C.idx[ v ] = nin++; }
        else C.idx[ v ] = -1;
    }
    if ( nin == nv ) return C3_UNCHANGED;
    if ( nin == 0 ) return C3_EMPTY;
    if ( C.nf >= SLOT_F ) return C3_FULL;
    const int fn = C.nf;
    C.fx[ fn ] = nx; C.fy[ fn ] = ny; C.fz[ fn ] = nz; C.fo[ fn ] = off; C.fid[ fn ] = id;
    for ( int v = 0; v < nv; ++v ) {                    // the inside vertices, their neighbours renumbered
        const int i = C.idx[ v ];
        if ( i < 0 ) continue;
        C.X[ i ] = C.x[ v ]; C.Y[ i ] = C.y[ v ]; C.Z[ i ] = C.z[ v ];
        for ( int j = 0; j < 3; ++j ) { C.PL[ 3 * i + j ] = C.pl[ 3 * v + j ]; C.NB[ 3 * i + j ] = C.idx[ C.nb[ 3 * v + j ] ]; }
    }
    int T = 0;
    for ( int b = 0; b < nv; ++b ) {                    // a new vertex per crossing edge, in the order ( outside vertex, edge )
        if ( C.idx[ b ] >= 0 ) continue;
        for ( int k = 0; k < 3; ++k ) {
            const int a = C.nb[ 3 * b + k ];
            if ( C.idx[ a ] < 0 ) continue;
            const int g = nin + T;
            if ( g >= SLOT_V ) return C3_FULL;
            const double sa = C.s[ a ], t = sa / ( sa - C.s[ b ] );
            C.X[ g ] = fma( C.x[ b ] - C.x[ a ], t, C.x[ a ] );
            C.Y[ g ] = fma( C.y[ b ] - C.y[ a ], t, C.y[ a ] );
            C.Z[ g ] = fma( C.z[ b ] - C.z[ a ], t, C.z[ a ] );
            const int p1 = C.pl[ 3 * b + ( k + 1 ) % 3 ], p2 = C.pl[ 3 * b + ( k + 2 ) % 3 ];
            C.PL[ 3 * g ] = fn; C.PL[ 3 * g + 1 ] = p1; C.PL[ 3 * g + 2 ] = p2;
            C.NB[ 3 * g ] = C.idx[ a ];
            C.np1[ T ] = p1; C.np2[ T ] = p2;
            const int ia = C.idx[ a ];
            for ( int j = 0; j < 3; ++j )
                if ( C.nb[ 3 * a + j ] == b ) C.NB[ 3 * ia + j ] = g;
            ++T;
        }
    }
    for ( int t0 = 0; t0 < T; ++t0 ) {                  // the new face closed
        const int p1 = C.np1[ t0 ], p2 = C.np2[ t0 ];
        int m1 = 0, m2 = 0, c1 = 0, c2 = 0;
        for ( int t = 0; t < T; ++t ) {
            if ( t == t0 ) continue;
            if ( C.np1[ t ] == p2 || C.np2[ t ] == p2 ) { m1 = nin + t; ++c1; }
            if ( C.np1[ t ] == p1 || C.np2[ t ] == p1 ) { m2 = nin + t; ++c2; }
        }
        if ( c1 != 1 || c2 != 1 ) return C3_FAILED;
        C.NB[ 3 * ( nin + t0 ) + 1 ] = m1;
        C.NB[ 3 * ( nin + t0 ) + 2 ] = m2;
    }
    const int m = nin + T;
    for ( int v = 0; v < m; ++v ) {
        C.x[ v ] = C.X[ v ]; C.y[ v ] = C.Y[ v ]; C.z[ v ] = C.Z[ v ];
        for ( int j = 0; j < 3; ++j ) { C.pl[ 3 * v + j ] = C.PL[ 3 * v + j ]; C.nb[ 3 * v + j ] = C.NB[ 3 * v + j ]; }
    }
    C.nv = m;
    C.nf = fn + 1;
    ev_new += unsigned( T );
    return C3_CUT;
}

__global__ __launch_bounds__( 64 ) void k_cells3_pass2( Params3 P, Slot3 *slots ) {
    const int tid = int( blockIdx.x ) * 64 + threadIdx.x;
    Slot3 &C = slots[ tid ];
    const int count = int( P.out.counters[ 1 ] );
    for ( int it = tid; it < count; it += PASS2_THREADS ) {
        const int k0 = P.out.ovf_list[ it ];
        const double4 p0 = P.seeds[ k0 ];
        const double w0 = p0.w;
        unsigned long long ev_side = 0, ev_new = 0;
        C.nv = 8; C.nf = 6;
        for ( int v = 0; v < 8; ++v ) {
            C.x[ v ] = ( v & 1 ? P.hi[ 0 ] : P.lo[ 0 ] ) - p0.x;
            C.y[ v ] = ( v & 2 ? P.hi[ 1 ] : P.lo[ 1 ] ) - p0.y;
            C.z[ v ] = ( v & 4 ? P.hi[ 2 ] : P.lo[ 2 ] ) - p0.z;
            for ( int d = 0; d < 3; ++d ) { C.pl[ 3 * v + d ] = 2 * d + ( ( v >> d ) & 1 ); C.nb[ 3 * v + d ] = v ^ ( 1 << d ); }
        }
        for ( int f = 0; f < 6; ++f ) {
            const int d = f >> 1, hi = f & 1;
            const double pc = d == 0 ? p0.x : d == 1 ? p0.y : p0.z;
            C.fx[ f ] = d == 0 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            C.fy[ f ] = d == 1 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            C.fz[ f ] = d == 2 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            C.fo[ f ] = hi ? P.hi[ d ] - pc : pc - P.lo[ d ];
            C.fid[ f ] = hi ? -2 - 2 * d : -1 - 2 * d;
        }
        int status = C3_CUT, top = 1;
        C.stk_m[ 0 ] = 0; C.stk_h[ 0 ] = P.depth;
        while ( top > 0 && status == C3_CUT ) {
            --top;
            const int m = C.stk_m[ top ], h = C.stk_h[ top ];
            const NodeRec3 r = P.nodes[ m ];
            if ( r.beg >= r.end ) continue;
            bool may = false;
            for ( int v = 0; v < C.nv && ! may; ++v ) {
                const double vv[ 3 ] = { C.x[ v ], C.y[ v ], C.z[ v ] }, pp[ 3 ] = { p0.x, p0.y, p0.z };
                double sv = w0 - r.wb;
                for ( int d = 0; d < 3; ++d ) {
                    const double X = vv[ d ] + pp[ d ];
                    double y = X + r.wa[ d ] / 2;
                    y = y < r.lo[ d ] ? r.lo[ d ] : ( y > r.hi[ d ] ? r.hi[ d ] : y );
                    const double u = y - X;
                    sv += u * u - vv[ d ] * vv[ d ] - r.wa[ d ] * y;
                }
                may = sv <= 0;
            }
            if ( ! may ) continue;
            if ( h <= 1 ) {
                for ( int k = r.beg; k < r.end; ++k ) {
                    if ( k == k0 ) continue;
                    const double4 q = P.seeds[ k ];
                    const double dx = q.x - p0.x, dy = q.y - p0.y, dz = q.z - p0.z;
                    const double off = ( ( w0 - q.w ) + ( dx * dx + dy * dy + dz * dz ) ) / 2;
                    const int res = cut_slot( C, dx, dy, dz, off, k, ev_side, ev_new );
                    if ( res == C3_EMPTY || res == C3_FAILED || res == C3_FULL ) { status = res; break; }
                }
                continue;
            }
            const int lc = m + 1, rc = m + ( 1 << ( h - 1 ) );
            const double dl = proximity3( P.nodes[ lc ], p0.x, p0.y, p0.z ), dr = proximity3( P.nodes[ rc ], p0.x, p0.y, p0.z );
            if ( top + 2 > STACK3 ) { status = C3_FAILED; break; }
            C.stk_m[ top ] = dl <= dr ? rc : lc; C.stk_h[ top ] = h - 1;
            C.stk_m[ top + 1 ] = dl <= dr ? lc : rc; C.stk_h[ top + 1 ] = h - 1;
            top += 2;
        }
        atomicAdd( &P.out.events[ 0 ], ev_side );
        atomicAdd( &P.out.events[ 2 ], ev_new );
        const I64 row0 = I64( k0 / 64 ) * 64 * P.room + ( k0 & 63 );
        if ( status == C3_FAILED || status == C3_FULL ) { record_failure( P, k0, row0, status == C3_FULL ? 1 : 2 ); continue; }
        if ( status == C3_EMPTY ) C.nv = 0;
        // the measures, face by face as the first pass
        const bool facets = P.mode == MODE_FACETS, moments = P.mode == MODE_MOMENTS;
        V3 Av{ 0, 0, 0 };
        for ( int v = 0; v < C.nv; ++v ) { Av.x += C.x[ v ]; Av.y += C.y[ v ]; Av.z += C.z[ v ]; }
        const double inv_nv = C.nv > 0 ? 1.0 / C.nv : 0.0;
        Av.x *= inv_nv; Av.y *= inv_nv; Av.z *= inv_nv;
        CellSums S;
        double dia = 0;
        int q = 0;
        for ( int f = 0; f < ( C.nv > 0 ? C.nf : 0 ); ++f ) {
            V3 g{ 0, 0, 0 };
            int cntf = 0;
            for ( int v = 0; v < C.nv; ++v )
                if ( C.pl[ 3 * v ] == f || C.pl[ 3 * v + 1 ] == f || C.pl[ 3 * v + 2 ] == f ) { g.x += C.x[ v ]; g.y += C.y[ v ]; g.z += C.z[ v ]; ++cntf; }
            if ( cntf < 3 ) continue;
            g.x /= cntf; g.y /= cntf; g.z /= cntf;
            const double dn = sqrt( C.fx[ f ] * C.fx[ f ] + C.fy[ f ] * C.fy[ f ] + C.fz[ f ] * C.fz[ f ] );
            const V3 u{ C.fx[ f ] / dn, C.fy[ f ] / dn, C.fz[ f ] / dn };
            S.part = 0;
            for ( int v = 0; v < C.nv; ++v ) {
                const int j = C.pl[ 3 * v ] == f ? 0 : C.pl[ 3 * v + 1 ] == f ? 1 : C.pl[ 3 * v + 2 ] == f ? 2 : -1;
                if ( j < 0 ) continue;
                const V3 vv{ C.x[ v ], C.y[ v ], C.z[ v ] };
                for ( int e = 1; e <= 2; ++e ) {
                    const int w = C.nb[ 3 * v + ( j + e ) % 3 ];
                    if ( ! ( v < w ) ) continue;
                    const V3 o{ C.x[ w ], C.y[ w ], C.z[ w ] };
                    edge_terms( P.dn, p0, facets, moments, C.fid[ f ] < 0, Av, g, vv, o, u, S );
                }
            }
            if ( facets && C.fid[ f ] >= 0 ) {
                const double cf = S.part / ( 2 * dn );
                if ( q < P.room ) {
                    P.out.col[ row0 + I64( q ) * 64 ] = C.fid[ f ];
                    P.out.val[ row0 + I64( q ) * 64 ] = cf;
                }
                dia += cf;
                ++q;
            }
        }
        write_cell( P, k0, p0, S, dia, q, row0 );
    }
}

__global__ void k_reset_counters3( unsigned *c ) {
    if ( threadIdx.x < 8 ) c[ threadIdx.x ] = 0;
}

} // namespace

void run_cells_3d( const DevTree3 &t, const double *lo, const double *hi, const Density3 &dn, int mode, DiagramOut &out, int room,
                   hipStream_t st ) {
    SDOT_CHECK( I64( room ) * ( ( t.n + 63 ) / 64 ) * 64 <= out.facet_cap || mode != MODE_FACETS, "3D rows: less room than the rows need" );
    SDOT_CHECK( I64( room ) * ( ( t.n + 63 ) / 64 ) * 64 < ( I64( 1 ) << 31 ), "3D rows: more entries than 32-bit offsets index" );
    Params3 P;
    P.nodes = t.nodes;
    P.seeds = t.seeds;
    P.depth = t.depth;
    P.n = int( t.n );
    for ( int d = 0; d < 3; ++d ) { P.lo[ d ] = lo[ d ]; P.hi[ d ] = hi[ d ]; }
    P.dn = dn;
    P.out = out;
    P.mode = mode;
    P.room = room;
    k_reset_counters3<<<1, 64, 0, st>>>( out.counters );
    check_launch( "k_reset_counters3" );
    Timing::Pending tp;
    const bool timed = timing().begin( TK_CELLS, st, tp );
    k_cells3<<<blocks_for( t.n, WAVES3 ), 64 * WAVES3, 0, st>>>( P );
    check_launch( "k_cells3" );
    if ( timed ) timing().end( tp, st );
    Slot3 *slots = reinterpret_cast<Slot3 *>( arena().get<char>( "cells3.slots", sizeof( Slot3 ) * size_t( PASS2_THREADS ) ) );
    k_cells3_pass2<<<PASS2_THREADS / 64, 64, 0, st>>>( P, slots );
    check_launch( "k_cells3_pass2" );
}

} // namespace sdot_amd
