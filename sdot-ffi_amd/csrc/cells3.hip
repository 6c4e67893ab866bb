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
// face's area centroid, moments by the 14-point degree-5 rule on the tetrahedra of those triangles and the cell's mean vertex.

#include "cells3.h"

#pragma clang fp contract( off )

namespace sdot_amd {

namespace {

constexpr int WAVES3 = 4;                    ///< waves ( cells ) per block
constexpr int STACK3 = 64;

/// the cells of the first pass kept for the measure kernel ( a sum of Gaussians ): per rank, 64 vertex and 64 face slots
/// ( `[ k ][ 64 ]`, a wave's lanes contiguous ), the vertex count ( -1: the cell is measured elsewhere ) and face count
struct PolyBuf {
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    int    *vpl = nullptr, *vnb = nullptr;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    int    *fid = nullptr;
    int    *nv = nullptr, *nf = nullptr;
};

struct Params3 {
    const NodeRec3 *nodes;
    const double4  *seeds;
    int             depth, n;
    double          lo[ 3 ], hi[ 3 ];
    Density3        dn;
    DiagramOut      out;
    int             mode, room;
    PolyBuf         pb;
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

/// lane `l`'s value ( `l` uniform ), in scalar registers
__device__ __forceinline__ double readlane_d( double v, int l ) {
    const long long b = __double_as_longlong( v );
    const int lo = __builtin_amdgcn_readlane( int( b ), l ), hi = __builtin_amdgcn_readlane( int( b >> 32 ), l );
    return __longlong_as_double( ( (long long) hi << 32 ) | (long long) (unsigned) lo );
}

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

/// `int_0^1 exp( -u t^2 ) t^2 dt`, `e = exp( -u )`: below 1 the series `sum_k ( -u )^k / ( k! ( 2 k + 3 ) )` to `k = 19` by
/// Horner ( the next term is below 3e-20 of the sum ), else `( sqrt( pi ) erf( sqrt u ) / 4 - sqrt( u ) e / 2 ) / u^1.5`
__device__ __forceinline__ double radial_h( double u, double e ) {
    if ( u < 1 ) {
        const double c[ 20 ] = { 0.3333333333333333, -0.2, 0.07142857142857142, -0.018518518518518517, 0.003787878787878788,
                                 -0.000641025641025641, 9.259259259259259e-05, -1.1671335200746965e-05, 1.3053467000835422e-06,
                                 -1.3122532963802806e-07, 1.1981443140863431e-08, -1.0020843354176688e-09, 7.732132217728926e-11,
                                 -5.537601323041936e-12, 3.7002405153966856e-13, -2.317322342975702e-14, 1.3655649521106815e-15,
                                 -7.598533119852758e-17, 4.0049248637400584e-18, -2.005032986981544e-19 };
        double r = c[ 19 ];
#pragma unroll
        for ( int k = 18; k >= 0; --k ) r = r * u + c[ k ];
        return r;
    }
    const double su = sqrt( u );
    return ( 0.44311346272637900 * erf( su ) - 0.5 * su * e ) / ( u * su );
}

// ---- what a face's edge contributes ----------------------------------------------------------------------------------

struct V3 { double x, y, z; };

constexpr int BOX_SUB = 4;                    ///< a box face's triangles: on `BOX_SUB x BOX_SUB` sub-triangles

/// the 7-point rule ( degree 5 ) on the triangle `( P0, P1, P2 )` ( seed frame, the seed at `p0` ) of area `ar` and unit
/// normal `n`, against a sum of Gaussians: adds `int rho` to `part` and the flux of `F` to `flux`, `F` the field of divergence
/// `rho` ( per Gaussian `peak ( x - mu ) h( |x - mu|^2 / ( 2 s^2 ) )` ). The Gaussians outside, so that each one's peak and
/// width divide once per triangle, and `exp( -u )` serves the density and the field
__device__ __forceinline__ void gauss_rule7( const Density3 &dn, const double4 &p0, V3 P0, V3 P1, V3 P2, double ar, V3 n, double &part,
                                             double &flux ) {
    const double s15 = sqrt( 15.0 );
    const double ta1 = ( 6 - s15 ) / 21, tw1 = ( 155 - s15 ) / 1200, ta2 = ( 6 + s15 ) / 21, tw2 = ( 155 + s15 ) / 1200;
    for ( int q = 0; q < dn.ng; ++q ) {
        const double *g = dn.gauss + 5 * q;
        const double s = g[ 3 ], peak = gauss_peak( g[ 4 ], s ), inv = 1 / ( 2 * s * s );
        const double mx = g[ 0 ] - p0.x, my = g[ 1 ] - p0.y, mz = g[ 2 ] - p0.z;
        double pr = 0, fl = 0;
#pragma unroll 1
        for ( int qq = 0; qq < 7; ++qq ) {
            double b0, b1, b2, wq;
            if ( qq == 0 ) { b0 = b1 = b2 = 1.0 / 3; wq = 9.0 / 40; }
            else {
                const double a = qq <= 3 ? ta1 : ta2;
                wq = qq <= 3 ? tw1 : tw2;
                const int jj = ( qq - 1 ) % 3;
                b0 = jj == 0 ? 1 - 2 * a : a; b1 = jj == 1 ? 1 - 2 * a : a; b2 = jj == 2 ? 1 - 2 * a : a;
            }
            const double dx = b0 * P0.x + b1 * P1.x + b2 * P2.x - mx, dy = b0 * P0.y + b1 * P1.y + b2 * P2.y - my;
            const double dz = b0 * P0.z + b1 * P1.z + b2 * P2.z - mz;
            const double u = ( dx * dx + dy * dy + dz * dz ) * inv, e = exp( -u );
            pr += wq * e;
            fl += wq * radial_h( u, e ) * ( dx * n.x + dy * n.y + dz * n.z );
        }
        part += ar * peak * pr;
        flux += ar * peak * fl;
    }
}

/// the sums a cell accumulates over the triangles ( face mean `g`, edge `v o` ) of its faces
struct CellSums {
    double part = 0;                 ///< `int rho` over the current face's triangles ( its facet )
    double mass = 0, mx = 0, my = 0, mz = 0, m2 = 0;
};

/// the area centroid of face `f` of a cell and the count of its vertices, in one walk over the vertices: the edges of `f` at
/// vertex `v`, `f` its `j`-th face ( `face_of( v, j )` ), lead to its neighbours `nb_of( v, j + 1 )` and `nb_of( v, j + 2 )`
/// ( mod 3 ), each edge taken from its lower vertex, and the triangles of those edges with the face's first vertex make the
/// convex face ( weighted by `| ( u x v ) . n |`, `n` the face's normal of any length: the areas up to a common factor, without
/// a square root ). It is the apex of the face's triangles for the rules: the mean of the vertices would jump when a vertex
/// appears or vanishes with an edge of zero length, and the rules' error with it ( ~`( h / sigma )^5` of a cell's mass for a
/// Gaussian, `h` the cell's size ), a jump of the masses in the weights that stalls a Newton solve asked for less than that
/// error; the centroid moves with the vertices
template<class FaceOf, class NbOf>
__device__ __forceinline__ int face_apex( int f, int nv, FaceOf face_of, NbOf nb_of, const double *px, const double *py, const double *pz,
                                          double nx, double ny, double nz, double *G ) {
    int cnt = 0, v0 = -1;
    double ox = 0, oy = 0, oz = 0, sa = 0, sx = 0, sy = 0, sz = 0;
    for ( int v = 0; v < nv; ++v )
        for ( int j = 0; j < 3; ++j ) {
            if ( face_of( v, j ) != f ) continue;
            ++cnt;
            if ( v0 < 0 ) { v0 = v; ox = px[ v ]; oy = py[ v ]; oz = pz[ v ]; }
            for ( int e = 1; e <= 2; ++e ) {
                const int w = nb_of( v, ( j + e ) % 3 );
                if ( ! ( v < w ) ) continue;
                const double ux = px[ v ] - ox, uy = py[ v ] - oy, uz = pz[ v ] - oz;
                const double vx = px[ w ] - ox, vy = py[ w ] - oy, vz = pz[ w ] - oz;
                const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
                const double a = fabs( cx * nx + cy * ny + cz * nz );
                sa += a; sx += a * ( ux + vx ); sy += a * ( uy + vy ); sz += a * ( uz + vz );
            }
        }
    G[ 0 ] = ox; G[ 1 ] = oy; G[ 2 ] = oz;
    if ( sa > 0 ) { G[ 0 ] += sx / ( 3 * sa ); G[ 1 ] += sy / ( 3 * sa ); G[ 2 ] += sz / ( 3 * sa ); }
    return cnt;
}

// ---- a voxel image: fluxes through the faces, cut by the voxel grid ------------------------------------------------------
//
// The divergence theorem with fields along x: `int_cell rho = flux of ( G, 0, 0 )`, `G( x, y, z ) = int_a^x rho( s, y, z ) ds`
// along the voxel row of `( y, z )` from a column `a` the same for the whole cell ( its mean vertex's: the terms stay of the
// cell's size ). In a voxel `G` is affine in `x`, so a face cut by the voxel grid integrates it exactly piece by piece. The
// moments are fluxes too: `( H, 0, 0 )`, `( y G, 0, 0 )`, `( z G, 0, 0 )`, `( K + ( y^2 + z^2 ) G, 0, 0 )` with
// `H = int_a^x s rho`, `K = int_a^x s^2 rho` ( degree 3 at most in a voxel: the 7-point degree-5 rule on each piece's fan ).
// The reference clips the cell by every voxel instead ( `Image.cxx:106-232`, in N-D ); the oracle does that.

constexpr int VPMAX = 12;                     ///< the vertices of a triangle's piece after six cuts ( 3 + 6 )
struct VPoly { int n = 0; double x[ VPMAX ], y[ VPMAX ], z[ VPMAX ]; };

/// the part of convex polygon `a` where coordinate `axis` is `< c` ( `below` ) or `>= c`: the slabs `[ c_i, c_i+1 )` share a
/// polygon out without a gap, and one lying on a plane goes to one slab only ( a box face on a grid plane, its flux counted once )
__device__ __forceinline__ void clip_axis( const VPoly &a, int axis, double c, bool below, VPoly &o ) {
    o.n = 0;
    for ( int q = 0; q < a.n; ++q ) {
        const int r = q + 1 < a.n ? q + 1 : 0;
        const double sq = ( axis == 0 ? a.x[ q ] : axis == 1 ? a.y[ q ] : a.z[ q ] ) - c;
        const double sr = ( axis == 0 ? a.x[ r ] : axis == 1 ? a.y[ r ] : a.z[ r ] ) - c;
        const bool iq = below ? sq < 0 : sq >= 0, ir = below ? sr < 0 : sr >= 0;
        if ( iq && o.n < VPMAX ) { o.x[ o.n ] = a.x[ q ]; o.y[ o.n ] = a.y[ q ]; o.z[ o.n ] = a.z[ q ]; ++o.n; }
        if ( iq != ir && o.n < VPMAX ) {
            const double t = sq / ( sq - sr );
            o.x[ o.n ] = a.x[ q ] + t * ( a.x[ r ] - a.x[ q ] );
            o.y[ o.n ] = a.y[ q ] + t * ( a.y[ r ] - a.y[ q ] );
            o.z[ o.n ] = a.z[ q ] + t * ( a.z[ r ] - a.z[ q ] );
            if ( axis == 0 ) o.x[ o.n ] = c; else if ( axis == 1 ) o.y[ o.n ] = c; else o.z[ o.n ] = c;
            ++o.n;
        }
    }
}

__device__ __forceinline__ int voxel_of( double v, double v0, double h, int n ) {
    const double c = floor( ( v - v0 ) / h );
    return c < 0 ? 0 : ( c >= double( n - 1 ) ? n - 1 : int( c ) );
}

/// the slabs `q0 .. q1` that a polygon spanning `[ lo, hi ]` along an axis may meet, decided by the comparisons `clip_axis`
/// makes ( `voxel_of`'s division rounds otherwise: a polygon on or next to a plane would miss the slab that takes it )
__device__ __forceinline__ void slabs_of( double lo, double hi, double v0, double h, int n, int &q0, int &q1 ) {
    q0 = voxel_of( lo, v0, h, n );
    q1 = voxel_of( hi, v0, h, n );
    while ( q0 > 0 && lo - ( v0 + q0 * h ) < 0 ) --q0;
    while ( q1 < n - 1 && hi - ( v0 + ( q1 + 1 ) * h ) >= 0 ) ++q1;
}

/// the triangle `( g, v, o )` of a face of outward unit normal `u` ( seed frame, the seed at `p0` ) against a voxel image:
/// the facet's `int rho`, the flux of `( G, 0, 0 )` ( the mass ), and with `moments` the moments' fluxes, `G` taken from the
/// column of `a` ( seed frame )
__device__ CellSums image_terms( const Density3 &dn, const double4 &p0, bool facets, bool moments, double a, V3 g, V3 v, V3 o, V3 u ) {
    CellSums S;
    const double X0 = dn.x0 - p0.x, Y0 = dn.y0 - p0.y, Z0 = dn.z0 - p0.z;   // ( the grid in the seed's frame )
    const int iref = voxel_of( a, X0, dn.hx, dn.nx );
    VPoly T;
    T.n = 3;
    T.x[ 0 ] = g.x; T.y[ 0 ] = g.y; T.z[ 0 ] = g.z;
    T.x[ 1 ] = v.x; T.y[ 1 ] = v.y; T.z[ 1 ] = v.z;
    T.x[ 2 ] = o.x; T.y[ 2 ] = o.y; T.z[ 2 ] = o.z;
    const double zmin = fmin( g.z, fmin( v.z, o.z ) ), zmax = fmax( g.z, fmax( v.z, o.z ) );
    int k0, k1;
    slabs_of( zmin, zmax, Z0, dn.hz, dn.nz, k0, k1 );
    const double s15 = sqrt( 15.0 );
    const double ta1 = ( 6 - s15 ) / 21, tw1 = ( 155 - s15 ) / 1200, ta2 = ( 6 + s15 ) / 21, tw2 = ( 155 + s15 ) / 1200;
    VPoly A, B, C, D, E;
    for ( int k = k0; k <= k1; ++k ) {
        // the slab of row `k` ( the first and last slabs unbounded outside the grid: the edge voxels continue the image )
        if ( k > 0 ) clip_axis( T, 2, Z0 + k * dn.hz, false, A ); else A = T;
        if ( k < dn.nz - 1 ) clip_axis( A, 2, Z0 + ( k + 1 ) * dn.hz, true, B ); else B = A;
        if ( B.n < 3 ) continue;
        double ymin = 1e300, ymax = -1e300;
        for ( int q = 0; q < B.n; ++q ) { ymin = fmin( ymin, B.y[ q ] ); ymax = fmax( ymax, B.y[ q ] ); }
        int j0, j1;
        slabs_of( ymin, ymax, Y0, dn.hy, dn.ny, j0, j1 );
        for ( int j = j0; j <= j1; ++j ) {
            if ( j > 0 ) clip_axis( B, 1, Y0 + j * dn.hy, false, C ); else C = B;
            if ( j < dn.ny - 1 ) clip_axis( C, 1, Y0 + ( j + 1 ) * dn.hy, true, D ); else D = C;
            if ( D.n < 3 ) continue;
            double xmin = 1e300, xmax = -1e300;
            for ( int q = 0; q < D.n; ++q ) { xmin = fmin( xmin, D.x[ q ] ); xmax = fmax( xmax, D.x[ q ] ); }
            int i0, i1;
            slabs_of( xmin, xmax, X0, dn.hx, dn.nx, i0, i1 );
            const double *row = dn.img + ( size_t( k ) * size_t( dn.ny ) + size_t( j ) ) * size_t( dn.nx );
            // the row integrals from column `iref` to the left side of column `i0`
            double s0 = 0, s1 = 0, s2 = 0;
            auto side = [&]( int q ) { return X0 + q * dn.hx; };
            if ( i0 >= iref )
                for ( int q = iref; q < i0; ++q ) {
                    const double lo = side( q ), hi = side( q + 1 ), r = row[ q ];
                    s0 += r * ( hi - lo ); s1 += r * ( hi * hi - lo * lo ) / 2; s2 += r * ( hi * hi * hi - lo * lo * lo ) / 3;
                }
            else
                for ( int q = i0; q < iref; ++q ) {
                    const double lo = side( q ), hi = side( q + 1 ), r = row[ q ];
                    s0 -= r * ( hi - lo ); s1 -= r * ( hi * hi - lo * lo ) / 2; s2 -= r * ( hi * hi * hi - lo * lo * lo ) / 3;
                }
            for ( int i = i0; i <= i1; ++i ) {
                if ( i > 0 ) clip_axis( D, 0, side( i ), false, E ); else E = D;
                VPoly F;
                if ( i < dn.nx - 1 ) clip_axis( E, 0, side( i + 1 ), true, F ); else F = E;
                const double r = row[ i ], xl = side( i );
                if ( F.n >= 3 ) {
                    // the piece's fan from its first vertex
                    for ( int q = 1; q + 1 < F.n; ++q ) {
                        const double ax = F.x[ 0 ], ay = F.y[ 0 ], az = F.z[ 0 ];
                        const double bx = F.x[ q ] - ax, by = F.y[ q ] - ay, bz = F.z[ q ] - az;
                        const double cx = F.x[ q + 1 ] - ax, cy = F.y[ q + 1 ] - ay, cz = F.z[ q + 1 ] - az;
                        const double nx = by * cz - bz * cy, ny = bz * cx - bx * cz, nz = bx * cy - by * cx;
                        const double ar = 0.5 * sqrt( nx * nx + ny * ny + nz * nz );
                        if ( facets ) S.part += r * ar;
                        if ( ! moments ) {       // `G` affine in the voxel: the centroid
                            const double xc = ax + ( bx + cx ) / 3;
                            S.mass += ar * u.x * ( s0 + r * ( xc - xl ) );
                            continue;
                        }
                        for ( int qq = 0; qq < 7; ++qq ) {
                            double b1, b2, wq;
                            if ( qq == 0 ) { b1 = b2 = 1.0 / 3; wq = 9.0 / 40; }
                            else {
                                const double t = qq <= 3 ? ta1 : ta2;
                                wq = qq <= 3 ? tw1 : tw2;
                                const int jj = ( qq - 1 ) % 3;
                                b1 = jj == 1 ? 1 - 2 * t : t; b2 = jj == 2 ? 1 - 2 * t : t;
                            }
                            const double x = ax + b1 * bx + b2 * cx, y = ay + b1 * by + b2 * cy, z = az + b1 * bz + b2 * cz;
                            const double G = s0 + r * ( x - xl ), H = s1 + r * ( x * x - xl * xl ) / 2, K = s2 + r * ( x * x * x - xl * xl * xl ) / 3;
                            const double w = wq * ar * u.x;
                            S.mass += w * G;
                            S.mx += w * H; S.my += w * y * G; S.mz += w * z * G;
                            S.m2 += w * ( K + ( y * y + z * z ) * G );
                        }
                    }
                }
                const double hi = side( i + 1 );
                s0 += r * ( hi - xl ); s1 += r * ( hi * hi - xl * xl ) / 2; s2 += r * ( hi * hi * hi - xl * xl * xl ) / 3;
            }
        }
    }
    return S;
}


/// the triangle `( g, v, o )` of a face of outward unit normal `u`, and the tetrahedron it makes with the cell's mean
/// vertex `A` ( seed frame; the seed at `p0` ): the facet's `int rho`, the mass ( a constant: the tetrahedron's volume; a
/// sum of Gaussians: the flux through the triangle ), the moments ( the tetrahedron's )
template<bool GAUSS>
__device__ __forceinline__ CellSums edge_terms( const Density3 &dn, const double4 &p0, bool facets, bool moments, bool box_face, V3 A, V3 g,
                                               V3 v, V3 o, V3 u ) {
    CellSums S;
    const double ux = v.x - g.x, uy = v.y - g.y, uz = v.z - g.z, tx = o.x - g.x, ty = o.y - g.y, tz = o.z - g.z;
    const double cx = uy * tz - uz * ty, cy = uz * tx - ux * tz, cz = ux * ty - uy * tx;
    const double ar = 0.5 * sqrt( cx * cx + cy * cy + cz * cz );
    const double bx = g.x - A.x, by = g.y - A.y, bz = g.z - A.z, ex = v.x - A.x, ey = v.y - A.y, ez = v.z - A.z;
    const double hx = o.x - A.x, hy = o.y - A.y, hz = o.z - A.z;
    const double det = bx * ( ey * hz - ez * hy ) - by * ( ex * hz - ez * hx ) + bz * ( ex * hy - ey * hx );
    const double vol = fabs( det ) / 6;
    if ( ! GAUSS || dn.kind != 2 ) {                     // ( `GAUSS` false: a constant density, the quadrature compiled out )
        if ( facets ) S.part += dn.rho * ar;
        if ( ! moments ) { S.mass += dn.rho * vol; return S; }
        const double w = dn.rho * vol;
        const double sx = A.x + g.x + v.x + o.x, sy = A.y + g.y + v.y + o.y, sz = A.z + g.z + v.z + o.z;
        const double sq = A.x * A.x + A.y * A.y + A.z * A.z + g.x * g.x + g.y * g.y + g.z * g.z +
                          v.x * v.x + v.y * v.y + v.z * v.z + o.x * o.x + o.y * o.y + o.z * o.z;
        S.mass += w;
        S.mx += w * sx / 4; S.my += w * sy / 4; S.mz += w * sz / 4;
        S.m2 += w * ( sq + ( sx * sx + sy * sy + sz * sz ) ) / 20;
        return S;
    }
    if ( ! moments ) {                                   // the 7-point rule: the facet's density and the mass's flux
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
                    double part = 0, flux = 0;
                    gauss_rule7( dn, p0, P0, P1, P2, sar, u, part, flux );
                    if ( facets ) S.part += part;
                    S.mass += flux;
                }
        return S;
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
    return S;
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

/// a cell's faces in the wave's LDS ( the first pass ) or in the polytope buffer ( the measure kernel )
struct LdsFaces {
    const WaveLds &L;
    __device__ double x( int f ) const { return L.fx[ f ]; }
    __device__ double y( int f ) const { return L.fy[ f ]; }
    __device__ double z( int f ) const { return L.fz[ f ]; }
    __device__ int id( int f ) const { return L.fid[ f ]; }
};
struct BufFaces {
    const PolyBuf &B;
    I64 base;
    __device__ double x( int f ) const { return B.fx[ base + f ]; }
    __device__ double y( int f ) const { return B.fy[ base + f ]; }
    __device__ double z( int f ) const { return B.fz[ base + f ]; }
    __device__ int id( int f ) const { return B.fid[ base + f ]; }
};

/// the measures of a built cell, one wave ( this lane's vertex in `c` ), written to the outputs
template<bool GAUSS, class Faces>
__device__ void measure_cell( const Params3 &P, const Faces &F, const Cell3 &c, int lane, int k0, const double4 &p0, I64 row0 ) {
    // the measures: face by face ( the faces in cut order, as the oracle lists the facets )
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
        const double fnx = F.x( f ), fny = F.y( f ), fnz = F.z( f );
        const int fid = F.id( f );
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
                const CellSums d = edge_terms<GAUSS>( P.dn, p0, facets, moments, fid < 0, Av, g, v, o, u );
                S.part += d.part; S.mass += d.mass; S.mx += d.mx; S.my += d.my; S.mz += d.mz; S.m2 += d.m2;
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

template<bool SPLIT>
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
            // the leaf's seeds in one vector load ( a lane each ), then broadcast in turn
            for ( int base = r.beg; base < r.end && status == C3_CUT; base += 64 ) {
                const int m = r.end - base < 64 ? r.end - base : 64;
                const double4 sq = lane < m ? P.seeds[ base + lane ] : make_double4( 0, 0, 0, 0 );
                for ( int j = 0; j < m; ++j ) {
                    const int k = base + j;
                    if ( k == k0 ) continue;
                    const double dx = readlane_d( sq.x, j ) - p0.x, dy = readlane_d( sq.y, j ) - p0.y, dz = readlane_d( sq.z, j ) - p0.z;
                    const double off = ( ( w0 - readlane_d( sq.w, j ) ) + ( dx * dx + dy * dy + dz * dz ) ) / 2;
                    const int res = cut_cell( L, c, lane, dx, dy, dz, off, k, ev_side, ev_new );
                    if ( res == C3_CUT ) ++ev_cut;
                    if ( res == C3_EMPTY || res == C3_FAILED || res == C3_FULL ) { status = res; break; }
                }
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
    if ( status == C3_FAILED || status == C3_FULL ) {
        if ( lane == 0 && SPLIT ) P.pb.nv[ k0 ] = -1;
        if ( lane == 0 && status == C3_FAILED ) record_failure( P, k0, row0, 2 );
        if ( lane == 0 && status == C3_FULL ) P.out.ovf_list[ atomicAdd( &P.out.counters[ 1 ], 1u ) ] = k0;   // the second pass
        return;
    }
    if ( status == C3_EMPTY ) c.nv = 0;

    if ( SPLIT ) {                                       // ( the Gaussians' measures: the next kernel )
        compact_faces( L, c, lane );
        const I64 at = I64( k0 ) * 64 + lane;
        if ( lane < c.nv ) { P.pb.vx[ at ] = c.x; P.pb.vy[ at ] = c.y; P.pb.vz[ at ] = c.z; P.pb.vpl[ at ] = c.pl; P.pb.vnb[ at ] = c.nb; }
        if ( lane < c.nf ) { P.pb.fx[ at ] = L.fx[ lane ]; P.pb.fy[ at ] = L.fy[ lane ]; P.pb.fz[ at ] = L.fz[ lane ]; P.pb.fid[ at ] = L.fid[ lane ]; }
        if ( lane == 0 ) { P.pb.nv[ k0 ] = status == C3_EMPTY ? 0 : c.nv; P.pb.nf[ k0 ] = c.nf; }
        return;
    }
    measure_cell<false>( P, LdsFaces{ L }, c, lane, k0, p0, row0 );
}

/// what a wave of the measure kernel keeps in LDS: the cell's vertices, its faces ( area centroid, unit normal, the norm of
/// the plane's normal, id, validity ), its triangles ( face, two vertices ) with their facet parts
constexpr int MTRI = 6 * CELL3_VERTICES / 2 + 64;
struct MeasLds {
    double px[ CELL3_VERTICES ], py[ CELL3_VERTICES ], pz[ CELL3_VERTICES ];
    int    pl[ CELL3_VERTICES ], nb[ CELL3_VERTICES ];
    double gx[ CELL3_FACES ], gy[ CELL3_FACES ], gz[ CELL3_FACES ], ux[ CELL3_FACES ], uy[ CELL3_FACES ], uz[ CELL3_FACES ], dn[ CELL3_FACES ];
    int    fid[ CELL3_FACES ], fok[ CELL3_FACES ];
    int    tf[ MTRI ], tv[ MTRI ], tw[ MTRI ];
    double tpart[ MTRI ];
};

/// the second kernel of a sum of Gaussians: the cells the first pass kept, one wave each. The triangles of the cell ( a
/// face's area centroid and one of its edges, `3 nv` of them ) are spread over the lanes, each face's sums taken in a fixed
/// order: the same terms as the face-by-face loop of the first pass, with every lane at work
template<bool IMAGE>
__global__ __launch_bounds__( 64 * WAVES3 ) void k_cells3_measure( Params3 P ) {
    __shared__ MeasLds lds[ WAVES3 ];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    MeasLds &M = lds[ wid ];
    const int k0 = int( blockIdx.x ) * WAVES3 + wid;
    if ( k0 >= P.n ) return;
    const int nv = P.pb.nv[ k0 ];
    if ( nv < 0 ) return;                                // ( measured by the second pass, or failed )
    const int nf = nv > 0 ? P.pb.nf[ k0 ] : 0;
    const I64 at = I64( k0 ) * 64 + lane;
    const double4 p0 = P.seeds[ k0 ];
    const bool act = lane < nv;
    double x = 0, y = 0, z = 0;
    int pl = 0, nb = 0;
    if ( act ) { x = P.pb.vx[ at ]; y = P.pb.vy[ at ]; z = P.pb.vz[ at ]; pl = P.pb.vpl[ at ]; nb = P.pb.vnb[ at ]; M.px[ lane ] = x; M.py[ lane ] = y; M.pz[ lane ] = z; M.pl[ lane ] = pl; M.nb[ lane ] = nb; }
    const double inv_nv = nv > 0 ? 1.0 / nv : 0.0;
    const V3 Av{ wave_total( x ) * inv_nv, wave_total( y ) * inv_nv, wave_total( z ) * inv_nv };
    wave_sync();
    // the faces, a lane each: the area centroid ( `face_apex` ), the unit normal
    if ( lane < nf ) {
        const double fx = P.pb.fx[ at ], fy = P.pb.fy[ at ], fz = P.pb.fz[ at ];
        const double d = sqrt( fx * fx + fy * fy + fz * fz );
        double G[ 3 ];
        const int cnt = face_apex( lane, nv, [&]( int v, int j ) { return byte_of( M.pl[ v ], j ); },
                                   [&]( int v, int j ) { return byte_of( M.nb[ v ], j ); }, M.px, M.py, M.pz, fx, fy, fz, G );
        M.gx[ lane ] = G[ 0 ]; M.gy[ lane ] = G[ 1 ]; M.gz[ lane ] = G[ 2 ];
        M.ux[ lane ] = fx / d; M.uy[ lane ] = fy / d; M.uz[ lane ] = fz / d; M.dn[ lane ] = d;
        M.fid[ lane ] = P.pb.fid[ at ];
        M.fok[ lane ] = cnt >= 3;
    }
    wave_sync();
    // the triangles: each edge of a face counted by its lower vertex
    int cnt = 0;
    if ( act )
        for ( int j = 0; j < 3; ++j ) {
            const int f = byte_of( pl, j );
            if ( ! M.fok[ f ] ) continue;
            cnt += int( lane < byte_of( nb, ( j + 1 ) % 3 ) ) + int( lane < byte_of( nb, ( j + 2 ) % 3 ) );
        }
    int inc = cnt;
    for ( int d = 1; d < 64; d <<= 1 ) {
        const int t = __shfl_up( inc, d, 64 );
        if ( lane >= d ) inc += t;
    }
    const int T = __shfl( inc, 63, 64 );
    if ( act ) {
        int t = inc - cnt;
        for ( int j = 0; j < 3; ++j ) {
            const int f = byte_of( pl, j );
            if ( ! M.fok[ f ] ) continue;
            for ( int e = 1; e <= 2; ++e ) {
                const int w = byte_of( nb, ( j + e ) % 3 );
                if ( lane < w ) { M.tf[ t ] = f; M.tv[ t ] = lane; M.tw[ t ] = w; ++t; }
            }
        }
    }
    wave_sync();
    const bool facets = P.mode == MODE_FACETS, moments = P.mode == MODE_MOMENTS;
    CellSums S;
    for ( int t = lane; t < T; t += 64 ) {
        const int f = M.tf[ t ], v = M.tv[ t ], w = M.tw[ t ];
        const CellSums d = IMAGE ? image_terms( P.dn, p0, facets, moments, Av.x, V3{ M.gx[ f ], M.gy[ f ], M.gz[ f ] },
                                                V3{ M.px[ v ], M.py[ v ], M.pz[ v ] }, V3{ M.px[ w ], M.py[ w ], M.pz[ w ] },
                                                V3{ M.ux[ f ], M.uy[ f ], M.uz[ f ] } )
                                 : edge_terms<true>( P.dn, p0, facets, moments, M.fid[ f ] < 0, Av, V3{ M.gx[ f ], M.gy[ f ], M.gz[ f ] },
                                             V3{ M.px[ v ], M.py[ v ], M.pz[ v ] }, V3{ M.px[ w ], M.py[ w ], M.pz[ w ] },
                                             V3{ M.ux[ f ], M.uy[ f ], M.uz[ f ] } );
        M.tpart[ t ] = d.part;
        S.mass += d.mass; S.mx += d.mx; S.my += d.my; S.mz += d.mz; S.m2 += d.m2;
    }
    wave_sync();
    // the facets: a face's parts summed in triangle order by its lane, the rows' entries in face order
    const I64 row0 = I64( k0 / 64 ) * 64 * P.room + ( k0 & 63 );
    double cf = 0;
    bool has = false;
    if ( facets && lane < nf && M.fok[ lane ] && M.fid[ lane ] >= 0 ) {
        double sp = 0;
        for ( int t = 0; t < T; ++t ) if ( M.tf[ t ] == lane ) sp += M.tpart[ t ];
        cf = sp / ( 2 * M.dn[ lane ] );
        has = true;
    }
    const unsigned long long mh = __ballot( has );
    const int q = __popcll( mh & ( ( 1ull << lane ) - 1 ) ), nq = __popcll( mh );
    if ( has && q < P.room ) {
        P.out.col[ row0 + I64( q ) * 64 ] = M.fid[ lane ];
        P.out.val[ row0 + I64( q ) * 64 ] = cf;
    }
    const double dia = wave_total( cf );
    S.mass = wave_total( S.mass );
    if ( moments ) { S.mx = wave_total( S.mx ); S.my = wave_total( S.my ); S.mz = wave_total( S.mz ); S.m2 = wave_total( S.m2 ); }
    if ( lane == 0 ) write_cell( P, k0, p0, S, dia, nq, row0 );
}

// ---- the second pass: cells past the lanes, one thread each in global slots ----------------------------------------

constexpr int SLOT_V = CELL3_SLOT_VERTICES, SLOT_F = CELL3_SLOT_FACES;   ///< a slot's vertices and faces
constexpr int PASS2_BLOCKS3 = 1024;                      ///< waves of the second pass ( each takes cells in turn )
constexpr int PASS2_VISITS = 16384;                      ///< the nodes a second-pass cell may visit ( ~100 for a typical cell )

/// an overflowed cell in the LDS of its wave: vertices with their faces and neighbours, a cut's staging, the faces, the
/// walk's stack; after the walk the staging holds the measures' faces and triangles
struct Slot3 {
    double x[ SLOT_V ], y[ SLOT_V ], z[ SLOT_V ], s[ SLOT_V ];
    int    pl[ 3 * SLOT_V ], nb[ 3 * SLOT_V ];
    double X[ SLOT_V ], Y[ SLOT_V ], Z[ SLOT_V ];
    int    PL[ 3 * SLOT_V ], NB[ 3 * SLOT_V ], idx[ SLOT_V ], np1[ SLOT_V ], np2[ SLOT_V ];
    double fx[ SLOT_F ], fy[ SLOT_F ], fz[ SLOT_F ], fo[ SLOT_F ];
    int    fid[ SLOT_F ];
    int    stk_m[ STACK3 ], stk_h[ STACK3 ];
    int    tf[ 3 * SLOT_V ], tv[ 3 * SLOT_V ], tw[ 3 * SLOT_V ];   ///< the measures' triangles ( face, vertices ), `3 nv` at most
    double tp[ 3 * SLOT_V ];                                       ///< their facet parts
    int    nv, nf, res;
};

/// the rest of a cut once the signs `s` are known and one is positive ( lane 0 alone, on the slot in LDS ): the
/// classification with the tie bound, the inside vertices renumbered, a new vertex per crossing edge, the new face
/// closed -- the oracle's `cut3` step by step
__device__ int cut_slot_rest( Slot3 &C, double nx, double ny, double nz, double off, int id, double vm, unsigned long long &ev_new ) {
    const int nv = C.nv;
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

/// the cut of an overflowed cell by its wave: the signs by every lane, the rest by lane 0 when a vertex is outside
__device__ int cut_slot( Slot3 &C, int lane, double nx, double ny, double nz, double off, int id, unsigned long long &ev_side,
                         unsigned long long &ev_new ) {
    const int nv = C.nv;
    ev_side += unsigned( nv );
    bool any = false;
    double vm = 0;
    for ( int v = lane; v < nv; v += 64 ) {
        const double sv = fma( nx, C.x[ v ], fma( ny, C.y[ v ], fma( nz, C.z[ v ], -off ) ) );
        C.s[ v ] = sv;
        any |= sv > 0;
        vm = fmax( vm, fmax( fabs( C.x[ v ] ), fmax( fabs( C.y[ v ] ), fabs( C.z[ v ] ) ) ) );
    }
    if ( __ballot( any ) == 0 ) return C3_UNCHANGED;
    vm = wave_maxd( vm );
    __syncthreads();                                     // ( the block is the wave )
    if ( lane == 0 ) C.res = cut_slot_rest( C, nx, ny, nz, off, id, vm, ev_new );
    __syncthreads();
    return C.res;
}

/// the second pass: the cells past the lanes of the first, one wave per cell with the cell in LDS ( up to `SLOT_V`
/// vertices ); a cell past that, or one whose walk visits more than `PASS2_VISITS` nodes, is a failure ( code 1 )
__global__ __launch_bounds__( 64 ) void k_cells3_pass2( Params3 P ) {
    __shared__ Slot3 C;
    const int lane = threadIdx.x;
    const int count = int( P.out.counters[ 1 ] );
    for ( int it = blockIdx.x; it < count; it += PASS2_BLOCKS3 ) {
        const int k0 = P.out.ovf_list[ it ];
        const double4 p0 = P.seeds[ k0 ];
        const double w0 = p0.w;
        unsigned long long ev_side = 0, ev_new = 0;
        __syncthreads();
        if ( lane < 8 ) {
            const int v = lane;
            C.x[ v ] = ( v & 1 ? P.hi[ 0 ] : P.lo[ 0 ] ) - p0.x;
            C.y[ v ] = ( v & 2 ? P.hi[ 1 ] : P.lo[ 1 ] ) - p0.y;
            C.z[ v ] = ( v & 4 ? P.hi[ 2 ] : P.lo[ 2 ] ) - p0.z;
            for ( int d = 0; d < 3; ++d ) { C.pl[ 3 * v + d ] = 2 * d + ( ( v >> d ) & 1 ); C.nb[ 3 * v + d ] = v ^ ( 1 << d ); }
        }
        if ( lane < 6 ) {
            const int f = lane, d = f >> 1, hi = f & 1;
            const double pc = d == 0 ? p0.x : d == 1 ? p0.y : p0.z;
            C.fx[ f ] = d == 0 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            C.fy[ f ] = d == 1 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            C.fz[ f ] = d == 2 ? ( hi ? 1.0 : -1.0 ) : 0.0;
            C.fo[ f ] = hi ? P.hi[ d ] - pc : pc - P.lo[ d ];
            C.fid[ f ] = hi ? -2 - 2 * d : -1 - 2 * d;
        }
        if ( lane == 0 ) { C.nv = 8; C.nf = 6; C.stk_m[ 0 ] = 0; C.stk_h[ 0 ] = P.depth; }
        __syncthreads();
        int status = C3_CUT, top = 1, visits = 0;
        while ( top > 0 && status == C3_CUT ) {
            --top;
            const int m = C.stk_m[ top ], h = C.stk_h[ top ];
            const NodeRec3 r = P.nodes[ m ];
            if ( r.beg >= r.end ) continue;
            // a cell that spans so much of the tree comes from weights no step keeps ( a trial's ): given up as too large
            if ( ++visits > PASS2_VISITS ) { status = C3_FULL; break; }
            bool may = false;
            for ( int v = lane; v < C.nv && ! may; v += 64 ) {
                const double vv[ 3 ] = { C.x[ v ], C.y[ v ], C.z[ v ] }, pp[ 3 ] = { p0.x, p0.y, p0.z };
                double sv = w0 - r.wb;
#pragma unroll
                for ( int d = 0; d < 3; ++d ) {
                    const double X = vv[ d ] + pp[ d ];
                    double y = X + r.wa[ d ] / 2;
                    y = y < r.lo[ d ] ? r.lo[ d ] : ( y > r.hi[ d ] ? r.hi[ d ] : y );
                    const double u = y - X;
                    sv += u * u - vv[ d ] * vv[ d ] - r.wa[ d ] * y;
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
                    const int res = cut_slot( C, lane, dx, dy, dz, off, k, ev_side, ev_new );
                    if ( res == C3_EMPTY || res == C3_FAILED || res == C3_FULL ) { status = res; break; }
                }
                continue;
            }
            const int lc = m + 1, rc = m + ( 1 << ( h - 1 ) );
            const double dl = proximity3( P.nodes[ lc ], p0.x, p0.y, p0.z ), dr = proximity3( P.nodes[ rc ], p0.x, p0.y, p0.z );
            if ( top + 2 > STACK3 ) { status = C3_FAILED; break; }
            __syncthreads();
            if ( lane == 0 ) {
                C.stk_m[ top ] = dl <= dr ? rc : lc; C.stk_h[ top ] = h - 1;
                C.stk_m[ top + 1 ] = dl <= dr ? lc : rc; C.stk_h[ top + 1 ] = h - 1;
            }
            top += 2;
            __syncthreads();
        }
        if ( lane == 0 ) { atomicAdd( &P.out.events[ 0 ], ev_side ); atomicAdd( &P.out.events[ 2 ], ev_new ); }
        const I64 row0 = I64( k0 / 64 ) * 64 * P.room + ( k0 & 63 );
        if ( status == C3_FAILED || status == C3_FULL ) {
            if ( lane == 0 ) record_failure( P, k0, row0, status == C3_FULL ? 1 : 2 );
            continue;
        }
        const int nv = status == C3_EMPTY ? 0 : C.nv, nf = nv > 0 ? C.nf : 0;
        // ---- the measures, the triangles spread over the lanes as in the measure kernel ( staging arrays reused: face
        // means in X, Y, Z, unit normals in fx, fy, fz after the norms in s, validity in idx )
        double ax = 0, ay = 0, az = 0;
        for ( int v = lane; v < nv; v += 64 ) { ax += C.x[ v ]; ay += C.y[ v ]; az += C.z[ v ]; }
        const double inv_nv = nv > 0 ? 1.0 / nv : 0.0;
        const V3 Av{ wave_total( ax ) * inv_nv, wave_total( ay ) * inv_nv, wave_total( az ) * inv_nv };
        __syncthreads();
        for ( int f = lane; f < nf; f += 64 ) {
            double G[ 3 ];                               // ( the area centroid, `face_apex` )
            const int cnt = face_apex( f, nv, [&]( int v, int j ) { return C.pl[ 3 * v + j ]; }, [&]( int v, int j ) { return C.nb[ 3 * v + j ]; },
                                       C.x, C.y, C.z, C.fx[ f ], C.fy[ f ], C.fz[ f ], G );
            C.X[ f ] = G[ 0 ]; C.Y[ f ] = G[ 1 ]; C.Z[ f ] = G[ 2 ];
            const double d = sqrt( C.fx[ f ] * C.fx[ f ] + C.fy[ f ] * C.fy[ f ] + C.fz[ f ] * C.fz[ f ] );
            C.s[ f ] = d;
            C.idx[ f ] = cnt >= 3;
        }
        __syncthreads();
        for ( int f = lane; f < nf; f += 64 ) { C.fx[ f ] /= C.s[ f ]; C.fy[ f ] /= C.s[ f ]; C.fz[ f ] /= C.s[ f ]; }
        // the triangles, each edge of a face counted by its lower vertex, in vertex order
        int T = 0;
        for ( int v0 = 0; v0 < nv; v0 += 64 ) {
            const int v = v0 + lane;
            int cnt = 0;
            if ( v < nv )
                for ( int j = 0; j < 3; ++j ) {
                    if ( ! C.idx[ C.pl[ 3 * v + j ] ] ) continue;
                    cnt += int( v < C.nb[ 3 * v + ( j + 1 ) % 3 ] ) + int( v < C.nb[ 3 * v + ( j + 2 ) % 3 ] );
                }
            int inc = cnt;
            for ( int d = 1; d < 64; d <<= 1 ) {
                const int t = __shfl_up( inc, d, 64 );
                if ( lane >= d ) inc += t;
            }
            if ( v < nv ) {
                int t = T + inc - cnt;
                for ( int j = 0; j < 3; ++j ) {
                    const int f = C.pl[ 3 * v + j ];
                    if ( ! C.idx[ f ] ) continue;
                    for ( int e = 1; e <= 2; ++e ) {
                        const int w = C.nb[ 3 * v + ( j + e ) % 3 ];
                        if ( v < w ) { C.tf[ t ] = f; C.tv[ t ] = v; C.tw[ t ] = w; ++t; }
                    }
                }
            }
            T += __shfl( inc, 63, 64 );
        }
        __syncthreads();
        const bool facets = P.mode == MODE_FACETS, moments = P.mode == MODE_MOMENTS;
        CellSums S;
        for ( int t = lane; t < T; t += 64 ) {
            const int f = C.tf[ t ], v = C.tv[ t ], w = C.tw[ t ];
            const V3 gf{ C.X[ f ], C.Y[ f ], C.Z[ f ] }, vv{ C.x[ v ], C.y[ v ], C.z[ v ] }, ww{ C.x[ w ], C.y[ w ], C.z[ w ] }, uf{ C.fx[ f ], C.fy[ f ], C.fz[ f ] };
            const CellSums d = P.dn.kind == 1 ? image_terms( P.dn, p0, facets, moments, Av.x, gf, vv, ww, uf )
                                              : edge_terms<true>( P.dn, p0, facets, moments, C.fid[ f ] < 0, Av, gf, vv, ww, uf );
            S.mass += d.mass; S.mx += d.mx; S.my += d.my; S.mz += d.mz; S.m2 += d.m2;
            C.tp[ t ] = d.part;
        }
        __syncthreads();
        // the facets in face order: a face's parts summed in triangle order by its lane
        double dsum = 0;
        int qbase = 0;
        for ( int f0 = 0; f0 < nf; f0 += 64 ) {
            const int f = f0 + lane;
            double cf = 0;
            bool has = false;
            if ( facets && f < nf && C.idx[ f ] && C.fid[ f ] >= 0 ) {
                double sp = 0;
                for ( int t = 0; t < T; ++t ) if ( C.tf[ t ] == f ) sp += C.tp[ t ];
                cf = sp / ( 2 * C.s[ f ] );
                has = true;
            }
            const unsigned long long mh = __ballot( has );
            const int q = qbase + __popcll( mh & ( ( 1ull << lane ) - 1 ) );
            if ( has && q < P.room ) {
                P.out.col[ row0 + I64( q ) * 64 ] = C.fid[ f ];
                P.out.val[ row0 + I64( q ) * 64 ] = cf;
            }
            dsum += cf;
            qbase += __popcll( mh );
        }
        const double dia = wave_total( dsum );
        S.mass = wave_total( S.mass );
        if ( moments ) { S.mx = wave_total( S.mx ); S.my = wave_total( S.my ); S.mz = wave_total( S.mz ); S.m2 = wave_total( S.m2 ); }
        if ( lane == 0 ) write_cell( P, k0, p0, S, dia, qbase, row0 );
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
    const bool split = dn.kind == 2 || dn.kind == 1;     // ( Gaussians or an image: the measures in a kernel of their own )
    if ( split ) {
        Arena &A = arena();
        const size_t m = size_t( t.n ) * 64;
        P.pb.vx = A.get<double>( "cells3.vx", m ); P.pb.vy = A.get<double>( "cells3.vy", m ); P.pb.vz = A.get<double>( "cells3.vz", m );
        P.pb.vpl = A.get<int>( "cells3.vpl", m ); P.pb.vnb = A.get<int>( "cells3.vnb", m );
        P.pb.fx = A.get<double>( "cells3.fx", m ); P.pb.fy = A.get<double>( "cells3.fy", m ); P.pb.fz = A.get<double>( "cells3.fz", m );
        P.pb.fid = A.get<int>( "cells3.fid", m );
        P.pb.nv = A.get<int>( "cells3.nv", size_t( t.n ) ); P.pb.nf = A.get<int>( "cells3.nf", size_t( t.n ) );
        k_cells3<true><<<blocks_for( t.n, WAVES3 ), 64 * WAVES3, 0, st>>>( P );
        check_launch( "k_cells3" );
        if ( dn.kind == 1 ) k_cells3_measure<true><<<blocks_for( t.n, WAVES3 ), 64 * WAVES3, 0, st>>>( P );
        else                k_cells3_measure<false><<<blocks_for( t.n, WAVES3 ), 64 * WAVES3, 0, st>>>( P );
        check_launch( "k_cells3_measure" );
    } else {
        k_cells3<false><<<blocks_for( t.n, WAVES3 ), 64 * WAVES3, 0, st>>>( P );
        check_launch( "k_cells3" );
    }
    if ( timed ) timing().end( tp, st );
    k_cells3_pass2<<<PASS2_BLOCKS3, 64, 0, st>>>( P );
    check_launch( "k_cells3_pass2" );
}

} // namespace sdot_amd
