// The Newton direction by flexible CG preconditioned with an aggregation multigrid.
//
// The reference solves `L d = b` with a multigrid whose aggregates follow the BSP tree's order
// ( `src/sdot/include/sdot/sdotplan/Multigrid.h:11-18`: consecutive tree ranks are neighbours in
// space, so packs of consecutive ranks are compact aggregates; the card's version
// `gpu/Linear2D.cuh:1104-1506`, packs of 4 in 2D, a flexible CG and a K-cycle ). This one is its own
// design for CDNA4:
//   * aggregates of `pack` consecutive ranks: restriction and prolongation are index arithmetic
//     ( `I = i / pack` ), nothing to store;
//   * Galerkin coarse Laplacians `P^T L P` ( P piecewise constant ), one lane per coarse row merging the
//     `pack` fine rows in LDS; coarse rows stored COLUMN-MAJOR ( entry `q` of row `I` at `I + q n` ), so the
//     lanes of a wave read consecutive words;
//   * the cycle: damped Jacobi before and after the coarse correction; on levels 1 .. `kcycle` the coarse
//     correction is two steps of flexible CG preconditioned by the next level's cycle ( Notay's K-cycle:
//     the iteration count stops growing with the number of levels ); below, one cycle; the coarsest level
//     ( <= 64 rows ) by a dense inverse of `L_c + alpha 1 1^T` ( the rank-one term removes the constants,
//     the kernel of a Laplacian );
//   * the outer flexible CG ( `alpha = p . r / p . q`, `beta = -z . q / p . q` ) in the zero-mean gauge:
//     constants are invisible to `L`, so the iteration needs no projection, and the solution is shifted at
//     the end to the reference's gauge `d[ gauge ] = 0`;
//   * every scalar stays on the device ( the K-cycle's and the CG's, the stop flag that turns the remaining
//     launches of a batch into no-ops ), the host reads the state once per batch.
// Deterministic: fixed-shape reductions, each coarse row merged in a fixed order.

#include "multigrid.h"

#include <algorithm>
#include <cstdint>

namespace sdot_amd {

namespace {

constexpr int MG_STRIDE = 96;      ///< neighbours of a coarse row, at most ( 3D aggregates of 8 reach ~60 )
constexpr int MG_BLOCK = 64;
constexpr int MAX_LEVELS = 24;
constexpr int DENSE_MAX = 64;      ///< the coarsest level, solved by a dense inverse

/// the outer CG's device state
enum { S_PQ = 0, S_PR, S_ALPHA, S_BETA, S_ZQ, S_RR, S_NB, S_ERR, S_DONE, S_IT, S_TOL, S_FAILED, S_BAD, S_TRUE, S_NB_STATE };
/// a K-cycle's partial sums, per level: `rho1 = c1 . v1`, `alpha1 = c1 . r`, `beta = c2 . v2`, `gamma = c2 . v1`, `alpha2 = c2 . r2`
enum { K_RHO1 = 0, K_ALPHA1, K_BETA2, K_GAMMA, K_ALPHA2, K_NB };

/// a level's matrix: entry `q` of row `i` at `base + q * step`, `base = row_off[ i ]` ( CSR, `step` 1 ) or `i`
/// ( column-major, `row_off` null, `step` the row count ); the off-diagonal values are the `c_ij > 0`
/// ( `R` the type the values are stored in: double, or float for the FP32 preconditioner's levels )
template<class R>
struct MatT {
    I64 n = 0, step = 1;
    const int *row_off = nullptr, *row_cnt = nullptr, *col = nullptr;
    const R *val = nullptr, *dia = nullptr;
};
using Mat = MatT<double>;

/// `( A x )_i`, accumulated in double whatever the storage
template<class R, class X>
__device__ __forceinline__ double row_times( const MatT<R> &A, I64 i, const X *x ) {
    double s = double( A.dia[ i ] ) * double( x[ i ] );
    const I64 b = A.row_off ? A.row_off[ i ] : i;
    const int c = A.row_cnt[ i ];
    for ( int q = 0; q < c; ++q ) {
        const I64 e = b + q * A.step;
        s -= double( A.val[ e ] ) * double( x[ A.col[ e ] ] );
    }
    return s;
}

/// the Galerkin product of one coarse row: the `pack` fine rows merged by coarse column, written column-major
template<class RF, class RC>
__global__ __launch_bounds__( MG_BLOCK ) void k_coarse_op( MatT<RF> F, int pack, I64 nc, int *row_cnt, int *col, RC *val, RC *dia, double *state ) {
    __shared__ int sj[ MG_STRIDE * MG_BLOCK ];
    __shared__ double sc[ MG_STRIDE * MG_BLOCK ];
    const I64 I = I64( blockIdx.x ) * MG_BLOCK + threadIdx.x;
    if ( I >= nc ) return;
    int *J = sj + threadIdx.x;
    double *C = sc + threadIdx.x;
    int u = 0;
    bool bad = false;
    double d = 0;
    const I64 i0 = I * pack, i1 = i0 + pack < F.n ? i0 + pack : F.n;
    for ( I64 i = i0; i < i1; ++i ) {
        d += double( F.dia[ i ] );
        const I64 b = F.row_off ? F.row_off[ i ] : i;
        const int cnt = F.row_cnt[ i ];
        for ( int q = 0; q < cnt; ++q ) {
            const I64 e = b + q * F.step;
            const int jc = F.col[ e ] / pack;
            const double c = double( F.val[ e ] );
            if ( jc == int( I ) ) { d -= c; continue; }
            int k = 0;
            while ( k < u && J[ k * MG_BLOCK ] != jc ) ++k;
            if ( k == u ) {
                if ( u == MG_STRIDE ) { bad = true; continue; }
                J[ u * MG_BLOCK ] = jc;
                C[ u * MG_BLOCK ] = 0;
                ++u;
            }
            C[ k * MG_BLOCK ] += c;
        }
    }
    row_cnt[ I ] = u;
    for ( int k = 0; k < u; ++k ) {
        col[ I + k * nc ] = J[ k * MG_BLOCK ];
        val[ I + k * nc ] = RC( C[ k * MG_BLOCK ] );
    }
    dia[ I ] = RC( d > 0 ? d : 1.0 );
    if ( bad ) state[ S_BAD ] = 1;
}

/// the inverse of `L_c + alpha 1 1^T` ( `alpha` the mean diagonal over `n`: the rank-one term removes the
/// constants, the kernel of a Laplacian ), by Gauss-Jordan in LDS, one workgroup; writes it row-major
template<class R>
__global__ __launch_bounds__( 256 ) void k_dense_inverse( MatT<R> C, double *inv ) {
    __shared__ double M[ DENSE_MAX ][ DENSE_MAX ], V[ DENSE_MAX ][ DENSE_MAX ];
    __shared__ double sh[ 256 ];
    const int n = int( C.n );
    double s = 0;
    for ( int i = threadIdx.x; i < n; i += blockDim.x ) s += double( C.dia[ i ] );
    const double alpha = block_sum<256>( s, sh ) / double( n ) / double( n );
    for ( int k = threadIdx.x; k < n * n; k += blockDim.x ) { M[ k / n ][ k % n ] = alpha; V[ k / n ][ k % n ] = ( k / n == k % n ) ? 1.0 : 0.0; }
    __syncthreads();
    if ( int( threadIdx.x ) < n ) {
        const int i = threadIdx.x;
        M[ i ][ i ] += double( C.dia[ i ] );
        const I64 b = C.row_off ? C.row_off[ i ] : i;
        for ( int q = 0; q < C.row_cnt[ i ]; ++q ) M[ i ][ C.col[ b + q * C.step ] ] -= double( C.val[ b + q * C.step ] );
    }
    __syncthreads();
    for ( int p = 0; p < n; ++p ) {
        const double piv = M[ p ][ p ];
        __syncthreads();
        for ( int c = threadIdx.x; c < n; c += blockDim.x ) { M[ p ][ c ] /= piv; V[ p ][ c ] /= piv; }
        __syncthreads();
        for ( int k = threadIdx.x; k < n * n; k += blockDim.x ) {
            const int r = k / n, c = k % n;
            if ( r == p ) continue;
            const double f = M[ r ][ p ];
            if ( c != p ) M[ r ][ c ] -= f * M[ p ][ c ];
            V[ r ][ c ] -= f * V[ p ][ c ];
        }
        __syncthreads();
        for ( int r = threadIdx.x; r < n; r += blockDim.x ) if ( r != p ) M[ r ][ p ] = 0;
        __syncthreads();
    }
    for ( int k = threadIdx.x; k < n * n; k += blockDim.x ) inv[ k ] = V[ k / n ][ k % n ];
}

// ---- the cycle --------------------------------------------------------------------------------------------

/// the damping of each sweep of a smoothing pass ( at most four sweeps; read by selects, never indexed per lane )
struct Damping {
    double w0 = 0.7, w1 = 0.7, w2 = 0.7, w3 = 0.7;
    __host__ __device__ __forceinline__ double operator()( int s ) const { return s == 0 ? w0 : s == 1 ? w1 : s == 2 ? w2 : w3; }
};

/// the sweeps' dampings for `o`: Chebyshev nodes of `[ 2 / ratio, 2 ]` ( largest first ), or `omega` for each
Damping damping_of( const MgOptions &o ) {
    Damping d;
    if ( ! o.chebyshev ) { d.w0 = d.w1 = d.w2 = d.w3 = o.omega; return d; }
    const double b = 2.0, a = b / o.ratio, th = ( b + a ) / 2, de = ( b - a ) / 2;
    const double pi = 3.14159265358979323846;
    double w[ 4 ] = { 0.7, 0.7, 0.7, 0.7 };
    for ( int i = 0; i < o.nu && i < 4; ++i ) w[ i ] = 1 / ( th + de * std::cos( ( 2 * i + 1 ) * pi / ( 2 * o.nu ) ) );
    d.w0 = w[ 0 ]; d.w1 = w[ 1 ]; d.w2 = w[ 2 ]; d.w3 = w[ 3 ];
    return d;
}

/// the row of the calling lane, the blocks of one XCD taking a contiguous range of rows: blocks are dealt round-robin to
/// the 8 XCDs, each with its own L2, and a row's neighbours ( near ranks ) are then read from the same L2
__device__ __forceinline__ I64 xcd_row() {
    const int nblk = int( gridDim.x ), b = int( blockIdx.x );
    const int per = nblk / 8, rem = nblk % 8, xcd = b % 8, slot = b / 8;
    const int lb = xcd < rem ? xcd * ( per + 1 ) + slot : rem * ( per + 1 ) + ( xcd - rem ) * per + slot;
    return I64( lb ) * blockDim.x + threadIdx.x;
}

template<class R>
__global__ void k_jacobi_first( I64 n, const R *r, const R *dia, double omega, R *x, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = xcd_row();
    if ( i < n ) x[ i ] = R( omega * double( r[ i ] ) / double( dia[ i ] ) );
}

template<class R>
__global__ void k_jacobi( MatT<R> A, const R *r, const R *x, R *xo, double omega, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = xcd_row();
    if ( i < A.n ) xo[ i ] = R( double( x[ i ] ) + omega * ( double( r[ i ] ) - row_times( A, i, x ) ) / double( A.dia[ i ] ) );
}

/// the first two sweeps from zero in one pass: `x1 = omega D^-1 r` read on the fly, `xo = x1 + omega2 D^-1 ( r - A x1 )`
template<class R>
__global__ void k_jacobi_two( MatT<R> A, const R *r, R *xo, double omega, double omega2, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = xcd_row();
    if ( i >= A.n ) return;
    const double di = double( A.dia[ i ] );
    // ( the first sweep's value of each row as the second reads it: stored in `R`, so rounded alike for every row )
    const double x1 = double( R( omega * double( r[ i ] ) / di ) );
    double s = di * x1;
    const I64 b = A.row_off ? A.row_off[ i ] : i;
    const int c = A.row_cnt[ i ];
    for ( int q = 0; q < c; ++q ) {
        const I64 e = b + q * A.step;
        const int j = A.col[ e ];
        s -= double( A.val[ e ] ) * double( R( omega * double( r[ j ] ) / double( A.dia[ j ] ) ) );
    }
    xo[ i ] = R( x1 + omega2 * ( double( r[ i ] ) - s ) / di );
}

/// the coarse correction prolonged and the first post-sweep in one pass: `x' = x + P xc` read on the fly,
/// `xo = x' + omega D^-1 ( r - A x' )`
template<class R>
__global__ void k_prolong_jacobi( MatT<R> A, const R *r, const R *x, const R *xc, int pack, R *xo, double omega, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = xcd_row();
    if ( i >= A.n ) return;
    const double di = double( A.dia[ i ] );
    const double xi = double( R( double( x[ i ] ) + double( xc[ i / pack ] ) ) );
    double s = di * xi;
    const I64 b = A.row_off ? A.row_off[ i ] : i;
    const int c = A.row_cnt[ i ];
    for ( int q = 0; q < c; ++q ) {
        const I64 e = b + q * A.step;
        const int j = A.col[ e ];
        s -= double( A.val[ e ] ) * double( R( double( x[ j ] ) + double( xc[ j / pack ] ) ) );
    }
    xo[ i ] = R( xi + omega * ( double( r[ i ] ) - s ) / di );
}

/// `xo = x + P xc`
template<class R>
__global__ void k_prolong_to( I64 n, int pack, const R *x, const R *xc, R *xo, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = xcd_row();
    if ( i < n ) xo[ i ] = R( double( x[ i ] ) + double( xc[ i / pack ] ) );
}

/// `rc[ I ] = sum_{ i in I } ( r - A x )_i`: one lane per fine row, the `pack` consecutive lanes of an aggregate
/// summed by shuffles ( `pack` divides 64 )
template<class R>
__global__ void k_restrict( MatT<R> A, const R *r, const R *x, int pack, R *rc, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = xcd_row();
    double v = i < A.n ? double( r[ i ] ) - row_times( A, i, x ) : 0.0;
    for ( int off = pack >> 1; off > 0; off >>= 1 ) v += __shfl_down( v, off, pack );
    if ( i < A.n && ( i % pack ) == 0 ) rc[ i / pack ] = R( v );
}


/// `x = Inv r`, one wave per row
template<class R>
__global__ void k_dense_apply( int n, const double *inv, const R *r, R *x, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const int row = blockIdx.x * ( blockDim.x / 64 ) + threadIdx.x / 64, lane = threadIdx.x & 63;
    if ( row >= n ) return;
    double s = lane < n ? inv[ size_t( row ) * n + lane ] * double( r[ lane ] ) : 0.0;
    for ( int off = 32; off > 0; off >>= 1 ) s += __shfl_down( s, off, 64 );
    if ( lane == 0 ) x[ row ] = R( s );
}

// ---- the bottom of the hierarchy in one workgroup ---------------------------------------------------------------
//
// Below a few thousand rows a level's kernels are too small to fill the chip, and a launch costs more than its work: the
// levels from `lb` down run as ONE kernel of one workgroup ( V-cycles: pre-smoothing, restriction, ... the dense
// coarsest solve ... prolongation, post-smoothing ), the steps separated by barriers. Their matrices and vectors stay in
// the L2 of the workgroup's XCD.

template<class R>
struct BottomLevel {
    I64 n = 0;
    const int *cnt = nullptr, *col = nullptr;            ///< column-major: entry `q` of row `i` at `i + q n`
    const R *val = nullptr, *dia = nullptr;
    R *x = nullptr, *t = nullptr, *r = nullptr;
    __device__ __forceinline__ double row_times( I64 i, const R *v ) const {
        double s = double( dia[ i ] ) * double( v[ i ] );
        for ( int q = 0; q < cnt[ i ]; ++q ) s -= double( val[ i + q * n ] ) * double( v[ col[ i + q * n ] ] );
        return s;
    }
};
template<class R>
struct Bottom {
    int nl = 0, pack = 4, nu = 2, nc = 0;
    Damping omega;
    const double *inv = nullptr;
    int rows = 0;                                        ///< the levels' rows together
    BottomLevel<R> L[ MAX_LEVELS ];
};
constexpr int BOTTOM_THREADS = 1024;
constexpr int BOTTOM_ROWS = 640, BOTTOM_ENTRIES = 6144;   ///< what the LDS copy of the bottom holds

/// one V-cycle over the levels `L[ 0 .. nl - 1 ]` ( `L[ 0 ].r` in, `L[ 0 ].x` out ), the steps separated by barriers
template<class R>
__device__ void bottom_cycle( const BottomLevel<R> *L, int nl, int pack, int nu, const Damping &omega, const double *inv, int nc ) {
    const int tid = threadIdx.x;
    for ( int l = 0; l + 1 < nl; ++l ) {
        const BottomLevel<R> &A = L[ l ];
        // pre-smoothing from zero, `nu` sweeps ping-ponging so that the last lands in `x`
        R *cur = ( nu - 1 ) % 2 ? A.t : A.x, *oth = ( nu - 1 ) % 2 ? A.x : A.t;
        for ( I64 i = tid; i < A.n; i += BOTTOM_THREADS ) cur[ i ] = R( omega( 0 ) * double( A.r[ i ] ) / double( A.dia[ i ] ) );
        __syncthreads();
        for ( int s = 1; s < nu; ++s ) {
            const double w = omega( s );
            for ( I64 i = tid; i < A.n; i += BOTTOM_THREADS )
                oth[ i ] = R( double( cur[ i ] ) + w * ( double( A.r[ i ] ) - A.row_times( i, cur ) ) / double( A.dia[ i ] ) );
            __syncthreads();
            R *sw = cur; cur = oth; oth = sw;
        }
        // the residual restricted
        const BottomLevel<R> &C = L[ l + 1 ];
        for ( I64 I = tid; I < C.n; I += BOTTOM_THREADS ) {
            double v = 0;
            const I64 e = ( I + 1 ) * pack < A.n ? ( I + 1 ) * pack : A.n;
            for ( I64 i = I * pack; i < e; ++i ) v += double( A.r[ i ] ) - A.row_times( i, A.x );
            C.r[ I ] = R( v );
        }
        __syncthreads();
    }
    {
        const BottomLevel<R> &D = L[ nl - 1 ];           // the coarsest: the dense inverse
        for ( int row = tid; row < nc; row += BOTTOM_THREADS ) {
            double v = 0;
            for ( int j = 0; j < nc; ++j ) v += inv[ size_t( row ) * nc + j ] * double( D.r[ j ] );
            D.x[ row ] = R( v );
        }
        __syncthreads();
    }
    for ( int l = nl - 2; l >= 0; --l ) {
        const BottomLevel<R> &A = L[ l ];
        const R *xc = L[ l + 1 ].x;
        // prolongation, then `nu` sweeps ending in `x`: from `t` when `nu` is odd, else from `x` through `t`
        R *cur = nu % 2 ? A.t : A.x, *oth = nu % 2 ? A.x : A.t;
        for ( I64 i = tid; i < A.n; i += BOTTOM_THREADS ) cur[ i ] = R( double( A.x[ i ] ) + double( xc[ i / pack ] ) );
        __syncthreads();
        for ( int s = 0; s < nu; ++s ) {
            const double w = omega( s );
            for ( I64 i = tid; i < A.n; i += BOTTOM_THREADS )
                oth[ i ] = R( double( cur[ i ] ) + w * ( double( A.r[ i ] ) - A.row_times( i, cur ) ) / double( A.dia[ i ] ) );
            __syncthreads();
            R *sw = cur; cur = oth; oth = sw;
        }
    }
}

/// `x ~ A^-1 r` on the bottom levels ( `B.L[ 0 ].r` in, `B.L[ 0 ].x` out ), one V-cycle. When the levels fit, they are
/// first copied in LDS as one CSR matrix over all their rows ( row `g` of the copy is row `g - base[ l ]` of level `l`,
/// columns shifted likewise; offsets from a block scan of the row counts ): the cycle's dependent reads then cost LDS
/// latencies instead of L2 ones. Every level index is uniform, so no per-thread array of level views is needed ( such an
/// array lands in scratch memory )
template<class R>
__global__ __launch_bounds__( BOTTOM_THREADS ) void k_bottom( Bottom<R> B, const double *state ) {
    __shared__ int s_base[ MAX_LEVELS + 1 ], s_off[ BOTTOM_ROWS + 1 ], s_col[ BOTTOM_ENTRIES ], s_wave[ BOTTOM_THREADS / 64 ];
    __shared__ R s_val[ BOTTOM_ENTRIES ], s_dia[ BOTTOM_ROWS ], s_r[ BOTTOM_ROWS ], s_xt[ 2 ][ BOTTOM_ROWS ];
    if ( state[ S_DONE ] != 0 ) return;
    if ( B.rows > BOTTOM_ROWS ) { bottom_cycle( B.L, B.nl, B.pack, B.nu, B.omega, B.inv, B.nc ); return; }
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    // this thread's row: level `lv`, index `li` there, `base` the level's first row in the copy
    int cnt = 0, lv = -1, li = 0, base = 0;
    for ( int l = 0, b = 0; l < B.nl; ++l ) {
        const int n = int( B.L[ l ].n );
        if ( tid >= b && tid < b + n ) { lv = l; li = tid - b; base = b; cnt = B.L[ l ].cnt[ li ]; }
        if ( tid == 0 ) s_base[ l ] = b;
        b += n;
    }
    // the offsets: an exclusive scan of the counts ( within the waves by shuffles, then over the wave totals )
    int inc = cnt;
    for ( int d = 1; d < 64; d <<= 1 ) {
        const int v = __shfl_up( inc, d, 64 );
        if ( lane >= d ) inc += v;
    }
    if ( lane == 63 ) s_wave[ wid ] = inc;
    __syncthreads();
    int before = 0, tot = 0;
    for ( int w = 0; w < BOTTOM_THREADS / 64; ++w ) {
        const int v = s_wave[ w ];
        before += w < wid ? v : 0;
        tot += v;
    }
    if ( tot > BOTTOM_ENTRIES ) { bottom_cycle( B.L, B.nl, B.pack, B.nu, B.omega, B.inv, B.nc ); return; }   // ( uniform )
    const int excl = before + inc - cnt;
    for ( int l = 0; l < B.nl; ++l ) {
        if ( lv != l ) continue;
        const BottomLevel<R> &A = B.L[ l ];
        s_off[ tid ] = excl;
        for ( int q = 0; q < cnt; ++q ) {
            s_col[ excl + q ] = base + A.col[ li + I64( q ) * A.n ];
            s_val[ excl + q ] = A.val[ li + I64( q ) * A.n ];
        }
        s_dia[ tid ] = A.dia[ li ];
        s_r[ tid ] = l == 0 ? A.r[ li ] : R( 0 );
    }
    if ( tid == 0 ) { s_off[ B.rows ] = tot; s_base[ B.nl ] = B.rows; }
    __syncthreads();

    auto row_times = [ & ]( int g, int v ) {
        double s = double( s_dia[ g ] ) * double( s_xt[ v ][ g ] );
        for ( int e = s_off[ g ]; e < s_off[ g + 1 ]; ++e ) s -= double( s_val[ e ] ) * double( s_xt[ v ][ s_col[ e ] ] );
        return s;
    };
    const int pack = B.pack, nu = B.nu;
    const Damping omega = B.omega;
    for ( int l = 0; l + 1 < B.nl; ++l ) {
        const int b = s_base[ l ], n = s_base[ l + 1 ] - b;
        // pre-smoothing from zero, `nu` sweeps ping-ponging between `t` ( 1 ) and `x` ( 0 ) so that the last lands in `x`
        int cur = ( nu - 1 ) % 2, oth = 1 - cur;
        for ( int i = tid; i < n; i += BOTTOM_THREADS ) s_xt[ cur ][ b + i ] = R( omega( 0 ) * double( s_r[ b + i ] ) / double( s_dia[ b + i ] ) );
        __syncthreads();
        for ( int s = 1; s < nu; ++s ) {
            const double w = omega( s );
            for ( int i = tid; i < n; i += BOTTOM_THREADS ) {
                const int g = b + i;
                s_xt[ oth ][ g ] = R( double( s_xt[ cur ][ g ] ) + w * ( double( s_r[ g ] ) - row_times( g, cur ) ) / double( s_dia[ g ] ) );
            }
            __syncthreads();
            cur = oth; oth = 1 - cur;
        }
        // the residual restricted
        const int bc = s_base[ l + 1 ], nc = s_base[ l + 2 ] - bc;
        for ( int I = tid; I < nc; I += BOTTOM_THREADS ) {
            double v = 0;
            const int e = ( I + 1 ) * pack < n ? ( I + 1 ) * pack : n;
            for ( int i = I * pack; i < e; ++i ) v += double( s_r[ b + i ] ) - row_times( b + i, 0 );
            s_r[ bc + I ] = R( v );
        }
        __syncthreads();
    }
    {
        const int bd = s_base[ B.nl - 1 ];              // the coarsest: the dense inverse
        for ( int row = tid; row < B.nc; row += BOTTOM_THREADS ) {
            double v = 0;
            for ( int j = 0; j < B.nc; ++j ) v += B.inv[ size_t( row ) * B.nc + j ] * double( s_r[ bd + j ] );
            s_xt[ 0 ][ bd + row ] = R( v );
        }
        __syncthreads();
    }
    for ( int l = B.nl - 2; l >= 0; --l ) {
        const int b = s_base[ l ], n = s_base[ l + 1 ] - b, bc = s_base[ l + 1 ];
        // prolongation, then `nu` sweeps ending in `x`: from `t` when `nu` is odd, else from `x` through `t`
        int cur = nu % 2, oth = 1 - cur;
        for ( int i = tid; i < n; i += BOTTOM_THREADS ) s_xt[ cur ][ b + i ] = R( double( s_xt[ 0 ][ b + i ] ) + double( s_xt[ 0 ][ bc + i / pack ] ) );
        __syncthreads();
        for ( int s = 0; s < nu; ++s ) {
            const double w = omega( s );
            for ( int i = tid; i < n; i += BOTTOM_THREADS ) {
                const int g = b + i;
                s_xt[ oth ][ g ] = R( double( s_xt[ cur ][ g ] ) + w * ( double( s_r[ g ] ) - row_times( g, cur ) ) / double( s_dia[ g ] ) );
            }
            __syncthreads();
            cur = oth; oth = 1 - cur;
        }
    }
    for ( int i = tid; i < int( B.L[ 0 ].n ); i += BOTTOM_THREADS ) B.L[ 0 ].x[ i ] = s_xt[ 0 ][ i ];
}

// ---- the K-cycle's two flexible CG steps ( Notay ) ------------------------------------------------------------
//
//   c1 = C( r ), v1 = A c1, rho1 = c1 . v1, alpha1 = c1 . r, r2 = r - alpha1 / rho1 v1,
//   c2 = C( r2 ), v2 = A c2, gamma = c2 . v1, beta = c2 . v2, alpha2 = c2 . r2, rho2 = beta - gamma^2 / rho1,
//   x = ( alpha1 / rho1 - gamma alpha2 / ( rho1 rho2 ) ) c1 + alpha2 / rho2 c2

template<int T>
__device__ __forceinline__ double block_sum_t( double v, double *sh ) { return block_sum<T>( v, sh ); }

/// the fixed-order sum of `nb` partials: every block that calls it gets the same value ( all its threads take part )
__device__ __forceinline__ double sum_parts( const double *p, int nb, double *sh ) {
    double a = 0;
    for ( int j = threadIdx.x; j < nb; j += RED_THREADS ) a += p[ j ];
    return block_sum<RED_THREADS>( a, sh );
}

/// `v = A c` and the partial sums of `c . v`, `c . u1` and ( `nd == 3` ) `c . u2`, over `nblk` blocks, into the slots
/// `part[ q * RED_BLOCKS ]`
template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_spmv_dots( MatT<R> A, const R *c, R *v, const R *u1, const R *u2, int nd, double *part,
                                                            const double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    double a0 = 0, a1 = 0, a2 = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < A.n; i += I64( gridDim.x ) * RED_THREADS ) {
        const double ci = double( c[ i ] );
        const R vr = R( row_times( A, i, c ) );
        v[ i ] = vr;
        const double vi = double( vr );
        a0 += ci * vi;
        a1 += ci * double( u1[ i ] );
        if ( nd == 3 ) a2 += ci * double( u2[ i ] );
    }
    const double r0 = block_sum<RED_THREADS>( a0, sh );
    const double r1 = block_sum<RED_THREADS>( a1, sh );
    const double r2 = nd == 3 ? block_sum<RED_THREADS>( a2, sh ) : 0.0;
    if ( threadIdx.x == 0 ) {
        part[ blockIdx.x ] = r0;
        part[ RED_BLOCKS + blockIdx.x ] = r1;
        part[ 2 * RED_BLOCKS + blockIdx.x ] = r2;
    }
}

/// `r2 = r - alpha1 / rho1 v1`, `rho1` and `alpha1` summed by every block from the first step's partials ( slots 0, 1 );
/// a grid-stride loop over `nblk` blocks
template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_kc_residual( I64 n, const R *r, const R *v1, const double *part, int nblk, R *r2,
                                                              const double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    const double rho1 = sum_parts( part, nblk, sh ), alpha1 = sum_parts( part + RED_BLOCKS, nblk, sh );
    const double f = rho1 > 0 ? alpha1 / rho1 : 0.0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < n; i += I64( gridDim.x ) * RED_THREADS )
        r2[ i ] = R( double( r[ i ] ) - f * double( v1[ i ] ) );
}

/// `x = a1 c1 + a2 c2` with the K-cycle's coefficients, summed by every block from the two steps' partials ( slots 0 .. 4:
/// `rho1`, `alpha1`, `beta`, `gamma`, `alpha2` ); one step only when the second is degenerate
template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_kc_combine( I64 n, const R *c1, const R *c2, const double *part, int nblk, R *x,
                                                             const double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    double k[ K_NB ];
    for ( int q = 0; q < K_NB; ++q ) k[ q ] = sum_parts( part + q * RED_BLOCKS, nblk, sh );
    const double rho1 = k[ 0 ], alpha1 = k[ 1 ], beta = k[ 2 ], gamma = k[ 3 ], alpha2 = k[ 4 ];
    double a1 = 1, a2 = 0;
    if ( rho1 > 0 ) {
        const double rho2 = beta - gamma * gamma / rho1;
        a1 = alpha1 / rho1;
        if ( rho2 > 0 ) {
            a2 = alpha2 / rho2;
            a1 -= gamma * a2 / rho1;
        }
    }
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < n; i += I64( gridDim.x ) * RED_THREADS )
        x[ i ] = R( a1 * double( c1[ i ] ) + a2 * double( c2[ i ] ) );
}

// ---- the outer flexible CG -------------------------------------------------------------------------------------

template<int T>
__device__ __forceinline__ void write_partial( double v, double *part, int slot ) {
    __shared__ double sh[ T ];
    const double r = block_sum<T>( v, sh );
    if ( threadIdx.x == 0 ) part[ slot ] = r;
}

template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_pcg_init( MatT<R> A, const R *b, R *d, R *r, double *part ) {
    double acc = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < A.n; i += I64( RED_BLOCKS ) * RED_THREADS ) {
        d[ i ] = 0;
        r[ i ] = b[ i ];
        acc += double( b[ i ] ) * double( b[ i ] );
    }
    write_partial<RED_THREADS>( acc, part, blockIdx.x );
}

/// partial sums of `| b - A d |^2`: the true residual, which the flexible CG's recurrence only estimates; with `res`, the
/// residual itself is written too
template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_true_res( MatT<R> A, const R *b, const R *d, double *part, R *res = nullptr ) {
    double acc = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < A.n; i += I64( RED_BLOCKS ) * RED_THREADS ) {
        const double e = double( b[ i ] ) - row_times( A, i, d );
        if ( res ) res[ i ] = R( e );
        acc += e * e;
    }
    write_partial<RED_THREADS>( acc, part, blockIdx.x );
}

__global__ __launch_bounds__( RED_THREADS ) void k_true_res_final( const double *part, double *state ) {
    __shared__ double sh[ RED_THREADS ];
    double a = 0;
    for ( int j = threadIdx.x; j < RED_BLOCKS; j += RED_THREADS ) a += part[ j ];
    const double ee = block_sum<RED_THREADS>( a, sh );
    if ( threadIdx.x == 0 ) state[ S_TRUE ] = state[ S_NB ] > 0 ? sqrt( ee ) / state[ S_NB ] : 0.0;
}

__global__ __launch_bounds__( RED_THREADS ) void k_pcg_init_final( const double *part, double *state, double tol ) {
    __shared__ double sh[ RED_THREADS ];
    double a = 0;
    for ( int j = threadIdx.x; j < RED_BLOCKS; j += RED_THREADS ) a += part[ j ];
    const double bb = block_sum<RED_THREADS>( a, sh );
    if ( threadIdx.x == 0 ) {
        state[ S_NB ] = sqrt( bb );
        state[ S_ERR ] = 1;
        state[ S_DONE ] = bb > 0 ? 0.0 : 1.0;
        state[ S_IT ] = 0;
        state[ S_TOL ] = tol;
        state[ S_FAILED ] = 0;
        state[ S_BETA ] = 0;
        state[ S_PQ ] = 0;
    }
}

/// `q = A p`, partial sums of `p . q` and `p . r`
template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_fcg_spmv( MatT<R> A, const R *p, const R *r, R *q, double *part, const double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    double pq = 0, pr = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < A.n; i += I64( RED_BLOCKS ) * RED_THREADS ) {
        const R sr = R( row_times( A, i, p ) );
        q[ i ] = sr;
        const double pi = double( p[ i ] );
        pq += pi * double( sr );
        pr += pi * double( r[ i ] );
    }
    const double a = block_sum<RED_THREADS>( pq, sh );
    const double b = block_sum<RED_THREADS>( pr, sh );
    if ( threadIdx.x == 0 ) { part[ blockIdx.x ] = a; part[ RED_BLOCKS + blockIdx.x ] = b; }
}

/// `alpha = p . r / p . q` summed by every block from the partials ( slots 0, 1 ), then `d += alpha p`, `r -= alpha q`
/// and the partials of `r . r` ( slot 2 ). `p . q <= 0` ( a breakdown ) stops the iteration as a failure
template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_pcg_update( I64 n, const R *p, const R *q, R *d, R *r, double *part, double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    const double pq = sum_parts( part, RED_BLOCKS, sh ), pr = sum_parts( part + RED_BLOCKS, RED_BLOCKS, sh );
    if ( ! ( pq > 0 ) ) {
        if ( blockIdx.x == 0 && threadIdx.x == 0 ) { state[ S_PQ ] = pq; state[ S_DONE ] = 1; state[ S_FAILED ] = 1; }
        return;
    }
    const double alpha = pr / pq;
    if ( blockIdx.x == 0 && threadIdx.x == 0 ) { state[ S_PQ ] = pq; state[ S_PR ] = pr; state[ S_ALPHA ] = alpha; }
    double acc = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < n; i += I64( RED_BLOCKS ) * RED_THREADS ) {
        d[ i ] = R( double( d[ i ] ) + alpha * double( p[ i ] ) );
        const R ri = R( double( r[ i ] ) - alpha * double( q[ i ] ) );
        r[ i ] = ri;
        acc += double( ri ) * double( ri );
    }
    const double rr = block_sum<RED_THREADS>( acc, sh );
    if ( threadIdx.x == 0 ) part[ 2 * RED_BLOCKS + blockIdx.x ] = rr;
}

__global__ __launch_bounds__( RED_THREADS ) void k_pcg_check( const double *part, double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    const double rr = sum_parts( part + 2 * RED_BLOCKS, RED_BLOCKS, sh );
    if ( threadIdx.x == 0 ) {
        state[ S_RR ] = rr;
        const double err = sqrt( rr ) / state[ S_NB ];
        state[ S_ERR ] = err;
        state[ S_IT ] += 1;
        if ( err <= state[ S_TOL ] ) state[ S_DONE ] = 1;
    }
}

/// partial sums of `z . q` ( slot 3: the flexible CG's `beta` )
template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_fcg_zq( I64 n, const R *z, const R *q, double *part, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    double acc = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < n; i += I64( RED_BLOCKS ) * RED_THREADS ) acc += double( z[ i ] ) * double( q[ i ] );
    write_partial<RED_THREADS>( acc, part, 3 * RED_BLOCKS + blockIdx.x );
}

/// `beta = - z . q / p . q` summed by every block from the partials ( slot 3 ), then `p = z + beta p`
template<class R>
__global__ __launch_bounds__( RED_THREADS ) void k_pcg_dir( I64 n, const R *z, R *p, const double *part, double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    const double zq = sum_parts( part + 3 * RED_BLOCKS, RED_BLOCKS, sh ), pq = state[ S_PQ ];
    const double beta = pq > 0 ? -zq / pq : 0.0;
    if ( blockIdx.x == 0 && threadIdx.x == 0 ) { state[ S_ZQ ] = zq; state[ S_BETA ] = beta; }
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < n; i += I64( RED_BLOCKS ) * RED_THREADS ) p[ i ] = R( double( z[ i ] ) + beta * double( p[ i ] ) );
}

/// the first direction, `p = z`
template<class R>
__global__ void k_pcg_first_dir( I64 n, const R *z, R *p, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) p[ i ] = z[ i ];
}

__global__ void k_shift_gauge( I64 n, double *d, const double *g, int gauge ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n && i != gauge ) d[ i ] -= g[ 0 ];
}

__global__ void k_read_gauge( const double *d, int gauge, double *g ) { if ( threadIdx.x == 0 ) g[ 0 ] = d[ gauge ]; }
__global__ void k_zero_state( double *state ) { if ( threadIdx.x < S_NB_STATE ) state[ threadIdx.x ] = 0; }
__global__ void k_zero_gauge( double *d, int gauge ) { if ( threadIdx.x == 0 ) d[ gauge ] = 0; }

// ---- the hierarchy and its cycles ----------------------------------------------------------------------------------

template<class R>
struct Hierarchy {
    int nl = 0;
    MatT<R> A[ MAX_LEVELS ];
    // per level: the cycle's right-hand side, solution and smoothing buffer; the K-cycle's vectors
    R *x[ MAX_LEVELS ] = {}, *t[ MAX_LEVELS ] = {}, *r[ MAX_LEVELS ] = {};
    R *c1[ MAX_LEVELS ] = {}, *c2[ MAX_LEVELS ] = {}, *v1[ MAX_LEVELS ] = {}, *v2[ MAX_LEVELS ] = {}, *r2[ MAX_LEVELS ] = {};
    double *kpart = nullptr;           ///< [ MAX_LEVELS ][ K_NB ][ RED_BLOCKS ] the K-cycles' partial sums
    double *part = nullptr;            ///< [ 3 ][ RED_BLOCKS ]
    double *inv = nullptr;
    int nc = 0;
    int lb = 0;                        ///< the first level of the bottom kernel ( `nl`: none )
    Bottom<R> bottom;                  ///< its levels `lb` .. `nl - 1`
};

unsigned red_blocks_for( I64 n ) {
    const unsigned b = blocks_for( n, RED_THREADS );
    return b < unsigned( RED_BLOCKS ) ? b : unsigned( RED_BLOCKS );
}

template<class R>
void cycle( Hierarchy<R> &H, int l, const R *r, R *x, const MgOptions &o, const double *state, hipStream_t st );

/// `x ~ A_l^-1 r` on level `l >= 1` by two flexible CG steps preconditioned by the cycle of level `l`
template<class R>
void kcycle( Hierarchy<R> &H, int l, const R *r, R *x, const MgOptions &o, const double *state, hipStream_t st ) {
    const MatT<R> &A = H.A[ l ];
    const unsigned rb = red_blocks_for( A.n );
    double *kp = H.kpart + size_t( l ) * K_NB * RED_BLOCKS;
    cycle( H, l, r, H.c1[ l ], o, state, st );
    k_spmv_dots<R><<<rb, RED_THREADS, 0, st>>>( A, H.c1[ l ], H.v1[ l ], r, nullptr, 2, kp, state );
    k_kc_residual<R><<<rb, RED_THREADS, 0, st>>>( A.n, r, H.v1[ l ], kp, int( rb ), H.r2[ l ], state );
    cycle( H, l, H.r2[ l ], H.c2[ l ], o, state, st );
    k_spmv_dots<R><<<rb, RED_THREADS, 0, st>>>( A, H.c2[ l ], H.v2[ l ], H.v1[ l ], H.r2[ l ], 3, kp + K_BETA2 * RED_BLOCKS, state );
    k_kc_combine<R><<<rb, RED_THREADS, 0, st>>>( A.n, H.c1[ l ], H.c2[ l ], kp, int( rb ), x, state );
}

/// one cycle on level `l`: `x ~ A_l^-1 r` from zero
template<class R>
void cycle( Hierarchy<R> &H, int l, const R *r, R *x, const MgOptions &o, const double *state, hipStream_t st ) {
    if ( l >= H.lb ) {                                   // the bottom levels, one workgroup ( `r` and `x` are its level's own )
        if ( l == H.lb ) k_bottom<R><<<1, BOTTOM_THREADS, 0, st>>>( H.bottom, state );
        else k_dense_apply<R><<<blocks_for( H.nc, 4 ), 256, 0, st>>>( H.nc, H.inv, r, x, state );   // ( `lb` past the coarsest )
        return;
    }
    if ( l == H.nl - 1 ) {
        k_dense_apply<R><<<blocks_for( H.nc, 4 ), 256, 0, st>>>( H.nc, H.inv, r, x, state );
        return;
    }
    const MatT<R> &A = H.A[ l ];
    const unsigned nb = blocks_for( A.n, 256 );
    // pre-smoothing from zero, `nu` sweeps ( the first two in one pass ) ping-ponging so that the last lands in `x`
    const int nk = o.nu >= 2 ? o.nu - 1 : 1;             // the launches
    R *cur = ( nk - 1 ) % 2 ? H.t[ l ] : x, *tmp = ( nk - 1 ) % 2 ? x : H.t[ l ];
    const Damping w = damping_of( o );
    if ( o.nu >= 2 ) k_jacobi_two<R><<<nb, 256, 0, st>>>( A, r, cur, w( 0 ), w( 1 ), state );
    else             k_jacobi_first<R><<<nb, 256, 0, st>>>( A.n, r, A.dia, w( 0 ), cur, state );
    for ( int s = 1; s < nk; ++s ) {
        k_jacobi<R><<<nb, 256, 0, st>>>( A, r, cur, tmp, w( s + 1 ), state );
        std::swap( cur, tmp );
    }
    // the coarse correction: the residual restricted, solved by the next level's K-cycle or cycle, prolonged
    k_restrict<R><<<nb, 256, 0, st>>>( A, r, x, o.pack, H.r[ l + 1 ], state );
    if ( l + 1 <= o.kcycle && l + 1 < H.nl - 1 && l + 1 < H.lb ) kcycle( H, l + 1, H.r[ l + 1 ], H.x[ l + 1 ], o, state, st );
    else                                                          cycle( H, l + 1, H.r[ l + 1 ], H.x[ l + 1 ], o, state, st );
    // post-smoothing: the prolongation fused with the first sweep ( into `t`, since it reads `x` around each row ), or for
    // an odd `nu` prolonged into `t` alone, then the sweeps ping-ponging so that the last lands in `x`
    if ( o.nu % 2 == 0 ) {
        k_prolong_jacobi<R><<<nb, 256, 0, st>>>( A, r, x, H.x[ l + 1 ], o.pack, H.t[ l ], w( 0 ), state );
        R *c2 = H.t[ l ], *t2 = x;
        for ( int s = 1; s < o.nu; ++s ) {
            k_jacobi<R><<<nb, 256, 0, st>>>( A, r, c2, t2, w( s ), state );
            std::swap( c2, t2 );
        }
    } else {
        k_prolong_to<R><<<nb, 256, 0, st>>>( A.n, o.pack, x, H.x[ l + 1 ], H.t[ l ], state );
        R *c2 = H.t[ l ], *t2 = x;
        for ( int s = 0; s < o.nu; ++s ) {
            k_jacobi<R><<<nb, 256, 0, st>>>( A, r, c2, t2, w( s ), state );
            std::swap( c2, t2 );
        }
    }
}

/// THE ITERATIONS AS A GRAPH: a batch of flexible CG iterations is ~100 launches of small kernels, whose issue costs
/// more than their work on the coarse levels; replayed from a HIP graph it is one launch. A graph is valid as long as
/// every kernel argument is: its key is the full record of them ( the hierarchy's levels, vectors and sizes, the options,
/// the storage type ), written field by field ( no padding bytes ) and compared whole, so that two different argument sets
/// never share a graph. The arena keeps the buffers of a solve of the same size, so Newton's successive systems ( two slots
/// of rows ) replay two graphs per storage type. A few per thread are kept ( least recently used out ).
struct GraphCache {
    struct Entry { std::vector<unsigned char> key; hipGraphExec_t exec; hipGraph_t graph; uint64_t used; };
    std::vector<Entry> entries;
    uint64_t clock = 0;
    ~GraphCache() {
        for ( auto &e : entries ) { (void) hipGraphExecDestroy( e.exec ); (void) hipGraphDestroy( e.graph ); }
    }
    hipGraphExec_t find( const std::vector<unsigned char> &key ) {
        for ( auto &e : entries ) if ( e.key == key ) { e.used = ++clock; return e.exec; }
        return nullptr;
    }
    void put( const std::vector<unsigned char> &key, hipGraph_t g, hipGraphExec_t x ) {
        if ( entries.size() >= 8 ) {
            auto old = std::min_element( entries.begin(), entries.end(), []( const Entry &a, const Entry &b ) { return a.used < b.used; } );
            (void) hipGraphExecDestroy( old->exec );
            (void) hipGraphDestroy( old->graph );
            entries.erase( old );
        }
        entries.push_back( Entry{ key, x, g, ++clock } );
    }
};
thread_local GraphCache g_graphs;

/// the stream graphs are captured on ( a capture needs a stream of its own: the legacy default stream cannot be
/// captured ); one per thread, default priority
struct CaptureStream {
    hipStream_t s = nullptr;
    ~CaptureStream() { if ( s ) (void) hipStreamDestroy( s ); }
    hipStream_t get() {
        if ( ! s && hipStreamCreateWithFlags( &s, hipStreamNonBlocking ) != hipSuccess ) { (void) hipGetLastError(); s = nullptr; }
        return s;
    }
};
thread_local CaptureStream g_capture;

/// a graph's key: the values of the kernels' arguments appended one by one ( scalars and pointers, never a struct with
/// padding )
struct Key {
    std::vector<unsigned char> b;
    template<class T> void val( const T &v ) {
        const unsigned char *c = reinterpret_cast<const unsigned char *>( &v );
        b.insert( b.end(), c, c + sizeof( T ) );
    }
    template<class R> void mat( const MatT<R> &A ) {
        val( A.n ); val( A.step ); val( A.row_off ); val( A.row_cnt ); val( A.col ); val( A.val ); val( A.dia );
    }
    template<class R> void bottom( const Bottom<R> &B ) {
        val( B.nl ); val( B.pack ); val( B.nu ); val( B.nc ); val( B.omega.w0 ); val( B.omega.w1 ); val( B.omega.w2 ); val( B.omega.w3 );
        val( B.inv ); val( B.rows );
        for ( int l = 0; l < B.nl; ++l ) {
            const BottomLevel<R> &L = B.L[ l ];
            val( L.n ); val( L.cnt ); val( L.col ); val( L.val ); val( L.dia ); val( L.x ); val( L.t ); val( L.r );
        }
    }
};

/// what one flexible-CG solve reports
struct FcgResult {
    int its = 0;
    bool bad = false, failed = false;    ///< no hierarchy ( a coarse row past its room ); a breakdown ( `p . q <= 0` )
    double err = 0, true_err = 0;        ///< relative residual by the recurrence, and the true one
};

/// `y = x` in another float
template<class X, class Y>
__global__ void k_convert( I64 n, const X *x, Y *y ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) y[ i ] = Y( x[ i ] );
}

/// THE FLEXIBLE CG preconditioned by one multigrid cycle: `A0 d = b` from `d = 0` to the relative residual `tol` in at most
/// `maxit` iterations, the outer vectors and `A0` in `R`, the cycle's levels and vectors in `P` ( from `A0p`, `A0` in `P` ),
/// the scalars in double. With `P = float` under `R = double` the preconditioner is an FP32 operator inside an FP64 Krylov
/// iteration: its rounding only makes it a slightly different preconditioner, which the flexible CG's explicit `z . q`
/// orthogonalisation absorbs. The arena's buffers are named after `pfx`, so the mixed and the FP64 solves keep their own.
template<class R, class P>
FcgResult fcg_mg( const MatT<R> &A0m, const MatT<P> &A0p, I64 nnz, const R *b, R *d, double tol, int maxit, const MgOptions &o,
                  hipStream_t st, int &nb_levels, const std::string &pfx ) {
    Arena &A = arena();
    double *state = A.get<double>( pfx + ".state", S_NB_STATE );
    double *hs = pinned<double>( pfx + ".hstate", S_NB_STATE );
    Hierarchy<P> H;
    H.part = A.get<double>( pfx + ".part", 4 * size_t( RED_BLOCKS ) );
    H.kpart = A.get<double>( pfx + ".kpart", size_t( MAX_LEVELS ) * K_NB * RED_BLOCKS );
    H.A[ 0 ] = A0p;
    H.nl = 1;
    k_zero_state<<<1, 64, 0, st>>>( state );
    // the hierarchy
    for ( int l = 0; H.A[ l ].n > DENSE_MAX && l + 1 < MAX_LEVELS; ++l ) {
        const I64 nc = ( H.A[ l ].n + o.pack - 1 ) / o.pack;
        const std::string tag = pfx + ".l" + std::to_string( l + 1 );
        int *rc = A.get<int>( tag + ".rc", size_t( nc ) );
        int *col = A.get<int>( tag + ".col", size_t( nc ) * MG_STRIDE );
        P *val = A.get<P>( tag + ".val", size_t( nc ) * MG_STRIDE ), *dia = A.get<P>( tag + ".dia", size_t( nc ) );
        k_coarse_op<P, P><<<blocks_for( nc, MG_BLOCK ), MG_BLOCK, 0, st>>>( H.A[ l ], o.pack, nc, rc, col, val, dia, state );
        check_launch( "k_coarse_op" );
        MatT<P> &C = H.A[ l + 1 ];
        C.n = nc; C.step = nc; C.row_off = nullptr; C.row_cnt = rc; C.col = col; C.val = val; C.dia = dia;
        H.nl = l + 2;
    }
    for ( int l = 0; l < H.nl; ++l ) {
        const std::string tag = pfx + ".v" + std::to_string( l );
        const size_t m = size_t( H.A[ l ].n );
        H.t[ l ] = A.get<P>( tag + ".t", m );
        if ( l ) {
            H.x[ l ] = A.get<P>( tag + ".x", m ); H.r[ l ] = A.get<P>( tag + ".r", m );
            if ( l <= o.kcycle && l < H.nl - 1 ) {
                H.c1[ l ] = A.get<P>( tag + ".c1", m ); H.c2[ l ] = A.get<P>( tag + ".c2", m );
                H.v1[ l ] = A.get<P>( tag + ".v1", m ); H.v2[ l ] = A.get<P>( tag + ".v2", m );
                H.r2[ l ] = A.get<P>( tag + ".r2", m );
            }
        }
    }
    H.nc = int( H.A[ H.nl - 1 ].n );
    SDOT_CHECK( H.nc <= DENSE_MAX, "multigrid: the coarsest level is too large for its dense inverse" );
    H.inv = A.get<double>( pfx + ".inv", size_t( DENSE_MAX ) * DENSE_MAX );
    // the bottom kernel: the levels from the first of at most `bottom` rows ( never the finest, whose rows are CSR )
    H.lb = H.nl;
    if ( o.bottom > 0 )
        for ( int l = 1; l < H.nl; ++l )
            if ( H.A[ l ].n <= o.bottom ) { H.lb = l; break; }
    if ( H.lb < H.nl - 1 ) {
        Bottom<P> &B = H.bottom;
        B.nl = H.nl - H.lb; B.pack = o.pack; B.nu = o.nu; B.omega = damping_of( o ); B.nc = H.nc; B.inv = H.inv; B.rows = 0;
        for ( int l = H.lb; l < H.nl; ++l ) {
            BottomLevel<P> &L = B.L[ l - H.lb ];
            L.n = H.A[ l ].n; L.cnt = H.A[ l ].row_cnt; L.col = H.A[ l ].col; L.val = H.A[ l ].val; L.dia = H.A[ l ].dia;
            L.x = H.x[ l ]; L.t = H.t[ l ]; L.r = H.r[ l ];
            B.rows += int( L.n );
        }
    } else
        H.lb = H.nl;                                     // ( a bottom of the coarsest alone: the dense apply as before )
    k_dense_inverse<P><<<1, 256, 0, st>>>( H.A[ H.nl - 1 ], H.inv );
    check_launch( "coarsest inverse" );
    nb_levels = H.nl;

    // the flexible CG
    const I64 n = A0m.n;
    R *r = A.get<R>( pfx + ".r", size_t( n ) ), *z = A.get<R>( pfx + ".z", size_t( n ) );
    R *p = A.get<R>( pfx + ".p", size_t( n ) ), *q = A.get<R>( pfx + ".q", size_t( n ) );
    constexpr bool same = std::is_same<R, P>::value;
    P *rp = nullptr, *zp = nullptr;                      // ( the cycle's own copies of `r` and `z` when it runs in another float )
    if ( ! same ) { rp = A.get<P>( pfx + ".rp", size_t( n ) ); zp = A.get<P>( pfx + ".zp", size_t( n ) ); }
    const MatT<R> &A0 = A0m;
    double *part = H.part;
    // z = M( r )
    auto precondition = [&]( hipStream_t s ) {
        if constexpr ( same ) cycle( H, 0, r, z, o, state, s );
        else {
            k_convert<R, P><<<blocks_for( n, 256 ), 256, 0, s>>>( n, r, rp );
            cycle( H, 0, rp, zp, o, state, s );
            k_convert<P, R><<<blocks_for( n, 256 ), 256, 0, s>>>( n, zp, z );
        }
    };
    k_pcg_init<R><<<RED_BLOCKS, RED_THREADS, 0, st>>>( A0, b, d, r, part );
    k_pcg_init_final<<<1, RED_THREADS, 0, st>>>( part, state, tol );
    // z = M( r ), p = z
    precondition( st );
    k_pcg_first_dir<R><<<blocks_for( n, 256 ), 256, 0, st>>>( n, z, p, state );
    check_launch( "pcg init" );
    const int batch = 4;
    // one batch of iterations on stream `st` ( the first launch of its SpMV timed when `timed` )
    auto iterations = [&]( hipStream_t st, bool timed_first ) {
        for ( int c = 0; c < batch; ++c ) {
            Timing::Pending tp;
            const bool timed = timed_first && c == 0 && timing().begin( TK_SPMV, st, tp );
            k_fcg_spmv<R><<<RED_BLOCKS, RED_THREADS, 0, st>>>( A0, p, r, q, part, state );
            if ( timed ) {
                timing().end( tp, st );
                const double s = double( sizeof( R ) );
                timing().spmv_bytes += double( n ) * ( 4 + 4 + s + 3 * s ) + double( nnz ) * ( 4 + s );
            }
            k_pcg_update<R><<<RED_BLOCKS, RED_THREADS, 0, st>>>( n, p, q, d, r, part, state );
            k_pcg_check<<<1, RED_THREADS, 0, st>>>( part, state );
            precondition( st );
            k_fcg_zq<R><<<RED_BLOCKS, RED_THREADS, 0, st>>>( n, z, q, part, state );
            k_pcg_dir<R><<<RED_BLOCKS, RED_THREADS, 0, st>>>( n, z, p, part, state );
        }
    };
    // the graph of a batch, found by the full record of every argument its kernels take
    hipGraphExec_t exec = nullptr;
    if ( o.graphs ) {
        Key key;
        key.val( int( sizeof( R ) ) ); key.val( int( sizeof( P ) ) );
        key.mat( A0 ); key.val( rp ); key.val( zp ); key.val( H.nl ); key.val( H.lb ); key.val( H.nc ); key.val( H.inv ); key.val( H.part ); key.val( H.kpart );
        for ( int l = 0; l < H.nl; ++l ) {
            key.mat( H.A[ l ] ); key.val( H.x[ l ] ); key.val( H.t[ l ] ); key.val( H.r[ l ] ); key.val( H.c1[ l ] ); key.val( H.c2[ l ] );
            key.val( H.v1[ l ] ); key.val( H.v2[ l ] ); key.val( H.r2[ l ] );
        }
        key.bottom( H.bottom ); key.val( o.pack ); key.val( o.nu ); key.val( o.omega ); key.val( o.chebyshev ); key.val( o.ratio );
        key.val( o.kcycle ); key.val( o.bottom );
        key.val( p ); key.val( q ); key.val( r ); key.val( z ); key.val( d ); key.val( state ); key.val( batch );
        exec = g_graphs.find( key.b );
        hipStream_t cs = exec ? nullptr : g_capture.get();
        if ( ! exec && cs && hipStreamBeginCapture( cs, hipStreamCaptureModeThreadLocal ) == hipSuccess ) {
            hipGraph_t g = nullptr;
            iterations( cs, false );
            const hipError_t e = hipStreamEndCapture( cs, &g );
            if ( e == hipSuccess && g && hipGraphInstantiate( &exec, g, nullptr, nullptr, 0 ) == hipSuccess ) {
                g_graphs.put( key.b, g, exec );
                timing().graph_captures += 1;
            }
            else {                                       // ( no graph: the batches are launched one by one )
                exec = nullptr;
                if ( g ) (void) hipGraphDestroy( g );
            }
        }
        (void) hipGetLastError();
    }
    FcgResult res;
    int launched = 0;
    for ( ;; ) {
        // the first batch launched one by one ( its SpMV timed live ); the next ones replayed from the graph
        if ( launched == 0 || ! exec ) iterations( st, launched == 0 );
        else {
            SDOT_HIP( hipGraphLaunch( exec, st ) );
            timing().graph_replays += 1;
        }
        check_launch( "pcg iterations" );
        launched += batch;
        SDOT_HIP( hipMemcpyAsync( hs, state, sizeof( double ) * S_NB_STATE, hipMemcpyDeviceToHost, st ) );
        SDOT_HIP( hipStreamSynchronize( st ) );
        if ( hs[ S_BAD ] != 0 ) { res.bad = true; return res; }
        if ( hs[ S_DONE ] != 0 || launched >= maxit ) break;
    }
    // the true residual ( the recurrence can drift from it )
    k_true_res<R><<<RED_BLOCKS, RED_THREADS, 0, st>>>( A0, b, d, part );
    k_true_res_final<<<1, RED_THREADS, 0, st>>>( part, state );
    check_launch( "k_true_res" );
    SDOT_HIP( hipMemcpyAsync( hs, state, sizeof( double ) * S_NB_STATE, hipMemcpyDeviceToHost, st ) );
    SDOT_HIP( hipStreamSynchronize( st ) );
    res.its = int( hs[ S_IT ] );
    res.err = hs[ S_ERR ];
    res.true_err = hs[ S_TRUE ];
    res.failed = hs[ S_FAILED ] != 0;
    return res;
}

// ---- the FP32 preconditioner's copy of the rows -----------------------------------------------------------------------

/// the FP32 copy of a Laplacian's values and diagonal ( the same layout: rows at `row_off`, `step` apart )
__global__ void k_to_float_rows( I64 n, const int *row_off, const int *row_cnt, I64 step, const double *val, const double *dia, float *valf,
                                 float *diaf ) {
    const I64 i = xcd_row();
    if ( i >= n ) return;
    diaf[ i ] = float( dia[ i ] );
    const I64 b = row_off[ i ];
    for ( int q = 0; q < row_cnt[ i ]; ++q ) valf[ b + q * step ] = float( val[ b + q * step ] );
}


} // namespace

int mg_solve( const Csr &L, int gauge, const double *b, double *d, double tol, int maxit, const MgOptions &o, hipStream_t st, CgStats &cs,
              int &nb_levels ) {
    SDOT_CHECK( o.pack >= 2 && o.pack <= 64 && ( o.pack & ( o.pack - 1 ) ) == 0, "multigrid: the aggregates must be a power of two from 2 to 64" );
    SDOT_CHECK( o.nu >= 1, "multigrid: at least one smoothing sweep" );
    SDOT_CHECK( ! o.chebyshev || o.nu <= 4, "multigrid: a Chebyshev smoother of at most four sweeps" );
    Arena &A = arena();
    const I64 n = L.n;
    Mat A0;
    A0.n = L.n; A0.step = L.step; A0.row_off = L.row_off; A0.row_cnt = L.row_cnt; A0.col = L.col; A0.val = L.val; A0.dia = L.dia;
    int total = 0;
    bool done = false;
    double err = 1, true_err = 1;
    bool breakdown = false;
    if ( o.fp32 ) {
        // THE FP32 PRECONDITIONER: the hierarchy, its smoothing and its coarse solves in FP32 ( from an FP32 copy of the rows ),
        // the flexible CG, its SpMV and its vectors in FP64; the FP64 preconditioner when that solve does not converge
        float *valf = A.get<float>( "mgm.valf", size_t( L.cap ) ), *diaf = A.get<float>( "mgm.diaf", size_t( n ) );
        k_to_float_rows<<<blocks_for( n, 256 ), 256, 0, st>>>( n, L.row_off, L.row_cnt, L.step, L.val, L.dia, valf, diaf );
        check_launch( "k_to_float_rows" );
        MatT<float> Af;
        Af.n = n; Af.step = L.step; Af.row_off = L.row_off; Af.row_cnt = L.row_cnt; Af.col = L.col; Af.val = valf; Af.dia = diaf;
        const FcgResult f = fcg_mg<double, float>( A0, Af, L.nnz, b, d, tol, maxit, o, st, nb_levels, "mgm" );
        if ( f.bad ) return -2;
        total += f.its;
        err = f.err;
        true_err = f.true_err;
        breakdown = f.failed;
        done = ! f.failed && f.err <= tol && f.true_err <= 10 * tol;
        ++timing().mg32_solves;
        timing().mg32_iterations += f.its;
        if ( ! done ) { cs.nb_fp64_fallbacks += 1; ++timing().mg32_fallbacks; }
    }
    if ( ! done ) {
        const FcgResult f = fcg_mg<double, double>( A0, A0, L.nnz, b, d, tol, std::max( maxit - total, 4 ), o, st, nb_levels, "mg" );
        if ( f.bad ) return -2;
        total += f.its;
        err = f.err;
        true_err = f.true_err;
        breakdown = f.failed;
        done = ! f.failed && f.err <= tol && f.true_err <= 10 * tol;
    }
    double *g = A.get<double>( "mg.g", 1 );
    k_read_gauge<<<1, 64, 0, st>>>( d, gauge, g );
    k_shift_gauge<<<blocks_for( n, 256 ), 256, 0, st>>>( n, d, g, gauge );
    k_zero_gauge<<<1, 64, 0, st>>>( d, gauge );
    check_launch( "k_shift_gauge" );
    timing().resolve();
    cs.nb_iter += total;
    cs.worst = std::max( cs.worst, err );
    cs.last_err = err; cs.last_true = true_err; cs.last_breakdown = breakdown;
    return done ? total : -1;
}

} // namespace sdot_amd
