// The Newton direction by CG preconditioned with an aggregation multigrid.
//
// The reference solves `L d = b` with a multigrid whose aggregates follow the BSP tree's order
// ( `src/sdot/include/sdot/sdotplan/Multigrid.h:11-18`: consecutive tree ranks are neighbours in
// space, so packs of consecutive ranks are compact aggregates; the card's version
// `gpu/Linear2D.cuh:1104-1506`, packs of 4 in 2D ). This one is its own design for CDNA4:
//   * aggregates of `pack` consecutive ranks: restriction and prolongation are index arithmetic
//     ( `I = i / pack` ), nothing to store;
//   * Galerkin coarse Laplacians `P^T L P` ( P piecewise constant ), one lane per coarse row merging the
//     `pack` fine rows in LDS, rows stored with a fixed stride so a level is built in one launch;
//   * a symmetric V-cycle: damped Jacobi before and after, the coarsest level ( <= 256 rows ) solved by
//     a dense inverse of `L_c + alpha 1 1^T` ( the rank-one term removes the constants, the kernel of a
//     Laplacian ), computed by Gauss-Jordan in one workgroup per Newton iteration;
//   * CG in the zero-mean gauge: constants are invisible to `L`, so the iteration needs no projection,
//     and the solution is shifted at the end to the reference's gauge `d[ gauge ] = 0`;
//   * every scalar of the CG stays on the device ( `alpha`, `beta`, the stop flag that turns the
//     remaining launches of a batch into no-ops ), the host reads the state once per batch.
// Deterministic: fixed-shape reductions, each coarse row merged in a fixed order.

#include "multigrid.h"

namespace sdot_amd {

namespace {

constexpr int MG_STRIDE = 48;      ///< neighbours of a coarse row, at most
constexpr int MG_BLOCK = 64;
constexpr int MAX_LEVELS = 24;

enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_NB, S_ERR, S_DONE, S_IT, S_TOL, S_FAILED, S_BAD, S_NB_STATE };

struct Mat {
    I64 n = 0;
    const int *row_off = nullptr, *row_cnt = nullptr, *col = nullptr;
    const double *val = nullptr, *dia = nullptr;
};

__device__ __forceinline__ double row_times( const Mat &A, I64 i, const double *x ) {
    double s = A.dia[ i ] * x[ i ];
    const int b = A.row_off[ i ], e = b + A.row_cnt[ i ];
    for ( int q = b; q < e; ++q ) s -= A.val[ q ] * x[ A.col[ q ] ];
    return s;
}

/// the Galerkin product of one coarse row: the `pack` fine rows merged by coarse column
__global__ __launch_bounds__( MG_BLOCK ) void k_coarse_op( Mat F, int pack, I64 nc, int *row_off, int *row_cnt, int *col, double *val, double *dia,
                                                         double *state ) {
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
        d += F.dia[ i ];
        const int b = F.row_off[ i ], e = b + F.row_cnt[ i ];
        for ( int q = b; q < e; ++q ) {
            const int jc = F.col[ q ] / pack;
            const double c = F.val[ q ];
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
    row_off[ I ] = int( I * MG_STRIDE );
    row_cnt[ I ] = u;
    for ( int k = 0; k < u; ++k ) {
        col[ I * MG_STRIDE + k ] = J[ k * MG_BLOCK ];
        val[ I * MG_STRIDE + k ] = C[ k * MG_BLOCK ];
    }
    dia[ I ] = d > 0 ? d : 1.0;
    if ( bad ) state[ S_BAD ] = 1;
}

constexpr int DENSE_MAX = 64;       ///< the coarsest level, solved by a dense inverse

/// the inverse of `L_c + alpha 1 1^T` ( `alpha` the mean diagonal over `n`: the rank-one term removes the
/// constants, the kernel of a Laplacian ), by Gauss-Jordan in LDS, one workgroup; writes it row-major
__global__ __launch_bounds__( 256 ) void k_dense_inverse( Mat C, double *inv ) {
    __shared__ double M[ DENSE_MAX ][ DENSE_MAX ], V[ DENSE_MAX ][ DENSE_MAX ];
    __shared__ double sh[ 256 ];
    const int n = int( C.n );
    double s = 0;
    for ( int i = threadIdx.x; i < n; i += blockDim.x ) s += C.dia[ i ];
    const double alpha = block_sum<256>( s, sh ) / double( n ) / double( n );
    for ( int k = threadIdx.x; k < n * n; k += blockDim.x ) { M[ k / n ][ k % n ] = alpha; V[ k / n ][ k % n ] = ( k / n == k % n ) ? 1.0 : 0.0; }
    __syncthreads();
    if ( int( threadIdx.x ) < n ) {
        const int i = threadIdx.x;
        M[ i ][ i ] += C.dia[ i ];
        const int b = C.row_off[ i ], e = b + C.row_cnt[ i ];
        for ( int q = b; q < e; ++q ) M[ i ][ C.col[ q ] ] -= C.val[ q ];
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

// ---- the V-cycle --------------------------------------------------------------------------------------------

__global__ void k_jacobi_first( I64 n, const double *r, const double *dia, double omega, double *x, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) x[ i ] = omega * r[ i ] / dia[ i ];
}

__global__ void k_jacobi( Mat A, const double *r, const double *x, double *xo, double omega, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < A.n ) xo[ i ] = x[ i ] + omega * ( r[ i ] - row_times( A, i, x ) ) / A.dia[ i ];
}

__global__ void k_copy( I64 n, const double *a, double *b, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) b[ i ] = a[ i ];
}

/// `rc[ I ] = sum_{ i in I } ( r - A x )_i`: one lane per fine row, the `pack` consecutive lanes of an aggregate
/// summed by shuffles ( `pack` divides 64 )
__global__ void k_restrict( Mat A, const double *r, const double *x, int pack, I64 nc, double *rc, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    double v = i < A.n ? r[ i ] - row_times( A, i, x ) : 0.0;
    for ( int off = pack >> 1; off > 0; off >>= 1 ) v += __shfl_down( v, off, pack );
    if ( i < A.n && ( i % pack ) == 0 ) rc[ i / pack ] = v;
}

__global__ void k_prolong( I64 n, int pack, const double *xc, double *x, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) x[ i ] += xc[ i / pack ];
}

/// `x = Inv r`, one wave per row
__global__ void k_dense_apply( int n, const double *inv, const double *r, double *x, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const int row = blockIdx.x * ( blockDim.x / 64 ) + threadIdx.x / 64, lane = threadIdx.x & 63;
    if ( row >= n ) return;
    double s = lane < n ? inv[ size_t( row ) * n + lane ] * r[ lane ] : 0.0;
    for ( int off = 32; off > 0; off >>= 1 ) s += __shfl_down( s, off, 64 );
    if ( lane == 0 ) x[ row ] = s;
}

// ---- the CG around it -------------------------------------------------------------------------------------------

template<int T>
__device__ __forceinline__ void write_partial( double v, double *part, int slot ) {
    __shared__ double sh[ T ];
    const double r = block_sum<T>( v, sh );
    if ( threadIdx.x == 0 ) part[ slot ] = r;
}

__global__ __launch_bounds__( RED_THREADS ) void k_pcg_init( Mat A, const double *b, double *d, double *r, double *part ) {
    double acc = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < A.n; i += I64( RED_BLOCKS ) * RED_THREADS ) {
        d[ i ] = 0;
        r[ i ] = b[ i ];
        acc += b[ i ] * b[ i ];
    }
    write_partial<RED_THREADS>( acc, part, blockIdx.x );
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
        state[ S_RZ ] = 0;
    }
}

/// `rz = r . z` ( partials ); `p = z` on the first use ( `first` ), else `p = z + beta p` comes after the final
__global__ __launch_bounds__( RED_THREADS ) void k_pcg_dot( I64 n, const double *r, const double *z, double *part, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    double acc = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < n; i += I64( RED_BLOCKS ) * RED_THREADS ) acc += r[ i ] * z[ i ];
    write_partial<RED_THREADS>( acc, part, blockIdx.x );
}

__global__ __launch_bounds__( RED_THREADS ) void k_pcg_beta( const double *part, double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    double a = 0;
    for ( int j = threadIdx.x; j < RED_BLOCKS; j += RED_THREADS ) a += part[ j ];
    const double rz2 = block_sum<RED_THREADS>( a, sh );
    if ( threadIdx.x == 0 ) {
        state[ S_BETA ] = state[ S_RZ ] > 0 ? rz2 / state[ S_RZ ] : 0.0;
        state[ S_RZ ] = rz2;
    }
}

__global__ void k_pcg_dir( I64 n, const double *z, double *p, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) p[ i ] = z[ i ] + state[ S_BETA ] * p[ i ];
}

__global__ __launch_bounds__( RED_THREADS ) void k_pcg_spmv( Mat A, const double *p, double *q, double *part, const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    double acc = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < A.n; i += I64( RED_BLOCKS ) * RED_THREADS ) {
        const double s = row_times( A, i, p );
        q[ i ] = s;
        acc += p[ i ] * s;
    }
    write_partial<RED_THREADS>( acc, part, blockIdx.x );
}

__global__ __launch_bounds__( RED_THREADS ) void k_pcg_alpha( const double *part, double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    double a = 0;
    for ( int j = threadIdx.x; j < RED_BLOCKS; j += RED_THREADS ) a += part[ j ];
    const double pq = block_sum<RED_THREADS>( a, sh );
    if ( threadIdx.x == 0 ) {
        state[ S_PQ ] = pq;
        if ( ! ( pq > 0 ) ) { state[ S_DONE ] = 1; state[ S_FAILED ] = 1; }
        else state[ S_ALPHA ] = state[ S_RZ ] / pq;
    }
}

__global__ __launch_bounds__( RED_THREADS ) void k_pcg_update( I64 n, const double *p, const double *q, double *d, double *r, double *part,
                                                             const double *state ) {
    if ( state[ S_DONE ] != 0 ) return;
    const double alpha = state[ S_ALPHA ];
    double acc = 0;
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < n; i += I64( RED_BLOCKS ) * RED_THREADS ) {
        d[ i ] += alpha * p[ i ];
        const double ri = r[ i ] - alpha * q[ i ];
        r[ i ] = ri;
        acc += ri * ri;
    }
    write_partial<RED_THREADS>( acc, part, blockIdx.x );
}

__global__ __launch_bounds__( RED_THREADS ) void k_pcg_check( const double *part, double *state ) {
    __shared__ double sh[ RED_THREADS ];
    if ( state[ S_DONE ] != 0 ) return;
    double a = 0;
    for ( int j = threadIdx.x; j < RED_BLOCKS; j += RED_THREADS ) a += part[ j ];
    const double rr = block_sum<RED_THREADS>( a, sh );
    if ( threadIdx.x == 0 ) {
        state[ S_RR ] = rr;
        const double err = sqrt( rr ) / state[ S_NB ];
        state[ S_ERR ] = err;
        state[ S_IT ] += 1;
        if ( err <= state[ S_TOL ] ) state[ S_DONE ] = 1;
    }
}

__global__ void k_shift_gauge( I64 n, double *d, const double *g, int gauge ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n && i != gauge ) d[ i ] -= g[ 0 ];
}

__global__ void k_read_gauge( const double *d, int gauge, double *g ) { if ( threadIdx.x == 0 ) g[ 0 ] = d[ gauge ]; }
__global__ void k_zero_gauge( double *d, int gauge ) { if ( threadIdx.x == 0 ) d[ gauge ] = 0; }

struct Hierarchy {
    int nl = 0;
    Mat A[ MAX_LEVELS ];
    double *x[ MAX_LEVELS ] = {}, *t[ MAX_LEVELS ] = {}, *r[ MAX_LEVELS ] = {};
    double *inv = nullptr;
    int nc = 0;
};

/// one cycle on level `l`: `x[ l ] ~ A_l^-1 r[ l ]` from zero; the next level is visited twice on the
/// top `o.w_levels` levels ( a W-cycle there ), once below
void cycle( Hierarchy &H, int l, const double *const *r, double *const *x, const MgOptions &o, const double *state, hipStream_t st ) {
    if ( l == H.nl - 1 ) {
        k_dense_apply<<<blocks_for( H.nc, 4 ), 256, 0, st>>>( H.nc, H.inv, r[ l ], x[ l ], state );
        return;
    }
    const Mat &A = H.A[ l ];
    const unsigned nb = blocks_for( A.n, 256 );
    // pre-smoothing from zero: the first sweep, then `nu - 1` sweeps ping-ponging so that the last lands in `x[ l ]`
    double *cur = ( o.nu - 1 ) % 2 ? H.t[ l ] : x[ l ], *tmp = ( o.nu - 1 ) % 2 ? x[ l ] : H.t[ l ];
    k_jacobi_first<<<nb, 256, 0, st>>>( A.n, r[ l ], A.dia, o.omega, cur, state );
    for ( int s = 1; s < o.nu; ++s ) {
        k_jacobi<<<nb, 256, 0, st>>>( A, r[ l ], cur, tmp, o.omega, state );
        std::swap( cur, tmp );
    }
    // the coarse corrections ( two in a row on the W levels: the square of an A-symmetric correction is one )
    const int visits = l < o.w_levels ? 2 : 1;
    for ( int v = 0; v < visits; ++v ) {
        const I64 nc = H.A[ l + 1 ].n;
        k_restrict<<<nb, 256, 0, st>>>( A, r[ l ], x[ l ], o.pack, nc, H.r[ l + 1 ], state );
        cycle( H, l + 1, r, x, o, state, st );
        k_prolong<<<nb, 256, 0, st>>>( A.n, o.pack, x[ l + 1 ], x[ l ], state );
    }
    // post-smoothing, the pre-smoothing's mirror
    double *c2 = x[ l ], *t2 = H.t[ l ];
    for ( int s = 0; s < o.nu; ++s ) {
        k_jacobi<<<nb, 256, 0, st>>>( A, r[ l ], c2, t2, o.omega, state );
        std::swap( c2, t2 );
    }
    if ( c2 != x[ l ] ) k_copy<<<nb, 256, 0, st>>>( A.n, c2, x[ l ], state );
}

/// `z = V( r )` on level 0, every launch a no-op once the CG has stopped
void vcycle( Hierarchy &H, const double *r0, double *z0, const MgOptions &o, const double *state, hipStream_t st ) {
    const double *r[ MAX_LEVELS ];
    double *x[ MAX_LEVELS ];
    for ( int l = 0; l < H.nl; ++l ) { r[ l ] = l ? H.r[ l ] : r0; x[ l ] = l ? H.x[ l ] : z0; }
    cycle( H, 0, r, x, o, state, st );
    check_launch( "vcycle" );
}

} // namespace

int mg_solve( const Csr &L, int gauge, const double *b, double *d, double tol, int maxit, const MgOptions &o, hipStream_t st, CgStats &cs,
              int &nb_levels ) {
    Arena &A = arena();
    double *state = A.get<double>( "mg.state", S_NB_STATE );
    double *part = A.get<double>( "mg.part", RED_BLOCKS );
    double *hs = pinned<double>( "mg.hstate", S_NB_STATE );
    Hierarchy H;
    H.A[ 0 ].n = L.n; H.A[ 0 ].row_off = L.row_off; H.A[ 0 ].row_cnt = L.row_cnt; H.A[ 0 ].col = L.col; H.A[ 0 ].val = L.val; H.A[ 0 ].dia = L.dia;
    H.nl = 1;
    SDOT_HIP( hipMemsetAsync( state, 0, sizeof( double ) * S_NB_STATE, st ) );
    // the hierarchy
    for ( int l = 0; H.A[ l ].n > DENSE_MAX && l + 1 < MAX_LEVELS; ++l ) {
        const I64 nc = ( H.A[ l ].n + o.pack - 1 ) / o.pack;
        const std::string tag = "mg.l" + std::to_string( l + 1 );
        int *ro = A.get<int>( tag + ".ro", size_t( nc ) ), *rc = A.get<int>( tag + ".rc", size_t( nc ) );
        int *col = A.get<int>( tag + ".col", size_t( nc ) * MG_STRIDE );
        double *val = A.get<double>( tag + ".val", size_t( nc ) * MG_STRIDE ), *dia = A.get<double>( tag + ".dia", size_t( nc ) );
        k_coarse_op<<<blocks_for( nc, MG_BLOCK ), MG_BLOCK, 0, st>>>( H.A[ l ], o.pack, nc, ro, rc, col, val, dia, state );
        check_launch( "k_coarse_op" );
        Mat &C = H.A[ l + 1 ];
        C.n = nc; C.row_off = ro; C.row_cnt = rc; C.col = col; C.val = val; C.dia = dia;
        H.nl = l + 2;
    }
    for ( int l = 0; l < H.nl; ++l ) {
        const std::string tag = "mg.v" + std::to_string( l );
        H.t[ l ] = A.get<double>( tag + ".t", size_t( H.A[ l ].n ) );
        if ( l ) { H.x[ l ] = A.get<double>( tag + ".x", size_t( H.A[ l ].n ) ); H.r[ l ] = A.get<double>( tag + ".r", size_t( H.A[ l ].n ) ); }
    }
    H.nc = int( H.A[ H.nl - 1 ].n );
    SDOT_CHECK( H.nc <= DENSE_MAX, "multigrid: the coarsest level is too large for its dense inverse" );
    H.inv = A.get<double>( "mg.inv", size_t( DENSE_MAX ) * DENSE_MAX );
    k_dense_inverse<<<1, 256, 0, st>>>( H.A[ H.nl - 1 ], H.inv );
    check_launch( "coarsest inverse" );
    nb_levels = H.nl;

    // the CG
    const I64 n = L.n;
    double *r = A.get<double>( "mg.r", size_t( n ) ), *z = A.get<double>( "mg.z", size_t( n ) );
    double *p = A.get<double>( "mg.p", size_t( n ) ), *q = A.get<double>( "mg.q", size_t( n ) );
    const Mat &A0 = H.A[ 0 ];
    k_pcg_init<<<RED_BLOCKS, RED_THREADS, 0, st>>>( A0, b, d, r, part );
    k_pcg_init_final<<<1, RED_THREADS, 0, st>>>( part, state, tol );
    // z = V( r ), rz = r . z, p = z ( beta = 0 on the first pass: `S_RZ` is 0 )
    vcycle( H, r, z, o, state, st );
    k_pcg_dot<<<RED_BLOCKS, RED_THREADS, 0, st>>>( n, r, z, part, state );
    k_pcg_beta<<<1, RED_THREADS, 0, st>>>( part, state );
    SDOT_HIP( hipMemsetAsync( p, 0, sizeof( double ) * size_t( n ), st ) );
    k_pcg_dir<<<blocks_for( n, 256 ), 256, 0, st>>>( n, z, p, state );
    check_launch( "pcg init" );
    const int batch = 4;
    int launched = 0;
    for ( ;; ) {
        for ( int c = 0; c < batch; ++c ) {
            Timing::Pending tp;
            const bool timed = c == 0 && timing().begin( TK_SPMV, st, tp );
            k_pcg_spmv<<<RED_BLOCKS, RED_THREADS, 0, st>>>( A0, p, q, part, state );
            if ( timed ) {
                timing().end( tp, st );
                timing().spmv_bytes += double( n ) * ( 4 + 4 + 8 + 8 + 8 ) + double( L.nnz ) * ( 4 + 8 );
            }
            k_pcg_alpha<<<1, RED_THREADS, 0, st>>>( part, state );
            k_pcg_update<<<RED_BLOCKS, RED_THREADS, 0, st>>>( n, p, q, d, r, part, state );
            k_pcg_check<<<1, RED_THREADS, 0, st>>>( part, state );
            vcycle( H, r, z, o, state, st );
            k_pcg_dot<<<RED_BLOCKS, RED_THREADS, 0, st>>>( n, r, z, part, state );
            k_pcg_beta<<<1, RED_THREADS, 0, st>>>( part, state );
            k_pcg_dir<<<blocks_for( n, 256 ), 256, 0, st>>>( n, z, p, state );
        }
        check_launch( "pcg iterations" );
        launched += batch;
        SDOT_HIP( hipMemcpyAsync( hs, state, sizeof( double ) * S_NB_STATE, hipMemcpyDeviceToHost, st ) );
        SDOT_HIP( hipStreamSynchronize( st ) );
        if ( hs[ S_BAD ] != 0 ) return -2;
        if ( hs[ S_DONE ] != 0 || launched >= maxit ) break;
    }
    double *g = A.get<double>( "mg.g", 1 );
    k_read_gauge<<<1, 64, 0, st>>>( d, gauge, g );
    k_shift_gauge<<<blocks_for( n, 256 ), 256, 0, st>>>( n, d, g, gauge );
    k_zero_gauge<<<1, 64, 0, st>>>( d, gauge );
    check_launch( "k_shift_gauge" );
    timing().resolve();
    const int its = int( hs[ S_IT ] );
    cs.nb_iter += its;
    cs.worst = std::max( cs.worst, hs[ S_ERR ] );
    return ( hs[ S_FAILED ] == 0 && hs[ S_ERR ] <= tol ) ? its : -1;
}

} // namespace sdot_amd
