// The extern "C" entry points declared in include/sdot_amd.h: argument checks, the error record of
// the calling thread, and the conversions between the reference's layouts ( int64 indices, tree in
// preorder ) and the library's own.

#include "../../include/sdot_amd.h"
#include "solver.h"
#include "multigrid.h"
#include "batch.h"
#include "server.h"
#include "memory.h"
#include "noalloc.h"

#include <algorithm>
#include <cmath>

namespace sdot_amd {

namespace {
thread_local Arena *tl_arena = nullptr;
}

Arena &arena() {
    if ( tl_arena ) return *tl_arena;
    static Arena a[ 16 ];
    int dev = 0;
    if ( hipGetDevice( &dev ) != hipSuccess || dev < 0 || dev >= 16 ) dev = 0;
    return a[ dev ];
}

void set_thread_arena( Arena *a ) { tl_arena = a; }
Arena *thread_arena() { return tl_arena; }

Timing &timing() {
    static Timing t;
    return t;
}

template<class T>
T *pinned( const std::string &name, size_t count ) { return arena().host<T>( name, count ); }
template double *pinned<double>( const std::string &, size_t );
template unsigned *pinned<unsigned>( const std::string &, size_t );

namespace {

struct LastError {
    int kind = 0, code = 0;
    long long value = 0;
    std::string msg;
};
thread_local LastError g_err;

template<class F>
int guarded( F &&f ) {
    try {
        f();
        g_err = LastError{};
        return 0;
    } catch ( const Failure &e ) {
        g_err.kind = e.kind; g_err.code = e.code; g_err.value = e.value; g_err.msg = e.what();
        return e.kind;
    } catch ( const std::exception &e ) {
        g_err.kind = 1; g_err.code = 0; g_err.value = 0; g_err.msg = e.what();
        return 1;
    }
}

Density make_density( int kind, double rho, const double *img, int nx, int ny, const double *geo ) {
    Density d;
    d.kind = kind;
    d.rho = rho;
    if ( kind == SDOT_AMD_IMAGE ) {
        SDOT_CHECK( img != nullptr && nx > 0 && ny > 0, "an image density needs its values and a positive shape" );
        SDOT_CHECK( geo[ 2 ] > 0 && geo[ 3 ] > 0, "an image density needs positive pixel sizes" );
        d.img = img; d.nx = nx; d.ny = ny;
        d.x0 = geo[ 0 ]; d.y0 = geo[ 1 ]; d.hx = geo[ 2 ]; d.hy = geo[ 3 ];
    } else if ( kind == SDOT_AMD_GAUSSIANS ) {
        SDOT_CHECK( img != nullptr && nx > 0, "a sum of Gaussians needs at least one Gaussian" );
        d.gauss = img; d.ng = nx;
    } else
        SDOT_CHECK( kind == SDOT_AMD_CONSTANT, "unknown density kind" );
    return d;
}

Density3 make_density3( int kind, double rho, const double *gauss, int ng ) {
    Density3 d;
    d.kind = kind;
    d.rho = rho;
    if ( kind == SDOT_AMD_GAUSSIANS ) {
        SDOT_CHECK( gauss != nullptr && ng > 0, "a sum of Gaussians needs at least one Gaussian" );
        d.gauss = gauss; d.ng = ng;
    } else
        SDOT_CHECK( kind == SDOT_AMD_CONSTANT, "a 3D density is a constant or a sum of Gaussians" );
    return d;
}

/// a voxel image `img[ ( k ny + j ) nx + i ]` ( device ), `geo` HOST [ 6 ]: origin x, y, z, then the voxel's sides
Density3 make_image3( const double *img, int nx, int ny, int nz, const double *geo ) {
    SDOT_CHECK( img && nx >= 1 && ny >= 1 && nz >= 1 && geo, "a 3D image needs its voxels, their counts and its geometry" );
    SDOT_CHECK( geo[ 3 ] > 0 && geo[ 4 ] > 0 && geo[ 5 ] > 0, "a 3D image's voxels have positive sides" );
    Density3 d;
    d.kind = 1;
    d.rho = 0;
    d.img = img; d.nx = nx; d.ny = ny; d.nz = nz;
    d.x0 = geo[ 0 ]; d.y0 = geo[ 1 ]; d.z0 = geo[ 2 ]; d.hx = geo[ 3 ]; d.hy = geo[ 4 ]; d.hz = geo[ 5 ];
    return d;
}

void check_box3( const double *box ) {
    SDOT_CHECK( box != nullptr, "the domain box is required" );
    SDOT_CHECK( box[ 3 ] > box[ 0 ] && box[ 4 ] > box[ 1 ] && box[ 5 ] > box[ 2 ], "the domain box is empty" );
}

void check_box( const double *box ) {
    SDOT_CHECK( box != nullptr, "the domain box is required" );
    SDOT_CHECK( box[ 2 ] > box[ 0 ] && box[ 3 ] > box[ 1 ], "the domain box is empty" );
}

/// An image is zero outside its grid ( the reference's `Image` has the grid as its support, `Image.cxx:106-160` ): the
/// integrals over the cells in `box` are those over the cells in the part of `box` that the grid covers, which the kernels then
/// take as their domain ( seeds may lie outside it ). `out` is that part for an image, `box` for the other densities.
const double *support_box( const Density &dn, const double *box, double *out ) {
    if ( dn.kind != SDOT_AMD_IMAGE ) return box;
    out[ 0 ] = std::max( box[ 0 ], dn.x0 );
    out[ 1 ] = std::max( box[ 1 ], dn.y0 );
    out[ 2 ] = std::min( box[ 2 ], dn.x0 + double( dn.nx ) * dn.hx );
    out[ 3 ] = std::min( box[ 3 ], dn.y0 + double( dn.ny ) * dn.hy );
    SDOT_CHECK( out[ 2 ] > out[ 0 ] && out[ 3 ] > out[ 1 ], "the image's grid and the domain box do not meet" );
    if ( dn.dom.nb ) {                                   // ( a polygon domain starts the cells: it has to lie on the grid )
        const double tx = 1e-12 * ( double( dn.nx ) * dn.hx ), ty = 1e-12 * ( double( dn.ny ) * dn.hy );
        for ( int i = 0; i < dn.dom.nb; ++i )
            SDOT_CHECK( dn.dom.x[ i ] >= dn.x0 - tx && dn.dom.x[ i ] <= dn.x0 + double( dn.nx ) * dn.hx + tx &&
                        dn.dom.y[ i ] >= dn.y0 - ty && dn.dom.y[ i ] <= dn.y0 + double( dn.ny ) * dn.hy + ty,
                        "a polygon domain with an image density lies on the image's grid" );
    }
    return out;
}

/// `support_box` of a 3D voxel image
const double *support_box3( const Density3 &dn, const double *box, double *out ) {
    if ( dn.kind != 1 ) return box;
    const double o[ 3 ] = { dn.x0, dn.y0, dn.z0 }, e[ 3 ] = { dn.x0 + dn.nx * dn.hx, dn.y0 + dn.ny * dn.hy, dn.z0 + dn.nz * dn.hz };
    for ( int d = 0; d < 3; ++d ) {
        out[ d ] = std::max( box[ d ], o[ d ] );
        out[ 3 + d ] = std::min( box[ 3 + d ], e[ d ] );
        SDOT_CHECK( out[ 3 + d ] > out[ d ], "the image's grid and the domain box do not meet" );
    }
    return out;
}

/// heap rank -> preorder node, and each preorder node's height
void preorder_tables( int depth, std::vector<int> &pre, std::vector<int> &height ) {
    const I64 N = ( I64( 1 ) << depth ) - 1;
    pre.assign( size_t( N ), 0 );
    height.assign( size_t( N ), 0 );
    for ( int level = 0; level + 1 < depth; ++level ) {
        const int h = depth - level;
        for ( I64 i = ( I64( 1 ) << level ) - 1; i < ( I64( 1 ) << ( level + 1 ) ) - 1; ++i ) {
            pre[ size_t( 2 * i + 1 ) ] = pre[ size_t( i ) ] + 1;
            pre[ size_t( 2 * i + 2 ) ] = pre[ size_t( i ) ] + int( I64( 1 ) << ( h - 1 ) );
        }
    }
    for ( int level = 0; level < depth; ++level )
        for ( I64 i = ( I64( 1 ) << level ) - 1; i < ( I64( 1 ) << ( level + 1 ) ) - 1; ++i ) height[ size_t( pre[ size_t( i ) ] ) ] = depth - level;
}

__global__ void k_i32_to_i64( I64 n, const int *a, long long *b ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) b[ i ] = a[ i ];
}
__global__ void k_i64_to_i32( I64 n, const long long *a, int *b ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) b[ i ] = int( a[ i ] );
}
__global__ void k_seeds_from( I64 n, const double *pos, const double *w, double4 *seeds ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) seeds[ i ] = make_double4( pos[ 2 * i ], pos[ 2 * i + 1 ], w[ i ], 0.0 );
}
__global__ void k_majorants_out( I64 N, const NodeRec *nodes, double *wa, double *wb ) {
    const I64 m = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( m >= N ) return;
    wa[ 2 * m ] = nodes[ m ].wa_x;
    wa[ 2 * m + 1 ] = nodes[ m ].wa_y;
    wb[ m ] = nodes[ m ].wb;
}
__global__ void k_csr_from( I64 n, const long long *row, const long long *col, int *row_off, int *row_cnt, int *col32, I64 nnz ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) { row_off[ i ] = int( row[ i ] ); row_cnt[ i ] = int( row[ i + 1 ] - row[ i ] ); }
    if ( i < nnz ) col32[ i ] = int( col[ i ] );
}

} // namespace
} // namespace sdot_amd

using namespace sdot_amd;

extern "C" {

int sdot_amd_abi_version( void ) { return SDOT_AMD_ABI_VERSION; }

const char *sdot_amd_last_error( int *kind, int *code, long long *value ) {
    if ( kind ) *kind = g_err.kind;
    if ( code ) *code = g_err.code;
    if ( value ) *value = g_err.value;
    return g_err.msg.c_str();
}

long long sdot_amd_device_bytes( void ) { return (long long) arena().bytes_held(); }

void sdot_amd_release( void ) {
    (void) hipDeviceSynchronize();
    arena().release();
}

int sdot_amd_set_timing( int on ) {
    return guarded( [&] {
        timing().resolve();
        timing().on = on != 0;
    } );
}

int sdot_amd_kernel_times( double *ms, long long *launches, double *spmv_bytes, unsigned long long *events, int reset ) {
    return guarded( [&] {
        Timing &T = timing();
        T.resolve();
        for ( int k = 0; k < TK_NB; ++k ) {
            if ( ms ) ms[ k ] = T.ms[ k ];
            if ( launches ) launches[ k ] = T.launches[ k ];
        }
        if ( spmv_bytes ) *spmv_bytes = T.spmv_bytes;
        unsigned long long *ev = arena().get<unsigned long long>( "events", 8 );
        if ( events ) copy_wait( events, ev, 5 * sizeof( unsigned long long ), hipMemcpyDeviceToHost, side_stream() );
        if ( reset ) {
            T.reset();
            SDOT_HIP( hipMemsetAsync( ev, 0, 8 * sizeof( unsigned long long ), side_stream() ) );
            SDOT_HIP( hipStreamSynchronize( side_stream() ) );
        }
    } );
}

int sdot_amd_fp32_counts( long long *diagrams, long long *redone, long long *dropped ) {
    return guarded( [&] {
        if ( diagrams ) *diagrams = timing().fp32_diagrams.load();
        if ( redone ) *redone = timing().fp32_redone.load();
        if ( dropped ) *dropped = timing().fp32_dropped.load();
    } );
}

int sdot_amd_mg32_counts( long long *solves, long long *iterations, long long *fallbacks ) {
    return guarded( [&] {
        if ( solves ) *solves = timing().mg32_solves.load();
        if ( iterations ) *iterations = timing().mg32_iterations.load();
        if ( fallbacks ) *fallbacks = timing().mg32_fallbacks.load();
    } );
}

int sdot_amd_graph_counts( long long *captures, long long *replays ) {
    return guarded( [&] {
        if ( captures ) *captures = timing().graph_captures.load();
        if ( replays ) *replays = timing().graph_replays.load();
    } );
}

int sdot_amd_measures_2d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                          double rho, const double *img, int nx, int ny, const double *img_geo, double *masses, long long *facet_i,
                          long long *facet_j, double *facet_c, long long facet_cap, long long *nb_facets, unsigned long long *events ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && masses, "measures: positions and masses are required" );
        check_box( box );
        SDOT_CHECK( ! facet_i || ( facet_j && facet_c ), "measures: facet_i, facet_j and facet_c go together" );
        const Density dn = make_density( density_kind, rho, img, nx, ny, img_geo );
        double sb[ 4 ];
        measures_2d( hipStream_t( stream ), n, positions, weights, support_box( dn, box, sb ), dn, masses, facet_i, facet_j, facet_c, facet_cap,
                     nb_facets, events );
    } );
}

static int measures_entry( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                           double rho, const double *img, int nx, int ny, const double *img_geo, double *masses, long long *facet_i,
                           long long *facet_j, double *facet_c, long long facet_cap, long long *nb_facets, unsigned long long *events,
                           int precision, int nb_domain, const double *domain_xy ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && masses, "measures: positions and masses are required" );
        SDOT_CHECK( precision == SDOT_AMD_FP64 || precision == SDOT_AMD_FP32, "measures: the precision is SDOT_AMD_FP64 or SDOT_AMD_FP32" );
        check_box( box );
        SDOT_CHECK( ! facet_i || ( facet_j && facet_c ), "measures: facet_i, facet_j and facet_c go together" );
        Density dn = make_density( density_kind, rho, img, nx, ny, img_geo );
        if ( nb_domain ) set_domain( dn, nb_domain, domain_xy );
        double sb[ 4 ];
        measures_2d( hipStream_t( stream ), n, positions, weights, support_box( dn, box, sb ), dn, masses, facet_i, facet_j, facet_c, facet_cap,
                     nb_facets, events, precision == SDOT_AMD_FP32 );
    } );
}

int sdot_amd_measures_2d_precision( void *stream, long long n, const double *positions, const double *weights, const double *box,
                                    int density_kind, double rho, const double *img, int nx, int ny, const double *img_geo, double *masses,
                                    long long *facet_i, long long *facet_j, double *facet_c, long long facet_cap, long long *nb_facets,
                                    unsigned long long *events, int precision ) {
    return measures_entry( stream, n, positions, weights, box, density_kind, rho, img, nx, ny, img_geo, masses, facet_i, facet_j, facet_c,
                           facet_cap, nb_facets, events, precision, 0, nullptr );
}

int sdot_amd_measures_2d_domain( void *stream, long long n, const double *positions, const double *weights, const double *box,
                                 int density_kind, double rho, const double *img, int nx, int ny, const double *img_geo, double *masses,
                                 long long *facet_i, long long *facet_j, double *facet_c, long long facet_cap, long long *nb_facets,
                                 unsigned long long *events, int precision, int nb_domain, const double *domain_xy ) {
    if ( nb_domain < 3 || ! domain_xy ) return guarded( [&] { SDOT_CHECK( false, "measures: a polygon domain of at least 3 vertices" ); } );
    return measures_entry( stream, n, positions, weights, box, density_kind, rho, img, nx, ny, img_geo, masses, facet_i, facet_j, facet_c,
                           facet_cap, nb_facets, events, precision, nb_domain, domain_xy );
}

int sdot_amd_moments_2d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                         double rho, const double *img, int nx, int ny, const double *img_geo, double *masses, double *first, double *second ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && masses && first && second, "moments: positions, masses, first and second are required" );
        check_box( box );
        const Density dn = make_density( density_kind, rho, img, nx, ny, img_geo );
        double sb[ 4 ];
        moments_2d( hipStream_t( stream ), n, positions, weights, support_box( dn, box, sb ), dn, masses, first, second );
    } );
}

int sdot_amd_moments_3d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                         double rho, const double *gauss, int ng, double *masses, double *first, double *second ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && masses && first && second, "moments: positions, masses, first and second are required" );
        check_box3( box );
        const Density3 dn = make_density3( density_kind, rho, gauss, ng );
        moments_3d( hipStream_t( stream ), n, positions, weights, box, dn, masses, first, second );
    } );
}

int sdot_amd_moments_3d_image( void *stream, long long n, const double *positions, const double *weights, const double *box, const double *img,
                               int nx, int ny, int nz, const double *img_geo, double *masses, double *first, double *second ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && masses && first && second, "moments: positions, masses, first and second are required" );
        check_box3( box );
        const Density3 dn = make_image3( img, nx, ny, nz, img_geo );
        double sb[ 6 ];
        moments_3d( hipStream_t( stream ), n, positions, weights, support_box3( dn, box, sb ), dn, masses, first, second );
    } );
}

int sdot_amd_cells_2d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                       double rho, const double *img, int nx, int ny, const double *img_geo, long long *offsets, double *vertices,
                       long long *cuts, long long vertex_cap, long long *nb_vertices ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && nb_vertices, "cells: positions and nb_vertices are required" );
        SDOT_CHECK( ! vertices || ( offsets && cuts ), "cells: offsets, vertices and cuts go together" );
        check_box( box );
        const Density dn = make_density( density_kind, rho, img, nx, ny, img_geo );
        const I64 nv = cells_2d( hipStream_t( stream ), n, positions, weights, box, dn, offsets, vertices, cuts, vertex_cap );
        *nb_vertices = nv;
        if ( vertices && nv > vertex_cap )
            throw Failure( 3, 0, nv, "cells: " + std::to_string( nv ) + " vertices, past vertex_cap ( " + std::to_string( vertex_cap ) + " )" );
    } );
}

int sdot_amd_measures_vjp_2d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                              double rho, const double *img, int nx, int ny, const double *img_geo, const double *grad_masses,
                              double *grad_weights, double *grad_positions ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && grad_masses && ( grad_weights || grad_positions ),
                    "measures_vjp: positions, grad_masses and a gradient to write are required" );
        check_box( box );
        const Density dn = make_density( density_kind, rho, img, nx, ny, img_geo );
        double sb[ 4 ];
        measures_vjp_2d( hipStream_t( stream ), n, positions, weights, support_box( dn, box, sb ), dn, grad_masses, grad_weights, grad_positions );
    } );
}

int sdot_amd_measures_2d_noalloc( void *stream, long long n, const double *positions, const double *weights, const double *box,
                                  int density_kind, double rho, const double *img, int nx, int ny, const double *img_geo, double *masses,
                                  unsigned long long *events ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && masses, "measures: positions and masses are required" );
        check_box( box );
        const Density dn = make_density( density_kind, rho, img, nx, ny, img_geo );
        double sb[ 4 ];
        measures_2d_noalloc( hipStream_t( stream ), n, positions, weights, support_box( dn, box, sb ), dn, masses, events );
    } );
}

long long sdot_amd_measures_2d_noalloc_max_seeds( void ) { return (long long) NOALLOC_MAX_SEEDS; }

static int solve_entry( void *stream, long long n, const double *positions, const double *nu, const double *w0, const long long *dups,
                        long long nb_dup_values, const double *options, const double *box, int density_kind, double rho, const double *img,
                        double *weights, double *history, long long *nb_steps, double *stats, double *cell_masses, double *barycenters,
                        double *cost, long long *clusters, int precision, int nb_domain, const double *domain_xy ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 2 && positions && nu && options && weights && cell_masses && barycenters, "solve: a required array is null" );
        SDOT_CHECK( precision == SDOT_AMD_FP64 || precision == SDOT_AMD_FP32 || precision == SDOT_AMD_MIXED,
                    "solve: the precision is SDOT_AMD_FP64, SDOT_AMD_FP32 or SDOT_AMD_MIXED" );
        check_box( box );
        SDOT_CHECK( options[ O_MAXIT ] >= 0, "solve: negative maxit" );
        SolveInputs in;
        in.n = n;
        in.positions = positions;
        in.nu = nu;
        in.w0 = w0;
        in.options = options;
        in.precision = precision;
        // the exact duplicates: the first `agg_nb_dups` pairs of `dups` ( the reference passes one dummy pair when there is none )
        in.nb_dups = dups ? std::min<long long>( (long long) options[ O_AGG_NB_DUPS ], nb_dup_values / 2 ) : 0;
        SDOT_CHECK( in.nb_dups >= 0, "solve: a negative count of duplicates" );
        SDOT_CHECK( in.nb_dups == 0 || options[ O_AGG_GAP ] > 0, "solve: duplicates need a positive gap ( option agg_gap )" );
        in.dups = dups;
        for ( int c = 0; c < 4; ++c ) in.box[ c ] = box[ c ];
        const double geo[ 4 ] = { options[ O_IMG_X0 ], options[ O_IMG_Y0 ], options[ O_IMG_HX ], options[ O_IMG_HY ] };
        in.dens = make_density( density_kind, rho, img, int( options[ O_IMG_NX ] ), int( options[ O_IMG_NY ] ), geo );
        if ( nb_domain ) set_domain( in.dens, nb_domain, domain_xy );
        double sb[ 4 ];
        const double *b = support_box( in.dens, box, sb );
        for ( int c = 0; c < 4; ++c ) in.box[ c ] = b[ c ];
        SolveOutputs out;
        out.weights = weights;
        out.cell_masses = cell_masses;
        out.barycenters = barycenters;
        out.clusters = clusters;
        out.history = history;
        out.nb_steps = nb_steps;
        out.stats = stats;
        out.cost = cost;
        solve_2d( in, out, hipStream_t( stream ) );
    } );
}

int sdot_amd_solve_2d_precision( void *stream, long long n, const double *positions, const double *nu, const double *w0, const long long *dups,
                                 long long nb_dup_values, const double *options, const double *box, int density_kind, double rho,
                                 const double *img, double *weights, double *history, long long *nb_steps, double *stats, double *cell_masses,
                                 double *barycenters, double *cost, long long *clusters, int precision ) {
    return solve_entry( stream, n, positions, nu, w0, dups, nb_dup_values, options, box, density_kind, rho, img, weights, history, nb_steps,
                        stats, cell_masses, barycenters, cost, clusters, precision, 0, nullptr );
}

int sdot_amd_solve_2d_domain( void *stream, long long n, const double *positions, const double *nu, const double *w0, const long long *dups,
                              long long nb_dup_values, const double *options, const double *box, int density_kind, double rho,
                              const double *img, double *weights, double *history, long long *nb_steps, double *stats, double *cell_masses,
                              double *barycenters, double *cost, long long *clusters, int precision, int nb_domain, const double *domain_xy ) {
    if ( nb_domain < 3 || ! domain_xy ) return guarded( [&] { SDOT_CHECK( false, "solve: a polygon domain of at least 3 vertices" ); } );
    return solve_entry( stream, n, positions, nu, w0, dups, nb_dup_values, options, box, density_kind, rho, img, weights, history, nb_steps,
                        stats, cell_masses, barycenters, cost, clusters, precision, nb_domain, domain_xy );
}

int sdot_amd_solve_2d( void *stream, long long n, const double *positions, const double *nu, const double *w0, const long long *dups,
                       long long nb_dup_values, const double *options, const double *box, int density_kind, double rho, const double *img,
                       double *weights, double *history, long long *nb_steps, double *stats, double *cell_masses, double *barycenters,
                       double *cost, long long *clusters ) {
    return sdot_amd_solve_2d_precision( stream, n, positions, nu, w0, dups, nb_dup_values, options, box, density_kind, rho, img, weights,
                                        history, nb_steps, stats, cell_masses, barycenters, cost, clusters, SDOT_AMD_FP64 );
}

int sdot_amd_measures_3d( void *stream, long long n, const double *positions, const double *weights, const double *box, int density_kind,
                          double rho, const double *gauss, int ng, double *masses, long long *facet_i, long long *facet_j, double *facet_c,
                          long long facet_cap, long long *nb_facets, unsigned long long *events ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && masses, "measures: positions and masses are required" );
        check_box3( box );
        SDOT_CHECK( ! facet_i || ( facet_j && facet_c ), "measures: facet_i, facet_j and facet_c go together" );
        measures_3d( hipStream_t( stream ), n, positions, weights, box, make_density3( density_kind, rho, gauss, ng ), masses, facet_i, facet_j,
                     facet_c, facet_cap, nb_facets, events );
    } );
}

int sdot_amd_measures_3d_image( void *stream, long long n, const double *positions, const double *weights, const double *box,
                                const double *img, int nx, int ny, int nz, const double *img_geo, double *masses, long long *facet_i,
                                long long *facet_j, double *facet_c, long long facet_cap, long long *nb_facets ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && masses, "measures: positions and masses are required" );
        check_box3( box );
        SDOT_CHECK( ! facet_i || ( facet_j && facet_c ), "measures: facet_i, facet_j and facet_c go together" );
        const Density3 dn = make_image3( img, nx, ny, nz, img_geo );
        double sb[ 6 ];
        measures_3d( hipStream_t( stream ), n, positions, weights, support_box3( dn, box, sb ), dn, masses, facet_i, facet_j, facet_c, facet_cap,
                     nb_facets, nullptr );
    } );
}

static int solve3_entry( void *stream, long long n, const double *positions, const double *nu, const double *w0, const double *options,
                         const double *box, const Density3 &dn, double *weights, double *history, long long *nb_steps, double *stats,
                         double *cell_masses, double *barycenters, double *cost ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 2 && positions && nu && options && weights && cell_masses && barycenters, "solve: a required array is null" );
        check_box3( box );
        SDOT_CHECK( options[ O_MAXIT ] >= 0, "solve: negative maxit" );
        SolveInputs in;
        in.dim = 3;
        in.n = n;
        in.positions = positions;
        in.nu = nu;
        in.w0 = w0;
        in.options = options;
        double sb[ 6 ];
        const double *b = support_box3( dn, box, sb );
        for ( int c = 0; c < 6; ++c ) in.box3[ c ] = b[ c ];
        in.dens3 = dn;
        SolveOutputs out;
        out.weights = weights;
        out.cell_masses = cell_masses;
        out.barycenters = barycenters;
        out.history = history;
        out.nb_steps = nb_steps;
        out.stats = stats;
        out.cost = cost;
        solve_2d( in, out, hipStream_t( stream ) );
    } );
}

int sdot_amd_solve_3d_image( void *stream, long long n, const double *positions, const double *nu, const double *w0, const double *options,
                             const double *box, const double *img, int nx, int ny, int nz, const double *img_geo, double *weights,
                             double *history, long long *nb_steps, double *stats, double *cell_masses, double *barycenters, double *cost ) {
    Density3 dn;
    const int e = guarded( [&] { dn = make_image3( img, nx, ny, nz, img_geo ); } );
    if ( e ) return e;
    return solve3_entry( stream, n, positions, nu, w0, options, box, dn, weights, history, nb_steps, stats, cell_masses, barycenters, cost );
}

int sdot_amd_solve_3d( void *stream, long long n, const double *positions, const double *nu, const double *w0, const double *options,
                       const double *box, int density_kind, double rho, const double *gauss, int ng, double *weights, double *history,
                       long long *nb_steps, double *stats, double *cell_masses, double *barycenters, double *cost ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 2 && positions && nu && options && weights && cell_masses && barycenters, "solve: a required array is null" );
        check_box3( box );
        SDOT_CHECK( options[ O_MAXIT ] >= 0, "solve: negative maxit" );
        SolveInputs in;
        in.dim = 3;
        in.n = n;
        in.positions = positions;
        in.nu = nu;
        in.w0 = w0;
        in.options = options;
        for ( int c = 0; c < 6; ++c ) in.box3[ c ] = box[ c ];
        in.dens3 = make_density3( density_kind, rho, gauss, ng );
        SolveOutputs out;
        out.weights = weights;
        out.cell_masses = cell_masses;
        out.barycenters = barycenters;
        out.history = history;
        out.nb_steps = nb_steps;
        out.stats = stats;
        out.cost = cost;
        solve_2d( in, out, hipStream_t( stream ) );
    } );
}

namespace {
int batch_call( void *stream, int nb_streams, long long k, const long long *offsets, const double *positions, const double *nu, const double *w0,
                const double *options, const double *boxes, int density_kind, const double *rho, const double *const *imgs,
                const double *img_geo, double *weights, double *cell_masses, double *barycenters, double *stats, double *costs,
                long long *errors, double *times, int precision ) {
    return guarded( [&] {
        SDOT_CHECK( precision == SDOT_AMD_FP64 || precision == SDOT_AMD_FP32 || precision == SDOT_AMD_MIXED,
                    "batch: the precision is SDOT_AMD_FP64, SDOT_AMD_FP32 or SDOT_AMD_MIXED" );
        SDOT_CHECK( k >= 1 && offsets && positions && nu && options && boxes && weights && cell_masses && barycenters && stats && costs && errors,
                    "batch: a required array is null" );
        SDOT_CHECK( offsets[ 0 ] == 0, "batch: offsets[ 0 ] must be 0" );
        for ( long long q = 0; q < k; ++q ) {
            SDOT_CHECK( offsets[ q + 1 ] - offsets[ q ] >= 2, "batch: every instance needs at least two seeds" );
            check_box( boxes + 4 * q );
        }
        SDOT_CHECK( options[ O_MAXIT ] >= 0, "batch: negative maxit" );
        SDOT_CHECK( density_kind == SDOT_AMD_CONSTANT || ( imgs && img_geo ), "batch: an image or Gaussians need imgs and img_geo" );
        SDOT_CHECK( density_kind != SDOT_AMD_CONSTANT || rho, "batch: a constant density needs rho" );
        SDOT_HIP( hipStreamSynchronize( hipStream_t( stream ) ) );
        std::vector<BatchJob> jobs( ( size_t( k ) ) );
        std::vector<std::vector<double>> opts( ( size_t( k ) ) );
        for ( long long q = 0; q < k; ++q ) {
            BatchJob &J = jobs[ size_t( q ) ];
            const I64 b = offsets[ q ], n = offsets[ q + 1 ] - b;
            opts[ size_t( q ) ].assign( options, options + NB_OPTS );
            double *o = opts[ size_t( q ) ].data();
            o[ O_AGG_NB_DUPS ] = 0;
            o[ O_AGG_MARGIN ] = 0;
            o[ O_MIN_SCALE ] = 0;                        // ( the continuation's smallest width: from each instance's own density )
            double geo[ 4 ] = { 0, 0, 1, 1 };
            int nx = 0, ny = 0;
            if ( density_kind != SDOT_AMD_CONSTANT ) {
                const double *g = img_geo + 6 * q;
                for ( int c = 0; c < 4; ++c ) { geo[ c ] = g[ c ]; o[ O_IMG_X0 + c ] = g[ c ]; }
                nx = int( g[ 4 ] ); ny = int( g[ 5 ] );
                o[ O_IMG_NX ] = g[ 4 ]; o[ O_IMG_NY ] = g[ 5 ];
            }
            J.in.n = n;
            J.in.positions = positions + 2 * b;
            J.in.nu = nu + b;
            J.in.w0 = w0 ? w0 + b : nullptr;
            J.in.options = o;
            J.in.precision = precision;
            for ( int c = 0; c < 4; ++c ) J.in.box[ c ] = boxes[ 4 * q + c ];
            J.in.dens = make_density( density_kind, rho ? rho[ q ] : 1.0, imgs ? imgs[ q ] : nullptr, nx, ny, geo );
            double sb[ 4 ];
            const double *bq = support_box( J.in.dens, J.in.box, sb );
            for ( int c = 0; c < 4; ++c ) J.in.box[ c ] = bq[ c ];
            J.out.weights = weights + b;
            J.out.cell_masses = cell_masses + b;
            J.out.barycenters = barycenters + 2 * b;
            J.out.stats = stats + size_t( NB_STATS ) * size_t( q );
            J.out.cost = costs + q;
        }
        solve_2d_batch( jobs, nb_streams );
        for ( long long q = 0; q < k; ++q ) {
            errors[ 3 * q ] = jobs[ size_t( q ) ].kind; errors[ 3 * q + 1 ] = jobs[ size_t( q ) ].code; errors[ 3 * q + 2 ] = jobs[ size_t( q ) ].value;
            if ( times ) { times[ 2 * q ] = jobs[ size_t( q ) ].t_start; times[ 2 * q + 1 ] = jobs[ size_t( q ) ].t_end; }
        }
    } );
}
} // namespace

int sdot_amd_solve_2d_batch( void *stream, int nb_streams, long long k, const long long *offsets, const double *positions,
                             const double *nu, const double *w0, const double *options, const double *boxes, int density_kind,
                             const double *rho, const double *const *imgs, const double *img_geo,
                             double *weights, double *cell_masses, double *barycenters, double *stats, double *costs,
                             long long *errors, double *times ) {
    return batch_call( stream, nb_streams, k, offsets, positions, nu, w0, options, boxes, density_kind, rho, imgs, img_geo, weights,
                       cell_masses, barycenters, stats, costs, errors, times, SDOT_AMD_FP64 );
}

int sdot_amd_solve_2d_batch_precision( void *stream, int nb_streams, long long k, const long long *offsets, const double *positions,
                                       const double *nu, const double *w0, const double *options, const double *boxes, int density_kind,
                                       const double *rho, const double *const *imgs, const double *img_geo,
                                       double *weights, double *cell_masses, double *barycenters, double *stats, double *costs,
                                       long long *errors, double *times, int precision ) {
    return batch_call( stream, nb_streams, k, offsets, positions, nu, w0, options, boxes, density_kind, rho, imgs, img_geo, weights,
                       cell_masses, barycenters, stats, costs, errors, times, precision );
}

namespace {
std::mutex g_servers_mtx;
std::map<long long, std::unique_ptr<Server>> g_servers;
long long g_next_server = 1;

Server *find_server( long long h ) {
    std::lock_guard<std::mutex> lk( g_servers_mtx );
    auto it = g_servers.find( h );
    SDOT_CHECK( it != g_servers.end(), "server: unknown handle" );
    return it->second.get();
}
}

long long sdot_amd_server_start( int nb_slots, double memory_budget ) {
    long long h = 0;
    const int rc = guarded( [&] {
        auto s = std::make_unique<Server>( nb_slots, memory_budget );
        std::lock_guard<std::mutex> lk( g_servers_mtx );
        h = g_next_server++;
        g_servers[ h ] = std::move( s );
    } );
    return rc == 0 ? h : 0;
}

/// a request to the server; `image3`: a 3D voxel image ( `density` its voxels, `image3[ 0 .. 2 ]` the counts, `image3_geo` the
/// origin and the sides ) in place of `density_kind`
static int server_submit( long long server, int dim, int precision, long long n, const double *positions, const double *nu, const double *w0,
                          const double *options, const double *box, int density_kind, double rho, const double *density, int nb_density,
                          const int *image3, const double *image3_geo, double *weights, double *cell_masses, double *barycenters,
                          double *stats, double *cost, long long *ticket ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 2 && positions && nu && options && box && weights && cell_masses && barycenters && stats && cost && ticket,
                    "server: a required array is null" );
        SDOT_CHECK( dim == 2 || dim == 3, "server: a 2D or 3D request" );
        SDOT_CHECK( precision == SDOT_AMD_FP64 || precision == SDOT_AMD_FP32 || precision == SDOT_AMD_MIXED,
                    "server: the precision is SDOT_AMD_FP64, SDOT_AMD_FP32 or SDOT_AMD_MIXED" );
        SDOT_CHECK( dim == 2 || precision != SDOT_AMD_FP32, "server: the 3D cells are FP64 ( SDOT_AMD_FP64 or SDOT_AMD_MIXED )" );
        if ( dim == 3 ) check_box3( box );
        else check_box( box );
        SDOT_CHECK( options[ O_MAXIT ] >= 0, "server: negative maxit" );
        Server *S = find_server( server );
        auto r = std::make_unique<Request>();
        r->opts.assign( options, options + NB_OPTS );
        r->opts[ O_AGG_NB_DUPS ] = 0;
        BatchJob &J = r->job;
        J.in.n = n; J.in.positions = positions; J.in.nu = nu; J.in.w0 = w0;
        J.in.dim = dim;
        J.in.precision = dim == 3 ? SDOT_AMD_FP64 : precision;
        if ( dim == 3 ) {
            J.in.dens3 = image3 ? make_image3( density, image3[ 0 ], image3[ 1 ], image3[ 2 ], image3_geo )
                                : make_density3( density_kind, rho, density, nb_density );
            double sb[ 6 ];
            const double *b = support_box3( J.in.dens3, box, sb );
            for ( int c = 0; c < 6; ++c ) J.in.box3[ c ] = b[ c ];
            r->opts[ O_STEP ] = 0;                       // ( 3D steps by trials )
        } else {
            for ( int c = 0; c < 4; ++c ) J.in.box[ c ] = box[ c ];
            const double geo[ 4 ] = { options[ O_IMG_X0 ], options[ O_IMG_Y0 ], options[ O_IMG_HX ], options[ O_IMG_HY ] };
            J.in.dens = make_density( density_kind, rho, density, int( options[ O_IMG_NX ] ), int( options[ O_IMG_NY ] ), geo );
            double sb[ 4 ];
            const double *b = support_box( J.in.dens, box, sb );
            for ( int c = 0; c < 4; ++c ) J.in.box[ c ] = b[ c ];
        }
        J.out.weights = weights; J.out.cell_masses = cell_masses; J.out.barycenters = barycenters; J.out.stats = stats; J.out.cost = cost;
        // the admission's footprint: the byte model of the solve, part by part
        r->bytes = solve_bytes( n, dim, image3 ? SDOT_AMD_IMAGE : density_kind, r->opts.data(), J.in.precision );
        *ticket = S->submit( std::move( r ) );
    } );
}

int sdot_amd_server_submit_nd( long long server, int dim, int precision, long long n, const double *positions, const double *nu,
                               const double *w0, const double *options, const double *box, int density_kind, double rho,
                               const double *density, int nb_density, double *weights, double *cell_masses, double *barycenters,
                               double *stats, double *cost, long long *ticket ) {
    return server_submit( server, dim, precision, n, positions, nu, w0, options, box, density_kind, rho, density, nb_density, nullptr, nullptr,
                          weights, cell_masses, barycenters, stats, cost, ticket );
}

int sdot_amd_server_submit_3d_image( long long server, long long n, const double *positions, const double *nu, const double *w0,
                                     const double *options, const double *box, const double *img, int nx, int ny, int nz,
                                     const double *img_geo, double *weights, double *cell_masses, double *barycenters, double *stats,
                                     double *cost, long long *ticket ) {
    const int counts[ 3 ] = { nx, ny, nz };
    return server_submit( server, 3, SDOT_AMD_FP64, n, positions, nu, w0, options, box, SDOT_AMD_IMAGE, 0.0, img, 0, counts, img_geo, weights,
                          cell_masses, barycenters, stats, cost, ticket );
}

int sdot_amd_server_submit( long long server, long long n, const double *positions, const double *nu, const double *w0,
                            const double *options, const double *box, int density_kind, double rho, const double *img,
                            double *weights, double *cell_masses, double *barycenters, double *stats, double *cost,
                            long long *ticket ) {
    return sdot_amd_server_submit_nd( server, 2, SDOT_AMD_FP64, n, positions, nu, w0, options, box, density_kind, rho, img,
                                      options ? int( options[ O_IMG_NX ] ) : 0, weights, cell_masses, barycenters, stats, cost, ticket );
}

int sdot_amd_server_info( long long server, double *info ) {
    return guarded( [&] {
        SDOT_CHECK( info, "server_info: info is required" );
        Server *S = find_server( server );
        double peak = 0;
        long long done = 0;
        S->info( peak, done );
        info[ 0 ] = peak;
        info[ 1 ] = double( done );
    } );
}

int sdot_amd_server_wait( long long server, long long ticket, long long *error, double *times ) {
    return guarded( [&] {
        Server *S = find_server( server );
        std::unique_ptr<Request> r = S->wait( ticket );
        if ( error ) { error[ 0 ] = r->job.kind; error[ 1 ] = r->job.code; error[ 2 ] = r->job.value; }
        if ( times ) { times[ 0 ] = r->t_arrival; times[ 1 ] = r->job.t_start; times[ 2 ] = r->job.t_end; }
    } );
}

int sdot_amd_server_stop( long long server ) {
    return guarded( [&] {
        std::unique_ptr<Server> s;
        {
            std::lock_guard<std::mutex> lk( g_servers_mtx );
            auto it = g_servers.find( server );
            SDOT_CHECK( it != g_servers.end(), "server: unknown handle" );
            s = std::move( it->second );
            g_servers.erase( it );
        }
        s.reset();
    } );
}

int sdot_amd_solve_bytes( long long n, int dim, int density_kind, const double *options, int precision, double *parts ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && ( dim == 2 || dim == 3 ) && parts, "solve_bytes: n >= 1, dim 2 or 3 and parts are required" );
        solve_parts( n, dim, density_kind, options, precision, parts );
    } );
}

long long sdot_amd_tree_nodes( long long n, int leaf ) { return ( 1ll << max_depth_for( n, leaf ) ) - 1; }

int sdot_amd_build_tree_2d( void *stream, long long n, const double *positions, int leaf, long long *seed_indices, long long *rank_of,
                            long long *node_begin, long long *node_end, long long *node_left, long long *node_right, double *node_box ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && positions && leaf >= 1, "build_tree: positions, n >= 1 and leaf >= 1 are required" );
        hipStream_t st = hipStream_t( stream );
        DevTree t;
        build_tree_device( t, n, positions, leaf, st, "abi_tree" );
        const I64 N = t.nb_nodes;
        std::vector<int> pre, height;
        preorder_tables( t.depth, pre, height );
        std::vector<long long> lft( ( size_t( N ) ) ), rgt( ( size_t( N ) ) );
        for ( I64 m = 0; m < N; ++m ) {
            const int h = height[ size_t( m ) ];
            lft[ size_t( m ) ] = h > 1 ? m + 1 : -1;
            rgt[ size_t( m ) ] = h > 1 ? m + ( 1ll << ( h - 1 ) ) : -1;
        }
        if ( seed_indices ) k_i32_to_i64<<<blocks_for( n, 256 ), 256, 0, st>>>( n, t.order, seed_indices );
        if ( rank_of ) k_i32_to_i64<<<blocks_for( n, 256 ), 256, 0, st>>>( n, t.rank_of, rank_of );
        if ( node_begin ) k_i32_to_i64<<<blocks_for( N, 256 ), 256, 0, st>>>( N, t.node_beg, node_begin );
        if ( node_end ) k_i32_to_i64<<<blocks_for( N, 256 ), 256, 0, st>>>( N, t.node_end, node_end );
        check_launch( "k_i32_to_i64" );
        if ( node_left ) SDOT_HIP( hipMemcpyAsync( node_left, lft.data(), sizeof( long long ) * size_t( N ), hipMemcpyHostToDevice, st ) );
        if ( node_right ) SDOT_HIP( hipMemcpyAsync( node_right, rgt.data(), sizeof( long long ) * size_t( N ), hipMemcpyHostToDevice, st ) );
        if ( node_box ) SDOT_HIP( hipMemcpyAsync( node_box, t.node_box, sizeof( double ) * size_t( N ) * 4, hipMemcpyDeviceToDevice, st ) );
        SDOT_HIP( hipStreamSynchronize( st ) );
    } );
}

int sdot_amd_refresh_majorants_2d( void *stream, long long n, long long nb_nodes, const double *node_box, const long long *node_begin,
                                   const long long *node_end, const double *sorted_positions, const double *sorted_weights, double *node_wa,
                                   double *node_wb ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 1 && node_box && node_begin && node_end && sorted_positions && sorted_weights && node_wa && node_wb,
                    "refresh_majorants: a required array is null" );
        int depth = 0;
        while ( ( 1ll << depth ) - 1 < nb_nodes ) ++depth;
        SDOT_CHECK( ( 1ll << depth ) - 1 == nb_nodes, "refresh_majorants: nb_nodes is not 2^depth - 1" );
        hipStream_t st = hipStream_t( stream );
        Arena &A = arena();
        DevTree t;
        t.n = n;
        t.nb_nodes = nb_nodes;
        t.depth = depth;
        t.nodes = A.get<NodeRec>( "rm.nodes", size_t( nb_nodes ) );
        t.node_box = const_cast<double *>( node_box );
        t.node_beg = A.get<int>( "rm.beg", size_t( nb_nodes ) );
        t.node_end = A.get<int>( "rm.end", size_t( nb_nodes ) );
        t.lvl_nodes = A.get<int>( "rm.lvl", size_t( nb_nodes ) );
        t.seeds = A.get<double4>( "rm.seeds", size_t( n ) );
        t.stat = A.get<double>( "rm.stat", size_t( nb_nodes ) * 11 );
        t.fit = A.get<double>( "rm.fit", size_t( nb_nodes ) * 2 );
        t.rkeys = A.get<unsigned long long>( "rm.rkeys", size_t( nb_nodes ) * 3 );
        std::vector<int> pre, height;
        preorder_tables( depth, pre, height );
        SDOT_HIP( hipMemcpyAsync( t.lvl_nodes, pre.data(), sizeof( int ) * size_t( nb_nodes ), hipMemcpyHostToDevice, st ) );
        k_i64_to_i32<<<blocks_for( nb_nodes, 256 ), 256, 0, st>>>( nb_nodes, node_begin, t.node_beg );
        k_i64_to_i32<<<blocks_for( nb_nodes, 256 ), 256, 0, st>>>( nb_nodes, node_end, t.node_end );
        k_seeds_from<<<blocks_for( n, 256 ), 256, 0, st>>>( n, sorted_positions, sorted_weights, t.seeds );
        check_launch( "refresh_majorants inputs" );
        refresh_majorants( t, st );
        k_majorants_out<<<blocks_for( nb_nodes, 256 ), 256, 0, st>>>( nb_nodes, t.nodes, node_wa, node_wb );
        check_launch( "k_majorants_out" );
        SDOT_HIP( hipStreamSynchronize( st ) );
    } );
}

int sdot_amd_linear_solves_2d( void *stream, long long n, const long long *row, const long long *col, const double *val, const double *dia,
                               const double *rhs, long long k, const double *options, double *solution, double *iterations ) {
    return guarded( [&] {
        SDOT_CHECK( n >= 2 && row && col && val && dia && rhs && solution && k >= 1, "linear_solves: a required array is null" );
        hipStream_t st = hipStream_t( stream );
        long long nnz = 0;
        SDOT_HIP( hipMemcpyAsync( &nnz, row + n, sizeof( long long ), hipMemcpyDeviceToHost, st ) );
        SDOT_HIP( hipStreamSynchronize( st ) );
        Arena &A = arena();
        int *row_off = A.get<int>( "ls.row_off", size_t( n ) ), *row_cnt = A.get<int>( "ls.row_cnt", size_t( n ) );
        int *col32 = A.get<int>( "ls.col", size_t( std::max<long long>( nnz, 1 ) ) );
        k_csr_from<<<blocks_for( std::max<long long>( n, nnz ), 256 ), 256, 0, st>>>( n, row, col, row_off, row_cnt, col32, nnz );
        check_launch( "k_csr_from" );
        Csr L;
        L.n = n; L.row_off = row_off; L.row_cnt = row_cnt; L.col = col32; L.val = val; L.dia = dia; L.nnz = nnz; L.cap = std::max<long long>( nnz, 1 );
        // options ( `SdotPlanNd.py:204-222` ): n, k, method ( 0 cg, 1 mg ), tol, smoothed, float, recycle, rebuild, kcycle, nu
        const double tol = options && options[ 3 ] > 0 ? options[ 3 ] : 1e-6;
        const bool mg = options && options[ 2 ] == 1;
        MgOptions mo;
        mo.fp32 = options && options[ 5 ] == 1;          // ( the multigrid's cycle in FP32, the flexible CG in FP64 )
        if ( options && options[ 8 ] >= 0 ) mo.kcycle = int( options[ 8 ] );
        if ( options && options[ 9 ] > 0 ) mo.nu = int( options[ 9 ] );
        for ( long long j = 0; j < k; ++j ) {
            CgStats cs;
            int levels = 0;
            int its = mg ? mg_solve( L, 0, rhs + j * n, solution + j * n, tol, 2000, mo, st, cs, levels ) : -2;
            if ( its == -2 ) its = cg_solve( L, 0, rhs + j * n, solution + j * n, tol, 20000, st, cs );
            if ( iterations ) iterations[ j ] = its;
        }
        SDOT_HIP( hipStreamSynchronize( st ) );
    } );
}

} // extern "C"
