// ONE DIAGRAM WITHOUT ALLOCATION: the masses of `sdot_amd_measures_2d` for callers that cannot allocate device memory
// ( code under HIP graph capture, tight serving loops, harnesses that forbid `hipMalloc` ). Everything the diagram needs
// besides its inputs and outputs -- the tree and its build's sort buffers, the majorants, the rank-order masses, the
// overflow lists and slots, the counters -- is carved from one file-scope device array, `scratch`, of `NOALLOC_BYTES`
// bytes, which each GPU the library runs on holds until the process ends. The host side carves its staging from a static
// host array the same way, so nothing is allocated, device or pinned host.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's contents: what must start at zero ( the cell
// counters, the event counters, a Voronoi diagram's weights ) is cleared by a kernel at each call. Calls on one GPU are
// serialized by a lock held for the whole call ( the call returns once its results are in place ), so two host threads
// cannot interleave their use of the array; work that the caller queues on another stream while a call runs does not
// touch it.

#include "noalloc.h"
#include "cells.h"

#include <mutex>

namespace sdot_amd {

/// the scratch of the calls of this file: one array per GPU ( a `__device__` array has a copy on each device ). In host code
/// its name is not a device address: only the pointer `hipGetSymbolAddress` gives for it is passed to kernels and copies
static __device__ __align__( 256 ) unsigned char scratch[ NOALLOC_BYTES ];

namespace {

constexpr int MAX_DEVICES = 16;
constexpr size_t HOST_STAGE_BYTES = 64 << 10;

struct PerDevice {
    std::mutex mtx;
    void      *ptr = nullptr;             ///< `scratch` on this device, as `hipGetSymbolAddress` gives it
    Arena      arena;                     ///< carves from `ptr` ( and from `stage` on the host )
    alignas( 256 ) unsigned char stage[ HOST_STAGE_BYTES ];
};
PerDevice g_dev[ MAX_DEVICES ];

__global__ void k_fill_f64( I64 n, double *v, double x ) {
    const I64 k = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( k < n ) v[ k ] = x;
}

__global__ void k_zero_events( unsigned long long *e ) {
    if ( threadIdx.x < 8 ) e[ threadIdx.x ] = 0;
}

__global__ void k_rank_to_user( I64 n, const int *order, const double *rank, double *user ) {
    const I64 k = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( k < n ) user[ order[ k ] ] = rank[ k ];
}

/// this device's slot, its `scratch` address looked up once ( with the device current )
PerDevice &device_slot() {
    int dev = 0;
    SDOT_HIP( hipGetDevice( &dev ) );
    SDOT_CHECK( dev >= 0 && dev < MAX_DEVICES, "measures without allocation: at most 16 devices" );
    return g_dev[ dev ];
}

} // namespace

void measures_2d_noalloc( hipStream_t st, I64 n, const double *positions, const double *weights, const double *box, const Density &dn,
                          double *masses, unsigned long long *events ) {
    SDOT_CHECK( n >= 1, "a diagram needs at least one seed" );
    SDOT_CHECK( n <= NOALLOC_MAX_SEEDS, "measures without allocation: at most " + std::to_string( NOALLOC_MAX_SEEDS ) + " seeds ( " +
                                            std::to_string( n ) + " given ); its scratch is a fixed array of " +
                                            std::to_string( NOALLOC_BYTES >> 20 ) + " MiB per GPU" );
    PerDevice &D = device_slot();
    std::lock_guard<std::mutex> lk( D.mtx );
    if ( ! D.ptr ) {
        void *p = nullptr;
        SDOT_HIP( hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) ) );
        SDOT_CHECK( p != nullptr, "measures without allocation: the scratch array has no device address" );
        D.ptr = p;
    }
    // every buffer carved afresh from the start of the array ( the sizes differ from call to call )
    D.arena.fixed( D.ptr, NOALLOC_BYTES, D.stage, HOST_STAGE_BYTES );
    Arena *prev = thread_arena();
    set_thread_arena( &D.arena );
    struct Restore { Arena *a; ~Restore() { set_thread_arena( a ); } } restore{ prev };

    Arena &A = D.arena;
    DevTree t;
    build_tree_device( t, n, positions, 10, st, "na" );
    const unsigned nbk = blocks_for( n, 256 );
    if ( weights ) set_weights_user( t, weights, st );
    else {
        double *z = A.get<double>( "na.zero", size_t( n ) );
        k_fill_f64<<<nbk, 256, 0, st>>>( n, z, 0.0 );
        check_launch( "k_fill_f64" );
        set_weights_rank( t, z, st );
    }
    refresh_majorants( t, st );
    DiagramOut out;
    out.mass = A.get<double>( "na.mass", size_t( n ) );
    out.counters = A.get<unsigned>( "na.counters", 8 );
    out.ovf_list = A.get<int>( "na.ovf_list", 2 * size_t( n ) );
    out.events = A.get<unsigned long long>( "na.events", 8 );
    k_zero_events<<<1, 64, 0, st>>>( out.events );
    check_launch( "k_zero_events" );
    const double lo[ 2 ] = { box[ 0 ], box[ 1 ] }, hi[ 2 ] = { box[ 2 ], box[ 3 ] };
    run_cells( t, lo, hi, dn, MODE_MASSES, out, st, false, NOALLOC_OVF_LANES );
    k_rank_to_user<<<nbk, 256, 0, st>>>( n, t.order, out.mass, masses );
    check_launch( "k_rank_to_user" );
    unsigned *cnt = A.host<unsigned>( "na.hcnt", 8 );
    SDOT_HIP( hipMemcpyAsync( cnt, out.counters, sizeof( unsigned ) * 8, hipMemcpyDeviceToHost, st ) );
    unsigned long long *hev = A.host<unsigned long long>( "na.hev", 8 );
    if ( events ) SDOT_HIP( hipMemcpyAsync( hev, out.events, sizeof( unsigned long long ) * 5, hipMemcpyDeviceToHost, st ) );
    SDOT_HIP( hipStreamSynchronize( st ) );
    if ( events ) for ( int q = 0; q < 5; ++q ) events[ q ] = hev[ q ];
    if ( cnt[ 2 ] ) {
        int *user = A.host<int>( "na.user", 1 );
        SDOT_HIP( hipMemcpyAsync( user, t.order + cnt[ 3 ], sizeof( int ), hipMemcpyDeviceToHost, st ) );
        SDOT_HIP( hipStreamSynchronize( st ) );
        throw Failure( 2, 1, *user, "a power cell needs more than " + std::to_string( CAP3 ) + " vertices ( seed " + std::to_string( *user ) + " )" );
    }
}

} // namespace sdot_amd
