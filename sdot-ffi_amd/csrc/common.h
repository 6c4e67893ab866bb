// Shared definitions of the MI355X semi-discrete OT library: error plumbing, the device arena,
// launch helpers and the small device utilities every kernel file uses.
#pragma once

#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace sdot_amd {

using I64 = long long;

// ---- errors ------------------------------------------------------------------------------------

/// what crosses the C ABI on failure: `kind` 1 = HIP / argument error, 2 = a kernel failure
/// ( `code` 1: a cell past the vertex limit, `value` its seed ), 3 = capacity, 4 = out of memory
struct Failure : std::runtime_error {
    int kind, code;
    I64 value;
    Failure( int kind, int code, I64 value, const std::string &msg ) : std::runtime_error( msg ), kind( kind ), code( code ), value( value ) {}
};

#define SDOT_HIP( expr )                                                                                     \
    do {                                                                                                     \
        hipError_t e_ = ( expr );                                                                            \
        if ( e_ != hipSuccess )                                                                              \
            throw ::sdot_amd::Failure( 1, int( e_ ), 0, std::string( #expr " -> " ) + hipGetErrorString( e_ ) ); \
    } while ( 0 )

#define SDOT_CHECK( cond, msg )                                      \
    do {                                                             \
        if ( ! ( cond ) ) throw ::sdot_amd::Failure( 1, 0, 0, msg ); \
    } while ( 0 )

inline void check_launch( const char *name ) {
    hipError_t e = hipGetLastError();
    if ( e != hipSuccess )
        throw Failure( 1, int( e ), 0, std::string( "launch of " ) + name + " -> " + hipGetErrorString( e ) );
}

/// a copy on `st`, waited for. The library makes no synchronous `hipMemcpy` / `hipMemset`: those go through the legacy
/// default stream, which HIP refuses ( hipErrorStreamCaptureImplicit ) while any thread of the process captures a graph,
/// as the solves of a batch or a server do on their own threads
inline void copy_wait( void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t st ) {
    SDOT_HIP( hipMemcpyAsync( dst, src, bytes, kind, st ) );
    SDOT_HIP( hipStreamSynchronize( st ) );
}

/// a non-blocking stream of the calling thread, for the one-off copies and fills made outside a solve
inline hipStream_t side_stream() {
    struct Owner {
        hipStream_t s = nullptr;
        ~Owner() { if ( s ) (void) hipStreamDestroy( s ); }
    };
    thread_local Owner o;
    if ( ! o.s ) SDOT_HIP( hipStreamCreateWithFlags( &o.s, hipStreamNonBlocking ) );
    return o.s;
}

// ---- the device arena --------------------------------------------------------------------------

/// Library-owned device memory, grow-only, one named buffer per role: a solve of `n` seeds takes
/// its buffers once and the next solve of the same size reuses them ( hipMalloc is a
/// synchronizing call and costs far more than a diagram ). `bytes_held()` is the footprint.
struct Arena {
    struct Buf { void *ptr = nullptr; size_t bytes = 0; };
    std::map<std::string, Buf> bufs, hbufs;          ///< device buffers; pinned host staging
    std::mutex mtx;
    // A FIXED REGION ( `fixed()` set it ): the buffers are then carved from caller-given memory, device and host, in the order
    // they are asked for, 256-byte aligned; nothing is allocated, and a request past the region is a capacity failure
    unsigned char *fix_dev = nullptr, *fix_host = nullptr;
    size_t fix_dev_bytes = 0, fix_dev_used = 0, fix_host_bytes = 0, fix_host_used = 0;

    void fixed( void *dev, size_t dev_bytes, void *host, size_t host_bytes ) {
        std::lock_guard<std::mutex> lk( mtx );
        bufs.clear(); hbufs.clear();
        fix_dev = static_cast<unsigned char *>( dev ); fix_dev_bytes = dev_bytes; fix_dev_used = 0;
        fix_host = static_cast<unsigned char *>( host ); fix_host_bytes = host_bytes; fix_host_used = 0;
    }
    static void *carve( unsigned char *base, size_t cap, size_t &used, size_t bytes, const std::string &name ) {
        const size_t at = ( used + 255 ) & ~size_t( 255 );
        if ( at + bytes > cap )
            throw Failure( 3, 0, I64( at + bytes ), "the fixed scratch region is too small for buffer '" + name + "' (" + std::to_string( at + bytes ) +
                                                        " of " + std::to_string( cap ) + " bytes)" );
        used = at + bytes;
        return base + at;
    }

    /// pinned host staging, grow-only ( readbacks of the few scalars Newton decides on )
    template<class T>
    T *host( const std::string &name, size_t count ) {
        std::lock_guard<std::mutex> lk( mtx );
        Buf &b = hbufs[ name ];
        const size_t bytes = std::max<size_t>( count * sizeof( T ), 64 );
        if ( b.bytes < bytes && fix_dev ) {
            b.ptr = carve( fix_host, fix_host_bytes, fix_host_used, bytes, name );
            b.bytes = bytes;
        }
        if ( b.bytes < bytes ) {
            if ( b.ptr ) (void) hipHostFree( b.ptr );
            b.ptr = nullptr;
            b.bytes = 0;
            SDOT_HIP( hipHostMalloc( &b.ptr, bytes, hipHostMallocDefault ) );
            b.bytes = bytes;
        }
        return reinterpret_cast<T *>( b.ptr );
    }

    /// the buffer `name` of at least `count` elements; `fresh`, when given, tells whether it was ( re )made by this call
    template<class T>
    T *get( const std::string &name, size_t count, bool *fresh = nullptr ) {
        std::lock_guard<std::mutex> lk( mtx );
        const size_t bytes = std::max<size_t>( count * sizeof( T ), 256 );
        Buf &b = bufs[ name ];
        if ( fresh ) *fresh = b.bytes < bytes;
        if ( b.bytes < bytes && fix_dev ) {
            b.ptr = carve( fix_dev, fix_dev_bytes, fix_dev_used, bytes, name );
            b.bytes = bytes;
        }
        if ( b.bytes < bytes ) {
            if ( b.ptr ) SDOT_HIP( hipFree( b.ptr ) );
            b.ptr = nullptr;
            b.bytes = 0;
            hipError_t e = hipMalloc( &b.ptr, bytes );
            if ( e != hipSuccess ) {
                (void) hipGetLastError();
                throw Failure( 4, int( e ), I64( bytes ), "out of device memory for buffer '" + name + "' (" + std::to_string( bytes ) + " bytes)" );
            }
            b.bytes = bytes;
        }
        return reinterpret_cast<T *>( b.ptr );
    }

    size_t bytes_held() {
        std::lock_guard<std::mutex> lk( mtx );
        size_t s = 0;
        for ( auto &kv : bufs ) s += kv.second.bytes;
        return s;
    }

    void release() {
        std::lock_guard<std::mutex> lk( mtx );
        if ( fix_dev ) { bufs.clear(); hbufs.clear(); fix_dev_used = fix_host_used = 0; return; }
        for ( auto &kv : bufs )
            if ( kv.second.ptr ) (void) hipFree( kv.second.ptr );
        bufs.clear();
        for ( auto &kv : hbufs )
            if ( kv.second.ptr ) (void) hipHostFree( kv.second.ptr );
        hbufs.clear();
    }
};

/// the arena of the calling thread: its own when `set_thread_arena` gave it one ( a worker of a batch ), else the
/// current device's
Arena &arena();
void set_thread_arena( Arena *a );
/// the calling thread's own arena ( null: it uses the device's )
Arena *thread_arena();

/// pinned host staging, grow-only ( readbacks of the few scalars Newton decides on )
template<class T>
T *pinned( const std::string &name, size_t count );

// ---- live kernel timing ------------------------------------------------------------------------

/// the kernels whose launches can be timed live ( HIP events on their own stream, resolved on demand )
enum TimedKernel { TK_CELLS = 0, TK_CELLS_OVF, TK_SPMV, TK_MAJORANTS, TK_NB };

struct Timing {
    bool on = false;
    struct Pending { int k; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double ms[ TK_NB ] = {};
    long long launches[ TK_NB ] = {};
    double spmv_bytes = 0;          ///< algorithmic bytes of the timed SpMV launches
    // the multigrid's HIP graphs ( counted whether timing is on or not ): batches of iterations captured, and replayed
    std::atomic<long long> graph_captures{ 0 }, graph_replays{ 0 };
    // the FP32 kernel's diagrams in solves ( counted whether timing is on or not ): the diagrams, the cells redone by the FP64
    // kernel for a facet only one side had, the facets still one-sided after that ( dropped )
    std::atomic<long long> fp32_diagrams{ 0 }, fp32_redone{ 0 }, fp32_dropped{ 0 };
    void fp32_cells( unsigned redone, unsigned dropped ) { ++fp32_diagrams; fp32_redone += redone; fp32_dropped += dropped; }
    // the multigrid's FP32 preconditioner ( counted whether timing is on or not ): the solves, their flexible-CG iterations, and
    // the solves redone with the FP64 preconditioner because they did not converge
    std::atomic<long long> mg32_solves{ 0 }, mg32_iterations{ 0 }, mg32_fallbacks{ 0 };
    std::mutex mtx;

    hipEvent_t take() {
        if ( pool.empty() ) { hipEvent_t e; if ( hipEventCreate( &e ) != hipSuccess ) return nullptr; return e; }
        hipEvent_t e = pool.back(); pool.pop_back(); return e;
    }
    /// a pair of events around one launch; `begin` returns false when timing is off
    bool begin( int k, hipStream_t st, Pending &p ) {
        if ( ! on ) return false;
        std::lock_guard<std::mutex> lk( mtx );
        p.k = k; p.a = take(); p.b = take();
        if ( ! p.a || ! p.b ) return false;
        (void) hipEventRecord( p.a, st );
        return true;
    }
    void end( Pending &p, hipStream_t st ) {
        (void) hipEventRecord( p.b, st );
        std::lock_guard<std::mutex> lk( mtx );
        pending.push_back( p );
    }
    /// waits for the recorded pairs and adds them up
    void resolve() {
        std::lock_guard<std::mutex> lk( mtx );
        for ( auto &p : pending ) {
            float t = 0;
            if ( hipEventSynchronize( p.b ) == hipSuccess && hipEventElapsedTime( &t, p.a, p.b ) == hipSuccess ) {
                ms[ p.k ] += t;
                launches[ p.k ] += 1;
            }
            pool.push_back( p.a );
            pool.push_back( p.b );
        }
        pending.clear();
    }
    void reset() {
        resolve();
        std::lock_guard<std::mutex> lk( mtx );
        for ( int k = 0; k < TK_NB; ++k ) { ms[ k ] = 0; launches[ k ] = 0; }
        spmv_bytes = 0;
        graph_captures = 0;
        fp32_diagrams = 0; fp32_redone = 0; fp32_dropped = 0;
        mg32_solves = 0; mg32_iterations = 0; mg32_fallbacks = 0;
        graph_replays = 0;
    }
};

Timing &timing();

// ---- launch shapes -----------------------------------------------------------------------------

constexpr int RED_BLOCKS  = 512;    ///< partial sums of every deterministic reduction ( fixed shape )
constexpr int RED_THREADS = 256;

inline unsigned blocks_for( I64 n, int threads ) { return unsigned( ( n + threads - 1 ) / threads ); }

// ---- device helpers ----------------------------------------------------------------------------

/// a monotone map from doubles to unsigned 64-bit integers: `a < b` iff `key( a ) < key( b )`,
/// so that an integer atomic min / max is a double min / max
__device__ __forceinline__ unsigned long long ordered_key( double x ) {
    unsigned long long u = __double_as_longlong( x );
    return ( u >> 63 ) ? ~u : ( u | 0x8000000000000000ull );
}
__device__ __forceinline__ double ordered_value( unsigned long long k ) {
    const unsigned long long u = ( k >> 63 ) ? ( k & 0x7fffffffffffffffull ) : ~k;
    return __longlong_as_double( u );
}
inline unsigned long long ordered_key_host( double x ) {
    unsigned long long u;
    std::memcpy( &u, &x, 8 );
    return ( u >> 63 ) ? ~u : ( u | 0x8000000000000000ull );
}

/// block-wide sum of a double, deterministic ( fixed tree over the block )
template<int THREADS>
__device__ __forceinline__ double block_sum( double v, double *sh ) {
    sh[ threadIdx.x ] = v;
    __syncthreads();
    for ( int s = THREADS / 2; s > 0; s >>= 1 ) {
        if ( int( threadIdx.x ) < s ) sh[ threadIdx.x ] += sh[ threadIdx.x + s ];
        __syncthreads();
    }
    const double r = sh[ 0 ];
    __syncthreads();
    return r;
}

template<int THREADS>
__device__ __forceinline__ double block_min( double v, double *sh ) {
    sh[ threadIdx.x ] = v;
    __syncthreads();
    for ( int s = THREADS / 2; s > 0; s >>= 1 ) {
        if ( int( threadIdx.x ) < s ) sh[ threadIdx.x ] = fmin( sh[ threadIdx.x ], sh[ threadIdx.x + s ] );
        __syncthreads();
    }
    const double r = sh[ 0 ];
    __syncthreads();
    return r;
}

template<int THREADS>
__device__ __forceinline__ double block_max( double v, double *sh ) {
    sh[ threadIdx.x ] = v;
    __syncthreads();
    for ( int s = THREADS / 2; s > 0; s >>= 1 ) {
        if ( int( threadIdx.x ) < s ) sh[ threadIdx.x ] = fmax( sh[ threadIdx.x ], sh[ threadIdx.x + s ] );
        __syncthreads();
    }
    const double r = sh[ 0 ];
    __syncthreads();
    return r;
}

} // namespace sdot_amd
