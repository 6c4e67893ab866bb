// Fixed-shape reductions on one stream: `RED_BLOCKS` partials in a grid-stride loop, then one block in a fixed
// order, so a result is bitwise the same run to run.
#pragma once

#include "common.h"

#include <algorithm>

namespace sdot_amd {

namespace {

// ---- fixed-shape reductions -------------------------------------------------------------------------

enum { OP_SUM = 0, OP_MIN = 1, OP_MAX = 2 };

__device__ __forceinline__ double op_init( int op ) { return op == OP_SUM ? 0.0 : ( op == OP_MIN ? INFINITY : -INFINITY ); }
__device__ __forceinline__ double op_apply( int op, double a, double b ) { return op == OP_SUM ? a + b : ( op == OP_MIN ? fmin( a, b ) : fmax( a, b ) ); }

template<int K>
struct Ops { int op[ K ]; };

__device__ __forceinline__ double block_op( int op, double v, double *sh ) {
    if ( op == OP_SUM ) return block_sum<RED_THREADS>( v, sh );
    if ( op == OP_MIN ) return block_min<RED_THREADS>( v, sh );
    return block_max<RED_THREADS>( v, sh );
}

template<int K, class F>
__global__ __launch_bounds__( RED_THREADS ) void k_partials( I64 n, F f, Ops<K> ops, double *part ) {
    __shared__ double sh[ RED_THREADS ];
    double acc[ K ];
    for ( int q = 0; q < K; ++q ) acc[ q ] = op_init( ops.op[ q ] );
    for ( I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x; i < n; i += I64( RED_BLOCKS ) * RED_THREADS ) {
        double v[ K ];
        f( i, v );
        for ( int q = 0; q < K; ++q ) acc[ q ] = op_apply( ops.op[ q ], acc[ q ], v[ q ] );
    }
    for ( int q = 0; q < K; ++q ) {
        const double r = block_op( ops.op[ q ], acc[ q ], sh );
        if ( threadIdx.x == 0 ) part[ q * RED_BLOCKS + blockIdx.x ] = r;
    }
}

/// one element per lane ( for functors heavy enough that a grid-stride loop over `RED_BLOCKS` blocks would leave the chip
/// half idle ): one partial per block, `nb` blocks, quantity `q`'s partials at `part[ q * nb + b ]`
template<int K, class F>
__global__ __launch_bounds__( RED_THREADS ) void k_partials_wide( I64 n, F f, Ops<K> ops, double *part, int nb ) {
    __shared__ double sh[ RED_THREADS ];
    double v[ K ];
    const I64 i = I64( blockIdx.x ) * RED_THREADS + threadIdx.x;
    if ( i < n ) f( i, v );
    else for ( int q = 0; q < K; ++q ) v[ q ] = op_init( ops.op[ q ] );
    for ( int q = 0; q < K; ++q ) {
        const double r = block_op( ops.op[ q ], v[ q ], sh );
        if ( threadIdx.x == 0 ) part[ I64( q ) * nb + blockIdx.x ] = r;
    }
}

/// the `nb` partials of every quantity reduced by one block; `E` then runs on thread 0 with the results
template<int K, class E>
__global__ __launch_bounds__( RED_THREADS ) void k_final( const double *part, Ops<K> ops, double *out, E epilogue, int nb = RED_BLOCKS ) {
    __shared__ double sh[ RED_THREADS ];
    __shared__ double res[ K ];
    for ( int q = 0; q < K; ++q ) {
        double a = op_init( ops.op[ q ] );
        for ( int j = threadIdx.x; j < nb; j += RED_THREADS ) a = op_apply( ops.op[ q ], a, part[ I64( q ) * nb + j ] );
        const double r = block_op( ops.op[ q ], a, sh );
        if ( threadIdx.x == 0 ) res[ q ] = r;
    }
    __syncthreads();
    if ( threadIdx.x == 0 ) {
        for ( int q = 0; q < K; ++q ) out[ q ] = res[ q ];
        epilogue( res );
    }
}

struct NoEpilogue { __device__ void operator()( const double * ) const {} };

template<int K, class F, class E = NoEpilogue>
void reduce( I64 n, const F &f, Ops<K> ops, double *out, hipStream_t st, E e = E{} ) {
    double *part = arena().get<double>( "red.part", size_t( K ) * RED_BLOCKS );
    k_partials<K, F><<<RED_BLOCKS, RED_THREADS, 0, st>>>( n, f, ops, part );
    check_launch( "k_partials" );
    k_final<K, E><<<1, RED_THREADS, 0, st>>>( part, ops, out, e );
    check_launch( "k_final" );
}

/// the same reduction with one element per lane ( `k_partials_wide` ): for per-element work of many loads and many
/// registers, where the chip needs every element in flight at once. Fixed shape for a given `n`, so still bitwise
/// reproducible
template<int K, class F>
void reduce_wide( I64 n, const F &f, Ops<K> ops, double *out, hipStream_t st ) {
    const int nb = int( std::max<I64>( 1, ( n + RED_THREADS - 1 ) / RED_THREADS ) );
    double *part = arena().get<double>( "red.wide", size_t( K ) * size_t( nb ) );
    k_partials_wide<K, F><<<nb, RED_THREADS, 0, st>>>( n, f, ops, part, nb );
    check_launch( "k_partials_wide" );
    k_final<K, NoEpilogue><<<1, RED_THREADS, 0, st>>>( part, ops, out, NoEpilogue{}, nb );
    check_launch( "k_final" );
}

} // namespace

} // namespace sdot_amd
