// The BSP tree built on the device, no host round trip: the same tree as the host build ( median cut on
// the longest axis, leaf <= `leaf`, perfect binary tree in preorder, each leaf's seeds in ascending
// user index ).
//
// How ( this build's own scheme; the reference's card builder is `gpu/Bsp2D.cuh:367` ):
//   1. two stable LSD radix sorts of the seeds, by x and by y ( 64-bit order-preserving keys of the
//      doubles, 8 passes of 8 bits; the in-tile rank of an element among its digit's is found with
//      wave64 ballots, so a pass is histogram / scan / scatter );
//   2. per level, every node's slice is a contiguous segment of BOTH sorted lists, so its box is read
//      off the segment's ends ( first and last of the x list, of the y list ); the median cut on the
//      longest axis takes the first half of that axis' list as is, and the other list is stably
//      partitioned by the flags the cut writes ( an exclusive scan over all positions ) -- both lists
//      stay sorted inside every child;
//   3. the leaves' seeds sorted by user index, the preorder numbering and the rank tables.
// Everything stays on the device; the host only issues launches ( their count is a function of `n` ).

#include "tree.h"
#include "tree3.h"

namespace sdot_amd {

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

__device__ __forceinline__ unsigned long long dkey( double x ) {
    unsigned long long u = __double_as_longlong( x );
    return ( u >> 63 ) ? ~u : ( u | 0x8000000000000000ull );
}

/// the keys of one axis ( `dim` coordinates a seed ), and the identity permutation
__global__ void k_keys( I64 n, const double *pos, int dim, int axis, unsigned long long *key, int *val ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i >= n ) return;
    key[ i ] = dkey( pos[ dim * i + axis ] );
    val[ i ] = int( i );
}

/// histogram of digit `shift` per tile: `hist[ digit * nt + tile ]`
__global__ __launch_bounds__( RS_THREADS ) void k_rs_hist( I64 n, const unsigned long long *key, int shift, int nt, unsigned *hist ) {
    __shared__ unsigned h[ 256 ];
    h[ threadIdx.x ] = 0;
    __syncthreads();
    const I64 base = I64( blockIdx.x ) * RS_TILE;
    for ( int j = 0; j < RS_ITEMS; ++j ) {
        const I64 i = base + j * RS_THREADS + threadIdx.x;
        if ( i < n ) atomicAdd( &h[ ( key[ i ] >> shift ) & 255 ], 1u );
    }
    __syncthreads();
    hist[ size_t( threadIdx.x ) * nt + blockIdx.x ] = h[ threadIdx.x ];
}

/// exclusive scan of `m` counts in one block ( digit-major: the offsets of each tile's digit run )
__global__ __launch_bounds__( 1024 ) void k_scan_single( unsigned *v, int m ) {
    __shared__ unsigned part[ 1024 ];
    const int per = ( m + 1023 ) / 1024;
    const int b = threadIdx.x * per, e = min( m, b + per );
    unsigned s = 0;
    for ( int i = b; i < e; ++i ) s += v[ i ];
    part[ threadIdx.x ] = s;
    __syncthreads();
    for ( int off = 1; off < 1024; off <<= 1 ) {
        const unsigned t = threadIdx.x >= unsigned( off ) ? part[ threadIdx.x - off ] : 0u;
        __syncthreads();
        part[ threadIdx.x ] += t;
        __syncthreads();
    }
    unsigned acc = threadIdx.x ? part[ threadIdx.x - 1 ] : 0u;
    for ( int i = b; i < e; ++i ) { const unsigned x = v[ i ]; v[ i ] = acc; acc += x; }
}

/// the stable scatter of one pass: a tile is read in rounds of 256 elements in order; inside a round the
/// rank of an element among the same digit comes from wave ballots ( 8 bits -> the mask of its peers )
/// plus the counts of the waves before it and of the earlier rounds
__global__ __launch_bounds__( RS_THREADS ) void k_rs_scatter( I64 n, const unsigned long long *key, const int *val, int shift, int nt,
                                                            const unsigned *offs, unsigned long long *key_out, int *val_out ) {
    __shared__ unsigned running[ 256 ];
    __shared__ unsigned wcount[ RS_THREADS / 64 ][ 256 ];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    running[ threadIdx.x ] = offs[ size_t( threadIdx.x ) * nt + blockIdx.x ];
    const I64 base = I64( blockIdx.x ) * RS_TILE;
    const unsigned long long lt = lane ? ( ~0ull >> ( 64 - lane ) ) : 0ull;
    for ( int j = 0; j < RS_ITEMS; ++j ) {
        for ( int w = 0; w < RS_THREADS / 64; ++w ) wcount[ w ][ threadIdx.x ] = 0;
        __syncthreads();
        const I64 i = base + j * RS_THREADS + threadIdx.x;
        const bool live = i < n;
        const unsigned long long k = live ? key[ i ] : 0ull;
        const unsigned d = unsigned( ( k >> shift ) & 255 );
        unsigned long long peers = __ballot( live );
        for ( int b = 0; b < 8; ++b ) {
            const unsigned long long m = __ballot( ( d >> b ) & 1 );
            peers &= ( ( d >> b ) & 1 ) ? m : ~m;
        }
        const unsigned r_in = __popcll( peers & lt );
        if ( live && r_in == 0 ) wcount[ wave ][ d ] = __popcll( peers );
        __syncthreads();
        if ( live ) {
            unsigned r = running[ d ] + r_in;
            for ( int w = 0; w < wave; ++w ) r += wcount[ w ][ d ];
            key_out[ r ] = k;
            val_out[ r ] = val[ i ];
        }
        __syncthreads();
        unsigned add = 0;
        for ( int w = 0; w < RS_THREADS / 64; ++w ) add += wcount[ w ][ threadIdx.x ];
        running[ threadIdx.x ] += add;
        __syncthreads();
    }
}

struct Sorted { unsigned long long *key; int *val; };

/// the stable sort of `n` ( key, id ) pairs by key; returns the buffers holding the result
Sorted radix_sort( I64 n, unsigned long long *k0, int *v0, unsigned long long *k1, int *v1, unsigned *hist, hipStream_t st ) {
    const int nt = int( ( n + RS_TILE - 1 ) / RS_TILE );
    for ( int pass = 0; pass < 8; ++pass ) {
        const int shift = 8 * pass;
        k_rs_hist<<<nt, RS_THREADS, 0, st>>>( n, k0, shift, nt, hist );
        k_scan_single<<<1, 1024, 0, st>>>( hist, 256 * nt );
        k_rs_scatter<<<nt, RS_THREADS, 0, st>>>( n, k0, v0, shift, nt, hist, k1, v1 );
        std::swap( k0, k1 );
        std::swap( v0, v1 );
    }
    check_launch( "radix_sort" );
    return Sorted{ k0, v0 };
}

// ---- the levels ------------------------------------------------------------------------------------------

/// one node of level `L` ( heap index `first + j` ): its slice, box and cut
__global__ void k_level_nodes( int count, int first, const int *beg, const int *end, const double *pos, const int *xid, const int *yid, int leaf,
                               double *hbox, int *axis, int *mid, int *hbeg, int *hend ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= count ) return;
    const int b = beg[ j ], e = end[ j ], node = first + j;
    hbeg[ node ] = b;
    hend[ node ] = e;
    double *bx = hbox + 4 * size_t( node );
    if ( e <= b ) {
        bx[ 0 ] = bx[ 1 ] = bx[ 2 ] = bx[ 3 ] = 0;
        axis[ j ] = 0;
        mid[ j ] = e;
        return;
    }
    const double lox = pos[ 2 * size_t( xid[ b ] ) ], hix = pos[ 2 * size_t( xid[ e - 1 ] ) ];
    const double loy = pos[ 2 * size_t( yid[ b ] ) + 1 ], hiy = pos[ 2 * size_t( yid[ e - 1 ] ) + 1 ];
    bx[ 0 ] = lox; bx[ 1 ] = loy; bx[ 2 ] = hix; bx[ 3 ] = hiy;
    const int ax = ( hiy - loy > hix - lox ) ? 1 : 0;
    axis[ j ] = ax;
    const double lo = ax ? loy : lox, hi = ax ? hiy : hix;
    mid[ j ] = ( e - b > leaf && hi > lo ) ? b + ( e - b ) / 2 : e;
}

/// which node of the level each position belongs to: `seg[ p ]`, from the slices
__global__ void k_fill_seg( int count, const int *beg, const int *end, int *seg ) {
    const int j = blockIdx.x;
    const int b = beg[ j ], e = end[ j ];
    for ( int p = b + threadIdx.x; p < e; p += blockDim.x ) seg[ p ] = j;
}

/// the side of every seed: from the cut axis' list, `flag[ id ] = position < mid`
__global__ void k_flags( I64 n, const int *seg, const int *axis, const int *mid, const int *xid, const int *yid, unsigned char *flag ) {
    const I64 p = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( p >= n ) return;
    const int j = seg[ p ];
    const int id = axis[ j ] ? yid[ p ] : xid[ p ];
    flag[ id ] = p < mid[ j ];
}

/// the left flags of the OTHER list, in its order ( what the scan counts )
__global__ void k_other_flags( I64 n, const int *seg, const int *axis, const int *xid, const int *yid, const unsigned char *flag, unsigned *f ) {
    const I64 p = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( p >= n ) return;
    const int j = seg[ p ];
    const int id = axis[ j ] ? xid[ p ] : yid[ p ];
    f[ p ] = flag[ id ];
}

/// block sums, then their scan, then the scan of each block: an exclusive scan of `f` in `s`
__global__ __launch_bounds__( 1024 ) void k_scan_blocks( I64 n, const unsigned *f, unsigned *s, unsigned *bsum ) {
    __shared__ unsigned sh[ 1024 ];
    const I64 i = I64( blockIdx.x ) * 1024 + threadIdx.x;
    const unsigned v = i < n ? f[ i ] : 0u;
    sh[ threadIdx.x ] = v;
    __syncthreads();
    for ( int off = 1; off < 1024; off <<= 1 ) {
        const unsigned t = threadIdx.x >= unsigned( off ) ? sh[ threadIdx.x - off ] : 0u;
        __syncthreads();
        sh[ threadIdx.x ] += t;
        __syncthreads();
    }
    if ( i < n ) s[ i ] = sh[ threadIdx.x ] - v;
    if ( threadIdx.x == 1023 ) bsum[ blockIdx.x ] = sh[ 1023 ];
}

__global__ void k_scan_add( I64 n, unsigned *s, const unsigned *bsum ) {
    const I64 i = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i < n ) s[ i ] += bsum[ i / 1024 ];
}

/// the stable partition of the other list inside every segment; the cut axis' list is already split
__global__ void k_partition( I64 n, const int *seg, const int *axis, const int *beg, const int *mid, const int *xid, const int *yid,
                             const unsigned char *flag, const unsigned *s, int *xo, int *yo ) {
    const I64 p = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( p >= n ) return;
    const int j = seg[ p ];
    const int ax = axis[ j ], b = beg[ j ], m = mid[ j ];
    const int id = ax ? xid[ p ] : yid[ p ];
    const unsigned left_before = s[ p ] - s[ b ];
    const int q = flag[ id ] ? b + int( left_before ) : m + int( ( p - b ) - left_before );
    if ( ax ) { yo[ p ] = yid[ p ]; xo[ q ] = id; }
    else      { xo[ p ] = xid[ p ]; yo[ q ] = id; }
}

/// the next level's slices: `[ b, mid )` and `[ mid, e )`
__global__ void k_children( int count, const int *beg, const int *end, const int *mid, int *nb, int *ne ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= count ) return;
    nb[ 2 * j ] = beg[ j ]; ne[ 2 * j ] = mid[ j ];
    nb[ 2 * j + 1 ] = mid[ j ]; ne[ 2 * j + 1 ] = end[ j ];
}

/// the leaves' seeds in ascending user index, and the outputs in rank order
__global__ void k_leaves( int count, int first, const int *hbeg, const int *hend, int *ids ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= count ) return;
    const int b = hbeg[ first + j ], e = hend[ first + j ];
    for ( int i = b + 1; i < e; ++i ) {                   // insertion sort, a handful of seeds
        const int v = ids[ i ];
        int k = i;
        for ( ; k > b && ids[ k - 1 ] > v; --k ) ids[ k ] = ids[ k - 1 ];
        ids[ k ] = v;
    }
}

__global__ void k_rank_tables( I64 n, const int *ids, const double *pos, int *order, int *rank_of, double4 *seeds ) {
    const I64 k = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( k >= n ) return;
    const int u = ids[ k ];
    order[ k ] = u;
    rank_of[ u ] = int( k );
    seeds[ k ] = make_double4( pos[ 2 * size_t( u ) ], pos[ 2 * size_t( u ) + 1 ], 0.0, 0.0 );
}

/// heap -> preorder for every node: `pre[ h ]`, computed level by level
__global__ void k_preorder( int L, int depth, int *pre ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= ( 1 << L ) ) return;
    // the path from the root: bit ( L - 1 - t ) of `j` is the side taken at level t
    int p = 0;
    for ( int t = 0; t < L; ++t ) {
        const int hgt = depth - t;
        p += ( ( j >> ( L - 1 - t ) ) & 1 ) ? ( 1 << ( hgt - 1 ) ) : 1;
    }
    pre[ ( 1 << L ) - 1 + j ] = p;
}

__global__ void k_to_preorder( I64 N, const int *pre, const int *hbeg, const int *hend, const double *hbox, int *beg, int *end, double *box ) {
    const I64 h = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( h >= N ) return;
    const int p = pre[ h ];
    beg[ p ] = hbeg[ h ];
    end[ p ] = hend[ h ];
    for ( int c = 0; c < 4; ++c ) box[ 4 * size_t( p ) + c ] = hbox[ 4 * size_t( h ) + c ];
}

// ---- 3D: the same scheme with three sorted lists ----------------------------------------------------------------------
//
// Every node's slice is a contiguous segment of each of the three lists. The cut list ( the longest axis' ) decides each
// seed's side; then EVERY list is stably partitioned by those sides ( for the cut list itself the partition is the identity:
// its first half is the left child ), one exclusive scan per list.

struct Lists3 { const int *id[ 3 ]; };

/// one node of level `L` in 3D: its slice, box ( the ends of the three lists' segments ) and cut
__global__ void k_level_nodes3( int count, int first, const int *beg, const int *end, const double *pos, Lists3 L, int leaf, double *hbox,
                                int *axis, int *mid, int *hbeg, int *hend ) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= count ) return;
    const int b = beg[ j ], e = end[ j ], node = first + j;
    hbeg[ node ] = b;
    hend[ node ] = e;
    double *bx = hbox + 6 * size_t( node );
    if ( e <= b ) {
        for ( int c = 0; c < 6; ++c ) bx[ c ] = 0;
        axis[ j ] = 0;
        mid[ j ] = e;
        return;
    }
    double lo[ 3 ], hi[ 3 ];
    for ( int d = 0; d < 3; ++d ) {
        lo[ d ] = pos[ 3 * size_t( L.id[ d ][ b ] ) + d ];
        hi[ d ] = pos[ 3 * size_t( L.id[ d ][ e - 1 ] ) + d ];
        bx[ d ] = lo[ d ];
        bx[ 3 + d ] = hi[ d ];
    }
    int ax = 0;                                          // ( the first longest axis, as the host build picks it )
    for ( int d = 1; d < 3; ++d ) if ( hi[ d ] - lo[ d ] > hi[ ax ] - lo[ ax ] ) ax = d;
    axis[ j ] = ax;
    mid[ j ] = ( e - b > leaf && hi[ ax ] > lo[ ax ] ) ? b + ( e - b ) / 2 : e;
}

/// the side of every seed from its node's cut list: `flag[ id ] = position < mid`
__global__ void k_flags3( I64 n, const int *seg, const int *axis, const int *mid, Lists3 L, unsigned char *flag ) {
    const I64 p = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( p >= n ) return;
    const int j = seg[ p ];
    flag[ L.id[ axis[ j ] ][ p ] ] = p < mid[ j ];
}

/// the left flags of list `ids`, in its order
__global__ void k_list_flags( I64 n, const int *ids, const unsigned char *flag, unsigned *f ) {
    const I64 p = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( p < n ) f[ p ] = flag[ ids[ p ] ];
}

/// the stable partition of list `ids` inside every segment by the flags ( `s` their exclusive scan in the list's order )
__global__ void k_partition_list( I64 n, const int *seg, const int *beg, const int *mid, const int *ids, const unsigned char *flag,
                                  const unsigned *s, int *out ) {
    const I64 p = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( p >= n ) return;
    const int j = seg[ p ];
    const int b = beg[ j ], m = mid[ j ], id = ids[ p ];
    const unsigned left_before = s[ p ] - s[ b ];
    out[ flag[ id ] ? b + int( left_before ) : m + int( ( p - b ) - left_before ) ] = id;
}

__global__ void k_rank_tables3( I64 n, const int *ids, const double *pos, int *order, int *rank_of, double4 *seeds ) {
    const I64 k = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( k >= n ) return;
    const int u = ids[ k ];
    order[ k ] = u;
    rank_of[ u ] = int( k );
    seeds[ k ] = make_double4( pos[ 3 * size_t( u ) ], pos[ 3 * size_t( u ) + 1 ], pos[ 3 * size_t( u ) + 2 ], 0.0 );
}

__global__ void k_to_preorder3( I64 N, const int *pre, const int *hbeg, const int *hend, const double *hbox, int *beg, int *end, double *box ) {
    const I64 h = I64( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( h >= N ) return;
    const int p = pre[ h ];
    beg[ p ] = hbeg[ h ];
    end[ p ] = hend[ h ];
    for ( int c = 0; c < 6; ++c ) box[ 6 * size_t( p ) + c ] = hbox[ 6 * size_t( h ) + c ];
}

} // namespace

void build_tree_device3( DevTree3 &t, I64 n, const double *pos, int leaf, hipStream_t st, const std::string &tag ) {
    SDOT_CHECK( n >= 1, "a tree needs at least one seed" );
    SDOT_CHECK( n < ( I64( 1 ) << 31 ) - 8, "more seeds than 32-bit ranks can index" );
    const int depth = max_depth_for( n, leaf );
    const I64 N = ( I64( 1 ) << depth ) - 1;
    t.n = n; t.depth = depth; t.leaf = leaf; t.nb_nodes = N;
    Arena &A = arena();
    t.nodes = A.get<NodeRec3>( tag + ".nodes3", size_t( N ) );
    t.node_box = A.get<double>( tag + ".node_box3", size_t( N ) * 6 );
    t.node_beg = A.get<int>( tag + ".node_beg3", size_t( N ) );
    t.node_end = A.get<int>( tag + ".node_end3", size_t( N ) );
    t.lvl_nodes = A.get<int>( tag + ".lvl_nodes3", size_t( N ) );
    t.order = A.get<int>( tag + ".order3", size_t( n ) );
    t.rank_of = A.get<int>( tag + ".rank_of3", size_t( n ) );
    t.seeds = A.get<double4>( tag + ".seeds3", size_t( n ) );
    t.stat = A.get<double>( tag + ".stat3", size_t( N ) * 16 );
    t.fit = A.get<double>( tag + ".fit3", size_t( N ) * 3 );
    t.rkeys = A.get<unsigned long long>( tag + ".rkeys3", size_t( N ) * 3 );
    // scratch ( shared with the 2D build's names: one build at a time on a thread's arena )
    unsigned long long *ka = A.get<unsigned long long>( "tb.ka", size_t( n ) ), *kb = A.get<unsigned long long>( "tb.kb", size_t( n ) );
    int *va = A.get<int>( "tb.va", size_t( n ) ), *vb = A.get<int>( "tb.vb", size_t( n ) );
    int *lst[ 3 ], *lso[ 3 ];
    for ( int d = 0; d < 3; ++d ) {
        lst[ d ] = A.get<int>( "tb3.l" + std::to_string( d ), size_t( n ) );
        lso[ d ] = A.get<int>( "tb3.o" + std::to_string( d ), size_t( n ) );
    }
    const int nt = int( ( n + RS_TILE - 1 ) / RS_TILE );
    unsigned *hist = A.get<unsigned>( "tb.hist", size_t( 256 ) * nt );
    const I64 half = I64( 1 ) << ( depth - 1 );
    int *beg = A.get<int>( "tb.beg", size_t( half ) ), *end = A.get<int>( "tb.end", size_t( half ) );
    int *nbeg = A.get<int>( "tb.nbeg", size_t( half ) ), *nend = A.get<int>( "tb.nend", size_t( half ) );
    int *axis = A.get<int>( "tb.axis", size_t( half ) ), *mid = A.get<int>( "tb.mid", size_t( half ) );
    int *hbeg = A.get<int>( "tb.hbeg", size_t( N ) ), *hend = A.get<int>( "tb.hend", size_t( N ) );
    double *hbox = A.get<double>( "tb3.hbox", size_t( N ) * 6 );
    int *seg = A.get<int>( "tb.seg", size_t( n ) );
    unsigned char *flag = A.get<unsigned char>( "tb.flag", size_t( n ) );
    unsigned *f = A.get<unsigned>( "tb.f", size_t( n ) ), *s = A.get<unsigned>( "tb.s", size_t( n ) );
    const I64 nsb = ( n + 1023 ) / 1024;
    unsigned *bsum = A.get<unsigned>( "tb.bsum", size_t( nsb ) );
    const unsigned nbk = blocks_for( n, 256 );

    // 1. the three sorted lists
    for ( int ax = 0; ax < 3; ++ax ) {
        k_keys<<<nbk, 256, 0, st>>>( n, pos, 3, ax, ka, va );
        const Sorted r = radix_sort( n, ka, va, kb, vb, hist, st );
        SDOT_HIP( hipMemcpyAsync( lst[ ax ], r.val, sizeof( int ) * size_t( n ), hipMemcpyDeviceToDevice, st ) );
    }
    // 2. the levels
    const int zero = 0, nn = int( n );
    SDOT_HIP( hipMemcpyAsync( beg, &zero, sizeof( int ), hipMemcpyHostToDevice, st ) );
    SDOT_HIP( hipMemcpyAsync( end, &nn, sizeof( int ), hipMemcpyHostToDevice, st ) );
    for ( int L = 0; L < depth; ++L ) {
        const int count = 1 << L, first = count - 1;
        k_level_nodes3<<<blocks_for( count, 256 ), 256, 0, st>>>( count, first, beg, end, pos, Lists3{ { lst[ 0 ], lst[ 1 ], lst[ 2 ] } }, leaf,
                                                                  hbox, axis, mid, hbeg, hend );
        if ( L + 1 == depth ) break;
        k_fill_seg<<<count, 256, 0, st>>>( count, beg, end, seg );
        k_flags3<<<nbk, 256, 0, st>>>( n, seg, axis, mid, Lists3{ { lst[ 0 ], lst[ 1 ], lst[ 2 ] } }, flag );
        for ( int d = 0; d < 3; ++d ) {
            k_list_flags<<<nbk, 256, 0, st>>>( n, lst[ d ], flag, f );
            k_scan_blocks<<<unsigned( nsb ), 1024, 0, st>>>( n, f, s, bsum );
            k_scan_single<<<1, 1024, 0, st>>>( bsum, int( nsb ) );
            k_scan_add<<<nbk, 256, 0, st>>>( n, s, bsum );
            k_partition_list<<<nbk, 256, 0, st>>>( n, seg, beg, mid, lst[ d ], flag, s, lso[ d ] );
        }
        for ( int d = 0; d < 3; ++d ) std::swap( lst[ d ], lso[ d ] );
        k_children<<<blocks_for( count, 256 ), 256, 0, st>>>( count, beg, end, mid, nbeg, nend );
        std::swap( beg, nbeg );
        std::swap( end, nend );
        check_launch( "3D tree level" );
    }
    // 3. leaves, rank tables, preorder
    const int L = depth - 1;
    k_leaves<<<blocks_for( I64( 1 ) << L, 256 ), 256, 0, st>>>( 1 << L, ( 1 << L ) - 1, hbeg, hend, lst[ 0 ] );
    k_rank_tables3<<<nbk, 256, 0, st>>>( n, lst[ 0 ], pos, t.order, t.rank_of, t.seeds );
    for ( int l = 0; l < depth; ++l ) k_preorder<<<blocks_for( I64( 1 ) << l, 256 ), 256, 0, st>>>( l, depth, t.lvl_nodes );
    k_to_preorder3<<<blocks_for( N, 256 ), 256, 0, st>>>( N, t.lvl_nodes, hbeg, hend, hbox, t.node_beg, t.node_end, t.node_box );
    check_launch( "3D tree finish" );
    // `beg` / `end` came from host constants ( `zero`, `nn` ): the copies must be done before they go
    SDOT_HIP( hipStreamSynchronize( st ) );
}

void build_tree_device( DevTree &t, I64 n, const double *pos, int leaf, hipStream_t st, const std::string &tag ) {
    SDOT_CHECK( n >= 1, "a tree needs at least one seed" );
    SDOT_CHECK( n < ( I64( 1 ) << 31 ) - 8, "more seeds than 32-bit ranks can index" );
    const int depth = max_depth_for( n, leaf );
    const I64 N = ( I64( 1 ) << depth ) - 1;
    t.n = n; t.depth = depth; t.leaf = leaf; t.nb_nodes = N;
    Arena &A = arena();
    t.nodes = A.get<NodeRec>( tag + ".nodes", size_t( N ) );
    t.nodes32 = A.get<NodeF>( tag + ".nodes32", size_t( N ) );
    t.node_box = A.get<double>( tag + ".node_box", size_t( N ) * 4 );
    t.node_beg = A.get<int>( tag + ".node_beg", size_t( N ) );
    t.node_end = A.get<int>( tag + ".node_end", size_t( N ) );
    t.lvl_nodes = A.get<int>( tag + ".lvl_nodes", size_t( N ) );
    t.order = A.get<int>( tag + ".order", size_t( n ) );
    t.rank_of = A.get<int>( tag + ".rank_of", size_t( n ) );
    t.seeds = A.get<double4>( tag + ".seeds", size_t( n ) );
    t.stat = A.get<double>( tag + ".stat", size_t( N ) * 11 );
    t.fit = A.get<double>( tag + ".fit", size_t( N ) * 2 );
    t.rkeys = A.get<unsigned long long>( tag + ".rkeys", size_t( N ) * 3 );
    // scratch
    unsigned long long *ka = A.get<unsigned long long>( "tb.ka", size_t( n ) ), *kb = A.get<unsigned long long>( "tb.kb", size_t( n ) );
    int *va = A.get<int>( "tb.va", size_t( n ) ), *vb = A.get<int>( "tb.vb", size_t( n ) );
    int *xid = A.get<int>( "tb.xid", size_t( n ) ), *yid = A.get<int>( "tb.yid", size_t( n ) );
    int *xo = A.get<int>( "tb.xo", size_t( n ) ), *yo = A.get<int>( "tb.yo", size_t( n ) );
    const int nt = int( ( n + RS_TILE - 1 ) / RS_TILE );
    unsigned *hist = A.get<unsigned>( "tb.hist", size_t( 256 ) * nt );
    const I64 half = I64( 1 ) << ( depth - 1 );
    int *beg = A.get<int>( "tb.beg", size_t( half ) ), *end = A.get<int>( "tb.end", size_t( half ) );
    int *nbeg = A.get<int>( "tb.nbeg", size_t( half ) ), *nend = A.get<int>( "tb.nend", size_t( half ) );
    int *axis = A.get<int>( "tb.axis", size_t( half ) ), *mid = A.get<int>( "tb.mid", size_t( half ) );
    int *hbeg = A.get<int>( "tb.hbeg", size_t( N ) ), *hend = A.get<int>( "tb.hend", size_t( N ) );
    double *hbox = A.get<double>( "tb.hbox", size_t( N ) * 4 );
    int *seg = A.get<int>( "tb.seg", size_t( n ) );
    unsigned char *flag = A.get<unsigned char>( "tb.flag", size_t( n ) );
    unsigned *f = A.get<unsigned>( "tb.f", size_t( n ) ), *s = A.get<unsigned>( "tb.s", size_t( n ) );
    const I64 nsb = ( n + 1023 ) / 1024;
    unsigned *bsum = A.get<unsigned>( "tb.bsum", size_t( nsb ) );
    const unsigned nbk = blocks_for( n, 256 );

    // 1. the two sorted lists
    for ( int ax = 0; ax < 2; ++ax ) {
        k_keys<<<nbk, 256, 0, st>>>( n, pos, 2, ax, ka, va );
        const Sorted r = radix_sort( n, ka, va, kb, vb, hist, st );
        SDOT_HIP( hipMemcpyAsync( ax ? yid : xid, r.val, sizeof( int ) * size_t( n ), hipMemcpyDeviceToDevice, st ) );
    }
    // 2. the levels
    const int zero = 0, nn = int( n );
    SDOT_HIP( hipMemcpyAsync( beg, &zero, sizeof( int ), hipMemcpyHostToDevice, st ) );
    SDOT_HIP( hipMemcpyAsync( end, &nn, sizeof( int ), hipMemcpyHostToDevice, st ) );
    for ( int L = 0; L < depth; ++L ) {
        const int count = 1 << L, first = count - 1;
        k_level_nodes<<<blocks_for( count, 256 ), 256, 0, st>>>( count, first, beg, end, pos, xid, yid, leaf, hbox, axis, mid, hbeg, hend );
        if ( L + 1 == depth ) break;
        k_fill_seg<<<count, 256, 0, st>>>( count, beg, end, seg );
        k_flags<<<nbk, 256, 0, st>>>( n, seg, axis, mid, xid, yid, flag );
        k_other_flags<<<nbk, 256, 0, st>>>( n, seg, axis, xid, yid, flag, f );
        k_scan_blocks<<<unsigned( nsb ), 1024, 0, st>>>( n, f, s, bsum );
        k_scan_single<<<1, 1024, 0, st>>>( bsum, int( nsb ) );
        k_scan_add<<<nbk, 256, 0, st>>>( n, s, bsum );
        k_partition<<<nbk, 256, 0, st>>>( n, seg, axis, beg, mid, xid, yid, flag, s, xo, yo );
        std::swap( xid, xo );
        std::swap( yid, yo );
        k_children<<<blocks_for( count, 256 ), 256, 0, st>>>( count, beg, end, mid, nbeg, nend );
        std::swap( beg, nbeg );
        std::swap( end, nend );
        check_launch( "tree level" );
    }
    // 3. leaves, rank tables, preorder
    const int L = depth - 1;
    k_leaves<<<blocks_for( I64( 1 ) << L, 256 ), 256, 0, st>>>( 1 << L, ( 1 << L ) - 1, hbeg, hend, xid );
    k_rank_tables<<<nbk, 256, 0, st>>>( n, xid, pos, t.order, t.rank_of, t.seeds );
    for ( int l = 0; l < depth; ++l ) k_preorder<<<blocks_for( I64( 1 ) << l, 256 ), 256, 0, st>>>( l, depth, t.lvl_nodes );
    k_to_preorder<<<blocks_for( N, 256 ), 256, 0, st>>>( N, t.lvl_nodes, hbeg, hend, hbox, t.node_beg, t.node_end, t.node_box );
    check_launch( "tree finish" );
    // `beg` / `end` came from host constants ( `zero`, `nn` ): the copies must be done before they go
    SDOT_HIP( hipStreamSynchronize( st ) );
    t.h_lvl_nodes.clear();
}

} // namespace sdot_amd
