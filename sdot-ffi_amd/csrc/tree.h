// The BSP tree of a 2D cloud and the affine majorants of its weights, as the cell kernels read them.
//
// Shape ( reference `src/sdot/AaBsp.py:264-279`, `cell/Providers.h:224-233` ): a perfect binary
// tree of `depth = ceil( log2( n / leaf ) ) + 1` levels in PREORDER, node `m` of height `h` having
// its children at `m + 1` and `m + 2^( h - 1 )`; a node's seeds are the contiguous rank slice
// `[ beg, end )`; a node that cannot be cut passes its whole slice to the left child.
#pragma once

#include "common.h"

namespace sdot_amd {

/// one node as the walk reads it: one 64-byte record ( box, majorant `w <= wa . y + wb`, slice )
struct alignas( 64 ) NodeRec {
    double lo_x, lo_y, hi_x, hi_y;
    double wa_x, wa_y, wb;
    int    beg, end;
};
static_assert( sizeof( NodeRec ) == 64, "one node, one 64-byte record" );

/// the same node for the FP32 walk ( the role of the reference card's float `Node`, `gpu/Cell2D.cuh:167-169, 1462-1489` ):
/// the box rounded OUTWARD to float, the slopes rounded to nearest and `wb` rounded UP after adding what the slopes' rounding
/// can change on the box, so that `w <= wa . y + wb` still holds for the node's seeds; `beg` the first seed's rank, -1 for a
/// node with no seed ( a leaf's end is read from `node_end` ). 32 bytes: two 16-byte loads
struct alignas( 32 ) NodeF {
    float lo_x, lo_y, hi_x, hi_y;
    float wa_x, wa_y, wb;
    int   beg;
};
static_assert( sizeof( NodeF ) == 32, "one float node, one 32-byte record" );

/// the tree on the device, plus what its host build keeps
struct DevTree {
    I64  n = 0, nb_nodes = 0;
    int  depth = 0, leaf = 10;
    // device
    NodeRec *nodes = nullptr;     ///< [ N ] the records the walk reads ( majorants refreshed per diagram )
    NodeF   *nodes32 = nullptr;   ///< [ N ] the FP32 walk's records ( refreshed with `nodes` when set )
    double  *node_box = nullptr;  ///< [ N ][ 4 ] lo x, lo y, hi x, hi y
    int     *node_beg = nullptr;  ///< [ N ]
    int     *node_end = nullptr;  ///< [ N ]
    int     *lvl_nodes = nullptr; ///< the nodes level by level ( level `L` at `[ 2^L - 1, 2^( L + 1 ) - 1 )` )
    int     *order = nullptr;     ///< [ n ] rank -> user index
    int     *rank_of = nullptr;   ///< [ n ] user index -> rank
    double4 *seeds = nullptr;     ///< [ n ] ( x, y, w, 0 ) in rank order
    // majorant scratch
    double  *stat = nullptr;      ///< [ 11 ][ N ] centred moments per node
    double  *fit = nullptr;       ///< [ 2 ][ N ]
    unsigned long long *rkeys = nullptr;  ///< [ 3 ][ N ] ordered keys: min r, max r, max |a.y|
    // host copies of the level structure
    std::vector<int> h_lvl_nodes;
};

/// `AaBsp.py:264`
int max_depth_for( I64 n, I64 leaf );

/// builds the tree of `pos_host[ n ][ 2 ]` ( user order ) on the host and uploads it; the arena
/// buffers are named with `tag` so two trees can live side by side
void build_tree( DevTree &t, I64 n, const double *pos_host, int leaf, hipStream_t st, const std::string &tag );

/// builds the same tree on the device from `pos_dev[ n ][ 2 ]` ( device, user order ): no host round trip
void build_tree_device( DevTree &t, I64 n, const double *pos_dev, int leaf, hipStream_t st, const std::string &tag );

/// rank-ordered weights are in `t.seeds[ k ].z`: the majorants of every node and the walk records
void refresh_majorants( DevTree &t, hipStream_t st );

/// writes `w_user[ order[ k ] ]` into `seeds[ k ].z`
void set_weights_user( DevTree &t, const double *w_user_dev, hipStream_t st );
/// writes `w_rank[ k ]` into `seeds[ k ].z`
void set_weights_rank( DevTree &t, const double *w_rank_dev, hipStream_t st );

} // namespace sdot_amd
