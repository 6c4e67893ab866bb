// The BSP tree of a 3D cloud and the affine majorants of its weights, as the 3D cell kernel reads them.
//
// Same shape as the 2D tree ( reference `src/sdot/AaBsp.py:264-279`, `cell/Providers.h:224-233` ): a perfect binary tree
// of `depth = ceil( log2( n / leaf ) ) + 1` levels in PREORDER, node `m` of height `h` having its children at `m + 1` and
// `m + 2^( h - 1 )`, a node's seeds the contiguous rank slice `[ beg, end )`, cut at the median of the longest axis of its
// box. Majorants `w <= wa . y + wb` as `bsp_build_level.h:84-209` with three coordinates ( `Majorant3D.cuh` on the card ).
#pragma once

#include "common.h"

namespace sdot_amd {

/// one node as the 3D walk reads it: box, majorant, slice ( 96 bytes )
struct alignas( 32 ) NodeRec3 {
    double lo[ 3 ], hi[ 3 ];
    double wa[ 3 ], wb;
    int    beg, end;
};
static_assert( sizeof( NodeRec3 ) == 96, "one 3D node, one 96-byte record" );

struct DevTree3 {
    I64  n = 0, nb_nodes = 0;
    int  depth = 0, leaf = 10;
    NodeRec3 *nodes = nullptr;    ///< [ N ]
    double  *node_box = nullptr;  ///< [ N ][ 6 ] lo x, y, z, hi x, y, z
    int     *node_beg = nullptr, *node_end = nullptr;
    int     *lvl_nodes = nullptr; ///< heap rank -> preorder node
    int     *order = nullptr;     ///< [ n ] rank -> user index
    int     *rank_of = nullptr;   ///< [ n ] user index -> rank
    double4 *seeds = nullptr;     ///< [ n ] ( x, y, z, w ) in rank order
    double  *stat = nullptr;      ///< [ 16 ][ N ] centred moments per node
    double  *fit = nullptr;       ///< [ 3 ][ N ]
    unsigned long long *rkeys = nullptr;   ///< [ 3 ][ N ]
};

/// builds the tree from `pos_dev[ n ][ 3 ]` ( device, user order ): three sorted lists partitioned level by
/// level, no copy of the cloud to the host
void build_tree_device3( DevTree3 &t, I64 n, const double *pos_dev, int leaf, hipStream_t st, const std::string &tag );

/// `w_user[ order[ k ] ]` / `w_rank[ k ]` into `seeds[ k ].w`
void set_weights_user3( DevTree3 &t, const double *w_user_dev, hipStream_t st );
void set_weights_rank3( DevTree3 &t, const double *w_rank_dev, hipStream_t st );

/// the majorants of every node and the walk records from the rank-ordered weights
void refresh_majorants3( DevTree3 &t, hipStream_t st );

} // namespace sdot_amd
