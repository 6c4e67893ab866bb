"""The cell a power diagram's cells START from when the domain is a convex polygon and not a box ( reference
`src/sdot/StartCell.py`: `start_cell( directions, offsets )` ): the polygon `{ x : direction . x <= offset }` as
counterclockwise vertices, laid down as each cell's first polygon by the kernels ( `sdot_amd_measures_2d_domain`,
`sdot_amd_solve_2d_domain` ), edge `i` ( vertex `i` to `i + 1` ) being the cut `-5 - i`.

2D only on this path. The polygon is found by clipping a square far larger than the offsets by each half-plane in turn
( convex clipping keeps the orientation and needs no interior point ); a vertex left on that square means the half-planes
do not bound the domain. The clipping is then done again from a square twice the polygon's box, so that the vertices carry
the rounding of the polygon's size, not of the first square's.
"""

import numpy as np

#: the vertices a start polygon may have ( the kernels' first passes keep 16 slots per cell: room is left for the cuts )
MAX_VERTICES = 12


def _clip( poly, n, c ):
    """the convex polygon `poly` ( `[ m, 2 ]`, counterclockwise ) cut by `n . x <= c`"""
    out = []
    m = len( poly )
    s = poly @ n - c
    for i in range( m ):
        j = ( i + 1 ) % m
        if s[ i ] <= 0:
            out.append( poly[ i ] )
        if ( s[ i ] < 0 < s[ j ] ) or ( s[ j ] < 0 < s[ i ] ):
            t = s[ i ] / ( s[ i ] - s[ j ] )
            out.append( poly[ i ] + t * ( poly[ j ] - poly[ i ] ) )
    return np.array( out ).reshape( -1, 2 )


def _simplified( poly, tol ):
    """the same polygon without repeated vertices nor vertices on the segment of their two neighbours"""
    changed = True
    while changed and len( poly ) >= 3:
        changed = False
        m = len( poly )
        for i in range( m ):
            a, b, c = poly[ i - 1 ], poly[ i ], poly[ ( i + 1 ) % m ]
            cross = ( b[ 0 ] - a[ 0 ] ) * ( c[ 1 ] - b[ 1 ] ) - ( b[ 1 ] - a[ 1 ] ) * ( c[ 0 ] - b[ 0 ] )
            if np.abs( b - a ).max() <= tol or abs( cross ) <= tol * ( np.abs( c - a ).max() + tol ):
                poly = np.delete( poly, i, axis = 0 )
                changed = True
                break
    return poly


def start_cell( directions, offsets ):
    """`dict( vertices [ m, 2 ] counterclockwise, bounding_box ( lo, hi ) )` of the 2D polygon `{ x : directions . x <= offsets }`,
    or `None` when it is empty, flat or unbounded"""
    dirs = np.asarray( directions, dtype = np.float64 )
    offs = np.asarray( offsets, dtype = np.float64 ).reshape( -1 )
    if dirs.ndim != 2 or dirs.shape[ 1 ] != 2 or len( dirs ) != len( offs ) or len( offs ) < 3:
        return None
    norm = np.sqrt( ( dirs ** 2 ).sum( axis = 1 ) )
    if not np.all( norm > 0 ) or not np.all( np.isfinite( offs ) ):
        return None
    dirs, offs = dirs / norm[ :, None ], offs / norm
    def clipped( square ):
        poly = square
        for n, c in zip( dirs, offs ):
            poly = _clip( poly, n, c )
            if len( poly ) < 3:
                return None
        return poly

    big = 1e6 * max( 1.0, float( np.abs( offs ).max() ) )
    poly = clipped( np.array( [ [ -big, -big ], [ big, -big ], [ big, big ], [ -big, big ] ] ) )
    if poly is None:
        return None
    if np.abs( poly ).max() >= 0.5 * big:
        return None                                      # ( a vertex of the far square: unbounded )
    c0, r = 0.5 * ( poly.min( axis = 0 ) + poly.max( axis = 0 ) ), float( ( poly.max( axis = 0 ) - poly.min( axis = 0 ) ).max() )
    poly = clipped( c0 + r * np.array( [ [ -1.0, -1.0 ], [ 1.0, -1.0 ], [ 1.0, 1.0 ], [ -1.0, 1.0 ] ] ) )
    if poly is None:
        return None
    scale = max( 1.0, float( np.abs( poly ).max() ) )
    poly = _simplified( poly, 1e-12 * scale )
    if len( poly ) < 3:
        return None
    area = 0.5 * float( np.sum( poly[ :, 0 ] * np.roll( poly[ :, 1 ], -1 ) - np.roll( poly[ :, 0 ], -1 ) * poly[ :, 1 ] ) )
    if not area > 1e-12 * scale * scale:
        return None
    return dict( vertices = poly, bounding_box = ( poly.min( axis = 0 ), poly.max( axis = 0 ) ), area = area )


def polygon_area( vertices ):
    v = np.asarray( vertices, dtype = np.float64 ).reshape( -1, 2 )
    return 0.5 * float( np.sum( v[ :, 0 ] * np.roll( v[ :, 1 ], -1 ) - np.roll( v[ :, 0 ], -1 ) * v[ :, 1 ] ) )
