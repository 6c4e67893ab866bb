"""Independent checks shared by the tests: clouds of the reference's test-suite, the power diagram
from scipy's convex hull of the lifted seeds, and exact polygon / pixel clipping in numpy."""

import numpy as np


# ---- clouds ( reference tests/test_CardCells.py:75-102 ) ------------------------------------------------------


def ring( k, rng, n_back = 2000 ):
    """a seed in the middle of a ring of `k` seeds ( its cell is a k-gon ), plus a uniform background"""
    a = 2 * np.pi * ( np.arange( k ) + rng.uniform( -0.1, 0.1, k ) ) / k
    rg = 0.5 + 0.2 * np.stack( [ np.cos( a ), np.sin( a ) ], axis = 1 )
    back = rng.uniform( 0.001, 0.999, size = ( n_back, 2 ) )
    back = back[ np.linalg.norm( back - 0.5, axis = 1 ) > 0.3 ]
    return np.concatenate( [ [ [ 0.5, 0.5 ] ], rg, back ] )


def lines( n, rng ):
    """five lines with a 0.005 normal spread ( thin, long cells )"""
    ends = rng.uniform( 0.05, 0.95, size = ( 5, 2, 2 ) )
    which = rng.integers( 0, 5, n )
    t = rng.uniform( 0, 1, n )[ :, None ]
    pos = ends[ which, 0 ] + t * ( ends[ which, 1 ] - ends[ which, 0 ] ) + rng.normal( 0, 0.005, size = ( n, 2 ) )
    return np.clip( pos, 0.001, 0.999 )


def clouds( rng, n = 3000 ):
    """`( label, positions, weights )`: uniform, lines, weighted lines"""
    uni = rng.uniform( 0.001, 0.999, size = ( n, 2 ) )
    lin = lines( n, rng )
    h2 = 1.0 / n
    return [ ( "uniform", uni, None ),
             ( "lines", lin, None ),
             ( "lines weighted", lin, rng.uniform( -0.5, 0.5, n ) * h2 + 0.05 * np.sin( 4 * lin[ :, 0 ] ) ) ]


# ---- the power diagram by lifting -------------------------------------------------------------------------------


def _clip_length( a, d, t1, lo, hi ):
    """the length of `{ a + t d : 0 <= t <= t1 } ∩ box` ( `t1` may be `inf`: a ray ), by Liang-Barsky"""
    t0 = 0.0
    for k in range( 2 ):
        if d[ k ] == 0:
            if a[ k ] < lo[ k ] or a[ k ] > hi[ k ]:
                return 0.0
            continue
        ta, tb = ( lo[ k ] - a[ k ] ) / d[ k ], ( hi[ k ] - a[ k ] ) / d[ k ]
        t0, t1 = max( t0, min( ta, tb ) ), min( t1, max( ta, tb ) )
        if t1 <= t0:
            return 0.0
    return ( t1 - t0 ) * float( np.hypot( d[ 0 ], d[ 1 ] ) )


def lifted_neighbours( pos, w, lo = ( 0, 0 ), hi = ( 1, 1 ), margin = 1e-10, ambiguous = False ):
    """the neighbour pairs `{ ( i, j ) }` of the power diagram restricted to the box `lo, hi`, from the lower convex hull of
    the lifted points `( p, |p|^2 - w )` ( a regular triangulation: scipy's Qhull, independent of the oracle ). A hull edge's
    power facet is the segment between the power centres of its two lower faces ( a face `z = a x + b y + c` has its centre
    at `( a / 2, b / 2 )` ), or the ray from its one face's centre away from that face's third vertex; a pair is kept when
    that facet meets the box over a length above `margin`. With `ambiguous`, also returns the pairs whose facet meets it over
    a length in `( 0, margin ]`, which rounding may keep or drop."""
    from scipy.spatial import ConvexHull
    pos = np.asarray( pos, dtype = np.float64 )
    w = np.zeros( len( pos ) ) if w is None else np.asarray( w, dtype = np.float64 )
    lift = np.column_stack( [ pos, ( pos ** 2 ).sum( 1 ) - w ] )
    hull = ConvexHull( lift )
    faces = {}                                           # edge -> [ ( centre, third vertex ) ]
    for simplex, eq in zip( hull.simplices, hull.equations ):
        if eq[ 2 ] >= 0:                                 # upper faces are not in the regular triangulation
            continue
        c = np.array( [ -eq[ 0 ] / eq[ 2 ], -eq[ 1 ] / eq[ 2 ] ] ) / 2
        for a in range( 3 ):
            i, j, k = int( simplex[ a ] ), int( simplex[ ( a + 1 ) % 3 ] ), int( simplex[ ( a + 2 ) % 3 ] )
            faces.setdefault( ( min( i, j ), max( i, j ) ), [] ).append( ( c, k ) )
    pairs, amb = set(), set()
    for ( i, j ), fl in faces.items():
        if len( fl ) >= 2:
            length = _clip_length( fl[ 0 ][ 0 ], fl[ 1 ][ 0 ] - fl[ 0 ][ 0 ], 1.0, lo, hi )
        else:
            c, k = fl[ 0 ]
            e = pos[ j ] - pos[ i ]
            nrm = np.array( [ e[ 1 ], -e[ 0 ] ] )
            if nrm @ ( pos[ k ] - pos[ i ] ) > 0:
                nrm = -nrm
            length = _clip_length( c, nrm, np.inf, lo, hi )
        if length > margin:
            pairs.add( ( i, j ) )
        elif length > 0:
            amb.add( ( i, j ) )
    return ( pairs, amb ) if ambiguous else pairs


def halfspace_cell( pos, w, i, lo = ( 0, 0 ), hi = ( 1, 1 ), extra = None ):
    """the vertices of cell `i` ( counter-clockwise ) by half-space intersection ( scipy ), or `None` if empty; `extra`: more
    half-planes `( directions, offsets )` ( `direction . x <= offset`, absolute coordinates ) the domain has"""
    from scipy.spatial import HalfspaceIntersection, ConvexHull
    pos = np.asarray( pos, dtype = np.float64 )
    w = np.zeros( len( pos ) ) if w is None else np.asarray( w, dtype = np.float64 )
    p0 = pos[ i ]
    hs = []
    for j in range( len( pos ) ):
        if j == i:
            continue
        d = pos[ j ] - p0
        off = ( d @ d + w[ i ] - w[ j ] ) / 2
        hs.append( [ d[ 0 ], d[ 1 ], -off ] )
    hs += [ [ 1, 0, -( hi[ 0 ] - p0[ 0 ] ) ], [ -1, 0, lo[ 0 ] - p0[ 0 ] ], [ 0, 1, -( hi[ 1 ] - p0[ 1 ] ) ], [ 0, -1, lo[ 1 ] - p0[ 1 ] ] ]
    if extra is not None:
        for n, c in zip( np.asarray( extra[ 0 ], dtype = np.float64 ), np.asarray( extra[ 1 ], dtype = np.float64 ) ):
            hs.append( [ n[ 0 ], n[ 1 ], -( c - n @ p0 ) ] )
    hs = np.asarray( hs )
    # an interior point: the seed's cell may not contain the seed; use a linear program
    from scipy.optimize import linprog
    A = hs[ :, :2 ]
    nrm = np.linalg.norm( A, axis = 1 )
    res = linprog( [ 0, 0, -1 ], A_ub = np.column_stack( [ A, nrm ] ), b_ub = -hs[ :, 2 ], bounds = [ ( None, None ), ( None, None ), ( 0, None ) ] )
    if not res.success or res.x[ 2 ] < 1e-14:
        return None
    hsi = HalfspaceIntersection( hs, res.x[ :2 ] )
    v = hsi.intersections
    hull = ConvexHull( v )
    return v[ hull.vertices ] + p0


def regular_polygon( m, centre = ( 0.5, 0.5 ), radius = 0.45, phase = 0.1 ):
    """the half-planes `( directions, offsets )` and the counterclockwise vertices of a regular `m`-gon"""
    a = phase + 2 * np.pi * np.arange( m ) / m
    v = np.column_stack( [ centre[ 0 ] + radius * np.cos( a ), centre[ 1 ] + radius * np.sin( a ) ] )
    e = np.roll( v, -1, axis = 0 ) - v
    n = np.column_stack( [ e[ :, 1 ], -e[ :, 0 ] ] )
    n /= np.linalg.norm( n, axis = 1 )[ :, None ]
    return ( n, ( n * v ).sum( axis = 1 ) ), v


def points_in_polygon( n, dirs, offs, rng, lo = ( 0, 0 ), hi = ( 1, 1 ) ):
    """`n` points uniform in the polygon `dirs . x <= offs` ( rejection from the box `[ lo, hi ]` )"""
    out = np.zeros( ( 0, 2 ) )
    while len( out ) < n:
        p = rng.uniform( lo, hi, ( 2 * n, 2 ) )
        out = np.concatenate( [ out, p[ np.all( p @ np.asarray( dirs ).T <= np.asarray( offs ), axis = 1 ) ] ] )
    return out[ :n ]


def polygon_area( xy ):
    x, y = xy[ :, 0 ], xy[ :, 1 ]
    return 0.5 * float( np.sum( x * np.roll( y, -1 ) - np.roll( x, -1 ) * y ) )


def clip_polygon( xy, nx, ny, off ):
    """Sutherland-Hodgman: the part of polygon `xy` where `nx x + ny y <= off`"""
    out = []
    m = len( xy )
    for k in range( m ):
        a, b = xy[ k ], xy[ ( k + 1 ) % m ]
        sa, sb = nx * a[ 0 ] + ny * a[ 1 ] - off, nx * b[ 0 ] + ny * b[ 1 ] - off
        if sa <= 0:
            out.append( a )
        if ( sa <= 0 ) != ( sb <= 0 ):
            t = sa / ( sa - sb )
            out.append( a + t * ( b - a ) )
    return np.asarray( out ).reshape( -1, 2 )


def image_mass_of_polygon( xy, img, x0, y0, hx, hy ):
    """`int_polygon rho` for the image `img[ j, i ]` by exact clipping of the polygon against every pixel it meets"""
    if len( xy ) < 3:
        return 0.0
    ny, nx = img.shape
    i0 = max( int( np.floor( ( xy[ :, 0 ].min() - x0 ) / hx ) ), 0 )
    i1 = min( int( np.floor( ( xy[ :, 0 ].max() - x0 ) / hx ) ), nx - 1 )
    j0 = max( int( np.floor( ( xy[ :, 1 ].min() - y0 ) / hy ) ), 0 )
    j1 = min( int( np.floor( ( xy[ :, 1 ].max() - y0 ) / hy ) ), ny - 1 )
    tot = 0.0
    for j in range( j0, j1 + 1 ):
        for i in range( i0, i1 + 1 ):
            p = clip_polygon( xy, 1, 0, x0 + ( i + 1 ) * hx )
            if len( p ) >= 3: p = clip_polygon( p, -1, 0, -( x0 + i * hx ) )
            if len( p ) >= 3: p = clip_polygon( p, 0, 1, y0 + ( j + 1 ) * hy )
            if len( p ) >= 3: p = clip_polygon( p, 0, -1, -( y0 + j * hy ) )
            if len( p ) >= 3:
                tot += img[ j, i ] * polygon_area( p )
    return tot


def bumps_image( nx, ny, rng, k = 8, floor = 0.1, sig = ( 0.03, 0.15 ) ):
    """`v[ i, j ] = floor + sum_k exp( -|x_ij - c_k|^2 / ( 2 s_k^2 ) )` at pixel centres ( first index along x ):
    the image density of the 1M-Dirac 2D benchmark, at any resolution"""
    c = rng.uniform( 0, 1, ( k, 2 ) )
    s = rng.uniform( sig[ 0 ], sig[ 1 ], k )
    xs = ( np.arange( nx ) + 0.5 ) / nx
    ys = ( np.arange( ny ) + 0.5 ) / ny
    X, Y = np.meshgrid( xs, ys, indexing = "ij" )
    v = np.full( ( nx, ny ), float( floor ) )
    for q in range( k ):
        v += np.exp( -( ( X - c[ q, 0 ] ) ** 2 + ( Y - c[ q, 1 ] ) ** 2 ) / ( 2 * s[ q ] ** 2 ) )
    return v


def degenerate_clouds( seed = 17, n = 400 ):
    """`{ name: ( positions, groups ) }`: a uniform cloud with seeds moved onto or next to others; `groups` the expected
    clusters as `( members, exact )`, `exact` when the positions are equal ( reference tests/test_SdotPlanNd.py:1377-1404 )"""
    rng = np.random.default_rng( seed )
    base = rng.uniform( 0.05, 0.95, size = ( n, 2 ) )
    unit = np.ones( 2 ) / 2 ** 0.5
    side = np.eye( 2 )[ 0 ]

    def moved( *moves ):
        pos = base.copy()
        groups = []
        for members, offsets in moves:
            for m, off in zip( members[ 1: ], offsets ):
                pos[ m ] = pos[ members[ 0 ] ] + off
            groups.append( ( list( members ), all( not np.any( off ) for off in offsets ) ) )
        return pos, groups

    return {
        "exact pair":         moved( ( [ 5, 7 ], [ 0.0 * unit ] ) ),
        "exact triple":       moved( ( [ 3, 11, 29 ], [ 0.0 * unit, 0.0 * unit ] ) ),
        "two pairs":          moved( ( [ 1, 6 ], [ 0.0 * unit ] ), ( [ 12, 15 ], [ 0.0 * unit ] ) ),
        "pair at 1e-9":       moved( ( [ 5, 7 ], [ 1e-9 * unit ] ) ),
        "pair at 1e-12":      moved( ( [ 8, 2 ], [ 1e-12 * unit ] ) ),
        "triple at 1e-10":    moved( ( [ 4, 9, 13 ], [ 1e-10 * unit, 2e-10 * side ] ) ),
        "pair and duplicate": moved( ( [ 1, 6 ], [ 3e-10 * unit ] ), ( [ 12, 15, 30 ], [ 0.0 * unit, 2e-10 * side ] ) ),
    }


EXACT_CLOUDS = ( "exact pair", "exact triple", "two pairs" )
NEAR_CLOUDS = ( "pair at 1e-9", "pair at 1e-12", "triple at 1e-10", "pair and duplicate" )


def brute_duplicates( pos ):
    """the exact duplicates by comparing every pair: `( dup, rep )` rows, `rep` the smallest index at that position"""
    n = len( pos )
    out = []
    for i in range( n ):
        same = np.nonzero( ( pos[ :i ] == pos[ i ] ).all( axis = 1 ) )[ 0 ]
        if len( same ):
            out.append( ( i, int( same[ 0 ] ) ) )
    return np.asarray( out, dtype = np.int64 ).reshape( -1, 2 )


def duplicate_gap( pos ):
    """`1e-6` of the larger of the squared norm and the squared extent of the cloud ( reference src/sdot/SdotPlanNd.py:671-674 )"""
    ext = pos.max( axis = 0 ) - pos.min( axis = 0 )
    s2 = max( float( ( pos * pos ).sum( axis = 1 ).max() ), float( ( ext * ext ).sum() ) )
    return 1e-6 * ( s2 if s2 > 0 else 1.0 )


def check_aggregated( name, st, weights, masses, nu, clusters, groups, tol ):
    """a plan of a cloud whose `groups` the doubles cannot separate ( reference tests/test_SdotPlanNd.py:1407-1433 ): converged
    ( aggregated, unless only exact duplicates ), the clusters found, each cluster's mass its target ( `k` members: `k tol` ),
    every other cell its own, an exact duplicate's cell empty and its weight under its representative's"""
    n = len( nu )
    only_exact = all( exact for _, exact in groups )
    assert st[ "status" ] in ( ( 1, "converged" ) if only_exact else ( 7, "converged (aggregated)" ) ), ( name, st[ "status" ] )
    assert st[ "residual" ] <= tol, ( name, st[ "residual" ] )
    alone = np.ones( n, dtype = bool )
    for members, exact in groups:
        if clusters is not None:
            assert ( clusters[ members ] == min( members ) ).all(), ( name, members, clusters[ members ] )
        alone[ members ] = False
        assert abs( masses[ members ].sum() - nu[ members ].sum() ) <= len( members ) * tol * ( 1 + 1e-9 ), name
        if exact:
            assert ( masses[ members[ 1: ] ] == 0 ).all() and ( weights[ members[ 1: ] ] < weights[ members[ 0 ] ] ).all(), name
    if clusters is not None:
        assert ( clusters[ alone ] == np.arange( n )[ alone ] ).all(), name
    assert np.abs( masses[ alone ] - nu[ alone ] ).max() <= tol, name
    assert st[ "nb_clusters" ] == len( groups ) and st[ "nb_aggregated" ] == sum( len( g ) for g, _ in groups ), name
