"""The 3D oracle pinned: quadrature rules exact on monomials, cells against scipy's half-space intersection ( volumes,
active faces, face areas ) and against the cut by every seed, the Laplacian against finite differences of the masses,
sums of Gaussians against the closed form over the box and an independent high-order integration, the tree's
invariants, and solves that reach the tolerance."""
import math
import itertools

import numpy as np
import pytest
from scipy.spatial import ConvexHull, Delaunay, HalfspaceIntersection

import oracle as O

LO, HI = np.zeros( 3 ), np.ones( 3 )


def gaussians( rng, ng = 4, sigma = 0.1 ):
    mu = rng.uniform( 0.2, 0.8, ( ng, 3 ) )
    mass = rng.dirichlet( np.ones( ng ) )
    return np.column_stack( [ mu, np.full( ng, sigma ), mass ] )


def density_at( g, x ):
    r = np.zeros( len( x ) )
    for cx, cy, cz, s, m in g:
        d2 = ( ( x - [ cx, cy, cz ] ) ** 2 ).sum( 1 )
        r += m / ( 2 * np.pi * s * s ) ** 1.5 * np.exp( -d2 / ( 2 * s * s ) )
    return r


# ---- the quadrature rules ------------------------------------------------------------------------------------------

def test_the_tetrahedron_rule_is_exact_to_degree_five():
    tet, _ = O.rules()
    b, w = tet[ :, :4 ], tet[ :, 4 ]
    assert np.all( w > 0 ) and abs( w.sum() - 1 ) < 1e-15
    for deg in range( 6 ):
        for e in itertools.product( range( deg + 1 ), repeat = 4 ):
            if sum( e ) != deg:
                continue
            exact = 6 * math.prod( math.factorial( k ) for k in e ) / math.factorial( deg + 3 )
            got = ( w * np.prod( b ** np.array( e ), axis = 1 ) ).sum()
            assert abs( got - exact ) < 1e-15 * max( 1, 1 / exact ), ( e, got, exact )


def test_the_triangle_rule_is_exact_to_degree_five():
    _, tri = O.rules()
    b, w = tri[ :, :3 ], tri[ :, 3 ]
    assert np.all( w > 0 ) and abs( w.sum() - 1 ) < 1e-15
    for deg in range( 6 ):
        for e in itertools.product( range( deg + 1 ), repeat = 3 ):
            if sum( e ) != deg:
                continue
            exact = 2 * math.prod( math.factorial( k ) for k in e ) / math.factorial( deg + 2 )
            got = ( w * np.prod( b ** np.array( e ), axis = 1 ) ).sum()
            assert abs( got - exact ) < 1e-15 * max( 1, 1 / exact ), ( e, got, exact )


# ---- cells ---------------------------------------------------------------------------------------------------------

def halfspaces_of( pos, w, i ):
    """the cell of seed `i` as scipy half-spaces `A x + b <= 0`, with the seed index of each ( -1 - f: box face f )"""
    p0 = pos[ i ]
    rows, ids = [], []
    for j in range( len( pos ) ):
        if j == i:
            continue
        d = pos[ j ] - p0
        off = ( ( w[ i ] - w[ j ] ) + d @ d ) / 2 + d @ p0      # d . x <= off in the global frame
        rows.append( [ *d, -off ] )
        ids.append( j )
    for ax in range( 3 ):
        e = np.eye( 3 )[ ax ]
        rows.append( [ *( -e ), LO[ ax ] ] ); ids.append( -1 - 2 * ax )
        rows.append( [ *e, -HI[ ax ] ] ); ids.append( -2 - 2 * ax )
    return np.array( rows ), np.array( ids )


def scipy_cell( pos, w, i, interior ):
    hs, ids = halfspaces_of( pos, w, i )
    h = HalfspaceIntersection( hs, interior )
    vol = ConvexHull( h.intersections ).volume
    areas = {}
    for f in range( len( hs ) ):
        pts = h.intersections[ [ f in df for df in h.dual_facets ] ]
        if len( pts ) < 3:
            continue
        n = hs[ f, :3 ] / np.linalg.norm( hs[ f, :3 ] )
        u = np.cross( n, [ 1, 0, 0 ] if abs( n[ 0 ] ) < 0.9 else [ 0, 1, 0 ] )
        u /= np.linalg.norm( u )
        v = np.cross( n, u )
        q = np.column_stack( [ pts @ u, pts @ v ] )
        try:
            a = ConvexHull( q ).volume
        except Exception:
            continue
        if a > 1e-14:
            areas[ int( ids[ f ] ) ] = a
    return vol, areas


@pytest.mark.parametrize( "seed", [ 0, 1 ] )
def test_cells_match_scipy_halfspace_intersection( seed ):
    rng = np.random.default_rng( seed )
    n = 60
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    w = rng.uniform( 0, 2e-3, n )                        # small weights: every seed keeps its cell and sits inside it
    m = O.measures_3d( pos, w )
    assert abs( m.sum() - 1 ) < 1e-13
    for i in range( 0, n, 3 ):
        xyz, ids, areas = O.cell_3d( pos, i, w )
        assert len( xyz ) > 0
        inside = xyz.mean( 0 )
        vol, ref = scipy_cell( pos, w, i, inside )
        assert abs( m[ i ] - vol ) < 1e-12, ( i, m[ i ], vol )
        got = { int( k ): a for k, a in zip( ids, areas ) if a > 1e-14 }
        assert set( got ) == set( ref ), ( i, sorted( got ), sorted( ref ) )
        for k in got:
            assert abs( got[ k ] - ref[ k ] ) < 1e-11 * max( 1, ref[ k ] ), ( i, k, got[ k ], ref[ k ] )


def test_the_tree_walk_finds_the_cells_of_the_cut_by_every_seed():
    rng = np.random.default_rng( 3 )
    n = 400
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    w = rng.normal( 0, 3e-3, n )
    for i in range( 0, n, 17 ):
        a = O.cell_3d( pos, i, w )
        b = O.cell_3d( pos, i, w, brute_force = True )
        assert sorted( a[ 1 ].tolist() ) == sorted( b[ 1 ].tolist() )
        assert np.allclose( np.sort( a[ 2 ] ), np.sort( b[ 2 ] ), rtol = 0, atol = 1e-14 )


def test_one_seed_takes_the_box_and_a_grid_cuts_equal_cells():
    assert np.allclose( O.measures_3d( np.array( [ [ 0.3, 0.6, 0.2 ] ] ) ), [ 1.0 ] )
    g = ( np.arange( 4 ) + 0.5 ) / 4
    pos = np.array( list( itertools.product( g, g, g ) ) )
    m = O.measures_3d( pos )
    assert np.allclose( m, 1 / 64, rtol = 0, atol = 1e-15 )
    lo, hi = ( -1.0, 0.0, 2.0 ), ( 3.0, 0.5, 2.5 )
    m = O.measures_3d( pos * [ 4, 0.5, 0.5 ] + [ -1, 0, 2 ], lo = lo, hi = hi )
    assert np.allclose( m, 4 * 0.5 * 0.5 / 64, rtol = 1e-14 )


def test_a_common_weight_shift_changes_nothing():
    rng = np.random.default_rng( 4 )
    pos = rng.uniform( 0, 1, ( 300, 3 ) )
    w = rng.normal( 0, 1e-3, 300 )
    a = O.measures_3d( pos, w )
    b = O.measures_3d( pos, w + 0.37 )
    assert np.allclose( a, b, rtol = 0, atol = 1e-15 )


@pytest.mark.parametrize( "kind", [ "lebesgue", "gaussians" ] )
def test_the_laplacian_is_the_derivative_of_the_masses( kind ):
    rng = np.random.default_rng( 5 )
    n = 120
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    w = rng.normal( 0, 1e-3, n )
    dens = None if kind == "lebesgue" else dict( gaussians = gaussians( rng, sigma = 0.25 ) )
    m, ( fi, fj, fc ) = O.measures_3d( pos, w, density = dens, facets = True )
    L = np.zeros( ( n, n ) )
    for i, j, c in zip( fi, fj, fc ):
        L[ i, j ] += c
    assert np.allclose( L, L.T, rtol = 1e-9, atol = 1e-13 )
    h = 1e-6
    for j in range( 0, n, 11 ):
        wp, wm = w.copy(), w.copy()
        wp[ j ] += h
        wm[ j ] -= h
        fd = ( O.measures_3d( pos, wp, density = dens ) - O.measures_3d( pos, wm, density = dens ) ) / ( 2 * h )
        expect = -L[ :, j ]                              # d m_i / d w_j = -c_ij, and d m_j / d w_j = sum_i c_ij
        expect[ j ] = L[ j ].sum()
        assert np.allclose( fd, expect, rtol = 0, atol = 2e-5 * np.abs( expect ).max() ), j


def test_the_gaussians_over_the_box_give_the_erf_closed_form():
    rng = np.random.default_rng( 6 )
    g = gaussians( rng, sigma = 0.12 )
    exact = 0.0
    for cx, cy, cz, s, mm in g:
        f = 1.0
        for c in ( cx, cy, cz ):
            f *= 0.5 * ( math.erf( ( 1 - c ) / ( s * math.sqrt( 2 ) ) ) - math.erf( ( 0 - c ) / ( s * math.sqrt( 2 ) ) ) )
        exact += mm * f
    pos = rng.uniform( 0, 1, ( 4000, 3 ) )
    m = O.measures_3d( pos, density = dict( gaussians = g ) )
    assert abs( m.sum() - exact ) < 2e-6 * exact, ( m.sum(), exact )


def conical_rule( q ):
    """a conical-product Gauss rule on the unit tetrahedron ( Stroud ): points `[ q^3, 3 ]`, weights summing to 1 / 6"""
    x, wx = np.polynomial.legendre.leggauss( q )
    x, wx = ( x + 1 ) / 2, wx / 2
    pts, wts = [], []
    for u, wu in zip( x, wx ):
        for v, wv in zip( x, wx ):
            for t, wt in zip( x, wx ):
                s1 = u
                s2 = ( 1 - u ) * v
                s3 = ( 1 - u ) * ( 1 - v ) * t
                pts.append( ( s1, s2, s3 ) )
                wts.append( wu * wv * wt * ( 1 - u ) ** 2 * ( 1 - v ) )
    return np.array( pts ), np.array( wts )


def test_gaussian_cell_masses_match_an_independent_integration():
    """the masses ( fluxes through the faces by the degree-5 triangle rule ) against a degree-19 conical product rule on
    scipy's Delaunay tetrahedra of the cell. The flux is a difference of face integrals: its error is absolute, a few
    1e-12 of the mean mass for cells of a third of sigma, so the bound is relative to the larger of the mass and the mean"""
    rng = np.random.default_rng( 7 )
    n = 30000
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    g = gaussians( rng, sigma = 0.1 )
    m = O.measures_3d( pos, density = dict( gaussians = g ) )
    P, W = conical_rule( 10 )
    for i in range( 0, n, n // 13 ):
        xyz, _, _ = O.cell_3d( pos, i )
        tri = Delaunay( xyz )
        tot = 0.0
        for s in tri.simplices:
            a, b, c, d = xyz[ s ]
            J = np.column_stack( [ b - a, c - a, d - a ] )
            x = a + P @ J.T
            tot += abs( np.linalg.det( J ) ) * ( W * density_at( g, x ) ).sum()
        assert abs( m[ i ] - tot ) < 1e-6 * max( tot, 1 / n ), ( i, m[ i ], tot )


# ---- the tree ------------------------------------------------------------------------------------------------------

def test_the_3d_tree_holds_every_seed_once_and_its_majorants_majorate():
    rng = np.random.default_rng( 8 )
    n = 5000
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    w = pos @ [ 0.3, -0.2, 0.1 ] + rng.normal( 0, 1e-3, n )
    t = O.build_tree_3d( pos, w )
    assert t[ "depth" ] == O.max_depth_for( n, 10 )
    assert sorted( t[ "order" ].tolist() ) == list( range( n ) )
    ranked = pos[ t[ "order" ] ]
    wr = w[ t[ "order" ] ]
    N = len( t[ "beg" ] )
    fitted = 0
    for m in range( N ):
        b, e = t[ "beg" ][ m ], t[ "end" ][ m ]
        if e <= b:
            continue
        p = ranked[ b:e ]
        assert np.all( p >= t[ "box" ][ m, :3 ] ) and np.all( p <= t[ "box" ][ m, 3: ] )
        assert np.all( wr[ b:e ] <= p @ t[ "wa" ][ m ] + t[ "wb" ][ m ] )
        fitted += np.any( t[ "wa" ][ m ] != 0 )
    assert fitted > N // 4                              # a clear linear trend: most nodes keep the fitted slope


# ---- solves --------------------------------------------------------------------------------------------------------

@pytest.mark.parametrize( "kind", [ "lebesgue", "gaussians" ] )
def test_the_3d_solve_reaches_the_tolerance( kind ):
    rng = np.random.default_rng( 9 )
    n = 1500
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    dens = None if kind == "lebesgue" else dict( gaussians = gaussians( rng, sigma = 0.15 ) )
    tol = 1e-9 * 1e6 / n / 1e3
    r = O.solve_3d( pos, density = dens, tol = tol )
    st = r[ "stats" ]
    assert st[ "status" ] == 1, st
    target = np.full( n, st[ "domain_mass" ] / n )
    assert np.abs( r[ "masses" ] - target ).max() <= tol
    # the masses are those of a fresh diagram of the returned weights
    m = O.measures_3d( pos, r[ "weights" ], density = dens )
    assert np.abs( m - r[ "masses" ] ).max() < 1e-15
    assert np.all( np.isfinite( r[ "barycenters" ] ) ) and r[ "cost" ] > 0


def test_a_grid_is_solved_at_the_start():
    g = ( np.arange( 5 ) + 0.5 ) / 5
    pos = np.array( list( itertools.product( g, g, g ) ) )
    r = O.solve_3d( pos, tol = 1e-13 )
    assert r[ "stats" ][ "status" ] == 1 and r[ "stats" ][ "nb_iter" ] == 0
    assert np.allclose( r[ "barycenters" ], pos, atol = 1e-14 )
    assert abs( r[ "cost" ] - 3 * ( 0.2 ** 2 / 12 ) ) < 1e-15      # each cube of side h: int |x - c|^2 = h^5 / 4, summed


# ---- voxel images --------------------------------------------------------------------------------------------------

def voxel_image( rng, shape = ( 4, 4, 5 ), origin = ( -0.1, 0.05, 0.0 ), h = ( 0.3, 0.25, 0.2 ) ):
    """random voxels on a grid that overhangs the unit box along x, leaves a strip of it empty along y and fits it along z"""
    return dict( voxels = rng.uniform( 0.5, 2.0, shape ), x0 = origin[ 0 ], y0 = origin[ 1 ], z0 = origin[ 2 ], hx = h[ 0 ], hy = h[ 1 ],
                 hz = h[ 2 ] )


def polytope_piece( hs, lo, hi ):
    """`( volume, centroid )` of the scipy half-spaces `hs` cut by the block `[ lo, hi ]`, or None when it is ( nearly ) empty:
    the interior point is the Chebyshev centre ( a linear program ), the volume a fan of tetrahedra from it"""
    from scipy.optimize import linprog
    A = np.concatenate( [ hs ] + [ np.array( [ [ *( -np.eye( 3 )[ ax ] ), lo[ ax ] ], [ *np.eye( 3 )[ ax ], -hi[ ax ] ] ] ) for ax in range( 3 ) ] )
    nrm = np.linalg.norm( A[ :, :3 ], axis = 1 )
    lp = linprog( [ 0, 0, 0, -1 ], A_ub = np.column_stack( [ A[ :, :3 ], nrm ] ), b_ub = -A[ :, 3 ], bounds = [ ( None, None ) ] * 3 + [ ( 0, None ) ] )
    if lp.status != 0 or lp.x[ 3 ] < 1e-9:
        return None
    pts = HalfspaceIntersection( A, lp.x[ :3 ] ).intersections
    hull = ConvexHull( pts )
    c = pts.mean( 0 )
    vol, mom = 0.0, np.zeros( 3 )
    for s in hull.simplices:
        a, b, d = pts[ s ]
        v = abs( np.linalg.det( np.array( [ a - c, b - c, d - c ] ) ) ) / 6
        vol += v
        mom += v * ( a + b + d + c ) / 4
    return vol, mom / vol


def image_integrals( pos, w, i, img ):
    """`( mass, first moment )` of cell `i` against the voxel image, voxel by voxel with scipy"""
    hs, _ = halfspaces_of( pos, w, i )
    xyz, _, _ = O.cell_3d( pos, i, w )
    v = img[ "voxels" ]
    o = np.array( [ img[ "x0" ], img[ "y0" ], img[ "z0" ] ] )
    h = np.array( [ img[ "hx" ], img[ "hy" ], img[ "hz" ] ] )
    a = np.clip( np.floor( ( xyz.min( 0 ) - o ) / h ).astype( int ), 0, np.array( v.shape ) - 1 )
    b = np.clip( np.floor( ( xyz.max( 0 ) - o ) / h ).astype( int ), 0, np.array( v.shape ) - 1 )
    mass, first = 0.0, np.zeros( 3 )
    for ijk in itertools.product( *[ range( a[ d ], b[ d ] + 1 ) for d in range( 3 ) ] ):
        lo = o + h * np.array( ijk )
        p = polytope_piece( hs, lo, lo + h )
        if p is not None:
            mass += v[ ijk ] * p[ 0 ]
            first += v[ ijk ] * p[ 0 ] * p[ 1 ]
    return mass, first


def test_a_constant_image_is_the_constant_density():
    rng = np.random.default_rng( 21 )
    n = 80
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    w = rng.normal( 0, 1e-3, n )
    img = dict( voxels = np.full( ( 3, 5, 4 ), 1.7 ), x0 = 0.0, y0 = 0.0, z0 = 0.0, hx = 1 / 3, hy = 0.2, hz = 0.25 )
    m, ( fi, fj, fc ) = O.measures_3d( pos, w, density = img, facets = True )
    mc, ( gi, gj, gc ) = O.measures_3d( pos, w, density = 1.7, facets = True )
    assert np.allclose( m, mc, rtol = 1e-13, atol = 0 )
    A = { ( i, j ): c for i, j, c in zip( fi, fj, fc ) }
    B = { ( i, j ): c for i, j, c in zip( gi, gj, gc ) }
    assert set( A ) == set( B )
    assert max( abs( A[ k ] - B[ k ] ) / abs( B[ k ] ) for k in B ) < 1e-12


def test_image_masses_match_scipy_voxel_by_voxel():
    rng = np.random.default_rng( 22 )
    n = 24
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    w = rng.uniform( 0, 2e-3, n )
    img = voxel_image( rng )
    m = O.measures_3d( pos, w, density = img )
    # the image's mass inside the box: x from 0 to 1 of the grid's -0.1 to 1.1, y from 0.05 to 1 ( zero below: the image is
    # zero outside its grid ), z all of it
    v = img[ "voxels" ]
    fx = np.array( [ 0.2, 0.3, 0.3, 0.2 ] ) / 0.3          # ( the share of each column and row inside the box )
    fy = np.array( [ 0.25, 0.25, 0.25, 0.2 ] ) / 0.25
    inside = np.einsum( "ijk,i,j->", v, fx, fy ) * 0.3 * 0.25 * 0.2
    assert abs( m.sum() - inside ) < 1e-13 * inside
    for i in range( n ):
        mass, _ = image_integrals( pos, w, i, img )
        assert abs( m[ i ] - mass ) < 1e-11 * max( mass, 1e-3 ), ( i, m[ i ], mass )


def test_image_masses_are_linear_in_the_image():
    rng = np.random.default_rng( 23 )
    pos = rng.uniform( 0, 1, ( 150, 3 ) )
    a, b = voxel_image( rng ), voxel_image( rng )
    ab = dict( a, voxels = 2.0 * a[ "voxels" ] - 0.5 * b[ "voxels" ] )
    ma, mb, mab = ( O.measures_3d( pos, density = d ) for d in ( a, b, ab ) )
    assert np.allclose( mab, 2.0 * ma - 0.5 * mb, rtol = 0, atol = 1e-15 * np.abs( ma ).max() * 10 )


def test_the_image_laplacian_is_the_derivative_of_the_masses():
    rng = np.random.default_rng( 24 )
    n = 100
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    w = rng.normal( 0, 1e-3, n )
    img = voxel_image( rng )
    m, ( fi, fj, fc ) = O.measures_3d( pos, w, density = img, facets = True )
    L = np.zeros( ( n, n ) )
    for i, j, c in zip( fi, fj, fc ):
        L[ i, j ] += c
    assert np.allclose( L, L.T, rtol = 1e-9, atol = 1e-13 )
    h = 1e-6
    for j in range( 0, n, 9 ):
        wp, wm = w.copy(), w.copy()
        wp[ j ] += h
        wm[ j ] -= h
        fd = ( O.measures_3d( pos, wp, density = img ) - O.measures_3d( pos, wm, density = img ) ) / ( 2 * h )
        expect = -L[ :, j ]
        expect[ j ] = L[ j ].sum()
        assert np.allclose( fd, expect, rtol = 0, atol = 2e-5 * np.abs( expect ).max() ), j


def test_the_3d_image_solve_reaches_the_tolerance_and_its_barycentres_are_the_voxel_integrals():
    rng = np.random.default_rng( 25 )
    n = 300
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    pos[ :, 1 ] = 0.05 + 0.95 * pos[ :, 1 ]               # ( no seed in the empty strip )
    img = voxel_image( rng )
    r = O.solve_3d( pos, density = img, tol = 1e-12 )
    st = r[ "stats" ]
    assert st[ "status" ] == 1, st
    assert np.abs( r[ "masses" ] - st[ "domain_mass" ] / n ).max() <= 1e-12
    for i in range( 0, n, 23 ):
        mass, first = image_integrals( pos, r[ "weights" ], i, img )
        assert abs( r[ "masses" ][ i ] - mass ) < 1e-10 * mass
        assert np.allclose( r[ "barycenters" ][ i ], first / mass, rtol = 0, atol = 1e-10 ), i
