"""Polygon domains on the CPU: the host's start polygon ( `StartCell.start_cell` ) against scipy's half-plane intersection,
and the oracle's cells started from a polygon against closed forms, scipy's half-space intersections and finite differences,
and its solves in a polygon."""

import numpy as np
import pytest

import helpers as H


def _scipy_polygon( dirs, offs ):
    from scipy.optimize import linprog
    from scipy.spatial import ConvexHull, HalfspaceIntersection
    nrm = np.linalg.norm( dirs, axis = 1 )
    res = linprog( [ 0, 0, -1 ], A_ub = np.column_stack( [ dirs, nrm ] ), b_ub = offs, bounds = [ ( None, None ), ( None, None ), ( 0, None ) ] )
    h = HalfspaceIntersection( np.column_stack( [ dirs, -offs ] ), res.x[ :2 ] )
    hull = ConvexHull( h.intersections )
    return h.intersections[ hull.vertices ], hull.volume


@pytest.mark.parametrize( "seed", range( 6 ) )
def test_start_cell_is_scipys_intersection( seed ):
    from sdot.StartCell import start_cell
    rng = np.random.default_rng( seed )
    k = rng.integers( 3, 10 )
    a = np.sort( rng.uniform( 0, 2 * np.pi, k ) )
    dirs = np.column_stack( [ np.cos( a ), np.sin( a ) ] ) * rng.uniform( 0.5, 2, ( k, 1 ) )
    offs = rng.uniform( 0.2, 1.0, k ) * np.linalg.norm( dirs, axis = 1 )
    # a bounded set: the normals must not leave a half-circle empty ( add a triangle around when they do )
    dirs = np.concatenate( [ dirs, [ [ 1, 0 ], [ -0.5, 0.8660254 ], [ -0.5, -0.8660254 ] ] ] )
    offs = np.concatenate( [ offs, [ 3, 3, 3 ] ] )
    cell = start_cell( dirs, offs )
    ref, area = _scipy_polygon( dirs, offs )
    v = cell[ "vertices" ]
    assert abs( H.polygon_area( v ) - area ) <= 1e-12 * area and H.polygon_area( v ) > 0
    assert len( v ) == len( ref )
    for p in ref:
        assert np.abs( v - p ).max( axis = 1 ).min() <= 1e-10


def test_start_cell_refuses_what_is_not_a_bounded_polygon():
    from sdot.StartCell import start_cell
    assert start_cell( [ [ 1, 0 ], [ -1, 0 ], [ 0, 1 ] ], [ 1, 1, 1 ] ) is None          # unbounded below
    assert start_cell( [ [ 1, 0 ], [ -1, 0 ], [ 0, 1 ], [ 0, -1 ] ], [ 1, -2, 1, 1 ] ) is None   # empty
    c = start_cell( [ [ 1, 0 ], [ -1, 0 ], [ 0, 1 ], [ 0, -1 ], [ 1, 1 ] ], [ 1, 0, 1, 0, 5 ] )   # a redundant half-plane
    assert len( c[ "vertices" ] ) == 4 and abs( c[ "area" ] - 1 ) < 1e-15


@pytest.mark.parametrize( "m", [ 3, 6, 11 ] )
def test_oracle_cells_in_a_polygon_are_scipys( oracle, m ):
    rng = np.random.default_rng( 40 + m )
    ( dirs, offs ), v = H.regular_polygon( m )
    n = 300
    pos = rng.uniform( 0.05, 0.95, ( n, 2 ) )
    w = rng.normal( 0, 1e-3, n )
    lo, hi = v.min( axis = 0 ), v.max( axis = 0 )
    mass = oracle.measures( pos, w, lo, hi, domain = v )
    assert abs( mass.sum() - H.polygon_area( v ) ) <= 1e-13                       # ( the Lebesgue measure: the polygon's area )
    for i in range( 0, n, 7 ):
        c = H.halfspace_cell( pos, w, i, lo, hi, extra = ( dirs, offs ) )
        ref = 0.0 if c is None else H.polygon_area( c )
        assert abs( mass[ i ] - ref ) <= 1e-12, ( i, mass[ i ], ref )
    # the domain's edges carry the cuts -5 - e ( each edge of the polygon bounds some cell ); the box's sides none
    seen = set()
    for i in range( n ):
        xy, ids = oracle.cell( pos, i, w, lo, hi, domain = v )
        seen |= { int( c ) for c in ids if c < 0 }
    assert seen == { -5 - e for e in range( m ) }


def test_oracle_polygon_of_the_box_is_the_box( oracle ):
    rng = np.random.default_rng( 44 )
    pos = rng.uniform( 0, 1, ( 500, 2 ) )
    w = rng.normal( 0, 1e-3, 500 )
    box = np.array( [ [ 0, 0 ], [ 1, 0 ], [ 1, 1 ], [ 0, 1 ] ], dtype = float )
    a = oracle.measures( pos, w )
    b = oracle.measures( pos, w, domain = box )
    assert np.abs( a - b ).max() <= 1e-17


def test_oracle_facets_in_a_polygon_are_the_masses_derivative( oracle ):
    rng = np.random.default_rng( 45 )
    ( dirs, offs ), v = H.regular_polygon( 7 )
    lo, hi = v.min( axis = 0 ), v.max( axis = 0 )
    n = 40
    pos = rng.uniform( 0.2, 0.8, ( n, 2 ) )
    w = rng.normal( 0, 1e-3, n )
    m, ( fi, fj, fc ) = oracle.measures( pos, w, lo, hi, facets = True, domain = v )
    L = np.zeros( ( n, n ) )
    for i, j, c in zip( fi, fj, fc ):
        L[ i, j ] -= c
        L[ i, i ] += c
    h = 1e-7
    for j in range( 0, n, 5 ):
        wp, wm = w.copy(), w.copy()
        wp[ j ] += h
        wm[ j ] -= h
        fd = ( oracle.measures( pos, wp, lo, hi, domain = v ) - oracle.measures( pos, wm, lo, hi, domain = v ) ) / ( 2 * h )
        assert np.abs( fd - L[ :, j ] ).max() <= 1e-6 * np.abs( L ).max(), j


@pytest.mark.parametrize( "step", [ "trials", "limits" ] )
def test_oracle_solves_in_a_polygon( oracle, step ):
    """a hexagon of an image density, Diracs in the hexagon: the masses reach the targets ( the targets rescaled to the hexagon's
    mass ), by both steps"""
    rng = np.random.default_rng( 46 )
    ( dirs, offs ), v = H.regular_polygon( 6 )
    lo, hi = v.min( axis = 0 ), v.max( axis = 0 )
    n = 2000
    pos = H.points_in_polygon( n, dirs, offs, rng )
    img = H.bumps_image( 64, 64, rng, k = 4 )
    dens = dict( values = img, x0 = 0.0, y0 = 0.0, hx = 1 / 64, hy = 1 / 64 )
    tol = 1e-11
    r = oracle.solve( pos, lo = lo, hi = hi, density = dens, tol = tol, step = step, domain = v )
    st = r[ "stats" ]
    assert st[ "status" ] == 1
    dm = st[ "domain_mass" ]
    assert abs( dm - H.image_mass_of_polygon( v, img, 0.0, 0.0, 1 / 64, 1 / 64 ) ) <= 1e-12
    m = oracle.measures( pos, r[ "weights" ], lo, hi, density = dens, domain = v )
    assert np.abs( m - dm / n ).max() <= 1.01 * tol
