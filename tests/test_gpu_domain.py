"""Polygon domains on the MI355X, through the C ABI ( `sdot_amd_measures_2d_domain`, `sdot_amd_solve_2d_domain` ): each
cell starts from a convex polygon instead of the box. Against the oracle ( which starts its cells from the same polygon ):
the FP64 masses and facets to the bit on a constant density ( the same cuts in the same order ), to 1e-12 on an image; the
FP32 kernel's to 1e-12 of the cell; the solve's decisions and plan. Through the API: a `Polytope` target, and a
`PowerDiagram` whose half-planes are not a box."""

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def say( *a ):
    print( *a, flush = True )


def _stats( s ):
    from sdot.SdotPlanNd import _STATS
    return dict( zip( _STATS, s ) )


def _rows( fi, fj, fc ):
    return { ( int( i ), int( j ) ): float( c ) for i, j, c in zip( fi, fj, fc ) }


@pytest.mark.parametrize( "m,kind", [ ( 3, "constant" ), ( 6, "constant" ), ( 11, "constant" ), ( 6, "image" ) ] )
def test_polygon_masses_and_facets_are_the_oracles( native, oracle, m, kind ):
    import sdot
    rng = np.random.default_rng( 60 + m )
    ( dirs, offs ), v = H.regular_polygon( m )
    lo, hi = v.min( axis = 0 ), v.max( axis = 0 )
    n = 20_000
    pos = H.points_in_polygon( n, dirs, offs, rng )
    w = rng.normal( 0, 1e-2 / n, n )
    box = [ lo[ 0 ], lo[ 1 ], hi[ 0 ], hi[ 1 ] ]
    if kind == "image":
        img = H.bumps_image( 128, 128, rng, k = 4 )
        dens = sdot.Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 128 )
        odens = dict( values = np.ascontiguousarray( np.asarray( dens.values ).T ), x0 = 0.0, y0 = 0.0, hx = 1 / 128, hy = 1 / 128 )
    else:
        dens = sdot.UniformBox( lo, hi, target_mass = float( np.prod( hi - lo ) ) )   # ( rho = 1 )
        odens = None
    say( f"case polygon domain: { m } vertices, { kind }, n = { n }" )
    got, ( fi, fj, fc ) = native.measures_2d( pos, w, box, dens, facets = True, domain = v )
    ref, ( ri, rj, rc ) = oracle.measures( pos, w, lo, hi, density = odens, facets = True, domain = v )
    got = got.cpu().numpy()
    g, r = _rows( fi.cpu().numpy(), fj.cpu().numpy(), fc.cpu().numpy() ), _rows( ri, rj, rc )
    assert set( g ) == set( r )
    if kind == "constant":
        assert np.array_equal( got, ref )
        assert all( g[ k ] == r[ k ] for k in r )
        assert abs( got.sum() - H.polygon_area( v ) ) <= 1e-12
    else:
        assert np.abs( got - ref ).max() <= 1e-12 * ref.max()
        assert max( abs( g[ k ] - r[ k ] ) for k in r ) <= 1e-12 * max( abs( x ) for x in r.values() )
    # the FP32 kernel ( its vertices re-solved in double from the planes, the domain's edges among them )
    m32 = native.measures_2d( pos, w, box, dens, precision = native.FP32, domain = v ).cpu().numpy()
    assert np.abs( m32 - ref ).max() <= 1e-12 * max( ref.max(), 1e-300 )


def test_polygon_solve_matches_the_oracles( native, oracle ):
    """the FP64 solve by the limits in a hexagon of an image: the oracle's decisions ( iterations, diagrams, backtracks, limit
    rounds ) and weights"""
    import sdot
    from sdot import Iterative, Tuning
    from sdot.SdotPlanNd import card_options
    rng = np.random.default_rng( 70 )
    ( dirs, offs ), v = H.regular_polygon( 6 )
    lo, hi = v.min( axis = 0 ), v.max( axis = 0 )
    n = 5_000
    pos = H.points_in_polygon( n, dirs, offs, rng )
    img = H.bumps_image( 64, 64, rng, k = 4 )
    dens = sdot.Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 64 )
    odens = dict( values = np.ascontiguousarray( np.asarray( dens.values ).T ), x0 = 0.0, y0 = 0.0, hx = 1 / 64, hy = 1 / 64 )
    tol = 1e-9 * 1e6 / n / 1e2
    opts = card_options( Iterative( precision = "fp64", tol = tol, continuation = "never", tuning = Tuning( step = "limits", linear_solver = "cg",
                                                                                                           linear_tol = 1e-11 ) ), dens )
    say( f"case polygon solve: hexagon, image, n = { n }" )
    r = native.solve_2d( pos, np.full( n, 1 / n ), None, opts, [ lo[ 0 ], lo[ 1 ], hi[ 0 ], hi[ 1 ] ], dens, domain = v )
    st = _stats( r[ "stats" ] )
    ro = oracle.solve( pos, lo = lo, hi = hi, density = odens, tol = tol, step = "limits", lin_tol = 1e-11, domain = v )
    assert st[ "status" ] == 1 and ro[ "stats" ][ "status" ] == 1
    for key in ( "nb_iter", "nb_diag", "nb_backtracks", "nb_limit_rounds" ):
        assert st[ key ] == ro[ "stats" ][ key ], key
    assert abs( st[ "domain_mass" ] - ro[ "stats" ][ "domain_mass" ] ) <= 1e-12
    w, wo = r[ "weights" ].cpu().numpy(), ro[ "weights" ]
    assert np.abs( w - wo ).max() <= 1e-9 * ( wo.max() - wo.min() )
    # the mixed kernel: the plan passes the tolerance on the oracle's FP64 cells of its weights
    r2 = native.solve_2d( pos, np.full( n, 1 / n ), None, opts, [ lo[ 0 ], lo[ 1 ], hi[ 0 ], hi[ 1 ] ], dens, precision = native.MIXED, domain = v )
    assert _stats( r2[ "stats" ] )[ "status" ] == 1
    mo = oracle.measures( pos, r2[ "weights" ].cpu().numpy(), lo, hi, density = odens, domain = v )
    assert np.abs( mo - ro[ "stats" ][ "domain_mass" ] / n ).max() <= 1.01 * tol


def test_polytope_target_and_polygon_diagram_through_the_api( native, oracle ):
    import sdot
    rng = np.random.default_rng( 71 )
    ( dirs, offs ), v = H.regular_polygon( 8 )
    n = 3_000
    pos = H.points_in_polygon( n, dirs, offs, rng )
    plan = sdot.OtProblem( sdot.SumOfDiracs( pos ), sdot.Polytope( dirs, offs ) ).solve( sdot.Iterative( tol = 1e-12, continuation = "never" ) )
    assert plan.converged
    m = plan.cell_masses.cpu().numpy()
    assert abs( m.sum() - 1 ) <= 1e-12 and np.abs( m - 1 / n ).max() <= 1e-12
    # the oracle's cells of the weights: the same masses ( the constant density 1 / area on the octagon )
    mo = oracle.measures( pos, plan.weights.cpu().numpy(), v.min( axis = 0 ), v.max( axis = 0 ), density = 1 / H.polygon_area( v ), domain = v )
    assert np.abs( mo - 1 / n ).max() <= 1.01e-12
    # a diagram of half-planes that are not a box: the volumes of its cells, which fill the octagon
    pd = sdot.PowerDiagram( pos, np.zeros( n ), boundaries = ( dirs, offs ) )
    vol = pd.measures.cpu().numpy()
    assert abs( vol.sum() - H.polygon_area( v ) ) <= 1e-12
    assert np.array_equal( vol, oracle.measures( pos, np.zeros( n ), pd.box_min, pd.box_max, domain = pd.domain ) )


def test_polygon_domains_the_kernels_refuse( native ):
    import sdot
    rng = np.random.default_rng( 72 )
    ( dirs, offs ), v = H.regular_polygon( 6 )
    pos = H.points_in_polygon( 100, dirs, offs, rng )
    box = [ *v.min( axis = 0 ), *v.max( axis = 0 ) ]
    dens = sdot.UniformBox( v.min( axis = 0 ), v.max( axis = 0 ) )
    with pytest.raises( native.SdotError ):
        native.measures_2d( pos, None, box, dens, domain = v[ ::-1 ] )           # clockwise
    with pytest.raises( native.SdotError ):
        native.measures_2d( pos, None, box, dens, domain = v[ [ 0, 2, 1, 3, 4, 5 ] ] )   # not convex
    ( d13, o13 ), v13 = H.regular_polygon( 13 )
    with pytest.raises( NotImplementedError ):
        sdot.Polytope( d13, o13 )
