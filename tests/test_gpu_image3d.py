"""Voxel images in 3D, and images that cover only part of the domain box, on the MI355X through the C ABI.

A 3D image's masses and facets are held to the oracle's at 1e-12 relative ( the cells far out get 1e-12 of the mean mass ),
its moments to the closed form of the voxels' integrals, its solve to the oracle's plan. An image is zero outside its grid
( the reference's `Image` has the grid as its support ): the grids here overhang the box on one side and stop short of it on
another, in 2D and in 3D."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX3 = [ 0.0, 0.0, 0.0, 1.0, 1.0, 1.0 ]


def say( *a ):
    print( *a, flush = True )


def voxels( rng, shape = ( 12, 10, 8 ), origin = ( -0.1, 0.05, 0.0 ), h = ( 0.1, 0.1, 0.125 ) ):
    """`( sdot.Image, the oracle's dict )` of random voxels on a grid over x in [-0.1, 1.1], y in [0.05, 1.05], z in [0, 1]"""
    from sdot import Image
    v = rng.uniform( 0.2, 2.0, shape )
    img = Image( v, origin = list( origin ), frame = np.diag( h ) )
    return img, dict( voxels = v, x0 = origin[ 0 ], y0 = origin[ 1 ], z0 = origin[ 2 ], hx = h[ 0 ], hy = h[ 1 ], hz = h[ 2 ] )


def rows_of( fi, fj, fc ):
    return { ( int( i ), int( j ) ): float( c ) for i, j, c in zip( fi, fj, fc ) }


def seeds( rng, n ):
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    pos[ :, 1 ] = 0.05 + 0.95 * pos[ :, 1 ]               # ( none in the strip the grid leaves empty )
    return pos


@pytest.mark.parametrize( "n,seed", [ ( 3000, 0 ), ( 20000, 1 ) ] )
def test_3d_image_masses_and_facets_match_the_oracle( native, oracle, n, seed ):
    say( f"3d image measures: n={ n } seed={ seed }" )
    rng = np.random.default_rng( seed )
    pos = seeds( rng, n )
    w = rng.normal( 0, 1e-2 * n ** ( -2 / 3 ), n )
    img, ref_img = voxels( rng )
    m, ( fi, fj, fc ) = native.measures_3d( pos, w, BOX3, img, facets = True )
    m = m.cpu().numpy()
    mr, ( ri, rj, rc ) = oracle.measures_3d( pos, w, density = ref_img, facets = True )
    bar = 1e-12 * np.maximum( np.abs( mr ), mr.mean() )
    assert np.all( np.abs( m - mr ) <= bar ), ( np.abs( m - mr ) / bar ).max()
    got, ref = rows_of( fi.cpu().numpy(), fj.cpu().numpy(), fc.cpu().numpy() ), rows_of( ri, rj, rc )
    assert set( got ) == set( ref )
    big = max( abs( c ) for c in ref.values() )
    assert max( abs( got[ k ] - ref[ k ] ) / max( abs( ref[ k ] ), 1e-6 * big ) for k in ref ) <= 1e-12


def test_3d_image_moments_are_the_voxel_integrals( native ):
    """the masses, first and second moments summed over the cells are the voxels' integrals over the box's part of the grid
    ( closed forms per voxel block ), and the masses are those of the measures"""
    rng = np.random.default_rng( 2 )
    n = 5000
    pos = seeds( rng, n )
    img, d = voxels( rng )
    say( "3d image moments" )
    m, f, s = ( t.cpu().numpy() for t in native.moments( pos, None, BOX3, img ) )
    mm = native.measures_3d( pos, None, BOX3, img ).cpu().numpy()
    assert np.abs( m - mm ).max() <= 1e-13 * m.max()
    v = d[ "voxels" ]
    o = np.array( [ d[ "x0" ], d[ "y0" ], d[ "z0" ] ] )
    h = np.array( [ d[ "hx" ], d[ "hy" ], d[ "hz" ] ] )
    # per axis, the part of each slab inside [ 0, 1 ]: its length, and the integrals of x and x^2 over it
    L, X, X2 = [], [], []
    for a in range( 3 ):
        lo = np.clip( o[ a ] + h[ a ] * np.arange( v.shape[ a ] ), 0, 1 )
        hi = np.clip( o[ a ] + h[ a ] * ( np.arange( v.shape[ a ] ) + 1 ), 0, 1 )
        L.append( hi - lo ); X.append( ( hi ** 2 - lo ** 2 ) / 2 ); X2.append( ( hi ** 3 - lo ** 3 ) / 3 )
    mass = np.einsum( "ijk,i,j,k->", v, L[ 0 ], L[ 1 ], L[ 2 ] )
    first = [ np.einsum( "ijk,i,j,k->", v, *[ X[ b ] if b == a else L[ b ] for b in range( 3 ) ] ) for a in range( 3 ) ]
    second = sum( np.einsum( "ijk,i,j,k->", v, *[ X2[ b ] if b == a else L[ b ] for b in range( 3 ) ] ) for a in range( 3 ) )
    assert abs( m.sum() - mass ) <= 1e-12 * mass
    assert np.allclose( f.sum( 0 ), first, rtol = 1e-12, atol = 0 )
    assert abs( s.sum() - second ) <= 1e-12 * second


def _opts( tol, linear ):
    from sdot import Iterative, Tuning
    from sdot.SdotPlanNd import card_options
    return card_options( Iterative( tol = tol, continuation = "never",
                                    tuning = Tuning( step = "trials", linear_solver = linear, linear_tol = 1e-12 if linear == "cg" else 1e-8 ) ), None )


@pytest.mark.parametrize( "linear", [ "cg", "mg" ] )
def test_3d_image_solve_matches_the_oracle( native, oracle, linear ):
    rng = np.random.default_rng( 3 )
    n = 3000
    pos = seeds( rng, n )
    img, d = voxels( rng )
    tol = 1e-9 * 1e6 / n / 1e2
    nu = np.full( n, 1.0 / n )
    say( f"3d image solve: { linear }" )
    r = native.solve_3d( pos, nu, None, _opts( tol, linear ), BOX3, img )
    st = r[ "stats" ]
    assert st[ 0 ] == 1, st[ :6 ]
    ref = oracle.solve_3d( pos, density = d, tol = tol, lin_tol = 1e-12, linear = linear )
    assert ref[ "stats" ][ "status" ] == 1
    target = nu * st[ 12 ]
    masses = r[ "cell_masses" ].cpu().numpy()
    assert np.abs( masses - target ).max() <= tol
    w = r[ "weights" ].cpu().numpy()
    assert np.abs( oracle.measures_3d( pos, w, density = d ) - target ).max() <= 1.01 * tol
    spread = ref[ "weights" ].max() - ref[ "weights" ].min()
    assert np.abs( w - ref[ "weights" ] ).max() <= 1e-5 * spread
    assert np.abs( r[ "barycenters" ].cpu().numpy() - ref[ "barycenters" ] ).max() < 1e-6
    assert abs( r[ "cost" ] - ref[ "cost" ] ) < 1e-6 * ref[ "cost" ]


def test_3d_image_through_the_reference_api( native ):
    """`OtProblem( SumOfDiracs, Image 3D ).solve( Iterative )`, then `PowerDiagram` of its weights"""
    import sdot
    rng = np.random.default_rng( 4 )
    n = 8000
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    v = rng.uniform( 0.5, 1.5, ( 16, 16, 16 ) )
    target = sdot.Image( v, origin = [ 0, 0, 0 ], frame = np.eye( 3 ) / 16 )
    say( "3d image through the api" )
    sol = sdot.OtProblem( sdot.SumOfDiracs( pos ), target ).solve( sdot.Iterative( tol = 1e-9 * 1e6 / n ) )
    assert sol.converged and sol.barycenters.shape == ( n, 3 )
    m = sdot.PowerDiagram( pos, sol.weights, distribution = target ).measures.cpu().numpy()
    assert np.abs( m - sol.target_masses ).max() <= 1.01e-9 * 1e6 / n
    assert abs( m.sum() - 1 ) < 1e-12


def _img2( values_xy, lo, hi ):
    nx, ny = values_xy.shape
    return dict( values = np.ascontiguousarray( values_xy.T ), x0 = lo[ 0 ], y0 = lo[ 1 ], hx = ( hi[ 0 ] - lo[ 0 ] ) / nx, hy = ( hi[ 1 ] - lo[ 1 ] ) / ny )


def test_2d_image_is_zero_outside_its_grid( native, oracle ):
    """a grid over [ -0.2, 0.7 ] x [ 0.1, 0.8 ] in the unit box, through the C ABI: the masses and facets are the oracle's ( which
    clips the cells by the pixels ), and a solve by the limits takes the oracle's steps to its plan"""
    from sdot import Image, Iterative, Tuning
    from sdot.SdotPlanNd import card_options
    rng = np.random.default_rng( 5 )
    n = 6000
    lo, hi = ( -0.2, 0.1 ), ( 0.7, 0.8 )
    v = rng.uniform( 0.5, 2.0, ( 90, 70 ) )
    img = Image( v, origin = list( lo ), frame = np.diag( [ 0.01, 0.01 ] ) ).normalized_version()
    d = _img2( img.values, lo, hi )
    pos = np.column_stack( [ rng.uniform( 0, 0.7, n ), rng.uniform( 0.1, 0.8, n ) ] )
    w = rng.normal( 0, 1e-6, n )
    say( "2d image zero outside its grid: measures" )
    m, ( fi, fj, fc ) = native.measures_2d( pos, w, [ 0, 0, 1, 1 ], img, facets = True )
    m = m.cpu().numpy()
    mr, ( ri, rj, rc ) = oracle.measures( pos, w, density = d, facets = True )
    assert np.abs( m - mr ).max() <= 1e-12 * mr.max()
    inside = img.values[ 20:, : ].sum() * 1e-4          # ( the columns from x = 0 on: the grid's mass in the box )
    assert abs( m.sum() - inside ) < 1e-12
    got, ref = rows_of( fi.cpu().numpy(), fj.cpu().numpy(), fc.cpu().numpy() ), rows_of( ri, rj, rc )
    assert set( got ) == set( ref )
    big = max( abs( c ) for c in ref.values() )
    assert max( abs( got[ k ] - ref[ k ] ) / max( abs( ref[ k ] ), 1e-6 * big ) for k in ref ) <= 1e-11
    say( "2d image zero outside its grid: solve by the limits" )
    tol = 1e-13                                          # ( masses of ~1.7e-4: the plan fixed to ~1e-9 of itself )
    opts = card_options( Iterative( precision = "fp64", tol = tol, continuation = "never", tuning = Tuning( step = "limits", linear_solver = "cg", linear_tol = 1e-10 ) ), img )
    r = native.solve_2d( pos, np.full( n, 1.0 / n ), None, opts, [ 0, 0, 1, 1 ], img )
    ro = oracle.solve( pos, tol = tol, lin_tol = 1e-10, density = d, step = "limits" )
    say( "  iterations", r[ "stats" ][ 3 ], ro[ "stats" ][ "nb_iter" ], "diagrams", r[ "stats" ][ 4 ], ro[ "stats" ][ "nb_diag" ] )
    assert r[ "stats" ][ 0 ] == 1 and ro[ "stats" ][ "status" ] == 1
    assert r[ "stats" ][ 3 ] == ro[ "stats" ][ "nb_iter" ]
    wg = r[ "weights" ].cpu().numpy()
    assert np.abs( wg - ro[ "weights" ] ).max() < 1e-9 * np.abs( ro[ "weights" ] ).max()


def test_served_3d_image_requests_are_single_solves( native, oracle ):
    """3D voxel-image requests mixed with 2D ones on a server: each plan is the one it gets alone, to the bit, and a sampled
    3D plan's masses are the oracle's at its weights"""
    import torch
    import sdot
    rng = np.random.default_rng( 6 )
    probs = []
    for q in range( 6 ):
        if q % 2:
            v = rng.uniform( 0.3, 1.5, ( 24, 24, 24 ) )
            pos = torch.as_tensor( rng.uniform( 0, 1, ( int( rng.integers( 2000, 9000 ) ), 3 ) ), device = "cuda" )
            probs.append( sdot.OtProblem( sdot.SumOfDiracs( pos ), sdot.Image( v, origin = [ 0, 0, 0 ], frame = np.eye( 3 ) / 24 ) ) )
        else:
            v = rng.uniform( 0.3, 1.5, ( 64, 64 ) )
            pos = torch.as_tensor( rng.uniform( 0, 1, ( int( rng.integers( 2000, 20000 ) ), 2 ) ), device = "cuda" )
            probs.append( sdot.OtProblem( sdot.SumOfDiracs( pos ), sdot.Image( v, origin = [ 0, 0 ], frame = np.eye( 2 ) / 64 ) ) )
    it = sdot.Iterative( precision = "fp64", tol = 1e-10, continuation = "never" )
    say( "served 3d image requests" )
    with sdot.Server( slots = 3 ) as srv:
        plans = [ t.result() for t in [ srv.submit( pb, it ) for pb in probs ] ]
    for pb, pl in zip( probs, plans ):
        alone = pb.solve( it )
        assert pl.converged and alone.converged
        assert torch.equal( pl.weights, alone.weights ) and torch.equal( pl.cell_masses, alone.cell_masses ) and pl.cost == alone.cost
    pb, pl = probs[ 1 ], plans[ 1 ]
    v = pb.target.values
    d = dict( voxels = v / ( v.sum() / 24 ** 3 ), x0 = 0.0, y0 = 0.0, z0 = 0.0, hx = 1 / 24, hy = 1 / 24, hz = 1 / 24 )
    mo = oracle.measures_3d( pb.source.positions.cpu().numpy(), pl.weights.cpu().numpy(), density = d )
    assert np.abs( mo - pl.cell_masses.cpu().numpy() ).max() <= 1e-12 * mo.max()
