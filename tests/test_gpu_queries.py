"""The diagram's other queries on the GPU, through the C ABI: the cells themselves ( the oracle's polygons to the bit ), the
moments ( polygon formulas on the oracle's cells; pixel clipping for an image; the box's closed forms in 3D ), the Hessian rows
( finite differences of the masses ), the adjoint of the masses ( finite differences in the weights and the positions; torch's
autograd through it ), the Laplacian solve ( its residual ), and the quantization caller ( Lloyd: the cost decreases )."""

import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def say( *a ):
    print( *a, flush = True )
    sys.stdout.flush()


def _polygon_moments( xy ):
    """`( area, int x, int |x|^2 )` of a counter-clockwise polygon, by its fan of triangles from the origin"""
    a = xy
    b = np.roll( xy, -1, axis = 0 )
    cr = a[ :, 0 ] * b[ :, 1 ] - b[ :, 0 ] * a[ :, 1 ]
    area = cr.sum() / 2
    first = np.array( [ ( cr * ( a[ :, 0 ] + b[ :, 0 ] ) ).sum() / 6, ( cr * ( a[ :, 1 ] + b[ :, 1 ] ) ).sum() / 6 ] )
    sq = lambda u, v: u * u + u * v + v * v
    second = ( cr * ( sq( a[ :, 0 ], b[ :, 0 ] ) + sq( a[ :, 1 ], b[ :, 1 ] ) ) ).sum() / 12
    return area, first, second


def _image_polygon_moments( xy, img, x0, y0, hx, hy ):
    """the same against an image `img[ j, i ]`, by exact clipping of the polygon against every pixel it meets"""
    ny, nx = img.shape
    i0 = max( int( np.floor( ( xy[ :, 0 ].min() - x0 ) / hx ) ), 0 )
    i1 = min( int( np.floor( ( xy[ :, 0 ].max() - x0 ) / hx ) ), nx - 1 )
    j0 = max( int( np.floor( ( xy[ :, 1 ].min() - y0 ) / hy ) ), 0 )
    j1 = min( int( np.floor( ( xy[ :, 1 ].max() - y0 ) / hy ) ), ny - 1 )
    m, f, s = 0.0, np.zeros( 2 ), 0.0
    for j in range( j0, j1 + 1 ):
        for i in range( i0, i1 + 1 ):
            p = H.clip_polygon( xy, 1, 0, x0 + ( i + 1 ) * hx )
            if len( p ) >= 3: p = H.clip_polygon( p, -1, 0, -( x0 + i * hx ) )
            if len( p ) >= 3: p = H.clip_polygon( p, 0, 1, y0 + ( j + 1 ) * hy )
            if len( p ) >= 3: p = H.clip_polygon( p, 0, -1, -( y0 + j * hy ) )
            if len( p ) >= 3:
                a, fi, si = _polygon_moments( p )
                m += img[ j, i ] * a
                f += img[ j, i ] * fi
                s += img[ j, i ] * si
    return m, f, s


def _cloud( n, seed ):
    """uniform seeds, weights of a hundredth of a cell ( a uniform cloud's close pairs may still leave a cell empty )"""
    rng = np.random.default_rng( seed )
    return rng.uniform( 0, 1, ( n, 2 ) ), rng.uniform( -0.01, 0.01, n ) / n


def test_cells_are_the_oracles_polygons( native, oracle ):
    from sdot import PowerDiagram, UniformBox
    pos, w = _cloud( 3000, 61 )
    say( "case cells: 3000 seeds" )
    off, xy, cut = ( t.cpu().numpy() for t in PowerDiagram( pos, w, distribution = UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).cells )
    nv = np.diff( off )
    assert off[ 0 ] == 0 and np.all( ( nv == 0 ) | ( nv >= 3 ) )
    for i in np.random.default_rng( 1 ).choice( 3000, 60, replace = False ):
        oxy, oid = oracle.cell( pos, int( i ), w )
        assert np.array_equal( xy[ off[ i ]: off[ i + 1 ] ], oxy ), i
        assert np.array_equal( cut[ off[ i ]: off[ i + 1 ] ], oid ), i
    # the cells tile the box
    area = sum( H.polygon_area( xy[ off[ i ]: off[ i + 1 ] ] ) for i in range( 3000 ) )
    assert abs( area - 1 ) < 1e-12


def test_moments_constant_density_are_the_polygons( native, oracle ):
    from sdot import PowerDiagram, UniformBox
    pos, w = _cloud( 2000, 62 )
    say( "case moments: constant" )
    m, f, s = ( t.cpu().numpy() for t in PowerDiagram( pos, w, distribution = UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).moments )
    for i in range( 0, 2000, 37 ):
        oxy, _ = oracle.cell( pos, i, w )
        if len( oxy ) < 3:
            assert m[ i ] == 0
            continue
        a, fi, si = _polygon_moments( oxy )
        assert abs( m[ i ] - a ) <= 1e-14
        assert np.abs( f[ i ] - fi ).max() <= 1e-14
        assert abs( s[ i ] - si ) <= 1e-14
    # the whole box: its mass, first moment and second moment in closed form
    assert abs( m.sum() - 1 ) < 1e-12 and np.abs( f.sum( 0 ) - 0.5 ).max() < 1e-12 and abs( s.sum() - 2 / 3 ) < 1e-12


def test_moments_image_density_are_the_pixel_pieces( native, oracle ):
    from sdot import Image, PowerDiagram
    pos, w = _cloud( 400, 63 )
    img = H.bumps_image( 64, 64, np.random.default_rng( 3 ) )
    dens = Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 64 ).normalized_version()
    v = np.asarray( dens.values ).T                      # [ j, i ]
    say( "case moments: image 64^2" )
    m, f, s = ( t.cpu().numpy() for t in PowerDiagram( pos, w, distribution = dens ).moments )
    for i in range( 0, 400, 23 ):
        oxy, _ = oracle.cell( pos, i, w )
        if len( oxy ) < 3:
            assert m[ i ] == 0 and np.all( f[ i ] == 0 ) and s[ i ] == 0
            continue
        mi, fi, si = _image_polygon_moments( oxy, v, 0.0, 0.0, 1 / 64, 1 / 64 )
        assert abs( m[ i ] - mi ) <= 1e-12 * max( mi, 1e-3 )
        assert np.abs( f[ i ] - fi ).max() <= 1e-12 * max( mi, 1e-3 )
        assert abs( s[ i ] - si ) <= 1e-12 * max( mi, 1e-3 )


def test_moments_3d_sum_to_the_box( native ):
    from sdot import PowerDiagram, UniformBox
    rng = np.random.default_rng( 64 )
    pos = rng.uniform( 0, 1, ( 3000, 3 ) )
    say( "case moments: 3D constant" )
    pd = PowerDiagram( pos, None, distribution = UniformBox( [ 0, 0, 0 ], [ 1, 1, 1 ] ) )
    m, f, s = ( t.cpu().numpy() for t in pd.moments )
    assert np.array_equal( m, pd.measures.cpu().numpy() )
    assert abs( m.sum() - 1 ) < 1e-12 and np.abs( f.sum( 0 ) - 0.5 ).max() < 1e-12 and abs( s.sum() - 1.0 ) < 1e-12
    assert np.all( m > 0 ) and np.all( np.abs( f / m[ :, None ] - 0.5 ) <= 0.5 )


def test_hessian_rows_are_the_jacobian_of_the_masses( native ):
    import torch
    from sdot import PowerDiagram, UniformBox
    pos, w = _cloud( 300, 65 )
    say( "case hessian rows" )
    box = UniformBox( [ 0, 0 ], [ 1, 1 ] )
    counts, ids, vals = PowerDiagram( pos, w, distribution = box ).hessian_rows()
    assert np.all( ids[ np.arange( ids.shape[ 1 ] )[ None, : ] < counts[ :, None ] ] >= 0 )
    for j in np.nonzero( counts > 0 )[ 0 ][ [ 0, 17, 123 ] ]:
        h = 1e-7 / 300
        wp, wm = w.copy(), w.copy()
        wp[ j ] += h
        wm[ j ] -= h
        dm = ( PowerDiagram( pos, wp, distribution = box ).measures - PowerDiagram( pos, wm, distribution = box ).measures ).cpu().numpy() / ( 2 * h )
        jac = np.zeros( 300 )
        jac[ j ] = vals[ j, :counts[ j ] ].sum()
        for r in range( counts[ j ] ):
            jac[ ids[ j, r ] ] -= vals[ j, r ]
        assert np.abs( dm - jac ).max() <= 1e-6 * np.abs( jac ).max()


@pytest.mark.parametrize( "kind", [ "constant", "image" ] )
def test_measures_vjp_is_the_derivative_of_the_masses( native, kind ):
    import torch
    from sdot import Image, PowerDiagram, UniformBox
    pos, w = _cloud( 500, 66 )
    if kind == "image":
        dens = Image( H.bumps_image( 64, 64, np.random.default_rng( 4 ) ), origin = [ 0, 0 ], frame = np.eye( 2 ) / 64 ).normalized_version()
    else:
        dens = UniformBox( [ 0, 0 ], [ 1, 1 ] )
    rng = np.random.default_rng( 5 )
    g = rng.normal( size = 500 )
    say( f"case vjp: { kind }" )
    gw, gp = ( t.cpu().numpy() for t in native.measures_vjp_2d( pos, w, [ 0, 0, 1, 1 ], dens, g ) )
    # directional derivatives by central differences, in the weights and in the positions
    dw, dp = rng.normal( size = 500 ), rng.normal( size = ( 500, 2 ) )
    hw, hp = 1e-6 / 500, 1e-7

    def mass( p, ww ):
        return PowerDiagram( p, ww, distribution = dens ).measures.cpu().numpy()

    fw = g @ ( mass( pos, w + hw * dw ) - mass( pos, w - hw * dw ) ) / ( 2 * hw )
    fp = g @ ( mass( pos + hp * dp, w ) - mass( pos - hp * dp, w ) ) / ( 2 * hp )
    assert abs( fw - gw @ dw ) <= 1e-6 * np.abs( gw ).sum()
    assert abs( fp - ( gp * dp ).sum() ) <= 1e-5 * np.abs( gp ).sum()
    # the same through torch's autograd
    pt = torch.tensor( pos, device = "cuda", requires_grad = True )
    wt = torch.tensor( w, device = "cuda", requires_grad = True )
    m = PowerDiagram( pt, wt, distribution = dens ).measures
    ( m * torch.as_tensor( g, device = "cuda" ) ).sum().backward()
    assert np.allclose( wt.grad.cpu().numpy(), gw, rtol = 1e-12, atol = 1e-14 * np.abs( gw ).max() )
    assert np.allclose( pt.grad.cpu().numpy(), gp, rtol = 1e-12, atol = 1e-14 * np.abs( gp ).max() )


def test_laplacian_solve( native ):
    import torch
    from sdot import PowerDiagram, UniformBox
    pos, w = _cloud( 800, 67 )
    say( "case laplacian solve" )
    pd = PowerDiagram( pos, w, distribution = UniformBox( [ 0, 0 ], [ 1, 1 ] ) )
    assert float( pd.measures.min() ) > 0                  # ( a connected Laplacian: no empty cell )
    rhs = np.random.default_rng( 6 ).normal( size = 800 )
    rhs -= rhs.mean()
    x = pd.laplacian_solve( rhs ).cpu().numpy()
    counts, ids, vals = pd.hessian_rows()
    L = np.zeros( ( 800, 800 ) )
    for i in range( 800 ):
        for r in range( counts[ i ] ):
            L[ i, ids[ i, r ] ] -= vals[ i, r ]
            L[ i, i ] += vals[ i, r ]
    assert x[ 0 ] == 0
    assert np.linalg.norm( L @ x - rhs ) <= 1e-9 * np.linalg.norm( rhs )


def test_quantization_decreases_the_cost( native ):
    import warnings
    import sdot
    img = H.bumps_image( 64, 64, np.random.default_rng( 8 ), k = 3 )
    target = sdot.Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 64 )
    costs = []
    say( "case quantization: 600 Diracs, 64^2 image" )
    with warnings.catch_warnings():
        warnings.simplefilter( "ignore" )
        pos = sdot.quantization( target, 600, seed = 3, tol = 2e-2, max_iter = 60, callback = lambda it, x, plan: costs.append( float( plan.cost ) ) )
    assert pos.shape == ( 600, 2 )
    assert all( b <= a * ( 1 + 1e-9 ) for a, b in zip( costs, costs[ 1: ] ) )
    assert costs[ -1 ] < 0.9 * costs[ 0 ]
