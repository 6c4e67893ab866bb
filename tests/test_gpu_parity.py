"""The HIP library against the oracle, through the C ABI, on the GPU.

* constant density: the masses and facets are the oracle's TO THE BIT ( same tree, same walk, same
  cut arithmetic );
* image density: within 1e-12 relative ( Green's theorem on the edges against the oracle's pixel
  pieces: two exact quadratures, different roundings );
* the solve: the same iterations and diagrams as the oracle, weights to 1e-10 of their scale, and at
  full size the properties that do not need the oracle.
"""

import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def say( *a ):
    print( *a, flush = True )
    sys.stdout.flush()


def _img( values_xy, lo = ( 0, 0 ), hi = ( 1, 1 ) ):
    """the oracle's image args from a `[ nx, ny ]` array ( first index along x )"""
    nx, ny = values_xy.shape
    return dict( values = np.ascontiguousarray( values_xy.T ), x0 = lo[ 0 ], y0 = lo[ 1 ], hx = ( hi[ 0 ] - lo[ 0 ] ) / nx, hy = ( hi[ 1 ] - lo[ 1 ] ) / ny )


def _gpu_measures( native, pos, w, density = None, facets = False ):
    from sdot.distributions import Image, UniformBox
    dens = UniformBox( [ 0, 0 ], [ 1, 1 ] ) if density is None else density
    res = native.measures_2d( pos, w, [ 0, 0, 1, 1 ], dens, facets = facets )
    if facets:
        m, ( fi, fj, fc ) = res
        return m.cpu().numpy(), ( fi.cpu().numpy(), fj.cpu().numpy(), fc.cpu().numpy() )
    return res.cpu().numpy()


CLOUDS = [ "uniform", "lines", "lines weighted", "ring", "grid", "tiny" ]


def _cloud( label, rng ):
    if label == "ring":
        return H.ring( 60, rng, 1500 ), None
    if label == "grid":
        g = ( np.arange( 40 ) + 0.5 ) / 40
        return np.array( [ [ x, y ] for x in g for y in g ] ), None
    if label == "tiny":
        return np.array( [ [ 0.2, 0.3 ], [ 0.7, 0.6 ] ] ), np.array( [ 0.01, -0.02 ] )
    for lab, pos, w in H.clouds( rng, 5000 ):
        if lab == label:
            return pos, w


@pytest.mark.parametrize( "label", CLOUDS )
def test_constant_density_masses_and_facets_are_the_oracles( native, oracle, label ):
    rng = np.random.default_rng( 100 + CLOUDS.index( label ) )
    pos, w = _cloud( label, rng )
    say( "case constant", label, len( pos ) )
    m, ( fi, fj, fc ) = _gpu_measures( native, pos, w, facets = True )
    mo, ( oi, oj, oc ) = oracle.measures( pos, w, facets = True )
    if label == "grid":
        # seeds tied at the medians: the two builds may send a tied seed to either side ( the reference's own card and
        # CPU trees differ there too, `AaBsp.py:227-229` ), so the cut order and the rounding differ; the cells do not
        assert np.abs( m - mo ).max() <= 1e-13 * mo.max() and np.allclose( m, 1 / 1600, rtol = 1e-12 )
        big = lambda i, j, c: { ( a, b ) for a, b, x in zip( i.tolist(), j.tolist(), c.tolist() ) if x > 1e-9 }
        assert big( fi, fj, fc ) == big( oi, oj, oc )
        return
    assert np.array_equal( m, mo ), ( label, np.abs( m - mo ).max() )
    # facets: the same rows ( user order ), each in its cell's edge order
    assert np.array_equal( fi, oi ) and np.array_equal( fj, oj ) and np.array_equal( fc, oc )


def test_event_counts_are_the_oracles( native, oracle ):
    rng = np.random.default_rng( 7 )
    pos = rng.uniform( 0, 1, ( 20000, 2 ) )
    say( "case events 20000" )
    _, ev = native.measures_2d( pos, None, [ 0, 0, 1, 1 ], __import__( "sdot" ).UniformBox( [ 0, 0 ], [ 1, 1 ] ), events = True )
    _, oev = oracle.measures( pos, events = True )
    assert ev == oev


@pytest.mark.parametrize( "res", [ 64, 1024 ] )
def test_image_masses_match_the_oracle( native, oracle, res ):
    from sdot import Image
    rng = np.random.default_rng( 200 + res )
    n = 4000 if res == 64 else 60000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    w = rng.uniform( -0.5, 0.5, n ) / n
    img = H.bumps_image( res, res, rng )
    dens = Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / res ).normalized_version()
    say( "case image", res, n )
    m = _gpu_measures( native, pos, w, dens )
    mo = oracle.measures( pos, w, density = _img( dens.values ) )
    rel = np.abs( m - mo ) / np.maximum( mo, 1e-3 / n )
    assert rel.max() < 1e-12, rel.max()
    assert abs( m.sum() - 1 ) < 1e-12


def test_image_facets_match_the_oracle( native, oracle ):
    from sdot import Image
    rng = np.random.default_rng( 300 )
    n = 3000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    img = H.bumps_image( 48, 48, rng )
    dens = Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 48 ).normalized_version()
    say( "case image facets" )
    m, ( fi, fj, fc ) = _gpu_measures( native, pos, None, dens, facets = True )
    mo, ( oi, oj, oc ) = oracle.measures( pos, None, density = _img( dens.values ), facets = True )
    assert np.array_equal( fi, oi ) and np.array_equal( fj, oj )
    assert np.abs( fc - oc ).max() <= 1e-12 * np.abs( oc ).max()


@pytest.mark.parametrize( "label", [ "uniform 300", "uniform", "lines", "lines weighted" ] )
def test_neighbour_sets_equal_the_lifted_hull( native, label ):
    """identical neighbour sets on generic inputs: the pairs whose power facet meets the box ( scipy's lifted hull ), no more
    and no fewer, up to facets shorter than 1e-10 that rounding may keep or drop"""
    rng = np.random.default_rng( 5 )
    if label == "uniform 300":
        pos = rng.uniform( 0.05, 0.95, ( 300, 2 ) )
        w = rng.uniform( -2e-3, 2e-3, 300 )
    else:
        pos, w = next( ( p, ww ) for lab, p, ww in H.clouds( rng, 3000 ) if lab == label )
    say( f"case neighbours: { label }" )
    _, ( fi, fj, _ ) = _gpu_measures( native, pos, w, facets = True )
    ours = { ( min( i, j ), max( i, j ) ) for i, j in zip( fi.tolist(), fj.tolist() ) }
    hull, amb = H.lifted_neighbours( pos, w, ambiguous = True )
    assert hull <= ours <= hull | amb, ( len( hull - ours ), len( ours - hull - amb ) )


def test_overflowed_cells_go_through_the_second_pass( native, oracle ):
    # a 300-gon: far past the first pass's LDS slots, redone with global slots
    rng = np.random.default_rng( 9 )
    pos = H.ring( 300, rng, 800 )
    say( "case overflow ring 300" )
    m = _gpu_measures( native, pos, None )
    mo = oracle.measures( pos )
    assert np.array_equal( m, mo )
    xy, _ = oracle.cell( pos, 0 )
    assert len( xy ) == 300


def test_a_cell_past_the_vertex_limit_is_a_clean_failure( native ):
    from sdot import KernelFailure
    # the centre of a ring of 2500 seeds: a 2500-gon, past the 2048-vertex limit of the last pass
    rng = np.random.default_rng( 10 )
    pos = H.ring( 2500, rng, 10 )
    say( "case failure ring 2500" )
    with pytest.raises( KernelFailure ) as e:
        _gpu_measures( native, pos, None )
    assert e.value.code == 1 and e.value.value == 0


def test_tree_through_the_abi_is_the_oracles( native, oracle ):
    rng = np.random.default_rng( 11 )
    pos = rng.uniform( 0, 1, ( 5000, 2 ) )
    say( "case tree" )
    t = { k: v.cpu().numpy() for k, v in native.build_tree_2d( pos ).items() }
    to = oracle.build_tree( pos )
    assert np.array_equal( t[ "seed_indices" ], to[ "order" ] )
    assert np.array_equal( t[ "node_begin" ], to[ "beg" ] ) and np.array_equal( t[ "node_end" ], to[ "end" ] )
    assert np.array_equal( t[ "node_box" ], to[ "box" ] )
    assert np.array_equal( t[ "rank_of" ][ t[ "seed_indices" ] ], np.arange( 5000 ) )


def test_majorants_through_the_abi_majorate( native ):
    rng = np.random.default_rng( 12 )
    pos = rng.uniform( 0, 1, ( 5000, 2 ) )
    t = native.build_tree_2d( pos )
    order = t[ "seed_indices" ].cpu().numpy()
    beg, end = t[ "node_begin" ].cpu().numpy(), t[ "node_end" ].cpu().numpy()
    sp = pos[ order ]
    say( "case majorants" )
    for w in ( 0.3 * pos[ :, 0 ] - 0.7 * pos[ :, 1 ], rng.uniform( -1, 1, 5000 ), 1e-2 * np.sin( 5 * pos[ :, 0 ] ) ):
        sw = w[ order ]
        wa, wb = native.refresh_majorants_2d( t, sp, sw )
        wa, wb = wa.cpu().numpy(), wb.cpu().numpy()
        for m in range( len( beg ) ):
            if end[ m ] > beg[ m ]:
                s = slice( beg[ m ], end[ m ] )
                assert np.all( sw[ s ] <= sp[ s ] @ wa[ m ] + wb[ m ] ), m
        # an affine weight is majorized by its own plane at the root, tightly
        if w is not None and np.allclose( w, 0.3 * pos[ :, 0 ] - 0.7 * pos[ :, 1 ] ):
            assert np.allclose( wa[ 0 ], [ 0.3, -0.7 ], atol = 1e-9 ) and abs( wb[ 0 ] ) < 1e-5


def test_linear_solves_match_a_dense_solve( native, oracle ):
    rng = np.random.default_rng( 13 )
    n = 500
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    _, ( fi, fj, fc ) = oracle.measures( pos, facets = True )
    row = np.zeros( n + 1, np.int64 )
    np.add.at( row, fi + 1, 1 )
    row = np.cumsum( row )
    dia = np.zeros( n )
    np.add.at( dia, fi, fc )
    L = -np.zeros( ( n, n ) )
    for i, j, c in zip( fi, fj, fc ):
        L[ i, j ] -= c
    L[ np.arange( n ), np.arange( n ) ] = dia
    b = rng.normal( size = ( 2, n ) )
    b -= b.mean( 1, keepdims = True )
    for method in ( "cg", "mg" ):
        say( "case linear", method )
        sol, its = native.linear_solves_2d( row, fj, fc, dia, b, tol = 1e-12, method = method )
        sol = sol.cpu().numpy()
        for k in range( 2 ):
            ref = np.zeros( n )
            ref[ 1: ] = np.linalg.solve( L[ 1:, 1: ], b[ k, 1: ] )
            assert its[ k ] > 0 and sol[ k, 0 ] == 0, method
            assert np.abs( sol[ k ] - ref ).max() < 1e-9 * np.abs( ref ).max(), method


def test_multigrid_solves_a_large_laplacian( native, oracle ):
    # the K-cycle multigrid on the Laplacian of a weighted 100k diagram: the relative residual it claims, the
    # Jacobi-PCG's solution, and far fewer iterations ( reference test_SdotPlanNd.py:680-727 checks the residual )
    import scipy.sparse as sps
    rng = np.random.default_rng( 14 )
    n = 100000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    pos = pos[ oracle.build_tree( pos )[ "order" ] ]      # ( in tree order: the aggregates of consecutive rows are compact )
    w = rng.normal( 0, 1e-10, n )
    _, ( fi, fj, fc ) = oracle.measures( pos, w, facets = True )
    row = np.zeros( n + 1, np.int64 )
    np.add.at( row, fi + 1, 1 )
    row = np.cumsum( row )
    dia = np.zeros( n )
    np.add.at( dia, fi, fc )
    assert np.all( dia > 0 )                             # ( no empty cell: the graph is connected, `b` of zero sum is in the range )
    L = sps.csr_matrix( ( -fc, ( fi, fj ) ), shape = ( n, n ) ) + sps.diags( dia )
    b = rng.normal( size = ( 1, n ) )
    b -= b.mean()
    say( "case multigrid 100k" )
    xm, im = native.linear_solves_2d( row, fj, fc, dia, b, tol = 1e-10, method = "mg" )
    xc, ic = native.linear_solves_2d( row, fj, fc, dia, b, tol = 1e-10, method = "cg" )
    xm, xc = xm.cpu().numpy()[ 0 ], xc.cpu().numpy()[ 0 ]
    # the residual in the gauge: row 0 is replaced by x_0 = 0, so compare on the other rows
    res = ( L @ xm - b[ 0 ] )[ 1: ]
    assert np.linalg.norm( res ) <= 3e-10 * np.linalg.norm( b[ 0 ] )
    assert np.abs( xm - xc ).max() <= 1e-7 * np.abs( xc ).max()
    assert 0 < im[ 0 ] < ic[ 0 ] / 10


def _solve_both( oracle, pos, density_gpu, density_oracle, tol, lin_tol, step = "trials", continuation = "never", maxit = 100 ):
    from sdot import Iterative, OtProblem, SumOfDiracs, Tuning
    sol = OtProblem( SumOfDiracs( pos ), density_gpu ).solve( Iterative( precision = "fp64", tol = tol, continuation = continuation, max_iter = maxit,
                                                                          tuning = Tuning( step = step, linear_solver = "cg", linear_tol = lin_tol ) ) )
    ro = oracle.solve( pos, tol = tol, lin_tol = lin_tol, density = density_oracle, step = step, continuation = continuation, maxit = maxit )
    return sol, ro


def test_solve_uniform_10k_matches_the_oracle( oracle, native ):
    # the plain method at 10 000 Diracs: same decisions, same iterations, same plan
    from sdot import UniformBox
    pos = np.random.default_rng( 0 ).uniform( 0, 1, ( 10000, 2 ) )
    say( "case solve uniform 10k" )
    sol, ro = _solve_both( oracle, pos, UniformBox( [ 0, 0 ], [ 1, 1 ] ), None, 1e-9, 1e-10 )
    assert sol.converged and ro[ "stats" ][ "status" ] == 1
    assert sol.stats[ "nb_iter" ] == ro[ "stats" ][ "nb_iter" ] and sol.stats[ "nb_diag" ] == ro[ "stats" ][ "nb_diag" ]
    w = sol.weights.cpu().numpy()
    assert np.abs( w - ro[ "weights" ] ).max() < 1e-10 * np.abs( ro[ "weights" ] ).max()
    m = sol.cell_masses.cpu().numpy()
    assert np.abs( m - 1e-4 ).max() <= 1e-9
    assert abs( sol.cost - ro[ "cost" ] ) < 1e-12 * ro[ "cost" ]
    assert np.abs( sol.barycenters.cpu().numpy() - ro[ "barycenters" ] ).max() < 1e-10
    # the cells of the returned weights, measured again, have the returned masses
    assert np.array_equal( _gpu_measures( native, pos, w ), m )


def test_solve_image_matches_the_oracle( oracle, native ):
    from sdot import Image
    rng = np.random.default_rng( 1 )
    n = 20000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    img = H.bumps_image( 128, 128, rng )
    dens = Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 128 ).normalized_version()
    say( "case solve image 20k" )
    sol, ro = _solve_both( oracle, pos, dens, _img( dens.values ), 1e-9 / 10, 1e-10 )
    assert sol.converged and ro[ "stats" ][ "status" ] == 1
    assert sol.stats[ "nb_iter" ] == ro[ "stats" ][ "nb_iter" ]
    w = sol.weights.cpu().numpy()
    assert np.abs( w - ro[ "weights" ] ).max() < 1e-9 * np.abs( ro[ "weights" ] ).max()
    assert abs( sol.cost - ro[ "cost" ] ) < 1e-10 * ro[ "cost" ]


def _stats( raw ):
    from sdot.SdotPlanNd import _STATS
    return dict( zip( _STATS, np.asarray( raw ).tolist() ) )


def test_limits_step_matches_the_oracle( oracle, native ):
    # the step by the limits ( the 2D default ): the forward limits of the accepted diagram, and the backward rounds on
    # the trials that fall under the floor ( which a bolder factor than 0.9 provokes ): the oracle's decisions
    from sdot import Iterative, Tuning, UniformBox
    from sdot.SdotPlanNd import card_options
    rng = np.random.default_rng( 24 )
    pos = np.clip( np.concatenate( [ rng.normal( c, 0.01, ( 250, 2 ) ) for c in rng.uniform( 0.2, 0.8, ( 8, 2 ) ) ] ), 1e-3, 1 - 1e-3 )
    n = len( pos )
    box = UniformBox( [ 0, 0 ], [ 1, 1 ] )
    for factor in ( 0.9, 1.5 ):
        opts = card_options( Iterative( precision = "fp64", tol = 1e-10, continuation = "never", tuning = Tuning( step = "limits", linear_solver = "cg", linear_tol = 1e-11 ) ), box )
        opts[ native.CARD_OPTIONS.index( "factor" ) ] = factor
        say( "case limits clusters factor", factor )
        r = native.solve_2d( pos, np.full( n, 1 / n ), None, opts, [ 0, 0, 1, 1 ], box )
        st = _stats( r[ "stats" ] )
        ro = oracle.solve( pos, tol = 1e-10, step = "limits", factor = factor, lin_tol = 1e-11 )
        assert st[ "status" ] == 1 and ro[ "stats" ][ "status" ] == 1
        for key in ( "nb_iter", "nb_diag", "nb_backtracks", "nb_limit_rounds", "nb_cell_lim" ):
            assert st[ key ] == ro[ "stats" ][ key ], ( factor, key )
        if factor > 1:
            assert st[ "nb_limit_rounds" ] > 0
        w = r[ "weights" ].cpu().numpy()
        assert np.abs( w - ro[ "weights" ] ).max() < 1e-9 * np.abs( ro[ "weights" ] ).max()


def test_limits_step_on_an_image_matches_the_oracle( oracle, native ):
    from sdot import Image
    rng = np.random.default_rng( 1 )
    n = 20000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    img = H.bumps_image( 128, 128, rng )
    dens = Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 128 ).normalized_version()
    say( "case limits image 20k" )
    sol, ro = _solve_both( oracle, pos, dens, _img( dens.values ), 1e-9 / 10, 1e-10, step = "limits" )
    assert sol.converged and ro[ "stats" ][ "status" ] == 1
    assert sol.stats[ "nb_iter" ] == ro[ "stats" ][ "nb_iter" ] and sol.stats[ "nb_diag" ] == ro[ "stats" ][ "nb_diag" ]
    w = sol.weights.cpu().numpy()
    assert np.abs( w - ro[ "weights" ] ).max() < 1e-9 * np.abs( ro[ "weights" ] ).max()


def test_continuation_matches_the_oracle( oracle, native ):
    # the width continuation: the same stages ( the image blurred on the device by the oracle's filter ), a Newton per
    # stage, the last one on the true density
    from sdot import Image
    rng = np.random.default_rng( 25 )
    n = 5000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    img = H.bumps_image( 64, 64, rng )
    dens = Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 64 ).normalized_version()
    say( "case continuation 5k" )
    sol, ro = _solve_both( oracle, pos, dens, _img( dens.values ), 1e-11, 1e-12, step = "limits", continuation = "always", maxit = 200 )
    assert sol.converged and ro[ "stats" ][ "status" ] == 1
    for key in ( "nb_continuation_steps", "nb_iter", "nb_diag" ):
        assert sol.stats[ key ] == ro[ "stats" ][ key ], key
    hist = np.array( [ [ row[ "step" ], row[ "s" ] ] for row in sol.history ] )
    assert np.array_equal( hist, ro[ "history" ][ :, [ 0, 7 ] ] )
    w = sol.weights.cpu().numpy()
    assert np.abs( w - ro[ "weights" ] ).max() < 1e-9 * np.abs( ro[ "weights" ] ).max()
    m = sol.cell_masses.cpu().numpy()
    assert np.abs( m - oracle.measures( pos, w, density = _img( dens.values ) ) ).max() < 1e-12 * m.max()


_GAUSS = np.array( [ [ 0.3, 0.35, 0.06, 1.0 ], [ 0.7, 0.6, 0.1, 1.5 ], [ 0.45, 0.8, 0.2, 0.5 ] ] )


def _sog( rows = _GAUSS ):
    from sdot import SumOfGaussians
    return SumOfGaussians( rows[ :, :2 ], rows[ :, 2 ], rows[ :, 3 ], support_box = ( [ 0, 0 ], [ 1, 1 ] ) ).normalized_version()


def test_gaussian_masses_and_facets_match_the_oracle( oracle, native ):
    # the corners of each edge ( polygon sums, device ) against the oracle's fan triangles, the erf facets
    rng = np.random.default_rng( 40 )
    dens = _sog()
    rows = np.concatenate( [ dens.positions, dens.sigmas[ :, None ], dens.weights[ :, None ] ], axis = 1 )
    for n, spread in ( ( 3000, 1e-4 ), ( 50000, 0.0 ) ):
        pos = rng.uniform( 0, 1, ( n, 2 ) )
        w = rng.normal( 0, spread, n ) if spread else np.zeros( n )
        say( "case gaussians", n )
        m, ( fi, fj, fc ) = _gpu_measures( native, pos, w, dens, facets = True )
        mo, ( oi, oj, oc ) = oracle.measures( pos, w, density = dict( gaussians = rows ), facets = True )
        assert np.abs( m - mo ).max() <= 1e-12 * mo.max()
        a = { ( i, j ): c for i, j, c in zip( fi.tolist(), fj.tolist(), fc.tolist() ) }
        b = { ( i, j ): c for i, j, c in zip( oi.tolist(), oj.tolist(), oc.tolist() ) }
        assert a.keys() == b.keys()
        ref = max( b.values() )
        assert max( abs( a[ k ] - b[ k ] ) for k in b ) <= 1e-12 * ref


def test_gaussian_solve_matches_the_oracle( oracle, native ):
    rng = np.random.default_rng( 41 )
    n = 4000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    dens = _sog()
    rows = dict( gaussians = np.concatenate( [ dens.positions, dens.sigmas[ :, None ], dens.weights[ :, None ] ], axis = 1 ) )
    for continuation in ( "never", "always" ):
        say( "case gaussian solve", continuation )
        sol, ro = _solve_both( oracle, pos, dens, rows, 1e-11, 1e-12, step = "limits", continuation = continuation, maxit = 200 )
        assert sol.converged and ro[ "stats" ][ "status" ] == 1
        for key in ( "nb_continuation_steps", "nb_iter" ):
            assert sol.stats[ key ] == ro[ "stats" ][ key ], ( continuation, key )
        w = sol.weights.cpu().numpy()
        assert np.abs( w - ro[ "weights" ] ).max() < 1e-9 * np.abs( ro[ "weights" ] ).max()
        assert np.abs( sol.barycenters.cpu().numpy() - ro[ "barycenters" ] ).max() < 1e-10
        assert abs( sol.cost - ro[ "cost" ] ) < 1e-10 * ro[ "cost" ]


@pytest.mark.parametrize( "step,lin", [ ( "trials", "cg" ), ( "limits", "cg" ), ( "limits", "mg" ) ] )
@pytest.mark.parametrize( "name", H.EXACT_CLOUDS + H.NEAR_CLOUDS )
def test_aggregation_matches_the_oracle( oracle, native, name, step, lin ):
    # exact duplicates and near-coincident seeds merged ( reference tests/test_SdotPlanNd.py:1436-1452 ): converged ( aggregated
    # unless only exact duplicates ), the clusters found, each cluster's mass its target, a duplicate's cell empty and its
    # weight `k gap` under its representative's; the oracle's iterations, weights and clusters
    from sdot import Iterative, OtProblem, SumOfDiracs, Tuning, UniformBox
    pos, groups = H.degenerate_clouds()[ name ]
    n = len( pos )
    tol = 1e-11 / n
    say( "case aggregation", name, step, lin )
    it = Iterative( precision = "fp64", tol = tol, max_iter = 60, continuation = "never", tuning = Tuning( step = step, linear_solver = lin, linear_tol = 1e-14 ) )
    sol = OtProblem( SumOfDiracs( pos ), UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).solve( it )
    st = sol.stats
    w = sol.weights.cpu().numpy()
    m = sol.cell_masses.cpu().numpy()
    cl = sol.clusters.cpu().numpy()
    H.check_aggregated( name, st, w, m, np.full( n, 1 / n ), cl, groups, tol )
    assert st[ "duplicates" ] == "checked"
    dups = H.brute_duplicates( pos )
    assert st[ "nb_duplicates" ] == len( dups )
    gap = H.duplicate_gap( pos )
    for k, ( d, r ) in enumerate( dups ):
        assert w[ d ] == w[ r ] - ( 1 + int( ( dups[ :k, 1 ] == r ).sum() ) ) * gap
    ro = oracle.solve( pos, tol = tol, maxit = 60, step = step, dups = dups if len( dups ) else None, gap = gap, margin = 4.0,
                       lin_tol = 1e-14, lin_maxit = 200000 )
    assert ro[ "stats" ][ "status" ] in ( 1, 7 )
    assert st[ "status" ] == ( "converged" if ro[ "stats" ][ "status" ] == 1 else "converged (aggregated)" )
    assert np.array_equal( cl, ro[ "clusters" ] )
    if lin == "cg":                                      # ( the multigrid's directions differ within the linear tolerance )
        assert st[ "nb_iter" ] == ro[ "stats" ][ "nb_iter" ], name
    # the weights agree where the doubles place them: a merged pair's split is only known to its floor
    alone = np.setdiff1d( np.arange( n ), [ i for g, _ in groups for i in g ] )
    assert np.abs( w[ alone ] - ro[ "weights" ][ alone ] ).max() < 1e-9 * np.abs( ro[ "weights" ] ).max(), name


def test_without_the_merge_duplicates_do_not_converge( native ):
    # two equal points cannot split a cell ( reference tests/test_SdotPlanNd.py:1449-1452 )
    from sdot import Iterative, OtProblem, SumOfDiracs, Tuning, UniformBox
    pos, _ = H.degenerate_clouds()[ "exact pair" ]
    say( "case duplicates without the merge" )
    it = Iterative( precision = "fp64", tol = 1e-11 / len( pos ), max_iter = 30, aggregate = False, on_failure = "ignore", continuation = "never",
                    tuning = Tuning( linear_solver = "cg" ) )
    off = OtProblem( SumOfDiracs( pos ), UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).solve( it )
    assert not off.converged and off.stats[ "aggregation" ] == "off" and off.clusters is None


def test_solve_is_bitwise_reproducible( native ):
    from sdot import Iterative, OtProblem, SumOfDiracs, UniformBox
    pos = np.random.default_rng( 2 ).uniform( 0, 1, ( 30000, 2 ) )
    say( "case determinism" )
    a = OtProblem( SumOfDiracs( pos ), UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).solve( Iterative( precision = "fp64", tol = 1e-10 ) )
    b = OtProblem( SumOfDiracs( pos ), UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).solve( Iterative( precision = "fp64", tol = 1e-10 ) )
    assert np.array_equal( a.weights.cpu().numpy(), b.weights.cpu().numpy() )
    assert a.stats[ "nb_diag" ] == b.stats[ "nb_diag" ]


def test_similarity_start_on_the_gpu( native ):
    from sdot import Iterative, OtProblem, SumOfDiracs, UniformBox
    pos = np.random.default_rng( 3 ).uniform( 2, 3, ( 2000, 2 ) )
    say( "case similarity" )
    sol = OtProblem( SumOfDiracs( pos ), UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).solve( Iterative( precision = "fp64", tol = 1e-10 ) )
    assert sol.stats[ "start" ] == "similarity" and sol.converged


def test_warm_restart_takes_few_steps( native ):
    from sdot import Iterative, OtProblem, SumOfDiracs, UniformBox
    rng = np.random.default_rng( 4 )
    pos = rng.uniform( 0, 1, ( 20000, 2 ) )
    pb = OtProblem( SumOfDiracs( pos ), UniformBox( [ 0, 0 ], [ 1, 1 ] ) )
    say( "case warm restart" )
    first = pb.solve( Iterative( precision = "fp64", tol = 1e-10 ) )
    pb.source = SumOfDiracs( pos + 1e-5 * rng.normal( size = pos.shape ).clip( -3, 3 ) )
    second = pb.solve( Iterative( precision = "fp64", tol = 1e-10 ) )
    assert second.stats[ "warm_start" ] == "ot_plan" and second.stats[ "nb_iter" ] <= 3 < first.stats[ "nb_iter" ]


def test_full_size_uniform_million( native ):
    # the 1M-Dirac shape: converged to 1e-9, masses summing to the domain, the gauge held
    from sdot import Iterative, OtProblem, SumOfDiracs, UniformBox
    pos = np.random.default_rng( 1 ).uniform( 0, 1, ( 1_000_000, 2 ) )
    say( "case 1M uniform" )
    sol = OtProblem( SumOfDiracs( pos ), UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).solve( Iterative( precision = "fp64", tol = 1e-9 ) )
    m = sol.cell_masses.cpu().numpy()
    assert sol.converged and np.abs( m - 1e-6 ).max() <= 1e-9
    assert abs( m.sum() - 1 ) < 1e-10
    assert sol.weights.cpu().numpy()[ 0 ] == 0


# ---- many instances at once ( the batched configuration: uneven sizes, one image each ) -------------------------------


def _batch_problems( k = 10, seed = 11 ):
    from sdot import Image, OtProblem, SumOfDiracs
    rng = np.random.default_rng( seed )
    sizes = np.round( np.exp( rng.uniform( np.log( 1e3 ), np.log( 4e4 ), k ) ) ).astype( int )
    out = []
    for q, n in enumerate( sizes ):
        img = H.bumps_image( 64, 64, rng, k = 4 )
        out.append( OtProblem( SumOfDiracs( rng.uniform( 0, 1, ( int( n ), 2 ) ) ), Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 64 ) ) )
    return out


def test_batch_results_are_those_of_single_solves( native ):
    # each instance's plan is the one it gets alone, to the bit: nothing is shared between the streams but the device
    from sdot import Iterative, solve_batch
    probs = _batch_problems()
    it = Iterative( precision = "fp64", tol = 1e-9, continuation = "never" )
    say( "case batch of", len( probs ), "sizes", [ p.nb_diracs for p in probs ] )
    plans = solve_batch( probs, it, streams = 4 )
    assert len( plans ) == len( probs )
    for pb, pl in zip( probs, plans ):
        alone = pb.solve( it )
        assert pl.converged and alone.converged
        assert pl.stats[ "nb_iter" ] == alone.stats[ "nb_iter" ] and pl.stats[ "nb_diag" ] == alone.stats[ "nb_diag" ]
        assert torch_equal( pl.weights, alone.weights ) and torch_equal( pl.cell_masses, alone.cell_masses )
        assert pl.cost == alone.cost
        assert np.abs( pl.cell_masses.cpu().numpy() - 1.0 / pb.nb_diracs ).max() <= 1e-9
    # and again with one stream, then with as many streams as instances: the same plans
    for streams in ( 1, len( probs ) ):
        again = solve_batch( probs, it, streams = streams )
        for a, b in zip( plans, again ):
            assert torch_equal( a.weights, b.weights )


def torch_equal( a, b ):
    return bool( ( a == b ).all().item() )


# ---- serving: requests admitted as they come ( continuous batching ) ---------------------------------------------------


def test_server_results_are_those_of_single_solves( native ):
    # requests arriving at Poisson times while others run: each plan is the one it gets alone, to the bit, and the
    # requests' work overlaps on the device
    import time
    from sdot import Iterative, Server
    probs = _batch_problems( k = 12, seed = 12 )
    it = Iterative( precision = "fp64", tol = 1e-9, continuation = "never" )
    gaps = np.random.default_rng( 3 ).exponential( 0.01, len( probs ) )
    say( "case server of", len( probs ), "requests" )
    with Server( slots = 4 ) as srv:
        tickets = []
        for pb, g in zip( probs, gaps ):
            time.sleep( float( g ) )
            tickets.append( srv.submit( pb, it ) )
        plans = [ t.result() for t in tickets ]
    for pb, pl in zip( probs, plans ):
        alone = pb.solve( it )
        assert pl.converged and alone.converged
        assert torch_equal( pl.weights, alone.weights ) and torch_equal( pl.cell_masses, alone.cell_masses ) and pl.cost == alone.cost
    times = np.asarray( [ t.times for t in tickets ] )
    assert ( times[ :, 0 ] <= times[ :, 1 ] ).all() and ( times[ :, 1 ] < times[ :, 2 ] ).all()
    overlap = any( times[ i, 1 ] < times[ j, 2 ] and times[ j, 1 ] < times[ i, 2 ] for i in range( len( times ) ) for j in range( i ) )
    assert overlap


def test_server_admission_keeps_to_the_memory_budget( native ):
    # a budget that holds one request at a time: the requests run one after the other, in arrival order, and the slots'
    # device bytes together ( sampled at every request's end ) never pass the budget
    from sdot import Iterative, Server, memory
    probs = _batch_problems( k = 4, seed = 13 )
    it = Iterative( precision = "fp64", tol = 1e-9, continuation = "never" )
    need = [ sum( memory.solve_bytes( pb.nb_diracs, 2, pb.target, it ).values() ) for pb in probs ]
    budget = 1.2 * max( need )
    assert 2 * min( need ) > budget
    say( f"case server with a budget of { budget / 1e6 :.0f} MB for requests of { [ round( b / 1e6 ) for b in need ] } MB" )
    for mem in ( 1.0, budget ):
        with Server( slots = 3, memory = mem ) as srv:
            tickets = [ srv.submit( pb, it ) for pb in probs ]
            plans = [ t.result() for t in tickets ]
            info = srv.info()
        assert all( p.converged for p in plans ) and info[ "done" ] == len( probs )
        times = sorted( t.times[ 1: ] for t in tickets )
        for ( s0, e0 ), ( s1, e1 ) in zip( times, times[ 1: ] ):
            assert e0 <= s1
        assert info[ "peak_bytes" ] <= max( budget, max( need ) ), ( info, budget )


def test_solve_bytes_holds_the_real_footprint( native ):
    # the byte model against what the library really took for a solve ( its arena, emptied first ): never under it, and
    # at most 10 % above on these generic clouds
    import sdot
    from sdot import memory
    rng = np.random.default_rng( 21 )
    img = sdot.Image( H.bumps_image( 128, 128, rng, k = 4 ), origin = [ 0, 0 ], frame = np.eye( 2 ) / 128 )
    cases = [ ( "2D image, mixed, limits", 2, 60_000, img, sdot.Iterative( tol = 1e-9, continuation = "never" ) ),
              ( "2D constant, fp64, trials", 2, 30_000, sdot.UniformBox( [ 0, 0 ], [ 1, 1 ] ),
                sdot.Iterative( tol = 1e-9, precision = "fp64", tuning = sdot.Tuning( step = "trials" ) ) ),
              ( "3D Gaussians", 3, 4_000, sdot.SumOfGaussians( rng.uniform( 0.3, 0.7, ( 2, 3 ) ), [ 0.15, 0.2 ], support_box = ( [ 0, 0, 0 ], [ 1, 1, 1 ] ) ),
                sdot.Iterative( tol = 1e-9 ) ) ]
    for label, d, n, dens, it in cases:
        say( f"case byte model: { label }, n = { n }" )
        native.release()
        pos = rng.uniform( 0, 1, ( n, d ) )
        sol = sdot.OtProblem( sdot.SumOfDiracs( pos ), dens ).solve( it )
        assert sol.converged
        real = native.device_bytes()
        model = sum( memory.solve_bytes( n, d, dens, it ).values() )
        assert real <= model <= 1.10 * real, ( label, real, model )
