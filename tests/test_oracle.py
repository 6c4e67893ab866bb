"""The oracle pinned: each function of `oracle/sdot_oracle.cpp` against what the mathematics fixes
( closed forms, invariants, an independent library, brute force, finite differences ), mirroring the
reference's own tests ( file:line cited per test ). CPU only."""

import numpy as np
import pytest

import helpers as H


# ---- the tree ( reference tests/test_PowerDiagram.py:773-900 ) -----------------------------------------------


def _check_tree( O, pos, w, leaf = 10 ):
    n = len( pos )
    t = O.build_tree( pos, w, leaf )
    depth = t[ "depth" ]
    assert depth == O.max_depth_for( n, leaf )
    N = 2 ** depth - 1
    assert len( t[ "beg" ] ) == N
    # every seed exactly once ( test_PowerDiagram.py:773 )
    assert np.array_equal( np.sort( t[ "order" ] ), np.arange( n ) )
    sp = pos[ t[ "order" ] ]
    ww = None if w is None else w[ t[ "order" ] ]

    def walk( m, h ):
        b, e = t[ "beg" ][ m ], t[ "end" ][ m ]
        if e > b:
            # the box contains the subtree ( test_PowerDiagram.py:835 )
            lo, hi = t[ "box" ][ m ]
            assert np.all( sp[ b:e ] >= lo ) and np.all( sp[ b:e ] <= hi )
            assert np.array_equal( lo, sp[ b:e ].min( 0 ) ) and np.array_equal( hi, sp[ b:e ].max( 0 ) )
            # the majorant majorates ( test_PowerDiagram.py:870 )
            if ww is not None:
                assert np.all( ww[ b:e ] <= sp[ b:e ] @ t[ "wa" ][ m ] + t[ "wb" ][ m ] )
        if h <= 1:
            assert e - b <= max( leaf, 1 ) or np.all( sp[ b:e ] == sp[ b ] )
            return
        lc, rc = m + 1, m + 2 ** ( h - 1 )
        # the children partition the slice
        assert t[ "beg" ][ lc ] == b and t[ "end" ][ rc ] == e and t[ "end" ][ lc ] == t[ "beg" ][ rc ]
        walk( lc, h - 1 )
        walk( rc, h - 1 )

    walk( 0, depth )
    return t


def test_tree_shape_depends_on_n_only( oracle ):
    # the tree's shape is a function of n alone, degenerate clouds included ( test_PowerDiagram.py:801 )
    rng = np.random.default_rng( 0 )
    for n in ( 1, 2, 10, 11, 57, 1000 ):
        for pos in ( rng.uniform( 0, 1, ( n, 2 ) ), np.full( ( n, 2 ), 0.3 ), np.column_stack( [ np.linspace( 0, 1, n ), np.zeros( n ) ] ) ):
            _check_tree( oracle, pos, None )
    assert oracle.max_depth_for( 10 ) == 1 and oracle.max_depth_for( 11 ) == 2 and oracle.max_depth_for( 10 ** 6 ) == 18


def test_tree_median_split( oracle ):
    # a median cut on the longest axis: the halves differ by at most one seed and are separated along that axis
    rng = np.random.default_rng( 1 )
    pos = rng.uniform( 0, 1, ( 500, 2 ) ) * [ 3.0, 1.0 ]
    t = _check_tree( oracle, pos, None )
    sp = pos[ t[ "order" ] ]
    h = t[ "depth" ]
    lc, rc = 1, 2 ** ( h - 1 )
    nl = t[ "end" ][ lc ] - t[ "beg" ][ lc ]
    nr = t[ "end" ][ rc ] - t[ "beg" ][ rc ]
    assert abs( nl - nr ) <= 1
    assert sp[ :nl, 0 ].max() <= sp[ nl:, 0 ].min()


def test_majorants_majorate( oracle ):
    rng = np.random.default_rng( 2 )
    pos = rng.uniform( 0, 1, ( 2000, 2 ) )
    for w in ( rng.uniform( -1, 1, 2000 ), 0.3 * pos[ :, 0 ] - 0.7 * pos[ :, 1 ] + 1e-3 * rng.normal( size = 2000 ),
               np.sin( 5 * pos[ :, 0 ] ) * 1e-2 ):
        _check_tree( oracle, pos, w )


def test_majorant_takes_an_affine_weight( oracle ):
    # an exactly affine weight: the least-squares slope is the slope, and the bound is tight up to its margin
    rng = np.random.default_rng( 3 )
    pos = rng.uniform( 0, 1, ( 40, 2 ) )
    w = 0.3 * pos[ :, 0 ] - 0.7 * pos[ :, 1 ] + 2.0
    wa, wb = oracle.weight_majorant( pos, w )
    assert np.allclose( wa, [ 0.3, -0.7 ], atol = 1e-12 )
    assert abs( wb - 2.0 ) < 1e-5
    # random weights: the slope is refused ( it does not tighten the spread by more than chance ), b = max w
    w = rng.uniform( -1, 1, 40 )
    wa, wb = oracle.weight_majorant( pos, w )
    assert np.all( wa == 0 ) and wb >= w.max() and wb - w.max() < 1e-5


# ---- the cells ( reference tests/test_PowerDiagram.py:46-330 ) --------------------------------------------------


def test_one_seed_takes_everything( oracle ):
    m = oracle.measures( np.array( [ [ 0.3, 0.3 ] ] ), lo = ( 0, 0 ), hi = ( 2, 2 ) )
    assert abs( m[ 0 ] - 4.0 ) < 1e-12


def test_sum_is_the_domain( oracle ):
    rng = np.random.default_rng( 14 )
    for label, pos, w in H.clouds( rng, 3000 ):
        m = oracle.measures( pos, w )
        assert abs( m.sum() - 1 ) < 1e-12, label
        assert np.all( m >= 0 )


def test_bisector_shifted_by_the_weight_gap( oracle ):
    # two seeds at x = 0 and 1: the plane is at ( 1 + w0 - w1 ) / 2 ( test_PowerDiagram.py:262 )
    for gap in ( -0.6, -0.25, 0.0, 0.25, 0.6 ):
        m = oracle.measures( np.array( [ [ 0.0, 0.5 ], [ 1.0, 0.5 ] ] ), np.array( [ gap, 0.0 ] ) )
        assert abs( m[ 0 ] - ( 1 + gap ) / 2 ) < 1e-14, gap


def test_seeds_on_a_grid( oracle ):
    # four cells meet at each grid point, the bisectors pass through vertices ( test_PowerDiagram.py:217 )
    xs = np.array( [ 1, 3, 5 ] ) / 6
    pos = np.array( [ [ x, y ] for x in xs for y in xs ] )
    assert np.allclose( oracle.measures( pos ), 1 / 9, atol = 1e-15 )
    g = ( np.arange( 32 ) + 0.5 ) / 32
    pos = np.array( [ [ x, y ] for x in g for y in g ] )
    assert np.allclose( oracle.measures( pos ), 1 / 1024, atol = 1e-17 )


def test_a_seed_outside_the_domain_has_no_cell( oracle ):
    m = oracle.measures( np.array( [ [ 0.5, 0.5 ], [ 0.2, 0.2 ], [ 5.0, 5.0 ] ] ) )
    assert abs( m[ 2 ] ) < 1e-14 and abs( m.sum() - 1 ) < 1e-12


def test_weight_shift_changes_nothing( oracle ):
    rng = np.random.default_rng( 33 )
    pos = rng.uniform( 0.1, 0.9, ( 25, 2 ) )
    w = rng.uniform( -0.05, 0.05, 25 )
    assert np.allclose( oracle.measures( pos, w ), oracle.measures( pos, w + 2.5 ), atol = 1e-12 )
    assert np.allclose( oracle.measures( pos ), oracle.measures( pos, np.full( 25, -3.0 ) ), atol = 1e-12 )


def test_tree_walk_equals_brute_force( oracle ):
    # pruning never loses a cut: the tree's cells are the all-seeds cells ( test_PowerDiagram.py:955-1036 )
    rng = np.random.default_rng( 4 )
    for label, pos, w in H.clouds( rng, 1500 ) + [ ( "ring", H.ring( 40, rng, 400 ), None ) ]:
        a = oracle.measures( pos, w )
        b = oracle.measures( pos, w, brute_force = True )
        assert np.abs( a - b ).max() <= 1e-15, label


def test_neighbours_and_areas_match_the_lifted_hull( oracle ):
    # an independent diagram: the lower convex hull of ( p, |p|^2 - w ) ( scipy / Qhull ), and half-space
    # intersections for the areas
    rng = np.random.default_rng( 5 )
    n = 300
    pos = rng.uniform( 0.05, 0.95, ( n, 2 ) )
    for w in ( None, rng.uniform( -2e-3, 2e-3, n ) ):
        m, ( fi, fj, fc ) = oracle.measures( pos, w, facets = True )
        ours = { ( min( i, j ), max( i, j ) ) for i, j in zip( fi.tolist(), fj.tolist() ) }
        hull, amb = H.lifted_neighbours( pos, w, ambiguous = True )
        # the same neighbour sets: every pair whose facet meets the box ( over more than 1e-10 ) and no other, up to the
        # pairs whose facet meets it over less, which rounding may keep or drop ( none on this cloud )
        assert hull <= ours <= hull | amb
        assert len( amb ) == 0
        for i in range( 0, n, 7 ):
            v = H.halfspace_cell( pos, w, i )
            ref = 0.0 if v is None else H.polygon_area( v )
            assert abs( m[ i ] - ref ) <= 1e-11 * max( ref, 1e-6 ), i


@pytest.mark.parametrize( "label", [ "uniform", "lines", "lines weighted" ] )
def test_neighbour_sets_equal_the_lifted_hull_on_the_reference_clouds( oracle, label ):
    """identical neighbour sets on generic clouds, the uniform one and the reference's thin lines, weighted or not"""
    rng = np.random.default_rng( 51 )
    for lab, pos, w in H.clouds( rng, 2000 ):
        if lab != label:
            continue
        _, ( fi, fj, _ ) = oracle.measures( pos, w, facets = True )
        ours = { ( min( i, j ), max( i, j ) ) for i, j in zip( fi.tolist(), fj.tolist() ) }
        hull, amb = H.lifted_neighbours( pos, w, ambiguous = True )
        assert hull <= ours <= hull | amb, ( len( hull - ours ), len( ours - hull - amb ) )


def test_the_lifted_hull_filter_catches_a_dropped_facet( oracle ):
    """the check above is two-sided: a facet taken out of the oracle's set is seen missing, a pair added is seen extra"""
    rng = np.random.default_rng( 52 )
    pos = rng.uniform( 0.05, 0.95, ( 300, 2 ) )
    _, ( fi, fj, _ ) = oracle.measures( pos, None, facets = True )
    ours = { ( min( i, j ), max( i, j ) ) for i, j in zip( fi.tolist(), fj.tolist() ) }
    hull = H.lifted_neighbours( pos, None )
    assert ours == hull
    dropped = set( list( ours )[ 1: ] )
    assert not ( hull <= dropped )
    far = ( int( np.argmin( pos[ :, 0 ] ) ), int( np.argmax( pos[ :, 0 ] ) ) )
    assert not ( ( ours | { tuple( sorted( far ) ) } ) <= hull )


def test_cell_vertices_and_ids( oracle ):
    rng = np.random.default_rng( 6 )
    pos = rng.uniform( 0, 1, ( 200, 2 ) )
    w = rng.uniform( -1e-3, 1e-3, 200 )
    for i in ( 0, 17, 123 ):
        xy, ids = oracle.cell( pos, i, w )
        xy2, ids2 = oracle.cell( pos, i, w, brute_force = True )
        # the same cyclic sequence ( the cut order changes where the list starts )
        assert len( xy ) == len( xy2 )
        r = int( np.argmin( np.abs( xy2 - xy[ 0 ] ).sum( 1 ) ) )
        assert np.allclose( xy, np.roll( xy2, -r, axis = 0 ), atol = 1e-15 ) and np.array_equal( ids, np.roll( ids2, -r ) )
        assert H.polygon_area( xy ) > 0
        # each vertex is on the bisectors of its two cuts ( cut q carries the edge [ v_q, v_q+1 ] )
        for q in range( len( xy ) ):
            for qq in ( q, q - 1 ):
                j = ids[ qq ]
                if j < 0:
                    continue
                lhs = ( xy[ q ] - pos[ i ] ) @ ( xy[ q ] - pos[ i ] ) - w[ i ]
                rhs = ( xy[ q ] - pos[ j ] ) @ ( xy[ q ] - pos[ j ] ) - w[ j ]
                assert abs( lhs - rhs ) < 1e-12


# ---- image ( reference tests/test_PowerDiagram.py:1527-1660 ) ---------------------------------------------------


def _img( values, lo = ( 0, 0 ), hi = ( 1, 1 ) ):
    ny, nx = values.shape
    return dict( values = values, x0 = lo[ 0 ], y0 = lo[ 1 ], hx = ( hi[ 0 ] - lo[ 0 ] ) / nx, hy = ( hi[ 1 ] - lo[ 1 ] ) / ny )


def test_constant_image_is_lebesgue( oracle ):
    rng = np.random.default_rng( 7 )
    pos = rng.uniform( 0, 1, ( 400, 2 ) )
    w = rng.uniform( -1e-3, 1e-3, 400 )
    a = oracle.measures( pos, w, density = 2.5 )
    b = oracle.measures( pos, w, density = _img( np.full( ( 13, 17 ), 2.5 ) ) )
    assert np.allclose( a, b, rtol = 1e-12, atol = 1e-16 )


def test_image_against_exact_clipping( oracle ):
    rng = np.random.default_rng( 8 )
    pos = rng.uniform( 0, 1, ( 120, 2 ) )
    w = rng.uniform( -2e-3, 2e-3, 120 )
    img = rng.uniform( 0.1, 2.0, ( 9, 11 ) )          # [ ny, nx ]
    d = _img( img )
    m = oracle.measures( pos, w, density = d )
    for i in range( 0, 120, 5 ):
        xy, _ = oracle.cell( pos, i, w )
        ref = H.image_mass_of_polygon( xy, img, 0, 0, d[ "hx" ], d[ "hy" ] )
        assert abs( m[ i ] - ref ) <= 1e-13 * max( ref, 1e-8 ), i
    assert abs( m.sum() - img.mean() ) < 1e-13


def test_one_pixel_moments_closed_form( oracle ):
    # the single seed in a box against a constant density: barycentre = the box centre, cost = int |x - p|^2
    r = oracle.solve( np.array( [ [ 0.3, 0.6 ], [ 0.7, 0.4 ] ] ), tol = 1e-12 )
    # two seeds symmetric about the centre: at the solution, each takes half of the square
    assert r[ "stats" ][ "status" ] == 1
    assert np.allclose( r[ "masses" ], 0.5, atol = 1e-12 )
    # closed form of the cost for a half-square cut by the bisector through the centre, by quadrature on a fine grid
    g = ( np.arange( 2000 ) + 0.5 ) / 2000
    X, Y = np.meshgrid( g, g, indexing = "ij" )
    d0 = ( X - 0.3 ) ** 2 + ( Y - 0.6 ) ** 2
    d1 = ( X - 0.7 ) ** 2 + ( Y - 0.4 ) ** 2
    w = r[ "weights" ]
    cost = np.where( d0 - w[ 0 ] <= d1 - w[ 1 ], d0, d1 ).mean()
    assert abs( r[ "cost" ] - cost ) < 1e-6


# ---- facets and the Hessian ( reference tests/test_SdotPlanNd.py:184-243 ) ---------------------------------------


def _dense_laplacian( n, fi, fj, fc ):
    L = np.zeros( ( n, n ) )
    for i, j, c in zip( fi, fj, fc ):
        L[ i, j ] -= c
        L[ i, i ] += c
    return L


@pytest.mark.parametrize( "image", [ False, True ] )
def test_hessian_is_the_jacobian_of_the_masses( oracle, image ):
    rng = np.random.default_rng( 9 )
    n = 40
    pos = rng.uniform( 0.05, 0.95, ( n, 2 ) )
    w = rng.uniform( -3e-3, 3e-3, n )
    dens = _img( rng.uniform( 0.2, 3.0, ( 7, 6 ) ) ) if image else None
    m, ( fi, fj, fc ) = oracle.measures( pos, w, density = dens, facets = True )
    L = _dense_laplacian( n, fi, fj, fc )
    h = 1e-6
    for j in range( 0, n, 3 ):
        e = np.zeros( n )
        e[ j ] = h
        J = ( oracle.measures( pos, w + e, density = dens ) - oracle.measures( pos, w - e, density = dens ) ) / ( 2 * h )
        assert np.abs( J - L[ :, j ] ).max() <= 1e-7 * np.abs( L ).max(), j
    # symmetric and zero row sums
    assert np.abs( L - L.T ).max() <= 1e-12 * np.abs( L ).max()
    assert np.abs( L.sum( 1 ) ).max() <= 1e-12 * np.abs( L ).max()


def test_cg_solves_the_gauged_laplacian( oracle ):
    rng = np.random.default_rng( 10 )
    n = 400
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    _, ( fi, fj, fc ) = oracle.measures( pos, facets = True )
    b = rng.normal( size = n )
    b -= b.mean()
    d, its = oracle.cg( n, fi, fj, fc, b, tol = 1e-13 )
    assert its > 0 and d[ 0 ] == 0
    L = _dense_laplacian( n, fi, fj, fc )
    ref = np.zeros( n )
    ref[ 1: ] = np.linalg.solve( L[ 1:, 1: ], b[ 1: ] )
    assert np.abs( d - ref ).max() <= 1e-9 * np.abs( ref ).max()


def _sparse_direct( n, fi, fj, fc, b ):
    """the gauged system `L d = b`, `d[ 0 ] = 0`, by a sparse LU ( SuperLU ) of `L` with row and column 0 struck out"""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sl
    L = sp.coo_matrix( ( -np.asarray( fc ), ( fi, fj ) ), shape = ( n, n ) ).tocsr()
    L = L + sp.diags( -np.asarray( L.sum( axis = 1 ) ).ravel() )
    ref = np.zeros( n )
    ref[ 1: ] = sl.spsolve( L[ 1:, 1: ].tocsc(), b[ 1: ] )
    return ref


@pytest.mark.parametrize( "by_rank", [ True, False ] )
def test_multigrid_solves_the_gauged_laplacian( oracle, by_rank ):
    """the multigrid-preconditioned CG ( the CPU baseline's linear solver ) against a sparse direct solve, on the Laplacian
    of an 8000-seed diagram ( a hierarchy of 2 levels ). Aggregated by tree rank, its preconditioner must cut the iterations
    of the Jacobi-PCG by a factor of 10 at least ( a wrong prolongation, Galerkin sign or smoother degrades it to CG or
    worse ); by index ( a random cloud's index order: aggregates scattered over the square ) it must still converge"""
    rng = np.random.default_rng( 12 )
    n = 8000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    _, ( fi, fj, fc ) = oracle.measures( pos, facets = True )
    b = rng.normal( size = n )
    b -= b.mean()
    rank_of = None
    if by_rank:
        order = oracle.build_tree( pos )[ "order" ]
        rank_of = np.empty( n, np.int64 )
        rank_of[ order ] = np.arange( n )
    d, its, levels = oracle.mg( n, fi, fj, fc, b, rank_of = rank_of, tol = 1e-12 )
    assert levels == 2 and its > 0 and d[ 0 ] == 0
    ref = _sparse_direct( n, fi, fj, fc, b )
    assert np.abs( d - ref ).max() <= 1e-8 * np.abs( ref ).max()
    _, its_cg = oracle.cg( n, fi, fj, fc, b, tol = 1e-12 )
    print( f"multigrid { its } iterations, Jacobi-PCG { its_cg }" )
    assert not by_rank or 10 * its <= its_cg, ( its, its_cg )


def test_multigrid_newton_solve_is_the_cg_one( oracle ):
    """a Newton solve whose systems go through the multigrid takes the same steps as the Jacobi-PCG one: the same
    iteration count, and weights equal to 1e-9 of their spread ( both linear solves run to 1e-12 )"""
    rng = np.random.default_rng( 13 )
    n = 6000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    img = 1.0 + 0.5 * np.sin( 6 * np.linspace( 0, 1, 64 ) )[ :, None ] * np.cos( 4 * np.linspace( 0, 1, 64 ) )[ None, : ]
    dens = dict( values = img, x0 = 0.0, y0 = 0.0, hx = 1 / 64, hy = 1 / 64 )
    a = oracle.solve( pos, density = dens, tol = 1e-12, step = "limits" )
    b = oracle.solve( pos, density = dens, tol = 1e-12, step = "limits", linear = "mg" )
    assert a[ "stats" ][ "status" ] == 1 and b[ "stats" ][ "status" ] == 1
    assert a[ "stats" ][ "nb_iter" ] == b[ "stats" ][ "nb_iter" ]
    w = a[ "weights" ]
    assert np.abs( w - b[ "weights" ] ).max() <= 1e-9 * ( w.max() - w.min() )
    assert b[ "stats" ][ "lin_nb_iter" ] * 5 <= a[ "stats" ][ "lin_nb_iter" ]


def test_multigrid_3d_newton_solve_is_the_cg_one( oracle ):
    rng = np.random.default_rng( 14 )
    n = 3000
    pos = rng.uniform( 0, 1, ( n, 3 ) )
    a = oracle.solve_3d( pos, tol = 1e-12 )
    b = oracle.solve_3d( pos, tol = 1e-12, linear = "mg" )
    assert a[ "stats" ][ "status" ] == 1 and b[ "stats" ][ "status" ] == 1
    assert a[ "stats" ][ "nb_iter" ] == b[ "stats" ][ "nb_iter" ]
    w = a[ "weights" ]
    assert np.abs( w - b[ "weights" ] ).max() <= 1e-9 * ( w.max() - w.min() )


# ---- the solve ( reference tests/test_SdotPlanNd.py:39-330 ) -----------------------------------------------------


def test_newton_matches_the_target_masses( oracle ):
    rng = np.random.default_rng( 11 )
    n = 2000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    nu = rng.uniform( 0.5, 1.5, n )
    nu /= nu.sum()
    r = oracle.solve( pos, nu, tol = 1e-10 / n * 1e3 )
    assert r[ "stats" ][ "status" ] == 1
    assert np.abs( r[ "masses" ] - nu ).max() <= 1e-10 / n * 1e3
    assert r[ "weights" ][ 0 ] == 0
    # the masses of the returned weights, measured again
    assert np.abs( oracle.measures( pos, r[ "weights" ] ) - r[ "masses" ] ).max() < 1e-15


def test_quadratic_tail_and_warm_start( oracle ):
    # the last steps gain an order of magnitude each, a warm start from the solution ends in a few steps
    # ( test_SdotPlanNd.py:245-276 )
    rng = np.random.default_rng( 12 )
    n = 3000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    img = H.bumps_image( 32, 32, rng ).T
    dens = _img( img / img.mean() )
    r = oracle.solve( pos, density = dens, tol = 1e-10 / n )
    assert r[ "stats" ][ "status" ] == 1
    res = r[ "history" ][ :, 4 ]
    assert res[ -1 ] < 1e-10 / n
    assert res[ -1 ] < 1e-2 * res[ -2 ]
    r2 = oracle.solve( pos, density = dens, tol = 1e-10 / n, w0 = r[ "weights" ] + 1e-7 * rng.normal( size = n ) * 0 )
    assert r2[ "stats" ][ "nb_iter" ] <= 1 and r2[ "stats" ][ "start" ] == 0 or r2[ "stats" ][ "nb_iter" ] == 0


def test_similarity_start( oracle ):
    # seeds outside the domain: the Voronoi start has empty cells, the similarity start has none ( test_SdotPlanNd.py:279-305 )
    rng = np.random.default_rng( 13 )
    pos = rng.uniform( 2, 3, ( 500, 2 ) )
    r = oracle.solve( pos, tol = 1e-9 )
    assert r[ "stats" ][ "start" ] == 2 and r[ "stats" ][ "status" ] == 1
    assert r[ "stats" ][ "min_start_mass" ] > 0


def test_residual_variants_reach_the_same_plan( oracle ):
    # log / lin residuals give the same plan ( test_SdotPlanNd.py:327-347 )
    rng = np.random.default_rng( 15 )
    pos = rng.uniform( 0, 1, ( 800, 2 ) )
    a = oracle.solve( pos, tol = 1e-12, residual = 1 )
    b = oracle.solve( pos, tol = 1e-12, residual = 0 )
    assert a[ "stats" ][ "status" ] == 1 and b[ "stats" ][ "status" ] == 1
    assert np.abs( a[ "weights" ] - b[ "weights" ] ).max() < 1e-9 * np.abs( a[ "weights" ] ).max()


# ---- the step by the limits ( reference gpu/Newton2D.cuh:505-770, sdotplan/Bounds.h:142-230 ) --------------------


def _weighted_cloud( seed, n = 400, spread = 1e-4 ):
    rng = np.random.default_rng( seed )
    return rng.uniform( 0, 1, ( n, 2 ) ), rng.normal( 0, spread, n ), rng.normal( 0, spread, n )


def test_area_polynomial_is_the_area_and_the_laplacian( oracle ):
    # a0 is the cell's area; a1 = d area_i / dt = sum_j c_ij ( d_i - d_j ) with the facets of the Hessian test
    # ( a wrong sign or a swapped operand in the moving offsets breaks it ); while the edges hold, the polynomial
    # is the exact area at w + t d
    pos, w, d = _weighted_cloud( 20 )
    r = oracle.along( pos, w, d )
    a, ( fi, fj, fc ) = oracle.measures( pos, w, facets = True )
    assert np.array_equal( r[ "ok" ], a > 0 ) and r[ "ok" ].sum() > 390   # ( an empty cell has no polynomial: area 0 )
    assert np.abs( r[ "poly" ][ :, 0 ] - a ).max() <= 1e-14 * a.max()
    a1 = np.zeros( len( pos ) )
    np.add.at( a1, fi, fc * ( d[ fi ] - d[ fj ] ) )
    assert np.abs( r[ "poly" ][ :, 1 ] - a1 ).max() <= 1e-12 * np.abs( a1 ).max()
    t = 1e-3
    at = oracle.measures( pos, w + t * d )
    pred = r[ "poly" ] @ np.array( [ 1, t, t * t ] )
    assert np.abs( pred - at ).max() <= 1e-13 * a.max()
    # the frozen cell's mass is the polynomial, whatever t ( constant density )
    ok = r[ "ok" ]
    fm = oracle.along( pos, w, d, t = 0.5 )[ "frozen_mass" ]
    assert np.abs( fm[ ok ] - r[ "poly" ][ ok ] @ np.array( [ 1, 0.5, 0.25 ] ) ).max() <= 1e-14 * a.max()


def test_frozen_image_mass_is_the_clipped_mass( oracle ):
    # the Green walk of a frozen cell against the pixel-clipping quadrature of the measures
    pos, w, d = _weighted_cloud( 21 )
    rng = np.random.default_rng( 22 )
    dens = _img( rng.uniform( 0.5, 2, ( 48, 64 ) ) )
    for t in ( 0.0, 1e-3 ):
        fm = oracle.along( pos, w, d, t = t, density = dens )[ "frozen_mass" ]
        ref = oracle.measures( pos, w + t * d, density = dens )
        assert np.abs( fm - ref ).max() <= 1e-13 * ref.max(), t


def test_an_image_is_zero_outside_its_grid( oracle ):
    # a grid on part of the box ( overhanging it on the left, short of it on the right and at the top ): the pixel clipping
    # of the measures counts nothing off the grid, the Green walk of the frozen cells ( which start from the box's part
    # the grid covers ) agrees, and a one-value image is the constant density on that part
    pos, w, d = _weighted_cloud( 24 )
    rng = np.random.default_rng( 25 )
    dens = _img( rng.uniform( 0.5, 2, ( 20, 30 ) ), lo = ( -0.2, 0.1 ), hi = ( 0.7, 0.8 ) )
    m = oracle.measures( pos, w, density = dens )
    v, hx, hy = dens[ "values" ], dens[ "hx" ], dens[ "hy" ]
    fx = np.clip( ( np.minimum( -0.2 + hx * ( np.arange( 30 ) + 1 ), 1 ) - np.maximum( -0.2 + hx * np.arange( 30 ), 0 ) ) / hx, 0, 1 )
    assert abs( m.sum() - ( v * fx[ None, : ] ).sum() * hx * hy ) < 1e-13
    for t in ( 0.0, 1e-3 ):
        fm = oracle.along( pos, w, d, t = t, density = dens )[ "frozen_mass" ]
        ref = oracle.measures( pos, w + t * d, density = dens )
        ok = np.isfinite( fm )                           # ( NaN: no frozen cell, the seed's cell being empty )
        assert np.all( ok | ( ref == 0 ) ) and ok.sum() > 100
        assert np.abs( fm[ ok ] - ref[ ok ] ).max() <= 1e-13 * ref.max(), t
    one = _img( np.full( ( 20, 30 ), 1.5 ), lo = ( -0.2, 0.1 ), hi = ( 0.7, 0.8 ) )
    a = oracle.measures( pos, w, density = one )
    b = oracle.measures( pos, w, lo = ( 0, 0.1 ), hi = ( 0.7, 0.8 ), density = 1.5 )
    assert np.abs( a - b ).max() < 1e-15


def test_forward_limit_is_the_first_root( oracle ):
    pos, w, d = _weighted_cloud( 23, spread = 2e-5 )
    a = oracle.measures( pos, w )
    eps = 0.9 * a.min()
    assert eps > 0
    r = oracle.along( pos, w, 30 * d, eps = eps )
    f = r[ "forward" ]
    hit = f < 1e299
    assert hit.any() and np.all( f > 0 )
    p = r[ "poly" ]
    # the mass at the limit is the floor, and above it before
    at = p[ hit, 0 ] + f[ hit ] * p[ hit, 1 ] + f[ hit ] ** 2 * p[ hit, 2 ]
    assert np.abs( at - eps ).max() <= 1e-12 * a.max()
    half = p[ hit, 0 ] + 0.5 * f[ hit ] * p[ hit, 1 ] + 0.25 * f[ hit ] ** 2 * p[ hit, 2 ]
    assert np.all( half > eps )
    # an image density bisects the frozen mass on ( 0, 1 ]: within 1e-2 under the same limit on a constant image of value 1
    ri = oracle.along( pos, w, 30 * d, eps = eps, density = _img( np.ones( ( 16, 16 ) ) ) )[ "forward" ]
    in1 = hit & ( f < 1 )
    assert in1.sum() > 10 and np.all( ri[ in1 ] <= f[ in1 ] ) and np.all( ri[ in1 ] >= 0.98 * f[ in1 ] )


def test_the_limits_step_solves_with_fewer_diagrams( oracle ):
    # the same plan as the trials, with fewer diagrams ( reference test_SdotPlanNd.py:917-920 compares the steps )
    rng = np.random.default_rng( 24 )
    pos = np.clip( np.concatenate( [ rng.normal( c, 0.01, ( 250, 2 ) ) for c in rng.uniform( 0.2, 0.8, ( 8, 2 ) ) ] ), 1e-3, 1 - 1e-3 )
    a = oracle.solve( pos, tol = 1e-10, step = "trials", lin_tol = 1e-11 )
    b = oracle.solve( pos, tol = 1e-10, step = "limits", lin_tol = 1e-11 )
    assert a[ "stats" ][ "status" ] == 1 and b[ "stats" ][ "status" ] == 1
    assert b[ "stats" ][ "nb_diag" ] < a[ "stats" ][ "nb_diag" ]
    assert np.abs( a[ "weights" ] - b[ "weights" ] ).max() < 1e-8 * np.abs( a[ "weights" ] ).max()
    # a bolder factor makes trials fall under the floor: the backward rounds then correct them
    c = oracle.solve( pos, tol = 1e-10, step = "limits", factor = 1.5, lin_tol = 1e-11 )
    assert c[ "stats" ][ "status" ] == 1 and c[ "stats" ][ "nb_limit_rounds" ] > 0 and c[ "stats" ][ "nb_cell_lim" ] > 0


# ---- the width continuation ( reference sdotplan/Continuation.h, sdotplan/convolved/Image.h ) ----------------------


def test_blur_keeps_constants_and_spreads_a_pixel( oracle ):
    ones = np.full( ( 40, 50 ), 3.0 )
    assert np.abs( oracle.blur( ones, 0.05, 1 / 50, 1 / 40 ) - ones ).max() <= 4e-15 * 3
    # one bright pixel far from the borders: the outer product of the normalized 1D Gaussians of s / h pixels
    img = np.zeros( ( 61, 61 ) )
    img[ 30, 30 ] = 1.0
    s, h = 0.02, 1 / 61
    b = oracle.blur( img, s, h, h )
    sp = s / h
    r = int( np.ceil( 4 * sp ) )
    k = np.exp( -0.5 * np.arange( -r, r + 1 ) ** 2 / sp ** 2 )
    k /= k.sum()
    ref = np.zeros_like( img )
    ref[ 30 - r:30 + r + 1, 30 - r:30 + r + 1 ] = np.outer( k, k )
    assert np.abs( b - ref ).max() <= 1e-15
    assert abs( b.sum() - 1 ) < 1e-14
    # no blur under a thousandth of a pixel
    assert np.array_equal( oracle.blur( img, 1e-4 * h, h, h ), img )


def test_continuation_ends_on_the_true_density( oracle ):
    # the stages s0, s0 / r, ... >= s_min, then 0 ( Continuation.h:31-39 ); the last stage is the density itself: the
    # returned masses are its measures at the returned weights, within the tolerance of the targets
    rng = np.random.default_rng( 25 )
    n = 1500
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    img = H.bumps_image( 32, 32, rng )
    dens = _img( np.ascontiguousarray( img.T ) / img.mean() )
    r = oracle.solve( pos, tol = 1e-11, density = dens, step = "limits", continuation = "always", lin_tol = 1e-12 )
    st = r[ "stats" ]
    s0, smin = 0.5 * np.sqrt( 2 ), ( 1 / 32 ) / 4
    assert st[ "nb_continuation_steps" ] == int( np.floor( np.log( s0 / smin ) / np.log( np.sqrt( 2 ) ) + 1e-9 ) ) + 2
    assert st[ "status" ] == 1
    m = oracle.measures( pos, r[ "weights" ], density = dens )
    assert np.abs( m - r[ "masses" ] ).max() <= 1e-15
    assert np.abs( m - m.sum() / n ).max() <= 1e-11
    hs = r[ "history" ][ :, 7 ]
    assert hs[ 0 ] == s0 and hs[ -1 ] == 0 and np.all( np.diff( hs ) <= 0 )
    # "never" reaches the same plan directly
    q = oracle.solve( pos, tol = 1e-11, density = dens, step = "limits", continuation = "never", lin_tol = 1e-12 )
    assert q[ "stats" ][ "status" ] == 1
    assert np.abs( q[ "weights" ] - r[ "weights" ] ).max() < 1e-7 * np.abs( q[ "weights" ] ).max()


# ---- a sum of Gaussians ( reference tests/test_SumOfGaussians.py:20-73 ) ------------------------------------------------


def _gauss( rows ):
    return dict( gaussians = np.asarray( rows, dtype = float ) )


def _rho_gauss( g ):
    def rho( x, y ):
        return sum( w * np.exp( -( ( x - cx ) ** 2 + ( y - cy ) ** 2 ) / ( 2 * s * s ) ) / ( 2 * np.pi * s * s ) for cx, cy, s, w in g )
    return rho


def test_gaussian_mass_of_a_box_is_a_product_of_erfs( oracle ):
    from scipy.special import erf
    g = np.array( [ [ 0.2, 0.7, 0.05, 2.0 ], [ 0.8, 0.3, 0.2, 0.5 ], [ 0.5, 0.5, 0.1, 1.0 ] ] )
    box = [ sum( w / 4 * ( erf( ( 1 - cx ) / ( np.sqrt( 2 ) * s ) ) + erf( cx / ( np.sqrt( 2 ) * s ) ) )
                 * ( erf( ( 1 - cy ) / ( np.sqrt( 2 ) * s ) ) + erf( cy / ( np.sqrt( 2 ) * s ) ) ) for cx, cy, s, w in g ) ]
    one = oracle.measures( np.array( [ [ 0.3, 0.6 ] ] ), density = _gauss( g ) )
    assert abs( one[ 0 ] - box[ 0 ] ) <= 1e-14 * box[ 0 ]
    # any diagram shares it out exactly
    rng = np.random.default_rng( 30 )
    pos = rng.uniform( 0, 1, ( 300, 2 ) )
    m = oracle.measures( pos, rng.normal( 0, 1e-4, 300 ), density = _gauss( g ) )
    assert np.all( m >= 0 ) and abs( m.sum() - box[ 0 ] ) <= 1e-13 * box[ 0 ]


def test_gaussian_cells_and_facets_against_quadrature( oracle ):
    from scipy import integrate
    g = np.array( [ [ 0.2, 0.7, 0.05, 2.0 ], [ 0.8, 0.3, 0.2, 0.5 ] ] )
    rho = _rho_gauss( g )
    rng = np.random.default_rng( 31 )
    pos = rng.uniform( 0, 1, ( 40, 2 ) )
    w = rng.normal( 0, 1e-3, 40 )
    a, ( fi, fj, fc ) = oracle.measures( pos, w, density = _gauss( g ), facets = True )
    for i in ( 0, 7, 19 ):
        xy, ids = oracle.cell( pos, i, w )
        ref = 0.0
        for v in range( 1, len( xy ) - 1 ):
            A, B, C = xy[ 0 ], xy[ v ], xy[ v + 1 ]
            jac = abs( ( B - A )[ 0 ] * ( C - A )[ 1 ] - ( B - A )[ 1 ] * ( C - A )[ 0 ] )
            ref += integrate.dblquad( lambda t, u: rho( *( A + u * ( B - A ) + t * ( C - A ) ) ) * jac, 0, 1, 0, lambda u: 1 - u,
                                      epsabs = 1e-15, epsrel = 1e-12 )[ 0 ]
        assert abs( a[ i ] - ref ) <= 1e-11 * ref, i
        # the facets of the cell: int rho ds / ( 2 |p_i - p_j| )
        for q in np.where( fi == i )[ 0 ]:
            j = fj[ q ]
            e = list( ids ).index( j )
            P, Q = xy[ e ], xy[ ( e + 1 ) % len( xy ) ]
            L = np.linalg.norm( Q - P )
            line = integrate.quad( lambda t: rho( *( P + t * ( Q - P ) ) ) * L, 0, 1, epsabs = 1e-16, epsrel = 1e-13 )[ 0 ]
            assert abs( fc[ q ] - line / ( 2 * np.linalg.norm( pos[ j ] - pos[ i ] ) ) ) <= 1e-12 * fc[ q ] + 1e-300


def test_gaussian_moments_against_quadrature( oracle ):
    # barycentres and cost from the boundary fluxes against a direct quadrature of the converged cells
    from scipy import integrate
    g = np.array( [ [ 0.35, 0.6, 0.15, 1.0 ], [ 0.7, 0.4, 0.2, 1.0 ] ] )
    rho = _rho_gauss( g )
    rng = np.random.default_rng( 32 )
    pos = rng.uniform( 0, 1, ( 60, 2 ) )
    r = oracle.solve( pos, density = _gauss( g ), tol = 1e-12, step = "limits", continuation = "auto", lin_tol = 1e-13 )
    assert r[ "stats" ][ "status" ] == 1
    cost = 0.0
    for i in range( len( pos ) ):
        xy, _ = oracle.cell( pos, i, r[ "weights" ] )
        m = mx = my = c2 = 0.0
        for v in range( 1, len( xy ) - 1 ):
            A, B, C = xy[ 0 ], xy[ v ], xy[ v + 1 ]
            jac = abs( ( B - A )[ 0 ] * ( C - A )[ 1 ] - ( B - A )[ 1 ] * ( C - A )[ 0 ] )
            def integral( f ):
                return integrate.dblquad( lambda t, u: f( *( A + u * ( B - A ) + t * ( C - A ) ) ) * jac, 0, 1, 0, lambda u: 1 - u,
                                          epsabs = 1e-15, epsrel = 1e-11 )[ 0 ]
            m += integral( rho )
            mx += integral( lambda x, y: x * rho( x, y ) )
            my += integral( lambda x, y: y * rho( x, y ) )
            c2 += integral( lambda x, y: ( ( x - pos[ i, 0 ] ) ** 2 + ( y - pos[ i, 1 ] ) ** 2 ) * rho( x, y ) )
        cost += c2
        if i % 7 == 0:
            assert np.abs( r[ "barycenters" ][ i ] - [ mx / m, my / m ] ).max() <= 1e-10, i
    assert abs( r[ "cost" ] - cost ) <= 1e-10 * cost


def test_gaussian_continuation_widens_the_gaussians( oracle ):
    # the stages of a sum of Gaussians stop at a quarter of the narrowest ( convolved/SumOfGaussians.h:39 ), and the plan is
    # the one the direct Newton reaches
    g = np.array( [ [ 0.3, 0.3, 0.06, 1.0 ], [ 0.7, 0.65, 0.1, 1.5 ] ] )
    rng = np.random.default_rng( 33 )
    pos = rng.uniform( 0, 1, ( 800, 2 ) )
    r = oracle.solve( pos, density = _gauss( g ), tol = 1e-11, step = "limits", continuation = "always", lin_tol = 1e-12 )
    st = r[ "stats" ]
    s0, smin = 0.5 * np.sqrt( 2 ), 0.06 / 4
    assert st[ "status" ] == 1
    assert st[ "nb_continuation_steps" ] == int( np.floor( np.log( s0 / smin ) / np.log( np.sqrt( 2 ) ) + 1e-9 ) ) + 2
    q = oracle.solve( pos, density = _gauss( g ), tol = 1e-11, step = "limits", continuation = "never", lin_tol = 1e-12, maxit = 200 )
    assert q[ "stats" ][ "status" ] == 1
    assert np.abs( q[ "weights" ] - r[ "weights" ] ).max() < 1e-7 * np.abs( q[ "weights" ] ).max()


# ---- exact duplicates ( reference sdotplan/Aggregation.h:107-220, tests/test_SdotPlanNd.py:1377-1470 ) --------------------


@pytest.mark.parametrize( "step", [ "trials", "limits" ] )
@pytest.mark.parametrize( "name", H.EXACT_CLOUDS )
def test_exact_duplicates_are_merged( oracle, name, step ):
    # each duplicate's weight tied under its representative's, its target carried by it: converged, the clusters' masses
    # their targets ( reference tests/test_SdotPlanNd.py:1436-1452 )
    pos, groups = H.degenerate_clouds()[ name ]
    n = len( pos )
    tol = 1e-11 / n
    dups = H.brute_duplicates( pos )
    r = oracle.solve( pos, tol = tol, maxit = 60, step = step, dups = dups, gap = H.duplicate_gap( pos ), lin_tol = 1e-13 )
    nu = np.full( n, 1.0 / n )
    H.check_aggregated( f"{ name } { step }", r[ "stats" ], r[ "weights" ], r[ "masses" ], nu, r[ "clusters" ], groups, tol )
    assert r[ "stats" ][ "nb_duplicates" ] == sum( len( g ) - 1 for g, _ in groups )
    # the k-th duplicate of a representative sits `k gap` under it ( Aggregation.h:111-117 )
    gap = H.duplicate_gap( pos )
    for members, _ in groups:
        for k, m in enumerate( members[ 1: ], 1 ):
            assert r[ "weights" ][ m ] == r[ "weights" ][ members[ 0 ] ] - k * gap
    assert r[ "stats" ][ "residual_full" ] <= 2 * tol


def test_exact_duplicates_solve_the_merged_problem( oracle ):
    # the representatives' weights are those of the problem without the duplicates, whose targets their representatives
    # carry: the same cells, the same equations ( an invariant of the merge, independent of how it is implemented )
    pos, groups = H.degenerate_clouds()[ "exact triple" ]
    n = len( pos )
    rng = np.random.default_rng( 5 )
    nu = rng.uniform( 0.5, 1.5, n )
    nu /= nu.sum()
    dups = H.brute_duplicates( pos )
    r = oracle.solve( pos, nu = nu, tol = 1e-14, maxit = 60, dups = dups, gap = H.duplicate_gap( pos ), lin_tol = 1e-13 )
    keep = np.setdiff1d( np.arange( n ), dups[ :, 0 ] )
    nu_m = nu.copy()
    for d, p in dups:
        nu_m[ p ] += nu[ d ]
    q = oracle.solve( pos[ keep ], nu = nu_m[ keep ], tol = 1e-14, maxit = 60, lin_tol = 1e-13 )
    assert r[ "stats" ][ "status" ] == 1 and q[ "stats" ][ "status" ] == 1
    assert np.abs( r[ "weights" ][ keep ] - q[ "weights" ] ).max() < 1e-9
    assert np.abs( r[ "masses" ][ keep ] - q[ "masses" ] ).max() < 1e-13
    assert ( r[ "masses" ][ dups[ :, 0 ] ] == 0 ).all()
    # the cost is blind to the duplicates: each of them holds nothing
    assert abs( r[ "cost" ] - q[ "cost" ] ) < 1e-9 * q[ "cost" ]


def test_without_the_merge_the_duplicates_cannot_converge( oracle ):
    # two equal points cannot split a cell: without the merge, no weight vector reaches the targets
    # ( reference tests/test_SdotPlanNd.py:1449-1452 )
    pos, _ = H.degenerate_clouds()[ "exact pair" ]
    r = oracle.solve( pos, tol = 1e-11 / len( pos ), maxit = 30 )
    assert r[ "stats" ][ "status" ] != 1


# ---- the aggregation of near-coincident seeds ( reference sdotplan/Aggregation.h, tests/test_SdotPlanNd.py:1436-1452 ) ------


@pytest.mark.parametrize( "step", [ "trials", "limits" ] )
@pytest.mark.parametrize( "name", H.NEAR_CLOUDS )
def test_near_coincident_seeds_are_aggregated( oracle, name, step ):
    # pairs the doubles cannot separate are merged ( margin 4, the reference's default ): converged ( aggregated ), the
    # clusters found, each cluster's mass its target, every other cell its own
    pos, groups = H.degenerate_clouds()[ name ]
    n = len( pos )
    tol = 1e-11 / n
    dups = H.brute_duplicates( pos )
    r = oracle.solve( pos, tol = tol, maxit = 60, step = step, dups = dups if len( dups ) else None, gap = H.duplicate_gap( pos ),
                      margin = 4.0, lin_tol = 1e-14, lin_maxit = 200000 )
    H.check_aggregated( f"{ name } { step }", r[ "stats" ], r[ "weights" ], r[ "masses" ], np.full( n, 1.0 / n ), r[ "clusters" ], groups, tol )
    # the full problem's residual is the floor of the doubles: above the tolerance, far below the cells' masses
    assert tol < r[ "stats" ][ "residual_full" ] < 1e-3 / n


def test_without_the_aggregation_near_pairs_stagnate( oracle ):
    # ( reference tests/test_SdotPlanNd.py:1449-1452: "without the aggregation: the floor" )
    pos, _ = H.degenerate_clouds()[ "pair at 1e-12" ]
    n = len( pos )
    r = oracle.solve( pos, tol = 1e-11 / n, maxit = 60, lin_tol = 1e-14, lin_maxit = 200000 )
    assert r[ "stats" ][ "status" ] != 1 and r[ "stats" ][ "nb_clusters" ] == 0
