"""The host logic of the Python mirror, on the CPU: the exact duplicates' search ( reference src/sdot/SdotPlanNd.py:124-160 )
against a brute force over every pair, and the gap it gives their weights."""

import numpy as np
import pytest

import helpers as H


def _planted( rng, n, k, d = 2 ):
    pos = rng.uniform( -1, 1, ( n, d ) )
    src = rng.integers( 0, n, k )
    dst = rng.integers( 0, n, k )
    pos[ dst ] = pos[ src ]
    return pos


@pytest.mark.parametrize( "seed", range( 6 ) )
def test_exact_duplicates_are_the_brute_force_pairs( seed ):
    from sdot.SdotPlanNd import exact_duplicates
    rng = np.random.default_rng( seed )
    pos = _planted( rng, 300 + 50 * seed, seed * 7 )
    dup, rep = exact_duplicates( pos )
    want = H.brute_duplicates( pos )
    assert np.array_equal( np.stack( [ dup, rep ], axis = 1 ).reshape( -1, 2 ), want )
    assert dup.dtype == np.int64 and rep.dtype == np.int64
    assert ( rep < dup ).all()


def test_exact_duplicates_edge_cases():
    from sdot.SdotPlanNd import exact_duplicates
    # none, one seed, an empty cloud
    for pos in ( np.random.default_rng( 0 ).uniform( 0, 1, ( 1000, 2 ) ), np.zeros( ( 1, 2 ) ), np.zeros( ( 0, 2 ) ) ):
        dup, rep = exact_duplicates( pos )
        assert len( dup ) == 0 and len( rep ) == 0
    # -0.0 and 0.0 are the same position; a one-ulp difference is not
    pos = np.array( [ [ 0.0, 1.0 ], [ -0.0, 1.0 ], [ 0.0, np.nextafter( 1.0, 2.0 ) ], [ 0.5, 0.5 ] ] )
    dup, rep = exact_duplicates( pos )
    assert dup.tolist() == [ 1 ] and rep.tolist() == [ 0 ]
    # a triple and a pair: each duplicate points at the smallest index of its position
    pos = np.array( [ [ 1, 2 ], [ 3, 4 ], [ 1, 2 ], [ 5, 6 ], [ 3, 4 ], [ 1, 2 ] ], dtype = float )
    dup, rep = exact_duplicates( pos )
    assert dup.tolist() == [ 2, 4, 5 ] and rep.tolist() == [ 0, 1, 0 ]
    # a cloud made only of one point
    dup, rep = exact_duplicates( np.ones( ( 5, 2 ) ) )
    assert dup.tolist() == [ 1, 2, 3, 4 ] and rep.tolist() == [ 0, 0, 0, 0 ]


def test_exact_duplicates_in_3d_and_on_lists():
    from sdot.SdotPlanNd import exact_duplicates
    rng = np.random.default_rng( 3 )
    pos = _planted( rng, 500, 20, d = 3 )
    dup, rep = exact_duplicates( pos.tolist() )
    want = H.brute_duplicates( pos )
    assert np.array_equal( np.stack( [ dup, rep ], axis = 1 ).reshape( -1, 2 ), want )


def test_duplicate_gap_is_a_millionth_of_the_squared_scale():
    from sdot.SdotPlanNd import duplicate_gap
    pos = np.array( [ [ 0.0, 0.0 ], [ 3.0, 4.0 ] ] )
    assert duplicate_gap( pos ) == 1e-6 * 25.0
    pos = np.array( [ [ 10.0, 0.0 ], [ 10.0, 1.0 ] ] )          # the norm dominates the extent
    assert duplicate_gap( pos ) == 1e-6 * 101.0
    assert duplicate_gap( np.zeros( ( 3, 2 ) ) ) == 1e-6
    assert H.duplicate_gap( pos ) == duplicate_gap( pos )


def test_card_options_carry_the_duplicates():
    from sdot import Iterative, UniformBox
    from sdot import _native
    from sdot.SdotPlanNd import card_options
    o = card_options( Iterative(), UniformBox( [ 0, 0 ], [ 1, 1 ] ), nb_dups = 3, gap = 2e-6 )
    assert o[ _native.CARD_OPTIONS.index( "agg_nb_dups" ) ] == 3 and o[ _native.CARD_OPTIONS.index( "agg_gap" ) ] == 2e-6


def test_precision_maps_to_the_kernels_float():
    """'auto' and 'mixed' run the FP32 kernel then FP64 ( as the reference's card ), 'fp32' and 'fp64' one kernel; 3D refuses
    'fp32' before any device work"""
    from sdot import Iterative, OtProblem, SumOfDiracs, UniformBox, _native
    from sdot.SdotPlanNd import _KERNEL_FLOAT
    assert _KERNEL_FLOAT == { "auto": _native.MIXED, "mixed": _native.MIXED, "fp32": _native.FP32, "fp64": _native.FP64 }
    rng = np.random.default_rng( 0 )
    pb3 = OtProblem( SumOfDiracs( rng.uniform( 0, 1, ( 50, 3 ) ) ), UniformBox( [ 0, 0, 0 ], [ 1, 1, 1 ] ) )
    with pytest.raises( NotImplementedError ):
        pb3.solve( Iterative( precision = "fp32" ) )


# ---- turned grids ----------------------------------------------------------------------------------------------------

def _turn( th ):
    return np.array( [ [ np.cos( th ), np.sin( th ) ], [ -np.sin( th ), np.cos( th ) ] ] )


def test_a_turned_grid_is_an_axis_aligned_one_in_its_frame():
    from sdot.distributions import Image
    rng = np.random.default_rng( 1 )
    F = np.diag( [ 0.5, 0.25 ] ) @ _turn( 0.3 )
    im = Image( rng.uniform( 0.5, 1.5, ( 4, 3 ) ), origin = [ 1.0, 2.0 ], frame = F )
    assert im.rotation is not None and np.allclose( im.steps, [ 0.5, 0.25 ] )
    x = rng.uniform( 0, 3, ( 50, 2 ) )
    assert np.abs( im.from_local( im.to_local( x ) ) - x ).max() < 1e-15
    # a pixel's centre goes to the axis-aligned one's, distances are kept
    assert np.allclose( im.to_local( np.array( [ [ 1.0, 2.0 ] ] ) + ( F.T @ [ 2.5, 1.5 ] )[ None ] ), [ [ 1.25, 0.375 ] ], atol = 1e-15 )
    y = im.to_local( x )
    assert np.allclose( np.linalg.norm( x[ :, None ] - x[ None ], axis = 2 ), np.linalg.norm( y[ :, None ] - y[ None ], axis = 2 ), atol = 1e-14 )
    # the half-planes hold exactly the grid's corners, the mass is the values times the pixels' area
    d, o = im.bounding_half_spaces()
    for ij in ( ( 0, 0 ), ( 4, 0 ), ( 0, 3 ), ( 4, 3 ) ):
        p = np.array( [ 1.0, 2.0 ] ) + F.T @ np.array( ij, dtype = float )
        assert np.all( d @ p <= o + 1e-12 ) and np.isclose( ( d @ p - o ).max(), 0, atol = 1e-12 )
    assert np.isclose( im.mass, im.values.sum() * 0.125 )
    loc = im.local_version()
    assert loc.rotation is None and np.allclose( loc.box()[ 1 ], [ 2.0, 0.75 ] )


def test_sheared_pixels_are_refused_and_a_turned_grid_crosses_in_its_frame_only():
    from sdot.distributions import Image
    with pytest.raises( NotImplementedError ):
        Image( np.ones( ( 3, 3 ) ), frame = [ [ 1.0, 0.2 ], [ 0.0, 1.0 ] ] )
    im = Image( np.ones( ( 3, 3 ) ), frame = _turn( 0.5 ) )
    with pytest.raises( ValueError ):
        im.native_args( device_tensor = lambda a: a )


def test_3d_duplicates_merge_on_the_host( monkeypatch ):
    """the host side of a 3D cloud's exact duplicates, the device solve replaced by a stand-in that records what it is given:
    the call sees the cloud without its duplicates, each representative carrying its duplicates' targets; the duplicates come
    back `k gap` under their representative, with empty cells, their seed as barycentre and their representative as cluster"""
    import torch
    from sdot import SumOfDiracs, _native
    from sdot.SdotPlanNd import _STATS, _solve_3d_merged, duplicate_gap, exact_duplicates
    rng = np.random.default_rng( 5 )
    base = rng.uniform( 0, 1, ( 50, 3 ) )
    pos = np.concatenate( [ base, base[ [ 7, 3, 7 ] ] ] )              # 7 twice ( k = 1, 2 ), 3 once
    nu = rng.uniform( 1, 2, len( pos ) )
    seen = {}

    def fake_solve_3d( positions, nu_r, w0, opts, box, density ):
        seen.update( positions = positions.numpy().copy(), nu = np.asarray( nu_r ).copy() )
        n = len( nu_r )
        return dict( weights = torch.arange( n, dtype = torch.float64 ), cell_masses = torch.full( ( n, ), 2.0, dtype = torch.float64 ),
                     barycenters = torch.zeros( ( n, 3 ), dtype = torch.float64 ), cost = 1.5,
                     stats = np.zeros( len( _STATS ) ), history = None )

    monkeypatch.setattr( _native, "on_device", lambda a, dtype = None: torch.as_tensor( np.asarray( a ), dtype = dtype or torch.float64 ) )
    monkeypatch.setattr( _native, "solve_3d", fake_solve_3d )
    dup, rep = exact_duplicates( pos )
    assert dup.tolist() == [ 50, 51, 52 ] and rep.tolist() == [ 7, 3, 7 ]
    res = _solve_3d_merged( SumOfDiracs( pos, nu ), dup, rep, None, None, [ 0, 0, 0, 1, 1, 1 ], None )
    assert np.array_equal( seen[ "positions" ], base )
    want = nu[ :50 ].copy()
    want[ 7 ] = ( want[ 7 ] + nu[ 50 ] ) + nu[ 52 ]                    # ( in the duplicates' order )
    want[ 3 ] += nu[ 51 ]
    assert np.array_equal( seen[ "nu" ], want )
    w = res[ "weights" ].numpy()
    gap = duplicate_gap( pos )
    assert np.array_equal( w[ :50 ], np.arange( 50.0 ) )
    assert w[ 50 ] == 7 - gap and w[ 51 ] == 3 - gap and w[ 52 ] == 7 - 2 * gap
    m = res[ "cell_masses" ].numpy()
    assert ( m[ :50 ] == 2 ).all() and ( m[ 50: ] == 0 ).all()
    assert np.array_equal( res[ "barycenters" ].numpy()[ 50: ], pos[ 50: ] )
    assert np.array_equal( res[ "clusters" ].numpy(), np.concatenate( [ np.arange( 50 ), [ 7, 3, 7 ] ] ) )
    st = res[ "stats" ]
    assert st[ _STATS.index( "nb_duplicates" ) ] == 3 and st[ _STATS.index( "nb_clusters" ) ] == 2
    assert st[ _STATS.index( "nb_aggregated" ) ] == 5 and res[ "cost" ] == 1.5
