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
