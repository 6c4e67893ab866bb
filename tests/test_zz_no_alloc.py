"""The entry point that allocates nothing ( `sdot_amd_measures_2d_noalloc` ): its masses are those of `sdot_amd_measures_2d`
to the bit, on two calls in a row ( the second smaller, reusing the same fixed scratch ), and a cloud past its limit is refused
before any device work."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say( *a ):
    print( *a, flush = True )
    sys.stdout.flush()


def _bumps( res, rng ):
    xs = ( np.arange( res ) + 0.5 ) / res
    X, Y = np.meshgrid( xs, xs, indexing = "ij" )
    v = np.full( ( res, res ), 0.1 )
    for c, s in zip( rng.uniform( 0, 1, ( 4, 2 ) ), rng.uniform( 0.03, 0.15, 4 ) ):
        v += np.exp( -( ( X - c[ 0 ] ) ** 2 + ( Y - c[ 1 ] ) ** 2 ) / ( 2 * s ** 2 ) )
    return v


def test_no_alloc_masses_equal_the_allocating_entry_points( native ):
    from sdot.distributions import Image, SumOfGaussians, UniformBox
    import torch
    rng = np.random.default_rng( 77 )
    nmax = native.lib().sdot_amd_measures_2d_noalloc_max_seeds()
    assert nmax == 1 << 17
    image = Image( _bumps( 256, rng ), origin = [ 0, 0 ], frame = np.eye( 2 ) / 256 ).normalized_version()
    gauss = SumOfGaussians( [ [ 0.3, 0.4 ], [ 0.7, 0.6 ] ], [ 0.1, 0.2 ], [ 0.3, 0.7 ], support_box = ( [ 0, 0 ], [ 1, 1 ] ) ).normalized_version()
    cases = [ ( "image, the largest cloud", nmax, image ), ( "constant, smaller", 5_000, UniformBox( [ 0, 0 ], [ 1, 1 ] ) ),
              ( "gaussians, smaller still", 1_500, gauss ) ]
    for label, n, dens in cases:
        say( f"no-alloc: { label }: n = { n }" )
        pos = rng.uniform( 0, 1, ( n, 2 ) )
        w = rng.uniform( -1e-2 / n, 1e-2 / n, n )
        ref = native.measures_2d( pos, w, [ 0, 0, 1, 1 ], dens ).cpu().numpy()
        before = native.device_bytes()
        got = native.measures_2d_noalloc( pos, w, [ 0, 0, 1, 1 ], dens ).cpu().numpy()
        assert native.device_bytes() == before           # ( the library's own arena did not grow )
        assert np.array_equal( got, ref ), ( label, float( np.abs( got - ref ).max() ) )
        # a Voronoi diagram ( null weights: zeroed in the scratch by a kernel, whatever the last call left there )
        ref0 = native.measures_2d( pos, None, [ 0, 0, 1, 1 ], dens ).cpu().numpy()
        got0 = native.measures_2d_noalloc( pos, None, [ 0, 0, 1, 1 ], dens ).cpu().numpy()
        assert np.array_equal( got0, ref0 ), ( label, float( np.abs( got0 - ref0 ).max() ) )
        torch.cuda.synchronize()


def test_no_alloc_refuses_a_cloud_past_its_limit( native ):
    from sdot._native import SdotError
    from sdot.distributions import UniformBox
    n = native.lib().sdot_amd_measures_2d_noalloc_max_seeds() + 1
    pos = np.random.default_rng( 3 ).uniform( 0, 1, ( n, 2 ) )
    with pytest.raises( SdotError ) as e:
        native.measures_2d_noalloc( pos, None, [ 0, 0, 1, 1 ], UniformBox( [ 0, 0 ], [ 1, 1 ] ) )
    assert e.value.kind == 1 and "at most" in str( e.value )
