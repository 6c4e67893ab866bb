"""The FP32 cell kernel ( the walk and the cuts in float, every vertex re-solved in double from its two planes, the rows made
symmetric ) and the mixed solve ( FP32, then FP64 from the first step the FP32 kernel cannot take ), on the GPU, through the
C ABI: the plan such a solve returns passes the tolerance on the ORACLE's FP64 cells of its weights, its weights are the
FP64 solve's, and its Laplacians keep the multigrid converging."""
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def say( *a ):
    print( *a, flush = True )
    sys.stdout.flush()


def _density( kind, rng ):
    """( the GPU density, the oracle's )"""
    from sdot import Image, UniformBox
    if kind == "uniform":
        return UniformBox( [ 0, 0 ], [ 1, 1 ] ), None
    img = H.bumps_image( 128, 128, rng )
    dens = Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 128 ).normalized_version()
    v = dens.values
    return dens, dict( values = np.ascontiguousarray( np.asarray( v ).T ), x0 = 0.0, y0 = 0.0, hx = 1 / 128, hy = 1 / 128 )


def _solve( pos, dens, precision, tol ):
    from sdot import Iterative, OtProblem, SumOfDiracs
    return OtProblem( SumOfDiracs( pos ), dens ).solve( Iterative( tol = tol, continuation = "never", precision = precision,
                                                                     on_failure = "ignore" ) )


@pytest.mark.parametrize( "kind", [ "uniform", "image" ] )
@pytest.mark.parametrize( "precision", [ "fp32", "mixed" ] )
def test_fp32_plan_passes_the_oracles_fp64_cells( native, oracle, kind, precision ):
    rng = np.random.default_rng( 40 + ( kind == "image" ) )
    n = 20000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    dens, dens_oracle = _density( kind, rng )
    tol = 1e-10                                          # ( 2e-6 of a mean cell )
    say( f"case precision { precision } { kind } n={ n }" )
    sol = _solve( pos, dens, precision, tol )
    ref = _solve( pos, dens, "fp64", tol )
    assert sol.converged and ref.converged, ( sol.stats[ "status" ], ref.stats[ "status" ] )
    assert sol.stats[ "it_double" ] == ( 0 if precision == "fp32" else -1 )
    # the FP64 cells of the returned weights ( the oracle's ) have the targets to the tolerance
    w = sol.weights.cpu().numpy()
    mo = oracle.measures( pos, w, density = dens_oracle )
    target = sol.target_masses.cpu().numpy() if hasattr( sol.target_masses, "cpu" ) else np.asarray( sol.target_masses )
    assert np.abs( mo - target ).max() <= 1.01 * tol
    # the masses the FP32 kernel reported are the oracle's cells' ( re-solved vertices ), far below the tolerance
    assert np.abs( sol.cell_masses.cpu().numpy() - mo ).max() <= 1e-3 * tol
    # the plan is the FP64 solve's
    w64 = ref.weights.cpu().numpy()
    assert np.abs( w - w64 ).max() <= 1e-6 * ( w64.max() - w64.min() )
    # the multigrid keeps converging on the FP32 rows ( asymmetric rows broke it down into the Jacobi-PCG fallback )
    assert sol.stats[ "lin_nb_iter" ] <= 1.5 * ref.stats[ "lin_nb_iter" ] + 50


def test_fp32_solve_at_1m( native ):
    """full size, through the properties: converged, and the FP64 cells of its weights pass the tolerance"""
    import sdot
    rng = np.random.default_rng( 42 )
    n = 1_000_000
    pos = rng.uniform( 0, 1, ( n, 2 ) )
    tol = 1e-12
    say( "case precision fp32 uniform 1M" )
    sol = _solve( pos, sdot.UniformBox( [ 0, 0 ], [ 1, 1 ] ), "fp32", tol )
    assert sol.converged
    m = sdot.PowerDiagram( pos, sol.weights, distribution = sdot.UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).measures.cpu().numpy()
    assert np.abs( m - 1.0 / n ).max() <= 1.01 * tol
    assert abs( m.sum() - 1 ) < 1e-10


def test_mixed_switches_to_fp64_where_fp32_stagnates( native ):
    """40 pairs of seeds 1e-7 apart ( no aggregation ) and a 1e-13 tolerance: the FP32 kernel stagnates where the floats
    decide those pairs' cuts; the mixed solve switches to the FP64 kernel there ( `it_double` the iteration ) and goes
    further than the FP32 solve"""
    import torch
    import sdot
    rng = np.random.default_rng( 7 )
    p = rng.uniform( 0.05, 0.95, ( 5000, 2 ) )
    pos = torch.as_tensor( np.concatenate( [ p, p[ :40 ] + 1e-7 * rng.normal( size = ( 40, 2 ) ) ] ), device = "cuda" )
    say( "case mixed switch: 5040 seeds, 40 pairs 1e-7 apart, tol 1e-13" )
    res = {}
    for precision in ( "fp32", "mixed" ):
        st = sdot.Iterative( tol = 1e-13, continuation = "never", on_failure = "ignore", precision = precision, aggregate = False )
        res[ precision ] = sdot.OtProblem( sdot.SumOfDiracs( pos ), sdot.UniformBox( [ 0, 0 ], [ 1, 1 ] ) ).solve( st ).stats
    f, m = res[ "fp32" ], res[ "mixed" ]
    assert f[ "it_double" ] == 0 and f[ "status" ] != "converged"
    # the reference's accounting ( `gpu/Newton2D.cuh:1701-1721` ): the switch uses up the iteration it happens in, which stays
    # counted, and `it_double` is that iteration's index ( one stage here, so `nb_iter` counts it and those after it )
    assert 0 < m[ "it_double" ] and m[ "it_double" ] + 1 <= m[ "nb_iter" ]
    assert m[ "residual" ] < f[ "residual" ]


@pytest.mark.parametrize( "label", [ "uniform", "lines", "lines weighted" ] )
def test_fp32_neighbour_sets_equal_the_lifted_hull( native, label ):
    """the FP32 kernel's facets ( float walk, vertices re-solved in double, one-sided facets redone in FP64 ): the same neighbour
    sets as the lifted hull on generic inputs, up to facets shorter than 1e-10, and its masses within 1e-12 of the FP64 kernel's"""
    rng = np.random.default_rng( 5 )
    pos, w = next( ( p, ww ) for lab, p, ww in H.clouds( rng, 3000 ) if lab == label )
    say( f"case fp32 neighbours: { label }" )
    native.kernel_times( reset = True )
    from sdot.distributions import UniformBox
    m32, ( fi, fj, _ ) = native.measures_2d( pos, w, [ 0, 0, 1, 1 ], UniformBox( [ 0, 0 ], [ 1, 1 ] ), facets = True, precision = native.FP32 )
    m64 = native.measures_2d( pos, w, [ 0, 0, 1, 1 ], UniformBox( [ 0, 0 ], [ 1, 1 ] ) )
    counts = native.fp32_counts()
    say( f"  fp32 repairs: { counts }" )
    assert counts[ "diagrams" ] == 1 and counts[ "dropped" ] == 0
    ours = { ( min( i, j ), max( i, j ) ) for i, j in zip( fi.cpu().tolist(), fj.cpu().tolist() ) }
    hull, amb = H.lifted_neighbours( pos, w, ambiguous = True )
    assert hull <= ours <= hull | amb, ( len( hull - ours ), len( ours - hull - amb ) )
    assert float( ( m32 - m64 ).abs().max() ) <= 1e-12 * float( m64.abs().max() )


def test_fp32_topology_repairs_at_1m_on_the_bench_image( native ):
    """a diagram of the bench's size ( 1M seeds, 1024^2 image of 8 bumps ) at weights of a cell's order by the FP32 kernel: no
    facet left one-sided, masses within 1e-6 of a mean cell of the FP64 kernel's ( 1e-3 of the bench's tolerance )"""
    import torch
    from sdot import Image
    rng = np.random.default_rng( 1 )
    n = 1_000_000
    pos = torch.as_tensor( rng.uniform( 0, 1, ( n, 2 ) ), device = "cuda" )
    img = H.bumps_image( 1024, 1024, np.random.default_rng( 2 ), k = 8 )
    dens = Image( img, origin = [ 0, 0 ], frame = np.eye( 2 ) / 1024 ).normalized_version()
    say( "case fp32 repairs: 1M, 1024^2 image" )
    w = torch.as_tensor( rng.uniform( -0.3, 0.3, n ) / n, device = "cuda" )
    native.kernel_times( reset = True )
    m32 = native.measures_2d( pos, w, [ 0, 0, 1, 1 ], dens, precision = native.FP32 )
    m64 = native.measures_2d( pos, w, [ 0, 0, 1, 1 ], dens )
    counts = native.fp32_counts()
    say( f"  fp32 repairs: { counts }" )
    assert counts[ "dropped" ] == 0
    assert float( ( m32 - m64 ).abs().max() ) <= 1e-6 / n
