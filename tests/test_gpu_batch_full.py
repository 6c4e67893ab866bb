"""The batched and served configurations at their full size, on the GPU, against the oracle.

* the batch: 256 independent 2D instances, Dirac counts log-uniform in 1k-200k ( default_rng(5) ), each with its own 256^2
  image of 4 Gaussian bumps + 0.1 floor, mixed precision ( the FP32 kernel, FP64 from the first step it cannot take ): every
  plan converges; the oracle's FP64 cells of the returned weights pass the tolerance on a sample of instances; the oracle's own
  solve of the three smallest gives the same weights; an instance solved alone gives the same plan to the bit;
* the server: a stream of 2D requests ( 128^2 images ) and 3D ones ( 32^3 voxel images ), each plan the one its solve gets
  alone, and a sample of both passing the tolerance on the oracle's FP64 cells.
"""

import os
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname( os.path.dirname( os.path.abspath( __file__ ) ) )


def say( *a ):
    print( *a, flush = True )
    sys.stdout.flush()


def _bench():
    sys.path.insert( 0, ROOT )
    import bench
    return bench


def _oracle_image( v, res ):
    return dict( values = np.ascontiguousarray( v.T ), x0 = 0.0, y0 = 0.0, hx = 1.0 / res, hy = 1.0 / res )


def test_the_batched_workload_at_full_size( native, oracle ):
    import torch
    import sdot
    bench = _bench()
    res = 256
    inst = bench.workload_batch( 256, res )
    total = sum( len( p ) for p, _ in inst )
    say( f"case batch: 256 instances, { total } Diracs, mixed" )
    dev = torch.device( "cuda", 0 )
    problems = [ sdot.OtProblem( sdot.SumOfDiracs( torch.as_tensor( p, device = dev ) ), sdot.Image( v, origin = [ 0, 0 ], frame = np.eye( 2 ) / res ) )
                 for p, v in inst ]
    tol = 1e-9
    settings = sdot.Iterative( tol = tol, continuation = "never", on_failure = "ignore", precision = "mixed" )
    plans = sdot.solve_batch( problems, settings, streams = 8 )
    assert all( p.converged for p in plans ), [ ( q, p.stats[ "status" ] ) for q, p in enumerate( plans ) if not p.converged ]
    sizes = np.array( [ len( p ) for p, _ in inst ] )
    order = np.argsort( sizes )
    # the oracle's FP64 cells of the returned weights: the tolerance, on the smallest, a middle and a large instance
    for q in ( order[ 0 ], order[ 1 ], order[ 128 ], order[ 200 ] ):
        pos, v = inst[ q ]
        img = problems[ q ].target
        w = plans[ q ].weights.cpu().numpy()
        mo = oracle.measures( pos, w, density = _oracle_image( np.asarray( img.values ), res ) )
        target = np.full( len( pos ), 1.0 / len( pos ) ) * ( mo.sum() )
        say( f"  instance { q }: n = { len( pos ) }, oracle's FP64 cells: max |m - nu| = { np.abs( mo - target ).max():.2e}" )
        assert np.abs( mo - target ).max() <= 1.01 * tol
    # the oracle's own solves of the three smallest: the same plan ( the weights to 1e-6 of their spread )
    for q in order[ :3 ]:
        pos, v = inst[ q ]
        ro = oracle.solve( pos, density = _oracle_image( np.asarray( problems[ q ].target.values ), res ), tol = tol, step = "limits" )
        wo = ro[ "weights" ]
        w = plans[ q ].weights.cpu().numpy()
        assert np.abs( w - wo ).max() <= 1e-5 * ( wo.max() - wo.min() ), q
    # alone, an instance gets the same plan to the bit
    for q in ( order[ 5 ], order[ 250 ] ):
        alone = problems[ q ].solve( settings )
        assert torch.equal( alone.weights, plans[ q ].weights ), q


def test_served_requests_match_their_solo_runs_and_the_oracle( native, oracle ):
    import torch
    import sdot
    bench = _bench()
    reqs = bench.workload_serve( 12, 50.0, n_min = 1_000, n_max = 30_000, res = 128, res3 = 32 )
    dev = torch.device( "cuda", 0 )
    tol = 1e-9
    s2 = sdot.Iterative( tol = tol, continuation = "never", on_failure = "ignore", precision = "mixed" )
    s3 = sdot.Iterative( tol = tol, continuation = "auto", on_failure = "ignore" )
    problems = []
    for _, d, pos, dens in reqs:
        p = torch.as_tensor( pos, device = dev )
        if d == 2:
            problems.append( sdot.OtProblem( sdot.SumOfDiracs( p ), sdot.Image( dens, origin = [ 0, 0 ], frame = np.eye( 2 ) / 128 ) ) )
        else:
            problems.append( sdot.OtProblem( sdot.SumOfDiracs( p ), sdot.Image( dens, origin = [ 0, 0, 0 ], frame = np.eye( 3 ) / 32 ) ) )
    say( f"case server: { len( reqs ) } requests ( { sum( d == 3 for _, d, _, _ in reqs ) } 3D )" )
    with sdot.Server( slots = 4 ) as srv:
        tickets = [ srv.submit( pb, s2 if r[ 1 ] == 2 else s3 ) for pb, r in zip( problems, reqs ) ]
        plans = [ t.result() for t in tickets ]
    for q, ( pb, r, pl ) in enumerate( zip( problems, reqs, plans ) ):
        assert pl.converged, ( q, pl.stats[ "status" ] )
        alone = pb.solve( s2 if r[ 1 ] == 2 else s3 )
        assert torch.equal( alone.weights, pl.weights ), q
        if r[ 1 ] == 2 and q % 2 == 0:
            pos = r[ 2 ]
            mo = oracle.measures( pos, pl.weights.cpu().numpy(), density = _oracle_image( np.asarray( pb.target.values ), 128 ) )
            assert np.abs( mo - mo.sum() / len( pos ) ).max() <= 1.01 * tol, q
        if r[ 1 ] == 3 and len( r[ 2 ] ) <= 12_000:     # ( the oracle's 3D cells of the served weights )
            v = np.asarray( pb.target.values )
            d = dict( voxels = v, x0 = 0.0, y0 = 0.0, z0 = 0.0, hx = 1 / 32, hy = 1 / 32, hz = 1 / 32 )
            mo = oracle.measures_3d( r[ 2 ], pl.weights.cpu().numpy(), density = d )
            assert np.abs( mo - mo.sum() / len( r[ 2 ] ) ).max() <= 1.01 * tol, q
