"""Throughput of the semi-discrete OT solve on one MI355X: Diracs solved per second to a max mass error
of 1e-9.

Workload ( one "step" = one complete solve, cold start, tree build included, inputs resident in HBM ):
1 000 000 Diracs uniform in [0,1]^2 ( numpy default_rng(1) ), masses 1/n, target a 1024 x 1024
piecewise-constant image density = 0.1 + a sum of 8 isotropic Gaussian bumps ( centres uniform in
[0,1]^2, sigma uniform in [0.03, 0.15], default_rng(2) ), normalized; FP64 cells, damped Newton,
stop at max_i |m_i - nu_i| <= 1e-9.

Prints ONE JSON line ( see the keys below ), with the roofline of the dominant kernel measured live
with HIP events on the library's stream, and a CPU baseline: the oracle's solve of a 100 000-Dirac
sample of the same workload on the host cores.

`--config gauss3d` measures the 3D workload instead ( not the default line ): 1 000 000 Diracs uniform in
[0,1]^3 ( default_rng(3) ), target 4 isotropic Gaussians ( means uniform in [0.2,0.8]^3, sigma 0.1, masses
Dirichlet(1), default_rng(4) ) truncated to the unit cube, width continuation 'auto', multigrid flexible-CG directions,
stop at max_i |m_i - nu_i| <= 1e-9; its CPU baseline is the oracle's 3D solve of a 20 000-Dirac sample.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname( os.path.abspath( __file__ ) )
sys.path.insert( 0, os.path.join( ROOT, "sdot-ffi_amd" ) )

METRIC = "Diracs/s to 1e-9 mass err (2D/3D 1M) vs plain GPU path; served p50/p99 latency"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# algorithmic flops of the cell kernel's events ( 2D ): a plane / vertex side test, one vertex of a pruning test,
# a new vertex ( intersection )
FLOP_SIDE, FLOP_PRUNE, FLOP_NEWV = 4, 11, 10
# the same in 3D
FLOP3_SIDE, FLOP3_PRUNE, FLOP3_NEWV = 6, 16, 40


def workload( n = 1_000_000, res = 1024 ):
    pos = np.random.default_rng( 1 ).uniform( 0, 1, ( n, 2 ) )
    rng = np.random.default_rng( 2 )
    c = rng.uniform( 0, 1, ( 8, 2 ) )
    s = rng.uniform( 0.03, 0.15, 8 )
    xs = ( np.arange( res ) + 0.5 ) / res
    X, Y = np.meshgrid( xs, xs, indexing = "ij" )
    v = np.full( ( res, res ), 0.1 )
    for k in range( 8 ):
        v += np.exp( -( ( X - c[ k, 0 ] ) ** 2 + ( Y - c[ k, 1 ] ) ** 2 ) / ( 2 * s[ k ] ** 2 ) )
    return pos, v


def bump_image( res, centres, sigmas, floor = 0.1 ):
    """`floor + sum_k exp( -|x - c_k|^2 / 2 s_k^2 )` at the centres of a `res x res` grid of the unit square, first index along x"""
    xs = ( np.arange( res ) + 0.5 ) / res
    X, Y = np.meshgrid( xs, xs, indexing = "ij" )
    v = np.full( ( res, res ), floor )
    for c, s in zip( centres, sigmas ):
        v += np.exp( -( ( X - c[ 0 ] ) ** 2 + ( Y - c[ 1 ] ) ** 2 ) / ( 2 * s ** 2 ) )
    return v


def bump_volume( res, centres, sigmas, floor = 0.1 ):
    """`bump_image` in 3D: a `res^3` voxel image of the unit cube, first index along x"""
    xs = ( np.arange( res ) + 0.5 ) / res
    X, Y, Z = np.meshgrid( xs, xs, xs, indexing = "ij" )
    v = np.full( ( res, res, res ), floor )
    for c, s in zip( centres, sigmas ):
        v += np.exp( -( ( X - c[ 0 ] ) ** 2 + ( Y - c[ 1 ] ) ** 2 + ( Z - c[ 2 ] ) ** 2 ) / ( 2 * s ** 2 ) )
    return v


def workload_batch( k = 256, res = 256, n_min = 1_000, n_max = 200_000 ):
    """THE BATCHED WORKLOAD: `k` independent 2D instances, Dirac counts log-uniform in [ n_min, n_max ] ( default_rng(5) ), each
    uniform in [0,1]^2 with its own `res^2` image of 4 Gaussian bumps ( sigma uniform in [ 0.03, 0.15 ] ) + a 0.1 floor
    ( default_rng( [ 5, q ] ) for instance q )"""
    sizes = np.round( np.exp( np.random.default_rng( 5 ).uniform( np.log( n_min ), np.log( n_max ), k ) ) ).astype( np.int64 )
    out = []
    for q, n in enumerate( sizes ):
        r = np.random.default_rng( [ 5, q ] )
        pos = r.uniform( 0, 1, ( int( n ), 2 ) )
        c = r.uniform( 0, 1, ( 4, 2 ) )
        s = r.uniform( 0.03, 0.15, 4 )
        out.append( ( pos, bump_image( res, c, s ) ) )
    return out


def bench_batch( args ):
    """the batched workload: one JSON line, `value` the Diracs solved per second by the batched solver, `vs_baseline` its
    ratio to the plain FP64 path ( one instance per call, one after the other, on the same instances )"""
    import torch
    import sdot
    dev = torch.device( "cuda", 0 )
    torch.cuda.set_device( dev )
    inst = workload_batch( args.batch_k, args.batch_res )
    problems = [ sdot.OtProblem( sdot.SumOfDiracs( torch.as_tensor( p, device = dev ) ),
                                 sdot.Image( v, origin = [ 0, 0 ], frame = np.eye( 2 ) / args.batch_res ) ) for p, v in inst ]
    total = int( sum( len( p ) for p, _ in inst ) )
    print( f"bench batch: { len( problems ) } instances, { total } Diracs", file = sys.stderr, flush = True )
    settings = sdot.Iterative( tol = args.tol, continuation = args.continuation, on_failure = "ignore", precision = args.precision,
                               tuning = sdot.Tuning( step = args.step, linear_solver = args.linear ) )
    plain = sdot.Iterative( tol = args.tol, continuation = args.continuation, on_failure = "ignore", precision = "fp64",
                            tuning = sdot.Tuning( step = args.step, linear_solver = args.linear ) )

    def batched():
        return sdot.solve_batch( problems, settings, streams = args.streams )

    for w in range( args.warmup ):
        batched()
        print( f"bench batch: warmup { w }", file = sys.stderr, flush = True )
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range( args.steps ):
        plans = batched()
        print( f"bench batch: step { k }", file = sys.stderr, flush = True )
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    conv = sum( p.converged for p in plans )
    worst = max( p.stats[ "residual" ] for p in plans )
    # the plain path: one solve per call, FP64 cells, one instance after the other
    t_plain = None
    if not args.no_plain:
        problems[ 0 ].solve( plain )
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pl = [ pb.solve( plain ) for pb in problems ]
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t1
        print( f"bench batch: plain FP64 path { t_plain:.2f} s, { sum( p.converged for p in pl ) } converged", file = sys.stderr, flush = True )
    value = total * args.steps / elapsed
    line = dict( metric = METRIC, value = value, unit = "Diracs/s", n_gpus = 1, steps = args.steps, warmup = args.warmup,
                 ms_per_step = 1e3 * elapsed / args.steps, higher_is_better = True, scaling = "weak",
                 vs_baseline = None if t_plain is None else value / ( total / t_plain ), dtype = "f64" if args.precision == "fp64" else "f32+f64",
                 data = "synthetic",
                 config = dict( workload = f"{ len( problems ) } independent 2D instances, Dirac counts log-uniform in 1k-200k ( { total } in all ), "
                                           f"each with its own { args.batch_res }^2 image of 4 Gaussian bumps + 0.1 floor, max mass error <= { args.tol:g} "
                                           f"( absolute ) per instance",
                                instances = len( problems ), diracs = total, tol_abs = args.tol, step = args.step, continuation = args.continuation,
                                precision = args.precision, streams = args.streams ),
                 batch = dict( converged = conv, worst_residual = worst,
                               plain_fp64_s = t_plain, plain_fp64_diracs_per_s = None if t_plain is None else total / t_plain ) )
    print( json.dumps( line ), flush = True )


def workload_serve( k = 40, rate = 1.0, n_min = 1_000, n_max = 500_000, res = 256, res3 = 64 ):
    """THE SERVED WORKLOAD: `k` requests, Poisson arrivals at `rate` per second ( exponential gaps, default_rng(6) ), 70 % 2D /
    30 % 3D, Dirac counts log-uniform in [ n_min, n_max ]; a 2D request is uniform in [0,1]^2 with a `res^2` image of 4 Gaussian
    bumps + 0.1 floor, a 3D one uniform in [0,1]^3 with a `res3^3` voxel image of 4 bumps ( sigma uniform in [ 0.05, 0.15 ],
    centres uniform in [0.2,0.8]^3 ) + 0.1 floor ( default_rng( [ 6, q ] ) for request q ). Returns `[ ( arrival, dim,
    positions, image ) ]`"""
    rng = np.random.default_rng( 6 )
    gaps = rng.exponential( 1.0 / rate, k )
    arrivals = np.cumsum( gaps ) - gaps[ 0 ]
    dims = np.where( rng.uniform( 0, 1, k ) < 0.7, 2, 3 )
    sizes = np.round( np.exp( rng.uniform( np.log( n_min ), np.log( n_max ), k ) ) ).astype( np.int64 )
    out = []
    for q in range( k ):
        r = np.random.default_rng( [ 6, q ] )
        n, d = int( sizes[ q ] ), int( dims[ q ] )
        pos = r.uniform( 0, 1, ( n, d ) )
        if d == 2:
            dens = bump_image( res, r.uniform( 0, 1, ( 4, 2 ) ), r.uniform( 0.03, 0.15, 4 ) )
        else:
            dens = bump_volume( res3, r.uniform( 0.2, 0.8, ( 4, 3 ) ), r.uniform( 0.05, 0.15, 4 ) )
        out.append( ( float( arrivals[ q ] ), d, pos, dens ) )
    return out


def bench_serve( args ):
    """the served workload: requests submitted at their arrival times to a `Server`, one JSON line: `value` the Diracs solved
    per second over the stream ( first arrival to last completion ), the solves per second and the p50 / p99 latency ( completion
    minus arrival ) beside it"""
    import torch
    import sdot
    dev = torch.device( "cuda", 0 )
    torch.cuda.set_device( dev )
    reqs = workload_serve( args.serve_k, args.serve_rate, res = args.batch_res )
    # the reference's default continuation ( 'auto' ): without it a 2D request whose Voronoi start leaves cells ~300 times
    # under their targets ( 400k Diracs on a 256^2 image of narrow bumps ) creeps for 100 damped steps and stops short
    settings2 = sdot.Iterative( tol = args.tol, continuation = "auto", on_failure = "ignore", precision = args.precision )
    settings3 = sdot.Iterative( tol = args.tol, continuation = "auto", on_failure = "ignore" )

    def problem( d, pos, dens ):
        p = torch.as_tensor( pos, device = dev )
        if d == 2:
            return sdot.OtProblem( sdot.SumOfDiracs( p ), sdot.Image( dens, origin = [ 0, 0 ], frame = np.eye( 2 ) / args.batch_res ) )
        return sdot.OtProblem( sdot.SumOfDiracs( p ), sdot.Image( dens, origin = [ 0, 0, 0 ], frame = np.eye( 3 ) / dens.shape[ 0 ] ) )

    problems = [ problem( d, pos, dens ) for _, d, pos, dens in reqs ]
    total = int( sum( len( pos ) for _, _, pos, _ in reqs ) )
    print( f"bench serve: { len( reqs ) } requests, { total } Diracs, rate { args.serve_rate }/s", file = sys.stderr, flush = True )
    # a warm-up stream of the first few requests ( the slots' buffers and the graphs made once )
    with sdot.Server( slots = args.streams ) as srv:
        for pb, ( _, d, _, _ ) in list( zip( problems, reqs ) )[ :4 ]:
            srv.submit( pb, settings2 if d == 2 else settings3 ).result()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tickets = []
        for q, ( pb, ( at, d, _, _ ) ) in enumerate( zip( problems, reqs ) ):
            wait = at - ( time.perf_counter() - t0 )
            if wait > 0:
                time.sleep( wait )
            tickets.append( srv.submit( pb, settings2 if d == 2 else settings3 ) )
        plans = [ t.result() for t in tickets ]
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        info = srv.info()
    lat = np.array( [ t.times[ 2 ] - t.times[ 0 ] for t in tickets ] )
    queue = np.array( [ t.times[ 1 ] - t.times[ 0 ] for t in tickets ] )
    conv = sum( p.converged for p in plans )
    worst = max( p.stats[ "residual" ] for p in plans )
    line = dict( metric = METRIC, value = total / elapsed, unit = "Diracs/s", n_gpus = 1, steps = 1, warmup = 1,
                 ms_per_step = 1e3 * elapsed, higher_is_better = True, scaling = "weak", vs_baseline = None, dtype = "f32+f64", data = "synthetic",
                 config = dict( workload = f"{ len( reqs ) } served solves, Poisson arrivals at { args.serve_rate }/s, 70 % 2D ( { args.batch_res }^2 "
                                           f"image of 4 bumps + 0.1 floor ) / 30 % 3D ( 64^3 voxel image of 4 bumps + 0.1 floor ), Dirac "
                                           f"counts log-uniform in 1k-500k ( { total } in all ), max mass error <= { args.tol:g} ( absolute )",
                                requests = len( reqs ), diracs = total, rate_per_s = args.serve_rate, slots = args.streams,
                                precision_2d = args.precision, cells_3d = "fp64", continuation = "auto" ),
                 serve = dict( solves_per_s = len( reqs ) / elapsed, p50_latency_s = float( np.percentile( lat, 50 ) ),
                               p99_latency_s = float( np.percentile( lat, 99 ) ), max_latency_s = float( lat.max() ),
                               p50_queue_s = float( np.percentile( queue, 50 ) ), converged = conv, worst_residual = worst,
                               peak_slot_bytes = info[ "peak_bytes" ] ) )
    print( json.dumps( line ), flush = True )


def workload3d( n = 1_000_000 ):
    pos = np.random.default_rng( 3 ).uniform( 0, 1, ( n, 3 ) )
    rng = np.random.default_rng( 4 )
    mu = rng.uniform( 0.2, 0.8, ( 4, 3 ) )
    mass = rng.dirichlet( np.ones( 4 ) )
    return pos, np.column_stack( [ mu, np.full( 4, 0.1 ), mass ] )


def bench_3d( args ):
    """the 3D four-Gaussian workload: one JSON line like the default one ( the roofline: the linear solver's SpMV, timed live )"""
    import torch
    import sdot
    from sdot import _native
    dev = torch.device( "cuda", 0 )
    torch.cuda.set_device( dev )
    pos_h, g = workload3d( args.n )
    pos = torch.as_tensor( pos_h, device = dev )
    target = sdot.SumOfGaussians( g[ :, :3 ], g[ :, 3 ], g[ :, 4 ], support_box = ( [ 0, 0, 0 ], [ 1, 1, 1 ] ) )
    settings = sdot.Iterative( tol = args.tol, continuation = "auto", on_failure = "ignore", tuning = sdot.Tuning( linear_solver = args.linear ) )

    def step():
        return sdot.OtProblem( sdot.SumOfDiracs( pos ), target ).solve( settings )

    for w in range( args.warmup ):
        sol = step()
        print( f"bench 3d: warmup { w }: { sol.stats[ 'status' ] }, { sol.stats[ 'nb_iter' ] } it", file = sys.stderr, flush = True )
    _native.set_timing( True )
    _native.kernel_times( reset = True )
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sols = []
    for k in range( args.steps ):
        sols.append( step() )
        print( f"bench 3d: step { k }: { sols[ -1 ].stats[ 'status' ] }, { sols[ -1 ].stats[ 'nb_iter' ] } it", file = sys.stderr, flush = True )
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    times, spmv_bytes, ev = _native.kernel_times()
    _native.set_timing( False )
    st = sols[ -1 ].stats
    spmv_ms, spmv_launches = times[ "spmv" ]
    spmv_avg_ms = spmv_ms / max( spmv_launches, 1 )
    spmv_gbs = ( spmv_bytes / max( spmv_launches, 1 ) ) / ( spmv_avg_ms * 1e-3 ) / 1e9 if spmv_launches else 0.0
    cells_ms, cells_launches = times[ "cells" ]
    # the cell kernels ( walk and cuts, then the Gaussian measures ) against the FP64 VALU peak: the algorithmic flops of their
    # events ( 3D: 6 a side test, 16 a pruning vertex test, 40 a new vertex ), over every diagram of the timed solves
    flops3 = FLOP3_SIDE * ev[ "side_tests" ] + FLOP3_PRUNE * ev[ "prune_vertex_tests" ] + FLOP3_NEWV * ev[ "new_vertices" ]
    cells_avg_ms = cells_ms / max( cells_launches, 1 )
    cells_tf = ( flops3 / max( cells_launches, 1 ) ) / ( cells_avg_ms * 1e-3 ) / 1e12 if cells_launches else 0.0
    roof_cells = dict( kernel = "k_cells3 + k_cells3_measure", bound = "fp64-valu", achieved = cells_tf, peak = FP64_PEAK_TFLOPS, unit = "TFLOP/s",
                       frac = cells_tf / FP64_PEAK_TFLOPS, traffic = None, avg_launch_ms = cells_avg_ms, launches = cells_launches,
                       flops_per_launch = flops3 / max( cells_launches, 1 ) )
    line = dict( metric = METRIC, value = args.n * args.steps / elapsed, unit = "Diracs/s", n_gpus = 1, steps = args.steps, warmup = args.warmup,
                 ms_per_step = 1e3 * elapsed / args.steps, higher_is_better = True, scaling = "weak", vs_baseline = None, dtype = "f64",
                 data = "synthetic",
                 config = dict( workload = f"3D { args.n } Diracs uniform in [0,1]^3, 4 Gaussians sigma 0.1 truncated to the unit cube, "
                                           f"max mass error <= { args.tol:g} ( absolute )",
                                n = args.n, tol_abs = args.tol, start = "cold ( Voronoi ), tree included", step = "trials", continuation = "auto",
                                linear = "jacobi-pcg" if args.linear == "cg" else "multigrid flexible CG", cells = "fp64" ),
                 roofline = roof_cells,
                 roofline_secondary = dict( kernel = "k_cg_spmv_dot" if args.linear == "cg" else "k_fcg_spmv", bound = "hbm", achieved = spmv_gbs, peak = HBM_PEAK_GBS, unit = "GB/s",
                                  frac = spmv_gbs / HBM_PEAK_GBS, traffic = None, avg_launch_ms = spmv_avg_ms, launches_timed = spmv_launches,
                                  bytes_per_launch = spmv_bytes / max( spmv_launches, 1 ) ),
                 cells = dict( kernels_ms_per_diagram = cells_ms / max( cells_launches, 1 ), diagrams = cells_launches,
                               events_per_cell = { k: v / max( cells_launches, 1 ) / args.n for k, v in ev.items() } ),
                 solve = dict( status = st[ "status" ], residual_abs = st[ "residual" ], residual_rel = st[ "residual" ] * args.n,
                               nb_iter = st[ "nb_iter" ], nb_diag = st[ "nb_diag" ], nb_backtracks = st[ "nb_backtracks" ],
                               lin_nb_iter = st[ "lin_nb_iter" ], nb_continuation_steps = st[ "nb_continuation_steps" ],
                               nb_overflowed = st[ "nb_overflowed" ], t_diag = st[ "t_diag" ], t_lin = st[ "t_lin" ], t_total = st[ "t_total" ],
                               device_bytes = _native.device_bytes() ) )
    if not args.no_cpu_baseline:
        sys.path.insert( 0, os.path.join( ROOT, "oracle" ) )
        import oracle
        m = min( 20_000, args.n )
        cores = min( os.cpu_count() or 1, 16 )
        oracle.set_threads( cores )
        tol = args.tol * args.n / m
        t0 = time.perf_counter()
        r = oracle.solve_3d( pos_h[ :m ], density = dict( gaussians = g ), tol = tol, lin_tol = 1e-6, continuation = "auto" )
        dt = time.perf_counter() - t0
        line[ "cpu_baseline" ] = dict( value = m / dt, unit = "Diracs/s", cores = cores, kind = "port",
                                       sample = f"the first { m } Diracs of the workload, tol { tol:g} ( the same relative accuracy ), "
                                                f"status { r[ 'stats' ][ 'status' ] }, { dt:.1f} s" )
    print( json.dumps( line ), flush = True )


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument( "--gpus", type = int, default = 1 )
    ap.add_argument( "--config", default = "image2d", choices = [ "image2d", "gauss3d", "batch", "serve" ],
                     help = "image2d: the 2D 1024^2-image workload ( the default line ); gauss3d: the 3D four-Gaussian workload; "
                            "batch: 256 independent 2D instances; serve: a Poisson stream of 2D / 3D requests to a server" )
    ap.add_argument( "--serve-k", type = int, default = 40 )
    ap.add_argument( "--serve-rate", type = float, default = 1.0 )
    ap.add_argument( "--batch-k", type = int, default = 256 )
    ap.add_argument( "--batch-res", type = int, default = 256 )
    ap.add_argument( "--streams", type = int, default = 8 )
    ap.add_argument( "--no-plain", action = "store_true", help = "skip the plain FP64 one-instance-per-call leg ( vs_baseline )" )
    ap.add_argument( "--no-lebesgue", action = "store_true", help = "skip the 1M uniform Lebesgue figures" )
    ap.add_argument( "--steps", type = int, default = 3 )
    ap.add_argument( "--warmup", type = int, default = 1 )
    ap.add_argument( "--n", type = int, default = 1_000_000 )
    ap.add_argument( "--res", type = int, default = 1024 )
    ap.add_argument( "--tol", type = float, default = 1e-9 )
    ap.add_argument( "--cpu-sample", type = int, default = 100_000 )
    ap.add_argument( "--no-cpu-baseline", action = "store_true" )
    ap.add_argument( "--step", default = "limits", choices = [ "limits", "trials" ] )
    ap.add_argument( "--continuation", default = "never", choices = [ "never", "auto", "always" ] )
    ap.add_argument( "--linear", default = "auto", choices = [ "auto", "cg" ] )
    ap.add_argument( "--precision", default = "auto", choices = [ "auto", "fp64", "fp32", "mixed" ],
                     help = "the cells' float ( 2D ): FP64, FP32 ( vertices re-solved in double ), or FP32 then FP64 from the first stagnating step" )
    ap.add_argument( "--linear-tol", type = float, default = None,
                     help = "the linear tolerance ( default the reference's 1e-6; negative: the inexact-Newton forcing term with that floor )" )
    ap.add_argument( "--copy-calibration", action = "store_true",
                     help = "first copy 512 MiB device to device ( the known-size kernel that scales the HBM counters )" )
    args = ap.parse_args()
    if args.gpus != 1:
        print( "bench.py: this build runs on ONE GPU ( the multi-GPU design is documented, not built ); use --gpus 1", file = sys.stderr )
        sys.exit( 2 )

    if args.config == "gauss3d":
        import torch
        assert torch.cuda.is_available(), "bench.py needs the GPU"
        return bench_3d( args )
    if args.config == "batch":
        import torch
        assert torch.cuda.is_available(), "bench.py needs the GPU"
        return bench_batch( args )
    if args.config == "serve":
        import torch
        assert torch.cuda.is_available(), "bench.py needs the GPU"
        return bench_serve( args )
    import torch
    import sdot
    from sdot import _native
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    dev = torch.device( "cuda", 0 )
    torch.cuda.set_device( dev )

    if args.copy_calibration:
        src = torch.ones( 64 << 20, dtype = torch.float64, device = dev )
        dst = torch.empty_like( src )
        dst.copy_( src )
        torch.cuda.synchronize()
        print( f"bench: copy calibration: { src.numel() * 8 } bytes read, as many written", file = sys.stderr, flush = True )
        del src, dst

    pos_h, img_h = workload( args.n, args.res )
    pos = torch.as_tensor( pos_h, device = dev )
    target = sdot.Image( img_h, origin = [ 0, 0 ], frame = np.eye( 2 ) / args.res )
    settings = sdot.Iterative( tol = args.tol, continuation = args.continuation, on_failure = "ignore", precision = args.precision,
                               tuning = sdot.Tuning( step = args.step, linear_solver = args.linear, linear_tol = args.linear_tol ) )
    linear = "jacobi-pcg" if args.linear == "cg" else "multigrid flexible CG ( aggregates of 8 tree ranks, degree-2 Chebyshev smoothing, K-cycle on levels 1-2 )"
    print( f"bench: n = { args.n }, image { args.res }^2, tol { args.tol }", file = sys.stderr, flush = True )

    def step():
        return sdot.OtProblem( sdot.SumOfDiracs( pos ), target ).solve( settings )

    for w in range( args.warmup ):
        sol = step()
        print( f"bench: warmup { w }: { sol.stats[ 'status' ] }, { sol.stats[ 'nb_iter' ] } it, t { sol.stats[ 't_total' ]:.3f} s", file = sys.stderr, flush = True )
    _native.set_timing( True )
    _native.kernel_times( reset = True )
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sols = []
    for k in range( args.steps ):
        sols.append( step() )
        print( f"bench: step { k }: { sols[ -1 ].stats[ 'status' ] }, { sols[ -1 ].stats[ 'nb_iter' ] } it", file = sys.stderr, flush = True )
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    times, spmv_bytes, ev = _native.kernel_times()
    fp32c = _native.fp32_counts()
    mg32 = _native.mg32_counts()
    _native.set_timing( False )

    st = sols[ -1 ].stats
    nu = 1.0 / args.n
    ms_step = 1e3 * elapsed / args.steps
    from sdot.SdotPlanNd import _KERNEL_FLOAT
    cells_float = { _native.FP64: "fp64", _native.FP32: "fp32 ( vertices re-solved in fp64 )",
                    _native.MIXED: "mixed ( fp32, fp64 from the first stagnating step )" }[ _KERNEL_FLOAT[ args.precision ] ]
    # ---- rooflines: the cell kernel ( FP64 or FP32 VALU ) and the CG's SpMV ( HBM ), the dominant one first
    cells_ms, cells_launches = times[ "cells" ]
    flops = FLOP_SIDE * ev[ "side_tests" ] + FLOP_PRUNE * ev[ "prune_vertex_tests" ] + FLOP_NEWV * ev[ "new_vertices" ]
    cells_avg_ms = cells_ms / max( cells_launches, 1 )
    cells_tf = ( flops / max( cells_launches, 1 ) ) / ( cells_avg_ms * 1e-3 ) / 1e12 if cells_launches else 0.0
    # the FP32 kernel ( 'fp32', or 'mixed' that never switched ): its walk and cuts in float, then the FP64 finish, timed together
    kf = _KERNEL_FLOAT[ args.precision ]
    k32 = kf == _native.FP32 or ( kf == _native.MIXED and st[ "it_double" ] < 0 )
    if k32:
        t_walk, t_fin = measured_traffic( "k_cells_walk32", args ), measured_traffic( "k_cells_finish32", args )
        roof_cells = dict( kernel = "k_cells_walk32 + k_cells_finish32", bound = "fp32-valu", achieved = cells_tf, peak = FP32_PEAK_TFLOPS,
                           unit = "TFLOP/s", frac = cells_tf / FP32_PEAK_TFLOPS,
                           traffic = None if t_walk is None or t_fin is None else t_walk + t_fin, avg_launch_ms = cells_avg_ms,
                           launches = cells_launches, flops_per_launch = flops / max( cells_launches, 1 ) )
    else:
        roof_cells = dict( kernel = "k_cells_pass1", bound = "fp64-valu", achieved = cells_tf, peak = FP64_PEAK_TFLOPS, unit = "TFLOP/s",
                           frac = cells_tf / FP64_PEAK_TFLOPS, traffic = measured_traffic( "k_cells_pass1", args ), avg_launch_ms = cells_avg_ms,
                           launches = cells_launches, flops_per_launch = flops / max( cells_launches, 1 ) )
    spmv_ms, spmv_launches = times[ "spmv" ]
    spmv_avg_ms = spmv_ms / max( spmv_launches, 1 )
    spmv_gbs = ( spmv_bytes / max( spmv_launches, 1 ) ) / ( spmv_avg_ms * 1e-3 ) / 1e9 if spmv_launches else 0.0
    roof_spmv = dict( kernel = "k_cg_spmv_dot" if args.linear == "cg" else "k_fcg_spmv", bound = "hbm", achieved = spmv_gbs, peak = HBM_PEAK_GBS, unit = "GB/s", frac = spmv_gbs / HBM_PEAK_GBS,
                      traffic = measured_traffic( "k_fcg_spmv" if args.linear != "cg" else "k_cg_spmv_dot", args ), avg_launch_ms = spmv_avg_ms,
                      launches_timed = spmv_launches,
                      bytes_per_launch = spmv_bytes / max( spmv_launches, 1 ) )
    # the SpMV is timed one launch in sixteen: its share of the step is estimated from the CG iteration count
    spmv_total_ms = spmv_avg_ms * st[ "lin_nb_iter" ] * args.steps
    dominant, other = ( roof_spmv, roof_cells ) if spmv_total_ms > cells_ms else ( roof_cells, roof_spmv )

    line = dict( metric = METRIC, value = args.n * args.steps / elapsed, unit = "Diracs/s", n_gpus = 1, steps = args.steps, warmup = args.warmup,
                 ms_per_step = ms_step, higher_is_better = True, scaling = "weak", vs_baseline = None,
                 dtype = "f32+f64" if kf != _native.FP64 else "f64", data = "synthetic",
                 config = dict( workload = f"2D { args.n } Diracs uniform in [0,1]^2, { args.res }^2 image of 8 Gaussian bumps + 0.1 floor, "
                                           f"max mass error <= { args.tol:g} ( absolute )",
                                n = args.n, image = [ args.res, args.res ], tol_abs = args.tol, start = "cold ( Voronoi ), tree included",
                                step = args.step, continuation = args.continuation, linear = linear, cells = cells_float ),
                 roofline = dominant, roofline_secondary = other,
                 solve = dict( status = st[ "status" ], residual_abs = st[ "residual" ], residual_rel = st[ "residual" ] / nu,
                               nb_iter = st[ "nb_iter" ], nb_diag = st[ "nb_diag" ], lin_nb_iter = st[ "lin_nb_iter" ],
                               nb_continuation_steps = st[ "nb_continuation_steps" ], t_lim = st[ "t_lim" ], it_double = st[ "it_double" ],
                               t_diag = st[ "t_diag" ], t_majorant = st[ "t_majorant" ], t_lin = st[ "t_lin" ], t_total = st[ "t_total" ],
                               device_bytes = _native.device_bytes(),
                               fp32_repairs = dict( diagrams = fp32c[ "diagrams" ], cells_redone_fp64 = fp32c[ "redone" ],
                                                    facets_dropped = fp32c[ "dropped" ] ),
                               mg_fp32_preconditioner = dict( systems = mg32[ "solves" ], iterations = mg32[ "iterations" ],
                                                              redone_in_fp64 = mg32[ "fallbacks" ] ) ) )

    # the FP64 cells of the returned weights: the error of the plan against the FP64 kernel ( the oracle's cells: equal to the
    # bit on a constant density, to 1e-14 relative on an image, by the parity tests )
    w_fin = sols[ -1 ].weights
    m64 = _native.measures_2d( pos, w_fin, [ 0, 0, 1, 1 ], target.normalized_version() )
    err64 = float( ( m64 - torch.as_tensor( sols[ -1 ].target_masses, device = dev ) ).abs().max() )
    line[ "solve" ][ "fp64_cells_residual_abs" ] = err64
    line[ "solve" ][ "fp64_cells_residual_rel" ] = err64 / nu
    line[ "solve" ][ "reported_vs_fp64_cells" ] = float( ( m64 - sols[ -1 ].cell_masses ).abs().max() )
    del m64
    # THE PLAIN PATH: the same solve with FP64 cells, one instance per call ( the metric's baseline, `vs_baseline` )
    if not args.no_plain:
        plain = sdot.Iterative( tol = args.tol, continuation = args.continuation, on_failure = "ignore", precision = "fp64",
                                tuning = sdot.Tuning( step = args.step, linear_solver = args.linear, linear_tol = args.linear_tol ) )
        tp = []
        for r in range( 2 ):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            sp = sdot.OtProblem( sdot.SumOfDiracs( pos ), target ).solve( plain )
            torch.cuda.synchronize()
            tp.append( time.perf_counter() - t1 )
            print( f"bench: plain FP64 path: { sp.stats[ 'status' ] }, { sp.stats[ 'nb_iter' ] } it, { tp[ -1 ]:.3f} s", file = sys.stderr, flush = True )
        t_plain = min( tp )
        line[ "vs_baseline" ] = line[ "value" ] / ( args.n / t_plain )
        line[ "plain_fp64" ] = dict( seconds = t_plain, diracs_per_s = args.n / t_plain, status = sp.stats[ "status" ], nb_iter = sp.stats[ "nb_iter" ],
                                     lin_nb_iter = sp.stats[ "lin_nb_iter" ], what = "the same workload, FP64 cells, one solve per call ( min of 2 )" )
    # the 1M uniform Lebesgue workload ( the reference's published 1M figures are on it ): the same 1M seeds, the unit square
    if not args.no_lebesgue and args.n == 1_000_000:
        leb = {}
        box = sdot.UniformBox( [ 0, 0 ], [ 1, 1 ] )
        for prec in ( "auto", "fp64" ):
            st_l = sdot.Iterative( tol = args.tol, continuation = "never", on_failure = "ignore", precision = prec,
                                   tuning = sdot.Tuning( linear_solver = args.linear ) )
            ts = []
            for r in range( 3 ):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                sl = sdot.OtProblem( sdot.SumOfDiracs( pos ), box ).solve( st_l )
                torch.cuda.synchronize()
                ts.append( time.perf_counter() - t1 )
            leb[ "mixed" if prec == "auto" else prec ] = dict( seconds = min( ts[ 1: ] ), status = sl.stats[ "status" ], nb_iter = sl.stats[ "nb_iter" ],
                                                            lin_nb_iter = sl.stats[ "lin_nb_iter" ], residual_abs = sl.stats[ "residual" ] )
            print( f"bench: 1M uniform Lebesgue ( { prec } ): { min( ts[ 1: ] ):.3f} s, { sl.stats[ 'nb_iter' ] } it", file = sys.stderr, flush = True )
        leb[ "what" ] = ( "the same 1M seeds, the Lebesgue measure of the unit square, tol 1e-9 absolute, cold start, min of 2 after a "
                          "warm-up; the reference publishes 0.383 s ( mixed, A6000 ) and 0.839 s ( FP64, 2080 Ti ) on this workload" )
        line[ "uniform_lebesgue_1M" ] = leb
    if not args.no_cpu_baseline:
        line[ "cpu_baseline" ] = cpu_baseline( pos_h, img_h, args )
    print( json.dumps( line ), flush = True )


def measured_traffic( kernel, args ):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary of this workload ( FETCH_SIZE and
    WRITE_SIZE, scaled by a 512 MiB copy: profiles/hbm_traffic_bench_1M_image.md ), for the default workload only"""
    if ( args.n, args.res, args.step, args.continuation ) != ( 1_000_000, 1024, "limits", "never" ):
        return None
    try:
        with open( os.path.join( ROOT, "profiles", "hbm_traffic_bench_1M_image.json" ) ) as f:
            k = json.load( f )[ "kernels" ].get( kernel )
    except OSError:
        return None
    return None if k is None else k[ "bytes" ]


def cpu_model():
    try:
        with open( "/proc/cpuinfo" ) as f:
            for line in f:
                if line.startswith( "model name" ):
                    return line.split( ":", 1 )[ 1 ].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline( pos_h, img_h, args ):
    """the oracle's solve of the first `cpu_sample` Diracs of the same workload on the host cores, with the linear solver the
    reference's CPU solve picks by default ( the multigrid ): one untimed run, then the best of three"""
    sys.path.insert( 0, os.path.join( ROOT, "oracle" ) )
    import oracle
    m = min( args.cpu_sample, len( pos_h ) )
    cores = min( os.cpu_count() or 1, 16 )
    oracle.set_threads( cores )
    v = img_h / ( img_h.mean() )                         # ( the unit square: mean value = mass )
    dens = dict( values = np.ascontiguousarray( v.T ), x0 = 0.0, y0 = 0.0, hx = 1.0 / args.res, hy = 1.0 / args.res )
    tol = args.tol * len( pos_h ) / m                    # the same relative accuracy per Dirac
    ts = []
    for k in range( 4 ):
        t0 = time.perf_counter()
        r = oracle.solve( pos_h[ :m ], density = dens, tol = tol, lin_tol = 1e-6, step = args.step, continuation = args.continuation,
                          linear = "mg" )
        ts.append( time.perf_counter() - t0 )
        print( f"bench: CPU baseline run { k }: { ts[ -1 ]:.2f} s, status { r[ 'stats' ][ 'status' ]:.0f}", file = sys.stderr, flush = True )
    dt = min( ts[ 1: ] )
    s = r[ "stats" ]
    return dict( value = m / dt, unit = "Diracs/s", cores = cores, kind = "port",
                 sample = f"the oracle ( C++ FP64 restatement built -O3 -march=x86-64-v3, multigrid-preconditioned CG as the reference's "
                          f"CPU default ) solving the first { m } Diracs of the same workload to max |m - nu| <= { tol:.1e} ( the same error "
                          f"relative to nu ), { s[ 'nb_iter' ]:.0f} Newton iterations, best of 3 after one warm-up: { dt:.2f} s "
                          f"( runs { ', '.join( f'{ t:.2f}' for t in ts[ 1: ] ) } s; linear { s[ 't_lin' ]:.2f} s, cells { s[ 't_diag' ]:.2f} s ), "
                          f"status { s[ 'status' ]:.0f}, on { cores } threads of { cpu_model() }" )


if __name__ == "__main__":
    main()
