"""SERVING: a `Server` takes solve requests as they come and runs them with continuous batching.

    with Server( slots = 8 ) as srv:
        tickets = [ srv.submit( problem, settings ) for problem in incoming ]
        plans = [ t.result() for t in tickets ]

The reference solves one problem per call and has no server. Here each slot is a HIP stream driven by its own host
thread with its own device buffers ( kept between its requests ); a free slot takes the oldest waiting request whose
footprint ( the byte model `sdot.memory.solve_bytes`, part by part ) fits the memory budget beside what the other slots hold. Requests that arrive while others run are
admitted as soon as a slot and the memory allow, and their work overlaps on the device. A request's plan does not
depend on what ran beside it ( the kernels are deterministic and nothing is shared ). `ticket.times` gives its arrival,
start and end ( seconds from the server's start ): the latency a caller sees is `end - arrival`.
"""

import ctypes

import numpy as np

from . import _native
from .distributions import Image, SumOfGaussians, UniformBox
from .SdotPlanNd import SdotPlanNd, card_options


class Ticket:
    """a submitted request: `result()` blocks until its plan is ready"""

    def __init__( self, server, handle, problem, settings, keep ):
        self._server, self._handle, self.problem, self._settings, self._keep = server, handle, problem, settings, keep
        self._plan = None
        self.times = None

    def result( self ):
        if self._plan is not None:
            return self._plan
        err = np.zeros( 3, np.int64 )
        times = np.zeros( 3 )
        _native._check( _native.lib().sdot_amd_server_wait( self._server._h, self._handle, err.ctypes.data_as( _native._P ),
                                                            times.ctypes.data_as( _native._P ) ) )
        self.times = tuple( float( t ) for t in times )
        self._server._open.discard( self )
        if err[ 0 ] != 0:
            raise _native.SdotError( f"server: the request failed ( kind { err[ 0 ] }, code { err[ 1 ] }, value { err[ 2 ] } )", *[ int( v ) for v in err ] )
        k = self._keep
        pl = SdotPlanNd.__new__( SdotPlanNd )
        pl.problem = self.problem
        pl.weights, pl.cell_masses, pl.barycenters = k[ "weights" ], k[ "masses" ], k[ "bary" ]
        pl._weights_fit = pl.weights
        pl.cost = float( k[ "cost" ][ 0 ] )
        pl._clusters = None
        pl._masses = np.asarray( self.problem.source.weights, dtype = np.float64 )
        pl._warm_start = "weights0" if self._settings.weights0 is not None else "none"
        pl._dup_note = "not checked ( server )"
        pl._read_stats( k[ "stats" ], self._settings )
        pl.stats[ "aggregation" ] = "off ( server )"
        pl.history = []
        self._keep = None
        self._plan = pl
        if getattr( self._settings, "on_failure", "raise" ) == "raise" and not pl.converged:
            from .OtProblem import OtNotConverged
            raise OtNotConverged( pl )
        return pl


class Server:
    """`slots` concurrent solves ( 1 .. 32 ), `memory` the device bytes they may hold together ( default: 80 % of the
    free device memory at start )"""

    def __init__( self, slots = 8, memory = None ):
        T = _native.torch()
        _native.device()
        if memory is None:
            memory = 0.8 * T.cuda.mem_get_info()[ 0 ]
        self._h = _native.lib().sdot_amd_server_start( int( slots ), float( memory ) )
        if self._h == 0:
            _native._check( 1 )
        self._open = set()

    def submit( self, problem, settings = None ):
        """queues the solve of `problem` ( a 2D or 3D `OtProblem` ) with `settings` ( an `Iterative` ); returns its `Ticket`.
        The cells' float follows `settings.precision` as a single solve's does ( 2D: 'auto' and 'mixed' the FP32 kernel then
        the FP64 one; 3D: FP64 cells ), so a served plan is the plan `problem.solve( settings )` returns"""
        from .OtProblem import Iterative
        from .SdotPlanNd import _KERNEL_FLOAT
        settings = settings if settings is not None else Iterative()
        d = problem.nb_dims
        if d not in ( 2, 3 ):
            raise NotImplementedError( f"Server: 2D or 3D problems on the MI355X path ( not { d }D )" )
        src, dst = problem.source, problem.target
        problem.domain
        T = _native.torch()
        dev = _native.device()
        pos = _native.on_device( src.positions ).reshape( -1, d ).contiguous()
        n = int( pos.shape[ 0 ] )
        nu = _native.on_device( src.weights ).reshape( -1 ).contiguous()
        w0 = None if settings.weights0 is None else _native.on_device( settings.weights0 ).reshape( -1 ).contiguous()
        lo, hi = dst.box()
        opts = np.ascontiguousarray( card_options( settings, dst ) )
        if d == 3:
            if settings.precision == "fp32":
                raise NotImplementedError( "Server: 3D cells are FP64 ( precision 'auto', 'fp64' or 'mixed' )" )
            if not isinstance( dst, ( UniformBox, SumOfGaussians, Image ) ):
                raise NotImplementedError( f"Server, 3D: a UniformBox, a SumOfGaussians or an Image target ( not { type( dst ).__name__ } )" )
            box = np.asarray( [ *lo[ :3 ], *hi[ :3 ] ], dtype = np.float64 )
            opts[ _native.CARD_OPTIONS.index( "step" ) ] = 0.0
            kind, rho, dens, nd = dst.native_args3( device_tensor = _native.on_device )
            precision = _native.FP64
        else:
            if not isinstance( dst, ( Image, UniformBox, SumOfGaussians ) ):
                raise NotImplementedError( f"Server: an Image, a UniformBox or a SumOfGaussians target ( not { type( dst ).__name__ } )" )
            box = np.asarray( [ lo[ 0 ], lo[ 1 ], hi[ 0 ], hi[ 1 ] ], dtype = np.float64 )
            kind, rho, dens, nx, ny, geo = dst.native_args( device_tensor = _native.on_device )
            nd = nx
            precision = _KERNEL_FLOAT[ settings.precision ]
        keep = dict( pos = pos, nu = nu, w0 = w0, img = dens, box = box, opts = opts,
                     weights = T.empty( n, dtype = T.float64, device = dev ), masses = T.empty( n, dtype = T.float64, device = dev ),
                     bary = T.empty( ( n, d ), dtype = T.float64, device = dev ), stats = np.zeros( 31 ), cost = np.zeros( 1 ) )
        T.cuda.current_stream().synchronize()           # ( the inputs are on the device before a slot reads them )
        ticket = ctypes.c_longlong( 0 )
        P = _native._ptr
        outs = ( P( keep[ "weights" ] ), P( keep[ "masses" ] ), P( keep[ "bary" ] ), keep[ "stats" ].ctypes.data_as( _native._P ),
                 keep[ "cost" ].ctypes.data_as( _native._P ), ctypes.byref( ticket ) )
        if d == 3 and kind == _native.IMAGE:             # ( a voxel image: its counts and geometry beside the voxels )
            nx, ny, nz, geo = nd
            keep[ "geo" ] = np.ascontiguousarray( geo, dtype = np.float64 )
            _native._check( _native.lib().sdot_amd_server_submit_3d_image( self._h, n, P( pos ), P( nu ), P( w0 ), opts.ctypes.data_as( _native._P ),
                                                                           box.ctypes.data_as( _native._P ), P( dens ), nx, ny, nz,
                                                                           keep[ "geo" ].ctypes.data_as( _native._P ), *outs ) )
        else:
            _native._check( _native.lib().sdot_amd_server_submit_nd( self._h, d, precision, n, P( pos ), P( nu ), P( w0 ), opts.ctypes.data_as( _native._P ),
                                                                     box.ctypes.data_as( _native._P ), kind, rho, P( dens ), int( nd ), *outs ) )
        t = Ticket( self, ticket.value, problem, settings, keep )
        self._open.add( t )
        return t

    def info( self ):
        """`{ "peak_bytes": the largest sum of the slots' device bytes seen at the end of a request, "done": requests done }`"""
        out = np.zeros( 2 )
        _native._check( _native.lib().sdot_amd_server_info( self._h, out.ctypes.data_as( _native._P ) ) )
        return dict( peak_bytes = float( out[ 0 ] ), done = int( out[ 1 ] ) )

    def close( self ):
        """waits for every open request, then stops the slots"""
        if self._h:
            for t in list( self._open ):
                try:
                    t.result()
                except Exception:
                    pass
            _native._check( _native.lib().sdot_amd_server_stop( self._h ) )
            self._h = 0

    def __enter__( self ):
        return self

    def __exit__( self, *exc ):
        self.close()
