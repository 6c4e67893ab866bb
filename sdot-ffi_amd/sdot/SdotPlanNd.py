"""`SdotPlanNd`: the solution of a semi-discrete transport in `d >= 2`, on the MI355X path.

Mirrors `src/sdot/SdotPlanNd.py` of the reference: the whole fit is ONE call of the C ABI
( `sdot_amd_solve_2d`, standing for `sdotplan_solve_card_2d[_image]`, `SdotPlanNd.py:864-890` ); the
options cross as the reference card's 41-double vector, `stats` / `history` come back in the
reference's layouts ( `_STATS`, `_HISTORY` ) and `OtNotConverged` is raised the same way.
"""

import numpy as np

from . import _native
from .distributions import Image, Polytope, SumOfDiracs, SumOfGaussians, UniformBox, _host

_STATS = [ "status", "residual", "residual0", "nb_iter", "nb_diag", "nb_backtracks", "t_majorant", "t_diag", "t_asm", "t_lin", "t_lim", "eps",
           "domain_mass", "nb_overflowed", "nb_cell_lim", "nb_limit_rounds", "lin_nb_hierarchies", "lin_nb_iter", "lin_worst", "start", "t_total",
           "nb_continuation_steps", "min_start_mass", "it_switch", "it_double", "scratch_bytes",
           "nb_clusters", "nb_aggregated", "nb_duplicates", "residual_full", "nb_polish" ]
_HISTORY = [ "step", "t", "residual_l2", "min_measure", "max_abs_residual", "nb_diag", "nb_evals", "s" ]
_STATUS = { 0: "running", 1: "converged", 2: "max iterations", 3: "stagnation", 4: "linear solver failure", 5: "card capacity",
            6: "failure", 7: "converged (aggregated)" }
_START = { 0: "weights0", 1: "voronoi", 2: "similarity" }
_LIN = { "auto": 0, "cholesky": 1, "amg": 2, "cg": 3, "mg": 4 }
_STEP = { "trials": 0, "limits": 1 }
#: the 2D cells' float per `precision` ( `src/sdot/SdotPlanNd.py:772-781`: 'mixed', and 'auto' as on the reference's card,
#: the FP32 kernel, then the FP64 one from the first FP32 step that stagnates )
_KERNEL_FLOAT = { "auto": _native.MIXED, "fp64": _native.FP64, "fp32": _native.FP32, "mixed": _native.MIXED }
_RESIDUAL = { "lin": 0, "log": 1, "power": 2 }
_CONTINUATION = { "never": 0, "auto": 1, "always": 2 }
#: THE AGGREGATION ( `sdotplan/Aggregation.h` ): a pair is merged when one ulp of its weights moves more than `tol / margin` of
#: mass between them ( `Tuning( aggregation_margin = ... )`, `SdotPlanNd.py:115-118` )
_AGGREGATE_MARGIN = 4.0
#: exact duplicates: `w_dup = w_rep - gap`, `gap` this fraction of the larger of the squared norm and the squared extent of the
#: cloud -- any positive gap empties a duplicate's cell, this one far above the rounding of the cuts ( `SdotPlanNd.py:119-121` )
_DUPLICATE_GAP = 1e-6


def _mix64( z ):
    """splitmix64's finalizer on a `uint64` array ( a bijection: distinct inputs, distinct outputs )"""
    z = ( z ^ ( z >> np.uint64( 30 ) ) ) * np.uint64( 0xBF58476D1CE4E5B9 )
    z = ( z ^ ( z >> np.uint64( 27 ) ) ) * np.uint64( 0x94D049BB133111EB )
    return z ^ ( z >> np.uint64( 31 ) )


def exact_duplicates( positions ):
    """THE EXACT DUPLICATES of a cloud `[ n, d ]`: `( dup, rep )`, each seed whose position an earlier one already holds and
    the smallest index at that position, both increasing with `dup` ( empty `int64` arrays when there is none; `-0.0` and
    `0.0` alike ). Reference `src/sdot/SdotPlanNd.py:124-155`. A 64-bit print of each row's bits, sorted: when no two prints
    are equal there is no duplicate; otherwise the rows whose prints collide are grouped exactly."""
    P = np.ascontiguousarray( _host( positions ), dtype = np.float64 )
    n = P.shape[ 0 ]
    none = ( np.zeros( 0, dtype = np.int64 ), np.zeros( 0, dtype = np.int64 ) )
    if n < 2:
        return none
    P = P.reshape( n, -1 ) + 0.0
    bits = P.view( np.uint64 )
    h = np.zeros( n, dtype = np.uint64 )
    for k in range( bits.shape[ 1 ] ):
        h = _mix64( h ^ bits[ :, k ] )
    if not ( np.diff( np.sort( h ) ) == 0 ).any():
        return none
    order = np.argsort( h, kind = "stable" )
    hs = h[ order ]
    eq = hs[ 1: ] == hs[ :-1 ]
    cand = np.unique( np.concatenate( [ order[ 1: ][ eq ], order[ :-1 ][ eq ] ] ) )
    _, inv = np.unique( P[ cand ], axis = 0, return_inverse = True )
    inv = inv.reshape( -1 )
    first = np.full( int( inv.max() ) + 1, n, dtype = np.int64 )
    np.minimum.at( first, inv, cand )
    rep = first[ inv ]
    keep = rep != cand
    return cand[ keep ].astype( np.int64 ), rep[ keep ].astype( np.int64 )


def duplicate_gap( positions ):
    """the gap of the exact duplicates' weights ( `SdotPlanNd.py:669-674` )"""
    P = np.asarray( _host( positions ), dtype = np.float64 )
    P = P.reshape( P.shape[ 0 ], -1 )
    ext = P.max( axis = 0 ) - P.min( axis = 0 )
    s2 = max( float( ( P * P ).sum( axis = 1 ).max() ), float( ( ext * ext ).sum() ) )
    return _DUPLICATE_GAP * ( s2 if s2 > 0 else 1.0 )


def _solve_3d_merged( src, dup, rep, w0, opts, box, dst ):
    """THE EXACT DUPLICATES OF A 3D CLOUD ( `dup[ q ]` at the place of `rep[ q ]` ): the solve of the cloud without them, each
    representative carrying the targets of its duplicates, brought back to the whole cloud: a duplicate's weight is its
    representative's less `k gap` ( the `k`-th duplicate of that representative, `gap = duplicate_gap`, as the 2D solve ties
    them, `SdotPlanNd.py:650-685` ), its cell empty ( mass 0, its seed as barycentre ), its cluster its representative. This
    is the plan the 2D solve's merge reaches on the device: the duplicates' empty cells take no part in the others'"""
    T = _native.torch()
    n = int( src.nb_diracs )
    keep = np.ones( n, dtype = bool )
    keep[ dup ] = False
    kept = np.flatnonzero( keep )
    nu = np.asarray( _host( src.weights ), dtype = np.float64 ).reshape( -1 ).copy()
    np.add.at( nu, rep, nu[ dup ] )
    pos = _native.on_device( src.positions ).reshape( n, 3 )
    ki = T.as_tensor( kept, device = pos.device )
    w0r = None if w0 is None else _native.on_device( w0 ).reshape( -1 )[ ki ].contiguous()
    res = _native.solve_3d( pos[ ki ].contiguous(), nu[ kept ], w0r, opts, box, dst )
    k = np.zeros( len( dup ) )
    count = {}
    for q, r in enumerate( rep.tolist() ):
        count[ r ] = count.get( r, 0 ) + 1
        k[ q ] = count[ r ]
    di, ri = T.as_tensor( dup, device = pos.device ), T.as_tensor( rep, device = pos.device )
    w = T.zeros( n, dtype = T.float64, device = pos.device )
    w[ ki ] = res[ "weights" ]
    w[ di ] = w[ ri ] - T.as_tensor( k * duplicate_gap( src.positions ), dtype = T.float64, device = pos.device )
    masses = T.zeros( n, dtype = T.float64, device = pos.device )
    masses[ ki ] = res[ "cell_masses" ]
    bary = pos.clone()
    bary[ ki ] = res[ "barycenters" ]
    clusters = T.arange( n, dtype = T.int64, device = pos.device )
    clusters[ di ] = ri
    stats = np.asarray( res[ "stats" ], dtype = np.float64 ).copy()
    nb_reps = len( np.unique( rep ) )
    stats[ _STATS.index( "nb_duplicates" ) ] = len( dup )
    stats[ _STATS.index( "nb_clusters" ) ] = nb_reps
    stats[ _STATS.index( "nb_aggregated" ) ] = len( dup ) + nb_reps
    return dict( weights = w, cell_masses = masses, barycenters = bary, clusters = clusters, cost = res[ "cost" ], stats = stats,
                 history = res[ "history" ] )


def card_options( settings, density, verbose = False, nb_dups = 0, gap = 0.0 ):
    """the 41-double option vector ( `SdotPlanNd.py:806-833` of the reference, same names and meanings )"""
    tun = settings.tuning
    step = { "auto": "limits" }.get( tun.step, tun.step )       # ( 2D: `SdotPlanNd.py:462` )
    if step not in _STEP:
        raise ValueError( f"unknown step: { tun.step!r} ( 'auto', 'trials' or 'limits' )" )
    if tun.residual not in _RESIDUAL:
        raise ValueError( f"unknown residual: { tun.residual!r}" )
    if tun.linear_solver not in _LIN:
        raise ValueError( f"unknown linear_solver: { tun.linear_solver!r}" )
    opts = [ 0.0 ] * len( _native.CARD_OPTIONS )

    def put( name, v ):
        opts[ _native.CARD_OPTIONS.index( name ) ] = float( v )

    put( "tol_abs", settings.tol ); put( "tol_rel", tun.mass_rtol ); put( "t_min", tun.t_min ); put( "mult_ok", tun.restart_factor )
    put( "factor", 0.9 ); put( "maxit", settings.max_iter ); put( "max_backtracks", tun.max_backtracks ); put( "step", _STEP[ step ] )
    put( "residual", _RESIDUAL[ tun.residual ] ); put( "power", tun.residual_power ); put( "switch", tun.residual_switch )
    put( "lin", 0 if tun.linear_solver == "cg" else 1 ); put( "host_method", _LIN[ tun.linear_solver ] ); put( "lin_tol", tun.linear_tol or 0.0 )
    put( "lin_maxit", 0 ); put( "trace", int( bool( verbose ) ) )
    # the multigrid: aggregates of `mg_pack` ranks ( a power of two ), `mg_nu` sweeps, K-cycle on the top `mg_kcycle` levels
    if tun.mg_pack:
        put( "mg_shift", int( tun.mg_pack ).bit_length() - 1 )
    put( "mg_nu", tun.mg_nu or 0 ); put( "mg_kcycle", -1 if tun.mg_kcycle is None else tun.mg_kcycle )
    # the multigrid's cycle in FP32 as the FP64 flexible CG's preconditioner ( `mg_precision` "float", the reference card's float levels,
    # `SdotPlanNd.py:241, 809-812` ) or all in FP64 ( "double" )
    mg_precision = tun.mg_precision or "float"
    if mg_precision not in ( "float", "double" ):
        raise ValueError( f"mg_precision: 'float' or 'double' ( got { mg_precision!r} )" )
    put( "mg_float", int( mg_precision == "float" ) )
    # the rows under which the hierarchy's bottom runs in one workgroup ( None: the default; 0 or less: never )
    if tun.mg_stop is not None:
        put( "mg_stop", tun.mg_stop if tun.mg_stop > 0 else -1 )
    put( "continuation", _CONTINUATION[ settings.continuation ] ); put( "conv_threshold", tun.conv_threshold )
    put( "conv_s0", tun.conv_start or 0.0 ); put( "conv_ratio", tun.conv_ratio ); put( "conv_min", tun.conv_min or 0.0 )
    put( "agg_gap", gap ); put( "agg_nb_dups", nb_dups )
    if settings.aggregate:
        put( "agg_margin", _AGGREGATE_MARGIN if tun.aggregation_margin is None else tun.aggregation_margin )
    if isinstance( density, Image ) and density.nb_dims == 2:     # ( a 3D image goes by its own arguments, no continuation )
        lo, hi = density.box()
        h = np.diag( density.frame )
        nx, ny = density.shape
        for name, v in zip( ( "img_x0", "img_y0", "img_hx", "img_hy", "img_nx", "img_ny" ), ( lo[ 0 ], lo[ 1 ], h[ 0 ], h[ 1 ], nx, ny ) ):
            put( name, v )
        # the blur of the width continuation stops at a quarter pixel ( `sdotplan/convolved/Image.h:49-53` )
        put( "conv_possible", 1 ); put( "min_scale", min( abs( h[ 0 ] ), abs( h[ 1 ] ) ) / 4 )
    if isinstance( density, SumOfGaussians ):
        # the Gaussians' count; the widening stops at a quarter of the narrowest ( `convolved/SumOfGaussians.h:39` )
        put( "img_nx", len( density.sigmas ) ); put( "conv_possible", 1 ); put( "min_scale", float( density.sigmas.min() ) / 4 )
    return np.asarray( opts )


class SdotPlanNd:
    """see the module docstring; obtained from `OtProblem.solve()`"""

    @classmethod
    def _solve( cls, problem, settings, verbose, warm = None ):
        self = cls.__new__( cls )
        self._build( problem, settings, verbose, warm )
        self._check_converged( settings )
        return self

    def _build( self, problem, settings, verbose, warm = None ):
        self.problem = problem
        src, dst = problem.source, problem.target
        self._turn = None
        if isinstance( dst, Image ) and dst.rotation is not None:
            # a turned grid: the problem solved in the grid's frame ( a move and a turn: the same power distances, so the same
            # weights and masses ), its barycentres brought back ( `_set_result` )
            self._turn = dst
            src = SumOfDiracs( dst.to_local( src.positions ), src.weights, src.target_mass )
            dst = dst.local_version()
        d = problem.nb_dims
        if d not in ( 2, 3 ):
            raise NotImplementedError( f"SdotPlanNd on the MI355X path: 2D or 3D ( not { d }D )" )
        if settings.precision == "fp32" and d == 3:
            raise NotImplementedError( "precision = 'fp32' is 2D on the MI355X path ( 3D cells are FP64: 'auto', 'fp64' or 'mixed' )" )
        plan = settings.ot_plan if settings.ot_plan is not None else warm
        if settings.weights0 is not None:
            w0, self._warm_start = settings.weights0, "weights0"
        elif plan is not None:
            n_plan = int( plan.weights.shape[ 0 ] )
            if n_plan != src.nb_diracs:
                if settings.ot_plan is not None:
                    raise ValueError( f"Iterative( ot_plan = ... ): this plan carries { n_plan } weights and the source has { src.nb_diracs }" )
                w0, self._warm_start = None, f"none ( the kept plan carries { n_plan } weights, the source has { src.nb_diracs } )"
            else:
                w0, self._warm_start = plan.weights, "ot_plan"
        else:
            w0, self._warm_start = None, "none"
        problem.domain
        lo, hi = dst.box()
        if d == 3:                                       # ( a constant, a sum of Gaussians or a voxel image; the step by trials, no duplicates )
            if not isinstance( dst, ( UniformBox, SumOfGaussians, Image ) ):
                raise NotImplementedError( f"SdotPlanNd on the MI355X path, 3D: a UniformBox, a SumOfGaussians or an Image target "
                                           f"( not { type( dst ).__name__ } )" )
            dup = rep = np.zeros( 0, dtype = np.int64 )
            if settings.aggregate:
                dup, rep = exact_duplicates( src.positions )
            self._dup_note = "checked" if settings.aggregate else "not checked ( aggregate = False )"
            step = settings.tuning.step
            if step == "limits":
                raise NotImplementedError( "SdotPlanNd on the MI355X path, 3D: the step is by trials ( the limits are a 2D step )" )
            opts = card_options( settings, dst, verbose, 0, 0.0 )
            opts[ _native.CARD_OPTIONS.index( "step" ) ] = 0.0   # ( 'auto': trials in 3D )
            box3 = [ *lo[ :3 ], *hi[ :3 ] ]
            if len( dup ):
                res = _solve_3d_merged( src, dup, rep, w0, opts, box3, dst )
            else:
                res = _native.solve_3d( src.positions, src.weights, w0, opts, box3, dst )
            self._set_result( res, src, settings )
            return
        box = [ lo[ 0 ], lo[ 1 ], hi[ 0 ], hi[ 1 ] ]
        if not isinstance( dst, ( Image, UniformBox, SumOfGaussians, Polytope ) ):
            raise NotImplementedError( f"SdotPlanNd on the MI355X path: an Image, a UniformBox, a SumOfGaussians or a Polytope target "
                                       f"( not { type( dst ).__name__ } )" )
        # a Polytope target IS the domain: the cells start from its polygon ( `StartCell.py` )
        domain = dst.domain_polygon() if isinstance( dst, Polytope ) else None
        # THE EXACT DUPLICATES ( `SdotPlanNd.py:650-685` ): found here on the host, merged by the solve
        dup = rep = np.zeros( 0, dtype = np.int64 )
        gap = 0.0
        self._dup_note = "not checked ( aggregate = False )"
        if settings.aggregate:
            dup, rep = exact_duplicates( src.positions )
            self._dup_note = "checked"
            if len( dup ):
                gap = duplicate_gap( src.positions )
        opts = card_options( settings, dst, verbose, len( dup ), gap )
        pairs = np.stack( [ dup, rep ], axis = 1 ) if len( dup ) else None
        res = _native.solve_2d( src.positions, src.weights, w0, opts, box, dst, dups = pairs, precision = _KERNEL_FLOAT[ settings.precision ],
                                domain = domain )
        self._set_result( res, src, settings )

    def _set_result( self, res, src, settings ):
        #: the fitted weights ( `weights[ 0 ] == 0` ), device tensor, user order
        self.weights = res[ "weights" ]
        self._weights_fit = self.weights
        #: the measure of each cell at the fitted weights
        self.cell_masses = res[ "cell_masses" ]
        #: the barycenter of each cell ( its seed if empty ), `[ n, d ]`
        self.barycenters = res[ "barycenters" ]
        if getattr( self, "_turn", None ) is not None:
            self.barycenters = self._turn.from_local( self.barycenters )
        #: the transport cost `sum_i int_{cell_i} |x - p_i|^2 rho`
        self.cost = res[ "cost" ]
        self._clusters = res[ "clusters" ]
        self._masses = np.asarray( src.weights, dtype = np.float64 )
        self._read_stats( res[ "stats" ], settings )
        self._read_history( res[ "history" ] )

    def _read_stats( self, st, settings ):
        self.stats = { name: float( st[ k ] ) for k, name in enumerate( _STATS ) }
        self.stats[ "status" ] = _STATUS.get( int( self.stats[ "status" ] ), "?" )
        self.stats[ "it_switch" ] = int( self.stats[ "it_switch" ] )
        self.stats[ "it_double" ] = int( self.stats[ "it_double" ] )
        self.stats[ "start" ] = _START.get( int( self.stats[ "start" ] ), "?" )
        for name in ( "nb_iter", "nb_diag", "nb_backtracks", "nb_overflowed", "nb_cell_lim", "nb_limit_rounds", "lin_nb_hierarchies",
                      "lin_nb_iter", "nb_continuation_steps", "scratch_bytes", "nb_clusters", "nb_aggregated", "nb_duplicates", "nb_polish" ):
            self.stats[ name ] = int( self.stats[ name ] )
        if not settings.aggregate:
            self.stats[ "aggregation" ] = "off"
        elif self.stats[ "nb_clusters" ] == 0:
            self.stats[ "aggregation" ] = "none"
        else:
            self.stats[ "aggregation" ] = ( f"{ self.stats[ 'nb_aggregated' ] } seeds in { self.stats[ 'nb_clusters' ] } clusters "
                                            f"( { self.stats[ 'nb_duplicates' ] } exact duplicates ), re-split in { self.stats[ 'nb_polish' ] } diagrams" )
        self.stats[ "duplicates" ] = self._dup_note
        self.stats[ "warm_start" ] = self._warm_start

    def _read_history( self, rows ):
        self.history = []
        for r in rows:
            entry = { name: float( r[ k ] ) for k, name in enumerate( _HISTORY ) }
            entry[ "step" ] = int( entry[ "step" ] )
            self.history.append( entry )

    def _check_converged( self, settings ):
        if getattr( settings, "on_failure", "raise" ) != "raise":
            return
        if not self.converged:
            from .OtProblem import OtNotConverged
            raise OtNotConverged( self )

    @property
    def converged( self ):
        return self.stats[ "status" ] in ( "converged", "converged (aggregated)" )

    @property
    def clusters( self ):
        return None if self.stats[ "nb_clusters" ] == 0 else self._clusters

    @property
    def target_masses( self ):
        """`nu` rescaled to the mass of the domain"""
        return self._masses * ( self.stats[ "domain_mass" ] / float( self._masses.sum() ) )

    def power_diagram( self, weights = None ):
        from .PowerDiagram import PowerDiagram
        return PowerDiagram( self.problem.source.positions, self._weights_fit if weights is None else weights, distribution = self.problem.target )
