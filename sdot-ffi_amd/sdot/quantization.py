"""Optimal quantization of a density by Lloyd iterations in their transport form ( the reference's
`src/sdot/quantization.py`: same function, arguments, stopping rule and warnings ).

    pos = sdot.quantization( target, pos0 )      # from given starting positions
    pos = sdot.quantization( target, 1000 )      # 1000 Diracs drawn uniformly in the support of `target`

Each iteration solves the transport from the Diracs ( their masses kept ) to `target`, then moves every Dirac to the
barycentre of its cell: a gradient step of size `1 / ( 2 m_i )` on `W_2^2`, which decreases. The loop stops when the root
mean square move is under `tol` times `sqrt( W_2^2 )` ( the typical distance from a point of the target to its Dirac: scale
free, independent of `n` and of the dimension ).

The problem is kept from one iteration to the next, so each solve starts from the previous weights ( a few Newton steps,
not a cold solve ); the transports are solved to `1e-4 / n` ( relative to a Dirac's mass ) unless `ot_tol` says otherwise, and
a solve that stagnates above it is used all the same ( counted, then reported in one warning ). Everything stays on the
device: the positions, the barycentres, the weights.
"""

import numbers
import warnings

import numpy as np

from .OtProblem import Iterative, OtNotConverged, OtProblem
from .distributions import SumOfDiracs


def uniform_in_support( target, n, seed = None ):
    """`n` points uniform in the support of `target` ( its `bounding_half_spaces`: a box here ), `numpy.random.default_rng( seed )`,
    as a host `[ n, d ]` array"""
    support = target.bounding_half_spaces()
    if support is None:
        raise ValueError( "sdot.quantization: the target declares no support ( `bounding_half_spaces` is None ), the starting positions "
                          "have to be given" )
    dirs, offs = ( np.asarray( a, dtype = np.float64 ) for a in support )
    offs = offs.reshape( -1 )
    lo, hi = target.box()
    lo, hi = np.asarray( lo, dtype = np.float64 ), np.asarray( hi, dtype = np.float64 )
    rng = np.random.default_rng( seed )
    out = np.zeros( ( 0, len( lo ) ) )
    while len( out ) < n:                                # ( rejection in the bounding box: exact for a box support )
        x = rng.uniform( lo, hi, ( 2 * ( n - len( out ) ) + 16, len( lo ) ) )
        out = np.concatenate( [ out, x[ np.all( x @ dirs.T <= offs + 1e-15 * np.abs( offs ), axis = 1 ) ] ] )
    return out[ :n ]


def quantization( target, positions, masses = None, *, tol = 2e-2, max_iter = 200, callback = None, verbose = False, ot_tol = None,
                  seed = None, **settings ):
    """the positions ( a device `[ n, d ]` tensor ) of Diracs that quantize `target` optimally, by Lloyd iterations from
    `positions` ( `[ n, d ]`, or a number of Diracs drawn uniformly in the support of `target` with `default_rng( seed )` ).

    `masses`: `[ n ]`, kept during the iterations ( equal by default ). `tol`: the stopping rule, the rms move under
    `tol sqrt( W_2^2 )`. `max_iter`: iterations at most ( a warning when reached ). `callback( iteration, positions, plan )`
    after each solve, with the positions the plan was solved for ( `iteration = 0`: the start ). `settings`: those of
    `Iterative`, but its `tol`, which is `ot_tol` here ( `1e-4 / n` by default )."""
    import torch
    if isinstance( positions, numbers.Integral ):
        positions = uniform_in_support( target.normalized_version(), int( positions ), seed )
    pb = OtProblem( SumOfDiracs( positions, masses ), target )
    n = int( pb.source.nb_diracs )
    if pb.regime != "iterative":
        raise NotImplementedError( "sdot.quantization: the direct regime ( d = 1 ) is not part of the MI355X path" )
    if "tol" in settings:
        raise TypeError( "sdot.quantization: the transports' tolerance is `ot_tol` ( `tol` is the stopping rule of the iterations )" )
    solver = Iterative( tol = 1e-4 / n if ot_tol is None else ot_tol, **settings )
    dev = torch.device( "cuda", torch.cuda.current_device() )
    x = torch.as_tensor( pb.source.positions, dtype = torch.float64, device = dev )
    stagnated = []
    ratio = float( "inf" )
    plan = None
    iteration = 0
    for iteration in range( max_iter + 1 ):
        try:
            plan = pb.solve( solver, verbose )
        except OtNotConverged as e:
            if e.solution.stats[ "status" ] != "stagnation":
                raise
            plan = e.solution
            stagnated.append( plan.stats[ "residual" ] )
        if callback is not None:
            callback( iteration, x, plan )
        b = plan.barycenters
        cost = float( plan.cost )
        move2 = float( ( ( b - x ) ** 2 ).sum() ) / n
        ratio = ( move2 / cost ) ** 0.5 if cost > 0 else 0.0
        if ratio <= tol or iteration == max_iter:
            break
        pb.source = SumOfDiracs( b, masses )              # ( the same count: the next solve starts from these weights )
        x = b
    if stagnated:
        warnings.warn( f"sdot.quantization: { len( stagnated ) } of { iteration + 1 } transports stagnated above their tolerance ( worst "
                       f"residual { max( stagnated ):.3g}, for a Dirac mass of { 1 / n:.3g} ) -- their solutions were used all the same",
                       stacklevel = 2 )
    if ratio > tol:
        warnings.warn( f"sdot.quantization: { max_iter } iterations without reaching tol = { tol } ( the last move is { ratio:.3g} times the "
                       "cell size )", stacklevel = 2 )
    return plan.barycenters
