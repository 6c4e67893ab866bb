"""MANY TRANSPORTS AT ONCE: `solve_batch( problems, settings )` solves independent `OtProblem`s of uneven sizes
together on the GPU and returns one `SdotPlanNd` per problem.

The reference solves one problem per call; a batch has no counterpart there. The instances run on several HIP
streams ( `streams` ), largest first, each stream driven by its own host thread with its own device buffers, so the
chip stays busy while each solve waits on its own decisions; an instance's result does not depend on what ran beside
it ( the kernels are deterministic, nothing is shared ). One call of `sdot_amd_solve_2d_batch`.

The cells' float follows `settings.precision` as a single solve's does ( 'auto' and 'mixed': the FP32 kernel, then the FP64
one from the first step it cannot take ), so each plan is the one `problem.solve( settings )` returns. What a batched plan does
not carry: the per-step `history` ( empty ) and the exact duplicates' merge ( `aggregate` is done by single solves ).
"""

import numpy as np

from . import _native
from .distributions import Image, SumOfGaussians, UniformBox
from .SdotPlanNd import SdotPlanNd, card_options


def solve_batch( problems, settings = None, streams = 8 ):
    """`problems`: `OtProblem`s ( 2D, every target of the same kind: `Image`, `UniformBox` or `SumOfGaussians` ),
    `settings`: one `Iterative` for all. Returns the list of plans, in the order of `problems`. A problem that did not
    converge is returned as such ( `plan.converged`, `plan.stats[ "status" ]` ), unless `settings.on_failure` is
    "raise", in which case the first one raises `OtNotConverged` after the whole batch ran."""
    from .OtProblem import Iterative, OtNotConverged
    settings = settings if settings is not None else Iterative()
    problems = list( problems )
    if not problems:
        return []
    instances = []
    for pb in problems:
        if pb.nb_dims != 2:
            raise NotImplementedError( f"solve_batch: 2D problems on the MI355X path ( not { pb.nb_dims }D )" )
        src, dst = pb.source, pb.target
        if not isinstance( dst, ( Image, UniformBox, SumOfGaussians ) ):
            raise NotImplementedError( f"solve_batch: an Image, a UniformBox or a SumOfGaussians target ( not { type( dst ).__name__ } )" )
        pb.domain
        lo, hi = dst.box()
        w0 = settings.weights0 if settings.weights0 is not None else None
        instances.append( ( src.positions, src.weights, w0, [ lo[ 0 ], lo[ 1 ], hi[ 0 ], hi[ 1 ] ], dst ) )
    opts = card_options( settings, problems[ 0 ].target )
    from .SdotPlanNd import _KERNEL_FLOAT
    res = _native.solve_2d_batch( instances, opts, streams = streams, precision = _KERNEL_FLOAT[ settings.precision ] )
    plans = []
    for pb, r in zip( problems, res ):
        if r[ "error" ][ 0 ] != 0:
            raise _native.SdotError( f"solve_batch: an instance failed ( kind { r[ 'error' ][ 0 ] }, code { r[ 'error' ][ 1 ] }, "
                                     f"value { r[ 'error' ][ 2 ] } )", *r[ "error" ] )
        plan = SdotPlanNd.__new__( SdotPlanNd )
        plan.problem = pb
        plan.weights = r[ "weights" ]
        plan._weights_fit = plan.weights
        plan.cell_masses = r[ "cell_masses" ]
        plan.barycenters = r[ "barycenters" ]
        plan.cost = r[ "cost" ]
        plan._clusters = None
        plan._masses = np.asarray( pb.source.weights, dtype = np.float64 )
        plan._warm_start = "weights0" if settings.weights0 is not None else "none"
        plan._dup_note = "not checked ( batch )"
        plan._read_stats( r[ "stats" ], settings )
        plan.stats[ "aggregation" ] = "off ( batch )"
        plan.history = []
        plan.batch_times = r[ "times" ]
        plans.append( plan )
    if getattr( settings, "on_failure", "raise" ) == "raise":
        for plan in plans:
            if not plan.converged:
                raise OtNotConverged( plan )
    return plans
