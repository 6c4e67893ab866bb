"""WHAT A SOLVE TAKES FROM THE DEVICE, and the check made before it starts.

The reference models what its card's solve takes, per buffer, and raises a `MemoryError` before launching a solve that
cannot fit ( `src/sdot/CardMemory.py:47-190`: `card_solve_bytes`, `check_card_memory` ). Here the model is the library's
own ( `sdot_amd_solve_bytes`, which follows the arena's takes of a solve buffer by buffer ) and the free memory is the
device's ( `torch.cuda.mem_get_info` ) plus what the library's arena already holds and would reuse. The server admits its
requests by the same model.
"""

from . import _native


def solve_bytes( n, dim = 2, density = None, settings = None, precision = None ):
    """the bytes a solve of `n` seeds takes, PER PART ( `{ part: bytes }`, the parts of `_native.MEMORY_PARTS` ); `density`
    the target ( its kind, and an image's size ), `settings` an `Iterative` ( the step, the linear solver, the continuation ),
    `precision` its 'auto' / 'fp64' / 'fp32' / 'mixed' ( default: the settings' )"""
    from .OtProblem import Iterative
    from .SdotPlanNd import _KERNEL_FLOAT, card_options
    from .distributions import UniformBox
    settings = settings if settings is not None else Iterative()
    density = density if density is not None else UniformBox( [ 0 ] * dim, [ 1 ] * dim )
    opts = card_options( settings, density )
    kind = { "Image": _native.IMAGE, "SumOfGaussians": _native.GAUSSIANS }.get( type( density ).__name__, _native.CONSTANT )
    if dim == 3:
        opts[ _native.CARD_OPTIONS.index( "step" ) ] = 0.0
        prec = _native.FP64
    else:
        prec = _KERNEL_FLOAT[ precision or settings.precision ]
    return _native.solve_bytes( n, dim, kind, opts, prec )


def check_memory( n, dim = 2, density = None, settings = None, precision = None ):
    """raises a `MemoryError` when the solve of `n` seeds ( `solve_bytes` ) does not fit in the device's free memory and
    what the library already holds; returns the parts otherwise"""
    parts = solve_bytes( n, dim, density, settings, precision )
    need = sum( parts.values() )
    free, total = _native.torch().cuda.mem_get_info()
    have = free + _native.device_bytes()
    if need <= have:
        return parts
    top = ", ".join( f"{ k } { v / n :.0f}" for k, v in sorted( parts.items(), key = lambda kv: -kv[ 1 ] )[ :4 ] )
    raise MemoryError( f"the solve of { n } seeds needs about { need / 1e9 :.2f} GB of the device ( { need / n :.0f} bytes per seed; the "
                       f"largest parts, bytes per seed: { top } ) and { have / 1e9 :.2f} GB are free or held by the library "
                       f"( { total / 1e9 :.0f} GB in all ): solve fewer seeds, or free what the program keeps on the device" )
