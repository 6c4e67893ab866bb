"""sdot on MI355X: the semi-discrete optimal transport solve of the reference `sdot`
( `OtProblem( SumOfDiracs, density ).solve( Iterative( ... ) )` -> `SdotPlanNd` ), with its
power diagrams, on hand-written HIP kernels behind the C ABI of `include/sdot_amd.h`.
"""

from .distributions import Image, Polytope, SumOfDiracs, SumOfGaussians, UniformBox
from .OtProblem import Direct, Iterative, OtNotConverged, OtProblem, Tuning, ot_solve
from .PowerDiagram import PowerDiagram, axis_aligned_box, box_half_spaces
from .SdotPlanNd import SdotPlanNd
from ._native import KernelFailure, SdotError
from .batch import solve_batch
from .serve import Server
from .quantization import quantization
from .StartCell import start_cell

__all__ = [ "Image", "Polytope", "start_cell", "SumOfDiracs", "SumOfGaussians", "UniformBox", "Direct", "Iterative", "OtNotConverged", "OtProblem", "Tuning",
            "ot_solve", "PowerDiagram", "axis_aligned_box", "box_half_spaces", "SdotPlanNd", "KernelFailure", "SdotError",
            "solve_batch", "Server", "quantization" ]
