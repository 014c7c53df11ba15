"""MI355X-native NFGS plain iterate: the NFGSSolverGPU drop-in over gfx950 HIP kernels."""
from .solver import NFGSSolverGPU

__all__ = ["NFGSSolverGPU"]
__version__ = "0.1.0"
