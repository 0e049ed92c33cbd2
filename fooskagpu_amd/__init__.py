"""MI355X (gfx950) explicit unstructured finite-volume step of
sDamianos/FooskaGPU: host mesh/case input and driver over a C ABI of
hand-written HIP kernels (include/fooska.h, include/fooska_mesh.h)."""
from . import _lib, cases
from ._lib import (FP64, MIXED, MODEL_SINGLE, MODEL_TWO_PHASE, RELAX_COLD, RELAX_MLP, RELAX_PLAIN, FkError,
                   FkParams)
from .mesh import Mesh
from .solver import GpuSolver

__all__ = ["FP64", "MIXED", "MODEL_SINGLE", "MODEL_TWO_PHASE", "RELAX_PLAIN", "RELAX_MLP", "RELAX_COLD", "FkError",
           "FkParams",
           "Mesh", "GpuSolver", "cases"]
