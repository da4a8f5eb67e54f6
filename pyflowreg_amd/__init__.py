"""pyflowreg_amd -- MI355X (gfx950) backend for pyflowreg's variational
optical-flow path (``get_displacement``), built on libflowreg_hip.so."""

from .backend_registry import (factory, get_backend, get_backend_executors, is_backend_available,
                               list_backends, register, register_backend)
from .optical_flow import (DisplacementSolver, get_displacement, get_motion_tensor_gc,
                           imregister_wrapper, imresize_fused_gauss_cubic, level_solver,
                           warpingDepth)

__all__ = [
    "DisplacementSolver", "get_displacement", "get_motion_tensor_gc", "imregister_wrapper",
    "imresize_fused_gauss_cubic", "level_solver", "warpingDepth", "factory", "register",
    "register_backend", "get_backend", "get_backend_executors", "list_backends",
    "is_backend_available",
]
