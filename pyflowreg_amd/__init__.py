"""pyflowreg_amd -- MI355X (gfx950) backend for pyflowreg's variational
optical-flow path (``get_displacement``), built on libflowreg_hip.so."""

from .backend_registry import (factory, get_backend, get_backend_executors, is_backend_available,
                               list_backends, register, register_backend)
from .live import FlowRegLive
from .of_options import OFOptions
from .compensate import compensate_arr
from .xcorr import estimate_rigid_xcorr_2d, estimate_rigid_xcorr_batch
from .optical_flow import (DisplacementSolver, WorkList, get_displacement, get_displacements, get_motion_tensor_cs,
                           get_motion_tensor_gc, get_motion_tensor_gray, imregister_wrapper,
                           imresize_fused_gauss_cubic, level_solver, normalize_gnc_schedule,
                           normalize_warping_steps, warpingDepth)

__all__ = [
    "DisplacementSolver", "WorkList", "FlowRegLive", "OFOptions", "compensate_arr", "get_displacement", "get_displacements", "get_motion_tensor_gc", "get_motion_tensor_gray",
    "get_motion_tensor_cs", "normalize_gnc_schedule", "normalize_warping_steps", "imregister_wrapper",
    "imresize_fused_gauss_cubic", "level_solver", "warpingDepth", "factory", "register",
    "register_backend", "get_backend", "get_backend_executors", "list_backends",
    "is_backend_available", "estimate_rigid_xcorr_2d", "estimate_rigid_xcorr_batch",
]
