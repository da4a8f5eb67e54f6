"""ctypes binding of libflowreg_hip.so (declarations in include/flowreg_hip.h).

The library is loaded from this package directory only; if it is missing or
cannot be loaded, every entry point raises -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflowreg_hip.so")

_lib = None

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
c_size = ctypes.c_size_t
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_f32p = ctypes.POINTER(ctypes.c_float)
c_f64p = ctypes.POINTER(ctypes.c_double)


class LevelDesc(ctypes.Structure):
    _fields_ = [
        ("h", c_int), ("w", c_int),
        ("h_row", c_double), ("h_col", c_double),
        ("alpha_x", c_double), ("alpha_y", c_double),
        ("P_img", c_int),
        ("img_ix", c_void_p), ("img_wx", c_void_p), ("img_iy", c_void_p), ("img_wy", c_void_p),
        ("P_up", c_int),
        ("up_ix", c_void_p), ("up_wx", c_void_p), ("up_iy", c_void_p), ("up_wy", c_void_p),
    ]


class SolverDesc(ctypes.Structure):
    _fields_ = [
        ("H", c_int), ("W", c_int), ("C", c_int),
        ("n_levels", c_int), ("levels", ctypes.POINTER(LevelDesc)),
        ("iterations", c_int), ("update_lag", c_int),
        ("a_smooth", c_double), ("a_data", c_void_p),
        ("max_batch", c_int),
        ("const_assumption", c_int),
    ]


# name -> (restype, argtypes); every entry point declared in the header.
SIGNATURES = {
    "frh_last_error": (ctypes.c_char_p, []),
    "frh_abi_version": (c_int, []),
    "frh_fill_resize_tables": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "frh_warping_depth": (c_int, [c_double, c_int, c_int, c_int]),
    "frh_resize": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_int, c_void_p,
                           c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "frh_warp_cubic": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p,
                               c_void_p, c_double, c_double, c_void_p, c_void_p]),
    "frh_warp_cubic_ex": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "frh_warp_linear_ex": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                   c_void_p, c_void_p, c_void_p, c_void_p]),
    "frh_warp_int16": (c_int, [c_int, c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p,
                               c_void_p, c_void_p, c_void_p]),
    "frh_warp_u8": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                            c_void_p, c_void_p]),
    "frh_prefilter_scratch_bytes": (c_size, [c_int, c_int, c_int, c_int, c_int]),
    "frh_prefilter_ex": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                 c_int, c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_size, c_void_p]),
    "frh_prefilter_noalloc": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                      c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_int, c_void_p]),
    "frh_prefilter_noalloc_max_bytes": (c_size, []),
    "frh_prefilter_3d": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                 c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_size,
                                 c_void_p]),
    "frh_motion_tensor": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_double, c_double,
                                  c_double, c_int, c_int] + [c_void_p] * 6 + [c_void_p]),
    "frh_level_solver": (c_int, [c_void_p] * 9 + [c_int, c_int, c_int, c_double, c_double, c_int,
                                                  c_int, c_void_p, c_double, c_double, c_double,
                                                  c_double, c_void_p, c_void_p, c_void_p, c_size,
                                                  c_void_p]),
    "frh_level_solver_scratch_bytes": (c_size, [c_int, c_int, c_int, c_double]),
    "frh_median5_mirror": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_size, c_void_p]),
    "frh_median5_mirror_scratch_bytes": (c_size, [c_int, c_int, c_int]),
    "frh_mean_frames_f32": (c_int, [c_void_p, c_int, ctypes.c_size_t, c_void_p, c_void_p]),
    "frh_resize_cubic_cv": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_int,
                                    c_void_p, c_size, c_void_p]),
    "frh_resize_cubic_cv_scratch_bytes": (c_size, [c_int, c_int, c_int, c_int]),
    "frh_solver_create": (c_int, [ctypes.POINTER(c_void_p), ctypes.POINTER(SolverDesc)]),
    "frh_solver_destroy": (None, [c_void_p]),
    "frh_solver_set_reference": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "frh_solver_run": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "frh_solver_run_pairs": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "frh_worklist_create": (c_int, [ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_int]),
    "frh_worklist_destroy": (None, [c_void_p]),
    "frh_worklist_run": (c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                 ctypes.POINTER(c_int), ctypes.POINTER(c_void_p), c_int, c_void_p]),
    "frh_solver_profile": (c_int, [c_void_p, c_int]),
    "frh_solver_set_graph": (c_int, [c_void_p, c_int]),
    "frh_solver_set_chains": (c_int, [c_void_p, c_int]),
    "frh_solver_level_info": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "frh_solver_set_pass_limit": (c_int, [c_void_p, c_int, c_int]),
    "frh_grad_skew_plane": (c_int, [c_int, c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "frh_grad_skew": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p]),
    "frh_grad_tensor": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
                                c_void_p, c_void_p, c_void_p, c_void_p]),
    "frh_grad_coef": (c_int, [c_void_p] * 4 + [c_int] * 4 + [c_double, c_double, c_void_p, c_int]
                      + [c_void_p] * 7),
    "frh_grad_sweep": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double,
                               c_void_p]),
    "frh_grad_sweeps_scratch_bytes": (ctypes.c_size_t, [c_int, c_int]),
    "frh_grad_sweeps": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t, c_int, c_int, c_int, c_int,
                                c_int, c_double, c_double, c_void_p, ctypes.c_size_t, c_void_p]),
    "frh_grad_sweep_adjoint": (c_int, [c_void_p] * 6 + [c_int, c_void_p]
                               + [c_int, c_int, c_int, c_double, c_double, c_void_p]),
    "frh_grad_sweeps_adjoint_scratch_bytes": (ctypes.c_size_t, [c_int, c_int, c_int, c_int]),
    "frh_grad_sweeps_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_size_t, c_int, c_void_p, c_void_p,
                                        c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_double, c_double,
                                        c_void_p, ctypes.c_size_t, c_void_p]),
    "frh_grad_median": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "frh_grad_warp": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "frh_solver_timings": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
}


class FlowregHipError(RuntimeError):
    """A libflowreg_hip entry point returned an error code."""


def lib() -> ctypes.CDLL:
    """Load (once) and return the library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FlowregHipError(
            f"{LIB_PATH} is missing: build it with `python -m pyflowreg_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().frh_last_error().decode(errors="replace")
        raise FlowregHipError(f"{what} failed (code {rc}): {msg}")


def ptr(t) -> int:
    """Device (or host) address of a torch tensor / numpy array."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def scratch(nbytes: int, device=None):
    """Caller-owned scratch for an entry point that takes one: a device
    byte tensor from torch's caching allocator (stream-ordered reuse, no
    device allocation once the pool holds a block of that size); keep it
    alive until the call's work is queued."""
    import torch
    return torch.empty((max(int(nbytes), 1),), dtype=torch.uint8,
                       device=device if device is not None else torch.device("cuda", torch.cuda.current_device()))


def stream_handle(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
