"""Backend registry with pyflowreg's interface (core/backend_registry.py:
12-105), plus the ``flowreg_hip`` backend factory.

When pyflowreg itself is importable, ``register(pyflowreg.core)`` adds
``flowreg_hip`` to pyflowreg's own registry; otherwise this module's
registry (same functions, same errors) is used.
"""

from __future__ import annotations

from typing import Callable, Dict, Optional, Set

_ALL_EXECUTORS = {"sequential", "threading", "multiprocessing", "multiprocessing_fork",
                  "multiprocessing_spawn"}

_BACKENDS: Dict[str, Callable[..., Callable]] = {}
_BACKEND_EXECUTORS: Dict[str, Set[str]] = {}


def register_backend(name: str, factory: Callable[..., Callable],
                     supported_executors: Optional[Set[str]] = None) -> None:
    """core/backend_registry.py:19-47."""
    _BACKENDS[name] = factory
    if supported_executors is None:
        supported_executors = set(_ALL_EXECUTORS)
    _BACKEND_EXECUTORS[name] = supported_executors


def get_backend(name: str) -> Callable[..., Callable]:
    """core/backend_registry.py:50-68."""
    if name not in _BACKENDS:
        available = list(_BACKENDS.keys())
        raise ValueError(f"Unknown flow backend: '{name}'. Available backends: {available}")
    return _BACKENDS[name]


def list_backends() -> list:
    return list(_BACKENDS.keys())


def is_backend_available(name: str) -> bool:
    return name in _BACKENDS


def get_backend_executors(name: str) -> Set[str]:
    """core/backend_registry.py:81-105."""
    if name not in _BACKEND_EXECUTORS:
        if name in _BACKENDS:
            return set(_ALL_EXECUTORS)
        raise ValueError(f"Unknown flow backend: '{name}'")
    return _BACKEND_EXECUTORS[name].copy()


SUPPORTED_EXECUTORS = {"sequential", "hip_batched"}


def factory(**backend_params) -> Callable:
    """Backend factory: returns ``get_displacement`` (the caller passes the
    flow options as keyword arguments, like the flowreg backend,
    core/__init__.py:72-77).  Unknown backend_params are ignored, as the
    torch and CuPy factories do (core/__init__.py:108,218)."""
    from .optical_flow import get_displacement
    from ._lib import lib
    lib()  # fail here, loudly, if the library is not built
    return get_displacement


def register(pyflowreg_core=None) -> bool:
    """Register ``flowreg_hip`` into pyflowreg's registry when given its
    ``pyflowreg.core`` module (or into this module's registry otherwise),
    if libflowreg_hip.so loads: like the reference's GPU backends, which are
    registered only when their library imports (core/__init__.py:97-104,
    199-213), a backend whose library is missing is simply not registered.
    Returns whether it was registered."""
    from ._lib import lib
    try:
        lib()
    except (OSError, RuntimeError):
        return False
    target = pyflowreg_core if pyflowreg_core is not None else None
    reg = getattr(target, "register_backend", None) if target is not None else register_backend
    reg("flowreg_hip", factory, supported_executors=set(SUPPORTED_EXECUTORS))
    return True


register()
