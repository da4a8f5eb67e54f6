"""The backend registry (core/backend_registry.py:12-105 interface) and the
registration rule: ``flowreg_hip`` is registered only when its library
loads, as the reference registers its GPU backends only when their library
imports (core/__init__.py:97-104,199-213).  CPU only (the library is loaded,
no compute entry point is called)."""

import types

import pytest

from pyflowreg_amd import backend_registry as R
from pyflowreg_amd import _lib


def test_registry_interface_and_errors():
    R.register_backend("test_dummy", lambda **p: (lambda f, m, **kw: None), supported_executors={"sequential"})
    assert R.is_backend_available("test_dummy") and "test_dummy" in R.list_backends()
    assert R.get_backend_executors("test_dummy") == {"sequential"}
    with pytest.raises(ValueError, match="Unknown flow backend"):
        R.get_backend("no_such_backend")
    with pytest.raises(ValueError, match="Unknown flow backend"):
        R.get_backend_executors("no_such_backend")
    # a backend registered without executors supports all of the reference's
    R.register_backend("test_dummy_all", lambda **p: None)
    assert R.get_backend_executors("test_dummy_all") == R._ALL_EXECUTORS


def test_flowreg_hip_registered_with_its_executors():
    assert R.is_backend_available("flowreg_hip")
    assert R.get_backend_executors("flowreg_hip") == {"sequential", "hip_batched"}


def test_register_into_an_external_registry():
    seen = {}
    core = types.SimpleNamespace(register_backend=lambda n, f, supported_executors=None: seen.update(
        {n: (f, supported_executors)}))
    assert R.register(core) is True
    assert seen["flowreg_hip"] == (R.factory, {"sequential", "hip_batched"})


def test_missing_library_is_not_registered(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libflowreg_hip.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    seen = {}
    core = types.SimpleNamespace(register_backend=lambda n, f, supported_executors=None: seen.update({n: f}))
    assert R.register(core) is False
    assert seen == {}
    # a compute path still fails loudly rather than falling back
    with pytest.raises(_lib.FlowregHipError, match="is missing"):
        R.factory()
