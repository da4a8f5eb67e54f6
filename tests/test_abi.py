"""The C ABI library loads and exports every entry point include/*.h declares
(no compute calls: runs without a GPU)."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h") and fn != "flowreg_detmath.h":
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(frh_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    from pyflowreg_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libflowreg_hip.so is not built (run __graft_entry__.build())")
    L = _lib.lib()
    declared = _declared()
    assert len(declared) >= 15
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) == declared


def test_host_entry_points_without_gpu():
    import numpy as np
    from pyflowreg_amd import _lib
    L = _lib.lib()
    assert L.frh_abi_version() == 10
    assert L.frh_warping_depth(0.8, 50, 512, 512) == 18
    g = np.array([1.0], np.float32)
    idx = np.zeros((5, 4), np.int32)
    wt = np.zeros((5, 4), np.float32)
    assert L.frh_fill_resize_tables(5, 5, 0, g.ctypes.data, idx.ctypes.data, wt.ctypes.data) == 0
    np.testing.assert_array_equal(wt[:, 2], 1.0)  # scale 1 is the identity
    rc = L.frh_fill_resize_tables(0, 5, 0, g.ctypes.data, idx.ctypes.data, wt.ctypes.data)
    assert rc == -1 and b"bad arguments" in L.frh_last_error()


def test_scratch_queries_and_rejection_without_gpu():
    """The stage entry points take caller scratch: the queries give its size
    and a missing or short buffer is rejected before any device work."""
    from pyflowreg_amd import _lib
    L = _lib.lib()
    assert L.frh_prefilter_scratch_bytes(4, 16, 20, 2, 0) == 8 * 4 * 16 * 20 * 2
    assert L.frh_prefilter_scratch_bytes(4, 16, 20, 2, 3) == 2 * 8 * 4 * 16 * 20 * 2
    assert L.frh_median5_mirror_scratch_bytes(3, 10, 12) == 8 * 3 * 10 * 12
    assert L.frh_resize_cubic_cv_scratch_bytes(2, 10, 30, 2) == 8 * 2 * 10 * 30 * 2
    # (m-2) x (n-2) = 14 x 18 interior: 64-row skew, 25 planes at K = 2 (12 J,
    # 2 weights, cu/cv, 5 coefficients, du/dv, the ring)
    assert L.frh_level_solver_scratch_bytes(16, 20, 2, 1.0) == 8 * 25 * 64 * (14 + 18 - 1)
    assert L.frh_level_solver_scratch_bytes(16, 20, 2, 0.5) == 8 * (29 * 64 * (14 + 18 - 1) + 16 * 20)
    dummy = 1 << 20  # never dereferenced: the checks fail first
    rc = L.frh_median5_mirror(dummy, 3, 10, 12, None, 0, None)
    assert rc == -1 and b"scratch of 2880 bytes needed, 0 given" in L.frh_last_error()
    rc = L.frh_median5_mirror(dummy, 3, 10, 12, dummy, 2879, None)
    assert rc == -1 and b"2879 given" in L.frh_last_error()


def test_pipelined_sweeps_reject_overlap_without_gpu():
    """frh_grad_sweeps with every state written: a stride under one state or
    an input inside the outputs is rejected before any device work."""
    from pyflowreg_amd import _lib
    L = _lib.lib()
    base = 1 << 24  # never dereferenced: the checks fail first
    state = 2 * 2 * 64 * (16 + 20 - 1)  # T = 2 frames, 2 planes of the 16 x 20 skew
    rc = L.frh_grad_sweeps(base, base + 8 * 10**6, base + 16 * 10**6, state - 1, 3, 1, 2, 16, 20, 1.0, 1.0, None, 0, None)
    assert rc == -1 and b"out_stride" in L.frh_last_error()
    out = base + 8 * 10**6
    rc = L.frh_grad_sweeps(base, out + 8 * state, out, state, 3, 1, 2, 16, 20, 1.0, 1.0, None, 0, None)
    assert rc == -1 and b"overlap" in L.frh_last_error()
    rc = L.frh_grad_sweeps(base, out, out, state, 0, 0, 2, 16, 20, 1.0, 1.0, None, 0, None)
    assert rc == -1 and b"bad arguments" in L.frh_last_error()
    # chained: a failure word and one 256-byte flag per stage and padded frame
    assert L.frh_grad_sweeps_scratch_bytes(5, 2) == 256 + 5 * 8 * 256
    assert L.frh_grad_sweeps_scratch_bytes(40, 9) == 256 + 32 * 16 * 256
    rc = L.frh_grad_sweeps(base, out, out, 0, 5, 0, 2, 16, 20, 1.0, 1.0, base, 256 + 5 * 8 * 256 - 1, None)
    assert rc == -1 and b"scratch of" in L.frh_last_error()


def test_chained_adjoint_scratch_without_gpu():
    """frh_grad_sweeps_adjoint's scratch: a failure word, one 256-byte flag per
    stage and padded frame, the lam planes between stages (and one hand-over
    plane above 32 sweeps); a short buffer is rejected before device work."""
    from pyflowreg_amd import _lib
    L = _lib.lib()
    state = 3 * 2 * 64 * (16 + 20 - 1)  # T = 3 frames (8 padded), 2 planes of the 16 x 20 skew
    assert L.frh_grad_sweeps_adjoint_scratch_bytes(5, 3, 16, 20) == 256 + 5 * 8 * 256 + 4 * state * 8
    assert L.frh_grad_sweeps_adjoint_scratch_bytes(1, 3, 16, 20) == 256 + 8 * 256
    assert L.frh_grad_sweeps_adjoint_scratch_bytes(40, 3, 16, 20) == 256 + 32 * 8 * 256 + 32 * state * 8
    assert L.frh_grad_sweeps_adjoint_scratch_bytes(5, 3, 2000, 20) == 0
    base = 1 << 24  # never dereferenced: the checks fail first
    need = L.frh_grad_sweeps_adjoint_scratch_bytes(5, 3, 16, 20)
    args = [base] * 3 + [state, 5, base, base + 8 * state, base, 1, base, 3, 16, 20, 1.0, 1.0, base]
    rc = L.frh_grad_sweeps_adjoint(*args, need - 1, None)
    assert rc == -1 and b"scratch of" in L.frh_last_error()
    args[6] = base  # lam_bottom aliases lam_top
    rc = L.frh_grad_sweeps_adjoint(*args, need, None)
    assert rc == -1 and b"alias" in L.frh_last_error()


def test_get_displacement_raises_without_gpu():
    import numpy as np
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import pyflowreg_amd as fr
    with pytest.raises(RuntimeError, match="GPU"):
        fr.get_displacement(np.zeros((16, 16)), np.zeros((16, 16)))


def test_unknown_constancy_assumption_is_value_error():
    import numpy as np
    import pyflowreg_amd as fr
    with pytest.raises(ValueError, match="Unknown constancy assumption"):
        fr.get_displacement(np.zeros((8, 8)), np.zeros((8, 8)), const_assumption="nope")


def test_gnc_options_are_validated_before_gpu_work():
    import numpy as np
    import pyflowreg_amd as fr
    f = np.zeros((8, 8))
    for bad in ([0.5, 1.0], [0.0], [0.0, 0.7, 0.5, 1.0], [[0.0, 1.0]]):
        with pytest.raises(ValueError):
            fr.get_displacement(f, f, gnc_schedule=bad)
    with pytest.raises(ValueError, match="warping_steps"):
        fr.get_displacement(f, f, gnc_schedule=[0.0, 1.0], warping_steps=0)
    np.testing.assert_array_equal(fr.normalize_gnc_schedule((0, 0.5, 1)), [0.0, 0.5, 1.0])
    assert fr.normalize_gnc_schedule(None) is None and fr.normalize_warping_steps(None) is None
