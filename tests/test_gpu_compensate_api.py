"""compensate_arr's argument surface against the reference's
(motion_correction/compensate_arr.py:16-231, compensate_recording.py:
200-267,361-440,786-867): callback schedule, displacement overrides,
executor selection, output dtype and input shapes."""

import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H, W, C = 32, 28, 2


def _video(T, seed, C=C):
    from pyflowreg_amd import synth
    v, _ = synth.video(T, H, W, n_channels=C, max_disp=2.0, sigma_d=8.0, seed=seed)
    return v.numpy()


def _opt(**kw):
    from pyflowreg_amd.compensate import RegistrationOptions
    base = dict(alpha=(2.0, 2.0), buffer_size=10, sigma=(1.0, 1.0, 0.1), iterations=20)
    base.update(kw)
    return RegistrationOptions(**base)


def test_callback_schedule_matches_reference(gpu):
    """26 frames in batches of 10: after each batch, progress(current,
    total), then registered_callback(frames, start, end), then
    w_callback(w, start, end); the w_init pre-pass reports nothing."""
    from pyflowreg_amd.compensate import compensate_arr
    v = _video(27, 41)
    calls = []
    reg, w = compensate_arr(
        v[1:], v[0], _opt(),
        progress_callback=lambda cur, tot: calls.append(("progress", cur, tot)),
        registered_callback=lambda b, s, e: calls.append(("registered", s, e, b.shape, b.dtype, b.copy())),
        w_callback=lambda ww, s, e: calls.append(("w", s, e, ww.shape, ww.dtype, ww.copy())))
    kinds = [c[0] for c in calls]
    assert kinds == ["progress", "registered", "w"] * 3
    assert [c[1:3] for c in calls if c[0] == "progress"] == [(10, 26), (20, 26), (26, 26)]
    regs = [c for c in calls if c[0] == "registered"]
    ws = [c for c in calls if c[0] == "w"]
    assert [(c[1], c[2]) for c in regs] == [(0, 10), (10, 20), (20, 26)]
    assert [(c[1], c[2]) for c in ws] == [(0, 10), (10, 20), (20, 26)]
    for c in regs:
        assert c[3] == (c[2] - c[1], H, W, C) and c[4] == np.float32
        np.testing.assert_array_equal(c[5], reg[c[1]:c[2]])
    for c in ws:
        assert c[3] == (c[2] - c[1], H, W, 2) and c[4] == np.float32
        np.testing.assert_array_equal(c[5], w[c[1]:c[2]])


def test_raising_callback_warns_and_run_continues(gpu):
    from pyflowreg_amd.compensate import compensate_arr
    v = _video(13, 42)

    def bad(*a):
        raise RuntimeError("boom")
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        reg, w = compensate_arr(v[1:], v[0], _opt(buffer_size=6), progress_callback=bad,
                                registered_callback=bad, w_callback=bad)
    msgs = [str(r.message) for r in rec]
    assert sum("Progress callback error: boom" in m for m in msgs) == 2
    assert sum("Registered frames callback error: boom" in m for m in msgs) == 2
    assert sum("Displacement field callback error: boom" in m for m in msgs) == 2
    assert w.shape == (12, H, W, 2)


def test_sequential_executor_reports_every_frame(gpu):
    """parallelization='sequential' reports progress per frame, as the
    sequential executor's progress_callback(1) does; an executor the
    backend does not declare falls back to sequential with a warning."""
    from pyflowreg_amd.compensate import compensate_arr
    v = _video(8, 43)
    seen = []
    reg_a, w_a = compensate_arr(v[1:], v[0], _opt(buffer_size=4),
                                progress_callback=lambda c, t: seen.append((c, t)),
                                registration_config={"parallelization": "sequential"})
    assert seen == [(k, 7) for k in range(1, 8)]
    with pytest.warns(UserWarning, match="does not support 'multiprocessing'"):
        reg_b, w_b = compensate_arr(v[1:], v[0], _opt(buffer_size=4),
                                    registration_config={"parallelization": "multiprocessing"})
    np.testing.assert_array_equal(w_a, w_b)
    np.testing.assert_array_equal(reg_a, reg_b)


def test_custom_get_displacement_is_called_per_frame(gpu):
    """get_displacement= overrides the solver: it is called frame by frame
    with float32 (H, W, C) arrays, uv = a copy of w_init (float32) and the
    flow options; the w_init pre-pass over 7 (> 4) frames passes uv too.
    A wrapper around this backend's own get_displacement gives exactly the
    batched path's result."""
    from pyflowreg_amd import get_displacement as own
    from pyflowreg_amd.compensate import compensate_arr
    v = _video(8, 44)
    log = []

    def custom(fixed, moving, **kw):
        log.append((fixed.dtype, fixed.shape, moving.dtype, moving.shape, kw.get("uv") is not None,
                    None if kw.get("uv") is None else kw["uv"].dtype, sorted(kw)))
        return own(fixed, moving, **kw)
    reg_c, w_c = compensate_arr(v[1:], v[0], _opt(), get_displacement=custom)
    reg_b, w_b = compensate_arr(v[1:], v[0], _opt())
    assert len(log) == 14  # 7 for w_init, 7 for the batch
    for dt1, s1, dt2, s2, has_uv, uvdt, keys in log:
        assert dt1 == np.float32 and dt2 == np.float32 and s1 == s2 == (H, W, C)
        assert has_uv and uvdt == np.float32
        assert {"alpha", "weight", "levels", "min_level", "eta", "iterations", "update_lag", "a_data",
                "a_smooth", "const_assumption"} <= set(keys)
    np.testing.assert_array_equal(w_c, w_b)
    np.testing.assert_array_equal(reg_c, reg_b)


def test_few_frames_w_init_passes_no_uv(gpu):
    """At most 4 frames: the reference computes w_init serially without uv
    (compensate_recording.py:664-671)."""
    from pyflowreg_amd import get_displacement as own
    from pyflowreg_amd.compensate import compensate_arr
    v = _video(4, 45)
    uvs = []

    def custom(fixed, moving, **kw):
        uvs.append("uv" in kw)
        return own(fixed, moving, **kw)
    compensate_arr(v[1:], v[0], _opt(), get_displacement=custom)
    assert uvs == [False] * 3 + [True] * 3


def test_factory_and_backend_resolution(gpu):
    """get_displacement_factory(**backend_params) is called once; a backend
    name is looked up in the registry (ValueError for an unknown name);
    this backend's own factory keeps the batched path."""
    from pyflowreg_amd import get_displacement as own
    from pyflowreg_amd import register_backend
    from pyflowreg_amd.compensate import compensate_arr
    v = _video(6, 46)
    made = []

    def factory(**params):
        made.append(params)
        return lambda f, m, **kw: own(f, m, **kw)
    _, w_f = compensate_arr(v[1:], v[0], _opt(), get_displacement_factory=factory,
                            backend_params={"device": "cuda:0"})
    assert made == [{"device": "cuda:0"}]
    _, w_b = compensate_arr(v[1:], v[0], _opt(), flow_backend="flowreg_hip")
    np.testing.assert_array_equal(w_f, w_b)
    # the reference's own backend name selects this backend (a drop-in for its default)
    _, w_r = compensate_arr(v[1:], v[0], _opt(), flow_backend="flowreg")
    np.testing.assert_array_equal(w_r, w_b)
    with pytest.raises(ValueError, match="Unknown flow backend"):
        compensate_arr(v[1:], v[0], _opt(), flow_backend="no_such_backend")
    register_backend("test_zero_flow", lambda **p: (lambda f, m, **kw: np.zeros(f.shape[:2] + (2,))),
                     supported_executors={"sequential"})
    _, w_z = compensate_arr(v[1:], v[0], _opt(), flow_backend="test_zero_flow")
    assert np.all(w_z == 0)


def test_output_typename_and_shapes(gpu):
    from pyflowreg_amd.compensate import compensate_arr
    v = _video(5, 47, C=1)[..., 0]                         # (T, H, W)
    reg, w = compensate_arr(v[1:], v[0], _opt())          # 2-D reference: (T, H, W) back
    assert reg.shape == (4, H, W) and reg.dtype == np.float64 and w.shape == (4, H, W, 2)
    reg1, _ = compensate_arr(v[1:], v[0], _opt(output_typename="single"))
    assert reg1.dtype == np.float32
    np.testing.assert_array_equal(reg1.astype(np.float64), reg)
    reg2, _ = compensate_arr(v[1:], v[0], _opt(output_typename="bogus"))
    assert reg2.dtype == np.float32                        # unknown names keep the dtype
    one, w1 = compensate_arr(v[1], v[0], _opt())           # (H, W): (H, W) and (H, W, 2)
    assert one.shape == (H, W) and w1.shape == (H, W, 2)
    with pytest.raises(ValueError, match="empty"):
        compensate_arr(np.zeros((0, H, W), np.float32), v[0], _opt())


def test_ofoptions_drive_compensate_arr(gpu):
    """An OFOptions (the reference's field names and defaults) gives the same
    result as the equivalent RegistrationOptions; its gnc_schedule reaches
    the displacement calls (the GNC stages change the flows)."""
    from pyflowreg_amd import OFOptions
    from pyflowreg_amd.compensate import compensate_arr
    v = _video(9, 48)
    o = OFOptions(alpha=2, buffer_size=10, iterations=20, sigma=[1.0, 1.0, 0.1], weight=[1, 1])
    reg_o, w_o = compensate_arr(v[1:], v[0], o)
    reg_r, w_r = compensate_arr(v[1:], v[0], _opt())
    np.testing.assert_array_equal(w_o, w_r)
    np.testing.assert_array_equal(reg_o, reg_r)
    assert reg_o.dtype == np.float64
    g = OFOptions(alpha=2, buffer_size=10, iterations=6, levels=3, sigma=[1.0, 1.0, 0.1],
                  gnc_schedule=[0.0, 1.0], warping_steps=1)
    _, w_g = compensate_arr(v[1:4], v[0], g)
    n = OFOptions(alpha=2, buffer_size=10, iterations=6, levels=3, sigma=[1.0, 1.0, 0.1])
    _, w_n = compensate_arr(v[1:4], v[0], n)
    assert np.isfinite(w_g).all() and w_g.shape == w_n.shape
    assert np.abs(w_g - w_n).max() > 1e-6
