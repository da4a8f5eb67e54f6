"""The motion corrector's per-frame statistics (compensate_recording.py:
808-821), computed on frame chunks in threads, equal the reference's serial
numpy loop exactly (runs on the CPU)."""

from types import SimpleNamespace

import numpy as np


def test_threaded_statistics_equal_serial_loop():
    from pyflowreg_amd.compensate import MotionCorrector
    rng = np.random.default_rng(5)
    w = (rng.standard_normal((23, 31, 27, 2)) * 2).astype(np.float32)
    mc = SimpleNamespace(mean_disp=[], max_disp=[], mean_div=[], mean_translation=[])
    MotionCorrector._statistics(mc, w)
    disp = np.sqrt(w[:, :, :, 0] ** 2 + w[:, :, :, 1] ** 2)
    assert mc.mean_disp == np.mean(disp, axis=(1, 2)).tolist()
    assert mc.max_disp == np.max(disp, axis=(1, 2)).tolist()
    div, tr = [], []
    for t in range(w.shape[0]):
        div.append(float(np.mean(np.gradient(w[t, :, :, 0], axis=1) + np.gradient(w[t, :, :, 1], axis=0))))
        um, vm = float(np.mean(w[t, :, :, 0])), float(np.mean(w[t, :, :, 1]))
        tr.append(float(np.sqrt(um ** 2 + vm ** 2)))
    assert mc.mean_div == div and mc.mean_translation == tr


def test_numpy_frame_mean_is_an_ordered_float32_sum():
    """The w_init means (compensate_recording.py:674,801-806) are
    np.mean(w, axis=0) of float32 flows: numpy adds the frames one after
    another in float32 (no pairwise summation over the outer axis) and
    rounds the quotient by the frame count to float32.  frh_mean_frames_f32
    computes exactly this; a pairwise or float64 sum gives other bits on
    these data."""
    rng = np.random.default_rng(6)
    for n in (1, 2, 3, 7, 20, 22):
        w = (rng.standard_normal((n, 19, 23, 2)) * rng.choice([1e-3, 1.0, 1e3], (n, 1, 1, 1))).astype(np.float32)
        s = w[0].copy()
        for i in range(1, n):
            s = s + w[i]
        np.testing.assert_array_equal(np.mean(w, axis=0), s / np.float32(n))
        assert (s / np.float32(n)).dtype == np.float32
    w = (rng.standard_normal((20, 64, 64, 2)) * rng.choice([1e-3, 1.0, 1e3], (20, 1, 1, 1))).astype(np.float32)
    f64 = (w.astype(np.float64).sum(axis=0) / 20).astype(np.float32)
    assert np.any(np.mean(w, axis=0) != f64)
