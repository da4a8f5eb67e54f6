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
