"""FlowRegLive on the GPU (pyflowreg_amd.live) against a frame-by-frame
restatement of motion_correction/flow_reg_live.py:111-323 with the oracle
(normalize / Gaussian / get_displacement / warp) and numpy: registered
frames and flows bit for bit, with the causal temporal filter active and the
reference mixed every few frames."""

from collections import deque

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _video(T, H, W, C, seed):
    from pyflowreg_amd import synth
    frames, _ = synth.video(T, H, W, n_channels=C, max_disp=2.0, sigma_d=H / 5, seed=seed)
    return frames.numpy()


def _oracle_live(oracle, ref_frame, frames, sigma, interval, weight=0.2, truncate=4.0,
                 channel_normalization="together"):
    sig = np.asarray(sigma, dtype=np.float64)
    sigma_2d, sigma_t = sig[:2], float(sig[2])
    ref_raw = ref_frame.copy()
    ref_proc = oracle.apply_gaussian_filter(oracle.normalize(ref_raw, channel_normalization=channel_normalization),
                                            sigma_2d, mode="reflect", truncate=truncate)
    lo, hi = ref_proc.min(), ref_proc.max()
    buf = deque(maxlen=max(1, int(truncate * sigma_t + 0.5) + 1))
    fp = dict(alpha=(2.0, 2.0), levels=100, min_level=6, eta=0.8, iterations=50, update_lag=5,
              a_data=0.45, a_smooth=1.0, weight=np.array([0.5, 0.5]))
    last, out = None, []
    for k, frame in enumerate(frames):
        norm = (frame - lo) / (hi - lo + 1e-8)
        f2d = oracle.apply_gaussian_filter(norm, sigma_2d, mode="reflect", truncate=truncate)
        buf.append(f2d)
        if sigma_t > 0 and len(buf) > 1:
            n = min(int(truncate * sigma_t + 0.5) + 1, len(buf))
            x = np.arange(n, dtype=np.float32)
            kern = np.exp(-0.5 * (x / sigma_t) ** 2)
            kern = kern / kern.sum()
            proc = np.zeros_like(buf[-1], dtype=np.float64)
            for i in range(n):
                proc += kern[i] * buf[-(i + 1)]
            proc = proc.astype(buf[-1].dtype)
        else:
            proc = f2d
        flow = oracle.get_displacement(ref_proc, proc, uv=last, **fp)
        last = flow.copy()
        reg = oracle.imregister_wrapper(frame, flow[..., 0], flow[..., 1], ref_raw)
        if (k + 1) % interval == 0:
            reg_proc = oracle.imregister_wrapper(proc, flow[..., 0], flow[..., 1], ref_proc)
            ref_proc = (1 - weight) * ref_proc + weight * reg_proc
            ref_raw = (1 - weight) * ref_raw + weight * reg
            lo, hi = ref_proc.min(), ref_proc.max()
        out.append((reg, flow))
    return out


@pytest.mark.timeout(600)
def test_flow_reg_live_matches_oracle(gpu, oracle):
    from pyflowreg_amd.compensate import RegistrationOptions
    from pyflowreg_amd.live import FlowRegLive
    v = _video(11, 48, 40, 2, 61)
    sigma = (1.0, 1.0, 0.6)  # temporal radius 2: a 3-frame causal buffer
    live = FlowRegLive(RegistrationOptions(alpha=(2.0, 2.0), sigma=sigma), reference_update_interval=4)
    live.set_reference(v[0])
    exp = _oracle_live(oracle, v[0], v[1:], sigma, 4)
    for k in range(1, v.shape[0]):
        reg, flow = live(v[k])
        oreg, oflow = exp[k - 1]
        np.testing.assert_array_equal(flow, oflow, err_msg=f"flow of frame {k}")
        np.testing.assert_array_equal(reg, oreg, err_msg=f"registered frame {k}")
    assert live.frame_count == 10


@pytest.mark.timeout(600)
def test_flow_reg_live_uint16_frames_match_oracle(gpu, oracle):
    """uint16 raw frames (flow_reg_live.py:216-323 with an integer
    recording): normalised in float64, registered with cv2's 16U remap
    (float sums, results rounded half to even and clamped), the raw
    reference becoming float64 at its first update."""
    from pyflowreg_amd.compensate import RegistrationOptions
    from pyflowreg_amd.live import FlowRegLive
    v = np.rint(_video(9, 40, 36, 2, 64) * 3000.0).astype(np.uint16)
    sigma = (1.0, 1.0, 0.0)
    live = FlowRegLive(RegistrationOptions(alpha=(2.0, 2.0), sigma=sigma), reference_update_interval=3)
    live.set_reference(v[0])
    exp = _oracle_live(oracle, v[0], v[1:], sigma, 3)
    for k in range(1, v.shape[0]):
        reg, flow = live(v[k])
        oreg, oflow = exp[k - 1]
        np.testing.assert_array_equal(flow, oflow, err_msg=f"flow of frame {k}")
        assert reg.dtype == np.float32
        np.testing.assert_array_equal(reg, oreg, err_msg=f"registered frame {k}")


def test_flow_reg_live_buffers_frames_before_a_reference(gpu):
    from pyflowreg_amd.live import FlowRegLive
    v = _video(3, 24, 20, 1, 62)
    live = FlowRegLive()
    frame = v[0]
    r, f = live(frame)
    assert r is frame and not f.any() and f.shape == (24, 20, 2) and len(live.reference_buffer) == 1
    with pytest.raises(ValueError):
        FlowRegLive().set_reference()


@pytest.mark.parametrize("mode", ["together", "separate"])
def test_float32_reference_preprocessing_matches_numpy(gpu, oracle, mode):
    """A float32 reference frame: normalize keeps float32 arithmetic
    (util/image_processing.py:12-98), and gaussian_filter of the float32
    array ("together") rounds each pass to float32, while the "separate"
    result is stored in a float64 array before filtering."""
    import torch
    from pyflowreg_amd.compensate import RegistrationOptions
    from pyflowreg_amd.live import FlowRegLive
    from pyflowreg_amd.preprocess import preprocess
    ref = _video(1, 36, 44, 2, 63)[0] * np.float32(3.7)
    live = FlowRegLive(RegistrationOptions(channel_normalization=mode))
    lo, den, sem = live._reference_bounds(ref)
    out = preprocess(ref, sigma=(1.3, 0.9, 0.0), out_dtype=torch.float64, norm_bounds=(lo, den),
                     semantics=sem)[0].permute(1, 2, 0).cpu().numpy()
    exp = oracle.apply_gaussian_filter(oracle.normalize(ref, channel_normalization=mode), np.array([1.3, 0.9]),
                                       mode="reflect", truncate=4.0)
    np.testing.assert_array_equal(out, exp)
