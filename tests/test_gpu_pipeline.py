"""GPU pipeline pieces against the oracle: the batched executor, the
compensate_arr batch loop (w_init from 22 frames, mean of the last 20),
min_level > 0 (final cv2.resize), registration warps and valid masks."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _video(T, H, W, C, seed):
    from pyflowreg_amd import synth
    frames, _ = synth.video(T, H, W, n_channels=C, max_disp=2.0, sigma_d=H / 5, seed=seed)
    return frames.numpy()


def _oracle_batch(oracle, batch, ref_raw, ref_proc, w_init, fp):
    """parallelization/sequential.py:97-169 with the oracle."""
    bp = oracle.preprocess(batch, ref_raw).astype(np.float32)
    flows, regs = [], []
    for t in range(batch.shape[0]):
        f = oracle.get_displacement(ref_proc, bp[t], uv=w_init.astype(np.float32).copy(),
                                    **fp).astype(np.float32)
        raw = batch[t] if batch.dtype in (np.uint8, np.uint16, np.int16) else batch[t].astype(np.float32)
        reg = oracle.imregister_wrapper(raw, f[..., 0].astype(np.float64),
                                        f[..., 1].astype(np.float64), ref_raw.astype(np.float32))
        flows.append(f)
        regs.append(reg)
    return np.stack(regs), np.stack(flows)


def test_executor_matches_sequential_oracle(gpu, oracle):
    from pyflowreg_amd.executor import HipBatchedExecutor
    v = _video(5, 48, 40, 2, 21)
    ref_raw = v[0].astype(np.float64)
    ref_proc = oracle.preprocess(ref_raw).astype(np.float32)
    batch = v[1:]
    w_init = np.zeros((48, 40, 2), np.float32)
    w_init[..., 0] = 0.25
    fp = dict(alpha=(2.0, 2.0), weight=np.ones((48, 40, 2)) * 0.5, levels=50, min_level=0,
              eta=0.8, iterations=50, update_lag=5, a_data=0.45, a_smooth=1.0)
    ex = HipBatchedExecutor(max_batch=3)
    bp = oracle.preprocess(batch, ref_raw).astype(np.float32)
    reg, w = ex.process_batch(batch, bp, ref_raw, ref_proc, w_init, flow_params=fp)
    oreg, ow = _oracle_batch(oracle, batch, ref_raw, ref_proc, w_init, fp)
    np.testing.assert_array_equal(w, ow)
    np.testing.assert_array_equal(reg, oreg)
    assert reg.dtype == batch.dtype and w.dtype == np.float32


def _int_video(T, H, W, C, seed, dt):
    """Synthetic frames scaled to 16-bit counts (signed frames straddle 0)."""
    v = _video(T, H, W, C, seed)
    off = 0.0 if dt == np.uint16 else -1200.0
    return np.rint(v * 3000.0 + off).astype(dt)


@pytest.mark.parametrize("dt", [np.uint16, np.int16])
@pytest.mark.parametrize("method", ["cubic", "linear"])
def test_warp_frames_int16_matches_oracle(gpu, oracle, dt, method):
    """16-bit raw frames warp in their depth (cv2.remap 16U / 16S): float
    sums of the integer samples, in-bounds results rounded half to even and
    clamped, the float64 fill rounded to float32 only."""
    import torch
    from pyflowreg_amd.executor import warp_frames
    rng = np.random.default_rng(25)
    T, H, W, C = 3, 31, 37, 2
    info = np.iinfo(dt)
    frames = rng.integers(info.min, info.max, (T, H, W, C), endpoint=True).astype(dt)
    frames[0, 10:20, 10:20] = info.max  # flat patches with steps: cubic overshoot is clamped
    frames[1, 5:15, 20:30] = info.min
    fill = rng.random((H, W, C)) * 1000.0
    flow = rng.uniform(-3, 3, (T, H, W, 2)).astype(np.float32)
    dev = torch.device("cuda")
    out = warp_frames(torch.from_numpy(frames.astype(np.float32)).to(dev), torch.from_numpy(flow).to(dev),
                      torch.from_numpy(fill).to(dev), method, raw_dtype=dt).cpu().numpy()
    for t in range(T):
        ref = oracle.imregister_wrapper(frames[t], flow[t, ..., 0].astype(np.float64),
                                        flow[t, ..., 1].astype(np.float64), fill, method)
        np.testing.assert_array_equal(out[t], ref, err_msg=f"frame {t}")
    f32 = oracle.imregister_wrapper(frames[0].astype(np.float32), flow[0, ..., 0].astype(np.float64),
                                    flow[0, ..., 1].astype(np.float64), fill, method)
    assert np.any(out[0] != f32)  # the rounding and clamping did something


@pytest.mark.parametrize("method", ["cubic", "linear"])
def test_warp_frames_u8_matches_oracle_and_stays_in_range(gpu, oracle, method):
    """uint8 raw frames warp in cv2's 8U depth: 15-bit fixed-point weights,
    exact integer sums, rounded and saturated results.  Step edges 0 -> 255
    make the cubic overshoot on both sides; every in-bounds sample must stay
    in [0, 255] and be an integer (a float path cast to uint8 would wrap)."""
    import torch
    from pyflowreg_amd.executor import warp_frames
    rng = np.random.default_rng(27)
    T, H, W, C = 3, 33, 41, 2
    frames = rng.integers(0, 256, (T, H, W, C), dtype=np.uint8)
    frames[0, :, : W // 2] = 0
    frames[0, :, W // 2:] = 255
    frames[1, 8:24, 8:24] = 255
    frames[1, 12:20, 12:20] = 0
    fill = rng.integers(0, 256, (H, W, C)).astype(np.float64) + 0.25
    flow = rng.uniform(-2.5, 2.5, (T, H, W, 2)).astype(np.float32)
    flow[0, ..., 0] = 0.4
    flow[0, ..., 1] = 0.0
    dev = torch.device("cuda")
    out = warp_frames(torch.from_numpy(frames.astype(np.float32)).to(dev), torch.from_numpy(flow).to(dev),
                      torch.from_numpy(fill).to(dev), method, raw_dtype=np.uint8).cpu().numpy()
    for t in range(T):
        ref = oracle.imregister_wrapper(frames[t], flow[t, ..., 0].astype(np.float64),
                                        flow[t, ..., 1].astype(np.float64), fill.astype(np.float32), method)
        np.testing.assert_array_equal(out[t], ref, err_msg=f"frame {t}")
    inb = out[0][:, :W - 1]
    assert inb.min() >= 0 and inb.max() <= 255 and np.all(inb == np.rint(inb))
    if method == "cubic":
        f32 = oracle.imregister_wrapper(frames[0].astype(np.float32), flow[0, ..., 0].astype(np.float64),
                                        flow[0, ..., 1].astype(np.float64), fill.astype(np.float32), method)
        assert f32.min() < 0 and f32.max() > 255   # what the float path would have stored


def test_imregister_wrapper_integer_depths(gpu, oracle):
    """The stand-alone imregister_wrapper keeps uint8 / uint16 / int16
    frames in their cv2.remap depth and rejects depths cv2.remap lacks."""
    rng = np.random.default_rng(28)
    H, W = 21, 26
    u = rng.uniform(-1.5, 1.5, (H, W))
    v = rng.uniform(-1.5, 1.5, (H, W))
    for dt in (np.uint8, np.uint16, np.int16):
        info = np.iinfo(dt)
        f2 = rng.integers(info.min, info.max, (H, W, 2), endpoint=True).astype(dt)
        f1 = rng.random((H, W, 2)).astype(np.float32) * 100
        for method in ("cubic", "linear"):
            a = gpu.imregister_wrapper(f2, u, v, f1, method)
            b = oracle.imregister_wrapper(f2, u, v, f1, method)
            np.testing.assert_array_equal(a, b, err_msg=f"{np.dtype(dt)} {method}")
    for dt in (np.int8, np.int32):
        with pytest.raises(ValueError):
            gpu.imregister_wrapper(np.zeros((H, W), dt), u, v, np.zeros((H, W), np.float32))


def test_executor_u8_frames_match_oracle(gpu, oracle):
    """process_batch with uint8 raw frames: registered frames are uint8,
    equal to cv2's 8U remap result (and the fill) as the reference stores
    them."""
    from pyflowreg_amd.executor import HipBatchedExecutor
    v = np.clip(np.rint(_video(5, 40, 36, 2, 29) * 200.0), 0, 255).astype(np.uint8)
    ref_raw = v[0].astype(np.float64)
    ref_proc = oracle.preprocess(ref_raw).astype(np.float32)
    batch = v[1:]
    w_init = np.zeros((40, 36, 2), np.float32)
    w_init[..., 1] = -1.25
    fp = dict(alpha=(2.0, 2.0), weight=np.ones((40, 36, 2)) * 0.5, levels=50, min_level=0,
              eta=0.8, iterations=50, update_lag=5, a_data=0.45, a_smooth=1.0)
    ex = HipBatchedExecutor(max_batch=4)
    bp = oracle.preprocess(batch, ref_raw).astype(np.float32)
    reg, w = ex.process_batch(batch, bp, ref_raw, ref_proc, w_init, flow_params=fp)
    oreg, ow = _oracle_batch(oracle, batch, ref_raw, ref_proc, w_init, fp)
    np.testing.assert_array_equal(w, ow)
    assert reg.dtype == np.uint8
    np.testing.assert_array_equal(reg, oreg.astype(np.uint8))
    with pytest.raises(ValueError):
        ex.process_batch(batch.astype(np.int8), bp, ref_raw, ref_proc, w_init, flow_params=fp)


@pytest.mark.parametrize("dt", [np.uint16, np.int16])
def test_executor_int16_frames_match_oracle(gpu, oracle, dt):
    """process_batch with 16-bit raw frames: registered frames in the
    batch's dtype, equal to the reference's float32 result of cv2's 16-bit
    remap (and of the float fill) cast to that dtype."""
    from pyflowreg_amd.executor import HipBatchedExecutor
    v = _int_video(5, 40, 36, 2, 26, dt)
    ref_raw = v[0].astype(np.float64) + 0.375  # a non-integer out-of-bounds fill
    ref_proc = oracle.preprocess(ref_raw).astype(np.float32)
    batch = v[1:]
    w_init = np.zeros((40, 36, 2), np.float32)
    w_init[..., 0] = 1.5
    fp = dict(alpha=(2.0, 2.0), weight=np.ones((40, 36, 2)) * 0.5, levels=50, min_level=0,
              eta=0.8, iterations=50, update_lag=5, a_data=0.45, a_smooth=1.0)
    ex = HipBatchedExecutor(max_batch=4)
    bp = oracle.preprocess(batch, ref_raw).astype(np.float32)
    reg, w = ex.process_batch(batch, bp, ref_raw, ref_proc, w_init, flow_params=fp)
    oreg, ow = _oracle_batch(oracle, batch, ref_raw, ref_proc, w_init, fp)
    np.testing.assert_array_equal(w, ow)
    assert reg.dtype == dt
    np.testing.assert_array_equal(reg, oreg.astype(dt))


@pytest.mark.parametrize("dt", [np.float32, np.uint16])
def test_compensate_arr_matches_oracle_batch_loop(gpu, oracle, dt):
    """compensate_recording.py:645-678,786-806 with buffer_size 10 over 26
    frames: w_init from the first min(22, 10) frames, then the mean of each
    batch's last (up to) 20 flows.  uint16 frames: staged as float32,
    registered in cv2's 16U depth, stored as uint16 and returned as float64
    (output_typename "double", compensate_arr.py:194-207)."""
    from pyflowreg_amd.compensate import RegistrationOptions, compensate_arr
    v = _video(27, 40, 32, 2, 31) if dt == np.float32 else _int_video(27, 40, 32, 2, 31, dt)
    ref = v[0].astype(np.float64)
    video = v[1:]
    opt = RegistrationOptions(alpha=(2.0, 2.0), buffer_size=10, sigma=(1.0, 1.0, 0.1))
    reg, w = compensate_arr(video, ref, opt)
    # oracle restatement of the batch loop
    ref_proc = oracle.preprocess(ref).astype(np.float32)
    fp = dict(alpha=(2.0, 2.0), weight=np.ones((40, 32, 2)) * 0.5, levels=100, min_level=0,
              eta=0.8, iterations=50, update_lag=5, a_data=0.45, a_smooth=1.0)
    w_init = None
    oregs, ows = [], []
    for s in range(0, video.shape[0], 10):
        b = video[s:s + 10]
        if w_init is None:
            _, w0 = _oracle_batch(oracle, b[:10], ref, ref_proc, np.zeros((40, 32, 2)), fp)
            w_init = np.mean(w0, axis=0)
        r, ww = _oracle_batch(oracle, b, ref, ref_proc, w_init, fp)
        w_init = np.mean(ww[-20:], axis=0) if ww.shape[0] > 20 else np.mean(ww, axis=0)
        oregs.append(r)
        ows.append(ww)
    np.testing.assert_array_equal(w, np.concatenate(ows))
    # the pipeline stores the frames in the batch's dtype; compensate_arr
    # then casts them per output_typename, "double" by default
    assert reg.dtype == np.float64
    np.testing.assert_array_equal(reg, np.concatenate(oregs).astype(dt).astype(np.float64))


def test_compensate_update_reference_matches_oracle(gpu, oracle):
    """compensate_recording.py:680-713,866-867: after each batch the
    processed reference becomes the float64 mean of the last min(100, T)
    preprocessed float64 frames warped by their flows (float64 images:
    cv2.remap's 64F path) with the old reference as fill; the next batch
    registers against it."""
    from pyflowreg_amd.compensate import MotionCorrector, RegistrationOptions
    v = _video(25, 40, 32, 2, 33)
    ref = v[0].astype(np.float64)
    video = v[1:]
    opt = RegistrationOptions(alpha=(2.0, 2.0), buffer_size=8, sigma=(1.0, 1.0, 0.1), update_reference=True)
    mc = MotionCorrector(ref, opt, max_batch=8)
    fp = dict(alpha=(2.0, 2.0), weight=np.ones((40, 32, 2)) * 0.5, levels=100, min_level=0,
              eta=0.8, iterations=50, update_lag=5, a_data=0.45, a_smooth=1.0)
    ref_proc = oracle.preprocess(ref)  # float64
    w_init = None
    try:
        for s in range(0, video.shape[0], 8):
            b = video[s:s + 8]
            reg, w = mc.process(b)
            rp32 = ref_proc.astype(np.float32)
            if w_init is None:
                _, w0 = _oracle_batch(oracle, b, ref, rp32, np.zeros((40, 32, 2)), fp)
                w_init = np.mean(w0, axis=0)
            oreg, ow = _oracle_batch(oracle, b, ref, rp32, w_init, fp)
            w_init = np.mean(ow[-20:], axis=0) if ow.shape[0] > 20 else np.mean(ow, axis=0)
            np.testing.assert_array_equal(w, ow)
            np.testing.assert_array_equal(reg, oreg)
            bp64 = oracle.preprocess(b, ref)
            n = min(100, b.shape[0])
            comp = np.stack([oracle.imregister_wrapper(bp64[t], ow[t, :, :, 0], ow[t, :, :, 1], ref_proc)
                             for t in range(b.shape[0] - n, b.shape[0])]).astype(np.float64)
            ref_proc = np.mean(comp, axis=0)
            np.testing.assert_array_equal(mc.reference_proc.cpu().numpy(), ref_proc)
    finally:
        mc.close()


def test_preprocess_float64_output_matches_oracle(gpu, oracle):
    from pyflowreg_amd.preprocess import preprocess
    v = _video(4, 36, 30, 2, 34)
    ref = v[0].astype(np.float64)
    out = preprocess(v[1:], normalization_ref=ref, out_dtype=__import__("torch").float64)
    exp = oracle.preprocess(v[1:], ref)
    np.testing.assert_array_equal(out.permute(0, 2, 3, 1).cpu().numpy(), exp)


def test_min_level_above_zero_matches_oracle(gpu, oracle):
    """quality 'balanced' (min_level 4): the flow stops at level 4 and is
    brought to full resolution with cv2.resize INTER_CUBIC."""
    import pyflowreg_amd as fr
    v = _video(2, 80, 64, 1, 41)
    ref = oracle.preprocess(v[0].astype(np.float64)).astype(np.float32)
    mov = oracle.preprocess(v[1], v[0].astype(np.float64)).astype(np.float32)
    a = fr.get_displacement(ref, mov, min_level=4)
    b = oracle.get_displacement(ref, mov, min_level=4)
    assert a.shape == (80, 64, 2)
    np.testing.assert_array_equal(a, b)


def test_hooked_level_solver_path_matches(gpu, oracle):
    """get_displacement(level_solver_backend=...) runs the reference's level
    loop with the caller's solver; with the oracle's solver it equals the
    oracle end to end."""
    import pyflowreg_amd as fr
    v = _video(2, 40, 48, 2, 51)
    ref = oracle.preprocess(v[0].astype(np.float64)).astype(np.float32)
    mov = oracle.preprocess(v[1], v[0].astype(np.float64)).astype(np.float32)
    calls = []

    def solver(*args):
        calls.append(args[0].shape)
        return oracle.level_solver(*args)
    a = fr.get_displacement(ref, mov, weight=[0.5, 0.5], level_solver_backend=solver)
    b = oracle.get_displacement(ref, mov, weight=[0.5, 0.5])
    assert len(calls) == len(oracle.level_plan(40, 48, 0.8, 50, 0)[0])
    np.testing.assert_array_equal(a, b)


def test_valid_masks():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test without a GPU")
    from pyflowreg_amd.executor import compute_batch_valid_masks
    w = np.zeros((2, 10, 12, 2), np.float32)
    w[0, ..., 0] = 2.5
    w[1, ..., 1] = -3.0
    m = compute_batch_valid_masks(w)
    assert m.dtype == np.uint8
    assert np.all(m[0][:, :9] == 255) and np.all(m[0][:, 10:] == 0)
    assert np.all(m[1][:3] == 0) and np.all(m[1][3:] == 255)


@pytest.mark.parametrize("sigma", [(1.0, 1.0, 0.6), (1.2, 0.8, 1.5), ((1.0, 1.0, 0.1), (2.0, 0.5, 0.7))])
def test_preprocess_temporal_and_per_channel_sigmas(gpu, oracle, sigma):
    """apply_gaussian_filter's 4-D branch (util/image_processing.py:147-161):
    the temporal kernel over the frames of the call (radius 2, and 6 > T = 5
    with repeated reflections), per-channel sigma rows; a single frame is
    filtered spatially only."""
    import torch
    from pyflowreg_amd.preprocess import preprocess
    v = _video(6, 36, 30, 2, 35)
    ref = v[0].astype(np.float64)
    sig = np.asarray(sigma, dtype=np.float64)
    exp = oracle.preprocess(v[1:], ref, sigma=sig)
    out64 = preprocess(v[1:], normalization_ref=ref, sigma=sig, out_dtype=torch.float64)
    np.testing.assert_array_equal(out64.permute(0, 2, 3, 1).cpu().numpy(), exp)
    out32 = preprocess(v[1:], normalization_ref=ref, sigma=sig)
    np.testing.assert_array_equal(out32.permute(0, 2, 3, 1).cpu().numpy(), exp.astype(np.float32))
    one = preprocess(ref, sigma=sig, out_dtype=torch.float64)
    np.testing.assert_array_equal(one.permute(0, 2, 3, 1)[0].cpu().numpy(), oracle.preprocess(ref, sigma=sig))


def test_compensate_temporal_prefilter_matches_oracle(gpu, oracle):
    """compensate_arr with sigma_t = 0.6 (temporal radius 2): each batch of
    5 frames is filtered along its own frames, as _preprocess_frames does."""
    from pyflowreg_amd.compensate import RegistrationOptions, compensate_arr
    v = _video(11, 32, 28, 2, 36)
    ref = v[0].astype(np.float64)
    video = v[1:]
    sig = (1.0, 1.0, 0.6)
    opt = RegistrationOptions(alpha=(2.0, 2.0), buffer_size=5, sigma=sig)
    reg, w = compensate_arr(video, ref, opt)
    ref_proc = oracle.preprocess(ref, sigma=sig).astype(np.float32)
    fp = dict(alpha=(2.0, 2.0), weight=np.ones((32, 28, 2)) * 0.5, levels=100, min_level=0,
              eta=0.8, iterations=50, update_lag=5, a_data=0.45, a_smooth=1.0)

    def batch(b, w_init):
        bp = oracle.preprocess(b, ref, sigma=sig).astype(np.float32)
        flows, regs = [], []
        for t in range(b.shape[0]):
            f = oracle.get_displacement(ref_proc, bp[t], uv=w_init.astype(np.float32).copy(), **fp).astype(np.float32)
            regs.append(oracle.imregister_wrapper(b[t].astype(np.float32), f[..., 0].astype(np.float64),
                                                  f[..., 1].astype(np.float64), ref.astype(np.float32)))
            flows.append(f)
        return np.stack(regs), np.stack(flows)
    w_init = None
    oregs, ows = [], []
    for s in range(0, video.shape[0], 5):
        b = video[s:s + 5]
        if w_init is None:
            _, w0 = batch(b, np.zeros((32, 28, 2)))
            w_init = np.mean(w0, axis=0)
        r, ww = batch(b, w_init)
        w_init = np.mean(ww[-20:], axis=0)
        oregs.append(r)
        ows.append(ww)
    np.testing.assert_array_equal(w, np.concatenate(ows))
    np.testing.assert_array_equal(reg, np.concatenate(oregs))


def _moving_disc_uint16(T, H, W, C, seed):
    """The reference's test recording (tests/fixtures.py:14-87, pattern
    "motion"): a bright disc moving on a circle over a textured, noisy
    background, clipped to [0, 1] and stored as uint16 counts."""
    rng = np.random.default_rng(seed)
    frames = np.zeros((T, H, W, C), np.float32)
    cy, cx, radius = H // 2, W // 2, min(H, W) // 8
    y, x = np.ogrid[:H, :W]
    for t in range(T):
        for c in range(C):
            ox = int(radius * 0.5 * np.cos(2 * np.pi * t / T))
            oy = int(radius * 0.5 * np.sin(2 * np.pi * t / T))
            mask = (x - (cx + ox)) ** 2 + (y - (cy + oy)) ** 2 <= radius ** 2
            frames[t, :, :, c] = mask.astype(np.float32) * 0.8
            frames[t, :, :, c] += 0.2 * rng.random((H, W))
            frames[t, :, :, c] += 0.1 * rng.random((H, W))
    return (np.clip(frames, 0.0, 1.0) * 65535).astype(np.uint16)


@pytest.mark.parametrize("preset", ["minimal", "fast"])
def test_compensate_arr_reference_test_options(gpu, oracle, preset):
    """The options of the reference's own test suite (tests/fixtures.py:
    133-165 get_minimal_of_options: alpha 100, levels 2, iterations 5, eta
    0.95, a_data 1, per-channel sigma rows with sigma_t 0.5, weights (1, 1),
    joint normalisation; tests/conftest.py:98-106 fast_of_options: levels 1,
    iterations 2, alpha 50) on its uint16 moving-disc recording, against the
    oracle's batch loop: flows and registered uint16 frames bit for bit."""
    from pyflowreg_amd.compensate import RegistrationOptions, compensate_arr
    T, H, W, C, bs = 12, 32, 64, 2, 6
    video = _moving_disc_uint16(T, H, W, C, 37)
    ref = np.mean(video[:5], axis=0)
    alpha, levels, iters = ((100.0, 100.0), 2, 5) if preset == "minimal" else ((50.0, 50.0), 1, 2)
    sig = ((1.0, 1.0, 0.5), (1.0, 1.0, 0.5))
    opt = RegistrationOptions(alpha=alpha, levels=levels, iterations=iters, eta=0.95, update_lag=5,
                              a_smooth=1.0, a_data=1.0, sigma=sig, weight=(1.0, 1.0), min_level=0,
                              quality_setting="fast", channel_normalization="joint", buffer_size=bs)
    reg, w = compensate_arr(video, ref, opt)
    ref_proc = oracle.preprocess(ref, sigma=sig).astype(np.float32)
    fp = dict(alpha=alpha, weight=np.ones((H, W, C)) * 0.5, levels=levels, min_level=0, eta=0.95,
              iterations=iters, update_lag=5, a_data=1.0, a_smooth=1.0)

    def batch(b, w_init):
        bp = oracle.preprocess(b, ref, sigma=sig).astype(np.float32)
        flows, regs = [], []
        for t in range(b.shape[0]):
            f = oracle.get_displacement(ref_proc, bp[t], uv=w_init.astype(np.float32).copy(),
                                        **fp).astype(np.float32)
            regs.append(oracle.imregister_wrapper(b[t], f[..., 0].astype(np.float64),
                                                  f[..., 1].astype(np.float64), ref))
            flows.append(f)
        return np.stack(regs), np.stack(flows)
    w_init = None
    oregs, ows = [], []
    for s in range(0, T, bs):
        b = video[s:s + bs]
        if w_init is None:
            _, w0 = batch(b, np.zeros((H, W, 2)))
            w_init = np.mean(w0, axis=0)
        r, ww = batch(b, w_init)
        w_init = np.mean(ww[-20:], axis=0)
        oregs.append(r)
        ows.append(ww)
    np.testing.assert_array_equal(w, np.concatenate(ows))
    assert reg.dtype == np.float64  # output_typename "double"
    np.testing.assert_array_equal(reg, np.concatenate(oregs).astype(np.uint16).astype(np.float64))
    assert np.abs(w).max() > 0.1  # the disc's motion is found


@pytest.mark.parametrize("n", [1, 2, 7, 20, 22])
def test_mean_frames_equals_numpy(gpu, n):
    """frh_mean_frames_f32 against np.mean(w, axis=0) on float32 flows of
    mixed magnitudes, bit for bit."""
    import torch
    from pyflowreg_amd.executor import mean_frames
    rng = np.random.default_rng(40 + n)
    w = (rng.standard_normal((n, 37, 41, 2)) * rng.choice([1e-3, 1.0, 1e3], (n, 1, 1, 1))).astype(np.float32)
    out = mean_frames(torch.from_numpy(w).cuda()).cpu().numpy()
    np.testing.assert_array_equal(out, np.mean(w, axis=0))


def _median_tie_inputs():
    """2-D float64 arrays that a median run on float32 keys gets wrong unless
    it resolves ties: distinct values one float64 ulp apart (one float32
    key), exact duplicates, signed zeros, a wide exponent range, subnormals."""
    rng = np.random.default_rng(11)
    h, w = 37, 41
    base = rng.choice([0.1, 1.0, -2.5, 3.0e5, 7.0e-6], size=(h, w))
    ulps = base + rng.integers(-3, 4, size=(h, w)) * np.spacing(np.abs(base))
    mixed = rng.standard_normal((h, w))
    mixed[rng.random((h, w)) < 0.3] = 0.25
    mixed[rng.random((h, w)) < 0.1] = 0.0
    mixed[rng.random((h, w)) < 0.1] = -0.0
    wide = np.sign(rng.standard_normal((h, w))) * 10.0 ** rng.uniform(-300, 300, (h, w))
    sub = rng.integers(-6, 7, size=(h, w)) * 5e-324 + rng.integers(0, 2, size=(h, w)) * 1e-310
    same = np.full((h, w), 3.0)
    # one float32 key everywhere, distinct doubles: every window takes the exact path
    one_key = 1.0 + rng.integers(0, 9, size=(h, w)) * 2.0 ** -52
    return {"ulps": ulps, "mixed": mixed, "wide": wide, "sub": sub, "same": same, "one_key": one_key}


@pytest.mark.parametrize("case", ["ulps", "mixed", "wide", "sub", "same", "one_key"])
def test_median_float32_key_ties_match_scipy(gpu, case):
    """The 5x5 mirror median (float32-keyed selection, float64 network for
    windows whose median key several distinct samples share) equals
    scipy.ndimage.median_filter on inputs built to break a keyed selection."""
    from scipy.ndimage import median_filter
    from pyflowreg_amd.optical_flow import median5_mirror
    f = _median_tie_inputs()[case]
    np.testing.assert_array_equal(median5_mirror(f), median_filter(f, size=(5, 5), mode="mirror"))
