"""Each benchmark configuration's own workload through the HIP path, against
the oracle on the same seeded inputs (bit-exact), one test per
configuration so that a failure names it:

* 256x256x1 single pair (the CPU-runnable case): preprocessing with the
  reference's range, then get_displacement;
* one 64-frame launch of 512x512x2 frames, frames 0, 31 and 63 against the
  oracle (frame 63's solver workspace lies past 4 GiB), and a 256-frame
  launch (the headline batch, ~25 GB of workspace) whose first and last
  frames equal one-frame solves and the oracle;
* uneven pairs: 256x256x1 and 1024x1024x2 pairs through get_displacements
  against per-pair oracle runs (1024 rows: the levels over 512 rows run the
  1024-lane sweep work-groups).
"""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _video(T, H, W, C, seed):
    from pyflowreg_amd import synth
    v, _ = synth.video(T, H, W, n_channels=C, max_disp=3.0, sigma_d=48.0, seed=seed)
    return v.numpy()


def test_single_pair_256x256x1_bit_exact(gpu, oracle):
    """256x256, 1 channel: 400 blobs, displacement sigma 32 px scaled to
    4 px, Poisson peak 200; normalised with the fixed frame's range and
    prefiltered (sigma (1, 1, 0.1)); default solver options (16 levels)."""
    import torch
    from pyflowreg_amd import synth
    from pyflowreg_amd.preprocess import preprocess
    fixed, moving, _ = synth.single_pair(256, 256, seed=0)
    ref_raw = fixed.astype(np.float64)
    ref = preprocess(ref_raw[..., None]).permute(0, 2, 3, 1)[0].cpu().numpy()
    mov = preprocess(moving[None, ..., None], normalization_ref=ref_raw[..., None]).permute(0, 2, 3, 1)[0]
    mov = mov.cpu().numpy()
    np.testing.assert_array_equal(ref, oracle.preprocess(ref_raw[..., None]).astype(np.float32))
    np.testing.assert_array_equal(mov, oracle.preprocess(moving[None, ..., None], ref_raw[..., None])
                                  .astype(np.float32)[0])
    a = gpu.get_displacement(ref, mov)
    b = oracle.get_displacement(ref, mov)
    assert a.shape == (256, 256, 2)
    d = np.abs(a - b)
    assert d.max() == 0.0, f"256x256x1: max |dflow| {d.max():.3e}"
    assert torch.cuda.is_available()


def test_batch_64_frames_512x512x2_bit_exact(gpu, oracle):
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver
    _say("64 x 512x512x2: generating")
    v = _video(65, 512, 512, 2, 1000)
    ref_raw = v[0].astype(np.float64)
    ref = oracle.preprocess(ref_raw).astype(np.float32)
    movs = oracle.preprocess(v[1:], ref_raw).astype(np.float32)
    s = DisplacementSolver(512, 512, 2, max_batch=64)
    s.set_reference(ref, np.array([0.5, 0.5]), None)
    _say("64 x 512x512x2: GPU launch")
    out = s.run(torch.from_numpy(movs).cuda(), out_dtype=torch.float64).cpu().numpy()
    s.close()
    for t in (0, 31, 63):
        _say(f"64 x 512x512x2: oracle frame {t}")
        exp = oracle.get_displacement(ref, movs[t], weight=[0.5, 0.5])
        d = np.abs(out[t] - exp)
        assert d.max() == 0.0, f"frame {t} of 64: max |dflow| {d.max():.3e}"


def test_batch_256_frames_512x512x2_first_and_last(gpu, oracle):
    """The headline launch: 256 frames in one batch.  A frame's flow does
    not depend on its batch, so frames 0 and 255 equal one-frame solves of
    the same frames, and frame 255 equals the oracle."""
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver
    _say("256 x 512x512x2: generating")
    v = _video(257, 512, 512, 2, 2000)
    ref_raw = v[0].astype(np.float64)
    ref = oracle.preprocess(ref_raw).astype(np.float32)
    movs = oracle.preprocess(v[1:], ref_raw).astype(np.float32)
    del v
    s = DisplacementSolver(512, 512, 2, max_batch=256)
    s.set_reference(ref, np.array([0.5, 0.5]), None)
    _say("256 x 512x512x2: GPU launch")
    out = s.run(torch.from_numpy(movs).cuda(), out_dtype=torch.float64).cpu().numpy()
    for t in (0, 255):
        one = s.run(torch.from_numpy(movs[t:t + 1]).cuda(), out_dtype=torch.float64).cpu().numpy()[0]
        np.testing.assert_array_equal(out[t], one, err_msg=f"frame {t} of 256 vs alone")
    s.close()
    _say("256 x 512x512x2: oracle frame 255")
    exp = oracle.get_displacement(ref, movs[255], weight=[0.5, 0.5])
    d = np.abs(out[255] - exp)
    assert d.max() == 0.0, f"frame 255 of 256: max |dflow| {d.max():.3e}"


def _pair(H, W, C, seed):
    v = _video(2, H, W, C, seed)
    return v[0], v[1]


def test_uneven_pairs_256x1_and_1024x2_bit_exact(gpu, oracle):
    """get_displacements over [256x256x1, 1024x1024x2, 256x256x1]: grouped
    by size, solved batched, returned in input order, each pair equal to
    its own oracle run."""
    _say("uneven: generating")
    pairs = [_pair(256, 256, 1, 3101), _pair(1024, 1024, 2, 3102), _pair(256, 256, 1, 3103)]
    fixed = [p[0] for p in pairs]
    moving = [p[1] for p in pairs]
    _say("uneven: GPU")
    got = gpu.get_displacements(fixed, moving)
    assert [g.shape for g in got] == [(256, 256, 2), (1024, 1024, 2), (256, 256, 2)]
    for k, (f, m) in enumerate(pairs):
        _say(f"uneven: oracle pair {k} {f.shape}")
        exp = oracle.get_displacement(f, m)
        d = np.abs(got[k] - exp)
        assert d.max() == 0.0, f"pair {k} {f.shape}: max |dflow| {d.max():.3e}"
