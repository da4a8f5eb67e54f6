"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the CPU oracle): the oracle must still reproduce them on the CPU, and
the HIP path must reproduce them bit for bit on the GPU.

Fixtures whose levels are downsampled store the float32 Gaussian taps the
oracle used; numpy's float32 exp may round differently on another host,
and then the end-to-end case is skipped there (the taps are an input of
both sides, so the comparison would not be like for like).
"""

import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    return json.loads(str(z["meta"])), z


def _check_taps(meta, z, gauss_fn):
    for k, s in enumerate(meta["sigmas"]):
        if not np.array_equal(gauss_fn(s)[1], z[f"gauss_{k}"]):
            pytest.skip(f"this host's float32 Gaussian taps for sigma={s} differ from the fixture's")


# ------------------------------------------------------------------ oracle (CPU)

def test_oracle_resize(oracle):
    meta, z = load("resize_down")
    np.testing.assert_array_equal(oracle.gauss_g(meta["sigma"])[1], z["gauss"])
    np.testing.assert_array_equal(oracle.resize(z["img"], tuple(meta["size"])), z["out"])
    meta, z = load("resize_up")
    np.testing.assert_array_equal(oracle.resize(z["img"], tuple(meta["size"])), z["out"])


def test_oracle_warp(oracle):
    _, z = load("warp")
    out = oracle.imregister_wrapper(z["f2"], z["u"].astype(np.float64), z["v"].astype(np.float64), z["f1"])
    np.testing.assert_array_equal(out, z["out"])


def test_oracle_tensor(oracle):
    meta, z = load("tensor_gc")
    J = oracle.get_motion_tensor_gc(z["f1"], z["f2"], meta["hy"], meta["hx"])
    np.testing.assert_array_equal(np.stack(J), z["J"])


@pytest.mark.parametrize("kind", ["gc", "gray", "cs"])
def test_oracle_tensor_mixed(oracle, kind):
    meta, z = load(f"tensor_{kind}_mixed")
    fn = {"gc": oracle.get_motion_tensor_gc, "gray": oracle.get_motion_tensor_gray,
          "cs": oracle.get_motion_tensor_cs}[kind]
    np.testing.assert_array_equal(np.stack(fn(z["f1"], z["f2"], meta["hy"], meta["hx"])), z["J"])


LEVEL_CASES = ["level_solver_a", "level_solver_b", "level_solver_a_gnc", "level_solver_b_gnc"]
DISPLACEMENT_CASES = ["displacement_32x32x1", "displacement_40x48x2", "displacement_gray_24x28x1",
                      "displacement_cs_24x28x2", "displacement_gnc_20x24x1"]


@pytest.mark.parametrize("name", LEVEL_CASES)
def test_oracle_level_solver(oracle, name):
    m, z = load(name)
    K = z["J"].shape[-1]
    du, dv = oracle.level_solver(*z["J"], z["weight"], z["u"], z["v"], m["alpha"], m["iterations"],
                                 m["update_lag"], 0, np.full(K, m["a_data"]), m["a_smooth"], m["hx"],
                                 m["hy"], m.get("gnc_beta"))
    np.testing.assert_array_equal(du, z["du"])
    np.testing.assert_array_equal(dv, z["dv"])


@pytest.mark.parametrize("name", DISPLACEMENT_CASES)
def test_oracle_displacement(oracle, name):
    meta, z = load(name)
    _check_taps(meta, z, oracle.gauss_g)
    ref = oracle.preprocess(z["fixed_raw"].astype(np.float64)).astype(np.float32)
    mov = oracle.preprocess(z["moving_raw"], z["fixed_raw"].astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(ref, z["fixed"])
    np.testing.assert_array_equal(mov, z["moving"])
    kw = dict(meta["opts"])
    if meta["weight"] is not None:
        kw["weight"] = meta["weight"]
    if z["uv"].size:
        kw["uv"] = z["uv"]
    np.testing.assert_array_equal(oracle.get_displacement(ref, mov, **kw), z["w"])


def test_oracle_preprocess_and_median(oracle):
    _, z = load("preprocess")
    out = oracle.preprocess(z["raw"], z["ref"].astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(out, z["out"])
    from scipy.ndimage import median_filter
    _, z = load("median")
    np.testing.assert_array_equal(median_filter(z["f"], size=(5, 5), mode="mirror"), z["out"])


# ---------------------------------------------------------------------- HIP (GPU)

@pytest.mark.gpu
def test_gpu_resize(gpu):
    meta, z = load("resize_down")
    from pyflowreg_amd.optical_flow import _gauss_taps
    if not np.array_equal(_gauss_taps(meta["sigma"])[1], z["gauss"]):
        pytest.skip("this host's float32 Gaussian taps differ from the fixture's")
    np.testing.assert_array_equal(gpu.imresize_fused_gauss_cubic(z["img"], tuple(meta["size"])), z["out"])
    meta, z = load("resize_up")
    np.testing.assert_array_equal(gpu.imresize_fused_gauss_cubic(z["img"], tuple(meta["size"])), z["out"])


@pytest.mark.gpu
def test_gpu_warp(gpu):
    _, z = load("warp")
    np.testing.assert_array_equal(gpu.imregister_wrapper(z["f2"], z["u"], z["v"], z["f1"]), z["out"])


@pytest.mark.gpu
def test_gpu_tensor(gpu):
    meta, z = load("tensor_gc")
    J = gpu.get_motion_tensor_gc(z["f1"], z["f2"], meta["hy"], meta["hx"])
    np.testing.assert_array_equal(np.stack(J), z["J"])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["gc", "gray", "cs"])
def test_gpu_tensor_mixed(gpu, kind):
    meta, z = load(f"tensor_{kind}_mixed")
    fn = {"gc": gpu.get_motion_tensor_gc, "gray": gpu.get_motion_tensor_gray,
          "cs": gpu.get_motion_tensor_cs}[kind]
    np.testing.assert_array_equal(np.stack(fn(z["f1"], z["f2"], meta["hy"], meta["hx"])), z["J"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", LEVEL_CASES)
def test_gpu_level_solver(gpu, name):
    m, z = load(name)
    K = z["J"].shape[-1]
    du, dv = gpu.level_solver(*z["J"], z["weight"], z["u"], z["v"], m["alpha"], m["iterations"],
                              m["update_lag"], 0, np.full(K, m["a_data"]), m["a_smooth"], m["hx"], m["hy"],
                              m.get("gnc_beta"))
    np.testing.assert_array_equal(du, z["du"])
    np.testing.assert_array_equal(dv, z["dv"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", DISPLACEMENT_CASES)
def test_gpu_displacement(gpu, name):
    meta, z = load(name)
    from pyflowreg_amd.optical_flow import _gauss_taps
    _check_taps(meta, z, _gauss_taps)
    from pyflowreg_amd.preprocess import preprocess
    fr = z["fixed_raw"].astype(np.float64)
    ref = preprocess(fr).permute(0, 2, 3, 1)[0].cpu().numpy()
    mov = preprocess(z["moving_raw"], normalization_ref=fr).permute(0, 2, 3, 1)[0].cpu().numpy()
    np.testing.assert_array_equal(ref, z["fixed"])
    np.testing.assert_array_equal(mov, z["moving"])
    kw = dict(meta["opts"])
    if meta["weight"] is not None:
        kw["weight"] = meta["weight"]
    if z["uv"].size:
        kw["uv"] = z["uv"]
    np.testing.assert_array_equal(gpu.get_displacement(ref, mov, **kw), z["w"])


@pytest.mark.gpu
def test_gpu_median(gpu):
    from pyflowreg_amd.optical_flow import median5_mirror
    _, z = load("median")
    np.testing.assert_array_equal(median5_mirror(z["f"]), z["out"])
