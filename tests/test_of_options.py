"""OFOptions' normalisation and resolution rules against the reference's
(motion_correction/OF_options.py:525-724,1021-1101) and the weight expansion
compensate_arr derives from them (compensate_recording.py:504-516).  CPU
only."""

import numpy as np
import pytest

from pyflowreg_amd.compensate import RegistrationOptions, _weight_maps
from pyflowreg_amd.of_options import OFOptions


def test_defaults_match_the_reference():
    o = OFOptions()
    assert o.alpha == (1.5, 1.5) and o.weight == [0.5, 0.5] and o.levels == 100
    assert o.min_level == -1 and o.quality_setting == "quality" and o.effective_min_level == 0
    assert (o.eta, o.update_lag, o.iterations, o.a_smooth, o.a_data) == (0.8, 5, 50, 1.0, 0.45)
    assert o.sigma == [[1.0, 1.0, 0.1], [1.0, 1.0, 0.1]] and o.buffer_size == 400
    assert o.output_typename == "double" and o.channel_normalization == "joint"
    assert o.interpolation_method == "cubic" and o.constancy_assumption == "gc"
    assert set(o.to_dict()) == {"alpha", "weight", "levels", "min_level", "eta", "iterations", "update_lag",
                                "a_data", "a_smooth", "gnc_schedule", "warping_steps", "const_assumption"}


def test_field_normalisation():
    assert OFOptions(alpha=3).alpha == (3.0, 3.0)
    assert OFOptions(alpha=[2]).alpha == (2.0, 2.0)
    assert OFOptions(alpha=(1, 4)).alpha == (1.0, 4.0)
    for bad in (0, -1, (1, 0), (1, 2, 3)):
        with pytest.raises(ValueError):
            OFOptions(alpha=bad)
    np.testing.assert_allclose(OFOptions(weight=[1, 3]).weight, [0.25, 0.75])
    np.testing.assert_allclose(OFOptions(weight=np.array([2.0, 2.0])).weight, [0.5, 0.5])
    m = np.ones((4, 5, 2))
    assert OFOptions(weight=m).weight is m
    assert OFOptions(sigma=[1, 2, 0.5]).sigma == [[1.0, 2.0, 0.5]]
    with pytest.raises(ValueError):
        OFOptions(sigma=[1, 2])
    assert OFOptions(gnc_schedule=[0, 0.5, 1]).gnc_schedule == (0.0, 0.5, 1.0)
    for bad in ([0.5, 1.0], [0.0], [0.0, 0.7, 0.5, 1.0], [0.0, 0.9]):
        with pytest.raises(ValueError):
            OFOptions(gnc_schedule=bad)
    assert OFOptions(constancy_assumption="Census").constancy_assumption == "cs"
    assert OFOptions(constancy_assumption="brightness").constancy_assumption == "gray"
    for kw in (dict(levels=0), dict(eta=0.0), dict(a_data=1.5), dict(iterations=2.5), dict(update_lag=0)):
        with pytest.raises(ValueError):
            OFOptions(**kw)


def test_quality_presets_and_custom_min_level():
    assert OFOptions(quality_setting="balanced").effective_min_level == 4
    assert OFOptions(quality_setting="fast").effective_min_level == 6
    o = OFOptions(quality_setting="fast", min_level=2)
    assert o.quality_setting == "custom" and o.effective_min_level == 2
    # validated construction: "custom" with min_level -1 takes the default preset
    assert OFOptions(quality_setting="custom").quality_setting == "quality"
    # model_copy does not re-validate (pydantic): still "custom", level max(-1, 0)
    c = o.model_copy(update={"min_level": -1})
    assert c.quality_setting == "custom" and c.effective_min_level == 0
    b = OFOptions().model_copy(update={"quality_setting": "balanced"})
    assert b.effective_min_level == 4
    with pytest.raises(ValueError):
        o.model_copy(update={"no_such_field": 1})


def test_get_weight_at_and_weight_maps():
    o = OFOptions(weight=[1, 1, 2])
    assert o.get_weight_at(2, 4) == 0.5 and o.get_weight_at(3, 4) == 0.25
    o2 = OFOptions(weight=[1, 1, 2])
    np.testing.assert_allclose([o2.get_weight_at(i, 2) for i in range(2)], [0.5, 0.5])  # cut, renormalised
    r = RegistrationOptions.from_any(OFOptions(weight=[1, 1, 2]))
    np.testing.assert_allclose(_weight_maps(r.weight, 3, 4, 2)[1, 2], [0.5, 0.5])
    np.testing.assert_allclose(_weight_maps(r.weight, 3, 4, 4)[0, 0], [0.25, 0.25, 0.5, 0.25])
    np.testing.assert_allclose(_weight_maps(RegistrationOptions(weight=[3.0]).weight, 2, 2, 2)[0, 0], [1.0, 1.0])
    wm = np.random.default_rng(0).random((3, 4))
    maps = _weight_maps(wm, 3, 4, 2)
    np.testing.assert_array_equal(maps[:, :, 0], wm)
    np.testing.assert_array_equal(maps[:, :, 1], 0.5)
    o3 = OFOptions(weight=wm)
    np.testing.assert_array_equal(o3.get_weight_at(0, 2), wm)


def test_resolve_get_displacement_priority():
    def impl(f, m, **kw):
        return "impl"
    made = []

    def factory(**p):
        made.append(p)
        return lambda f, m, **kw: "factory"
    assert OFOptions(get_displacement_impl=impl, get_displacement_factory=factory).resolve_get_displacement() is impl
    fn = OFOptions(get_displacement_factory=factory, backend_params={"x": 1}).resolve_get_displacement()
    assert fn(None, None) == "factory" and made == [{"x": 1}]
    from pyflowreg_amd.optical_flow import get_displacement
    import pyflowreg_amd._lib as L
    import os
    if os.path.exists(L.LIB_PATH):
        assert OFOptions().resolve_get_displacement() is get_displacement
    with pytest.raises(ValueError, match="Unknown flow backend"):
        OFOptions(flow_backend="nope").resolve_get_displacement()


def test_registration_options_from_ofoptions():
    o = OFOptions(alpha=2, gnc_schedule=[0, 1], warping_steps=3, quality_setting="balanced",
                  interpolation_method="linear", output_typename="single")
    r = RegistrationOptions.from_any(o)
    fp = r.flow_params()
    assert fp["alpha"] == (2.0, 2.0) and fp["gnc_schedule"] == (0.0, 1.0) and fp["warping_steps"] == 3
    assert fp["min_level"] == 4 and r.interpolation_method == "linear" and r.output_typename == "single"
    assert r.channel_normalization == "together"
    with pytest.raises(ValueError, match="bin_size"):
        RegistrationOptions.from_any(OFOptions(bin_size=2))
