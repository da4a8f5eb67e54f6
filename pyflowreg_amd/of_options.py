"""``OFOptions`` with the reference's fields, defaults and normalisation
(motion_correction/OF_options.py:401-500 fields, :525-646 validators,
:649-661 effective_min_level, :679-724 get_weight_at, :1021-1101
resolve_get_displacement / to_dict), for code written against pyflowreg's
options object.

String values stand in for the reference's enums (an enum member with a
``.value`` is accepted too).  The I/O fields (input and output paths and
formats, reference frame selection, save flags, naming) are kept as given:
this package's ``compensate_arr`` works on in-memory arrays and reads only
the registration fields.  The default ``flow_backend`` is ``"flowreg_hip"``.
"""

from __future__ import annotations

import copy as _copy
from dataclasses import dataclass, field, fields
from pathlib import Path
from typing import Any, Callable, Dict, Optional

import numpy as np

_QUALITY_MIN_LEVEL = {"quality": 0, "balanced": 4, "fast": 6}
_CONSTANCY_ALIASES = {"gc": "gc", "gradient": "gc", "gray": "gray", "brightness": "gray", "cs": "cs",
                      "census": "cs"}


def _value(v):
    return getattr(v, "value", v)


def normalize_alpha(v):
    """OF_options.py:525-545: a positive scalar, 1- or 2-sequence -> a
    2-tuple of floats."""
    if isinstance(v, (int, float, np.integer, np.floating)):
        if v <= 0:
            raise ValueError("Alpha must be positive")
        return (float(v), float(v))
    v = tuple(v)
    if len(v) == 1:
        if v[0] <= 0:
            raise ValueError("Alpha must be positive")
        return (float(v[0]), float(v[0]))
    if len(v) == 2:
        if v[0] <= 0 or v[1] <= 0:
            raise ValueError("All alpha values must be positive")
        return (float(v[0]), float(v[1]))
    raise ValueError("Alpha must be a scalar or a 2-element sequence")


def normalize_weight(v):
    """OF_options.py:547-584: 1-D channel weights are divided by their sum
    (when positive); (H, W) and (H, W, C) weight maps are kept."""
    if isinstance(v, np.ndarray):
        if v.ndim == 1:
            s = v.sum()
            return (v / s).tolist() if s > 0 else v.tolist()
        if v.ndim <= 3:
            return v
        raise ValueError(f"Weight array cannot exceed 3 dimensions (got {v.ndim}D array).")
    if isinstance(v, (list, tuple)):
        arr = np.asarray(v, dtype=float)
        if arr.ndim == 1:
            s = arr.sum()
            if s > 0:
                return (arr / s).tolist()
        return list(v)
    return v


def normalize_sigma(v):
    """OF_options.py:586-601: [sx, sy, st] or (n_channels, 3)."""
    sig = np.asarray(v, dtype=float)
    if sig.ndim == 1:
        if sig.size != 3:
            raise ValueError("1D sigma must be [sx, sy, st]")
        return sig.reshape(1, 3).tolist()
    if sig.ndim == 2:
        if sig.shape[1] != 3:
            raise ValueError("2D sigma must be (n_channels, 3)")
        return sig.tolist()
    raise ValueError("Sigma must be [sx,sy,st] or (n_channels, 3)")


def normalize_gnc_schedule(v):
    """OF_options.py:603-623."""
    if v is None:
        return None
    s = np.asarray(v, dtype=float)
    if s.ndim != 1:
        raise ValueError("gnc_schedule must be a 1D sequence")
    if s.size < 2:
        raise ValueError("gnc_schedule must contain at least two stages")
    if np.any(s < 0.0) or np.any(s > 1.0):
        raise ValueError("gnc_schedule entries must lie in [0, 1]")
    if not np.all(np.diff(s) >= 0.0):
        raise ValueError("gnc_schedule must be monotone nondecreasing")
    if not np.isclose(s[0], 0.0):
        raise ValueError("gnc_schedule must start at 0.0")
    if not np.isclose(s[-1], 1.0):
        raise ValueError("gnc_schedule must end at 1.0")
    return tuple(float(x) for x in s)


def normalize_constancy_assumption(v):
    """OF_options.py:218-230: aliases to "gc" / "gray" / "cs"; other values
    pass through (get_displacement rejects them)."""
    v = _value(v)
    if isinstance(v, str):
        key = v.strip().lower()
        return _CONSTANCY_ALIASES.get(key, key)
    return v


@dataclass
class OFOptions:
    """The reference's option set (defaults as OF_options.py:401-500)."""
    # I/O
    input_file: Any = None
    output_path: Any = Path("results")
    output_format: Any = "MAT"
    output_file_name: Optional[str] = None
    channel_idx: Optional[list] = None
    # flow parameters
    alpha: Any = (1.5, 1.5)
    weight: Any = field(default_factory=lambda: [0.5, 0.5])
    levels: int = 100
    min_level: int = -1
    quality_setting: Any = "quality"
    eta: float = 0.8
    update_lag: int = 5
    iterations: int = 50
    a_smooth: float = 1.0
    a_data: float = 0.45
    gnc_schedule: Any = None
    warping_steps: Optional[int] = None
    # preprocessing
    sigma: Any = field(default_factory=lambda: [[1.0, 1.0, 0.1], [1.0, 1.0, 0.1]])
    bin_size: int = 1
    buffer_size: int = 400
    # reference
    reference_frames: Any = field(default_factory=lambda: list(range(50, 500)))
    update_reference: bool = False
    n_references: int = 1
    min_frames_per_reference: int = 20
    # processing
    verbose: bool = False
    save_meta_info: bool = True
    save_w: bool = False
    save_valid_mask: bool = False
    save_valid_idx: bool = False
    output_typename: Optional[str] = "double"
    channel_normalization: Any = "joint"
    interpolation_method: Any = "cubic"
    cc_initialization: bool = False
    cc_hw: Any = 256
    cc_up: int = 1
    update_initialization_w: bool = True
    naming_convention: Any = "default"
    constancy_assumption: Any = "gc"
    # backend
    flow_backend: str = "flowreg_hip"
    backend_params: Dict[str, Any] = field(default_factory=dict)
    preproc_funct: Optional[Callable] = None
    get_displacement_impl: Optional[Callable] = None
    get_displacement_factory: Optional[Callable[..., Callable]] = None

    def __post_init__(self):
        self.alpha = normalize_alpha(self.alpha)
        self.weight = normalize_weight(self.weight)
        self.sigma = normalize_sigma(self.sigma)
        self.gnc_schedule = normalize_gnc_schedule(self.gnc_schedule)
        self.constancy_assumption = normalize_constancy_assumption(self.constancy_assumption)
        for name, lo in (("levels", 1), ("min_level", -1), ("update_lag", 1), ("iterations", 1),
                         ("bin_size", 1), ("buffer_size", 1), ("n_references", 1),
                         ("min_frames_per_reference", 1), ("cc_up", 1)):
            v = getattr(self, name)
            if isinstance(v, bool) or not isinstance(v, (int, np.integer)):
                raise ValueError(f"{name} must be an integer")
            if v < lo:
                raise ValueError(f"{name} must be >= {lo}")
        if self.warping_steps is not None and (not isinstance(self.warping_steps, (int, np.integer))
                                               or self.warping_steps < 1):
            raise ValueError("warping_steps must be an integer >= 1")
        if not 0 < self.eta <= 1:
            raise ValueError("eta must be in (0, 1]")
        if self.a_smooth < 0:
            raise ValueError("a_smooth must be >= 0")
        if not 0 < self.a_data <= 1:
            raise ValueError("a_data must be in (0, 1]")
        self.output_path = Path(self.output_path)
        self.quality_setting = str(_value(self.quality_setting)).lower()
        self.channel_normalization = str(_value(self.channel_normalization)).lower()
        self.interpolation_method = str(_value(self.interpolation_method)).lower()
        # OF_options.py:638-646: a set min_level makes the preset "custom";
        # "custom" with min_level -1 falls back to the previous preset
        old = getattr(self, "_quality_setting_old", "quality")
        if self.quality_setting != "custom":
            old = self.quality_setting
        self._quality_setting_old = old
        if self.min_level >= 0:
            self.quality_setting = "custom"
        elif self.quality_setting == "custom":
            self.quality_setting = old

    @property
    def effective_min_level(self) -> int:
        """OF_options.py:649-661."""
        if self.min_level >= 0:
            return self.min_level
        if self.quality_setting == "custom":
            return max(self.min_level, 0)
        return _QUALITY_MIN_LEVEL.get(self.quality_setting, 0)

    def get_weight_at(self, i: int, n_channels: int):
        """OF_options.py:679-724: channel i's weight, a constant or an
        (H, W) map."""
        w = np.asarray(self.weight, dtype=float)
        if w.ndim <= 1:
            if w.size == 1:
                return float(w.item())
            if w.size > n_channels:
                w = w[:n_channels]
                w = w / w.sum()
                self.weight = w.tolist()
            if i >= w.size:
                return 1.0 / n_channels
            return float(w[i])
        if w.ndim == 2:
            return w if i == 0 else np.ones_like(w) / n_channels
        if w.ndim == 3:
            if i >= w.shape[2]:
                return np.ones(w.shape[:2]) / n_channels
            return w[:, :, i]
        raise ValueError(f"Unexpected weight array with {w.ndim} dimensions")

    def to_dict(self) -> dict:
        """OF_options.py:1084-1101: the keyword arguments of the
        displacement callable."""
        return {"alpha": self.alpha, "weight": self.weight, "levels": self.levels,
                "min_level": self.effective_min_level, "eta": self.eta, "iterations": self.iterations,
                "update_lag": self.update_lag, "a_data": self.a_data, "a_smooth": self.a_smooth,
                "gnc_schedule": self.gnc_schedule, "warping_steps": self.warping_steps,
                "const_assumption": self.constancy_assumption}

    def resolve_get_displacement(self) -> Callable:
        """OF_options.py:1021-1082: get_displacement_impl, else
        get_displacement_factory(**backend_params), else the registered
        backend's factory(**backend_params)."""
        if self.get_displacement_impl is not None:
            return self.get_displacement_impl
        if self.get_displacement_factory is not None:
            return self.get_displacement_factory(**self.backend_params)
        from .backend_registry import get_backend
        return get_backend(self.flow_backend)(**self.backend_params)

    def copy(self) -> "OFOptions":
        return _copy.deepcopy(self)

    def model_copy(self, update: Optional[dict] = None, deep: bool = False) -> "OFOptions":
        """pydantic's model_copy as the reference uses it (flow_reg_live.py:
        128-130): a copy with fields replaced, without re-validation (as
        pydantic's model_copy)."""
        new = _copy.deepcopy(self) if deep else _copy.copy(self)
        for k, v in (update or {}).items():
            if k not in {f.name for f in fields(self)}:
                raise ValueError(f"OFOptions has no field {k!r}")
            setattr(new, k, v)
        return new
