"""Configuration: a six-option public ``Config`` and a private ``Tuning``.

Restates the two layers of the reference (``src/mlxlearn/_common/config.py``):

``Config`` is public, immutable and thread-local, and ``config_context`` scopes
it, so a library that sets a policy cannot leak it into its caller's thread. It
keeps the reference's six options and their meaning on this path:

- ``device``: ``"auto"`` and ``"gpu"`` both compute on the GPU. ``"cpu"`` is
  rejected, because there is no CPU route and a silent one would hide a missing
  device.
- ``fallback_policy``: accepted for compatibility. Every capability mismatch
  raises ``UnsupportedParameterError`` under each policy, because nothing here
  falls back.
- ``output_type``: ``"numpy"`` only, as in the reference.
- ``deterministic``: every path is already deterministic. Results never depend
  on plans, streams or thread count.
- ``random_state``: seeds KMeans initialisation when the estimator's own
  ``random_state`` is ``None``.
- ``diagnostics``: whether backend events are recorded.

``Tuning`` is private. It holds the native screen-plan knobs, a plan-cache file
loaded at start-up and the diagnostics capacity. ``MLXLEARN_<FIELD>``
environment variables override it; a malformed value raises rather than being
ignored (reference ``config.py:251-275``). The plan knobs live in the native
library and are process-wide, unlike the thread-local ``Config``, and
``tuning_context`` is for tests and debugging only. Results never depend on
tuning, only the time does.
"""
from __future__ import annotations

import os
import threading
from contextlib import contextmanager
from dataclasses import dataclass, fields, replace
from typing import Any, Literal, get_args

__all__ = ["Config", "Tuning", "get_config", "set_config", "config_context", "reset_config", "get_tuning",
           "tuning_context", "PUBLIC_CONFIG_FIELDS"]

Device = Literal["auto", "gpu", "cpu"]
FallbackPolicy = Literal["warn", "raise", "silent"]
OutputType = Literal["numpy", "input", "mlx"]


def _choice(name, value, allowed):
    if value not in allowed:
        raise ValueError(f"{name} must be one of {', '.join(repr(a) for a in allowed)}; got {value!r}")


@dataclass(frozen=True)
class Config:
    """User-visible configuration; replace it, do not mutate it."""

    device: Device = "auto"
    fallback_policy: FallbackPolicy = "warn"
    output_type: OutputType = "numpy"
    deterministic: bool = True
    random_state: int | None = None
    diagnostics: bool = True

    def __post_init__(self):
        _choice("device", self.device, get_args(Device))
        _choice("fallback_policy", self.fallback_policy, get_args(FallbackPolicy))
        _choice("output_type", self.output_type, get_args(OutputType))
        if self.device == "cpu":
            raise ValueError("device='cpu' is not available: this library computes exact neighbours on an "
                             "MI355X GPU only and has no CPU route.")
        if self.output_type != "numpy":
            raise ValueError(f"output_type={self.output_type!r} is not implemented; estimators take NumPy in "
                             "and return NumPy out.")
        if not isinstance(self.deterministic, bool):
            raise TypeError(f"deterministic must be a bool, got {type(self.deterministic).__name__}")
        if self.random_state is not None and (isinstance(self.random_state, bool)
                                              or not isinstance(self.random_state, int)):
            raise TypeError(f"random_state must be an int or None, got {type(self.random_state).__name__}")
        if not isinstance(self.diagnostics, bool):
            raise TypeError(f"diagnostics must be a bool, got {type(self.diagnostics).__name__}")


PUBLIC_CONFIG_FIELDS: tuple[str, ...] = tuple(f.name for f in fields(Config))


@dataclass(frozen=True)
class Tuning:
    """Private knobs; 0 / empty means automatic (the autotuned or measured default plan)."""

    screen_splits: int = 0         # database splits per screen round
    candidate_capacity: int = 0    # candidate slots per (query, split, lane half)
    query_tiles: int = 0           # query tiles of 32 per screen wave (1-2)
    screen_stages: int = 0         # stages of the screen's LDS tile ring (2-4)
    screen_rounds: str = ""        # cumulative round fractions, comma-separated, increasing, ending at 1
    plan_file: str = ""            # plan-cache text (as exported by the autotuner) imported when the library loads
    diagnostics_capacity: int = 512

    def __post_init__(self):
        for f in fields(self):
            value = getattr(self, f.name)
            if isinstance(value, int) and not isinstance(value, bool) and value < 0:
                raise ValueError(f"tuning value {f.name} must be non-negative, got {value}")
        if self.query_tiles > 2:
            raise ValueError(f"query_tiles must be 0, 1 or 2, got {self.query_tiles}")
        if self.screen_stages == 1 or self.screen_stages > 4:
            raise ValueError(f"screen_stages must be 0 or 2-4, got {self.screen_stages}")
        if self.diagnostics_capacity < 1:
            raise ValueError("diagnostics_capacity must be at least 1")
        self.rounds()

    def rounds(self) -> list[float]:
        if not self.screen_rounds.strip():
            return []
        try:
            out = [float(x) for x in self.screen_rounds.split(",")]
        except ValueError as exc:
            raise ValueError(f"screen_rounds must be comma-separated numbers, got {self.screen_rounds!r}") from exc
        prev = 0.0
        for f in out:
            if not (prev < f <= 1.0):
                raise ValueError(f"screen_rounds must increase within (0, 1], got {self.screen_rounds!r}")
            prev = f
        return out


_ENV_PREFIX = "MLXLEARN_"


def _tuning_from_env(base: Tuning) -> Tuning:
    overrides: dict[str, Any] = {}
    for f in fields(Tuning):
        key = _ENV_PREFIX + f.name.upper()
        raw = os.environ.get(key)
        if raw is None:
            continue
        kind = f.type if isinstance(f.type, str) else f.type.__name__
        try:
            overrides[f.name] = int(raw) if kind == "int" else raw
        except ValueError as exc:
            raise ValueError(f"{key}={raw!r} is not a valid value for {f.name}") from exc
    try:
        return replace(base, **overrides) if overrides else base
    except ValueError as exc:
        raise ValueError(f"{_ENV_PREFIX}* tuning override rejected: {exc}") from exc


class _State(threading.local):
    def __init__(self):
        self.config = Config()


_state = _State()
_tuning_lock = threading.Lock()
_tuning: Tuning | None = None
_applied: Tuning | None = None


def get_config() -> Config:
    """This thread's configuration."""
    return _state.config


def set_config(**kwargs: Any) -> Config:
    """Change the named fields of this thread's configuration; unknown names raise TypeError."""
    known = set(PUBLIC_CONFIG_FIELDS)
    unknown = set(kwargs) - known
    if unknown:
        raise TypeError(f"unknown configuration option(s): {', '.join(sorted(unknown))}. "
                        f"Valid options are: {', '.join(sorted(known))}.")
    _state.config = replace(_state.config, **kwargs)
    return _state.config


@contextmanager
def config_context(**kwargs: Any):
    """Override this thread's configuration inside the block; restored even if it raises."""
    previous = _state.config
    try:
        set_config(**kwargs)
        yield _state.config
    finally:
        _state.config = previous


def reset_config() -> None:
    """Restore this thread's configuration, and the process tuning, to their defaults (environment included)."""
    global _tuning
    _state.config = Config()
    with _tuning_lock:
        _tuning = None
    get_tuning()


def get_tuning() -> Tuning:
    """The process tuning: defaults overridden by MLXLEARN_* variables on first use."""
    global _tuning
    with _tuning_lock:
        if _tuning is None:
            _tuning = _tuning_from_env(Tuning())
        current = _tuning
    _resize_diagnostics(current.diagnostics_capacity)
    return current


@contextmanager
def tuning_context(**kwargs: Any):
    """Override tuning inside the block (tests and debugging; the native plan knobs are process-wide)."""
    global _tuning
    known = {f.name for f in fields(Tuning)}
    unknown = set(kwargs) - known
    if unknown:
        raise TypeError(f"unknown tuning option(s): {', '.join(sorted(unknown))}")
    previous = get_tuning()
    new = replace(previous, **kwargs)
    with _tuning_lock:
        _tuning = new
    _resize_diagnostics(new.diagnostics_capacity)
    try:
        if _native_loaded():
            apply_tuning(force=True)
        yield new
    finally:
        with _tuning_lock:
            _tuning = previous
        _resize_diagnostics(previous.diagnostics_capacity)
        if _native_loaded():
            apply_tuning(force=True)


def _resize_diagnostics(capacity: int) -> None:
    from . import _diagnostics
    _diagnostics.resize(capacity)


def _native_loaded() -> bool:
    from . import _lib
    return _lib._LIB is not None


def apply_tuning(force: bool = False) -> None:
    """Push the plan knobs (and, once per file, the plan cache) into the native library."""
    global _applied
    from . import _lib
    t = get_tuning()
    if not force and _applied == t:
        return
    _lib.set_tuning(t.screen_splits, t.candidate_capacity, t.query_tiles, t.screen_stages, t.rounds())
    if t.plan_file and (_applied is None or _applied.plan_file != t.plan_file):
        with open(t.plan_file) as fh:
            _lib.import_plans(fh.read())
    _applied = t
