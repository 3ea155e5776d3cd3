"""Exception hierarchy of the neighbour-search surface.

Mirrors the reference's public classes (``src/mlxlearn/exceptions.py:45-140``):
capability mismatches are ``ValueError``/``TypeError`` subclasses so code written
against scikit-learn's error contract keeps working; device failures are
``RuntimeError`` subclasses and are never hidden by a rerun elsewhere.
"""
from __future__ import annotations

__all__ = [
    "MLXLearnError", "CapabilityError", "UnsupportedParameterError", "UnsupportedInputError",
    "SparseInputError", "DeviceError", "BackendExecutionError",
]


class MLXLearnError(Exception):
    """Base class of every error raised on purpose."""


class CapabilityError(MLXLearnError):
    """The request is understood but this implementation does not serve it."""

    def __init__(self, message: str, *, reason: str = "capability") -> None:
        super().__init__(message)
        self.reason = reason


class UnsupportedParameterError(CapabilityError, ValueError):
    """A constructor parameter value the GPU path does not implement."""

    def __init__(self, message: str, *, parameter: str, value: object = None) -> None:
        super().__init__(message, reason=f"unsupported-parameter:{parameter}")
        self.parameter = parameter
        self.value = value


class UnsupportedInputError(CapabilityError, ValueError):
    """Input data the GPU path cannot represent (e.g. float64 overflowing float32)."""

    def __init__(self, message: str, *, reason: str = "unsupported-input") -> None:
        super().__init__(message, reason=reason)


class SparseInputError(CapabilityError, TypeError):
    """Sparse input given to a dense-only path (message contains "sparse")."""

    def __init__(self, message: str) -> None:
        super().__init__(message, reason="sparse-input")


class DeviceError(MLXLearnError, RuntimeError):
    """No usable MI355X / HIP device, or the native library is missing."""


class BackendExecutionError(MLXLearnError, RuntimeError):
    """A device computation failed; never resolved by silently rerunning elsewhere."""

    def __init__(self, message: str, *, original: BaseException | None = None) -> None:
        super().__init__(message)
        self.original = original
