"""Backend events: what ran each estimator call, with the native per-call statistics.

Restates the recorder of the reference (``src/mlxlearn/_common/diagnostics.py``):
a process-wide, lock-guarded ring buffer of ``BackendEvent`` records, bounded by
the ``diagnostics_capacity`` tuning value, written only while
``Config.diagnostics`` is true. Every event here has backend ``"gpu"`` and
reason ``"selected"``, because there is no fallback route. Where the call ran a
native search, ``extra`` carries that search's statistics:
- the screen precision actually used;
- candidates appended;
- candidates re-ranked in float64;
- queries resolved by the exact fallback scan;
- rounds, splits and the candidate capacity;
- device milliseconds of the screen and of the whole search.
"""
from __future__ import annotations

import threading
from collections import deque
from dataclasses import asdict, dataclass, field
from typing import Any

__all__ = ["BackendEvent", "record_event", "get_backend_diagnostics", "get_last_backend_event",
           "clear_backend_diagnostics", "diagnostics_summary"]


@dataclass(frozen=True)
class BackendEvent:
    """One estimator call: ``estimator.operation`` on ``shape`` ran on ``backend``."""

    estimator: str
    operation: str
    backend: str = "gpu"
    reason: str = "selected"
    detail: str = ""
    shape: tuple[int, ...] | None = None
    fallback: bool = False
    extra: dict[str, Any] = field(default_factory=dict)

    def as_dict(self) -> dict[str, Any]:
        return asdict(self)

    def __str__(self) -> str:
        where = f" {self.shape}" if self.shape else ""
        return f"{self.estimator}.{self.operation}{where} -> {self.backend}"


class _Recorder:
    def __init__(self, capacity: int = 512):
        self._lock = threading.Lock()
        self._events: deque[BackendEvent] = deque(maxlen=capacity)

    def record(self, event):
        with self._lock:
            self._events.append(event)

    def events(self):
        with self._lock:
            return list(self._events)

    def last(self):
        with self._lock:
            return self._events[-1] if self._events else None

    def clear(self):
        with self._lock:
            self._events.clear()

    def resize(self, capacity):
        with self._lock:
            if self._events.maxlen != capacity:
                self._events = deque(self._events, maxlen=capacity)


_recorder = _Recorder()


def resize(capacity: int) -> None:
    _recorder.resize(int(capacity))


def native_stats() -> dict[str, Any]:
    """Statistics of the calling thread's last native search, keyed for an event's ``extra``."""
    from . import _lib
    s = _lib.stats()
    names = {code: name for name, code in _lib.PRECISIONS.items()}
    return {"screen_precision": names.get(int(s["screen_precision"]), str(s["screen_precision"])),
            "queries": int(s["queries"]), "candidates": int(s["candidates"]), "reranked": int(s["reranked"]),
            "fallback_queries": int(s["fallback_queries"]), "rounds": int(s["rounds"]), "splits": int(s["splits"]),
            "capacity": int(s["capacity"]), "screen_ms": float(s["screen_ms"]), "total_ms": float(s["total_ms"])}


def record_event(estimator: str, operation: str, *, shape=None, detail: str = "", search: bool = True,
                 **extra: Any) -> BackendEvent | None:
    """Record one call when diagnostics are on; ``search`` attaches the thread's last native statistics."""
    from ._config import get_config, get_tuning
    if not get_config().diagnostics:
        return None
    get_tuning()  # sizes the ring from the tuning on first use
    info = dict(native_stats()) if search else {}
    info.update(extra)
    event = BackendEvent(estimator=estimator, operation=operation, detail=detail,
                         shape=tuple(int(v) for v in shape) if shape is not None else None, extra=info)
    _recorder.record(event)
    return event


def get_backend_diagnostics() -> list[BackendEvent]:
    """Recorded events, oldest first (empty when diagnostics are off)."""
    return _recorder.events()


def get_last_backend_event() -> BackendEvent | None:
    return _recorder.last()


def clear_backend_diagnostics() -> None:
    _recorder.clear()


def diagnostics_summary() -> dict[str, int]:
    """Event counts by ``"estimator.operation -> backend"``."""
    counts: dict[str, int] = {}
    for event in _recorder.events():
        key = f"{event.estimator}.{event.operation} -> {event.backend}"
        counts[key] = counts.get(key, 0) + 1
    return counts
