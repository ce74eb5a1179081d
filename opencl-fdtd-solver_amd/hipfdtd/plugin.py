"""Source / monitor callback registration (mirrors the reference's plugin.py:12-53).

Sources run after the H update with ``fdtd.t = t_n + dt/2``; monitors run
after the full step with ``t = t_{n+1}`` and ``step_num`` incremented.
"""

from __future__ import annotations

from typing import Any, Protocol, runtime_checkable


@runtime_checkable
class StepCallback(Protocol):
    """Callable invoked once per time step with the solver instance."""

    def __call__(self, fdtd: Any) -> None: ...


class SourceMonitorMixin:
    """``add_source`` / ``add_monitor`` / ``clear_*`` with the reference's errors."""

    _sources: list
    _monitors: list

    def add_source(self, source: StepCallback) -> StepCallback:
        """Register ``source(fdtd)``, called after each H update."""
        if not callable(source):
            raise TypeError(f"source must be callable, got {type(source)!r}")
        self._sources.append(source)
        return source

    def add_monitor(self, monitor: StepCallback) -> StepCallback:
        """Register ``monitor(fdtd)``, called after each full step."""
        if not callable(monitor):
            raise TypeError(f"monitor must be callable, got {type(monitor)!r}")
        self._monitors.append(monitor)
        return monitor

    def clear_sources(self) -> None:
        self._sources.clear()

    def clear_monitors(self) -> None:
        self._monitors.clear()
