"""HIP-event timing of individual kernel launches.

The reference wraps OpenCL profiling events (``sumpy/tools.py:720-765``);
here :func:`kernel_span` brackets a launch with two events on the current
stream when recording is switched on, and :func:`summary` reports per-kernel
launch counts, total time and work units.
"""

from __future__ import annotations

import contextlib

import torch

_state = {"enabled": False, "records": []}


def enable(on: bool = True) -> None:
    _state["enabled"] = on
    if on:
        _state["records"] = []


@contextlib.contextmanager
def kernel_span(name: str, units: float = 0.0):
    if not _state["enabled"]:
        yield
        return
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    yield
    end.record()
    _state["records"].append((name, start, end, units))


def summary() -> dict:
    """{name: {"launches", "ms", "units"}} over the recorded spans."""
    torch.cuda.synchronize()
    out: dict = {}
    for name, s, e, units in _state["records"]:
        d = out.setdefault(name, {"launches": 0, "ms": 0.0, "units": 0.0})
        d["launches"] += 1
        d["ms"] += s.elapsed_time(e)
        d["units"] += units
    return out
