"""Opt-in per-kernel timing with HIP events recorded on the launching stream.

`enable()` turns recording on; each instrumented launch records a (start, end) event pair under
its name; `summary()` synchronises and returns per-name launch counts and total/mean milliseconds.
Off by default (no events are created).
"""
from __future__ import annotations

import contextlib
from collections import defaultdict

import torch

_enabled = False
_events = defaultdict(list)


def enable(on: bool = True):
    global _enabled
    _enabled = on


def reset():
    _events.clear()


@contextlib.contextmanager
def span(name: str):
    if not _enabled:
        yield
        return
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record(torch.cuda.current_stream())
    try:
        yield
    finally:
        e.record(torch.cuda.current_stream())
        _events[name].append((s, e))


def summary():
    torch.cuda.synchronize()
    out = {}
    for name, pairs in _events.items():
        ms = [s.elapsed_time(e) for s, e in pairs]
        out[name] = {"launches": len(ms), "total_ms": sum(ms), "mean_ms": sum(ms) / max(len(ms), 1)}
    return out
