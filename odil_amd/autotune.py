"""Launch-shape autotuner.

Some kernels take a launch parameter whose best value depends on the grid
and operator shape (threads per cell of the stored-stencil sweep, tile
shapes of batched launches). `pick` times every candidate on the caller's
own operator and buffers with HIP events, once per shape key, and caches the
fastest for the rest of the process. Each timed batch is queued behind a
short device-side spin so that the host has enqueued every launch before the
first one starts: the events then bracket back-to-back kernels and measure
device time, not the host's launch rate (which would hide the differences
between candidates on small grids). Set ODIL_AUTOTUNE=0 to skip tuning and
use the first candidate (the kernel's size-based default).
"""

import os

import torch

_CACHE = {}


def enabled():
    return os.environ.get('ODIL_AUTOTUNE', '1') != '0'


def cached(key):
    return _CACHE.get(key)


_SPIN_CYCLES = 4_000_000  # about 2 ms of device spin ahead of each batch


def _hold_device():
    spin = getattr(torch.cuda, '_sleep', None)
    if spin is not None:
        spin(_SPIN_CYCLES)


def pick(key, candidates, run, reps=10):
    """Fastest candidate for `key`. `run(c)` launches the kernel with
    candidate c on torch's current stream; it must be safe to repeat (write
    only scratch outputs). Not callable under HIP graph capture."""
    if key in _CACHE:
        return _CACHE[key]
    candidates = list(candidates)
    if not enabled() or len(candidates) == 1:
        _CACHE[key] = candidates[0]
        return candidates[0]
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError('autotune.pick called during graph capture')
    s = torch.cuda.current_stream()
    times = {}
    for c in candidates:
        run(c)  # warm-up (code object load, caches)
        _hold_device()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            run(c)
        e1.record(s)
        e1.synchronize()
        times[c] = e0.elapsed_time(e1) / reps
    best = min(candidates, key=lambda c: times[c])
    _CACHE[key] = best
    return best


def table():
    """Copy of the tuned choices (for reports)."""
    return dict(_CACHE)
