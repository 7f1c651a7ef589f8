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

ODIL_AUTOTUNE_FILE names a JSON table of choices (key and choice as their
Python repr): keys found there are not timed again, and every new choice is
written back. A second process with the same table launches the same shapes
(e.g. under a counter profiler, whose serialised dispatches would distort
the timing).
"""

import json
import os

import torch

_CACHE = {}
_FILE = {'path': None, 'table': {}}


def enabled():
    return os.environ.get('ODIL_AUTOTUNE', '1') != '0'


def cached(key):
    return _CACHE.get(key)


_SPIN_CYCLES = 4_000_000  # about 2 ms of device spin ahead of each batch


def _hold_device():
    spin = getattr(torch.cuda, '_sleep', None)
    if spin is not None:
        spin(_SPIN_CYCLES)


def _file_table():
    path = os.environ.get('ODIL_AUTOTUNE_FILE')
    if path != _FILE['path']:
        table = {}
        if path and os.path.exists(path):
            with open(path) as f:
                table = json.load(f)
        _FILE['path'], _FILE['table'] = path, table
    return _FILE['table'] if _FILE['path'] else None


def _from_file(key, candidates):
    table = _file_table()
    if table is None or repr(key) not in table:
        return None
    for c in candidates:
        if repr(c) == table[repr(key)]:
            return c
    return None


def _to_file(key, choice):
    table = _file_table()
    if table is None:
        return
    table[repr(key)] = repr(choice)
    tmp = _FILE['path'] + '.tmp'
    with open(tmp, 'w') as f:
        json.dump(table, f, indent=0, sort_keys=True)
    os.replace(tmp, _FILE['path'])


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
    stored = _from_file(key, candidates)
    if stored is not None:
        _CACHE[key] = stored
        return stored
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
    _to_file(key, best)
    return best


def table():
    """Copy of the tuned choices (for reports)."""
    return dict(_CACHE)
