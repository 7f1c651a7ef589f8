"""Launch-shape autotuner.

Some kernels take a launch parameter whose best value depends on the grid
and operator shape (threads per cell of the stored-stencil sweep, tile
shapes of batched launches). `pick` times every candidate on the caller's
own operator and buffers with HIP events, once per shape key, and caches the
fastest for the rest of the process. Each timed batch is queued behind a
short device-side spin so that the host has enqueued every launch before the
first one starts: the events then bracket back-to-back kernels and measure
device time, not the host's launch rate (which would hide the differences
between candidates on small grids). `pick_in_situ` re-times a kernel's
candidates through a larger workload that contains it (a whole multigrid
cycle), where the other kernels' traffic between its launches changes what
the caches hold, and replaces the isolated choice. Set ODIL_AUTOTUNE=0 to
skip tuning and use the first candidate (the kernel's size-based default).

ODIL_AUTOTUNE_FILE names a JSON table of choices (key and choice as their
Python repr): keys found there are not timed again, and every new choice is
written back. A second process with the same table launches the same shapes
(e.g. under a counter profiler, whose serialised dispatches would distort
the timing).
"""

import json
import os
import time

import torch

_CACHE = {}
_FILE = {'path': None, 'table': {}}


def enabled():
    return os.environ.get('ODIL_AUTOTUNE', '1') != '0'


def cached(key):
    return _CACHE.get(key)


_SPIN_CYCLES = 4_000_000  # about 2 ms of device spin ahead of each batch


def _hold_device(seconds=None):
    """Queue a device-side spin of about 2 ms, or of `seconds`."""
    spin = getattr(torch.cuda, '_sleep', None)
    if spin is not None:
        cycles = _SPIN_CYCLES if seconds is None else \
            max(_SPIN_CYCLES, int(seconds * _SPIN_CYCLES / 2e-3))
        spin(cycles)


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


def pick_in_situ(key, candidates, install, run, reps=4, batches=3):
    """Fastest candidate for `key` timed inside a larger workload.

    `install(c)` makes candidate c the kernel's launch shape and `run()`
    executes a workload that contains that kernel among others (e.g. one
    multigrid cycle, in which the other levels evict this level's data from
    the caches between its launches, as they do in the solver). The kernel
    alone, relaunched on the same buffers, runs from the caches and can
    rank the candidates differently. Each candidate is timed over `batches`
    batches of `reps` workloads (the fastest batch counts); the choice
    replaces any choice timed in isolation under `key`, in this process and
    in the ODIL_AUTOTUNE_FILE table (where it is marked so that a second
    process installs it without timing). Returns the choice, or None when
    tuning is off. Not callable under HIP graph capture."""
    candidates = list(candidates)
    if not enabled() or not candidates:
        return None
    mark = ('in_situ', ) + tuple(key)
    done = _CACHE.get(mark)
    if done is None:
        done = _from_file(mark, candidates)
    if done is not None:
        install(done)
        _CACHE[key] = _CACHE[mark] = done
        return done
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError('autotune.pick_in_situ called during graph capture')
    s = torch.cuda.current_stream()
    times = {}
    for c in candidates:
        install(c)
        run()  # warm-up
        # The spin ahead of each batch outlasts the host's enqueueing of the
        # whole batch (twice the measured time of one workload's launches).
        t0 = time.perf_counter()
        run()
        host = time.perf_counter() - t0
        best = float('inf')
        for _ in range(batches):
            _hold_device(2 * reps * host)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                run()
            e1.record(s)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1) / reps)
        times[c] = best
    choice = min(candidates, key=lambda c: times[c])
    install(choice)
    _CACHE[key] = _CACHE[mark] = choice
    _to_file(key, choice)
    _to_file(mark, choice)
    return choice


def table():
    """Copy of the tuned choices (for reports)."""
    return dict(_CACHE)
