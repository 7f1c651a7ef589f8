#!/usr/bin/env python3
"""Extract golden vectors from the reference's shipped data files.

Run in a container that has the reference checkout at /root/reference; the
outputs (JSON / npz data only) are committed next to this script.

* `wave_fd_newton.json`: RMS error against the exact cosine solution after 2
  Newton steps of the 1D wave FD case for 36 grid sizes N=Nt=Nx, from
  `code/cases/wave_cost/precomputed/fd_newton/lines.pickle`. The pickle is
  NOT unpickled: its opcodes are walked with `pickletools.genops` and the raw
  little-endian payloads of the `N` (int64) and `error` (float64) arrays read.
* `cavity_newton.json`: per-epoch max|du|, max|dv|, max|dp| of the N=64
  lid-driven cavity Newton runs at Re=100 and Re=400, from
  `code/cases/opt_cavity/precomputed/data/cavity_N64_Re*_newton.csv`.
* `cavity_N128_Re*.npz`: converged (u, v, p) states of the N=128 cavity from
  `code/cases/cavity_stream/cavity_N128_Re*.pickle`, again read by opcode walk
  (float64 arrays of shape (128, 128) stored as raw bytes).
"""

import csv
import json
import os
import pickletools

import numpy as np

REF = '/root/reference/code/cases'
HERE = os.path.dirname(os.path.abspath(__file__))


def walk_arrays(path, keys):
    """Map each key string in `keys` to the raw bytes of the first BINBYTES
    payload that follows it in the opcode stream."""
    data = open(path, 'rb').read()
    out = {}
    pending = None
    for op, arg, _ in pickletools.genops(data):
        if op.name in ('SHORT_BINUNICODE', 'BINUNICODE') and arg in keys:
            pending = arg
        elif op.name in ('BINBYTES', 'SHORT_BINBYTES',
                         'BINBYTES8') and pending is not None:
            if len(arg) > 16:
                out[pending] = arg
                pending = None
    return out


def wave():
    path = os.path.join(REF, 'wave_cost/precomputed/fd_newton/lines.pickle')
    raw = walk_arrays(path, ('N', 'error'))
    n = np.frombuffer(raw['N'], '<i8')
    err = np.frombuffer(raw['error'], '<f8')
    nx = [int(round(np.sqrt(v))) for v in n]
    assert all(x * x == v for x, v in zip(nx, n))
    res = {'source': 'code/cases/wave_cost/precomputed/fd_newton/lines.pickle',
           'nx': nx, 'error': [float(e) for e in err]}
    with open(os.path.join(HERE, 'wave_fd_newton.json'), 'w') as f:
        json.dump(res, f, indent=1)


def cavity_csv():
    res = {}
    for re_ in (100, 400):
        name = 'cavity_N64_Re{}_newton.csv'.format(re_)
        path = os.path.join(REF, 'opt_cavity/precomputed/data', name)
        rows = list(csv.DictReader(open(path)))
        res['Re{}'.format(re_)] = {
            'source': 'code/cases/opt_cavity/precomputed/data/' + name,
            'epoch': [int(r['epoch']) for r in rows],
            'du': [float(r['du']) for r in rows],
            'dv': [float(r['dv']) for r in rows],
            'dp': [float(r['dp']) for r in rows],
            'tt_opt': [float(r['tt_opt']) for r in rows],
        }
    with open(os.path.join(HERE, 'cavity_newton.json'), 'w') as f:
        json.dump(res, f, indent=1)


def cavity_states():
    for re_ in (100, 1000):
        name = 'cavity_N128_Re{}.pickle'.format(re_)
        raw = walk_arrays(os.path.join(REF, 'cavity_stream', name),
                          ('u', 'v', 'p'))
        arrs = {
            k: np.frombuffer(raw[k], '<f8').reshape(128, 128)
            for k in ('u', 'v', 'p')
        }
        np.savez_compressed(
            os.path.join(HERE, 'cavity_N128_Re{}.npz'.format(re_)), **arrs)


if __name__ == '__main__':
    wave()
    cavity_csv()
    cavity_states()
