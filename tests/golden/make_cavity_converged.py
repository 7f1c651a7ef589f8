#!/usr/bin/env python3
"""Converged lid-driven cavity states from the oracle's direct Newton.

The oracle's `optimize_newton` restates the reference's Newton driver with
the direct normal-equation solve (`util.py:242-268`, `linsolver.py:35-45`;
SuperLU with MMD_ATA ordering). Starting from the zero state it runs until
the max per-equation RMS residual is <= 1e-13 (or stops improving) and
stores (u, v, p) with the residual history, so GPU tests can compare a
Newton-multigrid solve at Re = 1000 (mu = 1/Re, `ns_steady.py:867`) with
the reference algorithm without re-running ~140 direct steps per test.

Usage: python tests/golden/make_cavity_converged.py N RE  (writes
cavity_converged_N{N}_Re{RE}.npz next to this script). Only the oracle is
used; no reference code is imported.
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import newton, tape  # noqa: E402
from oracle import operators as ops  # noqa: E402


def main(n, re, tol=1e-13, max_epochs=400):
    dom = ops.cavity_domain(n)
    operator = ops.make_cavity_operator(1.0 / re)
    packed = tape.pack_fields(dom, {k: np.zeros((n, n)) for k in 'uvp'})
    hist = []
    t0 = time.time()
    best = None
    for epoch in range(max_epochs + 1):
        state = tape.unpack_fields(dom, packed)
        const, m = tape.linearize(operator, dom, state, epoch=epoch)
        rms = newton.rms_residuals(const)
        hist.append(rms)
        print(epoch, ['%.3e' % x for x in rms], '%.0fs' % (time.time() - t0),
              flush=True)
        if best is None or max(rms) < best[0]:
            best = (max(rms), packed.copy(), epoch)
        if max(rms) <= tol or epoch == max_epochs or (
                epoch > 20 and max(rms) > 0.9 * best[0] and
                epoch - best[2] > 10):
            break
        c = np.hstack([eq.flatten() for eq in const])
        packed = packed + newton.solve_direct(m, -c)
    state = tape.unpack_fields(dom, best[1])
    out = os.path.join(HERE, 'cavity_converged_N{}_Re{:g}.npz'.format(n, re))
    np.savez_compressed(out, u=state['u'], v=state['v'], p=state['p'],
                        rms=np.array(hist), epoch=best[2], re=re)
    print('wrote', out, 'epoch', best[2], 'rms', best[0])


if __name__ == '__main__':
    main(int(sys.argv[1]), float(sys.argv[2]))
