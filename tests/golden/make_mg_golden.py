#!/usr/bin/env python3
"""Golden V-cycle iterates from the reference's own `MultigridOp`.

The recursion defaults of the reference V-cycle (the recursive call does not
forward `pre`, `post`, `ndirect`, multigrid.py:589-593) and its smoothing
order are not pinned by any closed form, so this script imports the
reference's `multigrid.py` (NumPy/SciPy only) in the build container and
records its outputs. Only the resulting arrays are committed
(`mg_vcycle.npz`); the reference itself never travels.

Cases:
  poisson17: the Dirichlet-row Laplacian of `test_multigrid_poisson`
             (test_multigrid.py:251-296) on 17x17 nodes, rhs = 10,
             3 V-cycles with omega = 0.7, default pre/post/ndirect.
  cavity17:  J^T J of the lid-driven cavity Jacobian (Re=100) at a seeded
             random state on 17x17 cells, built by the oracle's tape, then
             3 V-cycles with omega = 0.6, pre = post = 2, ndirect = 3 and
             4 levels, as `util.optimize_multigridop_base` calls it
             (util.py:577-595).
"""

import os
import sys
from functools import partial

import numpy as np

# Importing the reference must not write bytecode next to its sources.
sys.dont_write_bytecode = True

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, '/root/reference/code')

import multigrid as refmg  # noqa: E402  (reference, NumPy/SciPy only)
import scipy.sparse as sp  # noqa: E402

from oracle import multigrid as omg  # noqa: E402
from oracle import operators as ops  # noqa: E402
from oracle import tape  # noqa: E402


def poisson17():
    mod = refmg.ModNumpy(np, sp)
    dtype = np.float64
    nw = np.array((17, 17))
    dx = 1 / (nw[0] - 1)
    stf = {(0, 0): 4 * np.ones(nw) / dx**2}
    pad = [[1, 1]] * 2
    for dw in [(1, 0), (-1, 0), (0, 1), (0, -1)]:
        stf[dw] = -np.pad(np.ones(nw - 2), pad) / dx**2
    A = refmg.SparseOperator(stf, nw, mod=mod, dtype=dtype)
    mg = refmg.MultigridOp(nw, mod=mod, dtype=dtype)
    mg.update_A([[A]])
    rhs = np.ones(nw).reshape(-1) * 10
    u = np.zeros_like(rhs)
    out = {'poisson17_rhs': rhs}
    for k in stf:
        out['poisson17_c_{}_{}'.format(*k)] = stf[k]
    for it in range(3):
        u = mg.step([u], [rhs],
                    smoother=partial(mg.smoother_jacobi, omega=0.7))[0]
        out['poisson17_u{}'.format(it + 1)] = np.array(u)
    out['poisson17_nlevels'] = np.array(mg.nlevels)
    return out


def cavity17():
    mod = refmg.ModNumpy(np, sp)
    n = 17
    rng = np.random.default_rng(7)
    st = {k: rng.uniform(-0.5, 0.5, (n, n)) for k in 'uvp'}
    const, grads, desc = tape.eval_grad(ops.make_cavity_operator(0.01),
                                        ops.cavity_domain(n), st)
    stf = omg.coeffs_from_tape(grads, desc, 3, 3)
    mm = [[refmg.SparseOperator(s, (n, n), mod=mod, dtype=np.float64)
           for s in row] for row in stf]
    rhs = [-sum(mm[i][j].mul_transpose_field(const[i]) for i in range(3))
           for j in range(3)]
    matr = [[None] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            matr[i][j] = mm[0][i].mul_transpose_op(mm[0][j])
            for e in range(1, 3):
                matr[i][j] = matr[i][j].add_elementwise(
                    mm[e][i].mul_transpose_op(mm[e][j]))
    mg = refmg.MultigridOp((n, n), nvar=3, mod=mod, dtype=np.float64,
                           nlevels=4)
    mg.update_A(matr)
    sol = [np.zeros_like(r) for r in rhs]
    out = {'cavity17_state_' + k: st[k] for k in 'uvp'}
    out['cavity17_rhs'] = np.array(rhs)
    for it in range(3):
        sol = mg.step(sol, rhs, pre=2, post=2, ndirect=3,
                      smoother=partial(mg.smoother_jacobi, omega=0.6))
        out['cavity17_u{}'.format(it + 1)] = np.array(sol)
    out['cavity17_nlevels'] = np.array(mg.nlevels)
    return out


if __name__ == '__main__':
    res = {}
    res.update(poisson17())
    res.update(cavity17())
    np.savez_compressed(os.path.join(HERE, 'mg_vcycle.npz'), **res)
    print('wrote', sorted(res))
