"""Direct normal-equation solves ('direct', linsolver.py:35-45) on the GPU:
the block-tridiagonal sparse path against the dense path and SciPy's
spsolve, the reference's own N = 64 Newton trajectories through it, and a
Newton step at the reference's N = 128 size (49 152 unknowns)."""

import json
import os

import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg
import torch

from oracle import operators as ops
from oracle import tape

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _jac(n, seed, dev, mu=0.01):
    from odil_amd.multigrid import BlockStencil
    from odil_amd.operators import Cavity
    rng = np.random.default_rng(seed)
    f = [torch.tensor(rng.uniform(-1, 1, (n, n)), device=dev)
         for _ in range(3)]
    jac, res = Cavity((n, n), mu).jacobian(f, want_residual=True)
    return BlockStencil.from_jacobian(jac), res


@pytest.mark.parametrize('beta,betadiag', [(0.3, 0.), (0.1, 0.5)])
def test_banded_direct_matches_dense_and_scipy(dev, beta, betadiag):
    from odil_amd import linsolver
    n = 20
    J, res = _jac(n, 3, dev)
    rhs = [-r for r in res]
    a = linsolver.default_args(beta=beta, betadiag=betadiag)
    a.direct_max_dense = 0
    xb = torch.stack(linsolver.solve(J, rhs, a, linsolver='direct'))
    a.direct_max_dense = 10**9
    xd = torch.stack(linsolver.solve(J, rhs, a, linsolver='direct'))
    m = J.tocsr()
    A = (m.T @ m).tocsr() + beta**2 * scipy.sparse.identity(m.shape[1])
    A = A + betadiag**2 * scipy.sparse.diags(A.diagonal())
    b = m.T @ np.concatenate([r.cpu().numpy().ravel() for r in rhs])
    xs = scipy.sparse.linalg.spsolve(A.tocsc(), b, permc_spec='MMD_ATA')
    for x in (xb, xd):
        got = x.cpu().numpy().ravel()
        assert np.abs(got - xs).max() <= 1e-10 * np.abs(xs).max()


def test_banded_direct_singular_cavity(dev):
    """beta = 0: the cavity's normal equations are singular (constant
    pressure); the banded and dense paths both return a solution of the
    normal equations (residual at round-off) that agree in u, v and in p up
    to its constant. J^T J has condition 2.5e4 outside its null space, so
    two FP64 solutions agree to ~kappa eps = 5e-12 (observed 2e-13)."""
    from odil_amd import linsolver
    n = 24
    J, res = _jac(n, 5, dev)
    rhs = [-r for r in res]
    a = linsolver.default_args()
    a.direct_max_dense = 0
    xb = [x.cpu().numpy() for x in linsolver.solve(J, rhs, a)]
    a.direct_max_dense = 10**9
    xd = [x.cpu().numpy() for x in linsolver.solve(J, rhs, a)]
    m = J.tocsr()
    b = m.T @ np.concatenate([r.cpu().numpy().ravel() for r in rhs])
    for x in (xb, xd):
        r = m.T @ (m @ np.concatenate([y.ravel() for y in x])) - b
        assert np.abs(r).max() <= 1e-12 * np.abs(b).max()
    for k in range(2):
        assert np.abs(xb[k] - xd[k]).max() <= 1e-10 * np.abs(xd[k]).max()
    dp = (xb[2] - xb[2].mean()) - (xd[2] - xd[2].mean())
    assert np.abs(dp).max() <= 1e-10 * np.abs(xd[2]).max()


@pytest.mark.parametrize('re_,epochs', [(100, 6), (400, 4)])
def test_newton_banded_reproduces_reference_trajectory(dev, re_, epochs):
    """The reference's N = 64 direct Newton runs (`cases/opt_cavity`,
    precomputed per-epoch max |du|, |dv|; pressure left out: its constant
    null mode depends on SuperLU's ordering) reproduced on the GPU with the
    block-tridiagonal direct path, to 1e-10 relative (observed at most
    5e-11, on the last Re = 400 step)."""
    from odil_amd import linsolver, odil
    from odil_amd.operators import Cavity
    with open(os.path.join(GOLD, 'cavity_newton.json')) as f:
        g = json.load(f)['Re{}'.format(re_)]
    n = 64
    domain = odil.Domain(ndim=2, shape=(n, n), varnames=['x', 'y'],
                         fieldnames=['u', 'v', 'p'])
    problem = odil.Problem(Cavity((n, n), 1.0 / re_), domain)
    state = odil.State()
    args = odil.util.default_args(epochs=epochs, linsolver='direct')
    args.direct_max_dense = 0
    prev = [None]
    got = []

    def callback(packed, epoch, dhistory=None, opt=None):
        if prev[0] is not None:
            d = packed - prev[0]
            got.append((epoch, d[:n * n].abs().max().item(),
                        d[n * n:2 * n * n].abs().max().item()))
        prev[0] = packed.clone()

    odil.util.optimize(args, 'newton', problem, state, callback)
    assert len(got) == epochs
    for epoch, du, dv in got:
        assert abs(du - g['du'][epoch]) <= 1e-10 * g['du'][epoch], (epoch,
                                                                     du)
        assert abs(dv - g['dv'][epoch]) <= 1e-10 * g['dv'][epoch], (epoch,
                                                                     dv)


def test_direct_newton_step_at_reference_size(dev):
    """One direct Newton step from the reference's N = 128, Re = 100 state
    (49 152 unknowns; the dense path could not hold it before): the update
    equals the oracle's SuperLU step in u, v (p up to its constant) and
    lowers the residual."""
    from odil_amd import linsolver
    from odil_amd.multigrid import BlockStencil
    from odil_amd.operators import Cavity
    from oracle import newton as onewton
    d = np.load(os.path.join(GOLD, 'cavity_N128_Re100.npz'))
    st = {k: d[k] for k in 'uvp'}
    n = 128
    op = Cavity((n, n), 0.01)
    f = [torch.tensor(st[k], device=dev) for k in 'uvp']
    jac, res = op.jacobian(f, want_residual=True)
    J = BlockStencil.from_jacobian(jac)
    dx = linsolver.solve(J, [-r for r in res], linsolver.default_args())
    const, m = tape.linearize(ops.make_cavity_operator(0.01),
                              ops.cavity_domain(n), st)
    c = np.hstack([e.ravel() for e in const])
    ref = onewton.solve_direct(m, -c).reshape(3, n, n)
    for k in range(2):
        got = dx[k].cpu().numpy()
        assert np.abs(got - ref[k]).max() <= 1e-6 * np.abs(ref[k]).max(), k
    new = [a + b for a, b in zip(f, dx)]
    _, s0 = op.residual(f, want_fields=False)
    _, s1 = op.residual(new, want_fields=False)
    assert s1.sum().item() < 0.5 * s0.sum().item()
