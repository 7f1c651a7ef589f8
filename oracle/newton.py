"""CPU oracle: normal-equation linear solve and the Newton driver.

TEST INFRASTRUCTURE ONLY (see `oracle/tape.py`).
"""

import numpy as np
import scipy.sparse
import scipy.sparse.linalg

from oracle import tape


def solve_direct(matr, rhs, beta=0., betadiag=0.):
    """`linsolver.solve(..., linsolver='direct')` (`linsolver.py:28-45`):
    A = J^T J (+beta^2 I, +betadiag^2 diag(A)), b = J^T rhs, SuperLU with
    MMD_ATA column ordering."""
    matr_reg = matr.T.dot(matr).tocsr()
    if beta:
        matr_reg += beta**2 * scipy.sparse.identity(matr.shape[1],
                                                    format='csr')
    if betadiag:
        matr_reg += betadiag**2 * scipy.sparse.diags(matr_reg.diagonal())
    rhs_reg = matr.T.dot(rhs)
    return scipy.sparse.linalg.spsolve(matr_reg, rhs_reg,
                                       permc_spec='MMD_ATA')


def optimize_newton(operator, domain, fields, epochs, callback=None,
                    solve=solve_direct):
    """`util.optimize_newton` (`util.py:242-268`): linearize, solve the normal
    equations for the correction, `packed += dpacked`, no line search.

    `callback(epoch, packed, const)` runs before each update like the
    reference's callback. Returns the final packed state.
    """
    packed = tape.pack_fields(domain, fields)
    for epoch in range(epochs + 1):
        state = tape.unpack_fields(domain, packed)
        const, m = tape.linearize(operator, domain, state, epoch=epoch)
        if callback is not None:
            callback(epoch, packed, const)
        if epoch == epochs:
            break
        c = np.hstack([eq.flatten() for eq in const])
        packed = packed + solve(m, -c)
    return packed


def optimize_multigrid(operator, domain, fields, epochs, nlevels=None,
                       maxiter=50, omega=0.6, ndirect=3, pre=2, post=2,
                       restriction='full', callback=None):
    """`util.optimize_multigrid_base` (`util.py:382-471`): linearize,
    rhs = -m^T c, matr = m^T m, `Multigrid.update_A(matr)`, `maxiter`
    V-cycles from zero with the point-Jacobi smoother, `packed += sol`.
    Returns the final packed state."""
    from oracle.multigrid import Multigrid
    packed = tape.pack_fields(domain, fields)
    mg = Multigrid(domain.shape, nvar=len(domain.fieldnames),
                   nlevels=nlevels, restriction=restriction)
    for epoch in range(epochs + 1):
        state = tape.unpack_fields(domain, packed)
        const, m = tape.linearize(operator, domain, state, epoch=epoch)
        if callback is not None:
            callback(epoch, packed, const)
        if epoch == epochs:
            break
        c = np.hstack([eq.flatten() for eq in const])
        rhs = -m.T @ c
        mg.update_A((m.T @ m).tocsr())
        sol = np.zeros_like(rhs)
        smooth = lambda level, u, f: mg.smoother_jacobi(level, u, f, omega)
        for _ in range(maxiter):
            sol = mg.step(sol, rhs, smooth, ndirect=ndirect, pre=pre,
                          post=post)
        packed = packed + sol
    return packed


def rms_residuals(const):
    """Per-equation RMS residual, `opt.last_residual` (`util.py:253-254`)."""
    return [float(np.mean(c**2)**0.5) for c in const]
