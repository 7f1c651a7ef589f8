"""Normal-equation solvers of `odil_amd.linsolver` on the GPU against the
routines the reference calls (linsolver.py:28-107): SciPy's `lsqr` and
`bicgstab` on the CSR copy of the same Jacobian (SciPy implements the
published LSQR and BiCGSTAB algorithms the GPU code restates), and the
dense direct solve for the converged answers."""

import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg
import torch

pytestmark = pytest.mark.gpu


def _cavity_jac(dev, n0, n1, seed):
    from odil_amd.multigrid import BlockStencil
    from odil_amd.operators import Cavity
    rng = np.random.default_rng(seed)
    f = [torch.tensor(rng.uniform(-0.5, 0.5, (n0, n1)), device=dev)
         for _ in range(3)]
    jac, res = Cavity((n0, n1), 0.01).jacobian(f, want_residual=True)
    return BlockStencil.from_jacobian(jac), [-r for r in res]


def _flat(fields):
    return np.concatenate([f.cpu().numpy().ravel() for f in fields])


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize('damp,iters', [(0., 8), (0.3, 8), (0.3, 16)])
def test_lsqr_iterates_match_scipy(dev, damp, iters):
    """Same iterate, iteration count and norm estimates as SciPy's lsqr
    after a fixed number of iterations (istop 7). Beyond ~20 iterations on
    this 297-unknown system the bidiagonalisation loses orthogonality and
    round-off differences of the two dot-product orders grow to 1e-3 in
    both directions, so the comparison stops at 16."""
    from odil_amd.linsolver import lsqr
    J, rhs = _cavity_jac(dev, 9, 11, 0)
    b = torch.stack(rhs)
    got = lsqr(J, b, damp=damp, atol=0., btol=0., conlim=0., iter_lim=iters)
    ref = scipy.sparse.linalg.lsqr(J.tocsr(), _flat(rhs), damp=damp, atol=0.,
                                   btol=0., conlim=0., iter_lim=iters)
    assert got[1] == ref[1] == 7
    assert got[2] == ref[2] == iters
    assert _rel(got[0].cpu().numpy().ravel(), ref[0]) <= 1e-9
    for k in (3, 4, 5, 6, 7, 8):  # r1norm r2norm anorm acond arnorm xnorm
        assert abs(got[k] - ref[k]) <= 1e-9 * abs(ref[k]), (k, got[k], ref[k])


def test_lsqr_stops_like_scipy(dev):
    """To a tight tolerance both stop on the same rule, at the same
    answer, after nearly the same number of iterations (several hundred:
    past the loss of orthogonality the two round-off histories differ, so
    the counts may differ by a few)."""
    from odil_amd.linsolver import lsqr
    J, rhs = _cavity_jac(dev, 9, 9, 1)
    got = lsqr(J, torch.stack(rhs), damp=0.5, atol=1e-10, btol=1e-10,
               iter_lim=2000)
    ref = scipy.sparse.linalg.lsqr(J.tocsr(), _flat(rhs), damp=0.5,
                                   atol=1e-10, btol=1e-10, iter_lim=2000)
    assert got[1] == ref[1] and got[1] in (1, 2)
    assert abs(got[2] - ref[2]) <= max(2, 0.02 * ref[2])
    assert _rel(got[0].cpu().numpy().ravel(), ref[0]) <= 1e-8


def test_bicgstab_iterates_match_scipy(dev):
    """The first BiCGSTAB iterates on A = J^T J + beta^2 I equal SciPy's
    (rtol 0 and atol as the reference passes)."""
    from odil_amd.linsolver import _regularised_normal, _stack, bicgstab
    J, rhs = _cavity_jac(dev, 9, 11, 2)
    beta = 0.3
    A = _regularised_normal(J, beta, 0.)
    b = _stack(J.apply_t(rhs))
    mine = []
    bicgstab(A, b, atol=0., maxiter=6,
             callback=lambda x: mine.append(x.cpu().numpy().ravel().copy()))
    Jc = J.tocsr()
    Ac = (Jc.T @ Jc + beta**2 * scipy.sparse.eye(Jc.shape[1])).tocsr()
    theirs = []
    scipy.sparse.linalg.bicgstab(Ac, Jc.T @ _flat(rhs), rtol=0., atol=0.,
                                 maxiter=6,
                                 callback=lambda x: theirs.append(x.copy()))
    assert len(mine) == len(theirs) == 6
    for a, c in zip(mine, theirs):
        assert _rel(a, c) <= 1e-8


def _direct(J, rhs, beta, betadiag=0.):
    Jc = J.tocsr()
    A = (Jc.T @ Jc + beta**2 * scipy.sparse.eye(Jc.shape[1])).tocsr()
    if betadiag:
        A = (A + betadiag**2 * scipy.sparse.diags(A.diagonal())).tocsr()
    return scipy.sparse.linalg.spsolve(A.tocsc(), Jc.T @ _flat(rhs))


@pytest.mark.parametrize('linsolver', ['direct', 'direct_cu', 'lsqr',
                                       'lsqr_cu', 'bicgstab', 'multigrid'])
def test_solve_choices_reach_the_direct_answer(dev, linsolver):
    """Every choice of `solve` returns the regularised normal-equation
    solution of SciPy's direct solve (the reference's default), within
    what its tolerance allows, and fills the reference's history keys."""
    from odil_amd.linsolver import default_args, solve
    J, rhs = _cavity_jac(dev, 17, 17, 3)
    args = default_args(beta=0.5, linsolver_tol=1e-12,
                        linsolver_maxiter=4000, nlvl=3)
    dh = {}
    dx = solve(J, rhs, args, dh, linsolver)
    ref = _direct(J, rhs, 0.5)
    assert _rel(_flat(dx), ref) <= 1e-8, linsolver
    assert dh['linsolver_niter'] >= 1
    if linsolver not in ('direct', 'direct_cu'):
        assert dh['linsolver_residual'] >= 0.
    if linsolver.startswith('lsqr'):
        assert dh['linsolver_anorm'] > 0 and dh['linsolver_acond'] > 1


def test_solve_betadiag(dev):
    from odil_amd.linsolver import default_args, solve
    J, rhs = _cavity_jac(dev, 9, 9, 4)
    args = default_args(beta=0.2, betadiag=0.7, linsolver_tol=1e-13,
                        linsolver_maxiter=3000)
    ref = _direct(J, rhs, 0.2, 0.7)
    for choice in ('direct', 'bicgstab', 'multigrid'):
        dx = solve(J, rhs, args, None, choice)
        assert _rel(_flat(dx), ref) <= 1e-8, choice


def test_multigrid_pcg_beats_plain_iterations(dev):
    """The V-cycle preconditioner: CG reaches 1e-10 relative residual on
    the 33x33 cavity normal equations (beta = 0.05) in under 60 iterations;
    unpreconditioned BiCGSTAB has not converged after as many."""
    from odil_amd.linsolver import (_regularised_normal, _stack, _Vec,
                                    bicgstab, pcg_multigrid)
    J, rhs = _cavity_jac(dev, 33, 33, 5)
    A = _regularised_normal(J, 0.05, 0.)
    b = _stack(J.apply_t(rhs))
    x, res = pcg_multigrid(A, b, tol=1e-10, maxiter=200)
    assert res[-1] <= 1e-10 * res[0]
    assert len(res) - 1 < 60
    vec = _Vec(b.device)
    r = torch.empty_like(b)
    A.residual(list(x), list(b), y=list(r))
    assert vec.norm(r) <= 2e-10 * vec.norm(b)
    _, info = bicgstab(A, b, atol=1e-10 * res[0], maxiter=len(res) - 1)
    assert info != 0


def test_newton_with_iterative_linsolver(dev):
    """optimize_newton with args.linsolver = 'lsqr' reaches the direct
    Newton answer of the 17x17 Gaussian wave case."""
    from odil_amd.drivers import optimize_newton
    from odil_amd.linsolver import default_args
    from odil_amd.operators import Wave
    from odil_amd.problem import Domain, Problem, State
    from oracle import operators as ops
    _, _, bnd = ops.wave_gaussian_setup(17, 17, seed=0)
    dom = Domain(shape=(17, 17), lower=(0, -1), upper=(1, 1),
                 varnames=('t', 'x'), fieldnames=['u'], device=dev)
    pb = Problem(Wave((17, 17), *bnd, device=dev), dom)
    st_d, st_l = State(), State()
    optimize_newton(pb, st_d, 2)
    optimize_newton(pb, st_l, 2, default_args(linsolver='lsqr',
                                              linsolver_tol=1e-14,
                                              linsolver_maxiter=5000))
    a = st_l.fields['u'].cpu().numpy()
    b = st_d.fields['u'].cpu().numpy()
    assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()


def test_solve_rejects(dev):
    from odil_amd.linsolver import default_args, solve
    J, rhs = _cavity_jac(dev, 8, 8, 6)
    with pytest.raises(ValueError):
        solve(J, rhs, default_args(), None, 'sparseqr')
    with pytest.raises(ValueError):
        solve(J, rhs, default_args(beta=0.1), None, 'multigrid')  # even
    with pytest.raises(ValueError):
        solve(J, rhs, default_args(), None, 'gmres')
