"""Normal-equation solves of the Newton step on the device
(`linsolver.solve`, linsolver.py:28-107).

`solve(J, rhs, linsolver=...)` solves (J^T J + beta^2 I +
betadiag^2 diag(J^T J)) dx = J^T rhs for a stencil Jacobian J
(`BlockStencil`) and residual-shaped rhs fields:

* 'direct': J^T J in stencil form (`odil_stencil_normal`), its dense FP64
  matrix (`odil_stencil_to_dense`) and a Cholesky factorisation on the GPU
  (rocSOLVER through torch.linalg) — the role of SuperLU with MMD_ATA in the
  reference, for systems up to `max_direct` unknowns. Singular normal
  equations (e.g. a pressure null mode) need beta > 0 here; the multigrid
  solver handles them without.
* 'multigrid': flexible CG on J^T J preconditioned by the GMG V-cycle
  (`odil_amd.multigrid.MultigridOp`), the solver of
  `odil_amd.newton.NewtonMultigrid`.
"""

from argparse import Namespace

import numpy as np
import torch

from odil_amd.multigrid import BlockStencil, normal_operator


def _arg(args, name, default):
    if args is None:
        return default
    v = getattr(args, name, None)
    return default if v is None else v


def solve(matr, rhs, args=None, dhistory=None, linsolver='direct',
          max_direct=40000):
    """Returns the correction fields dx (list of device tensors, FP64).

    matr: `BlockStencil` Jacobian (equations x fields), FP64.
    rhs: list of residual-shaped device tensors (one per equation), usually
         -f as in `optimize_newton` (util.py:261-267).
    args: optional namespace with `beta`, `betadiag`, `linsolver_maxiter`,
          `nlvl`, `omega` (linsolver.py:110-170)."""
    if not isinstance(matr, BlockStencil):
        raise ValueError('solve expects a BlockStencil Jacobian')
    beta = _arg(args, 'beta', 0.)
    betadiag = _arg(args, 'betadiag', 0.)
    b = matr.apply_t(rhs)
    if linsolver == 'direct':
        nvar = matr.nvar_in
        n = nvar * int(np.prod(matr.nw))
        if n > max_direct:
            raise ValueError('direct solve limited to {} unknowns, got {}; '
                             "use linsolver='multigrid'".format(max_direct,
                                                                n))
        N = normal_operator(matr, torch.float64, sym=True)
        M = N.to_dense()
        if beta:
            M.diagonal().add_(beta**2)
        if betadiag:  # diag of the beta-regularised matrix (linsolver.py:39)
            M.diagonal().mul_(1. + betadiag**2)
        L, info = torch.linalg.cholesky_ex(M)
        if int(info.item()) != 0:
            raise ValueError('normal equations not positive definite '
                             '(pass beta > 0 or use multigrid)')
        sol = torch.cholesky_solve(
            torch.cat([x.reshape(-1) for x in b]).unsqueeze(1), L)
        sol = sol.reshape(nvar, *matr.nw)
        if dhistory is not None:
            dhistory['linsolver_niter'] = 1
        return [sol[k].contiguous() for k in range(nvar)]
    if linsolver == 'multigrid':
        raise ValueError("linsolver='multigrid' runs through "
                         'odil_amd.newton.NewtonMultigrid (it keeps the '
                         'hierarchy and graphs across Newton steps)')
    raise ValueError('Unknown linsolver=' + str(linsolver))


def add_arguments(parser):
    """Solver options with the reference's names (linsolver.py:110-170)."""
    parser.add_argument('--linsolver', default='direct',
                        choices=['direct', 'multigrid'])
    parser.add_argument('--linsolver_maxiter', type=int, default=None)
    parser.add_argument('--beta', type=float, default=0.)
    parser.add_argument('--betadiag', type=float, default=0.)
    parser.add_argument('--nlvl', type=int, default=None)
    parser.add_argument('--omega', type=float, default=0.5)
    parser.add_argument('--smooth_pre', type=int, default=2)
    parser.add_argument('--smooth_post', type=int, default=2)
    parser.add_argument('--ndirect', type=int, default=3)
    return parser


def default_args(**kw):
    a = Namespace(linsolver='direct', linsolver_maxiter=None, beta=0.,
                  betadiag=0., nlvl=None, omega=0.5, smooth_pre=2,
                  smooth_post=2, ndirect=3)
    for k, v in kw.items():
        setattr(a, k, v)
    return a
