"""Optimisation drivers under the reference's names and signatures
(util.py:109-240, 242-372, 382-604): `optimize(args, optname, problem,
state, callback)` and the loops it dispatches to, each calling
`callback(packed, epoch, dhistory=..., opt=...)` before its updates with
`opt.last_loss` / `opt.last_residual` set, and updating `packed` (and
through `domain.assign_active_state` the state) in place.

Every evaluation is a kernel of the problem's operator: the stencil
Jacobian and residual for the Newton loops, the fused loss adjoint for the
gradient optimisers. `optimize_multigridop` runs the reference's loop
literally (split-shift Jacobian -> per-block `SparseOperator`s ->
right-hand side and block normal matrix -> `MultigridOp.update_A` ->
`linsolver_maxiter` V-cycles from zero with the Jacobi smoother), with the
reference's restriction and stencil coarsening, on the device. The faster
solver of this package for the same Newton step (flexible CG with a V-cycle
preconditioner, mixed precision) is `odil_amd.newton.NewtonMultigrid`
(`optimize_multigridop(..., krylov=True)` selects it).
"""

import argparse
from collections import defaultdict
from functools import partial

import numpy as np
import torch

from odil_amd import linsolver
from odil_amd import optimizer
from odil_amd.multigrid import Multigrid, MultigridOp, SparseOperator, \
    normal_operator


def add_arguments(parser):
    """The optimisation options of the reference's `add_arguments`
    (util.py:109-240; plotting and output options are not on this path)
    and the linear-solver options (linsolver.py:108-170)."""
    parser.add_argument('--epochs', type=int, default=None)
    parser.add_argument('--optimizer', type=str, default='newton')
    parser.add_argument('--seed', type=int, default=0)
    parser.add_argument('--double', type=int, default=0)
    parser.add_argument('--epoch_start', type=int, default=0)
    parser.add_argument('--checkpoint', type=str)
    parser.add_argument('--bfgs_m', type=int, default=50)
    parser.add_argument('--adam_epsilon', type=float)
    parser.add_argument('--adam_beta_1', type=float)
    parser.add_argument('--adam_beta_2', type=float)
    parser.add_argument('--multigrid', type=int, default=0)
    parser.add_argument('--mg_interp', type=str, default='conv',
                        choices=['roll', 'conv', 'manual'])
    parser.add_argument('--nn_initializer', type=str, default='legacy',
                        choices=['legacy', 'glorot', 'lecun', 'he'])
    parser.add_argument('--lr', type=float, default=1e-3)
    linsolver.add_arguments(parser)
    return parser


def default_args(**kw):
    """Namespace with the defaults of `add_arguments`."""
    a = add_arguments(argparse.ArgumentParser()).parse_args([])
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def _rms(const):
    return [float(torch.sqrt(torch.mean(c.double()**2))) for c in const]


def _weights_tail(problem, packed):
    n = problem.domain.fields_size()
    return packed.numel() - n


def optimize_newton(args, problem, state, callback):
    """util.py:242-268: linearize (stencil Jacobian and residual kernels),
    solve the normal equations with `args.linsolver` (every choice on the
    device, `linsolver.solve`), `packed += dpacked`."""
    opt = optimizer.Optimizer(name='newton', displayname='Newton')
    packed = problem.pack_state(state)
    dhistory = defaultdict(list)
    for epoch in range(args.epoch_start, args.epochs + 1):
        s = problem.unpack_state(packed)
        problem.domain.assign_active_state(state, s)
        J, const = problem.jacobian(state, epoch)
        rms = _rms(const)
        opt.last_loss = sum(r * r for r in rms)
        opt.last_residual = rms
        callback(packed, epoch, dhistory=dhistory, opt=opt)
        if epoch == args.epochs:
            break
        opt.evals += 1
        dx = linsolver.solve(J, [-c for c in const], args, dhistory,
                             args.linsolver)
        d = torch.cat([x.reshape(-1) for x in dx])
        packed[:d.numel()] += d
    return packed


def optimize_opt(args, optname, problem, state, callback, **kwargs):
    """util.py:271-317: a gradient optimiser over the packed state with the
    fused loss adjoint (field and weight gradients concatenated)."""

    def loss_grad(packed, epoch):
        s = problem.unpack_state(packed)
        problem.domain.assign_active_state(state, s)
        loss, grads, wgrads, loss_split = problem.eval_loss_grad(state,
                                                                 epoch=epoch)
        last_residual = [v**0.5 if v >= 0 else v for v in loss_split]
        g = torch.cat([problem.pack_fields(grads),
                       problem.pack_weights(wgrads)])
        return loss, g, last_residual

    packed = problem.pack_state(state)
    if getattr(args, 'bfgs_m', None) is not None:
        kwargs['m'] = args.bfgs_m
    if getattr(args, 'adam_epsilon', None) is not None:
        kwargs['epsilon'] = args.adam_epsilon
    if getattr(args, 'adam_beta_1', None) is not None:
        kwargs['beta_1'] = args.adam_beta_1
    if getattr(args, 'adam_beta_2', None) is not None:
        kwargs['beta_2'] = args.adam_beta_2
    opt = optimizer.make_optimizer(optname, dtype=problem.domain.dtype,
                                   **kwargs)
    loss, _, residual = loss_grad(packed, args.epoch_start)
    opt.last_loss = float(loss)
    opt.last_residual = [float(r) for r in residual]
    callback(packed, epoch=args.epoch_start, opt=opt)
    packed, info = opt.run(packed, loss_grad=loss_grad,
                           epochs=args.epochs - args.epoch_start,
                           callback=callback, epoch_start=args.epoch_start,
                           lr=args.lr, **kwargs)
    problem.domain.assign_active_state(state, problem.unpack_state(packed))
    return packed, info


def optimize_multigrid(args, problem, state, callback):
    """util.py:382-482 (`optimize_multigrid_base`): stored normal matrix
    J^T J, the sparse-matrix `Multigrid` hierarchy, `linsolver_maxiter`
    V-cycles from zero with the Jacobi smoother."""
    opt = optimizer.Optimizer(name='multigrid', displayname='Multigrid')
    packed = problem.pack_state(state)
    dhistory = defaultdict(list)
    domain = problem.domain
    nvar = len(domain.fieldnames)
    mg = Multigrid(domain.shape, nvar=nvar, nlevels=args.nlvl,
                   restriction=args.restriction, dtype=domain.dtype,
                   device=domain.device)
    for epoch in range(args.epoch_start, args.epochs + 1):
        s = problem.unpack_state(packed)
        domain.assign_active_state(state, s)
        J, const = problem.jacobian(state, epoch)
        opt.last_residual = _rms(const)
        opt.last_loss = sum(r * r for r in opt.last_residual)
        callback(packed, epoch, dhistory=dhistory, opt=opt)
        if epoch == args.epochs:
            break
        opt.evals += 1
        rhs = torch.stack(J.apply_t([-c for c in const])).reshape(-1)
        mg.update_A(normal_operator(J, torch.float64))
        sol = torch.zeros_like(rhs)
        smoother = partial(mg.smoother_jacobi, omega=args.omega)
        for _ in range(args.linsolver_maxiter or 50):
            sol = mg.step(sol, rhs, smoother, ndirect=args.ndirect,
                          pre=args.smooth_pre, post=args.smooth_post)
        packed[:sol.numel()] += sol
    return packed


def optimize_multigridop(args, problem, state, callback, krylov=False,
                         precision='fp64'):
    """util.py:509-604 (`optimize_multigridop_base`), the same sequence of
    calls on device arrays. krylov=True runs `NewtonMultigrid` instead
    (flexible CG preconditioned by one V-cycle, `precision` 'fp64' or
    'mixed')."""
    domain = problem.domain
    nvar = len(domain.fieldnames)
    dhistory = defaultdict(list)
    opt = optimizer.Optimizer(name='multigridop', displayname='MultigridOp')
    packed = problem.pack_state(state)
    if krylov:
        return _multigrid_krylov(args, problem, state, callback, opt,
                                 packed, dhistory, precision)
    mg = MultigridOp(domain.shape, nvar=nvar, restriction=args.restriction,
                     mod=None, dtype=domain.dtype, nlevels=args.nlvl,
                     device=domain.device)
    nw = domain.shape
    for epoch in range(args.epoch_start, args.epochs + 1):
        s = problem.unpack_state(packed)
        domain.assign_active_state(state, s)
        callback(packed, epoch, dhistory=dhistory, opt=opt)
        if epoch == args.epochs:
            break
        const, grads, field_desc, wgrads = problem._eval_grad(
            state.fields, state.weights, epoch)
        opt.last_residual = _rms(const)
        neqn = len(const)
        stf = [[dict() for _ in range(nvar)] for _ in range(neqn)]
        for i in range(neqn):
            for j, (varindex, dw) in enumerate(field_desc):
                g = grads[i][j]
                if g is None or not bool((g != 0).any()):
                    continue
                stf[i][varindex][tuple(np.array(dw))] = g
        mm = [[SparseOperator(st, nw, dtype=domain.dtype) for st in row]
              for row in stf]
        opt.evals += 1
        rhs = [-sum(mm[i][j].mul_transpose_field(const[i])
                    for i in range(neqn)) for j in range(nvar)]
        matr = [[None] * nvar for _ in range(nvar)]
        for i in range(nvar):
            for j in range(nvar):
                matr[i][j] = mm[0][i].mul_transpose_op(mm[0][j])
                for e in range(1, neqn):
                    matr[i][j] = matr[i][j].add_elementwise(
                        mm[e][i].mul_transpose_op(mm[e][j]))
        mg.update_A(matr)
        sol = [torch.zeros_like(r) for r in rhs]
        for _ in range(args.linsolver_maxiter or 50):
            sol = mg.step(sol, rhs, pre=args.smooth_pre,
                          post=args.smooth_post, ndirect=args.ndirect,
                          smoother=partial(mg.smoother_jacobi,
                                           omega=args.omega))
        dpacked = torch.stack([x.reshape(-1) for x in sol]).reshape(-1)
        packed[:dpacked.numel()] += dpacked
    return packed


def _multigrid_krylov(args, problem, state, callback, opt, packed, dhistory,
                      precision):
    from odil_amd.newton import NewtonMultigrid
    domain = problem.domain
    nm = NewtonMultigrid(problem.operator, nlvl=args.nlvl, omega=args.omega,
                         smooth_pre=args.smooth_pre,
                         smooth_post=args.smooth_post, ndirect=args.ndirect,
                         maxiter=args.linsolver_maxiter or 10,
                         precision=precision, setup_every=1)
    for epoch in range(args.epoch_start, args.epochs + 1):
        s = problem.unpack_state(packed)
        domain.assign_active_state(state, s)
        fields = [state.fields[k] for k in domain.fieldnames]
        if epoch == args.epochs:
            opt.last_residual = nm.residual_rms(fields)
            callback(packed, epoch, dhistory=dhistory, opt=opt)
            break
        # The residual of this state (before the update), as the reference
        # reports it; the callback then sees it.
        opt.last_residual = nm.residual_rms(fields)
        callback(packed, epoch, dhistory=dhistory, opt=opt)
        nm.step(fields)
        opt.evals += 1
        packed.copy_(problem.pack_state(state))
    return packed


def optimize(args, optname, problem, state, callback, **kwargs):
    """util.py:365-370."""
    if optname == 'newton':
        return optimize_newton(args, problem, state, callback)
    if optname == 'multigrid':
        return optimize_multigrid(args, problem, state, callback)
    if optname == 'multigridop':
        return optimize_multigridop(args, problem, state, callback)
    if optname == 'jacobi':
        raise ValueError("optimizer 'jacobi' (a sampled-diagonal method) is "
                         'not provided')
    return optimize_opt(args, optname, problem, state, callback, **kwargs)
