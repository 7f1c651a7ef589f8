"""Optimisation drivers with the reference's structure (util.py:240-317,
382-604): Newton with a direct normal-equation solve, Newton with the
geometric-multigrid solvers (plain V-cycles on the stored normal matrix,
and flexible CG with a V-cycle preconditioner), and gradient optimisers
over the packed state.

Every evaluation runs on the device: residual and Jacobian kernels for
Newton, the fused loss adjoint for the gradient optimisers.
"""

import time

import numpy as np
import torch

from odil_amd import linsolver as ls
from odil_amd import optimizer as opt_mod
from odil_amd.newton import NewtonMultigrid


def _rms(problem, state):
    """Per-equation RMS residual (util.py:253-254) from the residual
    kernel's sums of squares; one host sync."""
    _, ssq = problem.operator.residual(problem._fields(state),
                                       want_fields=False)
    return [float(np.sqrt(s / problem.operator.ncell))
            for s in ssq.tolist()]


def optimize_newton(problem, state, epochs, args=None, callback=None,
                    tol=None):
    """util.optimize_newton (util.py:240-268): at each epoch linearize
    (residual and stencil Jacobian on the device), solve the normal
    equations for the correction (`linsolver.solve` with `args.linsolver`,
    default 'direct', every choice on the GPU) and
    update the fields in place (`packed += dpacked`). `callback(epoch,
    state, rms)` runs before each update. Stops early when
    max(rms) <= tol. Returns the history [(epoch, rms, seconds)]."""
    problem.init_missing(state)
    names = problem.domain.fieldnames
    hist = []
    t0 = time.perf_counter()
    for epoch in range(epochs + 1):
        J, res = problem.jacobian(state, epoch)
        rms = _rms(problem, state)
        hist.append((epoch, rms, time.perf_counter() - t0))
        if callback is not None:
            callback(epoch, state, rms)
        if epoch == epochs or (tol is not None and max(rms) <= tol):
            break
        dx = ls.solve(J, [-r for r in res], args,
                      linsolver=getattr(args, 'linsolver', None) or 'direct')
        for k, d in zip(names, dx):
            state.fields[k] += d
    return hist


def optimize_multigrid(problem, state, epochs, args=None, callback=None,
                       tol=None):
    """util.optimize_multigrid -> optimize_multigrid_base (util.py:382-471)
    on the device: per epoch the stencil Jacobian J and residual f,
    rhs = -J^T f, the stored normal matrix J^T J (`normal_operator`, stencil
    form instead of the reference's host CSR product), `Multigrid.update_A`
    (Galerkin R A T levels) and `linsolver_maxiter` plain V-cycles from zero
    with the point-Jacobi smoother (omega, smooth_pre, smooth_post, ndirect,
    restriction and nlvl from `args`), then `packed += sol`.
    `callback(epoch, state, rms)` runs before each update; stops early when
    max(rms) <= tol. Returns the history [(epoch, rms, seconds)]."""
    from odil_amd.multigrid import Multigrid, normal_operator
    problem.init_missing(state)
    names = problem.domain.fieldnames
    a = args or ls.default_args()
    mg = None
    hist = []
    t0 = time.perf_counter()
    for epoch in range(epochs + 1):
        J, res = problem.jacobian(state, epoch)
        rms = _rms(problem, state)
        hist.append((epoch, rms, time.perf_counter() - t0))
        if callback is not None:
            callback(epoch, state, rms)
        if epoch == epochs or (tol is not None and max(rms) <= tol):
            break
        rhs = torch.stack(J.apply_t([-r for r in res])).reshape(-1)
        N = normal_operator(J, torch.float64)
        if mg is None:
            mg = Multigrid(problem.domain.shape, nvar=len(names),
                           nlevels=getattr(a, 'nlvl', None),
                           restriction=getattr(a, 'restriction', None) or
                           'full', device=rhs.device)
        mg.update_A(N)
        omega = a.omega

        def smoother(level, u, f):
            return mg.smoother_jacobi(level, u, f, omega=omega)
        sol = torch.zeros_like(rhs)
        for _ in range(a.linsolver_maxiter or 50):
            sol = mg.step(sol, rhs, smoother, ndirect=a.ndirect,
                          pre=a.smooth_pre, post=a.smooth_post)
        sol = sol.view((len(names), ) + tuple(problem.domain.shape))
        for k, name in enumerate(names):
            state.fields[name] += sol[k]
    return hist


def optimize_multigridop(problem, state, epochs, args=None, callback=None,
                         tol=None, precision='fp64', maxiter=None):
    """util.optimize_multigridop_base (util.py:509-604) on the device:
    Newton steps whose normal equations are solved by flexible CG with a
    GMG V-cycle preconditioner (`NewtonMultigrid`)."""
    problem.init_missing(state)
    names = problem.domain.fieldnames
    fields = [state.fields[k] for k in names]
    a = args or ls.default_args()
    nm = NewtonMultigrid(
        problem.operator, nlvl=a.nlvl or 7, omega=a.omega,
        smooth_pre=a.smooth_pre, smooth_post=a.smooth_post,
        ndirect=a.ndirect,
        maxiter=maxiter or a.linsolver_maxiter or 10, precision=precision)
    cb = None
    if callback is not None:
        def cb(epoch, f, rms):
            callback(epoch, state, rms)
    return nm.run(fields, epochs, tol=tol, callback=cb)


def make_loss_grad(problem):
    """loss_grad(packed, epoch) -> (loss, grad, residual_list) over the
    packed fields (util.py:273-286)."""

    def loss_grad(packed, epoch):
        from odil_amd.problem import State
        st = State(problem.unpack_fields(packed))
        loss, grads, _, split = problem.eval_loss_grad(st, epoch)
        last_residual = [torch.sqrt(torch.clamp(v, min=0)) for v in split]
        return loss, problem.pack_fields(grads), last_residual

    return loss_grad


def optimize_opt(problem, state, optname, epochs, callback=None,
                 epoch_start=0, lr=1e-3, **kwargs):
    """util.optimize_opt (util.py:271-317): run `make_optimizer(optname)`
    on the packed state with the device loss and adjoint. Returns
    (state, info); `state.fields` hold the final fields."""
    packed = problem.pack_state(state)
    opt = opt_mod.make_optimizer(optname, dtype=problem.domain.dtype,
                                 **kwargs)
    x, info = opt.run(packed, make_loss_grad(problem), epochs=epochs,
                      callback=callback, epoch_start=epoch_start, lr=lr,
                      **kwargs)
    for k, v in problem.unpack_fields(x).items():
        state.fields[k] = v.clone()
    info.optimizer = opt
    return state, info
