"""Newton drivers with a geometric-multigrid linear solver on device.

Each Newton step follows `util.optimize_multigridop_base` (util.py:509-604):
linearize at the current state (residual f and stencil Jacobian J, frozen
fields excluded, so for the cavity this is the reference's Picard-type
Newton), form the normal equations J^T J dx = -J^T f in stencil form, build
the Galerkin hierarchy, solve with V-cycles, and update the state in place
(`packed += dpacked`, util.py:604).

Precision modes
  'fp64'  : everything in FP64 (the plain path).
  'mixed' : residual, J and the right-hand side in FP64; the normal operator,
            its Galerkin hierarchy and all smoothing in FP32; the coarsest
            level solve in FP64; the correction is accumulated in FP64.
            `ir_steps` > 0 adds FP64 iterative refinement of the linear
            system: r = b - J^T (J x) in FP64, solve again in FP32, add.

Coarse operators default to the exact Galerkin product R A T with the
variational restriction R = T^T / 2^dim ('transpose'). With the reference's
restriction ('full', injection at coarse boundary nodes) the coarse normal
equations of the cavity are unsymmetric, singular and inconsistent, the
V-cycle is not a symmetric preconditioner and CG stalls. The
reference's stencil coarsening (`galerkin_mode='reference'`) is exact only
for reflection-symmetric stencils; on the cavity normal equations its V-cycle
diverges, so it is available but not the default here.

Linear solvers
  'vcycle': `maxiter` stationary V-cycles from zero (the reference driver).
  'pcg'   : conjugate gradients on J^T J preconditioned by one V-cycle,
            stopped after `maxiter` iterations or at relative residual `rtol`
            (checked every `check_every` iterations).
"""

import time

import numpy as np
import torch

from odil_amd import _lib
from odil_amd._lib import c_double, c_int, c_vp, check, ptrs, stream
from odil_amd.multigrid import BlockStencil, MultigridOp, normal_operator


class DeviceScalars:
    """Small FP64 device buffer for Krylov scalars."""

    def __init__(self, n, device):
        self.t = torch.zeros(n, dtype=torch.float64, device=device)

    def p(self, k):
        return c_vp(self.t.data_ptr() + 8 * k)


class VectorOps:
    """Dot / axpy kernels on lists of fields (HIP, FP64 accumulation)."""

    def __init__(self, device):
        self.partial = torch.zeros(_lib.lib().odil_dot_partials(),
                                   dtype=torch.float64, device=device)

    def dot(self, x, y, out_ptr, keep_prev=False):
        n = x[0].numel()
        check(_lib.lib().odil_dot(c_int(_lib.dtype_tag(x[0].dtype)),
                                  c_int(len(x)), ctypes_ll(n), ptrs(x),
                                  ptrs(y), c_vp(self.partial.data_ptr()),
                                  out_ptr, c_int(1 if keep_prev else 0),
                                  stream()), 'dot')

    def axpy(self, y, x, num, den, sign):
        check(_lib.lib().odil_axpy_ratio(c_int(_lib.dtype_tag(y[0].dtype)),
                                         c_int(len(y)),
                                         ctypes_ll(y[0].numel()), ptrs(y),
                                         ptrs(x), num, den, c_double(sign),
                                         stream()), 'axpy_ratio')

    def xpay(self, p, z, num, den):
        check(_lib.lib().odil_xpay_ratio(c_int(_lib.dtype_tag(p[0].dtype)),
                                         c_int(len(p)),
                                         ctypes_ll(p[0].numel()), ptrs(p),
                                         ptrs(z), num, den, stream()),
              'xpay_ratio')


def ctypes_ll(n):
    import ctypes
    return ctypes.c_longlong(int(n))


class NewtonMultigrid:
    """Newton iteration for an `odil_amd.operators` operator with a GMG
    normal-equation solver. `fields` are FP64 device tensors updated in
    place by `step()`."""

    def __init__(self, op, nlvl=7, smooth_pre=2, smooth_post=2, omega=0.6,
                 ndirect=3, maxiter=50, restriction='transpose',
                 galerkin_mode='exact', precision='fp64',
                 linsolver='vcycle', rtol=0., check_every=10, ir_steps=0,
                 coarse_reg='auto', use_graph=True, device=None):
        if precision not in ('fp64', 'mixed'):
            raise ValueError('precision must be fp64 or mixed')
        if linsolver not in ('vcycle', 'pcg'):
            raise ValueError('linsolver must be vcycle or pcg')
        self.op = op
        self.device = device or torch.device('cuda')
        self.wdtype = torch.float64 if precision == 'fp64' else torch.float32
        self.precision = precision
        self.linsolver = linsolver
        self.maxiter = int(maxiter)
        self.rtol = float(rtol)
        self.check_every = max(1, int(check_every))
        self.ir_steps = int(ir_steps)
        self.omega = float(omega)
        self.pre, self.post, self.ndirect = smooth_pre, smooth_post, ndirect
        self.use_graph = use_graph
        self.nvar = len(op.fieldnames)
        if coarse_reg == 'auto':
            # Tikhonov least squares on the coarsest level, with a damping
            # above the working precision's round-off (a singular value that
            # is zero in exact arithmetic is ~eps * sigma_max in storage).
            coarse_reg = 1e-14 if precision == 'fp64' else 1e-10
        self.mg = MultigridOp(op.shape, nvar=self.nvar, nlevels=nlvl,
                              restriction=restriction, dtype=self.wdtype,
                              device=self.device, coarse_reg=coarse_reg,
                              galerkin_mode=galerkin_mode)
        self.vec = VectorOps(self.device)
        self.sc = DeviceScalars(8, self.device)
        self.N = None
        self._graph = None
        self.timers = {}
        self.last_linear = None

    # -- pieces ----------------------------------------------------------
    def _w(self):
        return [torch.zeros(self.op.shape, dtype=self.wdtype,
                            device=self.device) for _ in range(self.nvar)]

    def _alloc(self):
        self.b = self._w()
        self.x = self._w()
        if self.linsolver == 'pcg':
            self.r, self.z, self.p, self.q = (self._w(), self._w(),
                                              self._w(), self._w())

    def _precond(self, r, z):
        for a in z:
            a.zero_()
        self.mg.vcycle(z, r, omega=self.omega, ndirect=self.ndirect,
                       pre=self.pre, post=self.post)

    def _solve_body(self, iters):
        """Linear solve on fixed buffers: b -> x (capturable)."""
        if self.linsolver == 'vcycle':
            for a in self.x:
                a.zero_()
            for _ in range(iters):
                self.mg.vcycle(self.x, self.b, omega=self.omega,
                               ndirect=self.ndirect, pre=self.pre,
                               post=self.post)
            return
        A = self.mg.AA[0]
        v, s = self.vec, self.sc
        for a, bb in zip(self.r, self.b):
            a.copy_(bb)
        for a in self.x:
            a.zero_()
        self._precond(self.r, self.z)
        for a, bb in zip(self.p, self.z):
            a.copy_(bb)
        v.dot(self.r, self.z, s.p(0))
        v.dot(self.b, self.b, s.p(4))  # |b|^2 for the stopping test
        for _ in range(iters):
            A.apply(self.p, self.q)
            v.dot(self.p, self.q, s.p(2))
            v.axpy(self.x, self.p, s.p(0), s.p(2), 1.)
            v.axpy(self.r, self.q, s.p(0), s.p(2), -1.)
            self._precond(self.r, self.z)
            v.dot(self.r, self.z, s.p(0), keep_prev=True)
            v.xpay(self.p, self.z, s.p(0), s.p(1))
        v.dot(self.r, self.r, s.p(6))

    def _linear_solve(self):
        if self.linsolver == 'vcycle' or self.rtol <= 0:
            self._run_chunk(self.maxiter)
            return self.maxiter
        done = 0
        while done < self.maxiter:
            k = min(self.check_every, self.maxiter - done)
            if done == 0:
                self._run_chunk(k)
            else:
                self._pcg_continue(k)
            done += k
            rr, bb = self.sc.t[6].item(), self.sc.t[4].item()
            if bb == 0 or (rr / bb)**0.5 <= self.rtol:
                break
        return done

    def _pcg_continue(self, iters):
        A = self.mg.AA[0]
        v, s = self.vec, self.sc
        for _ in range(iters):
            A.apply(self.p, self.q)
            v.dot(self.p, self.q, s.p(2))
            v.axpy(self.x, self.p, s.p(0), s.p(2), 1.)
            v.axpy(self.r, self.q, s.p(0), s.p(2), -1.)
            self._precond(self.r, self.z)
            v.dot(self.r, self.z, s.p(0), keep_prev=True)
            v.xpay(self.p, self.z, s.p(0), s.p(1))
        v.dot(self.r, self.r, s.p(6))

    def _run_chunk(self, iters):
        if not self.use_graph:
            self._solve_body(iters)
            return
        key = iters
        if self._graph is None:
            self._graph = {}
        if key not in self._graph:
            # First use: run eagerly (allocates level buffers; the result
            # stays in x), then record the same work. Capture does not
            # execute, so x keeps the eager result.
            self._solve_body(iters)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._solve_body(iters)
            self._graph[key] = g
            return
        self._graph[key].replay()

    # -- Newton ------------------------------------------------------------
    def linearize(self, fields):
        """(residual fields, Jacobian BlockStencil) at `fields` (FP64)."""
        jac, res = self.op.jacobian(fields, want_residual=True)
        return res, BlockStencil.from_jacobian(jac)

    def step(self, fields, tol=None):
        """One Newton step in place. Returns the per-equation RMS residual at
        the state before the update (`opt.last_residual`, util.py:253). With
        `tol`, no update is made when max(rms) <= tol."""
        t = {}
        t0 = time.perf_counter()
        res, J = self.linearize(fields)
        rms = [torch.sqrt(torch.mean(r * r)) for r in res]
        if tol is not None:
            rms = [r.item() for r in rms]
            if max(rms) <= tol:
                self.last_linear = 0
                return rms
        # rhs = -J^T f in FP64
        rhs = J.apply_t([-r for r in res])
        if self.N is None:
            self.N = normal_operator(J, self.wdtype)
            self._alloc()
        else:
            normal_operator(J, out=self.N)
        self.mg.update_A(self.N)
        self.mg.prepare(self.ndirect)
        for a, bb in zip(self.b, rhs):
            a.copy_(bb)
        t['setup'] = time.perf_counter() - t0
        its = self._linear_solve()
        dx = [a.to(torch.float64) for a in self.x]
        for _ in range(self.ir_steps):
            # FP64 refinement: b - J^T J x, solved again in working precision.
            Jx = J.apply(dx)
            rl = J.apply_t(Jx)
            for a, bb, c in zip(self.b, rhs, rl):
                a.copy_(bb - c)
            self._linear_solve()
            for a, b in zip(dx, self.x):
                a += b.to(torch.float64)
        for f, d in zip(fields, dx):
            f += d
        self.last_linear = its
        self.timers = t
        return [float(r) for r in rms]

    def run(self, fields, epochs, tol=None, callback=None):
        """Newton iterations until the max per-equation RMS residual is
        <= tol (checked before each update) or `epochs` updates.
        Returns a list of (epoch, rms list, wall seconds)."""
        hist = []
        t0 = time.perf_counter()
        for epoch in range(epochs + 1):
            if epoch == epochs:
                _, ssq = self.op.residual(fields, want_fields=False)
                rms = [(x / self.op.ncell)**0.5 for x in ssq.tolist()]
                hist.append((epoch, rms, time.perf_counter() - t0))
                break
            rms = self.step(fields, tol=tol)
            hist.append((epoch, rms, time.perf_counter() - t0))
            if callback:
                callback(epoch, fields, rms)
            if tol is not None and max(rms) <= tol:
                break
        return hist
