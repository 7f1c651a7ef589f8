"""Newton drivers with a geometric-multigrid linear solver on device.

Each Newton step follows `util.optimize_multigridop_base` (util.py:509-604):
linearize at the current state (residual f and stencil Jacobian J, frozen
fields excluded, so for the cavity this is the reference's Picard-type
Newton), form the normal equations J^T J dx = -J^T f, solve them with
geometric multigrid, and update the state in place (`packed += dpacked`,
util.py:604).

Kernels per step
  * rhs = -J^T f and the residual's sums of squares: one fused adjoint kernel
    (FP64);
  * J as stored coefficient fields (FP64) and the stored normal operator
    N = J^T J in the working precision, from which the Galerkin hierarchy is
    built (coarse levels only);
  * on the finest level N is never read from memory: products, residuals
    and Jacobi sweeps recompute J from the state inside one tile kernel
    (`Operator.normal_apply`), which moves ~15 values per cell instead of the
    ~100 coefficient fields of a stored N.

Precision modes
  'fp64'  : everything in FP64 (the plain path).
  'mixed' : rhs, Krylov vectors and the Krylov operator N p in FP64 (so the
            linear system is refined in FP64); the multigrid preconditioner
            (FP32 state copy, FP32 Galerkin hierarchy, FP32 smoothing) in
            FP32; the coarsest level solve in FP64.

Linear solvers
  'pcg'   : flexible conjugate gradients on N (Polak-Ribiere beta, which
            tolerates a preconditioner that is not exactly linear, such as an
            FP32 V-cycle), preconditioned by one V(pre, post) cycle, `maxiter`
            iterations (optionally stopped at relative residual `rtol`,
            checked every `check_every` iterations).
  'vcycle': `maxiter` stationary V-cycles from zero in the working precision
            (the reference driver's loop, util.py:588-595).

Coarse operators default to the exact Galerkin product R A T with the
variational restriction R = T^T / 2^dim ('transpose'). With the reference's
restriction ('full', injection at coarse boundary nodes) the coarse normal
equations of the cavity are unsymmetric, singular and inconsistent, the
V-cycle is not a symmetric preconditioner and CG stalls; with the reference's
stencil coarsening (`galerkin_mode='reference'`), which is exact only for
reflection-symmetric stencils, the cavity V-cycle diverges. Both remain
available.
"""

import ctypes
import time

import torch

from odil_amd import _lib
from odil_amd._lib import c_double, c_int, c_vp, check, ptrs, stream
from odil_amd.graphs import capture
from odil_amd.multigrid import (BlockStencil, MultigridOp, jacobi_zero,
                               normal_operator)


class DeviceScalars:
    """Small FP64 device buffer for Krylov scalars."""

    def __init__(self, n, device):
        self.t = torch.zeros(n, dtype=torch.float64, device=device)

    def p(self, k):
        return c_vp(self.t.data_ptr() + 8 * k)


def _ll(n):
    return ctypes.c_longlong(int(n))


class VectorOps:
    """Dot / axpy kernels on lists of fields (HIP, FP64 accumulation)."""

    def __init__(self, device):
        self.partial = torch.zeros(_lib.lib().odil_dot_partials(),
                                   dtype=torch.float64, device=device)
        self.partial2 = torch.zeros(2 * _lib.lib().odil_dot_partials(),
                                    dtype=torch.float64, device=device)

    def dot(self, x, y, out_ptr, keep_prev=False):
        check(_lib.lib().odil_dot(c_int(_lib.dtype_tag(x[0].dtype)),
                                  c_int(len(x)), _ll(x[0].numel()), ptrs(x),
                                  ptrs(y), c_vp(self.partial.data_ptr()),
                                  out_ptr, c_int(1 if keep_prev else 0),
                                  stream()), 'dot')

    def axpy(self, y, x, num, den, sign):
        check(_lib.lib().odil_axpy_ratio(c_int(_lib.dtype_tag(y[0].dtype)),
                                         c_int(len(y)), _ll(y[0].numel()),
                                         ptrs(y), ptrs(x), num, den,
                                         c_double(sign), stream()),
              'axpy_ratio')

    def axpy2(self, x, p, r, q, num, den):
        """x += (num / den) p and r -= (num / den) q in one pass."""
        check(_lib.lib().odil_axpy2_ratio(c_int(_lib.dtype_tag(x[0].dtype)),
                                          c_int(len(x)), _ll(x[0].numel()),
                                          ptrs(x), ptrs(p), ptrs(r), ptrs(q),
                                          num, den, stream()), 'axpy2_ratio')

    def axpy2_jacobi_zero(self, x, p, r, q, num, den, d, y, omega):
        """axpy2, then y = omega r / d with the updated r, in one pass."""
        check(_lib.lib().odil_axpy2_jacobi_zero(
            c_int(_lib.dtype_tag(x[0].dtype)), c_int(len(x)),
            _ll(x[0].numel()), ptrs(x), ptrs(p), ptrs(r), ptrs(q), num, den,
            ptrs(d), ptrs(y), c_double(omega), stream()),
            'axpy2_jacobi_zero')

    def fcg_z(self, zw, r, zo, rz, rz_prev, rzo):
        """z = zw in FP64; rz_prev = rz, rz = r.z, rzo = r.zo; zo = z."""
        check(_lib.lib().odil_fcg_z(c_int(_lib.dtype_tag(zw[0].dtype)),
                                    c_int(_lib.dtype_tag(r[0].dtype)),
                                    c_int(len(zw)), _ll(zw[0].numel()),
                                    ptrs(zw), ptrs(r), ptrs(zo),
                                    c_vp(self.partial2.data_ptr()), rz,
                                    rz_prev, rzo, stream()), 'fcg_z')

    def fcg_dots(self, z, r, zo, rz, rz_prev, rzo):
        """rz_prev = rz, rz = r.z, rzo = r.zo (fcg_z's sums, no copy)."""
        check(_lib.lib().odil_fcg_dots(c_int(_lib.dtype_tag(z[0].dtype)),
                                       c_int(len(z)), _ll(z[0].numel()),
                                       ptrs(z), ptrs(r), ptrs(zo),
                                       c_vp(self.partial2.data_ptr()), rz,
                                       rz_prev, rzo, stream()), 'fcg_dots')

    def xpay(self, p, z, num, den, sub=None):
        """p = z + ((num - sub) / den) p with device scalars."""
        check(_lib.lib().odil_xpay_ratio(c_int(_lib.dtype_tag(p[0].dtype)),
                                         c_int(len(p)), _ll(p[0].numel()),
                                         ptrs(p), ptrs(z), num, sub, den,
                                         stream()), 'xpay_ratio')


class FusedNormal:
    """Finest-level normal operator N = J^T J applied matrix-free from the
    state (same interface as `BlockStencil`: apply, residual, jacobi)."""

    def __init__(self, op, state):
        self.op = op
        self.state = state
        self.diag = [torch.empty_like(a) for a in state]

    def refresh(self):
        """Recompute diag(N) after the state changed."""
        self.op.normal_apply(self.state, None, self.diag, mode=3)

    def apply(self, x, y):
        return self.op.normal_apply(self.state, x, y, mode=0)

    def residual(self, x, f, y):
        return self.op.normal_apply(self.state, x, y, b=f, mode=1)

    def jacobi(self, x, f, y, omega):
        return self.op.normal_apply(self.state, x, y, b=f, d=self.diag,
                                    omega=omega, mode=2)

    def jacobi_zero(self, f, y, omega, f_src=None):
        """Jacobi sweep from x = 0 (y = omega f / diag N); with f_src, reads
        it and stores its conversion into f."""
        return jacobi_zero(f if f_src is None else f_src, self.diag, y, omega,
                           fw=None if f_src is None else f)

    def diagonal(self):
        return self.diag


class NewtonMultigrid:
    """Newton iteration for an `odil_amd.operators` 2D operator with a GMG
    normal-equation solver. `fields` are FP64 device tensors updated in
    place by `step()`."""

    def __init__(self, op, nlvl=7, smooth_pre=2, smooth_post=2, omega=0.6,
                 ndirect=3, maxiter=10, restriction='transpose',
                 galerkin_mode='exact', precision='fp64', linsolver='pcg',
                 rtol=0., check_every=10, coarse_reg='auto',
                 matrix_free=True, use_graph=True, setup_every=1,
                 krylov='auto', refine=0, overlap_setup=False,
                 setup_first=0, coarse_smooth=(2, 2), omega_auto=True,
                 device=None):
        if precision not in ('fp64', 'mixed'):
            raise ValueError('precision must be fp64 or mixed')
        if linsolver not in ('vcycle', 'pcg'):
            raise ValueError('linsolver must be vcycle or pcg')
        if krylov not in ('auto', 'fp64', 'fp32'):
            raise ValueError('krylov must be auto, fp64 or fp32')
        self.op = op
        self.device = device or torch.device('cuda')
        self.wdtype = torch.float64 if precision == 'fp64' else torch.float32
        self.precision = precision
        # Krylov precision: 'auto' is that of the preconditioner. An FP32
        # flexible CG solves the correction equation of each FP64 Newton
        # residual; `refine` extra passes each recompute the FP64 linear
        # residual b - N x and add an FP32 solve for it (iterative
        # refinement).
        if krylov == 'auto':
            krylov = 'fp64' if precision == 'fp64' else 'fp32'
        if krylov == 'fp32' and precision == 'fp64':
            raise ValueError('an FP32 Krylov solve needs precision=mixed')
        self.kdtype = torch.float64 if krylov == 'fp64' else torch.float32
        self.refine = max(0, int(refine))
        if self.refine and self.kdtype == torch.float64:
            raise ValueError('refine needs krylov=fp32')
        self.linsolver = linsolver
        self.maxiter = int(maxiter)
        self.rtol = float(rtol)
        self.check_every = max(1, int(check_every))
        self.omega = float(omega)
        # Cap the damping per level at 1.8 / lambda_max(D^-1 A), estimated
        # at every hierarchy build (MultigridOp.estimate_damping): a fixed
        # omega diverges on the normal equations of high cell Reynolds
        # numbers.
        self.omega_auto = bool(omega_auto)
        self.pre, self.post, self.ndirect = smooth_pre, smooth_post, ndirect
        self.matrix_free = matrix_free
        self.use_graph = use_graph
        # Rebuild the stored hierarchy (J, J^T J, Galerkin levels, coarsest
        # solution operator) every `setup_every` Newton steps; the Krylov
        # operator and the finest level are always those of the current
        # state, so a lagged hierarchy only changes the preconditioner.
        self.setup_every = max(1, int(setup_every))
        # ... and at each of the first `setup_first` steps, while the state
        # still moves far from the one the hierarchy was built for.
        self.setup_first = max(0, int(setup_first))
        self._nstep = 0
        # J^T J and its Galerkin products R N T with R = T^T / 2^dim are
        # symmetric: store upper halves.
        self.sym = galerkin_mode == 'exact' and restriction == 'transpose'
        self.nvar = len(op.fieldnames)
        if coarse_reg == 'auto':
            # Tikhonov least squares on the coarsest level. The damping sits
            # above the working precision's round-off of an exactly singular
            # direction (~eps * sigma_max in storage) but below the smallest
            # genuine singular values of the 17x17 normal equations
            # (~1e-5 sigma_max, squared 1e-10): 1e-10 would damp the
            # smoothest coarse modes and slows Newton down threefold.
            coarse_reg = 1e-14 if precision == 'fp64' else 1e-12
        self.coarse_smooth = tuple(int(x) for x in coarse_smooth)
        if len(self.coarse_smooth) != 2 or min(self.coarse_smooth) < 0:
            raise ValueError('coarse_smooth must be (pre, post) >= 0')

        def mk():
            m = MultigridOp(op.shape, nvar=self.nvar, nlevels=nlvl,
                            restriction=restriction, dtype=self.wdtype,
                            device=self.device, coarse_reg=coarse_reg,
                            galerkin_mode=galerkin_mode)
            m.coarse_pre, m.coarse_post = self.coarse_smooth
            return m
        self.mg = mk()
        # overlap_setup: rebuilds after the first run on a second stream
        # into a second hierarchy while the current step's linear solve
        # runs, and are adopted at the next step (one step more lag).
        self.overlap_setup = bool(overlap_setup) and matrix_free
        self._mgs = [self.mg, mk()] if self.overlap_setup else [self.mg]
        self._side = torch.cuda.Stream(self.device) \
            if self.overlap_setup else None
        self._pending = None
        self.vec = VectorOps(self.device)
        self.sc = DeviceScalars(8, self.device)
        self.N = None
        self._ndot = None
        self._graph = {}
        self._cycle_tuned = set()
        self._zi = 0
        self._ready = False
        self.timers = {}
        self.last_linear = None

    # -- buffers ---------------------------------------------------------
    def _vecs(self, dtype):
        return [torch.zeros(self.op.shape, dtype=dtype, device=self.device)
                for _ in range(self.nvar)]

    def _setup(self, fields):
        f64 = torch.float64
        self.rhs = self._vecs(f64)
        self.x = self._vecs(f64)
        if self.wdtype == f64:
            self.state_w = fields
        else:
            self.state_w = self._vecs(self.wdtype)
        self.fine = FusedNormal(self.op, self.state_w) \
            if self.matrix_free else None
        self.b_w = self._vecs(self.wdtype)
        self.x_w = self._vecs(self.wdtype)
        if self.linsolver == 'pcg':
            kd = self.kdtype
            self.r, self.z, self.p, self.q, self.zo = (
                self._vecs(kd), self._vecs(kd), self._vecs(kd),
                self._vecs(kd), self._vecs(kd))
            # Krylov solution (the FP64 solution itself for an FP64 solve)
            # and the FP64 linear residual of the refinement passes.
            self.xk = self.x if kd == f64 else self._vecs(kd)
            self.r64 = self._vecs(f64) if self.refine else None
            # Partials of the product launch that also forms p.Np.
            self._ndot = None
            has = getattr(self.op, 'has_normal_dot', None)
            if self.matrix_free and has is not None and has():
                self._ndot = torch.empty(
                    self.op.nblocks(2, _lib.dtype_tag(kd)),
                    dtype=f64, device=self.device)
        self._fields = fields
        self._ready = True

    # -- linear solve on fixed buffers (capturable) ----------------------
    def _precond(self, r, first_done=False, out=None):
        """M r: one V-cycle from zero in the working precision (an FP32
        cycle converts an FP64 r in its first sweep). Returns the result in
        the working precision (`out`, default self.z, or self.x_w).
        first_done: the finest level's first sweep is already in its buffer
        (see _fused_sweep)."""
        if self.wdtype == r[0].dtype:
            u = self.z if out is None else out
            self.mg.vcycle(u, r, omega=self.omega, ndirect=self.ndirect,
                           pre=self.pre, post=self.post, zero=True,
                           first_done=first_done)
            return u
        self.mg.vcycle(self.x_w, self.b_w, omega=self.omega,
                       ndirect=self.ndirect, pre=self.pre, post=self.post,
                       zero=True, f_src=r)
        return self.x_w

    def _apply_N(self, p, q, dot=None):
        """q = N p on the finest level in the Krylov precision (an FP32
        solve applies N matrix-free from the FP32 copy of the state). With
        `dot` (device scalar pointer), a matrix-free operator that fuses it
        also writes p.q there in the same launch; returns whether it did."""
        fuse = dot is not None and self._ndot is not None
        if self.kdtype == torch.float32:
            if self.matrix_free:
                if fuse:
                    self.op.normal_apply_dot(self.state_w, p, q, self._ndot,
                                             dot)
                    return True
                self.op.normal_apply(self.state_w, p, q, mode=0)
            else:
                self.N.apply(p, q)
        elif self.matrix_free:
            if fuse:
                self.op.normal_apply_dot(self._fields, p, q, self._ndot, dot)
                return True
            self.op.normal_apply(self._fields, p, q, mode=0)
        else:
            self._N64.apply(p, q)
        return False

    def _fused_sweep(self):
        """True when the residual update of each iteration also makes the
        preconditioner cycle's first sweep from zero (y = omega r / d on
        the finest level): r in the working precision, a pre-sweep and a
        finest level that is smoothed, not solved directly."""
        return (self.wdtype == self.kdtype and self.pre > 0
                and not self.mg._is_direct(0, self.ndirect))

    def _pingpong(self):
        """True when the preconditioned residuals go to z and zo in turn
        (no copy of the new z into zo): z already in the Krylov precision
        and a fixed iteration budget, so that every solve (and every replay
        of its captured graph) runs the same buffer sequence from
        _fcg_pass on."""
        return self.rtol <= 0 and self.wdtype == self.kdtype

    def _pcg_iters(self, iters):
        # Scalars: s0 = r.z, s1 = previous r.z, s2 = p.N p, s3 = r.z_old,
        # s4 = |b|^2, s6 = |r|^2. zo holds the latest preconditioned
        # residual z in the Krylov precision (fcg_z moves the new z there
        # after using the old one for s3); with _pingpong the latest z is
        # in (z, zo)[self._zi] and the next goes to the other buffer.
        v, s = self.vec, self.sc
        fused = self._fused_sweep()
        pp = self._pingpong()
        bufs = (self.z, self.zo)
        for _ in range(iters):
            if not self._apply_N(self.p, self.q, dot=s.p(2)):
                v.dot(self.p, self.q, s.p(2))
            if fused:
                d, y0, om = self.mg.first_sweep(self.omega)
                v.axpy2_jacobi_zero(self.xk, self.p, self.r, self.q, s.p(0),
                                    s.p(2), d, y0, om)
            else:
                v.axpy2(self.xk, self.p, self.r, self.q, s.p(0), s.p(2))
            if pp:
                new, old = bufs[1 - self._zi], bufs[self._zi]
                self._precond(self.r, first_done=fused, out=new)
                v.fcg_dots(new, self.r, old, s.p(0), s.p(1), s.p(3))
                v.xpay(self.p, new, s.p(0), s.p(1), sub=s.p(3))
                self._zi = 1 - self._zi
                continue
            zw = self._precond(self.r, first_done=fused)
            v.fcg_z(zw, self.r, self.zo, s.p(0), s.p(1), s.p(3))
            v.xpay(self.p, self.zo, s.p(0), s.p(1), sub=s.p(3))
        v.dot(self.r, self.r, s.p(6))

    def _solve_body(self, iters, finish=True):
        if self.linsolver == 'vcycle':
            for a, b in zip(self.b_w, self.rhs):
                a.copy_(b)
            for a in self.x_w:
                a.zero_()
            for _ in range(iters):
                self.mg.vcycle(self.x_w, self.b_w, omega=self.omega,
                               ndirect=self.ndirect, pre=self.pre,
                               post=self.post)
            for a, b in zip(self.x, self.x_w):
                a.copy_(b)
            return
        self._fcg_pass(self.rhs, iters)
        if finish:
            self._finish()

    def _finish(self):
        """x = xk in FP64, then the refinement passes."""
        if self.xk is not self.x:
            for a, b in zip(self.x, self.xk):
                a.copy_(b)
        iters = self.maxiter
        for _ in range(self.refine):
            # (With rtol the passes run eagerly with the full budget.)
            # FP64 residual of the current solution, solved again in FP32.
            self.op.normal_apply(self._fields, self.x, self.r64, b=self.rhs,
                                 mode=1)
            self._fcg_pass(self.r64, iters)
            for a, b in zip(self.x, self.xk):
                a.add_(b)

    def _fcg_pass(self, b, iters):
        """Flexible CG from x = 0 on N x = b (b FP64, converted to the
        Krylov precision): xk holds the result."""
        v, s = self.vec, self.sc
        for a, c in zip(self.r, b):
            a.copy_(c)
        for a in self.xk:
            a.zero_()
        zw = self._precond(self.r)
        # s0 = r.z and zo = z (s1, s3 are not used before being set again);
        # with _pingpong z stays where it is (zw is self.z).
        if self._pingpong():
            v.fcg_dots(zw, self.r, self.zo, s.p(0), s.p(1), s.p(3))
            self._zi = 0
            src = zw
        else:
            v.fcg_z(zw, self.r, self.zo, s.p(0), s.p(1), s.p(3))
            src = self.zo
        for a, c in zip(self.p, src):
            a.copy_(c)
        v.dot(b, b, s.p(4))
        self._pcg_iters(iters)

    def _tune_cycle(self):
        """Launch shapes of the stored levels timed inside the cycle this
        solver applies (once per hierarchy object, before its solve is
        captured; later hierarchies of the same shapes reuse the choice)."""
        if id(self.mg) in self._cycle_tuned:
            return
        if self.linsolver == 'pcg':
            run = lambda: self._precond(self.r)
        else:
            run = lambda: self.mg.vcycle(self.x_w, self.b_w,
                                         omega=self.omega,
                                         ndirect=self.ndirect, pre=self.pre,
                                         post=self.post)
        self.mg.tune_in_cycle(run)
        self._cycle_tuned.add(id(self.mg))

    def _run_chunk(self, iters, fresh=True, finish=True):
        if not fresh:
            self._pcg_iters(iters)
            return
        self._tune_cycle()
        key = (iters, finish, id(self.mg))
        if not self.use_graph:
            self._solve_body(iters, finish)
            return
        # The damping of each level is a kernel argument inside the graph:
        # recapture when the hierarchy's caps changed.
        om = self.mg.omega_key()
        if key in self._graph and self._graph[key][0] != om:
            del self._graph[key]
        if key not in self._graph:
            # First use: record the solve (capture does not execute it), then
            # replay it like every later step.
            self._graph[key] = (om, capture(
                lambda: self._solve_body(iters, finish), self.device))
        self._graph[key][1].replay()

    def _linear_solve(self):
        """Krylov iterations run (refinement passes included)."""
        if self.linsolver == 'vcycle' or self.rtol <= 0:
            self._run_chunk(self.maxiter)
            return self.maxiter * (1 + self.refine)
        done = 0
        while done < self.maxiter:
            k = min(self.check_every, self.maxiter - done)
            self._run_chunk(k, fresh=(done == 0), finish=False)
            done += k
            rr, bb = self.sc.t[6].item(), self.sc.t[4].item()
            if bb == 0 or (rr / bb)**0.5 <= self.rtol:
                break
        # Same FP64 solution copy and refinement passes as a fixed budget.
        self._finish()
        return done + self.refine * self.maxiter

    # -- Newton ------------------------------------------------------------
    def residual_rms(self, fields):
        _, ssq = self.op.residual(fields, want_fields=False)
        return [(x / self.op.ncell)**0.5 for x in ssq.tolist()]

    def step(self, fields, tol=None):
        """One Newton step in place. Returns the per-equation RMS residual at
        the state before the update (`opt.last_residual`, util.py:253). With
        `tol`, no update is made when max(rms) <= tol."""
        if not self._ready:
            self._setup(fields)
        elif any(a.data_ptr() != b.data_ptr()
                 for a, b in zip(fields, self._fields)):
            raise ValueError('NewtonMultigrid steps one set of field tensors '
                             '(updated in place)')
        t0 = time.perf_counter()
        ssq = self.op.jtf(fields, self.rhs, -1.0)
        rms = [(x / self.op.ncell)**0.5 for x in ssq.tolist()]
        if tol is not None and max(rms) <= tol:
            self.last_linear = 0
            return rms
        main = torch.cuda.current_stream()
        need_damping = False
        if self._pending is not None:
            ev, mg = self._pending
            main.wait_event(ev)
            self.mg = mg
            self._pending = None
            need_damping = True
        side = None
        if (self._nstep % self.setup_every == 0 or not self.mg.AA
                or self._nstep < self.setup_first
                or not self.matrix_free):
            if self.overlap_setup and self.mg.AA:
                side = torch.cuda.Event()
                side.record()
            else:
                self._build(fields, self.mg)
                need_damping = True
                for mg in self._mgs:
                    if not mg.AA:
                        # First build of the other hierarchy here, on this
                        # stream: it allocates and tunes its launch shapes.
                        mg.update_A(self.N)
                        mg.prepare(self.ndirect)
        self._nstep += 1
        if self.matrix_free:
            if self.state_w is not fields:
                for a, b in zip(self.state_w, fields):
                    a.copy_(b)
            self.fine.refresh()
            self.mg.fine = self.fine
        if need_damping and self.omega_auto:
            # On the hierarchy the solve uses, finest level included (the
            # matrix-free operator of the current state).
            self.mg.estimate_damping()
        self.timers = {'setup': time.perf_counter() - t0}
        its = self._linear_solve()
        if side is not None:
            # Queued after the solve so that its launches do not delay the
            # solve's; the side Jacobian reads the fields before the update.
            main.wait_event(self._setup_side(fields, side))
        for f, d in zip(fields, self.x):
            f += d
        self.last_linear = its
        return rms

    def _build(self, fields, mg):
        """J, N = J^T J (stored), Galerkin hierarchy and coarsest solution
        operator of `mg`, on the current stream; returns after queueing the
        Jacobian and the rest (the Jacobian is the only reader of fields)."""
        jac, _ = self.op.jacobian(fields)
        ev = torch.cuda.Event()
        ev.record()
        J = BlockStencil.from_jacobian(jac)
        if self.N is None:
            self.N = normal_operator(J, self.wdtype, sym=self.sym)
        else:
            normal_operator(J, out=self.N)
        if not self.matrix_free:
            self._N64 = self.N if self.wdtype == torch.float64 else \
                normal_operator(J, torch.float64, sym=self.sym)
        mg.update_A(self.N)
        mg.prepare(self.ndirect)
        return ev

    def _setup_side(self, fields, after):
        """Rebuild the inactive hierarchy on the side stream after event
        `after`; it is adopted at the next step. Returns the event after
        which the fields may change."""
        target = self._mgs[1] if self.mg is self._mgs[0] else self._mgs[0]
        target.fine = self.fine
        self._side.wait_event(after)
        with torch.cuda.stream(self._side):
            jac_ready = self._build(fields, target)
            done = torch.cuda.Event()
            done.record()
        self._pending = (done, target)
        return jac_ready

    def run(self, fields, epochs, tol=None, callback=None):
        """Newton iterations until the max per-equation RMS residual is
        <= tol (checked before each update) or `epochs` updates. Returns a
        list of (epoch, rms list, wall seconds)."""
        hist = []
        t0 = time.perf_counter()
        for epoch in range(epochs + 1):
            if epoch == epochs:
                rms = self.residual_rms(fields)
                hist.append((epoch, rms, time.perf_counter() - t0))
                break
            rms = self.step(fields, tol=tol)
            hist.append((epoch, rms, time.perf_counter() - t0))
            if callback:
                callback(epoch, fields, rms)
            if tol is not None and max(rms) <= tol:
                break
        return hist
