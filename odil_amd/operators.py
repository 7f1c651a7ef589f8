"""Discrete ODIL operators backed by fused HIP kernels.

Each operator evaluates, on device tensors of the domain's shape,
* the residual fields f_e and their sums of squares,
* the stencil-Jacobian coefficient fields (entries (equation, field, shift),
  the tape Jacobian of `Problem._eval_grad0`, util_op.py:585-614),
* the loss sum_e mean(f_e^2) and its gradient (util_op.py:627-654) in one
  fused adjoint kernel.

The operators are the reference's PDE callbacks: lid-driven cavity
(ns_steady.py:142-394), 1D wave (wave.py:67-122), 2D advection
(advection_2d.py:103-223) and Laplace/Poisson (util_op.py:1038-1079).
"""

import ctypes

import numpy as np
import torch

from odil_amd import _lib
from odil_amd._lib import c_double, c_int, c_vp, check, ptr, ptrs

c_dp = ctypes.POINTER(ctypes.c_double)


class StencilJacobian:
    """Stencil Jacobian J: coefficient fields `coef[q]` for entries
    `entries[q] = (equation, field, shift)`; entry q contributes
    coef[q][c] * x_field[c + shift] to equation e at cell c (zero outside)."""

    def __init__(self, coef, entries, shape, neq, nvar):
        self.coef = coef
        self.entries = entries
        self.shape = tuple(shape)
        self.neq = neq
        self.nvar = nvar

    def blocks(self):
        """Dict {(e, k): {shift: coef_field}} like the reference's
        `stf[eq][var][shift]` (util.py:551-560)."""
        out = {}
        for q, (e, k, s) in enumerate(self.entries):
            out.setdefault((e, k), {})[s] = self.coef[q]
        return out


class Operator:
    """Base class: subclasses set `fieldnames`, `neq`, `prefix` and
    implement `_args()` returning the ctypes parameter list that follows
    (dtype, shape...) in the C entry points."""

    fieldnames = ()
    neq = 0
    prefix = ''

    def __init__(self, shape):
        self.shape = tuple(int(s) for s in shape)
        self.ncell = int(np.prod(self.shape))
        self._entries = None

    # -- helpers ---------------------------------------------------------
    def _shape_args(self):
        return [c_int(s) for s in self.shape]

    def _fn(self, what):
        return getattr(_lib.lib(), 'odil_{}_{}'.format(self.prefix, what))

    def _check_fields(self, fields):
        if len(fields) != len(self.fieldnames):
            raise ValueError('{} expects {} fields, got {}'.format(
                type(self).__name__, len(self.fieldnames), len(fields)))
        dt = fields[0].dtype
        for f in fields:
            _lib.require_device(f)
            if tuple(f.shape) != self.shape:
                raise ValueError('field shape {} != domain shape {}'.format(
                    tuple(f.shape), self.shape))
            if f.dtype != dt:
                raise ValueError('fields must share one dtype')
        return _lib.dtype_tag(dt)

    def nblocks(self, kind, dtag):
        return self._fn('nblocks')(c_int(kind), c_int(dtag),
                                   *self._shape_args())

    # -- raw launches on preallocated buffers (no allocation, capturable) --
    def launch_residual(self, fields, res, partial):
        """Residual kernel only: `res` fields (or None) and per-block sums of
        squares into `partial` ((neq, nblocks(0)) float64)."""
        dtag = self._check_fields(fields)
        st = self._fn('residual')(c_int(dtag), *self._shape_args(),
                                  *self._args_for(dtag), ptrs(fields),
                                  ptrs(res) if res else None, ptr(partial),
                                  _lib.stream())
        check(st, self.prefix + '_residual')

    def launch_loss_grad(self, fields, grads, partial):
        """Fused loss-adjoint kernel only; `partial` is
        (neq, nblocks(1)) float64."""
        dtag = self._check_fields(fields)
        nbo = c_int(0)
        st = self._fn('loss_grad')(c_int(dtag), *self._shape_args(),
                                   *self._args_for(dtag), ptrs(fields),
                                   ptrs(grads), ptr(partial),
                                   ctypes.byref(nbo), _lib.stream())
        check(st, self.prefix + '_loss_grad')

    def normal_apply(self, state, x, y, b=None, d=None, omega=0., mode=0):
        """Matrix-free N = J^T J with J recomputed from `state` (same dtype
        as x, y): mode 0 y = N x, 1 y = b - N x, 2 Jacobi sweep
        y = x + omega (b - N x) / d, 3 y = diag(N)."""
        dtag = self._check_fields(state)
        st = self._fn('normal_apply')(
            c_int(dtag), *self._shape_args(), *self._args_for(dtag),
            ptrs(state), ptrs(x) if x is not None else None,
            ptrs(b) if b is not None else None,
            ptrs(d) if d is not None else None, ptrs(y), c_double(omega),
            c_int(mode), _lib.stream())
        check(st, self.prefix + '_normal_apply')
        return y

    def has_normal_dot(self):
        """Whether the library fuses p.Np into the product (normal_apply_dot)."""
        return hasattr(_lib.lib(), 'odil_{}_normal_apply_dot'.format(
            self.prefix))

    def normal_apply_dot(self, state, x, y, partial, dot):
        """y = N x (mode 0 of normal_apply) and the FP64 sum of x * y into
        the device scalar at `dot` (a pointer), in one launch; `partial` is
        a float64 device tensor of nblocks(2, dtype) values."""
        dtag = self._check_fields(state)
        st = self._fn('normal_apply_dot')(
            c_int(dtag), *self._shape_args(), *self._args_for(dtag),
            ptrs(state), ptrs(x), ptrs(y), ptr(partial), dot, _lib.stream())
        check(st, self.prefix + '_normal_apply_dot')
        return y

    def jtf(self, state, out, scale, partial=None):
        """out = scale * J^T f(state); returns per-equation sums of f^2
        (float64 device tensor)."""
        dtag = self._check_fields(state)
        nb = self.nblocks(1, dtag)
        if partial is None:
            partial = torch.empty((self.neq, nb), dtype=torch.float64,
                                  device=state[0].device)
        nbo = c_int(0)
        st = self._fn('jtf')(c_int(dtag), *self._shape_args(),
                             *self._args_for(dtag), ptrs(state), ptrs(out),
                             c_double(scale), ptr(partial),
                             ctypes.byref(nbo), _lib.stream())
        check(st, self.prefix + '_jtf')
        ssq = torch.empty(self.neq, dtype=torch.float64,
                          device=state[0].device)
        check(_lib.lib().odil_reduce_partials(ptr(partial), c_int(self.neq),
                                              c_int(nb), ptr(ssq),
                                              _lib.stream()),
              'reduce_partials')
        return ssq

    def _args_for(self, dtag):
        """Parameters for a launch in dtype `dtag` (operators whose
        parameters hold field-typed arrays override this)."""
        return self._args()

    # -- API -------------------------------------------------------------
    def residual(self, fields, want_fields=True):
        """Returns (list of residual fields or None, ssq) where `ssq` is a
        float64 device tensor of per-equation sums of f_e^2."""
        dtag = self._check_fields(fields)
        dev = fields[0].device
        res = [torch.empty(self.shape, dtype=fields[0].dtype, device=dev)
               for _ in range(self.neq)] if want_fields else None
        nb = self.nblocks(0, dtag)
        partial = torch.empty((self.neq, nb), dtype=torch.float64,
                              device=dev)
        st = self._fn('residual')(c_int(dtag), *self._shape_args(),
                                  *self._args_for(dtag), ptrs(fields),
                                  ptrs(res) if res else None, ptr(partial),
                                  _lib.stream())
        check(st, self.prefix + '_residual')
        ssq = torch.empty(self.neq, dtype=torch.float64, device=dev)
        check(_lib.lib().odil_reduce_partials(ptr(partial), c_int(self.neq),
                                              c_int(nb), ptr(ssq),
                                              _lib.stream()),
              'reduce_partials')
        return res, ssq

    def entries(self):
        """Jacobian entry table [(equation, field, shift tuple)]."""
        if self._entries is None:
            ent = (ctypes.c_int * 3 * 256)()
            sh = (ctypes.c_int * (3 * 64))()
            n = self._fn('jac_entries')(ent, sh)
            ndim = len(self.shape)
            flat_e = np.ctypeslib.as_array(ent).reshape(-1)[:3 * n]
            flat_s = np.ctypeslib.as_array(sh)
            out = []
            for q in range(n):
                e, k, s = (int(x) for x in flat_e[3 * q:3 * q + 3])
                shift = tuple(int(x) for x in flat_s[ndim * s:ndim * s + ndim])
                out.append((e, k, shift))
            self._entries = out
        return self._entries

    def jacobian(self, fields, want_residual=False):
        """Returns (StencilJacobian, residual fields or None)."""
        dtag = self._check_fields(fields)
        ent = self.entries()
        dev = fields[0].device
        coef = torch.empty((len(ent), ) + self.shape, dtype=fields[0].dtype,
                           device=dev)
        res = [torch.empty(self.shape, dtype=fields[0].dtype, device=dev)
               for _ in range(self.neq)] if want_residual else None
        st = self._fn('jac')(c_int(dtag), *self._shape_args(),
                             *self._args_for(dtag),
                             ptrs(fields), ptr(coef),
                             ptrs(res) if res else None, _lib.stream())
        check(st, self.prefix + '_jac')
        jac = StencilJacobian(coef, ent, self.shape, self.neq,
                              len(self.fieldnames))
        return jac, res

    def loss_grad(self, fields, grads=None):
        """Fused loss and adjoint. Returns (loss, grads, loss_split) with
        loss and loss_split as float64 device tensors (scalar and (neq,))."""
        dtag = self._check_fields(fields)
        dev = fields[0].device
        if grads is None:
            grads = [torch.empty_like(f) for f in fields]
        nb = self.nblocks(1, dtag)
        partial = torch.empty((self.neq, nb), dtype=torch.float64,
                              device=dev)
        nbo = c_int(0)
        st = self._fn('loss_grad')(c_int(dtag), *self._shape_args(),
                                   *self._args_for(dtag), ptrs(fields),
                                   ptrs(grads), ptr(partial),
                                   ctypes.byref(nbo), _lib.stream())
        check(st, self.prefix + '_loss_grad')
        if nbo.value != nb:
            raise _lib.OdilError('block count mismatch {} != {}'.format(
                nbo.value, nb))
        ssq = torch.empty(self.neq, dtype=torch.float64, device=dev)
        check(_lib.lib().odil_reduce_partials(ptr(partial), c_int(self.neq),
                                              c_int(nb), ptr(ssq),
                                              _lib.stream()),
              'reduce_partials')
        loss_split = ssq / self.ncell
        return loss_split.sum(), grads, loss_split


class Cavity(Operator):
    """Steady lid-driven cavity, unknowns (u, v, p) on an (nx, ny) grid over
    the unit square; mu = 1/Re (ns_steady.py:867)."""

    fieldnames = ('u', 'v', 'p')
    neq = 3
    prefix = 'cavity'

    def __init__(self, shape, mu, bc=0, rhiechow=1, lower=(0., 0.),
                 upper=(1., 1.)):
        super().__init__(shape)
        if len(self.shape) != 2 or min(self.shape) < 3:
            raise ValueError('cavity needs a 2D grid with at least 3 cells '
                             'per axis')
        if bc not in (0, 1, 2):
            raise ValueError('Unknown bc={:}'.format(bc))
        self.mu = float(mu)
        self.bc = int(bc)
        self.rhiechow = int(rhiechow)
        self.dx = (upper[0] - lower[0]) / self.shape[0]
        self.dy = (upper[1] - lower[1]) / self.shape[1]

    def _args(self):
        return [c_double(self.dx), c_double(self.dy), c_double(self.mu),
                c_int(self.bc), c_int(self.rhiechow)]


class CavityExtras(Operator):
    """Extra loss terms of the cavity's inverse problems
    (ns_steady.py:396-411) on the cavity's unknowns: imposed points
    mask (u - u_imp) cimp, mask (v - v_imp) cimp with cimp = alpha N / n_imp
    (`imposed` = (mask, u_imp, v_imp) arrays of the grid shape, n_imp the
    nonzeros of mask), and kreg (u_xx, u_yy, v_xx, v_yy) on the
    halo-corrected stencils (`kreg` already annealed). Its adjoint is added
    to the gradients it is given; `cavity_inverse_loss_grad` combines it
    with the cavity loss in the reference's equation order."""

    fieldnames = ('u', 'v', 'p')
    neq = 6
    prefix = 'cavity_extras'

    def __init__(self, cavity, imposed=None, alpha=1., kreg=0.,
                 device=None):
        super().__init__(cavity.shape)
        self.cavity = cavity
        self.kreg = float(kreg)
        device = device or torch.device('cuda')
        self.n_imp = 0
        self._imp = {}
        if imposed is not None:
            mask, uimp, vimp = (np.asarray(a, dtype=np.float64)
                                for a in imposed)
            if any(a.shape != self.shape for a in (mask, uimp, vimp)):
                raise ValueError('imposed arrays must have the grid shape')
            self.n_imp = int(np.count_nonzero(mask))
            for dt, tag in ((torch.float64, _lib.F64),
                            (torch.float32, _lib.F32)):
                self._imp[tag] = [torch.as_tensor(a, dtype=dt,
                                                  device=device)
                                  for a in (mask, uimp, vimp)]
        self.cimp = (float(alpha) * self.ncell / self.n_imp
                     if self.n_imp else 0.)

    def enabled(self):
        """Indices of the extra equations the reference would append."""
        return ([0, 1] if self.n_imp else []) + \
            ([2, 3, 4, 5] if self.kreg else [])

    def _args_for(self, dtag):
        c = self.cavity
        imp = self._imp.get(dtag)
        return [c_double(c.dx), c_double(c.dy), c_int(c.bc),
                c_double(self.kreg), c_double(self.cimp)] + (
                    [ptr(a) for a in imp] if imp else [None, None, None])

    def _args(self):
        return self._args_for(_lib.F64)

    def loss_grad(self, fields, grads=None):
        """Adds the extra terms' adjoint to `grads` (zeros if None)."""
        if grads is None:
            grads = [torch.zeros_like(f) for f in fields]
        return super().loss_grad(fields, grads)


def cavity_inverse_loss_grad(cavity, extras, fields):
    """Loss, gradients and loss split of the cavity residuals with the
    enabled extra equations appended (ns_steady.py:394-411)."""
    loss, grads, split = cavity.loss_grad(fields)
    lx, grads, sx = extras.loss_grad(fields, grads)
    idx = torch.tensor(extras.enabled(), dtype=torch.long,
                       device=split.device)
    return loss + lx, grads, torch.cat([split, sx[idx]])


def _dev64(a, device):
    return torch.as_tensor(np.asarray(a, dtype=np.float64),
                           device=device).contiguous()


class Wave(Operator):
    """1D wave equation u_tt = u_xx on a (nt, nx) space-time grid over
    t in [lower[0], upper[0]], x in [lower[1], upper[1]] (wave.py:67-122).
    `init_u`, `init_ut` are arrays over x at t = lower[0]; `left_u`,
    `right_u` boundary values over t."""

    fieldnames = ('u', )
    neq = 1
    prefix = 'wave'

    def __init__(self, shape, init_u, init_ut, left_u, right_u,
                 lower=(0., -1.), upper=(1., 1.), device=None):
        super().__init__(shape)
        nt, nx = self.shape
        if nt < 3 or nx < 2:
            raise ValueError('wave needs nt >= 3 and nx >= 2')
        device = device or torch.device('cuda')
        self.dt = (upper[0] - lower[0]) / nt
        self.dx = (upper[1] - lower[1]) / nx
        self.bnd = [_dev64(init_u, device), _dev64(init_ut, device),
                    _dev64(left_u, device), _dev64(right_u, device)]
        for a, n in zip(self.bnd, (nx, nx, nt, nt)):
            if a.numel() != n:
                raise ValueError('boundary array of size {} != {}'.format(
                    a.numel(), n))

    def _args(self):
        return [c_double(self.dt), c_double(self.dx)] + [
            ptr(a) for a in self.bnd]


class Poisson(Operator):
    """f = lap(u) - src on an (n0, n1) grid over the unit square (or
    [lower, upper]) with Dirichlet u = 0 walls."""

    fieldnames = ('u', )
    neq = 1
    prefix = 'poisson'

    def __init__(self, shape, src, lower=(0., 0.), upper=(1., 1.)):
        super().__init__(shape)
        self.src = src
        _lib.require_device(src)
        if tuple(src.shape) != self.shape:
            raise ValueError('src shape {} != {}'.format(tuple(src.shape),
                                                         self.shape))
        self.dx = (upper[0] - lower[0]) / self.shape[0]
        self.dy = (upper[1] - lower[1]) / self.shape[1]

    def _args(self):
        return [c_double(self.dx), c_double(self.dy), ptr(self.src)]

    def _args_for(self, dtag):
        want = torch.float64 if dtag == _lib.F64 else torch.float32
        if self.src.dtype != want:
            if getattr(self, '_src_cast', None) is None or \
                    self._src_cast.dtype != want:
                self._src_cast = self.src.to(want)
            return [c_double(self.dx), c_double(self.dy),
                    ptr(self._src_cast)]
        return self._args()

    def _check_fields(self, fields):
        return super()._check_fields(fields)


class Laplace(Operator):
    """(lap u, lap v, 0) with linear-extrapolation halos
    (operator_laplace, util_op.py:1038-1079)."""

    fieldnames = ('u', 'v')
    neq = 3
    prefix = 'laplace'

    def __init__(self, shape, lower=(0., 0.), upper=(1., 1.)):
        super().__init__(shape)
        self.dx = (upper[0] - lower[0]) / self.shape[0]
        self.dy = (upper[1] - lower[1]) / self.shape[1]

    def _args(self):
        return [c_double(self.dx), c_double(self.dy)]


class Advection:
    """p_t + a . grad p = 0 on a (nt, nx, ny) space-time grid, t in
    [0, tmax], x, y in [0, 1], periodic in x and y, constant velocity
    (ax, ay), initial field `p_init` (advection_2d.py:103-223)."""

    fieldnames = ('p', )
    neq = 1

    def __init__(self, shape, ax, ay, p_init, tmax=1.):
        self.shape = tuple(int(s) for s in shape)
        self.ncell = int(np.prod(self.shape))
        nt, nx, ny = self.shape
        self.ax, self.ay = float(ax), float(ay)
        self.dt = tmax / nt
        self.dx = 1. / nx
        self.dy = 1. / ny
        _lib.require_device(p_init)
        if tuple(p_init.shape) != (nx, ny):
            raise ValueError('p_init must have shape (nx, ny)')
        self.p_init = p_init

    def _args(self, dtag):
        return [c_int(dtag)] + [c_int(s) for s in self.shape] + [
            c_double(self.dt), c_double(self.dx), c_double(self.dy),
            c_double(self.ax), c_double(self.ay), ptr(self.p_init)]

    def _check(self, fields):
        (p, ) = fields
        _lib.require_device(p)
        if tuple(p.shape) != self.shape or p.dtype != self.p_init.dtype:
            raise ValueError('field must have shape {} and dtype {}'.format(
                self.shape, self.p_init.dtype))
        return _lib.dtype_tag(p.dtype)

    def residual(self, fields, want_fields=True):
        dtag = self._check(fields)
        p = fields[0]
        res = torch.empty_like(p) if want_fields else None
        nb = _lib.lib().odil_advection_nblocks(*[c_int(s)
                                                 for s in self.shape])
        partial = torch.empty((1, nb), dtype=torch.float64, device=p.device)
        check(_lib.lib().odil_advection_residual(*self._args(dtag), ptr(p),
                                                 ptr(res), ptr(partial),
                                                 _lib.stream()),
              'advection_residual')
        ssq = torch.empty(1, dtype=torch.float64, device=p.device)
        check(_lib.lib().odil_reduce_partials(ptr(partial), c_int(1),
                                              c_int(nb), ptr(ssq),
                                              _lib.stream()),
              'reduce_partials')
        return ([res] if want_fields else None), ssq

    def loss_grad(self, fields, grads=None):
        dtag = self._check(fields)
        p = fields[0]
        g = grads[0] if grads is not None else torch.empty_like(p)
        nb = _lib.lib().odil_advection_nblocks(*[c_int(s)
                                                 for s in self.shape])
        partial = torch.empty((1, nb), dtype=torch.float64, device=p.device)
        check(_lib.lib().odil_advection_loss_grad(*self._args(dtag), ptr(p),
                                                  ptr(g), ptr(partial),
                                                  _lib.stream()),
              'advection_loss_grad')
        ssq = torch.empty(1, dtype=torch.float64, device=p.device)
        check(_lib.lib().odil_reduce_partials(ptr(partial), c_int(1),
                                              c_int(nb), ptr(ssq),
                                              _lib.stream()),
              'reduce_partials')
        split = ssq / self.ncell
        return split.sum(), [g], split



class AdvectionInverse:
    """Inverse advection problem of advection_2d.py:103-250
    (problem='inverse', central=0, implicit=0) on an (nt, nx, ny)
    space-time grid, t in [0, tmax], x, y in [0, 1]: unknowns u, v, p;
    equations [fe, fp] + [fsu, fsv] (kreg != 0) + [ftregu, ftregv]
    (ktreg != 0) + [fdiv] (divfree != 0) in the reference's order. fe is
    the misfit to `p_exact` annealed by 0.5^(epoch / anneal_half) (only at
    the first and last time rows with onlyfinal); fp is the advection
    residual with the live velocity and first-order upwinding on the sign of
    the frozen one; x, y periodic or zero outside (periodic=False). One
    kernel each for the residuals (`odil_advection_inverse_residual`) and
    the fused loss + adjoint (`odil_advection_inverse_loss_grad`)."""

    fieldnames = ('u', 'v', 'p')

    def __init__(self, shape, p_exact, alpha=1., anneal_half=0.,
                 onlyfinal=False, kreg=1e-4, ktreg=0., divfree=0.,
                 periodic=True, tmax=1.):
        self.shape = tuple(int(s) for s in shape)
        if len(self.shape) != 3 or min(self.shape[1:]) < 3:
            raise ValueError('shape must be (nt, nx, ny) with nx, ny >= 3')
        self.ncell = int(np.prod(self.shape))
        nt, nx, ny = self.shape
        self.dt, self.dx, self.dy = tmax / nt, 1. / nx, 1. / ny
        _lib.require_device(p_exact)
        if tuple(p_exact.shape) != self.shape:
            raise ValueError('p_exact must have shape {}'.format(self.shape))
        self.p_exact = p_exact
        self.alpha, self.anneal_half = float(alpha), float(anneal_half)
        self.onlyfinal, self.periodic = bool(onlyfinal), bool(periodic)
        self.kreg, self.ktreg = float(kreg), float(ktreg)
        self.divfree = float(divfree)
        self.neq = _lib.lib().odil_advection_inverse_neq(
            c_double(self.kreg), c_double(self.ktreg),
            c_double(self.divfree))

    def anneal(self, epoch):
        """get_anneal_factor (advection_2d.py:25-26)."""
        return 0.5**(epoch / self.anneal_half) if self.anneal_half else 1.

    def _args(self, dtag, epoch):
        return [c_int(dtag)] + [c_int(s) for s in self.shape] + [
            c_double(self.dt), c_double(self.dx), c_double(self.dy),
            c_double(self.alpha * self.anneal(epoch)),
            c_int(int(self.onlyfinal)), c_double(self.kreg),
            c_double(self.ktreg), c_double(self.divfree),
            c_int(int(self.periodic)), ptr(self._pex(dtag))]

    def _pex(self, dtag):
        want = torch.float64 if dtag == _lib.F64 else torch.float32
        if self.p_exact.dtype == want:
            return self.p_exact
        if getattr(self, '_pex_cast', None) is None or \
                self._pex_cast.dtype != want:
            self._pex_cast = self.p_exact.to(want)
        return self._pex_cast

    def _check(self, fields):
        if len(fields) != 3:
            raise ValueError('fields are (u, v, p)')
        dt = fields[0].dtype
        for f in fields:
            _lib.require_device(f)
            if tuple(f.shape) != self.shape or f.dtype != dt:
                raise ValueError('fields must have shape {} and one dtype'
                                 .format(self.shape))
        return _lib.dtype_tag(dt)

    def _reduce(self, partial, nb, dev):
        ssq = torch.empty(self.neq, dtype=torch.float64, device=dev)
        check(_lib.lib().odil_reduce_partials(ptr(partial), c_int(self.neq),
                                              c_int(nb), ptr(ssq),
                                              _lib.stream()),
              'reduce_partials')
        return ssq

    def _nb(self):
        return _lib.lib().odil_advection_inverse_nblocks(
            *[c_int(s) for s in self.shape])

    def residual(self, fields, epoch=0, want_fields=True):
        """(residual fields or None, per-equation sums of squares)."""
        dtag = self._check(fields)
        dev = fields[0].device
        res = [torch.empty_like(fields[0]) for _ in range(self.neq)] \
            if want_fields else None
        nb = self._nb()
        partial = torch.empty((self.neq, nb), dtype=torch.float64,
                              device=dev)
        check(_lib.lib().odil_advection_inverse_residual(
            *self._args(dtag, epoch), ptrs(fields),
            ptrs(res) if res else None, ptr(partial), _lib.stream()),
            'advection_inverse_residual')
        return res, self._reduce(partial, nb, dev)

    def loss_grad(self, fields, grads=None, epoch=0):
        """(loss, [grad u, grad v, grad p], loss_split): loss =
        sum_e mean(f_e^2), its gradient with the upwind direction held."""
        dtag = self._check(fields)
        dev = fields[0].device
        grads = grads or [torch.empty_like(f) for f in fields]
        nb = self._nb()
        partial = torch.empty((self.neq, nb), dtype=torch.float64,
                              device=dev)
        check(_lib.lib().odil_advection_inverse_loss_grad(
            *self._args(dtag, epoch), ptrs(fields), ptrs(grads),
            ptr(partial), _lib.stream()), 'advection_inverse_loss_grad')
        split = self._reduce(partial, nb, dev) / self.ncell
        return split.sum(), grads, split


class WaveInverse:
    """Inverse wave problem: unknown field u on the (nt, nx) space-time grid
    and unknown speed c(x) [nx]. Loss = mean(f_wave^2) + mean(f_data^2) +
    k_reg^2 mean((dc/dx)^2): the wave residual with c(x)^2 u_xx (c = 1 is
    wave.py:67-122), the misfit k_data (u - data) on the sensor columns
    (imposed-point form of ns_steady.py:396-411) and a Tikhonov term on the
    speed gradient. `loss_grad` is one call of
    `odil_wave_inverse_loss_grad`."""

    def __init__(self, shape, init_u, init_ut, left_u, right_u, sensors,
                 data, k_data=1.0, k_reg=1e-2, lower=(0., -1.),
                 upper=(1., 1.), device=None):
        self.shape = tuple(int(s) for s in shape)
        nt, nx = self.shape
        self.ncell = nt * nx
        device = device or torch.device('cuda')
        self.device = device
        self.dt = (upper[0] - lower[0]) / nt
        self.dx = (upper[1] - lower[1]) / nx
        self.bnd = [_dev64(a, device) for a in (init_u, init_ut, left_u,
                                                 right_u)]
        sensors = np.asarray(sensors, dtype=np.int64)
        sof = np.full(nx, -1, dtype=np.int32)
        sof[sensors] = np.arange(len(sensors), dtype=np.int32)
        self.sensor_of = torch.tensor(sof, device=device)
        self.nsensor = len(sensors)
        self._data64 = torch.as_tensor(np.asarray(data), dtype=torch.float64,
                                       device=device).contiguous()
        if tuple(self._data64.shape) != (nt, self.nsensor):
            raise ValueError('data must have shape (nt, nsensor)')
        self._data = {torch.float64: self._data64}
        self.k_data, self.k_reg = float(k_data), float(k_reg)

    def _data_for(self, dtype):
        if dtype not in self._data:
            self._data[dtype] = self._data64.to(dtype)
        return self._data[dtype]

    def loss_grad(self, u, c, grad_u=None, grad_c=None):
        """Returns (loss, grad_u, grad_c, split): loss and split (3 terms)
        as float64 device tensors."""
        _lib.require_device(u)
        _lib.require_device(c)
        nt, nx = self.shape
        if tuple(u.shape) != self.shape or c.numel() != nx \
                or c.dtype != u.dtype:
            raise ValueError('u must be {} and c ({},) of one dtype'.format(
                self.shape, nx))
        dtag = _lib.dtype_tag(u.dtype)
        gu = torch.empty_like(u) if grad_u is None else grad_u
        gc = torch.empty_like(c) if grad_c is None else grad_c
        nwork = _lib.lib().odil_wave_inverse_work(c_int(dtag), c_int(nt),
                                                  c_int(nx))
        work = torch.empty(nwork, dtype=torch.float64, device=u.device)
        split = torch.empty(3, dtype=torch.float64, device=u.device)
        st = _lib.lib().odil_wave_inverse_loss_grad(
            c_int(dtag), c_int(nt), c_int(nx), c_double(self.dt),
            c_double(self.dx), *[ptr(a) for a in self.bnd], ptr(c),
            ptr(self.sensor_of), ptr(self._data_for(u.dtype)),
            c_int(self.nsensor), c_double(self.k_data), c_double(self.k_reg),
            ptr(u), ptr(gu), ptr(gc), ptr(work), ptr(split), _lib.stream())
        check(st, 'wave_inverse_loss_grad')
        return split.sum(), gu, gc, split

    def march(self, c, dtype=torch.float64):
        """Solution of f_wave = 0 for speed c (None: c = 1)."""
        nt, nx = self.shape
        u = torch.empty(self.shape, dtype=dtype, device=self.device)
        if c is not None:
            c = c.to(dtype).contiguous()
        st = _lib.lib().odil_wave_march(
            c_int(_lib.dtype_tag(dtype)), c_int(nt), c_int(nx),
            c_double(self.dt), c_double(self.dx),
            *[ptr(a) for a in self.bnd], ptr(c), ptr(u), _lib.stream())
        check(st, 'wave_march')
        return u

    # -- packed state [u, c] for the gradient optimizers ------------------
    def pack(self, u, c):
        return torch.cat([u.reshape(-1), c.reshape(-1)])

    def unpack(self, x):
        n = self.ncell
        return x[:n].view(self.shape), x[n:]

    def make_loss_grad(self):
        """loss_grad(x, epoch) -> (loss, grad, residual_list) over the
        packed [u, c] (the optimizer interface, optimizer.py:18-28)."""

        def loss_grad(x, epoch):
            u, c = self.unpack(x)
            g = torch.empty_like(x)
            gu, gc = self.unpack(g)
            loss, _, _, split = self.loss_grad(u, c, gu, gc)
            return loss, g, [torch.sqrt(v) for v in split]

        return loss_grad
