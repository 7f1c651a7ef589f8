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
                                  *self._args(), ptrs(fields),
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
        st = self._fn('jac')(c_int(dtag), *self._shape_args(), *self._args(),
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
                                   *self._args(), ptrs(fields), ptrs(grads),
                                   ptr(partial), ctypes.byref(nbo),
                                   _lib.stream())
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
