"""Gradient optimizers with the reference's interface (optimizer.py:8-338):
`make_optimizer(name, dtype, **kw).run(x0, loss_grad, epochs, callback,
epoch_start, lr, ...) -> (x, info)` with
`loss_grad(x, epoch) -> (loss, grad, residual_list)`.

`x` and `grad` are flat device tensors (the packed state). The vector work
runs in HIP kernels of libodil_amd.so: the fused Adam update
(`odil_adam_step`), trial points (`odil_waxpy`), and the L-BFGS two-loop
recursion on device-resident scalars (`odil_dot`, `odil_axpy_ratio`), so an
L-BFGS iteration synchronises with the host only for the line-search
values. `lbfgsb` keeps the reference's SciPy L-BFGS-B on the host around the
device loss and gradient, as the reference does.
"""

import ctypes
import math
from argparse import Namespace

import numpy as np
import torch

from odil_amd import _lib
from odil_amd._lib import c_double, c_int, check, ptr, stream
from odil_amd.newton import DeviceScalars, VectorOps


def ctypes_ll(n):
    return ctypes.c_longlong(int(n))


def _waxpy(out, x, y, a):
    """out = x + a y (flat device tensors)."""
    check(_lib.lib().odil_waxpy(c_int(_lib.dtype_tag(x.dtype)),
                                ctypes_ll(x.numel()), ptr(out), ptr(x),
                                ptr(y), c_double(a), stream()), 'waxpy')
    return out


def _host(v):
    return v.item() if isinstance(v, torch.Tensor) else float(v)


class Optimizer:
    """Base class (optimizer.py:8-28)."""

    def __init__(self, name=None, displayname=None, dtype=None):
        self.name = name
        self.displayname = displayname if displayname is not None else name
        self.dtype = dtype
        self.last_loss = None
        self.last_residual = None
        self.evals = 0

    def _record(self, loss, residual):
        self.evals += 1
        self.last_loss = _host(loss)
        self.last_residual = [_host(r) for r in residual]

    def run(self, x0, loss_grad, epochs, callback=None, epoch_start=0,
            **kwargs):
        info = Namespace()
        info.evals = 0
        info.epochs = 0
        return x0, info


class GdOptimizer(Optimizer):
    """x -= lr grad (optimizer.py:241-268)."""

    def __init__(self, dtype=None, **kwargs):
        super().__init__(name='gd', displayname='GD', dtype=dtype)

    def run(self, x0, loss_grad, epochs=None, callback=None, lr=1e-3,
            epoch_start=0, **kwargs):
        x = x0.clone()
        for epoch in range(epoch_start + 1, epoch_start + epochs + 1):
            loss, grad, res = loss_grad(x, epoch)
            _waxpy(x, x, grad, -lr)
            self._record(loss, res)
            if epoch > 0 and callback is not None:
                callback(x, epoch=epoch, opt=self)
        return x, Namespace(epochs=epochs, evals=self.evals)


class AdamNativeOptimizer(Optimizer):
    """Adam with the update of optimizer.py:271-322 (Keras 2.12 form),
    one fused kernel per step."""

    def __init__(self, dtype=None, **kwargs):
        super().__init__(name='adamn', displayname='AdamNative', dtype=dtype)

    def run(self, x0, loss_grad, epochs=None, callback=None, lr=1e-3,
            epoch_start=0, beta_1=0.9, beta_2=0.999, epsilon=1e-7,
            **kwargs):
        x = x0.clone()
        m = torch.zeros_like(x)
        v = torch.zeros_like(x)
        dt = c_int(_lib.dtype_tag(x.dtype))
        for epoch in range(epoch_start + 1, epoch_start + epochs + 1):
            loss, grad, res = loss_grad(x, epoch)
            t = epoch - epoch_start
            alpha = lr * math.sqrt(1 - beta_2**t) / (1 - beta_1**t)
            check(_lib.lib().odil_adam_step(
                dt, ctypes_ll(x.numel()), ptr(x), ptr(grad), ptr(m), ptr(v),
                c_double(alpha), c_double(beta_1), c_double(beta_2),
                c_double(epsilon), stream()), 'adam_step')
            self._record(loss, res)
            if epoch > 0 and callback is not None:
                callback(x, epoch=epoch, opt=self)
        return x, Namespace(epochs=epochs, evals=self.evals)


class AdamOptimizer(AdamNativeOptimizer):
    """`adam` (Keras Adam in the reference, optimizer.py:183-238): the same
    update rule as AdamNative."""

    def __init__(self, dtype=None, **kwargs):
        super().__init__(dtype=dtype)
        self.name, self.displayname = 'adam', 'Adam'


class LbfgsOptimizer(Optimizer):
    """Limited-memory BFGS on the device (`lbfgs`, optimizer.py:102-180,
    which calls tfp.optimizer.lbfgs_minimize with m correction pairs and
    maxls line-search evaluations and no tolerance stop).

    Direction: two-loop recursion with H0 = (s.y / y.y) I over the last m
    pairs; all inner products stay on the device (FP64 accumulation).
    Step: strong-Wolfe line search (c1 = 1e-4, c2 = 0.9) with cubic
    interpolation (Nocedal & Wright, Algorithms 3.5-3.6); pairs with
    s.y <= 0 are skipped. The line search differs from TFP's Hager-Zhang
    search, so iterates are not expected to match the reference's."""

    def __init__(self, m=50, maxls=50, dtype=None, c1=1e-4, c2=0.9,
                 **kwargs):
        super().__init__(name='lbfgs', displayname='L-BFGS', dtype=dtype)
        self.m = int(m)
        self.maxls = int(maxls)
        self.c1, self.c2 = c1, c2

    # -- two-loop recursion on device scalars --------------------------
    def _direction(self, g, r, q, hist, sc, vec):
        """r = H g; returns r (the search direction is -r)."""
        q.copy_(g)
        n = len(hist)
        for j in range(n - 1, -1, -1):
            s, y, slot = hist[j]
            vec.dot([s], [q], sc.p(slot + 1))           # a_j = s.q
            vec.axpy([q], [y], sc.p(slot + 1), sc.p(slot), -1.)
        r.zero_()
        if n:
            _, _, slot = hist[-1]
            vec.axpy([r], [q], sc.p(slot), sc.p(slot + 2), 1.)  # (s.y/y.y) q
        else:
            r.copy_(q)
        for j in range(n):
            s, y, slot = hist[j]
            vec.dot([y], [r], sc.p(self._tmp))           # b = y.r
            vec.axpy([r], [s], sc.p(slot + 1), sc.p(slot), 1.)
            vec.axpy([r], [s], sc.p(self._tmp), sc.p(slot), -1.)
        return r

    def _line_search(self, fg, x, d, f0, g0d, t0):
        """Strong-Wolfe search along d from x; fg(t) -> (f, g.d, grad,
        loss, residual). Returns (t, f, grad, loss, residual) of the
        accepted point or None."""
        c1, c2 = self.c1, self.c2
        evals = 0

        def cubic(a, fa, da, b, fb, db):
            d1 = da + db - 3 * (fa - fb) / (a - b)
            s = d1 * d1 - da * db
            if s < 0:
                return 0.5 * (a + b)
            d2 = math.copysign(math.sqrt(s), b - a)
            t = b - (b - a) * (db + d2 - d1) / (db - da + 2 * d2)
            lo, hi = min(a, b), max(a, b)
            if not (lo + 0.1 * (hi - lo) <= t <= hi - 0.1 * (hi - lo)):
                t = 0.5 * (a + b)
            return t

        def zoom(lo, flo, dlo, hi, fhi, dhi):
            nonlocal evals
            while evals < self.maxls:
                t = cubic(lo, flo, dlo, hi, fhi, dhi)
                f, dd, gr, ls, rs = fg(t)
                evals += 1
                if f > f0 + c1 * t * g0d or f >= flo:
                    hi, fhi, dhi = t, f, dd
                else:
                    if abs(dd) <= -c2 * g0d:
                        return t, f, gr, ls, rs
                    if dd * (hi - lo) >= 0:
                        hi, fhi, dhi = lo, flo, dlo
                    lo, flo, dlo = t, f, dd
            return None

        tprev, fprev, dprev = 0., f0, g0d
        t = t0
        while evals < self.maxls:
            f, dd, gr, ls, rs = fg(t)
            evals += 1
            if not math.isfinite(f):
                t = 0.5 * (tprev + t)
                continue
            if f > f0 + c1 * t * g0d or (evals > 1 and f >= fprev):
                return zoom(tprev, fprev, dprev, t, f, dd)
            if abs(dd) <= -c2 * g0d:
                return t, f, gr, ls, rs
            if dd >= 0:
                return zoom(t, f, dd, tprev, fprev, dprev)
            tprev, fprev, dprev = t, f, dd
            t = 2 * t
        return None

    def run(self, x0, loss_grad, epochs=None, callback=None, epoch_start=0,
            **kwargs):
        dev = x0.device
        x = x0.clone()
        vec = VectorOps(dev)
        # Scalar slots: per history entry 3 (s.y, a = s.q, y.y), 1 temp,
        # 2 for line-search dots.
        sc = DeviceScalars(3 * self.m + 4, dev)
        self._tmp = 3 * self.m
        q = torch.empty_like(x)
        r = torch.empty_like(x)
        xt = torch.empty_like(x)
        pool = [(torch.empty_like(x), torch.empty_like(x))
                for _ in range(self.m)]
        hist = []      # [(s, y, slot)], oldest first
        free = list(range(self.m))
        self.epoch = epoch_start
        loss, g, res = loss_grad(x, self.epoch)
        self._record(loss, res)
        f = _host(loss)
        g = g.clone()
        dslot = sc.p(self._tmp + 1)

        def gdot(a, b):
            vec.dot([a], [b], dslot)
            return sc.t[self._tmp + 1].item()

        for it in range(epochs):
            self._direction(g, r, q, hist, sc, vec)
            g0d = -gdot(g, r)              # direction d = -r
            if not (g0d < 0):
                hist, free = [], list(range(self.m))
                r.copy_(g)
                g0d = -gdot(g, g)
                if g0d == 0:
                    break
            t0 = 1.0 if hist else min(1.0, 1.0 / max(1e-300,
                                                      math.sqrt(-g0d)))

            def fg(t):
                _waxpy(xt, x, r, -t)
                ls, gr, rs = loss_grad(xt, self.epoch + 1)
                return (_host(ls), -gdot(gr, r), gr, ls, rs)

            out = self._line_search(fg, x, r, f, g0d, t0)
            if out is None:
                break
            t, fn, gn, ls, rs = out
            self._record(ls, rs)
            # The accepted point is the last one evaluated (xt, gn).
            # New pair s = xt - x, y = gn - g (oldest pair recycled).
            slot_i = free.pop(0) if free else hist.pop(0)[2] // 3
            s, y = pool[slot_i]
            base = 3 * slot_i
            _waxpy(s, xt, x, -1.)
            _waxpy(y, gn, g, -1.)
            x.copy_(xt)
            g.copy_(gn)
            f = fn
            vec.dot([s], [y], sc.p(base))
            vec.dot([y], [y], sc.p(base + 2))
            if sc.t[base].item() > 1e-30 * max(1e-300, sc.t[base + 2].item()):
                hist.append((s, y, base))
            else:
                free.append(slot_i)
            self.epoch += 1
            if callback is not None:
                callback(x, epoch=self.epoch, opt=self)
        return x, Namespace(epochs=self.epoch - epoch_start,
                            evals=self.evals)


class LbfgsbOptimizer(Optimizer):
    """SciPy L-BFGS-B on the host around the device loss and gradient
    (optimizer.py:31-99, same options and defaults)."""

    def __init__(self, pgtol=1e-16, m=50, maxls=50, factr=0, dtype=None,
                 **kwargs):
        super().__init__(name='lbfgsb', displayname='L-BFGS-B', dtype=dtype)
        self.pgtol, self.m, self.maxls, self.factr = pgtol, m, maxls, factr

    def run(self, x0, loss_grad, epochs=None, callback=None, epoch_start=0,
            **kwargs):
        import scipy.optimize
        self.epoch = epoch_start
        dev, dt = x0.device, x0.dtype

        def callback_wrap(x):
            self.epoch += 1
            if callback:
                callback(torch.as_tensor(x, dtype=dt, device=dev),
                         self.epoch, opt=self)

        def func_wrap(x):
            xt = torch.as_tensor(x, dtype=dt, device=dev)
            loss, grad, res = loss_grad(xt, self.epoch)
            self._record(loss, res)
            return (np.float64(self.last_loss),
                    grad.double().cpu().numpy())

        x, f, sinfo = scipy.optimize.fmin_l_bfgs_b(
            func=func_wrap, x0=x0.double().cpu().numpy(), maxiter=epochs,
            pgtol=self.pgtol, m=self.m, maxls=self.maxls, factr=self.factr,
            maxfun=math.inf, callback=callback_wrap)
        info = Namespace(epochs=sinfo['nit'], evals=sinfo['funcalls'],
                         task=sinfo['task'])
        return torch.as_tensor(x, dtype=dt, device=dev), info


def make_optimizer(name, dtype=None, **kwargs):
    """optimizer.py:325-338."""
    if name == 'lbfgsb':
        return LbfgsbOptimizer(dtype=dtype, **kwargs)
    if name == 'lbfgs':
        return LbfgsOptimizer(dtype=dtype, **kwargs)
    if name == 'adam':
        return AdamOptimizer(dtype=dtype, **kwargs)
    if name == 'adamn':
        return AdamNativeOptimizer(dtype=dtype, **kwargs)
    if name == 'gd':
        return GdOptimizer(dtype=dtype, **kwargs)
    raise ValueError("Unknown optimizer '{}'".format(name))
