"""CPU oracle: discrete-loss core of ODIL restated with torch-CPU autograd.

TEST INFRASTRUCTURE ONLY. Nothing in the shipped package imports this
module; only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg
of `bench.py` use it, and only as the checker.

The reference evaluates operators under a TensorFlow gradient tape; here the
same tape semantics are restated with torch autograd on CPU in float64:

* `Context.field(name, *shift, freeze)` returns the field rolled by `-shift`
  (periodic, `util_op.py:541-563`); `freeze=True` is `detach()`, the
  equivalent of `tf.stop_gradient` (`util_op.py:541-543`).
* With `split_shifts=True` every distinct (field, shift) becomes its own tape
  argument (`util_op.py:545-554`); the gradient of `sum(f_i)` with respect to
  that argument is the stencil-Jacobian coefficient field
  (`Problem._eval_grad0`, `util_op.py:585-614`).
* With `split_shifts=False` one argument per field is used and the loss is
  `sum_i mean(f_i**2)` (`Problem._eval_loss`, `util_op.py:627-654`).
* `linearize` assembles the CSR Jacobian with the reference's row/column
  order and periodic column roll (`util_op.py:859-937`).
"""

import numpy as np
import scipy.sparse
import torch


class TorchMod:
    """The `mod` object handed to operators (the reference passes `tf`)."""

    @staticmethod
    def where(cond, a, b):
        cond = torch.as_tensor(cond)
        a = torch.as_tensor(a, dtype=torch.float64)
        b = torch.as_tensor(b, dtype=torch.float64)
        return torch.where(cond, a, b)

    @staticmethod
    def roll(u, shift, axis):
        u = torch.as_tensor(u)
        if isinstance(shift, (list, tuple)):
            return torch.roll(u, tuple(int(s) for s in shift), tuple(axis))
        return torch.roll(u, int(shift), int(axis))

    @staticmethod
    def abs(u):
        return torch.abs(u)

    @staticmethod
    def sqrt(u):
        return torch.sqrt(u)

    @staticmethod
    def tanh(u):
        return torch.tanh(u)


class Domain:
    """Cell-centred grid geometry (`util_op.py:15-262`)."""

    def __init__(self, ndim=None, lower=0., upper=1., shape=None,
                 varnames=('x', 'y', 'z'), fieldnames=(), dtype=np.float64):
        if ndim is None:
            ndim = len(varnames)
        self.ndim = ndim
        self.shape = tuple(int(s) for s in shape)
        self.lower = np.ones(ndim, dtype=np.float64) * lower
        self.upper = np.ones(ndim, dtype=np.float64) * upper
        self.varnames = tuple(varnames[:ndim])
        self.fieldnames = list(fieldnames)
        self.dtype = dtype

    def _dim(self, i):
        return self.varnames.index(i) if isinstance(i, str) else i

    def step(self, i):
        i = self._dim(i)
        return (self.upper[i] - self.lower[i]) / self.shape[i]

    def size(self, i):
        return self.shape[self._dim(i)]

    def cell_center_1d(self, i):
        i = self._dim(i)
        x = np.linspace(self.lower[i], self.upper[i], self.shape[i],
                        endpoint=False, dtype=np.float64)
        x += (x[1] - x[0]) * 0.5
        return x

    def cell_center_all(self):
        xx = [self.cell_center_1d(i) for i in range(self.ndim)]
        return np.meshgrid(*xx, indexing='ij')

    def cell_center(self, i):
        return self.cell_center_all()[self._dim(i)]

    def cell_index(self, i):
        ii = np.meshgrid(*[np.arange(s) for s in self.shape], indexing='ij')
        return ii[self._dim(i)]

    def fields_size(self):
        return len(self.fieldnames) * int(np.prod(self.shape))


class Context:
    """Restatement of `util_op.Context` (`util_op.py:432-563`)."""

    def __init__(self, domain, state_fields, split_shifts, epoch=0,
                 extra=None):
        self.domain = domain
        self.state_fields = state_fields  # name -> torch tensor (float64)
        self.split_shifts = split_shifts
        self.args = []
        self.field_desc = []
        self.epoch = epoch
        self.extra = extra

    def step(self, i):
        return self.domain.step(i)

    def size(self, i):
        return self.domain.size(i)

    def cell_index(self, i):
        return torch.as_tensor(self.domain.cell_index(i))

    def cell_center(self, i):
        return torch.as_tensor(self.domain.cell_center(i))

    def field(self, name, *shift, freeze=False):
        d = self.domain
        shift = tuple(shift) + (0, ) * (d.ndim - len(shift))
        if name == 'zeros':
            return torch.zeros(d.shape, dtype=torch.float64)
        if name == 'ones':
            return torch.ones(d.shape, dtype=torch.float64)
        try:
            j = d.fieldnames.index(name)
        except ValueError:
            raise ValueError("Unknown field name '" + name + "'")
        axes = tuple(range(d.ndim))
        neg = tuple(-s for s in shift)
        if freeze:
            return torch.roll(self.state_fields[name].detach(), neg, axes)
        if self.split_shifts:
            if (j, shift) in self.field_desc:
                return self.args[self.field_desc.index((j, shift))]
            u = torch.roll(self.state_fields[name].detach(), neg,
                           axes).clone().requires_grad_(True)
            self.args.append(u)
            self.field_desc.append((j, shift))
            return u
        szero = (0, ) * d.ndim
        if (j, szero) in self.field_desc:
            uc = self.args[self.field_desc.index((j, szero))]
        else:
            uc = self.state_fields[name].detach().clone().requires_grad_(
                True)
            self.args.append(uc)
            self.field_desc.append((j, szero))
        return torch.roll(uc, neg, axes)


def _to_torch(fields, domain):
    return {
        k: torch.as_tensor(np.asarray(v, dtype=np.float64)).reshape(
            domain.shape)
        for k, v in fields.items()
    }


def eval_operator(operator, domain, fields, epoch=0, extra=None):
    """Forward residual, `Problem._eval_operator` (`util_op.py:656-673`)."""
    ctx = Context(domain, _to_torch(fields, domain), split_shifts=False,
                  epoch=epoch, extra=extra)
    with torch.no_grad():
        ff = operator(TorchMod, ctx)
    if not isinstance(ff, (list, tuple)):
        ff = (ff, )
    return [f.detach().numpy().copy() for f in ff]


def eval_grad(operator, domain, fields, epoch=0, extra=None):
    """Stencil-Jacobian coefficient fields, `Problem._eval_grad0`
    (`util_op.py:585-614`).

    Returns (ff, grads, field_desc) with `grads[i][j]` the coefficient field of
    equation i for tape argument j (None when unused).
    """
    ctx = Context(domain, _to_torch(fields, domain), split_shifts=True,
                  epoch=epoch, extra=extra)
    ff = operator(TorchMod, ctx)
    if not isinstance(ff, (list, tuple)):
        ff = (ff, )
    grads = []
    for f in ff:
        if f.requires_grad and ctx.args:
            g = torch.autograd.grad(f.sum(), ctx.args, allow_unused=True,
                                    retain_graph=True)
        else:
            g = [None] * len(ctx.args)
        grads.append([None if x is None else x.numpy().copy() for x in g])
    ff = [f.detach().numpy().copy() for f in ff]
    return ff, grads, list(ctx.field_desc)


def eval_loss_grad(operator, domain, fields, epoch=0, extra=None):
    """Loss and its gradient, `Problem._eval_loss` + `eval_loss_grad`
    (`util_op.py:627-654,939-966`).

    Returns (loss, grads{name: ndarray}, loss_split).
    """
    ctx = Context(domain, _to_torch(fields, domain), split_shifts=False,
                  epoch=epoch, extra=extra)
    ff = operator(TorchMod, ctx)
    if not isinstance(ff, (list, tuple)):
        ff = (ff, )
    loss_split = [torch.mean(f**2) for f in ff]
    loss = sum(loss_split)
    if ctx.args and loss.requires_grad:
        g = torch.autograd.grad(loss, ctx.args, allow_unused=True)
    else:
        g = [None] * len(ctx.args)
    grads = {}
    for gi, (j, shift) in zip(g, ctx.field_desc):
        assert shift == (0, ) * domain.ndim
        if gi is not None:
            grads[domain.fieldnames[j]] = gi.numpy().copy()
    for name in domain.fieldnames:
        if name not in grads:
            grads[name] = np.zeros(domain.shape)
    return (float(loss.detach()), grads,
            [float(v.detach()) for v in loss_split])


def linearize(operator, domain, fields, epoch=0, extra=None):
    """CSR Jacobian and residuals, `Problem.linearize` (`util_op.py:859-937`).

    Rows are `row + i*N` (equation-major), columns
    `roll(col_diag, -shift) + index*N` with periodic wrap (`:877-894`).
    """
    const, grads, field_desc = eval_grad(operator, domain, fields, epoch,
                                         extra)
    Nu = len(domain.fieldnames)
    N = int(np.prod(domain.shape))
    Ne = sum(int(np.prod(c.shape)) for c in const)
    mshape = (Ne, Nu * N)
    linear = scipy.sparse.csr_matrix(mshape, dtype=np.float64)
    row = np.arange(N)
    col_diag = row.reshape(domain.shape)
    for i in range(len(const)):
        for j, (index, shift) in enumerate(field_desc):
            col = np.roll(col_diag, shift=np.negative(shift),
                          axis=range(domain.ndim)).flatten()
            g = grads[i][j]
            if g is not None:
                a = scipy.sparse.csr_matrix(
                    (g.flatten(), (row + i * N, col + index * N)),
                    dtype=np.float64, shape=mshape)
                linear += a
    return const, linear


def pack_fields(domain, fields):
    """Flat layout: concatenation over `fieldnames` of C-order arrays
    (`util_op.py:733-739`)."""
    return np.concatenate(
        [np.asarray(fields[n], dtype=np.float64).ravel()
         for n in domain.fieldnames])


def unpack_fields(domain, packed):
    """Inverse of `pack_fields` (`util_op.py:741-749`)."""
    N = int(np.prod(domain.shape))
    return {
        n: np.asarray(packed[i * N:(i + 1) * N]).reshape(domain.shape)
        for i, n in enumerate(domain.fieldnames)
    }
