"""Reference-shaped problem API: `Domain`, `State`, `Problem` with
`eval_operator`, `eval_loss_grad`, `linearize`, `pack_*`/`unpack_*` and
`init_missing` (util_op.py:15-333, 566-974).

The reference's `Problem` evaluates a Python operator callback under a
TensorFlow tape. Here the operator is one of the kernel-backed operators of
`odil_amd.operators` (the same discretisations as the reference callbacks),
fields are device tensors, and every evaluation is a HIP kernel: the
residual (`odil_<op>_residual`), the fused loss adjoint
(`odil_<op>_loss_grad`), the split-shift Jacobian (`odil_<op>_jac`) and its
CSR (`odil_stencil_to_csr`).
"""

import numpy as np
import torch

from odil_amd.multigrid import BlockStencil


class Domain:
    """Uniform cell-centred grid (util_op.py:15-230): `shape` cells between
    `lower` and `upper` along the axes named `varnames`; `fieldnames` are
    the unknown fields."""

    def __init__(self, ndim=None, lower=0., upper=1., shape=None,
                 varnames=('x', 'y', 'z'), fieldnames=(),
                 dtype=torch.float64, device=None):
        if ndim is None:
            ndim = len(shape) if shape is not None else len(varnames)
        if shape is not None and len(shape) != ndim:
            raise ValueError('shape {} does not have {} axes'.format(
                shape, ndim))
        self.ndim = ndim
        self.lower = np.ones(ndim) * lower
        self.upper = np.ones(ndim) * upper
        self.varnames = tuple(varnames)[:ndim]
        self.shape = tuple(int(n) for n in shape) if shape is not None \
            else None
        self.fieldnames = list(fieldnames)
        self.dtype = dtype
        self.device = torch.device(device) if device is not None else \
            torch.device('cuda')

    def _axis(self, i):
        return self.varnames.index(i) if isinstance(i, str) else i

    def step(self, i):
        i = self._axis(i)
        return (self.upper[i] - self.lower[i]) / self.shape[i]

    def size(self, i):
        return self.shape[self._axis(i)]

    def cell_center_1d(self, i):
        i = self._axis(i)
        h = self.step(i)
        return self.lower[i] + h * (np.arange(self.shape[i]) + 0.5)

    def cell_center_all(self):
        return np.meshgrid(*[self.cell_center_1d(i)
                             for i in range(self.ndim)], indexing='ij')

    def fields_size(self):
        return len(self.fieldnames) * int(np.prod(self.shape))

    def zeros(self):
        return torch.zeros(self.shape, dtype=self.dtype, device=self.device)


class State:
    """Fields by name (device tensors); weights are unused here (no neural
    networks on this path)."""

    def __init__(self, fields=None, weights=None):
        self.fields = dict(fields or {})
        self.weights = dict(weights or {})


class Problem:
    """`Problem(operator, domain)` with the reference's methods
    (util_op.py:566-974). `operator` is an `odil_amd.operators` operator
    whose `fieldnames` are the domain's."""

    def __init__(self, operator, domain):
        if tuple(operator.fieldnames) != tuple(domain.fieldnames):
            raise ValueError('operator fields {} != domain fields {}'.format(
                operator.fieldnames, domain.fieldnames))
        if tuple(operator.shape) != tuple(domain.shape):
            raise ValueError('operator shape {} != domain shape {}'.format(
                operator.shape, domain.shape))
        self.operator = operator
        self.domain = domain

    # -- state handling (util_op.py:733-857) ------------------------------
    def init_missing(self, state):
        """Zero fields for names missing from `state` (util_op.py:812-815);
        host arrays are moved to the device in the domain dtype."""
        d = self.domain
        for name in d.fieldnames:
            f = state.fields.get(name)
            if f is None:
                state.fields[name] = d.zeros()
            elif not isinstance(f, torch.Tensor) or f.device != d.device \
                    or f.dtype != d.dtype:
                state.fields[name] = torch.as_tensor(
                    np.asarray(f) if not isinstance(f, torch.Tensor) else f,
                    dtype=d.dtype, device=d.device).reshape(d.shape)
        return state

    def _fields(self, state):
        self.init_missing(state)
        return [state.fields[k].contiguous() for k in self.domain.fieldnames]

    def pack_fields(self, fields):
        """Concatenation of the C-order flattened fields in `fieldnames`
        order (util_op.py:733-739)."""
        return torch.cat([fields[k].reshape(-1)
                          for k in self.domain.fieldnames])

    def unpack_fields(self, packed):
        """Views of `packed` as fields (util_op.py:741-749)."""
        N = int(np.prod(self.domain.shape))
        return {k: packed[i * N:(i + 1) * N].view(self.domain.shape)
                for i, k in enumerate(self.domain.fieldnames)}

    def pack_state(self, state, with_frozen=False):
        return self.pack_fields(self._state_fields(state))

    def unpack_state(self, packed, with_frozen=False):
        return State(self.unpack_fields(packed))

    def _state_fields(self, state):
        self.init_missing(state)
        return state.fields

    # -- evaluations ------------------------------------------------------
    def eval_operator(self, state, epoch=0):
        """Residual fields, one per equation (util_op.py:968-974)."""
        res, _ = self.operator.residual(self._fields(state))
        return res

    def eval_loss_grad(self, state, epoch=0):
        """(loss, grads {field name}, wgrads {}, loss_split) with
        loss = sum_e mean(f_e^2) and grads its gradient with respect to each
        field, frozen paths excluded (util_op.py:939-966)."""
        loss, grads, split = self.operator.loss_grad(self._fields(state))
        return (loss, dict(zip(self.domain.fieldnames, grads)), {},
                list(split))

    def jacobian(self, state, epoch=0, want_residual=True):
        """Stencil Jacobian (split-shift tape gradient, util_op.py:585-617)
        as a `BlockStencil` and the residual fields."""
        jac, res = self.operator.jacobian(self._fields(state),
                                          want_residual=want_residual)
        return BlockStencil.from_jacobian(jac), res

    def linearize(self, state, epoch=0, mod=np, modsp=None, device=False):
        """(const, J): residual fields and the Jacobian as a sparse matrix
        in the layout of util_op.py:859-937 (row e N + c, column
        k N + flat((c + s) mod shape)). device=False returns NumPy arrays and
        a SciPy CSR (explicit zeros removed, columns sorted, as the
        reference's sum of CSR matrices); device=True returns the residual
        tensors and a torch sparse CSR tensor on the device (structural
        zeros kept)."""
        J, res = self.jacobian(state, epoch)
        rowptr, col, val = J.to_csr_arrays()
        N = int(np.prod(self.domain.shape))
        shape = (J.nvar_out * N, J.nvar_in * N)
        if device:
            m = torch.sparse_csr_tensor(rowptr.long(), col.long(), val,
                                        size=shape)
            return res, m
        import scipy.sparse
        m = scipy.sparse.csr_matrix(
            (val.cpu().numpy(), col.cpu().numpy(), rowptr.cpu().numpy()),
            shape=shape)
        m.sum_duplicates()
        m.eliminate_zeros()
        m.sort_indices()
        return [r.cpu().numpy() for r in res], m
