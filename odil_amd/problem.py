"""Reference-shaped problem API: `Domain`, `State`, `Problem` with
`eval_operator`, `eval_loss_grad`, `_eval_grad`, `linearize`,
`pack_*`/`unpack_*` and `init_missing` (util_op.py:15-333, 566-974), plus
`checkpoint_save` / `checkpoint_load` and the halo extrapolation helpers
under the reference's names.

The reference's `Problem` evaluates a Python operator callback under a
TensorFlow tape. Here the operator is one of the kernel-backed operators of
`odil_amd.operators` (the same discretisations as the reference callbacks),
fields are device tensors, and every evaluation is a HIP kernel: the
residual (`odil_<op>_residual`), the fused loss adjoint
(`odil_<op>_loss_grad`), the split-shift Jacobian (`odil_<op>_jac`) and its
CSR (`odil_stencil_to_csr`). Arbitrary callbacks are not evaluated.

Multigrid field parameterisation (`Domain(multigrid=True, mg_*)`,
util_op.py:17-110, 271-333): each unknown field named in `fieldnames` is
represented by weights on the node (or cell) hierarchy `mg_nnw`, stored like
the reference as a zero-input "neural network" [0, netsize] whose bias
vector holds the weights (`state.weights[name][1]`). The operator then sees
the field `multigrid_to_field(weights)` (interpolation kernels of
`odil_amd.decomp`), and the weights' gradient is the adjoint of that map.
"""

import numpy as np
import torch

from odil_amd.checkpoint import checkpoint_load, checkpoint_save  # noqa
from odil_amd.multigrid import BlockStencil, _torch_dtype


def extrap_quadh(u0, u1, u1p):
    """Quadratic extrapolation from points 0, 1 and the half point 1.5 to
    point 2 (util_op.py:1015-1021)."""
    return (u0 - 6 * u1 + 8 * u1p) / 3


def extrap_quad(u0, u1, u2):
    """Quadratic extrapolation from points 0, 1, 2 to 3 (util_op.py:
    1024-1027)."""
    return u0 - 3 * u1 + 3 * u2


def extrap_linear(u0, u1):
    """Linear extrapolation from points 0, 1 to 2 (util_op.py:1030-1033)."""
    return 2 * u1 - u0


class Domain:
    """Uniform cell-centred grid (util_op.py:15-262): `shape` cells between
    `lower` and `upper` along the axes named `varnames`; `fieldnames` are
    the unknown fields. `dtype` may be a NumPy or torch dtype (stored as the
    torch dtype). `neuralnets` ({name: layers}) holds only the zero-input
    networks of the multigrid parameterisation: neural-network fields are
    not on this path. `multigrid`, `mg_nlvl`, `mg_factors`, `mg_axes`,
    `mg_interp`, `mg_cell` as the reference (decomposition along every
    axis; every interpolation method of the reference computes the same
    multilinear map, done here by one set of kernels)."""

    def __init__(self, ndim=None, lower=0., upper=1., shape=None,
                 varnames=('x', 'y', 'z'), fieldnames=(), neuralnets=None,
                 frozen_weights=(), frozen_fields=(), dtype=torch.float64,
                 multigrid=False, mg_nlvl=None, mg_factors=None,
                 mg_axes=None, mg_interp=None, mg_cell=False, device=None):
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
        self.dtype = _torch_dtype(dtype)
        self.device = torch.device(device) if device is not None else \
            torch.device('cuda')
        neuralnets = dict(neuralnets or {})
        for name, layers in neuralnets.items():
            if len(layers) != 2 or layers[0] != 0:
                raise ValueError(
                    "neural network '{}' with layers {}: only zero-input "
                    'networks (bias vectors) are on this path'.format(
                        name, layers))
        self.multigrid = bool(multigrid)
        self.mg_fieldnames = []
        if self.multigrid:
            from odil_amd.decomp import MultigridDecomp
            mg_axes = [True] * ndim if mg_axes is None else list(mg_axes)
            if not all(mg_axes):
                raise ValueError('multigrid decomposition along a subset of '
                                 'the axes is not provided (mg_axes={})'
                                 .format(mg_axes))
            if mg_interp not in (None, 'conv', 'manual', 'roll'):
                raise ValueError('unknown mg_interp {}'.format(mg_interp))
            nlvl_max = min(int(round(np.log2(n if mg_cell else n - 1)))
                           for n in self.shape)
            self.mg_nlvl = nlvl_max if mg_nlvl is None else min(
                int(mg_nlvl), nlvl_max)
            self.mg_cell = bool(mg_cell)
            self.mg_nnw = [
                tuple((n >> lv) if self.mg_cell else ((n - 1) >> lv) + 1
                      for n in self.shape) for lv in range(self.mg_nlvl)]
            MultigridDecomp.check_nnw(self.mg_nnw, use_axes=mg_axes,
                                      cell=self.mg_cell)
            self.mg_axes = mg_axes
            self.mg_interp = mg_interp
            self.mg_factors = mg_factors
            netsize = sum(int(np.prod(nw)) for nw in self.mg_nnw)
            for name in self.fieldnames:
                if name in neuralnets:
                    raise RuntimeError(
                        "Name collision of field and neuralnet '{}'".format(
                            name))
                neuralnets[name] = [0, netsize]
                self.mg_fieldnames.append(name)
        self.neuralnets = neuralnets
        self.frozen_weights = list(frozen_weights)
        self.frozen_fields = list(frozen_fields)
        self.aweights = [k for k in neuralnets if k not in
                         self.frozen_weights]
        self.weight_shapes = {
            name: [(layers[0], layers[1]), (layers[1], )]
            for name, layers in neuralnets.items()}

    def _axis(self, i):
        return self.varnames.index(i) if isinstance(i, str) else i

    def step(self, i):
        i = self._axis(i)
        return (self.upper[i] - self.lower[i]) / self.shape[i]

    def step_by_dim(self, i):
        return self.step(i)

    def size(self, i):
        return self.shape[self._axis(i)]

    def cell_center_1d(self, i):
        i = self._axis(i)
        h = self.step(i)
        return self.lower[i] + h * (np.arange(self.shape[i]) + 0.5)

    def cell_center_all(self):
        return np.meshgrid(*[self.cell_center_1d(i)
                             for i in range(self.ndim)], indexing='ij')

    def cell_center_by_dim(self, i):
        return self.cell_center_all()[self._axis(i)]

    def cell_index(self, i):
        i = self._axis(i)
        return np.meshgrid(*[np.arange(s) for s in self.shape],
                           indexing='ij')[i]

    def fields_size(self):
        return len(self.fieldnames) * int(np.prod(self.shape))

    def weights_size(self):
        return sum(int(np.prod(s)) for n in self.neuralnets
                   for s in self.weight_shapes[n])

    def aweights_size(self):
        return sum(int(np.prod(s)) for n in self.aweights
                   for s in self.weight_shapes[n])

    def zeros(self):
        return torch.zeros(self.shape, dtype=self.dtype, device=self.device)

    def assign_active_state(self, dst, src):
        """Copies the fields and active weights of `src` into `dst` in place
        (util_op.py:263-269)."""
        for name in self.fieldnames:
            dst.fields[name].copy_(src.fields[name].reshape(self.shape))
        for name in self.aweights:
            for a, b in zip(dst.weights[name], src.weights[name]):
                a.copy_(b.reshape(a.shape))

    # -- multigrid parameterisation (util_op.py:271-333) -------------------
    def multigrid_to_field(self, uw, mod=None, factors=None, use_axes=None,
                           method=None, cell=None):
        """Field on the finest grid from weights on `mg_nnw`."""
        from odil_amd.decomp import MultigridDecomp
        return MultigridDecomp.weights_to_field(
            uw, self.mg_nnw,
            factors=self.mg_factors if factors is None else factors,
            cell=self.mg_cell if cell is None else cell)

    def field_to_multigrid(self, u, mod=None, factors=None):
        """Weights whose field is `u` (MultigridDecomp.field_to_weights)."""
        from odil_amd.decomp import MultigridDecomp
        return MultigridDecomp.field_to_weights(
            u, self.mg_nnw, factors=self.mg_factors if factors is None
            else factors, cell=self.mg_cell)

    def state_to_field(self, key, state, mod=None):
        if key in state.weights:
            return self.multigrid_to_field(state.weights[key][1])
        return state.fields[key]

    def add_field_to_state(self, u, key, state, mod=None):
        if key in state.weights:
            state.weights[key][1] += self.field_to_multigrid(u)
        else:
            state.fields[key] += u


class State:
    """Fields (device tensors) and weights ({name: [W, b]}) by name."""

    def __init__(self, fields=None, weights=None):
        self.fields = dict(fields or {})
        self.weights = dict(weights or {})


class Problem:
    """`Problem(operator, domain, extra=None, nn_initializer='legacy')` with
    the reference's methods (util_op.py:566-974). `operator` is a
    kernel-backed operator of `odil_amd.operators` whose `fieldnames` are the
    domain's. `extra` is kept for callers that read it back
    (`problem.extra`, the reference hands it to the operator context);
    kernel operators take their data at construction. `nn_initializer`
    names the reference's initialisers; the zero-input networks of the
    multigrid parameterisation start at zero whichever is named."""

    _INITIALIZERS = ('legacy', 'glorot', 'lecun', 'he')

    def __init__(self, operator, domain, extra=None,
                 nn_initializer='legacy'):
        if tuple(operator.fieldnames) != tuple(domain.fieldnames):
            raise ValueError('operator fields {} != domain fields {}'.format(
                operator.fieldnames, domain.fieldnames))
        if tuple(operator.shape) != tuple(domain.shape):
            raise ValueError('operator shape {} != domain shape {}'.format(
                operator.shape, domain.shape))
        if nn_initializer not in self._INITIALIZERS:
            raise ValueError('Unknown initializer=' + str(nn_initializer))
        self.operator = operator
        self.domain = domain
        self.extra = extra
        self.nn_initializer = nn_initializer

    # -- state handling (util_op.py:733-857) ------------------------------
    def init_missing(self, state):
        """Zero fields (and zero multigrid weights) for names missing from
        `state` (util_op.py:812-857); host arrays are moved to the device in
        the domain dtype."""
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
        for name in d.neuralnets:
            shapes = d.weight_shapes[name]
            ww = state.weights.get(name)
            if ww is None:
                ww = [torch.zeros(s, dtype=d.dtype, device=d.device)
                      for s in shapes]
            else:
                if len(ww) != len(shapes):
                    raise ValueError(
                        'Invalid number of weights, expected {}, got {}'
                        .format(len(shapes), len(ww)))
                ww = [torch.as_tensor(
                    np.asarray(w) if not isinstance(w, torch.Tensor) else w,
                    dtype=d.dtype, device=d.device).reshape(s)
                    for w, s in zip(ww, shapes)]
            state.weights[name] = ww
        return state

    def _fields(self, state):
        """Operator inputs: the state fields, or for multigrid fields the
        field of their weights."""
        self.init_missing(state)
        d = self.domain
        return [d.multigrid_to_field(state.weights[k][1]).contiguous()
                if k in d.mg_fieldnames else state.fields[k].contiguous()
                for k in d.fieldnames]

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

    def pack_weights(self, weights, with_frozen=False):
        """Active weights flattened in order (util_op.py:751-765)."""
        d = self.domain
        names = d.neuralnets if with_frozen else d.aweights
        out = []
        for name in names:
            for i, s in enumerate(d.weight_shapes[name]):
                w = weights[name][i] if weights.get(name) is not None and \
                    weights[name][i] is not None else torch.zeros(
                        s, dtype=d.dtype, device=d.device)
                out.append(w.reshape(-1))
        if not out:
            return torch.zeros(0, dtype=d.dtype, device=d.device)
        return torch.cat(out)

    def unpack_weights(self, packed, with_frozen=False):
        d = self.domain
        names = d.neuralnets if with_frozen else d.aweights
        out, i = {}, 0
        for name in names:
            ww = []
            for s in d.weight_shapes[name]:
                n = int(np.prod(s))
                ww.append(packed[i:i + n].view(s))
                i += n
            out[name] = ww
        return out

    def pack_state(self, state, with_frozen=False):
        self.init_missing(state)
        return torch.cat([self.pack_fields(state.fields),
                          self.pack_weights(state.weights, with_frozen)])

    def unpack_state(self, packed, with_frozen=False):
        n = self.domain.fields_size()
        return State(self.unpack_fields(packed[:n]),
                     self.unpack_weights(packed[n:], with_frozen))

    # -- evaluations ------------------------------------------------------
    def eval_operator(self, state, epoch=0):
        """Residual fields, one per equation (util_op.py:968-974)."""
        res, _ = self.operator.residual(self._fields(state))
        return res

    def eval_loss_grad(self, state, epoch=0):
        """(loss, grads {field name}, wgrads {name: [dW, db]}, loss_split)
        with loss = sum_e mean(f_e^2) and grads its gradient with respect to
        each field, frozen paths excluded (util_op.py:939-966). Fields of a
        multigrid domain get zero gradients and their weights the adjoint
        of the interpolation (T^T per level)."""
        d = self.domain
        loss, grads, split = self.operator.loss_grad(self._fields(state))
        out, wout = {}, {}
        for k, g in zip(d.fieldnames, grads):
            if k in d.mg_fieldnames:
                from odil_amd.decomp import MultigridDecomp
                gw = MultigridDecomp.weights_to_field_t(
                    g, d.mg_nnw, factors=d.mg_factors, cell=d.mg_cell)
                wout[k] = [torch.zeros(d.weight_shapes[k][0], dtype=d.dtype,
                                       device=d.device), gw]
                out[k] = torch.zeros_like(g)
            else:
                out[k] = g
        wgrads = {k: wout[k] for k in d.aweights if k in wout}
        return loss, out, wgrads, list(split)

    def _eval_grad(self, state_fields, state_weights, epoch=0):
        """(ff, grads, field_desc, wgrads) of the split-shift tape Jacobian
        (util_op.py:585-617): ff the residual fields, field_desc the list of
        (field index, shift) arguments, grads[i][j] the coefficient field
        d ff_i / d (field_desc[j]) or None where equation i does not depend
        on it. Multigrid weights are not differentiated here (the reference
        builds a dense Jacobian for them): wgrads is [[] per equation]."""
        if self.domain.mg_fieldnames:
            raise ValueError('_eval_grad: the split-shift Jacobian of a '
                             'multigrid-parameterised domain is not provided')
        st = State(state_fields, state_weights)
        jac, res = self.operator.jacobian(self._fields(st),
                                          want_residual=True)
        field_desc = []
        index = {}
        for (e, k, s) in jac.entries:
            key = (int(k), tuple(int(x) for x in s))
            if key not in index:
                index[key] = len(field_desc)
                field_desc.append(key)
        grads = [[None] * len(field_desc) for _ in range(jac.neq)]
        for q, (e, k, s) in enumerate(jac.entries):
            grads[e][index[(int(k), tuple(int(x) for x in s))]] = jac.coef[q]
        return res, grads, field_desc, [[] for _ in res]

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
