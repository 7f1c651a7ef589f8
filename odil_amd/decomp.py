"""Multigrid field parameterisation (MultigridDecomp, multigrid.py:
1217-1426) on the GPU.

A field on the finest grid is represented by weights on a hierarchy of
grids nnw[0] (finest) ... nnw[-1] (coarsest), the field being the sum of
every level's weights interpolated to the finest grid. Node grids refine
n -> 2n - 1 with multilinear interpolation ([1,2,1]/2 per axis; the
prolongation kernel `odil_mg_prolong_add`); cell grids refine n -> 2n with
[1,3,3,1]/4 over a linearly extrapolated halo (`odil_interp_cell`). The
adjoints (gradients with respect to the weights) are T^T = 2^dim R with R
the transpose restriction (`odil_mg_restrict`), and `odil_interp_cell_t`.
`WeightsToField` puts the map behind torch autograd.
"""

import ctypes

import numpy as np
import torch

from odil_amd import _lib
from odil_amd._lib import c_int, check, ptr, stream
from odil_amd.multigrid import prolong_add, restrict, restriction_weights


def _grid(shape):
    return (ctypes.c_int * 3)(*([int(s) for s in shape] +
                                [1] * (3 - len(shape))))


def _as2(shape):
    """Node kernels take 2D/3D grids; a 1D field runs as (n, 1)."""
    return tuple(shape) + (1, ) if len(shape) == 1 else tuple(shape)


_RW = {}


def _rw(dim, device):
    key = (dim, str(device))
    if key not in _RW:
        _RW[key] = torch.tensor(restriction_weights('transpose', dim),
                                device=device)
    return _RW[key]


def interp_node(u, out=None, add=None):
    """Fine field 2n - 1 per axis from node field u; `add` (fine shape)
    is added (weights_to_field's `interp(res) + u`)."""
    shape = tuple(2 * s - 1 for s in u.shape)
    if out is None:
        out = torch.empty(shape, dtype=u.dtype, device=u.device)
    if add is None:
        out.zero_()
    else:
        out.copy_(add)
    s2 = _as2(shape)
    prolong_add([u.reshape(_as2(u.shape))], [out.reshape(s2)], s2)
    return out


def interp_node_t(g):
    """Adjoint: coarse (n + 1) / 2 per axis from fine gradient g."""
    s2 = _as2(g.shape)
    dim = len(s2)
    c = restrict([g.reshape(s2)], s2, _rw(dim, g.device),
                 restriction='transpose')[0]
    c = c.reshape(tuple((s + 1) // 2 for s in g.shape))
    return c * float(2**dim)


def interp_cell(u, out=None, add=None):
    """Fine field 2n per axis from cell field u (+ `add`)."""
    shape = tuple(2 * s for s in u.shape)
    if out is None:
        out = torch.empty(shape, dtype=u.dtype, device=u.device)
    acc = 0
    if add is not None:
        out.copy_(add)
        acc = 1
    u = u.contiguous()
    st = _lib.lib().odil_interp_cell(
        c_int(_lib.dtype_tag(u.dtype)), c_int(u.dim()), _grid(u.shape),
        ptr(u), ptr(out), c_int(acc), stream())
    check(st, 'interp_cell')
    return out


def interp_cell_t(g):
    shape = tuple(s // 2 for s in g.shape)
    out = torch.empty(shape, dtype=g.dtype, device=g.device)
    g = g.contiguous()
    st = _lib.lib().odil_interp_cell_t(
        c_int(_lib.dtype_tag(g.dtype)), c_int(g.dim()), _grid(shape), ptr(g),
        ptr(out), stream())
    check(st, 'interp_cell_t')
    return out


class MultigridDecomp:
    """Static methods with the reference's names and arguments (the `mod`
    and `method` arguments do not apply: one GPU implementation).

    `check_nnw` and `field_to_weights` are called by the reference's
    `Domain` (util_op.py:77,303) but not defined in its `MultigridDecomp`;
    here `check_nnw` raises ValueError unless every level refines to the
    next finer one (n -> 2n - 1 for nodes, n -> 2n for cells), and
    `field_to_weights` returns the weights that put the field on the finest
    level (divided by its factor) and zeros on the coarser ones, the exact
    right inverse of `weights_to_field`."""

    @staticmethod
    def check_nnw(nnw, use_axes=None, cell=False):
        if use_axes is not None and not all(use_axes):
            raise ValueError('decomposition along a subset of the axes is '
                             'not provided')
        for a, b in zip(nnw[:-1], nnw[1:]):
            want = tuple(2 * x if cell else 2 * x - 1 for x in b)
            if len(a) != len(b) or tuple(a) != want:
                raise ValueError('level {} does not refine to {} ({} grids '
                                 'need {})'.format(tuple(b), tuple(a),
                                                   'cell' if cell else
                                                   'node', want))

    @staticmethod
    def field_to_weights(u, nnw, mod=None, factors=None, cell=False):
        f0 = float(factors[0]) if factors else 1.
        if tuple(u.shape) != tuple(nnw[0]) and u.numel() != int(
                np.prod(nnw[0])):
            raise ValueError('field of shape {} for the finest level {}'
                             .format(tuple(u.shape), tuple(nnw[0])))
        rest = sum(int(np.prod(nw)) for nw in nnw[1:])
        return torch.cat([u.reshape(-1) / f0,
                          torch.zeros(rest, dtype=u.dtype, device=u.device)])

    @staticmethod
    def interp_node_field(u, depth=1):
        for _ in range(depth):
            u = interp_node(u)
        return u

    @staticmethod
    def interp_cell_field(u, depth=1):
        for _ in range(depth):
            u = interp_cell(u)
        return u

    @staticmethod
    def interp_field(u, depth=1, cell=False):
        return (MultigridDecomp.interp_cell_field(u, depth) if cell else
                MultigridDecomp.interp_node_field(u, depth))

    @staticmethod
    def _split(uw, nnw, factors, cell=False):
        for a, b in zip(nnw[:-1], nnw[1:]):
            want = tuple(2 * x if cell else 2 * x - 1 for x in b)
            if tuple(a) != want or len(a) != len(b):
                raise ValueError('level {} does not refine to {} ({} grids '
                                 'need {})'.format(tuple(b), tuple(a),
                                                   'cell' if cell else
                                                   'node', want))
        uu, s = [], 0
        factors = factors or [1.] * len(nnw)
        for nw, f in zip(nnw, factors):
            n = int(np.prod(nw))
            uu.append(uw[s:s + n].reshape(tuple(nw)) * float(f))
            s += n
        if s != uw.numel():
            raise ValueError('weights size {} != {}'.format(uw.numel(), s))
        return uu

    @staticmethod
    def weights_to_fields(uw, nnw, factors=None, cell=False):
        """Each level's weights interpolated to the finest grid."""
        return [MultigridDecomp.interp_field(u, i, cell) for i, u in
                enumerate(MultigridDecomp._split(uw, nnw, factors, cell))]

    @staticmethod
    def weights_to_field(uw, nnw, factors=None, cell=False):
        """Sum of the interpolated levels (coarsest first)."""
        res = None
        step = interp_cell if cell else interp_node
        for u in reversed(MultigridDecomp._split(uw, nnw, factors, cell)):
            res = u.clone() if res is None else step(res, add=u)
        return res

    @staticmethod
    def weights_to_field_t(g, nnw, factors=None, cell=False):
        """Gradient with respect to the weights of <weights_to_field, g>."""
        step_t = interp_cell_t if cell else interp_node_t
        factors = factors or [1.] * len(nnw)
        out = []
        for lv in range(len(nnw)):
            if lv:
                g = step_t(g)
            out.append(g.reshape(-1) * float(factors[lv]))
        return torch.cat(out)

    @staticmethod
    def get_weights_coarse_ones(nnw, dtype, device=None):
        res = [torch.zeros(int(np.prod(nw)), dtype=dtype, device=device)
               for nw in nnw]
        res[-1].fill_(1)
        return torch.cat(res)


class WeightsToField(torch.autograd.Function):
    """field = weights_to_field(uw, nnw, factors, cell) with the adjoint
    kernels as backward."""

    @staticmethod
    def forward(ctx, uw, nnw, factors=None, cell=False):
        ctx.nnw, ctx.factors, ctx.cell = nnw, factors, cell
        return MultigridDecomp.weights_to_field(uw.detach(), nnw, factors,
                                                cell)

    @staticmethod
    def backward(ctx, g):
        return (MultigridDecomp.weights_to_field_t(g.contiguous(), ctx.nnw,
                                                   ctx.factors, ctx.cell),
                None, None, None)
