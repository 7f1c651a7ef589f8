"""CPU restatement of MultigridDecomp (multigrid.py:1217-1426), float64.

Test infrastructure only: imported by tests/ (never by the product).

interp_node_conv / interp_cell_conv follow the reference's 'conv' method
(transposed convolution with stride 2 and weights [1,2,1]/2 resp.
[1,3,3,1]/4 per axis, multigrid.py:1262-1279 and 1337-1354), here as one
1D transposed convolution per axis (the tensor-product weights separate).
interp_node_manual / interp_cell_manual follow the 'manual' method
(averages of rolled copies interleaved by batch_to_space,
multigrid.py:1231-1261 and 1316-1336). The two forms are independent
restatements of the same map and pin each other (tests/test_oracle.py),
together with the exactness of both interpolations on linear fields.
"""

import numpy as np


def _conv_t_axis(u, w, axis):
    """out[2 i + k] += u[i] w[k] along `axis` (stride-2 transposed conv)."""
    u = np.moveaxis(u, axis, 0)
    n = u.shape[0]
    out = np.zeros((2 * (n - 1) + len(w), ) + u.shape[1:])
    for k, wk in enumerate(w):
        out[k:k + 2 * n:2] += wk * u
    return np.moveaxis(out, 0, axis)


def pad_linear(u):
    """One halo cell per side: 2 * SYMMETRIC pad - REFLECT pad
    (multigrid.py:1307-1310)."""
    us = np.pad(u, 1, mode='symmetric')
    ur = np.pad(u, 1, mode='reflect')
    return 2 * us - ur


def interp_node_conv(u, depth=1):
    for _ in range(depth):
        for ax in range(u.ndim):
            u = _conv_t_axis(u, [0.5, 1., 0.5], ax)
        u = u[(slice(1, -1), ) * u.ndim]
    return u


def interp_cell_conv(u, depth=1):
    for _ in range(depth):
        u = pad_linear(u)
        for ax in range(u.ndim):
            u = _conv_t_axis(u, [0.25, 0.75, 0.75, 0.25], ax)
        u = u[(slice(3, -3), ) * u.ndim]
    return u


def _batch_to_space(uu, dim, crop_lo, crop_hi):
    """Interleave 2^dim arrays (index bits in axis order) into one grid of
    twice the size, then crop."""
    shape = uu[0].shape
    out = np.zeros(tuple(2 * s for s in shape))
    for m, a in enumerate(uu):
        bits = [(m >> (dim - 1 - d)) & 1 for d in range(dim)]
        out[tuple(slice(b, None, 2) for b in bits)] = a
    return out[tuple(slice(crop_lo, 2 * s - crop_hi) for s in shape)]


def _rolled(u, d):
    return np.roll(u, tuple(-x for x in d), tuple(range(u.ndim)))


def interp_node_manual(u, depth=1):
    for _ in range(depth):
        dim = u.ndim
        offs = np.array(np.meshgrid(*[[0, 1]] * dim, indexing='ij'))
        offs = offs.reshape(dim, -1).T
        uu = []
        for dw in offs:
            terms = [r for r in offs if np.all(r <= dw)]
            uu.append(sum(_rolled(u, r) for r in terms) / len(terms))
        u = _batch_to_space(uu, dim, 0, 1)
    return u


def interp_cell_manual(u, depth=1):
    for _ in range(depth):
        dim = u.ndim
        up = pad_linear(u)
        offs = np.array(np.meshgrid(*[[0, 1]] * dim, indexing='ij'))
        offs = offs.reshape(dim, -1).T
        uu = []
        for dw in offs:
            ww = [3**(dim - np.sum(np.abs(r - dw))) for r in offs]
            uu.append(sum(w * _rolled(up, r) for r, w in zip(offs, ww)) /
                      sum(ww))
        u = _batch_to_space(uu, dim, 1, 3)
    return u


def interp_field(u, depth=1, cell=False):
    return interp_cell_conv(u, depth) if cell else interp_node_conv(u, depth)


def split_weights(uw, nnw, factors=None):
    uu, s = [], 0
    factors = factors or [1.] * len(nnw)
    for nw, f in zip(nnw, factors):
        n = int(np.prod(nw))
        uu.append(np.reshape(uw[s:s + n], nw) * f)
        s += n
    return uu


def weights_to_fields(uw, nnw, factors=None, cell=False):
    """multigrid.py:1380-1396."""
    return [interp_field(u, i, cell) for i, u in
            enumerate(split_weights(uw, nnw, factors))]


def weights_to_field(uw, nnw, factors=None, cell=False):
    """multigrid.py:1410-1426: coarsest first, interpolated and added."""
    res = None
    for u in reversed(split_weights(uw, nnw, factors)):
        res = u if res is None else interp_field(res, 1, cell) + u
    return res
