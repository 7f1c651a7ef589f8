"""Allocation-free cavity loss + adjoint (`odil_cavity_loss_grad_noalloc`)
against the workspace entry point `odil_cavity_loss_grad` followed by
`odil_reduce_partials`, called through the C ABI: gradients and sums of
squares must be bitwise equal (same kernels, same fixed-order reduction),
on two calls in a row, the second on a smaller grid."""

import ctypes

import numpy as np
import pytest
import torch

from odil_amd import _lib
from odil_amd._lib import c_int, ptr, ptrs

pytestmark = pytest.mark.gpu


def _state(n0, n1, seed, dtype, dev):
    g = torch.Generator(device='cpu').manual_seed(seed)
    return [torch.rand((n0, n1), generator=g, dtype=torch.float64)
            .mul_(2).sub_(1).to(dtype=dtype, device=dev) for _ in range(3)]


def _with_workspace(op, fields):
    dtag = _lib.dtype_tag(fields[0].dtype)
    grads = [torch.full_like(f, np.nan) for f in fields]
    nb = op.nblocks(1, dtag)
    part = torch.empty((3, nb), dtype=torch.float64, device=fields[0].device)
    op.launch_loss_grad(fields, grads, part)
    ssq = torch.empty(3, dtype=torch.float64, device=fields[0].device)
    _lib.check(_lib.lib().odil_reduce_partials(ptr(part), c_int(3), c_int(nb),
                                               ptr(ssq), _lib.stream()),
               'reduce_partials')
    return grads, ssq


def _noalloc(op, fields):
    dtag = _lib.dtype_tag(fields[0].dtype)
    grads = [torch.full_like(f, np.nan) for f in fields]
    ssq = torch.full((3,), np.nan, dtype=torch.float64,
                     device=fields[0].device)
    st = _lib.lib().odil_cavity_loss_grad_noalloc(
        c_int(dtag), *op._shape_args(), *op._args_for(dtag), ptrs(fields),
        ptrs(grads), ptr(ssq), _lib.stream())
    _lib.check(st, 'cavity_loss_grad_noalloc')
    return grads, ssq


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_noalloc_matches_workspace_entry(dev, dtype):
    from odil_amd.operators import Cavity
    for n0, n1, seed in ((300, 257, 1), (45, 70, 2)):
        op = Cavity((n0, n1), 0.01)
        fields = _state(n0, n1, seed, dtype, dev)
        g_ref, s_ref = _with_workspace(op, fields)
        g, s = _noalloc(op, fields)
        torch.cuda.synchronize()
        assert torch.equal(s, s_ref), (n0, n1, s, s_ref)
        for a, b in zip(g, g_ref):
            assert torch.equal(a, b), (n0, n1)
        loss, _, split = op.loss_grad(fields)
        np.testing.assert_allclose(
            (s / (n0 * n1)).cpu().numpy(), split.cpu().numpy(), rtol=0,
            atol=0)


def test_noalloc_rejects_oversize(dev):
    """A grid needing more partials than the library's array holds is
    refused on the host before any launch."""
    from odil_amd.operators import Cavity
    op = Cavity((4096, 4096), 0.01)
    small = [torch.zeros((8, 8), dtype=torch.float64, device=dev)
             for _ in range(3)]
    ssq = torch.zeros(3, dtype=torch.float64, device=dev)
    st = _lib.lib().odil_cavity_loss_grad_noalloc(
        c_int(_lib.F64), *op._shape_args(), *op._args_for(_lib.F64),
        ptrs(small), ptrs(small), ptr(ssq), _lib.stream())
    assert st == 3
    assert b'scratch' in _lib.lib().odil_last_error()
    assert ctypes.c_int(st).value == 3
