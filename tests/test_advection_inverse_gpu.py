"""Inverse advection operator (advection_2d.py:103-250, problem='inverse')
on the GPU against the oracle: residuals of every equation, loss, loss
split and the adjoint gradients of u, v, p."""

import numpy as np
import pytest
import torch

from oracle import operators as ops
from oracle import tape

pytestmark = pytest.mark.gpu

CASES = [
    # shape, kwargs, epoch
    ((4, 5, 6), dict(alpha=1.3, kreg=0.3, ktreg=0.2, divfree=0.7), 0),
    ((5, 8, 7), dict(alpha=0.8, anneal_half=4., kreg=1e-4), 3),
    ((3, 6, 9), dict(onlyfinal=True, kreg=0., periodic=False), 0),
    ((6, 7, 5), dict(kreg=0.5, ktreg=0.4, divfree=0.3, periodic=False,
                     onlyfinal=True, anneal_half=2.), 5),
    ((16, 32, 24), dict(kreg=2e-3, divfree=0.1), 1),
]


@pytest.mark.parametrize('shape,kw,epoch', CASES)
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_advection_inverse_matches_oracle(dev, shape, kw, epoch, dtype):
    from odil_amd.operators import AdvectionInverse
    rng = np.random.default_rng(sum(shape))
    st = {k: rng.normal(size=shape) for k in 'uvp'}
    pex = rng.normal(size=shape)
    dom = ops.advection_inverse_domain(*shape)
    op_ref = ops.make_advection_inverse_operator(pex, **kw)
    ff = tape.eval_operator(op_ref, dom, st, epoch=epoch)
    loss, grads, split = tape.eval_loss_grad(op_ref, dom, st, epoch=epoch)
    op = AdvectionInverse(shape, torch.tensor(pex, dtype=dtype, device=dev),
                          **kw)
    assert op.neq == len(ff)
    fields = [torch.tensor(st[k], dtype=dtype, device=dev) for k in 'uvp']
    res, ssq = op.residual(fields, epoch=epoch)
    rtol = 1e-12 if dtype == torch.float64 else 3e-5
    for e, f in enumerate(ff):
        scale = np.abs(f).max() + 1e-300
        err = np.abs(res[e].double().cpu().numpy() - f).max() / scale
        assert err <= rtol, (e, err)
        np.testing.assert_allclose(ssq[e].item(), (f**2).sum(),
                                   rtol=1e-10 if dtype == torch.float64
                                   else 1e-4)
    l, g, ls = op.loss_grad(fields, epoch=epoch)
    np.testing.assert_allclose(l.item(), loss,
                               rtol=1e-11 if dtype == torch.float64 else 1e-4)
    np.testing.assert_allclose(ls.cpu().numpy(), split,
                               rtol=1e-10 if dtype == torch.float64 else 1e-4,
                               atol=1e-300)
    for k, name in enumerate('uvp'):
        ref = grads[name]
        err = np.abs(g[k].double().cpu().numpy() - ref).max() / np.abs(
            ref).max()
        assert err <= rtol, (name, err)


def test_advection_inverse_rejects(dev):
    from odil_amd.operators import AdvectionInverse
    pex = torch.zeros((3, 2, 5), dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        AdvectionInverse((3, 2, 5), pex)
    op = AdvectionInverse((3, 4, 5), torch.zeros((3, 4, 5),
                                                 dtype=torch.float64,
                                                 device=dev))
    with pytest.raises(ValueError):
        op.residual([torch.zeros((3, 4, 5), device=dev)] * 2)
