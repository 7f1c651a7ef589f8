"""The kernel adjoint behind torch.autograd: gradients equal the kernel's
and the oracle's, pass torch's finite-difference gradcheck, and drive a
torch optimiser."""

import numpy as np
import pytest
import torch

from oracle import operators as ops
from oracle import tape

pytestmark = pytest.mark.gpu


def test_discrete_loss_gradcheck_and_oracle(dev):
    from odil_amd.autograd import discrete_loss
    from odil_amd.operators import Cavity
    n0, n1 = 9, 11
    rng = np.random.default_rng(4)
    st = {k: rng.uniform(-1, 1, (n0, n1)) for k in 'uvp'}
    op = Cavity((n0, n1), 0.05)
    fields = [torch.tensor(st[k], device=dev, requires_grad=True)
              for k in 'uvp']
    loss = discrete_loss(op, *fields)
    loss.backward()
    ol, og, _ = tape.eval_loss_grad(ops.make_cavity_operator(0.05),
                                    ops.cavity_domain(n0, n1), st)
    np.testing.assert_allclose(loss.item(), ol, rtol=1e-12)
    for f, k in zip(fields, 'uvp'):
        np.testing.assert_allclose(f.grad.cpu().numpy(), og[k], rtol=1e-10,
                                   atol=1e-12 * np.abs(og[k]).max())


def test_gradcheck_wave_and_inverse(dev):
    """torch.autograd.gradcheck (central FD in FP64) on the wave loss and
    on the inverse-wave loss in (u, c): both operators are free of frozen
    paths, so the kernel adjoint is the full derivative."""
    from odil_amd.autograd import InverseWaveLoss, discrete_loss
    from odil_amd.operators import Wave, WaveInverse
    from oracle import inverse as inv
    dom, op_ref, bnd = ops.wave_gaussian_setup(8, 9, seed=1)
    op = Wave((8, 9), *bnd, device=dev)
    u = torch.randn(8, 9, dtype=torch.float64, device=dev,
                    requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: discrete_loss(op, x), (u, ),
                                    eps=1e-6, atol=1e-5, rtol=1e-6)
    case = inv.wave_inverse_setup(10, 8, nsensor=3, seed=2)
    wi = WaveInverse((10, 8), case['init_u'], case['init_ut'],
                     case['left_u'], case['right_u'], case['sensors'],
                     case['data'], k_data=2.0, k_reg=0.3, device=dev)
    uu = torch.randn(10, 8, dtype=torch.float64, device=dev,
                     requires_grad=True)
    cc = (1 + 0.1 * torch.randn(8, dtype=torch.float64, device=dev)
          ).requires_grad_()
    assert torch.autograd.gradcheck(
        lambda a, b: InverseWaveLoss.apply(wi, a, b), (uu, cc), eps=1e-6,
        atol=1e-4, rtol=1e-6)


def test_torch_optimizer_on_kernel_loss(dev):
    """torch.optim.LBFGS on the Poisson discrete loss through the
    autograd wrapper converges like the device L-BFGS."""
    from odil_amd.autograd import discrete_loss
    from odil_amd.operators import Poisson
    n = 16
    src = ops.gaussian_mixture_source(n, seed=4)
    op = Poisson((n, n), torch.tensor(src, device=dev))
    u = torch.zeros(n, n, dtype=torch.float64, device=dev,
                    requires_grad=True)
    opt = torch.optim.LBFGS([u], lr=1, max_iter=400, history_size=10,
                            tolerance_grad=0, tolerance_change=0,
                            line_search_fn='strong_wolfe')

    def closure():
        opt.zero_grad()
        loss = discrete_loss(op, u)
        loss.backward()
        return loss

    l0 = closure().item()
    opt.step(closure)
    assert closure().item() <= 1e-6 * l0
