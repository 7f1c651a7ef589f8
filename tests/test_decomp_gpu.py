"""MultigridDecomp on the GPU against the oracle (multigrid.py:1217-1426)."""

import numpy as np
import pytest
import torch

from oracle import decomp as od

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def dev():
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch.device('cuda')


def _levels(shape, n, cell):
    nnw = [tuple(shape)]
    for _ in range(n - 1):
        nnw.append(tuple(s // 2 if cell else (s + 1) // 2 for s in nnw[-1]))
    return nnw


@pytest.mark.parametrize('shape,cell', [((17, ), False), ((9, 17), False),
                                        ((9, 5, 17), False), ((16, ), True),
                                        ((8, 12), True), ((8, 4, 12), True)])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_weights_to_field_matches_oracle(dev, shape, cell, dtype):
    from odil_amd.decomp import MultigridDecomp as MD
    nnw = _levels(shape, 3, cell)
    rng = np.random.default_rng(len(shape) + cell)
    n = sum(int(np.prod(s)) for s in nnw)
    uw = rng.standard_normal(n)
    factors = [1., 0.5, 2.]
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    t = torch.tensor(uw, dtype=dtype, device=dev)
    got = MD.weights_to_field(t, nnw, factors, cell).double().cpu().numpy()
    ref = od.weights_to_field(uw, nnw, factors, cell)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    for a, b in zip(MD.weights_to_fields(t, nnw, factors, cell),
                    od.weights_to_fields(uw, nnw, factors, cell)):
        assert np.abs(a.double().cpu().numpy() - b).max() <= \
            tol * np.abs(b).max()
    # adjoint: <W uw, g> = <uw, W^T g> and the transpose of the oracle map
    g = rng.standard_normal(ref.shape)
    wt = MD.weights_to_field_t(torch.tensor(g, dtype=dtype, device=dev),
                               nnw, factors, cell).double().cpu().numpy()
    lhs = float(np.sum(ref * g))
    assert abs(lhs - float(uw @ wt)) <= tol * 10 * (abs(lhs) + 1)
    if dtype == torch.float64 and n <= 400:
        cols = np.stack([od.weights_to_field(e, nnw, factors, cell).ravel()
                         for e in np.eye(n)], axis=1)
        np.testing.assert_allclose(wt, cols.T @ g.ravel(), rtol=1e-11,
                                   atol=1e-12)


def test_weights_to_field_autograd(dev):
    from odil_amd.decomp import MultigridDecomp as MD, WeightsToField
    nnw = _levels((9, 17), 3, False)
    n = sum(int(np.prod(s)) for s in nnw)
    uw = torch.randn(n, dtype=torch.float64, device=dev, requires_grad=True)
    target = torch.randn(nnw[0], dtype=torch.float64, device=dev)
    loss = ((WeightsToField.apply(uw, nnw) - target)**2).sum()
    loss.backward()
    f = MD.weights_to_field(uw.detach(), nnw)
    want = MD.weights_to_field_t(2 * (f - target), nnw)
    assert torch.allclose(uw.grad, want, rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        MD.weights_to_field(uw.detach(), [(9, 17), (5, 8)])
    ones = MD.get_weights_coarse_ones(nnw, torch.float64, dev)
    np.testing.assert_allclose(MD.weights_to_field(ones, nnw).cpu().numpy(),
                               1., rtol=1e-13)
