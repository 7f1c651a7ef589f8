"""Batched Poisson instances: the persistent launch against the single-grid
kernel and the oracle, and the graph-captured batched normal-equation
solve against direct solves."""

import numpy as np
import pytest
import scipy.sparse.linalg
import torch

from oracle import operators as ops
from oracle import tape

pytestmark = pytest.mark.gpu


def _batch(dev, dtype, count=10, seed=0, lo=64, hi=256, snap=True,
           rect=False):
    from odil_amd.batched import PoissonBatch, batch_sizes, gaussian_mixture
    sizes = batch_sizes(count, seed, lo, hi, snap=snap)
    srcs = [gaussian_mixture(n, seed=1000 + i) for i, n in enumerate(sizes)]
    if rect:  # n x m instances, m drawn independently (sources cropped)
        m = batch_sizes(count, seed + 1, lo, hi, snap=False)
        sizes = [(n, min(n, k)) for n, k in zip(sizes, m)]
        srcs = [s[:, :k] for s, (_, k) in zip(srcs, sizes)]
    return PoissonBatch(sizes, srcs, dtype=dtype, device=dev), srcs


def test_sizes_and_sources():
    from odil_amd.batched import batch_sizes, gaussian_mixture
    s = batch_sizes(256, 0)
    assert set(s) <= {65, 129, 257, 513, 1025} and len(set(s)) == 5
    r = batch_sizes(256, 0, snap=False)
    assert min(r) >= 64 and max(r) <= 1024 and len(set(r)) > 150
    np.testing.assert_array_equal(gaussian_mixture(33, 5),
                                  ops.gaussian_mixture_source(33, 5))


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64])
@pytest.mark.parametrize('sizes', ['snapped', 'uneven', 'rectangular'])
def test_batch_launch_matches_single_grid_and_oracle(dev, dtype, sizes):
    """Snapped (2^k + 1), uneven (any n) and rectangular (n x m) instances:
    every tile shape of the persistent launch equals the single-grid kernel
    bit for bit and the oracle to FP64 round-off."""
    b, srcs = _batch(dev, dtype, snap=sizes == 'snapped',
                     rect=sizes == 'rectangular')
    rng = np.random.default_rng(3)
    b.u.copy_(torch.tensor(rng.normal(size=b.total), dtype=dtype))
    for shape in b.shapes():
        loss, grad = b.launch(shape)
        loss = loss.cpu().numpy().copy()
        grad = grad.clone()
        ops_ = b.instance_ops()
        g1 = torch.empty_like(b.u)
        for i, op in enumerate(ops_):
            l, gi, _ = op.loss_grad([b.view(b.u, i)], [b.view(g1, i)])
            np.testing.assert_allclose(loss[i], l.item(), rtol=1e-12)
        # same per-cell arithmetic as the single-grid tile kernel
        assert torch.equal(grad, g1), shape
    # FP64 oracle on two instances
    for i in (0, b.count - 1):
        n = b.sizes[i]
        u = b.view(b.u, i).double().cpu().numpy()
        dom = tape.Domain(ndim=2, shape=n, varnames=('x', 'y'),
                          fieldnames=['u'])
        src = b.view(b.src, i).double().cpu().numpy()
        ol, og, _ = tape.eval_loss_grad(ops.make_poisson_operator(src), dom,
                                        {'u': u})
        rtol = 1e-11 if dtype == torch.float64 else 1e-4
        np.testing.assert_allclose(loss[i], ol, rtol=rtol)
        e = np.abs(b.view(grad, i).double().cpu().numpy() - og['u']).max()
        assert e <= (1e-11 if dtype == torch.float64 else 3e-5) * \
            np.abs(og['u']).max()


def test_autotune_picks_an_available_shape(dev):
    b, _ = _batch(dev, torch.float32, count=6)
    shape = b.tune()
    assert shape in b.shapes()


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_batched_solve_matches_direct(dev, precision):
    """Three graph-replayed solves of the whole batch (10 FCG iterations
    each, V(2,2) preconditioner shared per size; each solve refines the
    previous one like a Newton step) reach the direct solution of every
    instance's discrete Poisson problem."""
    from odil_amd.batched import PoissonBatchSolver
    b, srcs = _batch(dev, torch.float32, count=8, seed=2, lo=64, hi=160)
    sol = PoissonBatchSolver(b, precision=precision, maxiter=10)
    for _ in range(3):
        sol.solve()
    rms = sol.residual_rms()
    for i, n in enumerate(b.sizes):
        src = b.view(b.src, i).double().cpu().numpy()
        dom = tape.Domain(ndim=2, shape=n, varnames=('x', 'y'),
                          fieldnames=['u'])
        const, m = tape.linearize(ops.make_poisson_operator(src), dom,
                                  {'u': np.zeros(n)})
        ref = scipy.sparse.linalg.spsolve(m.tocsc(), -const[0].ravel())
        u = sol.u64[i].cpu().numpy().ravel()
        assert np.abs(u - ref).max() <= 1e-6 * np.abs(ref).max(), (i, n)
        # the residual carries the 1/h^2 scale of the operator
        assert rms[i] <= 1e-5 * np.sqrt(np.mean(src**2)), (i, rms[i])
