"""Parity of the wave, Poisson, Laplace and advection HIP kernels with the
CPU oracle (float64 tape)."""

import numpy as np
import pytest
import torch

from oracle import operators as ops
from oracle import tape
from oracle.tape import Domain

pytestmark = pytest.mark.gpu


def _close(a, b, rtol, scale):
    err = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max()
    assert err <= rtol * scale, 'error {:.3e} > {:.1e} * {:.3e}'.format(
        err, rtol, scale)


def _check_jac(jac, grads, desc, neq, rtol):
    coef = jac.coef.double().cpu().numpy()
    seen = set()
    for q, (e, k, s) in enumerate(jac.entries):
        key = (k, tuple(s))
        g = grads[e][desc.index(key)] if key in desc else None
        if g is None:
            assert np.abs(coef[q]).max() == 0, (e, k, s)
        else:
            _close(coef[q], g, rtol, np.abs(g).max() + 1)
            seen.add((e, key))
    for e in range(neq):
        for j, key in enumerate(desc):
            g = grads[e][j]
            if g is not None and np.any(g):
                assert (e, key) in seen, (e, key)


def _wave_case(nt, nx, seed):
    dom, op_ref, (iu, iut, lu, ru) = ops.wave_gaussian_setup(nt, nx, seed)
    rng = np.random.default_rng(seed)
    lu = rng.normal(size=nt)  # non-zero boundary values exercise the halo
    ru = rng.normal(size=nt)
    iut = rng.normal(size=nx)
    op_ref = ops.make_wave_operator(iu, iut, lu, ru)
    u = rng.normal(size=(nt, nx))
    return dom, op_ref, (iu, iut, lu, ru), u


@pytest.mark.parametrize('nt,nx', [(8, 9), (33, 70), (64, 64)])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_wave(dev, nt, nx, dtype):
    from odil_amd.operators import Wave
    dom, op_ref, bnd, u = _wave_case(nt, nx, nt + nx)
    ff, grads, desc = tape.eval_grad(op_ref, dom, {'u': u})
    loss, lg, split = tape.eval_loss_grad(op_ref, dom, {'u': u})
    op = Wave((nt, nx), *bnd, device=dev)
    x = [torch.tensor(u, dtype=dtype, device=dev)]
    rtol = 1e-12 if dtype == torch.float64 else 2e-5
    scale = 1. / min(op.dt, op.dx)**2
    res, ssq = op.residual(x)
    _close(res[0].cpu(), ff[0], rtol, scale)
    jac, _ = op.jacobian(x)
    _check_jac(jac, grads, desc, 1, rtol)
    l, g, _ = op.loss_grad(x)
    np.testing.assert_allclose(l.item(), loss,
                               rtol=1e-11 if dtype == torch.float64 else 1e-4)
    _close(g[0].cpu(), lg['u'], rtol, np.abs(lg['u']).max())


@pytest.mark.parametrize('n0,n1', [(7, 9), (40, 70)])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_poisson(dev, n0, n1, dtype):
    from odil_amd.operators import Poisson
    rng = np.random.default_rng(n0)
    src = rng.normal(size=(n0, n1))
    u = rng.normal(size=(n0, n1))
    dom = Domain(ndim=2, shape=(n0, n1), varnames=('x', 'y'),
                 fieldnames=['u'])
    op_ref = ops.make_poisson_operator(src)
    ff, grads, desc = tape.eval_grad(op_ref, dom, {'u': u})
    loss, lg, _ = tape.eval_loss_grad(op_ref, dom, {'u': u})
    op = Poisson((n0, n1), torch.tensor(src, dtype=dtype, device=dev))
    x = [torch.tensor(u, dtype=dtype, device=dev)]
    rtol = 1e-12 if dtype == torch.float64 else 2e-5
    scale = max(n0, n1)**2
    res, _ = op.residual(x)
    _close(res[0].cpu(), ff[0], rtol, scale)
    jac, _ = op.jacobian(x)
    _check_jac(jac, grads, desc, 1, rtol)
    l, g, _ = op.loss_grad(x)
    np.testing.assert_allclose(l.item(), loss,
                               rtol=1e-11 if dtype == torch.float64 else 1e-4)
    _close(g[0].cpu(), lg['u'], rtol, np.abs(lg['u']).max())


def test_laplace(dev):
    from odil_amd.operators import Laplace
    n0, n1 = 12, 15
    rng = np.random.default_rng(0)
    st = {k: rng.normal(size=(n0, n1)) for k in 'uv'}
    dom = Domain(ndim=2, shape=(n0, n1), varnames=('x', 'y'),
                 fieldnames=['u', 'v'])
    ff, grads, desc = tape.eval_grad(ops.operator_laplace, dom, st)
    loss, lg, split = tape.eval_loss_grad(ops.operator_laplace, dom, st)
    op = Laplace((n0, n1))
    x = [torch.tensor(st[k], device=dev) for k in 'uv']
    res, _ = op.residual(x)
    for e in range(3):
        _close(res[e].cpu(), ff[e], 1e-12, 300.)
    jac, _ = op.jacobian(x)
    _check_jac(jac, grads, desc, 3, 1e-12)
    l, g, ls = op.loss_grad(x)
    np.testing.assert_allclose(ls.cpu().numpy(), split, rtol=1e-11,
                               atol=1e-300)
    for k, name in enumerate('uv'):
        _close(g[k].cpu(), lg[name], 1e-12, np.abs(lg[name]).max())


@pytest.mark.parametrize('shape,seed', [((5, 6, 7), 1), ((16, 32, 24), 2),
                                        ((9, 8, 8), 5)])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_advection(dev, shape, seed, dtype):
    from odil_amd.operators import Advection
    dom, op_ref, (ax, ay, p0) = ops.advection_setup(*shape, seed=seed)
    rng = np.random.default_rng(seed)
    p = rng.normal(size=shape)
    ff = tape.eval_operator(op_ref, dom, {'p': p})
    loss, lg, _ = tape.eval_loss_grad(op_ref, dom, {'p': p})
    op = Advection(shape, ax, ay, torch.tensor(p0, dtype=dtype, device=dev))
    x = [torch.tensor(p, dtype=dtype, device=dev)]
    rtol = 1e-12 if dtype == torch.float64 else 2e-5
    res, ssq = op.residual(x)
    _close(res[0].cpu(), ff[0], rtol, np.abs(ff[0]).max())
    l, g, _ = op.loss_grad(x)
    np.testing.assert_allclose(l.item(), loss,
                               rtol=1e-11 if dtype == torch.float64 else 1e-4)
    _close(g[0].cpu(), lg['p'], rtol, np.abs(lg['p']).max())


def _ops_for_normal(dev, dtype):
    """(name, operator, state) for every 2D operator, random states."""
    from odil_amd.operators import Cavity, Laplace, Poisson, Wave
    rng = np.random.default_rng(7)
    out = []
    n0, n1 = 37, 70
    st = [torch.tensor(rng.uniform(-1, 1, (n0, n1)), dtype=dtype, device=dev)
          for _ in range(3)]
    out.append(('cavity', Cavity((n0, n1), 0.01), st))
    out.append(('cavity_bc2', Cavity((n0, n1), 0.03, bc=2), st))
    out.append(('cavity_rc0', Cavity((n0, n1), 0.02, rhiechow=0), st))
    st2 = [torch.tensor(rng.uniform(-1, 1, (70, 100)), dtype=dtype,
                        device=dev) for _ in range(3)]
    out.append(('cavity_bc1_70x100', Cavity((70, 100), 0.005, bc=1), st2))
    dom, op_ref, bnd, u = _wave_case(33, 45, 3)
    out.append(('wave', Wave((33, 45), *bnd, device=dev),
                [torch.tensor(u, dtype=dtype, device=dev)]))
    src = torch.tensor(rng.normal(size=(n0, n1)), dtype=dtype, device=dev)
    out.append(('poisson', Poisson((n0, n1), src),
                [torch.tensor(rng.normal(size=(n0, n1)), dtype=dtype,
                              device=dev)]))
    out.append(('laplace', Laplace((n0, n1)), st[:2]))
    return out


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_normal_apply_dot_matches_product_and_dot(dev, dtype):
    """odil_cavity_normal_apply_dot: y equals mode 0 of normal_apply bit for
    bit and the dot equals the FP64 sum of x * y up to summation order, on
    grids with interior tiles and on one without (every cell a frame
    cell); the partials start as NaN, so every slot must be written."""
    from odil_amd import _lib
    from odil_amd.newton import DeviceScalars
    from odil_amd.operators import Cavity
    rng = np.random.default_rng(11)
    for shape, mu, bc in (((37, 70), 0.01, 0), ((70, 100), 0.005, 1),
                          ((6, 9), 0.02, 2), ((129, 65), 0.01, 0)):
        op = Cavity(shape, mu, bc=bc)
        assert op.has_normal_dot()
        mk = lambda: [torch.tensor(rng.uniform(-1, 1, shape), dtype=dtype,
                                   device=dev) for _ in range(3)]
        st, x = mk(), mk()
        y1 = [torch.empty_like(a) for a in x]
        op.normal_apply(st, x, y1, mode=0)
        part = torch.full((op.nblocks(2, _lib.dtype_tag(dtype)), ),
                          float('nan'), dtype=torch.float64, device=dev)
        sc = DeviceScalars(1, dev)
        y2 = [torch.full_like(a, float('nan')) for a in x]
        op.normal_apply_dot(st, x, y2, part, sc.p(0))
        for a, b in zip(y1, y2):
            assert torch.equal(a, b), shape
        prod = [a.double() * b.double() for a, b in zip(x, y1)]
        want = sum(q.sum() for q in prod).item()
        mag = sum(q.abs().sum() for q in prod).item()
        assert abs(sc.t[0].item() - want) <= 1e-13 * mag, shape


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_normal_apply_matches_stored_normal(dev, dtype):
    from odil_amd.multigrid import BlockStencil, normal_operator
    tol = 1e-12 if dtype == torch.float64 else 3e-5
    rng = np.random.default_rng(3)
    for name, op, st in _ops_for_normal(dev, dtype):
        jac, res = op.jacobian(st, want_residual=True)
        J = BlockStencil.from_jacobian(jac)
        N = normal_operator(J)
        nf = len(st)
        mk = lambda: [torch.tensor(rng.uniform(-1, 1, op.shape), dtype=dtype,
                                   device=dev) for _ in range(nf)]
        x, b = mk(), mk()
        y = [torch.empty_like(a) for a in x]
        ref = N.apply(x)
        scale = max(a.abs().max().item() for a in ref)
        op.normal_apply(st, x, y, mode=0)
        for a, r in zip(y, ref):
            assert (a - r).abs().max().item() <= tol * scale, name
        op.normal_apply(st, x, y, b=b, mode=1)
        for a, r, bb in zip(y, ref, b):
            assert (a - (bb - r)).abs().max().item() <= tol * scale, name
        d = [torch.empty_like(a) for a in x]
        op.normal_apply(st, None, d, mode=3)
        zero = (0, ) * len(op.shape)
        for k in range(nf):
            q = N.entries.index((k, k, zero))
            dr = N.coef[q]
            assert (d[k] - dr).abs().max().item() <= \
                tol * dr.abs().max().item(), name
        op.normal_apply(st, x, y, b=b, d=d, omega=0.6, mode=2)
        for k in range(nf):
            want = x[k] + 0.6 * (b[k] - ref[k]) / d[k]
            assert (y[k] - want).abs().max().item() <= \
                tol * 10 * want.abs().max().item(), name
        # rhs = -J^T f with the residual's sums of squares
        out = [torch.empty_like(a) for a in st]
        ssq = op.jtf(st, out, -1.0)
        ref = J.apply_t([-r for r in res])
        sc = max(a.abs().max().item() for a in ref)
        for a, r in zip(out, ref):
            assert (a - r).abs().max().item() <= tol * sc, name
        for e in range(len(res)):
            np.testing.assert_allclose(
                ssq[e].item(), (res[e].double()**2).sum().item(),
                rtol=1e-12 if dtype == torch.float64 else 1e-4)
