"""Parity of the HIP cavity kernels with the CPU oracle (float64 tape)."""

import numpy as np
import pytest
import torch

from oracle import operators as ops
from oracle import tape

pytestmark = pytest.mark.gpu

# FP64 kernels must agree with the oracle to round-off: terms are O(mu/h^2),
# so the bound is relative to the largest term magnitude of each field.
RTOL64 = 1e-12
RTOL32 = 2e-5


def _state(n0, n1, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.uniform(-1, 1, (n0, n1)) for k in 'uvp'}


def _dev(fields, dev, dtype):
    return [torch.tensor(fields[k], dtype=dtype, device=dev) for k in 'uvp']


def _close(a, b, rtol, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = scale if scale is not None else max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / s
    assert err <= rtol, 'relative error {:.3e} > {:.1e}'.format(err, rtol)


CASES = [(16, 16, 100., 0, 1), (17, 23, 400., 0, 1), (9, 12, 10., 1, 1),
         (11, 10, 50., 2, 1), (12, 9, 30., 0, 0)]


@pytest.mark.parametrize('n0,n1,re,bc,rc', CASES)
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_residual_and_jacobian(dev, n0, n1, re, bc, rc, dtype):
    from odil_amd.operators import Cavity
    st = _state(n0, n1, seed=n0 * 7 + n1)
    dom = ops.cavity_domain(n0, n1)
    op_ref = ops.make_cavity_operator(1 / re, bc=bc, rhiechow=rc)
    ff, grads, desc = tape.eval_grad(op_ref, dom, st)
    op = Cavity((n0, n1), 1 / re, bc=bc, rhiechow=rc)
    rtol = RTOL64 if dtype == torch.float64 else RTOL32
    res, ssq = op.residual(_dev(st, dev, dtype))
    for e in range(3):
        _close(res[e].cpu(), ff[e], rtol, scale=np.abs(ff[e]).max() + 1 / re)
        np.testing.assert_allclose(ssq[e].item(), (ff[e]**2).sum(),
                                   rtol=1e-10 if dtype == torch.float64 else
                                   1e-4)
    jac, _ = op.jacobian(_dev(st, dev, dtype))
    coef = jac.coef.cpu().numpy()
    seen = set()
    for q, (e, k, s) in enumerate(jac.entries):
        key = (k, tuple(s))
        g = None
        if key in desc:
            g = grads[e][desc.index(key)]
        if g is None:
            assert np.abs(coef[q]).max() == 0, (e, k, s)
        else:
            _close(coef[q], g, rtol, scale=np.abs(g).max() + 1)
            seen.add((e, key))
    # Every non-zero tape coefficient is represented by a kernel entry.
    for e in range(3):
        for j, key in enumerate(desc):
            g = grads[e][j]
            if g is not None and np.any(g):
                assert (e, key) in seen, (e, key)


@pytest.mark.parametrize('n0,n1,re,bc,rc', CASES + [(40, 70, 100., 0, 1)])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_loss_grad(dev, n0, n1, re, bc, rc, dtype):
    from odil_amd.operators import Cavity
    st = _state(n0, n1, seed=3 * n0 + n1)
    dom = ops.cavity_domain(n0, n1)
    op_ref = ops.make_cavity_operator(1 / re, bc=bc, rhiechow=rc)
    loss, grads, split = tape.eval_loss_grad(op_ref, dom, st)
    op = Cavity((n0, n1), 1 / re, bc=bc, rhiechow=rc)
    l, g, ls = op.loss_grad(_dev(st, dev, dtype))
    rtol = RTOL64 if dtype == torch.float64 else RTOL32
    np.testing.assert_allclose(l.item(), loss,
                               rtol=1e-11 if dtype == torch.float64 else 1e-4)
    np.testing.assert_allclose(ls.cpu().numpy(), split,
                               rtol=1e-11 if dtype == torch.float64 else 1e-4)
    for k, name in enumerate('uvp'):
        _close(g[k].cpu(), grads[name], rtol,
               scale=np.abs(grads[name]).max())


def test_fixture_state_residual(dev):
    """Converged Re=100 state shipped by the reference (128x128): the FP64
    residual kernel agrees with the oracle and is small."""
    import os
    from odil_amd.operators import Cavity
    path = os.path.join(os.path.dirname(__file__), 'golden',
                        'cavity_N128_Re100.npz')
    d = np.load(path)
    st = {k: d[k] for k in 'uvp'}
    ff = tape.eval_operator(ops.make_cavity_operator(0.01),
                            ops.cavity_domain(128), st)
    res, ssq = Cavity((128, 128), 0.01).residual(_dev(st, dev,
                                                      torch.float64))
    # Round-off is relative to the largest term, mu/h^2 * |u| ~ 164.
    for e in range(3):
        _close(res[e].cpu(), ff[e], 1e-12, scale=0.01 * 128**2)
        assert np.sqrt(ssq[e].item() / 128**2) < 1e-5


# -- end to end: Newton with the GMG normal-equation solver ------------------

@pytest.fixture(scope='module')
def oracle_cavity65():
    """Converged Re=100 cavity on 65x65 cells from the oracle's Newton with
    the reference's direct normal-equation solve (linsolver.py:35-45), 40
    steps from zero (max RMS residual ~2e-12)."""
    from oracle import newton
    n = 65
    dom = ops.cavity_domain(n)
    zero = {k: np.zeros((n, n)) for k in 'uvp'}
    packed = newton.optimize_newton(ops.make_cavity_operator(0.01), dom,
                                    zero, 40)
    return tape.unpack_fields(dom, packed)


@pytest.mark.parametrize('precision,tol,rtol,kw', [
    ('fp64', 1e-11, 1e-8, {}),
    ('mixed', 1e-9, 1e-6, {}),                      # FP32 Krylov
    ('mixed', 1e-9, 1e-6, {'krylov': 'fp64'}),      # FP64 Krylov, FP32 MG
    ('mixed', 1e-9, 1e-6, {'refine': 1}),           # + FP64 refinement
    ('mixed', 1e-9, 1e-6, {'setup_every': 4}),      # lagged hierarchy
    ('mixed', 1e-9, 1e-6, {'setup_every': 4, 'setup_first': 2}),
    ('mixed', 1e-9, 1e-6, {'setup_every': 12, 'setup_first': 2,
                           'coarse_smooth': (1, 1)}),
    ('fp64', 1e-11, 1e-8, {'coarse_smooth': (0, 1)}),
    ('mixed', 1e-9, 1e-6, {'overlap_setup': True}),  # built on a 2nd stream
    ('fp64', 1e-11, 1e-8, {'setup_every': 2, 'overlap_setup': True}),
])
def test_newton_multigrid_matches_oracle_direct(dev, oracle_cavity65,
                                                precision, tol, rtol, kw):
    """The GPU Newton-multigrid solve (FCG on J^T J, V(2,2) preconditioner,
    5 levels) reaches the same converged cavity solution as the oracle's
    direct Newton; pressure is compared up to its constant (null mode)."""
    from odil_amd.newton import NewtonMultigrid
    from odil_amd.operators import Cavity
    n = 65
    op = Cavity((n, n), 0.01)
    nm = NewtonMultigrid(op, nlvl=5, maxiter=10, precision=precision, **kw)
    fields = [torch.zeros(n, n, dtype=torch.float64, device=dev)
              for _ in range(3)]
    hist = nm.run(fields, 80, tol=tol)
    assert max(hist[-1][1]) <= tol, hist[-1]
    ref = oracle_cavity65
    for k, name in enumerate('uv'):
        _close(fields[k].cpu().numpy(), ref[name], rtol)
    p = fields[2].cpu().numpy()
    _close(p - p.mean(), ref['p'] - ref['p'].mean(), rtol)


def test_newton_multigrid_options_rejected(dev):
    from odil_amd.newton import NewtonMultigrid
    from odil_amd.operators import Cavity
    op = Cavity((17, 17), 0.01)
    with pytest.raises(ValueError):
        NewtonMultigrid(op, nlvl=3, precision='fp64', krylov='fp32')
    with pytest.raises(ValueError):
        NewtonMultigrid(op, nlvl=3, precision='mixed', krylov='fp64',
                        refine=1)
    with pytest.raises(ValueError):
        NewtonMultigrid(op, nlvl=3, precision='mixed', krylov='bf16')
    with pytest.raises(ValueError):
        NewtonMultigrid(op, nlvl=3, coarse_smooth=(1, -1))


# -- inverse-problem terms (imposed points, kreg) ----------------------------
def _imposed(n0, n1, seed):
    rng = np.random.default_rng(seed)
    mask = (rng.uniform(size=(n0, n1)) < 0.1).astype(np.float64)
    mask[0, 0] = 1.
    return mask, rng.uniform(-1, 1, (n0, n1)), rng.uniform(-1, 1, (n0, n1))


@pytest.mark.parametrize('n0,n1,re,bc,rc', CASES)
@pytest.mark.parametrize('terms', ['imposed', 'kreg', 'both'])
@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_cavity_extras_match_oracle(dev, n0, n1, re, bc, rc, terms, dtype):
    """Loss, split and gradient of the cavity with the extra equations
    appended, the extras' residuals and Jacobian, against the oracle."""
    from odil_amd.operators import (Cavity, CavityExtras,
                                    cavity_inverse_loss_grad)
    st = _state(n0, n1, seed=5 * n0 + n1)
    imp = _imposed(n0, n1, n0) if terms in ('imposed', 'both') else None
    kreg = 0.3 if terms in ('kreg', 'both') else 0.
    dom = ops.cavity_domain(n0, n1)
    op_ref = ops.make_cavity_operator(1 / re, bc=bc, rhiechow=rc,
                                      imposed=imp, alpha=0.8, kreg=kreg)
    loss, grads, split = tape.eval_loss_grad(op_ref, dom, st)
    cav = Cavity((n0, n1), 1 / re, bc=bc, rhiechow=rc)
    ext = CavityExtras(cav, imposed=imp, alpha=0.8, kreg=kreg)
    f = _dev(st, dev, dtype)
    l, g, ls = cavity_inverse_loss_grad(cav, ext, f)
    tol = 1e-11 if dtype == torch.float64 else 1e-4
    np.testing.assert_allclose(l.item(), loss, rtol=tol)
    np.testing.assert_allclose(ls.cpu().numpy(), split, rtol=tol,
                               atol=tol * max(split))
    rtol = RTOL64 if dtype == torch.float64 else RTOL32
    for k, name in enumerate('uvp'):
        _close(g[k].cpu(), grads[name], rtol,
               scale=np.abs(grads[name]).max())
    # residuals and Jacobian of the extra equations
    ff, jg, desc = tape.eval_grad(op_ref, dom, st)
    res, _ = ext.residual(f)
    on = ext.enabled()
    for i, e in enumerate(on):
        _close(res[e].cpu(), ff[3 + i], rtol,
               scale=np.abs(ff[3 + i]).max() + 1)
    jac, _ = ext.jacobian(f)
    coef = jac.coef.cpu().numpy()
    for q, (e, k, s) in enumerate(jac.entries):
        key = (k, tuple(s))
        gref = None
        if e in on and key in desc:
            gref = jg[3 + on.index(e)][desc.index(key)]
        if gref is None:
            assert np.abs(coef[q]).max() == 0, (e, k, s)
        else:
            _close(coef[q], gref, rtol, scale=np.abs(gref).max() + 1)


def test_cavity_extras_rejects_bad_shape(dev):
    from odil_amd.operators import Cavity, CavityExtras
    cav = Cavity((8, 8), 0.01)
    with pytest.raises(ValueError):
        CavityExtras(cav, imposed=(np.zeros((8, 7)), np.zeros((8, 8)),
                                   np.zeros((8, 8))))
