"""Pins of the CPU oracle against the reference's golden data and against
closed forms (CPU only, no GPU)."""

import json
import os

import numpy as np
import pytest
import scipy.sparse

from oracle import multigrid as omg
from oracle import newton
from oracle import operators as ops
from oracle import tape

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# -- wave: 2 Newton steps reproduce the reference's discretisation error ----
@pytest.mark.parametrize('nx', [3, 4, 5, 8, 10, 16, 21, 33, 40, 64])
def test_wave_newton_matches_reference_error(nx):
    g = _load('wave_fd_newton.json')
    ref = g['error'][g['nx'].index(nx)]
    dom, op, exact = ops.wave_reference_setup(nx, nx)
    packed = newton.optimize_newton(op, dom, {'u': np.zeros((nx, nx))}, 2)
    u = tape.unpack_fields(dom, packed)['u']
    err = np.mean((exact - u)**2)**0.5
    assert abs(err - ref) <= 1e-11 * ref


# -- cavity: Newton trajectory max|du|, max|dv| of the reference CSV --------
@pytest.mark.parametrize('re_,epochs', [(100, 6), (400, 4)])
def test_cavity_newton_trajectory(re_, epochs):
    g = _load('cavity_newton.json')['Re{}'.format(re_)]
    dom = ops.cavity_domain(64)
    op = ops.make_cavity_operator(1 / re_)
    prev = [None]
    got = []

    def cb(epoch, packed, const):
        if prev[0] is not None:
            d = tape.unpack_fields(dom, packed - prev[0])
            got.append((epoch, abs(d['u']).max(), abs(d['v']).max()))
        prev[0] = packed.copy()

    zero = {k: np.zeros((64, 64)) for k in 'uvp'}
    newton.optimize_newton(op, dom, zero, epochs, callback=cb)
    for epoch, du, dv in got:
        assert abs(du - g['du'][epoch]) <= 1e-9 * g['du'][epoch]
        assert abs(dv - g['dv'][epoch]) <= 1e-9 * g['dv'][epoch]


@pytest.mark.parametrize('re_,bound', [(100, 1e-5), (1000, 1e-4)])
def test_cavity_fixture_is_converged(re_, bound):
    """The reference's own N=128 states (20 direct Newton steps at Re 100 and
    1000, cases/cavity_stream/run) are near zeros of the oracle's operator
    at mu = 1/Re (ns_steady.py:867): RMS residual 1e-6 / 4e-5 against ~1 at
    the zero state. With mu off by a factor 2 it would be ~1e-2 / 1e-3."""
    d = np.load(os.path.join(GOLD, 'cavity_N128_Re{}.npz'.format(re_)))
    st = {k: d[k] for k in 'uvp'}
    ff = tape.eval_operator(ops.make_cavity_operator(1.0 / re_),
                            ops.cavity_domain(128), st)
    for f in ff:
        assert np.sqrt(np.mean(f**2)) < bound
    wrong = tape.eval_operator(ops.make_cavity_operator(2.0 / re_),
                               ops.cavity_domain(128), st)
    assert max(np.sqrt(np.mean(f**2)) for f in wrong) > 10 * bound


@pytest.mark.parametrize('n,re_', [(65, 100), (65, 1000), (129, 1000)])
def test_cavity_converged_fixtures(n, re_):
    """The direct-Newton solutions the GPU Newton tests compare against
    (written by tests/golden/make_cavity_converged.py) are zeros of the
    oracle's operator to round-off and have the cavity's flow pattern."""
    d = np.load(os.path.join(GOLD, 'cavity_converged_N{}_Re{}.npz'.format(
        n, re_)))
    st = {k: d[k] for k in 'uvp'}
    ff = tape.eval_operator(ops.make_cavity_operator(1.0 / re_),
                            ops.cavity_domain(n), st)
    for f in ff:
        assert np.sqrt(np.mean(f**2)) <= 1e-13
    # The lid drives a clockwise primary vortex: u > 0 near the lid, < 0
    # below the centre.
    assert st['u'][:, -1].mean() > 0.5 and st['u'][n // 2, n // 4] < 0


# -- tape gradients against central finite differences ---------------------
def _fd_grad(loss_fn, x, eps):
    g = np.zeros_like(x)
    it = np.nditer(x, flags=['multi_index'])
    for _ in it:
        i = it.multi_index
        xp = x.copy()
        xm = x.copy()
        xp[i] += eps
        xm[i] -= eps
        g[i] = (loss_fn(xp) - loss_fn(xm)) / (2 * eps)
    return g


def test_wave_loss_gradient_fd():
    dom, op, _ = ops.wave_gaussian_setup(8, 9, seed=3)
    rng = np.random.default_rng(1)
    u = rng.normal(size=(8, 9))
    _, grads, _ = tape.eval_loss_grad(op, dom, {'u': u})
    fd = _fd_grad(lambda x: tape.eval_loss_grad(op, dom, {'u': x})[0], u,
                  1e-6)
    np.testing.assert_allclose(grads['u'], fd, rtol=1e-6,
                               atol=1e-8 * np.abs(fd).max())


def test_cavity_loss_gradient_frozen_fd():
    """Frozen fields are held at the base state while live fields move:
    the tape gradient equals FD of that partially frozen loss."""
    n = 6
    rng = np.random.default_rng(2)
    base = {k: rng.uniform(-1, 1, (n, n)) for k in 'uvp'}
    dom = ops.cavity_domain(n)
    op = ops.make_cavity_operator(0.05)
    _, grads, _ = tape.eval_loss_grad(op, dom, base)

    import torch

    class Split(tape.Context):
        # live values from `live`, frozen values from the base state
        def __init__(self, live, *a, **k):
            super().__init__(*a, **k)
            self.live = live

        def field(self, name, *shift, freeze=False):
            if freeze or name in ('zeros', 'ones'):
                return super().field(name, *shift, freeze=freeze)
            d = self.domain
            shift = tuple(shift) + (0, ) * (d.ndim - len(shift))
            return torch.roll(self.live[name], tuple(-s for s in shift),
                              tuple(range(d.ndim)))

    def loss(name, x):
        live = {k: torch.tensor(base[k]) for k in 'uvp'}
        live[name] = torch.tensor(x)
        ctx = Split(live, dom, {k: torch.tensor(base[k]) for k in 'uvp'},
                    split_shifts=False)
        ff = op(tape.TorchMod, ctx)
        return float(sum(torch.mean(f**2) for f in ff))

    for name in 'uvp':
        fd = _fd_grad(lambda x: loss(name, x), base[name], 1e-6)
        np.testing.assert_allclose(grads[name], fd, rtol=1e-5,
                                   atol=1e-7 * np.abs(fd).max())


def test_advection_gradient_fd():
    dom, op, _ = ops.advection_setup(4, 5, 6, seed=4)
    rng = np.random.default_rng(3)
    p = rng.normal(size=dom.shape)
    _, grads, _ = tape.eval_loss_grad(op, dom, {'p': p})
    fd = _fd_grad(lambda x: tape.eval_loss_grad(op, dom, {'p': x})[0], p,
                  1e-6)
    np.testing.assert_allclose(grads['p'], fd, rtol=1e-6,
                               atol=1e-8 * np.abs(fd).max())


# -- multigrid: transfer operators, Galerkin, V-cycle ------------------------
def test_interpolation_exact_for_linear_fields():
    nw = (17, 9)
    tt = omg.get_TT(nw)
    xc, yc = np.meshgrid(np.arange(9) * 2., np.arange(5) * 2., indexing='ij')
    xf, yf = np.meshgrid(np.arange(17.), np.arange(9.), indexing='ij')
    uc = 3 * xc - 2 * yc + 1
    uf = tt.mul_transpose_field(uc.ravel()).reshape(nw)
    np.testing.assert_allclose(uf, 3 * xf - 2 * yf + 1, rtol=0, atol=1e-12)


def test_restriction_weights():
    R = omg.get_R((9, 9)).tocsr().toarray()
    sums = R.sum(axis=1).reshape(5, 5)
    np.testing.assert_allclose(sums, np.ones((5, 5)))
    # boundary coarse nodes inject
    row = np.ravel_multi_index((0, 2), (5, 5))
    assert R[row, np.ravel_multi_index((0, 4), (9, 9))] == 1.
    assert np.count_nonzero(R[row]) == 1


def test_galerkin_modes_agree_for_symmetric_stencil():
    nw = (11, 11)
    pad = lambda: np.pad(np.ones((9, 9)), 1)
    A = omg.StencilOp({(0, 0): 4 * np.ones(nw), (1, 0): -pad(),
                       (-1, 0): -pad(), (0, 1): -pad(), (0, -1): -pad()}, nw)
    R, TT = omg.get_R(nw), omg.get_TT(nw)
    a = omg.galerkin(A, R, TT, 'exact').tocsr()
    b = omg.galerkin(A, R, TT, 'reference').tocsr()
    assert abs(a - b).max() < 1e-14
    c = R.tocsr() @ A.tocsr() @ TT.tocsr().T
    assert abs(a - c).max() < 1e-14


def test_vcycle_matches_reference_golden():
    g = np.load(os.path.join(GOLD, 'mg_vcycle.npz'))
    n = 17
    stf = {k: g['poisson17_c_{}_{}'.format(*k)]
           for k in [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)]}
    M = omg.MultigridOp((n, n), nvar=1)
    M.update_A([[omg.StencilOp(stf, (n, n))]])
    u = [np.zeros(n * n)]
    rhs = g['poisson17_rhs']
    for it in range(3):
        u = M.step(u, [rhs.copy()], 0.7)
        ref = g['poisson17_u{}'.format(it + 1)]
        assert np.abs(u[0] - ref).max() <= 1e-13 * np.abs(ref).max()

    st = {k: g['cavity17_state_' + k] for k in 'uvp'}
    const, grads, desc = tape.eval_grad(ops.make_cavity_operator(0.01),
                                        ops.cavity_domain(n), st)
    stf = omg.coeffs_from_tape(grads, desc, 3, 3)
    mm, matr = omg.normal_operator(stf, (n, n), 3)
    rhs = [-sum(mm[i][j].mul_transpose_field(const[i].ravel())
                for i in range(3)) for j in range(3)]
    np.testing.assert_allclose(np.array(rhs), g['cavity17_rhs'], rtol=0,
                               atol=1e-10 * np.abs(g['cavity17_rhs']).max())
    M = omg.MultigridOp((n, n), nvar=3, nlevels=4)
    M.update_A(matr)
    sol = [np.zeros(n * n) for _ in range(3)]
    for it in range(3):
        sol = M.step(sol, [r.copy() for r in rhs], 0.6, ndirect=3)
        ref = g['cavity17_u{}'.format(it + 1)]
        for k in range(3):
            d = sol[k] - ref[k]
            if k == 2:  # pressure is defined up to a constant
                d = d - d.mean()
            assert np.abs(d).max() <= 1e-10 * np.abs(ref[k]).max()


def _cavity17_normal(beta2):
    """Stencil blocks of J^T J + beta2 I on the 17x17 cavity golden state."""
    g = np.load(os.path.join(GOLD, 'mg_vcycle.npz'))
    n = 17
    st = {k: g['cavity17_state_' + k] for k in 'uvp'}
    const, grads, desc = tape.eval_grad(ops.make_cavity_operator(0.01),
                                        ops.cavity_domain(n), st)
    stf = omg.coeffs_from_tape(grads, desc, 3, 3)
    mm, matr = omg.normal_operator(stf, (n, n), 3)
    for i in range(3):
        c = dict(matr[i][i].coeffs)
        c[(0, 0)] = c[(0, 0)] + beta2
        matr[i][i] = omg.StencilOp(c, (n, n))
    rhs = [-sum(mm[i][j].mul_transpose_field(const[i].ravel())
                for i in range(3)) for j in range(3)]
    return matr, rhs


def test_sparse_multigrid_equals_stencil_multigrid():
    """The sparse-matrix Multigrid (multigrid.py:72-339) and the stencil
    MultigridOp with exact Galerkin products (multigrid.py:342-602) are
    independent restatements of the same V-cycle when the top level is
    called with the recursion defaults (ndirect 5, pre = post = 2): their
    iterates agree to round-off (the two forms sum the Galerkin products in
    different orders) on the 17x17 Poisson case and on the regularised
    cavity normal equations."""
    n = 17
    dx = 1 / (n - 1)
    pad = lambda: np.pad(np.ones((n - 2, n - 2)), 1) / dx**2
    A = omg.StencilOp({(0, 0): 4 * np.ones((n, n)) / dx**2, (1, 0): -pad(),
                       (-1, 0): -pad(), (0, 1): -pad(), (0, -1): -pad()},
                      (n, n))
    cases = [([[A]], [np.full(n * n, 10.)], 1, None, 0.7)]
    matr, rhs = _cavity17_normal(1e-2)
    cases.append((matr, rhs, 3, 4, 0.6))
    for blocks, f, nvar, nlv, omega in cases:
        M = omg.MultigridOp((n, n), nvar=nvar, nlevels=nlv,
                            galerkin_mode='exact')
        M.update_A(blocks)
        S = omg.Multigrid((n, n), nvar=nvar, nlevels=nlv)
        S.update_A(scipy.sparse.bmat([[blocks[i][j].tocsr()
                                       for j in range(nvar)]
                                      for i in range(nvar)]))
        smooth = lambda level, u, ff: S.smoother_jacobi(level, u, ff, omega)
        u = [np.zeros(n * n) for _ in range(nvar)]
        v = np.zeros(nvar * n * n)
        for _ in range(3):
            u = M.step(u, [x.copy() for x in f], omega)
            v = S.step(v, np.concatenate(f), smooth)
            ref = np.concatenate(u)
            assert np.abs(v - ref).max() <= 1e-10 * np.abs(ref).max()


def _random_block_csr(nw, nvar, reach, seed, shift0_only=False):
    """Random square block stencil matrix on node grid `nw` (fields
    concatenated, C order): random couplings for every shift within `reach`
    (about 70 % of the blocks present), made diagonally dominant."""
    import itertools
    rng = np.random.default_rng(seed)
    nw = tuple(nw)
    N = int(np.prod(nw))
    idx = np.arange(N).reshape(nw)
    zero = (0, ) * len(nw)
    shifts = [zero] if shift0_only else list(
        itertools.product(range(-reach, reach + 1), repeat=len(nw)))
    rows, cols, vals = [], [], []
    for o in range(nvar):
        for i in range(nvar):
            for s in shifts:
                if rng.random() < 0.3 and not (o == i and s == zero):
                    continue
                c = rng.uniform(-1, 1, nw)
                src = tuple(slice(max(0, -d), n - max(0, d))
                            for d, n in zip(s, nw))
                dst = tuple(slice(max(0, d), n - max(0, -d))
                            for d, n in zip(s, nw))
                rows.append(o * N + idx[src].ravel())
                cols.append(i * N + idx[dst].ravel())
                vals.append(c[src].ravel())
    A = scipy.sparse.coo_matrix(
        (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
        shape=(nvar * N, nvar * N)).tocsr()
    dom = np.asarray(abs(A).sum(axis=1)).ravel() + 1.
    return (A + scipy.sparse.diags(dom)).tocsr()


def test_gauss_seidel_oracle_is_the_sequential_sweep():
    """The oracle's Gauss-Seidel (spsolve_triangular of the lower triangle,
    multigrid.py:301-305) equals the textbook forward sweep written as a
    loop over unknowns in flat order: x_k = (f_k - sum_(j<k) A_kj x_j -
    sum_(j>k) A_kj u_j) / A_kk, then u + omega (x - u). A 2-field
    non-symmetric 5x5-stencil matrix, so a transposed triangle, a wrong
    field order or a dropped coupling changes the result."""
    nw = (5, 7)
    A = _random_block_csr(nw, 2, 2, seed=3)
    rng = np.random.default_rng(4)
    u0, f = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    S = omg.Multigrid(nw, nvar=2, nlevels=1)
    S.update_A(A)
    for omega in (1., 0.8):
        got = S.smoother_gauss_seidel(0, u0.copy(), f, omega)
        D = A.toarray()
        x = np.zeros_like(u0)
        for k in range(len(u0)):
            t = f[k] - D[k, :k] @ x[:k] - D[k, k + 1:] @ u0[k + 1:]
            x[k] = t / D[k, k]
        assert np.allclose(got, u0 + omega * (x - u0), rtol=1e-13, atol=0)


def test_block_jacobi_oracle_closed_forms():
    """full=True Jacobi (multigrid.py:289-294): on a matrix whose couplings
    are all at shift 0 (a block diagonal of per-cell nvar x nvar blocks)
    one sweep with omega = 1 is the exact solve A^-1 f from any start; with
    one field it is point Jacobi; a cell whose block is singular raises."""
    nw = (5, 7)
    A = _random_block_csr(nw, 3, 0, seed=5, shift0_only=True)
    rng = np.random.default_rng(6)
    f = rng.standard_normal(A.shape[0])
    S = omg.Multigrid(nw, nvar=3, nlevels=1)
    S.update_A(A)
    got = S.smoother_jacobi(0, rng.standard_normal(A.shape[0]), f, 1.,
                            full=True)
    ref = scipy.sparse.linalg.spsolve(A.tocsc(), f)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14)
    A1 = _random_block_csr(nw, 1, 2, seed=7)
    S1 = omg.Multigrid(nw, nvar=1, nlevels=1)
    S1.update_A(A1)
    u0 = rng.standard_normal(A1.shape[0])
    f1 = rng.standard_normal(A1.shape[0])
    a = S1.smoother_jacobi(0, u0.copy(), f1, 0.7, full=True)
    b = S1.smoother_jacobi(0, u0.copy(), f1, 0.7)
    assert np.allclose(a, b, rtol=1e-14, atol=1e-15)
    Z = A.tolil()
    N = int(np.prod(nw))
    for k in range(3):  # cell 4: block row of field 1 is zero
        Z[N + 4, k * N + 4] = 0.
    S.update_A(Z.tocsr())
    with pytest.raises(np.linalg.LinAlgError):
        S.smoother_jacobi(0, np.zeros(A.shape[0]), f, 1., full=True)


def test_poisson_multigrid_converges():
    """Acceptance of the reference's test_multigrid_poisson
    (test_multigrid.py:251-368): 10 V-cycles, |rhs - A u| <= 1e-6."""
    n = 17
    dx = 1 / (n - 1)
    pad = lambda: np.pad(np.ones((n - 2, n - 2)), 1) / dx**2
    A = omg.StencilOp({(0, 0): 4 * np.ones((n, n)) / dx**2, (1, 0): -pad(),
                       (-1, 0): -pad(), (0, 1): -pad(), (0, -1): -pad()},
                      (n, n))
    M = omg.MultigridOp((n, n))
    M.update_A([[A]])
    rhs = np.full(n * n, 10.)
    u = [np.zeros(n * n)]
    for _ in range(10):
        u = M.step(u, [rhs], 0.7)
    assert np.linalg.norm(rhs - A.mul_field(u[0])) <= 1e-6


def test_normal_operator_is_jtj():
    n = 7
    rng = np.random.default_rng(5)
    st = {k: rng.uniform(-1, 1, (n, n)) for k in 'uvp'}
    const, grads, desc = tape.eval_grad(ops.make_cavity_operator(0.02),
                                        ops.cavity_domain(n), st)
    stf = omg.coeffs_from_tape(grads, desc, 3, 3)
    mm, matr = omg.normal_operator(stf, (n, n), 3)
    J = scipy.sparse.bmat([[mm[i][j].tocsr() for j in range(3)]
                           for i in range(3)])
    N = scipy.sparse.bmat([[matr[i][j].tocsr() for j in range(3)]
                           for i in range(3)])
    assert abs(J.T @ J - N).max() < 1e-9
    # and the CSR Jacobian of `linearize` equals the stencil form
    _, lin = tape.linearize(ops.make_cavity_operator(0.02),
                            ops.cavity_domain(n), st)
    assert abs(lin - J).max() < 1e-12


def test_transpose_restriction_is_scaled_interpolation_transpose():
    for nw in [(9, 17), (5, 9, 9)]:
        R = omg.get_R(nw, 'transpose').tocsr()
        T = omg.get_TT(nw).tocsr()
        assert abs(R - T / 2**len(nw)).max() < 1e-15


# -- inverse wave oracle ----------------------------------------------------
def test_inverse_wave_oracle_pins():
    """c = 1 reproduces the reference wave operator exactly; the forward
    march zeroes the residual; u and c gradients match central FD."""
    import torch
    from oracle import inverse as inv
    case = inv.wave_inverse_setup(24, 20, nsensor=6, seed=1)
    rng = np.random.default_rng(0)
    u = rng.normal(size=(24, 20))
    fw = inv.wave_c_residual(torch.tensor(u), torch.ones(20,
                                                         dtype=torch.float64),
                             case).numpy()
    op = ops.make_wave_operator(case['init_u'], case['init_ut'],
                                case['left_u'], case['right_u'])
    dom = tape.Domain(ndim=2, shape=(24, 20), lower=(0, -1), upper=(1, 1),
                      varnames=('t', 'x'), fieldnames=['u'])
    ff = tape.eval_operator(op, dom, {'u': u})
    np.testing.assert_array_equal(fw, ff[0])
    fw = inv.wave_c_residual(torch.tensor(case['u_true']),
                             torch.tensor(case['c_true']), case).numpy()
    assert np.abs(fw).max() <= 1e-12 * np.abs(case['u_true']).max() * 24**2
    c = case['c_true'] * (1 + 0.1 * rng.normal(size=20))
    _, _, gu, gc = inv.inverse_loss_grad(u, c, case, k_data=3.0, k_reg=0.1)
    du, dc, h = rng.normal(size=u.shape), rng.normal(size=20), 1e-6
    lp = inv.inverse_loss_grad(u + h * du, c + h * dc, case, 3.0, 0.1)[0]
    lm = inv.inverse_loss_grad(u - h * du, c - h * dc, case, 3.0, 0.1)[0]
    fd = (lp - lm) / (2 * h)
    assert abs(np.sum(gu * du) + np.sum(gc * dc) - fd) <= 1e-7 * abs(fd)
    with pytest.raises(ValueError):
        inv.wave_inverse_setup(70, 130)  # explicit march CFL violated


# -- cavity inverse-problem terms (ns_steady.py:396-411) ---------------------
def _imposed(n, rng):
    mask = np.zeros((n, n))
    for i, j in ((1, 2), (3, 3), (n - 1, 0)):
        mask[i, j] = 1.
    return mask, rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, n))


def test_cavity_extras_closed_forms():
    """kreg terms of a quadratic field: u = x^2 has u_xx = 2, u_yy = 0 and
    v = y^2 has v_xx = 0, v_yy = 2 at every cell away from the walls; unit
    misfits on the imposed cells give a loss term alpha^2 N / n_imp."""
    n = 8
    dom = ops.cavity_domain(n)
    xc = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xc, xc, indexing='ij')
    st = {'u': X**2, 'v': Y**2, 'p': np.zeros((n, n))}
    kreg, alpha = 0.3, 2.5
    ff = tape.eval_operator(ops.make_cavity_operator(0.01, kreg=kreg),
                            dom, st)
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(ff[3][inner], 2 * kreg, rtol=1e-10)
    np.testing.assert_allclose(ff[4][inner], 0, atol=1e-9)
    np.testing.assert_allclose(ff[5][inner], 0, atol=1e-9)
    np.testing.assert_allclose(ff[6][inner], 2 * kreg, rtol=1e-10)
    rng = np.random.default_rng(0)
    mask, uimp, vimp = _imposed(n, rng)
    st = {'u': uimp + 1, 'v': vimp - 1, 'p': np.zeros((n, n))}
    ff = tape.eval_operator(
        ops.make_cavity_operator(0.01, imposed=(mask, uimp, vimp),
                                 alpha=alpha), dom, st)
    for f in ff[3:5]:
        np.testing.assert_allclose(np.mean(f**2), alpha**2 * n * n / 3,
                                   rtol=1e-12)


def test_cavity_extras_gradient_fd():
    """The extra terms depend on live fields only: their share of the tape
    gradient (with minus without them) equals FD of their loss."""
    n = 6
    rng = np.random.default_rng(4)
    base = {k: rng.uniform(-1, 1, (n, n)) for k in 'uvp'}
    dom = ops.cavity_domain(n)
    imp = _imposed(n, rng)
    for bc in (0, 1, 2):
        full = ops.make_cavity_operator(0.05, bc=bc, imposed=imp, alpha=0.7,
                                        kreg=0.2)
        plain = ops.make_cavity_operator(0.05, bc=bc)
        _, g1, _ = tape.eval_loss_grad(full, dom, base)
        _, g0, _ = tape.eval_loss_grad(plain, dom, base)

        def xloss(name, x):
            st = dict(base)
            st[name] = x
            ff = tape.eval_operator(full, dom, st)
            return float(sum(np.mean(f**2) for f in ff[3:]))

        for name in 'uvp':
            fd = _fd_grad(lambda x: xloss(name, x), base[name], 1e-6)
            np.testing.assert_allclose(g1[name] - g0[name], fd, rtol=1e-5,
                                       atol=1e-7 * (np.abs(fd).max() + 1))


# -- MultigridDecomp restatement (multigrid.py:1217-1426) --------------------
@pytest.mark.parametrize('shape', [(5, ), (4, 6), (3, 4, 5)])
def test_decomp_conv_equals_manual(shape):
    """The 'conv' and 'manual' forms of the reference are independent
    restatements of one map: they agree (node grids dims 1-3, cell grids
    dims 1-2 as the reference's manual form supports)."""
    from oracle import decomp as od
    rng = np.random.default_rng(len(shape))
    u = rng.standard_normal(shape)
    for depth in (1, 2):
        np.testing.assert_allclose(od.interp_node_conv(u, depth),
                                   od.interp_node_manual(u, depth),
                                   rtol=1e-13, atol=1e-13)
        if len(shape) <= 2:
            np.testing.assert_allclose(od.interp_cell_conv(u, depth),
                                       od.interp_cell_manual(u, depth),
                                       rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize('shape', [(5, ), (4, 6), (3, 4, 5)])
def test_decomp_linear_exact(shape):
    """Both interpolations reproduce linear fields exactly: node values at
    the fine nodes x_f = x_c / 2, cell values at the fine centres."""
    from oracle import decomp as od
    a = np.array([0.3, -1.2, 0.7])[:len(shape)]
    nodes = np.meshgrid(*[np.arange(n, dtype=float) for n in shape],
                        indexing='ij')
    lin = 0.5 + sum(ai * x for ai, x in zip(a, nodes))
    fine_n = np.meshgrid(*[np.arange(2 * n - 1) / 2. for n in shape],
                         indexing='ij')
    np.testing.assert_allclose(od.interp_node_conv(lin),
                               0.5 + sum(ai * x for ai, x in
                                         zip(a, fine_n)), atol=1e-13)
    centres = np.meshgrid(*[np.arange(n) + 0.5 for n in shape],
                          indexing='ij')
    lin = 0.5 + sum(ai * x for ai, x in zip(a, centres))
    fine_c = np.meshgrid(*[(np.arange(2 * n) + 0.5) / 2. for n in shape],
                         indexing='ij')
    np.testing.assert_allclose(od.interp_cell_conv(lin),
                               0.5 + sum(ai * x for ai, x in
                                         zip(a, fine_c)), atol=1e-13)


# -- closed forms of the Laplace, Poisson and advection operators ------------
def _grid2(n0, n1):
    dom = tape.Domain(ndim=2, shape=(n0, n1), varnames=('x', 'y'),
                      fieldnames=['u', 'v'])
    x, y = dom.cell_center_all()
    return dom, x, y


@pytest.mark.parametrize('n0,n1', [(7, 11), (12, 5)])
def test_laplace_closed_form(n0, n1):
    """`operator_laplace` (util_op.py:1038-1079) of quadratic fields: the
    5-point second difference is exact (2a, 2b) in interior cells, and the
    linear-extrapolation halo (util_op.py:1032-1035) makes the normal second
    difference of any field vanish in wall cells. So
    lap(a x^2 + b y^2 + c x y + d x + e y) = 2a [x interior] + 2b [y
    interior] cell by cell, and the third equation is zero."""
    dom, x, y = _grid2(n0, n1)
    u = 3 * x**2 - 5 * y**2 + 0.7 * x * y + 2 * x - y + 1
    v = -x**2 + 4 * y**2 + 1.3 * x * y - 3 * x + 0.5
    ff = tape.eval_operator(ops.operator_laplace, dom, {'u': u, 'v': v})
    ix = np.arange(n0)[:, None] * np.ones((1, n1))
    iy = np.ones((n0, 1)) * np.arange(n1)[None, :]
    inx = ((ix > 0) & (ix < n0 - 1)).astype(float)
    iny = ((iy > 0) & (iy < n1 - 1)).astype(float)
    np.testing.assert_allclose(ff[0], 2 * 3 * inx + 2 * (-5) * iny,
                               atol=1e-9)
    np.testing.assert_allclose(ff[1], 2 * (-1) * inx + 2 * 4 * iny,
                               atol=1e-9)
    np.testing.assert_array_equal(ff[2], 0)
    # A linear field has zero Laplacian everywhere, walls included.
    lin = {'u': 2 * x - 3 * y + 1, 'v': -x + 0.25 * y}
    for f in tape.eval_operator(ops.operator_laplace, dom, lin):
        np.testing.assert_allclose(f, 0, atol=1e-9)


@pytest.mark.parametrize('n0,n1', [(8, 13), (17, 6)])
def test_poisson_closed_form(n0, n1):
    """`make_poisson_operator`: u = x(1-x) y(1-y) vanishes on every wall of
    the unit square and is quadratic along each axis, so both the interior
    second differences and the wall halo (quadratic through the wall value,
    util_op.py:1017-1023) are exact: f = -2 (x(1-x) + y(1-y)) - src in every
    cell, wall cells included."""
    dom, x, y = _grid2(n0, n1)
    dom.fieldnames = ['u']
    src = np.random.default_rng(n0).normal(size=(n0, n1))
    u = x * (1 - x) * y * (1 - y)
    ff = tape.eval_operator(ops.make_poisson_operator(src), dom, {'u': u})
    np.testing.assert_allclose(ff[0], -2 * (x * (1 - x) + y * (1 - y)) - src,
                               atol=1e-10)
    # A field that is quadratic but not zero on the walls shows the wall rule:
    # u = 1 (constant) has halo (1 - 6 + 0)/3 = -5/3 at each wall, so a wall
    # cell's second difference along that axis is (-5/3 - 2 + 1)/h^2.
    ff = tape.eval_operator(ops.make_poisson_operator(0 * src), dom,
                            {'u': np.ones((n0, n1))})
    hx, hy = 1. / n0, 1. / n1
    want = np.zeros((n0, n1))
    want[0, :] += (-5 / 3 - 1) / hx**2
    want[-1, :] += (-5 / 3 - 1) / hx**2
    want[:, 0] += (-5 / 3 - 1) / hy**2
    want[:, -1] += (-5 / 3 - 1) / hy**2
    np.testing.assert_allclose(ff[0], want, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize('ax,ay', [(0.7, -0.4), (-0.9, 0.3), (0.5, 0.8),
                                   (-0.2, -0.6)])
def test_advection_closed_form(ax, ay):
    """Advection residual (advection_2d.py:129-138,200-220) of
    p = beta t + x^2 + y^2 with first-order upwinding on the velocity sign
    s = sign(a): (x_i^2 - x_{i-1}^2)/dx = 2x - dx for a > 0 and
    (x_{i+1}^2 - x_i^2)/dx = 2x + dx otherwise, and (p - p(t - dt))/dt =
    beta exactly, so f = beta + ax (2x - s_x dx) + ay (2y - s_y dy) in every
    cell except on the periodic seam; the t = 0 row uses p_init, here
    p(t_0 - dt). On the seam the upwind neighbour wraps around (x_{-1} :=
    x_{n-1})."""
    nt, nx, ny = 5, 9, 7
    beta = 1.7
    from oracle.tape import Domain
    dom = Domain(ndim=3, shape=(nt, nx, ny), upper=(1., 1., 1.),
                 varnames=('t', 'x', 'y'), fieldnames=['p'])
    t, x, y = dom.cell_center_all()
    dt, dx, dy = 1. / nt, 1. / nx, 1. / ny
    p = beta * t + x**2 + y**2
    p_init = beta * (t[0] - dt) + x[0]**2 + y[0]**2
    ff = tape.eval_operator(ops.make_advection_operator(ax, ay, p_init),
                            dom, {'p': p})[0]
    sx = 1. if ax > 0 else -1.
    sy = 1. if ay > 0 else -1.
    want = beta + ax * (2 * x - sx * dx) + ay * (2 * y - sy * dy)
    seam_x = 0 if ax > 0 else nx - 1
    seam_y = 0 if ay > 0 else ny - 1
    keep = np.ones(p.shape, bool)
    keep[:, seam_x, :] = False
    keep[:, :, seam_y] = False
    np.testing.assert_allclose(ff[keep], want[keep], rtol=1e-12, atol=1e-11)
    # Seam cells: the upwind neighbour is the periodic image.
    xs = x[:, seam_x, :]
    xn = x[:, nx - 1 if ax > 0 else 0, :]
    px = (xs**2 - xn**2) / dx if ax > 0 else (xn**2 - xs**2) / dx
    ys = y[:, :, seam_y]
    yn = y[:, :, ny - 1 if ay > 0 else 0]
    py = (ys**2 - yn**2) / dy if ay > 0 else (yn**2 - ys**2) / dy
    wx = beta + ax * px + ay * (2 * y[:, seam_x, :] -
                                sy * dy)
    sel = np.ones(ny, bool)
    sel[seam_y] = False
    np.testing.assert_allclose(ff[:, seam_x, :][:, sel], wx[:, sel],
                               rtol=1e-12, atol=1e-11)
    wy = beta + ax * (2 * x[:, :, seam_y] - sx * dx) + ay * py
    sel = np.ones(nx, bool)
    sel[seam_x] = False
    np.testing.assert_allclose(ff[:, :, seam_y][:, sel], wy[:, sel],
                               rtol=1e-12, atol=1e-11)


# -- inverse advection (advection_2d.py:224-250) -----------------------------
def test_advection_inverse_closed_form():
    """u = a + b x (a > |b|, so upwinding is backward everywhere),
    v = c y + g t, p = beta t + x^2 + y^2, p_exact = p: off the periodic
    seams fp = beta + u (2x - dx) + v (2y - s_v dy), fe = 0, the kreg
    Laplacians of the linear fields vanish, ktreg gives g ktreg (0 at
    t = 0) for v and 0 for u, and fdiv = divfree (b + c)."""
    nt, nx, ny = 4, 7, 9
    dom = ops.advection_inverse_domain(nt, nx, ny)
    t, x, y = dom.cell_center_all()
    dt, dx, dy = 1. / nt, 1. / nx, 1. / ny
    a, b, c, g, beta = 2.0, 0.5, -0.8, 0.3, 1.7
    u = a + b * x
    v = c * y + g * t
    p = beta * t + x**2 + y**2
    pex = p.copy()
    pex[0] = beta * (t[0] - dt) + x[0]**2 + y[0]**2  # previous level at t0
    op = ops.make_advection_inverse_operator(pex, alpha=1.3, kreg=0.4,
                                             ktreg=0.6, divfree=0.9)
    fe, fp, fsu, fsv, ftu, ftv, fdiv = tape.eval_operator(
        op, dom, {'u': u, 'v': v, 'p': p})
    sv = np.where(v > 0, 1., -1.)
    keep = np.ones(p.shape, bool)
    keep[:, 0, :] = keep[:, -1, :] = False
    keep[:, :, 0] = keep[:, :, -1] = False
    want = beta + u * (2 * x - dx) + v * (2 * y - sv * dy)
    np.testing.assert_allclose(fp[keep], want[keep], rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(fe[1:], 0, atol=1e-13)
    np.testing.assert_allclose(fe[0], 1.3 * beta * dt, rtol=1e-12)
    np.testing.assert_allclose(fsu[keep], 0, atol=1e-9)
    np.testing.assert_allclose(fsv[keep], 0, atol=1e-9)
    np.testing.assert_allclose(ftu, 0, atol=1e-12)
    np.testing.assert_allclose(ftv[1:], 0.6 * g, rtol=1e-12)
    np.testing.assert_array_equal(ftv[0], 0)
    np.testing.assert_allclose(fdiv[keep], 0.9 * (b + c), rtol=1e-12)


def test_advection_inverse_matches_direct_operator():
    """With a constant velocity field the inverse operator's fp equals the
    residual of the direct advection oracle (pinned above) whose initial
    level is p_exact at t = 0."""
    dom_i = ops.advection_inverse_domain(5, 6, 7)
    dom_d, op_d, (ax, ay, p0) = ops.advection_setup(5, 6, 7, seed=2)
    rng = np.random.default_rng(1)
    p = rng.normal(size=dom_i.shape)
    pex = rng.normal(size=dom_i.shape)
    pex[0] = p0
    op_i = ops.make_advection_inverse_operator(pex, kreg=0.)
    ff = tape.eval_operator(op_i, dom_i, {
        'u': np.full(dom_i.shape, ax), 'v': np.full(dom_i.shape, ay),
        'p': p})
    fd = tape.eval_operator(op_d, dom_d, {'p': p})[0]
    np.testing.assert_allclose(ff[1], fd, rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(ff[0], p - pex, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize('periodic,onlyfinal', [(True, False), (False, True)])
def test_advection_inverse_gradient_fd(periodic, onlyfinal):
    """Tape gradient (frozen upwind direction) against central differences:
    velocities kept away from zero so the direction does not flip."""
    dom = ops.advection_inverse_domain(3, 5, 4)
    rng = np.random.default_rng(7)
    sgn = np.where(rng.uniform(size=dom.shape) < 0.5, -1., 1.)
    st = {'u': sgn * rng.uniform(0.5, 1.5, dom.shape),
          'v': -sgn * rng.uniform(0.5, 1.5, dom.shape),
          'p': rng.normal(size=dom.shape)}
    op = ops.make_advection_inverse_operator(
        rng.normal(size=dom.shape), alpha=0.7, anneal_half=3.,
        onlyfinal=onlyfinal, kreg=0.2, ktreg=0.5, divfree=0.4,
        periodic=periodic)
    _, grads, _ = tape.eval_loss_grad(op, dom, st, epoch=2)

    def loss(name, x):
        s = dict(st)
        s[name] = x
        return tape.eval_loss_grad(op, dom, s, epoch=2)[0]

    for name in 'uvp':
        fd = _fd_grad(lambda x: loss(name, x), st[name], 1e-6)
        np.testing.assert_allclose(grads[name], fd, rtol=1e-6,
                                   atol=1e-7 * np.abs(fd).max())
