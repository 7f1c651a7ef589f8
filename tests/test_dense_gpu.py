"""Blocked Cholesky solver of the coarsest multigrid level against NumPy."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def dev():
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch.device('cuda')


def _spd(rng, n, cond):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.logspace(0, -np.log10(cond), n)
    return (Q * lam) @ Q.T


@pytest.mark.parametrize('n', [1, 5, 64, 65, 130, 300])
def test_spd_solve_matches_numpy(dev, n):
    from odil_amd.dense import SpdSolver
    rng = np.random.default_rng(n)
    s = SpdSolver(n, dev)
    # Graph replays with fresh operands, eager work and syncs in between.
    for rep in range(3):
        G = _spd(rng, n, 1e6)
        B = rng.standard_normal((n, n))
        X = s.solve(torch.tensor(G, device=dev), torch.tensor(B, device=dev))
        torch.cuda.synchronize()
        ref = np.linalg.solve(G, B)
        err = np.abs(X.cpu().numpy() - ref).max() / np.abs(ref).max()
        # G has condition 1e6: two backward-stable FP64 solves agree to
        # ~kappa eps = 2e-10; asserted with a factor 100 for the growth.
        assert err < 100 * 1e6 * np.finfo(np.float64).eps, (n, rep, err)
        _ = torch.ones(1000, device=dev).sum().item()


def test_spd_solve_tikhonov_coarse(dev):
    """(M^T M + d I)^-1 M^T for a symmetric singular M (constant null mode,
    like the cavity's coarsest level) with d = 1e-12 max diag, against the
    FP64 CPU result."""
    from odil_amd.dense import SpdSolver
    rng = np.random.default_rng(7)
    n = 867
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([np.logspace(0, -4, n - 1), [0.]])
    M = (Q * lam) @ Q.T                      # rank n - 1, symmetric
    G = M.T @ M
    G[np.diag_indices(n)] += 1e-12 * G.diagonal().max()
    ref = np.linalg.solve(G, M.T)
    s = SpdSolver(n, dev)
    X = s.solve(torch.tensor(G, device=dev),
                torch.tensor(M.T.copy(), device=dev)).cpu().numpy()
    # The operator is applied to right-hand sides; compare its action off
    # the null vector v (round-off along v is amplified by 1/d in any
    # solver, as for the cavity's constant pressure).
    b = rng.standard_normal(n)
    v = Q[:, -1]
    proj = lambda y: y - v * (v @ y)
    assert np.abs(proj(X @ b) - proj(ref @ b)).max() <= \
        1e-6 * np.abs(proj(ref @ b)).max()
    # and it is a least-squares solution operator: M (X b) ~ projection of b
    r1 = M @ (X @ b) - b
    r2 = M @ (ref @ b) - b
    assert abs(np.linalg.norm(r1) - np.linalg.norm(r2)) <= \
        1e-6 * np.linalg.norm(b)


def test_spd_solve_rejects_bad_shape(dev):
    from odil_amd.dense import SpdSolver
    s = SpdSolver(8, dev)
    with pytest.raises(ValueError):
        s.solve(torch.zeros(8, 9, dtype=torch.float64, device=dev),
                torch.zeros(8, 8, dtype=torch.float64, device=dev))
