"""Normal-equation solves of the Newton step on the device
(`linsolver.solve`, linsolver.py:28-107).

`solve(J, rhs, args, dhistory, linsolver)` returns dx minimising
|J dx - rhs|^2 (+ regularisation) for a stencil Jacobian J (`BlockStencil`,
equations x fields) and residual-shaped rhs fields. The reference's choices,
all on the GPU:

* 'direct' / 'direct_cu' (SuperLU / cuSolver `spsolve` of
  A = J^T J + beta^2 I + betadiag^2 diag, linsolver.py:35-53): A in stencil
  form (`odil_stencil_normal`); up to `max_dense` unknowns its dense FP64
  matrix (`odil_stencil_to_dense`) and a Cholesky factorisation, above that
  a sparse direct solve: the unknowns interleaved per cell make A
  block-tridiagonal with blocks of nvar (max stencil shift + 1) rows, which
  a block Cholesky factors in n B^2 work (the reference's N = 128 cavity,
  49 152 unknowns, B = 774). Singular normal equations (the cavity's
  constant pressure) are solved with a 1e-14 max(diag) Tikhonov shift, the
  minimum-norm-like answer (SuperLU's null component is set by round-off).
* 'lsqr' / 'lsqr_cu' (SciPy / CuPy `lsqr` on J itself with damp = beta,
  atol = btol = linsolver_tol, linsolver.py:57-75): the Paige-Saunders
  LSQR bidiagonalisation with J and J^T applied by the stencil kernels and
  the recurrences in `odil_axpby` / `odil_dot`; records the residual norm
  of the normal equations, the estimates of |J| and cond(J) and the
  iteration count like the reference.
* 'bicgstab' (SciPy `bicgstab` on A with tol = 0, atol = linsolver_tol,
  linsolver.py:88-101): van der Vorst's BiCGSTAB on the stored symmetric A.
* 'multigrid' (PyAMG smoothed aggregation as a CG preconditioner,
  tol = linsolver_tol relative to |b|, linsolver.py:77-87): CG on the stored
  A preconditioned by one geometric V(smooth_pre, smooth_post) cycle
  (`odil_amd.multigrid.MultigridOp`: Galerkin levels, weighted Jacobi,
  dense coarsest solve). Node grids of odd size only (2^k + 1 for full
  depth). The Newton driver with matrix-free operators and captured graphs
  is `odil_amd.newton.NewtonMultigrid`.

'sparseqr' (SuiteSparseQR) has no GPU counterpart here and raises
ValueError.
"""

import ctypes
import math
from argparse import Namespace

import numpy as np
import torch

from odil_amd import _lib
from odil_amd._lib import c_double, c_int, c_vp, check, ptr, ptrs, stream
from odil_amd.multigrid import BlockStencil, MultigridOp, normal_operator


def _arg(args, name, default):
    if args is None:
        return default
    v = getattr(args, name, None)
    return default if v is None else v


class _Vec:
    """FP64 device vectors of k fields stored as one (k, *grid) buffer; the
    recurrences run in `odil_axpby`, the norms in `odil_dot` (fixed-order
    FP64 sums)."""

    def __init__(self, device):
        self.partial = torch.zeros(_lib.lib().odil_dot_partials(),
                                   dtype=torch.float64, device=device)
        self.out = torch.zeros(2, dtype=torch.float64, device=device)

    def dot(self, x, y):
        check(_lib.lib().odil_dot(
            c_int(_lib.dtype_tag(x.dtype)), c_int(1),
            ctypes.c_longlong(x.numel()), ptrs([x]), ptrs([y]),
            c_vp(self.partial.data_ptr()), c_vp(self.out.data_ptr()),
            c_int(0), stream()), 'dot')
        return self.out[0].item()

    def norm(self, x):
        return math.sqrt(self.dot(x, x))

    def axpby(self, out, x, y, a, b):
        """out = a x + b y (out may alias x or y)."""
        check(_lib.lib().odil_axpby(
            c_int(_lib.dtype_tag(x.dtype)), ctypes.c_longlong(x.numel()),
            ptr(out), ptr(x), ptr(y), c_double(a), c_double(b), stream()),
            'axpby')
        return out

    def scale(self, x, a):
        return self.axpby(x, x, x, a, 0.)


def _stack(fields):
    return torch.stack([f.to(torch.float64) for f in fields]).contiguous()


def _vector(b, nvar, nw, what):
    """b as a contiguous FP64 (nvar, *nw) device tensor (ValueError if the
    shape differs)."""
    shape = (nvar, ) + tuple(nw)
    if tuple(b.shape) != shape:
        raise ValueError('{}: b has shape {}, expected {}'.format(
            what, tuple(b.shape), shape))
    return b.to(torch.float64).contiguous()


def _regularised_normal(matr, beta, betadiag):
    """A = J^T J (stored upper half, FP64) with the diagonal regularisation
    of linsolver.py:37-40: (diag + beta^2) (1 + betadiag^2)."""
    N = normal_operator(matr, torch.float64, sym=True)
    if beta or betadiag:
        for d in N.diagonal():
            d.add_(beta**2).mul_(1. + betadiag**2)
    return N


def lsqr(A, b, damp=0., atol=1e-6, btol=1e-6, conlim=1e8, iter_lim=None):
    """LSQR (Paige and Saunders, ACM TOMS 8:43-71, 1982) for
    min |A x - b|^2 + damp^2 |x|^2 with a stencil operator A (fields in,
    equations out) and b stacked as (neq, *grid). Returns
    (x, istop, itn, r1norm, r2norm, anorm, acond, arnorm, xnorm) with x
    stacked as (nvar, *grid); the stopping rules and estimates are those of
    the paper (istop 1-2: residual / least-squares tolerance met, 3: cond
    limit, 4-6: the same at machine precision, 7: iteration limit)."""
    b = _vector(b, A.nvar_out, A.nwh, 'lsqr')
    vec = _Vec(b.device)
    nvar = A.nvar_in
    n = nvar * int(np.prod(A.nw))
    if iter_lim is None:
        iter_lim = 2 * n
    eps = np.finfo(np.float64).eps
    ctol = 1. / conlim if conlim > 0 else 0.
    shape_x = (nvar, ) + tuple(A.nw)
    x = torch.zeros(shape_x, dtype=torch.float64, device=b.device)
    u = b.clone()
    v = torch.empty_like(x)
    tmp_u = torch.empty_like(u)
    tmp_v = torch.empty_like(x)
    beta = vec.norm(u)
    alpha = 0.
    if beta > 0:
        vec.scale(u, 1. / beta)
        A.apply_t(list(u), y=list(v))
        alpha = vec.norm(v)
    if alpha > 0:
        vec.scale(v, 1. / alpha)
    w = v.clone()
    anorm = acond = ddnorm = res2 = xnorm = xxnorm = z = 0.
    cs2, sn2 = -1., 0.
    dampsq = damp**2
    rhobar, phibar = alpha, beta
    bnorm = rnorm = r1norm = r2norm = beta
    arnorm = alpha * beta
    istop, itn = 0, 0
    if arnorm == 0:
        return x, istop, itn, r1norm, r2norm, anorm, acond, arnorm, xnorm
    while itn < iter_lim:
        itn += 1
        # Bidiagonalisation: beta u = A v - alpha u, alpha v = A^T u - beta v.
        A.apply(list(v), y=list(tmp_u))
        vec.axpby(u, tmp_u, u, 1., -alpha)
        beta = vec.norm(u)
        if beta > 0:
            vec.scale(u, 1. / beta)
            anorm = math.sqrt(anorm**2 + alpha**2 + beta**2 + dampsq)
            A.apply_t(list(u), y=list(tmp_v))
            vec.axpby(v, tmp_v, v, 1., -beta)
            alpha = vec.norm(v)
            if alpha > 0:
                vec.scale(v, 1. / alpha)
        # Plane rotation eliminating the damping term, then the one that
        # eliminates the subdiagonal beta of the bidiagonal matrix.
        if damp > 0:
            rhobar1 = math.sqrt(rhobar**2 + dampsq)
            cs1, sn1 = rhobar / rhobar1, damp / rhobar1
            psi = sn1 * phibar
            phibar = cs1 * phibar
        else:
            rhobar1, psi = rhobar, 0.
        rho = math.sqrt(rhobar1**2 + beta**2)
        cs, sn = rhobar1 / rho, beta / rho
        theta = sn * alpha
        rhobar = -cs * alpha
        phi = cs * phibar
        phibar = sn * phibar
        tau = sn * phi
        # x += (phi / rho) w; w = v - (theta / rho) w; |w / rho|^2 feeds
        # the condition estimate.
        ddnorm += vec.dot(w, w) / rho**2
        vec.axpby(x, x, w, 1., phi / rho)
        vec.axpby(w, v, w, 1., -theta / rho)
        # Norm of x from the rotations of the lower bidiagonal system.
        delta = sn2 * rho
        gambar = -cs2 * rho
        rhs = phi - delta * z
        zbar = rhs / gambar
        xnorm = math.sqrt(xxnorm + zbar**2)
        gamma = math.sqrt(gambar**2 + theta**2)
        cs2, sn2 = gambar / gamma, theta / gamma
        z = rhs / gamma
        xxnorm += z**2
        acond = anorm * math.sqrt(ddnorm)
        res2 += psi**2
        rnorm = math.sqrt(phibar**2 + res2)
        arnorm = alpha * abs(tau)
        r1sq = rnorm**2 - dampsq * xxnorm
        r1norm = math.copysign(math.sqrt(abs(r1sq)), r1sq)
        r2norm = rnorm
        test1 = rnorm / bnorm
        test2 = arnorm / (anorm * rnorm + eps)
        test3 = 1. / (acond + eps)
        t1 = test1 / (1. + anorm * xnorm / bnorm)
        rtol = btol + atol * anorm * xnorm / bnorm
        if itn >= iter_lim:
            istop = 7
        if 1. + test3 <= 1.:
            istop = 6
        if 1. + test2 <= 1.:
            istop = 5
        if 1. + t1 <= 1.:
            istop = 4
        if test3 <= ctol:
            istop = 3
        if test2 <= atol:
            istop = 2
        if test1 <= rtol:
            istop = 1
        if istop:
            break
    return x, istop, itn, r1norm, r2norm, anorm, acond, arnorm, xnorm


def bicgstab(A, b, atol=1e-6, maxiter=50, callback=None):
    """BiCGSTAB (van der Vorst, SIAM J. Sci. Stat. Comput. 13:631-644, 1992)
    on a square stencil operator A from x = 0, unpreconditioned, stopping
    when |r| < atol (relative tolerance 0, as the reference passes).
    `callback(x)` after every iteration. Returns (x, info): info 0 when
    converged, else the iterations run."""
    b = _vector(b, A.nvar_out, A.nw, 'bicgstab')
    vec = _Vec(b.device)
    x = torch.zeros_like(b)
    r = b.clone()
    rt = b.clone()
    p = torch.zeros_like(b)
    v = torch.zeros_like(b)
    t = torch.empty_like(b)
    rho_prev = alpha = omega = 1.
    if vec.norm(r) < atol:
        return x, 0
    for it in range(maxiter):
        rho = vec.dot(rt, r)
        if rho == 0.:
            return x, -10
        if it == 0:
            p.copy_(r)
        else:
            # p = r + (rho / rho_prev) (alpha / omega) (p - omega v)
            bb = (rho / rho_prev) * (alpha / omega)
            vec.axpby(p, p, v, 1., -omega)
            vec.axpby(p, r, p, 1., bb)
        A.apply(list(p), y=list(v))
        alpha = rho / vec.dot(rt, v)
        vec.axpby(r, r, v, 1., -alpha)  # r holds s = r - alpha v
        if vec.norm(r) < atol:
            vec.axpby(x, x, p, 1., alpha)
            if callback is not None:
                callback(x)
            return x, 0
        A.apply(list(r), y=list(t))
        omega = vec.dot(t, r) / vec.dot(t, t)
        vec.axpby(x, x, p, 1., alpha)
        vec.axpby(x, x, r, 1., omega)
        vec.axpby(r, r, t, 1., -omega)
        if callback is not None:
            callback(x)
        if vec.norm(r) < atol:
            return x, 0
        if omega == 0.:
            return x, -11
        rho_prev = rho
    return x, maxiter


def _mg_levels(nw, nlvl):
    """Levels of a node hierarchy on `nw`: halve while every size is odd
    and above 3 (at most nlvl levels when given)."""
    levels, n = 1, list(nw)
    while all(k % 2 == 1 and k > 3 for k in n):
        if nlvl and levels >= nlvl:
            break
        n = [(k - 1) // 2 + 1 for k in n]
        levels += 1
    return levels


def pcg_multigrid(A, b, tol=1e-6, maxiter=50, nlvl=None, omega=0.6,
                  pre=2, post=2, ndirect=3):
    """CG on the stored symmetric A preconditioned by one V(pre, post) cycle
    from zero per iteration; stops when |r| <= tol |b|. Returns
    (x, residual norms including the initial one)."""
    b = _vector(b, A.nvar_out, A.nw, 'pcg_multigrid')
    vec = _Vec(b.device)
    if any(k % 2 == 0 for k in A.nw):
        raise ValueError("linsolver='multigrid' needs odd node grids "
                         '(2^k + 1 for full depth), got {}'.format(A.nw))
    mg = MultigridOp(A.nw, nvar=A.nvar_in, nlevels=_mg_levels(A.nw, nlvl),
                     restriction='transpose', dtype=torch.float64,
                     device=b.device, coarse_reg=1e-14,
                     galerkin_mode='exact')
    mg.update_A(A)
    mg.prepare(ndirect)
    x = torch.zeros_like(b)
    r = b.clone()
    z = torch.zeros_like(b)
    q = torch.empty_like(b)
    bnorm = vec.norm(b)
    res = [bnorm]
    if bnorm == 0.:
        return x, res
    zl = list(z)
    mg.vcycle(zl, list(r), omega=omega, ndirect=ndirect, pre=pre,
              post=post, zero=True)
    p = z.clone()
    rz = vec.dot(r, z)
    for _ in range(maxiter):
        A.apply(list(p), y=list(q))
        a = rz / vec.dot(p, q)
        vec.axpby(x, x, p, 1., a)
        vec.axpby(r, r, q, 1., -a)
        res.append(vec.norm(r))
        if res[-1] <= tol * bnorm:
            break
        mg.vcycle(zl, list(r), omega=omega, ndirect=ndirect, pre=pre,
                  post=post, zero=True)
        rz_new = vec.dot(r, z)
        vec.axpby(p, z, p, 1., rz_new / rz)
        rz = rz_new
    return x, res


# Relative Tikhonov shift of a retried direct solve (singular normal
# equations: the cavity's constant-pressure mode).
DIRECT_SHIFT = 1e-14


def _normal_apply_reg(N, beta, betadiag):
    """x -> A x with A = J^T J + beta^2 I, its diagonal scaled by
    1 + betadiag^2 (linsolver.py:36-40), from the stored normal operator."""
    diag = torch.stack(N.diagonal()) if betadiag else None

    def apply(x):
        y = torch.stack(N.apply([a.contiguous() for a in x]))
        if beta:
            y += beta**2 * x
        if betadiag:
            y += betadiag**2 * (diag + beta**2) * x
        return y
    return apply


def _refine(apply, solve, b, x, passes=2):
    """FP64 iterative refinement x += M^-1 (b - A x) with M the shifted
    factorisation: removes the O(shift / lambda) bias the Tikhonov shift of
    a singular system leaves on the smoothest modes."""
    for _ in range(passes):
        x = x + solve(b - apply(x))
    return x


def _direct_dense(matr, b, beta, betadiag):
    """Dense Cholesky of J^T J (+beta^2 I, diagonal scaled by 1 +
    betadiag^2, linsolver.py:35-45); if the matrix is singular (Cholesky
    breaks down) it is factored with DIRECT_SHIFT * max(diag) added and the
    solution refined against the unshifted matrix."""
    nvar = matr.nvar_in
    N = normal_operator(matr, torch.float64, sym=True)
    M = N.to_dense()
    if beta:
        M.diagonal().add_(beta**2)
    if betadiag:  # diag of the beta-regularised matrix (linsolver.py:39)
        M.diagonal().mul_(1. + betadiag**2)
    L, info = torch.linalg.cholesky_ex(M)
    shifted = int(info.item()) != 0
    if shifted:
        M.diagonal().add_(DIRECT_SHIFT * float(M.diagonal().max()))
        L, info = torch.linalg.cholesky_ex(M)
        if int(info.item()) != 0:
            raise ValueError('normal equations not positive semi-definite')
    shape = (nvar, ) + tuple(matr.nw)
    solve = lambda r: torch.cholesky_solve(r.reshape(-1, 1), L).reshape(
        shape)
    bb = torch.stack(list(b))
    x = solve(bb)
    if shifted:
        x = _refine(_normal_apply_reg(N, beta, betadiag), solve, bb, x)
    return x


def _band_blocks(N, nvar, beta, betadiag):
    """Block-tridiagonal form of the stored (full) normal operator N with
    the unknowns interleaved cell by cell (index nvar c + field): a
    coupling of flat cell shift s lies within nvar (|s| + 1) of the
    diagonal, so blocks of B = nvar (max|s| + 1) rows couple only to their
    neighbours. Returns (D [nb][B][B] diagonal blocks, Lo [nb-1][B][B]
    blocks below the diagonal, B, n)."""
    nw = N.nw
    ncell = int(np.prod(nw))
    strides = [int(np.prod(nw[d + 1:])) for d in range(len(nw))]
    dev = N.coef.device
    smax = max(abs(sum(a * b for a, b in zip(sh, strides)))
               for _, _, sh in N.entries)
    B = nvar * (smax + 1)
    n = nvar * ncell
    nb = (n + B - 1) // B
    D = torch.zeros((nb, B, B), dtype=torch.float64, device=dev)
    Lo = torch.zeros((max(nb - 1, 1), B, B), dtype=torch.float64,
                     device=dev)
    cells = torch.arange(ncell, device=dev)
    idx = []
    rem = cells
    for d in range(len(nw)):
        idx.append(rem // strides[d])
        rem = rem % strides[d]
    for q, (o, i, sh) in enumerate(N.entries):
        ok = torch.ones(ncell, dtype=torch.bool, device=dev)
        for d, sd in enumerate(sh):
            ok &= (idx[d] + sd >= 0) & (idx[d] + sd < nw[d])
        c = cells[ok]
        sf = sum(a * b for a, b in zip(sh, strides))
        rows = nvar * c + o
        cols = nvar * (c + sf) + i
        vals = N.coef[q].reshape(-1)[ok].double()
        rb, cb = rows // B, cols // B
        on = rb == cb
        D.index_put_((rb[on], rows[on] % B, cols[on] % B), vals[on],
                     accumulate=True)
        lo = cb == rb - 1
        Lo.index_put_((rb[lo] - 1, rows[lo] % B, cols[lo] % B), vals[lo],
                      accumulate=True)
    diag = D.diagonal(dim1=1, dim2=2)
    if beta:
        diag.add_(beta**2)
    if betadiag:
        diag.mul_(1. + betadiag**2)
    # Rows past the last unknown (padding of the last block): identity.
    pad = nb * B - n
    if pad:
        D[-1, B - pad:, B - pad:].diagonal().fill_(1.)
    return D, Lo, B, n


def _direct_banded(matr, b, beta, betadiag):
    """Sparse direct solve of J^T J x = b: block Cholesky of the
    block-tridiagonal interleaved form (`_band_blocks`),
    L_kk L_kk^T = D_k - L_k,k-1 L_k,k-1^T, L_k+1,k = Lo_k L_kk^-T, then the
    two block substitutions. Work n B^2 and memory 2 n B (B ~ nvar *
    reach * row length) instead of the dense n^3 and n^2. A singular system
    is factored with DIRECT_SHIFT * max(diag) on the diagonal and the
    solution refined against the unshifted operator."""
    nvar = matr.nvar_in
    N = normal_operator(matr, torch.float64, sym=False)
    D, Lo, B, n = _band_blocks(N, nvar, beta, betadiag)
    nb = D.shape[0]
    shift = DIRECT_SHIFT * float(D.diagonal(dim1=1, dim2=2).abs().max())
    for attempt in range(2):
        Lk, Lsub = [], []
        ok = True
        for k in range(nb):
            A = D[k].clone()
            if attempt:
                A.diagonal().add_(shift)
            if k:
                A -= Lsub[-1] @ Lsub[-1].T
            L, info = torch.linalg.cholesky_ex(A)
            if int(info.item()) != 0:
                ok = False
                break
            Lk.append(L)
            if k + 1 < nb:
                # L_{k+1,k} = Lo_k L_kk^-T  <=>  L_kk L_{k+1,k}^T = Lo_k^T
                Lsub.append(torch.linalg.solve_triangular(
                    L, Lo[k].T, upper=False).T)
        if ok:
            break
    else:
        raise ValueError('normal equations not positive semi-definite')
    shape = (nvar, ) + tuple(matr.nw)

    def solve(r):
        # fields (nvar, *nw) -> interleaved, padded blocks and back
        rr = r.reshape(nvar, -1).T.reshape(-1)
        rr = torch.cat([rr, rr.new_zeros(nb * B - n)]).view(nb, B, 1)
        y = []
        for k in range(nb):
            t = rr[k] if not k else rr[k] - Lsub[k - 1] @ y[-1]
            y.append(torch.linalg.solve_triangular(Lk[k], t, upper=False))
        x = [None] * nb
        for k in reversed(range(nb)):
            t = y[k] if k + 1 == nb else y[k] - Lsub[k].T @ x[k + 1]
            x[k] = torch.linalg.solve_triangular(Lk[k].T, t, upper=True)
        sol = torch.cat(x).reshape(-1)[:n]
        return sol.view(-1, nvar).T.contiguous().reshape(shape)

    bb = torch.stack(list(b))
    x = solve(bb)
    if attempt:
        x = _refine(_normal_apply_reg(N, beta, betadiag), solve, bb, x)
    return x


def solve(matr, rhs, args=None, dhistory=None, linsolver='direct',
          max_dense=16384):
    """Returns the correction fields dx (list of device tensors, FP64).

    matr: `BlockStencil` Jacobian (equations x fields), FP64.
    rhs: list of residual-shaped device tensors (one per equation), usually
         -f as in `optimize_newton` (util.py:261-267).
    args: optional namespace with `beta`, `betadiag`, `direct_max_dense`
          (largest system given to the dense direct path, default
          `max_dense`), `linsolver_maxiter`
          (default 1000 for 'lsqr', else 50, linsolver.py:29-33),
          `linsolver_tol`, `nlvl`, `omega`, `smooth_pre`, `smooth_post`,
          `ndirect` (linsolver.py:108-170).
    dhistory: optional dict receiving `linsolver_niter` and, for the
          iterative solvers, `linsolver_residual` (and `linsolver_anorm`,
          `linsolver_acond` for 'lsqr') as in the reference."""
    if not isinstance(matr, BlockStencil):
        raise ValueError('solve expects a BlockStencil Jacobian')
    beta = _arg(args, 'beta', 0.)
    betadiag = _arg(args, 'betadiag', 0.)
    tol = _arg(args, 'linsolver_tol', 1e-6)
    lsqr_like = linsolver in ('lsqr', 'lsqr_cu')
    maxiter = _arg(args, 'linsolver_maxiter', 1000 if lsqr_like else 50)
    nvar = matr.nvar_in
    dh = dhistory if dhistory is not None else {}
    if lsqr_like:
        out = lsqr(matr, _stack(rhs), damp=beta, atol=tol, btol=tol,
                   iter_lim=maxiter)
        x, _, itn, _, _, anorm, acond, arnorm = out[:8]
        dh['linsolver_residual'] = arnorm
        dh['linsolver_anorm'] = anorm
        dh['linsolver_acond'] = acond
        dh['linsolver_niter'] = itn
        return [x[k].contiguous() for k in range(nvar)]
    if linsolver == 'sparseqr':
        raise ValueError("linsolver='sparseqr' (SuiteSparseQR) is not "
                         "provided; 'lsqr' solves the same least-squares "
                         'problem on the GPU')
    if linsolver not in ('direct', 'direct_cu', 'bicgstab', 'multigrid'):
        raise ValueError('Unknown linsolver=' + str(linsolver))
    b = matr.apply_t(rhs)
    if linsolver in ('direct', 'direct_cu'):
        n = nvar * int(np.prod(matr.nw))
        if n <= _arg(args, 'direct_max_dense', max_dense):
            sol = _direct_dense(matr, b, beta, betadiag)
        else:
            sol = _direct_banded(matr, b, beta, betadiag)
        dh['linsolver_niter'] = 1
        return [sol[k].contiguous() for k in range(nvar)]
    A = _regularised_normal(matr, beta, betadiag)
    bb = _stack(b)
    if linsolver == 'bicgstab':
        vec = _Vec(bb.device)
        res = []
        tmp = torch.empty_like(bb)

        def callback(x):
            # RMS of A x - b (linsolver.py:90-92)
            A.residual(list(x), list(bb), y=list(tmp))
            res.append(vec.norm(tmp) / math.sqrt(tmp.numel()))
        x, _ = bicgstab(A, bb, atol=tol, maxiter=maxiter, callback=callback)
        if res:
            dh['linsolver_residual'] = res[-1]
        dh['linsolver_niter'] = len(res)
        return [x[k].contiguous() for k in range(nvar)]
    x, res = pcg_multigrid(
        A, bb, tol=tol, maxiter=maxiter, nlvl=_arg(args, 'nlvl', None),
        omega=_arg(args, 'omega', 0.6), pre=_arg(args, 'smooth_pre', 2),
        post=_arg(args, 'smooth_post', 2), ndirect=_arg(args, 'ndirect', 3))
    dh['linsolver_residual'] = res[-1]
    dh['linsolver_niter'] = len(res)
    return [x[k].contiguous() for k in range(nvar)]


def add_arguments(parser):
    """Solver options with the reference's names and defaults
    (linsolver.py:108-170)."""
    parser.add_argument('--linsolver', default='direct',
                        choices=['multigrid', 'direct', 'direct_cu',
                                 'sparseqr', 'lsqr', 'lsqr_cu', 'bicgstab'])
    parser.add_argument('--linsolver_maxiter', type=int, default=None)
    parser.add_argument('--linsolver_tol', type=float, default=1e-6)
    parser.add_argument('--beta', type=float, default=0.)
    parser.add_argument('--betadiag', type=float, default=0.)
    parser.add_argument('--linsolver_verbose', type=int, default=0)
    parser.add_argument('--nlvl', type=int, default=5)
    parser.add_argument('--omega', type=float, default=0.6)
    parser.add_argument('--smooth_pre', type=int, default=2)
    parser.add_argument('--smooth_post', type=int, default=2)
    parser.add_argument('--ndirect', type=int, default=3)
    parser.add_argument('--restriction', default='full',
                        choices=('full', 'half', 'injection', 'transpose'))
    return parser


def default_args(**kw):
    a = Namespace(linsolver='direct', linsolver_maxiter=None,
                  linsolver_tol=1e-6, beta=0., betadiag=0.,
                  linsolver_verbose=0, nlvl=5, omega=0.6, smooth_pre=2,
                  smooth_post=2, ndirect=3, restriction='full')
    for k, v in kw.items():
        setattr(a, k, v)
    return a
