"""CPU oracle: stencil sparse operators and the geometric multigrid V-cycle.

TEST INFRASTRUCTURE ONLY (see `oracle/tape.py`).

Restates `SparseOperator`, `MultigridOp` and the matrix-free Newton driver of
the reference (`multigrid.py:342-602,877-1214`, `util.py:509-604`) with NumPy
and SciPy. A stencil operator maps a field `u` on a grid of `nw` points to
`(A u)[I] = sum_s c_s[I] * u[stride*I + s]` with zero fill outside the grid
(`SparseOperator.mul_field` + `apply_shift`, `multigrid.py:941-969`).
Galerkin coarse operators are formed from exact CSR products R A T and
converted back to stencil form. The reference's stencil coarsening
(`coarsen_sparse_operator`) pairs each fine coefficient c_s(a) with the
neighbour a - s instead of a + s; algebraically it equals
reflect(R reflect(A) T), where reflect maps c_s(x) to c_{-s}(x). This agrees
with R A T for reflection-symmetric stencils (the constant Laplacian of
`test_coarsen`, `test_multigrid.py:197-249`) and differs otherwise; both forms
are provided (mode 'reference' and 'exact').
"""

import itertools

import numpy as np
import scipy.sparse
import scipy.sparse.linalg


def _shift_slices(s, n):
    """Index ranges so that dst[a:b] pairs with src[a+s:b+s] inside [0,n)."""
    lo = max(0, -s)
    hi = n - max(0, s)
    return lo, hi


class StencilOp:
    """Stencil operator with coefficient fields of shape `nwh`."""

    def __init__(self, coeffs, nw, stride=1):
        self.nw = tuple(int(n) for n in nw)
        self.stride = int(stride)
        self.nwh = tuple((n - 1) // self.stride + 1 for n in self.nw)
        self.coeffs = {
            tuple(int(x) for x in k): np.asarray(v, dtype=np.float64).reshape(
                self.nwh)
            for k, v in coeffs.items()
        }

    def mul_field(self, u):
        u = np.asarray(u).reshape(self.nw)
        res = np.zeros(self.nwh)
        st = self.stride
        for s, c in self.coeffs.items():
            # Gather u[st*I + s] for every coarse index I, zero outside.
            idx = []
            valid = np.ones(self.nwh, dtype=bool)
            for d in range(len(self.nw)):
                k = st * np.arange(self.nwh[d]) + s[d]
                ok = (k >= 0) & (k < self.nw[d])
                shape = [1] * len(self.nw)
                shape[d] = -1
                valid = valid & ok.reshape(shape)
                idx.append(np.clip(k, 0, self.nw[d] - 1).reshape(shape))
            g = u[tuple(np.broadcast_arrays(*idx))]
            res += np.where(valid, c * g, 0.)
        return res.ravel()

    def tocsr(self):
        n = int(np.prod(self.nw))
        nh = int(np.prod(self.nwh))
        rows, cols, vals = [], [], []
        grids = np.meshgrid(*[np.arange(m) for m in self.nwh], indexing='ij')
        flat_row = np.arange(nh).reshape(self.nwh)
        for s, c in self.coeffs.items():
            valid = np.ones(self.nwh, dtype=bool)
            kk = []
            for d in range(len(self.nw)):
                k = self.stride * grids[d] + s[d]
                valid &= (k >= 0) & (k < self.nw[d])
                kk.append(k)
            col = np.ravel_multi_index(
                tuple(np.where(valid, k, 0) for k in kk), self.nw)
            rows.append(flat_row[valid])
            cols.append(col[valid])
            vals.append(c[valid])
        if not vals:
            return scipy.sparse.csr_matrix((nh, n))
        m = scipy.sparse.csr_matrix(
            (np.concatenate(vals), (np.concatenate(rows),
                                    np.concatenate(cols))), shape=(nh, n))
        m.sum_duplicates()
        m.eliminate_zeros()
        return m

    def mul_transpose_field(self, uh):
        return self.tocsr().T @ np.asarray(uh).ravel()

    @staticmethod
    def from_csr(m, nw):
        """Stride-1 stencil form of a square CSR matrix on grid `nw`."""
        m = m.tocoo()
        dim = len(nw)
        ri = np.unravel_index(m.row, nw)
        ci = np.unravel_index(m.col, nw)
        shifts = np.stack([ci[d] - ri[d] for d in range(dim)], axis=1)
        coeffs = {}
        keys, inv = np.unique(shifts, axis=0, return_inverse=True)
        inv = inv.ravel()
        for k, key in enumerate(keys):
            sel = inv == k
            c = np.zeros(nw)
            c[tuple(r[sel] for r in ri)] = m.data[sel]
            coeffs[tuple(int(x) for x in key)] = c
        return StencilOp(coeffs, nw)


def mul_transpose_op(a, b):
    """Stencil form of A^T B for stride-1 operators
    (`SparseOperator.mul_transpose_op`, `multigrid.py:1035-1071`)."""
    return StencilOp.from_csr((a.tocsr().T @ b.tocsr()).tocsr(), a.nw)


def get_R(nw, restriction='full'):
    """Node-based restriction to `nw//2+1` points: tensor weights
    0.5**(|dw|+dim) at inner coarse nodes, injection at coarse boundary nodes
    (`multigrid.py:496-534`)."""
    nw = np.array(nw)
    dim = len(nw)
    nwh = nw // 2 + 1
    if restriction in ('full', 'transpose'):
        ddw = list(itertools.product([-1, 0, 1], repeat=dim))
        coeff = lambda dw: 0.5**(sum(abs(x) for x in dw) + dim)
    elif restriction == 'half':
        ddw = [(0, ) * dim]
        for i in range(dim):
            ddw.append(tuple(1 if j == i else 0 for j in range(dim)))
            ddw.append(tuple(-1 if j == i else 0 for j in range(dim)))
        coeff = lambda dw: (0.5 if sum(abs(x) for x in dw) == 0 else 0.25 /
                            dim)
    elif restriction == 'injection':
        ddw = [(0, ) * dim]
        coeff = lambda dw: 1.
    else:
        raise ValueError("Unknown restriction=" + restriction)
    coeffs = {}
    for dw in ddw:
        c = np.full(nwh, 1. if all(x == 0 for x in dw) else 0.)
        inner = tuple(slice(1, -1) for _ in range(dim))
        c[inner] = coeff(dw)
        if restriction == 'transpose':
            # Same weights on every coarse node: R = T^T / 2^dim.
            c[...] = coeff(dw)
        coeffs[dw] = c
    return StencilOp(coeffs, tuple(nw), stride=2)


def get_TT(nw):
    """Transpose of node-based multilinear interpolation as a stride-2
    stencil, weights 0.5**|dw| (`multigrid.py:471-494`)."""
    nw = np.array(nw)
    dim = len(nw)
    nwh = nw // 2 + 1
    coeffs = {}
    for dw in itertools.product([-1, 0, 1], repeat=dim):
        c = np.zeros(nwh)
        sel = tuple(
            slice(0, n - 1) if s > 0 else slice(1, None) if s < 0 else slice(
                None) for s, n in zip(dw, nwh))
        c[sel] = 0.5**sum(abs(x) for x in dw)
        coeffs[dw] = c
    return StencilOp(coeffs, tuple(nw), stride=2)


def reflect(a):
    """Stencil with every shift negated at the same coefficient location."""
    return StencilOp({tuple(-x for x in s): c
                      for s, c in a.coeffs.items()}, a.nw, a.stride)


def galerkin(a, r, tt, mode='reference'):
    """Coarse operator on grid r.nwh. mode 'exact': R A T;
    mode 'reference': the arithmetic of `coarsen_sparse_operator`
    (`multigrid.py:390-427`), reflect(R reflect(A) T)."""
    if mode == 'reference':
        return reflect(galerkin(reflect(a), r, tt, 'exact'))
    if mode != 'exact':
        raise ValueError('Unknown galerkin mode ' + mode)
    m = r.tocsr() @ a.tocsr() @ tt.tocsr().T
    return StencilOp.from_csr(m.tocsr(), r.nwh)


class MultigridOp:
    """V-cycle over nvar x nvar blocks of stencil operators
    (`multigrid.py:342-602`)."""

    def __init__(self, nw, nvar=1, nlevels=None, restriction='full',
                 galerkin_mode='reference', coarse_reg=None):
        nw = np.array(nw)
        max_nlevels = 0
        while 2**(max_nlevels + 1) < min(nw):
            max_nlevels += 1
        nlevels = max_nlevels if nlevels is None else min(
            nlevels, max_nlevels)
        nnw = [nw]
        for i in range(1, nlevels):
            nnw.append((nnw[i - 1] - 1) // 2 + 1)
        self.nnw = [tuple(int(x) for x in n) for n in nnw]
        self.nlevels = nlevels
        self.nvar = nvar
        self.RR = [get_R(n, restriction) for n in self.nnw]
        self.restriction = restriction
        self.TT = [get_TT(n) for n in self.nnw]
        self.galerkin_mode = galerkin_mode
        self.coarse_reg = coarse_reg

    def update_A(self, A0):
        AA = [A0]
        for level in range(1, self.nlevels):
            R = self.RR[level - 1]
            T = self.TT[level - 1]
            AA.append([[galerkin(AA[-1][i][j], R, T, self.galerkin_mode)
                        for j in range(self.nvar)]
                       for i in range(self.nvar)])
        self.AA = AA
        zero = (0, ) * len(self.nnw[0])
        self.DD = [[
            A[i][i].coeffs.get(zero, np.zeros(A[i][i].nw)).ravel()
            for i in range(self.nvar)
        ] for A in AA]

    def residual(self, A, u, f):
        return [
            f[i] - sum(A[i][j].mul_field(u[j]) for j in range(self.nvar))
            for i in range(self.nvar)
        ]

    def smoother_jacobi(self, level, u, f, omega=1.):
        r = self.residual(self.AA[level], u, f)
        D = self.DD[level]
        for i in range(self.nvar):
            u[i] += r[i] / D[i] * omega
        return u

    def smoother_direct(self, level, u, f):
        """Coarsest-level solve. The reference calls SuperLU
        (`multigrid.py:554-565`); here small systems use the dense
        minimum-norm least-squares solution, identical for non-singular
        systems and well defined for the cavity's constant-pressure null
        space, where SuperLU's null component is set by round-off."""
        A = self.AA[level]
        m = scipy.sparse.bmat([[A[i][j].tocsr() for j in range(self.nvar)]
                               for i in range(self.nvar)])
        b = np.concatenate(f)
        if self.coarse_reg:
            # Tikhonov least squares (M^T M + d I)^-1 M^T b with
            # d = coarse_reg * max diag(M^T M).
            md = m.toarray()
            g = md.T @ md
            g[np.diag_indices_from(g)] += self.coarse_reg * g.diagonal().max()
            x = np.linalg.solve(g, md.T @ b)
        elif m.shape[0] <= 4096:
            x = np.linalg.lstsq(m.toarray(), b, rcond=1e-13)[0]
        else:
            x = scipy.sparse.linalg.spsolve(m.tocsc(), b)
        return list(x.reshape(self.nvar, -1))

    def step(self, u, f, omega, level=0, ndirect=5, pre=2, post=2):
        """One V-cycle. As in the reference (`multigrid.py:589-593`) the
        recursive call does not forward `ndirect`, `pre` or `post`, so levels
        below the finest use the defaults 5, 2, 2."""
        if level + 1 >= self.nlevels or min(self.nnw[level]) <= ndirect:
            return self.smoother_direct(level, u, f)
        for _ in range(pre):
            u = self.smoother_jacobi(level, u, f, omega)
        A = self.AA[level]
        R = self.RR[level]
        T = self.TT[level]
        r = [R.mul_field(x) for x in self.residual(A, u, f)]
        du = self.step([np.zeros_like(x) for x in r], r, omega,
                       level=level + 1)
        du = [T.mul_transpose_field(x) for x in du]
        u = [a + b for a, b in zip(u, du)]
        for _ in range(post):
            u = self.smoother_jacobi(level, u, f, omega)
        return u


def normal_operator(stf, nw, nvar):
    """Stencil form of J^T J and the per-(eq, var) stencil blocks of J from
    tape coefficient fields `stf[i][j][shift]`
    (`util.py:551-583`)."""
    neqn = len(stf)
    mm = [[StencilOp(stf[i][j], nw) for j in range(nvar)]
          for i in range(neqn)]
    matr = [[None] * nvar for _ in range(nvar)]
    for i in range(nvar):
        for j in range(nvar):
            acc = None
            for e in range(neqn):
                prod = (mm[e][i].tocsr().T @ mm[e][j].tocsr()).tocsr()
                acc = prod if acc is None else acc + prod
            matr[i][j] = StencilOp.from_csr(acc, nw)
    return mm, matr


def coeffs_from_tape(grads, field_desc, neqn, nvar):
    """Group tape gradients into `stf[eq][var][shift]`, skipping unused and
    all-zero coefficient fields (`util.py:551-560`)."""
    stf = [[dict() for _ in range(nvar)] for _ in range(neqn)]
    for i in range(neqn):
        for j, (var, dw) in enumerate(field_desc):
            g = grads[i][j]
            if g is None or not np.any(g):
                continue
            stf[i][var][tuple(int(x) for x in dw)] = g
    return stf


class Multigrid:
    """The sparse-matrix `Multigrid` (`multigrid.py:72-339`): block-diagonal
    R and T over nvar fields, coarse matrices R A T by sparse products
    (`update_A`, `:116-157`), Jacobi smoother on the diagonal blocks
    (`:284-299`: full=False point Jacobi, full=True a dense nvar x nvar
    solve per cell with every block's diagonal), Gauss-Seidel with the
    lower triangle (`:301-305`, SciPy's sparse triangular solve), sparse
    direct coarsest solve (`:307-310`) and the V-cycle whose recursion
    forwards `ndirect` but not `pre`/`post` (`:312-339`). Vectors are flat,
    fields concatenated."""

    def __init__(self, nw, nvar=1, nlevels=None, restriction='full'):
        nw = np.array(nw)
        max_nlevels = 0
        while 2**(max_nlevels + 1) < min(nw):
            max_nlevels += 1
        nlevels = max_nlevels if nlevels is None else min(
            nlevels, max_nlevels)
        nnw = [nw]
        for i in range(1, nlevels):
            nnw.append((nnw[i - 1] - 1) // 2 + 1)
        self.nnw = [tuple(int(x) for x in n) for n in nnw]
        self.nlevels = nlevels
        self.nvar = nvar
        bd = lambda m: scipy.sparse.block_diag([m] * nvar).tocsr()
        self.RR = [bd(get_R(n, restriction).tocsr()) for n in self.nnw]
        self.TT = [bd(get_TT(n).tocsr().T) for n in self.nnw]

    def update_A(self, A0):
        AA = [A0.tocsr()]
        for level in range(1, self.nlevels):
            AA.append((self.RR[level - 1] @ AA[-1] @
                       self.TT[level - 1]).tocsr())
        self.AA = AA
        self.DD = [A.diagonal() for A in AA]
        # extdiag (`:133-140`): diagonal of block (i, j), shape (nv, nv, n).
        nv = self.nvar
        self.DB = []
        for A in AA:
            ns = A.shape[0] // nv
            self.DB.append(np.array([[
                A[i * ns:(i + 1) * ns, j * ns:(j + 1) * ns].diagonal(0)
                for j in range(nv)] for i in range(nv)]))
        # tril including the diagonal (`:148-149`).
        self.LL = [scipy.sparse.tril(A).tocsr() for A in AA]

    def smoother_jacobi(self, level, u, f, omega=1., full=False):
        r = f - self.AA[level] @ u
        if full:
            d = np.moveaxis(self.DB[level], 2, 0)  # (n, nv, nv)
            rr = r.reshape(self.nvar, -1).T[..., None]  # (n, nv, 1)
            du = np.linalg.solve(d, rr)[..., 0].T.ravel()
        else:
            du = r / self.DD[level]
        u += du * omega
        return u

    def smoother_gauss_seidel(self, level, u, f, omega=1.):
        A, L = self.AA[level], self.LL[level]
        x = scipy.sparse.linalg.spsolve_triangular(L, f - (A @ u - L @ u))
        u += (x - u) * omega
        return u

    def smoother_direct(self, level, u, f):
        return scipy.sparse.linalg.spsolve(self.AA[level].tocsc(), f)

    def step(self, u, f, smoother, level=0, ndirect=5, pre=2, post=2):
        if level + 1 >= self.nlevels or min(self.nnw[level]) <= ndirect:
            return self.smoother_direct(level, u, f)
        for _ in range(pre):
            u = smoother(level, u, f)
        r = self.RR[level] @ (f - self.AA[level] @ u)
        u = u + self.TT[level] @ self.step(np.zeros_like(r), r, smoother,
                                           level=level + 1, ndirect=ndirect)
        for _ in range(post):
            u = smoother(level, u, f)
        return u
