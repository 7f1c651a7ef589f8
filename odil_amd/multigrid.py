"""Stencil operators and geometric multigrid on device (reference:
`SparseOperator`, `MultigridOp`, multigrid.py:342-602,877-1214).

`BlockStencil` is the device form of an nvar_out x nvar_in block of
`SparseOperator`s: coefficient fields `coef[nent, *nwh]` and an entry table
(out, in, shift). All arithmetic runs in the HIP kernels of
`libodil_amd.so`; torch only owns the memory.
"""

import ctypes
import math
import itertools

import numpy as np
import torch

from odil_amd import _lib
from odil_amd._lib import c_double, c_int, check, ptr, ptrs, stream
from odil_amd.dense import SpdSolver

# `lanes` value selecting the LDS-tiled symmetric sweep (`per` is then the
# tile shape, odil_stencil_apply_sym_tiled).
TILED = -3


def _grid_arr(nw):
    return (ctypes.c_int * 3)(*(list(nw) + [1] * (3 - len(nw))))


def _shift3(s):
    s = tuple(int(x) for x in s)
    return s + (0, ) * (3 - len(s))


def is_upper(o, i, s):
    """Entry (o, i, s) belongs to the stored upper half of a symmetric block
    operator: o < i, or o == i and s >= 0 (lexicographic)."""
    return o < i or (o == i and tuple(s) >= (0, ) * len(s))


def _mirror(o, i, s):
    return (i, o, tuple(-x for x in s))


class BlockStencil:
    """Block stencil operator from nvar_in fields on grid `nw` to nvar_out
    fields on grid `nwh = (nw - 1) // stride + 1`:
    (A x)_o[I] = sum_{e: out_e = o} coef_e[I] x_{in_e}[stride*I + s_e],
    zero outside the grid (non-circular, multigrid.py:941-969).

    sym=True: a symmetric square stride-1 operator stored as its upper half
    (`is_upper` entries); the lower half is A_(i,o,-s)(c) = A_(o,i,s)(c-s).
    Halves the coefficient memory and traffic of J^T J and its Galerkin
    coarse operators."""

    def __init__(self, coef, entries, nw, stride=1, nvar_in=None,
                 nvar_out=None, sym=False):
        self.sym = bool(sym)
        self.lanes = 0  # threads per cell of the symmetric sweep, 0 = auto
        self.per = 1    # 256-thread tiles per workgroup of that sweep
        self.nw = tuple(int(n) for n in nw)
        self.ndim = len(self.nw)
        self.stride = int(stride)
        self.nwh = tuple((n - 1) // self.stride + 1 for n in self.nw)
        entries = [(int(o), int(i), tuple(int(x) for x in s))
                   for o, i, s in entries]
        if len(entries) != coef.shape[0]:
            raise ValueError('entries and coefficient count differ')
        if tuple(coef.shape[1:]) != self.nwh:
            raise ValueError('coefficient shape {} != {}'.format(
                tuple(coef.shape[1:]), self.nwh))
        order = sorted(range(len(entries)), key=lambda q: entries[q])
        if order != list(range(len(entries))):
            idx = torch.tensor(order, device=coef.device)
            coef = coef.index_select(0, idx)
            entries = [entries[q] for q in order]
        if len(set(entries)) != len(entries):
            raise ValueError('duplicate stencil entries')
        self.coef = coef.contiguous()
        self.entries = entries
        self.nvar_out = nvar_out if nvar_out is not None else (
            max(e[0] for e in entries) + 1 if entries else 1)
        self.nvar_in = nvar_in if nvar_in is not None else (
            max(e[1] for e in entries) + 1 if entries else 1)
        if self.sym:
            if self.stride != 1:
                raise ValueError('symmetric storage needs stride 1')
            self.nvar_in = self.nvar_out = max(self.nvar_in, self.nvar_out)
            bad = [e for e in entries if not is_upper(*e)]
            if bad:
                raise ValueError('symmetric storage holds upper entries '
                                 'only, got {}'.format(bad[:3]))
        self._tables()
        if self.sym:
            self.zero_outside()

    def zero_outside(self):
        """Zero the coefficients that couple a node to a point outside the
        grid (they never enter a product: zero fill, multigrid.py:941-947).
        Symmetric operators keep this invariant, which lets their sweeps skip
        the per-product bounds test along the contiguous axis; the
        constructor, normal_operator, galerkin and update_A apply it, and
        code that writes `coef` in place must call it again."""
        st = _lib.lib().odil_stencil_zero_outside(
            c_int(_lib.dtype_tag(self.dtype)), c_int(self.ndim),
            _grid_arr(self.nw), ptr(self.t_ent), c_int(len(self.entries)),
            self._reach_all(), ptr(self.coef), stream())
        check(st, 'stencil_zero_outside')
        return self

    def _reach_all(self):
        r = [0, 0, 0]
        for _, _, s in self.entries:
            for d, v in enumerate(_shift3(s)):
                r[d] = max(r[d], abs(v))
        return (ctypes.c_int * 3)(*r)

    @property
    def dtype(self):
        return self.coef.dtype

    @property
    def device(self):
        return self.coef.device

    def _tables(self):
        dev = self.coef.device
        ent = np.array([[o, i] + list(_shift3(s)) for o, i, s in self.entries],
                       dtype=np.int32).reshape(-1, 5)
        ostart = np.zeros(self.nvar_out + 1, dtype=np.int32)
        for o, _, _ in self.entries:
            ostart[o + 1] += 1
        ostart = np.cumsum(ostart).astype(np.int32)
        zero = (0, ) * self.ndim
        dent = np.full(self.nvar_out, -1, dtype=np.int32)
        for q, (o, i, s) in enumerate(self.entries):
            if o == i and s == zero:
                dent[o] = q
        by_in = sorted(range(len(self.entries)),
                       key=lambda q: (self.entries[q][1], q))
        tent = np.array(by_in, dtype=np.int32)
        istart = np.zeros(self.nvar_in + 1, dtype=np.int32)
        for _, i, _ in self.entries:
            istart[i + 1] += 1
        istart = np.cumsum(istart).astype(np.int32)
        self._dent_host = dent
        self._max_entries = int(np.max(np.diff(ostart))) if len(ostart) > 1 \
            else 0
        t = lambda a: torch.tensor(a, dtype=torch.int32, device=dev)
        if self.sym:
            self._sym_tables(t)
        self.t_ent = t(ent)
        self.t_ostart = t(ostart)
        self.t_dent = t(dent)
        self.t_tent = t(tent if len(tent) else np.zeros(1, np.int32))
        self.t_istart = t(istart)

    def _sym_tables(self, t):
        """Per-output product list of the symmetric apply (k_apply_sym)."""
        n3 = list(self.nw) + [1] * (3 - self.ndim)
        lin = lambda v: (v[0] * n3[1] + v[1]) * n3[2] + v[2]
        per = [[] for _ in range(self.nvar_out)]
        zero3 = (0, 0, 0)
        for q, (o, i, s) in enumerate(self.entries):
            s3 = _shift3(s)
            per[o].append((q, i, zero3, s3))
            if not (o == i and all(x == 0 for x in s)):
                m3 = tuple(-x for x in s3)
                per[i].append((q, o, m3, m3))
        rows, ostart = [], [0]
        reach = [0, 0, 0]
        for lst in per:
            for q, var, cs, xs in lst:
                rows.append([q, var, lin(cs), lin(xs)] + list(cs) + list(xs))
                reach = [max(r, abs(a), abs(b))
                         for r, a, b in zip(reach, cs, xs)]
            ostart.append(len(rows))
        self.t_items = t(np.array(rows, dtype=np.int32).reshape(-1, 10))
        self.t_iostart = t(np.array(ostart, dtype=np.int32))
        self._reach = (ctypes.c_int * 3)(*reach)
        # Interior test of the sweeps once the coefficients that couple to
        # outside points are zero (zero_outside): nothing along the
        # contiguous axis, one more line of margin on the axis before it, so
        # that a flat read past the end of a line stays inside the arrays
        # and lands on a zero coefficient.
        fast = list(reach)
        la = self.ndim - 1
        if reach[la] < self.nw[la]:
            fast[la] = 0
            fast[la - 1] += 1
        self._reach_fast = (ctypes.c_int * 3)(*fast)
        self._max_items = max(b - a for a, b in zip(ostart[:-1], ostart[1:]))

    # -- products --------------------------------------------------------
    def _fields(self, x, n, shape):
        if isinstance(x, torch.Tensor):
            x = [x[k] for k in range(n)] if x.dim() == len(shape) + 1 else [x]
        x = list(x)
        if len(x) != n:
            raise ValueError('expected {} fields, got {}'.format(n, len(x)))
        for a in x:
            _lib.require_device(a)
            if a.dtype != self.dtype:
                raise ValueError('dtype mismatch {} vs {}'.format(
                    a.dtype, self.dtype))
            if a.numel() != int(np.prod(shape)):
                raise ValueError('field size {} != grid {}'.format(
                    a.numel(), shape))
        return x

    def _new(self, n, shape):
        return [torch.empty(shape, dtype=self.dtype, device=self.device)
                for _ in range(n)]

    def _contig(self, x):
        return all(a.data_ptr() == x[0].data_ptr() +
                   k * x[0].numel() * x[0].element_size()
                   for k, a in enumerate(x))

    def _call_apply(self, x, y, f, mode, omega):
        if self.sym and self.lanes == TILED:
            st = _lib.lib().odil_stencil_apply_sym_tiled(
                c_int(_lib.dtype_tag(self.dtype)), _grid_arr(self.nw),
                c_int(self.nvar_out), ptr(self.t_ent), c_int(len(self.entries)),
                ptr(self.t_dent), self._reach, c_int(self.per),
                ptr(self.coef), ptrs(x), ptrs(y),
                ptrs(f) if f is not None else None, c_int(mode),
                c_double(omega), stream())
            check(st, 'stencil_apply_sym_tiled')
            return
        if self.sym:
            st = _lib.lib().odil_stencil_apply_sym(
                c_int(_lib.dtype_tag(self.dtype)), c_int(self.ndim),
                _grid_arr(self.nw), c_int(self.nvar_out), ptr(self.t_items),
                ptr(self.t_iostart), ptr(self.t_dent), self._reach_fast,
                c_int(self._max_items), c_int(len(self.entries)),
                c_int(self.lanes), c_int(self.per if self.lanes >= 0 else 1),
                ptr(self.coef), ptrs(x), ptrs(y),
                ptrs(f) if f is not None else None, c_int(mode),
                c_double(omega), stream())
            check(st, 'stencil_apply_sym')
            return
        st = _lib.lib().odil_stencil_apply(
            c_int(_lib.dtype_tag(self.dtype)), c_int(self.ndim),
            _grid_arr(self.nw), c_int(self.stride), c_int(self.nvar_in),
            c_int(self.nvar_out), ptr(self.t_ent), ptr(self.t_ostart),
            ptr(self.t_dent), c_int(self._max_entries), ptr(self.coef),
            ptrs(x), ptrs(y),
            ptrs(f) if f is not None else None, c_int(mode),
            c_double(omega), stream())
        check(st, 'stencil_apply')

    def tile_shapes(self):
        """Tile shapes of the LDS-tiled symmetric sweep that apply to this
        operator (2D, shifts of at most 2, 1, 3 or 4 fields, at most 256
        stored entries and 64 KiB of LDS per workgroup); empty otherwise.
        The autotuner times them beside the per-cell launch shapes."""
        if not self.sym or self.ndim != 2 or self.nvar_out not in (1, 3, 4):
            return []
        if len(self.entries) > 256:
            return []
        if max(self._reach[0], self._reach[1]) > 2:
            return []
        lib = _lib.lib()
        dt = _lib.dtype_tag(self.dtype)
        return [t for t in range(1, 7)
                if 0 < lib.odil_stencil_tile_lds(dt, self.nvar_out, t)
                <= 65536]

    def apply(self, x, y=None):
        """y = A x (mul_field)."""
        x = self._fields(x, self.nvar_in, self.nw)
        y = self._new(self.nvar_out, self.nwh) if y is None else y
        self._call_apply(x, y, None, 0, 0.)
        return y

    def residual(self, x, f, y=None):
        """y = f - A x."""
        x = self._fields(x, self.nvar_in, self.nw)
        f = self._fields(f, self.nvar_out, self.nwh)
        y = self._new(self.nvar_out, self.nwh) if y is None else y
        self._call_apply(x, y, f, 1, 0.)
        return y

    def jacobi(self, x, f, y, omega):
        """One weighted-Jacobi sweep y = x + omega (f - A x) / diag(A)
        (smoother_jacobi, multigrid.py:543-552). y must not alias x."""
        if np.any(self._dent_host < 0):
            raise ValueError('Jacobi needs a diagonal entry in every block')
        self._call_apply(x, y, f, 2, omega)
        return y

    def diagonal(self):
        """Views of the diagonal coefficient array of each field."""
        if np.any(self._dent_host < 0):
            raise ValueError('Jacobi needs a diagonal entry in every block')
        return [self.coef[int(q)] for q in self._dent_host]

    def jacobi_zero(self, f, y, omega, f_src=None):
        """The Jacobi sweep from x = 0: y = omega f / diag(A), one pass
        without the operator product. With `f_src` (same shape, e.g. the
        FP64 vector a working-precision V-cycle starts from), reads f_src
        and stores its conversion into f."""
        return jacobi_zero(f if f_src is None else f_src, self.diagonal(), y,
                           omega, fw=None if f_src is None else f)

    def tune_key(self, contig):
        """Autotune key of the symmetric sweep's launch shape."""
        return ('apply_sym', self.nw, len(self.entries), self._max_items,
                str(self.dtype), contig)

    def sweep_candidates(self, contig, per=(1, 2, 4, 8)):
        """Launch shapes (lanes, per) of the symmetric sweep: threads per
        cell with `per` tiles per workgroup; with contiguous fields also the
        16-byte vector form, two cells per thread and the LDS tiles."""
        cand = [(L, p) for L in (0, 1, 2, 4, 8, 16) for p in per]
        if contig:
            cand += [(-1, 1), (-2, 1)]
            cand += [(TILED, t) for t in self.tile_shapes()]
        return cand

    def tune(self, x, f, y, omega=0.5):
        """Pick the threads per cell of the symmetric sweep by timing the
        Jacobi sweep on these buffers (cached per operator shape)."""
        if not self.sym:
            return self
        from odil_amd import autotune
        contig = self._contig(x)

        def run(c):
            self.lanes, self.per = c
            self._call_apply(x, y, f, 2, omega)
        self.lanes, self.per = autotune.pick(self.tune_key(contig),
                                             self.sweep_candidates(contig),
                                             run)
        return self

    def apply_t(self, x, y=None, accumulate=False):
        """y = A^T x (mul_transpose_field); accumulate adds into y."""
        x = self._fields(x, self.nvar_out, self.nwh)
        if self.sym:
            if not accumulate or y is None:
                return self.apply(x, y)
            t = self.apply(x)
            for a, b in zip(y, t):
                a.add_(b)
            return y
        if y is None:
            y = self._new(self.nvar_in, self.nw)
            accumulate = False
        st = _lib.lib().odil_stencil_apply_t(
            c_int(_lib.dtype_tag(self.dtype)), c_int(self.ndim),
            _grid_arr(self.nw), c_int(self.stride), c_int(self.nvar_in),
            c_int(self.nvar_out), ptr(self.t_ent), ptr(self.t_tent),
            ptr(self.t_istart), ptr(self.coef), ptrs(x), ptrs(y),
            c_int(1 if accumulate else 0), stream())
        check(st, 'stencil_apply_t')
        return y

    def to_dense(self):
        """Dense FP64 matrix (nvar*N)^2 of a square stride-1 operator."""
        if self.stride != 1 or self.nvar_in != self.nvar_out:
            raise ValueError('to_dense needs a square stride-1 operator')
        n = self.nvar_out * int(np.prod(self.nw))
        M = torch.zeros((n, n), dtype=torch.float64, device=self.device)
        st = _lib.lib().odil_stencil_to_dense(
            c_int(_lib.dtype_tag(self.dtype)), c_int(self.ndim),
            _grid_arr(self.nw), c_int(self.nvar_out), ptr(self.t_ent),
            c_int(len(self.entries)), ptr(self.coef),
            c_int(1 if self.sym else 0), ptr(M), stream())
        check(st, 'stencil_to_dense')
        return M

    def to_csr_arrays(self):
        """Device CSR (rowptr, col, val) in the reference linearize layout
        (util_op.py:877-894): row out N + c, column in N + flat((c + s) mod
        shape), entries of each row in entry order, explicit zeros kept.
        Stride-1 operators only."""
        if self.stride != 1:
            raise ValueError('to_csr_arrays needs a stride-1 operator')
        if self.sym:
            raise ValueError('to_csr_arrays needs full (non-symmetric) '
                             'storage')
        N = int(np.prod(self.nw))
        dev = self.device
        rowptr = torch.empty(self.nvar_out * N + 1, dtype=torch.int32,
                             device=dev)
        col = torch.empty(len(self.entries) * N, dtype=torch.int32,
                          device=dev)
        val = torch.empty(len(self.entries) * N, dtype=self.dtype,
                          device=dev)
        st = _lib.lib().odil_stencil_to_csr(
            c_int(_lib.dtype_tag(self.dtype)), c_int(self.ndim),
            _grid_arr(self.nw), c_int(self.nvar_out), ptr(self.t_ent),
            ptr(self.t_ostart), c_int(len(self.entries)), ptr(self.coef),
            ptr(rowptr), ptr(col), ptr(val), stream())
        check(st, 'stencil_to_csr')
        return rowptr, col, val

    def astype(self, dtype):
        if dtype == self.dtype:
            return self
        out = BlockStencil(self.coef.to(dtype), self.entries, self.nw,
                           self.stride, self.nvar_in, self.nvar_out,
                           sym=self.sym)
        out.lanes, out.per = self.lanes, self.per
        return out

    def tocsr(self):
        """SciPy CSR copy on the host (tests and the CSR export)."""
        import scipy.sparse
        coef = self.coef.double().cpu().numpy()
        n = int(np.prod(self.nw))
        nh = int(np.prod(self.nwh))
        grids = np.meshgrid(*[np.arange(m) for m in self.nwh], indexing='ij')
        rows, cols, vals = [], [], []
        for q, (o, i, s) in enumerate(self.entries):
            valid = np.ones(self.nwh, dtype=bool)
            kk = []
            for d in range(self.ndim):
                k = self.stride * grids[d] + s[d]
                valid &= (k >= 0) & (k < self.nw[d])
                kk.append(np.where(valid, k, 0))
            col = np.ravel_multi_index(tuple(np.where(valid, k, 0)
                                             for k in kk), self.nw)
            r = np.arange(nh).reshape(self.nwh)
            rows.append((r + o * nh)[valid])
            cols.append((col + i * n)[valid])
            vals.append(coef[q][valid])
        shape = (self.nvar_out * nh, self.nvar_in * n)
        if not vals:
            return scipy.sparse.csr_matrix(shape)
        m = scipy.sparse.csr_matrix(
            (np.concatenate(vals), (np.concatenate(rows),
                                    np.concatenate(cols))), shape=shape)
        m.sum_duplicates()
        if self.sym:
            d = scipy.sparse.diags(m.diagonal())
            m = (m + m.T - d).tocsr()
        return m

    @staticmethod
    def from_jacobian(jac):
        """J (equations x fields) as a block stencil."""
        return BlockStencil(jac.coef, jac.entries, jac.shape,
                            nvar_in=jac.nvar, nvar_out=jac.neq)

    @staticmethod
    def from_csr(m, nw, nvar_out, nvar_in=None, device=None,
                 dtype=torch.float64):
        """Stride-1 block stencil of a sparse matrix whose rows are
        nvar_out fields and columns nvar_in fields on node grid `nw`, each
        field-major and C-order (the layout of `Problem.linearize`,
        util_op.py:877-894, and of the reference's `Multigrid` vectors).
        Entry (o, i, s) holds row (o, c) -> column (i, c + s); the grid
        offset s is taken from the actual column, so a coupling that wraps
        around a periodic grid becomes a long shift (exact as a matrix)."""
        nvar_in = nvar_out if nvar_in is None else nvar_in
        nw = tuple(int(n) for n in nw)
        N = int(np.prod(nw))
        if m.shape != (nvar_out * N, nvar_in * N):
            raise ValueError('matrix shape {} does not match {} x {} fields '
                             'on {}'.format(m.shape, nvar_out, nvar_in, nw))
        m = m.tocoo()
        m.sum_duplicates()
        o, c = np.divmod(m.row.astype(np.int64), N)
        i, cc = np.divmod(m.col.astype(np.int64), N)
        ri = np.unravel_index(c, nw)
        ci = np.unravel_index(cc, nw)
        key = np.stack([o, i] + [ci[d] - ri[d] for d in range(len(nw))],
                       axis=1)
        keys, inv = np.unique(key, axis=0, return_inverse=True)
        coef = np.zeros((len(keys), N))
        coef[inv.ravel(), c] = m.data
        entries = [(int(k[0]), int(k[1]), tuple(int(x) for x in k[2:]))
                   for k in keys]
        t = torch.tensor(coef.reshape((len(keys), ) + nw), dtype=dtype,
                         device=device or torch.device('cuda'))
        return BlockStencil(t, entries, nw, nvar_in=nvar_in,
                            nvar_out=nvar_out)



def _torch_dtype(dtype):
    """torch dtype for a torch, NumPy or string dtype (None: FP64)."""
    if dtype is None:
        return torch.float64
    if isinstance(dtype, torch.dtype):
        return dtype
    name = np.dtype(dtype).name
    if name == 'float64':
        return torch.float64
    if name == 'float32':
        return torch.float32
    raise ValueError('unsupported dtype {}'.format(dtype))


class SparseOperator:
    """The reference's `SparseOperator` (multigrid.py:877-1214) on the GPU:
    A = sum_s diag(c_s) Shift_s over one field, zero outside the grid
    (non-circular, `apply_shift`, multigrid.py:941-947), optional output
    stride. `shift_to_field` maps shift tuples to coefficient fields (device
    tensors of the output grid, any layout of that size). Backed by a
    single-block `BlockStencil`: `mul_field` / `mul_transpose_field` are the
    stencil apply kernels, `mul_transpose_op` the stored J^T J kernel on the
    two-field block [A B] (its (0, 1) block is A^T B), `add_elementwise`
    sums coefficient fields on the device. Vectors are returned flat like the
    reference's. `mod` is accepted for the reference's signature and unused.
    """

    def __init__(self, shift_to_field, nw, stride=1, mod=None, dtype=None,
                 device=None):
        self.mod = mod
        self.nw = np.array([int(n) for n in nw])
        self.stride = int(stride)
        self.dim = len(self.nw)
        self.nwh = (self.nw - 1) // self.stride + 1
        self._tdtype = _torch_dtype(dtype)
        self.dtype = dtype if dtype is not None else self._tdtype
        nwh = tuple(int(n) for n in self.nwh)
        keys, vals = [], []
        for dw, c in dict(shift_to_field).items():
            dw = tuple(int(x) for x in np.atleast_1d(dw))
            if len(dw) != self.dim:
                raise ValueError('shift {} has not {} axes'.format(
                    dw, self.dim))
            c = torch.as_tensor(c, dtype=self._tdtype, device=device)
            if c.numel() != int(np.prod(nwh)):
                raise ValueError('coefficient of {} has {} values, the '
                                 'output grid {}'.format(dw, c.numel(), nwh))
            keys.append(dw)
            vals.append(c.reshape(nwh))
        self.device = vals[0].device if vals else (
            torch.device(device) if device is not None else
            torch.device('cuda'))
        self.shift_to_field = dict(zip(keys, vals))
        if vals:
            coef = torch.stack(vals).contiguous()
        else:
            coef = torch.zeros((0, ) + nwh, dtype=self._tdtype,
                               device=self.device)
        self._op = BlockStencil(coef, [(0, 0, k) for k in keys],
                                tuple(int(n) for n in self.nw),
                                stride=self.stride, nvar_in=1, nvar_out=1) \
            if keys else None

    def _vec(self, u, shape):
        u = torch.as_tensor(u, dtype=self._tdtype, device=self.device)
        if u.numel() != int(np.prod(shape)):
            raise ValueError('vector of {} values for a grid {}'.format(
                u.numel(), tuple(shape)))
        return u.contiguous().reshape(tuple(int(n) for n in shape))

    def mul_field(self, u):
        """A u, flat (multigrid.py:958-969)."""
        u = self._vec(u, self.nw)
        if self._op is None:
            return torch.zeros(int(np.prod(self.nwh)), dtype=self._tdtype,
                               device=self.device)
        return self._op.apply([u])[0].reshape(-1)

    def mul_transpose_field(self, uh):
        """A^T uh, flat (multigrid.py:1006-1033)."""
        uh = self._vec(uh, self.nwh)
        if self._op is None:
            return torch.zeros(int(np.prod(self.nw)), dtype=self._tdtype,
                               device=self.device)
        return self._op.apply_t([uh])[0].reshape(-1)

    def mul_transpose_op(self, op):
        """A^T B as a SparseOperator, coefficient of shift sb - sa:
        sum c^A_sa(x - sa) c^B_sb(x - sa) (multigrid.py:1035-1071)."""
        if self.stride != 1 or op.stride != 1:
            raise ValueError('mul_transpose_op needs stride 1')
        if tuple(self.nw) != tuple(op.nw):
            raise ValueError('grids differ')
        if self._op is None or op._op is None:
            return SparseOperator({}, self.nw, mod=self.mod,
                                  dtype=self.dtype, device=self.device)
        same = op is self
        ents = [(0, 0, k) for k in self.shift_to_field]
        coefs = list(self.shift_to_field.values())
        if not same:
            ents += [(0, 1, k) for k in op.shift_to_field]
            coefs += list(op.shift_to_field.values())
        J = BlockStencil(torch.stack(coefs).contiguous(), ents,
                         tuple(int(n) for n in self.nw), nvar_in=1 + (
                             not same), nvar_out=1)
        N = normal_operator(J, torch.float64)
        want = 0 if same else 1
        out = {}
        for q, (o, i, sh) in enumerate(N.entries):
            if o == 0 and i == want:
                out[sh] = N.coef[q].to(self._tdtype)
        return SparseOperator(out, self.nw, mod=self.mod, dtype=self.dtype,
                              device=self.device)

    def mul_self_transpose(self):
        return self.mul_transpose_op(self)

    def add_elementwise(self, op):
        """A + B (union of the shifts, multigrid.py:1138-1184)."""
        if tuple(self.nw) != tuple(op.nw) or self.stride != op.stride:
            raise ValueError('operators on different grids')
        out = {k: v.clone() for k, v in self.shift_to_field.items()}
        for k, v in op.shift_to_field.items():
            if k in out:
                out[k] += v
            else:
                out[k] = v.clone()
        return SparseOperator(out, self.nw, stride=self.stride,
                              mod=self.mod, dtype=self.dtype,
                              device=self.device)

    def eliminate_zeros(self):
        """Drops shifts whose coefficient field is zero."""
        keep = {k: v for k, v in self.shift_to_field.items()
                if bool((v != 0).any())}
        self.__init__(keep, self.nw, stride=self.stride, mod=self.mod,
                      dtype=self.dtype, device=self.device)

    def tocsr(self):
        """SciPy CSR copy (multigrid.py:906-920), explicit zeros dropped."""
        if self._op is None:
            import scipy.sparse
            return scipy.sparse.csr_matrix(
                (int(np.prod(self.nwh)), int(np.prod(self.nw))))
        m = self._op.tocsr()
        m.eliminate_zeros()
        return m

    @staticmethod
    def eye(nw, stride=1, mod=None, dtype=None, device=None):
        nwh = tuple((int(n) - 1) // stride + 1 for n in nw)
        return SparseOperator(
            {(0, ) * len(nw): torch.ones(nwh, dtype=_torch_dtype(dtype),
                                        device=device)},
            nw, stride=stride, mod=mod, dtype=dtype, device=device)

    @staticmethod
    def diag(c, nw, stride=1, mod=None):
        c = torch.as_tensor(c)
        return SparseOperator({(0, ) * len(nw): c}, nw, stride=stride,
                              mod=mod, dtype=c.dtype, device=c.device)


def blocks_to_stencil(A, nw, dtype=None):
    """nvar x nvar list of SparseOperators (the reference's block matrix of
    `MultigridOp.update_A`, util.py:577-586) as one BlockStencil with
    entries (i, j, shift) for block A[i][j]."""
    nvar = len(A)
    ents, coefs = [], []
    for i in range(nvar):
        if len(A[i]) != nvar:
            raise ValueError('update_A expects an nvar x nvar block list')
        for j in range(nvar):
            for k, c in A[i][j].shift_to_field.items():
                ents.append((i, j, tuple(k)))
                coefs.append(c)
    if not coefs:
        raise ValueError('empty operator')
    coef = torch.stack([c.to(dtype or coefs[0].dtype) for c in coefs])
    return BlockStencil(coef.contiguous(), ents, tuple(int(n) for n in nw),
                        nvar_in=nvar, nvar_out=nvar)


_PLANS = {}


def _normal_plan(J, sym=False):
    key = ('normal', tuple(J.entries), str(J.device), sym)
    if key not in _PLANS:
        out = {}
        for a, (ea, ka, sa) in enumerate(J.entries):
            for b, (eb, kb, sb) in enumerate(J.entries):
                if ea != eb:
                    continue
                d = tuple(y - x for x, y in zip(sa, sb))
                out.setdefault((ka, kb, d), []).append((a, b) + _shift3(sa))
        keys = sorted(k for k in out if not sym or is_upper(*k))
        terms = []
        tstart = [0]
        for k in keys:
            terms.extend(out[k])
            tstart.append(len(terms))
        t_terms = torch.tensor(np.array(terms, dtype=np.int32).reshape(-1, 5),
                               device=J.device)
        t_start = torch.tensor(np.array(tstart, dtype=np.int32),
                               device=J.device)
        _PLANS[key] = (keys, t_terms, t_start)
    return _PLANS[key]


def normal_operator(J, dtype=None, out=None, sym=False):
    """N = J^T J in stencil form (the sum over equations of
    `mul_transpose_op`, util.py:577-583 / multigrid.py:1035-1071):
    N_o[m] = sum_{(a,b,sa)} J_a[m - sa] J_b[m - sa], out var = field of a,
    in var = field of b, shift = s_b - s_a. Accumulated in FP64.
    With `out` (a BlockStencil of the same structure) the result is written
    into its coefficient buffer. sym=True returns the upper half only
    (symmetric storage)."""
    if J.stride != 1:
        raise ValueError('normal_operator needs a stride-1 Jacobian')
    if out is not None:
        sym = out.sym
    keys, t_terms, t_start = _normal_plan(J, sym)
    if out is not None:
        if out.entries != keys or out.nw != J.nw:
            raise ValueError('out has a different structure')
        coef = out.coef
        dtype = coef.dtype
    else:
        dtype = dtype or J.dtype
        coef = torch.empty((len(keys), ) + J.nw, dtype=dtype,
                           device=J.device)
    st = _lib.lib().odil_stencil_normal(
        c_int(_lib.dtype_tag(J.dtype)), c_int(_lib.dtype_tag(dtype)),
        c_int(J.ndim), _grid_arr(J.nw), ptr(t_terms), ptr(t_start),
        c_int(len(keys)), ptr(J.coef), ptr(coef), stream())
    check(st, 'stencil_normal')
    if out is not None:
        if out.sym:
            out.zero_outside()
        return out
    return BlockStencil(coef, keys, J.nw, nvar_in=J.nvar_in,
                        nvar_out=J.nvar_in, sym=sym)


def restriction_weights(restriction, dim):
    """Inner-node restriction weights indexed by offset r in {-1,0,1}^3
    (get_R, multigrid.py:496-534). 'transpose' has the weights of 'full' on
    every coarse node (R = T^T / 2^dim) instead of injecting at coarse
    boundary nodes."""
    rw = np.zeros(27)
    for r in itertools.product([-1, 0, 1], repeat=3):
        if dim == 2 and r[2] != 0:
            continue
        a = sum(abs(x) for x in r)
        if restriction in ('full', 'transpose'):
            w = 0.5**(a + dim)
        elif restriction == 'half':
            w = 0.5 if a == 0 else (0.25 / dim if a == 1 else 0.)
        elif restriction == 'injection':
            w = 1. if a == 0 else 0.
        else:
            raise ValueError('Unknown restriction=' + restriction)
        rw[(r[0] + 1) * 9 + (r[1] + 1) * 3 + (r[2] + 1)] = w
    return rw


def _galerkin_plan(A, sigma):
    key = ('galerkin', tuple(A.entries), A.nw, sigma, A.sym, str(A.device))
    if key not in _PLANS:
        dim = A.ndim
        ofs = list(itertools.product([-1, 0, 1], repeat=dim))
        dcs = list(itertools.product([-2, -1, 0, 1, 2], repeat=dim))
        # Logical entries (o, i, s) -> (stored array q, transposed flag).
        blocks = {}
        for q, (o, i, s) in enumerate(A.entries):
            blocks.setdefault((o, i), []).append((q, s, 0))
            if A.sym and not (o == i and all(x == 0 for x in s)):
                mo, mi, ms = _mirror(o, i, s)
                blocks.setdefault((mo, mi), []).append((q, ms, 1))
        keys, terms, tstart = [], [], [0]
        for (o, i) in sorted(blocks):
            for dc in dcs:
                if A.sym and not is_upper(o, i, dc):
                    continue
                tl = []
                for q, s, tr in blocks[(o, i)]:
                    for r in ofs:
                        t = [r[d] + sigma * (s[d] - 2 * dc[d])
                             for d in range(dim)]
                        if all(abs(x) <= 1 for x in t):
                            tl.append((q, ) + _shift3(r) +
                                      (sum(abs(x) for x in t), ) +
                                      _shift3(s) + (tr, ))
                if tl:
                    keys.append((o, i, dc))
                    terms.extend(tl)
                    tstart.append(len(terms))
        dev = A.device
        t = lambda a: torch.tensor(np.array(a, dtype=np.int32), device=dev)
        oent = t([[o, i] + list(_shift3(dc)) for o, i, dc in keys])
        wtab = torch.tensor([1., 0.5, 0.25, 0.125], dtype=torch.float64,
                            device=dev)
        _PLANS[key] = (keys, oent, t(terms).reshape(-1, 9), t(tstart), wtab)
    return _PLANS[key]


def _rmode(restriction):
    return 1 if restriction == 'transpose' else 0


def galerkin(A, restriction='full', rw_dev=None, mode='reference',
             out=None):
    """Coarse operator on the node grid nw//2+1.

    mode 'exact': R A T (the CSR `Multigrid.update_A`, multigrid.py:131).
    mode 'reference': the arithmetic of the stencil coarsening
    `coarsen_sparse_operator` (multigrid.py:390-427), which pairs the fine
    coefficient c_s(a) with node a - s; it equals reflect(R reflect(A) T)
    with reflect: c_s(x) -> c_{-s}(x), and R A T when the stencil is
    reflection-symmetric.
    With `out` the result is written into its coefficient buffer.
    A symmetric (upper-half) A gives a symmetric coarse operator; this needs
    mode 'exact' and restriction 'transpose' (R = T^T / 2^dim)."""
    if mode not in ('reference', 'exact'):
        raise ValueError('Unknown galerkin mode ' + str(mode))
    if A.stride != 1 or A.nvar_in != A.nvar_out:
        raise ValueError('galerkin needs a square stride-1 operator')
    if A.sym and (mode != 'exact' or restriction != 'transpose'):
        raise ValueError("symmetric Galerkin needs mode='exact' and "
                         "restriction='transpose'")
    sigma = 1 if mode == 'exact' else -1
    keys, oent, terms, tstart, wtab = _galerkin_plan(A, sigma)
    nwh = tuple(n // 2 + 1 for n in A.nw)
    if rw_dev is None:
        rw_dev = torch.tensor(restriction_weights(restriction, A.ndim),
                              device=A.device)
    if out is not None:
        if out.entries != keys or out.nw != nwh:
            raise ValueError('out has a different structure')
        coef = out.coef
    else:
        coef = torch.empty((len(keys), ) + nwh, dtype=A.dtype,
                           device=A.device)
    st = _lib.lib().odil_stencil_galerkin(
        c_int(_lib.dtype_tag(A.dtype)), c_int(A.ndim), _grid_arr(A.nw),
        ptr(oent), ptr(terms), ptr(tstart), c_int(len(keys)),
        ptr(rw_dev), ptr(wtab), ptr(A.coef), ptr(coef), c_int(sigma),
        c_int(_rmode(restriction)), stream())
    check(st, 'stencil_galerkin')
    if out is not None:
        if out.sym:
            out.zero_outside()
        return out
    return BlockStencil(coef, keys, nwh, nvar_in=A.nvar_in,
                        nvar_out=A.nvar_out, sym=A.sym)


def jacobi_zero(f, d, y, omega, fw=None):
    """y_k = omega f_k / d_k for each field (a weighted-Jacobi sweep from a
    zero iterate); fw_k = f_k converted to y's dtype when fw is given."""
    n = f[0].numel()
    st = _lib.lib().odil_jacobi_zero(
        c_int(_lib.dtype_tag(f[0].dtype)), c_int(_lib.dtype_tag(y[0].dtype)),
        c_int(len(f)), ctypes.c_longlong(int(n)), ptrs(f), ptrs(d), ptrs(y),
        ptrs(fw) if fw is not None else None, c_double(omega), stream())
    check(st, 'jacobi_zero')
    return y


def restrict(fine, nw, rw_dev, coarse=None, restriction='full'):
    """Restriction R of a list of fields from node grid `nw` to nw//2+1."""
    nwh = tuple(n // 2 + 1 for n in nw)
    if coarse is None:
        coarse = [torch.empty(nwh, dtype=f.dtype, device=f.device)
                  for f in fine]
    st = _lib.lib().odil_mg_restrict(
        c_int(_lib.dtype_tag(fine[0].dtype)), c_int(len(nw)), _grid_arr(nw),
        c_int(len(fine)), ptr(rw_dev), c_int(_rmode(restriction)),
        ptrs(fine), ptrs(coarse), stream())
    check(st, 'mg_restrict')
    return coarse


def restrict_smooth(fine, nw, rw_dev, coarse, d, u, omega,
                    restriction='full'):
    """coarse = R fine, then the coarse level's first Jacobi sweep from a
    zero iterate, u = omega coarse / d (d: the coarse diagonal), in one
    launch."""
    st = _lib.lib().odil_mg_restrict_smooth(
        c_int(_lib.dtype_tag(fine[0].dtype)), c_int(len(nw)), _grid_arr(nw),
        c_int(len(fine)), ptr(rw_dev), c_int(_rmode(restriction)),
        ptrs(fine), ptrs(coarse), ptrs(d), ptrs(u), c_double(omega),
        stream())
    check(st, 'mg_restrict_smooth')
    return coarse


def prolong_add(coarse, fine, nw):
    """fine += T coarse (multilinear interpolation, node grids)."""
    st = _lib.lib().odil_mg_prolong_add(
        c_int(_lib.dtype_tag(fine[0].dtype)), c_int(len(nw)), _grid_arr(nw),
        c_int(len(fine)), ptrs(coarse), ptrs(fine), stream())
    check(st, 'mg_prolong_add')
    return fine


def level_sizes(nw, nlevels=None):
    """Grid sizes per level, (n - 1)//2 + 1 (multigrid.py:362-370)."""
    nw = np.array(nw)
    max_nlevels = 0
    while 2**(max_nlevels + 1) < min(nw):
        max_nlevels += 1
    nlevels = max_nlevels if nlevels is None else min(nlevels, max_nlevels)
    nnw = [tuple(int(x) for x in nw)]
    for _ in range(1, nlevels):
        nnw.append(tuple((n - 1) // 2 + 1 for n in nnw[-1]))
    return nnw


class MultigridOp:
    """Geometric multigrid on node-based grids of odd size (2^k+1 for full
    depth), mirroring the reference `MultigridOp` (multigrid.py:342-602):
    Galerkin coarse operators R A T, weighted-Jacobi smoothing, and a
    direct solve on the coarsest level, here an FP64 dense inverse applied by
    a GEMV kernel.

    `coarse_reg` > 0 switches the coarsest solve from the inverse to
    Tikhonov-regularised least squares (see `_invert`); use a tiny value
    (e.g. 1e-14) for operators with a null space such as the cavity's
    constant pressure, 0 otherwise.
    """

    def __init__(self, nw, nvar=1, nlevels=None, mod=None, dtype=None,
                 restriction='full', device=None, coarse_reg=0.,
                 galerkin_mode='reference'):
        # Positional order of the reference (nw, nvar, nlevels, mod, dtype,
        # restriction); `mod` is accepted and unused, `dtype` may be a NumPy
        # or torch dtype (default FP64).
        self.mod = mod
        dtype = _torch_dtype(dtype)
        self.nnw = level_sizes(nw, nlevels)
        self.nlevels = len(self.nnw)
        self._spd = {}
        for n in self.nnw[:-1]:
            if any(x % 2 == 0 for x in n):
                raise ValueError('grid sizes {} must be odd on every level '
                                 'but the coarsest'.format(self.nnw))
        self.nvar = nvar
        self.dim = len(self.nnw[0])
        self.restriction = restriction
        self.dtype = dtype
        self.device = device or torch.device('cuda')
        self.coarse_reg = float(coarse_reg)
        self.galerkin_mode = galerkin_mode
        self.rw = torch.tensor(restriction_weights(restriction, self.dim),
                               device=self.device)
        self.AA = None
        self._buf = None
        # Sweeps of the levels below the finest in `vcycle`: the reference's
        # recursion uses its defaults 2 and 2 (multigrid.py:590-593).
        self.coarse_pre, self.coarse_post = 2, 2
        # Optional finest-level operator used instead of AA[0] for
        # smoothing and residuals (same apply/residual/jacobi interface),
        # e.g. a matrix-free J^T J; AA[0] still seeds the Galerkin levels.
        self.fine = None
        # Per-level upper bounds of the Jacobi damping (see
        # estimate_damping); None: the caller's omega on every level.
        self.omega_cap = None

    # -- Jacobi damping bounds -------------------------------------------
    def estimate_damping(self, limit=1.8, iters=20, seed=0):
        """Cap the Jacobi damping of each level at limit / lambda_l, with
        lambda_l the largest eigenvalue of D^-1 A on that level (A
        symmetric positive semi-definite, D its diagonal).

        Weighted Jacobi on such an A is stable only for omega lambda_max <
        2. The normal equations J^T J of the cavity have lambda_max of
        3.0-3.3 at cell Reynolds numbers below ~4 (omega = 0.6 is stable)
        but 3.7 and 4.4 at Re = 1000 on 129^2 and 65^2 cells, where a fixed
        omega = 0.6 smoother diverges on the highest modes and the
        preconditioned Krylov solve stalls.

        Estimate: `iters` power steps x <- (I + D^-1 A) x from a seeded
        random vector, each one Jacobi sweep with omega = -1 and a zero
        right-hand side (no normalisation: (1 + lambda)^iters stays far
        below the FP32 range), then 1 + lambda_l = |x_k| / |x_(k-1)|. The
        estimate approaches lambda_max from below; a limit of 1.8 leaves
        10 % for that. One host synchronisation for all levels. The caps
        are rounded down to steps of 2^(1/16) so that small changes between
        hierarchy rebuilds keep the same value (a captured solve that
        bakes the damping in stays valid). Returns the caps (also stored in
        omega_cap; the coarsest level, solved directly, has none)."""
        buf = self._buffers()
        g = torch.Generator(device=self.device).manual_seed(seed)
        nrm = []
        for level in range(self.nlevels - 1):
            # The finest level's own operator (the matrix-free one when
            # set): the one the V-cycle smooths with.
            A = self._op(level)
            b = buf[level]
            zero = b['r']
            for z in zero:
                z.zero_()
            x, y = b['u'], b['v']
            for a in x:
                a.copy_(torch.rand(a.shape, dtype=a.dtype, device=a.device,
                                   generator=g))
            for _ in range(iters):
                A.jacobi(x, zero, y, -1.0)
                x, y = y, x
            nrm.append(torch.stack([
                sum((a.double()**2).sum() for a in x),
                sum((a.double()**2).sum() for a in y)]))
        vals = torch.stack(nrm).cpu().tolist() if nrm else []
        lam, caps = [], []
        for last, prev in vals:
            v = (last / prev)**0.5 - 1 if prev > 0 else 0.
            lam.append(v)
            c = limit / max(v, 1e-300)
            caps.append(2.0**(math.floor(16 * math.log2(c)) / 16))
        self.lambda_max = lam + [None]
        self.omega_cap = caps + [None]
        return self.omega_cap

    def level_omega(self, level, omega):
        """Damping used on `level` for the caller's omega."""
        if self.omega_cap is None or self.omega_cap[level] is None:
            return omega
        return min(omega, self.omega_cap[level])

    def omega_key(self):
        return None if self.omega_cap is None else tuple(self.omega_cap)

    # -- setup -----------------------------------------------------------
    def update_A(self, A0, set_lower=False):
        """A0: BlockStencil of the finest operator (cast to self.dtype), or
        the reference's nvar x nvar list of SparseOperators (util.py:
        577-586).

        A second call with an operator of the same structure refills the
        level operators and coarse inverses in place, so device addresses
        stay fixed (a captured V-cycle graph stays valid)."""
        if isinstance(A0, (list, tuple)):
            A0 = blocks_to_stencil(A0, self.nnw[0], self.dtype)
        if A0.nw != self.nnw[0] or A0.nvar_in != self.nvar:
            raise ValueError('operator grid/fields do not match the levels')
        same = (self.AA is not None and self.AA[0].entries == A0.entries
                and self.AA[0].sym == A0.sym)
        if not same:
            AA = [A0.astype(self.dtype)]
            if AA[0] is A0:
                AA[0] = BlockStencil(A0.coef.clone(), A0.entries, A0.nw,
                                     nvar_in=A0.nvar_in,
                                     nvar_out=A0.nvar_out, sym=A0.sym)
            for level in range(1, self.nlevels):
                AA.append(galerkin(AA[-1], self.restriction, self.rw,
                                   self.galerkin_mode))
            self.AA = AA
            self._coarse_inv = {}
            self._tune()
        else:
            self.AA[0].coef.copy_(A0.coef)
            if self.AA[0].sym:
                self.AA[0].zero_outside()
            for level in range(1, self.nlevels):
                galerkin(self.AA[level - 1], self.restriction, self.rw,
                         self.galerkin_mode, out=self.AA[level])
            for level in list(self._coarse_inv):
                self._coarse_inv[level].copy_(self._invert(level))
        return self

    def _tune(self):
        """Autotune the sweep launch shape of every stored level on the
        level buffers."""
        buf = self._buffers()
        for level, A in enumerate(self.AA):
            if A.sym:
                b = buf[level]
                A.tune(b['u'], b['f'], b['v'])

    def tune_in_cycle(self, run):
        """Re-pick the sweep launch shape of every stored symmetric level
        (finest first) by timing `run()`, one application of the cycle as
        the solver makes it, with each candidate of that level in turn:
        between a level's launches the other levels' traffic evicts its
        coefficients from the caches, which a sweep timed alone on its own
        buffers does not see (on the 1025^2 cavity the FP64 513^2 level
        times fastest with two cells per thread alone and slower than one
        cell per thread inside the cycle). Shapes are kernel arguments:
        call before the cycle is captured into a graph."""
        from odil_amd import autotune
        buf = self._buffers()
        for level in range(len(self.AA) - 1):
            A = self._op(level)
            if not isinstance(A, BlockStencil) or not A.sym:
                continue
            contig = A._contig(buf[level]['u'])
            cand = A.sweep_candidates(contig, per=(1, ))
            if (A.lanes, A.per) not in cand:
                cand.append((A.lanes, A.per))

            def install(c, A=A):
                A.lanes, A.per = c
            autotune.pick_in_situ(A.tune_key(contig), cand, install, run)
        return self

    def _invert(self, level):
        """Dense FP64 solution operator of the coarsest level.

        coarse_reg == 0: the inverse. coarse_reg > 0: the Tikhonov
        least-squares operator (M^T M + d I)^-1 M^T with
        d = coarse_reg * max diag(M^T M). Singular coarse systems (the
        cavity's constant pressure) are also inconsistent under the
        reference restriction (injection rows at coarse boundary nodes), so
        an inverse would amplify the incompatible part; least squares
        discards it, like the minimum-norm solution."""
        M = self.AA[level].to_dense()
        if not self.coarse_reg:
            # LAPACK-style results may come back column-major; the GEMV
            # kernel reads row-major.
            return torch.linalg.inv(M).contiguous()
        G = M.T @ M
        d = self.coarse_reg * G.diagonal().max()
        G.diagonal().add_(d)
        n = G.shape[0]
        if n not in self._spd:
            self._spd[n] = SpdSolver(n, G.device)
        return self._spd[n].solve(G, M.T.contiguous()).clone()

    def _coarse(self, level):
        if level not in self._coarse_inv:
            self._coarse_inv[level] = self._invert(level)
        return self._coarse_inv[level]

    def prepare(self, ndirect=5):
        """Precompute the coarse inverse used by step() (outside any graph
        capture)."""
        lv = self._direct_level(ndirect)
        self._coarse(lv)
        return lv

    def _direct_level(self, ndirect):
        for level in range(self.nlevels):
            if level + 1 >= self.nlevels or min(self.nnw[level]) <= ndirect:
                return level
        return self.nlevels - 1

    def _op(self, level):
        if level == 0 and self.fine is not None:
            return self.fine
        return self.AA[level]

    # -- reference-shaped operations -------------------------------------
    def _lvl(self, v, level):
        """Views of a list of level vectors (flat or shaped) as fields."""
        shp = tuple(self.nnw[level])
        return [x.reshape(shp) for x in v]

    def residual(self, A, u, f):
        """r = f - A u (multigrid.py:536-541); A is a level's BlockStencil
        or the reference's block list. Returns vectors shaped like f."""
        if isinstance(A, (list, tuple)):
            A = blocks_to_stencil(A, tuple(A[0][0].nw), self.dtype)
        level = self.nnw.index(tuple(A.nw))
        r = A.residual(self._lvl(u, level), self._lvl(f, level),
                       [torch.empty(tuple(self.nnw[level]), dtype=A.dtype,
                                    device=self.device) for _ in f])
        return [a.reshape(b.shape) for a, b in zip(r, f)]

    def smoother_jacobi(self, level, u, f, omega=1.):
        """In-place weighted Jacobi u += omega (f - A u) / D
        (multigrid.py:543-552): u is modified and returned."""
        uu = self._lvl(u, level)
        tmp = [torch.empty_like(x) for x in uu]
        self._op(level).jacobi(uu, self._lvl(f, level), tmp, omega)
        for a, b in zip(uu, tmp):
            a.copy_(b)
        return u

    def smoother_direct(self, level, u, f):
        """Direct solve on `level` (multigrid.py:554-565): the dense
        solution operator applied by the GEMV kernel. Returns new vectors
        shaped like f."""
        Minv = self._coarse(level)
        ff = [x.contiguous() for x in self._lvl(f, level)]
        out = [torch.empty_like(x) for x in ff]
        N = int(np.prod(self.nnw[level]))
        dt = _lib.dtype_tag(ff[0].dtype)
        st = _lib.lib().odil_dense_gemv(c_int(dt), c_int(dt),
                                        c_int(self.nvar), c_int(N), ptr(Minv),
                                        ptrs(ff), ptrs(out), stream())
        check(st, 'dense_gemv')
        return [a.reshape(b.shape) for a, b in zip(out, f)]

    def step(self, u, f, smoother=None, level=0, ndirect=5, pre=2, post=2,
             omega=None):
        """One V-cycle (multigrid.py:567-602). `smoother(level, u, f)`
        returns the smoothed u, e.g. partial(mg.smoother_jacobi, omega=w)
        as util.py:588-595 passes it; `omega=` alone builds that smoother.
        Like the reference, the recursive call does not pass `ndirect`,
        `pre`, `post`: coarser levels use 5, 2, 2. u and f are lists of
        level vectors (flat or shaped); the result has their shape."""
        if smoother is None:
            w = 0.6 if omega is None else omega
            smoother = lambda lv, a, b: self.smoother_jacobi(lv, a, b, w)
        elif not callable(smoother):
            raise ValueError('smoother must be callable(level, u, f)')
        if level + 1 >= self.nlevels or min(self.nnw[level]) <= ndirect:
            return self.smoother_direct(level, u, f)
        for _ in range(pre):
            u = smoother(level, u, f)
        A = self._op(level)
        r = A.residual(self._lvl(u, level), self._lvl(f, level),
                       [torch.empty(tuple(self.nnw[level]), dtype=self.dtype,
                                    device=self.device) for _ in f])
        rc = restrict(r, self.nnw[level], self.rw,
                      restriction=self.restriction)
        du = self.step([torch.zeros_like(x) for x in rc], rc, smoother,
                       level=level + 1)
        u = [x.clone() for x in u]
        prolong_add(self._lvl(du, level + 1), self._lvl(u, level),
                    self.nnw[level])
        for _ in range(post):
            u = smoother(level, u, f)
        return u

    # -- allocation-free V-cycle on preallocated buffers ----------------
    def _buffers(self):
        if self._buf is None:
            b = []
            for n in self.nnw:
                # One allocation per vector (fields x_k = x_0 + k N): the
                # stencil kernels read such x through a single base pointer.
                mk = lambda: list(torch.zeros((self.nvar, ) + tuple(n),
                                              dtype=self.dtype,
                                              device=self.device).unbind(0))
                b.append({'u': mk(), 'v': mk(), 'f': mk(), 'r': mk()})
            self._buf = b
        return self._buf

    def vcycle(self, u, f, omega=0.6, ndirect=5, pre=2, post=2, zero=False,
               f_src=None, first_done=False):
        """V-cycle with the same arithmetic as step() but on preallocated
        level buffers (no allocation, capturable in a HIP graph). `u` is
        updated in place; `f` is read only. zero=True: the cycle starts from
        u = 0 whatever u holds (the first pre-smoothing sweep is then one
        pass, y = omega f / d). f_src (with zero=True): the right-hand side
        in another precision, converted into `f` by that first pass.
        first_done (with zero=True, pre > 0, no f_src): the caller has
        already stored that first pass in the buffer named by
        first_sweep()."""
        buf = self._buffers()
        if f_src is not None and not (zero and pre > 0):
            for a, b in zip(f, f_src):
                a.copy_(b)
            f_src = None
        if first_done and not (zero and pre > 0 and f_src is None
                               and not self._is_direct(0, ndirect)):
            raise ValueError('first_done needs zero=True, pre > 0, no f_src '
                             'and a finest level that is not solved directly')
        self._vc(0, u, f, buf[0]['v'], omega, ndirect, pre, post, zero, f_src,
                 first_done=first_done)
        return u

    def first_sweep(self, omega):
        """(diagonal, output fields, damping) of the finest level's first
        sweep from zero in vcycle(zero=True): y = damping f / diagonal into
        the output fields, which vcycle(first_done=True) then takes as
        done."""
        return (self._op(0).diagonal(), self._buffers()[0]['v'],
                self.level_omega(0, omega))

    def _is_direct(self, level, ndirect):
        return level + 1 >= self.nlevels or min(self.nnw[level]) <= ndirect

    def _vc(self, level, u, f, spare, omega, ndirect, pre, post, zero=False,
            f_src=None, first_done=False):
        # `u` holds the iterate (ignored with zero=True); `spare` is a
        # same-shape scratch. first_done (with zero=True): the first sweep
        # from zero is already in `spare`. Returns u.
        if self._is_direct(level, ndirect):
            Minv = self._coarse(level)
            N = int(np.prod(self.nnw[level]))
            dt = _lib.dtype_tag(f[0].dtype)
            if zero:  # u = Minv f directly
                out = u
            else:
                out = spare
            check(_lib.lib().odil_dense_gemv(c_int(dt), c_int(dt),
                                             c_int(self.nvar), c_int(N),
                                             ptr(Minv), ptrs(f), ptrs(out),
                                             stream()), 'dense_gemv')
            if out is not u:
                for a, b in zip(u, out):
                    a.copy_(b)
            return u
        A = self._op(level)
        buf = self._buffers()
        omega_in = omega
        omega = self.level_omega(level, omega_in)
        cur, oth = u, spare
        k0 = 0
        if zero and pre > 0:
            if not first_done:
                A.jacobi_zero(f, oth, omega, f_src=f_src)
            cur, oth = oth, cur
            k0 = 1
        elif zero:
            for x in u:
                x.zero_()
        for _ in range(k0, pre):
            A.jacobi(cur, f, oth, omega)
            cur, oth = oth, cur
        r = buf[level]['r']
        A.residual(cur, f, r)
        nxt = buf[level + 1]
        # The coarse cycle starts from zero (pre = 2 there): unless it is
        # the direct solve, the restriction also does its first sweep.
        fused = not self._is_direct(level + 1, 5) and self.coarse_pre > 0
        if fused:
            restrict_smooth(r, self.nnw[level], self.rw, nxt['f'],
                            self.AA[level + 1].diagonal(), nxt['v'],
                            self.level_omega(level + 1, omega_in),
                            restriction=self.restriction)
        else:
            restrict(r, self.nnw[level], self.rw, nxt['f'],
                     restriction=self.restriction)
        res = self._vc(level + 1, nxt['u'], nxt['f'], nxt['v'], omega_in, 5,
                       self.coarse_pre, self.coarse_post, zero=True,
                       first_done=fused and self.coarse_pre > 0)
        prolong_add(res, cur, self.nnw[level])
        for _ in range(post):
            A.jacobi(cur, f, oth, omega)
            cur, oth = oth, cur
        if cur is not u:
            for a, b in zip(u, cur):
                a.copy_(b)
        return u


class Multigrid:
    """The reference's sparse-matrix `Multigrid` (multigrid.py:72-339) on
    the GPU: same constructor (`nw`, `nvar`, `nlevels`, `restriction`; `mod`
    and `dtype` are accepted and unused), `update_A` with the finest matrix
    as a SciPy sparse matrix (or a `BlockStencil`), `smoother_jacobi`,
    `smoother_direct` and `step(u, f, smoother, level, ndirect, pre,
    post)` on flat FP64 device vectors (fields concatenated, C-order).

    The matrix is converted once to block-stencil form
    (`BlockStencil.from_csr`), the coarse matrices are R A T
    (`galerkin_mode='exact'`, what the reference's sparse products compute),
    smoothing, transfers and the coarsest solve (a dense inverse applied by
    a GEMV kernel) run in the HIP kernels of `MultigridOp`. Like the
    reference, `step` passes `ndirect` to the coarser levels but not `pre`
    and `post` (multigrid.py:331-332). Node grids of odd size only.

    The three smoothers of the reference: point Jacobi, block-diagonal
    Jacobi (`full=True`, the per-cell nvar x nvar blocks factored once per
    `update_A`, an exactly singular block raising numpy's LinAlgError like
    `numpy.linalg.solve`) and forward Gauss-Seidel in the flat numbering
    (lexicographic within a field, fields in order), swept in wavefronts of
    independent cells, one launch per wavefront: a sequential method, about
    nvar (3 n0 + n1) launches per sweep on a 2D grid, meant for small
    levels and reference checks rather than the Newton hot path.
    """

    def __init__(self, nw, nvar=1, nlevels=None, restriction='full',
                 mod=None, dtype=None, device=None):
        self._mg = MultigridOp(nw, nvar=nvar, nlevels=nlevels,
                               restriction=restriction, dtype=torch.float64,
                               device=device, coarse_reg=0.,
                               galerkin_mode='exact')
        self.nnw = self._mg.nnw
        self.nlevels = self._mg.nlevels
        self.nvar = nvar
        self.dim = len(self.nnw[0])
        self.restriction = restriction
        self.AA = None
        self._buf = {}
        self._factors = {}
        self._gs_tab = {}
        self._levels_given = False

    def update_A(self, A0, DD=None, compute_lower=False):
        """Set the operator (multigrid.py:116-157): the finest matrix (SciPy
        sparse or `BlockStencil`), coarsened by R A T, or a list of one
        matrix per level, used as given. `DD` is recomputed from the
        matrices, as in the reference; `compute_lower` is accepted and not
        needed (Gauss-Seidel reads the lower triangle from the stored
        operator)."""
        self._factors = {}
        self._gs_tab = {}
        if isinstance(A0, list):
            if len(A0) != self.nlevels:
                raise ValueError('Expected len(A0)={:}, got {:}'.format(
                    self.nlevels, len(A0)))
            AA = [self._stencil(a, level) for level, a in enumerate(A0)]
            self._mg.AA = AA
            self._mg._coarse_inv = {}
            self._mg._tune()
            self._levels_given = True
        else:
            A0 = self._stencil(A0, 0)
            if self._levels_given:
                self._mg.AA = None  # rebuild the Galerkin levels
                self._levels_given = False
            self._mg.update_A(A0)
        self.AA = self._mg.AA
        return self

    def _stencil(self, A, level):
        if not isinstance(A, BlockStencil):
            A = BlockStencil.from_csr(A, self.nnw[level], self.nvar,
                                      device=self._mg.device)
        if A.nw != tuple(self.nnw[level]) or A.nvar_in != self.nvar:
            raise ValueError('level {} operator grid/fields do not match '
                             '{} x {}'.format(level, self.nvar,
                                              self.nnw[level]))
        if any(abs(x) > 2 for _, _, sh in A.entries for x in sh):
            raise ValueError('Multigrid.update_A: couplings beyond 2 nodes '
                             '(e.g. a periodic wrap) are not supported by '
                             'the stencil Galerkin product')
        return A.astype(torch.float64)

    def _work(self, level, key):
        k = (level, key)
        if k not in self._buf:
            self._buf[k] = [torch.empty(tuple(self.nnw[level]),
                                        dtype=torch.float64,
                                        device=self._mg.device)
                            for _ in range(self.nvar)]
        return self._buf[k]

    def _fields(self, v, level):
        n = (self.nvar, ) + tuple(self.nnw[level])
        if v.numel() != int(np.prod(n)) or not v.is_contiguous():
            raise ValueError('level {} vectors hold {} values (contiguous), '
                             'got {}'.format(level, int(np.prod(n)),
                                             v.numel()))
        return list(v.view(n))

    def smoother_jacobi(self, level, u, f, omega=1., full=False):
        """u += omega D^-1 (f - A u) in place (multigrid.py:284-299): D the
        diagonal of the diagonal blocks, or with full=True the per-cell
        nvar x nvar matrix of every block's diagonal."""
        uu, ff = self._fields(u, level), self._fields(f, level)
        if not full:
            self._mg.smoother_jacobi(level, uu, ff, omega)
            return u
        A = self._mg.AA[level]
        r = A.residual(uu, ff, self._work(level, 'r'))
        lu, piv = self._block_factors(level)
        st = _lib.lib().odil_block_jacobi(
            c_int(_lib.F64), c_int(self.nvar), c_int(uu[0].numel()), ptr(lu),
            ptr(piv), ptrs(r), ptrs(uu), c_double(omega), stream())
        check(st, 'block_jacobi')
        return u

    def _block_factors(self, level):
        if level in self._factors:
            return self._factors[level]
        A = self._mg.AA[level]
        nv = self.nvar
        zero = (0, ) * self.dim
        at = {(o, i): q for q, (o, i, s) in enumerate(A.entries) if s == zero}
        dq = []
        for o in range(nv):
            for i in range(nv):
                q = at.get((o, i))
                if q is None and A.sym:
                    q = at.get((i, o))  # mirror at shift 0: same cell
                dq.append(-1 if q is None else q)
        N = int(np.prod(self.nnw[level]))
        dev = self._mg.device
        lu = torch.empty((nv * nv, N), dtype=torch.float64, device=dev)
        piv = torch.empty((nv, N), dtype=torch.int32, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        st = _lib.lib().odil_block_diag_factor(
            c_int(_lib.F64), c_int(nv), c_int(N),
            (ctypes.c_int * len(dq))(*dq), c_int(len(A.entries)),
            ptr(A.coef), ptr(lu), ptr(piv), ptr(flag), stream())
        check(st, 'block_diag_factor')
        if int(flag.item()):
            raise np.linalg.LinAlgError(
                'Singular matrix (a diagonal block on level {})'.format(level))
        self._factors[level] = (lu, piv)
        return lu, piv

    def smoother_gauss_seidel(self, level, u, f, omega=1.):
        """u += omega (L^-1 (f - (A - L) u) - u) in place, L the lower
        triangle of A with its diagonal (multigrid.py:301-305)."""
        A = self._mg.AA[level]
        if A.sym:
            raise ValueError('Gauss-Seidel needs the operator in full '
                             '(non-symmetric) storage')
        uu, ff = self._fields(u, level), self._fields(f, level)
        if level not in self._gs_tab:
            ent = [v for o, i, s in A.entries for v in (o, i) + _shift3(s)]
            self._gs_tab[level] = (ctypes.c_int * len(ent))(*ent)
        st = _lib.lib().odil_stencil_gauss_seidel(
            c_int(_lib.F64), c_int(self.dim), _grid_arr(A.nw),
            c_int(self.nvar), ptr(A.t_ent), self._gs_tab[level],
            c_int(len(A.entries)), ptr(A.coef), ptrs(uu), ptrs(ff),
            ptrs(self._work(level, 'x')), c_double(omega), stream())
        check(st, 'gauss_seidel')
        return u

    def smoother_direct(self, level, u, f):
        out = self._mg.smoother_direct(level, self._fields(u, level),
                                       self._fields(f, level))
        return torch.stack(out).reshape(-1)

    def step(self, u, f, smoother, level=0, ndirect=5, pre=2, post=2):
        """One V-cycle from u (multigrid.py:312-339); returns the new u."""
        if level + 1 >= self.nlevels or min(self.nnw[level]) <= ndirect:
            return self.smoother_direct(level, u, f)
        for _ in range(pre):
            u = smoother(level, u, f)
        A = self._mg.AA[level]
        r = A.residual(self._fields(u, level), self._fields(f, level),
                       [torch.empty(tuple(self.nnw[level]), dtype=u.dtype,
                                    device=u.device)
                        for _ in range(self.nvar)])
        rc = torch.stack(restrict(r, self.nnw[level], self._mg.rw,
                                  restriction=self.restriction)).reshape(-1)
        du = self.step(torch.zeros_like(rc), rc, smoother, level=level + 1,
                       ndirect=ndirect)
        u = u.clone()
        prolong_add(self._fields(du, level + 1), self._fields(u, level),
                    self.nnw[level])
        for _ in range(post):
            u = smoother(level, u, f)
        return u
