"""Dense symmetric positive definite solves for the coarsest multigrid level.

`SpdSolver(n)` computes X = G^-1 B for an n x n SPD G and n-column B (FP64)
by a right-looking blocked Cholesky factorisation: each diagonal block of
64 columns is factored and inverted in LDS by `odil_chol_block`; the panel
(L_ik = A_ik W_kk^T) and trailing updates (A_22 -= L_21 L_21^T) and the
block forward and backward substitutions for L L^T X = B are GEMMs. The
sequence is captured once per solver into a HIP graph and replayed. It
replaces the library Cholesky and triangular solves of the coarsest
level's Tikhonov solution operator (the factorisation dominates the
hierarchy setup otherwise).
"""

import ctypes

import torch

from odil_amd import _lib
from odil_amd._lib import c_int, check, stream
from odil_amd.graphs import capture


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class SpdSolver:
    """X = G^-1 B for SPD G (n x n, FP64); `solve(G, B)` returns a view of
    an internal buffer valid until the next call."""

    def __init__(self, n, device, use_graph=True):
        self.n = int(n)
        self.nb = _lib.lib().odil_chol_block_size()
        self.nblk = (self.n + self.nb - 1) // self.nb
        f64 = torch.float64
        self.A = torch.zeros((self.n, self.n), dtype=f64, device=device)
        self.Y = torch.zeros((self.n, self.n), dtype=f64, device=device)
        self.W = torch.zeros((self.nblk, self.nb, self.nb), dtype=f64,
                             device=device)
        self.T = torch.zeros((self.nb, self.n), dtype=f64, device=device)
        self.P = torch.zeros((self.n, self.nb), dtype=f64, device=device)
        self.use_graph = use_graph
        self._graph = None

    def _blocks(self):
        for k in range(self.nblk):
            k0 = k * self.nb
            yield k, k0, min(self.n, k0 + self.nb)

    def _body(self):
        A, Y, n = self.A, self.Y, self.n
        lib = _lib.lib()
        # Factorisation: A = L L^T (lower, in place), W_k = L_kk^-1.
        for k, k0, k1 in self._blocks():
            check(lib.odil_chol_block(_p(A), c_int(n), c_int(k0),
                                      c_int(k1 - k0), _p(self.W[k]),
                                      stream()), 'chol_block')
            if k1 < n:
                w = self.W[k, :k1 - k0, :k1 - k0]
                P = self.P[:n - k1, :k1 - k0]
                torch.mm(A[k1:, k0:k1], w.t(), out=P)
                A[k1:, k0:k1].copy_(P)
                A[k1:, k1:].addmm_(P, P.t(), alpha=-1.0)
        # Forward substitution: Y <- L^-1 Y.
        for k, k0, k1 in self._blocks():
            w = self.W[k, :k1 - k0, :k1 - k0]
            t = self.T[:k1 - k0]
            torch.mm(w, Y[k0:k1], out=t)
            Y[k0:k1].copy_(t)
            if k1 < n:
                Y[k1:].addmm_(A[k1:, k0:k1], Y[k0:k1], alpha=-1.0)
        # Backward substitution: Y <- L^-T Y.
        for k, k0, k1 in reversed(list(self._blocks())):
            w = self.W[k, :k1 - k0, :k1 - k0]
            t = self.T[:k1 - k0]
            torch.mm(w.t(), Y[k0:k1], out=t)
            Y[k0:k1].copy_(t)
            if k0 > 0:
                Y[:k0].addmm_(A[k0:k1, :k0].t(), Y[k0:k1], alpha=-1.0)

    def solve(self, G, B):
        if G.shape != (self.n, self.n) or B.shape != (self.n, self.n):
            raise ValueError('SpdSolver: expected {0}x{0} operands'.format(
                self.n))
        self.A.copy_(G)
        self.Y.copy_(B)
        if not self.use_graph or torch.cuda.is_current_stream_capturing():
            self._body()
            return self.Y
        if self._graph is None:
            # First use: run eagerly (the result stays in Y), then record
            # the same work; capture does not execute it.
            self._body()
            self._graph = capture(self._body, self.A.device)
            return self.Y
        self._graph.replay()
        return self.Y
