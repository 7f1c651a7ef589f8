"""Many independent Poisson instances of uneven size on one GPU.

* `PoissonBatch` packs the instances (u, src, grad) into single buffers and
  evaluates every instance's residual, loss mean(f^2) and adjoint
  (2/N) J^T f in ONE persistent launch over a worklist of tiles
  (`odil_poisson_batch_loss_grad`); the tile shape is chosen by the
  autotuner on the batch itself.
* `PoissonBatchSolver` solves every instance's normal equations
  J^T J u = J^T src (one Newton step of the linear problem, the
  `optimize_multigridop_base` step of util.py:509-604) by flexible CG with
  a GMG V-cycle preconditioner. The Poisson Jacobian does not depend on u or
  src, so all instances of one size share one multigrid hierarchy. The
  batch solve (every instance, every iteration: ~700 kernels per instance)
  is captured into HIP graphs of at most `graph_instances` instances each,
  so graph size stays bounded as the batch grows, and the graphs replay
  back to back with no host work in between.

`batch_sizes` draws sizes log-uniformly between 64 and 1024. The batched
loss/adjoint launch takes any sizes (rectangular too); the solver needs
node grids of 2^k + 1 points per axis (multigrid.py:362-370), so for it the
sizes are snapped to the nearest 2^k + 1 (`snap=True`).
"""

import ctypes

import numpy as np
import torch

from odil_amd import _lib, autotune
from odil_amd._lib import c_int, check, ptr, stream
from odil_amd.multigrid import MultigridOp, normal_operator, BlockStencil
from odil_amd.newton import DeviceScalars, FusedNormal, VectorOps
from odil_amd.operators import Poisson


def batch_sizes(count, seed, lo=64, hi=1024, snap=True):
    """`count` grid sizes n = round(lo (hi/lo)^U), U ~ U(0, 1); with snap,
    moved to the nearest 2^k + 1 in log scale (what the multigrid solver
    needs)."""
    rng = np.random.default_rng(seed)
    raw = lo * (hi / lo)**rng.uniform(0, 1, count)
    if not snap:
        return [int(round(n)) for n in raw]
    ks = np.arange(int(np.log2(lo)), int(np.log2(hi)) + 1)
    cand = 2.0**ks + 1
    idx = np.abs(np.log(raw[:, None]) - np.log(cand[None, :])).argmin(1)
    return [int(cand[i]) for i in idx]


def gaussian_mixture(n, seed, nblobs=3):
    """Synthetic source on the unit square: `nblobs` Gaussians with centres
    U(0.2, 0.8)^2, widths U(0.05, 0.15), amplitudes U(-1, 1)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    xx, yy = np.meshgrid(x, x, indexing='ij')
    s = np.zeros((n, n))
    for _ in range(nblobs):
        cx, cy = rng.uniform(0.2, 0.8, 2)
        w = rng.uniform(0.05, 0.15)
        a = rng.uniform(-1, 1)
        s += a * np.exp(-((xx - cx)**2 + (yy - cy)**2) / (2 * w**2))
    return s


class PoissonBatch:
    """Instances i = 0..count-1 of f = lap u - src on n_i x n_i cells of the
    unit square, Dirichlet u = 0, packed back to back in C order."""

    def __init__(self, sizes, sources, dtype=torch.float32, device=None):
        self.device = device or torch.device('cuda')
        self.dtype = dtype
        self.sizes = [(int(n), int(n)) if np.isscalar(n) else tuple(n)
                      for n in sizes]
        self.count = len(self.sizes)
        ncell = [a * b for a, b in self.sizes]
        self.offsets = np.concatenate([[0], np.cumsum(ncell)]).astype(np.int64)
        self.total = int(self.offsets[-1])
        self.src = torch.empty(self.total, dtype=dtype, device=self.device)
        for i, s in enumerate(sources):
            self.view(self.src, i).copy_(torch.as_tensor(np.asarray(s)))
        self.u = torch.zeros(self.total, dtype=dtype, device=self.device)
        self.grad = torch.empty_like(self.u)
        t = lambda a, dt: torch.tensor(a, dtype=dt, device=self.device)
        self.t_off = t(self.offsets[:-1], torch.int64)
        self.t_n0 = t([a for a, _ in self.sizes], torch.int32)
        self.t_n1 = t([b for _, b in self.sizes], torch.int32)
        self.t_dx = t([1. / a for a, _ in self.sizes], torch.float64)
        self.t_dy = t([1. / b for _, b in self.sizes], torch.float64)
        self.loss = torch.empty(self.count, dtype=torch.float64,
                                device=self.device)
        self._tiles = {}
        self.shape = None
        self.nblocks = 4 * torch.cuda.get_device_properties(
            self.device).multi_processor_count

    def view(self, buf, i):
        a, b = self.sizes[i]
        return buf[int(self.offsets[i]):int(self.offsets[i + 1])].view(a, b)

    def shapes(self):
        arr = (ctypes.c_int * 32)()
        n = _lib.lib().odil_poisson_batch_shapes(
            c_int(_lib.dtype_tag(self.dtype)), arr, c_int(16))
        return [(arr[2 * k], arr[2 * k + 1]) for k in range(n)]

    def tiles(self, tx, ty):
        """Tile worklist (instance, x0, y0, 0) grouped by instance."""
        key = (tx, ty)
        if key not in self._tiles:
            rows, tstart = [], [0]
            for i, (a, b) in enumerate(self.sizes):
                for x0 in range(0, a, tx):
                    for y0 in range(0, b, ty):
                        rows.append((i, x0, y0, 0))
                tstart.append(len(rows))
            tl = torch.tensor(np.array(rows, dtype=np.int32),
                              device=self.device)
            ts = torch.tensor(np.array(tstart, dtype=np.int32),
                              device=self.device)
            part = torch.empty(len(rows), dtype=torch.float64,
                               device=self.device)
            self._tiles[key] = (tl, ts, part)
        return self._tiles[key]

    def launch(self, shape=None, u=None, grad=None):
        """One persistent launch: loss of every instance into self.loss and
        the adjoints into `grad` (default self.grad)."""
        tx, ty = shape or self.shape or self.shapes()[0]
        tl, ts, part = self.tiles(tx, ty)
        u = self.u if u is None else u
        grad = self.grad if grad is None else grad
        st = _lib.lib().odil_poisson_batch_loss_grad(
            c_int(_lib.dtype_tag(self.dtype)), c_int(self.count),
            ptr(self.t_off), ptr(self.t_n0), ptr(self.t_n1), ptr(self.t_dx),
            ptr(self.t_dy), ptr(tl), c_int(tl.shape[0]), ptr(ts), c_int(tx),
            c_int(ty), c_int(self.nblocks), ptr(u), ptr(self.src), ptr(grad),
            ptr(part), ptr(self.loss), stream())
        check(st, 'poisson_batch_loss_grad')
        return self.loss, grad

    def tune(self):
        """Pick the tile shape of the batched launch by timing it."""
        key = ('poisson_batch', tuple(self.sizes), str(self.dtype))
        self.shape = autotune.pick(key, self.shapes(),
                                   lambda s: self.launch(s))
        return self.shape

    def instance_ops(self):
        """Single-grid `Poisson` operators viewing each instance."""
        return [Poisson(self.sizes[i], self.view(self.src, i))
                for i in range(self.count)]

    def loss_grad_loop(self, ops=None):
        """Reference point: one single-grid launch per instance."""
        ops = ops or self.instance_ops()
        out = []
        for i, op in enumerate(ops):
            nb = op.nblocks(1, _lib.dtype_tag(self.dtype))
            if not hasattr(op, '_part'):
                op._part = torch.empty((1, nb), dtype=torch.float64,
                                       device=self.device)
            op.launch_loss_grad([self.view(self.u, i)],
                                [self.view(self.grad, i)], op._part)
            out.append(op._part)
        return out

    def algorithmic_bytes(self):
        """Read u and src, write grad, per cell."""
        return 3 * self.total * self.u.element_size()


class _SizeClass:
    """Shared multigrid hierarchy and work vectors of one grid size."""

    def __init__(self, shape, wdtype, nlvl, device, coarse_reg):
        f64 = torch.float64
        z = torch.zeros(shape, dtype=f64, device=device)
        self.op = Poisson(shape, z)
        jac, _ = self.op.jacobian([z])
        N = normal_operator(BlockStencil.from_jacobian(jac), wdtype,
                            sym=True)
        self.mg = MultigridOp(shape, nvar=1, nlevels=nlvl,
                              restriction='transpose', dtype=wdtype,
                              device=device, coarse_reg=coarse_reg,
                              galerkin_mode='exact')
        self.mg.update_A(N)
        self.state64 = [z]
        self.state_w = [torch.zeros(shape, dtype=wdtype, device=device)]
        self.fine = FusedNormal(self.op, self.state_w)
        self.fine.refresh()
        self.mg.fine = self.fine
        mk = lambda dt: [torch.zeros(shape, dtype=dt, device=device)]
        self.r, self.z, self.p, self.q, self.zo, self.x = (
            mk(f64), mk(f64), mk(f64), mk(f64), mk(f64), mk(f64))
        self.b_w, self.x_w = mk(wdtype), mk(wdtype)
        self.sc = DeviceScalars(8, device)


class PoissonBatchSolver:
    """Normal-equation solve of every instance: `maxiter` flexible-CG
    iterations on J^T J (FP64, matrix-free) preconditioned by one V(2,2)
    cycle of the shared hierarchy of the instance's size (FP32 in 'mixed').
    `solve()` replays the batch's HIP graphs (at most `graph_instances`
    instances per graph).

    A graph of the whole 256-instance batch (about 180 000 kernel nodes in
    one dependency chain) ended the process with a host segmentation fault
    during its first use; graphs of at most MAX_GRAPH_INSTANCES instances
    (about 6 000 nodes) are the largest verified, so larger values of
    `graph_instances` are refused."""

    MAX_GRAPH_INSTANCES = 8

    def __init__(self, batch, precision='mixed', maxiter=10, omega=0.5,
                 nlvl=None, ndirect=3, use_graph=True, graph_instances=8):
        self.batch = batch
        gi = int(graph_instances)
        if gi < 1 or gi > self.MAX_GRAPH_INSTANCES:
            raise ValueError(
                'graph_instances must be 1..{} (larger single graphs are '
                'not verified: a whole-batch graph crashed the host '
                'process)'.format(self.MAX_GRAPH_INSTANCES))
        self.graph_instances = gi
        self.maxiter = int(maxiter)
        self.omega = float(omega)
        self.ndirect = ndirect
        self.use_graph = use_graph
        self.wdtype = torch.float32 if precision == 'mixed' else \
            torch.float64
        dev = batch.device
        self.vec = VectorOps(dev)
        self.classes = {}
        for shp in sorted(set(batch.sizes)):
            self.classes[shp] = _SizeClass(
                shp, self.wdtype, nlvl, dev,
                1e-12 if precision == 'mixed' else 1e-14)
            self.classes[shp].mg.prepare(ndirect)
        # FP64 solution and source per instance (the batch buffers may be
        # FP32); the instance operator only supplies the residual.
        self.u64 = [torch.zeros(s, dtype=torch.float64, device=dev)
                    for s in batch.sizes]
        self.ops = [Poisson(s, batch.view(batch.src, i).double())
                    for i, s in enumerate(batch.sizes)]
        self._graphs = None

    def _precond(self, k):
        """One V-cycle from zero on k.r; returns the working-precision
        result (an FP32 cycle converts k.r in its first sweep)."""
        if self.wdtype == torch.float64:
            k.mg.vcycle(k.z, k.r, omega=self.omega, ndirect=self.ndirect,
                        zero=True)
            return k.z
        k.mg.vcycle(k.x_w, k.b_w, omega=self.omega, ndirect=self.ndirect,
                    zero=True, f_src=k.r)
        return k.x_w

    def _solve_instance(self, i):
        # Flexible CG as NewtonMultigrid._pcg_iters: zo holds the latest
        # preconditioned residual in FP64.
        k = self.classes[self.batch.sizes[i]]
        op, u = self.ops[i], self.u64[i]
        v, s = self.vec, k.sc
        op.jtf([u], k.r, -1.0)             # r = -J^T f(u)
        k.x[0].zero_()
        v.fcg_z(self._precond(k), k.r, k.zo, s.p(0), s.p(1), s.p(3))
        k.p[0].copy_(k.zo[0])
        for _ in range(self.maxiter):
            k.op.normal_apply(k.state64, k.p, k.q, mode=0)
            v.dot(k.p, k.q, s.p(2))
            v.axpy2(k.x, k.p, k.r, k.q, s.p(0), s.p(2))
            v.fcg_z(self._precond(k), k.r, k.zo, s.p(0), s.p(1), s.p(3))
            v.xpay(k.p, k.zo, s.p(0), s.p(1), sub=s.p(3))
        u.add_(k.x[0])

    def _chunks(self):
        n, k = self.batch.count, self.graph_instances
        return [range(a, min(a + k, n)) for a in range(0, n, k)]

    def solve(self):
        """One solve of every instance (u += correction)."""
        if not self.use_graph:
            for i in range(self.batch.count):
                self._solve_instance(i)
            return
        if self._graphs is None:
            graphs = []
            for chunk in self._chunks():
                for i in chunk:          # eager run (the first solve)
                    self._solve_instance(i)
                torch.cuda.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):  # recorded, not executed
                    for i in chunk:
                        self._solve_instance(i)
                graphs.append(g)
            self._graphs = graphs
            return
        for g in self._graphs:
            g.replay()

    def reset(self):
        for u in self.u64:
            u.zero_()

    def residual_rms(self):
        """Per-instance RMS residual of the current solutions."""
        out = []
        for op, u in zip(self.ops, self.u64):
            _, ssq = op.residual([u], want_fields=False)
            out.append((ssq[0].item() / op.ncell)**0.5)
        return out
