#!/usr/bin/env python3
"""Benchmark: Newton steps/s of the lid-driven cavity at 1025x1025 (the
1024^2-class grid of the reference's node-based multigrid, 2^10 + 1 cells per
axis), mixed precision vs FP64, plus residual / adjoint kernel bandwidth.

One step = one Newton iteration: residual and right-hand side -J^T f (FP64),
`maxiter` iterations of flexible CG on J^T J (FP64, matrix-free)
preconditioned by one V(2,2) cycle, state update; every `setup_every`-th step
also rebuilds the stencil Jacobian, the stored normal operator J^T J, the
Galerkin hierarchy (7 levels, down to 17x17) and the coarsest FP64
least-squares operator. Mixed precision runs the preconditioner in FP32.
Inputs start resident in HBM (zero state, Re = 100, lid u = 1).

Prints one JSON line (see the keys below). `--gpus N` with N > 1 is not
supported: this build runs on one GPU.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--n', type=int, default=1025)
    p.add_argument('--re', type=float, default=100.)
    p.add_argument('--maxiter', type=int, default=10)
    p.add_argument('--linsolver', default='pcg', choices=['vcycle', 'pcg'])
    p.add_argument('--omega', type=float, default=0.6)
    p.add_argument('--nlvl', type=int, default=7)
    p.add_argument('--setup_every', type=int, default=12,
                   help='rebuild the stored multigrid hierarchy every k '
                   'Newton steps (the Krylov operator and finest level are '
                   'always current)')
    p.add_argument('--smooth', default='2,2',
                   help='pre,post Jacobi sweeps on the finest level')
    p.add_argument('--coarse_smooth', default='2,2',
                   help='pre,post Jacobi sweeps on the levels below the '
                   'finest (2,2: the V(2,2) cycle of the benchmarked '
                   'configuration and of the reference recursion)')
    p.add_argument('--setup_first', type=int, default=2,
                   help='also rebuild the hierarchy at each of the first k '
                   'Newton steps')
    p.add_argument('--overlap_setup', type=int, default=0,
                   help='1: build the next hierarchy on a second stream '
                   'during the linear solve, adopted one step later')
    p.add_argument('--krylov', default='auto', choices=['auto', 'fp32', 'fp64'],
                   help='Krylov precision of the mixed solve (auto: FP32)')
    p.add_argument('--refine', type=int, default=0,
                   help='FP64 iterative-refinement passes per linear solve')
    p.add_argument('--tol', type=float, default=1e-8)
    p.add_argument('--tol_maxiter', type=int, default=18,
                   help='linear iterations per Newton step of the '
                        'time-to-tolerance runs (both precisions)')
    p.add_argument('--max_epochs', type=int, default=150)
    p.add_argument('--kernel_reps', type=int, default=50)
    p.add_argument('--no_tol_run', action='store_true')
    p.add_argument('--no_cpu', action='store_true')
    p.add_argument('--cpu_n', type=int, default=129)
    p.add_argument('--copy_probe', action='store_true',
                   help='first run a plain streaming kernel over a known '
                   'size (torch.neg on 256 MiB of FP32) whose PMC '
                   'FETCH_SIZE / WRITE_SIZE calibrate the traffic counters')
    p.add_argument('--no_graph', action='store_true',
                   help='launch the linear solve eagerly instead of from a '
                   'captured HIP graph (kernel tracers see every launch)')
    return p.parse_args()


def make_solver(args, precision):
    import torch
    from odil_amd.newton import NewtonMultigrid
    from odil_amd.operators import Cavity
    op = Cavity((args.n, args.n), 1.0 / args.re)
    sm = [int(x) for x in getattr(args, 'smooth', '2,2').split(',')]
    nm = NewtonMultigrid(op, nlvl=args.nlvl, maxiter=args.maxiter,
                         smooth_pre=sm[0], smooth_post=sm[1],
                         omega=args.omega, precision=precision,
                         linsolver=args.linsolver, galerkin_mode='exact',
                         restriction='transpose',
                         use_graph=not args.no_graph,
                         setup_every=args.setup_every,
                         overlap_setup=bool(getattr(args, 'overlap_setup',
                                                    0)),
                         setup_first=getattr(args, 'setup_first', 0),
                         coarse_smooth=tuple(
                             int(x) for x in
                             getattr(args, 'coarse_smooth', '2,2').split(',')),
                         krylov=args.krylov if precision == 'mixed' else 'auto',
                         refine=args.refine if precision == 'mixed' else 0)
    fields = [torch.zeros(args.n, args.n, dtype=torch.float64,
                          device='cuda') for _ in range(3)]
    return op, nm, fields


def steps_per_second(args, precision):
    import torch
    _, nm, fields = make_solver(args, precision)
    for _ in range(args.warmup):
        nm.step(fields)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nm.step(fields)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return args.steps / (t1 - t0), (t1 - t0) / args.steps, nm


def _event_time(fn, reps, use_graph):
    """Average milliseconds of `fn` over `reps` launches, timed with HIP
    events on torch's current stream (the stream the kernels run on). With
    use_graph the launches are captured once and replayed, so host launch
    gaps do not enter."""
    import torch
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    if use_graph:
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(reps):
                fn()
        graph.replay()
        torch.cuda.synchronize()
        s = torch.cuda.current_stream()
        e0.record(s)
        graph.replay()
        e1.record(s)
    else:
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def level1_sweep_kernel(args, nm):
    """The kernel with the largest share of the mixed Newton step: the
    weighted-Jacobi sweep of the stored (upper-half symmetric) Galerkin
    operator on multigrid level 1 (513 x 513 nodes, FP32), run on the
    solver's own operator and level buffers. Algorithmic bytes per launch:
    every stored coefficient array once, x, f read and y written
    (nent + 3 * 3) * N1 * 4."""
    A = nm.mg.AA[1]
    buf = nm.mg._buffers()[1]
    n1 = 1
    for n in A.nw:
        n1 *= n
    B = A.coef.element_size()
    ms = _event_time(lambda: A.jacobi(buf['u'], buf['f'], buf['v'],
                                      args.omega),
                     args.kernel_reps, not args.no_graph)
    nbytes = (len(A.entries) + 3 * A.nvar_out) * n1 * B
    return {'ms': ms, 'GBps': nbytes / (ms * 1e-3) / 1e9, 'bytes': nbytes,
            'nent': len(A.entries), 'grid': list(A.nw)}


def normal_sweep_kernel(args, nm):
    """The kernel with the second largest share of the mixed Newton step:
    the matrix-free weighted-Jacobi sweep of N = J^T J on the finest level
    (`k_normal_apply<float>`, mode 2, 1025 x 1025 cells), run on the
    solver's FP32 state copy and diagonal with synthetic x and b. Algorithmic
    bytes per launch: state, x, b and diag(N) read and y written, three
    fields each: 15 * N * 4."""
    import torch
    st = nm.state_w
    g = torch.Generator(device='cuda').manual_seed(1)
    x = [torch.rand(a.shape, dtype=a.dtype, device='cuda', generator=g)
         for a in st]
    b = [torch.rand(a.shape, dtype=a.dtype, device='cuda', generator=g)
         for a in st]
    y = [torch.empty_like(a) for a in st]
    ms = _event_time(lambda: nm.op.normal_apply(st, x, y, b=b,
                                                d=nm.fine.diag,
                                                omega=args.omega, mode=2),
                     args.kernel_reps, not args.no_graph)
    n = st[0].numel()
    nbytes = 15 * n * st[0].element_size()
    return {'ms': ms, 'GBps': nbytes / (ms * 1e-3) / 1e9, 'bytes': nbytes,
            'grid': list(st[0].shape)}


TRAFFIC_FILE = os.path.join(ROOT, 'profiles',
                            'traffic_apply_sym_level1_cavity1025.txt')


def measured_traffic():
    """HBM bytes per launch of the dominant kernel from the committed PMC
    summary (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known
    size), or None when that summary is absent."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith('hbm_bytes_per_launch:'):
            return float(line.split(':', 1)[1])
    return None


def copy_ceiling():
    """Measured device-copy bandwidth (STREAM-copy-like, 512 MiB read +
    512 MiB written by torch's copy kernel), as a practical ceiling beside
    the 8 TB/s peak."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.float32, device='cuda')
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        y.copy_(x)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / 10
    del x, y
    return 2 * (1 << 27) * 4 / (ms * 1e-3) / 1e9


def copy_probe():
    """Plain streaming kernel of known size for counter calibration:
    y = -x over 2^26 FP32 values (256 MiB read, 256 MiB written), 5 times."""
    import torch
    x = torch.ones(1 << 26, dtype=torch.float32, device='cuda')
    y = torch.empty_like(x)
    for _ in range(5):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    del x, y


def time_to_tol(args, precision):
    """Wall-clock of Newton steps from the zero state until the max RMS
    residual is <= tol, `tol_maxiter` linear iterations per step (the
    fastest budget for both precisions; the step rate is quoted at
    `maxiter`)."""
    import copy
    import torch
    a = copy.copy(args)
    a.maxiter = args.tol_maxiter
    _, nm, fields = make_solver(a, precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = nm.run(fields, args.max_epochs, tol=args.tol)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ep, rms, _ = hist[-1]
    return {'seconds': t1 - t0, 'newton_steps': ep, 'final_rms': max(rms),
            'reached': max(rms) <= args.tol}, fields


def kernel_bandwidth(args):
    """Average duration of the cavity residual and fused loss-adjoint kernels
    at n x n. `kernel_reps` launches are captured in a HIP graph (eager
    with --no_graph) and timed with events on the launch stream.
    Algorithmic bytes: residual reads u, v, p and writes fu, fv, fp (6 N B);
    the adjoint reads u, v, p and writes three gradients (6 N B)."""
    import torch
    from odil_amd import _lib
    from odil_amd.operators import Cavity
    n = args.n
    N = n * n
    op = Cavity((n, n), 1.0 / args.re)
    out = {}
    for dt, name in ((torch.float32, 'fp32'), (torch.float64, 'fp64')):
        g = torch.Generator(device='cuda').manual_seed(0)
        fields = [torch.rand(n, n, dtype=dt, device='cuda', generator=g)
                  for _ in range(3)]
        res = [torch.empty_like(f) for f in fields]
        grads = [torch.empty_like(f) for f in fields]
        dtag = _lib.dtype_tag(dt)
        B = 4 if dt == torch.float32 else 8
        for kind in ('residual', 'adjoint'):
            k = 0 if kind == 'residual' else 1
            partial = torch.empty((3, op.nblocks(k, dtag)),
                                  dtype=torch.float64, device='cuda')

            def call():
                if kind == 'residual':
                    op.launch_residual(fields, res, partial)
                else:
                    op.launch_loss_grad(fields, grads, partial)
            ms = _event_time(call, args.kernel_reps, not args.no_graph)
            nbytes = 6 * N * B
            out['{}_{}'.format(kind, name)] = {
                'ms': ms, 'GBps': nbytes / (ms * 1e-3) / 1e9,
                'bytes': nbytes}
    return out


def cpu_baseline(args):
    """One Newton step of the same algorithm in the CPU oracle (NumPy/SciPy,
    float64) on a bounded sample: the cavity at cpu_n x cpu_n with the same
    level count rule and V-cycle count, scaled to 1025^2 by cell count (the
    multigrid step is linear in the number of cells). Timed with one thread
    and again with all the host's threads."""
    import torch
    ncpu = min(os.cpu_count() or 1, 16)
    one = _cpu_step(args, 1)
    allc = _cpu_step(args, ncpu)
    one['all_cores'] = {'value': allc['value'], 'cores': ncpu,
                        'seconds': allc['seconds']}
    torch.set_num_threads(ncpu)
    return one


def _cpu_step(args, threads):
    import numpy as np
    import torch
    torch.set_num_threads(threads)
    from oracle import multigrid as omg
    from oracle import operators as ops
    from oracle import tape
    n = args.cpu_n
    lv = len(omg.MultigridOp((n, n)).nnw)
    st = {k: np.zeros((n, n)) for k in 'uvp'}
    t0 = time.perf_counter()
    const, grads, desc = tape.eval_grad(ops.make_cavity_operator(1 / args.re),
                                        ops.cavity_domain(n), st)
    stf = omg.coeffs_from_tape(grads, desc, 3, 3)
    mm, matr = omg.normal_operator(stf, (n, n), 3)
    rhs = [-sum(mm[i][j].mul_transpose_field(const[i].ravel())
                for i in range(3)) for j in range(3)]
    M = omg.MultigridOp((n, n), nvar=3, nlevels=min(lv, args.nlvl),
                        galerkin_mode='exact', coarse_reg=1e-14)
    M.update_A(matr)
    sol = [np.zeros(n * n) for _ in range(3)]
    for _ in range(args.maxiter):
        sol = M.step(sol, rhs, args.omega, ndirect=3)
    t1 = time.perf_counter()
    sec = (t1 - t0) * (args.n**2) / (n * n)
    return {'value': 1.0 / sec, 'unit': 'Newton steps/s', 'cores': threads,
            'kind': 'port', 'seconds': t1 - t0,
            'sample': ('one oracle Newton step (tape Jacobian, stencil J^T J, '
                       'exact Galerkin, {} V(2,2) cycles) at {}x{}, {:.2f} s, '
                       'scaled to {}x{} by cell count').format(
                           args.maxiter, n, n, t1 - t0, args.n, args.n)}


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only (this build has no multi-GPU '
              'path); got --gpus {}'.format(args.gpus), file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print('no GPU visible', file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        copy_probe()
    sps_mixed, ms_mixed, nm_mixed = steps_per_second(args, 'mixed')
    # The level-1 stored sweep and the finest-level matrix-free sweep take
    # the two largest shares of the mixed step (about equal, the first by
    # ~1 % in the kernel trace): the roofline is quoted on the first.
    dom = level1_sweep_kernel(args, nm_mixed)
    nsw = normal_sweep_kernel(args, nm_mixed)
    del nm_mixed
    sps_64, ms_64, _ = steps_per_second(args, 'fp64')
    tol = {}
    if not args.no_tol_run:
        tol['mixed'], f_mixed = time_to_tol(args, 'mixed')
        tol['fp64'], f_64 = time_to_tol(args, 'fp64')
        # Error of the mixed-precision solution against the FP64 one.
        d = [torch.max(torch.abs(a - b)).item()
             for a, b in zip(f_mixed[:2], f_64[:2])]
        pm = f_mixed[2] - f_mixed[2].mean()
        p6 = f_64[2] - f_64[2].mean()
        d.append(torch.max(torch.abs(pm - p6)).item())
        tol['mixed_vs_fp64_maxabs'] = {'u': d[0], 'v': d[1], 'p': d[2]}
    kb = kernel_bandwidth(args)
    kb['mg_level1_jacobi_fp32'] = dom
    kb['normal_jacobi_fp32'] = nsw
    line = {
        'metric': 'Newton steps/s + residual/adjoint GB/s at 1024^2, '
                  'mixed vs FP64, same tol',
        'value': sps_mixed,
        'unit': 'Newton steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_mixed * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp32+fp64',
        'data': 'synthetic',
        'config': {
            'workload': 'lid-driven cavity {0}x{0}, Re={1:g}, Newton with '
                        'GMG normal-equation solve ({5}, {3} iterations/step, '
                        '{2} levels, V(2,2) on the finest and V({9}) on '
                        'the coarser levels, omega={4}, hierarchy rebuilt '
                        'every {6} steps and at the first {8}{7})'.format(
                            args.n, args.re, args.nlvl, args.maxiter,
                            args.omega, 'FCG+V-cycle' if args.linsolver ==
                            'pcg' else 'V-cycles', args.setup_every,
                            ' on a second stream' if args.overlap_setup
                            else '', args.setup_first, args.coarse_smooth),
            'grid': [args.n, args.n],
            'mixed': 'FP32 flexible CG and V-cycle, FP64 Newton residual, '
                     'update and coarsest solve' + (
                         ', {} FP64 refinement pass(es)'.format(args.refine)
                         if args.refine else ''),
            'time_to_tol_iterations_per_step': args.tol_maxiter,
            'parallelism': 'single GPU'},
        'fp64_steps_per_s': sps_64,
        'fp64_ms_per_step': ms_64 * 1e3,
        'mixed_over_fp64_step_speedup': sps_mixed / sps_64,
        'time_to_tol': tol,
        'kernels': kb,
        'roofline': {
            'bound': 'hbm',
            'kernel': 'k_apply_sym<float> (level-1 Galerkin Jacobi sweep, '
                      '{}x{}, {} stored coefficient fields, FP32)'.format(
                          dom['grid'][0], dom['grid'][1], dom['nent']),
            'achieved': dom['GBps'], 'peak': HBM_PEAK_GBPS, 'unit': 'GB/s',
            'frac': dom['GBps'] / HBM_PEAK_GBPS,
            'traffic': measured_traffic(),
            'measured_copy_GBps': copy_ceiling()},
    }
    if 'mixed' in tol and 'fp64' in tol:
        line['time_to_tol_speedup'] = (tol['fp64']['seconds'] /
                                       tol['mixed']['seconds'])
    if not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
