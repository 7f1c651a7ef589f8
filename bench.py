#!/usr/bin/env python3
"""Benchmark: Newton steps/s of the lid-driven cavity at 1025x1025 (the
1024^2-class grid of the reference's node-based multigrid, 2^10 + 1 cells per
axis), mixed precision vs FP64 at the same tolerance, plus residual /
adjoint kernel bandwidth.

One step = one Newton iteration: residual and right-hand side -J^T f (FP64),
`maxiter` iterations of flexible CG on J^T J (matrix-free) preconditioned by
one V(1,1) cycle, state update; the stored hierarchy (stencil Jacobian,
J^T J, Galerkin levels down to 17x17, coarsest FP64 least-squares operator)
is rebuilt at the first 2 steps and then every 12. Mixed precision runs the
Krylov solve and the preconditioner in FP32 with the FP64 Newton residual as
iterative refinement. `maxiter` and the cycle shape default to the fastest
to the 1e-8 RMS residual of those measured (V(1,1) and V(2,2), 18-26
iterations; the same for both precisions), so `value` is the step rate of
the same solve that `time_to_tol` times to tolerance.
Inputs start resident in HBM (zero state, Re = 100, lid u = 1); Re = 1000 is
timed to tolerance beside it.

The CPU baseline (the reference's own algorithm: Newton with the direct
normal-equation solve, `linsolver.py:35-45`, restated by the oracle) runs in
child processes started before this process touches the GPU, with the BLAS
and OpenMP thread counts set in their environment.

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
    p.add_argument('--maxiter', type=int, default=22,
                   help='linear iterations per Newton step (22 with V(1,1) '
                   'cycles: the fastest to tol of the budgets and cycle '
                   'shapes measured, for both precisions and Reynolds '
                   'numbers)')
    p.add_argument('--linsolver', default='pcg', choices=['vcycle', 'pcg'])
    p.add_argument('--omega', type=float, default=0.6)
    p.add_argument('--nlvl', type=int, default=7)
    p.add_argument('--setup_every', type=int, default=12,
                   help='rebuild the stored multigrid hierarchy every k '
                   'Newton steps (the Krylov operator and finest level are '
                   'always current)')
    p.add_argument('--smooth', default='1,1',
                   help='pre,post Jacobi sweeps on the finest level (the '
                   "reference's --smooth_pre/--smooth_post, default 2,2 "
                   'there)')
    p.add_argument('--coarse_smooth', default='1,1',
                   help='pre,post Jacobi sweeps on the levels below the '
                   'finest (1,1 with --smooth 1,1: the V(1,1) cycle of the '
                   'benchmarked configuration)')
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
    p.add_argument('--tol_maxiter', type=int, default=None,
                   help='linear iterations per Newton step of the '
                        'time-to-tolerance runs (default: --maxiter)')
    p.add_argument('--re_tol', default='1000',
                   help='further Reynolds numbers timed to tolerance '
                   '(comma list, empty for none)')
    p.add_argument('--max_epochs', type=int, default=150)
    p.add_argument('--kernel_reps', type=int, default=50)
    p.add_argument('--no_tol_run', action='store_true')
    p.add_argument('--no_cpu', action='store_true')
    p.add_argument('--cpu_sizes', default='65,129',
                   help='grids at which the CPU reference algorithm is '
                   'timed to tolerance with one thread')
    p.add_argument('--cpu_all_sizes', default='65',
                   help='grids timed again with every host thread')
    p.add_argument('--cpu_leg', default=None, help=argparse.SUPPRESS)
    p.add_argument('--rotate_mb', type=float, default=600.,
                   help='kernel timing loops cycle through buffer sets of '
                   'at least this many MB (above the 256 MB Infinity '
                   'Cache) so the quoted GB/s come from HBM')
    p.add_argument('--copy_probe', action='store_true',
                   help='first run a plain streaming kernel over a known '
                   'size (torch.neg on 256 MiB of FP32) whose PMC '
                   'FETCH_SIZE / WRITE_SIZE calibrate the traffic counters')
    p.add_argument('--no_graph', action='store_true',
                   help='launch the linear solve eagerly instead of from a '
                   'captured HIP graph (kernel tracers see every launch)')
    a = p.parse_args()
    if a.tol_maxiter is None:
        a.tol_maxiter = a.maxiter
    return a


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


def _event_time(fn, reps, use_graph, nrot=1):
    """Average milliseconds of one launch of `fn(k)` over `reps` launches,
    k = 0, 1, ... mod nrot cycling through the caller's buffer sets, timed
    with HIP events on torch's current stream (the stream the kernels run
    on). With use_graph the launches are captured once and replayed, so host
    launch gaps do not enter."""
    import torch
    for k in range(max(3, nrot)):
        fn(k % nrot)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    if use_graph:
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for k in range(reps):
                fn(k % nrot)
        graph.replay()
        torch.cuda.synchronize()
        s = torch.cuda.current_stream()
        e0.record(s)
        graph.replay()
        e1.record(s)
    else:
        e0.record(s)
        for k in range(reps):
            fn(k % nrot)
        e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _nrot(args, nbytes):
    """Buffer sets to cycle through so one cycle moves >= rotate_mb."""
    return max(1, -(-int(args.rotate_mb * 1e6) // int(nbytes)))


def level1_sweep_kernel(args, nm):
    """The kernel with the largest share of the mixed Newton step: the
    weighted-Jacobi sweep of the stored (upper-half symmetric) Galerkin
    operator on multigrid level 1 (513 x 513 nodes, FP32), on copies of the
    solver's own level-1 operator and buffers (enough copies that one cycle
    of launches reads more than the Infinity Cache holds). Algorithmic bytes
    per launch: every stored coefficient array once, x, f read and y written
    (nent + 3 * 3) * N1 * 4."""
    import copy
    import torch
    A = nm.mg.AA[1]
    buf = nm.mg._buffers()[1]
    n1 = 1
    for n in A.nw:
        n1 *= n
    B = A.coef.element_size()
    nbytes = (len(A.entries) + 3 * A.nvar_out) * n1 * B
    def dup(fields):
        # Same layout as the level buffers: one allocation, field k at k N.
        return list(torch.stack(fields).unbind(0))
    sets = [(A, buf['u'], buf['f'], buf['v'])]
    for k in range(1, _nrot(args, nbytes)):
        Ak = copy.copy(A)
        Ak.coef = A.coef.clone()
        sets.append((Ak, dup(buf['u']), dup(buf['f']), dup(buf['v'])))
    ms = _event_time(lambda k: sets[k][0].jacobi(sets[k][1], sets[k][2],
                                                 sets[k][3], args.omega),
                     args.kernel_reps, not args.no_graph, len(sets))
    return {'ms': ms, 'GBps': nbytes / (ms * 1e-3) / 1e9, 'bytes': nbytes,
            'nent': len(A.entries), 'grid': list(A.nw),
            'lanes': A.lanes, 'per': A.per, 'buffer_sets': len(sets)}


def normal_sweep_kernel(args, nm):
    """The kernel with the second largest share of the mixed Newton step:
    the matrix-free weighted-Jacobi sweep of N = J^T J on the finest level
    (`k_normal_apply<float>`, mode 2, 1025 x 1025 cells), on copies of the
    solver's FP32 state and diagonal with synthetic x and b. Algorithmic
    bytes per launch: state, x, b and diag(N) read and y written, three
    fields each: 15 * N * 4."""
    import torch
    st0 = nm.state_w
    n = st0[0].numel()
    nbytes = 15 * n * st0[0].element_size()
    g = torch.Generator(device='cuda').manual_seed(1)
    sets = []
    for k in range(_nrot(args, nbytes)):
        st = st0 if k == 0 else [a.clone() for a in st0]
        d = nm.fine.diag if k == 0 else [a.clone() for a in nm.fine.diag]
        x = [torch.rand(a.shape, dtype=a.dtype, device='cuda', generator=g)
             for a in st0]
        b = [torch.rand(a.shape, dtype=a.dtype, device='cuda', generator=g)
             for a in st0]
        sets.append((st, x, b, d, [torch.empty_like(a) for a in st0]))

    def call(k):
        st, x, b, d, y = sets[k]
        nm.op.normal_apply(st, x, y, b=b, d=d, omega=args.omega, mode=2)
    ms = _event_time(call, args.kernel_reps, not args.no_graph, len(sets))
    return {'ms': ms, 'GBps': nbytes / (ms * 1e-3) / 1e9, 'bytes': nbytes,
            'grid': list(st0[0].shape), 'buffer_sets': len(sets)}


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


def time_to_tol(args, precision, re=None, n=None, nlvl=None):
    """Wall-clock of Newton steps from the zero state until the max RMS
    residual is <= tol, `tol_maxiter` linear iterations per step (the same
    settings for both precisions), at Reynolds number `re` on an n x n grid
    (defaults: the bench's)."""
    import copy
    import torch
    a = copy.copy(args)
    a.maxiter = args.tol_maxiter
    if re is not None:
        a.re = re
    if n is not None:
        a.n, a.nlvl = n, nlvl
    _, nm, fields = make_solver(a, precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = nm.run(fields, args.max_epochs if re is None or re <= 100
                  else 4 * args.max_epochs, tol=args.tol)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ep, rms, _ = hist[-1]
    return {'seconds': t1 - t0, 'newton_steps': ep, 'final_rms': max(rms),
            'reached': max(rms) <= args.tol}, fields


def _solution_diff(fa, fb):
    """Max |a - b| of u, v and of p up to its constant (null mode)."""
    import torch
    d = [torch.max(torch.abs(a - b)).item() for a, b in zip(fa[:2], fb[:2])]
    pa = fa[2] - fa[2].mean()
    pb = fb[2] - fb[2].mean()
    d.append(torch.max(torch.abs(pa - pb)).item())
    return {'u': d[0], 'v': d[1], 'p': d[2]}


def kernel_bandwidth(args):
    """Average duration of the cavity residual and fused loss-adjoint kernels
    at n x n. `kernel_reps` launches cycle through enough input/output sets
    that one cycle moves more than the Infinity Cache holds; they are
    captured in a HIP graph (eager with --no_graph) and timed with events on
    the launch stream. Algorithmic bytes: residual reads u, v, p and writes
    fu, fv, fp (6 N B); the adjoint reads u, v, p and writes three
    gradients (6 N B)."""
    import torch
    from odil_amd import _lib
    from odil_amd.operators import Cavity
    n = args.n
    N = n * n
    op = Cavity((n, n), 1.0 / args.re)
    out = {}
    for dt, name in ((torch.float32, 'fp32'), (torch.float64, 'fp64')):
        g = torch.Generator(device='cuda').manual_seed(0)
        B = 4 if dt == torch.float32 else 8
        nbytes = 6 * N * B
        nrot = _nrot(args, nbytes)
        sets = []
        for _ in range(nrot):
            fields = [torch.rand(n, n, dtype=dt, device='cuda', generator=g)
                      for _ in range(3)]
            sets.append((fields, [torch.empty_like(f) for f in fields]))
        dtag = _lib.dtype_tag(dt)
        for kind in ('residual', 'adjoint'):
            k = 0 if kind == 'residual' else 1
            partial = torch.empty((3, op.nblocks(k, dtag)),
                                  dtype=torch.float64, device='cuda')

            def call(j):
                fields, outs = sets[j]
                if kind == 'residual':
                    op.launch_residual(fields, outs, partial)
                else:
                    op.launch_loss_grad(fields, outs, partial)
            ms = _event_time(call, args.kernel_reps, not args.no_graph,
                             nrot)
            out['{}_{}'.format(kind, name)] = {
                'ms': ms, 'GBps': nbytes / (ms * 1e-3) / 1e9,
                'bytes': nbytes, 'buffer_sets': nrot}
        del sets
    return out


# -- CPU baseline --------------------------------------------------------------
def _cpu_leg(spec):
    """Child-process body: the reference's Newton with the direct
    normal-equation solve (oracle restatement of util.py:242-268 and
    linsolver.py:35-45, SuperLU with MMD_ATA ordering) from the zero state
    to max RMS residual <= tol on each grid of `spec`; prints one JSON
    line. Imports only NumPy/SciPy/torch-CPU (no device)."""
    import numpy as np
    from oracle import newton, tape
    from oracle import operators as ops
    out = []
    for n in spec['sizes']:
        dom = ops.cavity_domain(n)
        op = ops.make_cavity_operator(1.0 / spec['re'])
        packed = tape.pack_fields(dom, {k: np.zeros((n, n)) for k in 'uvp'})
        t0 = time.perf_counter()
        for ep in range(spec['max_epochs'] + 1):
            st = tape.unpack_fields(dom, packed)
            const, m = tape.linearize(op, dom, st, epoch=ep)
            rms = max(newton.rms_residuals(const))
            if rms <= spec['tol'] or ep == spec['max_epochs']:
                break
            c = np.hstack([e.flatten() for e in const])
            packed = packed + newton.solve_direct(m, -c)
        sec = time.perf_counter() - t0
        out.append({'n': n, 'seconds': sec, 'newton_steps': ep,
                    'steps_per_s': ep / sec, 'final_rms': rms,
                    'reached': rms <= spec['tol']})
    print(json.dumps(out), flush=True)


def cpu_baseline(args):
    """Runs the CPU legs in child processes (started before this process
    touches the GPU), one with a single thread and one with every host
    thread (at most 16, this process's share of the GPU host), the thread
    counts set in the children's environment before NumPy loads."""
    import subprocess
    ncpu = os.cpu_count() or 1
    nall = min(ncpu, 16)
    legs = {}
    for key, threads, sizes in (('one', 1, args.cpu_sizes),
                                ('all', nall, args.cpu_all_sizes)):
        sizes = [int(x) for x in sizes.split(',') if x]
        if not sizes:
            continue
        env = dict(os.environ)
        for var in ('OMP_NUM_THREADS', 'OPENBLAS_NUM_THREADS',
                    'MKL_NUM_THREADS', 'NUMEXPR_NUM_THREADS',
                    'VECLIB_MAXIMUM_THREADS'):
            env[var] = str(threads)
        env['CUDA_VISIBLE_DEVICES'] = ''
        env['HIP_VISIBLE_DEVICES'] = ''
        spec = {'sizes': sizes, 're': args.re, 'tol': args.tol,
                'max_epochs': args.max_epochs}
        r = subprocess.run([sys.executable, os.path.abspath(__file__),
                            '--cpu_leg', json.dumps(spec)], env=env,
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError('CPU leg failed:\n' + r.stderr[-2000:])
        legs[key] = {'threads': threads,
                     'runs': json.loads(r.stdout.strip().splitlines()[-1])}
    one = legs['one']['runs']
    head = one[-1]
    res = {'value': head['steps_per_s'], 'unit': 'Newton steps/s',
           'cores': 1, 'kind': 'port', 'host_cpu_count': ncpu,
           'sample': ('the reference algorithm (Newton + direct '
                      'normal-equation solve, SuperLU MMD_ATA) restated by '
                      'the oracle, Re={:g}, zero state to RMS residual {:g}, '
                      'measured (not scaled) at {}; value = Newton steps/s '
                      'at {}x{} (a direct solve at 1025^2 does not fit a '
                      'bounded CPU run)').format(
                          args.re, args.tol,
                          ', '.join('{0}x{0}'.format(r['n']) for r in one),
                          head['n'], head['n']),
           'time_to_tol': {str(r['n']): r for r in one}}
    if 'all' in legs:
        res['all_cores'] = {'cores': legs['all']['threads'],
                            'time_to_tol': {str(r['n']): r
                                            for r in legs['all']['runs']}}
    return res


def main():
    args = parse()
    if args.cpu_leg is not None:
        _cpu_leg(json.loads(args.cpu_leg))
        return
    if args.gpus != 1:
        print('bench.py runs on one GPU only (this build has no multi-GPU '
              'path); got --gpus {}'.format(args.gpus), file=sys.stderr)
        sys.exit(2)
    # CPU legs first, in child processes, before this process touches the
    # GPU (counting devices below does not initialise it).
    cpu = None if args.no_cpu else cpu_baseline(args)
    import torch
    if not torch.cuda.is_available():
        print('no GPU visible', file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        copy_probe()
    sps_mixed, ms_mixed, nm_mixed = steps_per_second(args, 'mixed')
    # The level-1 stored sweep and the finest-level matrix-free sweep take
    # the two largest shares of the mixed step: the roofline is quoted on
    # the first.
    dom = level1_sweep_kernel(args, nm_mixed)
    nsw = normal_sweep_kernel(args, nm_mixed)
    del nm_mixed
    sps_64, ms_64, _ = steps_per_second(args, 'fp64')
    tol = {}
    if not args.no_tol_run:
        tol['mixed'], f_mixed = time_to_tol(args, 'mixed')
        tol['fp64'], f_64 = time_to_tol(args, 'fp64')
        tol['mixed_vs_fp64_maxabs'] = _solution_diff(f_mixed, f_64)
        tol['speedup'] = tol['fp64']['seconds'] / tol['mixed']['seconds']
        del f_mixed, f_64
        for re in [float(x) for x in args.re_tol.split(',') if x]:
            r = {}
            r['mixed'], fa = time_to_tol(args, 'mixed', re=re)
            r['fp64'], fb = time_to_tol(args, 'fp64', re=re)
            r['mixed_vs_fp64_maxabs'] = _solution_diff(fa, fb)
            r['speedup'] = r['fp64']['seconds'] / r['mixed']['seconds']
            tol['re{:g}'.format(re)] = r
            del fa, fb
        if cpu is not None:
            # The GPU solve at the CPU legs' grids (same settings otherwise).
            for r in cpu['time_to_tol'].values():
                n = r['n']
                nl = max(2, (n - 1).bit_length() - 4)
                g, _ = time_to_tol(args, 'mixed', n=n, nlvl=nl)
                g64, _ = time_to_tol(args, 'fp64', n=n, nlvl=nl)
                r['gpu_mixed_seconds'] = g['seconds']
                r['gpu_fp64_seconds'] = g64['seconds']
                r['gpu_newton_steps'] = [g['newton_steps'],
                                         g64['newton_steps']]
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
                        'GMG normal-equation solve ({5}, {3} iterations/step '
                        '(the budget of the time-to-tolerance runs), {2} '
                        'levels, V({10}) on the finest and V({9}) on '
                        'the coarser levels, omega={4}, hierarchy rebuilt '
                        'every {6} steps and at the first {8}{7})'.format(
                            args.n, args.re, args.nlvl, args.maxiter,
                            args.omega, 'FCG+V-cycle' if args.linsolver ==
                            'pcg' else 'V-cycles', args.setup_every,
                            ' on a second stream' if args.overlap_setup
                            else '', args.setup_first, args.coarse_smooth,
                            args.smooth),
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
    if 'speedup' in tol:
        line['time_to_tol_speedup'] = tol['speedup']
    if cpu is not None:
        line['cpu_baseline'] = cpu
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
