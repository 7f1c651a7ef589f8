"""bench.py pieces that run without a GPU: the one-GPU guard and the CPU
baseline leg (the reference's direct Newton restated by the oracle, timed
to tolerance in a child process with its thread counts set)."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'),
                        '--gpus', '2'], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr
    assert r.stdout.strip() == ''


def test_cpu_baseline_leg_reaches_tolerance():
    sys.path.insert(0, ROOT)
    import bench
    old = sys.argv
    sys.argv = ['bench.py', '--cpu_sizes', '9,13', '--cpu_all_sizes', '9']
    try:
        args = bench.parse()
    finally:
        sys.argv = old
    res = bench.cpu_baseline(args)
    assert res['kind'] == 'port' and res['cores'] == 1
    assert res['host_cpu_count'] == os.cpu_count()
    runs = res['time_to_tol']
    assert set(runs) == {'9', '13'}
    for r in runs.values():
        assert r['reached'] and r['final_rms'] <= 1e-8
        assert r['newton_steps'] > 2 and r['seconds'] > 0
    assert res['value'] == runs['13']['steps_per_s']
    assert res['all_cores']['time_to_tol']['9']['reached']
