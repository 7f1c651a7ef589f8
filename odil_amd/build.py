"""Build the HIP/C-ABI library `libodil_amd.so` in-tree for gfx950.

Every `csrc/*.hip` file is compiled to an object with hipcc, then linked into
one shared library next to this file. Run `python -m odil_amd.build`.
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libodil_amd.so')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-munsafe-fp-atomics', '-Wno-unused-result', '-I' + CSRC, '-I' + INCLUDE
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for {}:\n{}\n{}'.format(
            src, ' '.join(cmd), r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))
    jobs = min(8, len(srcs))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950', '-o', tmp
           ] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n{}\n{}'.format(' '.join(cmd),
                                                        r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
