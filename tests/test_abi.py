"""The C-ABI library loads without a GPU, exports exactly the entry points
declared in include/odil_amd.h, and its host-only calls and argument checks
behave as documented (no kernel is launched here)."""

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'odil_amd.h')


def _declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', ' ', text, flags=re.S)
    return sorted(set(re.findall(r'\b(odil_\w+)\s*\(', text)))


@pytest.fixture(scope='module')
def lib():
    from odil_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('library not built (python -m odil_amd.build)')
    return _lib.lib()


def test_header_compiles_as_c():
    r = subprocess.run(['gcc', '-fsyntax-only', '-std=c99', '-Wall',
                        '-Werror', '-x', 'c', HEADER], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


def test_every_declared_symbol_is_exported(lib):
    names = _declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_undeclared_exports(lib):
    from odil_amd import _lib
    r = subprocess.run(['nm', '-D', '--defined-only', _lib.LIB_PATH],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip('nm unavailable')
    exported = sorted(set(re.findall(r'\bT (odil_\w+)', r.stdout)))
    assert exported == _declared()


def test_host_calls(lib):
    from odil_amd import _lib
    assert lib.odil_abi_version() == _lib.ABI_VERSION
    assert lib.odil_dot_partials() > 0
    ent = (ctypes.c_int * (3 * 64))()
    shifts = (ctypes.c_int * (2 * 16))()
    # cavity: 27 coefficient fields, 5 shifts (centre and the 4 neighbours)
    assert lib.odil_cavity_jac_entries(ent, shifts) == 27
    got = {(shifts[2 * s], shifts[2 * s + 1]) for s in range(5)}
    assert got == {(0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)}
    # (equation, field) blocks: the momentum equations freeze the advecting
    # velocity, so fu has no v block and fv no u block.
    blocks = {}
    for q in range(27):
        key = (ent[3 * q], ent[3 * q + 1])
        blocks[key] = blocks.get(key, 0) + 1
    assert blocks == {(0, 0): 5, (0, 2): 3, (1, 1): 5, (1, 2): 3,
                      (2, 0): 3, (2, 1): 3, (2, 2): 5}
    assert lib.odil_wave_jac_entries(ent, shifts) == 5
    assert lib.odil_poisson_jac_entries(ent, shifts) == 5
    assert lib.odil_laplace_jac_entries(ent, shifts) == 10
    assert lib.odil_cavity_nblocks(0, 1, 1025, 1025) > 0
    assert lib.odil_advection_nblocks(8, 16, 16) > 0


def test_argument_errors_do_not_launch(lib):
    """Invalid arguments return status 1 with a message, before any HIP
    call (so this runs on a CPU-only host)."""
    from odil_amd import _lib
    lib.odil_last_error.restype = ctypes.c_char_p
    nw = (ctypes.c_int * 3)(9, 9, 1)
    st = lib.odil_stencil_apply(ctypes.c_int(1), ctypes.c_int(4), nw,
                                ctypes.c_int(1), ctypes.c_int(1),
                                ctypes.c_int(1), None, None, None,
                                ctypes.c_int(1), None, None, None, None,
                                ctypes.c_int(0), ctypes.c_double(0.), None)
    assert st == 1
    assert b'ndim' in lib.odil_last_error()
    st = lib.odil_stencil_apply(ctypes.c_int(1), ctypes.c_int(2), nw,
                                ctypes.c_int(1), ctypes.c_int(1),
                                ctypes.c_int(1), None, None, None,
                                ctypes.c_int(100000), None, None, None, None,
                                ctypes.c_int(0), ctypes.c_double(0.), None)
    assert st == 3  # capacity
    st = lib.odil_dot(ctypes.c_int(1), ctypes.c_int(9), ctypes.c_longlong(4),
                      None, None, None, None, ctypes.c_int(0), None)
    assert st == 1
    with pytest.raises(ValueError):
        _lib.check(1, 'probe')
    with pytest.raises(_lib.OdilError):
        _lib.check(3, 'probe')


def _param_counts():
    text = open(HEADER).read()
    out = {}
    for m in re.finditer(r'\bint\s+(odil_\w+)\s*\(([^;]*?)\)\s*;', text,
                         re.S):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ('', 'void') else \
            params.count(',') + 1
    return out


def test_operator_entry_points_match_python_calls():
    """The shared Python operator base passes (dtype, shape..., operator
    parameters..., then fields, out, partial[, nblocks], stream): every
    operator's residual / loss_grad / jac declaration takes exactly that
    many arguments (ctypes does not check arity)."""
    import torch
    from odil_amd import _lib
    from odil_amd.operators import Cavity, CavityExtras
    counts = _param_counts()
    cav = Cavity((8, 9), 0.01)
    ops = [cav, CavityExtras(cav, kreg=0.1, device=torch.device('cpu'))]
    for op in ops:
        base = 1 + len(op.shape) + len(op._args_for(_lib.F64))
        for what, tail in (('residual', 4), ('loss_grad', 5), ('jac', 4)):
            name = 'odil_{}_{}'.format(op.prefix, what)
            assert counts[name] == base + tail, (name, counts[name])


def test_sweep_and_dense_limits_are_checked_on_the_host(lib):
    """Limits of the symmetric sweep (tiles per workgroup, products per
    field, shift reach, 32-bit offsets, lanes), of the Cholesky block and of
    axpby return their status before any launch (CPU-only host)."""
    lib.odil_last_error.restype = ctypes.c_char_p
    i3 = lambda *v: (ctypes.c_int * 3)(*v)
    ok_items = (ctypes.c_int * 10)()
    one = ctypes.c_void_p(1)   # never dereferenced: every case fails first
    fake = (ctypes.c_void_p * 3)(16, 32, 48)

    def sym(nw=(9, 9, 1), reach=(1, 1, 0), max_items=8, nent=5, lanes=1,
            per=1, mode=0, x=fake):
        return lib.odil_stencil_apply_sym(
            ctypes.c_int(0), ctypes.c_int(2), i3(*nw), ctypes.c_int(3),
            ok_items, ok_items, ok_items, i3(*reach), ctypes.c_int(max_items),
            ctypes.c_int(nent), ctypes.c_int(lanes), ctypes.c_int(per), one,
            x, fake, None, ctypes.c_int(mode), ctypes.c_double(0.), None)
    assert sym(x=None) == 1 and b'null' in lib.odil_last_error()
    assert sym(per=0) == 1 and b'tiles per workgroup' in lib.odil_last_error()
    assert sym(max_items=600) == 3
    assert sym(reach=(600, 0, 0)) == 1
    assert b'shifts beyond 511' in lib.odil_last_error()
    assert sym(reach=(1, 1, 1)) == 1          # 2D with a z reach
    assert sym(nw=(40000, 40000, 1), nent=72) == 3   # 32-bit offsets
    assert sym(nw=(50000, 50000, 1)) == 1            # 2^31 points
    assert sym(lanes=3) == 1

    assert sym(mode=2) in (1, 3) and sym(mode=1) == 1   # no f given
    assert lib.odil_chol_block_size() == 64
    assert lib.odil_chol_block(one, ctypes.c_int(128), ctypes.c_int(0),
                               ctypes.c_int(65), one, None) == 1
    assert lib.odil_chol_block(one, ctypes.c_int(32), ctypes.c_int(0),
                               ctypes.c_int(64), one, None) == 1
    assert lib.odil_axpby(ctypes.c_int(1), ctypes.c_longlong(-1), one, one,
                          one, ctypes.c_double(1.), ctypes.c_double(1.),
                          None) == 1
