"""CPU checks of the reference-named facade (`odil_amd.odil`, odil.py:1-18)
and of the raw + XDMF field files (util_io.py)."""

import numpy as np
import pytest


def test_facade_names():
    from odil_amd import odil
    for name in ('Domain', 'State', 'Problem', 'History', 'MultigridDecomp',
                 'util_op', 'util', 'util_history', 'linsolver', 'optimizer',
                 'parse_raw_xmf', 'read_raw', 'write_raw_xmf',
                 'write_raw_with_xmf', 'tf'):
        assert hasattr(odil, name), name
    for name in ('optimize', 'optimize_newton', 'optimize_opt',
                 'optimize_multigrid', 'optimize_multigridop',
                 'add_arguments'):
        assert callable(getattr(odil.util, name)), name
    for name in ('checkpoint_save', 'checkpoint_load', 'extrap_quadh',
                 'extrap_quad', 'extrap_linear'):
        assert callable(getattr(odil.util_op, name)), name
    a = odil.util.default_args()
    assert (a.linsolver, a.omega, a.smooth_pre, a.smooth_post, a.ndirect,
            a.restriction, a.bfgs_m, a.epoch_start) == (
                'direct', 0.6, 2, 2, 3, 'full', 50, 0)


def test_extrapolations_exact_for_polynomials():
    """util_op.py:1015-1033: quadratic through points 0, 1, 1.5 and 0, 1, 2,
    linear through 0, 1."""
    from odil_amd.problem import extrap_linear, extrap_quad, extrap_quadh
    q = lambda x: 2 - 3 * x + 0.5 * x * x
    assert extrap_quadh(q(0), q(1), q(1.5)) == pytest.approx(q(2))
    assert extrap_quad(q(0), q(1), q(2)) == pytest.approx(q(3))
    assert extrap_linear(1.0, 4.0) == 7.0


@pytest.mark.parametrize('dtype,cell', [(np.float64, True),
                                        (np.float32, False)])
def test_raw_xmf_round_trip(tmp_path, dtype, cell):
    from odil_amd import rawio
    u = np.random.default_rng(0).normal(size=(3, 4, 5)).astype(dtype)
    xmf = str(tmp_path / 'u.xmf')
    assert rawio.write_raw_with_xmf(u, xmf, spacing=(0.1, 0.2, 0.3),
                                    cell=cell, name='u') == xmf
    shape, raw = rawio.parse_raw_xmf(xmf)
    assert shape == (3, 4, 5) and raw == str(tmp_path / 'u.raw')
    back = rawio.read_raw(xmf)
    assert back.dtype == dtype and np.array_equal(back, u)
    text = open(xmf).read()
    nodes = '4 5 6' if cell else '3 4 5'
    assert 'Dimensions="{}"'.format(nodes) in text
    assert 'Center="{}"'.format('Cell' if cell else 'Node') in text
    assert '0.3 0.2 0.1' in text  # spacing reversed to x y z
    # A 2D field is one z layer.
    rawio.write_raw_with_xmf(u[0].astype(np.float64), xmf)
    assert rawio.read_raw(xmf).shape == (1, 4, 5)
