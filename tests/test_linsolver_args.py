"""Solver options keep the reference's names and defaults
(linsolver.py:108-170)."""

import argparse

from odil_amd import linsolver


def test_arguments_match_reference_choices_and_defaults():
    p = linsolver.add_arguments(argparse.ArgumentParser())
    a = p.parse_args([])
    assert a.linsolver == 'direct'
    assert a.linsolver_tol == 1e-6 and a.linsolver_maxiter is None
    assert (a.beta, a.betadiag, a.nlvl, a.omega) == (0., 0., 5, 0.6)
    assert (a.smooth_pre, a.smooth_post, a.ndirect) == (2, 2, 3)
    choices = next(x.choices for x in p._actions if x.dest == 'linsolver')
    assert set(choices) == {'multigrid', 'direct', 'direct_cu', 'sparseqr',
                            'lsqr', 'lsqr_cu', 'bicgstab'}
    d = linsolver.default_args(beta=0.5)
    assert d.beta == 0.5 and d.linsolver_tol == 1e-6 and d.omega == 0.6


def test_multigrid_levels():
    assert linsolver._mg_levels((33, 33), None) == 5   # 33 17 9 5 3
    assert linsolver._mg_levels((33, 33), 3) == 3
    assert linsolver._mg_levels((17, 16), None) == 1   # even: one level
    assert linsolver._mg_levels((3, 3), None) == 1


def test_iterative_solvers_check_the_vector_shape():
    """A right-hand side of the wrong shape is refused before any device
    work (CPU: the check precedes every kernel call)."""
    import pytest
    import torch
    from odil_amd.multigrid import BlockStencil
    coef = torch.ones((1, 5, 5), dtype=torch.float64)
    A = BlockStencil(coef, [(0, 0, (0, 0))], (5, 5))
    bad = torch.zeros((2, 5, 5), dtype=torch.float64)
    for fn in (linsolver.lsqr, linsolver.bicgstab, linsolver.pcg_multigrid):
        with pytest.raises(ValueError):
            fn(A, bad)
