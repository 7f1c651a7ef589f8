"""The names of the reference's `odil` package (odil.py:1-18), bound to this
package's device implementations, so a driver written against the reference
imports `from odil_amd import odil` instead of `import odil`:

  util_op          Domain, State, Problem, checkpoint_save/load,
                   extrap_quadh/quad/linear (`odil_amd.problem`)
  util             optimize, optimize_newton, optimize_opt,
                   optimize_multigrid, optimize_multigridop, add_arguments
  util_history     History (`odil_amd.checkpoint`)
  linsolver        solve, add_arguments and the solver choices
  optimizer        make_optimizer and the optimiser classes
  multigrid        SparseOperator, MultigridOp, Multigrid (device)
  MultigridDecomp  multigrid field parameterisation (`odil_amd.decomp`)
  parse_raw_xmf, read_raw, write_raw_xmf, write_raw_with_xmf
                   raw + XDMF field files (`odil_amd.rawio`)

Operators are the kernel-backed classes of `odil_amd.operators` (the
reference's Python operator callbacks are not evaluated). `tf` is None:
there is no TensorFlow on this path.
"""

from odil_amd import checkpoint as util_history  # noqa: F401
from odil_amd import linsolver  # noqa: F401
from odil_amd import multigrid  # noqa: F401
from odil_amd import optimizer  # noqa: F401
from odil_amd import problem as util_op
from odil_amd import util  # noqa: F401
from odil_amd.decomp import MultigridDecomp  # noqa: F401
from odil_amd.rawio import (parse_raw_xmf, read_raw,  # noqa: F401
                            write_raw_with_xmf, write_raw_xmf)

Domain = util_op.Domain
State = util_op.State
Problem = util_op.Problem
History = util_history.History
tf = None
