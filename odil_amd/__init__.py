"""odil_amd: MI355X-native hot path of ODIL (Optimizing a DIscrete Loss).

Fused HIP kernels for discrete-loss residuals, stencil Jacobians, loss
adjoints, matrix-free normal operators and a geometric-multigrid Newton solver,
behind the reference's problem/operator API (`odil_amd.odil` holds the
reference-shaped facade: Domain, Problem, MultigridOp, the linear solvers and
optimizers under the names of the reference's `odil` package).
"""

__version__ = '0.1.0'
