"""odil_amd: MI355X-native hot path of ODIL (Optimizing a DIscrete Loss).

Fused HIP kernels for discrete-loss residuals, stencil Jacobians, loss
adjoints, matrix-free normal operators and a geometric-multigrid Newton solver,
behind the reference's problem/operator API. See `odil_amd.odil` for the
reference-shaped facade.
"""

__version__ = '0.1.0'
