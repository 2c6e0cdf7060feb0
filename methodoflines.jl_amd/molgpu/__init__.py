"""molgpu — MI355X-native method-of-lines right-hand sides and stage loops.

The Python surface mirrors MethodOfLines.jl (``MOLFiniteDifference``, ``discretize``,
``UpwindScheme``, ``WENOScheme``) over a HIP library with a C ABI (include/mol.h).
"""
from .pdesys import (PDESystem, Differential, Interval, Eq, ifelse, parameters, variables,
                     grid_field, GridField)
from .discretization import (MOLFiniteDifference, UpwindScheme, WENOScheme, Lowered, lower,
                             uniform_grid, center_align, edge_align)
from .fornberg import calculate_weights
from .problem import (ODEProblem, ODESolution, RK4, SSPRK33, codegen, discretize,
                      symbolic_discretize, solve)
from . import native

__all__ = [
    "PDESystem", "Differential", "Interval", "Eq", "ifelse", "parameters", "variables",
    "grid_field", "GridField", "MOLFiniteDifference", "UpwindScheme", "WENOScheme", "Lowered",
    "lower", "uniform_grid", "center_align", "edge_align", "calculate_weights", "ODEProblem", "ODESolution", "RK4", "SSPRK33",
    "codegen", "discretize", "symbolic_discretize", "solve", "native",
]
