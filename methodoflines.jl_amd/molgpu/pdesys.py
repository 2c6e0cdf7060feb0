"""Symbolic PDE-system front-end (the Python counterpart of ModelingToolkit's PDESystem).

A ``PDESystem`` holds equations, boundary/initial conditions, domains, independent and
dependent variables and parameters, written with sympy:

    t, x = parameters("t x")
    u = variables("u")
    Dt, Dxx = Differential(t), Differential(x) ** 2
    eq = Eq(Dt(u(t, x)), D * Dxx(u(t, x)))
    bcs = [Eq(u(0, x), sin(pi * x)), Eq(u(t, 0), 0), Eq(u(t, 1), 0)]
    domains = [Interval(t, 0.0, 0.1), Interval(x, 0.0, 1.0)]
    pdesys = PDESystem([eq], bcs, domains, [t, x], [u(t, x)], {D: 1.0})

Julia's ``lhs ~ rhs`` is ``Eq(lhs, rhs)``; ``x ∈ Interval(a, b)`` is ``Interval(x, a, b)``.
Grid-valued parameters (a learnable per-node field such as a diffusivity) are declared with
``GridField("d0")`` and used as ``d0(x, y, z)``; their values are bound on the problem.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Sequence

import sympy as sp


def parameters(names: str):
    """Declare independent variables / scalar parameters (``@parameters``)."""
    out = sp.symbols(names, real=True)
    return out


def variables(names: str):
    """Declare dependent variables (``@variables u(..)``): returns sympy function heads."""
    heads = [sp.Function(n) for n in names.replace(",", " ").split()]
    return heads[0] if len(heads) == 1 else tuple(heads)


class GridField(sp.Function):
    """Base of grid-valued parameter fields (see ``grid_field``)."""


def grid_field(name: str):
    """A parameter that varies over the grid, e.g. a learnable diffusivity ``d0(x, y, z)``."""
    return type(name, (GridField,), {})


class Differential:
    """``Differential(x)`` as in Symbolics: ``Differential(x)(expr)`` is d(expr)/dx and
    ``Differential(x) ** 2`` the second derivative."""

    def __init__(self, x, order: int = 1):
        self.x = x
        self.order = order

    def __call__(self, expr):
        e = expr
        for _ in range(self.order):
            e = sp.Derivative(e, self.x)
        return e

    def __pow__(self, n: int):
        return Differential(self.x, self.order * n)

    def __mul__(self, other):
        if isinstance(other, Differential):
            return _ComposedDifferential([self, other])
        return NotImplemented


class _ComposedDifferential:
    def __init__(self, ds):
        self.ds = ds

    def __call__(self, expr):
        e = expr
        for d in reversed(self.ds):
            e = d(e)
        return e


def Eq(lhs, rhs):
    """``lhs ~ rhs``."""
    return sp.Eq(sp.sympify(lhs), sp.sympify(rhs), evaluate=False)


def ifelse(cond, a, b):
    """Branch-free conditional, ``ifelse(cond, a, b)``."""
    return sp.Piecewise((sp.sympify(a), cond), (sp.sympify(b), True))


@dataclass
class Interval:
    """Domain of one independent variable: ``x ∈ Interval(lo, hi)``."""
    var: sp.Symbol
    lo: float
    hi: float


@dataclass
class PDESystem:
    """Equations + boundary/initial conditions + domains (ModelingToolkit.PDESystem)."""
    eqs: Sequence
    bcs: Sequence
    domain: Sequence[Interval]
    ivs: Sequence
    dvs: Sequence
    ps: Dict = field(default_factory=dict)
    name: str = "pdesys"

    def __post_init__(self):
        if not isinstance(self.eqs, (list, tuple)):
            self.eqs = [self.eqs]
        self.eqs = list(self.eqs)
        self.bcs = list(self.bcs)
        self.ivs = list(self.ivs)
        self.dvs = list(self.dvs)
        self.ps = dict(self.ps or {})
