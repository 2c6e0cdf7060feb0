"""``MOLFiniteDifference`` and the lowering of a PDESystem to the stencil IR.

This mirrors MethodOfLines.jl's discretization pipeline (PDEBase symbolic_discretize ->
MOL hooks), but instead of emitting one symbolic equation per grid node it emits a
compact descriptor that the GPU kernels interpret:

* grids per axis ............ discretize_space / prepare_dx (src/discretization/discretize_vars.jl:219-233,283-285)
* unknown box per variable .. construct_var_equation_mapping / clip_interior!! / calculate_stencil_extents
                              (src/system_parsing/interior_map.jl:1-10,53-103,117-139)
* operator tables ........... construct_differential_discretizer (src/discretization/differential_discretizer.jl:14-109)
* derivative rules .......... generate_finite_difference_rules precedence
                              (src/discretization/generate_finite_difference_rules.jl:47-116)
* boundary nodes ............ generate_bc_eqs / boundary_value_maps / generate_extrap_eqs! / generate_corner_eqs!
                              (src/discretization/generate_bc_eqs.jl:24-58,238-311,336-416), solved for the
                              boundary unknowns the way mtkcompile eliminates them (affine closures)
* initial values ............ generate_ic_defaults (src/discretization/generate_ic_defaults.jl:1-27)
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import sympy as sp

from . import fornberg
from .bytecode import Compiler, ConstPool, CompileError
from .pdesys import GridField, PDESystem

SLOT_CENTERED, SLOT_UPWIND, SLOT_WENO, SLOT_NONLINLAP, SLOT_SPHERICAL, SLOT_MIXED, SLOT_WENO_NU = range(7)


class UpwindScheme:
    """First-order-derivative (advection) scheme: upwind differences (scheme_types.jl:8-13)."""

    def __init__(self, order: int = 1):
        self.order = order


class WENOScheme:
    """Jiang-Shu WENO5 advection scheme (schemes/WENO/WENO.jl:83-89); uniform grids."""

    def __init__(self, epsilon: float = 1e-6):
        self.epsilon = float(epsilon)


# grid alignments (src/interface/grid_types.jl:1-25): nodes on the domain ends, or cell centres
# with one ghost node past each end (the boundary lies half-way between the ghost and the first
# cell centre)
center_align = "center"
edge_align = "edge"


class MOLFiniteDifference:
    """Discretization settings (src/interface/MOLFiniteDifference.jl:31-76).

    ``dxs`` maps each spatial variable to a step (float) or a node count (int; with
    ``grid_align=edge_align`` a cell count, prepare_dx, discretize_vars.jl:283-285).
    ``dtype`` ("f64" / "f32") is the state and arithmetic type of the GPU path.
    """

    def __init__(self, dxs, time=None, approx_order: int = 2, advection_scheme=None,
                 grid_align: str = center_align, dtype: str = "f64", **kwargs):
        if approx_order % 2 != 0:
            warnings.warn("Discretization approx_order must be even, rounding up to %d" % (approx_order + 1))
            approx_order += 1
        if approx_order < 1:
            raise ValueError("approx_order must be at least 1")
        if grid_align not in (center_align, edge_align):
            raise NotImplementedError("grid_align must be center_align or edge_align (staggered grids are not supported)")
        if dtype not in ("f64", "f32"):
            raise ValueError("dtype must be 'f64' or 'f32'")
        self.dxs = dict(dxs)
        self.time = time
        self.approx_order = approx_order
        self.advection_scheme = advection_scheme if advection_scheme is not None else UpwindScheme()
        self.grid_align = grid_align
        self.dtype = dtype
        self.kwargs = kwargs


# --------------------------------------------------------------------------------------
# grids

def uniform_grid(lo: float, hi: float, dx) -> np.ndarray:
    """Nodes lo:dx:hi as Julia's float range builds them: node k is the correctly rounded
    value of lo + k*step where step is the rational the float step stands for
    (discretize_vars.jl:219-233).  An integer ``dx`` is a node count (prepare_dx, :283)."""
    lo, hi = float(lo), float(hi)
    if isinstance(dx, (int, np.integer)) and not isinstance(dx, bool):
        n = int(dx)
        if n < 2:
            raise ValueError("need at least 2 nodes")
        step_f = (hi - lo) / (n - 1)
        fr_lo, fr_hi = Fraction(lo), Fraction(hi)
        nodes = [float(fr_lo + (fr_hi - fr_lo) * k / (n - 1)) for k in range(n)]
        return np.array(nodes, dtype=np.float64)
    step_f = float(dx)
    if step_f <= 0:
        raise ValueError("dx must be positive")
    step = Fraction(step_f).limit_denominator(10 ** 9)
    if float(step) != step_f:
        step = Fraction(step_f)
    start = Fraction(lo).limit_denominator(10 ** 9)
    if float(start) != lo:
        start = Fraction(lo)
    stop = Fraction(hi).limit_denominator(10 ** 9)
    if float(stop) != hi:
        stop = Fraction(hi)
    m = int((stop - start) / step)
    nodes = [float(start + k * step) for k in range(m + 1)]
    # guard against an overshoot from the rational reconstruction
    while nodes and nodes[-1] > hi:
        nodes.pop()
    if nodes[-1] != hi:
        warnings.warn("d%s does not divide domain exactly, adding grid point at %s" % ("x", hi))
        nodes.append(hi)
    return np.array(nodes, dtype=np.float64)


def edge_step(lo: float, hi: float, dx) -> float:
    """Step of an edge-aligned axis: a float dx as given, an integer one a cell count
    (prepare_dx(dx::Integer, xdomain, ::EdgeAlignedGrid), discretize_vars.jl:284)."""
    if isinstance(dx, (int, np.integer)) and not isinstance(dx, bool):
        if dx < 1:
            raise ValueError("need at least one cell")
        return (float(hi) - float(lo)) / int(dx)
    return float(dx)


def edge_grid(lo: float, hi: float, h: float) -> np.ndarray:
    """Nodes (lo - h/2):h:(hi + h/2) of an edge-aligned axis as Julia's float range builds
    them (discretize_vars.jl:380-382): the ends computed in FP64, as many steps as fit (with a
    relative tolerance of 1e-10 on the quotient), node k the correctly rounded start + k*h.
    Unlike the centre-aligned grid no node is appended at the end."""
    start, stop = float(lo) - h / 2, float(hi) + h / 2
    q = (stop - start) / h
    m = int(np.floor(q + 1e-10 * max(1.0, q)))
    if m < 2:
        raise ValueError("an edge-aligned axis needs at least one cell")
    fs, fh = Fraction(start), Fraction(h)
    return np.array([float(fs + k * fh) for k in range(m + 1)], dtype=np.float64)


def half_row(op: "fornberg.DerivativeOperator", i: int, n: int):
    """(weights, first tap) of a half-point operator at index i (the half point between nodes i
    and i+1) on a non-periodic axis of n nodes (get_half_offset_weights_and_stencil,
    half_offset_centred_difference.jl:9-36)."""
    bpc, BL, L = op.boundary_point_count, op.boundary_stencil_length, op.stencil_length
    if i <= bpc:
        return op.low_boundary_coefs[i - 1], 1
    if i > n - bpc:
        return op.high_boundary_coefs[n - i - 1], n - BL + 1
    return op.stencil_coefs, i + 1 - L // 2


# --------------------------------------------------------------------------------------
# lowered problem

@dataclass
class Table:
    op: fornberg.DerivativeOperator
    ref_shift: int = 0


@dataclass
class Closure:
    field: int
    axis: int
    side: int
    C: np.ndarray          # nghost x K
    c: np.ndarray          # nghost
    gprog: int             # program id or -1
    src: int = -1          # interface: the variable read past this side (C, c empty)


@dataclass
class Lowered:
    """The stencil IR of one PDE system (what ``mol_desc`` carries)."""
    ndim: int
    axes: List[sp.Symbol]
    n: List[int]
    periodic: List[bool]
    coords: List[np.ndarray]
    fields: List[str]
    field_heads: List
    lo: List[List[int]]
    hi: List[List[int]]
    halo: List[int]
    tables: List[Table]
    slots: List[dict]
    closures: List[Closure]
    programs: List[List[int]]
    cpool: List[float]
    rhs_prog: List[int]
    params: List[sp.Symbol]
    param_values: List[float]
    aux: List[str]
    aux_heads: List
    dtype: str
    u0: np.ndarray
    time: Optional[sp.Symbol]
    tspan: Tuple[float, float]
    rhs_exprs: List = field(default_factory=list)   # the lowered expressions (for inspection)
    # variables on adjacent ranges of one axis joined by interface conditions: axis, and per
    # variable its range of combined-grid nodes (1-based, inclusive) and its own spatial symbol
    interface: Optional[dict] = None

    @property
    def nfields(self) -> int:
        return len(self.fields)

    def box_shape(self, f: int) -> Tuple[int, ...]:
        return tuple(self.hi[f][a] - self.lo[f][a] + 1 for a in range(self.ndim))

    def field_slices(self) -> List[slice]:
        out, off = [], 0
        for f in range(self.nfields):
            sz = int(np.prod(self.box_shape(f)))
            out.append(slice(off, off + sz))
            off += sz
        return out

    @property
    def n_unknowns(self) -> int:
        return sum(int(np.prod(self.box_shape(f))) for f in range(self.nfields))

    def grid_shape(self) -> Tuple[int, ...]:
        return tuple(self.n)


# --------------------------------------------------------------------------------------
# helpers over sympy trees

def _applied_heads(expr, heads) -> List:
    return [a for a in expr.atoms(sp.Function) if a.func in heads]


def _derivs(expr) -> List[sp.Derivative]:
    return list(expr.atoms(sp.Derivative))


def _deriv_spec(d: sp.Derivative):
    """Return (inner expression, [(var, count), ...]) of a derivative."""
    return d.expr, [(v, int(c)) for v, c in d.variable_count]


class _Lowering:
    def __init__(self, pdesys: PDESystem, disc: MOLFiniteDifference):
        self.sys = pdesys
        self.disc = disc
        self.t = disc.time
        self.p = disc.approx_order
        self.scheme = disc.advection_scheme
        self.cpool = ConstPool()
        self.programs: List[List[int]] = []
        self.tables: List[Table] = []
        self.table_index: Dict[tuple, int] = {}
        self.slots: List[dict] = []
        self.slot_index: Dict[tuple, int] = {}
        self.rhs_exprs = []

    # ---------------------------------------------------------------- variables / grid
    def setup(self):
        sys, disc = self.sys, self.disc
        if self.t is None:
            raise NotImplementedError("time-independent systems (NonlinearProblem) are not supported")
        self.heads = []
        fargs = []
        for dv in sys.dvs:
            if not isinstance(dv, sp.Function) or not isinstance(dv, sp.core.function.AppliedUndef):
                raise ValueError("dependent variables must be applied functions such as u(t, x)")
            self.heads.append(dv.func)
            if self.t not in dv.args:
                raise ValueError("dependent variable %s does not depend on time" % dv)
            fargs.append([a for a in dv.args if a != self.t])
        dom = {iv.var: iv for iv in sys.domain}
        self.domain = dom
        self.edge = disc.grid_align == edge_align
        self.iface = None
        self.sym_map = {}
        if all(a == fargs[0] for a in fargs):
            self.field_args = fargs[0]
        else:
            self._setup_interface(fargs)
        self.axes = list(self.field_args)
        self.ndim = len(self.axes)
        if self.ndim < 1 or self.ndim > 3:
            raise NotImplementedError("1 to 3 spatial dimensions are supported")
        self.coords, self.n = [], []
        for ax, x in enumerate(self.axes):
            if x not in disc.dxs:
                raise ValueError("Variable %s has no step size" % x)
            dx = disc.dxs[x]
            if self.iface is not None and ax == self.iface["axis"]:
                g = self.iface["grid"]
            elif isinstance(dx, (list, tuple, np.ndarray)):
                # nonuniform grid: the nodes themselves (generate_grid, discretize_vars.jl:219-233)
                if disc.grid_align == edge_align:
                    raise NotImplementedError("edge-aligned grids are supported on uniform axes only")
                g = np.asarray(dx, dtype=np.float64).copy()
                if g.ndim != 1 or len(g) < 5 or not np.all(np.diff(g) > 0):
                    raise ValueError("a nonuniform grid must be at least 5 strictly increasing nodes")
            elif disc.grid_align == edge_align:
                # cell centres plus a ghost past each end: (lo - h/2):h:(hi + h/2)
                # (generate_grid, discretize_vars.jl:373-390; prepare_dx, :283-285)
                g = edge_grid(dom[x].lo, dom[x].hi, edge_step(dom[x].lo, dom[x].hi, dx))
            else:
                g = uniform_grid(dom[x].lo, dom[x].hi, dx)
            self.coords.append(g)
            self.n.append(len(g))
        self.dx = []
        self.nonuniform = []
        for x, g in zip(self.axes, self.coords):
            dxv = self.disc.dxs[x]
            nu = isinstance(dxv, (list, tuple, np.ndarray))
            self.nonuniform.append(nu)
            if nu:
                dxv = float("nan")
            elif self.edge:
                dxv = edge_step(dom[x].lo, dom[x].hi, dxv)
            elif isinstance(dxv, (int, np.integer)) and not isinstance(dxv, bool):
                dxv = (dom[x].hi - dom[x].lo) / (dxv - 1)
            self.dx.append(float(dxv))
        self.tspan = (float(dom[self.t].lo), float(dom[self.t].hi)) if self.t in dom else (0.0, 1.0)
        # parameters: scalar values and grid fields
        self.params, self.param_values = [], []
        self.aux_heads, self.aux_values = [], []
        for k, v in sys.ps.items():
            if isinstance(k, type) and issubclass(k, GridField):
                self.aux_heads.append(k)
                self.aux_values.append(v)
            else:
                self.params.append(k)
                self.param_values.append(float(v))
        # grid fields used but not given values
        for eq in sys.eqs:
            for a in sorted(eq.atoms(sp.Function), key=str):
                if isinstance(a, GridField) and a.func not in self.aux_heads:
                    self.aux_heads.append(a.func)
                    self.aux_values.append(None)
        # a fixed order (by name): the atoms of an expression come as an unordered set
        order = sorted(range(len(self.aux_heads)), key=lambda i: self.aux_heads[i].__name__)
        self.aux_heads = [self.aux_heads[i] for i in order]
        self.aux_values = [self.aux_values[i] for i in order]

    def _setup_interface(self, fargs):
        """Variables over different spatial variables: supported when they differ along one axis
        only, each variable owns one interval along it, the intervals meet end to end and carry
        grids of one spacing.  The intervals become one combined axis (the first interval's
        symbol over their union); a variable's range is its own nodes, the shared end node of
        two neighbours counted in both (interface_boundary.jl:79-111, generate_bc_eqs.jl:35-58)."""
        nd = len(fargs[0])
        if any(len(a) != nd for a in fargs):
            raise NotImplementedError("all dependent variables must have the same number of spatial variables")
        diff = [k for k in range(nd) if len({a[k] for a in fargs}) > 1]
        if len(diff) != 1 or any(a[k] != fargs[0][k] for a in fargs for k in range(nd) if k != diff[0]):
            raise NotImplementedError("variables over different spatial variables must differ along one axis "
                                      "(an interface between adjacent intervals)")
        ia = diff[0]
        dom = self.domain
        syms = sorted({a[ia] for a in fargs}, key=lambda x: float(dom[x].lo))
        owners = {x: [f for f, a in enumerate(fargs) if a[ia] == x] for x in syms}
        if any(len(v) != 1 for v in owners.values()):
            raise NotImplementedError("one variable per interval along an interface axis")
        for x0, x1 in zip(syms, syms[1:]):
            if abs(float(dom[x0].hi) - float(dom[x1].lo)) > 1e-12 * max(1.0, abs(float(dom[x0].hi))):
                raise NotImplementedError("intervals of %s and %s do not meet" % (x0, x1))
        grids, dxs = [], []
        for x in syms:
            dx = self.disc.dxs.get(x)
            if dx is None:
                raise ValueError("Variable %s has no step size" % x)
            if isinstance(dx, (list, tuple, np.ndarray)):
                raise NotImplementedError("interfaces between nonuniform grids are not supported (%s)" % x)
            g = uniform_grid(dom[x].lo, dom[x].hi, dx)
            grids.append(g)
            dxs.append((float(dom[x].hi) - float(dom[x].lo)) / (len(g) - 1))
        if max(dxs) - min(dxs) > 1e-12 * max(dxs):
            # the reference's interior stencils wrap by index across an interface
            # (interface_boundary.jl:79-111): only one spacing keeps them consistent
            raise NotImplementedError("interfaces between grids of different spacing are not supported")
        combined = np.concatenate([grids[0]] + [g[1:] for g in grids[1:]])
        start, end, k = [], [], 1
        for g in grids:
            start.append(k)
            end.append(k + len(g) - 1)
            k += len(g) - 1
        order = [owners[x][0] for x in syms]
        X = syms[0]
        self.sym_map = {x: X for x in syms[1:]}
        self.field_args = [X if k == ia else fargs[order[0]][k] for k in range(nd)]

        class _Iv:
            pass
        iv = _Iv()
        iv.var, iv.lo, iv.hi = X, dom[syms[0]].lo, dom[syms[-1]].hi
        self.domain = dict(dom)
        self.domain[X] = iv
        nf = len(fargs)
        rng = [None] * nf
        lohi = [None] * nf
        for j, f in enumerate(order):
            rng[f] = (start[j], end[j])
            lohi[f] = (float(dom[syms[j]].lo), float(dom[syms[j]].hi))
        self.iface = dict(axis=ia, grid=combined, range=rng, coord_range=lohi,
                          symbols=[fargs[f][ia] for f in range(nf)], dx=dxs[0])

    def _interface_bc(self, bc, lhs, rhs, spos, axis):
        """u1(t, .., m, ..) ~ u2(t, .., m, ..) at the point m where the interval of one variable
        ends and the other's begins.  The lower variable keeps m as an unknown, the upper one
        takes it as the alias u2[first] = u1[last] (clip_interior!!, interior_map.jl:1-10;
        generate_bc_eqs.jl:35-58)."""
        if self.iface is None or axis != self.iface["axis"]:
            raise NotImplementedError("interface condition away from an interface axis: %s" % bc)
        fa, fb = self.heads.index(lhs.func), self.heads.index(rhs.func)
        if any(lhs.args[i] != rhs.args[i] for i in spos if i != spos[axis]):
            raise NotImplementedError("unsupported interface condition %s" % bc)
        m = float(lhs.args[spos[axis]])
        if abs(float(rhs.args[spos[axis]]) - m) > 1e-12 * max(1.0, abs(m)):
            raise NotImplementedError("interface condition at two different points: %s" % bc)
        cr = self.iface["coord_range"]
        tol = 1e-12 * max(1.0, abs(m))
        if abs(cr[fa][1] - m) <= tol and abs(cr[fb][0] - m) <= tol:
            lower, upper = fa, fb
        elif abs(cr[fb][1] - m) <= tol and abs(cr[fa][0] - m) <= tol:
            lower, upper = fb, fa
        else:
            raise NotImplementedError("interface condition away from the meeting point of %s and %s: %s"
                                      % (self.heads[fa], self.heads[fb], bc))
        self.bc[lower][axis][1] = ("iface", upper)
        self.bc[upper][axis][0] = ("iface", lower)

    def _is_iface(self, f, a, side):
        b = self.bc[f][a][side]
        return isinstance(b, tuple) and len(b) == 2 and b[0] == "iface"

    def _range(self, f, a):
        """Nodes (1-based, inclusive) of variable f along axis a."""
        if self.iface is not None and a == self.iface["axis"]:
            return self.iface["range"][f]
        return 1, self.n[a]

    # ---------------------------------------------------------------- boundary conditions
    def parse_bcs(self):
        """Classify each condition: initial value, Dirichlet/Neumann/Robin (affine), periodic."""
        nf = len(self.heads)
        self.ic = [None] * nf
        self.bc = [[[None, None] for _ in range(self.ndim)] for _ in range(nf)]
        self.periodic = [False] * self.ndim
        per_vars = [[False] * self.ndim for _ in range(nf)]
        t0 = self.tspan[0]
        for bc in self.sys.bcs:
            lhs, rhs = bc.lhs, bc.rhs
            if self.sym_map:
                lhs, rhs = sp.sympify(lhs).xreplace(self.sym_map), sp.sympify(rhs).xreplace(self.sym_map)
            apps = _applied_heads(sp.Add(lhs, rhs, evaluate=False), self.heads)
            if not apps:
                raise ValueError("condition %s references no dependent variable" % bc)
            a0 = apps[0]
            dv = next(d for d in self.sys.dvs if d.func == a0.func)
            tpos = list(dv.args).index(self.t)
            spos = [i for i in range(len(dv.args)) if i != tpos]
            # initial condition: u(t0, x...) ~ expr(x...)
            if lhs in apps and len(apps) == 1 and lhs.args[tpos] != self.t and not (
                    isinstance(rhs, sp.core.function.AppliedUndef) and rhs.func in self.heads):
                tv = lhs.args[tpos]
                if not tv.is_number or abs(float(tv) - t0) > 1e-12:
                    raise ValueError("initial condition not at the domain start: %s" % bc)
                if [lhs.args[i] for i in spos] != self.axes:
                    raise ValueError("initial condition must be given over all spatial variables: %s" % bc)
                self.ic[self.heads.index(lhs.func)] = rhs
                continue
            # boundary condition: exactly one spatial argument held at a domain end
            f = self.heads.index(a0.func)
            fixed = [(k, a0.args[i]) for k, i in enumerate(spos) if a0.args[i].is_number]
            if len(fixed) != 1:
                raise NotImplementedError("unsupported boundary condition %s" % bc)
            axis, val = fixed[0]
            val = float(val)
            x = self.axes[axis]
            lo, hi = self.domain[x].lo, self.domain[x].hi
            if self.iface is not None and axis == self.iface["axis"]:
                lo, hi = self.iface["coord_range"][f]
            # periodic: u(.., lo, ..) ~ u(.., hi, ..)  (an interface of a variable with itself)
            if (isinstance(lhs, sp.core.function.AppliedUndef) and isinstance(rhs, sp.core.function.AppliedUndef)
                    and lhs.func in self.heads and rhs.func in self.heads):
                if lhs.func != rhs.func:
                    self._interface_bc(bc, lhs, rhs, spos, axis)
                    continue
                av = sorted([float(lhs.args[spos[axis]]), float(rhs.args[spos[axis]])])
                if av != sorted([float(lo), float(hi)]):
                    raise NotImplementedError("unsupported interface condition %s" % bc)
                self.periodic[axis] = True
                per_vars[f][axis] = True
                continue
            side = 0 if abs(val - lo) <= 1e-12 * max(1.0, abs(lo)) else (1 if abs(val - hi) <= 1e-12 * max(1.0, abs(hi)) else -1)
            if side < 0:
                raise NotImplementedError("boundary condition away from the domain ends: %s" % bc)
            # affine form a0*u_b + sum_d a_d D^d u_b = g
            ub = a0
            R = sp.expand(lhs - rhs)
            subs, Ukeys = {}, {}
            for dd in R.atoms(sp.Derivative):
                inner, vc = _deriv_spec(dd)
                if inner != ub or len(vc) != 1 or vc[0][0] != x:
                    raise NotImplementedError("unsupported derivative in boundary condition %s" % bc)
                key = vc[0][1]
                sym = sp.Symbol("__U%d" % key)
                subs[dd] = sym
                Ukeys[key] = sym
            R2 = R.xreplace(subs)
            U0 = sp.Symbol("__U0")
            R2 = R2.xreplace({ub: U0})
            Ukeys[0] = U0
            if R2.has(*self.heads):
                raise NotImplementedError("boundary condition couples other variables: %s" % bc)
            coefs = {}
            for k, s in Ukeys.items():
                c = sp.diff(R2, s)
                if c.free_symbols & set(Ukeys.values()):
                    raise NotImplementedError("nonlinear boundary condition: %s" % bc)
                c = c.xreplace(dict(zip(self.params, self.param_values)))
                if not c.is_number:
                    raise NotImplementedError("boundary condition coefficients must be constants: %s" % bc)
                coefs[k] = float(c)
            g = sp.sympify(-R2.xreplace({s: 0 for s in Ukeys.values()}))
            self.bc[f][axis][side] = (coefs, g)
        if self.iface is not None:
            ax = self.iface["axis"]
            for f in range(nf):
                s0, s1 = self.iface["range"][f]
                if (s0 > 1 and self.bc[f][ax][0] is None) or (s1 < self.n[ax] and self.bc[f][ax][1] is None):
                    raise NotImplementedError("%s meets another variable without an interface condition" % self.heads[f])
        for f in range(nf):
            if self.ic[f] is None:
                raise ValueError("no initial condition for %s" % self.heads[f])
            for a in range(self.ndim):
                if self.periodic[a] and not per_vars[f][a]:
                    raise NotImplementedError("all variables must share periodic axes")

    # ---------------------------------------------------------------- tables and slots
    def table(self, key) -> int:
        if key in self.table_index:
            return self.table_index[key]
        kind, axis = key[0], key[1]
        if self.nonuniform[axis]:
            ref_shift = key[4] if key[0] == "half" and len(key) > 4 else 0
            self.tables.append(Table(self._nonuniform_table(key), ref_shift))
            self.table_index[key] = len(self.tables) - 1
            return self.table_index[key]
        dx = self.dx[axis]
        ref_shift = 0
        if kind == "centered":
            op = fornberg.centered_difference(key[2], self.p, dx)
        elif kind == "upwind":
            d, order, offside = key[2], key[3], key[4]
            op = fornberg.upwind_difference(d, order, dx, offside)
        elif kind == "half":
            op = fornberg.half_centered_difference(key[2], key[3], dx)
            ref_shift = key[4] if len(key) > 4 else 0
        else:
            raise ValueError(kind)
        self.tables.append(Table(op, ref_shift))
        self.table_index[key] = len(self.tables) - 1
        return self.table_index[key]

    def _nonuniform_table(self, key):
        """Per-node operator rows on a nonuniform axis: centered derivatives
        (centered_diff_weights.jl:77-153; no interface/periodic boundaries, centered_difference.jl:36),
        first-order upwind derivatives (upwind_diff_weights.jl:93-171) and the half-point
        operators of the nonlinear Laplacian (half_offset_weights.jl:72-150; the outer one on the
        midpoints of the nodes, differential_discretizer.jl:44-51)."""
        kind, axis = key[0], key[1]
        name = self.axes[axis]
        if self.periodic[axis]:
            raise NotImplementedError("operators on a periodic nonuniform axis (%s)" % name)
        x = self.coords[axis]
        if kind == "centered":
            return fornberg.centered_difference_nonuniform(key[2], self.p, x)
        if kind == "upwind":
            d, order, offside = key[2], key[3], key[4]
            if d + order > 2:
                raise NotImplementedError("upwind stencils wider than two nodes on a nonuniform axis (%s): the "
                                          "reference's backward rows end before the last nodes" % name)
            return fornberg.upwind_difference_nonuniform(d, order, x, offside)
        if kind == "half":
            d, p = key[2], key[3]
            outer = len(key) > 4 and key[4] == -1
            return fornberg.half_centered_difference_nonuniform(d, p, (x[:-1] + x[1:]) / 2 if outer else x)
        raise NotImplementedError("%s operators on a nonuniform axis (%s)" % (kind, name))

    def slot(self, desc: dict) -> int:
        key = tuple(sorted((k, v if not isinstance(v, list) else tuple(v)) for k, v in desc.items()))
        if key in self.slot_index:
            return self.slot_index[key]
        self.slots.append(desc)
        self.slot_index[key] = len(self.slots) - 1
        return self.slot_index[key]

    # ---------------------------------------------------------------- programs
    def _leaf_node(self, extra: Dict):
        def leaf(e):
            if e in extra:
                return extra[e]
            if e == self.t:
                return ("TIME", 0)
            if e in self.axes:
                return ("COORD", self.axes.index(e))
            if e in self.params:
                return ("PARAM", self.params.index(e))
            raise CompileError("unresolved symbol %s" % e)
        return leaf

    def add_program(self, expr, extra: Dict) -> int:
        comp = Compiler(self._leaf_node(extra), self.cpool)
        code = comp.compile(expr)
        self.programs.append(code)
        return len(self.programs) - 1

    def _field_subs(self, expr):
        """Replace u_f(t, x..) by field placeholders and grid fields by aux placeholders."""
        extra, rep = {}, {}
        for f, hd in enumerate(self.heads):
            s = sp.Symbol("__F%d" % f)
            extra[s] = ("FIELD", f)
        for a, hd in enumerate(self.aux_heads):
            s = sp.Symbol("__A%d" % a)
            extra[s] = ("AUX", a)
        for app in expr.atoms(sp.Function):
            if app.func in self.heads:
                f = self.heads.index(app.func)
                sargs = [a for a in app.args if a != self.t]
                if sargs != self.axes or self.t not in app.args:
                    raise NotImplementedError("boundary values inside a PDE are not supported: %s" % app)
                rep[app] = sp.Symbol("__F%d" % f)
            elif isinstance(app, GridField):
                a = self.aux_heads.index(app.func)
                if list(app.args) != self.axes:
                    raise NotImplementedError("grid fields must be evaluated at the node: %s" % app)
                rep[app] = sp.Symbol("__A%d" % a)
        return expr.xreplace(rep), extra

    # ---------------------------------------------------------------- derivative rules
    def _var_of(self, e):
        if isinstance(e, sp.Function) and e.func in self.heads:
            return self.heads.index(e.func)
        return None

    def _match_nonlinlap(self, d: sp.Derivative):
        """d = Dx(a * Dx(u)) -> (field, axis, a) else None (nonlinear_laplacian.jl:149-290)."""
        inner, vc = _deriv_spec(d)
        if len(vc) != 1 or vc[0][1] != 1:
            return None
        x = vc[0][0]
        if x not in self.axes:
            return None
        factors = sp.Mul.make_args(inner)
        dus = [f for f in factors if isinstance(f, sp.Derivative)]
        if len(dus) != 1:
            return None
        din, dvc = _deriv_spec(dus[0])
        fvar = self._var_of(din)
        if fvar is None or len(dvc) != 1 or dvc[0] != (x, 1):
            return None
        a = sp.Mul(*[f for f in factors if f is not dus[0]])
        if a.has(sp.Derivative):
            # derivatives of u along x inside the coefficient use the inner half-point stencil
            for dd in a.atoms(sp.Derivative):
                i2, v2 = _deriv_spec(dd)
                if not (self._var_of(i2) == fvar and v2 == [(x, 1)]):
                    return None
        return fvar, self.axes.index(x), a

    def _coef_program(self, a, fvar: int, half: bool) -> int:
        a2, extra = self._field_subs(a)
        if half:
            din = sp.Symbol("__DIN")
            extra[din] = ("DINNER", 0)
            reps = {}
            for dd in a2.atoms(sp.Derivative):
                reps[dd] = din
            a2 = a2.xreplace(reps)
        elif a2.has(sp.Derivative):
            raise NotImplementedError("derivatives inside a coefficient are not supported: %s" % a)
        return self.add_program(a2, extra)

    def _nonlinlap_slot(self, fvar: int, axis: int, a) -> dict:
        p = self.p
        return dict(kind=SLOT_NONLINLAP, field=fvar, axis=axis, axis2=0,
                    tab=[self.table(("half", axis, 1, p, -1)), self.table(("half", axis, 1, p)),
                         self.table(("half", axis, 0, max(4, p)))],
                    prog=self._coef_program(a, fvar, True), eps=0.0, inv_dx=0.0)

    def _match_spherical(self, term):
        """term = c * r^-2 * Dr(r^2 * a * Dr(u)) -> (c, field, axis, a)  (spherical_laplacian.jl:44-140)."""
        factors = list(sp.Mul.make_args(term))
        for i, f in enumerate(factors):
            if not isinstance(f, sp.Derivative):
                continue
            inner, vc = _deriv_spec(f)
            if len(vc) != 1 or vc[0][1] != 1 or vc[0][0] not in self.axes:
                continue
            r = vc[0][0]
            others = factors[:i] + factors[i + 1:]
            rm2 = [o for o in others if o == r ** -2]
            if not rm2:
                continue
            ifs = list(sp.Mul.make_args(inner))
            if r ** 2 not in ifs:
                continue
            dus = [g for g in ifs if isinstance(g, sp.Derivative)]
            if len(dus) != 1:
                continue
            din, dvc = _deriv_spec(dus[0])
            fvar = self._var_of(din)
            if fvar is None or dvc != [(r, 1)]:
                continue
            rest_in = list(ifs)
            rest_in.remove(r ** 2)
            rest_in.remove(dus[0])
            a = sp.Mul(*rest_in)
            if a.has(sp.Derivative):
                continue
            rest_out = list(others)
            rest_out.remove(r ** -2)
            return sp.Mul(*rest_out), fvar, self.axes.index(r), a
        return None

    def _replace_atoms(self, e):
        """Replace every remaining derivative of a dependent variable by an operator slot
        (centered even orders, mixed, advection for odd orders)."""
        reps = {}
        for d in sorted(_derivs(e), key=lambda z: -sp.count_ops(z)):
            if d in reps:
                continue
            nl = self._match_nonlinlap(d)
            if nl is not None:
                k = self.slot(self._nonlinlap_slot(*nl))
                reps[d] = sp.Symbol("__S%d" % k)
                continue
            inner, vc = _deriv_spec(d)
            fvar = self._var_of(inner)
            if fvar is None:
                raise NotImplementedError("derivative of a non-variable expression: %s" % d)
            if any(v == self.t for v, _ in vc):
                raise NotImplementedError("time derivative inside the right-hand side: %s" % d)
            if len(vc) == 2 and vc[0][1] == 1 and vc[1][1] == 1:
                # on a nonuniform axis the first-derivative table has a row per node, as
                # mixed_central_difference takes it (2nd_order_mixed_deriv.jl:5-22 through
                # central_difference_weights_and_stencil on nodes, centered_difference.jl:32-57)
                ax, ay = self.axes.index(vc[0][0]), self.axes.index(vc[1][0])
                k = self.slot(dict(kind=SLOT_MIXED, field=fvar, axis=ax, axis2=ay,
                                   tab=[self.table(("centered", ax, 1)), self.table(("centered", ay, 1))],
                                   prog=-1, eps=0.0, inv_dx=0.0))
                reps[d] = sp.Symbol("__S%d" % k)
                continue
            if len(vc) != 1:
                raise NotImplementedError("unsupported mixed derivative %s" % d)
            x, order = vc[0]
            ax = self.axes.index(x)
            if order % 2 == 0:
                k = self.slot(dict(kind=SLOT_CENTERED, field=fvar, axis=ax, axis2=0,
                                   tab=[self.table(("centered", ax, order))], prog=-1, eps=0.0, inv_dx=0.0))
            elif order == 1 and isinstance(self.scheme, WENOScheme) and self.nonuniform[ax]:
                # node-centred nonuniform WENO5 (weno_f_nonuniform, schemes/WENO/nonuniform_weno.jl)
                k = self.slot(dict(kind=SLOT_WENO_NU, field=fvar, axis=ax, axis2=0, tab=[0] * 0, prog=-1,
                                   eps=self.scheme.epsilon, inv_dx=0.0))
            elif order == 1 and isinstance(self.scheme, WENOScheme):
                k = self.slot(dict(kind=SLOT_WENO, field=fvar, axis=ax, axis2=0, tab=[0] * 0, prog=-1,
                                   eps=self.scheme.epsilon, inv_dx=1.0 / self.dx[ax]))
            else:
                k = self.slot(self._upwind_slot(fvar, ax, order, None))
            reps[d] = sp.Symbol("__S%d" % k)
        return e.xreplace(reps)

    def _upwind_slot(self, fvar, ax, order, coef):
        uo = self.scheme.order if isinstance(self.scheme, UpwindScheme) else 1
        L = order + uo
        fwd = self.table(("upwind", ax, order, uo, 0))
        bwd = self.table(("upwind", ax, order, uo, L - 1))
        prog = -1
        if coef is not None:
            prog = self._coef_program(coef, fvar, False)
        return dict(kind=SLOT_UPWIND, field=fvar, axis=ax, axis2=0, tab=[fwd, bwd], prog=prog,
                    eps=0.0, inv_dx=0.0)

    def _match_winding(self, term):
        """term = a * D^odd_x u * b  ->  coefficient-signed upwind slot (upwind_difference.jl:201-276)."""
        skip1 = isinstance(self.scheme, WENOScheme)
        num, den = [], []
        for f in sp.Mul.make_args(term):
            if f.is_Pow and f.exp.is_number and float(f.exp) < 0 and not f.base.has(sp.Derivative):
                den.append(f)
            else:
                num.append(f)
        cands = [f for f in num if isinstance(f, sp.Derivative)]
        for dd in cands:
            inner, vc = _deriv_spec(dd)
            fvar = self._var_of(inner)
            if fvar is None or len(vc) != 1:
                continue
            x, order = vc[0]
            if order % 2 == 0 or (skip1 and order == 1):
                continue
            rest = [f for f in num if f is not dd] + den
            coef = sp.Mul(*rest)
            if coef.has(sp.Derivative):
                continue
            ax = self.axes.index(x)
            return self._upwind_slot(fvar, ax, order, coef)
        return None

    def lower_equation(self, eq) -> Tuple[int, sp.Expr]:
        """Cardinalize lhs - rhs ~ 0, find the time derivative, apply the rules term by term."""
        R = eq.lhs - eq.rhs
        terms = list(sp.Add.make_args(R))
        dt_terms = [(i, tt) for i, tt in enumerate(terms) if any(
            isinstance(d, sp.Derivative) and any(v == self.t for v, _ in d.variable_count) for d in tt.atoms(sp.Derivative))]
        if len(dt_terms) != 1:
            raise NotImplementedError("each equation needs exactly one time-derivative term: %s" % eq)
        i, dtt = dt_terms[0]
        dts = [d for d in dtt.atoms(sp.Derivative) if any(v == self.t for v, _ in d.variable_count)]
        dterm = dts[0]
        inner, vc = _deriv_spec(dterm)
        fvar = self._var_of(inner)
        if fvar is None or vc != [(self.t, 1)]:
            raise NotImplementedError("only first-order time derivatives of variables are supported: %s" % dterm)
        c = sp.simplify(dtt / dterm)
        if c.has(sp.Derivative) or not c.is_number:
            raise NotImplementedError("time derivative must have a constant coefficient: %s" % eq)
        rest = terms[:i] + terms[i + 1:]
        out = []
        for term in rest:
            sph = self._match_spherical(term)
            if sph is not None:
                cfac, sv, ax, a = sph
                p = self.p
                slot = dict(kind=SLOT_SPHERICAL, field=sv, axis=ax, axis2=0,
                            tab=[self.table(("centered", ax, 1)), self.table(("centered", ax, 2)),
                                 self.table(("half", ax, 1, p, -1)), self.table(("half", ax, 1, p)),
                                 self.table(("half", ax, 0, max(4, p)))],
                            prog=self._coef_program(a, sv, False) if a != 1 else -1, eps=0.0, inv_dx=0.0)
                # the same program is evaluated at the node (prefactor) and at half points
                k = self.slot(slot)
                out.append(cfac * sp.Symbol("__S%d" % k))
                continue
            # coefficient-signed upwinding of odd derivatives (for WENO, orders >= 3 only)
            w = self._match_winding(term) if not _has_nonlinlap(term, self) else None
            if w is not None:
                k = self.slot(w)
                out.append(sp.Symbol("__S%d" % k))
                continue
            out.append(self._replace_atoms(term))
        rhs = -sp.Add(*out) / c
        return fvar, rhs

    # ---------------------------------------------------------------- unknown boxes
    def boxes(self):
        nf = len(self.heads)
        self.lo = [[1] * self.ndim for _ in range(nf)]
        self.hi = [list(self.n) for _ in range(nf)]
        self.vlower = [[0] * self.ndim for _ in range(nf)]
        self.vupper = [[0] * self.ndim for _ in range(nf)]
        self.extent = [[(0, 0)] * self.ndim for _ in range(nf)]
        for f in range(nf):
            pde = self.eq_of[f]
            for a in range(self.ndim):
                # an interface clips the lower node of the upper variable (its alias) and keeps the
                # upper node of the lower one (clip_interior!!, interior_map.jl:1-10)
                up_if = self._is_iface(f, a, 1)
                cl = 1 if (self.periodic[a] or self.bc[f][a][0] is not None) else 0
                cu = 1 if (not self.periodic[a] and self.bc[f][a][1] is not None and not up_if) else 0
                self.vlower[f][a], self.vupper[f][a] = cl, cu
                ext = 0
                x = self.axes[a]
                if not self.periodic[a]:
                    for d in pde.atoms(sp.Derivative):
                        for v, cnt in d.variable_count:
                            if v == x and cnt % 2 == 1:
                                # WENO extent: 2 on uniform grids, 0 on nonuniform ones (WENO.jl:94-95)
                                if cnt == 1 and isinstance(self.scheme, WENOScheme) and not self.nonuniform[a]:
                                    ext = max(ext, 2)
                el, eu = (ext if not self.periodic[a] else 0), (ext if not self.periodic[a] else 0)
                if ext and (self._is_iface(f, a, 0) or up_if):
                    raise NotImplementedError("WENO extrapolated nodes next to an interface are not supported")
                if self.edge and (ext or self.periodic[a] or self._is_iface(f, a, 0) or up_if):
                    raise NotImplementedError("edge-aligned grids are supported with boundary conditions at both ends "
                                              "and no extrapolated nodes (no periodic axes, interfaces or WENO)")
                self.extent[f][a] = (el, eu)
                r0, r1 = self._range(f, a)
                self.lo[f][a] = r0 + max(cl, el)
                self.hi[f][a] = r1 - max(cu, eu)
                if self.lo[f][a] + max(cu, eu) > r1:
                    raise ValueError("The domain is too small to support the requested discretization")

    # ---------------------------------------------------------------- ghost closures
    def _central_row(self, op: fornberg.DerivativeOperator, i: int, n: int):
        """(weights, first tap) of central_difference at node i, non-periodic axis
        (centered_difference.jl:5-30)."""
        bpc, BL = op.boundary_point_count, op.boundary_stencil_length
        if i <= bpc:
            return op.low_boundary_coefs[i - 1], 1
        if i > n - bpc:
            return op.high_boundary_coefs[n - i], n - BL + 1
        return op.stencil_coefs, i - op.stencil_length // 2

    def closures(self) -> List[Closure]:
        out = []
        nf = len(self.heads)
        for f in range(nf):
            for a in range(self.ndim):
                if self.periodic[a]:
                    continue
                n = self.n[a]
                r0, r1 = self._range(f, a)
                for side in (0, 1):
                    lo, hi = self.lo[f][a], self.hi[f][a]
                    if self._is_iface(f, a, side):
                        # positions past the box on this side hold the other variable
                        reach = max(1, self.reach[a])
                        out.append(Closure(f, a, side, np.zeros((reach, 0)), np.zeros(reach), -1,
                                           src=self.bc[f][a][side][1]))
                        continue
                    ghosts = list(range(r0, lo)) if side == 0 else list(range(hi + 1, r1 + 1))
                    if not ghosts:
                        continue
                    rows = {g: [] for g in ghosts}   # g -> list of (coef dict node->w, gcoef)
                    bc = self.bc[f][a][side]
                    bnode = r0 if side == 0 else r1
                    if bc is not None and self.edge:
                        # the condition holds at the cell edge between the ghost and the first cell
                        # centre: values there by half-point interpolation (order max(4, p)), derivatives
                        # by half-point differences (boundary_value_maps, generate_bc_eqs.jl:79-161;
                        # the upper edge is half point n - 1, newindex(shift = true), MOL_utils.jl:87)
                        coefs, gexpr = bc
                        row = {}
                        hp = 1 if side == 0 else n - 1
                        for d, cval in coefs.items():
                            if cval == 0.0:
                                continue
                            op = fornberg.half_centered_difference(d, max(4, self.p) if d == 0 else self.p, self.dx[a])
                            w, first = half_row(op, hp, n)
                            for k, wk in enumerate(w):
                                row[first + k] = row.get(first + k, 0.0) + cval * wk
                        rows[bnode].append((row, 1.0))
                    elif bc is not None:
                        coefs, gexpr = bc
                        row = {}
                        for d, cval in coefs.items():
                            if cval == 0.0:
                                continue
                            if d == 0:
                                row[bnode] = row.get(bnode, 0.0) + cval
                            else:
                                if self.nonuniform[a]:
                                    op = fornberg.centered_difference_nonuniform(d, self.p, self.coords[a])
                                else:
                                    op = fornberg.centered_difference(d, self.p, self.dx[a])
                                w, first = self._central_row(op, bnode, n)
                                for k, wk in enumerate(w):
                                    row[first + k] = row.get(first + k, 0.0) + cval * wk
                        rows[bnode].append((row, 1.0))
                    # extrapolated nodes (generate_extrap_eqs!, generate_bc_eqs.jl:336-392)
                    el, eu = self.extent[f][a]
                    vl = self.vlower[f][a] if side == 0 else self.vupper[f][a]
                    ext = el if side == 0 else eu
                    ninterp = ext - vl
                    eop = fornberg.boundary_interp_extrap(max(6, self.p), self.dx[a]) if ext > 0 else None
                    while ninterp >= vl:
                        node = (1 + ninterp) if side == 0 else (n - ninterp)
                        if node in rows:
                            w, first = self._central_row(eop, node, n)
                            row = {node: 1.0}
                            for k, wk in enumerate(w):
                                row[first + k] = row.get(first + k, 0.0) - wk
                            rows[node].append((row, 0.0))
                        ninterp -= 1
                    # assemble A z = B u_int + cg g
                    missing = [g for g in ghosts if not rows[g]]
                    if missing:
                        raise NotImplementedError(
                            "boundary node(s) %s of %s along %s have no equation" % (missing, self.heads[f], self.axes[a]))
                    gi = {g: k for k, g in enumerate(ghosts)}
                    ng = len(ghosts)
                    first_int = lo if side == 0 else hi
                    dirn = 1 if side == 0 else -1
                    K = 0
                    for g in ghosts:
                        for row, _ in rows[g]:
                            for node in row:
                                if node not in gi:
                                    m = (node - first_int) * dirn
                                    if m < 0:
                                        raise NotImplementedError("boundary closure reaches the opposite side")
                                    K = max(K, m + 1)
                    A = np.zeros((ng, ng))
                    B = np.zeros((ng, K))
                    cg = np.zeros(ng)
                    for g in ghosts:
                        eqs = rows[g]
                        nr = len(eqs)
                        for row, gc in eqs:
                            for node, w in row.items():
                                if node in gi:
                                    A[gi[g], gi[node]] += w / nr
                                else:
                                    B[gi[g], (node - first_int) * dirn] -= w / nr
                            cg[gi[g]] += gc / nr
                    Ainv = np.linalg.inv(A)
                    C = Ainv @ B
                    c = Ainv @ cg
                    # order ghosts from the box edge outwards
                    order = sorted(ghosts, key=lambda g: abs(g - first_int))
                    C = np.array([C[gi[g]] for g in order])
                    c = np.array([c[gi[g]] for g in order])
                    gprog = -1
                    if bc is not None and bc[1] != 0:
                        gexpr, extra = self._field_subs(sp.sympify(bc[1]))
                        gprog = self.add_program(gexpr, extra)
                    out.append(Closure(f, a, side, C, c, gprog))
        return out

    def _check_row(self, op, i, lref, ax):
        """A node (or half point) i whose interior row is missing from a per-node operator: the
        reference's operator has no row there either."""
        if op.pn_lo > 0 and not (i <= op.nlow or i > lref - op.nhigh or op.pn_lo <= i <= op.pn_hi):
            raise NotImplementedError("index %d along %s has no row in the reference's nonuniform %s operator"
                                      % (i, self.axes[ax], op.kind))

    # ---------------------------------------------------------------- halo
    def halo(self) -> List[int]:
        reach = [0] * self.ndim

        def row_span(op, i, n, per, lref_shift=0):
            if op.kind == "half":
                lref = n + lref_shift
                if not per and i <= op.nlow:
                    return 1, op.boundary_stencil_length
                if not per and i > lref - op.nhigh:
                    return lref - op.boundary_stencil_length + 1, lref
                return i + op.tap_offset, i + op.tap_offset + op.stencil_length - 1
            if not per and i <= op.nlow:
                return 1, op.boundary_stencil_length
            if not per and i > n - op.nhigh:
                return n - op.boundary_stencil_length + 1, n
            return i + op.tap_offset, i + op.tap_offset + op.stencil_length - 1

        for s in self.slots:
            axes = [s["axis"]] + ([s["axis2"]] if s["kind"] == SLOT_MIXED else [])
            for ax in axes:
                n, per = self.n[ax], self.periodic[ax]
                f = s["field"]
                los = [self.lo[g][ax] for g in range(len(self.heads))]
                his = [self.hi[g][ax] for g in range(len(self.heads))]
                lo, hi = min(los), max(his)
                nodes = sorted(set(list(range(lo, min(hi, lo + 12) + 1)) + list(range(max(lo, hi - 12), hi + 1))))
                for i in nodes:
                    spans = []
                    if s["kind"] == SLOT_WENO:
                        spans.append((i - 2, i + 2))
                    elif s["kind"] == SLOT_WENO_NU:
                        if not per and i <= 2:
                            spans.append((1, 5))
                        elif not per and i > n - 2:
                            spans.append((n - 4, n))
                        else:
                            spans.append((i - 2, i + 2))
                    elif s["kind"] in (SLOT_CENTERED, SLOT_UPWIND, SLOT_MIXED):
                        tabs = s["tab"][:2] if s["kind"] == SLOT_UPWIND else [s["tab"][0 if ax == s["axis"] else 1]]
                        for tb in tabs:
                            op = self.tables[tb].op
                            self._check_row(op, i, n, ax)
                            spans.append(row_span(op, i, n, per))
                    else:
                        if s["kind"] == SLOT_SPHERICAL:
                            for tb in s["tab"][:2]:
                                self._check_row(self.tables[tb].op, i, n, ax)
                                spans.append(row_span(self.tables[tb].op, i, n, per))
                            to, ti, tp = s["tab"][2:5]
                        else:
                            to, ti, tp = s["tab"][:3]
                        self._check_row(self.tables[to].op, i - 1, n - 1, ax)
                        h0, h1 = row_span(self.tables[to].op, i - 1, n, per, -1)
                        for h in range(h0, h1 + 1):
                            for tb in (ti, tp):
                                self._check_row(self.tables[tb].op, h, n, ax)
                                spans.append(row_span(self.tables[tb].op, h, n, per))
                    for a0, a1 in spans:
                        reach[ax] = max(reach[ax], i - a0, a1 - i)
        return reach


def _has_nonlinlap(term, low: _Lowering) -> bool:
    for d in term.atoms(sp.Derivative):
        if low._match_nonlinlap(d) is not None:
            return True
    return False


def lower(pdesys: PDESystem, disc: MOLFiniteDifference) -> Lowered:
    """Lower a PDE system to the stencil IR (the counterpart of symbolic_discretize)."""
    L = _Lowering(pdesys, disc)
    L.setup()
    L.parse_bcs()
    nf = len(L.heads)
    rhs = [None] * nf
    L.eq_of = [None] * nf
    # map equations to variables by their time derivative (the tearing of interior_map.jl:53-103)
    for eq in pdesys.eqs:
        if L.sym_map:
            eq = sp.Eq(sp.sympify(eq.lhs).xreplace(L.sym_map), sp.sympify(eq.rhs).xreplace(L.sym_map), evaluate=False)
        fvar, r = L.lower_equation(eq)
        if rhs[fvar] is not None:
            raise ValueError("two equations solve for %s" % L.heads[fvar])
        rhs[fvar] = r
        L.eq_of[fvar] = eq.lhs - eq.rhs
    if any(r is None for r in rhs):
        raise ValueError("There must be the same number of equations and unknowns")
    L.boxes()
    rhs_prog = []
    for f in range(nf):
        expr, extra = L._field_subs(rhs[f])
        for k in range(len(L.slots)):
            extra[sp.Symbol("__S%d" % k)] = ("SLOT", k)
        L.rhs_exprs.append(expr)
        rhs_prog.append(L.add_program(expr, extra))
    halo = L.halo()
    L.reach = halo
    closures = L.closures()
    # initial values (generate_ic_defaults.jl:1-27), evaluated in FP64 at every unknown
    u0 = []
    for f in range(nf):
        ic = L.ic[f]
        ic = ic.xreplace(dict(zip(L.params, [sp.Float(v) for v in L.param_values])))
        fn = sp.lambdify(L.axes, ic, modules="numpy")
        grids = [L.coords[a][L.lo[f][a] - 1:L.hi[f][a]] for a in range(L.ndim)]
        mesh = np.meshgrid(*grids, indexing="ij")
        vals = np.asarray(fn(*mesh), dtype=np.float64)
        vals = np.broadcast_to(vals, mesh[0].shape)
        # flat order: first axis fastest
        u0.append(np.asarray(vals).reshape(-1, order="F"))
    u0 = np.concatenate(u0)
    return Lowered(
        ndim=L.ndim, axes=L.axes, n=L.n, periodic=L.periodic, coords=L.coords,
        fields=[str(h) for h in L.heads], field_heads=L.heads, lo=L.lo, hi=L.hi, halo=halo,
        tables=L.tables, slots=L.slots, closures=closures, programs=L.programs,
        cpool=L.cpool.values, rhs_prog=rhs_prog, params=L.params, param_values=L.param_values,
        aux=[h.__name__ for h in L.aux_heads], aux_heads=L.aux_heads, dtype=disc.dtype,
        u0=u0, time=L.t, tspan=L.tspan, rhs_exprs=L.rhs_exprs,
        interface=None if L.iface is None else dict(axis=L.iface["axis"], range=L.iface["range"],
                                                   symbols=L.iface["symbols"]))
