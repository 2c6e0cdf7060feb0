"""Host-side logic on the CPU: weights, grids, lowering, bytecode, and the C ABI surface."""
import ctypes as C
import math
import os
import re
from fractions import Fraction

import numpy as np
import pytest
import sympy as sp

import oracle as orc
from molgpu import (Eq, Differential, Interval, MOLFiniteDifference, PDESystem, UpwindScheme,
                    WENOScheme, ifelse, lower, parameters, uniform_grid, variables, native)
from molgpu import fornberg
from molgpu.bytecode import OP, NREG, Compiler, ConstPool

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- weights
@pytest.mark.parametrize("order,x0,x", [
    (2, 0.0, [-1, 0, 1]), (1, 0.0, [-1, 1]), (3, 0.0, range(6)), (4, 0.5, range(-3, 4)),
    (1, 1.5, [1, 2, 3]), (2, 0.37, [-2, -1, 0, 1, 2, 3]), (0, 2.5, [1, 2, 3, 4]), (6, 0.0, range(-4, 5)),
])
def test_fornberg_matches_oracle_bitwise(order, x0, x):
    a = fornberg.calculate_weights(order, x0, list(x))
    b = orc.calculate_weights(order, x0, list(x))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("d,p,dx", [(1, 2, 0.1), (2, 2, 1 / 512), (2, 4, 0.03), (3, 4, 0.5), (1, 6, 0.2)])
def test_centered_tables_match_oracle(d, p, dx):
    op = fornberg.centered_difference(d, p, dx)
    st, low, high = orc.centered_weights(d, p, dx)
    assert np.array_equal(op.stencil_coefs, st)
    assert np.array_equal(np.array(op.low_boundary_coefs), low)
    assert np.array_equal(np.array(op.high_boundary_coefs), high)


@pytest.mark.parametrize("d,p,dx", [(1, 2, 0.1), (0, 4, 0.25), (1, 4, 1 / 127), (0, 6, 0.1)])
def test_half_tables_match_oracle(d, p, dx):
    op = fornberg.half_centered_difference(d, p, dx)
    st, low, high = orc.half_weights(d, p, dx)
    assert np.array_equal(op.stencil_coefs, st)
    assert np.array_equal(np.array(op.low_boundary_coefs), low)
    assert np.array_equal(np.array(op.high_boundary_coefs), high)


def test_upwind_tables():
    op = fornberg.upwind_difference(1, 1, 0.1, 0)
    np.testing.assert_allclose(op.stencil_coefs, [-10, 10])
    assert op.tap_offset == 0 and op.nlow == 0 and op.nhigh == 1
    op = fornberg.upwind_difference(1, 1, 0.1, 1)
    np.testing.assert_allclose(op.stencil_coefs, [-10, 10])
    assert op.tap_offset == -1 and op.nlow == 1 and op.nhigh == 0


def test_extrapolator_rows_are_lagrange():
    op = fornberg.boundary_interp_extrap(6, 0.1)
    # interior row excludes the node itself and is exact on quintics
    xs = np.arange(-2, 3)
    for deg in range(4):
        assert abs(op.stencil_coefs @ xs.astype(float) ** deg - (0.0 if deg else 1.0)) < 1e-12
    assert op.stencil_coefs[2] == 0.0


# ---------------------------------------------------------------- grids
def test_uniform_grid_matches_julia_ranges():
    g = uniform_grid(0.0, 1.0, 0.01)
    assert len(g) == 101
    for k in range(101):
        assert g[k] == float(Fraction(k, 100))
    g = uniform_grid(0.0, 1.0, 1 / 512)
    assert len(g) == 513 and np.array_equal(g, np.arange(513) / 512)
    g = uniform_grid(0.0, 1.0, 128)
    assert len(g) == 128 and g[-1] == 1.0 and g[1] == float(Fraction(1, 127))


# ---------------------------------------------------------------- lowering
t, x, y, z = parameters("t x y z")
u, v = variables("u v")
Dt, Dx, Dy, Dz = Differential(t), Differential(x), Differential(y), Differential(z)


def heat_sys(bcl, bch):
    eq = Eq(Dt(u(t, x)), (Dx ** 2)(u(t, x)))
    bcs = [Eq(u(0, x), sp.sin(sp.pi * x)), bcl, bch]
    return PDESystem([eq], bcs, [Interval(t, 0, 1), Interval(x, 0, 1)], [t, x], [u(t, x)])


def test_lower_heat_dirichlet():
    L = lower(heat_sys(Eq(u(t, 0), 0), Eq(u(t, 1), 0)), MOLFiniteDifference({x: 0.01}, t))
    assert L.n == [101] and L.lo == [[2]] and L.hi == [[100]] and L.halo == [1]
    assert len(L.closures) == 2
    for cl in L.closures:
        assert cl.C.shape == (1, 0) and cl.c.tolist() == [1.0]
    np.testing.assert_allclose(L.u0, np.sin(np.pi * np.arange(1, 100) / 100), rtol=1e-14)


def test_lower_neumann_closure_closed_form():
    L = lower(heat_sys(Eq(Dx(u(t, 0)), 0), Eq(Dx(u(t, 1)), 0)), MOLFiniteDifference({x: 0.05}, t))
    for cl in L.closures:
        np.testing.assert_allclose(cl.C, [[4 / 3, -1 / 3]], rtol=1e-14)
        assert cl.gprog == -1 or True


def test_lower_robin_closure_closed_form():
    a0, a1, g = 2.0, 0.5, 0.3
    dx = 0.05
    L = lower(heat_sys(Eq(a0 * u(t, 0) + a1 * Dx(u(t, 0)), g), Eq(a0 * u(t, 1) + a1 * Dx(u(t, 1)), g)),
              MOLFiniteDifference({x: dx}, t))
    lo = [c for c in L.closures if c.side == 0][0]
    den = a0 - 1.5 * a1 / dx
    np.testing.assert_allclose(lo.C[0], [-a1 * 2 / dx / den, a1 * 0.5 / dx / den], rtol=1e-13)
    np.testing.assert_allclose(lo.c, [1 / den], rtol=1e-13)


def test_lower_periodic_and_weno_boxes():
    eq = Eq(Dt(u(t, x)), -Dx(u(t, x)))
    bcs = [Eq(u(0, x), sp.sin(2 * sp.pi * x)), Eq(u(t, 0), u(t, 1))]
    L = lower(PDESystem([eq], bcs, [Interval(t, 0, 1), Interval(x, 0, 1)], [t, x], [u(t, x)]),
              MOLFiniteDifference({x: 0.05}, t, advection_scheme=WENOScheme()))
    assert L.periodic == [True] and L.lo == [[2]] and L.hi == [[21]] and L.halo == [2]
    assert L.slots[0]["kind"] == 2
    # non-periodic WENO: extent 2 clips the unknowns, with one extrapolated node per side
    bcs = [Eq(u(0, x), sp.sin(2 * sp.pi * x)), Eq(u(t, 0), 0), Eq(u(t, 1), 0)]
    L = lower(PDESystem([eq], bcs, [Interval(t, 0, 1), Interval(x, 0, 1)], [t, x], [u(t, x)]),
              MOLFiniteDifference({x: 0.05}, t, advection_scheme=WENOScheme()))
    assert L.lo == [[3]] and L.hi == [[19]]
    assert all(cl.C.shape[0] == 2 for cl in L.closures)


def test_lower_upwind_coefficient_slot():
    eq = Eq(Dt(u(t, x)), -(sp.sin(x) + 0.1) * Dx(u(t, x)))
    bcs = [Eq(u(0, x), 0), Eq(u(t, 0), 0), Eq(u(t, 1), 0)]
    L = lower(PDESystem([eq], bcs, [Interval(t, 0, 1), Interval(x, 0, 1)], [t, x], [u(t, x)]),
              MOLFiniteDifference({x: 0.1}, t, advection_scheme=UpwindScheme()))
    s = L.slots[0]
    assert s["kind"] == 1 and s["prog"] >= 0
    # the program is -(slot): the coefficient lives inside the slot (cardinalized sign)
    assert str(L.rhs_exprs[0]) == "-__S0"


def test_lower_brusselator_program_and_boxes():
    alpha = 10.0
    f = ifelse((x - 0.3) ** 2 + (y - 0.6) ** 2 <= 0.1 ** 2, 1.0, 0.0) * ifelse(t >= 1.1, 1.0, 0.0) * 5.0
    lap = lambda w: (Dx ** 2)(w) + (Dy ** 2)(w)
    U, V = u(x, y, t), v(x, y, t)
    eqs = [Eq(Dt(U), 1.0 + V * U ** 2 - 4.4 * U + alpha * lap(U) + f), Eq(Dt(V), 3.4 * U - V * U ** 2 + alpha * lap(V))]
    bcs = [Eq(u(x, y, 0), 22 * (y * (1 - y)) ** 1.5), Eq(u(0, y, t), u(1, y, t)), Eq(u(x, 0, t), u(x, 1, t)),
           Eq(v(x, y, 0), 27 * (x * (1 - x)) ** 1.5), Eq(v(0, y, t), v(1, y, t)), Eq(v(x, 0, t), v(x, 1, t))]
    sysb = PDESystem(eqs, bcs, [Interval(x, 0, 1), Interval(y, 0, 1), Interval(t, 0, 11.5)], [x, y, t], [U, V])
    L = lower(sysb, MOLFiniteDifference({x: 1 / 512, y: 1 / 512}, t))
    assert L.n == [513, 513] and L.lo == [[2, 2]] * 2 and L.hi == [[513, 513]] * 2
    assert L.n_unknowns == 2 * 512 * 512
    assert len(L.slots) == 4 and all(s["kind"] == 0 for s in L.slots)


# ---------------------------------------------------------------- bytecode
def _decode_run(code, cpool, leaves):
    r = [0.0] * NREG
    inv = {v: k for k, v in OP.items()}
    for ins in code:
        op = inv[ins & 0xFF]
        d, a, b = (ins >> 8) & 15, (ins >> 12) & 15, (ins >> 16) & 15
        bconst = (ins >> 20) & 1
        imm = ins >> 21
        simm = imm - 2048 if imm & 0x400 else imm
        A = r[a]
        B = cpool[imm] if bconst else r[b]
        Cc = r[d]
        if op == "CONST": res = cpool[imm]
        elif op in ("FIELD", "SLOT", "COORD", "TIME", "PARAM", "AUX", "DINNER"): res = leaves[(op, imm)]
        elif op == "ADD": res = A + B
        elif op == "SUB": res = A - B
        elif op == "MUL": res = A * B
        elif op == "DIV": res = A / B
        elif op == "POW": res = A ** B
        elif op == "MIN": res = min(A, B)
        elif op == "MAX": res = max(A, B)
        elif op == "LT": res = float(A < B)
        elif op == "LE": res = float(A <= B)
        elif op == "GT": res = float(A > B)
        elif op == "GE": res = float(A >= B)
        elif op == "AND": res = float(A != 0 and B != 0)
        elif op == "OR": res = float(A != 0 or B != 0)
        elif op == "SELECT": res = A if Cc != 0 else B
        elif op == "NEG": res = -A
        elif op == "INV": res = 1.0 / A
        elif op == "SQ": res = A * A
        elif op == "POWI": res = A ** simm
        elif op == "SQRT": res = math.sqrt(A)
        elif op == "EXP": res = math.exp(A)
        elif op == "SIN": res = math.sin(A)
        elif op == "COS": res = math.cos(A)
        elif op == "TANH": res = math.tanh(A)
        elif op == "ABS": res = abs(A)
        elif op == "LOG": res = math.log(A)
        else: raise AssertionError(op)
        r[d] = res
    return r[0]


def test_bytecode_matches_sympy():
    a, b, c, tt = sp.symbols("a b c tt")
    leafmap = {a: ("FIELD", 0), b: ("FIELD", 1), c: ("SLOT", 0), tt: ("TIME", 0)}
    exprs = [
        1.0 + b * a ** 2 - 4.4 * a + 10.0 * c,
        3.4 * a - b * a ** 2 + 10.0 * c,
        -a * c - b * c / (1 + a ** 2),
        sp.exp(-a) * sp.sin(b) + sp.sqrt(a + 2) - sp.tanh(c) ** 3,
        sp.Piecewise((5.0, (a - 0.3) ** 2 + (b - 0.6) ** 2 <= 0.01), (0.0, True)) * sp.Piecewise((1.0, tt >= 1.1), (0.0, True)),
        sp.Max(a, b, c) - sp.Min(a, 0.5) + abs(c - a),
        (a + b) * (b + c) * (c + a) * (a - b) / (1 + c * c),
        a ** -2 + b ** 2.5 + (1 + a) ** (-0.5),
    ]
    rng = np.random.default_rng(1)
    for e in exprs:
        cp = ConstPool()
        code = Compiler(lambda s: leafmap[s], cp).compile(e)
        fn = sp.lambdify((a, b, c, tt), e, "math")
        for _ in range(25):
            va, vb, vc, vt = rng.uniform(0.2, 1.5), rng.uniform(0.2, 1.5), rng.uniform(-1, 1), rng.uniform(0, 2)
            got = _decode_run(code, cp.values, {("FIELD", 0): va, ("FIELD", 1): vb, ("SLOT", 0): vc, ("TIME", 0): vt})
            ref = fn(va, vb, vc, vt)
            assert abs(got - ref) <= 1e-13 * max(1.0, abs(ref)), (e, got, ref)


# ---------------------------------------------------------------- C ABI surface
def _declared_functions():
    hdr = open(os.path.join(ROOT, "include", "mol.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(mol_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(native.LIB_PATH)
    names = _declared_functions()
    assert "mol_create" in names and "mol_rhs" in names and "mol_integrate" in names
    for n in names:
        assert hasattr(lib, n), n


def test_create_rejects_bad_descriptors_without_a_device():
    lib = native.load_library()
    d = native.mol_desc()
    d.abi_version = 999
    h = C.c_void_p()
    rc = lib.mol_create(C.byref(d), C.byref(h))
    assert rc == -1 and b"ABI" in lib.mol_last_error(None)
    L = lower(heat_sys(Eq(u(t, 0), 0), Eq(u(t, 1), 0)), MOLFiniteDifference({x: 0.01}, t))
    d, keep = native.pack_descriptor(L)
    d.lo[0][0] = 0
    rc = lib.mol_create(C.byref(d), C.byref(h))
    assert rc == -2
    d, keep = native.pack_descriptor(L)
    d.nslots = 10 ** 6
    assert lib.mol_create(C.byref(d), C.byref(h)) == -1


# ---------------------------------------------------------------- code generation (no GPU)
def _codegen_systems():
    from molgpu import systems as S
    return {
        "heat_p2": lambda: S.heat1d(n=1200, D=0.8),
        "heat_robin_p4": lambda: S.heat1d(n=1200, D=0.8, bcl=(2.0, 0.5, 0.3), bch=(1.5, -0.25, -0.2), approx_order=4),
        "brusselator": lambda: S.brusselator2d(65),
        "burgers_weno": lambda: S.burgers2d(65),
        "weno_dirichlet": lambda: S.weno1d_dirichlet(1200, 1.0, 1.0, 0.1, 0.05),
        "upwind_o1": lambda: S.upwind1d(1200, 0.1, 1.0, 0.3, -0.2, 1),
        "upwind_o2": lambda: S.upwind1d(1200, 0.1, 1.0, 0.3, -0.2, 2),
        "mixed_p4": lambda: S.mixed2d(200, 0.7, 4),
        "spherical": lambda: S.spherical1d(1200, 0.25, 2),
        "nonlinlap_periodic_p4": lambda: S.nonlinlap1d(1200, True, 0.2, -0.4, 4),
        "nonlindiff3d_aux": lambda: S.nonlindiff3d(40),
        "diffreact_mixed_bcs": lambda: S.diffreact2d(200, 0.05, 2.5, (0, 1, 2, 0)),
    }


@pytest.mark.parametrize("name", sorted(_codegen_systems()))
def test_generated_kernel_compiles_for_gfx950(name):
    """Every operator-slot kind lowers to straight-line code that hiprtc compiles for gfx950."""
    import molgpu
    sys_, disc = _codegen_systems()[name]()
    src = molgpu.codegen(lower(sys_, disc), compile=True)
    assert 'extern "C" __global__' in src and "struct GeoB" in src
    # the interpreter is not part of the generated kernel
    assert "stack_prog" not in src and "cpool" not in src


@pytest.mark.parametrize("name", sorted(_codegen_systems()))
def test_generated_vjp_kernel_compiles_for_gfx950(name):
    """The reverse-mode kernels of every slot kind compile: the tile kernel (LDS tap adjoints
    with LDS atomics, ghost-node adjoints through the transposed closures, tile + halo adjoints
    to per-tile buffers) and the gather kernel that sums the buffers."""
    import molgpu
    sys_, disc = _codegen_systems()[name]()
    src = molgpu.codegen(lower(sys_, disc), compile=True, vjp=True)
    assert "mol_vjp_tile<GeoB, VOLX>" in src and "static __device__ void scatter(" in src
    assert "mol_vjp_gather(VArgs va)" in src


def test_generated_kernel_folds_exact_weights():
    """Interior weights appear as exact hexadecimal literals, collected per tap as the reference's
    generated code has them (docs/src/generated/bruss_code.md:46-118: 160.0 per neighbour and
    -644.4 at the node for alpha = 10, dx = 1/4): alpha/dx^2 = 10 * 64^2 per neighbour and
    -4 alpha/dx^2 - 4.4 at the node for dx = 1/64."""
    import molgpu
    sys_, disc = systems_bruss = __import__("molgpu").systems.brusselator2d(65)
    src = molgpu.codegen(lower(sys_, disc))
    nb = "T(%s)" % _c_hex(10.0 * 64 ** 2)
    for tap in ("L[b - 1]", "L[b + 1]"):
        assert "%s * %s" % (nb, tap) in src, tap
    assert "T(%s) * L[b]" % _c_hex(-4 * 10.0 * 64 ** 2 - 4.4) in src
    assert "static constexpr int PER0 = 1;" in src and "static constexpr int N0 = 65;" in src


def test_folded_weights_match_reference_fixture():
    """At the reference's own fixture size (n = 5, dx = 1/4) the collected weights are the
    numbers of its generated code: 160.0 per neighbour, -644.4 at the node for u and -640.0 for
    v, 3.4 for u in the v row (docs/src/generated/bruss_code.md:46-118)."""
    import molgpu
    sys_, disc = molgpu.systems.brusselator2d(5)
    src = molgpu.codegen(lower(sys_, disc))
    for w in (160.0, -644.4, -640.0, 3.4):
        assert "T(%s) * L[" % _c_hex(w) in src, w


def _c_hex(x: float) -> str:
    """printf("%a") spelling of a double (trailing zero hex digits dropped)."""
    m, e = float(x).hex().split("p")
    m = m.rstrip("0").rstrip(".")
    return m + "p" + e


def test_generated_kernel_covers_boundary_rows_and_closures():
    """A small Robin problem has no interior tile: the generated kernel selects the boundary
    rows per node index and stages ghost nodes from the closure (affine in the unknowns plus
    c * g)."""
    import molgpu
    from molgpu import systems as S
    sys_, disc = S.heat1d(n=21, D=0.8, bcl=(2.0, 0.5, 0.3), bch=(1.5, -0.25, -0.2), approx_order=4)
    low = lower(sys_, disc)
    src = molgpu.codegen(low, compile=True)
    assert "if constexpr (CORE)" not in src          # no core tile at n = 21
    assert "I0 == 2)" in src and "I0 == 20)" in src
    # the ghost node left of node 2 combines the unknowns at nodes 2..5 with the closure weights
    cl = [c for c in low.closures if c.side == 0][0]
    for m in range(4):
        assert "T(%s) * u[uidx(0, %d, p1, p2)]" % (_c_hex(cl.C[0, m]), 2 + m) in src
    assert _c_hex(cl.c[0]) in src


def test_batched_kernel_has_one_class_per_boundary_pattern():
    """Instances of one structure share generated code whatever their size and coefficients;
    the batched kernel compiles for gfx950."""
    from molgpu import systems as S
    from molgpu.batch import codegen_batched
    rng = np.random.default_rng(0)
    lows, pats = [], set()
    for _ in range(10):
        n = int(rng.integers(10, 60))
        bct = tuple(int(b) for b in rng.integers(0, 2, 4))
        pats.add(bct)
        sys_, disc = S.diffreact2d(n, float(rng.uniform(0.01, 0.1)), float(rng.uniform(0.5, 5)), bct, dtype="f32")
        lows.append(lower(sys_, disc))
    src = codegen_batched(lows, compile=True)
    assert src.count("struct Geo") == len(pats)
    assert "hd.W[" in src and "hd.K[" in src and "mol_batch_tiles" in src


def test_solution_interpolation_is_multilinear():
    """sol(t, x, y; dv) interpolates the saved full grids linearly in t and in each axis
    (Gridded(Linear()), src/interface/solution/solution_utils.jl:3-32): exact on data that is
    linear in each variable separately, equal to the grid values at grid points."""
    import numpy as np
    import molgpu
    from molgpu import systems
    from molgpu.discretization import lower
    low = lower(*systems.diffreact2d(9, 0.05, 2.5, (0, 1, 1, 0)))
    prob = molgpu.ODEProblem(low)
    x, y = low.coords[0], low.coords[1]
    ts = [0.0, 0.5, 1.5]
    f = lambda t, X, Y: (1 + 2 * t) * (0.3 + X) * (2 - Y) + 0.7 * t * Y
    X, Y = np.meshgrid(x, y, indexing="ij")
    full = [np.stack([f(t, X, Y)]) for t in ts]
    sol = molgpu.ODESolution(ts, [None] * 3, prob, full)
    rng = np.random.default_rng(0)
    tq, xq, yq = rng.uniform(0, 1.5, 50), rng.uniform(0, 1, 50), rng.uniform(0, 1, 50)
    np.testing.assert_allclose(sol(tq, xq, yq, dv="u"), f(tq, xq, yq), rtol=1e-13, atol=1e-13)
    assert sol(0.5, x[3], y[5]) == f(0.5, x[3], y[5])
    np.testing.assert_array_equal(sol["u"], np.stack([g[0] for g in full]))
    np.testing.assert_array_equal(sol[low.axes[1]], y)
