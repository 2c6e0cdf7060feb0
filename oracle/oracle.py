"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the benchmark's
CPU-baseline leg.  The product path never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sig = {
            "orc_calculate_weights": [C.c_int, C.c_double, _dp, C.c_int, _dp],
            "orc_centered_weights": [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _ip],
            "orc_half_weights": [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _ip],
            "orc_weno_f": [_dp, C.c_double, C.c_double, _dp],
            "orc_weno_nu": [_dp, _dp, C.c_double, C.c_int, _dp],
            "orc_weno_nu1d_rhs": [C.c_int, _dp, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, _dp,
                                  C.c_double, _dp],
            "orc_nonuniform1d_rhs": [C.c_int, _dp, C.c_double, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp,
                                     C.c_double, _dp],
            "orc_nonlinlap1d_nu_rhs": [C.c_int, _dp, C.c_int, C.c_double, C.c_double, _dp, C.c_double, _dp],
            "orc_spherical1d_nu_rhs": [C.c_int, _dp, C.c_int, C.c_double, _dp, C.c_double, _dp],
            "orc_diffreact_nd_rhs": [C.c_int, _ip, C.POINTER(C.c_void_p), C.c_int, C.c_double, C.c_double, _ip,
                                     _dp, C.c_double, _dp],
            "orc_diffusion1d_rhs": [C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double, _dp],
            "orc_diffusion1d_edge_rhs": [C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double,
                                         _dp],
            "orc_diffusion1d_integrate": [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double, C.c_double, C.c_longlong, C.c_int],
            "orc_bruss2d_rhs": [C.c_int, C.c_int, C.c_double, _dp, C.c_double, _dp, C.c_int],
            "orc_bruss2d_integrate": [C.c_int, C.c_int, C.c_double, _dp, C.c_double, C.c_double, C.c_longlong, C.c_int],
            "orc_burgers2d_rhs": [C.c_int, C.c_double, _dp, C.c_double, _dp, C.c_int],
            "orc_burgers2d_integrate": [C.c_int, C.c_double, _dp, C.c_double, C.c_double, C.c_longlong, C.c_int],
            "orc_weno1d_dirichlet_rhs": [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp, C.c_double, _dp],
            "orc_diffreact2d_rhs": [C.c_int, C.c_double, C.c_double, _ip, _dp, C.c_double, _dp],
            "orc_nonlindiff3d_rhs": [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, _dp, C.c_int],
            "orc_nonlindiff3d_jvp": [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_longlong, C.POINTER(C.c_double), C.POINTER(C.c_double)],
            "orc_upwind1d_rhs": [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp, C.c_double, _dp],
            "orc_mixed2d_rhs": [C.c_int, C.c_int, C.c_double, _dp, C.c_double, _dp],
            "orc_interface1d_rhs": [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, _dp, C.c_double, _dp],
            "orc_mixed2d_grid_rhs": [_ip, C.POINTER(C.c_void_p), C.c_int, C.c_double, _dp, _dp, C.c_double, _dp],
            "orc_spherical1d_rhs": [C.c_int, C.c_int, C.c_double, _dp, C.c_double, _dp],
            "orc_nonlinlap1d_rhs": [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp, C.c_double, _dp],
            "orc_num_threads": [],
            "orc_expect": [C.c_longlong],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _state(u):
    """A state vector for an oracle right-hand side: its length is checked against the system's
    number of unknowns by the call that follows."""
    u = _f64(u)
    lib().orc_expect(len(u))
    return u


def _ok(rc):
    if rc != 0:
        raise RuntimeError("oracle call failed")


def calculate_weights(order, x0, x):
    x = _f64(x)
    out = np.empty_like(x)
    _ok(lib().orc_calculate_weights(order, float(x0), x, len(x), out))
    return out


def centered_weights(d, p, dx):
    L = d + p - 1 + (d + p) % 2
    BL = d + p
    st = np.zeros(L)
    low = np.zeros(L * BL + 1)
    high = np.zeros(L * BL + 1)
    dims = np.zeros(3, dtype=np.int32)
    lib().orc_centered_weights(d, p, float(dx), st, low, high, dims)
    bpc = dims[2]
    return st, low[:bpc * BL].reshape(bpc, BL), high[:bpc * BL].reshape(bpc, BL)


def half_weights(d, p, dx):
    L = p + 2 * (d // 2) + p % 2
    BL = d + p
    st = np.zeros(L)
    low = np.zeros(L * BL + 1)
    high = np.zeros(L * BL + 1)
    dims = np.zeros(3, dtype=np.int32)
    lib().orc_half_weights(d, p, float(dx), st, low, high, dims)
    bpc = dims[2]
    return st, low[:bpc * BL].reshape(bpc, BL), high[:bpc * BL].reshape(bpc, BL)


def weno_f(u5, eps, dx):
    out = np.zeros(1)
    lib().orc_weno_f(_f64(u5), float(eps), float(dx), out)
    return out[0]


def weno_nu(u5, x5, eps, target):
    """Nonuniform WENO-5 first derivative at target node 1..5 of the 5-node stencil."""
    out = np.zeros(1)
    _ok(lib().orc_weno_nu(_f64(u5), _f64(x5), float(eps), int(target), out))
    return float(out[0])


def weno_nu1d_rhs(x, c, eps, periodic, gl, gh, u, t=0.0):
    """u_t = -c u_x with nonuniform WENO5 on nodes x (Dirichlet gl/gh or periodic)."""
    x = _f64(x)
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_weno_nu1d_rhs(len(x), x, float(c), float(eps), 1 if periodic else 0, float(gl), float(gh),
                                u, float(t), du))
    return du


def nonuniform1d_rhs(x, D, c, p, uo, bcl, bch, u, t=0.0):
    """u_t = D u_xx - c u_x on the nodes x: centred (uo = 0) or upwind (order uo) advection,
    a0 u + a1 u_x = g at each end (bcl, bch = (a0, a1, g))."""
    x = _f64(x)
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_nonuniform1d_rhs(len(x), x, float(D), float(c), int(p), int(uo), _f64(bcl), _f64(bch), u,
                                   float(t), du))
    return du


def nonlinlap1d_nu_rhs(x, p, gl, gh, u, t=0.0):
    """u_t = d/dx((1 + u^2) u_x) on the nodes x, Dirichlet gl / gh."""
    x, u = _f64(x), _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_nonlinlap1d_nu_rhs(len(x), x, int(p), float(gl), float(gh), u, float(t), du))
    return du


def spherical1d_nu_rhs(x, p, g1, u, t=0.0):
    """Spherical Laplacian on the nodes x (r from 0), u_r(0) = 0, u(1) = g1."""
    x, u = _f64(x), _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_spherical1d_nu_rhs(len(x), x, int(p), float(g1), u, float(t), du))
    return du


def diffreact_nd_rhs(grids, p, D, k, bctype, u, t=0.0):
    """u_t = D Lap u + k u (1 - u) on [0,1]^d; grids[a]: a node count (uniform) or the nodes;
    bctype per side (x-lo, x-hi, y-lo, ...): 0 Dirichlet 0, 1 Neumann 0."""
    nd = len(grids)
    keep = [None if np.isscalar(g) else _f64(g) for g in grids]
    n = np.array([int(g) if np.isscalar(g) else len(g) for g in grids] + [1] * (3 - nd), dtype=np.int32)
    xs = (C.c_void_p * 3)(*[(k_.ctypes.data if k_ is not None else None) for k_ in keep] + [None] * (3 - nd))
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_diffreact_nd_rhs(nd, n, xs, int(p), float(D), float(k),
                                   np.ascontiguousarray(bctype, dtype=np.int32), u, float(t), du))
    return du


def diffusion1d_rhs(n, D, bcl, bch, u, t=0.0, p=2, xlo=0.0, xhi=1.0):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_diffusion1d_rhs(n, xlo, xhi, p, D, _f64(bcl), _f64(bch), u, float(t), du))
    return du


def diffusion1d_edge_rhs(ncell, D, bcl, bch, u, t=0.0, p=2, xlo=0.0, xhi=1.0):
    """u_t = D u_xx on the edge-aligned grid of ncell cells (state: the ncell cell centres)."""
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_diffusion1d_edge_rhs(ncell, xlo, xhi, p, D, _f64(bcl), _f64(bch), u, float(t), du))
    return du


def diffusion1d_integrate(n, D, bcl, bch, u, t0, dt, nsteps, scheme=0, p=2):
    u = _state(u).copy()
    _ok(lib().orc_diffusion1d_integrate(n, p, D, _f64(bcl), _f64(bch), u, t0, dt, nsteps, scheme))
    return u


def bruss2d_rhs(n, u, t, alpha=10.0, p=2, threads=1):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_bruss2d_rhs(n, p, alpha, u, float(t), du, threads))
    return du


def bruss2d_integrate(n, u, t0, dt, nsteps, alpha=10.0, p=2, threads=1):
    u = _state(u).copy()
    _ok(lib().orc_bruss2d_integrate(n, p, alpha, u, t0, dt, nsteps, threads))
    return u


def burgers2d_rhs(n, u, eps=1e-6, t=0.0, threads=1):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_burgers2d_rhs(n, eps, u, float(t), du, threads))
    return du


def burgers2d_integrate(n, u, t0, dt, nsteps, eps=1e-6, threads=1):
    u = _state(u).copy()
    _ok(lib().orc_burgers2d_integrate(n, eps, u, t0, dt, nsteps, threads))
    return u


def weno1d_dirichlet_rhs(n, L, c, eps, gl, gh, u, t=0.0):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_weno1d_dirichlet_rhs(n, L, c, eps, gl, gh, u, float(t), du))
    return du


def diffreact2d_rhs(n, D, k, bctype, u, t=0.0):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_diffreact2d_rhs(n, D, k, np.ascontiguousarray(bctype, dtype=np.int32), u, float(t), du))
    return du


def nonlindiff3d_rhs(n, d0, src, u, t=0.0, p=2, threads=1):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_nonlindiff3d_rhs(n, p, _f64(d0), _f64(src), u, float(t), du, threads))
    return du


def nonlindiff3d_jvp(n, d0, vdir, src, u0, target, dt, nsteps, p=2):
    L, dL = C.c_double(), C.c_double()
    _ok(lib().orc_nonlindiff3d_jvp(n, p, _f64(d0), _f64(vdir), _f64(src), _f64(u0), _f64(target), dt, nsteps,
                                   C.byref(L), C.byref(dL)))
    return L.value, dL.value


def upwind1d_rhs(n, uo, c0, c1, gl, gh, u, t=0.0):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_upwind1d_rhs(n, uo, c0, c1, gl, gh, u, float(t), du))
    return du


def mixed2d_rhs(n, c, u, p=2, t=0.0):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_mixed2d_rhs(n, p, c, u, float(t), du))
    return du


def mixed2d_grid_rhs(grids, c, u, p=2, q=(0.0, 1.0, 2.0, 0.0, 0.0, 0.0), t=0.0):
    """u_t = u_xx + u_yy + c u_xy on [0,1]^2 (grids[a]: a node count or the nodes), Dirichlet
    u = q0 + q1 x + q2 y + q3 x^2 + q4 x y + q5 y^2 on all sides."""
    keep = [None if np.isscalar(g) else _f64(g) for g in grids]
    n = np.array([int(g) if np.isscalar(g) else len(g) for g in grids], dtype=np.int32)
    xs = (C.c_void_p * 2)(*[(k_.ctypes.data if k_ is not None else None) for k_ in keep])
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_mixed2d_grid_rhs(n, xs, int(p), float(c), _f64(q), u, float(t), du))
    return du


def interface1d_rhs(n1, n2, dx, D1, D2, u, p=2, gl=0.0, gr=0.0, t=0.0):
    """c1 on n1 nodes and c2 on n2 nodes of one spacing dx, joined at c1's last / c2's first node
    by an interface condition; Dt c_k = D_k Dxx c_k, Dirichlet gl and gr at the outer ends.
    State: c1 at nodes 2..n1, then c2 at nodes 2..n2-1."""
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_interface1d_rhs(int(n1), int(n2), float(dx), int(p), float(D1), float(D2), float(gl), float(gr),
                                  u, float(t), du))
    return du


def spherical1d_rhs(n, g1, u, p=2, t=0.0):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_spherical1d_rhs(n, p, g1, u, float(t), du))
    return du


def nonlinlap1d_rhs(n, u, periodic, gl=0.0, gh=0.0, p=2, t=0.0):
    u = _state(u)
    du = np.empty_like(u)
    _ok(lib().orc_nonlinlap1d_rhs(n, p, int(periodic), gl, gh, u, float(t), du))
    return du


def num_threads():
    return lib().orc_num_threads()
