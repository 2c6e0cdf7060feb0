"""``discretize`` / ``ODEProblem`` / ``solve``: the drop-in surface of the GPU path.

``discretize(pdesys, MOLFiniteDifference(...))`` returns an ``ODEProblem`` whose ``f`` is
the HIP right-hand side (the reference returns an ODEProblem whose f is ModelingToolkit's
generated function, src/discretization/staggered_discretize.jl:19-23 shows the shape).
``solve(prob, RK4() | SSPRK33(), dt=...)`` runs the device-resident fixed-step stage loop.
Device memory and streams come from PyTorch; the arithmetic is the HIP library's.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import native
from .discretization import Lowered, MOLFiniteDifference, lower
from .pdesys import PDESystem


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise native.NativeLibraryError("no GPU: the MI355X path needs a HIP device")
    return torch


class RK4:
    """Classic 4-stage Runge-Kutta (OrdinaryDiffEq RK4, fixed step)."""
    scheme = native.MOL_RK4
    stages = 4


class SSPRK33:
    """Shu-Osher 3-stage SSP Runge-Kutta (OrdinaryDiffEq SSPRK33, fixed step)."""
    scheme = native.MOL_SSPRK33
    stages = 3


class GPUHandle:
    """Owns a ``mol_handle*`` for a lowered system."""

    def __init__(self, low: Lowered):
        self.lib = native.load_library()
        self.desc, self._keep = native.pack_descriptor(low)
        h = C.c_void_p()
        rc = self.lib.mol_create(C.byref(self.desc), C.byref(h))
        native.check(self.lib, rc, None)
        self.h = h
        self.n_unknowns = int(self.lib.mol_num_unknowns(h))

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and self.h.value:
                self.lib.mol_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class ODEProblem:
    """Semi-discrete problem du/dt = f(u, p, t) with f evaluated on the GPU."""

    def __init__(self, low: Lowered, tspan=None, p=None):
        self.lowered = low
        self.u0 = low.u0.copy()
        self.tspan = tuple(tspan) if tspan is not None else low.tspan
        self.p = np.asarray(low.param_values if p is None else p, dtype=np.float64)
        self._handle: Optional[GPUHandle] = None
        self._aux: Dict[str, object] = {}
        self._aux_host: Dict[str, np.ndarray] = {}

    # ---------------------------------------------------------------- device resources
    @property
    def dtype_np(self):
        return np.float64 if self.lowered.dtype == "f64" else np.float32

    def torch_dtype(self):
        torch = _torch()
        return torch.float64 if self.lowered.dtype == "f64" else torch.float32

    @property
    def handle(self) -> GPUHandle:
        if self._handle is None:
            self._handle = GPUHandle(self.lowered)
            if self.lowered.aux:
                missing = [a for a in self.lowered.aux if a not in self._aux_host]
                if missing:
                    raise ValueError("grid fields %s have no values: call set_grid_field" % missing)
                self._bind_aux()
        return self._handle

    def set_grid_field(self, name: str, values) -> None:
        """Bind the full-grid values (shape = grid shape, first axis fastest in memory) of a
        grid-valued parameter.  A contiguous CUDA tensor of the state dtype holding the grid in
        the library's layout (first axis fastest, flat) is bound in place, without a copy."""
        if name not in self.lowered.aux:
            raise KeyError(name)
        if not isinstance(values, np.ndarray) and hasattr(values, "is_cuda") and values.is_cuda:
            torch = _torch()
            if values.dtype != self.torch_dtype() or not values.is_contiguous() or \
                    values.numel() != int(np.prod(self.lowered.n)):
                raise ValueError("grid-field tensor must be contiguous %s with %d values"
                                 % (self.torch_dtype(), int(np.prod(self.lowered.n))))
            self._aux_dev = getattr(self, "_aux_dev", {})
            self._aux_dev[name] = values
            self._aux_host[name] = None
            if self._handle is not None:
                self._bind_aux()
            return
        v = np.asarray(values, dtype=np.float64)
        if v.shape != tuple(self.lowered.n) and v.size != int(np.prod(self.lowered.n)):
            raise ValueError("grid field must have the grid shape %s" % (tuple(self.lowered.n),))
        self._aux_host[name] = v.reshape(tuple(self.lowered.n))
        if self._handle is not None:
            self._bind_aux()

    def _bind_aux(self):
        torch = _torch()
        ptrs = (C.c_void_p * max(1, len(self.lowered.aux)))()
        dev = getattr(self, "_aux_dev", {})
        for i, name in enumerate(self.lowered.aux):
            if self._aux_host.get(name) is None and name in dev:
                t = dev[name]
            else:
                host = self._aux_host[name].reshape(-1, order="F")
                t = torch.tensor(host, dtype=self.torch_dtype(), device="cuda")
            self._aux[name] = t
            ptrs[i] = t.data_ptr()
        native.check(self._handle.lib, self._handle.lib.mol_set_aux(self._handle.h, ptrs), self._handle.h)

    def set_tile(self, tx: int, ty: int = 1, tz: int = 1) -> None:
        """Tile shape (nodes per work-group) of the stencil kernels; (0, 0, 0) restores the default."""
        h = self.handle
        native.check(h.lib, h.lib.mol_set_tile(h.h, tx, ty, tz), h.h)

    def plan(self) -> dict:
        h = self.handle
        info = (C.c_int32 * 10)()
        h.lib.mol_get_plan(h.h, info, 10)
        return {"tile": (info[0], info[1], info[2]), "core_tiles": info[3], "generic_tiles": info[4],
                "weno_tiles": info[5], "generated": info[6], "interp_core_ok": info[7], "fused_tiles": info[8],
                "fused_steps_per_launch": info[9]}

    def aux_tensor(self, name: str):
        self.handle
        return self._aux[name]

    def device_params(self):
        """Device copy of the problem's scalar parameters (cached while they are unchanged)."""
        torch = _torch()
        if len(self.p) == 0:
            return None
        key = (self.p.tobytes(), self.lowered.dtype)
        if getattr(self, "_p_key", None) != key:
            self._p_dev = torch.tensor(self.p, dtype=self.torch_dtype(), device="cuda")
            self._p_key = key
        return self._p_dev

    # ---------------------------------------------------------------- f(du, u, p, t)
    def f(self, du, u, p, t: float) -> None:
        """In-place right-hand side.  ``u``/``du``: CUDA tensors (asynchronous on torch's current
        stream) or numpy arrays (host path: copy in, run, copy out)."""
        h = self.handle
        if isinstance(u, np.ndarray):
            uu = np.ascontiguousarray(u, dtype=self.dtype_np)
            out = np.empty_like(uu)
            pp = np.ascontiguousarray(np.asarray(p if p is not None else self.p), dtype=self.dtype_np)
            rc = h.lib.mol_rhs_h(h.h, out.ctypes.data, uu.ctypes.data,
                                 pp.ctypes.data if pp.size else None, float(t))
            native.check(h.lib, rc, h.h)
            du[...] = out
            return
        torch = _torch()
        if p is None:
            p = self.device_params()
        elif isinstance(p, np.ndarray):
            p = torch.tensor(p, dtype=self.torch_dtype(), device="cuda")
        self._check_tensor(u)
        self._check_tensor(du)
        stream = torch.cuda.current_stream().cuda_stream
        rc = h.lib.mol_rhs(h.h, du.data_ptr(), u.data_ptr(), p.data_ptr() if p is not None else None,
                           float(t), C.c_void_p(stream))
        native.check(h.lib, rc, h.h)

    def expand(self, u, t: float, p=None):
        """Full grids of every variable from the state ``u`` (CUDA tensor) at time ``t``: a tensor
        (nfields, n_1, ..., n_d) indexed [f, i_1, ..., i_d] (0-based node indices) holding the
        unknowns, the observed boundary values of the boundary equations, periodic aliases and 0
        at corners, as the reference's solution arrays (PDETimeSeriesSolution,
        src/interface/solution/timedep.jl:19-93)."""
        torch = _torch()
        h = self.handle
        self._check_tensor(u)
        if p is None:
            p = self.device_params()
        elif isinstance(p, np.ndarray):
            p = torch.tensor(p, dtype=self.torch_dtype(), device="cuda")
        low = self.lowered
        shape = [int(v) for v in low.n[:low.ndim]]
        full = torch.empty([low.nfields] + shape[::-1], dtype=self.torch_dtype(), device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        rc = h.lib.mol_expand(h.h, full.data_ptr(), u.data_ptr(), p.data_ptr() if p is not None else None,
                              float(t), C.c_void_p(stream))
        native.check(h.lib, rc, h.h)
        # the device layout has the first axis fastest
        return full.permute([0] + list(range(low.ndim, 0, -1)))

    def _check_tensor(self, x):
        if x.dtype != self.torch_dtype() or not x.is_cuda or not x.is_contiguous():
            raise ValueError("state tensors must be contiguous CUDA tensors of %s" % self.torch_dtype())
        if x.numel() != self.lowered.n_unknowns:
            raise ValueError("state has %d entries, the system has %d unknowns" % (x.numel(), self.lowered.n_unknowns))

    def __call__(self, u, p, t):
        """Out-of-place f(u, p, t)."""
        if isinstance(u, np.ndarray):
            du = np.empty_like(u)
        else:
            du = u.new_empty(u.shape)
        self.f(du, u, p, t)
        return du

    # ---------------------------------------------------------------- integration
    def integrate(self, u, t0: float, dt: float, nsteps: int, alg=RK4, kahan: bool = False,
                  graph: bool = False, p=None, fused: bool = True) -> None:
        """Advance the device state ``u`` in place by ``nsteps`` fixed steps.  ``fused=False``
        runs RK4 as four stage launches per step even where the fused step kernel exists."""
        torch = _torch()
        h = self.handle
        self._check_tensor(u)
        if p is None:
            p = self.device_params()
        flags = (native.MOL_FLAG_KAHAN if kahan else 0) | (native.MOL_FLAG_GRAPH if graph else 0) | \
            (0 if fused else native.MOL_FLAG_UNFUSED)
        stream = torch.cuda.current_stream().cuda_stream
        rc = h.lib.mol_integrate(h.h, u.data_ptr(), p.data_ptr() if p is not None else None, float(t0),
                                 float(dt), int(nsteps), alg.scheme, flags, C.c_void_p(stream))
        native.check(h.lib, rc, h.h)

    # ---------------------------------------------------------------- reverse mode
    def vjp(self, u, lam, t: float = 0.0, p=None, g_aux=None):
        """Vector-Jacobian product of the right-hand side: returns (df/du)^T lam at (u, t).
        With ``g_aux`` (float64 CUDA tensor of naux * prod(n) values) the gradients with respect
        to the grid fields, (df/daux)^T lam, are added to it."""
        torch = _torch()
        h = self.handle
        self._check_tensor(u)
        self._check_tensor(lam)
        if p is None:
            p = self.device_params()
        out = torch.empty_like(lam)
        gptr = self._check_gaux(g_aux)
        stream = torch.cuda.current_stream().cuda_stream
        rc = h.lib.mol_vjp(h.h, u.data_ptr(), p.data_ptr() if p is not None else None, float(t), lam.data_ptr(),
                           out.data_ptr(), gptr, C.c_void_p(stream))
        native.check(h.lib, rc, h.h)
        return out

    def _check_gaux(self, g_aux):
        if g_aux is None:
            return None
        torch = _torch()
        ng = int(np.prod(self.lowered.n))
        if g_aux.dtype != torch.float64 or not g_aux.is_cuda or not g_aux.is_contiguous() or \
                g_aux.numel() != ng * len(self.lowered.aux):
            raise ValueError("g_aux must be a contiguous float64 CUDA tensor of %d values"
                             % (ng * len(self.lowered.aux)))
        return C.c_void_p(g_aux.data_ptr())

    def checkpoint_interval(self, nsteps: int, budget: float = 0.5) -> int:
        """``ckpt_every`` for a run of ``nsteps`` RK4 steps: 0 (keep every stage state, no
        recomputation in the adjoint) when 4 * nsteps state vectors fit in ``budget`` of the free
        device memory, else about sqrt(nsteps) steps between checkpoints."""
        torch = _torch()
        free, _ = torch.cuda.mem_get_info()
        need = 4 * int(nsteps) * self.lowered.n_unknowns * (8 if self.lowered.dtype == "f64" else 4)
        if need <= budget * free:
            return 0
        return max(1, int(round(float(nsteps) ** 0.5)))

    def forward_checkpointed(self, u, t0: float, dt: float, nsteps: int, ckpt_every: Optional[int] = None,
                             p=None) -> None:
        """RK4 from t0 (u advanced in place) keeping a checkpoint every ``ckpt_every`` steps for
        a later :meth:`adjoint`; ``ckpt_every=0`` keeps every stage state instead (the adjoint
        recomputes nothing), ``None`` picks with :meth:`checkpoint_interval`."""
        torch = _torch()
        h = self.handle
        self._check_tensor(u)
        if ckpt_every is None:
            ckpt_every = self.checkpoint_interval(nsteps)
        if p is None:
            p = self.device_params()
        stream = torch.cuda.current_stream().cuda_stream
        rc = h.lib.mol_forward_checkpointed(h.h, u.data_ptr(), p.data_ptr() if p is not None else None, float(t0),
                                            float(dt), int(nsteps), int(ckpt_every), C.c_void_p(stream))
        native.check(h.lib, rc, h.h)

    def adjoint(self, lam, g_aux=None, p=None) -> None:
        """Discrete adjoint of the last checkpointed RK4 run: on entry ``lam`` = dL/du(T), on
        exit dL/du(t0); grid-field gradients are added to ``g_aux`` (float64)."""
        torch = _torch()
        h = self.handle
        self._check_tensor(lam)
        if p is None:
            p = self.device_params()
        gptr = self._check_gaux(g_aux)
        stream = torch.cuda.current_stream().cuda_stream
        rc = h.lib.mol_adjoint(h.h, p.data_ptr() if p is not None else None, lam.data_ptr(), gptr,
                               C.c_void_p(stream))
        native.check(h.lib, rc, h.h)

    # ---------------------------------------------------------------- error checks
    def error_norms(self, u, ref=None):
        """FP64 norms of the device state against a float64 reference of the same length:
        a float64 CUDA tensor [max |u - ref|, ||u - ref||_2, ||ref||_2] (the norms of u itself
        when ``ref`` is None), computed on the device without a host copy of the state."""
        torch = _torch()
        h = self.handle
        self._check_tensor(u)
        if ref is not None and (ref.dtype != torch.float64 or not ref.is_cuda or not ref.is_contiguous()
                                or ref.numel() != self.lowered.n_unknowns):
            raise ValueError("ref must be a contiguous float64 CUDA tensor of %d values" % self.lowered.n_unknowns)
        out = torch.empty(3, dtype=torch.float64, device=u.device)
        stream = torch.cuda.current_stream().cuda_stream
        rc = h.lib.mol_error_norms(h.h, u.data_ptr(), ref.data_ptr() if ref is not None else None,
                                   out.data_ptr(), C.c_void_p(stream))
        native.check(h.lib, rc, h.h)
        return out

    # ---------------------------------------------------------------- layout helpers
    def unflatten(self, u) -> List[np.ndarray]:
        """Split a state vector into per-variable grids over their unknown boxes (first axis first)."""
        low = self.lowered
        u = np.asarray(u)
        out = []
        for f, sl in enumerate(low.field_slices()):
            out.append(u[sl].reshape(low.box_shape(f), order="F"))
        return out


def codegen(low: Lowered, compile: bool = False, tile=None, vjp: bool = False, fused: bool = False) -> str:
    """Source of the kernel generated for a lowered system (no GPU needed): the right-hand
    side, with ``vjp=True`` its vector-Jacobian product, with ``fused=True`` the fused RK4 step;
    ``compile=True`` also compiles it with hiprtc for gfx950."""
    lib = native.load_library()
    desc, keep = native.pack_descriptor(low)
    tl = (C.c_int32 * 3)(*tile) if tile is not None else None
    what = 2 if vjp else (4 if fused else 0)
    n = lib.mol_codegen(C.byref(desc), tl, None, 0, what)
    native.check(lib, n if n < 0 else 0, None)
    buf = C.create_string_buffer(int(n) + 1)
    n = lib.mol_codegen(C.byref(desc), tl, buf, len(buf), what | (1 if compile else 0))
    native.check(lib, n if n < 0 else 0, None)
    return buf.value.decode()


def symbolic_discretize(pdesys: PDESystem, disc: MOLFiniteDifference) -> Lowered:
    """Lower the PDE system to the stencil IR (no device work)."""
    return lower(pdesys, disc)


def discretize(pdesys: PDESystem, disc: MOLFiniteDifference) -> ODEProblem:
    """``discretize(pdesys, MOLFiniteDifference(...))`` -> ``ODEProblem``."""
    low = lower(pdesys, disc)
    prob = ODEProblem(low)
    for h, name in zip(low.aux_heads, low.aux):
        v = pdesys.ps.get(h)
        if v is not None:
            prob.set_grid_field(name, v)
    return prob


@dataclass
class ODESolution:
    """Saved states of a fixed-step solve, indexed as the reference's PDETimeSeriesSolution
    (src/interface/solution/timedep.jl:19-93): ``sol[u]`` is the full-grid trajectory of a
    variable, shape (nt, n_1, ..., n_d), boundary nodes included; ``sol[x]`` the grid of an
    axis; ``sol.t`` the save times.  ``sol.u`` keeps the state vectors (unknowns only)."""
    t: List[float]
    u: List[np.ndarray]
    prob: ODEProblem
    full: List[np.ndarray] = None

    def _field(self, var) -> int:
        low = self.prob.lowered
        name = var if isinstance(var, str) else str(getattr(var, "func", var))
        return low.fields.index(name)

    def __getitem__(self, var):
        low = self.prob.lowered
        if var in low.axes:
            return np.asarray(low.coords[low.axes.index(var)])
        f = self._field(var)
        return np.stack([g[f] for g in self.full])

    def interior(self, var):
        """Per-variable trajectory over the unknown box: array (nt, *box_shape)."""
        f = self._field(var)
        return np.stack([self.prob.unflatten(u)[f] for u in self.u])

    def __call__(self, t, *x, dv=None):
        """Value of variable ``dv`` (default: the first) at time t and position x by multilinear
        interpolation over the save times and the grid, as the reference's solution interpolates
        (Gridded(Linear()), src/interface/solution/solution_utils.jl:3-32).  Arguments broadcast
        against each other."""
        from scipy.interpolate import RegularGridInterpolator
        low = self.prob.lowered
        if len(x) != low.ndim:
            raise ValueError("expected %d spatial coordinates" % low.ndim)
        f = 0 if dv is None else self._field(dv)
        if len(self.t) < 2:
            raise ValueError("interpolation needs at least two saved times")
        key = ("interp", f)
        cache = self.__dict__.setdefault("_interp", {})
        if key not in cache:
            grid = (np.asarray(self.t, dtype=np.float64),) + tuple(np.asarray(c, dtype=np.float64)
                                                                   for c in low.coords[:low.ndim])
            cache[key] = RegularGridInterpolator(grid, np.stack([g[f] for g in self.full]), method="linear")
        pts = np.broadcast_arrays(np.asarray(t, dtype=np.float64), *[np.asarray(v, dtype=np.float64) for v in x])
        out = cache[key](np.stack([p_.reshape(-1) for p_ in pts], axis=-1))
        return out.reshape(pts[0].shape) if pts[0].ndim else float(out[0])


def solve(prob: ODEProblem, alg=RK4, dt: float = None, adaptive: bool = False, saveat=None,
          kahan: bool = False, graph: bool = False, tspan=None) -> ODESolution:
    """Fixed-step solve on the GPU (adaptive stepping is not supported)."""
    if adaptive:
        raise NotImplementedError("adaptive time stepping is not supported; pass dt and adaptive=False")
    if dt is None:
        raise ValueError("dt is required")
    torch = _torch()
    t0, t1 = tspan if tspan is not None else prob.tspan
    nsteps_total = int(round((t1 - t0) / dt))
    u = torch.tensor(prob.u0, dtype=prob.torch_dtype(), device="cuda")
    ts, us = [t0], [prob.u0.astype(prob.dtype_np).copy()]
    fulls = [prob.expand(u, t0).cpu().numpy().astype(np.float64)]
    if saveat is None:
        marks = [nsteps_total]
    else:
        every = int(round(saveat / dt))
        marks = list(range(every, nsteps_total + 1, every))
        if not marks or marks[-1] != nsteps_total:
            marks.append(nsteps_total)
    done = 0
    for m in marks:
        prob.integrate(u, t0 + done * dt, dt, m - done, alg, kahan=kahan, graph=graph)
        done = m
        ts.append(t0 + done * dt)
        us.append(u.detach().cpu().numpy().astype(np.float64))
        fulls.append(prob.expand(u, ts[-1]).cpu().numpy().astype(np.float64))
    return ODESolution(ts, us, prob, fulls)
