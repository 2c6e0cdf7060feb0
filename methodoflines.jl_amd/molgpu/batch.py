"""Batches of PDE instances advanced together on the GPU.

The reference discretizes and solves one problem at a time (an ensemble of MOL problems is a
loop of ``discretize``/``solve``).  A ``BatchProblem`` takes many lowered instances — uneven
grids, coefficients and boundary kinds — and runs their right-hand sides in one launch per
Runge–Kutta stage (a work-group per tile of an instance, largest tiles first): instances of one
structure share generated code, per-instance numbers are read from instance headers
(``mol_create_batched`` in include/mol.h).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import native
from .discretization import Lowered
from .problem import RK4, ODEProblem, _torch


def _descs(lows: Sequence[Lowered]):
    keep = []
    ds = (C.POINTER(native.mol_desc) * len(lows))()
    ps = (C.POINTER(C.c_double) * len(lows))()
    for i, low in enumerate(lows):
        d, k = native.pack_descriptor(low)
        keep += [d, k]
        ds[i] = C.pointer(d)
        pv = np.ascontiguousarray(np.asarray(low.param_values, dtype=np.float64).reshape(-1))
        if pv.size == 0:
            pv = np.zeros(1)
        keep.append(pv)
        ps[i] = pv.ctypes.data_as(C.POINTER(C.c_double))
    return ds, ps, keep


def codegen_batched(lows: Sequence[Lowered], compile: bool = False, tile=None) -> str:
    """Source of the kernel of a batch (no GPU needed)."""
    lib = native.load_library()
    ds, ps, keep = _descs(lows)
    tl = (C.c_int32 * 3)(*tile) if tile is not None else None
    n = lib.mol_codegen_batched(ds, ps, len(lows), tl, None, 0, 0)
    native.check(lib, n if n < 0 else 0, None)
    buf = C.create_string_buffer(int(n) + 1)
    n = lib.mol_codegen_batched(ds, ps, len(lows), tl, buf, len(buf), 1 if compile else 0)
    native.check(lib, n if n < 0 else 0, None)
    return buf.value.decode()


class BatchProblem:
    """Many instances in one state vector (instance-major, each in its own unknown order)."""

    def __init__(self, instances: Sequence, tile: Optional[Sequence[int]] = None):
        self.lowered: List[Lowered] = [p.lowered if isinstance(p, ODEProblem) else p for p in instances]
        if not self.lowered:
            raise ValueError("empty batch")
        self.lib = native.load_library()
        ds, ps, keep = _descs(self.lowered)
        tl = (C.c_int32 * 3)(*tile) if tile is not None else None
        h = C.c_void_p()
        rc = self.lib.mol_create_batched(ds, ps, len(self.lowered), tl, C.byref(h))
        native.check(self.lib, rc, None)
        self.h = h
        self.n_unknowns = int(self.lib.mol_num_unknowns(h))
        offs = (C.c_int64 * len(self.lowered))()
        self.lib.mol_batch_offsets(h, offs, len(self.lowered))
        self.offsets = [int(o) for o in offs] + [self.n_unknowns]
        self.dtype = self.lowered[0].dtype

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and self.h.value:
                self.lib.mol_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def torch_dtype(self):
        torch = _torch()
        return torch.float64 if self.dtype == "f64" else torch.float32

    @property
    def u0(self) -> np.ndarray:
        return np.concatenate([low.u0 for low in self.lowered])

    def plan(self) -> dict:
        info = (C.c_int32 * 8)()
        self.lib.mol_get_plan(self.h, info, 8)
        return {"tile": (info[0], info[1], info[2]), "work_items": info[3], "classes": info[7]}

    def split(self, u) -> List[np.ndarray]:
        u = np.asarray(u)
        return [u[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.lowered))]

    def _check(self, x):
        if x.dtype != self.torch_dtype() or not x.is_cuda or not x.is_contiguous() or x.numel() != self.n_unknowns:
            raise ValueError("batch state must be a contiguous CUDA tensor of %d %s values"
                             % (self.n_unknowns, self.torch_dtype()))

    def f(self, du, u, t: float) -> None:
        torch = _torch()
        self._check(u)
        self._check(du)
        rc = self.lib.mol_rhs(self.h, du.data_ptr(), u.data_ptr(), None, float(t),
                              C.c_void_p(torch.cuda.current_stream().cuda_stream))
        native.check(self.lib, rc, self.h)

    def integrate(self, u, t0: float, dt: float, nsteps: int, alg=RK4, kahan: bool = False,
                  graph: bool = False) -> None:
        torch = _torch()
        self._check(u)
        flags = (native.MOL_FLAG_KAHAN if kahan else 0) | (native.MOL_FLAG_GRAPH if graph else 0)
        rc = self.lib.mol_integrate(self.h, u.data_ptr(), None, float(t0), float(dt), int(nsteps), alg.scheme,
                                    flags, C.c_void_p(torch.cuda.current_stream().cuda_stream))
        native.check(self.lib, rc, self.h)
