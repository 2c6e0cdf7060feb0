"""ctypes binding of the C ABI in include/mol.h and packing of the stencil IR into ``mol_desc``.

The shared library ``libmolgpu.so`` is built in-tree (see build.py).  There is no CPU
fallback: every entry point raises ``NativeLibraryError`` when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List

import numpy as np

MOL_ABI_VERSION = 3
MOL_MAX_DIMS = 3
MOL_MAX_FIELDS = 8
MOL_MAX_AUX = 8
MOL_F64, MOL_F32 = 0, 1
MOL_RK4, MOL_SSPRK33 = 0, 1
MOL_FLAG_KAHAN, MOL_FLAG_GRAPH, MOL_FLAG_UNFUSED = 1, 2, 4

LIB_NAME = "libmolgpu.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)


class NativeLibraryError(RuntimeError):
    pass


class MolError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("mol error %d: %s" % (code, msg))
        self.code = code


class mol_optab(C.Structure):
    _fields_ = [(n, C.c_int32) for n in
                ("len", "off", "w", "nlow", "nhigh", "blen", "wlow", "whigh", "hoff", "ref_shift", "pn_lo", "pn_hi")]


class mol_slot(C.Structure):
    _fields_ = [("kind", C.c_int32), ("field", C.c_int32), ("axis", C.c_int32), ("axis2", C.c_int32),
                ("tab", C.c_int32 * 5), ("prog", C.c_int32), ("eps", C.c_double), ("inv_dx", C.c_double)]


class mol_closure(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("field", "axis", "side", "nghost", "K", "wC", "wc", "gprog", "src")]


class mol_desc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("ndim", C.c_int32), ("nfields", C.c_int32), ("naux", C.c_int32),
        ("dtype", C.c_int32), ("nparams", C.c_int32),
        ("n", C.c_int64 * MOL_MAX_DIMS),
        ("periodic", C.c_int32 * MOL_MAX_DIMS),
        ("coord_h", C.POINTER(C.c_double) * MOL_MAX_DIMS),
        ("lo", (C.c_int64 * MOL_MAX_DIMS) * MOL_MAX_FIELDS),
        ("hi", (C.c_int64 * MOL_MAX_DIMS) * MOL_MAX_FIELDS),
        ("halo", C.c_int32 * MOL_MAX_DIMS),
        ("nslots", C.c_int32), ("slots_h", C.POINTER(mol_slot)),
        ("ntabs", C.c_int32), ("tabs_h", C.POINTER(mol_optab)),
        ("nclosures", C.c_int32), ("closures_h", C.POINTER(mol_closure)),
        ("npool", C.c_int32), ("pool_h", C.POINTER(C.c_double)),
        ("ncode", C.c_int32), ("code_h", C.POINTER(C.c_uint32)),
        ("ncpool", C.c_int32), ("cpool_h", C.POINTER(C.c_double)),
        ("nprogs", C.c_int32), ("prog_h", C.POINTER(C.c_int32)),
        ("rhs_prog", C.c_int32 * MOL_MAX_FIELDS),
    ]


# exported symbols and their signatures (kept in step with include/mol.h)
_SIGNATURES = {
    "mol_create": (C.c_int, [C.POINTER(mol_desc), C.POINTER(C.c_void_p)]),
    "mol_rhs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "mol_rhs_h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    "mol_expand": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "mol_set_aux": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "mol_integrate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int64,
                                C.c_int32, C.c_int32, C.c_void_p]),
    "mol_integrate_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int64,
                                        C.c_int32, C.c_int32, C.c_void_p]),
    "mol_num_unknowns": (C.c_int64, [C.c_void_p]),
    "mol_error_norms": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mol_set_tile": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "mol_get_plan": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32]),
    "mol_vjp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                          C.c_void_p, C.c_void_p]),
    "mol_forward_checkpointed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                           C.c_int64, C.c_int32, C.c_void_p]),
    "mol_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mol_last_error": (C.c_char_p, [C.c_void_p]),
    "mol_destroy": (None, [C.c_void_p]),
    "mol_device_info": (C.c_int, [C.c_char_p, C.c_int]),
    "mol_create_batched": (C.c_int, [C.POINTER(C.POINTER(mol_desc)), C.POINTER(C.POINTER(C.c_double)), C.c_int32,
                                     C.POINTER(C.c_int32), C.POINTER(C.c_void_p)]),
    "mol_batch_offsets": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int32]),
    "mol_codegen_batched": (C.c_int64, [C.POINTER(C.POINTER(mol_desc)), C.POINTER(C.POINTER(C.c_double)), C.c_int32,
                                        C.POINTER(C.c_int32), C.c_char_p, C.c_int64, C.c_int32]),
    "mol_codegen": (C.c_int64, [C.POINTER(mol_desc), C.POINTER(C.c_int32), C.c_char_p, C.c_int64, C.c_int32]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libmolgpu.so (raises NativeLibraryError when it has not been built)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise NativeLibraryError(
            "%s not found: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')" % path)
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def check(lib, rc: int, handle=None):
    if rc != 0:
        msg = lib.mol_last_error(handle)
        raise MolError(rc, msg.decode() if msg else "")


def _arr(a, ctype):
    a = np.ascontiguousarray(a)
    return a, a.ctypes.data_as(C.POINTER(ctype))


def weight_pool(low) -> np.ndarray:
    """The weight pool of ``pack_descriptor(low)`` (tables, then closures), as an array."""
    d, keep = pack_descriptor(low)
    return np.ctypeslib.as_array(d.pool_h, shape=(max(1, d.npool),)).copy()


def pack_descriptor(low) -> tuple:
    """Build a ``mol_desc`` for a lowered system.  Returns (desc, keepalive)."""
    keep: List = []
    d = mol_desc()
    d.abi_version = MOL_ABI_VERSION
    d.ndim = low.ndim
    d.nfields = low.nfields
    d.naux = len(low.aux)
    d.dtype = MOL_F64 if low.dtype == "f64" else MOL_F32
    d.nparams = len(low.params)
    for a in range(low.ndim):
        d.n[a] = low.n[a]
        d.periodic[a] = int(low.periodic[a])
        c, ptr = _arr(np.asarray(low.coords[a], dtype=np.float64), C.c_double)
        keep.append(c)
        d.coord_h[a] = ptr
        d.halo[a] = low.halo[a]
    for f in range(low.nfields):
        for a in range(low.ndim):
            d.lo[f][a] = low.lo[f][a]
            d.hi[f][a] = low.hi[f][a]
    pool: List[float] = []
    tabs = (mol_optab * max(1, len(low.tables)))()
    for i, tb in enumerate(low.tables):
        op = tb.op
        t = tabs[i]
        t.len = op.stencil_length
        t.off = op.tap_offset
        t.w = len(pool)
        t.pn_lo, t.pn_hi = op.pn_lo, op.pn_hi
        pool.extend(float(v) for v in np.asarray(op.stencil_coefs, dtype=np.float64).reshape(-1))
        t.blen = op.boundary_stencil_length
        t.nlow = op.nlow
        t.wlow = len(pool)
        for row in op.low_boundary_coefs[:op.nlow]:
            pool.extend(float(v) for v in row)
        t.nhigh = op.nhigh
        t.whigh = len(pool)
        for row in op.high_boundary_coefs[:op.nhigh]:
            pool.extend(float(v) for v in row)
        t.hoff = op.hoff
        t.ref_shift = tb.ref_shift
        if len(op.low_boundary_coefs) < op.nlow or len(op.high_boundary_coefs) < op.nhigh:
            raise ValueError("operator table has fewer boundary rows than it selects")
    slots = (mol_slot * max(1, len(low.slots)))()
    for i, s in enumerate(low.slots):
        sl = slots[i]
        sl.kind, sl.field, sl.axis, sl.axis2 = s["kind"], s["field"], s["axis"], s["axis2"]
        for q, tb in enumerate(s["tab"]):
            sl.tab[q] = tb
        sl.prog = s["prog"]
        sl.eps = s["eps"]
        sl.inv_dx = s["inv_dx"]
    clos = (mol_closure * max(1, len(low.closures)))()
    for i, cl in enumerate(low.closures):
        c = clos[i]
        c.field, c.axis, c.side = cl.field, cl.axis, cl.side
        c.nghost, c.K = cl.C.shape[0], cl.C.shape[1]
        c.wC = len(pool)
        pool.extend(float(v) for v in cl.C.reshape(-1))
        c.wc = len(pool)
        pool.extend(float(v) for v in cl.c)
        c.gprog = cl.gprog
        c.src = getattr(cl, "src", -1)
    code: List[int] = []
    progs: List[int] = []
    for pr in low.programs:
        progs.extend([len(code), len(pr)])
        code.extend(pr)
    pool_a, pool_p = _arr(np.asarray(pool if pool else [0.0], dtype=np.float64), C.c_double)
    code_a, code_p = _arr(np.asarray(code if code else [0], dtype=np.uint32), C.c_uint32)
    cp_a, cp_p = _arr(np.asarray(low.cpool if low.cpool else [0.0], dtype=np.float64), C.c_double)
    pr_a, pr_p = _arr(np.asarray(progs if progs else [0, 0], dtype=np.int32), C.c_int32)
    keep += [tabs, slots, clos, pool_a, code_a, cp_a, pr_a]
    d.nslots, d.slots_h = len(low.slots), C.cast(slots, C.POINTER(mol_slot))
    d.ntabs, d.tabs_h = len(low.tables), C.cast(tabs, C.POINTER(mol_optab))
    d.nclosures, d.closures_h = len(low.closures), C.cast(clos, C.POINTER(mol_closure))
    d.npool, d.pool_h = len(pool), pool_p
    d.ncode, d.code_h = len(code), code_p
    d.ncpool, d.cpool_h = len(low.cpool), cp_p
    d.nprogs, d.prog_h = len(low.programs), pr_p
    for f in range(low.nfields):
        d.rhs_prog[f] = low.rhs_prog[f]
    return d, keep
