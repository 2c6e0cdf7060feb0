"""Compile generated kernel source for the CPU and run it (test infrastructure).

The source that the library generates for the GPU (``molgpu.codegen``) is plain C++ apart from
a few HIP keywords; with the shim in this directory, the thread count set to one and the
kernel called once per work-group, it runs on the host.  CPU tests use this to check the
generated forward and reverse-mode code against the oracle without a GPU.
"""
import ctypes as C
import hashlib
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PRELUDE = os.path.join(ROOT, "methodoflines.jl_amd", "csrc", "mol_jit_prelude.h")
BUILD = os.path.join(tempfile.gettempdir(), "molgpu_genhost")


def _host_source(src: str, entry: str) -> str:
    prelude = open(PRELUDE).read()
    shim = open(os.path.join(HERE, "shim.h")).read()
    src = src.replace('#include "mol_jit_prelude.h"', shim + "\n" + prelude)
    src = re.sub(r"static constexpr int NTHR = \d+;", "static constexpr int NTHR = 1;", src)
    # rows of the nonuniform WENO5 coefficient tables, as the library builds them on the device
    wtab = """
extern "C" __attribute__((visibility("default"))) void mol_host_wtab_row(const double* x, int n, int per, int i,
                                                                          double* c) {
  mol_weno_nu_row(x, n, per != 0, i, c);
}
"""
    return src + entry + wtab


class Epi(C.Structure):
    _fields_ = [("mode", C.c_int32), ("subA", C.c_int32), ("out", C.c_void_p), ("A", C.c_void_p), ("B", C.c_void_p),
                ("acc", C.c_void_p), ("comp", C.c_void_p), ("ca", C.c_double), ("cb", C.c_double),
                ("cf", C.c_double), ("acc_keep", C.c_double), ("acc_f", C.c_double), ("tnow", C.c_void_p),
                ("tofs", C.c_double), ("p", C.c_void_p), ("tnext", C.c_void_p), ("tadv", C.c_double)]


class JPtrs(C.Structure):
    _fields_ = [("coord", C.c_void_p * 3), ("aux", C.c_void_p * 8), ("pool", C.c_void_p), ("wtab", C.c_void_p * 3)]

NU_NC = 56   # coefficients per node of a nonuniform WENO5 table row


class VArgs(C.Structure):
    _fields_ = [("u", C.c_void_p), ("s1", C.c_void_p), ("s2", C.c_void_p), ("c1", C.c_double), ("c2", C.c_double),
                ("out", C.c_void_p), ("g_aux", C.c_void_p), ("p", C.c_void_p), ("tnow", C.c_void_p),
                ("tofs", C.c_double), ("bf", C.c_void_p), ("ba", C.c_void_p), ("mode", C.c_int32),
                ("lam", C.c_void_p), ("y0", C.c_void_p), ("y1", C.c_void_p), ("y2", C.c_void_p)]


def _compile(src: str) -> C.CDLL:
    os.makedirs(BUILD, exist_ok=True)
    key = hashlib.sha1((src + "|O1 hidden Bsymbolic").encode()).hexdigest()[:16]
    so = os.path.join(BUILD, "gen_%s.so" % key)
    if not os.path.exists(so):
        cpp = os.path.join(BUILD, "gen_%s.cpp" % key)
        with open(cpp, "w") as f:
            f.write(src)
        # each generated library binds its own (identically named) inline symbols
        r = subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-w", "-fvisibility=hidden",
                            "-fvisibility-inlines-hidden", "-Wl,-Bsymbolic", cpp, "-o", so + ".tmp"],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-4000:])
        os.replace(so + ".tmp", so)
    return C.CDLL(so)


def _nblocks(src: str) -> int:
    return _const(src, "static constexpr int NBLK = ")


def _const(src: str, key: str, start: str = "") -> int:
    i = src.index(start) if start else 0
    j = src.index(key, i) + len(key)
    return int(src[j:src.index(";", j)].split(",")[0])


def _ptrs(low, aux, lib=None):
    from molgpu import native
    from molgpu.discretization import SLOT_WENO_NU
    jp = JPtrs()
    keep = []
    pool = np.ascontiguousarray(native.weight_pool(low), dtype=np.float64)
    keep.append(pool)
    jp.pool = pool.ctypes.data
    for a in range(low.ndim):
        c = np.ascontiguousarray(low.coords[a], dtype=np.float64)
        keep.append(c)
        jp.coord[a] = c.ctypes.data
    for i, name in enumerate(low.aux):
        v = np.ascontiguousarray(np.asarray(aux[name], dtype=np.float64).reshape(-1, order="F"))
        keep.append(v)
        jp.aux[i] = v.ctypes.data
    for a in range(low.ndim):
        if lib is None or not any(s["kind"] == SLOT_WENO_NU and s["axis"] == a for s in low.slots):
            continue
        x = np.ascontiguousarray(low.coords[a], dtype=np.float64)
        n = len(x)
        tab = np.zeros(n * NU_NC)
        f = lib.mol_host_wtab_row
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        for i in range(1, n + 1):
            f(x.ctypes.data, n, int(bool(low.periodic[a])), i, tab.ctypes.data + 8 * NU_NC * (i - 1))
        keep += [x, tab]
        jp.wtab[a] = tab.ctypes.data
    return jp, keep


def rhs(low, u, t=0.0, p=None, aux=None, tile=None, grid=None):
    """Generated forward kernel on the CPU: du = f(u, p, t) (FP64 systems).  ``grid``: number of
    work-groups of the persistent launch (default one per tile)."""
    import molgpu
    src = molgpu.codegen(low, tile=tile)
    entry = """
extern "C" __attribute__((visibility("default"))) void run(const T* u, Epi ep, JPtrs jp, int nblk) {
  gridDim.x = nblk;
  for (int b = 0; b < nblk; ++b) { blockIdx.x = b; mol_rhs_tiles(u, ep, jp); }
}
"""
    lib = _compile(_host_source(src, entry))
    u = np.ascontiguousarray(u, dtype=np.float64)
    du = np.zeros_like(u)
    pv = np.ascontiguousarray(np.asarray(low.param_values if p is None else p, dtype=np.float64).reshape(-1))
    ep = Epi()
    ep.mode = 0
    ep.out = du.ctypes.data
    ep.tofs = t
    ep.p = pv.ctypes.data if pv.size else None
    jp, keep = _ptrs(low, aux or {}, lib)
    lib.run.argtypes = [C.c_void_p, Epi, JPtrs, C.c_int]
    lib.run(u.ctypes.data, ep, jp, grid if grid is not None else _nblocks(src))
    return du


def vjp(low, u, lam, t=0.0, p=None, aux=None, tile=None):
    """Generated reverse-mode kernel on the CPU: (df/du)^T lam and the grid-field gradients."""
    import molgpu
    src = molgpu.codegen(low, tile=tile, vjp=True)
    entry = """
extern "C" __attribute__((visibility("default"))) void run(VArgs va, JPtrs jp, int nblk, long long items) {
  gridDim.x = nblk;
  threadIdx.x = 0;
  for (int b = 0; b < nblk; ++b) { blockIdx.x = b; mol_vjp_tiles(va, jp); }
  for (long long i = 0; i < items; ++i) {
    blockIdx.x = (unsigned)(i / 256);
    threadIdx.x = (unsigned)(i % 256);
    mol_vjp_gather(va);
  }
  threadIdx.x = 0;
}
"""
    lib = _compile(_host_source(src, entry))
    u = np.ascontiguousarray(u, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.zeros_like(u)
    ng = int(np.prod(low.n))
    g = np.zeros(max(1, len(low.aux)) * ng)
    pv = np.ascontiguousarray(np.asarray(low.param_values if p is None else p, dtype=np.float64).reshape(-1))
    va = VArgs()
    va.u = u.ctypes.data
    va.s1 = lam.ctypes.data
    va.c1 = 1.0
    va.out = out.ctypes.data
    va.g_aux = g.ctypes.data
    va.p = pv.ctypes.data if pv.size else None
    va.tofs = t
    jp, keep = _ptrs(low, aux or {}, lib)
    nblk = _nblocks(src)
    volx = _const(src, "constexpr int VOLX = ")
    bf = np.zeros(max(1, nblk * volx * _const(src, "static constexpr int NSTF = ", "struct GeoB")))
    ba = np.zeros(max(1, nblk * volx * _const(src, "static constexpr int NSTA = ", "struct GeoB")))
    va.bf = bf.ctypes.data
    va.ba = ba.ctypes.data
    va.mode = 3
    items = low.n_unknowns + (len(low.aux) * ng if low.aux else 0)
    lib.run.argtypes = [VArgs, JPtrs, C.c_int, C.c_longlong]
    lib.run(va, jp, nblk, items)
    return out, g[: len(low.aux) * ng].reshape(len(low.aux), ng) if low.aux else None


def rk4_fused(low, u, t0, dt, p=None, aux=None, tile=None, nsub=1):
    """Generated fused RK4 step kernel on the CPU: one classic RK4 step from (u, t0), or nsub
    steps when the kernel holds the system in one resident tile."""
    import molgpu
    src = molgpu.codegen(low, tile=tile, fused=True)
    entry = """
extern "C" __attribute__((visibility("default"))) void run(const T* u, T* unew, Epi ep, JPtrs jp, int nblk,
                                                            int nsub) {
  gridDim.x = nblk;
  threadIdx.x = 0;
  for (int b = 0; b < nblk; ++b) { blockIdx.x = b; mol_rk4_fused(u, unew, ep, jp, nsub); }
}
"""
    lib = _compile(_host_source(src, entry))
    u = np.ascontiguousarray(u, dtype=np.float64)
    unew = np.zeros_like(u)
    clock = np.array([t0, 0.0])
    pv = np.ascontiguousarray(np.asarray(low.param_values if p is None else p, dtype=np.float64).reshape(-1))
    ep = Epi()
    ep.tnow = clock.ctypes.data
    ep.tnext = clock.ctypes.data + 8
    ep.tadv = dt
    ep.p = pv.ctypes.data if pv.size else None
    jp, keep = _ptrs(low, aux or {}, lib)
    lib.run.argtypes = [C.c_void_p, C.c_void_p, Epi, JPtrs, C.c_int, C.c_int]
    lib.run(u.ctypes.data, unew.ctypes.data, ep, jp, _nblocks(src), nsub)
    return unew, clock[1]
