// Core-tile path of the right-hand side.
//
// Tiles whose nodes use only interior stencil rows and whose halo holds no ghost node (every
// tile of a periodic problem, the bulk of the others) are evaluated by a kernel whose whole
// stencil IR — bytecode, constants, slot descriptors and interior weights — lives in the
// kernel-argument segment.  Every table read is then a uniform scalar load, the tile is
// staged with periodic wrap only, and slots are applied inline from LDS.
#pragma once
#include "mol_device.h"

namespace mol {

#define CORE_MAX_CODE 192
#define CORE_MAX_CONST 40
#define CORE_MAX_SLOTS 8
#define CORE_MAX_PROGS 12
#define CORE_MAX_POOL 80

struct CoreSlot {
  int32_t kind, field, axis, axis2, prog;
  int32_t len[5], off[5], w[5];
  double eps, inv_dx;
};

template <typename T>
struct KCore {
  int32_t ndim, nfields, naux, nslots;
  int32_t n[3], periodic[3];
  int32_t lo[3], hi[3];            // shared unknown box of all variables
  int32_t tile[3], halo[3], ext[3], vol;
  int32_t ct_lo[3], ct_n[3];       // core tiles: first tile index and count per axis
  int32_t nd_lo[3], nd_hi[3];     // nodes whose rows are all interior and whose halo lies in the box
  uint32_t magic0, magic1;         // floor(q / ext0) = umulhi(q, magic0) (0: divisor 1), same for ext0*ext1
  int64_t box;                     // unknowns per variable
  int32_t rhs_prog[MOL_MAX_FIELDS];
  int32_t prog[2 * CORE_MAX_PROGS];
  uint32_t code[CORE_MAX_CODE];
  double cpool[CORE_MAX_CONST];
  CoreSlot slots[CORE_MAX_SLOTS];
  double pool[CORE_MAX_POOL];
  const double* coord[3];
  const T* aux[MOL_MAX_AUX];
};

// floor(q / d) for q, d < 2^16 with magic = ceil(2^32 / d): the error term q*(magic*d - 2^32)
// stays below 2^32, so the high word of q*magic is exact.  magic == 0 encodes d == 1.
__device__ __forceinline__ uint32_t fdiv_magic(uint32_t q, uint32_t magic) {
  return magic == 0u ? q : __umulhi(q, magic);
}

// Interpreter of the core path: code and constants are read from the kernel arguments.
template <typename T, class Ctx>
__device__ __forceinline__ T core_prog(const KCore<T>& kc, int prog, const Ctx& ctx) {
  return stack_prog<T>(kc, prog, ctx);
}

// Leaf values of the core contexts.  `base` is the LDS index of the node in array 0; moving
// one node along axis d moves the index by st[d].
template <typename T>
struct CoreView {
  const KCore<T>* kc;
  const T* lds;
  int base;
  int st[3];
  int I[3];
  double t;
  const T* p;
  // selections instead of indexed member arrays keep the view in registers
  __device__ __forceinline__ int stride(int d) const { return d == 0 ? st[0] : (d == 1 ? st[1] : st[2]); }
  __device__ __forceinline__ int idx(int d) const { return d == 0 ? I[0] : (d == 1 ? I[1] : I[2]); }
  __device__ __forceinline__ T at(int arr, int d, int off) const {
    return lds[arr * kc->vol + base + off * stride(d)];
  }
  __device__ __forceinline__ T at2(int arr, int d0, int o0, int d1, int o1) const {
    return lds[arr * kc->vol + base + o0 * stride(d0) + o1 * stride(d1)];
  }
  __device__ __forceinline__ T coord_node(int d) const { return (T)kc->coord[d][idx(d) - 1]; }
  __device__ __forceinline__ T coord_tap(int d, int off) const {
    int k = idx(d) + off;
    if (kc->periodic[d]) k = wrap_periodic(k, kc->n[d]);
    return (T)kc->coord[d][k - 1];
  }
  // interior-row stencil: sum_k w[k] * arr[node + (off + k) e_d]
  __device__ __forceinline__ T stencil(int arr, int d, int len, int off, int w) const {
    T s = T(0);
    for (int k = 0; k < len; ++k) s += (T)kc->pool[w + k] * at(arr, d, off + k);
    return s;
  }
};

// Half-point context of the core nonlinear Laplacian: interior interpolation rows only.
template <typename T>
struct CoreHalfCtx {
  const CoreView<T>* v;
  int ax, h;            // half point between node+h and node+h+1 along ax (offset from the node)
  int ilen, ioff, iw;   // interpolation row
  int dlen, doff, dw;   // inner derivative row
  int dfield;
  __device__ __forceinline__ T interp(int arr) const {
    T s = T(0);
    for (int k = 0; k < ilen; ++k) s += (T)v->kc->pool[iw + k] * v->at(arr, ax, h + ioff + k);
    return s;
  }
  __device__ __forceinline__ T field(int f) const { return interp(f); }
  __device__ __forceinline__ T aux(int a) const { return interp(v->kc->nfields + a); }
  __device__ __forceinline__ T coord(int d) const {
    if (d != ax) return v->coord_node(d);
    T s = T(0);
    for (int k = 0; k < ilen; ++k) s += (T)v->kc->pool[iw + k] * v->coord_tap(d, h + ioff + k);
    return s;
  }
  __device__ __forceinline__ T slot(int) const { return T(0); }
  __device__ __forceinline__ T time() const { return (T)v->t; }
  __device__ __forceinline__ T param(int i) const { return v->p[i]; }
  __device__ __forceinline__ T dinner() const {
    T s = T(0);
    for (int k = 0; k < dlen; ++k) s += (T)v->kc->pool[dw + k] * v->at(dfield, ax, h + doff + k);
    return s;
  }
};

template <typename T>
struct CorePlainCtx {
  const CoreView<T>* v;
  __device__ __forceinline__ T field(int f) const { return v->at(f, 0, 0); }
  __device__ __forceinline__ T aux(int a) const { return v->at(v->kc->nfields + a, 0, 0); }
  __device__ __forceinline__ T coord(int d) const { return v->coord_node(d); }
  __device__ __forceinline__ T slot(int) const { return T(0); }
  __device__ __forceinline__ T time() const { return (T)v->t; }
  __device__ __forceinline__ T param(int i) const { return v->p[i]; }
  __device__ __forceinline__ T dinner() const { return T(0); }
};

template <typename T>
__device__ __forceinline__ T core_nonlinlap(const CoreView<T>& v, const CoreSlot& s, int to, int ti, int tp) {
  // outer operator at half-point index i-1: taps h = (i-1) + off + k  ->  offset from node i is
  // h - i = off + k - 1 ; the half point between node i+o and i+o+1 has offset o.
  CoreHalfCtx<T> hc;
  hc.v = &v;
  hc.ax = s.axis;
  hc.ilen = s.len[tp]; hc.ioff = s.off[tp]; hc.iw = s.w[tp];
  hc.dlen = s.len[ti]; hc.doff = s.off[ti]; hc.dw = s.w[ti];
  hc.dfield = s.field;
  T acc = T(0);
  for (int k = 0; k < s.len[to]; ++k) {
    hc.h = s.off[to] + k - 1;
    const T a = s.prog >= 0 ? core_prog(*v.kc, s.prog, hc) : T(1);
    acc += (T)v.kc->pool[s.w[to] + k] * (a * hc.dinner());
  }
  return acc;
}

template <typename T>
struct CoreNodeCtx {
  const CoreView<T>* v;
  __device__ __forceinline__ T field(int f) const { return v->at(f, 0, 0); }
  __device__ __forceinline__ T aux(int a) const { return v->at(v->kc->nfields + a, 0, 0); }
  __device__ __forceinline__ T coord(int d) const { return v->coord_node(d); }
  __device__ __forceinline__ T time() const { return (T)v->t; }
  __device__ __forceinline__ T param(int i) const { return v->p[i]; }
  __device__ __forceinline__ T dinner() const { return T(0); }
  __device__ __forceinline__ T slot(int k) const {
    const CoreSlot& s = v->kc->slots[k];
    const int ax = s.axis;
    switch (s.kind) {
      case MOL_SLOT_CENTERED:
        return v->stencil(s.field, ax, s.len[0], s.off[0], s.w[0]);
      case MOL_SLOT_WENO: {
        const T* q = v->lds + s.field * v->kc->vol + v->base;
        const int sd = v->stride(ax);
        return weno5_ref(q[-2 * sd], q[-sd], q[0], q[sd], q[2 * sd], (T)s.eps, (T)s.inv_dx);
      }
      case MOL_SLOT_UPWIND: {
        if (s.prog < 0) return v->stencil(s.field, ax, s.len[1], s.off[1], s.w[1]);
        CorePlainCtx<T> pc{v};
        const T coef = core_prog(*v->kc, s.prog, pc);
        const int q = coef > T(0) ? 1 : 0;
        return coef * v->stencil(s.field, ax, s.len[q], s.off[q], s.w[q]);
      }
      case MOL_SLOT_MIXED: {
        const int ay = s.axis2;
        T acc = T(0);
        for (int i = 0; i < s.len[0]; ++i) {
          T sy = T(0);
          for (int j = 0; j < s.len[1]; ++j)
            sy += (T)(v->kc->pool[s.w[0] + i] * v->kc->pool[s.w[1] + j]) *
                  v->at2(s.field, ax, s.off[0] + i, ay, s.off[1] + j);
          acc += sy;
        }
        return acc;
      }
      case MOL_SLOT_NONLINLAP:
        return core_nonlinlap(*v, s, 0, 1, 2);
      case MOL_SLOT_SPHERICAL: {
        CorePlainCtx<T> pc{v};
        const T a = s.prog >= 0 ? core_prog(*v->kc, s.prog, pc) : T(1);
        const T r = v->coord_node(ax);
        if (fabs((double)r) <= 1e-6) return T(6) * a * v->stencil(s.field, ax, s.len[1], s.off[1], s.w[1]);
        const T d1 = v->stencil(s.field, ax, s.len[0], s.off[0], s.w[0]);
        return a * (d1 / r + core_nonlinlap(*v, s, 2, 3, 4));
      }
      default:
        return T(0);
    }
  }
};

}  // namespace mol
