// Device-side building blocks of the generic stencil-IR right-hand side.
//
// A work-group owns a box-shaped tile of grid nodes.  It stages every dependent variable
// (and every auxiliary grid array) over the tile plus a halo into LDS, filling halo nodes
// that fall outside a variable's unknown box from the periodic ring or from the ghost-node
// closures, then evaluates each variable's right-hand-side program per node.  Operator
// slots read their taps from LDS.  Everything is templated on the arithmetic type T
// (double or float).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mol.h"

namespace mol {

struct Tab {
  int32_t len, off, w, nlow, nhigh, blen, wlow, whigh, hoff, ref_shift, pn_lo, pn_hi;
};
static_assert(sizeof(Tab) == sizeof(mol_optab), "table layout");

struct Slot {
  int32_t kind, field, axis, axis2, tab[5], prog;
  double eps, inv_dx;
};
static_assert(sizeof(Slot) == sizeof(mol_slot), "slot layout");

struct Clos {
  int32_t field, axis, side, nghost, K, wC, wc, gprog, src;
};
static_assert(sizeof(Clos) == sizeof(mol_closure), "closure layout");

// Kernel arguments shared by all stencil kernels (passed by value).
template <typename T>
struct KArgs {
  int32_t ndim, nfields, naux, nparams;
  int32_t n[3];
  int32_t periodic[3];
  int32_t lo[MOL_MAX_FIELDS][3], hi[MOL_MAX_FIELDS][3];
  int64_t fofs[MOL_MAX_FIELDS];
  int32_t ulo[3], uhi[3];      // union of the unknown boxes
  int32_t tile[3];             // tile extent in nodes
  int32_t ntile[3];            // tiles per axis
  int32_t halo[3];
  int32_t ext[3];              // tile + 2*halo
  int32_t vol;                 // ext[0]*ext[1]*ext[2]
  int32_t clos_id[MOL_MAX_FIELDS][3][2];
  int32_t rhs_prog[MOL_MAX_FIELDS];
  const double* coord[3];
  const Tab* tabs;
  const Slot* slots;
  const Clos* clos;
  const double* pool;
  const uint32_t* code;
  const double* cpool;
  const int32_t* prog;
  const T* aux[MOL_MAX_AUX];
};

// Fused epilogue of a right-hand-side launch (one Runge-Kutta stage):
//   mode 0:  out = f
//   mode 1:  out = ca*A + cb*B + cf*f ;  acc = acc_keep*acc + acc_f*f  (if acc)
//   mode 2:  compensated update  out = A (+) [cb*(B - subA*A) + cf*f]  with Kahan term comp
template <typename T>
struct Epi {
  int32_t mode;
  int32_t subA;
  T* out;
  const T* A;
  const T* B;
  T* acc;
  T* comp;
  double ca, cb, cf, acc_keep, acc_f;
  const double* tnow;   // device time of the step start (nullptr: 0)
  double tofs;          // stage offset added to *tnow (or the absolute time if tnow is null)
  const T* p;           // scalar parameters
  double* tnext;        // last stage of a step: the next step's clock, *tnext = *tnow + tadv
  double tadv;
};

// Device clock of the stage loop: the last stage of a step writes the next step's start time
// into the other buffer of a double-buffered clock (no kernel of the launch reads it).
template <typename T>
__device__ __forceinline__ void advance_clock(const Epi<T>& ep) {
  if (ep.tnext && blockIdx.x == 0 && threadIdx.x == 0) ep.tnext[0] = ep.tnow[0] + ep.tadv;
}

template <typename T>
__device__ __forceinline__ void epilogue(const Epi<T>& ep, long long idx, T f) {
  if (ep.mode == 0) {
    ep.out[idx] = f;
    return;
  }
  if (ep.mode == 1) {
    T v = (T)ep.cf * f;
    if (ep.A) v += (T)ep.ca * ep.A[idx];
    if (ep.B) v += (T)ep.cb * ep.B[idx];
    if (ep.acc) {
      T a = (T)ep.acc_f * f;
      if (ep.acc_keep != 0.0) a += (T)ep.acc_keep * ep.acc[idx];
      ep.acc[idx] = a;
    }
    ep.out[idx] = v;
    return;
  }
  // mode 2: compensated (Kahan) accumulation of the step update into A
  const T a = ep.A[idx];
  T incr = (T)ep.cf * f;
  if (ep.B) incr += (T)ep.cb * (ep.subA ? (ep.B[idx] - a) : ep.B[idx]);
  const T y = incr - ep.comp[idx];
  const T s = a + y;
  ep.comp[idx] = (s - a) - y;
  ep.out[idx] = s;
}

__device__ __forceinline__ int wrap_periodic(int k, int n) {
  // single wrap of the reference's bwrap: k <= 1 -> k + (n-1); k > n -> k - (n-1)
  // (src/discretization/interface_boundary.jl:33-42)
  return k <= 1 ? k + (n - 1) : (k > n ? k - (n - 1) : k);
}

// Row selection of a weight table at (node or half-point) index i along an axis.
// Returns the weight pointer, the absolute index of the first tap and the tap count.
struct Row {
  const double* w;
  int a0;
  int cnt;
};

__device__ __forceinline__ Row select_row(const Tab& tb, const double* pool, int i, int n,
                                          bool per) {
  Row r;
  int lref = n + tb.ref_shift;
  if (!per && i <= tb.nlow) {
    r.w = pool + tb.wlow + (i - 1) * tb.blen;
    r.a0 = 1;
    r.cnt = tb.blen;
  } else if (!per && i > lref - tb.nhigh) {
    int row = lref - i + tb.hoff;
    r.w = pool + tb.whigh + (row - 1) * tb.blen;
    r.a0 = lref - tb.blen + 1;
    r.cnt = tb.blen;
  } else {
    r.w = pool + tb.w;
    r.a0 = i + tb.off;
    r.cnt = tb.len;
  }
  return r;
}

template <typename T>
__device__ __forceinline__ T powi(T x, int e) {
  // power by repeated squaring, exact for |e| <= 3 the way x*x, x*x*x are written
  bool neg = e < 0;
  unsigned m = neg ? (unsigned)(-e) : (unsigned)e;
  T r;
  if (m == 0) r = T(1);
  else if (m == 1) r = x;
  else if (m == 2) r = x * x;
  else if (m == 3) r = x * x * x;
  else {
    T base = x;
    r = T(1);
    while (m) {
      if (m & 1u) r *= base;
      base *= base;
      m >>= 1;
    }
  }
  return neg ? T(1) / r : r;
}

// Library math kept out of line: these are rare in right-hand sides and large when inlined.
template <typename T>
__device__ __noinline__ T math_unary(uint32_t op, T a) {
  switch (op) {
    case MOL_OP_EXP: return exp(a);
    case MOL_OP_LOG: return log(a);
    case MOL_OP_SIN: return sin(a);
    case MOL_OP_COS: return cos(a);
    case MOL_OP_TAN: return tan(a);
    case MOL_OP_TANH: return tanh(a);
    case MOL_OP_SINH: return sinh(a);
    case MOL_OP_COSH: return cosh(a);
    case MOL_OP_ATAN: return atan(a);
    case MOL_OP_ASIN: return asin(a);
    case MOL_OP_ACOS: return acos(a);
    case MOL_OP_ERF: return erf(a);
    case MOL_OP_LOG10: return log10(a);
    case MOL_OP_CBRT: return cbrt(a);
    case MOL_OP_EXPM1: return expm1(a);
    case MOL_OP_LOG1P: return log1p(a);
    case MOL_OP_FLOOR: return floor(a);
    case MOL_OP_CEIL: return ceil(a);
    default: return a;
  }
}

template <typename T>
__device__ __noinline__ T math_pow(T a, T b) { return pow(a, b); }

// Interpreter of the pointwise bytecode.  The host compiler allocates registers as a stack
// (an instruction at depth d reads r[d] and r[d+1] and writes r[d]), so the register file is
// kept as a shift-register stack s0 (top) .. s5 in named variables: no indexed array, nothing
// in scratch.  Leaves push, binary operations pop (or use a constant operand), unary
// operations rewrite the top.  Programs deeper than MOL_STACK are rejected by the host.
#define MOL_STACK 6

template <typename T>
__device__ __forceinline__ T binop(uint32_t op, T a, T b) {
  switch (op) {
    case MOL_OP_ADD: return a + b;
    case MOL_OP_SUB: return a - b;
    case MOL_OP_MUL: return a * b;
    case MOL_OP_DIV: return a / b;
    case MOL_OP_POW: return math_pow(a, b);
    case MOL_OP_MIN: return fmin(a, b);
    case MOL_OP_MAX: return fmax(a, b);
    case MOL_OP_LT: return a < b ? T(1) : T(0);
    case MOL_OP_LE: return a <= b ? T(1) : T(0);
    case MOL_OP_GT: return a > b ? T(1) : T(0);
    case MOL_OP_GE: return a >= b ? T(1) : T(0);
    case MOL_OP_EQ: return a == b ? T(1) : T(0);
    case MOL_OP_NE: return a != b ? T(1) : T(0);
    case MOL_OP_AND: return (a != T(0) && b != T(0)) ? T(1) : T(0);
    case MOL_OP_OR: return (a != T(0) || b != T(0)) ? T(1) : T(0);
    default: return a;
  }
}

template <typename T>
__device__ __forceinline__ T unop(uint32_t op, T a, int imm) {
  switch (op) {
    case MOL_OP_NEG: return -a;
    case MOL_OP_SQ: return a * a;
    case MOL_OP_INV: return T(1) / a;
    case MOL_OP_ABS: return fabs(a);
    case MOL_OP_SQRT: return sqrt(a);
    case MOL_OP_SIGN: return a > T(0) ? T(1) : (a < T(0) ? T(-1) : T(0));
    case MOL_OP_HEAVISIDE: return a > T(0) ? T(1) : (a < T(0) ? T(0) : T(0.5));
    case MOL_OP_POWI: return powi(a, (imm & 0x400) ? imm - 0x800 : imm);
    default: return math_unary(op, a);
  }
}

// code/cpool/prog come from `Src` (kernel arguments or device memory).
template <typename T, class Src, class Ctx>
__device__ __forceinline__ T stack_prog(const Src& src, int prog, const Ctx& ctx) {
  const int start = src.prog[2 * prog];
  const int len = src.prog[2 * prog + 1];
  T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0), s4 = T(0), s5 = T(0);
  for (int pc = start; pc < start + len; ++pc) {
    const uint32_t ins = src.code[pc];
    const uint32_t op = ins & 0xffu;
    const bool bconst = (ins >> 20) & 1u;
    const int imm = (int)(ins >> 21);
    if (op < 16u) {
      T v;
      switch (op) {
        case MOL_OP_CONST: v = (T)src.cpool[imm]; break;
        case MOL_OP_FIELD: v = ctx.field(imm); break;
        case MOL_OP_SLOT: v = ctx.slot(imm); break;
        case MOL_OP_COORD: v = ctx.coord(imm); break;
        case MOL_OP_TIME: v = ctx.time(); break;
        case MOL_OP_PARAM: v = ctx.param(imm); break;
        case MOL_OP_AUX: v = ctx.aux(imm); break;
        default: v = ctx.dinner(); break;
      }
      s5 = s4; s4 = s3; s3 = s2; s2 = s1; s1 = s0; s0 = v;
    } else if (op < 48u) {
      if (bconst) {
        const T c = (T)src.cpool[imm];
        if (op == MOL_OP_SELECT) {
          s0 = (s1 != T(0)) ? s0 : c;
          s1 = s2; s2 = s3; s3 = s4; s4 = s5;
        } else {
          s0 = binop(op, s0, c);
        }
      } else if (op == MOL_OP_SELECT) {
        s0 = (s2 != T(0)) ? s1 : s0;
        s1 = s3; s2 = s4; s3 = s5;
      } else {
        s0 = binop(op, s1, s0);
        s1 = s2; s2 = s3; s3 = s4; s4 = s5;
      }
    } else {
      s0 = unop(op, s0, imm);
    }
  }
  return s0;
}

// Vectorised interpreter: one instruction dispatch serves B nodes held by the same lane
// (register blocking), which divides the scalar-unit cost of interpretation by B.  The stack
// is DEPTH levels of B values; the host only routes programs of depth <= DEPTH here.
// Ctx::leaf(op, imm, v) fills v[0..B) with the leaf values of the B nodes.
template <typename T, int B, int DEPTH, class Src, class Ctx>
__device__ __forceinline__ void stack_prog_vec(const Src& src, int prog, const Ctx& ctx, T* out) {
  static_assert(DEPTH >= 2 && DEPTH <= 4, "vector stack depth");
  const int start = src.prog[2 * prog];
  const int len = src.prog[2 * prog + 1];
  T s0[B], s1[B], s2[B], s3[B];
#pragma unroll
  for (int b = 0; b < B; ++b) s0[b] = s1[b] = s2[b] = s3[b] = T(0);
  for (int pc = start; pc < start + len; ++pc) {
    const uint32_t ins = src.code[pc];
    const uint32_t op = ins & 0xffu;
    const bool bconst = (ins >> 20) & 1u;
    const int imm = (int)(ins >> 21);
    if (op < 16u) {
      T v[B];
      if (op == MOL_OP_CONST) {
        const T c = (T)src.cpool[imm];
#pragma unroll
        for (int b = 0; b < B; ++b) v[b] = c;
      } else {
        ctx.leaf(op, imm, v);
      }
#pragma unroll
      for (int b = 0; b < B; ++b) {
        if (DEPTH > 3) s3[b] = s2[b];
        if (DEPTH > 2) s2[b] = s1[b];
        s1[b] = s0[b];
        s0[b] = v[b];
      }
    } else if (op < 48u) {
      if (bconst) {
        const T c = (T)src.cpool[imm];
        if (op == MOL_OP_SELECT) {
#pragma unroll
          for (int b = 0; b < B; ++b) {
            s0[b] = (s1[b] != T(0)) ? s0[b] : c;
            s1[b] = s2[b];
            s2[b] = s3[b];
          }
        } else {
#pragma unroll
          for (int b = 0; b < B; ++b) s0[b] = binop(op, s0[b], c);
        }
      } else if (op == MOL_OP_SELECT) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
          s0[b] = (s2[b] != T(0)) ? s1[b] : s0[b];
          s1[b] = s3[b];
        }
      } else {
#pragma unroll
        for (int b = 0; b < B; ++b) {
          s0[b] = binop(op, s1[b], s0[b]);
          s1[b] = s2[b];
          s2[b] = s3[b];
        }
      }
    } else {
#pragma unroll
      for (int b = 0; b < B; ++b) s0[b] = unop(op, s0[b], imm);
    }
  }
#pragma unroll
  for (int b = 0; b < B; ++b) out[b] = s0[b];
}

// Tables of the generic path in device memory, in the layout stack_prog reads.
struct GlobalProgs {
  const int32_t* prog;
  const uint32_t* code;
  const double* cpool;
};

template <typename T, class Ctx>
__device__ __forceinline__ T run_prog(const KArgs<T>& ka, int prog, Ctx& ctx) {
  GlobalProgs g{ka.prog, ka.code, ka.cpool};
  return stack_prog<T>(g, prog, ctx);
}

// Context used for ghost-node programs G(t, boundary coordinates): no variables.
template <typename T>
struct BoundaryCtx {
  const KArgs<T>* ka;
  int idx[3];      // full-grid node index of the boundary node
  double t;
  const T* p;
  __device__ T field(int) const { return T(0); }
  __device__ T slot(int) const { return T(0); }
  __device__ T coord(int d) const { return (T)ka->coord[d][idx[d] - 1]; }
  __device__ T time() const { return (T)t; }
  __device__ T param(int i) const { return p[i]; }
  __device__ T aux(int a) const {
    // auxiliary arrays are full-grid: x fastest
    long long li = (long long)(idx[0] - 1);
    if (ka->ndim > 1) li += (long long)ka->n[0] * (idx[1] - 1);
    if (ka->ndim > 2) li += (long long)ka->n[0] * ka->n[1] * (idx[2] - 1);
    return ka->aux[a][li];
  }
  __device__ T dinner() const { return T(0); }
};

// Value of variable f at full-grid node p (already periodic-wrapped), read from global memory.
template <typename T>
__device__ __forceinline__ T load_unknown(const KArgs<T>& ka, const T* __restrict__ u, int f,
                                          const int* p) {
  long long li = (long long)(p[0] - ka.lo[f][0]);
  long long stride = ka.hi[f][0] - ka.lo[f][0] + 1;
  if (ka.ndim > 1) {
    li += stride * (p[1] - ka.lo[f][1]);
    stride *= ka.hi[f][1] - ka.lo[f][1] + 1;
  }
  if (ka.ndim > 2) li += stride * (p[2] - ka.lo[f][2]);
  return u[ka.fofs[f] + li];
}

template <typename T>
__device__ __noinline__ T closure_value(const KArgs<T>& ka, const T* __restrict__ u, int f,
                                        const int* q, int d, double t, const T* p);

// Value of variable f at an arbitrary (unwrapped) full-grid position: ring wrap on periodic
// axes, unknown, ghost closure (outside the box along exactly one axis) or 0 (corner /
// outside the domain), the latter as the reference's corner equations do
// (src/discretization/generate_bc_eqs.jl:396-416).
template <typename T>
__device__ T field_at(const KArgs<T>& ka, const T* __restrict__ u, int f, const int* pin,
                      double t, const T* p) {
  int q[3] = {1, 1, 1};
  int out_axis = -1, nout = 0;
  for (int d = 0; d < ka.ndim; ++d) {
    int k = pin[d];
    if (ka.periodic[d]) k = wrap_periodic(k, ka.n[d]);
    if (k < 1 || k > ka.n[d]) return T(0);
    q[d] = k;
    if (k < ka.lo[f][d] || k > ka.hi[f][d]) {
      ++nout;
      out_axis = d;
    }
  }
  if (nout == 0) return load_unknown(ka, u, f, q);
  if (nout > 1) return T(0);
  return closure_value(ka, u, f, q, out_axis, t, p);
}

// Ghost-node value from the closure of (variable f, axis d) at wrapped position q.
template <typename T>
__device__ __noinline__ T closure_value(const KArgs<T>& ka, const T* __restrict__ u, int f,
                                        const int* q, int d, double t, const T* p) {
  const int side = q[d] < ka.lo[f][d] ? 0 : 1;
  const int cid = ka.clos_id[f][d][side];
  if (cid < 0) return T(0);
  const Clos cl = ka.clos[cid];
  const int g = side == 0 ? ka.lo[f][d] - 1 - q[d] : q[d] - ka.hi[f][d] - 1;
  if (g >= cl.nghost) return T(0);
  const int first = side == 0 ? ka.lo[f][d] : ka.hi[f][d];
  const int dir = side == 0 ? 1 : -1;
  T acc = T(0);
  int qq[3] = {q[0], q[1], q[2]};
  for (int m = 0; m < cl.K; ++m) {
    qq[d] = first + m * dir;
    acc += (T)ka.pool[cl.wC + g * cl.K + m] * load_unknown(ka, u, f, qq);
  }
  if (cl.gprog >= 0) {
    BoundaryCtx<T> bc;
    bc.ka = &ka;
    bc.idx[0] = q[0];
    bc.idx[1] = q[1];
    bc.idx[2] = q[2];
    bc.idx[d] = side == 0 ? 1 : ka.n[d];
    bc.t = t;
    bc.p = p;
    acc += (T)ka.pool[cl.wc + g] * run_prog(ka, cl.gprog, bc);
  }
  return acc;
}

template <typename T>
__device__ T aux_at(const KArgs<T>& ka, int a, const int* pin) {
  long long li = 0, stride = 1;
  for (int d = 0; d < ka.ndim; ++d) {
    int k = pin[d];
    if (ka.periodic[d]) k = wrap_periodic(k, ka.n[d]);
    if (k < 1 || k > ka.n[d]) return T(0);
    li += stride * (k - 1);
    stride *= ka.n[d];
  }
  return ka.aux[a][li];
}

// ---------------------------------------------------------------------------------------
// Tile view: LDS arrays over the tile + halo, addressed by local coordinates.
template <typename T>
struct TileView {
  const KArgs<T>* ka;
  const T* lds;
  int org[3];      // full-grid index of local coordinate (halo, halo, halo)
  __device__ __forceinline__ int lin(int lx, int ly, int lz) const {
    return lx + ka->ext[0] * (ly + ka->ext[1] * lz);
  }
  // value of array `arr` (variable index, or nfields + aux index) at local coordinates
  __device__ __forceinline__ T at(int arr, int lx, int ly, int lz) const {
    return lds[arr * ka->vol + lin(lx, ly, lz)];
  }
};

// Coordinate value at (unwrapped) index k along axis d, as the reference reads grid[x][bwrap(k)].
template <typename T>
__device__ __forceinline__ T coord_at(const KArgs<T>& ka, int d, int k) {
  if (ka.periodic[d]) k = wrap_periodic(k, ka.n[d]);
  return (T)ka.coord[d][k - 1];
}

// Half-point context of the nonlinear Laplacian: variables, coordinates and auxiliary arrays
// interpolated at half point h along axis `ax` with the 0th-order half-offset operator
// (src/discretization/schemes/nonlinear_laplacian/nonlinear_laplacian.jl:54-83).
template <typename T>
struct HalfCtx {
  const KArgs<T>* ka;
  const TileView<T>* tv;
  int L[3];        // local coordinates of the node
  int I[3];        // full-grid index of the node
  int ax;
  Row ri;          // interpolation row at the half point
  Row rd;          // inner derivative row at the half point
  int dfield;      // variable differentiated by the inner derivative
  double t;
  const T* p;
  __device__ T interp_arr(int arr) const {
    T s = T(0);
    int l[3] = {L[0], L[1], L[2]};
    for (int k = 0; k < ri.cnt; ++k) {
      l[ax] = L[ax] + (ri.a0 + k - I[ax]);
      s += (T)ri.w[k] * tv->at(arr, l[0], l[1], l[2]);
    }
    return s;
  }
  __device__ T field(int f) const { return interp_arr(f); }
  __device__ T aux(int a) const { return interp_arr(ka->nfields + a); }
  __device__ T coord(int d) const {
    if (d != ax) return (T)ka->coord[d][I[d] - 1];
    T s = T(0);
    for (int k = 0; k < ri.cnt; ++k) s += (T)ri.w[k] * coord_at(*ka, d, ri.a0 + k);
    return s;
  }
  __device__ T slot(int) const { return T(0); }
  __device__ T time() const { return (T)t; }
  __device__ T param(int i) const { return p[i]; }
  __device__ T dinner() const {
    T s = T(0);
    int l[3] = {L[0], L[1], L[2]};
    for (int k = 0; k < rd.cnt; ++k) {
      l[ax] = L[ax] + (rd.a0 + k - I[ax]);
      s += (T)rd.w[k] * tv->at(dfield, l[0], l[1], l[2]);
    }
    return s;
  }
};

// Node context without slots (upwind coefficients, spherical prefactor).
template <typename T>
struct PlainNodeCtx {
  const KArgs<T>* ka;
  const TileView<T>* tv;
  int L[3];
  int I[3];
  double t;
  const T* p;
  __device__ T field(int f) const { return tv->at(f, L[0], L[1], L[2]); }
  __device__ T aux(int a) const { return tv->at(ka->nfields + a, L[0], L[1], L[2]); }
  __device__ T coord(int d) const { return (T)ka->coord[d][I[d] - 1]; }
  __device__ T slot(int) const { return T(0); }
  __device__ T time() const { return (T)t; }
  __device__ T param(int i) const { return p[i]; }
  __device__ T dinner() const { return T(0); }
};

// Apply a weight row along axis d to array arr around the node.
template <typename T>
__device__ __forceinline__ T apply_row(const TileView<T>& tv, int arr, const int* L, const int* I,
                                       int d, const Row& r) {
  T s = T(0);
  int l[3] = {L[0], L[1], L[2]};
  for (int k = 0; k < r.cnt; ++k) {
    l[d] = L[d] + (r.a0 + k - I[d]);
    s += (T)r.w[k] * tv.at(arr, l[0], l[1], l[2]);
  }
  return s;
}

// Jiang-Shu WENO5 node derivative (û⁻_{i+1/2} - û⁺_{i-1/2})/dx, restating weno_f_uniform
// (src/discretization/schemes/WENO/WENO.jl:6-57) term by term.
template <typename T>
__device__ __forceinline__ T weno5_ref(T um2, T um1, T u0, T up1, T up2, T eps, T inv_dx) {
  const T b1 = T(13) * (u0 - T(2) * up1 + up2) * (u0 - T(2) * up1 + up2) / T(12) +
               (T(3) * u0 - T(4) * up1 + up2) * (T(3) * u0 - T(4) * up1 + up2) / T(4);
  const T b2 = T(13) * (um1 - T(2) * u0 + up1) * (um1 - T(2) * u0 + up1) / T(12) +
               (um1 - up1) * (um1 - up1) / T(4);
  const T b3 = T(13) * (um2 - T(2) * um1 + u0) * (um2 - T(2) * um1 + u0) / T(12) +
               (um2 - T(4) * um1 + T(3) * u0) * (um2 - T(4) * um1 + T(3) * u0) / T(4);
  const T q1 = (eps + b1) * (eps + b1), q2 = (eps + b2) * (eps + b2), q3 = (eps + b3) * (eps + b3);
  const T om1 = T(0.1) / q1, om2 = T(0.6) / q2, om3 = T(0.3) / q3;
  const T wmd = om1 + om2 + om3;
  const T op1 = T(0.3) / q1, op2 = T(0.6) / q2, op3 = T(0.1) / q3;
  const T wpd = op1 + op2 + op3;
  const T hm1 = (T(11) * u0 - T(7) * up1 + T(2) * up2) / T(6);
  const T hm2 = (T(5) * u0 - up1 + T(2) * um1) / T(6);
  const T hm3 = (T(2) * u0 + T(5) * um1 - um2) / T(6);
  const T hp1 = (T(2) * u0 + T(5) * up1 - up2) / T(6);
  const T hp2 = (T(5) * u0 + T(2) * up1 - um1) / T(6);
  const T hp3 = (T(11) * u0 - T(7) * um1 + T(2) * um2) / T(6);
  const T hp = (op1 / wpd) * hp1 + (op2 / wpd) * hp2 + (op3 / wpd) * hp3;
  const T hm = (om1 / wmd) * hm1 + (om2 / wmd) * hm2 + (om3 / wmd) * hm3;
  return (hp - hm) * inv_dx;
}

// Node context with operator slots (right-hand-side programs).
template <typename T>
struct NodeCtx {
  const KArgs<T>* ka;
  const TileView<T>* tv;
  int L[3];
  int I[3];
  double t;
  const T* p;
  __device__ T field(int f) const { return tv->at(f, L[0], L[1], L[2]); }
  __device__ T aux(int a) const { return tv->at(ka->nfields + a, L[0], L[1], L[2]); }
  __device__ T coord(int d) const { return (T)ka->coord[d][I[d] - 1]; }
  __device__ T time() const { return (T)t; }
  __device__ T param(int i) const { return p[i]; }
  __device__ T dinner() const { return T(0); }

  __device__ __noinline__ T nonlinlap(const Slot& s, int tab_outer, int tab_inner,
                                      int tab_interp) const {
    const int ax = s.axis;
    const int n = ka->n[ax];
    const bool per = ka->periodic[ax] != 0;
    const Tab to = ka->tabs[tab_outer];
    const Tab tin = ka->tabs[tab_inner];
    const Tab tip = ka->tabs[tab_interp];
    // outer operator on half points, evaluated at II - e_ax with len = n - 1
    const Row ro = select_row(to, ka->pool, I[ax] - 1, n, per);
    HalfCtx<T> hc;
    hc.ka = ka;
    hc.tv = tv;
    hc.L[0] = L[0]; hc.L[1] = L[1]; hc.L[2] = L[2];
    hc.I[0] = I[0]; hc.I[1] = I[1]; hc.I[2] = I[2];
    hc.ax = ax;
    hc.dfield = s.field;
    hc.t = t;
    hc.p = p;
    T acc = T(0);
    for (int k = 0; k < ro.cnt; ++k) {
      const int h = ro.a0 + k;   // half-point index (between nodes h and h+1)
      hc.ri = select_row(tip, ka->pool, h, n, per);
      hc.rd = select_row(tin, ka->pool, h, n, per);
      const T a = s.prog >= 0 ? run_prog(*ka, s.prog, hc) : T(1);
      acc += (T)ro.w[k] * (a * hc.dinner());
    }
    return acc;
  }

  __device__ T slot(int k) const;
  __device__ T slot_impl(int k) const {
    const Slot s = ka->slots[k];
    const int ax = s.axis;
    const int n = ka->n[ax];
    const bool per = ka->periodic[ax] != 0;
    switch (s.kind) {
      case MOL_SLOT_CENTERED: {
        const Row r = select_row(ka->tabs[s.tab[0]], ka->pool, I[ax], n, per);
        return apply_row(*tv, s.field, L, I, ax, r);
      }
      case MOL_SLOT_UPWIND: {
        if (s.prog < 0) {
          const Row r = select_row(ka->tabs[s.tab[1]], ka->pool, I[ax], n, per);
          return apply_row(*tv, s.field, L, I, ax, r);
        }
        PlainNodeCtx<T> pc{ka, tv, {L[0], L[1], L[2]}, {I[0], I[1], I[2]}, t, p};
        const T coef = run_prog(*ka, s.prog, pc);
        const int tb = coef > T(0) ? s.tab[1] : s.tab[0];
        const Row r = select_row(ka->tabs[tb], ka->pool, I[ax], n, per);
        return coef * apply_row(*tv, s.field, L, I, ax, r);
      }
      case MOL_SLOT_WENO: {
        int l[3] = {L[0], L[1], L[2]};
        T v[5];
        for (int q = 0; q < 5; ++q) {
          l[ax] = L[ax] + q - 2;
          v[q] = tv->at(s.field, l[0], l[1], l[2]);
        }
        return weno5_ref(v[0], v[1], v[2], v[3], v[4], (T)s.eps, (T)s.inv_dx);
      }
      case MOL_SLOT_NONLINLAP:
        return nonlinlap(s, s.tab[0], s.tab[1], s.tab[2]);
      case MOL_SLOT_SPHERICAL: {
        PlainNodeCtx<T> pc{ka, tv, {L[0], L[1], L[2]}, {I[0], I[1], I[2]}, t, p};
        const T a = s.prog >= 0 ? run_prog(*ka, s.prog, pc) : T(1);
        const T r = (T)ka->coord[ax][I[ax] - 1];
        if (fabs((double)r) <= 1e-6) {
          // r ~ 0: 6 a D2 u (spherical_laplacian.jl:30-32)
          const Row r2 = select_row(ka->tabs[s.tab[1]], ka->pool, I[ax], n, per);
          return T(6) * a * apply_row(*tv, s.field, L, I, ax, r2);
        }
        const Row r1 = select_row(ka->tabs[s.tab[0]], ka->pool, I[ax], n, per);
        const T d1 = apply_row(*tv, s.field, L, I, ax, r1);
        return a * (d1 / r + nonlinlap(s, s.tab[2], s.tab[3], s.tab[4]));
      }
      case MOL_SLOT_MIXED: {
        const int ay = s.axis2;
        const Row rx = select_row(ka->tabs[s.tab[0]], ka->pool, I[ax], n, per);
        const Row ry = select_row(ka->tabs[s.tab[1]], ka->pool, I[ay], ka->n[ay],
                                  ka->periodic[ay] != 0);
        T acc = T(0);
        int l[3] = {L[0], L[1], L[2]};
        for (int i = 0; i < rx.cnt; ++i) {
          l[ax] = L[ax] + (rx.a0 + i - I[ax]);
          T sy = T(0);
          for (int j = 0; j < ry.cnt; ++j) {
            l[ay] = L[ay] + (ry.a0 + j - I[ay]);
            sy += (T)(rx.w[i] * ry.w[j]) * tv->at(s.field, l[0], l[1], l[2]);
          }
          acc += sy;
        }
        return acc;
      }
      default:
        return T(0);
    }
  }
};

template <typename T>
__device__ __noinline__ T eval_slot(const NodeCtx<T>* nc, int k) { return nc->slot_impl(k); }

template <typename T>
__device__ T NodeCtx<T>::slot(int k) const { return eval_slot<T>(this, k); }

}  // namespace mol
