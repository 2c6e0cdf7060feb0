// Device-side skeleton of the generated right-hand-side kernels.
//
// This header is not compiled with the library: its text is embedded into libmolgpu.so and
// handed to hiprtc together with the source that mol_jit.cpp generates for one PDE system.
// The generated source defines MOL_T (double or float) before including this header, then a
// struct `Geo` with the tile geometry as constants and the per-node right-hand side
// `Geo::node`, whose operator weights, constants and programs are folded into straight-line
// code.  The tile kernel below stages the tile plus halo into LDS with periodic wrap and runs
// the node function with the fused Runge-Kutta stage epilogue.
#pragma once

typedef MOL_T T;
typedef long long i64;

// Layout identical to mol::Epi<T> of the library (passed by value as a kernel argument).
struct Epi {
  int mode;
  int subA;
  T* out;
  const T* A;
  const T* B;
  T* acc;
  T* comp;
  double ca, cb, cf, acc_keep, acc_f;
  const double* tnow;
  double tofs;
  const T* p;
  double* tnext;   // last stage of a step: *tnext = *tnow + tadv (double-buffered device clock)
  double tadv;
};

__device__ __forceinline__ void mol_advance_clock(const Epi& ep) {
  if (ep.tnext && blockIdx.x == 0 && threadIdx.x == 0) ep.tnext[0] = ep.tnow[0] + ep.tadv;
}

// Node coordinates and auxiliary grid arrays (layout identical to mol::JitPtrs<T>).
struct JPtrs {
  const double* coord[3];
  const T* aux[8];
  const double* pool;      // weight pool: per-node rows of nonuniform axes
  const double* wtab[3];   // per axis: nonuniform WENO5 coefficient rows (MOL_NU_NC per node)
};

// Fused epilogue of one Runge-Kutta stage:
//   mode 0:  out = f
//   mode 1:  out = ca*A + cb*B + cf*f ;  acc = acc_keep*acc + acc_f*f  (if acc)
//   mode 2:  compensated update  out = A (+) [cb*(B - subA*A) + cf*f]  with Kahan term comp
template <class I>
__device__ __forceinline__ void mol_epilogue(const Epi& ep, const I idx, T f) {
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
  const T a = ep.A[idx];
  T incr = (T)ep.cf * f;
  if (ep.B) incr += (T)ep.cb * (ep.subA ? (ep.B[idx] - a) : ep.B[idx]);
  const T y = incr - ep.comp[idx];
  const T s = a + y;
  ep.comp[idx] = (s - a) - y;
  ep.out[idx] = s;
}

// The epilogue of CH nodes at once (ok[k]: node k exists; out[k][f] at index idx[k] + dv[f]):
// every operand load of the CH nodes is issued before the first store, so one memory latency
// is exposed per chunk instead of one per node.
template <int CH, int NF>
__device__ __forceinline__ void mol_epilogue_chunk(const Epi& ep, const i64 (&idx)[CH], const bool (&ok)[CH],
                                                   const i64 (&dv)[NF], const T (&out)[CH][NF]) {
  if (ep.mode == 0) {
#pragma unroll
    for (int k = 0; k < CH; ++k)
      if (ok[k])
#pragma unroll
        for (int f = 0; f < NF; ++f) ep.out[idx[k] + dv[f]] = out[k][f];
    return;
  }
  T a[CH][NF], b[CH][NF], c[CH][NF];
  const bool keep = ep.mode == 1 && ep.acc && ep.acc_keep != 0.0;
#pragma unroll
  for (int k = 0; k < CH; ++k)
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const i64 i = idx[k] + dv[f];
      a[k][f] = ok[k] && ep.A ? ep.A[i] : T(0);
      b[k][f] = ok[k] && ep.B ? ep.B[i] : T(0);
      c[k][f] = ok[k] && (ep.mode == 2 || keep) ? (ep.mode == 2 ? ep.comp[i] : ep.acc[i]) : T(0);
    }
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    if (!ok[k]) continue;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const i64 i = idx[k] + dv[f];
      const T fv = out[k][f];
      if (ep.mode == 1) {
        T v = (T)ep.cf * fv;
        if (ep.A) v += (T)ep.ca * a[k][f];
        if (ep.B) v += (T)ep.cb * b[k][f];
        if (ep.acc) {
          T ac = (T)ep.acc_f * fv;
          if (ep.acc_keep != 0.0) ac += (T)ep.acc_keep * c[k][f];
          ep.acc[i] = ac;
        }
        ep.out[i] = v;
      } else {
        T incr = (T)ep.cf * fv;
        if (ep.B) incr += (T)ep.cb * (ep.subA ? (b[k][f] - a[k][f]) : b[k][f]);
        const T y = incr - c[k][f];
        const T sm = a[k][f] + y;
        ep.comp[i] = (sm - a[k][f]) - y;
        ep.out[i] = sm;
      }
    }
  }
}

// Single periodic wrap of a node index onto the ring 2..n (node 1 is the image of node n).
__device__ __forceinline__ int mol_wrap(int k, int n) {
  return k <= 1 ? k + (n - 1) : (k > n ? k - (n - 1) : k);
}

__device__ __forceinline__ T mol_sign(T a) { return a > T(0) ? T(1) : (a < T(0) ? T(-1) : T(0)); }
__device__ __forceinline__ T mol_heaviside(T a) { return a > T(0) ? T(1) : (a < T(0) ? T(0) : T(0.5)); }

// x^e for |e| > 3 by repeated squaring (small exponents are expanded by the generator).
__device__ __noinline__ T mol_powi(T x, int e) {
  const bool neg = e < 0;
  unsigned m = neg ? (unsigned)(-e) : (unsigned)e;
  T r = T(1), base = x;
  while (m) {
    if (m & 1u) r *= base;
    base *= base;
    m >>= 1;
  }
  return neg ? T(1) / r : r;
}

// FP64 quotient n / d for the WENO5 weights: the hardware reciprocal (v_rcp_f64) refined by
// two Newton steps and the quotient corrected by its residual (Markstein): 8 instructions
// against about 11 for the IEEE division sequence, within an ulp or two of the IEEE quotient.
// d is a positive normal number here (a product of sums of squares offset by eps).
__device__ __forceinline__ double mol_quot(const double n, const double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  const double q = n * r;
  return fma(r, fma(-d, q, n), q);
}

// Jiang-Shu WENO5 node derivative (u^-_{i+1/2} - u^+_{i-1/2}) / dx from the five values
// u(i-2) .. u(i+2).  The candidate reconstructions are written as u(i) + difference/6 so u(i)
// cancels, and the normalised nonlinear weights gamma_k/q_k / sum_j gamma_j/q_j are cleared
// to gamma_k P_k / sum_j gamma_j P_j (P_k = product of the other two q_j): one division.  The
// weights are unchanged when every eps + beta_k is scaled by one factor, so the smoothness
// sums are formed as 4 (eps + beta_k) = 13/3 D2^2 + a^2 + 4 eps (no 1/4 and 13/12 products).
template <class U>
__device__ __forceinline__ U mol_weno5_t(U um2, U um1, U u0, U up1, U up2, T eps, T inv_dx) {
  const U dmm = um1 - um2, dm = u0 - um1, dp = up1 - u0, dpp = up2 - up1;
  const U D2m = dm - dmm, D20 = dp - dm, D2p = dpp - dp;
  const U a1 = T(-2) * dp + D2p;
  const U a2 = dm + dp;
  const U a3 = T(2) * dm + D2m;
  const T eps4 = T(4) * eps;
  const U b1 = (T(13.0 / 3.0) * D2p) * D2p + (a1 * a1 + eps4);
  const U b2 = (T(13.0 / 3.0) * D20) * D20 + (a2 * a2 + eps4);
  const U b3 = (T(13.0 / 3.0) * D2m) * D2m + (a3 * a3 + eps4);
  const U q1 = b1 * b1, q2 = b2 * b2, q3 = b3 * b3;
  const U P1 = q2 * q3, P2 = q1 * q3, P3 = q1 * q2;
  const U A1 = T(0.3) * P1, A2 = T(0.6) * P2, A3 = T(0.1) * P3;
  const U B1 = T(0.1) * P1, B3 = T(0.3) * P3;
  const U Dp = A1 + A2 + A3;
  const U Dm = B1 + A2 + B3;
  const U ep1 = T(4) * dp - dpp, ep2 = T(2) * dp + dm, ep3 = T(5) * dm - T(2) * dmm;
  const U em1 = T(-5) * dp + T(2) * dpp, em2 = T(-2) * dm - dp, em3 = T(-4) * dm + dmm;
  const U Np = A1 * ep1 + A2 * ep2 + A3 * ep3;
  const U Nm = B1 * em1 + A2 * em2 + B3 * em3;
  return mol_quot(Np * Dm - Nm * Dp, Dp * Dm) * (inv_dx * T(1.0 / 6.0));
}

// Reciprocal and quotient for the FP32 WENO5 weights: the hardware reciprocal (v_rcp_f32,
// within 1 ulp) refined by one Newton step, and the quotient corrected by its residual
// (Markstein), which is the correctly rounded result but for rare ties, in 3 and 5
// instructions against about 10 for the IEEE division sequence.  FP64: plain divisions.
__device__ __forceinline__ float mol_rcp(const float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ double mol_rcp(const double x) { return 1.0 / x; }
__device__ __forceinline__ float mol_div(const float n, const float d) {
  const float r = mol_rcp(d);
  const float q = n * r;
  return fmaf(fmaf(-d, q, n), r, q);
}
__device__ __forceinline__ double mol_div(const double n, const double d) { return n / d; }

// The same derivative in the reference's normalised-weight form (weno_f_uniform,
// src/discretization/schemes/WENO/WENO.jl:6-57).  Used in FP32, where the products of four
// squared smoothness indicators of the cleared form (~eps^8) underflow.
template <class U>
__device__ __forceinline__ U mol_weno5_ref_t(U um2, U um1, U u0, U up1, U up2, T eps, T inv_dx) {
  const U c1 = u0 - T(2) * up1 + up2, c2 = T(3) * u0 - T(4) * up1 + up2;
  const U c3 = um1 - T(2) * u0 + up1, c4 = um1 - up1;
  const U c5 = um2 - T(2) * um1 + u0, c6 = um2 - T(4) * um1 + T(3) * u0;
  const U b1 = T(13) * c1 * c1 / T(12) + c2 * c2 / T(4);
  const U b2 = T(13) * c3 * c3 / T(12) + c4 * c4 / T(4);
  const U b3 = T(13) * c5 * c5 / T(12) + c6 * c6 / T(4);
  const U q1 = (eps + b1) * (eps + b1), q2 = (eps + b2) * (eps + b2), q3 = (eps + b3) * (eps + b3);
  // the six weights share three reciprocals and each normalisation one division
  const U r1 = mol_rcp(q1), r2 = mol_rcp(q2), r3 = mol_rcp(q3);
  const U om1 = T(0.1) * r1, om2 = T(0.6) * r2, om3 = T(0.3) * r3;
  const U op1 = T(0.3) * r1, op2 = T(0.6) * r2, op3 = T(0.1) * r3;
  const U wmd = om1 + om2 + om3, wpd = op1 + op2 + op3;
  // candidate values times 6 (the 1/6 goes into the final scale)
  const U hm1 = T(11) * u0 - T(7) * up1 + T(2) * up2;
  const U hm2 = T(5) * u0 - up1 + T(2) * um1;
  const U hm3 = T(2) * u0 + T(5) * um1 - um2;
  const U hp1 = T(2) * u0 + T(5) * up1 - up2;
  const U hp2 = T(5) * u0 + T(2) * up1 - um1;
  const U hp3 = T(11) * u0 - T(7) * um1 + T(2) * um2;
  const U hp = mol_div(op1 * hp1 + op2 * hp2 + op3 * hp3, wpd);
  const U hm = mol_div(om1 * hm1 + om2 * hm2 + om3 * hm3, wmd);
  return (hp - hm) * (inv_dx * T(1.0 / 6.0));
}

__device__ __forceinline__ T mol_weno5(T um2, T um1, T u0, T up1, T up2, T eps, T inv_dx) {
  return mol_weno5_t<T>(um2, um1, u0, up1, up2, eps, inv_dx);
}
__device__ __forceinline__ T mol_weno5_ref(T um2, T um1, T u0, T up1, T up2, T eps, T inv_dx) {
  return mol_weno5_ref_t<T>(um2, um1, u0, up1, up2, eps, inv_dx);
}

// Forward-mode dual number (value, derivative) for the gradient of the WENO reconstruction.
struct MolDual {
  T v, d;
};
__device__ __forceinline__ MolDual operator+(MolDual a, MolDual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ MolDual operator-(MolDual a, MolDual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ MolDual operator*(MolDual a, MolDual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ MolDual operator/(MolDual a, MolDual b) {
  const T q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ MolDual operator+(MolDual a, T b) { return {a.v + b, a.d}; }
__device__ __forceinline__ MolDual operator+(T a, MolDual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ MolDual operator*(T a, MolDual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ MolDual operator*(MolDual a, T b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ MolDual operator/(MolDual a, T b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ MolDual operator/(T a, MolDual b) {
  const T q = a / b.v;
  return {q, -q * b.d / b.v};
}
__device__ __forceinline__ MolDual mol_rcp(const MolDual x) {
  const T r = mol_rcp(x.v);
  return {r, -(x.d * r) * r};
}
__device__ __forceinline__ MolDual mol_div(const MolDual n, const MolDual d) { return n / d; }
__device__ __forceinline__ MolDual mol_quot(const MolDual n, const MolDual d) { return n / d; }

// Gradient of the WENO5 node derivative with respect to its five inputs.
__device__ __noinline__ void mol_weno5_grad(T um2, T um1, T u0, T up1, T up2, T eps, T inv_dx, T* g, bool cleared) {
  const T x[5] = {um2, um1, u0, up1, up2};
  for (int k = 0; k < 5; ++k) {
    MolDual d[5];
    for (int j = 0; j < 5; ++j) d[j] = MolDual{x[j], j == k ? T(1) : T(0)};
    const MolDual r = cleared ? mol_weno5_t<MolDual>(d[0], d[1], d[2], d[3], d[4], eps, inv_dx)
                              : mol_weno5_ref_t<MolDual>(d[0], d[1], d[2], d[3], d[4], eps, inv_dx);
    g[k] = r.d;
  }
}

// ---- node-centred WENO5 on nonuniform coordinates (weno_f_nonuniform,
//      src/discretization/schemes/WENO/nonuniform_weno.jl): Fornberg 3-point sub-stencil weights
//      (:5-46), Simpson smoothness indicators (:54-73), target geometry (:76-82), closed-form
//      ideal weights (:85-118), Shi-Hu-Shu splitting with theta = 3 (:120-163).  Coordinates and
//      weights in FP64, values in U (T, or MolDual for the gradient).
__device__ __forceinline__ void mol_fornberg3(const double a0, const double a1, const double a2, const double xt,
                                              double* m0, double* m1, double* m2) {
  double n1m0 = 1.0, n1m1 = 0.0, n1m2 = 0.0, c1 = 1.0;
  double c2 = a1 - a0;
  const double r1 = c1 / c2;
  const double tA = a1 - xt;
  const double a1m0 = (tA * n1m0) / c2, a1m1 = (tA * n1m1 - n1m0) / c2, a1m2 = (tA * n1m2 - 2.0 * n1m1) / c2;
  const double sB = a0 - xt;
  const double a2m0 = r1 * (-(sB)*n1m0), a2m1 = r1 * (n1m0 - sB * n1m1), a2m2 = r1 * (2.0 * n1m1 - sB * n1m2);
  c1 = c2;
  n1m0 = a1m0;
  n1m1 = a1m1;
  n1m2 = a1m2;
  c2 = (a2 - a0) * (a2 - a1);
  const double r2 = c1 / c2;
  const double c3a = a2 - a0, c3b = a2 - a1, tA2 = a2 - xt;
  m0[0] = (tA2 * n1m0) / c3a;
  m1[0] = (tA2 * n1m1 - n1m0) / c3a;
  m2[0] = (tA2 * n1m2 - 2.0 * n1m1) / c3a;
  m0[1] = (tA2 * a2m0) / c3b;
  m1[1] = (tA2 * a2m1 - a2m0) / c3b;
  m2[1] = (tA2 * a2m2 - 2.0 * a2m1) / c3b;
  const double sB2 = a1 - xt;
  m0[2] = r2 * (-(sB2)*a2m0);
  m1[2] = r2 * (a2m0 - sB2 * a2m1);
  m2[2] = r2 * (2.0 * a2m1 - sB2 * a2m2);
}

__device__ __forceinline__ double mol_val(double v) { return v; }
__device__ __forceinline__ float mol_val(float v) { return v; }
__device__ __forceinline__ T mol_val(MolDual v) { return v.v; }

template <class U>
__device__ __forceinline__ U mol_dot3(const double* w, U a, U b, U c) {
  return (T)w[0] * a + (T)w[1] * b + (T)w[2] * c;
}

// The coordinate-only part of the nonuniform WENO5 derivative at one node (target TG of the
// stencil x1..x5), MOL_NU_NC values: for each sub-stencil k = 0, 1, 2 the weights of r_k
// (derivative at the target) and of the interval values pL, pM, pR (first derivative at the left
// end, midpoint and right end) and pp (second derivative), 15 per sub-stencil; then dx/6, dx,
// dx^3 of the Simpson smoothness indicator, the split ideal weights dp_k/sgp and dm_k/sgm, and
// sgp, sgm.  mol_weno5_nu_apply turns them and the five values into the derivative.
constexpr int MOL_NU_NC = 56;

__device__ __forceinline__ void mol_nu_sub(const double a0, const double a1, const double a2, const double xi,
                                           const double xL, const double xM, const double xph, double* c) {
  double m0[3], t2[3];
  mol_fornberg3(a0, a1, a2, xi, m0, c + 0, t2);
  mol_fornberg3(a0, a1, a2, xL, m0, c + 3, t2);
  mol_fornberg3(a0, a1, a2, xM, m0, c + 6, c + 12);
  mol_fornberg3(a0, a1, a2, xph, m0, c + 9, t2);
}

template <int TG>
__device__ __forceinline__ void mol_weno5_nu_coefs(const double x1, const double x2, const double x3, const double x4,
                                                   const double x5, double* c) {
  double xi, xL, xph, d0, d2;
  if (TG == 1) { xi = x1; xL = x1; xph = (x1 + x2) / 2; }
  else if (TG == 2) { xi = x2; xL = (x1 + x2) / 2; xph = (x2 + x3) / 2; }
  else if (TG == 4) { xi = x4; xL = (x3 + x4) / 2; xph = (x4 + x5) / 2; }
  else if (TG == 5) { xi = x5; xL = (x4 + x5) / 2; xph = x5; }
  else { xi = x3; xL = (x2 + x3) / 2; xph = (x3 + x4) / 2; }
  if (TG == 1) {
    d0 = ((2 * x1 - x2 - x3) * (x1 - x4) * (x1 - x5) + (x1 - x3) * (x1 - x5) * (x1 - x2) +
          (x1 - x3) * (x1 - x4) * (x1 - x2)) / ((2 * x1 - x2 - x3) * (x1 - x4) * (x1 - x5));
    d2 = ((x1 - x3) * (x1 - x4) * (x1 - x2)) / ((-x1 + x5) * (2 * x1 - x3 - x4) * (-x2 + x5));
  } else if (TG == 2) {
    d0 = ((x2 - x4) * (x2 - x5)) / ((x1 - x4) * (x1 - x5));
    d2 = ((-x1 + x2) * (x2 - x3) * (x2 - x4)) / ((-x1 + x5) * (2 * x2 - x3 - x4) * (-x2 + x5));
  } else if (TG == 4) {
    d0 = ((-x2 + x4) * (-x3 + x4) * (x4 - x5)) / ((x1 - x4) * (x1 - x5) * (-x2 - x3 + 2 * x4));
    d2 = ((-x1 + x4) * (-x2 + x4)) / ((-x1 + x5) * (-x2 + x5));
  } else if (TG == 5) {
    d0 = ((-x2 + x5) * (-x3 + x5) * (-x4 + x5)) / ((x1 - x4) * (x1 - x5) * (-x2 - x3 + 2 * x5));
    d2 = ((-x1 - x4 + 2 * x5) * (-x2 + x5) * (-x3 + x5) + (-x1 + x5) * (-x2 - x3 + 2 * x5) * (-x4 + x5)) /
         ((-x1 + x5) * (-x2 + x5) * (-x3 - x4 + 2 * x5));
  } else {
    d0 = ((x3 - x4) * (x3 - x5)) / ((x1 - x4) * (x1 - x5));
    d2 = ((x3 - x1) * (x3 - x2)) / ((x5 - x1) * (x5 - x2));
  }
  const double dx = xph - xL, xM = (xL + xph) / 2;
  mol_nu_sub(x1, x2, x3, xi, xL, xM, xph, c + 0);
  mol_nu_sub(x2, x3, x4, xi, xL, xM, xph, c + 15);
  mol_nu_sub(x3, x4, x5, xi, xL, xM, xph, c + 30);
  c[45] = dx / 6;
  c[46] = dx;
  c[47] = dx * dx * dx;
  const double d1 = 1.0 - d0 - d2, th = 3.0, half = 0.5;
  const double dp0 = half * (d0 + th * fabs(d0)), dp1 = half * (d1 + th * fabs(d1)), dp2 = half * (d2 + th * fabs(d2));
  const double dm0 = dp0 - d0, dm1 = dp1 - d1, dm2 = dp2 - d2;
  const double sgp = dp0 + dp1 + dp2, sgm = dm0 + dm1 + dm2;
  c[48] = dp0 / sgp;
  c[49] = dp1 / sgp;
  c[50] = dp2 / sgp;
  c[51] = dm0 / sgm;
  c[52] = dm1 / sgm;
  c[53] = dm2 / sgm;
  c[54] = sgp;
  c[55] = sgm;
}

template <class U>
__device__ __forceinline__ void mol_nu_beta_r(const double* c, const double* k, U ua, U ub, U uc, U& beta, U& r) {
  r = mol_dot3(c + 0, ua, ub, uc);
  const U pL = mol_dot3(c + 3, ua, ub, uc), pM = mol_dot3(c + 6, ua, ub, uc), pR = mol_dot3(c + 9, ua, ub, uc);
  const U pp = mol_dot3(c + 12, ua, ub, uc);
  const U I1 = (T)k[0] * (pL * pL + T(4) * (pM * pM) + pR * pR);
  const U I2 = (T)k[1] * (pp * pp);
  const U v = (T)k[1] * I1 + (T)k[2] * I2;
  beta = mol_val(v) > T(0) ? v : U{T(0)};
}

template <class U>
__device__ __forceinline__ U mol_weno5_nu_apply(U u1, U u2, U u3, U u4, U u5, const double* c, const T eps) {
  U b0, b1, b2, r0, r1, r2;
  mol_nu_beta_r(c + 0, c + 45, u1, u2, u3, b0, r0);
  mol_nu_beta_r(c + 15, c + 45, u2, u3, u4, b1, r1);
  mol_nu_beta_r(c + 30, c + 45, u3, u4, u5, b2, r2);
  const U e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  const U q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const U ap0 = (T)c[48] / q0, ap1 = (T)c[49] / q1, ap2 = (T)c[50] / q2;
  const U sp = ap0 + ap1 + ap2;
  const U am0 = (T)c[51] / q0, am1 = (T)c[52] / q1, am2 = (T)c[53] / q2;
  const U sm = am0 + am1 + am2;
  const U Rp = (ap0 / sp) * r0 + (ap1 / sp) * r1 + (ap2 / sp) * r2;
  const U Rm = (am0 / sm) * r0 + (am1 / sm) * r1 + (am2 / sm) * r2;
  return (T)c[54] * Rp - (T)c[55] * Rm;
}

template <int TG, class U>
__device__ __forceinline__ U mol_weno5_nu_t(U u1, U u2, U u3, U u4, U u5, const double x1, const double x2,
                                            const double x3, const double x4, const double x5, const T eps) {
  double c[MOL_NU_NC];
  mol_weno5_nu_coefs<TG>(x1, x2, x3, x4, x5, c);
  return mol_weno5_nu_apply<U>(u1, u2, u3, u4, u5, c, eps);
}

template <int TG>
__device__ __forceinline__ T mol_weno5_nu(T u1, T u2, T u3, T u4, T u5, double x1, double x2, double x3, double x4,
                                          double x5, T eps) {
  return mol_weno5_nu_t<TG, T>(u1, u2, u3, u4, u5, x1, x2, x3, x4, x5, eps);
}

// The FP64 derivative from the coefficients with the normalised nonlinear weights cleared, as
// in the uniform case: alpha_k / sum_j alpha_j = g_k P_k / sum_j g_j P_j with P_k the product of
// the other two q_j, so sgp Rp - sgm Rm = (sgp Np Dm - sgm Nm Dp) / (Dp Dm) costs one quotient
// (mol_quot) instead of twelve divisions.  The q_j are at least eps^2, so the products of four
// stay far above the FP64 underflow threshold (FP32 keeps the normalised form).
__device__ __forceinline__ double mol_weno5_nu_cleared(double u1, double u2, double u3, double u4, double u5,
                                                       const double* c, double eps) {
  double b0, b1, b2, r0, r1, r2;
  mol_nu_beta_r<double>(c + 0, c + 45, u1, u2, u3, b0, r0);
  mol_nu_beta_r<double>(c + 15, c + 45, u2, u3, u4, b1, r1);
  mol_nu_beta_r<double>(c + 30, c + 45, u3, u4, u5, b2, r2);
  const double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const double P0 = q1 * q2, P1 = q0 * q2, P2 = q0 * q1;
  const double A0 = c[48] * P0, A1 = c[49] * P1, A2 = c[50] * P2;
  const double B0 = c[51] * P0, B1 = c[52] * P1, B2 = c[53] * P2;
  const double Dp = A0 + A1 + A2, Dm = B0 + B1 + B2;
  const double Np = A0 * r0 + A1 * r1 + A2 * r2, Nm = B0 * r0 + B1 * r1 + B2 * r2;
  return mol_quot(c[54] * Np * Dm - c[55] * Nm * Dp, Dp * Dm);
}

// The same derivative from a node's row of precomputed coefficients (mol_weno_nu_table).
__device__ __forceinline__ T mol_weno5_nu_tab(T u1, T u2, T u3, T u4, T u5, const double* c, T eps) {
  if constexpr (sizeof(T) == 8) return (T)mol_weno5_nu_cleared(u1, u2, u3, u4, u5, c, eps);
  else return mol_weno5_nu_apply<T>(u1, u2, u3, u4, u5, c, eps);
}

// Coordinate of (unwrapped) node k of a periodic axis: the image across an end is continued by
// the period x[n] - x[1] (bcoord, src/discretization/interface_boundary.jl:120-151).
__device__ __forceinline__ double mol_pcoord(const double* x, const int k, const int n) {
  if (k <= 1) return x[k + n - 2] - (x[n - 1] - x[0]);
  if (k > n) return x[k - n] + (x[n - 1] - x[0]);
  return x[k - 1];
}

// Gradient of the nonuniform WENO5 derivative with respect to its five values.
template <int TG>
__device__ __noinline__ void mol_weno5_nu_grad(T u1, T u2, T u3, T u4, T u5, double x1, double x2, double x3,
                                               double x4, double x5, T eps, T* g) {
  const T x[5] = {u1, u2, u3, u4, u5};
  for (int k = 0; k < 5; ++k) {
    MolDual d[5];
    for (int j = 0; j < 5; ++j) d[j] = MolDual{x[j], j == k ? T(1) : T(0)};
    g[k] = mol_weno5_nu_t<TG, MolDual>(d[0], d[1], d[2], d[3], d[4], x1, x2, x3, x4, x5, eps).d;
  }
}

// The gradient from a node's row of precomputed coefficients.
__device__ __noinline__ void mol_weno5_nu_tab_grad(T u1, T u2, T u3, T u4, T u5, const double* c, T eps, T* g) {
  const T x[5] = {u1, u2, u3, u4, u5};
  for (int k = 0; k < 5; ++k) {
    MolDual d[5];
    for (int j = 0; j < 5; ++j) d[j] = MolDual{x[j], j == k ? T(1) : T(0)};
    g[k] = mol_weno5_nu_apply<MolDual>(d[0], d[1], d[2], d[3], d[4], c, eps).d;
  }
}

// Row of coefficients (MOL_NU_NC values) of node i = 1..n of an axis with coordinates x: the
// stencil and target of the generated code (one-sided at nodes 1, 2, n-1, n of a non-periodic
// axis, centred elsewhere; periodic images through mol_pcoord).
__device__ __forceinline__ void mol_weno_nu_row(const double* x, const int n, const bool periodic, const int i,
                                                double* c) {
  int first = -2, target = 3;
  if (!periodic) {
    if (i <= 2) {
      target = i;
      first = 1 - i;
    } else if (i >= n - 1) {
      target = 5 - (n - i);
      first = n - 4 - i;
    }
  }
  double xs[5];
  for (int q = 0; q < 5; ++q) {
    const int k = i + first + q;
    xs[q] = periodic ? mol_pcoord(x, k, n) : x[k - 1];
  }
  switch (target) {
    case 1: mol_weno5_nu_coefs<1>(xs[0], xs[1], xs[2], xs[3], xs[4], c); break;
    case 2: mol_weno5_nu_coefs<2>(xs[0], xs[1], xs[2], xs[3], xs[4], c); break;
    case 4: mol_weno5_nu_coefs<4>(xs[0], xs[1], xs[2], xs[3], xs[4], c); break;
    case 5: mol_weno5_nu_coefs<5>(xs[0], xs[1], xs[2], xs[3], xs[4], c); break;
    default: mol_weno5_nu_coefs<3>(xs[0], xs[1], xs[2], xs[3], xs[4], c); break;
  }
}

// Hardware floating-point atomic add on global memory.
__device__ __forceinline__ void mol_atomic_add(T* p, T v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void mol_atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }

// Work-groups are dealt to the 8 XCDs round-robin; remap so that each XCD runs one contiguous
// range of tiles (neighbouring tiles share halo lines in that XCD's L2).
template <int NBLK>
__device__ __forceinline__ int mol_xcd_tile(int b) {
  constexpr int NX = 8;
  if (NBLK < 2 * NX) return b;
  constexpr int q = NBLK / NX, rem = NBLK % NX;
  const int x = b % NX, i = b / NX;
  return x * q + (x < rem ? x : rem) + i;
}

// Coordinates (x, y, z) of position q of an EX x EY x ... box, advanced by a fixed stride
// without divisions (the staging loops visit q, q + S, q + 2S, ...).
template <int EX, int EY, int S>
struct MolWalk {
  int x, y, z;
  __device__ __forceinline__ explicit MolWalk(const int q) {
    z = q / (EX * EY);
    const int r = q - z * (EX * EY);
    y = r / EX;
    x = r - y * EX;
  }
  __device__ __forceinline__ void step() {
    constexpr int DZ = S / (EX * EY), R = S % (EX * EY), DY = R / EX, DX = R % EX;
    x += DX;
    y += DY;
    z += DZ;
    if (x >= EX) { x -= EX; ++y; }
    if (y >= EY) { y -= EY; ++z; }
  }
};

// Staging mode of a tile in variant G: 0 plain loads, 1 plain loads after the periodic wrap,
// 2 through the ghost-node closures (tile + halo leaves a variable's unknown box).
template <class G>
__device__ __forceinline__ int mol_stage_mode(const int tx, const int ty, const int tz) {
  if (tx >= G::FLO0 && tx <= G::FHI0 && ty >= G::FLO1 && ty <= G::FHI1 && tz >= G::FLO2 && tz <= G::FHI2) return 0;
  if (tx >= G::SLO0 && tx <= G::SHI0 && ty >= G::SLO1 && ty <= G::SHI1 && tz >= G::SLO2 && tz <= G::SHI2) return 1;
  return 2;
}

// Position (x, y, z) of the tile + halo box lies outside the tile along two or more axes and
// variant G reads no such position (G::CROSS): it is not staged.
template <class G>
__device__ __forceinline__ bool mol_unread(const int x, const int y, const int z) {
  if (!G::CROSS) return false;
  return (int)(x < G::HX || x >= G::HX + G::TX) + (int)(y < G::HY || y >= G::HY + G::TY) +
             (int)(z < G::HZ || z >= G::HZ + G::TZ) >= 2;
}

// Tile + halo of the staged arrays into LDS.
template <class G>
__device__ __forceinline__ void mol_stage_tile(T* L, const int o0, const int o1, const int o2, const int mode,
                                               const double t, const T* __restrict__ P, const T* __restrict__ u,
                                               const JPtrs& jp) {
  constexpr int IT = (G::VOL + G::NTHR - 1) / G::NTHR;   // positions per thread
  MolWalk<G::EX, G::EY, G::NTHR> w((int)threadIdx.x);
  if constexpr (G::NST > 0 && IT * G::NST <= 16) {
    // every load of the thread is issued before its first LDS store: one exposed memory
    // latency per tile instead of one per position
    T v[IT][G::NST];
    bool ok[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k, w.step()) {
      const int q = (int)threadIdx.x + k * G::NTHR;
      ok[k] = (G::VOL % G::NTHR == 0 || q < G::VOL) && !mol_unread<G>(w.x, w.y, w.z);
      if (ok[k]) G::stage_vals(v[k], o0 + w.x - G::HX, o1 + w.y - G::HY, o2 + w.z - G::HZ, mode, t, P, u, jp);
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int q = (int)threadIdx.x + k * G::NTHR;
      if (ok[k])
#pragma unroll
        for (int s = 0; s < G::NST; ++s) L[s * G::VOL + q] = v[k][s];
    }
  } else {
    for (int q = threadIdx.x; q < G::VOL; q += G::NTHR, w.step())
      if (!mol_unread<G>(w.x, w.y, w.z))
        G::stage_pos(L, q, o0 + w.x - G::HX, o1 + w.y - G::HY, o2 + w.z - G::HZ, mode, t, P, u, jp);
  }
}

// Node phase of one staged tile in variant G: each thread runs G::RB consecutive nodes along
// axis G::BA; their node functions see the same LDS tile at offsets that differ by a constant,
// so shared taps and subexpressions are loaded and computed once.  Nodes of a ragged tile past
// the unknown box run at the last node of the box (every load stays inside the tile + halo
// and the arrays) and store nothing.
template <class G>
__device__ __forceinline__ void mol_tile_nodes(const T* L, const int tx, const int ty, const int tz,
                                               const T* __restrict__ u, const Epi& ep, const JPtrs& jp,
                                               const double t) {
  const int o0 = G::ULO0 + tx * G::TX, o1 = G::ULO1 + ty * G::TY, o2 = G::ULO2 + tz * G::TZ;
  constexpr int RB = G::RB;
  constexpr int PX = G::TX, PY = G::BA == 1 ? G::TY / RB : G::TY, PZ = G::BA == 2 ? G::TZ / RB : G::TZ;
  constexpr int PXY = PX * PY, PV = PXY * PZ;
  constexpr int STEP = G::BA == 1 ? G::EX : (G::BA == 2 ? G::EX * G::EY : 1);
#pragma unroll
  for (int r = 0; r < (PV + G::NTHR - 1) / G::NTHR; ++r) {
    const int q = (int)threadIdx.x + r * G::NTHR;
    if (PV % G::NTHR != 0 && q >= PV) break;
    const int lz0 = q / PXY;
    const int rr = q - lz0 * PXY;
    const int ly0 = rr / PX;
    const int lx = rr - ly0 * PX;
    const int ly = G::BA == 1 ? ly0 * RB : ly0, lz = G::BA == 2 ? lz0 * RB : lz0;
    const int base = (lx + G::HX) + G::EX * (ly + G::HY) + G::EX * G::EY * (lz + G::HZ);
    T out[RB][G::NF];
    int J[RB][3];
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      J[k][0] = o0 + lx;
      J[k][1] = o1 + ly + (G::BA == 1 ? k : 0);
      J[k][2] = o2 + lz + (G::BA == 2 ? k : 0);
      G::node(L, base + k * STEP, min(J[k][0], (int)G::UHI0), min(J[k][1], (int)G::UHI1), min(J[k][2], (int)G::UHI2), t,
              ep.p, u, jp, out[k]);
    }
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      if (J[k][0] > G::UHI0 || J[k][1] > G::UHI1 || J[k][2] > G::UHI2) continue;
#pragma unroll
      for (int f = 0; f < G::NF; ++f)
        if (G::inbox(f, J[k][0], J[k][1], J[k][2])) mol_epilogue(ep, G::uidx(f, J[k][0], J[k][1], J[k][2]), out[k][f]);
    }
  }
}

// Loads of the tile + halo of variant G into registers R[i][slot] (position threadIdx.x + i *
// NTHR): issued for the next tile while the current one runs its node phase.
template <class G, int NPF, int NSR>
__device__ __forceinline__ void mol_fetch(T (&R)[NPF][NSR], const int tx, const int ty, const int tz, const double t,
                                          const T* __restrict__ P, const T* __restrict__ u, const JPtrs& jp) {
  const int o0 = G::ULO0 + tx * G::TX, o1 = G::ULO1 + ty * G::TY, o2 = G::ULO2 + tz * G::TZ;
  const int mode = mol_stage_mode<G>(tx, ty, tz);
  MolWalk<G::EX, G::EY, G::NTHR> w((int)threadIdx.x);
#pragma unroll
  for (int i = 0; i < NPF; ++i) {
    if (i * G::NTHR >= G::VOL) break;
    if ((int)threadIdx.x + i * G::NTHR < G::VOL) {
      if (mol_unread<G>(w.x, w.y, w.z)) {
#pragma unroll
        for (int s = 0; s < NSR; ++s) R[i][s] = T(0);
      } else {
        G::stage_vals(R[i], o0 + w.x - G::HX, o1 + w.y - G::HY, o2 + w.z - G::HZ, mode, t, P, u, jp);
      }
    }
    w.step();
  }
}

template <class G, int NPF, int NSR>
__device__ __forceinline__ void mol_put(T* L, const T (&R)[NPF][NSR]) {
#pragma unroll
  for (int i = 0; i < NPF; ++i) {
    const int q = (int)threadIdx.x + i * G::NTHR;
    if (i * G::NTHR >= G::VOL) break;
    if (q < G::VOL)
#pragma unroll
      for (int s = 0; s < G::NST; ++s) L[s * G::VOL + q] = R[i][s];
  }
}

// One tile of the union unknown box, in the tile variant G (interior-rows core variant or the
// boundary variant, each with its own halo): stage (mol_stage_tile), then the node phase.
template <class G>
__device__ __forceinline__ void mol_full_tile(T* L, const int tx, const int ty, const int tz, const T* __restrict__ u,
                                              const Epi& ep, const JPtrs& jp, const double t) {
  if (G::NST > 0) {
    const int o0 = G::ULO0 + tx * G::TX, o1 = G::ULO1 + ty * G::TY, o2 = G::ULO2 + tz * G::TZ;
    mol_stage_tile<G>(L, o0, o1, o2, mol_stage_mode<G>(tx, ty, tz), t, ep.p, u, jp);
    __syncthreads();
  }
  mol_tile_nodes<G>(L, tx, ty, tz, u, ep, jp, t);
}

// Persistent loop over the tiles of a launch of gridDim.x work-groups (a multiple of 8 when
// smaller than the tile count, so that tile k + gridDim.x stays on the XCD of tile k): the
// staging loads of the next tile are in flight while the current tile runs its node phase.
// GC: interior-rows variant (used when HASC and the tile is a core tile), GB: boundary variant;
// both stage the same arrays.
template <class GC, class GB, bool HASC, int NPF>
__device__ __forceinline__ void mol_tiles_persistent(T* L, const T* __restrict__ u, const Epi& ep, const JPtrs& jp,
                                                     const double t) {
  constexpr int NSR = GB::NST > 0 ? GB::NST : 1;
  constexpr int NBLK = GB::NBLK;
  int k = (int)blockIdx.x;
  if (k >= NBLK) return;
  auto decode = [](const int lin, int& x, int& y, int& z) {
    int b = mol_xcd_tile<NBLK>(lin);
    x = b % GB::NT0;
    b /= GB::NT0;
    y = b % GB::NT1;
    z = b / GB::NT1;
  };
  // more than 16 staging values per thread would cost occupancy: stage each tile in turn
  if (NPF * NSR > 16) {
    for (; k < NBLK; k += (int)gridDim.x) {
      int tx, ty, tz;
      decode(k, tx, ty, tz);
      if (HASC && GB::is_core(tx, ty, tz)) mol_full_tile<GC>(L, tx, ty, tz, u, ep, jp, t);
      else mol_full_tile<GB>(L, tx, ty, tz, u, ep, jp, t);
      if (GB::NST > 0 && k + (int)gridDim.x < NBLK) __syncthreads();
    }
    return;
  }
  T R[NPF][NSR];
  int tx, ty, tz;
  decode(k, tx, ty, tz);
  bool core = HASC && GB::is_core(tx, ty, tz);
  if (GB::NST > 0) {
    if (HASC && core) mol_fetch<GC, NPF, NSR>(R, tx, ty, tz, t, ep.p, u, jp);
    else mol_fetch<GB, NPF, NSR>(R, tx, ty, tz, t, ep.p, u, jp);
  }
  while (true) {
    if (GB::NST > 0) {
      if (HASC && core) mol_put<GC, NPF, NSR>(L, R);
      else mol_put<GB, NPF, NSR>(L, R);
      __syncthreads();
    }
    const int kn = k + (int)gridDim.x;
    int nx = 0, ny = 0, nz = 0;
    bool ncore = false;
    if (kn < NBLK) {
      decode(kn, nx, ny, nz);
      ncore = HASC && GB::is_core(nx, ny, nz);
      if (GB::NST > 0) {
        if (HASC && ncore) mol_fetch<GC, NPF, NSR>(R, nx, ny, nz, t, ep.p, u, jp);
        else mol_fetch<GB, NPF, NSR>(R, nx, ny, nz, t, ep.p, u, jp);
      }
    }
    if (HASC && core) mol_tile_nodes<GC>(L, tx, ty, tz, u, ep, jp, t);
    else mol_tile_nodes<GB>(L, tx, ty, tz, u, ep, jp, t);
    if (kn >= NBLK) break;
    if (GB::NST > 0) __syncthreads();
    k = kn;
    tx = nx;
    ty = ny;
    tz = nz;
    core = ncore;
  }
}

// ---------------------------------------------------------------------------------------
// Fused classic RK4 step: stage S of tile (o0, o1, o2) on the tile grown by (4 - S) reach per
// axis.  The items of a stage are the tile's nodes (0 .. TVOL-1) and then the frame of the
// grown region; item q belongs to thread q mod NTHR, so every tile node keeps its thread across
// the four stages and its step-initial value Ur and running sum Ar = k1 + 2 k2 + 2 k3 stay in
// that thread's registers.  Lin holds the stage input (u, y1, y2, y3), Lout receives
// y_S = u + c_S dt k_S; stage 4 writes u + dt/6 k4 + dt/6 Ar.  The arithmetic is that of the
// stage kernels' epilogue (mode 1 with ca = 1, acc_f = 1 or 2, acc_keep = 0 or 1), so a fused
// step equals four stage launches.
template <class G>
struct MolFusedItems {
  static constexpr int TVOL = G::TX * G::TY * G::TZ;
  static constexpr int NJ = (TVOL + G::NTHR - 1) / G::NTHR;   // tile nodes per thread, NTHR apart
  static __device__ __forceinline__ void pos(const int j, int& lx, int& ly) {
    const int q = (int)threadIdx.x + j * G::NTHR;
    ly = q / G::TX;
    lx = q - ly * G::TX;
  }
  static __device__ __forceinline__ bool valid(const int j) {
    return TVOL % G::NTHR == 0 || (int)threadIdx.x + j * G::NTHR < TVOL;
  }
};

// Frame items of the region grown by M: the square (tile + M reach per axis) less the tile,
// or, for 2D cross-shaped stencils of reach one on periodic axes (G::DIAM), the positions
// within M steps of the tile along the axes (L1 distance), which are all that the remaining
// stages read: 2 M (TX + TY) + 2 M (M - 1) items instead of 2 M (TX + TY) + 4 M^2.
template <class G, int M>
struct MolFrame2 {
  static constexpr int N = G::DIAM ? 2 * M * (G::TX + G::TY) + 2 * M * (M - 1)
                                   : (G::TX + 2 * M * G::RX) * (G::TY + 2 * M * G::RY) - G::TX * G::TY;
};

// Position (tile-relative) of frame item h of the region grown by M (2D / 1D: RZ = 0).
template <class G, int M>
__device__ __forceinline__ void mol_frame_pos(const int h, int& lx, int& ly) {
  if constexpr (G::DIAM) {
    // rows above the tile from the outermost (ly = -M, TX wide) inwards (ly = -d spans
    // TX + 2 (M - d) positions centred on the tile), the same rows below, then M positions on
    // either side of each tile row
    constexpr int TOP = M * G::TX + M * (M - 1);
    if (h < 2 * TOP) {
      const bool below = h >= TOP;
      int r = below ? h - TOP : h;
      int d = M;
#pragma unroll
      for (int k = M; k > 1; --k) {
        const int w = G::TX + 2 * (M - k);
        if (d == k && r >= w) {
          r -= w;
          d = k - 1;
        }
      }
      ly = below ? G::TY - 1 + d : -d;
      lx = -(M - d) + r;
    } else {
      const int h3 = h - 2 * TOP;
      const int c = h3 % (2 * M);
      ly = h3 / (2 * M);
      lx = c < M ? -M + c : G::TX + (c - M);
    }
    return;
  }
  constexpr int MR0 = M * G::RX, MR1 = M * G::RY, W0 = G::TX + 2 * MR0;
  if (h < MR1 * W0) {
    ly = -MR1 + h / W0;
    lx = -MR0 + h % W0;
  } else if (h < 2 * MR1 * W0) {
    const int h2 = h - MR1 * W0;
    ly = G::TY + h2 / W0;
    lx = -MR0 + h2 % W0;
  } else {
    const int h3 = h - 2 * MR1 * W0;
    const int c = h3 % (2 * MR0);
    ly = h3 / (2 * MR0);
    lx = c < MR0 ? -MR0 + c : G::TX + (c - MR0);
  }
}

// Node-invariant values of the tile nodes (G::NINV per node: what depends only on constants,
// parameters and coordinates), computed once per launch and read by G::node_h in every stage.
template <class G>
struct MolFusedInv {
  static constexpr int NI = G::NINV > 0 ? G::NINV : 1;
  T v[MolFusedItems<G>::NJ][NI];
  __device__ __forceinline__ void init(const int o0, const int o1, const int o2, const T* __restrict__ P,
                                       const JPtrs& jp) {
    if constexpr (G::NINV > 0) {
#pragma unroll
      for (int j = 0; j < MolFusedItems<G>::NJ; ++j) {
        int lx, ly;
        MolFusedItems<G>::pos(j, lx, ly);
        if (MolFusedItems<G>::valid(j)) G::node_inv(o0 + lx, o1 + ly, o2, P, jp, v[j]);
      }
    }
  }
};

template <class G, int S, int M = 4 - S, int ME = 0>
__device__ __forceinline__ void mol_rk4_stage(const T* Lin, T* Lout, T* U, T (&Ur)[MolFusedItems<G>::NJ][G::NF],
                                              T (&Ar)[MolFusedItems<G>::NJ][G::NF], const MolFusedInv<G>& Iv,
                                              const int o0, const int o1,
                                              const int o2, const T* __restrict__ u, T* __restrict__ unew,
                                              const Epi& ep, const JPtrs& jp, const double t, const double dt,
                                              T* Ukeep = nullptr, const bool store = true, T* Ac = nullptr) {
  // M: the stage region is the tile grown by M reach per axis.  ME: the step's result is
  // wanted on the tile grown by ME (ME > 0: the first of two steps in one launch, whose result
  // is the next step's tile + halo); the frame positions inside it keep their running sum in
  // Ac, and stage 4 (M == ME) leaves the new state in U, in place, at every position of its
  // region and in Ur for the tile nodes.
  // Ukeep (resident steps): stage 4 also leaves the new state in Ukeep and in Ur for the next
  // step; store: stage 4 writes it to unew
  static_assert(S < 4 ? M == ME + 4 - S : M == ME, "stage regions shrink by one reach per stage");
  constexpr int TVOL = MolFusedItems<G>::TVOL, NJ = MolFusedItems<G>::NJ;
  constexpr int FR = MolFrame2<G, M>::N;
  static_assert(G::TZ == 1 && G::RZ == 0, "fused RK4 steps are 1D / 2D");
  static_assert(!G::DIAM || (G::RX == 1 && G::RY == 1 && !G::FIX), "L1 stage regions: reach one, periodic");
  const double ts = t + (S == 1 ? 0.0 : (S == 4 ? dt : dt / 2));
  const double cf = S == 1 || S == 2 ? dt / 2 : (S == 3 ? dt : dt / 6);
  if (ME > 0 && S == 4) Ukeep = U;
  // tile nodes: registers carry u and the running sum
  T out[NJ][G::NF];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    if (!MolFusedItems<G>::valid(j)) break;
    int lx, ly;
    MolFusedItems<G>::pos(j, lx, ly);
    if constexpr (G::NINV > 0)
      G::node_h(Lin, (lx + G::HX) + G::EX * (ly + G::HY), o0 + lx, o1 + ly, o2, ts, ep.p, u, jp, Iv.v[j], out[j]);
    else
      G::node(Lin, (lx + G::HX) + G::EX * (ly + G::HY), o0 + lx, o1 + ly, o2, ts, ep.p, u, jp, out[j]);
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    if (!MolFusedItems<G>::valid(j)) break;
    int lx, ly;
    MolFusedItems<G>::pos(j, lx, ly);
    const int b = (lx + G::HX) + G::EX * (ly + G::HY);
    const int I0 = o0 + lx, I1 = o1 + ly, I2 = o2;
#pragma unroll
    for (int f = 0; f < G::NF; ++f) {
      T v = (T)cf * out[j][f];
      v += Ur[j][f];
      if (S < 4) {
        Lout[f * G::VOL + b] = v;
        T a = (S == 1 ? T(1) : T(2)) * out[j][f];
        if (S > 1) a += Ar[j][f];
        Ar[j][f] = a;
      } else {
        v += (T)cf * Ar[j][f];
        if (ME == 0 && store && I0 <= G::UHI0 && I1 <= G::UHI1 && I2 <= G::UHI2) unew[G::uidx(f, I0, I1, I2)] = v;
        if (Ukeep) {
          Ukeep[f * G::VOL + b] = v;
          Ur[j][f] = v;
          Ar[j][f] = T(0);
        }
      }
    }
  }
  // frame of the grown region: u from LDS, and inside the step's result region the running
  // sum from Ac
  if (M > 0) {
    for (int q = (int)threadIdx.x; q < TVOL + FR; q += G::NTHR) {
      if (q < TVOL) continue;
      int lx, ly;
      mol_frame_pos<G, M>(q - TVOL, lx, ly);
      const int b = (lx + G::HX) + G::EX * (ly + G::HY);
      T out[G::NF];
      G::node(Lin, b, o0 + lx, o1 + ly, o2, ts, ep.p, u, jp, out);
      const bool keep = ME > 0 && lx >= -ME * G::RX && lx < G::TX + ME * G::RX && ly >= -ME * G::RY &&
                        ly < G::TY + ME * G::RY;
#pragma unroll
      for (int f = 0; f < G::NF; ++f) {
        T v = (T)cf * out[f];
        v += U[f * G::VOL + b];
        if (S < 4) {
          Lout[f * G::VOL + b] = v;
          if (keep) {
            T a = (S == 1 ? T(1) : T(2)) * out[f];
            if (S > 1) a += Ac[f * G::VOL + b];
            Ac[f * G::VOL + b] = a;
          }
        } else {
          v += (T)cf * Ac[f * G::VOL + b];
          U[f * G::VOL + b] = v;
        }
      }
    }
  }
}

// After a stage of a fused step on a non-periodic system: the positions of the stage region
// (tile + M reach) outside the unknown box along a non-periodic axis get the value of their
// boundary equations from the stage values just written, at the next stage's time (what
// staging does for each stage launch of the unfused path).  Returns whether it wrote.
template <class G, int M>
__device__ __forceinline__ bool mol_rk4_fix_ghosts(T* L, const int o0, const int o1, const int o2, const double tn,
                                                   const T* __restrict__ P, const JPtrs& jp) {
  constexpr int W0 = G::TX + 2 * M * G::RX, W1 = G::TY + 2 * M * G::RY, W2 = G::TZ + 2 * M * G::RZ;
  constexpr int NV = W0 * W1 * W2;
  // the whole work-group takes the same branch: the region reaches outside the box or not
  const bool out0 = !G::PER0 && (o0 - M * G::RX < (int)G::ULO0 || o0 + G::TX - 1 + M * G::RX > (int)G::UHI0);
  const bool out1 = !G::PER1 && G::RY > 0 && (o1 - M * G::RY < (int)G::ULO1 || o1 + G::TY - 1 + M * G::RY > (int)G::UHI1);
  if (!out0 && !out1) return false;
  for (int q = threadIdx.x; q < NV; q += G::NTHR) {
    const int z = q / (W0 * W1);
    const int r = q - z * (W0 * W1);
    const int y = r / W0;
    const int x = r - y * W0;
    const int lx = x - M * G::RX, ly = y - M * G::RY, lz = z - M * G::RZ;
    const int p0 = o0 + lx, p1 = o1 + ly, p2 = o2 + lz;
    const bool ghost = (!G::PER0 && (p0 < (int)G::ULO0 || p0 > (int)G::UHI0)) ||
                       (!G::PER1 && G::RY > 0 && (p1 < (int)G::ULO1 || p1 > (int)G::UHI1));
    if (!ghost) continue;
    const int b = (lx + G::HX) + G::EX * ((ly + G::HY) + G::EY * (lz + G::HZ));
    T v[G::NF];
#pragma unroll
    for (int f = 0; f < G::NF; ++f) v[f] = G::stage_lds(f, p0, p1, p2, tn, P, L, o0 - G::HX, o1 - G::HY, o2 - G::HZ, jp);
#pragma unroll
    for (int f = 0; f < G::NF; ++f) L[f * G::VOL + b] = v[f];
  }
  return true;
}

// One RK4 step of a fused launch from U (tile + halo) at time t: stage regions of margin
// ME + 3 .. ME; ME = 4: the first of two steps, whose result (tile + 4 reach) stays in U with its
// ghosts refilled at t + dt; ME = 0: the last step, stored to unew.
template <class G, int ME>
__device__ __forceinline__ void mol_rk4_step(T* U, T* Ya, T* Yb, T* Ac, T (&Ur)[MolFusedItems<G>::NJ][G::NF],
                                             T (&Ar)[MolFusedItems<G>::NJ][G::NF], const MolFusedInv<G>& Iv,
                                             const int o0, const int o1, const int o2, const T* __restrict__ u,
                                             T* __restrict__ unew, const Epi& ep, const JPtrs& jp, const double t,
                                             const double dt) {
  mol_rk4_stage<G, 1, ME + 3, ME>(U, Ya, U, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt, nullptr, true, Ac);
  __syncthreads();
  if constexpr (G::FIX) {
    if (mol_rk4_fix_ghosts<G, ME + 3>(Ya, o0, o1, o2, t + dt / 2, ep.p, jp)) __syncthreads();
  }
  mol_rk4_stage<G, 2, ME + 2, ME>(Ya, Yb, U, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt, nullptr, true, Ac);
  __syncthreads();
  if constexpr (G::FIX) {
    if (mol_rk4_fix_ghosts<G, ME + 2>(Yb, o0, o1, o2, t + dt / 2, ep.p, jp)) __syncthreads();
  }
  mol_rk4_stage<G, 3, ME + 1, ME>(Yb, Ya, U, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt, nullptr, true, Ac);
  __syncthreads();
  if constexpr (G::FIX) {
    if (mol_rk4_fix_ghosts<G, ME + 1>(Ya, o0, o1, o2, t + dt, ep.p, jp)) __syncthreads();
  }
  mol_rk4_stage<G, 4, ME, ME>(Ya, nullptr, U, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt, nullptr, true, Ac);
  if constexpr (ME > 0) {
    __syncthreads();
    if constexpr (G::FIX) {
      if (mol_rk4_fix_ghosts<G, ME>(U, o0, o1, o2, t + dt, ep.p, jp)) __syncthreads();
    }
  }
}

// A fused launch of nsteps (1, or 2 when the halo is 8 reach wide) RK4 steps on one tile.  The
// second step starts at t + dt, the device clock's next value.
template <class G>
__device__ __forceinline__ void mol_rk4_fused_tile(T* U, T* Ya, T* Yb, T* Ac, double* fc, const int fcs, int* fo,
                                                   const int tx, const int ty, const int tz, const T* __restrict__ u,
                                                   T* __restrict__ unew, const Epi& ep, const JPtrs& jp,
                                                   const double t, const int nsteps) {
  const int o0 = G::ULO0 + tx * G::TX, o1 = G::ULO1 + ty * G::TY, o2 = G::ULO2 + tz * G::TZ;
  const double dt = ep.tadv;
  // coordinates of the tile + halo positions (periodic images wrapped), per axis: position c of
  // the concatenated extents EX, EY, EZ; loaded before the staging loads are awaited, so both
  // are in flight together
  constexpr int NCP = G::EX + G::EY + G::EZ, NCI = (NCP + G::NTHR - 1) / G::NTHR;
  double cv[NCI];
  int cdst[NCI];
#pragma unroll
  for (int k = 0; k < NCI; ++k) {
    const int c = (int)threadIdx.x + k * G::NTHR;
    const int a = c < G::EX ? 0 : (c < G::EX + G::EY ? 1 : 2);
    const int e = c - (a == 0 ? 0 : (a == 1 ? G::EX : G::EX + G::EY));
    const int org = a == 0 ? o0 - G::HX : (a == 1 ? o1 - G::HY : o2 - G::HZ);
    const int nn = a == 0 ? G::N0 : (a == 1 ? G::N1 : G::N2);
    const bool per = a == 0 ? G::PER0 != 0 : (a == 1 ? G::PER1 != 0 : G::PER2 != 0);
    cdst[k] = -1;
    cv[k] = 0.0;
    if (c < NCP && jp.coord[a]) {
      int p = org + e;
      if (per) p = mol_wrap(mol_wrap(p, nn), nn);
      cv[k] = jp.coord[a][min(max(p, 1), nn) - 1];
      cdst[k] = a * fcs + e;
    }
  }
  for (int a = (int)threadIdx.x; a < 3; a += G::NTHR) fo[a] = a == 0 ? o0 - G::HX : (a == 1 ? o1 - G::HY : o2 - G::HZ);
  mol_stage_tile<G>(U, o0, o1, o2, mol_stage_mode<G>(tx, ty, tz), t, ep.p, u, jp);
#pragma unroll
  for (int k = 0; k < NCI; ++k)
    if (cdst[k] >= 0) fc[cdst[k]] = cv[k];
  __syncthreads();
  constexpr int NJ = MolFusedItems<G>::NJ;
  T Ur[NJ][G::NF], Ar[NJ][G::NF];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    int lx, ly;
    MolFusedItems<G>::pos(j, lx, ly);
    const bool ok = MolFusedItems<G>::valid(j);
    const int b = ok ? (lx + G::HX) + G::EX * (ly + G::HY) : 0;
#pragma unroll
    for (int f = 0; f < G::NF; ++f) {
      Ur[j][f] = ok ? U[f * G::VOL + b] : T(0);
      Ar[j][f] = T(0);
    }
  }
  MolFusedInv<G> Iv;
  Iv.init(o0, o1, o2, ep.p, jp);
  if constexpr (G::HX >= 8 * G::RX && G::HY >= 8 * G::RY) {
    if (nsteps == 2) {
      mol_rk4_step<G, 4>(U, Ya, Yb, Ac, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt);
      mol_rk4_step<G, 0>(U, Ya, Yb, Ac, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t + dt, dt);
      return;
    }
  }
  mol_rk4_step<G, 0>(U, Ya, Yb, Ac, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt);
}

// Resident steps of a 1D system held by a single tile: nsub RK4 steps in one launch with the
// state in LDS (and, for the tile nodes, in registers).  Between steps the halo of U is
// refreshed: periodic images copied from the tile (which covers the whole ring), ghosts of a
// non-periodic axis from their closures at the new time.  The time advances by the same
// sequence of additions as the device clock of one launch per step.
template <class G>
__device__ __forceinline__ void mol_rk4_fused_resident(T* U, T* Ya, T* Yb, double* fc, const int fcs, int* fo,
                                                       const T* __restrict__ u, T* __restrict__ unew, const Epi& ep,
                                                       const JPtrs& jp, const double t0, const int nsub) {
  static_assert(G::NBLK == 1 && G::RY == 0 && G::TY == 1, "resident steps: one 1D tile");
  const int o0 = G::ULO0, o1 = G::ULO1, o2 = G::ULO2;
  const double dt = ep.tadv;
  mol_stage_tile<G>(U, o0, o1, o2, mol_stage_mode<G>(0, 0, 0), t0, ep.p, u, jp);
  {
    const int org = o0 - G::HX;
    if (threadIdx.x == 0) fo[0] = org;
    if (jp.coord[0])
      for (int e = threadIdx.x; e < G::EX; e += G::NTHR) {
        int p = org + e;
        if (G::PER0) p = mol_wrap(mol_wrap(p, G::N0), G::N0);
        fc[e] = jp.coord[0][min(max(p, 1), (int)G::N0) - 1];
      }
  }
  __syncthreads();
  constexpr int NJ = MolFusedItems<G>::NJ;
  T Ur[NJ][G::NF], Ar[NJ][G::NF];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    int lx, ly;
    MolFusedItems<G>::pos(j, lx, ly);
    const bool ok = MolFusedItems<G>::valid(j);
    const int b = ok ? (lx + G::HX) + G::EX * (ly + G::HY) : 0;
#pragma unroll
    for (int f = 0; f < G::NF; ++f) {
      Ur[j][f] = ok ? U[f * G::VOL + b] : T(0);
      Ar[j][f] = T(0);
    }
  }
  MolFusedInv<G> Iv;
  Iv.init(o0, o1, o2, ep.p, jp);
  double t = t0;
  for (int k = 0; k < nsub; ++k) {
    const bool last = k == nsub - 1;
    mol_rk4_stage<G, 1>(U, Ya, U, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt);
    __syncthreads();
    if constexpr (G::FIX) {
      if (mol_rk4_fix_ghosts<G, 3>(Ya, o0, o1, o2, t + dt / 2, ep.p, jp)) __syncthreads();
    }
    mol_rk4_stage<G, 2>(Ya, Yb, U, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt);
    __syncthreads();
    if constexpr (G::FIX) {
      if (mol_rk4_fix_ghosts<G, 2>(Yb, o0, o1, o2, t + dt / 2, ep.p, jp)) __syncthreads();
    }
    mol_rk4_stage<G, 3>(Yb, Ya, U, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt);
    __syncthreads();
    if constexpr (G::FIX) {
      if (mol_rk4_fix_ghosts<G, 1>(Ya, o0, o1, o2, t + dt, ep.p, jp)) __syncthreads();
    }
    // the stage-4 frame is empty: U is only read at tile positions by stage 4 (through Ur)
    mol_rk4_stage<G, 4>(Ya, nullptr, U, Ur, Ar, Iv, o0, o1, o2, u, unew, ep, jp, t, dt, U, last);
    t = t + dt;
    if (last) break;
    __syncthreads();
    // refresh the halo of U for the next step
    if constexpr (G::PER0) {
      constexpr int P = G::TX;   // the tile is the whole ring of unknowns
      for (int q = threadIdx.x; q < 2 * G::HX; q += G::NTHR) {
        const int lx = q < G::HX ? q - G::HX : G::TX + (q - G::HX);
        const int src = lx < 0 ? lx + P : lx - P;
#pragma unroll
        for (int f = 0; f < G::NF; ++f) U[f * G::VOL + lx + G::HX] = U[f * G::VOL + src + G::HX];
      }
    } else {
      mol_rk4_fix_ghosts<G, 4>(U, o0, o1, o2, t, ep.p, jp);
    }
    __syncthreads();
  }
  if (ep.tnext && threadIdx.x == 0) ep.tnext[0] = t;
}

// Arguments of the generated vector-Jacobian product (layout identical to mol::VArgs<T>).
// The seed is lam = c1*s1 + c2*s2 (s2 may be null); out receives (df/du)^T lam and g_aux the
// FP64 auxiliary-array gradients (may be null); both are accumulated, not overwritten.  bf and
// ba hold each tile's tile + halo adjoints of the staged variables and grid fields between the
// tile kernel and the gather kernel.
struct VArgs {
  const T* u;
  const T* s1;
  const T* s2;
  double c1, c2;
  T* out;
  double* g_aux;
  const T* p;
  const double* tnow;
  double tofs;
  T* bf;
  double* ba;
  int mode;   // 1: gather the variables, 2: gather the grid fields, 4: add the tile kernel's
              // grid-field adjoints to the buffers instead of storing them, 8: the gather ends
              // an RK4 step of the adjoint sweep: lam = lam + y0 + y1 + y2 + (this gather)
  T* lam;
  const T* y0;
  const T* y1;
  const T* y2;
};

// Floor of a / b for b > 0.
__device__ __forceinline__ int mol_floordiv(const int a, const int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Tiles along one axis whose tile + halo (halo H, the larger of the two variants) contains
// position p or, on a periodic axis with period P, one of its images p + kP: tile indices and
// the position relative to each tile's origin.  At most 10 entries.
template <int TL, int H, int NT, int ULO, bool PER, int P>
__device__ __forceinline__ int mol_cover(const int p, int* ts, int* ls) {
  int n = 0;
  if (!PER && H <= TL) {
    // own tile and at most one neighbour on each side
    const int r = p - ULO, t = mol_floordiv(r, TL), l = r - t * TL;
    if (l < H && t - 1 >= 0 && t - 1 < NT) { ts[n] = t - 1; ls[n] = l + TL; ++n; }
    if (t >= 0 && t < NT) { ts[n] = t; ls[n] = l; ++n; }
    if (l >= TL - H && t + 1 >= 0 && t + 1 < NT) { ts[n] = t + 1; ls[n] = l - TL; ++n; }
    return n;
  }
#pragma unroll
  for (int k = PER ? -2 : 0; k <= (PER ? 2 : 0); ++k) {
    const int pp = p + k * P;
    const int tlo = max(0, -mol_floordiv(-(pp - ULO - TL + 1 - H), TL));
    const int thi = min(NT - 1, mol_floordiv(pp - ULO + H, TL));
    for (int t = tlo; t <= thi && n < 10; ++t) {
      ts[n] = t;
      ls[n] = pp - (ULO + t * TL);
      ++n;
    }
  }
  return n;
}

// The (at most) three tiles along one axis whose tile + halo H contains position p: the left
// neighbour, the own tile and the right neighbour (slots 0, 1, 2), with p relative to each
// tile's origin; on a periodic axis of period P the neighbours across the ends hold the images
// p + P and p - P.  Exact when H <= TL and, on a periodic axis, the last tile is at least H wide.
struct MolCov {
  int t[3], l[3];
  bool v[3];
};
template <int TL, int H, int NT, int ULO, bool PER, int P>
__device__ __forceinline__ MolCov mol_cover3(const int p) {
  MolCov c;
  const int r = p - ULO, t = mol_floordiv(r, TL), l = r - t * TL;
  c.t[1] = t;
  c.l[1] = l;
  c.v[1] = t >= 0 && t < NT;
  if (t - 1 >= 0 || !PER) {
    c.t[0] = t - 1;
    c.l[0] = l + TL;
    c.v[0] = l < H && t - 1 >= 0 && t - 1 < NT;
  } else {
    c.t[0] = NT - 1;
    c.l[0] = r + P - (NT - 1) * TL;
    c.v[0] = c.l[0] < TL + H;
  }
  if (t + 1 < NT || !PER) {
    c.t[2] = t + 1;
    c.l[2] = l - TL;
    c.v[2] = l >= TL - H && t + 1 >= 0 && t + 1 < NT;
  } else {
    c.t[2] = 0;
    c.l[2] = r - P;
    c.v[2] = c.l[2] >= -H;
  }
  return c;
}

template <class G, int A>
struct MolAxis {
  static constexpr int TL = A == 0 ? G::TX : (A == 1 ? G::TY : G::TZ);
  static constexpr int H = A == 0 ? G::HX : (A == 1 ? G::HY : G::HZ);
  static constexpr int NT = A == 0 ? G::NT0 : (A == 1 ? G::NT1 : G::NT2);
  static constexpr int ULO = A == 0 ? G::ULO0 : (A == 1 ? G::ULO1 : G::ULO2);
  static constexpr int UHI = A == 0 ? G::UHI0 : (A == 1 ? G::UHI1 : G::UHI2);
  static constexpr bool PER = (A == 0 ? G::PER0 : (A == 1 ? G::PER1 : G::PER2)) != 0;
  static constexpr int P = (A == 0 ? G::N0 : (A == 1 ? G::N1 : G::N2)) - 1;
  static constexpr bool FAST = H <= TL && (!PER || (UHI - ULO + 1) - (NT - 1) * TL >= H);
};

// Adjoint-buffer value of slot s at tile-relative position (l0, l1, l2) of tile (t0, t1, t2), in
// the tile's own variant layout, or 0 outside that variant's tile + halo.
template <class GC, class GB, bool HASC, int VOLX, int NS, typename B>
__device__ __forceinline__ double mol_tile_buf(const B* __restrict__ base, const int s, const int t0, const int t1,
                                               const int t2, const int l0, const int l1, const int l2) {
  // cross-shaped staging: positions outside the tile along two or more axes hold no adjoint
  // (no diagonal taps) and the tile kernel leaves their buffer entries unwritten
  if (GB::CROSS && (int)(l0 < 0 || l0 >= GB::TX) + (int)(l1 < 0 || l1 >= GB::TY) + (int)(l2 < 0 || l2 >= GB::TZ) >= 2)
    return 0.0;
  const int tl = t0 + GB::NT0 * (t1 + GB::NT1 * t2);
  int q;
  if (HASC && GB::is_core(t0, t1, t2)) {
    if (l0 < -GC::HX || l0 >= GC::TX + GC::HX || l1 < -GC::HY || l1 >= GC::TY + GC::HY || l2 < -GC::HZ ||
        l2 >= GC::TZ + GC::HZ)
      return 0.0;
    q = (l0 + GC::HX) + GC::EX * ((l1 + GC::HY) + GC::EY * (l2 + GC::HZ));
  } else {
    if (l0 < -GB::HX || l0 >= GB::TX + GB::HX || l1 < -GB::HY || l1 >= GB::TY + GB::HY || l2 < -GB::HZ ||
        l2 >= GB::TZ + GB::HZ)
      return 0.0;
    q = (l0 + GB::HX) + GB::EX * ((l1 + GB::HY) + GB::EY * (l2 + GB::HZ));
  }
  return (double)base[(i64)tl * (NS * VOLX) + (i64)s * VOLX + q];
}

// Sum over the tiles covering position (p0, p1, p2) of slot s of their adjoint buffers (base:
// buffer of tile 0, stride NS * VOLX per tile), in a fixed order: each tile in its own
// variant's layout.
template <class GC, class GB, bool HASC, int VOLX, int NS, typename B>
__device__ __forceinline__ double mol_gather_pos(const B* __restrict__ base, const int s, const int p0, const int p1,
                                                 const int p2) {
  using A0 = MolAxis<GB, 0>;
  using A1 = MolAxis<GB, 1>;
  using A2 = MolAxis<GB, 2>;
  double sum = 0.0;
  if (A0::FAST && A1::FAST && A2::FAST) {
    const MolCov c0 = mol_cover3<A0::TL, A0::H, A0::NT, A0::ULO, A0::PER, A0::P>(p0);
    const MolCov c1 = mol_cover3<A1::TL, A1::H, A1::NT, A1::ULO, A1::PER, A1::P>(p1);
    const MolCov c2 = mol_cover3<A2::TL, A2::H, A2::NT, A2::ULO, A2::PER, A2::P>(p2);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          // slot 1 is the tile owning the position along that axis; with cross-shaped staging a
          // neighbour along two or more axes holds the position outside its tile along those
          // axes, where its buffer is never written: those combinations drop out at compile time
          if (GB::CROSS && (int)(a != 1) + (int)(b != 1) + (int)(c != 1) >= 2) continue;
          if (c0.v[a] && c1.v[b] && c2.v[c])
            sum += mol_tile_buf<GC, GB, HASC, VOLX, NS>(base, s, c0.t[a], c1.t[b], c2.t[c], c0.l[a], c1.l[b], c2.l[c]);
        }
    return sum;
  }
  int t0[10], l0[10], t1[10], l1[10], t2[10], l2[10];
  const int n0 = mol_cover<GB::TX, GB::HX, GB::NT0, GB::ULO0, GB::PER0 != 0, GB::N0 - 1>(p0, t0, l0);
  const int n1 = mol_cover<GB::TY, GB::HY, GB::NT1, GB::ULO1, GB::PER1 != 0, GB::N1 - 1>(p1, t1, l1);
  const int n2 = mol_cover<GB::TZ, GB::HZ, GB::NT2, GB::ULO2, GB::PER2 != 0, GB::N2 - 1>(p2, t2, l2);
  for (int c = 0; c < n2; ++c)
    for (int b = 0; b < n1; ++b)
      for (int a = 0; a < n0; ++a)
        sum += mol_tile_buf<GC, GB, HASC, VOLX, NS>(base, s, t0[a], t1[b], t2[c], l0[a], l1[b], l2[c]);
  return sum;
}

// One tile of the reverse-mode sweep in tile variant G: stage as the forward kernel, run each
// node's adjoint (tap adjoints summed in FP64 in LDS with LDS atomics, whatever the state's
// type; arrays read only at the node go
// straight to global memory), then add the LDS adjoints of tile + halo to the unknowns (through
// the transposed ghost closures) and to the grid-field gradients.
template <class G, int VOLX>
__device__ __forceinline__ void mol_vjp_tile(T* L, double* Ab, double* Aa, const int tx, const int ty, const int tz,
                                             const VArgs& va, const JPtrs& jp, const double t) {
  const int o0 = G::ULO0 + tx * G::TX, o1 = G::ULO1 + ty * G::TY, o2 = G::ULO2 + tz * G::TZ;
  constexpr int EXY = G::EX * G::EY;
  {
    const int mode = mol_stage_mode<G>(tx, ty, tz);
    MolWalk<G::EX, G::EY, G::NTHR> w((int)threadIdx.x);
    for (int q = threadIdx.x; q < G::VOL; q += G::NTHR, w.step()) {
      if (!mol_unread<G>(w.x, w.y, w.z)) {
        G::stage_pos(L, q, o0 + w.x - G::HX, o1 + w.y - G::HY, o2 + w.z - G::HZ, mode, t, va.p, va.u, jp);
        G::zero_pos(Ab, Aa, q);
      }
    }
  }
  __syncthreads();
  constexpr int TXY = G::TX * G::TY, TVOL = TXY * G::TZ;
  for (int q = threadIdx.x; q < TVOL; q += G::NTHR) {
    const int lz = q / TXY;
    const int rr = q - lz * TXY;
    const int ly = rr / G::TX;
    const int lx = rr - ly * G::TX;
    const int I0 = o0 + lx, I1 = o1 + ly, I2 = o2 + lz;
    if (I0 > G::UHI0 || I1 > G::UHI1 || I2 > G::UHI2) continue;
    T lam[G::NF];
    bool any = false;
#pragma unroll
    for (int f = 0; f < G::NF; ++f) {
      T s = T(0);
      if (G::inbox(f, I0, I1, I2)) {
        const auto k = G::uidx(f, I0, I1, I2);
        s = (T)va.c1 * va.s1[k];
        if (va.s2) s += (T)va.c2 * va.s2[k];
      }
      lam[f] = s;
      any = any || s != T(0);
    }
    if (!any) continue;
    const int base = (lx + G::HX) + G::EX * (ly + G::HY) + EXY * (lz + G::HZ);
    G::vjp(L, Ab, Aa, base, I0, I1, I2, t, va.p, va.u, jp, lam, va.out, va.g_aux);
  }
  __syncthreads();
  // tile + halo adjoints to this tile's buffers (plain stores; the gather kernel sums the
  // buffers of the tiles covering each position), ghost positions through the closures
  const i64 tl = tx + (i64)G::NT0 * (ty + (i64)G::NT1 * tz);
  T* bft = va.bf + tl * (G::NSTF * VOLX);
  double* bat = va.g_aux ? va.ba + tl * (G::NSTA * VOLX) : nullptr;
  const bool acc = (va.mode & 4) != 0;
  MolWalk<G::EX, G::EY, G::NTHR> w((int)threadIdx.x);
  for (int q = threadIdx.x; q < G::VOL; q += G::NTHR, w.step())
    if (!mol_unread<G>(w.x, w.y, w.z))
      G::flush_pos(Ab, Aa, q, o0 + w.x - G::HX, o1 + w.y - G::HY, o2 + w.z - G::HZ, va.out, bft, bat, VOLX, acc);
}

// ---------------------------------------------------------------------------------------
// Batched instances.  Each instance of a batch has a header with its grid sizes, unknown
// boxes (offsets into the batch state vector), tile counts and tile classes, and pointers to
// its weight pool W (the descriptor's pool), constants K (the constant pool, then eps and
// 1/dx of every slot) and coordinates.  Instances of one structure share a class body
// (struct with the node function); a persistent kernel walks the work list of (instance, tile)
// items and dispatches on the class.  Layout identical to mol::BInst.
struct BInst {
  int n[3];
  int lo[8][3];
  int hi[8][3];
  i64 fofs[8];
  int ulo[3], uhi[3], nt[3];
  int slo[3], shi[3], clo[3], chi[3];
  int ilo[3], ihi[3];   // nodes where the interior-rows code applies
  int cls, first;
  const double* coord[3];
  const double* W;
  const double* K;
};

// Index of node (p0, p1, p2) of variable f in the batch state: the instance's offset plus a
// 32-bit offset inside the variable's box (each box holds fewer than 2^31 unknowns).
__device__ __forceinline__ i64 mol_b_uidx(const BInst& hd, const int f, const int p0, const int p1, const int p2) {
  return hd.fofs[f] + (i64)((p0 - hd.lo[f][0]) +
                            (hd.hi[f][0] - hd.lo[f][0] + 1) *
                                ((p1 - hd.lo[f][1]) + (hd.hi[f][1] - hd.lo[f][1] + 1) * (p2 - hd.lo[f][2])));
}

__device__ __forceinline__ bool mol_b_inbox(const BInst& hd, const int f, const int I0, const int I1, const int I2) {
  return I0 >= hd.lo[f][0] && I0 <= hd.hi[f][0] && I1 >= hd.lo[f][1] && I1 <= hd.hi[f][1] && I2 >= hd.lo[f][2] &&
         I2 <= hd.hi[f][2];
}

// Position walker over a box of run-time extents (x fastest), advanced by the work-group's
// thread count: one division per axis when a tile starts, none per position.
// Positions of a box of run-time extents (x fastest), visited by the threads of a work-group
// with stride NTHR: one division per axis when a walk starts, none per position.  Z: the box
// has more than one plane.
template <bool Z>
struct MolRtWalk {
  int x, y, z, dx, dy, dz, ex, ey;
  __device__ __forceinline__ MolRtWalk(const int q, const int ex_, const int ey_, const int stride) : ex(ex_), ey(ey_) {
    if constexpr (Z) {
      const int exy = ex * ey;
      z = q / exy;
      const int r = q - z * exy;
      y = r / ex;
      x = r - y * ex;
      dz = stride / exy;
      const int rs = stride - dz * exy;
      dy = rs / ex;
      dx = rs - dy * ex;
    } else {
      z = dz = 0;
      y = q / ex;
      x = q - y * ex;
      dy = stride / ex;
      dx = stride - dy * ex;
    }
  }
  __device__ __forceinline__ void step() {
    x += dx;
    y += dy;
    if (x >= ex) { x -= ex; ++y; }
    if constexpr (Z) {
      z += dz;
      if (y >= ey) { y -= ey; ++z; }
    }
  }
};

// Inclusive integer box.
struct MolBox {
  int lo[3], hi[3];
  __device__ __forceinline__ bool empty() const { return hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]; }
  __device__ __forceinline__ int ext(const int a) const { return hi[a] - lo[a] + 1; }
  __device__ __forceinline__ int vol() const { return ext(0) * ext(1) * ext(2); }
  __device__ __forceinline__ MolBox meet(const MolBox& o) const {
    MolBox m;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      m.lo[a] = max(lo[a], o.lo[a]);
      m.hi[a] = min(hi[a], o.hi[a]);
    }
    return m;
  }
};

// The positions of box B outside its sub-box A (A non-empty), as up to six slabs: below and
// above A along z (whole planes of B), then along y (within A's planes), then along x (within
// A's rows).  count(): their number; at(k): the k-th position.
struct MolFrame {
  MolBox B, A;
  int n[6];
  __device__ __forceinline__ MolFrame(const MolBox& b, const MolBox& a) : B(b), A(a) {
    const int ex = B.ext(0), ey = B.ext(1), ay = A.ext(1), az = A.ext(2);
    n[0] = (A.lo[2] - B.lo[2]) * ex * ey;
    n[1] = (B.hi[2] - A.hi[2]) * ex * ey;
    n[2] = (A.lo[1] - B.lo[1]) * ex * az;
    n[3] = (B.hi[1] - A.hi[1]) * ex * az;
    n[4] = (A.lo[0] - B.lo[0]) * ay * az;
    n[5] = (B.hi[0] - A.hi[0]) * ay * az;
  }
  __device__ __forceinline__ int count() const { return n[0] + n[1] + n[2] + n[3] + n[4] + n[5]; }
  __device__ __forceinline__ void at(int k, int& p0, int& p1, int& p2) const {
    const int ex = B.ext(0), ey = B.ext(1);
    int s = 0;
    while (s < 5 && k >= n[s]) k -= n[s++];
    if (s < 2) {   // planes of B
      const int exy = ex * ey, pz = k / exy, r = k - pz * exy, py = r / ex;
      p2 = (s == 0 ? B.lo[2] : A.hi[2] + 1) + pz;
      p1 = B.lo[1] + py;
      p0 = B.lo[0] + (r - py * ex);
    } else if (s < 4) {   // rows of B within A's planes
      const int w = (s == 2 ? A.lo[1] - B.lo[1] : B.hi[1] - A.hi[1]);
      const int pz = k / (ex * w), r = k - pz * ex * w, py = r / ex;
      p2 = A.lo[2] + pz;
      p1 = (s == 2 ? B.lo[1] : A.hi[1] + 1) + py;
      p0 = B.lo[0] + (r - py * ex);
    } else {   // columns within A's rows
      const int w = (s == 4 ? A.lo[0] - B.lo[0] : B.hi[0] - A.hi[0]), ay = A.ext(1);
      const int pz = k / (w * ay), r = k - pz * w * ay, py = r / w;
      p2 = A.lo[2] + pz;
      p1 = A.lo[1] + py;
      p0 = (s == 4 ? B.lo[0] : A.hi[0] + 1) + (r - py * w);
    }
  }
};

// Position inside every variable's unknown box.
template <int NF>
__device__ __forceinline__ bool mol_b_inall(const BInst& hd, const int p0, const int p1, const int p2) {
  bool in = true;
#pragma unroll
  for (int f = 0; f < NF; ++f) in = in && mol_b_inbox(hd, f, p0, p1, p2);
  return in;
}

// One node of a batched tile whose LDS box starts at (s0, s1, s2), and its stage epilogue.
// INTERIOR: the interior-rows code, at a node inside every variable's box.
template <class G, bool INTERIOR>
__device__ __forceinline__ void mol_batch_node(const T* L, const BInst& hd, const int I0, const int I1, const int I2,
                                               const int s0, const int s1, const int s2, const T* __restrict__ u,
                                               const Epi& ep, const JPtrs& jp, const double t) {
  const int base = (I0 - s0) + G::EX * ((I1 - s1) + G::EY * (I2 - s2));
  T out[G::NF];
  G::template node<INTERIOR>(hd, L, base, I0, I1, I2, t, nullptr, u, jp, out);
#pragma unroll
  for (int f = 0; f < G::NF; ++f)
    if (INTERIOR || mol_b_inbox(hd, f, I0, I1, I2)) mol_epilogue(ep, mol_b_uidx(hd, f, I0, I1, I2), out[f]);
}

// One tile of one instance (L: the kernel's LDS tile, sized for the largest class).  Only the
// part of the tile inside the instance's union box is staged and evaluated, so tiles at least
// as wide as an instance cover whole rows of it (contiguous in memory: staging loads and the
// epilogue touch no cache line of another row) and narrow instances leave no idle lanes in a
// wide tile.  Staging: plain loads where the tile + halo lies inside every variable's box,
// the ghost closures (and periodic images) on the frame around it.  Nodes: the interior-rows
// code inside the instance's interior node range, the boundary-rows code on the frame around it.
template <class G>
__device__ __forceinline__ void mol_batch_tile(T* L, const BInst& hd, const int tile, const T* __restrict__ u,
                                               const Epi& ep, const double t) {
  constexpr bool Z = G::TZ > 1 || G::HZ > 0;
  const int tx = tile % hd.nt[0];
  const int r1 = tile / hd.nt[0];
  const int ty = r1 % hd.nt[1];
  const int tz = r1 / hd.nt[1];
  MolBox N;   // the tile's nodes inside the union box
  N.lo[0] = hd.ulo[0] + tx * G::TX;
  N.lo[1] = hd.ulo[1] + ty * G::TY;
  N.lo[2] = hd.ulo[2] + tz * G::TZ;
  N.hi[0] = min(N.lo[0] + G::TX - 1, hd.uhi[0]);
  N.hi[1] = min(N.lo[1] + G::TY - 1, hd.uhi[1]);
  N.hi[2] = min(N.lo[2] + G::TZ - 1, hd.uhi[2]);
  MolBox S = N;   // with the halo: the LDS box
  S.lo[0] -= G::HX; S.hi[0] += G::HX;
  S.lo[1] -= G::HY; S.hi[1] += G::HY;
  S.lo[2] -= G::HZ; S.hi[2] += G::HZ;
  JPtrs jp;
  jp.coord[0] = hd.coord[0];
  jp.coord[1] = hd.coord[1];
  jp.coord[2] = hd.coord[2];
  if constexpr (G::NST > 0) {
    MolBox V = S;   // inside every variable's box: plain loads
#pragma unroll
    for (int f = 0; f < G::NF; ++f) {
      MolBox bf;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        bf.lo[a] = hd.lo[f][a];
        bf.hi[a] = hd.hi[f][a];
      }
      V = V.meet(bf);
    }
    constexpr int IT = (G::VOL + G::NTHR - 1) / G::NTHR;
    constexpr int NS = G::NST > 0 ? G::NST : 1;
    if (!V.empty()) {
      // variable boxes and the first plain position, per staged slot
      i64 vb[NS];
      int sx[NS], sxy[NS];
#pragma unroll
      for (int s = 0; s < G::NST; ++s) {
        const int f = G::SF[s];
        sx[s] = hd.hi[f][0] - hd.lo[f][0] + 1;
        sxy[s] = sx[s] * (hd.hi[f][1] - hd.lo[f][1] + 1);
        vb[s] = mol_b_uidx(hd, f, V.lo[0], V.lo[1], V.lo[2]);
      }
      const int nv = V.vol();
      const int lofs = (V.lo[0] - S.lo[0]) + G::EX * ((V.lo[1] - S.lo[1]) + G::EY * (V.lo[2] - S.lo[2]));
      MolRtWalk<Z> w((int)threadIdx.x, V.ext(0), V.ext(1), G::NTHR);
      // chunks of CH positions per thread: all loads of a chunk are issued before its LDS
      // stores (one exposed memory latency per chunk, not per position)
      constexpr int CH = IT * G::NST <= 16 ? IT : (16 / G::NST > 0 ? 16 / G::NST : 1);
      for (int q0 = (int)threadIdx.x; q0 < nv; q0 += CH * G::NTHR) {
        T v[CH][NS];
        int dst[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k, w.step()) {
          dst[k] = q0 + k * G::NTHR < nv ? lofs + w.x + G::EX * (w.y + G::EY * w.z) : -1;
          if (dst[k] >= 0)
#pragma unroll
            for (int s = 0; s < G::NST; ++s) v[k][s] = u[vb[s] + (w.x + sx[s] * w.y + sxy[s] * w.z)];
        }
#pragma unroll
        for (int k = 0; k < CH; ++k)
          if (dst[k] >= 0)
#pragma unroll
            for (int s = 0; s < G::NST; ++s) L[s * G::VOL + dst[k]] = v[k][s];
      }
    }
    // the frame of ghost positions and periodic images (everything when nothing is plain)
    const MolFrame F(S, V.empty() ? S : V);
    const int nf = V.empty() ? S.vol() : F.count();
    for (int q = threadIdx.x; q < nf; q += G::NTHR) {
      int p0, p1, p2;
      if (V.empty()) {
        const int e0 = S.ext(0), e1 = S.ext(1), pz = q / (e0 * e1), r = q - pz * e0 * e1, py = r / e0;
        p0 = S.lo[0] + (r - py * e0);
        p1 = S.lo[1] + py;
        p2 = S.lo[2] + pz;
      } else {
        F.at(q, p0, p1, p2);
      }
      T v[NS];
      G::stage_vals(hd, v, p0, p1, p2, false, t, nullptr, u, jp);
      const int d = (p0 - S.lo[0]) + G::EX * ((p1 - S.lo[1]) + G::EY * (p2 - S.lo[2]));
#pragma unroll
      for (int s = 0; s < G::NST; ++s) L[s * G::VOL + d] = v[s];
    }
    __syncthreads();
  }
  MolBox I;   // nodes where the interior-rows code applies
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    I.lo[a] = hd.ilo[a];
    I.hi[a] = hd.ihi[a];
  }
  I = I.meet(N);
  if (!I.empty()) {
    // interior nodes lie in every variable's box (the interior range is empty otherwise), so
    // the variables' indices differ by their offsets only; chunks of CN nodes per thread share
    // one exposed latency of the epilogue's loads
    constexpr int CN = 4;
    i64 dv[G::NF];
#pragma unroll
    for (int f = 0; f < G::NF; ++f) dv[f] = hd.fofs[f] - hd.fofs[0];
    const int ni = I.vol();
    MolRtWalk<Z> w((int)threadIdx.x, I.ext(0), I.ext(1), G::NTHR);
    for (int q0 = (int)threadIdx.x; q0 < ni; q0 += CN * G::NTHR) {
      T out[CN][G::NF];
      i64 idx[CN];
      bool ok[CN];
#pragma unroll
      for (int k = 0; k < CN; ++k, w.step()) {
        ok[k] = q0 + k * G::NTHR < ni;
        idx[k] = 0;
        if (!ok[k]) continue;
        const int I0 = I.lo[0] + w.x, I1 = I.lo[1] + w.y, I2 = I.lo[2] + w.z;
        const int base = (I0 - S.lo[0]) + G::EX * ((I1 - S.lo[1]) + G::EY * (I2 - S.lo[2]));
        G::template node<true>(hd, L, base, I0, I1, I2, t, nullptr, u, jp, out[k]);
        idx[k] = mol_b_uidx(hd, 0, I0, I1, I2);
      }
      mol_epilogue_chunk<CN, G::NF>(ep, idx, ok, dv, out);
    }
    const MolFrame F(N, I);
    const int nf = F.count();
    for (int q = threadIdx.x; q < nf; q += G::NTHR) {
      int p0, p1, p2;
      F.at(q, p0, p1, p2);
      mol_batch_node<G, false>(L, hd, p0, p1, p2, S.lo[0], S.lo[1], S.lo[2], u, ep, jp, t);
    }
  } else {
    const int nn = N.vol();
    MolRtWalk<Z> w((int)threadIdx.x, N.ext(0), N.ext(1), G::NTHR);
    for (int q = threadIdx.x; q < nn; q += G::NTHR, w.step())
      mol_batch_node<G, false>(L, hd, N.lo[0] + w.x, N.lo[1] + w.y, N.lo[2] + w.z, S.lo[0], S.lo[1], S.lo[2], u,
                               ep, jp, t);
  }
  __syncthreads();
}
