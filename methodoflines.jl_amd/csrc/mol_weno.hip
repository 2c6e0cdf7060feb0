// WENO5 fast path for 2D systems whose operator slots are all Jiang-Shu WENO5 derivatives
// (e.g. inviscid Burgers u_t = -u u_x - u u_y on a periodic grid).
//
// The derivative at node i along an axis restates weno_f_uniform
// (src/discretization/schemes/WENO/WENO.jl:6-57) in an equivalent form with fewer operations:
//   * first differences  d(i+1/2) = u(i+1) - u(i)  and second differences  D2(i) = d(i+1/2) - d(i-1/2)
//     are computed once per node and shared by the neighbours of a register strip;
//   * the smoothness indicators become  beta1 = 13/12 D2(i+1)^2 + 1/4 (D2(i+1) - 2 d(i+1/2))^2,
//     beta2 = 13/12 D2(i)^2 + 1/4 (d(i-1/2) + d(i+1/2))^2,  beta3 = 13/12 D2(i-1)^2 + 1/4 (D2(i-1) + 2 d(i-1/2))^2
//     (identical polynomials to the reference's);
//   * every candidate reconstruction is written as u(i) + (a difference)/6, so u(i) cancels
//     exactly in  û-(i+1/2) - û+(i-1/2);
//   * the normalised weights gamma_k/q_k / sum_j gamma_j/q_j  (q = (eps + beta)^2) are cleared to
//     gamma_k P_k / sum_j gamma_j P_j with P_k = prod_{j != k} q_j, leaving one reciprocal per node
//     and axis.
// A work-group owns a 64 x 32 tile.  Pass 1: lanes run along rows, each thread one 8-node strip
// in x; the x-derivatives go to LDS.  Pass 2: lanes run along columns, 8-node strips in y; the
// y-derivatives stay in registers, the right-hand-side program combines both and the RK stage
// epilogue writes the result.
#include "mol_core.h"
#include "mol_device.h"
#include "mol_internal.h"

namespace mol {

constexpr int WT_X = WENO_TILE_X, WT_Y = WENO_TILE_Y, WSTRIP = 8, WHALO = 3;
constexpr int WEXT_X = WT_X + 2 * WHALO, WEXT_Y = WT_Y + 2 * WHALO;
constexpr int WRS = WEXT_X + 1;        // odd row stride (in elements): conflict-free row-parallel reads
constexpr int WRS_OUT = WT_X + 1;      // row stride of the x-derivative buffer

template <typename T>
__device__ __forceinline__ T fast_rcp(T x) {
  return T(1) / x;
}

// Numerator/denominator of the WENO5 node derivative for B consecutive nodes, from the B+4
// values v[0..B+3] (nodes i0-2 .. i0+B+1).  W(i) = num[i] / den[i] * (inv_dx / 6).
template <typename T, int B>
__device__ __forceinline__ void weno_strip(const T* v, T eps, T* num, T* den) {
  T d[B + 3];   // d[k] = v[k+1] - v[k]   (difference at half point i0-2+k+1/2)
#pragma unroll
  for (int k = 0; k < B + 3; ++k) d[k] = v[k + 1] - v[k];
  T D2[B + 2];  // D2 at node i0-1+k
  T S[B + 2];   // 13/12 D2^2 + eps
#pragma unroll
  for (int k = 0; k < B + 2; ++k) {
    D2[k] = d[k + 1] - d[k];
    S[k] = fma(T(13.0 / 12.0) * D2[k], D2[k], eps);
  }
#pragma unroll
  for (int j = 0; j < B; ++j) {
    // node i = i0 + j: d(i-3/2) = d[j], d(i-1/2) = d[j+1], d(i+1/2) = d[j+2], d(i+3/2) = d[j+3]
    // D2(i-1) = D2[j], D2(i) = D2[j+1], D2(i+1) = D2[j+2]
    const T dmm = d[j], dm = d[j + 1], dp = d[j + 2], dpp = d[j + 3];
    const T a1 = fma(T(-2), dp, D2[j + 2]);
    const T a2 = dm + dp;
    const T a3 = fma(T(2), dm, D2[j]);
    const T b1 = fma(T(0.25) * a1, a1, S[j + 2]);
    const T b2 = fma(T(0.25) * a2, a2, S[j + 1]);
    const T b3 = fma(T(0.25) * a3, a3, S[j]);
    const T q1 = b1 * b1, q2 = b2 * b2, q3 = b3 * b3;
    const T P1 = q2 * q3, P2 = q1 * q3, P3 = q1 * q2;
    const T A1 = T(0.3) * P1, A2 = T(0.6) * P2, A3 = T(0.1) * P3;
    const T B1 = T(0.1) * P1, B3 = T(0.3) * P3;
    const T Dp = A1 + A2 + A3;
    const T Dm = B1 + A2 + B3;
    // 6 (hp_k - u_i) and 6 (hm_k - u_i)
    const T ep1 = fma(T(4), dp, -dpp), ep2 = fma(T(2), dp, dm), ep3 = fma(T(5), dm, T(-2) * dmm);
    const T em1 = fma(T(-5), dp, T(2) * dpp), em2 = fma(T(-2), dm, -dp), em3 = fma(T(-4), dm, dmm);
    const T Np = fma(A1, ep1, fma(A2, ep2, A3 * ep3));
    const T Nm = fma(B1, em1, fma(A2, em2, B3 * em3));
    num[j] = fma(Np, Dm, -(Nm * Dp));
    den[j] = Dp * Dm;
  }
}

// Leaves of the right-hand-side program for the WSTRIP nodes of a column strip.
template <typename T>
struct WenoStripCtx {
  const KCore<T>* kc;
  const T* lds;
  int lin0;             // LDS index of the strip's first node; the next node is +WRS
  int ldsvol;
  const T* sv0;         // slot values by slot id (<= 4 slots), WSTRIP each
  const T* sv1;
  const T* sv2;
  const T* sv3;
  int I0, I1;           // full-grid index of the first node
  double t;
  const T* p;
  __device__ __forceinline__ void leaf(uint32_t op, int imm, T* v) const {
#pragma unroll
    for (int j = 0; j < WSTRIP; ++j) {
      T x;
      switch (op) {
        case MOL_OP_FIELD: x = lds[imm * ldsvol + lin0 + j * WRS]; break;
        case MOL_OP_SLOT: x = imm == 0 ? sv0[j] : (imm == 1 ? sv1[j] : (imm == 2 ? sv2[j] : sv3[j])); break;
        case MOL_OP_COORD: x = (T)kc->coord[imm][(imm == 0 ? I0 : I1 + j) - 1]; break;
        case MOL_OP_TIME: x = (T)t; break;
        case MOL_OP_PARAM: x = p[imm]; break;
        default: x = T(0); break;
      }
      v[j] = x;
    }
  }
};

template <typename T>
__global__ void __launch_bounds__(256) weno2d_kernel(const KCore<T> kc_arg, WenoPlan plan,
                                                     const T* __restrict__ u, Epi<T> ep) {
#if defined(__HIP_DEVICE_COMPILE__)
  const KCore<T>& kc = *(const KCore<T>*)(__builtin_amdgcn_kernarg_segment_ptr());
  (void)kc_arg;
#else
  const KCore<T>& kc = kc_arg;
#endif
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const int ldsvol = WRS * WEXT_Y;
  T* wx = lds + kc.nfields * ldsvol;                 // x-derivative buffers, WT_Y x WRS_OUT each
  const int ntx = kc.ct_n[0];
  const int tx = kc.ct_lo[0] + (int)blockIdx.x % ntx;
  const int ty = kc.ct_lo[1] + (int)blockIdx.x / ntx;
  const int org0 = kc.lo[0] + tx * WT_X;
  const int org1 = kc.lo[1] + ty * WT_Y;
  const int bx = kc.hi[0] - kc.lo[0] + 1;
  const double t = (ep.tnow ? ep.tnow[0] : 0.0) + ep.tofs;
  advance_clock(ep);

  // stage tile + halo (rows of WEXT_X, padded to WRS)
  for (int q = threadIdx.x; q < WEXT_X * WEXT_Y; q += blockDim.x) {
    const int ly = q / WEXT_X;
    const int lx = q - ly * WEXT_X;
    int p0 = org0 + lx - WHALO, p1 = org1 + ly - WHALO;
    // second wrap: a ragged last tile reaches up to tile - 1 + halo nodes past n
    if (kc.periodic[0]) p0 = wrap_periodic(wrap_periodic(p0, kc.n[0]), kc.n[0]);
    if (kc.periodic[1]) p1 = wrap_periodic(wrap_periodic(p1, kc.n[1]), kc.n[1]);
    const long long li = (long long)(p0 - kc.lo[0]) + (long long)bx * (p1 - kc.lo[1]);
    for (int f = 0; f < kc.nfields; ++f) lds[f * ldsvol + ly * WRS + lx] = u[f * kc.box + li];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // pass 1: x-derivatives; lane -> row (32 rows x 2 strips per wave), 4 waves x 2 = 8 strips
  {
    const int row = lane & 31;
    const int strip = wave * 2 + (lane >> 5);
    const int c0 = strip * WSTRIP;
    for (int s = 0; s < plan.nx; ++s) {
      const CoreSlot& cs = kc.slots[plan.sx[s]];
      const T* src = lds + plan.fx[s] * ldsvol + (row + WHALO) * WRS + (c0 + WHALO - 2);
      T v[WSTRIP + 4];
#pragma unroll
      for (int k = 0; k < WSTRIP + 4; ++k) v[k] = src[k];
      T num[WSTRIP], den[WSTRIP];
      weno_strip<T, WSTRIP>(v, (T)cs.eps, num, den);
      const T sc = (T)(cs.inv_dx / 6.0);
      T* dst = wx + s * (WT_Y * WRS_OUT) + row * WRS_OUT + c0;
#pragma unroll
      for (int j = 0; j < WSTRIP; ++j) dst[j] = num[j] * fast_rcp(den[j]) * sc;
    }
  }
  __syncthreads();

  // pass 2: y-derivatives; lane -> column, wave -> 8-row strip; then the program and epilogue
  {
    const int col = lane;
    const int r0 = wave * WSTRIP;
    T wy[2][WSTRIP];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (s >= plan.ny) break;
      const CoreSlot& cs = kc.slots[plan.sy[s]];
      const T* src = lds + plan.fy[s] * ldsvol + (r0 + WHALO - 2) * WRS + (col + WHALO);
      T v[WSTRIP + 4];
#pragma unroll
      for (int k = 0; k < WSTRIP + 4; ++k) v[k] = src[k * WRS];
      T num[WSTRIP], den[WSTRIP];
      weno_strip<T, WSTRIP>(v, (T)cs.eps, num, den);
      const T sc = (T)(cs.inv_dx / 6.0);
#pragma unroll
      for (int j = 0; j < WSTRIP; ++j) wy[s][j] = num[j] * fast_rcp(den[j]) * sc;
    }
    // slot values by slot id, WSTRIP nodes each
    T sv[4][WSTRIP];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < WSTRIP; ++j) sv[k][j] = T(0);
    for (int s = 0; s < plan.nx; ++s) {
      const int id = plan.sx[s];
#pragma unroll
      for (int j = 0; j < WSTRIP; ++j) {
        const T w = wx[s * (WT_Y * WRS_OUT) + (r0 + j) * WRS_OUT + col];
        if (id == 0) sv[0][j] = w; else if (id == 1) sv[1][j] = w; else if (id == 2) sv[2][j] = w; else sv[3][j] = w;
      }
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (s >= plan.ny) break;
      const int id = plan.sy[s];
#pragma unroll
      for (int j = 0; j < WSTRIP; ++j) {
        const T w = wy[s][j];
        if (id == 0) sv[0][j] = w; else if (id == 1) sv[1][j] = w; else if (id == 2) sv[2][j] = w; else sv[3][j] = w;
      }
    }
    WenoStripCtx<T> sc;
    sc.kc = &kc;
    sc.lds = lds;
    sc.lin0 = (r0 + WHALO) * WRS + (col + WHALO);
    sc.ldsvol = ldsvol;
    sc.sv0 = sv[0]; sc.sv1 = sv[1]; sc.sv2 = sv[2]; sc.sv3 = sv[3];
    sc.I0 = org0 + col;
    sc.I1 = org1 + r0;
    sc.t = t;
    sc.p = ep.p;
    const bool colok = sc.I0 <= kc.hi[0];
    for (int f = 0; f < kc.nfields; ++f) {
      T val[WSTRIP];
      stack_prog_vec<T, WSTRIP, 4>(kc, kc.rhs_prog[f], sc, val);
#pragma unroll
      for (int j = 0; j < WSTRIP; ++j) {
        const int I1 = org1 + r0 + j;
        if (!colok || I1 > kc.hi[1]) continue;
        const long long li = (long long)(sc.I0 - kc.lo[0]) + (long long)bx * (I1 - kc.lo[1]);
        epilogue(ep, f * kc.box + li, val[j]);
      }
    }
  }
}

// Plan the fast path: 2D, every slot WENO5, at most two per axis and four in all, core tiles
// covering the box in 64 x 32 tiles.
int program_depth(const mol_desc* d, int prog);

bool weno_plan(const mol_desc* d, WenoPlan& plan) {
  for (int f = 0; f < d->nfields; ++f)
    if (program_depth(d, d->rhs_prog[f]) > 4) return false;
  // FP64 only: the cleared weight denominators (products of squared smoothness indicators,
  // ~eps^8) underflow in FP32
  if (d->dtype != MOL_F64) return false;
  if (d->ndim != 2 || d->naux != 0 || d->nslots < 1 || d->nslots > 4 || d->nfields > 2) return false;
  if (!d->periodic[0] || !d->periodic[1]) return false;
  if (d->hi[0][0] - d->lo[0][0] + 1 < WT_X + WHALO || d->hi[0][1] - d->lo[0][1] + 1 < WT_Y + WHALO) return false;
  plan.nx = plan.ny = 0;
  for (int k = 0; k < d->nslots; ++k) {
    const mol_slot& s = d->slots_h[k];
    if (s.kind != MOL_SLOT_WENO) return false;
    if (s.axis == 0) {
      if (plan.nx == 2) return false;
      plan.sx[plan.nx] = k;
      plan.fx[plan.nx++] = s.field;
    } else {
      if (plan.ny == 2) return false;
      plan.sy[plan.ny] = k;
      plan.fy[plan.ny++] = s.field;
    }
  }
  return true;
}

template <typename T>
size_t weno_lds_bytes(int nfields, int nx) {
  return ((size_t)nfields * WRS * WEXT_Y + (size_t)nx * WT_Y * WRS_OUT) * sizeof(T);
}

template <typename T>
int launch_weno2d(const KCore<T>& kc, const WenoPlan& plan, int nblocks, const T* u, const Epi<T>& ep,
                  hipStream_t stream) {
  const size_t lds = weno_lds_bytes<T>(kc.nfields, plan.nx);
  hipLaunchKernelGGL(weno2d_kernel<T>, dim3(nblocks), dim3(256), lds, stream, kc, plan, u, ep);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}

template int launch_weno2d<double>(const KCore<double>&, const WenoPlan&, int, const double*,
                                   const Epi<double>&, hipStream_t);
template int launch_weno2d<float>(const KCore<float>&, const WenoPlan&, int, const float*,
                                  const Epi<float>&, hipStream_t);
template size_t weno_lds_bytes<double>(int, int);
template size_t weno_lds_bytes<float>(int, int);

}  // namespace mol
