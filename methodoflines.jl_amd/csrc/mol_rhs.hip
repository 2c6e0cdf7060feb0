// Generic stencil-IR right-hand-side kernel with fused Runge-Kutta stage epilogues.
#include "mol_core.h"
#include "mol_device.h"
#include "mol_internal.h"

namespace mol {

template <typename T>
__device__ __forceinline__ void tile_origin(const KArgs<T>& ka, const int32_t* tiles, int* org) {
  int b = tiles ? tiles[blockIdx.x] : (int)blockIdx.x;
  const int tx = b % ka.ntile[0];
  b /= ka.ntile[0];
  const int ty = b % ka.ntile[1];
  const int tz = b / ka.ntile[1];
  org[0] = ka.ulo[0] + tx * ka.tile[0];
  org[1] = ka.ulo[1] + ty * ka.tile[1];
  org[2] = ka.ulo[2] + tz * ka.tile[2];
}

// Stage every variable and auxiliary array over tile + halo into LDS.
template <typename T>
__device__ void stage_tile(const KArgs<T>& ka, const T* __restrict__ u, T* lds, const int* org,
                           double t, const T* p) {
  const int nload = ka.nfields + ka.naux;
  const int total = nload * ka.vol;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    const int arr = q / ka.vol;
    int r = q - arr * ka.vol;
    const int lx = r % ka.ext[0];
    r /= ka.ext[0];
    const int ly = r % ka.ext[1];
    const int lz = r / ka.ext[1];
    int pos[3] = {org[0] + lx - ka.halo[0], org[1] + ly - ka.halo[1], org[2] + lz - ka.halo[2]};
    T v;
    if (arr < ka.nfields) v = field_at(ka, u, arr, pos, t, p);
    else v = aux_at(ka, arr - ka.nfields, pos);
    lds[q] = v;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) rhs_generic_kernel(const KArgs<T>* __restrict__ kap,
                                                          const int32_t* __restrict__ tiles,
                                                          const T* __restrict__ u, Epi<T> ep) {
  const KArgs<T>& ka = *kap;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  int org[3];
  tile_origin(ka, tiles, org);
  const double t = (ep.tnow ? ep.tnow[0] : 0.0) + ep.tofs;
  advance_clock(ep);
  stage_tile(ka, u, lds, org, t, ep.p);
  __syncthreads();

  TileView<T> tv;
  tv.ka = &ka;
  tv.lds = lds;
  const int tvol = ka.tile[0] * ka.tile[1] * ka.tile[2];
  for (int q = threadIdx.x; q < tvol; q += blockDim.x) {
    int r = q;
    const int lx = r % ka.tile[0];
    r /= ka.tile[0];
    const int ly = r % ka.tile[1];
    const int lz = r / ka.tile[1];
    NodeCtx<T> nc;
    nc.ka = &ka;
    nc.tv = &tv;
    nc.I[0] = org[0] + lx;
    nc.I[1] = org[1] + ly;
    nc.I[2] = org[2] + lz;
    if (nc.I[0] > ka.uhi[0] || nc.I[1] > ka.uhi[1] || nc.I[2] > ka.uhi[2]) continue;
    nc.L[0] = lx + ka.halo[0];
    nc.L[1] = ly + ka.halo[1];
    nc.L[2] = lz + ka.halo[2];
    nc.t = t;
    nc.p = ep.p;
    for (int f = 0; f < ka.nfields; ++f) {
      bool in = true;
      for (int d = 0; d < ka.ndim; ++d)
        in = in && nc.I[d] >= ka.lo[f][d] && nc.I[d] <= ka.hi[f][d];
      if (!in) continue;
      const T val = run_prog(ka, ka.rhs_prog[f], nc);
      long long li = nc.I[0] - ka.lo[f][0];
      long long stride = ka.hi[f][0] - ka.lo[f][0] + 1;
      if (ka.ndim > 1) {
        li += stride * (nc.I[1] - ka.lo[f][1]);
        stride *= ka.hi[f][1] - ka.lo[f][1] + 1;
      }
      if (ka.ndim > 2) li += stride * (nc.I[2] - ka.lo[f][2]);
      epilogue(ep, ka.fofs[f] + li, val);
    }
  }
}

// Core tiles: the stencil IR in the kernel arguments, interior rows, wrap-only staging.
template <typename T>
__global__ void __launch_bounds__(256) rhs_core_kernel(const KCore<T> kc_arg, const T* __restrict__ u,
                                                       Epi<T> ep) {
  // The argument block is read in place through the kernel-argument segment (constant address
  // space, scalar loads); taking the address of the by-value parameter would copy it to scratch.
#if defined(__HIP_DEVICE_COMPILE__)
  const KCore<T>& kc = *(const KCore<T>*)(__builtin_amdgcn_kernarg_segment_ptr());
  (void)kc_arg;
#else
  const KCore<T>& kc = kc_arg;  // host pass: never executed
#endif
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  int b = blockIdx.x;
  const int tx = kc.ct_lo[0] + b % kc.ct_n[0];
  b /= kc.ct_n[0];
  const int ty = kc.ct_lo[1] + b % kc.ct_n[1];
  const int tz = kc.ct_lo[2] + b / kc.ct_n[1];
  const int org0 = kc.lo[0] + tx * kc.tile[0];
  const int org1 = kc.lo[1] + ty * kc.tile[1];
  const int org2 = kc.lo[2] + tz * kc.tile[2];
  const double t = (ep.tnow ? ep.tnow[0] : 0.0) + ep.tofs;
  advance_clock(ep);
  const int bx = kc.hi[0] - kc.lo[0] + 1;
  const int by = kc.hi[1] - kc.lo[1] + 1;
  const int ext01 = kc.ext[0] * kc.ext[1];
  // stage: every position of tile + halo is an unknown (or its periodic image)
  for (int q = threadIdx.x; q < kc.vol; q += blockDim.x) {
    const int lz = (int)fdiv_magic((uint32_t)q, kc.magic1);
    const int r = q - lz * ext01;
    const int ly = (int)fdiv_magic((uint32_t)r, kc.magic0);
    const int lx = r - ly * kc.ext[0];
    int p0 = org0 + lx - kc.halo[0], p1 = org1 + ly - kc.halo[1], p2 = org2 + lz - kc.halo[2];
    // second wrap: a ragged last tile reaches up to tile - 1 + halo nodes past n
    if (kc.periodic[0]) p0 = wrap_periodic(wrap_periodic(p0, kc.n[0]), kc.n[0]);
    if (kc.periodic[1]) p1 = wrap_periodic(wrap_periodic(p1, kc.n[1]), kc.n[1]);
    if (kc.periodic[2]) p2 = wrap_periodic(wrap_periodic(p2, kc.n[2]), kc.n[2]);
    const long long li = (long long)(p0 - kc.lo[0]) + (long long)bx * ((p1 - kc.lo[1]) + (long long)by * (p2 - kc.lo[2]));
    for (int f = 0; f < kc.nfields; ++f) lds[f * kc.vol + q] = u[f * kc.box + li];
    if (kc.naux) {
      const long long gi = (long long)(p0 - 1) + (long long)kc.n[0] * ((p1 - 1) + (long long)kc.n[1] * (p2 - 1));
      for (int a = 0; a < kc.naux; ++a) lds[(kc.nfields + a) * kc.vol + q] = kc.aux[a][gi];
    }
  }
  __syncthreads();
  CoreView<T> v;
  v.kc = &kc;
  v.lds = lds;
  v.st[0] = 1;
  v.st[1] = kc.ext[0];
  v.st[2] = ext01;
  v.t = t;
  v.p = ep.p;
  const int tvol = kc.tile[0] * kc.tile[1] * kc.tile[2];
  const int t01 = kc.tile[0] * kc.tile[1];
  for (int q = threadIdx.x; q < tvol; q += blockDim.x) {
    const int lz = q / t01;
    const int r = q - lz * t01;
    const int ly = r / kc.tile[0];
    const int lx = r - ly * kc.tile[0];
    v.I[0] = org0 + lx;
    v.I[1] = org1 + ly;
    v.I[2] = org2 + lz;
    if (v.I[0] > kc.hi[0] || v.I[1] > kc.hi[1] || v.I[2] > kc.hi[2]) continue;
    v.base = (lx + kc.halo[0]) + kc.ext[0] * (ly + kc.halo[1]) + ext01 * (lz + kc.halo[2]);
    const long long li = (long long)(v.I[0] - kc.lo[0]) + (long long)bx * ((v.I[1] - kc.lo[1]) + (long long)by * (v.I[2] - kc.lo[2]));
    CoreNodeCtx<T> nc{&v};
    for (int f = 0; f < kc.nfields; ++f) {
      const T val = core_prog(kc, kc.rhs_prog[f], nc);
      epilogue(ep, f * kc.box + li, val);
    }
  }
}

template <typename T>
int launch_rhs(const Launch<T>& L, const T* u, const Epi<T>& ep, hipStream_t stream) {
  const size_t lds = (size_t)(L.ka->nfields + L.ka->naux) * L.ka->vol * sizeof(T);
  if (L.batch) {
    const T* uu = u;
    Epi<T> e = ep;
    const BInst* ins = L.insts;
    const int32_t* wl = L.wl;
    int nw = L.nwork;
    void* args[] = {(void*)&uu, (void*)&e, (void*)&ins, (void*)&wl, (void*)&nw};
    if (hipModuleLaunchKernel(L.batch, L.bgrid, 1, 1, 256, 1, 1, 0, stream, args, nullptr) != hipSuccess)
      return MOL_ERR_HIP;
    return 0;
  }
  if (L.jit) {
    // generated kernel: every tile, boundary rows and ghost closures included
    const T* uu = u;
    Epi<T> e = ep;
    JitPtrs<T> jp = L.jp;
    void* args[] = {(void*)&uu, (void*)&e, (void*)&jp};
    if (hipModuleLaunchKernel(L.jit, L.jgrid, 1, 1, L.jthr, 1, 1, 0, stream, args, nullptr) != hipSuccess)
      return MOL_ERR_HIP;
    return 0;
  }
  // the device clock advances in the first kernel of the launch only
  Epi<T> e2 = ep;
  if (L.ncore > 0) {
    hipLaunchKernelGGL(rhs_core_kernel<T>, dim3(L.ncore), dim3(256), lds, stream, *L.kc, u, ep);
    e2.tnext = nullptr;
  }
  if (L.nbound > 0)
    hipLaunchKernelGGL(rhs_generic_kernel<T>, dim3(L.nbound), dim3(256), lds, stream, L.dka, L.tiles, u, e2);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}

template int launch_rhs<double>(const Launch<double>&, const double*, const Epi<double>&, hipStream_t);
template int launch_rhs<float>(const Launch<float>&, const float*, const Epi<float>&, hipStream_t);

// Full-grid values of every variable (variable-major, first axis fastest): unknowns, the
// observed boundary values of the ghost closures, periodic aliases and the corner zeros, as
// field_at gives them.  One thread per full-grid node and variable.
template <typename T>
__global__ void __launch_bounds__(256) expand_kernel(const KArgs<T>* __restrict__ kap, const T* __restrict__ u,
                                                     const T* __restrict__ p, double t, T* __restrict__ full,
                                                     int64_t ngrid) {
  const KArgs<T>& ka = *kap;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ngrid * ka.nfields) return;
  const int f = (int)(q / ngrid);
  int64_t r = q - (int64_t)f * ngrid;
  int I[3] = {1, 1, 1};
  for (int d = 0; d < ka.ndim; ++d) {
    I[d] = (int)(r % ka.n[d]) + 1;
    r /= ka.n[d];
  }
  full[q] = field_at(ka, u, f, I, t, p);
}

template <typename T>
int launch_expand(const Launch<T>& L, const T* u, const T* p, double t, T* full, hipStream_t stream) {
  const KArgs<T>& ka = *L.ka;
  int64_t ngrid = 1;
  for (int d = 0; d < ka.ndim; ++d) ngrid *= ka.n[d];
  const int64_t items = ngrid * ka.nfields;
  if (items <= 0) return 0;
  const int64_t blocks = (items + 255) / 256;
  if (blocks > 0x7fffffffLL) return MOL_ERR_INVALID;
  hipLaunchKernelGGL(expand_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, stream, L.dka, u, p, t, full, ngrid);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}
template int launch_expand<double>(const Launch<double>&, const double*, const double*, double, double*, hipStream_t);
template int launch_expand<float>(const Launch<float>&, const float*, const float*, double, float*, hipStream_t);

template <typename T>
__global__ void fill_zero_kernel(T* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

template <typename T>
int launch_fill_zero(T* p, int64_t n, hipStream_t stream) {
  if (n <= 0) return 0;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(blocks), dim3(256), 0, stream, p, n);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}

// lam += y0 + y1 + y2 + y3 in the order the stage adjoints are produced; the stage-adjoint
// buffers are cleared for the next step (they accumulate with atomics)
template <typename T>
__global__ void adjoint_sum_kernel(T* __restrict__ lam, T* __restrict__ y0, T* __restrict__ y1, T* __restrict__ y2,
                                   T* __restrict__ y3, int64_t n, bool zero) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    lam[i] = (((lam[i] + y0[i]) + y1[i]) + y2[i]) + y3[i];
    if (zero) {
      y0[i] = T(0);
      y1[i] = T(0);
      y2[i] = T(0);
      y3[i] = T(0);
    }
  }
}

template <typename T>
int launch_adjoint_sum(T* lam, T* y0, T* y1, T* y2, T* y3, int64_t n, bool zero, hipStream_t stream) {
  if (n <= 0) return 0;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(adjoint_sum_kernel<T>, dim3(blocks), dim3(256), 0, stream, lam, y0, y1, y2, y3, n, zero);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}

// The device clock slot takes the start time from the kernel argument: the value travels in
// the launch packet, so no host memory has to outlive the call.
__global__ void set_clock_kernel(double* __restrict__ slot, double t) { slot[0] = t; }

int launch_set_clock(double* slot, double t, hipStream_t stream) {
  hipLaunchKernelGGL(set_clock_kernel, dim3(1), dim3(64), 0, stream, slot, t);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}

// ---- FP64 error norms ---------------------------------------------------------------------
// One 64-lane wavefront reduction step: every lane adds (or takes the larger of) its value and
// the value DPP control CTRL moves into it; lanes whose source is masked off by ROW_MASK see 0,
// the identity of both reductions here (the values are sums of squares or absolute values).
// The two 32-bit halves of the double move with the same control.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(const double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum (MAX = false) or maximum (MAX = true) over the 64 lanes of a wavefront, in lane 63:
// quad_perm [1,0,3,2] and [2,3,0,1] combine lanes within quads, row_ror 4 and 8 within rows of
// 16, row_bcast 15 adds row r's lane 15 into row r + 1 (rows 1 and 3), row_bcast 31 lane 31
// into rows 2 and 3.  Returned to every lane by readlane 63.
template <bool MAX>
__device__ __forceinline__ double wave_reduce(double v) {
  auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
  v = op(v, dpp_move<0xb1, 0xf>(v));    // quad_perm [1,0,3,2]
  v = op(v, dpp_move<0x4e, 0xf>(v));    // quad_perm [2,3,0,1]
  v = op(v, dpp_move<0x124, 0xf>(v));   // row_ror 4
  v = op(v, dpp_move<0x128, 0xf>(v));   // row_ror 8
  v = op(v, dpp_move<0x142, 0xa>(v));   // row_bcast 15 into rows 1, 3
  v = op(v, dpp_move<0x143, 0xc>(v));   // row_bcast 31 into rows 2, 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Partial norms of blocks of 256 threads: part[3 b + k] for block b, k = 0 max |u - r|,
// 1 sum (u - r)^2, 2 sum r^2 (r = 0 when ref is null), accumulated in FP64 in a fixed order.
template <typename T>
__global__ void __launch_bounds__(256) norms_partial_kernel(const T* __restrict__ u, const double* __restrict__ ref,
                                                            int64_t n, double* __restrict__ part) {
  double mx = 0.0, sd = 0.0, sr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double r = ref ? ref[i] : 0.0;
    const double d = (double)u[i] - r;
    mx = fmax(mx, fabs(d));
    sd = fma(d, d, sd);
    sr = fma(r, r, sr);
  }
  mx = wave_reduce<true>(mx);
  sd = wave_reduce<false>(sd);
  sr = wave_reduce<false>(sr);
  __shared__ double w[3][4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    w[0][wave] = mx;
    w[1][wave] = sd;
    w[2][wave] = sr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = fmax(fmax(w[0][0], w[0][1]), fmax(w[0][2], w[0][3]));
    part[3 * blockIdx.x + 1] = ((w[1][0] + w[1][1]) + w[1][2]) + w[1][3];
    part[3 * blockIdx.x + 2] = ((w[2][0] + w[2][1]) + w[2][2]) + w[2][3];
  }
}

// The partials of nb blocks (one work-group of 256 threads): out = max, sqrt of both sums.
__global__ void __launch_bounds__(256) norms_final_kernel(const double* __restrict__ part, int nb,
                                                          double* __restrict__ out) {
  double mx = 0.0, sd = 0.0, sr = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    mx = fmax(mx, part[3 * b]);
    sd += part[3 * b + 1];
    sr += part[3 * b + 2];
  }
  mx = wave_reduce<true>(mx);
  sd = wave_reduce<false>(sd);
  sr = wave_reduce<false>(sr);
  __shared__ double w[3][4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    w[0][wave] = mx;
    w[1][wave] = sd;
    w[2][wave] = sr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = fmax(fmax(w[0][0], w[0][1]), fmax(w[0][2], w[0][3]));
    out[1] = sqrt(((w[1][0] + w[1][1]) + w[1][2]) + w[1][3]);
    out[2] = sqrt(((w[2][0] + w[2][1]) + w[2][2]) + w[2][3]);
  }
}

template <typename T>
int launch_norms(const T* u, const double* ref, int64_t n, double* part, int nb, double* out, hipStream_t stream) {
  hipLaunchKernelGGL(norms_partial_kernel<T>, dim3(nb), dim3(256), 0, stream, u, ref, n, part);
  hipLaunchKernelGGL(norms_final_kernel, dim3(1), dim3(256), 0, stream, part, nb, out);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}
template int launch_norms<double>(const double*, const double*, int64_t, double*, int, double*, hipStream_t);
template int launch_norms<float>(const float*, const double*, int64_t, double*, int, double*, hipStream_t);

template int launch_fill_zero<double>(double*, int64_t, hipStream_t);
template int launch_fill_zero<float>(float*, int64_t, hipStream_t);
template int launch_adjoint_sum<double>(double*, double*, double*, double*, double*, int64_t, bool, hipStream_t);
template int launch_adjoint_sum<float>(float*, float*, float*, float*, float*, int64_t, bool, hipStream_t);

}  // namespace mol
