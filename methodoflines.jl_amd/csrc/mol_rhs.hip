// Generic stencil-IR right-hand-side kernel with fused Runge-Kutta stage epilogues.
#include "mol_device.h"
#include "mol_internal.h"

namespace mol {

template <typename T>
__device__ __forceinline__ void tile_origin(const KArgs<T>& ka, int* org) {
  int b = blockIdx.x;
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

template <typename T>
__global__ void __launch_bounds__(256) rhs_generic_kernel(const KArgs<T>* __restrict__ kap,
                                                          const T* __restrict__ u, Epi<T> ep) {
  const KArgs<T>& ka = *kap;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  int org[3];
  tile_origin(ka, org);
  const double t = (ep.tnow ? ep.tnow[0] : 0.0) + ep.tofs;
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

template <typename T>
int launch_rhs_generic(const KArgs<T>& ka, const KArgs<T>* dka, const T* u, const Epi<T>& ep,
                       hipStream_t stream) {
  const int nblocks = ka.ntile[0] * ka.ntile[1] * ka.ntile[2];
  const size_t lds = (size_t)(ka.nfields + ka.naux) * ka.vol * sizeof(T);
  hipLaunchKernelGGL(rhs_generic_kernel<T>, dim3(nblocks), dim3(256), lds, stream, dka, u, ep);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}

template int launch_rhs_generic<double>(const KArgs<double>&, const KArgs<double>*, const double*,
                                        const Epi<double>&, hipStream_t);
template int launch_rhs_generic<float>(const KArgs<float>&, const KArgs<float>*, const float*,
                                       const Epi<float>&, hipStream_t);

__global__ void advance_time_kernel(double* tnow, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) tnow[0] += dt;
}

int launch_advance_time(double* tnow, double dt, hipStream_t stream) {
  hipLaunchKernelGGL(advance_time_kernel, dim3(1), dim3(64), 0, stream, tnow, dt);
  return hipGetLastError() == hipSuccess ? 0 : MOL_ERR_HIP;
}

}  // namespace mol
