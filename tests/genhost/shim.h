// Host shim for running generated kernel source on the CPU (test infrastructure).  Each kernel
// body runs once per work-group with a single "thread" whose loops cover the whole tile
// (NTHR is set to 1), so the phases separated by __syncthreads() keep their order.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
using std::max;
using std::min;
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__
#define __global__
#define __shared__ static
#define __launch_bounds__(x)
struct MolDim3 {
  unsigned x = 0, y = 0, z = 0;
};
static MolDim3 threadIdx, blockIdx, gridDim, blockDim;
inline void __syncthreads() {}
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
template <class T>
inline T atomicAdd(T* p, T v) {
  const T o = *p;
  *p += v;
  return o;
}
template <class T>
inline void unsafeAtomicAdd(T* p, T v) {
  *p += v;
}
