// Host-side internals shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "mol.h"
#include "mol_device.h"

namespace mol {

template <typename T>
int launch_rhs_generic(const KArgs<T>& ka, const KArgs<T>* dka, const T* u, const Epi<T>& ep,
                       hipStream_t stream);
int launch_advance_time(double* tnow, double dt, hipStream_t stream);

// Set the message of the calling thread's last failure (used when no handle exists yet).
void set_global_error(const std::string& msg);
const char* global_error();

}  // namespace mol

struct mol_handle {
  int device = 0;
  int dtype = MOL_F64;
  int ndim = 1, nfields = 0, naux = 0, nparams = 0;
  int64_t nunk = 0;
  int64_t ngrid = 1;   // full-grid node count
  mol::KArgs<double> ka64{};
  mol::KArgs<float> ka32{};
  mol::KArgs<double>* d_ka64 = nullptr;   // device copies read by the kernels
  mol::KArgs<float>* d_ka32 = nullptr;
  // device copies of the descriptor tables
  void* d_tabs = nullptr;
  void* d_slots = nullptr;
  void* d_clos = nullptr;
  void* d_pool = nullptr;
  void* d_code = nullptr;
  void* d_cpool = nullptr;
  void* d_prog = nullptr;
  void* d_coord[3] = {nullptr, nullptr, nullptr};
  // integrator workspace
  void* w_y1 = nullptr;
  void* w_y2 = nullptr;
  void* w_acc = nullptr;
  void* w_comp = nullptr;
  double* d_tnow = nullptr;
  // rhs_h staging
  void* h_u = nullptr;
  void* h_du = nullptr;
  void* h_p = nullptr;
  std::string err;
  size_t esize() const { return dtype == MOL_F64 ? sizeof(double) : sizeof(float); }
};
