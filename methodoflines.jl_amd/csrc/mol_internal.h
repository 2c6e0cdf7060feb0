// Host-side internals shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "mol.h"
#include "mol_core.h"
#include "mol_device.h"

namespace mol {

// Geometry of the core-tile region handed to the code generator.
struct JitGeom {
  int n[3], periodic[3];
  int lo[MOL_MAX_FIELDS][3], hi[MOL_MAX_FIELDS][3];   // unknown box per variable
  long long fofs[MOL_MAX_FIELDS];                     // offset of each variable in the state
  int ulo[3], uhi[3];                                 // union box (tile origin)
  int tile[3], halo[3], ext[3], vol, ntile[3];
  int ct_lo[3], ct_n[3], ncore;                       // core tiles (interior rows only)
  int nthr;                                           // threads per work-group (multiple of 64)
  int fsteps = 1;                                     // fused RK4: steps per launch (1 or 2)
};

// Pointer arguments of the generated kernel (layout identical to JPtrs of mol_jit_prelude.h).
template <typename T>
struct JitPtrs {
  const double* coord[3];
  const T* aux[MOL_MAX_AUX];
  const double* pool;   // descriptor weight pool (per-node rows of nonuniform axes)
  const double* wtab[3];   // per axis: nonuniform WENO5 coefficient rows (generated kernels)
};

// Code generation and in-process compilation of the tile kernel (mol_jit.cpp).
std::string jit_source(const mol_desc* d, const JitGeom& g, std::string& err);
std::string jit_vjp_source(const mol_desc* d, const JitGeom& g, std::string& err);
// Fused RK4 step kernel (mol_rk4_fused); empty with err set when the system does not qualify.
std::string jit_fused_source(const mol_desc* d, const JitGeom& g, std::string& err);
// LDS bytes per tile + halo position of the generated forward (vjp = false) or reverse-mode
// kernel: the arrays it stages (and their adjoints); 0 when the system cannot be generated.
size_t jit_lds_per_pos(const mol_desc* d, const JitGeom& g, size_t esize, bool vjp);
std::string jit_class_source(const mol_desc* d, const JitGeom& g, const std::string& name, std::string& err);

// Header of one instance of a batch (layout identical to BInst of mol_jit_prelude.h).
struct BInst {
  int32_t n[3];
  int32_t lo[MOL_MAX_FIELDS][3];
  int32_t hi[MOL_MAX_FIELDS][3];
  int64_t fofs[MOL_MAX_FIELDS];
  int32_t ulo[3], uhi[3], nt[3];
  int32_t slo[3], shi[3], clo[3], chi[3];
  int32_t ilo[3], ihi[3];   // nodes where the interior-rows code applies
  int32_t cls, first;
  const double* coord[3];
  const double* W;
  const double* K;
};
int jit_compile(const std::string& src, std::vector<char>& code, std::string& log);
int jit_kernel(const std::string& src, const char* name, int device, hipFunction_t* fn, std::string& err);

// Arguments of the generated vector-Jacobian product kernel (layout identical to VArgs of
// mol_jit_prelude.h): seed c1*s1 + c2*s2, result accumulated into out, FP64 aux gradients.
template <typename T>
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
  T* bf;            // per-tile adjoint buffers of the staged variables
  double* ba;       // per-tile adjoint buffers of the staged grid fields
  int mode;         // VJP_GATHER_F | VJP_GATHER_A | VJP_ACC_A | VJP_SUM
  T* lam;           // VJP_SUM: the gather (which stores) finishes the step, lam = lam + y0 + y1 + y2 + gathered
  const T* y0;
  const T* y1;
  const T* y2;
};
enum { VJP_GATHER_F = 1, VJP_GATHER_A = 2, VJP_ACC_A = 4, VJP_SUM = 8 };

// What one right-hand-side launch needs: the core-tile kernel arguments and the list of tiles
// left to the generic kernel.
template <typename T>
struct Launch {
  const KArgs<T>* ka = nullptr;      // host copy (sizes)
  const KArgs<T>* dka = nullptr;     // device copy
  const KCore<T>* kc = nullptr;      // core kernel arguments (host)
  const int32_t* tiles = nullptr;    // device list of generic tiles
  int ncore = 0, nbound = 0;
  hipFunction_t jit = nullptr;       // generated kernel over all tiles (nullptr: interpreter)
  int ntiles = 0;
  int jgrid = 0;                  // work-groups of the persistent generated kernel
  int jthr = 256;                 // threads per work-group of the generated kernel
  hipFunction_t batch = nullptr;     // persistent kernel over the work list of a batch
  const BInst* insts = nullptr;
  const int32_t* wl = nullptr;
  int nwork = 0, bgrid = 0;
  JitPtrs<T> jp{};
};

template <typename T>
int launch_rhs(const Launch<T>& L, const T* u, const Epi<T>& ep, hipStream_t stream);

// Full grid of every variable from the unknowns at time t (mol_expand).
template <typename T>
int launch_expand(const Launch<T>& L, const T* u, const T* p, double t, T* full, hipStream_t stream);

template <typename T>
int launch_fill_zero(T* p, int64_t n, hipStream_t stream);
// lam += y0 + y1 + y2 + y3 (the adjoint of one RK4 step), then y0..y3 = 0 when `zero`
template <typename T>
int launch_adjoint_sum(T* lam, T* y0, T* y1, T* y2, T* y3, int64_t n, bool zero, hipStream_t stream);
// *slot = t on `stream` (t passed by value in the launch)
int launch_set_clock(double* slot, double t, hipStream_t stream);
// FP64 norms of u - ref over n values (ref may be null): out[0..2] = max |u - ref|,
// ||u - ref||_2, ||ref||_2; part holds 3 * nb doubles of per-block partials
template <typename T>
int launch_norms(const T* u, const double* ref, int64_t n, double* part, int nb, double* out, hipStream_t stream);
constexpr int kNormBlocks = 1024;

// Set the message of the calling thread's last failure (used when no handle exists yet).
void set_global_error(const std::string& msg);
const char* global_error();

}  // namespace mol

// Owned copy of a descriptor (the handle re-plans launches from it).
struct DescCopy {
  mol_desc d{};
  std::vector<double> coord[3];
  std::vector<mol_slot> slots;
  std::vector<mol_optab> tabs;
  std::vector<mol_closure> clos;
  std::vector<double> pool, cpool;
  std::vector<uint32_t> code;
  std::vector<int32_t> prog;
  void assign(const mol_desc* s) {
    d = *s;
    for (int a = 0; a < s->ndim; ++a) {
      coord[a].assign(s->coord_h[a], s->coord_h[a] + s->n[a]);
      d.coord_h[a] = coord[a].data();
    }
    slots.assign(s->slots_h, s->slots_h + s->nslots);
    tabs.assign(s->tabs_h, s->tabs_h + s->ntabs);
    clos.assign(s->closures_h, s->closures_h + s->nclosures);
    pool.assign(s->pool_h, s->pool_h + s->npool);
    cpool.assign(s->cpool_h, s->cpool_h + s->ncpool);
    code.assign(s->code_h, s->code_h + s->ncode);
    prog.assign(s->prog_h, s->prog_h + 2 * (size_t)s->nprogs);
    d.slots_h = slots.data();
    d.tabs_h = tabs.data();
    d.closures_h = clos.data();
    d.pool_h = pool.data();
    d.cpool_h = cpool.data();
    d.code_h = code.data();
    d.prog_h = prog.data();
  }
};

struct mol_handle {
  DescCopy desc;
  int tile_req[3] = {0, 0, 0};            // requested tile shape (0: default)
  int device = 0;
  int dtype = MOL_F64;
  int ndim = 1, nfields = 0, naux = 0, nparams = 0;
  int64_t nunk = 0;
  int64_t ngrid = 1;   // full-grid node count
  mol::KArgs<double> ka64{};
  mol::KArgs<float> ka32{};
  mol::KArgs<double>* d_ka64 = nullptr;   // device copies read by the kernels
  mol::KArgs<float>* d_ka32 = nullptr;
  mol::KCore<double> kc64{};
  mol::KCore<float> kc32{};
  int32_t* d_tiles = nullptr;             // tiles handled by the generic kernel
  int ncore = 0, nbound = 0;
  std::string core_reason;                // why the core path is off, if it is
  bool core_interp_ok = false;            // the stencil IR fits the interpreter's argument block
  hipFunction_t jit_fn = nullptr;         // generated kernel over all tiles
  std::string jit_reason;                 // why it is not used, if it is not
  // reverse mode
  // batch of instances (mol_create_batched)
  bool batch = false;
  hipFunction_t batch_fn = nullptr;
  int nwork = 0, batch_grid = 0, nclasses = 0, ninst = 0;
  int batch_tile[3] = {0, 0, 0};
  void* d_insts = nullptr;
  void* d_wl = nullptr;
  void* d_bdata = nullptr;                // weights, constants and coordinates of all instances
  std::vector<int64_t> inst_off;          // first unknown of each instance
  int jit_grid = 0;                       // work-groups of the persistent generated RHS kernel
  int jit_nthr = 256;                     // threads per work-group of the generated RHS kernel
  hipFunction_t fused_fn = nullptr;       // generated fused RK4 step (periodic 1D/2D systems)
  int fused_tiles = 0;
  int fused_nthr = 256;
  bool fused_resident = false;     // one 1D tile: many steps per fused launch
  int fused_steps = 1;             // steps per fused launch otherwise (1, or 2 with an 8-reach halo)
  std::string fused_reason;               // why there is no fused kernel
  bool no_fuse = false;                   // this integration runs the stage kernels (MOL_FLAG_UNFUSED)
  int vjp_nthr = 256;                     // threads per work-group of the generated VJP kernel
  double* d_norm_part = nullptr;          // per-block partials of mol_error_norms
  double* d_wtab[3] = {nullptr, nullptr, nullptr};   // nonuniform WENO5 coefficient rows per axis
  hipFunction_t vjp_fn = nullptr;         // generated vector-Jacobian product kernel
  bool vjp_store = false;                 // its gather stores every unknown (no zeroing of the output)
  hipFunction_t vjp_gather_fn = nullptr;  // its gather of the per-tile adjoint buffers
  void* vjp_bf = nullptr;                 // per-tile adjoint buffers: staged variables (T)
  void* vjp_ba = nullptr;                 //   and staged grid fields (FP64)
  size_t vjp_ba_count = 0;                //   (values in vjp_ba)
  int vjp_tiles = 0;
  bool vjp_aux_ok = true;                 // no boundary value reads an auxiliary array
  void* ck_u = nullptr;                   // checkpoints of the last mol_forward_checkpointed run
  double* ck_t = nullptr;                 // device clock at each checkpoint
  size_t ck_cap = 0;
  int64_t ck_nsteps = 0;
  int32_t ck_every = 0;
  double ck_dt = 0.0;
  bool ck_valid = false;
  bool ck_stored = false;                 // the run kept every stage state (no recomputation)
  void* seg_u = nullptr;                  // stage states of one recomputed segment (or the whole run)
  double* seg_t = nullptr;                // device clock at each step of the segment
  size_t seg_cap = 0;
  void* adj_y[4] = {nullptr, nullptr, nullptr, nullptr};
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
  double* d_tnow = nullptr;                // double-buffered device clock: d_tnow[clk] is current
  int clk = 0;
  // Instantiated step chunks of mol_integrate (MOL_FLAG_GRAPH), captured once per key and
  // replayed: the kernels of a chunk take their pointers and dt by value, so a chunk is
  // valid only for the state, parameters and step it was captured with.  Every chunk is an
  // even number of clock flips and starts at clock slot 0.
  struct GraphChunk {
    const void* u;
    const void* p;
    double dt;
    int scheme, kahan, nofuse;
    int64_t steps;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<GraphChunk> graphs;
  hipStream_t cap_stream = nullptr;       // capture-only stream (nothing executes on it)
  // mol_rhs_h: device copies of u, du and p, allocated on the first call
  void* r_u = nullptr;
  void* r_du = nullptr;
  void* r_p = nullptr;
  hipStream_t rhs_stream = nullptr;
  std::string err;
  size_t esize() const { return dtype == MOL_F64 ? sizeof(double) : sizeof(float); }
};
