// C ABI of the stencil library: descriptor validation/upload, RHS and stage-loop drivers.
#include <array>
#include <cstdio>
#include <map>
#include <mutex>
#include <sstream>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "mol_internal.h"

namespace mol {

static thread_local std::string g_err;
void set_global_error(const std::string& msg) { g_err = msg; }
const char* global_error() { return g_err.c_str(); }

}  // namespace mol

using namespace mol;

// Workspace of mol_integrate_noalloc: a device array of this library's code object, reserved on
// each device when the code object is loaded there and held until the process ends.  Its name
// is a host-side stand-in; the device address comes from hipGetSymbolAddress per device.
static __device__ __align__(256) unsigned char scratch[MOL_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};          // device address of `scratch` on each device
hipError_t g_scratch_err[kMaxDevices] = {};     // result of the lookup on each device
bool g_scratch_done[kMaxDevices] = {};

// Look up the current device's address of `scratch` once per device (called by mol_create and
// mol_create_batched with the handle's device current).
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
  g_scratch_done[device] = true;
}

void* scratch_of(int device, std::string& err) {
  if (device < 0 || device >= kMaxDevices) {
    err = "device index out of range for the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[device]) {
    err = "scratch of this device was not looked up (no handle created on it)";
    return nullptr;
  }
  if (!g_scratch_ptr[device]) err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_err[device]);
  return g_scratch_ptr[device];
}

#define HIPCHK(h, call)                                                           \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
      return MOL_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

template <typename T>
KArgs<T>& kargs(mol_handle* h);
template <typename T>
KArgs<T>* dkargs(mol_handle* h);

int fail(mol_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else set_global_error(msg);
  return code;
}

template <typename T>
int upload(mol_handle* h, void** dst, const T* src, size_t count) {
  if (count == 0) {
    *dst = nullptr;
    return 0;
  }
  HIPCHK(h, hipMalloc(dst, count * sizeof(T)));
  HIPCHK(h, hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Release the captured step chunks (after mol_set_aux / mol_set_tile they are stale).  A
// replay may still be queued, so the device is drained first; this runs only when the
// handle is reconfigured, destroyed, or more than eight step keys have been captured.
void drop_graphs(mol_handle* h) {
  if (h->graphs.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& c : h->graphs) {
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
  }
  h->graphs.clear();
}

void free_handle(mol_handle* h) {
  if (!h) return;
  drop_graphs(h);
  if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
  if (h->rhs_stream) (void)hipStreamDestroy(h->rhs_stream);
  void* ptrs[] = {h->d_tabs, h->d_slots, h->d_clos, h->d_pool, h->d_code, h->d_cpool,
                  h->d_prog, h->d_coord[0], h->d_coord[1], h->d_coord[2], h->w_y1,
                  h->w_y2, h->w_acc, h->w_comp, h->d_tnow, h->d_ka64, h->d_ka32, h->d_tiles,
                  h->ck_u, h->ck_t, h->seg_u, h->seg_t, h->adj_y[0], h->adj_y[1], h->adj_y[2], h->adj_y[3],
                  h->d_insts, h->d_wl, h->d_bdata, h->vjp_bf, h->vjp_ba, h->r_u, h->r_du, h->r_p,
                  h->d_norm_part, h->d_wtab[0], h->d_wtab[1], h->d_wtab[2]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete h;
}

}  // namespace


namespace {

// Structural validation of a descriptor against the limits the kernels assume.
int validate(const mol_desc* d, std::string& msg) {
  if (!d) { msg = "descriptor is NULL"; return MOL_ERR_INVALID; }
  if (d->abi_version != MOL_ABI_VERSION) { msg = "ABI version mismatch"; return MOL_ERR_INVALID; }
  if (d->ndim < 1 || d->ndim > MOL_MAX_DIMS) { msg = "ndim must be 1..3"; return MOL_ERR_INVALID; }
  if (d->nfields < 1 || d->nfields > MOL_MAX_FIELDS) { msg = "nfields out of range"; return MOL_ERR_INVALID; }
  if (d->naux < 0 || d->naux > MOL_MAX_AUX) { msg = "naux out of range"; return MOL_ERR_INVALID; }
  if (d->dtype != MOL_F64 && d->dtype != MOL_F32) { msg = "dtype must be MOL_F64 or MOL_F32"; return MOL_ERR_INVALID; }
  if (d->nparams < 0) { msg = "nparams < 0"; return MOL_ERR_INVALID; }
  for (int a = 0; a < d->ndim; ++a) {
    if (d->n[a] < 2 || d->n[a] > (1 << 28)) { msg = "grid size out of range"; return MOL_ERR_SHAPE; }
    if (!d->coord_h[a]) { msg = "missing coordinates"; return MOL_ERR_INVALID; }
    if (d->halo[a] < 0 || d->halo[a] > 16) { msg = "halo out of range"; return MOL_ERR_INVALID; }
    if (d->periodic[a] && d->n[a] - 1 < 2 * d->halo[a] + 1) { msg = "periodic ring shorter than the stencil"; return MOL_ERR_SHAPE; }
  }
  for (int f = 0; f < d->nfields; ++f)
    for (int a = 0; a < d->ndim; ++a) {
      if (d->lo[f][a] < 1 || d->hi[f][a] > d->n[a] || d->lo[f][a] > d->hi[f][a]) {
        msg = "unknown box out of the grid";
        return MOL_ERR_SHAPE;
      }
      if (d->periodic[a] && (d->lo[f][a] != 2 || d->hi[f][a] != d->n[a])) {
        msg = "periodic axis must hold nodes 2..n as unknowns";
        return MOL_ERR_SHAPE;
      }
    }
  if (d->nslots < 0 || d->nslots > MOL_MAX_SLOTS) { msg = "nslots out of range"; return MOL_ERR_INVALID; }
  if (d->nprogs < 0 || (d->nprogs > 0 && !d->prog_h)) { msg = "programs missing"; return MOL_ERR_INVALID; }
  for (int i = 0; i < d->nprogs; ++i) {
    int s = d->prog_h[2 * i], l = d->prog_h[2 * i + 1];
    if (s < 0 || l < 0 || s + l > d->ncode) { msg = "program range out of the code pool"; return MOL_ERR_INVALID; }
  }
  for (int pc = 0; pc < d->ncode; ++pc) {
    uint32_t ins = d->code_h[pc];
    uint32_t op = ins & 0xff;
    int imm = (int)(ins >> 21);
    bool bconst = (ins >> 20) & 1u;
    if (bconst && imm >= d->ncpool) { msg = "constant index out of range"; return MOL_ERR_INVALID; }
    if (op == MOL_OP_CONST && imm >= d->ncpool) { msg = "constant index out of range"; return MOL_ERR_INVALID; }
    if (op == MOL_OP_FIELD && imm >= d->nfields) { msg = "variable index out of range"; return MOL_ERR_INVALID; }
    if (op == MOL_OP_SLOT && imm >= d->nslots) { msg = "slot index out of range"; return MOL_ERR_INVALID; }
    if (op == MOL_OP_COORD && imm >= d->ndim) { msg = "axis index out of range"; return MOL_ERR_INVALID; }
    if (op == MOL_OP_PARAM && imm >= d->nparams) { msg = "parameter index out of range"; return MOL_ERR_INVALID; }
    if (op == MOL_OP_AUX && imm >= d->naux) { msg = "aux index out of range"; return MOL_ERR_INVALID; }
  }
  for (int f = 0; f < d->nfields; ++f)
    if (d->rhs_prog[f] < 0 || d->rhs_prog[f] >= d->nprogs) { msg = "rhs program missing"; return MOL_ERR_INVALID; }
  // every program must follow the stack discipline the interpreter relies on: an instruction
  // at depth k writes register k-1 (binary: reads k-2, k-1), leaves push, depth <= MOL_STACK
  for (int i = 0; i < d->nprogs; ++i) {
    const int st = d->prog_h[2 * i], ln = d->prog_h[2 * i + 1];
    int depth = 0;
    for (int pc = st; pc < st + ln; ++pc) {
      const uint32_t ins = d->code_h[pc];
      const int op = ins & 0xff, dst = (ins >> 8) & 15, ra = (ins >> 12) & 15, rb = (ins >> 16) & 15;
      const bool bc = (ins >> 20) & 1u;
      bool ok;
      if (op < 16) { ok = op != MOL_OP_MOV && dst == depth; depth += 1; }
      else if (op == MOL_OP_FMA) ok = false;
      else if (op == MOL_OP_SELECT) {
        if (bc) { ok = dst == depth - 2 && ra == depth - 1; depth -= 1; }
        else { ok = dst == depth - 3 && ra == depth - 2 && rb == depth - 1; depth -= 2; }
      } else if (op < 48) {
        if (bc) ok = dst == depth - 1 && ra == depth - 1;
        else { ok = dst == depth - 2 && ra == depth - 2 && rb == depth - 1; depth -= 1; }
      } else ok = dst == depth - 1 && ra == depth - 1;
      if (!ok || depth < 0 || depth > MOL_STACK) { msg = "program does not follow the stack discipline or is deeper than MOL_STACK"; return MOL_ERR_UNSUPPORTED; }
    }
    if (depth != 1) { msg = "program does not leave exactly one value"; return MOL_ERR_INVALID; }
  }
  auto tab_ok = [&](int t) { return t >= 0 && t < d->ntabs; };
  for (int k = 0; k < d->nslots; ++k) {
    const mol_slot& s = d->slots_h[k];
    if (s.field < 0 || s.field >= d->nfields || s.axis < 0 || s.axis >= d->ndim) { msg = "slot operand out of range"; return MOL_ERR_INVALID; }
    if (s.prog >= d->nprogs) { msg = "slot program out of range"; return MOL_ERR_INVALID; }
    int ntab = 0;
    switch (s.kind) {
      case MOL_SLOT_CENTERED: ntab = 1; break;
      case MOL_SLOT_UPWIND: ntab = 2; break;
      case MOL_SLOT_WENO: ntab = 0; break;
      case MOL_SLOT_WENO_NU:
        ntab = 0;
        if (d->n[s.axis] < 5) { msg = "nonuniform WENO needs at least 5 nodes"; return MOL_ERR_INVALID; }
        break;
      case MOL_SLOT_NONLINLAP: ntab = 3; break;
      case MOL_SLOT_SPHERICAL: ntab = 5; break;
      case MOL_SLOT_MIXED:
        ntab = 2;
        if (s.axis2 < 0 || s.axis2 >= d->ndim || s.axis2 == s.axis) { msg = "mixed slot axes"; return MOL_ERR_INVALID; }
        break;
      default: msg = "unknown slot kind"; return MOL_ERR_INVALID;
    }
    for (int q = 0; q < ntab; ++q) {
      if (!tab_ok(s.tab[q])) { msg = "slot table out of range"; return MOL_ERR_INVALID; }
      const mol_optab& tb = d->tabs_h[s.tab[q]];
      if (tb.pn_hi > 0) {
        const int tax = s.kind == MOL_SLOT_MIXED && q == 1 ? s.axis2 : s.axis;   // the table's axis
        if (d->periodic[tax] || tb.pn_hi > d->n[tax]) { msg = "per-node rows on this axis"; return MOL_ERR_INVALID; }
      }
    }
  }
  for (int t = 0; t < d->ntabs; ++t) {
    const mol_optab& tb = d->tabs_h[t];
    if (tb.len < 0 || tb.w < 0 || tb.w + tb.len > d->npool) { msg = "table weights out of the pool"; return MOL_ERR_INVALID; }
    if (tb.nlow < 0 || tb.nhigh < 0 || tb.blen < 0) { msg = "table rows"; return MOL_ERR_INVALID; }
    if (tb.pn_hi != 0 || tb.pn_lo != 0) {
      if (tb.pn_lo < 1 || tb.pn_hi < tb.pn_lo) { msg = "per-node row range"; return MOL_ERR_INVALID; }
      if ((int64_t)tb.w + (int64_t)(tb.pn_hi - tb.pn_lo + 1) * tb.len > d->npool) {
        msg = "per-node rows out of the pool";
        return MOL_ERR_INVALID;
      }
    }
    if (tb.nlow && tb.wlow + tb.nlow * tb.blen > d->npool) { msg = "table low rows out of the pool"; return MOL_ERR_INVALID; }
    if (tb.nhigh && tb.whigh + tb.nhigh * tb.blen > d->npool) { msg = "table high rows out of the pool"; return MOL_ERR_INVALID; }
  }
  for (int c = 0; c < d->nclosures; ++c) {
    const mol_closure& cl = d->closures_h[c];
    if (cl.field < 0 || cl.field >= d->nfields || cl.axis < 0 || cl.axis >= d->ndim || cl.side < 0 || cl.side > 1) { msg = "closure target"; return MOL_ERR_INVALID; }
    if (cl.K < 0 || cl.K > MOL_MAX_CLOSURE_K || cl.nghost < 0 || cl.nghost > MOL_MAX_GHOST) { msg = "closure size"; return MOL_ERR_INVALID; }
    if (cl.src < -1 || cl.src >= d->nfields || cl.src == cl.field) { msg = "interface variable out of range"; return MOL_ERR_INVALID; }
    if (cl.src >= 0 && (cl.K != 0 || cl.gprog >= 0 || cl.nghost < 1)) { msg = "interface closure"; return MOL_ERR_INVALID; }
    int64_t boxlen = d->hi[cl.field][cl.axis] - d->lo[cl.field][cl.axis] + 1;
    if (cl.K > boxlen) { msg = "closure reaches past the unknown box"; return MOL_ERR_SHAPE; }
    if (cl.wC + cl.nghost * cl.K > d->npool || cl.wc + cl.nghost > d->npool) { msg = "closure weights out of the pool"; return MOL_ERR_INVALID; }
    if (cl.gprog >= d->nprogs) { msg = "closure program out of range"; return MOL_ERR_INVALID; }
    if (d->periodic[cl.axis]) { msg = "closure on a periodic axis"; return MOL_ERR_INVALID; }
  }
  return 0;
}

// lds_pos: LDS bytes per tile + halo position (0: the forward kernels' arrays)
// clamp: shrink the tile to the box (a batch keeps one tile shape for all its instances)
template <typename T>
void fill_kargs(mol_handle* h, const mol_desc* d, KArgs<T>& ka, size_t lds_pos = 0, bool clamp = true,
                int jit_tiles = 0, size_t lds_cap = 60000) {
  std::memset(&ka, 0, sizeof(ka));
  ka.ndim = d->ndim;
  ka.nfields = d->nfields;
  ka.naux = d->naux;
  ka.nparams = d->nparams;
  int64_t off = 0;
  for (int a = 0; a < 3; ++a) {
    ka.n[a] = a < d->ndim ? (int32_t)d->n[a] : 1;
    ka.periodic[a] = a < d->ndim ? d->periodic[a] : 0;
    ka.halo[a] = a < d->ndim ? d->halo[a] : 0;
    ka.ulo[a] = 1 << 30;
    ka.uhi[a] = 0;
  }
  for (int f = 0; f < d->nfields; ++f) {
    int64_t sz = 1;
    for (int a = 0; a < 3; ++a) {
      ka.lo[f][a] = a < d->ndim ? (int32_t)d->lo[f][a] : 1;
      ka.hi[f][a] = a < d->ndim ? (int32_t)d->hi[f][a] : 1;
      sz *= ka.hi[f][a] - ka.lo[f][a] + 1;
      ka.ulo[a] = std::min(ka.ulo[a], ka.lo[f][a]);
      ka.uhi[a] = std::max(ka.uhi[a], ka.hi[f][a]);
    }
    ka.fofs[f] = off;
    off += sz;
    ka.rhs_prog[f] = d->rhs_prog[f];
    for (int a = 0; a < 3; ++a) ka.clos_id[f][a][0] = ka.clos_id[f][a][1] = -1;
  }
  h->nunk = off;
  for (int c = 0; c < d->nclosures; ++c) {
    const mol_closure& cl = d->closures_h[c];
    ka.clos_id[cl.field][cl.axis][cl.side] = c;
  }
  // tile shape: contiguous along x, sized so that tile + halo of every staged array fits LDS
  // default: one node per thread (256-node tiles), contiguous along x
  int tile[3] = {256, 1, 1};
  if (d->ndim == 2) { tile[0] = 64; tile[1] = 4; }
  if (d->ndim == 3) { tile[0] = 32; tile[1] = 4; tile[2] = 2; }
  if (h->tile_req[0] > 0)
    for (int a = 0; a < 3; ++a) tile[a] = a < d->ndim ? std::max(1, h->tile_req[a]) : 1;
  const int nload = d->nfields + d->naux;
  auto vol_of = [&](const int* tl) {
    int v = 1;
    for (int a = 0; a < 3; ++a) v *= tl[a] + 2 * ka.halo[a];
    return v;
  };
  for (int a = 0; a < 3; ++a)
    tile[a] = a >= d->ndim ? 1 : (clamp ? std::max(1, std::min(tile[a], ka.uhi[a] - ka.ulo[a] + 1)) : tile[a]);
  while (tile[0] * tile[1] * tile[2] > 4096) {
    int big = 0;
    for (int a = 1; a < 3; ++a)
      if (tile[a] > tile[big]) big = a;
    tile[big] = (tile[big] + 1) / 2;
  }
  if (lds_pos == 0) lds_pos = (size_t)nload * sizeof(T);
  if (jit_tiles && h->tile_req[0] <= 0 && d->ndim >= 2) {
    // generated kernels run several nodes per thread along the slowest axis: the first
    // candidate whose tile + halo of the staged arrays fits the LDS budget (forward: about
    // three to four work-groups of 256 threads per compute unit; reverse mode (jit_tiles 2):
    // one work-group of up to 1024 threads and 160 KB) and that still gives every compute unit
    // four tiles (reverse mode: one); small grids fall back to the candidate with the most tiles
    // 2D forward stencils of reach two or more along both axes (WENO5, fourth-order rows)
    // start from 64x32 tiles of 512 threads, two nodes per thread in registers and two node
    // groups in turn, wanted on two work-groups per compute unit (Burgers WENO5 FP64: 106 us
    // against 116 us at 4096^2, 9.1 against 9.2 us at 1024^2)
    static const int cand2[][3] = {{64, 32, 1}, {64, 16, 1}, {32, 16, 1}, {64, 8, 1}, {64, 4, 1}, {32, 8, 1}, {32, 4, 1}};
    static const int cand3[][3] = {{16, 8, 8}, {16, 8, 4}, {16, 4, 8}, {16, 4, 4}, {32, 4, 2}, {16, 4, 2}, {8, 4, 4}};
    const bool rev = jit_tiles == 2;
    const bool wide = d->ndim == 2 && !rev && ka.halo[0] >= 2 && ka.halo[1] >= 2;
    const int (*cand)[3] = d->ndim == 2 ? (wide ? cand2 : cand2 + 1) : cand3;
    const int nc = d->ndim == 2 ? (wide ? 7 : 6) : 7;
    const size_t budget = rev ? 163000 : (d->ndim == 2 ? 40960 : 81920);
    const int64_t enough = rev ? 256 : 1024;
    int best = -1;
    int64_t best_tiles = -1;
    for (int i = 0; i < nc; ++i) {
      int tl[3];
      int64_t nt = 1;
      for (int a = 0; a < 3; ++a) {
        const int ext = ka.uhi[a] - ka.ulo[a] + 1;
        tl[a] = a >= d->ndim ? 1 : std::max(1, std::min(cand[i][a], ext));
        nt *= (ext + tl[a] - 1) / tl[a];
      }
      if (lds_pos * vol_of(tl) > budget) continue;
      if (nt >= (wide && i == 0 ? enough / 2 : enough)) {
        best = i;
        break;
      }
      if (nt > best_tiles) {
        best_tiles = nt;
        best = i;
      }
    }
    if (best >= 0)
      for (int a = 0; a < 3; ++a)
        tile[a] = a >= d->ndim ? 1 : std::max(1, std::min(cand[best][a], ka.uhi[a] - ka.ulo[a] + 1));
  }
  if (jit_tiles == 1 && d->ndim == 3) lds_cap = std::max<size_t>(lds_cap, 81920);
  while (lds_pos * vol_of(tile) > lds_cap) {
    int big = 0;
    for (int a = 1; a < 3; ++a)
      if (tile[a] > tile[big] / 4) big = a;
    if (tile[big] <= 1) break;
    tile[big] = (tile[big] + 1) / 2;
  }
  for (int a = 0; a < 3; ++a) {
    ka.tile[a] = tile[a];
    ka.ntile[a] = (ka.uhi[a] - ka.ulo[a] + 1 + tile[a] - 1) / tile[a];
    ka.ext[a] = tile[a] + 2 * ka.halo[a];
  }
  ka.vol = ka.ext[0] * ka.ext[1] * ka.ext[2];
  for (int a = 0; a < 3; ++a) ka.coord[a] = (const double*)h->d_coord[a];
  ka.tabs = (const Tab*)h->d_tabs;
  ka.slots = (const Slot*)h->d_slots;
  ka.clos = (const Clos*)h->d_clos;
  ka.pool = (const double*)h->d_pool;
  ka.code = (const uint32_t*)h->d_code;
  ka.cpool = (const double*)h->d_cpool;
  ka.prog = (const int32_t*)h->d_prog;
}

template <typename T>
KCore<T>& kcore(mol_handle* h);
template <>
KCore<double>& kcore<double>(mol_handle* h) { return h->kc64; }
template <>
KCore<float>& kcore<float>(mol_handle* h) { return h->kc32; }

template <typename T>
Launch<T> launch_of(mol_handle* h) {
  Launch<T> L;
  L.ka = &kargs<T>(h);
  L.dka = dkargs<T>(h);
  L.kc = &kcore<T>(h);
  L.tiles = h->d_tiles;
  L.ncore = h->ncore;
  L.nbound = h->nbound;
  L.jit = h->jit_fn;
  L.jgrid = h->jit_grid > 0 ? h->jit_grid : h->ka64.ntile[0] * h->ka64.ntile[1] * h->ka64.ntile[2];
  L.jthr = h->jit_nthr;
  if (h->batch) {
    L.batch = h->batch_fn;
    L.insts = (const BInst*)h->d_insts;
    L.wl = (const int32_t*)h->d_wl;
    L.nwork = h->nwork;
    L.bgrid = h->batch_grid;
  }
  L.ntiles = h->ka64.ntile[0] * h->ka64.ntile[1] * h->ka64.ntile[2];
  const KArgs<T>& ka = kargs<T>(h);
  for (int a = 0; a < 3; ++a) L.jp.coord[a] = ka.coord[a];
  for (int a = 0; a < MOL_MAX_AUX; ++a) L.jp.aux[a] = ka.aux[a];
  L.jp.pool = ka.pool;
  for (int a = 0; a < 3; ++a) L.jp.wtab[a] = h->d_wtab[a];
  return L;
}

// Magic multiplier of fdiv_magic: ceil(2^32 / d), 0 for d == 1 (d < 2^16).
uint32_t magic32(uint32_t dv) { return dv <= 1 ? 0u : (uint32_t)(((1ull << 32) + dv - 1) / dv); }

// Node range along an axis where a table uses only its interior row (node-centred index i).
void interior_range(const mol_optab& tb, int n, int& lo, int& hi) {
  lo = std::max(lo, tb.nlow + 1);
  hi = std::min(hi, n + tb.ref_shift - tb.nhigh);
  if (tb.pn_hi > 0) {
    lo = std::max(lo, tb.pn_lo);
    hi = std::min(hi, tb.pn_hi);
  }
}

// The descriptor uses operators only the generated kernels implement: nonuniform WENO5 or
// per-node weight rows (nonuniform centered / upwind).
// Any closure joining two variables (interface boundaries run in generated kernels only).
bool has_interface(const mol_desc* d) {
  for (int c = 0; c < d->nclosures; ++c)
    if (d->closures_h[c].src >= 0) return true;
  return false;
}

bool generated_only(const mol_desc* d) {
  if (has_interface(d)) return true;
  for (int k = 0; k < d->nslots; ++k)
    if (d->slots_h[k].kind == MOL_SLOT_WENO_NU) return true;
  for (int t = 0; t < d->ntabs; ++t)
    if (d->tabs_h[t].pn_hi > 0) return true;
  return false;
}

// Build the core-tile kernel arguments and the list of tiles left to the generic kernel.
template <typename T>
void build_core(mol_handle* h, const mol_desc* d, const KArgs<T>& ka, KCore<T>& kc,
                std::vector<int32_t>& gen_tiles, int& ncore) {
  std::memset(&kc, 0, sizeof(kc));
  ncore = 0;
  const int ntot = ka.ntile[0] * ka.ntile[1] * ka.ntile[2];
  auto all_generic = [&](const char* why) {
    gen_tiles.resize(ntot);
    for (int i = 0; i < ntot; ++i) gen_tiles[i] = i;
    h->core_reason = why;
    h->core_interp_ok = false;
  };
  for (int f = 1; f < d->nfields; ++f)
    for (int a = 0; a < d->ndim; ++a)
      if (d->lo[f][a] != d->lo[0][a] || d->hi[f][a] != d->hi[0][a]) return all_generic("variables with different unknown boxes");
  // the interpreter's core kernel carries the IR in its argument block; the generated kernel
  // has no such limit, so the geometry below is computed either way
  bool fits = !(d->ncode > CORE_MAX_CODE || d->ncpool > CORE_MAX_CONST || d->nslots > CORE_MAX_SLOTS ||
                d->nprogs > CORE_MAX_PROGS);
  kc.ndim = d->ndim;
  kc.nfields = d->nfields;
  kc.naux = d->naux;
  kc.nslots = d->nslots;
  int64_t box = 1;
  for (int a = 0; a < 3; ++a) {
    kc.n[a] = ka.n[a];
    kc.periodic[a] = ka.periodic[a];
    kc.lo[a] = ka.lo[0][a];
    kc.hi[a] = ka.hi[0][a];
    kc.tile[a] = ka.tile[a];
    kc.halo[a] = ka.halo[a];
    kc.ext[a] = ka.ext[a];
    box *= kc.hi[a] - kc.lo[a] + 1;
  }
  kc.vol = ka.vol;
  kc.box = box;
  if (kc.vol >= (1 << 16)) return all_generic("tile too large for the index arithmetic");
  kc.magic0 = magic32((uint32_t)kc.ext[0]);
  kc.magic1 = magic32((uint32_t)(kc.ext[0] * kc.ext[1]));
  for (int f = 0; f < d->nfields; ++f) kc.rhs_prog[f] = d->rhs_prog[f];
  if (fits) {
    for (int i = 0; i < 2 * d->nprogs; ++i) kc.prog[i] = d->prog_h[i];
    for (int i = 0; i < d->ncode; ++i) kc.code[i] = d->code_h[i];
    for (int i = 0; i < d->ncpool; ++i) kc.cpool[i] = d->cpool_h[i];
  }
  for (int a = 0; a < 3; ++a) kc.coord[a] = ka.coord[a];
  for (int a = 0; a < MOL_MAX_AUX; ++a) kc.aux[a] = ka.aux[a];
  // core node range per axis: tile + halo inside the box, interior rows for every slot table
  int clo[3], chi[3];
  for (int a = 0; a < 3; ++a) {
    clo[a] = kc.lo[a];
    chi[a] = kc.hi[a];
    if (a < d->ndim && !kc.periodic[a]) {
      clo[a] = kc.lo[a] + kc.halo[a];
      chi[a] = kc.hi[a] - kc.halo[a];
    }
  }
  int np = 0;
  for (int k = 0; k < d->nslots; ++k) {
    const mol_slot& sl = d->slots_h[k];
    CoreSlot dummy;
    CoreSlot& cs = k < CORE_MAX_SLOTS ? kc.slots[k] : dummy;
    cs.kind = sl.kind; cs.field = sl.field; cs.axis = sl.axis; cs.axis2 = sl.axis2; cs.prog = sl.prog;
    cs.eps = sl.eps; cs.inv_dx = sl.inv_dx;
    int ntab = 0;
    switch (sl.kind) {
      case MOL_SLOT_CENTERED: ntab = 1; break;
      case MOL_SLOT_UPWIND: case MOL_SLOT_MIXED: ntab = 2; break;
      case MOL_SLOT_NONLINLAP: ntab = 3; break;
      case MOL_SLOT_SPHERICAL: ntab = 5; break;
      default: ntab = 0;
    }
    if (sl.kind == MOL_SLOT_WENO_NU) {
      // generated kernels only; one-sided targets at nodes 1, 2, n-1, n of a non-periodic axis
      fits = false;
      if (!kc.periodic[sl.axis]) {
        clo[sl.axis] = std::max(clo[sl.axis], 3);
        chi[sl.axis] = std::min(chi[sl.axis], kc.n[sl.axis] - 2);
      }
    }
    for (int q = 0; q < ntab; ++q) {
      const mol_optab& tb = d->tabs_h[sl.tab[q]];
      if (np + tb.len > CORE_MAX_POOL || tb.pn_hi > 0) fits = false;
      cs.len[q] = tb.len;
      cs.off[q] = tb.off;
      cs.w[q] = np;
      if (fits)
        for (int j = 0; j < tb.len; ++j) kc.pool[np++] = d->pool_h[tb.w + j];
      const int ax = (sl.kind == MOL_SLOT_MIXED && q == 1) ? sl.axis2 : sl.axis;
      if (kc.periodic[ax]) continue;
      const int n = kc.n[ax];
      const bool half_op = (sl.kind == MOL_SLOT_NONLINLAP && q < 3) || (sl.kind == MOL_SLOT_SPHERICAL && q >= 2);
      if (!half_op) {
        interior_range(tb, n, clo[ax], chi[ax]);
      } else {
        const int qo = sl.kind == MOL_SLOT_NONLINLAP ? 0 : 2;
        const mol_optab& to = d->tabs_h[sl.tab[qo]];
        if (q == qo) {
          // outer row at index i-1 with lref = n + ref_shift
          clo[ax] = std::max(clo[ax], to.nlow + 2);
          chi[ax] = std::min(chi[ax], n + to.ref_shift - to.nhigh + 1);
          if (to.pn_hi > 0) {
            clo[ax] = std::max(clo[ax], to.pn_lo + 1);
            chi[ax] = std::min(chi[ax], to.pn_hi + 1);
          }
        } else {
          // inner / interpolation rows at the outer taps h = i-1+off_o .. i-1+off_o+len_o-1
          clo[ax] = std::max(clo[ax], tb.nlow + 2 - to.off);
          chi[ax] = std::min(chi[ax], n - tb.nhigh + 2 - to.off - to.len);
          if (tb.pn_hi > 0) {
            clo[ax] = std::max(clo[ax], tb.pn_lo + 1 - to.off);
            chi[ax] = std::min(chi[ax], tb.pn_hi + 2 - to.off - to.len);
          }
        }
      }
    }
  }
  for (int a = 0; a < 3; ++a) {
    kc.nd_lo[a] = clo[a];
    kc.nd_hi[a] = chi[a];
  }
  int ct_lo[3], ct_hi[3];
  for (int a = 0; a < 3; ++a) {
    const int T_ = kc.tile[a];
    if (a >= d->ndim || kc.periodic[a]) {
      ct_lo[a] = 0;
      ct_hi[a] = ka.ntile[a] - 1;
      continue;
    }
    const int first = (clo[a] - kc.lo[a] + T_ - 1) / T_;
    const int last = (chi[a] - kc.lo[a] + 1) / T_ - 1;
    ct_lo[a] = std::max(0, first);
    ct_hi[a] = std::min(ka.ntile[a] - 1, last);
  }
  bool empty = false;
  for (int a = 0; a < 3; ++a) empty = empty || ct_hi[a] < ct_lo[a];
  gen_tiles.clear();
  if (empty) return all_generic("no tile away from the boundaries");
  for (int a = 0; a < 3; ++a) {
    kc.ct_lo[a] = ct_lo[a];
    kc.ct_n[a] = ct_hi[a] - ct_lo[a] + 1;
  }
  ncore = kc.ct_n[0] * kc.ct_n[1] * kc.ct_n[2];
  for (int tz = 0; tz < ka.ntile[2]; ++tz)
    for (int ty = 0; ty < ka.ntile[1]; ++ty)
      for (int tx = 0; tx < ka.ntile[0]; ++tx) {
        const bool core = tx >= ct_lo[0] && tx <= ct_hi[0] && ty >= ct_lo[1] && ty <= ct_hi[1] &&
                          tz >= ct_lo[2] && tz <= ct_hi[2];
        if (!core) gen_tiles.push_back(tx + ka.ntile[0] * (ty + ka.ntile[1] * tz));
      }
  h->core_reason = fits ? "" : "stencil IR larger than the interpreter's kernel-argument block";
  h->core_interp_ok = fits;
}

// Geometry handed to the code generator: the tiling of the union box and the core tiles.
template <typename T>
JitGeom geom_of(const KArgs<T>& ka, const KCore<T>& kc, int ncore) {
  JitGeom g{};
  g.nthr = 256;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = ka.n[a];
    g.periodic[a] = ka.periodic[a];
    g.ulo[a] = ka.ulo[a];
    g.uhi[a] = ka.uhi[a];
    g.tile[a] = ka.tile[a];
    g.halo[a] = ka.halo[a];
    g.ext[a] = ka.ext[a];
    g.ntile[a] = ka.ntile[a];
    g.ct_lo[a] = ncore > 0 ? kc.ct_lo[a] : 0;
    g.ct_n[a] = ncore > 0 ? kc.ct_n[a] : 0;
  }
  for (int f = 0; f < MOL_MAX_FIELDS; ++f) {
    for (int a = 0; a < 3; ++a) {
      g.lo[f][a] = ka.lo[f][a];
      g.hi[f][a] = ka.hi[f][a];
    }
    g.fofs[f] = ka.fofs[f];
  }
  g.vol = ka.vol;
  g.ncore = ncore;
  return g;
}

bool jit_enabled() {
  const char* e = std::getenv("MOL_JIT");
  return !(e && e[0] == '0');
}

int sync_kargs(mol_handle* h);

// Work-groups of a persistent launch over `ntiles` tiles: every one that fits on the device at
// once (MOL_BPC=k: k per compute unit), a multiple of 8 (tile k + grid stays on the XCD of tile
// k), at most one per tile.
int persistent_grid(hipFunction_t fn, int ntiles, int nthr) {
  int dev = 0, ncu = 0, occ = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
    return ntiles;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, nthr, 0) != hipSuccess || occ <= 0) occ = 1;
  if (const char* e = std::getenv("MOL_BPC")) occ = std::max(1, std::atoi(e));
  int64_t g = (int64_t)occ * ncu;
  if (g >= ntiles) return ntiles;
  g = std::max<int64_t>(8, g / 8 * 8);
  return (int)std::min<int64_t>(g, ntiles);
}

// Threads per work-group of a generated kernel: 256 for the right-hand side; the reverse-mode
// kernel runs one thread per tile node (256 to 1024).  MOL_NTHR / MOL_VJP_NTHR set them.
int kernel_threads(const JitGeom& g, bool vjp) {
  const char* e = std::getenv(vjp ? "MOL_VJP_NTHR" : "MOL_NTHR");
  int n = 256;
  if (vjp) n = g.tile[0] * g.tile[1] * g.tile[2];
  // 3D forward tiles: two nodes per thread (16x8x8: 512 threads, two work-groups per CU)
  else if (g.tile[2] > 1) n = std::max(256, g.tile[0] * g.tile[1] * g.tile[2] / 2);
  // 2D tiles of 2048 nodes or more under a stencil of reach two along both axes: 512 threads
  else if (g.tile[1] > 1 && g.halo[0] >= 2 && g.halo[1] >= 2 && g.tile[0] * g.tile[1] >= 2048) n = 512;
  if (e) n = std::atoi(e);
  n = (std::max(64, std::min(1024, n)) + 63) / 64 * 64;
  return vjp ? std::max(256, n) : n;
}

// Fused RK4 step kernel for periodic 1D/2D systems of reach 1 (MOL_FUSED=0 disables it): the
// tile + 4-node halo of u and two stage buffers in LDS, one launch per step.
bool fused_geometry(mol_handle* h, const mol_desc* d, const KArgs<double>& ka, const KCore<double>& kc, int ncore,
                    JitGeom& g, std::string& why) {
  if (d->ndim > 2) { why = "not a 1D/2D system"; return false; }
  if (has_interface(d)) { why = "interface boundaries between variables"; return false; }
  for (int a = 0; a < d->ndim; ++a)
    if (ka.halo[a] > 2 || (ka.halo[a] > 1 && !ka.periodic[a])) { why = "reach above 1 (2 on periodic axes)"; return false; }
  // a ghost's closure reads unknowns inward from the box edge: they must lie in the tile region
  int kmax = 0;
  for (int c = 0; c < d->nclosures; ++c) kmax = std::max(kmax, d->closures_h[c].K + d->closures_h[c].nghost);
  // the largest tile that leaves at least one tile per CU (fewer halo positions per node);
  // three stage buffers of the tile + halo, two work-groups per CU within 160 KB
  static const int cand2[][3] = {{32, 32, 1}, {32, 16, 1}, {32, 8, 1}, {64, 8, 1}, {16, 16, 1}, {16, 8, 1}};
  static const int cand1[][3] = {{256, 1, 1}, {128, 1, 1}, {64, 1, 1}};
  const int (*cand)[3] = d->ndim == 2 ? cand2 : cand1;
  const int nc = d->ndim == 2 ? 6 : 3;
  int best = -1;
  int64_t best_n = -1;
  for (int i = 0; i < nc; ++i) {
    int64_t vol = 1, tv = 1, nt = 1;
    for (int a = 0; a < 3; ++a) {
      const int ext = ka.uhi[a] - ka.ulo[a] + 1;
      const int tl = a < d->ndim ? std::max(1, std::min(cand[i][a], ext)) : 1;
      vol *= tl + 8 * (a < d->ndim ? ka.halo[a] : 0);
      tv *= tl;
      nt *= (ext + tl - 1) / tl;
    }
    (void)tv;
    const size_t lds = (size_t)3 * vol * d->nfields * h->esize();
    if (lds > 81920) continue;
    if (nt >= 256) { best = i; break; }
    if (nt > best_n) { best_n = nt; best = i; }
  }
  if (best < 0) { why = "tile does not fit LDS"; return false; }
  // a 1D box of up to 1024 unknowns is one tile: the kernel then runs many steps per launch
  // with the state resident in LDS
  static const int whole1[3] = {1024, 1, 1};
  const bool whole = d->ndim == 1 && ka.uhi[0] - ka.ulo[0] + 1 <= whole1[0];
  if (whole) cand = &whole1, best = 0;
  g = geom_of(ka, kc, ncore);
  g.vol = 1;
  for (int a = 0; a < 3; ++a) {
    const int ext = ka.uhi[a] - ka.ulo[a] + 1;
    g.tile[a] = a < d->ndim ? std::max(1, std::min(cand[best][a], ext)) : 1;
    if (h->tile_req[0] > 0) g.tile[a] = a < d->ndim ? std::max(1, std::min(h->tile_req[a], ext)) : 1;
    int ft[3];
    const char* fe = std::getenv("MOL_FUSED_TILE");
    if (fe && std::sscanf(fe, "%d,%d,%d", &ft[0], &ft[1], &ft[2]) == 3 && ft[0] > 0)
      g.tile[a] = a < d->ndim ? std::max(1, std::min(ft[a], ext)) : 1;
    g.ntile[a] = (ext + g.tile[a] - 1) / g.tile[a];
    g.ext[a] = g.tile[a] + 2 * g.halo[a];
    g.vol *= g.ext[a];
    if (a < d->ndim && !ka.periodic[a] && (g.tile[a] < kmax || ext < kmax)) {
      why = "tile narrower than the boundary closures";
      return false;
    }
  }
  // two threads per tile node (every stage region is larger than the tile), at most 1024
  g.nthr = std::max(256, std::min(1024, (2 * g.tile[0] * g.tile[1] * g.tile[2] + 63) / 64 * 64));
  if (const char* ne = std::getenv("MOL_FUSED_NTHR")) g.nthr = (std::max(64, std::min(1024, std::atoi(ne))) + 63) / 64 * 64;
  // Two steps per launch (an 8-reach halo, a fourth LDS array for the first step's running
  // sums) when every tile is resident at once: the launch overhead, the staging latency and the
  // store drain then come once per two steps, against stage regions about a quarter larger
  // (Brusselator 512^2, 32x32 tiles: 6168 against 4920 node evaluations per tile and step).
  // With more tiles than compute units, work-groups of the one-step kernel overlap each
  // other's phases instead.  MOL_FUSED_STEPS=1 or 2 overrides.
  g.fsteps = 1;
  if (d->ndim > 1 || g.ntile[0] > 1) {   // not one resident 1D tile
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess) ncu = 0;
    int64_t vol2 = 1, nt = 1;
    bool reach1 = true;
    for (int a = 0; a < 3; ++a) {
      vol2 *= g.tile[a] + 16 * (a < d->ndim ? ka.halo[a] : 0);
      nt *= g.ntile[a];
      if (a < d->ndim) reach1 = reach1 && ka.halo[a] <= 1;
    }
    const size_t lds2 = (size_t)4 * vol2 * d->nfields * h->esize();
    bool two = reach1 && lds2 <= 150 * 1024 && ncu > 0 && nt <= ncu;
    if (const char* se = std::getenv("MOL_FUSED_STEPS")) two = std::atoi(se) == 2 && reach1 && lds2 <= 150 * 1024;
    if (two) g.fsteps = 2;
  }
  return true;
}

// Coefficient rows of the nonuniform WENO5 derivative (coordinates only) at every node of each
// axis that has such a slot, computed once on the device by the generated mol_wenonu_table.
int build_wenonu_tables(mol_handle* h, const mol_desc* d, const std::string& src, std::string& err) {
  hipFunction_t fn = nullptr;
  if (jit_kernel(src, "mol_wenonu_table", h->device, &fn, err)) return MOL_ERR_HIP;
  constexpr int kNC = 56;   // coefficients per node (MOL_NU_NC of the generated code)
  for (int a = 0; a < d->ndim; ++a) {
    bool used = false;
    for (int k = 0; k < d->nslots; ++k) used = used || (d->slots_h[k].kind == MOL_SLOT_WENO_NU && d->slots_h[k].axis == a);
    if (!used) continue;
    const int n = d->n[a];
    if (!h->d_wtab[a] && hipMalloc((void**)&h->d_wtab[a], (size_t)n * kNC * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      err = "not enough device memory for the nonuniform WENO5 coefficient rows";
      return MOL_ERR_HIP;
    }
    const double* x = (const double*)h->d_coord[a];
    int nn = n, per = h->ka64.periodic[a] ? 1 : 0;
    double* out = h->d_wtab[a];
    void* args[] = {(void*)&x, (void*)&nn, (void*)&per, (void*)&out};
    if (hipModuleLaunchKernel(fn, (unsigned)((n + 255) / 256), 1, 1, 256, 1, 1, 0, nullptr, args, nullptr) != hipSuccess ||
        hipStreamSynchronize(nullptr) != hipSuccess) {
      err = std::string("nonuniform WENO5 table launch: ") + hipGetErrorString(hipGetLastError());
      return MOL_ERR_HIP;
    }
  }
  return 0;
}

void setup_fused(mol_handle* h, const mol_desc* d, const KArgs<double>& ka, const KCore<double>& kc, int ncore) {
  h->fused_fn = nullptr;
  h->fused_tiles = 0;
  h->fused_reason.clear();
  const char* e = std::getenv("MOL_FUSED");
  if (e && e[0] == '0') { h->fused_reason = "disabled (MOL_FUSED=0)"; return; }
  JitGeom g;
  if (!fused_geometry(h, d, ka, kc, ncore, g, h->fused_reason)) return;
  std::string err;
  const std::string src = jit_fused_source(d, g, err);
  if (src.empty()) { h->fused_reason = err; return; }
  if (jit_kernel(src, "mol_rk4_fused", h->device, &h->fused_fn, err)) {
    h->fused_fn = nullptr;
    h->fused_reason = err;
    return;
  }
  h->fused_tiles = g.ntile[0] * g.ntile[1] * g.ntile[2];
  h->fused_nthr = g.nthr;
  h->fused_resident = src.find("// resident steps per launch: any") != std::string::npos;
  h->fused_steps = src.find("// resident steps per launch: 2") != std::string::npos ? 2 : 1;
}

int setup_paths(mol_handle* h, const mol_desc* d) {
  std::vector<int32_t> t64, t32;
  int n64 = 0, n32 = 0;
  build_core<double>(h, d, h->ka64, h->kc64, t64, n64);
  build_core<float>(h, d, h->ka32, h->kc32, t32, n32);
  h->jit_fn = nullptr;
  h->jit_reason.clear();
  bool retiled = false;
  const void* aux_bound[MOL_MAX_AUX];
  for (int a = 0; a < MOL_MAX_AUX; ++a) aux_bound[a] = h->ka64.aux[a];
  auto refill = [&](size_t pos) {
    const bool jt = pos > 0;
    fill_kargs<double>(h, d, h->ka64, pos, true, jt);
    fill_kargs<float>(h, d, h->ka32, pos, true, jt);
    for (int a = 0; a < MOL_MAX_AUX; ++a) {
      h->ka64.aux[a] = (const double*)aux_bound[a];
      h->ka32.aux[a] = (const float*)aux_bound[a];
    }
    t64.clear();
    t32.clear();
    build_core<double>(h, d, h->ka64, h->kc64, t64, n64);
    build_core<float>(h, d, h->ka32, h->kc32, t32, n32);
  };
  if (jit_enabled()) {
    // tiles sized for the arrays the generated kernel stages (the interpreter stages them all)
    const size_t pos = jit_lds_per_pos(d, geom_of(h->ka64, h->kc64, n64), h->esize(), false);
    refill(pos > 0 ? pos : h->esize());
    retiled = true;
    std::string err;
    JitGeom jg = geom_of(h->ka64, h->kc64, n64);
    jg.nthr = kernel_threads(jg, false);
    h->jit_nthr = jg.nthr;
    const std::string src = jit_source(d, jg, err);
    if (src.empty()) h->jit_reason = err;
    else if (jit_kernel(src, "mol_rhs_tiles", h->device, &h->jit_fn, err)) {
      h->jit_fn = nullptr;
      h->jit_reason = err;
    }
    if (h->jit_fn && src.find("mol_wenonu_table(") != std::string::npos && build_wenonu_tables(h, d, src, err)) {
      h->jit_fn = nullptr;
      h->jit_reason = err;
    }
    const int ntl = h->ka64.ntile[0] * h->ka64.ntile[1] * h->ka64.ntile[2];
    if (h->jit_fn)
      h->jit_grid =
          src.find("mol_tiles_persistent<") != std::string::npos ? persistent_grid(h->jit_fn, ntl, jg.nthr) : ntl;
    if (!h->jit_fn && retiled) refill(0);   // back to the interpreter's tiles
    if (h->jit_fn) setup_fused(h, d, h->ka64, h->kc64, n64);
  } else {
    h->jit_reason = "code generation disabled (MOL_JIT=0)";
  }
  if (int rc = sync_kargs(h)) return rc;
  if (!h->jit_fn && generated_only(d))
    return fail(h, MOL_ERR_UNSUPPORTED,
                "nonuniform-grid operators run in generated kernels only (" + h->jit_reason + ")");
  if (n64 > 0 && !h->jit_fn && !h->core_interp_ok) {
    // neither core kernel can run this system: every tile goes to the generic kernel
    const KArgs<double>& ka = h->ka64;
    t64.clear();
    for (int i = 0; i < ka.ntile[0] * ka.ntile[1] * ka.ntile[2]; ++i) t64.push_back(i);
    n64 = 0;
  }
  h->ncore = n64;
  h->nbound = (int)t64.size();
  if (h->d_tiles) { (void)hipFree(h->d_tiles); h->d_tiles = nullptr; }
  if (!t64.empty()) {
    HIPCHK(h, hipMalloc((void**)&h->d_tiles, t64.size() * sizeof(int32_t)));
    HIPCHK(h, hipMemcpy(h->d_tiles, t64.data(), t64.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  return 0;
}

int ensure_workspace(mol_handle* h) {
  const size_t bytes = (size_t)h->nunk * h->esize();
  if (!h->w_y1) HIPCHK(h, hipMalloc(&h->w_y1, bytes));
  if (!h->w_y2) HIPCHK(h, hipMalloc(&h->w_y2, bytes));
  if (!h->w_acc) HIPCHK(h, hipMalloc(&h->w_acc, bytes));
  if (!h->w_comp) HIPCHK(h, hipMalloc(&h->w_comp, bytes));
  if (!h->d_tnow) HIPCHK(h, hipMalloc((void**)&h->d_tnow, 2 * sizeof(double)));
  return 0;
}

template <typename T>
KArgs<T>& kargs(mol_handle* h);
template <>
KArgs<double>& kargs<double>(mol_handle* h) { return h->ka64; }
template <>
KArgs<float>& kargs<float>(mol_handle* h) { return h->ka32; }
template <typename T>
KArgs<T>* dkargs(mol_handle* h);
template <>
KArgs<double>* dkargs<double>(mol_handle* h) { return h->d_ka64; }
template <>
KArgs<float>* dkargs<float>(mol_handle* h) { return h->d_ka32; }

// Upload the host kernel-argument blocks to their device copies.
int sync_kargs(mol_handle* h) {
  if (!h->d_ka64) HIPCHK(h, hipMalloc((void**)&h->d_ka64, sizeof(KArgs<double>)));
  if (!h->d_ka32) HIPCHK(h, hipMalloc((void**)&h->d_ka32, sizeof(KArgs<float>)));
  HIPCHK(h, hipMemcpy(h->d_ka64, &h->ka64, sizeof(KArgs<double>), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->d_ka32, &h->ka32, sizeof(KArgs<float>), hipMemcpyHostToDevice));
  return 0;
}

// Workspace of one integration: stage vectors (acc: RK4 stage launches; comp: Kahan) and the
// double-buffered device clock.
struct Work {
  void* y1;
  void* y2;
  void* acc;
  void* comp;
  double* tnow;
};

Work handle_work(const mol_handle* h) { return Work{h->w_y1, h->w_y2, h->w_acc, h->w_comp, h->d_tnow}; }

// t0v: the start time, passed by value to the first fused launch instead of read from the
// device clock (which that launch then sets for the next); null: the clock holds it.
template <typename T>
int enqueue_steps(mol_handle* h, const Work& w, T* u, const T* p, double dt, int64_t nsteps, int scheme,
                  bool kahan, hipStream_t s, const double* t0v = nullptr) {
  const Launch<T> L = launch_of<T>(h);
  T* y1 = (T*)w.y1;
  T* y2 = (T*)w.y2;
  T* acc = (T*)w.acc;
  T* comp = (T*)w.comp;
  if (scheme == MOL_RK4 && !kahan && h->fused_fn && !h->no_fuse) {
    // one launch per step (or per chunk of resident steps), the state alternating between u
    // and the workspace
    T* src = u;
    T* dst = y1;
    for (int64_t step = 0; step < nsteps;) {
      const int nsub = h->fused_resident ? (int)std::min<int64_t>(nsteps - step, 1024)
                                         : (h->fused_steps == 2 && nsteps - step >= 2 ? 2 : 1);
      Epi<T> e{};
      e.tnow = w.tnow + h->clk;
      e.tnext = w.tnow + (1 - h->clk);
      e.tadv = dt;
      e.p = p;
      if (t0v && step == 0 && !h->fused_resident) {
        e.tnow = nullptr;
        e.tofs = *t0v;
      }
      JitPtrs<T> jp = L.jp;
      const T* sp = src;
      int ns = nsub;
      void* args[] = {(void*)&sp, (void*)&dst, (void*)&e, (void*)&jp, (void*)&ns};
      if (hipModuleLaunchKernel(h->fused_fn, h->fused_tiles, 1, 1, h->fused_nthr, 1, 1, 0, s, args, nullptr) !=
          hipSuccess)
        return fail(h, MOL_ERR_HIP, "fused RK4 launch failed");
      h->clk ^= 1;
      std::swap(src, dst);
      step += nsub;
    }
    if (src != u) HIPCHK(h, hipMemcpyAsync(u, src, (size_t)h->nunk * sizeof(T), hipMemcpyDeviceToDevice, s));
    return 0;
  }
  for (int64_t step = 0; step < nsteps; ++step) {
    double* tcur = w.tnow + h->clk;
    double* tnxt = w.tnow + (1 - h->clk);
    Epi<T> e{};
    e.tnow = tcur;
    e.p = p;
    int rc = 0;
    if (scheme == MOL_RK4) {
      e.mode = 1; e.out = y1; e.A = u; e.ca = 1.0; e.cf = dt / 2; e.acc = acc; e.acc_keep = 0.0; e.acc_f = 1.0; e.tofs = 0.0;
      rc |= launch_rhs<T>(L, u, e, s);
      e.out = y2; e.cf = dt / 2; e.acc_keep = 1.0; e.acc_f = 2.0; e.tofs = dt / 2;
      rc |= launch_rhs<T>(L, y1, e, s);
      e.out = y1; e.cf = dt; e.tofs = dt / 2;
      rc |= launch_rhs<T>(L, y2, e, s);
      Epi<T> l{};
      l.tnow = tcur; l.p = p; l.tofs = dt; l.out = u; l.A = u; l.B = acc; l.cb = dt / 6; l.cf = dt / 6;
      l.tnext = tnxt; l.tadv = dt;
      if (kahan) { l.mode = 2; l.subA = 0; l.comp = comp; }
      else { l.mode = 1; l.ca = 1.0; }
      rc |= launch_rhs<T>(L, y1, l, s);
    } else {
      e.mode = 1; e.out = y1; e.A = u; e.ca = 1.0; e.cf = dt; e.tofs = 0.0;
      rc |= launch_rhs<T>(L, u, e, s);
      e.out = y2; e.ca = 0.75; e.B = y1; e.cb = 0.25; e.cf = dt / 4; e.tofs = dt;
      rc |= launch_rhs<T>(L, y1, e, s);
      Epi<T> l{};
      l.tnow = tcur; l.p = p; l.tofs = dt / 2; l.out = u; l.A = u; l.B = y2;
      l.tnext = tnxt; l.tadv = dt;
      l.cb = 2.0 / 3.0; l.cf = 2.0 * dt / 3.0;
      if (kahan) { l.mode = 2; l.subA = 1; l.comp = comp; }
      else { l.mode = 1; l.ca = 1.0 / 3.0; }
      rc |= launch_rhs<T>(L, y2, l, s);
    }
    h->clk ^= 1;
    if (rc) return fail(h, MOL_ERR_HIP, std::string("kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
  }
  return 0;
}

// Step counts of the captured chunks, largest first.  Each is even: a chunk flips the device
// clock and the fused step's state buffers an even number of times, so it starts and ends at
// clock slot 0 with the state in u, and any sequence of chunks can be replayed.
constexpr int64_t kGraphChunks[3] = {64, 16, 4};

template <typename T>
int capture_chunk(mol_handle* h, T* u, const T* p, double dt, int64_t steps, int scheme, bool kahan,
                  mol_handle::GraphChunk& c) {
  if (!h->cap_stream) HIPCHK(h, hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
  const int clk0 = h->clk;
  hipGraph_t g = nullptr;
  HIPCHK(h, hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_steps<T>(h, handle_work(h), u, p, dt, steps, scheme, kahan, h->cap_stream);
  const hipError_t ee = hipStreamEndCapture(h->cap_stream, &g);
  h->clk = clk0;
  if (rc || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : fail(h, MOL_ERR_HIP, "graph capture failed");
  }
  hipGraphExec_t ge = nullptr;
  if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(h, MOL_ERR_HIP, "graph instantiation failed");
  }
  c = mol_handle::GraphChunk{u, p, dt, scheme, kahan ? 1 : 0, h->no_fuse ? 1 : 0, steps, g, ge};
  return 0;
}

template <typename T>
int integrate_impl(mol_handle* h, T* u, const T* p, double t0, double dt, int64_t nsteps,
                   int scheme, int flags, hipStream_t s) {
  int rc = ensure_workspace(h);
  if (rc) return rc;
  const bool kahan = (flags & MOL_FLAG_KAHAN) != 0;
  h->no_fuse = (flags & MOL_FLAG_UNFUSED) != 0;
  h->clk = 0;
  // a resident fused kernel already runs up to 1024 steps per launch and flips the clock once
  // per launch, whatever the step count: it is launched directly
  const bool fused = scheme == MOL_RK4 && !kahan && h->fused_fn && !h->no_fuse;
  const bool resident = fused && h->fused_resident;
  const bool direct = !(flags & MOL_FLAG_GRAPH) || resident || nsteps < kGraphChunks[2];
  // direct fused steps take the start time by value in their first launch: no clock launch
  if (direct && fused && !resident && nsteps > 0)
    return enqueue_steps<T>(h, handle_work(h), u, p, dt, nsteps, scheme, kahan, s, &t0);
  if (launch_set_clock(h->d_tnow, t0, s)) return fail(h, MOL_ERR_HIP, "clock launch failed");
  if (kahan && launch_fill_zero<T>((T*)h->w_comp, h->nunk, s)) return fail(h, MOL_ERR_HIP, "zero fill launch failed");
  if (direct) return enqueue_steps<T>(h, handle_work(h), u, p, dt, nsteps, scheme, kahan, s);
  hipGraphExec_t ex[3] = {nullptr, nullptr, nullptr};
  for (const auto& c : h->graphs)
    if (c.u == u && c.p == p && c.dt == dt && c.scheme == scheme && c.kahan == (kahan ? 1 : 0) &&
        c.nofuse == (h->no_fuse ? 1 : 0))
      for (int k = 0; k < 3; ++k)
        if (c.steps == kGraphChunks[k]) ex[k] = c.exec;
  if (!ex[0] || !ex[1] || !ex[2]) {
    if (h->graphs.size() >= 24) drop_graphs(h);
    for (int k = 0; k < 3; ++k) {
      if (ex[k]) continue;
      mol_handle::GraphChunk c{};
      if ((rc = capture_chunk<T>(h, u, p, dt, kGraphChunks[k], scheme, kahan, c))) return rc;
      h->graphs.push_back(c);
      ex[k] = c.exec;
    }
  }
  int64_t left = nsteps;
  for (int k = 0; k < 3; ++k)
    for (; left >= kGraphChunks[k]; left -= kGraphChunks[k]) HIPCHK(h, hipGraphLaunch(ex[k], s));
  return enqueue_steps<T>(h, handle_work(h), u, p, dt, left, scheme, kahan, s);
}

// mol_integrate without any allocation: the stage vectors and the device clock live in this
// device's `scratch`; no graphs (instantiation allocates), no memset nodes (a caller may
// capture the launches).  The arrays a path reads before writing (the Kahan compensation) are
// cleared by a fill kernel, since the scratch keeps the last call's contents.
template <typename T>
int integrate_noalloc_impl(mol_handle* h, T* u, const T* p, double t0, double dt, int64_t nsteps, int scheme,
                           int flags, hipStream_t s) {
  std::string err;
  char* base = (char*)scratch_of(h->device, err);
  if (!base) return fail(h, MOL_ERR_HIP, err);
  const bool kahan = (flags & MOL_FLAG_KAHAN) != 0;
  h->no_fuse = (flags & MOL_FLAG_UNFUSED) != 0;
  const bool fused = scheme == MOL_RK4 && !kahan && h->fused_fn && !h->no_fuse;
  // stage vectors each path touches: fused RK4 y1; RK4 stage launches y1, y2, acc; SSPRK33 y1,
  // y2; Kahan adds the compensation
  const size_t vb = ((size_t)h->nunk * sizeof(T) + 255) / 256 * 256;
  const int narr = fused ? 1 : (scheme == MOL_RK4 ? 3 : 2) + (kahan ? 1 : 0);
  const size_t need = 256 + (size_t)narr * vb;
  if (need > (size_t)MOL_NOALLOC_SCRATCH_BYTES)
    return fail(h, MOL_ERR_SHAPE, "mol_integrate_noalloc: this integration needs " + std::to_string(need) +
                                      " bytes of scratch, more than its " +
                                      std::to_string((size_t)MOL_NOALLOC_SCRATCH_BYTES) + " bytes");
  Work w{};
  w.tnow = (double*)base;
  char* q = base + 256;
  w.y1 = q;
  if (!fused) {
    w.y2 = q + vb;
    if (scheme == MOL_RK4) w.acc = q + 2 * vb;
    if (kahan) w.comp = q + (size_t)(narr - 1) * vb;
  }
  h->clk = 0;
  if (launch_set_clock(w.tnow, t0, s)) return fail(h, MOL_ERR_HIP, "clock launch failed");
  if (kahan && launch_fill_zero<T>((T*)w.comp, h->nunk, s)) return fail(h, MOL_ERR_HIP, "zero fill launch failed");
  return enqueue_steps<T>(h, w, u, p, dt, nsteps, scheme, kahan, s);
}

// Does program `prog` read opcode `op` (e.g. an auxiliary array)?
bool prog_uses(const mol_desc* d, int prog, uint32_t op) {
  if (prog < 0) return false;
  const int st = d->prog_h[2 * prog], ln = d->prog_h[2 * prog + 1];
  for (int pc = st; pc < st + ln; ++pc)
    if ((d->code_h[pc] & 0xffu) == op) return true;
  return false;
}

// Generate and load the vector-Jacobian product kernel (first use).  Its tiles are sized for
// the state, the auxiliary arrays and their adjoint buffers in LDS.
// Tiling of the reverse-mode kernel: LDS holds the state, the grid fields and their adjoint
// buffers (FP64 for every array).
JitGeom vjp_geometry(mol_handle* h, const mol_desc* d, int* ntiles) {
  KArgs<double> kv;
  // staged arrays of the reverse-mode kernel, estimated first with every array staged
  fill_kargs<double>(h, d, kv, ((size_t)d->nfields + d->naux) * (h->esize() + sizeof(double)));
  {
    KCore<double> kc0;
    std::vector<int32_t> gen0;
    int nc0 = 0;
    const std::string reason = h->core_reason;
    const bool interp_ok = h->core_interp_ok;
    build_core<double>(h, d, kv, kc0, gen0, nc0);
    h->core_reason = reason;
    h->core_interp_ok = interp_ok;
    const size_t pos = jit_lds_per_pos(d, geom_of(kv, kc0, nc0), h->esize(), true);
    // MOL_VJP_TILE=tx,ty,tz asks for a tile of the reverse-mode kernel (LDS up to 160 KB)
    int req[3] = {0, 0, 0};
    const char* e = std::getenv("MOL_VJP_TILE");
    if (e && std::sscanf(e, "%d,%d,%d", &req[0], &req[1], &req[2]) == 3 && req[0] > 0) {
      int keep[3];
      for (int a = 0; a < 3; ++a) {
        keep[a] = h->tile_req[a];
        h->tile_req[a] = std::max(1, req[a]);
      }
      fill_kargs<double>(h, d, kv, pos > 0 ? pos : h->esize(), true, 0, 163000);
      for (int a = 0; a < 3; ++a) h->tile_req[a] = keep[a];
    } else if (pos > 0) {
      fill_kargs<double>(h, d, kv, pos, true, 2, 163000);
    }
  }
  KCore<double> kc;
  std::vector<int32_t> gen;
  int ncore = 0;
  const std::string reason = h->core_reason;
  const bool interp_ok = h->core_interp_ok;
  build_core<double>(h, d, kv, kc, gen, ncore);
  h->core_reason = reason;
  h->core_interp_ok = interp_ok;
  if (ntiles) *ntiles = kv.ntile[0] * kv.ntile[1] * kv.ntile[2];
  JitGeom g = geom_of(kv, kc, ncore);
  g.nthr = kernel_threads(g, true);
  return g;
}

int ensure_vjp(mol_handle* h) {
  if (h->vjp_fn) return 0;
  const mol_desc* d = &h->desc.d;
  int ntiles = 0;
  const JitGeom g = vjp_geometry(h, d, &ntiles);
  std::string err;
  const std::string src = jit_vjp_source(d, g, err);
  if (src.empty()) return fail(h, MOL_ERR_UNSUPPORTED, "vector-Jacobian product: " + err);
  h->vjp_store = src.find("// adjoint gather stores every unknown") != std::string::npos;
  if (jit_kernel(src, "mol_vjp_tiles", h->device, &h->vjp_fn, err) ||
      jit_kernel(src, "mol_vjp_gather", h->device, &h->vjp_gather_fn, err)) {
    h->vjp_fn = nullptr;
    return fail(h, MOL_ERR_HIP, "vector-Jacobian product: " + err);
  }
  // per-tile buffers: VOLX positions per staged array and tile (numbers of the generated code)
  auto num_after = [&](const std::string& key, size_t from) -> long long {
    const size_t k = src.find(key, from);
    return k == std::string::npos ? -1 : std::atoll(src.c_str() + k + key.size());
  };
  const size_t gb = src.find("struct GeoB");
  const long long volx = num_after("constexpr int VOLX = ", 0);
  const long long nstf = num_after("static constexpr int NSTF = ", gb);
  const long long nsta = num_after("static constexpr int NSTA = ", gb);
  if (gb == std::string::npos || volx <= 0 || nstf < 0 || nsta < 0) {
    h->vjp_fn = nullptr;
    return fail(h, MOL_ERR_UNSUPPORTED, "vector-Jacobian product: buffer layout not found");
  }
  if (h->vjp_bf) (void)hipFree(h->vjp_bf);
  if (h->vjp_ba) (void)hipFree(h->vjp_ba);
  h->vjp_bf = h->vjp_ba = nullptr;
  if (nstf > 0 && hipMalloc(&h->vjp_bf, (size_t)ntiles * volx * nstf * h->esize()) != hipSuccess) {
    h->vjp_fn = nullptr;
    return fail(h, MOL_ERR_HIP, "vector-Jacobian product: tile buffer allocation failed");
  }
  h->vjp_ba_count = (size_t)ntiles * volx * nsta;
  if (nsta > 0 && hipMalloc(&h->vjp_ba, (size_t)ntiles * volx * nsta * sizeof(double)) != hipSuccess) {
    h->vjp_fn = nullptr;
    return fail(h, MOL_ERR_HIP, "vector-Jacobian product: tile buffer allocation failed");
  }
  h->vjp_tiles = ntiles;
  h->vjp_nthr = g.nthr;
  h->vjp_aux_ok = true;
  for (int c = 0; c < d->nclosures; ++c)
    if (prog_uses(d, d->closures_h[c].gprog, MOL_OP_AUX)) h->vjp_aux_ok = false;
  return 0;
}

// Tile kernel (when `tiles`) then the gather kernel for the parts va.mode selects.
template <typename T>
int launch_vjp(mol_handle* h, const VArgs<T>& va, hipStream_t s, bool tiles = true) {
  const KArgs<T>& ka = kargs<T>(h);
  JitPtrs<T> jp{};
  for (int a = 0; a < 3; ++a) jp.coord[a] = ka.coord[a];
  for (int a = 0; a < MOL_MAX_AUX; ++a) jp.aux[a] = ka.aux[a];
  jp.pool = ka.pool;
  for (int a = 0; a < 3; ++a) jp.wtab[a] = h->d_wtab[a];
  VArgs<T> v = va;
  v.bf = (T*)h->vjp_bf;
  v.ba = (double*)h->vjp_ba;
  void* args[] = {(void*)&v, (void*)&jp};
  hipError_t e = hipSuccess;
  if (tiles) {
    e = hipModuleLaunchKernel(h->vjp_fn, h->vjp_tiles, 1, 1, h->vjp_nthr, 1, 1, 0, s, args, nullptr);
    if (e != hipSuccess) return fail(h, MOL_ERR_HIP, std::string("vector-Jacobian product launch: ") + hipGetErrorString(e));
  }
  if (!(v.mode & (VJP_GATHER_F | VJP_GATHER_A))) return 0;
  // gather: every unknown, then every grid position of every grid field (when g_aux is given)
  const int64_t items = h->nunk + (v.g_aux ? (int64_t)h->desc.d.naux * ka.n[0] * ka.n[1] * ka.n[2] : 0);
  void* gargs[] = {(void*)&v};
  e = hipModuleLaunchKernel(h->vjp_gather_fn, (unsigned)((items + 255) / 256), 1, 1, 256, 1, 1, 0, s, gargs, nullptr);
  if (e != hipSuccess) return fail(h, MOL_ERR_HIP, std::string("vector-Jacobian gather launch: ") + hipGetErrorString(e));
  return 0;
}

// One RK4 step from uin that keeps the stage states y1, y2, y3 and writes u_{n+1} to uout;
// the same arithmetic as the stage loop of enqueue_steps (time from the device clock).
template <typename T>
int rk4_step_store(mol_handle* h, const T* p, const T* uin, T* y1, T* y2, T* y3, T* uout, double dt,
                   hipStream_t s) {
  const Launch<T> L = launch_of<T>(h);
  T* acc = (T*)h->w_acc;
  double* tcur = h->d_tnow + h->clk;
  Epi<T> e{};
  e.tnow = tcur;
  e.p = p;
  int rc = 0;
  e.mode = 1; e.out = y1; e.A = uin; e.ca = 1.0; e.cf = dt / 2; e.acc = acc; e.acc_keep = 0.0; e.acc_f = 1.0; e.tofs = 0.0;
  rc |= launch_rhs<T>(L, uin, e, s);
  e.out = y2; e.cf = dt / 2; e.acc_keep = 1.0; e.acc_f = 2.0; e.tofs = dt / 2;
  rc |= launch_rhs<T>(L, y1, e, s);
  e.out = y3; e.cf = dt; e.tofs = dt / 2;
  rc |= launch_rhs<T>(L, y2, e, s);
  Epi<T> l{};
  l.tnow = tcur; l.p = p; l.tofs = dt; l.mode = 1; l.out = uout; l.A = uin; l.ca = 1.0; l.B = acc;
  l.cb = dt / 6; l.cf = dt / 6; l.tnext = h->d_tnow + (1 - h->clk); l.tadv = dt;
  rc |= launch_rhs<T>(L, y3, l, s);
  h->clk ^= 1;
  if (rc) return fail(h, MOL_ERR_HIP, std::string("kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
  return 0;
}

template <typename T>
int forward_ckpt_impl(mol_handle* h, T* u, const T* p, double t0, double dt, int64_t nsteps, int32_t every,
                      hipStream_t s) {
  int rc = ensure_workspace(h);
  if (rc) return rc;
  const size_t vbytes = (size_t)h->nunk * h->esize();
  h->ck_valid = false;
  if (every == 0) {
    // keep u and the three stage states of every step: the adjoint then runs no forward
    // recomputation (4 * nsteps state vectors of device memory)
    const size_t need = (size_t)(4 * nsteps + 1);
    if (h->seg_cap < need) {
      if (h->seg_u) (void)hipFree(h->seg_u);
      if (h->seg_t) (void)hipFree(h->seg_t);
      h->seg_u = nullptr;
      h->seg_t = nullptr;
      h->seg_cap = 0;
      if (hipMalloc(&h->seg_u, need * vbytes) != hipSuccess) {
        (void)hipGetLastError();
        h->seg_u = nullptr;
        return fail(h, MOL_ERR_HIP, "not enough device memory to keep every stage state (use ckpt_every >= 1)");
      }
      HIPCHK(h, hipMalloc((void**)&h->seg_t, (size_t)(nsteps + 1) * sizeof(double)));
      h->seg_cap = need;
    }
    const int64_t n = h->nunk;
    T* S = (T*)h->seg_u;
    auto st = [&](int64_t j, int k) { return S + (size_t)(4 * j + k) * n; };
    if (launch_set_clock(h->d_tnow + h->clk, t0, s)) return fail(h, MOL_ERR_HIP, "clock launch failed");
    HIPCHK(h, hipMemcpyAsync(st(0, 0), u, vbytes, hipMemcpyDeviceToDevice, s));
    for (int64_t j = 0; j < nsteps; ++j) {
      HIPCHK(h, hipMemcpyAsync(h->seg_t + j, h->d_tnow + h->clk, sizeof(double), hipMemcpyDeviceToDevice, s));
      rc = rk4_step_store<T>(h, p, st(j, 0), st(j, 1), st(j, 2), st(j, 3), j + 1 < nsteps ? st(j + 1, 0) : u, dt, s);
      if (rc) return rc;
    }
    h->ck_nsteps = nsteps;
    h->ck_every = (int32_t)std::min<int64_t>(nsteps, INT32_MAX);
    h->ck_dt = dt;
    h->ck_stored = true;
    h->ck_valid = true;
    return 0;
  }
  const int64_t nseg = (nsteps + every - 1) / every;
  if (h->ck_cap < (size_t)nseg) {
    if (h->ck_u) (void)hipFree(h->ck_u);
    if (h->ck_t) (void)hipFree(h->ck_t);
    h->ck_u = nullptr;
    h->ck_t = nullptr;
    h->ck_cap = 0;
    HIPCHK(h, hipMalloc(&h->ck_u, (size_t)nseg * vbytes));
    HIPCHK(h, hipMalloc((void**)&h->ck_t, (size_t)nseg * sizeof(double)));
    h->ck_cap = (size_t)nseg;
  }
  h->ck_valid = false;
  if (launch_set_clock(h->d_tnow + h->clk, t0, s)) return fail(h, MOL_ERR_HIP, "clock launch failed");
  for (int64_t g = 0; g < nseg; ++g) {
    const int64_t len = std::min<int64_t>(every, nsteps - g * every);
    HIPCHK(h, hipMemcpyAsync((char*)h->ck_u + g * vbytes, u, vbytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(h, hipMemcpyAsync(h->ck_t + g, h->d_tnow + h->clk, sizeof(double), hipMemcpyDeviceToDevice, s));
    rc = enqueue_steps<T>(h, handle_work(h), u, p, dt, len, MOL_RK4, false, s);
    if (rc) return rc;
  }
  h->ck_nsteps = nsteps;
  h->ck_every = every;
  h->ck_dt = dt;
  h->ck_stored = false;
  h->ck_valid = true;
  return 0;
}

template <typename T>
int adjoint_impl(mol_handle* h, const T* p, T* lam, double* g_aux, hipStream_t s) {
  int rc = ensure_vjp(h);
  if (rc) return rc;
  if (g_aux && !h->vjp_aux_ok)
    return fail(h, MOL_ERR_UNSUPPORTED, "auxiliary-array gradient through boundary values that read auxiliary arrays");
  rc = ensure_workspace(h);
  if (rc) return rc;
  const int64_t every = h->ck_every, nsteps = h->ck_nsteps;
  const int64_t nseg = (nsteps + every - 1) / every;
  const double dt = h->ck_dt;
  const int64_t n = h->nunk;
  const size_t vbytes = (size_t)n * h->esize();
  const size_t need = (size_t)(4 * every + 1);
  if (h->ck_stored && h->seg_cap < need) return fail(h, MOL_ERR_INVALID, "stored stage states were released");
  if (h->seg_cap < need) {
    if (h->seg_u) (void)hipFree(h->seg_u);
    if (h->seg_t) (void)hipFree(h->seg_t);
    h->seg_u = nullptr;
    h->seg_t = nullptr;
    h->seg_cap = 0;
    HIPCHK(h, hipMalloc(&h->seg_u, need * vbytes));
    HIPCHK(h, hipMalloc((void**)&h->seg_t, (size_t)(every + 1) * sizeof(double)));
    h->seg_cap = need;
  }
  for (int k = 0; k < 4; ++k)
    if (!h->adj_y[k]) HIPCHK(h, hipMalloc(&h->adj_y[k], vbytes));
  T* S = (T*)h->seg_u;
  auto st = [&](int64_t j, int k) { return S + (size_t)(4 * j + k) * n; };   // state k of step j
  T* Y[4] = {(T*)h->adj_y[0], (T*)h->adj_y[1], (T*)h->adj_y[2], (T*)h->adj_y[3]};
  // the stage adjoints accumulate unless the gather stores them: cleared here, then by every
  // step's sum kernel; the tile buffers of the grid-field adjoints accumulate over the whole
  // sweep and are gathered once
  for (int k = 0; k < 4 && !h->vjp_store; ++k)
    if ((rc = launch_fill_zero<T>(Y[k], n, s))) return fail(h, rc, "zero fill launch failed");
  const bool acc_aux = g_aux && h->vjp_ba && h->vjp_ba_count > 0;
  if (acc_aux && (rc = launch_fill_zero<double>((double*)h->vjp_ba, (int64_t)h->vjp_ba_count, s)))
    return fail(h, rc, "zero fill launch failed");
  for (int64_t g = nseg - 1; g >= 0; --g) {
    const int64_t len = std::min<int64_t>(every, nsteps - g * every);
    // recompute the segment from its checkpoint, keeping every stage state and step time
    // (a run that kept them all is one segment with nothing to recompute)
    if (!h->ck_stored) {
      HIPCHK(h, hipMemcpyAsync(st(0, 0), (char*)h->ck_u + g * vbytes, vbytes, hipMemcpyDeviceToDevice, s));
      HIPCHK(h, hipMemcpyAsync(h->d_tnow + h->clk, h->ck_t + g, sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    for (int64_t j = 0; !h->ck_stored && j < len; ++j) {
      HIPCHK(h, hipMemcpyAsync(h->seg_t + j, h->d_tnow + h->clk, sizeof(double), hipMemcpyDeviceToDevice, s));
      // the state after the segment is not needed: it goes to the integrator's scratch
      rc = rk4_step_store<T>(h, p, st(j, 0), st(j, 1), st(j, 2), st(j, 3), j + 1 < len ? st(j + 1, 0) : (T*)h->w_y2,
                             dt, s);
      if (rc) return rc;
    }
    // backward sweep: k4 = f(y3, t+dt), k3 = f(y2, t+dt/2), k2 = f(y1, t+dt/2), k1 = f(u, t)
    for (int64_t j = len - 1; j >= 0; --j) {
      const int stage_state[4] = {3, 2, 1, 0};
      const double tofs[4] = {dt, dt / 2, dt / 2, 0.0};
      const double c1[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
      const double c2[4] = {0.0, dt, dt / 2, dt / 2};
      for (int q = 0; q < 4; ++q) {
        VArgs<T> va{};
        va.u = st(j, stage_state[q]);
        va.s1 = lam;
        va.c1 = c1[q];
        va.s2 = q > 0 ? Y[q - 1] : nullptr;
        va.c2 = c2[q];
        va.out = Y[q];
        va.g_aux = g_aux;
        va.p = p;
        va.tnow = h->seg_t + j;
        va.tofs = tofs[q];
        va.mode = acc_aux ? (VJP_GATHER_F | VJP_ACC_A) : (VJP_GATHER_F | VJP_GATHER_A);
        if (q == 3 && h->vjp_store) {
          // the last gather of the step stores lam + Y0 + Y1 + Y2 + its own sum into lam
          va.mode |= VJP_SUM;
          va.lam = lam;
          va.y0 = Y[0];
          va.y1 = Y[1];
          va.y2 = Y[2];
        }
        if ((rc = launch_vjp<T>(h, va, s))) return rc;
      }
      if (!h->vjp_store && (rc = launch_adjoint_sum<T>(lam, Y[0], Y[1], Y[2], Y[3], n, true, s)))
        return fail(h, rc, "adjoint update launch failed");
    }
  }
  if (acc_aux) {
    VArgs<T> va{};
    va.g_aux = g_aux;
    va.mode = VJP_GATHER_A;
    if ((rc = launch_vjp<T>(h, va, s, false))) return rc;
  }
  return 0;
}

// Tiles of one instance whose tile + halo lies inside every variable's box (plain staging).
void simple_range(const mol_desc* d, const KArgs<double>& ka, int a, int& lo_t, int& hi_t) {
  if (a >= d->ndim || ka.periodic[a]) {
    lo_t = 0;
    hi_t = ka.ntile[a] - 1;
    return;
  }
  int lo = -(1 << 30), hi = 1 << 30;
  for (int f = 0; f < d->nfields; ++f) {
    lo = std::max(lo, ka.lo[f][a]);
    hi = std::min(hi, ka.hi[f][a]);
  }
  lo_t = ka.ntile[a];
  hi_t = -1;
  for (int t = 0; t < ka.ntile[a]; ++t) {
    const int o0 = ka.ulo[a] + t * ka.tile[a];
    if (o0 - ka.halo[a] >= lo && o0 + ka.tile[a] - 1 + ka.halo[a] <= hi) {
      lo_t = std::min(lo_t, t);
      hi_t = std::max(hi_t, t);
    }
  }
}

// Host side of a batch: instance headers (device pointers left as offsets into `data`), the
// work list, the weights/constants/coordinates of every instance and the generated source.
struct BatchPlan {
  std::vector<BInst> hdr;
  std::vector<std::array<int64_t, 5>> dofs;   // offsets of W, K, coord0..2 in data
  std::vector<double> data;
  std::vector<int32_t> wl;
  std::vector<int64_t> inst_off;
  int64_t nunk = 0;
  int nclasses = 0;
  int tile[3] = {1, 1, 1};
  std::string src;
};

int plan_batched(const mol_desc* const* ds, const double* const* params, int32_t ninst, const int32_t* tile,
                 BatchPlan& P) {
  std::string msg;
  for (int i = 0; i < ninst; ++i) {
    if (!ds[i]) return fail(nullptr, MOL_ERR_INVALID, "descriptor " + std::to_string(i) + " is NULL");
    if (int rc = validate(ds[i], msg)) return fail(nullptr, rc, "instance " + std::to_string(i) + ": " + msg);
    if (ds[i]->dtype != ds[0]->dtype || ds[i]->ndim != ds[0]->ndim || ds[i]->nfields != ds[0]->nfields)
      return fail(nullptr, MOL_ERR_INVALID, "instances differ in dtype, dimension or number of variables");
    if (ds[i]->naux != 0) return fail(nullptr, MOL_ERR_UNSUPPORTED, "batched instances take no grid fields");
    if (generated_only(ds[i]))
      return fail(nullptr, MOL_ERR_UNSUPPORTED, "batched instances take no nonuniform-grid operators or interfaces");
    if (ds[i]->nparams > 0 && (!params || !params[i]))
      return fail(nullptr, MOL_ERR_INVALID, "parameter values of instance " + std::to_string(i) + " missing");
  }
  const size_t esize = ds[0]->dtype == MOL_F64 ? sizeof(double) : sizeof(float);
  std::map<std::string, int> cls_of;
  std::vector<std::string> cls_src;
  P.hdr.assign(ninst, BInst{});
  P.dofs.assign(ninst, {});
  int ldsn = 1;
  for (int i = 0; i < ninst; ++i) {
    mol_handle th;
    th.desc.assign(ds[i]);
    const mol_desc* d = &th.desc.d;
    th.dtype = d->dtype;
    if (tile)
      for (int a = 0; a < 3; ++a) th.tile_req[a] = tile[a];
    fill_kargs<double>(&th, d, th.ka64, (size_t)d->nfields * esize, false);
    std::vector<int32_t> gen;
    int ncore = 0;
    build_core<double>(&th, d, th.ka64, th.kc64, gen, ncore);
    const KArgs<double>& ka = th.ka64;
    ldsn = std::max(ldsn, d->nfields * ka.vol);
    if (i == 0)
      for (int a = 0; a < 3; ++a) P.tile[a] = ka.tile[a];
    std::string err;
    const std::string src = jit_class_source(d, geom_of(ka, th.kc64, ncore), "GeoX", err);
    if (src.empty()) return fail(nullptr, MOL_ERR_UNSUPPORTED, "instance " + std::to_string(i) + ": " + err);
    auto it = cls_of.find(src);
    int cls;
    if (it == cls_of.end()) {
      cls = (int)cls_src.size();
      cls_of.emplace(src, cls);
      cls_src.push_back(src);
    } else {
      cls = it->second;
    }
    BInst& b = P.hdr[i];
    std::memset(&b, 0, sizeof(b));
    for (int a = 0; a < 3; ++a) {
      b.n[a] = ka.n[a];
      b.ulo[a] = ka.ulo[a];
      b.uhi[a] = ka.uhi[a];
      b.nt[a] = ka.ntile[a];
      simple_range(d, ka, a, b.slo[a], b.shi[a]);
      b.clo[a] = ncore > 0 ? th.kc64.ct_lo[a] : 1;
      b.chi[a] = ncore > 0 ? th.kc64.ct_lo[a] + th.kc64.ct_n[a] - 1 : 0;
      b.ilo[a] = th.kc64.nd_lo[a];
      b.ihi[a] = th.kc64.nd_hi[a];
    }
    for (int f = 0; f < MOL_MAX_FIELDS; ++f) {
      for (int a = 0; a < 3; ++a) {
        b.lo[f][a] = ka.lo[f][a];
        b.hi[f][a] = ka.hi[f][a];
      }
      b.fofs[f] = P.nunk + ka.fofs[f];
    }
    b.cls = cls;
    b.first = (int32_t)(P.wl.size() / 2);
    const int64_t nt = (int64_t)ka.ntile[0] * ka.ntile[1] * ka.ntile[2];
    if ((int64_t)P.wl.size() / 2 + nt > (int64_t)INT32_MAX / 4) return fail(nullptr, MOL_ERR_SHAPE, "batch too large");
    {
      int64_t box = 1;
      for (int f = 0; f < d->nfields; ++f) {
        int64_t v = 1;
        for (int a = 0; a < 3; ++a) v *= ka.hi[f][a] - ka.lo[f][a] + 1;
        box = std::max(box, v);
      }
      if (box >= (int64_t)INT32_MAX) return fail(nullptr, MOL_ERR_SHAPE, "a batched instance holds 2^31 or more unknowns of one variable");
    }
    for (int64_t k = 0; k < nt; ++k) {
      P.wl.push_back(i);
      P.wl.push_back((int32_t)k);
    }
    P.inst_off.push_back(P.nunk);
    P.nunk += th.nunk;
    std::vector<double>& data = P.data;
    P.dofs[i][0] = (int64_t)data.size();
    data.insert(data.end(), d->pool_h, d->pool_h + d->npool);
    P.dofs[i][1] = (int64_t)data.size();
    data.insert(data.end(), d->cpool_h, d->cpool_h + d->ncpool);
    for (int k = 0; k < d->nslots; ++k) {
      data.push_back(d->slots_h[k].eps);
      data.push_back(d->slots_h[k].inv_dx);
    }
    for (int k = 0; k < d->nparams; ++k) data.push_back(params[i][k]);
    for (int a = 0; a < 3; ++a) {
      P.dofs[i][2 + a] = (int64_t)data.size();
      if (a < d->ndim) data.insert(data.end(), d->coord_h[a], d->coord_h[a] + d->n[a]);
      else data.push_back(0.0);
    }
  }
  P.nclasses = (int)cls_src.size();
  // Work items largest first (nodes of the tile inside its instance's box): the kernel runs one
  // work-group per item, so the dispatcher hands the small items of the tail to the compute
  // units that finish first.
  {
    const size_t nw = P.wl.size() / 2;
    std::vector<std::pair<int64_t, size_t>> order(nw);
    for (size_t k = 0; k < nw; ++k) {
      const BInst& b = P.hdr[P.wl[2 * k]];
      int64_t t = P.wl[2 * k + 1], nodes = 1;
      for (int a = 0; a < 3; ++a) {
        const int ta = (int)(t % b.nt[a]);
        t /= b.nt[a];
        const int lo = b.ulo[a] + ta * P.tile[a];
        nodes *= std::min(P.tile[a], b.uhi[a] - lo + 1);
      }
      order[k] = {-nodes, k};
    }
    std::stable_sort(order.begin(), order.end());
    std::vector<int32_t> wl(P.wl.size());
    for (size_t k = 0; k < nw; ++k) {
      wl[2 * k] = P.wl[2 * order[k].second];
      wl[2 * k + 1] = P.wl[2 * order[k].second + 1];
    }
    P.wl.swap(wl);
  }
  // one kernel: the class bodies and the loop over the work list
  std::ostringstream src;
  src << "// generated kernel of a batch of " << ninst << " instances, " << cls_src.size() << " classes\n";
  src << "#define MOL_T " << (ds[0]->dtype == MOL_F64 ? "double" : "float") << "\n#include \"mol_jit_prelude.h\"\n\n";
  src << "static_assert(sizeof(BInst) == " << sizeof(BInst) << " && sizeof(Epi) == "
      << (ds[0]->dtype == MOL_F64 ? sizeof(Epi<double>) : sizeof(Epi<float>))
      << ", \"argument layout differs from the library\");\n\n";
  for (size_t c = 0; c < cls_src.size(); ++c) {
    std::string body = cls_src[c];
    const std::string from = "struct GeoX {";
    body.replace(body.find(from), from.size(), "struct Geo" + std::to_string(c) + " {");
    src << body << "\n";
  }
  src << "constexpr int LDSN = " << ldsn << ";\n";
  src << "extern \"C\" __global__ void __launch_bounds__(256) mol_batch_tiles(const T* __restrict__ u, Epi ep,\n"
         "    const BInst* __restrict__ insts, const int* __restrict__ wl, const int nwork) {\n"
         "  __shared__ T L[LDSN];\n"
         "  const double t = (ep.tnow ? ep.tnow[0] : 0.0) + ep.tofs;\n"
         "  mol_advance_clock(ep);\n"
         "  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {\n"
         "    const BInst& hd = insts[wl[2 * w]];\n"
         "    const int tile = wl[2 * w + 1];\n"
         "    switch (hd.cls) {\n";
  for (size_t c = 0; c < cls_src.size(); ++c)
    src << "    case " << c << ": mol_batch_tile<Geo" << c << ">(L, hd, tile, u, ep, t); break;\n";
  src << "    default: break;\n    }\n  }\n}\n";
  P.src = src.str();
  return 0;
}

int create_batched(const mol_desc* const* ds, const double* const* params, int32_t ninst, const int32_t* tile,
                   mol_handle** out) {
  BatchPlan P;
  if (int rc = plan_batched(ds, params, ninst, tile, P)) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, MOL_ERR_HIP, "no HIP device available");
  mol_handle* h = new mol_handle();
  h->batch = true;
  h->desc.assign(ds[0]);
  (void)hipGetDevice(&h->device);
  scratch_setup(h->device);
  h->dtype = ds[0]->dtype;
  h->ndim = ds[0]->ndim;
  h->nfields = ds[0]->nfields;
  h->ninst = ninst;
  for (int a = 0; a < 3; ++a) h->ka64.tile[a] = P.tile[a];
  h->nunk = P.nunk;
  h->inst_off = P.inst_off;
  h->nwork = (int)(P.wl.size() / 2);
  h->nclasses = P.nclasses;
  int r = upload(h, &h->d_bdata, P.data.data(), P.data.size());
  const double* base = (const double*)h->d_bdata;
  for (int i = 0; i < ninst; ++i) {
    P.hdr[i].W = base + P.dofs[i][0];
    P.hdr[i].K = base + P.dofs[i][1];
    for (int a = 0; a < 3; ++a) P.hdr[i].coord[a] = base + P.dofs[i][2 + a];
  }
  r |= upload(h, &h->d_insts, P.hdr.data(), P.hdr.size());
  r |= upload(h, &h->d_wl, P.wl.data(), P.wl.size());
  if (r) {
    set_global_error(h->err);
    free_handle(h);
    return MOL_ERR_HIP;
  }
  std::string err;
  if (jit_kernel(P.src, "mol_batch_tiles", h->device, &h->batch_fn, err)) {
    set_global_error("batched kernel: " + err);
    free_handle(h);
    return MOL_ERR_HIP;
  }
  h->batch_grid = std::min(h->nwork, 1 << 24);
  *out = h;
  return 0;
}

}  // namespace

extern "C" {

int mol_create(const mol_desc* d, mol_handle** out) {
  if (!out) return fail(nullptr, MOL_ERR_INVALID, "out is NULL");
  *out = nullptr;
  std::string msg;
  int rc = validate(d, msg);
  if (rc) return fail(nullptr, rc, msg);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(nullptr, MOL_ERR_HIP, "no HIP device available");
  mol_handle* h = new mol_handle();
  h->desc.assign(d);
  d = &h->desc.d;
  (void)hipGetDevice(&h->device);
  scratch_setup(h->device);
  h->dtype = d->dtype;
  h->ndim = d->ndim;
  h->nfields = d->nfields;
  h->naux = d->naux;
  h->nparams = d->nparams;
  for (int a = 0; a < d->ndim; ++a) h->ngrid *= d->n[a];
  int r = 0;
  r |= upload(h, &h->d_tabs, d->tabs_h, d->ntabs);
  r |= upload(h, &h->d_slots, d->slots_h, d->nslots);
  r |= upload(h, &h->d_clos, d->closures_h, d->nclosures);
  r |= upload(h, &h->d_pool, d->pool_h, d->npool);
  r |= upload(h, &h->d_code, d->code_h, d->ncode);
  r |= upload(h, &h->d_cpool, d->cpool_h, d->ncpool);
  r |= upload(h, &h->d_prog, d->prog_h, 2 * (size_t)d->nprogs);
  for (int a = 0; a < d->ndim; ++a) r |= upload(h, &h->d_coord[a], d->coord_h[a], (size_t)d->n[a]);
  if (r) {
    set_global_error(h->err);
    free_handle(h);
    return MOL_ERR_HIP;
  }
  fill_kargs<double>(h, d, h->ka64);
  fill_kargs<float>(h, d, h->ka32);
  if (sync_kargs(h) || setup_paths(h, d)) {
    set_global_error(h->err);
    free_handle(h);
    return MOL_ERR_HIP;
  }
  *out = h;
  return 0;
}

int mol_create_batched(const mol_desc* const* ds, const double* const* params, int32_t ninst, const int32_t* tile,
                       mol_handle** out) {
  if (!out) return fail(nullptr, MOL_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (!ds || ninst < 1) return fail(nullptr, MOL_ERR_INVALID, "no instances");
  if (tile && ((int64_t)std::max(tile[0], 1) * std::max(tile[1], 1) * std::max(tile[2], 1) > 4096 || tile[0] < 0 ||
               tile[1] < 0 || tile[2] < 0))
    return fail(nullptr, MOL_ERR_INVALID, "tile shape out of range (at most 4096 nodes)");
  return create_batched(ds, params, ninst, tile, out);
}

int64_t mol_codegen_batched(const mol_desc* const* ds, const double* const* params, int32_t ninst, const int32_t* tile,
                            char* buf, int64_t buflen, int32_t flags) {
  if (!ds || ninst < 1) return fail(nullptr, MOL_ERR_INVALID, "no instances");
  BatchPlan P;
  if (int rc = plan_batched(ds, params, ninst, tile, P)) return rc;
  if (flags & MOL_CODEGEN_COMPILE) {
    std::vector<char> code;
    std::string log;
    if (jit_compile(P.src, code, log)) return fail(nullptr, MOL_ERR_HIP, "hiprtc: " + log);
  }
  if (buf && buflen > 0) {
    const size_t n = std::min<size_t>(P.src.size(), (size_t)buflen - 1);
    std::memcpy(buf, P.src.data(), n);
    buf[n] = '\0';
  }
  return (int64_t)P.src.size();
}

int mol_batch_offsets(const mol_handle* h, int64_t* offs, int32_t n) {
  if (!h || !offs) return MOL_ERR_INVALID;
  if (!h->batch) return MOL_ERR_INVALID;
  for (int i = 0; i < n && i < h->ninst; ++i) offs[i] = h->inst_off[i];
  return 0;
}

int mol_set_tile(mol_handle* h, int32_t tx, int32_t ty, int32_t tz) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (h->batch) return fail(h, MOL_ERR_UNSUPPORTED, "batched handles take their tile at creation");
  if (tx < 0 || ty < 0 || tz < 0 || (int64_t)std::max(tx, 1) * std::max(ty, 1) * std::max(tz, 1) > 4096)
    return fail(h, MOL_ERR_INVALID, "tile shape out of range (at most 4096 nodes)");
  drop_graphs(h);
  h->tile_req[0] = tx;
  h->tile_req[1] = ty;
  h->tile_req[2] = tz;
  const mol_desc* d = &h->desc.d;
  const void* aux64[MOL_MAX_AUX];
  for (int a = 0; a < MOL_MAX_AUX; ++a) aux64[a] = h->ka64.aux[a];
  fill_kargs<double>(h, d, h->ka64);
  fill_kargs<float>(h, d, h->ka32);
  for (int a = 0; a < MOL_MAX_AUX; ++a) {
    h->ka64.aux[a] = (const double*)aux64[a];
    h->ka32.aux[a] = (const float*)aux64[a];
  }
  if (sync_kargs(h) || setup_paths(h, d)) return MOL_ERR_HIP;
  for (int a = 0; a < MOL_MAX_AUX; ++a) {
    h->kc64.aux[a] = (const double*)aux64[a];
    h->kc32.aux[a] = (const float*)aux64[a];
  }
  return 0;
}

int mol_get_plan(const mol_handle* h, int32_t* info, int32_t ninfo) {
  if (!h || !info) return MOL_ERR_INVALID;
  if (h->batch) {
    // batch: tile, work items, 0, 0, generated, number of classes
    const int32_t b[8] = {h->ka64.tile[0], h->ka64.tile[1], h->ka64.tile[2], h->nwork, 0, 0, 1, h->nclasses};
    for (int i = 0; i < ninfo && i < 8; ++i) info[i] = b[i];
    return 0;
  }
  int32_t v[10] = {h->ka64.tile[0], h->ka64.tile[1], h->ka64.tile[2], h->ncore, h->nbound,
                   0, h->jit_fn != nullptr ? 1 : 0,
                   h->core_interp_ok ? 1 : 0, h->fused_fn ? h->fused_tiles : 0,
                   h->fused_fn ? (h->fused_resident ? 0 : h->fused_steps) : 0};
  for (int i = 0; i < ninfo && i < 10; ++i) info[i] = v[i];
  return 0;
}

int mol_set_aux(mol_handle* h, void* const* aux) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (h->naux > 0 && !aux) return fail(h, MOL_ERR_INVALID, "aux pointers missing");
  drop_graphs(h);
  for (int a = 0; a < h->naux; ++a) {
    if (!aux[a]) return fail(h, MOL_ERR_INVALID, "aux pointer is NULL");
    h->ka64.aux[a] = (const double*)aux[a];
    h->ka32.aux[a] = (const float*)aux[a];
    h->kc64.aux[a] = (const double*)aux[a];
    h->kc32.aux[a] = (const float*)aux[a];
  }
  return sync_kargs(h);
}

static int check_aux(mol_handle* h) {
  for (int a = 0; a < h->naux; ++a)
    if (!h->ka64.aux[a]) return fail(h, MOL_ERR_INVALID, "auxiliary arrays not bound (mol_set_aux)");
  return 0;
}

int mol_rhs(mol_handle* h, void* du, const void* u, const void* p, double t, void* stream) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (!du || !u) return fail(h, MOL_ERR_INVALID, "du/u is NULL");
  if (h->nparams > 0 && !p) return fail(h, MOL_ERR_INVALID, "parameters missing");
  if (int rc = check_aux(h)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == MOL_F64) {
    Epi<double> e{};
    e.mode = 0; e.out = (double*)du; e.tofs = t; e.p = (const double*)p;
    if (launch_rhs<double>(launch_of<double>(h), (const double*)u, e, s))
      return fail(h, MOL_ERR_HIP, std::string("rhs launch: ") + hipGetErrorString(hipGetLastError()));
  } else {
    Epi<float> e{};
    e.mode = 0; e.out = (float*)du; e.tofs = t; e.p = (const float*)p;
    if (launch_rhs<float>(launch_of<float>(h), (const float*)u, e, s))
      return fail(h, MOL_ERR_HIP, std::string("rhs launch: ") + hipGetErrorString(hipGetLastError()));
  }
  return 0;
}

int mol_expand(mol_handle* h, void* full, const void* u, const void* p, double t, void* stream) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (h->batch) return fail(h, MOL_ERR_UNSUPPORTED, "batched handles have no full-grid expansion");
  if (has_interface(&h->desc.d))
    return fail(h, MOL_ERR_UNSUPPORTED, "full-grid expansion of variables joined by an interface");
  if (!full || !u) return fail(h, MOL_ERR_INVALID, "full/u is NULL");
  if (h->nparams > 0 && !p) return fail(h, MOL_ERR_INVALID, "parameters missing");
  if (int rc = check_aux(h)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int rc = h->dtype == MOL_F64
                     ? launch_expand<double>(launch_of<double>(h), (const double*)u, (const double*)p, t,
                                             (double*)full, s)
                     : launch_expand<float>(launch_of<float>(h), (const float*)u, (const float*)p, t,
                                            (float*)full, s);
  if (rc) return fail(h, rc, std::string("expand launch: ") + hipGetErrorString(hipGetLastError()));
  return 0;
}

int mol_rhs_h(mol_handle* h, void* du_h, const void* u_h, const void* p_h, double t) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (!du_h || !u_h) return fail(h, MOL_ERR_INVALID, "du/u is NULL");
  if (h->nparams > 0 && !p_h) return fail(h, MOL_ERR_INVALID, "parameters missing");
  const size_t bytes = (size_t)h->nunk * h->esize();
  const size_t pbytes = (size_t)h->nparams * h->esize();
  // device copies owned by the handle (kept for later calls, released by mol_destroy) and a
  // stream of its own, ordered after the legacy default stream like the caller's work
  if (!h->rhs_stream) HIPCHK(h, hipStreamCreate(&h->rhs_stream));
  if (!h->r_u) HIPCHK(h, hipMalloc(&h->r_u, bytes));
  if (!h->r_du) HIPCHK(h, hipMalloc(&h->r_du, bytes));
  if (pbytes && !h->r_p) HIPCHK(h, hipMalloc(&h->r_p, pbytes));
  hipStream_t s = h->rhs_stream;
  HIPCHK(h, hipMemcpyAsync(h->r_u, u_h, bytes, hipMemcpyHostToDevice, s));
  if (pbytes) HIPCHK(h, hipMemcpyAsync(h->r_p, p_h, pbytes, hipMemcpyHostToDevice, s));
  const int rc = mol_rhs(h, h->r_du, h->r_u, pbytes ? h->r_p : nullptr, t, s);
  if (rc == 0) HIPCHK(h, hipMemcpyAsync(du_h, h->r_du, bytes, hipMemcpyDeviceToHost, s));
  // the host arrays are the caller's: every copy has finished when the call returns
  HIPCHK(h, hipStreamSynchronize(s));
  return rc;
}

int mol_integrate(mol_handle* h, void* u, const void* p, double t0, double dt, int64_t nsteps,
                  int32_t scheme, int32_t flags, void* stream) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (!u) return fail(h, MOL_ERR_INVALID, "u is NULL");
  if (nsteps < 0) return fail(h, MOL_ERR_INVALID, "nsteps < 0");
  if (scheme != MOL_RK4 && scheme != MOL_SSPRK33) return fail(h, MOL_ERR_INVALID, "unknown scheme");
  if (h->nparams > 0 && !p) return fail(h, MOL_ERR_INVALID, "parameters missing");
  if (int rc = check_aux(h)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == MOL_F64)
    return integrate_impl<double>(h, (double*)u, (const double*)p, t0, dt, nsteps, scheme, flags, s);
  return integrate_impl<float>(h, (float*)u, (const float*)p, t0, dt, nsteps, scheme, flags, s);
}

int mol_integrate_noalloc(mol_handle* h, void* u, const void* p, double t0, double dt, int64_t nsteps,
                          int32_t scheme, int32_t flags, void* stream) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (!u) return fail(h, MOL_ERR_INVALID, "u is NULL");
  if (nsteps < 0) return fail(h, MOL_ERR_INVALID, "nsteps < 0");
  if (scheme != MOL_RK4 && scheme != MOL_SSPRK33) return fail(h, MOL_ERR_INVALID, "unknown scheme");
  if (flags & MOL_FLAG_GRAPH)
    return fail(h, MOL_ERR_INVALID, "mol_integrate_noalloc takes no MOL_FLAG_GRAPH (instantiating a graph allocates)");
  if (h->nparams > 0 && !p) return fail(h, MOL_ERR_INVALID, "parameters missing");
  if (int rc = check_aux(h)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == MOL_F64)
    return integrate_noalloc_impl<double>(h, (double*)u, (const double*)p, t0, dt, nsteps, scheme, flags, s);
  return integrate_noalloc_impl<float>(h, (float*)u, (const float*)p, t0, dt, nsteps, scheme, flags, s);
}

int mol_vjp(mol_handle* h, const void* u, const void* p, double t, const void* lam, void* lam_dot, double* g_aux,
            void* stream) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (h->batch) return fail(h, MOL_ERR_UNSUPPORTED, "reverse mode of batched handles");
  if (!u || !lam || !lam_dot) return fail(h, MOL_ERR_INVALID, "u/lam/lam_dot is NULL");
  if (h->nparams > 0 && !p) return fail(h, MOL_ERR_INVALID, "parameters missing");
  if (int rc = check_aux(h)) return rc;
  if (int rc = ensure_vjp(h)) return rc;
  if (g_aux && !h->vjp_aux_ok)
    return fail(h, MOL_ERR_UNSUPPORTED, "auxiliary-array gradient through boundary values that read auxiliary arrays");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == MOL_F64) {
    VArgs<double> va{(const double*)u, (const double*)lam, nullptr, 1.0, 0.0, (double*)lam_dot, g_aux,
                     (const double*)p, nullptr, t};
    va.mode = VJP_GATHER_F | VJP_GATHER_A;
    if (!h->vjp_store && launch_fill_zero<double>((double*)lam_dot, h->nunk, s)) return fail(h, MOL_ERR_HIP, "zero fill launch failed");
    return launch_vjp<double>(h, va, s);
  }
  VArgs<float> va{(const float*)u, (const float*)lam, nullptr, 1.0, 0.0, (float*)lam_dot, g_aux,
                  (const float*)p, nullptr, t};
  va.mode = VJP_GATHER_F | VJP_GATHER_A;
  if (!h->vjp_store && launch_fill_zero<float>((float*)lam_dot, h->nunk, s)) return fail(h, MOL_ERR_HIP, "zero fill launch failed");
  return launch_vjp<float>(h, va, s);
}

int mol_forward_checkpointed(mol_handle* h, void* u, const void* p, double t0, double dt, int64_t nsteps,
                             int32_t ckpt_every, void* stream) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (h->batch) return fail(h, MOL_ERR_UNSUPPORTED, "reverse mode of batched handles");
  if (!u) return fail(h, MOL_ERR_INVALID, "u is NULL");
  if (nsteps < 1 || ckpt_every < 0) return fail(h, MOL_ERR_INVALID, "nsteps must be >= 1 and ckpt_every >= 0");
  if (h->nparams > 0 && !p) return fail(h, MOL_ERR_INVALID, "parameters missing");
  if (int rc = check_aux(h)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == MOL_F64)
    return forward_ckpt_impl<double>(h, (double*)u, (const double*)p, t0, dt, nsteps, ckpt_every, s);
  return forward_ckpt_impl<float>(h, (float*)u, (const float*)p, t0, dt, nsteps, ckpt_every, s);
}

int mol_adjoint(mol_handle* h, const void* p, void* lam, double* g_aux, void* stream) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (h->batch) return fail(h, MOL_ERR_UNSUPPORTED, "reverse mode of batched handles");
  if (!lam) return fail(h, MOL_ERR_INVALID, "lam is NULL");
  if (!h->ck_valid) return fail(h, MOL_ERR_INVALID, "no checkpointed run (call mol_forward_checkpointed first)");
  if (h->nparams > 0 && !p) return fail(h, MOL_ERR_INVALID, "parameters missing");
  if (int rc = check_aux(h)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == MOL_F64) return adjoint_impl<double>(h, (const double*)p, (double*)lam, g_aux, s);
  return adjoint_impl<float>(h, (const float*)p, (float*)lam, g_aux, s);
}

int mol_error_norms(mol_handle* h, const void* u, const double* ref, double* out, void* stream) {
  if (!h) return fail(nullptr, MOL_ERR_INVALID, "handle is NULL");
  if (!u || !out) return fail(h, MOL_ERR_INVALID, "u/out is NULL");
  if (!h->d_norm_part) HIPCHK(h, hipMalloc((void**)&h->d_norm_part, 3 * kNormBlocks * sizeof(double)));
  hipStream_t s = (hipStream_t)stream;
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kNormBlocks, (h->nunk + 255) / 256));
  const int rc = h->dtype == MOL_F64
                     ? launch_norms<double>((const double*)u, ref, h->nunk, h->d_norm_part, nb, out, s)
                     : launch_norms<float>((const float*)u, ref, h->nunk, h->d_norm_part, nb, out, s);
  if (rc) return fail(h, rc, std::string("norm launch: ") + hipGetErrorString(hipGetLastError()));
  return 0;
}

int64_t mol_num_unknowns(const mol_handle* h) { return h ? h->nunk : -1; }

const char* mol_last_error(const mol_handle* h) {
  if (h) return h->err.c_str();
  return global_error();
}

void mol_destroy(mol_handle* h) { free_handle(h); }

JitGeom fwd_geom(const mol_handle& h, int ncore) {
  JitGeom g = geom_of(h.ka64, h.kc64, ncore);
  g.nthr = kernel_threads(g, false);
  return g;
}

int64_t mol_codegen(const mol_desc* d, const int32_t* tile, char* buf, int64_t buflen, int32_t flags) {
  std::string msg;
  int rc = validate(d, msg);
  if (rc) return fail(nullptr, rc, msg);
  mol_handle tmp;
  tmp.desc.assign(d);
  d = &tmp.desc.d;
  if (tile)
    for (int a = 0; a < 3; ++a) tmp.tile_req[a] = tile[a];
  tmp.dtype = d->dtype;
  fill_kargs<double>(&tmp, d, tmp.ka64);
  fill_kargs<float>(&tmp, d, tmp.ka32);
  std::vector<int32_t> gen;
  int ncore = 0;
  build_core<double>(&tmp, d, tmp.ka64, tmp.kc64, gen, ncore);
  {
    // the tiles mol_create uses with the generated kernel (sized for its staged arrays)
    const size_t pos = jit_lds_per_pos(d, geom_of(tmp.ka64, tmp.kc64, ncore), tmp.esize(), false);
    fill_kargs<double>(&tmp, d, tmp.ka64, pos > 0 ? pos : tmp.esize(), true, true);
    gen.clear();
    build_core<double>(&tmp, d, tmp.ka64, tmp.kc64, gen, ncore);
  }
  std::string err;
  std::string src;
  if (flags & MOL_CODEGEN_FUSED) {
    JitGeom g;
    if (fused_geometry(&tmp, d, tmp.ka64, tmp.kc64, ncore, g, err)) src = jit_fused_source(d, g, err);
  } else {
    src = (flags & MOL_CODEGEN_VJP) ? jit_vjp_source(d, vjp_geometry(&tmp, d, nullptr), err)
                                    : jit_source(d, fwd_geom(tmp, ncore), err);
  }
  if (src.empty()) return fail(nullptr, MOL_ERR_UNSUPPORTED, err);
  if (flags & MOL_CODEGEN_COMPILE) {
    std::vector<char> code;
    std::string log;
    if (jit_compile(src, code, log)) return fail(nullptr, MOL_ERR_HIP, "hiprtc: " + log);
  }
  if (buf && buflen > 0) {
    const size_t n = std::min<size_t>(src.size(), (size_t)buflen - 1);
    std::memcpy(buf, src.data(), n);
    buf[n] = '\0';
  }
  return (int64_t)src.size();
}

int mol_device_info(char* buf, int buflen) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return MOL_ERR_HIP;
  if (buf && buflen > 0)
    std::snprintf(buf, buflen, "%s gcn=%s cus=%d", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount);
  return 0;
}

}  // extern "C"
