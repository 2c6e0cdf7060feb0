// Code generation of the core-tile right-hand-side kernel from the stencil IR, and its
// in-process compilation with hiprtc for gfx950.
//
// The reference turns the discretized system into straight-line code with ModelingToolkit's
// build_function (src/MOL_discretization.jl:171-201; the generated Brusselator RHS is shown in
// docs/src/generated/bruss_code.md:46-118).  This file does the same from the descriptor:
// every operator slot becomes a fixed-weight stencil over LDS, every bytecode program becomes
// SSA statements, constants are printed as exact hexadecimal literals, and the result is
// combined with the hand-written tile skeleton (mol_jit_prelude.h) into one HIP kernel per
// system.  The interpreter kernels stay as the path for boundary tiles and as the fallback.
#include <hip/hiprtc.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sys/stat.h>
#include <unistd.h>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "mol_internal.h"

extern "C" const char mol_jit_prelude_text[];   // embedded text of mol_jit_prelude.h

namespace mol {

namespace {

std::string lit(double x) {
  char buf[64];
  if (std::isnan(x)) return "T(__builtin_nan(\"\"))";
  if (std::isinf(x)) return x > 0 ? "T(__builtin_inf())" : "T(-__builtin_inf())";
  std::snprintf(buf, sizeof(buf), "T(%a)", x);
  return buf;
}

std::string istr(long long v) { return std::to_string(v); }

// ---------------------------------------------------------------------------------------
// Expression IR of one node's right-hand side.  Leaves: literals, LDS reads (variable or
// auxiliary array at a displacement from the node), and opaque expressions that depend on no
// variable (coordinates, time, parameters).  Operations: the bytecode opcodes plus stencil
// linear combinations, WENO5 reconstructions, selections and a branch node whose value is
// chosen by the node index along one axis (boundary rows).
enum NodeKind { N_LIT, N_LDS, N_EXPR, N_OP, N_BRANCH };
enum : int { X_LINCOMB = 1000, X_WENO = 1001, X_SPHCOND = 1002, X_SELECT = 1003, X_WENONU = 1004 };

struct Node {
  int kind = N_LIT;
  int op = 0;
  std::vector<int> a;        // operands
  std::vector<double> w;     // X_LINCOMB weights; X_WENO: eps, inv_dx
  double val = 0;            // N_LIT
  int src = 0;               // N_LIT: 0 structural constant, 2 constant-pool entry sidx
  long long sidx = -1;
  std::vector<long long> wi, wj;   // X_LINCOMB: pool index of each weight (wj >= 0: product wi*wj)
  std::vector<std::string> wx;     // X_LINCOMB: weights read at run time (per-node rows)
  int arr = 0;               // N_LDS: array (variable, or nfields + aux index)
  int dv[3] = {0, 0, 0};     //        displacement from the node in nodes per axis
  bool core = false;         //        built for the interior-rows variant
  std::string text;          // N_EXPR
  int imm = 0;               // MOL_OP_POWI exponent; X_WENO / X_WENONU: slot
  int target = 3;            // X_WENONU: reconstruction node 1..5 of the stencil (operands 0..4: u);
                             //   ax: its axis, text: the node's index along it (row of the table)
  // N_BRANCH: region reg[0] when lo <= I_ax <= hi, region reg[k] when I_ax == cval[k], else 0
  int ax = 0, lo = 0, hi = 0;
  std::vector<int> cval, reg;
};

struct Region {
  std::vector<int> nodes;    // N_OP / N_BRANCH nodes in evaluation order
  int result = -1;
};

// A weight row resolved for one node: weights and tap positions relative to the node (in
// nodes along the row's axis).  `valid` is false where the reference would read outside the
// table (such nodes are never evaluated).
struct RowR {
  std::vector<double> w;
  std::vector<long long> widx;   // pool index of each weight
  std::vector<std::string> wx;   // per-node rows at a symbolic node index: weights read from jp.pool
  std::vector<int> rel;
  bool valid = true;
};

// Node-index variant of one axis: the interior-row code (index symbolic) or one concrete index.
struct Var {
  bool interior = true;
  int i = 0;
};

// Leaf context of a program.
enum CtxKind { CTX_NODE, CTX_PLAIN, CTX_HALF, CTX_BOUNDARY };
struct Ctx {
  CtxKind kind = CTX_NODE;
  int ax = 0;                  // half-point axis, or the closure axis (CTX_BOUNDARY)
  const RowR* ip = nullptr;    // interpolation row at the half point
  const RowR* dp = nullptr;    // inner derivative row at the half point
  int dfield = 0;
  int side = 0;                // CTX_BOUNDARY: 0 low (index 1), 1 high (index n)
};

class Gen {
 public:
  Gen(const mol_desc* d, const JitGeom& g) : d_(d), g_(g) {}

  std::string source(bool f64);       // forward kernel mol_rhs_tiles
  std::string source_vjp(bool f64);   // reverse-mode kernel mol_vjp_tiles
  // One whole classic RK4 step per launch (mol_rk4_fused): periodic systems in 1D/2D, every
  // variable staged; empty (with error) when the system does not qualify.
  std::string source_fused(bool f64);
  // Parametric class body for batched instances: weights, constants, grid sizes and boxes are
  // read from the instance header `hd`; the text is identical for instances of one structure.
  std::string source_class(const std::string& name);
  std::string error;

 private:
  const mol_desc* d_;
  JitGeom g_;
  std::vector<Node> nodes_;
  std::vector<Region> regs_;
  std::vector<char> live_;            // node reaches an output of a built variant
  int cur_ = -1;                      // region being built
  std::map<int, int> slot_node_;      // slot -> node id (per node function)
  bool core_ = true;                  // building the interior-rows variant
  bool ok_ = true;
  bool param_ = false;                // parametric (batched) printing
  bool fused_ = false;                // fused RK4 step printing (coordinates from LDS)
  bool stage_noif_ = false;           // printing stage_ni (interface sides read nothing)
  // printing overrides: nodes not printed, and nodes referenced by another expression
  std::vector<char> skip_;
  std::map<int, std::string> subst_;

  // instance-dependent numbers: literal in a specialised kernel, header loads in a class body
  std::string pw(double v, long long idx) const { return param_ ? "(T)hd.W[" + istr(idx) + "]" : lit(v); }
  std::string pk(double v, long long idx) const { return param_ ? "(T)hd.K[" + istr(idx) + "]" : lit(v); }
  std::string NN(int a) const { return param_ ? "hd.n[" + istr(a) + "]" : istr(g_.n[a]); }
  // node index c along axis a: indices in the upper half are written relative to n
  std::string IX(int a, long long c) const {
    if (!param_ || c <= g_.n[a] / 2) return istr(c);
    return "(hd.n[" + istr(a) + "] - " + istr(g_.n[a] - c) + ")";
  }
  std::string LO(int f, int a) const { return param_ ? "hd.lo[" + istr(f) + "][" + istr(a) + "]" : istr(g_.lo[f][a]); }
  std::string HI(int f, int a) const { return param_ ? "hd.hi[" + istr(f) + "][" + istr(a) + "]" : istr(g_.hi[f][a]); }
  // position of the m-th unknown a ghost closure combines (counted inwards from the box edge)
  std::string closure_pos(int f, int a, int side, int m) const {
    if (!param_) return istr(side == 0 ? g_.lo[f][a] + m : g_.hi[f][a] - m);
    return side == 0 ? "(" + LO(f, a) + " + " + istr(m) + ")" : "(" + HI(f, a) + " - " + istr(m) + ")";
  }
  std::string UIDX(int f, const std::string& q0, const std::string& q1, const std::string& q2) const {
    return param_ ? "mol_b_uidx(hd, " + istr(f) + ", " + q0 + ", " + q1 + ", " + q2 + ")"
                  : "uidx(" + istr(f) + ", " + q0 + ", " + q1 + ", " + q2 + ")";
  }
  std::string weight(const Node& n, size_t k) const {
    if (!n.wx.empty()) return n.wx[k];
    if (!param_ || n.wi.empty()) return lit(n.w[k]);
    if (n.wj[k] >= 0) return "(T)(hd.W[" + istr(n.wi[k]) + "] * hd.W[" + istr(n.wj[k]) + "])";
    return "(T)hd.W[" + istr(n.wi[k]) + "]";
  }

  void bad(const std::string& m) {
    if (ok_) error = m;
    ok_ = false;
  }
  // ---- IR construction
  int add(Node n) {
    nodes_.push_back(std::move(n));
    const int id = (int)nodes_.size() - 1;
    if (nodes_[id].kind == N_OP || nodes_[id].kind == N_BRANCH) regs_[cur_].nodes.push_back(id);
    return id;
  }
  int lit_node(double v) {
    Node n;
    n.kind = N_LIT;
    n.val = v;
    return add(n);
  }
  int lds(int arr, int dx, int dy, int dz) {
    Node n;
    n.kind = N_LDS;
    n.arr = arr;
    n.dv[0] = dx;
    n.dv[1] = dy;
    n.dv[2] = dz;
    n.core = core_;
    return add(n);
  }
  int lds_ax(int arr, int ax, int rel) {
    int d[3] = {0, 0, 0};
    d[ax] = rel;
    return lds(arr, d[0], d[1], d[2]);
  }
  int klit_node(int idx) {
    Node n;
    n.kind = N_LIT;
    n.val = d_->cpool_h[idx];
    n.src = 2;
    n.sidx = idx;
    return add(n);
  }
  int expr(const std::string& t) {
    Node n;
    n.kind = N_EXPR;
    n.text = t;
    return add(n);
  }
  int op(int code, std::vector<int> args, int imm = 0) {
    Node n;
    n.kind = N_OP;
    n.op = code;
    n.a = std::move(args);
    n.imm = imm;
    return add(n);
  }
  // sum_k w[k] * args[k]; wi/wj: pool indices of the weights (wj[k] >= 0 for products)
  int lincomb(const std::vector<double>& w, const std::vector<int>& args, const std::vector<long long>& wi = {},
              const std::vector<long long>& wj = {}) {
    Node n;
    n.kind = N_OP;
    n.op = X_LINCOMB;
    for (size_t k = 0; k < w.size(); ++k)
      if (w[k] != 0.0 || param_) {
        n.w.push_back(w[k]);
        n.a.push_back(args[k]);
        if (!wi.empty()) {
          n.wi.push_back(wi[k]);
          n.wj.push_back(wj.empty() ? -1 : wj[k]);
        }
      }
    if (n.a.empty()) return lit_node(0.0);
    return add(n);
  }
  int new_region() {
    regs_.push_back(Region{});
    return (int)regs_.size() - 1;
  }

  // Print geometry: LDS extent and the slot of each staged array (-1: read from global memory).
  struct PrintGeo {
    int ext[3] = {1, 1, 1};
    int vol = 1;
    std::vector<int> slot;     // per array: LDS slot among the staged arrays
    std::vector<int> fslot;    // per variable: adjoint slot (reverse mode)
    std::vector<int> aslot;    // per grid field: adjoint slot (reverse mode)
  };
  PrintGeo pg_;
  // reverse mode: every variable is staged, so the tile kernel writes only the tile buffers and
  // the gather kernel stores each unknown's whole adjoint (no zeroing, no atomics)
  bool gather_stores() const {
    if (param_ || pg_.fslot.size() < (size_t)d_->nfields) return false;
    for (int f = 0; f < d_->nfields; ++f)
      if (pg_.fslot[f] < 0) return false;
    return true;
  }
 public:
  // Arrays read away from the node (these are staged in LDS) and their reach per axis, over
  // all variants and over the interior-rows variant only.
  struct Usage {
    std::vector<bool> staged;
    int reach[3] = {0, 0, 0};
    int core_reach[3] = {0, 0, 0};
    bool separable = true;   // every tap offset lies along one axis (no diagonal taps)
  };
  // Build the node IR (both variants) and report the array usage; empty on failure.
  Usage analyze() {
    std::vector<int> o1, o2;
    int r1, r2;
    if (g_.ncore > 0) build_node(true, o1, r1);
    build_node(false, o2, r2);
    return ok_ ? usage() : Usage{};
  }
  Usage usage() const {
    Usage u;
    u.staged.assign(d_->nfields + d_->naux, false);
    for (size_t id = 0; id < nodes_.size(); ++id) {
      const Node& n = nodes_[id];
      if (n.kind != N_LDS || (id < live_.size() && !live_[id])) continue;
      int naxes = 0;
      for (int a = 0; a < 3; ++a) {
        const int r = n.dv[a] < 0 ? -n.dv[a] : n.dv[a];
        if (r) u.staged[n.arr] = true;
        if (r) ++naxes;
        u.reach[a] = std::max(u.reach[a], r);
        if (n.core) u.core_reach[a] = std::max(u.core_reach[a], r);
      }
      if (naxes > 1) u.separable = false;
    }
    return u;
  }
 private:
  std::string coord_node(int d) const {
    // fused step kernel: coordinates of the tile + halo staged in LDS, indexed by the
    // unwrapped node index relative to the first staged position
    if (fused_) return "(T)mol_fc[" + istr(d) + "][I" + istr(d) + " - mol_fo[" + istr(d) + "]]";
    return "(T)jp.coord[" + istr(d) + "][I" + istr(d) + " - 1]";
  }
  std::string coord_tap(int d, int off) const {
    if (off == 0) return coord_node(d);
    if (fused_)
      return "(T)mol_fc[" + istr(d) + "][I" + istr(d) + (off > 0 ? " + " : " - ") + istr(off < 0 ? -off : off) +
             " - mol_fo[" + istr(d) + "]]";
    const std::string k = "I" + istr(d) + (off > 0 ? " + " + istr(off) : " - " + istr(-off));
    if (g_.periodic[d]) return "(T)jp.coord[" + istr(d) + "][mol_wrap(" + k + ", " + NN(d) + ") - 1]";
    // ghost positions past the ends read the end coordinate (never an address outside the array)
    return "(T)jp.coord[" + istr(d) + "][min(max(" + k + ", 1), " + NN(d) + ") - 1]";
  }

  RowR row(const mol_optab& tb, int ax, const Var& v, int idx_rel);
  int apply(const RowR& r, int arr, int ax) {
    if (!r.valid) return lit_node(0.0);
    std::vector<int> args;
    for (size_t k = 0; k < r.w.size(); ++k) args.push_back(lds_ax(arr, ax, r.rel[k]));
    if (!r.wx.empty()) {
      Node n;
      n.kind = N_OP;
      n.op = X_LINCOMB;
      n.w = r.w;
      n.a = args;
      n.wx = r.wx;
      return add(n);
    }
    return lincomb(r.w, args, r.widx);
  }
  void slot_range(const mol_slot& s, int ax, int& lo, int& hi) const;
  int leaf(uint32_t op, int imm, const Ctx& c);
  int prog(int p, const Ctx& c);
  int slot(int k);
  int slot_body(int k, const Var* v);
  int slot_branch(int k, const std::vector<int>& axes, size_t a, Var* v);
  int nonlinlap(const mol_slot& s, int qo, int qi, int qp, const Var& v);
  void build_node(bool core, std::vector<int>& outs, int& region);
  void fold_linear(int first_region, const std::vector<int>& outs);

  // ---- printing
  std::string ref(int id) const;
  // row of the nonuniform-WENO coefficient table of the axis of X_WENONU node n at its node
  std::string wrow(const Node& n) const {
    return "(jp.wtab[" + istr(n.ax) + "] + MOL_NU_NC * (" + n.text + " - 1))";
  }
  std::string op_expr(const Node& n, int id) const;
  void print_region(std::ostream& o, int r) const;
  void print_node(std::ostream& o, int id) const;
  void print_reverse(std::ostream& o, int r, const std::vector<std::pair<int, std::string>>& seeds) const;
  void partials(const Node& n, int id, std::vector<std::pair<int, std::string>>& out) const;
  std::string stage_fn(bool lds = false);
  std::string scatter_fn();
  PrintGeo make_pg(const int* halo, const Usage& us) const;
  // Node-invariant part of region r (fused step): values that depend only on constants,
  // parameters and coordinates.  inv[id] marks them; the returned frontier lists those that a
  // time- or state-dependent node (or an output) uses.
  std::vector<int> invariant_frontier(int r, const std::vector<int>& outs, std::vector<char>& inv) const;
  std::string node_hoisted(int region, const std::vector<int>& outs);
  std::string geo_variant(const std::string& name, const int* halo, const Usage& us, bool core, int region,
                          const std::vector<int>& outs, bool vjp);
  std::string source_kernel(bool f64, bool vjp);
  // sizes of the argument structs as the library lays them out, checked in the generated code
  static std::string layout_asserts(bool f64) {
    const size_t epi = f64 ? sizeof(Epi<double>) : sizeof(Epi<float>);
    const size_t jp = f64 ? sizeof(JitPtrs<double>) : sizeof(JitPtrs<float>);
    const size_t va = f64 ? sizeof(VArgs<double>) : sizeof(VArgs<float>);
    return "static_assert(sizeof(Epi) == " + istr((long long)epi) + " && sizeof(JPtrs) == " + istr((long long)jp) +
           " && sizeof(VArgs) == " + istr((long long)va) + " && sizeof(BInst) == " +
           istr((long long)sizeof(BInst)) + ", \"argument layout differs from the library\");\n\n";
  }
  std::string geo_common();
  // the state vector or the full grid has 2^31 entries or more
  bool wide() const {
    long long nunk = 0;
    for (int f = 0; f < d_->nfields; ++f) {
      long long sz = 1;
      for (int a = 0; a < 3; ++a) sz *= g_.hi[f][a] - g_.lo[f][a] + 1;
      nunk = std::max(nunk, (long long)g_.fofs[f] + sz);
    }
    const long long ngrid = (long long)g_.n[0] * g_.n[1] * g_.n[2];
    return nunk >= (1LL << 31) - 1 || ngrid >= (1LL << 31) - 1;
  }
  // blocking axis and nodes per thread along it: the slowest axis with more than one tile
  // node; RB divides the tile along it and leaves at least one node group per thread
  void blocking(int& ba, int& rb) const {
    ba = 0;
    rb = 1;
    for (int a = 1; a < d_->ndim; ++a)
      if (g_.tile[a] > 1) ba = a;
    if (ba == 0) return;
    const long long vol = (long long)g_.tile[0] * g_.tile[1] * g_.tile[2];
    // MOL_RB=k asks for k nodes per thread whenever k divides the tile (tests of small tiles)
    const char* e = std::getenv("MOL_RB");
    // 3D tiles: no blocking (128^3 nonlinear diffusion, 16x8x8 tile, 512 threads: 31.3 us with
    // two nodes per thread in turn, 37.7 us with two z-neighbours blocked in registers)
    if (!e && d_->ndim == 3) return;
    // 2D work-groups of 512 threads (64x32 tiles under reach-two stencils): two nodes per
    // thread in registers, two node groups in turn (four in registers: 10.3 against 9.1 us for
    // the Burgers WENO5 right-hand side at 1024^2)
    const int want = e ? std::max(1, std::atoi(e)) : (g_.nthr >= 512 ? 2 : 4);
    for (int r = want; r > 1; r /= 2)
      if (g_.tile[ba] % r == 0 && (e || vol / r >= g_.nthr)) {
        rb = r;
        return;
      }
  }
};

// Row of table `tb` along axis `ax` selected at index (node index + idx_rel), as select_row
// of the interpreter: boundary rows near non-periodic ends, the interior row elsewhere.
RowR Gen::row(const mol_optab& tb, int ax, const Var& v, int idx_rel) {
  RowR r;
  const int n = g_.n[ax];
  if ((v.interior || g_.periodic[ax]) && tb.pn_hi > 0) {
    // per-node row of a nonuniform axis at the symbolic node index: weights read from the pool
    if (param_) {
      bad("per-node weight rows in a batched kernel");
      r.valid = false;
      return r;
    }
    const std::string I = "I" + istr(ax);
    const std::string pos = idx_rel == 0 ? I : "(" + I + (idx_rel > 0 ? " + " : " - ") + istr(std::abs(idx_rel)) + ")";
    for (int k = 0; k < tb.len; ++k) {
      r.w.push_back(1.0);
      r.widx.push_back(-1);
      r.wx.push_back("(T)jp.pool[" + istr((long long)tb.w + k) + " + (" + pos + " - " + istr(tb.pn_lo) + ") * " +
                     istr(tb.len) + "]");
      r.rel.push_back(idx_rel + tb.off + k);
    }
    return r;
  }
  if (v.interior || g_.periodic[ax]) {
    for (int k = 0; k < tb.len; ++k) {
      r.w.push_back(d_->pool_h[tb.w + k]);
      r.widx.push_back(tb.w + k);
      r.rel.push_back(idx_rel + tb.off + k);
    }
    return r;
  }
  const int i = v.i + idx_rel;
  const int lref = n + tb.ref_shift;
  int a0, cnt;
  long long w0;
  if (i <= tb.nlow) {
    if (i < 1) { r.valid = false; return r; }
    w0 = tb.wlow + (long long)(i - 1) * tb.blen;
    a0 = 1;
    cnt = tb.blen;
  } else if (i > lref - tb.nhigh) {
    const int rw = lref - i + tb.hoff;
    if (rw < 1 || rw > tb.nhigh) { r.valid = false; return r; }
    w0 = tb.whigh + (long long)(rw - 1) * tb.blen;
    a0 = lref - tb.blen + 1;
    cnt = tb.blen;
  } else {
    w0 = tb.w;
    if (tb.pn_hi > 0) {
      if (i < tb.pn_lo || i > tb.pn_hi) { r.valid = false; return r; }
      w0 = tb.w + (long long)(i - tb.pn_lo) * tb.len;
    }
    a0 = i + tb.off;
    cnt = tb.len;
  }
  if (w0 < 0 || w0 + cnt > d_->npool) { r.valid = false; return r; }
  for (int k = 0; k < cnt; ++k) {
    r.w.push_back(d_->pool_h[w0 + k]);
    r.widx.push_back(w0 + k);
    r.rel.push_back(a0 + k - v.i);
  }
  return r;
}

void Gen::slot_range(const mol_slot& s, int ax, int& lo, int& hi) const {
  lo = -(1 << 30);
  hi = 1 << 30;
  if (g_.periodic[ax]) return;
  const int n = g_.n[ax];
  auto node_tab = [&](int t) {
    const mol_optab& tb = d_->tabs_h[t];
    lo = std::max(lo, tb.nlow + 1);
    hi = std::min(hi, n + tb.ref_shift - tb.nhigh);
    if (tb.pn_hi > 0) {
      lo = std::max(lo, tb.pn_lo);
      hi = std::min(hi, tb.pn_hi);
    }
  };
  auto half_tabs = [&](int to_i, int ti_i, int tp_i) {
    const mol_optab& to = d_->tabs_h[to_i];
    // outer row at index i-1
    lo = std::max(lo, to.nlow + 2);
    hi = std::min(hi, n + to.ref_shift - to.nhigh + 1);
    if (to.pn_hi > 0) {
      lo = std::max(lo, to.pn_lo + 1);
      hi = std::min(hi, to.pn_hi + 1);
    }
    // inner / interpolation rows at the outer taps h = i-1+off .. i-1+off+len-1
    for (int t : {ti_i, tp_i}) {
      const mol_optab& tb = d_->tabs_h[t];
      lo = std::max(lo, tb.nlow + 2 - to.off);
      hi = std::min(hi, n + tb.ref_shift - tb.nhigh + 2 - to.off - to.len);
      if (tb.pn_hi > 0) {
        lo = std::max(lo, tb.pn_lo + 1 - to.off);
        hi = std::min(hi, tb.pn_hi + 2 - to.off - to.len);
      }
    }
  };
  switch (s.kind) {
    case MOL_SLOT_CENTERED: node_tab(s.tab[0]); break;
    case MOL_SLOT_UPWIND: node_tab(s.tab[0]); node_tab(s.tab[1]); break;
    case MOL_SLOT_MIXED: node_tab(ax == s.axis ? s.tab[0] : s.tab[1]); break;
    case MOL_SLOT_NONLINLAP: half_tabs(s.tab[0], s.tab[1], s.tab[2]); break;
    case MOL_SLOT_SPHERICAL:
      node_tab(s.tab[0]);
      node_tab(s.tab[1]);
      half_tabs(s.tab[2], s.tab[3], s.tab[4]);
      break;
    case MOL_SLOT_WENO_NU:
      lo = std::max(lo, 3);
      hi = std::min(hi, n - 2);
      break;
    default: break;
  }
}

int Gen::leaf(uint32_t code, int imm, const Ctx& c) {
  switch (code) {
    case MOL_OP_CONST: return klit_node(imm);
    case MOL_OP_TIME: return expr("(T)t");
    case MOL_OP_PARAM:
      // batched instances carry their parameter values after the constants, eps and 1/dx
      if (param_) return expr("(T)hd.K[" + istr(d_->ncpool + 2LL * d_->nslots + imm) + "]");
      return expr("P[" + istr(imm) + "]");
    default: break;
  }
  if (c.kind == CTX_BOUNDARY) {
    // ghost-node program G(t, coordinates of the boundary node): no variables
    const std::string q[3] = {"p0", "p1", "p2"};
    std::string idx[3] = {q[0], q[1], q[2]};
    idx[c.ax] = c.side == 0 ? "1" : NN(c.ax);
    switch (code) {
      case MOL_OP_COORD: return expr("(T)jp.coord[" + istr(imm) + "][" + idx[imm] + " - 1]");
      case MOL_OP_AUX:
        return expr("jp.aux[" + istr(imm) + "][(i64)(" + idx[0] + " - 1) + (i64)" + NN(0) + " * ((" +
                    idx[1] + " - 1) + (i64)" + NN(1) + " * (" + idx[2] + " - 1))]");
      default: return lit_node(0.0);
    }
  }
  if (c.kind == CTX_HALF) {
    switch (code) {
      case MOL_OP_FIELD: return apply(*c.ip, imm, c.ax);
      case MOL_OP_AUX: return apply(*c.ip, d_->nfields + imm, c.ax);
      case MOL_OP_COORD: {
        if (imm != c.ax) return expr(coord_node(imm));
        if (!c.ip->valid) return lit_node(0.0);
        std::string e;
        for (size_t k = 0; k < c.ip->w.size(); ++k) {
          const std::string wk = c.ip->wx.empty() ? pw(c.ip->w[k], c.ip->widx[k]) : c.ip->wx[k];
          const std::string term = wk + " * " + coord_tap(imm, c.ip->rel[k]);
          e = k == 0 ? term : e + " + " + term;
        }
        return e.empty() ? lit_node(0.0) : expr("(" + e + ")");
      }
      case MOL_OP_DINNER: return apply(*c.dp, c.dfield, c.ax);
      case MOL_OP_SLOT: return lit_node(0.0);
      default: bad("unknown leaf opcode"); return lit_node(0.0);
    }
  }
  switch (code) {
    case MOL_OP_FIELD: return lds(imm, 0, 0, 0);
    case MOL_OP_AUX: return lds(d_->nfields + imm, 0, 0, 0);
    case MOL_OP_COORD: return expr(coord_node(imm));
    case MOL_OP_SLOT: return c.kind == CTX_NODE ? slot(imm) : lit_node(0.0);
    case MOL_OP_DINNER: return lit_node(0.0);
    default: bad("unknown leaf opcode"); return lit_node(0.0);
  }
}

bool known_unary(uint32_t code) {
  switch (code) {
    case MOL_OP_NEG: case MOL_OP_ABS: case MOL_OP_SQRT: case MOL_OP_EXP: case MOL_OP_LOG:
    case MOL_OP_SIN: case MOL_OP_COS: case MOL_OP_TAN: case MOL_OP_TANH: case MOL_OP_SINH:
    case MOL_OP_COSH: case MOL_OP_ATAN: case MOL_OP_ASIN: case MOL_OP_ACOS: case MOL_OP_SIGN:
    case MOL_OP_POWI: case MOL_OP_ERF: case MOL_OP_LOG10: case MOL_OP_CBRT: case MOL_OP_INV:
    case MOL_OP_SQ: case MOL_OP_EXPM1: case MOL_OP_LOG1P: case MOL_OP_FLOOR: case MOL_OP_CEIL:
    case MOL_OP_HEAVISIDE:
      return true;
    default:
      return false;
  }
}

bool known_binary(uint32_t code) { return code >= MOL_OP_ADD && code <= MOL_OP_OR; }

// Translation of a stack program into IR (same evaluation order as stack_prog).
int Gen::prog(int p, const Ctx& c) {
  const int start = d_->prog_h[2 * p], len = d_->prog_h[2 * p + 1];
  std::vector<int> st;
  auto pop = [&]() {
    const int v = st.back();
    st.pop_back();
    return v;
  };
  for (int pc = start; pc < start + len; ++pc) {
    const uint32_t ins = d_->code_h[pc];
    const uint32_t code = ins & 0xffu;
    const bool bconst = (ins >> 20) & 1u;
    const int imm = (int)(ins >> 21);
    if (code < 16u) {
      if (code > MOL_OP_DINNER) { bad("opcode " + istr(code) + " not supported by the generator"); return lit_node(0); }
      st.push_back(leaf(code, imm, c));
      continue;
    }
    if (code == MOL_OP_SELECT) {
      if (bconst) {
        if (st.size() < 2) { bad("stack underflow"); return lit_node(0); }
        const int s0 = pop(), s1 = pop();
        st.push_back(op(X_SELECT, {s1, s0, klit_node(imm)}));
      } else {
        if (st.size() < 3) { bad("stack underflow"); return lit_node(0); }
        const int s0 = pop(), s1 = pop(), s2 = pop();
        st.push_back(op(X_SELECT, {s2, s1, s0}));
      }
      continue;
    }
    if (code < 48u) {
      if (!known_binary(code)) { bad("opcode " + istr(code) + " not supported by the generator"); return lit_node(0); }
      int a, b;
      if (bconst) {
        if (st.empty()) { bad("stack underflow"); return lit_node(0); }
        a = pop();
        b = klit_node(imm);
      } else {
        if (st.size() < 2) { bad("stack underflow"); return lit_node(0); }
        b = pop();
        a = pop();
      }
      st.push_back(op((int)code, {a, b}));
      continue;
    }
    if (!known_unary(code)) { bad("opcode " + istr(code) + " not supported by the generator"); return lit_node(0); }
    if (st.empty()) { bad("stack underflow"); return lit_node(0); }
    const int a = pop();
    st.push_back(op((int)code, {a}, code == MOL_OP_POWI ? ((imm & 0x400) ? imm - 0x800 : imm) : 0));
  }
  if (st.size() != 1) { bad("program leaves " + istr((long long)st.size()) + " values"); return lit_node(0); }
  return st.back();
}

// Nonlinear Laplacian at a node: outer half-point operator at index i-1, the coefficient
// program and the inner derivative at each outer tap h
// (src/discretization/schemes/nonlinear_laplacian/nonlinear_laplacian.jl:20-83).
int Gen::nonlinlap(const mol_slot& s, int qo, int qi, int qp, const Var& v) {
  const RowR ro = row(d_->tabs_h[s.tab[qo]], s.axis, v, -1);
  if (!ro.valid) return lit_node(0.0);
  std::vector<int> terms;
  for (size_t k = 0; k < ro.w.size(); ++k) {
    const int h = ro.rel[k];   // half point between node+h and node+h+1
    const RowR rp = row(d_->tabs_h[s.tab[qp]], s.axis, v, h);
    const RowR rd = row(d_->tabs_h[s.tab[qi]], s.axis, v, h);
    Ctx c;
    c.kind = CTX_HALF;
    c.ax = s.axis;
    c.ip = &rp;
    c.dp = &rd;
    c.dfield = s.field;
    const int di = leaf(MOL_OP_DINNER, 0, c);
    terms.push_back(s.prog >= 0 ? op(MOL_OP_MUL, {prog(s.prog, c), di}) : di);
  }
  if (!ro.wx.empty()) {
    Node n;
    n.kind = N_OP;
    n.op = X_LINCOMB;
    n.w = ro.w;
    n.a = terms;
    n.wx = ro.wx;
    return add(n);
  }
  return lincomb(ro.w, terms, ro.widx);
}

// Value of slot k for the node-index variants v[axis].
int Gen::slot_body(int k, const Var* v) {
  const mol_slot& s = d_->slots_h[k];
  const int ax = s.axis;
  Ctx plain;
  plain.kind = CTX_PLAIN;
  switch (s.kind) {
    case MOL_SLOT_CENTERED:
      return apply(row(d_->tabs_h[s.tab[0]], ax, v[ax], 0), s.field, ax);
    case MOL_SLOT_UPWIND: {
      if (s.prog < 0) return apply(row(d_->tabs_h[s.tab[1]], ax, v[ax], 0), s.field, ax);
      const int coef = prog(s.prog, plain);
      const int back = apply(row(d_->tabs_h[s.tab[1]], ax, v[ax], 0), s.field, ax);
      const int fwd = apply(row(d_->tabs_h[s.tab[0]], ax, v[ax], 0), s.field, ax);
      const int cond = op(MOL_OP_GT, {coef, lit_node(0.0)});
      return op(MOL_OP_MUL, {coef, op(X_SELECT, {cond, back, fwd})});
    }
    case MOL_SLOT_WENO: {
      Node n;
      n.kind = N_OP;
      n.op = X_WENO;
      for (int q = -2; q <= 2; ++q) n.a.push_back(lds_ax(s.field, ax, q));
      n.w = {s.eps, s.inv_dx};
      n.imm = k;
      return add(n);
    }
    case MOL_SLOT_WENO_NU: {
      // stencil and target (get_f_taps_coords, function_scheme.jl:1-34): one-sided at nodes
      // 1, 2 and n-1, n of a non-periodic axis, centred elsewhere
      const int n = g_.n[ax];
      int first = -2, target = 3;
      if (!g_.periodic[ax] && !v[ax].interior) {
        const int c = v[ax].i;
        if (c <= 2) {
          target = c;
          first = 1 - c;
        } else if (c >= n - 1) {
          target = 5 - (n - c);
          first = n - 4 - c;
        }
      }
      // the coordinate-only coefficients of the node come from its row of the axis table
      // (mol_weno_nu_row, built once per handle by mol_wenonu_table)
      Node nd;
      nd.kind = N_OP;
      nd.op = X_WENONU;
      nd.target = target;
      nd.ax = ax;
      nd.text = (!g_.periodic[ax] && !v[ax].interior) ? IX(ax, v[ax].i) : "I" + istr(ax);
      for (int q = 0; q < 5; ++q) nd.a.push_back(lds_ax(s.field, ax, first + q));
      nd.w = {s.eps};
      nd.imm = k;
      return add(nd);
    }
    case MOL_SLOT_NONLINLAP:
      return nonlinlap(s, 0, 1, 2, v[ax]);
    case MOL_SLOT_SPHERICAL: {
      const int a = s.prog >= 0 ? prog(s.prog, plain) : lit_node(1.0);
      const int rr = expr(coord_node(ax));
      const int d2 = apply(row(d_->tabs_h[s.tab[1]], ax, v[ax], 0), s.field, ax);
      const int d1 = apply(row(d_->tabs_h[s.tab[0]], ax, v[ax], 0), s.field, ax);
      const int nl = nonlinlap(s, 2, 3, 4, v[ax]);
      // r ~ 0: 6 a D2 u (src/discretization/schemes/spherical_laplacian/spherical_laplacian.jl:30-32)
      const int at0 = op(MOL_OP_MUL, {op(MOL_OP_MUL, {lit_node(6.0), a}), d2});
      const int gen = op(MOL_OP_MUL, {a, op(MOL_OP_ADD, {op(MOL_OP_DIV, {d1, rr}), nl})});
      return op(X_SELECT, {op(X_SPHCOND, {rr}), at0, gen});
    }
    case MOL_SLOT_MIXED: {
      // sum_i sum_j (wx_i * wy_j) u[i, j] (2nd_order_mixed_deriv.jl:5-22); on a nonuniform
      // axis the interior row of the node is read from the pool at run time
      const RowR rx = row(d_->tabs_h[s.tab[0]], ax, v[ax], 0);
      const RowR ry = row(d_->tabs_h[s.tab[1]], s.axis2, v[s.axis2], 0);
      if (!rx.valid || !ry.valid) return lit_node(0.0);
      const bool runtime = !rx.wx.empty() || !ry.wx.empty();
      std::vector<double> w;
      std::vector<long long> wi, wj;
      std::vector<std::string> wx;
      std::vector<int> args;
      for (size_t i = 0; i < rx.w.size(); ++i)
        for (size_t j = 0; j < ry.w.size(); ++j) {
          w.push_back(runtime ? 1.0 : rx.w[i] * ry.w[j]);
          wi.push_back(rx.widx[i]);
          wj.push_back(ry.widx[j]);
          if (runtime)
            wx.push_back("(" + (rx.wx.empty() ? lit(rx.w[i]) : rx.wx[i]) + " * " +
                         (ry.wx.empty() ? lit(ry.w[j]) : ry.wx[j]) + ")");
          int dd[3] = {0, 0, 0};
          dd[ax] += rx.rel[i];
          dd[s.axis2] += ry.rel[j];
          args.push_back(lds(s.field, dd[0], dd[1], dd[2]));
        }
      if (!runtime) return lincomb(w, args, wi, wj);
      Node n;
      n.kind = N_OP;
      n.op = X_LINCOMB;
      n.w = w;
      n.a = args;
      n.wx = wx;
      return add(n);
    }
    default:
      bad("slot kind " + istr(s.kind) + " not supported by the generator");
      return lit_node(0.0);
  }
}

// Slot k as branch nodes over the node-index variants of its axes.
int Gen::slot_branch(int k, const std::vector<int>& axes, size_t a, Var* v) {
  if (a == axes.size()) return slot_body(k, v);
  const int ax = axes[a];
  int lo, hi;
  slot_range(d_->slots_h[k], ax, lo, hi);
  lo = std::max(lo, g_.ulo[ax]);
  hi = std::min(hi, g_.uhi[ax]);
  std::vector<int> cases;
  for (int i = g_.ulo[ax]; i <= g_.uhi[ax]; ++i)
    if (i < lo || i > hi) cases.push_back(i);
  v[ax] = Var{};
  if (cases.empty()) return slot_branch(k, axes, a + 1, v);
  Node br;
  br.kind = N_BRANCH;
  br.ax = ax;
  br.lo = lo;
  br.hi = hi;
  const int parent = cur_;
  // interior region
  {
    const int r = new_region();
    cur_ = r;
    regs_[r].result = slot_branch(k, axes, a + 1, v);
    br.reg.push_back(r);
    br.cval.push_back(0);
  }
  for (int c : cases) {
    const int r = new_region();
    cur_ = r;
    v[ax] = Var{false, c};
    regs_[r].result = slot_branch(k, axes, a + 1, v);
    br.reg.push_back(r);
    br.cval.push_back(c);
  }
  v[ax] = Var{};
  cur_ = parent;
  return add(br);
}

int Gen::slot(int k) {
  auto it = slot_node_.find(k);
  if (it != slot_node_.end()) return it->second;
  const mol_slot& s = d_->slots_h[k];
  Var v[3];
  int r;
  if (core_ || s.kind == MOL_SLOT_WENO) {
    r = slot_body(k, v);
  } else {
    std::vector<int> axes = {s.axis};
    if (s.kind == MOL_SLOT_MIXED) axes.push_back(s.axis2);
    r = slot_branch(k, axes, 0, v);
  }
  slot_node_[k] = r;
  return r;
}

void Gen::build_node(bool core, std::vector<int>& outs, int& region) {
  core_ = core;
  slot_node_.clear();
  region = new_region();
  cur_ = region;
  Ctx node;
  node.kind = CTX_NODE;
  outs.clear();
  for (int f = 0; f < d_->nfields; ++f) outs.push_back(prog(d_->rhs_prog[f], node));
  if (ok_) fold_linear(region, outs);
  // drop the operations no output depends on (e.g. the terms of a boundary row whose weight is
  // zero, built before the row drops them): their taps would widen the staged halo
  live_.resize(nodes_.size(), 0);
  std::vector<int> todo(outs.begin(), outs.end());
  while (!todo.empty()) {
    const int id = todo.back();
    todo.pop_back();
    if (live_[id]) continue;
    live_[id] = 1;
    const Node& n = nodes_[id];
    for (int a : n.a) todo.push_back(a);
    if (n.kind == N_BRANCH)
      for (int r : n.reg)
        if (regs_[r].result >= 0) todo.push_back(regs_[r].result);
  }
  for (size_t r = (size_t)region; r < regs_.size(); ++r) {
    std::vector<int>& v = regs_[r].nodes;
    v.erase(std::remove_if(v.begin(), v.end(), [&](int id) { return !live_[id]; }), v.end());
  }
}

// Linear forms collected as mtkcompile's expand-and-collect does before build_function: the
// reference's generated f carries one coefficient per distinct tap (the Brusselator row reads
// 1.0 + (-644.4) u[i,j] + 160.0 u[i+1,j] + ... + u[i,j]^2 v[i,j],
// docs/src/generated/bruss_code.md:46-118), where the operator tree of the equation has a
// stencil sum per derivative, a coefficient product per term and a sum over terms.  Chains of
// additions, subtractions, negations, products with (or divisions by) a literal and stencil
// combinations with numeric weights, whose intermediate values have one use, become one
// X_LINCOMB: a constant plus one weight per distinct operand (taps, and values the chain does
// not absorb).  Values that read no LDS tap are never absorbed, so what depends only on
// constants, coordinates and time keeps its own node (hoisted by the fused step).  The weights
// are products and sums of the literals in FP64; the result differs from the unfolded tree by
// rounding only.  Batched class bodies read their weights at run time and are not folded.
void Gen::fold_linear(int first_region, const std::vector<int>& outs) {
  if (param_) return;
  const int N = (int)nodes_.size();
  std::vector<int> reg_of(N, -1), uses(N, 0), consumer(N, -1);
  for (size_t r = (size_t)first_region; r < regs_.size(); ++r)
    for (int id : regs_[r].nodes) reg_of[id] = (int)r;
  auto use = [&](int a, int by) {
    uses[a]++;
    consumer[a] = by;
  };
  for (size_t r = (size_t)first_region; r < regs_.size(); ++r)
    for (int id : regs_[r].nodes) {
      const Node& n = nodes_[id];
      for (int a : n.a) use(a, id);
      if (n.kind == N_BRANCH)
        for (int sr : n.reg)
          if (regs_[sr].result >= 0) use(regs_[sr].result, -1);
    }
  for (int f : outs) use(f, -1);
  auto is_lit = [&](int id) { return nodes_[id].kind == N_LIT; };
  auto foldable = [&](int id) {
    const Node& n = nodes_[id];
    if (n.kind != N_OP || reg_of[id] < 0) return false;
    switch (n.op) {
      case MOL_OP_ADD: case MOL_OP_SUB: case MOL_OP_NEG: return true;
      case MOL_OP_MUL: return is_lit(n.a[0]) || is_lit(n.a[1]);
      case MOL_OP_DIV: return is_lit(n.a[1]) && nodes_[n.a[1]].val != 0.0;
      case X_LINCOMB: return n.wx.empty();
      default: return false;
    }
  };
  // ids grow from operands to users, so one pass in id order is a topological order
  std::vector<char> dyn(N, 0), fold(N, 0), absorbed(N, 0);
  for (int id = 0; id < N; ++id) {
    const Node& n = nodes_[id];
    if (n.kind == N_LDS) dyn[id] = 1;
    for (int a : n.a) dyn[id] |= dyn[a];
    if (n.kind == N_BRANCH)
      for (int sr : n.reg)
        if (regs_[sr].result >= 0) dyn[id] |= dyn[regs_[sr].result];
    fold[id] = foldable(id);
  }
  for (int id = 0; id < N; ++id)
    absorbed[id] = fold[id] && dyn[id] && uses[id] == 1 && consumer[id] >= 0 && fold[consumer[id]] &&
                   reg_of[consumer[id]] == reg_of[id];
  struct Form {
    double c = 0.0;
    std::vector<std::pair<int, double>> t;   // operand, weight (first-appearance order)
  };
  std::map<int, Form> F;
  auto axpy = [](Form& y, double s, const Form& x) {
    y.c += s * x.c;
    for (const auto& tw : x.t) {
      auto it = std::find_if(y.t.begin(), y.t.end(), [&](const std::pair<int, double>& p) { return p.first == tw.first; });
      if (it == y.t.end()) y.t.push_back({tw.first, s * tw.second});
      else it->second += s * tw.second;
    }
  };
  // reads of one array at one displacement are one operand, whichever node made them
  std::map<std::array<int, 5>, int> tap;
  std::vector<int> canon(N);
  for (int id = 0; id < N; ++id) {
    const Node& n = nodes_[id];
    canon[id] = id;
    if (n.kind == N_LDS)
      canon[id] = tap.emplace(std::array<int, 5>{n.arr, n.dv[0], n.dv[1], n.dv[2], n.core ? 1 : 0}, id).first->second;
  }
  auto G = [&](int a) {
    Form g;
    if (is_lit(a)) g.c = nodes_[a].val;
    else if (absorbed[a]) g = F[a];
    else g.t.push_back({canon[a], 1.0});
    return g;
  };
  for (int id = 0; id < N; ++id) {
    if (!fold[id]) continue;
    const Node& n = nodes_[id];
    Form f;
    switch (n.op) {
      case MOL_OP_ADD: axpy(f, 1.0, G(n.a[0])); axpy(f, 1.0, G(n.a[1])); break;
      case MOL_OP_SUB: axpy(f, 1.0, G(n.a[0])); axpy(f, -1.0, G(n.a[1])); break;
      case MOL_OP_NEG: axpy(f, -1.0, G(n.a[0])); break;
      case MOL_OP_MUL:
        if (is_lit(n.a[0])) axpy(f, nodes_[n.a[0]].val, G(n.a[1]));
        else axpy(f, nodes_[n.a[1]].val, G(n.a[0]));
        break;
      case MOL_OP_DIV: axpy(f, 1.0 / nodes_[n.a[1]].val, G(n.a[0])); break;
      default:   // numeric stencil combination
        for (size_t k = 0; k < n.a.size(); ++k) axpy(f, n.w[k], G(n.a[k]));
        break;
    }
    F[id] = std::move(f);
  }
  // each chain's last value becomes one combination; absorbed values lose their last user
  for (int id = 0; id < N; ++id) {
    if (!fold[id] || absorbed[id] || !dyn[id]) continue;
    const Form& f = F[id];
    std::vector<int> args;
    std::vector<double> w;
    if (f.c != 0.0) {
      args.push_back(lit_node(f.c));
      w.push_back(1.0);
    }
    for (const auto& tw : f.t)
      if (tw.second != 0.0) {
        args.push_back(tw.first);
        w.push_back(tw.second);
      }
    if (args.empty()) {
      args.push_back(lit_node(0.0));
      w.push_back(1.0);
    }
    Node& n = nodes_[id];
    n.op = X_LINCOMB;
    n.a = std::move(args);
    n.w = std::move(w);
    n.wi.clear();
    n.wj.clear();
    n.wx.clear();
  }
}

// ---------------------------------------------------------------------------------------
// printing

std::string Gen::ref(int id) const {
  const Node& n = nodes_[id];
  if (!subst_.empty()) {
    const auto it = subst_.find(id);
    if (it != subst_.end()) return it->second;
  }
  switch (n.kind) {
    case N_LIT: return n.src == 2 ? pk(n.val, n.sidx) : lit(n.val);
    case N_LDS: {
      const int sl = pg_.slot.empty() ? n.arr : pg_.slot[n.arr];
      if (sl < 0) {
        // read at the node only: straight from global memory
        if (n.arr < d_->nfields) return "u[" + UIDX(n.arr, "I0", "I1", "I2") + "]";
        return "jp.aux[" + istr(n.arr - d_->nfields) + "][gidx(I0, I1, I2)]";
      }
      const long long o = (long long)sl * pg_.vol + n.dv[0] + (long long)pg_.ext[0] * (n.dv[1] + (long long)pg_.ext[1] * n.dv[2]);
      if (o == 0) return "L[b]";
      return o > 0 ? "L[b + " + istr(o) + "]" : "L[b - " + istr(-o) + "]";
    }
    case N_EXPR: return n.text;
    default: return "v" + istr(id);
  }
}

const char* unary_fn(uint32_t code) {
  switch (code) {
    case MOL_OP_ABS: return "fabs";
    case MOL_OP_SQRT: return "sqrt";
    case MOL_OP_EXP: return "exp";
    case MOL_OP_LOG: return "log";
    case MOL_OP_SIN: return "sin";
    case MOL_OP_COS: return "cos";
    case MOL_OP_TAN: return "tan";
    case MOL_OP_TANH: return "tanh";
    case MOL_OP_SINH: return "sinh";
    case MOL_OP_COSH: return "cosh";
    case MOL_OP_ATAN: return "atan";
    case MOL_OP_ASIN: return "asin";
    case MOL_OP_ACOS: return "acos";
    case MOL_OP_ERF: return "erf";
    case MOL_OP_LOG10: return "log10";
    case MOL_OP_CBRT: return "cbrt";
    case MOL_OP_EXPM1: return "expm1";
    case MOL_OP_LOG1P: return "log1p";
    case MOL_OP_FLOOR: return "floor";
    case MOL_OP_CEIL: return "ceil";
    case MOL_OP_SIGN: return "mol_sign";
    case MOL_OP_HEAVISIDE: return "mol_heaviside";
    default: return nullptr;
  }
}

std::string powi_expr(const std::string& a, int ex) {
  const int m = ex < 0 ? -ex : ex;
  std::string r;
  if (m == 0) r = "T(1)";
  else if (m == 1) r = a;
  else if (m == 2) r = "(" + a + " * " + a + ")";
  else if (m == 3) r = "(" + a + " * " + a + " * " + a + ")";
  else return "mol_powi(" + a + ", " + istr(ex) + ")";
  return ex < 0 ? "(T(1) / " + r + ")" : r;
}

std::string Gen::op_expr(const Node& n, int id) const {
  (void)id;
  auto A = [&](int k) { return ref(n.a[k]); };
  switch (n.op) {
    case MOL_OP_ADD: return "(" + A(0) + " + " + A(1) + ")";
    case MOL_OP_SUB: return "(" + A(0) + " - " + A(1) + ")";
    case MOL_OP_MUL: return "(" + A(0) + " * " + A(1) + ")";
    case MOL_OP_DIV: return "(" + A(0) + " / " + A(1) + ")";
    case MOL_OP_POW: return "pow(" + A(0) + ", " + A(1) + ")";
    case MOL_OP_MIN: return "fmin(" + A(0) + ", " + A(1) + ")";
    case MOL_OP_MAX: return "fmax(" + A(0) + ", " + A(1) + ")";
    case MOL_OP_LT: return "(" + A(0) + " < " + A(1) + " ? T(1) : T(0))";
    case MOL_OP_LE: return "(" + A(0) + " <= " + A(1) + " ? T(1) : T(0))";
    case MOL_OP_GT: return "(" + A(0) + " > " + A(1) + " ? T(1) : T(0))";
    case MOL_OP_GE: return "(" + A(0) + " >= " + A(1) + " ? T(1) : T(0))";
    case MOL_OP_EQ: return "(" + A(0) + " == " + A(1) + " ? T(1) : T(0))";
    case MOL_OP_NE: return "(" + A(0) + " != " + A(1) + " ? T(1) : T(0))";
    case MOL_OP_AND: return "((" + A(0) + " != T(0) && " + A(1) + " != T(0)) ? T(1) : T(0))";
    case MOL_OP_OR: return "((" + A(0) + " != T(0) || " + A(1) + " != T(0)) ? T(1) : T(0))";
    case X_SELECT: return "(" + A(0) + " != T(0) ? " + A(1) + " : " + A(2) + ")";
    case X_SPHCOND: return "(fabs((double)" + A(0) + ") <= 1e-6 ? T(1) : T(0))";
    case MOL_OP_NEG: return "(-" + A(0) + ")";
    case MOL_OP_SQ: return "(" + A(0) + " * " + A(0) + ")";
    case MOL_OP_INV: return "(T(1) / " + A(0) + ")";
    case MOL_OP_POWI: return powi_expr(A(0), n.imm);
    case X_LINCOMB: {
      std::string e;
      const bool numeric = n.wx.empty() && (!param_ || n.wi.empty());
      for (size_t k = 0; k < n.a.size(); ++k) {
        // unit weights (folded linear forms) multiply nothing
        if (numeric && (n.w[k] == 1.0 || n.w[k] == -1.0)) {
          const bool neg = n.w[k] < 0.0;
          e = k == 0 ? (neg ? "(-" + A((int)k) + ")" : A((int)k)) : e + (neg ? " - " : " + ") + A((int)k);
          continue;
        }
        const std::string term = weight(n, k) + " * " + A((int)k);
        e = k == 0 ? term : e + " + " + term;
      }
      return e;
    }
    case X_WENONU:
      return "mol_weno5_nu_tab(" + A(0) + ", " + A(1) + ", " + A(2) + ", " + A(3) + ", " + A(4) + ", " + wrow(n) +
             ", " + pk(n.w[0], d_->ncpool + 2LL * n.imm) + ")";
    case X_WENO:
      // FP64: one-division cleared form; FP32: normalised weights (the cleared form underflows)
      return std::string(d_->dtype == MOL_F64 ? "mol_weno5(" : "mol_weno5_ref(") + A(0) + ", " + A(1) + ", " +
             A(2) + ", " + A(3) + ", " + A(4) + ", " + pk(n.w[0], d_->ncpool + 2LL * n.imm) + ", " +
             pk(n.w[1], d_->ncpool + 2LL * n.imm + 1) + ")";
    default: {
      const char* fn = unary_fn(n.op);
      return fn ? std::string(fn) + "(" + A(0) + ")" : "T(0)";
    }
  }
}

void Gen::print_node(std::ostream& o, int id) const {
  const Node& n = nodes_[id];
  if (id < (int)skip_.size() && skip_[id]) return;
  if (n.kind == N_OP) {
    o << "    const T v" << id << " = " << op_expr(n, id) << ";\n";
    return;
  }
  if (n.kind != N_BRANCH) return;
  const std::string v = "v" + istr(id), I = "I" + istr(n.ax);
  o << "    T " << v << " = T(0);\n";
  o << "    if (" << I << " >= " << IX(n.ax, n.lo) << " && " << I << " <= " << IX(n.ax, n.hi) << ") {\n";
  print_region(o, n.reg[0]);
  o << "    " << v << " = " << ref(regs_[n.reg[0]].result) << ";\n";
  o << "    }";
  for (size_t k = 1; k < n.reg.size(); ++k) {
    o << " else if (" << I << " == " << IX(n.ax, n.cval[k]) << ") {\n";
    print_region(o, n.reg[k]);
    o << "    " << v << " = " << ref(regs_[n.reg[k]].result) << ";\n    }";
  }
  o << "\n";
}

void Gen::print_region(std::ostream& o, int r) const {
  for (int id : regs_[r].nodes) print_node(o, id);
}

// Partial derivatives of a node with respect to its operands, times the node's adjoint g<id>.
void Gen::partials(const Node& n, int id, std::vector<std::pair<int, std::string>>& out) const {
  const std::string g = "g" + istr(id), v = "v" + istr(id);
  auto A = [&](int k) { return ref(n.a[k]); };
  auto dep = [&](int k) { return nodes_[n.a[k]].kind == N_LDS || nodes_[n.a[k]].kind == N_OP || nodes_[n.a[k]].kind == N_BRANCH; };
  auto put = [&](int k, const std::string& e) {
    if (dep(k)) out.push_back({n.a[k], e});
  };
  switch (n.op) {
    case MOL_OP_ADD: put(0, g); put(1, g); break;
    case MOL_OP_SUB: put(0, g); put(1, "(-" + g + ")"); break;
    case MOL_OP_MUL: put(0, g + " * " + A(1)); put(1, g + " * " + A(0)); break;
    case MOL_OP_DIV: put(0, g + " / " + A(1)); put(1, "(-" + g + " * " + v + " / " + A(1) + ")"); break;
    case MOL_OP_POW:
      put(0, g + " * " + A(1) + " * pow(" + A(0) + ", " + A(1) + " - T(1))");
      put(1, g + " * " + v + " * log(" + A(0) + ")");
      break;
    case MOL_OP_MIN:
      put(0, "(" + A(0) + " <= " + A(1) + " ? " + g + " : T(0))");
      put(1, "(" + A(0) + " <= " + A(1) + " ? T(0) : " + g + ")");
      break;
    case MOL_OP_MAX:
      put(0, "(" + A(0) + " >= " + A(1) + " ? " + g + " : T(0))");
      put(1, "(" + A(0) + " >= " + A(1) + " ? T(0) : " + g + ")");
      break;
    case X_SELECT:
      put(1, "(" + A(0) + " != T(0) ? " + g + " : T(0))");
      put(2, "(" + A(0) + " != T(0) ? T(0) : " + g + ")");
      break;
    case MOL_OP_NEG: put(0, "(-" + g + ")"); break;
    case MOL_OP_SQ: put(0, "T(2) * " + A(0) + " * " + g); break;
    case MOL_OP_INV: put(0, "(-" + g + " * " + v + " * " + v + ")"); break;
    case MOL_OP_ABS: put(0, g + " * mol_sign(" + A(0) + ")"); break;
    case MOL_OP_SQRT: put(0, g + " * T(0.5) / " + v); break;
    case MOL_OP_EXP: put(0, g + " * " + v); break;
    case MOL_OP_LOG: put(0, g + " / " + A(0)); break;
    case MOL_OP_SIN: put(0, g + " * cos(" + A(0) + ")"); break;
    case MOL_OP_COS: put(0, "(-" + g + " * sin(" + A(0) + "))"); break;
    case MOL_OP_TAN: put(0, g + " * (T(1) + " + v + " * " + v + ")"); break;
    case MOL_OP_TANH: put(0, g + " * (T(1) - " + v + " * " + v + ")"); break;
    case MOL_OP_SINH: put(0, g + " * cosh(" + A(0) + ")"); break;
    case MOL_OP_COSH: put(0, g + " * sinh(" + A(0) + ")"); break;
    case MOL_OP_ATAN: put(0, g + " / (T(1) + " + A(0) + " * " + A(0) + ")"); break;
    case MOL_OP_ASIN: put(0, g + " / sqrt(T(1) - " + A(0) + " * " + A(0) + ")"); break;
    case MOL_OP_ACOS: put(0, "(-" + g + " / sqrt(T(1) - " + A(0) + " * " + A(0) + "))"); break;
    case MOL_OP_ERF: put(0, g + " * T(0x1.20dd750429b6dp+0) * exp(-" + A(0) + " * " + A(0) + ")"); break;
    case MOL_OP_LOG10: put(0, g + " / (" + A(0) + " * T(0x1.26bb1bbb55516p+1))"); break;
    case MOL_OP_CBRT: put(0, g + " / (T(3) * " + v + " * " + v + ")"); break;
    case MOL_OP_EXPM1: put(0, g + " * (" + v + " + T(1))"); break;
    case MOL_OP_LOG1P: put(0, g + " / (T(1) + " + A(0) + ")"); break;
    case MOL_OP_POWI:
      if (n.imm != 0) put(0, g + " * T(" + istr(n.imm) + ") * " + powi_expr(A(0), n.imm - 1));
      break;
    case X_LINCOMB:
      for (size_t k = 0; k < n.a.size(); ++k) put((int)k, weight(n, k) + " * " + g);
      break;
    case X_WENO:
    case X_WENONU:
      for (int k = 0; k < 5; ++k) put(k, g + " * wd" + istr(id) + "[" + istr(k) + "]");
      break;
    default: break;   // comparisons, sign, floor, ...: zero derivative
  }
}

// Reverse sweep of region r: `seeds` are (node, adjoint expression) pairs of its outputs.
// Adjoints of LDS leaves are summed per position and added to the LDS adjoint buffers
// (Ab for variables, Aa in FP64 for auxiliary arrays) with LDS atomics.
void Gen::print_reverse(std::ostream& o, int r, const std::vector<std::pair<int, std::string>>& seeds) const {
  const Region& R = regs_[r];
  std::map<std::array<int, 4>, std::string> acc;   // (array, dx, dy, dz) -> adjoint terms
  for (int id : R.nodes) o << "    T g" << id << " = T(0);\n";
  auto add_adj = [&](int target, const std::string& e) {
    const Node& t = nodes_[target];
    if (t.kind == N_OP || t.kind == N_BRANCH) {
      o << "    g" << target << " += " << e << ";\n";
    } else if (t.kind == N_LDS) {
      std::string& s = acc[{t.arr, t.dv[0], t.dv[1], t.dv[2]}];
      s = s.empty() ? e : s + " + " + e;
    }
  };
  for (const auto& sd : seeds) add_adj(sd.first, sd.second);
  for (auto it = R.nodes.rbegin(); it != R.nodes.rend(); ++it) {
    const int id = *it;
    const Node& n = nodes_[id];
    if (n.kind == N_OP) {
      if (n.op == X_WENO) {
        o << "    T wd" << id << "[5];\n    mol_weno5_grad(" << ref(n.a[0]) << ", " << ref(n.a[1]) << ", " << ref(n.a[2])
          << ", " << ref(n.a[3]) << ", " << ref(n.a[4]) << ", " << pk(n.w[0], d_->ncpool + 2LL * n.imm) << ", "
          << pk(n.w[1], d_->ncpool + 2LL * n.imm + 1) << ", wd" << id
          << ", " << (d_->dtype == MOL_F64 ? "true" : "false") << ");\n";
      }
      if (n.op == X_WENONU) {
        o << "    T wd" << id << "[5];\n    mol_weno5_nu_tab_grad(";
        for (int q = 0; q < 5; ++q) o << ref(n.a[q]) << ", ";
        o << wrow(n) << ", " << pk(n.w[0], d_->ncpool + 2LL * n.imm) << ", wd" << id << ");\n";
      }
      std::vector<std::pair<int, std::string>> ps;
      partials(n, id, ps);
      for (const auto& p : ps) add_adj(p.first, p.second);
      continue;
    }
    // branch: the region of the taken case runs forward again, then backwards
    const std::string I = "I" + istr(n.ax), g = "g" + istr(id);
    o << "    if (" << I << " >= " << IX(n.ax, n.lo) << " && " << I << " <= " << IX(n.ax, n.hi) << ") {\n";
    print_region(o, n.reg[0]);
    print_reverse(o, n.reg[0], {{regs_[n.reg[0]].result, g}});
    o << "    }";
    for (size_t k = 1; k < n.reg.size(); ++k) {
      o << " else if (" << I << " == " << IX(n.ax, n.cval[k]) << ") {\n";
      print_region(o, n.reg[k]);
      print_reverse(o, n.reg[k], {{regs_[n.reg[k]].result, g}});
      o << "    }";
    }
    o << "\n";
  }
  std::ostringstream aux_adds;
  for (const auto& kv : acc) {
    const int arr = kv.first[0];
    const bool field = arr < d_->nfields;
    const int sl = field ? pg_.fslot[arr] : pg_.aslot[arr - d_->nfields];
    if (sl < 0) {
      // array read at the node only: its adjoint goes straight to global memory (this node is
      // the only writer of its position in a launch)
      if (field)
        o << "    outp[" << UIDX(arr, "I0", "I1", "I2") << "] += " << kv.second << ";\n";
      else
        o << "    if (gaux) gaux[" << (arr - d_->nfields) << " * NGRID + gidx(I0, I1, I2)] += (double)(" << kv.second
          << ");\n";
      continue;
    }
    const long long off = kv.first[1] + (long long)pg_.ext[0] * (kv.first[2] + (long long)pg_.ext[1] * kv.first[3]);
    const std::string pos = off == 0 ? "b" : (off > 0 ? "b + " + istr(off) : "b - " + istr(-off));
    if (field)
      o << "    atomicAdd(&Ab[" << (long long)sl * pg_.vol << " + " << pos << "], (double)(" << kv.second << "));\n";
    else
      aux_adds << "      atomicAdd(&Aa[" << (long long)sl * pg_.vol << " + " << pos << "], (double)(" << kv.second
               << "));\n";
  }
  // grid-field adjoints only when the caller asks for them (a uniform branch)
  if (!aux_adds.str().empty()) o << "    if (gaux) {\n" << aux_adds.str() << "    }\n";
}

// Slow staging of one position: ring wrap, zero outside the grid and at corners, the unknown
// inside the variable's box, else the ghost-node closure of the axis it leaves the box along
// (field_at / closure_value of the interpreter; src/discretization/generate_bc_eqs.jl).
std::string Gen::stage_fn(bool lds) {
  std::ostringstream o;
  // interfaces between variables: the ghost of one variable is the other variable at the same
  // position, staged by stage_ni (the same function with interface sides returning 0: the
  // other variable's far side is never an interface back)
  bool iface = false;
  for (int c = 0; c < d_->nclosures; ++c) iface = iface || d_->closures_h[c].src >= 0;
  if (iface && !lds && !stage_noif_) {
    stage_noif_ = true;
    o << stage_fn(false);
    stage_noif_ = false;
  }
  // lds: the same values from the stage values in LDS (tile + halo with first position o0, o1,
  // o2), positions unwrapped; only non-periodic axes have ghosts
  auto rd = [&](int f, const std::string& q0, const std::string& q1, const std::string& q2) {
    if (!lds) return "u[" + UIDX(f, q0, q1, q2) + "]";
    return "L[" + istr(f) + " * VOL + ((" + q0 + ") - o0) + EX * (((" + q1 + ") - o1) + EY * ((" + q2 + ") - o2))]";
  };
  if (lds)
    o << "  static __device__ T stage_lds(const int f, int p0, int p1, int p2, const double t, const T* __restrict__ P,\n"
         "                               const T* __restrict__ L, const int o0, const int o1, const int o2,\n"
         "                               const JPtrs& jp) {\n"
         "    (void)t; (void)P; (void)jp;\n";
  else
    o << "  static __device__ T " << (stage_noif_ ? "stage_ni(" : "stage(") << (param_ ? "const BInst& hd, " : "")
      << "const int f, int p0, int p1, int p2, const double t, const T* __restrict__ P,\n"
         "                           const T* __restrict__ u, const JPtrs& jp) {\n"
         "    (void)t; (void)P; (void)jp;\n";
  for (int a = 0; a < 3; ++a)
    if (g_.periodic[a] && !lds) o << "    p" << a << " = mol_wrap(p" << a << ", " << NN(a) << ");\n";
  {
    std::string cond;
    for (int a = 0; a < 3; ++a) {
      if (lds && g_.periodic[a]) continue;
      const std::string c = "p" + istr(a) + " < 1 || p" + istr(a) + " > " + NN(a);
      cond = cond.empty() ? c : cond + " || " + c;
    }
    if (!cond.empty()) o << "    if (" << cond << ") return T(0);\n";
  }
  o << "    switch (f) {\n";
  for (int f = 0; f < d_->nfields; ++f) {
    o << "    case " << f << ": {\n";
    o << "    const int nout = 0";
    for (int a = 0; a < 3; ++a)
      if (!(lds && g_.periodic[a]))
        o << " + (p" << a << " < " << LO(f, a) << " || p" << a << " > " << HI(f, a) << ")";
    o << ";\n    if (nout == 0) return " << rd(f, "p0", "p1", "p2") << ";\n    if (nout > 1) return T(0);\n";
    for (int a = 0; a < d_->ndim; ++a)
      for (int side = 0; side < 2; ++side) {
        if (lds && g_.periodic[a]) continue;
        int cid = -1;
        for (int c = 0; c < d_->nclosures; ++c) {
          const mol_closure& cl = d_->closures_h[c];
          if (cl.field == f && cl.axis == a && cl.side == side) cid = c;
        }
        const std::string pa = "p" + istr(a);
        o << "    if (" << pa << (side == 0 ? " < " + LO(f, a) : " > " + HI(f, a)) << ") {\n";
        if (cid < 0) {
          o << "    return T(0);\n    }\n";
          continue;
        }
        const mol_closure& cl = d_->closures_h[cid];
        if (cl.src >= 0) {
          if (stage_noif_ || lds) o << "    return T(0);\n    }\n";
          else o << "    return stage_ni(" << cl.src << ", p0, p1, p2, t, P, u, jp);\n    }\n";
          continue;
        }
        std::string G;
        if (cl.gprog >= 0) {
          const int r = new_region();
          cur_ = r;
          Ctx bc;
          bc.kind = CTX_BOUNDARY;
          bc.ax = a;
          bc.side = side;
          regs_[r].result = prog(cl.gprog, bc);
          print_region(o, r);
          G = ref(regs_[r].result);
        }
        o << "    const int g = " << (side == 0 ? "(" + LO(f, a) + " - 1) - " + pa : pa + " - (" + HI(f, a) + " + 1)")
          << ";\n";
        for (int gi = 0; gi < cl.nghost; ++gi) {
          std::string e;
          for (int m = 0; m < cl.K; ++m) {
            const double c = d_->pool_h[cl.wC + gi * cl.K + m];
            if (c == 0.0 && !param_) continue;
            std::string q[3] = {"p0", "p1", "p2"};
            q[a] = closure_pos(f, a, side, m);
            const std::string term = pw(c, cl.wC + gi * cl.K + m) + " * " + rd(f, q[0], q[1], q[2]);
            e = e.empty() ? term : e + " + " + term;
          }
          if (cl.gprog >= 0) {
            const std::string term = pw(d_->pool_h[cl.wc + gi], cl.wc + gi) + " * " + G;
            e = e.empty() ? term : e + " + " + term;
          }
          o << "    if (g == " << gi << ") return " << (e.empty() ? "T(0)" : e) << ";\n";
        }
        o << "    return T(0);\n    }\n";
      }
    o << "    return T(0);\n    }\n";
  }
  o << "    default: return T(0);\n    }\n  }\n";
  return o.str();
}

// Transpose of the staging map for one LDS position holding adjoint v of variable f: the
// unknown itself, or the unknowns its ghost-node closure combines (c * G carries no variable).
std::string Gen::scatter_fn() {
  std::ostringstream o;
  o << "  static __device__ void scatter(const int f, int p0, int p1, int p2, const T v, T* __restrict__ out) {\n";
  for (int a = 0; a < 3; ++a)
    if (g_.periodic[a]) o << "    p" << a << " = mol_wrap(p" << a << ", " << NN(a) << ");\n";
  o << "    if (p0 < 1 || p0 > " << NN(0) << " || p1 < 1 || p1 > " << NN(1) << " || p2 < 1 || p2 > " << NN(2)
    << ") return;\n";
  o << "    switch (f) {\n";
  for (int f = 0; f < d_->nfields; ++f) {
    o << "    case " << f << ": {\n";
    o << "    const int nout = ";
    for (int a = 0; a < 3; ++a)
      o << (a ? " + " : "") << "(p" << a << " < " << LO(f, a) << " || p" << a << " > " << HI(f, a) << ")";
    o << ";\n    if (nout == 0) { mol_atomic_add(&out[" << UIDX(f, "p0", "p1", "p2") << "], v); return; }\n"
      << "    if (nout > 1) return;\n";
    for (int a = 0; a < d_->ndim; ++a)
      for (int side = 0; side < 2; ++side) {
        int cid = -1;
        for (int c = 0; c < d_->nclosures; ++c) {
          const mol_closure& cl = d_->closures_h[c];
          if (cl.field == f && cl.axis == a && cl.side == side) cid = c;
        }
        if (cid < 0) continue;
        const mol_closure& cl = d_->closures_h[cid];
        const std::string pa = "p" + istr(a);
        o << "    if (" << pa << (side == 0 ? " < " + LO(f, a) : " > " + HI(f, a)) << ") {\n";
        o << "    const int g = " << (side == 0 ? "(" + LO(f, a) + " - 1) - " + pa : pa + " - (" + HI(f, a) + " + 1)")
          << ";\n";
        for (int gi = 0; gi < cl.nghost; ++gi) {
          o << "    if (g == " << gi << ") {\n";
          for (int m = 0; m < cl.K; ++m) {
            const double c = d_->pool_h[cl.wC + gi * cl.K + m];
            if (c == 0.0 && !param_) continue;
            std::string q[3] = {"p0", "p1", "p2"};
            q[a] = closure_pos(f, a, side, m);
            o << "      mol_atomic_add(&out[" << UIDX(f, q[0], q[1], q[2]) << "], " << pw(c, cl.wC + gi * cl.K + m)
              << " * v);\n";
          }
          o << "      return;\n    }\n";
        }
        o << "    return;\n    }\n";
      }
    o << "    return;\n    }\n";
  }
  o << "    default: return;\n    }\n  }\n";
  return o.str();
}

// Geometry-independent part of the generated structs: sizes, boxes, index helpers and the
// slow staging / scatter functions.
std::string Gen::geo_common() {
  std::ostringstream o;
  auto c = [&](const std::string& name, long long v) { o << "  static constexpr int " << name << " = " << v << ";\n"; };
  c("NF", d_->nfields);
  c("NAUX", d_->naux);
  for (int a = 0; a < 3; ++a) {
    const std::string s = istr(a);
    c("N" + s, g_.n[a]);
    c("PER" + s, g_.periodic[a]);
    c("ULO" + s, g_.ulo[a]);
    c("UHI" + s, g_.uhi[a]);
    c("NT" + s, g_.ntile[a]);
  }
  c("TX", g_.tile[0]); c("TY", g_.tile[1]); c("TZ", g_.tile[2]);
  c("NBLK", g_.ntile[0] * g_.ntile[1] * g_.ntile[2]);
  c("NTHR", g_.nthr);
  // register blocking: each thread runs RB consecutive nodes along axis BA (the slowest tiled
  // axis), so taps and subexpressions shared by neighbours along BA are loaded and computed once
  int ba = 0, rb = 1;
  blocking(ba, rb);
  c("BA", ba);
  c("RB", rb);
  o << "  static constexpr i64 NGRID = " << (long long)g_.n[0] * g_.n[1] * g_.n[2] << ";\n";
  // 32-bit index arithmetic unless the state vector or the grid reaches 2^31 entries
  o << "  using ix = " << (wide() ? "i64" : "int") << ";\n";
  // core tiles: interior operator rows for every node
  for (int a = 0; a < 3; ++a) {
    c("CLO" + istr(a), g_.ncore > 0 ? g_.ct_lo[a] : 1);
    c("CHI" + istr(a), g_.ncore > 0 ? g_.ct_lo[a] + g_.ct_n[a] - 1 : 0);
  }
  o << "  static __device__ __forceinline__ bool is_core(const int tx, const int ty, const int tz) {\n"
       "    return tx >= CLO0 && tx <= CHI0 && ty >= CLO1 && ty <= CHI1 && tz >= CLO2 && tz <= CHI2;\n  }\n";
  o << "  static __device__ __forceinline__ ix uidx(const int f, const int p0, const int p1, const int p2) {\n"
       "    switch (f) {\n";
  for (int f = 0; f < d_->nfields; ++f) {
    const long long bx = g_.hi[f][0] - g_.lo[f][0] + 1, by = g_.hi[f][1] - g_.lo[f][1] + 1;
    o << "    case " << f << ": return (ix)" << g_.fofs[f] << " + (p0 - " << g_.lo[f][0] << ") + (ix)" << bx
      << " * ((p1 - " << g_.lo[f][1] << ") + (ix)" << by << " * (p2 - " << g_.lo[f][2] << "));\n";
  }
  o << "    default: return 0;\n    }\n  }\n";
  o << "  static __device__ __forceinline__ bool inbox(const int f, const int I0, const int I1, const int I2) {\n"
       "    switch (f) {\n";
  for (int f = 0; f < d_->nfields; ++f)
    o << "    case " << f << ": return I0 >= " << g_.lo[f][0] << " && I0 <= " << g_.hi[f][0] << " && I1 >= "
      << g_.lo[f][1] << " && I1 <= " << g_.hi[f][1] << " && I2 >= " << g_.lo[f][2] << " && I2 <= " << g_.hi[f][2]
      << ";\n";
  o << "    default: return false;\n    }\n  }\n";
  // full-grid index of an (unwrapped) position, -1 outside the grid
  o << "  static __device__ __forceinline__ ix gidx(int p0, int p1, int p2) {\n";
  for (int a = 0; a < 3; ++a)
    if (g_.periodic[a]) o << "    p" << a << " = mol_wrap(mol_wrap(p" << a << ", " << g_.n[a] << "), " << g_.n[a] << ");\n";
  o << "    if (p0 < 1 || p0 > " << g_.n[0] << " || p1 < 1 || p1 > " << g_.n[1] << " || p2 < 1 || p2 > " << g_.n[2]
    << ") return -1;\n"
    << "    return gidx_in(p0, p1, p2);\n  }\n";
  // full-grid index of a position inside the grid (no wrap, no check)
  o << "  static __device__ __forceinline__ ix gidx_in(const int p0, const int p1, const int p2) {\n"
    << "    return (ix)(p0 - 1) + (ix)" << g_.n[0] << " * ((p1 - 1) + (ix)" << g_.n[1] << " * (p2 - 1));\n  }\n";
  o << "  static __device__ __forceinline__ T stage_aux(const int a, int p0, int p1, int p2, const JPtrs& jp) {\n"
       "    const ix gi = gidx(p0, p1, p2);\n"
       "    return gi < 0 ? T(0) : jp.aux[a][gi];\n  }\n";
  return o.str();
}

// Print geometry of a tile variant with halo `halo`: staged arrays get consecutive LDS slots.
Gen::PrintGeo Gen::make_pg(const int* halo, const Usage& us) const {
  PrintGeo pg;
  pg.vol = 1;
  for (int a = 0; a < 3; ++a) {
    pg.ext[a] = g_.tile[a] + 2 * halo[a];
    pg.vol *= pg.ext[a];
  }
  const int narr = d_->nfields + d_->naux;
  pg.slot.assign(narr, -1);
  pg.fslot.assign(d_->nfields, -1);
  pg.aslot.assign(d_->naux, -1);
  int ns = 0, nf = 0, na = 0;
  for (int k = 0; k < narr; ++k) {
    if (!us.staged[k]) continue;
    pg.slot[k] = ns++;
    if (k < d_->nfields) pg.fslot[k] = nf++;
    else pg.aslot[k - d_->nfields] = na++;
  }
  return pg;
}

// One tile variant: geometry constants, staging of one position, node function (forward or
// reverse mode) and, for reverse mode, the flush of one position's adjoints.
std::vector<int> Gen::invariant_frontier(int r, const std::vector<int>& outs, std::vector<char>& inv) const {
  inv.assign(nodes_.size(), 0);
  for (size_t id = 0; id < nodes_.size(); ++id) {
    const Node& n = nodes_[id];
    if (n.kind == N_LIT) inv[id] = 1;
    else if (n.kind == N_EXPR) inv[id] = n.text != "(T)t";
  }
  // operations of the top-level region in evaluation order (operands come first); nodes of
  // branch regions stay in place (boundary rows)
  for (int id : regs_[r].nodes) {
    const Node& n = nodes_[id];
    if (n.kind != N_OP) continue;
    bool all = true;
    for (int a : n.a) all = all && inv[a];
    inv[id] = all;
  }
  std::vector<char> used(nodes_.size(), 0);
  std::vector<int> stack{r};
  while (!stack.empty()) {
    const int q = stack.back();
    stack.pop_back();
    for (int id : regs_[q].nodes) {
      const Node& n = nodes_[id];
      if (n.kind == N_BRANCH) {
        for (int sr : n.reg) {
          stack.push_back(sr);
          if (regs_[sr].result >= 0) used[regs_[sr].result] = 1;
        }
        continue;
      }
      if (inv[id] && std::find(regs_[r].nodes.begin(), regs_[r].nodes.end(), id) != regs_[r].nodes.end()) continue;
      for (int a : n.a) used[a] = 1;
    }
  }
  for (int f : outs) used[f] = 1;
  // hoist the top-level invariant operations and the coordinate reads a dependent node uses
  // (literals cost nothing and parameters are uniform)
  std::vector<int> front;
  for (size_t id = 0; id < nodes_.size(); ++id) {
    if (!used[id] || !inv[id]) continue;
    const Node& n = nodes_[id];
    if (n.kind == N_LIT) continue;
    if (n.kind == N_EXPR && n.text.compare(0, 2, "P[") == 0) continue;
    if (n.kind == N_OP && std::find(regs_[r].nodes.begin(), regs_[r].nodes.end(), (int)id) == regs_[r].nodes.end())
      continue;
    front.push_back((int)id);
  }
  return front;
}

// node_inv (the frontier values of a node, computed once) and node_h (the node function reading
// them) of the fused step's variant; NINV = 0 when nothing is worth hoisting.
std::string Gen::node_hoisted(int region, const std::vector<int>& outs) {
  std::vector<char> inv;
  const std::vector<int> front = invariant_frontier(region, outs, inv);
  std::ostringstream o;
  o << "  static constexpr int NINV = " << front.size() << ";\n";
  if (front.empty()) return o.str();
  o << "  static __device__ __forceinline__ void node_inv(const int I0, const int I1, const int I2,\n"
       "      const T* __restrict__ P, const JPtrs& jp, T* inv) {\n"
       "    (void)I0; (void)I1; (void)I2; (void)P; (void)jp;\n";
  for (int id : regs_[region].nodes)
    if (nodes_[id].kind == N_OP && inv[id]) print_node(o, id);
  for (size_t k = 0; k < front.size(); ++k) o << "    inv[" << k << "] = " << ref(front[k]) << ";\n";
  o << "  }\n";
  o << "  static __device__ __forceinline__ void node_h(const T* __restrict__ L, const int b, const int I0, const int I1,\n"
       "      const int I2, const double t, const T* __restrict__ P, const T* __restrict__ u, const JPtrs& jp,\n"
       "      const T* inv, T* out) {\n"
       "    (void)I0; (void)I1; (void)I2; (void)t; (void)P; (void)jp; (void)L; (void)b; (void)u; (void)inv;\n";
  skip_.assign(nodes_.size(), 0);
  for (int id : regs_[region].nodes)
    if (nodes_[id].kind == N_OP && inv[id]) skip_[id] = 1;
  for (size_t k = 0; k < front.size(); ++k) subst_[front[k]] = "inv[" + istr((long long)k) + "]";
  print_region(o, region);
  for (int f = 0; f < d_->nfields; ++f) o << "    out[" << f << "] = " << ref(outs[f]) << ";\n";
  skip_.clear();
  subst_.clear();
  o << "  }\n";
  return o.str();
}

std::string Gen::geo_variant(const std::string& name, const int* halo, const Usage& us, bool core, int region,
                             const std::vector<int>& outs, bool vjp) {
  pg_ = make_pg(halo, us);
  std::ostringstream o;
  int nst = 0, nstf = 0, nsta = 0;
  for (int k = 0; k < d_->nfields + d_->naux; ++k)
    if (pg_.slot[k] >= 0) (k < d_->nfields ? nstf : nsta)++, ++nst;
  o << "struct " << name << " : GeoBase {\n";
  auto c = [&](const std::string& nm, long long v) { o << "  static constexpr int " << nm << " = " << v << ";\n"; };
  c("CORE", core ? 1 : 0);
  c("HX", halo[0]); c("HY", halo[1]); c("HZ", halo[2]);
  c("EX", pg_.ext[0]); c("EY", pg_.ext[1]); c("EZ", pg_.ext[2]);
  c("VOL", pg_.vol);
  c("NST", nst); c("NSTF", nstf); c("NSTA", nsta);
  // no diagonal taps: halo positions outside the tile along two or more axes are never read
  // and are not staged (3D, where they are 27-40 % of the box; not for the fused step, whose
  // stage regions grow along every axis)
  o << "  static constexpr bool CROSS = " << (us.separable && !fused_ && !param_ && d_->ndim >= 3 ? "true" : "false")
    << ";\n";
  // tiles staged with plain loads: tile + halo inside every variable's box (after the wrap)
  for (int a = 0; a < 3; ++a) {
    int slo = 0, shi = g_.ntile[a] - 1;
    if (!g_.periodic[a]) {
      int lo = -(1 << 30), hi = 1 << 30;
      for (int f = 0; f < d_->nfields; ++f) {
        lo = std::max(lo, g_.lo[f][a]);
        hi = std::min(hi, g_.hi[f][a]);
      }
      slo = g_.ntile[a];
      shi = -1;
      for (int t = 0; t < g_.ntile[a]; ++t) {
        const int o0 = g_.ulo[a] + t * g_.tile[a];
        if (o0 - halo[a] >= lo && o0 + g_.tile[a] - 1 + halo[a] <= hi) {
          slo = std::min(slo, t);
          shi = std::max(shi, t);
        }
      }
    }
    c("SLO" + istr(a), slo);
    c("SHI" + istr(a), shi);
  }
  // tiles staged without the periodic wrap: tile + halo inside every variable's box unwrapped
  for (int a = 0; a < 3; ++a) {
    int lo = -(1 << 30), hi = 1 << 30;
    for (int f = 0; f < d_->nfields; ++f) {
      lo = std::max(lo, g_.lo[f][a]);
      hi = std::min(hi, g_.hi[f][a]);
    }
    int flo = g_.ntile[a], fhi = -1;
    for (int t = 0; t < g_.ntile[a]; ++t) {
      const int o0 = g_.ulo[a] + t * g_.tile[a];
      if (o0 - halo[a] >= lo && o0 + g_.tile[a] - 1 + halo[a] <= hi) {
        flo = std::min(flo, t);
        fhi = std::max(fhi, t);
      }
    }
    c("FLO" + istr(a), flo);
    c("FHI" + istr(a), fhi);
  }
  // values of the staged arrays at one tile + halo position (p: unwrapped full-grid index),
  // v[slot]; mode 0: plain loads, 1: plain loads after the periodic wrap, 2: through the ghost
  // closures
  o << "  static __device__ __forceinline__ void stage_vals(T* v, int p0, int p1, int p2, const int mode,\n"
       "      const double t, const T* __restrict__ P, const T* __restrict__ u, const JPtrs& jp) {\n"
       "    (void)t; (void)P; (void)jp; (void)u; (void)v;\n";
  o << "    if (mode == 2) {\n";
  for (int f = 0; f < d_->nfields; ++f)
    if (pg_.slot[f] >= 0) o << "      v[" << pg_.slot[f] << "] = stage(" << f << ", p0, p1, p2, t, P, u, jp);\n";
  for (int a = 0; a < d_->naux; ++a)
    if (pg_.slot[d_->nfields + a] >= 0)
      o << "      v[" << pg_.slot[d_->nfields + a] << "] = stage_aux(" << a << ", p0, p1, p2, jp);\n";
  o << "      return;\n    }\n";
  bool any_per = false;
  for (int a = 0; a < 3; ++a) any_per = any_per || g_.periodic[a];
  if (any_per) {
    o << "    if (mode == 1) {\n";
    for (int a = 0; a < 3; ++a)
      if (g_.periodic[a])
        o << "      p" << a << " = mol_wrap(mol_wrap(p" << a << ", " << g_.n[a] << "), " << g_.n[a] << ");\n";
    o << "    }\n";
  }
  for (int f = 0; f < d_->nfields; ++f)
    if (pg_.slot[f] >= 0) o << "    v[" << pg_.slot[f] << "] = u[uidx(" << f << ", p0, p1, p2)];\n";
  for (int a = 0; a < d_->naux; ++a)
    if (pg_.slot[d_->nfields + a] >= 0)
      o << "    v[" << pg_.slot[d_->nfields + a] << "] = jp.aux[" << a << "][gidx_in(p0, p1, p2)];\n";
  o << "  }\n";
  o << "  static __device__ __forceinline__ void stage_pos(T* L, const int q, int p0, int p1, int p2, const int mode,\n"
       "      const double t, const T* __restrict__ P, const T* __restrict__ u, const JPtrs& jp) {\n"
       "    T v[NST > 0 ? NST : 1];\n"
       "    stage_vals(v, p0, p1, p2, mode, t, P, u, jp);\n"
       "#pragma unroll\n"
       "    for (int s = 0; s < NST; ++s) L[s * VOL + q] = v[s];\n"
       "  }\n";
  if (!vjp) {
    o << "  static __device__ __forceinline__ void node(const T* __restrict__ L, const int b, const int I0, const int I1,\n"
         "      const int I2, const double t, const T* __restrict__ P, const T* __restrict__ u, const JPtrs& jp, T* out) {\n"
         "    (void)I0; (void)I1; (void)I2; (void)t; (void)P; (void)jp; (void)L; (void)b; (void)u;\n";
    print_region(o, region);
    for (int f = 0; f < d_->nfields; ++f) o << "    out[" << f << "] = " << ref(outs[f]) << ";\n";
    o << "  }\n";
    if (fused_) o << node_hoisted(region, outs);
  } else {
    o << "  static __device__ __forceinline__ void zero_pos(double* Ab, double* Aa, const int q) {\n"
         "    (void)Ab; (void)Aa;\n";
    for (int k = 0; k < nstf; ++k) o << "    Ab[" << (long long)k * pg_.vol << " + q] = 0.0;\n";
    for (int k = 0; k < nsta; ++k) o << "    Aa[" << (long long)k * pg_.vol << " + q] = 0.0;\n";
    o << "  }\n";
    o << "  static __device__ __forceinline__ void vjp(const T* __restrict__ L, double* Ab, double* Aa, const int b,\n"
         "      const int I0, const int I1, const int I2, const double t, const T* __restrict__ P,\n"
         "      const T* __restrict__ u, const JPtrs& jp, const T* lam, T* __restrict__ outp, double* __restrict__ gaux) {\n"
         "    (void)I0; (void)I1; (void)I2; (void)t; (void)P; (void)jp; (void)L; (void)b; (void)Ab; (void)Aa; (void)u;\n"
         "    (void)outp; (void)gaux;\n";
    print_region(o, region);
    std::vector<std::pair<int, std::string>> seeds;
    for (int f = 0; f < d_->nfields; ++f) seeds.push_back({outs[f], "lam[" + istr(f) + "]"});
    print_reverse(o, region, seeds);
    o << "  }\n";
    // adjoints of one tile + halo position: positions of a variable's unknown box (after the
    // periodic wrap) and of the grid go to the tile's buffers (slot stride vs; zero elsewhere),
    // ghost positions through the transposed closures to the unknowns (atomics)
    o << "  static __device__ __forceinline__ void flush_pos(const double* Ab, const double* Aa, const int q, const int p0,\n"
         "      const int p1, const int p2, T* __restrict__ outp, T* __restrict__ bft, double* __restrict__ bat,\n"
         "      const int vs, const bool acc) {\n"
         "    (void)Ab; (void)Aa; (void)bat; (void)vs; (void)outp; (void)acc;\n";
    if (nstf > 0) {
      o << "    int w0 = p0, w1 = p1, w2 = p2;\n";
      for (int a = 0; a < 3; ++a)
        if (g_.periodic[a])
          o << "    w" << a << " = mol_wrap(mol_wrap(w" << a << ", " << g_.n[a] << "), " << g_.n[a] << ");\n";
    }
    for (int f = 0; f < d_->nfields; ++f)
      if (pg_.fslot[f] >= 0) {
        const long long k = pg_.fslot[f];
        if (gather_stores()) {
          // the gather applies the transposed closures to the ghost positions' buffer entries
          o << "    bft[" << k << " * vs + q] = (T)Ab[" << k * pg_.vol << " + q];\n";
          continue;
        }
        o << "    { const T v = (T)Ab[" << k * pg_.vol << " + q];\n"
          << "      if (inbox(" << f << ", w0, w1, w2)) bft[" << k << " * vs + q] = v;\n"
          << "      else { bft[" << k << " * vs + q] = T(0); if (v != T(0)) scatter(" << f << ", p0, p1, p2, v, outp); } }\n";
      }
    if (nsta > 0) {
      o << "    if (bat) {\n      const bool in = gidx(p0, p1, p2) >= 0;\n";
      for (int a = 0; a < d_->naux; ++a)
        if (pg_.aslot[a] >= 0)
          o << "      { double* b = &bat[" << pg_.aslot[a] << " * vs + q]; const double v = in ? Aa["
            << (long long)pg_.aslot[a] * pg_.vol << " + q] : 0.0; *b = acc ? *b + v : v; }\n";
      o << "    }\n";
    }
    o << "  }\n";
  }
  o << "};\n\n";
  return o.str();
}

// Forward (vjp = false) or reverse-mode kernel of one system.
std::string Gen::source_kernel(bool f64, bool vjp) {
  std::ostringstream o;
  o << "// generated " << (vjp ? "vector-Jacobian product of the " : "") << "right-hand side of one PDE system\n";
  o << "#define MOL_T " << (f64 ? "double" : "float") << "\n";
  o << "#include \"mol_jit_prelude.h\"\n\n";
  o << layout_asserts(f64);
  const std::string stage = stage_fn();
  const std::string scatter = vjp ? scatter_fn() : "";
  std::vector<int> out_core, out_bound;
  int r_core = -1, r_bound = -1;
  if (g_.ncore > 0) build_node(true, out_core, r_core);
  build_node(false, out_bound, r_bound);
  if (!ok_) return "";
  const Usage us = usage();
  o << "struct GeoBase {\n" << geo_common() << stage << scatter << "};\n\n";
  int lds = 1, ldsf = 1, ldsa = 1, npf = 1;
  auto sizes = [&](const PrintGeo& pg) {
    int ns = 0, nf = 0, na = 0;
    for (size_t k = 0; k < pg.slot.size(); ++k)
      if (pg.slot[k] >= 0) ((int)k < d_->nfields ? nf : na)++, ++ns;
    lds = std::max(lds, ns * pg.vol);
    ldsf = std::max(ldsf, nf * pg.vol);
    ldsa = std::max(ldsa, na * pg.vol);
    npf = std::max(npf, pg.vol);
  };
  if (g_.ncore > 0) {
    o << geo_variant("GeoC", us.core_reach, us, true, r_core, out_core, vjp);
    sizes(pg_);
  }
  o << geo_variant("GeoB", us.reach, us, false, r_bound, out_bound, vjp);
  sizes(pg_);
  o << "constexpr int LDSN = " << lds << ", LDSF = " << ldsf << ", LDSA = " << ldsa << ";\n";
  // staging registers per thread of the persistent loop: positions of the larger tile + halo
  o << "constexpr int NPF = (" << npf << " + GeoBase::NTHR - 1) / GeoBase::NTHR;\n\n";
  const std::string decode =
      "  int b = mol_xcd_tile<GeoBase::NBLK>((int)blockIdx.x);\n"
      "  const int tx = b % GeoBase::NT0;\n  b /= GeoBase::NT0;\n"
      "  const int ty = b % GeoBase::NT1;\n  const int tz = b / GeoBase::NT1;\n";
  int nst = 0;
  for (bool b : us.staged) nst += b ? 1 : 0;
  // (3D tiles run one work-group per tile: their persistent loop measured 53 vs 41 us on 128^3)
  const bool persistent = nst > 0 && d_->ndim <= 2 && ((npf + g_.nthr - 1) / g_.nthr) * nst <= 16;
  if (!vjp && persistent) {
    // persistent: gridDim.x work-groups walk the tiles, the next tile's staging loads in flight
    // during the current tile's node phase (few staging values per thread)
    o << "extern \"C\" __global__ void __launch_bounds__(" << g_.nthr << ") mol_rhs_tiles(const T* __restrict__ u, Epi ep, JPtrs jp) {\n"
         "  __shared__ T L[LDSN];\n"
         "  const double t = (ep.tnow ? ep.tnow[0] : 0.0) + ep.tofs;\n"
         "  mol_advance_clock(ep);\n"
      << "  mol_tiles_persistent<" << (g_.ncore > 0 ? "GeoC" : "GeoB") << ", GeoB, " << (g_.ncore > 0 ? "true" : "false")
      << ", NPF>(L, u, ep, jp, t);\n}\n";
  } else if (!vjp) {
    // one work-group per tile
    o << "extern \"C\" __global__ void __launch_bounds__(" << g_.nthr << ") mol_rhs_tiles(const T* __restrict__ u, Epi ep, JPtrs jp) {\n"
         "  __shared__ T L[LDSN];\n" << decode
      << "  const double t = (ep.tnow ? ep.tnow[0] : 0.0) + ep.tofs;\n"
         "  mol_advance_clock(ep);\n";
    if (g_.ncore > 0)
      o << "  if (GeoBase::is_core(tx, ty, tz)) mol_full_tile<GeoC>(L, tx, ty, tz, u, ep, jp, t);\n  else ";
    else
      o << "  ";
    o << "mol_full_tile<GeoB>(L, tx, ty, tz, u, ep, jp, t);\n}\n";
  } else {
    // tile kernel: reverse sweep of every node, tile + halo adjoints to per-tile buffers
    o << "constexpr int VOLX = " << npf << ";\n\n";
    o << "extern \"C\" __global__ void __launch_bounds__(" << g_.nthr << ") mol_vjp_tiles(VArgs va, JPtrs jp) {\n"
         "  __shared__ T L[LDSN];\n  __shared__ double Ab[LDSF];\n  __shared__ double Aa[LDSA];\n" << decode
      << "  const double t = (va.tnow ? va.tnow[0] : 0.0) + va.tofs;\n";
    if (g_.ncore > 0)
      o << "  if (GeoBase::is_core(tx, ty, tz)) mol_vjp_tile<GeoC, VOLX>(L, Ab, Aa, tx, ty, tz, va, jp, t);\n  else ";
    else
      o << "  ";
    o << "mol_vjp_tile<GeoB, VOLX>(L, Ab, Aa, tx, ty, tz, va, jp, t);\n}\n\n";
    // gather kernel: each unknown of a staged variable (then each grid position of a staged
    // grid field, when g_aux is given) sums the buffers of the tiles whose tile + halo covers it
    const std::string gc = g_.ncore > 0 ? "GeoC" : "GeoB", hasc = g_.ncore > 0 ? "true" : "false";
    long long nunk = 0;
    for (int f = 0; f < d_->nfields; ++f) {
      long long sz = 1;
      for (int a = 0; a < 3; ++a) sz *= g_.hi[f][a] - g_.lo[f][a] + 1;
      nunk = std::max(nunk, (long long)g_.fofs[f] + sz);
    }
    if (gather_stores()) o << "// adjoint gather stores every unknown\n";
    o << "extern \"C\" __global__ void __launch_bounds__(256) mol_vjp_gather(VArgs va) {\n"
         "  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;\n"
         "  (void)va;\n";
    o << "  if (i < " << nunk << "LL) {\n"
         "    if (!(va.mode & 1)) return;\n";
    for (int f = 0; f < d_->nfields; ++f) {
      if (pg_.fslot[f] < 0) continue;
      long long sz = 1;
      for (int a = 0; a < 3; ++a) sz *= g_.hi[f][a] - g_.lo[f][a] + 1;
      const int bx = g_.hi[f][0] - g_.lo[f][0] + 1, by = g_.hi[f][1] - g_.lo[f][1] + 1;
      const std::string gp = "mol_gather_pos<" + gc + ", GeoB, " + hasc + ", VOLX, GeoB::NSTF>(va.bf, " +
                             istr(pg_.fslot[f]) + ", ";
      o << "    if (i >= " << g_.fofs[f] << "LL && i < " << (long long)g_.fofs[f] + sz << "LL) {\n"
        << "      const int r = (int)(i - " << g_.fofs[f] << "LL), r1 = r / " << bx << ";\n"
        << "      const int p0 = " << g_.lo[f][0] << " + (r - r1 * " << bx << "), p1 = " << g_.lo[f][1] << " + r1 % " << by
        << ", p2 = " << g_.lo[f][2] << " + r1 / " << by << ";\n"
        << "      double s = " << gp << "p0, p1, p2);\n";
      if (gather_stores()) {
        // transposed ghost closures: the unknown at source m of a closure receives
        // sum_g w[g][m] * (adjoint of ghost g at the same position along the other axes)
        for (int c = 0; c < d_->nclosures; ++c) {
          const mol_closure& cl = d_->closures_h[c];
          if (cl.field != f) continue;
          const int a = cl.axis;
          for (int m = 0; m < cl.K; ++m) {
            std::ostringstream terms;
            for (int gi = 0; gi < cl.nghost; ++gi) {
              const double w = d_->pool_h[cl.wC + gi * cl.K + m];
              const int gpos = cl.side == 0 ? g_.lo[f][a] - 1 - gi : g_.hi[f][a] + 1 + gi;
              if (w == 0.0 || gpos < 1 || gpos > g_.n[a]) continue;
              std::string q[3] = {"p0", "p1", "p2"};
              q[a] = istr(gpos);
              terms << "        s += (double)" << lit(w) << " * " << gp << q[0] << ", " << q[1] << ", " << q[2] << ");\n";
            }
            if (terms.str().empty()) continue;
            o << "      if (p" << a << " == " << closure_pos(f, a, cl.side, m) << ") {\n" << terms.str() << "      }\n";
          }
        }
      }
      if (gather_stores())
        // the last stage of an adjoint RK4 step: the step's update of lam (the sum kernel's
        // association, with this stage's adjoint rounded to T first) instead of a store
        o << "      if (va.mode & 8) {\n"
             "        const T y3 = (T)s;\n"
             "        va.lam[i] = (((va.lam[i] + va.y0[i]) + va.y1[i]) + va.y2[i]) + y3;\n"
             "      } else {\n"
             "        va.out[i] = (T)s;\n"
             "      }\n    }\n";
      else
        o << "      va.out[i] += (T)s;\n    }\n";
    }
    o << "    return;\n  }\n";
    bool any_aux = false;
    for (int a = 0; a < d_->naux; ++a) any_aux = any_aux || pg_.aslot[a] >= 0;
    if (any_aux) {
      o << "  if (!va.g_aux || !(va.mode & 2)) return;\n"
        << "  const i64 j = i - " << nunk << "LL;\n"
        << "  if (j >= " << d_->naux << " * GeoBase::NGRID) return;\n"
        << "  const int a = (int)(j / GeoBase::NGRID), g = (int)(j - (i64)a * GeoBase::NGRID);\n"
        << "  const int g1 = g / " << g_.n[0] << ";\n"
        << "  const int p0 = 1 + (g - g1 * " << g_.n[0] << "), p1 = 1 + g1 % " << g_.n[1] << ", p2 = 1 + g1 / " << g_.n[1]
        << ";\n";
      for (int a = 0; a < 3; ++a)
        if (g_.periodic[a]) o << "  if (p" << a << " == 1) return;   // image of node n\n";
      for (int a = 0; a < d_->naux; ++a)
        if (pg_.aslot[a] >= 0)
          o << "  if (a == " << a << ") va.g_aux[j] += mol_gather_pos<" << gc << ", GeoB, " << hasc
            << ", VOLX, GeoB::NSTA>(va.ba, " << pg_.aslot[a] << ", p0, p1, p2);\n";
    }
    o << "}\n";
  }
  bool nu = false;
  for (int k = 0; k < d_->nslots; ++k) nu = nu || d_->slots_h[k].kind == MOL_SLOT_WENO_NU;
  if (!vjp && nu)
    // coefficient rows of the nonuniform WENO5 derivative at every node of an axis: built once
    // per handle, read by the node functions (jp.wtab)
    o << "\nextern \"C\" __global__ void __launch_bounds__(256) mol_wenonu_table(const double* __restrict__ x, int n,\n"
         "                                                                 int periodic, double* __restrict__ out) {\n"
         "  const int i = (int)(blockIdx.x * 256 + threadIdx.x) + 1;\n"
         "  if (i <= n) mol_weno_nu_row(x, n, periodic != 0, i, out + (i64)MOL_NU_NC * (i - 1));\n"
         "}\n";
  return o.str();
}

// Fused RK4 step: the tile + 4 * reach halo of u is staged once; the four stages run on
// regions shrinking by the reach (tile + 3r, + 2r, + r, tile) with the stage states in LDS,
// and only the tile's new state is written.  The update arithmetic is the stage-kernel
// epilogue's, so the step equals four stage launches.
std::string Gen::source_fused(bool f64) {
  if (d_->ndim > 2) { bad("fused RK4: more than two dimensions"); return ""; }
  bool allper = true;
  for (int a = 0; a < d_->ndim; ++a) allper = allper && g_.periodic[a];
  for (int f = 1; f < d_->nfields; ++f)
    for (int a = 0; a < 3; ++a)
      if (g_.lo[f][a] != g_.lo[0][a] || g_.hi[f][a] != g_.hi[0][a]) { bad("fused RK4: unequal unknown boxes"); return ""; }
  std::vector<int> out_core;
  int r_core = -1;
  fused_ = true;
  // non-periodic axes: the node code with boundary rows (branches on the node index); the
  // ghost positions of each stage's values are refilled from their closures (stage_lds)
  build_node(allper, out_core, r_core);
  if (!ok_) return "";
  Usage us = usage();
  for (int k = d_->nfields; k < d_->nfields + d_->naux; ++k)
    if (us.staged[k]) { bad("fused RK4: a grid field read away from the node"); return ""; }
  for (int f = 0; f < d_->nfields; ++f) us.staged[f] = true;
  int halo[3] = {0, 0, 0}, reach[3] = {0, 0, 0};
  for (int a = 0; a < d_->ndim; ++a) {
    // reach 2 only on periodic axes: the ghost refill reads inward from the box edge by the
    // boundary rows' width, which the reach-1 geometry keeps inside the tile
    if (us.reach[a] > 2 || (us.reach[a] > 1 && !g_.periodic[a])) { bad("fused RK4: reach too large"); return ""; }
    reach[a] = us.reach[a];
    halo[a] = 4 * g_.fsteps * us.reach[a];
  }
  std::ostringstream o;
  o << "// generated fused RK4 step of one PDE system\n";
  o << "#define MOL_T " << (f64 ? "double" : "float") << "\n";
  o << "#include \"mol_jit_prelude.h\"\n\n";
  o << layout_asserts(f64);
  // coordinates of the tile + halo (per axis) and the unwrapped index of its first position
  int emax = 1;
  for (int a = 0; a < 3; ++a) emax = std::max(emax, g_.tile[a] + 2 * halo[a]);
  o << "__shared__ double mol_fc[3][" << emax << "];\n__shared__ int mol_fo[3];\n\n";
  const std::string stage = stage_fn();
  o << "struct GeoBase {\n" << geo_common() << stage << "};\n\n";
  std::string var = geo_variant("GeoF", halo, us, allper, r_core, out_core, false);
  // stage regions: the reach per axis
  const std::string close = "};\n\n";
  var = var.substr(0, var.size() - close.size());
  o << var;
  o << "  static constexpr int RX = " << reach[0] << ", RY = " << reach[1] << ", RZ = " << reach[2] << ";\n";
  o << "  static constexpr bool FIX = " << (allper ? "false" : "true") << ";\n";
  // stage regions by L1 distance from the tile: every tap lies along one axis, reach one,
  // periodic axes (no ghost refill, whose closures read along the square region's rows)
  const bool diam = us.separable && allper && d_->ndim == 2 && reach[0] == 1 && reach[1] == 1;
  o << "  static constexpr bool DIAM = " << (diam ? "true" : "false") << ";\n";
  if (!allper) o << stage_fn(true);
  o << close;
  // one 1D tile holding the whole system: nsub resident steps per launch; otherwise one step
  const bool resident = g_.ntile[0] * g_.ntile[1] * g_.ntile[2] == 1 && d_->ndim == 1 && g_.tile[0] >= halo[0];
  o << "extern \"C\" __global__ void __launch_bounds__(" << g_.nthr
    << ") mol_rk4_fused(const T* __restrict__ u, T* __restrict__ unew, Epi ep, JPtrs jp, int nsub) {\n"
       "  __shared__ T U[GeoF::NST * GeoF::VOL];\n  __shared__ T Ya[GeoF::NST * GeoF::VOL];\n"
       "  __shared__ T Yb[GeoF::NST * GeoF::VOL];\n";
  // two steps per launch: the running sums of the first step's frame positions
  if (g_.fsteps == 2 && !resident) o << "  __shared__ T Ac[GeoF::NST * GeoF::VOL];\n";
  else o << "  T* const Ac = nullptr;\n";
  o << "  // the step's time from the device clock, or (first launch of a call) by value in tofs\n"
       "  const double t = ep.tnow ? ep.tnow[0] : ep.tofs;\n";
  if (resident) {
    o << "  mol_rk4_fused_resident<GeoF>(U, Ya, Yb, &mol_fc[0][0], (int)(sizeof(mol_fc[0]) / sizeof(double)), mol_fo,\n"
         "                               u, unew, ep, jp, t, nsub);\n}\n";
  } else {
    // nsub steps (1, or 2 with the 8-reach halo); the clock advances as by nsub launches
    o << "  int b = mol_xcd_tile<GeoBase::NBLK>((int)blockIdx.x);\n"
         "  const int tx = b % GeoBase::NT0;\n  b /= GeoBase::NT0;\n"
         "  const int ty = b % GeoBase::NT1;\n  const int tz = b / GeoBase::NT1;\n"
         "  if (ep.tnext && blockIdx.x == 0 && threadIdx.x == 0) ep.tnext[0] = nsub == 2 ? (t + ep.tadv) + ep.tadv : t + ep.tadv;\n"
         "  mol_rk4_fused_tile<GeoF>(U, Ya, Yb, Ac, &mol_fc[0][0], (int)(sizeof(mol_fc[0]) / sizeof(double)), mol_fo,\n"
         "                           tx, ty, tz, u, unew, ep, jp, t, nsub);\n}\n";
  }
  o << "// resident steps per launch: " << (resident ? "any" : (g_.fsteps == 2 ? "2" : "1")) << "\n";
  return o.str();
}

std::string Gen::source(bool f64) { return source_kernel(f64, false); }
std::string Gen::source_vjp(bool f64) {
  for (int c = 0; c < d_->nclosures; ++c)
    if (d_->closures_h[c].src >= 0) {
      bad("reverse mode across interface boundaries between variables");
      return "";
    }
  return source_kernel(f64, true);
}

// Class body of batched instances: one geometry (the full reach as halo), numbers from `hd`.
std::string Gen::source_class(const std::string& name) {
  param_ = true;
  std::ostringstream o;
  const std::string stage = stage_fn();
  std::vector<int> out_core, out_bound;
  int r_core = -1, r_bound = -1;
  build_node(true, out_core, r_core);
  build_node(false, out_bound, r_bound);
  if (!ok_) return "";
  const Usage us = usage();
  pg_ = make_pg(us.reach, us);
  int nst = 0;
  for (int v : pg_.slot) nst += v >= 0;
  o << "struct " << name << " {\n";
  auto c = [&](const std::string& nm, long long v) { o << "  static constexpr int " << nm << " = " << v << ";\n"; };
  c("NF", d_->nfields);
  c("NST", nst);
  for (int a = 0; a < 3; ++a) c("PER" + istr(a), g_.periodic[a]);
  c("TX", g_.tile[0]); c("TY", g_.tile[1]); c("TZ", g_.tile[2]);
  c("HX", us.reach[0]); c("HY", us.reach[1]); c("HZ", us.reach[2]);
  c("EX", pg_.ext[0]); c("EY", pg_.ext[1]); c("EZ", pg_.ext[2]);
  c("VOL", pg_.vol);
  c("NTHR", 256);
  // the variable staged in each LDS slot
  {
    std::vector<int> sf(std::max(nst, 1), 0);
    for (int f = 0; f < d_->nfields; ++f)
      if (pg_.slot[f] >= 0) sf[pg_.slot[f]] = f;
    o << "  static constexpr int SF[" << sf.size() << "] = {";
    for (size_t k = 0; k < sf.size(); ++k) o << (k ? ", " : "") << sf[k];
    o << "};\n";
  }
  o << stage;
  // values of the staged variables at one tile + halo position, v[slot]: plain loads (after the
  // periodic wrap) in simple tiles, else through the ghost closures
  o << "  static __device__ __forceinline__ void stage_vals(const BInst& hd, T* v, int p0, int p1, int p2,\n"
       "      const bool simple, const double t, const T* __restrict__ P, const T* __restrict__ u, const JPtrs& jp) {\n"
       "    (void)t; (void)P; (void)jp; (void)u; (void)simple; (void)hd; (void)v;\n"
       "    if (simple) {\n";
  for (int a = 0; a < 3; ++a)
    if (g_.periodic[a]) o << "      p" << a << " = mol_wrap(mol_wrap(p" << a << ", hd.n[" << a << "]), hd.n[" << a << "]);\n";
  for (int f = 0; f < d_->nfields; ++f)
    if (pg_.slot[f] >= 0) o << "      v[" << pg_.slot[f] << "] = u[mol_b_uidx(hd, " << f << ", p0, p1, p2)];\n";
  o << "    } else {\n";
  for (int f = 0; f < d_->nfields; ++f)
    if (pg_.slot[f] >= 0) o << "      v[" << pg_.slot[f] << "] = stage(hd, " << f << ", p0, p1, p2, t, P, u, jp);\n";
  o << "    }\n  }\n";
  o << "  template <bool CORE>\n"
       "  static __device__ __forceinline__ void node(const BInst& hd, const T* __restrict__ L, const int b, const int I0,\n"
       "      const int I1, const int I2, const double t, const T* __restrict__ P, const T* __restrict__ u,\n"
       "      const JPtrs& jp, T* out) {\n"
       "    (void)I0; (void)I1; (void)I2; (void)t; (void)P; (void)jp; (void)L; (void)b; (void)hd; (void)u;\n"
       "    if constexpr (CORE) {\n";
  print_region(o, r_core);
  for (int f = 0; f < d_->nfields; ++f) o << "    out[" << f << "] = " << ref(out_core[f]) << ";\n";
  o << "    } else {\n";
  print_region(o, r_bound);
  for (int f = 0; f < d_->nfields; ++f) o << "    out[" << f << "] = " << ref(out_bound[f]) << ";\n";
  o << "    }\n  }\n};\n";
  return o.str();
}

uint64_t fnv1a(const std::string& s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 1099511628211ull;
  }
  return h;
}

struct CacheEntry {
  std::vector<char> code;
  std::map<int, hipModule_t> modules;   // per device
};
std::mutex g_jit_mu;
std::map<std::string, CacheEntry> g_jit_cache;   // key: generated source

// On-disk cache of compiled code objects, shared by processes: one file per generated source,
// named by its hash, holding a header (magic, hash of the embedded prelude and of the compiler
// version and options, source length), the source itself and the code object.  A file is used
// only when the whole source and the header match, so a hash collision or a library with
// another prelude compiles afresh.  Directory: $MOL_JIT_CACHE ("0" disables the cache), else
// $XDG_CACHE_HOME/molgpu/hiprtc, else $HOME/.cache/molgpu/hiprtc; failures to read or write
// it are ignored.
constexpr uint64_t kDiskMagic = 0x31657661636c6f6dull;   // "molcache1" tag

std::string disk_cache_dir() {
  const char* env = std::getenv("MOL_JIT_CACHE");
  if (env && env[0] == '0' && env[1] == '\0') return std::string();
  if (env && env[0]) return env;
  if (const char* x = std::getenv("XDG_CACHE_HOME"); x && x[0]) return std::string(x) + "/molgpu/hiprtc";
  if (const char* h = std::getenv("HOME"); h && h[0]) return std::string(h) + "/.cache/molgpu/hiprtc";
  return std::string();
}

uint64_t compiler_tag() {
  static const uint64_t tag = [] {
    int major = 0, minor = 0;
    (void)hiprtcVersion(&major, &minor);
    return fnv1a(std::string(mol_jit_prelude_text)) ^ (((uint64_t)major << 40) | ((uint64_t)minor << 24)) ^
           fnv1a("--offload-arch=gfx950 -O3 -std=c++17 -packed-fp32-ops");
  }();
  return tag;
}

std::string disk_cache_file(const std::string& dir, const std::string& src) {
  char name[48];
  std::snprintf(name, sizeof(name), "/%016llx.co", (unsigned long long)fnv1a(src));
  return dir + name;
}

bool disk_cache_load(const std::string& src, std::vector<char>& code) {
  const std::string dir = disk_cache_dir();
  if (dir.empty()) return false;
  FILE* f = std::fopen(disk_cache_file(dir, src).c_str(), "rb");
  if (!f) return false;
  uint64_t hdr[3] = {0, 0, 0};
  bool ok = std::fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == kDiskMagic && hdr[1] == compiler_tag() &&
            hdr[2] == src.size();
  if (ok) {
    std::string s(src.size(), '\0');
    ok = std::fread(&s[0], 1, s.size(), f) == s.size() && s == src;
  }
  if (ok) {
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t n;
    while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    ok = !buf.empty();
    if (ok) code.swap(buf);
  }
  std::fclose(f);
  return ok;
}

void disk_cache_store(const std::string& src, const std::vector<char>& code) {
  const std::string dir = disk_cache_dir();
  if (dir.empty()) return;
  // create the directory chain (mkdir -p); errors surface as a failed open below
  for (size_t i = 1; i <= dir.size(); ++i)
    if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
  const std::string path = disk_cache_file(dir, src);
  char tmp_suffix[40];
  std::snprintf(tmp_suffix, sizeof(tmp_suffix), ".tmp%d", (int)getpid());
  const std::string tmp = path + tmp_suffix;
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  const uint64_t hdr[3] = {kDiskMagic, compiler_tag(), (uint64_t)src.size()};
  bool ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1 && std::fwrite(src.data(), 1, src.size(), f) == src.size() &&
            std::fwrite(code.data(), 1, code.size(), f) == code.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) (void)std::remove(tmp.c_str());
}

}  // namespace

std::string jit_fused_source(const mol_desc* d, const JitGeom& g, std::string& err) {
  Gen gen(d, g);
  std::string s = gen.source_fused(d->dtype == MOL_F64);
  err = gen.error;
  return s;
}

std::string jit_source(const mol_desc* d, const JitGeom& g, std::string& err) {
  Gen gen(d, g);
  std::string s = gen.source(d->dtype == MOL_F64);
  err = gen.error;
  return s;
}

size_t jit_lds_per_pos(const mol_desc* d, const JitGeom& g, size_t esize, bool vjp) {
  Gen gen(d, g);
  const Gen::Usage us = gen.analyze();
  if (us.staged.empty()) return 0;
  size_t nf = 0, na = 0;
  for (int k = 0; k < d->nfields + d->naux; ++k)
    if (us.staged[k]) (k < d->nfields ? nf : na)++;
  const size_t bytes = vjp ? (nf + na) * (esize + sizeof(double)) : (nf + na) * esize;
  return bytes > 0 ? bytes : esize;
}

std::string jit_class_source(const mol_desc* d, const JitGeom& g, const std::string& name, std::string& err) {
  Gen gen(d, g);
  std::string s = gen.source_class(name);
  err = gen.error;
  return s;
}

std::string jit_vjp_source(const mol_desc* d, const JitGeom& g, std::string& err) {
  Gen gen(d, g);
  std::string s = gen.source_vjp(d->dtype == MOL_F64);
  err = gen.error;
  return s;
}

int jit_compile(const std::string& src, std::vector<char>& code, std::string& log) {
  hiprtcProgram prog;
  const char* hdr_names[] = {"mol_jit_prelude.h"};
  const char* hdrs[] = {mol_jit_prelude_text};
  if (hiprtcCreateProgram(&prog, src.c_str(), "mol_rhs_gen.hip", 1, hdrs, hdr_names) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return MOL_ERR_HIP;
  }
  // packed FP32 instructions stay off, as in the library build
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return MOL_ERR_HIP;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return 0;
}

// Compile (or fetch from the process-wide cache) and load the generated kernel on `device`.
int jit_kernel(const std::string& src, const char* name, int device, hipFunction_t* fn, std::string& err) {
  std::lock_guard<std::mutex> lk(g_jit_mu);
  CacheEntry& e = g_jit_cache[src];
  if (e.code.empty() && !disk_cache_load(src, e.code)) {
    std::string log;
    if (jit_compile(src, e.code, log)) {
      g_jit_cache.erase(src);
      err = "hiprtc: " + log;
      return MOL_ERR_HIP;
    }
    disk_cache_store(src, e.code);
    if (const char* dir = std::getenv("MOL_JIT_DUMP")) {
      char fname[64];
      std::snprintf(fname, sizeof(fname), "/mol_gen_%016llx.hip", (unsigned long long)fnv1a(src));
      if (FILE* f = std::fopen((std::string(dir) + fname).c_str(), "w")) {
        std::fwrite(src.data(), 1, src.size(), f);
        std::fclose(f);
      }
    }
  }
  auto it = e.modules.find(device);
  if (it == e.modules.end()) {
    hipModule_t m;
    const hipError_t he = hipModuleLoadData(&m, e.code.data());
    if (he != hipSuccess) {
      err = std::string("hipModuleLoadData: ") + hipGetErrorString(he);
      return MOL_ERR_HIP;
    }
    it = e.modules.emplace(device, m).first;
  }
  const hipError_t he = hipModuleGetFunction(fn, it->second, name);
  if (he != hipSuccess) {
    err = std::string("hipModuleGetFunction: ") + hipGetErrorString(he);
    return MOL_ERR_HIP;
  }
  return 0;
}

}  // namespace mol
