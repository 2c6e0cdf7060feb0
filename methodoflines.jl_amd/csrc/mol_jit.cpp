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

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
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

// Leaf context of a program.
enum CtxKind { CTX_NODE, CTX_PLAIN, CTX_HALF };
struct Ctx {
  CtxKind kind = CTX_NODE;
  int ax = 0, h = 0;              // half point between node+h and node+h+1 along ax
  const mol_optab* ip = nullptr;  // interpolation row (interior)
  const mol_optab* dp = nullptr;  // inner derivative row (interior)
  int dfield = 0;
};

class Gen {
 public:
  Gen(const mol_desc* d, const JitGeom& g) : d_(d), g_(g) {}

  std::string source(bool f64);
  std::string error;

 private:
  const mol_desc* d_;
  JitGeom g_;
  std::ostringstream body_;
  int nv_ = 0;
  std::map<int, std::string> slot_var_;

  std::string var() { return "v" + istr(nv_++); }
  std::string emit(const std::string& expr) {
    const std::string v = var();
    body_ << "    const T " << v << " = " << expr << ";\n";
    return v;
  }
  int stride(int ax) const {
    return ax == 0 ? 1 : (ax == 1 ? g_.ext[0] : g_.ext[0] * g_.ext[1]);
  }
  // LDS value of array `arr` at the node displaced by `off` (in elements)
  std::string at(int arr, long long off) const {
    long long o = (long long)arr * g_.vol + off;
    if (o == 0) return "L[b]";
    return o > 0 ? "L[b + " + istr(o) + "]" : "L[b - " + istr(-o) + "]";
  }
  std::string coord_node(int d) const {
    return "(T)jp.coord[" + istr(d) + "][I" + istr(d) + " - 1]";
  }
  std::string coord_tap(int d, int off) const {
    if (off == 0) return coord_node(d);
    const std::string k = "I" + istr(d) + (off > 0 ? " + " + istr(off) : " - " + istr(-off));
    if (g_.periodic[d]) return "(T)jp.coord[" + istr(d) + "][mol_wrap(" + k + ", " + istr(g_.n[d]) + ") - 1]";
    return "(T)jp.coord[" + istr(d) + "][" + k + " - 1]";
  }
  // interior-row stencil sum_k w[k] * arr[node + (off + k) e_ax]
  std::string stencil(const mol_optab& tb, int arr, int ax, int shift = 0) {
    std::string e;
    for (int k = 0; k < tb.len; ++k) {
      const double w = d_->pool_h[tb.w + k];
      const std::string term = lit(w) + " * " + at(arr, (long long)(shift + tb.off + k) * stride(ax));
      e = k == 0 ? term : e + " + " + term;
    }
    return e.empty() ? "T(0)" : emit(e);
  }
  std::string leaf(uint32_t op, int imm, const Ctx& c);
  std::string prog(int p, const Ctx& c);
  std::string slot(int k);
  std::string nonlinlap(const mol_slot& s, int qo, int qi, int qp);
  bool ok_ = true;
  void bad(const std::string& m) {
    if (ok_) error = m;
    ok_ = false;
  }
};

std::string Gen::leaf(uint32_t op, int imm, const Ctx& c) {
  switch (op) {
    case MOL_OP_CONST: return lit(d_->cpool_h[imm]);
    case MOL_OP_TIME: return "(T)t";
    case MOL_OP_PARAM: return "P[" + istr(imm) + "]";
    default: break;
  }
  if (c.kind == CTX_HALF) {
    auto interp = [&](int arr) {
      std::string e;
      for (int k = 0; k < c.ip->len; ++k) {
        const std::string term = lit(d_->pool_h[c.ip->w + k]) + " * " +
                                 at(arr, (long long)(c.h + c.ip->off + k) * stride(c.ax));
        e = k == 0 ? term : e + " + " + term;
      }
      return emit(e);
    };
    switch (op) {
      case MOL_OP_FIELD: return interp(imm);
      case MOL_OP_AUX: return interp(d_->nfields + imm);
      case MOL_OP_COORD: {
        if (imm != c.ax) return coord_node(imm);
        std::string e;
        for (int k = 0; k < c.ip->len; ++k) {
          const std::string term = lit(d_->pool_h[c.ip->w + k]) + " * " + coord_tap(imm, c.h + c.ip->off + k);
          e = k == 0 ? term : e + " + " + term;
        }
        return emit(e);
      }
      case MOL_OP_DINNER: {
        std::string e;
        for (int k = 0; k < c.dp->len; ++k) {
          const std::string term = lit(d_->pool_h[c.dp->w + k]) + " * " +
                                   at(c.dfield, (long long)(c.h + c.dp->off + k) * stride(c.ax));
          e = k == 0 ? term : e + " + " + term;
        }
        return emit(e);
      }
      case MOL_OP_SLOT: return "T(0)";
      default: bad("unknown leaf opcode"); return "T(0)";
    }
  }
  switch (op) {
    case MOL_OP_FIELD: return at(imm, 0);
    case MOL_OP_AUX: return at(d_->nfields + imm, 0);
    case MOL_OP_COORD: return coord_node(imm);
    case MOL_OP_SLOT: return c.kind == CTX_NODE ? slot(imm) : "T(0)";
    case MOL_OP_DINNER: return "T(0)";
    default: bad("unknown leaf opcode"); return "T(0)";
  }
}

const char* unary_fn(uint32_t op) {
  switch (op) {
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

std::string binexpr(uint32_t op, const std::string& a, const std::string& b) {
  switch (op) {
    case MOL_OP_ADD: return "(" + a + " + " + b + ")";
    case MOL_OP_SUB: return "(" + a + " - " + b + ")";
    case MOL_OP_MUL: return "(" + a + " * " + b + ")";
    case MOL_OP_DIV: return "(" + a + " / " + b + ")";
    case MOL_OP_POW: return "pow(" + a + ", " + b + ")";
    case MOL_OP_MIN: return "fmin(" + a + ", " + b + ")";
    case MOL_OP_MAX: return "fmax(" + a + ", " + b + ")";
    case MOL_OP_LT: return "(" + a + " < " + b + " ? T(1) : T(0))";
    case MOL_OP_LE: return "(" + a + " <= " + b + " ? T(1) : T(0))";
    case MOL_OP_GT: return "(" + a + " > " + b + " ? T(1) : T(0))";
    case MOL_OP_GE: return "(" + a + " >= " + b + " ? T(1) : T(0))";
    case MOL_OP_EQ: return "(" + a + " == " + b + " ? T(1) : T(0))";
    case MOL_OP_NE: return "(" + a + " != " + b + " ? T(1) : T(0))";
    case MOL_OP_AND: return "((" + a + " != T(0) && " + b + " != T(0)) ? T(1) : T(0))";
    case MOL_OP_OR: return "((" + a + " != T(0) || " + b + " != T(0)) ? T(1) : T(0))";
    default: return "";
  }
}

// Straight-line translation of a stack program (same evaluation order as stack_prog).
std::string Gen::prog(int p, const Ctx& c) {
  const int start = d_->prog_h[2 * p], len = d_->prog_h[2 * p + 1];
  std::vector<std::string> st;
  for (int pc = start; pc < start + len; ++pc) {
    const uint32_t ins = d_->code_h[pc];
    const uint32_t op = ins & 0xffu;
    const bool bconst = (ins >> 20) & 1u;
    const int imm = (int)(ins >> 21);
    if (op < 16u) {
      if (op > MOL_OP_DINNER) { bad("opcode " + istr(op) + " not supported by the generator"); return "T(0)"; }
      st.push_back(leaf(op, imm, c));
      continue;
    }
    if (op < 48u) {
      if (op == MOL_OP_SELECT) {
        if (bconst) {
          if (st.size() < 2) { bad("stack underflow"); return "T(0)"; }
          const std::string s0 = st.back(); st.pop_back();
          const std::string s1 = st.back(); st.pop_back();
          st.push_back(emit("(" + s1 + " != T(0)) ? " + s0 + " : " + lit(d_->cpool_h[imm])));
        } else {
          if (st.size() < 3) { bad("stack underflow"); return "T(0)"; }
          const std::string s0 = st.back(); st.pop_back();
          const std::string s1 = st.back(); st.pop_back();
          const std::string s2 = st.back(); st.pop_back();
          st.push_back(emit("(" + s2 + " != T(0)) ? " + s1 + " : " + s0));
        }
        continue;
      }
      std::string a, b;
      if (bconst) {
        if (st.empty()) { bad("stack underflow"); return "T(0)"; }
        a = st.back(); st.pop_back();
        b = lit(d_->cpool_h[imm]);
      } else {
        if (st.size() < 2) { bad("stack underflow"); return "T(0)"; }
        b = st.back(); st.pop_back();
        a = st.back(); st.pop_back();
      }
      const std::string e = binexpr(op, a, b);
      if (e.empty()) { bad("opcode " + istr(op) + " not supported by the generator"); return "T(0)"; }
      st.push_back(emit(e));
      continue;
    }
    if (st.empty()) { bad("stack underflow"); return "T(0)"; }
    const std::string a = st.back(); st.pop_back();
    std::string e;
    if (op == MOL_OP_NEG) e = "(-" + a + ")";
    else if (op == MOL_OP_SQ) e = "(" + a + " * " + a + ")";
    else if (op == MOL_OP_INV) e = "(T(1) / " + a + ")";
    else if (op == MOL_OP_POWI) {
      const int ex = (imm & 0x400) ? imm - 0x800 : imm;
      const int m = ex < 0 ? -ex : ex;
      std::string r;
      if (m == 0) r = "T(1)";
      else if (m == 1) r = a;
      else if (m == 2) r = "(" + a + " * " + a + ")";
      else if (m == 3) r = "(" + a + " * " + a + " * " + a + ")";
      e = m > 3 ? "mol_powi(" + a + ", " + istr(ex) + ")" : (ex < 0 ? "(T(1) / " + r + ")" : r);
    } else {
      const char* fn = unary_fn(op);
      if (!fn) { bad("opcode " + istr(op) + " not supported by the generator"); return "T(0)"; }
      e = std::string(fn) + "(" + a + ")";
    }
    st.push_back(emit(e));
  }
  if (st.size() != 1) { bad("program leaves " + istr((long long)st.size()) + " values"); return "T(0)"; }
  return st.back();
}

// Nonlinear Laplacian at an interior node: outer half-point operator at index i-1, the
// coefficient program and the inner derivative at each outer tap h
// (src/discretization/schemes/nonlinear_laplacian/nonlinear_laplacian.jl:20-83).
std::string Gen::nonlinlap(const mol_slot& s, int qo, int qi, int qp) {
  const mol_optab& to = d_->tabs_h[s.tab[qo]];
  Ctx c;
  c.kind = CTX_HALF;
  c.ax = s.axis;
  c.dp = &d_->tabs_h[s.tab[qi]];
  c.ip = &d_->tabs_h[s.tab[qp]];
  c.dfield = s.field;
  std::string acc;
  for (int k = 0; k < to.len; ++k) {
    c.h = to.off + k - 1;
    const std::string a = s.prog >= 0 ? prog(s.prog, c) : "";
    const std::string di = leaf(MOL_OP_DINNER, 0, c);
    const std::string term = emit(lit(d_->pool_h[to.w + k]) + " * " + (a.empty() ? di : "(" + a + " * " + di + ")"));
    acc = k == 0 ? term : emit(acc + " + " + term);
  }
  return acc.empty() ? "T(0)" : acc;
}

std::string Gen::slot(int k) {
  auto it = slot_var_.find(k);
  if (it != slot_var_.end()) return it->second;
  const mol_slot& s = d_->slots_h[k];
  const int ax = s.axis;
  Ctx plain;
  plain.kind = CTX_PLAIN;
  std::string r;
  switch (s.kind) {
    case MOL_SLOT_CENTERED:
      r = stencil(d_->tabs_h[s.tab[0]], s.field, ax);
      break;
    case MOL_SLOT_UPWIND: {
      if (s.prog < 0) {
        r = stencil(d_->tabs_h[s.tab[1]], s.field, ax);
        break;
      }
      const std::string coef = prog(s.prog, plain);
      const std::string back = stencil(d_->tabs_h[s.tab[1]], s.field, ax);
      const std::string fwd = stencil(d_->tabs_h[s.tab[0]], s.field, ax);
      r = emit(coef + " * (" + coef + " > T(0) ? " + back + " : " + fwd + ")");
      break;
    }
    case MOL_SLOT_WENO: {
      const int sd = stride(ax);
      // FP64: one-division cleared form; FP32: normalised weights (the cleared form underflows)
      r = emit(std::string(d_->dtype == MOL_F64 ? "mol_weno5(" : "mol_weno5_ref(") + at(s.field, -2LL * sd) + ", " + at(s.field, -sd) + ", " + at(s.field, 0) + ", " +
               at(s.field, sd) + ", " + at(s.field, 2LL * sd) + ", " + lit(s.eps) + ", " + lit(s.inv_dx) + ")");
      break;
    }
    case MOL_SLOT_NONLINLAP:
      r = nonlinlap(s, 0, 1, 2);
      break;
    case MOL_SLOT_SPHERICAL: {
      const std::string a = s.prog >= 0 ? prog(s.prog, plain) : "T(1)";
      const std::string rr = emit(coord_node(ax));
      const std::string d2 = stencil(d_->tabs_h[s.tab[1]], s.field, ax);
      const std::string d1 = stencil(d_->tabs_h[s.tab[0]], s.field, ax);
      const std::string nl = nonlinlap(s, 2, 3, 4);
      // r ~ 0: 6 a D2 u (src/discretization/schemes/spherical_laplacian/spherical_laplacian.jl:30-32)
      r = emit("(fabs((double)" + rr + ") <= 1e-6) ? T(6) * " + a + " * " + d2 + " : " + a + " * (" + d1 + " / " +
               rr + " + " + nl + ")");
      break;
    }
    case MOL_SLOT_MIXED: {
      const mol_optab& tx = d_->tabs_h[s.tab[0]];
      const mol_optab& ty = d_->tabs_h[s.tab[1]];
      std::string acc;
      for (int i = 0; i < tx.len; ++i) {
        std::string sy;
        for (int j = 0; j < ty.len; ++j) {
          const double w = d_->pool_h[tx.w + i] * d_->pool_h[ty.w + j];
          const long long off = (long long)(tx.off + i) * stride(ax) + (long long)(ty.off + j) * stride(s.axis2);
          const std::string term = lit(w) + " * " + at(s.field, off);
          sy = j == 0 ? term : sy + " + " + term;
        }
        const std::string v = emit(sy.empty() ? "T(0)" : sy);
        acc = i == 0 ? v : emit(acc + " + " + v);
      }
      r = acc.empty() ? "T(0)" : acc;
      break;
    }
    default:
      bad("slot kind " + istr(s.kind) + " not supported by the generator");
      r = "T(0)";
  }
  slot_var_[k] = r;
  return r;
}

std::string Gen::source(bool f64) {
  std::ostringstream o;
  o << "// generated right-hand side: core tiles\n";
  o << "#define MOL_T " << (f64 ? "double" : "float") << "\n";
  o << "#include \"mol_jit_prelude.h\"\n\n";
  Ctx node;
  node.kind = CTX_NODE;
  std::vector<std::string> outs;
  for (int f = 0; f < d_->nfields; ++f) outs.push_back(prog(d_->rhs_prog[f], node));
  if (!ok_) return "";
  int nblk = g_.ct_n[0] * g_.ct_n[1] * g_.ct_n[2];
  o << "struct Geo {\n";
  auto c = [&](const char* name, long long v) { o << "  static constexpr " << (std::strcmp(name, "BOX") ? "int " : "i64 ") << name << " = " << v << ";\n"; };
  c("NF", d_->nfields); c("NAUX", d_->naux); c("NARR", d_->nfields + d_->naux);
  c("N0", g_.n[0]); c("N1", g_.n[1]); c("N2", g_.n[2]);
  c("PER0", g_.periodic[0]); c("PER1", g_.periodic[1]); c("PER2", g_.periodic[2]);
  c("LO0", g_.lo[0]); c("LO1", g_.lo[1]); c("LO2", g_.lo[2]);
  c("HI0", g_.hi[0]); c("HI1", g_.hi[1]); c("HI2", g_.hi[2]);
  c("TX", g_.tile[0]); c("TY", g_.tile[1]); c("TZ", g_.tile[2]);
  c("HX", g_.halo[0]); c("HY", g_.halo[1]); c("HZ", g_.halo[2]);
  c("EX", g_.ext[0]); c("EY", g_.ext[1]); c("EZ", g_.ext[2]);
  c("VOL", g_.vol);
  c("CTLO0", g_.ct_lo[0]); c("CTLO1", g_.ct_lo[1]); c("CTLO2", g_.ct_lo[2]);
  c("CTN0", g_.ct_n[0]); c("CTN1", g_.ct_n[1]);
  c("NBLK", nblk);
  c("NTHR", 256);
  c("BOX", g_.box);
  bool ragged = false;
  for (int a = 0; a < 3; ++a) {
    const int last = g_.lo[a] + (g_.ct_lo[a] + g_.ct_n[a]) * g_.tile[a] - 1;
    ragged = ragged || last > g_.hi[a];
  }
  c("RAGGED", ragged ? 1 : 0);
  o << "  static __device__ __forceinline__ void node(const T* __restrict__ L, const int b, const int I0, const int I1,\n"
       "                                              const int I2, const double t, const T* __restrict__ P,\n"
       "                                              const JPtrs& jp, T* out) {\n";
  o << "    (void)I0; (void)I1; (void)I2; (void)t; (void)P; (void)jp;\n";
  o << body_.str();
  for (int f = 0; f < d_->nfields; ++f) o << "    out[" << f << "] = " << outs[f] << ";\n";
  o << "  }\n};\n\n";
  o << "extern \"C\" __global__ void __launch_bounds__(256) mol_core(const T* __restrict__ u, Epi ep, JPtrs jp) {\n"
       "  mol_core_tile<Geo>(u, ep, jp);\n}\n";
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

}  // namespace

std::string jit_core_source(const mol_desc* d, const JitGeom& g, std::string& err) {
  Gen gen(d, g);
  std::string s = gen.source(d->dtype == MOL_F64);
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
int jit_kernel(const std::string& src, int device, hipFunction_t* fn, std::string& err) {
  std::lock_guard<std::mutex> lk(g_jit_mu);
  CacheEntry& e = g_jit_cache[src];
  if (e.code.empty()) {
    std::string log;
    if (jit_compile(src, e.code, log)) {
      g_jit_cache.erase(src);
      err = "hiprtc: " + log;
      return MOL_ERR_HIP;
    }
    if (const char* dir = std::getenv("MOL_JIT_DUMP")) {
      char name[64];
      std::snprintf(name, sizeof(name), "/mol_core_%016llx.hip", (unsigned long long)fnv1a(src));
      if (FILE* f = std::fopen((std::string(dir) + name).c_str(), "w")) {
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
  const hipError_t he = hipModuleGetFunction(fn, it->second, "mol_core");
  if (he != hipSuccess) {
    err = std::string("hipModuleGetFunction: ") + hipGetErrorString(he);
    return MOL_ERR_HIP;
  }
  return 0;
}

}  // namespace mol
