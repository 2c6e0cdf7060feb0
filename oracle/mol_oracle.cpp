// mol_oracle.cpp — CPU restatement of MethodOfLines.jl's semi-discrete right-hand side.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is linked by the tests, by
// the smoke check and by the benchmark's CPU-baseline leg, never by the product path.
//
// It restates, node by node and in the reference's own terms, what the generated f(du,u,p,t)
// computes: the full grid of each variable is rebuilt, its boundary nodes are obtained by
// solving the boundary-condition and extrapolation equations of every grid line (what
// ModelingToolkit's mtkcompile does symbolically), and every unknown node applies the
// derivative rules with the reference's boundary-row selection.  It shares no code with the
// GPU library or its host front-end (weights, row selection, closures are all recomputed).
//
// Reference functions restated (paths relative to the reference root):
//   calculate_weights ............... src/discretization/schemes/fornberg_calculate_weights.jl:20-78
//   CompleteCenteredDifference ...... schemes/centered_difference/centered_diff_weights.jl:4-75
//   CompleteUpwindDifference ........ schemes/upwind_difference/upwind_diff_weights.jl:4-91
//   CompleteHalfCenteredDifference .. schemes/half_offset_weights.jl:4-70
//   BoundaryInterpolatorExtrapolator  schemes/extrapolation_weights.jl:1-72
//   central_difference .............. schemes/centered_difference/centered_difference.jl:5-63
//   _upwind_difference / winding .... schemes/upwind_difference/upwind_difference.jl:1-28,169-199
//   weno_f_uniform .................. schemes/WENO/WENO.jl:6-57
//   cartesian_nonlinear_laplacian ... schemes/nonlinear_laplacian/nonlinear_laplacian.jl:28-103
//   get_half_offset_weights_and_stencil schemes/half_offset_centred_difference.jl:9-37
//   spherical_diffusion ............. schemes/spherical_laplacian/spherical_laplacian.jl:10-42
//   mixed_central_difference ........ schemes/2nd_order_mixed_deriv/2nd_order_mixed_deriv.jl:5-22
//   bwrap / _wrapperiodic ........... src/discretization/interface_boundary.jl:26-42
//   BC / extrapolation / corner eqs . src/discretization/generate_bc_eqs.jl:24-58,238-311,336-416
//   interior (unknown set) .......... src/system_parsing/interior_map.jl:1-10,53-139
//   RK4 / SSPRK33 ................... textbook tableaus (OrdinaryDiffEq is not vendored)
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <tuple>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ------------------------------------------------------------------ forward-mode dual
struct Dual {
  double v, d;
  Dual(double a = 0.0, double b = 0.0) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
inline bool operator>(Dual a, double b) { return a.v > b; }
inline double val(double a) { return a; }
inline double val(Dual a) { return a.v; }

// ------------------------------------------------------------------ Fornberg weights
std::vector<double> calculate_weights(int order, double x0, const std::vector<double>& x) {
  const int N = (int)x.size();
  if (order >= N) throw std::runtime_error("Not enough points for the requested order.");
  const int M = order;
  double c1 = 1.0, c4 = x[0] - x0;
  std::vector<double> C((size_t)N * (M + 1), 0.0);
  auto at = [&](int i, int s) -> double& { return C[(size_t)i * (M + 1) + s]; };
  at(0, 0) = 1.0;
  for (int i = 1; i < N; ++i) {
    const int mn = std::min(i, M);
    double c2 = 1.0;
    const double c5 = c4;
    c4 = x[i] - x0;
    for (int j = 0; j < i; ++j) {
      const double c3 = x[i] - x[j];
      c2 *= c3;
      if (j == i - 1) {
        for (int s = mn; s >= 1; --s) at(i, s) = c1 * (s * at(i - 1, s - 1) - c5 * at(i - 1, s)) / c2;
        at(i, 0) = -c1 * c5 * at(i - 1, 0) / c2;
      }
      for (int s = mn; s >= 1; --s) at(j, s) = (c4 * at(j, s) - s * at(j, s - 1)) / c3;
      at(j, 0) = c4 * at(j, 0) / c3;
    }
    c1 = c2;
  }
  std::vector<double> w(N);
  for (int k = 0; k < N; ++k) w[k] = at(k, M);
  if (order != 0) {
    double s = 0.0;
    for (double v : w) s += v;
    w[N / 2] -= s;
  }
  return w;
}

// ------------------------------------------------------------------ derivative operators
struct Op {
  int d = 0, L = 0, BL = 0, bpc = 0, offside = 0;
  std::vector<double> st;
  std::vector<std::vector<double>> low, high;
  int pn_lo = 0;                              // nonuniform: interior row of node i is pn[i - pn_lo]
  std::vector<std::vector<double>> pn;
  const std::vector<double>& row(int i) const {
    if (pn.empty()) return st;
    if (i < pn_lo || i >= pn_lo + (int)pn.size()) throw std::runtime_error("node without an operator row");
    return pn[i - pn_lo];
  }
};

static std::vector<double> scaled(const std::vector<double>& w, double s) {
  std::vector<double> o(w.size());
  for (size_t k = 0; k < w.size(); ++k) o[k] = s * w[k];
  return o;
}

Op centered(int d, int p, double dx) {
  Op o;
  o.d = d;
  o.L = d + p - 1 + (d + p) % 2;
  o.BL = d + p;
  const int half = o.L / 2;
  std::vector<double> dummy;
  for (int k = -half; k <= half; ++k) dummy.push_back(k);
  const double sc = 1.0 / std::pow(dx, d);
  o.st = scaled(calculate_weights(d, 0.0, dummy), sc);
  std::vector<double> lb;
  for (int k = 0; k < o.BL; ++k) lb.push_back(k);
  o.bpc = half;
  for (int r = 0; r < half; ++r) {
    o.low.push_back(scaled(calculate_weights(d, (double)r, lb), sc));
    std::vector<double> h(o.low.back().rbegin(), o.low.back().rend());
    o.high.push_back(scaled(h, (d % 2) ? -1.0 : 1.0));
  }
  return o;
}

Op upwind(int d, int p, double dx, int offside) {
  Op o;
  o.d = d;
  o.L = d + p;
  o.BL = d + p;
  o.offside = offside;
  const int hc = o.L - 1 - offside;
  o.bpc = hc;
  std::vector<double> dummy;
  for (int k = 0; k < o.L; ++k) dummy.push_back(0.0 - offside + k);
  o.st = scaled(calculate_weights(d, 0.0, dummy), 1.0 / std::pow(dx, d));
  std::vector<double> lx, hx;
  for (int k = 0; k < o.BL; ++k) { lx.push_back(k); hx.push_back(-k); }
  for (int r = 0; r < offside; ++r) o.low.push_back(scaled(calculate_weights(d, (double)r, lx), 1.0 / std::pow(dx, d)));
  std::vector<std::vector<double>> h;
  for (int r = 0; r < hc; ++r) h.push_back(scaled(calculate_weights(d, -(double)r, hx), std::pow(-1.0 / dx, d)));
  o.high.assign(h.rbegin(), h.rend());
  return o;
}

// CompleteCenteredDifference(d, p, x) on nodes (centered_diff_weights.jl:77-153): node i in
// bpc+1 .. n-bpc takes Fornberg weights at x[i] on x[i-half .. i+half]; nodes 1..bpc and
// n-bpc+1..n take weights on the first / last BL nodes, measured from x[1] and from x[n-BL]
// by running sums of the spacings (:87-89); high row r belongs to node n - r (0-based r).
Op centered_nu(int d, int p, const std::vector<double>& x) {
  Op o;
  o.d = d;
  o.L = d + p - 1 + (d + p) % 2;
  o.BL = d + p;
  const int half = o.L / 2, n = (int)x.size();
  o.bpc = half;
  std::vector<double> dx(n - 1), lx(o.BL), hx(o.BL);
  for (int i = 0; i + 1 < n; ++i) dx[i] = x[i + 1] - x[i];
  lx[0] = 0.0;
  for (int k = 1; k < o.BL; ++k) lx[k] = lx[k - 1] + dx[k - 1];
  double acc = 0.0;
  for (int k = 0; k < o.BL; ++k) { acc += dx[n - 1 - o.BL + k]; hx[k] = acc; }
  o.pn_lo = half + 1;
  for (int i = half + 1; i <= n - half; ++i) {
    std::vector<double> pts(x.begin() + (i - 1 - half), x.begin() + (i + half));
    o.pn.push_back(calculate_weights(d, x[i - 1], pts));
  }
  for (int r = 0; r < half; ++r) {
    o.low.push_back(calculate_weights(d, lx[r], lx));
    o.high.push_back(calculate_weights(d, hx[o.BL - 1 - r], hx));
  }
  return o;
}

// CompleteUpwindDifference(d, p, x, offside) on nodes (upwind_diff_weights.jl:93-171) with the row
// selection of _upwind_difference (upwind_difference.jl:131-162).  Rows are built for nodes
// i = offside+1 .. n-hc at x[i] on x[i-offside .. i+L-1-offside] and the operator is stored with
// offside 0, so node II reads the row built for node II + offside: its own row for the forward
// operator (offside 0; high rows at x[n-hc+1..n] in that order for the last hc nodes), the row of
// node II+L-1 for the backward one (offside L-1; taps II-L+1..II, no low rows).
Op upwind_nu(int d, int p, const std::vector<double>& x, int offside) {
  Op o;
  o.d = d;
  o.L = d + p;
  o.BL = d + p;
  const int n = (int)x.size(), hc = o.L - 1 - offside;
  o.offside = 0;
  o.bpc = hc;
  std::vector<std::vector<double>> built;
  for (int i = offside + 1; i <= n - hc; ++i) {
    std::vector<double> pts(x.begin() + (i - 1 - offside), x.begin() + (i - 1 + o.L - offside));
    built.push_back(calculate_weights(d, x[i - 1], pts));
  }
  std::vector<double> hx(x.end() - o.BL, x.end());
  for (int k = 0; k < hc; ++k) o.high.push_back(calculate_weights(d, x[n - hc + k], hx));
  // node II (1-based) reads built[II - 1]
  o.pn_lo = 1;
  o.pn = built;
  return o;
}

Op half_op(int d, int p, double dx) {
  Op o;
  o.d = d;
  o.L = p + 2 * (d / 2) + (p % 2);
  o.BL = d + p;
  const int ep = o.L / 2;
  std::vector<double> dummy, lb;
  for (int k = 1 - ep; k <= ep; ++k) dummy.push_back(k);
  for (int k = 1; k <= o.BL; ++k) lb.push_back(k);
  o.bpc = o.L / 2;
  const double sc = 1.0 / std::pow(dx, d);
  o.st = scaled(calculate_weights(d, 0.5, dummy), sc);
  for (int r = 0; r < o.bpc; ++r) {
    o.low.push_back(scaled(calculate_weights(d, 1.5 + r, lb), sc));
    std::vector<double> h(o.low.back().rbegin(), o.low.back().rend());
    o.high.push_back(scaled(h, (d % 2) ? -1.0 : 1.0));
  }
  return o;
}

// CompleteHalfCenteredDifference(d, p, x) on nodes (half_offset_weights.jl:72-150): index i is the
// half point hx[i] = (x[i] + x[i+1]) / 2; half points ep+1 .. n-ep take weights at hx[i] on
// x[i-ep+1 .. i+ep] (the row selection reads row i - bpc, half_offset_centred_difference.jl:40-69),
// low rows at hx[1..ep] on the first BL nodes, high row k at hx[len-k] on the last BL nodes.
Op half_nu(int d, int p, const std::vector<double>& x) {
  Op o;
  o.d = d;
  o.L = p + 2 * (d / 2) + (p % 2);
  o.BL = d + p;
  const int ep = o.L / 2, n = (int)x.size();
  o.bpc = ep;
  std::vector<double> hx(n - 1);
  for (int i = 0; i + 1 < n; ++i) hx[i] = (x[i] + x[i + 1]) / 2;
  o.pn_lo = ep + 1;
  for (int i = ep + 1; i <= n - ep; ++i) {
    std::vector<double> pts(x.begin() + (i - ep), x.begin() + (i + ep));
    o.pn.push_back(calculate_weights(d, hx[i - 1], pts));
  }
  std::vector<double> lx(x.begin(), x.begin() + o.BL), rx(x.end() - o.BL, x.end());
  for (int k = 0; k < ep; ++k) {
    o.low.push_back(calculate_weights(d, hx[k], lx));
    o.high.push_back(calculate_weights(d, hx[n - 2 - k], rx));
  }
  return o;
}

Op extrap_op(int p) {
  Op o;
  o.L = p - 1 + p % 2;
  o.BL = p;
  const int half = o.L / 2;
  o.bpc = half;
  std::vector<double> dummy;
  for (int k = -half; k <= half; ++k)
    if (k != 0) dummy.push_back(k);
  o.st = calculate_weights(0, 0.0, dummy);
  o.st.insert(o.st.begin() + half, 0.0);
  for (int r = 0; r < half; ++r) {
    std::vector<double> pts;
    for (int k = 0; k < o.BL; ++k)
      if (k != r) pts.push_back(k);
    auto w = calculate_weights(0, (double)r, pts);
    w.insert(w.begin() + r, 0.0);
    o.low.push_back(w);
    o.high.push_back(std::vector<double>(w.rbegin(), w.rend()));
  }
  return o;
}

// ------------------------------------------------------------------ grids and fields
struct Axis {
  int n = 0;
  double dx = 0.0;
  std::vector<double> x;
  bool periodic = false;
  bool nonuniform = false;   // x holds arbitrary nodes: operators are built on them
};

struct Space {
  int nd = 1;
  Axis ax[3];
  int n(int a) const { return a < nd ? ax[a].n : 1; }
  size_t size() const { return (size_t)n(0) * n(1) * n(2); }
  size_t idx(const int* I) const {  // 1-based
    return (size_t)(I[0] - 1) + (size_t)n(0) * ((size_t)(I[1] - 1) + (size_t)n(1) * (size_t)(I[2] - 1));
  }
  int wrap(int a, int k) const {
    if (!ax[a].periodic) return k;
    const int l = ax[a].n;
    if (k <= 1) return k + (l - 1);
    if (k > l) return k - (l - 1);
    return k;
  }
};

// One boundary side: sum_d a[d] * D^d u (at the boundary node) = g(t, node coords).
struct Side {
  bool has = false;
  std::vector<double> a;                       // a[0] u + a[1] Du + a[2] D2u ...
  std::function<double(double, const int*)> g; // value at (t, node index)
};

template <typename S>
struct Field {
  const Space* sp = nullptr;
  int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
  std::vector<S> U;   // full grid
  int p = 2;          // approx order (for BC derivative operators)
  int ext_lo[3] = {0, 0, 0}, ext_hi[3] = {0, 0, 0};
  int vlo[3] = {0, 0, 0}, vhi[3] = {0, 0, 0};
  Side side[3][2];
  S& at(const int* I) { return U[sp->idx(I)]; }
  S get(int i, int j, int k) const {
    int I[3] = {i, j, k};
    return U[sp->idx(I)];
  }
  size_t box() const {
    size_t s = 1;
    for (int a = 0; a < sp->nd; ++a) s *= (size_t)(hi[a] - lo[a] + 1);
    return s;
  }
};

// unknown (compact, first axis fastest) <-> full grid
template <typename S>
void scatter(Field<S>& F, const S* u) {
  const Space& sp = *F.sp;
  F.U.assign(sp.size(), S(0.0));
  int I[3] = {1, 1, 1};
  size_t q = 0;
  for (I[2] = F.lo[2]; I[2] <= F.hi[2]; ++I[2])
    for (I[1] = F.lo[1]; I[1] <= F.hi[1]; ++I[1])
      for (I[0] = F.lo[0]; I[0] <= F.hi[0]; ++I[0]) F.at(I) = u[q++];
}

// Dense Gaussian elimination with partial pivoting (tiny per-line systems).
template <typename S>
void solve_dense(std::vector<double> A, std::vector<S>& b, int n) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0.0) throw std::runtime_error("singular boundary system");
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      if (f == 0.0) continue;
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      b[r] = b[r] - S(f) * b[c];
    }
  }
  for (int c = n - 1; c >= 0; --c) {
    S s = b[c];
    for (int k = c + 1; k < n; ++k) s = s - S(A[c * n + k]) * b[k];
    b[c] = s / S(A[c * n + c]);
  }
}

// central_difference row at node i on a non-periodic axis: (weights, first tap)
static void central_row(const Op& op, int i, int n, const std::vector<double>** w, int* first) {
  if (i <= op.bpc) { *w = &op.low[i - 1]; *first = 1; }
  else if (i > n - op.bpc) { *w = &op.high[n - i]; *first = n - op.BL + 1; }
  else { *w = &op.row(i); *first = i - op.L / 2; }
}

// Solve the boundary equations of every grid line outside the unknown box along each
// non-periodic axis: the BC equation at the domain edge node and the extrapolation
// equations of the nodes between it and the box (generate_bc_eqs.jl:238-311, 336-392).
// Nodes outside the box along two or more axes are the corner equations ~ 0 (:396-416).
template <typename S>
void fill_boundaries(Field<S>& F, double t) {
  const Space& sp = *F.sp;
  for (int a = 0; a < sp.nd; ++a) {
    if (sp.ax[a].periodic) continue;
    const int n = sp.ax[a].n;
    for (int s = 0; s < 2; ++s) {
      std::vector<int> ghosts;
      if (s == 0) for (int k = 1; k < F.lo[a]; ++k) ghosts.push_back(k);
      else for (int k = F.hi[a] + 1; k <= n; ++k) ghosts.push_back(k);
      if (ghosts.empty()) continue;
      const int ng = (int)ghosts.size();
      auto gidx = [&](int node) {
        for (int q = 0; q < ng; ++q) if (ghosts[q] == node) return q;
        return -1;
      };
      const Side& sd = F.side[a][s];
      const int bnode = s == 0 ? 1 : n;
      // equations: list of (coefficient per node, has_g)
      struct Eq { std::vector<std::pair<int, double>> terms; bool g; };
      std::vector<std::vector<Eq>> eqs(ng);
      if (sd.has) {
        Eq e; e.g = true;
        for (size_t dd = 0; dd < sd.a.size(); ++dd) {
          if (sd.a[dd] == 0.0) continue;
          if (dd == 0) { e.terms.push_back({bnode, sd.a[0]}); continue; }
          Op op = sp.ax[a].nonuniform ? centered_nu((int)dd, F.p, sp.ax[a].x) : centered((int)dd, F.p, sp.ax[a].dx);
          const std::vector<double>* w; int first;
          central_row(op, bnode, n, &w, &first);
          for (size_t k = 0; k < w->size(); ++k) e.terms.push_back({first + (int)k, sd.a[dd] * (*w)[k]});
        }
        eqs[gidx(bnode)].push_back(e);
      }
      const int ext = s == 0 ? F.ext_lo[a] : F.ext_hi[a];
      const int vl = s == 0 ? F.vlo[a] : F.vhi[a];
      Op eop = extrap_op(std::max(6, F.p));
      for (int ni = ext - vl; ni >= vl; --ni) {
        const int node = s == 0 ? 1 + ni : n - ni;
        if (gidx(node) < 0) continue;
        const std::vector<double>* w; int first;
        central_row(eop, node, n, &w, &first);
        Eq e; e.g = false;
        e.terms.push_back({node, 1.0});
        for (size_t k = 0; k < w->size(); ++k) e.terms.push_back({first + (int)k, -(*w)[k]});
        eqs[gidx(node)].push_back(e);
      }
      for (int q = 0; q < ng; ++q)
        if (eqs[q].empty()) throw std::runtime_error("boundary node without equation");
      // every line of the face: the other axes run over the unknown box
      int oth[2] = {-1, -1}, no = 0;
      for (int b = 0; b < sp.nd; ++b)
        if (b != a) oth[no++] = b;
      const int lo0 = no > 0 ? F.lo[oth[0]] : 1, hi0 = no > 0 ? F.hi[oth[0]] : 1;
      const int lo1 = no > 1 ? F.lo[oth[1]] : 1, hi1 = no > 1 ? F.hi[oth[1]] : 1;
      for (int j1 = lo1; j1 <= hi1; ++j1)
        for (int j0 = lo0; j0 <= hi0; ++j0) {
          int I[3] = {1, 1, 1};
          if (no > 0) I[oth[0]] = j0;
          if (no > 1) I[oth[1]] = j1;
          std::vector<double> A((size_t)ng * ng, 0.0);
          std::vector<S> rhs(ng, S(0.0));
          for (int q = 0; q < ng; ++q) {
            const double inv = 1.0 / eqs[q].size();
            for (const Eq& e : eqs[q]) {
              for (auto& tm : e.terms) {
                const int gi = gidx(tm.first);
                if (gi >= 0) A[q * ng + gi] += tm.second * inv;
                else {
                  I[a] = tm.first;
                  rhs[q] = rhs[q] - S(tm.second * inv) * F.at(I);
                }
              }
              if (e.g) {
                I[a] = bnode;
                rhs[q] = rhs[q] + S(sd.g ? sd.g(t, I) * inv : 0.0);
              }
            }
          }
          solve_dense(A, rhs, ng);
          for (int q = 0; q < ng; ++q) {
            I[a] = ghosts[q];
            F.at(I) = rhs[q];
          }
        }
    }
  }
  // periodic alias u[1] ~ u[n] (never read through a wrapped tap, kept for completeness)
  for (int a = 0; a < sp.nd; ++a) {
    if (!sp.ax[a].periodic) continue;
    int I[3], J[3];
    for (I[2] = 1; I[2] <= sp.n(2); ++I[2])
      for (I[1] = 1; I[1] <= sp.n(1); ++I[1])
        for (I[0] = 1; I[0] <= sp.n(0); ++I[0]) {
          if (I[a] != 1) continue;
          J[0] = I[0]; J[1] = I[1]; J[2] = I[2];
          J[a] = sp.ax[a].n;
          F.at(I) = F.at(J);
        }
  }
}

// ------------------------------------------------------------------ operators at a node
template <typename S>
S central(const Field<S>& F, const Op& op, const int* I, int a) {
  const Space& sp = *F.sp;
  const int n = sp.ax[a].n;
  const bool per = sp.ax[a].periodic;
  const std::vector<double>* w;
  int first;
  if (!per && I[a] <= op.bpc) { w = &op.low[I[a] - 1]; first = 1; }
  else if (!per && I[a] > n - op.bpc) { w = &op.high[n - I[a]]; first = n - op.BL + 1; }
  else { w = &op.row(I[a]); first = I[a] - op.L / 2; }
  S s(0.0);
  int J[3] = {I[0], I[1], I[2]};
  for (size_t k = 0; k < w->size(); ++k) {
    J[a] = sp.wrap(a, first + (int)k);
    s += S((*w)[k]) * F.get(J[0], J[1], J[2]);
  }
  return s;
}

template <typename S>
S upwind_at(const Field<S>& F, const Op& fwd, const Op& bwd, const int* I, int a, bool positive) {
  const Space& sp = *F.sp;
  const int n = sp.ax[a].n;
  const bool per = sp.ax[a].periodic;
  const std::vector<double>* w;
  int first;
  if (!positive) {
    if (!per && I[a] > n - fwd.bpc) { w = &fwd.high[n - I[a]]; first = n - fwd.BL + 1; }
    else { w = &fwd.row(I[a]); first = I[a]; }
  } else {
    if (!per && I[a] <= bwd.offside) { w = &bwd.low[I[a] - 1]; first = 1; }
    else { w = &bwd.row(I[a]); first = I[a] - bwd.L + 1; }
  }
  if (!per && (first < 1 || first + (int)w->size() - 1 > n)) throw std::runtime_error("upwind taps outside the axis");
  S s(0.0);
  int J[3] = {I[0], I[1], I[2]};
  for (size_t k = 0; k < w->size(); ++k) {
    J[a] = sp.wrap(a, first + (int)k);
    s += S((*w)[k]) * F.get(J[0], J[1], J[2]);
  }
  return s;
}

template <typename S>
S weno_f(S u_m2, S u_m1, S u_0, S u_p1, S u_p2, double eps, double dx) {
  const S b1 = S(13.0) * ((u_0 - S(2.0) * u_p1 + u_p2) * (u_0 - S(2.0) * u_p1 + u_p2)) / S(12.0) +
               ((S(3.0) * u_0 - S(4.0) * u_p1 + u_p2) * (S(3.0) * u_0 - S(4.0) * u_p1 + u_p2)) / S(4.0);
  const S b2 = S(13.0) * ((u_m1 - S(2.0) * u_0 + u_p1) * (u_m1 - S(2.0) * u_0 + u_p1)) / S(12.0) +
               ((u_m1 - u_p1) * (u_m1 - u_p1)) / S(4.0);
  const S b3 = S(13.0) * ((u_m2 - S(2.0) * u_m1 + u_0) * (u_m2 - S(2.0) * u_m1 + u_0)) / S(12.0) +
               ((u_m2 - S(4.0) * u_m1 + S(3.0) * u_0) * (u_m2 - S(4.0) * u_m1 + S(3.0) * u_0)) / S(4.0);
  const S e(eps);
  const S wm1 = S(1.0 / 10) / ((e + b1) * (e + b1)), wm2 = S(3.0 / 5) / ((e + b2) * (e + b2)),
          wm3 = S(3.0 / 10) / ((e + b3) * (e + b3));
  const S wmd = wm1 + wm2 + wm3;
  const S wp1 = S(3.0 / 10) / ((e + b1) * (e + b1)), wp2 = S(3.0 / 5) / ((e + b2) * (e + b2)),
          wp3 = S(1.0 / 10) / ((e + b3) * (e + b3));
  const S wpd = wp1 + wp2 + wp3;
  const S hm1 = (S(11.0) * u_0 - S(7.0) * u_p1 + S(2.0) * u_p2) / S(6.0);
  const S hm2 = (S(5.0) * u_0 - u_p1 + S(2.0) * u_m1) / S(6.0);
  const S hm3 = (S(2.0) * u_0 + S(5.0) * u_m1 - u_m2) / S(6.0);
  const S hp1 = (S(2.0) * u_0 + S(5.0) * u_p1 - u_p2) / S(6.0);
  const S hp2 = (S(5.0) * u_0 + S(2.0) * u_p1 - u_m1) / S(6.0);
  const S hp3 = (S(11.0) * u_0 - S(7.0) * u_m1 + S(2.0) * u_m2) / S(6.0);
  const S hp = (wp1 / wpd) * hp1 + (wp2 / wpd) * hp2 + (wp3 / wpd) * hp3;
  const S hm = (wm1 / wmd) * hm1 + (wm2 / wmd) * hm2 + (wm3 / wmd) * hm3;
  return (hp - hm) / S(dx);
}

// ---- nonuniform node-centred WENO-5 first derivative (weno_f_nonuniform,
//      src/discretization/schemes/WENO/nonuniform_weno.jl): Fornberg 3-point sub-stencil weights
//      (:5-46), Simpson smoothness indicators (:54-73), target geometry (:76-82), closed-form
//      ideal weights (:85-118), Shi-Hu-Shu positive/negative splitting with theta = 3 (:120-163).
static void fornberg3(const double al[3], double xt, double m0[3], double m1[3], double m2[3]) {
  const double a0 = al[0], a1 = al[1], a2 = al[2];
  double n1m0 = 1.0, n1m1 = 0.0, n1m2 = 0.0, c1 = 1.0;
  double c2 = a1 - a0;
  const double r1 = c1 / c2;
  const double tA = a1 - xt;
  const double a1m0 = (tA * n1m0) / c2, a1m1 = (tA * n1m1 - n1m0) / c2, a1m2 = (tA * n1m2 - 2.0 * n1m1) / c2;
  const double sB = a0 - xt;
  const double a2m0 = r1 * (-(sB)*n1m0), a2m1 = r1 * (n1m0 - sB * n1m1), a2m2 = r1 * (2.0 * n1m1 - sB * n1m2);
  c1 = c2;
  n1m0 = a1m0; n1m1 = a1m1; n1m2 = a1m2;
  const double n2m0 = a2m0, n2m1 = a2m1, n2m2 = a2m2;
  c2 = (a2 - a0) * (a2 - a1);
  const double r2 = c1 / c2;
  const double c3a = a2 - a0, c3b = a2 - a1, tA2 = a2 - xt;
  m0[0] = (tA2 * n1m0) / c3a; m1[0] = (tA2 * n1m1 - n1m0) / c3a; m2[0] = (tA2 * n1m2 - 2.0 * n1m1) / c3a;
  m0[1] = (tA2 * n2m0) / c3b; m1[1] = (tA2 * n2m1 - n2m0) / c3b; m2[1] = (tA2 * n2m2 - 2.0 * n2m1) / c3b;
  const double sB2 = a1 - xt;
  m0[2] = r2 * (-(sB2)*n2m0); m1[2] = r2 * (n2m0 - sB2 * n2m1); m2[2] = r2 * (2.0 * n2m1 - sB2 * n2m2);
}

template <typename S>
static S dot3(const double w[3], S a, S b, S c) { return S(w[0]) * a + S(w[1]) * b + S(w[2]) * c; }

template <typename S>
static void substencil_beta_r(const double al[3], S ua, S ub, S uc, double xi, double xL, double xM, double xph,
                              double dx, S* beta, S* r) {
  double m0[3], m1i[3], m1L[3], m1M[3], m2M[3], m1R[3], t[3];
  fornberg3(al, xi, m0, m1i, t);
  fornberg3(al, xL, m0, m1L, t);
  fornberg3(al, xM, m0, m1M, m2M);
  fornberg3(al, xph, m0, m1R, t);
  *r = dot3(m1i, ua, ub, uc);
  const S pL = dot3(m1L, ua, ub, uc), pM = dot3(m1M, ua, ub, uc), pR = dot3(m1R, ua, ub, uc);
  const S pp = dot3(m2M, ua, ub, uc);
  const S I1 = S(dx / 6) * (pL * pL + S(4.0) * (pM * pM) + pR * pR);
  const S I2 = S(dx) * (pp * pp);
  const S v = S(dx) * I1 + S(dx * dx * dx) * I2;
  *beta = val(v) > 0.0 ? v : S(0.0);
}

static void weno_ideal(int target, double x1, double x2, double x3, double x4, double x5, double* d0, double* d2) {
  switch (target) {
    case 1:
      *d0 = ((2 * x1 - x2 - x3) * (x1 - x4) * (x1 - x5) + (x1 - x3) * (x1 - x5) * (x1 - x2) +
             (x1 - x3) * (x1 - x4) * (x1 - x2)) / ((2 * x1 - x2 - x3) * (x1 - x4) * (x1 - x5));
      *d2 = ((x1 - x3) * (x1 - x4) * (x1 - x2)) / ((-x1 + x5) * (2 * x1 - x3 - x4) * (-x2 + x5));
      return;
    case 2:
      *d0 = ((x2 - x4) * (x2 - x5)) / ((x1 - x4) * (x1 - x5));
      *d2 = ((-x1 + x2) * (x2 - x3) * (x2 - x4)) / ((-x1 + x5) * (2 * x2 - x3 - x4) * (-x2 + x5));
      return;
    case 4:
      *d0 = ((-x2 + x4) * (-x3 + x4) * (x4 - x5)) / ((x1 - x4) * (x1 - x5) * (-x2 - x3 + 2 * x4));
      *d2 = ((-x1 + x4) * (-x2 + x4)) / ((-x1 + x5) * (-x2 + x5));
      return;
    case 5:
      *d0 = ((-x2 + x5) * (-x3 + x5) * (-x4 + x5)) / ((x1 - x4) * (x1 - x5) * (-x2 - x3 + 2 * x5));
      *d2 = ((-x1 - x4 + 2 * x5) * (-x2 + x5) * (-x3 + x5) + (-x1 + x5) * (-x2 - x3 + 2 * x5) * (-x4 + x5)) /
            ((-x1 + x5) * (-x2 + x5) * (-x3 - x4 + 2 * x5));
      return;
    default:
      *d0 = ((x3 - x4) * (x3 - x5)) / ((x1 - x4) * (x1 - x5));
      *d2 = ((x3 - x1) * (x3 - x2)) / ((x5 - x1) * (x5 - x2));
  }
}

template <typename S>
S weno_nu(const S* u, const double* x, double eps, int target) {
  const double x1 = x[0], x2 = x[1], x3 = x[2], x4 = x[3], x5 = x[4];
  double xi, xL, xph;
  switch (target) {
    case 1: xi = x1; xL = x1; xph = (x1 + x2) / 2; break;
    case 2: xi = x2; xL = (x1 + x2) / 2; xph = (x2 + x3) / 2; break;
    case 4: xi = x4; xL = (x3 + x4) / 2; xph = (x4 + x5) / 2; break;
    case 5: xi = x5; xL = (x4 + x5) / 2; xph = x5; break;
    default: xi = x3; xL = (x2 + x3) / 2; xph = (x3 + x4) / 2;
  }
  const double dx = xph - xL, xM = (xL + xph) / 2;
  const double s0[3] = {x1, x2, x3}, s1[3] = {x2, x3, x4}, s2[3] = {x3, x4, x5};
  S b0, b1, b2, r0, r1, r2;
  substencil_beta_r(s0, u[0], u[1], u[2], xi, xL, xM, xph, dx, &b0, &r0);
  substencil_beta_r(s1, u[1], u[2], u[3], xi, xL, xM, xph, dx, &b1, &r1);
  substencil_beta_r(s2, u[2], u[3], u[4], xi, xL, xM, xph, dx, &b2, &r2);
  double d0, d2;
  weno_ideal(target, x1, x2, x3, x4, x5, &d0, &d2);
  const double d1 = 1.0 - d0 - d2, th = 3.0, half = 0.5;
  const double dp0 = half * (d0 + th * std::fabs(d0)), dp1 = half * (d1 + th * std::fabs(d1)),
               dp2 = half * (d2 + th * std::fabs(d2));
  const double dm0 = dp0 - d0, dm1 = dp1 - d1, dm2 = dp2 - d2;
  const double sgp = dp0 + dp1 + dp2, sgm = dm0 + dm1 + dm2;
  const S e(eps);
  const S ap0 = S(dp0 / sgp) / ((e + b0) * (e + b0)), ap1 = S(dp1 / sgp) / ((e + b1) * (e + b1)),
          ap2 = S(dp2 / sgp) / ((e + b2) * (e + b2));
  const S sp = ap0 + ap1 + ap2;
  const S am0 = S(dm0 / sgm) / ((e + b0) * (e + b0)), am1 = S(dm1 / sgm) / ((e + b1) * (e + b1)),
          am2 = S(dm2 / sgm) / ((e + b2) * (e + b2));
  const S sm = am0 + am1 + am2;
  const S Rp = (ap0 / sp) * r0 + (ap1 / sp) * r1 + (ap2 / sp) * r2;
  const S Rm = (am0 / sm) * r0 + (am1 / sm) * r1 + (am2 / sm) * r2;
  return S(sgp) * Rp - S(sgm) * Rm;
}

// Nonuniform WENO at node I along axis a (get_f_taps_coords, function_scheme.jl:1-34): near a
// non-periodic end the one-sided targets 1, 2 (taps 1..5) and 4, 5 (taps n-4..n); elsewhere
// target 3 on taps I-2..I+2, wrapped on a periodic axis with the coordinates continued across
// the end (bcoord, interface_boundary.jl:120-151).
template <typename S>
S weno_nu_at(const Field<S>& F, const int* I, int a, double eps) {
  const Space& sp = *F.sp;
  const Axis& ax = sp.ax[a];
  const int n = ax.n;
  int first = I[a] - 2, target = 3;
  if (!ax.periodic && I[a] <= 2) { first = 1; target = I[a]; }
  else if (!ax.periodic && I[a] > n - 2) { first = n - 4; target = 5 - (n - I[a]); }
  S v[5];
  double xs[5];
  int J[3] = {I[0], I[1], I[2]};
  for (int q = 0; q < 5; ++q) {
    const int k = first + q;
    J[a] = sp.wrap(a, k);
    v[q] = F.get(J[0], J[1], J[2]);
    if (k <= 1 && ax.periodic) xs[q] = ax.x[k + n - 1 - 1] - (ax.x[n - 1] - ax.x[0]);
    else if (k > n && ax.periodic) xs[q] = ax.x[k + 1 - n - 1] + (ax.x[n - 1] - ax.x[0]);
    else xs[q] = ax.x[k - 1];
  }
  return weno_nu(v, xs, eps, target);
}

template <typename S>
S weno_at(const Field<S>& F, const int* I, int a, double eps) {
  const Space& sp = *F.sp;
  S v[5];
  int J[3] = {I[0], I[1], I[2]};
  for (int q = 0; q < 5; ++q) {
    J[a] = sp.wrap(a, I[a] - 2 + q);
    v[q] = F.get(J[0], J[1], J[2]);
  }
  return weno_f(v[0], v[1], v[2], v[3], v[4], eps, sp.ax[a].dx);
}

// get_half_offset_weights_and_stencil at (half-point or node) index i with length len
static void half_row(const Op& op, int i, int len, bool per, const std::vector<double>** w, int* first, int* cnt) {
  if (!per && i <= op.bpc) { *w = &op.low[i - 1]; *first = 1; *cnt = op.BL; }
  else if (!per && i > len - op.bpc) { *w = &op.high[len - i - 1]; *first = len - op.BL + 1; *cnt = op.BL; }
  else { *w = &op.row(i); *first = i + 1 - op.L / 2; *cnt = op.L; }
}

// The three half-point operators of the nonlinear Laplacian on a nonuniform axis: the outer one
// on the midpoints of the nodes (differential_discretizer.jl:44-51), the inner derivative and
// the max(4, p)-point interpolation on the nodes (:97-98); cached per axis.
struct HalfNU { Op outer, inner, interp; };
static const HalfNU& cached_half_nu(const std::vector<double>& x, int p) {
  thread_local std::vector<std::pair<std::vector<double>, std::pair<int, std::unique_ptr<HalfNU>>>> cache;
  for (auto& e : cache)
    if (e.second.first == p && e.first == x) return *e.second.second;
  auto h = std::make_unique<HalfNU>();
  std::vector<double> mid(x.size() - 1);
  for (size_t i = 0; i + 1 < x.size(); ++i) mid[i] = (x[i] + x[i + 1]) / 2;
  h->outer = half_nu(1, p, mid);
  h->inner = half_nu(1, p, x);
  h->interp = half_nu(0, std::max(4, p), x);
  cache.emplace_back(x, std::make_pair(p, std::move(h)));
  return *cache.back().second.second;
}

// Half-offset operators are rebuilt for every node by the reference's rules; cache them per
// (order, approximation order, dx) so the oracle stays usable at tens of thousands of nodes.
static const Op& cached_half(int d, int p, double dx) {
  thread_local std::vector<std::pair<std::tuple<int, int, double>, std::unique_ptr<Op>>> cache;
  for (auto& e : cache)
    if (e.first == std::make_tuple(d, p, dx)) return *e.second;
  cache.emplace_back(std::make_tuple(d, p, dx), std::make_unique<Op>(half_op(d, p, dx)));
  return *cache.back().second;
}

// Coefficient callback of the nonlinear Laplacian: a(u_interp[nvars], x_half, node index, aux_interp[naux])
template <typename S>
using CoefFn = std::function<S(const S* u, double xh, const int* I, const S* aux)>;

// cartesian_nonlinear_laplacian at node I along axis a for variable `var` of `fields`.
template <typename S>
S nonlinlap(const std::vector<const Field<S>*>& fields, int var, const std::vector<const Field<S>*>& aux,
            const CoefFn<S>& coef, const int* I, int a, int p) {
  const Field<S>& F = *fields[var];
  const Space& sp = *F.sp;
  const int n = sp.ax[a].n;
  const bool per = sp.ax[a].periodic;
  const double dx = sp.ax[a].dx;
  const HalfNU* nu = sp.ax[a].nonuniform ? &cached_half_nu(sp.ax[a].x, p) : nullptr;
  const Op& outer = nu ? nu->outer : cached_half(1, p, dx);
  const Op& inner = nu ? nu->inner : outer;
  const Op& interp = nu ? nu->interp : cached_half(0, std::max(4, p), dx);
  const std::vector<double>* wo; int fo, co;
  half_row(outer, I[a] - 1, n - 1, per, &wo, &fo, &co);
  S acc(0.0);
  int J[3] = {I[0], I[1], I[2]};
  std::vector<S> uv(fields.size()), av(aux.size());
  for (int k = 0; k < co; ++k) {
    const int h = fo + k;
    const std::vector<double>* wi; int fi, ci;
    half_row(interp, h, n, per, &wi, &fi, &ci);
    for (size_t v = 0; v < fields.size(); ++v) {
      S s(0.0);
      for (int q = 0; q < ci; ++q) { J[a] = sp.wrap(a, fi + q); s += S((*wi)[q]) * fields[v]->get(J[0], J[1], J[2]); }
      uv[v] = s;
    }
    for (size_t v = 0; v < aux.size(); ++v) {
      S s(0.0);
      for (int q = 0; q < ci; ++q) { J[a] = sp.wrap(a, fi + q); s += S((*wi)[q]) * aux[v]->get(J[0], J[1], J[2]); }
      av[v] = s;
    }
    double xh = 0.0;
    for (int q = 0; q < ci; ++q) xh += (*wi)[q] * sp.ax[a].x[sp.wrap(a, fi + q) - 1];
    const std::vector<double>* wd; int fd, cd;
    half_row(inner, h, n, per, &wd, &fd, &cd);
    S du(0.0);
    for (int q = 0; q < cd; ++q) { J[a] = sp.wrap(a, fd + q); du += S((*wd)[q]) * F.get(J[0], J[1], J[2]); }
    const S av_ = coef(uv.data(), xh, I, av.data());
    acc += S((*wo)[k]) * (av_ * du);
  }
  return acc;
}

template <typename S>
S mixed_at(const Field<S>& F, const Op& Dx, const Op& Dy, const int* I, int ax, int ay) {
  const Space& sp = *F.sp;
  auto row = [&](const Op& op, int a, std::vector<double>& w, std::vector<int>& taps) {
    const int n = sp.ax[a].n;
    const bool per = sp.ax[a].periodic;
    const std::vector<double>* ww; int first;
    if (!per && I[a] <= op.bpc) { ww = &op.low[I[a] - 1]; first = 1; }
    else if (!per && I[a] > n - op.bpc) { ww = &op.high[n - I[a]]; first = n - op.BL + 1; }
    else { ww = &op.row(I[a]); first = I[a] - op.L / 2; }   // per-node row on a nonuniform axis
    w = *ww;
    taps.clear();
    for (size_t k = 0; k < w.size(); ++k) taps.push_back(sp.wrap(a, first + (int)k));
  };
  std::vector<double> wx, wy;
  std::vector<int> tx, ty;
  row(Dx, ax, wx, tx);
  row(Dy, ay, wy, ty);
  S out(0.0);
  for (size_t i = 0; i < wx.size(); ++i) {
    S sy(0.0);
    for (size_t j = 0; j < wy.size(); ++j) {
      int J[3] = {I[0], I[1], I[2]};
      J[ax] = tx[i];
      J[ay] = ty[j];
      sy += S(wx[i] * wy[j]) * F.get(J[0], J[1], J[2]);
    }
    out += sy;
  }
  return out;
}

// ------------------------------------------------------------------ systems
// A system maps the unknown vector to du (compact ordering, variables concatenated).
// Length of the next state array passed in (set by orc_expect, checked and cleared by the
// first System::rhs of the call; -1: not stated).
static thread_local long long g_expect = -1;

template <typename S>
struct System {
  Space sp;
  std::vector<Field<S>> F;
  std::vector<Field<S>> aux;   // full-grid parameter fields (d0, s)
  int p = 2;
  std::function<void(System&, const int* I, double t, S* out)> node_rhs;  // writes one value per variable
  size_t nunk() const {
    size_t s = 0;
    for (auto& f : F) s += f.box();
    return s;
  }
  void rhs(const S* u, double t, S* du, int threads) {
    if (g_expect >= 0) {
      // the caller stated the length of u (orc_expect): it must be the number of unknowns
      const long long e = g_expect;
      g_expect = -1;
      if (e != (long long)nunk()) throw std::runtime_error("state length differs from the number of unknowns");
    }
    size_t off = 0;
    for (auto& f : F) { scatter(f, u + off); off += f.box(); }
    for (auto& f : F) fill_boundaries(f, t);
    // every variable shares the unknown box in these systems
    const Field<S>& f0 = F[0];
    const long long nx = f0.hi[0] - f0.lo[0] + 1, ny = f0.hi[1] - f0.lo[1] + 1, nz = f0.hi[2] - f0.lo[2] + 1;
    const long long tot = nx * ny * nz;
    const size_t box = f0.box();
    (void)threads;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (long long q = 0; q < tot; ++q) {
      int I[3] = {(int)(q % nx) + f0.lo[0], (int)((q / nx) % ny) + f0.lo[1], (int)(q / (nx * ny)) + f0.lo[2]};
      S out[8];
      node_rhs(*this, I, t, out);
      for (size_t v = 0; v < F.size(); ++v) du[v * box + q] = out[v];
    }
  }
};

static void make_axis(Axis& ax, int n, double lo, double hi, bool per) {
  ax.n = n;
  ax.periodic = per;
  ax.dx = (hi - lo) / (n - 1);
  ax.x.resize(n);
  for (int k = 0; k < n; ++k) ax.x[k] = lo + (hi - lo) * k / (n - 1);
}

template <typename S>
void box_default(System<S>& sys, Field<S>& f) {
  f.sp = &sys.sp;
  f.p = sys.p;
  for (int a = 0; a < 3; ++a) {
    f.lo[a] = 1;
    f.hi[a] = sys.sp.n(a);
  }
  for (int a = 0; a < sys.sp.nd; ++a) {
    const bool per = sys.sp.ax[a].periodic;
    f.vlo[a] = (per || f.side[a][0].has) ? 1 : 0;
    f.vhi[a] = (!per && f.side[a][1].has) ? 1 : 0;
    f.lo[a] = 1 + std::max(f.vlo[a], f.ext_lo[a]);
    f.hi[a] = sys.sp.ax[a].n - std::max(f.vhi[a], f.ext_hi[a]);
  }
}

}  // namespace orc

using namespace orc;

// ====================================================================== C API
extern "C" {

int orc_calculate_weights(int order, double x0, const double* x, int N, double* out) {
  try {
    auto w = calculate_weights(order, x0, std::vector<double>(x, x + N));
    std::memcpy(out, w.data(), sizeof(double) * N);
    return 0;
  } catch (...) {
    return -1;
  }
}

// Centered operator rows: writes stencil (L), low rows (bpc*BL), high rows (bpc*BL); returns L,BL,bpc.
int orc_centered_weights(int d, int p, double dx, double* st, double* low, double* high, int* dims) {
  Op o = centered(d, p, dx);
  dims[0] = o.L; dims[1] = o.BL; dims[2] = o.bpc;
  std::memcpy(st, o.st.data(), sizeof(double) * o.L);
  for (int r = 0; r < o.bpc; ++r) {
    std::memcpy(low + r * o.BL, o.low[r].data(), sizeof(double) * o.BL);
    std::memcpy(high + r * o.BL, o.high[r].data(), sizeof(double) * o.BL);
  }
  return 0;
}

int orc_half_weights(int d, int p, double dx, double* st, double* low, double* high, int* dims) {
  Op o = half_op(d, p, dx);
  dims[0] = o.L; dims[1] = o.BL; dims[2] = o.bpc;
  std::memcpy(st, o.st.data(), sizeof(double) * o.L);
  for (int r = 0; r < o.bpc; ++r) {
    std::memcpy(low + r * o.BL, o.low[r].data(), sizeof(double) * o.BL);
    std::memcpy(high + r * o.BL, o.high[r].data(), sizeof(double) * o.BL);
  }
  return 0;
}

int orc_expect(long long n) {
  g_expect = n;
  return 0;
}

int orc_weno_nu(const double* u5, const double* x5, double eps, int target, double* out) {
  if (target < 1 || target > 5) return -1;
  *out = weno_nu(u5, x5, eps, target);
  return 0;
}

// u_t = -c u_x with nonuniform WENO on the given nodes x[0..n-1]: Dirichlet u = gl, gh at the
// ends (unknowns 2..n-1, WENO extent 0 on nonuniform grids: WENO.jl:94-95) or periodic.
int orc_weno_nu1d_rhs(int n, const double* x, double c, double eps, int periodic, double gl, double gh,
                      const double* u, double t, double* du) {
  try {
    System<double> s;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, x[0], x[n - 1], periodic != 0);
    for (int i = 0; i < n; ++i) s.sp.ax[0].x[i] = x[i];
    s.F.resize(1);
    Field<double>& f = s.F[0];
    f.ext_lo[0] = f.ext_hi[0] = 0;
    if (!periodic) {
      f.side[0][0].has = true; f.side[0][0].a = {1.0}; f.side[0][0].g = [gl](double, const int*) { return gl; };
      f.side[0][1].has = true; f.side[0][1].a = {1.0}; f.side[0][1].g = [gh](double, const int*) { return gh; };
    }
    box_default(s, f);
    s.node_rhs = [c, eps](System<double>& S, const int* I, double, double* out) {
      out[0] = -(c * weno_nu_at(S.F[0], I, 0, eps));
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

int orc_weno_f(const double* u5, double eps, double dx, double* out) {
  *out = weno_f(u5[0], u5[1], u5[2], u5[3], u5[4], eps, dx);
  return 0;
}

// ---- 1D diffusion u_t = D u_xx on [xlo, xhi], n nodes, general affine BCs
//      a0*u + a1*u_x = g at each end (a0=1,a1=0: Dirichlet; a0=0,a1=1: Neumann).
int orc_diffusion1d_rhs(int n, double xlo, double xhi, int p, double D, const double* bcl, const double* bch,
                        const double* u, double t, double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, xlo, xhi, false);
    s.F.resize(1);
    Field<double>& f = s.F[0];
    for (int side = 0; side < 2; ++side) {
      const double* b = side ? bch : bcl;
      f.side[0][side].has = true;
      f.side[0][side].a = {b[0], b[1]};
      const double g = b[2];
      f.side[0][side].g = [g](double, const int*) { return g; };
    }
    box_default(s, f);
    Op op = centered(2, p, s.sp.ax[0].dx);
    s.node_rhs = [D, op](System<double>& S, const int* I, double, double* out) {
      out[0] = D * central(S.F[0], op, I, 0);
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- 1D diffusion u_t = D u_xx on an edge-aligned grid of [xlo, xhi]: ncell cells of width
//      h = (xhi - xlo) / ncell, nodes x_i = xlo - h/2 + (i - 1) h for i = 1..n = ncell + 2
//      (generate_grid, discretize_vars.jl:373-390), unknowns at the cell centres 2..n-1.  The
//      condition a0 u + a1 u_x = g of each end holds at the cell edge: u there by the half-point
//      interpolation of order max(4, p), u_x by the half-point difference of order p
//      (boundary_value_maps, generate_bc_eqs.jl:79-161), each row chosen by
//      get_half_offset_weights_and_stencil (half_offset_centred_difference.jl:9-36) at half point
//      1 (lower edge) or n - 1 (upper edge, newindex(shift = true), MOL_utils.jl:87); the row is
//      solved for the ghost node.  Interior nodes take the node-centred rows of D_xx
//      (centered_difference.jl:5-30).  Constant g.
extern "C" int orc_diffusion1d_edge_rhs(int ncell, double xlo, double xhi, int p, double D, const double* bcl,
                                        const double* bch, const double* u, double t, double* du) {
  (void)t;
  try {
    if (ncell < 4) return -1;
    const int n = ncell + 2;
    const double h = (xhi - xlo) / ncell;
    std::vector<double> v(n + 1, 0.0);
    for (int i = 2; i <= n - 1; ++i) v[i] = u[i - 2];
    auto hrow = [n](const Op& o, int i, int& first) -> const std::vector<double>& {
      if (i <= o.bpc) {
        first = 1;
        return o.low[i - 1];
      }
      if (i > n - o.bpc) {
        first = n - o.BL + 1;
        return o.high[n - i - 1];
      }
      first = i + 1 - o.L / 2;
      return o.st;
    };
    const Op oi = half_op(0, std::max(4, p), h), od = half_op(1, p, h);
    for (int side = 0; side < 2; ++side) {
      const double* b = side ? bch : bcl;
      const int ghost = side ? n : 1, hp = side ? n - 1 : 1;
      std::vector<double> c(n + 1, 0.0);
      int f = 0;
      if (b[0] != 0.0) {
        const std::vector<double>& w = hrow(oi, hp, f);
        for (size_t k = 0; k < w.size(); ++k) c[f + k] += b[0] * w[k];
      }
      if (b[1] != 0.0) {
        const std::vector<double>& w = hrow(od, hp, f);
        for (size_t k = 0; k < w.size(); ++k) c[f + k] += b[1] * w[k];
      }
      if (c[side ? 1 : n] != 0.0) return -1;   // a row reaching the other ghost: grid too small
      double rest = b[2];
      for (int k = 1; k <= n; ++k)
        if (k != ghost) rest -= c[k] * v[k];
      v[ghost] = rest / c[ghost];
    }
    const Op oc = centered(2, p, h);
    for (int i = 2; i <= n - 1; ++i) {
      const std::vector<double>* w;
      int first;
      if (i <= oc.bpc) {
        w = &oc.low[i - 1];
        first = 1;
      } else if (i > n - oc.bpc) {
        w = &oc.high[n - i];
        first = n - oc.BL + 1;
      } else {
        w = &oc.st;
        first = i - oc.L / 2;
      }
      double acc = 0.0;
      for (size_t k = 0; k < w->size(); ++k) acc += (*w)[k] * v[first + k];
      du[i - 2] = D * acc;
    }
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- two variables on adjacent intervals joined by an interface condition c1(t, m) ~ c2(t, m):
//      Dt c_k = D_k Dxx c_k, c1 on n1 nodes of [a, m], c2 on n2 nodes of [m, b] (same spacing
//      dx), Dirichlet c1 = gl at a and c2 = gr at b.  The interface keeps the upper node of c1 as
//      an unknown (clip_interior!!, interior_map.jl:1-10: an upper interface clips nothing) and
//      eliminates the lower node of c2 by c2[1] = c1[n1] (generate_bc_eqs.jl:35-58).  Stencils
//      reaching past the interface read the other variable (_wrapinterface,
//      interface_boundary.jl:79-111): c1 at node i > n1 is c2 at i - n1 + 1, c2 at node k <= 1 is
//      c1 at k + n1 - 1; next to an interface the interior row is used (no boundary rows,
//      centered_difference.jl:5-30 with haslower / hasupper).  State: c1 at 2..n1, then c2 at
//      2..n2-1.
extern "C" int orc_interface1d_rhs(int n1, int n2, double dx, int p, double D1, double D2, double gl, double gr,
                                   const double* u, double t, double* du) {
  (void)t;
  try {
    if (n1 < 3 || n2 < 3) return -1;
    const Op op = centered(2, p, dx);
    const double* u1 = u;             // c1 nodes 2..n1
    const double* u2 = u + (n1 - 1);  // c2 nodes 2..n2-1
    std::function<double(int)> get1, get2;
    get1 = [&](int i) -> double {
      if (i == 1) return gl;
      if (i >= 2 && i <= n1) return u1[i - 2];
      if (i > n1) return get2(i - n1 + 1);
      throw std::runtime_error("c1 read outside its grid");
    };
    get2 = [&](int k) -> double {
      if (k <= 1) return get1(k + n1 - 1);
      if (k <= n2 - 1) return u2[k - 2];
      if (k == n2) return gr;
      throw std::runtime_error("c2 read outside its grid");
    };
    const int bpc = op.bpc, half = op.L / 2;
    for (int i = 2; i <= n1; ++i) {
      // lower end Dirichlet: boundary rows; upper end an interface: interior rows with wrapping
      const std::vector<double>& w = i <= bpc ? op.low[i - 1] : op.st;
      const int first = i <= bpc ? 1 : i - half;
      double s = 0.0;
      for (size_t k = 0; k < w.size(); ++k) s += w[k] * get1(first + (int)k);
      du[i - 2] = D1 * s;
    }
    for (int k = 2; k <= n2 - 1; ++k) {
      // lower end an interface: interior rows; upper end Dirichlet: boundary rows
      const bool high = k > n2 - bpc;
      const std::vector<double>& w = high ? op.high[n2 - k] : op.st;
      const int first = high ? n2 - op.BL + 1 : k - half;
      double s = 0.0;
      for (size_t m = 0; m < w.size(); ++m) s += w[m] * get2(first + (int)m);
      du[(n1 - 1) + (k - 2)] = D2 * s;
    }
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- u_t = D u_xx - c u_x on the nodes x[0..n-1] (nonuniform), approximation order p; the
//      advection term upwinded with order uo (cardinalized coefficient c: backward rows where
//      c > 0, upwind_difference.jl:184-199) or, with uo = 0, centred; a0 u + a1 u_x = g at each end.
int orc_nonuniform1d_rhs(int n, const double* x, double D, double c, int p, int uo, const double* bcl,
                         const double* bch, const double* u, double t, double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, x[0], x[n - 1], false);
    for (int i = 0; i < n; ++i) s.sp.ax[0].x[i] = x[i];
    s.sp.ax[0].nonuniform = true;
    s.sp.ax[0].dx = std::nan("");
    s.F.resize(1);
    Field<double>& f = s.F[0];
    for (int side = 0; side < 2; ++side) {
      const double* b = side ? bch : bcl;
      f.side[0][side].has = true;
      f.side[0][side].a = {b[0], b[1]};
      const double g = b[2];
      f.side[0][side].g = [g](double, const int*) { return g; };
    }
    box_default(s, f);
    f.p = p;
    const std::vector<double> xs(x, x + n);
    auto d2 = std::make_shared<Op>(centered_nu(2, p, xs));
    auto d1 = std::make_shared<Op>(centered_nu(1, p, xs));
    auto fw = std::make_shared<Op>(uo > 0 ? upwind_nu(1, uo, xs, 0) : Op());
    auto bw = std::make_shared<Op>(uo > 0 ? upwind_nu(1, uo, xs, uo) : Op());
    s.node_rhs = [D, c, uo, d2, d1, fw, bw](System<double>& S, const int* I, double, double* out) {
      const double diff = D == 0.0 ? 0.0 : D * central(S.F[0], *d2, I, 0);
      double adv = 0.0;
      if (c != 0.0) {
        if (uo > 0) adv = c > 0 ? c * upwind_at(S.F[0], *fw, *bw, I, 0, true) : c * upwind_at(S.F[0], *fw, *bw, I, 0, false);
        else adv = c * central(S.F[0], *d1, I, 0);
      }
      out[0] = diff - adv;
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- 2D Brusselator on [0,1]^2 periodic, n nodes per axis (docs brusselator; brusselator_eq.jl:8-52)
static System<double>* make_bruss(int n, int p, double alpha) {
  auto* s = new System<double>();
  s->p = p;
  s->sp.nd = 2;
  make_axis(s->sp.ax[0], n, 0.0, 1.0, true);
  make_axis(s->sp.ax[1], n, 0.0, 1.0, true);
  s->F.resize(2);
  for (auto& f : s->F) box_default(*s, f);
  auto op2x = std::make_shared<Op>(centered(2, p, s->sp.ax[0].dx));
  auto op2y = std::make_shared<Op>(centered(2, p, s->sp.ax[1].dx));
  s->node_rhs = [alpha, op2x, op2y](System<double>& S, const int* I, double t, double* out) {
    const double u = S.F[0].get(I[0], I[1], I[2]), v = S.F[1].get(I[0], I[1], I[2]);
    const double x = S.sp.ax[0].x[I[0] - 1], y = S.sp.ax[1].x[I[1] - 1];
    const double lapu = central(S.F[0], *op2x, I, 0) + central(S.F[0], *op2y, I, 1);
    const double lapv = central(S.F[1], *op2x, I, 0) + central(S.F[1], *op2y, I, 1);
    const double forcing = (((x - 0.3) * (x - 0.3) + (y - 0.6) * (y - 0.6)) <= 0.1 * 0.1) * (t >= 1.1) * 5.0;
    out[0] = 1.0 + v * u * u - 4.4 * u + alpha * lapu + forcing;
    out[1] = 3.4 * u - v * u * u + alpha * lapv;
  };
  return s;
}

int orc_bruss2d_rhs(int n, int p, double alpha, const double* u, double t, double* du, int threads) {
  try {
    std::unique_ptr<System<double>> s(make_bruss(n, p, alpha));
    s->rhs(u, t, du, threads);
    return 0;
  } catch (...) {
    return -1;
  }
}

}  // extern "C"

// ---- 2D inviscid Burgers u_t + u u_x + u u_y = 0, WENO5, periodic [0,1]^2
template <typename S>
static System<S>* make_burgers(int n, double eps) {
  auto* s = new System<S>();
  s->sp.nd = 2;
  make_axis(s->sp.ax[0], n, 0.0, 1.0, true);
  make_axis(s->sp.ax[1], n, 0.0, 1.0, true);
  s->F.resize(1);
  box_default(*s, s->F[0]);
  s->node_rhs = [eps](System<S>& Sy, const int* I, double, S* out) {
    const S u = Sy.F[0].get(I[0], I[1], I[2]);
    out[0] = -(u * weno_at(Sy.F[0], I, 0, eps)) - u * weno_at(Sy.F[0], I, 1, eps);
  };
  return s;
}

extern "C" {

int orc_burgers2d_rhs(int n, double eps, const double* u, double t, double* du, int threads) {
  try {
    std::unique_ptr<System<double>> s(make_burgers<double>(n, eps));
    s->rhs(u, t, du, threads);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- 1D WENO advection u_t + c u_x = 0 on [0, L] with Dirichlet u = gl at x = 0 and
//      u = gh at x = L (non-periodic: extent-2 clipping + extrapolated nodes)
int orc_weno1d_dirichlet_rhs(int n, double L, double c, double eps, double gl, double gh, const double* u,
                             double t, double* du) {
  try {
    System<double> s;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, 0.0, L, false);
    s.F.resize(1);
    Field<double>& f = s.F[0];
    f.ext_lo[0] = f.ext_hi[0] = 2;
    f.side[0][0].has = true; f.side[0][0].a = {1.0}; f.side[0][0].g = [gl](double, const int*) { return gl; };
    f.side[0][1].has = true; f.side[0][1].a = {1.0}; f.side[0][1].g = [gh](double, const int*) { return gh; };
    box_default(s, f);
    s.node_rhs = [c, eps](System<double>& S, const int* I, double, double* out) {
      out[0] = -(c * weno_at(S.F[0], I, 0, eps));
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- 2D diffusion-reaction u_t = D Lap u + k u (1 - u) on [0,1]^2, n nodes per axis.
//      bctype[4] = (x-low, x-high, y-low, y-high): 0 Dirichlet u = 0, 1 Neumann du/dn = 0.
int orc_diffreact2d_rhs(int n, double D, double k, const int* bctype, const double* u, double t, double* du) {
  try {
    System<double> s;
    s.sp.nd = 2;
    make_axis(s.sp.ax[0], n, 0.0, 1.0, false);
    make_axis(s.sp.ax[1], n, 0.0, 1.0, false);
    s.F.resize(1);
    Field<double>& f = s.F[0];
    for (int a = 0; a < 2; ++a)
      for (int sd = 0; sd < 2; ++sd) {
        f.side[a][sd].has = true;
        f.side[a][sd].a = bctype[2 * a + sd] == 0 ? std::vector<double>{1.0} : std::vector<double>{0.0, 1.0};
        f.side[a][sd].g = [](double, const int*) { return 0.0; };
      }
    box_default(s, f);
    Op ox = centered(2, 2, s.sp.ax[0].dx), oy = centered(2, 2, s.sp.ax[1].dx);
    s.node_rhs = [D, k, ox, oy](System<double>& S, const int* I, double, double* out) {
      const double uu = S.F[0].get(I[0], I[1], I[2]);
      out[0] = D * (central(S.F[0], ox, I, 0) + central(S.F[0], oy, I, 1)) + k * uu * (1.0 - uu);
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- u_t = D sum_a Daa u + k u (1 - u) in nd = 1..3 dimensions on [0,1]^nd; axis a has n[a]
//      nodes, uniform, or the given nodes x[a] (nonuniform) when x[a] is not null; approximation
//      order p; side (a, s) Dirichlet 0 (bctype 0) or Neumann 0 (bctype 1).
int orc_diffreact_nd_rhs(int nd, const int* n, const double* const* x, int p, double D, double k,
                         const int* bctype, const double* u, double t, double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = nd;
    std::vector<std::shared_ptr<Op>> ops;
    for (int a = 0; a < nd; ++a) {
      make_axis(s.sp.ax[a], n[a], 0.0, 1.0, false);
      if (x && x[a]) {
        for (int i = 0; i < n[a]; ++i) s.sp.ax[a].x[i] = x[a][i];
        s.sp.ax[a].nonuniform = true;
        s.sp.ax[a].dx = std::nan("");
        ops.push_back(std::make_shared<Op>(centered_nu(2, p, s.sp.ax[a].x)));
      } else {
        ops.push_back(std::make_shared<Op>(centered(2, p, s.sp.ax[a].dx)));
      }
    }
    s.F.resize(1);
    Field<double>& f = s.F[0];
    for (int a = 0; a < nd; ++a)
      for (int sd = 0; sd < 2; ++sd) {
        f.side[a][sd].has = true;
        f.side[a][sd].a = bctype[2 * a + sd] == 0 ? std::vector<double>{1.0} : std::vector<double>{0.0, 1.0};
        f.side[a][sd].g = [](double, const int*) { return 0.0; };
      }
    box_default(s, f);
    f.p = p;
    s.node_rhs = [D, k, ops, nd](System<double>& S, const int* I, double, double* out) {
      const double uu = S.F[0].get(I[0], I[1], I[2]);
      double lap = 0.0;
      for (int a = 0; a < nd; ++a) lap += central(S.F[0], *ops[a], I, a);
      out[0] = D * lap + k * uu * (1.0 - uu);
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

}  // extern "C"

// ---- 3D nonlinear diffusion u_t = sum_a Da(d0 (1 + u^2) Da u) + s on [0,1]^3, Neumann 0.
//      d0, src: full-grid arrays (n^3, first axis fastest).
template <typename S>
static System<S>* make_nonlin3d(int n, int p, const S* d0, const double* src) {
  auto* s = new System<S>();
  s->p = p;
  s->sp.nd = 3;
  for (int a = 0; a < 3; ++a) make_axis(s->sp.ax[a], n, 0.0, 1.0, false);
  s->F.resize(1);
  Field<S>& f = s->F[0];
  for (int a = 0; a < 3; ++a)
    for (int sd = 0; sd < 2; ++sd) {
      f.side[a][sd].has = true;
      f.side[a][sd].a = {0.0, 1.0};
      f.side[a][sd].g = [](double, const int*) { return 0.0; };
    }
  box_default(*s, f);
  s->aux.resize(2);
  for (auto& ax : s->aux) { ax.sp = &s->sp; ax.U.resize(s->sp.size()); }
  for (size_t q = 0; q < s->sp.size(); ++q) { s->aux[0].U[q] = d0[q]; s->aux[1].U[q] = S(src[q]); }
  s->node_rhs = [](System<S>& Sy, const int* I, double, S* out) {
    std::vector<const Field<S>*> fl{&Sy.F[0]}, ax{&Sy.aux[0]};
    CoefFn<S> coef = [](const S* u, double, const int*, const S* a) { return a[0] * (S(1.0) + u[0] * u[0]); };
    S acc(0.0);
    for (int a = 0; a < 3; ++a) acc += nonlinlap(fl, 0, ax, coef, I, a, Sy.p);
    out[0] = acc + Sy.aux[1].get(I[0], I[1], I[2]);
  };
  return s;
}

extern "C" {

int orc_nonlindiff3d_rhs(int n, int p, const double* d0, const double* src, const double* u, double t, double* du,
                         int threads) {
  try {
    std::unique_ptr<System<double>> s(make_nonlin3d<double>(n, p, d0, src));
    s->rhs(u, t, du, threads);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- 1D advection u_t + c(x) u_x = 0 with c(x) = c0 + c1 sin(2 pi x) on [0,1], upwind of order `uo`,
//      Dirichlet u = gl at 0 and u = gh at 1
int orc_upwind1d_rhs(int n, int uo, double c0, double c1, double gl, double gh, const double* u, double t, double* du) {
  try {
    System<double> s;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, 0.0, 1.0, false);
    s.F.resize(1);
    Field<double>& f = s.F[0];
    f.side[0][0].has = true; f.side[0][0].a = {1.0}; f.side[0][0].g = [gl](double, const int*) { return gl; };
    f.side[0][1].has = true; f.side[0][1].a = {1.0}; f.side[0][1].g = [gh](double, const int*) { return gh; };
    box_default(s, f);
    const double dx = s.sp.ax[0].dx;
    Op fwd = upwind(1, uo, dx, 0), bwd = upwind(1, uo, dx, uo);
    s.node_rhs = [c0, c1, fwd, bwd](System<double>& S, const int* I, double, double* out) {
      const double x = S.sp.ax[0].x[I[0] - 1];
      const double c = c0 + c1 * std::sin(2.0 * M_PI * x);
      // cardinalized residual term c * Dx(u): ifelse(c > 0, c*D_backward, c*D_forward)
      const double term = c > 0 ? c * upwind_at(S.F[0], fwd, bwd, I, 0, true) : c * upwind_at(S.F[0], fwd, bwd, I, 0, false);
      out[0] = -term;
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- 2D u_t = Dxx u + Dyy u + c Dx Dy u on [0,1]^2, Dirichlet u = g(x,y) = x + 2y on all sides
int orc_mixed2d_rhs(int n, int p, double c, const double* u, double t, double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = 2;
    make_axis(s.sp.ax[0], n, 0.0, 1.0, false);
    make_axis(s.sp.ax[1], n, 0.0, 1.0, false);
    s.F.resize(1);
    Field<double>& f = s.F[0];
    Space* spp = &s.sp;
    for (int a = 0; a < 2; ++a)
      for (int sd = 0; sd < 2; ++sd) {
        f.side[a][sd].has = true;
        f.side[a][sd].a = {1.0};
        f.side[a][sd].g = [spp](double, const int* I) { return spp->ax[0].x[I[0] - 1] + 2.0 * spp->ax[1].x[I[1] - 1]; };
      }
    box_default(s, f);
    Op d2x = centered(2, p, s.sp.ax[0].dx), d2y = centered(2, p, s.sp.ax[1].dx);
    Op d1x = centered(1, p, s.sp.ax[0].dx), d1y = centered(1, p, s.sp.ax[1].dx);
    s.node_rhs = [c, d2x, d2y, d1x, d1y](System<double>& S, const int* I, double, double* out) {
      out[0] = central(S.F[0], d2x, I, 0) + central(S.F[0], d2y, I, 1) + c * mixed_at(S.F[0], d1x, d1y, I, 0, 1);
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- the same system on any mix of uniform and nonuniform axes (x[a] NULL: n[a] uniform nodes on
//      [0,1], else the nodes), Dirichlet u = q(x, y) = q0 + q1 x + q2 y + q3 x^2 + q4 x y + q5 y^2
//      on all sides.  Interior rows of a nonuniform axis per node (centered_nu), as
//      mixed_central_difference takes them (2nd_order_mixed_deriv.jl:5-22 through
//      central_difference_weights_and_stencil on nodes, centered_difference.jl:32-57).
int orc_mixed2d_grid_rhs(const int* n, const double* const* x, int p, double c, const double* q, const double* u,
                         double t, double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = 2;
    Op d2[2], d1[2];
    for (int a = 0; a < 2; ++a) {
      make_axis(s.sp.ax[a], n[a], 0.0, 1.0, false);
      if (x && x[a]) {
        for (int i = 0; i < n[a]; ++i) s.sp.ax[a].x[i] = x[a][i];
        s.sp.ax[a].nonuniform = true;
        s.sp.ax[a].dx = std::nan("");
        d2[a] = centered_nu(2, p, s.sp.ax[a].x);
        d1[a] = centered_nu(1, p, s.sp.ax[a].x);
      } else {
        d2[a] = centered(2, p, s.sp.ax[a].dx);
        d1[a] = centered(1, p, s.sp.ax[a].dx);
      }
    }
    s.F.resize(1);
    Field<double>& f = s.F[0];
    Space* spp = &s.sp;
    const std::array<double, 6> qc = {q[0], q[1], q[2], q[3], q[4], q[5]};
    for (int a = 0; a < 2; ++a)
      for (int sd = 0; sd < 2; ++sd) {
        f.side[a][sd].has = true;
        f.side[a][sd].a = {1.0};
        f.side[a][sd].g = [spp, qc](double, const int* I) {
          const double X = spp->ax[0].x[I[0] - 1], Y = spp->ax[1].x[I[1] - 1];
          return qc[0] + qc[1] * X + qc[2] * Y + qc[3] * X * X + qc[4] * X * Y + qc[5] * Y * Y;
        };
      }
    box_default(s, f);
    f.p = p;
    s.node_rhs = [c, d2, d1](System<double>& S, const int* I, double, double* out) {
      out[0] = central(S.F[0], d2[0], I, 0) + central(S.F[0], d2[1], I, 1) + c * mixed_at(S.F[0], d1[0], d1[1], I, 0, 1);
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- 1D spherical Laplacian u_t = 1/r^2 Dr(r^2 Dr u) on [0,1]: Neumann Dr u = 0 at r = 0,
//      Dirichlet u = g1 at r = 1
int orc_spherical1d_rhs(int n, int p, double g1, const double* u, double t, double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, 0.0, 1.0, false);
    s.F.resize(1);
    Field<double>& f = s.F[0];
    f.side[0][0].has = true; f.side[0][0].a = {0.0, 1.0}; f.side[0][0].g = [](double, const int*) { return 0.0; };
    f.side[0][1].has = true; f.side[0][1].a = {1.0}; f.side[0][1].g = [g1](double, const int*) { return g1; };
    box_default(s, f);
    Op d1 = centered(1, p, s.sp.ax[0].dx), d2 = centered(2, p, s.sp.ax[0].dx);
    s.node_rhs = [d1, d2](System<double>& S, const int* I, double, double* out) {
      const double r = S.sp.ax[0].x[I[0] - 1];
      const double a = 1.0;
      if (std::fabs(r) <= 1e-6) { out[0] = 6.0 * a * central(S.F[0], d2, I, 0); return; }
      std::vector<const Field<double>*> fl{&S.F[0]}, ax;
      CoefFn<double> coef = [](const double*, double, const int*, const double*) { return 1.0; };
      out[0] = a * (central(S.F[0], d1, I, 0) / r + nonlinlap(fl, 0, ax, coef, I, 0, S.p));
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- the same two systems on nodes x[0..n-1] spanning [0, 1] (nonuniform; Dirichlet only for the
//      nonlinear diffusion: the reference has no periodic nonuniform half-point operators)
int orc_nonlinlap1d_nu_rhs(int n, const double* x, int p, double gl, double gh, const double* u, double t,
                           double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, x[0], x[n - 1], false);
    for (int i = 0; i < n; ++i) s.sp.ax[0].x[i] = x[i];
    s.sp.ax[0].nonuniform = true;
    s.sp.ax[0].dx = std::nan("");
    s.F.resize(1);
    Field<double>& f = s.F[0];
    f.side[0][0].has = true; f.side[0][0].a = {1.0}; f.side[0][0].g = [gl](double, const int*) { return gl; };
    f.side[0][1].has = true; f.side[0][1].a = {1.0}; f.side[0][1].g = [gh](double, const int*) { return gh; };
    box_default(s, f);
    s.node_rhs = [](System<double>& S, const int* I, double, double* out) {
      std::vector<const Field<double>*> fl{&S.F[0]}, ax;
      CoefFn<double> coef = [](const double* uu, double, const int*, const double*) { return 1.0 + uu[0] * uu[0]; };
      out[0] = nonlinlap(fl, 0, ax, coef, I, 0, S.p);
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

int orc_spherical1d_nu_rhs(int n, const double* x, int p, double g1, const double* u, double t, double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, x[0], x[n - 1], false);
    for (int i = 0; i < n; ++i) s.sp.ax[0].x[i] = x[i];
    s.sp.ax[0].nonuniform = true;
    s.sp.ax[0].dx = std::nan("");
    s.F.resize(1);
    Field<double>& f = s.F[0];
    f.side[0][0].has = true; f.side[0][0].a = {0.0, 1.0}; f.side[0][0].g = [](double, const int*) { return 0.0; };
    f.side[0][1].has = true; f.side[0][1].a = {1.0}; f.side[0][1].g = [g1](double, const int*) { return g1; };
    box_default(s, f);
    const std::vector<double> xs(x, x + n);
    auto d1 = std::make_shared<Op>(centered_nu(1, p, xs));
    auto d2 = std::make_shared<Op>(centered_nu(2, p, xs));
    s.node_rhs = [d1, d2](System<double>& S, const int* I, double, double* out) {
      const double r = S.sp.ax[0].x[I[0] - 1];
      if (std::fabs(r) <= 1e-6) { out[0] = 6.0 * central(S.F[0], *d2, I, 0); return; }
      std::vector<const Field<double>*> fl{&S.F[0]}, ax;
      CoefFn<double> coef = [](const double*, double, const int*, const double*) { return 1.0; };
      out[0] = central(S.F[0], *d1, I, 0) / r + nonlinlap(fl, 0, ax, coef, I, 0, S.p);
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- 1D nonlinear diffusion u_t = Dx((1 + u^2) Dx u) on [0,1]; periodic (per=1) or Dirichlet
//      u(0) = gl, u(1) = gh
int orc_nonlinlap1d_rhs(int n, int p, int per, double gl, double gh, const double* u, double t, double* du) {
  try {
    System<double> s;
    s.p = p;
    s.sp.nd = 1;
    make_axis(s.sp.ax[0], n, 0.0, 1.0, per != 0);
    s.F.resize(1);
    Field<double>& f = s.F[0];
    if (!per) {
      f.side[0][0].has = true; f.side[0][0].a = {1.0}; f.side[0][0].g = [gl](double, const int*) { return gl; };
      f.side[0][1].has = true; f.side[0][1].a = {1.0}; f.side[0][1].g = [gh](double, const int*) { return gh; };
    }
    box_default(s, f);
    s.node_rhs = [](System<double>& S, const int* I, double, double* out) {
      std::vector<const Field<double>*> fl{&S.F[0]}, ax;
      CoefFn<double> coef = [](const double* uu, double, const int*, const double*) { return 1.0 + uu[0] * uu[0]; };
      out[0] = nonlinlap(fl, 0, ax, coef, I, 0, S.p);
    };
    s.rhs(u, t, du, 1);
    return 0;
  } catch (...) {
    return -1;
  }
}

// ---- Runge-Kutta drivers over the systems above (textbook tableaus)
typedef int (*rhs_fn)(void* ctx, const double* u, double t, double* du);

static int rk_drive(rhs_fn f, void* ctx, size_t N, double* u, double t0, double dt, long long nsteps, int scheme) {
  std::vector<double> k1(N), k2(N), k3(N), k4(N), y(N);
  double t = t0;
  for (long long s = 0; s < nsteps; ++s) {
    if (scheme == 0) {
      if (f(ctx, u, t, k1.data())) return -1;
      for (size_t i = 0; i < N; ++i) y[i] = u[i] + dt / 2 * k1[i];
      if (f(ctx, y.data(), t + dt / 2, k2.data())) return -1;
      for (size_t i = 0; i < N; ++i) y[i] = u[i] + dt / 2 * k2[i];
      if (f(ctx, y.data(), t + dt / 2, k3.data())) return -1;
      for (size_t i = 0; i < N; ++i) y[i] = u[i] + dt * k3[i];
      if (f(ctx, y.data(), t + dt, k4.data())) return -1;
      for (size_t i = 0; i < N; ++i) u[i] = u[i] + dt / 6 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
    } else {
      if (f(ctx, u, t, k1.data())) return -1;
      for (size_t i = 0; i < N; ++i) y[i] = u[i] + dt * k1[i];
      if (f(ctx, y.data(), t + dt, k2.data())) return -1;
      for (size_t i = 0; i < N; ++i) k3[i] = 0.75 * u[i] + 0.25 * (y[i] + dt * k2[i]);
      if (f(ctx, k3.data(), t + dt / 2, k4.data())) return -1;
      for (size_t i = 0; i < N; ++i) u[i] = u[i] / 3.0 + 2.0 / 3.0 * (k3[i] + dt * k4[i]);
    }
    t += dt;
  }
  return 0;
}

struct Ctx1 { int n, p; double D; double bcl[3], bch[3]; };
static int heat_cb(void* c, const double* u, double t, double* du) {
  Ctx1* x = (Ctx1*)c;
  return orc_diffusion1d_rhs(x->n, 0.0, 1.0, x->p, x->D, x->bcl, x->bch, u, t, du);
}
int orc_diffusion1d_integrate(int n, int p, double D, const double* bcl, const double* bch, double* u, double t0,
                              double dt, long long nsteps, int scheme) {
  Ctx1 c{n, p, D, {bcl[0], bcl[1], bcl[2]}, {bch[0], bch[1], bch[2]}};
  int nu = 0;
  {
    // number of unknowns: nodes 2..n-1 (both ends carry a condition)
    nu = n - 2;
  }
  return rk_drive(heat_cb, &c, (size_t)nu, u, t0, dt, nsteps, scheme);
}

struct CtxB { int n, p; double alpha; int threads; };
static int bruss_cb(void* c, const double* u, double t, double* du) {
  CtxB* x = (CtxB*)c;
  return orc_bruss2d_rhs(x->n, x->p, x->alpha, u, t, du, x->threads);
}
int orc_bruss2d_integrate(int n, int p, double alpha, double* u, double t0, double dt, long long nsteps, int threads) {
  CtxB c{n, p, alpha, threads};
  const size_t N = 2 * (size_t)(n - 1) * (n - 1);
  return rk_drive(bruss_cb, &c, N, u, t0, dt, nsteps, 0);
}

struct CtxW { int n; double eps; int threads; };
static int burgers_cb(void* c, const double* u, double t, double* du) {
  CtxW* x = (CtxW*)c;
  return orc_burgers2d_rhs(x->n, x->eps, u, t, du, x->threads);
}
int orc_burgers2d_integrate(int n, double eps, double* u, double t0, double dt, long long nsteps, int threads) {
  CtxW c{n, eps, threads};
  const size_t N = (size_t)(n - 1) * (n - 1);
  return rk_drive(burgers_cb, &c, N, u, t0, dt, nsteps, 1);
}

// ---- forward-mode directional derivative of the 3D nonlinear-diffusion final state with
//      respect to d0 along direction v (the ForwardDiff check of the adjoint gradient):
//      integrates (u, du/de) with RK4 for nsteps and returns L = 1/2 |u - target|^2 and dL/de.
int orc_nonlindiff3d_jvp(int n, int p, const double* d0, const double* vdir, const double* src,
                         const double* u0, const double* target, double dt, long long nsteps, double* loss,
                         double* dloss) {
  try {
    const size_t G = (size_t)n * n * n;
    std::vector<Dual> d0d(G);
    for (size_t q = 0; q < G; ++q) d0d[q] = Dual(d0[q], vdir[q]);
    std::unique_ptr<System<Dual>> s(make_nonlin3d<Dual>(n, p, d0d.data(), src));
    const size_t N = s->nunk();
    std::vector<Dual> u(N), k1(N), k2(N), k3(N), k4(N), y(N);
    for (size_t i = 0; i < N; ++i) u[i] = Dual(u0[i], 0.0);
    double t = 0.0;
    for (long long st = 0; st < nsteps; ++st) {
      s->rhs(u.data(), t, k1.data(), 1);
      for (size_t i = 0; i < N; ++i) y[i] = u[i] + Dual(dt / 2) * k1[i];
      s->rhs(y.data(), t + dt / 2, k2.data(), 1);
      for (size_t i = 0; i < N; ++i) y[i] = u[i] + Dual(dt / 2) * k2[i];
      s->rhs(y.data(), t + dt / 2, k3.data(), 1);
      for (size_t i = 0; i < N; ++i) y[i] = u[i] + Dual(dt) * k3[i];
      s->rhs(y.data(), t + dt, k4.data(), 1);
      for (size_t i = 0; i < N; ++i)
        u[i] = u[i] + Dual(dt / 6) * (k1[i] + Dual(2.0) * k2[i] + Dual(2.0) * k3[i] + k4[i]);
      t += dt;
    }
    double L = 0.0, dL = 0.0;
    for (size_t i = 0; i < N; ++i) {
      const double r = u[i].v - target[i];
      L += 0.5 * r * r;
      dL += r * u[i].d;
    }
    *loss = L;
    *dloss = dL;
    return 0;
  } catch (...) {
    return -1;
  }
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
