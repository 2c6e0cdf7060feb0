/*
 * mol.h — C ABI of the MI355X (gfx950) method-of-lines stencil library.
 *
 * The library evaluates the semi-discrete right-hand side f(du, u, p, t) that
 * MethodOfLines.jl's `discretize(pdesys, MOLFiniteDifference(...))` produces, and
 * drives it with explicit Runge-Kutta stage loops, adjoints and batched launches.
 *
 * A problem is handed over as a plain "stencil IR" descriptor (mol_desc): the grid,
 * the unknown box of every dependent variable, per-axis operator weight tables,
 * operator slots (centered / upwind / WENO5 / nonlinear Laplacian / spherical /
 * mixed), ghost-node closures for Dirichlet / Neumann / Robin / extrapolated
 * boundary nodes, and small bytecode programs for the pointwise part of each
 * equation.  The host front-end builds the descriptor from a PDE system, the way
 * the reference builds its symbolic equations (src/scalar_discretization.jl:1-64).
 *
 * Conventions
 *   - every function returns 0 on success and a negative mol_status otherwise;
 *     mol_last_error() gives the message.  No exception crosses this boundary.
 *   - pointers named *_h are host pointers; all other array pointers are device
 *     pointers.  `stream` is a hipStream_t passed as void* (NULL = default stream).
 *   - calls are asynchronous on `stream` unless named *_h (those synchronize).
 *   - one handle must not be used from two threads at once, and work it enqueues on two
 *     streams must not overlap: its workspace (stage vectors, reverse-mode buffers, norm
 *     partials) is shared by all of its calls.
 *   - state vectors use the reference's unknown ordering: variable-major, and within
 *     a variable the first spatial axis fastest (docs/src/generated/bruss_code.md:48-79,
 *     src/interface/solution/timedep.jl:30-47).
 */
#ifndef MOL_H
#define MOL_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOL_ABI_VERSION 3
#define MOL_MAX_DIMS 3
#define MOL_MAX_FIELDS 8
#define MOL_MAX_AUX 8
#define MOL_MAX_SLOTS 64
#define MOL_MAX_CLOSURE_K 24
#define MOL_MAX_GHOST 6

typedef enum {
  MOL_OK = 0,
  MOL_ERR_INVALID = -1,   /* malformed descriptor or argument */
  MOL_ERR_SHAPE = -2,     /* size mismatch / out-of-range sizes */
  MOL_ERR_HIP = -3,       /* a HIP runtime call failed */
  MOL_ERR_UNSUPPORTED = -4,
  MOL_ERR_NOMEM = -5
} mol_status;

typedef enum { MOL_F64 = 0, MOL_F32 = 1 } mol_dtype;

/* ---- operator weight tables -------------------------------------------------
 * One table per (operator, axis), weights already scaled by 1/dx^d.
 * Interior rows: taps at i + off + k, k = 0..len-1, weights pool[w .. w+len).
 * Rows near a non-periodic lower side: for i <= nlow, row r = i, taps 1..blen,
 * weights pool[wlow + (r-1)*blen ..).  Near a non-periodic upper side, with
 * lref = n + ref_shift: for i > lref - nhigh, row r = lref - i + hoff, taps
 * lref-blen+1 .. lref, weights pool[whigh + (r-1)*blen ..).
 * This is the row-selection logic of central_difference_weights_and_stencil
 * (src/discretization/schemes/centered_difference/centered_difference.jl:5-30),
 * _upwind_difference (schemes/upwind_difference/upwind_difference.jl:1-28) and
 * get_half_offset_weights_and_stencil (schemes/half_offset_centred_difference.jl:9-37).
 * Indices are 1-based full-grid node (or half-point) indices, as in the reference.
 * Nonuniform axes (pn_hi > 0): the interior row differs per node; node i in pn_lo..pn_hi
 * takes weights pool[w + (i - pn_lo)*len ..), taps at i + off + k as above; interior nodes
 * outside that range have no row (the reference's operator has none there).  Built as
 * CompleteCenteredDifference / CompleteUpwindDifference on nodes
 * (centered_diff_weights.jl:77-153, upwind_diff_weights.jl:93-171); generated kernels only. */
typedef struct {
  int32_t len, off, w;
  int32_t nlow, nhigh, blen, wlow, whigh;
  int32_t hoff;       /* 1: node-centred high rows (r = lref-i+1); 0: half-offset (r = lref-i) */
  int32_t ref_shift;  /* lref = n + ref_shift (0 normally; -1 for the nonlinear-Laplacian outer operator) */
  int32_t pn_lo, pn_hi; /* per-node interior rows (nonuniform axis) for nodes pn_lo..pn_hi; 0, 0: one row */
} mol_optab;

/* ---- operator slots ---------------------------------------------------------- */
typedef enum {
  MOL_SLOT_CENTERED = 0,  /* tab[0]: D^d along axis (generate_cartesian_rules, centered_difference.jl:68-97) */
  MOL_SLOT_UPWIND = 1,    /* tab[0]: forward (coef <= 0), tab[1]: backward (coef > 0); prog: coefficient,
                             or -1 for a bare odd derivative (always backward) (upwind_difference.jl:169-199,278-302);
                             the slot value is coef*D (coef is part of the slot) unless prog == -1 */
  MOL_SLOT_WENO = 2,      /* Jiang-Shu WENO5 first derivative, eps = eps (WENO.jl:6-57) */
  MOL_SLOT_NONLINLAP = 3, /* d/dx(a(...) d/dx u): tab[0]=outer, tab[1]=inner, tab[2]=interp; prog = a
                             (nonlinear_laplacian.jl:28-103) */
  MOL_SLOT_SPHERICAL = 4, /* r^-2 d/dr(r^2 a d/dr u): tab[0]=D1, tab[1]=D2, tab[2..4]=outer,inner,interp;
                             prog = a (spherical_laplacian.jl:10-42) */
  MOL_SLOT_MIXED = 5,     /* d2/dx dy u: tab[0]=D1 along axis, tab[1]=D1 along axis2 (2nd_order_mixed_deriv.jl:5-22) */
  MOL_SLOT_WENO_NU = 6    /* node-centred WENO5 first derivative on the axis's (nonuniform) coordinates, eps = eps
                             (weno_f_nonuniform, schemes/WENO/nonuniform_weno.jl:120-212): targets 1, 2 on taps
                             1..5 and 4, 5 on taps n-4..n at non-periodic ends, target 3 on i-2..i+2 elsewhere
                             (get_f_taps_coords, schemes/function_scheme/function_scheme.jl:1-34); generated
                             kernels only */
} mol_slot_kind;

typedef struct {
  int32_t kind;
  int32_t field;          /* operand dependent variable */
  int32_t axis, axis2;
  int32_t tab[5];
  int32_t prog;           /* program id or -1 */
  double eps;             /* WENO epsilon */
  double inv_dx;          /* WENO 1/dx */
} mol_slot;

/* ---- ghost-node closures ------------------------------------------------------
 * For a non-periodic side of axis `axis` of variable `field`, the nodes between the
 * domain edge and the unknown box (the boundary node, plus extrapolated nodes when the
 * unknown box is clipped further) are affine in the first K unknowns of the same line:
 *   u[ghost g] = sum_{m<K} C[g*K+m] * u[first + m*dir] + c[g] * G(t, boundary coords)
 * where `first` is the unknown next to the side and dir = +1 (lower) / -1 (upper).
 * ghost g = 0 is the node adjacent to the box, g = nghost-1 the domain edge node.
 * G is program `gprog` (or 0 when gprog = -1).  This is what mtkcompile's alias
 * elimination leaves of the BC / extrapolation equations of
 * src/discretization/generate_bc_eqs.jl:24-58,238-311,336-392.
 * An interface between two variables on adjacent ranges of one axis (src >= 0, K = 0,
 * gprog = -1): the nghost positions past this side of the box hold variable `src` at the same
 * position (_wrapinterface, src/discretization/interface_boundary.jl:79-111; the eliminated
 * lower node of the upper variable is the alias of generate_bc_eqs.jl:35-58). */
typedef struct {
  int32_t field, axis, side;  /* side 0 = lower, 1 = upper */
  int32_t nghost, K;
  int32_t wC;                 /* pool offset of C (nghost*K) */
  int32_t wc;                 /* pool offset of c (nghost) */
  int32_t gprog;              /* program id of G or -1 */
  int32_t src;                /* interface: the variable read past this side, else -1 */
} mol_closure;

/* ---- pointwise programs ---------------------------------------------------------
 * Register-machine bytecode, one 32-bit word per instruction:
 *   bits 0..7 opcode, 8..11 dst register, 12..15 a register, 16..19 b register,
 *   bit 20: b is constant pool[imm] instead of a register, bits 21..31 imm.
 * Programs are [start, start+len) ranges of the code pool; the result is register 0. */
typedef enum {
  MOL_OP_CONST = 0,  /* r[d] = cpool[imm] */
  MOL_OP_FIELD = 1,  /* r[d] = value of variable imm in the current context */
  MOL_OP_SLOT = 2,   /* r[d] = operator slot imm at the node */
  MOL_OP_COORD = 3,  /* r[d] = coordinate along axis imm */
  MOL_OP_TIME = 4,   /* r[d] = t */
  MOL_OP_PARAM = 5,  /* r[d] = p[imm] */
  MOL_OP_AUX = 6,    /* r[d] = auxiliary grid array imm */
  MOL_OP_DINNER = 7, /* r[d] = inner half-point derivative (nonlinear Laplacian context) */
  MOL_OP_MOV = 8,
  MOL_OP_ADD = 16, MOL_OP_SUB = 17, MOL_OP_MUL = 18, MOL_OP_DIV = 19,
  MOL_OP_POW = 20, MOL_OP_MIN = 21, MOL_OP_MAX = 22,
  MOL_OP_LT = 23, MOL_OP_LE = 24, MOL_OP_GT = 25, MOL_OP_GE = 26, MOL_OP_EQ = 27, MOL_OP_NE = 28,
  MOL_OP_AND = 29, MOL_OP_OR = 30,
  MOL_OP_SELECT = 31, /* r[d] = r[d] != 0 ? r[a] : b   (d holds the condition) */
  MOL_OP_FMA = 32,    /* r[d] = r[a] * b + r[d] */
  MOL_OP_NEG = 48, MOL_OP_ABS = 49, MOL_OP_SQRT = 50, MOL_OP_EXP = 51, MOL_OP_LOG = 52,
  MOL_OP_SIN = 53, MOL_OP_COS = 54, MOL_OP_TAN = 55, MOL_OP_TANH = 56, MOL_OP_SINH = 57,
  MOL_OP_COSH = 58, MOL_OP_ATAN = 59, MOL_OP_ASIN = 60, MOL_OP_ACOS = 61, MOL_OP_SIGN = 62,
  MOL_OP_POWI = 63,   /* r[d] = r[a]^imm (imm signed 11-bit) */
  MOL_OP_ERF = 64, MOL_OP_LOG10 = 65, MOL_OP_CBRT = 66, MOL_OP_INV = 67, MOL_OP_SQ = 68,
  MOL_OP_EXPM1 = 69, MOL_OP_LOG1P = 70, MOL_OP_FLOOR = 71, MOL_OP_CEIL = 72,
  MOL_OP_HEAVISIDE = 73
} mol_opcode;

#define MOL_NREG 8

/* ---- the descriptor -------------------------------------------------------------- */
typedef struct {
  int32_t abi_version;       /* MOL_ABI_VERSION */
  int32_t ndim;              /* 1..3 spatial axes */
  int32_t nfields;           /* dependent variables (ODE states) */
  int32_t naux;              /* auxiliary full-grid arrays (e.g. a learnable diffusivity field) */
  int32_t dtype;             /* mol_dtype of state and arithmetic */
  int32_t nparams;           /* scalar runtime parameters p[] */
  int64_t n[MOL_MAX_DIMS];   /* grid nodes per axis, both ends included */
  int32_t periodic[MOL_MAX_DIMS];
  const double* coord_h[MOL_MAX_DIMS]; /* node coordinates per axis (n[d] values) */
  int64_t lo[MOL_MAX_FIELDS][MOL_MAX_DIMS]; /* unknown box per variable, 1-based inclusive */
  int64_t hi[MOL_MAX_FIELDS][MOL_MAX_DIMS];
  int32_t halo[MOL_MAX_DIMS];               /* largest tap reach of any slot along each axis */
  int32_t nslots;        const mol_slot* slots_h;
  int32_t ntabs;         const mol_optab* tabs_h;
  int32_t nclosures;     const mol_closure* closures_h;
  int32_t npool;         const double* pool_h;      /* weights / closure coefficients */
  int32_t ncode;         const uint32_t* code_h;    /* bytecode pool */
  int32_t ncpool;        const double* cpool_h;     /* program constants */
  int32_t nprogs;        const int32_t* prog_h;     /* nprogs pairs (start, len) */
  int32_t rhs_prog[MOL_MAX_FIELDS];                 /* program of each variable's right-hand side */
} mol_desc;

typedef struct mol_handle mol_handle;

/* Replaces: the ODEFunction built at src/MOL_discretization.jl:171-201 (f = the
 * RuntimeGeneratedFunction of ModelingToolkit's codegen).  Validates and uploads the
 * descriptor, selects kernels. */
int mol_create(const mol_desc* d, mol_handle** out);

/* Replaces: f(du, u, p, t) — the in-place generated RHS (docs/src/generated/bruss_code.md:46-118,
 * called by OrdinaryDiffEq through prob.f).  u, du: device arrays of unknowns;
 * p: device array of nparams scalars (may be NULL when nparams == 0), of the state dtype;
 * aux: device pointers of the naux full-grid arrays (host array of pointers, may be NULL). */
int mol_rhs(mol_handle* h, void* du, const void* u, const void* p, double t, void* stream);

/* Host-pointer convenience for the literal ODEFunction drop-in (the integrator owns u and du as
 * host arrays and calls f serially, src/MOL_discretization.jl:171-201): copies u (and p) in,
 * runs, copies du out.  Uses device copies owned by the handle (allocated on the first call,
 * released by mol_destroy) and a stream of the handle that is ordered after the legacy default
 * stream; returns when du_h holds the result (only that stream is synchronized).  Calls on one
 * handle must not overlap. */
int mol_rhs_h(mol_handle* h, void* du_h, const void* u_h, const void* p_h, double t);

/* Replaces: the reshaping of a solution into full-grid arrays per variable in
 * PDETimeSeriesSolution (src/interface/solution/timedep.jl:19-93), whose boundary nodes are
 * observed values of the eliminated boundary equations (timedep.jl:46, generate_bc_eqs.jl:238-392)
 * and whose corner / edge nodes are 0 (generate_bc_eqs.jl:396-416).  Writes, on `stream`, the
 * full grid of every variable from the unknowns u at time t: full[f*G + i0 + n0*(i1 + n1*i2)]
 * (G = n0*n1*n2): node 1 of a periodic axis is first replaced by node n (the ring of unknowns
 * is 2..n), then unknown-box nodes are copied, ghost nodes outside the box along one
 * non-periodic axis come from their closures and nodes outside it along two or more axes are
 * 0.  Device pointers; full holds nfields*G values of the handle's type. */
int mol_expand(mol_handle* h, void* full, const void* u, const void* p, double t, void* stream);

/* Batch of instances (config: many uneven PDE instances advanced together).  No counterpart in
 * the reference, which solves one ODEProblem per discretize call; replaces a loop of
 * discretize/solve over instances (e.g. an EnsembleProblem of MOL problems).  The handle's state
 * vector is the concatenation of the instances' state vectors (instance-major, each in its own
 * unknown order, first offsets from mol_batch_offsets); mol_rhs, mol_rhs_h and mol_integrate
 * then act on all instances with one launch per stage (a work-group per tile of an instance,
 * largest tiles first).  Instances must share dtype, dimension and number of variables and take no grid fields; params[i] holds the nparams values
 * of instance i (host, FP64; fixed for the handle's life, the p argument of mol_rhs/mol_integrate
 * is ignored).  Instances of one structure (same operators and boundary kinds) share generated
 * code; numbers (grid sizes, weights, coefficients) are read per instance.  tile = tx,ty,tz for
 * every instance, or NULL. */
int mol_create_batched(const mol_desc* const* ds, const double* const* params, int32_t ninst, const int32_t* tile,
                       mol_handle** out);

/* Generated source of the kernel of a batch (no GPU needed); flags as mol_codegen. */
int64_t mol_codegen_batched(const mol_desc* const* ds, const double* const* params, int32_t ninst, const int32_t* tile,
                            char* buf, int64_t buflen, int32_t flags);

/* First unknown of each of the first n instances of a batched handle. */
int mol_batch_offsets(const mol_handle* h, int64_t* offs, int32_t n);

/* Tile shape of the stencil kernels (nodes per work-group along x, y, z; 0,0,0 = default).
 * No counterpart in the reference; used by the autotuner.  At most 4096 nodes per tile. */
int mol_set_tile(mol_handle* h, int32_t tx, int32_t ty, int32_t tz);

/* Launch plan: info[0..2] tile shape, info[3] core tiles, info[4] generic tiles, info[5] WENO tiles,
 * info[6] 1 when the core tiles run the generated kernel, info[7] 1 when the IR fits the
 * interpreter's core kernel, info[8] tiles of the fused RK4 step kernel (0: none), info[9] RK4 steps
 * per fused launch (1, or 2 with an 8-reach halo; 0: none, or a resident 1D tile). */
int mol_get_plan(const mol_handle* h, int32_t* info, int32_t ninfo);

/* Bind auxiliary full-grid arrays (device pointers, naux of them, each prod(n) values). */
int mol_set_aux(mol_handle* h, void* const* aux);

typedef enum { MOL_RK4 = 0, MOL_SSPRK33 = 1 } mol_scheme;
typedef enum { MOL_FLAG_KAHAN = 1, MOL_FLAG_GRAPH = 2, MOL_FLAG_UNFUSED = 4 } mol_flags;

/* Replaces: solve(prob, RK4() | SSPRK33(); dt, adaptive = false) of OrdinaryDiffEq as used in
 * benchmark/weno/suite.jl:50-55 — the fixed-step stage loop, device-resident.  u is advanced in
 * place from t0 by nsteps steps of size dt.  Asynchronous: everything is enqueued on `stream` and
 * the call returns without waiting for the device (t0 travels as a kernel argument).
 * MOL_FLAG_KAHAN: compensated (Kahan) accumulation of the step update (FP32 state).
 * MOL_FLAG_GRAPH: chunks of 64, 16 and 4 steps are captured as hipGraphs on the first call for a
 * given (u, p, dt, scheme, flags) and kept on the handle; later calls with the same key replay
 * them (fewer than 4 leftover steps are launched directly).  The chunks are dropped by
 * mol_set_aux, mol_set_tile and mol_destroy, and all of them when more than eight keys are in
 * use (that drop waits for the device).  A 1D system held by one resident tile runs up to 1024
 * steps per launch and is launched directly.
 * Periodic 1D/2D systems of reach 1 run RK4 (without Kahan) as one fused launch per step (the
 * four stages on a tile with a 4-node halo, same arithmetic as the stage launches);
 * MOL_FLAG_UNFUSED runs the four stage launches instead. */
int mol_integrate(mol_handle* h, void* u, const void* p, double t0, double dt, int64_t nsteps,
                  int32_t scheme, int32_t flags, void* stream);

/* No counterpart in the reference: mol_integrate for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, arithmetic and results as mol_integrate; the stage vectors and the device clock
 * live in a library-owned device array of MOL_NOALLOC_SCRATCH_BYTES per device (reserved when
 * the library is loaded on the device, held until the process ends; its address is looked up by
 * mol_create on the handle's device).  It calls no hipMalloc/hipFree, creates no graph
 * (MOL_FLAG_GRAPH is MOL_ERR_INVALID) and records no memset.  Needs 256 bytes plus, per stage
 * vector rounded up to 256 bytes, one vector for fused RK4 steps, three for RK4 stage launches,
 * two for SSPRK33, one more with MOL_FLAG_KAHAN; larger integrations fail with MOL_ERR_SHAPE
 * before anything is enqueued.  Every call on one device shares that device's array: calls that
 * overlap on one device (any handle, any stream) race. */
#define MOL_NOALLOC_SCRATCH_BYTES (64u << 20)
int mol_integrate_noalloc(mol_handle* h, void* u, const void* p, double t0, double dt, int64_t nsteps,
                          int32_t scheme, int32_t flags, void* stream);

/* ---- adjoint ------------------------------------------------------------------------
 * Replaces: reverse-mode AD of the generated f (what SciMLSensitivity's adjoint methods
 * obtain by differentiating the RuntimeGeneratedFunction of src/MOL_discretization.jl:171-201
 * with Zygote/ReverseDiff on the CPU); the reference itself ships no hand-written adjoint.
 * Every operator slot kind and every program of the generated code is supported (the reverse
 * sweep is generated from the same IR); batched handles are not, and a grid-field gradient
 * through a boundary value that reads a grid field is MOL_ERR_UNSUPPORTED.
 *
 * mol_vjp: lam_dot = (df/du)^T lam, and g_aux[a*ngrid + node] += (df/daux_a)^T lam in FP64
 * (g_aux may be NULL).  u, lam, lam_dot: unknown vectors; g_aux: naux full-grid FP64 arrays.
 * Each tile writes its tile + halo adjoints to its own buffer and a gather kernel adds the
 * buffers of the tiles covering each node in a fixed order; only adjoints of ghost nodes reach
 * the unknowns with atomics (through the transposed closures). */
int mol_vjp(mol_handle* h, const void* u, const void* p, double t, const void* lam, void* lam_dot,
            double* g_aux, void* stream);

/* mol_forward_checkpointed: mol_integrate (RK4) that also keeps the state every ckpt_every
 * steps (and u0) in the handle, for a later mol_adjoint of the same run.  ckpt_every = 0 keeps
 * u and the three stage states of every step instead (4 * nsteps state vectors of device
 * memory; MOL_ERR_HIP if they do not fit): the adjoint then recomputes nothing. */
int mol_forward_checkpointed(mol_handle* h, void* u, const void* p, double t0, double dt,
                             int64_t nsteps, int32_t ckpt_every, void* stream);

/* mol_adjoint: discrete adjoint of the last checkpointed RK4 run.  On entry lam = dL/du(T);
 * on exit lam = dL/du(t0) and g_aux += dL/daux (FP64).  Each segment between checkpoints is
 * recomputed forward once (stage states kept on the device), then swept backwards with four
 * VJPs per step; after a ckpt_every = 0 run the sweep uses the stored stage states. */
int mol_adjoint(mol_handle* h, const void* p, void* lam, double* g_aux, void* stream);

/* FP64 error norms of a state against a reference, the counterpart of the final-time error checks
 * the reference's tests make on the CPU (isapprox against the analytic solution,
 * test/Diffusion/MOL_1D_Linear_Diffusion.jl:82,128): out (device,
 * 3 doubles) = max |u - ref|, ||u - ref||_2, ||ref||_2 over the unknowns, accumulated in FP64 with
 * 64-lane wavefront reductions in a fixed order.  u: device state of the handle's dtype; ref: device
 * FP64 array of the same length, or NULL (the norms of u).  Asynchronous on `stream`. */
int mol_error_norms(mol_handle* h, const void* u, const double* ref, double* out, void* stream);

/* Number of unknowns (sum over variables of the unknown box sizes). */
int64_t mol_num_unknowns(const mol_handle* h);

/* Per-handle error message of the last failing call (or the last mol_create failure when h is NULL). */
const char* mol_last_error(const mol_handle* h);

/* Replaces: nothing to free in the reference (GC); releases device memory of the handle. */
void mol_destroy(mol_handle* h);

/* Generated source of the core-tile kernel for a descriptor (the counterpart of the reference's
 * build_function codegen, src/MOL_discretization.jl:171-201).  Needs no GPU.  Writes at most
 * buflen bytes (NUL-terminated) into buf (may be NULL) and returns the source length, or a
 * negative mol_status (message in mol_last_error(NULL)); tile = tx,ty,tz or NULL for the default.
 * flags: MOL_CODEGEN_COMPILE also compiles it with hiprtc for gfx950 and fails when that does;
 * MOL_CODEGEN_VJP returns the reverse-mode (vector-Jacobian product) kernel instead,
 * MOL_CODEGEN_FUSED the fused RK4 step kernel (MOL_ERR_UNSUPPORTED when the system has none). */
enum { MOL_CODEGEN_COMPILE = 1, MOL_CODEGEN_VJP = 2, MOL_CODEGEN_FUSED = 4 };
int64_t mol_codegen(const mol_desc* d, const int32_t* tile, char* buf, int64_t buflen, int32_t flags);

/* Library / device information (for test harnesses). */
int mol_device_info(char* buf, int buflen);

#ifdef __cplusplus
}
#endif
#endif /* MOL_H */
