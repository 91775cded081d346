/*
 * libqtt -- MI355X (gfx950) engine for the quantics tensor-network hot path of
 * ITensorNumericalAnalysis.jl (JoeyT1994, v0.2.1).  Plain C ABI: no C++ or
 * torch types cross it.  A Julia `ccall` shim and the Python ctypes harness
 * both bind exactly these symbols.
 *
 * Network convention (identical for every entry point):
 *   - a rooted tree of nv vertices renumbered so that parent[0] = -1 and
 *     parent[v] < v (a chain is parent[v] = v - 1);
 *   - vertex v carries at most one digit site of dimension sdim[v] (2, or 1
 *     for a vertex without digit); the digit is number vdigit[v] (1 = most
 *     significant) of coordinate vcoord[v] (0-based, -1 = none); coordinate c
 *     has ndigits[c] digits;
 *   - bond[v] is the dimension of edge (v, parent[v]); bond[0] = 1;
 *   - core v is a dense row-major float64 array
 *       [bond[v]][sdim[v]][bond[c1]]...[bond[cm]]   (children c1 < ... < cm);
 *   - an operator core is [bond[v]][s_in][s_out][bond[c1]]... and acts as
 *       phi[s_out] = sum_{s_in} O[s_in, s_out] psi[s_in].
 * Julia/ITensors are column-major: a Julia binding permutes each ITensor into
 * this order (the row-major dims (d0, ..., dn) are a Julia array's (dn, ..., d0)).
 *
 * Status codes: QTT_OK, QTT_EINVAL (shape/topology/argument), QTT_ERANGE
 * (a coordinate < 0 or NaN -- the reference never terminates there; x >= 1
 * saturates every digit to 1 exactly as the reference), QTT_ENOMEM,
 * QTT_EHIP (a HIP runtime error), QTT_ENOCONV (SVD did not converge).
 * The message of the last failure on a context is qtt_last_error().
 *
 * Threading: one context owns one HIP stream; calls on a context are
 * serialised by the caller.  Device pointers passed to the *_dev / async entry
 * points must be valid on the context's device; host buffers are only touched
 * during the call.
 */
#ifndef QTT_H
#define QTT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QTT_OK 0
#define QTT_EINVAL 1
#define QTT_ERANGE 2
#define QTT_ENOMEM 3
#define QTT_EHIP 4
#define QTT_ENOCONV 5

/* arithmetic of the bulk contractions: the exact contraction and the mode
 * products of the truncation sweep (its QR, SVD and cutoff are always FP64, and
 * so is the apply of a kept Q: a QR move of a matrix beyond LDS keeps its
 * reflectors and the move consuming that core applies them in FP64) */
#define QTT_F64 0       /* float64 everywhere */
#define QTT_F32_SPLIT 1 /* float32 hi/lo split operands, float-float accumulation */
#define QTT_F32 2       /* plain float32 operands and accumulation */

typedef struct qtt_ctx qtt_ctx;
typedef struct qtt_net qtt_net;
typedef struct qtt_mpo qtt_mpo;

typedef struct {
  int32_t nv;
  const int32_t* parent;
  const int32_t* sdim;
  const int32_t* vcoord;
  const int32_t* vdigit;
  int32_t ncoord;
  const int32_t* ndigits;
} qtt_topology;

/* ABI version (major * 100 + minor). */
int qtt_version(void);

/* Context on HIP device `device`. */
int qtt_ctx_create(int device, qtt_ctx** out);
void qtt_ctx_destroy(qtt_ctx* ctx);
const char* qtt_last_error(const qtt_ctx* ctx);
/* Run subsequent work on an existing HIP stream (hipStream_t, NULL = the
 * context's own stream).  The caller keeps the stream alive. */
int qtt_ctx_set_stream(qtt_ctx* ctx, void* hip_stream);
/* Wait for queued work; reports QTT_ERANGE if an asynchronous evaluate met an
 * invalid coordinate since the last sync. */
int qtt_sync(qtt_ctx* ctx);

/* ---- networks -----------------------------------------------------------
 * Replaces the ITensorNetworkFunction storage (src/itensornetworkfunction.jl:18-57):
 * cores are uploaded once and stay resident in HBM. */
int qtt_net_create(qtt_ctx* ctx, const qtt_topology* topo, const int64_t* bond,
                   const double* const* cores, qtt_net** out);
int qtt_net_bonds(qtt_ctx* ctx, const qtt_net* net, int64_t* bond_out);
int qtt_net_read(qtt_ctx* ctx, const qtt_net* net, double* const* cores_out);
int qtt_net_copy(qtt_ctx* ctx, const qtt_net* net, qtt_net** out);
void qtt_net_destroy(qtt_net* net);

/* Operator networks (the TTN returned by ttn(::OpSum), src/elementary_operators.jl:151). */
int qtt_mpo_create(qtt_ctx* ctx, const qtt_topology* topo, const int64_t* bond,
                   const double* const* cores, qtt_mpo** out);
void qtt_mpo_destroy(qtt_mpo* mpo);
/* bonds and cores of an operator network (as qtt_net_bonds / qtt_net_read;
 * cores in the operator layout [bond][s_in][s_out][children...]) */
int qtt_mpo_bonds(qtt_ctx* ctx, const qtt_mpo* mpo, int64_t* bond_out);
int qtt_mpo_read(qtt_ctx* ctx, const qtt_mpo* mpo, double* const* cores_out);
/* operator_proj(f) (src/elementary_operators.jl:246-258): the diagonal
 * operator |f><delta| of a network, O[s_in][s_out] = f[s] when s_in = s_out = s,
 * so operate(O, g) = f g.  Built on the device from the network's cores;
 * same layout and bonds.  The operator of map_to_zero_operator
 * (src/elementary_operators.jl:213-229) is this projection of
 * 1 - sum of point deltas.  Synchronous. */
int qtt_operator_proj(qtt_ctx* ctx, const qtt_net* f, qtt_mpo** out);

/* ---- network files (no reference counterpart: the reference has no on-disk
 * format; this is the interchange format of a binding and of test fixtures) --
 * Little-endian: "QTTN", uint32 version (1), kind (0 function, 1 operator),
 * nv, ncoord, flags (bit 0: names); int32 parent, sdim, vcoord, vdigit [nv],
 * ndigits [ncoord]; int64 bond [nv]; names (uint32 length + UTF-8 bytes per
 * vertex, when flagged); float64 cores in vertex order in the layouts above.
 * names (nullable) are stored as given and ignored on load.  Save writes a
 * temporary file and renames it; load validates the header and the size. */
int qtt_net_save(qtt_ctx* ctx, const qtt_net* net, const char* path, const char* const* names);
int qtt_net_load(qtt_ctx* ctx, const char* path, qtt_net** out);
int qtt_mpo_save(qtt_ctx* ctx, const qtt_mpo* mpo, const char* path, const char* const* names);
int qtt_mpo_load(qtt_ctx* ctx, const char* path, qtt_mpo** out);

/* ---- networks on graphs with loops ----------------------------------------
 * evaluate(f, xs; alg="bp") on a loopy layout (src/itensornetworkfunction.jl:96-106
 * -> ITensorNetworks.scalar(tn; alg="bp"); the grids of
 * test/test_realitensorfunction.jl:41).  The graph has nv vertices and ne
 * edges (eu[e], ev[e]) with bond ebond[e] (1..32); vertex v carries a binary
 * digit (sdim 2, vcoord, vdigit as in qtt_topology) or none (sdim 1).  Core v
 * is row-major [sdim[v]][legs...], its legs being its edges in increasing
 * edge id.  Per point, belief propagation on the projected network: sweeps
 * over the vertices in order, each vertex sending its messages over its legs
 * in edge order, every message normalised to unit 2-norm and used at once;
 * uniform start; stop when a sweep changes no message entry by more than tol,
 * or after maxiter sweeps.  The value is prod_v Z_v / prod_e Z_e (Z_v: vertex
 * v contracted with its incoming messages; Z_e: the two messages of edge e
 * contracted), exact on trees.  iters_out (nullable) receives the largest
 * sweep count over the points.  Host buffers; synchronous; invalid points
 * give QTT_ERANGE. */
typedef struct {
  int32_t nv, ne;
  const int32_t* eu;
  const int32_t* ev;
  const int64_t* ebond;
  const int32_t* sdim;
  const int32_t* vcoord;
  const int32_t* vdigit;
  int32_t ncoord;
  const int32_t* ndigits;
} qtt_graph;
int qtt_bp_eval_host(qtt_ctx* ctx, const qtt_graph* g, const double* const* cores, const double* xs,
                     int64_t npts, double* out, int maxiter, double tol, int* iters_out);

/* ---- evaluate ------------------------------------------------------------
 * Replaces evaluate(f, xs, dims; alg) (src/itensornetworkfunction.jl:96-106:
 * calculate_ind_values + project + scalar), batched over points.
 * xs: device pointer, row-major [npts][ncoord]; out: device pointer [npts].
 * Asynchronous on the context stream; invalid points yield NaN and a
 * QTT_ERANGE at the next qtt_sync. */
int qtt_eval(qtt_ctx* ctx, const qtt_net* net, const double* xs, int64_t npts, double* out,
             int precision);
/* Same with host buffers; synchronous, returns QTT_ERANGE directly. */
int qtt_eval_host(qtt_ctx* ctx, const qtt_net* net, const double* xs, int64_t npts,
                  double* out, int precision);

/* ---- integrate / inner ---------------------------------------------------
 * integrate(f; take_sum) (src/integration.jl:6-17): weight 0.5 per binary digit
 * gives the left Riemann sum, weight 1.0 the plain sum.  Synchronous. */
int qtt_integrate(qtt_ctx* ctx, const qtt_net* net, double weight, double* out);
/* <a|op|b> with op = NULL meaning identity; integrate(op, f)
 * (src/integration.jl:20-32) is qtt_integrate of qtt_apply(op, f) without
 * truncation, or qtt_inner with a uniform product state.  Synchronous. */
int qtt_inner(qtt_ctx* ctx, const qtt_net* a, const qtt_mpo* op, const qtt_net* b, double* out);

/* ---- operate / truncate --------------------------------------------------
 * operate(op, f; cutoff, maxdim) (src/elementary_operators.jl:294-299):
 * exact operator application followed by the truncation sweep below.
 * cutoff < 0 disables truncation (exact product, bonds multiply).
 * *out receives a new device-resident network; trunc_err (nullable) the summed
 * discarded relative weight. */
int qtt_apply(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, double cutoff,
              int64_t maxdim, int precision, qtt_net** out, double* trunc_err);
/* truncate(f; cutoff, maxdim) (src/itensornetworkfunction.jl:114-118 ->
 * ITensorNetworks.truncate): for every edge in post-order from vertex 0,
 * orthogonalise toward the child (Householder QR moves), SVD the edge
 * (one-sided Jacobi, FP64) and keep singular values by the ITensors rule on
 * squared values (relative cutoff, then maxdim).  An edge that can drop
 * nothing (cutoff <= 0, maxdim >= its rank, R of its QR nonsingular) keeps
 * its QR factors instead: the same function, bonds and error, the bond's
 * basis differing by an orthogonal change.  In place. */
int qtt_truncate(qtt_ctx* ctx, qtt_net* net, double cutoff, int64_t maxdim, double* trunc_err);

/* ---- adjoint (no reference counterpart: the rrule of evaluate) ------------
 * grad = sum_m ybar[m] d f(x_m) / d cores, written as the concatenation of
 * all cores in vertex order (same sizes as the network's cores).
 * xs, ybar, grad: device pointers.  Asynchronous.  Any tree in QTT_F64;
 * chains also in QTT_F32 (float32 environments and per-tile sums, float64
 * accumulation across tiles; about 1e-6 relative).  QTT_F32_SPLIT: EINVAL. */
int qtt_eval_grad(qtt_ctx* ctx, const qtt_net* net, const double* xs, const double* ybar,
                  int64_t npts, double* grad, int precision);
/* Host-buffer variant, synchronous; grads[v] receives core v's gradient. */
int qtt_eval_grad_host(qtt_ctx* ctx, const qtt_net* net, const double* xs, const double* ybar,
                       int64_t npts, double* const* grads, int precision);

/* Pullback of the exact operator application (no reference counterpart: the
 * rrule of operate(op, f) without truncation, src/elementary_operators.jl:294-299).
 * gbar: the cotangent of the exact product's cores (the network that
 * qtt_apply(op, net, cutoff < 0) returns), concatenated in vertex order;
 * fbar receives the cotangent of net's cores, wbar (nullable) that of op's
 * cores, both concatenated in vertex order.  Device pointers; returns once
 * the work is done.  The truncation that qtt_apply applies after the product
 * has no pullback here. */
int qtt_apply_grad(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, const double* gbar, double* fbar,
                   double* wbar);

/* ---- function algebra ------------------------------------------------------
 * Hadamard product (g f)(x) = g(x) f(x), multiply(g, f)
 * (src/elementary_operators.jl:260-282).  The product lives on `topo`, the
 * union of the two layouts: g_of[o] / f_of[o] give the vertex of g / f at
 * output vertex o (-1: that function has no vertex there and contributes 1).
 * Every vertex and edge of both operands must appear in topo, digits at the
 * same vertices.  Bond of an output edge: g's bond times f's (1 where an
 * operand lacks the edge); output index i = i_g * bond_f + i_f.  Synchronous;
 * *out receives a new network. */
int qtt_multiply(qtt_ctx* ctx, const qtt_net* g, const qtt_net* f, const qtt_topology* topo, const int32_t* g_of,
                 const int32_t* f_of, qtt_net** out);
/* (f + g)(x), the direct sum of two networks on the same layout
 * (ITensorNetworks.add as overloaded in src/fixes.jl:40-95): bonds add, f's
 * block first.  Synchronous. */
int qtt_add(qtt_ctx* ctx, const qtt_net* f, const qtt_net* g, qtt_net** out);
/* partial_integrate(f, dims; merge_vertices, take_sum) (src/integration.jl:35-54):
 * sums the digits of every coordinate c with integrate[c] != 0 (each digit
 * weighted by `weight`: 0.5 for the left Riemann sum, 1.0 for the plain sum)
 * and contracts the vertices left without a digit into the remaining ones
 * (merge_internal_tensors, src/utils.jl:99-113).  topo is the output layout,
 * out_vertex[o] the network vertex that output vertex o keeps; its edges must
 * be network edges (so the kept vertices are connected) and its digits those
 * that remain.  An output layout with every vertex (integrated ones without
 * digit) is merge_vertices = false.  Synchronous. */
int qtt_partial_integrate(qtt_ctx* ctx, const qtt_net* net, const int32_t* integrate, double weight,
                          const qtt_topology* topo, const int32_t* out_vertex, qtt_net** out);

/* number of float64 elements of all cores of a network / an operator */
int64_t qtt_net_size(const qtt_net* net);
int64_t qtt_mpo_size(const qtt_mpo* op);

/* ---- batches (no reference counterpart: the reference applies operators to
 * one function at a time, src/elementary_operators.jl:305-319) -------------
 * n independent operate(op[i], nets[i]; cutoff, maxdim): the exact
 * contractions, then ONE truncation launch with one workgroup per network.
 * outs[i] receives new networks; errs (nullable) the truncation errors. */
int qtt_apply_batch(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets,
                    double cutoff, int64_t maxdim, int precision, qtt_net** outs, double* errs);
/* n integrals in one launch (one workgroup per network); out[i] on the host. */
int qtt_integrate_batch(qtt_ctx* ctx, int n, const qtt_net* const* nets, double weight, double* out);
/* evaluate(nets[i], xs[i]) for a batch of functions, each at its own points:
 * xs[i] device [npts[i]][ncoord_i] float64, out[i] device [npts[i]].  Float64
 * chains with bonds <= 32 run in one launch per bond class; other networks are
 * evaluated one launch each.  Returns once the launches are queued (their
 * argument tables copied); the values are ready after qtt_sync. */
int qtt_eval_batch(qtt_ctx* ctx, int n, const qtt_net* const* nets, const double* const* xs,
                   const int64_t* npts, double* const* out, int precision);
/* One step of a batch of kernel-operator solves, the loop body of
 * examples/fredholm_solver.jl:52-63 for n functions at once:
 *   outs[i] = truncate(contract(ops[i], nets[i]); cutoff, maxdim)   (operate)
 *   integrals[i] = integrate(outs[i]), weight per digit as qtt_integrate
 *   vals[i][p] = outs[i](xs[i][p])                                    (evaluate)
 * Every stage is one batched launch; the stages follow each other on the
 * stream with no host synchronisation (bond-dependent evaluate layouts are
 * computed by a kernel) and are captured into one HIP graph unless the
 * context's stream cannot be captured, a single network needs helper
 * workgroups for its truncation (a cooperative launch), or QTT_BATCH_GRAPH=0.
 * Bonds, truncation errors and integrals are copied back once, after the
 * graph.  cutoff < 0: no truncation (errs 0).  integrals (host [n]) and the
 * evaluate arguments (xs, npts, vals: device points and values as in
 * qtt_eval_batch) are nullable.  Evaluation is float64; chains whose final
 * bonds exceed 32, trees, and more than 4 coordinates are evaluated after the
 * graph by the single-network path.  precision: as qtt_apply.  Synchronous:
 * returns when everything is done; invalid points report QTT_ERANGE (the
 * outputs are then destroyed). */
int qtt_batch_run(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets, double cutoff,
                  int64_t maxdim, int precision, double weight, const double* const* xs, const int64_t* npts,
                  double* const* vals, qtt_net** outs, double* errs, double* integrals);

/* ---- evaluation without allocation (no reference counterpart) -------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same result
 * as qtt_eval with QTT_F64 on a chain network, but the packed copy of the
 * cores, its offsets and its site descriptors are staged in a fixed device
 * array of the library (one per GPU, held until the process ends) instead of
 * a per-network allocation: the call allocates nothing and copies nothing
 * from the host.  Limits: at most QTT_NOALLOC_MAX_SITES sites and
 * QTT_NOALLOC_PACK_DOUBLES doubles of packed cores (each site takes
 * 2 * round4(left bond) * round4(right bond)); bonds as the LDS chain kernel
 * allows (<= 160).  Trees, larger networks and other precisions return
 * QTT_EINVAL.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams or contexts) race and must be ordered by
 * the caller.  Asynchronous, as qtt_eval. */
#define QTT_NOALLOC_MAX_SITES 1024
#define QTT_NOALLOC_PACK_DOUBLES (1 << 20)
int qtt_eval_noalloc(qtt_ctx* ctx, const qtt_net* net, const double* xs, int64_t npts, double* out);

/* ---- autotuning (no reference counterpart) --------------------------------
 * Times every launch shape of the float64 register chain evaluation on these
 * device points (xs [npts][ncoord]; out [npts] is overwritten by the runs):
 * digit sites per environment table (0 = no tables, else 8, 10, ... up to 16
 * and half of the chain's digit sites), points per lane 1 or 2, workgroups of
 * 64, 128 or 256 threads.  The fastest is kept in the context for networks of
 * the same bond class and point counts of the same octave; later qtt_eval
 * calls use it.  cfg[3] receives {points per lane, workgroup size, table
 * digit sites} ({0, 0, 0} if the network does not take that kernel: not a
 * chain, bonds over 32), ms the time of one call once the network's tables
 * exist.  Synchronous.
 * Environment tables (the prefix and suffix products of a chain for every
 * value of their digits) are built by the first qtt_eval of a network with a
 * given table size and kept with the network, like its packed copy, until its
 * cores change (a truncation) or it is destroyed: later calls only run the
 * chain kernel. */
int qtt_tune_eval(qtt_ctx* ctx, qtt_net* net, const double* xs, int64_t npts, double* out, int* cfg,
                  double* ms);

#ifdef __cplusplus
}
#endif
#endif
