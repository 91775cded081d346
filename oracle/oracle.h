/*
 * CPU FP64 restatement of the quantics tensor-network hot path of
 * ITensorNumericalAnalysis.jl (v0.2.1).  TEST INFRASTRUCTURE ONLY: it is the
 * checker for the HIP engine and the CPU baseline timed beside it.  Nothing in
 * the product library links or calls it.
 *
 * Network convention (shared with the engine's C-ABI, see include/qtt.h):
 *   - a rooted tree of nv vertices, parent[0] = -1 and parent[v] < v;
 *   - vertex v carries one site index of dimension sdim[v] (1 = no site);
 *   - bond[v] is the dimension of edge (v, parent[v]); bond[0] must be 1;
 *   - core v is a dense row-major array [bond[v]][sdim[v]][bond[c1]]...[bond[cm]]
 *     with children c1 < ... < cm;
 *   - an operator (MPO) core is [Dp][s_in][s_out][Dc1]...: the application rule
 *     is phi[s_out] = sum_{s_in} O[s_in, s_out] psi[s_in]
 *     (src/digit_inds.jl:35-58, src/elementary_operators.jl:296).
 *
 * Each function is pinned by tests of its own: closed forms, brute force over
 * the whole grid, numpy's LAPACK and the reference's own test values.
 */
#ifndef QTT_ORACLE_H
#define QTT_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define OR_OK 0
#define OR_EINVAL 1
#define OR_ERANGE 2
#define OR_ENOMEM 3
#define OR_ENOCONV 5

/* Digit j (1 = most significant) of coordinate x with L binary digits,
 * following the greedy map of src/IndexMaps/abstractindexmap.jl:121-138
 * (exact for base 2: the bits of floor(x 2^L); x >= 1 saturates to all ones).
 * Returns -1 for x < 0 or NaN (the reference never terminates there). */
int or_digit(double x, int L, int j);

/* f(x) for npts points, xs row-major [npts][ncoord].  vcoord[v] is the
 * coordinate (0-based) whose digit vdigit[v] (1-based) sits on v, or -1.
 * Lc[c] is the digit count of coordinate c.  Invalid points give NaN and the
 * return value OR_ERANGE. */
int or_eval(int nv, const int* parent, const int* sdim, const long* bond,
            const double* const* cores, const int* vcoord, const int* vdigit,
            int ncoord, const int* Lc, const double* xs, long npts, double* out,
            int nthreads);

/* Contract every core with a uniform weight on its site (src/integration.jl:6-17):
 * weight = 1/base gives the left Riemann sum, weight = 1 the plain sum. */
int or_integrate(int nv, const int* parent, const int* sdim, const long* bond,
                 const double* const* cores, double weight, double* out);

/* <A|B> = sum over all site configurations of A * B (real networks). */
int or_inner(int nv, const int* parent, const int* sdim, const long* bondA,
             const double* const* coresA, const long* bondB,
             const double* const* coresB, double* out);

/* Exact operator application, out core bonds = bondW * bondA with merged
 * index (w, a) -> w * bondA + a.  out[v] must hold the full product size. */
int or_contract(int nv, const int* parent, const int* sdim, const long* bondW,
                const double* const* coresW, const long* bondA,
                const double* const* coresA, double* const* out);

/* Pullback of or_contract: abar[v] (A's core shapes) and wbar[v] (W's, nullable)
 * from gbar[v] (the shapes of or_contract's output). */
int or_contract_adjoint(int nv, const int* parent, const int* sdim, const long* bondW,
                        const double* const* coresW, const long* bondA,
                        const double* const* coresA, const double* const* gbar,
                        double* const* abar, double* const* wbar);

/* In-place truncation (post-order edges from vertex 0; for each edge child->parent
 * orthogonalise toward the child, SVD the edge, keep per the ITensors relative
 * cutoff rule on squared singular values and maxdim, push S V^T into the parent).
 * Bonds only shrink, so each core buffer keeps its original capacity.
 * trunc_err receives the sum over edges of the discarded relative weight. */
int or_truncate(int nv, const int* parent, const int* sdim, long* bond,
                double* const* cores, double cutoff, long maxdim,
                double* trunc_err);

/* Cotangent-weighted gradient: grads[v] = sum_m ybar[m] d f(x_m) / d core_v. */
int or_eval_grad(int nv, const int* parent, const int* sdim, const long* bond,
                 const double* const* cores, const int* vcoord, const int* vdigit,
                 int ncoord, const int* Lc, const double* xs, long npts,
                 const double* ybar, double* const* grads, int nthreads);

/* Dense factorizations used above, exposed for checks against LAPACK.
 * or_svd: M (r x c, row-major) = U (r x k) diag(S) Vt (k x c), k = min(r, c),
 * singular values descending. */
int or_svd(int r, int c, const double* M, double* U, double* S, double* Vt, double cut);
/* or_qr: M (r x c) = Q (r x k) R (k x c), k = min(r, c). */
int or_qr(int r, int c, const double* M, double* Q, double* R);

/* Truncation rule alone: p[0..n) squared singular values sorted descending.
 * Returns kept count; *err = discarded weight / total. */
long or_trunc_rule(const double* p, long n, double cutoff, long maxdim, double* err);

#ifdef __cplusplus
}
#endif
#endif
