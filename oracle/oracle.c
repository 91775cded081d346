/*
 * CPU FP64 restatement of the quantics tensor-network hot path of
 * ITensorNumericalAnalysis.jl v0.2.1 -- TEST INFRASTRUCTURE ONLY (the checker
 * for the HIP engine, and the CPU baseline it is timed against).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Functions and the reference code they restate:
 *   or_digit      src/IndexMaps/abstractindexmap.jl:121-138, realindexmap.jl:13-15,67-76
 *   or_eval       src/itensornetworkfunction.jl:84-106 (project + scalar)
 *   or_integrate  src/integration.jl:6-17
 *   or_inner      src/integration.jl:20-32 (inner of two networks)
 *   or_contract   src/elementary_operators.jl:294-299 (exact operator application)
 *   or_truncate   src/itensornetworkfunction.jl:114-118 -> ITensorNetworks.truncate
 *                 (third-party, ITensorNetworks 0.13, restated: post-order edges,
 *                 orthogonalise toward the edge source, SVD truncation with the
 *                 ITensors/NDTensors 0.9 relative cutoff rule on squared singular
 *                 values)
 *   or_eval_grad  no reference counterpart (adjoint of or_eval)
 *   or_contract_adjoint  no reference counterpart (adjoint of or_contract)
 * The parity status of each is recorded with its tests.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* tree helpers                                                               */
/* ------------------------------------------------------------------------- */
typedef struct {
  int nv;
  const int* parent;
  int* nch;    /* number of children */
  int* chs;    /* children, flattened */
  int* choff;  /* offset of v's children in chs */
} tree_t;

static void tree_free(tree_t* t) {
  free(t->nch);
  free(t->chs);
  free(t->choff);
}

static int tree_build(tree_t* t, int nv, const int* parent) {
  t->nv = nv;
  t->parent = parent;
  t->nch = (int*)calloc((size_t)nv + 1, sizeof(int));
  t->chs = (int*)calloc((size_t)nv + 1, sizeof(int));
  t->choff = (int*)calloc((size_t)nv + 1, sizeof(int));
  if (!t->nch || !t->chs || !t->choff) {
    tree_free(t);
    return OR_ENOMEM;
  }
  if (nv < 1 || parent[0] != -1) {
    tree_free(t);
    return OR_EINVAL;
  }
  for (int v = 1; v < nv; v++) {
    if (parent[v] < 0 || parent[v] >= v) {
      tree_free(t);
      return OR_EINVAL;
    }
    t->nch[parent[v]]++;
  }
  int off = 0;
  for (int v = 0; v < nv; v++) {
    t->choff[v] = off;
    off += t->nch[v];
  }
  int* fill = (int*)calloc((size_t)nv, sizeof(int));
  if (!fill) {
    tree_free(t);
    return OR_ENOMEM;
  }
  for (int v = 1; v < nv; v++) {
    int p = parent[v];
    t->chs[t->choff[p] + fill[p]++] = v;
  }
  free(fill);
  return OR_OK;
}

static inline const int* tree_ch(const tree_t* t, int v) { return t->chs + t->choff[v]; }

/* dims of core v: [bond[v], sdim[v], bond[c1], ...]; returns ndim */
static int core_dims(const tree_t* t, const int* sdim, const long* bond, int v, long* dims) {
  int m = t->nch[v];
  dims[0] = bond[v];
  dims[1] = sdim[v];
  const int* ch = tree_ch(t, v);
  for (int k = 0; k < m; k++) dims[2 + k] = bond[ch[k]];
  return 2 + m;
}

static long prod(const long* d, int n) {
  long p = 1;
  for (int i = 0; i < n; i++) p *= d[i];
  return p;
}

/* ------------------------------------------------------------------------- */
/* digit map                                                                  */
/* ------------------------------------------------------------------------- */
int or_digit(double x, int L, int j) {
  if (!(x >= 0.0)) return -1; /* negative or NaN */
  if (L < 1 || L > 62 || j < 1 || j > L) return -1;
  uint64_t k;
  if (x >= 1.0)
    k = (((uint64_t)1) << L) - 1; /* greedy loop saturates at base-1 */
  else
    k = (uint64_t)ldexp(x, L); /* x * 2^L is exact; truncation = floor */
  return (int)((k >> (L - j)) & 1u);
}

/* ------------------------------------------------------------------------- */
/* contraction of a core slice with child vectors                              */
/* ------------------------------------------------------------------------- */
static double contract_children(const double* base, int m, const int* ch, const long* bond,
                                const double* vec, const long* voff) {
  if (m == 0) return base[0];
  long stride = 1;
  for (int k = 1; k < m; k++) stride *= bond[ch[k]];
  const double* u = vec + voff[ch[0]];
  double acc = 0.0;
  for (long i = 0; i < bond[ch[0]]; i++) {
    if (u[i] != 0.0) acc += u[i] * contract_children(base + i * stride, m - 1, ch + 1, bond, vec, voff);
  }
  return acc;
}

static int check_bonds(int nv, const long* bond) {
  if (bond[0] != 1) return OR_EINVAL;
  for (int v = 0; v < nv; v++)
    if (bond[v] < 1) return OR_EINVAL;
  return OR_OK;
}

int or_eval(int nv, const int* parent, const int* sdim, const long* bond, const double* const* cores,
            const int* vcoord, const int* vdigit, int ncoord, const int* Lc, const double* xs,
            long npts, double* out, int nthreads) {
  tree_t t;
  int rc = tree_build(&t, nv, parent);
  if (rc) return rc;
  if (check_bonds(nv, bond)) {
    tree_free(&t);
    return OR_EINVAL;
  }
  for (int v = 0; v < nv; v++) {
    if (vcoord[v] >= ncoord || (vcoord[v] >= 0 && (vdigit[v] < 1 || vdigit[v] > Lc[vcoord[v]]))) {
      tree_free(&t);
      return OR_EINVAL;
    }
    if (vcoord[v] >= 0 && sdim[v] != 2) {
      tree_free(&t);
      return OR_EINVAL;
    }
  }
  long* voff = (long*)malloc(sizeof(long) * (size_t)(nv + 1));
  long tot = 0;
  for (int v = 0; v < nv; v++) {
    voff[v] = tot;
    tot += bond[v];
  }
  int bad = 0;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : bad)
  {
    double* vec = (double*)malloc(sizeof(double) * (size_t)tot);
    int* dig = (int*)malloc(sizeof(int) * (size_t)nv);
#pragma omp for schedule(static)
    for (long i = 0; i < npts; i++) {
      const double* x = xs + i * ncoord;
      int ok = 1;
      for (int v = 0; v < nv; v++) {
        if (vcoord[v] < 0) {
          dig[v] = 0;
          continue;
        }
        int d = or_digit(x[vcoord[v]], Lc[vcoord[v]], vdigit[v]);
        if (d < 0) {
          ok = 0;
          break;
        }
        dig[v] = d;
      }
      if (!ok) {
        out[i] = NAN;
        bad = 1;
        continue;
      }
      for (int v = nv - 1; v >= 0; v--) {
        int m = t.nch[v];
        const int* ch = tree_ch(&t, v);
        long pc = 1;
        for (int k = 0; k < m; k++) pc *= bond[ch[k]];
        for (long p = 0; p < bond[v]; p++) {
          const double* base = cores[v] + (p * sdim[v] + dig[v]) * pc;
          vec[voff[v] + p] = contract_children(base, m, ch, bond, vec, voff);
        }
      }
      out[i] = vec[voff[0]];
    }
    free(vec);
    free(dig);
  }
  free(voff);
  tree_free(&t);
  return bad ? OR_ERANGE : OR_OK;
}

int or_integrate(int nv, const int* parent, const int* sdim, const long* bond,
                 const double* const* cores, double weight, double* out) {
  tree_t t;
  int rc = tree_build(&t, nv, parent);
  if (rc) return rc;
  if (check_bonds(nv, bond)) {
    tree_free(&t);
    return OR_EINVAL;
  }
  long* voff = (long*)malloc(sizeof(long) * (size_t)(nv + 1));
  long tot = 0, maxpc = 1;
  for (int v = 0; v < nv; v++) {
    voff[v] = tot;
    tot += bond[v];
  }
  for (int v = 0; v < nv; v++) {
    long pc = 1;
    for (int k = 0; k < t.nch[v]; k++) pc *= bond[tree_ch(&t, v)[k]];
    if (pc > maxpc) maxpc = pc;
  }
  double* vec = (double*)malloc(sizeof(double) * (size_t)tot);
  double* fold = (double*)malloc(sizeof(double) * (size_t)maxpc);
  for (int v = nv - 1; v >= 0; v--) {
    int m = t.nch[v];
    const int* ch = tree_ch(&t, v);
    long pc = 1;
    for (int k = 0; k < m; k++) pc *= bond[ch[k]];
    /* the weight is per site index: a vertex without a digit has none */
    const double wv = sdim[v] > 1 ? weight : 1.0;
    for (long p = 0; p < bond[v]; p++) {
      for (long j = 0; j < pc; j++) {
        double s = 0.0;
        for (int si = 0; si < sdim[v]; si++) s += cores[v][(p * sdim[v] + si) * pc + j];
        fold[j] = wv * s;
      }
      vec[voff[v] + p] = contract_children(fold, m, ch, bond, vec, voff);
    }
  }
  *out = vec[voff[0]];
  free(vec);
  free(fold);
  free(voff);
  tree_free(&t);
  return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* generic dense tensor helpers                                               */
/* ------------------------------------------------------------------------- */
/* Y = X with axis `ax` multiplied by M (nnew x nold): Y[a][j][b] = sum_i M[j][i] X[a][i][b] */
static void mode_product(const double* X, const long* dims, int nd, int ax, const double* M, long nnew,
                         double* Y) {
  long A = prod(dims, ax), nold = dims[ax], B = prod(dims + ax + 1, nd - ax - 1);
  for (long a = 0; a < A; a++)
    for (long j = 0; j < nnew; j++) {
      double* y = Y + (a * nnew + j) * B;
      for (long b = 0; b < B; b++) y[b] = 0.0;
      for (long i = 0; i < nold; i++) {
        double mji = M[j * nold + i];
        if (mji == 0.0) continue;
        const double* x = X + (a * nold + i) * B;
        for (long b = 0; b < B; b++) y[b] += mji * x[b];
      }
    }
}

/* move axis ax to the last position: Y[a][b][i] = X[a][i][b] */
static void axis_to_last(const double* X, const long* dims, int nd, int ax, double* Y) {
  long A = prod(dims, ax), n = dims[ax], B = prod(dims + ax + 1, nd - ax - 1);
  for (long a = 0; a < A; a++)
    for (long i = 0; i < n; i++)
      for (long b = 0; b < B; b++) Y[(a * B + b) * n + i] = X[(a * n + i) * B + b];
}

/* inverse of axis_to_last with the moved axis now of size n */
static void last_to_axis(const double* Y, const long* dims, int nd, int ax, long n, double* X) {
  long A = prod(dims, ax), B = prod(dims + ax + 1, nd - ax - 1);
  for (long a = 0; a < A; a++)
    for (long i = 0; i < n; i++)
      for (long b = 0; b < B; b++) X[(a * n + i) * B + b] = Y[(a * B + b) * n + i];
}

int or_inner(int nv, const int* parent, const int* sdim, const long* bondA, const double* const* coresA,
             const long* bondB, const double* const* coresB, double* out) {
  tree_t t;
  int rc = tree_build(&t, nv, parent);
  if (rc) return rc;
  if (check_bonds(nv, bondA) || check_bonds(nv, bondB)) {
    tree_free(&t);
    return OR_EINVAL;
  }
  double** E = (double**)calloc((size_t)nv, sizeof(double*));
  long dims[64], dimsA[64];
  for (int v = nv - 1; v >= 0; v--) {
    int nd = core_dims(&t, sdim, bondB, v, dims);
    core_dims(&t, sdim, bondA, v, dimsA);
    long sz = prod(dims, nd);
    double* Y = (double*)malloc(sizeof(double) * (size_t)sz);
    memcpy(Y, coresB[v], sizeof(double) * (size_t)sz);
    const int* ch = tree_ch(&t, v);
    for (int k = 0; k < t.nch[v]; k++) {
      int c = ch[k];
      long nsz = sz / dims[2 + k] * bondA[c];
      double* Z = (double*)malloc(sizeof(double) * (size_t)nsz);
      mode_product(Y, dims, nd, 2 + k, E[c], bondA[c], Z);
      free(Y);
      Y = Z;
      sz = nsz;
      dims[2 + k] = bondA[c];
    }
    /* E_v[pa][pb] = sum_rest A[pa][rest] Y[pb][rest] */
    long rest = sz / bondB[v];
    double* Ev = (double*)calloc((size_t)(bondA[v] * bondB[v]), sizeof(double));
    for (long pa = 0; pa < bondA[v]; pa++)
      for (long pb = 0; pb < bondB[v]; pb++) {
        double s = 0.0;
        const double* a = coresA[v] + pa * rest;
        const double* y = Y + pb * rest;
        for (long r = 0; r < rest; r++) s += a[r] * y[r];
        Ev[pa * bondB[v] + pb] = s;
      }
    free(Y);
    E[v] = Ev;
    for (int k = 0; k < t.nch[v]; k++) {
      free(E[ch[k]]);
      E[ch[k]] = NULL;
    }
  }
  *out = E[0][0];
  free(E[0]);
  free(E);
  tree_free(&t);
  return OR_OK;
}

int or_contract(int nv, const int* parent, const int* sdim, const long* bondW, const double* const* coresW,
                const long* bondA, const double* const* coresA, double* const* out) {
  tree_t t;
  int rc = tree_build(&t, nv, parent);
  if (rc) return rc;
  if (check_bonds(nv, bondW) || check_bonds(nv, bondA)) {
    tree_free(&t);
    return OR_EINVAL;
  }
  for (int v = 0; v < nv; v++) {
    int m = t.nch[v];
    const int* ch = tree_ch(&t, v);
    int d = sdim[v];
    long pcw = 1, pca = 1;
    for (int k = 0; k < m; k++) {
      pcw *= bondW[ch[k]];
      pca *= bondA[ch[k]];
    }
    long pco = pcw * pca;
    long Dp = bondW[v], Ap = bondA[v];
    /* precompute merged child offset for every (jw, ja) pair via radix decomposition */
    long* offw = (long*)malloc(sizeof(long) * (size_t)pcw);
    long* offa = (long*)malloc(sizeof(long) * (size_t)pca);
    /* output child flat index = sum_k (cw_k * bondA[c_k] + ca_k) * S_k with
       S_k = prod_{l>k} bondW[c_l] * bondA[c_l] */
    for (long jw = 0; jw < pcw; jw++) {
      long r = jw, off = 0, S = 1;
      for (int k = m - 1; k >= 0; k--) {
        long bw = bondW[ch[k]], ba = bondA[ch[k]];
        long cw = r % bw;
        r /= bw;
        off += cw * ba * S;
        S *= bw * ba;
      }
      offw[jw] = off;
    }
    for (long ja = 0; ja < pca; ja++) {
      long r = ja, off = 0, S = 1;
      for (int k = m - 1; k >= 0; k--) {
        long bw = bondW[ch[k]], ba = bondA[ch[k]];
        long ca = r % ba;
        r /= ba;
        off += ca * S;
        S *= bw * ba;
      }
      offa[ja] = off;
    }
    double* C = out[v];
    memset(C, 0, sizeof(double) * (size_t)(Dp * Ap * d * pco));
    for (long pw = 0; pw < Dp; pw++)
      for (long pa = 0; pa < Ap; pa++)
        for (int so = 0; so < d; so++) {
          double* c = C + ((pw * Ap + pa) * d + so) * pco;
          for (int si = 0; si < d; si++) {
            const double* w = coresW[v] + ((pw * d + si) * d + so) * pcw;
            const double* a = coresA[v] + (pa * d + si) * pca;
            for (long jw = 0; jw < pcw; jw++) {
              double wv = w[jw];
              if (wv == 0.0) continue;
              for (long ja = 0; ja < pca; ja++) c[offw[jw] + offa[ja]] += wv * a[ja];
            }
          }
        }
    free(offw);
    free(offa);
  }
  tree_free(&t);
  return OR_OK;
}

/* Pullback of or_contract (no reference counterpart: the rrule of operate
 * without truncation, src/elementary_operators.jl:294-299).  gbar holds the
 * cotangent of the cores of C = W A in or_contract's layout; then
 *   abar_v[pa][si][ca..]     = sum_{pw, so, cw..} W_v[pw][si][so][cw..] gbar_v[(pw, pa)][so][(cw, ca)..]
 *   wbar_v[pw][si][so][cw..] = sum_{pa, ca..}     A_v[pa][si][ca..]     gbar_v[(pw, pa)][so][(cw, ca)..]
 * abar and wbar (nullable) are overwritten. */
int or_contract_adjoint(int nv, const int* parent, const int* sdim, const long* bondW,
                        const double* const* coresW, const long* bondA, const double* const* coresA,
                        const double* const* gbar, double* const* abar, double* const* wbar) {
  tree_t t;
  int rc = tree_build(&t, nv, parent);
  if (rc) return rc;
  if (check_bonds(nv, bondW) || check_bonds(nv, bondA)) {
    tree_free(&t);
    return OR_EINVAL;
  }
  for (int v = 0; v < nv; v++) {
    int m = t.nch[v];
    const int* ch = tree_ch(&t, v);
    int d = sdim[v];
    long pcw = 1, pca = 1;
    for (int k = 0; k < m; k++) {
      pcw *= bondW[ch[k]];
      pca *= bondA[ch[k]];
    }
    long pco = pcw * pca;
    long Dp = bondW[v], Ap = bondA[v];
    long* offw = (long*)malloc(sizeof(long) * (size_t)pcw);
    long* offa = (long*)malloc(sizeof(long) * (size_t)pca);
    for (long jw = 0; jw < pcw; jw++) {
      long r = jw, off = 0, S = 1;
      for (int k = m - 1; k >= 0; k--) {
        long bw = bondW[ch[k]], ba = bondA[ch[k]];
        off += (r % bw) * ba * S;
        r /= bw;
        S *= bw * ba;
      }
      offw[jw] = off;
    }
    for (long ja = 0; ja < pca; ja++) {
      long r = ja, off = 0, S = 1;
      for (int k = m - 1; k >= 0; k--) {
        long bw = bondW[ch[k]], ba = bondA[ch[k]];
        off += (r % ba) * S;
        r /= ba;
        S *= bw * ba;
      }
      offa[ja] = off;
    }
    memset(abar[v], 0, sizeof(double) * (size_t)(Ap * d * pca));
    if (wbar) memset(wbar[v], 0, sizeof(double) * (size_t)(Dp * d * d * pcw));
    for (long pw = 0; pw < Dp; pw++)
      for (long pa = 0; pa < Ap; pa++)
        for (int so = 0; so < d; so++) {
          const double* g = gbar[v] + ((pw * Ap + pa) * d + so) * pco;
          for (int si = 0; si < d; si++) {
            const double* w = coresW[v] + ((pw * d + si) * d + so) * pcw;
            const double* a = coresA[v] + (pa * d + si) * pca;
            double* ab = abar[v] + (pa * d + si) * pca;
            double* wb = wbar ? wbar[v] + ((pw * d + si) * d + so) * pcw : NULL;
            for (long jw = 0; jw < pcw; jw++)
              for (long ja = 0; ja < pca; ja++) {
                double gv = g[offw[jw] + offa[ja]];
                ab[ja] += w[jw] * gv;
                if (wb) wb[jw] += a[ja] * gv;
              }
          }
        }
    free(offw);
    free(offa);
  }
  tree_free(&t);
  return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* dense factorizations                                                       */
/* ------------------------------------------------------------------------- */
/* Householder QR of M (r x c): Q (r x k), R (k x c), k = min(r, c). */
int or_qr(int r, int c, const double* M, double* Q, double* R) {
  int k = r < c ? r : c;
  double* A = (double*)malloc(sizeof(double) * (size_t)r * c);
  double* V = (double*)calloc((size_t)r * k, sizeof(double)); /* reflectors, column j in V[:, j] */
  double* beta = (double*)calloc((size_t)k, sizeof(double));
  memcpy(A, M, sizeof(double) * (size_t)r * c);
  for (int j = 0; j < k; j++) {
    double nrm = 0.0;
    for (int i = j; i < r; i++) nrm += A[i * c + j] * A[i * c + j];
    nrm = sqrt(nrm);
    if (nrm == 0.0) {
      beta[j] = 0.0;
      continue;
    }
    double x0 = A[j * c + j];
    double alpha = x0 >= 0 ? -nrm : nrm;
    /* v = x - alpha e1, H = I - 2 v v^T / (v^T v) */
    double vnorm2 = 0.0;
    for (int i = j; i < r; i++) {
      double vi = A[i * c + j] - (i == j ? alpha : 0.0);
      V[i * k + j] = vi;
      vnorm2 += vi * vi;
    }
    beta[j] = vnorm2 > 0 ? 2.0 / vnorm2 : 0.0;
    for (int col = j; col < c; col++) {
      double s = 0.0;
      for (int i = j; i < r; i++) s += V[i * k + j] * A[i * c + col];
      s *= beta[j];
      for (int i = j; i < r; i++) A[i * c + col] -= s * V[i * k + j];
    }
  }
  for (int i = 0; i < k; i++)
    for (int col = 0; col < c; col++) R[i * c + col] = col >= i ? A[i * c + col] : 0.0;
  /* Q = H_0 ... H_{k-1} [I_k; 0] */
  for (int i = 0; i < r; i++)
    for (int j = 0; j < k; j++) Q[i * k + j] = (i == j) ? 1.0 : 0.0;
  for (int j = k - 1; j >= 0; j--) {
    if (beta[j] == 0.0) continue;
    for (int col = 0; col < k; col++) {
      double s = 0.0;
      for (int i = j; i < r; i++) s += V[i * k + j] * Q[i * k + col];
      s *= beta[j];
      for (int i = j; i < r; i++) Q[i * k + col] -= s * V[i * k + j];
    }
  }
  free(A);
  free(V);
  free(beta);
  return OR_OK;
}

/* One-sided Jacobi on the rows of X (n x c, n <= c) with the rotations
 * accumulated in J (n x n, starts as identity): J X_in = X_out, rows of X_out
 * mutually orthogonal.  Cyclic row ordering.
 * cut in (0, 1): the relative cutoff of the truncation that follows.  Two rows
 * whose squared norms are both at most cut ||X||_F^2 / n (the noise level) are
 * not rotated against each other.  At convergence every row above the level is
 * orthogonal to all rows, so it is an exact singular triplet; the rows below
 * it (the noise block) weigh W <= cut ||X||_F^2 together, but are not yet
 * mutually orthogonal.  svd_factor then checks whether the noise block can
 * change what the truncation keeps (see noise_block_exact) and, if it can,
 * calls this again with cut = 0 and init_j = 0 to finish the SVD.  Any other
 * cut rotates every pair.  init_j = 0 continues from the rotations in J. */
static int jacobi_rows(int n, int c, double* X, double* J, double cut, int init_j) {
  if (init_j)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) J[i * n + j] = (i == j) ? 1.0 : 0.0;
  if (n < 2) return OR_OK;
  /* a computed dot product of length c carries rounding error up to ~c eps,
     so the stopping test must not be tighter than that */
  const double tol = DBL_EPSILON * (double)(c > n ? c : n);
  /* rows below eps * ||X||_F are numerically zero: rotating them against each
     other only stirs rounding noise (and can reach subnormal range) */
  double fro = 0.0;
  for (long e = 0; e < (long)n * c; e++) fro += X[e] * X[e];
  const double small = fro * DBL_EPSILON * DBL_EPSILON;
  const double noise = (cut > 0.0 && cut < 1.0) ? cut * fro / (double)n : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    int rot = 0;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double a = 0.0, b = 0.0, g = 0.0;
        const double* xp = X + (long)p * c;
        const double* xq = X + (long)q * c;
        for (int k = 0; k < c; k++) {
          a += xp[k] * xp[k];
          b += xq[k] * xq[k];
          g += xp[k] * xq[k];
        }
        if (g == 0.0 || a <= small || b <= small || fabs(g) <= tol * sqrt(a) * sqrt(b)) continue;
        if (a <= noise && b <= noise) continue;
        double zeta = (b - a) / (2.0 * g);
        double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
        double* yp = X + (long)p * c;
        double* yq = X + (long)q * c;
        for (int k = 0; k < c; k++) {
          double u = yp[k], w = yq[k];
          yp[k] = cs * u - sn * w;
          yq[k] = sn * u + cs * w;
        }
        double* jp = J + (long)p * n;
        double* jq = J + (long)q * n;
        for (int k = 0; k < n; k++) {
          double u = jp[k], w = jq[k];
          jp[k] = cs * u - sn * w;
          jq[k] = sn * u + cs * w;
        }
        rot++;
      }
    if (!rot) return OR_OK;
  }
  return OR_ENOCONV;
}

/* After jacobi_rows with the noise rule: 1 if the unfinished noise block
 * cannot change the truncation, i.e. the factorization already keeps what a
 * full SVD keeps.  Every singular value of the noise block has a square of at
 * most its weight W; when W is below the smallest squared norm q of the rows
 * above the level, all of the block's singular values sort after those rows,
 * and the truncation discards all of them whatever their split (W <= cut
 * ||X||_F^2), so the kept triplets, their number and the discarded weight are
 * those of the full SVD.  The level is widened by 1e-9 relative, so a row on
 * the boundary counts as noise. */
static int noise_block_exact(int n, int c, const double* X, double cut) {
  double fro = 0.0, W = 0.0, q = -1.0;
  int ns = 0;
  double* a = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int j = 0; j < c; j++) s += X[(long)i * c + j] * X[(long)i * c + j];
    a[i] = s;
    fro += s;
  }
  const double level = cut * fro / (double)n * (1.0 + 1e-9);
  for (int i = 0; i < n; i++) {
    if (a[i] <= level) {
      W += a[i];
      ns++;
    } else if (q < 0.0 || a[i] < q) {
      q = a[i];
    }
  }
  free(a);
  if (ns <= 1) return 1; /* a single noise row is orthogonal to all others */
  if (q < 0.0) return 0;
  return W <= q * (1.0 - 1e-9);
}

long or_trunc_rule(const double* p, long n, double cutoff, long maxdim, double* err) {
  double scale = 0.0;
  for (long i = 0; i < n; i++) scale += p[i];
  long m = n;
  double truncerr = 0.0;
  if (maxdim < 1) maxdim = 1;
  while (m > maxdim) {
    truncerr += p[m - 1];
    m--;
  }
  double sc = scale == 0.0 ? 1.0 : scale;
  if (cutoff >= 0.0) {
    while (m > 1 && truncerr + p[m - 1] <= cutoff * sc) {
      truncerr += p[m - 1];
      m--;
    }
  }
  if (err) *err = truncerr / sc;
  return m;
}

/* Factor M (r x c) = U (r x k) Y (k x c) with U orthonormal and the rows of Y
 * orthogonal, sorted by norm (singular value) descending.  k = min(r, c).
 * When r > c a Householder QR reduces M to its c x c triangle first.  cut:
 * the truncation cutoff that follows (see jacobi_rows; 0 for a full SVD). */
static int svd_factor(int r, int c, const double* M, double* U, double* Y, double* sig, double cut) {
  int k = r < c ? r : c;
  double* Qm = NULL;
  double* X = (double*)malloc(sizeof(double) * (size_t)k * c);
  if (r > c) {
    Qm = (double*)malloc(sizeof(double) * (size_t)r * c);
    or_qr(r, c, M, Qm, X);
  } else {
    memcpy(X, M, sizeof(double) * (size_t)r * c);
  }
  double* J = (double*)malloc(sizeof(double) * (size_t)k * k);
  int rc = jacobi_rows(k, c, X, J, cut, 1);
  /* the noise block could change what is kept: finish the SVD among its rows
     (pairs with a row above the level are converged and are skipped) */
  if (rc == OR_OK && cut > 0.0 && cut < 1.0 && !noise_block_exact(k, c, X, cut))
    rc = jacobi_rows(k, c, X, J, 0.0, 0);
  /* sort rows by norm */
  int* idx = (int*)malloc(sizeof(int) * (size_t)k);
  double* nrm = (double*)malloc(sizeof(double) * (size_t)k);
  for (int i = 0; i < k; i++) {
    double s = 0.0;
    for (int j = 0; j < c; j++) s += X[(long)i * c + j] * X[(long)i * c + j];
    nrm[i] = sqrt(s);
    idx[i] = i;
  }
  for (int i = 1; i < k; i++) { /* insertion sort, stable, descending */
    int t = idx[i];
    int j = i - 1;
    while (j >= 0 && nrm[idx[j]] < nrm[t]) {
      idx[j + 1] = idx[j];
      j--;
    }
    idx[j + 1] = t;
  }
  for (int i = 0; i < k; i++) {
    sig[i] = nrm[idx[i]];
    memcpy(Y + (long)i * c, X + (long)idx[i] * c, sizeof(double) * (size_t)c);
  }
  /* X_in = J^T Y  =>  U = (Qm or I) J^T, column i <- J row idx[i] */
  for (int a = 0; a < r; a++)
    for (int i = 0; i < k; i++) {
      const double* jr = J + (long)idx[i] * k;
      double s;
      if (Qm) {
        s = 0.0;
        for (int b = 0; b < k; b++) s += Qm[(long)a * c + b] * jr[b];
      } else {
        s = jr[a];
      }
      U[(long)a * k + i] = s;
    }
  free(Qm);
  free(X);
  free(J);
  free(idx);
  free(nrm);
  return rc;
}

int or_svd(int r, int c, const double* M, double* U, double* S, double* Vt, double cut) {
  int rc = svd_factor(r, c, M, U, Vt, S, cut);
  int k = r < c ? r : c;
  for (int i = 0; i < k; i++)
    if (S[i] > 0)
      for (int j = 0; j < c; j++) Vt[(long)i * c + j] /= S[i];
  return rc;
}

/* ------------------------------------------------------------------------- */
/* truncation sweep                                                           */
/* ------------------------------------------------------------------------- */
typedef struct {
  tree_t* t;
  const int* sdim;
  long* bond;
  double* const* cores;
  double* w1;
  double* w2;
  double* w3;
  double* w4;
  double* sig;
} sweep_t;

/* axis index of the bond to neighbour u in core x */
static int axis_of(const tree_t* t, int x, int u) {
  if (t->parent[x] == u) return 0;
  const int* ch = tree_ch(t, x);
  for (int k = 0; k < t->nch[x]; k++)
    if (ch[k] == u) return 2 + k;
  return -1;
}

static long* edge_bond(sweep_t* s, int x, int u) {
  return s->t->parent[x] == u ? &s->bond[x] : &s->bond[u];
}

/* factor core x along its bond to u; keep the isometry in x, push the rest into u.
 * truncate != 0 applies the cutoff/maxdim rule. */
static int move_center(sweep_t* s, int x, int u, int truncate, double cutoff, long maxdim, double* err) {
  long dx[64], du[64];
  int ndx = core_dims(s->t, s->sdim, s->bond, x, dx);
  int ndu = core_dims(s->t, s->sdim, s->bond, u, du);
  int ax = axis_of(s->t, x, u), au = axis_of(s->t, u, x);
  long b = dx[ax];
  long rows = prod(dx, ndx) / b;
  axis_to_last(s->cores[x], dx, ndx, ax, s->w1); /* M: rows x b */
  long k = rows < b ? rows : b;
  long m;
  double* Qx = s->w2; /* rows x m */
  double* Rx = s->w3; /* m x b */
  if (!truncate) {
    or_qr((int)rows, (int)b, s->w1, Qx, Rx);
    m = k;
  } else {
    int rc = svd_factor((int)rows, (int)b, s->w1, Qx, Rx, s->sig, cutoff);
    if (rc) return rc;
    double* p = (double*)malloc(sizeof(double) * (size_t)k);
    for (long i = 0; i < k; i++) p[i] = s->sig[i] * s->sig[i];
    double e = 0.0;
    m = or_trunc_rule(p, k, cutoff, maxdim, &e);
    free(p);
    if (err) *err += e;
    if (m < k) { /* compact U columns */
      for (long a = 0; a < rows; a++)
        for (long i = 0; i < m; i++) Qx[a * m + i] = Qx[a * k + i];
    }
  }
  /* write back x with new bond m */
  last_to_axis(Qx, dx, ndx, ax, m, s->cores[x]);
  /* u <- R x_{au} u */
  long nsz = prod(du, ndu) / du[au] * m;
  mode_product(s->cores[u], du, ndu, au, Rx, m, s->w4);
  memcpy(s->cores[u], s->w4, sizeof(double) * (size_t)nsz);
  *edge_bond(s, x, u) = m;
  return OR_OK;
}

int or_truncate(int nv, const int* parent, const int* sdim, long* bond, double* const* cores,
                double cutoff, long maxdim, double* trunc_err) {
  tree_t t;
  int rc = tree_build(&t, nv, parent);
  if (rc) return rc;
  if (check_bonds(nv, bond)) {
    tree_free(&t);
    return OR_EINVAL;
  }
  if (trunc_err) *trunc_err = 0.0;
  if (nv == 1) {
    tree_free(&t);
    return OR_OK;
  }
  long maxsz = 1;
  long dims[64];
  for (int v = 0; v < nv; v++) {
    int nd = core_dims(&t, sdim, bond, v, dims);
    long sz = prod(dims, nd);
    if (sz > maxsz) maxsz = sz;
  }
  sweep_t s = {&t, sdim, bond, cores, NULL, NULL, NULL, NULL, NULL};
  s.w1 = (double*)malloc(sizeof(double) * (size_t)maxsz);
  s.w2 = (double*)malloc(sizeof(double) * (size_t)maxsz);
  s.w3 = (double*)malloc(sizeof(double) * (size_t)maxsz);
  s.w4 = (double*)malloc(sizeof(double) * (size_t)maxsz);
  s.sig = (double*)malloc(sizeof(double) * (size_t)maxsz);

  /* post-order DFS edges from vertex 0: children in increasing order */
  int* src = (int*)malloc(sizeof(int) * (size_t)nv);
  int* stack = (int*)malloc(sizeof(int) * (size_t)nv);
  int* it = (int*)calloc((size_t)nv, sizeof(int));
  int ne = 0, sp = 0;
  stack[sp++] = 0;
  while (sp) {
    int v = stack[sp - 1];
    if (it[v] < t.nch[v]) {
      stack[sp++] = tree_ch(&t, v)[it[v]++];
    } else {
      sp--;
      if (v != 0) src[ne++] = v; /* edge v -> parent[v] */
    }
  }
  /* orthogonalise toward src[0] from no gauge: every vertex moves its weight
     one step toward the target, farthest first (BFS from the target). */
  int target = src[0];
  int* order = (int*)malloc(sizeof(int) * (size_t)nv);
  int* via = (int*)malloc(sizeof(int) * (size_t)nv); /* neighbour toward target */
  int* seen = (int*)calloc((size_t)nv, sizeof(int));
  int head = 0, tail = 0;
  order[tail++] = target;
  seen[target] = 1;
  via[target] = -1;
  while (head < tail) {
    int v = order[head++];
    int nb[66];
    int nn = 0;
    if (parent[v] >= 0) nb[nn++] = parent[v];
    for (int k = 0; k < t.nch[v]; k++) nb[nn++] = tree_ch(&t, v)[k];
    for (int k = 0; k < nn; k++)
      if (!seen[nb[k]]) {
        seen[nb[k]] = 1;
        via[nb[k]] = v;
        order[tail++] = nb[k];
      }
  }
  for (int i = nv - 1; i >= 1 && !rc; i--) rc = move_center(&s, order[i], via[order[i]], 0, 0, 0, NULL);
  int center = target;
  for (int e = 0; e < ne && !rc; e++) {
    int c = src[e], p = parent[c];
    /* move the centre to c along the tree path */
    while (center != c && !rc) {
      /* next vertex from center toward c: if c is in center's subtree go down, else up */
      int step;
      int x = c;
      int below = 0;
      while (x != -1) {
        if (parent[x] == center) {
          below = 1;
          break;
        }
        x = parent[x];
      }
      step = below ? x : parent[center];
      rc = move_center(&s, center, step, 0, 0, 0, NULL);
      center = step;
    }
    if (!rc) rc = move_center(&s, c, p, 1, cutoff, maxdim, trunc_err);
    center = p;
  }
  free(src);
  free(stack);
  free(it);
  free(order);
  free(via);
  free(seen);
  free(s.w1);
  free(s.w2);
  free(s.w3);
  free(s.w4);
  free(s.sig);
  tree_free(&t);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* gradient                                                                   */
/* ------------------------------------------------------------------------- */
int or_eval_grad(int nv, const int* parent, const int* sdim, const long* bond, const double* const* cores,
                 const int* vcoord, const int* vdigit, int ncoord, const int* Lc, const double* xs,
                 long npts, const double* ybar, double* const* grads, int nthreads) {
  tree_t t;
  int rc = tree_build(&t, nv, parent);
  if (rc) return rc;
  if (check_bonds(nv, bond)) {
    tree_free(&t);
    return OR_EINVAL;
  }
  long* voff = (long*)malloc(sizeof(long) * (size_t)(nv + 1));
  long* goff = (long*)malloc(sizeof(long) * (size_t)(nv + 1));
  long tot = 0, gtot = 0, dims[64];
  for (int v = 0; v < nv; v++) {
    voff[v] = tot;
    tot += bond[v];
    goff[v] = gtot;
    int nd = core_dims(&t, sdim, bond, v, dims);
    gtot += prod(dims, nd);
  }
  if (nthreads < 1) nthreads = 1;
  double* gall = (double*)calloc((size_t)gtot * nthreads, sizeof(double));
  int bad = 0;
#pragma omp parallel num_threads(nthreads) reduction(| : bad)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* G = gall + (size_t)tid * gtot;
    double* vec = (double*)malloc(sizeof(double) * (size_t)tot);
    double* env = (double*)malloc(sizeof(double) * (size_t)tot);
    int* dig = (int*)malloc(sizeof(int) * (size_t)nv);
    long idx[64];
#pragma omp for schedule(static)
    for (long i = 0; i < npts; i++) {
      const double* x = xs + i * ncoord;
      int ok = 1;
      for (int v = 0; v < nv; v++) {
        if (vcoord[v] < 0) {
          dig[v] = 0;
          continue;
        }
        int d = or_digit(x[vcoord[v]], Lc[vcoord[v]], vdigit[v]);
        if (d < 0) {
          ok = 0;
          break;
        }
        dig[v] = d;
      }
      if (!ok) {
        bad = 1;
        continue;
      }
      for (int v = nv - 1; v >= 0; v--) {
        int m = t.nch[v];
        const int* ch = tree_ch(&t, v);
        long pc = 1;
        for (int k = 0; k < m; k++) pc *= bond[ch[k]];
        for (long p = 0; p < bond[v]; p++)
          vec[voff[v] + p] =
              contract_children(cores[v] + (p * sdim[v] + dig[v]) * pc, m, ch, bond, vec, voff);
      }
      memset(env, 0, sizeof(double) * (size_t)tot);
      env[voff[0]] = 1.0;
      double w = ybar[i];
      for (int v = 0; v < nv; v++) {
        int m = t.nch[v];
        const int* ch = tree_ch(&t, v);
        long pc = 1;
        for (int k = 0; k < m; k++) pc *= bond[ch[k]];
        for (long p = 0; p < bond[v]; p++) {
          double ep = env[voff[v] + p];
          const double* T = cores[v] + (p * sdim[v] + dig[v]) * pc;
          double* g = G + goff[v] + (p * sdim[v] + dig[v]) * pc;
          for (long j = 0; j < pc; j++) {
            long r = j;
            for (int k = m - 1; k >= 0; k--) {
              idx[k] = r % bond[ch[k]];
              r /= bond[ch[k]];
            }
            double wp = 1.0;
            for (int k = 0; k < m; k++) wp *= vec[voff[ch[k]] + idx[k]];
            g[j] += w * ep * wp;
            if (ep != 0.0 && T[j] != 0.0) {
              for (int k = 0; k < m; k++) {
                double o = 1.0;
                for (int l = 0; l < m; l++)
                  if (l != k) o *= vec[voff[ch[l]] + idx[l]];
                env[voff[ch[k]] + idx[k]] += ep * T[j] * o;
              }
            }
          }
        }
      }
    }
    free(vec);
    free(env);
    free(dig);
  }
  for (int v = 0; v < nv; v++) {
    int nd = core_dims(&t, sdim, bond, v, dims);
    long sz = prod(dims, nd);
    for (long j = 0; j < sz; j++) {
      double s = 0.0;
      for (int th = 0; th < nthreads; th++) s += gall[(size_t)th * gtot + goff[v] + j];
      grads[v][j] = s;
    }
  }
  free(gall);
  free(voff);
  free(goff);
  tree_free(&t);
  return bad ? OR_ERANGE : OR_OK;
}
