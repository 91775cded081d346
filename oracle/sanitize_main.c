/*
 * Drives every oracle entry point on small networks (a chain, a comb, a
 * vertex without digit, bond-1 edges, rank-deficient and zero matrices, an
 * invalid point) for the AddressSanitizer / UndefinedBehaviorSanitizer build
 * (`make -C oracle sanitize`, run by tests/test_sanitize.py).  TEST
 * INFRASTRUCTURE ONLY.  Exit status 0 when every call returned the expected
 * status; the sanitizers abort on the first memory or undefined-behaviour
 * error.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static unsigned long long rng_state = 88172645463325252ULL;
static double urand(void) {  /* xorshift64*, uniform in [-1, 1) */
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return (double)((rng_state * 2685821657736338717ULL) >> 11) / 4503599627370496.0 - 1.0;
}

#define MAXV 16
typedef struct {
  int nv, ncoord;
  int parent[MAXV], sdim[MAXV], vcoord[MAXV], vdigit[MAXV], Lc[4];
  long bond[MAXV];
  double* cores[MAXV];
} Net;

static long core_size(const Net* n, int v, int op) {
  long s = n->bond[v] * n->sdim[v] * (op ? n->sdim[v] : 1);
  for (int c = v + 1; c < n->nv; c++)
    if (n->parent[c] == v) s *= n->bond[c];
  return s;
}

static void fill(Net* n, int op) {
  for (int v = 0; v < n->nv; v++) {
    long s = core_size(n, v, op);
    n->cores[v] = (double*)malloc(sizeof(double) * (size_t)s);
    for (long i = 0; i < s; i++) n->cores[v][i] = urand();
  }
}

static void release(Net* n) {
  for (int v = 0; v < n->nv; v++) free(n->cores[v]);
}

static int fails = 0;
#define EXPECT(cond, what)                               \
  do {                                                   \
    if (!(cond)) {                                       \
      fprintf(stderr, "sanitize: %s failed\n", what);    \
      fails++;                                           \
    }                                                    \
  } while (0)

static void run(Net* f, Net* w) {
  const int nv = f->nv;
  double xs[3 * 4] = {0.3, 0.7, 0.1, 0.999, 0.0, 0.5, 1.0, 0.25, 0.0, 0.0, 0.0, 0.0};
  double out[4];
  int rc = or_eval(nv, f->parent, f->sdim, f->bond, (const double* const*)f->cores, f->vcoord, f->vdigit,
                   f->ncoord, f->Lc, xs, 3, out, 1);
  EXPECT(rc == OR_OK, "or_eval");
  double bad[4] = {NAN, 0.5, -1.0, 0.5};
  rc = or_eval(nv, f->parent, f->sdim, f->bond, (const double* const*)f->cores, f->vcoord, f->vdigit, f->ncoord,
               f->Lc, bad, 2 / (f->ncoord ? f->ncoord : 1), out, 1);
  EXPECT(rc == OR_ERANGE, "or_eval invalid point");
  double integral = 0.0, ip = 0.0;
  EXPECT(or_integrate(nv, f->parent, f->sdim, f->bond, (const double* const*)f->cores, 0.5, &integral) == OR_OK,
         "or_integrate");
  EXPECT(or_inner(nv, f->parent, f->sdim, f->bond, (const double* const*)f->cores, f->bond,
                  (const double* const*)f->cores, &ip) == OR_OK && ip >= 0.0, "or_inner");
  /* exact product, its adjoint, then truncation of the product */
  long pb[MAXV];
  double* C[MAXV];
  double* gbar[MAXV];
  double* abar[MAXV];
  double* wbar[MAXV];
  for (int v = 0; v < nv; v++) pb[v] = f->bond[v] * w->bond[v];
  Net P = *f;
  for (int v = 0; v < nv; v++) P.bond[v] = pb[v];
  for (int v = 0; v < nv; v++) {
    long s = core_size(&P, v, 0);
    C[v] = (double*)malloc(sizeof(double) * (size_t)s);
    gbar[v] = (double*)malloc(sizeof(double) * (size_t)s);
    for (long i = 0; i < s; i++) gbar[v][i] = urand();
    abar[v] = (double*)malloc(sizeof(double) * (size_t)core_size(f, v, 0));
    wbar[v] = (double*)malloc(sizeof(double) * (size_t)core_size(w, v, 1));
  }
  EXPECT(or_contract(nv, f->parent, f->sdim, w->bond, (const double* const*)w->cores, f->bond,
                     (const double* const*)f->cores, C) == OR_OK, "or_contract");
  EXPECT(or_contract_adjoint(nv, f->parent, f->sdim, w->bond, (const double* const*)w->cores, f->bond,
                             (const double* const*)f->cores, (const double* const*)gbar, abar, wbar) == OR_OK,
         "or_contract_adjoint");
  double err = -1.0;
  for (int pass = 0; pass < 3; pass++) {  /* cutoff rule, maxdim rule, then truncating an already truncated net */
    const double cut = pass == 0 ? 1e-3 : 0.0;
    const long md = pass == 1 ? 2 : 1 << 20;
    EXPECT(or_truncate(nv, f->parent, f->sdim, pb, C, cut, md, &err) == OR_OK && err >= 0.0, "or_truncate");
  }
  /* gradient of the truncated product at the points */
  double ybar[3] = {1.0, -0.5, 2.0};
  double* grads[MAXV];
  for (int v = 0; v < nv; v++) {
    Net T = P;
    for (int u = 0; u < nv; u++) T.bond[u] = pb[u];
    grads[v] = (double*)calloc((size_t)core_size(&T, v, 0), sizeof(double));
  }
  EXPECT(or_eval_grad(nv, f->parent, f->sdim, pb, (const double* const*)C, f->vcoord, f->vdigit, f->ncoord, f->Lc,
                      xs, 3, ybar, grads, 1) == OR_OK, "or_eval_grad");
  for (int v = 0; v < nv; v++) {
    free(C[v]);
    free(gbar[v]);
    free(abar[v]);
    free(wbar[v]);
    free(grads[v]);
  }
}

static void factorizations(void) {
  const int shapes[][2] = {{1, 1}, {5, 3}, {3, 5}, {17, 8}, {8, 17}};
  for (size_t i = 0; i < sizeof(shapes) / sizeof(shapes[0]); i++) {
    const int r = shapes[i][0], c = shapes[i][1], k = r < c ? r : c;
    double* M = (double*)malloc(sizeof(double) * (size_t)(r * c));
    double* U = (double*)malloc(sizeof(double) * (size_t)(r * k));
    double* S = (double*)malloc(sizeof(double) * (size_t)k);
    double* Vt = (double*)malloc(sizeof(double) * (size_t)(k * c));
    for (int z = 0; z < 3; z++) {  /* random, rank one, zero */
      for (int a = 0; a < r; a++)
        for (int b = 0; b < c; b++) M[a * c + b] = z == 0 ? urand() : z == 1 ? (a + 1.0) * (b - 2.0) : 0.0;
      EXPECT(or_svd(r, c, M, U, S, Vt, z == 0 ? 0.0 : 0.1) == OR_OK, "or_svd");
      EXPECT(or_qr(r, c, M, U, Vt) == OR_OK, "or_qr");
    }
    double err = 0.0;
    S[0] = 1.0;
    for (int j = 1; j < k; j++) S[j] = S[j - 1] * 0.1;
    EXPECT(or_trunc_rule(S, k, 1e-4, 1 << 20, &err) >= 1 && err >= 0.0, "or_trunc_rule");
    EXPECT(or_trunc_rule(S, k, 0.0, 1, &err) == 1, "or_trunc_rule maxdim");
    free(M);
    free(U);
    free(S);
    free(Vt);
  }
}

int main(void) {
  /* chain of 6 binary digits of one coordinate, bonds 1,2,3,4,3,2; operator bonds 1,3,3,3,3,3 */
  Net f = {.nv = 6, .ncoord = 1, .Lc = {6}};
  Net w = {.nv = 6, .ncoord = 1, .Lc = {6}};
  const long fb[6] = {1, 2, 3, 4, 3, 2}, wb[6] = {1, 3, 3, 3, 3, 3};
  for (int v = 0; v < 6; v++) {
    f.parent[v] = w.parent[v] = v - 1;
    f.sdim[v] = w.sdim[v] = 2;
    f.vcoord[v] = w.vcoord[v] = 0;
    f.vdigit[v] = w.vdigit[v] = v + 1;
    f.bond[v] = fb[v];
    w.bond[v] = wb[v];
  }
  fill(&f, 0);
  fill(&w, 1);
  run(&f, &w);
  release(&f);
  release(&w);
  /* comb: backbone 0 - 3, teeth 0 -> 1 -> 2 and 3 -> 4 -> 5, vertex 3 without digit,
     two coordinates; one bond-1 edge */
  Net g = {.nv = 6, .ncoord = 2, .Lc = {3, 2}};
  Net o = {.nv = 6, .ncoord = 2, .Lc = {3, 2}};
  const int gp[6] = {-1, 0, 1, 0, 3, 4}, gs[6] = {2, 2, 2, 1, 2, 2};
  const int gc[6] = {0, 0, 0, -1, 1, 1}, gd[6] = {1, 2, 3, 0, 1, 2};
  const long gb[6] = {1, 2, 2, 1, 3, 2}, ob[6] = {1, 2, 2, 1, 2, 2};
  for (int v = 0; v < 6; v++) {
    g.parent[v] = o.parent[v] = gp[v];
    g.sdim[v] = o.sdim[v] = gs[v];
    g.vcoord[v] = o.vcoord[v] = gc[v];
    g.vdigit[v] = o.vdigit[v] = gd[v];
    g.bond[v] = gb[v];
    o.bond[v] = ob[v];
  }
  fill(&g, 0);
  fill(&o, 1);
  run(&g, &o);
  release(&g);
  release(&o);
  factorizations();
  EXPECT(or_digit(0.625, 3, 1) == 1 && or_digit(NAN, 3, 1) == -1 && or_digit(2.0, 4, 4) == 1, "or_digit");
  if (fails) {
    fprintf(stderr, "sanitize: %d checks failed\n", fails);
    return 1;
  }
  printf("sanitize: all oracle entry points ran clean\n");
  return 0;
}
