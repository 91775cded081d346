// Evaluation of networks on graphs with loops by belief propagation (gfx950).
//
// Replaces evaluate(f, xs; alg="bp") on a loopy layout
// (src/itensornetworkfunction.jl:96-106 -> ITensorNetworks.scalar(tn; alg="bp"),
// the grids of test/test_realitensorfunction.jl:41): after the digits of a
// point are projected, every vertex tensor has only bond legs, and BP
// estimates the scalar of that network as prod_v Z_v / prod_e Z_e from the
// fixed point of its messages.  On a tree the estimate is exact.
//
// One lane per point.  Messages live in global memory as [entry][point]
// (coalesced across the lanes of a wavefront); the projected core slices are
// read in place.  Schedule: a sweep visits the vertices in order, and each
// vertex sends its messages over its legs in edge order; every message is
// replaced as soon as it is computed and normalised to unit 2-norm (so later
// messages of the sweep already use it).  Messages start uniform (all ones)
// and sweeps stop when the largest change of a sweep is at most tol, or after
// maxiter sweeps.
#include "qtt_internal.h"

#include <algorithm>
#include <cstring>

namespace {

constexpr int kMaxLegs = QTT_MAX_NDIM - 1;  // bond legs per vertex
constexpr int kMaxBond = 32;                // message entries kept in registers

struct BpVertex {
  int64_t core;                // offset of the core [sdim][legs...]
  int64_t slice;               // elements per digit value (product of the leg bonds)
  int nleg, sdim, coord, shift;
  int leg_dim[kMaxLegs];
  int64_t leg_stride[kMaxLegs];
  int in_msg[kMaxLegs];        // message arriving over each leg (offset into the message entries)
  int out_msg[kMaxLegs];       // message leaving over each leg
};

struct BpArgs {
  const double* cores;
  const BpVertex* vert;
  int nv, ne;
  const int* edge_msg;         // per edge: offsets of its two messages; bond
  double* msg;                 // [entry][point]
  int nmsg;                    // message entries per point
  int ncoord;
  const int* ndigits;
  const double* xs;
  int64_t npts;
  double* out;
  int maxiter;
  double tol;
  int* flag;                   // status word: invalid coordinate
  int* iters;                  // largest sweep count over the points
};

__device__ inline uint64_t bp_digits(double x, int L, bool& bad) {
  if (!(x >= 0.0)) {
    bad = true;
    return 0;
  }
  if (x >= 1.0) return (uint64_t(1) << L) - 1;
  return (uint64_t)(x * (double)(uint64_t(1) << L));
}

// sum over the slice of vertex v at digit value d, all legs but `skip`
// contracted with their incoming messages; acc[j] collects leg `skip`'s
// entries (skip < 0: every leg contracted, the total in acc[0])
__device__ void bp_contract(const BpArgs& a, const BpVertex& V, int d, int skip, int64_t p, double* acc) {
  const double* T = a.cores + V.core + (int64_t)d * V.slice;
  const int bs = skip >= 0 ? V.leg_dim[skip] : 1;
  for (int j = 0; j < bs; ++j) acc[j] = 0.0;
  for (int64_t t = 0; t < V.slice; ++t) {
    int64_t r = t;
    double w = T[t];
    int js = 0;
    for (int k = V.nleg - 1; k >= 0; --k) {
      const int i = (int)(r % V.leg_dim[k]);
      r /= V.leg_dim[k];
      if (k == skip) js = i;
      else w *= a.msg[(int64_t)(V.in_msg[k] + i) * a.npts + p];
    }
    acc[js] += w;
  }
}

__global__ void __launch_bounds__(256) bp_eval_kernel(BpArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npts) return;
  // digit of every vertex
  uint64_t words[QTT_MAX_COORD];
  bool bad = false;
  for (int c = 0; c < a.ncoord && c < QTT_MAX_COORD; ++c) words[c] = bp_digits(a.xs[p * a.ncoord + c], a.ndigits[c], bad);
  if (bad) {
    a.out[p] = __builtin_nan("");
    atomicOr(a.flag, 1);
    return;
  }
  auto digit = [&](const BpVertex& V) -> int { return V.coord < 0 ? 0 : (int)((words[V.coord] >> V.shift) & 1); };
  // messages start uniform
  for (int m = 0; m < a.nmsg; ++m) a.msg[(int64_t)m * a.npts + p] = 1.0;
  double acc[kMaxBond];
  int it = 0;
  for (; it < a.maxiter; ++it) {
    double change = 0.0;
    for (int v = 0; v < a.nv; ++v) {
      const BpVertex& V = a.vert[v];
      const int d = digit(V);
      for (int k = 0; k < V.nleg; ++k) {
        bp_contract(a, V, d, k, p, acc);
        double nrm = 0.0;
        for (int j = 0; j < V.leg_dim[k]; ++j) nrm += acc[j] * acc[j];
        const double s = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
        for (int j = 0; j < V.leg_dim[k]; ++j) {
          double* m = a.msg + (int64_t)(V.out_msg[k] + j) * a.npts + p;
          const double nv = acc[j] * s;
          change = fmax(change, fabs(nv - *m));
          *m = nv;
        }
      }
    }
    if (change <= a.tol) {
      ++it;
      break;
    }
  }
  atomicMax(a.iters, it);
  // prod_v Z_v / prod_e Z_e
  double val = 1.0;
  for (int v = 0; v < a.nv; ++v) {
    const BpVertex& V = a.vert[v];
    bp_contract(a, V, digit(V), -1, p, acc);
    val *= acc[0];
  }
  for (int e = 0; e < a.ne; ++e) {
    const int m0 = a.edge_msg[3 * e], m1 = a.edge_msg[3 * e + 1], b = a.edge_msg[3 * e + 2];
    double z = 0.0;
    for (int j = 0; j < b; ++j) z += a.msg[(int64_t)(m0 + j) * a.npts + p] * a.msg[(int64_t)(m1 + j) * a.npts + p];
    val /= z;
  }
  a.out[p] = val;
}

}  // namespace

extern "C" {

int qtt_bp_eval_host(qtt_ctx* ctx, const qtt_graph* g, const double* const* cores, const double* xs, int64_t npts,
                     double* out, int maxiter, double tol, int* iters_out) {
  if (!ctx || !g || !cores || (npts && (!xs || !out))) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  if (npts < 0 || maxiter < 1 || !(tol >= 0.0)) return qtt_fail(ctx, QTT_EINVAL, "bp: bad npts, maxiter or tol");
  const int nv = g->nv, ne = g->ne;
  if (nv < 1 || ne < 0 || (ne && (!g->eu || !g->ev || !g->ebond)) || !g->sdim || !g->vcoord || !g->vdigit)
    return qtt_fail(ctx, QTT_EINVAL, "bp: bad graph");
  if (g->ncoord < 0 || g->ncoord > QTT_MAX_COORD || (g->ncoord && !g->ndigits))
    return qtt_fail(ctx, QTT_EINVAL, "bp: at most 4 coordinates");
  for (int c = 0; c < g->ncoord; c++)
    if (g->ndigits[c] < 0 || g->ndigits[c] > 62) return qtt_fail(ctx, QTT_EINVAL, "bp: digits per coordinate in 0..62");
  // legs of each vertex: its edges in increasing edge id
  std::vector<BpVertex> vert(nv);
  std::vector<int> edge_msg(3 * std::max(ne, 1));
  int nmsg = 0;
  for (int v = 0; v < nv; v++) {
    std::memset(&vert[v], 0, sizeof(BpVertex));
    if (g->sdim[v] < 1 || g->sdim[v] > 2) return qtt_fail(ctx, QTT_EINVAL, "bp: sites are binary or absent");
    vert[v].sdim = g->sdim[v];
    const int c = g->vcoord[v];
    if (c >= 0) {
      if (c >= g->ncoord || g->sdim[v] != 2 || g->vdigit[v] < 1 || g->vdigit[v] > g->ndigits[c])
        return qtt_fail(ctx, QTT_EINVAL, "bp: digit out of range");
      vert[v].coord = c;
      vert[v].shift = g->ndigits[c] - g->vdigit[v];
    } else {
      if (g->sdim[v] != 1) return qtt_fail(ctx, QTT_EINVAL, "bp: a binary site needs a digit");
      vert[v].coord = -1;
    }
  }
  for (int e = 0; e < ne; e++) {
    const int u = g->eu[e], w = g->ev[e];
    const int64_t b = g->ebond[e];
    if (u < 0 || u >= nv || w < 0 || w >= nv || u == w) return qtt_fail(ctx, QTT_EINVAL, "bp: bad edge");
    if (b < 1 || b > kMaxBond) return qtt_fail(ctx, QTT_EINVAL, "bp: bonds must be in 1..32");
    // message u -> w at nmsg, w -> u after it
    edge_msg[3 * e] = nmsg;
    edge_msg[3 * e + 1] = nmsg + (int)b;
    edge_msg[3 * e + 2] = (int)b;
    for (int side = 0; side < 2; side++) {
      BpVertex& V = vert[side == 0 ? u : w];
      if (V.nleg >= kMaxLegs) return qtt_fail(ctx, QTT_EINVAL, "bp: vertex degree too large");
      V.leg_dim[V.nleg] = (int)b;
      V.out_msg[V.nleg] = side == 0 ? nmsg : nmsg + (int)b;
      V.in_msg[V.nleg] = side == 0 ? nmsg + (int)b : nmsg;
      V.nleg++;
    }
    nmsg += 2 * (int)b;
  }
  int64_t tot = 0;
  std::vector<int64_t> size(nv);
  for (int v = 0; v < nv; v++) {
    int64_t s = 1;
    for (int k = vert[v].nleg - 1; k >= 0; --k) {
      vert[v].leg_stride[k] = s;
      s *= vert[v].leg_dim[k];
    }
    if (s > (int64_t(1) << 24)) return qtt_fail(ctx, QTT_EINVAL, "bp: core too large");
    vert[v].slice = s;
    vert[v].core = tot;
    size[v] = s * vert[v].sdim;
    tot += size[v];
    if (!cores[v]) return qtt_fail(ctx, QTT_EINVAL, "null core");
  }
  if (npts == 0) {
    if (iters_out) *iters_out = 0;
    return QTT_OK;
  }
  // one allocation: cores | vertices | edge table | points | values | messages | counter
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t b_cores = al(tot * sizeof(double)), b_vert = al(nv * sizeof(BpVertex)),
               b_edge = al(edge_msg.size() * sizeof(int)), b_xs = al((size_t)npts * std::max(g->ncoord, 1) * sizeof(double)),
               b_out = al((size_t)npts * sizeof(double)), b_msg = al((size_t)npts * std::max(nmsg, 1) * sizeof(double)),
               b_nd = al(std::max(g->ncoord, 1) * sizeof(int));
  char* d = nullptr;
  int rc = qtt_hip_check(ctx, hipMalloc(&d, b_cores + b_vert + b_edge + b_xs + b_out + b_msg + b_nd + 256), "hipMalloc bp");
  if (rc) return rc;
  char* q = d;
  double* d_cores = (double*)q; q += b_cores;
  BpVertex* d_vert = (BpVertex*)q; q += b_vert;
  int* d_edge = (int*)q; q += b_edge;
  double* d_xs = (double*)q; q += b_xs;
  double* d_out = (double*)q; q += b_out;
  double* d_msg = (double*)q; q += b_msg;
  int* d_nd = (int*)q; q += b_nd;
  int* d_iters = (int*)q;
  {
    int rs = QTT_OK;
    for (int v = 0; v < nv && !rs; v++)
      rs = qtt_hip_check(ctx, hipMemcpyAsync(d_cores + vert[v].core, cores[v], size[v] * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "bp cores");
    if (!rs) rs = qtt_hip_check(ctx, hipMemcpyAsync(d_vert, vert.data(), nv * sizeof(BpVertex), hipMemcpyHostToDevice, ctx->stream), "bp vertices");
    if (!rs) rs = qtt_hip_check(ctx, hipMemcpyAsync(d_edge, edge_msg.data(), edge_msg.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream), "bp edges");
    if (!rs && g->ncoord)
      rs = qtt_hip_check(ctx, hipMemcpyAsync(d_xs, xs, (size_t)npts * g->ncoord * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "bp points");
    if (!rs && g->ncoord)
      rs = qtt_hip_check(ctx, hipMemcpyAsync(d_nd, g->ndigits, g->ncoord * sizeof(int), hipMemcpyHostToDevice, ctx->stream), "bp digits");
    const int zero = 0;
    if (!rs) rs = qtt_hip_check(ctx, hipMemcpyAsync(d_iters, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream), "bp counter");
    rc = rs;
  }
  int iters = 0;
  if (!rc) {
    BpArgs a;
    a.cores = d_cores;
    a.vert = d_vert;
    a.nv = nv;
    a.ne = ne;
    a.edge_msg = d_edge;
    a.msg = d_msg;
    a.nmsg = nmsg;
    a.ncoord = g->ncoord;
    a.ndigits = d_nd;
    a.xs = d_xs;
    a.npts = npts;
    a.out = d_out;
    a.maxiter = maxiter;
    a.tol = tol;
    a.flag = ctx->d_flag + 2;
    a.iters = d_iters;
    const int64_t nb = (npts + 255) / 256;
    if (nb > INT32_MAX) rc = qtt_fail(ctx, QTT_EINVAL, "bp: too many points");
    if (!rc) {
      bp_eval_kernel<<<dim3((unsigned)nb), 256, 0, ctx->stream>>>(a);
      rc = qtt_hip_check(ctx, hipGetLastError(), "bp_eval_kernel");
    }
    if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(out, d_out, npts * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "bp values");
    if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(&iters, d_iters, sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "bp sweeps");
  }
  if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "bp sync");
  if (!rc) {
    int fl = 0;
    rc = qtt_hip_check(ctx, hipMemcpy(&fl, ctx->d_flag + 2, sizeof(int), hipMemcpyDeviceToHost), "bp status");
    if (!rc && fl) {
      (void)hipMemset(ctx->d_flag + 2, 0, sizeof(int));
      rc = qtt_fail(ctx, QTT_ERANGE, "a coordinate was negative or NaN");
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (!rc && iters_out) *iters_out = iters;
  return rc;
}

}  // extern "C"
