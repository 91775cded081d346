// Function algebra on device-resident networks (gfx950): the Hadamard product
// of two functions, their direct sum, and partial integration over some
// coordinates with the integrated vertices merged away.
//
//   qtt_multiply          multiply(g, f)           src/elementary_operators.jl:260-282
//   qtt_add               ITensorNetworks.add      src/fixes.jl:40-95
//   qtt_partial_integrate partial_integrate        src/integration.jl:35-54
//                         (+ merge_internal_tensors, src/utils.jl:99-113)
//   qtt_operator_proj     operator_proj            src/elementary_operators.jl:246-258
//
// Every output core is one job of an elementwise kernel over the output
// core's elements.  The host plans the jobs from the two layouts (which
// operand leg feeds which output leg, with which stride); the kernels only
// gather.  Partial integration first contracts the merged vertices into
// vectors on the edges to the remaining vertices (one workgroup, vertices in
// dependency order, one wavefront per vector entry), then forms each
// remaining core contracted with those vectors.  All of this is memory bound
// and small next to operate; the kernels are plain coalesced gathers.
#include "qtt_internal.h"

#include <algorithm>
#include <cstring>

namespace {

constexpr int kL = QTT_MAX_NDIM;

// One output core of a product or sum.  Leg k of the output has dim[k]
// entries; split[k] tells how an output index i maps to the operands':
//   product: f index i / split, g index i % split (split = g's bond); 0: the
//            same index in both (the site);
//   sum:     f's bond: i < split is f's block (index i), else g's (i - split);
//            0: the same index in both (the site, and the root's unit bond).
// fs/gs: the operands' strides of that leg (0 where the operand lacks it: a
// missing vertex, a missing edge, or no digit at the vertex).
struct PairJob {
  const double* f;
  const double* g;
  double* out;
  int nleg;
  int dim[kL];
  int split[kL];
  int64_t fs[kL], gs[kL];
  int64_t nelem;
};

template <bool ADD>
__global__ void __launch_bounds__(256) pair_kernel(const PairJob* __restrict__ jobs) {
  const PairJob& J = jobs[blockIdx.y];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < J.nelem;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t rest = e, fo = 0, go = 0;
    bool inf = false, ing = false;
    for (int k = J.nleg - 1; k >= 0; --k) {
      const int i = (int)(rest % J.dim[k]);
      rest /= J.dim[k];
      const int sp = J.split[k];
      if (sp == 0) {
        fo += i * J.fs[k];
        go += i * J.gs[k];
      } else if (ADD) {
        if (i < sp) {
          inf = true;
          fo += i * J.fs[k];
        } else {
          ing = true;
          go += (i - sp) * J.gs[k];
        }
      } else {
        fo += (i / sp) * J.fs[k];
        go += (i % sp) * J.gs[k];
      }
    }
    double v;
    if (ADD) {
      v = (inf && ing) ? 0.0 : inf ? J.f[fo] : ing ? J.g[go] : J.f[fo] + J.g[go];
    } else {
      v = (J.f ? J.f[fo] : 1.0) * (J.g ? J.g[go] : 1.0);
    }
    J.out[e] = v;
  }
}

// A core contracted with vectors on some legs: dst[o...] =
// sum_{c...} src[o... c...] * prod_k cv[k][c_k].  Output legs are row-major
// in the order given (os: their strides in src).
struct LegJob {
  const double* src;
  double* dst;
  int nout, ncon;
  int odim[kL];
  int64_t os[kL];
  int cdim[kL];
  int64_t cs[kL];
  const double* cv[kL];
  int64_t nelem, ncomb;
};

__device__ double leg_sum(const LegJob& J, int64_t e, int64_t t0, int64_t tstep) {
  int64_t base = 0, rest = e;
  for (int k = J.nout - 1; k >= 0; --k) {
    base += (rest % J.odim[k]) * J.os[k];
    rest /= J.odim[k];
  }
  double acc = 0.0;
  for (int64_t t = t0; t < J.ncomb; t += tstep) {
    int64_t off = base, r = t;
    double w = 1.0;
    for (int k = J.ncon - 1; k >= 0; --k) {
      const int c = (int)(r % J.cdim[k]);
      r /= J.cdim[k];
      off += c * J.cs[k];
      w *= J.cv[k][c];
    }
    acc += J.src[off] * w;
  }
  return acc;
}

// The environment vectors, in dependency order, by one workgroup: one
// wavefront per entry, its lanes splitting the contracted sum.
__global__ void __launch_bounds__(512) leg_env_kernel(const LegJob* __restrict__ jobs, int njobs) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int j = 0; j < njobs; ++j) {
    const LegJob& J = jobs[j];
    for (int64_t e = wid; e < J.nelem; e += nw) {
      double acc = leg_sum(J, e, lane, 64);
      for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
      if (lane == 0) J.dst[e] = acc;
    }
    __syncthreads();  // the next vertex reads these entries
  }
}

// The remaining cores, one job per core, one thread per element.
__global__ void __launch_bounds__(256) leg_out_kernel(const LegJob* __restrict__ jobs) {
  const LegJob& J = jobs[blockIdx.y];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < J.nelem;
       e += (int64_t)gridDim.x * blockDim.x)
    J.dst[e] = leg_sum(J, e, 0, 1);
}

// Legs of core v: dims and row-major strides ([parent][site][children...]).
int core_legs(const HostTopo& t, const std::vector<int64_t>& bond, int v, int* dims, int64_t* st) {
  const int r = 2 + t.nch[v];
  dims[0] = (int)bond[v];
  dims[1] = t.sdim[v];
  for (int k = 0; k < t.nch[v]; k++) dims[2 + k] = (int)bond[t.chs[t.choff[v] + k]];
  int64_t s = 1;
  for (int k = r - 1; k >= 0; --k) {
    st[k] = s;
    s *= dims[k];
  }
  return r;
}

// Leg of vertex a that joins neighbour b (0: the parent edge, 2 + k: child
// k), or -1 when they are not adjacent.
int leg_to(const HostTopo& t, int a, int b) {
  if (a < 0 || b < 0) return -1;
  if (t.parent[a] == b) return 0;
  for (int k = 0; k < t.nch[a]; k++)
    if (t.chs[t.choff[a] + k] == b) return 2 + k;
  return -1;
}

// bond of the edge a - b of a network (1 when there is no such edge)
int64_t edge_bond(const qtt_net* n, int a, int b) {
  if (a < 0 || b < 0) return 1;
  if (n->t.parent[a] == b) return n->bond[a];
  if (n->t.parent[b] == a) return n->bond[b];
  return 1;
}

// neighbour of output vertex o across its leg k (-1 for the root's parent leg)
int out_neighbour(const HostTopo& t, int o, int k) {
  if (k == 0) return t.parent[o];
  return t.chs[t.choff[o] + (k - 2)];
}

// The diagonal operator of a function, |f><delta|: operator core
// [p][s_in][s_out][children] = f[p][s][children] when s_in = s_out = s, else 0.
struct ProjJob {
  const double* f;
  double* out;
  int sd;
  int64_t rest;  // product of the child bonds
  int64_t nelem;
};

__global__ void __launch_bounds__(256) proj_kernel(const ProjJob* __restrict__ jobs) {
  const ProjJob& J = jobs[blockIdx.y];
  const int64_t ss = (int64_t)J.sd * J.sd;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < J.nelem;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % J.rest;
    const int64_t q = e / J.rest;  // p * sd * sd + s_in * sd + s_out
    const int so = (int)(q % J.sd), si = (int)((q / J.sd) % J.sd);
    const int64_t p = q / ss;
    J.out[e] = si == so ? J.f[(p * J.sd + si) * J.rest + r] : 0.0;
  }
}

int launch_pairs(qtt_ctx* ctx, const std::vector<PairJob>& jobs, bool add) {
  if (jobs.empty()) return QTT_OK;
  if (jobs.size() > 65535) return qtt_fail(ctx, QTT_EINVAL, "too many vertices");
  PairJob* d = (PairJob*)qtt_scratch(ctx, jobs.size() * sizeof(PairJob) + 256);
  if (!d) return qtt_fail(ctx, QTT_ENOMEM, "algebra scratch");
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d, jobs.data(), jobs.size() * sizeof(PairJob), hipMemcpyHostToDevice, ctx->stream), "jobs");
  if (rc) return rc;
  int64_t maxe = 1;
  for (const auto& j : jobs) maxe = std::max(maxe, j.nelem);
  const dim3 grid((unsigned)std::min<int64_t>((maxe + 255) / 256, 4096), (unsigned)jobs.size());
  if (add)
    pair_kernel<true><<<grid, 256, 0, ctx->stream>>>(d);
  else
    pair_kernel<false><<<grid, 256, 0, ctx->stream>>>(d);
  rc = qtt_hip_check(ctx, hipGetLastError(), "pair_kernel");
  if (rc) return rc;
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "algebra sync");  // the job table is host memory
}

}  // namespace

extern "C" {

int qtt_multiply(qtt_ctx* ctx, const qtt_net* f, const qtt_net* g, const qtt_topology* topo, const int32_t* f_of,
                 const int32_t* g_of, qtt_net** out) {
  if (!ctx || !f || !g || !topo || !f_of || !g_of || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  if (f->ctx != ctx || g->ctx != ctx) return qtt_fail(ctx, QTT_EINVAL, "multiply: networks of another context");
  HostTopo T;
  int rc = topo_from_abi(ctx, topo, T);
  if (rc) return rc;
  const int nv = T.nv;
  const qtt_net* ops[2] = {f, g};
  const int32_t* maps[2] = {f_of, g_of};
  std::vector<int> inv[2];
  for (int q = 0; q < 2; q++) {
    const HostTopo& F = ops[q]->t;
    inv[q].assign(F.nv, -1);
    for (int o = 0; o < nv; o++) {
      const int a = maps[q][o];
      if (a < -1 || a >= F.nv) return qtt_fail(ctx, QTT_EINVAL, "multiply: vertex map out of range");
      if (a < 0) continue;
      if (inv[q][a] >= 0) return qtt_fail(ctx, QTT_EINVAL, "multiply: a vertex is mapped twice");
      inv[q][a] = o;
      if (F.sdim[a] == 2 && (T.sdim[o] != 2 || T.vcoord[o] != F.vcoord[a] || T.vdigit[o] != F.vdigit[a]))
        return qtt_fail(ctx, QTT_EINVAL, "multiply: a digit is not at its vertex in the product layout");
    }
    for (int a = 0; a < F.nv; a++) {
      if (inv[q][a] < 0) return qtt_fail(ctx, QTT_EINVAL, "multiply: a vertex is missing from the product layout");
      if (a > 0 && leg_to(T, inv[q][a], inv[q][F.parent[a]]) < 0)
        return qtt_fail(ctx, QTT_EINVAL, "multiply: an edge is missing from the product layout");
    }
  }
  std::vector<int64_t> bond(nv, 1), bq[2];
  for (int q = 0; q < 2; q++) {
    bq[q].assign(nv, 1);
    for (int o = 1; o < nv; o++) bq[q][o] = edge_bond(ops[q], maps[q][o], maps[q][T.parent[o]]);
  }
  for (int o = 1; o < nv; o++) {
    bond[o] = bq[0][o] * bq[1][o];
    if (bond[o] > (1 << 20)) return qtt_fail(ctx, QTT_EINVAL, "multiply: product bond out of range");
  }
  qtt_net* n = nullptr;
  if ((rc = alloc_net_bonds(ctx, T, bond, &n))) return rc;
  std::vector<PairJob> jobs(nv);
  for (int o = 0; o < nv; o++) {
    PairJob& J = jobs[o];
    std::memset(&J, 0, sizeof(J));
    int dims[kL];
    int64_t st[kL];
    J.nleg = core_legs(T, bond, o, dims, st);
    J.out = n->d_data + n->off[o];
    J.nelem = n->cap[o];
    for (int k = 0; k < J.nleg; k++) J.dim[k] = dims[k];
    for (int q = 0; q < 2; q++) {
      const qtt_net* F = ops[q];
      const int a = maps[q][o];
      int fd[kL];
      int64_t fst[kL];
      if (a >= 0) core_legs(F->t, F->bond, a, fd, fst);
      if (q == 0) J.f = a >= 0 ? F->d_data + F->off[a] : nullptr;
      else J.g = a >= 0 ? F->d_data + F->off[a] : nullptr;
      int64_t* S = q == 0 ? J.fs : J.gs;
      for (int k = 0; k < J.nleg; k++) {
        S[k] = 0;
        if (a < 0) continue;
        if (k == 1) {
          if (F->t.sdim[a] == 2) S[k] = fst[1];
          continue;
        }
        const int nb = out_neighbour(T, o, k);
        const int fl = nb >= 0 ? leg_to(F->t, a, maps[q][nb]) : -1;
        if (fl >= 0) S[k] = fst[fl];
      }
    }
    for (int k = 0; k < J.nleg; k++) {
      if (k == 1) J.split[k] = 0;
      else if (k == 0) J.split[k] = (int)bq[1][o];
      else J.split[k] = (int)bq[1][out_neighbour(T, o, k)];
    }
  }
  rc = launch_pairs(ctx, jobs, false);
  if (rc) {
    qtt_net_destroy(n);
    return rc;
  }
  *out = n;
  return QTT_OK;
}

int qtt_add(qtt_ctx* ctx, const qtt_net* f, const qtt_net* g, qtt_net** out) {
  if (!ctx || !f || !g || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  if (f->ctx != ctx || g->ctx != ctx) return qtt_fail(ctx, QTT_EINVAL, "add: networks of another context");
  const HostTopo& T = f->t;
  const HostTopo& G = g->t;
  if (T.nv != G.nv || T.parent != G.parent || T.sdim != G.sdim || T.vcoord != G.vcoord || T.vdigit != G.vdigit)
    return qtt_fail(ctx, QTT_EINVAL, "add: networks on different layouts");
  const int nv = T.nv;
  std::vector<int64_t> bond(nv, 1);
  for (int o = 1; o < nv; o++) {
    bond[o] = f->bond[o] + g->bond[o];
    if (bond[o] > (1 << 20)) return qtt_fail(ctx, QTT_EINVAL, "add: bond out of range");
  }
  qtt_net* n = nullptr;
  int rc = alloc_net_bonds(ctx, T, bond, &n);
  if (rc) return rc;
  std::vector<PairJob> jobs(nv);
  for (int o = 0; o < nv; o++) {
    PairJob& J = jobs[o];
    std::memset(&J, 0, sizeof(J));
    int dims[kL], fd[kL], gd[kL];
    int64_t st[kL];
    J.nleg = core_legs(T, bond, o, dims, st);
    core_legs(T, f->bond, o, fd, J.fs);
    core_legs(G, g->bond, o, gd, J.gs);
    J.f = f->d_data + f->off[o];
    J.g = g->d_data + g->off[o];
    J.out = n->d_data + n->off[o];
    J.nelem = n->cap[o];
    for (int k = 0; k < J.nleg; k++) {
      J.dim[k] = dims[k];
      J.split[k] = (k == 1 || (k == 0 && o == 0)) ? 0 : fd[k];  // site and the root's unit bond: shared
    }
  }
  rc = launch_pairs(ctx, jobs, true);
  if (rc) {
    qtt_net_destroy(n);
    return rc;
  }
  *out = n;
  return QTT_OK;
}

int qtt_partial_integrate(qtt_ctx* ctx, const qtt_net* net, const int32_t* integrate, double weight,
                          const qtt_topology* topo, const int32_t* out_vertex, qtt_net** out) {
  if (!ctx || !net || !integrate || !topo || !out_vertex || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  if (net->ctx != ctx) return qtt_fail(ctx, QTT_EINVAL, "partial_integrate: network of another context");
  const HostTopo& F = net->t;
  HostTopo T;
  int rc = topo_from_abi(ctx, topo, T);
  if (rc) return rc;
  const int nf = F.nv, nt = T.nv;
  // a vertex stays when its digit is not integrated; the others merge away
  auto summed = [&](int a) { return F.vcoord[a] >= 0 && integrate[F.vcoord[a]] != 0; };
  std::vector<int> pos(nf, -1);
  for (int o = 0; o < nt; o++) {
    const int a = out_vertex[o];
    if (a < 0 || a >= nf || pos[a] >= 0) return qtt_fail(ctx, QTT_EINVAL, "partial_integrate: bad vertex map");
    pos[a] = o;
    const bool keep_digit = F.vcoord[a] >= 0 && !summed(a);
    if (keep_digit ? (T.sdim[o] != 2 || T.vcoord[o] != F.vcoord[a] || T.vdigit[o] != F.vdigit[a]) : T.sdim[o] != 1)
      return qtt_fail(ctx, QTT_EINVAL, "partial_integrate: the output layout's digits do not match");
    if (o > 0 && leg_to(F, a, out_vertex[T.parent[o]]) < 0)
      return qtt_fail(ctx, QTT_EINVAL, "partial_integrate: an output edge is not an edge of the network");
  }
  for (int a = 0; a < nf; a++)
    if (pos[a] < 0 && F.vcoord[a] >= 0 && !summed(a))
      return qtt_fail(ctx, QTT_EINVAL, "partial_integrate: a vertex with a remaining digit is missing");
  // The output layout is a tree on the kept vertices whose edges are the
  // network's, so the kept vertices are connected and every merged component
  // hangs off exactly one kept vertex: below it (a subtree, contracted into a
  // vector bottom-up) or above it (the rest of the tree above the topmost
  // kept vertex, contracted top-down along the path from the root).
  std::vector<char> sub_merged(nf, 1);  // the whole subtree of a is merged
  for (int a = nf - 1; a >= 0; --a) {
    if (pos[a] >= 0) sub_merged[a] = 0;
    if (a > 0 && !sub_merged[a]) sub_merged[F.parent[a]] = 0;
  }
  int top = -1;  // the kept vertex nearest the root
  for (int a = 0; a < nf && top < 0; a++)
    if (pos[a] >= 0) top = a;
  std::vector<int> path;  // root ... parent of top, all merged
  for (int a = F.parent[top]; a >= 0; a = F.parent[a]) path.push_back(a);
  std::reverse(path.begin(), path.end());
  // scratch: jobs | environment vectors (bottom-up per vertex, top-down per path vertex + top) | site weights
  std::vector<int64_t> voff(nf, -1), uoff(nf, -1);
  int64_t vtot = 0;
  for (int a = 0; a < nf; a++)
    if (sub_merged[a]) {
      voff[a] = vtot;
      vtot += net->bond[a];
    }
  for (int a : path) {
    uoff[a] = vtot;
    vtot += net->bond[a];
  }
  if (!path.empty()) {
    uoff[top] = vtot;
    vtot += net->bond[top];
  }
  std::vector<int64_t> bond(nt, 1);
  for (int o = 1; o < nt; o++) bond[o] = edge_bond(net, out_vertex[o], out_vertex[T.parent[o]]);
  qtt_net* n = nullptr;
  if ((rc = alloc_net_bonds(ctx, T, bond, &n))) return rc;
  const size_t njob_max = (size_t)nf + path.size() + 1 + nt;
  const size_t jb = (njob_max * sizeof(LegJob) + 255) & ~size_t(255);
  char* scr = (char*)qtt_scratch(ctx, jb + (vtot + 4) * sizeof(double) + 256);
  if (!scr) {
    qtt_net_destroy(n);
    return qtt_fail(ctx, QTT_ENOMEM, "partial_integrate scratch");
  }
  LegJob* d_jobs = (LegJob*)scr;
  double* d_w = (double*)(scr + jb);  // [weight, weight, 1, 1]
  double* d_vec = d_w + 4;
  std::vector<LegJob> env, fin;
  bool plan_ok = true;
  auto core_of = [&](int a) { return net->d_data + net->off[a]; };
  // contract the site of a (summed digit: weights; no digit: its unit leg)
  auto add_site = [&](LegJob& J, int a, const int64_t* st) {
    if (F.sdim[a] == 1) return;
    J.cdim[J.ncon] = 2;
    J.cs[J.ncon] = st[1];
    J.cv[J.ncon] = d_w;
    J.ncon++;
  };
  auto finish = [&](LegJob& J) {
    J.nelem = 1;
    for (int k = 0; k < J.nout; k++) J.nelem *= J.odim[k];
    J.ncomb = 1;
    for (int k = 0; k < J.ncon; k++) J.ncomb *= J.cdim[k];
  };
  // bottom-up vectors of merged subtrees (children before parents)
  for (int a = nf - 1; a >= 0; --a) {
    if (!sub_merged[a]) continue;
    LegJob J;
    std::memset(&J, 0, sizeof(J));
    int dims[kL];
    int64_t st[kL];
    core_legs(F, net->bond, a, dims, st);
    J.src = core_of(a);
    J.dst = d_vec + voff[a];
    J.nout = 1;
    J.odim[0] = dims[0];
    J.os[0] = st[0];
    add_site(J, a, st);
    for (int k = 0; k < F.nch[a]; k++) {
      const int c = F.chs[F.choff[a] + k];
      J.cdim[J.ncon] = dims[2 + k];
      J.cs[J.ncon] = st[2 + k];
      if (voff[c] < 0) plan_ok = false;
      J.cv[J.ncon] = d_vec + std::max<int64_t>(voff[c], 0);
      J.ncon++;
    }
    finish(J);
    env.push_back(J);
  }
  // top-down vectors along the path from the root to the topmost kept vertex
  for (size_t i = 0; i < path.size(); i++) {
    const int a = path[i];
    const int next = i + 1 < path.size() ? path[i + 1] : top;
    LegJob J;
    std::memset(&J, 0, sizeof(J));
    int dims[kL];
    int64_t st[kL];
    core_legs(F, net->bond, a, dims, st);
    J.src = core_of(a);
    J.dst = d_vec + uoff[next];
    const int nl = leg_to(F, a, next);
    J.nout = 1;
    J.odim[0] = dims[nl];
    J.os[0] = st[nl];
    if (i > 0) {  // the root's parent leg is a unit leg
      J.cdim[J.ncon] = dims[0];
      J.cs[J.ncon] = st[0];
      J.cv[J.ncon] = d_vec + uoff[a];
      J.ncon++;
    }
    add_site(J, a, st);
    for (int k = 0; k < F.nch[a]; k++) {
      const int c = F.chs[F.choff[a] + k];
      if (c == next) continue;
      J.cdim[J.ncon] = dims[2 + k];
      J.cs[J.ncon] = st[2 + k];
      if (voff[c] < 0) plan_ok = false;
      J.cv[J.ncon] = d_vec + std::max<int64_t>(voff[c], 0);
      J.ncon++;
    }
    finish(J);
    env.push_back(J);
  }
  // the kept cores in the output layout, merged neighbours contracted
  for (int o = 0; o < nt; o++) {
    const int a = out_vertex[o];
    LegJob J;
    std::memset(&J, 0, sizeof(J));
    int dims[kL], od[kL];
    int64_t st[kL], ost[kL];
    const int r = core_legs(F, net->bond, a, dims, st);
    const int ro = core_legs(T, bond, o, od, ost);
    std::vector<char> used(r, 0);
    J.src = core_of(a);
    J.dst = n->d_data + n->off[o];
    J.nout = ro;
    for (int k = 0; k < ro; k++) {
      J.odim[k] = od[k];
      int fl = -1;
      if (k == 1) fl = T.sdim[o] == 2 ? 1 : -1;
      else {
        const int nb = out_neighbour(T, o, k);
        if (nb >= 0) fl = leg_to(F, a, out_vertex[nb]);
      }
      J.os[k] = fl >= 0 ? st[fl] : 0;
      if (fl >= 0) used[fl] = 1;
    }
    for (int k = 0; k < r; k++) {
      if (used[k]) continue;
      if (k == 1) {
        add_site(J, a, st);
        continue;
      }
      // the root's unit parent leg stays at index 0; an edge to a merged
      // neighbour is contracted with its vector even when its bond is 1 (the
      // one entry is that part's integral)
      if (k == 0 && F.parent[a] < 0) continue;
      const int nb = k == 0 ? F.parent[a] : F.chs[F.choff[a] + (k - 2)];
      const int64_t vo = k == 0 ? uoff[a] : voff[nb];
      if (vo < 0) plan_ok = false;  // cannot happen for a connected output layout
      J.cdim[J.ncon] = dims[k];
      J.cs[J.ncon] = st[k];
      J.cv[J.ncon] = d_vec + std::max<int64_t>(vo, 0);
      J.ncon++;
    }
    finish(J);
    fin.push_back(J);
  }
  if (!plan_ok) rc = qtt_fail(ctx, QTT_EINVAL, "partial_integrate: a merged part touches several kept vertices");
  if (fin.size() > 65535) rc = qtt_fail(ctx, QTT_EINVAL, "too many vertices");
  std::vector<LegJob> all(env);
  all.insert(all.end(), fin.begin(), fin.end());
  const double hw[4] = {weight, weight, 1.0, 1.0};
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d_w, hw, sizeof(hw), hipMemcpyHostToDevice, ctx->stream), "weights");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d_jobs, all.data(), all.size() * sizeof(LegJob), hipMemcpyHostToDevice, ctx->stream), "jobs");
  if (!rc && !env.empty()) {
    leg_env_kernel<<<1, 512, 0, ctx->stream>>>(d_jobs, (int)env.size());
    rc = qtt_hip_check(ctx, hipGetLastError(), "leg_env_kernel");
  }
  if (!rc) {
    int64_t maxe = 1;
    for (const auto& J : fin) maxe = std::max(maxe, J.nelem);
    const dim3 grid((unsigned)std::min<int64_t>((maxe + 255) / 256, 4096), (unsigned)fin.size());
    leg_out_kernel<<<grid, 256, 0, ctx->stream>>>(d_jobs + env.size());
    rc = qtt_hip_check(ctx, hipGetLastError(), "leg_out_kernel");
  }
  if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "partial_integrate sync");
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    qtt_net_destroy(n);
    return rc;
  }
  *out = n;
  return QTT_OK;
}

int qtt_operator_proj(qtt_ctx* ctx, const qtt_net* f, qtt_mpo** out) {
  if (!ctx || !f || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  if (f->ctx != ctx) return qtt_fail(ctx, QTT_EINVAL, "operator_proj: network of another context");
  const HostTopo& T = f->t;
  qtt_mpo* m = nullptr;
  int rc = alloc_mpo(ctx, T, f->bond, &m);
  if (rc) return rc;
  std::vector<ProjJob> jobs(T.nv);
  int64_t maxe = 1;
  for (int v = 0; v < T.nv; v++) {
    int64_t rest = 1;
    for (int k = 0; k < T.nch[v]; k++) rest *= f->bond[T.chs[T.choff[v] + k]];
    jobs[v] = {f->d_data + f->off[v], m->d_data + m->off[v], T.sdim[v], rest, mpo_core_size(m, v)};
    maxe = std::max(maxe, jobs[v].nelem);
  }
  if (jobs.size() > 65535) rc = qtt_fail(ctx, QTT_EINVAL, "too many vertices");
  ProjJob* d = rc ? nullptr : (ProjJob*)qtt_scratch(ctx, jobs.size() * sizeof(ProjJob) + 256);
  if (!rc && !d) rc = qtt_fail(ctx, QTT_ENOMEM, "algebra scratch");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d, jobs.data(), jobs.size() * sizeof(ProjJob), hipMemcpyHostToDevice, ctx->stream), "jobs");
  if (!rc) {
    const dim3 grid((unsigned)std::min<int64_t>((maxe + 255) / 256, 4096), (unsigned)jobs.size());
    proj_kernel<<<grid, 256, 0, ctx->stream>>>(d);
    rc = qtt_hip_check(ctx, hipGetLastError(), "proj_kernel");
  }
  if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "operator_proj sync");
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    qtt_mpo_destroy(m);
    return rc;
  }
  *out = m;
  return QTT_OK;
}

}  // extern "C"
