// Integration and inner products as environment sweeps over the tree.
//
// integrate(f) (src/integration.jl:6-17) contracts every core with a uniform
// weight on its site: vec_v[p] = w * sum_s sum_children T[p][s][c...] prod vec_c.
// The fold over s is fused into the contraction, so each core is read once.
// inner(a, b) (src/integration.jl:20-32 for the <Sigma|op|f> form) builds
// E_v[pa][pb] = sum A[pa][s][ca..] B[pb][s][cb..] prod E_c[ca][cb] by mode
// products (one child at a time) followed by a row-by-row dot product.
#include "qtt_internal.h"

#include <algorithm>
#include <cstring>

namespace {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Post-order contraction of all cores with the site weight, one workgroup.
// Wave per output row p, lanes over the flattened child index j.  The weight
// belongs to a site index, so a vertex without a digit (sdim 1) takes none.
__device__ void integrate_body(const double* __restrict__ data, const int64_t* __restrict__ off,
                               const int* __restrict__ bond, const TopoView& t, double weight,
                               const int64_t* __restrict__ voff, double* vec, double* out);

struct IntegrateArgs {
  const double* data;
  const int64_t* off;
  const int* bond;
  TopoView t;
  const int64_t* voff;
  double* vec;
  double* out;
};

__global__ void __launch_bounds__(512) integrate_batch_kernel(const IntegrateArgs* __restrict__ args, double weight) {
  const IntegrateArgs& a = args[blockIdx.x];
  integrate_body(a.data, a.off, a.bond, a.t, weight, a.voff, a.vec, a.out);
}

__global__ void __launch_bounds__(512)
integrate_tree_kernel(const double* __restrict__ data, const int64_t* __restrict__ off,
                      const int* __restrict__ bond, TopoView t, double weight,
                      const int64_t* __restrict__ voff, double* vec, double* out) {
  integrate_body(data, off, bond, t, weight, voff, vec, out);
}

__device__ void integrate_body(const double* __restrict__ data, const int64_t* __restrict__ off,
                               const int* __restrict__ bond, const TopoView& t, double weight,
                               const int64_t* __restrict__ voff, double* vec, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int v = t.nv - 1; v >= 0; --v) {
    const int m = t.nch[v];
    const int* ch = t.chs + t.choff[v];
    int64_t pc = 1;
    for (int k = 0; k < m; ++k) pc *= bond[ch[k]];
    const int sd = t.sdim[v];
    for (int pp = wid; pp < bond[v]; pp += nw) {
      double acc = 0.0;
      const double* base = data + off[v] + (int64_t)pp * sd * pc;
      for (int64_t j = lane; j < pc; j += 64) {
        double f = 0.0;
        for (int s = 0; s < sd; ++s) f += base[(int64_t)s * pc + j];
        int64_t r = j;
        for (int k = m - 1; k >= 0; --k) {
          const int b = bond[ch[k]];
          f *= vec[voff[ch[k]] + (r % b)];
          r /= b;
        }
        acc += f;
      }
      acc = wave_sum(acc);
      if (lane == 0) vec[voff[v] + pp] = (sd > 1 ? weight : 1.0) * acc;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = vec[voff[0]];
}

// ---------------------------------------------------------------------------
// Dataflow integrate over many workgroups (networks of many MB: the 3D comb
// with bond 128 holds 45 MB of cores).
//
// A unit is one row segment of a vertex's core folded over its site:
// F[j] = sum_s A_v[p][s][j] for j in [ci*FLOW_CL, (ci+1)*FLOW_CL) of the
// flattened child index.  Its partial sum sum_j F[j] prod_k vec_k[c_k(j)] goes
// to part[poff_v + p*nchunk_v + ci]; vec_v[p] is the sum of row p's partials
// times the site weight.  A task is up to FLOW_UPT consecutive units of one
// vertex, run by one workgroup of four wavefronts (two units each).
//
// Every task first loads and folds its core segment into registers, then
// waits until the child vertices' unit counters are complete, so the HBM
// stream of all cores overlaps the dependency chain from the leaves to the
// root.  Tasks are handed out by an atomic ticket in the order of the task
// table, where every vertex comes after its children; a task therefore only
// waits on tasks that were handed out earlier to running workgroups, and the
// waits always end.  A wait is also bounded by FLOW_SPIN_LIMIT polls (it sets
// an error flag and carries on), so every wave exits in any case.  The
// workgroup that completes the root zeroes the counters for the next launch.
// ---------------------------------------------------------------------------
// Per-task timestamps for profiling builds (-DQTT_PHASE_TIMING): ticket taken,
// segments loaded, children ready, published; read back with qtt_debug_flow.
#ifdef QTT_PHASE_TIMING
__device__ unsigned long long qtt_flow_t[8192 * 8];
extern "C" int qtt_debug_flow(unsigned long long* out, int n) {
  return (int)hipMemcpyFromSymbol(out, HIP_SYMBOL(qtt_flow_t), (size_t)n * 8 * sizeof(unsigned long long));
}
#define QTT_FLOW_T(task, i) \
  do { if (threadIdx.x == 0 && (task) < 8192) qtt_flow_t[(task) * 8 + (i)] = wall_clock64(); } while (0)
#else
#define QTT_FLOW_T(task, i) do { } while (0)
#endif

constexpr int FLOW_EPL = 16;                  // elements per lane per unit
constexpr int FLOW_CL = 64 * FLOW_EPL;        // unit length (flattened child index)
constexpr int FLOW_UPW = 2;                   // units per wavefront
constexpr int FLOW_UPT = 4 * FLOW_UPW;        // units per task
constexpr int FLOW_SPIN_LIMIT = 1 << 22;      // polls of a child counter (about 0.2 s)

struct FlowTask {
  int v, u0, nu, pad;
};

struct FlowArgs {
  const double* data;
  const int64_t* off;
  const int* bond;
  TopoView t;
  const FlowTask* tasks;
  const int64_t* vinfo;  // per vertex: poff, pc, nchunk, units
  double* part;
  int* sync;  // [0] ticket, [2] wait limit hit, [4 + v] units done of vertex v
  double* out;  // [0] integral, [1] wait-limit flag
  double weight;
  int ntasks;
};

__global__ void __launch_bounds__(256) integrate_flow_kernel(FlowArgs a) {
  extern __shared__ double s_ch[];  // child vectors, one after the other
  __shared__ int s_i[2];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) s_i[0] = __hip_atomic_fetch_add(a.sync, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int ti = __builtin_amdgcn_readfirstlane(s_i[0]);
  if (ti >= a.ntasks) return;  // the grid has exactly ntasks workgroups
  QTT_FLOW_T(ti, 0);
  const FlowTask tk = a.tasks[ti];
  const int v = tk.v;
  const int64_t poff = a.vinfo[4 * v];
  const int pc = (int)a.vinfo[4 * v + 1], nchunk = (int)a.vinfo[4 * v + 2];
  const int sd = a.t.sdim[v];
  const double* core = a.data + a.off[v];
  // 1. this task's core segments, folded over the site, into registers
  double F[FLOW_UPW][FLOW_EPL];
#pragma unroll
  for (int i = 0; i < FLOW_UPW; ++i) {
    const int u = tk.u0 + wid + 4 * i;
    const bool on = u < tk.u0 + tk.nu;
    const int p = on ? u / nchunk : 0, ci = on ? u - p * nchunk : 0;
    const double* row = core + (int64_t)p * sd * pc;
    const int j0 = ci * FLOW_CL + lane;
    if (sd == 2) {
#pragma unroll
      for (int e = 0; e < FLOW_EPL; ++e) {
        const int j = j0 + 64 * e;
        F[i][e] = (on && j < pc) ? row[j] + row[(int64_t)pc + j] : 0.0;
      }
    } else {
#pragma unroll
      for (int e = 0; e < FLOW_EPL; ++e) {
        const int j = j0 + 64 * e;
        double f = 0.0;
        if (on && j < pc)
          for (int s = 0; s < sd; ++s) f += row[(int64_t)s * pc + j];
        F[i][e] = f;
      }
    }
  }
  // 2. wait for the children
  const int m = a.t.nch[v];
  const int* ch = a.t.chs + a.t.choff[v];
#ifdef QTT_PHASE_TIMING
  {  // the loads have landed
    double z = 0.0;
#pragma unroll
    for (int i = 0; i < FLOW_UPW; ++i)
#pragma unroll
      for (int e = 0; e < FLOW_EPL; ++e) z += F[i][e];
    if (z == 12345.678) a.out[1] = z;
    QTT_FLOW_T(ti, 1);
  }
#endif
  if (tid == 0) {
    for (int k = 0; k < m; ++k) {
      const int c = ch[k];
      const int need = (int)a.vinfo[4 * c + 3];
      int spins = 0, nap = 0;
      while (__hip_atomic_load(a.sync + 4 + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        // back off (64 to 1024 cycles between polls) to keep the polling of
        // hundreds of waiting workgroups off the memory system
        if (nap < 4) __builtin_amdgcn_s_sleep(1);
        else if (nap < 12) __builtin_amdgcn_s_sleep(4);
        else __builtin_amdgcn_s_sleep(16);
        ++nap;
        if (++spins > FLOW_SPIN_LIMIT) {
          __hip_atomic_store(a.sync + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  QTT_FLOW_T(ti, 2);
  // 3. the child vectors into LDS: vec_c[x] = w_c * sum of row x's partials
  int lo = 0;
  for (int k = 0; k < m; ++k) {
    const int c = ch[k], bc = a.bond[c];
    const int64_t pof = a.vinfo[4 * c];
    const int nck = (int)a.vinfo[4 * c + 2];
    const double wc = a.t.sdim[c] > 1 ? a.weight : 1.0;
    for (int x = tid; x < bc; x += 256) {
      double s = 0.0;
      for (int q = 0; q < nck; ++q) s += a.part[pof + (int64_t)x * nck + q];
      s_ch[lo + x] = wc * s;
    }
    lo += bc;
  }
  __syncthreads();
  QTT_FLOW_T(ti, 4);
  // 4. contract the folded segments with the child vectors
  const int b1 = m > 0 ? a.bond[ch[m - 1]] : 1;  // fastest child index
  const int b0 = m > 1 ? a.bond[ch[m - 2]] : 1;
  const int lo1 = lo - b1, lo0 = lo1 - b0;
#pragma unroll
  for (int i = 0; i < FLOW_UPW; ++i) {
    const int u = tk.u0 + wid + 4 * i;
    const bool on = u < tk.u0 + tk.nu;
    const int p = on ? u / nchunk : 0, ci = on ? u - p * nchunk : 0;
    (void)p;
    const int j0 = ci * FLOW_CL + lane;
    double acc = 0.0;
    if (on) {
      if (m == 0) {
#pragma unroll
        for (int e = 0; e < FLOW_EPL; ++e) acc += F[i][e];
      } else if (m == 1) {
#pragma unroll
        for (int e = 0; e < FLOW_EPL; ++e) {
          const int j = j0 + 64 * e;
          if (j < pc) acc += F[i][e] * s_ch[lo1 + j];
        }
      } else if (m == 2) {
        int c0 = j0 / b1, c1 = j0 - c0 * b1;
#pragma unroll
        for (int e = 0; e < FLOW_EPL; ++e) {
          const int j = j0 + 64 * e;
          if (j < pc) acc += F[i][e] * s_ch[lo0 + c0] * s_ch[lo1 + c1];
          c1 += 64;  // advance (c0, c1) by 64 without a division
          while (c1 >= b1) {
            c1 -= b1;
            ++c0;
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < FLOW_EPL; ++e) {
          const int j = j0 + 64 * e;
          if (j < pc) {
            double w = 1.0;
            int r = j, o = lo;
            for (int k = m - 1; k >= 0; --k) {
              const int bk = a.bond[ch[k]];
              o -= bk;
              const int q = r / bk;
              w *= s_ch[o + (r - q * bk)];
              r = q;
            }
            acc += F[i][e] * w;
          }
        }
      }
    }
    acc = wave_sum(acc);
    if (on && lane == 0) a.part[poff + u] = acc;
  }
  // 5. publish: the barrier orders every wave's partials before thread 0's
  // release, then the units are counted
  QTT_FLOW_T(ti, 5);
  __syncthreads();
  QTT_FLOW_T(ti, 6);
  if (tid == 0) {
    const int prev = __hip_atomic_fetch_add(a.sync + 4 + v, tk.nu, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_i[1] = v == 0 && prev + tk.nu == (int)a.vinfo[3];
    if (s_i[1]) {  // the root (bond 1) is complete
      double s = 0.0;
      for (int q = 0; q < nchunk; ++q) s += a.part[poff + q];
      a.out[0] = (sd > 1 ? a.weight : 1.0) * s;
      a.out[1] = (double)__hip_atomic_load(a.sync + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    QTT_FLOW_T(ti, 3);
  }
  __syncthreads();
  // Every task counts its units after its own waits, so once the root is
  // complete no workgroup waits any more (every ticket is taken, every wait
  // has ended): zero the counters for the next launch.  (After a wait that hit
  // its limit the host zeroes them instead.)
  if (s_i[1]) {
    for (int x = tid; x < a.t.nv; x += 256) a.sync[4 + x] = 0;
    if (tid == 0) {
      a.sync[0] = 0;
      a.sync[2] = 0;
    }
  }
}

// Two-layer environment sweep for <a|b>, one workgroup.
__global__ void __launch_bounds__(512)
inner_tree_kernel(const double* __restrict__ A, const int64_t* __restrict__ aoff,
                  const int* __restrict__ abond, const double* __restrict__ B,
                  const int64_t* __restrict__ boff, const int* __restrict__ bbond, TopoView t,
                  const int64_t* __restrict__ eoff, double* E, double* Y0, double* Y1, double* out) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = bs >> 6;
  for (int v = t.nv - 1; v >= 0; --v) {
    const int m = t.nch[v];
    const int* ch = t.chs + t.choff[v];
    int64_t dims[QTT_MAX_NDIM];
    const int nd = 2 + m;
    dims[0] = bbond[v];
    dims[1] = t.sdim[v];
    for (int k = 0; k < m; ++k) dims[2 + k] = bbond[ch[k]];
    const double* cur = B + boff[v];
    double* bufs[2] = {Y0, Y1};
    int which = 0;
    for (int k = 0; k < m; ++k) {
      const int ax = 2 + k;
      int64_t Apre = 1, Bpost = 1;
      for (int i = 0; i < ax; ++i) Apre *= dims[i];
      for (int i = ax + 1; i < nd; ++i) Bpost *= dims[i];
      const int64_t nold = dims[ax];
      const int64_t nnew = abond[ch[k]];
      const double* Ec = E + eoff[ch[k]];  // [ca][cb]
      double* nxt = bufs[which];
      const int64_t total = Apre * nnew * Bpost;
      for (int64_t e = tid; e < total; e += bs) {
        const int64_t bb = e % Bpost;
        const int64_t rest = e / Bpost;
        const int64_t j = rest % nnew;
        const int64_t a = rest / nnew;
        double s = 0.0;
        for (int64_t i = 0; i < nold; ++i) s += Ec[j * nold + i] * cur[(a * nold + i) * Bpost + bb];
        nxt[e] = s;
      }
      __syncthreads();
      cur = nxt;
      which ^= 1;
      dims[ax] = nnew;
    }
    // E_v[pa][pb] = sum_rest A[pa][rest] Y[pb][rest]
    int64_t rest = 1;
    for (int i = 1; i < nd; ++i) rest *= dims[i];
    const double* Av = A + aoff[v];
    const int na = abond[v], nb = bbond[v];
    for (int pr = wid; pr < na * nb; pr += nw) {
      const int pa = pr / nb, pb = pr % nb;
      double s = 0.0;
      for (int64_t r = lane; r < rest; r += 64) s += Av[(int64_t)pa * rest + r] * cur[(int64_t)pb * rest + r];
      s = wave_sum(s);
      if (lane == 0) E[eoff[v] + pr] = s;
    }
    __syncthreads();
  }
  if (tid == 0) out[0] = E[eoff[0]];
}

}  // namespace

// Counters of the dataflow kernels: zeroed once when allocated; every launch
// leaves them zero again.
static int* sync_counters(qtt_ctx* ctx, int n) {
  if (n <= ctx->sync_cap) return ctx->d_sync;
  if (ctx->d_sync) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->d_sync);
    ctx->d_sync = nullptr;
    ctx->sync_cap = 0;
  }
  const int cap = std::max(n, 4096);
  if (hipMalloc(&ctx->d_sync, cap * sizeof(int)) != hipSuccess ||
      hipMemsetAsync(ctx->d_sync, 0, cap * sizeof(int), ctx->stream) != hipSuccess) {
    (void)hipGetLastError();
    if (ctx->d_sync) (void)hipFree(ctx->d_sync);
    ctx->d_sync = nullptr;
    return nullptr;
  }
  ctx->sync_cap = cap;
  return ctx->d_sync;
}

// Networks from this many core elements up use the dataflow kernel.
constexpr int64_t FLOW_MIN_ELEMS = int64_t(1) << 15;

static int launch_integrate_flow(qtt_ctx* ctx, const qtt_net* net, double weight, double* out_host, bool* used) {
  *used = false;
  const HostTopo& t = net->t;
  std::vector<int64_t> info(4 * (size_t)t.nv);
  std::vector<FlowTask> tasks;
  int64_t ptot = 0, maxch = 0;
  for (int v = t.nv - 1; v >= 0; --v) {  // children (larger ids) before parents
    int64_t pc = 1, chs = 0;
    for (int k = 0; k < t.nch[v]; k++) {
      const int64_t b = net->bond[t.chs[t.choff[v] + k]];
      pc *= b;
      chs += b;
    }
    const int64_t nchunk = (pc + FLOW_CL - 1) / FLOW_CL;
    const int64_t units = net->bond[v] * nchunk;
    if (pc >= (int64_t(1) << 30) || units >= (int64_t(1) << 30)) return QTT_OK;
    maxch = std::max(maxch, chs);
    info[4 * v] = ptot;
    info[4 * v + 1] = pc;
    info[4 * v + 2] = nchunk;
    info[4 * v + 3] = units;
    ptot += units;
    for (int64_t u0 = 0; u0 < units; u0 += FLOW_UPT)
      tasks.push_back({v, (int)u0, (int)std::min<int64_t>(FLOW_UPT, units - u0), 0});
  }
  if (maxch * sizeof(double) > 64 * 1024 || tasks.size() >= (size_t(1) << 30)) return QTT_OK;
  int* sync = sync_counters(ctx, t.nv + 4);
  if (!sync) return qtt_fail(ctx, QTT_ENOMEM, "integrate counters");
  const size_t tab = info.size() * sizeof(int64_t) + tasks.size() * sizeof(FlowTask);
  double* scr = qtt_scratch(ctx, (2 + ptot) * sizeof(double) + tab + 64);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "integrate scratch");
  *used = true;
  std::vector<char> h(tab);
  std::memcpy(h.data(), info.data(), info.size() * sizeof(int64_t));
  std::memcpy(h.data() + info.size() * sizeof(int64_t), tasks.data(), tasks.size() * sizeof(FlowTask));
  FlowArgs a;
  a.data = net->d_data;
  a.off = net->d_off;
  a.bond = net->d_bond;
  a.t = net->view();
  a.out = scr;
  a.part = scr + 2;
  a.vinfo = (const int64_t*)(a.part + ptot);
  a.tasks = (const FlowTask*)(a.vinfo + info.size());
  a.sync = sync;
  a.weight = weight;
  a.ntasks = (int)tasks.size();
  int rc = qtt_hip_check(ctx, hipMemcpyAsync((void*)a.vinfo, h.data(), tab, hipMemcpyHostToDevice, ctx->stream), "integrate tables");
  if (rc) return rc;
  integrate_flow_kernel<<<a.ntasks, 256, maxch * sizeof(double), ctx->stream>>>(a);
  rc = qtt_hip_check(ctx, hipGetLastError(), "integrate_flow_kernel");
  if (rc) return rc;
  double r[2];
  rc = qtt_hip_check(ctx, hipMemcpyAsync(r, scr, sizeof(r), hipMemcpyDeviceToHost, ctx->stream), "integral");
  if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "integrate sync");
  if (rc) return rc;
  if (r[1] != 0.0) {
    (void)hipMemsetAsync(sync, 0, (t.nv + 4) * sizeof(int), ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    return qtt_fail(ctx, QTT_EHIP, "integrate: a dependency wait exceeded its limit");
  }
  *out_host = r[0];
  return QTT_OK;
}

int launch_integrate(qtt_ctx* ctx, const qtt_net* net, double weight, double* out_host) {
  const HostTopo& t = net->t;
  if (net->total >= FLOW_MIN_ELEMS) {
    bool used = false;
    const int rc = launch_integrate_flow(ctx, net, weight, out_host, &used);
    if (rc || used) return rc;
  }
  std::vector<int64_t> voff(t.nv);
  int64_t tot = 0;
  for (int v = 0; v < t.nv; v++) {
    voff[v] = tot;
    tot += net->bond[v];
  }
  size_t bytes = (tot + 2) * sizeof(double) + t.nv * sizeof(int64_t) + 64;
  double* scr = qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "integrate scratch");
  double* d_out = scr;
  double* vec = scr + 2;
  int64_t* d_voff = (int64_t*)(vec + tot);
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_voff, voff.data(), t.nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "voff");
  if (rc) return rc;
  integrate_tree_kernel<<<1, 512, 0, ctx->stream>>>(net->d_data, net->d_off, net->d_bond, net->view(), weight,
                                                    d_voff, vec, d_out);
  rc = qtt_hip_check(ctx, hipGetLastError(), "integrate_tree_kernel");
  if (rc) return rc;
  rc = qtt_hip_check(ctx, hipMemcpyAsync(out_host, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "integral");
  if (rc) return rc;
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "integrate sync");
}

// Integrals of a batch, one workgroup each, in one launch.  The vectors are
// laid out from the host bonds, which may be upper bounds of the device
// bonds the kernel reads (a pipeline stages this before the truncation that
// lowers them).  Staged like the other batch launches: tables at `scr`,
// results left on the device at st.d_out.
size_t integrate_batch_bytes(int n, const qtt_net* const* nets) {
  if (n <= 0) return 0;
  int64_t vall = 0, oall = 0;
  for (int i = 0; i < n; i++) {
    oall += nets[i]->t.nv;
    for (int v = 0; v < nets[i]->t.nv; v++) vall += nets[i]->bond[v];
  }
  return (n + vall + 8) * sizeof(double) + oall * sizeof(int64_t) + (size_t)n * sizeof(IntegrateArgs) + 256;
}

int integrate_batch_stage(qtt_ctx* ctx, int n, const qtt_net* const* nets, double weight, char* scr,
                          BatchStage& st) {
  st.n = n;
  st.weight = weight;
  if (n <= 0) return QTT_OK;
  std::vector<int64_t> voff;
  std::vector<int64_t> vbase(n), obase(n);
  int64_t vall = 0, oall = 0;
  for (int i = 0; i < n; i++) {
    const qtt_net* net = nets[i];
    obase[i] = oall;
    oall += net->t.nv;
    vbase[i] = vall;
    int64_t tot = 0;
    for (int v = 0; v < net->t.nv; v++) {
      voff.push_back(tot);
      tot += net->bond[v];
    }
    vall += tot;
  }
  const size_t args_bytes = (size_t)n * sizeof(IntegrateArgs);
  double* d_out = (double*)scr;
  double* vec = d_out + n + 8;
  int64_t* d_voff = (int64_t*)(vec + vall);
  IntegrateArgs* d_args = (IntegrateArgs*)(d_voff + oall);
  // host copy: voff, then the argument table, contiguous as on the device
  st.host.assign(oall * sizeof(int64_t) + args_bytes, 0);
  std::memcpy(st.host.data(), voff.data(), oall * sizeof(int64_t));
  IntegrateArgs* h = (IntegrateArgs*)(st.host.data() + oall * sizeof(int64_t));
  for (int i = 0; i < n; i++) {
    h[i].data = nets[i]->d_data;
    h[i].off = nets[i]->d_off;
    h[i].bond = nets[i]->d_bond;
    h[i].t = nets[i]->view();
    h[i].voff = d_voff + obase[i];
    h[i].vec = vec + vbase[i];
    h[i].out = d_out + i;
  }
  st.dev = (char*)d_args;
  st.d_out = d_out;
  return qtt_hip_check(ctx, hipMemcpyAsync(d_voff, st.host.data(), st.host.size(), hipMemcpyHostToDevice, ctx->stream),
                       "integrate args");
}

int integrate_batch_launch(qtt_ctx* ctx, const BatchStage& st) {
  if (st.n <= 0) return QTT_OK;
  integrate_batch_kernel<<<st.n, 512, 0, ctx->stream>>>((const IntegrateArgs*)st.dev, st.weight);
  return qtt_hip_check(ctx, hipGetLastError(), "integrate_batch_kernel");
}

int launch_integrate_batch(qtt_ctx* ctx, int n, const qtt_net* const* nets, double weight, double* out_host) {
  if (n <= 0) return QTT_OK;
  char* scr = (char*)qtt_scratch(ctx, integrate_batch_bytes(n, nets));
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "integrate scratch");
  BatchStage st;
  int rc = integrate_batch_stage(ctx, n, nets, weight, scr, st);
  if (!rc) rc = integrate_batch_launch(ctx, st);
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(out_host, st.d_out, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "integrals");
  if (rc) return rc;
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "integrate batch sync");
}

int launch_inner(qtt_ctx* ctx, const qtt_net* a, const qtt_net* b, double* out_host) {
  const HostTopo& t = a->t;
  std::vector<int64_t> eoff(t.nv);
  int64_t etot = 0, ymax = 1;
  for (int v = 0; v < t.nv; v++) {
    eoff[v] = etot;
    etot += a->bond[v] * b->bond[v];
    // largest intermediate: the B core with any subset of child bonds replaced by A's
    int64_t s = b->bond[v] * t.sdim[v];
    for (int k = 0; k < t.nch[v]; k++) {
      int c = t.chs[t.choff[v] + k];
      s *= std::max(a->bond[c], b->bond[c]);
    }
    ymax = std::max(ymax, s);
  }
  size_t bytes = (2 + etot + 2 * ymax) * sizeof(double) + t.nv * sizeof(int64_t) + 64;
  double* scr = qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "inner scratch");
  double* d_out = scr;
  double* E = scr + 2;
  double* Y0 = E + etot;
  double* Y1 = Y0 + ymax;
  int64_t* d_eoff = (int64_t*)(Y1 + ymax);
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_eoff, eoff.data(), t.nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "eoff");
  if (rc) return rc;
  inner_tree_kernel<<<1, 512, 0, ctx->stream>>>(a->d_data, a->d_off, a->d_bond, b->d_data, b->d_off, b->d_bond,
                                                a->view(), d_eoff, E, Y0, Y1, d_out);
  rc = qtt_hip_check(ctx, hipGetLastError(), "inner_tree_kernel");
  if (rc) return rc;
  rc = qtt_hip_check(ctx, hipMemcpyAsync(out_host, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "inner");
  if (rc) return rc;
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "inner sync");
}
