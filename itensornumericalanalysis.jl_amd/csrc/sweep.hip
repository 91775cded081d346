// Integration and inner products as environment sweeps over the tree.
//
// integrate(f) (src/integration.jl:6-17) contracts every core with a uniform
// weight on its site: vec_v[p] = w * sum_s sum_children T[p][s][c...] prod vec_c.
// The fold over s is fused into the contraction, so each core is read once.
// inner(a, b) (src/integration.jl:20-32 for the <Sigma|op|f> form) builds
// E_v[pa][pb] = sum A[pa][s][ca..] B[pb][s][cb..] prod E_c[ca][cb] by mode
// products (one child at a time) followed by a row-by-row dot product.
#include "qtt_internal.h"

namespace {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Post-order contraction of all cores with the site weight, one workgroup.
// Wave per output row p, lanes over the flattened child index j.
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
      if (lane == 0) vec[voff[v] + pp] = weight * acc;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = vec[voff[0]];
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

int launch_integrate(qtt_ctx* ctx, const qtt_net* net, double weight, double* out_host) {
  const HostTopo& t = net->t;
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

int launch_integrate_batch(qtt_ctx* ctx, int n, const qtt_net* const* nets, double weight, double* out_host) {
  if (n <= 0) return QTT_OK;
  std::vector<int64_t> voff;
  std::vector<int64_t> vbase(n), obase(n), vtot(n);
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
    vtot[i] = tot;
    vall += tot;
  }
  const size_t args_bytes = (size_t)n * sizeof(IntegrateArgs);
  const size_t bytes = (n + vall + 8) * sizeof(double) + oall * sizeof(int64_t) + args_bytes + 256;
  double* scr = qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "integrate scratch");
  double* d_out = scr;
  double* vec = scr + n + 8;
  int64_t* d_voff = (int64_t*)(vec + vall);
  IntegrateArgs* d_args = (IntegrateArgs*)(d_voff + oall);
  std::vector<IntegrateArgs> h(n);
  for (int i = 0; i < n; i++) {
    h[i].data = nets[i]->d_data;
    h[i].off = nets[i]->d_off;
    h[i].bond = nets[i]->d_bond;
    h[i].t = nets[i]->view();
    h[i].voff = d_voff + obase[i];
    h[i].vec = vec + vbase[i];
    h[i].out = d_out + i;
  }
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_voff, voff.data(), oall * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "voff");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d_args, h.data(), args_bytes, hipMemcpyHostToDevice, ctx->stream), "args");
  if (rc) return rc;
  integrate_batch_kernel<<<n, 512, 0, ctx->stream>>>(d_args, weight);
  rc = qtt_hip_check(ctx, hipGetLastError(), "integrate_batch_kernel");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(out_host, d_out, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "integrals");
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
