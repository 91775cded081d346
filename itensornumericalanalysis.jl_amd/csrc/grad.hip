// Adjoint of batched evaluation (no reference counterpart; the rrule of
// evaluate): grad_k[l][s][r] = sum_m ybar_m [s == d_k(x_m)] L_k(x_m)[l] R_k(x_m)[r]
// where L_k is the product of the slices left of site k and R_k the product
// right of it.
//
// Chain kernel: one lane per point, tiles of 128 points per workgroup.
//   1. backward pass (last site to first) computes every R_k and parks it in
//      a per-workgroup global buffer (the environments are recomputed per
//      tile, never stored for the whole batch);
//   2. forward pass rebuilds L_k site by site; the tile's outer products are
//      reduced over its 128 points in LDS and added to the workgroup's own
//      partial gradient (owner-computes, no atomics);
//   3. a second kernel sums the partials in a fixed order, so the result is
//      bitwise reproducible run to run.
#include "qtt_internal.h"

namespace {

constexpr int GB = 128;  // points per tile = threads per workgroup

__device__ inline uint64_t digit_word_g(double x, int L, bool& bad) {
  if (!(x >= 0.0)) {
    bad = true;
    return 0;
  }
  if (x >= 1.0) return (uint64_t(1) << L) - 1;
  return (uint64_t)(x * (double)(uint64_t(1) << L));
}

__global__ void __launch_bounds__(GB)
grad_chain_kernel(const double* __restrict__ pack, const int64_t* __restrict__ poff,
                  const int* __restrict__ bond, TopoView t, const int64_t* __restrict__ goff,
                  const int64_t* __restrict__ roff, int64_t rtot, const double* __restrict__ xs,
                  const double* __restrict__ ybar, int64_t npts, int64_t gsize, double* Rscr,
                  double* gpart, int* __restrict__ flag, int vstride) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x;
  // LDS: Lv[vstride][GB], Nv[vstride][GB], Rv[vstride][GB], digit per lane
  double* Lv = lds;
  double* Nv = lds + (size_t)vstride * GB;
  double* Rv = lds + (size_t)2 * vstride * GB;
  int* dig = (int*)(lds + (size_t)3 * vstride * GB);
  double* R = Rscr + (size_t)blockIdx.x * rtot * GB;  // [roff[k] + r][lane]
  double* G = gpart + (size_t)blockIdx.x * gsize;
  for (int64_t e = tid; e < gsize; e += GB) G[e] = 0.0;
  const int nc = t.ncoord;
  const int64_t ntiles = (npts + GB - 1) / GB;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t p = tile * GB + tid;
    const bool active = p < npts;
    bool bad = false;
    uint64_t kw[QTT_MAX_COORD] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < QTT_MAX_COORD; ++c)
      if (c < nc) kw[c] = digit_word_g(active ? xs[p * nc + c] : 0.0, t.ndigits[c], bad);
    if (bad) flag[0] = 1;
    const double yb = (active && !bad) ? ybar[p] : 0.0;
    // ---- backward: R_in(k) = vector over the child bond of site k -------
    double* v = Lv + tid;
    double* w = Nv + tid;
    v[0] = 1.0;
    for (int k = t.nv - 1; k >= 0; --k) {
      const int chil = bond[k];
      const int chir = (k == t.nv - 1) ? 1 : bond[k + 1];
      for (int r = 0; r < chir; ++r) R[(roff[k] + r) * GB + tid] = v[(size_t)r * GB];
      const int c = t.vcoord[k];
      int d = 0;
      if (c >= 0) {
        const uint64_t kk = c == 0 ? kw[0] : c == 1 ? kw[1] : c == 2 ? kw[2] : kw[3];
        d = (int)((kk >> (t.ndigits[c] - t.vdigit[k])) & 1u);
      }
      const double* __restrict__ P = pack + poff[k];
      const int lpad = (chil + 3) & ~3;
      for (int l0 = 0; l0 < chil; l0 += 4) {
        double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
        const double* __restrict__ q = P + 2 * l0;
        for (int r = 0; r < chir; ++r) {
          const double vr = v[(size_t)r * GB];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            a0[i] = __builtin_fma(q[2 * i], vr, a0[i]);
            a1[i] = __builtin_fma(q[2 * i + 1], vr, a1[i]);
          }
          q += 2 * lpad;
        }
        const uint64_t dmask = 0 - (uint64_t)d;
        double sel[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint64_t ua = (uint64_t)__double_as_longlong(a0[i]);
          const uint64_t ub = (uint64_t)__double_as_longlong(a1[i]);
          sel[i] = __longlong_as_double((long long)((ua & ~dmask) | (ub & dmask)));
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (l0 + i < chil) w[(size_t)(l0 + i) * GB] = sel[i];
      }
      double* tmp = v;
      v = w;
      w = tmp;
    }
    // ---- forward: accumulate outer products, advance L ------------------
    Lv[tid] = 1.0;  // L_0 over the root's parent bond (dimension 1)
    for (int k = 0; k < t.nv; ++k) {
      const int chil = bond[k];
      const int chir = (k == t.nv - 1) ? 1 : bond[k + 1];
      const int c = t.vcoord[k];
      int d = 0;
      if (c >= 0) {
        const uint64_t kk = c == 0 ? kw[0] : c == 1 ? kw[1] : c == 2 ? kw[2] : kw[3];
        d = (int)((kk >> (t.ndigits[c] - t.vdigit[k])) & 1u);
      }
      for (int r = 0; r < chir; ++r) Rv[(size_t)r * GB + tid] = yb * R[(roff[k] + r) * GB + tid];
      dig[tid] = d;
      __syncthreads();
      // owner-computes: entry (l, s, r) of the site's gradient
      const int sd = t.sdim[k];
      const int64_t ne = (int64_t)chil * sd * chir;
      double* Gk = G + goff[k];
      for (int64_t e = tid; e < ne; e += GB) {
        const int r = (int)(e % chir);
        const int64_t ls = e / chir;
        const int s = (int)(ls % sd);
        const int l = (int)(ls / sd);
        double acc = 0.0;
        const double* lrow = Lv + (size_t)l * GB;
        const double* rrow = Rv + (size_t)r * GB;
        for (int j = 0; j < GB; ++j) acc += (dig[j] == s) ? lrow[j] * rrow[j] : 0.0;
        Gk[e] += acc;
      }
      // L_{k+1}[r] = sum_l L_k[l] T[l][d][r]
      const double* __restrict__ P = pack + poff[k];
      const int lpad = (chil + 3) & ~3;
      for (int r = 0; r < chir; ++r) {
        double s0 = 0.0, s1 = 0.0;
        const double* __restrict__ q = P + (size_t)r * 2 * lpad;
        for (int l = 0; l < chil; ++l) {
          const double lv = Lv[(size_t)l * GB + tid];
          s0 = __builtin_fma(q[2 * l], lv, s0);
          s1 = __builtin_fma(q[2 * l + 1], lv, s1);
        }
        Nv[(size_t)r * GB + tid] = d ? s1 : s0;
      }
      __syncthreads();
      for (int r = 0; r < chir; ++r) Lv[(size_t)r * GB + tid] = Nv[(size_t)r * GB + tid];
      __syncthreads();
    }
  }
}

__global__ void grad_reduce_kernel(const double* __restrict__ gpart, int nparts, int64_t gsize,
                                   double* __restrict__ grad) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < gsize;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nparts; ++i) s += gpart[(size_t)i * gsize + e];
    grad[e] = s;
  }
}

}  // namespace

// implemented in eval.hip
int qtt_ensure_pack(qtt_ctx* ctx, qtt_net* net);

int launch_grad(qtt_ctx* ctx, const qtt_net* cnet, const double* xs, const double* ybar,
                int64_t npts, double* grad, int precision) {
  qtt_net* net = const_cast<qtt_net*>(cnet);
  const HostTopo& t = net->t;
  if (!t.chain || t.ncoord > QTT_MAX_COORD)
    return qtt_fail(ctx, QTT_EINVAL, "gradient: implemented for chain networks");
  if (precision != QTT_F64)
    return qtt_fail(ctx, QTT_EINVAL, "gradient: float64 only");
  int rc = qtt_ensure_pack(ctx, net);
  if (rc) return rc;
  const int nv = t.nv;
  std::vector<int64_t> goff(nv), roff(nv);
  int64_t gsize = 0, rtot = 0, maxb = 1;
  for (int k = 0; k < nv; k++) {
    goff[k] = gsize;
    gsize += net->core_size(k);
    roff[k] = rtot;
    rtot += (k == nv - 1) ? 1 : net->bond[k + 1];
    maxb = std::max(maxb, net->bond[k]);
  }
  const size_t lds = (size_t)3 * maxb * GB * sizeof(double) + GB * sizeof(int);
  if (lds > 160 * 1024) return qtt_fail(ctx, QTT_EINVAL, "gradient: bond dimension too large");
  const int64_t ntiles = (npts + GB - 1) / GB;
  int nwg = (int)std::max<int64_t>(1, std::min<int64_t>(ntiles, 2 * ctx->cus));
  size_t bytes = ((size_t)nwg * (rtot * GB + gsize) + 2 * nv + 16) * sizeof(double);
  double* scr = qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "gradient scratch");
  double* Rscr = scr;
  double* gpart = Rscr + (size_t)nwg * rtot * GB;
  int64_t* d_goff = (int64_t*)(gpart + (size_t)nwg * gsize);
  int64_t* d_roff = d_goff + nv;
  rc = qtt_hip_check(ctx, hipMemcpyAsync(d_goff, goff.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "goff");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d_roff, roff.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "roff");
  if (rc) return rc;
  if (lds > 64 * 1024) {
    rc = qtt_hip_check(ctx, hipFuncSetAttribute((const void*)grad_chain_kernel,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                       "grad LDS attribute");
    if (rc) return rc;
  }
  if (npts > 0) {
    grad_chain_kernel<<<nwg, GB, lds, ctx->stream>>>(net->d_pack, net->d_pack_off, net->d_bond, net->view(), d_goff,
                                                     d_roff, rtot, xs, ybar, npts, gsize, Rscr, gpart, ctx->d_flag,
                                                     (int)maxb);
    rc = qtt_hip_check(ctx, hipGetLastError(), "grad_chain_kernel");
    if (rc) return rc;
  } else {
    nwg = 1;
    rc = qtt_hip_check(ctx, hipMemsetAsync(gpart, 0, gsize * sizeof(double), ctx->stream), "zero");
    if (rc) return rc;
  }
  int64_t nb = std::min<int64_t>((gsize + 255) / 256, 4096);
  grad_reduce_kernel<<<(unsigned)std::max<int64_t>(nb, 1), 256, 0, ctx->stream>>>(gpart, nwg, gsize, grad);
  rc = qtt_hip_check(ctx, hipGetLastError(), "grad_reduce_kernel");
  if (rc) return rc;
  // the host offset vectors must outlive their asynchronous copies
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "grad sync");
}
