// Batched point evaluation: f(x) = product of the core slices selected by the
// digits of x (src/itensornetworkfunction.jl:96-106 -- calculate_ind_values,
// project, scalar -- for a whole batch of points in one launch).
//
// Chain kernel (the hot path): one lane per point.  The digit map is fused
// into the prologue as integer bit extraction of floor(x 2^L)
// (src/IndexMaps/abstractindexmap.jl:121-138 for base 2).  The chain is
// reduced from the last site to the first: w[l] = sum_r T[l][d][r] v[r].
// Core entries are wave-uniform, so they are read through the scalar data
// path from a packed copy [r][lpad][s] (8 doubles = one 64-byte scalar load
// per r and 4 rows); each lane forms both digit slices with FMAs and keeps
// the one its digit selects.  The per-lane vectors live in LDS ([i][lane]
// layout, conflict free).
#include "qtt_internal.h"

#include <cstring>

namespace {

__device__ inline uint64_t digit_word(double x, int L, bool& bad) {
  if (!(x >= 0.0)) {  // negative or NaN
    bad = true;
    return 0;
  }
  if (x >= 1.0) return (uint64_t(1) << L) - 1;  // the greedy map saturates at 1
  return (uint64_t)(x * (double)(uint64_t(1) << L));  // exact scaling, truncation = floor
}

// branch-free select of a (mask 0) or b (mask all ones), kept in registers
__device__ inline double bitsel(double a, double b, uint64_t mask) {
  const uint64_t ua = (uint64_t)__double_as_longlong(a), ub = (uint64_t)__double_as_longlong(b);
  return __longlong_as_double((long long)((ua & ~mask) | (ub & mask)));
}

__device__ inline float bitself(float a, float b, uint32_t mask) {
  return __uint_as_float((__float_as_uint(a) & ~mask) | (__float_as_uint(b) & mask));
}

// pack chain site k as P[r][lpad][2] (lpad = bond[k] rounded up to 4, zero padded)
__global__ void pack_chain_kernel(const double* __restrict__ data, const int64_t* __restrict__ off,
                                  const int* __restrict__ bond, const int* __restrict__ sdim,
                                  const int64_t* __restrict__ poff, int nv,
                                  double* __restrict__ pack) {
  const int k = blockIdx.y;
  const int chil = bond[k];
  const int chir = (k == nv - 1) ? 1 : bond[k + 1];
  const int lpad = (chil + 3) & ~3;
  const int rpad = (chir + 3) & ~3;
  const int sd = sdim[k];
  const int64_t n = (int64_t)rpad * lpad * 2;
  const double* T = data + off[k];
  double* P = pack + poff[k];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int s = (int)(e & 1);
    int64_t rest = e >> 1;
    int l = (int)(rest % lpad);
    int r = (int)(rest / lpad);
    double val = 0.0;
    if (l < chil && r < chir) {
      int ss = sd == 1 ? 0 : s;
      val = T[((int64_t)l * sd + ss) * chir + r];
    }
    P[e] = val;
  }
}

__global__ void __launch_bounds__(256)
eval_chain_f64_kernel(const double* __restrict__ pack, const int64_t* __restrict__ poff,
                      const int* __restrict__ bond, TopoView t, const double* __restrict__ xs,
                      int64_t npts, double* __restrict__ out, int* __restrict__ flag, int vstride) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x;
  const int bs = blockDim.x;
  const int64_t p = (int64_t)blockIdx.x * bs + tid;
  const bool active = p < npts;
  const int nc = t.ncoord;
  bool bad = false;
  uint64_t kw[QTT_MAX_COORD] = {0, 0, 0, 0};
#pragma unroll
  for (int c = 0; c < QTT_MAX_COORD; ++c)
    if (c < nc) kw[c] = digit_word(active ? xs[p * nc + c] : 0.0, t.ndigits[c], bad);
  double* v = lds + tid;
  double* w = lds + (size_t)vstride * bs + tid;
  v[0] = 1.0;
  for (int k = t.nv - 1; k >= 0; --k) {
    const int chil = bond[k];
    const int chir = (k == t.nv - 1) ? 1 : bond[k + 1];
    const int c = t.vcoord[k];
    int d = 0;
    if (c >= 0) {
      const uint64_t kk = c == 0 ? kw[0] : c == 1 ? kw[1] : c == 2 ? kw[2] : kw[3];
      d = (int)((kk >> (t.ndigits[c] - t.vdigit[k])) & 1u);
    }
    const uint64_t dmask = 0 - (uint64_t)d;
    const double* __restrict__ P = pack + poff[k];
    const int lpad = (chil + 3) & ~3;
    for (int l0 = 0; l0 < chil; l0 += 4) {
      double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
      const double* __restrict__ q = P + 2 * l0;
      for (int r = 0; r < chir; ++r) {
        const double vr = v[(size_t)r * bs];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a0[i] = __builtin_fma(q[2 * i], vr, a0[i]);
          a1[i] = __builtin_fma(q[2 * i + 1], vr, a1[i]);
        }
        q += 2 * lpad;
      }
      double sel[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) sel[i] = bitsel(a0[i], a1[i], dmask);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (l0 + i < chil) w[(size_t)(l0 + i) * bs] = sel[i];
    }
    double* tmp = v;
    v = w;
    w = tmp;
  }
  if (active) out[p] = bad ? __builtin_nan("") : v[0];
  if (bad) flag[0] = 1;
}

// Register-resident chain kernel for bonds <= CHI (4, 8, 16 or 32).  Same
// arithmetic as eval_chain_f64_kernel; the per-lane vectors stay in VGPRs
// (fully unrolled, compile-time indices) and rows of the packed core are zero
// padded to a multiple of 4, so four 64-byte scalar loads are issued back to
// back for each group of 4 rows before their 32 FMAs.

// One site of the one-point chain kernel: y = T[:, d, :] x over the child
// bond for this lane's digit d (mask dmask), written in blocks of 4 up to the
// parent bond (zero past the bond inside the last block).
template <int CHI>
__device__ __forceinline__ void chain1_site(const double* __restrict__ P, const SiteDesc sd, uint64_t dmask,
                                            const double (&x)[CHI], double (&y)[CHI]) {
  const int lpad = (sd.chil + 3) & ~3;
#pragma unroll
  for (int l0 = 0; l0 < CHI; l0 += 4) {
    if (l0 < sd.chil) {
      double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
      // rows in pairs, the next pair's scalar loads issued before this pair's
      // multiply-adds (rows up to the bond rounded up to 4 exist in the pack)
      const double* __restrict__ Pl = P + l0 * 2;
      const int rs = lpad * 2;
      double qa[8], qb[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        qa[e] = Pl[e];
        qb[e] = Pl[rs + e];
      }
#pragma unroll
      for (int r = 0; r < CHI; r += 2) {
        if (r != 0 && r >= sd.chir) break;
        const int rn = r + 2 < sd.chir ? r + 2 : r;
        double na[8], nb[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          na[e] = Pl[rn * rs + e];
          nb[e] = Pl[(rn + 1) * rs + e];
        }
        const double u = x[r], w = x[r + 1];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a0[i] = __builtin_fma(qa[2 * i], u, a0[i]);
          a1[i] = __builtin_fma(qa[2 * i + 1], u, a1[i]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a0[i] = __builtin_fma(qb[2 * i], w, a0[i]);
          a1[i] = __builtin_fma(qb[2 * i + 1], w, a1[i]);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          qa[e] = na[e];
          qb[e] = nb[e];
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) y[l0 + i] = bitsel(a0[i], a1[i], dmask);
    }
    // blocks past the bond are left as they are (never read by the next site)
  }
}

// One lane per point; sites two at a time (v -> w, w -> v), so no register
// copy follows a site.
template <int CHI>
__device__ __forceinline__ void eval_chain_reg_body(const double* __restrict__ pack, const int64_t* __restrict__ poff,
                                                    const SiteDesc* __restrict__ desc, int nv, int nc,
                                                    const int* __restrict__ ndigits, const double* __restrict__ xs,
                                                    int64_t npts, double* __restrict__ out, int* __restrict__ flag,
                                                    int64_t p) {
  const bool active = p < npts;
  bool bad = false;
  uint64_t kw[QTT_MAX_COORD] = {0, 0, 0, 0};
#pragma unroll
  for (int c = 0; c < QTT_MAX_COORD; ++c)
    if (c < nc) kw[c] = digit_word(active ? xs[p * nc + c] : 0.0, ndigits[c], bad);
  double v[CHI], w[CHI];
#pragma unroll
  for (int i = 0; i < CHI; ++i) v[i] = 0.0;
  v[0] = 1.0;
  auto mask = [&](const SiteDesc& sd) -> uint64_t {
    if (sd.coord < 0) return 0;
    const uint64_t kk = sd.coord == 0 ? kw[0] : sd.coord == 1 ? kw[1] : sd.coord == 2 ? kw[2] : kw[3];
    return 0 - ((kk >> sd.shift) & 1u);
  };
  int k = nv - 1;
  for (; k >= 1; k -= 2) {
    const SiteDesc sa = desc[k];
    chain1_site<CHI>(pack + poff[k], sa, mask(sa), v, w);
    const SiteDesc sb = desc[k - 1];
    chain1_site<CHI>(pack + poff[k - 1], sb, mask(sb), w, v);
  }
  double r = v[0];
  if (k == 0) {  // an odd number of sites: the last one lands in w
    const SiteDesc sa = desc[0];
    chain1_site<CHI>(pack + poff[0], sa, mask(sa), v, w);
    r = w[0];
  }
  if (active) out[p] = bad ? __builtin_nan("") : r;
  if (bad) flag[0] = 1;
}

template <int CHI>
__global__ void __launch_bounds__(256)
eval_chain_reg_kernel(const double* __restrict__ pack, const int64_t* __restrict__ poff,
                      const SiteDesc* __restrict__ desc, int nv, int nc, const int* __restrict__ ndigits,
                      const double* __restrict__ xs, int64_t npts, double* __restrict__ out,
                      int* __restrict__ flag) {
  eval_chain_reg_body<CHI>(pack, poff, desc, nv, nc, ndigits, xs, npts, out, flag,
                           (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// A batch of chains in one launch: workgroup b evaluates 256 points of the
// item item_of[b] (its first workgroup is items[i].block0).
struct EvalItem {
  const double* pack;
  const int64_t* poff;
  const SiteDesc* desc;
  const int* ndigits;
  const double* xs;
  double* out;
  int64_t npts;
  int64_t block0;
  int nv, nc;
};

template <int CHI>
__global__ void __launch_bounds__(256)
eval_chain_batch_kernel(const EvalItem* __restrict__ items, const int* __restrict__ item_of, int* __restrict__ flag) {
  const EvalItem& it = items[item_of[blockIdx.x]];
  eval_chain_reg_body<CHI>(it.pack, it.poff, it.desc, it.nv, it.nc, it.ndigits, it.xs, it.npts, it.out, flag,
                           (int64_t)(blockIdx.x - it.block0) * blockDim.x + threadIdx.x);
}

// One site of the two-point chain kernel: y = T[:, d, :] x for each point's
// digit d (masks m0, m1), over the child bond.  y is written in blocks of 4
// up to the parent bond; inside the last block the entries past the bond are
// zero (the packed core is zero there).
template <int CHI>
__device__ __forceinline__ void chain2_site(const double* __restrict__ P, const SiteDesc sd, uint64_t m0, uint64_t m1,
                                            const double (&x0)[CHI], const double (&x1)[CHI], double (&y0)[CHI],
                                            double (&y1)[CHI]) {
  const int lpad = (sd.chil + 3) & ~3;
#pragma unroll
  for (int l0 = 0; l0 < CHI; l0 += 4) {
    if (l0 < sd.chil) {
      double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
      double b0[4] = {0.0, 0.0, 0.0, 0.0}, b1[4] = {0.0, 0.0, 0.0, 0.0};
      // rows in pairs, the next pair's scalar loads issued before this pair's
      // multiply-adds (rows up to the bond rounded up to 4 exist in the pack)
      const double* __restrict__ Pl = P + l0 * 2;
      const int rs = lpad * 2;
      double qa[8], qb[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        qa[e] = Pl[e];
        qb[e] = Pl[rs + e];
      }
#pragma unroll
      for (int r = 0; r < CHI; r += 2) {
        if (r != 0 && r >= sd.chir) break;
        const int rn = r + 2 < sd.chir ? r + 2 : r;
        double na[8], nb[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          na[e] = Pl[rn * rs + e];
          nb[e] = Pl[(rn + 1) * rs + e];
        }
        const double u0 = x0[r], u1 = x1[r], w0 = x0[r + 1], w1 = x1[r + 1];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a0[i] = __builtin_fma(qa[2 * i], u0, a0[i]);
          a1[i] = __builtin_fma(qa[2 * i + 1], u0, a1[i]);
          b0[i] = __builtin_fma(qa[2 * i], u1, b0[i]);
          b1[i] = __builtin_fma(qa[2 * i + 1], u1, b1[i]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a0[i] = __builtin_fma(qb[2 * i], w0, a0[i]);
          a1[i] = __builtin_fma(qb[2 * i + 1], w0, a1[i]);
          b0[i] = __builtin_fma(qb[2 * i], w1, b0[i]);
          b1[i] = __builtin_fma(qb[2 * i + 1], w1, b1[i]);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          qa[e] = na[e];
          qb[e] = nb[e];
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        y0[l0 + i] = bitsel(a0[i], a1[i], m0);
        y1[l0 + i] = bitsel(b0[i], b1[i], m1);
      }
    }
    // blocks past the bond are left as they are: the next site reads rows only
    // up to the bond rounded up to 2, which the last computed block covers
  }
}

// Two points per lane (p and p + blockDim.x of the block's span): every core
// entry loaded through the scalar path feeds four FMAs instead of two, which
// halves the scalar loads, the address arithmetic and the loop overhead per
// point.  Same arithmetic per point as eval_chain_reg_kernel.  Sites are taken
// two at a time, v -> w then w -> v, so no register copy follows a site.
template <int CHI>
__global__ void __launch_bounds__(256)
eval_chain_reg2_kernel(const double* __restrict__ pack, const int64_t* __restrict__ poff,
                       const SiteDesc* __restrict__ desc, int nv, int nc, const int* __restrict__ ndigits,
                       const double* __restrict__ xs, int64_t npts, double* __restrict__ out,
                       int* __restrict__ flag) {
  const int64_t p0 = (int64_t)blockIdx.x * (2 * blockDim.x) + threadIdx.x;
  const int64_t p1 = p0 + blockDim.x;
  const bool act0 = p0 < npts, act1 = p1 < npts;
  bool bad0 = false, bad1 = false;
  uint64_t kw0[QTT_MAX_COORD] = {0, 0, 0, 0}, kw1[QTT_MAX_COORD] = {0, 0, 0, 0};
#pragma unroll
  for (int c = 0; c < QTT_MAX_COORD; ++c) {
    if (c < nc) {
      kw0[c] = digit_word(act0 ? xs[p0 * nc + c] : 0.0, ndigits[c], bad0);
      kw1[c] = digit_word(act1 ? xs[p1 * nc + c] : 0.0, ndigits[c], bad1);
    }
  }
  double v0[CHI], v1[CHI], w0[CHI], w1[CHI];
#pragma unroll
  for (int i = 0; i < CHI; ++i) {
    v0[i] = 0.0;
    v1[i] = 0.0;
  }
  v0[0] = 1.0;
  v1[0] = 1.0;
  // the digit masks of site k for both points
  auto masks = [&](const SiteDesc& sd, uint64_t& m0, uint64_t& m1) {
    m0 = 0;
    m1 = 0;
    if (sd.coord >= 0) {
      const int c = sd.coord;
      const uint64_t k0 = c == 0 ? kw0[0] : c == 1 ? kw0[1] : c == 2 ? kw0[2] : kw0[3];
      const uint64_t k1 = c == 0 ? kw1[0] : c == 1 ? kw1[1] : c == 2 ? kw1[2] : kw1[3];
      m0 = 0 - ((k0 >> sd.shift) & 1u);
      m1 = 0 - ((k1 >> sd.shift) & 1u);
    }
  };
  int k = nv - 1;
  for (; k >= 1; k -= 2) {
    uint64_t m0, m1;
    const SiteDesc sa = desc[k];
    masks(sa, m0, m1);
    chain2_site<CHI>(pack + poff[k], sa, m0, m1, v0, v1, w0, w1);
    const SiteDesc sb = desc[k - 1];
    masks(sb, m0, m1);
    chain2_site<CHI>(pack + poff[k - 1], sb, m0, m1, w0, w1, v0, v1);
  }
  double r0 = v0[0], r1 = v1[0];
  if (k == 0) {  // an odd number of sites: the last one lands in w
    uint64_t m0, m1;
    const SiteDesc sa = desc[0];
    masks(sa, m0, m1);
    chain2_site<CHI>(pack + poff[0], sa, m0, m1, v0, v1, w0, w1);
    r0 = w0[0];
    r1 = w1[0];
  }
  if (act0) out[p0] = bad0 ? __builtin_nan("") : r0;
  if (act1) out[p1] = bad1 ? __builtin_nan("") : r1;
  if (bad0 || bad1) flag[0] = 1;
}

// float32 variants of the chain kernel.  SPLIT keeps each operand as a
// (hi, lo) float pair and accumulates in float-float (Dekker/Knuth error-free
// transformations), PLAIN is ordinary float32.
struct ff {
  float h, l;
};

// The error-free transformations need every + and * rounded on its own:
// contraction into FMA (hipcc's default) would silently drop the low word.
__device__ inline ff two_sum(float a, float b) {
#pragma clang fp contract(off)
  float s = a + b;
  float bb = s - a;
  float e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

__device__ inline ff ff_fma(float ah, float al, float bh, float bl, ff c) {
  // c + (ah + al) * (bh + bl), dropping al * bl
#pragma clang fp contract(off)
  float p = ah * bh;
  float pe = __builtin_fmaf(ah, bh, -p);
  pe = __builtin_fmaf(ah, bl, pe);
  pe = __builtin_fmaf(al, bh, pe);
  ff s = two_sum(c.h, p);
  float e = s.l + c.l + pe;
  float h = s.h + e;
  return {h, e - (h - s.h)};
}

template <bool SPLIT>
__global__ void __launch_bounds__(256)
eval_chain_f32_kernel(const float* __restrict__ pack, const int64_t* __restrict__ poff,
                      const int* __restrict__ bond, TopoView t, const double* __restrict__ xs,
                      int64_t npts, double* __restrict__ out, int* __restrict__ flag, int vstride) {
  extern __shared__ float ldsf[];
  const int tid = threadIdx.x;
  const int bs = blockDim.x;
  const int64_t p = (int64_t)blockIdx.x * bs + tid;
  const bool active = p < npts;
  const int nc = t.ncoord;
  bool bad = false;
  uint64_t kw[QTT_MAX_COORD] = {0, 0, 0, 0};
#pragma unroll
  for (int c = 0; c < QTT_MAX_COORD; ++c)
    if (c < nc) kw[c] = digit_word(active ? xs[p * nc + c] : 0.0, t.ndigits[c], bad);
  // vectors as (hi, lo) pairs: [i][lane][2]
  float* v = ldsf + 2 * tid;
  float* w = ldsf + (size_t)2 * vstride * bs + 2 * tid;
  v[0] = 1.0f;
  v[1] = 0.0f;
  for (int k = t.nv - 1; k >= 0; --k) {
    const int chil = bond[k];
    const int chir = (k == t.nv - 1) ? 1 : bond[k + 1];
    const int c = t.vcoord[k];
    int d = 0;
    if (c >= 0) {
      const uint64_t kk = c == 0 ? kw[0] : c == 1 ? kw[1] : c == 2 ? kw[2] : kw[3];
      d = (int)((kk >> (t.ndigits[c] - t.vdigit[k])) & 1u);
    }
    // packed float layout: [r][lpad][s][hi, lo]
    const float* __restrict__ P = pack + 2 * poff[k];
    const int lpad = (chil + 3) & ~3;
    for (int l0 = 0; l0 < chil; l0 += 4) {
      ff a0[4], a1[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a0[i] = a1[i] = {0.0f, 0.0f};
      const float* __restrict__ q = P + 4 * l0;
      for (int r = 0; r < chir; ++r) {
        const float vh = v[(size_t)2 * r * bs];
        const float vl = SPLIT ? v[(size_t)2 * r * bs + 1] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (SPLIT) {
            a0[i] = ff_fma(q[4 * i], q[4 * i + 1], vh, vl, a0[i]);
            a1[i] = ff_fma(q[4 * i + 2], q[4 * i + 3], vh, vl, a1[i]);
          } else {
            a0[i].h = __builtin_fmaf(q[4 * i], vh, a0[i].h);
            a1[i].h = __builtin_fmaf(q[4 * i + 2], vh, a1[i].h);
          }
        }
        q += 4 * lpad;
      }
      const uint32_t fmask = 0u - (uint32_t)d;
      float sh[4], sl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        sh[i] = bitself(a0[i].h, a1[i].h, fmask);
        sl[i] = bitself(a0[i].l, a1[i].l, fmask);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (l0 + i < chil) {
          w[(size_t)2 * (l0 + i) * bs] = sh[i];
          w[(size_t)2 * (l0 + i) * bs + 1] = sl[i];
        }
    }
    float* tmp = v;
    v = w;
    w = tmp;
  }
  if (active) out[p] = bad ? __builtin_nan("") : (double)v[0] + (double)v[1];
  if (bad) flag[0] = 1;
}

// float32 (hi, lo) pack of the float64 pack: [r][lpad][s] -> [r][lpad][s][2]
__global__ void pack_split_kernel(const double* __restrict__ pack, int64_t n, float* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double a = pack[e];
    float h = (float)a;
    float l = (float)(a - (double)h);
    out[2 * e] = h;
    out[2 * e + 1] = l;
  }
}

// Generic tree evaluation: one thread per point, per-vertex vectors in a
// global buffer laid out [vector element][thread] (coalesced).  Used for tree
// topologies; the vertex contraction sums over all child index combinations.
__global__ void __launch_bounds__(256)
eval_tree_kernel(const double* __restrict__ data, const int64_t* __restrict__ off,
                 const int* __restrict__ bond, TopoView t, const int64_t* __restrict__ voff,
                 const double* __restrict__ xs, int64_t npts, double* __restrict__ out,
                 int* __restrict__ flag, double* __restrict__ vs, int64_t nthr) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nthr) return;
  const int nc = t.ncoord;
  for (int64_t p = gid; p < npts; p += nthr) {
    bool bad = false;
    for (int c = 0; c < nc; ++c) {
      double x = xs[p * nc + c];
      if (!(x >= 0.0)) bad = true;
    }
    if (bad) {
      out[p] = __builtin_nan("");
      flag[0] = 1;
      continue;
    }
    for (int v = t.nv - 1; v >= 0; --v) {
      const int c = t.vcoord[v];
      int d = 0;
      if (c >= 0) {
        bool b2 = false;
        uint64_t kk = digit_word(xs[p * nc + c], t.ndigits[c], b2);
        d = (int)((kk >> (t.ndigits[c] - t.vdigit[v])) & 1u);
      }
      const int m = t.nch[v];
      const int* ch = t.chs + t.choff[v];
      int64_t pc = 1;
      for (int k = 0; k < m; ++k) pc *= bond[ch[k]];
      const int sd = t.sdim[v];
      for (int pp = 0; pp < bond[v]; ++pp) {
        const double* base = data + off[v] + ((int64_t)pp * sd + d) * pc;
        double acc = 0.0;
        for (int64_t j = 0; j < pc; ++j) {
          double wgt = base[j];
          int64_t r = j;
          for (int k = m - 1; k >= 0 && wgt != 0.0; --k) {
            const int b = bond[ch[k]];
            const int idx = (int)(r % b);
            r /= b;
            wgt *= vs[(voff[ch[k]] + idx) * nthr + gid];
          }
          acc += wgt;
        }
        vs[(voff[v] + pp) * nthr + gid] = acc;
      }
    }
    out[p] = vs[voff[0] * nthr + gid];
  }
}

// ---------------------------------------------------------------------------
// Chain evaluation with environment tables.
//
// f(x) = e^T T_0[d_0] ... T_{n-1}[d_{n-1}] e.  The product of the first ml
// sites depends only on their digits, so for a batch of many points every one
// of its 2^bl values (bl = digit sites among them) is computed once, as a
// table L[p] of row vectors; likewise the last mr sites give a table R[s] of
// column vectors.  A point then gathers L[p(x)] and R[s(x)] and runs only the
// middle sites ml .. n-mr-1 on the register chain body: f = L[p] M R[s].
// p takes the prefix digits in site order (first digit = most significant
// bit), s the suffix digits from the last site backwards.
// ---------------------------------------------------------------------------

// One table: the sites first, first + step, ... (count sites; step +1 for L,
// -1 for R), nbits of them with a digit.  Workgroup b fixes the first
// nbits - kb digits to the bits of b and expands the last kb in LDS, so it
// writes the 2^kb consecutive rows (b << kb) ... of the table.
struct EnvTableArgs {
  double* tab;   // [2^nbits][w]
  int first, count, step;
  int nbits, kb;
  int w;         // row width: the end bond rounded up to 4 (zero padded)
  int maxw;      // LDS row stride: the largest bond the build passes through
  int nblk;      // workgroups of this table (2^(nbits - kb))
};

// What one step of a table build reads of site k: both digit slices when the
// step expands a digit, else the slice d ([chil][ns][chir] staged).
struct EnvStep {
  const double* T;  // the core, [chil][sd][chir]
  int chil, chir, sd, d, expand;
};

__device__ __forceinline__ EnvStep env_step(const double* __restrict__ data, const int64_t* __restrict__ off,
                                            const int* __restrict__ bond, const TopoView& t, const EnvTableArgs& a,
                                            int i, int blk, int& seen) {
  const int k = a.first + i * a.step;
  EnvStep s;
  s.T = data + off[k];
  s.chil = bond[k];
  s.chir = (k == t.nv - 1) ? 1 : bond[k + 1];
  s.sd = t.sdim[k];
  s.d = 0;
  s.expand = 0;
  if (t.vcoord[k] >= 0) {
    const int top = a.nbits - a.kb;
    if (seen < top) s.d = (blk >> (top - 1 - seen)) & 1;
    else s.expand = 1;
    ++seen;
  }
  return s;
}

constexpr int kEnvPre = 8;  // staged core entries per thread: 2 * 32 * 32 / 256

// the entries of step s this thread stages: e = tid + bs * q, as [l][slice][o]
__device__ __forceinline__ void env_fetch(const EnvStep& s, int tid, int bs, double (&pre)[kEnvPre]) {
  const int ns = s.expand ? 2 : 1;
  const int n = s.chil * ns * s.chir;
#pragma unroll
  for (int q = 0; q < kEnvPre; ++q) {
    const int e = tid + bs * q;
    if (e < n) {
      const int l = e / (ns * s.chir), rest = e - l * ns * s.chir;
      const int sl = rest / s.chir, o = rest - sl * s.chir;
      pre[q] = s.T[((int64_t)l * s.sd + (s.expand ? sl : s.d)) * s.chir + o];
    }
  }
}

__global__ void __launch_bounds__(256)
env_table_kernel(const double* __restrict__ data, const int64_t* __restrict__ off, const int* __restrict__ bond,
                 TopoView t, EnvTableArgs ta, EnvTableArgs tb) {
  extern __shared__ double lds[];
  const bool isb = (int)blockIdx.x >= ta.nblk;
  const EnvTableArgs& a = isb ? tb : ta;
  const int blk = isb ? (int)blockIdx.x - ta.nblk : (int)blockIdx.x;
  const int tid = threadIdx.x, bs = blockDim.x;
  const bool left = a.step > 0;
  double* A = lds;
  double* B = lds + ((size_t)1 << a.kb) * a.maxw;
  double* Ts = B + ((size_t)1 << a.kb) * a.maxw;  // the step's core slices, [l][slice][o]
  if (tid == 0) A[0] = 1.0;  // the chain's end bonds are 1
  int cnt = 1, len = 1, seen = 0;
  // the first site's slices are fetched before the loop; every step then
  // fetches the next site's into registers while it computes
  double pre[kEnvPre];
  EnvStep cur;
  if (a.count > 0) {
    cur = env_step(data, off, bond, t, a, 0, blk, seen);
    env_fetch(cur, tid, bs, pre);
  }
  for (int i = 0; i < a.count; ++i) {
    const int ns = cur.expand ? 2 : 1;
    const int n = cur.chil * ns * cur.chir;
#pragma unroll
    for (int q = 0; q < kEnvPre; ++q)
      if (tid + bs * q < n) Ts[tid + bs * q] = pre[q];
    __syncthreads();
    EnvStep nxt = cur;
    if (i + 1 < a.count) {
      nxt = env_step(data, off, bond, t, a, i + 1, blk, seen);
      env_fetch(nxt, tid, bs, pre);
    }
    const int ncnt = cur.expand ? 2 * cnt : cnt;
    const int olen = left ? cur.chir : cur.chil;
    const int chir = cur.chir;
    for (int e = tid; e < ncnt * olen; e += bs) {
      const int j = e / olen, o = e - j * olen;
      const int src = cur.expand ? (j >> 1) : j;
      const int sl = cur.expand ? (j & 1) : 0;
      const double* x = A + (size_t)src * a.maxw;
      double s = 0.0;
      if (left) {
        for (int l = 0; l < len; ++l) s = __builtin_fma(x[l], Ts[(l * ns + sl) * chir + o], s);
      } else {
        const double* row = Ts + (o * ns + sl) * chir;
        for (int r = 0; r < len; ++r) s = __builtin_fma(row[r], x[r], s);
      }
      B[(size_t)j * a.maxw + o] = s;
    }
    __syncthreads();
    double* tmp = A;
    A = B;
    B = tmp;
    cnt = ncnt;
    len = olen;
    cur = nxt;
  }
  double* out = a.tab + ((size_t)blk << a.kb) * a.w;
  for (int e = tid; e < cnt * a.w; e += bs) {
    const int j = e / a.w, o = e - j * a.w;
    out[(size_t)j * a.w + o] = o < len ? A[(size_t)j * a.maxw + o] : 0.0;
  }
}

// What a point needs of the tables.
struct EnvView {
  const double* L;  // [2^bl][lw]
  const double* R;  // [2^br][rw]
  int lw, rw;
  int ml, mr;       // sites covered by L (0 .. ml-1) and by R (n-mr .. n-1)
};

__device__ __forceinline__ uint64_t word_of(const uint64_t (&kw)[QTT_MAX_COORD], int c) {
  return c == 0 ? kw[0] : c == 1 ? kw[1] : c == 2 ? kw[2] : kw[3];
}

// table indices of a point from its digit words
__device__ __forceinline__ void env_index(const SiteDesc* __restrict__ desc, int nv, const EnvView& ev,
                                          const uint64_t (&kw)[QTT_MAX_COORD], uint32_t& pi, uint32_t& si) {
  pi = 0;
  si = 0;
  for (int k = 0; k < ev.ml; ++k) {
    const SiteDesc sd = desc[k];
    if (sd.coord >= 0) pi = (pi << 1) | (uint32_t)((word_of(kw, sd.coord) >> sd.shift) & 1u);
  }
  for (int k = nv - 1; k >= nv - ev.mr; --k) {
    const SiteDesc sd = desc[k];
    if (sd.coord >= 0) si = (si << 1) | (uint32_t)((word_of(kw, sd.coord) >> sd.shift) & 1u);
  }
}

template <int CHI>
__device__ __forceinline__ void load_env(const double* __restrict__ row, int w, double (&v)[CHI]) {
#pragma unroll
  for (int i = 0; i < CHI; i += 2) {
    if (i < w) {
      const double2 q = *(const double2*)(row + i);
      v[i] = q.x;
      v[i + 1] = q.y;
    } else {
      v[i] = 0.0;
      v[i + 1] = 0.0;
    }
  }
}

template <int CHI>
__device__ __forceinline__ double dot_env(const double* __restrict__ row, int w, const double (&v)[CHI]) {
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int i = 0; i < CHI; i += 2) {
    if (i < w) {
      const double2 q = *(const double2*)(row + i);
      s0 = __builtin_fma(q.x, v[i], s0);
      s1 = __builtin_fma(q.y, v[i + 1], s1);
    }
  }
  return s0 + s1;
}

// One point per lane: the middle sites two at a time (v -> w, w -> v).
template <int CHI>
__global__ void __launch_bounds__(256)
eval_chain_env_kernel(const double* __restrict__ pack, const int64_t* __restrict__ poff,
                      const SiteDesc* __restrict__ desc, int nv, int nc, const int* __restrict__ ndigits, EnvView ev,
                      const double* __restrict__ xs, int64_t npts, double* __restrict__ out, int* __restrict__ flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < npts;
  bool bad = false;
  uint64_t kw[QTT_MAX_COORD] = {0, 0, 0, 0};
#pragma unroll
  for (int c = 0; c < QTT_MAX_COORD; ++c)
    if (c < nc) kw[c] = digit_word(active ? xs[p * nc + c] : 0.0, ndigits[c], bad);
  uint32_t pi, si;
  env_index(desc, nv, ev, kw, pi, si);
  double v[CHI], w[CHI];
  load_env<CHI>(ev.R + (size_t)si * ev.rw, ev.rw, v);
  auto mask = [&](const SiteDesc& sd) -> uint64_t {
    if (sd.coord < 0) return 0;
    return 0 - ((word_of(kw, sd.coord) >> sd.shift) & 1u);
  };
  int k = nv - ev.mr - 1;
  for (; k >= ev.ml + 1; k -= 2) {
    const SiteDesc sa = desc[k];
    chain1_site<CHI>(pack + poff[k], sa, mask(sa), v, w);
    const SiteDesc sb = desc[k - 1];
    chain1_site<CHI>(pack + poff[k - 1], sb, mask(sb), w, v);
  }
  double r;
  if (k == ev.ml) {  // an odd number of middle sites: the last one lands in w
    const SiteDesc sa = desc[k];
    chain1_site<CHI>(pack + poff[k], sa, mask(sa), v, w);
    r = dot_env<CHI>(ev.L + (size_t)pi * ev.lw, ev.lw, w);
  } else {
    r = dot_env<CHI>(ev.L + (size_t)pi * ev.lw, ev.lw, v);
  }
  if (active) out[p] = bad ? __builtin_nan("") : r;
  if (bad) flag[0] = 1;
}

// Two points per lane (p and p + blockDim.x of the block's span), as
// eval_chain_reg2_kernel: every core entry loaded feeds four multiply-adds.
template <int CHI>
__global__ void __launch_bounds__(256)
eval_chain_env2_kernel(const double* __restrict__ pack, const int64_t* __restrict__ poff,
                       const SiteDesc* __restrict__ desc, int nv, int nc, const int* __restrict__ ndigits,
                       EnvView ev, const double* __restrict__ xs, int64_t npts, double* __restrict__ out,
                       int* __restrict__ flag) {
  const int64_t p0 = (int64_t)blockIdx.x * (2 * blockDim.x) + threadIdx.x;
  const int64_t p1 = p0 + blockDim.x;
  const bool act0 = p0 < npts, act1 = p1 < npts;
  bool bad0 = false, bad1 = false;
  uint64_t kw0[QTT_MAX_COORD] = {0, 0, 0, 0}, kw1[QTT_MAX_COORD] = {0, 0, 0, 0};
#pragma unroll
  for (int c = 0; c < QTT_MAX_COORD; ++c) {
    if (c < nc) {
      kw0[c] = digit_word(act0 ? xs[p0 * nc + c] : 0.0, ndigits[c], bad0);
      kw1[c] = digit_word(act1 ? xs[p1 * nc + c] : 0.0, ndigits[c], bad1);
    }
  }
  uint32_t pi0, si0, pi1, si1;
  env_index(desc, nv, ev, kw0, pi0, si0);
  env_index(desc, nv, ev, kw1, pi1, si1);
  double v0[CHI], v1[CHI], w0[CHI], w1[CHI];
  load_env<CHI>(ev.R + (size_t)si0 * ev.rw, ev.rw, v0);
  load_env<CHI>(ev.R + (size_t)si1 * ev.rw, ev.rw, v1);
  auto masks = [&](const SiteDesc& sd, uint64_t& m0, uint64_t& m1) {
    m0 = 0;
    m1 = 0;
    if (sd.coord >= 0) {
      m0 = 0 - ((word_of(kw0, sd.coord) >> sd.shift) & 1u);
      m1 = 0 - ((word_of(kw1, sd.coord) >> sd.shift) & 1u);
    }
  };
  int k = nv - ev.mr - 1;
  for (; k >= ev.ml + 1; k -= 2) {
    uint64_t m0, m1;
    const SiteDesc sa = desc[k];
    masks(sa, m0, m1);
    chain2_site<CHI>(pack + poff[k], sa, m0, m1, v0, v1, w0, w1);
    const SiteDesc sb = desc[k - 1];
    masks(sb, m0, m1);
    chain2_site<CHI>(pack + poff[k - 1], sb, m0, m1, w0, w1, v0, v1);
  }
  double r0, r1;
  if (k == ev.ml) {
    uint64_t m0, m1;
    const SiteDesc sa = desc[k];
    masks(sa, m0, m1);
    chain2_site<CHI>(pack + poff[k], sa, m0, m1, v0, v1, w0, w1);
    r0 = dot_env<CHI>(ev.L + (size_t)pi0 * ev.lw, ev.lw, w0);
    r1 = dot_env<CHI>(ev.L + (size_t)pi1 * ev.lw, ev.lw, w1);
  } else {
    r0 = dot_env<CHI>(ev.L + (size_t)pi0 * ev.lw, ev.lw, v0);
    r1 = dot_env<CHI>(ev.L + (size_t)pi1 * ev.lw, ev.lw, v1);
  }
  if (act0) out[p0] = bad0 ? __builtin_nan("") : r0;
  if (act1) out[p1] = bad1 ? __builtin_nan("") : r1;
  if (bad0 || bad1) flag[0] = 1;
}

}  // namespace

static int ensure_pack(qtt_ctx* ctx, qtt_net* net) {
  if (net->pack_valid) return QTT_OK;
  const int nv = net->t.nv;
  // host tables kept on the network: they outlive the asynchronous copies below
  std::vector<int64_t>& poff = net->h_pack_off;
  std::vector<int>& desc = net->h_desc;
  poff.assign(nv, 0);
  desc.assign(4 * nv, 0);
  int64_t o = 0;
  for (int k = 0; k < nv; k++) {
    int64_t chil = net->bond[k];
    int64_t chir = (k == nv - 1) ? 1 : net->bond[k + 1];
    int64_t lpad = (chil + 3) & ~int64_t(3);
    int64_t rpad = (chir + 3) & ~int64_t(3);
    poff[k] = o;
    o += rpad * lpad * 2;
    const int c = net->t.vcoord[k];
    desc[4 * k + 0] = (int)chil;
    desc[4 * k + 1] = (int)chir;
    desc[4 * k + 2] = c;
    desc[4 * k + 3] = c >= 0 ? net->t.ndigits[c] - net->t.vdigit[k] : 0;
  }
  if (!net->d_desc) {
    int rc = qtt_hip_check(ctx, hipMalloc(&net->d_desc, 4 * nv * sizeof(int)), "hipMalloc site descriptors");
    if (rc) return rc;
  }
  {
    int rc = qtt_hip_check(ctx, hipMemcpyAsync(net->d_desc, desc.data(), 4 * nv * sizeof(int), hipMemcpyHostToDevice, ctx->stream), "site descriptors");
    if (rc) return rc;
  }
  if (net->d_pack && o > net->pack_elems) {
    (void)hipFree(net->d_pack);
    net->d_pack = nullptr;
    if (net->d_pack32) (void)hipFree(net->d_pack32);
    net->d_pack32 = nullptr;
  }
  if (!net->d_pack) {
    int rc = qtt_hip_check(ctx, hipMalloc(&net->d_pack, std::max<int64_t>(o, 1) * sizeof(double)), "hipMalloc pack");
    if (rc) return rc;
    net->pack_elems = o;
  }
  if (!net->d_pack_off) {
    int rc = qtt_hip_check(ctx, hipMalloc(&net->d_pack_off, nv * sizeof(int64_t)), "hipMalloc pack offsets");
    if (rc) return rc;
  }
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(net->d_pack_off, poff.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "pack offsets");
  if (rc) return rc;
  TopoView tv = net->view();
  dim3 grid(64, nv);
  pack_chain_kernel<<<grid, 256, 0, ctx->stream>>>(net->d_data, net->d_off, net->d_bond, tv.sdim,
                                                   net->d_pack_off, nv, net->d_pack);
  rc = qtt_hip_check(ctx, hipGetLastError(), "pack_chain_kernel");
  if (rc) return rc;
  net->pack_valid = true;
  net->pack32_valid = false;
  net->env_valid = false;
  return QTT_OK;
}

int qtt_ensure_pack(qtt_ctx* ctx, qtt_net* net) { return ensure_pack(ctx, net); }

static int ensure_pack32(qtt_ctx* ctx, qtt_net* net);
int qtt_ensure_pack32(qtt_ctx* ctx, qtt_net* net) { return ensure_pack32(ctx, net); }

static int ensure_pack32(qtt_ctx* ctx, qtt_net* net) {
  int rc = ensure_pack(ctx, net);
  if (rc) return rc;
  if (net->pack32_valid) return QTT_OK;
  if (!net->d_pack32) {
    rc = qtt_hip_check(ctx, hipMalloc(&net->d_pack32, std::max<int64_t>(net->pack_elems, 1) * 2 * sizeof(float)), "hipMalloc pack32");
    if (rc) return rc;
  }
  pack_split_kernel<<<256, 256, 0, ctx->stream>>>(net->d_pack, net->pack_elems, net->d_pack32);
  rc = qtt_hip_check(ctx, hipGetLastError(), "pack_split_kernel");
  if (rc) return rc;
  net->pack32_valid = true;
  return QTT_OK;
}

// Key of a tuned launch shape: the register kernel's bond class and the octave
// of the point count.
static uint32_t eval_tune_key(int64_t maxb, int64_t npts) {
  const uint32_t cls = maxb <= 4 ? 0 : maxb <= 8 ? 1 : maxb <= 16 ? 2 : 3;
  uint32_t oct = 0;
  while (oct < 62 && (int64_t(1) << (oct + 1)) <= npts) ++oct;
  return cls | (oct << 4);
}

// The register chain kernel (float64 chains, bonds <= 32) with ppl points per
// lane (2 only for bonds <= 16) and wg threads per workgroup (64..256).
static int launch_eval_reg_on(qtt_ctx* ctx, const qtt_net* net, const double* pack, const int64_t* pack_off,
                              const SiteDesc* dd, const double* xs, int64_t npts, double* out, int64_t maxb, int ppl,
                              int wg);

static int launch_eval_reg(qtt_ctx* ctx, qtt_net* net, const double* xs, int64_t npts, double* out, int64_t maxb,
                           int ppl, int wg) {
  return launch_eval_reg_on(ctx, net, net->d_pack, net->d_pack_off, (const SiteDesc*)net->d_desc, xs, npts, out,
                            maxb, ppl, wg);
}

// The register chain kernels on a packed copy of the chain (pack, pack_off,
// dd: as ensure_pack lays them out).
static int launch_eval_reg_on(qtt_ctx* ctx, const qtt_net* net, const double* pack, const int64_t* pack_off,
                              const SiteDesc* dd, const double* xs, int64_t npts, double* out, int64_t maxb, int ppl,
                              int wg) {
  const HostTopo& t = net->t;
  TopoView tv = net->view();
  const int bsr = wg;
  if (ppl == 2 && maxb <= 16) {
    const int64_t nb2 = (npts + 2 * bsr - 1) / (2 * bsr);
    if (nb2 > INT32_MAX) return qtt_fail(ctx, QTT_EINVAL, "evaluate: too many points");
#define QTT_REG2_LAUNCH(C)                                                                         \
  eval_chain_reg2_kernel<C><<<dim3((unsigned)nb2), bsr, 0, ctx->stream>>>(                         \
      pack, pack_off, dd, t.nv, t.ncoord, tv.ndigits, xs, npts, out, ctx->flags())
    if (maxb <= 4) QTT_REG2_LAUNCH(4);
    else if (maxb <= 8) QTT_REG2_LAUNCH(8);
    else QTT_REG2_LAUNCH(16);
#undef QTT_REG2_LAUNCH
    return qtt_hip_check(ctx, hipGetLastError(), "eval_chain_reg2_kernel");
  }
  const int64_t nbr = (npts + bsr - 1) / bsr;
  if (nbr > INT32_MAX) return qtt_fail(ctx, QTT_EINVAL, "evaluate: too many points");
#define QTT_REG_LAUNCH(C)                                                                          \
  eval_chain_reg_kernel<C><<<dim3((unsigned)nbr), bsr, 0, ctx->stream>>>(                          \
      pack, pack_off, dd, t.nv, t.ncoord, tv.ndigits, xs, npts, out, ctx->flags())
  if (maxb <= 4) QTT_REG_LAUNCH(4);
  else if (maxb <= 8) QTT_REG_LAUNCH(8);
  else if (maxb <= 16) QTT_REG_LAUNCH(16);
  else QTT_REG_LAUNCH(32);
#undef QTT_REG_LAUNCH
  return qtt_hip_check(ctx, hipGetLastError(), "eval_chain_reg_kernel");
}

// ---------------------------------------------------------------------------
// Environment tables (see env_table_kernel): plan, build, launch.
// ---------------------------------------------------------------------------
struct EnvPlan {
  int ml = 0, mr = 0;    // sites under the left / right table
  int bl = 0, br = 0;    // digit sites among them (table rows 2^bl, 2^br)
  int64_t lw = 4, rw = 4;  // row widths: end bonds rounded up to 4
};

// tb digit sites per table, at most half of the chain's digit sites each
static EnvPlan env_plan(const qtt_net* net, int tb) {
  const HostTopo& t = net->t;
  int D = 0;
  for (int k = 0; k < t.nv; k++) D += t.vcoord[k] >= 0;
  tb = std::max(0, std::min(tb, D / 2));
  EnvPlan p;
  int seen = 0, k = 0;
  while (seen < tb) seen += t.vcoord[k++] >= 0;
  p.ml = k;
  p.bl = tb;
  seen = 0;
  k = t.nv - 1;
  while (seen < tb) {
    seen += t.vcoord[k--] >= 0;
    p.mr++;
  }
  p.br = tb;
  p.lw = (net->bond[p.ml] + 3) & ~int64_t(3);
  p.rw = ((p.mr ? net->bond[t.nv - p.mr] : 1) + 3) & ~int64_t(3);
  return p;
}

// Digit sites per table when nothing is tuned: enough points that a table
// row is reused about 32 times, at most 16.
static int default_table_bits(int64_t npts) {
  if (npts < (int64_t(1) << 15)) return 0;
  int lg = 0;
  while ((int64_t(1) << (lg + 1)) <= npts) ++lg;
  return std::min(16, lg - 5);
}

static EnvTableArgs env_table_args(const qtt_net* net, const EnvPlan& pl, bool left, double* tab) {
  const HostTopo& t = net->t;
  EnvTableArgs e;
  e.tab = tab;
  e.first = left ? 0 : t.nv - 1;
  e.count = left ? pl.ml : pl.mr;
  e.step = left ? 1 : -1;
  e.nbits = left ? pl.bl : pl.br;
  e.w = (int)(left ? pl.lw : pl.rw);
  int maxw = 1;
  for (int i = 0; i < e.count; i++) {
    const int k = e.first + i * e.step;
    maxw = std::max<int>(maxw, (int)net->bond[k]);
    maxw = std::max<int>(maxw, k == t.nv - 1 ? 1 : (int)net->bond[k + 1]);
  }
  e.maxw = maxw;
  // rows expanded per workgroup: both row buffers and the staged core slices
  // (2 maxw^2) within 64 KiB of LDS
  const size_t stage = 2 * size_t(maxw) * maxw;
  int kb = 0;
  while (kb < e.nbits && (2 * (size_t(2) << kb) * maxw + stage) * sizeof(double) <= 64 * 1024) ++kb;
  e.kb = kb;
  e.nblk = 1 << (e.nbits - kb);
  return e;
}

static size_t env_lds(const EnvTableArgs& e) {
  return (2 * (size_t(1) << e.kb) * e.maxw + 2 * size_t(e.maxw) * e.maxw) * sizeof(double);
}

// Build the two tables in scratch and evaluate with them: two launches.
static int launch_eval_env(qtt_ctx* ctx, qtt_net* net, const double* xs, int64_t npts, double* out, int64_t maxb,
                           int ppl, int wg, int tb) {
  const HostTopo& t = net->t;
  const EnvPlan pl = env_plan(net, tb);
  if (pl.bl == 0) return launch_eval_reg(ctx, net, xs, npts, out, maxb, ppl, wg);
  const int64_t nL = (int64_t(1) << pl.bl) * pl.lw, nR = (int64_t(1) << pl.br) * pl.rw;
  const TopoView tv = net->view();
  // The tables depend only on the cores: build them on the first call with
  // this size and keep them with the network (as the packed copy) until the
  // cores change.
  if (!(net->env_valid && net->env_tb == tb)) {
    net->env_valid = false;
    const int64_t need = nL + nR;
    if (net->d_env && (net->env_elems < need || net->env_elems > 4 * need)) {
      (void)hipFree(net->d_env);
      net->d_env = nullptr;
      net->env_elems = 0;
    }
    if (!net->d_env) {
      int rc = qtt_hip_check(ctx, hipMalloc(&net->d_env, need * sizeof(double)), "hipMalloc environment tables");
      if (rc) return rc;
      net->env_elems = need;
    }
    const EnvTableArgs a = env_table_args(net, pl, true, net->d_env), b = env_table_args(net, pl, false, net->d_env + nL);
    env_table_kernel<<<a.nblk + b.nblk, 256, std::max(env_lds(a), env_lds(b)), ctx->stream>>>(
        net->d_data, net->d_off, net->d_bond, tv, a, b);
    int rc = qtt_hip_check(ctx, hipGetLastError(), "env_table_kernel");
    if (rc) return rc;
    net->env_valid = true;
    net->env_tb = tb;
  }
  double* scr = net->d_env;
  EnvView ev{scr, scr + nL, (int)pl.lw, (int)pl.rw, pl.ml, pl.mr};
  const SiteDesc* dd = (const SiteDesc*)net->d_desc;
  const int bsr = wg;
  if (ppl == 2 && maxb <= 16) {
    const int64_t nb2 = (npts + 2 * bsr - 1) / (2 * bsr);
    if (nb2 > INT32_MAX) return qtt_fail(ctx, QTT_EINVAL, "evaluate: too many points");
#define QTT_ENV2_LAUNCH(C)                                                                         \
  eval_chain_env2_kernel<C><<<dim3((unsigned)nb2), bsr, 0, ctx->stream>>>(                         \
      net->d_pack, net->d_pack_off, dd, t.nv, t.ncoord, tv.ndigits, ev, xs, npts, out, ctx->flags())
    if (maxb <= 4) QTT_ENV2_LAUNCH(4);
    else if (maxb <= 8) QTT_ENV2_LAUNCH(8);
    else QTT_ENV2_LAUNCH(16);
#undef QTT_ENV2_LAUNCH
    return qtt_hip_check(ctx, hipGetLastError(), "eval_chain_env2_kernel");
  }
  const int64_t nbr = (npts + bsr - 1) / bsr;
  if (nbr > INT32_MAX) return qtt_fail(ctx, QTT_EINVAL, "evaluate: too many points");
#define QTT_ENV_LAUNCH(C)                                                                          \
  eval_chain_env_kernel<C><<<dim3((unsigned)nbr), bsr, 0, ctx->stream>>>(                          \
      net->d_pack, net->d_pack_off, dd, t.nv, t.ncoord, tv.ndigits, ev, xs, npts, out, ctx->flags())
  if (maxb <= 4) QTT_ENV_LAUNCH(4);
  else if (maxb <= 8) QTT_ENV_LAUNCH(8);
  else if (maxb <= 16) QTT_ENV_LAUNCH(16);
  else QTT_ENV_LAUNCH(32);
#undef QTT_ENV_LAUNCH
  return qtt_hip_check(ctx, hipGetLastError(), "eval_chain_env_kernel");
}

// Time every launch shape of the float64 register chain evaluation on these
// points (one warm-up and five timed launches each, HIP events on the
// context's stream) and keep the fastest for this bond class and point-count
// octave.  Shapes: digit sites per environment table (0: none, else 8 .. 16,
// at most half of the digit sites), points per lane 1 or 2, workgroups of 64,
// 128 or 256 threads.  cfg receives {points per lane, workgroup size, table
// digit sites} ({0, 0, 0}: the network does not take the register kernel), ms
// the time of one call once the network's tables exist (the warm-up builds
// them).
int tune_eval(qtt_ctx* ctx, qtt_net* net, const double* xs, int64_t npts, double* out, int* cfg, double* ms) {
  cfg[0] = cfg[1] = cfg[2] = 0;
  *ms = 0.0;
  const HostTopo& t = net->t;
  int64_t maxb = 1;
  for (int v = 0; v < t.nv; v++) maxb = std::max(maxb, net->bond[v]);
  if (!t.chain || t.ncoord > QTT_MAX_COORD || maxb > 32 || npts <= 0) return QTT_OK;
  int rc = ensure_pack(ctx, net);
  if (rc) return rc;
  int D = 0;
  for (int k = 0; k < t.nv; k++) D += t.vcoord[k] >= 0;
  // tables of at most 2^16 rows each (8 MiB at bond 16): with the tables kept
  // on the network the timed launches below measure the steady state, where
  // larger tables would always win by turning the chain into two lookups
  std::vector<int> tbs{0};
  for (int tb = 8; tb <= std::min(16, D / 2); tb += 2) tbs.push_back(tb);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = qtt_hip_check(ctx, hipEventCreate(&e0), "tune event");
  if (!rc) rc = qtt_hip_check(ctx, hipEventCreate(&e1), "tune event");
  float best = 0.0f;
  int bp = 0, bw = 0, bt = 0;
  for (int tb : tbs) {
    for (int ppl = 1; ppl <= (maxb <= 16 ? 2 : 1) && !rc; ++ppl) {
      for (int wg = 64; wg <= 256 && !rc; wg *= 2) {
        rc = launch_eval_env(ctx, net, xs, npts, out, maxb, ppl, wg, tb);
        if (!rc) rc = qtt_hip_check(ctx, hipEventRecord(e0, ctx->stream), "tune event");
        for (int r = 0; r < 5 && !rc; ++r) rc = launch_eval_env(ctx, net, xs, npts, out, maxb, ppl, wg, tb);
        if (!rc) rc = qtt_hip_check(ctx, hipEventRecord(e1, ctx->stream), "tune event");
        if (!rc) rc = qtt_hip_check(ctx, hipEventSynchronize(e1), "tune sync");
        float f = 0.0f;
        if (!rc) rc = qtt_hip_check(ctx, hipEventElapsedTime(&f, e0, e1), "tune time");
        if (!rc && (bp == 0 || f < best)) {
          best = f;
          bp = ppl;
          bw = wg;
          bt = tb;
        }
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  ctx->eval_tuned[eval_tune_key(maxb, npts)] = {bp, bw, bt};
  cfg[0] = bp;
  cfg[1] = bw;
  cfg[2] = bt;
  *ms = best / 5.0;
  return QTT_OK;
}

int launch_eval(qtt_ctx* ctx, qtt_net* net, const double* xs, int64_t npts, double* out,
                int precision) {
  const HostTopo& t = net->t;
  TopoView tv = net->view();
  int64_t maxb = 1;
  for (int v = 0; v < t.nv; v++) maxb = std::max(maxb, net->bond[v]);
  if (t.chain && t.ncoord <= QTT_MAX_COORD) {
    int rc = precision == QTT_F64 ? ensure_pack(ctx, net) : ensure_pack32(ctx, net);
    if (rc) return rc;
    if (precision == QTT_F64 && maxb <= 32 && !ctx->force_lds_eval) {
      // tuned shape if qtt_tune_eval measured one for this bond class and point
      // count; else two points per lane for bonds <= 16 (register room) when
      // there are enough points to fill the GPU at half the workgroup count
      int ppl = maxb <= 16 && npts >= (int64_t(1) << 19) ? 2 : 1, wg = 256;
      int tb = default_table_bits(npts);
      const auto it = ctx->eval_tuned.find(eval_tune_key(maxb, npts));
      if (it != ctx->eval_tuned.end()) {
        ppl = it->second.ppl;
        wg = it->second.wg;
        tb = it->second.tbits;
      }
      if (ctx->eval_ppl) ppl = ctx->eval_ppl;
      if (ctx->eval_tbits >= 0) tb = ctx->eval_tbits;
      return launch_eval_env(ctx, net, xs, npts, out, maxb, ppl, wg, tb);
    }
    // LDS: two vectors of maxb entries per lane (double, or a float pair)
    const size_t per_lane = 2 * size_t(maxb) * sizeof(double);
    int bs = 256;
    while (bs > 64 && per_lane * bs > 160 * 1024) bs /= 2;
    if (per_lane * bs > 160 * 1024)
      return qtt_fail(ctx, QTT_EINVAL, "evaluate: bond dimension too large for the chain kernel");
    const int64_t nb = (npts + bs - 1) / bs;
    if (nb > INT32_MAX) return qtt_fail(ctx, QTT_EINVAL, "evaluate: too many points");
    const size_t lds = per_lane * bs;
    if (lds > 64 * 1024) {
      const void* fn = precision == QTT_F64 ? (const void*)eval_chain_f64_kernel
                       : precision == QTT_F32_SPLIT ? (const void*)eval_chain_f32_kernel<true>
                                                    : (const void*)eval_chain_f32_kernel<false>;
      rc = qtt_hip_check(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                         "evaluate LDS attribute");
      if (rc) return rc;
    }
    if (precision == QTT_F64) {
      eval_chain_f64_kernel<<<dim3((unsigned)nb), bs, lds, ctx->stream>>>(
          net->d_pack, net->d_pack_off, net->d_bond, tv, xs, npts, out, ctx->flags(), (int)maxb);
    } else if (precision == QTT_F32_SPLIT) {
      eval_chain_f32_kernel<true><<<dim3((unsigned)nb), bs, lds, ctx->stream>>>(
          (const float*)net->d_pack32, net->d_pack_off, net->d_bond, tv, xs, npts, out, ctx->flags(), (int)maxb);
    } else {
      eval_chain_f32_kernel<false><<<dim3((unsigned)nb), bs, lds, ctx->stream>>>(
          (const float*)net->d_pack32, net->d_pack_off, net->d_bond, tv, xs, npts, out, ctx->flags(), (int)maxb);
    }
    return qtt_hip_check(ctx, hipGetLastError(), "eval_chain kernel");
  }
  if (precision != QTT_F64)
    return qtt_fail(ctx, QTT_EINVAL, "evaluate: float32 paths are implemented for chains only");
  // generic tree path
  std::vector<int64_t> voff(t.nv);
  int64_t tot = 0;
  for (int v = 0; v < t.nv; v++) {
    voff[v] = tot;
    tot += net->bond[v];
  }
  int64_t nthr = std::min<int64_t>(npts, 65536);
  while (nthr > 256 && (size_t)tot * nthr * sizeof(double) > (size_t(1) << 31)) nthr /= 2;
  size_t bytes = (size_t)tot * nthr * sizeof(double) + t.nv * sizeof(int64_t) + 64;
  double* scr = qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "evaluate: scratch");
  int64_t* d_voff = (int64_t*)(scr + (size_t)tot * nthr);
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_voff, voff.data(), t.nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "voff");
  if (rc) return rc;
  int bs = 256;
  int64_t nb = (nthr + bs - 1) / bs;
  eval_tree_kernel<<<dim3((unsigned)nb), bs, 0, ctx->stream>>>(net->d_data, net->d_off, net->d_bond, tv, d_voff, xs, npts, out,
                                                               ctx->flags(), scr, nthr);
  rc = qtt_hip_check(ctx, hipGetLastError(), "eval_tree_kernel");
  if (rc) return rc;
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "eval_tree sync");
}

// f_i at its own points for a batch of networks.  Float64 chains with bonds
// <= 32 run in one launch per bond class (4, 8, 16, 32); the others take the
// single-network launcher one by one.  Blocks until the launches are queued
// and their argument tables copied.
int launch_eval_batch(qtt_ctx* ctx, int n, qtt_net* const* nets, const double* const* xs, const int64_t* npts,
                      double* const* out, int precision) {
  std::vector<EvalItem> items[4];
  std::vector<int> item_of[4];
  for (int i = 0; i < n; i++) {
    qtt_net* net = nets[i];
    if (npts[i] == 0) continue;
    const HostTopo& t = net->t;
    int64_t maxb = 1;
    for (int v = 0; v < t.nv; v++) maxb = std::max(maxb, net->bond[v]);
    if (precision != QTT_F64 || !t.chain || t.ncoord > QTT_MAX_COORD || maxb > 32 || ctx->force_lds_eval) {
      int rc = launch_eval(ctx, net, xs[i], npts[i], out[i], precision);
      if (rc) return rc;
      continue;
    }
    int rc = ensure_pack(ctx, net);
    if (rc) return rc;
    const int cls = maxb <= 4 ? 0 : maxb <= 8 ? 1 : maxb <= 16 ? 2 : 3;
    EvalItem it;
    it.pack = net->d_pack;
    it.poff = net->d_pack_off;
    it.desc = (const SiteDesc*)net->d_desc;
    it.ndigits = net->view().ndigits;
    it.xs = xs[i];
    it.out = out[i];
    it.npts = npts[i];
    it.block0 = (int64_t)item_of[cls].size();
    it.nv = t.nv;
    it.nc = t.ncoord;
    const int64_t nb = (npts[i] + 255) / 256;
    if (it.block0 + nb > INT32_MAX) return qtt_fail(ctx, QTT_EINVAL, "evaluate batch: too many points");
    item_of[cls].insert(item_of[cls].end(), (size_t)nb, (int)items[cls].size());
    items[cls].push_back(it);
  }
  size_t bytes = 256;
  for (int c = 0; c < 4; c++)
    bytes += items[c].size() * sizeof(EvalItem) + item_of[c].size() * sizeof(int) + 512;
  char* scr = (char*)qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "evaluate batch scratch");
  for (int c = 0; c < 4; c++) {
    if (items[c].empty()) continue;
    EvalItem* d_items = (EvalItem*)scr;
    scr += (items[c].size() * sizeof(EvalItem) + 255) & ~size_t(255);
    int* d_map = (int*)scr;
    scr += (item_of[c].size() * sizeof(int) + 255) & ~size_t(255);
    int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_items, items[c].data(), items[c].size() * sizeof(EvalItem), hipMemcpyHostToDevice, ctx->stream), "batch items");
    if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d_map, item_of[c].data(), item_of[c].size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream), "batch map");
    if (rc) return rc;
    const dim3 grid((unsigned)item_of[c].size());
    if (c == 0) eval_chain_batch_kernel<4><<<grid, 256, 0, ctx->stream>>>(d_items, d_map, ctx->flags());
    else if (c == 1) eval_chain_batch_kernel<8><<<grid, 256, 0, ctx->stream>>>(d_items, d_map, ctx->flags());
    else if (c == 2) eval_chain_batch_kernel<16><<<grid, 256, 0, ctx->stream>>>(d_items, d_map, ctx->flags());
    else eval_chain_batch_kernel<32><<<grid, 256, 0, ctx->stream>>>(d_items, d_map, ctx->flags());
    rc = qtt_hip_check(ctx, hipGetLastError(), "eval_chain_batch_kernel");
    if (rc) return rc;
  }
  // the host tables must outlive their asynchronous copies
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "evaluate batch sync");
}

// ---------------------------------------------------------------------------
// Evaluate inside a pipeline (qtt_batch_run): the chains' bonds are known only
// on the device (a truncation kernel earlier on the stream sets them), so the
// packing layout and each network's bond class are computed by kernels.  Each
// network gets a pack region sized for bonds clipped to 32 (the largest the
// register kernels take); one class-filtered launch per register class then
// evaluates the networks of that class, and networks whose bonds exceed 32 are
// left to the single-network path once the bonds are back on the host.
// ---------------------------------------------------------------------------
struct PipeItem {
  EvalItem e;              // pack, poff, desc point into this network's region
  const double* data;      // source cores
  const int64_t* off;
  const int* bond;         // device bonds, final only when the pipeline runs
  const int* sdim;
  const int* vcoord;
  const int* vdigit;
  const int* ndigits;
  int64_t pack_cap;        // doubles in the pack region
};

// Packing offsets and site descriptors of a chain from its bonds (one thread);
// returns the packed size in doubles.
__device__ int64_t chain_layout(const int* __restrict__ bond, int nv, const int* __restrict__ vcoord,
                                const int* __restrict__ vdigit, const int* __restrict__ ndigits,
                                int64_t* __restrict__ poff, int* __restrict__ desc) {
  int64_t o = 0;
  for (int k = 0; k < nv; ++k) {
    const int chil = bond[k];
    const int chir = (k == nv - 1) ? 1 : bond[k + 1];
    const int64_t lpad = (chil + 3) & ~3, rpad = (chir + 3) & ~3;
    if (poff) poff[k] = o;
    o += rpad * lpad * 2;
    const int c = vcoord[k];
    if (desc) {
      desc[4 * k + 0] = chil;
      desc[4 * k + 1] = chir;
      desc[4 * k + 2] = c;
      desc[4 * k + 3] = c >= 0 ? ndigits[c] - vdigit[k] : 0;
    }
  }
  return o;
}

// One thread per network: bond class (0..3 for bonds <= 4, 8, 16, 32; 4 when
// larger or when the packed chain exceeds its region) and the layout tables.
__global__ void pipe_meta_kernel(const PipeItem* __restrict__ items, int n, int* __restrict__ cls) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PipeItem& it = items[i];
  int maxb = 1;
  for (int k = 0; k < it.e.nv; ++k) maxb = max(maxb, it.bond[k]);
  const int64_t need = chain_layout(it.bond, it.e.nv, it.vcoord, it.vdigit, it.ndigits, nullptr, nullptr);
  int c = maxb <= 4 ? 0 : maxb <= 8 ? 1 : maxb <= 16 ? 2 : maxb <= 32 ? 3 : 4;
  if (need > it.pack_cap) c = 4;
  if (c < 4)
    chain_layout(it.bond, it.e.nv, it.vcoord, it.vdigit, it.ndigits, const_cast<int64_t*>(it.e.poff),
                 (int*)const_cast<SiteDesc*>(it.e.desc));
  cls[i] = c;
}

// grid (x-blocks, site, networks): the packed copy of site k of each network
// z, blockIdx.z + j * gridDim.z
__global__ void pipe_pack_kernel(const PipeItem* __restrict__ items, const int* __restrict__ cls, int n) {
  const int k = blockIdx.y;
  for (int z = blockIdx.z; z < n; z += gridDim.z) {
    const PipeItem& it = items[z];
    if (cls[z] >= 4 || k >= it.e.nv) continue;
    const int chil = it.bond[k];
    const int chir = (k == it.e.nv - 1) ? 1 : it.bond[k + 1];
    const int lpad = (chil + 3) & ~3;
    const int rpad = (chir + 3) & ~3;
    const int sd = it.sdim[k];
    const int64_t ne = (int64_t)rpad * lpad * 2;
    const double* T = it.data + it.off[k];
    double* P = const_cast<double*>(it.e.pack) + it.e.poff[k];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
      const int s = (int)(e & 1);
      const int64_t rest = e >> 1;
      const int l = (int)(rest % lpad);
      const int r = (int)(rest / lpad);
      double val = 0.0;
      if (l < chil && r < chir) val = T[((int64_t)l * sd + (sd == 1 ? 0 : s)) * chir + r];
      P[e] = val;
    }
  }
}

// Blocks of every network's points; a block whose network is not of class
// `want` returns at once.
template <int CHI>
__global__ void __launch_bounds__(256)
eval_pipe_kernel(const PipeItem* __restrict__ items, const int* __restrict__ item_of, const int* __restrict__ cls,
                 int want, int* __restrict__ flag) {
  const int i = item_of[blockIdx.x];
  if (cls[i] != want) return;
  const EvalItem& it = items[i].e;
  eval_chain_reg_body<CHI>(it.pack, it.poff, it.desc, it.nv, it.nc, it.ndigits, it.xs, it.npts, it.out, flag,
                           (int64_t)(blockIdx.x - it.block0) * blockDim.x + threadIdx.x);
}

namespace {
int64_t pipe_pack_bound(const qtt_net* net) {
  const int nv = net->t.nv;
  int64_t elems = 0;
  for (int k = 0; k < nv; k++) {
    const int64_t chil = std::min<int64_t>(net->bond[k], 32);
    const int64_t chir = k == nv - 1 ? 1 : std::min<int64_t>(net->bond[k + 1], 32);
    elems += ((chil + 3) & ~int64_t(3)) * ((chir + 3) & ~int64_t(3)) * 2;
  }
  return elems;
}

bool pipe_eligible(const qtt_net* net, int64_t npts, int precision) {
  return npts > 0 && precision == QTT_F64 && net->t.chain && net->t.ncoord <= QTT_MAX_COORD;
}

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

size_t eval_pipe_bytes(int n, qtt_net* const* nets, const int64_t* npts) {
  size_t bytes = al256((size_t)n * sizeof(PipeItem)) + al256((size_t)n * sizeof(int)) + 512;
  int64_t blocks = 0;
  for (int i = 0; i < n; i++) {
    if (npts[i] <= 0) continue;
    const int nv = nets[i]->t.nv;
    bytes += al256(pipe_pack_bound(nets[i]) * sizeof(double)) + al256(nv * sizeof(int64_t)) +
             al256(4 * nv * sizeof(int));
    blocks += (npts[i] + 255) / 256;
  }
  return bytes + al256((size_t)blocks * sizeof(int));
}

int eval_pipe_stage(qtt_ctx* ctx, int n, qtt_net* const* nets, const double* const* xs, const int64_t* npts,
                    double* const* out, int precision, char* scr, BatchStage& st) {
  st.n = 0;
  st.eval_host.clear();
  st.cls.clear();
  std::vector<PipeItem> items;
  std::vector<int> item_of;
  char* d_items = scr;
  char* p = scr + al256((size_t)n * sizeof(PipeItem));
  int* d_cls = (int*)p;
  p += al256((size_t)n * sizeof(int));
  for (int i = 0; i < n; i++) {
    if (npts[i] <= 0) continue;
    qtt_net* net = nets[i];
    if (!pipe_eligible(net, npts[i], precision)) {
      st.eval_host.push_back(i);
      continue;
    }
    const int nv = net->t.nv;
    const TopoView tv = net->view();
    PipeItem it;
    it.pack_cap = pipe_pack_bound(net);
    it.e.pack = (const double*)p;
    p += al256(it.pack_cap * sizeof(double));
    it.e.poff = (const int64_t*)p;
    p += al256(nv * sizeof(int64_t));
    it.e.desc = (const SiteDesc*)p;
    p += al256(4 * nv * sizeof(int));
    it.e.ndigits = tv.ndigits;
    it.e.xs = xs[i];
    it.e.out = out[i];
    it.e.npts = npts[i];
    it.e.block0 = (int64_t)item_of.size();
    it.e.nv = nv;
    it.e.nc = net->t.ncoord;
    it.data = net->d_data;
    it.off = net->d_off;
    it.bond = net->d_bond;
    it.sdim = tv.sdim;
    it.vcoord = tv.vcoord;
    it.vdigit = tv.vdigit;
    it.ndigits = tv.ndigits;
    const int64_t nb = (npts[i] + 255) / 256;
    if (it.e.block0 + nb > INT32_MAX) return qtt_fail(ctx, QTT_EINVAL, "evaluate batch: too many points");
    item_of.insert(item_of.end(), (size_t)nb, (int)items.size());
    items.push_back(it);
    st.cls.push_back(i);  // network index of each item, replaced by classes at readback
  }
  int* d_map = (int*)p;
  st.n = (int)items.size();
  st.total = (int64_t)item_of.size();
  st.dev = d_items;
  st.aux = (char*)d_map;
  st.d_out = (double*)d_cls;
  st.maxnv = 1;
  for (const auto& it : items) st.maxnv = std::max(st.maxnv, it.e.nv);
  if (st.n == 0) return QTT_OK;
  const size_t ib = items.size() * sizeof(PipeItem), mb = item_of.size() * sizeof(int);
  st.host.assign(ib + mb, 0);
  std::memcpy(st.host.data(), items.data(), ib);
  std::memcpy(st.host.data() + ib, item_of.data(), mb);
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_items, st.host.data(), ib, hipMemcpyHostToDevice, ctx->stream), "pipeline items");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d_map, st.host.data() + ib, mb, hipMemcpyHostToDevice, ctx->stream), "pipeline map");
  return rc;
}

int eval_pipe_launch(qtt_ctx* ctx, const BatchStage& st) {
  if (st.n <= 0) return QTT_OK;
  const PipeItem* items = (const PipeItem*)st.dev;
  const int* d_map = (const int*)st.aux;
  int* d_cls = (int*)st.d_out;
  pipe_meta_kernel<<<(st.n + 63) / 64, 64, 0, ctx->stream>>>(items, st.n, d_cls);
  int rc = qtt_hip_check(ctx, hipGetLastError(), "pipe_meta_kernel");
  if (rc) return rc;
  pipe_pack_kernel<<<dim3(8, (unsigned)st.maxnv, (unsigned)std::min(st.n, 65535)), 256, 0, ctx->stream>>>(items, d_cls,
                                                                                                       st.n);
  rc = qtt_hip_check(ctx, hipGetLastError(), "pipe_pack_kernel");
  if (rc) return rc;
  const dim3 grid((unsigned)st.total);
  eval_pipe_kernel<4><<<grid, 256, 0, ctx->stream>>>(items, d_map, d_cls, 0, ctx->flags());
  eval_pipe_kernel<8><<<grid, 256, 0, ctx->stream>>>(items, d_map, d_cls, 1, ctx->flags());
  eval_pipe_kernel<16><<<grid, 256, 0, ctx->stream>>>(items, d_map, d_cls, 2, ctx->flags());
  eval_pipe_kernel<32><<<grid, 256, 0, ctx->stream>>>(items, d_map, d_cls, 3, ctx->flags());
  return qtt_hip_check(ctx, hipGetLastError(), "eval_pipe_kernel");
}

int eval_pipe_readback(qtt_ctx* ctx, BatchStage& st) {
  if (st.n <= 0) return QTT_OK;
  // st.cls holds the network index of each item; the classes go after them
  st.cls.resize(2 * (size_t)st.n);
  return qtt_hip_check(ctx, hipMemcpyAsync(st.cls.data() + st.n, st.d_out, st.n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
                       "pipeline classes");
}

int eval_pipe_rest(qtt_ctx* ctx, const BatchStage& st, int n, qtt_net* const* nets, const double* const* xs,
                   const int64_t* npts, double* const* out, int precision) {
  std::vector<int> rest = st.eval_host;
  for (int j = 0; j < st.n; j++)
    if (st.cls[st.n + j] >= 4) rest.push_back(st.cls[j]);
  for (int i : rest) {
    if (i < 0 || i >= n) return qtt_fail(ctx, QTT_EINVAL, "pipeline: bad network index");
    int rc = launch_eval(ctx, nets[i], xs[i], npts[i], out[i], precision);
    if (rc) return rc;
  }
  return QTT_OK;
}

// ---------------------------------------------------------------------------
// Evaluation without allocation (qtt_eval_noalloc): the packed chain, its
// offsets and its site descriptors are staged in `scratch`, a fixed array of
// the library, instead of the per-network copy that ensure_pack allocates.
// Layout in doubles: [pack: QTT_NOALLOC_PACK_DOUBLES][offsets: one int64 per
// site][descriptors: 4 ints per site], QTT_NOALLOC_MAX_SITES sites.  Every
// call rewrites what it reads, so nothing needs clearing; calls that overlap
// on one GPU race on the array.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[QTT_NOALLOC_PACK_DOUBLES + 3 * QTT_NOALLOC_MAX_SITES];

namespace {

// Device address of `scratch` on each device (hipGetSymbolAddress, taken once
// per device with that device current); host code uses only these.
constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Offsets and descriptors of a packed chain, as ensure_pack computes them on
// the host (one thread: a prefix sum over at most QTT_NOALLOC_MAX_SITES sites).
__global__ void chain_meta_kernel(const int* __restrict__ bond, TopoView t, int64_t* __restrict__ poff,
                                  int* __restrict__ desc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t o = 0;
  for (int k = 0; k < t.nv; ++k) {
    const int chil = bond[k];
    const int chir = (k == t.nv - 1) ? 1 : bond[k + 1];
    const int64_t lpad = (chil + 3) & ~3, rpad = (chir + 3) & ~3;
    poff[k] = o;
    o += rpad * lpad * 2;
    const int c = t.vcoord[k];
    desc[4 * k + 0] = chil;
    desc[4 * k + 1] = chir;
    desc[4 * k + 2] = c;
    desc[4 * k + 3] = c >= 0 ? t.ndigits[c] - t.vdigit[k] : 0;
  }
}

}  // namespace

int qtt_noalloc_init(qtt_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return qtt_fail(ctx, QTT_EINVAL, "no-allocation scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[ctx->device]) return QTT_OK;
  void* p = nullptr;
  int rc = qtt_hip_check(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (rc) return rc;
  g_scratch_dev[ctx->device] = (double*)p;
  return QTT_OK;
}

int launch_eval_noalloc(qtt_ctx* ctx, const qtt_net* net, const double* xs, int64_t npts, double* out) {
  const HostTopo& t = net->t;
  if (!t.chain || t.ncoord > QTT_MAX_COORD)
    return qtt_fail(ctx, QTT_EINVAL, "evaluate (no allocation): chains only");
  if (t.nv > QTT_NOALLOC_MAX_SITES)
    return qtt_fail(ctx, QTT_EINVAL, "evaluate (no allocation): more sites than QTT_NOALLOC_MAX_SITES");
  int64_t maxb = 1, elems = 0;
  for (int k = 0; k < t.nv; k++) {
    maxb = std::max(maxb, net->bond[k]);
    const int64_t chil = net->bond[k], chir = (k == t.nv - 1) ? 1 : net->bond[k + 1];
    elems += ((chir + 3) & ~int64_t(3)) * ((chil + 3) & ~int64_t(3)) * 2;
  }
  if (elems > QTT_NOALLOC_PACK_DOUBLES)
    return qtt_fail(ctx, QTT_EINVAL, "evaluate (no allocation): packed cores exceed QTT_NOALLOC_PACK_DOUBLES");
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch_dev[ctx->device];
  }
  if (!base) return qtt_fail(ctx, QTT_EINVAL, "evaluate (no allocation): scratch not initialised");
  if (npts == 0) return QTT_OK;
  double* pack = base;
  int64_t* poff = (int64_t*)(base + QTT_NOALLOC_PACK_DOUBLES);
  int* desc = (int*)(base + QTT_NOALLOC_PACK_DOUBLES + QTT_NOALLOC_MAX_SITES);
  const TopoView tv = net->view();
  chain_meta_kernel<<<1, 64, 0, ctx->stream>>>(net->d_bond, tv, poff, desc);
  int rc = qtt_hip_check(ctx, hipGetLastError(), "chain_meta_kernel");
  if (rc) return rc;
  pack_chain_kernel<<<dim3(64, t.nv), 256, 0, ctx->stream>>>(net->d_data, net->d_off, net->d_bond, tv.sdim, poff,
                                                             t.nv, pack);
  rc = qtt_hip_check(ctx, hipGetLastError(), "pack_chain_kernel");
  if (rc) return rc;
  if (maxb <= 32) {
    int ppl = maxb <= 16 && npts >= (int64_t(1) << 19) ? 2 : 1, wg = 256;
    const auto it = ctx->eval_tuned.find(eval_tune_key(maxb, npts));
    if (it != ctx->eval_tuned.end()) {
      ppl = it->second.ppl;
      wg = it->second.wg;
    }
    return launch_eval_reg_on(ctx, net, pack, poff, (const SiteDesc*)desc, xs, npts, out, maxb, ppl, wg);
  }
  const size_t per_lane = 2 * size_t(maxb) * sizeof(double);
  int bs = 256;
  while (bs > 64 && per_lane * bs > 160 * 1024) bs /= 2;
  if (per_lane * bs > 160 * 1024)
    return qtt_fail(ctx, QTT_EINVAL, "evaluate (no allocation): bond dimension too large for the chain kernel");
  const int64_t nb = (npts + bs - 1) / bs;
  if (nb > INT32_MAX) return qtt_fail(ctx, QTT_EINVAL, "evaluate: too many points");
  const size_t lds = per_lane * bs;
  if (lds > 64 * 1024) {
    rc = qtt_hip_check(ctx, hipFuncSetAttribute((const void*)eval_chain_f64_kernel,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                       "evaluate LDS attribute");
    if (rc) return rc;
  }
  eval_chain_f64_kernel<<<dim3((unsigned)nb), bs, lds, ctx->stream>>>(pack, poff, net->d_bond, tv, xs, npts, out,
                                                                      ctx->flags(), (int)maxb);
  return qtt_hip_check(ctx, hipGetLastError(), "eval_chain_f64_kernel");
}
