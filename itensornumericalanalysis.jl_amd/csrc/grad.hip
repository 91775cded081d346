// Adjoint of batched evaluation (no reference counterpart; the rrule of
// evaluate): grad_k[l][s][r] = sum_m ybar_m [s == d_k(x_m)] L_k(x_m)[l] R_k(x_m)[r]
// where L_k is the product of the slices left of site k and R_k the product
// right of it.
//
// Register kernel (chains of binary sites, bonds <= 32), tiles of 256 points
// per workgroup, one lane per point:
//   1. backward pass (last site to first) with the vector in registers; every
//      R_k is parked in the workgroup's own global buffer (environments are
//      recomputed per tile, never stored for the whole batch);
//   2. forward pass rebuilds L_k in registers.  At each site the tile's L_k and
//      ybar R_k are staged in LDS and reduced over the points: wavefront w
//      takes site value s = w & 1 and half of the parent-bond rows, walks only
//      the points whose digit is s (a scalar loop over the ballot of digits)
//      and keeps a small block of (l, r) sums per lane; the block is added to
//      the workgroup's own partial gradient (owner-computes, no atomics);
//   3. a second kernel sums the partials in a fixed order, so the result is
//      bitwise reproducible run to run.
// A general kernel (LDS vectors, any bond) covers the other chains.
#include "qtt_internal.h"

namespace {

constexpr int GB = 128;  // general kernel: points per tile = threads per workgroup
constexpr int GT = 256;  // register kernel: points per tile = threads per workgroup

__device__ inline uint64_t digit_word_g(double x, int L, bool& bad) {
  if (!(x >= 0.0)) {
    bad = true;
    return 0;
  }
  if (x >= 1.0) return (uint64_t(1) << L) - 1;
  return (uint64_t)(x * (double)(uint64_t(1) << L));
}

// F: the arithmetic of the environments and of the per-tile sums (double, or
// float for the float32 gradient); S: the stride of a pack entry in F units
// (1 for the float64 pack, 2 for the float (hi, lo) pack, whose hi is read).
// The partial gradients accumulate in float64 across tiles either way.
template <typename F, int S>
__global__ void __launch_bounds__(GB)
grad_chain_kernel(const F* __restrict__ pack, const int64_t* __restrict__ poff,
                  const int* __restrict__ bond, TopoView t, const int64_t* __restrict__ goff,
                  const int64_t* __restrict__ roff, int64_t rtot, const double* __restrict__ xs,
                  const double* __restrict__ ybar, int64_t npts, int64_t gsize, F* Rscr,
                  double* gpart, int* __restrict__ flag, int vstride) {
  extern __shared__ double lds_raw[];
  F* lds = (F*)lds_raw;
  const int tid = threadIdx.x;
  // LDS: Lv[vstride][GB], Nv[vstride][GB], Rv[vstride][GB], digit per lane
  F* Lv = lds;
  F* Nv = lds + (size_t)vstride * GB;
  F* Rv = lds + (size_t)2 * vstride * GB;
  int* dig = (int*)(lds + (size_t)3 * vstride * GB);
  F* R = Rscr + (size_t)blockIdx.x * rtot * GB;  // [roff[k] + r][lane]
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
    const F yb = (F)((active && !bad) ? ybar[p] : 0.0);
    // ---- backward: R_in(k) = vector over the child bond of site k -------
    F* v = Lv + tid;
    F* w = Nv + tid;
    v[0] = (F)1;
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
      const F* __restrict__ P = pack + (size_t)S * poff[k];
      const int lpad = (chil + 3) & ~3;
      for (int l0 = 0; l0 < chil; l0 += 4) {
        F a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0};
        const F* __restrict__ q = P + 2 * S * l0;
        for (int r = 0; r < chir; ++r) {
          const F vr = v[(size_t)r * GB];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            a0[i] = __builtin_fma(q[(2 * i) * S], vr, a0[i]);
            a1[i] = __builtin_fma(q[(2 * i + 1) * S], vr, a1[i]);
          }
          q += 2 * S * lpad;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (l0 + i < chil) w[(size_t)(l0 + i) * GB] = d ? a1[i] : a0[i];
      }
      F* tmp = v;
      v = w;
      w = tmp;
    }
    // ---- forward: accumulate outer products, advance L ------------------
    Lv[tid] = (F)1;  // L_0 over the root's parent bond (dimension 1)
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
        F acc = 0;
        const F* lrow = Lv + (size_t)l * GB;
        const F* rrow = Rv + (size_t)r * GB;
        for (int j = 0; j < GB; ++j) acc += (dig[j] == s) ? lrow[j] * rrow[j] : (F)0;
        Gk[e] += (double)acc;
      }
      // L_{k+1}[r] = sum_l L_k[l] T[l][d][r]
      const F* __restrict__ P = pack + (size_t)S * poff[k];
      const int lpad = (chil + 3) & ~3;
      for (int r = 0; r < chir; ++r) {
        F s0 = 0, s1 = 0;
        const F* __restrict__ q = P + (size_t)r * 2 * S * lpad;
        for (int l = 0; l < chil; ++l) {
          const F lv = Lv[(size_t)l * GB + tid];
          s0 = __builtin_fma(q[(2 * l) * S], lv, s0);
          s1 = __builtin_fma(q[(2 * l + 1) * S], lv, s1);
        }
        Nv[(size_t)r * GB + tid] = d ? s1 : s0;
      }
      __syncthreads();
      for (int r = 0; r < chir; ++r) Lv[(size_t)r * GB + tid] = Nv[(size_t)r * GB + tid];
      __syncthreads();
    }
  }
}

__device__ inline double bsel(double a, double b, uint64_t mask) {
  const uint64_t ua = (uint64_t)__double_as_longlong(a), ub = (uint64_t)__double_as_longlong(b);
  return __longlong_as_double((long long)((ua & ~mask) | (ub & mask)));
}

__device__ inline uint64_t uniform64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint64_t site_digit_mask(const SiteDesc& sd, const uint64_t (&kw)[QTT_MAX_COORD]) {
  if (sd.coord < 0) return 0;
  const uint64_t kk = sd.coord == 0 ? kw[0] : sd.coord == 1 ? kw[1] : sd.coord == 2 ? kw[2] : kw[3];
  return 0 - ((kk >> sd.shift) & 1u);
}

// v <- T[:, d, :] v over the child bond (packed site P[r][lpad][2]); entries
// past the parent bond are left at zero.
template <int CHI>
__device__ __forceinline__ void step_to_parent(const double* __restrict__ P, const SiteDesc& sd, uint64_t dmask,
                                               double (&v)[CHI]) {
  const int lpad = (sd.chil + 3) & ~3;
  double w[CHI];
#pragma unroll
  for (int l0 = 0; l0 < CHI; l0 += 4) {
    if (l0 < sd.chil) {
      double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
      // rows in pairs, the next pair's scalar loads issued before this pair's
      // multiply-adds (the packed core has rows up to the bond rounded up to 4,
      // zero past the bond)
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
        const double vu = v[r], vw = v[r + 1];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a0[i] = __builtin_fma(qa[2 * i], vu, a0[i]);
          a1[i] = __builtin_fma(qa[2 * i + 1], vu, a1[i]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a0[i] = __builtin_fma(qb[2 * i], vw, a0[i]);
          a1[i] = __builtin_fma(qb[2 * i + 1], vw, a1[i]);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          qa[e] = na[e];
          qb[e] = nb[e];
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) w[l0 + i] = bsel(a0[i], a1[i], dmask);
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) w[l0 + i] = 0.0;
    }
  }
#pragma unroll
  for (int i = 0; i < CHI; ++i) v[i] = w[i];
}

// L <- L T[:, d, :] over the parent bond; entries past the child bond stay zero.
template <int CHI>
__device__ __forceinline__ void step_to_child(const double* __restrict__ P, const SiteDesc& sd, uint64_t dmask,
                                              double (&L)[CHI]) {
  const int lpad = (sd.chil + 3) & ~3;
  double o[CHI];
#pragma unroll
  for (int r0 = 0; r0 < CHI; r0 += 4) {
    if (r0 < sd.chir) {
      double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int l0 = 0; l0 < CHI; l0 += 4) {
        if (l0 == 0 || l0 < sd.chil) {  // at least one parent-bond entry: the first block always runs
          // i outer, rr inner: consecutive multiply-adds go to 8 different
          // accumulators (r, digit), so none waits on the one before
#pragma unroll
          for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
              const double* __restrict__ q = P + ((r0 + rr) * lpad + l0) * 2;
              a0[rr] = __builtin_fma(q[2 * i], L[l0 + i], a0[rr]);
              a1[rr] = __builtin_fma(q[2 * i + 1], L[l0 + i], a1[rr]);
            }
          }
        }
      }
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) o[r0 + rr] = bsel(a0[rr], a1[rr], dmask);
    } else {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) o[r0 + rr] = 0.0;
    }
  }
#pragma unroll
  for (int i = 0; i < CHI; ++i) L[i] = o[i];
}

// Reduction block of one lane: BL parent-bond rows x BR child-bond columns.
template <int CHI>
struct RedBlock {
  static constexpr int BR = CHI >= 16 ? 2 : 1;
  static constexpr int BL = CHI >= 32 ? 4 : 1;
  static constexpr int NRB = CHI / BR;  // lanes along the child bond
};

template <int CHI>
__global__ void __launch_bounds__(GT)
grad_chain_reg_kernel(const double* __restrict__ pack, const int64_t* __restrict__ poff,
                      const SiteDesc* __restrict__ desc, const int* __restrict__ sdim, int nv, int nc,
                      const int* __restrict__ ndigits, const int64_t* __restrict__ goff,
                      const int64_t* __restrict__ roff, int64_t rtot, const double* __restrict__ xs,
                      const double* __restrict__ ybar, int64_t npts, int64_t gsize, double* __restrict__ Rscr,
                      double* __restrict__ gpart, int* __restrict__ flag) {
  // The tile's points are staged half at a time (HP = 128 points, two
  // wavefronts), which halves the LDS footprint so that two workgroups share a
  // compute unit.
  constexpr int HP = GT / 2;
  constexpr int LS = HP + 1;  // odd row stride: the rows read together sit in different banks
  constexpr int BL = RedBlock<CHI>::BL, BR = RedBlock<CHI>::BR, NRB = RedBlock<CHI>::NRB;
  extern __shared__ double lds[];
  double* Ls = lds;                                    // [CHI][LS]  L_k per point of the half
  double* Rs = lds + CHI * LS;                         // [CHI][LS]  ybar R_k per point of the half
  uint64_t* bits = (uint64_t*)(lds + 2 * CHI * LS);    // [0, 2): digit ballots, [2, 4): active ballots
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int s_w = wv & 1;
  const int lb = lane / NRB, rb = lane % NRB;
  const bool lane_on = lb * BL < CHI / 2;
  const int lrow0 = lane_on ? (wv >> 1) * (CHI / 2) + lb * BL : 0;
  const int rcol0 = rb * BR;
  double* R = Rscr + (size_t)blockIdx.x * rtot * GT;
  double* G = gpart + (size_t)blockIdx.x * gsize;
  for (int64_t e = tid; e < gsize; e += GT) G[e] = 0.0;
  if (tid < CHI) {  // column HP is a zero point, paired with the last point of an odd count
    Ls[tid * LS + HP] = 0.0;
    Rs[tid * LS + HP] = 0.0;
  }
  const int64_t ntiles = (npts + GT - 1) / GT;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t p = tile * GT + tid;
    const bool active = p < npts;
    bool bad = false;
    uint64_t kw[QTT_MAX_COORD] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < QTT_MAX_COORD; ++c)
      if (c < nc) kw[c] = digit_word_g(active ? xs[p * nc + c] : 0.0, ndigits[c], bad);
    if (bad) flag[0] = 1;
    const double yb = (active && !bad) ? ybar[p] : 0.0;
    // ---- backward: R_k over the child bond of site k, parked per tile ----
    double v[CHI];
#pragma unroll
    for (int i = 0; i < CHI; ++i) v[i] = 0.0;
    v[0] = 1.0;
    for (int k = nv - 1; k >= 0; --k) {
      const SiteDesc sd = desc[k];
      double* Rk = R + roff[k] * GT + tid;
#pragma unroll
      for (int r = 0; r < CHI; ++r)
        if (r < sd.chir) Rk[(int64_t)r * GT] = v[r];
      if (k == 0) break;
      step_to_parent<CHI>(pack + poff[k], sd, site_digit_mask(sd, kw), v);
    }
    // ---- forward: stage, reduce, advance L ----
    double L[CHI];
#pragma unroll
    for (int i = 0; i < CHI; ++i) L[i] = 0.0;
    L[0] = 1.0;
    for (int k = 0; k < nv; ++k) {
      const SiteDesc sd = desc[k];
      const uint64_t dmask = site_digit_mask(sd, kw);
      const double* Rk = R + roff[k] * GT + tid;
      double rv[CHI];
#pragma unroll
      for (int i = 0; i < CHI; ++i) rv[i] = i < sd.chir ? yb * Rk[(int64_t)i * GT] : 0.0;
      const uint64_t db = __ballot(dmask != 0 && active);
      const uint64_t ab = __ballot(active);
      const int sk = sdim[k];
      double acc[BL][BR];
#pragma unroll
      for (int i = 0; i < BL; ++i)
#pragma unroll
        for (int q = 0; q < BR; ++q) acc[i][q] = 0.0;
      for (int h = 0; h < 2; ++h) {
        if ((wv >> 1) == h) {  // this half's two wavefronts stage their points
          const int col = tid & (HP - 1);
#pragma unroll
          for (int i = 0; i < CHI; ++i) {
            Ls[i * LS + col] = L[i];
            Rs[i * LS + col] = rv[i];
          }
          if (lane == 0) {
            bits[wv & 1] = db;
            bits[2 + (wv & 1)] = ab;
          }
        }
        __syncthreads();
        if (s_w < sk) {
          for (int c = 0; c < HP / 64; ++c) {
            uint64_t m = bits[2 + c];
            if (sk == 2) m = s_w ? (m & bits[c]) : (m & ~bits[c]);
            m = uniform64(m);
            while (m) {
              const int ja = c * 64 + __builtin_ctzll(m);
              m &= m - 1;
              int jb = HP;
              if (m) {
                jb = c * 64 + __builtin_ctzll(m);
                m &= m - 1;
              }
              double la[BL], lq[BL], ra[BR], rq[BR];
#pragma unroll
              for (int i = 0; i < BL; ++i) {
                la[i] = Ls[(lrow0 + i) * LS + ja];
                lq[i] = Ls[(lrow0 + i) * LS + jb];
              }
#pragma unroll
              for (int q = 0; q < BR; ++q) {
                ra[q] = Rs[(rcol0 + q) * LS + ja];
                rq[q] = Rs[(rcol0 + q) * LS + jb];
              }
#pragma unroll
              for (int i = 0; i < BL; ++i)
#pragma unroll
                for (int q = 0; q < BR; ++q) {
                  acc[i][q] = __builtin_fma(la[i], ra[q], acc[i][q]);
                  acc[i][q] = __builtin_fma(lq[i], rq[q], acc[i][q]);
                }
            }
          }
        }
        __syncthreads();
      }
      if (s_w < sk && lane_on) {
        double* Gk = G + goff[k];
#pragma unroll
        for (int i = 0; i < BL; ++i)
#pragma unroll
          for (int q = 0; q < BR; ++q) {
            const int l = lrow0 + i, r = rcol0 + q;
            if (l < sd.chil && r < sd.chir) Gk[((int64_t)l * sk + s_w) * sd.chir + r] += acc[i][q];
          }
      }
      if (k + 1 < nv) step_to_child<CHI>(pack + poff[k], sd, dmask, L);
    }
  }
}

// ---------------------------------------------------------------------------
// Gradient on a general tree (float64), tiles of GTR points per workgroup.
//   phase A, a lane per point: the up pass gives every vertex's subtree vector
//     u_v over bond[v] (children before parents: vertices in decreasing id);
//     the down pass gives every vertex's environment e_v over bond[v] (the
//     rest of the network contracted, parents first).  Both go to the
//     workgroup's scratch as [element][lane], with the digit of every vertex;
//   phase B, a lane per gradient entry (v, p, s, c_1 .. c_m):
//     G_v[p][s][c] += sum over the tile's points with digit s at v of
//     ybar e_v[p] prod_i u_{ch_i}[c_i]  (owner-computes, no atomics).
// The partials are summed by grad_reduce_kernel in a fixed order.
// ---------------------------------------------------------------------------
constexpr int GTR = 64;

__global__ void __launch_bounds__(256)
grad_tree_kernel(const double* __restrict__ data, const int64_t* __restrict__ off, const int* __restrict__ bond,
                 TopoView t, const int64_t* __restrict__ voff, int64_t vtot, const int64_t* __restrict__ goff,
                 const double* __restrict__ xs, const double* __restrict__ ybar, int64_t npts, int64_t gsize,
                 double* __restrict__ scr, double* __restrict__ gpart, int* __restrict__ flag) {
  const int tid = threadIdx.x, bs = blockDim.x;
  // scratch of this workgroup: U [vtot][GTR] | E [vtot][GTR] | ybar [GTR] | digits [nv][GTR]
  double* U = scr + (size_t)blockIdx.x * (2 * vtot + 1 + t.nv) * GTR;
  double* E = U + (size_t)vtot * GTR;
  double* Y = E + (size_t)vtot * GTR;
  int* D = (int*)(Y + GTR);
  double* G = gpart + (size_t)blockIdx.x * gsize;
  for (int64_t e = tid; e < gsize; e += bs) G[e] = 0.0;
  const int nc = t.ncoord;
  const int64_t ntiles = (npts + GTR - 1) / GTR;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    __syncthreads();
    if (tid < GTR) {
      const int lane = tid;
      const int64_t p = tile * GTR + lane;
      const bool active = p < npts;
      bool bad = false;
      for (int c = 0; c < nc; ++c) (void)digit_word_g(active ? xs[p * nc + c] : 0.0, t.ndigits[c], bad);
      if (bad) flag[0] = 1;
      Y[lane] = (active && !bad) ? ybar[p] : 0.0;
      for (int v = 0; v < t.nv; ++v) {  // any number of coordinates
        const int c = t.vcoord[v];
        int d = 0;
        if (c >= 0 && active && !bad) {
          bool b2 = false;
          const uint64_t kk = digit_word_g(xs[p * nc + c], t.ndigits[c], b2);
          d = (int)((kk >> (t.ndigits[c] - t.vdigit[v])) & 1u);
        }
        D[(size_t)v * GTR + lane] = d;
      }
      // up pass
      for (int v = t.nv - 1; v >= 0; --v) {
        const int m = t.nch[v];
        const int* ch = t.chs + t.choff[v];
        int64_t pc = 1;
        for (int k = 0; k < m; ++k) pc *= bond[ch[k]];
        const int sd = t.sdim[v];
        const int d = D[(size_t)v * GTR + lane];
        for (int pp = 0; pp < bond[v]; ++pp) {
          const double* base = data + off[v] + ((int64_t)pp * sd + d) * pc;
          double acc = 0.0;
          for (int64_t j = 0; j < pc; ++j) {
            double w = base[j];
            int64_t r = j;
            for (int k = m - 1; k >= 0; --k) {
              const int b = bond[ch[k]];
              w *= U[(voff[ch[k]] + r % b) * GTR + lane];
              r /= b;
            }
            acc += w;
          }
          U[(voff[v] + pp) * GTR + lane] = acc;
        }
      }
      // down pass
      E[voff[0] * GTR + lane] = 1.0;
      for (int v = 0; v < t.nv; ++v) {
        const int m = t.nch[v];
        const int* ch = t.chs + t.choff[v];
        int64_t pc = 1;
        for (int k = 0; k < m; ++k) pc *= bond[ch[k]];
        const int sd = t.sdim[v];
        const int d = D[(size_t)v * GTR + lane];
        for (int kc = 0; kc < m; ++kc) {
          const int c = ch[kc];
          for (int j = 0; j < bond[c]; ++j) E[(voff[c] + j) * GTR + lane] = 0.0;
          for (int pp = 0; pp < bond[v]; ++pp) {
            const double ev = E[(voff[v] + pp) * GTR + lane];
            const double* base = data + off[v] + ((int64_t)pp * sd + d) * pc;
            for (int64_t j = 0; j < pc; ++j) {
              double w = base[j] * ev;
              int64_t r = j;
              int jc = 0;
              for (int k = m - 1; k >= 0; --k) {
                const int b = bond[ch[k]];
                const int idx = (int)(r % b);
                r /= b;
                if (k == kc) jc = idx;
                else w *= U[(voff[ch[k]] + idx) * GTR + lane];
              }
              E[(voff[c] + jc) * GTR + lane] += w;
            }
          }
        }
      }
    }
    __syncthreads();
    // phase B
    for (int v = 0; v < t.nv; ++v) {
      const int m = t.nch[v];
      const int* ch = t.chs + t.choff[v];
      int64_t pc = 1;
      for (int k = 0; k < m; ++k) pc *= bond[ch[k]];
      const int sd = t.sdim[v];
      const int64_t ne = (int64_t)bond[v] * sd * pc;
      for (int64_t e = tid; e < ne; e += bs) {
        const int64_t j = e % pc;
        const int s = (int)((e / pc) % sd);
        const int pp = (int)(e / (pc * sd));
        double acc = 0.0;
        for (int lane = 0; lane < GTR; ++lane) {
          if (D[(size_t)v * GTR + lane] != s) continue;
          double w = Y[lane] * E[(voff[v] + pp) * GTR + lane];
          int64_t r = j;
          for (int k = m - 1; k >= 0; --k) {
            const int b = bond[ch[k]];
            w *= U[(voff[ch[k]] + r % b) * GTR + lane];
            r /= b;
          }
          acc += w;
        }
        G[goff[v] + e] += acc;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Pullback of the exact operator contraction g = W f (contract_kernel):
//   g_v[(pw, pa)][so][(cw_1, ca_1) ...] = sum_si W_v[pw][si][so][cw..] f_v[pa][si][ca..]
// so, for a cotangent gbar on g's cores,
//   fbar_v[pa][si][ca..] = sum_{pw, so, cw..} W_v[pw][si][so][cw..] gbar_v[(pw, pa)][so][(cw, ca)..]
//   wbar_v[pw][si][so][cw..] = sum_{pa, ca..} f_v[pa][si][ca..] gbar_v[(pw, pa)][so][(cw, ca)..]
// One thread per output entry; a combined child index is cw * bondA + ca.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t comb_offset(const int* ch, int m, const int* bw, const int* ba, int64_t jw,
                                               int64_t ja) {
  // offset in g's child block of the combined indices (cw_k * ba_k + ca_k)
  int64_t o = 0, S = 1;
  for (int k = m - 1; k >= 0; --k) {
    const int64_t bwk = bw[ch[k]], bak = ba[ch[k]];
    const int64_t cw = jw % bwk, ca = ja % bak;
    jw /= bwk;
    ja /= bak;
    o += (cw * bak + ca) * S;
    S *= bwk * bak;
  }
  return o;
}

__global__ void __launch_bounds__(256)
contract_adjoint_kernel(const double* __restrict__ W, const int64_t* __restrict__ woff, const int* __restrict__ bw,
                        const double* __restrict__ A, const int64_t* __restrict__ aoff, const int* __restrict__ ba,
                        TopoView t, const double* __restrict__ gbar, const int64_t* __restrict__ gstart,
                        double* __restrict__ fbar, const int64_t* __restrict__ fstart, double* __restrict__ wbar,
                        const int64_t* __restrict__ wstart) {
  // fbar entries first, then wbar entries (if wanted), over one flat index
  const int64_t nf = fstart[t.nv], nw = wbar ? wstart[t.nv] : 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nf + nw;
       e += (int64_t)gridDim.x * blockDim.x) {
    const bool isf = e < nf;
    const int64_t ee = isf ? e : e - nf;
    const int64_t* st = isf ? fstart : wstart;
    int v = 0;
    while (st[v + 1] <= ee) ++v;
    const int64_t loc = ee - st[v];
    const int m = t.nch[v];
    const int* ch = t.chs + t.choff[v];
    const int sd = t.sdim[v];
    int64_t pcw = 1, pca = 1;
    for (int k = 0; k < m; ++k) {
      pcw *= bw[ch[k]];
      pca *= ba[ch[k]];
    }
    const int64_t Ap = ba[v], Dp = bw[v];
    const int64_t pco = pcw * pca;
    const double* Wv = W + woff[v];
    const double* Av = A + aoff[v];
    const double* Gv = gbar + gstart[v];
    double s = 0.0;
    if (isf) {  // loc = (pa * sd + si) * pca + ja
      const int64_t ja = loc % pca;
      const int si = (int)((loc / pca) % sd);
      const int64_t pa = loc / (pca * sd);
      for (int64_t pw = 0; pw < Dp; ++pw)
        for (int so = 0; so < sd; ++so) {
          const double* wrow = Wv + ((pw * sd + si) * sd + so) * pcw;
          const double* grow = Gv + ((pw * Ap + pa) * sd + so) * pco;
          for (int64_t jw = 0; jw < pcw; ++jw) {
            const double wv = wrow[jw];
            if (wv != 0.0) s = __builtin_fma(wv, grow[comb_offset(ch, m, bw, ba, jw, ja)], s);
          }
        }
      fbar[e] = s;
    } else {  // loc = ((pw * sd + si) * sd + so) * pcw + jw
      const int64_t jw = loc % pcw;
      const int so = (int)((loc / pcw) % sd);
      const int si = (int)((loc / (pcw * sd)) % sd);
      const int64_t pw = loc / (pcw * sd * sd);
      for (int64_t pa = 0; pa < Ap; ++pa) {
        const double* arow = Av + (pa * sd + si) * pca;
        const double* grow = Gv + ((pw * Ap + pa) * sd + so) * pco;
        for (int64_t ja = 0; ja < pca; ++ja) s = __builtin_fma(arow[ja], grow[comb_offset(ch, m, bw, ba, jw, ja)], s);
      }
      wbar[ee] = s;
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
int qtt_ensure_pack32(qtt_ctx* ctx, qtt_net* net);

// The tree kernel (float64, any tree): tiles of GTR points, partials summed
// in a fixed order.
static int launch_grad_tree(qtt_ctx* ctx, qtt_net* net, const double* xs, const double* ybar, int64_t npts,
                            double* grad) {
  const HostTopo& t = net->t;
  const int nv = t.nv;
  std::vector<int64_t> voff(nv), goff(nv);
  int64_t vtot = 0, gsize = 0;
  for (int v = 0; v < nv; v++) {
    voff[v] = vtot;
    vtot += net->bond[v];
    goff[v] = gsize;
    gsize += net->core_size(v);
  }
  const int64_t ntiles = (npts + GTR - 1) / GTR;
  int nwg = (int)std::max<int64_t>(1, std::min<int64_t>(ntiles, 2 * (int64_t)ctx->cus));
  const size_t per_wg = (size_t)(2 * vtot + 1 + nv) * GTR;
  const size_t bytes = ((size_t)nwg * (per_wg + gsize) + 2 * nv + 16) * sizeof(double);
  double* scr = qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "gradient scratch");
  double* wscr = scr;
  double* gpart = wscr + (size_t)nwg * per_wg;
  int64_t* d_voff = (int64_t*)(gpart + (size_t)nwg * gsize);
  int64_t* d_goff = d_voff + nv;
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_voff, voff.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "voff");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d_goff, goff.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "goff");
  if (rc) return rc;
  grad_tree_kernel<<<nwg, 256, 0, ctx->stream>>>(net->d_data, net->d_off, net->d_bond, net->view(), d_voff, vtot,
                                                 d_goff, xs, ybar, npts, gsize, wscr, gpart, ctx->flags());
  rc = qtt_hip_check(ctx, hipGetLastError(), "grad_tree_kernel");
  if (rc) return rc;
  int64_t nb = std::min<int64_t>((gsize + 255) / 256, 4096);
  grad_reduce_kernel<<<(unsigned)std::max<int64_t>(nb, 1), 256, 0, ctx->stream>>>(gpart, nwg, gsize, grad);
  rc = qtt_hip_check(ctx, hipGetLastError(), "grad_reduce_kernel");
  if (rc) return rc;
  // the host offset vectors must outlive their asynchronous copies
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "grad sync");
}

int launch_apply_grad(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, const double* gbar, double* fbar,
                      double* wbar) {
  const HostTopo& t = net->t;
  const int nv = t.nv;
  std::vector<int64_t> tab(3 * (nv + 1), 0);  // fstart | wstart | gstart
  int64_t* fs = tab.data();
  int64_t* ws = fs + nv + 1;
  int64_t* gs = ws + nv + 1;
  for (int v = 0; v < nv; v++) {
    int64_t pa = 1, pw = 1, pg = 1;
    for (int k = 0; k < t.nch[v]; k++) {
      const int c = t.chs[t.choff[v] + k];
      pa *= net->bond[c];
      pw *= op->bond[c];
      pg *= net->bond[c] * op->bond[c];
    }
    const int64_t sd = t.sdim[v];
    fs[v + 1] = fs[v] + net->bond[v] * sd * pa;
    ws[v + 1] = ws[v] + op->bond[v] * sd * sd * pw;
    gs[v + 1] = gs[v] + op->bond[v] * net->bond[v] * sd * pg;
  }
  double* scr = qtt_scratch(ctx, tab.size() * sizeof(int64_t) + 64);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "apply gradient scratch");
  int64_t* d = (int64_t*)scr;
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d, tab.data(), tab.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "offsets");
  if (rc) return rc;
  const int64_t total = fs[nv] + (wbar ? ws[nv] : 0);
  const int64_t nb = std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
  contract_adjoint_kernel<<<(unsigned)nb, 256, 0, ctx->stream>>>(op->d_data, op->d_off, op->d_bond, net->d_data,
                                                                 net->d_off, net->d_bond, net->view(), gbar,
                                                                 d + 2 * (nv + 1), fbar, d, wbar, d + (nv + 1));
  rc = qtt_hip_check(ctx, hipGetLastError(), "contract_adjoint_kernel");
  if (rc) return rc;
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "apply gradient sync");
}

int launch_grad(qtt_ctx* ctx, const qtt_net* cnet, const double* xs, const double* ybar,
                int64_t npts, double* grad, int precision) {
  qtt_net* net = const_cast<qtt_net*>(cnet);
  const HostTopo& t = net->t;
  if (precision != QTT_F64 && precision != QTT_F32)
    return qtt_fail(ctx, QTT_EINVAL, "gradient: QTT_F64 or QTT_F32");
  if (!t.chain || t.ncoord > QTT_MAX_COORD) {
    if (precision != QTT_F64) return qtt_fail(ctx, QTT_EINVAL, "gradient: float32 is implemented for chains");
    return launch_grad_tree(ctx, net, xs, ybar, npts, grad);
  }
  const bool f32 = precision == QTT_F32;
  int rc = f32 ? qtt_ensure_pack32(ctx, net) : qtt_ensure_pack(ctx, net);
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
  bool binary = true;
  for (int k = 0; k < nv; k++) binary = binary && t.sdim[k] <= 2;
  const int chi = maxb <= 8 ? 8 : maxb <= 16 ? 16 : 32;
  const bool reg = binary && maxb <= 32 && !f32;
  const int tilep = reg ? GT : GB;
  const size_t fsz = f32 ? sizeof(float) : sizeof(double);
  const size_t lds = reg ? ((size_t)2 * chi * (GT / 2 + 1) + 4) * sizeof(double)
                         : (size_t)3 * maxb * GB * fsz + GB * sizeof(int);
  if (lds > 160 * 1024) return qtt_fail(ctx, QTT_EINVAL, "gradient: bond dimension too large");
  const int64_t ntiles = (npts + tilep - 1) / tilep;
  const int per_cu = reg ? std::max(1, std::min(4, (int)((160 * 1024) / lds))) : 2;
  int nwg = (int)std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)per_cu * ctx->cus));
  size_t bytes = ((size_t)nwg * (rtot * tilep + gsize) + 2 * nv + 16) * sizeof(double);
  double* scr = qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "gradient scratch");
  double* Rscr = scr;
  double* gpart = Rscr + (size_t)nwg * rtot * tilep;  // each workgroup parks rtot x tilep values
  int64_t* d_goff = (int64_t*)(gpart + (size_t)nwg * gsize);
  int64_t* d_roff = d_goff + nv;
  rc = qtt_hip_check(ctx, hipMemcpyAsync(d_goff, goff.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "goff");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(d_roff, roff.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "roff");
  if (rc) return rc;
  const TopoView tv = net->view();
  const void* fn = !reg ? (f32 ? (const void*)grad_chain_kernel<float, 2> : (const void*)grad_chain_kernel<double, 1>)
                   : chi == 8 ? (const void*)grad_chain_reg_kernel<8>
                   : chi == 16 ? (const void*)grad_chain_reg_kernel<16>
                               : (const void*)grad_chain_reg_kernel<32>;
  if (lds > 64 * 1024) {
    rc = qtt_hip_check(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                       "grad LDS attribute");
    if (rc) return rc;
  }
  if (npts > 0) {
    if (reg) {
      const SiteDesc* dd = (const SiteDesc*)net->d_desc;
#define QTT_GRAD_LAUNCH(C)                                                                                   \
  grad_chain_reg_kernel<C><<<nwg, GT, lds, ctx->stream>>>(net->d_pack, net->d_pack_off, dd, tv.sdim, nv,     \
                                                          tv.ncoord, tv.ndigits, d_goff, d_roff, rtot, xs, ybar, \
                                                          npts, gsize, Rscr, gpart, ctx->flags())
      if (chi == 8) QTT_GRAD_LAUNCH(8);
      else if (chi == 16) QTT_GRAD_LAUNCH(16);
      else QTT_GRAD_LAUNCH(32);
#undef QTT_GRAD_LAUNCH
    } else {
      if (f32)  // the float (hi, lo) pack, hi read; float vectors parked in the scratch
        grad_chain_kernel<float, 2><<<nwg, GB, lds, ctx->stream>>>(
            (const float*)net->d_pack32, net->d_pack_off, net->d_bond, tv, d_goff, d_roff, rtot, xs, ybar, npts,
            gsize, (float*)Rscr, gpart, ctx->flags(), (int)maxb);
      else
        grad_chain_kernel<double, 1><<<nwg, GB, lds, ctx->stream>>>(net->d_pack, net->d_pack_off, net->d_bond, tv,
                                                                    d_goff, d_roff, rtot, xs, ybar, npts, gsize, Rscr,
                                                                    gpart, ctx->flags(), (int)maxb);
    }
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
