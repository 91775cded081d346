// operate + truncate on device.
//
// contract_kernel: exact operator application (src/elementary_operators.jl:294-299),
//   C[(w,a)][s_out][(cw,ca)...] = sum_{s_in} W[w][s_in][s_out][cw...] A[a][s_in][ca...],
//   all vertices in one launch, one thread per output element.
//
// truncate_kernel: the truncation sweep of ITensorNetworks.truncate(::TTN)
//   (called from src/itensornetworkfunction.jl:114-118), run by one workgroup
//   per network over a host-built schedule of moves; a batch of networks is one
//   launch with one workgroup each.  A move (x -> u) views core x as a matrix
//   (all other indices) x (bond to u) and
//     * QR move:  Householder QR, x <- Q, u <- R (x)_bond u;
//     * SVD move: Householder QR if the matrix is tall, then one-sided Jacobi
//       on the rows of the (square or wide) factor with the rotations kept in J,
//       singular values sorted, the ITensors cutoff rule applied, x <- U (an
//       exact product of orthogonal factors), u <- S V^T (x)_bond u.  A tall
//       or square move that can drop nothing (cutoff <= 0, maxdim >= its
//       rank, R nonsingular) keeps x <- Q, u <- R (x)_bond u instead.
//   The matrix lives in LDS when it fits (158 KiB), otherwise in a global
//   scratch area; all arithmetic is FP64.
//   Work is spread over lane groups of G lanes (G = 8 for matrices of up to
//   256 rows, a whole wavefront above): one group per column in the QR, one
//   group per row pair in each Jacobi round (round-robin pairing, all pairs of
//   a round independent).  Group sums use DPP lane permutes, not LDS.
#include <cfloat>
#include <cstring>

#include "qtt_internal.h"

// Phase timestamps of the sweep for profiling builds (-DQTT_PHASE_TIMING; step
// -1 = the step opened last, read back with qtt_debug_times); empty otherwise.
#ifdef QTT_PHASE_TIMING
__device__ unsigned long long qtt_phase_t[8192 * 8];
__device__ int qtt_phase_step;
__device__ inline void qtt_phase_mark(int step, int phase, unsigned long long v) {
  if (step >= 0) qtt_phase_step = step;
  if (qtt_phase_step < 8192) qtt_phase_t[qtt_phase_step * 8 + phase] = v;
}
extern "C" int qtt_debug_times(unsigned long long* out, int n) {
  return (int)hipMemcpyFromSymbol(out, HIP_SYMBOL(qtt_phase_t), (size_t)n * 8 * sizeof(unsigned long long));
}
#define QTT_PHASE(step, phase) \
  do { if (threadIdx.x == 0) qtt_phase_mark((step), (phase), wall_clock64()); } while (0)
// the move's shape in the two spare slots of its record
#define QTT_SHAPE(step, rows, b, tr) \
  do { if (threadIdx.x == 0 && (step) < 8192) { qtt_phase_t[(step) * 8 + 6] = (rows); qtt_phase_t[(step) * 8 + 7] = (b) | ((tr) << 20); } } while (0)
// Jacobi sweeps (low 16 bits) and calls (high bits) of the step opened last
__device__ unsigned long long qtt_phase_sweeps[8192];
extern "C" int qtt_debug_sweeps(unsigned long long* out, int n) {
  return (int)hipMemcpyFromSymbol(out, HIP_SYMBOL(qtt_phase_sweeps), (size_t)n * sizeof(unsigned long long));
}
#define QTT_SWEEPS(s) \
  do { if (threadIdx.x == 0 && qtt_phase_step < 8192) qtt_phase_sweeps[qtt_phase_step] += (unsigned long long)(s) + (1ull << 16); } while (0)
#define QTT_SWEEPS_CLEAR(step) do { if (threadIdx.x == 0 && (step) < 8192) qtt_phase_sweeps[step] = 0; } while (0)
#endif
// Cycle counters inside the Jacobi rounds and QR panels (-DQTT_JACOBI_TIMING,
// on top of the phase timestamps); each mark is a read-modify-write of global
// memory by lane 0, so they slow the rounds they measure.
#ifdef QTT_JACOBI_TIMING
// cycles of lane 0 per segment of a Jacobi round, plus round / sweep counts
__device__ unsigned long long qtt_jac_t[8];
extern "C" int qtt_debug_jacobi(unsigned long long* out) {
  return (int)hipMemcpyFromSymbol(out, HIP_SYMBOL(qtt_jac_t), 8 * sizeof(unsigned long long));
}
#define QTT_JT_INIT unsigned long long jt_last_ = clock64()
#define QTT_JT(i)                                                   \
  do {                                                              \
    const unsigned long long c_ = clock64();                        \
    if (threadIdx.x == 0) qtt_jac_t[i] += c_ - jt_last_;            \
    jt_last_ = c_;                                                  \
  } while (0)
#define QTT_JC(i) do { if (threadIdx.x == 0) qtt_jac_t[i] += 1; } while (0)
// cycles of thread 0 per segment of a blocked-QR panel
__device__ unsigned long long qtt_pnl_t[16];
extern "C" int qtt_debug_panel(unsigned long long* out) {
  return (int)hipMemcpyFromSymbol(out, HIP_SYMBOL(qtt_pnl_t), 16 * sizeof(unsigned long long));
}
#define QTT_PT_INIT unsigned long long pt_last_ = clock64()
#define QTT_PANEL_COUNT (qtt_pnl_t[7] += 1)
#define QTT_WY_COUNT(coop) (qtt_pnl_t[(coop) ? 0 : 6] += 1)
#define QTT_WY_COUNT_COLS (qtt_pnl_t[12] += 1)
#define QTT_PT(i)                                                   \
  do {                                                              \
    const unsigned long long c_ = clock64();                        \
    if (threadIdx.x == 0) qtt_pnl_t[i] += c_ - pt_last_;            \
    pt_last_ = c_;                                                  \
  } while (0)
#else
#define QTT_PT_INIT do { } while (0)
#define QTT_PT(i) do { } while (0)
#define QTT_PANEL_COUNT (void)0
#define QTT_WY_COUNT(coop) (void)0
#define QTT_WY_COUNT_COLS (void)0
#define QTT_JT_INIT do { } while (0)
#define QTT_JT(i) do { } while (0)
#define QTT_JC(i) do { } while (0)
#endif
#ifndef QTT_PHASE_TIMING
#define QTT_PHASE(step, phase) do { } while (0)
#define QTT_SHAPE(step, rows, b, tr) do { } while (0)
#define QTT_SWEEPS(s) do { } while (0)
#define QTT_SWEEPS_CLEAR(step) do { } while (0)
#endif

namespace {

// ---------------------------------------------------------------------------
// cross-lane sums with DPP (double = two 32-bit halves)
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ inline double readfirstlane_d(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

__device__ inline double readlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// sum over aligned groups of 8 lanes (result in every lane of the group)
__device__ inline double sum8(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);  // row_half_mirror
  return v;
}

__device__ inline double sum64(double v) {
  v = sum8(v);
  v += dpp_d<0x140>(v);  // row_mirror: 16-lane sums
  return readlane_d(v, 0) + readlane_d(v, 16) + readlane_d(v, 32) + readlane_d(v, 48);
}

template <int G>
__device__ inline double gsum(double v) {
  if constexpr (G == 8) return sum8(v);
  else return sum64(v);
}

__device__ inline double block_sum(double v, double* red) {
  v = sum64(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

// ---------------------------------------------------------------------------
// exact operator application
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Arithmetic of the bulk products (the exact contraction and the truncation's
// mode products) for the three precisions of qtt_apply:
//   QTT_F64        float64;
//   QTT_F32_SPLIT  float-float: operands split into (hi, lo) floats, products
//                  and sums by error-free transformations (Dekker / Knuth),
//                  about 48 bits;
//   QTT_F32        float32 operands and accumulation.
// The QR, the SVD and the cutoff decision stay float64 in all three.
// ---------------------------------------------------------------------------
struct ffv {
  float h, l;
};

__device__ __forceinline__ ffv ff_split(double a) {
  const float h = (float)a;
  return {h, (float)(a - (double)h)};
}

// c += a * b (dropping a.l * b.l); every + and * rounded on its own, since
// contraction into FMA would lose the low words
__device__ __forceinline__ void ff_madd(ffv& c, ffv a, ffv b) {
#pragma clang fp contract(off)
  const float p = a.h * b.h;
  float pe = __builtin_fmaf(a.h, b.h, -p);
  pe = __builtin_fmaf(a.h, b.l, pe);
  pe = __builtin_fmaf(a.l, b.h, pe);
  const float s = c.h + p;
  const float bb = s - c.h;
  float e = (c.h - (s - bb)) + (p - bb);
  e = e + c.l + pe;
  const float h = s + e;
  c.l = e - (h - s);
  c.h = h;
}

__device__ __forceinline__ double ff_value(ffv c) { return (double)c.h + (double)c.l; }

// element e (< total) of the contracted network: vertex by binary search over
// cstart, then the flattened indices of the output core
template <int MODE>
__device__ __forceinline__ void contract_elem(const double* __restrict__ W, const int64_t* __restrict__ woff,
                                              const int* __restrict__ wbond, const double* __restrict__ A,
                                              const int64_t* __restrict__ aoff, const int* __restrict__ abond,
                                              const TopoView& t, double* __restrict__ C,
                                              const int64_t* __restrict__ coff, const int64_t* __restrict__ cstart,
                                              int64_t e) {
  int lo = 0, hi = t.nv;  // find v with cstart[v] <= e < cstart[v+1]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (cstart[mid] <= e) lo = mid; else hi = mid;
  }
  const int v = lo;
  int64_t local = e - cstart[v];
  const int m = t.nch[v];
  const int* ch = t.chs + t.choff[v];
  int64_t jw = 0, ja = 0, sw = 1, sa = 1;
  for (int k = m - 1; k >= 0; --k) {
    const int64_t bw = wbond[ch[k]], ba = abond[ch[k]];
    const int64_t idx = local % (bw * ba);
    local /= bw * ba;
    jw += (idx / ba) * sw;
    ja += (idx % ba) * sa;
    sw *= bw;
    sa *= ba;
  }
  const int d = t.sdim[v];
  const int so = (int)(local % d);
  local /= d;
  const int64_t ap = abond[v];
  const int64_t pw = local / ap, pa = local % ap;
  double out;
  if (MODE == QTT_F64) {
    double s = 0.0;
    for (int si = 0; si < d; ++si)
      s += W[woff[v] + ((pw * d + si) * d + so) * sw + jw] * A[aoff[v] + (pa * d + si) * sa + ja];
    out = s;
  } else {
    // each product of float words on operands scaled by powers of two to
    // near 1 (exact), scaled back and summed over the d site values in float64
    double s = 0.0;
    for (int si = 0; si < d; ++si) {
      const double wv = W[woff[v] + ((pw * d + si) * d + so) * sw + jw];
      const double av = A[aoff[v] + (pa * d + si) * sa + ja];
      if (wv == 0.0 || av == 0.0) continue;
      const int ew = ilogb(wv), ea = ilogb(av);
      double p;
      if (MODE == QTT_F32_SPLIT) {
        ffv t = {0.0f, 0.0f};
        ff_madd(t, ff_split(ldexp(wv, -ew)), ff_split(ldexp(av, -ea)));
        p = ff_value(t);
      } else {
        p = (double)((float)ldexp(wv, -ew) * (float)ldexp(av, -ea));
      }
      s += ldexp(p, ew + ea);
    }
    out = s;
  }
  C[coff[v] + (e - cstart[v])] = out;
}

template <int MODE>
__global__ void __launch_bounds__(256)
contract_kernel(const double* __restrict__ W, const int64_t* __restrict__ woff, const int* __restrict__ wbond,
                const double* __restrict__ A, const int64_t* __restrict__ aoff, const int* __restrict__ abond,
                TopoView t, double* __restrict__ C, const int64_t* __restrict__ coff,
                const int64_t* __restrict__ cstart, int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    contract_elem<MODE>(W, woff, wbond, A, aoff, abond, t, C, coff, cstart, e);
}

// A batch of contractions in one launch: the elements of all networks form one
// index space, item i owning [estart[i], estart[i+1]).
struct ContractItem {
  const double* W;
  const int64_t* woff;
  const int* wbond;
  const double* A;
  const int64_t* aoff;
  const int* abond;
  TopoView t;
  double* C;
  const int64_t* coff;
  const int64_t* cstart;
};

template <int MODE>
__global__ void __launch_bounds__(256)
contract_batch_kernel(const ContractItem* __restrict__ items, const int64_t* __restrict__ estart, int n,
                      int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = n;  // find i with estart[i] <= e < estart[i+1]
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (estart[mid] <= e) lo = mid; else hi = mid;
    }
    const ContractItem& it = items[lo];
    contract_elem<MODE>(it.W, it.woff, it.wbond, it.A, it.aoff, it.abond, it.t, it.C, it.coff, it.cstart,
                        e - estart[lo]);
  }
}

// The batch again, split into segments of at most kSegElems elements of one
// core: a workgroup takes one segment, so the network and the vertex are
// uniform (no searches) and the index arithmetic is 32-bit.
constexpr int kSegElems = 4096;
struct ContractSeg {
  int item, v, start, count;
};

template <int MODE>
__global__ void __launch_bounds__(256)
contract_seg_kernel(const ContractItem* __restrict__ items, const ContractSeg* __restrict__ segs) {
  const ContractSeg sg = segs[blockIdx.x];
  const ContractItem& it = items[sg.item];
  const TopoView& t = it.t;
  const int v = sg.v;
  const int m = t.nch[v];
  const int* ch = t.chs + t.choff[v];
  const int d = t.sdim[v];
  const int ap = it.abond[v];
  const double* __restrict__ Wv = it.W + it.woff[v];
  const double* __restrict__ Av = it.A + it.aoff[v];
  double* __restrict__ Cv = it.C + it.coff[v];
  int bwk[QTT_MAX_NDIM], bak[QTT_MAX_NDIM];
  for (int k = 0; k < m; ++k) {
    bwk[k] = it.wbond[ch[k]];
    bak[k] = it.abond[ch[k]];
  }
  for (int q = threadIdx.x; q < sg.count; q += blockDim.x) {
    unsigned local = (unsigned)(sg.start + q);
    unsigned jw = 0, ja = 0, sw = 1, sa = 1;
    for (int k = m - 1; k >= 0; --k) {
      const unsigned bw = bwk[k], ba = bak[k], bb = bw * ba;
      const unsigned idx = local % bb;
      local /= bb;
      jw += (idx / ba) * sw;
      ja += (idx % ba) * sa;
      sw *= bw;
      sa *= ba;
    }
    const unsigned so = local % (unsigned)d;
    local /= (unsigned)d;
    const unsigned pw = local / (unsigned)ap, pa = local % (unsigned)ap;
    double out = 0.0;
    if (MODE == QTT_F64) {
      for (int si = 0; si < d; ++si)
        out += Wv[((int64_t)(pw * d + si) * d + so) * sw + jw] * Av[(int64_t)(pa * d + si) * sa + ja];
    } else {
      for (int si = 0; si < d; ++si) {
        const double wv = Wv[((int64_t)(pw * d + si) * d + so) * sw + jw];
        const double av = Av[(int64_t)(pa * d + si) * sa + ja];
        if (wv == 0.0 || av == 0.0) continue;
        const int ew = ilogb(wv), ea = ilogb(av);
        double p;
        if (MODE == QTT_F32_SPLIT) {
          ffv tt = {0.0f, 0.0f};
          ff_madd(tt, ff_split(ldexp(wv, -ew)), ff_split(ldexp(av, -ea)));
          p = ff_value(tt);
        } else {
          p = (double)((float)ldexp(wv, -ew) * (float)ldexp(av, -ea));
        }
        out += ldexp(p, ew + ea);
      }
    }
    Cv[sg.start + q] = out;
  }
}

// ---------------------------------------------------------------------------
// truncation sweep
// ---------------------------------------------------------------------------
struct SweepArgs {
  double* data;
  const int64_t* off;
  int* bond_g;
  TopoView t;
  const int* steps;
  int nsteps;
  double cutoff;
  long long maxdim;
  int max_sweeps;  // Jacobi sweeps before an SVD is reported as not converged
  int prec;        // arithmetic of the mode products (QTT_F64, QTT_F32_SPLIT, QTT_F32)
  double* gscr;  // global workspace (used when a matrix does not fit LDS)
  double* gY;    // global buffer for the mode-product result
  int ws_lds;    // doubles of LDS workspace
  double* err_out;
  int* status;   // [1] <- non-convergence
  struct CoopJob* job;  // jobs for helper workgroups (single-network launches), or null
  int nwg;              // workgroups taking part in a job (1: no helpers)
  double* gT;           // T factors of the QR moves whose Q is kept as reflectors: vertex v's at gT + v tstride
  int tstride;
};

// A job posted by workgroup 0 of a single-network launch to the helper
// workgroups that follow it in the grid (a cooperative launch, so all are
// resident).  Workgroup 0 writes the job, then releases seq; every workgroup
// processes its share and counts itself in done; workgroup 0 waits until done
// reaches nwg per job before it posts the next one, so no workgroup reads a
// job while it changes.  quit ends the helpers.  Zeroed by the host before
// the launch.  Every wait is bounded: a helper that waits too long for a job
// exits, and workgroup 0 sets err (reported by the host) if a job does not
// complete in time.
struct CoopJob {
  int seq;
  int quit;
  // the posted job: seq << 32 | nact, one atomic word, so a helper reads both
  // together; workgroups 0 .. nact-1 take part, the others skip it and do not
  // count it.  seq and dtarget (the completions posted so far) are workgroup
  // 0's own.
  unsigned long long post;
  unsigned int dtarget;
  int rot;   // a rotation happened in the current Jacobi sweep
  int err;   // a wait exceeded its limit
  unsigned int done;
  int pad;
  int type;  // 1: Jacobi round, 2: mode product columns, 3: copy, ... (below)
  // type 1: one round t of a one-sided Jacobi sweep on the rows of X (n x b,
  // ld ldx) with the rotations kept in J (n x n, ld ldj), both in global memory
  double* X;
  double* J;
  int n, b, ldx, ldj, t, nn, e;  // e: entries per lane of the row loads (2, 4, 6 or 8)
  double tol, small, noise;
  // type 2: Y = R (x) U over the columns (aa, bb), each workgroup a range of
  // them (mode_product_mfma); type 3: D[0, ytot) <- Y, each a range
  // type 10: a move's matrix gathered from its core, D <- U (gather_tiles
  // with Au blocks, rb = b, Bu = bx)
  const double* R;
  const int* perm;
  const double* U;
  double* Y;
  double* D;
  long long ytot;
  int use_perm, m, rb, Au, nu, Bu;
  int prec;  // arithmetic of the mode product (QTT_F64, QTT_F32_SPLIT, QTT_F32)
  // type 4: the first sums of a tall panel (X = the panel, ldx, n = its rows,
  // nb columns); type 5: pass t of a tall panel (tol = the reflector's scale,
  // small = its product with column t+1, dcol per column); both leave 16 sums
  // per workgroup in Gp.  type 6 / 7: the two MFMA phases of a WY update
  // (wy_* fields), the first leaving per-unit partial tiles in Pg, the second
  // reading -U from Ug.
  int nb;
  int bsz, nblk;  // type 8: block round t of a block Jacobi sweep (blocks of bsz rows)
  // type 9: a whole block Jacobi sweep; each workgroup takes its block pair of
  // every round as soon as the round before has written both blocks back:
  // ver[b] = rbase + (rounds of this sweep that block b has been through)
  int rbase;
  int* ver;
  double dcol[16];
  double* Gp;  // [workgroup][16]
  double* Pg;  // [unit][256]
  double* Ug;  // [tile][16][16]
  double* wy_M;
  const double* wy_V;
  int wy_ldm, wy_ldv, wy_mr, wy_cg, wy_c1, wy_nct, wy_nch, wy_chunk;
  // type 11: a whole WY update (wy_cols) over columns [wy_cg, wy_c1), T in
  // Ug[0, 256), op(T) = T^T if wy_trans
  int wy_trans;
};
constexpr int COOP_PG_UNITS = 512;  // partial tiles Pg holds

constexpr int COOP_SPIN_MASTER = 1 << 22;  // polls of workgroup 0 for a job's completion
constexpr int COOP_SPIN_HELPER = 1 << 24;  // polls of a helper for the next job (long moves in between)

__device__ inline int core_dims(const TopoView& t, const int* bond, int v, int64_t* dims) {
  const int m = t.nch[v];
  dims[0] = bond[v];
  dims[1] = t.sdim[v];
  for (int k = 0; k < m; ++k) dims[2 + k] = bond[t.chs[t.choff[v] + k]];
  return 2 + m;
}

__device__ inline int axis_of(const TopoView& t, int x, int u) {
  if (t.parent[x] == u) return 0;
  const int* ch = t.chs + t.choff[x];
  for (int k = 0; k < t.nch[x]; ++k)
    if (ch[k] == u) return 2 + k;
  return -1;
}

// Core v viewed as (A, n, B) around its bond axis to w: dims [parent bond,
// site, child bonds...], n = the bond to w, A / B the products before / after.
__device__ inline void move_shape(const TopoView& t, const int* bond, int v, int w, int* out) {
  int A = 1, n = 1, B = 1;
  bool seen = false;
  auto visit = [&](int d, bool axis) {
    if (axis) {
      n = d;
      seen = true;
    } else if (seen) {
      B *= d;
    } else {
      A *= d;
    }
  };
  visit(bond[v], t.parent[v] == w);
  visit(t.sdim[v], false);
  const int* ch = t.chs + t.choff[v];
  for (int k = 0; k < t.nch[v]; ++k) visit(bond[ch[k]], ch[k] == w);
  out[0] = A;
  out[1] = n;
  out[2] = B;
}

__device__ inline void split_dims(const int64_t* dims, int nd, int ax, int64_t& A, int64_t& n, int64_t& B) {
  A = 1;
  B = 1;
  for (int i = 0; i < ax; ++i) A *= dims[i];
  n = dims[ax];
  for (int i = ax + 1; i < nd; ++i) B *= dims[i];
}

// Householder reflector of column j (LAPACK dlarfg), computed by one group:
// v = [1; W[j+1:, j] * scal] stored below the diagonal, beta on the diagonal.
// The strided loops below run 4 elements per iteration (loads first, then the
// arithmetic) so several LDS/L2 loads are in flight per wait; the summation
// order is fixed, so results stay deterministic.
template <int G, typename P = double*>
__device__ inline void grp_reflector(P W, int rows, int ld, int j, P tau) {
  const int gl = threadIdx.x & (G - 1);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = j + 1 + gl;
  for (; i + 3 * G < rows; i += 4 * G) {
    const double x0 = W[(int64_t)i * ld + j], x1 = W[(int64_t)(i + G) * ld + j];
    const double x2 = W[(int64_t)(i + 2 * G) * ld + j], x3 = W[(int64_t)(i + 3 * G) * ld + j];
    s0 = __builtin_fma(x0, x0, s0);
    s1 = __builtin_fma(x1, x1, s1);
    s2 = __builtin_fma(x2, x2, s2);
    s3 = __builtin_fma(x3, x3, s3);
  }
  for (; i < rows; i += G) {
    const double x = W[(int64_t)i * ld + j];
    s0 = __builtin_fma(x, x, s0);
  }
  const double xn2 = gsum<G>((s0 + s1) + (s2 + s3));
  const double x0 = W[(int64_t)j * ld + j];
  double tj = 0.0, beta = x0;
  if (xn2 > 0.0) {
    const double nrm = sqrt(x0 * x0 + xn2);
    beta = x0 >= 0.0 ? -nrm : nrm;
    tj = (beta - x0) / beta;
    const double scal = 1.0 / (x0 - beta);
    i = j + 1 + gl;
    for (; i + 3 * G < rows; i += 4 * G) {
      P p0 = W + (int64_t)i * ld + j;
      P p1 = W + (int64_t)(i + G) * ld + j;
      P p2 = W + (int64_t)(i + 2 * G) * ld + j;
      P p3 = W + (int64_t)(i + 3 * G) * ld + j;
      const double y0 = *p0, y1 = *p1, y2 = *p2, y3 = *p3;
      *p0 = y0 * scal;
      *p1 = y1 * scal;
      *p2 = y2 * scal;
      *p3 = y3 * scal;
    }
    for (; i < rows; i += G) W[(int64_t)i * ld + j] *= scal;
  }
  if (gl == 0) {
    tau[j] = tj;
    W[(int64_t)j * ld + j] = beta;
  }
}

// apply H = I - tau v v^T (v[0] = 1 implicit, v[i] = vcol[i * vstride] for i > j) to column col
template <int G, typename P = double*, typename CP = const double*>
__device__ inline void grp_apply(P W, int rows, int ld, int j, int col, CP vcol, int vstride, double tj) {
  const int gl = threadIdx.x & (G - 1);
  double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
  int i = j + 1 + gl;
  for (; i + 3 * G < rows; i += 4 * G) {
    const double v0 = vcol[(int64_t)i * vstride], v1 = vcol[(int64_t)(i + G) * vstride];
    const double v2 = vcol[(int64_t)(i + 2 * G) * vstride], v3 = vcol[(int64_t)(i + 3 * G) * vstride];
    const double w0 = W[(int64_t)i * ld + col], w1 = W[(int64_t)(i + G) * ld + col];
    const double w2 = W[(int64_t)(i + 2 * G) * ld + col], w3 = W[(int64_t)(i + 3 * G) * ld + col];
    d0 = __builtin_fma(v0, w0, d0);
    d1 = __builtin_fma(v1, w1, d1);
    d2 = __builtin_fma(v2, w2, d2);
    d3 = __builtin_fma(v3, w3, d3);
  }
  for (; i < rows; i += G) d0 = __builtin_fma(vcol[(int64_t)i * vstride], W[(int64_t)i * ld + col], d0);
  const double d = (gsum<G>((d0 + d1) + (d2 + d3)) + W[(int64_t)j * ld + col]) * tj;
  i = j + 1 + gl;
  for (; i + 3 * G < rows; i += 4 * G) {
    const double v0 = vcol[(int64_t)i * vstride], v1 = vcol[(int64_t)(i + G) * vstride];
    const double v2 = vcol[(int64_t)(i + 2 * G) * vstride], v3 = vcol[(int64_t)(i + 3 * G) * vstride];
    P p0 = W + (int64_t)i * ld + col;
    P p1 = W + (int64_t)(i + G) * ld + col;
    P p2 = W + (int64_t)(i + 2 * G) * ld + col;
    P p3 = W + (int64_t)(i + 3 * G) * ld + col;
    const double w0 = *p0, w1 = *p1, w2 = *p2, w3 = *p3;
    *p0 = __builtin_fma(-d, v0, w0);
    *p1 = __builtin_fma(-d, v1, w1);
    *p2 = __builtin_fma(-d, v2, w2);
    *p3 = __builtin_fma(-d, v3, w3);
  }
  for (; i < rows; i += G) W[(int64_t)i * ld + col] = __builtin_fma(-d, vcol[(int64_t)i * vstride], W[(int64_t)i * ld + col]);
  if (gl == 0) W[(int64_t)j * ld + col] -= d;
}

__device__ inline int first_owned(int start, int gid, int ng) {
  return start + ((gid - start % ng) + ng) % ng;
}

// Householder QR in place (LAPACK dgeqr2): W is rows x cols with leading
// dimension ld; reflectors below the diagonal, R on and above it, tau[k].
// Column c belongs to group c % ng; the owner of column j+1 updates it first
// and builds the next reflector right away (look-ahead): one barrier per column.
template <int G, typename P = double*>
__device__ void geqr2(P W, int rows, int cols, int ld, P tau) {
  const int gid = threadIdx.x / G, ng = blockDim.x / G;
  const int k = rows < cols ? rows : cols;
  if (k <= 0) return;
  if (gid == 0) grp_reflector<G>(W, rows, ld, 0, tau);
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    const double tj = tau[j];
    for (int col = first_owned(j + 1, gid, ng); col < cols; col += ng) {
      if (tj != 0.0) grp_apply<G>(W, rows, ld, j, col, W + j, ld, tj);
      if (col == j + 1 && j + 1 < k) grp_reflector<G>(W, rows, ld, j + 1, tau);
    }
    __syncthreads();
  }
}

// Form Q (rows x k) in place from the reflectors (LAPACK dorg2r).  The
// reflector of step j is staged in vbuf (double buffered) so the owner of
// column j can overwrite it with the Q column while the other groups still
// apply H_j: one barrier per column.
template <int G>
__device__ void org2r(double* W, int rows, int k, int ld, const double* tau, double* vbuf) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int gl = tid & (G - 1), gid = tid / G, ng = bs / G;
  if (k <= 0) return;
  for (int i = tid; i < rows; i += bs) vbuf[((k - 1) & 1) * rows + i] = W[(int64_t)i * ld + (k - 1)];
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const double tj = tau[j];
    const double* v = vbuf + (j & 1) * rows;
    for (int col = first_owned(j + 1, gid, ng); col < k; col += ng)
      if (tj != 0.0) grp_apply<G>(W, rows, ld, j, col, v, 1, tj);
    if (gid == j % ng) {
      for (int i = gl; i < rows; i += G) {
        double* a = W + (int64_t)i * ld + j;
        if (i < j) *a = 0.0;
        else if (i == j) *a = 1.0 - tj;
        else *a = -tj * v[i];
      }
    }
    if (j > 0 && gid == (j - 1) % ng) {
      for (int i = gl; i < rows; i += G) vbuf[((j - 1) & 1) * rows + i] = W[(int64_t)i * ld + (j - 1)];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// register-resident Householder QR for matrices of up to GL * RPL rows and
// NS * groups columns (groups of GL lanes).  Lane gl of a group holds rows
// gl + GL t (t < RPL) of each of its columns in registers; column c belongs
// to group c % groups, slot c / groups.  Only the reflectors travel through
// memory: reflector j is the row VT[j * GL RPL, (j + 1) * GL RPL), zero past
// the last matrix row, so every step reads it with fixed offsets from one
// base address (one contiguous load per group and register).
// ---------------------------------------------------------------------------

// Lane masks built with integer arithmetic (all ones / zero), so the per-row
// tests of the unrolled loops need no condition registers.
__device__ __forceinline__ uint64_t lt_mask(int a, int b) {  // a < b
  return (uint64_t)(int64_t)((a - b) >> 31);
}
__device__ __forceinline__ uint64_t eq_mask(int a, int b) {  // a == b
  const int d = a - b;
  return (uint64_t)(int64_t)(~((d | -d) >> 31));
}
__device__ __forceinline__ double keep(double x, uint64_t m) {
  return __longlong_as_double((long long)((uint64_t)__double_as_longlong(x) & m));
}
// The argument of a non-inlined function is uniform across the wavefront but
// arrives in a vector register; reading it into a scalar register makes the
// loops and tests that depend on it scalar (no per-lane masks).
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// The lane index made opaque to the optimizer: the per-row masks are then
// recomputed where they are used instead of being hoisted out of the column
// loop and held (dozens of condition registers) across it.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// sum over aligned groups of GL (8 or 16) lanes
template <int GL>
__device__ __forceinline__ double lanesum(double v) {
  v = sum8(v);
  if constexpr (GL >= 16) v += dpp_d<0x140>(v);  // row_mirror: 16-lane sums
  return v;
}

template <int GL, int RPL, typename CP>
__device__ __forceinline__ void reg_load_v(CP VT, int j, int gl, double (&v)[RPL]) {
  CP p = VT + j * (GL * RPL) + gl;
#pragma unroll
  for (int t = 0; t < RPL; ++t) v[t] = p[GL * t];
}

// a <- (I - tau v v^T) a for the column held in a
template <int GL, int RPL>
__device__ __forceinline__ void reg_apply(double (&a)[RPL], const double (&v)[RPL], double tj) {
  double d0 = 0.0, d1 = 0.0;
#pragma unroll
  for (int t = 0; t + 1 < RPL; t += 2) {
    d0 = __builtin_fma(v[t], a[t], d0);
    d1 = __builtin_fma(v[t + 1], a[t + 1], d1);
  }
  const double d = lanesum<GL>(d0 + d1) * tj;
#pragma unroll
  for (int t = 0; t < RPL; ++t) a[t] = __builtin_fma(-d, v[t], a[t]);
}

// the same with the reflector read from memory in each pass (fewer live registers)
template <int GL, int RPL, typename CP>
__device__ __forceinline__ void reg_apply_mem(double (&a)[RPL], CP VT, int j, int gl, double tj) {
  CP p = VT + j * (GL * RPL) + gl;
  double d0 = 0.0, d1 = 0.0;
#pragma unroll
  for (int t = 0; t + 1 < RPL; t += 2) {
    d0 = __builtin_fma(p[GL * t], a[t], d0);
    d1 = __builtin_fma(p[GL * t + GL], a[t + 1], d1);
  }
  const double d = lanesum<GL>(d0 + d1) * tj;
#pragma unroll
  for (int t = 0; t < RPL; ++t) a[t] = __builtin_fma(-d, p[GL * t], a[t]);
}

// Reflector of column j held in a (LAPACK dlarfg): a keeps its R part (beta at
// row j, zeros below), the reflector goes to VT (zero above row j, one at row
// j), tau[j] the scalar factor.
template <int GL, int RPL, typename DP>
__device__ __forceinline__ void reg_reflector(double (&a)[RPL], int j, int gl_in, DP VT, DP tau) {
  const int gl = opaque(gl_in);
  double s0 = 0.0, s1 = 0.0, xj = 0.0;
#pragma unroll
  for (int t = 0; t + 1 < RPL; t += 2) {
    const int r0 = gl + GL * t, r1 = r0 + GL;
    const double x0 = keep(a[t], lt_mask(j, r0)), x1 = keep(a[t + 1], lt_mask(j, r1));
    s0 = __builtin_fma(x0, x0, s0);
    s1 = __builtin_fma(x1, x1, s1);
    xj += keep(a[t], eq_mask(r0, j)) + keep(a[t + 1], eq_mask(r1, j));
  }
  const double xn2 = lanesum<GL>(s0 + s1);
  const double x0 = lanesum<GL>(xj);
  double tj = 0.0, beta = x0, scal = 0.0;
  if (xn2 > 0.0) {
    const double nrm = sqrt(x0 * x0 + xn2);
    beta = x0 >= 0.0 ? -nrm : nrm;
    tj = (beta - x0) / beta;
    scal = 1.0 / (x0 - beta);
  }
  DP p = VT + j * (GL * RPL) + gl;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int row = gl + GL * t;
    const uint64_t below = lt_mask(j, row), diag = eq_mask(row, j), above = lt_mask(row, j);
    p[GL * t] = keep(a[t] * scal, below) + keep(1.0, diag);  // rows past the matrix hold a = 0
    a[t] = keep(a[t], above) + keep(beta, diag);
  }
  if (gl == 0) tau[j] = tj;
}

// QR of W (rows x cols, leading dimension cols, rows <= GL RPL, cols <= NS
// groups): R (k x cols) into Rb, the reflectors into VT, which overwrites W
// (k * GL RPL doubles from W).  The owner of column j + 1 updates it first and
// builds the next reflector right away, so there is one barrier per column.
template <int GL, int RPL, int NS, typename DP>
__device__ __forceinline__ void qr_body(DP W, int rows, int cols, DP tau, DP Rb) {
  const int tid = threadIdx.x, bs = blockDim.x, gl = tid & (GL - 1), gid = tid / GL, ng = bs / GL;
  const int k = rows < cols ? rows : cols;
  double a[NS][RPL];
  const int gll = opaque(gl);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int c = gid + ng * s;
    const int cc = c < cols ? c : cols - 1;
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int row = gll + GL * t;
      const double x = W[(row < rows ? row : rows - 1) * cols + cc];  // clamped: always in the matrix
      a[s][t] = c < cols ? keep(x, lt_mask(row, rows)) : 0.0;
    }
  }
  __syncthreads();
  DP VT = W;
  if (gid == 0) reg_reflector<GL, RPL>(a[0], 0, gl, VT, tau);
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    const double tj = tau[j];
    if (gid + ng * (NS - 1) > j) {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int c = gid + ng * s;
        if (c > j && c < cols) {
          // (tau 0 leaves the column as it is: the reflector is finite, so the
          // update subtracts zeros; not testing it lets tau and the reflector
          // load together)
          reg_apply_mem<GL, RPL>(a[s], VT, j, gl, tj);
          if (c == j + 1 && c < k) reg_reflector<GL, RPL>(a[s], c, gl, VT, tau);
        }
      }
    }
    __syncthreads();
  }
  const int glr = opaque(gl);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int c = gid + ng * s;
    if (c < cols) {
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int row = glr + GL * t;
        if (row < k) Rb[row * cols + c] = a[s][t];
      }
    }
  }
}

// u <- H_0 ... H_jtop u
template <int GL, int RPL, typename CP>
__device__ __forceinline__ void reg_apply_chain(double (&u)[RPL], CP VT, CP tau, int gl, int jtop) {
  for (int j = jtop; j >= 0; --j) {
    const double tj = tau[j];
    if (tj != 0.0) reg_apply_mem<GL, RPL>(u, VT, j, gl, tj);
  }
}

// Offset in the core of row gl + GL t of the matrix (rows split as (aa, bb),
// bb < bx, bond index 0), -1 past the last row.
template <int GL, int RPL>
__device__ __forceinline__ void reg_dest(int rows, int gl, int m, int bx, int (&dst)[RPL]) {
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int row = gl + GL * t;
    const int aa = row / bx, bb = row - aa * bx;
    const int valid = (int)lt_mask(row, rows);
    dst[t] = ((aa * m * bx + bb) & valid) | ~valid;
  }
}

template <int RPL>
__device__ __forceinline__ void reg_scatter(const double (&u)[RPL], const int (&dst)[RPL], double* Xc, int col,
                                            int bx) {
  double* xc = Xc + (int64_t)col * bx;
#pragma unroll
  for (int t = 0; t < RPL; ++t)
    if (dst[t] >= 0) xc[dst[t]] = u[t];
}

// x <- Q (rows x k): column c = H_0 ... H_c e_c, no barriers.
template <int GL, int RPL, typename CP>
__device__ __forceinline__ void q_form_body(CP VT, int rows, int k, CP tau, double* Xc, int bx) {
  const int tid = threadIdx.x, gl = tid & (GL - 1), gid = tid / GL, ng = blockDim.x / GL;
  int dst[RPL];
  reg_dest<GL, RPL>(rows, gl, k, bx, dst);
  for (int c = gid; c < k; c += ng) {
    const int glo = opaque(gl);
    double u[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) u[t] = keep(1.0, eq_mask(glo + GL * t, c));
    reg_apply_chain<GL, RPL>(u, VT, tau, gl, c);
    reg_scatter<RPL>(u, dst, Xc, c, bx);
  }
}

// x <- U = Q [J_perm^T; 0] (rows x m): column i starts as row perm[i] of J
// (length n) and takes all k reflectors.
template <int GL, int RPL, typename CP, typename CIP>
__device__ __forceinline__ void u_form_body(CP VT, int rows, int k, CP tau, CP Jb, int n, CIP perm, int m,
                                            double* Xc, int bx) {
  const int tid = threadIdx.x, gl = tid & (GL - 1), gid = tid / GL, ng = blockDim.x / GL;
  int dst[RPL];
  reg_dest<GL, RPL>(rows, gl, m, bx, dst);
  for (int i = gid; i < m; i += ng) {
    CP jr = Jb + perm[i] * n;
    const int glo = opaque(gl);
    double u[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int row = glo + GL * t;
      u[t] = keep(jr[row < n ? row : n - 1], lt_mask(row, n));  // clamped read, zero past n
    }
    reg_apply_chain<GL, RPL>(u, VT, tau, gl, k - 1);
    reg_scatter<RPL>(u, dst, Xc, i, bx);
  }
}

// The register QR pieces run as separate (not inlined) functions on an LDS
// workspace: their registers then do not add to the sweep kernel's, and the
// LDS-qualified pointers keep the memory operations on LDS instructions.
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(3))) int lds_i32;

template <int GL, int RPL, int NS>
__device__ __noinline__ void qr_reg(lds_f64* W, int rows, int cols, lds_f64* tau, lds_f64* Rb) {
  qr_body<GL, RPL, NS, lds_f64*>(W, uni(rows), uni(cols), tau, Rb);
}

template <int GL, int RPL>
__device__ __noinline__ void q_form_reg(const lds_f64* VT, int rows, int k, const lds_f64* tau, double* Xc, int bx) {
  q_form_body<GL, RPL, const lds_f64*>(VT, uni(rows), uni(k), tau, Xc, uni(bx));
}

template <int GL, int RPL>
__device__ __noinline__ void u_form_reg(const lds_f64* VT, int rows, int k, const lds_f64* tau, const lds_f64* Jb,
                                        int n, const lds_i32* perm, int m, double* Xc, int bx) {
  u_form_body<GL, RPL, const lds_f64*, const lds_i32*>(VT, uni(rows), uni(k), tau, Jb, uni(n), perm, uni(m), Xc,
                                                       uni(bx));
}

__device__ inline int rr_player(int pos, int t, int nn) {
  // pos < nn and t < nn - 1, so one conditional subtraction is the modulo
  int x = pos - 1 + t;
  x = x >= nn - 1 ? x - (nn - 1) : x;
  return pos == 0 ? 0 : 1 + x;
}

// Rows whose squared norm is at most cutoff * ||X||_F^2 / n (the noise level)
// are not rotated against each other.  At convergence every row above the
// level is orthogonal to all rows, so it is an exact singular triplet; the
// rows below it (the noise block) weigh at most cutoff * ||X||_F^2 together
// but are not yet mutually orthogonal.  noise_block_exact below decides
// whether that block can change what the truncation keeps; if it can, the
// Jacobi runs again with the rule off (cutoff 0, rotations continued in J),
// which rotates only the unconverged pairs of the block.  A cutoff outside
// (0, 1) rotates every pair.
__device__ inline double noise_level(double fro2, double cut, int n) {
  return (cut > 0.0 && cut < 1.0) ? cut * fro2 / (double)n : 0.0;
}

// true if the noise block of the rows with norms sig[0..n) cannot change the
// truncation: each of its singular values has a square of at most its weight
// W, so when W is below the smallest squared norm q of the rows above the
// level, all of them sort after those rows and are all discarded (W <= cut
// ||X||_F^2), and what is kept and the discarded weight are those of the full
// SVD.  The level is widened by 1e-9 relative so that a row on the boundary
// counts as noise.  Block-wide; slot is a broadcast int in shared memory.
__device__ bool noise_block_exact(const double* sig, int n, double cut, int* slot) {
  if (threadIdx.x == 0) {
    double fro = 0.0, W = 0.0, q = -1.0;
    int ns = 0;
    for (int i = 0; i < n; ++i) fro += sig[i] * sig[i];
    const double level = cut * fro / (double)n * (1.0 + 1e-9);
    for (int i = 0; i < n; ++i) {
      const double a = sig[i] * sig[i];
      if (a <= level) {
        W += a;
        ++ns;
      } else if (q < 0.0 || a < q) {
        q = a;
      }
    }
    *slot = (ns <= 1 || (q >= 0.0 && W <= q * (1.0 - 1e-9))) ? 1 : 0;
  }
  __syncthreads();
  const int r = *slot;
  __syncthreads();
  return r != 0;
}

// Stopping rule of every Jacobi variant here: a sweep whose rotations all had
// |cos(angle between the rows)| at most sqrt(tol) / 8 ends the iteration.
// One-sided Jacobi converges quadratically, so the next sweep's angles would
// be about the square of these, below the rotation threshold tol (tol =
// eps max(n, b)): that sweep would rotate nothing, or by angles of order
// eps, and is not run.  Returns the bound on g^2 / (a c).
__device__ __forceinline__ double jacobi_stop2(double tol) { return tol * (1.0 / 64.0); }

// One-sided Jacobi on the rows of X (n x b, ld b), rotations accumulated in J
// (n x n; set to the identity first when init_j).  Returns false if it did
// not converge.
template <int G>
__device__ bool jacobi_rows(double* X, int n, int b, double* J, int* flag, double* red, double cut, bool init_j,
                            int max_sweeps) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int gl = tid & (G - 1), gid = tid / G, ng = bs / G;
  if (init_j)
    for (int e = tid; e < n * n; e += bs) J[e] = (e / n == e % n) ? 1.0 : 0.0;
  __syncthreads();
  if (n < 2) return true;
  const int nn = n + (n & 1);
  // a computed dot product of length b carries rounding error up to ~b eps
  const double tol = DBL_EPSILON * (double)(b > n ? b : n);
  // rows below eps * ||X||_F are numerically zero and are left alone
  double fro = 0.0;
  for (int64_t e = tid; e < (int64_t)n * b; e += bs) fro += X[e] * X[e];
  const double fsum = block_sum(fro, red);
  const double small = fsum * DBL_EPSILON * DBL_EPSILON;
  const double noise = noise_level(fsum, cut, n);
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) *flag = 0;
    __syncthreads();
    for (int t = 0; t < nn - 1; ++t) {
      for (int pi = gid; pi < nn / 2; pi += ng) {
        int p = rr_player(pi, t, nn), q = rr_player(nn - 1 - pi, t, nn);
        if (p > q) {
          const int tmp = p;
          p = q;
          q = tmp;
        }
        if (q >= n) continue;
        double* xp = X + (int64_t)p * b;
        double* xq = X + (int64_t)q * b;
        double a0 = 0.0, c0 = 0.0, g0 = 0.0, a1 = 0.0, c1 = 0.0, g1 = 0.0;
        int i = gl;
        for (; i + G < b; i += 2 * G) {
          const double u0 = xp[i], w0 = xq[i], u1 = xp[i + G], w1 = xq[i + G];
          a0 = __builtin_fma(u0, u0, a0);
          c0 = __builtin_fma(w0, w0, c0);
          g0 = __builtin_fma(u0, w0, g0);
          a1 = __builtin_fma(u1, u1, a1);
          c1 = __builtin_fma(w1, w1, c1);
          g1 = __builtin_fma(u1, w1, g1);
        }
        if (i < b) {
          const double u0 = xp[i], w0 = xq[i];
          a0 = __builtin_fma(u0, u0, a0);
          c0 = __builtin_fma(w0, w0, c0);
          g0 = __builtin_fma(u0, w0, g0);
        }
        const double a = gsum<G>(a0 + a1);
        const double c = gsum<G>(c0 + c1);
        const double g = gsum<G>(g0 + g1);
        if (g != 0.0 && a > small && c > small && (a > noise || c > noise) && fabs(g) > tol * sqrt(a) * sqrt(c)) {
          const double zeta = (c - a) / (2.0 * g);
          const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
          i = gl;
          for (; i + G < b; i += 2 * G) {
            const double u0 = xp[i], w0 = xq[i], u1 = xp[i + G], w1 = xq[i + G];
            xp[i] = cs * u0 - sn * w0;
            xq[i] = sn * u0 + cs * w0;
            xp[i + G] = cs * u1 - sn * w1;
            xq[i + G] = sn * u1 + cs * w1;
          }
          if (i < b) {
            const double u0 = xp[i], w0 = xq[i];
            xp[i] = cs * u0 - sn * w0;
            xq[i] = sn * u0 + cs * w0;
          }
          double* jp = J + (int64_t)p * n;
          double* jq = J + (int64_t)q * n;
          i = gl;
          for (; i + G < n; i += 2 * G) {
            const double u0 = jp[i], w0 = jq[i], u1 = jp[i + G], w1 = jq[i + G];
            jp[i] = cs * u0 - sn * w0;
            jq[i] = sn * u0 + cs * w0;
            jp[i + G] = cs * u1 - sn * w1;
            jq[i + G] = sn * u1 + cs * w1;
          }
          if (i < n) {
            const double u0 = jp[i], w0 = jq[i];
            jp[i] = cs * u0 - sn * w0;
            jq[i] = sn * u0 + cs * w0;
          }
          if (gl == 0 && g * g > jacobi_stop2(tol) * (a * c)) *flag = 1;
        }
      }
      __syncthreads();
    }
    const int rot = *flag;
    __syncthreads();
    if (!rot) return true;
  }
  return false;
}

// One-sided Jacobi as jacobi_rows<G> with G lanes per pair sized to the row
// length (b <= 4 G): each group loads both rows of its pair with all loads in
// flight at once (4 entries per lane), rotates them in registers and stores
// them back, then does the same for the two rows of J (length n <= b).  A
// round then costs about one memory latency plus the rotation instead of one
// latency per 8 entries.  Rows are read in whole blocks from one base
// address; entries past a row's end (fewer than 4 G, inside the workspace,
// which is padded) are read and replaced by zeros.
constexpr int JEPL = 4;

// The rounds run in a separate (not inlined) function on LDS-qualified
// pointers, like the register QR above.

template <int G>
__device__ __forceinline__ double jsum(double v) {
  v = sum8(v);
  if constexpr (G >= 16) v += dpp_d<0x140>(v);  // row_mirror: 16-lane sums
  if constexpr (G >= 32) v += __shfl_xor(v, 16);
  if constexpr (G >= 64) v += __shfl_xor(v, 32);
  return v;
}

// F: the element type, LDS (lds_f64) or global (double) memory; JE entries per lane
template <int G, int JE, typename F>
__device__ __forceinline__ void load_row(const F* r, int len, int gl, double (&x)[JE]) {
  const F* p = r + gl;
#pragma unroll
  for (int e = 0; e < JE; ++e) x[e] = keep(p[G * e], lt_mask(gl + G * e, len));
}

template <int G, int JE, typename F>
__device__ __forceinline__ void store_row(F* r, int len, int gl, const double (&x)[JE]) {
#pragma unroll
  for (int e = 0; e < JE; ++e) {
    const int i = gl + G * e;
    if (i < len) r[i] = x[e];
  }
}

// 1/x and 1/sqrt(x) from the hardware estimates refined by two Newton steps
// (about 1 ulp); they replace the IEEE division and square root sequences,
// whose latency dominates a Jacobi round.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = __builtin_fma(-x * y, y, 1.0);
  y = __builtin_fma(0.5 * y, e, y);
  e = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(0.5 * y, e, y);
}

// Rotation of a row pair with squared norms a, c and inner product g != 0
// that makes the rows orthogonal: tan = t with
//   t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),  zeta = (c - a) / (2 g),
// evaluated as t = +-|2 g| / (|c - a| + sqrt((c - a)^2 + 4 g^2)) (the same
// value with one reciprocal fewer on the dependent chain of a round); then
// cs = 1 / sqrt(1 + t^2), sn = cs t.  (a, c and g come from row entries, so
// (c - a)^2 + 4 g^2 stays finite as a c does.)
__device__ __forceinline__ void jacobi_rotation(double a, double c, double g, double& cs, double& sn) {
  const double d = c - a, g2 = 2.0 * g;
  const double r2 = __builtin_fma(d, d, g2 * g2);
  const double h = r2 * rsq_nr(r2);                 // sqrt((c - a)^2 + 4 g^2)
  const double t0 = fabs(g2) * rcp_nr(fabs(d) + h);  // |t|
  const double tt = d * g >= 0.0 ? t0 : -t0;         // sign of zeta (+ at zeta = 0)
  cs = rsq_nr(__builtin_fma(tt, tt, 1.0));
  sn = cs * tt;
}

template <int JE>
__device__ __forceinline__ void rotate4(double (&u)[JE], double (&w)[JE], double cs, double sn) {
#pragma unroll
  for (int e = 0; e < JE; ++e) {
    const double u0 = u[e], w0 = w[e];
    u[e] = cs * u0 - sn * w0;
    w[e] = sn * u0 + cs * w0;
  }
}

// (JE and F generalise it to longer rows and to matrices in global memory,
// where a row's tail past its end lies inside the padded workspace too.)
template <int G, int JE = JEPL, typename F = lds_f64>
__device__ __noinline__ bool jacobi_rows_reg(F* X, int ldx_in, int n_in, int b_in, F* J, int ldj_in,
                                             lds_i32* flag, double* red, double cut, int init_j_in,
                                             int max_sweeps_in) {
  const int ldx = uni(ldx_in), n = uni(n_in), b = uni(b_in), ldj = uni(ldj_in), init_j = uni(init_j_in);
  const int max_sweeps = uni(max_sweeps_in);
  const int tid = threadIdx.x, bs = blockDim.x;
  const int gid = tid / G, ng = bs / G;
  if (init_j)
    for (int e = tid; e < n * n; e += bs) J[(e / n) * ldj + e % n] = (e / n == e % n) ? 1.0 : 0.0;
  __syncthreads();
  if (n < 2) return true;
  const int nn = n + (n & 1);
  const double tol = DBL_EPSILON * (double)(b > n ? b : n);
  double fro = 0.0;
  for (int e = tid; e < n * b; e += bs) {
    const double x = X[(e / b) * ldx + e % b];
    fro += x * x;
  }
  const double fsum = block_sum(fro, red);
  const double small = fsum * DBL_EPSILON * DBL_EPSILON;
  const double noise = noise_level(fsum, cut, n);
  const double stop2 = jacobi_stop2(tol);
  QTT_JT_INIT;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) *flag = 0;
    __syncthreads();
    QTT_JC(6);
    for (int t = 0; t < nn - 1; ++t) {
      QTT_JT(4);
      QTT_JC(7);
      for (int pi = gid; pi < nn / 2; pi += ng) {
        int p = rr_player(pi, t, nn), q = rr_player(nn - 1 - pi, t, nn);
        if (p > q) {
          const int tmp = p;
          p = q;
          q = tmp;
        }
        if (q >= n) continue;
        const int gl = opaque(tid & (G - 1));
        F* xp = X + (int64_t)p * ldx;
        F* xq = X + (int64_t)q * ldx;
        F* jp = J + (int64_t)p * ldj;
        F* jq = J + (int64_t)q * ldj;
        double u[JE], w[JE], ju[JE], jw[JE];
        load_row<G, JE>(xp, b, gl, u);
        load_row<G, JE>(xq, b, gl, w);
        load_row<G, JE>(jp, n, gl, ju);  // read with X so both rows arrive in one wait
        load_row<G, JE>(jq, n, gl, jw);
        double a0 = 0.0, c0 = 0.0, g0 = 0.0, a1 = 0.0, c1 = 0.0, g1 = 0.0;
#pragma unroll
        for (int e = 0; e < JE; e += 2) {
          a0 = __builtin_fma(u[e], u[e], a0);
          c0 = __builtin_fma(w[e], w[e], c0);
          g0 = __builtin_fma(u[e], w[e], g0);
          a1 = __builtin_fma(u[e + 1], u[e + 1], a1);
          c1 = __builtin_fma(w[e + 1], w[e + 1], c1);
          g1 = __builtin_fma(u[e + 1], w[e + 1], g1);
        }
        const double a = jsum<G>(a0 + a1), c = jsum<G>(c0 + c1), g = jsum<G>(g0 + g1);
        QTT_JT(0);
        // |g| > tol sqrt(a c), compared squared (a c stays finite for core entries)
        if (g != 0.0 && a > small && c > small && (a > noise || c > noise) && g * g > tol * tol * (a * c)) {
          double cs, sn;
          jacobi_rotation(a, c, g, cs, sn);
          QTT_JT(1);
          rotate4(u, w, cs, sn);
          rotate4(ju, jw, cs, sn);
          store_row<G, JE>(xp, b, gl, u);
          store_row<G, JE>(xq, b, gl, w);
          store_row<G, JE>(jp, n, gl, ju);
          store_row<G, JE>(jq, n, gl, jw);
          if (gl == 0 && g * g > stop2 * (a * c)) *flag = 1;
          QTT_JT(2);
        }
      }
      QTT_JT(5);
      __syncthreads();
      QTT_JT(3);
    }
    const int rot = *flag;
    __syncthreads();
    if (!rot) {
      QTT_SWEEPS(sweep + 1);
      return true;
    }
  }
  QTT_SWEEPS(max_sweeps);
  return false;
}

// The one-sided Jacobi of jacobi_rows_reg for small matrices, run by the first
// wavefront alone: pair pi of each round goes to lanes [G pi, G pi + G), so
// every pair of a round is in flight at once when n <= 128 / G.  LDS
// operations of one wavefront complete in the order they are issued, so each
// round reads the rows the round before it wrote with no barrier between
// rounds; the other wavefronts wait at the barrier that ends the call.  A
// round then costs one LDS round trip plus the rotation, not a block-wide
// barrier behind the slowest wavefront.  The rows are padded with zeros to bp
// (X) and np (J) entries, multiples of G of at most G JE and G JJ, so every
// lane loads, rotates and stores whole blocks with no masks; the zeros stay
// zero.  Rotations, tests and stopping rule are those of jacobi_rows_reg;
// J must hold its start (the identity or the rotations so far).
template <int G, int JE>
__device__ __forceinline__ void wave_load(const lds_f64* r, int len, int gl, double (&x)[JE]) {
#pragma unroll
  for (int e = 0; e < JE; ++e) x[e] = (G * e < len) ? r[gl + G * e] : 0.0;
}

template <int G, int JE>
__device__ __forceinline__ void wave_store(lds_f64* r, int len, int gl, const double (&x)[JE]) {
#pragma unroll
  for (int e = 0; e < JE; ++e)
    if (G * e < len) r[gl + G * e] = x[e];
}

template <int G, int JE, int JJ>
__device__ __noinline__ bool jacobi_wave(lds_f64* X, int ldx_in, int n_in, int b_in, int bp_in, lds_f64* J,
                                         int ldj_in, int np_in, lds_i32* flag, double* red, double cut,
                                         int max_sweeps_in) {
  static_assert(JE % 2 == 0, "two partial sums per row");
  const int ldx = uni(ldx_in), n = uni(n_in), b = uni(b_in), bp = uni(bp_in), ldj = uni(ldj_in), np = uni(np_in);
  const int max_sweeps = uni(max_sweeps_in);
  const int tid = threadIdx.x, bs = blockDim.x;
  if (n < 2) return true;
  const int nn = n + (n & 1);
  const double tol = DBL_EPSILON * (double)(b > n ? b : n);
  double fro = 0.0;
  for (int e = tid; e < n * b; e += bs) {
    const double x = X[(e / b) * ldx + e % b];
    fro += x * x;
  }
  const double fsum = block_sum(fro, red);
  const double small = fsum * DBL_EPSILON * DBL_EPSILON;
  const double noise = noise_level(fsum, cut, n);
  const double stop2 = jacobi_stop2(tol);
  if (tid < 64) {
    const int pi = tid / G;
    const bool player = pi < nn / 2;
    int sweeps = max_sweeps, conv = 0;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
      int rot = 0;
      for (int t = 0; t < nn - 1; ++t) {
        int p = rr_player(pi, t, nn), q = rr_player(nn - 1 - pi, t, nn);
        if (p > q) {
          const int tmp = p;
          p = q;
          q = tmp;
        }
        if (player && q < n) {
          const int gl = tid & (G - 1);
          lds_f64* xp = X + p * ldx;
          lds_f64* xq = X + q * ldx;
          lds_f64* jp = J + p * ldj;
          lds_f64* jq = J + q * ldj;
          double u[JE], w[JE], ju[JJ], jw[JJ];
          wave_load<G, JE>(xp, bp, gl, u);
          wave_load<G, JE>(xq, bp, gl, w);
          wave_load<G, JJ>(jp, np, gl, ju);
          wave_load<G, JJ>(jq, np, gl, jw);
          double a0 = 0.0, c0 = 0.0, g0 = 0.0, a1 = 0.0, c1 = 0.0, g1 = 0.0;
#pragma unroll
          for (int e = 0; e < JE; e += 2) {
            a0 = __builtin_fma(u[e], u[e], a0);
            c0 = __builtin_fma(w[e], w[e], c0);
            g0 = __builtin_fma(u[e], w[e], g0);
            a1 = __builtin_fma(u[e + 1], u[e + 1], a1);
            c1 = __builtin_fma(w[e + 1], w[e + 1], c1);
            g1 = __builtin_fma(u[e + 1], w[e + 1], g1);
          }
          double a = a0 + a1, c = c0 + c1, g = g0 + g1;
          if constexpr (G == 4) {  // sums over the 4 lanes of a quad
            a += dpp_d<0xB1>(a);
            c += dpp_d<0xB1>(c);
            g += dpp_d<0xB1>(g);
            a += dpp_d<0x4E>(a);
            c += dpp_d<0x4E>(c);
            g += dpp_d<0x4E>(g);
          } else {
            a = jsum<G>(a);
            c = jsum<G>(c);
            g = jsum<G>(g);
          }
          if (g != 0.0 && a > small && c > small && (a > noise || c > noise) && g * g > tol * tol * (a * c)) {
            double cs, sn;
            jacobi_rotation(a, c, g, cs, sn);
            rotate4(u, w, cs, sn);
            rotate4(ju, jw, cs, sn);
            wave_store<G, JE>(xp, bp, gl, u);
            wave_store<G, JE>(xq, bp, gl, w);
            wave_store<G, JJ>(jp, np, gl, ju);
            wave_store<G, JJ>(jq, np, gl, jw);
            rot |= g * g > stop2 * (a * c) ? 1 : 0;
          }
        }
        // keep the next round's loads after this round's stores (the LDS
        // itself serves one wavefront's operations in order)
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      }
      if (__builtin_amdgcn_ballot_w64(rot != 0) == 0) {
        sweeps = sweep + 1;
        conv = 1;
        break;
      }
    }
    if (tid == 0) *flag = conv;
    QTT_SWEEPS(sweeps);
  }
  __syncthreads();
  const int cv = *flag;
  __syncthreads();
  return cv != 0;
}

// Lanes per pair of the single-wavefront Jacobi for an n x b matrix, 0 when it
// does not apply: 8 lanes with 4 entries of a row each, up to 16 rows of at
// most 32 entries.  (Longer rows make a round issue-bound on the one SIMD that
// runs the wavefront; the rounds spread over the block are then faster.)
__device__ __forceinline__ int wave_group(int n, int b) {
  return n <= 16 && b <= 32 ? 8 : 0;
}

// LDS the single-wavefront Jacobi needs for its padded copy of X and J
__device__ __forceinline__ int64_t wave_jacobi_lds(int n, int b) {
  const int gw = wave_group(n, b);
  if (!gw) return 0;
  const int bp = (b + gw - 1) / gw * gw, np = (n + gw - 1) / gw * gw;
  return (int64_t)n * ((bp + 1) + (np + 1)) + 16;
}

// ---------------------------------------------------------------------------
// Blocked Householder QR for matrices in global memory (the move's workspace
// does not fit LDS): panels of PNB columns are factored in registers (up to
// 1024 rows) or in place in global memory (taller), and the rest of the
// matrix is updated once per panel in the compact WY form
// H_0 ... H_{nb-1} = I - V T V^T (LAPACK dlarft / dlarfb) with FP64 MFMA
// tiles.  Q (or U of a tall SVD move) is formed the same way, panels in
// reverse order.
// ---------------------------------------------------------------------------
constexpr int PNB = 16;
constexpr int PLD = PNB + 1;  // row stride of a panel staged in LDS: odd, so no bank conflicts

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) f64x2 lds_f64x2;

// ---------------------------------------------------------------------------
// Blocked Householder QR for matrices too tall for an LDS panel.
//
// The panel (PNB columns) is factored in place in global memory, one pass over
// its rows per column: the pass applies reflector j and, in the same read,
// sums what reflector j+1 needs (the norm of column j+1, its dot products with
// the later columns, and with the earlier reflectors for T).  The rows are
// split over all threads and the sums reduced once per column.  The trailing
// columns, and later the Q (or U) being formed, are updated in the WY form
// with the rows split over the wavefronts and one column per lane; each
// wavefront stages 64 rows of V at a time in its own LDS region and reads
// them back as broadcasts.  T of every panel is kept in global memory for the
// Q formation.
//
// LDS (TALL_LDS doubles): per-wave regions 8 x 2048 | W final 16 x 128 |
// T 16 x 16 | sums 16 | reduction 8 x 16.
// ---------------------------------------------------------------------------
constexpr int TALL_REGION = 2048;
constexpr int TALL_MF = 16;  // rows of M in flight per wait in the WY updates (divides 64)
constexpr int TALL_LDS = 8 * TALL_REGION + 16 * 128 + PNB * PNB + 16 + 8 * 16;

// Sum the slot value of every lane over the lanes of its column (lane & 15)
// and over the workgroup into out[0..15] (all threads see them).
__device__ __forceinline__ void tall_reduce16(double acc, lds_f64* buf, lds_f64* out) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  acc += __shfl_xor(acc, 16);
  acc += __shfl_xor(acc, 32);
  if (lane < PNB) buf[wv * PNB + lane] = acc;
  __syncthreads();
  if (tid < PNB) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += buf[w * PNB + tid];
    out[tid] = t;
  }
  __syncthreads();
}

__device__ bool coop_post_run(CoopJob* job, int nwg, int* iflag, lds_f64* lw, int nact = 0);
__device__ int coop_post(CoopJob* job, int* iflag, int nact);
__device__ bool coop_wait(CoopJob* job, int s, int nwg, int* iflag);
__device__ void coop_wy_cols(CoopJob* job, int wg, int nwg, lds_f64* lw);
__device__ void coop_wy_rows(CoopJob* job, int wg, int nwg);
__device__ void gather_tiles(const double* X, double* W, int na, int b, int bx, int wg, int nwg, lds_f64* lw);

// The two kinds of pass over panel rows [r0, r1) (lane layout: a 16-lane group
// per row, a lane per column c = lane & 15; rows step by 4 per wavefront).
constexpr int PANEL_U = 8;  // row steps in flight per wait

// sums for column 0: slot 0 its norm^2, slot c its dot with column c
__device__ __forceinline__ double panel_pre_rows(const double* P, int ld, int nb, int r0, int r1) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, bs = blockDim.x;
  const int c = lane & 15, rstep = (bs >> 6) * 4, rsub = wv * 4 + (lane >> 4);
  const bool cin = c < nb;
  double acc = 0.0;
  for (int rb = r0; rb < r1; rb += PANEL_U * rstep) {
    double x[PANEL_U], x0[PANEL_U];  // this lane's column and column 0 (the same load for the whole group)
#pragma unroll
    for (int u = 0; u < PANEL_U; ++u) {
      const int r = rb + u * rstep + rsub;
      const bool in = r < r1;
      x[u] = (in && cin) ? P[(int64_t)r * ld + c] : 0.0;
      x0[u] = in ? P[(int64_t)r * ld] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < PANEL_U; ++u) acc = __builtin_fma(x0[u], x[u], acc);
  }
  return acc;
}

// apply reflector j (scale, this lane's dc; dn for column j+1) to rows
// [r0, r1) (all below row j) and return this lane's slot of the sums for
// column j+1: slot c > j+1 column c, c == j+1 its norm^2, c == j reflector j,
// c < j reflector c (rows below j+1 only)
__device__ __forceinline__ double panel_pass_rows(double* P, int ld, int nb, int j, double scale, double dc,
                                                  double dn, int r0, int r1) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, bs = blockDim.x;
  const int c = lane & 15, rstep = (bs >> 6) * 4, rsub = wv * 4 + (lane >> 4);
  const bool cin = c < nb, nxt = j + 1 < nb;
  double acc = 0.0;
  for (int rb = r0; rb < r1; rb += PANEL_U * rstep) {
    // this lane's column, and columns j and j+1 of the same row (one load for
    // the whole group each): no cross-lane traffic in the pass
    double x[PANEL_U], xj[PANEL_U], xj1[PANEL_U];
#pragma unroll
    for (int u = 0; u < PANEL_U; ++u) {
      const int r = rb + u * rstep + rsub;
      const bool in = r < r1;
      const double* row = P + (int64_t)r * ld;
      x[u] = (in && cin) ? row[c] : 0.0;
      xj[u] = in ? row[j] : 0.0;
      xj1[u] = (in && nxt) ? row[j + 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < PANEL_U; ++u) {
      const int r = rb + u * rstep + rsub;
      const double v = scale * xj[u];  // reflector j at row r
      double y = x[u];
      if (c > j) y = __builtin_fma(-dc, v, y);
      if (r < r1 && cin && c >= j) P[(int64_t)r * ld + c] = c == j ? v : y;
      const double xn = __builtin_fma(-dn, v, xj1[u]);
      const double o = c == j ? v : y;
      if (r >= j + 2 && r < r1 && nxt) acc = __builtin_fma(xn, o, acc);
    }
  }
  return acc;
}

// Factor the panel P = W[p0:, p0:p0+nb] (global, ld): R entries on and above
// its diagonal, reflectors below (unit diagonal implied), tau[p0 + j], and the
// panel's T into Tl (LDS, PNB x PNB; padding rows and columns zero).  Each
// wavefront reads four panel rows per load, a 16-lane group per row and a lane
// per column, so the loads are coalesced and each lane keeps the sum of its
// own column's slot; the group's lanes also read columns j and j+1 of their
// row (one broadcast load each) instead of exchanging them.  With helper
// workgroups (job), every pass is a job: each workgroup takes a range of rows
// and leaves its 16 sums in job->Gp, which workgroup 0 adds in a fixed order.
__device__ __noinline__ void tall_panel(double* W, int rows_in, int ld_in, int p0_in, int nb_in, double* tau,
                                        lds_f64* Tl, lds_f64* S, lds_f64* buf, CoopJob* job, int nwg_in,
                                        int* iflag, lds_f64* lw) {
  const int tid = threadIdx.x, bs = blockDim.x, lane = tid & 63;
  const int rows = uni(rows_in), ld = uni(ld_in), p0 = uni(p0_in), nb = uni(nb_in), nwg = uni(nwg_in);
  double* P = W + (int64_t)p0 * ld + p0;
  const int mr = rows - p0;
  const int c = lane & 15;  // this lane's column
  const bool cin = c < nb;
  const bool coop = job != nullptr && nwg > 1;
  // the sums over all rows of one pass into S
  auto gather = [&]() {
    if (tid < PNB) {
      double t = 0.0;
      for (int w = 0; w < nwg; ++w) t += job->Gp[w * PNB + tid];
      S[tid] = t;
    }
    __syncthreads();
  };
  for (int e = tid; e < PNB * PNB; e += bs) Tl[e] = 0.0;
  if (coop) {
    if (tid == 0) {
      job->type = 4;
      job->X = P;
      job->ldx = ld;
      job->n = mr;
      job->nb = nb;
    }
    coop_post_run(job, nwg, iflag, lw);
    gather();
  } else {
    tall_reduce16(panel_pre_rows(P, ld, nb, 1, mr), buf, S);
  }
  for (int j = 0; j < nb; ++j) {
    // slot j: norm^2 of column j below row j; slot c > j: its dot with column c;
    // slot i < j: its dot with reflector i (all over the rows below row j)
    const double* rj = P + (int64_t)j * ld;
    const double alpha = rj[j], sg = S[j];
    double tj = 0.0, beta = alpha, scale = 0.0;
    if (sg != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sg), alpha);
      tj = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    // this lane's column: tau_j (v_j . a_c) for c > j; and for column j+1,
    // needed by every lane (row j is rewritten below, so both are read now)
    const double dc = (c > j && cin) ? tj * (rj[c] + scale * S[c]) : 0.0;
    const double rjc = cin ? rj[c] : 0.0;
    const bool nxt = j + 1 < nb;
    const double dn = nxt ? tj * (rj[j + 1] + scale * S[j + 1]) : 0.0;
    // T[:, j] = -tau_j T[:, 0:j] y, y_i = v_i . v_j = v_i[j] + scale * slot i
    if (tid < j) {
      double acc = 0.0;
      for (int l = tid; l < j; ++l) acc = __builtin_fma(Tl[tid * PNB + l], rj[l] + scale * S[l], acc);
      Tl[tid * PNB + j] = -tj * acc;
    }
    __syncthreads();  // row j and the sums are read: both are rewritten below
    if (tid == 0) {
      Tl[j * PNB + j] = tj;
      tau[p0 + j] = tj;
      P[(int64_t)j * ld + j] = beta;
    }
    // row j of R right of the diagonal: a_c[j] - tau_j (v_j . a_c) (v_j[j] = 1)
    if (tid < PNB && c > j && cin) P[(int64_t)j * ld + c] = rjc - dc;
    if (coop) {
      if (tid < PNB) job->dcol[tid] = dc;
      if (tid == 0) {
        job->type = 5;
        job->X = P;
        job->ldx = ld;
        job->n = mr;
        job->nb = nb;
        job->t = j;
        job->tol = scale;
        job->small = dn;
      }
      coop_post_run(job, nwg, iflag, lw);
      if (nxt) gather();
    } else {
      const double acc = panel_pass_rows(P, ld, nb, j, scale, dc, dn, j + 1, mr);
      if (nxt) tall_reduce16(acc, buf, S);
    }
  }
  __syncthreads();
}

// M[0:mr, c0:c1] <- (I - V op(T) V^T) M[0:mr, c0:c1], op(T) = T^T (trans) or T.
// V = the panel's reflectors (global, ldv; nb columns, unit diagonal implied,
// zero above), M global with leading dimension ldm.  Both products are FP64
// MFMA tiles (v_mfma_f64_16x16x4f64: lane l holds A[l%16][l/16] and
// B[l/16][l%16], and D[l/16 + 4e][l%16] in element e), so the operands come
// from coalesced loads of 16 consecutive doubles per 16-lane group, not from
// LDS broadcasts.
//   W = V^T M: per (16-column tile, row chunk) unit, A = V^T, B = M, 4 rows a step;
//   U = op(T) W (16 x 16 per tile, in LDS);
//   M -= V U: per (tile, 16-row block), C = M, A = V, B = -U, 4 reflectors a step.

__device__ __forceinline__ double vref(const double* V, int ldv, int r, int j, int mr, int nb) {
  return (j < nb && r < mr) ? (r > j ? V[(int64_t)r * ldv + j] : (r == j ? 1.0 : 0.0)) : 0.0;
}

// W partial of one (tile, row chunk) unit: rows [rs, re), 16 columns from
// col0 (columns past c1 read as zero).  Returns this lane's 4 entries.
__device__ __forceinline__ f64x4 wy_p1_unit(const double* M, int ldm, int mr, int col0, int c1, const double* V,
                                            int ldv, int nb, int rs, int re) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int col = col0 + li;
  const bool con = col < c1;
  f64x4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  int r = rs;
  for (; r + 16 <= re; r += 16) {  // 16 rows (four steps) in flight per wait
    double a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = r + 4 * q + lk;
      a[q] = vref(V, ldv, rr, li, mr, nb);
      b[q] = con ? M[(int64_t)rr * ldm + col] : 0.0;
    }
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b[1], acc1, 0, 0, 0);
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[2], b[2], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[3], b[3], acc1, 0, 0, 0);
  }
  for (; r < re; r += 4) {
    const int rr = r + lk;
    const bool rin = rr < re;
    const double a = rin ? vref(V, ldv, rr, li, mr, nb) : 0.0;
    const double b = (rin && con) ? M[(int64_t)rr * ldm + col] : 0.0;
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc0, 0, 0, 0);
  }
  return acc0 + acc1;
}

// M -= V U on one (tile, 16-row block) unit; Ut holds -U of the tile as
// [j][c] (16 x 16), in LDS or global memory.
template <typename UP>
__device__ __forceinline__ void wy_p3_unit(double* M, int ldm, int mr, int col0, int c1, const double* V, int ldv,
                                           int nb, int rb, UP Ut) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int col = col0 + li;
  const bool con = col < c1;
  double bu[4], a[4];
  f64x4 cm;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    bu[q] = Ut[(4 * q + lk) * 16 + li];               // B = -U[4q + k][c]
    a[q] = vref(V, ldv, rb + li, 4 * q + lk, mr, nb);  // A = V[row][4q + k]
    const int rr = rb + lk + 4 * q;
    cm[q] = (con && rr < mr) ? M[(int64_t)rr * ldm + col] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) cm = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], bu[q], cm, 0, 0, 0);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int rr = rb + lk + 4 * e;
    if (con && rr < mr) M[(int64_t)rr * ldm + col] = cm[e];
  }
}

// Two wy_p3_unit's with all loads of both issued before the MFMAs.
template <typename UP>
__device__ __forceinline__ void wy_p3_unit2(double* M, int ldm, int mr, int col0a, int rba, UP Uta, int col0b,
                                            int rbb, UP Utb, int c1, const double* V, int ldv, int nb) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int cola = col0a + li, colb = col0b + li;
  const bool cona = cola < c1, conb = colb < c1;
  double bua[4], aa[4], bub[4], ab[4];
  f64x4 cma, cmb;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    bua[q] = Uta[(4 * q + lk) * 16 + li];
    bub[q] = Utb[(4 * q + lk) * 16 + li];
    aa[q] = vref(V, ldv, rba + li, 4 * q + lk, mr, nb);
    ab[q] = vref(V, ldv, rbb + li, 4 * q + lk, mr, nb);
    const int ra = rba + lk + 4 * q, rb = rbb + lk + 4 * q;
    cma[q] = (cona && ra < mr) ? M[(int64_t)ra * ldm + cola] : 0.0;
    cmb[q] = (conb && rb < mr) ? M[(int64_t)rb * ldm + colb] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    cma = __builtin_amdgcn_mfma_f64_16x16x4f64(aa[q], bua[q], cma, 0, 0, 0);
    cmb = __builtin_amdgcn_mfma_f64_16x16x4f64(ab[q], bub[q], cmb, 0, 0, 0);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int ra = rba + lk + 4 * e, rb = rbb + lk + 4 * e;
    if (cona && ra < mr) M[(int64_t)ra * ldm + cola] = cma[e];
    if (conb && rb < mr) M[(int64_t)rb * ldm + colb] = cmb[e];
  }
}

constexpr int64_t WY_COOP_MIN = int64_t(1) << 15;  // entries of a row-chunk (two-job) shared update
constexpr int64_t WY_COLS_MIN = int64_t(1) << 13;  // entries of a column-tile (one-job) shared update

__device__ __noinline__ void tall_wy(double* M, int ldm_in, int mr_in, int c0_in, int c1_in, const double* V,
                                     int ldv_in, int nb_in, const lds_f64* Tl, bool trans, lds_f64* L,
                                     CoopJob* job, int nwg_in, int* iflag) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwv = blockDim.x >> 6;
  const int ldm = uni(ldm_in), mr = uni(mr_in), c0 = uni(c0_in), c1 = uni(c1_in), ldv = uni(ldv_in),
            nb = uni(nb_in), nwg = uni(nwg_in);
  // helper workgroups only for a large trailing matrix: below WY_COOP_MIN
  // (WY_COLS_MIN for a column-tile job) entries the hand-offs of a shared
  // update cost more than the update on workgroup 0 alone
  const int64_t ent = (int64_t)mr * (c1 - c0);
  const bool coop = job != nullptr && nwg > 1 && ent >= (mr <= 1024 ? WY_COLS_MIN : WY_COOP_MIN);
  if (coop && mr <= 1024) {  // whole column tiles per workgroup: one job
    if (tid == 0) {
      job->type = 11;
      job->wy_M = M;
      job->wy_ldm = ldm;
      job->wy_mr = mr;
      job->wy_cg = c0;
      job->wy_c1 = c1;
      job->wy_V = V;
      job->wy_ldv = ldv;
      job->nb = nb;
      job->wy_trans = trans ? 1 : 0;
    }
    for (int e = tid; e < PNB * PNB; e += blockDim.x) job->Ug[e] = Tl[e];
    QTT_PT_INIT;
    const int s = coop_post(job, iflag, min(nwg, (c1 - c0 + 15) >> 4));
    coop_wy_cols(job, 0, nwg, L);
    coop_wait(job, s, nwg, iflag);
    QTT_PT(11);
    if (threadIdx.x == 0) QTT_WY_COUNT_COLS;
    return;
  }
  lds_f64* Pw = L;                  // partial W tiles: [unit][e * 64 + lane]
  lds_f64* Wt = L + 32 * 256;       // W tiles [tile][j][c]
  lds_f64* Ut = Wt + 8 * 256;       // U tiles [tile][j][c]
  QTT_PT_INIT;
  if (threadIdx.x == 0) QTT_WY_COUNT(coop);
  for (int cg = c0; cg < c1; cg += 128) {
    const int nct = (min(128, c1 - cg) + 15) >> 4;  // 16-column tiles in this group
    // row chunks per tile: enough units for every wavefront (of all workgroups with helpers)
    const int nch = coop ? max(1, min(min((nwg * nwv + nct - 1) / nct, (mr + 63) / 64), COOP_PG_UNITS / nct))
                         : max(1, min(32 / nct, (mr + 63) / 64));
    const int nunit = nct * nch, chunk = (((mr + nch - 1) / nch) + 3) & ~3;
    // W = V^T M
    if (coop) {
      if (tid == 0) {
        job->type = 6;
        job->wy_M = M;
        job->wy_ldm = ldm;
        job->wy_mr = mr;
        job->wy_cg = cg;
        job->wy_c1 = c1;
        job->wy_V = V;
        job->wy_ldv = ldv;
        job->nb = nb;
        job->wy_nct = nct;
        job->wy_nch = nch;
        job->wy_chunk = chunk;
      }
      {
        const int s = coop_post(job, iflag, nwg);
        coop_wy_rows(job, 0, nwg);
        coop_wait(job, s, nwg, iflag);
      }
    } else {
      for (int u = wv; u < nunit; u += nwv) {
        const int t = u / nch, h = u - t * nch;
        const int rs = h * chunk, re = min(mr, rs + chunk);
        const f64x4 acc = wy_p1_unit(M, ldm, mr, cg + t * 16, c1, V, ldv, nb, rs, re);
#pragma unroll
        for (int e = 0; e < 4; ++e) Pw[u * 256 + e * 64 + lane] = acc[e];
      }
      __syncthreads();
    }
    // reduce the row chunks: W[t][j][c] with j = lane/16 + 4e, c = lane%16
    for (int x = tid; x < nct * 256; x += blockDim.x) {
      const int t = x >> 8, y = x & 255;
      const int e = y >> 6, ln = y & 63;
      double sum = 0.0;
      if (coop)
        for (int h = 0; h < nch; ++h) sum += job->Pg[(int64_t)(t * nch + h) * 256 + y];
      else
        for (int h = 0; h < nch; ++h) sum += Pw[(t * nch + h) * 256 + y];
      Wt[t * 256 + ((ln >> 4) + 4 * e) * 16 + (ln & 15)] = sum;
    }
    __syncthreads();
    QTT_PT(coop ? 2 : 8);
    // U = op(T) W, stored negated for the update
    for (int x = tid; x < nct * 256; x += blockDim.x) {
      const int t = x >> 8, i = (x >> 4) & 15, c = x & 15;
      double sum = 0.0;
#pragma unroll
      for (int j = 0; j < PNB; ++j) sum = __builtin_fma(trans ? Tl[j * PNB + i] : Tl[i * PNB + j], Wt[t * 256 + j * 16 + c], sum);
      Ut[t * 256 + i * 16 + c] = -sum;
      if (coop) job->Ug[t * 256 + i * 16 + c] = -sum;
    }
    __syncthreads();
    QTT_PT(coop ? 3 : 9);
    // M -= V U, one (tile, 16-row block) at a time
    const int nrb = (mr + 15) >> 4;
    if (coop) {
      if (tid == 0) job->type = 7;  // the other wy_* fields stay from the first phase
      {
        const int s = coop_post(job, iflag, nwg);
        coop_wy_rows(job, 0, nwg);
        coop_wait(job, s, nwg, iflag);
      }
    } else {
      for (int x = wv; x < nct * nrb; x += nwv) {
        const int t = x / nrb, rb = (x - t * nrb) * 16;
        wy_p3_unit(M, ldm, mr, cg + t * 16, c1, V, ldv, nb, rb, (const lds_f64*)(Ut + t * 256));
      }
      __syncthreads();
    }
    QTT_PT(coop ? 5 : 10);
  }
}

// Two panels' trailing update in one pass over M, on this workgroup alone:
// M <- (I - V2 T2^T V2^T)(I - V1 T1^T V1^T) M for M[0:mr, c0:c1] (global,
// ldm).  V1: the first panel's reflectors (rows from 0), V2: the second's
// (rows from 16, so zero above), both with leading dimension ldv.  With
// W1 = V1^T M and W2 = V2^T M read in the same pass, U1 = T1^T W1,
// U2 = T2^T (W2 - G U1) with G = V2^T V1, and M -= V1 U1 + V2 U2 in a second
// pass: M is read twice and written once per 32 columns instead of per 16.
// L: LDS [0, 16384) for the partial and final tiles; T1, G: 256 doubles each.
__device__ __forceinline__ void wy2_p1_unit(const double* M, int ldm, int mr, int col0, int c1, const double* V1,
                                            int nb1, const double* V2, int nb2, int ldv, int rs, int re, f64x4& w1,
                                            f64x4& w2) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int col = col0 + li;
  const bool con = col < c1;
  f64x4 a1 = {0.0, 0.0, 0.0, 0.0}, a2 = {0.0, 0.0, 0.0, 0.0};
  for (int r = rs; r < re; r += 16) {  // 16 rows (four steps) in flight per wait
    double v1[4], v2[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = r + 4 * q + lk;
      const bool rin = rr < re;
      v1[q] = rin ? vref(V1, ldv, rr, li, mr, nb1) : 0.0;
      v2[q] = rin ? vref(V2, ldv, rr - PNB, li, mr - PNB, nb2) : 0.0;
      b[q] = (rin && con) ? M[(int64_t)rr * ldm + col] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(v1[q], b[q], a1, 0, 0, 0);
      a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(v2[q], b[q], a2, 0, 0, 0);
    }
  }
  w1 = a1;
  w2 = a2;
}

__device__ __noinline__ void tall_wy2(double* M, int ldm_in, int mr_in, int c0_in, int c1_in, const double* V1,
                                      const double* V2, int ldv_in, int nb1_in, int nb2_in, const lds_f64* T1,
                                      const lds_f64* T2, lds_f64* G, lds_f64* L, bool trans, bool rev) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwv = blockDim.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int ldm = uni(ldm_in), mr = uni(mr_in), c0 = uni(c0_in), c1 = uni(c1_in), ldv = uni(ldv_in),
            nb1 = uni(nb1_in), nb2 = uni(nb2_in);
  lds_f64* P1 = L;                // [unit][e * 64 + lane], 16 units
  lds_f64* P2 = L + 16 * 256;
  lds_f64* W1 = L + 32 * 256;     // [tile][j][c], 8 tiles
  lds_f64* W2 = L + 40 * 256;
  lds_f64* U1 = L + 48 * 256;     // -U1 [tile][j][c]
  lds_f64* U2 = L + 56 * 256;     // -U2
  // G = V2^T V1 over the rows from 16: partials per wavefront in P1, summed
  {
    f64x4 g = {0.0, 0.0, 0.0, 0.0};
    for (int r = PNB + 4 * wv; r < mr; r += 4 * nwv) {
      const int rr = r + lk;
      const double a = rr < mr ? vref(V2, ldv, rr - PNB, li, mr - PNB, nb2) : 0.0;
      const double bb = rr < mr ? vref(V1, ldv, rr, li, mr, nb1) : 0.0;
      g = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, g, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) P1[wv * 256 + e * 64 + lane] = g[e];
    __syncthreads();
    if (tid < 256) {  // element e of lane ln is G[ln / 16 + 4 e][ln % 16]
      const int e = tid >> 6, ln = tid & 63;
      double sum = 0.0;
      for (int w = 0; w < nwv; ++w) sum += P1[w * 256 + tid];
      G[((ln >> 4) + 4 * e) * 16 + (ln & 15)] = sum;
    }
    __syncthreads();
  }
  for (int cg = c0; cg < c1; cg += 128) {
    const int nct = (min(128, c1 - cg) + 15) >> 4;
    const int nch = max(1, min(16 / nct, (mr + 63) / 64));
    const int nunit = nct * nch, chunk = (((mr + nch - 1) / nch) + 15) & ~15;
    for (int u = wv; u < nunit; u += nwv) {
      const int t = u / nch, h = u - t * nch;
      const int rs = h * chunk, re = min(mr, rs + chunk);
      f64x4 w1, w2;
      wy2_p1_unit(M, ldm, mr, cg + t * 16, c1, V1, nb1, V2, nb2, ldv, rs, re, w1, w2);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        P1[u * 256 + e * 64 + lane] = w1[e];
        P2[u * 256 + e * 64 + lane] = w2[e];
      }
    }
    __syncthreads();
    for (int x = tid; x < nct * 256; x += blockDim.x) {
      const int t = x >> 8, y = x & 255;
      const int e = y >> 6, ln = y & 63;
      double s1 = 0.0, s2 = 0.0;
      for (int h = 0; h < nch; ++h) {
        s1 += P1[(t * nch + h) * 256 + y];
        s2 += P2[(t * nch + h) * 256 + y];
      }
      const int o = t * 256 + ((ln >> 4) + 4 * e) * 16 + (ln & 15);
      W1[o] = s1;
      W2[o] = s2;
    }
    __syncthreads();
    // the panel applied first (V1, or V2 if rev): U_f = op(T_f) W_f; then
    // W_s <- W_s - (V_s^T V_f) U_f and U_s = op(T_s) W_s for the second, with
    // V_s^T V_f = G (or G^T if rev); op(T) = T^T if trans.  U stored negated.
    const lds_f64* Tf = rev ? T2 : T1;
    const lds_f64* Ts = rev ? T1 : T2;
    lds_f64* Wf = rev ? W2 : W1;
    lds_f64* Ws = rev ? W1 : W2;
    lds_f64* Uf = rev ? U2 : U1;
    lds_f64* Us = rev ? U1 : U2;
    for (int x = tid; x < nct * 256; x += blockDim.x) {
      const int t = x >> 8, i = (x >> 4) & 15, c = x & 15;
      double sum = 0.0;
#pragma unroll
      for (int j = 0; j < PNB; ++j)
        sum = __builtin_fma(trans ? Tf[j * PNB + i] : Tf[i * PNB + j], Wf[t * 256 + j * 16 + c], sum);
      Uf[t * 256 + i * 16 + c] = -sum;
    }
    __syncthreads();
    for (int x = tid; x < nct * 256; x += blockDim.x) {  // Uf is stored negated
      const int t = x >> 8, i = (x >> 4) & 15, c = x & 15;
      double sum = Ws[t * 256 + i * 16 + c];
#pragma unroll
      for (int j = 0; j < PNB; ++j)
        sum = __builtin_fma(rev ? G[j * 16 + i] : G[i * 16 + j], Uf[t * 256 + j * 16 + c], sum);
      P2[t * 256 + i * 16 + c] = sum;
    }
    __syncthreads();
    for (int x = tid; x < nct * 256; x += blockDim.x) {
      const int t = x >> 8, i = (x >> 4) & 15, c = x & 15;
      double sum = 0.0;
#pragma unroll
      for (int j = 0; j < PNB; ++j)
        sum = __builtin_fma(trans ? Ts[j * PNB + i] : Ts[i * PNB + j], P2[t * 256 + j * 16 + c], sum);
      Us[t * 256 + i * 16 + c] = -sum;
    }
    __syncthreads();
    // M -= V1 U1 + V2 U2, one (tile, 16-row block) at a time
    const int nrb = (mr + 15) >> 4;
    for (int x = wv; x < nct * nrb; x += nwv) {
      const int t = x / nrb, rb = (x - t * nrb) * 16;
      const int col = cg + t * 16 + li;
      const bool con = col < c1;
      double b1[4], b2[4], a1[4], a2[4];
      f64x4 cm;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        b1[q] = U1[t * 256 + (4 * q + lk) * 16 + li];
        b2[q] = U2[t * 256 + (4 * q + lk) * 16 + li];
        a1[q] = vref(V1, ldv, rb + li, 4 * q + lk, mr, nb1);
        a2[q] = vref(V2, ldv, rb + li - PNB, 4 * q + lk, mr - PNB, nb2);
        const int rr = rb + lk + 4 * q;
        cm[q] = (con && rr < mr) ? M[(int64_t)rr * ldm + col] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) cm = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[q], b1[q], cm, 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 4; ++q) cm = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[q], b2[q], cm, 0, 0, 0);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int rr = rb + lk + 4 * e;
        if (con && rr < mr) M[(int64_t)rr * ldm + col] = cm[e];
      }
    }
    __syncthreads();
  }
}

// QR of W (global rows x cols, ld) for the tall case: as qr_panel, plus every
// panel's T in Tg (PNB x PNB each).
__device__ __noinline__ void qr_tall(double* W, int rows_in, int cols_in, int ld_in, double* tau, double* Tg,
                                     lds_f64* L, CoopJob* job, int nwg_in, int* iflag) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int rows = uni(rows_in), cols = uni(cols_in), ld = uni(ld_in), nwg = uni(nwg_in);
  const int k = rows < cols ? rows : cols;
  lds_f64* Tl = L + 8 * TALL_REGION + 16 * 128;
  lds_f64* S = Tl + PNB * PNB;
  lds_f64* buf = S + 16;
  for (int p0 = 0; p0 < k; p0 += PNB) {
    const int nb = k - p0 < PNB ? k - p0 : PNB;
    tall_panel(W, rows, ld, p0, nb, tau, Tl, S, buf, job, nwg, iflag, L);
    for (int e = tid; e < PNB * PNB; e += bs) Tg[(p0 / PNB) * PNB * PNB + e] = Tl[e];
    if (p0 + nb < cols)
      tall_wy(W + (int64_t)p0 * ld, ld, rows - p0, p0 + nb, cols, W + (int64_t)p0 * ld + p0, ld, nb, Tl, true, L,
              job, nwg, iflag);
    __syncthreads();
  }
}

// x <- Q (rows x k) of qr_tall's reflectors, or with Jb x <- U = Q [J_perm^T; 0]
// (rows x m; perm null: the rows of J in order), formed row-major in Mq
// (global) and then written into the core Xc (rows split as (aa, bb), bb < bx).
// W is read until the last panel is applied, so Xc may be W's memory.
__device__ __noinline__ void q_form_tall(const double* W, int rows_in, int k_in, int ld_in, const double* Tg,
                                         double* Mq, double* Xc, int bx_in, const double* Jb, int n_in,
                                         const int* perm, int m_in, lds_f64* L, CoopJob* job, int nwg_in,
                                         int* iflag) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int rows = uni(rows_in), k = uni(k_in), ld = uni(ld_in), bx = uni(bx_in), n = uni(n_in), m = uni(m_in),
            nwg = uni(nwg_in);
  const int ncol = Jb ? m : k;
  lds_f64* Tl = L + 8 * TALL_REGION + 16 * 128;
  {  // a row per wavefront, the columns over its lanes (no divisions per entry)
    const int lane = tid & 63, wv = tid >> 6, nwv = bs >> 6;
    for (int r = wv; r < rows; r += nwv)
      for (int c = lane; c < ncol; c += 64)
        Mq[(int64_t)r * ncol + c] =
            Jb ? (r < n ? Jb[(int64_t)(perm ? perm[c] : c) * n + r] : 0.0) : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  // panels in pairs (tall_wy2, one pass over Mq per two panels) where every
  // update would run on this workgroup alone anyway
  const bool pairs = !(job != nullptr && nwg > 1) || (int64_t)rows * ncol < WY_COLS_MIN;
  lds_f64* T1 = L + 8 * TALL_REGION + 512;  // the earlier panel's T of a pair
  lds_f64* Gp = L + 8 * TALL_REGION + 768;  // V_later^T V_earlier
  for (int p0 = ((k - 1) / PNB) * PNB; p0 >= 0;) {
    const int nb = k - p0 < PNB ? k - p0 : PNB;
    for (int e = tid; e < PNB * PNB; e += bs) Tl[e] = Tg[(p0 / PNB) * PNB * PNB + e];
    if (pairs && p0 >= PNB) {
      // panels pa = p0 - 16 (V1) and p0 (V2, 16 rows further), p0's applied first
      const int pa = p0 - PNB;
      for (int e = tid; e < PNB * PNB; e += bs) T1[e] = Tg[(pa / PNB) * PNB * PNB + e];
      __syncthreads();
      tall_wy2(Mq + (int64_t)pa * ncol, ncol, rows - pa, Jb ? 0 : pa, ncol, W + (int64_t)pa * ld + pa,
               W + (int64_t)p0 * ld + p0, ld, PNB, nb, T1, Tl, Gp, L, false, true);
      p0 = pa - PNB;
    } else {
      __syncthreads();
      // Q: only columns >= p0 differ from the identity yet; U: every column
      tall_wy(Mq + (int64_t)p0 * ncol, ncol, rows - p0, Jb ? 0 : p0, ncol, W + (int64_t)p0 * ld + p0, ld, nb, Tl,
              false, L, job, nwg, iflag);
      p0 -= PNB;
    }
  }
  // Xc[aa] (ncol x bx) = Mq[aa]^T (Mq[aa]: bx x ncol): tiled transposes, with
  // the helper workgroups for a large result
  if (bx == 1) {
    for (int e = tid; e < rows * ncol; e += bs) Xc[e] = Mq[e];
  } else if (job != nullptr && nwg > 1 && (int64_t)rows * ncol >= (int64_t(1) << 15)) {
    if (tid == 0) {
      job->type = 10;
      job->U = Mq;
      job->D = Xc;
      job->Au = rows / bx;
      job->rb = bx;
      job->Bu = ncol;
    }
    const int64_t ntile = (int64_t)(rows / bx) * ((bx + 63) >> 6) * ((ncol + 63) >> 6);
    coop_post_run(job, nwg, iflag, L, (int)min<int64_t>(ntile, nwg));
  } else {
    gather_tiles(Mq, Xc, rows / bx, bx, ncol, 0, 1, L);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Panel factorization in registers (panels of up to 512 RPT rows, 512
// threads): thread t holds rows t + 512 k (k < RPT) of the panel's PNB
// columns.  Step j applies reflector j to those rows and, in the same pass,
// forms this thread's share of what step j+1 needs: s_c = the sum over the
// rows r > j+1 of a[r][j+1] a[r][c], for every column c -- for c > j+1 the
// dot products with column j+1, for c == j+1 its norm^2, and for c <= j its
// dot products with the earlier reflectors (stored in place), which give
// column j+1 of T.  The 16 sums are reduced within each wavefront by four
// halving exchanges (every lane ends with the sum of one column over its
// quad) and across the wavefronts through LDS: one barrier per column.
// Row j+1, read by every thread in step j+1, travels through LDS with them.
// ---------------------------------------------------------------------------

// column whose sum lane l holds after wave_sum16: 8 b5 + 4 b4 + 2 b3 + b2
__device__ __forceinline__ int sum16_col(int lane) {
  return ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
}

// v_permlane32_swap on both halves of a and b, then the sum: lanes 0-31 get
// a + a of lanes 32-63, lanes 32-63 get b + b of lanes 0-31.
__device__ __forceinline__ double swap32_add(double a, double b) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
  return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}
// v_permlane16_swap (odd 16-lane rows of a with even rows of b), then the
// sum: rows 0 and 2 get a + a of the next row, rows 1 and 3 b + b of the row before.
__device__ __forceinline__ double swap16_add(double a, double b) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
  return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// Sum p[0..15] over the wavefront; lane l returns the sum of column
// sum16_col(l).  Each exchange pairs lanes that hold the same column set and
// differ in one lane bit (5, 4, 3, 2): the lane keeps the half of its set
// that the bit selects and adds the partner's copy of that half.  The last
// two steps sum the quad.  Fixed order: the same sums on every run.
__device__ __forceinline__ double wave_sum16(const double (&p)[PNB]) {
  const int lane = threadIdx.x & 63;
  const bool h3 = lane & 8, h2 = lane & 4;
  double q8[8], q4[4], q2[2];
#pragma unroll
  for (int i = 0; i < 8; ++i) q8[i] = swap32_add(p[i], p[i + 8]);   // bit 5: columns i / i + 8
#pragma unroll
  for (int i = 0; i < 4; ++i) q4[i] = swap16_add(q8[i], q8[i + 4]);  // bit 4: + 0 / + 4
#pragma unroll
  for (int i = 0; i < 2; ++i) {  // row_mirror: lane 15 - m of the 16-lane row (bit 3 differs)
    const double keep = h3 ? q4[i + 2] : q4[i], send = h3 ? q4[i] : q4[i + 2];
    q2[i] = keep + dpp_d<0x140>(send);
  }
  // row_half_mirror: lane 7 - m of the 8-lane half (bit 2 differs, bit 3 kept)
  double q = (h2 ? q2[1] : q2[0]) + dpp_d<0x141>(h2 ? q2[0] : q2[1]);
  q += dpp_d<0xB1>(q);  // quad_perm [1,0,3,2]
  q += dpp_d<0x4E>(q);  // quad_perm [2,3,0,1]
  return q;
}

// The 16 column sums of p over the workgroup (8 wavefronts); lane l returns
// the sum of column l & 15.  red: 8 x PNB doubles of LDS.  Every wavefront
// adds the 8 partials in the same order, so all threads see the same sums.
__device__ __forceinline__ double panel_sums16(const double (&p)[PNB], lds_f64* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double q = wave_sum16(p);
  if ((lane & 3) == 0) red[wv * PNB + sum16_col(lane)] = q;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < 8; ++w) t += red[w * PNB + (lane & 15)];
  return t;
}

// Factor the panel P = W[p0:, p0:p0+nb] (global, ld; mr = rows - p0 rows,
// nb <= mr): R on and above its diagonal, the reflectors below (unit diagonal
// implied), tau[0..nb), and the panel's T (PNB x PNB, zero rows and columns
// past nb) into Tl (LDS) and Tg.  L: LDS, the panel staged through its first
// mr * PLD doubles (coalesced global reads and writes); red: 2 x 8 x PNB sums
// and 2 x PNB row doubles of LDS, disjoint from the staging area and Tl.
template <int RPT>
__device__ __noinline__ void reg_panel(double* W, int rows_in, int ld_in, int p0_in, int nb_in, double* tau,
                                       double* Tg, lds_f64* Tl, lds_f64* L, lds_f64* red) {
  const int tid = threadIdx.x;
  const int rows = uni(rows_in), ld = uni(ld_in), p0 = uni(p0_in), nb = uni(nb_in);
  const int mr = rows - p0, ne = mr * PNB;
  double* P = W + (int64_t)p0 * ld + p0;
  lds_f64* rowb = red + 2 * 8 * PNB;
  // global -> LDS (8 loads in flight per thread) -> registers
  for (int e0 = 0; e0 < ne; e0 += 8 * 512) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * 512 + tid, r = e >> 4, c = e & 15;
      t[u] = (e < ne && c < nb) ? P[(int64_t)r * ld + c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * 512 + tid;
      if (e < ne) L[(e >> 4) * PLD + (e & 15)] = t[u];
    }
  }
  __syncthreads();
  double x[RPT][PNB];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int r = tid + 512 * k;
#pragma unroll
    for (int c = 0; c < PNB; ++c) x[k][c] = r < mr ? L[r * PLD + c] : 0.0;
  }
  // column layout: lane l holds, for column l & 15, the sum S (over the
  // rows below the current row j) and row j's entry R
  double S, R, Trow[PNB];
  const int cl = threadIdx.x & 15;
  {
    double p[PNB];
#pragma unroll
    for (int c = 0; c < PNB; ++c) p[c] = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const double a = (tid + 512 * k > 0) ? x[k][0] : 0.0;
#pragma unroll
      for (int c = 0; c < PNB; ++c) p[c] = __builtin_fma(a, x[k][c], p[c]);
    }
    if (tid == 0) {
#pragma unroll
      for (int c = 0; c < PNB; ++c) rowb[c] = x[0][c];
    }
    S = panel_sums16(p, red);
    R = rowb[cl];
  }
  const int ti = tid & 15;  // the row of T this thread keeps (written by threads 0..15)
#pragma unroll
  for (int j = 0; j < PNB; ++j) {
    // reflector j from alpha = a[j][j] and S[j] = the norm^2 below it (zero
    // columns past nb give tau 0)
    const double alpha = readlane_d(R, j), sg = readlane_d(S, j);
    double tj = 0.0, beta = alpha, scale = 0.0;
    if (sg != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sg), alpha);
      tj = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    // z_c = v_j . a_c = a[j][c] + scale * S[c]: for c < j the y_c of T (a_c
    // holds reflector c there), for c > j, times tau_j, what reflector j
    // takes from column c
    double zl = __builtin_fma(scale, S, R);
    zl = cl > j ? tj * zl : (cl < j ? zl : 0.0);
    double z[PNB];
#pragma unroll
    for (int c = 0; c < PNB; ++c) z[c] = readlane_d(zl, c);
    // T[:, j] = -tau_j T[:, 0:j] y
    {
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < PNB; ++l)
        if (l < j) acc = __builtin_fma(Trow[l], z[l], acc);
      Trow[j] = ti == j ? tj : (ti < j ? -tj * acc : 0.0);
    }
    if (tid == 0 && j < nb) tau[j] = tj;
    double p[PNB];
#pragma unroll
    for (int c = 0; c < PNB; ++c) p[c] = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int r = tid + 512 * k;
      const double v = r > j ? scale * x[k][j] : (r == j ? 1.0 : 0.0);
#pragma unroll
      for (int c = 0; c < PNB; ++c)  // constant trip counts: every loop unrolls, the arrays stay in registers
        if (c > j) x[k][c] = __builtin_fma(-z[c], v, x[k][c]);
      x[k][j] = r > j ? v : (r == j ? beta : x[k][j]);
      if (j + 1 < PNB) {
        const double a = r > j + 1 ? x[k][j + 1] : 0.0;
#pragma unroll
        for (int c = 0; c < PNB; ++c) p[c] = __builtin_fma(a, x[k][c], p[c]);
      }
    }
    if (j + 1 < PNB) {
      lds_f64* rb = rowb + ((j + 1) & 1) * PNB;
      if (tid == j + 1) {
#pragma unroll
        for (int c = 0; c < PNB; ++c) rb[c] = x[0][c];
      }
      S = panel_sums16(p, red + ((j + 1) & 1) * 8 * PNB);
      R = rb[cl];
    }
  }
  if (tid < PNB) {
#pragma unroll
    for (int c = 0; c < PNB; ++c) {
      Tl[tid * PNB + c] = Trow[c];
      Tg[tid * PNB + c] = Trow[c];
    }
  }
  // registers -> LDS -> global
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int r = tid + 512 * k;
    if (r < mr) {
#pragma unroll
      for (int c = 0; c < PNB; ++c) L[r * PLD + c] = x[k][c];
    }
  }
  __syncthreads();
  for (int e = tid; e < ne; e += 512) {
    const int r = e >> 4, c = e & 15;
    if (c < nb) P[(int64_t)r * ld + c] = L[r * PLD + c];
  }
  __syncthreads();
}

// QR of W (global rows x cols, ld) whose panels fit LDS (rows * PLD <= 16384):
// each panel is factored in registers (reg_panel), its reflectors and T
// written back (T of every panel to Tg, for q_form_tall), and the trailing
// columns updated by tall_wy with the reflectors read from W.  L: the tall
// layout (TALL_LDS doubles); the panel is staged through its first part,
// which tall_wy reuses once the panel is written back.
__device__ __noinline__ void qr_panel(double* W, int rows_in, int cols_in, int ld_in, double* tau, double* Tg,
                                      lds_f64* L, CoopJob* job, int nwg_in, int* iflag) {
  const int nwg = uni(nwg_in);
  const int rows = uni(rows_in), cols = uni(cols_in), ld = uni(ld_in);
  const int k = rows < cols ? rows : cols;
  lds_f64* red = L + 8 * TALL_REGION;          // 2 x 8 x PNB sums, 2 x PNB rows
  lds_f64* Tl = L + 8 * TALL_REGION + 16 * 128;
  // on this workgroup alone, panels go in pairs: the second panel's columns
  // are updated by the first, and the rest by both in one pass (tall_wy2)
  const bool pairs = !(job != nullptr && nwg > 1);
  lds_f64* T1 = red + 512;   // the first panel's T while the second is factored
  lds_f64* Gp = red + 768;   // V2^T V1
  QTT_PT_INIT;
  auto panel = [&](int p0, int nb) {
    if (rows - p0 <= 512)
      reg_panel<1>(W, rows, ld, p0, nb, tau + p0, Tg + (p0 / PNB) * PNB * PNB, Tl, L, red);
    else
      reg_panel<2>(W, rows, ld, p0, nb, tau + p0, Tg + (p0 / PNB) * PNB * PNB, Tl, L, red);
  };
  for (int p0 = 0; p0 < k;) {
    const int nb = k - p0 < PNB ? k - p0 : PNB, mr = rows - p0;
    panel(p0, nb);
    QTT_PT(1);
    const int p1 = p0 + nb;
    double* M = W + (int64_t)p0 * ld;
    const double* V1 = W + (int64_t)p0 * ld + p0;
    if (pairs && p1 < k && p1 + PNB < cols) {
      const int nb2 = k - p1 < PNB ? k - p1 : PNB;
      tall_wy(M, ld, mr, p1, p1 + nb2, V1, ld, nb, Tl, true, L, nullptr, 1, iflag);
      __syncthreads();
      for (int e = threadIdx.x; e < PNB * PNB; e += blockDim.x) T1[e] = Tl[e];
      __syncthreads();
      panel(p1, nb2);
      tall_wy2(M, ld, mr, p1 + nb2, cols, V1, W + (int64_t)p1 * ld + p1, ld, nb, nb2, T1, Tl, Gp, L, true, false);
      p0 = p1 + nb2;
    } else {
      if (p1 < cols) tall_wy(M, ld, mr, p1, cols, V1, ld, nb, Tl, true, L, job, nwg, iflag);
      p0 = p1;
    }
    __syncthreads();
    QTT_PT(4);
    if (threadIdx.x == 0) QTT_PANEL_COUNT;
  }
}

template <int G>
__device__ inline void row_norms(const double* X, int n, int b, double* sig) {
  const int gl = threadIdx.x & (G - 1), gid = threadIdx.x / G, ng = blockDim.x / G;
  for (int i = gid; i < n; i += ng) {
    double s = 0.0;
    for (int c = gl; c < b; c += G) s += X[(int64_t)i * b + c] * X[(int64_t)i * b + c];
    s = gsum<G>(s);
    if (gl == 0) sig[i] = sqrt(s);
  }
}

// Mode product Y[(aa*m + i)*Bu + bb] = sum_c R[row_i][c] U[(aa*nu + c)*Bu + bb]
// (row_i = perm[i] or i).  One task = one (aa, bb) column and 8 output rows:
// u is read once per 8 rows, eight of its entries in flight per wait.
// MODE: QTT_F64, QTT_F32_SPLIT or QTT_F32 (see ff_madd); output columns
// j = aa * Bu + bb in [j0, j1).
template <int MODE, typename RP, typename IP>
__device__ __forceinline__ void mode_product_body(RP Rrows, IP perm, bool use_perm, int m, int b,
                                                  const double* __restrict__ Uc, double* __restrict__ gY, int Au,
                                                  int nu, int Bu, int j0, int j1) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int ncol = j1 - j0;
  const int nib = (m + 7) >> 3;
  (void)Au;
  for (int task = tid; task < ncol * nib; task += bs) {
    const int j = j0 + task % ncol, ib = task / ncol;
    const int aa = j / Bu, bb = j - aa * Bu;
    RP rr[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int i = min(ib * 8 + q, m - 1);
      rr[q] = Rrows + (use_perm ? perm[i] : i) * b;
    }
    const double* up = Uc + (int64_t)aa * nu * Bu + bb;
    double res[8];
    if constexpr (MODE == QTT_F64) {
      double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int c = 0;
      for (; c + 8 <= nu; c += 8) {
        double uv[8];
#pragma unroll
        for (int q2 = 0; q2 < 8; ++q2) uv[q2] = up[(int64_t)(c + q2) * Bu];
#pragma unroll
        for (int q2 = 0; q2 < 8; ++q2)
#pragma unroll
          for (int q = 0; q < 8; ++q) acc[q] = __builtin_fma(rr[q][c + q2], uv[q2], acc[q]);
      }
      for (; c < nu; ++c) {
        const double uv = up[(int64_t)c * Bu];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = __builtin_fma(rr[q][c], uv, acc[q]);
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) res[q] = acc[q];
    } else {
      // Float words cover a far narrower exponent range than doubles (a
      // network's norm can sit near 1e-33): the R rows and the u column are
      // scaled by powers of two (exact) to a maximum near 1 first, and the
      // result scaled back in float64.
      int eu = 0, er[8];
      {
        double mx = 0.0;
        for (int c = 0; c < nu; ++c) mx = fmax(mx, fabs(up[(int64_t)c * Bu]));
        eu = mx > 0.0 ? ilogb(mx) : 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          double mr = 0.0;
          for (int c = 0; c < nu; ++c) mr = fmax(mr, fabs((double)rr[q][c]));
          er[q] = mr > 0.0 ? ilogb(mr) : 0;
        }
      }
      if constexpr (MODE == QTT_F32_SPLIT) {
        ffv acc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = {0.0f, 0.0f};
        for (int c = 0; c < nu; ++c) {
          const ffv uv = ff_split(ldexp(up[(int64_t)c * Bu], -eu));
#pragma unroll
          for (int q = 0; q < 8; ++q) ff_madd(acc[q], ff_split(ldexp((double)rr[q][c], -er[q])), uv);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) res[q] = ldexp(ff_value(acc[q]), er[q] + eu);
      } else {
        float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (int c = 0; c < nu; ++c) {
          const float uv = (float)ldexp(up[(int64_t)c * Bu], -eu);
#pragma unroll
          for (int q = 0; q < 8; ++q) acc[q] = __builtin_fmaf((float)ldexp((double)rr[q][c], -er[q]), uv, acc[q]);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) res[q] = ldexp((double)acc[q], er[q] + eu);
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int i = ib * 8 + q;
      if (i < m) gY[((int64_t)aa * m + i) * Bu + bb] = res[q];
    }
  }
}

// LDS-resident R
template <int MODE>
__device__ __forceinline__ void mode_product_lds(const lds_f64* Rrows, const lds_i32* perm, bool use_perm, int m, int b,
                                              const double* Uc, double* gY, int Au, int nu, int Bu) {
  mode_product_body<MODE>(Rrows, perm, use_perm, m, b, Uc, gY, Au, nu, Bu, 0, Au * Bu);
}

template <int MODE>
__device__ __forceinline__ void mode_product(const double* Rrows, const int* perm, bool use_perm, int m, int b,
                                             const double* Uc, double* gY, int Au, int nu, int Bu) {
  mode_product_body<MODE>(Rrows, perm, use_perm, m, b, Uc, gY, Au, nu, Bu, 0, Au * Bu);
}

// The mode product in float32 on MFMA tiles, v_mfma_f32_16x16x4f32: the same
// units as mode_product_mfma below (32 rows by 16 columns per wavefront);
// operands rounded to float, float accumulation.  D element e of lane =
// row 4 lk + e, column li.
__device__ __noinline__ void mode_product_mfma32(const double* Rrows, const int* perm, bool use_perm, int m_in,
                                                 int b_in, const double* Uc, double* gY, int nu_in, int Bu_in,
                                                 int j0_in, int j1_in) {
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m = uni(m_in), b = uni(b_in), nu = uni(nu_in), Bu = uni(Bu_in), j0 = uni(j0_in), j1 = uni(j1_in);
  const int nit = (m + 31) >> 5, njt = (j1 - j0 + 15) >> 4;
  for (int u = wv; u < nit * njt; u += nwv) {
    const int it = u / njt, jt = u - it * njt;
    const int i0 = it * 32, jc = j0 + jt * 16 + li;
    const bool jon = jc < j1;
    const int aa = jon ? jc / Bu : 0, bb = jon ? jc - aa * Bu : 0;
    const double* up = Uc + (int64_t)aa * nu * Bu + bb;
    const int ra = i0 + li, rb = i0 + 16 + li;
    const bool aon = ra < m, bon = rb < m;
    const double* rpa = Rrows + (int64_t)(aon ? (use_perm ? perm[ra] : ra) : 0) * b;
    const double* rpb = Rrows + (int64_t)(bon ? (use_perm ? perm[rb] : rb) : 0) * b;
    // powers of two bringing each R row and u column to a maximum near 1 (the
    // float range is far narrower than the double one); the four lanes that
    // share a row / column (lk = 0..3) each scan a quarter of it
    double mu = 0.0, ma = 0.0, mb = 0.0;
    for (int c = lk; c < nu; c += 4) {
      if (jon) mu = fmax(mu, fabs(up[(int64_t)c * Bu]));
      if (aon) ma = fmax(ma, fabs(rpa[c]));
      if (bon) mb = fmax(mb, fabs(rpb[c]));
    }
    mu = fmax(mu, __shfl_xor(mu, 16));
    mu = fmax(mu, __shfl_xor(mu, 32));
    ma = fmax(ma, __shfl_xor(ma, 16));
    ma = fmax(ma, __shfl_xor(ma, 32));
    mb = fmax(mb, __shfl_xor(mb, 16));
    mb = fmax(mb, __shfl_xor(mb, 32));
    const int eu = mu > 0.0 ? ilogb(mu) : 0, ea = ma > 0.0 ? ilogb(ma) : 0, eb = mb > 0.0 ? ilogb(mb) : 0;
    f32x4 da = {0.0f, 0.0f, 0.0f, 0.0f}, db = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int c0 = 0; c0 < nu; c0 += 32) {
      float bv[8], av[8], aw[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int c = c0 + 4 * q + lk;
        const bool con = c < nu;
        bv[q] = (jon && con) ? (float)ldexp(up[(int64_t)c * Bu], -eu) : 0.0f;
        av[q] = (aon && con) ? (float)ldexp(rpa[c], -ea) : 0.0f;
        aw[q] = (bon && con) ? (float)ldexp(rpb[c], -eb) : 0.0f;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        da = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q], bv[q], da, 0, 0, 0);
        db = __builtin_amdgcn_mfma_f32_16x16x4f32(aw[q], bv[q], db, 0, 0, 0);
      }
    }
    // the row exponents of this lane's output rows 4 lk + e live in lanes 4 lk + e
    int era[4], erb[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      era[e] = __shfl(ea, 4 * lk + e);
      erb[e] = __shfl(eb, 4 * lk + e);
    }
    if (jon) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i0 + 4 * lk + e;
        if (i < m) gY[((int64_t)aa * m + i) * Bu + bb] = ldexp((double)da[e], era[e] + eu);
        if (i + 16 < m) gY[((int64_t)aa * m + i + 16) * Bu + bb] = ldexp((double)db[e], erb[e] + eu);
      }
    }
  }
}

// Mode product with R in global memory (the global-workspace path) as FP64
// MFMA tiles, v_mfma_f64_16x16x4f64: Y[(aa*m + i)*Bu + bb] = sum_c
// R[perm(i)][c] U[(aa*nu + c)*Bu + bb] over the output columns j = aa*Bu + bb
// in [j0, j1).  A unit is 32 output rows (two 16-row tiles sharing each B
// operand) by 16 columns; A = R rows (lane: row li, column lk of the k-step),
// B = U (lane: k-row lk, column li), D element e = row lk + 4e, column li.
__device__ __noinline__ void mode_product_mfma(const double* Rrows, const int* perm, bool use_perm, int m_in,
                                               int b_in, const double* Uc, double* gY, int Au_in, int nu_in,
                                               int Bu_in, int j0_in, int j1_in) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m = uni(m_in), b = uni(b_in), nu = uni(nu_in), Bu = uni(Bu_in), j0 = uni(j0_in), j1 = uni(j1_in);
  (void)Au_in;
  const int nit = (m + 31) >> 5, njt = (j1 - j0 + 15) >> 4;
  for (int u = wv; u < nit * njt; u += nwv) {
    const int it = u / njt, jt = u - it * njt;
    const int i0 = it * 32, jc = j0 + jt * 16 + li;
    const bool jon = jc < j1;
    const int aa = jon ? jc / Bu : 0, bb = jon ? jc - aa * Bu : 0;
    const double* up = Uc + (int64_t)aa * nu * Bu + bb;
    const int ra = i0 + li, rb = i0 + 16 + li;
    const bool aon = ra < m, bon = rb < m;
    const double* rpa = Rrows + (int64_t)(aon ? (use_perm ? perm[ra] : ra) : 0) * b;
    const double* rpb = Rrows + (int64_t)(bon ? (use_perm ? perm[rb] : rb) : 0) * b;
    f64x4 da = {0.0, 0.0, 0.0, 0.0}, db = {0.0, 0.0, 0.0, 0.0};
    // without a permutation the rows are a QR's R (upper triangular, its zeros
    // stored): the k-steps left of the tile's first row add nothing
    for (int c0 = use_perm ? 0 : i0; c0 < nu; c0 += 32) {  // 8 k-steps: all 24 loads in flight before the MFMAs
      double bv[8], av[8], aw[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int c = c0 + 4 * q + lk;
        const bool con = c < nu;
        bv[q] = (jon && con) ? up[(int64_t)c * Bu] : 0.0;
        av[q] = (aon && con) ? rpa[c] : 0.0;
        aw[q] = (bon && con) ? rpb[c] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        da = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], da, 0, 0, 0);
        db = __builtin_amdgcn_mfma_f64_16x16x4f64(aw[q], bv[q], db, 0, 0, 0);
      }
    }
    if (jon) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i0 + lk + 4 * e;
        if (i < m) gY[((int64_t)aa * m + i) * Bu + bb] = da[e];
        if (i + 16 < m) gY[((int64_t)aa * m + i + 16) * Bu + bb] = db[e];
      }
    }
  }
}

// One round of a job on this workgroup: pair pi = wg + nwg * slot goes to
// wavefront slot (64 lanes per pair, E entries per lane of a row of X and of
// J), as in jacobi_rows_reg.  Returns whether any of its pairs rotated by
// more than the stopping rule's angle (jacobi_stop2).
template <int E>
__device__ __noinline__ bool coop_pairs(double* X, double* J, int n_in, int b_in, int ldx_in, int ldj_in, int t_in,
                                        int nn_in, double tol, double small, double noise, int wg_in, int nwg_in) {
  const int n = uni(n_in), b = uni(b_in), ldx = uni(ldx_in), ldj = uni(ldj_in), t = uni(t_in), nn = uni(nn_in),
            wg = uni(wg_in), nwg = uni(nwg_in);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  bool any = false;
  for (int pi = wg + nwg * wv; pi < nn / 2; pi += nwg * nwv) {
    int p = rr_player(pi, t, nn), q = rr_player(nn - 1 - pi, t, nn);
    if (p > q) {
      const int tmp = p;
      p = q;
      q = tmp;
    }
    if (q >= n) continue;
    double* xp = X + (int64_t)p * ldx;
    double* xq = X + (int64_t)q * ldx;
    double* jp = J + (int64_t)p * ldj;
    double* jq = J + (int64_t)q * ldj;
    double u[E], w[E], ju[E], jw[E];
    const int gl = opaque(lane);
    load_row<64, E>(xp, b, gl, u);
    load_row<64, E>(xq, b, gl, w);
    load_row<64, E>(jp, n, gl, ju);
    load_row<64, E>(jq, n, gl, jw);
    double a0 = 0.0, c0 = 0.0, g0 = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      a0 = __builtin_fma(u[e], u[e], a0);
      c0 = __builtin_fma(w[e], w[e], c0);
      g0 = __builtin_fma(u[e], w[e], g0);
    }
    const double a = jsum<64>(a0), c = jsum<64>(c0), g = jsum<64>(g0);
    if (g != 0.0 && a > small && c > small && (a > noise || c > noise) && g * g > tol * tol * (a * c)) {
      double cs, sn;
      jacobi_rotation(a, c, g, cs, sn);
      rotate4<E>(u, w, cs, sn);
      rotate4<E>(ju, jw, cs, sn);
      store_row<64, E>(xp, b, gl, u);
      store_row<64, E>(xq, b, gl, w);
      store_row<64, E>(jp, n, gl, ju);
      store_row<64, E>(jq, n, gl, jw);
      any |= g * g > jacobi_stop2(tol) * (a * c);
    }
  }
  return any;
}

// Type 8: this workgroup's block pair of block round t: the 2 bsz rows of X
// and of J are brought into LDS, their pairs rotated (one wavefront per pair,
// both rows of X and J in registers), and written back.  Returns whether any
// pair rotated by more than the stopping rule's angle (jacobi_stop2).
template <int E>
__device__ __noinline__ bool coop_block(double* X, double* J, int n_in, int b_in, int bsz_in, int nblk_in, int t_in,
                                        double tol, double small, double noise, int wg_in, lds_f64* lw) {
  const int n = uni(n_in), b = uni(b_in), bsz = uni(bsz_in), nblk = uni(nblk_in), t = uni(t_in), wg = uni(wg_in);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, bs = blockDim.x;
  if (wg >= nblk / 2) return false;
  const int P = rr_player(wg, t, nblk), Q = rr_player(nblk - 1 - wg, t, nblk);
  const int nl = 2 * bsz;  // local rows: P's then Q's
  auto grow = [&](int i) { return i < bsz ? P * bsz + i : Q * bsz + (i - bsz); };
  const int lx = (b + 1) & ~1, lj = (n + 1) & ~1;  // LDS row strides
  lds_f64* Xl = lw;
  lds_f64* Jl = lw + nl * lx;
  lds_i32* fl = (lds_i32*)(Jl + nl * lj);
  // the two blocks of X and of J into LDS, 8 loads per thread in flight per wait
  auto stage_in = [&](lds_f64* dst, const double* src, int ld, int len, int cols) {
    for (int e0 = tid; e0 < nl * ld; e0 += 8 * bs) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * bs;
        const int i = e / ld, c = e - i * ld, r = e < nl * ld ? grow(i) : n;
        v[u] = (r < n && c < cols) ? src[(int64_t)r * len + c] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (e0 + u * bs < nl * ld) dst[e0 + u * bs] = v[u];
    }
  };
  stage_in(Xl, X, lx, b, b);
  stage_in(Jl, J, lj, n, n);
  if (tid == 0) fl[0] = 0;
  __syncthreads();
  // round 0 sweeps all local pairs (so every pair inside a block is met once
  // per sweep); the later rounds take only the pairs across the two blocks,
  // bsz steps of bsz disjoint pairs (P row w with Q row (w + step) mod bsz)
  const int nsteps = t == 0 ? nl - 1 : bsz;
  const int nwv = bs >> 6;
  for (int lt = 0; lt < nsteps; ++lt) {
    for (int pw = wv; pw < bsz; pw += nwv) {  // bsz pairs per step, one wavefront each
      int i = t == 0 ? rr_player(pw, lt, nl) : pw;
      int k = t == 0 ? rr_player(nl - 1 - pw, lt, nl) : bsz + (pw + lt) % bsz;
      if (grow(i) > grow(k)) {  // the lower global row first, as in the other Jacobi routines
        const int tmp = i;
        i = k;
        k = tmp;
      }
      if (grow(k) < n) {
        const int gl = opaque(lane);
        double u[E], w[E], ju[E], jw[E];
        load_row<64, E>(Xl + i * lx, b, gl, u);
        load_row<64, E>(Xl + k * lx, b, gl, w);
        load_row<64, E>(Jl + i * lj, n, gl, ju);
        load_row<64, E>(Jl + k * lj, n, gl, jw);
        double a0 = 0.0, c0 = 0.0, g0 = 0.0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          a0 = __builtin_fma(u[e], u[e], a0);
          c0 = __builtin_fma(w[e], w[e], c0);
          g0 = __builtin_fma(u[e], w[e], g0);
        }
        const double a = jsum<64>(a0), c = jsum<64>(c0), g = jsum<64>(g0);
        if (g != 0.0 && a > small && c > small && (a > noise || c > noise) && g * g > tol * tol * (a * c)) {
          double cs, sn;
          jacobi_rotation(a, c, g, cs, sn);
          rotate4<E>(u, w, cs, sn);
          rotate4<E>(ju, jw, cs, sn);
          store_row<64, E>(Xl + i * lx, b, gl, u);
          store_row<64, E>(Xl + k * lx, b, gl, w);
          store_row<64, E>(Jl + i * lj, n, gl, ju);
          store_row<64, E>(Jl + k * lj, n, gl, jw);
          if (lane == 0 && g * g > jacobi_stop2(tol) * (a * c)) fl[0] = 1;
        }
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < nl * lx; e += bs) {
    const int i = e / lx, c = e - i * lx, r = grow(i);
    if (r < n && c < b) X[(int64_t)r * b + c] = Xl[e];
  }
  for (int e = tid; e < nl * lj; e += bs) {
    const int i = e / lj, c = e - i * lj, r = grow(i);
    if (r < n && c < n) J[(int64_t)r * n + c] = Jl[e];
  }
  const bool any = fl[0] != 0;
  __syncthreads();
  return any;
}

// W[(aa bx + bb) b + i] = X[(aa b + i) bx + bb] for aa < na: the move's
// matrix gathered from its core, for each aa a transpose of the b x bx block
// X[aa] into W[aa].  64 x 64 tiles through LDS (lw: 64 x 65 doubles), so the
// reads (along bb) and the writes (along i) are both coalesced; this
// workgroup takes tiles wg, wg + nwg, ...
__device__ void gather_tiles(const double* X, double* W, int na, int b, int bx, int wg, int nwg, lds_f64* lw) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int ti = (b + 63) >> 6, tb = (bx + 63) >> 6;
  const int64_t per = (int64_t)ti * tb, ntile = (int64_t)na * per;
  for (int64_t t = wg; t < ntile; t += nwg) {
    const int aa = (int)(t / per), rem = (int)(t - (int64_t)aa * per);
    const int i0 = (rem / tb) * 64, b0 = (rem - (rem / tb) * tb) * 64;
    const double* S = X + (int64_t)aa * b * bx;
    double* D = W + (int64_t)aa * bx * b;
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int i = i0 + wv + 8 * q, c = b0 + lane;
      v[q] = (i < b && c < bx) ? S[(int64_t)i * bx + c] : 0.0;
    }
    __syncthreads();  // the previous tile's reads of lw are done
#pragma unroll
    for (int q = 0; q < 8; ++q) lw[(wv + 8 * q) * 65 + lane] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int c = b0 + wv + 8 * q, i = i0 + lane;
      if (c < bx && i < b) D[(int64_t)c * b + i] = lw[lane * 65 + wv + 8 * q];
    }
  }
  __syncthreads();
}

// M[0:mr, c0:c1] <- (I - V op(T) V^T) M[0:mr, c0:c1] for mr <= 1024, whole
// 16-column tiles per workgroup (tiles wg, wg + nwg, ...): W = V^T M_tile
// with the rows split over the wavefronts (16-row blocks w, w + 8, ...),
// the 8 partial tiles summed through LDS, U = op(T) W, then M_tile -= V U
// (the tile read again, from L2).  One job per update, no partial tiles
// through global memory.  Few registers (two blocks in flight): the sweep
// kernel keeps values across calls in the registers its callees leave
// alone.  T: 16 x 16 row-major; L: 2560 doubles of LDS.
__device__ __noinline__ void wy_cols(double* M, int ldm, int mr, int c0, int c1, const double* V, int ldv, int nb, const double* T,
                        bool trans, int wg, int nwg, lds_f64* L) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, li = lane & 15, lk = lane >> 4;
  const int ntile = (c1 - c0 + 15) >> 4;
  lds_f64* Pw = L;              // [wavefront][e * 64 + lane]
  lds_f64* Wt = L + 8 * 256;    // W [j][c]
  lds_f64* Ut = Wt + 256;       // -U [j][c]
  for (int t = wg; t < ntile; t += nwg) {
    const int col0 = c0 + 16 * t;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
    for (int s = 0; s < 8; ++s) {
      const int rb = 16 * (wv + 8 * s);
      if (rb < mr) acc += wy_p1_unit(M, ldm, mr, col0, c1, V, ldv, nb, rb, min(mr, rb + 16));
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) Pw[wv * 256 + e * 64 + lane] = acc[e];
    __syncthreads();
    if (tid < 256) {  // element e of lane ln is W[ln / 16 + 4 e][ln % 16]
      const int e = tid >> 6, ln = tid & 63;
      double w = 0.0;
#pragma unroll
      for (int h = 0; h < 8; ++h) w += Pw[h * 256 + tid];
      Wt[((ln >> 4) + 4 * e) * 16 + (ln & 15)] = w;
    }
    __syncthreads();
    if (tid < 256) {
      const int i = tid >> 4, c = tid & 15;
      double u = 0.0;
#pragma unroll
      for (int j = 0; j < PNB; ++j) u = __builtin_fma(trans ? T[j * PNB + i] : T[i * PNB + j], Wt[j * 16 + c], u);
      Ut[i * 16 + c] = -u;
    }
    __syncthreads();
#pragma unroll 2
    for (int s = 0; s < 8; ++s) {
      const int rb = 16 * (wv + 8 * s);
      if (rb < mr) wy_p3_unit(M, ldm, mr, col0, c1, V, ldv, nb, rb, (const lds_f64*)Ut);
    }
    __syncthreads();  // Pw, Wt and Ut are reused by the next tile
  }
}

// This workgroup's share of the posted job, then its completion count.  lw:
// LDS the job may use (the row blocks of a block Jacobi round).
__device__ void coop_work(CoopJob* job, int wg, int nwg, lds_f64* lw) {
  const int tid = threadIdx.x;
  const int type = job->type;
  if (type == 1) {
    const int e = job->e;
    bool r;
    if (e <= 2) r = coop_pairs<2>(job->X, job->J, job->n, job->b, job->ldx, job->ldj, job->t, job->nn, job->tol, job->small, job->noise, wg, nwg);
    else if (e <= 4) r = coop_pairs<4>(job->X, job->J, job->n, job->b, job->ldx, job->ldj, job->t, job->nn, job->tol, job->small, job->noise, wg, nwg);
    else if (e <= 6) r = coop_pairs<6>(job->X, job->J, job->n, job->b, job->ldx, job->ldj, job->t, job->nn, job->tol, job->small, job->noise, wg, nwg);
    else r = coop_pairs<8>(job->X, job->J, job->n, job->b, job->ldx, job->ldj, job->t, job->nn, job->tol, job->small, job->noise, wg, nwg);
    if (r && (tid & 63) == 0) __hip_atomic_store(&job->rot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (type == 2) {
    const int ncol = job->Au * job->Bu;
    const int chunk = ((ncol + nwg - 1) / nwg + 63) & ~63;
    const int j0 = min(ncol, wg * chunk), j1 = min(ncol, j0 + chunk);
    if (j0 < j1) {
      if (job->prec == QTT_F32)
        mode_product_mfma32(job->R, job->perm, job->use_perm != 0, job->m, job->rb, job->U, job->Y, job->nu,
                            job->Bu, j0, j1);
      else if (job->prec == QTT_F32_SPLIT)
        mode_product_body<QTT_F32_SPLIT>(job->R, job->perm, job->use_perm != 0, job->m, job->rb, job->U, job->Y,
                                         job->Au, job->nu, job->Bu, j0, j1);
      else
        mode_product_mfma(job->R, job->perm, job->use_perm != 0, job->m, job->rb, job->U, job->Y, job->Au, job->nu,
                          job->Bu, j0, j1);
    }
  } else if (type == 3) {
    const long long tot = job->ytot;
    const long long chunk = ((tot + nwg - 1) / nwg + 63) & ~63LL;
    const long long e0 = min(tot, (long long)wg * chunk), e1 = min(tot, e0 + chunk);
    for (long long x = e0 + tid; x < e1; x += blockDim.x) job->D[x] = job->Y[x];
  } else if (type == 10) {
    gather_tiles(job->U, job->D, job->Au, job->rb, job->Bu, wg, nwg, lw);
  } else if (type == 8) {  // a block round of a block Jacobi sweep
    const int e = job->e;
    bool r;
    if (e <= 2) r = coop_block<2>(job->X, job->J, job->n, job->b, job->bsz, job->nblk, job->t, job->tol, job->small, job->noise, wg, lw);
    else if (e <= 4) r = coop_block<4>(job->X, job->J, job->n, job->b, job->bsz, job->nblk, job->t, job->tol, job->small, job->noise, wg, lw);
    else if (e <= 6) r = coop_block<6>(job->X, job->J, job->n, job->b, job->bsz, job->nblk, job->t, job->tol, job->small, job->noise, wg, lw);
    else r = coop_block<8>(job->X, job->J, job->n, job->b, job->bsz, job->nblk, job->t, job->tol, job->small, job->noise, wg, lw);
    if (r && tid == 0) __hip_atomic_store(&job->rot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (type == 9) {  // a whole block Jacobi sweep, rounds ordered block by block
    const int nblk = job->nblk, rbase = job->rbase, e = job->e;
    bool r = false;
    for (int t = 0; wg < nblk / 2 && t < nblk - 1; ++t) {
      const int P = rr_player(wg, t, nblk), Q = rr_player(nblk - 1 - wg, t, nblk);
      if (tid == 0) {  // both blocks written back by round t - 1 (bounded wait)
        int spins = 0;
        while (__hip_atomic_load(&job->ver[P], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < rbase + t ||
               __hip_atomic_load(&job->ver[Q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < rbase + t) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins > COOP_SPIN_MASTER) {
            __hip_atomic_store(&job->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      bool rt;
      if (e <= 2) rt = coop_block<2>(job->X, job->J, job->n, job->b, job->bsz, nblk, t, job->tol, job->small, job->noise, wg, lw);
      else if (e <= 4) rt = coop_block<4>(job->X, job->J, job->n, job->b, job->bsz, nblk, t, job->tol, job->small, job->noise, wg, lw);
      else if (e <= 6) rt = coop_block<6>(job->X, job->J, job->n, job->b, job->bsz, nblk, t, job->tol, job->small, job->noise, wg, lw);
      else rt = coop_block<8>(job->X, job->J, job->n, job->b, job->bsz, nblk, t, job->tol, job->small, job->noise, wg, lw);
      r = r || rt;
      // (coop_block ends with a barrier after its stores: every wavefront's
      // writes are done before thread 0 releases the two blocks)
      if (tid == 0) {
        __hip_atomic_store(&job->ver[P], rbase + t + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&job->ver[Q], rbase + t + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (r && tid == 0) __hip_atomic_store(&job->rot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (type == 4 || type == 5) {  // a tall-panel pass over this workgroup's rows
    const int mr = job->n, nb = job->nb, j = job->t;
    const int start = type == 4 ? 1 : j + 1;
    const int rc = (mr - start + nwg - 1) / nwg;
    const int r0 = min(mr, start + wg * rc), r1 = min(mr, r0 + rc);
    double acc = 0.0;
    if (r0 < r1) {
      if (type == 4) {
        acc = panel_pre_rows(job->X, job->ldx, nb, r0, r1);
      } else {
        const int c = tid & 15;
        acc = panel_pass_rows(job->X, job->ldx, nb, j, job->tol, job->dcol[c], job->small, r0, r1);
      }
    }
    tall_reduce16(acc, lw, lw + 128);
    if (tid < PNB) job->Gp[wg * PNB + tid] = lw[128 + tid];
  }
  __syncthreads();  // every wavefront's writes are done before thread 0 releases them
  if (tid == 0) __hip_atomic_fetch_add(&job->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup 0: post the job whose fields this workgroup has written, take its own
// share, and wait until every workgroup has done its share.  Returns false if
// the wait exceeded its limit.
// nact: the workgroups taking part (0: all).
__device__ bool coop_post_run(CoopJob* job, int nwg, int* iflag, lds_f64* lw, int nact) {
  const int s = coop_post(job, iflag, nact > 0 ? min(nact, nwg) : nwg);
  coop_work(job, 0, nwg, lw);
  return coop_wait(job, s, nwg, iflag);
}

// Post the job whose fields this workgroup has written, for workgroups
// 0 .. nact-1; returns its sequence number.
__device__ int coop_post(CoopJob* job, int* iflag, int nact) {
  const int tid = threadIdx.x;
  lds_i32* sflag = (lds_i32*)iflag;
  __syncthreads();  // job fields written by any thread of this workgroup come before the release
  if (tid == 0) {
    const int s = job->seq + 1;
    job->seq = s;
    job->dtarget += (unsigned int)nact;
    __hip_atomic_store(&job->post, ((unsigned long long)s << 32) | (unsigned int)nact, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
    sflag[1] = s;
  }
  __syncthreads();
  return sflag[1];
}

// Wait until every workgroup taking part has done its share of the posted
// job (workgroup 0's included); false if the wait exceeded its limit.
__device__ bool coop_wait(CoopJob* job, int s, int nwg, int* iflag) {
  const int tid = threadIdx.x;
  lds_i32* sflag = (lds_i32*)iflag;
  (void)s;
  (void)nwg;
  if (tid == 0) {
    const unsigned int target = job->dtarget;
    int spins = 0, bad = 0;
    while (__hip_atomic_load(&job->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > COOP_SPIN_MASTER) {
        __hip_atomic_store(&job->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bad = 1;
        break;
      }
    }
    sflag[3] = bad;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  const int bad = sflag[3];
  __syncthreads();
  return !bad;
}

// This workgroup's share of a type-11 job (kept out of coop_work, which the
// sweep calls in many places, so that its registers stay few), then its
// completion count.
__device__ void coop_wy_cols(CoopJob* job, int wg, int nwg, lds_f64* lw) {
  wy_cols(job->wy_M, job->wy_ldm, job->wy_mr, job->wy_cg, job->wy_c1, job->wy_V, job->wy_ldv, job->nb, job->Ug,
          job->wy_trans != 0, wg, nwg, lw);
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(&job->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// W partial of one (tile, row chunk) unit as wy_p1_unit, 32 rows (eight
// steps) in flight per wait: the row-chunk jobs stream tall matrices from
// memory, one unit per wavefront.
__device__ __forceinline__ f64x4 wy_p1_unit32(const double* M, int ldm, int mr, int col0, int c1, const double* V,
                                              int ldv, int nb, int rs, int re) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int col = col0 + li;
  const bool con = col < c1;
  f64x4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  int r = rs;
  for (; r + 32 <= re; r += 32) {
    double a[8], b[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int rr = r + 4 * q + lk;
      a[q] = vref(V, ldv, rr, li, mr, nb);
      b[q] = con ? M[(int64_t)rr * ldm + col] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; q += 2) {
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b[q], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q + 1], b[q + 1], acc1, 0, 0, 0);
    }
  }
  return acc0 + acc1 + wy_p1_unit(M, ldm, mr, col0, c1, V, ldv, nb, r, re);
}

// This workgroup's share of a row-chunk WY job (type 6: the W partials of its
// wavefronts' units; type 7: M -= V U on its (tile, 16-row block) units, two
// per wait), then its completion count.  Kept out of coop_work, which the
// sweep calls in many places, so that its registers stay few.
__device__ __noinline__ void coop_wy_rows(CoopJob* job, int wg, int nwg) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwv = blockDim.x >> 6;
  const int mr = job->wy_mr;
  if (job->type == 6) {
    const int nch = job->wy_nch, chunk = job->wy_chunk;
    const int nunit = job->wy_nct * nch;
    for (int u = wg * nwv + wv; u < nunit; u += nwg * nwv) {
      const int t = u / nch, h = u - t * nch;
      const int rs = h * chunk, re = min(mr, rs + chunk);
      const f64x4 acc = wy_p1_unit32(job->wy_M, job->wy_ldm, mr, job->wy_cg + t * 16, job->wy_c1, job->wy_V,
                                     job->wy_ldv, job->nb, rs, re);
#pragma unroll
      for (int e = 0; e < 4; ++e) job->Pg[(int64_t)u * 256 + e * 64 + lane] = acc[e];
    }
  } else {
    const int nrb = (mr + 15) >> 4, nct = job->wy_nct, nu = nct * nrb, stride = nwg * nwv;
    for (int x = wg * nwv + wv; x < nu; x += 2 * stride) {
      const int t = x / nrb, rb = (x - t * nrb) * 16, x2 = x + stride;
      if (x2 < nu) {
        const int t2 = x2 / nrb, rb2 = (x2 - t2 * nrb) * 16;
        wy_p3_unit2(job->wy_M, job->wy_ldm, mr, job->wy_cg + t * 16, rb, (const double*)(job->Ug + t * 256),
                    job->wy_cg + t2 * 16, rb2, (const double*)(job->Ug + t2 * 256), job->wy_c1, job->wy_V,
                    job->wy_ldv, job->nb);
      } else {
        wy_p3_unit(job->wy_M, job->wy_ldm, mr, job->wy_cg + t * 16, job->wy_c1, job->wy_V, job->wy_ldv, job->nb,
                   rb, (const double*)(job->Ug + t * 256));
      }
    }
  }
  __syncthreads();
  if (tid == 0) __hip_atomic_fetch_add(&job->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Helper workgroups: take every posted job they take part in until
// workgroup 0 quits.  A job of workgroups 0 .. nact-1 is not followed by the
// next before they have all counted it, so a helper that takes part always
// sees its job; one that does not may see a later job and skips to it.
__device__ void coop_helper(CoopJob* job, int wg, int nwg, lds_i32* sflag, lds_f64* lw) {
  const int tid = threadIdx.x;
  for (int expected = 1;;) {
    if (tid == 0) {
      int spins = 0, st = 0, seen = 0, part = 0;
      for (;;) {
        if (__hip_atomic_load(&job->quit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { st = 1; break; }
        const unsigned long long w = __hip_atomic_load(&job->post, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        seen = (int)(w >> 32);
        if (seen >= expected) {
          part = wg < (int)(unsigned int)w;
          break;
        }
        if (spins < 64) __builtin_amdgcn_s_sleep(2);
        else __builtin_amdgcn_s_sleep(64);
        if (++spins > COOP_SPIN_HELPER) {  // give up; a later job would then time out in workgroup 0
          st = 1;
          break;
        }
      }
      sflag[0] = st;
      sflag[4] = seen;
      sflag[5] = part;
    }
    __syncthreads();
    const int st = sflag[0], seen = sflag[4], part = sflag[5];
    __syncthreads();
    if (st) return;
    expected = seen + 1;
    if (!part) continue;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (job->type == 11) coop_wy_cols(job, wg, nwg, lw);
    else if (job->type == 6 || job->type == 7) coop_wy_rows(job, wg, nwg);
    else coop_work(job, wg, nwg, lw);
  }
}

// One-sided Jacobi on rows in global memory with the helper workgroups (run
// by workgroup 0; same rules as jacobi_rows_reg, round-robin rounds, each
// round one job).
__device__ bool jacobi_coop(double* X, int n, int b, double* J, CoopJob* job, int nwg, double* red, int* iflag,
                            lds_f64* lw, int lw_cap, double cut, bool init_j, int max_sweeps) {
  const int tid = threadIdx.x, bs = blockDim.x;
  if (init_j)
    for (int e2 = tid; e2 < n * n; e2 += bs) J[e2] = (e2 / n == e2 % n) ? 1.0 : 0.0;
  __syncthreads();
  if (n < 2) return true;
  const int nn = n + (n & 1);
  const double tol = DBL_EPSILON * (double)(b > n ? b : n);
  double fro = 0.0;
  for (int e2 = tid; e2 < n * b; e2 += bs) fro += X[e2] * X[e2];
  const double fsum = block_sum(fro, red);
  const double small = fsum * DBL_EPSILON * DBL_EPSILON;
  const double noise = noise_level(fsum, cut, n);
  const int len = b > n ? b : n;
  const int e = len <= 128 ? 2 : len <= 256 ? 4 : len <= 384 ? 6 : 8;
  lds_i32* sflag = (lds_i32*)iflag;
  // block rounds: pairs of blocks of bsz rows, one pair per workgroup, their
  // rows in LDS (fewer rounds, each with more work); else row-pair rounds.
  // A round costs a handoff and the blocks' trip through LDS, a local step
  // costs one row pair per wavefront: bsz is 8 when the block pair fits LDS
  // and leaves at least 4 blocks (one pair per wavefront per step), else 4, or
  // more when the workgroups could not take all blocks
  const int ldsum = ((b + 1) & ~1) + ((n + 1) & ~1);
  const int cap = min(lw_cap, QTT_LDS_DOUBLES - 64);
  int bsz = max(4, (n + 2 * nwg - 1) / (2 * nwg));
  for (int c = 8; c > bsz; c >>= 1)
    if (2 * c * ldsum + 8 <= cap && (n + c - 1) / c >= 4) {
      bsz = c;
      break;
    }
  int nblk = (n + bsz - 1) / bsz;
  nblk += nblk & 1;
  const int need = 2 * bsz * ldsum + 8;
  const bool blocked = bsz <= 16 && nblk >= 4 && nblk <= 256 && need <= cap;  // (ver holds 256 blocks)
  const int rounds = blocked ? nblk - 1 : nn - 1;
  // block sweeps are one job each (type 9), the rounds ordered by per-block
  // counters instead of a hand-off through workgroup 0 per round
  int rbase = 0;
  if (blocked)
    for (int b2 = tid; b2 < nblk; b2 += bs) job->ver[b2] = 0;  // published by the first post
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) __hip_atomic_store(&job->rot, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int t = 0; t < (blocked ? 1 : rounds); ++t) {
      if (tid == 0) {
        job->type = blocked ? 9 : 1;
        job->rbase = rbase;
        job->X = X;
        job->J = J;
        job->n = n;
        job->b = b;
        job->ldx = b;
        job->ldj = n;
        job->t = t;
        job->nn = nn;
        job->e = e;
        job->tol = tol;
        job->small = small;
        job->noise = noise;
        job->bsz = bsz;
        job->nblk = nblk;
      }
      if (!coop_post_run(job, nwg, iflag, lw, blocked ? nblk / 2 : 0)) break;
    }
    rbase += rounds;
    if (tid == 0) sflag[2] = __hip_atomic_load(&job->rot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) |
                             __hip_atomic_load(&job->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) * 2;
    __syncthreads();
    const int st = sflag[2];
    __syncthreads();
    if (st & 2) return false;  // a wait failed: report non-convergence
    if (!st) {
      QTT_SWEEPS(sweep + 1);
      return true;
    }
  }
  QTT_SWEEPS(max_sweeps);
  return false;
}

// One-sided Jacobi of a wide matrix X (n x b, n <= b, ld b, global memory)
// preconditioned by two QRs (Drmac and Veselic; here without pivoting):
// X^T = Q1 R1 and R1^T = Q2 R2, so X = Q2 R2 Q1^T.  The rounds run on the rows
// of R2 (n x n) with J starting from Q2^T; X's rotated rows are then
// (J R2) Q1^T, formed as (Q1 (J R2)^T)^T.  A truncation sweep's wide matrices
// are numerically rank deficient (their smallest singular values are round-off
// of the bigger ones) and their rows keep rotating for many sweeps unless
// preconditioned: on a 256 x 384 matrix of the 3D comb operate a CPU replay of
// the block rounds takes 35 sweeps on X, 8 on R2.  A (b x n) receives X^T and
// keeps Q1's reflectors; Tg1, Tg2 the panels' T of the two QRs; Mq (b x n) is
// scratch.  Same result as the Jacobi on X up to round-off (the rotations act
// on the same row space).
__device__ __noinline__ bool jacobi_precond(double* X, int n_in, int b_in, double* A, double* J, double* tau,
                                            double* Tg1, double* Tg2, double* Mq, lds_f64* pnl, int pnl_cap,
                                            CoopJob* job, int nwg_in, double* red, int* iflag, double cut,
                                            int max_sweeps) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int n = uni(n_in), b = uni(b_in), nwg = uni(nwg_in);
  // A = X^T: tiled transposes shared with the helper workgroups
  if (tid == 0) {
    job->type = 10;
    job->U = X;
    job->D = A;
    job->Au = 1;
    job->rb = n;
    job->Bu = b;
  }
  coop_post_run(job, nwg, iflag, pnl, (int)min<int64_t>((int64_t)((n + 63) >> 6) * ((b + 63) >> 6), nwg));
  __syncthreads();
  if (b * PLD <= 8 * TALL_REGION) qr_panel(A, b, n, n, tau, Tg1, pnl, job, nwg, iflag);
  else qr_tall(A, b, n, n, tau, Tg1, pnl, job, nwg, iflag);
  __syncthreads();
  for (int e = tid; e < n * n; e += bs) {  // X <- R1^T (n x n, ld n)
    const int i = e / n, j = e - i * n;
    X[e] = j <= i ? A[(int64_t)j * n + i] : 0.0;
  }
  __syncthreads();
  if (n * PLD <= 8 * TALL_REGION) qr_panel(X, n, n, n, tau, Tg2, pnl, job, nwg, iflag);
  else qr_tall(X, n, n, n, tau, Tg2, pnl, job, nwg, iflag);
  __syncthreads();
  // J <- Q2^T: Q2 formed and written transposed (rows of Q2 as columns, bx = n)
  q_form_tall(X, n, n, n, Tg2, Mq, J, n, nullptr, 0, nullptr, 0, pnl, job, nwg, iflag);
  for (int e = tid; e < n * n; e += bs) {  // X <- R2 (upper triangle of the second QR)
    const int i = e / n, j = e - i * n;
    if (j < i) X[e] = 0.0;
  }
  __syncthreads();
  const bool cv = jacobi_coop(X, n, n, J, job, nwg, red, iflag, pnl, pnl_cap, cut, false, max_sweeps);
  // X (n x b) <- (J R2) Q1^T: Q1 applied to [(J R2)^T; 0], written transposed (bx = b)
  q_form_tall(A, b, n, n, Tg1, Mq, X, b, X, n, nullptr, n, pnl, job, nwg, iflag);
  return cv;
}

// u <- R (x) u with the helper workgroups: the columns (aa, bb) split over all
// workgroups into Y, then Y copied into u's core (a second job: the core's
// new layout overlaps columns of other workgroups).
__device__ bool mode_product_coop(const double* Rrows, const int* perm, bool use_perm, int m, int b, const double* Uc,
                                  double* gY, double* Ud, int Au, int nu, int Bu, CoopJob* job, int nwg, int* iflag,
                                  lds_f64* lw, int prec) {
  if (threadIdx.x == 0) {
    job->type = 2;
    job->prec = prec;
    job->R = Rrows;
    job->perm = perm;
    job->use_perm = use_perm ? 1 : 0;
    job->m = m;
    job->rb = b;
    job->U = Uc;
    job->Y = gY;
    job->Au = Au;
    job->nu = nu;
    job->Bu = Bu;
  }
  if (!coop_post_run(job, nwg, iflag, lw)) return false;
  if (threadIdx.x == 0) {
    job->type = 3;
    job->D = Ud;
    job->ytot = (long long)Au * m * Bu;
  }
  return coop_post_run(job, nwg, iflag, lw);
}

// One move of the sweep, with the workspace `ws` either in LDS or in global
// scratch.  Two separate instantiations (one per memory) let the compiler use
// LDS instructions on the LDS path instead of generic flat addressing.
template <int G, bool REG>
__device__ __forceinline__ int run_move(double* ws, const bool gws, CoopJob* job, const int nwg, double* jpad, const bool jwave,
                                        lds_f64* pnl, const int pnl_cap, const int pmode, const int rows,
                                        const int b, const int tr, const int rpl, const int bx,
                                        const int64_t Au, const int64_t nu, const int64_t Bu, double* Xc,
                                        const double* Uc, double* Ud, double* gY, double cutoff, long long maxdim, int max_sweeps, int prec,
                                        int* status, int* iflag, double* red, double* dbc, double* gscr_any,
                                        lds_f64* pnl_any, double* lazyT, const double* uT) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int k = rows < b ? rows : b;
  // register QR for the shapes it covers (the QR of a tall SVD matrix included)
  // register QR for the QR steps: rpl = rows per lane bound / 8 (8: rows <= 64,
  // 16: rows <= 128, both 8-lane groups with 2 column slots; 20: rows <= 160,
  // 24: rows <= 192, 16-lane groups with 3 column slots)
  const bool use_reg = REG && G == 8 && rpl != 0;
  // workspace: W rows x b (or the reflector rows) | R k x b | J k x k (SVD only) |
  //            tau, sig, perm | 2 staging rows
  const int64_t jsz = tr ? (int64_t)k * k : 0;
  const int64_t wsz = use_reg ? max((int64_t)rows * b, (int64_t)k * 8 * rpl) : (int64_t)rows * b;
  // A tall QR move on the blocked path whose bond to u is the core's last leg
  // (bx = 1) has its matrix already row-major in x's core: it is factored
  // there (no copy into the workspace, and a kept Q's reflectors are then in
  // place too), which saves two passes over the core in global memory.
  const bool inplace = gws && !tr && pmode && bx == 1 && rows >= b;
  double* W = inplace ? Xc : ws;                // rows x b
  double* Rb = ws + wsz;                        // k x b
  double* Jb = Rb + (int64_t)k * b;             // k x k
  double* tau = Jb + jsz;                       // k
  double* sig = tau + k;                        // k
  int* perm = (int*)(sig + k);                  // k ints (2k doubles reserved)
  double* vbuf = sig + 3 * (int64_t)k;          // 2 rows
  // gather the matrix: M[a*Bx + bb][i] = X[(a*b + i)*Bx + bb] (sizes < 2^31, checked on the host)
  if (!inplace && gws && bx > 1) {
    // tiled transposes, with the helper workgroups when the matrix is large
    if (job != nullptr && nwg > 1 && (int64_t)rows * b >= (int64_t(1) << 15)) {
      if (tid == 0) {
        job->type = 10;
        job->U = Xc;
        job->D = W;
        job->Au = rows / bx;
        job->rb = b;
        job->Bu = bx;
      }
      const int64_t ntile = (int64_t)(rows / bx) * ((b + 63) >> 6) * ((bx + 63) >> 6);
      coop_post_run(job, nwg, iflag, pnl_any, (int)min<int64_t>(ntile, nwg));
    } else {
      gather_tiles(Xc, W, rows / bx, b, bx, 0, 1, pnl_any);
    }
  } else if (!inplace) {
    for (int e = tid; e < rows * b; e += bs) {
      const int row = e / b, i = e - row * b;
      const int aa = row / bx, bb = row - aa * bx;
      W[e] = Xc[((int64_t)aa * b + i) * bx + bb];
    }
  }
  __syncthreads();
  QTT_PHASE(-1, 1);  // gathered
  int m, n;
  const double* Rrows;
  bool use_perm = false, tall = false;
  if (!tr) {
    m = k;
    n = k;
    Rrows = Rb;
    if (use_reg) {
      lds_f64 *Wl = (lds_f64*)W, *taul = (lds_f64*)tau, *Rl = (lds_f64*)Rb;
      if (rpl == 8) qr_reg<8, 8, 2>(Wl, rows, b, taul, Rl);
      else if (rpl == 16) qr_reg<8, 16, 2>(Wl, rows, b, taul, Rl);
      else if (rpl == 20) qr_reg<16, 10, 3>(Wl, rows, b, taul, Rl);
      else qr_reg<16, 12, 3>(Wl, rows, b, taul, Rl);
      __syncthreads();
      QTT_PHASE(-1, 2);  // factored
      // x <- Q (bond m)
      if (rpl == 8) q_form_reg<8, 8>(Wl, rows, k, taul, Xc, bx);
      else if (rpl == 16) q_form_reg<8, 16>(Wl, rows, k, taul, Xc, bx);
      else if (rpl == 20) q_form_reg<16, 10>(Wl, rows, k, taul, Xc, bx);
      else q_form_reg<16, 12>(Wl, rows, k, taul, Xc, bx);
    } else if (pmode) {  // workspace in global memory: blocked QR, WY updates as MFMA tiles
      double* Tg = vbuf;
      double* Mq = Tg + (((int64_t)k + PNB) * PNB);
      if (pmode == 3) qr_tall(W, rows, b, b, tau, Tg, pnl, job, nwg, iflag);  // panels in global memory
      else qr_panel(W, rows, b, b, tau, Tg, pnl, job, nwg, iflag);            // panels in LDS
      for (int e = tid; e < k * b; e += bs) {
        const int i = e / b, c = e - i * b;
        Rb[e] = c >= i ? W[(int64_t)i * b + c] : 0.0;
      }
      __syncthreads();
      QTT_PHASE(-1, 2);  // factored
      if (lazyT) {
        // x keeps the reflectors (rows x k, ld k) and lazyT the panels' T: the
        // move that consumes x applies Q to its own thin factor (4 rows k m
        // flops for m columns) instead of Q being formed here (about 2 rows
        // k^2) and multiplied later (2 rows k m).  (In place, k = b: they are
        // there already.)
        if (!inplace && k == b && job != nullptr && nwg > 1 && (int64_t)rows * k >= (int64_t(1) << 15)) {
          // a plain copy, spread over the helper workgroups
          if (tid == 0) {
            job->type = 3;
            job->D = Xc;
            job->Y = W;
            job->ytot = (long long)rows * k;
          }
          coop_post_run(job, nwg, iflag, pnl_any);
        } else if (!inplace) {
          const int lane = tid & 63, wv = tid >> 6, nwv = bs >> 6;
          for (int r = wv; r < rows; r += nwv)
            for (int c = lane; c < k; c += 64) Xc[(int64_t)r * k + c] = W[(int64_t)r * b + c];
        }
        for (int e = tid; e < (k + PNB - 1) / PNB * PNB * PNB; e += bs) lazyT[e] = Tg[e];
      } else {
        q_form_tall(W, rows, k, b, Tg, Mq, Xc, bx, nullptr, 0, nullptr, 0, pnl, job, nwg, iflag);  // x <- Q (bond m)
      }
    } else {
      geqr2<G>(W, rows, b, b, tau);
      for (int e = tid; e < k * b; e += bs) {
        const int i = e / b, c = e - i * b;
        Rb[e] = c >= i ? W[(int64_t)i * b + c] : 0.0;
      }
      __syncthreads();
      QTT_PHASE(-1, 2);  // factored
      org2r<G>(W, rows, k, b, tau, vbuf);
      // x <- Q (bond m)
      for (int e = tid; e < rows * m; e += bs) {
        const int row = e / m, j = e - row * m;
        const int aa = row / bx, bb = row - aa * bx;
        Xc[((int64_t)aa * m + j) * bx + bb] = W[(int64_t)row * b + j];
      }
    }
  } else {
    // a square move that can drop nothing takes the tall path as well (QR
    // first), so that it can keep Q R below
    tall = rows > b || (rows == b && cutoff <= 0.0 && maxdim >= b);
    double* X = W;
    n = rows;
    if (tall) {
      if (use_reg) {  // the reflectors stay in W for U below
        lds_f64 *Wl = (lds_f64*)W, *taul = (lds_f64*)tau, *Rl = (lds_f64*)Rb;
        if (rpl == 8) qr_reg<8, 8, 2>(Wl, rows, b, taul, Rl);
        else if (rpl == 16) qr_reg<8, 16, 2>(Wl, rows, b, taul, Rl);
        else if (rpl == 20) qr_reg<16, 10, 3>(Wl, rows, b, taul, Rl);
        else qr_reg<16, 12, 3>(Wl, rows, b, taul, Rl);
      } else if (pmode) {  // blocked QR; the reflectors stay in W for U below
        if (pmode == 3) qr_tall(W, rows, b, b, tau, vbuf, pnl, job, nwg, iflag);
        else qr_panel(W, rows, b, b, tau, vbuf, pnl, job, nwg, iflag);
        for (int e = tid; e < b * b; e += bs) {
          const int i = e / b, c = e - i * b;
          Rb[e] = c >= i ? W[(int64_t)i * b + c] : 0.0;
        }
      } else {
        geqr2<G>(W, rows, b, b, tau);
        for (int e = tid; e < b * b; e += bs) {
          const int i = e / b, c = e - i * b;
          Rb[e] = c >= i ? W[(int64_t)i * b + c] : 0.0;
        }
        __syncthreads();
        org2r<G>(W, rows, b, b, tau, vbuf);
      }
      __syncthreads();
      X = Rb;
      n = b;
    }
    QTT_PHASE(-1, 2);  // factored (tall) / gathered (wide)
    // lanes per row pair sized so a row is at most 4 entries per lane.
    // cut: pairs of rows below its noise level are left alone; init_j: start
    // J from the identity (else continue the rotations already in it)
    auto jacobi = [&](const double cut, const bool init_j) -> bool {
      bool cv;
      if (REG && G == 8 && jwave) {
        // small matrices: the single-wavefront rounds on a zero-padded copy
        const int gw = wave_group(n, b);
        const int bp = (b + gw - 1) / gw * gw, np = (n + gw - 1) / gw * gw;
        const int ldx = bp + 1, ldj = np + 1;  // odd strides: the rows of a round fall in different banks
        lds_f64* Xl = (lds_f64*)jpad;
        lds_f64* Jl = Xl + n * ldx;
        for (int e = tid; e < n * bp; e += bs) {
          const int r = e / bp, c = e - r * bp;
          Xl[r * ldx + c] = c < b ? X[r * b + c] : 0.0;
        }
        for (int e = tid; e < n * np; e += bs) {
          const int r = e / np, c = e - r * np;
          Jl[r * ldj + c] = c >= n ? 0.0 : init_j ? (r == c ? 1.0 : 0.0) : Jb[r * n + c];
        }
        __syncthreads();
        cv = jacobi_wave<8, 4, 2>((lds_f64*)Xl, ldx, n, b, bp, Jl, ldj, np, (lds_i32*)iflag, red, cut, max_sweeps);
        __syncthreads();
        for (int e = tid; e < n * b; e += bs) X[e] = Xl[(e / b) * ldx + e % b];
        for (int e = tid; e < n * n; e += bs) Jb[e] = Jl[(e / n) * ldj + e % n];
        __syncthreads();
      } else if (REG && G == 8 && b <= 256) {
        // (REG: the workspace is in LDS.)  With room for it, the rounds run on a
        // copy whose row strides are odd numbers of doubles, so the rows that
        // the groups of a wavefront read together fall in different banks.
        lds_f64 *Xl = (lds_f64*)X, *Jl = (lds_f64*)Jb;
        int ldx = b, ldj = n;
        if (jpad) {
          ldx = b | 1;
          ldj = n | 1;
          Xl = (lds_f64*)jpad;
          Jl = Xl + n * ldx;
          for (int e = tid; e < n * b; e += bs) Xl[(e / b) * ldx + e % b] = X[e];
          if (!init_j)
            for (int e = tid; e < n * n; e += bs) Jl[(e / n) * ldj + e % n] = Jb[e];
          __syncthreads();
        }
        lds_i32* fl = (lds_i32*)iflag;
        const int ij = init_j ? 1 : 0;
        if (b <= 32) cv = jacobi_rows_reg<8>(Xl, ldx, n, b, Jl, ldj, fl, red, cut, ij, max_sweeps);
        else if (b <= 64) cv = jacobi_rows_reg<16>(Xl, ldx, n, b, Jl, ldj, fl, red, cut, ij, max_sweeps);
        else if (b <= 128) cv = jacobi_rows_reg<32>(Xl, ldx, n, b, Jl, ldj, fl, red, cut, ij, max_sweeps);
        else cv = jacobi_rows_reg<64>(Xl, ldx, n, b, Jl, ldj, fl, red, cut, ij, max_sweeps);
        if (jpad) {
          __syncthreads();
          for (int e = tid; e < n * b; e += bs) X[e] = Xl[(e / b) * ldx + e % b];
          for (int e = tid; e < n * n; e += bs) Jb[e] = Jl[(e / n) * ldj + e % n];
          __syncthreads();
        }
      } else if (gws && job && nwg > 1 && n >= 32 && b <= 512) {
        // rows in global memory and helper workgroups present: each round's
        // pairs are spread over all of them
        cv = jacobi_coop(X, n, b, Jb, job, nwg, red, iflag, pnl, pnl ? pnl_cap : 0, cut, init_j, max_sweeps);
      } else if (gws && b <= 512) {
        // workspace in global memory: both rows of a pair (and of J) in
        // registers, 32 lanes per pair (two pairs per wavefront; 64 for rows
        // over 384), all loads of a pair in flight at once
        lds_i32* fl = (lds_i32*)iflag;
        const int ij = init_j ? 1 : 0;
        if (b <= 128) cv = jacobi_rows_reg<32, 4, double>(X, b, n, b, Jb, n, fl, red, cut, ij, max_sweeps);
        else if (b <= 256) cv = jacobi_rows_reg<32, 8, double>(X, b, n, b, Jb, n, fl, red, cut, ij, max_sweeps);
        else if (b <= 384) cv = jacobi_rows_reg<32, 12, double>(X, b, n, b, Jb, n, fl, red, cut, ij, max_sweeps);
        else cv = jacobi_rows_reg<64, 8, double>(X, b, n, b, Jb, n, fl, red, cut, ij, max_sweeps);
      } else {
        // eight lanes per pair: every group of a wavefront has a pair of its
        // own, so the rows of 64 pairs are read at once
        cv = jacobi_rows<8>(X, n, b, Jb, iflag, red, cut, init_j, max_sweeps);
      }
      return cv;
    };
    // wide matrices of at least 192 rows on the global workspace with helper
    // workgroups: the QR-preconditioned rounds (3D comb operate: 256 x 384
    // moves 25.3 -> 12.7 ms; at 128 x 128 the two QRs cost what the saved
    // sweeps do)
    const bool precond = !REG && gws && job && nwg > 1 && !tall && n >= 192 && b <= 512 && pnl != nullptr;
    // A tall move that can drop nothing (cutoff <= 0: at most exactly zero singular
    // values would go; maxdim >= n) and whose R has no zero on its diagonal
    // (nonsingular, so none of its singular values is zero) keeps all n: Q R
    // is then already the factorization the truncation would produce, up to
    // an orthogonal change of basis on the new bond.  The rows of R stand for
    // S V^T (J = I), sorted by norm below, and the Jacobi is not run.
    bool keep_all = false;
    if (tall && cutoff <= 0.0 && maxdim >= n) {
      if (tid == 0) iflag[11] = 1;
      __syncthreads();
      for (int i = tid; i < n; i += bs)
        if (X[(int64_t)i * b + i] == 0.0) iflag[11] = 0;
      __syncthreads();
      keep_all = iflag[11] != 0;
      if (keep_all)
        for (int e = tid; e < n * n; e += bs) Jb[e] = (e / n == e - (e / n) * n) ? 1.0 : 0.0;
      __syncthreads();
    }
    bool conv = keep_all ? true
                : precond ? jacobi_precond(X, n, b, Rb, Jb, tau, vbuf, vbuf + ((int64_t)n + PNB) * PNB, gY, pnl,
                                           pnl_cap, job, nwg, red, iflag, cutoff, max_sweeps)
                          : jacobi(cutoff, true);
    row_norms<G>(X, n, b, sig);
    __syncthreads();
    if (conv && cutoff > 0.0 && cutoff < 1.0 && !noise_block_exact(sig, n, cutoff, iflag + 10)) {
      // the noise block could change what is kept: finish the SVD among its
      // rows (pairs with a row above the level are converged and skipped)
      conv = jacobi(0.0, false);
      row_norms<G>(X, n, b, sig);
      __syncthreads();
    }
    if (!conv) {
      if (tid == 0) status[1] = 1;
    }
    for (int i = tid; i < n; i += bs) {  // by decreasing norm (any order when all are kept)
      int r = 0;
      const double si = sig[i];
      if (!keep_all)
        for (int j = 0; j < n; ++j) r += (sig[j] > si || (sig[j] == si && j < i)) ? 1 : 0;
      perm[keep_all ? i : r] = i;
    }
    __syncthreads();
    if (tid == 0) {
      // ITensors/NDTensors truncate!!: relative cutoff on squared singular values
      double scale = 0.0;
      for (int i = 0; i < n; ++i) scale += sig[perm[i]] * sig[perm[i]];
      int mm = n;
      double terr = 0.0;
      const long long md = maxdim < 1 ? 1 : maxdim;
      while (mm > md) {
        terr += sig[perm[mm - 1]] * sig[perm[mm - 1]];
        --mm;
      }
      const double sc = scale == 0.0 ? 1.0 : scale;
      if (cutoff >= 0.0) {
        while (mm > 1) {
          const double pm = sig[perm[mm - 1]] * sig[perm[mm - 1]];
          if (terr + pm <= cutoff * sc) {
            terr += pm;
            --mm;
          } else {
            break;
          }
        }
      }
      iflag[1] = mm;
      dbc[0] += terr / sc;
    }
    __syncthreads();
    m = iflag[1];
    Rrows = X;
    use_perm = !keep_all;  // kept all: U = Q and the rows of R in order (triangular)
    // x <- U (bond m): U[row][i] = sum_j Q[row][j] J[perm[i]][j]  (or J[perm[i]][row])
    if (tall && pmode) {
      q_form_tall(W, rows, b, b, vbuf, vbuf + (((int64_t)b + PNB) * PNB), Xc, bx, keep_all ? nullptr : Jb, n,
                  keep_all ? nullptr : perm, m, pnl, job, nwg, iflag);
    } else if (tall && use_reg) {
      const lds_f64 *Wl = (const lds_f64*)W, *taul = (const lds_f64*)tau, *Jl = (const lds_f64*)Jb;
      const lds_i32* pl = (const lds_i32*)perm;
      if (rpl == 8) u_form_reg<8, 8>(Wl, rows, b, taul, Jl, n, pl, m, Xc, bx);
      else if (rpl == 16) u_form_reg<8, 16>(Wl, rows, b, taul, Jl, n, pl, m, Xc, bx);
      else if (rpl == 20) u_form_reg<16, 10>(Wl, rows, b, taul, Jl, n, pl, m, Xc, bx);
      else u_form_reg<16, 12>(Wl, rows, b, taul, Jl, n, pl, m, Xc, bx);
    } else {
    for (int e = tid; e < rows * m; e += bs) {
      const int row = e / m, i = e - row * m;
      const double* jr = Jb + (int64_t)perm[i] * n;
      double s;
      if (tall) {
        s = 0.0;
        const double* wr = W + (int64_t)row * b;
        for (int j = 0; j < n; ++j) s = __builtin_fma(wr[j], jr[j], s);
      } else {
        s = jr[row];
      }
      const int aa = row / bx, bb = row - aa * bx;
      Xc[((int64_t)aa * m + i) * bx + bb] = s;
    }
    }
  }
  __syncthreads();
  QTT_PHASE(-1, 3);  // x written
  // u <- R (x)_{au} u : Y[(aa*m + i)*Bu + bb] = sum_c R[i][c] U[(aa*nu + c)*Bu + bb]
  const int ytot = (int)(Au * m * Bu);
  if (uT) {
    // u holds the reflectors of its own QR toward x (Q kept, not formed):
    // u <- Q_u [Z; 0], Z = R^T (nu x m, the rows of R in perm order), in FP64
    const double* zs = Rrows;
    const int* zp = use_perm ? perm : nullptr;
    if (!gws) {  // R in LDS, which the apply works in: copied to global scratch first
      for (int e = tid; e < m * b; e += bs) {
        const int i = e / b, c = e - i * b;
        gscr_any[e] = Rrows[(int64_t)(use_perm ? perm[i] : i) * b + c];
      }
      __syncthreads();
      zs = gscr_any;
      zp = nullptr;
    }
    q_form_tall(Uc, (int)(Au * Bu), (int)nu, (int)nu, uT, gY, Ud, (int)Bu, zs, b, zp, m, pnl_any, job, nwg, iflag);
  } else if (!REG && gws && job && nwg > 1 && Au * Bu * nu * (int64_t)m >= (int64_t(1) << 22)) {
    // a large product in global memory: columns spread over the helper workgroups
    if (!mode_product_coop(Rrows, perm, use_perm, m, b, Uc, gY, Ud, (int)Au, (int)nu, (int)Bu, job, nwg, iflag, pnl,
                           prec)) {
      if (tid == 0) status[1] = 1;
    }
  } else {
    // the bulk product in the precision asked for (QR, SVD and cutoff stay float64)
    if (REG) {
      const lds_f64* Rl = (const lds_f64*)Rrows;
      const lds_i32* pl = (const lds_i32*)perm;
      if (prec == QTT_F32) mode_product_lds<QTT_F32>(Rl, pl, use_perm, m, b, Uc, gY, (int)Au, (int)nu, (int)Bu);
      else if (prec == QTT_F32_SPLIT)
        mode_product_lds<QTT_F32_SPLIT>(Rl, pl, use_perm, m, b, Uc, gY, (int)Au, (int)nu, (int)Bu);
      else mode_product_lds<QTT_F64>(Rl, pl, use_perm, m, b, Uc, gY, (int)Au, (int)nu, (int)Bu);
    } else if (gws) {  // global workspace: MFMA tiles (float-float on the VALU)
      if (prec == QTT_F32)
        mode_product_mfma32(Rrows, perm, use_perm, m, b, Uc, gY, (int)nu, (int)Bu, 0, (int)(Au * Bu));
      else if (prec == QTT_F32_SPLIT)
        mode_product<QTT_F32_SPLIT>(Rrows, perm, use_perm, m, b, Uc, gY, (int)Au, (int)nu, (int)Bu);
      else
        mode_product_mfma(Rrows, perm, use_perm, m, b, Uc, gY, (int)Au, (int)nu, (int)Bu, 0, (int)(Au * Bu));
    } else {
      if (prec == QTT_F32) mode_product<QTT_F32>(Rrows, perm, use_perm, m, b, Uc, gY, (int)Au, (int)nu, (int)Bu);
      else if (prec == QTT_F32_SPLIT)
        mode_product<QTT_F32_SPLIT>(Rrows, perm, use_perm, m, b, Uc, gY, (int)Au, (int)nu, (int)Bu);
      else mode_product<QTT_F64>(Rrows, perm, use_perm, m, b, Uc, gY, (int)Au, (int)nu, (int)Bu);
    }
    __syncthreads();
    for (int e = tid; e < ytot; e += bs) Ud[e] = gY[e];
  }
  QTT_PHASE(-1, 4);  // u written
  return m;
}

__global__ void __launch_bounds__(512) truncate_kernel(const SweepArgs* __restrict__ all) {
  extern __shared__ double lds[];
  // a single-network launch may carry helper workgroups after workgroup 0
  if (all[0].nwg > 1 && blockIdx.x > 0) {
    coop_helper(all[0].job, (int)blockIdx.x, all[0].nwg, (lds_i32*)(lds + 16), (lds_f64*)(lds + 64));
    return;
  }
  const SweepArgs& a = all[blockIdx.x];
  const TopoView& t = a.t;
  const int tid = threadIdx.x, bs = blockDim.x;
  // LDS layout: [misc 64 doubles][bonds: nv ints][kept-Q state: nv ints][workspace]
  double* red = lds;                      // 16 doubles
  int* iflag = (int*)(lds + 16);          // flags and broadcast ints
  double* dbc = lds + 24;                 // broadcast doubles
  const int nvi = (t.nv + 1) & ~1;
  int* bond = (int*)(lds + 64);
  // kq[v] = w: v's core holds the reflectors of its QR toward w (Q not formed
  // yet), its T factors at a.gT + v a.tstride; -1: v's core is plain
  int* kq = bond + nvi;
  double* wsl = lds + 64 + nvi;
  // LDS of the blocked QR's panels and WY updates (TALL_LDS doubles), 16-byte aligned
  const int poff = (64 + nvi + 1) & ~1;
  lds_f64* pnl_any = (lds_f64*)(lds + poff);
  for (int v = tid; v < t.nv; v += bs) {
    bond[v] = a.bond_g[v];
    kq[v] = -1;
  }
  if (tid == 0) dbc[0] = 0.0;
  __syncthreads();
  // Q of v formed from its kept reflectors, in v's core (gY as the scratch:
  // a core's size is at most its capacity)
  auto form_kept_q = [&](int v) {
    int sh[3];
    move_shape(t, bond, v, kq[v], sh);
    double* Xv = a.data + a.off[v];
    q_form_tall(Xv, sh[0] * sh[2], sh[1], sh[1], a.gT + (int64_t)v * a.tstride, a.gY, Xv, sh[2], nullptr, 0, nullptr,
                0, pnl_any, a.job, a.nwg, iflag);
    __syncthreads();
    if (tid == 0) kq[v] = -1;
    __syncthreads();
  };
  for (int st = 0; st < a.nsteps; ++st) {
    const int x = a.steps[3 * st], u = a.steps[3 * st + 1], tr = a.steps[3 * st + 2];
    QTT_PHASE(st, 0);
    // the move's shape, computed once and kept in scalar registers (all sizes
    // are < 2^31, checked on the host)
    if (tid == 0) {
      move_shape(t, bond, x, u, iflag + 4);
      move_shape(t, bond, u, x, iflag + 7);
    }
    __syncthreads();
    int sx[3], su[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      sx[i] = __builtin_amdgcn_readfirstlane(iflag[4 + i]);
      su[i] = __builtin_amdgcn_readfirstlane(iflag[7 + i]);
    }
    const int64_t Bx = sx[2], Au = su[0], nu = su[1], Bu = su[2];
    const int rows = sx[0] * sx[2], b = sx[1];
    const int k = rows < b ? rows : b;
    // A kept Q is applied by the move along its edge that takes it as u, when
    // that move's new bond is at most half the kept one (the apply then costs
    // at most what forming Q and multiplying it would); any other use forms it
    // first.
    if (kq[x] >= 0) form_kept_q(x);
    const long long mbound = tr ? min((long long)k, a.maxdim < 1 ? 1LL : a.maxdim) : (long long)k;
    if (kq[u] >= 0 && (kq[u] != x || 2 * mbound > nu)) form_kept_q(u);
    const double* uT = kq[u] == x ? a.gT + (int64_t)u * a.tstride : nullptr;
    QTT_SHAPE(st, rows, b, tr);
    QTT_SWEEPS_CLEAR(st);
    const int64_t jsz = tr ? (int64_t)k * k : 0;
    const bool small = rows <= 256 && b <= 256;
    // register QR: rows <= 64 or <= 128 with columns <= 2 per 8-lane group, or
    // rows <= 160 or <= 192 with columns <= 3 per 16-lane group (the reflector
    // rows take k * 8 rpl doubles)
    int rpl = !small ? 0
                    : rows <= 64 && b <= 2 * (bs / 8) ? 8
                    : rows <= 128 && b <= 2 * (bs / 8) ? 16
                    : rows <= 160 && b <= 3 * (bs / 16) ? 20
                    : rows <= 192 && b <= 3 * (bs / 16) ? 24 : 0;
    const int64_t base = (int64_t)k * b + jsz + 4 * (int64_t)k + 2 * (int64_t)rows + 16;
    int64_t need = max((int64_t)rows * b, rpl ? (int64_t)k * 8 * rpl : 0) + base;
    if (rpl && need > a.ws_lds) {  // the reflector rows do not fit: plain LDS path
      rpl = 0;
      need = (int64_t)rows * b + base;
    }
    // global workspace: the blocked QR's LDS (TALL_LDS doubles) holds the
    // panel when rows * PLD fit its first part (pmode 1), else the panels stay
    // in global memory (pmode 3)
    lds_f64* pnl = nullptr;
    int pnl_cap = 0;  // doubles of LDS at pnl
    int pmode = 0;
    if (need > a.ws_lds && TALL_LDS + 64 <= a.ws_lds) pmode = (int64_t)rows * PLD <= 8 * TALL_REGION ? 1 : 3;
    if (pmode) {
      pnl = pnl_any;
      pnl_cap = QTT_LDS_DOUBLES - poff;
    }
    // a QR move on the blocked path keeps its reflectors instead of forming Q
    double* lazyT = (pmode && !tr && a.gT) ? a.gT + (int64_t)x * a.tstride : nullptr;
    // padded copy for the Jacobi rounds of an SVD move, when it fits after the workspace
    double* jpad = nullptr;
    bool jwave = false;  // room for the single-wavefront Jacobi's copy at jpad
    if (tr) {
      const int nj = rows > b ? b : rows;
      if (need + (int64_t)nj * ((b | 1) + (nj | 1)) + 16 <= a.ws_lds) jpad = wsl + need;
      const int64_t wl = wave_jacobi_lds(nj, b);
      jwave = jpad && wl && need + wl <= a.ws_lds;
    }
    double* Xc = a.data + a.off[x];
    const double* Uc = a.data + a.off[u];
    double* Ud = a.data + a.off[u];
    int m;
    if (need <= a.ws_lds) {
      if (small)
        m = run_move<8, true>(wsl, false, nullptr, 1, jpad, jwave, nullptr, 0, 0, rows, b, tr, rpl, (int)Bx, Au, nu, Bu, Xc, Uc, Ud, a.gY, a.cutoff, a.maxdim, a.max_sweeps, a.prec, a.status,
                        iflag, red, dbc, a.gscr, pnl_any, lazyT, uT);
      else
        m = run_move<64, false>(wsl, false, nullptr, 1, nullptr, false, nullptr, 0, 0, rows, b, tr, 0, (int)Bx, Au, nu, Bu, Xc, Uc, Ud, a.gY, a.cutoff, a.maxdim, a.max_sweeps, a.prec, a.status,
                         iflag, red, dbc, a.gscr, pnl_any, lazyT, uT);
    } else {
      if (small)
        m = run_move<8, false>(a.gscr, true, a.job, a.nwg, nullptr, false, pnl, pnl_cap, pmode, rows, b, tr, 0, (int)Bx, Au, nu, Bu, Xc, Uc, Ud, a.gY, a.cutoff, a.maxdim, a.max_sweeps, a.prec, a.status,
                        iflag, red, dbc, a.gscr, pnl_any, lazyT, uT);
      else
        m = run_move<64, false>(a.gscr, true, a.job, a.nwg, nullptr, false, pnl, pnl_cap, pmode, rows, b, tr, 0, (int)Bx, Au, nu, Bu, Xc, Uc, Ud, a.gY, a.cutoff, a.maxdim, a.max_sweeps, a.prec, a.status,
                         iflag, red, dbc, a.gscr, pnl_any, lazyT, uT);
    }
    QTT_PHASE(st, 5);
    if (tid == 0) {
      if (t.parent[x] == u) bond[x] = m; else bond[u] = m;
      if (uT) kq[u] = -1;
      if (lazyT) kq[x] = u;
    }
    __syncthreads();
  }
  for (int v = 0; v < t.nv; ++v)
    if (kq[v] >= 0) form_kept_q(v);
  for (int v = tid; v < t.nv; v += bs) a.bond_g[v] = bond[v];
  if (tid == 0) a.err_out[0] = dbc[0];
  if (a.nwg > 1 && tid == 0) __hip_atomic_store(&a.job->quit, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

int launch_contract(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, qtt_net* out, int precision) {
  const HostTopo& t = net->t;
  std::vector<int64_t> cstart(t.nv + 1, 0);
  for (int v = 0; v < t.nv; v++) cstart[v + 1] = cstart[v] + out->cap[v];
  double* scr = qtt_scratch(ctx, (t.nv + 1) * sizeof(int64_t) + 64);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "contract scratch");
  int64_t* d_cstart = (int64_t*)scr;
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_cstart, cstart.data(), (t.nv + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream), "cstart");
  if (rc) return rc;
  const int64_t total = cstart[t.nv];
  int64_t nb = std::min<int64_t>((total + 255) / 256, 65536);
  const dim3 grid((unsigned)std::max<int64_t>(nb, 1));
#define QTT_CONTRACT_LAUNCH(M)                                                                              \
  contract_kernel<M><<<grid, 256, 0, ctx->stream>>>(op->d_data, op->d_off, op->d_bond, net->d_data,        \
                                                    net->d_off, net->d_bond, net->view(), out->d_data, out->d_off, \
                                                    d_cstart, total)
  if (precision == QTT_F32) QTT_CONTRACT_LAUNCH(QTT_F32);
  else if (precision == QTT_F32_SPLIT) QTT_CONTRACT_LAUNCH(QTT_F32_SPLIT);
  else QTT_CONTRACT_LAUNCH(QTT_F64);
#undef QTT_CONTRACT_LAUNCH
  rc = qtt_hip_check(ctx, hipGetLastError(), "contract_kernel");
  if (rc) return rc;
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "contract sync");
}

// Exact contraction of a batch in one launch, in two phases so that a
// pipeline can queue the kernel without host synchronisation: stage (item
// table, element starts and core starts into device memory at `scr`,
// asynchronously from st.host, which must live until the stream passes the
// copy) and launch.
size_t contract_batch_bytes(int n, qtt_net* const* outs) {
  if (n <= 0) return 0;
  int64_t ctot = 0, nseg = 0;
  for (int i = 0; i < n; i++) {
    ctot += outs[i]->t.nv + 1;
    for (int v = 0; v < outs[i]->t.nv; v++) nseg += (outs[i]->cap[v] + kSegElems - 1) / kSegElems;
  }
  const size_t ib = ((size_t)n * sizeof(ContractItem) + 15) & ~size_t(15);
  return ib + (size_t)(n + 1 + ctot) * sizeof(int64_t) + (size_t)nseg * sizeof(ContractSeg) + 64;
}

int contract_batch_stage(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets,
                         qtt_net* const* outs, int precision, char* scr, BatchStage& st) {
  st.n = n;
  st.precision = precision;
  st.dev = scr;
  if (n <= 0) return QTT_OK;
  // scratch: items | estart (n + 1) | cstart of every network (nv + 1 each)
  std::vector<int64_t> estart(n + 1, 0), cbase(n);
  int64_t ctot = 0;
  for (int i = 0; i < n; i++) {
    estart[i + 1] = estart[i];
    for (int v = 0; v < outs[i]->t.nv; v++) estart[i + 1] += outs[i]->cap[v];
    cbase[i] = ctot;
    ctot += outs[i]->t.nv + 1;
  }
  const size_t ib = ((size_t)n * sizeof(ContractItem) + 15) & ~size_t(15);
  const size_t bytes = ib + (size_t)(n + 1 + ctot) * sizeof(int64_t);
  st.host.assign(bytes, 0);
  ContractItem* hi = (ContractItem*)st.host.data();
  int64_t* hes = (int64_t*)(st.host.data() + ib);
  int64_t* hcs = hes + n + 1;
  const int64_t* d_es = (const int64_t*)(scr + ib);
  const int64_t* d_cs = d_es + n + 1;
  std::copy(estart.begin(), estart.end(), hes);
  for (int i = 0; i < n; i++) {
    const qtt_net* o = outs[i];
    int64_t* cs = hcs + cbase[i];
    cs[0] = 0;
    for (int v = 0; v < o->t.nv; v++) cs[v + 1] = cs[v] + o->cap[v];
    hi[i] = {ops[i]->d_data, ops[i]->d_off, ops[i]->d_bond, nets[i]->d_data, nets[i]->d_off, nets[i]->d_bond,
             nets[i]->view(), o->d_data, o->d_off, d_cs + cbase[i]};
  }
  st.total = estart[n];
  st.aux = (char*)d_es;
  // segments of the 32-bit kernel (cores below 2^31 elements), after the tables
  bool small = true;
  int64_t nseg = 0;
  for (int i = 0; i < n && small; i++)
    for (int v = 0; v < outs[i]->t.nv; v++) {
      if (outs[i]->cap[v] >= (int64_t(1) << 31)) small = false;
      nseg += (outs[i]->cap[v] + kSegElems - 1) / kSegElems;
    }
  st.nseg = 0;  // segments in use (0: the searching kernel)
  if (small && nseg > 0 && nseg < (int64_t(1) << 31)) {  // contract_batch_bytes left room for them
    const size_t off = st.host.size();
    st.host.resize(off + (size_t)nseg * sizeof(ContractSeg));
    ContractSeg* hs = (ContractSeg*)(st.host.data() + off);
    int64_t j = 0;
    for (int i = 0; i < n; i++)
      for (int v = 0; v < outs[i]->t.nv; v++)
        for (int64_t e0 = 0; e0 < outs[i]->cap[v]; e0 += kSegElems)
          hs[j++] = {i, v, (int)e0, (int)std::min<int64_t>(kSegElems, outs[i]->cap[v] - e0)};
    st.nseg = nseg;
    st.seg = scr + bytes;
  }
  return qtt_hip_check(ctx, hipMemcpyAsync(scr, st.host.data(), st.host.size(), hipMemcpyHostToDevice, ctx->stream),
                       "contract items");
}

int contract_batch_launch(qtt_ctx* ctx, const BatchStage& st) {
  if (st.n <= 0) return QTT_OK;
  const ContractItem* items = (const ContractItem*)st.dev;
  if (st.nseg > 0) {
    const ContractSeg* segs = (const ContractSeg*)st.seg;
    const dim3 g((unsigned)st.nseg);
    if (st.precision == QTT_F32)
      contract_seg_kernel<QTT_F32><<<g, 256, 0, ctx->stream>>>(items, segs);
    else if (st.precision == QTT_F32_SPLIT)
      contract_seg_kernel<QTT_F32_SPLIT><<<g, 256, 0, ctx->stream>>>(items, segs);
    else
      contract_seg_kernel<QTT_F64><<<g, 256, 0, ctx->stream>>>(items, segs);
    return qtt_hip_check(ctx, hipGetLastError(), "contract_seg_kernel");
  }
  const int64_t* d_es = (const int64_t*)st.aux;
  const int64_t nb = std::min<int64_t>((st.total + 255) / 256, 65536);
  const dim3 grid((unsigned)std::max<int64_t>(nb, 1));
  if (st.precision == QTT_F32)
    contract_batch_kernel<QTT_F32><<<grid, 256, 0, ctx->stream>>>(items, d_es, st.n, st.total);
  else if (st.precision == QTT_F32_SPLIT)
    contract_batch_kernel<QTT_F32_SPLIT><<<grid, 256, 0, ctx->stream>>>(items, d_es, st.n, st.total);
  else
    contract_batch_kernel<QTT_F64><<<grid, 256, 0, ctx->stream>>>(items, d_es, st.n, st.total);
  return qtt_hip_check(ctx, hipGetLastError(), "contract_batch_kernel");
}

int launch_contract_batch(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets,
                          qtt_net* const* outs, int precision) {
  if (n <= 0) return QTT_OK;
  char* scr = (char*)qtt_scratch(ctx, contract_batch_bytes(n, outs));
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "contract scratch");
  BatchStage st;
  int rc = contract_batch_stage(ctx, n, ops, nets, outs, precision, scr, st);
  if (!rc) rc = contract_batch_launch(ctx, st);
  if (rc) return rc;
  // the host tables must outlive the copy
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "contract batch sync");
}

// Truncate a batch of networks, one workgroup each, in one launch.  In
// phases, so that a pipeline can queue the launch between other kernels
// without host synchronisation (qtt_batch_run): stage (plan, argument tables
// into device memory at `scr`), launch (the kernel only), readback (queued
// copies of errors and bonds), finish (after the stream passed the copies:
// host bond mirrors, errors, status).
struct TruncStaged::Impl {
  std::vector<SweepArgs> hargs;
  std::vector<int> allsteps;
  CoopJob hjob;
  std::vector<double> herr;
  std::vector<std::vector<int>> b32;
  SweepArgs* d_args = nullptr;
  double* d_err = nullptr;
  CoopJob* d_job = nullptr;
};

TruncStaged::TruncStaged() : impl(new Impl) {}
TruncStaged::~TruncStaged() { delete impl; }

// doubles of T factors per vertex for a kept Q: one PNB x PNB block per panel
// of at most maxb columns
static int64_t trunc_tstride(int64_t maxb) { return (maxb + PNB - 1) / PNB * PNB * PNB; }

static size_t trunc_layout(qtt_net* const* nets, int nnet, int nwg, std::vector<int64_t>* ws_out,
                           std::vector<int64_t>* soff_out, std::vector<std::vector<int>>* steps_out,
                           int64_t* tot_doubles_out, int64_t* tot_steps_out) {
  std::vector<std::vector<int>> steps(nnet);
  std::vector<int64_t> ws(nnet), soff(nnet);
  int64_t tot_doubles = 0, tot_steps = 0;
  for (int i = 0; i < nnet; i++) {
    const qtt_net* net = nets[i];
    steps[i] = truncation_schedule(net->t);
    int64_t maxcap = 1, maxb = 1;
    for (int v = 0; v < net->t.nv; v++) {
      maxcap = std::max(maxcap, net->cap[v]);
      maxb = std::max(maxb, net->bond[v]);
    }
    // workspace bound: W (or the reflector rows, <= max(2 core, 64 maxb)), R, J,
    // staging rows (<= 2 core), vectors; then the T factors of kept Q's
    // (trunc_tstride per vertex) and the Y buffer
    ws[i] = 4 * maxcap + 2 * maxb * maxb + 68 * maxb + 64 + (int64_t)net->t.nv * trunc_tstride(maxb) + maxcap;
    soff[i] = tot_steps;
    tot_doubles += ws[i];
    tot_steps += (int64_t)steps[i].size();
  }
  const size_t args_bytes = (size_t)nnet * sizeof(SweepArgs);
  // the tail slack covers the whole-block row reads of the Jacobi rounds (< 512 doubles past a row)
  // helpers' exchange buffers after the job: Gp [QTT_MAX_HELPERS + 1][16], Pg [512][256], Ug [8][256],
  // then the block counters of the Jacobi sweeps (at most 128 blocks; 256 ints)
  const size_t coop_doubles = (QTT_MAX_HELPERS + 1) * 16 + (size_t)COOP_PG_UNITS * 256 + 8 * 256 + 128;
  const size_t bytes = (tot_doubles + nnet + 8) * sizeof(double) + ((args_bytes + 255) & ~size_t(255)) +
                       ((tot_steps * sizeof(int) + 255) & ~size_t(255)) + 4096 + sizeof(CoopJob) + 256 +
                       (nwg > 1 ? coop_doubles * sizeof(double) : 0);
  if (ws_out) *ws_out = ws;
  if (soff_out) *soff_out = soff;
  if (steps_out) *steps_out = std::move(steps);
  if (tot_doubles_out) *tot_doubles_out = tot_doubles;
  if (tot_steps_out) *tot_steps_out = tot_steps;
  return bytes;
}

static int trunc_nwg(qtt_ctx* ctx, qtt_net* const* nets, int nnet) {
  // A single network with cores beyond LDS gets helper workgroups for the
  // Jacobi rounds of its largest moves.
  int nwg = 1;
  if (nnet == 1) {
    int64_t maxcap = 1;
    for (int v = 0; v < nets[0]->t.nv; v++) maxcap = std::max(maxcap, nets[0]->cap[v]);
    if (maxcap > QTT_LDS_DOUBLES) nwg = std::max(1, std::min(ctx->trunc_helpers + 1, ctx->cus));
  }
  return nwg;
}

size_t truncate_batch_bytes(qtt_ctx* ctx, qtt_net* const* nets, int nnet) {
  if (nnet <= 0) return 0;
  return trunc_layout(nets, nnet, trunc_nwg(ctx, nets, nnet), nullptr, nullptr, nullptr, nullptr, nullptr);
}

int truncate_batch_stage(qtt_ctx* ctx, qtt_net* const* nets, int nnet, double cutoff, int64_t maxdim,
                         int precision, double* scr, TruncStaged& st) {
  st.nnet = nnet;
  if (nnet <= 0) return QTT_OK;
  TruncStaged::Impl* im = st.impl;
  for (int i = 0; i < nnet; i++) {
    int64_t maxcap = 1;
    for (int v = 0; v < nets[i]->t.nv; v++) maxcap = std::max(maxcap, nets[i]->cap[v]);
    if (maxcap >= (int64_t(1) << 31) / 8)
      return qtt_fail(ctx, QTT_EINVAL, "truncate: a core is too large for the sweep kernel");
  }
  const int nwg = trunc_nwg(ctx, nets, nnet);
  std::vector<int64_t> ws, soff;
  std::vector<std::vector<int>> steps;
  int64_t tot_doubles = 0, tot_steps = 0;
  (void)trunc_layout(nets, nnet, nwg, &ws, &soff, &steps, &tot_doubles, &tot_steps);
  std::vector<double> cost(nnet);
  for (int i = 0; i < nnet; i++) {
    cost[i] = 0.0;
    for (int v = 0; v < nets[i]->t.nv; v++)
      cost[i] += (double)nets[i]->cap[v] * (double)nets[i]->bond[v];  // ~ a move's factorization work
  }
  const size_t args_bytes = (size_t)nnet * sizeof(SweepArgs);
  double* d_err = scr;
  double* wsp = scr + nnet + 8;
  SweepArgs* d_args = (SweepArgs*)(wsp + tot_doubles);
  int* d_steps = (int*)((char*)d_args + ((args_bytes + 255) & ~size_t(255)));
  CoopJob* d_job = (CoopJob*)((char*)d_steps + ((tot_steps * sizeof(int) + 255) & ~size_t(255)));
  im->allsteps.clear();
  im->allsteps.reserve(tot_steps);
  for (int i = 0; i < nnet; i++) im->allsteps.insert(im->allsteps.end(), steps[i].begin(), steps[i].end());
  const int lds_doubles = QTT_LDS_DOUBLES;
  // Workgroups are dispatched in index order and the networks of a batch are
  // independent: the costliest networks get the first workgroups, so that a
  // large one does not start last (longest-first list scheduling).
  std::vector<int> ord(nnet);
  for (int i = 0; i < nnet; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  im->hargs.assign(nnet, SweepArgs{});
  double* w = wsp;
  std::vector<double*> wbase(nnet);
  for (int i = 0; i < nnet; i++) {  // workspaces in network order, as the layout counted them
    wbase[i] = w;
    w += ws[i];
  }
  for (int j = 0; j < nnet; j++) {
    const int i = ord[j];
    qtt_net* net = nets[i];
    SweepArgs& sa = im->hargs[j];
    sa.data = net->d_data;
    sa.off = net->d_off;
    sa.bond_g = net->d_bond;
    sa.t = net->view();
    sa.steps = d_steps + soff[i];
    sa.nsteps = (int)steps[i].size() / 3;
    sa.cutoff = cutoff;
    sa.maxdim = (long long)maxdim;
    sa.max_sweeps = ctx->jacobi_max_sweeps;
    sa.prec = precision;
    sa.gscr = wbase[i];
    int64_t maxcap = 1, maxb = 1;
    for (int v = 0; v < net->t.nv; v++) {
      maxcap = std::max(maxcap, net->cap[v]);
      maxb = std::max(maxb, net->bond[v]);
    }
    sa.gY = wbase[i] + ws[i] - maxcap;
    sa.tstride = (int)trunc_tstride(maxb);
    sa.gT = sa.gY - (int64_t)net->t.nv * sa.tstride;
    sa.ws_lds = lds_doubles - 64 - ((net->t.nv + 1) & ~1);
    sa.err_out = d_err + i;
    sa.status = ctx->d_flag + 2;  // truncation is synchronous: its own status slots
    sa.job = nwg > 1 ? d_job : nullptr;
    sa.nwg = nwg;
    if (sa.ws_lds < 0) return qtt_fail(ctx, QTT_EINVAL, "truncate: too many vertices");
  }
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_args, im->hargs.data(), args_bytes, hipMemcpyHostToDevice, ctx->stream), "sweep args");
  std::memset(&im->hjob, 0, sizeof(im->hjob));  // zero state, plus the exchange buffers
  if (nwg > 1) {
    double* cb = (double*)((char*)d_job + ((sizeof(CoopJob) + 255) & ~size_t(255)));
    im->hjob.Gp = cb;
    im->hjob.Pg = cb + (QTT_MAX_HELPERS + 1) * 16;
    im->hjob.Ug = im->hjob.Pg + (size_t)COOP_PG_UNITS * 256;
    im->hjob.ver = (int*)(im->hjob.Ug + 8 * 256);
  }
  if (!rc && nwg > 1)
    rc = qtt_hip_check(ctx, hipMemcpyAsync(d_job, &im->hjob, sizeof(CoopJob), hipMemcpyHostToDevice, ctx->stream), "coop job");
  if (!rc && tot_steps)
    rc = qtt_hip_check(ctx, hipMemcpyAsync(d_steps, im->allsteps.data(), tot_steps * sizeof(int), hipMemcpyHostToDevice, ctx->stream), "steps");
  if (rc) return rc;
  const size_t lds_bytes = (size_t)lds_doubles * sizeof(double);
  rc = qtt_hip_check(ctx, hipFuncSetAttribute((const void*)truncate_kernel,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes),
                     "truncate LDS attribute");
  if (rc) return rc;
  st.nwg = nwg;
  st.coop = false;
  im->d_args = d_args;
  im->d_err = d_err;
  im->d_job = d_job;
  return QTT_OK;
}

int truncate_batch_launch(qtt_ctx* ctx, TruncStaged& st) {
  if (st.nnet <= 0) return QTT_OK;
  TruncStaged::Impl* im = st.impl;
  const size_t lds_bytes = (size_t)QTT_LDS_DOUBLES * sizeof(double);
  st.coop = false;
  if (st.nwg > 1) {  // helpers: one cooperative launch, so every workgroup is resident
    void* kargs[] = {(void*)&im->d_args};
    hipError_t e = hipLaunchCooperativeKernel((const void*)truncate_kernel, dim3(st.nwg), dim3(512), kargs, lds_bytes,
                                              ctx->stream);
    if (e == hipSuccess) {
      st.coop = true;
    } else {  // not available: run without helpers
      (void)hipGetLastError();
      im->hargs[0].nwg = 1;
      im->hargs[0].job = nullptr;
      int rc = qtt_hip_check(ctx, hipMemcpyAsync(im->d_args, im->hargs.data(), sizeof(SweepArgs), hipMemcpyHostToDevice, ctx->stream), "sweep args");
      if (rc) return rc;
    }
  }
  if (!st.coop) truncate_kernel<<<st.nnet, 512, lds_bytes, ctx->stream>>>(im->d_args);
  return qtt_hip_check(ctx, hipGetLastError(), "truncate_kernel");
}

int truncate_batch_readback(qtt_ctx* ctx, TruncStaged& st, qtt_net* const* nets) {
  if (st.nnet <= 0) return QTT_OK;
  TruncStaged::Impl* im = st.impl;
  im->herr.assign(st.nnet, 0.0);
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(im->herr.data(), im->d_err, st.nnet * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "err");
  im->b32.assign(st.nnet, std::vector<int>());
  for (int i = 0; i < st.nnet && !rc; i++) {
    im->b32[i].resize(nets[i]->t.nv);
    rc = qtt_hip_check(ctx, hipMemcpyAsync(im->b32[i].data(), nets[i]->d_bond, nets[i]->t.nv * sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "bonds");
  }
  return rc;
}

int truncate_batch_finish(qtt_ctx* ctx, TruncStaged& st, qtt_net* const* nets, double* errs) {
  if (st.nnet <= 0) return QTT_OK;
  TruncStaged::Impl* im = st.impl;
  for (int i = 0; i < st.nnet; i++) {
    for (int v = 0; v < nets[i]->t.nv; v++) nets[i]->bond[v] = im->b32[i][v];
    nets[i]->pack_valid = false;
    nets[i]->pack32_valid = false;
    nets[i]->env_valid = false;
    if (errs) errs[i] = im->herr[i];
  }
  int fl[2];
  int rc = qtt_hip_check(ctx, hipMemcpy(fl, ctx->d_flag + 2, sizeof(fl), hipMemcpyDeviceToHost), "flags");
  if (rc) return rc;
  if (st.coop) {
    CoopJob hj;
    rc = qtt_hip_check(ctx, hipMemcpy(&hj, im->d_job, sizeof(hj), hipMemcpyDeviceToHost), "coop job");
    if (rc) return rc;
    if (hj.err) {
      (void)hipMemset(ctx->d_flag + 2, 0, 2 * sizeof(int));
      return qtt_fail(ctx, QTT_EHIP, "truncate: a helper workgroup wait exceeded its limit");
    }
  }
  if (fl[1]) {
    (void)hipMemset(ctx->d_flag + 2, 0, 2 * sizeof(int));
    return qtt_fail(ctx, QTT_ENOCONV, "truncate: Jacobi SVD did not converge");
  }
  return QTT_OK;
}

int launch_truncate_batch(qtt_ctx* ctx, qtt_net* const* nets, int nnet, double cutoff, int64_t maxdim,
                          double* errs, int precision) {
  if (nnet <= 0) return QTT_OK;
  double* scr = qtt_scratch(ctx, truncate_batch_bytes(ctx, nets, nnet));
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "truncate scratch");
  TruncStaged st;
  int rc = truncate_batch_stage(ctx, nets, nnet, cutoff, maxdim, precision, scr, st);
  if (!rc) rc = truncate_batch_launch(ctx, st);
  if (!rc) rc = truncate_batch_readback(ctx, st, nets);
  if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "truncate sync");
  if (rc) return rc;
  return truncate_batch_finish(ctx, st, nets, errs);
}

int launch_truncate(qtt_ctx* ctx, qtt_net* net, double cutoff, int64_t maxdim, double* err, int precision) {
  if (err) *err = 0.0;
  if (net->t.nv < 2) return QTT_OK;
  double e = 0.0;
  int rc = launch_truncate_batch(ctx, &net, 1, cutoff, maxdim, &e, precision);
  if (err) *err = e;
  return rc;
}
