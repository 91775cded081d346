// operate + truncate on device.
//
// contract_kernel: exact operator application (src/elementary_operators.jl:294-299),
//   C[(w,a)][s_out][(cw,ca)...] = sum_{s_in} W[w][s_in][s_out][cw...] A[a][s_in][ca...],
//   all vertices in one launch, one thread per output element.
//
// truncate_kernel: the truncation sweep of ITensorNetworks.truncate(::TTN)
//   (called from src/itensornetworkfunction.jl:114-118), run by one workgroup
//   per network over a host-built schedule of moves.  A move (x -> u) views
//   core x as a matrix (all other indices) x (bond to u) and
//     * QR move:  Householder QR, x <- Q, u <- R (x)_bond u;
//     * SVD move: Householder QR if the matrix is tall, then one-sided Jacobi
//       on the rows of the (square or wide) factor with the rotations kept in J,
//       singular values sorted, the ITensors cutoff rule applied, x <- U (an
//       exact product of orthogonal factors), u <- S V^T (x)_bond u.
//   The matrix lives in LDS when it fits (158 KiB), otherwise in a global
//   scratch area; all arithmetic is FP64.  Rounds of the Jacobi sweep use the
//   round-robin (circle) pairing so every pair in a round is independent: one
//   wavefront per pair, lanes over the row.
#include <cfloat>

#include "qtt_internal.h"

namespace {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
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
__global__ void __launch_bounds__(256)
contract_kernel(const double* __restrict__ W, const int64_t* __restrict__ woff, const int* __restrict__ wbond,
                const double* __restrict__ A, const int64_t* __restrict__ aoff, const int* __restrict__ abond,
                TopoView t, double* __restrict__ C, const int64_t* __restrict__ coff,
                const int64_t* __restrict__ cstart, int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
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
    double s = 0.0;
    for (int si = 0; si < d; ++si)
      s += W[woff[v] + ((pw * d + si) * d + so) * sw + jw] * A[aoff[v] + (pa * d + si) * sa + ja];
    C[coff[v] + (e - cstart[v])] = s;
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
  double* gscr;      // global workspace (used when a matrix does not fit LDS)
  int64_t gscr_ws;   // doubles available for the workspace part
  double* gY;        // global buffer for the mode-product result
  int ws_lds;        // doubles of LDS workspace
  double* err_out;
  int* status;       // [1] <- non-convergence
};

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

__device__ inline void split_dims(const int64_t* dims, int nd, int ax, int64_t& A, int64_t& n, int64_t& B) {
  A = 1;
  B = 1;
  for (int i = 0; i < ax; ++i) A *= dims[i];
  n = dims[ax];
  for (int i = ax + 1; i < nd; ++i) B *= dims[i];
}

// Householder QR in place (LAPACK dgeqr2): W is rows x cols with leading
// dimension ld; reflectors below the diagonal, R on and above it, tau[k].
__device__ void geqr2(double* W, int rows, int cols, int ld, double* tau, double* red) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = bs >> 6;
  const int k = rows < cols ? rows : cols;
  for (int j = 0; j < k; ++j) {
    double s = 0.0;
    for (int i = j + 1 + tid; i < rows; i += bs) {
      const double x = W[(int64_t)i * ld + j];
      s += x * x;
    }
    const double xn2 = block_sum(s, red);
    const double x0 = W[(int64_t)j * ld + j];
    double tj = 0.0, beta = x0, scal = 0.0;
    if (xn2 > 0.0) {
      const double nrm = sqrt(x0 * x0 + xn2);
      beta = x0 >= 0.0 ? -nrm : nrm;
      tj = (beta - x0) / beta;
      scal = 1.0 / (x0 - beta);
      for (int i = j + 1 + tid; i < rows; i += bs) W[(int64_t)i * ld + j] *= scal;
    }
    __syncthreads();
    if (tj != 0.0) {
      for (int col = j + 1 + wid; col < cols; col += nw) {
        double d = 0.0;
        for (int i = j + 1 + lane; i < rows; i += 64) d += W[(int64_t)i * ld + j] * W[(int64_t)i * ld + col];
        d = (wave_sum(d) + W[(int64_t)j * ld + col]) * tj;
        for (int i = j + 1 + lane; i < rows; i += 64) W[(int64_t)i * ld + col] -= d * W[(int64_t)i * ld + j];
        if (lane == 0) W[(int64_t)j * ld + col] -= d;
      }
    }
    if (tid == 0) {
      tau[j] = tj;
      W[(int64_t)j * ld + j] = beta;
    }
    __syncthreads();
  }
}

// Form Q (rows x k) in place from the reflectors (LAPACK dorg2r).
__device__ void org2r(double* W, int rows, int k, int ld, const double* tau) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = bs >> 6;
  for (int j = k - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj != 0.0) {
      for (int col = j + 1 + wid; col < k; col += nw) {
        double d = 0.0;
        for (int i = j + 1 + lane; i < rows; i += 64) d += W[(int64_t)i * ld + j] * W[(int64_t)i * ld + col];
        d = (wave_sum(d) + W[(int64_t)j * ld + col]) * tj;
        for (int i = j + 1 + lane; i < rows; i += 64) W[(int64_t)i * ld + col] -= d * W[(int64_t)i * ld + j];
        if (lane == 0) W[(int64_t)j * ld + col] -= d;
      }
    }
    __syncthreads();
    for (int i = tid; i < rows; i += bs) {
      double* a = W + (int64_t)i * ld + j;
      if (i < j) *a = 0.0;
      else if (i == j) *a = 1.0 - tj;
      else *a *= -tj;
    }
    __syncthreads();
  }
}

__device__ inline int rr_player(int pos, int t, int nn) {
  return pos == 0 ? 0 : 1 + (pos - 1 + t) % (nn - 1);
}

// One-sided Jacobi on the rows of X (n x b, ld b), rotations accumulated in J
// (n x n).  Returns false if it did not converge.
__device__ bool jacobi_rows(double* X, int n, int b, double* J, int* flag, double* red) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = bs >> 6;
  for (int e = tid; e < n * n; e += bs) J[e] = (e / n == e % n) ? 1.0 : 0.0;
  __syncthreads();
  if (n < 2) return true;
  const int nn = n + (n & 1);
  // a computed dot product of length b carries rounding error up to ~b eps
  const double tol = DBL_EPSILON * (double)(b > n ? b : n);
  // rows below eps * ||X||_F are numerically zero and are left alone
  double fro = 0.0;
  for (int64_t e = tid; e < (int64_t)n * b; e += bs) fro += X[e] * X[e];
  const double small = block_sum(fro, red) * DBL_EPSILON * DBL_EPSILON;
  for (int sweep = 0; sweep < 100; ++sweep) {
    if (tid == 0) *flag = 0;
    __syncthreads();
    for (int t = 0; t < nn - 1; ++t) {
      for (int pi = wid; pi < nn / 2; pi += nw) {
        int p = rr_player(pi, t, nn), q = rr_player(nn - 1 - pi, t, nn);
        if (p > q) {
          int tmp = p;
          p = q;
          q = tmp;
        }
        if (q >= n) continue;
        double* xp = X + (int64_t)p * b;
        double* xq = X + (int64_t)q * b;
        double a = 0.0, c = 0.0, g = 0.0;
        for (int i = lane; i < b; i += 64) {
          const double u = xp[i], w = xq[i];
          a += u * u;
          c += w * w;
          g += u * w;
        }
        a = wave_sum(a);
        c = wave_sum(c);
        g = wave_sum(g);
        if (g != 0.0 && a > small && c > small && fabs(g) > tol * sqrt(a) * sqrt(c)) {
          const double zeta = (c - a) / (2.0 * g);
          const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
          for (int i = lane; i < b; i += 64) {
            const double u = xp[i], w = xq[i];
            xp[i] = cs * u - sn * w;
            xq[i] = sn * u + cs * w;
          }
          double* jp = J + (int64_t)p * n;
          double* jq = J + (int64_t)q * n;
          for (int i = lane; i < n; i += 64) {
            const double u = jp[i], w = jq[i];
            jp[i] = cs * u - sn * w;
            jq[i] = sn * u + cs * w;
          }
          if (lane == 0) *flag = 1;
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

__global__ void __launch_bounds__(512) truncate_kernel(SweepArgs a) {
  extern __shared__ double lds[];
  const TopoView& t = a.t;
  const int tid = threadIdx.x, bs = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = bs >> 6;
  // LDS layout: [misc 64 doubles][bonds: nv ints][workspace]
  double* red = lds;                      // 16 doubles
  int* iflag = (int*)(lds + 16);          // flags and broadcast ints
  double* dbc = lds + 24;                 // broadcast doubles
  int* bond = (int*)(lds + 64);
  double* wsl = lds + 64 + ((t.nv + 1) >> 1);
  for (int v = tid; v < t.nv; v += bs) bond[v] = a.bond_g[v];
  if (tid == 0) dbc[0] = 0.0;
  __syncthreads();
  int64_t dx[QTT_MAX_NDIM], du[QTT_MAX_NDIM];
  for (int st = 0; st < a.nsteps; ++st) {
    const int x = a.steps[3 * st], u = a.steps[3 * st + 1], tr = a.steps[3 * st + 2];
    const int ndx = core_dims(t, bond, x, dx);
    const int ndu = core_dims(t, bond, u, du);
    const int ax = axis_of(t, x, u), au = axis_of(t, u, x);
    int64_t Ax, nx, Bx, Au, nu, Bu;
    split_dims(dx, ndx, ax, Ax, nx, Bx);
    split_dims(du, ndu, au, Au, nu, Bu);
    const int rows = (int)(Ax * Bx), b = (int)nx;
    const int k = rows < b ? rows : b;
    const int64_t need = (int64_t)rows * b + (int64_t)k * b + (int64_t)k * k + 4 * (int64_t)k + 16;
    double* ws = need <= a.ws_lds ? wsl : a.gscr;
    double* W = ws;                               // rows x b
    double* Rb = W + (int64_t)rows * b;           // k x b
    double* Jb = Rb + (int64_t)k * b;             // k x k
    double* tau = Jb + (int64_t)k * k;            // k
    double* sig = tau + k;                        // k
    int* perm = (int*)(sig + k);                  // k ints (2k doubles reserved)
    double* Xc = a.data + a.off[x];
    const double* Uc = a.data + a.off[u];
    // gather the matrix: M[a*Bx + bb][i] = X[(a*nx + i)*Bx + bb]
    for (int64_t e = tid; e < (int64_t)rows * b; e += bs) {
      const int64_t row = e / b, i = e % b;
      const int64_t aa = row / Bx, bb = row % Bx;
      W[e] = Xc[(aa * nx + i) * Bx + bb];
    }
    __syncthreads();
    int m;
    const double* Rrows;  // m rows of length b pushed into u, row i = Rrows[rowidx(i) * b]
    bool use_perm = false;
    if (!tr) {
      geqr2(W, rows, b, b, tau, red);
      for (int64_t e = tid; e < (int64_t)k * b; e += bs) {
        const int64_t i = e / b, c = e % b;
        Rb[e] = c >= i ? W[i * b + c] : 0.0;
      }
      __syncthreads();
      org2r(W, rows, k, b, tau);
      m = k;
      // x <- Q (bond m)
      for (int64_t e = tid; e < (int64_t)rows * m; e += bs) {
        const int64_t row = e / m, j = e % m;
        const int64_t aa = row / Bx, bb = row % Bx;
        Xc[(aa * m + j) * Bx + bb] = W[row * b + j];
      }
      Rrows = Rb;
    } else {
      const bool tall = rows > b;
      double* X = W;
      int n = rows;
      if (tall) {
        geqr2(W, rows, b, b, tau, red);
        for (int64_t e = tid; e < (int64_t)b * b; e += bs) {
          const int64_t i = e / b, c = e % b;
          Rb[e] = c >= i ? W[i * b + c] : 0.0;
        }
        __syncthreads();
        org2r(W, rows, b, b, tau);
        X = Rb;
        n = b;
      }
      if (!jacobi_rows(X, n, b, Jb, iflag, red)) {
        if (tid == 0) a.status[1] = 1;
      }
      // singular values = row norms; descending order with index tie-break
      for (int i = wid; i < n; i += nw) {
        double s = 0.0;
        for (int c = lane; c < b; c += 64) s += X[(int64_t)i * b + c] * X[(int64_t)i * b + c];
        s = wave_sum(s);
        if (lane == 0) sig[i] = sqrt(s);
      }
      __syncthreads();
      for (int i = tid; i < n; i += bs) {
        int r = 0;
        const double si = sig[i];
        for (int j = 0; j < n; ++j) r += (sig[j] > si || (sig[j] == si && j < i)) ? 1 : 0;
        perm[r] = i;
      }
      __syncthreads();
      if (tid == 0) {
        // ITensors/NDTensors truncate!!: relative cutoff on squared singular values
        double scale = 0.0;
        for (int i = 0; i < n; ++i) scale += sig[perm[i]] * sig[perm[i]];
        int mm = n;
        double terr = 0.0;
        const long long md = a.maxdim < 1 ? 1 : a.maxdim;
        while (mm > md) {
          terr += sig[perm[mm - 1]] * sig[perm[mm - 1]];
          --mm;
        }
        const double sc = scale == 0.0 ? 1.0 : scale;
        if (a.cutoff >= 0.0) {
          while (mm > 1) {
            const double pm = sig[perm[mm - 1]] * sig[perm[mm - 1]];
            if (terr + pm <= a.cutoff * sc) {
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
      // x <- U (bond m): U[row][i] = sum_j Q[row][j] J[perm[i]][j]  (or J[perm[i]][row])
      for (int64_t e = tid; e < (int64_t)rows * m; e += bs) {
        const int64_t row = e / m, i = e % m;
        const double* jr = Jb + (int64_t)perm[i] * n;
        double s;
        if (tall) {
          s = 0.0;
          for (int j = 0; j < n; ++j) s += W[row * b + j] * jr[j];
        } else {
          s = jr[row];
        }
        const int64_t aa = row / Bx, bb = row % Bx;
        Xc[(aa * m + i) * Bx + bb] = s;
      }
      Rrows = X;
      use_perm = true;
    }
    // u <- R (x)_{au} u : Y[(aa*m + i)*Bu + bb] = sum_c R[i][c] U[(aa*nu + c)*Bu + bb]
    const int64_t ytot = Au * m * Bu;
    for (int64_t e = tid; e < ytot; e += bs) {
      const int64_t bb = e % Bu;
      const int64_t rest = e / Bu;
      const int64_t i = rest % m, aa = rest / m;
      const double* rr = Rrows + (int64_t)(use_perm ? perm[i] : i) * b;
      double s = 0.0;
      for (int64_t c = 0; c < nu; ++c) s += rr[c] * Uc[(aa * nu + c) * Bu + bb];
      a.gY[e] = s;
    }
    __syncthreads();
    double* Ud = a.data + a.off[u];
    for (int64_t e = tid; e < ytot; e += bs) Ud[e] = a.gY[e];
    if (tid == 0) {
      if (t.parent[x] == u) bond[x] = m; else bond[u] = m;
    }
    __syncthreads();
  }
  for (int v = tid; v < t.nv; v += bs) a.bond_g[v] = bond[v];
  if (tid == 0) a.err_out[0] = dbc[0];
}

}  // namespace

int launch_contract(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, qtt_net* out) {
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
  contract_kernel<<<dim3((unsigned)std::max<int64_t>(nb, 1)), 256, 0, ctx->stream>>>(
      op->d_data, op->d_off, op->d_bond, net->d_data, net->d_off, net->d_bond, net->view(), out->d_data,
      out->d_off, d_cstart, total);
  rc = qtt_hip_check(ctx, hipGetLastError(), "contract_kernel");
  if (rc) return rc;
  return qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "contract sync");
}

int launch_truncate(qtt_ctx* ctx, qtt_net* net, double cutoff, int64_t maxdim, double* err) {
  const HostTopo& t = net->t;
  if (err) *err = 0.0;
  if (t.nv < 2) return QTT_OK;
  std::vector<int> steps = truncation_schedule(t);
  const int nsteps = (int)steps.size() / 3;
  int64_t maxcap = 1, maxb = 1;
  for (int v = 0; v < t.nv; v++) {
    maxcap = std::max(maxcap, net->cap[v]);
    maxb = std::max(maxb, net->bond[v]);
  }
  // workspace bound: rows*b <= core size; k <= b
  const int64_t ws = maxcap + 2 * maxb * maxb + 4 * maxb + 64;
  const size_t bytes = (ws + maxcap + 4) * sizeof(double) + steps.size() * sizeof(int) + 64;
  double* scr = qtt_scratch(ctx, bytes);
  if (!scr) return qtt_fail(ctx, QTT_ENOMEM, "truncate scratch");
  double* gws = scr;
  double* gY = gws + ws;
  double* d_err = gY + maxcap;
  int* d_steps = (int*)(d_err + 4);
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d_steps, steps.data(), steps.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream), "steps");
  if (rc) return rc;
  SweepArgs sa;
  sa.data = net->d_data;
  sa.off = net->d_off;
  sa.bond_g = net->d_bond;
  sa.t = net->view();
  sa.steps = d_steps;
  sa.nsteps = nsteps;
  sa.cutoff = cutoff;
  sa.maxdim = (long long)maxdim;
  sa.gscr = gws;
  sa.gscr_ws = ws;
  sa.gY = gY;
  sa.err_out = d_err;
  sa.status = ctx->d_flag;
  const int lds_doubles = QTT_LDS_DOUBLES;
  sa.ws_lds = lds_doubles - 64 - (t.nv + 1) / 2;
  if (sa.ws_lds < 0) return qtt_fail(ctx, QTT_EINVAL, "truncate: too many vertices");
  const size_t lds_bytes = (size_t)lds_doubles * sizeof(double);
  rc = qtt_hip_check(ctx, hipFuncSetAttribute((const void*)truncate_kernel,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes),
                     "truncate LDS attribute");
  if (rc) return rc;
  truncate_kernel<<<1, 512, lds_bytes, ctx->stream>>>(sa);
  rc = qtt_hip_check(ctx, hipGetLastError(), "truncate_kernel");
  if (rc) return rc;
  std::vector<int> b32(t.nv);
  double herr = 0.0;
  rc = qtt_hip_check(ctx, hipMemcpyAsync(b32.data(), net->d_bond, t.nv * sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "bonds");
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(&herr, d_err, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "err");
  if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "truncate sync");
  if (rc) return rc;
  for (int v = 0; v < t.nv; v++) net->bond[v] = b32[v];
  net->pack_valid = false;
  net->pack32_valid = false;
  if (err) *err = herr;
  int fl[2];
  rc = qtt_hip_check(ctx, hipMemcpy(fl, ctx->d_flag, sizeof(fl), hipMemcpyDeviceToHost), "flags");
  if (rc) return rc;
  if (fl[1]) {
    (void)hipMemset(ctx->d_flag, 0, 2 * sizeof(int));
    return qtt_fail(ctx, QTT_ENOCONV, "truncate: Jacobi SVD did not converge");
  }
  return QTT_OK;
}
