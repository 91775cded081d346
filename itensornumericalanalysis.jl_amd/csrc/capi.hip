// C-ABI of libqtt: contexts, device-resident networks and operators, and the
// entry points declared in include/qtt.h.  Host-side C++; the arithmetic runs
// in the kernels of eval.hip, sweep.hip, factor.hip and grad.hip.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>

#include "qtt_internal.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
int qtt_fail(qtt_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int qtt_hip_check(qtt_ctx* ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return QTT_OK;
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    return qtt_fail(ctx, QTT_ENOMEM, std::string(what) + ": out of device memory");
  }
  return qtt_fail(ctx, QTT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, expr)                                   \
  do {                                                      \
    int _rc = qtt_hip_check((ctx), (expr), #expr);          \
    if (_rc) return _rc;                                    \
  } while (0)

double* qtt_scratch(qtt_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->scratch_bytes) return ctx->d_scratch;
  if (ctx->d_scratch) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->d_scratch);
    ctx->d_scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  size_t sz = std::max<size_t>(bytes, size_t(1) << 20);
  if (hipMalloc(&ctx->d_scratch, sz) != hipSuccess) {
    (void)hipGetLastError();
    ctx->d_scratch = nullptr;
    return nullptr;
  }
  ctx->scratch_bytes = sz;
  return ctx->d_scratch;
}

// ---------------------------------------------------------------------------
// topology
// ---------------------------------------------------------------------------
static int build_topo(qtt_ctx* ctx, const qtt_topology* tp, HostTopo& t) {
  if (!tp || tp->nv < 1 || !tp->parent || !tp->sdim || !tp->vcoord || !tp->vdigit)
    return qtt_fail(ctx, QTT_EINVAL, "topology: missing arrays or nv < 1");
  if (tp->ncoord < 0 || (tp->ncoord > 0 && !tp->ndigits))
    return qtt_fail(ctx, QTT_EINVAL, "topology: bad ncoord");
  int nv = tp->nv;
  t.nv = nv;
  t.ncoord = tp->ncoord;
  t.parent.assign(tp->parent, tp->parent + nv);
  t.sdim.assign(tp->sdim, tp->sdim + nv);
  t.vcoord.assign(tp->vcoord, tp->vcoord + nv);
  t.vdigit.assign(tp->vdigit, tp->vdigit + nv);
  t.ndigits.assign(tp->ndigits, tp->ndigits + tp->ncoord);
  if (t.parent[0] != -1) return qtt_fail(ctx, QTT_EINVAL, "topology: parent[0] must be -1");
  for (int v = 1; v < nv; v++)
    if (t.parent[v] < 0 || t.parent[v] >= v)
      return qtt_fail(ctx, QTT_EINVAL, "topology: parent[v] must satisfy 0 <= parent[v] < v");
  for (int c = 0; c < t.ncoord; c++)
    if (t.ndigits[c] < 0 || t.ndigits[c] > 62)  // 0: a coordinate the function does not depend on
      return qtt_fail(ctx, QTT_EINVAL, "topology: digits per coordinate must be in 0..62");
  std::vector<std::vector<char>> seen(t.ncoord);
  for (int c = 0; c < t.ncoord; c++) seen[c].assign(t.ndigits[c] + 1, 0);
  for (int v = 0; v < nv; v++) {
    if (t.sdim[v] < 1) return qtt_fail(ctx, QTT_EINVAL, "topology: sdim < 1");
    if (t.vcoord[v] >= 0) {
      if (t.vcoord[v] >= t.ncoord || t.vdigit[v] < 1 || t.vdigit[v] > t.ndigits[t.vcoord[v]])
        return qtt_fail(ctx, QTT_EINVAL, "topology: digit out of range");
      if (t.sdim[v] != 2) return qtt_fail(ctx, QTT_EINVAL, "topology: digit sites must be binary");
      char& s = seen[t.vcoord[v]][t.vdigit[v]];
      if (s) return qtt_fail(ctx, QTT_EINVAL, "topology: a digit sits on two vertices");
      s = 1;
    }
  }
  t.nch.assign(nv, 0);
  for (int v = 1; v < nv; v++) t.nch[t.parent[v]]++;
  t.choff.assign(nv, 0);
  int o = 0;
  for (int v = 0; v < nv; v++) {
    t.choff[v] = o;
    o += t.nch[v];
    if (2 + t.nch[v] > QTT_MAX_NDIM) return qtt_fail(ctx, QTT_EINVAL, "topology: vertex degree too large");
  }
  t.chs.assign(std::max(o, 1), 0);
  std::vector<int> fill(nv, 0);
  for (int v = 1; v < nv; v++) {
    int p = t.parent[v];
    t.chs[t.choff[p] + fill[p]++] = v;
  }
  t.chain = true;
  for (int v = 1; v < nv; v++)
    if (t.parent[v] != v - 1) t.chain = false;
  return QTT_OK;
}

static bool same_topo(const HostTopo& a, const HostTopo& b) {
  return a.nv == b.nv && a.parent == b.parent && a.sdim == b.sdim && a.vcoord == b.vcoord &&
         a.vdigit == b.vdigit;
}

// the packed int arrays of a topology, as topo_view reads them
static std::vector<int> topo_ints(const HostTopo& t) {
  int nv = t.nv;
  size_t n = size_t(7) * nv + std::max(t.ncoord, 1);
  std::vector<int> h(n, 0);
  std::copy(t.parent.begin(), t.parent.end(), h.begin() + 0 * nv);
  std::copy(t.sdim.begin(), t.sdim.end(), h.begin() + 1 * nv);
  std::copy(t.vcoord.begin(), t.vcoord.end(), h.begin() + 2 * nv);
  std::copy(t.vdigit.begin(), t.vdigit.end(), h.begin() + 3 * nv);
  std::copy(t.nch.begin(), t.nch.end(), h.begin() + 4 * nv);
  std::copy(t.choff.begin(), t.choff.end(), h.begin() + 5 * nv);
  std::copy(t.chs.begin(), t.chs.begin() + std::max(0, nv - 1), h.begin() + 6 * nv);
  std::copy(t.ndigits.begin(), t.ndigits.end(), h.begin() + 7 * nv);
  return h;
}

static int upload_topo(qtt_ctx* ctx, const HostTopo& t, int** d_topo) {
  int nv = t.nv;
  size_t n = size_t(7) * nv + std::max(t.ncoord, 1);
  std::vector<int> h(n, 0);
  std::copy(t.parent.begin(), t.parent.end(), h.begin() + 0 * nv);
  std::copy(t.sdim.begin(), t.sdim.end(), h.begin() + 1 * nv);
  std::copy(t.vcoord.begin(), t.vcoord.end(), h.begin() + 2 * nv);
  std::copy(t.vdigit.begin(), t.vdigit.end(), h.begin() + 3 * nv);
  std::copy(t.nch.begin(), t.nch.end(), h.begin() + 4 * nv);
  std::copy(t.choff.begin(), t.choff.end(), h.begin() + 5 * nv);
  std::copy(t.chs.begin(), t.chs.begin() + std::max(0, nv - 1), h.begin() + 6 * nv);
  std::copy(t.ndigits.begin(), t.ndigits.end(), h.begin() + 7 * nv);
  HIPCHK(ctx, hipMalloc(d_topo, n * sizeof(int)));
  HIPCHK(ctx, hipMemcpy(*d_topo, h.data(), n * sizeof(int), hipMemcpyHostToDevice));
  return QTT_OK;
}

TopoView topo_view(const int* d, const HostTopo& t);
static TopoView make_view(const int* d, const HostTopo& t) { return topo_view(d, t); }

TopoView topo_view(const int* d, const HostTopo& t) {
  TopoView v;
  int nv = t.nv;
  v.parent = d;
  v.sdim = d + nv;
  v.vcoord = d + 2 * nv;
  v.vdigit = d + 3 * nv;
  v.nch = d + 4 * nv;
  v.choff = d + 5 * nv;
  v.chs = d + 6 * nv;
  v.ndigits = d + 7 * nv;
  v.nv = nv;
  v.ncoord = t.ncoord;
  return v;
}

TopoView qtt_net::view() const { return make_view(d_topo, t); }

int64_t qtt_net::core_size(int v) const {
  int64_t s = bond[v] * t.sdim[v];
  for (int k = 0; k < t.nch[v]; k++) s *= bond[t.chs[t.choff[v] + k]];
  return s;
}

// ---------------------------------------------------------------------------
// truncation schedule: post-order edges from vertex 0; orthogonalise toward
// the first edge source from no gauge (farthest vertices first), then for each
// edge (c -> parent) move the centre to c along the tree path and SVD the edge.
// ---------------------------------------------------------------------------
std::vector<int> truncation_schedule(const HostTopo& t) {
  std::vector<int> steps;
  int nv = t.nv;
  if (nv < 2) return steps;
  std::vector<int> src;
  std::vector<int> stack{0}, it(nv, 0);
  while (!stack.empty()) {
    int v = stack.back();
    if (it[v] < t.nch[v]) {
      stack.push_back(t.chs[t.choff[v] + it[v]++]);
    } else {
      stack.pop_back();
      if (v != 0) src.push_back(v);
    }
  }
  int target = src[0];
  std::vector<int> order{target}, via(nv, -1), seen(nv, 0);
  seen[target] = 1;
  for (size_t h = 0; h < order.size(); h++) {
    int v = order[h];
    std::vector<int> nb;
    if (t.parent[v] >= 0) nb.push_back(t.parent[v]);
    for (int k = 0; k < t.nch[v]; k++) nb.push_back(t.chs[t.choff[v] + k]);
    for (int u : nb)
      if (!seen[u]) {
        seen[u] = 1;
        via[u] = v;
        order.push_back(u);
      }
  }
  for (int i = nv - 1; i >= 1; i--) {
    steps.push_back(order[i]);
    steps.push_back(via[order[i]]);
    steps.push_back(0);
  }
  int center = target;
  for (int c : src) {
    while (center != c) {
      int x = c;
      bool below = false;
      while (x != -1) {
        if (t.parent[x] == center) {
          below = true;
          break;
        }
        x = t.parent[x];
      }
      int step = below ? x : t.parent[center];
      steps.push_back(center);
      steps.push_back(step);
      steps.push_back(0);
      center = step;
    }
    steps.push_back(c);
    steps.push_back(t.parent[c]);
    steps.push_back(1);
    center = t.parent[c];
  }
  return steps;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int qtt_version(void) { return 200; }

int qtt_ctx_create(int device, qtt_ctx** out) {
  if (!out) return QTT_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return QTT_EHIP;
  }
  qtt_ctx* c = new (std::nothrow) qtt_ctx;
  if (!c) return QTT_ENOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->own_stream) != hipSuccess ||
      hipMalloc(&c->d_flag, 4 * sizeof(int)) != hipSuccess ||
      hipMemset(c->d_flag, 0, 4 * sizeof(int)) != hipSuccess) {
    (void)hipGetLastError();
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return QTT_EHIP;
  }
  if (qtt_noalloc_init(c) != QTT_OK) {  // the device address of the fixed scratch, with the device current
    (void)hipFree(c->d_flag);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return QTT_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
  c->stream = c->own_stream;
  const char* ek = getenv("QTT_EVAL_KERNEL");
  c->force_lds_eval = ek && std::strcmp(ek, "lds") == 0;
  const char* ppl = std::getenv("QTT_EVAL_PPL");  // 1 or 2 (default: chosen by bond size)
  c->eval_ppl = ppl ? std::atoi(ppl) : 0;
  // QTT_EVAL_TABLE_BITS: digit sites per environment table of the chain evaluate
  const char* tbits = std::getenv("QTT_EVAL_TABLE_BITS");
  if (tbits) c->eval_tbits = std::max(0, std::min(24, std::atoi(tbits)));
  // QTT_TRUNC_HELPERS: helper workgroups of a single-network truncation whose
  // cores exceed LDS (default 31, at most QTT_MAX_HELPERS and one per other
  // CU; 0: workgroup 0 alone)
  const char* th = std::getenv("QTT_TRUNC_HELPERS");
  if (th) c->trunc_helpers = std::min(QTT_MAX_HELPERS, std::max(0, std::atoi(th)));
  // QTT_JACOBI_MAX_SWEEPS: sweeps before an SVD reports QTT_ENOCONV (default 100)
  const char* js = std::getenv("QTT_JACOBI_MAX_SWEEPS");
  if (js) c->jacobi_max_sweeps = std::min(1000, std::max(1, std::atoi(js)));
  // QTT_BATCH_GRAPH=0: qtt_batch_run queues its stages directly instead of as one graph
  const char* bg = std::getenv("QTT_BATCH_GRAPH");
  c->batch_graph = !(bg && std::strcmp(bg, "0") == 0);
  *out = c;
  return QTT_OK;
}

static void ctx_free(qtt_ctx* ctx) {
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_scratch) (void)hipFree(ctx->d_scratch);
  if (ctx->d_sync) (void)hipFree(ctx->d_sync);
  if (ctx->d_flag) (void)hipFree(ctx->d_flag);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
}

void qtt_ctx_destroy(qtt_ctx* ctx) {
  if (!ctx) return;
  ctx->closed = true;
  if (ctx->refs == 0) ctx_free(ctx);  // else the last network or operator destroyed frees it
}

// a network or operator released its hold on the context
static void ctx_release(qtt_ctx* ctx) {
  if (!ctx) return;
  if (--ctx->refs == 0 && ctx->closed) ctx_free(ctx);
}

const char* qtt_last_error(const qtt_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int qtt_ctx_set_stream(qtt_ctx* ctx, void* s) {
  if (!ctx) return QTT_EINVAL;
  ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
  return QTT_OK;
}

// Wait for the stream, then report and clear the status words of slot pair
// `base` (0: asynchronous calls, 2: the synchronous call in progress).
static int collect_status(qtt_ctx* ctx, int base) {
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  int h[2];
  HIPCHK(ctx, hipMemcpy(h, ctx->d_flag + base, sizeof(h), hipMemcpyDeviceToHost));
  if (h[0] || h[1]) {
    HIPCHK(ctx, hipMemset(ctx->d_flag + base, 0, sizeof(h)));
    if (h[1]) return qtt_fail(ctx, QTT_ENOCONV, "SVD did not converge");
    return qtt_fail(ctx, QTT_ERANGE, "a coordinate was negative or NaN");
  }
  return QTT_OK;
}

int qtt_sync(qtt_ctx* ctx) {
  if (!ctx) return QTT_EINVAL;
  return collect_status(ctx, 0);
}

namespace {
// Kernels launched while this guard lives write the synchronous status slots.
struct SyncSlots {
  qtt_ctx* c;
  int old;
  explicit SyncSlots(qtt_ctx* ctx) : c(ctx), old(ctx->flag_base) { c->flag_base = 2; }
  ~SyncSlots() { c->flag_base = old; }
};
}  // namespace

static int alloc_net(qtt_ctx* ctx, const HostTopo& t, const std::vector<int64_t>& bond,
                     const std::vector<int64_t>& cap, qtt_net** out) {
  qtt_net* n = new (std::nothrow) qtt_net;
  if (!n) return qtt_fail(ctx, QTT_ENOMEM, "host allocation");
  n->ctx = ctx;
  ctx->refs++;
  n->t = t;
  n->bond = bond;
  n->cap = cap;
  n->off.resize(t.nv);
  int64_t o = 0;
  for (int v = 0; v < t.nv; v++) {
    n->off[v] = o;
    o += (cap[v] + 1) & ~int64_t(1);  // 16-byte aligned cores
  }
  n->total = o;
  int rc = QTT_OK;
  std::vector<int> b32(t.nv);
  for (int v = 0; v < t.nv; v++) b32[v] = (int)bond[v];
  if ((rc = qtt_hip_check(ctx, hipMalloc(&n->d_data, std::max<int64_t>(o, 2) * sizeof(double)), "hipMalloc cores")) ||
      (rc = qtt_hip_check(ctx, hipMalloc(&n->d_off, t.nv * sizeof(int64_t)), "hipMalloc offsets")) ||
      (rc = qtt_hip_check(ctx, hipMalloc(&n->d_bond, t.nv * sizeof(int)), "hipMalloc bonds")) ||
      (rc = upload_topo(ctx, t, &n->d_topo)) ||
      (rc = qtt_hip_check(ctx, hipMemcpy(n->d_off, n->off.data(), t.nv * sizeof(int64_t), hipMemcpyHostToDevice), "offsets")) ||
      (rc = qtt_hip_check(ctx, hipMemcpy(n->d_bond, b32.data(), t.nv * sizeof(int), hipMemcpyHostToDevice), "bonds"))) {
    qtt_net_destroy(n);
    return rc;
  }
  *out = n;
  return QTT_OK;
}

extern "C++" {
int topo_from_abi(qtt_ctx* ctx, const qtt_topology* tp, HostTopo& t) { return build_topo(ctx, tp, t); }

int alloc_net_bonds(qtt_ctx* ctx, const HostTopo& t, const std::vector<int64_t>& bond, qtt_net** out) {
  std::vector<int64_t> cap(t.nv);
  for (int v = 0; v < t.nv; v++) {
    int64_t s = bond[v] * t.sdim[v];
    for (int k = 0; k < t.nch[v]; k++) s *= bond[t.chs[t.choff[v] + k]];
    cap[v] = s;
  }
  return alloc_net(ctx, t, bond, cap, out);
}
}  // extern "C++"

int qtt_net_create(qtt_ctx* ctx, const qtt_topology* topo, const int64_t* bond,
                   const double* const* cores, qtt_net** out) {
  if (!ctx || !out || !bond || !cores) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  HostTopo t;
  int rc = build_topo(ctx, topo, t);
  if (rc) return rc;
  if (bond[0] != 1) return qtt_fail(ctx, QTT_EINVAL, "bond[0] must be 1");
  std::vector<int64_t> b(bond, bond + t.nv), cap(t.nv);
  for (int v = 0; v < t.nv; v++) {
    if (b[v] < 1 || b[v] > (1 << 20)) return qtt_fail(ctx, QTT_EINVAL, "bond out of range");
    if (!cores[v]) return qtt_fail(ctx, QTT_EINVAL, "null core");
  }
  for (int v = 0; v < t.nv; v++) {
    int64_t s = b[v] * t.sdim[v];
    for (int k = 0; k < t.nch[v]; k++) s *= b[t.chs[t.choff[v] + k]];
    cap[v] = s;
  }
  qtt_net* n = nullptr;
  if ((rc = alloc_net(ctx, t, b, cap, &n))) return rc;
  for (int v = 0; v < t.nv; v++) {
    rc = qtt_hip_check(ctx, hipMemcpy(n->d_data + n->off[v], cores[v], cap[v] * sizeof(double), hipMemcpyHostToDevice), "upload core");
    if (rc) {
      qtt_net_destroy(n);
      return rc;
    }
  }
  *out = n;
  return QTT_OK;
}

int qtt_net_bonds(qtt_ctx* ctx, const qtt_net* net, int64_t* bond_out) {
  if (!net || !bond_out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  std::copy(net->bond.begin(), net->bond.end(), bond_out);
  return QTT_OK;
}

int64_t qtt_net_size(const qtt_net* net) {
  if (!net) return 0;
  int64_t s = 0;
  for (int v = 0; v < net->t.nv; v++) s += net->core_size(v);
  return s;
}

int64_t qtt_mpo_size(const qtt_mpo* op) {
  if (!op) return 0;
  int64_t s = 0;
  for (int v = 0; v < op->t.nv; v++) {
    int64_t c = op->bond[v] * op->t.sdim[v] * op->t.sdim[v];
    for (int k = 0; k < op->t.nch[v]; k++) c *= op->bond[op->t.chs[op->t.choff[v] + k]];
    s += c;
  }
  return s;
}

int qtt_net_read(qtt_ctx* ctx, const qtt_net* net, double* const* cores_out) {
  if (!ctx || !net || !cores_out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int v = 0; v < net->t.nv; v++)
    HIPCHK(ctx, hipMemcpy(cores_out[v], net->d_data + net->off[v], net->core_size(v) * sizeof(double),
                          hipMemcpyDeviceToHost));
  return QTT_OK;
}

int qtt_net_copy(qtt_ctx* ctx, const qtt_net* net, qtt_net** out) {
  if (!ctx || !net || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  qtt_net* n = nullptr;
  int rc = alloc_net(ctx, net->t, net->bond, net->cap, &n);
  if (rc) return rc;
  rc = qtt_hip_check(ctx, hipMemcpyAsync(n->d_data, net->d_data, net->total * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "copy cores");
  if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  if (rc) {
    qtt_net_destroy(n);
    return rc;
  }
  *out = n;
  return QTT_OK;
}

void qtt_net_destroy(qtt_net* n) {
  if (!n) return;
  if (n->ctx) (void)hipStreamSynchronize(n->ctx->stream);
  if (!n->pool) {  // pooled networks release their share of the batch allocation below
    if (n->d_data) (void)hipFree(n->d_data);
    if (n->d_off) (void)hipFree(n->d_off);
    if (n->d_topo) (void)hipFree(n->d_topo);
    if (n->d_bond) (void)hipFree(n->d_bond);
  }
  if (n->d_pack) (void)hipFree(n->d_pack);
  if (n->d_pack_off) (void)hipFree(n->d_pack_off);
  if (n->d_pack32) (void)hipFree(n->d_pack32);
  if (n->d_desc) (void)hipFree(n->d_desc);
  if (n->d_env) (void)hipFree(n->d_env);
  qtt_ctx* c = n->ctx;
  delete n;
  ctx_release(c);
}

extern "C++" {
// An operator network with these bonds, cores allocated but not written.
int alloc_mpo(qtt_ctx* ctx, const HostTopo& t, const std::vector<int64_t>& bond, qtt_mpo** out) {
  *out = nullptr;
  if (bond.empty() || bond[0] != 1) return qtt_fail(ctx, QTT_EINVAL, "bond[0] must be 1");
  qtt_mpo* m = new (std::nothrow) qtt_mpo;
  if (!m) return qtt_fail(ctx, QTT_ENOMEM, "host allocation");
  m->t = t;
  m->ctx = ctx;
  ctx->refs++;
  const int nv = m->t.nv;
  m->bond = bond;
  m->off.resize(nv);
  int64_t o = 0;
  for (int v = 0; v < nv; v++) {
    if (m->bond[v] < 1 || m->bond[v] > (1 << 20)) {
      qtt_mpo_destroy(m);
      return qtt_fail(ctx, QTT_EINVAL, "operator bond out of range");
    }
    m->off[v] = o;
    o += (mpo_core_size(m, v) + 1) & ~int64_t(1);
  }
  m->total = o;
  std::vector<int> b32(nv);
  for (int v = 0; v < nv; v++) b32[v] = (int)m->bond[v];
  int rc;
  if ((rc = qtt_hip_check(ctx, hipMalloc(&m->d_data, std::max<int64_t>(o, 2) * sizeof(double)), "hipMalloc operator")) ||
      (rc = qtt_hip_check(ctx, hipMalloc(&m->d_off, nv * sizeof(int64_t)), "hipMalloc")) ||
      (rc = qtt_hip_check(ctx, hipMalloc(&m->d_bond, nv * sizeof(int)), "hipMalloc")) ||
      (rc = qtt_hip_check(ctx, hipMemcpy(m->d_off, m->off.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice), "upload")) ||
      (rc = qtt_hip_check(ctx, hipMemcpy(m->d_bond, b32.data(), nv * sizeof(int), hipMemcpyHostToDevice), "upload"))) {
    qtt_mpo_destroy(m);
    return rc;
  }
  *out = m;
  return QTT_OK;
}

int64_t mpo_core_size(const qtt_mpo* m, int v) {
  int64_t s = m->bond[v] * m->t.sdim[v] * m->t.sdim[v];
  for (int k = 0; k < m->t.nch[v]; k++) s *= m->bond[m->t.chs[m->t.choff[v] + k]];
  return s;
}
}  // extern "C++"

int qtt_mpo_create(qtt_ctx* ctx, const qtt_topology* topo, const int64_t* bond,
                   const double* const* cores, qtt_mpo** out) {
  if (!ctx || !out || !bond || !cores) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  HostTopo t;
  int rc = build_topo(ctx, topo, t);
  if (rc) return rc;
  for (int v = 0; v < t.nv; v++)
    if (!cores[v]) return qtt_fail(ctx, QTT_EINVAL, "null operator core");
  qtt_mpo* m = nullptr;
  if ((rc = alloc_mpo(ctx, t, std::vector<int64_t>(bond, bond + t.nv), &m))) return rc;
  for (int v = 0; v < t.nv; v++) {
    rc = qtt_hip_check(ctx, hipMemcpy(m->d_data + m->off[v], cores[v], mpo_core_size(m, v) * sizeof(double), hipMemcpyHostToDevice), "upload operator core");
    if (rc) {
      qtt_mpo_destroy(m);
      return rc;
    }
  }
  *out = m;
  return QTT_OK;
}

int qtt_mpo_bonds(qtt_ctx* ctx, const qtt_mpo* m, int64_t* bond_out) {
  if (!ctx || !m || !bond_out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  std::copy(m->bond.begin(), m->bond.end(), bond_out);
  return QTT_OK;
}

int qtt_mpo_read(qtt_ctx* ctx, const qtt_mpo* m, double* const* cores_out) {
  if (!ctx || !m || !cores_out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int v = 0; v < m->t.nv; v++)
    HIPCHK(ctx, hipMemcpy(cores_out[v], m->d_data + m->off[v], mpo_core_size(m, v) * sizeof(double), hipMemcpyDeviceToHost));
  return QTT_OK;
}

void qtt_mpo_destroy(qtt_mpo* m) {
  if (!m) return;
  if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
  if (m->d_data) (void)hipFree(m->d_data);
  if (m->d_off) (void)hipFree(m->d_off);
  if (m->d_bond) (void)hipFree(m->d_bond);
  qtt_ctx* c = m->ctx;
  delete m;
  ctx_release(c);
}

static int check_precision(qtt_ctx* ctx, int precision) {
  if (precision != QTT_F64 && precision != QTT_F32_SPLIT && precision != QTT_F32)
    return qtt_fail(ctx, QTT_EINVAL, "unknown precision");
  return QTT_OK;
}

// operator application: the bulk products (exact contraction, mode products)
// run in the precision asked for; QR, SVD and the cutoff stay float64
static int check_apply_precision(qtt_ctx* ctx, int precision) { return check_precision(ctx, precision); }

int qtt_eval(qtt_ctx* ctx, const qtt_net* net, const double* xs, int64_t npts, double* out,
             int precision) {
  if (!ctx || !net || (!xs && npts) || (!out && npts) || npts < 0)
    return qtt_fail(ctx, QTT_EINVAL, "null argument");
  int rc = check_precision(ctx, precision);
  if (rc) return rc;
  if (npts == 0) return QTT_OK;
  return launch_eval(ctx, const_cast<qtt_net*>(net), xs, npts, out, precision);
}

int qtt_eval_noalloc(qtt_ctx* ctx, const qtt_net* net, const double* xs, int64_t npts, double* out) {
  if (!ctx || !net || (!xs && npts) || (!out && npts) || npts < 0)
    return qtt_fail(ctx, QTT_EINVAL, "null argument");
  return launch_eval_noalloc(ctx, net, xs, npts, out);
}

int qtt_eval_host(qtt_ctx* ctx, const qtt_net* net, const double* xs, int64_t npts, double* out,
                  int precision) {
  if (!ctx || !net || (!xs && npts) || (!out && npts) || npts < 0)
    return qtt_fail(ctx, QTT_EINVAL, "null argument");
  if (npts == 0) return QTT_OK;
  int nc = std::max(net->t.ncoord, 1);
  double* d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, size_t(npts) * (nc + 1) * sizeof(double)));
  SyncSlots slots(ctx);
  int rc = qtt_hip_check(ctx, hipMemcpyAsync(d, xs, size_t(npts) * nc * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "upload points");
  if (!rc) rc = qtt_eval(ctx, net, d, npts, d + size_t(npts) * nc, precision);
  if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(out, d + size_t(npts) * nc, npts * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "download values");
  int rs = collect_status(ctx, 2);
  (void)hipFree(d);
  return rc ? rc : rs;
}

int qtt_integrate(qtt_ctx* ctx, const qtt_net* net, double weight, double* out) {
  if (!ctx || !net || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  return launch_integrate(ctx, net, weight, out);
}

int qtt_inner(qtt_ctx* ctx, const qtt_net* a, const qtt_mpo* op, const qtt_net* b, double* out) {
  if (!ctx || !a || !b || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  if (!same_topo(a->t, b->t)) return qtt_fail(ctx, QTT_EINVAL, "inner: networks on different layouts");
  if (!op) return launch_inner(ctx, a, b, out);
  qtt_net* wb = nullptr;
  int rc = qtt_apply(ctx, op, b, -1.0, 0, QTT_F64, &wb, nullptr);
  if (rc) return rc;
  rc = launch_inner(ctx, a, wb, out);
  qtt_net_destroy(wb);
  return rc;
}

int qtt_apply(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, double cutoff, int64_t maxdim,
              int precision, qtt_net** out, double* trunc_err) {
  if (!ctx || !op || !net || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  int rc = check_apply_precision(ctx, precision);
  if (rc) return rc;
  if (!same_topo(op->t, net->t)) return qtt_fail(ctx, QTT_EINVAL, "apply: operator and network layouts differ");
  const HostTopo& t = net->t;
  std::vector<int64_t> b(t.nv), cap(t.nv);
  for (int v = 0; v < t.nv; v++) b[v] = op->bond[v] * net->bond[v];
  for (int v = 0; v < t.nv; v++) {
    int64_t s = b[v] * t.sdim[v];
    for (int k = 0; k < t.nch[v]; k++) s *= b[t.chs[t.choff[v] + k]];
    cap[v] = s;
  }
  qtt_net* n = nullptr;
  if ((rc = alloc_net(ctx, t, b, cap, &n))) return rc;
  if ((rc = launch_contract(ctx, op, net, n, precision))) {
    qtt_net_destroy(n);
    return rc;
  }
  if (trunc_err) *trunc_err = 0.0;
  if (cutoff >= 0.0) {
    rc = launch_truncate(ctx, n, cutoff, maxdim > 0 ? maxdim : (int64_t(1) << 40), trunc_err, precision);
    if (rc) {
      qtt_net_destroy(n);
      return rc;
    }
  }
  *out = n;
  return QTT_OK;
}

// Output networks of an apply batch (bond = operator bond x network bond): all
// cores in one allocation and offsets + topology + bonds in another, whose
// contents are uploaded asynchronously from `meta` (the caller keeps it until
// the stream has passed the copy).  On failure the networks made so far are
// destroyed and `made` is cleared.
static int alloc_batch_outputs(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets,
                               std::vector<qtt_net*>& made, std::vector<char>& meta) {
  made.assign(n, nullptr);
  auto cleanup = [&]() {
    for (auto* m : made) qtt_net_destroy(m);
    made.assign(n, nullptr);
  };
  if (n <= 0) return QTT_OK;
  int rc = QTT_OK;
  std::vector<int64_t> dbase(n), mbase(n);
  int64_t dtot = 0, mtot = 0;  // doubles of cores; bytes of metadata
  for (int i = 0; i < n && !rc; i++) {
    const HostTopo& t = nets[i]->t;
    qtt_net* m = new (std::nothrow) qtt_net;
    if (!m) {
      rc = qtt_fail(ctx, QTT_ENOMEM, "host allocation");
      break;
    }
    made[i] = m;
    m->ctx = ctx;
    ctx->refs++;
    m->t = t;
    m->bond.resize(t.nv);
    m->cap.resize(t.nv);
    m->off.resize(t.nv);
    for (int v = 0; v < t.nv; v++) m->bond[v] = ops[i]->bond[v] * nets[i]->bond[v];
    int64_t o = 0;
    for (int v = 0; v < t.nv; v++) {
      int64_t sz = m->bond[v] * t.sdim[v];
      for (int k = 0; k < t.nch[v]; k++) sz *= m->bond[t.chs[t.choff[v] + k]];
      m->cap[v] = sz;
      m->off[v] = o;
      o += (sz + 1) & ~int64_t(1);  // 16-byte aligned cores
    }
    m->total = o;
    dbase[i] = dtot;
    dtot += std::max<int64_t>(o, 2);
    mbase[i] = mtot;
    const size_t nti = size_t(7) * t.nv + std::max(t.ncoord, 1);
    mtot += (int64_t)((t.nv * sizeof(int64_t) + nti * sizeof(int) + t.nv * sizeof(int) + 15) & ~size_t(15));
  }
  double* dpool = nullptr;
  char* mpool = nullptr;
  if (!rc) rc = qtt_hip_check(ctx, hipMalloc(&dpool, dtot * sizeof(double)), "hipMalloc batch cores");
  if (!rc) rc = qtt_hip_check(ctx, hipMalloc(&mpool, std::max<int64_t>(mtot, 16)), "hipMalloc batch metadata");
  if (rc) {
    if (dpool) (void)hipFree(dpool);
    cleanup();
    return rc;
  }
  std::shared_ptr<void> pool(dpool, [mpool](void* p) {
    (void)hipFree(p);
    (void)hipFree(mpool);
  });
  meta.assign(mtot, 0);
  for (int i = 0; i < n; i++) {
    qtt_net* m = made[i];
    const HostTopo& t = m->t;
    char* h = meta.data() + mbase[i];
    char* d = mpool + mbase[i];
    const std::vector<int> ti = topo_ints(t);
    std::vector<int> b32(t.nv);
    for (int v = 0; v < t.nv; v++) b32[v] = (int)m->bond[v];
    std::memcpy(h, m->off.data(), t.nv * sizeof(int64_t));
    std::memcpy(h + t.nv * sizeof(int64_t), ti.data(), ti.size() * sizeof(int));
    std::memcpy(h + t.nv * sizeof(int64_t) + ti.size() * sizeof(int), b32.data(), t.nv * sizeof(int));
    m->pool = pool;
    m->d_data = dpool + dbase[i];
    m->d_off = (int64_t*)d;
    m->d_topo = (int*)(d + t.nv * sizeof(int64_t));
    m->d_bond = (int*)(d + t.nv * sizeof(int64_t) + ti.size() * sizeof(int));
  }
  rc = qtt_hip_check(ctx, hipMemcpyAsync(mpool, meta.data(), mtot, hipMemcpyHostToDevice, ctx->stream), "batch metadata");
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    cleanup();
  }
  return rc;
}

static int check_apply_batch(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets,
                             qtt_net** outs) {
  for (int i = 0; i < n; i++) {
    outs[i] = nullptr;
    if (!ops[i] || !nets[i]) return qtt_fail(ctx, QTT_EINVAL, "null operator or network in batch");
    if (!same_topo(ops[i]->t, nets[i]->t))
      return qtt_fail(ctx, QTT_EINVAL, "apply batch: operator and network layouts differ");
  }
  return QTT_OK;
}

int qtt_apply_batch(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets, double cutoff,
                    int64_t maxdim, int precision, qtt_net** outs, double* errs) {
  if (!ctx || n < 0 || (n && (!ops || !nets || !outs))) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  int rc = check_apply_precision(ctx, precision);
  if (!rc) rc = check_apply_batch(ctx, n, ops, nets, outs);
  if (rc) return rc;
  std::vector<qtt_net*> made;
  std::vector<char> meta;
  // All outputs in two allocations with one upload of the metadata, and the
  // exact contractions in one launch.
  rc = alloc_batch_outputs(ctx, n, ops, nets, made, meta);
  if (rc) return rc;
  auto cleanup = [&]() {
    for (auto* m : made) qtt_net_destroy(m);
  };
  if (n > 0) {
    rc = launch_contract_batch(ctx, n, ops, nets, made.data(), precision);
    if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "batch contract sync");  // meta outlives the copy
  }
  if (!rc && cutoff >= 0.0)
    rc = launch_truncate_batch(ctx, made.data(), n, cutoff, maxdim > 0 ? maxdim : (int64_t(1) << 40), errs,
                               precision);
  else if (!rc && errs)
    for (int i = 0; i < n; i++) errs[i] = 0.0;
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    cleanup();
    return rc;
  }
  for (int i = 0; i < n; i++) outs[i] = made[i];
  return QTT_OK;
}

// The whole batch pipeline of a kernel-operator solve step: contract ->
// truncate -> integrate -> evaluate, every stage one batched launch, queued
// back to back with no host synchronisation and, where the stream allows,
// captured into one HIP graph.  Bonds, truncation errors, integrals and the
// evaluation classes are copied back once, after the graph.
int qtt_batch_run(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets, double cutoff,
                  int64_t maxdim, int precision, double weight, const double* const* xs, const int64_t* npts,
                  double* const* vals, qtt_net** outs, double* errs, double* integrals) {
  if (!ctx || n < 0 || (n && (!ops || !nets || !outs))) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  int rc = check_apply_precision(ctx, precision);
  if (!rc) rc = check_apply_batch(ctx, n, ops, nets, outs);
  if (rc) return rc;
  const bool want_eval = xs && npts && vals;
  if (want_eval)
    for (int i = 0; i < n; i++)
      if (npts[i] < 0 || (npts[i] && (!xs[i] || !vals[i]))) return qtt_fail(ctx, QTT_EINVAL, "bad batch entry");
  if (n == 0) return QTT_OK;
  SyncSlots slots(ctx);
  std::vector<qtt_net*> made;
  std::vector<char> meta;
  rc = alloc_batch_outputs(ctx, n, ops, nets, made, meta);
  if (rc) return rc;
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto* m : made) qtt_net_destroy(m);
    return code;
  };
  const bool trunc = cutoff >= 0.0;
  const int64_t md = maxdim > 0 ? maxdim : (int64_t(1) << 40);
  // one scratch allocation for every stage's tables and workspaces
  std::vector<int64_t> zero_npts;
  const int64_t* np = npts;
  if (!want_eval) {
    zero_npts.assign(n, 0);
    np = zero_npts.data();
  }
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t cb = al(contract_batch_bytes(n, made.data()));
  const size_t tb = trunc ? al(truncate_batch_bytes(ctx, made.data(), n)) : 0;
  const size_t ib = integrals ? al(integrate_batch_bytes(n, made.data())) : 0;
  const size_t eb = al(eval_pipe_bytes(n, made.data(), np));
  char* scr = (char*)qtt_scratch(ctx, cb + tb + ib + eb);
  if (!scr) return fail(qtt_fail(ctx, QTT_ENOMEM, "pipeline scratch"));
  BatchStage cst, ist, est;
  TruncStaged tst;
  rc = contract_batch_stage(ctx, n, ops, nets, made.data(), precision, scr, cst);
  if (!rc && trunc) rc = truncate_batch_stage(ctx, made.data(), n, cutoff, md, precision, (double*)(scr + cb), tst);
  if (!rc && integrals) rc = integrate_batch_stage(ctx, n, made.data(), weight, scr + cb + tb, ist);
  if (!rc) rc = eval_pipe_stage(ctx, n, made.data(), want_eval ? xs : nullptr, np, want_eval ? vals : nullptr,
                                QTT_F64, scr + cb + tb + ib, est);
  if (rc) return fail(rc);
  auto queue = [&]() {
    int r = contract_batch_launch(ctx, cst);
    if (!r && trunc) r = truncate_batch_launch(ctx, tst);
    if (!r && integrals) r = integrate_batch_launch(ctx, ist);
    if (!r) r = eval_pipe_launch(ctx, est);
    return r;
  };
  // Graph capture needs a stream other than the null stream and kernels that
  // capture (a truncation with helper workgroups is a cooperative launch).
  const hipStream_t exec = ctx->stream;
  const hipStream_t cap = exec ? exec : ctx->own_stream;
  const bool graph = ctx->batch_graph && cap && tst.nwg <= 1;
  if (graph) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    rc = qtt_hip_check(ctx, hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture");
    if (!rc) {
      ctx->stream = cap;
      const int rq = queue();
      ctx->stream = exec;
      const int re = qtt_hip_check(ctx, hipStreamEndCapture(cap, &g), "end capture");
      rc = rq ? rq : re;
    }
    if (!rc) rc = qtt_hip_check(ctx, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "graph instantiate");
    if (!rc) rc = qtt_hip_check(ctx, hipGraphLaunch(ge, exec), "graph launch");
    // a graph may be destroyed once launched; its work still runs
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
  } else {
    rc = queue();
  }
  if (!rc && trunc) rc = truncate_batch_readback(ctx, tst, made.data());
  if (!rc && integrals)
    rc = qtt_hip_check(ctx, hipMemcpyAsync(integrals, ist.d_out, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "integrals");
  if (!rc) rc = eval_pipe_readback(ctx, est);
  if (!rc) rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "pipeline sync");
  if (!rc && trunc) rc = truncate_batch_finish(ctx, tst, made.data(), errs);
  if (!rc && !trunc && errs)
    for (int i = 0; i < n; i++) errs[i] = 0.0;
  if (!rc && want_eval) rc = eval_pipe_rest(ctx, est, n, made.data(), xs, npts, vals, QTT_F64);
  {
    const int rs = collect_status(ctx, 2);
    if (!rc) rc = rs;
  }
  if (rc) return fail(rc);
  for (int i = 0; i < n; i++) outs[i] = made[i];
  return QTT_OK;
}

int qtt_integrate_batch(qtt_ctx* ctx, int n, const qtt_net* const* nets, double weight, double* out) {
  if (!ctx || n < 0 || (n && (!nets || !out))) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  for (int i = 0; i < n; i++)
    if (!nets[i]) return qtt_fail(ctx, QTT_EINVAL, "null network in batch");
  return launch_integrate_batch(ctx, n, nets, weight, out);
}

int qtt_eval_batch(qtt_ctx* ctx, int n, const qtt_net* const* nets, const double* const* xs,
                   const int64_t* npts, double* const* out, int precision) {
  if (!ctx || n < 0 || (n && (!nets || !xs || !npts || !out))) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  int rc = check_precision(ctx, precision);
  if (rc) return rc;
  for (int i = 0; i < n; i++) {
    if (!nets[i] || npts[i] < 0 || (npts[i] && (!xs[i] || !out[i])))
      return qtt_fail(ctx, QTT_EINVAL, "bad batch entry");
  }
  return launch_eval_batch(ctx, n, const_cast<qtt_net* const*>(nets), xs, npts, out, precision);
}

int qtt_tune_eval(qtt_ctx* ctx, qtt_net* net, const double* xs, int64_t npts, double* out, int* cfg,
                  double* ms) {
  if (!ctx || !net || !cfg || !ms || npts < 0 || (npts && (!xs || !out)))
    return qtt_fail(ctx, QTT_EINVAL, "null argument");
  return tune_eval(ctx, net, xs, npts, out, cfg, ms);
}

int qtt_truncate(qtt_ctx* ctx, qtt_net* net, double cutoff, int64_t maxdim, double* trunc_err) {
  if (!ctx || !net) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  if (maxdim < 0) return qtt_fail(ctx, QTT_EINVAL, "maxdim < 0");
  return launch_truncate(ctx, net, cutoff < 0 ? 0.0 : cutoff, maxdim > 0 ? maxdim : (int64_t(1) << 40), trunc_err,
                         QTT_F64);
}

int qtt_eval_grad(qtt_ctx* ctx, const qtt_net* net, const double* xs, const double* ybar,
                  int64_t npts, double* grad, int precision) {
  if (!ctx || !net || !grad || npts < 0 || (npts && (!xs || !ybar)))
    return qtt_fail(ctx, QTT_EINVAL, "null argument");
  int rc = check_precision(ctx, precision);
  if (rc) return rc;
  return launch_grad(ctx, net, xs, ybar, npts, grad, precision);
}

int qtt_apply_grad(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, const double* gbar, double* fbar,
                   double* wbar) {
  if (!ctx || !op || !net || !gbar || !fbar) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  if (!same_topo(op->t, net->t)) return qtt_fail(ctx, QTT_EINVAL, "apply gradient: operator and network layouts differ");
  return launch_apply_grad(ctx, op, net, gbar, fbar, wbar);
}

int qtt_eval_grad_host(qtt_ctx* ctx, const qtt_net* net, const double* xs, const double* ybar,
                       int64_t npts, double* const* grads, int precision) {
  if (!ctx || !net || !grads || npts < 0 || (npts && (!xs || !ybar)))
    return qtt_fail(ctx, QTT_EINVAL, "null argument");
  int nc = std::max(net->t.ncoord, 1);
  int64_t gsz = qtt_net_size(net);
  double* d = nullptr;
  size_t bytes = (size_t(npts) * (nc + 1) + gsz) * sizeof(double);
  HIPCHK(ctx, hipMalloc(&d, std::max<size_t>(bytes, 16)));
  SyncSlots slots(ctx);
  int rc = QTT_OK;
  double* dx = d;
  double* dy = d + size_t(npts) * nc;
  double* dg = dy + npts;
  if (npts) {
    rc = qtt_hip_check(ctx, hipMemcpyAsync(dx, xs, size_t(npts) * nc * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "upload");
    if (!rc) rc = qtt_hip_check(ctx, hipMemcpyAsync(dy, ybar, size_t(npts) * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "upload");
  }
  if (!rc) rc = qtt_eval_grad(ctx, net, dx, dy, npts, dg, precision);
  {
    const int rs = collect_status(ctx, 2);
    if (!rc) rc = rs;
  }
  int64_t o = 0;
  for (int v = 0; v < net->t.nv && !rc; v++) {
    int64_t s = net->core_size(v);
    rc = qtt_hip_check(ctx, hipMemcpy(grads[v], dg + o, s * sizeof(double), hipMemcpyDeviceToHost), "download");
    o += s;
  }
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
