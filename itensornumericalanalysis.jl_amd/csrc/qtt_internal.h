// Internal declarations shared by the libqtt translation units (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qtt.h"

#define QTT_MAX_COORD 4     // coordinates handled by the fused digit prologue
#define QTT_MAX_NDIM 10     // core rank (parent + site + up to 8 children)
#define QTT_LDS_DOUBLES 20224  // 158 KiB of the 160 KiB LDS per CU

// most helper workgroups a single-network truncation can take (the size of
// their exchange buffers)
constexpr int QTT_MAX_HELPERS = 255;

struct qtt_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  // Status words written by kernels: slots [0] (invalid coordinate) and [1]
  // (SVD non-convergence) belong to asynchronous calls and are collected by
  // qtt_sync; slots [2] and [3] hold the same for synchronous entry points,
  // which read and clear them before returning, so a pending asynchronous
  // status is never reported by an unrelated synchronous call.
  int* d_flag = nullptr;
  int flag_base = 0;  // 0 (asynchronous) or 2 (inside a synchronous entry point)
  int* flags() const { return d_flag + flag_base; }
  // Networks and operators keep their context alive: qtt_ctx_destroy only
  // marks it closed while handles remain, and the last handle destroyed frees it.
  int refs = 0;
  bool closed = false;
  double* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  int* d_sync = nullptr;  // counters of the dataflow kernels, zero between launches
  int sync_cap = 0;
  int cus = 256;
  bool force_lds_eval = false;  // QTT_EVAL_KERNEL=lds selects the LDS-vector chain kernel
  int eval_ppl = 0;             // points per lane of the register chain kernel (0: by bond size)
  int jacobi_max_sweeps = 100;  // Jacobi sweeps before an SVD reports QTT_ENOCONV
  int trunc_helpers = 31;       // helper workgroups of a single-network truncation (QTT_TRUNC_HELPERS)
  bool batch_graph = true;      // qtt_batch_run captures its stages into one HIP graph (QTT_BATCH_GRAPH)
  // launch shapes of the register chain kernel picked by qtt_tune_eval, keyed
  // by eval_tune_key (bond class, point-count octave); absent keys use the
  // built-in rule
  struct EvalShape {
    int ppl, wg, tbits;  // tbits: digit sites per environment table (0: none)
  };
  int eval_tbits = -1;  // QTT_EVAL_TABLE_BITS overrides the table size (-1: tuned or default)
  std::map<uint32_t, EvalShape> eval_tuned;
};

// Topology view passed by value to kernels (device pointers).
struct TopoView {
  const int* parent;
  const int* sdim;
  const int* vcoord;
  const int* vdigit;
  const int* nch;
  const int* choff;
  const int* chs;
  const int* ndigits;
  int nv;
  int ncoord;
};

// Per-site descriptor of a packed chain (qtt_net::d_desc): bonds left/right,
// the coordinate whose digit selects the slice (< 0: no digit) and the bit
// position of that digit in the coordinate's digit word.
struct SiteDesc {
  int chil, chir, coord, shift;
};

struct HostTopo {
  int nv = 0, ncoord = 0;
  std::vector<int> parent, sdim, vcoord, vdigit, ndigits, nch, choff, chs;
  bool chain = false;
};

struct qtt_net {
  qtt_ctx* ctx = nullptr;
  HostTopo t;
  std::vector<int64_t> bond;  // host mirror of the current bonds
  std::vector<int64_t> cap;   // capacity (elements) per vertex
  std::vector<int64_t> off;   // element offset per vertex
  int64_t total = 0;
  double* d_data = nullptr;
  int64_t* d_off = nullptr;
  int* d_topo = nullptr;  // packed int arrays, see topo_view()
  int* d_bond = nullptr;  // int32 bonds on device
  // packed copy for the chain evaluate kernel: site k as [r][lpad][s]
  double* d_pack = nullptr;
  int64_t* d_pack_off = nullptr;
  int64_t pack_elems = 0;
  int* d_desc = nullptr;  // per site {chil, chir, coord, shift} for the register kernel
  std::vector<int64_t> h_pack_off;  // host sources of the pack offsets / descriptors
  std::vector<int> h_desc;
  bool pack_valid = false;
  // float32 copies for the mixed-precision evaluate
  float* d_pack32 = nullptr;
  bool pack32_valid = false;
  // environment tables of the chain evaluate (left then right), built by the
  // first evaluate with table size env_tb and reused while the cores are
  // unchanged (cleared with pack_valid)
  double* d_env = nullptr;
  int64_t env_elems = 0;
  int env_tb = -1;
  bool env_valid = false;
  // Networks made by one apply batch share two device allocations (cores;
  // offsets, topology and bonds): d_data, d_off, d_topo and d_bond point into
  // them, and the last network destroyed frees them.  Null: own allocations.
  std::shared_ptr<void> pool;
  TopoView view() const;
  int64_t core_size(int v) const;
};

struct qtt_mpo {
  qtt_ctx* ctx = nullptr;
  HostTopo t;
  std::vector<int64_t> bond;
  std::vector<int64_t> off;
  int64_t total = 0;
  double* d_data = nullptr;
  int64_t* d_off = nullptr;
  int* d_bond = nullptr;
};

// helpers implemented in capi.hip
int qtt_fail(qtt_ctx* ctx, int code, const std::string& msg);
TopoView topo_view(const int* d_topo, const HostTopo& t);
int qtt_hip_check(qtt_ctx* ctx, hipError_t e, const char* what);
double* qtt_scratch(qtt_ctx* ctx, size_t bytes);
// a checked HostTopo from the C-ABI description; a new network with these
// bonds (cores uninitialised)
int topo_from_abi(qtt_ctx* ctx, const qtt_topology* tp, HostTopo& t);
int alloc_net_bonds(qtt_ctx* ctx, const HostTopo& t, const std::vector<int64_t>& bond, qtt_net** out);
int alloc_mpo(qtt_ctx* ctx, const HostTopo& t, const std::vector<int64_t>& bond, qtt_mpo** out);
int64_t mpo_core_size(const qtt_mpo* m, int v);

// kernel launchers (eval.hip, sweep.hip, factor.hip, grad.hip)
int launch_eval(qtt_ctx* ctx, qtt_net* net, const double* xs, int64_t npts, double* out,
                int precision);
int tune_eval(qtt_ctx* ctx, qtt_net* net, const double* xs, int64_t npts, double* out, int* cfg, double* ms);
int launch_eval_batch(qtt_ctx* ctx, int n, qtt_net* const* nets, const double* const* xs, const int64_t* npts,
                      double* const* out, int precision);
// the library's fixed evaluate scratch: its device address, once per device (eval.hip)
int qtt_noalloc_init(qtt_ctx* ctx);
int launch_eval_noalloc(qtt_ctx* ctx, const qtt_net* net, const double* xs, int64_t npts, double* out);
int launch_integrate(qtt_ctx* ctx, const qtt_net* net, double weight, double* out_host);
int launch_inner(qtt_ctx* ctx, const qtt_net* a, const qtt_net* b, double* out_host);
int launch_integrate_batch(qtt_ctx* ctx, int n, const qtt_net* const* nets, double weight, double* out_host);
// precision: the arithmetic of the bulk products (QTT_F64, QTT_F32_SPLIT, QTT_F32)
int launch_contract(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, qtt_net* out, int precision);
int launch_contract_batch(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets,
                          qtt_net* const* outs, int precision);
int launch_truncate(qtt_ctx* ctx, qtt_net* net, double cutoff, int64_t maxdim, double* err, int precision);
int launch_truncate_batch(qtt_ctx* ctx, qtt_net* const* nets, int nnet, double cutoff, int64_t maxdim,
                          double* errs, int precision);
// Staged batch launches, for pipelines that queue several batch kernels with
// no host synchronisation between them (qtt_batch_run): *_bytes sizes the
// device tables, *_stage fills them at `scr` with asynchronous copies from
// `host` (which must live until the stream passes them), *_launch queues the
// kernel only, so it can be captured into a HIP graph.
struct BatchStage {
  int n = 0;
  int precision = QTT_F64;
  double weight = 1.0;
  int64_t total = 0;
  int maxnv = 1;
  char* dev = nullptr;  // device tables
  char* aux = nullptr;  // a second table inside them
  char* seg = nullptr;  // contraction: the segment table (nseg segments)
  int64_t nseg = 0;
  double* d_out = nullptr;
  std::vector<char> host;
  std::vector<int> cls;  // evaluate: class per network read back after the run
  std::vector<int> eval_host;  // evaluate: networks left to the single-network path
};
size_t contract_batch_bytes(int n, qtt_net* const* outs);
int contract_batch_stage(qtt_ctx* ctx, int n, const qtt_mpo* const* ops, const qtt_net* const* nets,
                         qtt_net* const* outs, int precision, char* scr, BatchStage& st);
int contract_batch_launch(qtt_ctx* ctx, const BatchStage& st);
struct TruncStaged {
  struct Impl;
  Impl* impl;
  int nnet = 0, nwg = 1;
  bool coop = false;
  TruncStaged();
  ~TruncStaged();
  TruncStaged(const TruncStaged&) = delete;
  TruncStaged& operator=(const TruncStaged&) = delete;
};
size_t truncate_batch_bytes(qtt_ctx* ctx, qtt_net* const* nets, int nnet);
int truncate_batch_stage(qtt_ctx* ctx, qtt_net* const* nets, int nnet, double cutoff, int64_t maxdim,
                         int precision, double* scr, TruncStaged& st);
int truncate_batch_launch(qtt_ctx* ctx, TruncStaged& st);
// queue the copies of errors and bonds; after the stream is synchronised,
// finish updates the host bond mirrors and reports the status words
int truncate_batch_readback(qtt_ctx* ctx, TruncStaged& st, qtt_net* const* nets);
int truncate_batch_finish(qtt_ctx* ctx, TruncStaged& st, qtt_net* const* nets, double* errs);
// integrate: bonds on the host may be upper bounds of the device bonds
size_t integrate_batch_bytes(int n, const qtt_net* const* nets);
int integrate_batch_stage(qtt_ctx* ctx, int n, const qtt_net* const* nets, double weight, char* scr,
                          BatchStage& st);
int integrate_batch_launch(qtt_ctx* ctx, const BatchStage& st);
// evaluate of chains whose bonds are known only on the device: packing and
// the bond class are decided by kernels; networks whose final bonds exceed
// the register kernels' (32) are evaluated by eval_pipe_rest afterwards
size_t eval_pipe_bytes(int n, qtt_net* const* nets, const int64_t* npts);
int eval_pipe_stage(qtt_ctx* ctx, int n, qtt_net* const* nets, const double* const* xs, const int64_t* npts,
                    double* const* out, int precision, char* scr, BatchStage& st);
int eval_pipe_launch(qtt_ctx* ctx, const BatchStage& st);
int eval_pipe_readback(qtt_ctx* ctx, BatchStage& st);
int eval_pipe_rest(qtt_ctx* ctx, const BatchStage& st, int n, qtt_net* const* nets, const double* const* xs,
                   const int64_t* npts, double* const* out, int precision);
int launch_grad(qtt_ctx* ctx, const qtt_net* net, const double* xs, const double* ybar,
                int64_t npts, double* grad, int precision);
int launch_apply_grad(qtt_ctx* ctx, const qtt_mpo* op, const qtt_net* net, const double* gbar, double* fbar,
                      double* wbar);

// schedule of the truncation sweep (capi.hip): triples (x, u, truncate)
std::vector<int> truncation_schedule(const HostTopo& t);
