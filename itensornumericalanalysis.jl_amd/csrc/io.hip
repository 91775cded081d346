// Network files: the on-disk interchange format of function and operator
// networks (host code; the cores go straight between the file and HBM).
//
// Layout, little-endian, no padding:
//   char     magic[4] = "QTTN"
//   uint32   version (1)
//   uint32   kind (0: function network, 1: operator network)
//   uint32   nv, ncoord, flags (bit 0: vertex names follow the bonds)
//   int32    parent[nv], sdim[nv], vcoord[nv], vdigit[nv], ndigits[ncoord]
//   int64    bond[nv]
//   names    (flag bit 0) per vertex: uint32 length, that many bytes (UTF-8)
//   float64  cores in vertex order, each row-major as in the C-ABI:
//            [bond][site][children...] or [bond][s_in][s_out][children...]
// The topology is the C-ABI's (parent[v] < v); the names are opaque here and
// let a binding map the vertices back to its graph (the ITensorNetworkFunction
// of src/itensornetworkfunction.jl:18-57 keys its tensors by vertex name).
#include "qtt_internal.h"

#include <cstdio>
#include <cstring>
#include <memory>

namespace {

constexpr uint32_t kVersion = 1;

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

bool put(FILE* f, const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
bool get(FILE* f, void* p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }

int write_file(qtt_ctx* ctx, const char* path, uint32_t kind, const HostTopo& t, const std::vector<int64_t>& bond,
               const double* d_data, const std::vector<int64_t>& off, const std::vector<int64_t>& size,
               const char* const* names) {
  if (!path) return qtt_fail(ctx, QTT_EINVAL, "null path");
  std::string tmp = std::string(path) + ".tmp";
  int rc = qtt_hip_check(ctx, hipStreamSynchronize(ctx->stream), "save sync");
  if (rc) return rc;
  {
    File f(std::fopen(tmp.c_str(), "wb"));
    if (!f) return qtt_fail(ctx, QTT_EINVAL, std::string("cannot open ") + tmp + " for writing");
    const uint32_t hdr[6] = {kVersion, kind, (uint32_t)t.nv, (uint32_t)t.ncoord, names ? 1u : 0u, 0u};
    bool ok = put(f.get(), "QTTN", 4) && put(f.get(), hdr, 5 * sizeof(uint32_t)) &&
              put(f.get(), t.parent.data(), t.nv * sizeof(int32_t)) &&
              put(f.get(), t.sdim.data(), t.nv * sizeof(int32_t)) &&
              put(f.get(), t.vcoord.data(), t.nv * sizeof(int32_t)) &&
              put(f.get(), t.vdigit.data(), t.nv * sizeof(int32_t)) &&
              put(f.get(), t.ndigits.data(), t.ncoord * sizeof(int32_t)) &&
              put(f.get(), bond.data(), t.nv * sizeof(int64_t));
    for (int v = 0; ok && names && v < t.nv; v++) {
      const char* nm = names[v] ? names[v] : "";
      const uint32_t len = (uint32_t)std::strlen(nm);
      ok = put(f.get(), &len, sizeof(len)) && put(f.get(), nm, len);
    }
    std::vector<double> buf;
    for (int v = 0; ok && v < t.nv; v++) {
      buf.resize(size[v]);
      rc = qtt_hip_check(ctx, hipMemcpy(buf.data(), d_data + off[v], size[v] * sizeof(double), hipMemcpyDeviceToHost),
                         "download core");
      if (rc) break;
      ok = put(f.get(), buf.data(), size[v] * sizeof(double));
    }
    if (!rc && !ok) rc = qtt_fail(ctx, QTT_EINVAL, std::string("write failed: ") + tmp);
    if (!rc && std::fflush(f.get()) != 0) rc = qtt_fail(ctx, QTT_EINVAL, std::string("write failed: ") + tmp);
  }
  if (!rc && std::rename(tmp.c_str(), path) != 0) rc = qtt_fail(ctx, QTT_EINVAL, std::string("cannot rename to ") + path);
  if (rc) std::remove(tmp.c_str());
  return rc;
}

// Reads the header into an ABI topology (arrays kept in `keep`) and the bonds;
// leaves the file positioned at the cores.
struct Header {
  uint32_t kind = 0, nv = 0, ncoord = 0, flags = 0;
  std::vector<int32_t> parent, sdim, vcoord, vdigit, ndigits;
  std::vector<int64_t> bond;
};

int read_header(qtt_ctx* ctx, FILE* f, const char* path, Header& h) {
  char magic[4];
  uint32_t hdr[5];
  if (!get(f, magic, 4) || std::memcmp(magic, "QTTN", 4) != 0 || !get(f, hdr, sizeof(hdr)))
    return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": not a network file");
  if (hdr[0] != kVersion) return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": unknown format version");
  h.kind = hdr[1];
  h.nv = hdr[2];
  h.ncoord = hdr[3];
  h.flags = hdr[4];
  if (h.kind > 1 || h.nv < 1 || h.nv > (1u << 24) || h.ncoord > 4096)
    return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": bad header");
  h.parent.resize(h.nv);
  h.sdim.resize(h.nv);
  h.vcoord.resize(h.nv);
  h.vdigit.resize(h.nv);
  h.ndigits.resize(std::max<uint32_t>(h.ncoord, 1));
  h.bond.resize(h.nv);
  bool ok = get(f, h.parent.data(), h.nv * 4) && get(f, h.sdim.data(), h.nv * 4) && get(f, h.vcoord.data(), h.nv * 4) &&
            get(f, h.vdigit.data(), h.nv * 4) && get(f, h.ndigits.data(), h.ncoord * 4) &&
            get(f, h.bond.data(), h.nv * 8);
  for (uint32_t v = 0; ok && (h.flags & 1) && v < h.nv; v++) {  // names: skipped here
    uint32_t len = 0;
    ok = get(f, &len, 4) && len < (1u << 20) && std::fseek(f, (long)len, SEEK_CUR) == 0;
  }
  if (!ok) return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": truncated header");
  return QTT_OK;
}

qtt_topology abi_topo(const Header& h) {
  return qtt_topology{(int32_t)h.nv, h.parent.data(), h.sdim.data(), h.vcoord.data(), h.vdigit.data(),
                      (int32_t)h.ncoord, h.ndigits.data()};
}

// Reads every core into one host buffer, in order, checking the file's size.
int read_cores(qtt_ctx* ctx, FILE* f, const char* path, const std::vector<int64_t>& size, std::vector<double>& all,
               std::vector<const double*>& ptr) {
  int64_t tot = 0;
  for (int64_t s : size) tot += s;
  all.resize(std::max<int64_t>(tot, 1));
  if (!get(f, all.data(), tot * sizeof(double))) return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": truncated cores");
  char extra;
  if (std::fread(&extra, 1, 1, f) != 0) return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": trailing bytes");
  ptr.resize(size.size());
  int64_t o = 0;
  for (size_t v = 0; v < size.size(); v++) {
    ptr[v] = all.data() + o;
    o += size[v];
  }
  return QTT_OK;
}

std::vector<int64_t> core_sizes(const HostTopo& t, const std::vector<int64_t>& bond, bool op) {
  std::vector<int64_t> sz(t.nv);
  for (int v = 0; v < t.nv; v++) {
    int64_t s = bond[v] * t.sdim[v] * (op ? t.sdim[v] : 1);
    for (int k = 0; k < t.nch[v]; k++) s *= bond[t.chs[t.choff[v] + k]];
    sz[v] = s;
  }
  return sz;
}

}  // namespace

extern "C" {

int qtt_net_save(qtt_ctx* ctx, const qtt_net* net, const char* path, const char* const* names) {
  if (!ctx || !net) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  std::vector<int64_t> size(net->t.nv);
  for (int v = 0; v < net->t.nv; v++) size[v] = net->core_size(v);
  return write_file(ctx, path, 0, net->t, net->bond, net->d_data, net->off, size, names);
}

int qtt_mpo_save(qtt_ctx* ctx, const qtt_mpo* op, const char* path, const char* const* names) {
  if (!ctx || !op) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  std::vector<int64_t> size(op->t.nv);
  for (int v = 0; v < op->t.nv; v++) size[v] = mpo_core_size(op, v);
  return write_file(ctx, path, 1, op->t, op->bond, op->d_data, op->off, size, names);
}

int qtt_net_load(qtt_ctx* ctx, const char* path, qtt_net** out) {
  if (!ctx || !path || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  File f(std::fopen(path, "rb"));
  if (!f) return qtt_fail(ctx, QTT_EINVAL, std::string("cannot open ") + path);
  Header h;
  int rc = read_header(ctx, f.get(), path, h);
  if (rc) return rc;
  if (h.kind != 0) return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": an operator network, not a function");
  const qtt_topology tp = abi_topo(h);
  HostTopo t;
  if ((rc = topo_from_abi(ctx, &tp, t))) return rc;
  for (uint32_t v = 0; v < h.nv; v++)
    if (h.bond[v] < 1 || h.bond[v] > (1 << 20)) return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": bad bond");
  std::vector<double> all;
  std::vector<const double*> ptr;
  if ((rc = read_cores(ctx, f.get(), path, core_sizes(t, h.bond, false), all, ptr))) return rc;
  return qtt_net_create(ctx, &tp, h.bond.data(), ptr.data(), out);
}

int qtt_mpo_load(qtt_ctx* ctx, const char* path, qtt_mpo** out) {
  if (!ctx || !path || !out) return qtt_fail(ctx, QTT_EINVAL, "null argument");
  *out = nullptr;
  File f(std::fopen(path, "rb"));
  if (!f) return qtt_fail(ctx, QTT_EINVAL, std::string("cannot open ") + path);
  Header h;
  int rc = read_header(ctx, f.get(), path, h);
  if (rc) return rc;
  if (h.kind != 1) return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": a function network, not an operator");
  const qtt_topology tp = abi_topo(h);
  HostTopo t;
  if ((rc = topo_from_abi(ctx, &tp, t))) return rc;
  for (uint32_t v = 0; v < h.nv; v++)
    if (h.bond[v] < 1 || h.bond[v] > (1 << 20)) return qtt_fail(ctx, QTT_EINVAL, std::string(path) + ": bad bond");
  std::vector<double> all;
  std::vector<const double*> ptr;
  if ((rc = read_cores(ctx, f.get(), path, core_sizes(t, h.bond, true), all, ptr))) return rc;
  return qtt_mpo_create(ctx, &tp, h.bond.data(), ptr.data(), out);
}

}  // extern "C"
