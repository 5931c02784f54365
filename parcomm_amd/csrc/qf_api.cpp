// C ABI of the product library (declared in include/parcomm_amd.h).
#include "../../include/parcomm_amd.h"
#include "qf_state.hpp"
#include <algorithm>
#include <cstring>
#include "qf_vec.hpp"
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <atomic>

using namespace qf;

struct qf_params : qf::ParamsState {};
struct qf_committed : qf::CommittedState {};
struct qf_proof {
  qf::Bytes bytes;
  qf::ProveTimings t;
};

namespace {
thread_local std::string g_error;
}
namespace qf {
const Layout& params_layout(const qf_params* p) { return p->layout; }
void set_last_error(const char* msg) { g_error = msg; }
}

namespace {
template <class F>
int guarded(F&& f) {
  g_error.clear();
  try {
    f();
    return QF_OK;
  } catch (const QfError& e) {
    g_error = e.what();
    return e.status;
  } catch (const std::exception& e) {
    g_error = e.what();
    return QF_ERR_ARGUMENT;
  }
}

}  // namespace

namespace qf {
// Interval list of entries that must be zero (layout.rs:136-155), split into
// chunks of at most 64 KiB so the device check spreads over many workgroups.
std::vector<uint64_t> padding_ranges(const Layout& l) {
  std::vector<uint64_t> out;
  auto add = [&](uint64_t start, uint64_t len) {
    for (uint64_t o = 0; o < len; o += 65536) { out.push_back(start + o); out.push_back(std::min<uint64_t>(65536, len - o)); }
  };
  for (auto& b : l.buckets) {
    if (b.physical_cols < b.cols)
      for (uint64_t r = 0; r < b.physical_rows; r++) add(b.offset + r * b.cols + b.physical_cols, b.cols - b.physical_cols);
    if (b.physical_rows < b.rows) add(b.offset + b.physical_rows * b.cols, (b.rows - b.physical_rows) * b.cols);
  }
  return out;
}

__global__ void k_check_zero(const Val* bva, const uint64_t* ranges, uint32_t* flag) {
  uint64_t start = ranges[2 * blockIdx.x], len = ranges[2 * blockIdx.x + 1];
  uint32_t bad = 0;
  for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) bad |= bva[start + i] != 0;
  if (bad) atomicOr(flag, 1u);
}

void launch_check_zero(const Val* bva, const uint64_t* d_ranges, size_t count, uint32_t* d_flag, hipStream_t st) {
  for (size_t first = 0; first < count; first += 65535) {
    size_t n = std::min<size_t>(65535, count - first);
    hipLaunchKernelGGL(k_check_zero, dim3(uint32_t(n)), dim3(256), 0, st, bva, d_ranges + 2 * first, d_flag);
  }
}
}  // namespace qf

namespace {

void validate_device(const ParamsState* p, const Val* d_bva, hipStream_t st) {
  {
    std::lock_guard<std::mutex> g(p->mu);
    if (p->zero_ranges.empty()) p->zero_ranges = padding_ranges(p->layout);
  }
  if (p->zero_ranges.empty()) return;
  size_t count = p->zero_ranges.size() / 2;
  uint64_t* d_ranges = static_cast<uint64_t*>(dmalloc(p->zero_ranges.size() * 8, st));
  uint32_t* d_flag = static_cast<uint32_t*>(dmalloc(4, st));
  QF_HIP(hipMemcpyAsync(d_ranges, p->zero_ranges.data(), p->zero_ranges.size() * 8, hipMemcpyHostToDevice, st));
  QF_HIP(hipMemsetAsync(d_flag, 0, 4, st));
  launch_check_zero(d_bva, d_ranges, count, d_flag, st);
  uint32_t flag = 0;
  QF_HIP(hipMemcpyAsync(&flag, d_flag, 4, hipMemcpyDeviceToHost, st));
  QF_HIP(hipStreamSynchronize(st));
  dfree(d_ranges, st);
  dfree(d_flag, st);
  require(flag == 0, QF_ERR_INVALID_PARAMETERS, "nonzero boundary padding");
}

// This thread's timing events, pinned root buffer and captured commit graphs,
// kept for the life of the thread.
struct CommitScratch {
  std::vector<hipEvent_t> ev;
  uint8_t* roots = nullptr;  // pinned, 128 x 32 bytes
  // captured block loops by everything their launches depend on: geometry,
  // BVA, cache mode and the device buffers of every block
  std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
  std::vector<std::vector<uint64_t>> graph_order;  // oldest first; at most 8 kept
  ~CommitScratch() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    for (auto e : ev) (void)hipEventDestroy(e);
    if (roots) (void)hipHostFree(roots);
  }
  hipEvent_t* events(size_t n) {
    while (ev.size() < n) {
      hipEvent_t e;
      QF_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
    return ev.data();
  }
};
thread_local CommitScratch t_commit;

// QF_COMMIT_GRAPH=0 issues every commit's launches one by one.
bool commit_graph_on() {
  const char* e = std::getenv("QF_COMMIT_GRAPH");
  return !(e && e[0] == '0');
}

// The block loop of protocol.rs:162-185 as a two-stream pipeline: the work
// stream encodes block b + 1 while the second stream hashes block b, so the
// tile passes and the leaf hash share the CUs. Roots are copied to pinned
// memory on the hashing stream and read once at the end. Recompute mode
// alternates two workspaces. From the second commit of a parameter set on
// (the first one runs the autotuner), the loop is captured once as a HIP
// graph per buffer set and replayed with one launch; the graph holds no
// memset nodes.
void commit_impl(const qf_params* p, const Val* d_bva, size_t len, int cache_mode, qf_committed** out) {
  require(len == p->layout.active, QF_ERR_INVALID_PARAMETERS, "active BVA length mismatch");
  require(cache_mode == QF_CACHE_CACHED || cache_mode == QF_CACHE_RECOMPUTE, QF_ERR_ARGUMENT, "cache mode");
  hipStream_t st = thread_stream(), st2 = thread_stream2();
  std::unique_ptr<qf_committed> c(new qf_committed);
  const bool cached = cache_mode == QF_CACHE_CACHED;
  E8* ws_cw[2] = {nullptr, nullptr};
  uint32_t* ws_tree[2] = {nullptr, nullptr};
  bool capturing = false;
  try {
    c->p = p;
    c->cache_mode = cache_mode;
    c->d_bva = d_bva;
    validate_device(p, d_bva, st);
    const uint64_t n = uint64_t(1) << p->geom.log_n, R = 2 * n;
    const size_t cw_bytes = size_t(32 * R) * 16, tree_bytes = size_t(2 * R - 1) * 32;
    const E8* alpha = p->alpha_encode(p->pp.codes.size() - 1);
    // blocks to encode; the all-zero blocks past the data share the first one's commitment
    std::vector<uint64_t> enc;
    std::vector<size_t> src(p->blocks);
    int zero_block = -1;
    for (uint64_t b = 0; b < p->blocks; b++) {
      const uint64_t first = b * p->block_len;
      if (first >= len && zero_block >= 0) { src[b] = src[zero_block]; continue; }
      src[b] = enc.size();
      enc.push_back(b);
      if (first >= len) zero_block = int(b);
    }
    if (!t_commit.roots) QF_HIP(hipHostMalloc(&t_commit.roots, 128 * 32, hipHostMallocDefault));
    // per encoded block: 0 start, 1 encoded, 2 leaf start, 3 leaf end, 4/5 around M_W, 6 tree and root done
    hipEvent_t* ev = t_commit.events(7 * enc.size() + 2);
    hipEvent_t e_start = ev[7 * enc.size()], e_end = ev[7 * enc.size() + 1];
    // every device buffer before any launch (allocations stay outside a captured graph)
    std::vector<E8*> blk_cw(enc.size());
    std::vector<uint32_t*> blk_tree(enc.size());
    for (size_t k = 0; k < enc.size(); k++) {
      if (cached) {  // owned by the handle from here on (freed with it on any failure)
        blk_cw[k] = static_cast<E8*>(dmalloc(cw_bytes, st));
        c->d_cw.push_back(blk_cw[k]);
        blk_tree[k] = static_cast<uint32_t*>(dmalloc(tree_bytes, st));
        c->d_tree.push_back(blk_tree[k]);
      } else {
        const int w = int(k % 2);
        if (!ws_cw[w]) {
          ws_cw[w] = static_cast<E8*>(dmalloc(cw_bytes, st));
          ws_tree[w] = static_cast<uint32_t*>(dmalloc(tree_bytes, st));
        }
        blk_cw[k] = ws_cw[w];
        blk_tree[k] = ws_tree[w];
      }
    }
    auto issue = [&] {
      for (size_t k = 0; k < enc.size(); k++) {
        const uint64_t b = enc[k];
        hipEvent_t* E = ev + 7 * k;
        if (!cached && k >= 2) QF_HIP(hipStreamWaitEvent(st, ev[7 * (k - 2) + 6], 0));  // workspace hashed
        QF_HIP(hipEventRecord(E[0], st));
        launch_commit_encode(p->geom, d_bva, b * p->block_len, len, blk_cw[k], alpha, st, E + 4);
        QF_HIP(hipEventRecord(E[1], st));
        QF_HIP(hipStreamWaitEvent(st2, E[1], 0));
        QF_HIP(hipEventRecord(E[2], st2));
        launch_tree(blk_cw[k], R, p->geom.leaf, p->geom.inner, blk_tree[k], st2, E[3]);
        QF_HIP(hipMemcpyAsync(t_commit.roots + 32 * k, blk_tree[k] + 8 * (2 * R - 2), 32, hipMemcpyDeviceToHost, st2));
        QF_HIP(hipEventRecord(E[6], st2));
        QF_HIP(hipGetLastError());
      }
      if (!enc.empty()) QF_HIP(hipStreamWaitEvent(st, ev[7 * (enc.size() - 1) + 6], 0));  // join
    };
    const bool graph = commit_graph_on() && p->commit_warm.load();
    c->graphed = graph;
    if (graph) {
      std::vector<uint64_t> key = {reinterpret_cast<uintptr_t>(p), reinterpret_cast<uintptr_t>(d_bva), len,
                                   uint64_t(cache_mode), reinterpret_cast<uintptr_t>(alpha), uint64_t(p->geom.log_n),
                                   p->blocks, p->block_len};
      for (size_t k = 0; k < enc.size(); k++) {
        key.push_back(reinterpret_cast<uintptr_t>(blk_cw[k]));
        key.push_back(reinterpret_cast<uintptr_t>(blk_tree[k]));
      }
      auto it = t_commit.graphs.find(key);
      if (it == t_commit.graphs.end()) {
        QF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        capturing = true;
        set_tuning_allowed(false);  // nothing may allocate or synchronise under capture
        struct Restore { ~Restore() { set_tuning_allowed(true); } } restore;
        issue();
        hipGraph_t g = nullptr;
        capturing = false;
        QF_HIP(hipStreamEndCapture(st, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        QF_HIP(ie);
        if (t_commit.graph_order.size() >= 8) {  // the oldest goes
          auto old = t_commit.graphs.find(t_commit.graph_order.front());
          if (old != t_commit.graphs.end()) { (void)hipGraphExecDestroy(old->second); t_commit.graphs.erase(old); }
          t_commit.graph_order.erase(t_commit.graph_order.begin());
        }
        it = t_commit.graphs.emplace(key, ex).first;
        t_commit.graph_order.push_back(key);
      }
      QF_HIP(hipEventRecord(e_start, st));
      QF_HIP(hipGraphLaunch(it->second, st));
      QF_HIP(hipEventRecord(e_end, st));
    } else {
      QF_HIP(hipEventRecord(e_start, st));
      issue();
      QF_HIP(hipEventRecord(e_end, st));
    }
    QF_HIP(hipEventSynchronize(e_end));
    // A replayed graph does not restamp the events recorded inside it (they
    // keep the times of the launch that captured them): such a commit reports
    // its total time only.
    for (size_t k = 0; k < enc.size() && !graph; k++) {
      hipEvent_t* E = ev + 7 * k;
      float a = 0, l = 0, m = 0;
      QF_HIP(hipEventElapsedTime(&a, E[0], E[1]));
      QF_HIP(hipEventElapsedTime(&l, E[2], E[3]));
      QF_HIP(hipEventElapsedTime(&m, E[4], E[5]));
      c->encode_ms += a;
      c->leaf_ms += l;
      c->mw_ms += m;
      c->timed_blocks++;
    }
    float ms = 0;
    QF_HIP(hipEventElapsedTime(&ms, e_start, e_end));
    c->total_ms = ms;
    if (cached) { c->d_cw.clear(); c->d_tree.clear(); }
    for (uint64_t b = 0; b < p->blocks; b++) {
      Digest root;
      std::memcpy(root.data(), t_commit.roots + 32 * src[b], 32);
      c->roots.push_back(root);
      if (cached) { c->d_cw.push_back(blk_cw[src[b]]); c->d_tree.push_back(blk_tree[src[b]]); }
    }
    for (int w = 0; w < 2; w++)
      if (ws_cw[w]) { dfree(ws_cw[w], st); dfree(ws_tree[w], st); ws_cw[w] = nullptr; ws_tree[w] = nullptr; }
    p->commit_warm = true;
  } catch (...) {
    if (capturing) {  // end the capture before the streams can be used again
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
    }
    (void)hipStreamSynchronize(st2);
    (void)hipStreamSynchronize(st);
    for (int w = 0; w < 2; w++)
      if (ws_cw[w]) { dfree(ws_cw[w], st); dfree(ws_tree[w], st); }
    throw;
  }
  *out = c.release();
}

std::vector<std::pair<size_t, size_t>> claim_dims(const ParamsState& p) {
  std::vector<std::pair<size_t, size_t>> d;
  for (auto& s : p.layout.shapes) {
    uint64_t rp = 1, cp = 1;
    while (rp < s.rows) rp <<= 1;
    while (cp < s.cols) cp <<= 1;
    d.push_back({size_t(ilog2u(rp)), size_t(ilog2u(cp))});
  }
  return d;
}
size_t claims_bytes(const ParamsState& p) {
  size_t n = 0;
  for (auto [r, c] : claim_dims(p)) n += 16 * (r + c + 1);
  return n;
}
E8 read_scalar(const uint8_t* b) {
  E8 v;
  require(e8_from_bytes(b, v), QF_ERR_ENCODING, "non-canonical field element in claims");
  return v;
}
std::vector<ClaimIn> parse_claims(const ParamsState& p, const uint8_t* q) {
  std::vector<ClaimIn> cl;
  for (auto [r, cc] : claim_dims(p)) {
    ClaimIn x;
    for (size_t i = 0; i < r; i++, q += 16) x.row.push_back(read_scalar(q));
    for (size_t i = 0; i < cc; i++, q += 16) x.col.push_back(read_scalar(q));
    x.value = read_scalar(q);
    q += 16;
    cl.push_back(std::move(x));
  }
  return cl;
}
}  // namespace

extern "C" {

const char* qf_last_error(void) { return g_error.c_str(); }
int qf_field(void) { return QF_FIELD; }

int qf_params_new(int field, const uint64_t* shapes, size_t n, uint32_t chunk_log, uint32_t packed_log,
                  uint32_t query_rf_bits, qf_params** out) {
  return guarded([&] {
    require(out && shapes, QF_ERR_ARGUMENT, "null argument");
    require(field == QF_FIELD, QF_ERR_INVALID_PARAMETERS, "this library is built for the other field backend");
    require(chunk_log >= 10 && chunk_log <= 30 && packed_log >= 1 && packed_log <= 6 && query_rf_bits >= 32 &&
                query_rf_bits <= 120,
            QF_ERR_INVALID_PARAMETERS, "invalid ParComm config");
    std::unique_ptr<qf_params> p(new qf_params);
    std::vector<Shape> s(n);
    for (size_t i = 0; i < n; i++) s[i] = {shapes[2 * i], shapes[2 * i + 1]};
    p->layout = make_layout(s);
    p->cfg_chunk_log = chunk_log;
    p->packed_log = packed_log;
    p->query_bits = query_rf_bits;
    p->chunk_log = std::min<int>(int(chunk_log), ilog2u(p->layout.logical));
    p->block_len = uint64_t(1) << p->chunk_log;
    p->blocks = p->layout.logical / p->block_len;
    require(p->blocks <= 128, QF_ERR_INVALID_PARAMETERS, "at most 128 PCS blocks are supported");
    uint8_t seed[32];
    std::memset(seed, 0x51, 32);  // protocol.rs:123
    p->pp = make_pcs_params(p->chunk_log, query_rf_bits, seed);
    p->geom.log_n = p->chunk_log - LOG_LANES - 5;  // 32 columns of packed elements
    p->geom.leaf = leaf_prefix(32);
    p->geom.inner = inner_prefix();
    p->leaf_ext = leaf_prefix(32, "qafold-short/extension-row/v1");
    *out = p.release();
  });
}

void qf_params_free(qf_params* p) { delete p; }

int qf_params_digest(const qf_params* p, uint8_t out[32]) {
  return guarded([&] { require(p && out, QF_ERR_ARGUMENT, "null argument"); std::memcpy(out, p->pp.parameter_digest.data(), 32); });
}

int qf_params_info(const qf_params* p, uint64_t* active, uint64_t* logical, uint64_t* block_len, uint64_t* blocks,
                   uint64_t* n_buckets) {
  return guarded([&] {
    require(p, QF_ERR_ARGUMENT, "null argument");
    if (active) *active = p->layout.active;
    if (logical) *logical = p->layout.logical;
    if (block_len) *block_len = p->block_len;
    if (blocks) *blocks = p->blocks;
    if (n_buckets) *n_buckets = p->layout.buckets.size();
  });
}

int qf_params_bucket(const qf_params* p, uint64_t i, uint64_t out[8]) {
  return guarded([&] {
    require(p && out && i < p->layout.buckets.size(), QF_ERR_ARGUMENT, "bucket index");
    const Bucket& b = p->layout.buckets[i];
    uint64_t v[8] = {b.matrix, b.row_offset, b.col_offset, b.rows, b.cols, b.physical_rows, b.physical_cols, b.offset};
    std::memcpy(out, v, sizeof v);
  });
}

int qf_commit(const qf_params* p, const void* values, size_t len, int cache_mode, qf_committed** out) {
  return guarded([&] {
    require(p && out && (values || len == 0), QF_ERR_ARGUMENT, "null argument");
    device_setup();
    Val* d = nullptr;
    hipError_t e = hipMalloc(&d, std::max<size_t>(len * sizeof(Val), 16));
    if (e == hipErrorOutOfMemory) {  // give the pool's idle blocks back and retry once
      (void)hipGetLastError();
      release_cached_blocks();
      e = hipMalloc(&d, std::max<size_t>(len * sizeof(Val), 16));
    }
    QF_HIP(e);
    try {
      QF_HIP(hipMemcpy(d, values, len * sizeof(Val), hipMemcpyHostToDevice));
      commit_impl(p, d, len, cache_mode, out);
    } catch (...) {
      (void)hipFree(d);
      throw;
    }
    (*out)->owns_bva = true;  // the opening reads the BVA again
  });
}

int qf_commit_device(const qf_params* p, const void* device_values, size_t len, int cache_mode, qf_committed** out) {
  return guarded([&] {
    require(p && out && device_values, QF_ERR_ARGUMENT, "null argument");
    require(reinterpret_cast<uintptr_t>(device_values) % 16 == 0, QF_ERR_ARGUMENT, "device BVA must be 16-byte aligned");
    commit_impl(p, static_cast<const Val*>(device_values), len, cache_mode, out);
  });
}

void qf_committed_free(qf_committed* c) { delete c; }

int qf_committed_roots(const qf_committed* c, uint8_t* out, size_t cap_blocks) {
  return guarded([&] {
    require(c && out && cap_blocks >= c->roots.size(), QF_ERR_ARGUMENT, "root buffer too small");
    for (size_t i = 0; i < c->roots.size(); i++) std::memcpy(out + 32 * i, c->roots[i].data(), 32);
  });
}

int qf_committed_codeword(const qf_committed* c, uint64_t block, void* out, size_t cap) {
  return guarded([&] {
    require(c && out && c->cache_mode == QF_CACHE_CACHED && block < c->d_cw.size(), QF_ERR_ARGUMENT, "codeword request");
    const size_t R = 2 * (size_t(1) << c->p->geom.log_n), bytes = 32 * R * 16;
    require(cap >= bytes, QF_ERR_ARGUMENT, "codeword buffer too small");
    QF_HIP(hipMemcpy(out, c->d_cw[block], bytes, hipMemcpyDeviceToHost));
  });
}

int qf_committed_timing(const qf_committed* c, double* total_ms, double* leaf_ms, double* encode_ms, double* mw_ms,
                        uint64_t* timed_blocks) {
  return guarded([&] {
    require(c, QF_ERR_ARGUMENT, "null argument");
    if (total_ms) *total_ms = c->total_ms;
    if (leaf_ms) *leaf_ms = c->leaf_ms;
    if (encode_ms) *encode_ms = c->encode_ms;
    if (mw_ms) *mw_ms = c->mw_ms;
    if (timed_blocks) *timed_blocks = c->timed_blocks;
  });
}

int qf_claims_size(const qf_params* p, size_t* bytes) {
  return guarded([&] { require(p && bytes, QF_ERR_ARGUMENT, "null argument"); *bytes = claims_bytes(*p); });
}

int qf_benchmark_claims(const qf_params* p, const qf_committed* c, uint8_t* claims, size_t cap) {
  return guarded([&] {
    require(p && c && claims && c->p == p, QF_ERR_ARGUMENT, "null or mismatched argument");
    require(cap >= claims_bytes(*p), QF_ERR_ARGUMENT, "claims buffer too small");
    auto cl = benchmark_claims(*p, *c);
    uint8_t* o = claims;
    for (auto& x : cl) {
      for (auto& v : x.row) { std::memcpy(o, &v, 16); o += 16; }
      for (auto& v : x.col) { std::memcpy(o, &v, 16); o += 16; }
      std::memcpy(o, &x.value, 16);
      o += 16;
    }
  });
}

int qf_evaluate_claims(const qf_params* p, const qf_committed* c, uint8_t* claims, size_t claims_len) {
  return guarded([&] {
    require(p && c && claims && c->p == p, QF_ERR_ARGUMENT, "null or mismatched argument");
    require(claims_len == claims_bytes(*p), QF_ERR_INVALID_PARAMETERS, "claims length mismatch");
    auto cl = parse_claims(*p, claims);
    evaluate_claims(*p, *c, cl);
    uint8_t* o = claims;
    for (auto& x : cl) {
      o += 16 * (x.row.size() + x.col.size());
      std::memcpy(o, &x.value, 16);
      o += 16;
    }
  });
}

int qf_prove(const qf_params* p, const qf_committed* c, const uint8_t* claims, size_t claims_len, qf_proof** out) {
  return guarded([&] {
    require(p && c && claims && out && c->p == p, QF_ERR_ARGUMENT, "null or mismatched argument");
    require(claims_len == claims_bytes(*p), QF_ERR_INVALID_PARAMETERS, "claims length mismatch");
    std::unique_ptr<qf_proof> pr(new qf_proof);
    pr->bytes = prove(*p, *c, parse_claims(*p, claims), pr->t);
    *out = pr.release();
  });
}

int qf_verify(const qf_params* p, const uint8_t* roots, size_t n_roots, const uint8_t* claims, size_t claims_len,
              const uint8_t* proof, size_t proof_len) {
  return guarded([&] {
    require(p && roots && claims && proof, QF_ERR_ARGUMENT, "null argument");
    require(claims_len == claims_bytes(*p), QF_ERR_INVALID_PARAMETERS, "claims length mismatch");
    std::vector<Digest> rs(n_roots);
    for (size_t i = 0; i < n_roots; i++) std::memcpy(rs[i].data(), roots + 32 * i, 32);
    verify(*p, rs, parse_claims(*p, claims), proof, proof_len);
  });
}

int qf_verify_batch(const qf_params* p, const uint8_t* roots, size_t n_roots, const uint8_t* const* claims,
                    const size_t* claims_lens, const uint8_t* const* proofs, const size_t* proof_lens, size_t count,
                    int threads, int* statuses) {
  return guarded([&] {
    require(p && roots && statuses && (count == 0 || (claims && claims_lens && proofs && proof_lens)), QF_ERR_ARGUMENT,
            "null argument");
    std::vector<Digest> rs(n_roots);
    for (size_t i = 0; i < n_roots; i++) std::memcpy(rs[i].data(), roots + 32 * i, 32);
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t workers = std::max<size_t>(1, std::min<size_t>(count, threads > 0 ? size_t(threads) : (hw ? hw : 1)));
    std::vector<std::string> msgs(count);
    std::atomic<size_t> next{0};
    auto work = [&] {
      set_verify_rows_parallel(workers == 1);  // one proof per thread: rows hash serially on each
      for (size_t i; (i = next++) < count;) {
        try {
          require(proofs[i] && claims[i], QF_ERR_ARGUMENT, "null proof or claims");
          require(claims_lens[i] == claims_bytes(*p), QF_ERR_INVALID_PARAMETERS, "claims length mismatch");
          verify(*p, rs, parse_claims(*p, claims[i]), proofs[i], proof_lens[i]);
          statuses[i] = QF_OK;
        } catch (const QfError& e) {
          statuses[i] = e.status;
          msgs[i] = e.what();
        } catch (const std::exception& e) {
          statuses[i] = QF_ERR_ARGUMENT;
          msgs[i] = e.what();
        }
      }
      set_verify_rows_parallel(true);
    };
    std::vector<std::thread> ts;
    for (size_t w = 1; w < workers; w++) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
    for (size_t i = 0; i < count; i++)
      if (statuses[i] != QF_OK) throw QfError(statuses[i], "proof " + std::to_string(i) + ": " + msgs[i]);
  });
}

int qf_proof_bytes(const qf_proof* pr, const uint8_t** data, size_t* len) {
  return guarded([&] {
    require(pr && data && len, QF_ERR_ARGUMENT, "null argument");
    *data = pr->bytes.data();
    *len = pr->bytes.size();
  });
}

int qf_proof_timing(const qf_proof* pr, double out[9]) {
  return guarded([&] {
    require(pr && out, QF_ERR_ARGUMENT, "null argument");
    const ProveTimings& t = pr->t;
    double v[9] = {t.conversion_ms, t.eo_ms, t.projection_ms, t.folding_ms, t.sumcheck_ms, t.binding_ms, t.native_ms, t.decide_ms, t.total_ms};
    std::memcpy(out, v, sizeof v);
  });
}

void qf_proof_free(qf_proof* pr) { delete pr; }

int qf_statement_json(const qf_params* p, const qf_committed* c, char* out, size_t cap, size_t* len) {
  return guarded([&] {
    require(p && c && len && c->p == p, QF_ERR_ARGUMENT, "null or mismatched argument");
    std::string s = statement_json(*p, *c);
    *len = s.size();
    require(out && cap > s.size(), QF_ERR_ARGUMENT, "statement buffer too small");
    std::memcpy(out, s.c_str(), s.size() + 1);
  });
}

int qf_test_tile_pass(int mode, const void* device_in, void* device_out1, void* device_out2,
                      const void* device_alpha, uint64_t vectors, uint64_t in_vstride, uint64_t out_vstride, int L,
                      int s, int t, int xform_vectors) {
  return guarded([&] {
    require(device_in && device_out1 && L >= 0 && L < 40, QF_ERR_ARGUMENT, "null argument");
    device_setup();
    const E8* alpha = static_cast<const E8*>(device_alpha);
    E8* enc = nullptr;  // alpha_1 in the encoding kernels' form
    if (alpha && LANES != 8) {
      QF_HIP(hipMalloc(&enc, (uint64_t(1) << L) * 16));
      launch_encode_form(alpha, enc, uint64_t(1) << L, nullptr);
      alpha = enc;
    }
    bool ok = tile_pass(mode, false, device_in, static_cast<E8*>(device_out1), static_cast<E8*>(device_out2), alpha,
                        vectors, in_vstride, out_vstride, L, s, t, 0, 0, nullptr, xform_vectors != 0);
    hipError_t e = hipDeviceSynchronize();
    if (enc) (void)hipFree(enc);
    QF_HIP(e);
    require(ok, QF_ERR_ARGUMENT, "geometry does not fit a tile pass");
  });
}

int qf_autotune_report(char* out, size_t cap, size_t* len) {
  return guarded([&] {
    require(len != nullptr, QF_ERR_ARGUMENT, "null argument");
    const std::string r = autotune_report();
    *len = r.size();
    if (out && cap) {
      const size_t m = std::min(cap - 1, r.size());
      std::memcpy(out, r.data(), m);
      out[m] = 0;
    }
  });
}

int qf_set_launch_mode(int mode) {
  return guarded([&] {
    require(mode == LAUNCH_BATCHED || mode == LAUNCH_PER_MATRIX, QF_ERR_ARGUMENT, "launch mode");
    set_launch_mode(mode);
  });
}

int qf_device_memory(size_t* free_bytes, size_t* total_bytes) {
  return guarded([&] {
    require(free_bytes && total_bytes, QF_ERR_ARGUMENT, "null argument");
    device_setup();
    QF_HIP(hipMemGetInfo(free_bytes, total_bytes));
    *free_bytes += cached_idle_bytes();  // the pool hands its idle blocks out again
  });
}

}  // extern "C"
