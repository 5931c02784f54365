// qf_commit_roots_noalloc: the roots of a commitment computed with no device
// allocation, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid it). Every device buffer
// of the commit lives in a fixed workspace of this file, one per device.
#include "../../include/parcomm_amd.h"
#include "qf_state.hpp"
#include <algorithm>
#include <cstring>
#include <mutex>

struct qf_params : qf::ParamsState {};

namespace qf {
void set_last_error(const char* msg);
}

using namespace qf;

namespace {
// Device workspace of qf_commit_roots_noalloc: one block's codeword and Merkle
// tree, alpha_1 (canonical and encode form), the padding-check flag and a
// window of padding ranges. Each device holds its own copy from module load
// until the process ends; its address comes from hipGetSymbolAddress, once
// per device (the name below is only a host-side stand-in).
constexpr size_t NOALLOC_BYTES = size_t(64) << 20;
constexpr size_t NOALLOC_RANGE_BYTES = size_t(1) << 20;  // 65536 (start, len) pairs per check launch
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

struct NoallocDevice {
  uint8_t* ptr = nullptr;  // device address of this device's `scratch`
  std::mutex mu;           // calls on one device run one at a time: they share the workspace
};
constexpr int NOALLOC_MAX_DEVICES = 64;
NoallocDevice g_noalloc[NOALLOC_MAX_DEVICES];
std::mutex g_noalloc_mu;

NoallocDevice& noalloc_device() {
  int dev = 0;
  QF_HIP(hipGetDevice(&dev));
  require(dev >= 0 && dev < NOALLOC_MAX_DEVICES, QF_ERR_DEVICE, "device index");
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  NoallocDevice& d = g_noalloc[dev];
  if (!d.ptr) {  // first call on this device, with it current
    void* ptr = nullptr;
    QF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    require(ptr != nullptr, QF_ERR_DEVICE, "no device address for the workspace");
    d.ptr = static_cast<uint8_t*>(ptr);
  }
  return d;
}

__global__ void k_zero_word(uint32_t* w) { *w = 0; }

template <class F>
int guarded_call(F&& f) {
  set_last_error("");
  try { f(); return QF_OK; } catch (const QfError& e) { set_last_error(e.what()); return e.status; }
  catch (const std::exception& e) { set_last_error(e.what()); return QF_ERR_ARGUMENT; }
}

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct NoallocLayout {
  size_t cw, tree, alpha, alpha_enc, flag, ranges, total;
};
NoallocLayout noalloc_layout(const ParamsState* p) {
  const uint64_t n = uint64_t(1) << p->geom.log_n, R = 2 * n;
  NoallocLayout o;
  o.cw = 0;
  o.tree = align256(size_t(32 * R) * 16);
  o.alpha = o.tree + align256(size_t(2 * R - 1) * 32);
  o.alpha_enc = o.alpha + align256(size_t(n) * 16);
  o.flag = o.alpha_enc + align256(size_t(n) * 16);
  o.ranges = o.flag + 256;
  o.total = o.ranges + NOALLOC_RANGE_BYTES;
  return o;
}

// Commitment.roots of the BVA (the roots qf_commit_device gives in either
// cache mode) with every device buffer in the workspace at `ws`.
void commit_roots_noalloc(const ParamsState* p, const Val* d_bva, size_t len, uint8_t* roots, uint8_t* ws,
                          hipStream_t st) {
  const NoallocLayout o = noalloc_layout(p);
  const uint64_t n = uint64_t(1) << p->geom.log_n, R = 2 * n;
  E8* cw = reinterpret_cast<E8*>(ws + o.cw);
  uint32_t* tree = reinterpret_cast<uint32_t*>(ws + o.tree);
  E8* alpha = reinterpret_cast<E8*>(ws + o.alpha);
  E8* alpha_enc = reinterpret_cast<E8*>(ws + o.alpha_enc);
  uint32_t* flag = reinterpret_cast<uint32_t*>(ws + o.flag);
  uint64_t* ranges = reinterpret_cast<uint64_t*>(ws + o.ranges);
  // padding check (layout.rs:136-155), through the range window
  {
    std::lock_guard<std::mutex> g(p->mu);
    if (p->zero_ranges.empty()) p->zero_ranges = padding_ranges(p->layout);
  }
  hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, st, flag);
  const size_t count = p->zero_ranges.size() / 2, window = NOALLOC_RANGE_BYTES / 16;
  for (size_t first = 0; first < count; first += window) {
    const size_t k = std::min(window, count - first);
    QF_HIP(hipMemcpyAsync(ranges, p->zero_ranges.data() + 2 * first, k * 16, hipMemcpyHostToDevice, st));
    launch_check_zero(d_bva, ranges, k, flag, st);
    QF_HIP(hipStreamSynchronize(st));  // the next window reuses the ranges area
  }
  uint32_t bad = 0;
  QF_HIP(hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, st));
  QF_HIP(hipStreamSynchronize(st));
  require(bad == 0, QF_ERR_INVALID_PARAMETERS, "nonzero boundary padding");
  const std::vector<E8>& a = p->pp.codes.back().alpha1;
  require(a.size() == n, QF_ERR_ARGUMENT, "alpha length");
  QF_HIP(hipMemcpyAsync(alpha, a.data(), n * 16, hipMemcpyHostToDevice, st));
  launch_encode_form(alpha, alpha_enc, n, st);
  set_tuning_allowed(false);
  struct Restore { ~Restore() { set_tuning_allowed(true); } } restore;
  int zero_block = -1;
  for (uint64_t b = 0; b < p->blocks; b++) {
    const uint64_t first = b * p->block_len;
    if (first >= len && zero_block >= 0) {  // identical all-zero blocks share one root
      std::memcpy(roots + 32 * b, roots + 32 * zero_block, 32);
      continue;
    }
    launch_commit_block(p->geom, d_bva, first, len, cw, alpha_enc, tree, st, nullptr);
    QF_HIP(hipGetLastError());
    QF_HIP(hipMemcpyAsync(roots + 32 * b, tree + 8 * (2 * R - 2), 32, hipMemcpyDeviceToHost, st));
    QF_HIP(hipStreamSynchronize(st));  // the next block reuses the workspace
    if (first >= len) zero_block = int(b);
  }
}

}  // namespace

extern "C" {

int qf_commit_roots_noalloc(const qf_params* p, const void* device_values, size_t len, uint8_t* roots,
                            size_t cap_blocks) {
  return guarded_call([&] {
    require(p && device_values && roots, QF_ERR_ARGUMENT, "null argument");
    require(reinterpret_cast<uintptr_t>(device_values) % 16 == 0, QF_ERR_ARGUMENT, "device BVA must be 16-byte aligned");
    require(len == p->layout.active, QF_ERR_INVALID_PARAMETERS, "active BVA length mismatch");
    require(cap_blocks >= p->blocks, QF_ERR_ARGUMENT, "root buffer too small");
    const NoallocLayout o = noalloc_layout(p);
    require(o.total <= NOALLOC_BYTES, QF_ERR_INVALID_PARAMETERS,
            "block too large for the fixed workspace of qf_commit_roots_noalloc (64 MiB: chunk_log <= 23 for "
            "F65521, <= 22 for KoalaBear)");
    device_setup();
    NoallocDevice& d = noalloc_device();
    std::lock_guard<std::mutex> g(d.mu);
    hipStream_t st = nullptr;  // the null stream: nothing is created for the call
    try {
      commit_roots_noalloc(p, static_cast<const Val*>(device_values), len, roots, d.ptr, st);
    } catch (...) {
      (void)hipStreamSynchronize(st);
      throw;
    }
  });
}

}  // extern "C"
