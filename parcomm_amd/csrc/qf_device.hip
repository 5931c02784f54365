// Device buffers and the synthetic BVA generator (plumbing for resident inputs).
#include "../../include/parcomm_amd.h"
#include "qf_kernels.hpp"
#include <cstdio>
#include <cstring>

using namespace qf;

struct qf_params;
namespace qf {
const Layout& params_layout(const qf_params* p);
}

namespace qf {
void set_last_error(const char* msg);
}

namespace {

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Entry i of the BVA: a Box-Muller normal z from two uniforms of
// mix64(seed, i), scaled by sigma. F65521 entries are round(z sigma) clipped
// to int8; KoalaBear entries are the bfloat16 bit pattern of z sigma (round to
// nearest even), the raw 16-bit payload of a BF16 weight.
__global__ void k_synthetic(Val* out, uint64_t n, uint64_t seed, float sigma) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) {
    uint64_t r = mix64(seed * 0x9e3779b97f4a7c15ull + i);
    float u1 = (float(uint32_t(r >> 40)) + 0.5f) * (1.0f / 16777216.0f);
    float u2 = float(uint32_t(r) & 0xffffffu) * (1.0f / 16777216.0f);
    float z = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
#ifdef QF_KB
    const uint32_t b = __float_as_uint(z * sigma);
    const uint32_t rb = b + 0x7fffu + ((b >> 16) & 1u);  // finite inputs only
    out[i] = Val(uint16_t(rb >> 16));
#else
    float v = rintf(z * sigma);
    v = fminf(fmaxf(v, -128.0f), 127.0f);
    out[i] = Val(int(v));
#endif
  }
}

// Plain copies, 16 bytes per thread per step: the calibration kernels of the
// HBM traffic counters. k_copy moves the whole buffer contiguously;
// k_copy_half moves only the 64-byte half `parity` of every 128-byte line (the
// row shape of the tile passes over high index bits, 4 contiguous elements),
// with the other half never touched in the launch, which shows what one
// isolated 64-byte read costs.
__global__ void k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n16) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += uint64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}
__global__ void k_copy_half(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n16, int parity) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16 / 2; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t j = 4 * (2 * (i / 4) + parity) + (i % 4);  // run i/4 -> line i/4, half `parity`
    dst[j] = src[j];
  }
}

// Zero a rows x width rectangle of a row-major bucket (padding region).
__global__ void k_zero_rect(Val* base, uint64_t rows, uint64_t stride, uint64_t width) {
  uint64_t total = rows * width;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += uint64_t(gridDim.x) * blockDim.x)
    base[(i / width) * stride + (i % width)] = 0;
}

// Layout::encode (layout.rs:115-135) on the device. A bucket's rows x cols
// entries (both powers of two) are cut into chunks of 16 bytes (CH entries),
// or of the whole bucket when it is smaller; entry (r, c) of a bucket is
// matrix entry (row_offset + r, col_offset + c) for r < physical_rows and
// c < physical_cols, and zero otherwise. Buckets are aligned to their length,
// so a 16-byte chunk never straddles two buckets, and a work unit of
// ENC_WG x ENC_PER chunks belongs to one bucket.
constexpr int ENC_WG = 256, ENC_PER = 8;
constexpr uint64_t ENC_UNIT = uint64_t(ENC_WG) * ENC_PER;
constexpr int ENC_CH = 16 / int(sizeof(Val));  // entries per 16-byte chunk
struct EncodeTask {
  uint64_t src;         // first source entry: row_offset * mcols + col_offset
  uint64_t mcols;       // source row stride (the matrix's columns)
  uint64_t dst;         // bucket offset in the BVA (entries)
  uint64_t chunks;      // chunks of the bucket
  uint64_t first_unit;  // first work unit of the bucket
  uint32_t log_cols, ch, prow, pcol, mat;
};

__global__ __launch_bounds__(ENC_WG) void k_encode(const Val* const* __restrict__ mats,
                                                 const EncodeTask* __restrict__ tasks,
                                                 const uint32_t* __restrict__ unit_task, Val* __restrict__ bva) {
  const EncodeTask t = tasks[unit_task[blockIdx.x]];
  const Val* __restrict__ m = mats[t.mat] + t.src;
  const uint64_t cmask = (uint64_t(1) << t.log_cols) - 1;
  const uint64_t c0 = (blockIdx.x - t.first_unit) * ENC_UNIT + threadIdx.x;
  for (int k = 0; k < ENC_PER; k++) {
    const uint64_t chunk = c0 + uint64_t(k) * ENC_WG;
    if (chunk >= t.chunks) break;
    const uint64_t e0 = chunk * t.ch;
    Val* d = bva + t.dst + e0;
    auto entry = [&](uint64_t e) -> Val {
      const uint64_t r = e >> t.log_cols, c = e & cmask;
      return (r < t.prow && c < t.pcol) ? m[r * t.mcols + c] : Val(0);
    };
    if (t.ch == ENC_CH) {  // a whole 16-byte store, inside one row when the bucket is at least CH wide
      const uint64_t r = e0 >> t.log_cols, c = e0 & cmask;
      const Val* s = m + r * t.mcols + c;
      if (cmask + 1 >= uint64_t(ENC_CH) && r < t.prow && c + ENC_CH <= t.pcol &&
          (reinterpret_cast<uintptr_t>(s) & 15) == 0) {
        *reinterpret_cast<uint4*>(d) = *reinterpret_cast<const uint4*>(s);
      } else {
        union { Val v[ENC_CH]; uint4 w; } u;
        for (int i = 0; i < ENC_CH; i++) u.v[i] = entry(e0 + i);
        *reinterpret_cast<uint4*>(d) = u.w;
      }
    } else {  // a bucket shorter than 16 bytes: entry stores
      for (uint32_t i = 0; i < t.ch; i++) d[i] = entry(e0 + i);
    }
  }
}

template <class F>
int guard(F&& f) {
  set_last_error("");
  try { f(); return QF_OK; } catch (const QfError& e) { set_last_error(e.what()); return e.status; }
  catch (const std::exception& e) { set_last_error(e.what()); return QF_ERR_ARGUMENT; }
}
}  // namespace

extern "C" {

int qf_device_alloc(size_t bytes, void** out) {
  return guard([&] {
    device_setup();
    require(out != nullptr, QF_ERR_ARGUMENT, "null argument");
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e == hipErrorOutOfMemory) {  // the pool's idle blocks may hold the memory: give them back and retry
      (void)hipGetLastError();
      release_cached_blocks();
      e = hipMalloc(out, bytes ? bytes : 1);
    }
    QF_HIP(e);
    QF_HIP(hipMemset(*out, 0, bytes ? bytes : 1));
  });
}
int qf_device_upload(void* dst, const void* src, size_t bytes) {
  return guard([&] { QF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int qf_device_download(void* dst, const void* src, size_t bytes) {
  return guard([&] { QF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int qf_device_free(void* p) {
  return guard([&] { if (p) QF_HIP(hipFree(p)); });
}
int qf_device_copy(void* dst, const void* src, size_t bytes, int pattern) {
  return guard([&] {
    require(dst && src, QF_ERR_ARGUMENT, "null argument");
    require(pattern >= 0 && pattern <= 2, QF_ERR_ARGUMENT, "copy pattern must be 0, 1 or 2");
    require(bytes % 16 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0,
            QF_ERR_ARGUMENT, "copy needs 16-byte aligned buffers and sizes");
    require(pattern == 0 || (bytes % 128 == 0 && reinterpret_cast<uintptr_t>(dst) % 128 == 0 &&
                             reinterpret_cast<uintptr_t>(src) % 128 == 0),
            QF_ERR_ARGUMENT, "half-line copies need 128-byte aligned buffers and sizes");
    if (pattern)
      hipLaunchKernelGGL(k_copy_half, dim3(65536), dim3(256), 0, 0, static_cast<uint4*>(dst),
                         static_cast<const uint4*>(src), uint64_t(bytes / 16), pattern - 1);
    else
      hipLaunchKernelGGL(k_copy, dim3(65536), dim3(256), 0, 0, static_cast<uint4*>(dst), static_cast<const uint4*>(src),
                         uint64_t(bytes / 16));
    QF_HIP(hipGetLastError());
    QF_HIP(hipDeviceSynchronize());
  });
}

}  // extern "C"

namespace {
// Scatters the buckets of the matrices in `which` (device pointers in `mats`,
// indexed by matrix) into the BVA on `st`.
void encode_launch(const Layout& l, const std::vector<uint64_t>& which, const std::vector<const Val*>& mats, Val* bva,
                   hipStream_t st) {
  std::vector<EncodeTask> tasks;
  std::vector<uint32_t> unit_task;
  std::vector<char> want(l.shapes.size(), 0);
  for (uint64_t m : which) want[m] = 1;
  for (const Bucket& b : l.buckets) {
    if (!want[b.matrix] || b.len() == 0) continue;
    EncodeTask t;
    t.src = b.row_offset * l.shapes[b.matrix].cols + b.col_offset;
    t.mcols = l.shapes[b.matrix].cols;
    t.dst = b.offset;
    t.ch = uint32_t(std::min<uint64_t>(ENC_CH, b.len()));
    t.chunks = b.len() / t.ch;
    t.first_unit = unit_task.size();
    t.log_cols = uint32_t(ilog2u(b.cols));
    t.prow = uint32_t(b.physical_rows);
    t.pcol = uint32_t(b.physical_cols);
    t.mat = uint32_t(b.matrix);
    require((uint64_t(1) << t.log_cols) == b.cols && b.len() % t.ch == 0 && b.offset % t.ch == 0, QF_ERR_ARGUMENT,
            "bucket geometry");
    const uint64_t units = (t.chunks + ENC_UNIT - 1) / ENC_UNIT;
    unit_task.insert(unit_task.end(), units, uint32_t(tasks.size()));
    tasks.push_back(t);
  }
  if (tasks.empty()) return;
  require(unit_task.size() < (1ull << 31), QF_ERR_ARGUMENT, "encode grid too large");
  auto* d_mats = static_cast<const Val**>(dmalloc(mats.size() * sizeof(Val*), st));
  auto* d_tasks = static_cast<EncodeTask*>(dmalloc(tasks.size() * sizeof(EncodeTask), st));
  auto* d_units = static_cast<uint32_t*>(dmalloc(unit_task.size() * 4, st));
  QF_HIP(hipMemcpyAsync(d_mats, mats.data(), mats.size() * sizeof(Val*), hipMemcpyHostToDevice, st));
  QF_HIP(hipMemcpyAsync(d_tasks, tasks.data(), tasks.size() * sizeof(EncodeTask), hipMemcpyHostToDevice, st));
  QF_HIP(hipMemcpyAsync(d_units, unit_task.data(), unit_task.size() * 4, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_encode, dim3(uint32_t(unit_task.size())), dim3(ENC_WG), 0, st, d_mats, d_tasks, d_units, bva);
  QF_HIP(hipGetLastError());
  QF_HIP(hipStreamSynchronize(st));  // the host tables above are released when this returns
  dfree(d_mats, st);
  dfree(d_tasks, st);
  dfree(d_units, st);
}

void check_matrix(const Layout& l, uint64_t m, uint64_t len) {
  require(m < l.shapes.size(), QF_ERR_INVALID_PARAMETERS, "matrix count mismatch");
  require(len == l.shapes[m].rows * l.shapes[m].cols, QF_ERR_INVALID_PARAMETERS, "matrix length mismatch");
}
}  // namespace

extern "C" {

int qf_encode(const qf_params* p, const void* const* matrices, const uint64_t* lens, size_t n, int on_device,
              void* device_bva) {
  return guard([&] {
    require(p && device_bva && (n == 0 || (matrices && lens)), QF_ERR_ARGUMENT, "null argument");
    require(reinterpret_cast<uintptr_t>(device_bva) % 16 == 0, QF_ERR_ARGUMENT, "device BVA must be 16-byte aligned");
    const Layout& l = params_layout(p);
    require(n == l.shapes.size(), QF_ERR_INVALID_PARAMETERS, "matrix count mismatch");  // layout.rs:116
    for (size_t i = 0; i < n; i++) check_matrix(l, i, lens[i]);                        // layout.rs:117-122
    device_setup();
    hipStream_t st = thread_stream();
    Val* bva = static_cast<Val*>(device_bva);
    if (on_device) {  // every bucket of every matrix in one launch
      std::vector<uint64_t> all(n);
      std::vector<const Val*> mats(n);
      for (size_t i = 0; i < n; i++) { all[i] = i; mats[i] = static_cast<const Val*>(matrices[i]); }
      encode_launch(l, all, mats, bva, st);
      return;
    }
    // host matrices stream through one device staging buffer, one matrix at a time
    uint64_t most = 1;
    for (size_t i = 0; i < n; i++) most = std::max<uint64_t>(most, lens[i]);
    Val* stage = static_cast<Val*>(dmalloc(most * sizeof(Val), st));
    std::vector<const Val*> mats(n, stage);
    try {
      for (size_t i = 0; i < n; i++) {
        if (lens[i]) QF_HIP(hipMemcpyAsync(stage, matrices[i], lens[i] * sizeof(Val), hipMemcpyHostToDevice, st));
        encode_launch(l, {uint64_t(i)}, mats, bva, st);
      }
    } catch (...) {
      (void)hipStreamSynchronize(st);
      dfree(stage, st);
      throw;
    }
    dfree(stage, st);
  });
}

int qf_encode_matrix(const qf_params* p, uint64_t matrix, const void* values, uint64_t len, int on_device,
                     void* device_bva) {
  return guard([&] {
    require(p && device_bva && (values || len == 0), QF_ERR_ARGUMENT, "null argument");
    require(reinterpret_cast<uintptr_t>(device_bva) % 16 == 0, QF_ERR_ARGUMENT, "device BVA must be 16-byte aligned");
    const Layout& l = params_layout(p);
    check_matrix(l, matrix, len);
    device_setup();
    hipStream_t st = thread_stream();
    std::vector<const Val*> mats(l.shapes.size(), nullptr);
    Val* stage = nullptr;
    if (on_device) {
      mats[matrix] = static_cast<const Val*>(values);
    } else {
      stage = static_cast<Val*>(dmalloc(std::max<uint64_t>(len, 1) * sizeof(Val), st));
      if (len) QF_HIP(hipMemcpyAsync(stage, values, len * sizeof(Val), hipMemcpyHostToDevice, st));
      mats[matrix] = stage;
    }
    try {
      encode_launch(l, {matrix}, mats, static_cast<Val*>(device_bva), st);
    } catch (...) {
      (void)hipStreamSynchronize(st);
      if (stage) dfree(stage, st);
      throw;
    }
    if (stage) dfree(stage, st);
  });
}

int qf_stream_bva(const qf_params* p, const char* path, void* device_bva, uint64_t* bytes_read) {
  return guard([&] {
    require(p && path && device_bva, QF_ERR_ARGUMENT, "null argument");
    const Layout& l = params_layout(p);
    const uint64_t total = l.active * sizeof(Val);
    device_setup();
    hipStream_t st = thread_stream();
    FILE* f = std::strcmp(path, "-") == 0 ? stdin : std::fopen(path, "rb");
    require(f != nullptr, QF_ERR_ARGUMENT, "cannot open the BVA input");
    // two pinned staging buffers: the file is read into one while the other uploads
    constexpr size_t CHUNK = size_t(64) << 20;
    uint8_t* stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    auto cleanup = [&] {
      (void)hipStreamSynchronize(st);
      for (int i = 0; i < 2; i++) {
        if (stage[i]) (void)hipHostFree(stage[i]);
        if (done[i]) (void)hipEventDestroy(done[i]);
      }
      if (f != stdin) std::fclose(f);
    };
    try {
      for (int i = 0; i < 2; i++) {
        QF_HIP(hipHostMalloc(&stage[i], CHUNK, hipHostMallocDefault));
        QF_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
      }
      uint64_t off = 0;
      for (int k = 0; off < total; k ^= 1) {
        QF_HIP(hipEventSynchronize(done[k]));  // the upload from this buffer has finished
        const size_t want = size_t(std::min<uint64_t>(CHUNK, total - off));
        const size_t got = std::fread(stage[k], 1, want, f);
        off += got;
        if (bytes_read) *bytes_read = off;
        require(got == want, QF_ERR_ENCODING, "BVA input ends early");  // main.rs:160-169 (read_exact)
        QF_HIP(hipMemcpyAsync(static_cast<uint8_t*>(device_bva) + off - got, stage[k], got, hipMemcpyHostToDevice, st));
        QF_HIP(hipEventRecord(done[k], st));
      }
      uint8_t extra = 0;
      require(std::fread(&extra, 1, 1, f) == 0, QF_ERR_ENCODING, "excess BVA input");  // main.rs:170-173
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int qf_synthetic_bva(const qf_params* p, uint64_t seed, float sigma, void* device_out) {
  return guard([&] {
    require(p && device_out, QF_ERR_ARGUMENT, "null argument");
    const Layout& l = params_layout(p);
    Val* out = static_cast<Val*>(device_out);
    hipLaunchKernelGGL(k_synthetic, dim3(65536), dim3(256), 0, 0, out, l.active, seed, sigma);
    for (const Bucket& b : l.buckets) {
      Val* base = out + b.offset;
      if (b.physical_cols < b.cols)
        hipLaunchKernelGGL(k_zero_rect, dim3(1024), dim3(256), 0, 0, base + b.physical_cols, b.physical_rows, b.cols,
                           b.cols - b.physical_cols);
      if (b.physical_rows < b.rows)
        hipLaunchKernelGGL(k_zero_rect, dim3(1024), dim3(256), 0, 0, base + b.physical_rows * b.cols,
                           b.rows - b.physical_rows, b.cols, b.cols);
    }
    QF_HIP(hipGetLastError());
    QF_HIP(hipDeviceSynchronize());
  });
}

}  // extern "C"
