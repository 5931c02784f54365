// Device buffers and the synthetic BVA generator (plumbing for resident inputs).
#include "../../include/parcomm_amd.h"
#include "qf_kernels.hpp"

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
