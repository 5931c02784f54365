// Device vector operations and BVA-scale kernels of the opening (see qf_vec.hpp).
#include "qf_vec.hpp"
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <type_traits>
#include <map>
#include <string>
#include <cstdio>
#include <mutex>
#include <unordered_map>

namespace qf {

namespace {
constexpr int WG = 256;
constexpr uint64_t TAB_CAP = 1u << 18;  // E8 entries of the small-table scratch
constexpr uint32_t PROJ_TAB_LOG = 12;   // low-bit weight table of a projection chunk: 2^12 entries
constexpr uint32_t MF_TAB_LOG = 11;     // the same for the MFMA projection: 32 KiB, three workgroups per CU

__device__ inline V8 v8_madd(const V8& acc, const V8& x) { return v8_add(acc, x); }

// ---------------------------------------------------------------- reductions
enum { R_INNER = 0, R_CROSS = 1, R_SC = 3 };

template <int MODE>
__global__ __launch_bounds__(WG) void k_reduce(const E8* __restrict__ a, const E8* __restrict__ b, uint64_t n,
                                               uint32_t* __restrict__ partials) {
  __shared__ uint32_t red[WG][16];
  V8 s0 = v8_zero(), s1 = v8_zero();
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG) {
    if (MODE == R_INNER) {
      s0 = v8_add(s0, v8_mul(v8_load(a[i]), v8_load(b[i])));
    } else if (MODE == R_CROSS) {
      s0 = v8_add(s0, v8_mul(v8_load(a[i]), v8_load(b[n + i])));
      s1 = v8_add(s1, v8_mul(v8_load(a[n + i]), v8_load(b[i])));
    } else {
      V8 vl = v8_load(a[i]), wl = v8_load(b[i]);
      s0 = v8_add(s0, v8_mul(vl, wl));
      s1 = v8_add(s1, v8_mul(v8_sub(v8_load(a[n + i]), vl), v8_sub(v8_load(b[n + i]), wl)));
    }
  }
#pragma unroll
  for (int k = 0; k < LANES; k++) { red[threadIdx.x][k] = s0.c[k]; red[threadIdx.x][LANES + k] = s1.c[k]; }
  __syncthreads();
  for (int w = WG / 2; w > 0; w /= 2) {
    if (int(threadIdx.x) < w)
      for (int k = 0; k < 2 * LANES; k++) red[threadIdx.x][k] = madd(red[threadIdx.x][k], red[threadIdx.x + w][k]);
    __syncthreads();
  }
  if (threadIdx.x < 2 * LANES) partials[blockIdx.x * 16 + threadIdx.x] = red[0][threadIdx.x];
}

// Block-wide sums of NS extension elements per thread; partials[group][32] u32.
constexpr int PSTRIDE = 32;
template <int NS>
__device__ inline void block_sums(const V8 (&s)[NS], uint32_t* __restrict__ partials) {
  __shared__ uint32_t red[WG][NS * LANES];
#pragma unroll
  for (int j = 0; j < NS; j++)
#pragma unroll
    for (int k = 0; k < LANES; k++) red[threadIdx.x][j * LANES + k] = s[j].c[k];
  __syncthreads();
  for (int w = WG / 2; w > 0; w /= 2) {
    if (int(threadIdx.x) < w)
      for (int k = 0; k < NS * LANES; k++) red[threadIdx.x][k] = madd(red[threadIdx.x][k], red[threadIdx.x + w][k]);
    __syncthreads();
  }
  if (threadIdx.x < NS * LANES) partials[blockIdx.x * PSTRIDE + threadIdx.x] = red[0][threadIdx.x];
}

// One LiteBullet fold in place (a[i] += a[h+i] x, b[i] += b[h+i] y, i < h;
// qafold_short.rs:2901-2926) fused with the next round's cross terms over the
// folded halves (h even, g = h/2): sum_{i<g} a'[i] b'[g+i] and sum a'[g+i] b'[i]
// (qafold_short.rs:2876-2898). Each thread owns positions i and g+i of both
// halves, so the in-place update reads nothing another thread writes.
__global__ __launch_bounds__(WG) void k_lb_fold_cross(E8* __restrict__ a, E8* __restrict__ b, uint64_t h, E8 x, E8 y,
                                                      uint32_t* __restrict__ partials) {
  const V8 vx = v8_encode_form(v8_load(x)), vy = v8_encode_form(v8_load(y));
  const uint64_t g = h / 2;
  V8 s[2] = {v8_zero(), v8_zero()};
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < g; i += uint64_t(gridDim.x) * WG) {
    const V8 a0 = v8_add(v8_load(a[i]), v8_mul_enc(v8_load(a[h + i]), vx));
    const V8 a1 = v8_add(v8_load(a[g + i]), v8_mul_enc(v8_load(a[h + g + i]), vx));
    const V8 b0 = v8_add(v8_load(b[i]), v8_mul_enc(v8_load(b[h + i]), vy));
    const V8 b1 = v8_add(v8_load(b[g + i]), v8_mul_enc(v8_load(b[h + g + i]), vy));
    a[i] = v8_store(a0);
    a[g + i] = v8_store(a1);
    b[i] = v8_store(b0);
    b[g + i] = v8_store(b1);
    s[0] = v8_add(s[0], v8_mul(a0, b1));
    s[1] = v8_add(s[1], v8_mul(a1, b0));
  }
  block_sums<2>(s, partials);
}

// One binary EO fold (eo.rs:57-79: w'[i] = w[2i] + w[2i+1] r^-1,
// q'[i] = q[2i] + q[2i+1] r) of length-2h vectors into length-h outputs,
// fused with the next round's cross terms over the outputs:
// sum_j w'[2j] q'[2j+1] and sum_j w'[2j+1] q'[2j] (h even).
// At most 128 registers, so the 1024 workgroups of a full launch are all resident.
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void k_eo_fold_cross(
    const E8* __restrict__ w, const E8* __restrict__ q, E8* __restrict__ ow, E8* __restrict__ oq, uint64_t h, E8 ri,
    E8 r, uint32_t* __restrict__ partials) {
  const V8 vri = v8_encode_form(v8_load(ri)), vr = v8_encode_form(v8_load(r));
  V8 s[2] = {v8_zero(), v8_zero()};
  const uint64_t n = h / 2, stride = uint64_t(gridDim.x) * WG;
  // the next iteration's eight inputs load while this one computes
  E8 nx[8];
  auto fetch = [&](uint64_t j) {
#pragma unroll
    for (int k = 0; k < 4; k++) { nx[k] = w[4 * j + k]; nx[4 + k] = q[4 * j + k]; }
  };
  uint64_t j = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (j < n) fetch(j);
  for (; j < n; j += stride) {
    E8 cx[8];
#pragma unroll
    for (int k = 0; k < 8; k++) cx[k] = nx[k];
    if (j + stride < n) fetch(j + stride);
    const V8 a0 = v8_add(v8_load(cx[0]), v8_mul_enc(v8_load(cx[1]), vri));
    const V8 a1 = v8_add(v8_load(cx[2]), v8_mul_enc(v8_load(cx[3]), vri));
    const V8 b0 = v8_add(v8_load(cx[4]), v8_mul_enc(v8_load(cx[5]), vr));
    const V8 b1 = v8_add(v8_load(cx[6]), v8_mul_enc(v8_load(cx[7]), vr));
    ow[2 * j] = v8_store(a0);
    ow[2 * j + 1] = v8_store(a1);
    oq[2 * j] = v8_store(b0);
    oq[2 * j + 1] = v8_store(b1);
    s[0] = v8_add(s[0], v8_mul(a0, b1));
    s[1] = v8_add(s[1], v8_mul(a1, b0));
  }
  block_sums<2>(s, partials);
}

uint32_t groups_for(uint64_t n) {
  uint64_t g = (n + WG - 1) / WG;
  if (g < 1) g = 1;
  return uint32_t(g < DevCtx::MAX_GROUPS ? g : DevCtx::MAX_GROUPS);
}

std::pair<E8, E8> finish_partials(DevCtx& c, uint32_t groups) {
  c.sync();  // the block sums are in the pinned buffer once the kernel has finished
  V8 s0 = v8_zero(), s1 = v8_zero();
  for (uint32_t g = 0; g < groups; g++)
    for (int k = 0; k < LANES; k++) {
      s0.c[k] = madd(s0.c[k], c.host_partials[g * 16 + k]);
      s1.c[k] = madd(s1.c[k], c.host_partials[g * 16 + LANES + k]);
    }
  return {v8_store(s0), v8_store(s1)};
}

// Sums of the first NS extension elements over `groups` partial rows (stride PSTRIDE).
template <int NS>
std::array<E8, NS> finish_sums(DevCtx& c, uint32_t groups) {
  c.sync();
  std::array<E8, NS> out;
  for (int j = 0; j < NS; j++) {
    V8 s = v8_zero();
    for (uint32_t g = 0; g < groups; g++)
      for (int k = 0; k < LANES; k++) s.c[k] = madd(s.c[k], c.host_partials[g * PSTRIDE + j * LANES + k]);
    out[j] = v8_store(s);
  }
  return out;
}

template <int MODE>
std::pair<E8, E8> reduce(DevCtx& c, const E8* a, const E8* b, uint64_t n) {
  if (n == 0) return {E8{}, E8{}};
  uint32_t g = groups_for(n);
  hipLaunchKernelGGL(k_reduce<MODE>, dim3(g), dim3(WG), 0, c.st, a, b, n, c.partials);
  QF_HIP(hipGetLastError());
  return finish_partials(c, g);
}

// ---------------------------------------------------------------- elementwise
__global__ void k_fill(E8* p, uint64_t n, E8 v) {
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG) p[i] = v;
}
__global__ void k_tensor(E8* out, uint64_t n, const E8* hi, const E8* lo, int a) {
  uint64_t mask = (uint64_t(1) << a) - 1;
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG)
    out[i] = v8_store(v8_mul(v8_load(hi[i >> a]), v8_load(lo[i & mask])));
}
// Kernels with a constant multiplier convert it once to the encode form
// (Montgomery for KoalaBear) and multiply with v8_mul_enc.
__global__ void k_lb_fold(E8* a, E8* b, uint64_t h, E8 x, E8 y) {
  const V8 vx = v8_encode_form(v8_load(x)), vy = v8_encode_form(v8_load(y));
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < h; i += uint64_t(gridDim.x) * WG) {
    a[i] = v8_store(v8_add(v8_load(a[i]), v8_mul_enc(v8_load(a[h + i]), vx)));
    b[i] = v8_store(v8_add(v8_load(b[i]), v8_mul_enc(v8_load(b[h + i]), vy)));
  }
}
__global__ void k_eo_fold(const E8* w, const E8* q, E8* ow, E8* oq, uint64_t h, E8 ri, E8 r) {
  const V8 vri = v8_encode_form(v8_load(ri)), vr = v8_encode_form(v8_load(r));
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < h; i += uint64_t(gridDim.x) * WG) {
    V8 w0 = v8_load(w[2 * i]), w1 = v8_load(w[2 * i + 1]), q0 = v8_load(q[2 * i]), q1 = v8_load(q[2 * i + 1]);
    ow[i] = v8_store(v8_add(w0, v8_mul_enc(w1, vri)));
    oq[i] = v8_store(v8_add(q0, v8_mul_enc(q1, vr)));
  }
}
__global__ void k_sc_fold(E8* v, E8* w, uint64_t h, E8 r) {
  const V8 vr = v8_encode_form(v8_load(r));
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < h; i += uint64_t(gridDim.x) * WG) {
    V8 vl = v8_load(v[i]), wl = v8_load(w[i]);
    v[i] = v8_store(v8_add(vl, v8_mul_enc(v8_sub(v8_load(v[h + i]), vl), vr)));
    w[i] = v8_store(v8_add(wl, v8_mul_enc(v8_sub(v8_load(w[h + i]), wl), vr)));
  }
}
// One sumcheck fold in place (x_lo += r (x_hi - x_lo), i < h; packed.rs:674-696)
// fused with the next round's message over the folded length h (h even, g =
// h/2): m0 = sum_{i<g} v'[i] w'[i], m1 = sum (v'[g+i] - v'[i]) (w'[g+i] - w'[i])
// (packed.rs:652-667). Each thread owns positions i and g+i of both vectors.
__global__ __launch_bounds__(WG) void k_sc_fold_msg(E8* __restrict__ v, E8* __restrict__ w, uint64_t h, E8 r,
                                                    uint32_t* __restrict__ partials) {
  const V8 vr = v8_encode_form(v8_load(r));
  const uint64_t g = h / 2;
  V8 s[2] = {v8_zero(), v8_zero()};
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < g; i += uint64_t(gridDim.x) * WG) {
    V8 x = v8_load(v[i]), y = v8_load(w[i]);
    const V8 v0 = v8_add(x, v8_mul_enc(v8_sub(v8_load(v[h + i]), x), vr));
    const V8 w0 = v8_add(y, v8_mul_enc(v8_sub(v8_load(w[h + i]), y), vr));
    x = v8_load(v[g + i]);
    y = v8_load(w[g + i]);
    const V8 v1 = v8_add(x, v8_mul_enc(v8_sub(v8_load(v[h + g + i]), x), vr));
    const V8 w1 = v8_add(y, v8_mul_enc(v8_sub(v8_load(w[h + g + i]), y), vr));
    v[i] = v8_store(v0);
    v[g + i] = v8_store(v1);
    w[i] = v8_store(w0);
    w[g + i] = v8_store(w1);
    s[0] = v8_add(s[0], v8_mul(v0, w0));
    s[1] = v8_add(s[1], v8_mul(v8_sub(v1, v0), v8_sub(w1, w0)));
  }
  block_sums<2>(s, partials);
}
__global__ void k_axpy(E8* y, const E8* x, uint64_t n, E8 s) {
  const V8 vs = v8_encode_form(v8_load(s));
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG)
    y[i] = v8_store(v8_add(v8_load(y[i]), v8_mul_enc(v8_load(x[i]), vs)));
}
__global__ void k_mul(E8* y, const E8* x, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG)
    y[i] = v8_store(v8_mul(v8_load(y[i]), v8_load(x[i])));
}
__global__ void k_scale(E8* y, uint64_t n, E8 s) {
  const V8 vs = v8_encode_form(v8_load(s));
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG)
    y[i] = v8_store(v8_mul_enc(v8_load(y[i]), vs));
}
// Column projection: grid (groups, cols); partials[(col * groups + g)][0..LANES)
__global__ __launch_bounds__(WG) void k_col_projection(const E8* __restrict__ cw, uint64_t R, uint64_t n,
                                                       const E8* __restrict__ w, uint32_t* __restrict__ partials) {
  __shared__ uint32_t red[WG][LANES];
  const E8* col = cw + uint64_t(blockIdx.y) * R;
  V8 s = v8_zero();
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG)
    s = v8_add(s, v8_mul(v8_load(col[i]), v8_load(w[i])));
  for (int k = 0; k < LANES; k++) red[threadIdx.x][k] = s.c[k];
  __syncthreads();
  for (int h = WG / 2; h > 0; h /= 2) {
    if (int(threadIdx.x) < h)
      for (int k = 0; k < LANES; k++) red[threadIdx.x][k] = madd(red[threadIdx.x][k], red[threadIdx.x + h][k]);
    __syncthreads();
  }
  if (threadIdx.x < LANES)
    partials[(uint64_t(blockIdx.y) * gridDim.x + blockIdx.x) * LANES + threadIdx.x] = red[0][threadIdx.x];
}
__global__ __launch_bounds__(WG) void k_col_fold(const E8* __restrict__ cw, uint64_t R, uint64_t n, uint64_t cols,
                                                 const E8* __restrict__ f, E8* __restrict__ out) {
  for (uint64_t r = uint64_t(blockIdx.x) * WG + threadIdx.x; r < n; r += uint64_t(gridDim.x) * WG) {
    V8 s = v8_zero();
    for (uint64_t c = 0; c < cols; c++) s = v8_add(s, v8_mul(v8_load(f[c]), v8_load(cw[c * R + r])));
    out[r] = v8_store(s);
  }
}
__global__ void k_gather_rows(const E8* cw, uint64_t R, uint64_t cols, const uint64_t* rows, uint64_t count, E8* out) {
  uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (i >= count * cols) return;
  uint64_t k = i / cols, c = i % cols;
  out[i] = cw[c * R + rows[k]];
}
__global__ void k_gather_weighted(const E8* __restrict__ a, uint64_t Ra, uint64_t ca, const E8* __restrict__ b,
                                  uint64_t Rb, uint64_t cb, const E8* __restrict__ w, const uint64_t* __restrict__ rows,
                                  uint64_t count, E8* __restrict__ out) {
  const uint64_t k = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (k >= count) return;
  const uint64_t r = rows[k];
  V8 s = v8_zero();
  for (uint64_t col = 0; col < ca; col++) s = v8_add(s, v8_mul(v8_load(a[col * Ra + r]), v8_load(w[col])));
  for (uint64_t col = 0; col < cb; col++) s = v8_add(s, v8_mul(v8_load(b[col * Rb + r]), v8_load(w[ca + col])));
  out[k] = v8_store(s);
}
// Serialized query rows, record k at out + k * per (per = 16 + 16 cols + 32
// depth): u64 cols, the cols elements of row rows[k] of a column-major
// codeword, u64 depth, then the depth sibling digests from the leaf level up
// (tree: the R leaf digests, then each level above). One thread per element,
// digest or header pair; records are 16-byte aligned, their fields 8-byte.
__global__ void k_row_records(const E8* __restrict__ cw, uint64_t R, uint64_t cols, const uint32_t* __restrict__ tree,
                              int depth, const uint64_t* __restrict__ rows, uint64_t count, uint8_t* __restrict__ out) {
  const uint64_t items = cols + uint64_t(depth) + 1, per = 16 + 16 * cols + 32 * uint64_t(depth);
  const uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (i >= count * items) return;
  const uint64_t k = i / items, j = i % items;
  uint8_t* rec = out + k * per;
  const uint64_t row = rows[k];
  if (j < cols) {
    const uint4 v = *reinterpret_cast<const uint4*>(&cw[j * R + row]);
    uint2* o = reinterpret_cast<uint2*>(rec + 8 + 16 * j);
    o[0] = make_uint2(v.x, v.y);
    o[1] = make_uint2(v.z, v.w);
  } else if (j < cols + uint64_t(depth)) {
    const int l = int(j - cols);
    uint64_t off = 0, cnt = R;
    for (int t = 0; t < l; t++) { off += cnt; cnt >>= 1; }
    const uint32_t* d = tree + 8 * (((row >> l) ^ 1) + off);
    uint2* o = reinterpret_cast<uint2*>(rec + 16 + 16 * cols + 32 * uint64_t(l));
#pragma unroll
    for (int w = 0; w < 4; w++) o[w] = make_uint2(d[2 * w], d[2 * w + 1]);
  } else {
    *reinterpret_cast<uint64_t*>(rec) = cols;
    *reinterpret_cast<uint64_t*>(rec + 8 + 16 * cols) = uint64_t(depth);
  }
}
__global__ void k_scatter_add(E8* dst, const uint64_t* idx, const E8* vals, uint64_t n) {
  uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (i >= n) return;
  dst[idx[i]] = v8_store(v8_add(v8_load(dst[idx[i]]), v8_load(vals[i])));
}

// ---------------------------------------------------------------- BVA-scale kernels
struct ProjTaskDev {
  uint64_t base, rows, chunk_first;  // chunk_first: first row of this chunk
  uint32_t nbits, pair_off, k, chunk_rows_log;
  E8 scale;
};

__device__ inline uint32_t mod_signed(int64_t x) {
  int64_t r = x % int64_t(P);
  return uint32_t(r < 0 ? r + P : r);
}

// Builds T[i] = prod_{b<k} pairs[b][(i >> b) & 1] in LDS (k <= 12).
template <int NT = WG>
__device__ inline void build_table(E8* T, const E8 (*pairs)[2], int k) {
  if (threadIdx.x == 0) T[0] = v8_store(v8_zero()), T[0].c[0] = 1;
  __syncthreads();
  for (int b = 0; b < k; b++) {
    const int half = 1 << b;
    V8 p0 = v8_load(pairs[b][0]), p1 = v8_load(pairs[b][1]);
    for (int i = threadIdx.x; i < half; i += NT) {
      V8 t = v8_load(T[i]);
      T[i + half] = v8_store(v8_mul(t, p1));
      T[i] = v8_store(v8_mul(t, p0));
    }
    __syncthreads();
  }
}

// Signed products of BVA entries with canonical coefficients accumulate
// exactly in Acc before one reduction: i8 x u16 products (< 2^23) fit 32 bits
// over the short runs summed here; i16 x u31 products (< 2^46) need 64.
#ifdef QF_KB
using Acc = int64_t;
// acc + v c, exact: a small entry and a canonical coefficient (< 2^31) are
// both signed 32-bit values, so the product is a sign-extended 32 x 32 one and
// compiles to a single v_mad_i64_i32 (the zero-extended coefficient would take
// two unsigned 64-bit multiply-adds).
__device__ inline void acc_mac(Acc& acc, int32_t v, uint32_t c) { acc += int64_t(v) * int64_t(int32_t(c)); }
#else
using Acc = int32_t;
__device__ inline void acc_mac(Acc& acc, int32_t v, uint32_t c) { acc += v * int32_t(c); }
#endif

// Q (2 or 4) consecutive BVA entries as one aligned word (idx a multiple of Q).
// The caller guarantees idx + Q <= available; there is no branch, so a
// thread can issue several of these loads before it consumes any.
template <int Q>
struct RawVals {
  using T = typename std::conditional<Q * sizeof(Val) == 8, uint2,
            typename std::conditional<Q * sizeof(Val) == 4, uint32_t, uint16_t>::type>::type;
};
template <int Q>
__device__ inline typename RawVals<Q>::T load_raw(const Val* __restrict__ bva, uint64_t idx) {
  return *reinterpret_cast<const typename RawVals<Q>::T*>(bva + idx);
}
template <int Q>
__device__ inline void unpack_vals(const typename RawVals<Q>::T& w, int32_t (&v)[Q]) {
  if constexpr (Q * sizeof(Val) == 8) {
#pragma unroll
    for (int j = 0; j < Q; j++) v[j] = int16_t((j < 2 ? w.x : w.y) >> (16 * (j & 1)));
  } else if constexpr (Q * sizeof(Val) == 4) {
#pragma unroll
    for (int j = 0; j < Q; j++) v[j] = sizeof(Val) == 1 ? int32_t(int8_t(w >> (8 * j))) : int32_t(int16_t(w >> (16 * j)));
  } else {
#pragma unroll
    for (int j = 0; j < Q; j++) v[j] = int8_t(w >> (8 * j));
  }
}
// Guarded form: entries at or past `available` read as zero.
template <int Q>
__device__ inline void load_vals(const Val* __restrict__ bva, uint64_t idx, uint64_t available, int32_t (&v)[Q]) {
  if (idx + Q <= available) {
    unpack_vals<Q>(load_raw<Q>(bva, idx), v);
    return;
  }
#pragma unroll
  for (int j = 0; j < Q; j++) v[j] = idx + j < available ? int32_t(bva[idx + j]) : 0;
}

// One workgroup = one chunk of one projection task (2^chunk_rows_log rows).
template <int L>
__global__ __launch_bounds__(WG) void k_tensor_project(const Val* __restrict__ bva, uint64_t available,
                                                       const ProjTaskDev* __restrict__ tasks,
                                                       const E8 (*__restrict__ pairs)[2], E8* __restrict__ out) {
  constexpr int Q = L < 4 ? L : 4;       // lanes per thread
  constexpr int TPR = L / Q;             // threads per row
  constexpr int RPP = WG / TPR;          // rows per pass
  constexpr int U = 8;                   // row loads in flight per thread
  __shared__ E8 T[1 << PROJ_TAB_LOG];
  __shared__ E8 Chi[16];                 // scale x high-bit weights of each hi group
  const ProjTaskDev t = tasks[blockIdx.x];
  const E8(*tp)[2] = pairs + t.pair_off;
  build_table(T, tp, int(t.k));
  const int q = threadIdx.x % TPR;
  const int r0 = threadIdx.x / TPR;
  const uint32_t lo_rows = 1u << t.k;
  const uint32_t hi_count = 1u << (t.chunk_rows_log - t.k);  // <= 16: chunks hold <= 2^16 rows, k = 12 then
  if (threadIdx.x < hi_count) {  // the hi-group scales, one per thread
    const uint64_t row_hi = t.chunk_first + uint64_t(threadIdx.x) * lo_rows;
    V8 c = v8_load(t.scale);
    for (uint32_t b = t.k; b < t.nbits; b++) c = v8_mul(c, v8_load(tp[b][(row_hi >> b) & 1]));
    Chi[threadIdx.x] = v8_store(c);
  }
  __syncthreads();
  V8 acc[Q];
  for (int l = 0; l < Q; l++) acc[l] = v8_zero();
  for (uint32_t h = 0; h < hi_count; h++) {
    const uint64_t row_hi = t.chunk_first + uint64_t(h) * lo_rows;  // first row of this hi group
    const uint64_t first = t.base + row_hi * L + Q * q;
    Acc ai[Q][LANES];
    for (int l = 0; l < Q; l++)
      for (int k = 0; k < LANES; k++) ai[l][k] = 0;
    const bool in_range = t.base + (row_hi + lo_rows) * L <= available;  // uniform: all but the last groups
#ifndef QF_KB
    // F65521 fast path over row pairs (2p, 2p+1): lanes biased to u = v + 128
    // and paired across the two rows, weights paired likewise from the table,
    // one v_dot2_u32_u16 per (lane, coefficient) and row pair; ts holds each
    // coefficient's weight sum over the thread's rows and removes the bias.
    // Per thread a group holds at most 32 row pairs: au < 2^30, 128 ts < 2^29.1.
    if constexpr (Q == 4) {
      constexpr int UP = 4;  // row pairs in flight per thread
      if (in_range && lo_rows >= uint32_t(2 * RPP * UP)) {
        uint32_t au[Q][LANES], ts[LANES];
#pragma unroll
        for (int k = 0; k < LANES; k++) {
          ts[k] = 0;
#pragma unroll
          for (int l = 0; l < Q; l++) au[l][k] = 0;
        }
        for (uint32_t p0 = r0; 2 * p0 < lo_rows; p0 += RPP * UP) {
          uint32_t raw0[UP], raw1[UP];
#pragma unroll
          for (int u = 0; u < UP; u++) {
            const uint64_t row = 2 * uint64_t(p0 + u * RPP);
            raw0[u] = load_raw<Q>(bva, first + row * L);
            raw1[u] = load_raw<Q>(bva, first + (row + 1) * L);
          }
#pragma unroll
          for (int u = 0; u < UP; u++) {
            const uint32_t pr = p0 + u * RPP;
            const uint4 t0 = *reinterpret_cast<const uint4*>(&T[2 * pr]);
            const uint4 t1 = *reinterpret_cast<const uint4*>(&T[2 * pr + 1]);
            const uint32_t w0[4] = {t0.x, t0.y, t0.z, t0.w}, w1[4] = {t1.x, t1.y, t1.z, t1.w};
            uint32_t tp[LANES];  // (T[2p].c[k], T[2p+1].c[k]) as a u16 pair
#pragma unroll
            for (int j = 0; j < 4; j++) {
              tp[2 * j] = __builtin_amdgcn_perm(w1[j], w0[j], 0x05040100u);
              tp[2 * j + 1] = __builtin_amdgcn_perm(w1[j], w0[j], 0x07060302u);
            }
            const uint32_t x0 = raw0[u] ^ 0x80808080u, x1 = raw1[u] ^ 0x80808080u;
#pragma unroll
            for (int l = 0; l < Q; l++) {
              const uint32_t up = __builtin_amdgcn_perm(x1, x0, uint32_t(l) | 0x0c00u | (uint32_t(4 + l) << 16) | 0x0c000000u);
#pragma unroll
              for (int k = 0; k < LANES; k++) au[l][k] = dot2_u16(up, tp[k], au[l][k]);
            }
#pragma unroll
            for (int k = 0; k < LANES; k++) ts[k] = dot2_u16(0x00010001u, tp[k], ts[k]);
          }
        }
        const V8 c = v8_load(Chi[h]);
        constexpr uint32_t OFF = uint32_t(P) * 16384u;  // a multiple of P above 128 ts
#pragma unroll
        for (int l = 0; l < Q; l++) {
          V8 x;
#pragma unroll
          for (int k = 0; k < LANES; k++) {
            const uint32_t y = fold16(fold16(au[l][k] + (OFF - (ts[k] << 7))));
            const uint32_t z = y - P;
            x.c[k] = y < z ? y : z;
          }
          acc[l] = v8_add(acc[l], v8_mul(x, c));
        }
        continue;
      }
    }
#endif
    if (in_range && lo_rows >= uint32_t(RPP * U)) {
      // U unguarded row loads in flight per thread, then U rows of products
      for (uint32_t lo0 = r0; lo0 < lo_rows; lo0 += RPP * U) {
        typename RawVals<Q>::T raw[U];
#pragma unroll
        for (int u = 0; u < U; u++) raw[u] = load_raw<Q>(bva, first + uint64_t(lo0 + u * RPP) * L);
#pragma unroll
        for (int u = 0; u < U; u++) {
          int32_t v[Q];
          unpack_vals<Q>(raw[u], v);
          const E8 wt = T[lo0 + u * RPP];
#pragma unroll
          for (int l = 0; l < Q; l++)
#pragma unroll
            for (int k = 0; k < LANES; k++) acc_mac(ai[l][k], v[l], wt.c[k]);
        }
      }
    } else {
      for (uint32_t lo = r0; lo < lo_rows; lo += RPP) {
        int32_t v[Q];
        load_vals<Q>(bva, first + uint64_t(lo) * L, available, v);
        const E8 wt = T[lo];
#pragma unroll
        for (int l = 0; l < Q; l++)
#pragma unroll
          for (int k = 0; k < LANES; k++) acc_mac(ai[l][k], v[l], wt.c[k]);
      }
    }
    V8 c = v8_load(Chi[h]);
#pragma unroll
    for (int l = 0; l < Q; l++) {
      V8 x;
#pragma unroll
      for (int k = 0; k < LANES; k++) x.c[k] = mod_signed(ai[l][k]);
      acc[l] = v8_add(acc[l], v8_mul(x, c));
    }
  }
  __syncthreads();
  // reduce acc over threads sharing the same lane quad; reuse the table as scratch
  uint32_t* red = reinterpret_cast<uint32_t*>(T);  // [WG][LANES Q]
  constexpr int RW = LANES * Q;
  static_assert(sizeof(T) >= sizeof(uint32_t) * WG * RW, "reduction scratch exceeds the table");
  for (int l = 0; l < Q; l++)
    for (int k = 0; k < LANES; k++) red[threadIdx.x * RW + l * LANES + k] = acc[l].c[k];
  __syncthreads();
  for (int h = WG / 2; h >= TPR; h /= 2) {
    if (int(threadIdx.x) < h)
      for (int j = 0; j < RW; j++) red[threadIdx.x * RW + j] = madd(red[threadIdx.x * RW + j], red[(threadIdx.x + h) * RW + j]);
    __syncthreads();
  }
  if (int(threadIdx.x) < TPR) {
    for (int l = 0; l < Q; l++) {
      V8 x;
      for (int k = 0; k < LANES; k++) x.c[k] = red[threadIdx.x * RW + l * LANES + k];
      out[uint64_t(blockIdx.x) * L + Q * threadIdx.x + l] = v8_store(x);
    }
  }
}

// MFMA projection: a wave stages MF_ROWS 16-byte BVA records per round in LDS;
// ds_read_tr8_b64 on 8 consecutive 16-byte rows gives lane i of each 16-lane
// group byte i of every row.
constexpr int MF_ROWS = 256;            // records per wave and round: 8 blocks of 32, 4 KiB
constexpr int MF_WAVES = WG / 64;
typedef int32_t qf_i32x4 __attribute__((ext_vector_type(4)));

#ifndef QF_KB
// k_col_fold on the matrix cores (F65521, 4 | cols <= 128): out[r] = sum_col
// f[col] cw[col][r] as an i8 contraction over k = 8 col + j.
// - Digits of the data: coefficient x = xl + 256 xh with bytes xl, xh, fed as
//   xl - 128 and xh - 128 (one xor per four bytes); the 128 sum_k B[k][n] this
//   takes away is added back per output column (`bias`, from the host).
// - B[k][n]: digit of M_col[c][j] (column j of M_col = f[col] X^j), low for
//   n = c < 8, high for n = 8 + c, balanced as in the block fold; step s
//   covers columns 4 s .. 4 s + 3 (lane n + 16 q: column 4 s + q).
// - A[m][k]: row m of the 16-row group (lane i + 16 q: row i, column 4 s + q,
//   one 16-byte load gives both digit operands). |D| < 2^23 per digit pair.
// - value[c] = Ll + 256 (Lh + Hl) + 2^16 Hh with L/H the data digit and l/h
//   the weight digit, each reduced first; a wave takes 64 rows per iteration
//   and stores them through LDS, one 16-byte element per lane.
// Element (col, r) is cw[col R + r rstride]; grid.y runs independent folds
// `in_y` and `out_y` elements apart (the column projection's inner contraction
// over 32 consecutive rows uses R = 1, rstride = 32, one fold per column).
__global__ __launch_bounds__(WG) void k_col_fold_mfma(const E8* __restrict__ cw, uint64_t R, uint64_t rstride,
                                                      uint64_t n, uint32_t steps, const long* __restrict__ bops,
                                                      const int32_t* __restrict__ bias, E8* __restrict__ out,
                                                      uint64_t in_y, uint64_t out_y) {
  cw += uint64_t(blockIdx.y) * in_y;
  out += uint64_t(blockIdx.y) * out_y;
  __shared__ long lbop[32][64];
  __shared__ __align__(16) uint16_t tbuf[MF_WAVES][64][8];
  for (uint32_t i = threadIdx.x; i < steps * 64; i += WG) lbop[i / 64][i % 64] = bops[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nn = lane & 15, q = lane >> 4;
  const int32_t bl = bias[nn & 7], bh = bias[8 + (nn & 7)];  // this lane's two weight digits
  for (uint64_t base = uint64_t(blockIdx.x) * WG; base < n; base += uint64_t(gridDim.x) * WG) {
    const uint64_t row0 = base + uint64_t(wave) * 64;
    qf_i32x4 dl[4], dh[4];  // data low / high digit sums
#pragma unroll
    for (int g = 0; g < 4; g++) dl[g] = dh[g] = qf_i32x4{0, 0, 0, 0};
    for (uint32_t s = 0; s < steps; s++) {
      uint4 v[4];
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const uint64_t r = row0 + 16 * g + nn;
        v[g] = r < n ? *reinterpret_cast<const uint4*>(cw + uint64_t(4 * s + q) * R + r * rstride)
                     : make_uint4(0, 0, 0, 0);
      }
      const long bop = lbop[s][lane];
#pragma unroll
      for (int g = 0; g < 4; g++) {
        // coefficient j's bytes: low byte of u16 j at byte 2 j (mod 4) of word j / 2
        const uint32_t lo0 = __builtin_amdgcn_perm(v[g].y, v[g].x, 0x06040200u) ^ 0x80808080u;
        const uint32_t lo1 = __builtin_amdgcn_perm(v[g].w, v[g].z, 0x06040200u) ^ 0x80808080u;
        const uint32_t hi0 = __builtin_amdgcn_perm(v[g].y, v[g].x, 0x07050301u) ^ 0x80808080u;
        const uint32_t hi1 = __builtin_amdgcn_perm(v[g].w, v[g].z, 0x07050301u) ^ 0x80808080u;
        const long al = long(uint64_t(lo0) | (uint64_t(lo1) << 32)), ah = long(uint64_t(hi0) | (uint64_t(hi1) << 32));
        dl[g] = __builtin_amdgcn_mfma_i32_16x16x32_i8(al, bop, dl[g], 0, 0, 0);
        dh[g] = __builtin_amdgcn_mfma_i32_16x16x32_i8(ah, bop, dh[g], 0, 0, 0);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the previous iteration's reads of tbuf are done
    const int32_t bown = nn < 8 ? bl : bh;  // bias of this lane's weight digit
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        // this lane's sums with 128 sum_k B added back, reduced mod p (nonnegative through the 16384 p bias)
        const uint32_t L = fold16(fold16(uint32_t(dl[g][j] + bown + 16384 * int32_t(P))));
        const uint32_t H = fold16(fold16(uint32_t(dh[g][j] + bown + 16384 * int32_t(P))));
        const uint32_t Lo = __shfl_xor(L, 8, 64), Ho = __shfl_xor(H, 8, 64);
        if ((j < 2) == (nn < 8)) {
          // Ll, Lh, Hl, Hh: data digit L/H times weight digit l/h
          const uint32_t Ll = nn < 8 ? L : Lo, Lh = nn < 8 ? Lo : L, Hl = nn < 8 ? H : Ho, Hh = nn < 8 ? Ho : H;
          const uint32_t mid = fold16(Lh + Hl);                          // < 2^17, folded below 2^16 + 2^5
          const uint32_t x = fold16(fold16(Ll + 256u * mid + 15u * Hh));  // 2^16 = 15 mod p
          const uint32_t y = x - P;
          tbuf[wave][16 * g + 4 * q + j][nn & 7] = uint16_t(x < y ? x : y);
        }
      }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const uint4 raw = *reinterpret_cast<const uint4*>(&tbuf[wave][lane][0]);
    if (row0 + lane < n) {
      E8 e;
      *reinterpret_cast<uint4*>(&e) = raw;
      out[row0 + lane] = e;
    }
  }
}
#endif
typedef int32_t qf_i32x2 __attribute__((ext_vector_type(2)));
__device__ inline long tr8_b64(const uint8_t* p) {
  const qf_i32x2 v = __builtin_amdgcn_ds_read_tr8_b64_v2i32(
      (__attribute__((address_space(3))) qf_i32x2*)(const_cast<uint8_t*>(p)));
  return (long(uint32_t(v.y)) << 32) | long(uint32_t(v.x));
}
#ifndef QF_KB
// Projection chunks of 8-, 16-, 32- and 64-lane rows (the batch lane
// projections, and the EO projections at stride 2s) as an exact i8 MFMA
// contraction of the BVA against the weight table, over 16-byte records: two
// consecutive 8-lane rows 2r', 2r'+1, whose weights pairs[0][b] T'[r'] share
// the table T' over the row bits above bit 0; one 16-lane row; or one of the
// NS = L / 16 records of a longer row, each accumulated as its own stream
// against the row's weight (lanes 16 s .. 16 s + 15 of the row in stream s).
// - Digits: a canonical table coefficient t is written as t' = t or t - p in
//   [-32881, 32639]; h = (t' + 128) >> 8 and l = t' - 256 h then both lie in
//   [-128, 127] and t' = 256 h + l. Record r of the digit table holds the low
//   digits of coefficients 0..7 in bytes 0..7 and the high digits in 8..15.
// - v_mfma_i32_16x16x32_i8 adds, per 32 records, the products entry x digit
//   (|.| <= 2^14) into D[record byte][digit byte]; a hi group of at most 2^11
//   records stays below 2^25 per wave and 2^27 per workgroup, exact in i32.
// - Operands: ds_read_tr8_b64 on 8 consecutive 16-byte records in LDS gives
//   lane i of each 16-lane group byte i of every record, which is the
//   K-contiguous layout of both operands: BVA records (staged through LDS by
//   the wave that loaded them, 4 KiB per round) for A, digit records for B.
// - Each hi group's per-wave sums are combined through LDS, reduced mod p
//   (low + 256 high), the two rows of an 8-lane record joined with pairs[0],
//   multiplied by the group's scale and accumulated by one thread per lane.
// The chunk's rows start 16-byte aligned and a hi group holds at least
// MF_ROWS records per wave (the host routes other chunks to k_tensor_project).
template <int L>
__global__ __launch_bounds__(WG) void k_tensor_project_mfma(const Val* __restrict__ bva, uint64_t available,
                                                            const ProjTaskDev* __restrict__ tasks,
                                                            const E8 (*__restrict__ pairs)[2], E8* __restrict__ out) {
  static_assert(L == 64 || L == 32 || L == 16 || L == 8, "16-byte records of two rows, one row, or part of a row");
  constexpr int RBITS = L == 8 ? 1 : 0;            // row bits inside a record
  constexpr int NS = L <= 16 ? 1 : L / 16;         // records (streams) per row
  constexpr int RT = MF_ROWS / NS;                 // table entries per wave and round
  __shared__ E8 T[1 << MF_TAB_LOG];                // table entry weights, then their digit bytes
  __shared__ uint4 stage[MF_WAVES][MF_ROWS];       // a round of BVA records per wave; the combine reuses it
  __shared__ uint32_t pv[16 * NS][8];              // a group's record-byte sums mod p, per stream
  __shared__ E8 Chi[32];                           // scale x high-bit weights of each hi group
  const ProjTaskDev t = tasks[blockIdx.x];
  const E8(*tp)[2] = pairs + t.pair_off;
  build_table(T, tp + RBITS, int(t.k) - RBITS);
  const uint32_t lo_rows = 1u << t.k, lo_recs = lo_rows >> RBITS;  // table entries of a hi group
  const uint32_t hi_count = 1u << (t.chunk_rows_log - t.k);  // <= 32
  if (threadIdx.x < hi_count) {
    const uint64_t row_hi = t.chunk_first + uint64_t(threadIdx.x) * lo_rows;
    V8 c = v8_load(t.scale);
    for (uint32_t b = t.k; b < t.nbits; b++) c = v8_mul(c, v8_load(tp[b][(row_hi >> b) & 1]));
    Chi[threadIdx.x] = v8_store(c);
  }
  for (uint32_t r = threadIdx.x; r < lo_recs; r += WG) {  // digits, in place
    const E8 e = T[r];
    uint32_t d[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int32_t tv = e.c[c] <= 32639 ? int32_t(e.c[c]) : int32_t(e.c[c]) - int32_t(P);
      const int32_t hd = (tv + 128) >> 8, ld = tv - 256 * hd;
      d[c / 4] |= uint32_t(uint8_t(ld)) << (8 * (c % 4));
      d[2 + c / 4] |= uint32_t(uint8_t(hd)) << (8 * (c % 4));
    }
    reinterpret_cast<uint4*>(T)[r] = make_uint4(d[0], d[1], d[2], d[3]);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint8_t* tb = reinterpret_cast<const uint8_t*>(T);
  uint8_t* st = reinterpret_cast<uint8_t*>(stage[wave]);
  const uint32_t wrecs = lo_recs / MF_WAVES, rounds = wrecs / RT, total = hi_count * rounds;
  // round j: hi group j / rounds, this wave's table entries [r0, r0 + RT) of it (MF_ROWS records)
  auto fetch = [&](uint32_t j, uint4 (&w)[4]) {
    const uint32_t r0 = wave * wrecs + (j % rounds) * RT;
    const uint64_t b0 = t.base + (t.chunk_first + uint64_t(j / rounds) * lo_rows) * L + uint64_t(r0) * 16 * NS;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint64_t b = b0 + 1024 * u + 16 * lane;
      if (b + 16 <= available) {
        w[u] = *reinterpret_cast<const uint4*>(bva + b);
      } else {  // entries at or past the active end read as zero
        uint32_t x[4] = {0, 0, 0, 0};
        for (int q = 0; q < 16; q++)
          if (b + q < available) x[q / 4] |= uint32_t(uint8_t(bva[b + q])) << (8 * (q % 4));
        w[u] = make_uint4(x[0], x[1], x[2], x[3]);
      }
    }
  };
  V8 acc = v8_zero();  // threads 0..L-1: lane threadIdx.x of the chunk
  qf_i32x4 dsum[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) dsum[s] = qf_i32x4{0, 0, 0, 0};
  uint4 cur[4], nxt[4];
  fetch(0, cur);
  for (uint32_t j = 0; j < total; j++) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the previous round's reads of st are done
#pragma unroll
    for (int u = 0; u < 4; u++) {  // record q = 64 u + lane goes to stream q % NS
      const uint32_t q = 64 * u + lane;
      *reinterpret_cast<uint4*>(st + (q % NS) * (16 * RT) + 16 * (q / NS)) = cur[u];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (j + 1 < total) fetch(j + 1, nxt);
    const uint32_t r0 = wave * wrecs + (j % rounds) * RT;
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int b = 0; b < RT / 32; b++) {
        const long a = tr8_b64(st + s * (16 * RT) + 512 * b + 8 * lane);
        const long w = tr8_b64(tb + 16 * (r0 + 32 * b) + 8 * lane);
        dsum[s] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a, w, dsum[s], 0, 0, 0);
      }
#pragma unroll
    for (int u = 0; u < 4; u++) cur[u] = nxt[u];
    if ((j + 1) % rounds) continue;
    // end of hi group j / rounds: D[m][n] sits in lane n + 16 (m / 4), element m % 4
    __syncthreads();  // every wave is done with its staging buffer
    int32_t* red = reinterpret_cast<int32_t*>(&stage[0][0]);  // [stream][wave][16 m][16 n]
#pragma unroll
    for (int s = 0; s < NS; s++) {
#pragma unroll
      for (int i = 0; i < 4; i++)
        red[(s * MF_WAVES + wave) * 256 + (4 * (lane / 16) + i) * 16 + lane % 16] = dsum[s][i];
      dsum[s] = qf_i32x4{0, 0, 0, 0};
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 128 * NS; i += WG) {  // stream s, record byte m, coefficient c
      const int s = i / 128, m = (i / 8) % 16, c = i % 8;
      const int32_t* rs = red + s * MF_WAVES * 256;
      int64_t lo = 0, hi = 0;
      for (int w2 = 0; w2 < MF_WAVES; w2++) {
        lo += rs[w2 * 256 + m * 16 + c];
        hi += rs[w2 * 256 + m * 16 + 8 + c];
      }
      pv[16 * s + m][c] = mod_signed(lo + 256 * hi);
    }
    __syncthreads();
    if (threadIdx.x < L) {
      V8 x[2];
#pragma unroll
      for (int p = 0; p <= RBITS; p++)
#pragma unroll
        for (int c = 0; c < 8; c++) x[p].c[c] = pv[p * L + threadIdx.x][c];
      V8 y = x[0];
      if (RBITS) y = v8_add(v8_mul(x[0], v8_load(tp[0][0])), v8_mul(x[1], v8_load(tp[0][1])));
      acc = v8_add(acc, v8_mul(y, v8_load(Chi[j / rounds])));
    }
    __syncthreads();  // the staging buffers are free again
  }
  if (threadIdx.x < L) out[uint64_t(blockIdx.x) * L + threadIdx.x] = v8_store(acc);
}
#else
// KoalaBear form of the MFMA projection, over 16-byte records of 8 i16
// entries: one of the NS = L / 8 records of a 16- or 32-lane row (each
// accumulates as its own stream against the row's weight), one 8-lane row, or
// two consecutive 4-lane rows joined with pairs[0] as in the F65521 form.
// - Entries: an i16 entry v is 256 vh + (vl - 128) + 128 with its high byte vh
//   and its offset low byte vl - 128 both in [-128, 127]; staging flips bit 7
//   of each low byte, which gives that byte, and the 128 sum_r t_r this leaves
//   out of a hi group is added back from ssum, the table's signed coefficient
//   sums (the same for every hi group; zero entries past the active end cancel
//   with it).
// - Digits: a canonical coefficient t (< 2^31) is written as t' = t or t - p in
//   (-2^30, 2^30) and split into four balanced digits d_i in [-128, 127] with
//   t' = sum_i 256^i d_i; record r of the digit table holds digit i of
//   coefficient c in byte 4 i + c.
// - D[record byte][digit byte] sums products of at most 2^14 over at most 2^9
//   records per wave and stream, 2^25 per workgroup: exact in i32; the
//   recombination sum_i 256^i (256 D_high + D_low) + 128 ssum stays below 2^58.
template <int L>
__global__ __launch_bounds__(WG) void k_tensor_project_mfma(const Val* __restrict__ bva, uint64_t available,
                                                            const ProjTaskDev* __restrict__ tasks,
                                                            const E8 (*__restrict__ pairs)[2], E8* __restrict__ out) {
  static_assert(L == 32 || L == 16 || L == 8 || L == 4, "records of part of a row, one row or two rows");
  constexpr int RBITS = L == 4 ? 1 : 0;            // row bits inside a record
  constexpr int NS = L >= 16 ? L / 8 : 1;          // records (streams) per row
  constexpr int RT = MF_ROWS / NS;                 // table entries per wave and round
  __shared__ E8 T[1 << MF_TAB_LOG];                // table entry weights, then their digit bytes
  __shared__ uint4 stage[MF_WAVES][MF_ROWS];       // a round of BVA records per wave; the combine reuses it
  __shared__ uint32_t pv[8 * NS < 16 ? 16 : 8 * NS][4];  // a group's record-entry sums mod p
  __shared__ E8 Chi[32];                           // scale x high-bit weights of each hi group
  __shared__ unsigned long long ssum[4];           // sum over the table of t', two's complement
  const ProjTaskDev t = tasks[blockIdx.x];
  const E8(*tp)[2] = pairs + t.pair_off;
  if (threadIdx.x < 4) ssum[threadIdx.x] = 0;
  build_table(T, tp + RBITS, int(t.k) - RBITS);   // its barriers order the clear above
  const uint32_t lo_rows = 1u << t.k, te = lo_rows >> RBITS;
  const uint32_t hi_count = 1u << (t.chunk_rows_log - t.k);  // <= 32
  if (threadIdx.x < hi_count) {
    const uint64_t row_hi = t.chunk_first + uint64_t(threadIdx.x) * lo_rows;
    V8 c = v8_load(t.scale);
    for (uint32_t b = t.k; b < t.nbits; b++) c = v8_mul(c, v8_load(tp[b][(row_hi >> b) & 1]));
    Chi[threadIdx.x] = v8_store(c);
  }
  int64_t ls[4] = {0, 0, 0, 0};
  for (uint32_t r = threadIdx.x; r < te; r += WG) {  // digits, in place
    const E8 e = T[r];
    uint32_t d[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int32_t x = e.c[c] <= P / 2 ? int32_t(e.c[c]) : int32_t(e.c[c] - P);
      ls[c] += x;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int32_t dd = ((x + 128) & 255) - 128;
        d[i] |= uint32_t(uint8_t(dd)) << (8 * c);
        x = (x - dd) >> 8;
      }
    }
    reinterpret_cast<uint4*>(T)[r] = make_uint4(d[0], d[1], d[2], d[3]);
  }
#pragma unroll
  for (int c = 0; c < 4; c++) atomicAdd(&ssum[c], (unsigned long long)ls[c]);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint8_t* tb = reinterpret_cast<const uint8_t*>(T);
  uint8_t* st = reinterpret_cast<uint8_t*>(stage[wave]);
  const uint32_t wte = te / MF_WAVES, rounds = wte / RT, total = hi_count * rounds;
  // round j: hi group j / rounds, this wave's table entries [r0, r0 + RT) of it
  auto fetch = [&](uint32_t j, uint4 (&w)[4]) {
    const uint32_t r0 = wave * wte + (j % rounds) * RT;
    const uint64_t e0 = t.base + (t.chunk_first + uint64_t(j / rounds) * lo_rows) * L + uint64_t(r0) * (8 * NS);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint64_t e = e0 + 8 * (64 * u + lane);
      if (e + 8 <= available) {
        w[u] = *reinterpret_cast<const uint4*>(bva + e);
      } else {  // entries at or past the active end read as zero
        uint32_t x[4] = {0, 0, 0, 0};
        for (int q = 0; q < 8; q++)
          if (e + q < available) x[q / 2] |= uint32_t(uint16_t(bva[e + q])) << (16 * (q % 2));
        w[u] = make_uint4(x[0], x[1], x[2], x[3]);
      }
    }
  };
  V8 acc = v8_zero();  // threads 0..L-1: lane threadIdx.x of the chunk
  qf_i32x4 dsum[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) dsum[s] = qf_i32x4{0, 0, 0, 0};
  uint4 cur[4], nxt[4];
  fetch(0, cur);
  for (uint32_t j = 0; j < total; j++) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the previous round's reads of st are done
#pragma unroll
    for (int u = 0; u < 4; u++) {  // record q = 64 u + lane goes to stream q % NS
      const uint32_t q = 64 * u + lane;
      const uint32_t at = (q % NS) * (16 * RT) + 16 * (q / NS);
      const uint32_t f = 0x00800080u;  // bit 7 of each low byte
      *reinterpret_cast<uint4*>(st + at) = make_uint4(cur[u].x ^ f, cur[u].y ^ f, cur[u].z ^ f, cur[u].w ^ f);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (j + 1 < total) fetch(j + 1, nxt);
    const uint32_t r0 = wave * wte + (j % rounds) * RT;
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int b = 0; b < RT / 32; b++) {
        const long a = tr8_b64(st + s * (16 * RT) + 512 * b + 8 * lane);
        const long w = tr8_b64(tb + 16 * (r0 + 32 * b) + 8 * lane);
        dsum[s] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a, w, dsum[s], 0, 0, 0);
      }
#pragma unroll
    for (int u = 0; u < 4; u++) cur[u] = nxt[u];
    if ((j + 1) % rounds) continue;
    // end of hi group j / rounds: D[m][n] sits in lane n + 16 (m / 4), element m % 4
    __syncthreads();  // every wave is done with its staging buffer
    int32_t* red = reinterpret_cast<int32_t*>(&stage[0][0]);  // [stream][wave][16 m][16 n]
#pragma unroll
    for (int s = 0; s < NS; s++) {
#pragma unroll
      for (int i = 0; i < 4; i++)
        red[(s * MF_WAVES + wave) * 256 + (4 * (lane / 16) + i) * 16 + lane % 16] = dsum[s][i];
      dsum[s] = qf_i32x4{0, 0, 0, 0};
    }
    __syncthreads();
    if (threadIdx.x < 32 * NS) {  // stream s, record entry e, coefficient c
      const int s = threadIdx.x / 32, e = (threadIdx.x / 4) % 8, c = threadIdx.x % 4;
      const int32_t* rs = red + s * MF_WAVES * 256;
      int64_t v = 0;
      for (int i = 3; i >= 0; i--) {
        int64_t lo = 0, hi = 0;
        for (int w2 = 0; w2 < MF_WAVES; w2++) {
          lo += rs[w2 * 256 + (2 * e) * 16 + 4 * i + c];
          hi += rs[w2 * 256 + (2 * e + 1) * 16 + 4 * i + c];
        }
        v = 256 * v + 256 * hi + lo;
      }
      pv[s * 8 + e][c] = mod_signed(v + 128 * int64_t(ssum[c]));
    }
    __syncthreads();
    if (threadIdx.x < L) {
      V8 x[2];
#pragma unroll
      for (int p = 0; p <= RBITS; p++)
#pragma unroll
        for (int c = 0; c < 4; c++) x[p].c[c] = pv[p * L + threadIdx.x][c];
      V8 y = x[0];
      if (RBITS) y = v8_add(v8_mul(x[0], v8_load(tp[0][0])), v8_mul(x[1], v8_load(tp[0][1])));
      acc = v8_add(acc, v8_mul(y, v8_load(Chi[j / rounds])));
    }
    __syncthreads();  // the staging buffers are free again
  }
  if (threadIdx.x < L) out[uint64_t(blockIdx.x) * L + threadIdx.x] = v8_store(acc);
}
#endif

// Witness rows row0 .. row0 + NR - 1 (eo.rs:119-132): w[row] = sum_lane
// v[row S + lane] pinv[lane], from S consecutive BVA entries per row (zero at
// or past `available`). On the fast path all loads of the NR rows are issued
// before any product.
template <int S, int NR>
__device__ inline void witness_rows(const Val* __restrict__ bva, uint64_t row0, uint64_t available,
                                    const E8 (&pinv)[S], V8 (&w)[NR]) {
  constexpr int Q = S < 4 ? S : 4;
  Acc acc[NR][LANES];
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int k = 0; k < LANES; k++) acc[r][k] = 0;
  const uint64_t first = row0 * S;
  if (first + NR * S <= available) {
    typename RawVals<Q>::T raw[NR][S / Q];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int g = 0; g < S / Q; g++) raw[r][g] = load_raw<Q>(bva, first + r * S + Q * g);
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int g = 0; g < S / Q; g++) {
        int32_t v[Q];
        unpack_vals<Q>(raw[r][g], v);
#pragma unroll
        for (int j = 0; j < Q; j++) {
          const E8 p = pinv[Q * g + j];
#pragma unroll
          for (int k = 0; k < LANES; k++) acc_mac(acc[r][k], v[j], p.c[k]);
        }
      }
  } else {
    for (int r = 0; r < NR; r++)
      for (int g = 0; g < S / Q; g++) {
        int32_t v[Q];
        load_vals<Q>(bva, first + r * S + Q * g, available, v);
#pragma unroll
        for (int j = 0; j < Q; j++) {
          const E8 p = pinv[Q * g + j];
#pragma unroll
          for (int k = 0; k < LANES; k++) acc_mac(acc[r][k], v[j], p.c[k]);
        }
      }
  }
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int k = 0; k < LANES; k++) w[r].c[k] = mod_signed(acc[r][k]);
}

template <int S>
__global__ __launch_bounds__(WG) void k_eo_witness(const Val* __restrict__ bva, uint64_t available, uint64_t n_rows,
                                                   const E8* __restrict__ pinv_g, E8* __restrict__ w) {
  __shared__ E8 pinv[S];
  for (int i = threadIdx.x; i < S; i += WG) pinv[i] = pinv_g[i];
  __syncthreads();
  for (uint64_t row = uint64_t(blockIdx.x) * WG + threadIdx.x; row < n_rows; row += uint64_t(gridDim.x) * WG) {
    V8 o[1];
    witness_rows<S, 1>(bva, row, available, pinv, o);
    w[row] = v8_store(o[0]);
  }
}

#ifndef QF_KB
// Lazy sums of F65521 extension products. A product a b splits a into 8-bit
// limbs and b into the wrap-folded sequence D of v8_mul_dot2; each product
// adds at most 4 x 2 x 255 x 65520 < 2^27.1 to every low-limb (sl) and
// high-limb (sh) coefficient sum, so a sum takes 31 products before it must
// be folded into a reduced element.
struct LazySum {
  uint32_t sl[8], sh[8];
};
__device__ inline void lazy_zero(LazySum& s) {
#pragma unroll
  for (int i = 0; i < 8; i++) s.sl[i] = s.sh[i] = 0;
}
__device__ inline void fold_seq(const V8& b, uint32_t (&D)[15]) {  // D[1..14]: u16 pairs of E (v8_mul_dot2)
  uint32_t E[15];
#pragma unroll
  for (int i = 0; i < 8; i++) E[7 + i] = b.c[i];
#pragma unroll
  for (int j = 1; j < 8; j++) {
    const uint32_t u = fold16(__umul24(b.c[j], 17u));
    const uint32_t w = u - P;
    E[j - 1] = u < w ? u : w;
  }
#pragma unroll
  for (int m = 1; m < 15; m++) D[m] = E[m] | (E[m - 1] << 16);
}
__device__ inline void limb_split(const V8& a, uint32_t (&lo)[4], uint32_t (&hi)[4]) {
#pragma unroll
  for (int jp = 0; jp < 4; jp++) {
    const uint32_t pr = a.c[2 * jp] | (a.c[2 * jp + 1] << 16);
    lo[jp] = pr & 0x00ff00ffu;
    hi[jp] = (pr >> 8) & 0x00ff00ffu;
  }
}
__device__ inline void lazy_mac(LazySum& s, const uint32_t (&lo)[4], const uint32_t (&hi)[4], const uint32_t (&D)[15]) {
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int jp = 0; jp < 4; jp++) {
      s.sl[i] = dot2_u16(lo[jp], D[7 + i - 2 * jp], s.sl[i]);
      s.sh[i] = dot2_u16(hi[jp], D[7 + i - 2 * jp], s.sh[i]);
    }
}
__device__ inline V8 lazy_value(const LazySum& s) {  // sl, sh < 2^32
  V8 r;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint32_t x = (fold16(fold16(s.sh[i])) << 8) + fold16(s.sl[i]);  // < 2^24.1 + 2^20.1
    x = fold16(fold16(x));                                          // < 2^16 + 15
    const uint32_t w = x - P;
    r.c[i] = x < w ? x : w;
  }
  return r;
}

#endif

// ---------------------------------------------------------------- EO rounds with the query on the fly
// Chunk ch of a level covers the output pairs [ch 2^OTF_LCP, (ch + 1) 2^OTF_LCP);
// pair i WG + t of a chunk belongs to thread t, and workgroup b takes chunks
// b, b + grid, ... (all workgroups stream through one window of the vector). A chunk inside a chunked term
// (OtfTermDev) carries the term and g0, g1 times the term's factors of the
// chunk-index bits; the term's tables hold the factors of the thread bits
// (tlo, WG entries) and of the i bits (thi, OTF_PPT entries), so the query
// value of pair m is tlo[t] thi[i] times the chunk's g.
struct OtfChunk {
  int32_t term;  // -1: no chunked term covers the chunk
  uint32_t pad[3];
  E8 g0, g1;
};
static_assert(sizeof(OtfChunk) == 48 && sizeof(OtfTermDev) == 48 && sizeof(OtfPieceDev) == 64 &&
                  sizeof(OtfItemDev) == 32, "device records are whole extension elements");
static_assert(WG == 256 && OTF_PPT == 8 && OTF_LCP == 11, "pair index bits: 8 thread bits, 3 i bits");

__device__ inline V8 v8_one() {
  V8 r = v8_zero();
  r.c[0] = 1;
  return r;
}
// prod_b f[base + b][(x >> b) & 1], b < nbits
__device__ inline V8 otf_tensor(const E8 (*__restrict__ f)[2], uint32_t base, uint32_t nbits, uint64_t x) {
  V8 g = v8_one();
  for (uint32_t b = 0; b < nbits; b++) g = v8_mul(g, v8_load(f[base + b][(x >> b) & 1]));
  return g;
}

// The level's chunk records (threads below `chunks`, each locating its chunk's
// term by bisection over the sorted terms) and the chunked terms' tables: tlo
// over the pair-index bits 0 .. LO-1 (2^LO entries per term), thi over bits
// LO .. OTF_LCP-1 (LO = 8: the fold kernels' thread and i bits; LO = 7: the
// MFMA first fold's wave/lane and iteration bits).
template <int LO>
__global__ __launch_bounds__(WG) void k_otf_desc(const OtfTermDev* __restrict__ terms, uint32_t nterms,
                                                 const E8 (*__restrict__ f)[2], uint64_t chunks,
                                                 OtfChunk* __restrict__ desc, E8* __restrict__ tlo,
                                                 E8* __restrict__ thi) {
  constexpr uint32_t NLO = 1u << LO, NHI = 1u << (OTF_LCP - LO);
  const uint64_t gid = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (gid < chunks) {
    const uint64_t p = gid << OTF_LCP;
    uint32_t lo = 0, hi = nterms;  // terms [0, lo) start at or before p
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (terms[mid].pair0 <= p) lo = mid + 1;
      else hi = mid;
    }
    OtfChunk d{};
    d.term = -1;
    if (lo > 0) {
      const OtfTermDev& T = terms[lo - 1];
      if (p - T.pair0 < (uint64_t(1) << T.lpairs)) {
        const V8 g = otf_tensor(f, T.fbase + OTF_LCP, T.lpairs - OTF_LCP, (p - T.pair0) >> OTF_LCP);
        d.term = int32_t(lo - 1);
        d.g0 = v8_store(v8_mul(g, v8_load(T.g0)));
        d.g1 = v8_store(v8_mul(g, v8_load(T.g1)));
      }
    }
    desc[gid] = d;
    return;
  }
  uint64_t e = gid - chunks;
  if (e < uint64_t(nterms) * NLO) {
    tlo[e] = v8_store(otf_tensor(f, terms[e / NLO].fbase, LO, e % NLO));
    return;
  }
  e -= uint64_t(nterms) * NLO;
  if (e < uint64_t(nterms) * NHI) thi[e] = v8_store(otf_tensor(f, terms[e / NHI].fbase + LO, OTF_LCP - LO, e % NHI));
}

// Per-chunk sums sum_i thi[i] y_i over one thread's OTF_PPT pairs: lazy dot2
// products for F65521 (8 < 31 products between folds), plain ones for KoalaBear.
#ifndef QF_KB
struct OtfAcc {
  LazySum l;
};
__device__ inline void otf_zero(OtfAcc& a) { lazy_zero(a.l); }
__device__ inline void otf_mac2(OtfAcc& a0, OtfAcc& a1, const V8& y0, const V8& y1, const V8& th) {
  uint32_t D[15], lo[4], hi[4];
  fold_seq(th, D);
  limb_split(y0, lo, hi);
  lazy_mac(a0.l, lo, hi, D);
  limb_split(y1, lo, hi);
  lazy_mac(a1.l, lo, hi, D);
}
__device__ inline V8 otf_value(const OtfAcc& a) { return lazy_value(a.l); }
#else
struct OtfAcc {
  V8 v;
};
__device__ inline void otf_zero(OtfAcc& a) { a.v = v8_zero(); }
__device__ inline void otf_mac2(OtfAcc& a0, OtfAcc& a1, const V8& y0, const V8& y1, const V8& th) {
  const V8 tm = v8_to_mont(th);
  a0.v = v8_add(a0.v, v8_mul_mont(y0, tm));
  a1.v = v8_add(a1.v, v8_mul_mont(y1, tm));
}
__device__ inline V8 otf_value(const OtfAcc& a) { return a.v; }
#endif

// A chunk's part of the cross terms: s0 += a0 tlo[t] g0, s1 += a1 tlo[t] g1.
__device__ inline void otf_chunk_sum(const OtfChunk& d, const E8* __restrict__ tlo, int32_t& cur, V8& tl,
                                     const OtfAcc& a0, const OtfAcc& a1, V8 (&s)[2]) {
  if (d.term != cur) {
    tl = v8_load(tlo[uint64_t(d.term) * WG + threadIdx.x]);
    cur = d.term;
  }
  s[0] = v8_add(s[0], v8_mul(otf_value(a0), v8_mul(tl, v8_load(d.g0))));
  s[1] = v8_add(s[1], v8_mul(otf_value(a1), v8_mul(tl, v8_load(d.g1))));
}

// One binary EO fold (eo.rs:57-79: w'[i] = w[2i] + w[2i+1] r^-1) of a length-2h
// witness into ow (h), with the chunked terms' part of the next round's cross
// terms over ow (the query is never read).
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void k_otf_fold(
    const E8* __restrict__ w, E8* __restrict__ ow, uint64_t chunks, E8 ri, const OtfChunk* __restrict__ desc,
    const E8* __restrict__ tlo, const E8* __restrict__ thi, uint32_t* __restrict__ partials) {
  const V8 vri = v8_encode_form(v8_load(ri));
  V8 s[2] = {v8_zero(), v8_zero()};
  int32_t cur = -1;
  V8 tl = v8_zero();
  // the next pair's four inputs load while this one computes
  E8 nx[4];
  auto fetch = [&](uint64_t ch, int i) {
    const uint64_t m = (ch << OTF_LCP) + uint64_t(i) * WG + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) nx[k] = w[4 * m + k];
  };
  if (blockIdx.x < chunks) fetch(blockIdx.x, 0);
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {  // the workgroups sweep the level together
    const OtfChunk d = desc[ch];
    OtfAcc a0, a1;
    otf_zero(a0);
    otf_zero(a1);
    for (int i = 0; i < OTF_PPT; i++) {
      const uint64_t m = (ch << OTF_LCP) + uint64_t(i) * WG + threadIdx.x;
      const E8 x0 = nx[0], x1 = nx[1], x2 = nx[2], x3 = nx[3];
      if (i + 1 < OTF_PPT) fetch(ch, i + 1);
      else if (ch + gridDim.x < chunks) fetch(ch + gridDim.x, 0);
      const V8 y0 = v8_add(v8_load(x0), v8_mul_enc(v8_load(x1), vri));
      const V8 y1 = v8_add(v8_load(x2), v8_mul_enc(v8_load(x3), vri));
      ow[2 * m] = v8_store(y0);
      ow[2 * m + 1] = v8_store(y1);
      if (d.term >= 0) otf_mac2(a0, a1, y0, y1, v8_load(thi[uint64_t(d.term) * OTF_PPT + i]));
    }
    if (d.term >= 0) otf_chunk_sum(d, tlo, cur, tl, a0, a1, s);
  }
  block_sums<2>(s, partials);
}

// ---------------------------------------------------------------- EO first round from the BVA
// The first binary EO round folds witness rows 2i and 2i + 1 of stride s
// (eo.rs:119-132, 140-163): w'[i] = w[2i] + w[2i+1] r^-1 is itself a witness
// row of stride S2 = 2s over entries [2 i s, 2 i s + 2s), with the weights
// pinv2[L] = c^-L (L < s) and r^-1 c^-(L-s) (L >= s). The pass therefore
// reads the BVA instead of a materialized witness, and the witness of stride s
// is never written; like k_otf_fold it writes w' and returns the chunked
// terms' part of the next round's cross terms over it.
#ifndef QF_KB
// F65521 form: the witness rows are biased-byte v_dot2_u32_u16 sums against u16
// pairs of pinv2 coefficients (entries biased to u = v + 128, the bias removed
// by a per-coefficient constant), four BVA words at a time.
template <int S2>
__global__ __launch_bounds__(WG) void k_otf_first_fold(const Val* __restrict__ bva, uint64_t available,
                                                       uint64_t chunks, const E8* __restrict__ pinv_g,
                                                       E8* __restrict__ ow, const OtfChunk* __restrict__ desc,
                                                       const E8* __restrict__ tlo, const E8* __restrict__ thi,
                                                       uint32_t* __restrict__ partials) {
  static_assert(S2 >= 4, "rows of at least one BVA word");
  constexpr int G = S2 / 4;   // BVA words (4 lanes) per witness row
  constexpr int NP = S2 / 2;  // lane pairs per row
  __shared__ E8 pinv[S2];
  __shared__ uint32_t PU[8][NP];
  __shared__ uint32_t negc[8];
  for (int i = threadIdx.x; i < S2; i += WG) pinv[i] = pinv_g[i];
  __syncthreads();
  for (int i = threadIdx.x; i < 8 * NP; i += WG) {
    const int k = i / NP, lp = i % NP;
    PU[k][lp] = uint32_t(pinv[2 * lp].c[k]) | (uint32_t(pinv[2 * lp + 1].c[k]) << 16);
  }
  if (threadIdx.x < 8) {
    uint32_t t = 0;
    for (int l = 0; l < S2; l++) t = madd(t, pinv[l].c[threadIdx.x]);
    negc[threadIdx.x] = msub(0, uint32_t((uint64_t(t) * 128) % P));
  }
  __syncthreads();
  // one witness row of S2 entries, all below `available`: each dot2 adds less
  // than 2^25, so the NP <= 64 of a row and the bias correction stay below 2^32
  // (the BVA is 16-byte aligned: rows of 16 entries or more load as 16-byte vectors)
  auto witness = [&](uint64_t row, V8& out) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(bva + row * S2);
    uint32_t acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = negc[k];
#pragma unroll
    for (int g0 = 0; g0 < G; g0 += 4) {
      uint32_t raw[4];
      if constexpr (G % 4 == 0) {
        const uint4 v = reinterpret_cast<const uint4*>(src)[g0 / 4];
        raw[0] = v.x; raw[1] = v.y; raw[2] = v.z; raw[3] = v.w;
      } else {
#pragma unroll
        for (int g = 0; g < 4; g++)
          if (g0 + g < G) raw[g] = src[g0 + g];
      }
#pragma unroll
      for (int g = 0; g < 4; g++) {
        if (g0 + g >= G) continue;
        const uint32_t x = raw[g] ^ 0x80808080u;
        const uint32_t u01 = __builtin_amdgcn_perm(0u, x, 0x0c010c00u);  // (u0, u1)
        const uint32_t u23 = __builtin_amdgcn_perm(0u, x, 0x0c030c02u);  // (u2, u3)
#pragma unroll
        for (int k = 0; k < 8; k++) {
          acc[k] = dot2_u16(u01, PU[k][2 * (g0 + g)], acc[k]);
          acc[k] = dot2_u16(u23, PU[k][2 * (g0 + g) + 1], acc[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t x = fold16(fold16(acc[k]));
      const uint32_t y = x - P;
      out.c[k] = x < y ? x : y;
    }
  };
  V8 s[2] = {v8_zero(), v8_zero()};
  int32_t cur = -1;
  V8 tl = v8_zero();
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {  // the workgroups sweep the level together
    const OtfChunk d = desc[ch];
    OtfAcc a0, a1;
    otf_zero(a0);
    otf_zero(a1);
#pragma unroll 1
    for (int i = 0; i < OTF_PPT; i++) {
      asm volatile("" ::: "memory");  // the coefficient pairs stay in LDS
      const uint64_t m = (ch << OTF_LCP) + uint64_t(i) * WG + threadIdx.x;
      V8 y[2];
      if ((2 * m + 2) * S2 <= available) {
        witness(2 * m, y[0]);
        witness(2 * m + 1, y[1]);
      } else {
        witness_rows<S2, 2>(bva, 2 * m, available, pinv, y);
      }
      ow[2 * m] = v8_store(y[0]);
      ow[2 * m + 1] = v8_store(y[1]);
      if (d.term >= 0) otf_mac2(a0, a1, y[0], y[1], v8_load(thi[uint64_t(d.term) * OTF_PPT + i]));
    }
    if (d.term >= 0) otf_chunk_sum(d, tlo, cur, tl, a0, a1, s);
  }
  block_sums<2>(s, partials);
}

// The same round on the matrix cores (S2 = 32, 64, 128): a witness row is S2
// BVA entries times the pinv2 coefficients, an exact i8 contraction.
// - Digits: coefficient t as t' = t or t - p in [-32881, 32639], t' = 256 h + l
//   with h, l in [-128, 127] (as the MFMA projection). B column n < 8 holds the
//   low digits of coefficient lane n, column 8 + n the high digits.
// - v_mfma_i32_16x16x32_i8 with A[m][k] = entry k of row m (lane i of 16-lane
//   group q: row i, entries 8q .. 8q + 7, one 8-byte load), B[k][n] constant
//   per lane, D[m][n] in lane n + 16 (m / 4), element m % 4; |D| <= S2 2^14, so
//   w[m] lane n = D[m][n] + 256 D[m][8 + n] (|.| < 2^29.1) is exact in i32.
// - A wave takes 64 rows per iteration (four groups of 16), stores the values
//   in LDS and reads back one row per lane: that row's 16-byte output.
// - Cross sums: a workgroup iteration covers 256 rows (128 pairs), a chunk 16
//   iterations. Pair-index bits 0..6 (wave, lane / 2) weigh tlo (128 entries),
//   bits 7..10 (iteration) thi (16 entries); even lanes hold the element 2m of
//   their pair (c0, with g0), odd lanes 2m + 1 (c1, with g1).
// (32-entry rows keep five waves per SIMD at a 12-byte spill, which measures faster)
template <int S2>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(S2 == 32 ? 5 : 1))) void k_otf_first_fold_mfma(const Val* __restrict__ bva, uint64_t available,
                                                            uint64_t chunks, const E8* __restrict__ pinv_g,
                                                            E8* __restrict__ ow, const OtfChunk* __restrict__ desc,
                                                            const E8* __restrict__ tlo, const E8* __restrict__ thi,
                                                            uint32_t* __restrict__ partials) {
  static_assert(S2 % 32 == 0 && S2 <= 128, "whole 32-entry k blocks");
  constexpr int KB = S2 / 32;
  __shared__ __align__(16) uint16_t tbuf[MF_WAVES][64][8];
  __shared__ __align__(16) uint32_t tds[16][16];  // the chunk term's thi[it] as fold_seq sequences
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = lane & 15, q = lane >> 4, par = lane & 1;
  long bop[KB];
#pragma unroll
  for (int kb = 0; kb < KB; kb++) {
    uint64_t x = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int32_t t = int32_t(pinv_g[32 * kb + 8 * q + j].c[n & 7]);
      const int32_t tv = t <= 32639 ? t : t - int32_t(P);
      const int32_t hd = (tv + 128) >> 8, ld = tv - 256 * hd;
      x |= uint64_t(uint8_t(int8_t(n < 8 ? ld : hd))) << (8 * j);
    }
    bop[kb] = long(x);
  }
  V8 sum = v8_zero();  // this lane's cross sum: c0 (even lanes) or c1 (odd lanes)
  int32_t cur = -1;
  V8 tl = v8_zero();
  // A operands (8 entries of one row per lane, four groups of 16 rows) of
  // iteration `it` of chunk `ch`; the next iteration's load while this one computes
  auto load_a = [&](uint64_t ch, int it, long (&a)[4][KB]) {
    const uint64_t row0 = (ch << (OTF_LCP + 1)) + uint64_t(it) * WG + uint64_t(wave) * 64;
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int kb = 0; kb < KB; kb++) {
        const uint64_t at = (row0 + 16 * g + n) * S2 + 32 * kb + 8 * q;
        if (at + 8 <= available) {
          a[g][kb] = *reinterpret_cast<const long*>(bva + at);
        } else {  // entries at or past the active end read as zero
          uint64_t x = 0;
          for (int j = 0; j < 8; j++)
            if (at + j < available) x |= uint64_t(uint8_t(bva[at + j])) << (8 * j);
          a[g][kb] = long(x);
        }
      }
  };
  // one iteration on operands `a`: the contraction, the digit join, the LDS
  // transpose, the store and the row's cross-sum product
  auto step = [&](const long (&a)[4][KB], uint64_t ch, int it, const OtfChunk& d, LazySum& acc) {
    const uint64_t row0 = (ch << (OTF_LCP + 1)) + uint64_t(it) * WG + uint64_t(wave) * 64;  // this wave's rows
    qf_i32x4 dsum[4];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      dsum[g] = qf_i32x4{0, 0, 0, 0};
#pragma unroll
      for (int kb = 0; kb < KB; kb++) dsum[g] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a[g][kb], bop[kb], dsum[g], 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the previous iteration's reads of tbuf are done
    // lanes n < 8 join rows 4q, 4q + 1 of each group, lanes n >= 8 rows 4q + 2, 4q + 3
    // (the low digit sum from lane n & 7, the high one from lane 8 + (n & 7)); a
    // multiple of p keeps the sum positive for two 16-bit folds
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int32_t own = dsum[g][j], other = __shfl_xor(own, 8, 64);
        if ((j < 2) == (n < 8)) {
          const int32_t lo = n < 8 ? own : other, hi = n < 8 ? other : own;
          const uint32_t x = fold16(fold16(uint32_t(lo + 256 * hi + 16384 * int32_t(P))));
          const uint32_t y = x - P;
          tbuf[wave][16 * g + 4 * q + j][n & 7] = uint16_t(x < y ? x : y);
        }
      }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const uint4 raw = *reinterpret_cast<const uint4*>(&tbuf[wave][lane][0]);
    E8 e;
    *reinterpret_cast<uint4*>(&e) = raw;
    ow[row0 + lane] = e;
    if (d.term >= 0) {  // the 8-bit limbs straight from the packed lane pairs, thi[it]'s sequence from LDS
      uint32_t D[15], l4[4], h4[4];
      const uint32_t wv[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
      for (int jp = 0; jp < 4; jp++) {
        l4[jp] = wv[jp] & 0x00ff00ffu;
        h4[jp] = (wv[jp] >> 8) & 0x00ff00ffu;
      }
#pragma unroll
      for (int m = 1; m < 15; m++) D[m] = tds[it][m];
      lazy_mac(acc, l4, h4, D);
    }
  };
  // two operand sets: the loads of the next iteration are in flight while one computes
  long a0[4][KB], a1[4][KB];
  if (blockIdx.x < chunks) load_a(blockIdx.x, 0, a0);
  int32_t seq_term = -1;  // the term whose sequences tds holds
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {  // the workgroups sweep the level together
    const OtfChunk d = desc[ch];
    if (d.term >= 0 && d.term != seq_term) {  // uniform: every wave takes this chunk
      __syncthreads();
      if (threadIdx.x < 16) {
        uint32_t D[15];
        fold_seq(v8_load(thi[uint64_t(d.term) * 16 + threadIdx.x]), D);
        for (int m = 1; m < 15; m++) tds[threadIdx.x][m] = D[m];
      }
      __syncthreads();
      seq_term = d.term;
    }
    LazySum acc;
    lazy_zero(acc);
    for (int it = 0; it < 16; it += 2) {
      load_a(ch, it + 1, a1);
      step(a0, ch, it, d, acc);
      if (it + 2 < 16) load_a(ch, it + 2, a0);
      else if (ch + gridDim.x < chunks) load_a(ch + gridDim.x, 0, a0);
      step(a1, ch, it + 1, d, acc);
    }
    if (d.term >= 0) {
      if (d.term != cur) {
        tl = v8_load(tlo[uint64_t(d.term) * 128 + wave * 32 + (lane >> 1)]);
        cur = d.term;
      }
      sum = v8_add(sum, v8_mul(lazy_value(acc), v8_mul(tl, v8_load(par ? d.g1 : d.g0))));
    }
  }
  V8 s[2] = {par ? v8_zero() : sum, par ? sum : v8_zero()};
  block_sums<2>(s, partials);
}
#else
template <int S2>
__global__ __launch_bounds__(WG) void k_otf_first_fold(const Val* __restrict__ bva, uint64_t available,
                                                       uint64_t chunks, const E8* __restrict__ pinv_g,
                                                       E8* __restrict__ ow, const OtfChunk* __restrict__ desc,
                                                       const E8* __restrict__ tlo, const E8* __restrict__ thi,
                                                       uint32_t* __restrict__ partials) {
  __shared__ E8 pinv[S2];
  for (int i = threadIdx.x; i < S2; i += WG) pinv[i] = pinv_g[i];
  __syncthreads();
  V8 s[2] = {v8_zero(), v8_zero()};
  int32_t cur = -1;
  V8 tl = v8_zero();
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {  // the workgroups sweep the level together
    const OtfChunk d = desc[ch];
    OtfAcc a0, a1;
    otf_zero(a0);
    otf_zero(a1);
    for (int i = 0; i < OTF_PPT; i++) {
      const uint64_t m = (ch << OTF_LCP) + uint64_t(i) * WG + threadIdx.x;
      V8 y[2];
      witness_rows<S2, 2>(bva, 2 * m, available, pinv, y);
      ow[2 * m] = v8_store(y[0]);
      ow[2 * m + 1] = v8_store(y[1]);
      if (d.term >= 0) otf_mac2(a0, a1, y[0], y[1], v8_load(thi[uint64_t(d.term) * OTF_PPT + i]));
    }
    if (d.term >= 0) otf_chunk_sum(d, tlo, cur, tl, a0, a1, s);
  }
  block_sums<2>(s, partials);
}
#endif

// The rest of a level's cross terms over the folded w (length h): each piece
// projects its pairs on its term's tensor (the factors of the pair-index bits
// multiplied per pair), each item adds v times the output it pairs with.
__global__ __launch_bounds__(WG) void k_otf_sparse(const E8* __restrict__ w, const OtfPieceDev* __restrict__ pieces,
                                                   uint32_t npieces, const OtfItemDev* __restrict__ items,
                                                   uint32_t nitems, const E8 (*__restrict__ f)[2],
                                                   uint32_t* __restrict__ partials) {
  V8 s[2] = {v8_zero(), v8_zero()};
  for (uint32_t pc = blockIdx.x; pc < npieces; pc += gridDim.x) {
    const OtfPieceDev pp = pieces[pc];
    V8 a0 = v8_zero(), a1 = v8_zero();
    for (uint32_t k = threadIdx.x; k < pp.count; k += WG) {
      const V8 q = otf_tensor(f, pp.fbase, pp.nbits, pp.rel0 + k);
      const uint64_t m = pp.pair0 + k;
      a0 = v8_add(a0, v8_mul(v8_load(w[2 * m]), q));
      a1 = v8_add(a1, v8_mul(v8_load(w[2 * m + 1]), q));
    }
    s[0] = v8_add(s[0], v8_mul(a0, v8_load(pp.g0)));
    s[1] = v8_add(s[1], v8_mul(a1, v8_load(pp.g1)));
  }
  for (uint64_t it = uint64_t(blockIdx.x) * WG + threadIdx.x; it < nitems; it += uint64_t(gridDim.x) * WG) {
    const OtfItemDev I = items[it];
    if (I.j & 1) s[0] = v8_add(s[0], v8_mul(v8_load(w[I.j - 1]), v8_load(I.v)));
    else s[1] = v8_add(s[1], v8_mul(v8_load(w[I.j + 1]), v8_load(I.v)));
  }
  block_sums<2>(s, partials);
}

// Joined lanes of stride-2s projection chunks: out[ch][l] = f[o][0] parts[ch][l]
// + f[o][1] parts[ch][s + l], o = owner[ch] (the stride-s projection, whose row
// bit 0 carries the term's factor f[o]).
__global__ __launch_bounds__(WG) void k_lane_join(const E8* __restrict__ parts, const uint32_t* __restrict__ owner,
                                                  uint32_t nch, int s, const E8 (*__restrict__ f)[2],
                                                  E8* __restrict__ out) {
  const uint64_t total = uint64_t(nch) * s;
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < total; i += uint64_t(gridDim.x) * WG) {
    const uint64_t ch = i / s, l = i % s;
    const uint32_t o = owner[ch];
    const E8* pr = parts + ch * 2 * s;
    out[i] = v8_store(v8_add(v8_mul(v8_load(pr[l]), v8_load(f[o][0])), v8_mul(v8_load(pr[s + l]), v8_load(f[o][1]))));
  }
}

// The first binary round's cross terms from the stride-2s projection chunks,
// once the packed fold challenge c is known: per chunk u0 = sum_{l<s} c^-l
// parts[ch][l] (the even rows of its superrows), u1 the same over the odd rows,
// and c0 = sum coef[o][1] u0, c1 = sum coef[o][0] u1, with coef[o][b] =
// (sum_l low_o[l] c^l) f_o[b] the folded query's weight of row bit b.
__global__ __launch_bounds__(WG) void k_first_cross(const E8* __restrict__ parts, const uint32_t* __restrict__ owner,
                                                    uint32_t nch, int s, const E8* __restrict__ pinv,
                                                    const E8 (*__restrict__ coef)[2], uint32_t* __restrict__ partials) {
  V8 acc[2] = {v8_zero(), v8_zero()};
  for (uint32_t ch = blockIdx.x * WG + threadIdx.x; ch < nch; ch += gridDim.x * WG) {
    const E8* pr = parts + uint64_t(ch) * 2 * s;
    V8 u0 = v8_zero(), u1 = v8_zero();
    for (int l = 0; l < s; l++) {
      const V8 pw = v8_load(pinv[l]);
      u0 = v8_add(u0, v8_mul(v8_load(pr[l]), pw));
      u1 = v8_add(u1, v8_mul(v8_load(pr[s + l]), pw));
    }
    const uint32_t o = owner[ch];
    acc[0] = v8_add(acc[0], v8_mul(u0, v8_load(coef[o][1])));
    acc[1] = v8_add(acc[1], v8_mul(u1, v8_load(coef[o][0])));
  }
  block_sums<2>(acc, partials);
}

struct FillTaskDev { uint64_t start, first; uint32_t nbits, pair_off, k, count_log; E8 scale; };
// out[start + j] (+)= scale * W(j), j in [first, first + 2^count_log): ADD
// accumulates, else the task's rows are written without being read. A chunk
// is at most 16 groups of 2^k rows; group g's scale (the task scale times
// the weights of the bits above k) is computed once, by thread g. 1024
// threads share the 64 KiB low-bit table, so a CU runs 32 waves of the kernel.
constexpr int FWG = 1024;
template <bool ADD>
__global__ __launch_bounds__(FWG) void k_tensor_fill(E8* __restrict__ out, const FillTaskDev* __restrict__ tasks,
                                                     const E8 (*__restrict__ pairs)[2]) {
  __shared__ E8 T[4096];
  __shared__ E8 Chi[16];
  const FillTaskDev t = tasks[blockIdx.x];
  const E8(*tp)[2] = pairs + t.pair_off;
  build_table<FWG>(T, tp, int(t.k));
  const uint32_t lo_n = 1u << t.k, groups = 1u << (t.count_log - t.k);
  if (threadIdx.x < groups) {
    const uint64_t j_hi = t.first + uint64_t(threadIdx.x) * lo_n;
    V8 c = v8_load(t.scale);
    for (uint32_t b = t.k; b < t.nbits; b++) c = v8_mul(c, v8_load(tp[b][(j_hi >> b) & 1]));
    Chi[threadIdx.x] = v8_store(c);
  }
  __syncthreads();
  for (uint32_t g = 0; g < groups; g++) {
    const V8 c = v8_load(Chi[g]);  // the uniform operand second: its product setup stays out of the loop
    E8* o = out + t.start + t.first + uint64_t(g) * lo_n;
    for (uint32_t lo = threadIdx.x; lo < lo_n; lo += FWG) {
      const V8 x = v8_mul(v8_load(T[lo]), c);
      o[lo] = v8_store(ADD ? v8_add(v8_load(o[lo]), x) : x);
    }
  }
}

// Zero out[start, start + len) for each (start, len) pair of the table, one
// workgroup per pair.
__global__ __launch_bounds__(WG) void k_zero_ranges(E8* __restrict__ out, const uint64_t* __restrict__ ranges) {
  const uint64_t start = ranges[2 * blockIdx.x], len = ranges[2 * blockIdx.x + 1];
  uint4* o = reinterpret_cast<uint4*>(out + start);
  for (uint64_t i = threadIdx.x; i < len; i += WG) o[i] = make_uint4(0, 0, 0, 0);
}

// Shifted lane correlations of the term projections (tbv.rs:361-392): block b
// < L-1 sums pos[d-1] = sum_ch sum_lane parts[ch][lane] low[owner][lane+d],
// block L-1+b the mirror neg[d-1] = sum parts[ch][lane+d] low[owner][lane],
// d = b + 1, over the projection chunks (parts) and the low-bit weights of
// each chunk's term.
__global__ __launch_bounds__(WG) void k_proj_cross(const E8* __restrict__ parts, const uint32_t* __restrict__ owner,
                                                   uint32_t nchunks, int L, const E8* __restrict__ lows,
                                                   E8* __restrict__ out) {
  __shared__ uint32_t red[WG][LANES];
  // block 2 (L - 1): the d = 0 correlation sum parts[ch][lane] low[owner][lane]
  const bool zero = blockIdx.x == uint32_t(2 * (L - 1));
  const int d = zero ? 0 : int(blockIdx.x % uint32_t(L - 1)) + 1;
  const bool neg = !zero && blockIdx.x >= uint32_t(L - 1);
  V8 acc = v8_zero();
  for (uint32_t ch = threadIdx.x; ch < nchunks; ch += WG) {
    const E8* pr = parts + uint64_t(ch) * L;
    const E8* lw = lows + uint64_t(owner[ch]) * L;
    for (int lane = 0; lane + d < L; lane++)
      acc = v8_add(acc, neg ? v8_mul(v8_load(pr[lane + d]), v8_load(lw[lane]))
                            : v8_mul(v8_load(pr[lane]), v8_load(lw[lane + d])));
  }
#pragma unroll
  for (int k = 0; k < LANES; k++) red[threadIdx.x][k] = acc.c[k];
  __syncthreads();
  for (int h = WG / 2; h > 0; h /= 2) {
    if (int(threadIdx.x) < h)
      for (int k = 0; k < LANES; k++) red[threadIdx.x][k] = madd(red[threadIdx.x][k], red[threadIdx.x + h][k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    V8 x;
    for (int k = 0; k < LANES; k++) x.c[k] = red[0][k];
    out[blockIdx.x] = v8_store(x);
  }
}

// values[i] = sum_b fw[b] * packed(b, i); `available` and `block_bytes` in bytes.
#ifndef QF_KB
// The F65521 block fold on the matrix cores: one exact i8 contraction over
// k = 8 b + j (block b, byte j of the element).
// - fw[b] * packed(b, i) = M_b v with v the element's 8 signed bytes and M_b
//   the matrix of multiplication by fw[b] (column j = fw[b] X^j), so
//   values[i][c] = sum_{b, j} M_b[c][j] v_{b,j}.
// - Digits: M_b[c][j] as t' = t or t - p in [-32881, 32639], t' = 256 h + l,
//   h and l in [-128, 127], built on the host. B column n < 8 holds the low
//   digits of coefficient n, column 8 + n the high digits; step s covers
//   blocks 4 s .. 4 s + 3 (lane n + 16 q: block 4 s + q), zero past the last.
// - v_mfma_i32_16x16x32_i8 with A[m][k] = byte k of position m (lane i of
//   16-lane group q: position i, block 4 s + q, one 8-byte load), D[m][n] in
//   lane n + 16 (m / 4), element m % 4; a step adds at most 32 x 2^14 = 2^19
//   and there are at most 32 steps, so |D| <= 2^24 and the high digit sum is
//   reduced before it is scaled by 256.
// - A wave takes 64 positions per iteration (four groups of 16), joins the
//   digits across lanes n and n ^ 8, transposes through LDS and stores one
//   16-byte element per lane.
// Bytes at or past `available` read as zero.
__global__ __launch_bounds__(WG) void k_block_fold_mfma(const uint8_t* __restrict__ bva, uint64_t available,
                                                        uint64_t block_bytes, uint32_t steps,
                                                        const long* __restrict__ bops, E8* __restrict__ values,
                                                        uint64_t n) {
  __shared__ long lbop[32][64];
  __shared__ __align__(16) uint16_t tbuf[MF_WAVES][64][8];
  for (uint32_t i = threadIdx.x; i < steps * 64; i += WG) lbop[i / 64][i % 64] = bops[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nn = lane & 15, q = lane >> 4;
  for (uint64_t base = uint64_t(blockIdx.x) * WG; base < n; base += uint64_t(gridDim.x) * WG) {
    const uint64_t pos0 = base + uint64_t(wave) * 64;  // this wave's 64 positions
    qf_i32x4 dsum[4];
#pragma unroll
    for (int g = 0; g < 4; g++) dsum[g] = qf_i32x4{0, 0, 0, 0};
    for (uint32_t s0 = 0; s0 < steps; s0 += 2) {  // two steps' loads in flight together
      long a[2][4];
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int g = 0; g < 4; g++) {
          const uint64_t pos = pos0 + 16 * g + nn;
          const uint64_t at = uint64_t(4 * (s0 + h) + q) * block_bytes + 8 * pos;
          uint64_t x = 0;
          if (s0 + h < steps && pos < n) {
            if (at + 8 <= available) {
              x = *reinterpret_cast<const uint64_t*>(bva + at);
            } else if (at < available) {
              for (int j = 0; j < 8; j++)
                if (at + j < available) x |= uint64_t(bva[at + j]) << (8 * j);
            }
          }
          a[h][g] = long(x);
        }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (s0 + h >= steps) break;
        const long bop = lbop[s0 + h][lane];
#pragma unroll
        for (int g = 0; g < 4; g++) dsum[g] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a[h][g], bop, dsum[g], 0, 0, 0);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the previous iteration's reads of tbuf are done
    // lanes nn < 8 join rows 4q, 4q + 1 of each group, lanes nn >= 8 rows 4q + 2, 4q + 3
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int32_t own = dsum[g][j], other = __shfl_xor(own, 8, 64);
        if ((j < 2) == (nn < 8)) {
          const int32_t lo = nn < 8 ? own : other, hi = nn < 8 ? other : own;
          const uint32_t h = fold16(fold16(uint32_t(hi + 16384 * int32_t(P))));  // < 2^16 + 135
          const uint32_t x = fold16(fold16(uint32_t(lo + 16384 * int32_t(P)) + 256u * h));
          const uint32_t y = x - P;
          tbuf[wave][16 * g + 4 * q + j][nn & 7] = uint16_t(x < y ? x : y);
        }
      }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const uint4 raw = *reinterpret_cast<const uint4*>(&tbuf[wave][lane][0]);
    if (pos0 + lane < n) {
      E8 e;
      *reinterpret_cast<uint4*>(&e) = raw;
      values[pos0 + lane] = e;
    }
  }
}

// F65521 form. packed(b, i) has the 8 signed bytes v_j of the element as
// coefficients; with u_j = v_j + 128 (one xor per four bytes) each is a single
// 8-bit limb of the v8_mul_dot2 product, so fw[b] u costs 32 v_dot2_u32_u16
// against fw[b]'s wrap-folded pair sequence, built once per block in LDS.
// Bytes at or past `available` count as v = 0 (u = 128). The bias of the
// first nb blocks, sum_{b < nb} fw[b] (128, ..., 128), is removed at the end
// (negb[nb]). Each block adds < 2^27.1 to a coefficient sum: folds every 16.
__global__ __launch_bounds__(WG) void k_block_fold(const uint8_t* __restrict__ bva, uint64_t available,
                                                   uint64_t block_bytes, uint32_t blocks, const E8* __restrict__ fw_g,
                                                   E8* __restrict__ values, uint64_t n) {
  __shared__ uint4 FD[128][4];       // D[0..15] of each block (D[0], D[15] unused)
  __shared__ uint32_t negb[129][8];  // -(sum_{b < nb} fw[b] (128, ..., 128)) mod P
  for (uint32_t b = threadIdx.x; b < blocks; b += WG) {
    const V8 f = v8_load(fw_g[b]);
    uint32_t D[15];
    fold_seq(f, D);
    uint32_t* fd = reinterpret_cast<uint32_t*>(FD[b]);
    fd[0] = 0;
    for (int m = 1; m < 15; m++) fd[m] = D[m];
    fd[15] = 0;
    V8 k128;
    for (int k = 0; k < 8; k++) k128.c[k] = 128;
    const V8 pb = v8_mul(f, k128);  // this block's bias, summed below
    for (int k = 0; k < 8; k++) negb[b + 1][k] = pb.c[k];
  }
  __syncthreads();
  if (threadIdx.x < 8) {  // prefix sums per coefficient, negated
    const int k = threadIdx.x;
    uint32_t run = 0;
    negb[0][k] = 0;
    for (uint32_t b = 1; b <= blocks; b++) {
      run = madd(run, negb[b][k]);
      negb[b][k] = msub(0, run);
    }
  }
  __syncthreads();
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG) {
    uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    V8 total = v8_zero();
    auto add_block = [&](uint32_t x0, uint32_t x1, uint32_t b) {
      x0 ^= 0x80808080u;
      x1 ^= 0x80808080u;
      const uint32_t lo[4] = {__builtin_amdgcn_perm(0u, x0, 0x0c010c00u), __builtin_amdgcn_perm(0u, x0, 0x0c030c02u),
                              __builtin_amdgcn_perm(0u, x1, 0x0c010c00u), __builtin_amdgcn_perm(0u, x1, 0x0c030c02u)};
      uint32_t D[16];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const uint4 v = FD[b][r];
        D[4 * r] = v.x; D[4 * r + 1] = v.y; D[4 * r + 2] = v.z; D[4 * r + 3] = v.w;
      }
#pragma unroll
      for (int c = 0; c < 8; c++)
#pragma unroll
        for (int jp = 0; jp < 4; jp++) acc[c] = dot2_u16(lo[jp], D[7 + c - 2 * jp], acc[c]);
    };
    auto fold_acc = [&]() {
      V8 r;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        uint32_t x = fold16(fold16(acc[c]));
        const uint32_t y = x - P;
        r.c[c] = x < y ? x : y;
        acc[c] = 0;
      }
      total = v8_add(total, r);
    };
    // Blocks [0, full) hold this element's 8 bytes entirely below `available`:
    // their loads are unguarded and issued four at a time.
    const uint64_t off = 8 * i;
    const uint32_t full = available >= off + 8
                              ? uint32_t(std::min<uint64_t>(blocks, (available - off - 8) / block_bytes + 1))
                              : 0u;
    uint32_t b = 0;
    for (; b + 4 <= full; b += 4) {
      uint2 u[4];
#pragma unroll
      for (int k = 0; k < 4; k++) u[k] = *reinterpret_cast<const uint2*>(bva + uint64_t(b + k) * block_bytes + off);
#pragma unroll
      for (int k = 0; k < 4; k++) add_block(u[k].x, u[k].y, b + k);
      if ((b + 4) % 16 == 0) fold_acc();
    }
    for (; b < full; b++) {
      const uint2 u = *reinterpret_cast<const uint2*>(bva + uint64_t(b) * block_bytes + off);
      add_block(u.x, u.y, b);
    }
    uint32_t nb = full;
    if (b < blocks) {  // the block that straddles `available`; later blocks are zero and skipped
      const uint64_t byte = uint64_t(b) * block_bytes + off;
      uint32_t x[2] = {0, 0};
      for (int j = 0; j < 8; j++) if (byte + j < available) x[j >> 2] |= uint32_t(bva[byte + j]) << (8 * (j & 3));
      add_block(x[0], x[1], b);
      nb = b + 1;
    }
    fold_acc();
    V8 bias;
#pragma unroll
    for (int c = 0; c < 8; c++) bias.c[c] = negb[nb][c];
    values[i] = v8_store(v8_add(total, bias));
  }
}
#else
__global__ __launch_bounds__(WG) void k_block_fold(const uint8_t* __restrict__ bva, uint64_t available,
                                                   uint64_t block_bytes, uint32_t blocks, const E8* __restrict__ fw_g,
                                                   E8* __restrict__ values, uint64_t n) {
  __shared__ E8 fw[128];
  for (uint32_t i = threadIdx.x; i < blocks; i += WG) fw[i] = fw_g[i];
  __syncthreads();
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG) {
    Acc S[LANES][LANES];
    for (int j = 0; j < LANES; j++)
      for (int k = 0; k < LANES; k++) S[j][k] = 0;
    auto add_block = [&](const uint32_t (&x)[2], const E8& f) {
#pragma unroll
      for (int j = 0; j < LANES; j++) {
#ifdef QF_KB
        const Acc v = int16_t(x[j >> 1] >> (16 * (j & 1)));
#else
        const Acc v = int8_t(x[j >> 2] >> (8 * (j & 3)));
#endif
#pragma unroll
        for (int k = 0; k < LANES; k++) acc_mac(S[j][k], int32_t(v), f.c[k]);
      }
    };
    // Blocks [0, full) hold this element's 8 bytes entirely below `available`:
    // their loads are unguarded and issued four at a time.
    const uint64_t off = 8 * i;
    const uint32_t full = available >= off + 8
                              ? uint32_t(std::min<uint64_t>(blocks, (available - off - 8) / block_bytes + 1))
                              : 0u;
    uint32_t b = 0;
    for (; b + 4 <= full; b += 4) {
      uint2 u[4];
#pragma unroll
      for (int k = 0; k < 4; k++) u[k] = *reinterpret_cast<const uint2*>(bva + uint64_t(b + k) * block_bytes + off);
#pragma unroll
      for (int k = 0; k < 4; k++) add_block({u[k].x, u[k].y}, fw[b + k]);
    }
    for (; b < full; b++) {
      const uint2 u = *reinterpret_cast<const uint2*>(bva + uint64_t(b) * block_bytes + off);
      add_block({u.x, u.y}, fw[b]);
    }
    if (b < blocks) {  // the block that straddles `available`; later blocks read as zero
      const uint64_t byte = uint64_t(b) * block_bytes + off;
      uint32_t x[2] = {0, 0};
      for (int j = 0; j < 8; j++) if (byte + j < available) x[j >> 2] |= uint32_t(bva[byte + j]) << (8 * (j & 3));
      add_block(x, fw[b]);
    }
    // value = sum_j X^j * S_j with X^LANES = EXT_W
    uint64_t acc[LANES];
#pragma unroll
    for (int k = 0; k < LANES; k++) acc[k] = 0;
#pragma unroll
    for (int j = 0; j < LANES; j++)
#pragma unroll
      for (int k = 0; k < LANES; k++) {
        const uint32_t s = mod_signed(S[j][k]);
        const int d = j + k;
        if (d < LANES) acc[d] += s; else acc[d - LANES] += uint64_t(EXT_W) * s;
      }
    V8 o;
#pragma unroll
    for (int k = 0; k < LANES; k++) o.c[k] = mred64(acc[k]);
    values[i] = v8_store(o);
  }
}
#endif

struct Images { E8 m[LANES]; };
__global__ __launch_bounds__(WG) void k_weight_images(E8* __restrict__ out, uint64_t n, const E8* __restrict__ hi,
                                                      const E8* __restrict__ lo, int a, Images im) {
  const uint64_t mask = (uint64_t(1) << a) - 1;
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG) {
    V8 e = v8_mul(v8_load(hi[i >> a]), v8_load(lo[i & mask]));
    uint64_t acc[LANES];  // LANES products of canonical coefficients: < 2^35 (F65521), < 2^64 (KoalaBear)
#pragma unroll
    for (int k = 0; k < LANES; k++) acc[k] = 0;
#pragma unroll
    for (int j = 0; j < LANES; j++)
#pragma unroll
      for (int k = 0; k < LANES; k++) acc[k] += uint64_t(im.m[j].c[k]) * e.c[j];
    V8 o;
#pragma unroll
    for (int k = 0; k < LANES; k++) o.c[k] = mred64(acc[k]);
    out[i] = v8_store(o);
  }
}

dim3 grid_for(uint64_t n) {
  uint64_t g = (n + WG - 1) / WG;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return dim3(uint32_t(g));
}

// Host-to-device copies of the opening's small tables and index lists go
// through this thread's pinned staging ring: the host data is copied into the
// ring at once and the device copy runs asynchronously in stream order, so
// the caller's buffer may go away and no synchronisation is needed. Only a
// wrap of the ring waits for the stream (everything staged before has then
// been consumed); copies larger than the ring are synchronous.
struct StagingRing {
  char* base = nullptr;
  size_t cap = 0, head = 0;
  hipStream_t st = nullptr;  // the stream the staged copies were issued on
  ~StagingRing() { if (base) (void)hipHostFree(base); }
};
thread_local StagingRing g_ring;
void h2d(DevCtx& c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  StagingRing& r = g_ring;
  constexpr size_t CAP = size_t(16) << 20;
  if (bytes > CAP) {
    QF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.st));
    c.sync();
    return;
  }
  if (!r.base) {
    QF_HIP(hipHostMalloc(&r.base, CAP, 0));
    r.cap = CAP;
  }
  if (r.st && r.st != c.st) {  // staged on another stream: let those copies finish first
    QF_HIP(hipStreamSynchronize(r.st));
    r.head = 0;
  }
  r.st = c.st;
  const size_t off = (r.head + 255) & ~size_t(255);
  if (off + bytes > r.cap) {
    QF_HIP(hipStreamSynchronize(c.st));
    r.head = 0;
    return h2d(c, dst, src, bytes);
  }
  std::memcpy(r.base + off, src, bytes);
  QF_HIP(hipMemcpyAsync(dst, r.base + off, bytes, hipMemcpyHostToDevice, c.st));
  r.head = off + bytes;
}

// Small host tables go to the scratch table in stream order (the next call's
// copy runs after this call's kernels).
E8* put_table(DevCtx& c, E8* scratch, const void* host, size_t bytes) {
  h2d(c, scratch, host, bytes);
  return scratch;
}

struct Scratch {
  E8* tab = nullptr;
  ~Scratch() { if (tab) (void)hipFree(tab); }
};
thread_local Scratch g_scratch;
E8* scratch_tab() {
  if (!g_scratch.tab) QF_HIP(hipMalloc(&g_scratch.tab, TAB_CAP * 16));
  return g_scratch.tab;
}
}  // namespace

// ---------------------------------------------------------------- allocation
// A caching allocator: released blocks are kept by exact (rounded) size and
// handed out again, so repeated commits and proofs of one shape never call
// hipMalloc/hipFree (both synchronise the device) after the first. A block
// released on stream S carries an event recorded on S; a later user on
// another stream waits for that event before touching it.
namespace {
thread_local hipStream_t g_current_stream = nullptr;
struct CachedBlock { void* p; hipEvent_t ready; };
struct BlockCache {
  std::mutex mu;
  std::unordered_map<void*, size_t> live;               // pointer -> rounded size
  std::multimap<size_t, CachedBlock> idle;              // rounded size -> block
  size_t idle_bytes = 0;
  size_t limit = 0;                                     // idle bytes kept before trimming: 3/4 of the device
  static size_t round(size_t b) {
    const size_t g = b >= (size_t(1) << 20) ? (size_t(2) << 20) : 256;
    return (std::max<size_t>(b, 16) + g - 1) / g * g;
  }
  void trim_locked(size_t keep) {
    while (idle_bytes > keep && !idle.empty()) {
      auto it = idle.begin();
      (void)hipEventSynchronize(it->second.ready);
      (void)hipEventDestroy(it->second.ready);
      (void)hipFree(it->second.p);
      idle_bytes -= it->first;
      idle.erase(it);
    }
  }
};
BlockCache& cache() {
  static BlockCache* c = new BlockCache;  // lives until process exit
  return *c;
}
}  // namespace
hipStream_t current_stream() { return g_current_stream; }
void device_setup() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char* w = std::getenv("QF_HOST_WAIT");  // "spin" (default), "yield" or "auto"
    const unsigned flag = !w || std::strcmp(w, "spin") == 0 ? hipDeviceScheduleSpin
                          : std::strcmp(w, "yield") == 0   ? hipDeviceScheduleYield
                                                           : hipDeviceScheduleAuto;
    if (hipSetDeviceFlags(flag) != hipSuccess) (void)hipGetLastError();
  });
}
size_t cached_idle_bytes() {
  BlockCache& bc = cache();
  std::lock_guard<std::mutex> g(bc.mu);
  return bc.idle_bytes;
}
void release_cached_blocks() {
  BlockCache& bc = cache();
  std::lock_guard<std::mutex> g(bc.mu);
  bc.trim_locked(0);
}
hipStream_t thread_stream() {
  device_setup();
  thread_local hipStream_t st = nullptr;
  if (!st) QF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  return st;
}
hipStream_t thread_stream2() {
  device_setup();
  thread_local hipStream_t st = nullptr;
  if (!st) QF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  return st;
}
void* dmalloc(size_t bytes, hipStream_t st) {
  device_setup();
  BlockCache& bc = cache();
  const size_t size = BlockCache::round(bytes);
  {
    std::lock_guard<std::mutex> g(bc.mu);
    auto it = bc.idle.find(size);
    if (it != bc.idle.end()) {
      CachedBlock b = it->second;
      bc.idle.erase(it);
      bc.idle_bytes -= size;
      bc.live[b.p] = size;
      QF_HIP(hipStreamWaitEvent(st, b.ready, 0));
      (void)hipEventDestroy(b.ready);
      return b.p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, size);
  if (e == hipErrorOutOfMemory) {  // give the idle blocks back and retry once
    (void)hipGetLastError();
    std::lock_guard<std::mutex> g(bc.mu);
    bc.trim_locked(0);
    e = hipMalloc(&p, size);
  }
  if (e != hipSuccess) throw QfError(QF_ERR_DEVICE, std::string("device allocation: ") + hipGetErrorString(e));
  std::lock_guard<std::mutex> g(bc.mu);
  bc.live[p] = size;
  return p;
}
void dfree(void* p, hipStream_t st) {
  if (!p) return;
  BlockCache& bc = cache();
  hipEvent_t ev;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, st) != hipSuccess) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(p);
    std::lock_guard<std::mutex> g(bc.mu);
    bc.live.erase(p);
    return;
  }
  std::lock_guard<std::mutex> g(bc.mu);
  auto it = bc.live.find(p);
  if (it == bc.live.end()) { (void)hipEventDestroy(ev); return; }
  const size_t size = it->second;
  bc.live.erase(it);
  bc.idle.insert({size, {p, ev}});
  bc.idle_bytes += size;
  if (!bc.limit) {
    size_t fr = 0, total = 0;
    bc.limit = hipMemGetInfo(&fr, &total) == hipSuccess ? total / 4 * 3 : (size_t(64) << 30);
  }
  bc.trim_locked(bc.limit);
}

// ---------------------------------------------------------------- DVec / DevCtx
DVec::DVec(uint64_t len) : n(len), st(current_stream()) {
  p = static_cast<E8*>(dmalloc(std::max<uint64_t>(len, 1) * 16, st));
}
DVec& DVec::operator=(DVec&& o) noexcept {
  if (this != &o) { dfree(p, st); p = o.p; n = o.n; st = o.st; o.p = nullptr; o.n = 0; }
  return *this;
}
DVec::~DVec() { dfree(p, st); }
// Reduction partials live in this thread's coherent pinned host buffer: the
// kernels write their block sums there directly, so a reduction costs one
// stream synchronisation and no copy.
namespace {
struct PinnedPartials {
  uint32_t* host = nullptr;
  ~PinnedPartials() { if (host) (void)hipHostFree(host); }
};
thread_local PinnedPartials g_pinned;
}  // namespace
DevCtx::DevCtx() {
  st = thread_stream();
  prev = g_current_stream;
  g_current_stream = st;
  if (!g_pinned.host)
    QF_HIP(hipHostMalloc(&g_pinned.host, size_t(MAX_GROUPS) * 64 * 32, hipHostMallocMapped | hipHostMallocCoherent));
  host_partials = g_pinned.host;
  QF_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&partials), host_partials, 0));
}
DevCtx::~DevCtx() {
  if (st) (void)hipStreamSynchronize(st);
  g_current_stream = prev;
}
bool sync_trace_on() {
  static const bool on = [] {
    const char* e = std::getenv("QF_TRACE_SYNC");
    return e && e[0] == '1';
  }();
  return on;
}
SyncTrace& sync_trace() {
  thread_local SyncTrace t;
  return t;
}
namespace {
thread_local std::map<std::string, std::pair<double, uint64_t>> t_scopes;
}
TraceLaps::TraceLaps() {
  if (sync_trace_on()) t = std::chrono::steady_clock::now();
}
void TraceLaps::mark(const char* name) {
  if (!sync_trace_on()) return;
  const auto now = std::chrono::steady_clock::now();
  auto& e = t_scopes[name];
  e.first += std::chrono::duration<double, std::milli>(now - t).count();
  e.second++;
  t = now;
}
void trace_scopes_flush() {
  for (auto& [n, v] : t_scopes) std::fprintf(stderr, "[qf]     %-28s %8.3f ms  x%llu\n", n.c_str(), v.first, (unsigned long long)v.second);
  t_scopes.clear();
}
void DevCtx::sync() {
  if (!sync_trace_on()) {
    QF_HIP(hipStreamSynchronize(st));
    return;
  }
  const auto t0 = std::chrono::steady_clock::now();
  QF_HIP(hipStreamSynchronize(st));
  SyncTrace& t = sync_trace();
  t.syncs++;
  t.wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

namespace {
// Workgroups of WG threads of `kernel` that are resident on the device at once:
// a persistent launch of more would run its last workgroups after the others.
template <typename K>
uint32_t resident_groups(K kernel) {
  int per = 0;
  QF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, WG, 0));
  return uint32_t(std::max(per, 1)) * uint32_t(device_cus());
}
#ifndef QF_KB
// QF_BLOCK_FOLD_MFMA=0 keeps the F65521 block fold and column fold on the vector ALUs.
bool block_fold_mfma_on() {
  const char* e = std::getenv("QF_BLOCK_FOLD_MFMA");
  return !(e && e[0] == '0');
}
#endif
}  // namespace

namespace dv {
DVec upload(DevCtx& c, const std::vector<E8>& v) {
  DVec d(v.size());
  h2d(c, d.p, v.data(), v.size() * 16);
  return d;
}
// Small downloads land in the thread's pinned buffer past the reduction rows
// (an asynchronous DMA and one synchronisation), larger ones in pageable memory.
constexpr uint64_t PIN_DOWN = 65536;                                            // u32 offset
constexpr uint64_t PIN_DOWN_BYTES = uint64_t(DevCtx::MAX_GROUPS) * 64 * 32 - 4 * PIN_DOWN;  // to the buffer's end
// Device-to-host copy through the thread's pinned buffer, in pieces of at most
// PIN_DOWN_BYTES with one synchronisation each: into pageable memory the
// runtime stages the copy itself, at about 1.6 GB/s for a few hundred KiB.
void d2h(DevCtx& c, void* dst, const void* src, size_t bytes) {
  char* pin = reinterpret_cast<char*>(c.host_partials + PIN_DOWN);
  for (size_t off = 0; off < bytes; off += PIN_DOWN_BYTES) {
    const size_t n = std::min<size_t>(PIN_DOWN_BYTES, bytes - off);
    QF_HIP(hipMemcpyAsync(pin, static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, c.st));
    c.sync();
    std::memcpy(static_cast<char*>(dst) + off, pin, n);
  }
}
std::vector<E8> download(DevCtx& c, const E8* p, uint64_t n) {
  return download2(c, p, nullptr, n).first;
}
std::pair<std::vector<E8>, std::vector<E8>> download2(DevCtx& c, const E8* a, const E8* b, uint64_t n) {
  std::vector<E8> va(n), vb(b ? n : 0);
  const uint64_t bytes = n * 16;
  if (!n) return {va, vb};
  if ((b ? 2 : 1) * bytes <= PIN_DOWN_BYTES) {
    char* pin = reinterpret_cast<char*>(c.host_partials + PIN_DOWN);
    QF_HIP(hipMemcpyAsync(pin, a, bytes, hipMemcpyDeviceToHost, c.st));
    if (b) QF_HIP(hipMemcpyAsync(pin + bytes, b, bytes, hipMemcpyDeviceToHost, c.st));
    c.sync();
    std::memcpy(va.data(), pin, bytes);
    if (b) std::memcpy(vb.data(), pin + bytes, bytes);
    return {va, vb};
  }
  d2h(c, va.data(), a, bytes);
  if (b) d2h(c, vb.data(), b, bytes);
  return {va, vb};
}
void fill(DevCtx& c, E8* p, uint64_t n, const E8& v) { hipLaunchKernelGGL(k_fill, grid_for(n), dim3(WG), 0, c.st, p, n, v); }
void copy(DevCtx& c, E8* dst, const E8* src, uint64_t n) {
  if (n) QF_HIP(hipMemcpyAsync(dst, src, n * 16, hipMemcpyDeviceToDevice, c.st));
}
void tensor(DevCtx& c, E8* out, const std::vector<E8>& hi, const std::vector<E8>& lo) {
  require(pow2(hi.size()) && pow2(lo.size()) && hi.size() + lo.size() <= TAB_CAP, QF_ERR_ARGUMENT, "tensor tables");
  E8* t = scratch_tab();
  put_table(c, t, hi.data(), hi.size() * 16);
  put_table(c, t + hi.size(), lo.data(), lo.size() * 16);
  uint64_t n = hi.size() * lo.size();
  hipLaunchKernelGGL(k_tensor, grid_for(n), dim3(WG), 0, c.st, out, n, t, t + hi.size(), ilog2u(lo.size()));
}
E8 inner(DevCtx& c, const E8* a, const E8* b, uint64_t n) { return reduce<R_INNER>(c, a, b, n).first; }
std::pair<E8, E8> cross_halves(DevCtx& c, const E8* a, const E8* b, uint64_t h) { return reduce<R_CROSS>(c, a, b, h); }
void lb_fold(DevCtx& c, E8* a, E8* b, uint64_t h, const E8& x, const E8& y) {
  hipLaunchKernelGGL(k_lb_fold, grid_for(h), dim3(WG), 0, c.st, a, b, h, x, y);
}
std::pair<E8, E8> lb_fold_cross(DevCtx& c, E8* a, E8* b, uint64_t h, const E8& x, const E8& y) {
  require(h >= 2 && h % 2 == 0, QF_ERR_ARGUMENT, "fold length");
  const uint32_t g = groups_for(h / 2);
  hipLaunchKernelGGL(k_lb_fold_cross, dim3(g), dim3(WG), 0, c.st, a, b, h, x, y, c.partials);
  QF_HIP(hipGetLastError());
  auto s = finish_sums<2>(c, g);
  return {s[0], s[1]};
}
void eo_fold(DevCtx& c, const E8* w, const E8* q, E8* ow, E8* oq, uint64_t h, const E8& ri, const E8& r) {
  hipLaunchKernelGGL(k_eo_fold, grid_for(h), dim3(WG), 0, c.st, w, q, ow, oq, h, ri, r);
}
std::pair<E8, E8> eo_fold_cross(DevCtx& c, const E8* w, const E8* q, E8* ow, E8* oq, uint64_t h, const E8& ri,
                                const E8& r) {
  require(h >= 2 && h % 2 == 0, QF_ERR_ARGUMENT, "fold length");
  const uint32_t g = groups_for(h / 2);
  hipLaunchKernelGGL(k_eo_fold_cross, dim3(g), dim3(WG), 0, c.st, w, q, ow, oq, h, ri, r, c.partials);
  QF_HIP(hipGetLastError());
  auto s = finish_sums<2>(c, g);
  return {s[0], s[1]};
}
std::pair<E8, E8> sc_msg(DevCtx& c, const E8* v, const E8* w, uint64_t h) { return reduce<R_SC>(c, v, w, h); }
void sc_fold(DevCtx& c, E8* v, E8* w, uint64_t h, const E8& r) {
  hipLaunchKernelGGL(k_sc_fold, grid_for(h), dim3(WG), 0, c.st, v, w, h, r);
}
std::pair<E8, E8> sc_fold_msg(DevCtx& c, E8* v, E8* w, uint64_t h, const E8& r) {
  require(h >= 2 && h % 2 == 0, QF_ERR_ARGUMENT, "fold length");
  const uint32_t g = groups_for(h / 2);
  hipLaunchKernelGGL(k_sc_fold_msg, dim3(g), dim3(WG), 0, c.st, v, w, h, r, c.partials);
  QF_HIP(hipGetLastError());
  auto s = finish_sums<2>(c, g);
  return {s[0], s[1]};
}
void axpy(DevCtx& c, E8* y, const E8* x, uint64_t n, const E8& s) {
  hipLaunchKernelGGL(k_axpy, grid_for(n), dim3(WG), 0, c.st, y, x, n, s);
}
void mul(DevCtx& c, E8* y, const E8* x, uint64_t n) { hipLaunchKernelGGL(k_mul, grid_for(n), dim3(WG), 0, c.st, y, x, n); }
void scale(DevCtx& c, E8* y, uint64_t n, const E8& s) { hipLaunchKernelGGL(k_scale, grid_for(n), dim3(WG), 0, c.st, y, n, s); }

std::vector<E8> col_projection(DevCtx& c, const E8* cw, uint64_t R, uint64_t n, uint64_t cols, const E8* w) {
  require(cols <= 64, QF_ERR_ARGUMENT, "column projection width");
  uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(DevCtx::MAX_GROUPS / uint32_t(cols), groups_for(n)));
  hipLaunchKernelGGL(k_col_projection, dim3(g, uint32_t(cols)), dim3(WG), 0, c.st, cw, R, n, w, c.partials);
  QF_HIP(hipGetLastError());
  c.sync();
  std::vector<E8> out(cols);
  for (uint64_t col = 0; col < cols; col++) {
    V8 s = v8_zero();
    for (uint32_t i = 0; i < g; i++)
      for (int k = 0; k < LANES; k++) s.c[k] = madd(s.c[k], c.host_partials[(col * g + i) * LANES + k]);
    out[col] = v8_store(s);
  }
  return out;
}
#ifndef QF_KB
namespace {
// B operands and bias of k_col_fold_mfma for the weights f[0 .. k) (4 | k <= 128), into
// the scratch table: per step and lane (nn + 16 q), digits of M_f[4 s + q][c][j]; bias[nn]
// = 128 x the sum of column nn's digits over every k.
std::pair<const long*, const int32_t*> col_fold_tables(DevCtx& c, const E8* f, uint64_t k) {
  const uint32_t steps = uint32_t(k / 4);
  std::vector<uint64_t> bops(size_t(steps) * 64, 0);
  int32_t bias[16] = {};
  for (uint64_t col = 0; col < k; col++) {
    E8 m[8];
    for (int j = 0; j < 8; j++) m[j] = f[col] * e8_basis(j);
    for (int nn = 0; nn < 16; nn++) {
      uint64_t x = 0;
      for (int j = 0; j < 8; j++) {
        const int32_t t = int32_t(m[j].c[nn & 7]);
        const int32_t tv = t <= 32639 ? t : t - int32_t(P);
        const int32_t hd = (tv + 128) >> 8, ld = tv - 256 * hd;
        const int32_t d = nn < 8 ? ld : hd;
        x |= uint64_t(uint8_t(int8_t(d))) << (8 * j);
        bias[nn] += 128 * d;
      }
      bops[(col / 4) * 64 + 16 * (col % 4) + nn] = x;
    }
  }
  E8* t = scratch_tab();
  put_table(c, t, bops.data(), bops.size() * 8);
  put_table(c, t + steps * 32, bias, sizeof(bias));
  return {reinterpret_cast<const long*>(t), reinterpret_cast<const int32_t*>(t + steps * 32)};
}
void launch_col_fold_mfma(DevCtx& c, const E8* cw, uint64_t R, uint64_t rstride, uint64_t n, uint64_t k,
                          const std::pair<const long*, const int32_t*>& tab, E8* out, uint32_t ny, uint64_t in_y,
                          uint64_t out_y) {
  static const uint32_t resident = resident_groups(k_col_fold_mfma);
  const uint64_t g = std::max<uint64_t>(1, std::min<uint64_t>(std::max<uint32_t>(resident / ny, 1), (n + WG - 1) / WG));
  hipLaunchKernelGGL(k_col_fold_mfma, dim3(uint32_t(g), ny), dim3(WG), 0, c.st, cw, R, rstride, n, uint32_t(k / 4),
                     tab.first, tab.second, out, in_y, out_y);
  QF_HIP(hipGetLastError());
}
}  // namespace
#endif

bool col_fold_rows32(DevCtx& c, const E8* cw, uint64_t R, uint64_t n, uint64_t cols, const std::vector<E8>& low,
                     E8* inner) {
#ifndef QF_KB
  if (n % 32 == 0 && n >= 32 && low.size() == 32 && cols <= 64 && block_fold_mfma_on()) {
    const uint64_t h = n / 32;
    launch_col_fold_mfma(c, cw, 1, 32, h, 32, col_fold_tables(c, low.data(), 32), inner, uint32_t(cols), R, h);
    return true;
  }
#endif
  return false;
}

void col_fold(DevCtx& c, const E8* cw, uint64_t R, uint64_t n, uint64_t cols, const std::vector<E8>& f, E8* out) {
#ifndef QF_KB
  if (cols % 4 == 0 && cols <= 128 && f.size() >= cols && block_fold_mfma_on()) {
    launch_col_fold_mfma(c, cw, R, 1, n, cols, col_fold_tables(c, f.data(), cols), out, 1, 0, 0);
    return;
  }
#endif
  E8* t = put_table(c, scratch_tab(), f.data(), f.size() * 16);
  hipLaunchKernelGGL(k_col_fold, grid_for(n), dim3(WG), 0, c.st, cw, R, n, cols, t, out);
}
std::vector<E8> gather_rows(DevCtx& c, const E8* cw, uint64_t R, uint64_t cols, const std::vector<uint64_t>& rows) {
  if (rows.empty()) return {};
  uint64_t* d_rows = static_cast<uint64_t*>(dmalloc(rows.size() * 8, c.st));
  DVec tmp(rows.size() * cols);
  h2d(c, d_rows, rows.data(), rows.size() * 8);
  uint64_t total = rows.size() * cols;
  hipLaunchKernelGGL(k_gather_rows, dim3(uint32_t((total + WG - 1) / WG)), dim3(WG), 0, c.st, cw, R, cols, d_rows,
                     uint64_t(rows.size()), tmp.p);
  auto out = download(c, tmp.p, total);
  dfree(d_rows, c.st);
  return out;
}
std::vector<E8> gather_weighted(DevCtx& c, const E8* a, uint64_t Ra, uint64_t ca, const E8* b, uint64_t Rb, uint64_t cb,
                                const std::vector<E8>& w, const std::vector<uint64_t>& rows) {
  require(w.size() >= ca + cb && (cb == 0 || b), QF_ERR_ARGUMENT, "gather weights");
  if (rows.empty()) return {};
  for (uint64_t r : rows) require(r < Ra && (cb == 0 || r < Rb), QF_ERR_ARGUMENT, "gather row out of range");
  uint64_t* d_rows = static_cast<uint64_t*>(dmalloc(rows.size() * 8, c.st));
  h2d(c, d_rows, rows.data(), rows.size() * 8);
  DVec d_w = upload(c, std::vector<E8>(w.begin(), w.begin() + (ca + cb)));
  DVec out(rows.size());
  hipLaunchKernelGGL(k_gather_weighted, dim3(uint32_t((rows.size() + WG - 1) / WG)), dim3(WG), 0, c.st, a, Ra, ca, b,
                     Rb, cb, d_w.p, d_rows, uint64_t(rows.size()), out.p);
  QF_HIP(hipGetLastError());
  auto r = download(c, out);
  dfree(d_rows, c.st);
  return r;
}
void append_row_records(DevCtx& c, Bytes& out, const E8* cw, uint64_t R, uint64_t cols, const uint32_t* tree,
                        const std::vector<uint64_t>& rows) {
  if (rows.empty()) return;
  for (uint64_t r : rows) require(r < R, QF_ERR_ARGUMENT, "query row out of range");
  const int depth = ilog2u(R);
  const uint64_t per = 16 + 16 * cols + 32 * uint64_t(depth), total = rows.size() * (cols + uint64_t(depth) + 1);
  uint64_t* d_rows = static_cast<uint64_t*>(dmalloc(rows.size() * 8, c.st));
  uint8_t* d_out = static_cast<uint8_t*>(dmalloc(rows.size() * per, c.st));
  h2d(c, d_rows, rows.data(), rows.size() * 8);
  hipLaunchKernelGGL(k_row_records, dim3(uint32_t((total + WG - 1) / WG)), dim3(WG), 0, c.st, cw, R, cols, tree, depth,
                     d_rows, uint64_t(rows.size()), d_out);
  QF_HIP(hipGetLastError());
  const size_t at = out.size();
  out.resize(at + rows.size() * per);
  d2h(c, out.data() + at, d_out, rows.size() * per);
  dfree(d_rows, c.st);
  dfree(d_out, c.st);
}
void scatter_add(DevCtx& c, E8* dst, const std::vector<uint64_t>& idx, const std::vector<E8>& vals) {
  if (idx.empty()) return;
  uint64_t* d_idx = static_cast<uint64_t*>(dmalloc(idx.size() * 8, c.st));
  DVec v = upload(c, vals);
  h2d(c, d_idx, idx.data(), idx.size() * 8);
  hipLaunchKernelGGL(k_scatter_add, dim3(uint32_t((idx.size() + WG - 1) / WG)), dim3(WG), 0, c.st, dst, d_idx, v.p,
                     uint64_t(idx.size()));
  dfree(d_idx, c.st);
}
}  // namespace dv

// ---------------------------------------------------------------- BVA-scale wrappers
static void upload_pairs(DevCtx& c, const std::vector<std::array<E8, 2>>& pairs, E8 (**out)[2]) {
  *out = reinterpret_cast<E8(*)[2]>(dmalloc(std::max<size_t>(pairs.size(), 1) * 32, c.st));
  h2d(c, *out, pairs.data(), pairs.size() * 32);
}

namespace {
// Each host thread has its own launch form, so a thread that times the
// per-matrix form never changes the form of proofs running on other threads.
thread_local int t_launch_mode = LAUNCH_BATCHED;
}
void set_launch_mode(int mode) { t_launch_mode = mode; }
int launch_mode() { return t_launch_mode; }

// Launches the batched projection: one workgroup per chunk of <= 2^16 rows of
// a task; returns the chunk lane sums (chunks x L, on the device) and each
// chunk's task in `owner`.
static DVec project_chunks(DevCtx& c, const Val* bva, uint64_t available, int L, const std::vector<ProjTask>& tasks,
                           const std::vector<std::array<E8, 2>>& pairs, std::vector<uint32_t>& owner) {
  // Chunks of 16-byte aligned rows that fill whole 16-byte records (F65521:
  // 16 and 8 lanes; KoalaBear: 16, 8 and 4) and give every wave at least one
  // round of MF_ROWS records per hi group run on k_tensor_project_mfma (listed
  // first), the rest on k_tensor_project.
  // QF_PROJECT_MFMA=0 keeps every chunk on k_tensor_project (same results);
  // read at every call, so a process can compare both forms.
  const bool use_mfma = [] {
    const char* e = std::getenv("QF_PROJECT_MFMA");
    return !(e && e[0] == '0');
  }();
  std::vector<ProjTaskDev> chunks, vchunks;
  std::vector<uint32_t> vowner;
  for (size_t i = 0; i < tasks.size(); i++) {
    const ProjTask& t = tasks[i];
    require(pow2(t.rows) && uint64_t(1) << t.nbits == t.rows, QF_ERR_ARGUMENT, "projection task rows");
    // records of REC entries: 2^rbits rows, or 1 / ns of a row; a hi group
    // holds 2^(k - rbits) <= 2^MF_TAB_LOG table entries, MF_WAVES MF_ROWS / ns
    // of them at least
    constexpr int REC = 16 / int(sizeof(Val));
    // F65521: 8 to 64 lanes (up to four records per row); KoalaBear: 4 to 32 (up to four)
    const bool shape = LANES == 8 ? (L >= 8 && L <= 64) : (L >= 4 && L <= 32);
    const uint32_t rbits = L < REC ? 1 : 0, nslog = L > REC ? uint32_t(ilog2u(uint64_t(L / REC))) : 0;
    const bool mfma = use_mfma && shape && t.base % REC == 0 && t.nbits + nslog >= 10 + rbits;
    uint32_t k = std::min<uint32_t>(t.nbits, mfma ? MF_TAB_LOG + rbits : PROJ_TAB_LOG);
    uint32_t clog = std::min<uint32_t>(t.nbits, 16);
    for (uint64_t first = 0; first < t.rows; first += uint64_t(1) << clog) {
      if (t.base + first * L >= available) break;  // rows past the active data contribute zero
      (mfma ? chunks : vchunks).push_back({t.base, t.rows, first, t.nbits, t.pair_off, k, clog, t.scale});
      (mfma ? owner : vowner).push_back(uint32_t(i));
    }
  }
  const size_t nm = chunks.size();
  chunks.insert(chunks.end(), vchunks.begin(), vchunks.end());
  owner.insert(owner.end(), vowner.begin(), vowner.end());
  if (chunks.empty()) return DVec();
  E8(*d_pairs)[2] = nullptr;
  upload_pairs(c, pairs, &d_pairs);
  auto* d_tasks = static_cast<ProjTaskDev*>(dmalloc(chunks.size() * sizeof(ProjTaskDev), c.st));
  h2d(c, d_tasks, chunks.data(), chunks.size() * sizeof(ProjTaskDev));
  DVec out(chunks.size() * L);
  auto* pr = (const E8(*)[2])d_pairs;
  for (size_t first = 0; first < chunks.size();) {  // at most 65535 workgroups per launch
    const bool mf = first < nm;
    const size_t end = std::min(mf ? nm : chunks.size(), first + 65535);
    const uint32_t n = uint32_t(end - first);
    auto* tk = d_tasks + first;
    E8* o = out.p + first * L;
    if (mf) {
      if (L == 16)
        hipLaunchKernelGGL(k_tensor_project_mfma<16>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o);
      else if (L == 8)
        hipLaunchKernelGGL(k_tensor_project_mfma<8>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o);
      else if (L == 32)
        hipLaunchKernelGGL(k_tensor_project_mfma<32>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o);
#ifndef QF_KB
      else if (L == 64)
        hipLaunchKernelGGL(k_tensor_project_mfma<64>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o);
#endif
#ifdef QF_KB
      else
        hipLaunchKernelGGL(k_tensor_project_mfma<4>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o);
#endif
    } else {
      switch (L) {
        case 2: hipLaunchKernelGGL(k_tensor_project<2>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o); break;
        case 4: hipLaunchKernelGGL(k_tensor_project<4>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o); break;
        case 8: hipLaunchKernelGGL(k_tensor_project<8>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o); break;
        case 16: hipLaunchKernelGGL(k_tensor_project<16>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o); break;
        case 32: hipLaunchKernelGGL(k_tensor_project<32>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o); break;
        case 64: hipLaunchKernelGGL(k_tensor_project<64>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o); break;
        default: hipLaunchKernelGGL(k_tensor_project<128>, dim3(n), dim3(WG), 0, c.st, bva, available, tk, pr, o); break;
      }
    }
    first = end;
  }
  QF_HIP(hipGetLastError());
  dfree(d_pairs, c.st);
  dfree(d_tasks, c.st);
  return out;
}

EoProjection eo_project(DevCtx& c, const Val* bva, uint64_t available, int s, const std::vector<ProjTask>& tasks,
                        const std::vector<std::array<E8, 2>>& pairs) {
  require(s >= 2 && s <= 64 && pow2(uint64_t(s)), QF_ERR_ARGUMENT, "packing stride");
  EoProjection pr;
  pr.s = s;
  if (launch_mode() != LAUNCH_PER_MATRIX || tasks.size() <= 1) {
    pr.parts = project_chunks(c, bva, available, 2 * s, tasks, pairs, pr.owner);
    return pr;
  }
  // the plain per-matrix form: one launch and one host round trip per bucket term
  std::vector<DVec> each;
  for (size_t i = 0; i < tasks.size(); i++) {
    ProjTask one = tasks[i];
    one.pair_off = 0;
    std::vector<std::array<E8, 2>> own(pairs.begin() + tasks[i].pair_off,
                                       pairs.begin() + tasks[i].pair_off + tasks[i].nbits);
    std::vector<uint32_t> ow;
    each.push_back(project_chunks(c, bva, available, 2 * s, {one}, own, ow));
    c.sync();
    pr.owner.insert(pr.owner.end(), ow.size(), uint32_t(i));
  }
  if (pr.owner.empty()) return pr;
  pr.parts = DVec(pr.owner.size() * 2 * uint64_t(s));
  uint64_t at = 0;
  for (auto& d : each)
    if (d.n) {
      dv::copy(c, pr.parts.p + at, d.p, d.n);
      at += d.n;
    }
  return pr;
}

EoCorr eo_correlations(DevCtx& c, const EoProjection& pr, const std::vector<std::array<E8, 2>>& f,
                       const std::vector<E8>& lows) {
  const int s = pr.s;
  EoCorr r;
  r.pos.assign(s - 1, E8{});
  r.neg.assign(s - 1, E8{});
  r.zero = E8{};
  if (pr.owner.empty()) return r;
  require(lows.size() == f.size() * size_t(s), QF_ERR_ARGUMENT, "projection low weights");
  const uint32_t nch = uint32_t(pr.owner.size());
  E8(*d_f)[2] = nullptr;
  upload_pairs(c, f, &d_f);
  auto* d_owner = static_cast<uint32_t*>(dmalloc(pr.owner.size() * 4, c.st));
  h2d(c, d_owner, pr.owner.data(), pr.owner.size() * 4);
  DVec joined(uint64_t(nch) * s);
  hipLaunchKernelGGL(k_lane_join, grid_for(uint64_t(nch) * s), dim3(WG), 0, c.st, pr.parts.p, d_owner, nch, s,
                     (const E8(*)[2])d_f, joined.p);
  DVec d_lows = dv::upload(c, lows);
  DVec out(2 * (s - 1) + 1);
  hipLaunchKernelGGL(k_proj_cross, dim3(2 * (s - 1) + 1), dim3(WG), 0, c.st, joined.p, d_owner, nch, s, d_lows.p,
                     out.p);
  QF_HIP(hipGetLastError());
  auto v = dv::download(c, out);
  dfree(d_owner, c.st);
  dfree(d_f, c.st);
  std::copy(v.begin(), v.begin() + (s - 1), r.pos.begin());
  std::copy(v.begin() + (s - 1), v.begin() + 2 * (s - 1), r.neg.begin());
  r.zero = v[2 * (s - 1)];
  return r;
}

std::pair<E8, E8> eo_first_cross(DevCtx& c, const EoProjection& pr, const std::vector<E8>& pinv,
                                 const std::vector<std::array<E8, 2>>& coef) {
  if (pr.owner.empty()) return {E8{}, E8{}};
  require(int(pinv.size()) == pr.s, QF_ERR_ARGUMENT, "witness weights");
  const uint32_t nch = uint32_t(pr.owner.size());
  auto* d_owner = static_cast<uint32_t*>(dmalloc(pr.owner.size() * 4, c.st));
  h2d(c, d_owner, pr.owner.data(), pr.owner.size() * 4);
  E8(*d_coef)[2] = nullptr;
  upload_pairs(c, coef, &d_coef);
  E8* t = put_table(c, scratch_tab(), pinv.data(), pinv.size() * 16);
  const uint32_t g = groups_for(nch);
  hipLaunchKernelGGL(k_first_cross, dim3(g), dim3(WG), 0, c.st, pr.parts.p, d_owner, nch, pr.s, t,
                     (const E8(*)[2])d_coef, c.partials);
  QF_HIP(hipGetLastError());
  auto sums = finish_sums<2>(c, g);
  dfree(d_owner, c.st);
  dfree(d_coef, c.st);
  return {sums[0], sums[1]};
}


OtfQuery::OtfQuery(DevCtx& c, const std::vector<std::array<E8, 2>>& factors, uint64_t max_chunks, size_t max_fast,
                   size_t max_pieces, size_t max_items)
    : f_(std::max<uint64_t>(2 * factors.size(), 2)),
      terms_(3 * std::max<uint64_t>(max_fast, 1)),
      desc_(3 * std::max<uint64_t>(max_chunks, 1)),
      tlo_(std::max<uint64_t>(max_fast, 1) * WG),
      thi_(std::max<uint64_t>(max_fast, 1) * 16),
      pieces_(4 * std::max<uint64_t>(max_pieces, 1)),
      items_(2 * std::max<uint64_t>(max_items, 1)),
      max_chunks_(max_chunks),
      max_fast_(max_fast),
      max_pieces_(std::max<size_t>(max_pieces, 1)),
      max_items_(std::max<size_t>(max_items, 1)) {
  h2d(c, f_.p, factors.data(), factors.size() * sizeof(factors[0]));
}

// Uploads the level's records, builds the chunk records and term tables, and
// returns the number of partial rows the sparse pass will use.
uint32_t OtfQuery::prepare(DevCtx& c, const OtfLevel& lv, uint64_t chunks, bool mfma_tables) {
  require(chunks <= max_chunks_ && lv.fast.size() <= max_fast_, QF_ERR_ARGUMENT, "on-the-fly query level size");
  if (lv.pieces.size() > max_pieces_) {  // later levels may split more terms than the first
    max_pieces_ = lv.pieces.size();
    pieces_ = DVec(4 * max_pieces_);
  }
  if (lv.items.size() > max_items_) {
    max_items_ = lv.items.size();
    items_ = DVec(2 * max_items_);
  }
  h2d(c, terms_.p, lv.fast.data(), lv.fast.size() * sizeof(OtfTermDev));
  h2d(c, pieces_.p, lv.pieces.data(), lv.pieces.size() * sizeof(OtfPieceDev));
  h2d(c, items_.p, lv.items.data(), lv.items.size() * sizeof(OtfItemDev));
  const uint32_t nt = uint32_t(lv.fast.size());
  const uint64_t threads = chunks + uint64_t(nt) * (256 + 16);
  const dim3 grid(uint32_t((threads + WG - 1) / WG));
  const auto* td = reinterpret_cast<const OtfTermDev*>(terms_.p);
  auto* dd = reinterpret_cast<OtfChunk*>(desc_.p);
  if (mfma_tables)
    hipLaunchKernelGGL(k_otf_desc<7>, grid, dim3(WG), 0, c.st, td, nt, (const E8(*)[2])f_.p, chunks, dd, tlo_.p, thi_.p);
  else
    hipLaunchKernelGGL(k_otf_desc<8>, grid, dim3(WG), 0, c.st, td, nt, (const E8(*)[2])f_.p, chunks, dd, tlo_.p, thi_.p);
  QF_HIP(hipGetLastError());
  if (lv.pieces.empty() && lv.items.empty()) return 0;
  const uint64_t want = std::max<uint64_t>(lv.pieces.size(), (lv.items.size() + WG - 1) / WG);
  return uint32_t(std::min<uint64_t>(want, 64));
}

// The sparse pass (gs workgroups, partial rows after the fold's `groups`) and the level's cross terms.
std::pair<E8, E8> OtfQuery::finish(DevCtx& c, const OtfLevel& lv, const E8* ow, uint32_t groups, uint32_t gs) {
  if (gs) {
    hipLaunchKernelGGL(k_otf_sparse, dim3(gs), dim3(WG), 0, c.st, ow, reinterpret_cast<const OtfPieceDev*>(pieces_.p),
                       uint32_t(lv.pieces.size()), reinterpret_cast<const OtfItemDev*>(items_.p),
                       uint32_t(lv.items.size()), (const E8(*)[2])f_.p, c.partials + uint64_t(groups) * PSTRIDE);
    QF_HIP(hipGetLastError());
  }
  auto sums = finish_sums<2>(c, groups + gs);
  return {sums[0], sums[1]};
}

std::pair<E8, E8> OtfQuery::first_fold(DevCtx& c, const Val* bva, uint64_t available, uint64_t rows, int s,
                                       const std::vector<E8>& pinv2, const OtfLevel& lv, E8* ow) {
  require(s >= 2 && s <= 64 && pow2(uint64_t(s)) && int(pinv2.size()) == 2 * s, QF_ERR_ARGUMENT, "witness stride");
  require(rows % (uint64_t(4) << OTF_LCP) == 0, QF_ERR_ARGUMENT, "witness rows");
  const uint64_t chunks = (rows / 4) >> OTF_LCP;
#ifndef QF_KB
  // rows of 32 to 128 entries run on the matrix cores (QF_PROJECT_MFMA=0: the vector-ALU form)
  const char* pm = std::getenv("QF_PROJECT_MFMA");
  const bool mfma = 2 * s >= 32 && !(pm && pm[0] == '0');
#else
  const bool mfma = false;
#endif
  const uint32_t gs = prepare(c, lv, chunks, mfma);
  E8* t = put_table(c, scratch_tab(), pinv2.data(), pinv2.size() * 16);
  const OtfChunk* d = reinterpret_cast<const OtfChunk*>(desc_.p);
  uint32_t g = 0;
#ifndef QF_KB
  if (mfma) {
#define QF_FIRST_FOLD_MF(S2)                                                                                      \
  {                                                                                                                \
    static const uint32_t resident = resident_groups(k_otf_first_fold_mfma<S2>);                                   \
    g = uint32_t(std::min<uint64_t>({chunks, resident, uint64_t(DevCtx::MAX_GROUPS - gs)}));                        \
    hipLaunchKernelGGL(k_otf_first_fold_mfma<S2>, dim3(g), dim3(WG), 0, c.st, bva, available, chunks, t, ow, d,    \
                       tlo_.p, thi_.p, c.partials);                                                                \
  }
    switch (2 * s) {
      case 32: QF_FIRST_FOLD_MF(32); break;
      case 64: QF_FIRST_FOLD_MF(64); break;
      default: QF_FIRST_FOLD_MF(128); break;
    }
#undef QF_FIRST_FOLD_MF
    QF_HIP(hipGetLastError());
    return finish(c, lv, ow, g, gs);
  }
#endif
#define QF_FIRST_FOLD(S2)                                                                                         \
  {                                                                                                                \
    static const uint32_t resident = resident_groups(k_otf_first_fold<S2>);                                        \
    g = uint32_t(std::min<uint64_t>({chunks, resident, uint64_t(DevCtx::MAX_GROUPS - gs)}));                        \
    hipLaunchKernelGGL(k_otf_first_fold<S2>, dim3(g), dim3(WG), 0, c.st, bva, available, chunks, t, ow, d, tlo_.p, \
                       thi_.p, c.partials);                                                                        \
  }
  switch (2 * s) {
    case 4: QF_FIRST_FOLD(4); break;
    case 8: QF_FIRST_FOLD(8); break;
    case 16: QF_FIRST_FOLD(16); break;
    case 32: QF_FIRST_FOLD(32); break;
    case 64: QF_FIRST_FOLD(64); break;
    default: QF_FIRST_FOLD(128); break;
  }
#undef QF_FIRST_FOLD
  QF_HIP(hipGetLastError());
  return finish(c, lv, ow, g, gs);
}

std::pair<E8, E8> OtfQuery::fold(DevCtx& c, const E8* w, E8* ow, uint64_t h, const E8& ri, const OtfLevel& lv) {
  require(h % (uint64_t(2) << OTF_LCP) == 0, QF_ERR_ARGUMENT, "on-the-fly fold length");
  const uint64_t chunks = (h / 2) >> OTF_LCP;
  const uint32_t gs = prepare(c, lv, chunks, false);
  static const uint32_t resident = resident_groups(k_otf_fold);
  const uint32_t g = uint32_t(std::min<uint64_t>({chunks, resident, uint64_t(DevCtx::MAX_GROUPS - gs)}));
  hipLaunchKernelGGL(k_otf_fold, dim3(g), dim3(WG), 0, c.st, w, ow, chunks, ri,
                     reinterpret_cast<const OtfChunk*>(desc_.p), tlo_.p, thi_.p, c.partials);
  QF_HIP(hipGetLastError());
  return finish(c, lv, ow, g, gs);
}

// One launch over every task (the batched form).
// out[t L + l] = sum of parts[ch L + l] over task t's chunks [r[2 t], r[2 t + 1]); one
// workgroup per task, WG / L chunk slots per lane.
__global__ __launch_bounds__(WG) void k_sum_chunks(const E8* __restrict__ parts, int L,
                                                   const uint32_t* __restrict__ ranges, E8* __restrict__ out) {
  __shared__ uint32_t red[WG][LANES];
  const uint32_t t = blockIdx.x, l = threadIdx.x % uint32_t(L), slot = threadIdx.x / uint32_t(L),
                 slots = WG / uint32_t(L);
  V8 s = v8_zero();
  for (uint32_t ch = ranges[2 * t] + slot; ch < ranges[2 * t + 1]; ch += slots)
    s = v8_add(s, v8_load(parts[uint64_t(ch) * L + l]));
  for (int k = 0; k < LANES; k++) red[threadIdx.x][k] = s.c[k];
  __syncthreads();
  for (uint32_t h = slots / 2; h > 0; h /= 2) {  // slots is a power of two
    if (slot < h)
      for (int k = 0; k < LANES; k++) red[threadIdx.x][k] = madd(red[threadIdx.x][k], red[threadIdx.x + h * L][k]);
    __syncthreads();
  }
  if (slot == 0) {
    V8 x;
    for (int k = 0; k < LANES; k++) x.c[k] = red[threadIdx.x][k];
    out[uint64_t(t) * L + l] = v8_store(x);
  }
}

static std::vector<E8> tensor_project_launch(DevCtx& c, const Val* bva, uint64_t available, int L,
                                             const std::vector<ProjTask>& tasks,
                                             const std::vector<std::array<E8, 2>>& pairs) {
  std::vector<uint32_t> owner;
  DVec out = project_chunks(c, bva, available, L, tasks, pairs, owner);
  std::vector<E8> result(tasks.size() * L, E8{});
  if (owner.empty()) return result;
  if (owner.size() > 4 * tasks.size() && WG % L == 0) {
    // many chunks per task: sum them on the device. A task's chunks are consecutive
    // (all of them on the matrix cores or all on the vector ALUs, appended in task order).
    std::vector<uint32_t> ranges(2 * tasks.size(), 0);
    std::vector<uint8_t> seen(tasks.size(), 0);
    for (size_t ch = 0; ch < owner.size(); ch++) {
      const uint32_t t = owner[ch];
      if (!seen[t]) {
        ranges[2 * t] = uint32_t(ch);
        seen[t] = 1;
      }
      require(ranges[2 * t + 1] == 0 || ranges[2 * t + 1] == ch, QF_ERR_ARGUMENT, "projection chunks out of order");
      ranges[2 * t + 1] = uint32_t(ch + 1);
    }
    auto* d_r = static_cast<uint32_t*>(dmalloc(ranges.size() * 4, c.st));
    h2d(c, d_r, ranges.data(), ranges.size() * 4);
    DVec sums(tasks.size() * L);
    hipLaunchKernelGGL(k_sum_chunks, dim3(uint32_t(tasks.size())), dim3(WG), 0, c.st, out.p, L, d_r, sums.p);
    QF_HIP(hipGetLastError());
    result = dv::download(c, sums);
    dfree(d_r, c.st);
    return result;
  }
  auto parts = dv::download(c, out);
  for (size_t ch = 0; ch < owner.size(); ch++)
    for (int l = 0; l < L; l++) result[owner[ch] * L + l] = e8_add(result[owner[ch] * L + l], parts[ch * L + l]);
  return result;
}

std::vector<E8> tensor_project(DevCtx& c, const Val* bva, uint64_t available, int L, const std::vector<ProjTask>& tasks,
                               const std::vector<std::array<E8, 2>>& pairs) {
  require(L >= 2 && L <= 64 && pow2(uint64_t(L)), QF_ERR_ARGUMENT, "projection lane count");
  if (launch_mode() != LAUNCH_PER_MATRIX || tasks.size() <= 1) return tensor_project_launch(c, bva, available, L, tasks, pairs);
  // the plain per-matrix form: one launch (and one round trip) per bucket term
  std::vector<E8> result;
  for (const ProjTask& t : tasks) {
    ProjTask one = t;
    one.pair_off = 0;
    std::vector<std::array<E8, 2>> own(pairs.begin() + t.pair_off, pairs.begin() + t.pair_off + t.nbits);
    auto r = tensor_project_launch(c, bva, available, L, {one}, own);
    result.insert(result.end(), r.begin(), r.end());
  }
  return result;
}

void eo_witness(DevCtx& c, const Val* bva, uint64_t available, uint64_t n_rows, int s, const std::vector<E8>& pinv,
                E8* w) {
  require(s >= 2 && s <= 64 && pow2(s) && int(pinv.size()) == s, QF_ERR_ARGUMENT, "witness stride");
  E8* t = put_table(c, scratch_tab(), pinv.data(), pinv.size() * 16);
  switch (s) {
    case 2: hipLaunchKernelGGL(k_eo_witness<2>, grid_for(n_rows), dim3(WG), 0, c.st, bva, available, n_rows, t, w); break;
    case 4: hipLaunchKernelGGL(k_eo_witness<4>, grid_for(n_rows), dim3(WG), 0, c.st, bva, available, n_rows, t, w); break;
    case 8: hipLaunchKernelGGL(k_eo_witness<8>, grid_for(n_rows), dim3(WG), 0, c.st, bva, available, n_rows, t, w); break;
    case 16: hipLaunchKernelGGL(k_eo_witness<16>, grid_for(n_rows), dim3(WG), 0, c.st, bva, available, n_rows, t, w); break;
    case 32: hipLaunchKernelGGL(k_eo_witness<32>, grid_for(n_rows), dim3(WG), 0, c.st, bva, available, n_rows, t, w); break;
    default: hipLaunchKernelGGL(k_eo_witness<64>, grid_for(n_rows), dim3(WG), 0, c.st, bva, available, n_rows, t, w); break;
  }
  QF_HIP(hipGetLastError());
}

// One launch over every chunk of `tasks`. With ADD each chunk reads and
// rewrites its rows, so the tasks of one launch must not share rows.
template <bool ADD>
static void fill_launch(DevCtx& c, E8* out, const std::vector<FillTask>& tasks,
                        const std::vector<std::array<E8, 2>>& pairs) {
  std::vector<FillTaskDev> chunks;
  for (const FillTask& t : tasks) {
    uint32_t k = std::min<uint32_t>(t.nbits, 12);
    uint32_t clog = std::min<uint32_t>(t.nbits, 16);
    for (uint64_t first = 0; first < (uint64_t(1) << t.nbits); first += uint64_t(1) << clog)
      chunks.push_back({t.start, first, t.nbits, t.pair_off, k, clog, t.scale});
  }
  if (chunks.empty()) return;
  E8(*d_pairs)[2] = nullptr;
  FillTaskDev* d_tasks = nullptr;
  upload_pairs(c, pairs, &d_pairs);
  d_tasks = static_cast<FillTaskDev*>(dmalloc(chunks.size() * sizeof(FillTaskDev), c.st));
  h2d(c, d_tasks, chunks.data(), chunks.size() * sizeof(FillTaskDev));
  for (size_t first = 0; first < chunks.size(); first += 65535) {
    uint32_t n = uint32_t(std::min<size_t>(65535, chunks.size() - first));
    hipLaunchKernelGGL(k_tensor_fill<ADD>, dim3(n), dim3(FWG), 0, c.st, out, d_tasks + first, (const E8(*)[2])d_pairs);
  }
  QF_HIP(hipGetLastError());
  dfree(d_pairs, c.st);
  dfree(d_tasks, c.st);
}

template <bool ADD>
static void fill_tasks(DevCtx& c, E8* out, const std::vector<FillTask>& tasks,
                       const std::vector<std::array<E8, 2>>& pairs) {
  if (launch_mode() == LAUNCH_PER_MATRIX && tasks.size() > 1) {
    // one launch per term; a single task never shares rows with another in its launch
    for (const FillTask& t : tasks) {
      FillTask one = t;
      one.pair_off = 0;
      std::vector<std::array<E8, 2>> own(pairs.begin() + t.pair_off, pairs.begin() + t.pair_off + t.nbits);
      fill_launch<ADD>(c, out, {one}, own);
    }
    return;
  }
  if (!ADD) {  // written tasks have disjoint rows (tensor_fill checks)
    fill_launch<ADD>(c, out, tasks, pairs);
    return;
  }
  // Added tasks can share rows: the column-padding rectangles of one bucket
  // each span the bucket's rows, with zero weights outside their own columns.
  // Tasks go to launches whose row ranges are pairwise disjoint (greedy
  // interval colouring by start), run one after another on the stream.
  std::vector<size_t> order(tasks.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return tasks[a].start < tasks[b].start; });
  std::vector<std::vector<FillTask>> groups;
  std::vector<uint64_t> group_end;
  for (size_t i : order) {
    const FillTask& t = tasks[i];
    size_t g = 0;
    while (g < groups.size() && group_end[g] > t.start) g++;
    if (g == groups.size()) {
      groups.emplace_back();
      group_end.push_back(0);
    }
    groups[g].push_back(t);
    group_end[g] = t.start + (uint64_t(1) << t.nbits);
  }
  for (auto& g : groups) fill_launch<ADD>(c, out, g, pairs);
}

void tensor_fill_add(DevCtx& c, E8* out, const std::vector<FillTask>& tasks, const std::vector<std::array<E8, 2>>& pairs) {
  fill_tasks<true>(c, out, tasks, pairs);
}

void tensor_fill(DevCtx& c, E8* out, uint64_t n, const std::vector<FillTask>& tasks,
                 const std::vector<std::array<E8, 2>>& pairs) {
  // the rows of [0, n) outside every task, in chunks of at most 2^16 rows
  std::vector<std::pair<uint64_t, uint64_t>> spans;
  for (const FillTask& t : tasks) spans.push_back({t.start, t.start + (uint64_t(1) << t.nbits)});
  std::sort(spans.begin(), spans.end());
  std::vector<uint64_t> gaps;
  uint64_t at = 0;
  auto gap = [&](uint64_t a, uint64_t b) {
    for (; a < b; a += std::min<uint64_t>(b - a, uint64_t(1) << 16)) {
      gaps.push_back(a);
      gaps.push_back(std::min<uint64_t>(b - a, uint64_t(1) << 16));
    }
  };
  for (auto& [a, b] : spans) {
    require(a >= at && b <= n, QF_ERR_ARGUMENT, "fill tasks overlap or leave the vector");
    gap(at, a);
    at = b;
  }
  gap(at, n);
  if (!gaps.empty()) {
    auto* d_gaps = static_cast<uint64_t*>(dmalloc(gaps.size() * 8, c.st));
    h2d(c, d_gaps, gaps.data(), gaps.size() * 8);
    for (size_t first = 0; first < gaps.size() / 2; first += 65535) {
      uint32_t cnt = uint32_t(std::min<size_t>(65535, gaps.size() / 2 - first));
      hipLaunchKernelGGL(k_zero_ranges, dim3(cnt), dim3(WG), 0, c.st, out, d_gaps + 2 * first);
    }
    QF_HIP(hipGetLastError());
    dfree(d_gaps, c.st);
  }
  fill_tasks<false>(c, out, tasks, pairs);
}

void block_fold(DevCtx& c, const Val* bva, uint64_t available, uint64_t block_len, uint64_t blocks,
                const std::vector<E8>& fw, E8* values, uint64_t n) {
  require(blocks <= 128 && fw.size() == blocks, QF_ERR_ARGUMENT, "block fold weights");
#ifndef QF_KB
  if (block_fold_mfma_on()) {
    // B operands per step and lane (n + 16 q): digits of M_b[c][j], b = 4 s + q
    const uint32_t steps = uint32_t((blocks + 3) / 4);
    std::vector<uint64_t> bops(size_t(steps) * 64, 0);
    for (uint64_t b = 0; b < blocks; b++) {
      E8 col[8];
      for (int j = 0; j < 8; j++) col[j] = fw[b] * e8_basis(j);
      for (int nn = 0; nn < 16; nn++) {
        uint64_t x = 0;
        for (int j = 0; j < 8; j++) {
          const int32_t t = int32_t(col[j].c[nn & 7]);
          const int32_t tv = t <= 32639 ? t : t - int32_t(P);
          const int32_t hd = (tv + 128) >> 8, ld = tv - 256 * hd;
          x |= uint64_t(uint8_t(int8_t(nn < 8 ? ld : hd))) << (8 * j);
        }
        bops[(b / 4) * 64 + 16 * (b % 4) + nn] = x;
      }
    }
    const long* t = reinterpret_cast<const long*>(put_table(c, scratch_tab(), bops.data(), bops.size() * 8));
    static const uint32_t resident = resident_groups(k_block_fold_mfma);
    const uint64_t g = std::min<uint64_t>(resident, (n + WG - 1) / WG);
    hipLaunchKernelGGL(k_block_fold_mfma, dim3(uint32_t(std::max<uint64_t>(g, 1))), dim3(WG), 0, c.st,
                       reinterpret_cast<const uint8_t*>(bva), available * sizeof(Val), block_len * sizeof(Val), steps,
                       t, values, n);
    QF_HIP(hipGetLastError());
    return;
  }
#endif
  E8* t = put_table(c, scratch_tab(), fw.data(), fw.size() * 16);
  hipLaunchKernelGGL(k_block_fold, grid_for(n), dim3(WG), 0, c.st, reinterpret_cast<const uint8_t*>(bva),
                     available * sizeof(Val), block_len * sizeof(Val), uint32_t(blocks), t, values, n);
  QF_HIP(hipGetLastError());
}

void weight_images(DevCtx& c, E8* out, const DVec& hi, const DVec& lo, const std::array<E8, LANES>& images,
                   hipStream_t st) {
  require(pow2(hi.n) && pow2(lo.n), QF_ERR_ARGUMENT, "weight tables");
  Images im;
  for (int j = 0; j < LANES; j++) im.m[j] = images[j];
  uint64_t n = hi.n * lo.n;
  hipLaunchKernelGGL(k_weight_images, grid_for(n), dim3(WG), 0, st ? st : c.st, out, n, hi.p, lo.p, ilog2u(lo.n),
                     im);
  QF_HIP(hipGetLastError());
}

namespace dv {
// The sumcheck message's block sums go past the rows every other reduction
// uses in the thread's pinned partial buffer (MAX_GROUPS rows of PSTRIDE), so
// kernels on the prover's own stream may write theirs meanwhile.
constexpr uint64_t SIDE_PARTIALS = uint64_t(DevCtx::MAX_GROUPS) * PSTRIDE;
SideMsg sc_msg_launch(DevCtx& c, hipStream_t st, const E8* v, const E8* w, uint64_t h) {
  SideMsg m;
  m.groups = groups_for(h);
  hipLaunchKernelGGL(k_reduce<R_SC>, dim3(m.groups), dim3(WG), 0, st, v, w, h, c.partials + SIDE_PARTIALS);
  QF_HIP(hipGetLastError());
  QF_HIP(hipEventCreateWithFlags(&m.done, hipEventDisableTiming));
  QF_HIP(hipEventRecord(m.done, st));
  return m;
}
std::pair<E8, E8> sc_msg_finish(DevCtx& c, SideMsg& m) {
  QF_HIP(hipEventSynchronize(m.done));
  QF_HIP(hipEventDestroy(m.done));
  m.done = nullptr;
  V8 s0 = v8_zero(), s1 = v8_zero();
  const uint32_t* hp = c.host_partials + SIDE_PARTIALS;
  for (uint32_t g = 0; g < m.groups; g++)
    for (int k = 0; k < LANES; k++) {
      s0.c[k] = madd(s0.c[k], hp[g * 16 + k]);
      s1.c[k] = madd(s1.c[k], hp[g * 16 + LANES + k]);
    }
  return {v8_store(s0), v8_store(s1)};
}
}  // namespace dv

}  // namespace qf
