// HIP kernels of the commitment phase for gfx950 (CDNA4).
//
// One PCS block of 2^c signed bytes becomes a 32-column codeword of Ext8
// elements, column-major in HBM: column k occupies entries [k*R, (k+1)*R),
// R = 2n rows, n = 2^(c-8) message rows. The message half of each column is
// the Mobius-transformed packed witness (packed.rs:220-258, :708-752); the
// parity half is WHT(alpha_1 . WHT(m)) (qafold_short.rs:969-1000). Rows are
// hashed into BLAKE2s leaves and a binary Merkle tree
// (qafold_short_pc.rs:2372-2516).
//
// Transforms are tensor products of 2x2 maps on the index bits, so a pass may
// apply any subset of bits: each workgroup stages a tile of C contiguous
// elements x 2^t elements spaced 2^s apart into LDS and runs t butterfly
// stages there. Columns of 2^22 elements take two passes (low 12 bits
// contiguous, high 10 bits strided with C = 4 -> 64-byte runs).
#include "qf_kernels.hpp"

namespace qf {

constexpr int TILE = 4096;   // elements per LDS tile (64 KiB)
constexpr int WG = 256;

enum { OP_MOBIUS = 0, OP_WHT = 1 };

template <int OP>
__device__ inline void lds_stages(E8* t, int C, int tbits) {
  const int pairs = (C << tbits) >> 1;
  for (int q = 0; q < tbits; q++) {
    for (int p = threadIdx.x; p < pairs; p += WG) {
      int l = p % C;
      int rest = p / C;                                   // h with bit q removed
      int h = ((rest >> q) << (q + 1)) | (rest & ((1 << q) - 1));
      int e0 = l + C * h, e1 = e0 + (C << q);
      V8 a = v8_load(t[e0]), b = v8_load(t[e1]);
      if (OP == OP_WHT) {
        t[e0] = v8_store(v8_add(a, b));
        t[e1] = v8_store(v8_sub(a, b));
      } else {
        t[e1] = v8_store(v8_sub(b, a));
      }
    }
    __syncthreads();
  }
}

// Geometry of a pass over `vectors` vectors of length 2^L, transforming bits [s, s+t).
struct PassGeom {
  uint64_t in_vstride, out_vstride;  // element strides between vectors
  int L, s, t, C;                    // C contiguous elements per row of the tile
};

__device__ inline void tile_coords(const PassGeom& g, uint64_t tile, uint64_t& vec, uint64_t& base) {
  uint64_t lo_chunks = (uint64_t(1) << g.s) / g.C;
  uint64_t hi_count = uint64_t(1) << (g.L - g.s - g.t);
  uint64_t lc = tile % lo_chunks;
  uint64_t rest = tile / lo_chunks;
  uint64_t hi = rest % hi_count;
  vec = rest / hi_count;
  base = lc * g.C + (hi << (g.s + g.t));
}

// Generic in-LDS butterfly pass; optionally multiplies by alpha[j] and repeats
// the same stages (the fused middle of WHT(alpha . WHT(.))).
template <int OP, bool ALPHA>
__global__ __launch_bounds__(WG) void k_pass(const E8* __restrict__ in, E8* __restrict__ out, PassGeom g,
                                             const E8* __restrict__ alpha) {
  __shared__ E8 t[TILE];
  uint64_t vec, base;
  tile_coords(g, blockIdx.x, vec, base);
  const int n = g.C << g.t;
  const E8* src = in + vec * g.in_vstride;
  E8* dst = out + vec * g.out_vstride;
  for (int e = threadIdx.x; e < n; e += WG) {
    uint64_t j = base + (e % g.C) + (uint64_t(e / g.C) << g.s);
    t[e] = src[j];
  }
  __syncthreads();
  lds_stages<OP>(t, g.C, g.t);
  if (ALPHA) {
    for (int e = threadIdx.x; e < n; e += WG) {
      uint64_t j = base + (e % g.C) + (uint64_t(e / g.C) << g.s);
      t[e] = v8_store(v8_mul(v8_load(t[e]), v8_load(alpha[j])));
    }
    __syncthreads();
    lds_stages<OP>(t, g.C, g.t);
  }
  for (int e = threadIdx.x; e < n; e += WG) {
    uint64_t j = base + (e % g.C) + (uint64_t(e / g.C) << g.s);
    dst[j] = t[e];
  }
}

// Pack 8 signed bytes per element (lane j = byte j mod p; bytes at or past
// `available` read as zero), run the Mobius stages on the low `t` bits of the
// element index inside a contiguous tile, and write the message half.
__global__ __launch_bounds__(WG) void k_pack_mobius(const int8_t* __restrict__ bva, uint64_t block_first,
                                                    uint64_t available, E8* __restrict__ cw, int log_n,
                                                    uint64_t rows, int tbits) {
  __shared__ E8 t[TILE];
  const int n = 1 << tbits;
  uint64_t first = uint64_t(blockIdx.x) << tbits;  // packed element index within the block
  for (int e = threadIdx.x; e < n; e += WG) {
    uint64_t i = first + e;
    uint64_t byte = block_first + 8 * i;
    E8 v;
    if (byte + 8 <= available) {
      uint2 w = *reinterpret_cast<const uint2*>(bva + byte);
      uint32_t b[2] = {w.x, w.y};
#pragma unroll
      for (int k = 0; k < 8; k++) {
        int x = int8_t(b[k >> 2] >> (8 * (k & 3)));
        v.c[k] = uint16_t(x < 0 ? x + int(P) : x);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) {
        int x = byte + k < available ? int(bva[byte + k]) : 0;
        v.c[k] = uint16_t(x < 0 ? x + int(P) : x);
      }
    }
    t[e] = v;
  }
  __syncthreads();
  lds_stages<OP_MOBIUS>(t, 1, tbits);
  uint64_t col = first >> log_n;
  uint64_t row0 = first & ((uint64_t(1) << log_n) - 1);
  for (int e = threadIdx.x; e < n; e += WG) cw[col * rows + row0 + e] = t[e];
}

// Mobius over the 5 column bits: for each message row, the 32 column entries.
// Thread = (row, 32-bit word of the 16-byte element).
__global__ __launch_bounds__(WG) void k_mobius_columns(E8* __restrict__ cw, uint64_t n, uint64_t rows) {
  uint64_t gid = uint64_t(blockIdx.x) * WG + threadIdx.x;
  uint64_t row = gid >> 2;
  int word = int(gid & 3);
  if (row >= n) return;
  uint32_t* base = reinterpret_cast<uint32_t*>(cw) + row * 4 + word;
  const uint64_t cstride = rows * 4;
  uint32_t lo[32], hi[32];
#pragma unroll
  for (int k = 0; k < 32; k++) {
    uint32_t w = base[k * cstride];
    lo[k] = w & 0xffffu;
    hi[k] = w >> 16;
  }
#pragma unroll
  for (int h = 1; h < 32; h *= 2)
#pragma unroll
    for (int k = 0; k < 32; k++)
      if (k & h) { lo[k] = msub(lo[k], lo[k ^ h]); hi[k] = msub(hi[k], hi[k ^ h]); }
#pragma unroll
  for (int k = 0; k < 32; k++) base[k * cstride] = lo[k] | (hi[k] << 16);
}

// Leaf hash of one codeword row: BLAKE2s(prefix75 || row bytes), 587 bytes,
// 10 compressions of which the first (all prefix) is precomputed on the host.
__global__ __launch_bounds__(WG) void k_leaf_hash(const E8* __restrict__ cw, uint64_t rows, LeafPrefix lp,
                                                  uint32_t* __restrict__ leaves) {
  uint64_t r = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (r >= rows) return;
  uint32_t R[129];
#pragma unroll
  for (int k = 0; k < 32; k++) {
    uint4 w = *reinterpret_cast<const uint4*>(cw + k * rows + r);
    R[4 * k] = w.x; R[4 * k + 1] = w.y; R[4 * k + 2] = w.z; R[4 * k + 3] = w.w;
  }
  R[128] = 0;
  uint32_t h[8];
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] = lp.h[i];
#pragma unroll
  for (int b = 1; b < 10; b++) {
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int w = 16 * b + i;
      uint32_t x;
      if (w == 16) x = lp.tail[0];
      else if (w == 17) x = lp.tail[1];
      else if (w == 18) x = lp.tail[2] | (R[0] << 24);
      else if (w <= 146) x = (R[w - 19] >> 8) | (R[w - 18] << 24);
      else x = 0;
      m[i] = x;
    }
    b2s_compress(h, m, b == 9 ? 587u : uint32_t(64 * (b + 1)), b == 9);
  }
  uint4* o = reinterpret_cast<uint4*>(leaves + 8 * r);
  o[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

// One Merkle level: parent i = BLAKE2s(prefix50 || child 2i || child 2i+1), 114 bytes.
__global__ __launch_bounds__(WG) void k_inner_level(const uint32_t* __restrict__ children, uint64_t parents,
                                                    InnerPrefix ip, uint32_t* __restrict__ out) {
  uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (i >= parents) return;
  uint32_t c[16];
  const uint4* s = reinterpret_cast<const uint4*>(children + 16 * i);
#pragma unroll
  for (int q = 0; q < 4; q++) { uint4 w = s[q]; c[4 * q] = w.x; c[4 * q + 1] = w.y; c[4 * q + 2] = w.z; c[4 * q + 3] = w.w; }
  uint32_t h[8];
  b2s_init(h);
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 12; k++) m[k] = ip.w[k];
  m[12] = ip.w[12] | (c[0] << 16);
  m[13] = (c[0] >> 16) | (c[1] << 16);
  m[14] = (c[1] >> 16) | (c[2] << 16);
  m[15] = (c[2] >> 16) | (c[3] << 16);
  b2s_compress(h, m, 64, false);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const int q = k + 4;  // message word 16 + k = 12 + q
    m[k] = q < 16 ? ((c[q - 1] >> 16) | (c[q] << 16)) : (q == 16 ? (c[15] >> 16) : 0u);
  }
  b2s_compress(h, m, 114, true);
  uint4* o = reinterpret_cast<uint4*>(out + 8 * i);
  o[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

// ---------------------------------------------------------------- launchers
static void launch_pass_geom(const E8* in, E8* out, uint64_t vectors, uint64_t in_vs, uint64_t out_vs, int L, int s,
                             int t, int op, const E8* alpha, hipStream_t st) {
  PassGeom g;
  g.in_vstride = in_vs;
  g.out_vstride = out_vs;
  g.L = L;
  g.s = s;
  g.t = t;
  int C = 1;
  while (s > 0 && C < (1 << s) && (C << 1) * (1 << t) <= TILE && C < 16) C <<= 1;
  g.C = C;
  require((C << t) <= TILE && t >= 1 && s + t <= L, QF_ERR_ARGUMENT, "pass geometry");
  uint64_t tiles = vectors * ((uint64_t(1) << s) / C) * (uint64_t(1) << (L - s - t));
  require(tiles < (1ull << 31), QF_ERR_ARGUMENT, "pass grid too large");
  dim3 grid{uint32_t(tiles), 1, 1}, block{uint32_t(WG), 1, 1};
  if (op == OP_MOBIUS) hipLaunchKernelGGL((k_pass<OP_MOBIUS, false>), grid, block, 0, st, in, out, g, alpha);
  else if (alpha) hipLaunchKernelGGL((k_pass<OP_WHT, true>), grid, block, 0, st, in, out, g, alpha);
  else hipLaunchKernelGGL((k_pass<OP_WHT, false>), grid, block, 0, st, in, out, g, alpha);
}

// Splits the L bits of each vector into LDS-sized passes (low bits contiguous).
static std::vector<std::pair<int, int>> bit_groups(int L) {
  std::vector<std::pair<int, int>> gs;
  int lo = L < 12 ? L : 12;
  if (lo > 0) gs.push_back({0, lo});
  int s = lo;
  while (s < L) { int t = L - s < 10 ? L - s : 10; gs.push_back({s, t}); s += t; }
  return gs;
}

void launch_commit_block(const CommitGeom& g, const int8_t* bva, uint64_t block_first, uint64_t available, E8* cw,
                         const E8* alpha, uint32_t* tree, hipStream_t st, hipEvent_t* ev) {
  const int L = g.log_n;
  if (ev) QF_HIP(hipEventRecord(ev[0], st));
  const uint64_t n = uint64_t(1) << L, R = 2 * n;
  const int tp = L < 12 ? L : 12;
  // 1. pack + Mobius on the low bits within each column
  uint64_t packed = n * 32;
  hipLaunchKernelGGL(k_pack_mobius, dim3(uint32_t(packed >> tp)), dim3(WG), 0, st, bva, block_first, available, cw, L,
                     R, tp);
  // 2. remaining Mobius bits within each column
  for (auto& gp : bit_groups(L))
    if (gp.first > 0) launch_pass_geom(cw, cw, 32, R, R, L, gp.first, gp.second, OP_MOBIUS, nullptr, st);
  // 3. Mobius over the column bits
  hipLaunchKernelGGL(k_mobius_columns, dim3(uint32_t((n * 4 + WG - 1) / WG)), dim3(WG), 0, st, cw, n, R);
  // 4. parity = WHT(alpha . WHT(m)): low bits copy m into the parity half, the
  //    high bits run forward, multiply, forward again, then the low bits again.
  auto groups = bit_groups(L);
  if (groups.size() == 1) {
    launch_pass_geom(cw, cw + n, 32, R, R, L, 0, groups[0].second, OP_WHT, alpha, st);
  } else {
    launch_pass_geom(cw, cw + n, 32, R, R, L, 0, groups[0].second, OP_WHT, nullptr, st);
    for (size_t i = 1; i + 1 < groups.size(); i++)
      launch_pass_geom(cw + n, cw + n, 32, R, R, L, groups[i].first, groups[i].second, OP_WHT, nullptr, st);
    launch_pass_geom(cw + n, cw + n, 32, R, R, L, groups.back().first, groups.back().second, OP_WHT, alpha, st);
    for (size_t i = groups.size() - 1; i-- > 1;)
      launch_pass_geom(cw + n, cw + n, 32, R, R, L, groups[i].first, groups[i].second, OP_WHT, nullptr, st);
    launch_pass_geom(cw + n, cw + n, 32, R, R, L, 0, groups[0].second, OP_WHT, nullptr, st);
  }
  // 5. leaves and inner levels; tree level l occupies (R >> l) digests after the previous levels.
  if (ev) QF_HIP(hipEventRecord(ev[1], st));
  hipLaunchKernelGGL(k_leaf_hash, dim3(uint32_t((R + WG - 1) / WG)), dim3(WG), 0, st, cw, R, g.leaf, tree);
  if (ev) QF_HIP(hipEventRecord(ev[2], st));
  uint32_t* level = tree;
  for (uint64_t cnt = R; cnt > 1; cnt /= 2) {
    uint32_t* next = level + 8 * cnt;
    hipLaunchKernelGGL(k_inner_level, dim3(uint32_t((cnt / 2 + WG - 1) / WG)), dim3(WG), 0, st, level, cnt / 2, g.inner,
                       next);
    level = next;
  }
  if (ev) QF_HIP(hipEventRecord(ev[3], st));
}

}  // namespace qf
