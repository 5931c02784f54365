// HIP kernels of the commitment phase for gfx950 (CDNA4).
//
// One PCS block of 2^c signed entries (i8 for F65521, i16 for KoalaBear)
// becomes a 32-column codeword of extension elements, column-major in HBM: column k occupies entries [k*R, (k+1)*R),
// R = 2n rows, n = 2^c / (32 LANES) message rows. The message half of each column is
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
      t[e] = v8_store(v8_mul_enc(v8_load(t[e]), v8_load(alpha[j])));
    }
    __syncthreads();
    lds_stages<OP>(t, g.C, g.t);
  }
  for (int e = threadIdx.x; e < n; e += WG) {
    uint64_t j = base + (e % g.C) + (uint64_t(e / g.C) << g.s);
    dst[j] = t[e];
  }
}

// The packed element whose LANES entries start at BVA byte `byte`: lane j =
// entry j mod p (field.rs:24-26, packed.rs:230-243); bytes at or past
// `available` read as zero. Both fields pack 8 bytes per element.
__device__ inline E8 packed_element(const uint8_t* __restrict__ bva, uint64_t byte, uint64_t available) {
  uint32_t w[2] = {0, 0};
  if (byte + 8 <= available) {
    const uint2 u = *reinterpret_cast<const uint2*>(bva + byte);
    w[0] = u.x;
    w[1] = u.y;
  } else {
    for (int j = 0; j < 8; j++)
      if (byte + j < available) w[j >> 2] |= uint32_t(bva[byte + j]) << (8 * (j & 3));
  }
  V8 v;
#ifdef QF_KB
#pragma unroll
  for (int k = 0; k < 4; k++) v.c[k] = from_val(int16_t(w[k >> 1] >> (16 * (k & 1))));
#else
#pragma unroll
  for (int k = 0; k < 8; k++) v.c[k] = from_val(int8_t(w[k >> 2] >> (8 * (k & 3))));
#endif
  return v8_store(v);
}

// Pack one element per 8 BVA bytes, run the Mobius stages on the low `t`
// bits of the element index inside a contiguous tile, and write the message
// half. `block_first` and `available` are byte offsets.
__global__ __launch_bounds__(WG) void k_pack_mobius(const uint8_t* __restrict__ bva, uint64_t block_first,
                                                    uint64_t available, E8* __restrict__ cw, int log_n,
                                                    uint64_t rows, int tbits) {
  __shared__ E8 t[TILE];
  const int n = 1 << tbits;
  uint64_t first = uint64_t(blockIdx.x) << tbits;  // packed element index within the block
  for (int e = threadIdx.x; e < n; e += WG) t[e] = packed_element(bva, block_first + 8 * (first + e), available);
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
#ifdef QF_KB  // one coefficient per word
  uint32_t c[32];
#pragma unroll
  for (int k = 0; k < 32; k++) c[k] = base[k * cstride];
#pragma unroll
  for (int h = 1; h < 32; h *= 2)
#pragma unroll
    for (int k = 0; k < 32; k++)
      if (k & h) c[k] = msub(c[k], c[k ^ h]);
#pragma unroll
  for (int k = 0; k < 32; k++) base[k * cstride] = c[k];
#else  // two u16 coefficients per word
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
#endif
}

// Leaf hash of one 32-column codeword row: BLAKE2s(prefix || row bytes), with
// the framed domain + width prefix of PLEN bytes (75 for base rows, 80 for
// extension rows). The first 64 prefix bytes are compressed on the host.
//
// The 512 row bytes stream through a register window of five columns (20
// words): message block b (b >= 1) covers row words 16 b - QB .. 16 b - QB + 16,
// which start at word A of the columns 4b - 5 .. 4b - 1, so each block loads
// four new columns and keeps the last one of the previous block. No array is
// indexed at run time, so nothing spills to scratch.
template <int PLEN>
__device__ inline void leaf_window_block(uint32_t (&W)[20], uint32_t (&h)[8], const uint4* __restrict__ col,
                                         uint64_t rows, const LeafPrefix& lp, int b, bool first, bool last) {
  constexpr int TOTAL = PLEN + 512;
  constexpr int OFF = (4 - PLEN % 4) % 4;    // byte offset of message words inside row words
  constexpr int QB = (PLEN + OFF) / 4;       // block b starts at row word 16 b - QB
  constexpr int A = ((-QB) % 4 + 4) % 4;     // ... which is word A of the window
  const int c0 = (16 * b - QB - A) / 4;      // first window column (exact division)
#pragma unroll
  for (int k = 0; k < 4; k++) W[k] = W[16 + k];  // column c0 was the previous block's last column
#pragma unroll
  for (int j = 1; j < 5; j++) {
    const int c = c0 + j;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (c < 32) v = col[uint64_t(c) * rows];
    W[4 * j] = v.x;
    W[4 * j + 1] = v.y;
    W[4 * j + 2] = v.z;
    W[4 * j + 3] = v.w;
  }
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t lo = W[A + i], hi = W[A + i + 1];
    uint32_t x = OFF ? (lo >> (8 * OFF)) | (hi << (32 - 8 * OFF)) : lo;
    if (first && 64 + 4 * i < PLEN) x |= lp.tail[i];  // prefix bytes 64 .. PLEN - 1
    m[i] = x;
  }
  b2s_compress(h, m, last ? uint32_t(TOTAL) : uint32_t(64 * (b + 1)), last);
}

// Rows [first, first + count) of a codeword whose columns are `rows` apart.
template <int PLEN>
__global__ __launch_bounds__(WG) void k_leaf_hash(const E8* __restrict__ cw, uint64_t rows, uint64_t first,
                                                  uint64_t count, LeafPrefix lp, uint32_t* __restrict__ leaves) {
  constexpr int BLOCKS = (PLEN + 512 + 63) / 64;
  static_assert(BLOCKS == 10 && (16 * 1 - (PLEN + (4 - PLEN % 4) % 4) / 4) < 0, "prefix ends inside block 1");
  for (uint64_t r = first + uint64_t(blockIdx.x) * WG + threadIdx.x; r < first + count;
       r += uint64_t(gridDim.x) * WG) {
    const uint4* col = reinterpret_cast<const uint4*>(cw) + r;  // column k of this row at col[k * rows]
    uint32_t h[8];
#pragma unroll
    for (int i = 0; i < 8; i++) h[i] = lp.h[i];
    uint32_t W[20];
#pragma unroll
    for (int k = 16; k < 20; k++) W[k] = 0;  // column -1: before the row
    leaf_window_block<PLEN>(W, h, col, rows, lp, 1, true, false);
#pragma unroll 1
    for (int b = 2; b < BLOCKS - 1; b++) leaf_window_block<PLEN>(W, h, col, rows, lp, b, false, false);
    leaf_window_block<PLEN>(W, h, col, rows, lp, BLOCKS - 1, false, true);
    uint4* o = reinterpret_cast<uint4*>(leaves + 8 * r);
    o[0] = make_uint4(h[0], h[1], h[2], h[3]);
    o[1] = make_uint4(h[4], h[5], h[6], h[7]);
  }
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

// Copies a flat packed vector (column k = entries [k n, (k+1) n)) into the
// message halves of a 32-column codeword.
__global__ __launch_bounds__(WG) void k_scatter_message(const E8* __restrict__ values, uint64_t n, uint64_t rows,
                                                        E8* __restrict__ cw) {
  uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (i >= 32 * n) return;
  cw[(i / n) * rows + (i % n)] = values[i];
}

// One-row messages: the WHT of length one is the identity, so parity = alpha_1[0] * m.
__global__ void k_encode_one_row(E8* cw, uint64_t cols, const E8* alpha) {
  uint64_t c = uint64_t(blockIdx.x) * WG + threadIdx.x;
  if (c < cols) cw[2 * c + 1] = v8_store(v8_mul_enc(v8_load(cw[2 * c]), v8_load(alpha[0])));
}

__global__ void k_encode_form(const E8* in, E8* out, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += uint64_t(gridDim.x) * WG)
    out[i] = v8_store(v8_encode_form(v8_load(in[i])));
}

// Index bits [first, last) in the fewest groups of at most 10, split as
// evenly as possible (larger groups first): tile passes with C = 4096 >> t
// contiguous positions, so fewer bits per pass mean longer contiguous runs per
// tile row at the same pass count, e.g. [12, 23) as (12, 6) + (18, 5) rather
// than (12, 10) + (22, 1).
static std::vector<std::pair<int, int>> even_groups(int first, int last) {
  std::vector<std::pair<int, int>> gs;
  const int m = last - first;
  if (m <= 0) return gs;
  const int passes = (m + 9) / 10, base = m / passes, extra = m % passes;
  for (int i = 0, s = first; i < passes; i++) {
    const int t = base + (i < extra ? 1 : 0);
    gs.push_back({s, t});
    s += t;
  }
  return gs;
}
// High index bits [12, L): the groups of every WHT over the row bits above a
// 4096-element tile.
static std::vector<std::pair<int, int>> hi_groups(int L) { return even_groups(12, L); }

// Mobius groups of a commit over the row bits [7, L) left after its first pass
// (row bits 0..6 with the column bits). The last group [a, L) also starts the
// parity WHT, whose pre-alpha bits [12, a) then take passes of their own (the
// second list). Either row bits 7..11 as one group and then the high groups
// (no pre-alpha passes beyond the high groups), or [7, L) split evenly,
// whichever needs fewer passes, with a >= 12 so that the low parity pass does
// not transform those bits again: L = 23 takes (7, 8) + (15, 8) and one
// pre-alpha pass over bits 12..14, seven passes per block instead of eight.
static std::pair<std::vector<std::pair<int, int>>, std::vector<std::pair<int, int>>> mobius_plan(int L) {
  std::vector<std::pair<int, int>> ma{{7, 5}};
  for (auto& g : hi_groups(L)) ma.push_back(g);
  std::vector<std::pair<int, int>> pa(ma.begin() + 1, ma.end() - 1);
  auto mb = even_groups(7, L);
  // the last group's WHT covers its own bits only once if it lies above the low tile
  if (mb.back().first < 12) return {ma, pa};
  auto pb = even_groups(12, mb.back().first);
  if (mb.size() + pb.size() < ma.size() + pa.size()) return {mb, pb};
  return {ma, pa};
}

// Tile-pass schedules (qf_tile.hip); each returns false when its first pass
// declines the geometry (vectors shorter than a tile), and the caller then
// runs the generic LDS-stage passes. Once a schedule has started, every later
// pass must run: a pass that declines would leave the data half transformed.
static void must(bool ran) { require(ran, QF_ERR_ARGUMENT, "tile pass declined a geometry mid-schedule"); }
static bool tiled_wht(E8* v, uint64_t vectors, uint64_t stride, int L, hipStream_t st) {
  if (!tile_pass(TP_W, false, v, v, nullptr, nullptr, vectors, stride, stride, L, 0, std::min(L, 12), 0, 0, st))
    return false;
  for (auto& g : hi_groups(L))
    must(tile_pass(TP_W, false, v, v, nullptr, nullptr, vectors, stride, stride, L, g.first, g.second, 0, 0, st));
  return true;
}
// Second half of the parity, in place: WHT on the low bits, alpha, WHT on the
// low bits (one pass), then WHT on the high groups. Precondition: the parity
// region holds the message with the WHT already applied on every high group.
// On the register-radix kernels F65521's one-pass WHT . alpha . WHT needs more
// registers than two workgroups per CU allow, so full-size blocks run it as two
// lean passes (WHT, then alpha . WHT); the quad-split kernels run it as one.
static void tiled_parity_tail(E8* par, uint64_t cols, uint64_t R, int L, const E8* alpha, hipStream_t st) {
  if (LANES == 8 && L >= 16 && !quad_tiles_enabled()) {
    must(tile_pass(TP_W, false, par, par, nullptr, nullptr, cols, R, R, L, 0, 12, 0, 0, st));
    must(tile_pass(TP_AW, false, par, par, nullptr, alpha, cols, R, R, L, 0, 12, 0, 0, st));
  } else {
    must(tile_pass(TP_WAW, false, par, par, nullptr, alpha, cols, R, R, L, 0, std::min(L, 12), 0, 0, st));
  }
  for (auto& g : hi_groups(L))
    must(tile_pass(TP_W, false, par, par, nullptr, nullptr, cols, R, R, L, g.first, g.second, 0, 0, st));
}
// Parity halves WHT(alpha . WHT(m)) from the message halves of `cols` columns.
static bool tiled_encode_parity(E8* cw, uint64_t cols, int L, const E8* alpha, hipStream_t st) {
  const uint64_t n = uint64_t(1) << L, R = 2 * n;
  auto gs = hi_groups(L);
  if (gs.empty())  // one pass: message -> WHT . alpha . WHT -> parity
    return tile_pass(TP_WAW, false, cw, cw + n, nullptr, alpha, cols, R, R, L, 0, std::min(L, 12), 0, 0, st);
  if (!tile_pass(TP_W, false, cw, cw + n, nullptr, nullptr, cols, R, R, L, gs[0].first, gs[0].second, 0, 0, st))
    return false;
  for (size_t i = 1; i < gs.size(); i++)
    must(tile_pass(TP_W, false, cw + n, cw + n, nullptr, nullptr, cols, R, R, L, gs[i].first, gs[i].second, 0, 0, st));
  tiled_parity_tail(cw + n, cols, R, L, alpha, st);
  return true;
}
// Message halves = Mobius over all 5 + L index bits of the packed block (from
// the BVA when pack, else from a flat E8 vector with column stride n), then the
// parity. Pass 1 holds 32 columns x 128 rows per tile (column bits and row bits
// 0..6), then the Mobius groups of mobius_plan, the last of which also starts
// the parity WHT.
// mw (optional, 2 events): recorded around that last Mobius+WHT pass.
// msg_ready (optional): recorded once the message halves are final, before the parity passes.
static bool tiled_commit_rows(bool pack, const void* src, uint64_t block_first, uint64_t available, E8* cw, int L,
                              const E8* alpha, hipStream_t st, hipEvent_t* mw = nullptr,
                              hipEvent_t msg_ready = nullptr) {
  if (L < 12) return false;
  const uint64_t n = uint64_t(1) << L, R = 2 * n;
  if (!tile_pass(TP_M, pack, src, cw, nullptr, nullptr, 32, n, R, L, 0, 7, block_first, available, st, true))
    return false;
  if (L == 12) {  // message complete after row bits 7..11; one pass gives the parity
    must(tile_pass(TP_M, false, cw, cw, nullptr, nullptr, 32, R, R, L, 7, 5, 0, 0, st));
    if (msg_ready) QF_HIP(hipEventRecord(msg_ready, st));
    must(tile_pass(TP_WAW, false, cw, cw + n, nullptr, alpha, 32, R, R, L, 0, 12, 0, 0, st));
    return true;
  }
  const auto plan = mobius_plan(L);
  const auto& mg = plan.first;
  for (size_t i = 0; i + 1 < mg.size(); i++)
    must(tile_pass(TP_M, false, cw, cw, nullptr, nullptr, 32, R, R, L, mg[i].first, mg[i].second, 0, 0, st));
  // last Mobius group completes the message; the same pass starts the WHT in the parity half
  if (mw) QF_HIP(hipEventRecord(mw[0], st));
  must(tile_pass(TP_M_W, false, cw, cw, cw + n, nullptr, 32, R, R, L, mg.back().first, mg.back().second, 0, 0, st));
  if (mw) QF_HIP(hipEventRecord(mw[1], st));
  if (msg_ready) QF_HIP(hipEventRecord(msg_ready, st));
  for (auto& g : plan.second)
    must(tile_pass(TP_W, false, cw + n, cw + n, nullptr, nullptr, 32, R, R, L, g.first, g.second, 0, 0, st));
  tiled_parity_tail(cw + n, 32, R, L, alpha, st);
  return true;
}

void launch_wht(E8* v, uint64_t vectors, uint64_t stride, int L, hipStream_t st) {
  if (L == 0) return;
  if (tiled_wht(v, vectors, stride, L, st)) return;
  for (auto& gp : bit_groups(L)) launch_pass_geom(v, v, vectors, stride, stride, L, gp.first, gp.second, OP_WHT, nullptr, st);
}

void launch_encode_parity(E8* cw, uint64_t cols, int L, const E8* alpha, hipStream_t st) {
  const uint64_t n = uint64_t(1) << L, R = 2 * n;
  if (L == 0) {
    hipLaunchKernelGGL(k_encode_one_row, dim3(uint32_t((cols + WG - 1) / WG)), dim3(WG), 0, st, cw, cols, alpha);
    return;
  }
  if (tiled_encode_parity(cw, cols, L, alpha, st)) return;
  // parity = WHT(alpha . WHT(m)): low bits copy m into the parity half, the
  // high bits run forward, multiply, forward again, then the low bits again.
  auto groups = bit_groups(L);
  if (groups.size() == 1) {
    launch_pass_geom(cw, cw + n, cols, R, R, L, 0, groups[0].second, OP_WHT, alpha, st);
    return;
  }
  launch_pass_geom(cw, cw + n, cols, R, R, L, 0, groups[0].second, OP_WHT, nullptr, st);
  for (size_t i = 1; i + 1 < groups.size(); i++)
    launch_pass_geom(cw + n, cw + n, cols, R, R, L, groups[i].first, groups[i].second, OP_WHT, nullptr, st);
  launch_pass_geom(cw + n, cw + n, cols, R, R, L, groups.back().first, groups.back().second, OP_WHT, alpha, st);
  for (size_t i = groups.size() - 1; i-- > 1;)
    launch_pass_geom(cw + n, cw + n, cols, R, R, L, groups[i].first, groups[i].second, OP_WHT, nullptr, st);
  launch_pass_geom(cw + n, cw + n, cols, R, R, L, 0, groups[0].second, OP_WHT, nullptr, st);
}

// Leaves of rows [first, first + count) of a codeword with R rows.
static void launch_leaves(const E8* cw, uint64_t R, uint64_t first, uint64_t count, const LeafPrefix& leaf,
                          uint32_t* tree, hipStream_t st) {
  const dim3 grid{uint32_t((count + WG - 1) / WG), 1, 1}, block{uint32_t(WG), 1, 1};
  if (leaf.prefix_len == 75)
    hipLaunchKernelGGL(k_leaf_hash<75>, grid, block, 0, st, cw, R, first, count, leaf, tree);
  else if (leaf.prefix_len == 80)
    hipLaunchKernelGGL(k_leaf_hash<80>, grid, block, 0, st, cw, R, first, count, leaf, tree);
  else
    throw QfError(QF_ERR_ARGUMENT, "unsupported leaf prefix");
}
static void launch_inner_levels(uint64_t R, const InnerPrefix& inner, uint32_t* tree, hipStream_t st) {
  uint32_t* level = tree;
  for (uint64_t cnt = R; cnt > 1; cnt /= 2) {
    uint32_t* next = level + 8 * cnt;
    hipLaunchKernelGGL(k_inner_level, dim3(uint32_t((cnt / 2 + WG - 1) / WG)), dim3(WG), 0, st, level, cnt / 2, inner,
                       next);
    level = next;
  }
}

void launch_tree(const E8* cw, uint64_t R, const LeafPrefix& leaf, const InnerPrefix& inner, uint32_t* tree,
                 hipStream_t st, hipEvent_t after_leaf) {
  launch_leaves(cw, R, 0, R, leaf, tree, st);
  if (after_leaf) QF_HIP(hipEventRecord(after_leaf, st));
  launch_inner_levels(R, inner, tree, st);
}

// This thread's events for a tree split over two streams.
struct SplitEvents {
  hipEvent_t msg = nullptr, side = nullptr;
  ~SplitEvents() {
    if (msg) (void)hipEventDestroy(msg);
    if (side) (void)hipEventDestroy(side);
  }
};
static SplitEvents& split_events() {
  thread_local SplitEvents e;
  if (!e.msg) QF_HIP(hipEventCreateWithFlags(&e.msg, hipEventDisableTiming));
  if (!e.side) QF_HIP(hipEventCreateWithFlags(&e.side, hipEventDisableTiming));
  return e;
}

// The message rows [0, R/2) hash on `side` once msg_ready (recorded on st when
// the message halves are final) has fired, beside the parity passes still
// running on st; the parity rows and then the inner levels follow on st.
static void launch_tree_split(const E8* cw, uint64_t R, const LeafPrefix& leaf, const InnerPrefix& inner,
                              uint32_t* tree, hipStream_t st, hipStream_t side, hipEvent_t msg_ready,
                              hipEvent_t side_done) {
  QF_HIP(hipStreamWaitEvent(side, msg_ready, 0));
  launch_leaves(cw, R, 0, R / 2, leaf, tree, side);
  QF_HIP(hipEventRecord(side_done, side));
  launch_leaves(cw, R, R / 2, R / 2, leaf, tree, st);
  QF_HIP(hipStreamWaitEvent(st, side_done, 0));
  launch_inner_levels(R, inner, tree, st);
}

// Mobius over the L row bits and the 5 column bits of the message halves (generic passes, in place).
static void launch_mobius_message(E8* cw, int L, hipStream_t st) {
  const uint64_t n = uint64_t(1) << L, R = 2 * n;
  if (L > 0)
    for (auto& gp : bit_groups(L)) launch_pass_geom(cw, cw, 32, R, R, L, gp.first, gp.second, OP_MOBIUS, nullptr, st);
  hipLaunchKernelGGL(k_mobius_columns, dim3(uint32_t((n * 4 + WG - 1) / WG)), dim3(WG), 0, st, cw, n, R);
}

void launch_commit_values(const CommitGeom& g, const E8* values, E8* cw, const E8* alpha, uint32_t* tree,
                          hipStream_t st, hipStream_t side) {
  const int L = g.log_n;
  const uint64_t n = uint64_t(1) << L, R = 2 * n;
  SplitEvents* ev = side ? &split_events() : nullptr;
  if (tiled_commit_rows(false, values, 0, 0, cw, L, alpha, st, nullptr, ev ? ev->msg : nullptr)) {
    if (ev) {
      launch_tree_split(cw, R, g.leaf, g.inner, tree, st, side, ev->msg, ev->side);
      return;
    }
  } else {
    hipLaunchKernelGGL(k_scatter_message, dim3(uint32_t((32 * n + WG - 1) / WG)), dim3(WG), 0, st, values, n, R, cw);
    launch_mobius_message(cw, L, st);
    launch_encode_parity(cw, 32, L, alpha, st);
  }
  launch_tree(cw, R, g.leaf, g.inner, tree, st);
}

void launch_commit_extension(const E8* values, E8* cw, int L, const E8* alpha, const LeafPrefix& leaf,
                             const InnerPrefix& inner, uint32_t* tree, hipStream_t st) {
  const uint64_t n = uint64_t(1) << L, R = 2 * n;
  auto gs = hi_groups(L);
  bool tiled;
  if (gs.empty()) {  // copy the message and encode the parity in one pass
    tiled = tile_pass(TP_COPY_WAW, false, values, cw, cw + n, alpha, 32, n, R, L, 0, std::min(L, 12), 0, 0, st);
  } else {  // copy the message while starting the WHT on the first high group
    tiled = tile_pass(TP_COPY_W, false, values, cw, cw + n, nullptr, 32, n, R, L, gs[0].first, gs[0].second, 0, 0, st);
    if (tiled) {
      for (size_t i = 1; i < gs.size(); i++)
        must(tile_pass(TP_W, false, cw + n, cw + n, nullptr, nullptr, 32, R, R, L, gs[i].first, gs[i].second, 0, 0, st));
      tiled_parity_tail(cw + n, 32, R, L, alpha, st);
    }
  }
  if (!tiled) {
    hipLaunchKernelGGL(k_scatter_message, dim3(uint32_t((32 * n + WG - 1) / WG)), dim3(WG), 0, st, values, n, R, cw);
    launch_encode_parity(cw, 32, L, alpha, st);
  }
  launch_tree(cw, R, leaf, inner, tree, st);
}

void launch_encode_form(const E8* in, E8* out, uint64_t n, hipStream_t st) {
  const uint64_t g = std::min<uint64_t>((n + WG - 1) / WG, 65536);
  hipLaunchKernelGGL(k_encode_form, dim3(uint32_t(std::max<uint64_t>(g, 1))), dim3(WG), 0, st, in, out, n);
}

void launch_commit_encode(const CommitGeom& g, const Val* values, uint64_t first_entry, uint64_t active, E8* cw,
                          const E8* alpha, hipStream_t st, hipEvent_t* mw) {
  const int L = g.log_n;
  // the packing kernels address the BVA in bytes
  const uint8_t* bva = reinterpret_cast<const uint8_t*>(values);
  const uint64_t block_first = first_entry * sizeof(Val), available = active * sizeof(Val);
  const uint64_t n = uint64_t(1) << L, R = 2 * n;
  if (mw) QF_HIP(hipEventRecord(mw[0], st));  // both M_W events stay here if no tiled M_W pass runs
  if (mw) QF_HIP(hipEventRecord(mw[1], st));
  if (!tiled_commit_rows(true, bva, block_first, available, cw, L, alpha, st, mw)) {
    const int tp = L < 12 ? L : 12;
    // 1. pack + Mobius on the low bits within each column
    uint64_t packed = n * 32;
    hipLaunchKernelGGL(k_pack_mobius, dim3(uint32_t(packed >> tp)), dim3(WG), 0, st, bva, block_first, available, cw,
                       L, R, tp);
    // 2. remaining Mobius bits within each column, then over the column bits
    for (auto& gp : bit_groups(L))
      if (gp.first > 0) launch_pass_geom(cw, cw, 32, R, R, L, gp.first, gp.second, OP_MOBIUS, nullptr, st);
    hipLaunchKernelGGL(k_mobius_columns, dim3(uint32_t((n * 4 + WG - 1) / WG)), dim3(WG), 0, st, cw, n, R);
    // 3. parity
    launch_encode_parity(cw, 32, L, alpha, st);
  }
}

void launch_commit_block(const CommitGeom& g, const Val* values, uint64_t first_entry, uint64_t active, E8* cw,
                         const E8* alpha, uint32_t* tree, hipStream_t st, hipEvent_t* ev) {
  const uint64_t R = 2 * (uint64_t(1) << g.log_n);
  if (ev) QF_HIP(hipEventRecord(ev[0], st));
  launch_commit_encode(g, values, first_entry, active, cw, alpha, st, ev ? ev + 4 : nullptr);
  // 4. leaves and inner levels
  if (ev) QF_HIP(hipEventRecord(ev[1], st));
  launch_tree(cw, R, g.leaf, g.inner, tree, st, ev ? ev[2] : nullptr);
  if (ev) QF_HIP(hipEventRecord(ev[3], st));
}

}  // namespace qf
