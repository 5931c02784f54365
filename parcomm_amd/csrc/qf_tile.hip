// Register-radix transform passes of the commitment (gfx950).
//
// A pass applies a tensor product of 2x2 maps (Mobius [[1,0],[-1,1]] or
// Walsh-Hadamard [[1,1],[1,-1]]) on up to 12 index bits of many vectors.
// Modes: M; M then WHT (second output); WHT; WHT . alpha . WHT; copy then WHT;
// copy then WHT . alpha . WHT. One
// workgroup (256 threads) owns a tile of 4096 elements: C contiguous positions
// x 2^t positions spaced 2^s apart (x 2^logV whole vectors when a vector is
// shorter than a tile). Each thread keeps 16 elements as eight signed 32-bit
// lanes without modular reduction (12 stages grow a canonical value by at
// most 2^12, so |x| < 2^28), runs four butterfly stages in registers, and
// trades elements through LDS to bring the next four tile bits into its
// registers (mappings A, B, C below). Values are reduced to canonical u16 only
// before they are stored or multiplied. Workgroups are persistent: each walks
// a contiguous range of tiles (neighbouring tiles share 128-byte lines, so
// they stay on one XCD's L2) and prefetches the next tile's input while the
// current one is transformed.
#include "qf_kernels.hpp"

namespace qf {
namespace {

constexpr int TWG = 256;               // threads per workgroup
constexpr int TQ = 16;                 // elements per thread
constexpr int TN = TWG * TQ;           // elements per tile
constexpr int TPAD = TN + TN / 16;     // LDS row with one pad slot per 16 elements
constexpr int32_t BIAS = 32768 * int32_t(P);  // multiple of P above every lazy magnitude

enum : int { TM_M = 0, TM_M_W = 1, TM_W = 2, TM_WAW = 3, TM_COPY_W = 4, TM_COPY_WAW = 5 };
enum : int { MAP_A = 0, MAP_B = 1, MAP_C = 2 };

struct TileArgs {
  const void* in;           // E8 vectors, or the BVA bytes when packing
  E8* out1;
  E8* out2;
  const E8* alpha;          // alpha_1 by position (TM_WAW)
  uint64_t in_vstride, out_vstride;
  uint64_t block_first, available, pack_n;  // packing: byte offset, active bytes, positions per packed vector
  uint32_t tiles, per_wg;
  int logC, t, s, logV, log_lo_chunks, log_hi_count;
};

// Tile element e held by thread `tid` in register slot q.
template <int MAP>
__device__ inline int tile_e(int tid, int q) {
  if (MAP == MAP_A) return tid + TWG * q;                         // bits 8..11 in registers
  if (MAP == MAP_B) return (tid & 15) + 16 * q + ((tid >> 4) << 8);  // bits 4..7
  return q + 16 * tid;                                            // bits 0..3
}
__device__ inline int lds_slot(int e) { return e + (e >> 4); }

// Position of tile T's element 0.
__device__ inline void tile_base(const TileArgs& a, uint64_t T, uint64_t& vec0, uint64_t& pos0) {
  const uint64_t lc = T & ((uint64_t(1) << a.log_lo_chunks) - 1);
  const uint64_t rest = T >> a.log_lo_chunks;
  const uint64_t hi = rest & ((uint64_t(1) << a.log_hi_count) - 1);
  vec0 = (rest >> a.log_hi_count) << a.logV;
  pos0 = (lc << a.logC) + (hi << (a.s + a.t));
}
// Element offset of tile element e from element 0 for vector stride vs. The
// map is additive over disjoint bit fields of e, so off(x + y) = off(x) + off(y)
// whenever x and y share no set bits; offsets stay below 2^32 elements.
__device__ inline uint32_t tile_off(const TileArgs& a, int e, uint64_t vs) {
  const uint32_t lowc = uint32_t(e) & ((1u << a.logC) - 1);
  const uint32_t h = (uint32_t(e) >> a.logC) & ((1u << a.t) - 1);
  const uint32_t v = uint32_t(e) >> (a.logC + a.t);
  return lowc + (h << a.s) + v * uint32_t(vs);
}

// Butterfly stages on the register bits whose tile bit (bit0 + b) lies in [lo, hi).
template <bool WHT>
__device__ inline void reg_stages(int32_t (&x)[TQ][8], int bit0, int lo, int hi) {
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const int tb = bit0 + b;
    if (tb < lo || tb >= hi) continue;
#pragma unroll
    for (int q = 0; q < TQ; q++) {
      if (q & (1 << b)) continue;
      const int r = q | (1 << b);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int32_t u = x[q][k], w = x[r][k];
        if (WHT) { x[q][k] = u + w; x[r][k] = u - w; }
        else x[r][k] = w - u;
      }
    }
  }
}

template <bool WHT, int MAP>
__device__ inline void phase(int32_t (&x)[TQ][8], int lo, int hi) {
  reg_stages<WHT>(x, MAP == MAP_A ? 8 : (MAP == MAP_B ? 4 : 0), lo, hi);
}

template <int FROM, int TO>
__device__ inline void exchange(int32_t (&x)[TQ][8], uint2 (*lds)[TPAD]) {
  const int tid = threadIdx.x;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < TQ; q++) {
    const int i = lds_slot(tile_e<FROM>(tid, q));
#pragma unroll
    for (int p = 0; p < 4; p++) lds[p][i] = make_uint2(uint32_t(x[q][2 * p]), uint32_t(x[q][2 * p + 1]));
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < TQ; q++) {
    const int i = lds_slot(tile_e<TO>(tid, q));
#pragma unroll
    for (int p = 0; p < 4; p++) {
      const uint2 v = lds[p][i];
      x[q][2 * p] = int32_t(v.x);
      x[q][2 * p + 1] = int32_t(v.y);
    }
  }
}

// Lazy signed value (|v| < 2^31 - BIAS margin) -> canonical residue mod P.
__device__ inline uint32_t reduce_lazy(int32_t v) {
  uint32_t u = uint32_t(v) + uint32_t(BIAS);
  u = u - uint32_t(__mul24(int32_t(u >> 16), int32_t(P)));   // (u & 0xffff) + 15 (u >> 16) < 2^20
  u = u - uint32_t(__mul24(int32_t(u >> 16), int32_t(P)));   // < 65536 + 225
  const uint32_t w = u - P;
  return u < w ? u : w;
}
__device__ inline void reduce_all(int32_t (&x)[TQ][8]) {
#pragma unroll
  for (int q = 0; q < TQ; q++)
#pragma unroll
    for (int k = 0; k < 8; k++) x[q][k] = int32_t(reduce_lazy(x[q][k]));
}

__device__ inline void unpack_e8(const uint4& w, int32_t (&c)[8]) {
  const uint32_t v[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
  for (int i = 0; i < 4; i++) { c[2 * i] = int32_t(v[i] & 0xffffu); c[2 * i + 1] = int32_t(v[i] >> 16); }
}
__device__ inline uint4 pack_e8(const int32_t (&c)[8]) {
  return make_uint4(uint32_t(c[0]) | (uint32_t(c[1]) << 16), uint32_t(c[2]) | (uint32_t(c[3]) << 16),
                    uint32_t(c[4]) | (uint32_t(c[5]) << 16), uint32_t(c[6]) | (uint32_t(c[7]) << 16));
}
__device__ inline void unpack_i8(const uint2& w, int32_t (&c)[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) c[k] = int32_t(int8_t(((k < 4 ? w.x : w.y) >> (8 * (k & 3))) & 0xffu));
}

// a * b in F_p[X]/(X^8 - 17) for canonical a, b: coefficient i sums exactly
// eight products a_j * b'_(i-j mod 8) with b' = b for j <= i and 17 b above.
// a is split into 8-bit limbs so every product and every sum of eight stays
// below 2^27 in a 24-bit multiply-add.
__device__ inline void mul_alpha(int32_t (&a)[8], const uint4& bw) {
  uint32_t b[8], b17[8];
  {
    const uint32_t v[4] = {bw.x, bw.y, bw.z, bw.w};
#pragma unroll
    for (int i = 0; i < 4; i++) { b[2 * i] = v[i] & 0xffffu; b[2 * i + 1] = v[i] >> 16; }
  }
#pragma unroll
  for (int j = 0; j < 8; j++) {
    uint32_t u = __umul24(b[j], 17u);  // < 2^21
    u = u - uint32_t(__mul24(int32_t(u >> 16), int32_t(P)));
    const uint32_t w = u - P;
    b17[j] = u < w ? u : w;
  }
  uint32_t lo[8], hi[8];
#pragma unroll
  for (int j = 0; j < 8; j++) { lo[j] = uint32_t(a[j]) & 0xffu; hi[j] = uint32_t(a[j]) >> 8; }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint32_t sl = 0, sh = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const uint32_t bb = j <= i ? b[i - j] : b17[i + 8 - j];
      sl = __umul24(lo[j], bb) + sl;
      sh = __umul24(hi[j], bb) + sh;
    }
    // sh < 2^27: reduce, shift by 8, add sl (< 2^27), reduce
    uint32_t r = sh - uint32_t(__mul24(int32_t(sh >> 16), int32_t(P)));   // < 2^16 + 15 * 2^11
    r = (r << 8) + sl;                                                     // < 2^25 + 2^27
    r = r - uint32_t(__mul24(int32_t(r >> 16), int32_t(P)));               // < 2^16 + 15 * 2^12
    r = r - uint32_t(__mul24(int32_t(r >> 16), int32_t(P)));               // < 2^16 + 45
    const uint32_t w = r - P;
    a[i] = int32_t(r < w ? r : w);
  }
}

// Per-thread element offsets of the mapping-A slots: slot q = thread part + q part.
struct SlotOffs {
  uint32_t in_t, out_t;      // thread part (input stride / output stride)
};
__device__ inline uint32_t q_part(const TileArgs& a, int q, uint64_t vs) { return tile_off(a, TWG * q, vs); }

template <bool PACK>
__device__ inline void load_tile(const TileArgs& a, uint64_t T, const SlotOffs& so, uint4 (&raw)[TQ]) {
  uint64_t vec0, pos0;
  tile_base(a, T, vec0, pos0);
  if (PACK) {
    const int8_t* bva = static_cast<const int8_t*>(a.in);
    const uint64_t b0 = a.block_first + 8 * (vec0 * a.pack_n + pos0);
#pragma unroll
    for (int q = 0; q < TQ; q++) {
      const uint64_t byte = b0 + 8 * uint64_t(so.in_t + q_part(a, q, a.pack_n));
      uint2 w;
      if (byte + 8 <= a.available) {
        w = *reinterpret_cast<const uint2*>(bva + byte);
      } else {
        w = make_uint2(0, 0);
        for (int j = 0; j < 8; j++)
          if (byte + j < a.available) {
            const uint32_t v = uint32_t(uint8_t(bva[byte + j])) << (8 * (j & 3));
            if (j < 4) w.x |= v; else w.y |= v;
          }
      }
      raw[q] = make_uint4(w.x, w.y, 0, 0);
    }
  } else {
    const uint4* base = reinterpret_cast<const uint4*>(static_cast<const E8*>(a.in) + vec0 * a.in_vstride + pos0);
#pragma unroll
    for (int q = 0; q < TQ; q++) raw[q] = base[so.in_t + q_part(a, q, a.in_vstride)];
  }
}

__device__ inline void store_tile(const TileArgs& a, uint64_t T, const SlotOffs& so, E8* out, const int32_t (&x)[TQ][8]) {
  uint64_t vec0, pos0;
  tile_base(a, T, vec0, pos0);
  uint4* base = reinterpret_cast<uint4*>(out + vec0 * a.out_vstride + pos0);
#pragma unroll
  for (int q = 0; q < TQ; q++) base[so.out_t + q_part(a, q, a.out_vstride)] = pack_e8(x[q]);
}

template <int MODE, bool PACK>
__global__ __launch_bounds__(TWG) void k_tile(TileArgs a) {
  __shared__ uint2 lds[4][TPAD];
  const uint64_t first = uint64_t(blockIdx.x) * a.per_wg;
  const uint64_t last = first + a.per_wg < a.tiles ? first + a.per_wg : a.tiles;
  if (first >= last) return;
  const int lo = a.logC, hi = a.logC + a.t;  // transformed tile bits
  SlotOffs so;
  so.in_t = tile_off(a, tile_e<MAP_A>(threadIdx.x, 0), PACK ? a.pack_n : a.in_vstride);
  so.out_t = tile_off(a, tile_e<MAP_A>(threadIdx.x, 0), a.out_vstride);
  uint4 raw[TQ];
  load_tile<PACK>(a, first, so, raw);
  for (uint64_t T = first; T < last; T++) {
    int32_t x[TQ][8];
#pragma unroll
    for (int q = 0; q < TQ; q++) {
      if (PACK) unpack_i8(make_uint2(raw[q].x, raw[q].y), x[q]);
      else unpack_e8(raw[q], x[q]);
    }
    if (MODE == TM_COPY_W || MODE == TM_COPY_WAW) store_tile(a, T, so, a.out1, x);
    constexpr bool WAW = MODE == TM_WAW || MODE == TM_COPY_WAW;
    if (!WAW && T + 1 < last) load_tile<PACK>(a, T + 1, so, raw);
    if (MODE == TM_M || MODE == TM_M_W) {
      phase<false, MAP_A>(x, lo, hi);
      exchange<MAP_A, MAP_B>(x, lds);
      phase<false, MAP_B>(x, lo, hi);
      exchange<MAP_B, MAP_C>(x, lds);
      phase<false, MAP_C>(x, lo, hi);
      exchange<MAP_C, MAP_A>(x, lds);
      reduce_all(x);
      store_tile(a, T, so, a.out1, x);
    }
    if (MODE == TM_M_W || MODE == TM_W || MODE == TM_COPY_W) {
      phase<true, MAP_A>(x, lo, hi);
      exchange<MAP_A, MAP_B>(x, lds);
      phase<true, MAP_B>(x, lo, hi);
      exchange<MAP_B, MAP_C>(x, lds);
      phase<true, MAP_C>(x, lo, hi);
      exchange<MAP_C, MAP_A>(x, lds);
      reduce_all(x);
      store_tile(a, T, so, MODE == TM_W ? a.out1 : a.out2, x);
    }
    if (WAW) {
      phase<true, MAP_A>(x, lo, hi);
      exchange<MAP_A, MAP_B>(x, lds);
      uint4 al[TQ];  // alpha_1 of the elements held in mapping C
      {
        uint64_t vec0, pos0;
        tile_base(a, T, vec0, pos0);
        const uint4* ab = reinterpret_cast<const uint4*>(a.alpha) + pos0;
        const uint32_t ct = tile_off(a, tile_e<MAP_C>(threadIdx.x, 0), 0);
#pragma unroll
        for (int q = 0; q < TQ; q++) al[q] = ab[ct + tile_off(a, q, 0)];
      }
      phase<true, MAP_B>(x, lo, hi);
      exchange<MAP_B, MAP_C>(x, lds);
      phase<true, MAP_C>(x, lo, hi);
      reduce_all(x);
#pragma unroll
      for (int q = 0; q < TQ; q++) mul_alpha(x[q], al[q]);
      phase<true, MAP_C>(x, lo, hi);
      exchange<MAP_C, MAP_B>(x, lds);
      phase<true, MAP_B>(x, lo, hi);
      exchange<MAP_B, MAP_A>(x, lds);
      phase<true, MAP_A>(x, lo, hi);
      reduce_all(x);
      store_tile(a, T, so, MODE == TM_WAW ? a.out1 : a.out2, x);
      if (T + 1 < last) load_tile<PACK>(a, T + 1, so, raw);
    }
    __syncthreads();  // the next tile's first exchange reuses the LDS rows
  }
}

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    QF_HIP(hipGetDevice(&dev));
    QF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus <= 0) cus = 256;
  }
  return cus;
}

}  // namespace

// One tile pass over `vectors` vectors of length 2^L, transforming bits [s, s+t).
// Requirements: a tile holds 4096 elements, so t + logC + logV = 12 with C = 2^logC
// contiguous positions (s >= logC) and 2^logV whole vectors (only when s = 0, t = L).
bool tile_pass(int mode, bool pack, const void* in, E8* out1, E8* out2, const E8* alpha, uint64_t vectors,
               uint64_t in_vstride, uint64_t out_vstride, int L, int s, int t, uint64_t block_first,
               uint64_t available, hipStream_t st) {
  TileArgs a{};
  a.in = in;
  a.out1 = out1;
  a.out2 = out2;
  a.alpha = alpha;
  a.in_vstride = in_vstride;
  a.out_vstride = out_vstride;
  a.block_first = block_first;
  a.available = available;
  a.pack_n = uint64_t(1) << L;
  a.s = s;
  a.t = t;
  if (s == 0) {
    if (t != std::min(L, 12) || t < 1) return false;
    a.logC = 0;
    a.logV = 12 - t;
    if (vectors % (uint64_t(1) << a.logV)) return false;
    a.log_lo_chunks = 0;
  } else {
    if (t < 1 || t > 10 || s + t > L) return false;
    a.logC = 12 - t;
    a.logV = 0;
    if (s < a.logC) return false;
    a.log_lo_chunks = s - a.logC;
  }
  a.log_hi_count = L - s - t;
  const uint64_t tiles = (vectors << L) / TN;
  if (tiles == 0 || tiles >= (1ull << 31)) return false;
  a.tiles = uint32_t(tiles);
  const uint64_t wgs = std::min<uint64_t>(tiles, uint64_t(device_cus()));
  a.per_wg = uint32_t((tiles + wgs - 1) / wgs);
  const uint32_t grid = uint32_t((tiles + a.per_wg - 1) / a.per_wg);
  dim3 g{grid, 1, 1}, b{uint32_t(TWG), 1, 1};
#define QF_TILE_LAUNCH(M, PK) hipLaunchKernelGGL((k_tile<M, PK>), g, b, 0, st, a)
  if (pack) {
    if (mode != TM_M) return false;
    QF_TILE_LAUNCH(TM_M, true);
  } else {
    switch (mode) {
      case TM_M: QF_TILE_LAUNCH(TM_M, false); break;
      case TM_M_W: QF_TILE_LAUNCH(TM_M_W, false); break;
      case TM_W: QF_TILE_LAUNCH(TM_W, false); break;
      case TM_WAW: QF_TILE_LAUNCH(TM_WAW, false); break;
      case TM_COPY_W: QF_TILE_LAUNCH(TM_COPY_W, false); break;
      case TM_COPY_WAW: QF_TILE_LAUNCH(TM_COPY_WAW, false); break;
      default: return false;
    }
  }
#undef QF_TILE_LAUNCH
  QF_HIP(hipGetLastError());
  return true;
}

}  // namespace qf
