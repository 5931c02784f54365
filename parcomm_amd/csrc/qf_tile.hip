// Register-radix transform passes of the commitment (gfx950).
//
// A pass applies a tensor product of 2x2 maps (Mobius [[1,0],[-1,1]] or
// Walsh-Hadamard [[1,1],[1,-1]]) on up to 12 index bits of many vectors.
// Modes: M; M then WHT (second output); WHT; WHT . alpha . WHT; copy then WHT;
// copy then WHT . alpha . WHT; alpha then WHT. One
// workgroup (512 threads) owns a tile of 4096 elements: C contiguous positions
// x 2^t positions spaced 2^s apart (x 2^logV whole vectors when a vector is
// shorter than a tile). Each thread keeps 8 elements in registers, one 32-bit
// lane per coefficient, runs three butterfly stages in registers, and
// trades elements through LDS to bring the next three tile bits into its
// registers (mappings with register bits 9..11, 6..8, 3..5, 0..2). F65521
// lanes are signed and unreduced (12 stages grow a canonical value by at most
// 2^12, so |x| < 2^28) and are reduced to canonical u16 only before they are
// stored or multiplied; KoalaBear lanes (p ~ 2^31) stay canonical through
// modular butterflies. A workgroup takes one tile (or, persistent, walks a
// contiguous range of tiles), except that in passes whose tile rows are 64
// bytes (4 contiguous positions) the two tiles that share every 128-byte line
// are taken at the same time by sibling workgroups on one XCD (workgroups are
// dealt to the 8 XCDs round robin), so the second half of each line is still
// in that XCD's L2 when its sibling asks for it; a line read by one tile long
// after its partner has been evicted is fetched from HBM twice.
#include "qf_kernels.hpp"
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace qf {
namespace {

constexpr int TWG = 512;               // threads per workgroup
constexpr int TQB = 3;                 // register bits per phase
constexpr int TQ = 1 << TQB;           // elements per thread
constexpr int TN = TWG * TQ;           // elements per tile
constexpr int TPAD = TN + TN / 32;     // LDS row with one pad slot per 32 elements
// An exchange moves the LANES 32-bit lanes of every tile element through NR
// LDS rows of uint2 (66 KiB for NR = 2, so two workgroups can share a CU);
// F65521's eight lanes cross in two rounds. A single round through four rows
// needs more registers than the alpha passes have and spills.
constexpr int NR = 2;

enum : int { TM_M = 0, TM_M_W = 1, TM_W = 2, TM_WAW = 3, TM_COPY_W = 4, TM_COPY_WAW = 5, TM_AW = 6 };
// Mappings are named by the first tile bit held in registers.
enum : int { MAP_A = 9, MAP_B = 6, MAP_C = 3, MAP_D = 0 };

struct TileArgs {
  const void* in;           // E8 vectors, or the BVA bytes when packing
  E8* out1;
  E8* out2;
  const E8* alpha;          // alpha_1 by position (TM_WAW)
  uint64_t in_vstride, out_vstride;
  uint64_t block_first, available, pack_n;  // packing: byte offset, active bytes, positions per packed vector
  uint32_t tiles, per_wg;
  uint32_t xcd_span;        // > 0: XCD-interleaved walk, tiles per XCD (even)
  int logC, t, s, logV, log_lo_chunks, log_hi_count;
  int xlo, xhi;             // transformed tile bits [xlo, xhi)
  int quad_ok;              // the quad-split kernels' per-thread offsets fit 32 bits
  int big_ok;               // the 8192-element tiles apply
  // quad-split kernels: element offset of slot 1 of each mapping (slot q is q
  // times that: the five slot bits lie in one field of the tile index), for the
  // input (BVA bytes / 8 when packing), the outputs and alpha (positions)
  uint64_t qs_in[2], qs_out[2], qs_alpha;
  uint64_t pack_span;       // packing: bytes from a tile's first input byte to past its last
};

// Tile element e held by thread `tid` in register slot q: q supplies tile bits
// [MAP, MAP+3), tid the other nine. Mappings A and B (the load/store layout
// and the first exchange) give them in order, so a wave's lanes cover
// consecutive elements; C and D, which only meet LDS, take them in the order
// that keeps their exchanges free of bank conflicts (below).
template <int MAP>
__device__ inline int tile_e(int tid, int q) {
  if (MAP == MAP_C)  // tid bits -> e bits 0 1 2 8 9 6 7 10 11
    return (tid & 7) | (q << 3) | (((tid >> 3) & 3) << 8) | (((tid >> 5) & 3) << 6) | ((tid >> 7) << 10);
  if (MAP == MAP_D)  // tid bits -> e bits 5 6 7 3 4 8 9 10 11
    return q | ((tid & 7) << 5) | (((tid >> 3) & 3) << 3) | ((tid >> 5) << 8);
  return (tid & ((1 << MAP) - 1)) | (q << MAP) | ((tid >> MAP) << (MAP + TQB));
}
// LDS slot (8 bytes) of tile element e: one pad slot per 32 elements, so bit
// i of e moves the slot by 2^i for i < 5 and by 2^i + 2^(i-5) above. The 8-byte
// accesses of 16 (or 32) lanes fall in distinct bank pairs when the weights
// of the element bits those lanes vary are 1, 2, 4, 8 (, 16) modulo 16 (32):
// A and B vary bits 0..4; C varies 0, 1, 2, 8 (264), 9 (528); D varies
// 5 (33), 6 (66), 7 (132), 3, 4.
__device__ inline int lds_slot(int e) { return e + (e >> 5); }

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
// whenever x and y share no set bits. The 32-bit form serves the per-thread
// parts, which the host bounds below 2^32 bytes; the uniform per-slot parts
// use the 64-bit form.
__host__ __device__ inline uint64_t tile_off64(const TileArgs& a, int e, uint64_t vs) {
  const uint64_t lowc = uint64_t(e) & ((uint64_t(1) << a.logC) - 1);
  const uint64_t h = (uint64_t(e) >> a.logC) & ((uint64_t(1) << a.t) - 1);
  const uint64_t v = uint64_t(e) >> (a.logC + a.t);
  return lowc + (h << a.s) + v * vs;
}
__device__ inline uint32_t tile_off(const TileArgs& a, int e, uint64_t vs) {
  const uint32_t lowc = uint32_t(e) & ((1u << a.logC) - 1);
  const uint32_t h = (uint32_t(e) >> a.logC) & ((1u << a.t) - 1);
  const uint32_t v = uint32_t(e) >> (a.logC + a.t);
  return lowc + (h << a.s) + v * uint32_t(vs);
}

// Butterfly stages on the register bits whose tile bit (MAP + b) lies in the
// transformed range: [LO, 12) when LO >= 0 (known at compile time, so the
// stages carry no branches), else the runtime range [lo, hi). The scheduler
// barriers keep the compiler from interleaving whole phases, which would
// lengthen live ranges.
// Butterfly arithmetic on register lanes.
#ifdef QF_KB
__device__ inline int32_t bf_add(int32_t u, int32_t w) { return int32_t(madd(uint32_t(u), uint32_t(w))); }
__device__ inline int32_t bf_sub(int32_t u, int32_t w) { return int32_t(msub(uint32_t(u), uint32_t(w))); }
#else
__device__ inline int32_t bf_add(int32_t u, int32_t w) { return u + w; }
__device__ inline int32_t bf_sub(int32_t u, int32_t w) { return u - w; }
#endif

template <bool WHT, int MAP, int LO>
__device__ inline void phase(int32_t (&x)[TQ][LANES], int lo, int hi) {
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < TQB; b++) {
    const int tb = MAP + b;
    if (LO >= 0) {
      if (tb < LO) continue;
    } else if (tb < lo || tb >= hi) {
      continue;
    }
#pragma unroll
    for (int q = 0; q < TQ; q++) {
      if (q & (1 << b)) continue;
      const int r = q | (1 << b);
#pragma unroll
      for (int k = 0; k < LANES; k++) {
        const int32_t u = x[q][k], w = x[r][k];
        if (WHT) { x[q][k] = bf_add(u, w); x[r][k] = bf_sub(u, w); }
        else x[r][k] = bf_sub(w, u);
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

template <int FROM, int TO, int LROWS>
__device__ inline void exchange(int32_t (&x)[TQ][LANES], uint2 (*lds)[TPAD]) {
  constexpr int XROUNDS = LANES / (2 * LROWS);
  const int tid = threadIdx.x;
#pragma unroll
  for (int rnd = 0; rnd < XROUNDS; rnd++) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TQ; q++) {
      const int i = lds_slot(tile_e<FROM>(tid, q));
#pragma unroll
      for (int p = 0; p < LROWS; p++) {
        const int k = 2 * (rnd * LROWS + p);
        lds[p][i] = make_uint2(uint32_t(x[q][k]), uint32_t(x[q][k + 1]));
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TQ; q++) {
      const int i = lds_slot(tile_e<TO>(tid, q));
#pragma unroll
      for (int p = 0; p < LROWS; p++) {
        const int k = 2 * (rnd * LROWS + p);
        const uint2 v = lds[p][i];
        x[q][k] = int32_t(v.x);
        x[q][k + 1] = int32_t(v.y);
      }
    }
  }
}

// One transform over the tile bits [lo, hi): the register phases of mappings
// A, B, C, D with LDS exchanges between them and back to A. Mappings whose
// three bits all lie below the lowest transformed bit are skipped with their
// exchanges: from bit 9 up A alone, from 6 A and B, from 3 A, B and C.
template <bool WHT, int LO>
__device__ inline void transform(int32_t (&x)[TQ][LANES], uint2 (*lds)[TPAD], int lo, int hi) {
  const int first = LO >= 0 ? LO : lo;  // uniform across the workgroup
  phase<WHT, MAP_A, LO>(x, lo, hi);
  if (first >= MAP_A) return;
  exchange<MAP_A, MAP_B, NR>(x, lds);
  phase<WHT, MAP_B, LO>(x, lo, hi);
  if (first >= MAP_B) {
    exchange<MAP_B, MAP_A, NR>(x, lds);
    return;
  }
  exchange<MAP_B, MAP_C, NR>(x, lds);
  phase<WHT, MAP_C, LO>(x, lo, hi);
  if (first >= MAP_C) {
    exchange<MAP_C, MAP_A, NR>(x, lds);
    return;
  }
  exchange<MAP_C, MAP_D, NR>(x, lds);
  phase<WHT, MAP_D, LO>(x, lo, hi);
  exchange<MAP_D, MAP_A, NR>(x, lds);
}

#ifdef QF_KB
// KoalaBear lanes are canonical throughout.
template <int LO>
__device__ inline void reduce_all(int32_t (&)[TQ][LANES]) {}

__device__ inline void unpack_e8(const uint4& w, int32_t (&c)[LANES]) {
  c[0] = int32_t(w.x); c[1] = int32_t(w.y); c[2] = int32_t(w.z); c[3] = int32_t(w.w);
}
__device__ inline uint4 pack_e8(const int32_t (&c)[LANES]) {
  return make_uint4(uint32_t(c[0]), uint32_t(c[1]), uint32_t(c[2]), uint32_t(c[3]));
}
// Four little-endian i16 BVA entries -> canonical lanes.
__device__ inline void unpack_bva(const uint2& w, int32_t (&c)[LANES]) {
  c[0] = int32_t(from_val(int16_t(w.x & 0xffffu)));
  c[1] = int32_t(from_val(int16_t(w.x >> 16)));
  c[2] = int32_t(from_val(int16_t(w.y & 0xffffu)));
  c[3] = int32_t(from_val(int16_t(w.y >> 16)));
}
// a * alpha in KoalaBear[X]/(X^4 - 3), alpha in Montgomery form.
__device__ inline void mul_alpha(int32_t (&a)[LANES], const uint4& bw) {
  V8 x, b;
#pragma unroll
  for (int k = 0; k < 4; k++) x.c[k] = uint32_t(a[k]);
  b.c[0] = bw.x; b.c[1] = bw.y; b.c[2] = bw.z; b.c[3] = bw.w;
  const V8 r = v8_mul_mont(x, b);
#pragma unroll
  for (int k = 0; k < 4; k++) a[k] = int32_t(r.c[k]);
}
#else
constexpr int32_t BIAS = 32768 * int32_t(P);  // multiple of P above every lazy magnitude

// Lazy signed value (|v| < 2^31 - BIAS margin) -> canonical residue mod P.
__device__ inline uint32_t reduce_lazy(int32_t v) {
  uint32_t u = uint32_t(v) + uint32_t(BIAS);
  u = u - uint32_t(__mul24(int32_t(u >> 16), int32_t(P)));   // (u & 0xffff) + 15 (u >> 16) < 2^20
  u = u - uint32_t(__mul24(int32_t(u >> 16), int32_t(P)));   // < 65536 + 225
  const uint32_t w = u - P;
  return u < w ? u : w;
}
// After at most 10 butterfly stages from canonical values |v| <= 2^10 (P - 1):
// v + 2^10 P < 2^27, and one fold leaves (u & 0xffff) + 15 (u >> 16) < 2 P.
constexpr int32_t BIAS10 = 1024 * int32_t(P);
__device__ inline uint32_t reduce_lazy10(int32_t v) {
  uint32_t u = uint32_t(v) + uint32_t(BIAS10);
  u = u - uint32_t(__mul24(int32_t(u >> 16), int32_t(P)));
  const uint32_t w = u - P;
  return u < w ? u : w;
}
// LO >= 2: the pass runs at most 10 stages between reductions.
template <int LO>
__device__ inline void reduce_all(int32_t (&x)[TQ][8]) {
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < TQ; q++)
#pragma unroll
    for (int k = 0; k < 8; k++) x[q][k] = int32_t(LO >= 2 ? reduce_lazy10(x[q][k]) : reduce_lazy(x[q][k]));
  __builtin_amdgcn_sched_barrier(0);
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
__device__ inline void unpack_bva(const uint2& w, int32_t (&c)[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) c[k] = int32_t(int8_t(((k < 4 ? w.x : w.y) >> (8 * (k & 3))) & 0xffu));
}

// a * b in F_p[X]/(X^8 - 17) for canonical a, b. Coefficient i sums eight
// products a_j * E[7 + i - j] over the sequence E = (17 b_1, ..., 17 b_7,
// b_0, ..., b_7) (the X^8 = 17 wrap folded into E). a is split into 8-bit
// limbs; each v_dot2_u32_u16 adds two limb x E products, so every partial
// sum of eight stays below 2^27.
typedef unsigned short qf_u16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t dot2(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(qf_u16x2, a), __builtin_bit_cast(qf_u16x2, b), c, false);
}
__device__ inline uint32_t fold_p(uint32_t u) {  // u < 2^31 -> (u & 0xffff) + 15 (u >> 16)
  return u - uint32_t(__mul24(int32_t(u >> 16), int32_t(P)));
}
__device__ inline void mul_alpha(int32_t (&a)[8], const uint4& bw) {
  uint32_t E[15];
  {
    const uint32_t v[4] = {bw.x, bw.y, bw.z, bw.w};
#pragma unroll
    for (int i = 0; i < 4; i++) { E[7 + 2 * i] = v[i] & 0xffffu; E[8 + 2 * i] = v[i] >> 16; }
  }
#pragma unroll
  for (int j = 1; j < 8; j++) {
    const uint32_t u = fold_p(__umul24(E[7 + j], 17u));  // 17 b_j < 2^21 -> < 2^16 + 255
    const uint32_t w = u - P;
    E[j - 1] = u < w ? u : w;
  }
  uint32_t D[15];  // D[m] = (E[m], E[m-1]) as a u16 pair
#pragma unroll
  for (int m = 1; m < 15; m++) D[m] = E[m] | (E[m - 1] << 16);
  uint32_t lo[4], hi[4];
#pragma unroll
  for (int jp = 0; jp < 4; jp++) {
    const uint32_t pr = uint32_t(a[2 * jp]) | (uint32_t(a[2 * jp + 1]) << 16);
    lo[jp] = pr & 0x00ff00ffu;
    hi[jp] = (pr >> 8) & 0x00ff00ffu;
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint32_t sl = 0, sh = 0;
#pragma unroll
    for (int jp = 0; jp < 4; jp++) {
      sl = dot2(lo[jp], D[7 + i - 2 * jp], sl);
      sh = dot2(hi[jp], D[7 + i - 2 * jp], sh);
    }
    uint32_t r = fold_p(sh);        // < 2^16 + 15 * 2^11
    r = (r << 8) + sl;              // < 2^25 + 2^27
    r = fold_p(fold_p(r));          // < 2^16 + 15
    const uint32_t w = r - P;
    a[i] = int32_t(r < w ? r : w);
  }
}
#endif

// Per-thread byte offsets of the mapping-A slots from the tile's first
// element: slot q = thread part + q part. The q part is uniform (64-bit
// scalar arithmetic); the thread part is a 32-bit offset against that uniform
// base, which keeps per-slot addresses out of 64-bit vector registers.
struct SlotOffs {
  uint32_t in_t, out_t;
};
__device__ inline uint64_t q_part(const TileArgs& a, int q, uint64_t vs) { return tile_off64(a, TWG * q, vs); }

template <bool PACK>
__device__ inline void load_tile(const TileArgs& a, uint64_t T, const SlotOffs& so, uint4 (&raw)[TQ]) {
  uint64_t vec0, pos0;
  tile_base(a, T, vec0, pos0);
  if (PACK) {
    const uint64_t b0 = a.block_first + 8 * (vec0 * a.pack_n + pos0);
    const char* base = static_cast<const char*>(a.in) + b0;
#pragma unroll
    for (int q = 0; q < TQ; q++) {
      const char* bq = base + 8 * q_part(a, q, a.pack_n);
      const uint64_t byte = uint64_t(bq - static_cast<const char*>(a.in)) + so.in_t;
      uint2 w;
      if (byte + 8 <= a.available) {
        w = *reinterpret_cast<const uint2*>(bq + so.in_t);
      } else {  // bytes at or past the active end read as zero
        const uint8_t* bva = static_cast<const uint8_t*>(a.in);
        w = make_uint2(0, 0);
        for (int j = 0; j < 8; j++)
          if (byte + j < a.available) {
            const uint32_t v = uint32_t(bva[byte + j]) << (8 * (j & 3));
            if (j < 4) w.x |= v; else w.y |= v;
          }
      }
      raw[q] = make_uint4(w.x, w.y, 0, 0);
    }
  } else {
    const char* base = reinterpret_cast<const char*>(static_cast<const E8*>(a.in) + vec0 * a.in_vstride + pos0);
#pragma unroll
    for (int q = 0; q < TQ; q++)
      raw[q] = *reinterpret_cast<const uint4*>(base + 16 * q_part(a, q, a.in_vstride) + so.in_t);
  }
}

__device__ inline void store_tile(const TileArgs& a, uint64_t T, const SlotOffs& so, E8* out, const int32_t (&x)[TQ][LANES]) {
  uint64_t vec0, pos0;
  tile_base(a, T, vec0, pos0);
  char* base = reinterpret_cast<char*>(out + vec0 * a.out_vstride + pos0);
#pragma unroll
  for (int q = 0; q < TQ; q++)
    *reinterpret_cast<uint4*>(base + 16 * q_part(a, q, a.out_vstride) + so.out_t) = pack_e8(x[q]);
}

// PF: prefetch the next tile's input while the current one is transformed
// (one workgroup per CU); without it each tile loads at the top of its
// iteration and the other workgroup on the CU hides the latency.
template <int MODE, bool PACK, int LO, bool PF>
__device__ inline void tile_body(const TileArgs& a) {
  __shared__ uint2 lds[NR][TPAD];
  uint64_t first, last, step;
  if (a.xcd_span) {  // XCD g = blockIdx % 8 owns tiles [g span, (g+1) span); its workgroups take consecutive tiles
    const uint64_t g = blockIdx.x % 8, slots = gridDim.x / 8;
    first = g * a.xcd_span + blockIdx.x / 8;
    last = (g + 1) * a.xcd_span < a.tiles ? (g + 1) * a.xcd_span : a.tiles;
    step = slots;
  } else {
    first = uint64_t(blockIdx.x) * a.per_wg;
    last = first + a.per_wg < a.tiles ? first + a.per_wg : a.tiles;
    step = 1;
  }
  if (first >= last) return;
  const int lo = a.xlo, hi = a.xhi;  // transformed tile bits (runtime form)
  SlotOffs so;
  so.in_t = (PACK ? 8u : 16u) * tile_off(a, tile_e<MAP_A>(threadIdx.x, 0), PACK ? a.pack_n : a.in_vstride);
  so.out_t = 16u * tile_off(a, tile_e<MAP_A>(threadIdx.x, 0), a.out_vstride);
  uint4 raw[TQ];
  if (PF) load_tile<PACK>(a, first, so, raw);
  for (uint64_t T = first; T < last; T += step) {
    int32_t x[TQ][LANES];
    if (!PF) load_tile<PACK>(a, T, so, raw);
#pragma unroll
    for (int q = 0; q < TQ; q++) {
      if (PACK) unpack_bva(make_uint2(raw[q].x, raw[q].y), x[q]);
      else unpack_e8(raw[q], x[q]);
    }
    if (MODE == TM_COPY_W || MODE == TM_COPY_WAW) store_tile(a, T, so, a.out1, x);
    constexpr bool WAW = MODE == TM_WAW || MODE == TM_COPY_WAW;
    if (PF && !WAW && T + step < last) load_tile<PACK>(a, T + step, so, raw);
    if (MODE == TM_AW) {  // x <- alpha . x (x canonical from memory), then the WHT below
      uint64_t vec0, pos0;
      tile_base(a, T, vec0, pos0);
      const char* ab = reinterpret_cast<const char*>(a.alpha + pos0);
      const uint32_t dt = 16u * tile_off(a, tile_e<MAP_A>(threadIdx.x, 0), 0);
      auto alpha_at = [&](int q) {
        return *reinterpret_cast<const uint4*>(ab + dt + 16u * tile_off(a, TWG * q, 0));
      };
      uint4 al0 = alpha_at(0), al1 = alpha_at(1);
#pragma unroll
      for (int q = 0; q < TQ; q += 2) {
        const uint4 c0 = al0, c1 = al1;
        if (q + 2 < TQ) { al0 = alpha_at(q + 2); al1 = alpha_at(q + 3); }
        mul_alpha(x[q], c0);
        __builtin_amdgcn_sched_barrier(0);
        mul_alpha(x[q + 1], c1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (MODE == TM_M || MODE == TM_M_W) {
      transform<false, LO>(x, lds, lo, hi);
      reduce_all<LO>(x);
      store_tile(a, T, so, a.out1, x);
    }
    if (MODE == TM_M_W || MODE == TM_W || MODE == TM_COPY_W || MODE == TM_AW) {
      transform<true, LO>(x, lds, lo, hi);
      reduce_all<LO>(x);
      store_tile(a, T, so, MODE == TM_W || MODE == TM_AW ? a.out1 : a.out2, x);
    }
    if (WAW) {
      phase<true, MAP_A, LO>(x, lo, hi);
      exchange<MAP_A, MAP_B, NR>(x, lds);
      phase<true, MAP_B, LO>(x, lo, hi);
      exchange<MAP_B, MAP_C, NR>(x, lds);
      // alpha_1 of the elements held in mapping D, two slots in flight
      const char* ab;
      uint32_t dt;
      {
        uint64_t vec0, pos0;
        tile_base(a, T, vec0, pos0);
        ab = reinterpret_cast<const char*>(a.alpha + pos0);
        dt = 16u * tile_off(a, tile_e<MAP_D>(threadIdx.x, 0), 0);
      }
      auto alpha_at = [&](int q) { return *reinterpret_cast<const uint4*>(ab + dt + 16u * tile_off(a, q, 0)); };
      uint4 al0 = alpha_at(0), al1 = alpha_at(1);
      phase<true, MAP_C, LO>(x, lo, hi);
      exchange<MAP_C, MAP_D, NR>(x, lds);
      phase<true, MAP_D, LO>(x, lo, hi);
      reduce_all<LO>(x);
#pragma unroll
      for (int q = 0; q < TQ; q += 2) {
        const uint4 c0 = al0, c1 = al1;
        if (q + 2 < TQ) { al0 = alpha_at(q + 2); al1 = alpha_at(q + 3); }
        mul_alpha(x[q], c0);
        __builtin_amdgcn_sched_barrier(0);
        mul_alpha(x[q + 1], c1);
        __builtin_amdgcn_sched_barrier(0);
      }
      phase<true, MAP_D, LO>(x, lo, hi);
      exchange<MAP_D, MAP_C, NR>(x, lds);
      phase<true, MAP_C, LO>(x, lo, hi);
      exchange<MAP_C, MAP_B, NR>(x, lds);
      phase<true, MAP_B, LO>(x, lo, hi);
      exchange<MAP_B, MAP_A, NR>(x, lds);
      phase<true, MAP_A, LO>(x, lo, hi);
      reduce_all<LO>(x);
      store_tile(a, T, so, MODE == TM_WAW ? a.out1 : a.out2, x);
      if (PF && T + step < last) load_tile<PACK>(a, T + step, so, raw);
    }
    __syncthreads();  // the next tile's first exchange reuses the LDS rows
  }
}

// Two workgroups share a CU (66 KiB of LDS each) when a wave stays within 128
// VGPRs. F65521's WHT . alpha . WHT modes need more registers and run one
// workgroup per CU rather than spill; KoalaBear's fit.
constexpr bool wide_mode(int mode) { return LANES == 8 && (mode == TM_WAW || mode == TM_COPY_WAW); }
template <int MODE, bool PACK, int LO>
__global__ __launch_bounds__(TWG) __attribute__((amdgpu_waves_per_eu(4))) void k_tile(TileArgs a) {
  tile_body<MODE, PACK, LO, false>(a);
}
template <int MODE, bool PACK, int LO>
__global__ __launch_bounds__(TWG) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_tile_wide(TileArgs a) {
  tile_body<MODE, PACK, LO, true>(a);
}

// ---------------------------------------------------------------- 8192-element tiles
// A pass over 10 strided index bits leaves a 4096-element tile 4 contiguous
// positions per tile row: 64-byte rows, which HBM serves at about 3.1 TB/s
// (128-byte rows 4.6, 256 bytes and up 5.0 TB/s; measured with the
// transforms switched off). The big tile holds 8192 elements, 8 contiguous
// positions (128 bytes) x 2^10 rows, in one workgroup of 1024 threads per CU
// (132 KiB of LDS). Each thread keeps 8 elements as in the 4096-element form;
// the tile's 13 bits go through mappings with register bits 10..12 (A),
// 7..9 (B), 4..6 (C) and 1..3 (D, of which only bit 3 is transformed). Tile
// bits 0..2 are the contiguous positions and are never transformed.
// Modes M, M_W and W.
constexpr int BWG = 1024;
constexpr int BN = BWG * TQ;            // 8192 elements
constexpr int BPAD = BN + BN / 32;
enum : int { BMAP_A = 10, BMAP_B = 7, BMAP_C = 4, BMAP_D = 1 };

// Tile element of slot q in thread tid. A and B keep the thread bits in order
// (a wave covers 8 contiguous positions x 8 rows in A, the load/store layout);
// C and D permute them so that the 8-byte LDS accesses of every 32 lanes hit
// distinct bank pairs under lds_slot (lane-bit weights 1, 2, 4, 8, 16 mod 32):
// C: tid bits -> e bits 0 1 2 3 9 7 8 10 11 12 (e bit 9 weighs 528 = 16 mod 32);
// D: tid bits -> e bits 0 6 7 8 4 5 9 10 11 12 (66, 132, 264 = 2, 4, 8 mod 32).
template <int MAP>
__device__ inline int btile_e(int tid, int q) {
  if (MAP == BMAP_C)
    return (tid & 15) | (q << 4) | (((tid >> 4) & 1) << 9) | (((tid >> 5) & 3) << 7) | ((tid >> 7) << 10);
  if (MAP == BMAP_D)
    return (tid & 1) | (q << 1) | (((tid >> 1) & 7) << 6) | (((tid >> 4) & 3) << 4) | (((tid >> 6) & 1) << 9) |
           ((tid >> 7) << 10);
  return (tid & ((1 << MAP) - 1)) | (q << MAP) | ((tid >> MAP) << (MAP + TQB));
}

// lds_slot is additive over disjoint bit fields (e + (e >> 5) with no carry
// between them), so slot q of a mapping is the thread's base slot plus a
// compile-time offset: one base register per mapping and LDS row.
template <int MAP>
__device__ inline int bslot_base(int tid) { return lds_slot(btile_e<MAP>(tid, 0)); }
template <int MAP>
constexpr int bslot_q(int q) {
  return MAP == BMAP_D ? (q << 1) + ((q << 1) >> 5) : (q << MAP) + ((q << MAP) >> 5);
}

template <int FROM, int TO>
__device__ inline void bexchange(int32_t (&x)[TQ][LANES], uint2 (*lds)[BPAD]) {
  constexpr int XROUNDS = LANES / (2 * NR);
  const int tid = threadIdx.x;
  const int bf = bslot_base<FROM>(tid), bt = bslot_base<TO>(tid);
#pragma unroll
  for (int rnd = 0; rnd < XROUNDS; rnd++) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NR; p++) {
      uint2* row = lds[p] + bf;
#pragma unroll
      for (int q = 0; q < TQ; q++) {
        const int k = 2 * (rnd * NR + p);
        row[bslot_q<FROM>(q)] = make_uint2(uint32_t(x[q][k]), uint32_t(x[q][k + 1]));
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NR; p++) {
      const uint2* row = lds[p] + bt;
#pragma unroll
      for (int q = 0; q < TQ; q++) {
        const int k = 2 * (rnd * NR + p);
        const uint2 v = row[bslot_q<TO>(q)];
        x[q][k] = int32_t(v.x);
        x[q][k + 1] = int32_t(v.y);
      }
    }
  }
}

// Transform over tile bits 3..12 (the 10 strided bits), from and back to A.
template <bool WHT>
__device__ inline void btransform(int32_t (&x)[TQ][LANES], uint2 (*lds)[BPAD]) {
  phase<WHT, BMAP_A, 3>(x, 3, 13);
  bexchange<BMAP_A, BMAP_B>(x, lds);
  phase<WHT, BMAP_B, 3>(x, 3, 13);
  bexchange<BMAP_B, BMAP_C>(x, lds);
  phase<WHT, BMAP_C, 3>(x, 3, 13);
  bexchange<BMAP_C, BMAP_D>(x, lds);
  phase<WHT, BMAP_D, 3>(x, 3, 13);
  bexchange<BMAP_D, BMAP_A>(x, lds);
}

// Workgroups take tiles blockIdx, blockIdx + grid, ... (one tile each unless
// persistent, one per CU): the tiles in flight at one time are neighbours
// along the contiguous bits.
template <int MODE>
__global__ __launch_bounds__(BWG) void k_btile(TileArgs a) {
  static_assert(MODE == TM_M || MODE == TM_M_W || MODE == TM_W, "big tiles run M, M_W and W");
  __shared__ uint2 lds[NR][BPAD];
  const uint32_t in_t = 16u * tile_off(a, btile_e<BMAP_A>(threadIdx.x, 0), a.in_vstride);
  const uint32_t out_t = 16u * tile_off(a, btile_e<BMAP_A>(threadIdx.x, 0), a.out_vstride);
  for (uint32_t T = blockIdx.x; T < a.tiles; T += gridDim.x) {
    uint64_t vec0, pos0;
    tile_base(a, T, vec0, pos0);
    int32_t x[TQ][LANES];
    {
      const char* base = reinterpret_cast<const char*>(static_cast<const E8*>(a.in) + vec0 * a.in_vstride + pos0);
      uint4 raw[TQ];
#pragma unroll
      for (int q = 0; q < TQ; q++)
        raw[q] = *reinterpret_cast<const uint4*>(base + 16 * tile_off64(a, BWG * q, a.in_vstride) + in_t);
#pragma unroll
      for (int q = 0; q < TQ; q++) unpack_e8(raw[q], x[q]);
    }
    auto store = [&](E8* out) {
      char* base = reinterpret_cast<char*>(out + vec0 * a.out_vstride + pos0);
#pragma unroll
      for (int q = 0; q < TQ; q++)
        *reinterpret_cast<uint4*>(base + 16 * tile_off64(a, BWG * q, a.out_vstride) + out_t) = pack_e8(x[q]);
    };
    if (MODE == TM_M || MODE == TM_M_W) {
      btransform<false>(x, lds);
      reduce_all<3>(x);
      store(a.out1);
    }
    if (MODE == TM_M_W || MODE == TM_W) {
      btransform<true>(x, lds);
      reduce_all<3>(x);
      store(MODE == TM_W ? a.out1 : a.out2);
    }
    __syncthreads();  // the next tile's first exchange reuses the LDS rows
  }
}

// ---------------------------------------------------------------- quad-split tiles
// A second kernel family for the same passes, with fewer LDS exchanges. A
// thread holds one 32-bit word d of 32 tile elements (the four words of a
// 16-byte element sit in the four lanes of a quad, lane bits 0..1 = d): two
// u16 coefficients, as lazy signed i32 lanes (F65521), or one canonical u32
// coefficient (KoalaBear). Five tile bits live in the register slots, so a
// 12-bit tile takes two mappings:
//   H: slots = tile bits 7..11; lane bits 2..5 = tile bits 0..3, wave bits = 4..6;
//   L: slots = tile bits 2..6;  lane bits 2..5 = tile bits 0, 1, 7, 8, wave bits = 9..11.
// Tile bits 0 and 1 are lane bits 2 and 3 in L and take DPP butterflies across
// each row of 16 lanes. A transform over all 12 bits is one H phase, one LDS
// exchange and one L phase (the register-radix form above takes four
// mappings and four exchanges); a pass whose bits all lie in 7..11 needs no
// exchange. Both mappings keep four consecutive elements (64 bytes) in 16
// consecutive lanes, so loads and stores coalesce in either.
constexpr int QB = 5;        // tile bits held in register slots
constexpr int QN = 1 << QB;  // slots per thread
constexpr int QV = LANES / 4;  // coefficient lanes per 32-bit word (F65521: two u16, KoalaBear: one u32)
enum : int { QMAP_H = 0, QMAP_L = 1 };

template <int MAP>
constexpr int qmap_lo() { return MAP == QMAP_H ? 7 : 2; }
// tile element of slot q held by thread tid: thread part | slot part (disjoint bits)
template <int MAP>
__device__ inline int qe_thread(int tid) {
  const int u = (tid >> 2) & 15, w = tid >> 6;
  return MAP == QMAP_H ? (u | (w << 4)) : ((u & 3) | ((u >> 2) << 7) | (w << 9));
}
template <int MAP>
__host__ __device__ constexpr int qe_slot(int q) { return MAP == QMAP_H ? (q << 7) : (q << 2); }

// Register butterflies on the slot bits of mapping MAP whose tile bit lies in
// the transformed range ([LO, 12) at compile time, or [lo, hi) when LO < 0).
template <bool WHT, int MAP, int LO>
__device__ inline void qphase(int32_t (&x)[QN][QV], int lo, int hi) {
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < QB; b++) {
    const int tb = qmap_lo<MAP>() + b;
    if (LO >= 0) {
      if (tb < LO) continue;
    } else if (tb < lo || tb >= hi) {
      continue;
    }
#pragma unroll
    for (int q = 0; q < QN; q++) {
      if (q & (1 << b)) continue;
      const int r = q | (1 << b);
#pragma unroll
      for (int k = 0; k < QV; k++) {
        const int32_t u = x[q][k], w = x[r][k];
        if (WHT) { x[q][k] = bf_add(u, w); x[r][k] = bf_sub(u, w); }
        else x[r][k] = bf_sub(w, u);
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

// Partner values across lane bit 2 (tile bit 0 in L) and lane bit 3 (tile bit
// 1): row_ror:N gives lane i the value of lane (i - N) mod 16 of its row, and
// the bank mask picks the groups of four lanes that take it.
__device__ inline int32_t dpp_partner4(int32_t v) {
  const int32_t up = __builtin_amdgcn_update_dpp(0, v, 0x12C, 0xf, 0x5, false);  // banks 0, 2: lane i + 4
  return __builtin_amdgcn_update_dpp(up, v, 0x124, 0xf, 0xa, false);            // banks 1, 3: lane i - 4
}
__device__ inline int32_t dpp_partner8(int32_t v) { return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xf, 0xf, false); }

// One butterfly with the partner value p; m = -1 in the lanes that hold the
// element with the bit set, else 0. WHT: (x + p, p - x); Mobius: (x, x - p).
template <bool WHT>
__device__ inline int32_t dpp_bfly(int32_t x, int32_t p, int32_t m) {
#ifdef QF_KB
  if (WHT) {
    const uint32_t v = uint32_t(p) + ((uint32_t(x) ^ uint32_t(m)) - uint32_t(m));  // x + p or p - x (wrapping)
    const uint32_t fix = m ? P : 0u - P;
    const uint32_t w = v + fix;
    return int32_t(v < w ? v : w);
  }
  const uint32_t v = uint32_t(x) - (uint32_t(p) & uint32_t(m));
  const uint32_t w = v + P;
  return int32_t(v < w ? v : w);
#else
  if (WHT) return p + ((x ^ m) - m);
  return x - (p & m);
#endif
}

// DPP butterflies on tile bits 0 and 1 (mapping L only) within the range.
template <bool WHT, int LO>
__device__ inline void qphase_dpp(int32_t (&x)[QN][QV], int lo, int hi) {
  const int lane = threadIdx.x & 63;
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < 2; b++) {
    if (LO >= 0) {
      if (b < LO) continue;
    } else if (b < lo || b >= hi) {
      continue;
    }
    const int32_t m = -int32_t((lane >> (2 + b)) & 1);
#pragma unroll
    for (int q = 0; q < QN; q++) {
#pragma unroll
      for (int k = 0; k < QV; k++) {
        const int32_t v = x[q][k];
        const int32_t p = b ? dpp_partner8(v) : dpp_partner4(v);
        x[q][k] = dpp_bfly<WHT>(v, p, m);
      }
      if (q % 2 == 1) __builtin_amdgcn_sched_barrier(0);  // bounds the partner values in flight
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

// LDS word of (element e, word d): W = 4 e + d with W bits 9..10 (tile bits 7,
// 8) folded into bits 4..5, A = W ^ (((W >> 9) & 3) << 4). The lanes of a wave
// then reach distinct banks in both mappings (H varies W bits 0..5, L varies
// W bits 0..3 and 9..10). Per thread: base[j] for slot bits 0..1 = j, and a
// slot offset for the rest (an immediate of the LDS instruction).
struct QLds {
  uint32_t h[4], l[4];
};
__device__ inline QLds qlds_bases() {
  const int tid = threadIdx.x, d = tid & 3;
  QLds o;
  const uint32_t wh = 4u * uint32_t(qe_thread<QMAP_H>(tid)) + d;  // below 512: slot bits 0..1 xor in
  const uint32_t wl = 4u * uint32_t(qe_thread<QMAP_L>(tid)) + d;
  const uint32_t kl = (wl >> 9) & 3;                                 // tile bits 7, 8 of the thread part
#pragma unroll
  for (int j = 0; j < 4; j++) {
    o.h[j] = 4u * (wh ^ (uint32_t(j) << 4));
    o.l[j] = 4u * (wl + ((uint32_t(j) ^ kl) << 4));
  }
  return o;
}
template <int MAP>
__device__ inline uint32_t qlds_addr(const QLds& b, int q) {  // byte address
  return MAP == QMAP_H ? b.h[q & 3] + 4u * (uint32_t(q) << 9) : b.l[q & 3] + 4u * (uint32_t(q >> 2) << 6);
}

template <int FROM, int TO>
__device__ inline void qexchange(int32_t (&x)[QN][QV], uint32_t* lds, const QLds& b) {
  char* base = reinterpret_cast<char*>(lds);
#pragma unroll
  for (int k = 0; k < QV; k++) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QN; q++) *reinterpret_cast<int32_t*>(base + qlds_addr<FROM>(b, q)) = x[q][k];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QN; q++) x[q][k] = *reinterpret_cast<const int32_t*>(base + qlds_addr<TO>(b, q));
  }
}

// A full transform over the range: the phases of the bits in each mapping,
// from mapping FROM to the other one (TO), with one exchange between them.
template <bool WHT, int MAP, int LO>
__device__ inline void qphase_all(int32_t (&x)[QN][QV], int lo, int hi) {
  qphase<WHT, MAP, LO>(x, lo, hi);
  if (MAP == QMAP_L) qphase_dpp<WHT, LO>(x, lo, hi);
}

#ifndef QF_KB
// LO >= 2: at most ten butterfly stages ran since the values were canonical.
template <int LO>
__device__ inline void qreduce(int32_t (&x)[QN][QV]) {
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < QN; q++)
#pragma unroll
    for (int k = 0; k < QV; k++) x[q][k] = int32_t(LO >= 2 ? reduce_lazy10(x[q][k]) : reduce_lazy(x[q][k]));
  __builtin_amdgcn_sched_barrier(0);
}
__device__ inline uint32_t qpack(const int32_t (&v)[QV]) { return uint32_t(v[0]) | (uint32_t(v[1]) << 16); }
__device__ inline void qunpack(uint32_t w, int32_t (&v)[QV]) { v[0] = int32_t(w & 0xffffu); v[1] = int32_t(w >> 16); }
// two i8 BVA entries (bytes 2d, 2d + 1 of the element) -> canonical lanes
__device__ inline void qunpack_bva(uint32_t w, int32_t (&v)[QV]) {
  v[0] = int32_t(from_val(int8_t(w & 0xffu)));
  v[1] = int32_t(from_val(int8_t((w >> 8) & 0xffu)));
}
#else
template <int LO>
__device__ inline void qreduce(int32_t (&)[QN][QV]) {}
__device__ inline uint32_t qpack(const int32_t (&v)[QV]) { return uint32_t(v[0]); }
__device__ inline void qunpack(uint32_t w, int32_t (&v)[QV]) { v[0] = int32_t(w); }
// one i16 BVA entry (bytes 2d, 2d + 1 of the element) -> a canonical lane
__device__ inline void qunpack_bva(uint32_t w, int32_t (&v)[QV]) { v[0] = int32_t(from_val(int16_t(w & 0xffffu))); }
#endif

// Offsets of a mapping's slots: a 32-bit per-thread part (bytes) and the
// uniform per-slot part (64-bit, from the slot's tile bits).
struct QOffs {
  uint32_t in_t, out_t, alpha_t;
};
template <int MAP>
__device__ inline QOffs qoffs(const TileArgs& a, bool pack) {
  const int tid = threadIdx.x, d = tid & 3, e = qe_thread<MAP>(tid);
  QOffs o;
  o.in_t = pack ? 8u * tile_off(a, e, a.pack_n) + 2u * d : 16u * tile_off(a, e, a.in_vstride) + 4u * d;
  o.out_t = 16u * tile_off(a, e, a.out_vstride) + 4u * d;
  o.alpha_t = 16u * tile_off(a, e, 0);
  return o;
}

// Slot q of a mapping is at byte (thread part + q x slot stride) from the
// tile's first element, below 2^32 (checked on the host): a buffer resource
// over the tile takes the thread part as the VGPR offset and the slot part as
// the scalar offset, so the 32 slots share one address register.
__device__ inline __amdgpu_buffer_rsrc_t qrsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0x7fffffff, 0x00020000);
}

template <int MAP, bool PACK>
__device__ inline void qload(const TileArgs& a, uint64_t T, const QOffs& o, uint32_t (&raw)[QN]) {
  uint64_t vec0, pos0;
  tile_base(a, T, vec0, pos0);
  if (PACK) {
    const uint64_t b0 = a.block_first + 8 * (vec0 * a.pack_n + pos0);
    const uint8_t* bva = static_cast<const uint8_t*>(a.in);
    const uint32_t sb = uint32_t(8 * a.qs_in[MAP]);
    if (b0 + a.pack_span <= a.available) {  // the whole tile lies before the active end
      const __amdgpu_buffer_rsrc_t r = qrsrc(bva + b0);
#pragma unroll
      for (int q = 0; q < QN; q++) raw[q] = __builtin_amdgcn_raw_buffer_load_b16(r, o.in_t, uint32_t(q) * sb, 0);
    } else {  // bytes at or past the active end read as zero
#pragma unroll
      for (int q = 0; q < QN; q++) {
        const uint64_t byte = b0 + o.in_t + uint64_t(q) * sb;
        if (byte + 2 <= a.available) raw[q] = *reinterpret_cast<const uint16_t*>(bva + byte);
        else raw[q] = byte < a.available ? uint32_t(bva[byte]) : 0u;
      }
    }
  } else {
    const char* base = reinterpret_cast<const char*>(static_cast<const E8*>(a.in) + vec0 * a.in_vstride + pos0);
    const __amdgpu_buffer_rsrc_t r = qrsrc(base);
    const uint32_t sb = uint32_t(16 * a.qs_in[MAP]);
#pragma unroll
    for (int q = 0; q < QN; q++) raw[q] = __builtin_amdgcn_raw_buffer_load_b32(r, o.in_t, uint32_t(q) * sb, 0);
  }
}

template <int MAP>
__device__ inline void qstore(const TileArgs& a, uint64_t T, const QOffs& o, E8* out, const int32_t (&x)[QN][QV]) {
  uint64_t vec0, pos0;
  tile_base(a, T, vec0, pos0);
  const __amdgpu_buffer_rsrc_t r = qrsrc(out + vec0 * a.out_vstride + pos0);
  const uint32_t sb = uint32_t(16 * a.qs_out[MAP]);
#pragma unroll
  for (int q = 0; q < QN; q++) __builtin_amdgcn_raw_buffer_store_b32(qpack(x[q]), r, o.out_t, uint32_t(q) * sb, 0);
}

// x <- alpha . x in mapping L for canonical x. The four words of slot q's
// element sit in the four lanes of a quad; two 2 x 2 transposes across the
// quad (quad_perm) give lane d the whole element of slot 4 g + d for g = 0..7,
// which is multiplied by its alpha as one extension product, and transposed back.
__device__ inline void quad_transpose(uint32_t (&y)[QN]) {
  const int lane = threadIdx.x & 63;
  const bool b0 = lane & 1, b1 = lane & 2;
#pragma unroll
  for (int g = 0; g < QN; g += 4) {
#pragma unroll
    for (int h = 0; h < 2; h++) {  // slot bit 0 <-> lane bit 0
      const uint32_t e = y[g + 2 * h], o = y[g + 2 * h + 1];
      const uint32_t pe = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(e), 0xB1, 0xf, 0xf, false));  // quad_perm(1,0,3,2)
      const uint32_t po = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(o), 0xB1, 0xf, 0xf, false));
      y[g + 2 * h] = b0 ? po : e;
      y[g + 2 * h + 1] = b0 ? o : pe;
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {  // slot bit 1 <-> lane bit 1
      const uint32_t e = y[g + h], o = y[g + h + 2];
      const uint32_t pe = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(e), 0x4E, 0xf, 0xf, false));  // quad_perm(2,3,0,1)
      const uint32_t po = uint32_t(__builtin_amdgcn_mov_dpp(int32_t(o), 0x4E, 0xf, 0xf, false));
      y[g + h] = b1 ? po : e;
      y[g + h + 2] = b1 ? o : pe;
    }
  }
}
__device__ inline void qmul_alpha(const TileArgs& a, uint64_t T, const QOffs& o, int32_t (&x)[QN][QV]) {
  uint64_t vec0, pos0;
  tile_base(a, T, vec0, pos0);
  const int d = threadIdx.x & 3;
  const __amdgpu_buffer_rsrc_t ar = qrsrc(a.alpha + pos0);
  const uint32_t at = o.alpha_t + uint32_t(d) * uint32_t(16 * a.qs_alpha);
  const uint32_t sa = uint32_t(16 * a.qs_alpha);
  uint32_t y[QN];
#pragma unroll
  for (int q = 0; q < QN; q++) y[q] = qpack(x[q]);
  quad_transpose(y);
#pragma unroll
  for (int g = 0; g < QN; g += 4) {  // element of slot g + d: words y[g .. g + 3]
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 bv = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(ar, at, uint32_t(g) * sa, 0));
    const uint4 bw = make_uint4(bv.x, bv.y, bv.z, bv.w);
    int32_t c[LANES];
    unpack_e8(make_uint4(y[g], y[g + 1], y[g + 2], y[g + 3]), c);
    mul_alpha(c, bw);
    const uint4 r = pack_e8(c);
    y[g] = r.x; y[g + 1] = r.y; y[g + 2] = r.z; y[g + 3] = r.w;
    __builtin_amdgcn_sched_barrier(0);
  }
  quad_transpose(y);
#pragma unroll
  for (int q = 0; q < QN; q++) qunpack(y[q], x[q]);
}

// Transform over the range starting in mapping FROM: returns in the other
// mapping when both hold transformed bits, else stays in FROM.
template <int MAP>
constexpr int qother() { return MAP == QMAP_H ? QMAP_L : QMAP_H; }

template <int MODE, bool PACK, int LO, bool PF>
__device__ inline void qtile_body(const TileArgs& a) {
  __shared__ uint32_t lds[16384];  // 4096 elements x 4 words
  uint64_t first, last, step;
  if (a.xcd_span) {
    const uint64_t g = blockIdx.x % 8, slots = gridDim.x / 8;
    first = g * a.xcd_span + blockIdx.x / 8;
    last = (g + 1) * a.xcd_span < a.tiles ? (g + 1) * a.xcd_span : a.tiles;
    step = slots;
  } else {
    first = uint64_t(blockIdx.x) * a.per_wg;
    last = first + a.per_wg < a.tiles ? first + a.per_wg : a.tiles;
    step = 1;
  }
  if (first >= last) return;
  const int lo = a.xlo, hi = a.xhi;
  // which mappings hold transformed bits (uniform)
  const bool in_h = (LO >= 0 ? true : hi > 7);
  const bool in_l = (LO >= 0 ? LO < 7 : lo < 7);
  const QLds lb = qlds_bases();
  const QOffs oh = qoffs<QMAP_H>(a, PACK), ol = qoffs<QMAP_L>(a, PACK);
  // The transform's bits may be taken in either order. alpha . WHT loads in
  // L, where the alpha product runs; WHT . alpha . WHT starts in H so that its
  // alpha product falls in L between the two halves; packing loads in H (128
  // contiguous BVA bytes per wave); every other mode starts in L when L holds
  // transformed bits (fewer registers live across its DPP stages).
  constexpr bool AW = MODE == TM_AW;
  constexpr bool WAW = MODE == TM_WAW || MODE == TM_COPY_WAW;
  const bool start_h = AW ? false : (WAW || PACK) ? in_h : !in_l;
  uint32_t raw[QN];
  auto load = [&](uint64_t T) {
    if (start_h) qload<QMAP_H, PACK>(a, T, oh, raw);
    else qload<QMAP_L, PACK>(a, T, ol, raw);
  };
  if (PF) load(first);
  for (uint64_t T = first; T < last; T += step) {
    int32_t x[QN][QV];
    if (!PF) load(T);
#pragma unroll
    for (int q = 0; q < QN; q++) {
      if (PACK) qunpack_bva(raw[q], x[q]);
      else qunpack(raw[q], x[q]);
    }
    if (PF && T + step < last) load(T + step);
    // cur: the mapping the data is in (uniform)
    bool cur_h = start_h;
    auto store = [&](E8* out) {
      if (cur_h) qstore<QMAP_H>(a, T, oh, out, x);
      else qstore<QMAP_L>(a, T, ol, out, x);
    };
    // one transform over the range: the current mapping's bits, then (if the
    // other mapping holds transformed bits) an exchange and its bits
    auto xform = [&](auto wht) {
      constexpr bool W = decltype(wht)::value;
      if (cur_h) {
        qphase_all<W, QMAP_H, LO>(x, lo, hi);
        if (in_l) { qexchange<QMAP_H, QMAP_L>(x, lds, lb); cur_h = false; qphase_all<W, QMAP_L, LO>(x, lo, hi); }
      } else {
        qphase_all<W, QMAP_L, LO>(x, lo, hi);
        if (in_h) { qexchange<QMAP_L, QMAP_H>(x, lds, lb); cur_h = true; qphase_all<W, QMAP_H, LO>(x, lo, hi); }
      }
    };
    using Mob = std::integral_constant<bool, false>;
    using Wht = std::integral_constant<bool, true>;
    if (MODE == TM_COPY_W || MODE == TM_COPY_WAW) store(a.out1);
    if (AW) {  // x canonical from memory, loaded in L
      qmul_alpha(a, T, ol, x);
      xform(Wht{});
      qreduce<LO>(x);
      store(a.out1);
    }
    if (MODE == TM_M || MODE == TM_M_W) {
      xform(Mob{});
      qreduce<LO>(x);
      store(a.out1);
    }
    if (MODE == TM_M_W || MODE == TM_W || MODE == TM_COPY_W) {
      xform(Wht{});
      qreduce<LO>(x);
      store(MODE == TM_W ? a.out1 : a.out2);
    }
    if (WAW) {
      xform(Wht{});
      qreduce<LO>(x);
      if (cur_h) {  // the alpha product runs in L
        qexchange<QMAP_H, QMAP_L>(x, lds, lb);
        cur_h = false;
      }
      qmul_alpha(a, T, ol, x);
      xform(Wht{});
      qreduce<LO>(x);
      store(MODE == TM_WAW ? a.out1 : a.out2);
    }
    __syncthreads();  // the next tile's first exchange reuses the LDS
  }
}

template <int MODE, bool PACK, int LO>
__global__ __launch_bounds__(TWG) __attribute__((amdgpu_waves_per_eu(4))) void k_qtile(TileArgs a) {
  qtile_body<MODE, PACK, LO, false>(a);
}
template <int MODE, bool PACK, int LO>
__global__ __launch_bounds__(TWG) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_qtile_wide(TileArgs a) {
  qtile_body<MODE, PACK, LO, true>(a);
}

}  // namespace

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

namespace {
// Launch shape of a tile pass: kernel variant (lean: at most 128 registers, no
// prefetch; wide: prefetches the next tile of a persistent workgroup),
// workgroups per CU when persistent, and the XCD-paired walk of 64-byte-row
// passes.
struct TileShape {
  bool wide;
  int per_cu;
  bool xcd;
  bool quad = false;  // the quad-split kernels (k_qtile) instead of the register-radix ones (k_tile)
  bool big = false;   // the 8192-element tiles (k_btile), one workgroup per CU
  bool one = false;   // one tile per workgroup instead of persistent workgroups (wide, per_cu unused)
};

bool xcd_applicable(const TileArgs& a, uint64_t wgs) {
  return a.logC == 2 && a.log_lo_chunks >= 1 && wgs % 16 == 0 && a.tiles >= 2 * wgs;
}

// The same pass in 8192-element tiles: 8 contiguous positions x the 10 bits.
TileArgs big_args(TileArgs a) {
  a.logC = 13 - a.t;
  a.log_lo_chunks = a.s - a.logC;
  a.tiles /= 2;
  a.xlo = a.logC;
  a.xhi = 13;
  return a;
}

// F65521's built-in shapes launch one tile per workgroup: the dispatcher then
// mixes workgroups of the commit's other stream (the leaf hash of the previous
// block) into the CUs between tiles (8B commit 79.3 -> 78.0 ms, folded block
// 0.3 ms faster). KoalaBear's keep persistent workgroups that each walk a
// range of tiles (sh.per_cu per CU; 8B commit 154 against 157 ms one tile
// each). QF_TILE_PERSISTENT=1 or 0 picks either form in both fields; the
// autotuner tries both.
bool tile_persistent() {
  static const bool on = [] {
    const char* e = std::getenv("QF_TILE_PERSISTENT");
#ifdef QF_KB
    return !(e && e[0] == '0');
#else
    return e && e[0] == '1';
#endif
  }();
  return on;
}

void launch_shape(int mode, bool pack, TileArgs a, const TileShape& sh, hipStream_t st) {
  const bool persistent = !sh.one;
  if (sh.big) {
    a = big_args(a);
    const uint32_t grid = persistent ? uint32_t(std::min<uint64_t>(a.tiles, uint64_t(device_cus()))) : a.tiles;
    dim3 g{grid, 1, 1}, b{uint32_t(BWG), 1, 1};
    switch (mode) {
      case TM_M: hipLaunchKernelGGL(k_btile<TM_M>, g, b, 0, st, a); break;
      case TM_M_W: hipLaunchKernelGGL(k_btile<TM_M_W>, g, b, 0, st, a); break;
      default: hipLaunchKernelGGL(k_btile<TM_W>, g, b, 0, st, a); break;
    }
    return;
  }
  const uint64_t tiles = a.tiles;
  const uint64_t wgs = std::min<uint64_t>(tiles, uint64_t(sh.per_cu) * uint64_t(device_cus()));
  a.per_wg = uint32_t((tiles + wgs - 1) / wgs);
  uint32_t grid = uint32_t((tiles + a.per_wg - 1) / a.per_wg);
  a.xcd_span = 0;
  // 64-byte tile rows: pair the tiles that share 128-byte lines across sibling
  // workgroups of one XCD (tiles 2m and 2m+1 differ in the lowest chunk bit)
  if (sh.xcd && xcd_applicable(a, wgs)) {
    a.xcd_span = uint32_t(((tiles + 7) / 8 + 1) & ~uint64_t(1));
    grid = uint32_t(wgs);
  }
  if (!persistent) {  // one tile per workgroup (with the same XCD pairing)
    a.per_wg = 1;
    grid = uint32_t(tiles);
    a.xcd_span = 0;
    if (sh.xcd && a.logC == 2 && a.log_lo_chunks >= 1) {
      a.xcd_span = uint32_t(((tiles + 7) / 8 + 1) & ~uint64_t(1));
      grid = 8 * a.xcd_span;
    }
  }
  dim3 g{grid, 1, 1}, b{uint32_t(TWG), 1, 1};
  // compile-time stage ranges for the full-width tiles ([logC, 12)), runtime otherwise
  const int lo_ct = a.xhi == 12 && (a.xlo == 0 || a.xlo == 2 || a.xlo == 7) ? a.xlo : -1;
#define QF_TILE_LAUNCH_LO(M, PK, LOV)                                   \
  do {                                                                  \
    if (sh.quad && sh.wide)                                             \
      hipLaunchKernelGGL((k_qtile_wide<M, PK, LOV>), g, b, 0, st, a);    \
    else if (sh.quad)                                                   \
      hipLaunchKernelGGL((k_qtile<M, PK, LOV>), g, b, 0, st, a);         \
    else if (sh.wide)                                                   \
      hipLaunchKernelGGL((k_tile_wide<M, PK, LOV>), g, b, 0, st, a);     \
    else                                                                \
      hipLaunchKernelGGL((k_tile<M, PK, LOV>), g, b, 0, st, a);          \
  } while (0)
#define QF_TILE_LAUNCH(M, PK)                        \
  do {                                               \
    switch (lo_ct) {                                 \
      case 0: QF_TILE_LAUNCH_LO(M, PK, 0); break;    \
      case 2: QF_TILE_LAUNCH_LO(M, PK, 2); break;    \
      case 7: QF_TILE_LAUNCH_LO(M, PK, 7); break;    \
      default: QF_TILE_LAUNCH_LO(M, PK, -1); break;  \
    }                                                \
  } while (0)
  if (pack) {
    QF_TILE_LAUNCH(TM_M, true);
  } else {
    switch (mode) {
      case TM_M: QF_TILE_LAUNCH(TM_M, false); break;
      case TM_M_W: QF_TILE_LAUNCH(TM_M_W, false); break;
      case TM_W: QF_TILE_LAUNCH(TM_W, false); break;
      case TM_WAW: QF_TILE_LAUNCH(TM_WAW, false); break;
      case TM_COPY_W: QF_TILE_LAUNCH(TM_COPY_W, false); break;
      case TM_COPY_WAW: QF_TILE_LAUNCH(TM_COPY_WAW, false); break;
      default: QF_TILE_LAUNCH(TM_AW, false); break;
    }
  }
#undef QF_TILE_LAUNCH
#undef QF_TILE_LAUNCH_LO
}

// QF_QTILE=0 keeps every pass on the register-radix kernels.
bool qtile_on() {
  const char* e = std::getenv("QF_QTILE");
  return !(e && e[0] == '0');
}
// QF_BTILE=0 keeps 10-bit strided passes on 4096-element tiles; QF_BTILE=1
// makes the 8192-element tiles the built-in shape of every pass they serve.
bool btile_on() {
  const char* e = std::getenv("QF_BTILE");
  return !(e && e[0] == '0');
}
// The 8192-element tiles for the M_W pass over 10 strided bits (2.16 against
// 2.33 ms on the 8B block; the W pass over the same bits is as fast in either
// form and keeps the 4096-element default unless the autotuner finds the big
// tile faster), else the quad-split kernels when they apply (their per-thread
// offsets fit 32 bits), else the register-radix ones.
TileShape builtin_shape(int mode, bool quad_ok, bool big_ok) {
  if (big_ok && btile_on() && (mode == TM_M_W || std::getenv("QF_BTILE"))) return {false, 1, false, false, true};
  if (quad_ok && qtile_on()) return {false, 2, true, true};  // quad_ok: a compile-time bit range
  const bool w = wide_mode(mode);
  return {w, w ? 1 : 2, true, false};
}
TileShape default_shape(int mode, bool quad_ok, bool big_ok) {
#ifndef QF_KB
  // F65521's WHT . alpha . WHT pass, bound by its vector ALU work, keeps two
  // persistent register-radix workgroups per CU beside the leaf hash (8B
  // commit 75.5 against 76.3 ms with the one-tile quad-split form)
  if ((mode == TM_WAW || mode == TM_COPY_WAW) && !std::getenv("QF_TILE_PERSISTENT"))
    return {false, 2, true, false, false, false};
#endif
  TileShape s = builtin_shape(mode, quad_ok, big_ok);
  if (!tile_persistent()) {  // one tile per workgroup: no prefetch, no per-CU count
    s.one = true;
    s.wide = false;
    s.per_cu = 1;
  }
  return s;
}

// ---------------------------------------------------------------- autotuner
// The first pass of each geometry in the process times the candidate launch
// shapes on scratch buffers laid out like the caller's (same strides, same
// overlap of input and outputs, zeroed), with HIP events on the pass's
// stream, and keeps the fastest; every candidate computes the same transform,
// so the choice never changes a result. The built-in shape (which also has
// the least HBM traffic: the XCD-paired walk fetches each line once) is kept
// unless another is faster by more than 5%, beyond run-to-run noise.
// QF_AUTOTUNE=0 keeps the built-in shapes. With QF_TUNE_FILE set, the picks
// are read from that file at first use and the file is rewritten after each
// new tuning, so a later process (a profiled run, say) launches exactly the
// shapes an earlier one chose instead of timing them again.
struct TuneKey {
  int mode, L, s, t;
  bool pack, xvec;
  uint64_t vectors, in_vs, out_vs;
  int out2_rel;  // out2 absent (0), out2 == out1 + 2^L (1), elsewhere (2)
  bool in_place;
  bool operator<(const TuneKey& o) const {
    return std::tie(mode, L, s, t, pack, xvec, vectors, in_vs, out_vs, out2_rel, in_place) <
           std::tie(o.mode, o.L, o.s, o.t, o.pack, o.xvec, o.vectors, o.in_vs, o.out_vs, o.out2_rel, o.in_place);
  }
};
struct TuneEntry {
  TileShape best;
  std::vector<std::pair<TileShape, float>> tried;  // candidate and its best time in ms
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry>& tune_table() {
  static auto* t = new std::map<TuneKey, TuneEntry>;  // lives until process exit
  return *t;
}
bool autotune_on() {
  const char* e = std::getenv("QF_AUTOTUNE");
  return !(e && e[0] == '0');
}

TuneEntry tune(int mode, bool pack, const TileArgs& a0, uint64_t vectors, int L, const TuneKey& key, hipStream_t st) {
  TuneEntry out{default_shape(mode, a0.quad_ok, a0.big_ok), {}};
  const uint64_t n = uint64_t(1) << L;
  const uint64_t out_ext = ((vectors - 1) * a0.out_vstride + n) * 16;
  // one scratch span for the E8 buffers, at the callers' relative offsets
  std::vector<std::pair<const char*, uint64_t>> bufs = {{reinterpret_cast<const char*>(a0.out1), out_ext}};
  if (a0.out2) bufs.push_back({reinterpret_cast<const char*>(a0.out2), out_ext});
  if (!pack) bufs.push_back({static_cast<const char*>(a0.in), ((vectors - 1) * a0.in_vstride + n) * 16});
  const char* lo = bufs[0].first;
  const char* hi = bufs[0].first + bufs[0].second;
  for (auto& b : bufs) { lo = std::min(lo, b.first); hi = std::max(hi, b.first + b.second); }
  const uint64_t span = uint64_t(hi - lo);
  const uint64_t pack_bytes = pack ? vectors * n * 8 : 0;
  if (span > (uint64_t(24) << 30)) return out;  // no scratch that large: keep the default
  char* scratch = nullptr;
  char* aux = nullptr;  // packed input bytes, or alpha
  const uint64_t aux_bytes = std::max<uint64_t>(pack_bytes, n * 16);
  if (hipMalloc(&scratch, span) != hipSuccess || hipMalloc(&aux, aux_bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (scratch) (void)hipFree(scratch);
    return out;
  }
  QF_HIP(hipMemsetAsync(scratch, 0, span, st));
  QF_HIP(hipMemsetAsync(aux, 0, aux_bytes, st));
  auto rebase = [&](const void* p) { return scratch + (static_cast<const char*>(p) - lo); };
  TileArgs a = a0;
  a.out1 = reinterpret_cast<E8*>(rebase(a0.out1));
  a.out2 = a0.out2 ? reinterpret_cast<E8*>(rebase(a0.out2)) : nullptr;
  a.alpha = reinterpret_cast<const E8*>(aux);
  if (pack) {
    a.in = aux;
    a.block_first = 0;
    a.available = pack_bytes;
  } else {
    a.in = rebase(a0.in);
  }
  // Every candidate keeps the XCD-paired walk where it applies: without it a
  // pass over 64-byte tile rows fetches every 128-byte line twice, for no
  // gain in time (the passes it applies to are bound by their VALU and LDS work).
  // Each kernel family persistent (lean or wide, 1 or 2 per CU) and one tile per workgroup.
  std::vector<TileShape> cands = {{false, 2, true}, {false, 1, true}, {true, 1, true}, {true, 2, true},
                                  {false, 1, true, false, false, true}};
  if (a0.quad_ok && qtile_on()) {
    cands.push_back({false, 2, true, true});
    cands.push_back({true, 1, true, true});
    cands.push_back({false, 1, true, true, false, true});
  }
  if (a0.big_ok && btile_on()) {
    cands.push_back({false, 1, false, false, true});
    cands.push_back({false, 1, false, false, true, true});
  }
  hipEvent_t e0, e1;
  QF_HIP(hipEventCreate(&e0));
  QF_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  const TileShape def = default_shape(mode, a0.quad_ok, a0.big_ok);
  float def_ms = -1;
  for (const TileShape& sh : cands) {
    launch_shape(mode, pack, a, sh, st);  // warm-up
    float t_best = 1e30f;
    for (int rep = 0; rep < 2; rep++) {
      QF_HIP(hipEventRecord(e0, st));
      launch_shape(mode, pack, a, sh, st);
      QF_HIP(hipEventRecord(e1, st));
      QF_HIP(hipEventSynchronize(e1));
      float ms = 0;
      QF_HIP(hipEventElapsedTime(&ms, e0, e1));
      t_best = std::min(t_best, ms);
    }
    QF_HIP(hipGetLastError());
    out.tried.push_back({sh, t_best});
    if (t_best < best) { best = t_best; out.best = sh; }
    if (sh.wide == def.wide && sh.per_cu == def.per_cu && sh.xcd == def.xcd && sh.quad == def.quad &&
        sh.big == def.big && sh.one == def.one)
      def_ms = t_best;
  }
  if (def_ms > 0 && best > def_ms / 1.05f) out.best = def;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  QF_HIP(hipStreamSynchronize(st));
  (void)hipFree(scratch);
  (void)hipFree(aux);
  (void)key;
  return out;
}

// The tuning file: one line per geometry, "key fields... wide per_cu xcd".
const char* tune_file() { return std::getenv("QF_TUNE_FILE"); }
void load_tune_file(std::map<TuneKey, TuneEntry>& tab) {
  const char* f = tune_file();
  if (!f) return;
  FILE* fp = std::fopen(f, "r");
  if (!fp) return;
  TuneKey k{};
  int pack = 0, xvec = 0, in_place = 0, wide = 0, per_cu = 0, xcd = 0, quad = 0, big = 0, one = 0;
  unsigned long long v = 0, iv = 0, ov = 0;
  while (std::fscanf(fp, "%d %d %d %d %d %d %llu %llu %llu %d %d %d %d %d %d %d %d", &k.mode, &k.L, &k.s, &k.t,
                     &pack, &xvec, &v, &iv, &ov, &k.out2_rel, &in_place, &wide, &per_cu, &xcd, &quad, &big,
                     &one) == 17) {
    k.pack = pack;
    k.xvec = xvec;
    k.in_place = in_place;
    k.vectors = v;
    k.in_vs = iv;
    k.out_vs = ov;
    if (per_cu == 1 || per_cu == 2)
      tab[k] = TuneEntry{{wide != 0, per_cu, xcd != 0, quad != 0, big != 0, one != 0}, {}};
  }
  std::fclose(fp);
}
void save_tune_file(const std::map<TuneKey, TuneEntry>& tab) {
  const char* f = tune_file();
  if (!f) return;
  FILE* fp = std::fopen(f, "w");
  if (!fp) return;
  for (auto& [k, e] : tab)
    std::fprintf(fp, "%d %d %d %d %d %d %llu %llu %llu %d %d %d %d %d %d %d %d\n", k.mode, k.L, k.s, k.t,
                 int(k.pack), int(k.xvec), (unsigned long long)k.vectors, (unsigned long long)k.in_vs,
                 (unsigned long long)k.out_vs, k.out2_rel, int(k.in_place), int(e.best.wide), e.best.per_cu,
                 int(e.best.xcd), int(e.best.quad), int(e.best.big), int(e.best.one));
  std::fclose(fp);
}

// Callers that may not allocate device memory turn tuning off on their thread:
// their passes take an earlier pick of the same geometry, else the built-in shape.
thread_local bool t_tune_allowed = true;

TileShape shape_for(int mode, bool pack, const TileArgs& a, uint64_t vectors, int L, bool xvec, hipStream_t st) {
  if (!autotune_on()) return default_shape(mode, a.quad_ok, a.big_ok);
  const uint64_t n = uint64_t(1) << L;
  TuneKey key{mode, L, a.s, a.t, pack, xvec, vectors, a.in_vstride, a.out_vstride,
              !a.out2 ? 0 : (a.out2 == a.out1 + n ? 1 : 2), !pack && a.in == static_cast<const void*>(a.out1)};
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto& tab = tune_table();
  static bool loaded = false;
  if (!loaded) {
    load_tune_file(tab);
    loaded = true;
  }
  auto it = tab.find(key);
  if (it == tab.end() && !t_tune_allowed) return default_shape(mode, a.quad_ok, a.big_ok);
  if (it == tab.end()) {
    it = tab.emplace(key, tune(mode, pack, a, vectors, L, key, st)).first;
    save_tune_file(tab);
  }
  return it->second.best;
}
}  // namespace

// One tile pass over `vectors` vectors of length 2^L, transforming bits [s, s+t)
// (and, with xform_vectors, also the index bits of the 2^logV vectors a tile
// holds). A tile holds 4096 elements: t + logC + logV = 12 with C = 2^logC
// contiguous positions (s >= logC) and 2^logV vectors (only when s = 0).
bool tile_pass(int mode, bool pack, const void* in, E8* out1, E8* out2, const E8* alpha, uint64_t vectors,
               uint64_t in_vstride, uint64_t out_vstride, int L, int s, int t, uint64_t block_first,
               uint64_t available, hipStream_t st, bool xform_vectors) {
  if (mode < TM_M || mode > TM_AW || (pack && mode != TM_M)) return false;
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
  if (t < 1 || t > 12 || s + t > L) return false;
  if (s == 0) {
    a.logC = 0;
    a.logV = 12 - t;
    if (!xform_vectors && t != std::min(L, 12)) return false;
    if (vectors % (uint64_t(1) << a.logV)) return false;
    a.log_lo_chunks = 0;
  } else {
    if (t > 10 || xform_vectors) return false;
    a.logC = 12 - t;
    a.logV = 0;
    if (s < a.logC) return false;
    a.log_lo_chunks = s - a.logC;
  }
  a.log_hi_count = L - s - t;
  // the per-thread byte offsets (tile bits 0..8 in mapping A, largest at
  // element 511 since the offset map is additive) must fit 32 bits
  const uint64_t thread_max = std::max(tile_off64(a, TWG - 1, in_vstride) * (pack ? 8 : 16),
                                       tile_off64(a, TWG - 1, out_vstride) * 16);
  if (thread_max + 16 >= (1ull << 32)) return false;
  a.xlo = a.logC;
  a.xhi = xform_vectors ? 12 : a.logC + a.t;
  // the quad-split mappings' thread parts: tile bits 0..6 (H) and 0, 1, 7..11 (L)
  const uint64_t qmax = std::max({tile_off64(a, 0x7f, in_vstride), tile_off64(a, 0xf83, in_vstride),
                                  tile_off64(a, 0x7f, out_vstride), tile_off64(a, 0xf83, out_vstride),
                                  tile_off64(a, 0xf83, 0)});
  a.quad_ok = qmax * 16 + 16 < (1ull << 32) && (!pack || tile_off64(a, 0xf83, a.pack_n) * 8 + 16 < (1ull << 32));
  // the quad kernels' runtime-range forms (partial ranges) run short of registers: the
  // register-radix kernels take those geometries
  a.quad_ok = a.quad_ok && a.xhi == 12 && (a.xlo == 0 || a.xlo == 2 || a.xlo == 7);
  // per-slot strides; the quad kernels address slot q as q x stride
  for (int m = 0; m < 2; m++) {
    const int one = m == 0 ? qe_slot<QMAP_H>(1) : qe_slot<QMAP_L>(1);
    a.qs_in[m] = tile_off64(a, one, pack ? a.pack_n : in_vstride);
    a.qs_out[m] = tile_off64(a, one, out_vstride);
    for (int q = 0; q < QN; q++) {
      const int e = m == 0 ? qe_slot<QMAP_H>(q) : qe_slot<QMAP_L>(q);
      if (tile_off64(a, e, pack ? a.pack_n : in_vstride) != uint64_t(q) * a.qs_in[m] ||
          tile_off64(a, e, out_vstride) != uint64_t(q) * a.qs_out[m] ||
          (m == 1 && tile_off64(a, e, 0) != uint64_t(q) * tile_off64(a, qe_slot<QMAP_L>(1), 0)))
        a.quad_ok = 0;
    }
  }
  a.qs_alpha = tile_off64(a, qe_slot<QMAP_L>(1), 0);
  // ... as byte offsets below 2^31 from the tile's base (the buffer range)
  {
    const uint64_t lim = 1ull << 31;
    const uint64_t in_w = pack ? 8 : 16, in_vs = pack ? a.pack_n : in_vstride;
    const uint64_t th[2] = {0x7f, 0xf83};
    for (int m = 0; m < 2; m++) {
      if (tile_off64(a, int(th[m]), in_vs) * in_w + 31 * in_w * a.qs_in[m] + 16 >= lim) a.quad_ok = 0;
      if (tile_off64(a, int(th[m]), out_vstride) * 16 + 31 * 16 * a.qs_out[m] + 16 >= lim) a.quad_ok = 0;
    }
    if (tile_off64(a, 0xf83, 0) * 16 + 31 * 16 * a.qs_alpha + 16 >= lim) a.quad_ok = 0;
    a.pack_span = pack ? 8 * tile_off64(a, TN - 1, a.pack_n) + 8 : 0;
  }
  const uint64_t tiles = (vectors << L) / TN;
  if (tiles == 0 || tiles >= (1ull << 31)) return false;
  a.tiles = uint32_t(tiles);
  // 8192-element tiles for the M, M_W and W passes over 10 strided bits: an
  // even number of 4096-element tiles, and per-thread offsets (tile bits 0..9
  // in mapping A) below 2^32 bytes
  {
    const TileArgs b = big_args(a);
    a.big_ok = !pack && s >= 3 && t == 10 && (mode == TM_M || mode == TM_M_W || mode == TM_W) && tiles % 2 == 0 &&
               std::max(tile_off64(b, BWG - 1, in_vstride), tile_off64(b, BWG - 1, out_vstride)) * 16 + 16 <
                   (1ull << 32);
  }
  launch_shape(mode, pack, a, shape_for(mode, pack, a, vectors, L, xform_vectors, st), st);
  QF_HIP(hipGetLastError());
  return true;
}

void set_tuning_allowed(bool allowed) { t_tune_allowed = allowed; }
bool quad_tiles_enabled() { return qtile_on(); }

// JSON array of the autotuner's choices so far (this library build).
std::string autotune_report() {
  std::lock_guard<std::mutex> g(g_tune_mu);
  std::string o = "[";
  auto shape = [](const TileShape& s) {
    return std::string("{\"wide\":") + (s.wide ? "true" : "false") + ",\"per_cu\":" + std::to_string(s.per_cu) +
           ",\"xcd\":" + (s.xcd ? "true" : "false") + ",\"quad\":" + (s.quad ? "true" : "false") +
           ",\"big\":" + (s.big ? "true" : "false") + ",\"one\":" + (s.one ? "true" : "false") + "}";
  };
  bool first = true;
  for (auto& [k, e] : tune_table()) {
    if (!first) o += ",";
    first = false;
    o += "{\"mode\":" + std::to_string(k.mode) + ",\"pack\":" + (k.pack ? "true" : "false") +
         ",\"L\":" + std::to_string(k.L) + ",\"s\":" + std::to_string(k.s) + ",\"t\":" + std::to_string(k.t) +
         ",\"vectors\":" + std::to_string(k.vectors) + ",\"best\":" + shape(e.best) + ",\"tried\":[";
    for (size_t i = 0; i < e.tried.size(); i++) {
      char ms[32];
      std::snprintf(ms, sizeof ms, "%.4f", e.tried[i].second);
      o += (i ? "," : "") + std::string("{\"shape\":") + shape(e.tried[i].first) + ",\"ms\":" + ms + "}";
    }
    o += "]}";
  }
  return o + "]";
}

}  // namespace qf
