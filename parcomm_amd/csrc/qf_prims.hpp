// Field and hash primitives shared by the HIP kernels and the host protocol code.
//
// The library is built once per field backend (the reference ships one crate
// per field):
//   default  F65521 = Z/(2^16 - 15) (qafold_short.rs:325-477) with the octic
//            extension F65521[X]/(X^8 - 17) (qafold_short.rs:482-660): eight
//            canonical u16 coefficients; BVA entries are i8.
//   QF_KB    KoalaBear = Z/(2^31 - 2^24 + 1) with the quartic extension
//            KoalaBear[X]/(X^4 - 3) (koalabear.rs:1-2, :131-145): four
//            canonical u32 coefficients; BVA entries are i16.
// Either way one extension element is 16 bytes, coefficient i = X^i, which is
// also its canonical wire encoding (full_extension.rs:157-164,
// koalabear.rs:229-246). The type keeps the name E8 in both builds.
//
// Device code keeps coefficients unpacked in 32-bit lanes. Sums use the
// unsigned-min trick: for s < 2P, min(s, s - P) (wrapping) is s mod P.
// F65521 products reduce with 2^16 = 15 (mod P); KoalaBear products use
// Montgomery reduction with R = 2^32.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>

namespace qf {

#ifdef QF_KB
constexpr uint32_t P = 2130706433;
constexpr int LANES = 4, LOG_LANES = 2;  // extension degree = BVA entries per packed element
using Coef = uint32_t;
using Val = int16_t;
constexpr uint32_t EXT_W = 3;             // X^4 = 3
constexpr uint32_t FROB_ROOT = 2113994754;  // x^p scales coefficient i by FROB_ROOT^i (koalabear.rs:155)
#else
constexpr uint32_t P = 65521;
constexpr int LANES = 8, LOG_LANES = 3;
using Coef = uint16_t;
using Val = int8_t;
constexpr uint32_t EXT_W = 17;            // X^8 = 17
constexpr uint32_t FROB_ROOT = 57852;     // 17^((p-1)/8) (qafold_short.rs:486-525)
#endif

struct alignas(16) E8 {
  Coef c[LANES];
};
static_assert(sizeof(E8) == 16 && LANES * sizeof(Val) == 8, "16-byte elements packed from 8 BVA bytes");

__host__ __device__ inline uint32_t madd(uint32_t a, uint32_t b) {
  uint32_t s = a + b;
  uint32_t t = s - P;
  return s < t ? s : t;
}
__host__ __device__ inline uint32_t msub(uint32_t a, uint32_t b) {
  uint32_t d = a - b;
  uint32_t e = d + P;
  return d < e ? d : e;
}
// Canonical residue of a signed BVA entry.
__host__ __device__ inline uint32_t from_val(int32_t v) { return uint32_t(v < 0 ? v + int32_t(P) : v); }

#ifdef QF_KB
// Montgomery constants: PINV = -P^-1 mod 2^32, R2 = 2^64 mod P, R32 = 2^32 mod P.
constexpr uint32_t MONT_PINV = 2130706431u;
constexpr uint32_t MONT_R2 = 402124772u;
constexpr uint32_t R32 = 33554430u;
// x < P * 2^32 -> x * 2^-32 mod P (canonical).
__host__ __device__ inline uint32_t redc(uint64_t x) {
  const uint32_t m = uint32_t(x) * MONT_PINV;
  const uint64_t t = (x + uint64_t(m) * P) >> 32;  // < 2P
  const uint32_t r = uint32_t(t);
  const uint32_t w = r - P;
  return r < w ? r : w;
}
// a -> a * 2^32 mod P (Montgomery form of a canonical a)
__host__ __device__ inline uint32_t to_mont(uint32_t a) { return redc(uint64_t(a) * MONT_R2); }
// any 64-bit x -> x mod P
__host__ __device__ inline uint32_t mred64(uint64_t x) {
  const uint64_t y = (x >> 32) * R32 + (x & 0xffffffffu);  // < 2^57 + 2^32 < P * 2^32
  return redc(uint64_t(redc(y)) * MONT_R2);
}
__host__ __device__ inline uint32_t mred32(uint32_t x) {
  x = x >= 2 * P ? x - 2 * P : x;
  return x >= P ? x - P : x;
}
#else
// x < 2^32 -> x mod P (2^16 = 15 mod P)
__host__ __device__ inline uint32_t mred32(uint32_t x) {
  uint32_t y = (x & 0xffffu) + 15u * (x >> 16);  // < 2^20
  y = (y & 0xffffu) + 15u * (y >> 16);           // < 65536 + 240
  uint32_t t = y - P;
  return y < t ? y : t;
}
__host__ __device__ inline uint32_t mred64(uint64_t x) {  // x < 2^48
  uint64_t y = (x & 0xffffu) + 15u * (x >> 16);   // < 2^36
  y = (y & 0xffffu) + 15u * (y >> 16);
  return mred32(uint32_t(y));
}
#endif

// Unpacked register form of one extension element.
struct V8 {
  uint32_t c[LANES];
};
__host__ __device__ inline V8 v8_load(const E8& e) {
  V8 r;
#ifdef QF_KB
#pragma unroll
  for (int i = 0; i < 4; i++) r.c[i] = e.c[i];
#else
  const uint32_t* w = reinterpret_cast<const uint32_t*>(&e);
#pragma unroll
  for (int i = 0; i < 4; i++) { r.c[2 * i] = w[i] & 0xffffu; r.c[2 * i + 1] = w[i] >> 16; }
#endif
  return r;
}
__host__ __device__ inline E8 v8_store(const V8& v) {
  E8 e;
#ifdef QF_KB
#pragma unroll
  for (int i = 0; i < 4; i++) e.c[i] = v.c[i];
#else
  uint32_t* w = reinterpret_cast<uint32_t*>(&e);
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = v.c[2 * i] | (v.c[2 * i + 1] << 16);
#endif
  return e;
}
__host__ __device__ inline V8 v8_add(const V8& a, const V8& b) { V8 r;
#pragma unroll
  for (int i = 0; i < LANES; i++) r.c[i] = madd(a.c[i], b.c[i]); return r; }
__host__ __device__ inline V8 v8_sub(const V8& a, const V8& b) { V8 r;
#pragma unroll
  for (int i = 0; i < LANES; i++) r.c[i] = msub(a.c[i], b.c[i]); return r; }
__host__ __device__ inline V8 v8_zero() { V8 r;
#pragma unroll
  for (int i = 0; i < LANES; i++) r.c[i] = 0; return r; }
#ifdef QF_KB
// Product in KoalaBear[X]/(X^4 - 3) with b given in Montgomery form: the
// X^4 = 3 wrap is folded into the sequence E = (3 b_1, 3 b_2, 3 b_3, b_0, ..,
// b_3), so coefficient i = sum_j a_j E[3 + i - j]; each pair of products
// (< 2 P^2 < P 2^32) takes one Montgomery reduction.
__host__ __device__ inline V8 v8_mul_mont(const V8& a, const V8& bm) {
  uint32_t E[7];
#pragma unroll
  for (int j = 0; j < 4; j++) E[3 + j] = bm.c[j];
#pragma unroll
  for (int j = 1; j < 4; j++) { const uint32_t d = madd(bm.c[j], bm.c[j]); E[j - 1] = madd(d, bm.c[j]); }
  V8 r;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint64_t s0 = uint64_t(a.c[0]) * E[3 + i] + uint64_t(a.c[1]) * E[2 + i];
    const uint64_t s1 = uint64_t(a.c[2]) * E[1 + i] + uint64_t(a.c[3]) * E[i];
    r.c[i] = madd(redc(s0), redc(s1));
  }
  return r;
}
__host__ __device__ inline V8 v8_to_mont(const V8& b) { V8 r;
#pragma unroll
  for (int i = 0; i < 4; i++) r.c[i] = to_mont(b.c[i]); return r; }
__host__ __device__ inline V8 v8_mul(const V8& a, const V8& b) { return v8_mul_mont(a, v8_to_mont(b)); }
// The QA generator alpha_1 is kept on the device in an "encode form" for the
// encoding kernels: Montgomery form here, canonical for F65521.
__host__ __device__ inline V8 v8_encode_form(const V8& b) { return v8_to_mont(b); }
__host__ __device__ inline V8 v8_mul_enc(const V8& a, const V8& b_enc) { return v8_mul_mont(a, b_enc); }
#else
// v_dot2_u32_u16: c + a.lo b.lo + a.hi b.hi.
__device__ inline uint32_t dot2_u16(uint32_t a, uint32_t b, uint32_t c) {
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b), c, false);
}
// u -> (u & 0xffff) + 15 (u >> 16), congruent mod P; below 2^20.1 for any u32.
__device__ inline uint32_t fold16(uint32_t u) { return u - __umul24(u >> 16, uint32_t(P)); }
#ifdef __HIP_DEVICE_COMPILE__
// Device product in F65521[X]/(X^8 - 17). Coefficient i sums a_j E[7 + i - j]
// over the sequence E = (17 b_1, ..., 17 b_7, b_0, ..., b_7), which folds the
// X^8 = 17 wrap. a is split into 8-bit limbs; each v_dot2_u32_u16 adds two
// limb x E products, so every partial sum of eight stays below 2^27.
__device__ inline V8 v8_mul_dot2(const V8& a, const V8& b) {
  uint32_t E[15];
#pragma unroll
  for (int i = 0; i < 8; i++) E[7 + i] = b.c[i];
#pragma unroll
  for (int j = 1; j < 8; j++) {
    const uint32_t u = fold16(__umul24(b.c[j], 17u));  // < 2^16 + 255
    const uint32_t w = u - P;
    E[j - 1] = u < w ? u : w;
  }
  uint32_t D[15];  // D[m] = (E[m], E[m-1]) as a u16 pair
#pragma unroll
  for (int m = 1; m < 15; m++) D[m] = E[m] | (E[m - 1] << 16);
  uint32_t lo[4], hi[4];
#pragma unroll
  for (int jp = 0; jp < 4; jp++) {
    const uint32_t pr = a.c[2 * jp] | (a.c[2 * jp + 1] << 16);
    lo[jp] = pr & 0x00ff00ffu;
    hi[jp] = (pr >> 8) & 0x00ff00ffu;
  }
  V8 r;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint32_t sl = 0, sh = 0;
#pragma unroll
    for (int jp = 0; jp < 4; jp++) {
      sl = dot2_u16(lo[jp], D[7 + i - 2 * jp], sl);
      sh = dot2_u16(hi[jp], D[7 + i - 2 * jp], sh);
    }
    uint32_t x = fold16(sh);  // < 2^16 + 15 * 2^11
    x = (x << 8) + sl;        // < 2^25 + 2^27
    x = fold16(fold16(x));    // < 2^16 + 15
    const uint32_t w = x - P;
    r.c[i] = x < w ? x : w;
  }
  return r;
}
#endif
// Product in F65521[X]/(X^8 - 17): the dot2 form on the device, 64-bit
// schoolbook sums with one reduction per coefficient on the host.
__host__ __device__ inline V8 v8_mul(const V8& a, const V8& b) {
#ifdef __HIP_DEVICE_COMPILE__
  return v8_mul_dot2(a, b);
#endif
  uint64_t t[15];
#pragma unroll
  for (int k = 0; k < 15; k++) t[k] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) t[i + j] += uint64_t(a.c[i] * b.c[j]);
  V8 r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.c[i] = mred64(t[i] + (i < 7 ? 17ull * t[i + 8] : 0ull));
  return r;
}
__host__ __device__ inline V8 v8_encode_form(const V8& b) { return b; }
__host__ __device__ inline V8 v8_mul_enc(const V8& a, const V8& b_enc) { return v8_mul(a, b_enc); }
#endif

// ---------------------------------------------------------------- BLAKE2s-256 (RFC 7693)
__host__ __device__ inline uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

#define QF_B2S_G(a, b, c, d, x, y)        \
  v[a] = v[a] + v[b] + (x);               \
  v[d] = rotr32(v[d] ^ v[a], 16);         \
  v[c] = v[c] + v[d];                     \
  v[b] = rotr32(v[b] ^ v[c], 12);         \
  v[a] = v[a] + v[b] + (y);               \
  v[d] = rotr32(v[d] ^ v[a], 8);          \
  v[c] = v[c] + v[d];                     \
  v[b] = rotr32(v[b] ^ v[c], 7);

// Compression of one 64-byte block given as 16 little-endian words.
// `t` is the byte count including this block; `last` sets the final flag.
__host__ __device__ inline void b2s_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, bool last) {
  constexpr uint32_t IV0 = 0x6A09E667, IV1 = 0xBB67AE85, IV2 = 0x3C6EF372, IV3 = 0xA54FF53A,
                     IV4 = 0x510E527F, IV5 = 0x9B05688C, IV6 = 0x1F83D9AB, IV7 = 0x5BE0CD19;
  uint32_t v[16] = {h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7],
                    IV0,  IV1,  IV2,  IV3,  IV4 ^ t, IV5, last ? ~IV6 : IV6, IV7};
  // Message schedule, RFC 7693 table SIGMA (rounds 0..9).
#define QF_B2S_ROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
  QF_B2S_G(0, 4, 8, 12, m[s0], m[s1]) QF_B2S_G(1, 5, 9, 13, m[s2], m[s3])                  \
  QF_B2S_G(2, 6, 10, 14, m[s4], m[s5]) QF_B2S_G(3, 7, 11, 15, m[s6], m[s7])               \
  QF_B2S_G(0, 5, 10, 15, m[s8], m[s9]) QF_B2S_G(1, 6, 11, 12, m[s10], m[s11])             \
  QF_B2S_G(2, 7, 8, 13, m[s12], m[s13]) QF_B2S_G(3, 4, 9, 14, m[s14], m[s15])
  QF_B2S_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
  QF_B2S_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3)
  QF_B2S_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4)
  QF_B2S_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8)
  QF_B2S_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13)
  QF_B2S_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9)
  QF_B2S_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11)
  QF_B2S_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10)
  QF_B2S_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5)
  QF_B2S_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0)
#undef QF_B2S_ROUND
#pragma unroll
  for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}

__host__ __device__ inline void b2s_init(uint32_t h[8]) {
  h[0] = 0x6A09E667u ^ 0x01010020u; h[1] = 0xBB67AE85; h[2] = 0x3C6EF372; h[3] = 0xA54FF53A;
  h[4] = 0x510E527F; h[5] = 0x9B05688C; h[6] = 0x1F83D9AB; h[7] = 0x5BE0CD19;
}

}  // namespace qf
