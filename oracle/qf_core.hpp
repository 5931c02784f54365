// CPU oracle for the QAFold prover: arithmetic and hashing core.
//
// TEST INFRASTRUCTURE ONLY. This directory is a plain CPU restatement of the
// reference prover's algorithm; it is used by tests/, by the smoke check and
// by the CPU-baseline leg of bench.py as the checker, never by the product
// library (parcomm_amd/), which must not link or call anything here.
//
// Pinning (what checks each part against something other than itself):
//   * BLAKE2s            : hashlib.blake2s on random inputs (tests/test_oracle_core.py).
//   * ChaCha12 keystream : RFC 8439 ChaCha20 block vector through the same core
//                          with 20 rounds; the 12-round setup seed is pinned only
//                          end to end through the reference's golden digests.
//   * F65521 / Ext8      : field axioms, Frobenius = x^p, inverse, X^8 = 17
//                          (tests/test_oracle_core.py, brute force in Python).
//   * KoalaBear / Ext4   : the same checks with X^4 = 3 plus the reference's
//                          Frobenius phase table (tests/test_oracle_kb.py).
// The field is chosen at compile time: the default build is F65521 / Ext8 with
// i8 entries (liboracle.so); -DORC_KB gives KoalaBear / Ext4 with i16 entries
// (liboracle_kb.so). Both builds reproduce the reference's demo proof digests.
// Anything not listed is pinned in the file that defines it.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <array>
#include <stdexcept>

namespace orc {

using u8 = uint8_t;
using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;
using Bytes = std::vector<u8>;
using Digest = std::array<u8, 32>;

struct Error : std::runtime_error {
  explicit Error(const char* m) : std::runtime_error(m) {}
};
inline void check(bool ok, const char* m) { if (!ok) throw Error(m); }

#ifdef ORC_KB
// ---------------------------------------------------------------- KoalaBear
// koalabear.rs:1-2, :85-120: p = 2^31 - 2^24 + 1, canonical u32; BVA entries are i16.
constexpr u32 P = 2130706433;
using Coef = u32;
using Val = int16_t;
constexpr int LANES = 4;                 // extension degree = packing lanes per element
constexpr int LOG_LANES = 2;
constexpr Coef FROB_ROOT = 2113994754;   // x -> x^p scales coefficient i by FROB_ROOT^i (koalabear.rs:155)
static const char* const FIELD_NAME = "FullKoalaBearExt4/X^4-3/scalar/v1";  // koalabear.rs:195
static const char* const OUTER_DOMAIN = "parcomm-koalabear-i16/public/v1";   // transcript.rs:15
static const char* const TERMINAL_DOMAIN = "parcomm/koalabear-i16/terminal/v2";  // protocol.rs:330
static const char* const PACKED_DOMAIN = "mikros/qafold-koalabear-ext4/packed-boolean-mle/v1";  // packed_koalabear.rs:478
static const char* const PACKED_MAGIC = "QAKOAPK1";                            // packed_koalabear.rs:24
#else
// ---------------------------------------------------------------- F65521
// qafold_short.rs:325-477 (Prime65521): p = 2^16 - 15, canonical u16; BVA entries are i8.
constexpr u32 P = 65521;
using Coef = u16;
using Val = int8_t;
constexpr int LANES = 8;
constexpr int LOG_LANES = 3;
constexpr Coef FROB_ROOT = 57852;        // 17^((p-1)/8) (qafold_short.rs:486-525)
static const char* const FIELD_NAME = "FullPrime65521Ext8/X^8-17/v1";  // full_extension.rs:130
static const char* const OUTER_DOMAIN = "parcomm-qafold/public/v1";
static const char* const TERMINAL_DOMAIN = "parcomm/qafold/terminal/v2";
static const char* const PACKED_DOMAIN = "mikros/qafold-f65521-ext8/packed-boolean-mle/v1";
static const char* const PACKED_MAGIC = "QA655PK1";
#endif
inline Coef fadd(Coef a, Coef b) { u32 s = u32(a) + b; return Coef(s >= P ? s - P : s); }
inline Coef fsub(Coef a, Coef b) { return Coef(a >= b ? u32(a) - b : u32(a) + P - b); }
inline Coef fneg(Coef a) { return Coef(a == 0 ? 0 : P - a); }
inline Coef fmul(Coef a, Coef b) { return Coef((u64(a) * b) % P); }
inline Coef fpow(Coef a, u64 e) {
  Coef r = 1;
  while (e) { if (e & 1) r = fmul(r, a); a = fmul(a, a); e >>= 1; }
  return r;
}
inline Coef finv(Coef a) { check(a != 0, "zero inverse"); return fpow(a, P - 2); }
inline Coef fsmall(long v) { long m = v % long(P); return Coef(m < 0 ? m + long(P) : m); }

// ---------------------------------------------------------------- extension field
// F65521[X]/(X^8 - 17) (qafold_short.rs:482-660) or KoalaBear[X]/(X^4 - 3)
// (koalabear.rs:1-2, 131-145), coefficient i = X^i, 16 canonical bytes.
// The type keeps the name E8 in both builds.
struct E8 {
  Coef c[LANES];
  bool operator==(const E8& o) const { return std::memcmp(c, o.c, sizeof c) == 0; }
  bool operator!=(const E8& o) const { return !(*this == o); }
};
static_assert(sizeof(E8) == 16, "16-byte field elements");
inline E8 e8_zero() { E8 r; std::memset(r.c, 0, sizeof r.c); return r; }
inline E8 e8_one() { E8 r = e8_zero(); r.c[0] = 1; return r; }
inline E8 e8_base(Coef v) { E8 r = e8_zero(); r.c[0] = v; return r; }
inline E8 e8_from_u64(u64 v) { return e8_base(Coef(v % P)); }
inline bool e8_is_zero(const E8& a) { for (int i = 0; i < LANES; i++) if (a.c[i]) return false; return true; }
inline E8 operator+(const E8& a, const E8& b) { E8 r; for (int i = 0; i < LANES; i++) r.c[i] = fadd(a.c[i], b.c[i]); return r; }
inline E8 operator-(const E8& a, const E8& b) { E8 r; for (int i = 0; i < LANES; i++) r.c[i] = fsub(a.c[i], b.c[i]); return r; }
inline E8 operator-(const E8& a) { E8 r; for (int i = 0; i < LANES; i++) r.c[i] = fneg(a.c[i]); return r; }
inline E8& operator+=(E8& a, const E8& b) { a = a + b; return a; }
inline E8& operator-=(E8& a, const E8& b) { a = a - b; return a; }
// Schoolbook product with the X^LANES = W fold (W = 17 or 3); any exact product
// gives the same canonical coefficients as the reference's Karatsuba / tower.
inline E8 operator*(const E8& a, const E8& b) {
#ifdef ORC_KB
  constexpr u64 W = 3;
#else
  constexpr u64 W = 17;
#endif
  u64 t[2 * LANES - 1] = {0};
  for (int i = 0; i < LANES; i++)
    for (int j = 0; j < LANES; j++) t[i + j] = (t[i + j] + u64(a.c[i]) * b.c[j]) % P;
  E8 r;
  for (int i = 0; i < LANES; i++) r.c[i] = Coef((t[i] + W * (i < LANES - 1 ? t[i + LANES] : 0)) % P);
  return r;
}
inline E8& operator*=(E8& a, const E8& b) { a = a * b; return a; }
inline E8 e8_mul_base(const E8& a, Coef s) { E8 r; for (int i = 0; i < LANES; i++) r.c[i] = fmul(a.c[i], s); return r; }
// Frobenius x -> x^(p^k): coefficient i scales by FROB_ROOT^(i k).
inline Coef frob_root_pow(int k) { return fpow(FROB_ROOT, u64(k % LANES)); }
inline E8 e8_frob(const E8& a, int power) {
  power = ((power % LANES) + LANES) % LANES;
  if (power == 0) return a;
  E8 r;
  for (int i = 0; i < LANES; i++) r.c[i] = fmul(a.c[i], frob_root_pow((i * power) % LANES));
  return r;
}
inline E8 e8_inv(const E8& a) {  // a^-1 = (product of the other conjugates) / norm
  check(!e8_is_zero(a), "zero inverse");
  E8 prod = e8_frob(a, 1);
  for (int k = 2; k < LANES; k++) prod = e8_frob(prod * a, 1);
  E8 norm = a * prod;
  return e8_mul_base(prod, finv(norm.c[0]));
}
inline void e8_write(const E8& a, u8* out) {
  for (int i = 0; i < LANES; i++)
    for (size_t b = 0; b < sizeof(Coef); b++) out[sizeof(Coef) * i + b] = u8(u32(a.c[i]) >> (8 * b));
}
inline bool e8_read(const u8* in, E8& a) {  // canonical bytes; rejects coefficients >= p
  for (int i = 0; i < LANES; i++) {
    u32 v = 0;
    for (size_t b = 0; b < sizeof(Coef); b++) v |= u32(in[sizeof(Coef) * i + b]) << (8 * b);
    if (v >= P) return false;
    a.c[i] = Coef(v);
  }
  return true;
}
// from_uniform_bytes: F65521 reads canonical u16s; KoalaBear first clears bit 7
// of bytes 3, 7, 11, 15 (koalabear.rs:248-255). Both reject values >= p.
inline bool e8_from_uniform(const u8* in, E8& a) {
#ifdef ORC_KB
  u8 b[16];
  std::memcpy(b, in, 16);
  for (int i : {3, 7, 11, 15}) b[i] &= 0x7f;
  return e8_read(b, a);
#else
  return e8_read(in, a);
#endif
}
inline std::array<u8, 16> e8_bytes(const E8& a) { std::array<u8, 16> b; e8_write(a, b.data()); return b; }

// ---------------------------------------------------------------- BLAKE2s-256
// RFC 7693, unkeyed, 32-byte output (the `blake2` and `blake2s_simd` crates).
struct Blake2s {
  u32 h[8];
  u8 buf[64];
  size_t buflen = 0;
  u64 t = 0;
  static constexpr u32 IV[8] = {0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
                                0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19};
  Blake2s() { for (int i = 0; i < 8; i++) h[i] = IV[i]; h[0] ^= 0x01010000u ^ 32u; }
  static u32 rotr(u32 x, int n) { return (x >> n) | (x << (32 - n)); }
  void compress(const u8* block, bool last) {
    static const u8 S[10][16] = {
        {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
        {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
        {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
        {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
        {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};
    u32 m[16], v[16];
    for (int i = 0; i < 16; i++) m[i] = block[4 * i] | (u32(block[4 * i + 1]) << 8) | (u32(block[4 * i + 2]) << 16) | (u32(block[4 * i + 3]) << 24);
    for (int i = 0; i < 8; i++) { v[i] = h[i]; v[i + 8] = IV[i]; }
    v[12] ^= u32(t); v[13] ^= u32(t >> 32);
    if (last) v[14] = ~v[14];
    auto G = [&](int a, int b, int c, int d, u32 x, u32 y) {
      v[a] = v[a] + v[b] + x; v[d] = rotr(v[d] ^ v[a], 16);
      v[c] = v[c] + v[d];     v[b] = rotr(v[b] ^ v[c], 12);
      v[a] = v[a] + v[b] + y; v[d] = rotr(v[d] ^ v[a], 8);
      v[c] = v[c] + v[d];     v[b] = rotr(v[b] ^ v[c], 7);
    };
    for (int r = 0; r < 10; r++) {
      const u8* s = S[r];
      G(0, 4, 8, 12, m[s[0]], m[s[1]]); G(1, 5, 9, 13, m[s[2]], m[s[3]]);
      G(2, 6, 10, 14, m[s[4]], m[s[5]]); G(3, 7, 11, 15, m[s[6]], m[s[7]]);
      G(0, 5, 10, 15, m[s[8]], m[s[9]]); G(1, 6, 11, 12, m[s[10]], m[s[11]]);
      G(2, 7, 8, 13, m[s[12]], m[s[13]]); G(3, 4, 9, 14, m[s[14]], m[s[15]]);
    }
    for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
  }
  void update(const void* data, size_t len) {
    const u8* p = static_cast<const u8*>(data);
    while (len > 0) {
      if (buflen == 64) { t += 64; compress(buf, false); buflen = 0; }
      size_t take = 64 - buflen; if (take > len) take = len;
      std::memcpy(buf + buflen, p, take); buflen += take; p += take; len -= take;
    }
  }
  void update(const Bytes& b) { update(b.data(), b.size()); }
  void update_u64(u64 v) { u8 b[8]; for (int i = 0; i < 8; i++) b[i] = u8(v >> (8 * i)); update(b, 8); }
  Digest finalize() const {
    Blake2s c = *this;
    c.t += c.buflen;
    std::memset(c.buf + c.buflen, 0, 64 - c.buflen);
    c.compress(c.buf, true);
    Digest out;
    for (int i = 0; i < 8; i++) for (int j = 0; j < 4; j++) out[4 * i + j] = u8(c.h[i] >> (8 * j));
    return out;
  }
  // u64le(len(label)) || label || u64le(len(bytes)) || bytes  (qafold_short_pc.rs:2186-2191)
  void frame(const char* label, const void* bytes, size_t n) {
    size_t l = std::strlen(label);
    update_u64(l); update(label, l); update_u64(n); update(bytes, n);
  }
  void frame(const char* label, const std::string& s) { frame(label, s.data(), s.size()); }
  void frame_u64(const char* label, u64 v) { u8 b[8]; for (int i = 0; i < 8; i++) b[i] = u8(v >> (8 * i)); frame(label, b, 8); }
};
inline Digest blake2s(const void* data, size_t len) { Blake2s h; h.update(data, len); return h.finalize(); }

// ---------------------------------------------------------------- ChaCha (rand_chacha 0.3 word order)
// Block function with a 64-bit block counter in words 12-13 and a 64-bit stream
// id in words 14-15. StdRng (rand 0.8) is the 12-round variant; fill_bytes
// emits the keystream words little-endian in order.
inline void chacha_block(const u8 key[32], u64 counter, u64 stream, int rounds, u32 out[16]) {
  u32 s[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};
  for (int i = 0; i < 8; i++) s[4 + i] = key[4 * i] | (u32(key[4 * i + 1]) << 8) | (u32(key[4 * i + 2]) << 16) | (u32(key[4 * i + 3]) << 24);
  s[12] = u32(counter); s[13] = u32(counter >> 32); s[14] = u32(stream); s[15] = u32(stream >> 32);
  u32 x[16]; std::memcpy(x, s, sizeof x);
  auto rotl = [](u32 v, int n) { return (v << n) | (v >> (32 - n)); };
  auto QR = [&](int a, int b, int c, int d) {
    x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16); x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
    x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);  x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
  };
  for (int r = 0; r < rounds; r += 2) {
    QR(0, 4, 8, 12); QR(1, 5, 9, 13); QR(2, 6, 10, 14); QR(3, 7, 11, 15);
    QR(0, 5, 10, 15); QR(1, 6, 11, 12); QR(2, 7, 8, 13); QR(3, 4, 9, 14);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}
// Sequential keystream reader (StdRng::from_seed + fill_bytes / next_u32 / next_u64).
struct ChaChaRng {
  u8 key[32];
  u64 block = 0;
  u32 words[16];
  int idx = 16;
  int rounds;
  ChaChaRng(const u8 k[32], int r = 12) : rounds(r) { std::memcpy(key, k, 32); }
  u32 next_u32() {
    if (idx == 16) { chacha_block(key, block++, 0, rounds, words); idx = 0; }
    return words[idx++];
  }
  u64 next_u64() { u64 lo = next_u32(); u64 hi = next_u32(); return lo | (hi << 32); }
  void fill_bytes(u8* out, size_t n) {
    for (size_t i = 0; i < n; i += 4) { u32 w = next_u32(); for (size_t j = 0; j < 4 && i + j < n; j++) out[i + j] = u8(w >> (8 * j)); }
  }
};

// rand_core 0.6 SeedableRng::seed_from_u64: PCG32 output words fill the 32-byte seed.
inline void seed_from_u64(u64 state, u8 seed[32]) {
  const u64 MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
  for (int w = 0; w < 8; w++) {
    state = state * MUL + INC;
    u32 xs = u32(((state >> 18) ^ state) >> 27);
    u32 rot = u32(state >> 59);
    u32 x = (xs >> rot) | (xs << ((32 - rot) & 31));
    for (int j = 0; j < 4; j++) seed[4 * w + j] = u8(x >> (8 * j));
  }
}

inline void put_u64(Bytes& out, u64 v) { for (int i = 0; i < 8; i++) out.push_back(u8(v >> (8 * i))); }
inline void put_e8(Bytes& out, const E8& v) { u8 b[16]; e8_write(v, b); out.insert(out.end(), b, b + 16); }
inline int ilog2(u64 x) { int r = 0; while ((u64(1) << (r + 1)) <= x && r < 63) r++; return r; }
inline bool is_pow2(u64 x) { return x && !(x & (x - 1)); }

}  // namespace orc
