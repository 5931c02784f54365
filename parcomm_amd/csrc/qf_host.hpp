// Host-side helpers of the product library: streaming BLAKE2s with the
// reference's frame encoding, extension-field host arithmetic and
// little-endian writers.
#pragma once
#include "qf_prims.hpp"
#include "../../include/parcomm_amd.h"
#include <array>
#include <stdexcept>
#include <string>
#include <vector>

namespace qf {

using Bytes = std::vector<uint8_t>;
using Digest = std::array<uint8_t, 32>;

// Field-specific names and domain separators of the reference crates.
#ifdef QF_KB
constexpr int QF_FIELD = QF_FIELD_KOALABEAR;
constexpr const char* FIELD_NAME = "FullKoalaBearExt4/X^4-3/scalar/v1";                // koalabear.rs:195
constexpr const char* OUTER_DOMAIN = "parcomm-koalabear-i16/public/v1";                // transcript.rs:15
constexpr const char* TERMINAL_DOMAIN = "parcomm/koalabear-i16/terminal/v2";           // protocol.rs:330
constexpr const char* PACKED_DOMAIN = "mikros/qafold-koalabear-ext4/packed-boolean-mle/v1";  // packed_koalabear.rs:478
constexpr const char* PACKED_MAGIC = "QAKOAPK1";                                       // packed_koalabear.rs:25
#else
constexpr int QF_FIELD = QF_FIELD_F65521;
constexpr const char* FIELD_NAME = "FullPrime65521Ext8/X^8-17/v1";                     // full_extension.rs:130
constexpr const char* OUTER_DOMAIN = "parcomm-qafold/public/v1";                       // transcript.rs:15
constexpr const char* TERMINAL_DOMAIN = "parcomm/qafold/terminal/v2";                  // protocol.rs:323
constexpr const char* PACKED_DOMAIN = "mikros/qafold-f65521-ext8/packed-boolean-mle/v1";  // packed.rs:473
constexpr const char* PACKED_MAGIC = "QA655PK1";                                       // packed.rs:19
#endif

// Status codes: the reference engine's error kinds (vendor/qafold/src/lib.rs:20-35)
// plus device and argument failures, as declared in the public header.
struct QfError : std::runtime_error {
  int status;
  QfError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};
inline void require(bool ok, int status, const char* msg) { if (!ok) throw QfError(status, msg); }

struct Hasher {
  uint32_t h[8];
  uint8_t buf[64];
  uint32_t buflen = 0;
  uint64_t total = 0;
  Hasher() { b2s_init(h); }
  void block(bool last, uint32_t t) {
    uint32_t m[16];
    for (int i = 0; i < 16; i++) m[i] = uint32_t(buf[4 * i]) | uint32_t(buf[4 * i + 1]) << 8 | uint32_t(buf[4 * i + 2]) << 16 | uint32_t(buf[4 * i + 3]) << 24;
    b2s_compress(h, m, t, last);
  }
  void update(const void* data, size_t n) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    while (n) {
      if (buflen == 64) { total += 64; block(false, uint32_t(total)); buflen = 0; }
      size_t k = 64 - buflen < n ? 64 - buflen : n;
      std::memcpy(buf + buflen, p, k);
      buflen += uint32_t(k); p += k; n -= k;
    }
  }
  void u64le(uint64_t v) { uint8_t b[8]; for (int i = 0; i < 8; i++) b[i] = uint8_t(v >> (8 * i)); update(b, 8); }
  void frame(const char* label, const void* bytes, size_t n) { size_t l = std::strlen(label); u64le(l); update(label, l); u64le(n); update(bytes, n); }
  void frame(const char* label, const char* text) { frame(label, text, std::strlen(text)); }
  void frame_u64(const char* label, uint64_t v) { uint8_t b[8]; for (int i = 0; i < 8; i++) b[i] = uint8_t(v >> (8 * i)); frame(label, b, 8); }
  Digest finish() const {
    Hasher c = *this;
    std::memset(c.buf + c.buflen, 0, 64 - c.buflen);
    c.block(true, uint32_t(c.total + c.buflen));
    Digest d;
    for (int i = 0; i < 8; i++) for (int j = 0; j < 4; j++) d[4 * i + j] = uint8_t(c.h[i] >> (8 * j));
    return d;
  }
};
inline Digest blake2s(const void* p, size_t n) { Hasher h; h.update(p, n); return h.finish(); }

// Host extension arithmetic on the canonical 16-byte form.
inline E8 e8(uint32_t c0) { E8 e{}; e.c[0] = Coef(c0 % P); return e; }
inline E8 e8_zero() { return E8{}; }
inline E8 e8_one() { return e8(1); }
inline bool e8_eq(const E8& a, const E8& b) { return std::memcmp(&a, &b, 16) == 0; }
inline bool e8_is_zero(const E8& a) { return e8_eq(a, E8{}); }
inline E8 e8_add(const E8& a, const E8& b) { return v8_store(v8_add(v8_load(a), v8_load(b))); }
inline E8 e8_sub(const E8& a, const E8& b) { return v8_store(v8_sub(v8_load(a), v8_load(b))); }
inline E8 e8_mul(const E8& a, const E8& b) { return v8_store(v8_mul(v8_load(a), v8_load(b))); }
inline E8 e8_neg(const E8& a) { return e8_sub(E8{}, a); }
inline uint32_t fp_mul(uint32_t a, uint32_t b) { return uint32_t((uint64_t(a) * b) % P); }
inline uint32_t fp_pow(uint32_t a, uint64_t e) { uint32_t r = 1; while (e) { if (e & 1) r = fp_mul(r, a); a = fp_mul(a, a); e >>= 1; } return r; }
inline uint32_t fp_inv(uint32_t a) { require(a != 0, QF_ERR_ENCODING, "zero inverse"); return fp_pow(a, P - 2); }
inline E8 e8_scale(const E8& a, uint32_t s) { E8 r; for (int i = 0; i < LANES; i++) r.c[i] = Coef(fp_mul(a.c[i], s)); return r; }
// x -> x^(p^k): coefficient i times FROB_ROOT^(i k mod LANES) (qafold_short.rs:486-525,
// koalabear.rs:153-163)
// frob_table()[k][i] = FROB_ROOT^((i k) mod LANES): the k-th Frobenius power scales coefficient i by it.
inline const std::array<std::array<uint32_t, LANES>, LANES>& frob_table() {
  static const auto t = [] {
    std::array<std::array<uint32_t, LANES>, LANES> m{};
    for (int k = 0; k < LANES; k++)
      for (int i = 0; i < LANES; i++) m[k][i] = uint32_t(fp_pow(FROB_ROOT, uint64_t((i * k) % LANES)));
    return m;
  }();
  return t;
}
inline E8 e8_frob(const E8& a, int k) {
  k = ((k % LANES) + LANES) % LANES;
  const auto& m = frob_table()[k];
  E8 r;
  for (int i = 0; i < LANES; i++) r.c[i] = Coef(fp_mul(a.c[i], m[i]));
  return r;
}
inline E8 e8_inv(const E8& a) {  // (product of the other conjugates) / norm
  require(!e8_is_zero(a), QF_ERR_ENCODING, "zero inverse");
  E8 prod = e8_frob(a, 1);
  for (int k = 2; k < LANES; k++) prod = e8_frob(e8_mul(prod, a), 1);
  E8 norm = e8_mul(a, prod);
  return e8_scale(prod, fp_inv(norm.c[0]));
}
inline E8 operator+(const E8& a, const E8& b) { return e8_add(a, b); }
inline E8 operator-(const E8& a, const E8& b) { return e8_sub(a, b); }
inline E8 operator-(const E8& a) { return e8_neg(a); }
inline E8 operator*(const E8& a, const E8& b) { return e8_mul(a, b); }
inline E8& operator+=(E8& a, const E8& b) { a = e8_add(a, b); return a; }
inline bool operator==(const E8& a, const E8& b) { return e8_eq(a, b); }
inline bool operator!=(const E8& a, const E8& b) { return !e8_eq(a, b); }
inline E8 e8_basis(int j) { E8 b{}; b.c[j] = 1; return b; }
// Canonical 16-byte encoding; rejects coefficients >= p.
inline bool e8_from_bytes(const uint8_t* b, E8& out) {
  for (int i = 0; i < LANES; i++) {
    uint32_t v = 0;
    for (size_t k = 0; k < sizeof(Coef); k++) v |= uint32_t(b[sizeof(Coef) * i + k]) << (8 * k);
    if (v >= P) return false;
    out.c[i] = Coef(v);
  }
  return true;
}
// from_uniform_bytes of the extension: F65521 reads the canonical form;
// KoalaBear first clears bit 7 of bytes 3, 7, 11 and 15 (koalabear.rs:248-255).
inline bool e8_from_uniform(const uint8_t* b, E8& out) {
#ifdef QF_KB
  uint8_t t[16];
  std::memcpy(t, b, 16);
  for (int i : {3, 7, 11, 15}) t[i] &= 0x7f;
  return e8_from_bytes(t, out);
#else
  return e8_from_bytes(b, out);
#endif
}
inline void put_u64(Bytes& o, uint64_t v) { for (int i = 0; i < 8; i++) o.push_back(uint8_t(v >> (8 * i))); }
// Overwrites the little-endian u64 at o[at, at + 8) (a length written as a placeholder).
inline void patch_u64(Bytes& o, size_t at, uint64_t v) { for (int i = 0; i < 8; i++) o[at + i] = uint8_t(v >> (8 * i)); }
inline void put_e8(Bytes& o, const E8& v) { const uint8_t* p = reinterpret_cast<const uint8_t*>(&v); o.insert(o.end(), p, p + 16); }
inline int ilog2u(uint64_t x) { int r = -1; while (x) { x >>= 1; r++; } return r; }
inline bool pow2(uint64_t x) { return x && !(x & (x - 1)); }

}  // namespace qf
