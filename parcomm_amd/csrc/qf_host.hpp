// Host-side helpers of the product library: streaming BLAKE2s with the
// reference's frame encoding, Ext8 host arithmetic and little-endian writers.
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

// Host Ext8 arithmetic on the canonical 16-byte form.
inline E8 e8(uint32_t c0) { E8 e{}; e.c[0] = uint16_t(c0 % P); return e; }
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
inline E8 e8_scale(const E8& a, uint32_t s) { E8 r; for (int i = 0; i < 8; i++) r.c[i] = uint16_t(fp_mul(a.c[i], s)); return r; }
// x -> x^(p^k): coefficient i times 57852^(i k mod 8) (qafold_short.rs:486-525)
inline E8 e8_frob(const E8& a, int k) {
  k = ((k % 8) + 8) % 8;
  E8 r;
  for (int i = 0; i < 8; i++) r.c[i] = uint16_t(fp_mul(a.c[i], fp_pow(57852, uint64_t((i * k) % 8))));
  return r;
}
inline E8 e8_inv(const E8& a) {
  require(!e8_is_zero(a), QF_ERR_ENCODING, "zero inverse");
  E8 prod = e8_frob(a, 1);
  for (int k = 2; k < 8; k++) prod = e8_frob(e8_mul(prod, a), 1);
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
inline bool e8_from_bytes(const uint8_t* b, E8& out) {
  for (int i = 0; i < 8; i++) { uint32_t v = b[2 * i] | uint32_t(b[2 * i + 1]) << 8; if (v >= P) return false; out.c[i] = uint16_t(v); }
  return true;
}
inline void put_u64(Bytes& o, uint64_t v) { for (int i = 0; i < 8; i++) o.push_back(uint8_t(v >> (8 * i))); }
inline void put_e8(Bytes& o, const E8& v) { const uint8_t* p = reinterpret_cast<const uint8_t*>(&v); o.insert(o.end(), p, p + 16); }
inline int ilog2u(uint64_t x) { int r = -1; while (x) { x >>= 1; r++; } return r; }
inline bool pow2(uint64_t x) { return x && !(x & (x - 1)); }

}  // namespace qf
