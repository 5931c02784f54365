// Host side of the QAFold protocol shared by the prover and the verifier:
// serde_json encodings of the public statement, the three Fiat-Shamir
// transcripts, query counts and row sampling, the TBV query compiler and the
// packed-opening algebra. Everything here is deterministic host code.
#pragma once
#include "qf_state.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <array>
#include <map>
#include <mutex>
#include <unordered_map>

namespace qf {

// ================================================================ statement JSON (serde_json compact)
inline void js_bytes(std::string& s, const uint8_t* b, size_t n) {
  char buf[4 * 64 + 2];  // "[", then up to 3 digits and a comma per byte, "]"
  for (size_t at = 0; at < n; at += 64) {
    size_t k = 0;
    if (at == 0) buf[k++] = '[';
    for (size_t i = at; i < n && i < at + 64; i++) {
      if (i) buf[k++] = ',';
      const unsigned v = b[i];
      if (v >= 100) buf[k++] = char('0' + v / 100);
      if (v >= 10) buf[k++] = char('0' + v / 10 % 10);
      buf[k++] = char('0' + v % 10);
    }
    if (at + 64 >= n) buf[k++] = ']';
    s.append(buf, k);
  }
  if (n == 0) s += "[]";
}
inline void js_e8(std::string& s, const E8& v) { js_bytes(s, reinterpret_cast<const uint8_t*>(&v), 16); }
inline std::string js_scalars(const std::vector<E8>& v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); i++) { if (i) s += ','; js_e8(s, v[i]); }
  return s + "]";
}

// ================================================================ transcripts
struct Outer {  // transcript.rs:8-70
  Digest state;
  uint64_t counter = 0;
  Outer() { state = blake2s(OUTER_DOMAIN, std::strlen(OUTER_DOMAIN)); }
  void bytes(const char* label, const void* d, size_t n) {
    Hasher h;
    h.update(state.data(), 32);
    h.u64le(std::strlen(label));
    h.update(label, std::strlen(label));
    h.u64le(n);
    h.update(d, n);
    state = h.finish();
  }
  void object(const char* label, const std::string& js) { bytes(label, js.data(), js.size()); }
  void scalar(const char* label, const E8& v) { bytes(label, &v, 16); }
  E8 challenge(const char* label) {
    for (;;) {
      uint64_t ct = counter++;
      uint8_t b[8];
      for (int i = 0; i < 8; i++) b[i] = uint8_t(ct >> (8 * i));
      bytes(label, b, 8);
      E8 v;
      if (e8_from_uniform(state.data(), v)) return v;
    }
  }
  E8 nonzero(const char* label) { for (;;) { E8 v = challenge(label); if (!e8_is_zero(v)) return v; } }
  E8 fold_challenge(uint64_t packed_log) {
    E8 m1 = e8_neg(e8_one());
    for (;;) {
      E8 v = nonzero("eo-fold"), p = v;
      bool ok = true;
      for (uint64_t i = 0; i < std::max<uint64_t>(packed_log, 1); i++) { ok = ok && p != m1; p = p * p; }
      if (ok) return v;
    }
  }
};

struct Pcs {  // qafold_short_pc.rs:2011-2229
  Hasher h;
  uint64_t counter = 0;
  std::vector<E8> responses;
  Pcs() { bytes("domain", "brakefold/qafold-short-pcs/v2"); }
  void bytes(const char* label, const void* b, size_t n) { h.frame(label, b, n); }
  void bytes(const char* label, const char* s) { h.frame(label, s); }
  void u64(const char* label, uint64_t v) { h.frame_u64(label, v); }
  void digest(const char* label, const Digest& d) { bytes(label, d.data(), 32); }
  void ext(const char* label, const E8& v) { bytes(label, &v, 16); }
  void exts(const char* label, const std::vector<E8>& v) {
    bytes("vector-label", label);
    u64("vector-length", v.size());
    for (auto& x : v) ext("vector-element", x);
  }
  void respond(const E8& v) { ext("response", v); responses.push_back(v); }
  Digest derive(const char* label, uint64_t ctr, uint64_t attempt) const {
    Hasher c = h;
    c.frame("challenge-label", label);
    c.u64le(ctr);
    c.u64le(0);
    c.u64le(attempt);
    return c.finish();
  }
  Digest challenge_digest(const char* label) {
    Digest d = derive(label, counter, 0);
    counter++;
    bytes("digest-challenge-label", label);
    digest("digest-challenge", d);
    return d;
  }
  E8 challenge(const char* label, bool nonzero = false) {
    uint64_t ctr = counter;
    for (uint64_t outer = 0;; outer++) {
      E8 v;
      for (uint64_t a = outer;; a++) if (e8_from_uniform(derive(label, ctr, a).data(), v)) break;
      if (!nonzero || !e8_is_zero(v)) {
        counter++;
        bytes("field-challenge-label", label);
        ext("field-challenge", v);
        return v;
      }
    }
  }
};

struct Packed {  // packed.rs:754-779
  Hasher h;
  void bytes(const void* b, size_t n) { h.u64le(n); h.update(b, n); }
  void field(const E8& v) { bytes(&v, 16); }
  E8 challenge() {
    for (uint64_t ctr = 0;; ctr++) {
      Hasher c = h;
      c.update("challenge", 9);
      c.u64le(ctr);
      Digest d = c.finish();
      E8 v;
      if (e8_from_uniform(d.data(), v)) { bytes(d.data(), 32); return v; }
    }
  }
};

// ================================================================ small host helpers
// MSB-first monomial weights: w[i] = prod over set bits b of p[len-1-b] (qafold_short_pc.rs:2846-2873)
inline std::vector<E8> tw_host(const E8* p, size_t m) {
  std::vector<E8> w(size_t(1) << m, E8{});
  w[0] = e8_one();
  size_t filled = 1;
  for (size_t k = m; k-- > 0;) {
    for (size_t i = 0; i < filled; i++) w[filled + i] = w[i] * p[k];
    filled *= 2;
  }
  return w;
}
inline std::vector<E8> tw_host(const std::vector<E8>& p) { return tw_host(p.data(), p.size()); }
// MSB-first equality table (packed.rs:525-545)
inline std::vector<E8> eq_host(const E8* p, size_t m) {
  std::vector<E8> t(1, e8_one());
  for (size_t k = 0; k < m; k++) {
    std::vector<E8> nx(t.size() * 2);
    for (size_t i = 0; i < t.size(); i++) { nx[2 * i + 1] = t[i] * p[k]; nx[2 * i] = t[i] - nx[2 * i + 1]; }
    t.swap(nx);
  }
  return t;
}
inline E8 inner_host(const std::vector<E8>& a, const std::vector<E8>& b) {
  require(a.size() == b.size(), QF_ERR_INVALID_VARIABLES, "inner product length");
  E8 s{};
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}
inline E8 tensor_inner(const std::vector<E8>& a, const std::vector<E8>& b) {
  E8 r = e8_one();
  for (size_t i = 0; i < a.size(); i++) r = r * (e8_one() + a[i] * b[i]);
  return r;
}
inline E8 weight_at(const std::vector<E8>& p, uint64_t pos) {
  E8 r = e8_one();
  for (size_t b = 0; b < p.size(); b++) if ((pos >> b) & 1) r = r * p[p.size() - b - 1];
  return r;
}
// Device tensor weights / equality table of an MSB-first point (tensor of two halves).

inline void put_part(Bytes& w, const Bytes& part) { put_u64(w, part.size()); w.insert(w.end(), part.begin(), part.end()); }
inline void put_tag(Bytes& w, const char* tag) { w.insert(w.end(), tag, tag + std::strlen(tag)); }
inline void put_digest(Bytes& w, const Digest& d) { w.insert(w.end(), d.begin(), d.end()); }

// ================================================================ query counts (qafold.rs:75-178, brakefold.rs:4-53)
inline bool log1m_exp(double x, double& out) {
  if (x >= 0.0) return false;
  out = x < -M_LN2 ? std::log1p(-std::exp(x)) : std::log(-std::expm1(x));
  return true;
}
inline double log_add_exp(double a, double b) { double m = std::max(a, b); return m + std::log(std::exp(a - m) + std::exp(b - m)); }
inline bool epsilon_at(uint64_t n, uint64_t index, double d, double ln_p, double& eps) {
  if (n == 0 || index <= 1 || !(d > 0.0 && d < 1.0) || !(ln_p > 0.0)) return false;
  double lpm1 = ln_p + std::log1p(-std::exp(-ln_p));
  double g = 1.0 - d * lpm1 / ln_p + d * std::log(d) / ln_p + (1.0 - d) * std::log(1.0 - d) / ln_p;
  eps = g - (1.0 + std::log(double(n)) / ln_p) / double(index);
  return eps > 0.0;
}
inline bool bound_at(uint64_t n, uint64_t index, double d, double ln_p, double& logf) {
  double eps;
  if (!epsilon_at(n, index, d, ln_p, eps)) return false;
  double lpm1 = ln_p + std::log1p(-std::exp(-ln_p)), lome;
  if (!log1m_exp(-double(index) * eps * ln_p, lome)) return false;
  double den = lome + lpm1;
  double fcl = std::log(double(index) * double(index - 1) * double(n) / 2.0) - 2.0 * ln_p;
  double ftl = -std::ceil(double(index - 1) / (double(index) * d)) * double(index) * eps * ln_p - den;
  double scl = std::log(double(index) * double(n)) - ln_p;
  double stl = -std::ceil(1.0 / d) * double(index) * eps * ln_p - den;
  logf = std::min(log_add_exp(fcl, ftl), log_add_exp(scl, stl));
  return true;
}
inline uint64_t num_queries_uncached(const QaCode& code, uint64_t bits, uint64_t rows) {
  unsigned __int128 card = 1;
  for (int i = 0; i < LANES; i++) card *= P;
  double ln_p = std::log2(double(card)) * M_LN2;
  double upper = 1.0 - 1.0 / double(code.index), target = -double(bits) * M_LN2;
  auto valid = [&](double d) { double lf; return bound_at(code.message_len, code.index, d, ln_p, lf) && lf <= target; };
  double lower = 0.0, ub = upper;
  for (uint64_t step = 1; step <= 4096; step++) {
    double cand = upper * double(step) / 4096.0;
    if (valid(cand)) lower = cand;
    else if (cand > lower) { ub = cand; break; }
  }
  if (lower == 0.0) return rows;  // no distance bound: every row is queried
  for (int i = 0; i < 96; i++) { double mid = (lower + ub) * 0.5; if (valid(mid)) lower = mid; else ub = mid; }
  double rel = lower * (1.0 - 1e-12);
  double q = -double(bits) / std::log2(1.0 - rel / 3.0);
  require(std::isfinite(q) && q >= 0.0, QF_ERR_ENCODING, "query count");
  return std::min<uint64_t>(uint64_t(std::ceil(q)), rows);
}
// The query count depends only on (message length, index, bits, rows); the
// search above costs a few thousand transcendental calls, so every distinct
// tuple is computed once per process.
inline uint64_t num_queries(const QaCode& code, uint64_t bits, uint64_t rows) {
  static std::mutex mu;
  static std::map<std::array<uint64_t, 4>, uint64_t> memo;
  const std::array<uint64_t, 4> key{code.message_len, code.index, bits, rows};
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
  }
  const uint64_t q = num_queries_uncached(code, bits, rows);
  std::lock_guard<std::mutex> g(mu);
  memo.emplace(key, q);
  return q;
}
inline std::vector<uint64_t> sample_rows(uint64_t count, uint64_t population, const Digest& seed) {  // qafold_short_pc.rs:3000-3047
  std::vector<uint64_t> out;
  if (count == population) { for (uint64_t i = 0; i < population; i++) out.push_back(i); return out; }
  std::unordered_map<uint64_t, uint64_t> swaps;  // the partial Fisher-Yates permutation
  swaps.reserve(2 * count);
  uint64_t counter = 0;
  for (uint64_t off = 0; off < count; off++) {
    uint64_t remaining = population - off;
    unsigned __int128 range = (unsigned __int128)1 << 64, limit = range - range % remaining;
    uint64_t s;
    for (;;) {
      Hasher h;
      h.update("qafold-short-index/v1", 21);
      h.update(seed.data(), 32);
      h.u64le(counter++);
      Digest d = h.finish();
      uint64_t cand = 0;
      for (int i = 0; i < 8; i++) cand |= uint64_t(d[i]) << (8 * i);
      if ((unsigned __int128)cand < limit) { s = cand % remaining; break; }
    }
    const uint64_t last = remaining - 1;
    const auto is = swaps.find(s), il = swaps.find(last);
    const uint64_t sv = is == swaps.end() ? s : is->second, lv = il == swaps.end() ? last : il->second;
    swaps[s] = lv;
    out.push_back(sv);
  }
  return out;
}


inline Digest hash_terminal(const std::vector<E8>& v) {
  Hasher h;
  h.frame("domain", "qafold-short/terminal-vector/v1");
  h.frame_u64("length", v.size());
  for (auto& x : v) h.frame("element", &x, 16);
  return h.finish();
}

struct MClaim { E8 value; std::vector<E8> left, right; };
inline MClaim reduce_generator(uint64_t message_len, const MClaim& g) {  // qafold_short.rs:1575-1625, index two
  E8 norm = e8_one();
  std::vector<E8> merged;
  for (size_t i = 0; i < g.right.size(); i++) {
    E8 l = g.left[1 + i], r = g.right[i];
    E8 dl = e8_one() + l, dr = e8_one() + r;
    norm = norm * dl * dr;
    merged.push_back(((e8_one() - l) * e8_inv(dl)) * ((e8_one() - r) * e8_inv(dr)));
  }
  (void)message_len;
  return {g.value * e8_inv(norm), merged, {g.left[0]}};
}
struct Sel { bool all; uint64_t count; std::vector<uint64_t> rows; };
inline void absorb_sel(Pcs& tr, const Sel& s) {
  tr.bytes("selection-kind", s.all ? "all" : "partial");
  tr.u64("selection-count", s.all ? s.count : s.rows.size());
  if (!s.all) for (auto r : s.rows) tr.u64("selection-row", r);
}


// ================================================================ TBV (tbv.rs)
struct Term {
  uint64_t offset = 0;
  E8 scale{};
  std::vector<std::array<E8, 2>> f;
  uint64_t len() const { return uint64_t(1) << f.size(); }
  E8 weight(uint64_t i) const { E8 v = scale; for (size_t b = 0; b < f.size(); b++) v = v * f[b][(i >> b) & 1]; return v; }
};
inline E8 monomial(uint64_t index, const E8* pl, size_t n) {
  E8 w = e8_one();
  for (size_t b = 0; b < n; b++) if ((index >> b) & 1) w = w * pl[b];
  return w;
}
inline std::vector<std::pair<uint64_t, uint64_t>> aligned(uint64_t start, uint64_t end) {
  std::vector<std::pair<uint64_t, uint64_t>> out;
  while (start < end) {
    int bits = ilog2u(end - start);
    if (start) bits = std::min(bits, __builtin_ctzll(start));
    out.push_back({start, uint64_t(1) << bits});
    start += uint64_t(1) << bits;
  }
  return out;
}
inline Term compact(Term t) {
  while (!t.f.empty()) {
    auto pr = t.f.back();
    uint64_t fixed;
    if (e8_is_zero(pr[0])) fixed = 1; else if (e8_is_zero(pr[1])) fixed = 0; else break;
    t.scale = t.scale * pr[fixed];
    t.offset += fixed << (t.f.size() - 1);
    t.f.pop_back();
  }
  return t;
}
struct Query { std::vector<Term> terms; size_t n_bucket_terms = 0; E8 claim{}; };
inline E8 query_eval(const Query& q, const std::vector<E8>& pl) {
  E8 s{};
  for (auto& t : q.terms) {
    E8 local = t.scale;
    size_t k = t.f.size();
    for (size_t i = 0; i < k; i++) local = local * (t.f[i][0] + t.f[i][1] * pl[i]);
    s += local * monomial(t.offset >> k, pl.data() + k, pl.size() - k);
  }
  return s;
}
inline E8 eq_weight(uint64_t index, const std::vector<E8>& p) {
  E8 w = e8_one();
  for (size_t c = 0; c < p.size(); c++) w = w * (((index >> (p.size() - 1 - c)) & 1) ? p[c] : e8_one() - p[c]);
  return w;
}
inline Query compile(const Layout& l, const std::vector<ClaimIn>& claims, Outer& tr) {  // tbv.rs:187-284
  uint64_t np = 1;
  while (np < claims.size()) np <<= 1;
  std::vector<E8> batch;
  for (int i = 0; i < ilog2u(np); i++) batch.push_back(tr.challenge("batch"));
  std::vector<E8> coef(claims.size());
  Query q;
  for (size_t i = 0; i < claims.size(); i++) { coef[i] = eq_weight(i, batch); q.claim += coef[i] * claims[i].value; }
  for (auto& b : l.buckets) {
    const ClaimIn& cl = claims[b.matrix];
    Term t;
    t.offset = b.offset;
    t.scale = coef[b.matrix] * monomial(b.row_offset, cl.row.data(), cl.row.size()) * monomial(b.col_offset, cl.col.data(), cl.col.size());
    for (int i = 0; i < ilog2u(b.cols); i++) t.f.push_back({e8_one(), cl.col[i]});
    for (int i = 0; i < ilog2u(b.rows); i++) t.f.push_back({e8_one(), cl.row[i]});
    q.terms.push_back(std::move(t));
  }
  q.n_bucket_terms = q.terms.size();
  std::vector<E8> pp;
  for (int i = 0; i < ilog2u(l.logical); i++) pp.push_back(tr.challenge("padding-point"));
  E8 mix = tr.nonzero("padding-mix");
  for (auto& b : l.buckets) {
    int cb = ilog2u(b.cols), bits = ilog2u(b.len()), rb = ilog2u(b.rows);
    std::vector<std::array<uint64_t, 4>> rects;
    for (auto [s, n] : aligned(b.physical_rows, b.rows)) rects.push_back({s, n, 0, b.cols});
    for (auto [rs, rl] : aligned(0, b.physical_rows))
      for (auto [cs, cl] : aligned(b.physical_cols, b.cols)) rects.push_back({rs, rl, cs, cl});
    for (auto& r : rects) {
      Term t;
      for (int i = 0; i < bits; i++) t.f.push_back({e8_one(), pp[i]});
      for (int bit = ilog2u(r[3]); bit < cb; bit++) t.f[bit][1 - ((r[2] >> bit) & 1)] = E8{};
      for (int bit = ilog2u(r[1]); bit < rb; bit++) t.f[cb + bit][1 - ((r[0] >> bit) & 1)] = E8{};
      t.offset = b.offset;
      t.scale = mix * monomial(b.offset >> bits, pp.data() + bits, pp.size() - bits);
      q.terms.push_back(compact(std::move(t)));
    }
  }
  for (auto [off, len] : aligned(l.active, l.logical)) {
    int bits = ilog2u(len);
    Term t;
    t.offset = off;
    t.scale = mix * monomial(off >> bits, pp.data() + bits, pp.size() - bits);
    for (int i = 0; i < bits; i++) t.f.push_back({e8_one(), pp[i]});
    q.terms.push_back(std::move(t));
  }
  return q;
}
inline std::vector<E8> tbv_w(const std::array<E8, 2>* f, size_t n) {
  std::vector<E8> r(1, e8_one());
  for (size_t k = 0; k < n; k++) {
    size_t len = r.size();
    r.resize(2 * len);
    for (size_t i = 0; i < len; i++) { r[len + i] = r[i] * f[k][1]; r[i] = r[i] * f[k][0]; }
  }
  return r;
}
inline E8 small_val(int v) { return e8(from_val(v)); }

// One extension element per packing lane (the conjugate evaluations, weight images, ...).
using Lanes = std::array<E8, LANES>;
// packed.rs:502-523, packed_koalabear.rs:507-528: inverse DFT of the conjugates.
inline E8 projected_claim(const Lanes& ev, const std::vector<E8>& low) {
  E8 x1 = e8_inv(e8_basis(1));
  Lanes ib;
  ib[0] = e8_one();
  for (int i = 1; i < LANES; i++) ib[i] = ib[i - 1] * x1;
  uint32_t omega_inv = fp_inv(FROB_ROOT), inv_lanes = fp_inv(LANES), step = 1;
  E8 claim{};
  for (int j = 0; j < LANES; j++) {
    uint32_t phase = 1;
    E8 sum{};
    for (int i = 0; i < LANES; i++) { sum += e8_scale(e8_frob(ev[i], i), phase); phase = fp_mul(phase, step); }
    claim += low[j] * e8_scale(sum * ib[j], inv_lanes);
    step = fp_mul(step, omega_inv);
  }
  return claim;
}
inline Packed packed_transcript(const ParamsState& p, const Digest& root, const std::vector<E8>& point, const E8& value,
                         const Bytes& context) {
  Packed t;
  const char* dom = PACKED_DOMAIN;
  t.bytes(dom, std::strlen(dom));
  t.bytes(p.pp.parameter_digest.data(), 32);
  uint8_t ln[8];
  for (int i = 0; i < 8; i++) ln[i] = uint8_t(uint64_t(p.chunk_log) >> (8 * i));
  t.bytes(ln, 8);
  t.bytes(root.data(), 32);
  t.bytes(context.data(), context.size());
  for (auto& x : point) t.field(x);
  t.field(value);
  return t;
}
inline Lanes lane_powers(const E8& v) { Lanes o; o[0] = e8_one(); for (int i = 1; i < LANES; i++) o[i] = o[i - 1] * v; return o; }
inline E8 batch_weight(const E8* point, size_t n, const std::vector<E8>& red, const Lanes& powers) {
  E8 sum{};
  for (int i = 0; i < LANES; i++) {
    E8 w = e8_one();
    for (size_t k = 0; k < n; k++) {
      E8 s = e8_frob(point[k], (LANES - i) % LANES);
      w = w * ((e8_one() - s) * (e8_one() - red[k]) + s * red[k]);
    }
    sum += powers[i] * w;
  }
  return sum;
}

}  // namespace qf
