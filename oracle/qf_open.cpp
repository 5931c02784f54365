// CPU oracle (test infrastructure only; see qf_core.hpp): the opening path.
//
// Pinning:
//   * ParComm composition (statement/claims JSON, TBV, EO, batched packed
//     opening with block EO, fold and binding, wire formats): the reference's
//     recorded demo proof (validation/qafold8-smoke-final/run.log: 26 974 bytes,
//     BLAKE2s 821aecd9...; tests/test_oracle_demo.py). That instance has one QA
//     level.
//   * Recursive native-opening rounds, partial-row Merkle payloads, f64 query
//     counts below the row count and the multi-level direct-interleaved Decide:
//     the reference engine's wire fixture (qafold_short_pc.rs:3605-3647,
//     BLAKE2s bb4993...; tests/test_oracle_demo.py) at n = 2^16.
//   * Every proof also runs the prover's own algebraic checks (EO terminal,
//     sumcheck terminal, LiteBullet products, terminal MLE) and throws on a
//     mismatch.
//   Parity unpinned: the exact bytes of proofs whose code sequences are deeper
//   than both goldens (chunk_log >= 22), and the f64 query counts of those
//   levels, rest on the composition of pinned pieces only.
#include "qf_open.hpp"
#include <cmath>
#include <map>
#include <memory>

namespace orc {

// ================================================================ JSON (serde_json compact)
static void json_bytes(std::string& s, const u8* b, size_t n) {
  s += '[';
  for (size_t i = 0; i < n; i++) { if (i) s += ','; s += std::to_string(unsigned(b[i])); }
  s += ']';
}
static void json_e8(std::string& s, const E8& v) { u8 b[16]; e8_write(v, b); json_bytes(s, b, 16); }
std::string json_scalars(const std::vector<E8>& v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); i++) { if (i) s += ','; json_e8(s, v[i]); }
  return s + "]";
}
std::string json_statement(const Statement& st) {  // protocol.rs:33-40, field order of the struct
  std::string s = "{\"version\":" + std::to_string(st.version) + ",\"config\":{\"chunk_log\":" +
                  std::to_string(st.config.chunk_log) + ",\"packed_log\":" + std::to_string(st.config.packed_log) +
                  ",\"query_rf_bits\":" + std::to_string(st.config.query_rf_bits) + "},\"shapes\":[";
  for (size_t i = 0; i < st.shapes.size(); i++) {
    if (i) s += ',';
    s += "{\"rows\":" + std::to_string(st.shapes[i].rows) + ",\"cols\":" + std::to_string(st.shapes[i].cols) + "}";
  }
  s += "],\"parameter_digest\":";
  json_bytes(s, st.parameter_digest.data(), 32);
  s += ",\"roots\":[";
  for (size_t i = 0; i < st.roots.size(); i++) { if (i) s += ','; json_bytes(s, st.roots[i].data(), 32); }
  return s + "]}";
}
std::string json_claims(const std::vector<Claim>& c) {  // tbv.rs:6-11
  std::string s = "[";
  for (size_t i = 0; i < c.size(); i++) {
    if (i) s += ',';
    s += "{\"row\":" + json_scalars(c[i].row) + ",\"col\":" + json_scalars(c[i].col) + ",\"value\":";
    json_e8(s, c[i].value);
    s += "}";
  }
  return s + "]";
}

// ================================================================ outer transcript (transcript.rs:8-70)
OuterTranscript::OuterTranscript() { state = blake2s(OUTER_DOMAIN, std::strlen(OUTER_DOMAIN)); }
void OuterTranscript::bytes(const char* label, const void* data, size_t n) {
  Blake2s h;
  h.update(state.data(), 32);
  size_t l = std::strlen(label);
  h.update_u64(l);
  h.update(label, l);
  h.update_u64(n);
  h.update(data, n);
  state = h.finalize();
}
E8 OuterTranscript::challenge(const char* label) {
  for (;;) {
    u64 count = counter++;
    u8 b[8];
    for (int i = 0; i < 8; i++) b[i] = u8(count >> (8 * i));
    bytes(label, b, 8);
    E8 v;
    if (e8_from_uniform(state.data(), v)) return v;
  }
}
E8 OuterTranscript::nonzero(const char* label) {
  for (;;) { E8 v = challenge(label); if (!e8_is_zero(v)) return v; }
}
E8 OuterTranscript::fold_challenge(u64 packed_log) {
  const E8 minus_one = -e8_one();
  for (;;) {
    E8 v = nonzero("eo-fold");
    E8 p = v;
    bool ok = true;
    for (u64 i = 0; i < std::max<u64>(packed_log, 1); i++) { ok = ok && p != minus_one; p = p * p; }
    if (ok) return v;
  }
}

// ================================================================ TBV (tbv.rs)
E8 monomial_weight(u64 index, const E8* p, size_t n) {  // field.rs:45-55
  E8 w = e8_one();
  for (size_t b = 0; b < n; b++) if ((index >> b) & 1) w = w * p[b];
  return w;
}
E8 equality_weight(u64 index, const std::vector<E8>& p) {  // field.rs:56-69
  E8 w = e8_one();
  for (size_t c = 0; c < p.size(); c++) w = w * (((index >> (p.size() - 1 - c)) & 1) ? p[c] : e8_one() - p[c]);
  return w;
}
E8 Term::weight(u64 index) const {
  E8 v = scale;
  for (size_t b = 0; b < factors.size(); b++) v = v * factors[b][(index >> b) & 1];
  return v;
}
E8 Term::evaluation(const std::vector<E8>& pl) const {  // tbv.rs:28-40
  E8 local = scale;
  size_t k = factors.size();
  for (size_t i = 0; i < k; i++) local = local * (factors[i][0] + factors[i][1] * pl[i]);
  return local * monomial_weight(offset >> k, pl.data() + k, pl.size() - k);
}
E8 Query::evaluation(const std::vector<E8>& pl) const {
  E8 s = e8_zero();
  for (auto& t : terms) s += t.evaluation(pl);
  return s;
}
// tbv.rs:45-76: result[i] = prod_bit pair_bit[(i >> bit) & 1]
std::vector<E8> tbv_weights(const std::array<E8, 2>* f, size_t n) {
  std::vector<E8> r(1, e8_one());
  for (size_t k = 0; k < n; k++) {
    size_t len = r.size();
    r.resize(2 * len);
    for (size_t i = 0; i < len; i++) { r[len + i] = r[i] * f[k][1]; r[i] = r[i] * f[k][0]; }
  }
  return r;
}
// tbv.rs:88-104
static std::vector<std::pair<u64, u64>> aligned_intervals(u64 start, u64 end) {
  std::vector<std::pair<u64, u64>> out;
  while (start < end) {
    int fit = ilog2(end - start);
    int bits = fit;
    if (start != 0) bits = std::min(fit, __builtin_ctzll(start));
    u64 len = u64(1) << bits;
    out.push_back({start, len});
    start += len;
  }
  return out;
}
// tbv.rs:106-123
static Term compact_term(Term t) {
  while (!t.factors.empty()) {
    auto pair = t.factors.back();
    u64 fixed;
    if (e8_is_zero(pair[0])) fixed = 1;
    else if (e8_is_zero(pair[1])) fixed = 0;
    else break;
    u64 bit = t.factors.size() - 1;
    t.scale = t.scale * pair[fixed];
    t.offset += fixed << bit;
    t.factors.pop_back();
  }
  return t;
}
static E8 small_val(long v) { return e8_base(fsmall(v)); }
// tbv.rs:288-352: lane sums of value * weight over the term's rows (exact integers mod p)
std::vector<E8> tbv_project(const Term& t, const Val* values, u64 n_values, u64 stride) {
  size_t low_bits = std::min<size_t>(ilog2(stride), t.factors.size());
  u64 lanes = u64(1) << low_bits;
  std::vector<E8> out(lanes, e8_zero());
  if (t.offset >= n_values || e8_is_zero(t.scale)) return out;
  std::vector<E8> hw = tbv_weights(t.factors.data() + low_bits, t.factors.size() - low_bits);
  u64 row_count = std::min<u64>((n_values - t.offset + lanes - 1) / lanes, t.len() / lanes);
  // i16 x u31 products summed over up to 2^c rows need more than 64 bits.
  std::vector<std::array<__int128, LANES>> acc(lanes);
  for (auto& a : acc) a.fill(0);
  for (u64 row = 0; row < row_count; row++) {
    E8 w = t.scale * hw[row];
    u64 idx = t.offset + row * lanes;
    for (u64 lane = 0; lane < lanes && idx + lane < n_values; lane++) {
      __int128 v = values[idx + lane];
      for (int k = 0; k < LANES; k++) acc[lane][k] += v * __int128(w.c[k]);
    }
  }
  for (u64 lane = 0; lane < lanes; lane++)
    for (int k = 0; k < LANES; k++) out[lane].c[k] = Coef(((acc[lane][k] % __int128(P)) + P) % P);
  return out;
}
// tbv.rs:361-392
void cross_terms(const Query& q, const Val* values, u64 n_values, u64 s, std::vector<E8>& pos, std::vector<E8>& neg) {
  pos.assign(s - 1, e8_zero());
  neg.assign(s - 1, e8_zero());
  auto val = [&](u64 i) { return i < n_values ? small_val(values[i]) : e8_zero(); };
  for (auto& t : q.terms) {
    if (t.len() >= s) {
      auto proj = tbv_project(t, values, n_values, s);
      auto low = tbv_weights(t.factors.data(), ilog2(s));
      for (u64 d = 1; d < s; d++)
        for (u64 lane = 0; lane + d < s; lane++) {
          pos[d - 1] += proj[lane] * low[lane + d];
          neg[d - 1] += proj[lane + d] * low[lane];
        }
    } else {
      u64 start = t.offset / s * s;
      std::vector<E8> query(s, e8_zero());
      for (u64 i = 0; i < t.len(); i++) query[t.offset - start + i] = t.weight(i);
      for (u64 d = 1; d < s; d++)
        for (u64 lane = 0; lane + d < s; lane++) {
          pos[d - 1] += query[lane + d] * val(start + lane);
          neg[d - 1] += query[lane] * val(start + lane + d);
        }
    }
  }
}
// tbv.rs:393-428
std::vector<E8> query_fold(const Query& q, u64 logical, u64 s, const E8& c) {
  std::vector<E8> result(logical / s, e8_zero());
  std::vector<E8> powers(s);
  powers[0] = e8_one();
  for (u64 i = 1; i < s; i++) powers[i] = powers[i - 1] * c;
  int ls = ilog2(s);
  for (auto& t : q.terms) {
    if (t.len() >= s) {
      auto low = tbv_weights(t.factors.data(), ls);
      E8 acc = e8_zero();
      for (u64 i = 0; i < s; i++) acc += low[i] * powers[i];
      E8 scale = t.scale * acc;
      auto hw = tbv_weights(t.factors.data() + ls, t.factors.size() - ls);
      u64 start = t.offset / s;
      for (u64 j = 0; j < hw.size(); j++) result[start + j] += scale * hw[j];
    } else {
      for (u64 i = 0; i < t.len(); i++) result[t.offset / s] += t.weight(i) * powers[(t.offset + i) % s];
    }
  }
  return result;
}
static void validate_claims(const Layout& l, const std::vector<Claim>& claims) {  // tbv.rs:124-142
  check(claims.size() == l.shapes.size(), "claim count mismatch");
  for (size_t i = 0; i < claims.size(); i++) {
    u64 rp = 1, cp = 1;
    while (rp < l.shapes[i].rows) rp <<= 1;
    while (cp < l.shapes[i].cols) cp <<= 1;
    check(claims[i].row.size() == size_t(ilog2(rp)) && claims[i].col.size() == size_t(ilog2(cp)), "point dimension mismatch");
  }
}
// tbv.rs:187-284
Query tbv_compile(const Layout& l, const std::vector<Claim>& claims, OuterTranscript& tr) {
  validate_claims(l, claims);
  u64 np = 1;
  while (np < claims.size()) np <<= 1;
  std::vector<E8> batch;
  for (int i = 0; i < ilog2(np); i++) batch.push_back(tr.challenge("batch"));
  std::vector<E8> coef(claims.size());
  for (size_t i = 0; i < claims.size(); i++) coef[i] = equality_weight(i, batch);
  Query q;
  q.claim = e8_zero();
  for (size_t i = 0; i < claims.size(); i++) q.claim += coef[i] * claims[i].value;
  for (auto& b : l.buckets) {
    const Claim& c = claims[b.matrix];
    Term t;
    t.offset = b.offset;
    t.scale = coef[b.matrix] * monomial_weight(b.row_offset, c.row.data(), c.row.size()) *
              monomial_weight(b.col_offset, c.col.data(), c.col.size());
    for (int i = 0; i < ilog2(b.cols); i++) t.factors.push_back({e8_one(), c.col[i]});
    for (int i = 0; i < ilog2(b.rows); i++) t.factors.push_back({e8_one(), c.row[i]});
    q.terms.push_back(std::move(t));
  }
  std::vector<E8> pp;
  for (int i = 0; i < ilog2(l.logical); i++) pp.push_back(tr.challenge("padding-point"));
  E8 mixing = tr.nonzero("padding-mix");
  for (auto& b : l.buckets) {
    int col_bits = ilog2(b.cols), bits = ilog2(b.len()), row_bits = ilog2(b.rows);
    std::vector<std::array<u64, 4>> rects;
    for (auto [s, n] : aligned_intervals(b.physical_rows, b.rows)) rects.push_back({s, n, 0, b.cols});
    for (auto [rs, rl] : aligned_intervals(0, b.physical_rows))
      for (auto [cs, cl] : aligned_intervals(b.physical_cols, b.cols)) rects.push_back({rs, rl, cs, cl});
    for (auto& r : rects) {
      Term t;
      for (int i = 0; i < bits; i++) t.factors.push_back({e8_one(), pp[i]});
      for (int bit = ilog2(r[3]); bit < col_bits; bit++) t.factors[bit][1 - ((r[2] >> bit) & 1)] = e8_zero();
      for (int bit = ilog2(r[1]); bit < row_bits; bit++) t.factors[col_bits + bit][1 - ((r[0] >> bit) & 1)] = e8_zero();
      t.offset = b.offset;
      t.scale = mixing * monomial_weight(b.offset >> bits, pp.data() + bits, pp.size() - bits);
      q.terms.push_back(compact_term(std::move(t)));
    }
  }
  for (auto [off, len] : aligned_intervals(l.active, l.logical)) {
    int bits = ilog2(len);
    Term t;
    t.offset = off;
    t.scale = mixing * monomial_weight(off >> bits, pp.data() + bits, pp.size() - bits);
    for (int i = 0; i < bits; i++) t.factors.push_back({e8_one(), pp[i]});
    q.terms.push_back(std::move(t));
  }
  return q;
}
// tbv.rs:143-185 (statement generation for benchmarks; never used by verification)
std::vector<Claim> evaluate_claims(const Layout& l, const Val* values, u64 n_values,
                                   const std::vector<std::pair<std::vector<E8>, std::vector<E8>>>& points) {
  check(n_values == l.active && points.size() == l.shapes.size(), "claim input dimensions");
  std::vector<Claim> claims(points.size());
  for (size_t i = 0; i < points.size(); i++) { claims[i].row = points[i].first; claims[i].col = points[i].second; claims[i].value = e8_zero(); }
  validate_claims(l, claims);
  for (auto& b : l.buckets) {
    auto& [row, col] = points[b.matrix];
    Term t;
    for (int i = 0; i < ilog2(b.cols); i++) t.factors.push_back({e8_one(), col[i]});
    for (int i = 0; i < ilog2(b.rows); i++) t.factors.push_back({e8_one(), row[i]});
    t.scale = monomial_weight(b.row_offset, row.data(), row.size()) * monomial_weight(b.col_offset, col.data(), col.size());
    t.offset = b.offset;
    auto partial = tbv_project(t, values, n_values, 16);
    auto low = tbv_weights(t.factors.data(), std::min<size_t>(t.factors.size(), 4));
    E8 v = e8_zero();
    for (size_t i = 0; i < partial.size(); i++) v += partial[i] * low[i];
    claims[b.matrix].value += v;
  }
  return claims;
}

// ================================================================ EO (eo.rs:97-181)
EoProof eo_prove(const Val* values, u64 n_values, u64 logical, u64 s, const Query& q, OuterTranscript& tr,
                 Terminal& terminal) {
  EoProof pf;
  cross_terms(q, values, n_values, s, pf.positive, pf.negative);
  tr.object("packed-positive", json_scalars(pf.positive));
  tr.object("packed-negative", json_scalars(pf.negative));
  E8 c = tr.fold_challenge(ilog2(s));
  std::vector<E8> challenges;
  E8 power = c;
  for (int i = 0; i < ilog2(s); i++) { challenges.push_back(power); power = power * power; }
  E8 inv = e8_inv(c);
  std::vector<E8> pinv(s);
  pinv[0] = e8_one();
  for (u64 i = 1; i < s; i++) pinv[i] = pinv[i - 1] * inv;
  std::vector<E8> w(logical / s);
  for (u64 row = 0; row < logical / s; row++) {
    int64_t acc[LANES] = {0};
    for (u64 lane = 0; lane < s; lane++) {
      u64 i = row * s + lane;
      int64_t v = i < n_values ? values[i] : 0;
      for (int k = 0; k < LANES; k++) acc[k] += v * int64_t(pinv[lane].c[k]);
    }
    for (int k = 0; k < LANES; k++) w[row].c[k] = Coef(((acc[k] % int64_t(P)) + P) % P);
  }
  std::vector<E8> qv = query_fold(q, logical, s, c);
  // update_claim (eo.rs:27-38)
  E8 claim = q.claim, fwd = c, bwd = inv;
  for (size_t d = 0; d < pf.positive.size(); d++) {
    claim += pf.positive[d] * fwd + pf.negative[d] * bwd;
    fwd = fwd * c;
    bwd = bwd * inv;
  }
  E8 dot = e8_zero();
  for (size_t i = 0; i < w.size(); i++) dot += w[i] * qv[i];
  check(dot == claim, "packed reduction claim mismatch");
  while (w.size() > 1) {
    E8 c0 = e8_zero(), c1 = e8_zero();
    for (size_t i = 0; i < w.size() / 2; i++) { c0 += w[2 * i] * qv[2 * i + 1]; c1 += w[2 * i + 1] * qv[2 * i]; }
    tr.object("binary-cross", json_scalars({c0, c1}));
    E8 r = tr.fold_challenge(1);
    E8 ri = e8_inv(r);
    claim += c0 * r + c1 * ri;
    challenges.push_back(r);
    for (size_t i = 0; i < w.size() / 2; i++) { w[i] = w[2 * i] + w[2 * i + 1] * ri; qv[i] = qv[2 * i] + qv[2 * i + 1] * r; }
    w.resize(w.size() / 2);
    qv.resize(qv.size() / 2);
    pf.rounds.push_back({c0, c1});
  }
  check(claim == w[0] * qv[0] && q.evaluation(challenges) == qv[0], "EO terminal mismatch");
  tr.scalar("terminal-value", w[0]);
  tr.scalar("terminal-query", qv[0]);
  // normalized (eo.rs:39-54)
  terminal.value = w[0];
  terminal.scale = e8_one();
  terminal.point_msb.clear();
  for (auto& ch : challenges) {
    E8 m = e8_inv(ch);
    E8 f = e8_one() + m;
    terminal.point_msb.push_back(m * e8_inv(f));
    terminal.scale = terminal.scale * f;
  }
  std::reverse(terminal.point_msb.begin(), terminal.point_msb.end());
  pf.value = w[0];
  pf.query = qv[0];
  return pf;
}

// ================================================================ PCS transcript (qafold_short_pc.rs:2011-2229)
static const char* PCS_DOMAIN = "brakefold/qafold-short-pcs/v2";
struct PcsTranscript {
  Blake2s h;
  u64 counter = 0;
  std::vector<E8> responses;
  PcsTranscript() { absorb_bytes("domain", PCS_DOMAIN, std::strlen(PCS_DOMAIN)); }
  void absorb_bytes(const char* label, const void* b, size_t n) { h.frame(label, b, n); }
  void absorb_str(const char* label, const char* s) { absorb_bytes(label, s, std::strlen(s)); }
  void absorb_u64(const char* label, u64 v) { h.frame_u64(label, v); }
  void absorb_digest(const char* label, const Digest& d) { absorb_bytes(label, d.data(), 32); }
  void absorb_ext(const char* label, const E8& v) { u8 b[16]; e8_write(v, b); absorb_bytes(label, b, 16); }
  void absorb_exts(const char* label, const std::vector<E8>& v) {
    absorb_str("vector-label", label);
    absorb_u64("vector-length", v.size());
    for (auto& x : v) absorb_ext("vector-element", x);
  }
  void push_response(const E8& v) { absorb_ext("response", v); responses.push_back(v); }
  Digest derive(const char* label, u64 ctr, u64 coefficient, u64 attempt) const {
    Blake2s c = h;
    c.frame("challenge-label", label, std::strlen(label));
    c.update_u64(ctr);
    c.update_u64(coefficient);
    c.update_u64(attempt);
    return c.finalize();
  }
  Digest challenge_digest(const char* label) {
    Digest d = derive(label, counter, 0, 0);
    counter++;
    absorb_str("digest-challenge-label", label);
    absorb_digest("digest-challenge", d);
    return d;
  }
  E8 challenge(const char* label, bool nonzero = false) {
    u64 ctr = counter;
    for (u64 outer = 0;; outer++) {
      E8 v;
      for (u64 attempt = outer;; attempt++) {
        Digest d = derive(label, ctr, 0, attempt);
        if (e8_from_uniform(d.data(), v)) break;
      }
      if (!nonzero || !e8_is_zero(v)) {
        counter++;
        absorb_str("field-challenge-label", label);
        absorb_ext("field-challenge", v);
        return v;
      }
    }
  }
};

// ================================================================ query counts (qafold.rs:75-178, brakefold.rs:4-53)
static bool log_one_minus_exp(double x, double& out) {
  if (x >= 0.0) return false;
  if (x < -M_LN2) out = std::log1p(-std::exp(x));
  else out = std::log(-std::expm1(x));
  return true;
}
static double log_add_exp(double a, double b) {
  double m = std::max(a, b);
  return m + std::log(std::exp(a - m) + std::exp(b - m));
}
static bool qa_epsilon_at(u64 n, u64 index, double d, double ln_p, double& eps) {
  if (n == 0 || index <= 1 || !(d > 0.0 && d < 1.0) || !(ln_p > 0.0)) return false;
  double ln_pm1 = ln_p + std::log1p(-std::exp(-ln_p));
  double g = 1.0 - d * ln_pm1 / ln_p + d * std::log(d) / ln_p + (1.0 - d) * std::log(1.0 - d) / ln_p;
  eps = g - (1.0 + std::log(double(n)) / ln_p) / double(index);
  return eps > 0.0;
}
static bool qa_bound_at(u64 n, u64 index, double d, double ln_p, double& eps, double& logf) {
  if (!qa_epsilon_at(n, index, d, ln_p, eps)) return false;
  double ln_pm1 = ln_p + std::log1p(-std::exp(-ln_p));
  double lome;
  if (!log_one_minus_exp(-double(index) * eps * ln_p, lome)) return false;
  double denom = lome + ln_pm1;
  double fc = double(index) * double(index - 1) * double(n) / 2.0;
  double fcl = std::log(fc) - 2.0 * ln_p;
  double fe = std::ceil(double(index - 1) / (double(index) * d)) * double(index) * eps;
  double ftl = -fe * ln_p - denom;
  double ffl = log_add_exp(fcl, ftl);
  double scl = std::log(double(index) * double(n)) - ln_p;
  double se = std::ceil(1.0 / d) * double(index) * eps;
  double stl = -se * ln_p - denom;
  double sfl = log_add_exp(scl, stl);
  logf = std::min(ffl, sfl);
  return true;
}
// Returns false where the reference reports InvalidParameters (all rows are queried).
static bool qa_distance_bound(u64 n, u64 index, u64 bits, double modulus_log2, double& rel) {
  double ln_p = modulus_log2 * M_LN2;
  double upper = 1.0 - 1.0 / double(index);
  double target = -double(bits) * M_LN2;
  auto valid = [&](double d) { double e, lf; return qa_bound_at(n, index, d, ln_p, e, lf) && lf <= target; };
  const u64 steps = 4096;
  double lower = 0.0, ub = upper;
  for (u64 step = 1; step <= steps; step++) {
    double cand = upper * double(step) / double(steps);
    if (valid(cand)) lower = cand;
    else if (cand > lower) { ub = cand; break; }
  }
  if (lower == 0.0) return false;
  for (int i = 0; i < 96; i++) {
    double mid = (lower + ub) * 0.5;
    if (valid(mid)) lower = mid; else ub = mid;
  }
  rel = lower * (1.0 - 1e-12);
  double e, lf;
  check(qa_bound_at(n, index, rel, ln_p, e, lf), "distance bound");
  return true;
}
// ShortQaCode::num_queries for the full extension field (FINITE_LENGTH_PROXIMITY, no RF layers)
u64 num_queries(const QaCode& code, u64 bits, u64 rows) {
  check(rows > 0, "no rows");
  unsigned __int128 card = 1;
  for (int i = 0; i < LANES; i++) card *= P;
  double log2card = std::log2(double(card));
  double rel;
  if (!qa_distance_bound(code.message_len, code.index, bits, log2card, rel)) return rows;
  check(std::isfinite(rel) && rel > 0.0 && rel < 1.0, "relative distance");
  double base = 1.0 - rel / 3.0;
  check(base > 0.0 && base < 1.0, "soundness base");
  double q = -double(bits) / std::log2(base);
  check(std::isfinite(q) && q >= 0.0, "query count");
  return std::min<u64>(u64(std::ceil(q)), rows);
}

// ================================================================ sampling (qafold_short_pc.rs:3000-3047)
static u64 sample_below(const Digest& seed, u64& counter, u64 upper) {
  unsigned __int128 range = (unsigned __int128)1 << 64;
  unsigned __int128 limit = range - range % upper;
  for (;;) {
    Blake2s h;
    h.update("qafold-short-index/v1", 21);
    h.update(seed.data(), 32);
    h.update_u64(counter);
    counter++;
    Digest d = h.finalize();
    u64 cand = 0;
    for (int i = 0; i < 8; i++) cand |= u64(d[i]) << (8 * i);
    if ((unsigned __int128)cand < limit) return u64(cand % upper);
  }
}
std::vector<u64> sample_without_replacement(u64 count, u64 population, const Digest& seed) {
  check(count > 0 && count <= population, "sample count");
  std::vector<u64> out;
  if (count == population) { for (u64 i = 0; i < population; i++) out.push_back(i); return out; }
  std::map<u64, u64> swaps;
  u64 counter = 0;
  for (u64 off = 0; off < count; off++) {
    u64 remaining = population - off;
    u64 s = sample_below(seed, counter, remaining);
    u64 last = remaining - 1;
    u64 sv = swaps.count(s) ? swaps[s] : s;
    u64 lv = swaps.count(last) ? swaps[last] : last;
    swaps[s] = lv;
    out.push_back(sv);
  }
  return out;
}

// ================================================================ wire helpers (qafold_short_pc.rs:3105-3212)
void write_base_rows_all(Bytes& out, const Committed& c) {
  out.push_back(1);
  put_u64(out, c.rows);
  put_u64(out, c.cols);
  put_u64(out, c.encoded.size());
  for (auto& v : c.encoded) put_e8(out, v);
}
void write_base_rows_partial(Bytes& out, const Committed& c, const std::vector<u64>& rows) {
  out.push_back(0);
  put_u64(out, rows.size());
  for (u64 r : rows) {
    put_u64(out, c.cols);
    for (size_t k = 0; k < c.cols; k++) put_e8(out, c.encoded[k * c.rows + r]);
    auto path = c.tree.path(r);
    put_u64(out, path.size());
    for (auto& d : path) out.insert(out.end(), d.begin(), d.end());
  }
}

// ================================================================ native opening
static std::vector<E8> tensor_weights(const std::vector<E8>& p, size_t from = 0, size_t to = SIZE_MAX) {
  if (to > p.size()) to = p.size();
  size_t n = to - from;
  std::vector<E8> w(size_t(1) << n, e8_zero());
  w[0] = e8_one();
  size_t filled = 1;
  for (size_t k = n; k-- > 0;) {  // reversed coordinates: MSB-first order
    E8 c = p[from + k];
    for (size_t i = 0; i < filled; i++) w[filled + i] = w[i] * c;
    filled *= 2;
  }
  return w;
}
static E8 tensor_weight_at(const std::vector<E8>& p, u64 pos) {
  E8 r = e8_one();
  for (size_t b = 0; b < p.size(); b++) if ((pos >> b) & 1) r = r * p[p.size() - b - 1];
  return r;
}
static E8 inner(const std::vector<E8>& a, const std::vector<E8>& b) {
  check(a.size() == b.size(), "inner product length");
  E8 s = e8_zero();
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}
static bool mle_eval(std::vector<E8> v, const std::vector<E8>& p, E8& out) {
  if (v.empty() || v.size() != (size_t(1) << p.size())) return false;
  size_t n = v.size();
  for (auto& c : p) { size_t h = n / 2; for (size_t i = 0; i < h; i++) v[i] += v[i + h] * c; n = h; }
  out = v[0];
  return true;
}
// G^T applied to a codeword-length vector: left[:n] + WHT(alpha_1 . WHT(left[n:])) (qafold_short.rs:1106-1163)
static std::vector<E8> generator_left(const QaCode& code, const std::vector<E8>& left) {
  size_t n = code.message_len;
  check(left.size() == 2 * n, "generator left length");
  std::vector<E8> v(left.begin() + n, left.end());
  wht(v.data(), n);
  for (size_t i = 0; i < n; i++) v[i] = v[i] * code.alphas[1][i];
  wht(v.data(), n);
  for (size_t i = 0; i < n; i++) v[i] += left[i];
  return v;
}
static std::vector<E8> generator_right(const QaCode& code, const std::vector<E8>& right) {
  std::vector<E8> out(code.codeword_len());
  code.encode(right.data(), out.data());
  return out;
}
static E8 generator_mle(const QaCode& code, const std::vector<E8>& lp, const std::vector<E8>& rp) {
  auto l = tensor_weights(lp), r = tensor_weights(rp);
  return inner(l, generator_right(code, r));
}

struct LiteBullet { E8 a_value, b_value; std::vector<E8> a_point, b_point; };
// qafold_short_pc.rs:1903-1948
static LiteBullet prove_litebullet(E8 claimed, std::vector<E8> left, std::vector<E8> right, PcsTranscript& tr,
                                   const char* label) {
  check(!left.empty() && is_pow2(left.size()) && left.size() == right.size() && inner(left, right) == claimed,
        "litebullet claim");
  tr.absorb_str("litebullet-domain", label);
  LiteBullet lb;
  E8 rhs = claimed;
  while (left.size() > 1) {
    size_t h = left.size() / 2;
    E8 lc = e8_zero(), rc = e8_zero();
    for (size_t i = 0; i < h; i++) { lc += left[i] * right[h + i]; rc += left[h + i] * right[i]; }
    tr.push_response(lc);
    tr.push_response(rc);
    E8 ch = tr.challenge("litebullet-fold", true);
    E8 ci = e8_inv(ch);
    rhs += lc * ch + rc * ci;
    for (size_t i = 0; i < h; i++) { left[i] += left[h + i] * ci; right[i] += right[h + i] * ch; }
    left.resize(h);
    right.resize(h);
    lb.a_point.push_back(ci);
    lb.b_point.push_back(ch);
  }
  lb.a_value = left[0];
  lb.b_value = right[0];
  tr.push_response(lb.a_value);
  tr.push_response(lb.b_value);
  check(lb.a_value * lb.b_value == rhs, "litebullet terminal");
  return lb;
}

struct RoundOut { E8 next_value, gen_value; std::vector<E8> next_point, gen_left, gen_right; };
// qafold_short_pc.rs:1454-1525
static RoundOut prove_round_reductions(const QaCode& code, const std::vector<E8>& point, const std::vector<E8>& freivalds,
                                       const E8& freivalds_value, const std::vector<E8>& folded,
                                       const std::vector<E8>& row_weights, const std::vector<u64>& selected,
                                       const std::vector<E8>& sampled, PcsTranscript& tr) {
  size_t n_ext = code.codeword_len();
  int col_log = ilog2(freivalds.size());
  u64 sp = 1;
  while (sp < selected.size()) sp <<= 1;
  std::vector<E8> sel_point;
  for (int i = 0; i < ilog2(sp); i++) sel_point.push_back(tr.challenge("selection-left"));
  std::vector<E8> sel_w(selected.size());
  for (size_t i = 0; i < selected.size(); i++) sel_w[i] = tensor_weight_at(sel_point, i);
  E8 selected_value = inner(sel_w, sampled);
  std::vector<E8> pi_left(n_ext, e8_zero());
  for (size_t i = 0; i < selected.size(); i++) pi_left[selected[i]] += sel_w[i];
  std::vector<E8> combined = generator_left(code, pi_left);
  E8 rho = tr.challenge("opening-batch");
  E8 combined_value = selected_value + rho * freivalds_value;
  for (size_t i = 0; i < combined.size(); i++) combined[i] += rho * row_weights[i];
  LiteBullet msg = prove_litebullet(combined_value, combined, folded, tr, "message-litebullet");
  E8 ole = e8_one();
  for (size_t i = 0; i < msg.a_point.size(); i++) ole = ole * (e8_one() + point[col_log + i] * msg.a_point[i]);
  E8 sle = msg.a_value - rho * ole;
  std::vector<E8> gr = generator_right(code, tensor_weights(msg.a_point));
  LiteBullet gen = prove_litebullet(sle, pi_left, gr, tr, "generator-litebullet");
  return {msg.b_value, gen.b_value, msg.b_point, gen.b_point, msg.a_point};
}

static Digest hash_terminal_vector(const std::vector<E8>& v) {
  Blake2s h;
  h.frame("domain", std::string("qafold-short/terminal-vector/v1"));
  h.frame_u64("length", v.size());
  for (auto& x : v) { u8 b[16]; e8_write(x, b); h.frame("element", b, 16); }
  return h.finalize();
}

// Commit an extension-level matrix (message column-major, ext_cols columns) (qafold_short_pc.rs:2327-2362).
static Committed commit_extension(const QaCode& code, const std::vector<E8>& entries, size_t cols) {
  Committed c;
  size_t n = code.message_len, R = code.codeword_len();
  check(entries.size() == n * cols, "extension matrix length");
  c.rows = R;
  c.cols = cols;
  c.encoded.assign(R * cols, e8_zero());
  for (size_t k = 0; k < cols; k++) code.encode(entries.data() + k * n, c.encoded.data() + k * R);
  c.tree = merkle_from_leaves(hash_rows(c.encoded.data(), R, cols, LEAF_EXTENSION_DOMAIN));
  return c;
}

// One opening round over a committed level (base round for the top level,
// extension rounds below; qafold_short_pc.rs:1120-1451).
static RoundOut prove_round(const PcsParams& pp, size_t level, const Digest& root, const Committed& cache,
                            const std::vector<E8>& point, const E8& value, bool top, PcsTranscript& tr, Bytes& rows_wire,
                            std::unique_ptr<Committed>& next_cache, std::vector<E8>& terminal) {
  const QaCode& code = pp.seq.codes[level];
  size_t n_rows = code.message_len, n_cols = cache.cols, R = cache.rows;
  int col_log = ilog2(n_cols);
  check(R == code.codeword_len() && point.size() == size_t(col_log + ilog2(n_rows)), "round dimensions");
  std::vector<E8> row_w = tensor_weights(point, col_log);
  std::vector<E8> opening(n_cols, e8_zero());
  for (size_t k = 0; k < n_cols; k++)
    for (size_t r = 0; r < n_rows; r++) opening[k] += cache.encoded[k * R + r] * row_w[r];
  if (!top) check(inner(opening, tensor_weights(point, 0, col_log)) == value, "extension round opening mismatch");
  tr.absorb_digest("round-root", root);
  for (auto& o : opening) tr.push_response(o);
  std::vector<E8> fr;
  for (size_t k = 0; k < n_cols; k++) fr.push_back(tr.challenge("freivalds"));
  E8 fv = inner(opening, fr);
  std::vector<E8> folded(n_rows, e8_zero());
  for (size_t r = 0; r < n_rows; r++)
    for (size_t k = 0; k < n_cols; k++) folded[r] += fr[k] * cache.encoded[k * R + r];
  if (level > 0) {
    next_cache.reset(new Committed(commit_extension(pp.seq.codes[level - 1], folded, pp.ext_cols)));
    tr.absorb_str("next-layer-kind", "extension-root");
    tr.absorb_digest("next-layer-binding", next_cache->root());
  } else {
    tr.absorb_str("next-layer-kind", "terminal-vector");
    tr.absorb_digest("next-layer-binding", hash_terminal_vector(folded));
    terminal = folded;
  }
  u64 count = num_queries(code, pp.security_bits, R);
  Digest seed = tr.challenge_digest("row-selection");
  std::vector<u64> selected;
  std::vector<E8> sampled;
  if (count == R) {
    for (u64 r = 0; r < R; r++) selected.push_back(r);
    write_base_rows_all(rows_wire, cache);
  } else {
    selected = sample_without_replacement(count, R, seed);
    write_base_rows_partial(rows_wire, cache, selected);
  }
  for (u64 r : selected) {
    E8 s = e8_zero();
    for (size_t k = 0; k < n_cols; k++) s += fr[k] * cache.encoded[k * R + r];
    sampled.push_back(s);
  }
  tr.absorb_exts("sampled-values", sampled);
  return prove_round_reductions(code, point, fr, fv, folded, row_w, selected, sampled, tr);
}

// qafold_short_pc.rs:578-725
E8 pcs_open_with_reduction(const PcsParams& pp, const Digest& root, const std::vector<E8>& point, const Committed& c,
                           Bytes& wire, OpeningReduction& red) {
  check(point.size() == size_t(ilog2(pp.n_witness)) && c.root() == root, "opening point or root");
  const QaCode& top = pp.seq.codes.back();
  size_t n_rows = top.message_len, cols = pp.ext_cols;
  int col_log = ilog2(cols);
  auto col_w = tensor_weights(point, 0, col_log), row_w = tensor_weights(point, col_log);
  std::vector<E8> opening(cols, e8_zero());
  for (size_t k = 0; k < cols; k++)
    for (size_t r = 0; r < n_rows; r++) opening[k] += c.encoded[k * c.rows + r] * row_w[r];
  E8 claimed = inner(opening, col_w);
  PcsTranscript tr;
  tr.absorb_digest("parameters", pp.parameter_digest);
  tr.absorb_digest("commitment", root);
  tr.absorb_exts("point", point);
  tr.absorb_ext("claimed-value", claimed);
  size_t levels = pp.seq.codes.size(), top_level = levels - 1;
  Bytes top_rows;
  std::vector<std::pair<Digest, Bytes>> ext_rounds;
  std::unique_ptr<Committed> next;
  std::vector<E8> terminal;
  RoundOut ro = prove_round(pp, top_level, root, c, point, claimed, true, tr, top_rows, next, terminal);
  red.pending.clear();
  auto settle = [&](size_t level, const RoundOut& o) {
    if (level == 0) check(generator_mle(pp.seq.codes[0], o.gen_left, o.gen_right) == o.gen_value, "generator claim");
    else red.pending.push_back({level, o.gen_value, o.gen_left, o.gen_right});
  };
  settle(top_level, ro);
  for (size_t level = top_level; level-- > 0;) {
    std::unique_ptr<Committed> cur = std::move(next);
    Digest r = cur->root();
    Bytes rows;
    ro = prove_round(pp, level, r, *cur, ro.next_point, ro.next_value, false, tr, rows, next, terminal);
    ext_rounds.push_back({r, std::move(rows)});
    settle(level, ro);
  }
  E8 tv;
  check(!terminal.empty() && mle_eval(terminal, ro.next_point, tv) && tv == ro.next_value, "terminal vector");
  wire.clear();
  const char* magic = "QFSHORT2";
  wire.insert(wire.end(), magic, magic + 8);
  put_u64(wire, tr.responses.size());
  for (auto& v : tr.responses) put_e8(wire, v);
  wire.insert(wire.end(), top_rows.begin(), top_rows.end());
  put_u64(wire, ext_rounds.size());
  for (auto& [r, rows] : ext_rounds) { wire.insert(wire.end(), r.begin(), r.end()); wire.insert(wire.end(), rows.begin(), rows.end()); }
  put_u64(wire, terminal.size());
  for (auto& v : terminal) put_e8(wire, v);
  return claimed;
}

// ---------------------------------------------------------------- Decide (direct-interleaved setup)
// qafold_short.rs:1575-1625: G_j claim -> H_j (alpha matrix) claim.
struct MatrixClaim { E8 value; std::vector<E8> left, right; };
static MatrixClaim reduce_generator_claim(u64 message_len, u64 index, const MatrixClaim& c) {
  size_t bv = ilog2(index);
  check(c.left.size() == size_t(ilog2(message_len)) + bv && c.right.size() == size_t(ilog2(message_len)), "generator claim dims");
  E8 norm = e8_one();
  std::vector<E8> merged;
  for (size_t i = 0; i < c.right.size(); i++) {
    E8 l = c.left[bv + i], r = c.right[i];
    E8 dl = e8_one() + l, dr = e8_one() + r;
    norm = norm * dl * dr;
    merged.push_back(((e8_one() - l) * e8_inv(dl)) * ((e8_one() - r) * e8_inv(dr)));
  }
  MatrixClaim out;
  out.value = c.value * e8_inv(norm);
  out.left = merged;  // row point of the alpha matrix
  out.right.assign(c.left.begin(), c.left.begin() + bv);  // block (column) point
  return out;
}
static E8 tensor_inner_product(const std::vector<E8>& a, const std::vector<E8>& b) {
  check(a.size() == b.size(), "tensor inner product");
  E8 r = e8_one();
  for (size_t i = 0; i < a.size(); i++) r = r * (e8_one() + a[i] * b[i]);
  return r;
}
struct Selection { bool all; u64 count; std::vector<u64> rows; };
static void absorb_selection(PcsTranscript& tr, const Selection& s) {  // qafold_short_decide_carry.rs:721-739
  tr.absorb_str("selection-kind", s.all ? "all" : "partial");
  tr.absorb_u64("selection-count", s.all ? s.count : s.rows.size());
  if (!s.all) for (u64 r : s.rows) tr.absorb_u64("selection-row", r);
}
static std::vector<u64> selection_rows(const Selection& s) {
  if (!s.all) return s.rows;
  std::vector<u64> v(s.count);
  for (u64 i = 0; i < s.count; i++) v[i] = i;
  return v;
}
// Multi-evaluation reduction on the mixed column-major matrix (qafold_short_decide.rs:486-551),
// restated as one LiteBullet over the flattened matrix and weight vectors.
struct EvalClaim { E8 value; std::vector<E8> row_point, col_weight; };
static std::pair<E8, std::vector<E8>> prove_multi_eval(const std::vector<EvalClaim>& claims,
                                                      const std::vector<const E8*>& columns, size_t log_rows,
                                                      size_t log_cols, PcsTranscript& tr, std::vector<E8>& folded_vector) {
  size_t rows = size_t(1) << log_rows, cols = size_t(1) << log_cols;
  tr.absorb_str("multi-eval-domain", "qafold-short-decide/multi-eval/v2");
  tr.absorb_u64("multi-eval-log-rows", log_rows);
  tr.absorb_u64("multi-eval-log-cols", log_cols);
  tr.absorb_u64("multi-eval-claim-count", claims.size());
  for (auto& c : claims) {
    tr.absorb_ext("multi-eval-claim-value", c.value);
    tr.absorb_exts("multi-eval-row-point", c.row_point);
    tr.absorb_exts("multi-eval-column-weight", c.col_weight);
  }
  E8 rho = tr.challenge("multi-eval-batch");
  E8 rhs = e8_zero(), pw = e8_one();
  std::vector<E8> left(rows * cols, e8_zero()), right(rows * cols, e8_zero());
  for (size_t k = 0; k < columns.size(); k++) std::memcpy(&left[k * rows], columns[k], rows * sizeof(E8));
  for (auto& c : claims) {
    rhs += pw * c.value;
    auto rw = tensor_weights(c.row_point);
    for (size_t k = 0; k < c.col_weight.size(); k++) {
      E8 s = pw * c.col_weight[k];
      for (size_t r = 0; r < rows; r++) right[k * rows + r] += s * rw[r];
    }
    pw = pw * rho;
  }
  std::vector<E8> a_point, b_point;
  size_t n = rows * cols;
  while (n > 1) {
    size_t h = n / 2;
    E8 lc = e8_zero(), rc = e8_zero();
    for (size_t i = 0; i < h; i++) { lc += left[i] * right[h + i]; rc += left[h + i] * right[i]; }
    tr.push_response(lc);
    tr.push_response(rc);
    E8 ch = tr.challenge("multi-eval-fold", true), ci = e8_inv(ch);
    rhs += lc * ch + rc * ci;
    for (size_t i = 0; i < h; i++) { left[i] += left[h + i] * ci; right[i] += right[h + i] * ch; }
    a_point.push_back(ci);
    b_point.push_back(ch);
    n = h;
    if (n == rows) folded_vector.assign(left.begin(), left.begin() + rows);
  }
  if (cols == 1) folded_vector.assign(columns[0], columns[0] + rows);
  tr.push_response(left[0]);
  tr.push_response(right[0]);
  check(left[0] * right[0] == rhs, "multi-eval terminal");
  return {left[0], a_point};
}
static void write_ext_rows(Bytes& out, const Committed& c, const Selection& s) {  // carry rows wire
  if (s.all) {
    out.push_back(1);
    put_u64(out, c.rows);
    put_u64(out, c.cols);
    put_u64(out, c.encoded.size());
    for (auto& v : c.encoded) put_e8(out, v);
  } else {
    out.push_back(0);
    put_u64(out, s.rows.size());
    for (u64 r : s.rows) {
      put_u64(out, c.cols);
      for (size_t k = 0; k < c.cols; k++) put_e8(out, c.encoded[k * c.rows + r]);
      auto path = c.tree.path(r);
      put_u64(out, path.size());
      for (auto& d : path) out.insert(out.end(), d.begin(), d.end());
    }
  }
}

// qafold_short_decide_driver.rs:51-318 (DirectInterleavedV1 branches)
Bytes pcs_prove_decide(const PcsParams& pp, const OpeningReduction& red) {
  size_t levels = pp.seq.codes.size();
  PcsTranscript tr;
  tr.absorb_str("decide-domain", "qafold-short/decide/v1");
  tr.absorb_digest("decide-parameters", pp.parameter_digest);
  tr.absorb_digest("decide-setup", pp.setup_digest);
  tr.absorb_u64("decide-levels", levels);
  u8 ok = 1;
  tr.absorb_bytes("opening-reduction-ok", &ok, 1);
  tr.absorb_u64("pending-h-count", red.pending.size());
  for (auto& p : red.pending) {
    tr.absorb_u64("pending-h-level", p.level);
    tr.absorb_ext("pending-g-value", p.value);
    tr.absorb_exts("pending-g-left", p.left);
    tr.absorb_exts("pending-g-right", p.right);
  }
  std::vector<std::vector<MatrixClaim>> h_claims(levels);
  for (auto& p : red.pending)
    h_claims[p.level].push_back(reduce_generator_claim(pp.seq.codes[p.level].message_len, pp.seq.codes[p.level].index,
                                                       {p.value, p.left, p.right}));
  Bytes rounds_wire;
  size_t n_rounds = 0;
  std::vector<E8> base_vector;
  const size_t rf = pp.ext_cols;
  bool have_w = false;
  EvalClaim w_claim;
  std::unique_ptr<Committed> carry;         // encoded carry of the current level
  std::vector<E8> carry_message;            // its logical message column
  for (size_t level = levels - 1; level >= 1; level--) {
    const SetupLevel& sl = pp.setup[level];
    const QaCode& code = pp.seq.codes[level];
    size_t rows = code.message_len, log_rows = ilog2(rows);
    std::vector<EvalClaim> claims;
    for (auto& h : h_claims[level]) claims.push_back({h.value, h.left, tensor_weights(h.right)});
    h_claims[level].clear();
    size_t setup_cols = code.alphas.size(), carry_cols = 0;
    if (have_w) {
      EvalClaim c = w_claim;
      c.col_weight.assign(setup_cols, e8_zero());
      c.col_weight.push_back(e8_one());
      claims.push_back(c);
      carry_cols = 1;
    }
    check(!claims.empty(), "no Decide claims");
    u64 real = setup_cols + carry_cols, padded = 1;
    while (padded < real) padded <<= 1;
    size_t log_cols = ilog2(padded);
    // round context (decide_driver.rs:716-749)
    tr.absorb_str("decide-round-domain", "qafold-short/decide-round/v1");
    tr.absorb_u64("decide-round-level", level);
    const SetupShape& sh = sl.shape;
    tr.absorb_u64("decide-round-shape-level", sh.level);
    tr.absorb_u64("decide-round-shape-encoded-rows", sh.encoded_rows);
    tr.absorb_u64("decide-round-shape-message-rows", sh.message_rows);
    tr.absorb_u64("decide-round-shape-t-cols", sh.t_cols);
    tr.absorb_u64("decide-round-shape-t-elements", sh.t_elements);
    tr.absorb_u64("decide-round-shape-t-last-len", sh.t_last_len);
    tr.absorb_u64("decide-round-shape-h-cols", sh.h_cols);
    Digest setup_root = sl.tree.root();
    tr.absorb_digest("decide-round-setup-root", setup_root);
    u8 present = carry ? 1 : 0;
    tr.absorb_bytes("decide-round-carry-present", &present, 1);
    if (carry) tr.absorb_digest("decide-round-carry-root", carry->root());
    std::vector<const E8*> columns;
    for (auto& a : code.alphas) columns.push_back(a.data());
    if (carry) columns.push_back(carry_message.data());
    std::vector<E8> folded_vector;
    auto [red_value, red_point] = prove_multi_eval(claims, columns, log_rows, log_cols, tr, folded_vector);
    std::vector<E8> col_point(red_point.begin(), red_point.begin() + log_cols);
    std::vector<E8> merged_left(red_point.begin() + log_cols, red_point.end());
    auto col_weights = tensor_weights(col_point);
    const QaCode& target = pp.seq.codes[level - 1];
    u64 n_enc = target.codeword_len();
    check(sh.encoded_rows == n_enc, "setup encoded rows");
    u64 count = num_queries(target, pp.security_bits, n_enc);
    Digest seed = tr.challenge_digest("decide-row-selection");
    Selection sel{count == n_enc, count, {}};
    if (!sel.all) sel.rows = sample_without_replacement(count, n_enc, seed);
    auto sel_rows = selection_rows(sel);
    // iota weights over physical columns (decide_driver.rs:617-653)
    size_t block_log = ilog2(rf);
    auto block_w = tensor_weights(std::vector<E8>(merged_left.begin(), merged_left.begin() + block_log));
    u64 phys = real * rf, padded_phys = 1;
    while (padded_phys < phys) padded_phys <<= 1;
    std::vector<E8> iw(padded_phys, e8_zero());
    for (u64 lc = 0; lc < real; lc++)
      for (size_t b = 0; b < rf; b++) iw[lc * rf + b] = col_weights[lc] * block_w[b];
    size_t phys_carry = carry_cols * rf;
    // iota statement and values (decide_carry.rs:336-380, 471-615)
    tr.absorb_str("iota-domain", "qafold-short/decide-iota/v1");
    tr.absorb_u64("iota-level", level);
    tr.absorb_u64("iota-shape-level", sh.level);
    tr.absorb_u64("iota-shape-encoded-rows", sh.encoded_rows);
    tr.absorb_u64("iota-shape-message-rows", sh.message_rows);
    tr.absorb_u64("iota-shape-t-cols", sh.t_cols);
    tr.absorb_u64("iota-shape-t-elements", sh.t_elements);
    tr.absorb_u64("iota-shape-t-last-len", sh.t_last_len);
    tr.absorb_u64("iota-shape-h-cols", sh.h_cols);
    tr.absorb_digest("iota-setup-root", setup_root);
    tr.absorb_bytes("iota-carry-present", &present, 1);
    if (carry) tr.absorb_digest("iota-carry-root", carry->root());
    tr.absorb_u64("iota-carry-cols", phys_carry);
    absorb_selection(tr, sel);
    tr.absorb_exts("iota-column-weights", iw);
    size_t scols = sh.t_cols + sh.h_cols;
    std::vector<E8> iota;
    for (u64 r : sel_rows) {
      E8 v = e8_zero();
      for (size_t c = 0; c < scols; c++) v += iw[c] * sl.encoded[c * n_enc + r];
      if (carry)
        for (size_t c = 0; c < phys_carry; c++) v += iw[scols + c] * carry->encoded[c * n_enc + r];
      iota.push_back(v);
    }
    for (auto& v : iota) tr.push_response(v);
    // round wire: level, next carry root, setup rows, carry rows
    Bytes round;
    put_u64(round, level);
    // fold the block coordinate (decide_carry.rs:182-222)
    size_t trows = target.message_len;
    std::vector<E8> rec(trows, e8_zero());
    for (size_t r = 0; r < trows; r++)
      for (size_t b = 0; b < rf; b++) rec[r] += folded_vector[b * trows + r] * block_w[b];
    std::unique_ptr<Committed> next_carry;
    if (level > 1) {
      const QaCode& nt = pp.seq.codes[level - 2];
      check(trows == nt.message_len * rf, "carry row factor");
      next_carry.reset(new Committed(commit_extension(nt, rec, rf)));
    }
    tr.absorb_str("decide-next-carry-domain", "qafold-short/decide-next-carry/v1");
    u8 np = next_carry ? 1 : 0;
    tr.absorb_bytes("decide-next-carry-present", &np, 1);
    if (next_carry) tr.absorb_digest("decide-next-carry-root", next_carry->root());
    round.push_back(np);
    if (next_carry) { Digest r = next_carry->root(); round.insert(round.end(), r.begin(), r.end()); }
    if (sel.all) {
      round.push_back(1);
      for (u64 v : {sh.level, sh.encoded_rows, sh.message_rows, sh.t_cols, sh.t_elements, sh.t_last_len, sh.h_cols}) put_u64(round, v);
      put_u64(round, sl.encoded.size());
      for (auto& v : sl.encoded) put_e8(round, v);
    } else {
      round.push_back(0);
      put_u64(round, sel.rows.size());
      for (u64 r : sel.rows) {
        put_u64(round, r);
        put_u64(round, scols);
        for (size_t c = 0; c < scols; c++) put_e8(round, sl.encoded[c * n_enc + r]);
        auto path = sl.tree.path(r);
        put_u64(round, path.size());
        for (auto& d : path) round.insert(round.end(), d.begin(), d.end());
      }
    }
    if (carry) { round.push_back(1); write_ext_rows(round, *carry, sel); } else round.push_back(0);
    // direct code switch (decide_carry.rs:940-1027)
    std::vector<E8> base_point(merged_left.begin() + block_log, merged_left.end());
    E8 mv;
    check(mle_eval(rec, base_point, mv) && mv == red_value, "merged carry value");
    tr.absorb_str("direct-code-switch-domain", "qafold-short/decide-direct-code-switch/v2");
    tr.absorb_u64("direct-code-switch-target-level", level - 1);
    tr.absorb_u64("direct-code-switch-row-factor", rf);
    tr.absorb_u64("direct-code-switch-message-len", target.message_len);
    tr.absorb_u64("direct-code-switch-encoded-len", n_enc);
    tr.absorb_u64("direct-code-switch-target-index", target.index);
    tr.absorb_digest("direct-code-switch-target-generator", target.generator_digest);
    absorb_selection(tr, sel);
    tr.absorb_exts("direct-code-switch-iota", iota);
    tr.absorb_ext("direct-code-switch-merged-value", red_value);
    tr.absorb_exts("direct-code-switch-merged-left-point", merged_left);
    u64 sp = 1;
    while (sp < sel_rows.size()) sp <<= 1;
    std::vector<E8> slp;
    for (int i = 0; i < ilog2(sp); i++) slp.push_back(tr.challenge("direct-code-switch-selection-left"));
    std::vector<E8> sw(sel_rows.size());
    for (size_t i = 0; i < sel_rows.size(); i++) sw[i] = tensor_weight_at(slp, i);
    E8 selected_value = inner(sw, iota);
    std::vector<E8> sparse(n_enc, e8_zero());
    for (size_t i = 0; i < sel_rows.size(); i++) sparse[sel_rows[i]] += sw[i];
    auto gl = generator_left(target, sparse);
    E8 rho = tr.challenge("direct-code-switch-opening-batch");
    auto ml = tensor_weights(base_point);
    for (size_t i = 0; i < gl.size(); i++) gl[i] += rho * ml[i];
    LiteBullet msg = prove_litebullet(selected_value + rho * red_value, gl, rec, tr, "decide-direct-carry-message-litebullet");
    E8 sle = msg.a_value - rho * tensor_inner_product(base_point, msg.a_point);
    LiteBullet gen = prove_litebullet(sle, sparse, generator_right(target, tensor_weights(msg.a_point)), tr,
                                      "decide-direct-carry-generator-litebullet");
    MatrixClaim gclaim{gen.b_value, gen.b_point, msg.a_point};
    if (level > 1) {
      h_claims[level - 1].push_back(reduce_generator_claim(target.message_len, target.index, gclaim));
      w_claim = {msg.b_value, msg.b_point, {}};
      have_w = true;
      carry = std::move(next_carry);
      carry_message = rec;
    } else {
      E8 tv;
      check(generator_mle(target, gclaim.left, gclaim.right) == gclaim.value, "base generator claim");
      check(mle_eval(rec, msg.b_point, tv) && tv == msg.b_value, "terminal carry");
      for (auto& v : rec) tr.push_response(v);
      base_vector = rec;
      carry.reset();
    }
    rounds_wire.insert(rounds_wire.end(), round.begin(), round.end());
    n_rounds++;
  }
  Bytes w;
  const char* magic = "QFSDCID1";
  w.insert(w.end(), magic, magic + 8);
  put_u64(w, tr.responses.size());
  for (auto& v : tr.responses) put_e8(w, v);
  put_u64(w, n_rounds);
  w.insert(w.end(), rounds_wire.begin(), rounds_wire.end());
  put_u64(w, base_vector.size());
  for (auto& v : base_vector) put_e8(w, v);
  return w;
}

// ================================================================ packed PCS
struct PackedTranscript {  // packed.rs:754-779
  Blake2s h;
  void bytes(const void* b, size_t n) { h.update_u64(n); h.update(b, n); }
  void field(const E8& v) { u8 b[16]; e8_write(v, b); bytes(b, 16); }
  E8 challenge() {
    for (u64 ctr = 0;; ctr++) {
      Blake2s c = h;
      c.update("challenge", 9);
      c.update_u64(ctr);
      Digest d = c.finalize();
      E8 v;
      if (e8_from_uniform(d.data(), v)) { bytes(d.data(), 32); return v; }
    }
  }
};
static PackedTranscript packed_transcript(const PcsParams& pp, int log_n, const Digest& root, const std::vector<E8>& point,
                                          const E8& value, const Bytes& context) {
  PackedTranscript t;
  t.bytes(PACKED_DOMAIN, std::strlen(PACKED_DOMAIN));
  t.bytes(pp.parameter_digest.data(), 32);
  u8 ln[8];
  for (int i = 0; i < 8; i++) ln[i] = u8(u64(log_n) >> (8 * i));
  t.bytes(ln, 8);
  t.bytes(root.data(), 32);
  t.bytes(context.data(), context.size());
  for (auto& c : point) t.field(c);
  t.field(value);
  return t;
}
static std::vector<E8> equality_table(const E8* p, size_t n) {  // packed.rs:525-545
  std::vector<E8> t(1, e8_one());
  for (size_t k = 0; k < n; k++) {
    std::vector<E8> nx(t.size() * 2);
    for (size_t i = 0; i < t.size(); i++) { nx[2 * i + 1] = t[i] * p[k]; nx[2 * i] = t[i] - nx[2 * i + 1]; }
    t = std::move(nx);
  }
  return t;
}
static E8 basis(int j) { E8 b = e8_zero(); b.c[j] = 1; return b; }
using Lanes = std::array<E8, LANES>;
static Lanes lane_powers(const E8& v) {
  Lanes o;
  o[0] = e8_one();
  for (int i = 1; i < LANES; i++) o[i] = o[i - 1] * v;
  return o;
}
// packed.rs:502-523
static E8 projected_claim(const Lanes& ev, const std::vector<E8>& low) {
  auto inv_basis = lane_powers(e8_inv(basis(1)));
  Coef omega_inv = finv(FROB_ROOT), inv_lanes = finv(LANES);
  Lanes conj;
  for (int i = 0; i < LANES; i++) conj[i] = e8_frob(ev[i], i);
  Coef step = 1;
  E8 claim = e8_zero();
  for (int j = 0; j < LANES; j++) {
    Coef phase = 1;
    E8 sum = e8_zero();
    for (int i = 0; i < LANES; i++) { sum += e8_mul_base(conj[i], phase); phase = fmul(phase, step); }
    E8 lane = e8_mul_base(sum * inv_basis[j], inv_lanes);
    claim += low[j] * lane;
    step = fmul(step, omega_inv);
  }
  return claim;
}
// lanes_j = sum_i coeff_j(values_i) * weights_i (packed.rs:569-592)
static Lanes project_lanes(const std::vector<E8>& values, const std::vector<E8>& w) {
  Lanes out;
  for (int j = 0; j < LANES; j++) {
    unsigned __int128 raw[LANES] = {0};
    for (size_t i = 0; i < values.size(); i++)
      for (int k = 0; k < LANES; k++) raw[k] += (unsigned __int128)(values[i].c[j]) * w[i].c[k];
    for (int k = 0; k < LANES; k++) out[j].c[k] = Coef(raw[k] % P);
  }
  return out;
}
static void apply_weight_images(std::vector<E8>& w, const Lanes& images) {  // packed.rs:634-650
  for (auto& x : w) {
    E8 r = e8_zero();
    for (int j = 0; j < LANES; j++) r += e8_mul_base(images[j], x.c[j]);
    x = r;
  }
}
static E8 batch_weight(const E8* point, size_t n, const std::vector<E8>& reduced, const Lanes& powers) {
  E8 sum = e8_zero();
  for (int i = 0; i < LANES; i++) {
    E8 w = e8_one();
    for (size_t k = 0; k < n; k++) {
      E8 s = e8_frob(point[k], (LANES - i) % LANES);
      w = w * ((e8_one() - s) * (e8_one() - reduced[k]) + s * reduced[k]);
    }
    sum += powers[i] * w;
  }
  return sum;
}
// Degree-2 sumcheck over (values, weights), MSB half first (packed.rs:652-696).
static void sumcheck(std::vector<E8> values, std::vector<E8> weights, E8& claim, PackedTranscript& t,
                     std::vector<std::array<E8, 2>>& rounds, std::vector<E8>& reduced, E8& final_value, E8& final_weight) {
  while (values.size() > 1) {
    size_t h = values.size() / 2;
    E8 m0 = e8_zero(), m1 = e8_zero();
    for (size_t i = 0; i < h; i++) {
      m0 += values[i] * weights[i];
      m1 += (values[i + h] - values[i]) * (weights[i + h] - weights[i]);
    }
    t.field(m0);
    t.field(m1);
    E8 r = t.challenge();
    E8 lin = claim - m0 - m0 - m1;
    claim = m0 + r * (lin + r * m1);
    for (size_t i = 0; i < h; i++) {
      values[i] = values[i] + r * (values[i + h] - values[i]);
      weights[i] = weights[i] + r * (weights[i + h] - weights[i]);
    }
    values.resize(h);
    weights.resize(h);
    rounds.push_back({m0, m1});
    reduced.push_back(r);
  }
  final_value = values[0];
  final_weight = weights[0];
}
static void put_part(Bytes& w, const Bytes& part) { put_u64(w, part.size()); w.insert(w.end(), part.begin(), part.end()); }

// packed_batch.rs:139-319 (recompute-mode codeword rebuild; the cached path writes the same bytes)
Bytes packed_open_batch(const PcsParams& pp, int log_n, const std::vector<Digest>& roots, const std::vector<E8>& point,
                        const Bytes& context, const BlockSource& src, E8& claimed_value) {
  size_t blocks = roots.size();
  check(blocks >= 2 && blocks <= 128 && is_pow2(blocks) && point.size() == size_t(log_n + ilog2(blocks)), "batch dims");
  size_t split = ilog2(blocks);
  size_t n_outer = point.size() - LOG_LANES - split;
  const E8* outer = point.data() + split;
  auto low = equality_table(point.data() + point.size() - LOG_LANES, LOG_LANES);
  auto high = equality_table(point.data(), split);
  auto weights = equality_table(outer, n_outer);
  size_t plen = size_t(1) << (log_n - LOG_LANES);
  auto packed_block = [&](size_t b) {
    std::vector<E8> v(plen);
    for (size_t i = 0; i < plen; i++)
      for (int j = 0; j < LANES; j++) v[i].c[j] = fsmall(src(b, i * LANES + j));
    return v;
  };
  std::vector<Lanes> ebb(blocks);
  for (size_t b = 0; b < blocks; b++) {
    auto lanes = project_lanes(packed_block(b), weights);
    for (int img = 0; img < LANES; img++) {
      E8 s = e8_zero();
      for (int lane = 0; lane < LANES; lane++) s += basis(lane) * e8_frob(lanes[lane], (LANES - img) % LANES);
      ebb[b][img] = s;
    }
  }
  Lanes ev;
  for (int img = 0; img < LANES; img++) {
    E8 s = e8_zero();
    for (size_t b = 0; b < blocks; b++) s += e8_frob(high[b], (LANES - img) % LANES) * ebb[b][img];
    ev[img] = s;
  }
  claimed_value = projected_claim(ev, low);
  PackedTranscript t = packed_transcript(pp, log_n, roots[0], point, claimed_value, context);
  const char* bd = "one-opening-block-packed-eo/v1";
  t.bytes(bd, std::strlen(bd));
  u8 nb[8];
  for (int i = 0; i < 8; i++) nb[i] = u8(u64(blocks) >> (8 * i));
  t.bytes(nb, 8);
  for (auto& r : roots) t.bytes(r.data(), 32);
  for (auto& v : ev) t.field(v);
  auto powers = lane_powers(t.challenge());
  E8 claim = e8_zero();
  for (int i = 0; i < LANES; i++) claim += powers[i] * ev[i];
  auto cross = [&](size_t w, size_t q) {
    E8 s = e8_zero();
    for (int i = 0; i < LANES; i++) s += powers[i] * e8_frob(high[q], (LANES - i) % LANES) * ebb[w][i];
    return s;
  };
  std::vector<E8> pos, neg;
  for (size_t d = 1; d < blocks; d++) {
    E8 p = e8_zero(), n = e8_zero();
    for (size_t b = 0; b + d < blocks; b++) { p += cross(b, b + d); n += cross(b + d, b); }
    pos.push_back(p);
    neg.push_back(n);
  }
  const char* cd = "block-packed-eo-cross";
  t.bytes(cd, std::strlen(cd));
  for (auto& v : pos) t.field(v);
  for (auto& v : neg) t.field(v);
  E8 omega;
  do { omega = t.challenge(); } while (e8_is_zero(omega));
  {
    E8 inv = e8_inv(omega), f = omega, bk = inv;
    for (size_t d = 0; d < pos.size(); d++) { claim += pos[d] * f + neg[d] * bk; f = f * omega; bk = bk * inv; }
  }
  std::vector<E8> fold_w(blocks);
  fold_w[0] = e8_one();
  E8 oinv = e8_inv(omega);
  for (size_t b = 1; b < blocks; b++) fold_w[b] = fold_w[b - 1] * oinv;
  Lanes coefs;
  for (int img = 0; img < LANES; img++) {
    E8 s = e8_zero(), fw = e8_one();
    for (size_t b = 0; b < blocks; b++) { s += fw * e8_frob(high[b], (LANES - img) % LANES); fw = fw * omega; }
    coefs[img] = powers[img] * s;
  }
  Lanes images;
  for (int lane = 0; lane < LANES; lane++) {
    E8 s = e8_zero();
    for (int img = 0; img < LANES; img++) s += coefs[img] * e8_frob(basis(lane), (LANES - img) % LANES);
    images[lane] = s;
  }
  apply_weight_images(weights, images);
  std::vector<E8> values(plen, e8_zero());
  for (size_t b = 0; b < blocks; b++) {
    auto pb = packed_block(b);
    for (size_t i = 0; i < plen; i++) values[i] += pb[i] * fold_w[b];
  }
  Committed folded = packed_commit(pp, values);
  Digest froot = folded.root();
  t.bytes(froot.data(), 32);
  std::vector<std::array<E8, 2>> rounds;
  std::vector<E8> reduced;
  E8 fval, fwt;
  sumcheck(values, weights, claim, t, rounds, reduced, fval, fwt);
  check(fwt == batch_weight(outer, n_outer, reduced, coefs) && claim == fval * fwt, "batch sumcheck terminal");
  const char* bsd = "authenticated-codeword-fold-queries/v1";
  t.bytes(bsd, std::strlen(bsd));
  t.field(fval);
  Digest seed = t.h.finalize();
  // binding payload (qafold_short_batch.rs:118-160)
  const QaCode& top = pp.seq.codes.back();
  u64 R = top.codeword_len();
  u64 count = num_queries(top, pp.security_bits, R);
  std::vector<u64> sel;
  if (count != R) sel = sample_without_replacement(count, R, seed);
  Bytes binding;
  const char* bm = "QAFOLDB1";
  binding.insert(binding.end(), bm, bm + 8);
  put_u64(binding, blocks);
  auto serialize = [&](const Committed& c) {
    Bytes b;
    if (count == R) write_base_rows_all(b, c); else write_base_rows_partial(b, c, sel);
    return b;
  };
  Bytes payload;
  for (size_t b = 0; b < blocks; b++) {
    if (b == 0 || roots[b] != roots[b - 1]) {
      Committed c = packed_commit(pp, packed_block(b));
      check(c.root() == roots[b], "recomputed block root mismatch");
      payload = serialize(c);
    }
    binding.insert(binding.end(), payload.begin(), payload.end());
  }
  Bytes fp = serialize(folded);
  binding.insert(binding.end(), fp.begin(), fp.end());
  Bytes opening;
  OpeningReduction red;
  E8 actual = pcs_open_with_reduction(pp, froot, reduced, folded, opening, red);
  check(actual == fval, "native opening value");
  Bytes decide = pcs_prove_decide(pp, red);
  Bytes w;
  const char* mg = "QAPKBAT1";
  w.insert(w.end(), mg, mg + 8);
  put_e8(w, claimed_value);
  for (auto& v : ev) put_e8(w, v);
  for (auto& v : pos) put_e8(w, v);
  for (auto& v : neg) put_e8(w, v);
  w.insert(w.end(), froot.begin(), froot.end());
  for (auto& r : rounds) { put_e8(w, r[0]); put_e8(w, r[1]); }
  put_e8(w, fval);
  put_part(w, binding);
  put_part(w, opening);
  put_part(w, decide);
  return w;
}

// packed.rs:294-393
Bytes packed_open_single(const PcsParams& pp, int log_n, const Committed& data, const std::vector<E8>& point,
                         const Bytes& context, E8& claimed_value) {
  check(point.size() == size_t(log_n), "packed point length");
  size_t n_outer = log_n - LOG_LANES;
  // The packed witness is recovered from the Mobius coefficients by the zeta transform.
  std::vector<E8> values = data.coeffs;
  for (size_t h = 1; h < values.size(); h *= 2)
    for (size_t b = 0; b < values.size(); b += 2 * h)
      for (size_t i = b; i < b + h; i++) values[i + h] += values[i];
  auto weights = equality_table(point.data(), n_outer);
  auto lanes = project_lanes(values, weights);
  auto low = equality_table(point.data() + n_outer, LOG_LANES);
  claimed_value = e8_zero();
  for (int j = 0; j < LANES; j++) claimed_value += lanes[j] * low[j];
  Lanes ev;
  for (int i = 0; i < LANES; i++) {
    E8 s = e8_zero();
    for (int j = 0; j < LANES; j++) s += basis(j) * e8_frob(lanes[j], (LANES - i) % LANES);
    ev[i] = s;
  }
  check(projected_claim(ev, low) == claimed_value, "projected claim");
  Digest root = data.root();
  PackedTranscript t = packed_transcript(pp, log_n, root, point, claimed_value, context);
  for (auto& v : ev) t.field(v);
  auto powers = lane_powers(t.challenge());
  E8 claim = e8_zero();
  for (int i = 0; i < LANES; i++) claim += powers[i] * ev[i];
  Lanes images;
  for (int j = 0; j < LANES; j++) {
    E8 s = e8_zero();
    for (int i = 0; i < LANES; i++) s += powers[i] * e8_frob(basis(j), (LANES - i) % LANES);
    images[j] = s;
  }
  apply_weight_images(weights, images);
  std::vector<std::array<E8, 2>> rounds;
  std::vector<E8> reduced;
  E8 fval, fwt;
  sumcheck(values, weights, claim, t, rounds, reduced, fval, fwt);
  check(fwt == batch_weight(point.data(), n_outer, reduced, powers) && claim == fval * fwt, "sumcheck terminal");
  Bytes opening;
  OpeningReduction red;
  E8 actual = pcs_open_with_reduction(pp, root, reduced, data, opening, red);
  check(actual == fval, "native opening value");
  Bytes decide = pcs_prove_decide(pp, red);
  Bytes w;
  w.insert(w.end(), PACKED_MAGIC, PACKED_MAGIC + 8);
  put_e8(w, claimed_value);
  for (auto& v : ev) put_e8(w, v);
  put_u64(w, rounds.size());
  for (auto& r : rounds) { put_e8(w, r[0]); put_e8(w, r[1]); }
  put_e8(w, fval);
  put_part(w, opening);
  put_part(w, decide);
  return w;
}

// ================================================================ ParComm prove (protocol.rs:186-263)
Bytes parcomm_prove(const Layout& l, const Config& cfg, const PcsParams& pp, int chunk_log, const Statement& st,
                    const Val* values, u64 n_values, const std::vector<Claim>& claims) {
  OuterTranscript tr;
  tr.object("collection", json_statement(st));
  tr.object("claims", json_claims(claims));
  Query q = tbv_compile(l, claims, tr);
  Terminal term;
  u64 stride = u64(1) << cfg.packed_log;
  EoProof eo = eo_prove(values, n_values, l.logical, stride, q, tr, term);
  Bytes context;
  context.insert(context.end(), TERMINAL_DOMAIN, TERMINAL_DOMAIN + std::strlen(TERMINAL_DOMAIN));
  context.insert(context.end(), tr.state.begin(), tr.state.end());
  put_u64(context, 0);
  u64 block_len = u64(1) << chunk_log;
  E8 actual;
  Bytes wire;
  if (st.roots.size() > 1) {
    BlockSource src = [&](u64 b, u64 i) -> Val { u64 k = b * block_len + i; return k < n_values ? values[k] : 0; };
    wire = packed_open_batch(pp, chunk_log, st.roots, term.point_msb, context, src, actual);
  } else {
    std::vector<Val> blk(block_len, 0);
    std::memcpy(blk.data(), values, std::min<u64>(n_values, block_len) * sizeof(Val));
    Committed c = packed_commit(pp, pack_block(blk.data(), block_len, block_len));
    wire = packed_open_single(pp, chunk_log, c, term.point_msb, context, actual);
  }
  check(actual * term.scale == term.value, "QAFold/EO terminal mismatch");
  // bincode fixint (protocol.rs:57-66, 89-95)
  Bytes out;
  put_u64(out, eo.positive.size());
  for (auto& v : eo.positive) put_e8(out, v);
  put_u64(out, eo.negative.size());
  for (auto& v : eo.negative) put_e8(out, v);
  put_u64(out, eo.rounds.size());
  for (auto& r : eo.rounds) { put_e8(out, r[0]); put_e8(out, r[1]); }
  put_e8(out, eo.value);
  put_e8(out, eo.query);
  put_u64(out, 1);
  put_e8(out, actual);
  put_u64(out, wire.size());
  out.insert(out.end(), wire.begin(), wire.end());
  return out;
}

}  // namespace orc
