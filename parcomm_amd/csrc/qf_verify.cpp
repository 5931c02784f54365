// The witness-free verifier: Parameters::verify (protocol.rs:266-319) with
// eo::verify (eo.rs:182-220), PackedPcs::verify / verify_batch
// (packed.rs:396-458, packed_batch.rs:321-423), the codeword-fold binding
// (qafold_short_batch.rs:162-260), the opening reduction
// (qafold_short_pc.rs:732-850, 1528-1760, 1860-1905, 1974-2010) and the
// direct-interleaved Decide (qafold_short_decide_driver.rs:320-518 with the
// helpers of qafold_short_decide*.rs). It reads only the parameters, the
// public roots, the claims and the proof bytes; all of it runs on the host.
#include <thread>
#include "qf_protocol.hpp"

namespace qf {
namespace {

struct Reject : QfError {
  explicit Reject(const std::string& m) : QfError(QF_ERR_INVALID_COMMITMENT, m) {}
};
void check(bool ok, const char* what) { if (!ok) throw Reject(what); }

// ---------------------------------------------------------------- wire reader
struct Reader {
  const uint8_t* p;
  size_t n, off = 0;
  Reader(const uint8_t* d, size_t len) : p(d), n(len) {}
  const uint8_t* take(size_t k) {
    if (k > n - off) throw QfError(QF_ERR_ENCODING, "truncated proof");
    const uint8_t* r = p + off;
    off += k;
    return r;
  }
  uint64_t u64() {
    const uint8_t* b = take(8);
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= uint64_t(b[i]) << (8 * i);
    return v;
  }
  uint64_t len(uint64_t max_items) {
    uint64_t v = u64();
    if (v > max_items) throw QfError(QF_ERR_ENCODING, "length out of range");
    return v;
  }
  uint8_t u8() { return *take(1); }
  E8 field() {
    E8 v;
    if (!e8_from_bytes(take(16), v)) throw QfError(QF_ERR_ENCODING, "non-canonical field element");
    return v;
  }
  Digest digest() { Digest d; std::memcpy(d.data(), take(32), 32); return d; }
  bool tag(const char* magic) { return std::memcmp(take(8), magic, 8) == 0; }
  Reader part() {
    uint64_t k = len(n - off);
    const uint8_t* b = take(k);
    return Reader(b, k);
  }
  bool done() const { return off == n; }
  size_t remaining() const { return n - off; }
};

// A row payload (qafold_short_pc.rs:3105-3212): every entry, or sampled rows with paths.
struct RowsPayload {
  bool all = false;
  uint64_t R = 0, cols = 0;
  std::vector<E8> entries;                  // all rows, column-major
  std::vector<std::vector<E8>> rows;        // partial: one row per query
  std::vector<std::vector<Digest>> paths;
  E8 entry(size_t query, uint64_t row, uint64_t col) const { return all ? entries[col * R + row] : rows[query][col]; }
};
RowsPayload read_rows(Reader& r) {
  RowsPayload o;
  uint8_t kind = r.u8();
  if (kind == 1) {
    o.all = true;
    o.R = r.len(uint64_t(1) << 40);
    o.cols = r.len(1 << 20);
    uint64_t count = r.len(r.remaining() / 16);
    check(count == o.R * o.cols, "all-rows payload size");
    for (uint64_t i = 0; i < count; i++) o.entries.push_back(r.field());
  } else {
    check(kind == 0, "row payload kind");
    uint64_t q = r.len(r.remaining());
    for (uint64_t i = 0; i < q; i++) {
      uint64_t c = r.len(r.remaining() / 16);
      std::vector<E8> row;
      for (uint64_t j = 0; j < c; j++) row.push_back(r.field());
      uint64_t d = r.len(64);
      std::vector<Digest> path;
      for (uint64_t j = 0; j < d; j++) path.push_back(r.digest());
      o.rows.push_back(std::move(row));
      o.paths.push_back(std::move(path));
    }
  }
  return o;
}

// ---------------------------------------------------------------- Merkle
Digest hash_row(const E8* row, uint64_t cols, const char* domain) {  // qafold_short_pc.rs:2437-2475
  Hasher h;
  h.frame("domain", domain);
  h.frame_u64("width", cols);
  h.update(row, cols * 16);
  return h.finish();
}
bool verify_siblings(Digest d, uint64_t index, uint64_t leaves, const std::vector<Digest>& sib, const Digest& root) {
  if (!pow2(leaves) || index >= leaves || sib.size() != size_t(ilog2u(leaves))) return false;
  for (auto& s : sib) {
    d = (index & 1) ? hash_inner(s, d) : hash_inner(d, s);
    index >>= 1;
  }
  return d == root;
}
// Row hashing and Merkle paths, the bulk of a proof's verification, are
// independent per row: they run on the host's threads. A batch verification,
// which already runs one proof per thread, turns this off on its threads.
thread_local bool t_rows_parallel = true;
template <class F>
void for_rows(uint64_t n, F&& body) {  // body(first, last)
  const unsigned hw = std::thread::hardware_concurrency();
  const uint64_t workers = std::min<uint64_t>(hw ? hw : 1, 16);
  if (!t_rows_parallel || n < 32 || workers <= 1) { body(uint64_t(0), n); return; }
  std::vector<std::thread> ts;
  const uint64_t chunk = (n + workers - 1) / workers;
  for (uint64_t w = 0; w < workers; w++) {
    const uint64_t a = w * chunk, b = std::min(n, a + chunk);
    if (a < b) ts.emplace_back([&body, a, b] { body(a, b); });
  }
  for (auto& t : ts) t.join();
}
Digest tree_root(const std::vector<E8>& entries, uint64_t R, uint64_t cols, const char* domain) {
  std::vector<Digest> leaves(R);
  for_rows(R, [&](uint64_t a, uint64_t b) {
    std::vector<E8> row(cols);
    for (uint64_t r = a; r < b; r++) {
      for (uint64_t c = 0; c < cols; c++) row[c] = entries[c * R + r];
      leaves[r] = hash_row(row.data(), cols, domain);
    }
  });
  return merkle_levels(std::move(leaves)).back()[0];
}
// Authenticates a payload against `root` for the selected rows; the first
// failing row (in row order) names the rejection.
void authenticate(const RowsPayload& pl, const Sel& sel, const std::vector<uint64_t>& rows, uint64_t R, uint64_t cols,
                  const Digest& root, const char* domain) {
  if (sel.all) {
    check(pl.all && pl.R == R && pl.cols == cols && tree_root(pl.entries, R, cols, domain) == root, "all-rows payload");
    return;
  }
  check(!pl.all && pl.rows.size() == rows.size(), "partial row count");
  std::vector<char> bad(rows.size(), 0);  // 1: row width, 2: Merkle path
  for_rows(rows.size(), [&](uint64_t a, uint64_t b) {
    for (uint64_t k = a; k < b; k++) {
      if (pl.rows[k].size() != cols) { bad[k] = 1; continue; }
      if (!verify_siblings(hash_row(pl.rows[k].data(), cols, domain), rows[k], R, pl.paths[k], root)) bad[k] = 2;
    }
  });
  for (char v : bad) {
    check(v != 1, "row width");
    check(v != 2, "Merkle path");
  }
}

const char* kBaseRow = "qafold-short/base-row/v1";
const char* kExtRow = "qafold-short/extension-row/v1";

// ---------------------------------------------------------------- PCS transcript reader
struct PcsReader {
  Pcs tr;
  std::vector<E8> resp;
  size_t next = 0;
  E8 read() {
    check(next < resp.size(), "transcript responses exhausted");
    E8 v = resp[next++];
    tr.ext("response", v);
    return v;
  }
  bool exhausted() const { return next == resp.size(); }
};

struct LB { E8 a_value, b_value; std::vector<E8> a_point, b_point; };
LB verify_litebullet(E8 claimed, uint64_t length, PcsReader& t, const char* label) {  // qafold_short_pc.rs:1974-2010
  check(pow2(length), "litebullet length");
  t.tr.bytes("litebullet-domain", label);
  LB lb;
  E8 rhs = claimed;
  for (uint64_t n = length; n > 1; n /= 2) {
    E8 lc = t.read(), rc = t.read();
    E8 ch = t.tr.challenge("litebullet-fold", true), ci = e8_inv(ch);
    rhs += lc * ch + rc * ci;
    lb.a_point.push_back(ci);
    lb.b_point.push_back(ch);
  }
  lb.a_value = t.read();
  lb.b_value = t.read();
  check(lb.a_value * lb.b_value == rhs, "litebullet terminal");
  return lb;
}

// MLE of the generator G (codeword x message) at (left, right) (qafold_short.rs:1166-1184)
E8 generator_mle(const QaCode& code, const std::vector<E8>& left, const std::vector<E8>& right) {
  check(left.size() == size_t(ilog2u(code.codeword_len())) && right.size() == size_t(ilog2u(code.message_len)),
        "generator point size");
  auto rw = tw_host(right);
  std::vector<E8> proj(code.codeword_len());
  host_encode(code, rw.data(), proj.data());
  return inner_host(tw_host(left), proj);
}

struct Pending { uint64_t level; E8 value; std::vector<E8> left, right; };
struct Round { E8 next_value, gen_value; std::vector<E8> next_point, gen_left, gen_right; };

// qafold_short_pc.rs:1528-1760 (base and extension rounds share the checks; only the leaf domain differs)
Round verify_round(const ParamsState& p, const QaCode& code, const Digest& root, const std::vector<E8>& point, E8 value,
                   bool next_is_root, const Digest& next_binding, const RowsPayload& rows, const char* domain,
                   PcsReader& t) {
  const uint64_t n = code.message_len, R = code.codeword_len(), cols = 32;
  check(point.size() == size_t(5 + ilog2u(n)), "round point size");
  t.tr.digest("round-root", root);
  std::vector<E8> opening;
  for (uint64_t i = 0; i < cols; i++) opening.push_back(t.read());
  check(inner_host(opening, tw_host(point.data(), 5)) == value, "round opening value");
  std::vector<E8> fr;
  for (uint64_t k = 0; k < cols; k++) fr.push_back(t.tr.challenge("freivalds"));
  E8 fv = inner_host(opening, fr);
  t.tr.bytes("next-layer-kind", next_is_root ? "extension-root" : "terminal-vector");
  t.tr.digest("next-layer-binding", next_binding);
  uint64_t count = num_queries(code, p.pp.security_bits, R);
  Digest seed = t.tr.challenge_digest("row-selection");
  Sel sel{count == R, count, {}};
  std::vector<uint64_t> srows = sample_rows(count, R, seed);
  authenticate(rows, sel, srows, R, cols, root, domain);
  std::vector<E8> sampled;
  for (size_t k = 0; k < srows.size(); k++) {
    E8 s{};
    for (uint64_t c = 0; c < cols; c++) s += rows.entry(k, srows[k], c) * fr[c];
    sampled.push_back(s);
  }
  t.tr.exts("sampled-values", sampled);
  // qafold_short_pc.rs:1660-1725
  uint64_t sp = 1;
  while (sp < srows.size()) sp <<= 1;
  std::vector<E8> slp;
  for (int i = 0; i < ilog2u(sp); i++) slp.push_back(t.tr.challenge("selection-left"));
  std::vector<E8> sw(srows.size());
  for (size_t i = 0; i < srows.size(); i++) sw[i] = weight_at(slp, i);
  E8 selected = inner_host(sw, sampled);
  E8 rho = t.tr.challenge("opening-batch");
  LB msg = verify_litebullet(selected + rho * fv, n, t, "message-litebullet");
  std::vector<E8> row_point(point.begin() + 5, point.end());
  E8 sle = msg.a_value - rho * tensor_inner(row_point, msg.a_point);
  LB gen = verify_litebullet(sle, R, t, "generator-litebullet");
  E8 sparse{};
  for (size_t k = 0; k < srows.size(); k++) sparse += sw[k] * weight_at(gen.a_point, srows[k]);
  check(gen.a_value == sparse, "generator sparse value");
  return {msg.b_value, gen.b_value, msg.b_point, gen.b_point, msg.a_point};
}

// qafold_short_pc.rs:732-850; returns the pending generator claims for Decide.
std::vector<Pending> verify_opening_reduction(const ParamsState& p, const Digest& root, const std::vector<E8>& point,
                                              E8 value, Reader r) {
  check(r.tag("QFSHORT2"), "opening magic");
  PcsReader t;
  uint64_t nr = r.len(r.remaining() / 16);
  for (uint64_t i = 0; i < nr; i++) t.resp.push_back(r.field());
  RowsPayload top_rows = read_rows(r);
  const size_t levels = p.pp.codes.size(), top = levels - 1;
  uint64_t n_ext = r.len(64);
  check(n_ext == top, "extension round count");
  std::vector<std::pair<Digest, RowsPayload>> ext;
  for (uint64_t i = 0; i < n_ext; i++) {
    Digest d = r.digest();
    ext.push_back({d, read_rows(r)});
  }
  uint64_t nb = r.len(r.remaining() / 16);
  std::vector<E8> base_vector;
  for (uint64_t i = 0; i < nb; i++) base_vector.push_back(r.field());
  check(r.done(), "trailing opening bytes");
  check(point.size() == size_t(5 + ilog2u(p.pp.codes[top].message_len)), "opening point size");
  t.tr.digest("parameters", p.pp.parameter_digest);
  t.tr.digest("commitment", root);
  t.tr.exts("point", point);
  t.tr.ext("claimed-value", value);
  std::vector<Pending> pending;
  auto binding = [&](size_t idx) -> std::pair<bool, Digest> {
    if (idx < ext.size()) return {true, ext[idx].first};
    return {false, hash_terminal(base_vector)};
  };
  auto [is_root, bind] = binding(0);
  Round ro = verify_round(p, p.pp.codes[top], root, point, value, is_root, bind, top_rows, kBaseRow, t);
  auto settle = [&](size_t level, const Round& o) {
    if (level == 0) check(generator_mle(p.pp.codes[0], o.gen_left, o.gen_right) == o.gen_value, "level-0 generator");
    else pending.push_back({level, o.gen_value, o.gen_left, o.gen_right});
  };
  settle(top, ro);
  for (size_t i = 0; i < ext.size(); i++) {
    size_t level = top - 1 - i;
    auto [nr2, nb2] = binding(i + 1);
    ro = verify_round(p, p.pp.codes[level], ext[i].first, ro.next_point, ro.next_value, nr2, nb2, ext[i].second, kExtRow,
                      t);
    settle(level, ro);
  }
  check(!base_vector.empty() && pow2(base_vector.size()) && ro.next_point.size() == size_t(ilog2u(base_vector.size())) &&
            inner_host(base_vector, tw_host(ro.next_point)) == ro.next_value,
        "terminal vector");
  check(t.exhausted(), "unused opening responses");
  return pending;
}

// qafold_short_decide_driver.rs:320-518 (direct-interleaved setup)
void verify_decide(const ParamsState& p, const std::vector<Pending>& pending, Reader r) {
  const auto& codes = p.pp.codes;
  const size_t levels = codes.size();
  check(r.tag("QFSDCID1"), "decide magic");
  PcsReader t;
  uint64_t nr = r.len(r.remaining() / 16);
  for (uint64_t i = 0; i < nr; i++) t.resp.push_back(r.field());
  uint64_t n_rounds = r.len(64);
  Pcs& tr = t.tr;
  tr.bytes("decide-domain", "qafold-short/decide/v1");
  tr.digest("decide-parameters", p.pp.parameter_digest);
  tr.digest("decide-setup", p.pp.setup_digest);
  tr.u64("decide-levels", levels);
  uint8_t one = 1, zero = 0;
  tr.bytes("opening-reduction-ok", &one, 1);
  tr.u64("pending-h-count", pending.size());
  for (auto& q : pending) {
    tr.u64("pending-h-level", q.level);
    tr.ext("pending-g-value", q.value);
    tr.exts("pending-g-left", q.left);
    tr.exts("pending-g-right", q.right);
  }
  std::vector<std::vector<MClaim>> h(levels);
  for (auto& q : pending) {
    check(q.level < levels && q.left.size() == size_t(ilog2u(codes[q.level].codeword_len())) &&
              q.right.size() == size_t(ilog2u(codes[q.level].message_len)),
          "pending claim shape");
    h[q.level].push_back(reduce_generator(codes[q.level].message_len, {q.value, q.left, q.right}));
  }
  if (levels <= 1) {
    for (auto& v : h) check(v.empty(), "pending claims without Decide levels");
    uint64_t nb = r.len(r.remaining() / 16);
    check(n_rounds == 0 && nb == 0 && r.done() && t.exhausted(), "single-level Decide");
    return;
  }
  check(n_rounds == levels - 1, "Decide round count");
  const uint64_t rf = 32;
  bool have_w = false, have_carry = false;
  E8 w_value{};
  std::vector<E8> w_point;
  Digest carry_root{};
  std::vector<E8> base_vector_checked;
  bool base_done = false;
  for (size_t level = levels - 1; level >= 1; level--) {
    const SetupLevel& sl = p.pp.setup[level];
    const SetupShape& sh = sl.shape;
    const QaCode& code = codes[level];
    const QaCode& target = codes[level - 1];
    const uint64_t rows = code.message_len, log_rows = ilog2u(rows), n_enc = target.codeword_len();
    check(sh.encoded_rows == n_enc, "setup shape");
    // ---- round wire
    check(r.u64() == level, "Decide round level");
    bool has_next = r.u8() != 0;
    Digest next_root{};
    if (has_next) next_root = r.digest();
    check(has_next == (level > 1), "next carry presence");
    uint8_t setup_kind = r.u8();
    const uint64_t scols = sh.t_cols + sh.h_cols;
    std::vector<uint64_t> setup_idx;
    std::vector<std::vector<E8>> setup_rows;
    std::vector<std::vector<Digest>> setup_paths;
    bool setup_all = false;
    std::vector<E8> setup_entries;
    if (setup_kind == 1) {
      setup_all = true;
      uint64_t f[7];
      for (auto& v : f) v = r.u64();
      check(f[0] == sh.level && f[1] == sh.encoded_rows && f[2] == sh.message_rows && f[3] == sh.t_cols &&
                f[4] == sh.t_elements && f[5] == sh.t_last_len && f[6] == sh.h_cols,
            "setup shape in proof");
      uint64_t cnt = r.len(r.remaining() / 16);
      for (uint64_t i = 0; i < cnt; i++) setup_entries.push_back(r.field());
    } else {
      check(setup_kind == 0, "setup row kind");
      uint64_t q = r.len(r.remaining());
      for (uint64_t i = 0; i < q; i++) {
        setup_idx.push_back(r.u64());
        uint64_t c = r.len(r.remaining() / 16);
        std::vector<E8> row;
        for (uint64_t j = 0; j < c; j++) row.push_back(r.field());
        uint64_t d = r.len(64);
        std::vector<Digest> path;
        for (uint64_t j = 0; j < d; j++) path.push_back(r.digest());
        setup_rows.push_back(std::move(row));
        setup_paths.push_back(std::move(path));
      }
    }
    bool carry_in_proof = r.u8() != 0;
    check(carry_in_proof == have_carry, "carry rows presence");
    RowsPayload carry_rows;
    if (carry_in_proof) carry_rows = read_rows(r);
    // ---- claims (qafold_short_decide_protocol.rs:260-360)
    struct EC { E8 value; std::vector<E8> row_point, col_weight; };
    std::vector<EC> claims;
    for (auto& hc : h[level]) {
      check(hc.left.size() == log_rows && hc.right.size() == 1, "alpha claim shape");
      claims.push_back({hc.value, hc.left, tw_host(hc.right)});
    }
    h[level].clear();
    const uint64_t setup_cols = 2, carry_cols = have_w ? 1 : 0;
    check(have_w == have_carry, "carry commitment / claim");
    if (have_w) {
      check(w_point.size() == log_rows, "carry claim shape");
      std::vector<E8> cwt(setup_cols, E8{});
      cwt.push_back(e8_one());
      claims.push_back({w_value, w_point, cwt});
    }
    check(!claims.empty(), "no Decide claims");
    uint64_t real = setup_cols + carry_cols, padded = 1;
    while (padded < real) padded <<= 1;
    const uint64_t log_cols = ilog2u(padded);
    tr.bytes("decide-round-domain", "qafold-short/decide-round/v1");
    tr.u64("decide-round-level", level);
    tr.u64("decide-round-shape-level", sh.level);
    tr.u64("decide-round-shape-encoded-rows", sh.encoded_rows);
    tr.u64("decide-round-shape-message-rows", sh.message_rows);
    tr.u64("decide-round-shape-t-cols", sh.t_cols);
    tr.u64("decide-round-shape-t-elements", sh.t_elements);
    tr.u64("decide-round-shape-t-last-len", sh.t_last_len);
    tr.u64("decide-round-shape-h-cols", sh.h_cols);
    const Digest& setup_root = sl.tree.back()[0];
    tr.digest("decide-round-setup-root", setup_root);
    tr.bytes("decide-round-carry-present", have_carry ? &one : &zero, 1);
    if (have_carry) tr.digest("decide-round-carry-root", carry_root);
    // ---- multi-evaluation (qafold_short_decide.rs:553-584)
    tr.bytes("multi-eval-domain", "qafold-short-decide/multi-eval/v2");
    tr.u64("multi-eval-log-rows", log_rows);
    tr.u64("multi-eval-log-cols", log_cols);
    tr.u64("multi-eval-claim-count", claims.size());
    for (auto& cl : claims) {
      tr.ext("multi-eval-claim-value", cl.value);
      tr.exts("multi-eval-row-point", cl.row_point);
      tr.exts("multi-eval-column-weight", cl.col_weight);
    }
    E8 rho = tr.challenge("multi-eval-batch");
    E8 rhs{}, pw = e8_one();
    std::vector<E8> powers;
    for (auto& cl : claims) { powers.push_back(pw); rhs += pw * cl.value; pw = pw * rho; }
    std::vector<E8> a_point, b_point;
    for (uint64_t i = 0; i < log_rows + log_cols; i++) {
      E8 lc = t.read(), rc = t.read();
      E8 ch = tr.challenge("multi-eval-fold", true), ci = e8_inv(ch);
      rhs += lc * ch + rc * ci;
      a_point.push_back(ci);
      b_point.push_back(ch);
    }
    E8 red_value = t.read(), bval = t.read();
    {
      std::vector<E8> bcol(b_point.begin(), b_point.begin() + log_cols), brow(b_point.begin() + log_cols, b_point.end());
      auto ct = tw_host(bcol);
      E8 expected{};
      for (size_t i = 0; i < claims.size(); i++) {
        E8 ce{};
        for (size_t k = 0; k < claims[i].col_weight.size(); k++) ce += claims[i].col_weight[k] * ct[k];
        expected += powers[i] * ce * tensor_inner(claims[i].row_point, brow);
      }
      check(red_value * bval == rhs && bval == expected, "multi-evaluation");
    }
    std::vector<E8> col_point(a_point.begin(), a_point.begin() + log_cols);
    std::vector<E8> merged(a_point.begin() + log_cols, a_point.end());
    auto col_w = tw_host(col_point);
    // ---- row selection and iota (qafold_short_decide_carry.rs:382-440)
    uint64_t count = num_queries(target, p.pp.security_bits, n_enc);
    Digest seed = tr.challenge_digest("decide-row-selection");
    Sel sel{count == n_enc, count, {}};
    std::vector<uint64_t> srows = sample_rows(count, n_enc, seed);
    if (!sel.all) sel.rows = srows;
    auto bw = tw_host(merged.data(), 5);
    uint64_t phys = real * rf, pphys = 1;
    while (pphys < phys) pphys <<= 1;
    std::vector<E8> iw(pphys, E8{});
    for (uint64_t lc = 0; lc < real; lc++)
      for (uint64_t b = 0; b < rf; b++) iw[lc * rf + b] = col_w[lc] * bw[b];
    const uint64_t phys_carry = carry_cols * rf;
    // setup rows: the public setup fixes them (and their Merkle paths under the setup root)
    auto setup_entry = [&](size_t k, uint64_t col) -> E8 {
      return setup_all ? setup_entries[col * n_enc + srows[k]] : setup_rows[k][col];
    };
    if (setup_all) {
      check(sel.all && setup_entries == sl.encoded, "setup rows (all)");
    } else {
      check(!sel.all && setup_rows.size() == srows.size(), "setup row count");
      for (size_t k = 0; k < srows.size(); k++) {
        check(setup_idx[k] == srows[k] && setup_rows[k].size() == scols, "setup row index");
        for (uint64_t col = 0; col < scols; col++)
          check(setup_rows[k][col] == sl.encoded[col * n_enc + srows[k]], "setup row entries");
        std::vector<Digest> want;
        uint64_t idx = srows[k];
        for (size_t l = 0; l + 1 < sl.tree.size(); l++) { want.push_back(sl.tree[l][idx ^ 1]); idx >>= 1; }
        check(setup_paths[k] == want, "setup row path");
      }
    }
    if (have_carry) authenticate(carry_rows, sel, srows, n_enc, phys_carry, carry_root, kExtRow);
    tr.bytes("iota-domain", "qafold-short/decide-iota/v1");
    tr.u64("iota-level", level);
    tr.u64("iota-shape-level", sh.level);
    tr.u64("iota-shape-encoded-rows", sh.encoded_rows);
    tr.u64("iota-shape-message-rows", sh.message_rows);
    tr.u64("iota-shape-t-cols", sh.t_cols);
    tr.u64("iota-shape-t-elements", sh.t_elements);
    tr.u64("iota-shape-t-last-len", sh.t_last_len);
    tr.u64("iota-shape-h-cols", sh.h_cols);
    tr.digest("iota-setup-root", setup_root);
    tr.bytes("iota-carry-present", have_carry ? &one : &zero, 1);
    if (have_carry) tr.digest("iota-carry-root", carry_root);
    tr.u64("iota-carry-cols", phys_carry);
    absorb_sel(tr, sel);
    tr.exts("iota-column-weights", iw);
    std::vector<E8> iota;
    for (size_t k = 0; k < srows.size(); k++) {
      E8 v{};
      for (uint64_t col = 0; col < scols; col++) v += iw[col] * setup_entry(k, col);
      if (have_carry)
        for (uint64_t col = 0; col < phys_carry; col++) v += iw[scols + col] * carry_rows.entry(k, srows[k], col);
      iota.push_back(v);
    }
    for (auto& v : iota) check(t.read() == v, "iota response");
    tr.bytes("decide-next-carry-domain", "qafold-short/decide-next-carry/v1");
    tr.bytes("decide-next-carry-present", has_next ? &one : &zero, 1);
    if (has_next) tr.digest("decide-next-carry-root", next_root);
    // ---- direct code switch (qafold_short_decide_carry.rs:1029-1094)
    const uint64_t trows = target.message_len;
    check(merged.size() == size_t(ilog2u(trows * rf)), "merged point size");
    std::vector<E8> base_point(merged.begin() + 5, merged.end());
    tr.bytes("direct-code-switch-domain", "qafold-short/decide-direct-code-switch/v2");
    tr.u64("direct-code-switch-target-level", level - 1);
    tr.u64("direct-code-switch-row-factor", rf);
    tr.u64("direct-code-switch-message-len", trows);
    tr.u64("direct-code-switch-encoded-len", n_enc);
    tr.u64("direct-code-switch-target-index", target.index);
    tr.digest("direct-code-switch-target-generator", target.generator_digest);
    absorb_sel(tr, sel);
    tr.exts("direct-code-switch-iota", iota);
    tr.ext("direct-code-switch-merged-value", red_value);
    tr.exts("direct-code-switch-merged-left-point", merged);
    uint64_t sp = 1;
    while (sp < srows.size()) sp <<= 1;
    std::vector<E8> slp;
    for (int i = 0; i < ilog2u(sp); i++) slp.push_back(tr.challenge("direct-code-switch-selection-left"));
    std::vector<E8> sw(srows.size());
    for (size_t i = 0; i < srows.size(); i++) sw[i] = weight_at(slp, i);
    E8 selected = inner_host(sw, iota);
    E8 rho2 = tr.challenge("direct-code-switch-opening-batch");
    LB msg = verify_litebullet(selected + rho2 * red_value, trows, t, "decide-direct-carry-message-litebullet");
    E8 sle = msg.a_value - rho2 * tensor_inner(base_point, msg.a_point);
    LB gen = verify_litebullet(sle, n_enc, t, "decide-direct-carry-generator-litebullet");
    E8 sparse{};
    for (size_t k = 0; k < srows.size(); k++) sparse += sw[k] * weight_at(gen.a_point, srows[k]);
    check(gen.a_value == sparse, "code-switch sparse value");
    if (level > 1) {
      h[level - 1].push_back(reduce_generator(trows, {gen.b_value, gen.b_point, msg.a_point}));
      w_value = msg.b_value;
      w_point = msg.b_point;
      have_w = true;
      carry_root = next_root;
      have_carry = true;
    } else {
      check(generator_mle(target, gen.b_point, msg.a_point) == gen.b_value, "terminal generator claim");
      uint64_t nb = r.len(r.remaining() / 16);
      check(nb == trows, "base vector length");
      std::vector<E8> base;
      for (uint64_t i = 0; i < nb; i++) base.push_back(r.field());
      for (auto& v : base) check(t.read() == v, "base vector response");
      check(inner_host(base, tw_host(msg.b_point)) == msg.b_value, "base vector evaluation");
      have_carry = false;
      base_done = true;
    }
  }
  for (auto& v : h) check(v.empty(), "undischarged generator claims");
  check(base_done && !have_carry && r.done() && t.exhausted(), "Decide transcript");
}

// packed.rs:396-458 (one block)
void verify_single(const ParamsState& p, const Digest& root, const std::vector<E8>& point, E8 claimed, Reader r,
                   const Bytes& context) {
  const size_t n_outer = size_t(p.chunk_log) - LOG_LANES;
  check(point.size() == size_t(p.chunk_log), "single opening point size");
  check(r.tag(PACKED_MAGIC) && r.field() == claimed, "single opening header");
  Lanes ev;
  for (auto& v : ev) v = r.field();
  auto low = eq_host(point.data() + n_outer, LOG_LANES);
  check(projected_claim(ev, low) == claimed, "projected claim");
  Packed t = packed_transcript(p, root, point, claimed, context);
  for (auto& v : ev) t.field(v);
  auto powers = lane_powers(t.challenge());
  E8 claim{};
  for (int i = 0; i < LANES; i++) claim += powers[i] * ev[i];
  check(r.u64() == n_outer, "sumcheck round count");
  std::vector<E8> reduced;
  for (size_t i = 0; i < n_outer; i++) {
    E8 m0 = r.field(), m1 = r.field();
    t.field(m0);
    t.field(m1);
    E8 c = t.challenge();
    claim = m0 + c * ((claim - m0 - m0 - m1) + c * m1);
    reduced.push_back(c);
  }
  E8 fval = r.field();
  check(claim == fval * batch_weight(point.data(), n_outer, reduced, powers), "sumcheck terminal");
  Reader opening = r.part(), decide = r.part();
  check(r.done(), "trailing single-opening bytes");
  auto pending = verify_opening_reduction(p, root, reduced, fval, opening);
  verify_decide(p, pending, decide);
}

// packed_batch.rs:321-423 with qafold_short_batch.rs:162-260
void verify_batch(const ParamsState& p, const std::vector<Digest>& roots, const std::vector<E8>& point, E8 claimed,
                  Reader r, const Bytes& context) {
  const uint64_t blocks = roots.size();
  check(blocks >= 2 && blocks <= 128 && pow2(blocks), "block count");
  const size_t split = ilog2u(blocks), n_outer = size_t(p.chunk_log) - LOG_LANES;
  check(point.size() == split + n_outer + LOG_LANES, "batch point size");
  std::vector<E8> outer(point.begin() + split, point.begin() + split + n_outer);
  auto low = eq_host(point.data() + point.size() - LOG_LANES, LOG_LANES);
  auto high = eq_host(point.data(), split);
  check(r.tag("QAPKBAT1") && r.field() == claimed, "batch header");
  Lanes ev;
  for (auto& v : ev) v = r.field();
  check(projected_claim(ev, low) == claimed, "projected claim");
  Packed t = packed_transcript(p, roots[0], point, claimed, context);
  const char* bd = "one-opening-block-packed-eo/v1";
  t.bytes(bd, std::strlen(bd));
  uint8_t nb[8];
  for (int i = 0; i < 8; i++) nb[i] = uint8_t(blocks >> (8 * i));
  t.bytes(nb, 8);
  for (auto& rt : roots) t.bytes(rt.data(), 32);
  for (auto& v : ev) t.field(v);
  auto powers = lane_powers(t.challenge());
  E8 claim{};
  for (int i = 0; i < LANES; i++) claim += powers[i] * ev[i];
  std::vector<E8> pos, neg;
  for (uint64_t d = 1; d < blocks; d++) pos.push_back(r.field());
  for (uint64_t d = 1; d < blocks; d++) neg.push_back(r.field());
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
  std::vector<E8> fw(blocks);
  fw[0] = e8_one();
  E8 oinv = e8_inv(omega);
  for (uint64_t b = 1; b < blocks; b++) fw[b] = fw[b - 1] * oinv;
  Lanes coefs;
  for (int img = 0; img < LANES; img++) {
    E8 s{}, f = e8_one();
    for (uint64_t b = 0; b < blocks; b++) { s += f * e8_frob(high[b], (LANES - img) % LANES); f = f * omega; }
    coefs[img] = powers[img] * s;
  }
  Digest folded = r.digest();
  t.bytes(folded.data(), 32);
  std::vector<E8> reduced;
  for (size_t i = 0; i < n_outer; i++) {
    E8 m0 = r.field(), m1 = r.field();
    t.field(m0);
    t.field(m1);
    E8 c = t.challenge();
    claim = m0 + c * ((claim - m0 - m0 - m1) + c * m1);
    reduced.push_back(c);
  }
  E8 fval = r.field();
  check(claim == fval * batch_weight(outer.data(), n_outer, reduced, coefs), "batch sumcheck terminal");
  const char* bsd = "authenticated-codeword-fold-queries/v1";
  t.bytes(bsd, std::strlen(bsd));
  t.field(fval);
  Digest seed = t.h.finish();
  // binding: every block's rows and the folded rows at the same positions, folded = sum_b w_b row_b
  {
    Reader b = r.part();
    const QaCode& top = p.pp.codes.back();
    const uint64_t R = top.codeword_len(), cols = 32;
    uint64_t count = num_queries(top, p.pp.security_bits, R);
    Sel sel{count == R, count, {}};
    std::vector<uint64_t> srows = sample_rows(count, R, seed);
    check(b.tag("QAFOLDB1") && b.u64() == blocks, "binding header");
    std::vector<RowsPayload> pl;
    for (uint64_t i = 0; i <= blocks; i++) {
      pl.push_back(read_rows(b));
      authenticate(pl.back(), sel, srows, R, cols, i < blocks ? roots[i] : folded, kBaseRow);
    }
    check(b.done(), "trailing binding bytes");
    for (size_t k = 0; k < srows.size(); k++)
      for (uint64_t c = 0; c < cols; c++) {
        E8 want{};
        for (uint64_t i = 0; i < blocks; i++) want += fw[i] * pl[i].entry(k, srows[k], c);
        check(pl[blocks].entry(k, srows[k], c) == want, "codeword fold");
      }
  }
  Reader opening = r.part(), decide = r.part();
  check(r.done(), "trailing batch bytes");
  auto pending = verify_opening_reduction(p, folded, reduced, fval, opening);
  verify_decide(p, pending, decide);
}

}  // namespace

// protocol.rs:266-319
void verify(const ParamsState& p, const std::vector<Digest>& roots, const std::vector<ClaimIn>& claims,
            const uint8_t* proof, size_t len) {
  require(len <= (256ull << 20), QF_ERR_ENCODING, "oversize ParComm proof");
  require(roots.size() == p.blocks, QF_ERR_INVALID_PARAMETERS, "commitment/parameter mismatch");
  require(claims.size() == p.layout.shapes.size(), QF_ERR_INVALID_PARAMETERS, "claim count mismatch");
  for (size_t i = 0; i < claims.size(); i++) {
    uint64_t rp = 1, cp = 1;
    while (rp < p.layout.shapes[i].rows) rp <<= 1;
    while (cp < p.layout.shapes[i].cols) cp <<= 1;
    require(claims[i].row.size() == size_t(ilog2u(rp)) && claims[i].col.size() == size_t(ilog2u(cp)),
            QF_ERR_INVALID_PARAMETERS, "point dimension mismatch");
  }
  Reader r(proof, len);
  const uint64_t s = uint64_t(1) << p.packed_log;
  const uint64_t nrounds = ilog2u(p.layout.logical / s);
  std::vector<E8> pos, neg;
  uint64_t np = r.len(1 << 20);
  for (uint64_t i = 0; i < np; i++) pos.push_back(r.field());
  uint64_t nn = r.len(1 << 20);
  for (uint64_t i = 0; i < nn; i++) neg.push_back(r.field());
  uint64_t nrr = r.len(1 << 20);
  std::vector<std::array<E8, 2>> rounds(nrr);
  for (auto& rd : rounds) { rd[0] = r.field(); rd[1] = r.field(); }
  E8 value = r.field(), query = r.field();
  require(pos.size() == s - 1 && neg.size() == s - 1 && rounds.size() == nrounds, QF_ERR_INVALID_COMMITMENT,
          "EO proof dimensions");
  require(r.u64() == 1, QF_ERR_INVALID_COMMITMENT, "PCS opening count");
  E8 actual = r.field();
  Reader wire = r.part();
  require(r.done(), QF_ERR_ENCODING, "trailing proof bytes");
  // statement and claims (protocol.rs:186-200), TBV compile, EO (eo.rs:182-220)
  CommittedState st;
  st.roots = roots;
  Outer tr;
  tr.object("collection", statement_json(p, st));
  tr.object("claims", claims_json(claims));
  Query q = compile(p.layout, claims, tr);
  tr.object("packed-positive", js_scalars(pos));
  tr.object("packed-negative", js_scalars(neg));
  const int ls = ilog2u(s);
  E8 ch = tr.fold_challenge(ls);
  E8 claim = q.claim;
  {
    E8 inv = e8_inv(ch), f = ch, b = inv;
    for (uint64_t d = 0; d + 1 < s; d++) { claim += pos[d] * f + neg[d] * b; f = f * ch; b = b * inv; }
  }
  std::vector<E8> challenges;
  E8 pw = ch;
  for (int i = 0; i < ls; i++) { challenges.push_back(pw); pw = pw * pw; }
  for (auto& rd : rounds) {
    tr.object("binary-cross", js_scalars({rd[0], rd[1]}));
    E8 c = tr.fold_challenge(1);
    claim += rd[0] * c + rd[1] * e8_inv(c);
    challenges.push_back(c);
  }
  require(claim == value * query, QF_ERR_INVALID_COMMITMENT, "EO scalar relation rejected");
  require(query_eval(q, challenges) == query, QF_ERR_INVALID_COMMITMENT, "TBV terminal evaluation rejected");
  tr.scalar("terminal-value", value);
  tr.scalar("terminal-query", query);
  std::vector<E8> point;
  E8 scale = e8_one();
  for (auto& x : challenges) {
    E8 m = e8_inv(x), f = e8_one() + m;
    point.push_back(m * e8_inv(f));
    scale = scale * f;
  }
  std::reverse(point.begin(), point.end());
  Bytes context;
  const char* cx = TERMINAL_DOMAIN;
  context.insert(context.end(), cx, cx + std::strlen(cx));
  context.insert(context.end(), tr.state.begin(), tr.state.end());
  put_u64(context, 0);
  if (roots.size() > 1) verify_batch(p, roots, point, actual, wire, context);
  else verify_single(p, roots[0], point, actual, wire, context);
  require(actual * scale == value, QF_ERR_INVALID_COMMITMENT, "PCS terminal aggregate rejected");
}

void set_verify_rows_parallel(bool on) { t_rows_parallel = on; }

}  // namespace qf
