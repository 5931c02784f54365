// The opening prover: Parameters::prove (protocol.rs:201-263) over a
// device-resident commitment. Fiat-Shamir transcripts are replayed on the
// host; every vector larger than a handful of elements lives in HBM and is
// processed by the kernels of qf_vec.hip and qf_commit_kernels.hip.
#include "qf_protocol.hpp"
#include "qf_rounds.hpp"
#include "qf_vec.hpp"
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <functional>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>

namespace qf {

const E8* ParamsState::alpha_device(size_t level) const {
  std::lock_guard<std::mutex> g(mu);
  if (d_alpha_levels.size() != pp.codes.size()) d_alpha_levels.assign(pp.codes.size(), nullptr);
  if (!d_alpha_levels[level]) {
    const auto& a = pp.codes[level].alpha1;
    E8* d = nullptr;
    QF_HIP(hipMalloc(&d, a.size() * 16));
    QF_HIP(hipMemcpy(d, a.data(), a.size() * 16, hipMemcpyHostToDevice));
    d_alpha_levels[level] = d;
    if (level + 1 == pp.codes.size()) d_alpha = d;
  }
  return d_alpha_levels[level];
}
const E8* ParamsState::setup_device(size_t level) const {
  std::lock_guard<std::mutex> g(mu);
  if (d_setup_enc.size() != pp.setup.size()) d_setup_enc.assign(pp.setup.size(), nullptr);
  if (!d_setup_enc[level]) {
    const auto& e = pp.setup[level].encoded;
    E8* d = nullptr;
    QF_HIP(hipMalloc(&d, std::max<size_t>(e.size(), 1) * 16));
    if (!e.empty()) QF_HIP(hipMemcpy(d, e.data(), e.size() * 16, hipMemcpyHostToDevice));
    d_setup_enc[level] = d;
  }
  return d_setup_enc[level];
}
const E8* ParamsState::alpha_encode(size_t level) const {
  const E8* a = alpha_device(level);
  if (LANES == 8) return a;  // F65521 encodes with the canonical form
  std::lock_guard<std::mutex> g(mu);
  if (d_alpha_enc.size() != pp.codes.size()) d_alpha_enc.assign(pp.codes.size(), nullptr);
  if (!d_alpha_enc[level]) {
    const uint64_t n = pp.codes[level].alpha1.size();
    E8* d = nullptr;
    QF_HIP(hipMalloc(&d, n * 16));
    launch_encode_form(a, d, n, nullptr);
    QF_HIP(hipDeviceSynchronize());
    d_alpha_enc[level] = d;
  }
  return d_alpha_enc[level];
}

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }
// QF_TRACE_SYNC=1: host waits and their time per prover phase, and the marked
// host intervals inside it, on stderr.
void trace_phase(const char* name, double ms) {
  if (!sync_trace_on()) return;
  SyncTrace& t = sync_trace();
  std::fprintf(stderr, "[qf] %-10s %8.3f ms, %4llu host waits, %8.3f ms waiting\n", name, ms,
               (unsigned long long)t.syncs, t.wait_ms);
  t = SyncTrace{};
  trace_scopes_flush();
}

void tw_device(DevCtx& c, E8* out, const std::vector<E8>& p) {
  size_t m = p.size(), a = m / 2;
  dv::tensor(c, out, tw_host(p.data(), m - a), tw_host(p.data() + (m - a), a));
}
// eq(p), MSB first (the first variables index the high bits), as two device
// half tables: hi over the first m - m/2 variables and lo over the rest, each
// the outer product of two host quarter tables of at most m/4 + 1 variables.
void eq_tables_device(DevCtx& c, const std::vector<E8>& p, DVec& hi, DVec& lo) {
  const size_t m = p.size(), a = m / 2;
  auto half = [&](const E8* q, size_t n, DVec& out) {
    const size_t b = n / 2;
    out = DVec(uint64_t(1) << n);
    dv::tensor(c, out.p, eq_host(q, n - b), eq_host(q + (n - b), b));
  };
  half(p.data(), m - a, hi);
  half(p.data() + (m - a), a, lo);
}
Digest download_root(DevCtx& c, const uint32_t* tree, uint64_t R) {
  Digest d;
  dv::d2h(c, d.data(), tree + 8 * (2 * R - 2), 32);
  return d;
}

// ================================================================ device codewords
// A 32-column codeword (column-major, R rows) and its Merkle tree in HBM.
struct DCode {
  DVec cw;
  uint32_t* tree = nullptr;
  bool own_tree = false;
  const E8* cwp = nullptr;  // borrowed codeword when cw is empty
  uint64_t R = 0, cols = 32;
  hipStream_t st = nullptr;  // stream the owned tree was allocated on
  Digest root{};
  const E8* data() const { return cw.p ? cw.p : cwp; }
  DCode() = default;
  DCode(DCode&& o) noexcept
      : cw(std::move(o.cw)), tree(o.tree), own_tree(o.own_tree), cwp(o.cwp), R(o.R), cols(o.cols), st(o.st),
        root(o.root) {
    o.tree = nullptr;
    o.own_tree = false;
  }
  DCode& operator=(DCode&& o) noexcept {
    if (this != &o) {
      if (own_tree) dfree(tree, st);
      cw = std::move(o.cw); tree = o.tree; own_tree = o.own_tree; cwp = o.cwp; R = o.R; cols = o.cols;
      st = o.st; root = o.root;
      o.tree = nullptr; o.own_tree = false;
    }
    return *this;
  }
  // 32 column projections of the codeword (qafold_short_pc.rs:2641-2671)
  std::vector<E8> projection(DevCtx& c, uint64_t n, const E8* w) const {
    return dv::col_projection(c, data(), R, n, cols, w);
  }
  // The same projections with w = tw(point): w[r] = top[r >> 5] low[r & 31], so the
  // low five coordinates first contract 32 consecutive rows of every column on the
  // matrix cores (F65521) and the rest project the n / 32 results.
  std::vector<E8> projection_tw(DevCtx& c, uint64_t n, const std::vector<E8>& point, const E8* w) const {
#ifndef QF_KB
    const size_t m = point.size();
    if (m >= 10 && n == uint64_t(1) << m && cols <= 64) {
      const uint64_t h = n / 32;
      DVec inner(cols * h);
      if (dv::col_fold_rows32(c, data(), R, n, cols, tw_host(point.data() + (m - 5), 5), inner.p)) {
        DVec top(h);
        tw_device(c, top.p, std::vector<E8>(point.begin(), point.end() - 5));
        return dv::col_projection(c, inner.p, h, h, cols, top.p);
      }
    }
#else
    (void)point;
#endif
    return projection(c, n, w);
  }
  // out[r] = sum_col f[col] * codeword[col][r], r < n
  void fold(DevCtx& c, uint64_t n, const std::vector<E8>& f, E8* out) const {
    dv::col_fold(c, data(), R, n, cols, f, out);
  }
  // rows[k * cols + col] of the codeword
  std::vector<E8> rows(DevCtx& c, const std::vector<uint64_t>& sel) const {
    return dv::gather_rows(c, data(), R, cols, sel);
  }
  ~DCode() { if (own_tree) dfree(tree, st); }
};
DCode alloc_code(uint64_t R) {
  DCode d;
  d.R = R;
  d.cw = DVec(32 * R);
  d.st = current_stream();
  d.tree = static_cast<uint32_t*>(dmalloc((2 * R - 1) * 32, d.st));
  d.own_tree = true;
  return d;
}
// Commit of an extension matrix under `code` (qafold_short_pc.rs:2327-2362).
DCode commit_extension(DevCtx& c, const ParamsState& p, size_t level, const E8* values) {
  const QaCode& code = p.pp.codes[level];
  DCode d = alloc_code(code.codeword_len());
  launch_commit_extension(values, d.cw.p, ilog2u(code.message_len), p.alpha_encode(level), p.leaf_ext, p.geom.inner,
                          d.tree, c.st);
  QF_HIP(hipGetLastError());
  d.root = download_root(c, d.tree, d.R);
  return d;
}
// G^T applied to a codeword-length vector: left[:n] + WHT(alpha_1 . WHT(left[n:])) (qafold_short.rs:1106-1163)
DVec generator_left(DevCtx& c, const ParamsState& p, size_t level, const DVec& left) {
  const QaCode& code = p.pp.codes[level];
  uint64_t n = code.message_len;
  DVec out(n);
  dv::copy(c, out.p, left.p + n, n);
  launch_wht(out.p, 1, n, ilog2u(n), c.st);
  dv::mul(c, out.p, p.alpha_device(level), n);
  launch_wht(out.p, 1, n, ilog2u(n), c.st);
  dv::axpy(c, out.p, left.p, n, e8_one());
  return out;
}
// G applied to a message-length vector: the codeword of `right` (qafold_short.rs:1091-1104)
DVec generator_right(DevCtx& c, const ParamsState& p, size_t level, const DVec& right) {
  const QaCode& code = p.pp.codes[level];
  uint64_t n = code.message_len;
  DVec out(2 * n);
  dv::copy(c, out.p, right.p, n);
  launch_encode_parity(out.p, 1, ilog2u(n), p.alpha_encode(level), c.st);
  QF_HIP(hipGetLastError());
  return out;
}
DVec tw_vec(DevCtx& c, const std::vector<E8>& point) {
  DVec v(uint64_t(1) << point.size());
  tw_device(c, v.p, point);
  return v;
}

struct LB { E8 a_value, b_value; std::vector<E8> a_point, b_point; };
// qafold_short_pc.rs:1903-1948 (label null: the multi-evaluation form without a domain frame)
// Folds of two device vectors over their top halves, a' = a_lo + x a_hi and
// b' = b_lo + y b_hi, with the cross terms (sum a_lo b_hi, sum a_hi b_lo) of
// each result (LiteBullet and the Decide multi-evaluation). Once the vectors
// hold at most HOST_TAIL elements they are copied to the host and the last
// rounds run there: one download replaces a launch and a host round trip per
// round, with the same field arithmetic.
class TwoFold {
 public:
  static constexpr uint64_t HOST_TAIL = 256;
  TwoFold(DevCtx& c, DVec& a, DVec& b) : c_(c), da_(a), db_(b) {}
  // Cross terms of the length-n vectors (n even).
  std::pair<E8, E8> cross(uint64_t n) {
    if (n <= HOST_TAIL) to_host(n);
    if (!host_) return dv::cross_halves(c_, da_.p, db_.p, n / 2);
    return host_cross(n / 2);
  }
  // Folds length n to n / 2; the cross terms of the result (zero below length 2).
  std::pair<E8, E8> fold(uint64_t n, const E8& x, const E8& y) {
    const uint64_t h = n / 2;
    if (n <= HOST_TAIL) to_host(n);
    if (!host_) {
      if (h >= 2) return dv::lb_fold_cross(c_, da_.p, db_.p, h, x, y);
      dv::lb_fold(c_, da_.p, db_.p, h, x, y);
      return {E8{}, E8{}};
    }
    for (uint64_t i = 0; i < h; i++) {
      a_[i] = a_[i] + x * a_[h + i];
      b_[i] = b_[i] + y * b_[h + i];
    }
    a_.resize(h);
    b_.resize(h);
    return h >= 2 ? host_cross(h / 2) : std::pair<E8, E8>{E8{}, E8{}};
  }
  // The first n elements of a into device memory.
  void copy_a(E8* dst, uint64_t n) {
    if (host_) QF_HIP(hipMemcpyAsync(dst, a_.data(), n * sizeof(E8), hipMemcpyHostToDevice, c_.st));
    else dv::copy(c_, dst, da_.p, n);
  }
  // The rounds from length n down to `stop` on the device, with the device
  // copy of the transcript (device_fold_rounds): every round's (lc, rc, ch)
  // in `out`, n = stop and (lc, rc) the next cross terms. False (nothing done)
  // when the vectors are on the host or the range holds no round.
  bool device_rounds(const Pcs& tr, const char* label, uint64_t& n, uint64_t stop, E8& lc, E8& rc,
                     std::vector<DevRound>& out) {
    if (host_ || !device_rounds_on() || stop < HOST_TAIL || n <= stop) return false;
    out = device_fold_rounds(c_, da_.p, db_.p, n, stop, tr, label, lc, rc);
    return true;
  }
  E8 a0() { return host_ ? a_[0] : dv::download(c_, da_.p, 1)[0]; }
  E8 b0() { return host_ ? b_[0] : dv::download(c_, db_.p, 1)[0]; }

 private:
  void to_host(uint64_t n) {
    if (host_) return;
    std::tie(a_, b_) = dv::download2(c_, da_.p, db_.p, n);
    host_ = true;
  }
  std::pair<E8, E8> host_cross(uint64_t g) const {
    E8 l{}, r{};
    for (uint64_t i = 0; i < g; i++) {
      l += a_[i] * b_[g + i];
      r += a_[g + i] * b_[i];
    }
    return {l, r};
  }
  DevCtx& c_;
  DVec& da_;
  DVec& db_;
  std::vector<E8> a_, b_;
  bool host_ = false;
};

// The reference checks <left, right> == claimed before the rounds
// (qafold_short_pc.rs:1910-1916); each fold changes the inner product by
// exactly the cross terms it adds to rhs (<a', b'> = <a, b> + c l + c^-1 r for
// a folded by c^-1 and b by c), so the terminal check a b == rhs holds if and
// only if that one does. It takes its place here (same error), which saves a
// full-length reduction and a host round trip per run; a failing prover state
// then returns after the rounds instead of before them.
LB litebullet(DevCtx& c, E8 claimed, DVec& left, DVec& right, Pcs& tr, const char* label, const char* fold_label) {
  uint64_t n = left.n;
  require(pow2(n) && right.n == n, QF_ERR_INVALID_COMMITMENT, "litebullet shape");
  if (label) tr.bytes("litebullet-domain", label);
  LB lb;
  E8 rhs = claimed;
  E8 lc{}, rc{};
  TraceLaps lap;
  TwoFold tf(c, left, right);
  if (n > 1) std::tie(lc, rc) = tf.cross(n);
  while (n > 1) {  // each fold also returns the next round's cross terms
    std::vector<DevRound> dr;
    if (tf.device_rounds(tr, fold_label, n, TwoFold::HOST_TAIL, lc, rc, dr)) {
      for (const DevRound& r : dr) {  // the host transcript replays the device's rounds
        tr.respond(r.lc);
        tr.respond(r.rc);
        E8 ch = tr.challenge(fold_label, true);
        require(ch == r.ch, QF_ERR_TRANSCRIPT, "device transcript mismatch");
        E8 ci = e8_inv(ch);
        rhs += r.lc * ch + r.rc * ci;
        lb.a_point.push_back(ci);
        lb.b_point.push_back(ch);
      }
      continue;
    }
    tr.respond(lc);
    tr.respond(rc);
    E8 ch = tr.challenge(fold_label, true), ci = e8_inv(ch);
    rhs += lc * ch + rc * ci;
    std::tie(lc, rc) = tf.fold(n, ci, ch);
    lb.a_point.push_back(ci);
    lb.b_point.push_back(ch);
    n /= 2;
  }
  lb.a_value = tf.a0();
  lb.b_value = tf.b0();
  tr.respond(lb.a_value);
  tr.respond(lb.b_value);
  require(lb.a_value * lb.b_value == rhs, QF_ERR_INVALID_COMMITMENT, "litebullet claim");
  lap.mark(".litebullet (inside the above)");
  return lb;
}


// Row payload of a codeword: all entries, or sampled rows with Merkle paths (qafold_short_pc.rs:3105-3212).
void write_rows(DevCtx& c, Bytes& out, const DCode& d, bool all, const std::vector<uint64_t>& rows) {
  if (all) {
    out.push_back(1);
    put_u64(out, d.R);
    put_u64(out, d.cols);
    put_u64(out, d.R * d.cols);
    const size_t at = out.size();  // the column-major codeword, element bytes as stored
    out.resize(at + d.R * d.cols * sizeof(E8));
    dv::d2h(c, out.data() + at, d.data(), d.R * d.cols * sizeof(E8));
    return;
  }
  out.push_back(0);
  put_u64(out, rows.size());
  // per row: u64 cols, the row, u64 depth, its Merkle siblings
  dv::append_row_records(c, out, d.data(), d.R, d.cols, d.tree, rows);
}

// ================================================================ native opening
struct RoundOut { E8 next_value, gen_value; std::vector<E8> next_point, gen_left, gen_right; };
struct Pending { uint64_t level; E8 value; std::vector<E8> left, right; };

// qafold_short_pc.rs:1454-1525
RoundOut round_reductions(DevCtx& c, const ParamsState& p, size_t level, const std::vector<E8>& point,
                          const E8& fv, DVec& folded, const DVec& row_w, const std::vector<uint64_t>& sel,
                          const std::vector<E8>& sampled, Pcs& tr) {
  const QaCode& code = p.pp.codes[level];
  uint64_t n = code.message_len, n_ext = 2 * n;
  int col_log = 5;
  uint64_t sp = 1;
  while (sp < sel.size()) sp <<= 1;
  std::vector<E8> sel_point;
  for (int i = 0; i < ilog2u(sp); i++) sel_point.push_back(tr.challenge("selection-left"));
  std::vector<E8> sw(sel.size());
  for (size_t i = 0; i < sel.size(); i++) sw[i] = weight_at(sel_point, i);
  E8 selected_value = inner_host(sw, sampled);
  DVec pi(n_ext);
  dv::fill(c, pi.p, n_ext, E8{});
  dv::scatter_add(c, pi.p, sel, sw);
  DVec comb = generator_left(c, p, level, pi);
  E8 rho = tr.challenge("opening-batch");
  dv::axpy(c, comb.p, row_w.p, n, rho);
  LB msg = litebullet(c, selected_value + rho * fv, comb, folded, tr, "message-litebullet", "litebullet-fold");
  E8 ole = e8_one();
  for (size_t i = 0; i < msg.a_point.size(); i++) ole = ole * (e8_one() + point[col_log + i] * msg.a_point[i]);
  E8 sle = msg.a_value - rho * ole;
  DVec tw = tw_vec(c, msg.a_point);
  DVec gr = generator_right(c, p, level, tw);
  LB gen = litebullet(c, sle, pi, gr, tr, "generator-litebullet", "litebullet-fold");
  return {msg.b_value, gen.b_value, msg.b_point, gen.b_point, msg.a_point};
}

// One opening round of the native PCS (qafold_short_pc.rs:1120-1451).
// The top round takes the row weights and the projection its caller already
// computed for the claimed value (`top_row_w`, `top_opening`).
RoundOut prove_round(DevCtx& c, const ParamsState& p, size_t level, const DCode& cache, const std::vector<E8>& point,
                     const E8& value, bool top, Pcs& tr, Bytes& rows_wire, DCode& next, std::vector<E8>& terminal,
                     DVec* top_row_w = nullptr, const std::vector<E8>* top_opening = nullptr) {
  const QaCode& code = p.pp.codes[level];
  uint64_t n = code.message_len, R = cache.R;
  require(R == code.codeword_len() && point.size() == size_t(5 + ilog2u(n)), QF_ERR_INVALID_VARIABLES, "round dims");
  TraceLaps lap;
  DVec own_w;
  const std::vector<E8> rp(point.begin() + 5, point.end());
  if (!top_row_w) own_w = tw_vec(c, rp);
  const DVec& row_w = top_row_w ? *top_row_w : own_w;
  const std::vector<E8> opening = top_opening ? *top_opening : cache.projection_tw(c, n, rp, row_w.p);
  if (!top) require(inner_host(opening, tw_host(point.data(), 5)) == value, QF_ERR_INVALID_COMMITMENT, "round opening");
  tr.digest("round-root", cache.root);
  for (auto& o : opening) tr.respond(o);
  std::vector<E8> fr;
  for (int k = 0; k < 32; k++) fr.push_back(tr.challenge("freivalds"));
  lap.mark("native.projection");
  E8 fv = inner_host(opening, fr);
  DVec folded(n);
  cache.fold(c, n, fr, folded.p);
  if (level > 0) {
    next = commit_extension(c, p, level - 1, folded.p);
    tr.bytes("next-layer-kind", "extension-root");
    tr.digest("next-layer-binding", next.root);
  } else {
    terminal = dv::download(c, folded);
    tr.bytes("next-layer-kind", "terminal-vector");
    tr.digest("next-layer-binding", hash_terminal(terminal));
  }
  lap.mark("native.fold+commit");
  uint64_t count = num_queries(code, p.pp.security_bits, R);
  Digest seed = tr.challenge_digest("row-selection");
  bool all = count == R;
  std::vector<uint64_t> sel = sample_rows(count, R, seed);
  write_rows(c, rows_wire, cache, all, sel);
  std::vector<E8> sampled;
  if (all) {
    DVec rv(R);
    cache.fold(c, R, fr, rv.p);
    sampled = dv::download(c, rv);
  } else {  // sampled[k] = sum_col fr[col] codeword[col][sel[k]]
    sampled = dv::gather_weighted(c, cache.data(), R, cache.cols, nullptr, 0, 0, fr, sel);
  }
  tr.exts("sampled-values", sampled);
  lap.mark("native.rows");
  RoundOut ro = round_reductions(c, p, level, point, fv, folded, row_w, sel, sampled, tr);
  lap.mark("native.reductions");
  return ro;
}

// qafold_short_pc.rs:578-725
E8 open_with_reduction(DevCtx& c, const ParamsState& p, const DCode& top_code, const std::vector<E8>& point, Bytes& wire,
                       std::vector<Pending>& pending) {
  size_t levels = p.pp.codes.size(), top = levels - 1;
  uint64_t n = p.pp.codes[top].message_len;
  std::vector<E8> rp(point.begin() + 5, point.end());
  DVec row_w = tw_vec(c, rp);
  auto opening = top_code.projection_tw(c, n, rp, row_w.p);
  E8 claimed = inner_host(opening, tw_host(point.data(), 5));
  Pcs tr;
  tr.digest("parameters", p.pp.parameter_digest);
  tr.digest("commitment", top_code.root);
  tr.exts("point", point);
  tr.ext("claimed-value", claimed);
  Bytes top_rows;
  std::vector<std::pair<Digest, Bytes>> ext_rounds;
  DCode next;
  std::vector<E8> terminal;
  RoundOut ro = prove_round(c, p, top, top_code, point, claimed, true, tr, top_rows, next, terminal, &row_w, &opening);
  pending.clear();
  auto settle = [&](size_t level, const RoundOut& o) {
    if (level > 0) pending.push_back({level, o.gen_value, o.gen_left, o.gen_right});
  };
  settle(top, ro);
  for (size_t level = top; level-- > 0;) {
    DCode cur = std::move(next);
    Bytes rows;
    ro = prove_round(c, p, level, cur, ro.next_point, ro.next_value, false, tr, rows, next, terminal);
    ext_rounds.push_back({cur.root, std::move(rows)});
    settle(level, ro);
  }
  require(!terminal.empty() && inner_host(terminal, tw_host(ro.next_point)) == ro.next_value, QF_ERR_INVALID_COMMITMENT,
          "terminal vector");
  wire.clear();
  put_tag(wire, "QFSHORT2");
  put_u64(wire, tr.responses.size());
  for (auto& v : tr.responses) put_e8(wire, v);
  wire.insert(wire.end(), top_rows.begin(), top_rows.end());
  put_u64(wire, ext_rounds.size());
  for (auto& [r, rows] : ext_rounds) { put_digest(wire, r); wire.insert(wire.end(), rows.begin(), rows.end()); }
  put_u64(wire, terminal.size());
  for (auto& v : terminal) put_e8(wire, v);
  return claimed;
}

// ================================================================ Decide (direct-interleaved setup)
// qafold_short_decide_driver.rs:51-318 and the direct-interleaved helpers of
// qafold_short_decide*.rs
Bytes prove_decide(DevCtx& c, const ParamsState& p, const std::vector<Pending>& pending) {
  const auto& codes = p.pp.codes;
  size_t levels = codes.size();
  Pcs tr;
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
  for (auto& q : pending) h[q.level].push_back(reduce_generator(codes[q.level].message_len, {q.value, q.left, q.right}));
  Bytes rounds;
  uint64_t n_rounds = 0;
  std::vector<E8> base_vector;
  const uint64_t rf = 32;
  bool have_w = false;
  E8 w_value{};
  std::vector<E8> w_point;
  DCode carry;
  DVec carry_msg;
  bool have_carry = false;
  TraceLaps lap;
  for (size_t level = levels - 1; level >= 1; level--) {
    const SetupLevel& sl = p.pp.setup[level];
    const QaCode& code = codes[level];
    uint64_t rows = code.message_len, log_rows = ilog2u(rows);
    struct EC { E8 value; std::vector<E8> row_point, col_weight; };
    std::vector<EC> claims;
    for (auto& hc : h[level]) claims.push_back({hc.value, hc.left, tw_host(hc.right)});
    h[level].clear();
    uint64_t setup_cols = 2, carry_cols = have_w ? 1 : 0;
    if (have_w) {
      std::vector<E8> cwt(setup_cols, E8{});
      cwt.push_back(e8_one());
      claims.push_back({w_value, w_point, cwt});
    }
    require(!claims.empty(), QF_ERR_INVALID_VARIABLES, "no Decide claims");
    uint64_t real = setup_cols + carry_cols, padded = 1;
    while (padded < real) padded <<= 1;
    uint64_t log_cols = ilog2u(padded);
    const SetupShape& sh = sl.shape;
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
    if (have_carry) tr.digest("decide-round-carry-root", carry.root);
    // multi-evaluation over the flattened mixed matrix (qafold_short_decide.rs:486-551)
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
    lap.mark("decide.transcript");
    DVec left(padded * rows), right(padded * rows);
    dv::fill(c, left.p, padded * rows, E8{});
    dv::fill(c, right.p, padded * rows, E8{});
    dv::fill(c, left.p, rows, e8_inv(e8(uint32_t(rows % P))));  // alpha_0 = 1/n
    dv::copy(c, left.p + rows, p.alpha_device(level), rows);
    if (have_carry) dv::copy(c, left.p + 2 * rows, carry_msg.p, rows);
    E8 rhs{}, pw = e8_one();
    for (auto& cl : claims) {
      rhs += pw * cl.value;
      DVec rw = tw_vec(c, cl.row_point);
      for (size_t k = 0; k < cl.col_weight.size(); k++)
        if (!e8_is_zero(cl.col_weight[k])) dv::axpy(c, right.p + k * rows, rw.p, rows, pw * cl.col_weight[k]);
      pw = pw * rho;
    }
    lap.mark("decide.setup vectors");
    std::vector<E8> red_point;
    DVec folded_vector(rows);
    uint64_t n = padded * rows;
    E8 lc{}, rc{};
    TwoFold tf(c, left, right);
    if (n > 1) std::tie(lc, rc) = tf.cross(n);
    while (n > 1) {
      std::vector<DevRound> dr;
      const uint64_t stop = rows > TwoFold::HOST_TAIL && rows < n ? rows : TwoFold::HOST_TAIL;
      if (tf.device_rounds(tr, "multi-eval-fold", n, stop, lc, rc, dr)) {
        for (const DevRound& r : dr) {  // the host transcript replays the device's rounds
          tr.respond(r.lc);
          tr.respond(r.rc);
          E8 ch = tr.challenge("multi-eval-fold", true);
          require(ch == r.ch, QF_ERR_TRANSCRIPT, "device transcript mismatch");
          E8 ci = e8_inv(ch);
          rhs += r.lc * ch + r.rc * ci;
          red_point.push_back(ci);
        }
        if (n == rows) tf.copy_a(folded_vector.p, rows);
        continue;
      }
      tr.respond(lc);
      tr.respond(rc);
      E8 ch = tr.challenge("multi-eval-fold", true), ci = e8_inv(ch);
      rhs += lc * ch + rc * ci;
      std::tie(lc, rc) = tf.fold(n, ci, ch);
      red_point.push_back(ci);
      n /= 2;
      if (n == rows) tf.copy_a(folded_vector.p, rows);
    }
    E8 red_value = tf.a0(), bval = tf.b0();
    tr.respond(red_value);
    tr.respond(bval);
    require(red_value * bval == rhs, QF_ERR_INVALID_COMMITMENT, "multi-eval terminal");
    lap.mark("decide.multieval");
    std::vector<E8> col_point(red_point.begin(), red_point.begin() + log_cols);
    std::vector<E8> merged(red_point.begin() + log_cols, red_point.end());
    auto col_w = tw_host(col_point);
    const QaCode& target = codes[level - 1];
    uint64_t n_enc = target.codeword_len();
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
    uint64_t phys_carry = carry_cols * rf, scols = sh.t_cols + sh.h_cols;
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
    if (have_carry) tr.digest("iota-carry-root", carry.root);
    tr.u64("iota-carry-cols", phys_carry);
    absorb_sel(tr, sel);
    tr.exts("iota-column-weights", iw);
    // iota[k] = sum_col iw[col] setup[col][row_k] + sum_col iw[scols + col] carry[col][row_k], on the device
    const E8* d_setup = p.setup_device(level);
    std::vector<E8> iota = dv::gather_weighted(c, d_setup, n_enc, scols, have_carry ? carry.data() : nullptr,
                                               have_carry ? carry.R : 0, have_carry ? phys_carry : 0, iw, srows);
    for (auto& v : iota) tr.respond(v);
    lap.mark("decide.iota");
    Bytes round;
    put_u64(round, level);
    uint64_t trows = target.message_len;
    DVec rec(trows);
    dv::col_fold(c, folded_vector.p, trows, trows, rf, bw, rec.p);  // fold the block coordinate
    DCode next_carry;
    bool have_next = level > 1;
    if (have_next) next_carry = commit_extension(c, p, level - 2, rec.p);
    tr.bytes("decide-next-carry-domain", "qafold-short/decide-next-carry/v1");
    tr.bytes("decide-next-carry-present", have_next ? &one : &zero, 1);
    if (have_next) tr.digest("decide-next-carry-root", next_carry.root);
    round.push_back(have_next ? 1 : 0);
    if (have_next) put_digest(round, next_carry.root);
    if (sel.all) {
      round.push_back(1);
      for (uint64_t v : {sh.level, sh.encoded_rows, sh.message_rows, sh.t_cols, sh.t_elements, sh.t_last_len, sh.h_cols})
        put_u64(round, v);
      put_u64(round, sl.encoded.size());
      for (auto& v : sl.encoded) put_e8(round, v);
    } else {
      round.push_back(0);
      put_u64(round, sel.rows.size());
      auto srow_vals = dv::gather_rows(c, d_setup, n_enc, scols, sel.rows);  // row-major, contiguous
      for (size_t k = 0; k < sel.rows.size(); k++) {
        const uint64_t r = sel.rows[k];
        put_u64(round, r);
        put_u64(round, scols);
        for (uint64_t col = 0; col < scols; col++) put_e8(round, srow_vals[k * scols + col]);
        uint64_t idx = r;
        put_u64(round, sl.tree.size() - 1);
        for (size_t l = 0; l + 1 < sl.tree.size(); l++) { put_digest(round, sl.tree[l][idx ^ 1]); idx >>= 1; }
      }
    }
    if (have_carry) { round.push_back(1); write_rows(c, round, carry, sel.all, srows); } else round.push_back(0);
    lap.mark("decide.carry+rows");
    // direct code switch (qafold_short_decide_carry.rs:940-1027)
    std::vector<E8> base_point(merged.begin() + 5, merged.end());
    DVec tbp = tw_vec(c, base_point);
    require(dv::inner(c, rec.p, tbp.p, trows) == red_value, QF_ERR_INVALID_COMMITMENT, "merged carry value");
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
    E8 selected_value = inner_host(sw, iota);
    DVec sparse(n_enc);
    dv::fill(c, sparse.p, n_enc, E8{});
    dv::scatter_add(c, sparse.p, srows, sw);
    DVec gl = generator_left(c, p, level - 1, sparse);
    E8 rho2 = tr.challenge("direct-code-switch-opening-batch");
    dv::axpy(c, gl.p, tbp.p, trows, rho2);
    DVec rec_copy(trows);
    dv::copy(c, rec_copy.p, rec.p, trows);
    LB msg = litebullet(c, selected_value + rho2 * red_value, gl, rec_copy, tr, "decide-direct-carry-message-litebullet",
                        "litebullet-fold");
    E8 sle = msg.a_value - rho2 * tensor_inner(base_point, msg.a_point);
    DVec tw = tw_vec(c, msg.a_point);
    DVec gr = generator_right(c, p, level - 1, tw);
    LB gen = litebullet(c, sle, sparse, gr, tr, "decide-direct-carry-generator-litebullet", "litebullet-fold");
    if (level > 1) {
      h[level - 1].push_back(reduce_generator(trows, {gen.b_value, gen.b_point, msg.a_point}));
      w_value = msg.b_value;
      w_point = msg.b_point;
      have_w = true;
      carry = std::move(next_carry);
      carry_msg = std::move(rec);
      have_carry = true;
    } else {
      auto term = dv::download(c, rec);
      require(inner_host(term, tw_host(msg.b_point)) == msg.b_value, QF_ERR_INVALID_COMMITMENT, "terminal carry");
      for (auto& v : term) tr.respond(v);
      base_vector = term;
      have_carry = false;
    }
    rounds.insert(rounds.end(), round.begin(), round.end());
    n_rounds++;
    lap.mark("decide.switch");
  }
  Bytes w;
  put_tag(w, "QFSDCID1");
  put_u64(w, tr.responses.size());
  for (auto& v : tr.responses) put_e8(w, v);
  put_u64(w, n_rounds);
  w.insert(w.end(), rounds.begin(), rounds.end());
  put_u64(w, base_vector.size());
  for (auto& v : base_vector) put_e8(w, v);
  return w;
}

std::vector<Val> read_values(DevCtx& c, const Val* bva, uint64_t active, uint64_t start, uint64_t n) {
  std::vector<Val> out(n, 0);
  if (start < active) {
    uint64_t m = std::min(n, active - start);
    dv::d2h(c, out.data(), bva + start, m * sizeof(Val));
  }
  return out;
}
// Lane projections of the terms with len >= s on the device (tbv.rs:288-352).
std::vector<std::vector<E8>> project_terms(DevCtx& c, const Query& q, const Val* bva, uint64_t active, uint64_t s,
                                           std::vector<size_t>& which) {
  std::vector<ProjTask> tasks;
  std::vector<std::array<E8, 2>> pairs;
  int ls = ilog2u(s);
  for (size_t i = 0; i < q.terms.size(); i++) {
    const Term& t = q.terms[i];
    if (t.len() < s || t.offset >= active || e8_is_zero(t.scale)) continue;
    ProjTask pt;
    pt.base = t.offset;
    pt.rows = t.len() / s;
    pt.nbits = uint32_t(t.f.size() - ls);
    pt.pair_off = uint32_t(pairs.size());
    pt.scale = t.scale;
    for (size_t b = ls; b < t.f.size(); b++) pairs.push_back(t.f[b]);
    tasks.push_back(pt);
    which.push_back(i);
  }
  auto flat = tensor_project(c, bva, active, int(s), tasks, pairs);
  std::vector<std::vector<E8>> out(tasks.size());
  for (size_t k = 0; k < tasks.size(); k++) out[k].assign(flat.begin() + k * s, flat.begin() + (k + 1) * s);
  return out;
}

// Output pairs from which an EO round projects on the query's terms instead of
// reading a materialized query (QF_EO_OTF_MIN; default 2^21: below that the
// vectors are cache-resident and the extra launches cost more than the query).
uint64_t eo_otf_min_pairs() {
  const char* e = std::getenv("QF_EO_OTF_MIN");
  return e ? uint64_t(std::strtoull(e, nullptr, 10)) : (uint64_t(1) << 21);
}

// ================================================================ EO (eo.rs:97-181)
// The packed reduction and its binary rounds. Every term of at least 2s
// entries is projected once at stride 2s (its offset is a multiple of its
// power-of-two length): the stride-s projections behind the packed cross
// terms, their d = 0 sum and, once the packed challenge is known, the first
// binary round's cross terms all follow from those chunk sums. Terms below 2s
// entries lie in one row of stride s and are handled on the host from the BVA
// entries around them. The reference's consistency check dot(w, qf) == claim
// (eo.rs:134-138) equals, term for term, sum_i v[i] q[i] == query.claim, since
// both sides add the same cross-term multiples of the challenge; the check is
// made on that sum. The witness of stride s is never written: the first round
// reads the BVA at stride 2s (OtfQuery::first_fold), and the long rounds
// project their outputs on the query's terms rather than reading a query.
struct EoOut { std::vector<E8> pos, neg; std::vector<std::array<E8, 2>> rounds; E8 value, query; std::vector<E8> point; E8 scale; };
EoOut eo_prove(DevCtx& c, const Layout& l, const Val* bva, uint64_t s, const Query& q, Outer& tr) {
  EoOut o;
  o.pos.assign(s - 1, E8{});
  o.neg.assign(s - 1, E8{});
  int ls = ilog2u(s);
  TraceLaps lap;
  // the lane weights of every term of at least s entries, used by the projection and the query fold
  // (the buckets of one matrix share their low factors, so equal factor lists share one table)
  std::vector<std::vector<E8>> term_low(q.terms.size());
  std::vector<size_t> low_of(q.terms.size(), SIZE_MAX);  // the first term with the same low factors
  {
    std::map<std::string, size_t> seen;
    for (size_t i = 0; i < q.terms.size(); i++) {
      if (q.terms[i].len() < s) continue;
      const std::string key(reinterpret_cast<const char*>(q.terms[i].f.data()), size_t(ls) * sizeof(q.terms[i].f[0]));
      auto [it, fresh] = seen.emplace(key, i);
      low_of[i] = it->second;
      term_low[i] = fresh ? tbv_w(q.terms[i].f.data(), ls) : term_low[it->second];
    }
  }
  std::vector<ProjTask> tasks;
  std::vector<std::array<E8, 2>> pairs, fbit;
  std::vector<E8> lows;
  std::vector<size_t> big;  // the term of each task
  for (size_t i = 0; i < q.terms.size(); i++) {
    const Term& t = q.terms[i];
    if (t.len() < 2 * s || t.offset >= l.active || e8_is_zero(t.scale)) continue;
    tasks.push_back({t.offset, t.len() / (2 * s), uint32_t(t.f.size() - ls - 1), uint32_t(pairs.size()), t.scale});
    for (size_t b = ls + 1; b < t.f.size(); b++) pairs.push_back(t.f[b]);
    lows.insert(lows.end(), term_low[i].begin(), term_low[i].end());
    fbit.push_back(t.f[ls]);
    big.push_back(i);
  }
  lap.mark("eo.terms");
  EoProjection proj = eo_project(c, bva, l.active, int(s), tasks, pairs);
  EoCorr corr = eo_correlations(c, proj, fbit, lows);
  o.pos = corr.pos;
  o.neg = corr.neg;
  E8 actual = corr.zero;  // sum_i v[i] q[i]
  lap.mark("eo.project+corr");
  // terms below 2s entries: one row k of stride s each
  struct SmallTerm { uint64_t k; std::vector<E8> query; };
  std::vector<SmallTerm> smalls;
  for (auto& t : q.terms) {
    if (t.len() >= 2 * s) continue;
    uint64_t start = t.offset / s * s;
    auto vals = read_values(c, bva, l.active, start, s);
    std::vector<E8> query(s, E8{});
    for (uint64_t i = 0; i < t.len(); i++) query[t.offset - start + i] = t.weight(i);
    for (uint64_t lane = 0; lane < s; lane++) actual += query[lane] * small_val(vals[lane]);
    for (uint64_t d = 1; d < s; d++)
      for (uint64_t lane = 0; lane + d < s; lane++) {
        o.pos[d - 1] += query[lane + d] * small_val(vals[lane]);
        o.neg[d - 1] += query[lane] * small_val(vals[lane + d]);
      }
    smalls.push_back({t.offset / s, std::move(query)});
  }
  tr.object("packed-positive", js_scalars(o.pos));
  tr.object("packed-negative", js_scalars(o.neg));
  E8 ch = tr.fold_challenge(ls);
  std::vector<E8> challenges;
  E8 pw = ch;
  for (int i = 0; i < ls; i++) { challenges.push_back(pw); pw = pw * pw; }
  E8 inv = e8_inv(ch);
  std::vector<E8> pinv(s), pows(s);
  pinv[0] = pows[0] = e8_one();
  for (uint64_t i = 1; i < s; i++) { pinv[i] = pinv[i - 1] * inv; pows[i] = pows[i - 1] * ch; }
  require(actual == q.claim, QF_ERR_INVALID_COMMITMENT, "packed reduction claim mismatch");
  lap.mark("eo.small+transcript");
  uint64_t rows = l.logical / s;
  // the query fold of each term of at least s entries: scale x sum_l low[l] c^l
  std::vector<E8> fold_scale(q.terms.size(), E8{});
  for (size_t i = 0; i < q.terms.size(); i++)
    if (q.terms[i].len() >= s) {
      if (low_of[i] != i) {
        fold_scale[i] = fold_scale[low_of[i]];
        continue;
      }
      E8 acc{};
      for (uint64_t j = 0; j < s; j++) acc += term_low[i][j] * pows[j];
      fold_scale[i] = acc;
    }
  // a[i]: term i folded over the entry bits below an element of the current
  // level (level L: the s lanes by c, then L binary rounds by their challenges;
  // q_L[k] = q_{L-1}[2k] + q_{L-1}[2k+1] r_L, tbv.rs:393-428). A term of at
  // least one element is a[i] times the tensor of its factors above those bits;
  // a shorter one is the single value a[i] at its element.
  std::vector<E8> a(q.terms.size());
  for (size_t i = 0; i < q.terms.size(); i++) {
    const Term& t = q.terms[i];
    if (t.len() >= s) a[i] = t.scale * fold_scale[i];
    else
      for (uint64_t j = 0; j < t.len(); j++) a[i] += t.weight(j) * pows[(t.offset + j) % s];
  }
  // a from level L to L + 1 by the binary challenge r: the factor of entry bit
  // ls + L folds to f[0] + r f[1]; a shorter term picks up r when that bit of its offset is set
  auto advance = [&](int level, const E8& r) {
    const size_t b = size_t(ls + level);
    for (size_t i = 0; i < q.terms.size(); i++) {
      const Term& t = q.terms[i];
      if (b < t.f.size()) a[i] = a[i] * (t.f[b][0] + r * t.f[b][1]);
      else if ((t.offset >> b) & 1) a[i] = a[i] * r;
    }
  };
  // The query of level L into out (rows >> L entries): bucket terms are
  // written (the entries outside them zeroed), then padding and suffix terms
  // added; each phase has disjoint supports. Terms shorter than an element go
  // in as single values.
  auto build_query = [&](E8* out, int level) {
    const uint64_t span = s << level;  // entries per element
    std::map<uint64_t, E8> single;
    for (int phase = 0; phase < 2; phase++) {
      std::vector<FillTask> ftasks;
      std::vector<std::array<E8, 2>> fpairs;
      size_t lo = phase ? q.n_bucket_terms : 0, hi = phase ? q.terms.size() : q.n_bucket_terms;
      for (size_t i = lo; i < hi; i++) {
        const Term& t = q.terms[i];
        if (t.len() >= span) {
          FillTask ft{t.offset / span, uint32_t(t.f.size() - ls - level), uint32_t(fpairs.size()), a[i]};
          for (size_t bb = ls + level; bb < t.f.size(); bb++) fpairs.push_back(t.f[bb]);
          ftasks.push_back(ft);
        } else {
          single[t.offset / span] += a[i];
        }
      }
      if (phase == 0) tensor_fill(c, out, rows >> level, ftasks, fpairs);
      else tensor_fill_add(c, out, ftasks, fpairs);
    }
    std::vector<uint64_t> idx;
    std::vector<E8> vals;
    for (auto& [k, v] : single) { idx.push_back(k); vals.push_back(v); }
    dv::scatter_add(c, out, idx, vals);
  };
  lap.mark("eo.fold_scale");
  E8 claim = q.claim, fwd = ch, bwd = inv;
  for (uint64_t d = 0; d + 1 < s; d++) { claim += o.pos[d] * fwd + o.neg[d] * bwd; fwd = fwd * ch; bwd = bwd * inv; }
  // the first binary round's cross terms sum_j w[2j] qf[2j+1], sum_j w[2j+1] qf[2j]
  E8 c0{}, c1{};
  if (rows >= 2) {
    // the terms' scales are in the projections already: coef carries the lane fold and row bit 0
    std::vector<std::array<E8, 2>> coef;
    for (size_t k = 0; k < big.size(); k++) {
      const E8 f = fold_scale[big[k]];
      coef.push_back({f * fbit[k][0], f * fbit[k][1]});
    }
    std::tie(c0, c1) = eo_first_cross(c, proj, pinv, coef);
    for (auto& st : smalls) {  // a term in row k pairs with the witness of row k ^ 1
      E8 qf{};
      for (uint64_t lane = 0; lane < s; lane++) qf += st.query[lane] * pows[lane];
      auto vals = read_values(c, bva, l.active, (st.k ^ 1) * s, s);
      E8 wv{};
      for (uint64_t lane = 0; lane < s; lane++) wv += pinv[lane] * small_val(vals[lane]);
      if (st.k & 1) c0 += wv * qf;
      else c1 += wv * qf;
    }
  }
  lap.mark("eo.first_cross");
  // Long vectors: the first round reads the BVA (no witness of stride s is
  // written) and, while a round's outputs have at least eo_otf_min_pairs()
  // pairs, it projects them on the query's terms instead of reading a query
  // (OtfQuery); the query is materialized at the first level below that.
  // Short vectors materialize the witness and the query at level 0.
  const bool from_bva = rows / 4 >= (uint64_t(1) << OTF_LCP);
  DVec w(from_bva ? rows / 4 : rows), w2(rows / 2 ? rows / 2 : 1), qa, qb;
  DVec *cw = &w, *nw = &w2, *cq = nullptr, *nq = nullptr;
  std::unique_ptr<OtfQuery> otf;
  // every term's factors in one table (fofs: the first of term i), and the
  // lowest bit at which a factor pair has a zero entry
  std::vector<uint32_t> fofs(q.terms.size());
  std::vector<size_t> zmin(q.terms.size(), SIZE_MAX);
  if (from_bva) {
    std::vector<std::array<E8, 2>> factors;
    for (size_t i = 0; i < q.terms.size(); i++) {
      fofs[i] = uint32_t(factors.size());
      for (size_t b = 0; b < q.terms[i].f.size(); b++) {
        const auto& fp = q.terms[i].f[b];
        factors.push_back(fp);
        if (zmin[i] == SIZE_MAX && (e8_is_zero(fp[0]) || e8_is_zero(fp[1]))) zmin[i] = b;
      }
    }
    otf = std::make_unique<OtfQuery>(c, factors, (rows / 4) >> OTF_LCP, q.n_bucket_terms, 0, 0);
  } else {
    qa = DVec(rows);
    qb = DVec(rows / 2 ? rows / 2 : 1);
    build_query(qa.p, 0);
    eo_witness(c, bva, l.active, rows, int(s), pinv, w.p);
    cq = &qa;
    nq = &qb;
  }
  // The records of output level L: bucket terms of at least one chunk of pairs
  // go whole, the other terms of at least one pair in pieces, and terms of at
  // most one element as items. A term other than a bucket term covers padding
  // only, whose entries are zero (validated at commit): while every factor
  // with a zero entry lies above the pair bit, each pair in its range is two
  // zero witness elements and it adds nothing; so does a term past the active entries.
  auto otf_level = [&](int level) {
    OtfLevel lv;
    const size_t pb = size_t(ls + level);  // the entry bit of output bit 0
    const uint64_t span = uint64_t(s) << level;
    for (size_t i = 0; i < q.terms.size(); i++) {
      const Term& t = q.terms[i];
      if (e8_is_zero(a[i])) continue;
      if (t.f.size() <= pb) {
        lv.items.push_back({t.offset / span, 0, a[i]});
        continue;
      }
      if (t.offset >= l.active) continue;
      const bool bucket = i < q.n_bucket_terms;
      if (!bucket && zmin[i] > pb) continue;
      const uint32_t lp = uint32_t(t.f.size() - pb - 1);  // log2 of its pairs
      const uint64_t pair0 = t.offset / (2 * span), np = uint64_t(1) << lp;
      const E8 g0 = a[i] * t.f[pb][1], g1 = a[i] * t.f[pb][0];
      const uint32_t fb = fofs[i] + uint32_t(pb) + 1;
      if (bucket && lp >= uint32_t(OTF_LCP)) {
        lv.fast.push_back({pair0, lp, fb, g0, g1});
      } else {
        for (uint64_t k = 0; k < np; k += OTF_PIECE)
          lv.pieces.push_back({pair0 + k, k, uint32_t(std::min<uint64_t>(OTF_PIECE, np - k)), fb, lp, 0, g0, g1});
      }
    }
    std::sort(lv.fast.begin(), lv.fast.end(), [](const OtfTermDev& x, const OtfTermDev& y) { return x.pair0 < y.pair0; });
    return lv;
  };
  const uint64_t otf_min = std::max<uint64_t>(eo_otf_min_pairs(), uint64_t(1) << OTF_LCP);
  uint64_t n = rows;
  int level = 0;
  // the last rounds run on host copies once the vectors are short (as TwoFold)
  std::vector<E8> hw, hq;
  while (n > 1) {
    uint64_t h = n / 2;
    tr.object("binary-cross", js_scalars({c0, c1}));
    E8 r = tr.fold_challenge(1), ri = e8_inv(r);
    claim += c0 * r + c1 * ri;
    challenges.push_back(r);
    o.rounds.push_back({c0, c1});
    if (!cq) {  // w' and the cross terms over it from the query's terms
      advance(level, r);
      const OtfLevel lv = otf_level(level + 1);
      lap.mark("eo.otf_records");
      if (level == 0) {  // w'[i] = sum_{L<2s} v[2is + L] pinv2[L], pinv2 = (c^-l, r^-1 c^-l)
        std::vector<E8> pinv2(2 * s);
        for (uint64_t lane = 0; lane < s; lane++) { pinv2[lane] = pinv[lane]; pinv2[s + lane] = ri * pinv[lane]; }
        std::tie(c0, c1) = otf->first_fold(c, bva, l.active, rows, int(s), pinv2, lv, w2.p);
        cw = &w2;  // the next round writes into w (rows / 4)
        nw = &w;
      } else {
        std::tie(c0, c1) = otf->fold(c, cw->p, nw->p, h, ri, lv);
        std::swap(cw, nw);
      }
      level++;
      n = h;
      if (n / 4 < otf_min) {  // the next round's outputs are short: materialize the query of this level
        qa = DVec(n);
        qb = DVec(n / 2);
        build_query(qa.p, level);
        cq = &qa;
        nq = &qb;
      }
      lap.mark(level == 1 ? "eo.first_fold" : "eo.otf_rounds");
      continue;
    }
    level++;
    if (hw.empty() && n <= TwoFold::HOST_TAIL) {
      std::tie(hw, hq) = dv::download2(c, cw->p, cq->p, n);
    }
    if (!hw.empty()) {  // w'[i] = w[2i] + w[2i+1] r^-1, q'[i] = q[2i] + q[2i+1] r, then the cross terms
      for (uint64_t i = 0; i < h; i++) {
        hw[i] = hw[2 * i] + ri * hw[2 * i + 1];
        hq[i] = hq[2 * i] + r * hq[2 * i + 1];
      }
      hw.resize(h);
      hq.resize(h);
      c0 = c1 = E8{};
      for (uint64_t j = 0; 2 * j + 1 < h; j++) {
        c0 += hw[2 * j] * hq[2 * j + 1];
        c1 += hw[2 * j + 1] * hq[2 * j];
      }
    } else {
      if (h >= 2) {
        std::tie(c0, c1) = dv::eo_fold_cross(c, cw->p, cq->p, nw->p, nq->p, h, ri, r);
      } else {
        dv::eo_fold(c, cw->p, cq->p, nw->p, nq->p, h, ri, r);
      }
      std::swap(cw, nw);
      std::swap(cq, nq);
    }
    n = h;
  }
  lap.mark("eo.rounds");
  o.value = hw.empty() ? dv::download(c, cw->p, 1)[0] : hw[0];
  o.query = hq.empty() ? dv::download(c, cq->p, 1)[0] : hq[0];
  require(claim == o.value * o.query && query_eval(q, challenges) == o.query, QF_ERR_INVALID_COMMITMENT, "EO terminal mismatch");
  tr.scalar("terminal-value", o.value);
  tr.scalar("terminal-query", o.query);
  o.scale = e8_one();
  for (auto& x : challenges) {
    E8 m = e8_inv(x), f = e8_one() + m;
    o.point.push_back(m * e8_inv(f));
    o.scale = o.scale * f;
  }
  std::reverse(o.point.begin(), o.point.end());
  return o;
}

// ================================================================ packed PCS openings (packed.rs, packed_batch.rs)
// `first`: the first message, when the caller has it already (computed beside other work).
void sumcheck(DevCtx& c, DVec& v, DVec& w, E8& claim, Packed& t, std::vector<std::array<E8, 2>>& rounds,
              std::vector<E8>& reduced, E8& fval, E8& fwt, const std::pair<E8, E8>* first = nullptr) {
  uint64_t n = v.n;
  E8 m0{}, m1{};
  if (first) std::tie(m0, m1) = *first;
  else if (n > 1) std::tie(m0, m1) = dv::sc_msg(c, v.p, w.p, n / 2);
  // the last rounds run on host copies once the vectors are short (as TwoFold)
  std::vector<E8> hv, hw;
  while (n > 1) {  // each fold also returns the next round's message
    uint64_t h = n / 2;
    t.field(m0);
    t.field(m1);
    E8 r = t.challenge();
    claim = m0 + r * ((claim - m0 - m0 - m1) + r * m1);
    rounds.push_back({m0, m1});
    if (hv.empty() && n <= TwoFold::HOST_TAIL) {
      std::tie(hv, hw) = dv::download2(c, v.p, w.p, n);
    }
    if (!hv.empty()) {  // v' = v_lo + r (v_hi - v_lo); message of the halves of v', w'
      for (uint64_t i = 0; i < h; i++) {
        hv[i] = hv[i] + r * (hv[h + i] - hv[i]);
        hw[i] = hw[i] + r * (hw[h + i] - hw[i]);
      }
      hv.resize(h);
      hw.resize(h);
      m0 = m1 = E8{};
      for (uint64_t g = h / 2, i = 0; i < g; i++) {
        m0 += hv[i] * hw[i];
        m1 += (hv[g + i] - hv[i]) * (hw[g + i] - hw[i]);
      }
    } else if (h >= 2) {
      std::tie(m0, m1) = dv::sc_fold_msg(c, v.p, w.p, h, r);
    } else {
      dv::sc_fold(c, v.p, w.p, h, r);
    }
    reduced.push_back(r);
    n = h;
  }
  fval = hv.empty() ? dv::download(c, v.p, 1)[0] : hv[0];
  fwt = hw.empty() ? dv::download(c, w.p, 1)[0] : hw[0];
}
// Lanes of every block against eq(outer): lanes[b][j] = sum_i entry(b, LANES i + j) * eq_i
std::vector<Lanes> block_lanes(DevCtx& c, const ParamsState& p, const Val* bva, uint64_t active,
                                           uint64_t blocks, const std::vector<E8>& outer) {
  std::vector<std::array<E8, 2>> pairs;
  size_t m = outer.size();
  for (size_t b = 0; b < m; b++) { const E8& pc = outer[m - 1 - b]; pairs.push_back({e8_one() - pc, pc}); }
  std::vector<ProjTask> tasks;
  for (uint64_t b = 0; b < blocks; b++) tasks.push_back({b * p.block_len, uint64_t(1) << m, uint32_t(m), 0, e8_one()});
  auto flat = tensor_project(c, bva, active, LANES, tasks, pairs);
  std::vector<Lanes> out(blocks);
  for (uint64_t b = 0; b < blocks; b++)
    for (int j = 0; j < LANES; j++) out[b][j] = flat[b * LANES + j];
  return out;
}
Lanes conjugates(const Lanes& lanes) {
  Lanes e;
  for (int img = 0; img < LANES; img++) {
    E8 s{};
    for (int j = 0; j < LANES; j++) s += e8_basis(j) * e8_frob(lanes[j], (LANES - img) % LANES);
    e[img] = s;
  }
  return e;
}

struct Timing { ProveTimings* t; };

// A persistent helper thread per prover thread, with its own stream and pinned
// buffer (the DevCtx it makes for each task), running one task at a time.
class SideThread {
 public:
  using Task = std::function<void(DevCtx&)>;
  SideThread() : th_([this] { loop(); }) {}
  ~SideThread() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    th_.join();
  }
  void start(int device, Task t) {
    std::lock_guard<std::mutex> g(mu_);
    require(!busy_, QF_ERR_ARGUMENT, "helper thread already has a task");
    device_ = device;
    task_ = std::move(t);
    err_ = nullptr;
    busy_ = true;
    cv_.notify_all();
  }
  // Waits for the task; rethrows its error.
  void wait() {
    std::unique_lock<std::mutex> g(mu_);
    cv_.wait(g, [this] { return !busy_; });
    if (err_) {
      std::exception_ptr e = err_;
      err_ = nullptr;
      std::rethrow_exception(e);
    }
  }

 private:
  void loop() {
    std::unique_lock<std::mutex> g(mu_);
    for (;;) {
      cv_.wait(g, [this] { return stop_ || (busy_ && task_); });
      if (stop_) return;
      Task t = std::move(task_);
      task_ = nullptr;
      const int dev = device_;
      g.unlock();
      std::exception_ptr e;
      try {
        QF_HIP(hipSetDevice(dev));
        DevCtx c;
        t(c);
      } catch (...) {
        e = std::current_exception();
      }
      g.lock();
      err_ = e;
      busy_ = false;
      cv_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_;
  Task task_;
  std::exception_ptr err_;
  int device_ = 0;
  bool busy_ = false, stop_ = false;
  std::thread th_;  // last: it starts after the members it uses
};
SideThread& side_thread() {
  thread_local SideThread s;
  return s;
}
// QF_SIDE_LEAVES=0 hashes all of the folded block's leaves on the prover's stream
// after its parity passes.
bool side_leaves_on() {
  const char* e = std::getenv("QF_SIDE_LEAVES");
  return !(e && e[0] == '0');
}
// QF_SIDE_BINDING=0 writes the binding rows on the prover's thread before the
// native opening.
bool side_binding_on() {
  const char* e = std::getenv("QF_SIDE_BINDING");
  return !(e && e[0] == '0');
}

// Codeword of block b: the resident cache, or a fresh commit in recompute mode.
DCode block_code(DevCtx& c, const ParamsState& p, const CommittedState& cs, uint64_t b, DCode& workspace) {
  uint64_t R = 2 * (uint64_t(1) << p.geom.log_n);
  DCode d;
  d.R = R;
  d.root = cs.roots[b];
  if (cs.cache_mode == QF_CACHE_CACHED && b < cs.d_cw.size()) {
    d.cwp = cs.d_cw[b];
    d.tree = cs.d_tree[b];
    return d;
  }
  if (!workspace.cw.p) workspace = alloc_code(R);
  launch_commit_block(p.geom, cs.d_bva, b * p.block_len, p.layout.active, workspace.cw.p,
                      p.alpha_encode(p.pp.codes.size() - 1), workspace.tree, c.st);
  QF_HIP(hipGetLastError());
  Digest r = download_root(c, workspace.tree, R);
  require(r == cs.roots[b], QF_ERR_INVALID_COMMITMENT, "recomputed block root mismatch");
  d.cwp = workspace.cw.p;
  d.tree = workspace.tree;
  return d;
}

// Appends the batched opening's wire to w (the proof being built), so the
// binding rows, most of a proof's bytes, are written once, in place.
void open_batch(DevCtx& c, const ParamsState& p, const CommittedState& cs, const std::vector<E8>& point,
                const Bytes& context, E8& claimed, ProveTimings& tm, Bytes& w) {
  auto t0 = Clock::now();
  TraceLaps lap;
  uint64_t blocks = cs.roots.size();
  size_t split = ilog2u(blocks), n_outer = point.size() - LOG_LANES - split;
  std::vector<E8> outer(point.begin() + split, point.begin() + split + n_outer);
  auto low = eq_host(point.data() + point.size() - LOG_LANES, LOG_LANES);
  auto high = eq_host(point.data(), split);
  uint64_t plen = uint64_t(1) << n_outer;
  lap.mark("projection.setup");
  auto lanes = block_lanes(c, p, cs.d_bva, p.layout.active, blocks, outer);
  lap.mark("projection.lanes");
  std::vector<Lanes> ebb(blocks);
  for (uint64_t b = 0; b < blocks; b++) ebb[b] = conjugates(lanes[b]);
  Lanes ev;
  for (int img = 0; img < LANES; img++) {
    E8 s{};
    for (uint64_t b = 0; b < blocks; b++) s += e8_frob(high[b], (LANES - img) % LANES) * ebb[b][img];
    ev[img] = s;
  }
  claimed = projected_claim(ev, low);
  Packed t = packed_transcript(p, cs.roots[0], point, claimed, context);
  const char* bd = "one-opening-block-packed-eo/v1";
  t.bytes(bd, std::strlen(bd));
  uint8_t nb[8];
  for (int i = 0; i < 8; i++) nb[i] = uint8_t(blocks >> (8 * i));
  t.bytes(nb, 8);
  for (auto& r : cs.roots) t.bytes(r.data(), 32);
  for (auto& v : ev) t.field(v);
  auto powers = lane_powers(t.challenge());
  E8 claim{};
  for (int i = 0; i < LANES; i++) claim += powers[i] * ev[i];
  auto cross = [&](uint64_t w, uint64_t q) {
    E8 s{};
    for (int i = 0; i < LANES; i++) s += powers[i] * e8_frob(high[q], (LANES - i) % LANES) * ebb[w][i];
    return s;
  };
  std::vector<E8> pos, neg;
  for (uint64_t d = 1; d < blocks; d++) {
    E8 a{}, b2{};
    for (uint64_t b = 0; b + d < blocks; b++) { a += cross(b, b + d); b2 += cross(b + d, b); }
    pos.push_back(a);
    neg.push_back(b2);
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
  Lanes images;
  for (int lane = 0; lane < LANES; lane++) {
    E8 s{};
    for (int img = 0; img < LANES; img++) s += coefs[img] * e8_frob(e8_basis(lane), (LANES - img) % LANES);
    images[lane] = s;
  }
  lap.mark("projection.transcript");
  DVec ehi, elo;
  eq_tables_device(c, outer, ehi, elo);
  DVec weights(plen), values(plen);
  lap.mark("projection.tables");
  // The weight images and then the sumcheck's first message (it needs the
  // folded values, not the folded root) run on the thread's second stream,
  // beside the block fold and the folded block's commit on c.st: the weights
  // are VALU work and the commit's leaf hash leaves HBM time free. c.st waits
  // for both before the sumcheck rounds.
  // (declared after the vectors the second stream reads: on an early exit it
  // drains that stream before they are released)
  struct SideWork {
    hipStream_t s2 = nullptr;
    hipEvent_t tables = nullptr, values = nullptr;
    dv::SideMsg msg;
    ~SideWork() {
      if (s2) (void)hipStreamSynchronize(s2);
      for (hipEvent_t e : {tables, values, msg.done})
        if (e) (void)hipEventDestroy(e);
    }
  } sw;
  const bool side = plen > 1;
  if (side) {
    sw.s2 = thread_stream2();
    QF_HIP(hipEventCreateWithFlags(&sw.tables, hipEventDisableTiming));
    QF_HIP(hipEventCreateWithFlags(&sw.values, hipEventDisableTiming));
    QF_HIP(hipEventRecord(sw.tables, c.st));
    QF_HIP(hipStreamWaitEvent(sw.s2, sw.tables, 0));
  }
  weight_images(c, weights.p, ehi, elo, images, sw.s2);
  tm.projection_ms = ms_since(t0);
  trace_phase("projection", tm.projection_ms);
  t0 = Clock::now();
  block_fold(c, cs.d_bva, p.layout.active, p.block_len, blocks, fw, values.p, plen);
  if (side) {
    QF_HIP(hipEventRecord(sw.values, c.st));
    QF_HIP(hipStreamWaitEvent(sw.s2, sw.values, 0));
    sw.msg = dv::sc_msg_launch(c, sw.s2, values.p, weights.p, plen / 2);
  }
  uint64_t R = 2 * (uint64_t(1) << p.geom.log_n);
  DCode folded = alloc_code(R);
  launch_commit_values(p.geom, values.p, folded.cw.p, p.alpha_encode(p.pp.codes.size() - 1), folded.tree, c.st,
                       side_leaves_on() ? sw.s2 : nullptr);
  QF_HIP(hipGetLastError());
  if (side) QF_HIP(hipStreamWaitEvent(c.st, sw.msg.done, 0));
  folded.root = download_root(c, folded.tree, R);
  t.bytes(folded.root.data(), 32);
  tm.folding_ms = ms_since(t0);
  trace_phase("folding", tm.folding_ms);
  t0 = Clock::now();
  std::vector<std::array<E8, 2>> rounds;
  std::vector<E8> reduced;
  E8 fval, fwt;
  std::pair<E8, E8> first_msg;
  if (side) first_msg = dv::sc_msg_finish(c, sw.msg);
  sumcheck(c, values, weights, claim, t, rounds, reduced, fval, fwt, side ? &first_msg : nullptr);
  require(fwt == batch_weight(outer.data(), n_outer, reduced, coefs) && claim == fval * fwt, QF_ERR_INVALID_COMMITMENT,
          "batch sumcheck terminal");
  tm.sumcheck_ms = ms_since(t0);
  trace_phase("sumcheck", tm.sumcheck_ms);
  t0 = Clock::now();
  const char* bsd = "authenticated-codeword-fold-queries/v1";
  t.bytes(bsd, std::strlen(bsd));
  t.field(fval);
  Digest seed = t.h.finish();
  const QaCode& top = p.pp.codes.back();
  uint64_t count = num_queries(top, p.pp.security_bits, R);
  bool all = count == R;
  std::vector<uint64_t> sel = sample_rows(count, R, seed);
  // the wire up to the binding, then the binding in place (its length patched after)
  const uint64_t rec = all ? 32 * R * sizeof(E8) : count * (16 + 32 * sizeof(E8) + 32 * uint64_t(ilog2u(R)));
  w.reserve(w.size() + (blocks + 1) * (rec + 32) + (4u << 20));
  put_tag(w, "QAPKBAT1");
  put_e8(w, claimed);
  for (auto& v : ev) put_e8(w, v);
  for (auto& v : pos) put_e8(w, v);
  for (auto& v : neg) put_e8(w, v);
  put_digest(w, folded.root);
  for (auto& r : rounds) { put_e8(w, r[0]); put_e8(w, r[1]); }
  put_e8(w, fval);
  const size_t bind_len_at = w.size();
  put_u64(w, 0);
  // The binding rows (queried rows and Merkle paths of every block and of the
  // folded block) depend on the transcript only through the row selection, and
  // the native opening and Decide only on the sumcheck point and the folded
  // code, so the binding runs on the thread's helper beside them: it appends
  // to w while this thread writes `opening` and `decide`, and w is not touched
  // here until the helper is done. Both only read the resident codewords.
  auto binding = [&](DevCtx& bx) {
    const auto tb = Clock::now();
    const size_t bind_at = w.size();
    put_tag(w, "QAFOLDB1");
    put_u64(w, blocks);
    DCode workspace;
    size_t prev_at = 0, prev_len = 0;
    for (uint64_t b = 0; b < blocks; b++) {
      if (b == 0 || cs.roots[b] != cs.roots[b - 1]) {
        DCode bc = block_code(bx, p, cs, b, workspace);
        prev_at = w.size();
        write_rows(bx, w, bc, all, sel);
        prev_len = w.size() - prev_at;
      } else {  // the same codeword as the previous block: its rows again
        const Bytes again(w.begin() + prev_at, w.begin() + prev_at + prev_len);
        prev_at = w.size();
        w.insert(w.end(), again.begin(), again.end());
      }
    }
    write_rows(bx, w, folded, all, sel);
    patch_u64(w, bind_len_at, w.size() - bind_at);
    tm.binding_ms = ms_since(tb);
    trace_phase("binding", tm.binding_ms);
  };
  SideThread* helper = side_binding_on() ? &side_thread() : nullptr;
  struct Join {  // on any exit the helper has finished with w and the codes first
    SideThread* s;
    ~Join() {
      if (s) try { s->wait(); } catch (...) {}
    }
  } join{nullptr};
  if (helper) {
    int dev = 0;
    QF_HIP(hipGetDevice(&dev));
    helper->start(dev, binding);
    join.s = helper;
  } else {
    binding(c);
  }
  t0 = Clock::now();
  Bytes opening;
  std::vector<Pending> pending;
  E8 actual = open_with_reduction(c, p, folded, reduced, opening, pending);
  require(actual == fval, QF_ERR_INVALID_COMMITMENT, "native opening value");
  tm.native_ms = ms_since(t0);
  trace_phase("native", tm.native_ms);
  t0 = Clock::now();
  Bytes decide = prove_decide(c, p, pending);
  tm.decide_ms = ms_since(t0);
  trace_phase("decide", tm.decide_ms);
  if (helper) {
    join.s = nullptr;
    helper->wait();  // rethrows an error of the binding
  }
  put_part(w, opening);
  put_part(w, decide);
}

// packed.rs:294-393 (one block)
Bytes open_single(DevCtx& c, const ParamsState& p, const CommittedState& cs, const std::vector<E8>& point,
                  const Bytes& context, E8& claimed, ProveTimings& tm) {
  auto t0 = Clock::now();
  size_t n_outer = point.size() - LOG_LANES;
  std::vector<E8> outer(point.begin(), point.begin() + n_outer);
  uint64_t plen = uint64_t(1) << n_outer;
  auto lanes = block_lanes(c, p, cs.d_bva, p.layout.active, 1, outer)[0];
  auto low = eq_host(point.data() + n_outer, LOG_LANES);
  claimed = E8{};
  for (int j = 0; j < LANES; j++) claimed += lanes[j] * low[j];
  auto ev = conjugates(lanes);
  require(projected_claim(ev, low) == claimed, QF_ERR_INVALID_COMMITMENT, "projected claim");
  Packed t = packed_transcript(p, cs.roots[0], point, claimed, context);
  for (auto& v : ev) t.field(v);
  auto powers = lane_powers(t.challenge());
  E8 claim{};
  for (int i = 0; i < LANES; i++) claim += powers[i] * ev[i];
  Lanes images;
  for (int j = 0; j < LANES; j++) {
    E8 s{};
    for (int i = 0; i < LANES; i++) s += powers[i] * e8_frob(e8_basis(j), (LANES - i) % LANES);
    images[j] = s;
  }
  DVec ehi, elo;
  eq_tables_device(c, outer, ehi, elo);
  DVec weights(plen), values(plen);
  weight_images(c, weights.p, ehi, elo, images);
  block_fold(c, cs.d_bva, p.layout.active, p.block_len, 1, {e8_one()}, values.p, plen);
  tm.projection_ms = ms_since(t0);
  trace_phase("projection", tm.projection_ms);
  t0 = Clock::now();
  std::vector<std::array<E8, 2>> rounds;
  std::vector<E8> reduced;
  E8 fval, fwt;
  sumcheck(c, values, weights, claim, t, rounds, reduced, fval, fwt);
  require(fwt == batch_weight(point.data(), n_outer, reduced, powers) && claim == fval * fwt, QF_ERR_INVALID_COMMITMENT,
          "sumcheck terminal");
  tm.sumcheck_ms = ms_since(t0);
  trace_phase("sumcheck", tm.sumcheck_ms);
  t0 = Clock::now();
  DCode workspace;
  DCode bc = block_code(c, p, cs, 0, workspace);
  Bytes opening;
  std::vector<Pending> pending;
  E8 actual = open_with_reduction(c, p, bc, reduced, opening, pending);
  require(actual == fval, QF_ERR_INVALID_COMMITMENT, "native opening value");
  tm.native_ms = ms_since(t0);
  trace_phase("native", tm.native_ms);
  t0 = Clock::now();
  Bytes decide = prove_decide(c, p, pending);
  tm.decide_ms = ms_since(t0);
  trace_phase("decide", tm.decide_ms);
  Bytes w;
  put_tag(w, PACKED_MAGIC);
  put_e8(w, claimed);
  for (auto& v : ev) put_e8(w, v);
  put_u64(w, rounds.size());
  for (auto& r : rounds) { put_e8(w, r[0]); put_e8(w, r[1]); }
  put_e8(w, fval);
  put_part(w, opening);
  put_part(w, decide);
  return w;
}

}  // namespace

std::string statement_json(const ParamsState& p, const CommittedState& c) {  // protocol.rs:33-40
  std::string s = "{\"version\":2,\"config\":{\"chunk_log\":" + std::to_string(p.cfg_chunk_log) +
                  ",\"packed_log\":" + std::to_string(p.packed_log) + ",\"query_rf_bits\":" +
                  std::to_string(p.query_bits) + "},\"shapes\":[";
  for (size_t i = 0; i < p.layout.shapes.size(); i++) {
    if (i) s += ',';
    s += "{\"rows\":" + std::to_string(p.layout.shapes[i].rows) + ",\"cols\":" + std::to_string(p.layout.shapes[i].cols) + "}";
  }
  s += "],\"parameter_digest\":";
  js_bytes(s, p.pp.parameter_digest.data(), 32);
  s += ",\"roots\":[";
  for (size_t i = 0; i < c.roots.size(); i++) { if (i) s += ','; js_bytes(s, c.roots[i].data(), 32); }
  return s + "]}";
}

std::string claims_json(const std::vector<ClaimIn>& claims) {  // tbv.rs:6-11
  std::string s = "[";
  for (size_t i = 0; i < claims.size(); i++) {
    if (i) s += ',';
    s += "{\"row\":" + js_scalars(claims[i].row) + ",\"col\":" + js_scalars(claims[i].col) + ",\"value\":";
    js_e8(s, claims[i].value);
    s += "}";
  }
  return s + "]";
}

// main.rs:188-209 with tbv.rs:143-185 evaluated on the device
std::vector<ClaimIn> benchmark_claims(const ParamsState& p, const CommittedState& cs) {
  Outer t;
  t.object("benchmark-points", statement_json(p, cs));
  std::vector<ClaimIn> claims;
  for (auto& sh : p.layout.shapes) {
    uint64_t rp = 1, cp = 1;
    while (rp < sh.rows) rp <<= 1;
    while (cp < sh.cols) cp <<= 1;
    ClaimIn cl;
    for (int i = 0; i < ilog2u(rp); i++) cl.row.push_back(t.challenge("row"));
    for (int i = 0; i < ilog2u(cp); i++) cl.col.push_back(t.challenge("col"));
    cl.value = E8{};
    claims.push_back(cl);
  }
  evaluate_claims(p, cs, claims);
  return claims;
}

// tbv.rs:143-185: values of the matrices at the claims' points, on the device BVA.
void evaluate_claims(const ParamsState& p, const CommittedState& cs, std::vector<ClaimIn>& claims) {
  require(claims.size() == p.layout.shapes.size(), QF_ERR_INVALID_PARAMETERS, "claim count mismatch");
  for (size_t i = 0; i < claims.size(); i++) {
    uint64_t rp = 1, cp = 1;
    while (rp < p.layout.shapes[i].rows) rp <<= 1;
    while (cp < p.layout.shapes[i].cols) cp <<= 1;
    require(claims[i].row.size() == size_t(ilog2u(rp)) && claims[i].col.size() == size_t(ilog2u(cp)),
            QF_ERR_INVALID_PARAMETERS, "point dimension mismatch");
    claims[i].value = E8{};
  }
  DevCtx c;
  Query q;
  for (auto& b : p.layout.buckets) {
    const ClaimIn& cl = claims[b.matrix];
    Term tm;
    for (int i = 0; i < ilog2u(b.cols); i++) tm.f.push_back({e8_one(), cl.col[i]});
    for (int i = 0; i < ilog2u(b.rows); i++) tm.f.push_back({e8_one(), cl.row[i]});
    tm.scale = monomial(b.row_offset, cl.row.data(), cl.row.size()) * monomial(b.col_offset, cl.col.data(), cl.col.size());
    tm.offset = b.offset;
    q.terms.push_back(std::move(tm));
  }
  std::vector<size_t> which;
  auto proj = project_terms(c, q, cs.d_bva, p.layout.active, 16, which);
  // a term's value is sum_i proj[i] low[i]; the buckets of one matrix share
  // their low factors, so their lane sums are added first and weighted once
  std::vector<bool> done(q.terms.size(), false);
  std::map<std::pair<size_t, std::string>, std::pair<size_t, std::array<E8, 16>>> groups;  // -> (a term, lane sums)
  for (size_t k = 0; k < which.size(); k++) {
    const Term& tm = q.terms[which[k]];
    const std::string key(reinterpret_cast<const char*>(tm.f.data()), 4 * sizeof(tm.f[0]));
    auto [it, fresh] = groups.try_emplace({p.layout.buckets[which[k]].matrix, key}, which[k], std::array<E8, 16>{});
    for (int i = 0; i < 16; i++) it->second.second[i] += proj[k][i];
    done[which[k]] = true;
  }
  for (auto& [mk, g] : groups) {
    const auto low = tbv_w(q.terms[g.first].f.data(), 4);
    E8 v{};
    for (int i = 0; i < 16; i++) v += g.second[i] * low[i];
    claims[mk.first].value += v;
  }
  for (size_t i = 0; i < q.terms.size(); i++) {
    const Term& tm = q.terms[i];
    if (done[i] || tm.len() >= 16) continue;  // big terms past the active data project to zero
    auto vals = read_values(c, cs.d_bva, p.layout.active, tm.offset, tm.len());
    E8 v{};
    for (uint64_t j = 0; j < tm.len(); j++) v += tm.weight(j) * small_val(vals[j]);
    claims[p.layout.buckets[i].matrix].value += v;
  }
}

Bytes prove(const ParamsState& p, const CommittedState& cs, const std::vector<ClaimIn>& claims, ProveTimings& tm) {
  auto start = Clock::now();
  require(claims.size() == p.layout.shapes.size(), QF_ERR_INVALID_PARAMETERS, "claim count mismatch");
  for (size_t i = 0; i < claims.size(); i++) {
    uint64_t rp = 1, cp = 1;
    while (rp < p.layout.shapes[i].rows) rp <<= 1;
    while (cp < p.layout.shapes[i].cols) cp <<= 1;
    require(claims[i].row.size() == size_t(ilog2u(rp)) && claims[i].col.size() == size_t(ilog2u(cp)),
            QF_ERR_INVALID_PARAMETERS, "point dimension mismatch");
  }
  require(cs.d_bva != nullptr, QF_ERR_ARGUMENT, "commitment has no resident BVA");
  uint64_t s = uint64_t(1) << p.packed_log;
  DevCtx c;
  auto t0 = Clock::now();
  Outer tr;
  tr.object("collection", statement_json(p, cs));
  tr.object("claims", claims_json(claims));
  Query q = compile(p.layout, claims, tr);
  tm.conversion_ms = ms_since(t0);
  trace_phase("conversion", tm.conversion_ms);
  t0 = Clock::now();
  EoOut eo = eo_prove(c, p.layout, cs.d_bva, s, q, tr);
  tm.eo_ms = ms_since(t0);
  trace_phase("eo", tm.eo_ms);
  Bytes context;
  const char* cx = TERMINAL_DOMAIN;
  context.insert(context.end(), cx, cx + std::strlen(cx));
  context.insert(context.end(), tr.state.begin(), tr.state.end());
  put_u64(context, 0);
  Bytes out;  // bincode fixint (protocol.rs:57-66, 89-95)
  put_u64(out, eo.pos.size());
  for (auto& v : eo.pos) put_e8(out, v);
  put_u64(out, eo.neg.size());
  for (auto& v : eo.neg) put_e8(out, v);
  put_u64(out, eo.rounds.size());
  for (auto& r : eo.rounds) { put_e8(out, r[0]); put_e8(out, r[1]); }
  put_e8(out, eo.value);
  put_e8(out, eo.query);
  put_u64(out, 1);
  // the opening's value and the wire's length follow; both are patched once the
  // opening, appended in place after them, is done
  const size_t actual_at = out.size();
  put_e8(out, E8{});
  const size_t wire_len_at = out.size();
  put_u64(out, 0);
  const size_t wire_at = out.size();
  E8 actual;
  if (cs.roots.size() > 1) {
    open_batch(c, p, cs, eo.point, context, actual, tm, out);
  } else {
    const Bytes wire = open_single(c, p, cs, eo.point, context, actual, tm);
    out.insert(out.end(), wire.begin(), wire.end());
  }
  require(actual * eo.scale == eo.value, QF_ERR_INVALID_COMMITMENT, "QAFold/EO terminal mismatch");
  std::memcpy(out.data() + actual_at, &actual, sizeof(E8));
  patch_u64(out, wire_len_at, out.size() - wire_at);
  require(out.size() <= (256ull << 20), QF_ERR_ENCODING, "oversize ParComm proof");
  tm.total_ms = ms_since(start);
  return out;
}

}  // namespace qf
