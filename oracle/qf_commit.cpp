// CPU oracle (test infrastructure only; see qf_core.hpp): commitment path.
//
// Pinning: the layout is pinned by the reference's llama8b layout fixture
// (qafold/f65521/tests/fixtures/llama8b-layout.json, checked by
// tests/test_layout.py). WHT and Mobius are pinned by closed forms
// (WHT o WHT = n I, Mobius inverse = zeta transform, brute force on tiny
// inputs in tests/test_oracle_core.py). Merkle leaves/inner nodes are pinned by
// hashlib.blake2s over the framed bytes. The QA generator sampling, setup
// commitments and parameter digest are pinned only end to end, through the
// reference's recorded demo proof digest.
#include "qf_commit.hpp"
#include <algorithm>
#include <tuple>

namespace orc {

const char* LEAF_BASE_DOMAIN = "qafold-short/base-row/v1";
const char* LEAF_EXTENSION_DOMAIN = "qafold-short/extension-row/v1";
static const char* INNER_DOMAIN = "qafold-short/merkle-inner/v1";

static void modulus_frame(Blake2s& h) { h.frame_u64("modulus", P); }

// ---------------------------------------------------------------- layout
// layout.rs:34-54
std::vector<std::pair<u64, u64>> intervals(u64 length, size_t budget) {
  std::vector<u64> sizes;
  u64 remaining = length;
  while (remaining > 0 && sizes.size() < budget) {
    u64 s = u64(1) << ilog2(remaining);
    sizes.push_back(s);
    remaining -= s;
  }
  if (remaining > 0) sizes.back() *= 2;
  std::vector<std::pair<u64, u64>> out;
  u64 off = 0;
  for (u64 s : sizes) { out.push_back({off, s}); off += s; }
  return out;
}

// layout.rs:56-114
Layout layout_new(const std::vector<Shape>& shapes) {
  check(!shapes.empty() && shapes.size() <= 65536, "invalid matrix count");
  Layout l;
  l.shapes = shapes;
  for (size_t m = 0; m < shapes.size(); m++) {
    const Shape& s = shapes[m];
    check(s.rows > 0 && s.cols > 0 && s.rows <= (u64(1) << 30) && s.cols <= (u64(1) << 30), "invalid matrix shape");
    for (auto [ro, rows] : intervals(s.rows, 4))
      for (auto [co, cols] : intervals(s.cols, 4)) {
        check(rows * cols <= (u64(1) << 37), "matrix exceeds supported capacity");
        l.buckets.push_back({m, ro, co, rows, cols, std::min(rows, s.rows - ro), std::min(cols, s.cols - co), 0});
      }
  }
  std::stable_sort(l.buckets.begin(), l.buckets.end(), [](const Bucket& a, const Bucket& b) {
    return std::make_tuple(-(int64_t)a.len(), a.matrix, a.row_offset, a.col_offset) <
           std::make_tuple(-(int64_t)b.len(), b.matrix, b.row_offset, b.col_offset);
  });
  u64 active = 0;
  for (auto& b : l.buckets) {
    b.offset = active;
    active += b.len();
    check(active <= (u64(1) << 37), "collection exceeds supported capacity");
    check(b.offset % b.len() == 0, "unaligned BVA bucket");
  }
  l.active = active;
  u64 lg = std::max<u64>(active, 1024);
  u64 p2 = 1;
  while (p2 < lg) p2 <<= 1;
  l.logical = p2;
  return l;
}

// layout.rs:115-135
std::vector<Val> layout_encode(const Layout& l, const std::vector<std::vector<Val>>& m) {
  check(m.size() == l.shapes.size(), "matrix count mismatch");
  for (size_t i = 0; i < m.size(); i++) check(m[i].size() == l.shapes[i].rows * l.shapes[i].cols, "matrix length mismatch");
  std::vector<Val> v(l.active, 0);
  for (auto& b : l.buckets)
    for (u64 r = 0; r < b.physical_rows; r++) {
      u64 src = (b.row_offset + r) * l.shapes[b.matrix].cols + b.col_offset;
      u64 dst = b.offset + r * b.cols;
      std::memcpy(&v[dst], &m[b.matrix][src], b.physical_cols);
    }
  return v;
}

// layout.rs:136-155
void layout_validate(const Layout& l, const Val* values, u64 n) {
  check(n == l.active, "active BVA length mismatch");
  for (auto& b : l.buckets)
    for (u64 r = 0; r < b.rows; r++) {
      u64 valid = r < b.physical_rows ? b.physical_cols : 0;
      for (u64 i = b.offset + r * b.cols + valid; i < b.offset + (r + 1) * b.cols; i++) check(values[i] == 0, "nonzero boundary padding");
    }
}

// ---------------------------------------------------------------- transforms
// qafold_short.rs:1911-2015 (any stage order gives the same canonical result)
void wht(E8* v, size_t n) {
  check(is_pow2(n), "wht length");
  for (size_t h = 1; h < n; h *= 2)
    for (size_t b = 0; b < n; b += 2 * h)
      for (size_t i = b; i < b + h; i++) { E8 a = v[i], c = v[i + h]; v[i] = a + c; v[i + h] = a - c; }
}
// packed.rs:708-752
void mobius(E8* v, size_t n) {
  for (size_t h = 1; h < n; h *= 2)
    for (size_t b = 0; b < n; b += 2 * h)
      for (size_t i = b; i < b + h; i++) v[i + h] -= v[i];
}

// ---------------------------------------------------------------- QA code
// qafold_short.rs:2211-2217
static Digest derive_level_seed(const Digest& seed, size_t level) {
  Blake2s h;
  h.frame("domain", std::string("qafold-short/qa-level-seed/v1"));
  h.frame("master-seed", seed.data(), 32);
  h.frame_u64("level", level);
  return h.finalize();
}

// qafold_short.rs:2219-2281, uniform (not nonzero-frequency) branch, 16-byte lanes
static E8 sample_generator_element(const Digest& seed, size_t message_len, size_t index, size_t block, size_t row) {
  for (u64 attempt = 0;; attempt++) {
    Blake2s h;
    h.frame("domain", std::string("qafold-short/qa-generator-coefficient/v1"));
    h.frame("field-name", std::string(FIELD_NAME));
    modulus_frame(h);
    h.frame("seed", seed.data(), 32);
    h.frame_u64("message-length", message_len);
    h.frame_u64("index", index);
    h.frame_u64("block", block);
    h.frame_u64("row", row);
    h.frame_u64("attempt", attempt);
    Digest d = h.finalize();
    for (int lane = 0; lane < 2; lane++) {
      E8 v;
      if (e8_from_uniform(d.data() + 16 * lane, v)) return v;
    }
  }
}

// qafold_short.rs:2283-2304
static Digest qa_generator_digest(const std::vector<std::vector<E8>>& alphas) {
  Blake2s h;
  h.frame("domain", std::string("qafold-short/qa-generator/v1"));
  h.frame("field-name", std::string(FIELD_NAME));
  modulus_frame(h);
  h.frame_u64("columns", alphas.size());
  h.frame_u64("rows", alphas.empty() ? 0 : alphas[0].size());
  for (size_t c = 0; c < alphas.size(); c++) {
    h.frame_u64("column", c);
    for (auto& v : alphas[c]) { u8 b[16]; e8_write(v, b); h.frame("coefficient", b, 16); }
  }
  return h.finalize();
}

// qafold_short.rs:699-801 (sample_systematic -> from_group_algebra)
QaCode qa_sample_systematic(size_t n, size_t index, const Digest& seed) {
  check(is_pow2(n) && index > 1 && is_pow2(index), "qa code shape");
  std::vector<std::vector<E8>> cols(index, std::vector<E8>(n, e8_zero()));
  cols[0][0] = e8_one();
  for (size_t b = 1; b < index; b++)
    for (size_t r = 0; r < n; r++) cols[b][r] = sample_generator_element(seed, n, index, b, r);
  E8 inv_n = e8_inv(e8_from_u64(n));
  for (auto& c : cols) { wht(c.data(), n); for (auto& v : c) v = v * inv_n; }
  QaCode code;
  code.message_len = n;
  code.index = index;
  code.alphas = std::move(cols);
  code.generator_digest = qa_generator_digest(code.alphas);
  for (auto& v : code.alphas[0]) check(v == inv_n, "non-systematic first block");
  return code;
}

// Systematic index-two encoding: [m | WHT(alpha_1 . WHT(m))] (qafold_short.rs:969-1000).
void QaCode::encode(const E8* m, E8* out) const {
  check(index == 2, "only index-two codes occur in this configuration");
  std::memcpy(out, m, message_len * sizeof(E8));
  E8* s = out + message_len;
  std::memcpy(s, m, message_len * sizeof(E8));
  wht(s, message_len);
  for (size_t i = 0; i < message_len; i++) s[i] = s[i] * alphas[1][i];
  wht(s, message_len);
}

// qafold_short.rs:2306-2325
static Digest qa_sequence_generator_digest(const std::vector<QaCode>& codes) {
  Blake2s h;
  h.frame("domain", std::string("qafold-short/qa-sequence-generators/v1"));
  h.frame_u64("levels", codes.size());
  for (size_t l = 0; l < codes.size(); l++) {
    h.frame_u64("level", l);
    h.frame_u64("message-length", codes[l].message_len);
    h.frame_u64("index", codes[l].index);
    h.frame("generator", codes[l].generator_digest.data(), 32);
  }
  return h.finalize();
}

// qafold_short.rs:1672-1734 (EXTENSION_DEGREE = 1 for the full octic field)
QaSequence qa_sequence_new(size_t witness, size_t ext_cols, size_t recursive_index, const Digest& seed) {
  check(is_pow2(witness) && ext_cols > 1 && is_pow2(ext_cols) && witness % ext_cols == 0, "qa sequence shape");
  size_t ml = witness / ext_cols;
  std::vector<size_t> desc;
  for (;;) {
    desc.push_back(ml);
    if (ml < ext_cols) break;
    check(ml % ext_cols == 0, "qa sequence divisibility");
    ml /= ext_cols;
  }
  std::reverse(desc.begin(), desc.end());
  QaSequence s;
  s.witness = witness;
  s.ext_cols = ext_cols;
  s.rf_layers = size_t(ilog2(ext_cols));  // brakefold.rs:37-51
  s.recursive_index = recursive_index;
  for (size_t l = 0; l < desc.size(); l++) {
    size_t idx = (desc.size() - l <= 1) ? 2 : recursive_index;  // qafold_short.rs:43-56
    s.codes.push_back(qa_sample_systematic(desc[l], idx, derive_level_seed(seed, l)));
  }
  s.generator_digest = qa_sequence_generator_digest(s.codes);
  return s;
}

// ---------------------------------------------------------------- Merkle
Digest hash_inner(const Digest& l, const Digest& r) {
  Blake2s h;
  h.frame("domain", std::string(INNER_DOMAIN));
  h.update(l.data(), 32);
  h.update(r.data(), 32);
  return h.finalize();
}
Merkle merkle_from_leaves(std::vector<Digest> leaves) {
  check(!leaves.empty() && is_pow2(leaves.size()), "merkle leaf count");
  Merkle m;
  m.levels.push_back(std::move(leaves));
  while (m.levels.back().size() > 1) {
    auto& c = m.levels.back();
    std::vector<Digest> p(c.size() / 2);
#pragma omp parallel for schedule(static) if (p.size() >= 4096)
    for (size_t i = 0; i < p.size(); i++) p[i] = hash_inner(c[2 * i], c[2 * i + 1]);
    m.levels.push_back(std::move(p));
  }
  return m;
}
std::vector<Digest> Merkle::path(size_t index) const {
  check(index < levels[0].size(), "merkle index");
  std::vector<Digest> s;
  for (size_t l = 0; l + 1 < levels.size(); l++) { s.push_back(levels[l][index ^ 1]); index >>= 1; }
  return s;
}
// qafold_short_pc.rs:2437-2475
std::vector<Digest> hash_rows(const E8* entries, size_t rows, size_t cols, const char* domain) {
  std::vector<Digest> out(rows);
  Blake2s prefix;
  prefix.frame("domain", std::string(domain));
  prefix.frame_u64("width", cols);
#pragma omp parallel for schedule(static) if (rows >= 4096)
  for (size_t r = 0; r < rows; r++) {
    Blake2s h = prefix;
    for (size_t k = 0; k < cols; k++) { u8 b[16]; e8_write(entries[k * rows + r], b); h.update(b, 16); }
    out[r] = h.finalize();
  }
  return out;
}

// ---------------------------------------------------------------- setup (qafold_short_setup.rs)
static void field_metadata(Blake2s& h) {  // qafold_short_setup.rs:873-887
  h.frame("field-name", std::string(FIELD_NAME));
  modulus_frame(h);
  h.frame_u64("canonical-bytes", 16);
  h.frame_u64("extension-degree", 1);
}
static void setup_shape_frames(Blake2s& h, const SetupShape& s) {  // :838-855
  h.frame_u64("level", s.level);
  h.frame_u64("encoded-rows", s.encoded_rows);
  h.frame_u64("message-rows", s.message_rows);
  h.frame_u64("t-columns", s.t_cols);
  h.frame_u64("t-elements", s.t_elements);
  h.frame_u64("t-last-length", s.t_last_len);
  h.frame_u64("h-columns", s.h_cols);
}
// :770-836 (DirectInterleavedV1 row hasher + raw canonical entries)
Digest setup_row_hash(const SetupShape& s, size_t row, const E8* row_vals) {
  Blake2s h;
  h.frame("domain", std::string("qafold-short/setup-base-row/v1"));
  field_metadata(h);
  h.frame("decide-setup-mode", std::string("direct-interleaved-v1"));
  setup_shape_frames(h, s);
  h.frame_u64("row", row);
  for (size_t c = 0; c < s.t_cols + s.h_cols; c++) { u8 b[16]; e8_write(row_vals[c], b); h.update(b, 16); }
  return h.finalize();
}

// qafold_short_pc.rs:2271-2314
static Digest parameter_digest(const PcsParams& pp) {
  Blake2s h;
  h.frame("domain", std::string("qafold-short-parameters/v4"));
  h.frame("field-name", std::string(FIELD_NAME));
  modulus_frame(h);
  h.frame_u64("extension-degree", 1);
  h.frame_u64("witness", pp.n_witness);
  h.frame_u64("extension-columns", pp.ext_cols);
  h.frame_u64("index", pp.index);
  h.frame_u64("security", pp.security_bits);
  h.frame_u64("rf-layers", pp.seq.rf_layers);
  h.frame("generator-sequence", pp.seq.generator_digest.data(), 32);
  if (pp.seq.recursive_index != 4) h.frame_u64("recursive-index-override", pp.seq.recursive_index);
  h.frame("derived-rfc-sequence", pp.derived_digest.data(), 32);
  h.frame("setup-roots-and-shapes", pp.setup_digest.data(), 32);
  return h.finalize();
}

// qafold_short_pc.rs:401-443 with qafold_short_setup.rs:203-221, 427-537, 655-741
PcsParams pcs_setup(size_t n_witness, size_t ext_cols, size_t recursive_index, size_t security_bits, const u8 rng_seed[32]) {
  ChaChaRng rng(rng_seed, 12);
  return pcs_setup_rng(n_witness, ext_cols, recursive_index, security_bits, rng);
}

PcsParams pcs_setup_rng(size_t n_witness, size_t ext_cols, size_t recursive_index, size_t security_bits, ChaChaRng& rng) {
  PcsParams pp;
  pp.n_witness = n_witness;
  pp.ext_cols = ext_cols;
  pp.recursive_index = recursive_index;
  pp.security_bits = security_bits;
  Digest seed;
  rng.fill_bytes(seed.data(), 32);
  pp.seq = qa_sequence_new(n_witness, ext_cols, recursive_index, seed);
  size_t L = pp.seq.codes.size();
  {  // derived parameters: DirectInterleavedV1 has no RFC sources
    Blake2s h;
    h.frame("domain", std::string("qafold-short/derived-rfc-parameters/v1"));
    field_metadata(h);
    h.frame("decide-setup-mode", std::string("direct-interleaved-v1"));
    h.frame_u64("levels", L);
    h.frame_u64("switch-ratio", ext_cols);
    for (size_t l = 1; l < L; l++) { h.frame_u64("level", l); u8 z = 0; h.frame("rfc-source-present", &z, 1); }
    pp.derived_digest = h.finalize();
  }
  pp.setup.resize(L);
  for (size_t level = 1; level < L; level++) {
    const QaCode& target = pp.seq.codes[level - 1];
    const QaCode& cur = pp.seq.codes[level];
    size_t mr = target.message_len, er = target.codeword_len(), rf = ext_cols;
    check(cur.message_len == mr * rf, "setup row factor");
    SetupLevel& sl = pp.setup[level];
    sl.shape = {level, er, mr, 0, 0, 0, cur.alphas.size() * rf};
    sl.encoded.assign(er * sl.shape.h_cols, e8_zero());
    for (size_t c = 0; c < sl.shape.h_cols; c++) {
      const auto& alpha = cur.alphas[c / rf];
      target.encode(alpha.data() + (c % rf) * mr, sl.encoded.data() + c * er);
    }
    std::vector<Digest> leaves(er);
#pragma omp parallel for schedule(static) if (er >= 4096)
    for (size_t r = 0; r < er; r++) {
      std::vector<E8> row(sl.shape.h_cols);
      for (size_t c = 0; c < sl.shape.h_cols; c++) row[c] = sl.encoded[c * er + r];
      leaves[r] = setup_row_hash(sl.shape, r, row.data());
    }
    sl.tree = merkle_from_leaves(std::move(leaves));
  }
  {
    Blake2s h;
    h.frame("domain", std::string("qafold-short/setup-aggregate/v1"));
    field_metadata(h);
    h.frame_u64("levels", L);
    for (size_t l = 0; l < L; l++) {
      h.frame_u64("slot", l);
      u8 present = l == 0 ? 0 : 1;
      h.frame("present", &present, 1);
      if (l > 0) {
        setup_shape_frames(h, pp.setup[l].shape);
        Digest r = pp.setup[l].tree.root();
        h.frame("root", r.data(), 32);
      }
    }
    pp.setup_digest = h.finalize();
  }
  pp.parameter_digest = parameter_digest(pp);
  return pp;
}

// ---------------------------------------------------------------- commit
// qafold_short_pc.rs:461-549
Committed pcs_commit_owned(const PcsParams& pp, std::vector<E8> witness) {
  const QaCode& top = pp.seq.codes.back();
  size_t n = top.message_len, R = top.codeword_len(), cols = pp.ext_cols;
  check(witness.size() == n * cols && witness.size() == pp.n_witness, "witness length");
  Committed c;
  c.rows = R;
  c.cols = cols;
  c.encoded.assign(R * cols, e8_zero());
#pragma omp parallel for schedule(dynamic) if (n >= 4096)
  for (size_t k = 0; k < cols; k++) top.encode(witness.data() + k * n, c.encoded.data() + k * R);
  c.coeffs = std::move(witness);
  c.tree = merkle_from_leaves(hash_rows(c.encoded.data(), R, cols, LEAF_BASE_DOMAIN));
  return c;
}

// field.rs:24-26 and packed.rs:220-243: lane j of element i = entry LANES*i + j
// (mod p), entries past `available` are zero.
std::vector<E8> pack_block(const Val* values, u64 available, u64 block_len) {
  std::vector<E8> out(block_len / LANES);
  for (u64 i = 0; i < block_len / LANES; i++)
    for (int j = 0; j < LANES; j++) {
      u64 idx = LANES * i + j;
      out[i].c[j] = fsmall(idx < available ? values[idx] : 0);
    }
  return out;
}

Committed packed_commit(const PcsParams& pp, const std::vector<E8>& packed) {
  std::vector<E8> coeffs = packed;
  mobius(coeffs.data(), coeffs.size());
  return pcs_commit_owned(pp, std::move(coeffs));
}

}  // namespace orc
