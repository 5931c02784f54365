// Host computation of the public parameters (see qf_params.hpp).
#include "qf_params.hpp"
#include <algorithm>
#include <thread>
#include <tuple>

namespace qf {


void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)>& body) {
  unsigned hw = std::thread::hardware_concurrency();
  uint64_t workers = std::min<uint64_t>(hw ? hw : 1, 16);
  if (n < 4096 || workers <= 1) { body(0, n); return; }
  std::vector<std::thread> ts;
  uint64_t chunk = (n + workers - 1) / workers;
  for (uint64_t w = 0; w < workers; w++) {
    uint64_t a = w * chunk, b = std::min(n, a + chunk);
    if (a < b) ts.emplace_back([&, a, b] { body(a, b); });
  }
  for (auto& t : ts) t.join();
}

// ---------------------------------------------------------------- layout (layout.rs:34-114)
static std::vector<std::pair<uint64_t, uint64_t>> dyadic_split(uint64_t length) {
  std::vector<uint64_t> sizes;
  uint64_t rem = length;
  while (rem > 0 && sizes.size() < 4) { uint64_t s = uint64_t(1) << ilog2u(rem); sizes.push_back(s); rem -= s; }
  if (rem > 0) sizes.back() *= 2;
  std::vector<std::pair<uint64_t, uint64_t>> out;
  uint64_t off = 0;
  for (uint64_t s : sizes) { out.emplace_back(off, s); off += s; }
  return out;
}

Layout make_layout(const std::vector<Shape>& shapes) {
  require(!shapes.empty() && shapes.size() <= 65536, QF_ERR_INVALID_PARAMETERS, "invalid matrix count");
  Layout l;
  l.shapes = shapes;
  for (uint64_t m = 0; m < shapes.size(); m++) {
    const Shape& s = shapes[m];
    require(s.rows > 0 && s.cols > 0 && s.rows <= (1ull << 30) && s.cols <= (1ull << 30), QF_ERR_INVALID_PARAMETERS, "invalid matrix shape");
    for (auto& r : dyadic_split(s.rows))
      for (auto& c : dyadic_split(s.cols)) {
        require(r.second * c.second <= (1ull << 37), QF_ERR_INVALID_PARAMETERS, "matrix exceeds supported capacity");
        l.buckets.push_back({m, r.first, c.first, r.second, c.second, std::min(r.second, s.rows - r.first), std::min(c.second, s.cols - c.first), 0});
      }
  }
  std::stable_sort(l.buckets.begin(), l.buckets.end(), [](const Bucket& a, const Bucket& b) {
    if (a.len() != b.len()) return a.len() > b.len();
    return std::tie(a.matrix, a.row_offset, a.col_offset) < std::tie(b.matrix, b.row_offset, b.col_offset);
  });
  for (auto& b : l.buckets) {
    b.offset = l.active;
    l.active += b.len();
    require(l.active <= (1ull << 37), QF_ERR_INVALID_PARAMETERS, "collection exceeds supported capacity");
    require(b.offset % b.len() == 0, QF_ERR_INVALID_PARAMETERS, "unaligned BVA bucket");
  }
  uint64_t m = std::max<uint64_t>(l.active, 1024);
  l.logical = 1;
  while (l.logical < m) l.logical <<= 1;
  return l;
}

// ---------------------------------------------------------------- QA codes
void host_wht(E8* v, uint64_t n) {
  for (uint64_t h = 1; h < n; h *= 2) {
    parallel_for(n / 2, [&](uint64_t a, uint64_t b) {
      for (uint64_t p = a; p < b; p++) {
        uint64_t i = (p / h) * 2 * h + (p % h);
        V8 x = v8_load(v[i]), y = v8_load(v[i + h]);
        v[i] = v8_store(v8_add(x, y));
        v[i + h] = v8_store(v8_sub(x, y));
      }
    });
  }
}

// [m | WHT(alpha_1 . WHT(m))], the systematic index-two encoding (qafold_short.rs:969-1000)
void host_encode(const QaCode& code, const E8* m, E8* out) {
  uint64_t n = code.message_len;
  std::memcpy(out, m, n * 16);
  E8* s = out + n;
  std::memcpy(s, m, n * 16);
  host_wht(s, n);
  for (uint64_t i = 0; i < n; i++) s[i] = e8_mul(s[i], code.alpha1[i]);
  host_wht(s, n);
}

static Digest level_seed(const Digest& seed, uint64_t level) {  // qafold_short.rs:2211-2217
  Hasher h;
  h.frame("domain", "qafold-short/qa-level-seed/v1");
  h.frame("master-seed", seed.data(), 32);
  h.frame_u64("level", level);
  return h.finish();
}

// One uniformly sampled generator coefficient (qafold_short.rs:2219-2281; the
// full octic field takes 16-byte canonical candidates, two per digest).
static E8 generator_element(const Hasher& prefix, uint64_t row) {
  for (uint64_t attempt = 0;; attempt++) {
    Hasher h = prefix;
    h.frame_u64("row", row);
    h.frame_u64("attempt", attempt);
    Digest d = h.finish();
    E8 v;
    if (e8_from_uniform(d.data(), v) || e8_from_uniform(d.data() + 16, v)) return v;
  }
}

static QaCode sample_code(uint64_t n, uint64_t index, const Digest& seed) {
  require(index == 2, QF_ERR_ENCODING, "only index-two QA codes are configured");
  QaCode c;
  c.message_len = n;
  c.index = index;
  Hasher prefix;
  prefix.frame("domain", "qafold-short/qa-generator-coefficient/v1");
  prefix.frame("field-name", FIELD_NAME);
  prefix.frame_u64("modulus", P);
  prefix.frame("seed", seed.data(), 32);
  prefix.frame_u64("message-length", n);
  prefix.frame_u64("index", index);
  prefix.frame_u64("block", 1);
  c.alpha1.resize(n);
  parallel_for(n, [&](uint64_t a, uint64_t b) { for (uint64_t r = a; r < b; r++) c.alpha1[r] = generator_element(prefix, r); });
  host_wht(c.alpha1.data(), n);
  E8 inv_n = e8_inv(e8(uint32_t(n % P)));
  parallel_for(n, [&](uint64_t a, uint64_t b) { for (uint64_t r = a; r < b; r++) c.alpha1[r] = e8_mul(c.alpha1[r], inv_n); });
  // Generator digest over alphas[0] = (1/n, ..., 1/n) and alphas[1] (qafold_short.rs:2283-2304).
  Hasher g;
  g.frame("domain", "qafold-short/qa-generator/v1");
  g.frame("field-name", FIELD_NAME);
  g.frame_u64("modulus", P);
  g.frame_u64("columns", index);
  g.frame_u64("rows", n);
  g.frame_u64("column", 0);
  for (uint64_t r = 0; r < n; r++) g.frame("coefficient", &inv_n, 16);
  g.frame_u64("column", 1);
  for (uint64_t r = 0; r < n; r++) g.frame("coefficient", &c.alpha1[r], 16);
  c.generator_digest = g.finish();
  return c;
}

static void field_metadata(Hasher& h) {  // qafold_short_setup.rs:873-887
  h.frame("field-name", FIELD_NAME);
  h.frame_u64("modulus", P);
  h.frame_u64("canonical-bytes", 16);
  h.frame_u64("extension-degree", 1);
}
static void shape_frames(Hasher& h, const SetupShape& s) {  // qafold_short_setup.rs:838-855
  h.frame_u64("level", s.level);
  h.frame_u64("encoded-rows", s.encoded_rows);
  h.frame_u64("message-rows", s.message_rows);
  h.frame_u64("t-columns", s.t_cols);
  h.frame_u64("t-elements", s.t_elements);
  h.frame_u64("t-last-length", s.t_last_len);
  h.frame_u64("h-columns", s.h_cols);
}

Digest hash_inner(const Digest& l, const Digest& r) {
  Hasher h;
  h.frame("domain", "qafold-short/merkle-inner/v1");
  h.update(l.data(), 32);
  h.update(r.data(), 32);
  return h.finish();
}
std::vector<std::vector<Digest>> merkle_levels(std::vector<Digest> leaves) {
  require(pow2(leaves.size()), QF_ERR_ENCODING, "merkle leaf count");
  std::vector<std::vector<Digest>> lv;
  lv.push_back(std::move(leaves));
  while (lv.back().size() > 1) {
    const auto& c = lv.back();
    std::vector<Digest> p(c.size() / 2);
    parallel_for(p.size(), [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) p[i] = hash_inner(c[2 * i], c[2 * i + 1]); });
    lv.push_back(std::move(p));
  }
  return lv;
}

PcsParams make_pcs_params(int log_n, uint64_t security_bits, const uint8_t seed32[32]) {
  require(log_n >= 4 && log_n <= 30, QF_ERR_INVALID_VARIABLES, "packed log_n out of range");
  PcsParams pp;
  pp.n_witness = 1ull << (log_n - LOG_LANES);
  pp.security_bits = security_bits;
  require(pp.n_witness % pp.ext_cols == 0, QF_ERR_ENCODING, "witness not divisible by columns");
  // StdRng::from_seed(seed).fill_bytes(32): the first 32 bytes of the ChaCha12
  // keystream, block counter 0, stream 0 (rand_chacha 0.3 word order).
  Digest seq_seed;
  {
    uint32_t s[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};
    for (int i = 0; i < 8; i++) s[4 + i] = seed32[4 * i] | uint32_t(seed32[4 * i + 1]) << 8 | uint32_t(seed32[4 * i + 2]) << 16 | uint32_t(seed32[4 * i + 3]) << 24;
    for (int i = 12; i < 16; i++) s[i] = 0;
    uint32_t x[16];
    std::memcpy(x, s, sizeof x);
    auto rl = [](uint32_t v, int n) { return (v << n) | (v >> (32 - n)); };
    auto qr = [&](int a, int b, int c, int d) {
      x[a] += x[b]; x[d] = rl(x[d] ^ x[a], 16); x[c] += x[d]; x[b] = rl(x[b] ^ x[c], 12);
      x[a] += x[b]; x[d] = rl(x[d] ^ x[a], 8); x[c] += x[d]; x[b] = rl(x[b] ^ x[c], 7);
    };
    for (int r = 0; r < 12; r += 2) { qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15); qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14); }
    for (int i = 0; i < 8; i++) { uint32_t w = x[i] + s[i]; for (int j = 0; j < 4; j++) seq_seed[4 * i + j] = uint8_t(w >> (8 * j)); }
  }
  // ShortQaFold::new_with_recursive_index (qafold_short.rs:1672-1734)
  std::vector<uint64_t> lens;
  for (uint64_t ml = pp.n_witness / pp.ext_cols;; ml /= pp.ext_cols) {
    lens.push_back(ml);
    if (ml < pp.ext_cols) break;
  }
  std::reverse(lens.begin(), lens.end());
  for (uint64_t l = 0; l < lens.size(); l++) pp.codes.push_back(sample_code(lens[l], 2, level_seed(seq_seed, l)));
  {
    Hasher h;  // qafold_short.rs:2306-2325
    h.frame("domain", "qafold-short/qa-sequence-generators/v1");
    h.frame_u64("levels", pp.codes.size());
    for (uint64_t l = 0; l < pp.codes.size(); l++) {
      h.frame_u64("level", l);
      h.frame_u64("message-length", pp.codes[l].message_len);
      h.frame_u64("index", pp.codes[l].index);
      h.frame("generator", pp.codes[l].generator_digest.data(), 32);
    }
    pp.sequence_digest = h.finish();
  }
  uint64_t L = pp.codes.size();
  {
    Hasher h;  // qafold_short_setup.rs:655-712, DirectInterleavedV1
    h.frame("domain", "qafold-short/derived-rfc-parameters/v1");
    field_metadata(h);
    h.frame("decide-setup-mode", "direct-interleaved-v1");
    h.frame_u64("levels", L);
    h.frame_u64("switch-ratio", pp.ext_cols);
    for (uint64_t l = 1; l < L; l++) { h.frame_u64("level", l); uint8_t z = 0; h.frame("rfc-source-present", &z, 1); }
    pp.derived_digest = h.finish();
  }
  // Direct-interleaved setup commitments: H_j columns encoded under G_{j-1}
  // (qafold_short_setup.rs:427-537), merged-row leaves (:770-836).
  pp.setup.resize(L);
  for (uint64_t level = 1; level < L; level++) {
    const QaCode& target = pp.codes[level - 1];
    const QaCode& cur = pp.codes[level];
    uint64_t mr = target.message_len, er = target.codeword_len(), rf = pp.ext_cols;
    require(cur.message_len == mr * rf, QF_ERR_ENCODING, "setup row factor");
    SetupLevel& sl = pp.setup[level];
    sl.shape = {level, er, mr, 0, 0, 0, 2 * rf};
    sl.encoded.assign(er * sl.shape.h_cols, E8{});
    E8 inv_n = e8_inv(e8(uint32_t(cur.message_len % P)));
    std::vector<E8> alpha0(mr, inv_n);
    for (uint64_t c = 0; c < sl.shape.h_cols; c++) {
      const E8* src = (c / rf == 0) ? alpha0.data() : cur.alpha1.data() + (c % rf) * mr;
      host_encode(target, src, sl.encoded.data() + c * er);
    }
    Hasher pre;
    pre.frame("domain", "qafold-short/setup-base-row/v1");
    field_metadata(pre);
    pre.frame("decide-setup-mode", "direct-interleaved-v1");
    shape_frames(pre, sl.shape);
    std::vector<Digest> leaves(er);
    parallel_for(er, [&](uint64_t a, uint64_t b) {
      for (uint64_t r = a; r < b; r++) {
        Hasher h = pre;
        h.frame_u64("row", r);
        for (uint64_t c = 0; c < sl.shape.h_cols; c++) h.update(&sl.encoded[c * er + r], 16);
        leaves[r] = h.finish();
      }
    });
    sl.tree = merkle_levels(std::move(leaves));
  }
  {
    Hasher h;  // qafold_short_setup.rs:714-741
    h.frame("domain", "qafold-short/setup-aggregate/v1");
    field_metadata(h);
    h.frame_u64("levels", L);
    for (uint64_t l = 0; l < L; l++) {
      h.frame_u64("slot", l);
      uint8_t present = l ? 1 : 0;
      h.frame("present", &present, 1);
      if (l) { shape_frames(h, pp.setup[l].shape); h.frame("root", pp.setup[l].tree.back()[0].data(), 32); }
    }
    pp.setup_digest = h.finish();
  }
  {
    Hasher h;  // qafold_short_pc.rs:2271-2314
    h.frame("domain", "qafold-short-parameters/v4");
    h.frame("field-name", FIELD_NAME);
    h.frame_u64("modulus", P);
    h.frame_u64("extension-degree", 1);
    h.frame_u64("witness", pp.n_witness);
    h.frame_u64("extension-columns", pp.ext_cols);
    h.frame_u64("index", 2);
    h.frame_u64("security", pp.security_bits);
    h.frame_u64("rf-layers", pp.rf_layers);
    h.frame("generator-sequence", pp.sequence_digest.data(), 32);
    h.frame_u64("recursive-index-override", pp.recursive_index);
    h.frame("derived-rfc-sequence", pp.derived_digest.data(), 32);
    h.frame("setup-roots-and-shapes", pp.setup_digest.data(), 32);
    pp.parameter_digest = h.finish();
  }
  return pp;
}

LeafPrefix leaf_prefix(uint64_t width, const char* dom) {
  Bytes pre;
  auto frame = [&](const char* label, const void* b, size_t n) {
    size_t l = std::strlen(label);
    put_u64(pre, l); pre.insert(pre.end(), label, label + l);
    put_u64(pre, n); pre.insert(pre.end(), (const uint8_t*)b, (const uint8_t*)b + n);
  };
  frame("domain", dom, std::strlen(dom));
  uint8_t w[8]; for (int i = 0; i < 8; i++) w[i] = uint8_t(width >> (8 * i));
  frame("width", w, 8);
  require(pre.size() == 75 || pre.size() == 80, QF_ERR_ENCODING, "unexpected leaf prefix length");
  LeafPrefix lp{};
  lp.prefix_len = uint32_t(pre.size());
  b2s_init(lp.h);
  uint32_t m[16];
  for (int i = 0; i < 16; i++) m[i] = pre[4 * i] | uint32_t(pre[4 * i + 1]) << 8 | uint32_t(pre[4 * i + 2]) << 16 | uint32_t(pre[4 * i + 3]) << 24;
  b2s_compress(lp.h, m, 64, false);
  pre.resize(80, 0);
  for (int i = 0; i < 4; i++)
    lp.tail[i] = pre[64 + 4 * i] | uint32_t(pre[65 + 4 * i]) << 8 | uint32_t(pre[66 + 4 * i]) << 16 | uint32_t(pre[67 + 4 * i]) << 24;
  return lp;
}

InnerPrefix inner_prefix() {
  Bytes pre;
  const char* dom = "qafold-short/merkle-inner/v1";
  put_u64(pre, 6); pre.insert(pre.end(), "domain", "domain" + 6);
  put_u64(pre, std::strlen(dom)); pre.insert(pre.end(), dom, dom + std::strlen(dom));
  require(pre.size() == 50, QF_ERR_ENCODING, "unexpected inner prefix length");
  pre.resize(52, 0);
  InnerPrefix ip{};
  for (int i = 0; i < 13; i++) ip.w[i] = pre[4 * i] | uint32_t(pre[4 * i + 1]) << 8 | uint32_t(pre[4 * i + 2]) << 16 | uint32_t(pre[4 * i + 3]) << 24;
  return ip;
}

}  // namespace qf
