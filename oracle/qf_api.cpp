// C entry points of the CPU oracle, loaded by tests/ and bench.py through
// ctypes (test infrastructure only; see qf_core.hpp).
#include "qf_commit.hpp"
#include <chrono>
#include <memory>

using namespace orc;

namespace {
thread_local std::string g_err;
// Wall seconds of the phases of the last orc_prove_* call on this thread:
// setup, commit, claims, prove (Parameters::prove: EO, opening, Decide).
thread_local double g_phase_s[4];
double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
template <class F>
int guarded(F&& f) {
  try { f(); return 0; } catch (const std::exception& e) { g_err = e.what(); return 1; }
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
// Phase times (seconds) of this thread's last orc_prove_* call: setup, commit, claims, prove.
void orc_last_phase_seconds(double* out4) { for (int i = 0; i < 4; i++) out4[i] = g_phase_s[i]; }
// Extension degree of this build: 8 (F65521 / Ext8) or 4 (KoalaBear / Ext4).
int orc_lanes() { return LANES; }

void orc_blake2s(const uint8_t* data, size_t len, uint8_t* out) {
  Digest d = blake2s(data, len);
  std::memcpy(out, d.data(), 32);
}

void orc_chacha_stream(const uint8_t* key, int rounds, uint8_t* out, size_t n) {
  ChaChaRng r(key, rounds);
  r.fill_bytes(out, n);
}

void orc_chacha_block(const uint8_t* key, uint64_t counter, uint64_t stream, int rounds, uint32_t* out16) {
  chacha_block(key, counter, stream, rounds, out16);
}

// Extension arithmetic for property tests on 16-byte canonical elements:
// op 0 add, 1 sub, 2 mul, 3 inv(a), 4 frob(a, k=b.c[0]).
int orc_e8_op(int op, const void* a, const void* b, void* out) {
  return guarded([&] {
    E8 x, y;
    std::memcpy(x.c, a, 16);
    std::memcpy(y.c, b, 16);
    E8 r;
    switch (op) {
      case 0: r = x + y; break;
      case 1: r = x - y; break;
      case 2: r = x * y; break;
      case 3: r = e8_inv(x); break;
      case 4: r = e8_frob(x, y.c[0]); break;
      default: throw Error("bad op");
    }
    std::memcpy(out, r.c, 16);
  });
}

void orc_wht(void* v, size_t n) { wht(reinterpret_cast<E8*>(v), n); }
void orc_mobius(void* v, size_t n) { mobius(reinterpret_cast<E8*>(v), n); }

// Layout: shapes[2*i] = rows, shapes[2*i+1] = cols. Writes up to `cap` buckets of
// 8 u64 (matrix,row_offset,col_offset,rows,cols,physical_rows,physical_cols,offset).
int orc_layout(const uint64_t* shapes, size_t n, uint64_t* buckets, size_t cap, size_t* n_buckets,
               uint64_t* active, uint64_t* logical) {
  return guarded([&] {
    std::vector<Shape> s(n);
    for (size_t i = 0; i < n; i++) s[i] = {shapes[2 * i], shapes[2 * i + 1]};
    Layout l = layout_new(s);
    *n_buckets = l.buckets.size();
    *active = l.active;
    *logical = l.logical;
    for (size_t i = 0; i < l.buckets.size() && i < cap; i++) {
      const Bucket& b = l.buckets[i];
      uint64_t v[8] = {b.matrix, b.row_offset, b.col_offset, b.rows, b.cols, b.physical_rows, b.physical_cols, b.offset};
      std::memcpy(buckets + 8 * i, v, sizeof v);
    }
  });
}

// PCS parameters of PackedPcs::new(chunk_log, 32, bits, seed) (packed.rs:134-158).
void* orc_pcs_new(int chunk_log, int security_bits, const uint8_t* seed32) {
  try {
    return new PcsParams(pcs_setup(size_t(1) << (chunk_log - LOG_LANES), 32, 2, size_t(security_bits), seed32));
  } catch (const std::exception& e) { g_err = e.what(); return nullptr; }
}
void orc_pcs_free(void* h) { delete static_cast<PcsParams*>(h); }
void orc_pcs_parameter_digest(void* h, uint8_t* out) { std::memcpy(out, static_cast<PcsParams*>(h)->parameter_digest.data(), 32); }
size_t orc_pcs_levels(void* h, size_t* lens) {
  auto* pp = static_cast<PcsParams*>(h);
  for (size_t i = 0; i < pp->seq.codes.size(); i++) lens[i] = pp->seq.codes[i].message_len;
  return pp->seq.codes.size();
}
// alpha_1 of the top code (length message_len), 16 canonical bytes per element.
void orc_pcs_top_alpha(void* h, void* out) {
  auto* pp = static_cast<PcsParams*>(h);
  const auto& a = pp->seq.codes.back().alphas[1];
  std::memcpy(out, a.data(), a.size() * 16);
}

// Commit one block of 2^chunk_log signed entries (i8 for F65521, i16 for
// KoalaBear; entries past `available` are 0).
// Optionally returns the Mobius-transformed witness and the codeword.
int orc_commit_block(void* h, const void* values, uint64_t available, uint8_t* root, void* coeffs_out,
                     void* codeword_out) {
  return guarded([&] {
    auto* pp = static_cast<PcsParams*>(h);
    u64 block_len = u64(pp->n_witness) * LANES;
    Committed c = packed_commit(*pp, pack_block(static_cast<const Val*>(values), available, block_len));
    Digest r = c.root();
    std::memcpy(root, r.data(), 32);
    if (coeffs_out) std::memcpy(coeffs_out, c.coeffs.data(), c.coeffs.size() * 16);
    if (codeword_out) std::memcpy(codeword_out, c.encoded.data(), c.encoded.size() * 16);
  });
}

// Merkle root over column-major extension entries with the base-row leaf domain.
void orc_merkle_root_rows(const void* entries, size_t rows, size_t cols, uint8_t* root) {
  Merkle m = merkle_from_leaves(hash_rows(reinterpret_cast<const E8*>(entries), rows, cols, LEAF_BASE_DOMAIN));
  Digest r = m.root();
  std::memcpy(root, r.data(), 32);
}

}  // extern "C"

// ---------------------------------------------------------------- full prover (main.rs:175-220 flow)
#include "qf_open.hpp"
// Next coordinate of a packed point list (16 canonical bytes); advances the cursor.
static E8 read_point_coord(const uint8_t*& q) {
  E8 v;
  check(e8_read(q, v), "non-canonical point coordinate");
  q += 16;
  return v;
}

extern "C" {
// Commits the BVA, derives the benchmark points from the statement
// (main.rs:190-204), evaluates the claims and proves. Writes the bincode proof
// into `proof` (capacity `cap`), its length into `proof_len`, and the compact
// JSON of the commitment and claims into the optional buffers.
// With `points` (per matrix its row point and then its column point, 16
// canonical bytes per coordinate) the claims are evaluated at those points
// instead of the benchmark points (tbv.rs:143-185), as the reference's own
// tests do with hand-picked coordinates (end_to_end.rs:170-192).
static int prove_flow(const uint64_t* shapes, size_t n, uint32_t chunk_log_cfg, uint32_t packed_log,
                      const void* values_in, uint64_t n_values, const uint8_t* points_in, uint8_t* proof, size_t cap,
                      size_t* proof_len, char* statement_json, size_t statement_cap, char* claims_json,
                      size_t claims_cap) {
  return guarded([&] {
    const Val* values = static_cast<const Val*>(values_in);
    std::vector<Shape> s(n);
    for (size_t i = 0; i < n; i++) s[i] = {shapes[2 * i], shapes[2 * i + 1]};
    Layout l = layout_new(s);
    layout_validate(l, values, n_values);
    int chunk_log = std::min<int>(int(chunk_log_cfg), ilog2(l.logical));
    u64 block_len = u64(1) << chunk_log;
    u64 blocks = l.logical / block_len;
    check(blocks <= 128, "at most 128 PCS blocks are supported");
    u8 seed[32];
    std::memset(seed, 0x51, 32);
    auto t0 = std::chrono::steady_clock::now();
    PcsParams pp = pcs_setup(size_t(1) << (chunk_log - LOG_LANES), 32, 2, 100, seed);
    g_phase_s[0] = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    Statement st;
    st.config = {chunk_log_cfg, packed_log, 100};
    st.shapes = s;
    st.parameter_digest = pp.parameter_digest;
    for (u64 b = 0; b < blocks; b++) {
      u64 first = b * block_len;
      u64 avail = first < n_values ? std::min<u64>(block_len, n_values - first) : 0;
      Committed c = packed_commit(pp, pack_block(values + (avail ? first : 0), avail, block_len));
      st.roots.push_back(c.root());
    }
    g_phase_s[1] = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    OuterTranscript pts;
    pts.object("benchmark-points", json_statement(st));
    std::vector<std::pair<std::vector<E8>, std::vector<E8>>> points;
    for (auto& sh : s) {
      u64 rp = 1, cp = 1;
      while (rp < sh.rows) rp <<= 1;
      while (cp < sh.cols) cp <<= 1;
      std::vector<E8> row, col;
      for (int i = 0; i < ilog2(rp); i++) row.push_back(points_in ? read_point_coord(points_in) : pts.challenge("row"));
      for (int i = 0; i < ilog2(cp); i++) col.push_back(points_in ? read_point_coord(points_in) : pts.challenge("col"));
      points.push_back({row, col});
    }
    auto claims = evaluate_claims(l, values, n_values, points);
    Config cfg{chunk_log_cfg, packed_log, 100};
    g_phase_s[2] = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    Bytes pf = parcomm_prove(l, cfg, pp, chunk_log, st, values, n_values, claims);
    g_phase_s[3] = seconds_since(t0);
    *proof_len = pf.size();
    check(pf.size() <= cap, "proof buffer too small");
    std::memcpy(proof, pf.data(), pf.size());
    auto put = [](const std::string& js, char* out, size_t c) {
      if (!out) return;
      check(js.size() + 1 <= c, "json buffer too small");
      std::memcpy(out, js.c_str(), js.size() + 1);
    };
    put(json_statement(st), statement_json, statement_cap);
    put(json_claims(claims), claims_json, claims_cap);
  });
}

int orc_prove_benchmark(const uint64_t* shapes, size_t n, uint32_t chunk_log_cfg, uint32_t packed_log,
                        const void* values_in, uint64_t n_values, uint8_t* proof, size_t cap, size_t* proof_len,
                        char* statement_json, size_t statement_cap, char* claims_json, size_t claims_cap) {
  return prove_flow(shapes, n, chunk_log_cfg, packed_log, values_in, n_values, nullptr, proof, cap, proof_len,
                    statement_json, statement_cap, claims_json, claims_cap);
}

int orc_prove_points(const uint64_t* shapes, size_t n, uint32_t chunk_log_cfg, uint32_t packed_log,
                     const void* values_in, uint64_t n_values, const uint8_t* points, uint8_t* proof, size_t cap,
                     size_t* proof_len, char* statement_json, size_t statement_cap, char* claims_json,
                     size_t claims_cap) {
  return prove_flow(shapes, n, chunk_log_cfg, packed_log, values_in, n_values, points, proof, cap, proof_len,
                    statement_json, statement_cap, claims_json, claims_cap);
}
}

extern "C" {
// qafold_short_pc.rs:3605-3647 (native_opening_and_decide_wire_fixture):
// ShortQaFoldPC<F>::new(n_witness, ext_cols, 2, bits) with StdRng::seed_from_u64(rng_seed);
// writes BLAKE2s(root || opening || decide) and the opening/decide lengths.
int orc_native_fixture(uint64_t n_witness, uint64_t ext_cols, uint64_t bits, uint64_t rng_seed, uint8_t* digest,
                       uint64_t* opening_len, uint64_t* decide_len) {
  return guarded([&] {
    u8 seed[32];
    seed_from_u64(rng_seed, seed);
    ChaChaRng rng(seed, 12);
    PcsParams pp = pcs_setup_rng(n_witness, ext_cols, 2, bits, rng);
    auto sample = [&]() {
      for (;;) {
        u8 b[16];
        u64 lo = rng.next_u64(), hi = rng.next_u64();
        for (int i = 0; i < 8; i++) { b[i] = u8(lo >> (8 * i)); b[8 + i] = u8(hi >> (8 * i)); }
        E8 v;
        if (e8_from_uniform(b, v)) return v;
      }
    };
    std::vector<E8> witness(n_witness);
    for (auto& w : witness) w = sample();
    std::vector<E8> point(ilog2(n_witness));
    for (auto& p : point) p = sample();
    Committed c = pcs_commit_owned(pp, witness);
    Bytes opening;
    OpeningReduction red;
    pcs_open_with_reduction(pp, c.root(), point, c, opening, red);
    Bytes decide = pcs_prove_decide(pp, red);
    Blake2s h;
    Digest r = c.root();
    h.update(r.data(), 32);
    h.update(opening);
    h.update(decide);
    Digest d = h.finalize();
    std::memcpy(digest, d.data(), 32);
    *opening_len = opening.size();
    *decide_len = decide.size();
  });
}
}
