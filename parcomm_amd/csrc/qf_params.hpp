// Public parameters of the QAFold prover: BVA layout, QA code
// sequence, setup commitments and the parameter digest. Computed once per
// context on the host (setup time is outside the prover's timed phases, as in
// the reference's Parameters::new, protocol.rs:109-130).
#pragma once
#include "qf_host.hpp"
#include <functional>

namespace qf {

struct Shape { uint64_t rows, cols; };
struct Bucket {
  uint64_t matrix, row_offset, col_offset, rows, cols, physical_rows, physical_cols, offset;
  uint64_t len() const { return rows * cols; }
};
struct Layout {
  std::vector<Shape> shapes;
  std::vector<Bucket> buckets;
  uint64_t active = 0, logical = 0;
};
Layout make_layout(const std::vector<Shape>& shapes);  // layout.rs:56-114

struct QaCode {
  uint64_t message_len = 0, index = 2;
  std::vector<E8> alpha1;  // alphas[1]; alphas[0] is the constant 1/n of a systematic code
  Digest generator_digest{};
  uint64_t codeword_len() const { return message_len * index; }
};

struct SetupShape { uint64_t level, encoded_rows, message_rows, t_cols, t_elements, t_last_len, h_cols; };
struct SetupLevel {
  SetupShape shape{};
  std::vector<E8> encoded;                 // column-major encoded_rows x h_cols
  std::vector<std::vector<Digest>> tree;   // levels[0] = leaves
};

struct PcsParams {
  uint64_t n_witness = 0, ext_cols = 32, security_bits = 100, rf_layers = 5, recursive_index = 2;
  std::vector<QaCode> codes;               // ascending; codes.back() = top code
  Digest sequence_digest{}, derived_digest{}, setup_digest{}, parameter_digest{};
  std::vector<SetupLevel> setup;           // index = level, [0] unused
};

// PackedPcs::new(log_n, 32, bits, seed) (packed.rs:134-158)
PcsParams make_pcs_params(int log_n, uint64_t security_bits, const uint8_t seed[32]);

// Host reference transforms used by setup and small recursive levels.
void host_wht(E8* v, uint64_t n);
void host_encode(const QaCode& code, const E8* message, E8* out);
Digest hash_inner(const Digest& l, const Digest& r);
std::vector<std::vector<Digest>> merkle_levels(std::vector<Digest> leaves);
void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)>& body);

// Prefix state of the base-row leaf hash: framed domain + width (qafold_short_pc.rs:2437-2475).
struct LeafPrefix {
  uint32_t h[8];        // chaining value after the first 64 prefix bytes
  uint32_t tail[4];     // prefix bytes 64.. as little-endian words (last one partial)
  uint32_t prefix_len;  // 75 (base rows) or 80 (extension rows)
};
// domain: "qafold-short/base-row/v1" or "qafold-short/extension-row/v1"
LeafPrefix leaf_prefix(uint64_t width, const char* domain = "qafold-short/base-row/v1");
struct InnerPrefix { uint32_t w[13]; };  // bytes 0..49 of the inner-node message, word 12 partial
InnerPrefix inner_prefix();

}  // namespace qf
