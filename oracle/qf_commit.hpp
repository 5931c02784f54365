// CPU oracle (test infrastructure only; see qf_core.hpp): layout, QA codes,
// setup, parameter digest, packing, Mobius transform, encoding and Merkle trees
// of the F65521 / Ext8 QAFold commitment.
#pragma once
#include "qf_core.hpp"

namespace orc {

// ---------------------------------------------------------------- layout (layout.rs:4-155)
struct Shape { u64 rows, cols; };
struct Bucket { u64 matrix, row_offset, col_offset, rows, cols, physical_rows, physical_cols, offset; u64 len() const { return rows * cols; } };
struct Layout {
  std::vector<Shape> shapes;
  std::vector<Bucket> buckets;
  u64 active = 0, logical = 0;
};
std::vector<std::pair<u64, u64>> intervals(u64 length, size_t budget);
Layout layout_new(const std::vector<Shape>& shapes);
std::vector<Val> layout_encode(const Layout& l, const std::vector<std::vector<Val>>& m);
void layout_validate(const Layout& l, const Val* values, u64 n);

// ---------------------------------------------------------------- QA codes over Ext8
void wht(E8* v, size_t n);          // unnormalised Walsh-Hadamard, in place
void mobius(E8* v, size_t n);       // Boolean -> monomial: high -= low at every level

struct QaCode {
  size_t message_len = 0, index = 0;
  std::vector<std::vector<E8>> alphas;  // alphas[0] = 1/n (systematic), alphas[1..]
  Digest generator_digest{};
  size_t codeword_len() const { return message_len * index; }
  void encode(const E8* message, E8* out) const;  // out has codeword_len entries
};
QaCode qa_sample_systematic(size_t message_len, size_t index, const Digest& seed);

struct QaSequence {
  size_t witness = 0, ext_cols = 0, rf_layers = 0, recursive_index = 2;
  std::vector<QaCode> codes;  // ascending message length; codes.back() is the top code
  Digest generator_digest{};
};
QaSequence qa_sequence_new(size_t witness, size_t ext_cols, size_t recursive_index, const Digest& seed);

// ---------------------------------------------------------------- Merkle (qafold_short_pc.rs:975-1021)
struct Merkle {
  std::vector<std::vector<Digest>> levels;  // levels[0] = leaves
  Digest root() const { return levels.back()[0]; }
  std::vector<Digest> path(size_t index) const;
};
Merkle merkle_from_leaves(std::vector<Digest> leaves);
Digest hash_inner(const Digest& l, const Digest& r);
// Leaves over column-major entries: row r = entries[k*rows + r], k < cols.
std::vector<Digest> hash_rows(const E8* entries, size_t rows, size_t cols, const char* domain);
extern const char* LEAF_BASE_DOMAIN;
extern const char* LEAF_EXTENSION_DOMAIN;

// ---------------------------------------------------------------- setup + PCS parameters
struct SetupShape { u64 level, encoded_rows, message_rows, t_cols, t_elements, t_last_len, h_cols; };
struct SetupLevel {
  SetupShape shape;
  std::vector<E8> encoded;  // column-major, encoded_rows x h_cols
  Merkle tree;
};
struct PcsParams {
  size_t n_witness = 0, ext_cols = 0, index = 2, recursive_index = 2, security_bits = 0;
  QaSequence seq;
  Digest derived_digest{}, setup_digest{}, parameter_digest{};
  std::vector<SetupLevel> setup;  // setup[level], level 0 unused
};
PcsParams pcs_setup(size_t n_witness, size_t ext_cols, size_t recursive_index, size_t security_bits, const u8 rng_seed[32]);
// Setup drawing the sequence seed from a caller-owned StdRng (ShortQaFoldPC::setup).
PcsParams pcs_setup_rng(size_t n_witness, size_t ext_cols, size_t recursive_index, size_t security_bits, ChaChaRng& rng);
Digest setup_row_hash(const SetupShape& s, size_t row, const E8* row_vals);

// ---------------------------------------------------------------- commit of one packed block
struct Committed {
  std::vector<E8> coeffs;   // Mobius-transformed packed witness (column-major message)
  std::vector<E8> encoded;  // column-major codeword, rows = codeword_len, cols = base columns
  size_t rows = 0, cols = 0;
  Merkle tree;
  Digest root() const { return tree.root(); }
};
// `witness` is the column-major message (n_witness entries). Encodes and hashes.
Committed pcs_commit_owned(const PcsParams& pp, std::vector<E8> witness);
// PackedPcs: 8 consecutive base values per Ext8 element, then Mobius (packed.rs:220-258).
std::vector<E8> pack_block(const Val* values, u64 available, u64 block_len);
Committed packed_commit(const PcsParams& pp, const std::vector<E8>& packed);

}  // namespace orc
