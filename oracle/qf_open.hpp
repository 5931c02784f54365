// CPU oracle (test infrastructure only; see qf_core.hpp): the opening path.
// Transcripts, JSON/bincode statement encodings, TBV compilation, packed EO,
// batched and single-block packed PCS openings, codeword-fold binding, the
// native QAFold opening and the Decide wire.
#pragma once
#include "qf_commit.hpp"
#include <functional>

namespace orc {

// ---------------------------------------------------------------- statement types
struct Config { u64 chunk_log = 30, packed_log = 4, query_rf_bits = 100; };
struct Claim { std::vector<E8> row, col; E8 value; };
struct Statement {  // protocol.rs:33-40 (Commitment)
  u32 version = 2;
  Config config;
  std::vector<Shape> shapes;
  Digest parameter_digest{};
  std::vector<Digest> roots;
};
std::string json_statement(const Statement& s);
std::string json_claims(const std::vector<Claim>& c);
std::string json_scalars(const std::vector<E8>& v);

// transcript.rs:8-70
struct OuterTranscript {
  Digest state;
  u64 counter = 0;
  OuterTranscript();
  void bytes(const char* label, const void* data, size_t n);
  void object(const char* label, const std::string& json) { bytes(label, json.data(), json.size()); }
  void scalar(const char* label, const E8& v) { u8 b[16]; e8_write(v, b); bytes(label, b, 16); }
  E8 challenge(const char* label);
  E8 nonzero(const char* label);
  E8 fold_challenge(u64 packed_log);
};

// ---------------------------------------------------------------- TBV (tbv.rs)
struct Term {
  u64 offset = 0;
  E8 scale;
  std::vector<std::array<E8, 2>> factors;
  u64 len() const { return u64(1) << factors.size(); }
  E8 weight(u64 index) const;
  E8 evaluation(const std::vector<E8>& point_low) const;
};
struct Query {
  std::vector<Term> terms;
  E8 claim;
  E8 evaluation(const std::vector<E8>& point_low) const;
};
E8 monomial_weight(u64 index, const E8* point_low, size_t n);
E8 equality_weight(u64 index, const std::vector<E8>& point_msb);
std::vector<E8> tbv_weights(const std::array<E8, 2>* factors, size_t n);
std::vector<E8> tbv_project(const Term& t, const Val* values, u64 n_values, u64 stride);
void cross_terms(const Query& q, const Val* values, u64 n_values, u64 stride, std::vector<E8>& pos, std::vector<E8>& neg);
std::vector<E8> query_fold(const Query& q, u64 logical, u64 stride, const E8& challenge);
Query tbv_compile(const Layout& l, const std::vector<Claim>& claims, OuterTranscript& t);
std::vector<Claim> evaluate_claims(const Layout& l, const Val* values, u64 n_values,
                                   const std::vector<std::pair<std::vector<E8>, std::vector<E8>>>& points);

// ---------------------------------------------------------------- EO (eo.rs)
struct EoProof { std::vector<E8> positive, negative; std::vector<std::array<E8, 2>> rounds; E8 value, query; };
struct Terminal { E8 value; std::vector<E8> point_msb; E8 scale; };
EoProof eo_prove(const Val* values, u64 n_values, u64 logical, u64 stride, const Query& q, OuterTranscript& t,
                 Terminal& terminal);

// ---------------------------------------------------------------- native PCS opening
struct OpeningReduction {  // ShortQaFoldVerificationReduction, pending claims of non-base levels
  struct Pending { u64 level; E8 value; std::vector<E8> left, right; };
  std::vector<Pending> pending;
};
// Returns (claimed value, opening wire) for a committed codeword at `point`.
E8 pcs_open_with_reduction(const PcsParams& pp, const Digest& root, const std::vector<E8>& point, const Committed& c,
                           Bytes& wire, OpeningReduction& red);
Bytes pcs_prove_decide(const PcsParams& pp, const OpeningReduction& red);
u64 num_queries(const QaCode& code, u64 security_bits, u64 available_rows);
std::vector<u64> sample_without_replacement(u64 count, u64 population, const Digest& seed);
void write_base_rows_all(Bytes& out, const Committed& c);
void write_base_rows_partial(Bytes& out, const Committed& c, const std::vector<u64>& rows);

// ---------------------------------------------------------------- packed PCS (packed.rs, packed_batch.rs)
using BlockSource = std::function<Val(u64 block, u64 index)>;
// Batched opening over 2..128 blocks (recompute mode: blocks are re-committed from source).
Bytes packed_open_batch(const PcsParams& pp, int log_n, const std::vector<Digest>& roots, const std::vector<E8>& point,
                        const Bytes& context, const BlockSource& src, E8& claimed_value);
Bytes packed_open_single(const PcsParams& pp, int log_n, const Committed& block, const std::vector<E8>& point,
                         const Bytes& context, E8& claimed_value);

// ---------------------------------------------------------------- ParComm prover (protocol.rs)
struct ProveResult { Bytes proof; E8 opening_value; };
Bytes parcomm_prove(const Layout& l, const Config& cfg, const PcsParams& pp, int chunk_log, const Statement& st,
                    const Val* values, u64 n_values, const std::vector<Claim>& claims);

}  // namespace orc
