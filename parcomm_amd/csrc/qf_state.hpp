// Resident state behind the C ABI handles: public parameters (host setup plus
// the device copies of the QA generators) and one commitment (per-block
// codewords and Merkle trees kept in HBM, or roots only).
#pragma once
#include "qf_kernels.hpp"
#include <algorithm>
#include <atomic>
#include <mutex>

namespace qf {

struct ParamsState {
  Layout layout;
  uint32_t cfg_chunk_log = 30, packed_log = 4, query_bits = 100;
  int chunk_log = 0;
  uint64_t block_len = 0, blocks = 0;
  PcsParams pp;
  CommitGeom geom{};
  LeafPrefix leaf_ext{};
  mutable E8* d_alpha = nullptr;              // top-code alpha_1, uploaded on first use
  mutable std::vector<E8*> d_alpha_levels;    // alpha_1 of every level (top included), on first prove
  mutable std::vector<E8*> d_alpha_enc;       // the same in encode form (KoalaBear: Montgomery)
  mutable std::vector<E8*> d_setup_enc;       // setup level l's encoded matrix (column-major), on first Decide
  mutable std::vector<uint64_t> zero_ranges;  // padding intervals (start, len) chunked for the device check
  mutable std::mutex mu;                      // guards the lazily filled members above
  mutable std::atomic<bool> commit_warm{false};  // a commit has run (and tuned its tile passes)
  ~ParamsState() {
    for (auto* a : d_setup_enc) if (a) (void)hipFree(a);
    for (auto* a : d_alpha_enc)
      if (a && std::find(d_alpha_levels.begin(), d_alpha_levels.end(), a) == d_alpha_levels.end()) (void)hipFree(a);
    for (auto* a : d_alpha_levels) if (a && a != d_alpha) (void)hipFree(a);
    if (d_alpha) (void)hipFree(d_alpha);
  }
  // alpha_1 of code `level`, canonical (for the opening algebra)
  const E8* alpha_device(size_t level) const;
  // alpha_1 of code `level` in the encoding kernels' form (v8_encode_form)
  const E8* alpha_encode(size_t level) const;
  // setup level `level`'s encoded matrix in device memory (same layout as pp.setup[level].encoded)
  const E8* setup_device(size_t level) const;
};

struct CommittedState {
  const ParamsState* p = nullptr;
  int cache_mode = 0;
  std::vector<Digest> roots;
  std::vector<E8*> d_cw;        // per block (cached mode; blocks with equal roots share one)
  std::vector<uint32_t*> d_tree;
  const Val* d_bva = nullptr;  // the committed BVA (active entries) in HBM
  bool owns_bva = false;
  double total_ms = 0, leaf_ms = 0, encode_ms = 0, mw_ms = 0;
  bool graphed = false;  // the block loop ran as a captured HIP graph
  uint64_t timed_blocks = 0;  // blocks actually encoded (zero blocks past the data reuse a root)
  ~CommittedState() {
    std::vector<void*> freed;
    auto release = [&](void* q) {  // codewords and trees come from the stream-ordered pool
      if (q && std::find(freed.begin(), freed.end(), q) == freed.end()) { freed.push_back(q); dfree(q, thread_stream()); }
    };
    for (auto* q : d_cw) release(q);
    for (auto* q : d_tree) release(q);
    if (owns_bva && d_bva) (void)hipFree(const_cast<Val*>(d_bva));
  }
};

struct ProveTimings {
  double conversion_ms = 0, eo_ms = 0, projection_ms = 0, folding_ms = 0, sumcheck_ms = 0, binding_ms = 0,
         native_ms = 0, decide_ms = 0, total_ms = 0;
};

// Parameters::prove (protocol.rs:201-263) on the device-resident commitment.
// `claims` holds n_matrices claims in the reference's order (row point,
// column point, value). Returns the bincode proof bytes.
struct ClaimIn { std::vector<E8> row, col; E8 value; };
Bytes prove(const ParamsState& p, const CommittedState& c, const std::vector<ClaimIn>& claims, ProveTimings& t);
// The benchmark statement of the reference CLI (main.rs:188-209): points from
// the commitment transcript, values evaluated on the device BVA.
std::vector<ClaimIn> benchmark_claims(const ParamsState& p, const CommittedState& c);
// tbv.rs:143-185: fills claims[i].value with matrix i's MLE at its (row, col) point.
void evaluate_claims(const ParamsState& p, const CommittedState& c, std::vector<ClaimIn>& claims);
std::string statement_json(const ParamsState& p, const CommittedState& c);
// Parameters::verify (protocol.rs:266-319): throws QfError on rejection.
void verify(const ParamsState& p, const std::vector<Digest>& roots, const std::vector<ClaimIn>& claims,
            const uint8_t* proof, size_t len);
// Whether verify() spreads its row hashing over the host's threads (the
// default) on the calling thread.
void set_verify_rows_parallel(bool on);
std::string claims_json(const std::vector<ClaimIn>& claims);

// Padding intervals of a layout as (start, len) pairs of at most 64 Ki entries
// (layout.rs:136-155), and the device check that ORs a nonzero entry in any of
// `count` such intervals (device table d_ranges) into *d_flag.
std::vector<uint64_t> padding_ranges(const Layout& l);
void launch_check_zero(const Val* bva, const uint64_t* d_ranges, size_t count, uint32_t* d_flag, hipStream_t st);

}  // namespace qf
