// Device vectors of extension elements and the vector operations of the opening
// (LiteBullet, even/odd and sumcheck rounds, tensor tables, codeword column
// projections). Reductions return their result to the host: each workgroup
// writes a partial sum and the host adds at most a few thousand partials.
#pragma once
#include <chrono>
#include "qf_kernels.hpp"
#include <memory>
#include <utility>

namespace qf {

// Owning device buffer of `n` extension elements.
// Allocated and released in order on the thread's current stream.
struct DVec {
  E8* p = nullptr;
  uint64_t n = 0;
  hipStream_t st = nullptr;
  DVec() = default;
  explicit DVec(uint64_t len);
  DVec(const DVec&) = delete;
  DVec& operator=(const DVec&) = delete;
  DVec(DVec&& o) noexcept : p(o.p), n(o.n), st(o.st) { o.p = nullptr; o.n = 0; }
  DVec& operator=(DVec&& o) noexcept;
  ~DVec();
  void resize_view(uint64_t len) { n = len; }  // shrink the logical length (storage kept)
};

// Host waits of this thread (QF_TRACE_SYNC=1: counted and timed in DevCtx::sync).
struct SyncTrace {
  uint64_t syncs = 0;
  double wait_ms = 0;
};
bool sync_trace_on();
SyncTrace& sync_trace();
// Wall time of named host intervals, accumulated per thread while
// QF_TRACE_SYNC=1: mark(name) books the time since the previous mark (or the
// construction) under name.
struct TraceLaps {
  std::chrono::steady_clock::time_point t;
  TraceLaps();
  void mark(const char* name);
};
void trace_scopes_flush();  // prints and clears this thread's interval totals

// Stream and scratch for reduction partials, owned by the prover of one
// proof; while it lives it is the thread's current stream. Its destructor
// drains the stream.
struct DevCtx {
  hipStream_t st = nullptr;
  hipStream_t prev = nullptr;
  uint32_t* partials = nullptr;       // device address of host_partials: kernels write block sums here
  uint32_t* host_partials = nullptr;  // pinned, coherent; the thread's buffer, read after sync()
  static constexpr uint32_t MAX_GROUPS = 1024;
  DevCtx();
  ~DevCtx();
  void sync();
};

namespace dv {
DVec upload(DevCtx& c, const std::vector<E8>& v);
// Device-to-host copy of `bytes` through the thread's pinned buffer (synchronous).
void d2h(DevCtx& c, void* dst, const void* src, size_t bytes);
std::vector<E8> download(DevCtx& c, const E8* p, uint64_t n);
// Two device vectors of n elements each (b may be null) with one synchronisation.
std::pair<std::vector<E8>, std::vector<E8>> download2(DevCtx& c, const E8* a, const E8* b, uint64_t n);
inline std::vector<E8> download(DevCtx& c, const DVec& v) { return download(c, v.p, v.n); }
void fill(DevCtx& c, E8* p, uint64_t n, const E8& value);
void copy(DevCtx& c, E8* dst, const E8* src, uint64_t n);
// out[i] = hi[i >> a] * lo[i & (2^a - 1)]  (tensor product of two host tables)
void tensor(DevCtx& c, E8* out, const std::vector<E8>& hi, const std::vector<E8>& lo);
E8 inner(DevCtx& c, const E8* a, const E8* b, uint64_t n);
// (sum_i a[i] b[h+i], sum_i a[h+i] b[i]), i < h
std::pair<E8, E8> cross_halves(DevCtx& c, const E8* a, const E8* b, uint64_t h);
// a[i] += a[h+i] x; b[i] += b[h+i] y
void lb_fold(DevCtx& c, E8* a, E8* b, uint64_t h, const E8& x, const E8& y);
// lb_fold, returning the next round's cross_halves over the folded length h (h even)
std::pair<E8, E8> lb_fold_cross(DevCtx& c, E8* a, E8* b, uint64_t h, const E8& x, const E8& y);
// ow[i] = w[2i] + w[2i+1] ri; oq[i] = q[2i] + q[2i+1] r
void eo_fold(DevCtx& c, const E8* w, const E8* q, E8* ow, E8* oq, uint64_t h, const E8& ri, const E8& r);
// eo_fold into (ow, oq) of length h, returning the next round's cross terms over the outputs (h even)
std::pair<E8, E8> eo_fold_cross(DevCtx& c, const E8* w, const E8* q, E8* ow, E8* oq, uint64_t h, const E8& ri,
                                const E8& r);
// (sum v_lo w_lo, sum (v_hi - v_lo)(w_hi - w_lo))
std::pair<E8, E8> sc_msg(DevCtx& c, const E8* v, const E8* w, uint64_t h);
// v_lo += r (v_hi - v_lo), same for w
void sc_fold(DevCtx& c, E8* v, E8* w, uint64_t h, const E8& r);
// sc_fold, returning the next round's sc_msg over the folded length h (h even)
std::pair<E8, E8> sc_fold_msg(DevCtx& c, E8* v, E8* w, uint64_t h, const E8& r);
void axpy(DevCtx& c, E8* y, const E8* x, uint64_t n, const E8& s);  // y += s x
void mul(DevCtx& c, E8* y, const E8* x, uint64_t n);                // y *= x
void scale(DevCtx& c, E8* y, uint64_t n, const E8& s);              // y *= s
// out[col] = sum_{r < n} cw[col*R + r] * w[r]
std::vector<E8> col_projection(DevCtx& c, const E8* cw, uint64_t R, uint64_t n, uint64_t cols, const E8* w);
// out[r] = sum_col f[col] * cw[col*R + r], r < n
void col_fold(DevCtx& c, const E8* cw, uint64_t R, uint64_t n, uint64_t cols, const std::vector<E8>& f, E8* out);
// inner[col n/32 + h] = sum_{l < 32} low[l] cw[col R + 32 h + l] (F65521 matrix cores; false
// where it does not apply, leaving inner untouched).
bool col_fold_rows32(DevCtx& c, const E8* cw, uint64_t R, uint64_t n, uint64_t cols, const std::vector<E8>& low,
                     E8* inner);
// Gather full rows (cols entries each) of a column-major matrix: out[k*cols + col].
std::vector<E8> gather_rows(DevCtx& c, const E8* cw, uint64_t R, uint64_t cols, const std::vector<uint64_t>& rows);
// out[k] = sum_{col < ca} w[col] a[col*Ra + rows[k]] + sum_{col < cb} w[ca + col] b[col*Rb + rows[k]]
// (column-major a and b; b may be null when cb = 0).
std::vector<E8> gather_weighted(DevCtx& c, const E8* a, uint64_t Ra, uint64_t ca, const E8* b, uint64_t Rb, uint64_t cb,
                                const std::vector<E8>& w, const std::vector<uint64_t>& rows);
// Appends one record per query row to `out`, built on the device and copied
// back once: u64 cols, the row's cols elements (column-major codeword with R
// rows), u64 depth = log2 R, then its depth Merkle siblings leaf to root.
void append_row_records(DevCtx& c, Bytes& out, const E8* cw, uint64_t R, uint64_t cols, const uint32_t* tree,
                        const std::vector<uint64_t>& rows);
void scatter_add(DevCtx& c, E8* dst, const std::vector<uint64_t>& idx, const std::vector<E8>& vals);
}  // namespace dv

// Launch form of the per-term opening kernels: one batched launch over all
// bucket terms (default), or the plain per-matrix form with one launch and one
// host round trip per term (kept for comparison). Process-wide.
enum : int { LAUNCH_BATCHED = 0, LAUNCH_PER_MATRIX = 1 };
void set_launch_mode(int mode);
int launch_mode();

// ---------------------------------------------------------------- BVA-scale kernels of the opening
// One projection task: lanes[l] = sum_row W(row) * v[base + row*L + l] over `rows`
// rows, W(row) = scale * prod_b pair_b[(row >> b) & 1] for the task's bit pairs
// (LSB first). Entries at or past `available` read as zero.
struct ProjTask {
  uint64_t base;      // entry offset of row 0
  uint64_t rows;      // power of two
  uint32_t nbits;     // log2(rows)
  uint32_t pair_off;  // index of this task's first pair in the pair table
  E8 scale;
};
// Returns tasks.size() x L lane sums.
std::vector<E8> tensor_project(DevCtx& c, const Val* bva, uint64_t available, int L, const std::vector<ProjTask>& tasks,
                               const std::vector<std::array<E8, 2>>& pairs);
// Packed EO (eo.rs:97-181) from projections at stride 2s. Each task is a
// query term of at least 2s entries (its offset a multiple of 2s): rows =
// len / 2s superrows, pairs = the term's factors above the row bit 0 of
// stride s. The chunk lane sums (chunks x 2s) stay on the device.
struct EoProjection {
  DVec parts;
  std::vector<uint32_t> owner;  // task of each chunk
  int s = 0;
};
EoProjection eo_project(DevCtx& c, const Val* bva, uint64_t available, int s, const std::vector<ProjTask>& tasks,
                        const std::vector<std::array<E8, 2>>& pairs);
// The stride-s projections (each chunk's two halves joined with its task's
// row-bit-0 factor f[task]) correlated with the tasks' lane weights (lows:
// tasks x s): pos[d-1] = sum proj[lane] low[lane+d], neg[d-1] = sum
// proj[lane+d] low[lane] (tbv.rs:361-392), and zero = sum proj[lane] low[lane],
// the terms' part of sum_i v[i] q[i].
struct EoCorr {
  std::vector<E8> pos, neg;
  E8 zero;
};
EoCorr eo_correlations(DevCtx& c, const EoProjection& pr, const std::vector<std::array<E8, 2>>& f,
                       const std::vector<E8>& lows);
// The terms' part of the first binary round's cross terms, sum_j w[2j] q[2j+1]
// and sum_j w[2j+1] q[2j], once the packed challenge c is known: pinv[l] =
// c^-l (l < s), coef[task] = (sum_l low[l] c^l) x the task's row-bit-0 factor.
std::pair<E8, E8> eo_first_cross(DevCtx& c, const EoProjection& pr, const std::vector<E8>& pinv,
                                 const std::vector<std::array<E8, 2>>& coef);
// The EO query of a binary level L is a sum of tensor terms (tbv.rs:187-284
// folded by eo.rs:57-79); the fold kernels below project each round's output
// pairs on it instead of reading a materialized query. At output level L a
// term of 2^nf entries (nf >= log2 s + L + 1) covers the pairs [pair0, pair0 +
// 2^(nf - log2 s - L - 1)); its value at output 2m + x of its range is
// a_L f[ls+L][x] prod_b f[ls+L+1+b][bit b of m - pair0] (a_L: the term's scale
// folded over the lower bits), so the round's cross terms sum_m w'[2m] q'[2m+1]
// and sum_m w'[2m+1] q'[2m] add g0 = a_L f[ls+L][1] and g1 = a_L f[ls+L][0]
// times the projections of w' on the term's pair tensor.
constexpr int OTF_PPT = 8;    // pairs per thread in one chunk
constexpr int OTF_LCP = 11;   // log2 pairs per chunk (256 threads x OTF_PPT)
constexpr uint32_t OTF_PIECE = 2048;  // pairs of one piece
// A term projected in whole chunks (at least one chunk long), disjoint from the others.
struct OtfTermDev {
  uint64_t pair0;
  uint32_t lpairs;  // log2 of the pair count, at least OTF_LCP
  uint32_t fbase;   // factor-pair index of the pair-index bit 0 (the term's bit ls + L + 1)
  E8 g0, g1;
};
// Up to OTF_PIECE pairs of a term projected pair by pair (terms shorter than a
// chunk, and terms that overlap the chunked ones).
struct OtfPieceDev {
  uint64_t pair0;  // first pair at the level
  uint64_t rel0;   // its pair index inside the term
  uint32_t count, fbase, nbits, pad;
  E8 g0, g1;
};
// A term of at most one output at the level: value v at output j, paired with output j ^ 1.
struct OtfItemDev {
  uint64_t j, pad;
  E8 v;
};
struct OtfLevel {
  std::vector<OtfTermDev> fast;  // sorted by pair0
  std::vector<OtfPieceDev> pieces;
  std::vector<OtfItemDev> items;
};
// Device state of one EO's on-the-fly rounds: every term's factor pairs and
// the per-level tables (chunk coefficients, the chunked terms' low-bit tables).
class OtfQuery {
 public:
  OtfQuery(DevCtx& c, const std::vector<std::array<E8, 2>>& factors, uint64_t max_chunks, size_t max_fast,
           size_t max_pieces, size_t max_items);
  // The first binary round read from the BVA (rows = logical / s, rows / 4 a
  // multiple of the chunk): ow[i] = sum_{L<2s} v[2is + L] pinv2[L] (the fold of
  // witness rows 2i, 2i+1 by r^-1) for i < rows / 2; returns the cross terms over ow.
  std::pair<E8, E8> first_fold(DevCtx& c, const Val* bva, uint64_t available, uint64_t rows, int s,
                               const std::vector<E8>& pinv2, const OtfLevel& lv, E8* ow);
  // ow[i] = w[2i] + w[2i+1] ri, i < h (h / 2 a multiple of the chunk); returns the cross terms over ow.
  std::pair<E8, E8> fold(DevCtx& c, const E8* w, E8* ow, uint64_t h, const E8& ri, const OtfLevel& lv);

 private:
  uint32_t prepare(DevCtx& c, const OtfLevel& lv, uint64_t chunks, bool mfma_tables);
  std::pair<E8, E8> finish(DevCtx& c, const OtfLevel& lv, const E8* ow, uint32_t groups, uint32_t gs);
  DVec f_, terms_, desc_, tlo_, thi_, pieces_, items_;
  uint64_t max_chunks_;
  size_t max_fast_, max_pieces_, max_items_;
};
// Packed EO witness: w[row] = sum_lane v[row*s + lane] * pinv[lane] for rows < n_rows.
void eo_witness(DevCtx& c, const Val* bva, uint64_t available, uint64_t n_rows, int s, const std::vector<E8>& pinv,
                E8* w);
// out[start + j] += scale * prod_b pair_b[(j >> b) & 1], j < 2^nbits, per task (tasks disjoint).
struct FillTask { uint64_t start; uint32_t nbits, pair_off; E8 scale; };
void tensor_fill_add(DevCtx& c, E8* out, const std::vector<FillTask>& tasks, const std::vector<std::array<E8, 2>>& pairs);
// The same with assignment, over out[0, n): task rows are written without
// being read and the rows outside every task are zeroed.
void tensor_fill(DevCtx& c, E8* out, uint64_t n, const std::vector<FillTask>& tasks,
                 const std::vector<std::array<E8, 2>>& pairs);
// values[i] = sum_b fw[b] * packed(b, i), lane j of packed(b, i) = entry b*block_len + LANES*i + j
void block_fold(DevCtx& c, const Val* bva, uint64_t available, uint64_t block_len, uint64_t blocks,
                const std::vector<E8>& fw, E8* values, uint64_t n);
// weights[i] = sum_j eq(i).c[j] * images[j], eq = hi (x) lo tables (on st, or on c.st when st is null)
void weight_images(DevCtx& c, E8* out, const DVec& hi, const DVec& lo, const std::array<E8, LANES>& images,
                   hipStream_t st = nullptr);
namespace dv {
// sc_msg launched on another stream `st` (its block sums in a partial area of
// their own); finish waits for it and returns the message.
struct SideMsg { uint32_t groups = 0; hipEvent_t done = nullptr; };
SideMsg sc_msg_launch(DevCtx& c, hipStream_t st, const E8* v, const E8* w, uint64_t h);
std::pair<E8, E8> sc_msg_finish(DevCtx& c, SideMsg& m);
}  // namespace dv

}  // namespace qf
