// Launch interfaces of the HIP kernels (device code lives in *.hip files).
#pragma once
#include "qf_params.hpp"
#include <string>

namespace qf {

#define QF_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) throw QfError(QF_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Stream-ordered device allocation from the device's default memory pool,
// which keeps freed blocks cached for reuse: no device-wide synchronisation
// and no page-table setup on the hot path.
void* dmalloc(size_t bytes, hipStream_t st);
void dfree(void* p, hipStream_t st);
// The stream of the innermost device context on this thread (null stream if none).
hipStream_t current_stream();
// This thread's persistent work stream (non-blocking, created on first use and
// kept for the life of the process): commits and proofs issued from one
// thread run on it, so pooled blocks are reused in stream order.
hipStream_t thread_stream();
// A second persistent stream of this thread (default priority): the commit
// hashes block b on it while the work stream encodes block b + 1.
hipStream_t thread_stream2();
// Once per process, before the library's first device call: host waits on
// the device spin instead of sleeping (the opening's round trips are short).
// Has no effect when the process already holds a device context.
void device_setup();
// Returns every idle block of the caching allocator to the device (before a
// large allocation outside the pool, e.g. a resident BVA).
void release_cached_blocks();
// Bytes held idle by the caching allocator (free for the library's next allocations).
size_t cached_idle_bytes();

struct CommitGeom {
  int log_n;          // log2 of message rows per column (chunk_log - 8)
  LeafPrefix leaf;
  InnerPrefix inner;
};

// Commits one PCS block (entries [first_entry, first_entry + 2^c) of the BVA;
// entries at or past `active` are zero): pack + Mobius into the message half
// of `cw`, encode the parity half, hash the rows and build all Merkle levels
// in `tree` (2R - 1 digests, leaves first). Everything is enqueued on `st`.
// Every `alpha` below is alpha_1 in its encode form (v8_encode_form).
// ev (optional, 6 events): [0] start, [1] before the leaf kernel, [2] after it,
// [3] end, [4]/[5] around the Mobius+WHT pass of the last high row group.
void launch_commit_block(const CommitGeom& g, const Val* bva, uint64_t first_entry, uint64_t active, E8* cw,
                         const E8* alpha, uint32_t* tree, hipStream_t st, hipEvent_t* ev = nullptr);
// The encoding half of launch_commit_block: pack + Mobius into the message half
// of `cw` and the parity half, on `st` (mw: optional 2 events around the M_W pass).
void launch_commit_encode(const CommitGeom& g, const Val* bva, uint64_t first_entry, uint64_t active, E8* cw,
                          const E8* alpha, hipStream_t st, hipEvent_t* mw = nullptr);
// out[i] = encode form of in[i] (Montgomery form for KoalaBear, a copy for F65521).
void launch_encode_form(const E8* in, E8* out, uint64_t n, hipStream_t st);
// ev (optional, 4 events): recorded at start, before the leaf kernel, after it, and at the end.

// Building blocks shared with the opening (all asynchronous on `st`):
// WHT over the L low index bits of `vectors` vectors spaced `stride` apart.
void launch_wht(E8* v, uint64_t vectors, uint64_t stride, int L, hipStream_t st);
// Parity halves of `cols` columns (R = 2^(L+1) rows each) from their message halves.
void launch_encode_parity(E8* cw, uint64_t cols, int L, const E8* alpha, hipStream_t st);
// Leaves (prefix chosen by `leaf`) and all inner levels of a 32-column codeword.
void launch_tree(const E8* cw, uint64_t R, const LeafPrefix& leaf, const InnerPrefix& inner, uint32_t* tree,
                 hipStream_t st, hipEvent_t after_leaf = nullptr);
// Commit of an extension matrix: flat 32 x 2^L message, no Mobius, extension leaves.
void launch_commit_extension(const E8* values, E8* cw, int L, const E8* alpha, const LeafPrefix& leaf,
                             const InnerPrefix& inner, uint32_t* tree, hipStream_t st);
// Commit of a flat packed vector of 32 * 2^L elements (the folded block of a batched opening).
// With `side`, the message rows' leaves hash on that stream beside the parity passes.
void launch_commit_values(const CommitGeom& g, const E8* values, E8* cw, const E8* alpha, uint32_t* tree,
                          hipStream_t st, hipStream_t side = nullptr);

// The tile-pass autotuner's choices so far, as a JSON array (qf_tile.hip).
std::string autotune_report();
// Whether tile passes issued from this thread may time new geometries (the
// autotuner allocates scratch buffers); when not, an untuned geometry takes
// the built-in launch shape.
void set_tuning_allowed(bool allowed);
// Compute units of the current device.
int device_cus();
// Whether tile passes may run on the quad-split kernels (QF_QTILE=0 keeps the
// register-radix ones); the commit schedule fuses WHT . alpha . WHT into one
// pass when they do.
bool quad_tiles_enabled();

// One register-radix tile pass (qf_tile.hip); false when the geometry does not
// fit a 4096-element tile. Its launch shape is the autotuner's pick for the
// geometry (timed on scratch buffers the first time the geometry runs).
enum : int { TP_M = 0, TP_M_W = 1, TP_W = 2, TP_WAW = 3, TP_COPY_W = 4, TP_COPY_WAW = 5, TP_AW = 6 };
// With xform_vectors the pass also transforms the index bits of the vectors a
// tile holds (s = 0, t < 12: 2^(12-t) vectors per tile).
bool tile_pass(int mode, bool pack, const void* in, E8* out1, E8* out2, const E8* alpha, uint64_t vectors,
               uint64_t in_vstride, uint64_t out_vstride, int L, int s, int t, uint64_t block_first,
               uint64_t available, hipStream_t st, bool xform_vectors = false);

}  // namespace qf
