// Launch interfaces of the HIP kernels (device code lives in *.hip files).
#pragma once
#include "qf_params.hpp"

namespace qf {

#define QF_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) throw QfError(QF_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct CommitGeom {
  int log_n;          // log2 of message rows per column (chunk_log - 8)
  LeafPrefix leaf;
  InnerPrefix inner;
};

// Commits one PCS block: pack + Mobius into the message half of `cw`,
// encode the parity half, hash the rows and build all Merkle levels in `tree`
// (2R - 1 digests, leaves first). Everything is enqueued on `st`.
void launch_commit_block(const CommitGeom& g, const int8_t* bva, uint64_t block_first, uint64_t available, E8* cw,
                         const E8* alpha, uint32_t* tree, hipStream_t st, hipEvent_t* ev = nullptr);
// ev (optional, 4 events): recorded at start, before the leaf kernel, after it, and at the end.

}  // namespace qf
