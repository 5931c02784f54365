// Device-resident fold rounds with a device copy of the PCS transcript
// (qf_rounds.hip).
#pragma once
#include "qf_vec.hpp"
#include <vector>

namespace qf {

struct Pcs;

// The PCS transcript's state as the device keeps it: the BLAKE2s hasher
// (chaining value, pending block, byte count) and the challenge counter.
struct DevPcsState {
  uint32_t h[8];
  uint8_t buf[64];
  uint32_t buflen;
  uint32_t pad;
  uint64_t total;
  uint64_t counter;
};

// One device round: the cross terms it absorbed and the challenge it drew.
struct DevRound { E8 lc, rc, ch; };

// QF_DEVICE_ROUNDS=1 runs the fold rounds on the device; by default they run
// on the host (a launch and one small download per round), which measures
// faster on MI355X: the device transcript step costs about as much as the
// host round trip it replaces.
bool device_rounds_on();

// Runs the fold rounds of the length-n device vectors a, b (a folded with
// ch^-1, b with ch) from n down to `stop`, starting from the transcript state
// of `tr` and the current cross terms (lc, rc); returns every round's
// (lc, rc, ch) for the host to replay, sets n = stop and (lc, rc) to the
// cross terms of the folded vectors. `tr` is not changed.
std::vector<DevRound> device_fold_rounds(DevCtx& c, E8* a, E8* b, uint64_t& n, uint64_t stop, const Pcs& tr,
                                         const char* label, E8& lc, E8& rc);

}  // namespace qf
