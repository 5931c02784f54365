// Device-resident fold rounds of LiteBullet (qafold_short.rs:1903-1948,
// prove_litebullet) and of the Decide multi-evaluation
// (qafold_short_decide.rs:486-551), with a device copy of the PCS transcript
// (qafold_short_pc.rs:2011-2229: ShortTranscriptWriter, derive_digest,
// derive_extension_challenge).
//
// A round reads the cross terms (lc, rc) of the current vectors, absorbs them
// as two responses, draws the fold challenge ch (rejection sampling over
// BLAKE2s digests, nonzero), inverts it and folds a' = a_lo + ch^-1 a_hi,
// b' = b_lo + ch b_hi, which also yields the next round's cross terms. On the
// host each round is a launch, a synchronisation and a transcript step. Here
// each round is ONE launch with no host wait: every workgroup folds its slice
// and writes its cross-term partials; the workgroup that finishes last (an
// atomic ticket after a device-scope release fence) sums the partials and runs
// the next round's transcript step on its first wave, leaving the next
// challenges in device memory for the next launch. The host uploads one
// argument block, downloads the round log once and replays the transcript
// from the same state; a challenge that differs from the replay fails the
// proof.
//
// The transcript wave keeps the BLAKE2s chaining value and counters in
// wave-uniform registers and the pending 64-byte block in LDS. The bytes a
// round absorbs come from frame templates built on the host (labels and
// lengths in place, zeros where the cross terms, counter and challenge go),
// patched in LDS and appended up to 64 bytes per step by the wave's lanes.
#include "qf_protocol.hpp"
#include "qf_rounds.hpp"
#include <cstring>

namespace qf {
namespace {

constexpr int FWG = 256;        // threads of the fold kernel
constexpr int RPSTRIDE = 32;    // partial row stride (two extension elements, up to 8 lanes each)
constexpr uint32_t MAX_RG = 1024;
constexpr int MAX_ATTEMPTS = 256;  // rejection sampling bound (acceptance >= 0.97 per attempt)
constexpr int STAGE = 192;         // LDS staging bytes for one template

// Argument block, uploaded once per call.
struct RoundArgs {
  DevPcsState init;            // the host transcript's state before the first round
  E8 cross0[2];                // cross terms of the input vectors
  uint32_t frob1[LANES];       // x -> x^p scales coefficient i by frob1[i]
  uint32_t rounds;
  uint32_t resp_len, derive_len, fc_len;
  uint32_t lc_off, rc_off, ctr_off, attempt_off, ch_off;
  uint8_t resp[STAGE];         // frame("response", lc) || frame("response", rc)
  uint8_t derive[STAGE];       // frame("challenge-label", label) || ctr || 0 || attempt
  uint8_t fc[STAGE];           // frame("field-challenge-label", label) || frame("field-challenge", ch)
};

// Device block: state, challenges (two buffers), output cross terms, control words, round log, partials.
struct RoundCtl {
  DevPcsState state;
  E8 xy[2][2];                 // round r folds with xy[r & 1] = (ch^-1, ch)
  E8 cross_out[2];             // cross terms after the last fold
  uint32_t fail;
  uint32_t ticket;
  uint32_t pad[2];
};

// ---------------------------------------------------------------- transcript on one wave
struct WaveHasher {   // chaining value and counts wave-uniform; the pending block in LDS
  uint32_t h[8];
  uint32_t buflen;
  uint64_t total;
  uint32_t* buf;      // 16 LDS words
};

__device__ inline uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ void wh_compress(WaveHasher& s, bool last, uint32_t t) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 16; i++) m[i] = uni(s.buf[i]);
  b2s_compress(s.h, m, t, last);
}
// Appends src[0..len) (LDS bytes) as the host Hasher::update does: a full
// pending block is compressed only when more bytes arrive.
__device__ __forceinline__ void wh_absorb(WaveHasher& s, const uint8_t* src, uint32_t len, int lane) {
  uint8_t* b = reinterpret_cast<uint8_t*>(s.buf);
  for (uint32_t pos = 0; pos < len;) {
    if (s.buflen == 64) {
      s.total += 64;
      wh_compress(s, false, uint32_t(s.total));
      s.buflen = 0;
    }
    const uint32_t k = min(64u - s.buflen, len - pos);
    __builtin_amdgcn_wave_barrier();
    if (uint32_t(lane) < k) b[s.buflen + lane] = src[pos + lane];
    __builtin_amdgcn_wave_barrier();
    s.buflen += k;
    pos += k;
  }
}
// The digest words of the absorbed bytes, computed on a copy (the copy's
// pending block goes to `tmp`, 16 LDS words).
__device__ __forceinline__ void wh_finish(const WaveHasher& s, uint32_t* tmp, int lane, uint32_t out[8]) {
  WaveHasher c = s;
  c.buf = tmp;
  const uint8_t* from = reinterpret_cast<const uint8_t*>(s.buf);
  uint8_t* to = reinterpret_cast<uint8_t*>(tmp);
  __builtin_amdgcn_wave_barrier();
  to[lane] = uint32_t(lane) < s.buflen ? from[lane] : 0;
  __builtin_amdgcn_wave_barrier();
  wh_compress(c, true, uint32_t(c.total + c.buflen));
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = c.h[i];
}
// Copies n template bytes to the stage (lanes in parallel).
__device__ __forceinline__ void stage_copy(uint8_t* stage, const uint8_t* tmpl, uint32_t n, int lane) {
  for (uint32_t j = lane; j < n; j += 64) stage[j] = tmpl[j];
}
__device__ __forceinline__ void stage_u64(uint8_t* stage, uint32_t off, uint64_t v, int lane) {
  if (lane < 8) stage[off + lane] = uint8_t(v >> (8 * lane));
}
// The 16 canonical bytes of v; each lane picks its byte out of the four words
// with constant indexes, so v stays in registers.
__device__ inline void stage_e8(uint8_t* stage, uint32_t off, const E8& v, int lane) {
  struct W4 { uint32_t w[4]; };
  const W4 x = __builtin_bit_cast(W4, v);
  const int q = lane >> 2;
  const uint32_t word = q == 0 ? x.w[0] : q == 1 ? x.w[1] : q == 2 ? x.w[2] : x.w[3];
  if (lane < 16) stage[off + lane] = uint8_t(word >> (8 * (lane & 3)));
}

// canonical digest bytes -> element; KoalaBear clears bit 7 of bytes 3, 7, 11, 15 first (koalabear.rs:248-255)
__device__ __forceinline__ bool from_digest(const uint32_t w[8], E8& out) {
  bool ok = true;
#ifdef QF_KB
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t v = w[i] & 0x7fffffffu;
    ok = ok && v < P;
    out.c[i] = Coef(v);
  }
#else
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t lo = w[i] & 0xffffu, hi = w[i] >> 16;
    ok = ok && lo < P && hi < P;
    out.c[2 * i] = Coef(lo);
    out.c[2 * i + 1] = Coef(hi);
  }
#endif
  return ok;
}

__device__ inline uint32_t fmul(uint32_t a, uint32_t b) { return mred64(uint64_t(a) * b); }
__device__ inline uint32_t fpow(uint32_t a, uint32_t e) {
  uint32_t r = 1;
  while (e) {
    if (e & 1) r = fmul(r, a);
    a = fmul(a, a);
    e >>= 1;
  }
  return r;
}
__device__ __forceinline__ E8 frob1(const E8& a, const uint32_t* f) {
  E8 r;
#pragma unroll
  for (int i = 0; i < LANES; i++) r.c[i] = Coef(fmul(a.c[i], f[i]));
  return r;
}
// (product of the other conjugates) / norm, as the host's e8_inv
__device__ __forceinline__ E8 inv(const E8& a, const uint32_t* f) {
  E8 prod = frob1(a, f);
  for (int k = 2; k < LANES; k++) prod = frob1(v8_store(v8_mul(v8_load(prod), v8_load(a))), f);
  const E8 norm = v8_store(v8_mul(v8_load(a), v8_load(prod)));
  const uint32_t ni = fpow(norm.c[0], P - 2);
  E8 r;
#pragma unroll
  for (int i = 0; i < LANES; i++) r.c[i] = Coef(fmul(prod.c[i], ni));
  return r;
}

// One round's transcript step on wave 0 of the calling workgroup:
// respond(lc), respond(rc), challenge(label, nonzero) (Pcs::respond /
// Pcs::challenge); writes xy[next] = (ch^-1, ch) and the round's log entry.
// lc, rc are read from cross (LDS).
__device__ __forceinline__ void transcript_round(const RoundArgs& ra, RoundCtl& ctl, uint32_t r, const E8* cross, E8* log,
                                 uint32_t* sbuf, uint32_t* dbuf, uint32_t* tmp, uint8_t* stage, int lane) {
  WaveHasher s;
  const DevPcsState& from = r == 0 ? ra.init : ctl.state;
#pragma unroll
  for (int i = 0; i < 8; i++) s.h[i] = uni(from.h[i]);
  s.buflen = uni(from.buflen);
  s.total = (uint64_t(uni(uint32_t(from.total >> 32))) << 32) | uni(uint32_t(from.total));
  uint64_t counter = (uint64_t(uni(uint32_t(from.counter >> 32))) << 32) | uni(uint32_t(from.counter));
  s.buf = sbuf;
  if (lane < 16) sbuf[lane] = reinterpret_cast<const uint32_t*>(from.buf)[lane];
  const E8 lc = cross[0], rc = cross[1];

  stage_copy(stage, ra.resp, ra.resp_len, lane);
  __builtin_amdgcn_wave_barrier();
  stage_e8(stage, ra.lc_off, lc, lane);
  stage_e8(stage, ra.rc_off, rc, lane);
  __builtin_amdgcn_wave_barrier();
  wh_absorb(s, stage, ra.resp_len, lane);

  E8 ch{};
  bool done = false;
  for (int outer = 0; outer < MAX_ATTEMPTS && !done; outer++) {
    bool got = false;
    for (int a = outer; a < MAX_ATTEMPTS && !got; a++) {
      WaveHasher d = s;  // derive(label, counter, attempt) on a copy of the state
      d.buf = dbuf;
      __builtin_amdgcn_wave_barrier();
      if (lane < 16) dbuf[lane] = sbuf[lane];
      stage_copy(stage, ra.derive, ra.derive_len, lane);
      __builtin_amdgcn_wave_barrier();
      stage_u64(stage, ra.ctr_off, counter, lane);
      stage_u64(stage, ra.attempt_off, uint64_t(a), lane);
      __builtin_amdgcn_wave_barrier();
      wh_absorb(d, stage, ra.derive_len, lane);
      uint32_t dig[8];
      wh_finish(d, tmp, lane, dig);
      got = from_digest(dig, ch);
    }
    if (!got) break;
    bool zero = true;
#pragma unroll
    for (int i = 0; i < LANES; i++) zero = zero && ch.c[i] == 0;
    done = !zero;
  }
  if (!done) {  // the host replay reports the mismatch
    if (lane == 0) ctl.fail = 1;
    ch = E8{};
    ch.c[0] = 1;
  }
  counter++;
  __builtin_amdgcn_wave_barrier();
  stage_copy(stage, ra.fc, ra.fc_len, lane);
  __builtin_amdgcn_wave_barrier();
  stage_e8(stage, ra.ch_off, ch, lane);
  __builtin_amdgcn_wave_barrier();
  wh_absorb(s, stage, ra.fc_len, lane);

  const E8 ci = inv(ch, ra.frob1);
  if (lane < 16) reinterpret_cast<uint32_t*>(ctl.state.buf)[lane] = sbuf[lane];
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 8; i++) ctl.state.h[i] = s.h[i];
    ctl.state.buflen = s.buflen;
    ctl.state.total = s.total;
    ctl.state.counter = counter;
    ctl.xy[r & 1][0] = ci;
    ctl.xy[r & 1][1] = ch;
    log[3 * r] = lc;
    log[3 * r + 1] = rc;
    log[3 * r + 2] = ch;
  }
}

// FOLD = false: one workgroup runs round 0's transcript on the input cross
// terms. FOLD = true: round r's fold of both vectors over their top halves
// with xy[r & 1], fused with the folded vectors' cross terms; the last
// workgroup to finish runs round r + 1's transcript (or, after the last
// round, stores the cross terms in cross_out).
template <bool FOLD>
__global__ __launch_bounds__(FWG) void k_round(E8* __restrict__ a, E8* __restrict__ b, uint64_t h,
                                               RoundCtl* __restrict__ ctl, const RoundArgs* __restrict__ ra,
                                               E8* __restrict__ log, uint32_t* __restrict__ partials, uint32_t r) {
  __shared__ uint32_t red[FWG][2 * LANES];
  __shared__ E8 cross[2];
  __shared__ uint32_t sbuf[16], dbuf[16], tmp[16];
  __shared__ uint8_t stage[STAGE];
  __shared__ uint32_t last;
  const int t = threadIdx.x;
  uint32_t next = 0;  // round whose transcript this workgroup runs
  if (FOLD) {
    const E8* xy = ctl->xy[r & 1];
    const V8 vx = v8_encode_form(v8_load(xy[0])), vy = v8_encode_form(v8_load(xy[1]));
    const uint64_t g = h / 2;
    V8 s0 = v8_zero(), s1 = v8_zero();
    for (uint64_t i = uint64_t(blockIdx.x) * FWG + t; i < g; i += uint64_t(gridDim.x) * FWG) {
      const V8 a0 = v8_add(v8_load(a[i]), v8_mul_enc(v8_load(a[h + i]), vx));
      const V8 a1 = v8_add(v8_load(a[g + i]), v8_mul_enc(v8_load(a[h + g + i]), vx));
      const V8 b0 = v8_add(v8_load(b[i]), v8_mul_enc(v8_load(b[h + i]), vy));
      const V8 b1 = v8_add(v8_load(b[g + i]), v8_mul_enc(v8_load(b[h + g + i]), vy));
      a[i] = v8_store(a0);
      a[g + i] = v8_store(a1);
      b[i] = v8_store(b0);
      b[g + i] = v8_store(b1);
      s0 = v8_add(s0, v8_mul(a0, b1));
      s1 = v8_add(s1, v8_mul(a1, b0));
    }
#pragma unroll
    for (int k = 0; k < LANES; k++) {
      red[t][k] = s0.c[k];
      red[t][LANES + k] = s1.c[k];
    }
    __syncthreads();
    for (int w = FWG / 2; w > 0; w /= 2) {
      if (t < w)
        for (int k = 0; k < 2 * LANES; k++) red[t][k] = madd(red[t][k], red[t + w][k]);
      __syncthreads();
    }
    if (t < 2 * LANES) partials[blockIdx.x * RPSTRIDE + t] = red[0][t];
    // release this workgroup's partials, then take a ticket; the last one in sums them
    __threadfence();
    __syncthreads();
    if (t == 0) last = atomicAdd(&ctl->ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    const uint32_t groups = gridDim.x;
    const int k = t & 15, slice = t >> 4;
    if (k < 2 * LANES) {  // red rows hold 2 x LANES words
      uint32_t acc = 0;
      for (uint32_t q = slice; q < groups; q += FWG / 16) acc = madd(acc, partials[q * RPSTRIDE + k]);
      red[slice][k] = acc;
    }
    __syncthreads();
    if (t < 2 * LANES) {
      uint32_t s = 0;
      for (int j = 0; j < FWG / 16; j++) s = madd(s, red[j][t]);
      reinterpret_cast<Coef*>(cross)[t] = Coef(s);  // cross[0] lanes, then cross[1] lanes
    }
    if (t == 0) ctl->ticket = 0;  // for the next launch
    __syncthreads();
    next = r + 1;
    if (next == ra->rounds) {
      if (t < 2) ctl->cross_out[t] = cross[t];
      return;
    }
  } else {
    if (t < 2) cross[t] = ra->cross0[t];
    __syncthreads();
  }
  if (t >= 64) return;
  transcript_round(*ra, *ctl, next, cross, log, sbuf, dbuf, tmp, stage, t);
}

uint32_t fold_groups(uint64_t g) {
  const uint64_t n = (g + FWG - 1) / FWG;
  return uint32_t(n < 1 ? 1 : (n > MAX_RG ? MAX_RG : n));
}

// frame(label, payload of n bytes) into out; returns the payload's offset
uint32_t put_frame(std::vector<uint8_t>& out, const char* label, const void* payload, size_t n) {
  auto u64 = [&](uint64_t v) { for (int i = 0; i < 8; i++) out.push_back(uint8_t(v >> (8 * i))); };
  const size_t l = std::strlen(label);
  u64(l);
  out.insert(out.end(), label, label + l);
  u64(n);
  const uint32_t at = uint32_t(out.size());
  const uint8_t* p = static_cast<const uint8_t*>(payload);
  out.insert(out.end(), p, p + n);
  return at;
}

}  // namespace

bool device_rounds_on() {
  const char* e = std::getenv("QF_DEVICE_ROUNDS");
  return e && e[0] == '1';
}

std::vector<DevRound> device_fold_rounds(DevCtx& c, E8* a, E8* b, uint64_t& n, uint64_t stop, const Pcs& tr,
                                         const char* label, E8& lc, E8& rc) {
  require(pow2(n) && pow2(stop) && stop >= 4 && stop < n, QF_ERR_ARGUMENT, "device rounds length");
  const int rounds = ilog2u(n / stop);
  RoundArgs ra{};
  std::memcpy(ra.init.h, tr.h.h, sizeof ra.init.h);
  std::memcpy(ra.init.buf, tr.h.buf, 64);
  ra.init.buflen = tr.h.buflen;
  ra.init.total = tr.h.total;
  ra.init.counter = tr.counter;
  ra.cross0[0] = lc;
  ra.cross0[1] = rc;
  const auto& ft = frob_table();
  for (int i = 0; i < LANES; i++) ra.frob1[i] = ft[1][i];
  ra.rounds = uint32_t(rounds);
  // the byte templates, framed as Pcs::respond, Pcs::derive and Pcs::challenge frame them
  const E8 zero{};
  const uint8_t z8[8] = {};
  std::vector<uint8_t> resp, der, fc;
  ra.lc_off = put_frame(resp, "response", &zero, 16);
  ra.rc_off = put_frame(resp, "response", &zero, 16);
  put_frame(der, "challenge-label", label, std::strlen(label));
  ra.ctr_off = uint32_t(der.size());
  der.insert(der.end(), z8, z8 + 8);  // counter
  der.insert(der.end(), z8, z8 + 8);  // 0
  ra.attempt_off = uint32_t(der.size());
  der.insert(der.end(), z8, z8 + 8);  // attempt
  put_frame(fc, "field-challenge-label", label, std::strlen(label));
  ra.ch_off = put_frame(fc, "field-challenge", &zero, 16);
  require(resp.size() <= STAGE && der.size() <= STAGE && fc.size() <= STAGE, QF_ERR_ARGUMENT, "round label");
  ra.resp_len = uint32_t(resp.size());
  ra.derive_len = uint32_t(der.size());
  ra.fc_len = uint32_t(fc.size());
  std::memcpy(ra.resp, resp.data(), resp.size());
  std::memcpy(ra.derive, der.data(), der.size());
  std::memcpy(ra.fc, fc.data(), fc.size());

  // one device block: arguments, control (zeroed), round log, partials
  const size_t log_bytes = size_t(3 * rounds) * 16;
  const size_t a_off = 0, ctl_off = (sizeof(RoundArgs) + 255) & ~size_t(255);
  const size_t log_off = ctl_off + ((sizeof(RoundCtl) + 255) & ~size_t(255));
  const size_t part_off = log_off + ((log_bytes + 255) & ~size_t(255));
  const size_t bytes = part_off + size_t(MAX_RG) * RPSTRIDE * 4;
  std::vector<uint8_t> up(log_off, 0);
  std::memcpy(up.data() + a_off, &ra, sizeof ra);
  char* mem = static_cast<char*>(dmalloc(bytes, c.st));
  RoundArgs* d_ra = reinterpret_cast<RoundArgs*>(mem + a_off);
  RoundCtl* ctl = reinterpret_cast<RoundCtl*>(mem + ctl_off);
  E8* log = reinterpret_cast<E8*>(mem + log_off);
  uint32_t* partials = reinterpret_cast<uint32_t*>(mem + part_off);
  QF_HIP(hipMemcpyAsync(mem, up.data(), up.size(), hipMemcpyHostToDevice, c.st));
  hipLaunchKernelGGL(k_round<false>, dim3(1), dim3(FWG), 0, c.st, a, b, uint64_t(0), ctl, d_ra, log, partials, 0u);
  for (int r = 0; r < rounds; r++) {
    const uint64_t h = n / 2;
    hipLaunchKernelGGL(k_round<true>, dim3(fold_groups(h / 2)), dim3(FWG), 0, c.st, a, b, h, ctl, d_ra, log,
                       partials, uint32_t(r));
    n = h;
  }
  QF_HIP(hipGetLastError());
  // the control block and the log are adjacent: one download
  std::vector<uint8_t> down(part_off - ctl_off);
  QF_HIP(hipMemcpyAsync(down.data(), mem + ctl_off, down.size(), hipMemcpyDeviceToHost, c.st));
  QF_HIP(hipStreamSynchronize(c.st));
  dfree(mem, c.st);
  RoundCtl out_ctl;
  std::memcpy(&out_ctl, down.data(), sizeof out_ctl);
  require(!out_ctl.fail, QF_ERR_TRANSCRIPT, "device transcript: challenge sampling did not finish");
  const E8* hl = reinterpret_cast<const E8*>(down.data() + (log_off - ctl_off));
  std::vector<DevRound> out(rounds);
  for (int r = 0; r < rounds; r++) out[r] = {hl[3 * r], hl[3 * r + 1], hl[3 * r + 2]};
  lc = out_ctl.cross_out[0];
  rc = out_ctl.cross_out[1];
  return out;
}

}  // namespace qf
