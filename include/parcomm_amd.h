/* parcomm_amd: MI355X-native QAFold prover, C ABI.
 *
 * The entry points replace the Rust library API of the reference crates
 * `parcomm-qafold` (qafold/f65521/src/protocol.rs, F65521 / Ext8, i8 entries)
 * and its KoalaBear twin (qafold/koalabear/src/protocol.rs, KoalaBear / Ext4,
 * i16 entries), which have no FFI of their own; each declaration cites the
 * item it stands in for (file:line in the F65521 crate; the KoalaBear crate has
 * the same items). Like the reference, which ships one crate per field, the
 * library is built once per field: libparcomm_amd.so (F65521) and
 * libparcomm_amd_kb.so (KoalaBear) export the same symbols; qf_field() names
 * the backend of the loaded library. "Entries" below are int8_t in the
 * F65521 build and int16_t in the KoalaBear build; extension elements are 16
 * bytes in both (8 little-endian u16 or 4 little-endian u32 coefficients). Buffers are
 * caller-owned host memory unless the name says `device`; device memory is
 * owned by the handles. Functions return 0 on success or one of the QF_ERR_*
 * codes (the reference engine's error kinds, vendor/qafold/src/lib.rs:20-35,
 * plus device and argument errors); qf_last_error() gives the message of the
 * last failure on the calling thread. No exception crosses this boundary.
 * Handles are independent: distinct handles may be used from distinct threads.
 */
#ifndef PARCOMM_AMD_H
#define PARCOMM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QF_OK 0
#define QF_ERR_ENCODING 1
#define QF_ERR_HASHING 2
#define QF_ERR_INVALID_COMMITMENT 3
#define QF_ERR_INVALID_VARIABLES 4
#define QF_ERR_INVALID_PARAMETERS 5
#define QF_ERR_TRANSCRIPT 6
#define QF_ERR_DEVICE 7
#define QF_ERR_ARGUMENT 8

#define QF_FIELD_F65521 0
#define QF_FIELD_KOALABEAR 1
#define QF_CACHE_CACHED 0
#define QF_CACHE_RECOMPUTE 1

typedef struct qf_params qf_params;
typedef struct qf_committed qf_committed;

/* Message of the last failure on this thread (empty after success). */
const char* qf_last_error(void);

/* Field backend of this library build: QF_FIELD_F65521 or QF_FIELD_KOALABEAR
 * (the reference selects it by crate: qafold/f65521 or qafold/koalabear). */
int qf_field(void);

/* Parameters::new(shapes, Config{chunk_log, packed_log, query_rf_bits})
 * (protocol.rs:109-130): layout, QA sequence, setup commitments. `shapes` holds
 * n (rows, cols) pairs; `field` must be this build's qf_field(). */
int qf_params_new(int field, const uint64_t* shapes, size_t n, uint32_t chunk_log, uint32_t packed_log,
                  uint32_t query_rf_bits, qf_params** out);
void qf_params_free(qf_params* p);

/* PackedPcs::parameter_digest (packed.rs:159-161). */
int qf_params_digest(const qf_params* p, uint8_t out[32]);

/* Layout fields (layout.rs:25-31): active and logical BVA length, effective
 * block length and number of PCS blocks (protocol.rs:116-122). */
int qf_params_info(const qf_params* p, uint64_t* active, uint64_t* logical, uint64_t* block_len, uint64_t* blocks,
                   uint64_t* n_buckets);

/* Bucket i as 8 u64: matrix, row_offset, col_offset, rows, cols,
 * physical_rows, physical_cols, offset (layout.rs:9-19). */
int qf_params_bucket(const qf_params* p, uint64_t i, uint64_t out[8]);

/* Parameters::commit / commit_recomputed (protocol.rs:131-185) over a BVA of
 * `len` == active entries in host memory. Non-physical padding must be
 * zero (layout.rs:136-155). cache_mode QF_CACHE_CACHED keeps every codeword
 * and tree resident in HBM; QF_CACHE_RECOMPUTE keeps roots only. */
int qf_commit(const qf_params* p, const void* values, size_t len, int cache_mode, qf_committed** out);

/* As qf_commit, with the BVA already resident in device memory (16-byte
 * aligned; borrowed: it must stay valid while the returned handle lives). */
int qf_commit_device(const qf_params* p, const void* device_values, size_t len, int cache_mode,
                     qf_committed** out);
void qf_committed_free(qf_committed* c);

/* Commitment.roots (protocol.rs:33-40): blocks x 32 bytes. */
int qf_committed_roots(const qf_committed* c, uint8_t* out, size_t cap_blocks);

/* Copy block `block`'s column-major codeword (32 x R extension elements, 16
 * bytes each) to host memory; only for cached commitments. `cap` is in bytes. */
int qf_committed_codeword(const qf_committed* c, uint64_t block, void* out, size_t cap);

/* Device time in milliseconds of the commit's kernels (HIP events on the
 * streams they run on), summed over the encoded blocks: all work, the
 * leaf-hash kernel, the encoding before the leaves, and the Mobius+WHT pass
 * of the last high row group; timed_blocks counts those blocks. A commit that
 * replayed a captured HIP graph (the second and later commits of a parameter
 * set, unless QF_COMMIT_GRAPH=0) reports total_ms only, with timed_blocks 0.
 * Any pointer may be null. */
int qf_committed_timing(const qf_committed* c, double* total_ms, double* leaf_ms, double* encode_ms, double* mw_ms,
                        uint64_t* timed_blocks);

typedef struct qf_proof qf_proof;

/* Size in bytes of a packed claim list for these parameters: for every matrix
 * in order, its row point (log2(next_pow2(rows)) extension elements), its column
 * point (log2(next_pow2(cols)) elements) and its value, each element its 16
 * canonical bytes (tbv.rs:6-11, Claim; layout of the reference's serde bytes
 * for Scalar). */
int qf_claims_size(const qf_params* p, size_t* bytes);

/* The benchmark statement of the reference CLI (main.rs:188-209): points drawn
 * from the transcript over the commitment statement, values evaluated on the
 * committed BVA (tbv.rs:143-185, evaluate_claims). Writes qf_claims_size bytes. */
int qf_benchmark_claims(const qf_params* p, const qf_committed* c, uint8_t* claims, size_t cap);

/* tbv.rs:143-185 (evaluate_claims): fills the value of every claim in the
 * packed list from its row and column points, on the committed BVA. The
 * values already in the buffer are ignored. */
int qf_evaluate_claims(const qf_params* p, const qf_committed* c, uint8_t* claims, size_t claims_len);

/* Parameters::prove (protocol.rs:201-263): TBV compile, packed even/odd
 * reduction, PCS opening (batched over blocks, or single-block) and Decide,
 * returning the bincode Proof (protocol.rs:57-95, Proof::to_bytes). The
 * commitment is not consumed. Invalid claims (wrong dimensions, non-canonical
 * elements, false values) fail with the reference's error kinds. */
int qf_prove(const qf_params* p, const qf_committed* c, const uint8_t* claims, size_t claims_len, qf_proof** out);
/* Parameters::verify (protocol.rs:266-319): the witness-free verifier over
 * the public roots (n_roots x 32 bytes), the packed claims and the proof
 * bytes. Returns QF_OK when the proof is accepted, else an error code with the
 * rejected check in qf_last_error(). Host only; needs no device. */
int qf_verify(const qf_params* p, const uint8_t* roots, size_t n_roots, const uint8_t* claims, size_t claims_len,
              const uint8_t* proof, size_t proof_len);

/* Batch verification (no reference counterpart; the reference verifies one
 * proof per call): `count` proofs against the same parameters and public
 * roots, proof i with its own packed claims, each checked exactly as by
 * qf_verify, spread over up to `threads` host threads (0: one per hardware
 * thread). statuses[i] receives proof i's qf_verify code. Returns QF_OK when
 * every proof is accepted, else the first rejected proof's code, with its
 * index and check in qf_last_error(). Host only. */
int qf_verify_batch(const qf_params* p, const uint8_t* roots, size_t n_roots, const uint8_t* const* claims,
                    const size_t* claims_lens, const uint8_t* const* proofs, const size_t* proof_lens, size_t count,
                    int threads, int* statuses);

/* Proof bytes, valid while the handle lives. */
int qf_proof_bytes(const qf_proof* pr, const uint8_t** data, size_t* len);
/* Host wall-clock phases of the proof in milliseconds (ProverProfile,
 * protocol.rs:68-80): conversion, packed_eo, pcs_projection, folding,
 * sumcheck, binding, native_opening, decide, total. */
int qf_proof_timing(const qf_proof* pr, double out[9]);
void qf_proof_free(qf_proof* pr);

/* The public statement as the reference serialises it (serde_json of
 * Commitment, protocol.rs:33-40; commitment.json of main.rs). NUL-terminated;
 * *len receives the length without the NUL. */
int qf_statement_json(const qf_params* p, const qf_committed* c, char* out, size_t cap, size_t* len);

/* Device memory for a resident BVA (no reference counterpart: the reference
 * keeps the BVA in host RAM). qf_device_alloc returns zeroed memory. */
int qf_device_alloc(size_t bytes, void** device_out);
int qf_device_upload(void* device_dst, const void* host_src, size_t bytes);
int qf_device_download(void* host_dst, const void* device_src, size_t bytes);
int qf_device_free(void* device_ptr);
/* Device-to-device copy by a plain 16-bytes-per-thread kernel (synchronous;
 * 16-byte aligned buffers and sizes). pattern 0 copies all `bytes`; pattern 1
 * (2) copies only the first (second) 64-byte half of every 128-byte line of
 * the range (128-byte aligned buffers and sizes) and leaves the rest of dst
 * as it was. Used to calibrate the HBM traffic counters against copies of
 * known size, whole lines and the isolated half lines of the commit kernels'
 * 64-byte tile rows. */
int qf_device_copy(void* device_dst, const void* device_src, size_t bytes, int pattern);

/* Free and total bytes of the current device (admission control of the
 * serving loop); free includes the blocks the library's caching allocator
 * holds idle, which its next allocations reuse. */
int qf_device_memory(size_t* free_bytes, size_t* total_bytes);

/* Layout::encode (layout.rs:115-135) on the device: scatters the n row-major
 * matrices (matrix i has lens[i] == rows_i x cols_i entries, in the order of
 * the shapes) into the BVA at device_bva (16-byte aligned, active entries),
 * bucket by bucket, writing zero on every non-physical padding entry. The
 * matrices are device pointers when on_device is nonzero (all buckets in one
 * launch), else host memory staged through the device one matrix at a time.
 * Count and length mismatches fail with QF_ERR_INVALID_PARAMETERS and the
 * reference's messages ("matrix count mismatch", "matrix length mismatch").
 * Synchronous. */
int qf_encode(const qf_params* p, const void* const* matrices, const uint64_t* lens, size_t n, int on_device,
              void* device_bva);
/* One matrix of Layout::encode, for callers that stream a model into HBM one
 * matrix at a time as the reference CLI streams its dense BVA input
 * (main.rs:145-173): writes matrix `matrix`'s buckets (its physical rows and
 * their zero padding) and leaves the rest of the BVA untouched. Encoding
 * every matrix once gives the same BVA as qf_encode. Synchronous. */
int qf_encode_matrix(const qf_params* p, uint64_t matrix, const void* values, uint64_t len, int on_device,
                     void* device_bva);

/* The reference CLI's dense BVA input (main.rs:145-173: the
 * parcomm-dense-bva-stream produced by modelconvert.py, entries in BVA order,
 * int8 or little-endian int16): reads exactly active entries from the file at
 * `path` ("-" for standard input) straight into the device BVA, through two
 * pinned staging buffers so reading and uploading overlap. Input that ends
 * early ("BVA input ends early") or runs past the active length ("excess BVA
 * input") fails with QF_ERR_ENCODING. *bytes_read (optional) receives the
 * bytes consumed. Synchronous. */
int qf_stream_bva(const qf_params* p, const char* path, void* device_bva, uint64_t* bytes_read);

/* Synthetic workload generator (no reference counterpart; the reference CLI
 * streams real checkpoints): fills the active BVA in device memory, from a
 * counter-based generator keyed by `seed`, with clip(round(N(0, sigma)),
 * -128, 127) (F65521 build) or the raw bfloat16 bit pattern of N(0, sigma),
 * rounded to nearest even (KoalaBear build), on every physical entry and zero
 * on boundary padding. */
int qf_synthetic_bva(const qf_params* p, uint64_t seed, float sigma, void* device_out);

/* The tile-pass autotuner's choices so far in this process (no reference
 * counterpart): a JSON array with, per pass geometry (mode, packing, L, s, t,
 * vectors), the launch shapes tried (kernel variant, workgroups per CU, XCD
 * pairing), their times in ms and the one kept. The first pass of each
 * geometry times the candidates on scratch buffers of the same layout; the
 * environment variable QF_AUTOTUNE=0 keeps the built-in shapes, and
 * QF_TUNE_FILE=path reads earlier picks from that file and records new ones
 * there. Writes at most cap bytes including the NUL; *len receives the full
 * length. */
int qf_autotune_report(char* out, size_t cap, size_t* len);

/* Launch form of the opening's per-term kernels (no reference counterpart; the
 * reference runs on CPU threads): QF_LAUNCH_BATCHED (default) issues one launch
 * per phase over every bucket term of every matrix; QF_LAUNCH_PER_MATRIX issues
 * one launch and one host round trip per term, the plain per-matrix GPU form
 * the benchmark compares against. Proofs are identical in both. The form is
 * per host thread: it applies to the calling thread's later calls only, so
 * proofs running on other threads (e.g. serving workers) keep their own. */
#define QF_LAUNCH_BATCHED 0
#define QF_LAUNCH_PER_MATRIX 1
int qf_set_launch_mode(int mode);

/* Kernel-level entry point for the parity tests (no reference counterpart):
 * one register-radix transform pass (mode 0 Mobius, 1 Mobius then WHT into
 * out2, 2 WHT, 3 WHT . alpha . WHT, 4 copy into out1 and WHT into out2, 5 copy
 * into out1 and WHT . alpha . WHT into out2, 6 alpha . (elementwise) then WHT)
 * over
 * `vectors` device vectors of 2^L extension elements, on index bits [s, s+t) and,
 * when xform_vectors is set (s = 0), also on the index bits of the 2^(12-t)
 * vectors that share a tile. Returns QF_ERR_ARGUMENT when the geometry does
 * not fit a tile. Synchronous. */
int qf_test_tile_pass(int mode, const void* device_in, void* device_out1, void* device_out2,
                      const void* device_alpha, uint64_t vectors, uint64_t in_vstride, uint64_t out_vstride, int L,
                      int s, int t, int xform_vectors);

/* Commitment.roots of a device-resident BVA (the roots qf_commit_device
 * gives, in either cache mode) computed without allocating device memory. No
 * counterpart in the reference: this is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Every device buffer of the commit (one block's codeword and
 * Merkle tree, alpha_1, the padding-check flag and ranges) lives in a fixed
 * 64 MiB workspace that each device holds from the library's load until the
 * process ends, so the effective chunk_log is limited to 23 (F65521) or 22
 * (KoalaBear); larger blocks fail with QF_ERR_INVALID_PARAMETERS. The call
 * runs on the null stream and returns when the roots are in `roots`
 * (cap_blocks x 32 bytes, at least the parameters' block count). Calls on one
 * device share the workspace: the library runs them one at a time, and the
 * workspace keeps the last call's contents. Tile passes take an earlier
 * autotuner pick of their geometry or the built-in launch shape (tuning
 * allocates). */
int qf_commit_roots_noalloc(const qf_params* p, const void* device_values, size_t len, uint8_t* roots,
                            size_t cap_blocks);

#ifdef __cplusplus
}
#endif
#endif
