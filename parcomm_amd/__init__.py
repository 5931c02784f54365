"""MI355X-native QAFold prover.

Python mirror of the reference crates' `Parameters` API
(qafold/f65521/src/protocol.rs:18-263 for F65521 / Ext8 with i8 entries,
qafold/koalabear/src/protocol.rs for KoalaBear / Ext4 with i16 entries):
`Config`, `Shape`, `Parameters.new`, `Parameters.commit` /
`commit_recomputed`, `prove`, `verify` and the `Commitment` statement. All work
runs in the HIP libraries built from `csrc/`, one per field
(`libparcomm_amd.so`, `libparcomm_amd_kb.so`); loading fails loudly when the
library of the requested field is missing.
"""
import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Sequence

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparcomm_amd.so")
LIB_PATH_KB = os.path.join(_HERE, "libparcomm_amd_kb.so")

QF_CACHE_CACHED = 0
QF_CACHE_RECOMPUTE = 1
QF_FIELD_F65521 = 0
QF_FIELD_KOALABEAR = 1
F65521 = QF_FIELD_F65521
KOALABEAR = QF_FIELD_KOALABEAR
# entry type and coefficient type of each backend
_FIELD_INFO = {
    QF_FIELD_F65521: {"path": LIB_PATH, "entry": "int8", "coef": "<u2", "lanes": 8, "p": 65521},
    QF_FIELD_KOALABEAR: {"path": LIB_PATH_KB, "entry": "int16", "coef": "<u4", "lanes": 4, "p": 2130706433},
}

_ERRORS = {
    1: "EncodingError",
    2: "HashingError",
    3: "InvalidCommitment",
    4: "InvalidNumberOfVariables",
    5: "InvalidParameters",
    6: "TranscriptError",
    7: "DeviceError",
    8: "ArgumentError",
}


class QfError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{_ERRORS.get(status, status)}: {message}")
        self.status = status


_libs = {}


def field_info(field=QF_FIELD_F65521):
    """Library path, entry dtype, coefficient dtype, lanes and modulus of a backend."""
    if field not in _FIELD_INFO:
        raise ValueError(f"unknown field backend {field}")
    return _FIELD_INFO[field]


def lib(field=QF_FIELD_F65521):
    """The loaded HIP library of a field backend; raises if it has not been built."""
    if field not in _libs:
        path = field_info(field)["path"]
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it with `python -m parcomm_amd.build`"
            )
        L = ctypes.CDLL(path)
        vp, u64, sz, i32, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32
        P = ctypes.POINTER
        L.qf_last_error.restype = ctypes.c_char_p
        L.qf_field.restype = i32
        L.qf_params_new.argtypes = [i32, P(u64), sz, u32, u32, u32, P(vp)]
        L.qf_params_free.argtypes = [vp]
        L.qf_params_free.restype = None
        L.qf_params_digest.argtypes = [vp, ctypes.c_char_p]
        L.qf_params_info.argtypes = [vp, P(u64), P(u64), P(u64), P(u64), P(u64)]
        L.qf_params_bucket.argtypes = [vp, u64, P(u64)]
        L.qf_commit.argtypes = [vp, vp, sz, i32, P(vp)]
        L.qf_commit_device.argtypes = [vp, vp, sz, i32, P(vp)]
        L.qf_committed_free.argtypes = [vp]
        L.qf_committed_free.restype = None
        L.qf_committed_roots.argtypes = [vp, ctypes.c_char_p, sz]
        L.qf_committed_codeword.argtypes = [vp, u64, vp, sz]
        L.qf_committed_timing.argtypes = [vp, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double),
                                          P(ctypes.c_double), P(u64)]
        L.qf_device_alloc.argtypes = [sz, P(vp)]
        L.qf_device_upload.argtypes = [vp, vp, sz]
        L.qf_device_download.argtypes = [vp, vp, sz]
        L.qf_device_free.argtypes = [vp]
        L.qf_device_copy.argtypes = [vp, vp, sz, i32]
        L.qf_synthetic_bva.argtypes = [vp, u64, ctypes.c_float, vp]
        L.qf_encode.argtypes = [vp, P(vp), P(u64), sz, i32, vp]
        L.qf_encode_matrix.argtypes = [vp, u64, vp, u64, i32, vp]
        L.qf_stream_bva.argtypes = [vp, ctypes.c_char_p, vp, P(u64)]
        L.qf_claims_size.argtypes = [vp, P(sz)]
        L.qf_benchmark_claims.argtypes = [vp, vp, ctypes.c_char_p, sz]
        L.qf_prove.argtypes = [vp, vp, ctypes.c_char_p, sz, P(vp)]
        L.qf_evaluate_claims.argtypes = [vp, vp, ctypes.c_char_p, sz]
        L.qf_verify.argtypes = [vp, ctypes.c_char_p, sz, ctypes.c_char_p, sz, ctypes.c_char_p, sz]
        L.qf_verify_batch.argtypes = [vp, ctypes.c_char_p, sz, P(ctypes.c_char_p), P(sz), P(ctypes.c_char_p), P(sz),
                                      sz, i32, P(i32)]
        L.qf_proof_bytes.argtypes = [vp, P(P(ctypes.c_uint8)), P(sz)]
        L.qf_proof_timing.argtypes = [vp, P(ctypes.c_double)]
        L.qf_proof_free.argtypes = [vp]
        L.qf_proof_free.restype = None
        L.qf_statement_json.argtypes = [vp, vp, ctypes.c_char_p, sz, P(sz)]
        L.qf_set_launch_mode.argtypes = [i32]
        L.qf_autotune_report.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        L.qf_device_memory.argtypes = [P(sz), P(sz)]
        L.qf_test_tile_pass.argtypes = [i32, vp, vp, vp, vp, u64, u64, u64, i32, i32, i32, i32]
        if L.qf_field() != field:
            raise RuntimeError(f"{path} implements field {L.qf_field()}, not {field}")
        _libs[field] = L
    return _libs[field]


def _check(status, field=QF_FIELD_F65521):
    if status != 0:
        raise QfError(status, lib(field).qf_last_error().decode())


LAUNCH_BATCHED = 0
LAUNCH_PER_MATRIX = 1


def set_launch_mode(mode, field=QF_FIELD_F65521):
    """Batched (one launch per phase over all matrices) or plain per-matrix launches."""
    _check(lib(field).qf_set_launch_mode(mode), field)


def autotune_report(field=QF_FIELD_F65521):
    """The tile-pass autotuner's choices in this process: one dict per pass
    geometry with the launch shapes tried, their times (ms) and the best."""
    import json
    L = lib(field)
    n = ctypes.c_size_t()
    _check(L.qf_autotune_report(None, 0, ctypes.byref(n)), field)
    buf = ctypes.create_string_buffer(n.value + 1)
    _check(L.qf_autotune_report(buf, n.value + 1, ctypes.byref(n)), field)
    return json.loads(buf.value.decode())


class DeviceBuffer:
    """Device memory owned by the library (resident BVA)."""

    def __init__(self, nbytes, field=QF_FIELD_F65521):
        self.nbytes = int(nbytes)
        self.field = field
        self._L = lib(field)
        p = ctypes.c_void_p()
        _check(self._L.qf_device_alloc(self.nbytes, ctypes.byref(p)), field)
        self.ptr = p

    @classmethod
    def from_host(cls, array, field=QF_FIELD_F65521):
        import numpy as np
        arr = np.ascontiguousarray(array)
        buf = cls(arr.nbytes, field)
        _check(buf._L.qf_device_upload(buf.ptr, arr.ctypes.data, arr.nbytes), field)
        return buf

    def to_host(self, dtype="int8", offset=0, count=None):
        """Copy `count` items of `dtype` from byte `offset` (default: all) to host."""
        import numpy as np
        item = np.dtype(dtype).itemsize
        if count is None:
            count = (self.nbytes - offset) // item
        if offset < 0 or offset + count * item > self.nbytes:
            raise ValueError("range outside the device buffer")
        out = np.empty(count, dtype=dtype)
        _check(self._L.qf_device_download(out.ctypes.data, ctypes.c_void_p(self.ptr.value + offset), out.nbytes),
               self.field)
        return out

    def upload(self, array, offset=0):
        """Copy a host array into the buffer at byte `offset`."""
        import numpy as np
        arr = np.ascontiguousarray(array)
        if offset < 0 or offset + arr.nbytes > self.nbytes:
            raise ValueError("range outside the device buffer")
        _check(self._L.qf_device_upload(ctypes.c_void_p(self.ptr.value + offset), arr.ctypes.data, arr.nbytes),
               self.field)

    def copy_from(self, other, nbytes=None, half=None):
        """Device-to-device copy of `nbytes` (default: all of `other`); with
        half = 0 or 1 only that 64-byte half of every 128-byte line."""
        n = other.nbytes if nbytes is None else int(nbytes)
        if n > self.nbytes or n > other.nbytes:
            raise ValueError("copy larger than a buffer")
        pattern = 0 if half is None else 1 + int(half)
        _check(self._L.qf_device_copy(self.ptr, other.ptr, n, pattern), self.field)

    def close(self):
        if self.ptr:
            self._L.qf_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass(frozen=True)
class Shape:
    rows: int
    cols: int


@dataclass
class Config:
    """protocol.rs:18-32"""
    chunk_log: int = 30
    packed_log: int = 4
    query_rf_bits: int = 100


@dataclass
class Commitment:
    """Public statement (protocol.rs:33-40)."""
    version: int
    config: Config
    shapes: List[Shape]
    parameter_digest: bytes
    roots: List[bytes] = field(default_factory=list)


@dataclass
class Claim:
    """tbv.rs:6-11: evaluation claim of one matrix. Scalars are the 16-byte
    canonical form of extension elements (8 little-endian u16 coefficients
    for F65521, 4 little-endian u32 for KoalaBear)."""
    row: List[bytes]
    col: List[bytes]
    value: bytes


PROFILE_KEYS = ("conversion", "packed_eo", "pcs_projection", "terminal_batch_folding", "pcs_sumcheck",
                "terminal_batch_binding", "pcs_native_opening", "pcs_decide", "total")


class Committed:
    """Device-resident commitment state (protocol.rs:46-50)."""

    def __init__(self, params, handle, statement):
        self._params = params
        self._L = params._L
        self._h = ctypes.c_void_p(handle)
        self.statement = statement

    def codeword(self, block):
        """Column-major codeword of a block: (32 columns, R rows, coefficients)."""
        import numpy as np
        info = field_info(self._params.field)
        n_rows = 2 * (self._params.block_len // (32 * info["lanes"]))
        out = np.empty((32, n_rows, info["lanes"]), dtype=info["coef"])
        _check(self._L.qf_committed_codeword(self._h, block, out.ctypes.data, out.nbytes), self._params.field)
        return out

    def statement_json(self) -> str:
        """serde_json of the Commitment statement (commitment.json of the CLI)."""
        n = ctypes.c_size_t()
        self._L.qf_statement_json(self._params._h, self._h, None, 0, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value + 1)
        _check(self._L.qf_statement_json(self._params._h, self._h, buf, n.value + 1, ctypes.byref(n)),
               self._params.field)
        return buf.value.decode()

    def timing(self):
        """Device milliseconds of the commit (HIP events), summed over encoded blocks."""
        t, l, e, m = (ctypes.c_double() for _ in range(4))
        nb = ctypes.c_uint64()
        _check(self._L.qf_committed_timing(self._h, ctypes.byref(t), ctypes.byref(l), ctypes.byref(e),
                                           ctypes.byref(m), ctypes.byref(nb)), self._params.field)
        return {"total_ms": t.value, "leaf_ms": l.value, "encode_ms": e.value, "mw_ms": m.value,
                "timed_blocks": nb.value}

    def close(self):
        if self._h:
            self._L.qf_committed_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Parameters:
    """protocol.rs:41-45, 109-130. `field` selects the backend crate:
    F65521 (i8 entries) or KOALABEAR (i16 entries)."""

    def __init__(self, shapes: Sequence[Shape], config: Config = None, field: int = QF_FIELD_F65521):
        config = config or Config()
        self.field = field
        self._L = lib(field)
        self.entry_dtype = field_info(field)["entry"]
        flat = (ctypes.c_uint64 * (2 * len(shapes)))(*[v for s in shapes for v in (s.rows, s.cols)])
        h = ctypes.c_void_p()
        self._check(self._L.qf_params_new(field, flat, len(shapes), config.chunk_log, config.packed_log,
                                          config.query_rf_bits, ctypes.byref(h)))
        self._h = h
        self.config = config
        self.shapes = list(shapes)
        a, l, b, n, nb = (ctypes.c_uint64() for _ in range(5))
        self._check(self._L.qf_params_info(h, ctypes.byref(a), ctypes.byref(l), ctypes.byref(b), ctypes.byref(n),
                                           ctypes.byref(nb)))
        self.active, self.logical, self.block_len, self.blocks, self.n_buckets = a.value, l.value, b.value, n.value, nb.value

    new = classmethod(lambda cls, shapes, config=None, field=QF_FIELD_F65521: cls(shapes, config, field))

    def _check(self, status):
        _check(status, self.field)

    def parameter_digest(self) -> bytes:
        out = ctypes.create_string_buffer(32)
        self._check(self._L.qf_params_digest(self._h, out))
        return out.raw

    def buckets(self):
        out = []
        row = (ctypes.c_uint64 * 8)()
        for i in range(self.n_buckets):
            self._check(self._L.qf_params_bucket(self._h, i, row))
            out.append(tuple(row))
        return out

    def _statement(self, handle):
        buf = ctypes.create_string_buffer(32 * self.blocks)
        self._check(self._L.qf_committed_roots(handle, buf, self.blocks))
        roots = [buf.raw[32 * i:32 * i + 32] for i in range(self.blocks)]
        return Commitment(2, self.config, self.shapes, self.parameter_digest(), roots)

    def _commit(self, values, mode):
        import numpy as np
        h = ctypes.c_void_p()
        if isinstance(values, DeviceBuffer):  # BVA resident in HBM
            entries = values.nbytes // np.dtype(self.entry_dtype).itemsize
            self._check(self._L.qf_commit_device(self._h, values.ptr, entries, mode, ctypes.byref(h)))
        else:
            arr = np.ascontiguousarray(values, dtype=self.entry_dtype)
            self._check(self._L.qf_commit(self._h, arr.ctypes.data, arr.size, mode, ctypes.byref(h)))
        committed = Committed(self, h.value, self._statement(h))
        if isinstance(values, DeviceBuffer):
            committed._bva = values  # borrowed by the handle: the opening reads it again
        return committed

    def commit(self, values):
        """Parameters::commit (protocol.rs:131-134): codewords stay resident."""
        return self._commit(values, QF_CACHE_CACHED)

    def commit_recomputed(self, values):
        """Parameters::commit_recomputed (protocol.rs:161-185): roots only."""
        return self._commit(values, QF_CACHE_RECOMPUTE)

    def _claim_dims(self):
        return [((s.rows - 1).bit_length(), (s.cols - 1).bit_length()) for s in self.shapes]

    def claims_to_bytes(self, claims: Sequence[Claim]) -> bytes:
        parts = []
        for c in claims:
            parts.extend(c.row)
            parts.extend(c.col)
            parts.append(c.value)
        if any(len(v) != 16 for v in parts):
            raise ValueError("scalars are 16 bytes")
        return b"".join(bytes(v) for v in parts)

    def claims_from_bytes(self, raw: bytes) -> List[Claim]:
        out, o = [], 0
        for r, c in self._claim_dims():
            row = [raw[o + 16 * i:o + 16 * i + 16] for i in range(r)]
            o += 16 * r
            col = [raw[o + 16 * i:o + 16 * i + 16] for i in range(c)]
            o += 16 * c
            out.append(Claim(row, col, raw[o:o + 16]))
            o += 16
        return out

    def benchmark_claims(self, committed: "Committed") -> List[Claim]:
        """The CLI's benchmark statement (main.rs:188-209)."""
        n = ctypes.c_size_t()
        self._check(self._L.qf_claims_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(self._L.qf_benchmark_claims(self._h, committed._h, buf, n.value))
        return self.claims_from_bytes(buf.raw)

    def evaluate_claims(self, committed: "Committed", points) -> List[Claim]:
        """tbv.rs:143-185: claims at the given (row point, col point) pairs, one per matrix."""
        zero = bytes(16)
        raw = b"".join(b"".join(row) + b"".join(col) + zero for row, col in points)
        buf = ctypes.create_string_buffer(raw, len(raw))
        self._check(self._L.qf_evaluate_claims(self._h, committed._h, buf, len(raw)))
        return self.claims_from_bytes(buf.raw)

    def prove(self, committed: "Committed", claims: Sequence[Claim]):
        """Parameters::prove (protocol.rs:201-263): (proof bytes, profile in seconds)."""
        raw = self.claims_to_bytes(claims)
        h = ctypes.c_void_p()
        self._check(self._L.qf_prove(self._h, committed._h, raw, len(raw), ctypes.byref(h)))
        try:
            data = ctypes.POINTER(ctypes.c_uint8)()
            n = ctypes.c_size_t()
            self._check(self._L.qf_proof_bytes(h, ctypes.byref(data), ctypes.byref(n)))
            proof = ctypes.string_at(data, n.value)
            t = (ctypes.c_double * 9)()
            self._check(self._L.qf_proof_timing(h, t))
            profile = {k + "_seconds": t[i] / 1000.0 for i, k in enumerate(PROFILE_KEYS)}
        finally:
            self._L.qf_proof_free(h)
        return proof, profile

    def verify(self, statement: "Commitment", claims: Sequence[Claim], proof: bytes):
        """Parameters::verify (protocol.rs:266-319); raises QfError on rejection."""
        raw = self.claims_to_bytes(claims)
        roots = b"".join(statement.roots)
        self._check(self._L.qf_verify(self._h, roots, len(statement.roots), raw, len(raw), proof, len(proof)))
        return True

    def encode(self, matrices, out=None):
        """Layout::encode (layout.rs:115-135) on the device: the BVA of the
        row-major matrices (numpy arrays in host memory, or DeviceBuffers),
        as a device-resident DeviceBuffer (`out`, or a new one)."""
        import numpy as np
        item = np.dtype(self.entry_dtype).itemsize
        if out is None:
            out = DeviceBuffer(self.active * item, self.field)
        on_device = bool(matrices) and all(isinstance(m, DeviceBuffer) for m in matrices)
        keep = []
        ptrs = (ctypes.c_void_p * max(len(matrices), 1))()
        lens = (ctypes.c_uint64 * max(len(matrices), 1))()
        for i, m in enumerate(matrices):
            if on_device:
                ptrs[i], lens[i] = m.ptr.value, m.nbytes // item
            else:
                arr = np.ascontiguousarray(m, dtype=self.entry_dtype)
                keep.append(arr)
                ptrs[i], lens[i] = arr.ctypes.data, arr.size
        self._check(self._L.qf_encode(self._h, ptrs, lens, len(matrices), int(on_device), out.ptr))
        return out

    def encode_matrix(self, index, matrix, out):
        """One matrix of Layout::encode into the device BVA `out` (streaming input)."""
        import numpy as np
        if isinstance(matrix, DeviceBuffer):
            n = matrix.nbytes // np.dtype(self.entry_dtype).itemsize
            self._check(self._L.qf_encode_matrix(self._h, index, matrix.ptr, n, 1, out.ptr))
        else:
            arr = np.ascontiguousarray(matrix, dtype=self.entry_dtype)
            self._check(self._L.qf_encode_matrix(self._h, index, arr.ctypes.data, arr.size, 0, out.ptr))
        return out

    def verify_batch(self, statement: "Commitment", items, threads=0):
        """Batch verification of (claims, proof) pairs against one statement on
        the host's threads: the list of qf_verify status codes (0 = accepted)."""
        n = len(items)
        raw_claims = [self.claims_to_bytes(c) for c, _ in items]
        proofs = [bytes(pr) for _, pr in items]
        cl = (ctypes.c_char_p * max(n, 1))(*raw_claims)
        cll = (ctypes.c_size_t * max(n, 1))(*[len(x) for x in raw_claims])
        pp = (ctypes.c_char_p * max(n, 1))(*proofs)
        ppl = (ctypes.c_size_t * max(n, 1))(*[len(x) for x in proofs])
        st = (ctypes.c_int * max(n, 1))()
        roots = b"".join(statement.roots)
        self._L.qf_verify_batch(self._h, roots, len(statement.roots), cl, cll, pp, ppl, n, threads, st)
        return [st[i] for i in range(n)]

    def synthetic_bva(self, seed, sigma=None):
        """Device-resident synthetic BVA on physical entries: clip(round(N(0, sigma)))
        as int8 (F65521, sigma 20 by default) or the BF16 bit patterns of
        N(0, sigma) (KoalaBear, sigma 0.02 by default)."""
        import numpy as np
        if sigma is None:
            sigma = 0.02 if self.field == QF_FIELD_KOALABEAR else 20.0
        buf = DeviceBuffer(self.active * np.dtype(self.entry_dtype).itemsize, self.field)
        self._check(self._L.qf_synthetic_bva(self._h, seed, sigma, buf.ptr))
        return buf

    def close(self):
        if getattr(self, "_h", None):
            self._L.qf_params_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
