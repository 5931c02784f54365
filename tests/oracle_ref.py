"""Helpers that run the CPU oracle (the checker) on the same inputs as the
product, for the parity tests."""
import ctypes

import numpy as np

DEMO_SHAPES = [(31, 37), (65, 17), (1, 7), (9, 1)]


def demo_bva(layout_encode, kb=False):
    """Demo matrices of the reference CLI: entry e of matrix i = (e*73 + i*19)
    as u8 as i8 (qafold/f65521/src/main.rs:133-144), or as u16 as i16 for the
    KoalaBear crate (qafold/koalabear/src/main.rs:137-148)."""
    mats = []
    for i, (r, c) in enumerate(DEMO_SHAPES):
        e = np.arange(r * c, dtype=np.int64)
        if kb:
            mats.append(((e * 73 + i * 19) % 65536).astype(np.uint16).view(np.int16))
        else:
            mats.append(((e * 73 + i * 19) % 256).astype(np.uint8).view(np.int8))
    return layout_encode(mats)


def layout_encode(params, mats):
    """layout.rs:115-135 on the host from the product's bucket table."""
    out = np.zeros(params.active, dtype=getattr(params, "entry_dtype", "int8"))
    for (m, ro, co, rows, cols, pr, pc, off) in params.buckets():
        src = mats[m].reshape(params.shapes[m].rows, params.shapes[m].cols)
        block = src[ro:ro + pr, co:co + pc]
        view = out[off:off + rows * cols].reshape(rows, cols)
        view[:pr, :pc] = block
    return out


def oracle_commit(oracle, chunk_log, values, blocks, want_codeword=False):
    """Per-block roots (and codewords) of the oracle; the entry type of `values`
    selects the field (int8: F65521 oracle, int16: KoalaBear oracle)."""
    kb = values.dtype == np.int16
    lanes, coef = (4, np.uint32) if kb else (8, np.uint16)
    h = oracle.orc_pcs_new(chunk_log, 100, bytes([0x51] * 32))
    assert h, oracle.orc_last_error()
    try:
        block_len = 1 << chunk_log
        roots, cws = [], []
        for b in range(blocks):
            first = b * block_len
            avail = max(0, min(block_len, len(values) - first))
            chunk = np.zeros(block_len, dtype=values.dtype)
            chunk[:avail] = values[first:first + avail]
            root = (ctypes.c_uint8 * 32)()
            cw = None
            if want_codeword:
                cw = np.empty((32, 2 * (block_len // (32 * lanes)), lanes), dtype=coef)
            rc = oracle.orc_commit_block(ctypes.c_void_p(h), chunk.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_uint64(block_len), root, None,
                                         cw.ctypes.data_as(ctypes.c_void_p) if cw is not None else None)
            assert rc == 0, oracle.orc_last_error()
            roots.append(bytes(root))
            cws.append(cw)
        return roots, cws
    finally:
        oracle.orc_pcs_free(h)


def synthetic_int8(shape_count, seed, sigma=20.0):
    """round(N(0, sigma)) clipped to [-128, 127], seeded."""
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(rng.normal(0.0, sigma, size=shape_count)), -128, 127).astype(np.int8)


def synthetic_bf16(shape_count, seed, sigma=0.02):
    """Raw bfloat16 bit patterns (as int16) of N(0, sigma) weights, seeded;
    round to nearest even from float32."""
    rng = np.random.default_rng(seed)
    b = rng.normal(0.0, sigma, size=shape_count).astype(np.float32).view(np.uint32).astype(np.uint64)
    r = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return r.astype(np.uint16).view(np.int16)


def synthetic_int16(shape_count, seed):
    """Uniform int16 entries over the full range (extreme values included), seeded."""
    rng = np.random.default_rng(seed)
    v = rng.integers(-32768, 32768, size=shape_count, dtype=np.int64)
    v[: min(4, shape_count)] = [-32768, 32767, -1, 0][: min(4, shape_count)]
    return v.astype(np.int16)


def shapes_flat(shapes):
    return (ctypes.c_uint64 * (2 * len(shapes)))(*[v for s in shapes for v in s])


def oracle_prove(oracle, flat, n_shapes, chunk_log, packed_log, values):
    """The reference CLI flow (main.rs:188-209) in the oracle: benchmark claims,
    Parameters::prove, Proof::to_bytes. Returns (proof, statement JSON, claims JSON)."""
    cap = 1 << 26
    proof = (ctypes.c_uint8 * cap)()
    plen = ctypes.c_size_t()
    sj = ctypes.create_string_buffer(1 << 22)
    cj = ctypes.create_string_buffer(1 << 22)
    rc = oracle.orc_prove_benchmark(flat, ctypes.c_size_t(n_shapes), chunk_log, packed_log,
                                    values.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(len(values)), proof,
                                    ctypes.c_size_t(cap), ctypes.byref(plen), sj, ctypes.c_size_t(1 << 22), cj,
                                    ctypes.c_size_t(1 << 22))
    assert rc == 0, oracle.orc_last_error()
    return bytes(proof[:plen.value]), sj.value, cj.value


def points_bytes(points):
    """Packed point list: per matrix its row coordinates, then its column
    coordinates, 16 canonical bytes each."""
    return b"".join(bytes(v) for row, col in points for v in list(row) + list(col))


def oracle_prove_points(oracle, flat, n_shapes, chunk_log, packed_log, values, points):
    """As oracle_prove, with the claims evaluated at the given (row, col) points."""
    cap = 1 << 26
    proof = (ctypes.c_uint8 * cap)()
    plen = ctypes.c_size_t()
    sj = ctypes.create_string_buffer(1 << 22)
    cj = ctypes.create_string_buffer(1 << 22)
    pts = points_bytes(points)
    rc = oracle.orc_prove_points(flat, ctypes.c_size_t(n_shapes), chunk_log, packed_log,
                                 values.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(len(values)), pts, proof,
                                 ctypes.c_size_t(cap), ctypes.byref(plen), sj, ctypes.c_size_t(1 << 22), cj,
                                 ctypes.c_size_t(1 << 22))
    assert rc == 0, oracle.orc_last_error()
    return bytes(proof[:plen.value]), sj.value, cj.value


# Magic tags that open the regions of a ParComm proof's wire
# (packed_batch.rs:285-305, qafold_short_batch.rs:86-87,
# qafold_short_pc.rs:240-261, qafold_short_decide_wire.rs:97-170).
REGION_TAGS = {"batch_opening": b"QAPKBAT1", "binding": b"QAFOLDB1", "native_opening": b"QFSHORT2",
               "decide": b"QFSDCID1"}


def proof_regions(proof):
    """Byte offset of one payload byte inside each region of a batched proof:
    the packed EO proof at the start (its first cross term), then the first
    payload byte past each region's tag and the 24 bytes that follow it."""
    out = {"eo": 8 + 3}
    for name, tag in REGION_TAGS.items():
        at = proof.find(tag)
        assert at >= 0, name
        out[name] = at + len(tag) + 24
    return out
