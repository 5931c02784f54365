"""Pins the KoalaBear / Ext4 build of the oracle (liboracle_kb.so, i16
entries): the quartic field against a pure-Python restatement of
KoalaBear[X]/(X^4 - 3) (koalabear.rs:1-2, :305-340 test), the Frobenius
against x^p by square-and-multiply and against the reference's phase table
(koalabear.rs:155), WHT / Mobius / Merkle closed forms at p = 2^31 - 2^24 + 1,
and the end-to-end demo golden: `parcomm-qafold demo --chunk-log 10
--packed-log 4` of the KoalaBear crate, proof 51 646 bytes, BLAKE2s
2a4a9dd7...0285, public commitment JSON 776 bytes
(validation/qafold16-smoke/run.log:7)."""
import ctypes
import hashlib
import random
import struct

import numpy as np

from oracle_ref import DEMO_SHAPES, oracle_prove, shapes_flat

P = 2130706433
KB_DEMO_PROOF_BLAKE2S = "2a4a9dd7db840a62c12e55c65ac700f82a7150f992886bad420dcab4d2fe0285"


def py_mul(a, b):
    t = [0] * 7
    for i in range(4):
        for j in range(4):
            t[i + j] += a[i] * b[j]
    return [(t[i] + 3 * (t[i + 4] if i < 3 else 0)) % P for i in range(4)]


def py_pow(a, e):
    acc, base = [1, 0, 0, 0], a
    while e:
        if e & 1:
            acc = py_mul(acc, base)
        base = py_mul(base, base)
        e >>= 1
    return acc


def e4(oracle, op, a, b=(0,) * 4):
    A = (ctypes.c_uint32 * 4)(*a)
    B = (ctypes.c_uint32 * 4)(*b)
    O = (ctypes.c_uint32 * 4)()
    assert oracle.orc_e8_op(op, A, B, O) == 0
    return list(O)


def test_kb_build_has_four_lanes(oracle_kb, oracle):
    assert oracle_kb.orc_lanes() == 4
    assert oracle.orc_lanes() == 8


def test_ext4_field_arithmetic(oracle_kb):
    rng = random.Random(11)
    one = [1, 0, 0, 0]
    assert py_pow([0, 1, 0, 0], 4) == [3, 0, 0, 0]
    for _ in range(200):
        a = [rng.randrange(P) for _ in range(4)]
        b = [rng.randrange(P) for _ in range(4)]
        assert e4(oracle_kb, 0, a, b) == [(x + y) % P for x, y in zip(a, b)]
        assert e4(oracle_kb, 1, a, b) == [(x - y) % P for x, y in zip(a, b)]
        assert e4(oracle_kb, 2, a, b) == py_mul(a, b)
        assert py_mul(a, e4(oracle_kb, 3, a)) == one
    for edge in ([P - 1] * 4, [0, 0, 0, P - 1], [1, P - 1, 1, P - 1]):
        assert e4(oracle_kb, 2, edge, edge) == py_mul(edge, edge)
    for _ in range(3):
        a = [rng.randrange(P) for _ in range(4)]
        assert e4(oracle_kb, 4, a, [1, 0, 0, 0]) == py_pow(a, P)
        assert e4(oracle_kb, 4, a, [2, 0, 0, 0]) == py_pow(py_pow(a, P), P)
        assert e4(oracle_kb, 4, a, [4, 0, 0, 0]) == a


def test_frobenius_phases_match_reference_table(oracle_kb):
    # frobenius(k) scales coefficient 1 by PHASES[k] (koalabear.rs:155).
    phases = [1, 2113994754, 2130706432, 16711679]
    for k, ph in enumerate(phases):
        assert e4(oracle_kb, 4, [0, 1, 0, 0], [k, 0, 0, 0]) == [0, ph, 0, 0]
        assert py_pow([0, 1, 0, 0], P ** k) == [0, ph, 0, 0]


def rand_e4(rng, n):
    return np.array([[rng.randrange(P) for _ in range(4)] for _ in range(n)], dtype=np.uint32)


def test_wht_and_mobius_closed_forms(oracle_kb):
    rng = random.Random(13)
    for logn in [1, 3, 6]:
        n = 1 << logn
        v = rand_e4(rng, n)
        w = v.copy()
        oracle_kb.orc_wht(w.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n))
        H = np.array([[(-1) ** bin(i & j).count("1") for j in range(n)] for i in range(n)], dtype=object)
        assert np.array_equal(w.astype(object), (H.dot(v.astype(object))) % P)
        oracle_kb.orc_wht(w.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n))
        assert np.array_equal(w.astype(object), (n * v.astype(object)) % P)
    n = 16
    v = rand_e4(rng, n)
    w = v.copy()
    oracle_kb.orc_mobius(w.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n))
    for i in range(n):
        acc = np.zeros(4, dtype=object)
        for j in range(n):
            if j & ~i == 0:
                acc += (-1) ** (bin(i).count("1") - bin(j).count("1")) * v[j].astype(object)
        assert np.array_equal(w[i].astype(object), acc % P)


def frame(label, data):
    return struct.pack("<Q", len(label)) + label + struct.pack("<Q", len(data)) + data


def test_merkle_rows_use_u32_canonical_bytes(oracle_kb):
    rng = random.Random(17)
    rows, cols = 8, 32
    e = rand_e4(rng, rows * cols)
    out = (ctypes.c_uint8 * 32)()
    oracle_kb.orc_merkle_root_rows(e.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(rows), ctypes.c_size_t(cols), out)
    leaves = []
    for r in range(rows):
        msg = frame(b"domain", b"qafold-short/base-row/v1") + frame(b"width", struct.pack("<Q", cols))
        msg += b"".join(e[k * rows + r].astype("<u4").tobytes() for k in range(cols))
        leaves.append(hashlib.blake2s(msg).digest())
    while len(leaves) > 1:
        leaves = [hashlib.blake2s(frame(b"domain", b"qafold-short/merkle-inner/v1") + leaves[2 * i] + leaves[2 * i + 1]).digest()
                  for i in range(len(leaves) // 2)]
    assert bytes(out) == leaves[0]


def kb_demo_values(oracle_kb):
    """Demo matrices of the KoalaBear CLI: entry e of matrix i =
    (e*73 + i*19) as u16 as i16 (qafold/koalabear/src/main.rs:137-148),
    scattered by the oracle's own layout."""
    flat = shapes_flat(DEMO_SHAPES)
    cap = 4096
    b = (ctypes.c_uint64 * (8 * cap))()
    nb, act, lg = ctypes.c_size_t(), ctypes.c_uint64(), ctypes.c_uint64()
    assert oracle_kb.orc_layout(flat, ctypes.c_size_t(len(DEMO_SHAPES)), b, ctypes.c_size_t(cap), ctypes.byref(nb),
                                ctypes.byref(act), ctypes.byref(lg)) == 0
    out = np.zeros(act.value, dtype=np.int16)
    for k in range(nb.value):
        m, ro, co, rows, cols, pr, pc, off = b[8 * k:8 * k + 8]
        e = np.arange(DEMO_SHAPES[m][0] * DEMO_SHAPES[m][1], dtype=np.int64)
        src = ((e * 73 + m * 19) % 65536).astype(np.uint16).view(np.int16).reshape(DEMO_SHAPES[m])
        out[off:off + rows * cols].reshape(rows, cols)[:pr, :pc] = src[ro:ro + pr, co:co + pc]
    return flat, out


def test_kb_demo_proof_matches_reference_digest(oracle_kb):
    flat, values = kb_demo_values(oracle_kb)
    assert len(values) == 2305
    proof, statement, _ = oracle_prove(oracle_kb, flat, len(DEMO_SHAPES), 10, 4, values)
    assert len(statement) == 776
    assert len(proof) == 51646
    assert hashlib.blake2s(proof).hexdigest() == KB_DEMO_PROOF_BLAKE2S


def test_kb_tampered_witness_changes_the_proof(oracle_kb):
    flat, values = kb_demo_values(oracle_kb)
    values = values.copy()
    values[7] += 1
    proof, _, _ = oracle_prove(oracle_kb, flat, len(DEMO_SHAPES), 10, 4, values)
    assert hashlib.blake2s(proof).hexdigest() != KB_DEMO_PROOF_BLAKE2S
