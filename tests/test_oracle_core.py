"""Pins the oracle's arithmetic and hashing core against independent
references: hashlib.blake2s, the RFC 8439 ChaCha20 block vector, and
pure-Python restatements of Ext8 arithmetic, WHT, Mobius and the framed
Merkle tree on small inputs."""
import ctypes
import hashlib
import random
import struct

import numpy as np

P = 65521


def test_blake2s_matches_hashlib(oracle):
    rng = random.Random(1)
    out = (ctypes.c_uint8 * 32)()
    for n in [0, 1, 55, 63, 64, 65, 114, 128, 587, 1000, 4097]:
        data = bytes(rng.getrandbits(8) for _ in range(n))
        oracle.orc_blake2s(data, ctypes.c_size_t(n), out)
        assert bytes(out) == hashlib.blake2s(data).digest(), n


def test_chacha20_rfc8439_block(oracle):
    # RFC 8439 section 2.3.2: key 00..1f, nonce 00:00:00:09:00:00:00:4a:00:00:00:00, counter 1.
    # rand_chacha keeps a 64-bit counter in words 12-13 and a 64-bit stream id in 14-15.
    w = (ctypes.c_uint32 * 16)()
    oracle.orc_chacha_block(bytes(range(32)), ctypes.c_uint64(1 | (0x09000000 << 32)),
                            ctypes.c_uint64(0x4A000000), 20, w)
    expected = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204,
                0x4E6CD4C3, 0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE,
                0xE883D0CB, 0x4E3C50A2]
    assert list(w) == expected


def py_mul(a, b):
    t = [0] * 15
    for i in range(8):
        for j in range(8):
            t[i + j] += a[i] * b[j]
    return [(t[i] + 17 * (t[i + 8] if i < 7 else 0)) % P for i in range(8)]


def e8(oracle, op, a, b=(0,) * 8):
    A = (ctypes.c_uint16 * 8)(*a)
    B = (ctypes.c_uint16 * 8)(*b)
    O = (ctypes.c_uint16 * 8)()
    assert oracle.orc_e8_op(op, A, B, O) == 0
    return list(O)


def test_ext8_field_arithmetic(oracle):
    rng = random.Random(7)
    one = [1] + [0] * 7
    for _ in range(200):
        a = [rng.randrange(P) for _ in range(8)]
        b = [rng.randrange(P) for _ in range(8)]
        assert e8(oracle, 0, a, b) == [(x + y) % P for x, y in zip(a, b)]
        assert e8(oracle, 1, a, b) == [(x - y) % P for x, y in zip(a, b)]
        assert e8(oracle, 2, a, b) == py_mul(a, b)
        inv = e8(oracle, 3, a)
        assert py_mul(a, inv) == one
    # X^8 = 17 and the Frobenius x -> x^p for a few elements (square and multiply in Python).
    x = [0, 1, 0, 0, 0, 0, 0, 0]
    r = one
    for _ in range(8):
        r = py_mul(r, x)
    assert r == [17] + [0] * 7
    for _ in range(3):
        a = [rng.randrange(P) for _ in range(8)]
        acc, base, e = one, a, P
        while e:
            if e & 1:
                acc = py_mul(acc, base)
            base = py_mul(base, base)
            e >>= 1
        assert e8(oracle, 4, a, [1] + [0] * 7) == acc


def rand_e8(rng, n):
    return np.array([[rng.randrange(P) for _ in range(8)] for _ in range(n)], dtype=np.uint16)


def test_wht_closed_forms(oracle):
    rng = random.Random(3)
    for logn in [1, 3, 6]:
        n = 1 << logn
        v = rand_e8(rng, n)
        w = v.copy()
        oracle.orc_wht(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), ctypes.c_size_t(n))
        H = np.array([[(-1) ** bin(i & j).count("1") for j in range(n)] for i in range(n)], dtype=np.int64)
        assert np.array_equal(w.astype(np.int64), (H @ v.astype(np.int64)) % P)
        oracle.orc_wht(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), ctypes.c_size_t(n))
        assert np.array_equal(w.astype(np.int64), (n * v.astype(np.int64)) % P)


def test_mobius_is_subset_sum_inverse(oracle):
    rng = random.Random(4)
    n = 32
    v = rand_e8(rng, n)
    w = v.copy()
    oracle.orc_mobius(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), ctypes.c_size_t(n))
    for i in range(n):
        acc = np.zeros(8, dtype=np.int64)
        for j in range(n):
            if j & ~i == 0:
                acc += (-1) ** (bin(i).count("1") - bin(j).count("1")) * v[j].astype(np.int64)
        assert np.array_equal(w[i].astype(np.int64), acc % P)


def frame(label, data):
    return struct.pack("<Q", len(label)) + label + struct.pack("<Q", len(data)) + data


def py_merkle_root(entries, rows, cols):
    leaves = []
    for r in range(rows):
        msg = frame(b"domain", b"qafold-short/base-row/v1") + frame(b"width", struct.pack("<Q", cols))
        for k in range(cols):
            msg += entries[k * rows + r].tobytes()
        leaves.append(hashlib.blake2s(msg).digest())
    while len(leaves) > 1:
        leaves = [hashlib.blake2s(frame(b"domain", b"qafold-short/merkle-inner/v1") + leaves[2 * i] + leaves[2 * i + 1]).digest()
                  for i in range(len(leaves) // 2)]
    return leaves[0]


def test_merkle_rows_match_framed_hashlib(oracle):
    rng = random.Random(5)
    for rows, cols in [(1, 1), (2, 3), (8, 32), (16, 4)]:
        e = rand_e8(rng, rows * cols)
        out = (ctypes.c_uint8 * 32)()
        oracle.orc_merkle_root_rows(e.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), ctypes.c_size_t(rows),
                                    ctypes.c_size_t(cols), out)
        assert bytes(out) == py_merkle_root(e, rows, cols)
