"""End-to-end golden: the oracle reproduces the reference CLI demo proof
(`parcomm-qafold demo --chunk-log 10 --packed-log 4`, recompute mode) byte for
byte. Golden values from validation/qafold8-smoke-final/run.log: proof 26 974
bytes, BLAKE2s 821aecd9...fb46, public commitment JSON 782 bytes."""
import ctypes
import hashlib

import numpy as np

from oracle_ref import DEMO_SHAPES, oracle_prove

DEMO_PROOF_BLAKE2S = "821aecd9a71dea52b8a29dc9ef8a810431a0605c05b3ec8f1783b47e15d1fb46"


def oracle_layout(oracle, shapes):
    flat = (ctypes.c_uint64 * (2 * len(shapes)))(*[v for s in shapes for v in s])
    cap = 4096
    b = (ctypes.c_uint64 * (8 * cap))()
    nb, act, lg = ctypes.c_size_t(), ctypes.c_uint64(), ctypes.c_uint64()
    assert oracle.orc_layout(flat, ctypes.c_size_t(len(shapes)), b, ctypes.c_size_t(cap), ctypes.byref(nb),
                             ctypes.byref(act), ctypes.byref(lg)) == 0
    return flat, [tuple(b[8 * i + j] for j in range(8)) for i in range(nb.value)], act.value


def oracle_encode(shapes, buckets, active, mats):
    out = np.zeros(active, dtype=np.int8)
    for (m, ro, co, rows, cols, pr, pc, off) in buckets:
        src = mats[m].reshape(shapes[m])
        out[off:off + rows * cols].reshape(rows, cols)[:pr, :pc] = src[ro:ro + pr, co:co + pc]
    return out


def demo_values(oracle):
    flat, buckets, active = oracle_layout(oracle, DEMO_SHAPES)
    mats = []
    for i, (r, c) in enumerate(DEMO_SHAPES):
        e = np.arange(r * c, dtype=np.int64)
        mats.append(((e * 73 + i * 19) % 256).astype(np.uint8).view(np.int8))
    return flat, oracle_encode(DEMO_SHAPES, buckets, active, mats)


def test_demo_proof_matches_reference_digest(oracle):
    flat, values = demo_values(oracle)
    proof, statement, _ = oracle_prove(oracle, flat, len(DEMO_SHAPES), 10, 4, values)
    assert len(statement) == 782
    assert len(proof) == 26974
    assert hashlib.blake2s(proof).hexdigest() == DEMO_PROOF_BLAKE2S


def test_tampered_witness_changes_the_proof(oracle):
    # A plausible mistake planted on the CPU (one witness byte) must change the digest.
    flat, values = demo_values(oracle)
    values = values.copy()
    values[5] ^= 1
    proof, _, _ = oracle_prove(oracle, flat, len(DEMO_SHAPES), 10, 4, values)
    assert hashlib.blake2s(proof).hexdigest() != DEMO_PROOF_BLAKE2S


NATIVE_FIXTURE_BLAKE2S = "bb4993475432e1fb1adb6e0181f80a48d625069a88182a6c721715b687f6d177"


def test_native_opening_and_decide_fixture(oracle):
    # qafold_short_pc.rs:3605-3647: ShortQaFoldPC<F>::new(2^16, 16, 2, 80),
    # StdRng::seed_from_u64(0x7061_7261_6c6c_656c); BLAKE2s(root || opening || decide).
    d = (ctypes.c_uint8 * 32)()
    ol, dl = ctypes.c_uint64(), ctypes.c_uint64()
    rc = oracle.orc_native_fixture(ctypes.c_uint64(1 << 16), ctypes.c_uint64(16), ctypes.c_uint64(80),
                                   ctypes.c_uint64(0x7061_7261_6c6c_656c), d, ctypes.byref(ol), ctypes.byref(dl))
    assert rc == 0, oracle.orc_last_error()
    assert bytes(d).hex() == NATIVE_FIXTURE_BLAKE2S


def test_multilevel_proof_runs_with_prover_self_checks(oracle):
    # 2^14-entry blocks give three QA levels and two Decide rounds; every
    # algebraic terminal check inside the prover must pass.
    shapes = [(300, 200), (129, 65)]
    flat, buckets, active = oracle_layout(oracle, shapes)
    rng = np.random.default_rng(3)
    mats = [np.clip(np.rint(rng.normal(0, 20, r * c)), -128, 127).astype(np.int8) for r, c in shapes]
    values = oracle_encode(shapes, buckets, active, mats)
    proof, _, _ = oracle_prove(oracle, flat, len(shapes), 14, 4, values)
    assert len(proof) > 10000
