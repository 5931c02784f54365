"""The witness-free verifier (Parameters::verify, protocol.rs:266-319) on the
host: it accepts the oracle's proofs (which reproduce the reference's golden
digests) and rejects altered proofs, statements and claims, as the
reference's end-to-end tests do (qafold/f65521/tests/end_to_end.rs:92-167)."""
import hashlib
import json

import numpy as np
import pytest

from oracle_ref import DEMO_SHAPES, demo_bva, layout_encode, oracle_prove, shapes_flat, synthetic_int8


def setup(oracle, shapes, chunk_log, packed_log, values):
    import parcomm_amd as pa
    p = pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log, packed_log=packed_log))
    proof, statement, claims = oracle_prove(oracle, shapes_flat(shapes), len(shapes), chunk_log, packed_log, values)
    st = json.loads(statement)
    commitment = pa.Commitment(2, p.config, p.shapes, bytes(st["parameter_digest"]), [bytes(r) for r in st["roots"]])
    cl = [pa.Claim([bytes(v) for v in c["row"]], [bytes(v) for v in c["col"]], bytes(c["value"]))
          for c in json.loads(claims)]
    return p, commitment, cl, proof


def demo(oracle):
    import parcomm_amd as pa
    p = pa.Parameters([pa.Shape(r, c) for r, c in DEMO_SHAPES], pa.Config(chunk_log=10))
    values = demo_bva(lambda mats: layout_encode(p, mats))
    return setup(oracle, DEMO_SHAPES, 10, 4, values)


def random_case(oracle, shapes, chunk_log, packed_log, seed):
    import parcomm_amd as pa
    p = pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log, packed_log=packed_log))
    values = layout_encode(p, [synthetic_int8(r * c, seed + i) for i, (r, c) in enumerate(shapes)])
    return setup(oracle, shapes, chunk_log, packed_log, values)


def test_demo_proof_is_accepted(oracle):
    p, st, claims, proof = demo(oracle)
    assert hashlib.blake2s(proof).hexdigest() == "821aecd9a71dea52b8a29dc9ef8a810431a0605c05b3ec8f1783b47e15d1fb46"
    assert p.verify(st, claims, proof)


@pytest.mark.parametrize("chunk_log,shapes,packed_log,seed", [
    (12, [(64, 64)], 4, 1),              # single block, one QA level
    (14, [(100, 163)], 4, 2),            # single block, two levels (extension round + Decide)
    (16, [(256, 256)], 3, 3),
    (14, [(300, 200), (7, 1)], 2, 4),    # four blocks
    (16, [(513, 129), (33, 65)], 5, 5),  # two blocks, two levels
    (12, [(64, 64)], 1, 6),
    (18, [(1000, 300)], 6, 7),           # three levels
])
def test_random_proofs_are_accepted(oracle, chunk_log, shapes, packed_log, seed):
    p, st, claims, proof = random_case(oracle, shapes, chunk_log, packed_log, seed)
    assert p.verify(st, claims, proof)


def test_2048_square_chunk22_proof_is_accepted(oracle):
    p, st, claims, proof = random_case(oracle, [(2048, 2048)], 22, 4, 2048)
    assert p.verify(st, claims, proof)


def rejects(p, st, claims, proof):
    import parcomm_amd as pa
    with pytest.raises(pa.QfError):
        p.verify(st, claims, proof)
    return True


@pytest.mark.parametrize("case", ["demo", "levels"])
def test_altered_proof_bytes_are_rejected(oracle, case):
    p, st, claims, proof = demo(oracle) if case == "demo" else random_case(oracle, [(513, 129), (33, 65)], 16, 5, 5)
    rng = np.random.default_rng(11)
    positions = sorted(set([0, 8, 16 * 5, len(proof) // 3, len(proof) // 2, 2 * len(proof) // 3, len(proof) - 40,
                            len(proof) - 1] + list(rng.integers(0, len(proof), 24))))
    for pos in positions:
        bad = bytearray(proof)
        bad[pos] ^= 0x01
        assert rejects(p, st, claims, bytes(bad)), pos
    assert rejects(p, st, claims, proof[:-1])
    assert rejects(p, st, claims, proof + b"\x00")


def test_altered_statement_and_claims_are_rejected(oracle):
    import parcomm_amd as pa
    p, st, claims, proof = demo(oracle)
    roots = list(st.roots)
    roots[1] = bytes([roots[1][0] ^ 1]) + roots[1][1:]
    assert rejects(p, pa.Commitment(2, st.config, st.shapes, st.parameter_digest, roots), claims, proof)
    v = bytearray(claims[2].value)
    v[1] ^= 1
    bad = list(claims)
    bad[2] = pa.Claim(claims[2].row, claims[2].col, bytes(v))
    assert rejects(p, st, bad, proof)
    r = bytearray(claims[0].row[0])
    r[0] ^= 1
    bad = list(claims)
    bad[0] = pa.Claim([bytes(r)] + claims[0].row[1:], claims[0].col, claims[0].value)
    assert rejects(p, st, bad, proof)


def test_batch_verification_matches_single_proof_verification(oracle):
    # qf_verify_batch: each proof checked as by qf_verify, on the host's threads;
    # one altered byte and one altered claim are rejected in the mix
    import parcomm_amd as pa
    p, st, claims, proof = random_case(oracle, [(300, 200), (7, 1)], 14, 2, 4)
    bad_proof = bytearray(proof)
    bad_proof[len(proof) // 2] ^= 1
    v = bytearray(claims[0].value)
    v[0] ^= 1
    bad_claims = [pa.Claim(claims[0].row, claims[0].col, bytes(v))] + claims[1:]
    items = [(claims, proof), (claims, bytes(bad_proof)), (claims, proof), (bad_claims, proof), (claims, proof)]
    for threads in (1, 3, 0):
        statuses = p.verify_batch(st, items, threads=threads)
        assert statuses[0] == statuses[2] == statuses[4] == 0
        assert statuses[1] != 0 and statuses[3] != 0
    for (c, pr), s in zip(items, p.verify_batch(st, items)):
        if s == 0:
            assert p.verify(st, c, pr)
        else:
            with pytest.raises(pa.QfError):
                p.verify(st, c, pr)
