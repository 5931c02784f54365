"""Opening parity: the HIP prover's proof bytes against the CPU oracle and the
reference's own golden digest (protocol.rs:201-263, eo.rs:97-181,
packed.rs:294-393, packed_batch.rs, qafold_short_pc.rs:578-725,
qafold_short_decide_driver.rs:51-318). Proofs are deterministic functions of
the statement and witness, so parity is byte equality of the whole proof."""
import hashlib
import json

import pytest

from oracle_ref import (DEMO_SHAPES, demo_bva, layout_encode, oracle_prove, shapes_flat,
                        synthetic_int8)

pytestmark = pytest.mark.gpu

# Reference CLI demo (chunk_log 10, packed_log 4), validation/qafold8-smoke-final/run.log
DEMO_PROOF_BLAKE2S = "821aecd9a71dea52b8a29dc9ef8a810431a0605c05b3ec8f1783b47e15d1fb46"


def params(shapes, chunk_log, packed_log=4):
    import parcomm_amd as pa
    return pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log, packed_log=packed_log))


def claims_from_json(text):
    return [(tuple(bytes(v) for v in c["row"]), tuple(bytes(v) for v in c["col"]), bytes(c["value"]))
            for c in json.loads(text)]


def check_against_oracle(oracle, shapes, chunk_log, values, packed_log=4, recompute=False):
    p = params(shapes, chunk_log, packed_log)
    committed = p.commit_recomputed(values) if recompute else p.commit(values)
    claims = p.benchmark_claims(committed)
    proof, profile = p.prove(committed, claims)
    want, statement, claims_json = oracle_prove(oracle, shapes_flat(shapes), len(shapes), chunk_log, packed_log,
                                                values)
    assert committed.statement_json().encode() == statement
    assert [(tuple(c.row), tuple(c.col), c.value) for c in claims] == claims_from_json(claims_json)
    assert len(proof) == len(want)
    assert proof == want
    assert profile["total_seconds"] > 0
    assert p.verify(committed.statement, claims, proof)
    return p, committed, claims, proof


@pytest.mark.parametrize("recompute", [True, False])
def test_demo_proof_matches_reference_digest(oracle, recompute):
    p = params(DEMO_SHAPES, 10)
    values = demo_bva(lambda mats: layout_encode(p, mats))
    print("demo proof, recompute", recompute, flush=True)
    _, _, _, proof = check_against_oracle(oracle, DEMO_SHAPES, 10, values, recompute=recompute)
    assert len(proof) == 26974
    assert hashlib.blake2s(proof).hexdigest() == DEMO_PROOF_BLAKE2S


@pytest.mark.parametrize("chunk_log,shapes,seed,packed_log", [
    (12, [(64, 64)], 1, 4),              # one PCS block: the single-block opening
    (14, [(100, 163)], 2, 4),            # two PCS levels
    (16, [(256, 256)], 3, 3),            # packed_log 3
    (14, [(300, 200), (7, 1)], 4, 2),    # several blocks, packed_log 2
    (16, [(513, 129), (33, 65)], 5, 5),  # ragged shapes, packed_log 5
    (12, [(64, 64)], 6, 1),              # packed_log 1 (two lanes)
    (14, [(300, 200)], 7, 6),            # packed_log 6
    (10, [(150, 150)], 8, 4),            # 22 blocks: block sums folded every 16 blocks
    (20, [(1000, 3000), (3, 5)], 9, 4),  # 16-lane terms of 2^10..2^17 rows on the MFMA projection, ragged end
    (18, [(500, 1000)], 10, 4),          # column-padding terms sharing query rows (added in separate launches)
])
def test_random_proofs_match_oracle(oracle, chunk_log, shapes, seed, packed_log):
    p = params(shapes, chunk_log, packed_log)
    mats = [synthetic_int8(r * c, seed + i) for i, (r, c) in enumerate(shapes)]
    values = layout_encode(p, mats)
    print("random proof", chunk_log, shapes, seed, packed_log, "blocks", p.blocks, flush=True)
    check_against_oracle(oracle, shapes, chunk_log, values, packed_log)


def test_2048_square_chunk22_proof_matches_oracle(oracle):
    shapes = [(2048, 2048)]
    p = params(shapes, 22)
    values = layout_encode(p, [synthetic_int8(2048 * 2048, 2048)])
    print("2048x2048 chunk 22 proof", flush=True)
    check_against_oracle(oracle, shapes, 22, values)


def test_false_claim_is_rejected():
    import parcomm_amd as pa
    p = params(DEMO_SHAPES, 10)
    committed = p.commit(demo_bva(lambda mats: layout_encode(p, mats)))
    claims = p.benchmark_claims(committed)
    v = bytearray(claims[1].value)
    v[0] = (v[0] + 1) % 256 if v[0] != 0xf0 else 0
    claims[1] = pa.Claim(claims[1].row, claims[1].col, bytes(v))
    with pytest.raises(pa.QfError) as e:
        p.prove(committed, claims)
    assert "packed reduction claim mismatch" in str(e.value)  # eo.rs:135-138


def test_malformed_claims_are_rejected():
    import parcomm_amd as pa
    p = params(DEMO_SHAPES, 10)
    committed = p.commit(demo_bva(lambda mats: layout_encode(p, mats)))
    claims = p.benchmark_claims(committed)
    with pytest.raises(pa.QfError) as e:  # missing a row coordinate
        p.prove(committed, [pa.Claim(claims[0].row[1:], claims[0].col, claims[0].value)] + claims[1:])
    assert e.value.status == 5  # InvalidParameters
    bad = b"\xff\xff" + claims[2].value[2:]  # coefficient 65535 >= p
    with pytest.raises(pa.QfError) as e:
        p.prove(committed, claims[:2] + [pa.Claim(claims[2].row, claims[2].col, bad)] + claims[3:])
    assert e.value.status == 1  # EncodingError (tbv.rs:131-138)


def test_per_matrix_launch_form_gives_identical_proofs(oracle):
    import parcomm_amd as pa
    shapes = [(300, 200), (7, 1), (64, 33)]
    p = params(shapes, 14, 4)
    values = layout_encode(p, [synthetic_int8(r * c, 30 + i) for i, (r, c) in enumerate(shapes)])
    committed = p.commit(values)
    claims = p.benchmark_claims(committed)
    batched, _ = p.prove(committed, claims)
    pa.set_launch_mode(pa.LAUNCH_PER_MATRIX)
    try:
        per_matrix, _ = p.prove(committed, claims)
    finally:
        pa.set_launch_mode(pa.LAUNCH_BATCHED)
    assert per_matrix == batched


def test_served_requests_match_their_proofs_alone():
    # the serving loop runs requests concurrently on separate streams; each
    # result must equal the proof of the same request computed alone
    import numpy as np
    from parcomm_amd.serve import request_claims, serve
    shapes = [(513, 129), (33, 65), (100, 7)]
    p = params(shapes, 16, 4)
    values = layout_encode(p, [synthetic_int8(r * c, 50 + i) for i, (r, c) in enumerate(shapes)])
    committed = p.commit(values)
    base = p.benchmark_claims(committed)
    print("serving 6 requests over 3 workers", flush=True)
    stats, reqs = serve(p, committed, base, n_requests=6, rate=500.0, seed=7, workers=3, verify=True)
    assert stats["concurrency"] >= 1 and stats["p99_ms"] >= stats["p50_ms"]
    for r in reqs:
        claims = request_claims(p, committed, base, r.subset, np.random.default_rng(7 * 1_000_003 + r.rid))
        alone, _ = p.prove(committed, claims)
        assert [c.value for c in claims] == [c.value for c in r.claims]
        assert alone == r.proof, r.rid
