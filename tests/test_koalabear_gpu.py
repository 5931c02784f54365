"""KoalaBear / Ext4 backend (libparcomm_amd_kb.so, i16 entries) against the
KoalaBear build of the CPU oracle: bit-exact roots and codewords, and proof
bytes equal to the oracle's and to the reference's KoalaBear demo digest
(qafold/koalabear: protocol.rs, packed_koalabear.rs, vendor/qafold
packed_batch.rs). Every proof is also checked by the KoalaBear verifier."""
import hashlib
import json

import numpy as np
import pytest

from oracle_ref import (DEMO_SHAPES, demo_bva, layout_encode, oracle_commit, oracle_prove, shapes_flat,
                        synthetic_bf16, synthetic_int16)

pytestmark = pytest.mark.gpu

# Reference KoalaBear CLI demo (chunk_log 10, packed_log 4), validation/qafold16-smoke/run.log:7
KB_DEMO_PROOF_BLAKE2S = "2a4a9dd7db840a62c12e55c65ac700f82a7150f992886bad420dcab4d2fe0285"


def params(shapes, chunk_log, packed_log=4):
    import parcomm_amd as pa
    return pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log, packed_log=packed_log),
                         field=pa.KOALABEAR)


def claims_from_json(text):
    return [(tuple(bytes(v) for v in c["row"]), tuple(bytes(v) for v in c["col"]), bytes(c["value"]))
            for c in json.loads(text)]


def check_against_oracle(oracle_kb, shapes, chunk_log, values, packed_log=4, recompute=False):
    p = params(shapes, chunk_log, packed_log)
    committed = p.commit_recomputed(values) if recompute else p.commit(values)
    claims = p.benchmark_claims(committed)
    proof, profile = p.prove(committed, claims)
    want, statement, claims_json = oracle_prove(oracle_kb, shapes_flat(shapes), len(shapes), chunk_log, packed_log,
                                                values)
    assert committed.statement_json().encode() == statement
    assert [(tuple(c.row), tuple(c.col), c.value) for c in claims] == claims_from_json(claims_json)
    assert len(proof) == len(want)
    assert proof == want
    assert p.verify(committed.statement, claims, proof)
    return p, committed, claims, proof


def test_kb_demo_roots_and_codewords_match_oracle(oracle_kb):
    p = params(DEMO_SHAPES, 10)
    values = demo_bva(lambda mats: layout_encode(p, mats), kb=True)
    print("koalabear demo commit, blocks", p.blocks, flush=True)
    committed = p.commit(values)
    roots, cws = oracle_commit(oracle_kb, 10, values, p.blocks, want_codeword=True)
    assert committed.statement.roots == roots
    for b in range(p.blocks):
        assert np.array_equal(committed.codeword(b), cws[b]), b


@pytest.mark.parametrize("chunk_log,rows,cols,seed,gen", [(12, 64, 64, 1, "int16"), (14, 100, 163, 2, "bf16"),
                                                          (16, 256, 256, 3, "int16"), (18, 1000, 300, 4, "bf16")])
def test_kb_random_roots_and_codewords_match_oracle(oracle_kb, chunk_log, rows, cols, seed, gen):
    p = params([(rows, cols)], chunk_log)
    make = synthetic_bf16 if gen == "bf16" else synthetic_int16
    values = layout_encode(p, [make(rows * cols, seed)])
    print("koalabear random commit", chunk_log, rows, cols, gen, "blocks", p.blocks, flush=True)
    committed = p.commit(values)
    roots, cws = oracle_commit(oracle_kb, p.block_len.bit_length() - 1, values, p.blocks, want_codeword=True)
    assert committed.statement.roots == roots
    for b in range(p.blocks):
        assert np.array_equal(committed.codeword(b), cws[b]), b


@pytest.mark.parametrize("recompute", [True, False])
def test_kb_demo_proof_matches_reference_digest(oracle_kb, recompute):
    p = params(DEMO_SHAPES, 10)
    values = demo_bva(lambda mats: layout_encode(p, mats), kb=True)
    print("koalabear demo proof, recompute", recompute, flush=True)
    _, _, _, proof = check_against_oracle(oracle_kb, DEMO_SHAPES, 10, values, recompute=recompute)
    assert len(proof) == 51646
    assert hashlib.blake2s(proof).hexdigest() == KB_DEMO_PROOF_BLAKE2S


@pytest.mark.parametrize("chunk_log,shapes,seed,packed_log,gen", [
    (12, [(64, 64)], 1, 4, "int16"),            # one PCS block: the single-block opening
    (14, [(100, 163)], 2, 4, "bf16"),           # two PCS levels
    (14, [(300, 200), (7, 1)], 4, 2, "int16"),  # several blocks, packed_log 2
    (16, [(513, 129), (33, 65)], 5, 5, "bf16"),  # ragged shapes, packed_log 5
    (12, [(64, 64)], 6, 1, "bf16"),             # packed_log 1
    (14, [(300, 200)], 7, 6, "int16"),          # packed_log 6
    (18, [(500, 1000)], 8, 4, "bf16"),          # column-padding terms sharing query rows
    (18, [(700, 1500), (3, 5)], 9, 4, "int16"),  # full-range i16 on the MFMA projection, ragged end
])
def test_kb_random_proofs_match_oracle(oracle_kb, chunk_log, shapes, seed, packed_log, gen):
    p = params(shapes, chunk_log, packed_log)
    make = synthetic_bf16 if gen == "bf16" else synthetic_int16
    values = layout_encode(p, [make(r * c, seed + i) for i, (r, c) in enumerate(shapes)])
    print("koalabear random proof", chunk_log, shapes, seed, packed_log, gen, "blocks", p.blocks, flush=True)
    check_against_oracle(oracle_kb, shapes, chunk_log, values, packed_log)


def test_kb_synthetic_bva_is_bf16_of_small_normals():
    p = params([(300, 200), (7, 5)], 14)
    buf = p.synthetic_bva(seed=5)
    v = buf.to_host("int16")
    b = v[v != 0].view(np.uint16).astype(np.uint32) << 16
    x = b.view(np.float32)
    assert np.isfinite(x).all() and 0.015 < np.std(x) < 0.025
    committed = p.commit(buf)
    assert len(committed.statement.roots) == p.blocks
