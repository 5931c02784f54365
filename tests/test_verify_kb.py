"""The KoalaBear build's witness-free verifier (libparcomm_amd_kb.so,
Parameters::verify of qafold/koalabear/src/protocol.rs) on the host: it
accepts the KoalaBear oracle's proofs (which reproduce the reference's
KoalaBear demo digest) and rejects altered proofs and claims."""
import hashlib
import json

import numpy as np
import pytest

from oracle_ref import DEMO_SHAPES, demo_bva, layout_encode, oracle_prove, shapes_flat, synthetic_bf16, synthetic_int16

KB_DEMO = "2a4a9dd7db840a62c12e55c65ac700f82a7150f992886bad420dcab4d2fe0285"


def setup(oracle_kb, shapes, chunk_log, packed_log, values):
    import parcomm_amd as pa
    p = pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log, packed_log=packed_log),
                      field=pa.KOALABEAR)
    proof, statement, claims = oracle_prove(oracle_kb, shapes_flat(shapes), len(shapes), chunk_log, packed_log, values)
    st = json.loads(statement)
    commitment = pa.Commitment(2, p.config, p.shapes, bytes(st["parameter_digest"]), [bytes(r) for r in st["roots"]])
    cl = [pa.Claim([bytes(v) for v in c["row"]], [bytes(v) for v in c["col"]], bytes(c["value"]))
          for c in json.loads(claims)]
    return p, commitment, cl, proof


def kb_params(shapes, chunk_log, packed_log):
    import parcomm_amd as pa
    return pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log, packed_log=packed_log),
                         field=pa.KOALABEAR)


def test_kb_parameter_digest_matches_oracle(oracle_kb):
    p = kb_params(DEMO_SHAPES, 10, 4)
    values = demo_bva(lambda mats: layout_encode(p, mats), kb=True)
    _, statement, _ = oracle_prove(oracle_kb, shapes_flat(DEMO_SHAPES), len(DEMO_SHAPES), 10, 4, values)
    assert bytes(json.loads(statement)["parameter_digest"]) == p.parameter_digest()


def test_kb_demo_proof_is_accepted(oracle_kb):
    p = kb_params(DEMO_SHAPES, 10, 4)
    values = demo_bva(lambda mats: layout_encode(p, mats), kb=True)
    p, st, claims, proof = setup(oracle_kb, DEMO_SHAPES, 10, 4, values)
    assert hashlib.blake2s(proof).hexdigest() == KB_DEMO
    assert p.verify(st, claims, proof)


@pytest.mark.parametrize("chunk_log,shapes,packed_log,seed,gen", [
    (12, [(64, 64)], 4, 1, "bf16"),             # single block
    (14, [(100, 163)], 4, 2, "int16"),          # single block, two QA levels
    (14, [(300, 200), (7, 1)], 2, 4, "bf16"),   # several blocks
    (16, [(513, 129), (33, 65)], 5, 5, "int16"),
    (12, [(64, 64)], 1, 6, "int16"),
    (16, [(1000, 100)], 6, 7, "bf16"),
])
def test_kb_random_proofs_are_accepted(oracle_kb, chunk_log, shapes, packed_log, seed, gen):
    p = kb_params(shapes, chunk_log, packed_log)
    make = synthetic_bf16 if gen == "bf16" else synthetic_int16
    values = layout_encode(p, [make(r * c, seed + i) for i, (r, c) in enumerate(shapes)])
    p, st, claims, proof = setup(oracle_kb, shapes, chunk_log, packed_log, values)
    assert p.verify(st, claims, proof)


def test_kb_altered_proofs_and_claims_are_rejected(oracle_kb):
    import parcomm_amd as pa
    p = kb_params(DEMO_SHAPES, 10, 4)
    values = demo_bva(lambda mats: layout_encode(p, mats), kb=True)
    p, st, claims, proof = setup(oracle_kb, DEMO_SHAPES, 10, 4, values)
    rng = np.random.default_rng(3)
    for pos in sorted(set([0, 8, 100, len(proof) // 2, len(proof) - 1] + list(rng.integers(0, len(proof), 16)))):
        bad = bytearray(proof)
        bad[pos] ^= 0x01
        with pytest.raises(pa.QfError):
            p.verify(st, claims, bytes(bad))
    wrong = list(claims)
    v = bytearray(wrong[1].value)
    v[0] ^= 1
    wrong[1] = pa.Claim(wrong[1].row, wrong[1].col, bytes(v))
    with pytest.raises(pa.QfError):
        p.verify(st, wrong, proof)
    # a coefficient >= p is not a canonical KoalaBear element
    big = bytearray(claims[0].value)
    big[0:4] = (2130706433).to_bytes(4, "little")
    noncanon = list(claims)
    noncanon[0] = pa.Claim(claims[0].row, claims[0].col, bytes(big))
    with pytest.raises(pa.QfError):
        p.verify(st, noncanon, proof)


def test_field_mismatch_is_rejected():
    import parcomm_amd as pa
    L = pa.lib(pa.KOALABEAR)
    import ctypes
    flat = (ctypes.c_uint64 * 2)(64, 64)
    h = ctypes.c_void_p()
    assert L.qf_params_new(pa.F65521, flat, 1, 12, 4, 100, ctypes.byref(h)) == 5
