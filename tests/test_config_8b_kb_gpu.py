"""The Llama-3.1-8B-shaped KoalaBear configuration at full size: 226 matrices
of BF16 bit patterns as int16 entries, chunk_log 30, packed_log 4, eight PCS
blocks of 2^30 entries, batched.

- The proof has the length the reference published for this configuration
  (results/reference/qafold-measurements.json:85), which fixes the KoalaBear
  level structure and query counts at chunk_log 30.
- The KoalaBear verifier accepts it; one altered byte in each region of the
  wire is rejected.
- Per-matrix launches give the same bytes.
- The root of block 7, which holds the active end, equals the KoalaBear
  oracle's root of the same 2^30 entries."""
import numpy as np
import pytest

from oracle_ref import oracle_commit, proof_regions

pytestmark = pytest.mark.gpu

SEED = 20260


@pytest.fixture(scope="module")
def run8b_kb():
    import parcomm_amd as pa
    from parcomm_amd.workloads import llama8b_shapes
    p = pa.Parameters([pa.Shape(r, c) for r, c in llama8b_shapes()], pa.Config(chunk_log=30, packed_log=4),
                      field=pa.KOALABEAR)
    assert (p.active, p.blocks) == (8053063680, 8)
    print("8B KoalaBear: synthetic BVA", flush=True)
    bva = p.synthetic_bva(SEED)
    committed = p.commit(bva)
    claims = p.benchmark_claims(committed)
    print("8B KoalaBear: batched proof", flush=True)
    proof, _ = p.prove(committed, claims)
    yield p, bva, committed, claims, proof
    committed.close()
    bva.close()


def test_8b_kb_proof_has_the_published_length_and_verifies(run8b_kb):
    from parcomm_amd.workloads import PUBLISHED_PROOF_BYTES
    p, _, committed, claims, proof = run8b_kb
    assert len(proof) == PUBLISHED_PROOF_BYTES["koalabear-8b"]
    assert p.verify(committed.statement, claims, proof)


def test_8b_kb_altered_proof_regions_are_rejected(run8b_kb):
    import parcomm_amd as pa
    p, _, committed, claims, proof = run8b_kb
    for name, at in proof_regions(proof).items():
        bad = bytearray(proof)
        bad[at] ^= 0x01
        with pytest.raises(pa.QfError):
            p.verify(committed.statement, claims, bytes(bad))
        print("rejected altered", name, "byte", at, flush=True)


def test_8b_kb_per_matrix_gives_the_same_proof(run8b_kb):
    import parcomm_amd as pa
    p, _, committed, claims, proof = run8b_kb
    pa.set_launch_mode(pa.LAUNCH_PER_MATRIX, pa.KOALABEAR)
    try:
        print("8B KoalaBear: per-matrix proof", flush=True)
        per_matrix, _ = p.prove(committed, claims)
    finally:
        pa.set_launch_mode(pa.LAUNCH_BATCHED, pa.KOALABEAR)
    assert per_matrix == proof


def test_8b_kb_last_data_block_root_matches_oracle(run8b_kb, oracle_kb):
    p, bva, committed, _, _ = run8b_kb
    b = p.active // p.block_len
    assert b == 7
    first = b * p.block_len
    chunk = np.zeros(p.block_len, dtype=np.int16)
    chunk[:p.active - first] = bva.to_host("int16", offset=2 * first, count=p.active - first)
    print("KoalaBear oracle commit of block 7 (2^30 entries) on the host", flush=True)
    roots, _ = oracle_commit(oracle_kb, 30, chunk, 1)
    assert roots[0] == committed.statement.roots[b]
