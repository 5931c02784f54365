"""The Llama-3.1-8B-shaped F65521 configuration at full size: 226 int8
matrices, chunk_log 30, packed_log 4, eight PCS blocks of 2^30 entries.

- The proof has the length the reference published for this configuration
  (results/reference/qafold-measurements.json:17), which fixes the level
  structure and the f64 query counts of every QA level at chunk_log 30.
- The verifier accepts it; one altered byte in each region of the wire is
  rejected (end_to_end.rs:92-167).
- Batched and per-matrix launches, and cached and recomputed commitments,
  give the same bytes (cpu_optimization.rs:5-30).
- The root of block 7, which holds the active end of the BVA, equals the CPU
  oracle's root of the same 2^30 entries."""
import numpy as np
import pytest

from oracle_ref import oracle_commit, proof_regions

pytestmark = pytest.mark.gpu

SEED = 20260


@pytest.fixture(scope="module")
def run8b():
    import parcomm_amd as pa
    from parcomm_amd.workloads import llama8b_shapes
    p = pa.Parameters([pa.Shape(r, c) for r, c in llama8b_shapes()], pa.Config(chunk_log=30, packed_log=4))
    assert (p.active, p.blocks, p.n_buckets) == (8053063680, 8, 424)
    print("8B F65521: synthetic BVA", flush=True)
    bva = p.synthetic_bva(SEED)
    committed = p.commit(bva)
    claims = p.benchmark_claims(committed)
    print("8B F65521: batched proof", flush=True)
    proof, _ = p.prove(committed, claims)
    yield p, bva, committed, claims, proof
    committed.close()
    bva.close()


def test_8b_proof_has_the_published_length_and_verifies(run8b):
    from parcomm_amd.workloads import PUBLISHED_PROOF_BYTES
    p, _, committed, claims, proof = run8b
    assert len(proof) == PUBLISHED_PROOF_BYTES["f65521-8b"]
    assert p.verify(committed.statement, claims, proof)


def test_8b_altered_proof_regions_are_rejected(run8b):
    import parcomm_amd as pa
    p, _, committed, claims, proof = run8b
    for name, at in proof_regions(proof).items():
        bad = bytearray(proof)
        bad[at] ^= 0x01
        with pytest.raises(pa.QfError):
            p.verify(committed.statement, claims, bytes(bad))
        print("rejected altered", name, "byte", at, flush=True)


def test_8b_per_matrix_and_recomputed_give_the_same_proof(run8b):
    import parcomm_amd as pa
    p, bva, committed, claims, proof = run8b
    pa.set_launch_mode(pa.LAUNCH_PER_MATRIX)
    try:
        print("8B F65521: per-matrix proof", flush=True)
        per_matrix, _ = p.prove(committed, claims)
    finally:
        pa.set_launch_mode(pa.LAUNCH_BATCHED)
    assert per_matrix == proof
    print("8B F65521: recomputed commitment and proof", flush=True)
    rc = p.commit_recomputed(bva)
    try:
        assert rc.statement.roots == committed.statement.roots
        again, _ = p.prove(rc, claims)
    finally:
        rc.close()
    assert again == proof


def test_8b_last_data_block_root_matches_oracle(run8b, oracle):
    p, bva, committed, _, _ = run8b
    b = p.active // p.block_len  # block 7: entries [7 2^30, 7.5 2^30) hold data, the rest is zero
    assert b == 7 and p.active % p.block_len
    first = b * p.block_len
    chunk = np.zeros(p.block_len, dtype=np.int8)
    chunk[:p.active - first] = bva.to_host("int8", offset=first, count=p.active - first)
    print("oracle commit of block 7 (2^30 entries) on the host", flush=True)
    roots, _ = oracle_commit(oracle, 30, chunk, 1)
    assert roots[0] == committed.statement.roots[b]
