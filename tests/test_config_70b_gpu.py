"""The Llama-3.1-70B-shaped F65521 configuration at full size: 562 int8
matrices (70.6 GB BVA), chunk_log 30, packed_log 6, 128 PCS blocks, roots-only
(recomputed) commitment. The proof has the length the reference published
(results/reference/qafold-measurements.json:160) and the verifier accepts it;
the blocks past the active data share one zero-block root (protocol.rs:169-175)."""
import pytest

pytestmark = pytest.mark.gpu


def test_70b_recomputed_proof_has_the_published_length_and_verifies():
    import parcomm_amd as pa
    from parcomm_amd.workloads import PUBLISHED_PROOF_BYTES, llama70b_shapes
    p = pa.Parameters([pa.Shape(r, c) for r, c in llama70b_shapes()], pa.Config(chunk_log=30, packed_log=6))
    assert (p.active, p.blocks) == (70598524928, 128)
    print("70B: synthetic BVA", flush=True)
    bva = p.synthetic_bva(70)
    try:
        committed = p.commit_recomputed(bva)
        roots = committed.statement.roots
        data_blocks = -(-p.active // p.block_len)
        assert len(set(roots[data_blocks:])) == 1 and roots[data_blocks] not in roots[:data_blocks]
        claims = p.benchmark_claims(committed)
        print("70B: proof", flush=True)
        proof, _ = p.prove(committed, claims)
        assert len(proof) == PUBLISHED_PROOF_BYTES["f65521-70b"]
        assert p.verify(committed.statement, claims, proof)
        committed.close()
    finally:
        bva.close()
