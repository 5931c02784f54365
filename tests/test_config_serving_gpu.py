"""Serving on the 8B-shaped F65521 commitment (chunk_log 30, packed_log 4):
opening requests over random subsets of the 226 matrices with fresh points,
served concurrently on the resident codewords, each response byte-identical to
the same request proved alone and accepted by the verifier."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_8b_served_requests_match_their_proofs_alone():
    import parcomm_amd as pa
    from parcomm_amd.serve import request_claims, serve
    from parcomm_amd.workloads import llama8b_shapes
    p = pa.Parameters([pa.Shape(r, c) for r, c in llama8b_shapes()], pa.Config(chunk_log=30, packed_log=4))
    bva = p.synthetic_bva(11)
    committed = p.commit(bva)
    try:
        base = p.benchmark_claims(committed)
        print("8B serving: 6 requests, 4 slots", flush=True)
        stats, reqs = serve(p, committed, base, n_requests=6, rate=200.0, seed=3, workers=4, verify=True)
        assert stats["completed"] == 6 and stats["failed"] == 0
        assert stats["max_in_flight"] >= 2, stats  # requests did overlap on the device
        for r in reqs:
            claims = request_claims(p, committed, base, r.subset, np.random.default_rng(3 * 1_000_003 + r.rid))
            alone, _ = p.prove(committed, claims)
            assert alone == r.proof, r.rid
            print("request", r.rid, "matrices", len(r.subset), "equal alone", flush=True)
        # the responses verify as a batch, and a tampered one is singled out
        items = [(r.claims, r.proof) for r in reqs]
        assert p.verify_batch(committed.statement, items) == [0] * len(items)
        bad = bytearray(reqs[2].proof)
        bad[len(bad) // 3] ^= 1
        items[2] = (reqs[2].claims, bytes(bad))
        statuses = p.verify_batch(committed.statement, items)
        assert statuses[2] != 0 and statuses[:2] + statuses[3:] == [0] * (len(items) - 1)
    finally:
        committed.close()
        bva.close()
