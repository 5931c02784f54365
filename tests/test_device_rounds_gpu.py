"""The device-resident fold rounds (QF_DEVICE_ROUNDS=1): LiteBullet and the
Decide multi-evaluation folded on the device with a device copy of the PCS
transcript (qafold_short_pc.rs:2011-2229), one launch per round, the host
replaying every challenge. The proof bytes must equal the CPU oracle's and
the host-round proof's, in both fields:

- shapes whose fold vectors need 1 to 128 workgroups per round (the partial
  sums of the last workgroup, 8 words per element row for KoalaBear and 16
  for F65521), with one and two Decide levels;
- the full-size 8B F65521 statement, whose largest rounds run 1024
  workgroups, against the host-round proof."""
import os

import numpy as np
import pytest

from oracle_ref import layout_encode, oracle_prove, shapes_flat, synthetic_int8, synthetic_int16

pytestmark = pytest.mark.gpu

FIELDS = ["f65521", "koalabear"]


def prove_both(p, committed, claims):
    os.environ["QF_DEVICE_ROUNDS"] = "1"
    try:
        device, _ = p.prove(committed, claims)
    finally:
        del os.environ["QF_DEVICE_ROUNDS"]
    host, _ = p.prove(committed, claims)
    return device, host


@pytest.mark.parametrize("field", FIELDS)
@pytest.mark.parametrize("shapes,chunk_log,packed_log", [
    ([(300, 200), (7, 1)], 14, 2),
    ([(513, 129), (33, 65)], 16, 5),
    ([(1000, 300)], 18, 6),
    ([(1024, 4096), (3, 5)], 22, 4),
])
def test_device_rounds_match_the_oracle(oracle, oracle_kb, field, shapes, chunk_log, packed_log):
    import parcomm_amd as pa
    kb = field == "koalabear"
    p = pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log, packed_log=packed_log),
                      field=pa.KOALABEAR if kb else pa.F65521)
    make = synthetic_int16 if kb else synthetic_int8
    values = layout_encode(p, [make(r * c, 40 + i) for i, (r, c) in enumerate(shapes)])
    print("device rounds", field, shapes, chunk_log, packed_log, flush=True)
    committed = p.commit(values)
    claims = p.benchmark_claims(committed)
    device, host = prove_both(p, committed, claims)
    want, _, _ = oracle_prove(oracle_kb if kb else oracle, shapes_flat(shapes), len(shapes), chunk_log, packed_log,
                              values)
    assert host == want
    assert device == want
    assert p.verify(committed.statement, claims, device)


def test_device_rounds_8b_f65521_match_host_rounds():
    import parcomm_amd as pa
    from parcomm_amd.workloads import PUBLISHED_PROOF_BYTES, llama8b_shapes
    p = pa.Parameters([pa.Shape(r, c) for r, c in llama8b_shapes()], pa.Config(chunk_log=30, packed_log=4))
    print("8B F65521 device rounds: synthetic BVA", flush=True)
    bva = p.synthetic_bva(7)
    committed = p.commit(bva)
    try:
        claims = p.benchmark_claims(committed)
        device, host = prove_both(p, committed, claims)
        assert len(host) == PUBLISHED_PROOF_BYTES["f65521-8b"]
        assert device == host
    finally:
        committed.close()
        bva.close()
